"""Arrays of 2 GiB or more, through the C-ABI: the paths whose index arithmetic
differs from the tested shapes below 2 GiB.

  * contiguous batches are split into launches whose arrays stay below 2 GiB;
  * interleaved batches take the 64-bit-addressed ("BIG") tiles;
  * the adjoint takes its composed form (solve on A^T plus element-wise passes).

Each batch replicates one seeded system, so the check needs no oracle pass over
gigabytes: every system must come out bit-identical to the first, and the first
must match the FP64 oracle's solve of that one system to 1e-5 (FP32)."""
import numpy as np
import pytest

import oracle
from gpu_util import TOL, announce

pytestmark = pytest.mark.gpu

TWO_GIB = 1 << 31


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tridiagonalsolver_amd as T
    T.lib()
    yield
    torch.cuda.empty_cache()


def _system(n, seed):
    return oracle.gen_dd(n, 1, layout=0, seed=seed)


def _replicate(torch, v, batch, layout):
    t = torch.from_numpy(v.astype(np.float32)).cuda()
    if layout == 0:
        return t.repeat(batch)                                   # s*n + i
    return t.repeat_interleave(batch)                            # i*batch + s


def _check_replicated(torch, x, xref, n, batch, layout):
    X = x.view(batch, n) if layout == 0 else x.view(n, batch).t()
    first = X[0].clone()
    assert bool((X == first).all()), "systems of a replicated batch differ"
    err = np.max(np.abs(first.cpu().numpy().astype(np.float64) - xref)) / np.max(np.abs(xref))
    assert err <= TOL[np.float32], err


def test_contiguous_split_below_2gib():
    import torch
    import tridiagonalsolver_amd as T
    n = 4096
    batch = TWO_GIB // (4 * n) + 1000                  # each array 2.16 GB
    assert n * batch * 4 > TWO_GIB
    announce("big contiguous", n, batch)
    sysv = _system(n, 61)
    xref = oracle.thomas(*sysv, n, 1, layout=0)
    a, b, c, d = (_replicate(torch, v, batch, 0) for v in sysv)
    T.solve(a, b, c, d, layout=0, n=n, batch=batch)
    torch.cuda.synchronize()
    del a, b, c
    _check_replicated(torch, d, xref, n, batch, 0)


def test_interleaved_big_tile():
    import torch
    import tridiagonalsolver_amd as T
    n = 1024
    batch = TWO_GIB // (4 * n) + 4096                  # each array 2.16 GB
    assert n * batch * 4 > TWO_GIB
    announce("big interleaved", n, batch, T.plan(0, -1, 2, n, batch))
    assert T.plan(0, -1, 2, n, batch)["kind"] == "tile"
    sysv = _system(n, 62)
    xref = oracle.thomas(*sysv, n, 1, layout=0)
    a, b, c, d = (_replicate(torch, v, batch, 2) for v in sysv)
    T.solve(a, b, c, d, layout=2, n=n, batch=batch)
    torch.cuda.synchronize()
    del a, b, c
    _check_replicated(torch, d, xref, n, batch, 2)


def test_adjoint_composed_beyond_2gib():
    import torch
    import tridiagonalsolver_amd as T
    n = 512
    batch = TWO_GIB // (4 * n) + 512                   # each array 2.15 GB
    announce("big adjoint", n, batch)
    sysv = _system(n, 63)
    a1, b1, c1, d1 = sysv
    x1 = oracle.thomas(a1, b1, c1, d1, n, 1, layout=0)
    g1 = np.random.default_rng(7).standard_normal(n)
    refs = oracle.adjoint(a1, b1, c1, x1, g1.astype(np.float32).astype(np.float64), n, 1)
    a, b, c = (_replicate(torch, v, batch, 0) for v in (a1, b1, c1))
    x = _replicate(torch, x1, batch, 0)
    g = _replicate(torch, g1, batch, 0)
    outs = T.solve_adjoint(a, b, c, x, g, layout=0, n=n, batch=batch)
    torch.cuda.synchronize()
    del a, b, c, x, g
    for o, r in zip(outs, refs):
        X = o.view(batch, n)
        first = X[0].clone()
        assert bool((X == first).all())
        err = np.max(np.abs(first.cpu().numpy().astype(np.float64) - r)) / np.max(np.abs(r))
        assert err <= TOL[np.float32], err
