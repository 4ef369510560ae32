"""Seeded random sweep over the solve's whole input space, each case checked against
the FP64 oracle (normwise relative error <= 1e-5 FP32 / 1e-12 FP64): precision,
layout (contiguous, packed, interleaved), every algorithm id valid for the layout,
n log-uniform in [1, 100000] (on-chip tiles exact and padded, lane kernels, chunked
solve, interleaved systems routed through the contiguous layout), batch up to
about 2^21 unknowns, and both dominance margins.  The adjoint is swept the same way
against the FP64 adjoint."""
import numpy as np
import pytest

import oracle
from gpu_util import TOL, announce, check_against_oracle, make_case, run_solve

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}
N_CASES = 120


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _cases(seed, count):
    import tridiagonalsolver_amd as T
    L = T.lib()
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        prec = int(rng.integers(0, 2))
        layout = int(rng.choice([0, 0, 1, 2, 2]))
        n = int(np.exp(rng.uniform(0, np.log(100000))))
        batch = int(max(1, min(2000, (1 << 21) // max(n, 1) * rng.uniform(0.1, 1.0))))
        algos = [-1] + [a for a in range(L.tridiag_algo_count(prec))
                        if L.tridiag_algo_layout(prec, a) == (1 if layout == 1 else 0)]
        if layout == 2:
            algos = [a for a in algos if a != 0]           # the vendor routine is layout 0
        if layout == 1:
            algos = [a for a in algos if a != 0] + [1, 9]  # BPLG ids plus two on-chip ids
        algo = int(rng.choice(algos))
        low = bool(rng.random() < 0.3)
        out.append((prec, layout, n, batch, algo, low))
    return out


@pytest.mark.parametrize("chunk", range(4))
def test_random_solves(chunk):
    cases = _cases(1000 + chunk, N_CASES // 4)
    fails = []
    for prec, layout, n, batch, algo, low in cases:
        if algo == 0 and not 3 <= n <= 65536:
            continue                                       # the vendor routine's own limits
        announce("sweep", prec, layout, n, batch, algo, low)
        blo, bhi = (0.05, 0.2) if low else (1.0, 2.0)
        arrs = make_case(n, batch, layout, DT[prec], seed=n * 7 + batch, blo=blo, bhi=bhi)
        x = run_solve(arrs, n, batch, layout, algo, DT[prec])
        st, _ = check_against_oracle(arrs, x, n, batch, layout, DT[prec])
        if not (st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]]):
            fails.append((prec, layout, n, batch, algo, low, st["relerr"]))
    assert not fails, fails


@pytest.mark.parametrize("chunk", range(2))
def test_random_adjoints(chunk):
    import torch
    import tridiagonalsolver_amd as T
    rng = np.random.default_rng(77 + chunk)
    fails = []
    for _ in range(12):
        prec = int(rng.integers(0, 2))
        layout = int(rng.choice([0, 2]))
        n = int(np.exp(rng.uniform(0, np.log(40000))))
        batch = int(max(1, min(500, (1 << 20) // max(n, 1))))
        dtype = DT[prec]
        announce("sweep adjoint", prec, layout, n, batch)
        a, b, c, d = make_case(n, batch, layout, dtype, seed=n + batch)
        g = rng.standard_normal(n * batch).astype(dtype)
        x64 = oracle.thomas(a, b, c, d, n, batch, layout=layout)
        ref = oracle.adjoint(a, b, c, x64, g, n, batch, layout=layout)
        ts = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (a, b, c, d, g)]
        outs = T.solve_fwd_bwd(*ts, layout=layout, n=n, batch=batch)
        torch.cuda.synchronize()
        for k, (o, r) in enumerate(zip(outs, (x64,) + tuple(ref))):
            err = np.max(np.abs(o.cpu().numpy().astype(np.float64) - r)) / max(np.max(np.abs(r)), 1e-300)
            if not err <= TOL[dtype]:
                fails.append((prec, layout, n, batch, k, err))
    assert not fails, fails


def test_random_ragged_batches():
    """Random ragged batches: sizes uniform in [0, n_max] (empty systems anywhere,
    the first and last included), both precisions, every system against the oracle."""
    import torch
    import tridiagonalsolver_amd as T
    rng = np.random.default_rng(4242)
    fails = []
    for it in range(12):
        prec = int(rng.integers(0, 2))
        dtype = DT[prec]
        nsys = int(rng.integers(1, 2000))
        nmax = int(rng.choice([16, 100, 1000, 8192]))
        sizes = rng.integers(0, nmax + 1, nsys)
        sizes[0] = 0
        sizes[-1] = 0
        off = oracle.offsets_from_sizes(sizes.astype(np.int64))
        announce("sweep ragged", prec, nsys, nmax)
        a, b, c, d = (v.astype(dtype) for v in oracle.gen_dd_ragged(off, seed=it))
        xr = oracle.thomas_ragged(*(v.astype(np.float64) for v in (a, b, c, d)), off)
        ts = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
        if int(off[-1]) == 0:
            continue
        T.RaggedPlan(off, dtype=ts[0].dtype, device="cuda").solve(*ts)
        torch.cuda.synchronize()
        x = ts[3].cpu().numpy().astype(np.float64)
        for s in np.nonzero(sizes)[0]:
            sl = slice(off[s], off[s + 1])
            err = np.max(np.abs(x[sl] - xr[sl])) / max(np.max(np.abs(xr[sl])), 1e-300)
            if not err <= TOL[dtype]:
                fails.append((it, int(s), float(err)))
                break
    assert not fails, fails
