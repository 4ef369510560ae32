"""GPU parity of the adjoint (backward) kernels and the autograd binding.

Checked against the FP64 CPU adjoint (oracle.adjoint: Thomas on A^T plus the
gradient formulas, itself pinned by dense solves and central differences) and
against FP64 central finite differences of the GPU forward solve.
Tolerances: normwise relative error <= 1e-5 (FP32) / 1e-12 (FP64) per gradient.
"""
import numpy as np
import pytest

import oracle
from gpu_util import announce, make_case

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}
TOL = {0: 1e-5, 1: 1e-12}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _rel(x, ref):
    den = max(np.max(np.abs(ref)), 1e-300)
    return np.max(np.abs(x - ref)) / den


def _to(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v)).cuda()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("layout", [0, 2])
@pytest.mark.parametrize("n", [2, 3, 7, 64, 255, 257, 512, 1000, 2049, 4096, 8191])
def test_adjoint_and_fused(prec, layout, n):
    import torch
    import tridiagonalsolver_amd as T
    # sizes beyond the fused adjoint kernels (4096 rows FP64, 1024 interleaved) take
    # the composed path: the plain solver on A^T plus element-wise gradients
    batch = 37
    dtype = DT[prec]
    announce("adjoint", prec, layout, n)
    a, b, c, d = make_case(n, batch, layout, dtype, seed=500 + n)
    g = np.random.default_rng(n).standard_normal(n * batch).astype(dtype)
    x64 = oracle.thomas(a, b, c, d, n, batch, layout=layout)
    ref = oracle.adjoint(a, b, c, x64, g, n, batch, layout=layout)
    # separate backward, given x (rounded to the working precision)
    ta, tb, tc, tg = (_to(v) for v in (a, b, c, g))
    tx = _to(x64.astype(dtype))
    outs = T.solve_adjoint(ta, tb, tc, tx, tg, layout=layout, n=n, batch=batch)
    torch.cuda.synchronize()
    for k, o in enumerate(outs):
        assert _rel(o.cpu().numpy().astype(np.float64), ref[k]) <= TOL[prec], k
    # fused forward + backward
    td = _to(d)
    xo, ga, gb, gc, gd = T.solve_fwd_bwd(ta, tb, tc, td, tg, layout=layout, n=n, batch=batch)
    torch.cuda.synchronize()
    assert _rel(xo.cpu().numpy().astype(np.float64), x64) <= TOL[prec]
    for k, o in enumerate((ga, gb, gc, gd)):
        assert _rel(o.cpu().numpy().astype(np.float64), ref[k]) <= TOL[prec], k


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [20001, 70000])
def test_adjoint_long_systems(prec, n):
    """Long contiguous systems: the composed adjoint over the chunked solve."""
    import torch
    import tridiagonalsolver_amd as T
    batch = 3
    dtype = DT[prec]
    announce("adjoint long", prec, n)
    a, b, c, d = make_case(n, batch, 0, dtype, seed=n, blo=0.05, bhi=0.2)
    g = np.random.default_rng(n).standard_normal(n * batch).astype(dtype)
    x64 = oracle.thomas(a, b, c, d, n, batch, layout=0)
    ref = oracle.adjoint(a, b, c, x64, g, n, batch, layout=0)
    ta, tb, tc, td, tg = (_to(v) for v in (a, b, c, d, g))
    xo, ga, gb, gc, gd = T.solve_fwd_bwd(ta, tb, tc, td, tg, layout=0, n=n, batch=batch)
    outs = T.solve_adjoint(ta, tb, tc, _to(x64.astype(dtype)), tg, layout=0, n=n, batch=batch)
    torch.cuda.synchronize()
    assert _rel(xo.cpu().numpy().astype(np.float64), x64) <= TOL[prec]
    for k, o in enumerate((ga, gb, gc, gd)):
        assert _rel(o.cpu().numpy().astype(np.float64), ref[k]) <= TOL[prec], ("fused", k)
    for k, o in enumerate(outs):
        assert _rel(o.cpu().numpy().astype(np.float64), ref[k]) <= TOL[prec], ("adjoint", k)
    # inputs are left as they were
    assert np.array_equal(tc.cpu().numpy(), c) and np.array_equal(td.cpu().numpy(), d)


@pytest.mark.parametrize("prec", [0, 1])
def test_autograd_matches_oracle(prec):
    import torch
    import tridiagonalsolver_amd as T
    n, batch = 512, 64
    dtype = DT[prec]
    a, b, c, d = make_case(n, batch, 0, dtype, seed=9)
    w = np.random.default_rng(9).standard_normal(n * batch).astype(dtype)
    ts = [_to(v).reshape(batch, n).requires_grad_(True) for v in (a, b, c, d)]
    x = T.tridiagonal_solve(*ts)
    loss = (x * _to(w).reshape(batch, n)).sum()
    loss.backward()
    x64 = oracle.thomas(a, b, c, d, n, batch)
    ref = oracle.adjoint(a, b, c, x64, w, n, batch)
    for k in range(4):
        got = ts[k].grad.detach().cpu().numpy().ravel().astype(np.float64)
        assert _rel(got, ref[k]) <= TOL[prec], k
    # inputs untouched by the forward pass
    assert np.array_equal(ts[3].detach().cpu().numpy().ravel(), d)


def test_full_batch_fp32():
    """The differentiable-solve workload at full size: FP32, n = 512, batch = 8192,
    L = sum w*x with w ~ N(0, 1); the fused forward+backward and the autograd path
    against the FP64 CPU adjoint."""
    import torch
    import tridiagonalsolver_amd as T
    n, batch = 512, 8192
    announce("adjoint full batch", 0, n, batch)
    a, b, c, d = make_case(n, batch, 0, np.float32, seed=512)
    w = np.random.default_rng(512).standard_normal(n * batch).astype(np.float32)
    x64 = oracle.thomas(a, b, c, d, n, batch)
    ref = oracle.adjoint(a, b, c, x64, w, n, batch)
    ta, tb, tc, td, tw = (_to(v) for v in (a, b, c, d, w))
    xo, ga, gb, gc, gd = T.solve_fwd_bwd(ta, tb, tc, td, tw, layout=0, n=n, batch=batch)
    torch.cuda.synchronize()
    assert _rel(xo.cpu().numpy().astype(np.float64), x64) <= TOL[0]
    for k, o in enumerate((ga, gb, gc, gd)):
        assert _rel(o.cpu().numpy().astype(np.float64), ref[k]) <= TOL[0], ("fused", k)
    ts = [t.reshape(batch, n).clone().requires_grad_(True) for t in (ta, tb, tc, td)]
    x = T.tridiagonal_solve(*ts)
    (x * tw.reshape(batch, n)).sum().backward()
    for k in range(4):
        got = ts[k].grad.detach().cpu().numpy().ravel().astype(np.float64)
        assert _rel(got, ref[k]) <= TOL[0], ("autograd", k)


def test_fd_fp64():
    """FP64 central differences (h=1e-6) of L = sum w*x on 64 sampled coefficients."""
    import torch
    import tridiagonalsolver_amd as T
    n, batch = 512, 8
    a, b, c, d = make_case(n, batch, 0, np.float64, seed=21)
    w = np.random.default_rng(21).standard_normal(n * batch)
    ts = [_to(v).reshape(batch, n).requires_grad_(True) for v in (a, b, c, d)]
    x = T.tridiagonal_solve(*ts)
    (x * _to(w).reshape(batch, n)).sum().backward()
    grads = [t.grad.cpu().numpy().ravel() for t in ts]
    rng = np.random.default_rng(5)
    h = 1e-6

    def loss(arrs):
        tt = [_to(v) for v in arrs]
        T.solve(*tt, n=n, batch=batch)
        torch.cuda.synchronize()
        return float(np.dot(w, tt[3].cpu().numpy()))
    for _ in range(64):
        k = int(rng.integers(4))
        i = int(rng.integers(n * batch))
        if (k == 0 and i % n == 0) or (k == 2 and i % n == n - 1):
            continue
        p = [v.copy() for v in (a, b, c, d)]
        m = [v.copy() for v in (a, b, c, d)]
        p[k][i] += h
        m[k][i] -= h
        fd = (loss(p) - loss(m)) / (2 * h)
        assert grads[k][i] == pytest.approx(fd, rel=1e-5, abs=1e-8), (k, i)
