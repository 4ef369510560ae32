"""GPU parity of every algorithm id against the FP64 oracle, through the C-ABI.

Tolerance: normwise relative error ||x - x_ref||_inf / ||x_ref||_inf
over every system <= 1e-5 (FP32) and <= 1e-12 (FP64) on diagonally dominant
inputs; the FP64 relative residual is checked too.
"""
import numpy as np
import pytest

from gpu_util import TOL, announce, check_against_oracle, make_case, oracle_x, run_solve

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}
TILE_ALGOS = list(range(2, 14))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tridiagonalsolver_amd as T
    T.lib()


def _solve_and_check(prec, algo, layout, n, batch, seed, blo=1.0, bhi=2.0):
    dtype = DT[prec]
    announce("solve", "prec", prec, "algo", algo, "layout", layout, "n", n, "batch", batch,
             "seed", seed)
    arrs = make_case(n, batch, layout, dtype, seed, blo, bhi)
    x = run_solve(arrs, n, batch, layout, algo, dtype)
    st, res = check_against_oracle(arrs, x, n, batch, layout, dtype)
    assert st["nonfinite"] == 0, st
    assert st["relerr"] <= TOL[dtype], (st, res)
    assert res <= 10 * TOL[dtype], (st, res)
    return x


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [1, 2, 3, 64, 100, 1000, 4096])
def test_thomas_layout0(prec, n):
    _solve_and_check(prec, 1, 0, n, 37, seed=n)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [1, 5, 4096])
def test_thomas_interleaved(prec, n):
    _solve_and_check(prec, 1, 2, n, 130, seed=n + 1)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo", TILE_ALGOS)
@pytest.mark.parametrize("n", [2, 3, 64, 100, 1000, 2048, 4096])
def test_tile_algos_layout0(prec, algo, n):
    _solve_and_check(prec, algo, 0, n, 9, seed=1000 * algo + n)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [1, 2, 7, 32, 64, 100, 257, 1000, 1024, 2048, 4096])
@pytest.mark.parametrize("batch", [1, 5, 129])
def test_auto_interleaved(prec, n, batch):
    _solve_and_check(prec, -1, 2, n, batch, seed=7 * n + batch)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo", TILE_ALGOS)
@pytest.mark.parametrize("n", [5, 64, 1000, 4096])
def test_tile_algos_interleaved(prec, algo, n):
    _solve_and_check(prec, algo, 2, n, 40, seed=3000 * algo + n)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo", [2, 3, 4, 5, 6, 7, 8, 10])
def test_interface_families_low_margin(prec, algo):
    """Every interface solver on the least dominant distribution (margin U[0.05, 0.2])."""
    _solve_and_check(prec, algo, 0, 2048, 16, seed=77 + algo, blo=0.05, bhi=0.2)


@pytest.mark.parametrize("algo", [14, 15, 16, 17])
@pytest.mark.parametrize("n", [64, 1024, 2000])
def test_bplg_packed_layout(algo, n):
    _solve_and_check(0, algo, 1, n, 11, seed=algo + n)


@pytest.mark.parametrize("prec", [0, 1])
def test_vendor_gtsv(prec):
    _solve_and_check(prec, 0, 0, 512, 16, seed=5)


@pytest.mark.parametrize("prec", [0, 1])
def test_benchmark_shape_interleaved_4096(prec):
    """n=4096 x batch=4096, interleaved: every system against the oracle."""
    _solve_and_check(prec, -1, 2, 4096, 4096, seed=2024)


@pytest.mark.parametrize("prec", [0, 1])
def test_benchmark_shape_layout0_4096(prec):
    _solve_and_check(prec, 9, 0, 4096, 1024, seed=2025)


def test_low_margin_fp64():
    """Dominance margin lowered to U[0.05, 0.2] (large-system workload distribution)."""
    _solve_and_check(1, -1, 2, 4096, 64, seed=31, blo=0.05, bhi=0.2)
    _solve_and_check(1, 9, 0, 4096, 64, seed=32, blo=0.05, bhi=0.2)


@pytest.mark.parametrize("prec,layout", [(0, 0), (0, 2), (1, 0), (1, 2)])
def test_manufactured_replicated_batch(prec, layout):
    """The reference's own test (test.cpp:19-38 + init_batch): one manufactured system
    replicated over the batch.  Every system must recover x and all must agree."""
    import oracle
    n, batch = 1024, 64
    dtype = DT[prec]
    a, b, c, d, x = oracle.gen_manufactured(n, single=(prec == 0))
    rep = lambda v: np.tile(v.astype(dtype), batch) if layout == 0 else \
        np.repeat(v.astype(dtype), batch)
    arrs = tuple(rep(v) for v in (a, b, c, d))
    announce("manufactured", prec, layout)
    xg = run_solve(arrs, n, batch, layout, -1 if layout == 2 else 9, dtype)
    per = xg.reshape(batch, n) if layout == 0 else xg.reshape(n, batch).T
    assert np.all(per == per[0:1])                      # identical systems, identical answers
    err = np.max(np.abs(per[0].astype(np.float64) - x)) / np.max(np.abs(x))
    assert err <= (1e-5 if prec == 0 else 1e-12)


@pytest.mark.parametrize("layout", [0, 2])
def test_ignores_a0_and_cn1(layout):
    n, batch = 300, 20
    arrs = make_case(n, batch, layout, np.float64, seed=3)
    x1 = run_solve(arrs, n, batch, layout, 9, np.float64)
    a, b, c, d = (v.copy() for v in arrs)
    for s in range(batch):
        i0 = s * n if layout == 0 else s
        il = s * n + n - 1 if layout == 0 else (n - 1) * batch + s
        a[i0] = 1e30 * (s + 1)
        c[il] = -1e30
    x2 = run_solve((a, b, c, d), n, batch, layout, 9, np.float64)
    assert np.array_equal(x1, x2)


@pytest.mark.parametrize("layout", [0, 2])
def test_deterministic(layout):
    n, batch = 1000, 33
    arrs = make_case(n, batch, layout, np.float32, seed=4)
    x1 = run_solve(arrs, n, batch, layout, -1, np.float32)
    x2 = run_solve(arrs, n, batch, layout, -1, np.float32)
    assert np.array_equal(x1, x2)


def test_nan_stays_in_its_system():
    n, batch = 512, 40
    arrs = list(make_case(n, batch, 2, np.float32, seed=5))
    arrs[3] = arrs[3].copy()
    arrs[3][100 * batch + 17] = np.nan          # system 17, row 100
    x = run_solve(tuple(arrs), n, batch, 2, -1, np.float32).reshape(n, batch)
    assert np.isnan(x[:, 17]).any()
    ok = np.delete(x, 17, axis=1)
    assert np.isfinite(ok).all()
    clean = [v.copy() for v in arrs]
    clean[3][100 * batch + 17] = 0.0
    xref = oracle_x(clean, n, batch, 2).reshape(n, batch)
    assert np.max(np.abs(np.delete(xref, 17, axis=1) - ok)) < 1e-4
