"""The restated reference CPU solvers (oracle/cpu_ref.cpp: Thomas, PCR-pThomas and
WM-pGE of Tridiagonal_AVX512) against the FP64 oracle, on the CPU.  They are the
measured CPU baseline, so their results must agree with the oracle to the same
bars as the GPU path: normwise relative error <= 1e-5 (FP32) / 1e-12 (FP64)."""
import numpy as np
import pytest

import oracle
from oracle import cpu_ref

TOL = {np.float32: 1e-5, np.float64: 1e-12}
ALGOS = [cpu_ref.THOMAS, cpu_ref.PCR_PTHOMAS, cpu_ref.WM_PGE]


@pytest.fixture(scope="module", autouse=True)
def _built():
    cpu_ref.build(force=False)


def _relerr(x, xref, n, batch):
    return oracle.errstats(x.astype(np.float64), xref, n, batch, 0)["relerr"]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("n", [32, 64, 96, 1024, 4096])
def test_against_oracle(dtype, algo, n):
    batch = 9
    a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=n + algo)
    xref = oracle.thomas(a, b, c, d, n, batch, layout=0)
    x = cpu_ref.solve(*(v.astype(dtype) for v in (a, b, c, d)), n, batch, algo)
    assert _relerr(x, xref, n, batch) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("algo", ALGOS)
def test_low_margin(dtype, algo):
    """The large-system distribution (b = |a|+|c|+U[0.05, 0.2]): long error chains."""
    n, batch = 2048, 5
    a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=77, blo=0.05, bhi=0.2)
    xref = oracle.thomas(a, b, c, d, n, batch, layout=0)
    x = cpu_ref.solve(*(v.astype(dtype) for v in (a, b, c, d)), n, batch, algo)
    assert _relerr(x, xref, n, batch) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("algo", ALGOS)
def test_poisson_closed_form(dtype, algo):
    """tridiag(-1, 2, -1) with d = e_1 + e_n has x = 1 (BPLGTridiagWM.hxx:28-54)."""
    n, batch = 64, 3
    a = np.full(n * batch, -1.0)
    c = np.full(n * batch, -1.0)
    b = np.full(n * batch, 2.0)
    d = np.zeros(n * batch)
    d[0::n] = 1.0
    d[n - 1::n] = 1.0
    x = cpu_ref.solve(*(v.astype(dtype) for v in (a, b, c, d)), n, batch, algo)
    # a_0 = c_(n-1) = -1 here: both must be ignored, as the oracle ignores them
    assert np.max(np.abs(x - 1.0)) <= 1e3 * TOL[dtype]


@pytest.mark.parametrize("algo", ALGOS)
def test_block_diagonal_split(algo):
    """c_k = a_(k+1) = 0 splits a system in two; each half is solved on its own,
    which a wrong neighbour index in a merge or PCR step would break."""
    n, batch = 128, 2
    a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=5)
    for k in (15, 16, 63, 64, 100):
        c[k] = 0.0
        a[k + 1] = 0.0
    x = cpu_ref.solve(a, b, c, d, n, batch, algo)
    xref = oracle.thomas(a, b, c, d, n, batch, layout=0)
    assert _relerr(x, xref, n, batch) <= 1e-12


def test_dense_solve_tiny():
    """n = 2L (the smallest WM-pGE chunk), against numpy's dense solve."""
    for dtype, L in ((np.float64, 8), (np.float32, 16)):
        n, batch = 2 * L, 4
        a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=8)
        for algo in ALGOS:
            x = cpu_ref.solve(*(v.astype(dtype) for v in (a, b, c, d)), n, batch, algo)
            for s in range(batch):
                sl = slice(s * n, (s + 1) * n)
                A = np.diag(b[sl]) + np.diag(a[sl][1:], -1) + np.diag(c[sl][:-1], 1)
                xs = np.linalg.solve(A, d[sl])
                err = np.max(np.abs(x[sl] - xs)) / np.max(np.abs(xs))
                assert err <= TOL[dtype], (dtype, algo, s, err)


def test_unsupported_sizes():
    a = np.ones(40)
    with pytest.raises(ValueError):
        cpu_ref.solve(a, a, a, a, 40, 1, cpu_ref.WM_PGE)       # FP64 WM-pGE needs n % 16
    with pytest.raises(ValueError):
        cpu_ref.solve(a, a, a, a, 20, 2, cpu_ref.PCR_PTHOMAS)  # FP64 PCR needs n % 8
    with pytest.raises(ValueError):
        cpu_ref.solve(a, a, a, a, 40, 1, 4)


def test_bench_harness_solves_the_input():
    n, batch = 512, 16
    a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=12)
    t, tmin, x = cpu_ref.bench(a, b, c, d, n, batch, cpu_ref.WM_PGE, nsets=2, rounds=2)
    assert 0 < tmin <= t
    xref = oracle.thomas(a, b, c, d, n, batch, layout=0)
    assert _relerr(x, xref, n, batch) <= 1e-12


def test_cpu_info_fields():
    info = cpu_ref.cpu_info()
    assert info["usable_cpus"] >= 1 and isinstance(info["avx512f"], bool)
    assert cpu_ref.simd() == ("avx512" if info["avx512f"] else "generic")


def test_generic_build_matches():
    """The build without AVX-512 (used on hosts that lack it) solves the same way."""
    import ctypes
    cpu_ref.build(force=False)
    L = ctypes.CDLL(cpu_ref._SO["generic"])
    L.cpuref_solve.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 4 + [ctypes.c_int]
    L.cpuref_avx512.restype = ctypes.c_int
    assert L.cpuref_avx512() == 0
    n, batch = 256, 3
    a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=21)
    xref = oracle.thomas(a, b, c, d, n, batch, layout=0)
    for algo in ALGOS:
        cc, x = c.copy(), d.copy()
        P = lambda v: ctypes.c_void_p(v.ctypes.data)
        assert L.cpuref_solve(1, algo, n, batch, P(a), P(b), P(cc), P(x), 0) == 0
        assert _relerr(x, xref, n, batch) <= 1e-12
