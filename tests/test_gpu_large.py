"""GPU parity of the chunked solve for systems above the on-chip capacity
(three passes: chunk reduction, chunk-interface merge, chunk solve), through the
C-ABI, against the FP64 oracle.  Tolerance: normwise relative error over every
system <= 1e-5 (FP32) / 1e-12 (FP64).
"""
import numpy as np
import pytest

from gpu_util import TOL, announce, check_against_oracle, make_case, run_solve

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}
CHUNK = {0: 8192, 1: 4096}       # rows per chunk of the default chunk shape
ONCHIP = {0: 32768, 1: 16384}    # largest on-chip tile of the contiguous layouts


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tridiagonalsolver_amd as T
    T.lib()


def _check(prec, algo, layout, n, batch, seed, blo=1.0, bhi=2.0):
    dtype = DT[prec]
    announce("large", "prec", prec, "algo", algo, "layout", layout, "n", n, "batch", batch)
    arrs = make_case(n, batch, layout, dtype, seed, blo, bhi)
    x = run_solve(arrs, n, batch, layout, algo, dtype)
    st, res = check_against_oracle(arrs, x, n, batch, layout, dtype)
    assert st["nonfinite"] == 0, st
    assert st["relerr"] <= TOL[dtype], (st, res)


@pytest.mark.parametrize("n,batch", [(1 << 20, 16), (1 << 16, 256)])
def test_large_fp64_low_margin(n, batch):
    """The large-system workloads: FP64, dominance margin U[0.05, 0.2]."""
    _check(1, -1, 0, n, batch, seed=n + batch, blo=0.05, bhi=0.2)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("extra", [1, 2, 3])
def test_just_above_onchip(prec, extra):
    n = ONCHIP[prec] + extra
    _check(prec, 9, 0, n, 5, seed=extra)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("k,delta", [(5, -1), (5, 0), (5, 1), (8, 7), (9, 1 - 8192)])
def test_chunk_boundaries(prec, k, delta):
    """n around multiples of the chunk size (a last chunk of 1 row, a full one, ...)."""
    n = k * CHUNK[prec] + delta
    if n <= ONCHIP[prec]:
        n += ONCHIP[prec]
    _check(prec, -1, 0, n, 3, seed=k * 100 + delta)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo", [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_every_interface_solver(prec, algo):
    n = 3 * CHUNK[prec] + ONCHIP[prec] + 5
    _check(prec, algo, 0, n, 2, seed=algo, blo=0.05, bhi=0.2)


def test_max_chunks_fp64():
    """1024 chunks: the largest system the chunked solve accepts."""
    _check(1, -1, 0, 1024 * CHUNK[1], 1, seed=77)


def test_packed_layout_fp32():
    _check(0, 14, 1, ONCHIP[0] + 1000, 3, seed=78)


def test_over_limit_is_rejected():
    import torch
    import tridiagonalsolver_amd as T
    n = 1025 * CHUNK[1]
    big = torch.zeros(n, dtype=torch.float64, device="cuda")
    with pytest.raises(T.TridiagError) as e:
        T.solve(big, big, big, big, algo=-1, layout=0, n=n, batch=1)
    assert e.value.code == 2


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [4097, 8191, 20000, 70001])
def test_long_interleaved_systems(prec, n):
    """Interleaved systems longer than the interleaved tiles hold: converted to the
    contiguous layout, solved there (chunked when long) and converted back."""
    _check(prec, -1, 2, n, 5, seed=n, blo=0.05, bhi=0.2)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("frac", [0, 1, 2])
@pytest.mark.parametrize("algo", [-1, 3, 13])
def test_chunked_below_onchip_capacity(prec, frac, algo):
    """Contiguous systems above two chunks take the chunked solve even though an
    on-chip tile would hold them (16384 < n <= 32768 FP32, 8192 < n <= 16384 FP64)."""
    lo = 2 * CHUNK[prec]
    n = [lo + 1, lo + lo // 2 + 3, ONCHIP[prec]][frac]
    _check(prec, algo, 0, n, 3, seed=n + 11 * frac, blo=0.05, bhi=0.2)


@pytest.mark.parametrize("prec", [0, 1])
def test_plan_reports_sizes(prec):
    import ctypes
    import tridiagonalsolver_amd as T
    from tridiagonalsolver_amd import _capi
    lo = 2 * CHUNK[prec]
    assert T.plan(prec, -1, 0, lo, 64)["kind"] == "tile"   # two chunks' worth: on-chip tile
    # above: K workgroups per system (Wang-Mou family) or the chunked solve (others)
    assert T.plan(prec, -1, 0, lo + 1, 64)["kind"] == "cooperative"
    assert T.plan(prec, 5, 0, lo + 1, 64)["kind"] == "chunked"
    i = ctypes.c_int()
    assert _capi.lib().tridiag_plan(prec, 5, 0, lo + 1, 64, ctypes.byref(i), ctypes.byref(i),
                                    ctypes.byref(i)) == _capi.EUNSUPPORTED
    assert T.plan(prec, -1, 2, 512, 64)["kind"] == "tile"
    p = T.plan(prec, -1, 2, 4096, 64)
    assert p["kind"] == "cooperative" and p["K"] * p["block_rows"] == 4096
    assert T.plan(prec, -1, 2, 5000, 64)["kind"] == "via_contiguous"
    assert T.plan(prec, 1, 0, 1000, 64)["kind"] == "thomas"
    assert T.plan(prec, -1, 0, 16, 64)["kind"] == "lane"


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo,n", [(11, 4097), (11, 6000), (12, 8193), (13, 16385), (11, 10638)])
def test_fixed_rows_algorithms_above_their_tiles(prec, algo, n):
    """CR_reg R fixes R rows per lane, so its tiles hold at most 64*16*R rows; longer
    contiguous systems take the chunked solve (one chunk or more) with its CR
    interface family instead of failing."""
    _check(prec, algo, 0, n, 3, seed=algo * 1000 + n)
