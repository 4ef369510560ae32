"""CPU-side checks of the C-ABI: the library loads, exports every symbol the header
declares, and its host-side logic (algorithm table, argument validation) works
without a GPU.  No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tridiagonal.h")


def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([A-Za-z_][A-Za-z_0-9]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def capi():
    from tridiagonalsolver_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        from tridiagonalsolver_amd import build
        build.build()
    return _capi


def test_header_symbols_exported(capi):
    L = capi.lib()
    names = _declared_symbols()
    assert len(names) >= 48
    for nm in names:
        assert hasattr(L, nm), nm
    assert set(capi.REFERENCE_SYMBOLS) <= set(names)
    assert set(capi.EXTENDED_SYMBOLS) <= set(names)


def test_reference_table(capi):
    s = capi.algo_names(capi.SINGLE)
    d = capi.algo_names(capi.DOUBLE)
    # order and names of Tridiagonal_CUDA/helpers.h:23-66 (vendor comparator renamed)
    assert s[1:14] == ["Thomas", "CR", "shmem PCR-pThomas", "shmem PCR-half-pThomas",
                       "reg PCR-pThomas", "reg PCR-half-pThomas", "shmem WM-pGE",
                       "shmem WM-half-pGE", "reg WM-pGE", "reg WM-half-pGE", "reg-4 CR",
                       "reg-8 CR", "reg-16 CR"]
    assert s[14:] == ["BPLG WM", "BPLG LF", "BPLG CR", "BPLG PCR"]
    assert d[1:] == s[1:14]
    L = capi.lib()
    assert L.tridiag_algo_layout(0, 14) == 1 and L.tridiag_algo_layout(0, 9) == 0
    assert L.tridiag_algo_name(1, 14) is None


def test_argument_validation_without_gpu(capi):
    L = capi.lib()
    nul = ctypes.c_void_p(0)
    assert L.tridiag_solve(7, 9, 0, 16, 4, nul, nul, nul, nul, nul) == capi.EINVAL
    assert L.tridiag_solve(0, 99, 0, 16, 4, nul, nul, nul, nul, nul) == capi.EINVAL
    assert L.tridiag_solve(0, 9, 0, 0, 4, nul, nul, nul, nul, nul) == capi.EINVAL
    assert L.tridiag_solve(0, 9, 0, 16, 4, nul, nul, nul, nul, nul) == capi.EINVAL
    assert b"null" in L.tridiag_last_error()
    assert L.tridiag_solve(0, 9, 5, 16, 4, nul, nul, nul, nul, nul) == capi.EINVAL


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("layout", [0, 2])
def test_planner_covers_benchmark_shapes(capi, prec, layout):
    for n in (1, 2, 3, 31, 64, 100, 512, 1000, 1024, 4096):
        p = capi.plan(prec, 9, layout, n, 4096)
        S, NW, M = p["S"], p["NW"], p["M"]
        cap = (64 // S) * NW * M if layout == 2 else 64 * (NW // S) * M
        assert cap >= n
    if layout == 2:
        # 32-byte row groups per workgroup in the interleaved layout
        assert capi.plan(prec, 9, 2, 4096, 4096)["S"] * (4 if prec == 0 else 8) == 32


def test_register_packed_cr_rows_per_lane(capi):
    for algo, R in ((11, 4), (12, 8), (13, 16)):
        assert capi.plan(0, algo, 0, 1024, 8)["M"] == R
