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
    # forced tile solves: interface solver id, precision and tile index are checked
    args = (0, 16, 4, nul, nul, nul, nul, nul)
    assert L.tridiag_solve_tile_iface(0, 0, 5, *args) == capi.EINVAL
    assert L.tridiag_solve_tile_iface(0, 0, -1, *args) == capi.EINVAL
    assert b"interface" in L.tridiag_last_error()
    assert L.tridiag_solve_tile_iface(3, 0, 2, *args) == capi.EINVAL
    assert L.tridiag_solve_tile_iface(0, -1, 2, *args) == capi.EINVAL
    # the vendor comparator (algo 0) needs layout 0 and at least 3 rows
    assert L.tridiag_solve(0, 0, 2, 16, 4, nul, nul, nul, nul, nul) == capi.EUNSUPPORTED
    assert L.tridiag_solve(1, 0, 0, 2, 4, nul, nul, nul, nul, nul) == capi.EUNSUPPORTED
    assert b">= 3" in L.tridiag_last_error()
    # the no-allocation solve refuses the vendor comparator before touching a GPU
    assert L.tridiag_solve_noalloc(0, 0, 0, 16, 4, nul, nul, nul, nul, nul) == capi.EUNSUPPORTED
    assert L.tridiag_solve_noalloc(7, 9, 0, 16, 4, nul, nul, nul, nul, nul) == capi.EINVAL
    assert L.tridiag_solve_noalloc(0, 99, 0, 16, 4, nul, nul, nul, nul, nul) == capi.EINVAL
    # a ragged solve needs a work buffer planned in this process
    fake = ctypes.c_void_p(0x1000)
    assert L.tridiag_ragged_solve(0, 4, fake, fake, fake, fake, fake, nul) == capi.EINVAL
    assert b"no plan" in L.tridiag_last_error()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("layout", [0, 2])
def test_planner_covers_benchmark_shapes(capi, prec, layout):
    for n in (1, 2, 3, 31, 64, 100, 512, 1000, 1024, 4096):
        p = capi.plan(prec, 9, layout, n, 4096)
        S, NW, M = p["S"], p["NW"], p["M"]
        if p["kind"] == "cooperative":      # with a GPU: K blocks of block_rows rows
            cap = p["block_rows"] * p["K"]
        else:
            cap = (64 // S) * NW * M if layout == 2 else 64 * (NW // S) * M
        assert cap >= n
    p = capi.plan(prec, 9, 2, 4096, 4096)
    if layout == 2 and p["kind"] == "tile":
        # 32-byte row groups per workgroup in the interleaved tiles
        assert p["S"] * (4 if prec == 0 else 8) == 32


def test_register_packed_cr_rows_per_lane(capi):
    for algo, R in ((11, 4), (12, 8), (13, 16)):
        assert capi.plan(0, algo, 0, 1024, 8)["M"] == R


@pytest.mark.parametrize("prec", [0, 1])
def test_ragged_buckets_hold_their_systems(capi, prec):
    shapes = {s["index"]: s for s in capi.ragged_shapes()}
    for n in (1, 2, 31, 32, 33, 100, 128, 129, 1000, 1500, 4096, 4097, 8192):
        shape, iface = capi.ragged_bucket(prec, n)
        assert shapes[shape]["cap"] >= n
        assert 0 <= iface <= 4
        assert shapes[shape]["lanes"] * shapes[shape]["S"] == 512
    with pytest.raises(capi.TridiagError):
        capi.ragged_bucket(prec, 8193)


def test_tuning_file_ragged_lines_are_honoured(tmp_path):
    """A ragged bucket line of a tuning file overrides the default shape (own process:
    the table is read once per process)."""
    import subprocess
    import sys
    table = tmp_path / "t.txt"
    table.write_text("# test table\nR 0 300 2 3\n")   # n <= 300: shape 2 (cap 512), PCR+Thomas
    code = ("import sys; sys.path.insert(0, %r); import tridiagonalsolver_amd as T; "
            "print(T.ragged_bucket(0, 200), T.ragged_bucket(0, 301), T.ragged_bucket(1, 200))"
            % str(ROOT))
    env = dict(os.environ, TRIDIAG_TUNING_FILE=str(table))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         check=True).stdout.split(") (")
    assert out[0].strip("( ") == "2, 3"
    assert out[1] != "2, 3"          # above the line's n_max: default choice


def _driver():
    path = os.path.join(ROOT, "tridiagonalsolver_amd", "lib", "tridiag_test")
    if not os.path.exists(path):
        pytest.skip("driver not built")
    return path


def test_driver_lists_reference_algorithms():
    import subprocess
    r = subprocess.run([_driver()], capture_output=True, text=True)
    assert r.returncode == 1                       # usage, like the reference program
    assert "single-precision algorithms:" in r.stdout
    assert " 9: reg WM-pGE" in r.stdout and "17: BPLG PCR" in r.stdout


@pytest.mark.parametrize("n", [2, 7, 40, 200])
def test_driver_condition_estimates(n):
    """Hager-Higham estimates of the driver vs the exact 1- and inf-norm condition
    numbers of the same generated system (dense numpy inverse)."""
    import subprocess
    import numpy as np
    out = subprocess.run([_driver(), "cond", str(n)], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    rows = np.array([[float(v) for v in ln.split()] for ln in out[:n]])
    c1, ci = (float(v) for v in out[n].split())
    a, b, c = rows[:, 0], rows[:, 1], rows[:, 2]
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    e1, ei = np.linalg.cond(A, 1), np.linalg.cond(A, np.inf)
    # the estimator is a lower bound, exact or within a small factor on these matrices
    assert e1 / 3 <= c1 <= e1 * (1 + 1e-12)
    assert ei / 3 <= ci <= ei * (1 + 1e-12)
    # diagonally dominant by construction: |b_i| exceeds the off-diagonal row sums
    assert np.all(np.abs(b) >= np.abs(a) + np.abs(c))


def test_convert_layout_validation_without_gpu(capi):
    """Argument checks of tridiag_convert_layout that fail before any device call."""
    L = capi.lib()
    nul = None
    p = [ctypes.c_void_p(0x100000 * (k + 1)) for k in range(8)]   # never dereferenced

    def conv(prec, sl, dl, n, batch, src, dst):
        return L.tridiag_convert_layout(prec, sl, dl, n, batch, *src, *dst, nul)

    assert conv(2, 0, 2, 8, 8, p[:4], p[4:]) == capi.EINVAL               # precision
    assert conv(0, 3, 2, 8, 8, p[:4], p[4:]) == capi.EINVAL               # layout
    assert conv(0, 0, -1, 8, 8, p[:4], p[4:]) == capi.EINVAL
    assert conv(0, 0, 2, 0, 8, p[:4], p[4:]) == capi.EINVAL               # n_eqt
    assert conv(0, 0, 2, 8, -1, p[:4], p[4:]) == capi.EINVAL              # n_batch
    assert conv(0, 0, 2, 1 << 40, 1 << 40, p[:4], p[4:]) == capi.EINVAL   # overflow
    # a coefficient given on one side only
    assert conv(0, 0, 2, 8, 8, p[:4], [p[4], p[5], p[6], nul]) == capi.EINVAL
    assert b"coefficient 3" in L.tridiag_last_error()
    assert conv(0, 0, 2, 8, 8, [nul] * 4, [nul] * 4) == capi.EINVAL       # nothing to do
    # packed side without its buffer
    assert conv(0, 1, 0, 8, 8, [nul] * 4, p[4:]) == capi.EINVAL
    assert conv(0, 0, 1, 8, 8, p[:4], [nul] * 4) == capi.EINVAL
    # overlap: same array, and a destination inside the packed source buffer
    assert conv(0, 0, 2, 8, 8, p[:4], p[:4]) == capi.EINVAL
    assert b"overlap" in L.tridiag_last_error()
    inside = ctypes.c_void_p(p[0].value + 4 * 8)
    assert conv(0, 1, 0, 8, 8, [p[0], nul, nul, nul], [nul, nul, nul, inside]) == capi.EINVAL
    # an empty batch after valid arguments is a no-op
    assert conv(1, 0, 2, 8, 0, p[:4], p[4:]) == capi.OK


@pytest.mark.parametrize("mode", ["refined", "ieee", "fast"])
def test_division_variants_load(capi, mode):
    # each division-policy library exports the header's symbols and reports its policy
    from tridiagonalsolver_amd import build
    path = os.path.join(os.path.dirname(capi.LIB_PATH), build.lib_name(mode))
    if not os.path.exists(path):
        build.build(div_mode=mode)
    L = ctypes.CDLL(path)
    for nm in _declared_symbols():
        assert hasattr(L, nm), (mode, nm)
    assert L.tridiag_div_mode() == capi.DIV_MODES[mode]
    assert capi.DIV_MODES[mode] == {"refined": 0, "ieee": 1, "fast": 2}[mode]


def test_division_mode_env_names_the_library():
    import subprocess
    import sys
    code = ("from tridiagonalsolver_amd import _capi; import os; "
            "print(os.path.basename(_capi.LIB_PATH))")
    for mode, name in (("ieee", "libtridiag_amd_div_ieee.so"),
                       ("fast", "libtridiag_amd_div_fast.so")):
        out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True,
                             text=True, env=dict(os.environ, TRIDIAG_DIV_MODE=mode), check=True)
        assert out.stdout.strip() == name
    bad = subprocess.run([sys.executable, "-c",
                          "from tridiagonalsolver_amd import _capi; _capi.lib()"],
                         cwd=ROOT, capture_output=True, text=True,
                         env=dict(os.environ, TRIDIAG_DIV_MODE="approx"))
    assert bad.returncode != 0 and "TRIDIAG_DIV_MODE" in bad.stderr
