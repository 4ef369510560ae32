"""ctypes binding of libtridiag_amd.so (the C-ABI in include/tridiagonal.h).

The HIP library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))

# Division policies (include/tridiagonal.h, TRIDIAG_DIV_*): one library file each,
# with the same symbols.  $TRIDIAG_DIV_MODE picks the file this binding loads.
DIV_MODES = {"refined": 0, "ieee": 1, "fast": 2}
_DIV_MODE = os.environ.get("TRIDIAG_DIV_MODE", "refined")
LIB_PATH = os.path.join(_PKG, "lib", "libtridiag_amd.so" if _DIV_MODE == "refined"
                        else "libtridiag_amd_div_%s.so" % _DIV_MODE)

SINGLE, DOUBLE = 0, 1
CONTIGUOUS, PACKED, INTERLEAVED = 0, 1, 2
ALGO_AUTO = -1
OK, EINVAL, EUNSUPPORTED, EHIP = 0, 1, 2, 3

# Reference-compatible void symbols (Tridiagonal_CUDA/tridiagonal.h:8-57).
REFERENCE_SYMBOLS = [
    "cuSparse_init_single", "cuSparse_init_double", "cuSparse_final", "cuSparse_single",
    "cuSparse_double", "BPLG_WM", "BPLG_LF", "BPLG_CR", "BPLG_PCR", "Thomas_single",
    "Thomas_double", "CR_init_single", "CR_init_double", "CR_final", "CR_single", "CR_double",
    "WM5_shmem_single", "WM4_shmem_single", "PCR5_shmem_single", "PCR4_shmem_single",
    "WM5_reg_single", "WM4_reg_single", "PCR5_reg_single", "PCR4_reg_single",
    "WM5_shmem_double", "WM4_shmem_double", "PCR5_shmem_double", "PCR4_shmem_double",
    "WM5_reg_double", "WM4_reg_double", "PCR5_reg_double", "PCR4_reg_double",
    "CR_reg4_double", "CR_reg8_double", "CR_reg16_double", "CR_reg4_single", "CR_reg8_single",
    "CR_reg16_single",
]
EXTENDED_SYMBOLS = [
    "tridiag_algo_count", "tridiag_algo_name", "tridiag_algo_layout", "tridiag_solve",
    "tridiag_ragged_work_size", "tridiag_solve_ragged", "tridiag_adjoint",
    "tridiag_solve_fwd_bwd", "tridiag_last_error", "tridiag_plan", "tridiag_tile_count",
    "tridiag_tile_info", "tridiag_solve_tile", "tridiag_solve_tile_iface", "tridiag_plan_kind",
    "tridiag_ragged_shape_count",
    "tridiag_ragged_shape_info", "tridiag_ragged_bucket", "tridiag_ragged_plan",
    "tridiag_ragged_solve", "tridiag_convert_layout", "tridiag_div_mode", "tridiag_solve_bulk",
]


class TridiagError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("tridiag error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if _DIV_MODE not in DIV_MODES:
        raise ValueError("TRIDIAG_DIV_MODE must be one of %s, not %r" % (sorted(DIV_MODES), _DIV_MODE))
    if not os.path.exists(LIB_PATH):
        raise ImportError("%s not built (run python -m tridiagonalsolver_amd.build%s)" %
                          (os.path.basename(LIB_PATH),
                           "" if _DIV_MODE == "refined" else " --div-mode " + _DIV_MODE))
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    L.tridiag_div_mode.argtypes = []
    L.tridiag_div_mode.restype = ci
    L.tridiag_algo_count.argtypes = [ci]
    L.tridiag_algo_count.restype = ci
    L.tridiag_algo_name.argtypes = [ci, ci]
    L.tridiag_algo_name.restype = ctypes.c_char_p
    L.tridiag_algo_layout.argtypes = [ci, ci]
    L.tridiag_algo_layout.restype = ci
    L.tridiag_solve.argtypes = [ci, ci, ci, i64, i64, vp, vp, vp, vp, vp]
    L.tridiag_solve.restype = ci
    L.tridiag_solve_bulk.argtypes = [ci, ci, ci, i64, i64, vp, vp, vp, vp, vp]
    L.tridiag_solve_bulk.restype = ci
    L.tridiag_solve_noalloc.argtypes = [ci, ci, ci, i64, i64, vp, vp, vp, vp, vp]
    L.tridiag_solve_noalloc.restype = ci
    L.tridiag_ragged_work_size.argtypes = [i64]
    L.tridiag_ragged_work_size.restype = i64
    L.tridiag_solve_ragged.argtypes = [ci, i64, vp, vp, vp, vp, vp, vp, vp]
    L.tridiag_solve_ragged.restype = ci
    L.tridiag_adjoint.argtypes = [ci, ci, i64, i64] + [vp] * 10
    L.tridiag_adjoint.restype = ci
    L.tridiag_solve_fwd_bwd.argtypes = [ci, ci, i64, i64] + [vp] * 11
    L.tridiag_solve_fwd_bwd.restype = ci
    L.tridiag_last_error.argtypes = []
    L.tridiag_last_error.restype = ctypes.c_char_p
    L.tridiag_plan.argtypes = [ci, ci, ci, i64, i64, ctypes.POINTER(ci), ctypes.POINTER(ci),
                               ctypes.POINTER(ci)]
    L.tridiag_plan.restype = ci
    L.tridiag_plan_kind.argtypes = [ci, ci, ci, i64, i64]
    L.tridiag_plan_kind.restype = ci
    L.tridiag_tile_count.argtypes = [ci]
    L.tridiag_tile_count.restype = ci
    pi = ctypes.POINTER(ci)
    L.tridiag_tile_info.argtypes = [ci, ci, pi, pi, pi, pi, pi, pi, ctypes.POINTER(i64), pi, pi,
                                    pi, pi]
    L.tridiag_tile_info.restype = ci
    L.tridiag_solve_tile.argtypes = [ci, ci, ci, i64, i64, vp, vp, vp, vp, vp]
    L.tridiag_solve_tile.restype = ci
    L.tridiag_solve_tile_iface.argtypes = [ci, ci, ci, ci, i64, i64, vp, vp, vp, vp, vp]
    L.tridiag_solve_tile_iface.restype = ci
    L.tridiag_ragged_shape_count.argtypes = []
    L.tridiag_ragged_shape_count.restype = ci
    L.tridiag_ragged_shape_info.argtypes = [ci, pi, pi, pi, ctypes.POINTER(i64)]
    L.tridiag_ragged_shape_info.restype = ci
    L.tridiag_ragged_bucket.argtypes = [ci, i64, pi, pi]
    L.tridiag_ragged_bucket.restype = ci
    L.tridiag_ragged_plan.argtypes = [ci, i64, vp, ci, ci, vp, vp]
    L.tridiag_ragged_plan.restype = ci
    L.tridiag_ragged_solve.argtypes = [ci, i64, vp, vp, vp, vp, vp, vp]
    L.tridiag_ragged_solve.restype = ci
    L.tridiag_convert_layout.argtypes = [ci, ci, ci, i64, i64] + [vp] * 9
    L.tridiag_convert_layout.restype = ci
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise TridiagError(rc, lib().tridiag_last_error().decode())
    return rc


def algo_names(precision):
    L = lib()
    return [L.tridiag_algo_name(precision, i).decode()
            for i in range(L.tridiag_algo_count(precision))]


PLAN_KINDS = ["lane", "thomas", "tile", "cooperative", "chunked", "via_contiguous", "bulk_row"]


def plan(precision, algo, layout, n, batch):
    """How tridiag_solve solves the batch: kind (PLAN_KINDS) and, for on-chip
    workgroup plans, rows per lane M, waves per workgroup NW, systems per workgroup S
    (cooperative plans: also the rows per block and blocks per system K)."""
    L = lib()
    k = L.tridiag_plan_kind(precision, algo, layout, n, batch)
    if k < 0:
        raise TridiagError(EINVAL, L.tridiag_last_error().decode())
    out = dict(kind=PLAN_KINDS[k])
    if PLAN_KINDS[k] in ("tile", "cooperative"):
        M, NW, S = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(L.tridiag_plan(precision, algo, layout, n, batch, ctypes.byref(M),
                             ctypes.byref(NW), ctypes.byref(S)))
        out.update(M=M.value, NW=NW.value, S=S.value)
        if PLAN_KINDS[k] == "cooperative":
            rows = (64 // S.value if layout == INTERLEAVED else 64) * NW.value * M.value
            out.update(block_rows=rows, K=-(-n // rows))
    return out


def tiles(precision):
    """All compiled on-chip kernel configurations of a precision."""
    L = lib()
    out = []
    for i in range(L.tridiag_tile_count(precision)):
        v = [ctypes.c_int() for _ in range(6)]
        cap = ctypes.c_int64()
        big, pad, nt, xp = (ctypes.c_int() for _ in range(4))
        check(L.tridiag_tile_info(precision, i, *(ctypes.byref(x) for x in v), ctypes.byref(cap),
                                  ctypes.byref(big), ctypes.byref(pad), ctypes.byref(nt),
                                  ctypes.byref(xp)))
        M, NW, S, il, vec, alds = (x.value for x in v)
        out.append(dict(index=i, M=M, NW=NW, S=S, interleaved=bool(il), vec=bool(vec),
                        lds_alpha=bool(alds), cap=cap.value, wide_addr=bool(big.value),
                        padded=bool(pad.value), nontemporal=bool(nt.value),
                        transposed=bool(xp.value)))
    return out


def ragged_shapes():
    """Shapes of the ragged persistent kernel: rows per lane M, waves per system NWS
    (0: a sub-wave shape of 512/S lanes per system), systems per workgroup S, lanes
    (interface segments) per system and capacity lanes*M."""
    L = lib()
    out = []
    for i in range(L.tridiag_ragged_shape_count()):
        M, NWS, S = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        cap = ctypes.c_int64()
        check(L.tridiag_ragged_shape_info(i, ctypes.byref(M), ctypes.byref(NWS), ctypes.byref(S),
                                          ctypes.byref(cap)))
        out.append(dict(index=i, M=M.value, NWS=NWS.value, S=S.value, cap=cap.value,
                        lanes=cap.value // M.value))
    return out


def ragged_bucket(precision, n):
    """(shape index, interface solver id) the planner picks for a system of n rows."""
    sh, fi = ctypes.c_int(), ctypes.c_int()
    check(lib().tridiag_ragged_bucket(precision, n, ctypes.byref(sh), ctypes.byref(fi)))
    return sh.value, fi.value


def div_mode():
    """Name of the division policy of the loaded library ("refined", "ieee", "fast")."""
    code = lib().tridiag_div_mode()
    return {v: k for k, v in DIV_MODES.items()}[code]
