"""numpy front end for oracle.c (TEST INFRASTRUCTURE ONLY).

The oracle is a CPU restatement of the reference's Thomas recurrence
(Tridiagonal_AVX512/tridiagonal_basic.cpp:6-26) in FP64, plus error statistics,
residuals, the adjoint and seeded input generators.  See oracle.c's header for
the reference file:line each function follows and what pins it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_SRC = os.path.join(_HERE, "oracle.c")

LAYOUT_CONTIGUOUS = 0
LAYOUT_PACKED = 1
LAYOUT_INTERLEAVED = 2


def build(force=True):
    """Compile oracle.c into liboracle.so (gcc, OpenMP, no FP contraction)."""
    if not force and os.path.exists(_SO) and os.path.getmtime(_SO) >= os.path.getmtime(_SRC):
        return _SO
    tmp = _SO + ".tmp%d" % os.getpid()
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", _SRC, "-o", tmp, "-lm"]
    subprocess.check_call(cmd)
    os.replace(tmp, _SO)
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        i64 = ctypes.c_int64
        L.oracle_thomas_f64.argtypes = [dp, dp, dp, dp, dp, i64, i64, ctypes.c_int, ctypes.c_int]
        L.oracle_thomas_f32.argtypes = [fp, fp, fp, fp, fp, i64, i64, ctypes.c_int, ctypes.c_int]
        L.oracle_thomas_ragged_f64.argtypes = [dp, dp, dp, dp, dp, ctypes.POINTER(i64), i64,
                                               ctypes.c_int]
        L.oracle_thomas_ragged_f32.argtypes = [fp, fp, fp, fp, fp, ctypes.POINTER(i64), i64,
                                               ctypes.c_int]
        L.oracle_adjoint_f64.argtypes = [dp, dp, dp, dp, dp, dp, dp, dp, dp, i64, i64,
                                         ctypes.c_int, ctypes.c_int]
        L.oracle_residual_f64.argtypes = [dp, dp, dp, dp, dp, i64, i64, ctypes.c_int]
        L.oracle_residual_f64.restype = ctypes.c_double
        L.oracle_errstats_f64.argtypes = [dp, dp, i64, i64, ctypes.c_int, dp]
        L.oracle_errstats_f64.restype = i64
        L.oracle_sm64_stream_seed.argtypes = [ctypes.c_uint64, i64]
        L.oracle_sm64_stream_seed.restype = ctypes.c_uint64
        L.oracle_sm64_fill.argtypes = [ctypes.c_uint64, dp, i64, ctypes.c_double, ctypes.c_double]
        L.oracle_gen_dd.argtypes = [dp, dp, dp, dp, i64, i64, ctypes.c_int, ctypes.c_uint64,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oracle_gen_dd_ragged.argtypes = [dp, dp, dp, dp, ctypes.POINTER(i64), i64,
                                           ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int]
        L.oracle_gen_manufactured.argtypes = [dp, dp, dp, dp, dp, i64, ctypes.c_uint64,
                                              ctypes.c_int]
        _lib = L
    return _lib


def _p(arr, ct=ctypes.c_double):
    return arr.ctypes.data_as(ctypes.POINTER(ct))


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def flat_size(n, batch):
    return int(n) * int(batch)


# ---------------------------------------------------------------- solvers

def thomas(a, b, c, d, n, batch, layout=LAYOUT_CONTIGUOUS, nthreads=0):
    """FP64 Thomas oracle; returns x in the same layout as d (flat array).

    For layout 1 (packed) pass the one buffer as `a` and None for b, c, d;
    the solution comes back in the full packed buffer (x occupies the d slot).
    """
    if layout == LAYOUT_PACKED:
        buf = _f64(a)
        x = buf.copy()
        rc = lib().oracle_thomas_f64(_p(buf), _p(buf), _p(buf), _p(buf), _p(x), n, batch, 1,
                                     nthreads)
    else:
        a, b, c, d = (_f64(v) for v in (a, b, c, d))
        x = np.empty_like(d)
        rc = lib().oracle_thomas_f64(_p(a), _p(b), _p(c), _p(d), _p(x), n, batch, layout,
                                     nthreads)
    if rc != 0:
        raise RuntimeError("oracle_thomas_f64 failed (%d)" % rc)
    return x


def thomas_f32(a, b, c, d, n, batch, layout=LAYOUT_CONTIGUOUS, nthreads=0):
    """The reference Thomas recurrence in FP32 arithmetic (CPU baseline port)."""
    a, b, c, d = (np.ascontiguousarray(v, dtype=np.float32) for v in (a, b, c, d))
    x = np.empty_like(d)
    fp = ctypes.c_float
    rc = lib().oracle_thomas_f32(_p(a, fp), _p(b, fp), _p(c, fp), _p(d, fp), _p(x, fp), n, batch,
                                 layout, nthreads)
    if rc != 0:
        raise RuntimeError("oracle_thomas_f32 failed (%d)" % rc)
    return x


def thomas_ragged(a, b, c, d, offsets, nthreads=0):
    a, b, c, d = (_f64(v) for v in (a, b, c, d))
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    x = np.empty_like(d)
    rc = lib().oracle_thomas_ragged_f64(_p(a), _p(b), _p(c), _p(d), _p(x),
                                        off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        len(off) - 1, nthreads)
    if rc != 0:
        raise RuntimeError("oracle_thomas_ragged_f64 failed (%d)" % rc)
    return x


def thomas_ragged_f32(a, b, c, d, offsets, nthreads=0):
    a, b, c, d = (np.ascontiguousarray(v, dtype=np.float32) for v in (a, b, c, d))
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    x = np.empty_like(d)
    fp = ctypes.c_float
    rc = lib().oracle_thomas_ragged_f32(_p(a, fp), _p(b, fp), _p(c, fp), _p(d, fp), _p(x, fp),
                                        off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        len(off) - 1, nthreads)
    if rc != 0:
        raise RuntimeError("oracle_thomas_ragged_f32 failed (%d)" % rc)
    return x


def adjoint(a, b, c, x, g, n, batch, layout=LAYOUT_CONTIGUOUS, nthreads=0):
    """FP64 analytic adjoint: returns (ga, gb, gc, gd) for L with dL/dx = g."""
    a, b, c, x, g = (_f64(v) for v in (a, b, c, x, g))
    outs = [np.empty_like(x) for _ in range(4)]
    rc = lib().oracle_adjoint_f64(_p(a), _p(b), _p(c), _p(x), _p(g), *(_p(o) for o in outs),
                                  n, batch, layout, nthreads)
    if rc != 0:
        raise RuntimeError("oracle_adjoint_f64 failed (%d)" % rc)
    return tuple(outs)


# ---------------------------------------------------------------- checks

def residual(a, b, c, d, x, n, batch, layout=LAYOUT_CONTIGUOUS):
    """max over systems of ||Ax-d||_inf / (||A||_inf ||x||_inf + ||d||_inf), in FP64."""
    a, b, c, d, x = (_f64(v) for v in (a, b, c, d, x))
    return float(lib().oracle_residual_f64(_p(a), _p(b), _p(c), _p(d), _p(x), n, batch, layout))


def errstats(x, xref, n, batch, layout=LAYOUT_CONTIGUOUS):
    """Returns dict(mean, rmse, max, relerr, nonfinite, argmax) over all systems."""
    x, xref = _f64(x), _f64(xref)
    out = np.zeros(5, dtype=np.float64)
    arg = lib().oracle_errstats_f64(_p(x), _p(xref), n, batch, layout, _p(out))
    if arg < 0:
        raise MemoryError("oracle_errstats_f64")
    return dict(mean=out[0], rmse=out[1], max=out[2], relerr=out[3], nonfinite=int(out[4]),
                argmax=int(arg))


# ---------------------------------------------------------------- generators

def sm64_fill(state, count, lo=0.0, hi=1.0):
    out = np.empty(count, dtype=np.float64)
    lib().oracle_sm64_fill(state, _p(out), count, lo, hi)
    return out


def stream_seed(seed, s):
    return int(lib().oracle_sm64_stream_seed(seed, s))


def gen_dd(n, batch, layout=LAYOUT_CONTIGUOUS, seed=1, blo=1.0, bhi=2.0, nthreads=0):
    """Diagonally dominant systems: a,c ~ U[-1,1] (a_0=c_{n-1}=0), b=|a|+|c|+U[blo,bhi],
    d ~ U[-1,1].  Returns (a, b, c, d) flat FP64 arrays in the requested layout
    (layout 1 returns the single packed buffer 4 times)."""
    if layout == LAYOUT_PACKED:
        buf = np.empty(4 * n * batch, dtype=np.float64)
        lib().oracle_gen_dd(_p(buf), _p(buf), _p(buf), _p(buf), n, batch, 1, seed, blo, bhi,
                            nthreads)
        return buf, buf, buf, buf
    arrs = [np.empty(n * batch, dtype=np.float64) for _ in range(4)]
    lib().oracle_gen_dd(*(_p(v) for v in arrs), n, batch, layout, seed, blo, bhi, nthreads)
    return tuple(arrs)


def gen_dd_ragged(offsets, seed=1, blo=1.0, bhi=2.0, nthreads=0):
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    total = int(off[-1])
    arrs = [np.empty(total, dtype=np.float64) for _ in range(4)]
    lib().oracle_gen_dd_ragged(*(_p(v) for v in arrs),
                               off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(off) - 1,
                               seed, blo, bhi, nthreads)
    return tuple(arrs)


def gen_manufactured(n, seed=13541532, single=False):
    """One system built with the recipe of Tridiagonal_CUDA/test.cpp:19-38.
    Returns (a, b, c, d, x) FP64 arrays of length n (values exactly representable
    in float when single=True)."""
    arrs = [np.zeros(n, dtype=np.float64) for _ in range(5)]
    lib().oracle_gen_manufactured(*(_p(v) for v in arrs), n, seed, 1 if single else 0)
    return tuple(arrs)


def ragged_sizes(nsys, nmin, nmax, seed):
    """Log-uniform system sizes n = round(exp(U[ln nmin, ln nmax])), seeded."""
    u = sm64_fill(stream_seed(seed, 0x5EED), nsys, 0.0, 1.0)
    n = np.rint(np.exp(np.log(nmin) + (np.log(nmax) - np.log(nmin)) * u)).astype(np.int64)
    return np.clip(n, nmin, nmax)


def offsets_from_sizes(sizes):
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    return off
