"""numpy front end for cpu_ref.cpp (TEST INFRASTRUCTURE ONLY).

The reference's CPU solvers (Tridiagonal_AVX512: Thomas, PCR-pThomas, WM-pGE;
OpenMP over systems, 64-byte SIMD vectors inside a system), restated in C++ and
used as the measured CPU baseline and pinned against the FP64 oracle by
tests/test_cpu_ref.py.  See cpu_ref.cpp's header for the reference file:line
each algorithm follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "cpu_ref.cpp")
_SO = {"avx512": os.path.join(_HERE, "libcpuref_avx512.so"),
       "generic": os.path.join(_HERE, "libcpuref_generic.so")}
_FLAGS = {"avx512": ["-mavx512f", "-mavx512dq", "-mavx512vl", "-mavx512bw"],
          "generic": []}

# Reference CPU solver table ids (Tridiagonal_AVX512/helpers.h:34-55).
THOMAS, PCR_PTHOMAS, WM_PGE = 1, 3, 5
NAMES = {THOMAS: "Thomas", PCR_PTHOMAS: "PCR-pThomas", WM_PGE: "WM-pGE"}
LANES = {np.float32: 16, np.float64: 8}


def build(force=True):
    """Compiles both variants (g++ -O3, OpenMP)."""
    for kind, so in _SO.items():
        if not force and os.path.exists(so) and os.path.getmtime(so) >= os.path.getmtime(_SRC):
            continue
        tmp = so + ".tmp%d" % os.getpid()
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wno-psabi"] + \
            _FLAGS[kind] + [_SRC, "-o", tmp]
        subprocess.check_call(cmd)
        os.replace(tmp, so)


def cpu_info():
    """Model name, logical CPUs usable by this process and the AVX-512F flag of the host."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "usable_cpus": usable, "avx512f": "avx512f" in flags}


_lib = None
_kind = None


def lib():
    """The AVX-512 build on a host with AVX-512F, else the generic one."""
    global _lib, _kind
    if _lib is None:
        kind = "avx512" if cpu_info()["avx512f"] else "generic"
        if not os.path.exists(_SO[kind]):
            build()
        L = ctypes.CDLL(_SO[kind])
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.cpuref_solve.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, ci]
        L.cpuref_solve.restype = ci
        L.cpuref_bench.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, ci, ci, ci,
                                   ctypes.POINTER(ctypes.c_double)]
        L.cpuref_bench.restype = ctypes.c_double
        L.cpuref_avx512.argtypes = []
        L.cpuref_avx512.restype = ci
        _lib, _kind = L, kind
    return _lib


def simd():
    """'avx512' or 'generic': the build in use."""
    lib()
    return _kind


def _prec(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return 0
    if dt == np.float64:
        return 1
    raise TypeError("float32 or float64 only")


def _ptr(x):
    return ctypes.c_void_p(x.ctypes.data)


def solve(a, b, c, d, n, batch, algo, threads=0):
    """Solves a batch in the reference layout (system s at s*n) and returns x; the
    inputs are not modified (c and d are copied first)."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    c, x = np.array(c, copy=True), np.array(d, copy=True)
    for v in (b, c, x):
        if v.dtype != a.dtype:
            raise TypeError("a, b, c, d must share one dtype")
    rc = lib().cpuref_solve(_prec(a.dtype), algo, n, batch, _ptr(a), _ptr(b), _ptr(c), _ptr(x),
                            threads)
    if rc == 1:
        raise ValueError("%s takes n divisible by %d" %
                         (NAMES.get(algo, algo), (2 if algo == WM_PGE else 1) * LANES[a.dtype.type]))
    if rc:
        raise ValueError("unknown CPU algorithm %r" % (algo,))
    return x


def bench(a, b, c, d, n, batch, algo, nsets, rounds, threads=0):
    """Times rounds x nsets solves of copies of (a, b, c, d), each starting from
    memory that the other copies have since pushed out of the host caches.
    Returns (seconds per solve, fastest solve, x of the last solve)."""
    a, b, c, d = (np.ascontiguousarray(v) for v in (a, b, c, d))
    x = np.empty_like(d)
    tmin = ctypes.c_double()
    t = lib().cpuref_bench(_prec(a.dtype), algo, n, batch, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                           _ptr(x), nsets, rounds, threads, ctypes.byref(tmin))
    if t < 0:
        raise RuntimeError("cpuref_bench failed (%g)" % t)
    return t, tmin.value, x
