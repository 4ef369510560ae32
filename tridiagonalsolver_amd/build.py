"""Builds libtridiag_amd.so (HIP, gfx950) in-tree.

    python -m tridiagonalsolver_amd.build          # or: python tridiagonalsolver_amd/build.py
    python -m tridiagonalsolver_amd.build --div-mode ieee     # one division-policy variant
    python -m tridiagonalsolver_amd.build --all-div-modes     # the library and both variants

Steps: regenerate the tile-kernel instantiation tables, compile every csrc/*.hip
to an object with hipcc in parallel, link the shared library into lib/.

Division policy (TRIDIAG_DIV_MODE, fixed per library as the reference fixes its
division macros per build): "refined" (default, libtridiag_amd.so), "ieee"
(libtridiag_amd_div_ieee.so) and "fast" (libtridiag_amd_div_fast.so).  The
variants export the same symbols.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIBNAME = "libtridiag_amd.so"
DIV_MODES = {"refined": 0, "ieee": 1, "fast": 2}


def lib_name(div_mode="refined"):
    """File name of the library built with a division policy."""
    return LIBNAME if div_mode == "refined" else "libtridiag_amd_div_%s.so" % div_mode
ARCH = os.environ.get("TRIDIAG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 VALU ops are disabled (the compiler still vectorises FP64 and
# scalar FP32 code normally).
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
            "-I" + os.path.join(os.path.dirname(PKG), "include")]
LDFLAGS = ["-shared", "-fPIC", "--offload-arch=" + ARCH, "-L/opt/rocm/lib", "-lhipsparse",
           "-Wl,-rpath,/opt/rocm/lib"]


# ------------------------------------------------------------------ tile tables
# Shapes (M rows per lane, NW waves per workgroup, S systems per workgroup).
# Flag expansion per shape:
#   PAD   both the exact-fit kernel (n == capacity) and the padded one;
#   NT    exact-fit kernels stream with nontemporal loads/stores (interleaved
#         shapes also keep a cached variant: neighbouring workgroups share lines);
#   ALDS  long segments also keep one coefficient in LDS;
#   VEC   contiguous shapes use 16-byte loads when M allows, with a scalar-load
#         variant for unaligned arrays.
IL_SHAPES = {
    "float": [(M, NW, 8) for M in (4, 8, 16, 32, 64) for NW in (1, 2, 4, 8, 16)
              if not (M == 64 and NW == 16)] + [(32, 8, 4), (16, 16, 4), (64, 4, 4)] +
             # whole 128-byte row pieces (32 systems) and 64-byte pieces (16 systems)
             [(16, 16, 32), (8, 16, 32), (16, 8, 32), (8, 8, 32), (16, 16, 16), (32, 8, 16)],
    "double": [(M, NW, 4) for M in (4, 8, 16, 32) for NW in (1, 2, 4, 8, 16)
               if not (M == 32 and NW == 16)] + [(16, 8, 2), (8, 16, 2)] +
              [(8, 16, 16), (16, 8, 16), (8, 8, 16), (16, 16, 8), (8, 16, 8)],
}
C_SHAPES = {
    "float": [(M, NW, S) for (NW, S) in ((4, 4), (4, 2), (4, 1), (8, 1), (16, 1))
              for M in (2, 4, 8, 16, 32) if not (M == 32 and NW == 16)] +
             [(32, 16, 1), (64, 8, 1)] +
             # 12 rows per lane: capacities of 1.5x a power of two (768 ... 12288)
             [(12, 4, 4), (12, 4, 2), (12, 4, 1), (12, 8, 1), (12, 16, 1)],
    "double": [(M, NW, S) for (NW, S) in ((4, 4), (4, 2), (4, 1), (8, 1), (16, 1))
               for M in (2, 4, 8, 16) if not (M == 16 and NW == 16)] +
              [(16, 16, 1), (32, 4, 1), (32, 8, 1)] +
              [(12, 4, 4), (12, 4, 2), (12, 4, 1), (12, 8, 1)],
}
ALDS_MIN_M = {"float": 32, "double": 16}
VEC_W = {"float": 4, "double": 2}
WORD = {"float": 4, "double": 8}
VEC_W_BYTES = {"float": 16, "double": 16}
# Interleaved shapes also built with 64-bit addressing, for arrays of 2 GiB or
# more (contiguous batches are split into chunks below 2 GiB instead).
BIG_IL = {(8, 8), (16, 8)}
N_TABLE_FILES = 6


def _entries(T):
    """(M, NW, S, IL, VEC, ALDS, BIG, PAD, NT, XP) tuples of every compiled kernel."""
    out = []
    for (M, NW, S) in IL_SHAPES[T]:
        heavy = M >= ALDS_MIN_M[T]
        for alds in ((False, True) if heavy else (False,)):
            if heavy and NW == 16 and not alds:
                continue                      # 1024-lane groups need the LDS variant
            out.append((M, NW, S, True, False, alds, False, False, True, False))
            out.append((M, NW, S, True, False, alds, False, False, False, False))
            out.append((M, NW, S, True, False, alds, False, True, False, False))
        if (M, NW) in BIG_IL:
            out.append((M, NW, S, True, False, False, True, True, False, False))
        if S * WORD[T] in (64, 128) and not heavy:
            # exact tiles reading whole row pieces, handed to lanes through LDS
            out.append((M, NW, S, True, False, False, False, False, True, True))
    for (M, NW, S) in C_SHAPES[T]:
        heavy = M >= ALDS_MIN_M[T]
        alds = heavy and (NW == 16 or (T == "float" and M == 64) or (T == "double" and M == 32))
        for vec in ((True, False) if M % VEC_W[T] == 0 else (False,)):
            out.append((M, NW, S, False, vec, alds, False, False, True, False))
            out.append((M, NW, S, False, vec, alds, False, False, False, False))
            out.append((M, NW, S, False, vec, alds, False, True, False, False))
            if vec and not alds and 64 * M * VEC_W_BYTES[T] // VEC_W[T] <= 8192:
                # whole-line loads through LDS, exact fit and padded (chunks bounded by n)
                out.append((M, NW, S, False, vec, alds, False, False, True, True))
                out.append((M, NW, S, False, vec, alds, False, True, True, True))
    return out


def _cap(M, NW, S, il):
    return (64 // S) * NW * M if il else 64 * (NW // S) * M


def _table_sources(T, suffix):
    ents = _entries(T)
    f = lambda x: "true" if x else "false"
    chunks = [ents[i::N_TABLE_FILES] for i in range(N_TABLE_FILES)]
    files = {}
    decls = []
    for k, part in enumerate(chunks):
        lines = ["// Compiled tile-kernel configurations (%s, part %d).  Generated by build.py;"
                 % (T, k), "// edit the shape lists there.", '#include "tri_host.hpp"', "",
                 "namespace tri {", "",
                 "TileTable<%s> tile_part_%s_%d() {" % (T, suffix, k),
                 "    static const TileEntry<%s> t[] = {" % T]
        for (M, NW, S, il, vec, alds, big, pad, nt, xp) in part:
            lines.append("    {%d, %d, %d, %s, %s, %s, %s, %s, %s, %s, %d, &launch_tile<%s, %d, %d, "
                         "%d, %s, %s, %s, %s, %s, %s, %s>}," % (
                             M, NW, S, f(il), f(vec), f(alds), f(big), f(pad), f(nt), f(xp),
                             _cap(M, NW, S, il), T, M, NW, S, f(il), f(vec), f(alds), f(big),
                             f(pad), f(nt), f(xp)))
        lines += ["    };", "    return {t, (int)(sizeof(t) / sizeof(t[0]))};", "}", "",
                  "}  // namespace tri", ""]
        files["tri_tiles_%s_%d.hip" % (suffix, k)] = "\n".join(lines)
        decls.append((k, len(part)))
    # table assembly + Thomas launcher
    lines = ["// Table of all compiled tile kernels (%s).  Generated by build.py." % T,
             '#include "tri_host.hpp"', "", "#include <vector>", "", "namespace tri {", ""]
    for k, _ in decls:
        lines.append("TileTable<%s> tile_part_%s_%d();" % (T, suffix, k))
    lines += ["", "TileTable<%s> tile_table_%s() {" % (T, suffix),
              "    static std::vector<TileEntry<%s>> all = [] {" % T,
              "        std::vector<TileEntry<%s>> v;" % T]
    for k, _ in decls:
        lines.append("        { TileTable<%s> p = tile_part_%s_%d(); v.insert(v.end(), p.entries, "
                     "p.entries + p.count); }" % (T, suffix, k))
    lines += ["        return v;", "    }();",
              "    return {all.data(), (int)all.size()};", "}", "",
              "hipError_t launch_thomas_%s(const %s *a, const %s *b, %s *c, %s *d, int64_t n, "
              "int64_t batch," % (suffix, T, T, T, T),
              "                             int64_t sstride, int64_t rstride, hipStream_t st) {",
              "    if (batch <= 0) return hipSuccess;",
              "    const int64_t blocks = (batch + 63) / 64;",
              "    hipLaunchKernelGGL((thomas_kernel<%s>), dim3((unsigned)blocks), dim3(64), 0, st, "
              "a, b, c, d, n," % T,
              "                       batch, sstride, rstride);",
              "    return hipGetLastError();", "}", "", "}  // namespace tri", ""]
    files["tri_tiles_%s.hip" % suffix] = "\n".join(lines)
    return files


# Adjoint / fused forward+adjoint kernels: (M, NW, S) per layout.
# Contiguous shapes keep 512-lane workgroups (8 waves: S systems of 8/S waves),
# plus 1024 lanes for one system of up to 8192 rows.
ADJ_C = [(M, 8, S) for M in (4, 8) for S in (8, 4, 2, 1)] + [(M, 16, 1) for M in (4, 8)]
ADJ_IL = {"float": [(M, NW, 8) for M in (4, 8) for NW in (1, 2, 4, 8, 16)],
          "double": [(M, NW, 4) for M in (4, 8) for NW in (1, 2, 4, 8, 16)]}


def _adjoint_source():
    lines = ["// Adjoint and fused forward+adjoint kernel tables.  Generated by build.py.",
             '#include "tri_adjoint.hpp"', "", "namespace tri {", ""]
    for T, suffix in (("float", "f32"), ("double", "f64")):
        ents = []
        for (M, NW, S) in ADJ_C:
            if T == "double" and NW == 16:
                continue                         # over the 128-VGPR budget of 1024 lanes
            ents.append((M, NW, S, False))
        for (M, NW, S) in ADJ_IL[T]:
            if T == "double" and NW == 16:
                continue
            ents.append((M, NW, S, True))
        lines.append("AdjTable<%s> adj_table_%s() {" % (T, suffix))
        lines.append("    static const AdjEntry<%s> t[] = {" % T)
        for (M, NW, S, il) in ents:
            cap = _cap(M, NW, S, il)
            for mode in (0, 1):
                lines.append("        {%d, %d, %d, %s, %d, %d, &launch_adj<%s, %d, %d, %d, %s, %d>},"
                             % (M, NW, S, "true" if il else "false", mode, cap, T, M, NW, S,
                                "true" if il else "false", mode))
        lines += ["    };", "    return {t, (int)(sizeof(t) / sizeof(t[0]))};", "}", ""]
    lines += ["}  // namespace tri", ""]
    return "\n".join(lines)


def _write_if_changed(path, text):
    if os.path.exists(path) and open(path).read() == text:
        return
    with open(path, "w") as f:
        f.write(text)


def generate():
    keep = set()
    for T, suffix in (("float", "f32"), ("double", "f64")):
        for name, text in _table_sources(T, suffix).items():
            _write_if_changed(os.path.join(CSRC, name), text)
            keep.add(name)
    _write_if_changed(os.path.join(CSRC, "tri_adjoint_tables.hip"), _adjoint_source())
    keep.add("tri_adjoint_tables.hip")
    for name in os.listdir(CSRC):      # drop stale generated parts
        if name.startswith("tri_tiles_") and name not in keep:
            os.remove(os.path.join(CSRC, name))


# ------------------------------------------------------------------ compile
def _compile(src, build_dir, defines=()):
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIPFLAGS + list(defines) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    return obj, r.stderr


def _headers_mtime():
    inc = os.path.join(os.path.dirname(PKG), "include")
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    hs += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return max(os.path.getmtime(h) for h in hs) if hs else 0.0


def build(verbose=False, jobs=None, incremental=False, div_mode="refined"):
    """Compiles every csrc/*.hip and links the library of one division policy
    (the default one also gets the command-line driver).  incremental=True keeps
    objects newer than their source and every header (a development shortcut; the
    default rebuilds everything)."""
    if div_mode not in DIV_MODES:
        raise ValueError("div_mode must be one of %s" % sorted(DIV_MODES))
    generate()
    build_dir = BUILD if div_mode == "refined" else BUILD + "_div_" + div_mode
    defines = ["-DTRIDIAG_DIV_MODE=%d" % DIV_MODES[div_mode]]
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    hdr_t = _headers_mtime() if incremental else 0.0
    todo, objs = [], []
    for src in srcs:
        obj = os.path.join(build_dir, os.path.basename(src) + ".o")
        if incremental and os.path.exists(obj) and \
                os.path.getmtime(obj) > max(os.path.getmtime(src), hdr_t):
            objs.append(obj)
        else:
            todo.append(src)
    jobs = jobs or min(max(1, len(todo)), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, log in ex.map(lambda src: _compile(src, build_dir, defines), todo):
            objs.append(obj)
            if verbose and log.strip():
                print(log, file=sys.stderr)
    out = os.path.join(LIBDIR, lib_name(div_mode))
    tmp = out + ".tmp%d" % os.getpid()
    r = subprocess.run([HIPCC] + objs + LDFLAGS + ["-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, out)
    if div_mode != "refined":
        return out
    # the command-line driver (reference test program interface) against the library
    drv = os.path.join(LIBDIR, "tridiag_test")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(os.path.dirname(PKG), "include"),
                        os.path.join(PKG, "driver", "tridiag_test.cpp"), "-L" + LIBDIR,
                        "-ltridiag_amd", "-Wl,-rpath,$ORIGIN", "-o", drv + ".tmp"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(drv + ".tmp", drv)
    return out


def build_all(verbose=False, jobs=None, incremental=False):
    """The library and its two division-policy variants."""
    return [build(verbose, jobs, incremental, m) for m in ("refined", "ieee", "fast")]


if __name__ == "__main__":
    inc = "--incremental" in sys.argv
    if "--all-div-modes" in sys.argv:
        print("\n".join(build_all(verbose="-v" in sys.argv, incremental=inc)))
    else:
        mode = sys.argv[sys.argv.index("--div-mode") + 1] if "--div-mode" in sys.argv else "refined"
        print(build(verbose="-v" in sys.argv, incremental=inc, div_mode=mode))
