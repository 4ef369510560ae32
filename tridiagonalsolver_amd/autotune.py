"""Autotuner: times every compiled on-chip kernel configuration for a
(precision, layout, n, batch) workload on the current GPU and keeps the fastest
one whose answer matches the FP64 reference solution.

    python -m tridiagonalsolver_amd.autotune --prec 0 1 --layout 2 --n 4096 --batch 4096 \
        --out results.json [--table tuning/tuning_gfx950.txt]

    python -m tridiagonalsolver_amd.autotune --ragged --prec 0 1 --table tuning/tuning_gfx950.txt

    python -m tridiagonalsolver_amd.autotune --prec 0 1 --layout 0 --iface 1 2 3 4 \
        --n 1024 4096 --unknowns 16777216 --table tuning/tuning_gfx950.txt

The table file holds one line per bucket,
    precision layout family n_max M NW S vec lds_alpha nt xp
(family 0 = the default on-chip solver; nt = nontemporal loads/stores;
xp = whole-line loads handed to lanes through LDS)
and is read by libtridiag_amd.so at first use (TRIDIAG_TUNING_FILE overrides its
location): for a solve of n rows the first bucket with n <= n_max whose kernel
exists is used.  family is the interface solver the row was tuned with (--iface:
0 Wang-Mou, 1 Wang-Mou half, 2 PCR, 3 PCR + Thomas, 4 CR); an algorithm looks at
the rows of its own interface solver first, then at family 0.  Tuning at sizes between powers of two (for example 768, 1536,
3072, 6144) gives those sizes buckets of their own, tuned on padded kernels.
Kernels are timed as back-to-back launches on rotated input sets, the way a
caller streaming solves sees them.  --ragged writes lines "R precision n_max shape iface": the shape
of the ragged persistent kernel and the interface solver for systems of up to
n_max rows (above the previous line's n_max).  iface is an interface-solver code:
the family in bits 0-3 and, for PCR + Thomas (family 3), the PCR depth in bits 4-7
(the PCR steps before Thomas runs along the remaining subsystems; 0 = the
default).  The ragged sweep times the PCR + Thomas family at every depth the
shape allows (--pcr-depths limits the list), so a bucket's line also carries its
PCR depth.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _capi

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
DEFAULT_TABLE = os.path.join(TUNING_DIR, "tuning_gfx950.txt")


def _log(*a):
    print("[autotune]", *a, file=sys.stderr, flush=True)


def _inputs(torch, prec, layout, n, batch, seed, nsets):
    """Seeded diagonally dominant systems (a, c ~ U[-1,1], b = |a|+|c|+U[1,2])."""
    dtype = torch.float32 if prec == 0 else torch.float64
    g = torch.Generator(device="cuda").manual_seed(seed)
    sets = []
    shape = (batch, n) if layout == 0 else (n, batch)
    for _ in range(nsets):
        a = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        c = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        b = a.abs() + c.abs() + 1 + torch.rand(shape, generator=g, device="cuda",
                                               dtype=torch.float64)
        d = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        sets.append([t.to(dtype).contiguous() for t in (a, b, c, d)])
    return sets


def _reference(sets, prec, layout, n, batch):
    """FP64 answer of the first set from the default (validated) solver path."""
    import torch
    a, b, c, d = (t.to(torch.float64) for t in sets[0])
    x = d.clone()
    _capi.check(_capi.lib().tridiag_solve(1, 9, layout, n, batch, ctypes.c_void_p(a.data_ptr()),
                                          ctypes.c_void_p(b.data_ptr()),
                                          ctypes.c_void_p(c.data_ptr()),
                                          ctypes.c_void_p(x.data_ptr()), None))
    return x


def sweep(prec, layout, n, batch, reps=20, warmup=3, nsets=3, seed=1, iface=0):
    import torch
    L = _capi.lib()
    tiles = [t for t in _capi.tiles(prec) if t["interleaved"] == (layout == 2)
             and t["cap"] >= n and not t["wide_addr"] and t["padded"] == (t["cap"] != n)]
    if not tiles:
        _log("no on-chip kernel for prec %d layout %d n %d" % (prec, layout, n))
        return []
    sets = _inputs(torch, prec, layout, n, batch, seed, nsets)
    xref = _reference(sets, prec, layout, n, batch)
    pristine = [s[3].clone() for s in sets]
    w = 4 if prec == 0 else 8
    out = []
    for t in tiles:
        if t["vec"] and layout == 2:
            continue
        stream = torch.cuda.current_stream()
        ptrs = [[ctypes.c_void_p(x.data_ptr()) for x in s] for s in sets]

        def run(k):
            return L.tridiag_solve_tile_iface(prec, t["index"], iface, layout, n, batch,
                                              *ptrs[k % nsets], ctypes.c_void_p(stream.cuda_stream))
        rc = run(0)
        if rc != 0:
            _log("skip", t, L.tridiag_last_error().decode())
            continue
        torch.cuda.synchronize()
        x = sets[0][3].to(torch.float64)
        err = float(((x - xref).abs().max() / xref.abs().max()).item())
        for k in range(warmup):
            run(k)
        # back-to-back launches on the rotated sets, as a caller streaming solves
        # (and the benchmark) sees them, after ~10 ms of untimed launches so the
        # clocks have left their idle state; three such runs, the median kept
        t_settle = time.perf_counter()
        k = 0
        while time.perf_counter() - t_settle < 0.01:
            run(k)
            k += 1
            if k % 8 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        ms = []
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for k in range(reps):
                run(k)
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1) / reps)
        ms.sort()
        med = ms[1]
        gbs = 5 * w * n * batch / (med * 1e-3) / 1e9
        rec = dict(t, ms=med, ms_min=ms[0], gbs=gbs, relerr=err)
        out.append(rec)
        _log("prec %d layout %d n %d: M=%d NW=%d S=%d vec=%d alds=%d nt=%d xp=%d  %.4f ms  "
             "%.0f GB/s  err %.2e" % (prec, layout, n, t["M"], t["NW"], t["S"], t["vec"],
                                      t["lds_alpha"], t["nontemporal"], t["transposed"], med, gbs,
                                      err))
        for s, p in zip(sets, pristine):      # in-place solves changed d; restore
            s[3].copy_(p)
    return out


def _ragged_inputs(torch, prec, sizes, seed):
    """Concatenated seeded systems of the given sizes (FP64 master copy)."""
    off = torch.zeros(len(sizes) + 1, dtype=torch.int64)
    off[1:] = torch.cumsum(torch.as_tensor(sizes, dtype=torch.int64), 0)
    rows = int(off[-1])
    g = torch.Generator(device="cuda").manual_seed(seed)
    a = torch.rand(rows, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    c = torch.rand(rows, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    b = a.abs() + c.abs() + 1 + torch.rand(rows, generator=g, device="cuda", dtype=torch.float64)
    d = torch.rand(rows, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return off, [a, b, c, d]


def iface_codes(segs, depths=None):
    """Interface-solver codes a ragged shape of `segs` lanes (segments) per system
    takes: the five families at their defaults, then PCR + Thomas at PCR depths
    1 .. log2(segs) (or the given depths, where valid)."""
    top = segs.bit_length() - 1
    ds = range(1, top + 1) if depths is None else [d for d in depths if 1 <= d <= top]
    return list(range(5)) + [3 | (d << 4) for d in ds]


def ragged_sweep(prec, kmin=5, kmax=13, rows_per_class=1 << 23, reps=10, seed=3,
                 pcr_depths=None):
    """Per size bucket: time every shape of the ragged kernel that holds the bucket's
    largest n with every interface solver (and PCR + Thomas at every PCR depth), on
    systems drawn log-uniformly inside the bucket; keep the fastest whose answer
    matches the FP64 ragged solve.  Bucket bounds are the powers of two
    2^kmin..2^kmax and every shape capacity between them, so each bucket ends at a
    capacity."""
    import torch
    from .solver import RaggedPlan
    dtype = torch.float32 if prec == 0 else torch.float64
    tol = 1e-5 if prec == 0 else 1e-12
    shapes = _capi.ragged_shapes()
    bounds = sorted({1 << k for k in range(kmin, kmax + 1)} |
                    {sh["cap"] for sh in shapes if (1 << kmin) <= sh["cap"] <= (1 << kmax)})
    out = []
    for k, hi in enumerate(bounds):
        lo = bounds[k - 1] + 1 if k > 0 else (hi >> 1) + 1
        gen = torch.Generator().manual_seed(seed + k)
        nsys = max(64, rows_per_class // int(0.7 * hi))
        u = torch.rand(nsys, generator=gen, dtype=torch.float64)
        sizes = torch.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u).round().clamp(lo, hi)
        sizes = sizes.to(torch.int64).tolist()
        off, m64 = _ragged_inputs(torch, prec, sizes, seed + 100 * k)
        ref = m64[3].clone()
        RaggedPlan(off, dtype=torch.float64).solve(m64[0], m64[1], m64[2], ref)
        work = [t.to(dtype) for t in m64]
        d0 = work[3].clone()
        res = []
        for sh in shapes:
            if sh["cap"] < hi:
                continue
            for iface in iface_codes(sh["lanes"], pcr_depths):
                try:
                    plan = RaggedPlan(off, dtype=dtype, shape=sh["index"], iface=iface)
                except _capi.TridiagError as e:
                    _log("skip", sh, iface, e)
                    continue
                work[3].copy_(d0)
                plan.solve(*work)
                torch.cuda.synchronize()
                err = float(((work[3].to(torch.float64) - ref).abs().max() / ref.abs().max()).item())
                stream = torch.cuda.current_stream()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(reps):
                    plan.solve(*work)
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                rows = int(off[-1])
                w = 4 if prec == 0 else 8
                res.append(dict(shape=sh["index"], M=sh["M"], NWS=sh["NWS"], S=sh["S"], iface=iface,
                                ms=ms, gbs=5 * w * rows / (ms * 1e-3) / 1e9, relerr=err))
                _log("ragged prec %d bucket <=%d: shape %d (M=%d NWS=%d S=%d) iface %d  %.4f ms  "
                     "%.0f GB/s  err %.2e" % (prec, hi, sh["index"], sh["M"], sh["NWS"], sh["S"],
                                              iface, ms, res[-1]["gbs"], err))
        ok = [r for r in res if r["relerr"] <= tol]
        bestr = min(ok, key=lambda r: r["ms"]) if ok else None
        out.append(dict(prec=prec, nmax=hi, systems=nsys, results=res, best=bestr))
        if bestr:
            _log("best ragged prec %d bucket <=%d -> shape %d iface %d" % (prec, hi, bestr["shape"],
                                                                          bestr["iface"]))
    return out


def merge_ragged_table(path, buckets):
    """Replaces the ragged lines ("R precision n_max shape iface") of a table file,
    keeping its other lines."""
    keep = []
    if os.path.exists(path):
        keep = [ln.rstrip("\n") for ln in open(path) if not ln.startswith("R ")]
    lines = keep + ["R %d %d %d %d" % (b["prec"], b["nmax"], b["best"]["shape"], b["best"]["iface"])
                    for b in buckets if b["best"]]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


_KEY = ("M", "NW", "S", "vec", "lds_alpha", "nontemporal", "transposed")


def read_table(path):
    """{(precision, layout, n_max, family): kernel key} of the tile lines of a table file."""
    out = {}
    if not path or not os.path.exists(path):
        return out
    for ln in open(path):
        f = ln.split()
        if not f or f[0].startswith("#") or f[0] == "R" or len(f) != 11:
            continue
        v = [int(x) for x in f]
        out[(v[0], v[1], v[3], v[2])] = tuple(v[4:])
    return out


def best(results, tol, tie=0.01, incumbent=None):
    """Fastest accurate result.  Among those within `tie` of it (run-to-run noise)
    the table's current choice for the bucket (`incumbent`, a kernel key) is kept;
    otherwise the one with the fewest waves per system, then per workgroup (a
    shorter cross-wave interface chain)."""
    ok = [r for r in results if r["relerr"] <= tol]
    if not ok:
        return None
    fastest = min(r["ms"] for r in ok)
    near = [r for r in ok if r["ms"] <= fastest * (1 + tie)]
    for r in near:
        if incumbent is not None and tuple(int(r[k]) for k in _KEY) == tuple(incumbent):
            return r
    # interleaved tiles spread every wave over all S systems
    per_sys = lambda r: r["NW"] if r.get("interleaved") else r["NW"] / r["S"]
    return min(near, key=lambda r: (per_sys(r), r["NW"], r["ms"]))


def write_table(path, rows):
    """rows: list of (prec, layout, n_max, entry) tuples; entry["family"] is the
    interface solver the entry was tuned with (0 when absent).  Ragged ("R") lines of
    an existing table are kept, and so are its tile lines of buckets (precision,
    layout, family, n_max) that `rows` does not cover, so that a sweep over some sizes
    or families retunes only their buckets."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    old = [ln.rstrip("\n") for ln in open(path)] if os.path.exists(path) else []
    ragged = [ln for ln in old if ln.startswith("R ")]
    tuned = {(r[0], r[1], r[3].get("family", 0), r[2]) for r in rows}
    for ln in old:
        f = ln.split()
        if len(f) == 11 and not f[0].startswith("#") and f[0] != "R" and \
                (int(f[0]), int(f[1]), int(f[2]), int(f[3])) not in tuned:
            v = [int(x) for x in f]
            rows = rows + [(v[0], v[1], v[3], dict(M=v[4], NW=v[5], S=v[6], vec=v[7],
                                                   lds_alpha=v[8], nontemporal=v[9],
                                                   transposed=v[10], family=v[2]))]
    with open(path, "w") as f:
        f.write("# precision layout family n_max M NW S vec lds_alpha nt xp  "
                "(written by autotune.py)\n")
        for ln in ragged:
            f.write(ln + "\n")
        for prec, layout, nmax, e in sorted(rows, key=lambda r: (r[0], r[1], r[3].get("family", 0),
                                                               r[2])):
            f.write("%d %d %d %d %d %d %d %d %d %d %d\n" % (
                prec, layout, e.get("family", 0), nmax, e["M"], e["NW"], e["S"], int(e["vec"]),
                int(e["lds_alpha"]), int(e["nontemporal"]), int(e["transposed"])))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--prec", type=int, nargs="+", default=[0, 1])
    ap.add_argument("--layout", type=int, nargs="+", default=[2])
    ap.add_argument("--n", nargs="+", default=["4096"],
                    help="sizes to tune; N:NMAX tunes at N and records the bucket as n <= NMAX "
                         "(e.g. 1792:2047 for sizes just below a power of two, on padded kernels)")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--unknowns", type=int, default=0,
                    help="if set, batch = unknowns // n for each n (constant work per point)")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--iface", type=int, nargs="+", default=[0],
                    help="interface solvers to tune with (0 Wang-Mou, the default family; "
                         "1 Wang-Mou half, 2 PCR, 3 PCR + Thomas, 4 CR): rows of family IFACE, "
                         "used by the algorithms of that interface solver")
    ap.add_argument("--out", default=None)
    ap.add_argument("--table", default=None, help="write the winners as a tuning table")
    ap.add_argument("--ragged", action="store_true",
                    help="tune the ragged kernel per size bucket instead (R lines of the table)")
    ap.add_argument("--pcr-depths", type=int, nargs="*", default=None,
                    help="ragged: PCR depths of the PCR + Thomas family to time (default: every "
                         "depth a shape allows)")
    ap.add_argument("--kmin", type=int, default=5, help="ragged: smallest bucket 2^kmin")
    ap.add_argument("--kmax", type=int, default=13, help="ragged: largest bucket 2^kmax")
    args = ap.parse_args(argv)
    if args.ragged:
        t0 = time.time()
        buckets = []
        for prec in args.prec:
            buckets += ragged_sweep(prec, kmin=args.kmin, kmax=args.kmax, reps=args.reps,
                                    pcr_depths=args.pcr_depths)
        if args.out:
            with open(args.out, "w") as f:
                json.dump(buckets, f, indent=1)
        if args.table:
            merge_ragged_table(args.table, buckets)
        _log("done in %.1f s" % (time.time() - t0))
        return
    allres, rows = [], []
    current = read_table(args.table)
    t0 = time.time()
    for prec in args.prec:
        for layout in args.layout:
            for fam in args.iface:
                for spec in args.n:
                    n, _, nmax = str(spec).partition(":")
                    n = int(n)
                    nmax = int(nmax) if nmax else n
                    # a multiple of 64 systems keeps interleaved rows line-aligned
                    batch = max(64, args.unknowns // n // 64 * 64) if args.unknowns else args.batch
                    res = sweep(prec, layout, n, batch, reps=args.reps, iface=fam)
                    allres.append(dict(prec=prec, layout=layout, family=fam, n=n, n_max=nmax,
                                       batch=batch, results=res))
                    b = best(res, 1e-5 if prec == 0 else 1e-12,
                             incumbent=current.get((prec, layout, nmax, fam)))
                    if b:
                        b = dict(b, family=fam)
                        rows.append((prec, layout, nmax, b))
                        _log("best prec %d layout %d family %d n %d -> M=%d NW=%d S=%d vec=%d "
                             "alds=%d nt=%d xp=%d %.4f ms" % (
                                 prec, layout, fam, n, b["M"], b["NW"], b["S"], b["vec"],
                                 b["lds_alpha"], b["nontemporal"], b["transposed"], b["ms"]))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(allres, f, indent=1)
    if args.table:
        write_table(args.table, rows)
    _log("done in %.1f s" % (time.time() - t0))


if __name__ == "__main__":
    main()
