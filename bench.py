#!/usr/bin/env python3
"""Benchmark of the batched tridiagonal solve on one MI355X.

Workload (the headline shape): n = 4096 rows x batch = 4096 systems, seeded
diagonally dominant inputs (a, c ~ U[-1,1], a_0 = c_{n-1} = 0, b = |a|+|c|+U[1,2],
d ~ U[-1,1]), in the reference's contiguous per-system layout (row i of system s
at s*n + i) and in the interleaved layout (i*batch + s).  One step = one
in-place solve of the whole batch, inputs already resident in HBM.  Three input
sets are rotated so the 256 MB MALL cannot serve a step from the previous one.

Prints ONE JSON line: value = unknowns/s of the FP32 contiguous solve (whole job),
with FP64 and the interleaved layout, the HBM roofline of each solve kernel
(algorithmic bytes 5*sizeof(T) per unknown / kernel time from HIP events) and
the CPU baseline alongside.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
N_EQT, N_BATCH = 4096, 4096
N_SETS = 3
SETTLE_MS = 30.0               # untimed solves before each timed region (see _settle)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_sets(torch, dtype, n, batch, nsets, seed0, layout):
    import oracle
    sets = []
    for k in range(nsets):
        a, b, c, d = oracle.gen_dd(n, batch, layout=layout, seed=seed0 + k)
        sets.append([torch.from_numpy(v.astype(dtype)).cuda() for v in (a, b, c, d)])
        del a, b, c, d
    return sets


def _settle(torch, fn, min_ms=SETTLE_MS):
    """Untimed: repeats fn() for at least min_ms of wall time, after the warm-up
    steps, so that the timed region starts with the GPU's clocks out of their idle
    state.  Measured on MI355X: after a host-side pause (input generation), the
    first ~60 solves of FP64 4096 x 4096 averaged 0.128-0.136 ms against a steady
    0.118 ms, which ten warm-up steps (about 1 ms) did not cover."""
    t0 = time.perf_counter()
    while True:
        for _ in range(4):
            fn()
        torch.cuda.synchronize()
        if (time.perf_counter() - t0) * 1e3 >= min_ms:
            return


def time_solver(torch, T, sets, n, batch, steps, warmup, algo, layout, bulk=False):
    """Times `steps` back-to-back solves through the C-ABI on torch's current stream
    (tridiag_solve, or tridiag_solve_bulk with bulk=True).  Returns (wall ms per
    step, ms per launch from HIP events around the timed region)."""
    import ctypes
    stream = torch.cuda.current_stream()
    prec = 0 if sets[0][0].dtype == torch.float32 else 1
    L = T.lib()
    entry = L.tridiag_solve_bulk if bulk else L.tridiag_solve
    h = ctypes.c_void_p(stream.cuda_stream)
    args = [(prec, algo, layout, n, batch) + tuple(ctypes.c_void_p(t.data_ptr()) for t in s) + (h,)
            for s in sets]
    for k in range(warmup):
        T.check(entry(*args[k % len(args)]))
    turn = [0]

    def rotate():                              # settle on the rotated sets, as timed
        turn[0] += 1
        return entry(*args[turn[0] % len(args)])
    _settle(torch, rotate)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rcs = 0
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(steps):
        rcs |= entry(*args[k % len(args)])
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    T.check(rcs)
    kern = e0.elapsed_time(e1) / steps
    return wall, kern


def check_parity(torch, T, dtype, n, batch, algo, layout, tol, bulk=False):
    """Solves one fresh set and compares every system with the FP64 oracle."""
    import numpy as np
    import oracle
    a, b, c, d = oracle.gen_dd(n, batch, layout=layout, seed=99)
    arrs = [v.astype(dtype) for v in (a, b, c, d)]
    ts = [torch.from_numpy(v).cuda() for v in arrs]
    if bulk:
        T.solve_bulk(*ts, algo, layout=layout, n=n, batch=batch)
    else:
        T.solve(*ts, algo=algo, layout=layout, n=n, batch=batch)
    x = ts[3].cpu().numpy().astype(np.float64)
    xref = oracle.thomas(*(v.astype(np.float64) for v in arrs), n, batch, layout=layout)
    st = oracle.errstats(x, xref, n, batch, layout)
    res = oracle.residual(*(v.astype(np.float64) for v in arrs), x, n, batch, layout)
    return dict(relerr=float(st["relerr"]), residual=float(res), tol=tol,
                ok=bool(st["relerr"] <= tol))


def _l3_bytes():
    """Total last-level cache of the host: the sum over distinct L3 instances."""
    seen, total = set(), 0
    base = "/sys/devices/system/cpu"
    try:
        cpus = [c for c in os.listdir(base) if c.startswith("cpu") and c[3:].isdigit()]
    except OSError:
        cpus = []
    for cpu in cpus:
        cdir = os.path.join(base, cpu, "cache")
        try:
            idx = os.listdir(cdir)
        except OSError:
            continue
        for ix in idx:
            try:
                with open(os.path.join(cdir, ix, "level")) as f:
                    if f.read().strip() != "3":
                        continue
                with open(os.path.join(cdir, ix, "shared_cpu_list")) as f:
                    key = f.read().strip()
                with open(os.path.join(cdir, ix, "size")) as f:
                    sz = f.read().strip()
            except OSError:
                continue
            if key in seen:
                continue
            seen.add(key)
            mult = {"K": 1 << 10, "M": 1 << 20, "G": 1 << 30}.get(sz[-1:], 1)
            total += int(sz.rstrip("KMG")) * mult
    return total or (256 << 20)


# Points of the CPU baseline: the reference's CPU solvers (restated, oracle/cpu_ref.cpp)
# at the reference's CPU-runnable case and at the headline shape, layout 0
# (the reference CPU kernels have no interleaved layout).  (name, dtype, n, batch, algo)
CPU_POINTS = [
    ("f64_n1024_b1024_thomas", "float64", 1024, 1024, 1),
    ("f64_n1024_b1024_wm_pge", "float64", 1024, 1024, 5),
    ("f32_n4096_b4096_pcr_pthomas", "float32", 4096, 4096, 3),
    ("f32_n4096_b4096_wm_pge", "float32", 4096, 4096, 5),
    ("f64_n4096_b4096_pcr_pthomas", "float64", 4096, 4096, 3),
    ("f64_n4096_b4096_wm_pge", "float64", 4096, 4096, 5),
]
CPU_HEADLINE = "f32_n4096_b4096_wm_pge"


def cpu_leg():
    """Runs in a child process (so OpenMP reads the placement below at start-up):
    the restated reference CPU solvers over every host CPU this process may use,
    threads spread over the cores.  Each point solves copies of one seeded batch;
    the copies together exceed twice the host's L3, and each timed solve starts
    from a copy the others have since pushed out of cache.  Parity of every point's
    last solve against the FP64 oracle is reported with it."""
    import numpy as np
    import oracle
    from oracle import cpu_ref
    info = cpu_ref.cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", info["usable_cpus"]))
    l3 = _l3_bytes()
    points = {}
    t_start = time.perf_counter()
    gen_cache = {}
    for name, dtname, n, batch, algo in CPU_POINTS:
        dtype = np.dtype(dtname).type
        key = (n, batch)
        if key not in gen_cache:
            gen_cache.clear()
            gen_cache[key] = oracle.gen_dd(n, batch, layout=0, seed=7)
        a, b, c, d = (v.astype(dtype) for v in gen_cache[key])
        set_bytes = 4 * n * batch * np.dtype(dtype).itemsize
        nsets = int(min(64, 1 + -(-2 * l3 // set_bytes)))
        # one round to size the run, then enough rounds for ~1.5 s of timed solves
        t1, _, _ = cpu_ref.bench(a, b, c, d, n, batch, algo, nsets, 1, threads)
        rounds = int(max(1, min(20, 1.5 / (t1 * nsets))))
        t, tmin, x = cpu_ref.bench(a, b, c, d, n, batch, algo, nsets, rounds, threads)
        xref = oracle.thomas(*gen_cache[key], n, batch, layout=0) if dtype == np.float64 else \
            oracle.thomas(*(v.astype(np.float64) for v in (a, b, c, d)), n, batch, layout=0)
        err = float(oracle.errstats(x.astype(np.float64), xref, n, batch, 0)["relerr"])
        tol = 1e-5 if dtype == np.float32 else 1e-12
        points[name] = {"algo": algo, "name": cpu_ref.NAMES[algo], "n": n, "batch": batch,
                        "dtype": "f32" if dtype == np.float32 else "f64",
                        "unknowns_per_s": n * batch / t, "ms_per_solve": t * 1e3,
                        "fastest_ms": tmin * 1e3, "solves": nsets * rounds, "sets": nsets,
                        "relerr": err, "ok": bool(err <= tol)}
        del a, b, c, d, x, xref
    head = points[CPU_HEADLINE]
    return {
        "value": head["unknowns_per_s"], "unit": "unknowns/s", "cores": threads,
        "kind": "port",
        "algorithm": "reference CPU WM-pGE (Tridiagonal_AVX512 WM4_single), restated",
        "sample": "FP32 WM-pGE n=4096 x batch=4096 layout 0, %d solves of %d input copies "
                  "(%.0f MB each, together > 2x the %.0f MB host L3), OpenMP %d threads "
                  "spread over cores, %s vectors" % (head["solves"], head["sets"],
                                                       4 * 4 * 4096 * 4096 / 1e6, l3 / 1e6,
                                                       threads, cpu_ref.simd()),
        "cpu": {"model": info["model"], "usable_cpus": info["usable_cpus"],
                "avx512f": info["avx512f"], "simd_build": cpu_ref.simd(), "l3_bytes": l3},
        "points": points,
        "leg_seconds": time.perf_counter() - t_start,
    }


def cpu_baseline():
    """The CPU leg in a child process with OpenMP placement set before start-up."""
    import subprocess
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", str(_host_threads()))
    env["OMP_PROC_BIND"] = "spread"
    env["OMP_PLACES"] = "cores"
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-only"], env=env,
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log("cpu leg failed:", r.stderr[-2000:])
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def plan_key(plan):
    """Short name of a solve plan (tridiagonalsolver_amd.plan): kind and shape."""
    if "M" in plan:
        return "%s M%d NW%d S%d" % (plan["kind"], plan["M"], plan["NW"], plan["S"])
    return plan["kind"]


def load_traffic(key):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE) for the
    workload and kernel plan in `key` ("<workload>|<plan>"), from
    profiles/traffic.json; None when that plan was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


WORKLOADS = [
    # name, numpy dtype name, layout
    ("f32_contiguous", "float32", 0),
    ("f64_contiguous", "float64", 0),
    ("f32_interleaved", "float32", 2),
    ("f64_interleaved", "float64", 2),
]


def reference_style_ms(torch, T, sets, n, batch, reps):
    """Host clock around the synchronous reference entry point (WM5_reg_single /
    WM5_reg_double: default stream, returns after the solve), averaged over `reps`
    calls on the rotated sets, as the reference driver times a solver
    (Tridiagonal_CUDA/test.cpp:94-100).  Includes launch and synchronisation."""
    import ctypes
    prec = 0 if sets[0][0].dtype == torch.float32 else 1
    fn = getattr(T.lib(), "WM5_reg_single" if prec == 0 else "WM5_reg_double")
    fn.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int]
    fn.restype = None
    ptrs = [[ctypes.c_void_p(t.data_ptr()) for t in s] for s in sets]
    for k in range(3):
        fn(*ptrs[k % len(ptrs)], n, batch)
    torch.cuda.synchronize()
    t = 0.0
    for k in range(reps):
        t0 = time.perf_counter()
        fn(*ptrs[k % len(ptrs)], n, batch)
        t += time.perf_counter() - t0
    return t * 1e3 / reps


def run_workload(torch, T, name, dtname, layout, steps, warmup):
    import numpy as np
    dtype = np.dtype(dtname).type
    n, batch = N_EQT, N_BATCH
    log("prepare", name)
    sets = make_sets(torch, dtype, n, batch, N_SETS, seed0=11, layout=layout)
    log("time", name)
    wall, kern = time_solver(torch, T, sets, n, batch, steps, warmup, -1, layout)
    ref_style = None
    if layout == 0:
        ref_style = reference_style_ms(torch, T, sets, n, batch, steps)
    # where the allocator put b, c, d relative to a (MiB): the interleaved solve is
    # sensitive to the spacing of the four arrays
    spacing = [round((t.data_ptr() - sets[0][0].data_ptr()) / 2**20, 3) for t in sets[0][1:]]
    del sets
    torch.cuda.empty_cache()
    w = np.dtype(dtype).itemsize
    alg_bytes = 5 * w * n * batch
    prec = 0 if w == 4 else 1
    pl = T.plan(prec, -1, layout, n, batch)
    log("parity", name)
    par = check_parity(torch, T, dtype, n, batch, -1, layout, 1e-5 if w == 4 else 1e-12)
    achieved = alg_bytes / (kern * 1e-3) / 1e9
    r = dict(ms_per_step=wall, kernel_ms=kern, unknowns_per_s=n * batch / (wall * 1e-3),
             reference_style_sync_ms=ref_style, array_spacing_mib=spacing,
             achieved_gbs=achieved, alg_bytes=alg_bytes, plan=pl, parity=par,
             layout={0: "contiguous", 2: "interleaved"}[layout],
             roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                       "traffic": load_traffic("%s_4096x4096|%s" % (name, plan_key(pl)))})
    log(name, json.dumps(r))
    return r


def run_pcr(torch, T, steps, warmup, algo=5):
    """The PCR-pThomas on-chip solver (reference id 5, reg PCR-pThomas) at n=4096 x
    4096, FP32 and FP64, contiguous and interleaved, timed and checked like the
    automatic plan."""
    import numpy as np
    n, batch = N_EQT, N_BATCH
    out = {}
    for name, dtname, layout in WORKLOADS:
        dtype = np.dtype(dtname).type
        log("prepare pcr", name)
        sets = make_sets(torch, dtype, n, batch, N_SETS, seed0=11, layout=layout)
        wall, kern = time_solver(torch, T, sets, n, batch, steps, warmup, algo, layout)
        del sets
        torch.cuda.empty_cache()
        w = np.dtype(dtype).itemsize
        par = check_parity(torch, T, dtype, n, batch, algo, layout, 1e-5 if w == 4 else 1e-12)
        achieved = 5 * w * n * batch / (kern * 1e-3) / 1e9
        out[name] = dict(kernel_ms=kern, ms_per_step=wall, unknowns_per_s=n * batch / (wall * 1e-3),
                         hbm_frac=round(achieved / HBM_PEAK_GBS, 4),
                         plan=T.plan(0 if w == 4 else 1, algo, layout, n, batch), parity=par)
        log("pcr", name, json.dumps(out[name]))
    return {"algo": algo, "name": "reg PCR-pThomas", **out}


# (id, name, words of memory traffic per unknown by design)
BULK_IDS = [(2, "CR, levels in global memory", 23), (3, "shmem PCR-pThomas", 9),
            (5, "reg PCR-pThomas", 9), (7, "shmem WM-pGE", 9), (9, "reg WM-pGE", 9),
            (12, "reg-8 CR, 8 rows per lane", 5)]


def run_bulk(torch, T, steps, warmup):
    """The reference's own streaming forms of PCR-pThomas and WM-pGE
    (tridiag_solve_bulk: one wave per system, c' and d' through memory, 9 words
    per unknown), of CR (its reduced levels written and read back through
    global memory: about 23 words per unknown, 11.5 per row of each level) and
    of register CR (8 rows per lane, one workgroup per system: 5 words) at
    n=4096 x 4096 contiguous, FP32 and FP64, beside the partition kernel
    tridiag_solve runs for the same ids."""
    import numpy as np
    n, batch = N_EQT, N_BATCH
    out = {}
    for name, dtname, layout in WORKLOADS[:2]:
        dtype = np.dtype(dtname).type
        w = np.dtype(dtype).itemsize
        log("prepare bulk", name)
        sets = make_sets(torch, dtype, n, batch, N_SETS, seed0=11, layout=layout)
        res = {}
        for algo, label, words in BULK_IDS:
            wall, kern = time_solver(torch, T, sets, n, batch, steps, warmup, algo, layout,
                                     bulk=True)
            res[str(algo)] = dict(name=label, kernel_ms=kern, ms_per_step=wall,
                                  unknowns_per_s=n * batch / (wall * 1e-3),
                                  hbm_frac=round(5 * w * n * batch / (kern * 1e-3) / 1e9 /
                                                 HBM_PEAK_GBS, 4),
                                  traffic_words=words,
                                  traffic_rate_gbs=round(words * w * n * batch / (kern * 1e-3) / 1e9,
                                                         1))
        del sets
        torch.cuda.empty_cache()
        for algo, _, _ in BULK_IDS:
            res[str(algo)]["parity"] = check_parity(torch, T, dtype, n, batch, algo, layout,
                                                    1e-5 if w == 4 else 1e-12, bulk=True)
        out[name] = res
        log("bulk", name, json.dumps(res))
    return out


def _event_time(torch, fn, steps, stream):
    """Runs fn() `steps` times between two HIP events on `stream`; returns
    (wall ms per call, event ms per call)."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, e0.elapsed_time(e1) / steps


def _ragged_relerr(x, xref, off):
    import numpy as np
    starts = off[:-1][np.diff(off) > 0]
    err = np.maximum.reduceat(np.abs(x - xref), starts)
    ref = np.maximum.reduceat(np.abs(xref), starts)
    return float(np.max(err / np.where(ref > 0, ref, 1.0)))


def run_ragged(torch, T, steps, warmup, n_sys=65536, nmin=32, nmax=8192):
    """Ragged batch, FP32: n_sys systems with n log-uniform in [nmin, nmax] (seeded),
    concatenated in order of size so that the comparison path can issue one plain
    launch per distinct n over the same buffers.  Reports the persistent single
    launch (eager and replayed from a captured graph) against the per-size plain
    launches (eager and captured)."""
    import ctypes
    import numpy as np
    import oracle
    log("prepare ragged")
    sizes = np.sort(oracle.ragged_sizes(n_sys, nmin, nmax, seed=2024))
    off = oracle.offsets_from_sizes(sizes)
    U = int(off[-1])
    arrs = [v.astype(np.float32) for v in oracle.gen_dd_ragged(off, seed=2025)]
    dev = torch.device("cuda:0")
    src = [torch.from_numpy(v).to(dev) for v in arrs]
    work = [t.clone() for t in src]
    stream = torch.cuda.current_stream()
    plan = T.RaggedPlan(off, dtype=torch.float32, device=dev)
    alg_bytes = 5 * 4 * U + 8 * (n_sys + 1)

    # parity of both paths on fresh inputs
    log("ragged parity")
    xref = oracle.thomas_ragged(*arrs, off)
    plan.solve(*work)
    torch.cuda.synchronize()
    err_ragged = _ragged_relerr(work[3].cpu().numpy().astype(np.float64), xref, off)
    L = T.lib()
    h = ctypes.c_void_p(stream.cuda_stream)
    uniq, first, counts = np.unique(sizes, return_index=True, return_counts=True)
    ptrs = [t.data_ptr() for t in work]

    def plain_args(handle):
        out = []
        for n, f, k in zip(uniq.tolist(), first.tolist(), counts.tolist()):
            o = int(off[f]) * 4
            out.append((0, -1, 0, n, k) + tuple(ctypes.c_void_p(p + o) for p in ptrs) +
                       (handle,))
        return out
    pargs = plain_args(h)

    def plain_sweep(args=pargs):
        rc = 0
        for a in args:
            rc |= L.tridiag_solve(*a)
        return rc
    work[3].copy_(src[3])
    T.check(plain_sweep())
    torch.cuda.synchronize()
    err_plain = _ragged_relerr(work[3].cpu().numpy().astype(np.float64), xref, off)
    del xref

    log("ragged timing")
    for _ in range(warmup):
        plan.solve(*work)
    _settle(torch, lambda: plan.solve(*work))
    wall_r, ev_r = _event_time(torch, lambda: plan.solve(*work), steps, stream)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(stream)
    with torch.cuda.graph(g, stream=s):
        plan.solve(*work)
    torch.cuda.synchronize()
    for _ in range(warmup):
        g.replay()
    _settle(torch, g.replay)
    wall_rg, ev_rg = _event_time(torch, g.replay, steps, stream)

    log("plain per-size timing (%d launches per sweep)" % len(pargs))
    psteps = max(2, min(steps, 10))
    for _ in range(min(warmup, 2)):
        plain_sweep()
    wall_p, ev_p = _event_time(torch, plain_sweep, psteps, stream)
    gp = torch.cuda.CUDAGraph()
    pargs_s = plain_args(ctypes.c_void_p(s.cuda_stream))
    with torch.cuda.graph(gp, stream=s):
        T.check(plain_sweep(pargs_s))
    torch.cuda.synchronize()
    for _ in range(min(warmup, 2)):
        gp.replay()
    wall_pg, ev_pg = _event_time(torch, gp.replay, psteps, stream)
    del g, gp, work
    torch.cuda.empty_cache()

    log("plain bucketed timing")
    bucketed = _ragged_bucketed(torch, T, arrs, sizes, off, steps, warmup, stream)
    del src
    torch.cuda.empty_cache()

    achieved = alg_bytes / (ev_r * 1e-3) / 1e9
    ratio = load_traffic("ragged_f32|%d shapes|traffic_per_alg_byte" %
                         T.lib().tridiag_ragged_shape_count())
    r = {
        "systems": n_sys, "unknowns": U, "distinct_n": int(len(uniq)), "dtype": "f32",
        "sizes": "n = round(exp(U[ln %d, ln %d])), seed 2024, ordered by size" % (nmin, nmax),
        "persistent": {"ms_per_step": wall_r, "kernel_ms": ev_r,
                       "unknowns_per_s": U / (wall_r * 1e-3), "relerr": err_ragged},
        "persistent_graph": {"ms_per_step": wall_rg, "kernel_ms": ev_rg,
                             "unknowns_per_s": U / (wall_rg * 1e-3)},
        "plain_per_size": {"launches": int(len(uniq)), "ms_per_step": wall_p, "gpu_ms": ev_p,
                           "unknowns_per_s": U / (wall_p * 1e-3), "relerr": err_plain},
        "plain_per_size_graph": {"ms_per_step": wall_pg, "gpu_ms": ev_pg,
                                 "unknowns_per_s": U / (wall_pg * 1e-3)},
        "plain_bucketed": bucketed,
        "speedup_vs_plain_graph": wall_pg / wall_rg,
        "speedup_vs_bucketed_graph": bucketed["graph"]["ms_per_step"] / wall_rg,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (round(ratio * alg_bytes) if ratio else None)},
    }
    log("ragged", json.dumps(r))
    return r


# Capacities of the padded buckets a caller without a ragged solver would use:
# powers of two and 1.5x powers of two from 32 to 8192 rows.
RAGGED_BUCKETS = sorted({c for k in range(9) for c in (32 << k, 48 << k) if c <= 8192})


def _ragged_bucketed(torch, T, arrs, sizes, off, steps, warmup, stream):
    """The baseline a caller would use without a ragged solver: systems grouped by
    size into RAGGED_BUCKETS, each padded with identity rows (a=0, b=1, c=0, d=0)
    to its bucket's capacity and laid out contiguously, then one plain
    tridiag_solve per bucket.  Timed eager and graph-replayed like the persistent
    launch; unknowns/s counts the real (unpadded) unknowns.  The padding is the
    caller's layout and happens before timing."""
    import ctypes
    import numpy as np
    import oracle
    U = int(off[-1])
    dev = torch.device("cuda:0")
    padded, meta = [], []
    prev = 0
    for cap in RAGGED_BUCKETS:
        lo, hi = np.searchsorted(sizes, prev, side="right"), np.searchsorted(sizes, cap,
                                                                             side="right")
        prev = cap
        if hi <= lo:
            continue
        ns = sizes[lo:hi]
        rows = np.arange(cap)
        mask = rows[None, :] < ns[:, None]
        srcix = np.where(mask, off[lo:hi][:, None] + rows[None, :], 0)
        fill = (0.0, 1.0, 0.0, 0.0)
        padded.append([torch.from_numpy(np.where(mask, v[srcix], f).astype(np.float32).ravel())
                       .to(dev) for v, f in zip(arrs, fill)])
        meta.append((cap, hi - lo, lo, hi, mask))
        del srcix
    L = T.lib()

    def args_for(handle):
        return [(0, -1, 0, cap, cnt) + tuple(ctypes.c_void_p(t.data_ptr()) for t in pb) + (handle,)
                for (cap, cnt, _, _, _), pb in zip(meta, padded)]
    bargs = args_for(ctypes.c_void_p(stream.cuda_stream))

    def sweep(a=bargs):
        rc = 0
        for x in a:
            rc |= L.tridiag_solve(*x)
        return rc
    # parity: the real rows of every bucket against the FP64 oracle
    keep = [t[3].clone() for t in padded]
    T.check(sweep())
    torch.cuda.synchronize()
    x = np.empty(U)
    for (cap, cnt, lo, hi, mask), pb in zip(meta, padded):
        xb = pb[3].cpu().numpy().reshape(cnt, cap)
        x[off[lo]:off[hi]] = xb[mask]
    for pb, k in zip(padded, keep):
        pb[3].copy_(k)
    del keep
    xref = oracle.thomas_ragged(*arrs, off)
    err = _ragged_relerr(x, xref, off)
    del xref, x
    for _ in range(warmup):
        sweep()
    _settle(torch, sweep)
    wall, ev = _event_time(torch, sweep, steps, stream)
    s = torch.cuda.Stream()
    s.wait_stream(stream)
    g = torch.cuda.CUDAGraph()
    sargs = args_for(ctypes.c_void_p(s.cuda_stream))
    with torch.cuda.graph(g, stream=s):
        T.check(sweep(sargs))
    torch.cuda.synchronize()
    for _ in range(warmup):
        g.replay()
    _settle(torch, g.replay)
    wall_g, ev_g = _event_time(torch, g.replay, steps, stream)
    padded_rows = int(sum(cap * cnt for cap, cnt, _, _, _ in meta))
    del g, padded
    return {"launches": len(meta), "buckets": [m[0] for m in meta],
            "padded_unknowns": padded_rows, "relerr": err,
            "eager": {"ms_per_step": wall, "gpu_ms": ev, "unknowns_per_s": U / (wall * 1e-3)},
            "graph": {"ms_per_step": wall_g, "gpu_ms": ev_g,
                      "unknowns_per_s": U / (wall_g * 1e-3)}}


def run_vendor(torch, T, steps, warmup, ours):
    """The reference's vendor comparator (cuSPARSE gtsv2StridedBatch behind the
    cuSparse_* symbols, Tridiagonal_CUDA/tridiagonal_basic.cu:205-247) in its ROCm
    form, hipSPARSE gtsv2StridedBatch (algo id 0), at 4096 x 4096 layout 0, timed
    the same way as the solve; `ours_speedup` = its kernel time / ours."""
    import numpy as np
    out = {}
    for name, dt in (("f32", np.float32), ("f64", np.float64)):
        sets = make_sets(torch, dt, N_EQT, N_BATCH, N_SETS, seed0=11, layout=0)
        k = max(2, min(steps, 20))
        wall, kern = time_solver(torch, T, sets, N_EQT, N_BATCH, k, min(warmup, 3), 0, 0)
        del sets
        torch.cuda.empty_cache()
        par = check_parity(torch, T, dt, N_EQT, N_BATCH, 0, 0, 1e-5 if dt == np.float32 else 1e-12)
        mine = ours[name + "_contiguous"]["kernel_ms"]
        out[name] = {"kernel_ms": kern, "ms_per_step": wall, "steps": k,
                     "unknowns_per_s": N_EQT * N_BATCH / (kern * 1e-3), "parity": par,
                     "ours_kernel_ms": mine, "ours_speedup": kern / mine}
        log("vendor gtsv", name, json.dumps(out[name]))
    return out


def run_convert(torch, T, steps, warmup, n=N_EQT, batch=N_BATCH):
    """Layout conversion of the four coefficient arrays at 4096 x 4096 (contiguous to
    interleaved and back), FP32 and FP64.  Algorithmic bytes: 2 * w per element
    (read once, write once) x 4 arrays."""
    import ctypes
    L = T.lib()
    stream = torch.cuda.current_stream()
    h = ctypes.c_void_p(stream.cuda_stream)
    out = {}
    for prec, dt in ((0, torch.float32), (1, torch.float64)):
        g = torch.Generator(device="cuda").manual_seed(8)
        src = [torch.rand(batch * n, generator=g, device="cuda", dtype=dt) for _ in range(4)]
        dst = [torch.empty_like(t) for t in src]
        P = lambda ts: [ctypes.c_void_p(t.data_ptr()) for t in ts]
        w = 4 if prec == 0 else 8
        alg = 2 * w * n * batch * 4
        for name, sl, dl, a_, b_ in (("contiguous_to_interleaved", 0, 2, src, dst),
                                     ("interleaved_to_contiguous", 2, 0, dst, src)):
            args = (prec, sl, dl, n, batch, *P(a_), *P(b_), h)
            for _ in range(warmup):
                T.check(L.tridiag_convert_layout(*args))
            _settle(torch, lambda: L.tridiag_convert_layout(*args))
            rc = [0]

            def f():
                rc[0] |= L.tridiag_convert_layout(*args)
            _, ev = _event_time(torch, f, steps, stream)
            T.check(rc[0])
            gbs = alg / (ev * 1e-3) / 1e9
            out["%s_%s" % ("f32" if prec == 0 else "f64", name)] = {
                "kernel_ms": ev, "achieved_gbs": round(gbs, 1),
                "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        # bit-exact round trip of the last pair of calls
        torch.cuda.synchronize()
        chk = [torch.empty_like(t) for t in src]
        T.check(L.tridiag_convert_layout(prec, 0, 2, n, batch, *P(src), *P(dst), h))
        T.check(L.tridiag_convert_layout(prec, 2, 0, n, batch, *P(dst), *P(chk), h))
        torch.cuda.synchronize()
        ok = all(torch.equal(x, y) for x, y in zip(src, chk)) and \
            torch.equal(dst[0].view(n, batch), src[0].view(batch, n).t())
        out["%s_round_trip_bit_exact" % ("f32" if prec == 0 else "f64")] = bool(ok)
        del src, dst, chk
        torch.cuda.empty_cache()
    return out


def run_fwd_bwd(torch, T, steps, warmup, n=512, batch=8192, layout=0, min_bytes=512 << 20):
    """Differentiable solve, FP32, n x batch: plain forward solve vs the fused
    forward + adjoint kernel (and the stand-alone adjoint).  Input sets (a, b, c,
    d, g and each call's outputs) are rotated, at least min_bytes of inputs in
    total, so that no call reads what an earlier one left in the 256 MB MALL: the
    ratio is between HBM-bound figures.  The forward solve runs in place on its
    set's own copy of d (the values drift from call to call; the work does not)."""
    import ctypes
    import numpy as np
    import oracle
    log("prepare fwd/bwd")
    dev = torch.device("cuda:0")
    U = n * batch
    nsets = max(3, -(-min_bytes // (5 * 4 * U)))
    w = np.random.default_rng(42).standard_normal(U)          # L = sum(w * x)
    sets = []
    for k in range(nsets):
        a, b, c, d = oracle.gen_dd(n, batch, layout=layout, seed=41 + k)
        ins = [torch.from_numpy(v.astype(np.float32)).to(dev) for v in (a, b, c, d)]
        g = torch.from_numpy(np.roll(w, 977 * k).astype(np.float32)).to(dev)
        outs = [torch.empty_like(g) for _ in range(5)]          # x, ga, gb, gc, gd
        adj = [torch.empty_like(g) for _ in range(4)]
        sets.append({"in": ins, "g": g, "out": outs, "adj": adj, "xf": ins[3].clone()})
    stream = torch.cuda.current_stream()
    h = ctypes.c_void_p(stream.cuda_stream)
    L = T.lib()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    fwd_args = [(0, -1, layout, n, batch, P(z["in"][0]), P(z["in"][1]), P(z["in"][2]),
                 P(z["xf"]), h) for z in sets]
    fb_args = [(0, layout, n, batch) + tuple(P(t) for t in z["in"]) + (P(z["g"]),) +
               tuple(P(o) for o in z["out"]) + (h,) for z in sets]
    adj_args = [(0, layout, n, batch, P(z["in"][0]), P(z["in"][1]), P(z["in"][2]),
                 P(z["out"][0]), P(z["g"])) + tuple(P(o) for o in z["adj"]) + (h,)
                for z in sets]
    for a_ in fb_args:
        T.check(L.tridiag_solve_fwd_bwd(*a_))
    torch.cuda.synchronize()
    # parity of set 0's fused outputs against the FP64 oracle
    arrs = [t.cpu().numpy().astype(np.float64) for t in sets[0]["in"]]
    xr = oracle.thomas(*arrs, n, batch, layout=layout)
    gr = oracle.adjoint(*arrs[:3], xr, sets[0]["g"].cpu().numpy().astype(np.float64), n, batch,
                        layout=layout)
    errs = [float(oracle.errstats(o.cpu().numpy().astype(np.float64), r, n, batch,
                                  layout)["relerr"]) for o, r in zip(sets[0]["out"], (xr,) + tuple(gr))]
    del arrs, xr, gr
    rc = [0]
    turn = [0]

    def rot(fn, args):
        def f():
            turn[0] += 1
            rc[0] |= fn(*args[turn[0] % nsets])
        return f
    f1 = rot(L.tridiag_solve, fwd_args)
    f2 = rot(L.tridiag_solve_fwd_bwd, fb_args)
    f3 = rot(L.tridiag_adjoint, adj_args)
    for _ in range(warmup):
        f1(); f2(); f3()
    _settle(torch, f2)
    wf, ef = _event_time(torch, f1, steps, stream)
    _settle(torch, f2)
    wb, eb = _event_time(torch, f2, steps, stream)
    _settle(torch, f3)
    wa, ea = _event_time(torch, f3, steps, stream)
    T.check(rc[0])
    del sets
    torch.cuda.empty_cache()
    r = {
        "n": n, "batch": batch, "layout": {0: "contiguous", 2: "interleaved"}[layout],
        "dtype": "f32", "input_sets": nsets,
        "input_bytes_rotated": nsets * 5 * 4 * U,
        "forward": {"kernel_ms": ef, "unknowns_per_s": U / (ef * 1e-3),
                    "gbs": 5 * 4 * U / (ef * 1e-3) / 1e9},
        "fused_fwd_bwd": {"kernel_ms": eb, "unknowns_per_s": U / (eb * 1e-3),
                          "gbs": 10 * 4 * U / (eb * 1e-3) / 1e9},
        "adjoint": {"kernel_ms": ea, "gbs": 9 * 4 * U / (ea * 1e-3) / 1e9},
        "fwd_bwd_over_forward": eb / ef,
        "relerr_x_ga_gb_gc_gd": errs,
    }
    log("fwd_bwd", json.dumps(r))
    return r


def _host_threads():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))


def run_n1024(torch, T, steps, warmup, with_cpu, budget_s=2.0):
    """FP64 n=1024 x batch=1024, layout 0: the reference's CPU-runnable case.  Its
    4 x 8 MiB of coefficients fit the 256 MB MALL, so the HBM fraction is reported,
    not graded.  Besides the event-timed solve, the reference-style figure: the
    synchronous Part-1 entry (WM5_reg_double, Thomas_double) timed by the host
    clock around each call with inputs restored untimed before it, as the reference
    driver does (Tridiagonal_CUDA/test.cpp:94-100).  With the CPU leg, LAPACK dgtsv
    (scipy; partial pivoting, the reference's CPU comparator LAPACKE_?gtsv at
    Tridiagonal_AVX512/tridiagonal_basic.cpp:131-143) on the same workload, reported
    only; the reference's own CPU solvers at this shape are in cpu_baseline.points."""
    import ctypes
    import numpy as np
    import oracle
    n = batch = 1024
    U = n * batch
    sets = make_sets(torch, np.float64, n, batch, N_SETS, 21, 0)
    wall, kern = time_solver(torch, T, sets, n, batch, steps, warmup, -1, 0)
    par = check_parity(torch, T, np.float64, n, batch, -1, 0, 1e-12)
    # the Thomas id itself (one wave per system, 9 words per unknown), event-timed
    twall, tkern = time_solver(torch, T, sets, n, batch, steps, warmup, 1, 0)
    thomas = {"algo": 1, "plan": T.plan(1, 1, 0, n, batch), "ms_per_step": twall,
              "kernel_ms": tkern, "unknowns_per_s": U / (tkern * 1e-3),
              "traffic_rate_gbs_9w": round(9 * 8 * U / (tkern * 1e-3) / 1e9, 1),
              "parity": check_parity(torch, T, np.float64, n, batch, 1, 0, 1e-12)}
    L = T.lib()
    src = sets[0]
    work = [t.clone() for t in src]
    ptrs = [ctypes.c_void_p(t.data_ptr()) for t in work]
    ref_style = {}
    for name in ("WM5_reg_double", "Thomas_double"):
        fn = getattr(L, name)
        fn.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int]
        fn.restype = None
        for _ in range(max(warmup, 1)):
            fn(*ptrs, n, batch)
        tt = 0.0
        for _ in range(steps):
            for w_, s_ in zip(work, src):
                w_.copy_(s_)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn(*ptrs, n, batch)
            tt += time.perf_counter() - t0
        ref_style[name] = {"wall_ms": tt * 1e3 / steps, "unknowns_per_s": U * steps / tt}
    r = {"n": n, "batch": batch, "layout": "contiguous", "dtype": "f64",
         "plan": T.plan(1, -1, 0, n, batch), "ms_per_step": wall, "kernel_ms": kern,
         "unknowns_per_s": U / (kern * 1e-3),
         "hbm_frac_mall_resident": 5 * 8 * U / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS,
         "parity": par, "thomas": thomas, "reference_style_sync_call": ref_style}
    del sets, work
    torch.cuda.empty_cache()
    if with_cpu:
        a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=21)
        x_or = oracle.thomas(a, b, c, d, n, batch, 0)
        try:
            from scipy.linalg import lapack
        except ImportError:
            lapack = None
        if lapack is not None:
            A, B, C, D = (v.reshape(batch, n) for v in (a, b, c, d))
            x_lp = np.empty_like(D)
            reps, t = 0, 0.0
            while t < budget_s:
                t0 = time.perf_counter()
                for s_ in range(batch):
                    out = lapack.dgtsv(A[s_, 1:], B[s_], C[s_, :-1], D[s_])
                    x_lp[s_] = out[3]
                t += time.perf_counter() - t0
                reps += 1
            err = float(np.max(np.abs(x_lp.ravel() - x_or)) / np.max(np.abs(x_or)))
            r["cpu_lapack_dgtsv"] = {"unknowns_per_s": U * reps / t, "threads": 1,
                                     "reps": reps, "relerr_vs_oracle": err}
    return r


def run_large(torch, T, steps, warmup):
    """Large single systems, FP64, dominance margin U[0.05, 0.2]: n = 2^20 x 16 and
    2^16 x 256 systems (contiguous layout), solved by the automatic plan: the chunked
    three-pass solve, or the cooperative row-block solve where its blocks fit."""
    import numpy as np
    out = {}
    for n, batch in ((1 << 20, 16), (1 << 16, 256)):
        name = "f64_n%d_b%d" % (n, batch)
        log("prepare", name)
        sets = []
        import oracle
        for k in range(2):
            a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=300 + k, blo=0.05, bhi=0.2)
            sets.append([torch.from_numpy(v).cuda() for v in (a, b, c, d)])
            del a, b, c, d
        log("time", name)
        wall, kern = time_solver(torch, T, sets, n, batch, steps, warmup, -1, 0)
        del sets
        torch.cuda.empty_cache()
        par = check_parity_lowmargin(torch, T, n, batch)
        U = n * batch
        achieved = 5 * 8 * U / (kern * 1e-3) / 1e9
        pl = T.plan(1, -1, 0, n, batch)
        out[name] = {"n": n, "batch": batch, "ms_per_step": wall, "kernel_ms": kern,
                     "unknowns_per_s": U / (wall * 1e-3), "achieved_gbs": achieved,
                     "hbm_frac": achieved / HBM_PEAK_GBS, "parity": par, "plan": pl,
                     "path": {"chunked": "chunked: reduce / chunk-interface merge / solve",
                              "cooperative": "cooperative row blocks (5 words per unknown)"}
                     .get(pl["kind"], pl["kind"])}
        log(name, json.dumps(out[name]))
    return out


def check_parity_lowmargin(torch, T, n, batch):
    import numpy as np
    import oracle
    a, b, c, d = oracle.gen_dd(n, batch, layout=0, seed=999, blo=0.05, bhi=0.2)
    ts = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    T.solve(*ts, algo=-1, layout=0, n=n, batch=batch)
    x = ts[3].cpu().numpy()
    xref = oracle.thomas(a, b, c, d, n, batch, layout=0)
    st = oracle.errstats(x, xref, n, batch, 0)
    return dict(relerr=float(st["relerr"]), tol=1e-12, ok=bool(st["relerr"] <= 1e-12))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only the CPU baseline leg (no GPU); prints its JSON object")
    args = ap.parse_args()
    if args.cpu_only:
        print(json.dumps(cpu_leg()), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (systems are "
              "independent; sharding the batch across GPUs needs no communication).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import tridiagonalsolver_amd as T
    T.lib()
    torch.cuda.set_device(0)
    res = {}
    for name, dtname, layout in WORKLOADS:
        res[name] = run_workload(torch, T, name, dtname, layout, args.steps, args.warmup)
    main_r = res["f32_contiguous"]
    pcr = run_pcr(torch, T, args.steps, args.warmup)
    bulk = run_bulk(torch, T, args.steps, args.warmup)
    large = run_large(torch, T, args.steps, args.warmup)
    ragged = run_ragged(torch, T, args.steps, args.warmup)
    fwdbwd = run_fwd_bwd(torch, T, args.steps, args.warmup)
    log("hipSPARSE gtsv2StridedBatch")
    vendor = run_vendor(torch, T, args.steps, args.warmup, res)
    log("layout conversion")
    convert = run_convert(torch, T, args.steps, args.warmup)
    log("n=1024 x 1024 FP64")
    n1024 = run_n1024(torch, T, args.steps, args.warmup, not args.no_cpu)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline()
    line = {
        "metric": "unknowns/s & HBM % of peak (4096\u00b2); ragged+fwd/bwd vs plain "
                  "@relerr\u22641e-5",
        "value": main_r["unknowns_per_s"], "unit": "unknowns/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "settle_ms": SETTLE_MS,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched tridiagonal solve (reg WM-pGE, auto plan), n=4096 x "
                               "batch=4096, FP32, contiguous per-system layout (reference "
                               "layout 0, whole-line loads); FP64 and the interleaved layout "
                               "in 'workloads'",
                   "n": N_EQT, "batch": N_BATCH, "layout": "contiguous", "plan": main_r["plan"]},
        "roofline": main_r["roofline"],
        "cpu_baseline": cpu,
        "workloads": {k: {kk: v[kk] for kk in ("unknowns_per_s", "ms_per_step", "kernel_ms",
                                               "reference_style_sync_ms", "layout", "plan",
                                               "array_spacing_mib", "parity", "roofline")}
                      for k, v in res.items()},
        "pcr_4096x4096": pcr,
        "bulk_row_forms_4096x4096": bulk,
        "large": large,
        "ragged": ragged,
        "n1024_f64": n1024,
        "layout_convert_4096x4096": convert,
        "vendor_gtsv2_4096x4096": vendor,
        "fwd_bwd": fwdbwd,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
