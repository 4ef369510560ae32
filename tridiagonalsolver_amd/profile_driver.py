"""Small fixed workloads for hardware-counter and kernel-trace runs (rocprofv3).

    rocprofv3 --pmc FETCH_SIZE -d OUT -o run --output-format csv -- \\
        python3 -m tridiagonalsolver_amd.profile_driver --prec 0 --layout 2 --n 4096 --batch 4096

Runs, in order: a plain device copy of --copy-mib MiB (its FETCH_SIZE / WRITE_SIZE
per dispatch calibrate the counters against a known byte count), then --reps
launches of one workload on seeded inputs (a, c ~ U[-1,1], b = |a|+|c|+U[blo,bhi],
d ~ U[-1,1], drawn with torch on the device):

    solve    tridiag_solve with the default plan (or --tile INDEX)
    ragged   one planned ragged solve of --nsys systems, n log-uniform in [--nmin, --nmax]
    fwd_bwd  tridiag_solve_fwd_bwd (fused forward + adjoint)
    adjoint  tridiag_adjoint
"""
import argparse
import ctypes
import sys

from . import _capi


def _coeffs(torch, shape, dtype, blo, bhi, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    a = (torch.rand(shape, generator=g, device="cuda", dtype=dtype) * 2 - 1).contiguous()
    c = (torch.rand(shape, generator=g, device="cuda", dtype=dtype) * 2 - 1).contiguous()
    b = a.abs() + c.abs() + blo + (bhi - blo) * torch.rand(shape, generator=g, device="cuda",
                                                            dtype=dtype)
    d = (torch.rand(shape, generator=g, device="cuda", dtype=dtype) * 2 - 1).contiguous()
    return a, b.contiguous(), c, d


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="solve", choices=["solve", "ragged", "fwd_bwd",
                                                            "adjoint"])
    ap.add_argument("--prec", type=int, default=0)
    ap.add_argument("--layout", type=int, default=2)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--algo", type=int, default=-1)
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--blo", type=float, default=1.0)
    ap.add_argument("--bhi", type=float, default=2.0)
    ap.add_argument("--nsys", type=int, default=65536)
    ap.add_argument("--nmin", type=int, default=32)
    ap.add_argument("--nmax", type=int, default=8192)
    ap.add_argument("--align", type=int, default=1,
                    help="ragged: round sizes up to a multiple (1 = as drawn)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=256)
    args = ap.parse_args(argv)
    import torch
    L = _capi.lib()
    torch.cuda.set_device(0)
    src = torch.ones(args.copy_mib << 18, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    dtype = torch.float32 if args.prec == 0 else torch.float64
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    desc = ""
    if args.workload == "ragged":
        from .solver import RaggedPlan
        gen = torch.Generator().manual_seed(2024)
        u = torch.rand(args.nsys, generator=gen, dtype=torch.float64)
        lo, hi = torch.log(torch.tensor(float(args.nmin))), torch.log(torch.tensor(float(args.nmax)))
        sizes = torch.exp(lo + (hi - lo) * u).round().clamp(args.nmin, args.nmax).to(torch.int64)
        if args.align > 1:
            sizes = ((sizes + args.align - 1) // args.align * args.align).clamp(max=args.nmax)
        off = torch.zeros(args.nsys + 1, dtype=torch.int64)
        off[1:] = torch.cumsum(sizes, 0)
        rows = int(off[-1])
        a, b, c, d = _coeffs(torch, (rows,), dtype, args.blo, args.bhi, 5)
        plan = RaggedPlan(off, dtype=dtype, device="cuda")
        for _ in range(args.reps):
            plan.solve(a, b, c, d)
        desc = "ragged %d systems, %d rows" % (args.nsys, rows)
    else:
        shape = (args.batch, args.n) if args.layout == 0 else (args.n, args.batch)
        a, b, c, d = _coeffs(torch, shape, dtype, args.blo, args.bhi, 5)
        ptr = [P(t) for t in (a, b, c, d)]
        if args.workload == "solve":
            for _ in range(args.reps):
                if args.tile >= 0:
                    rc = L.tridiag_solve_tile(args.prec, args.tile, args.layout, args.n,
                                              args.batch, *ptr, st)
                else:
                    rc = L.tridiag_solve(args.prec, args.algo, args.layout, args.n, args.batch,
                                         *ptr, st)
                _capi.check(rc)
            try:
                desc = "plan %s" % (_capi.plan(args.prec, args.algo, args.layout, args.n,
                                               args.batch) if args.tile < 0
                                    else _capi.tiles(args.prec)[args.tile])
            except _capi.TridiagError:
                desc = "chunked solve (n above the on-chip capacity)"
        else:
            g = torch.randn(shape, device="cuda", dtype=dtype)
            outs = [torch.empty_like(d) for _ in range(5)]
            for _ in range(args.reps):
                if args.workload == "fwd_bwd":
                    rc = L.tridiag_solve_fwd_bwd(args.prec, args.layout, args.n, args.batch,
                                                 *ptr, P(g), *(P(o) for o in outs), st)
                else:
                    rc = L.tridiag_adjoint(args.prec, args.layout, args.n, args.batch, ptr[0],
                                           ptr[1], ptr[2], ptr[3], P(g),
                                           *(P(o) for o in outs[:4]), st)
                _capi.check(rc)
            desc = args.workload
    torch.cuda.synchronize()
    print("profile_driver: copy %d MiB, %d x %s, prec %d layout %d n %d batch %d: %s"
          % (args.copy_mib, args.reps, args.workload, args.prec, args.layout, args.n,
             args.batch, desc), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
