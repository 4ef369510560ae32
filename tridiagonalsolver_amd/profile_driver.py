"""Small fixed workload for hardware-counter runs (rocprofv3 --pmc / --kernel-trace).

    rocprofv3 --pmc FETCH_SIZE -d OUT -o run --output-format csv -- \
        python3 -m tridiagonalsolver_amd.profile_driver --prec 0 --layout 2 --n 4096 --batch 4096

Runs, in order: a plain device copy of --copy-mib MiB (its FETCH_SIZE / WRITE_SIZE
per dispatch calibrate the counters against a known byte count), then --reps
solves of the seeded workload with the default kernel plan (or --tile INDEX).
"""
import argparse
import ctypes
import sys

from . import _capi


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--prec", type=int, default=0)
    ap.add_argument("--layout", type=int, default=2)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--algo", type=int, default=-1)
    ap.add_argument("--tile", type=int, default=-1)
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
    shape = (args.batch, args.n) if args.layout == 0 else (args.n, args.batch)
    g = torch.Generator(device="cuda").manual_seed(5)
    a = (torch.rand(shape, generator=g, device="cuda", dtype=dtype) * 2 - 1).contiguous()
    c = (torch.rand(shape, generator=g, device="cuda", dtype=dtype) * 2 - 1).contiguous()
    b = (a.abs() + c.abs() + 1 + torch.rand(shape, generator=g, device="cuda", dtype=dtype))
    d = (torch.rand(shape, generator=g, device="cuda", dtype=dtype) * 2 - 1).contiguous()
    ptr = [ctypes.c_void_p(t.data_ptr()) for t in (a, b, c, d)]
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(args.reps):
        if args.tile >= 0:
            rc = L.tridiag_solve_tile(args.prec, args.tile, args.layout, args.n, args.batch, *ptr,
                                      st)
        else:
            rc = L.tridiag_solve(args.prec, args.algo, args.layout, args.n, args.batch, *ptr, st)
        _capi.check(rc)
    torch.cuda.synchronize()
    plan = _capi.plan(args.prec, args.algo, args.layout, args.n, args.batch) \
        if args.tile < 0 else _capi.tiles(args.prec)[args.tile]
    print("profile_driver: copy %d MiB, %d solves prec %d layout %d n %d batch %d plan %s"
          % (args.copy_mib, args.reps, args.prec, args.layout, args.n, args.batch, plan),
          flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
