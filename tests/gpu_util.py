"""Helpers for the GPU parity tests: seeded inputs from the oracle's generator,
upload through torch, one call through the C-ABI, comparison with the FP64 oracle."""
import sys

import numpy as np

import oracle

TOL = {np.float32: 1e-5, np.float64: 1e-12}


def announce(*parts):
    print("[gpu-case]", *parts, flush=True)
    sys.stdout.flush()


def make_case(n, batch, layout, dtype, seed, blo=1.0, bhi=2.0):
    a, b, c, d = oracle.gen_dd(n, batch, layout=layout, seed=seed, blo=blo, bhi=bhi)
    if layout == 1:
        buf = a.astype(dtype)
        return (buf,)
    return tuple(v.astype(dtype) for v in (a, b, c, d))


def oracle_x(arrs, n, batch, layout):
    if layout == 1:
        return oracle.thomas(arrs[0].astype(np.float64), None, None, None, n, batch, layout=1)
    return oracle.thomas(*(v.astype(np.float64) for v in arrs), n, batch, layout=layout)


def run_solve(arrs, n, batch, layout, algo, dtype):
    import torch
    import tridiagonalsolver_amd as T
    dev = torch.device("cuda:0")
    ts = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in arrs]
    if layout == 1:
        T.solve(ts[0], None, None, None, algo=algo, layout=1, n=n, batch=batch)
    else:
        T.solve(*ts, algo=algo, layout=layout, n=n, batch=batch)
    torch.cuda.synchronize()
    return ts[0].cpu().numpy() if layout == 1 else ts[3].cpu().numpy()


def check_against_oracle(arrs, x_gpu, n, batch, layout, dtype):
    xref = oracle_x(arrs, n, batch, layout)
    st = oracle.errstats(x_gpu.astype(np.float64), xref, n, batch, layout)
    if layout == 1:
        a64 = arrs[0].astype(np.float64)
        res = oracle.residual(a64, a64, a64, a64, x_gpu.astype(np.float64), n, batch, 1)
    else:
        res = oracle.residual(*(v.astype(np.float64) for v in arrs), x_gpu.astype(np.float64),
                              n, batch, layout)
    return st, res
