"""Solves a fixed set of cases with the library that $TRIDIAG_DIV_MODE selects and
prints one JSON line per case: the policy the library reports, the plan, and the
normwise relative error against the FP64 oracle.  Run as a child process by
tests/test_gpu_div_modes.py (the binding loads one library per process)."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
from gpu_util import check_against_oracle, make_case, run_solve  # noqa: E402

DT = {0: np.float32, 1: np.float64}

# (precision, algo, layout, n, batch): the lane kernels, Thomas, the tiles with each
# interface family, the cooperative and chunked solves
CASES = [
    (p, algo, layout, n, batch)
    for p in (0, 1)
    for (algo, layout, n, batch) in (
        (-1, 0, 16, 300), (1, 0, 1000, 64), (1, 2, 1000, 64), (9, 0, 4096, 16),
        (9, 2, 4096, 64), (5, 0, 1000, 32), (2, 0, 1000, 32), (12, 0, 2048, 16),
        (-1, 2, 4096, 512), (-1, 0, 70000, 2), (-1, 0, 20000, 8))
]


def main():
    import torch
    import tridiagonalsolver_amd as T
    out = []
    mode = T.div_mode()
    for k, (p, algo, layout, n, batch) in enumerate(CASES):
        print(json.dumps({"case": [p, algo, layout, n, batch]}), file=sys.stderr, flush=True)
        arrs = make_case(n, batch, layout, DT[p], seed=100 + k)
        x = run_solve(arrs, n, batch, layout, algo, DT[p])
        st, res = check_against_oracle(arrs, x, n, batch, layout, DT[p])
        out.append({"mode": mode, "prec": p, "algo": algo, "layout": layout, "n": n,
                    "batch": batch, "plan": T.plan(p, algo, layout, n, batch)["kind"],
                    "relerr": float(st["relerr"]), "nonfinite": int(st["nonfinite"])})
    # ragged and the fused forward + adjoint
    off = np.concatenate([[0], np.cumsum(np.random.default_rng(5).integers(1, 3000, 200))])
    arrs = [v.astype(np.float32) for v in oracle.gen_dd_ragged(off, seed=9)]
    xr = oracle.thomas_ragged(*(v.astype(np.float64) for v in arrs), off)
    dev = torch.device("cuda:0")
    ts = [torch.from_numpy(v).to(dev) for v in arrs]
    T.solve_ragged(*ts, off)
    torch.cuda.synchronize()
    xg = ts[3].cpu().numpy().astype(np.float64)
    out.append({"mode": mode, "case": "ragged_f32", "relerr":
                float(np.max(np.abs(xg - xr)) / np.max(np.abs(xr)))})
    n, batch = 512, 64
    g0 = np.random.default_rng(4).standard_normal(n * batch)
    for p in (0, 1):
        a, b, c, d = make_case(n, batch, 0, DT[p], seed=3)
        g = g0.astype(DT[p])
        x64 = oracle.thomas(*(v.astype(np.float64) for v in (a, b, c, d)), n, batch)
        ref = oracle.adjoint(*(v.astype(np.float64) for v in (a, b, c)), x64,
                             g.astype(np.float64), n, batch)
        ts = [torch.from_numpy(v).to(dev) for v in (a, b, c, d, g)]
        x, *grads = T.solve_fwd_bwd(*ts, n=n, batch=batch)
        torch.cuda.synchronize()
        errs = [float(np.max(np.abs(u.cpu().numpy().astype(np.float64) - r)) /
                      np.max(np.abs(r))) for u, r in zip([x] + list(grads), [x64] + list(ref))]
        out.append({"mode": mode, "case": "fwd_bwd", "prec": p, "relerr": max(errs)})
    for r in out:
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
