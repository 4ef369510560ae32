"""GPU parity of the reference's own forms of ids 2-13 (tridiag_solve_bulk:
PCR-pThomas and WM-pGE streamed through one wave per system, full and half forms,
LDS or ds_bpermute exchanges; CR with its levels in global memory; CR over R
rows per lane), through the C-ABI, against the FP64 oracle:
normwise relative error <= 1e-5 (FP32) / 1e-12 (FP64).  Sizes straddle the
64-row blocks and 128-row chunks and the prefetch rings (4 blocks forward, 8
back), CR's odd and even level lengths and its 64-row last level, and include
sizes the reference itself does not take (n % 64 != 0)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from gpu_util import TOL, announce, check_against_oracle, make_case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DT = {0: np.float32, 1: np.float64}
ALGOS = [2, 3, 4, 5, 6, 7, 8, 9, 10]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _bulk(arrs, n, batch, layout, algo, dtype):
    import torch
    import tridiagonalsolver_amd as T
    ts = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in arrs]
    if layout == 1:
        T.solve_bulk(ts[0], None, None, None, algo, layout=1, n=n, batch=batch)
    else:
        T.solve_bulk(*ts, algo, layout=layout, n=n, batch=batch)
    torch.cuda.synchronize()
    return ts[0].cpu().numpy() if layout == 1 else ts[3].cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo", ALGOS)
def test_bulk_sizes(prec, algo):
    for n, batch in ((1, 5), (2, 3), (63, 7), (64, 9), (65, 4), (127, 3), (128, 5), (129, 6),
                     (255, 2), (256, 3), (1000, 11), (1024, 6), (4096, 5), (5000, 3)):
        announce("bulk", prec, algo, n, batch)
        arrs = make_case(n, batch, 0, DT[prec], seed=n * 13 + algo)
        x = _bulk(arrs, n, batch, 0, algo, DT[prec])
        st, res = check_against_oracle(arrs, x, n, batch, 0, DT[prec])
        assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], (n, batch, st)
        assert res <= 10 * TOL[DT[prec]], (n, batch, res)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo,R", [(11, 4), (12, 8), (13, 16)])
def test_bulk_register_cr(prec, algo, R):
    """CR over R rows per lane (CR_reg4/8/16), one workgroup per system of at most
    1024 R rows: sizes around one lane, one wave and the largest system."""
    for n, batch in ((1, 3), (2, 5), (R - 1, 4), (R, 3), (R + 1, 2), (64 * R - 1, 3),
                     (64 * R, 2), (64 * R + 1, 3), (1000, 7), (1024 * R - 3, 2), (1024 * R, 3)):
        announce("bulk register CR", prec, algo, n, batch)
        arrs = make_case(n, batch, 0, DT[prec], seed=n * 7 + algo)
        x = _bulk(arrs, n, batch, 0, algo, DT[prec])
        st, res = check_against_oracle(arrs, x, n, batch, 0, DT[prec])
        assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], (n, batch, st)
        assert res <= 10 * TOL[DT[prec]], (n, batch, res)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("algo", [2, 5, 6, 9, 10, 12])
def test_bulk_low_margin(prec, algo):
    n, batch = 4096, 20
    announce("bulk low margin", prec, algo)
    arrs = make_case(n, batch, 0, DT[prec], seed=algo, blo=0.05, bhi=0.2)
    x = _bulk(arrs, n, batch, 0, algo, DT[prec])
    st, _ = check_against_oracle(arrs, x, n, batch, 0, DT[prec])
    assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], st


@pytest.mark.parametrize("algo", [2, 3, 9, 12])
def test_bulk_packed_layout(algo):
    n, batch = 1000, 33
    announce("bulk packed", algo)
    (buf,) = make_case(n, batch, 1, np.float32, seed=algo)
    x = _bulk([buf], n, batch, 1, algo, np.float32)
    st, _ = check_against_oracle([buf], x, n, batch, 1, np.float32)
    assert st["relerr"] <= TOL[np.float32], st


def test_bulk_refuses_other_ids_and_layouts():
    import ctypes
    import tridiagonalsolver_amd as T
    from tridiagonalsolver_amd import _capi
    L = T.lib()
    nul = ctypes.c_void_p(0)
    for algo in (0, 1, 14, 17):
        assert L.tridiag_solve_bulk(0, algo, 0, 64, 1, nul, nul, nul, nul, nul) == \
            _capi.EUNSUPPORTED, algo
    # register CR holds at most 1024 R rows in its one workgroup
    assert L.tridiag_solve_bulk(0, 11, 0, 4097, 1, nul, nul, nul, nul, nul) == _capi.EUNSUPPORTED
    assert b"4096 rows" in L.tridiag_last_error()
    assert L.tridiag_solve_bulk(0, 9, 2, 64, 1, nul, nul, nul, nul, nul) == _capi.EUNSUPPORTED
    assert L.tridiag_solve_bulk(0, 99, 0, 64, 1, nul, nul, nul, nul, nul) == _capi.EINVAL


_ROUTED = r"""
import sys, numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(root)r + '/tests')
import tridiagonalsolver_amd as T
from gpu_util import TOL, check_against_oracle, make_case, run_solve
for prec, algo in ((0, 5), (1, 9), (0, 8), (1, 4), (0, 2), (1, 2), (0, 12), (1, 13)):
    dtype = np.float32 if prec == 0 else np.float64
    assert T.plan(prec, algo, 0, 4096, 16)["kind"] == "bulk_row"
    assert T.plan(prec, 11, 0, 8192, 16)["kind"] != "bulk_row"   # beyond 1024 R rows
    assert T.plan(prec, algo, 2, 4096, 16)["kind"] != "bulk_row"
    arrs = make_case(4096, 16, 0, dtype, seed=algo)
    x = run_solve(arrs, 4096, 16, 0, algo, dtype)
    st, _ = check_against_oracle(arrs, x, 4096, 16, 0, dtype)
    assert st["relerr"] <= TOL[dtype], st
# the CR form needs its scratch: the no-allocation entry refuses it
import ctypes, torch
from tridiagonalsolver_amd import _capi
t = torch.zeros(4 * 4096, device="cuda")
p = [ctypes.c_void_p(t.data_ptr())] * 4
assert T.lib().tridiag_solve_noalloc(0, 2, 0, 4096, 1, *p, None) == _capi.EUNSUPPORTED
assert T.lib().tridiag_solve_noalloc(0, 9, 0, 4096, 1, *p, None) == _capi.OK
torch.cuda.synchronize()
print("ok", flush=True)
"""


def test_form_env_routes_reference_ids():
    """TRIDIAG_FORM=bulk: tridiag_solve (and the reference symbols) run ids 2-13 in
    their bulk-row forms in layouts 0 and 1 (child process: read once)."""
    env = dict(os.environ, TRIDIAG_FORM="bulk")
    announce("bulk routing (child process)")
    r = subprocess.run([sys.executable, "-c", _ROUTED % {"root": ROOT}], env=env,
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
