"""Every compiled on-chip kernel configuration against the FP64 oracle.

Each tile (rows per lane M, waves NW, systems per workgroup S, layout and load
variant) is run through tridiag_solve_tile at its capacity, and padded tiles also
at a size that leaves part of the last lane segment empty; the batch leaves the
last workgroup partly idle.  Every tile also runs under each of the other interface
solvers (tridiag_solve_tile_iface), or reports that the solver's LDS does not fit.  Tolerance: normwise relative error <= 1e-5 (FP32) /
<= 1e-12 (FP64) on every system.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import TOL, announce, check_against_oracle, make_case
from tridiagonalsolver_amd import _capi

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _run_tile(prec, t, n, batch, seed, iface=None):
    import torch
    import tridiagonalsolver_amd as T
    layout = 2 if t["interleaved"] else 0
    arrs = make_case(n, batch, layout, DT[prec], seed)
    ts = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in arrs]
    ptrs = [ctypes.c_void_p(x.data_ptr()) for x in ts]
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if iface is None:
        rc = T.lib().tridiag_solve_tile(prec, t["index"], layout, n, batch, *ptrs, st)
    else:
        rc = T.lib().tridiag_solve_tile_iface(prec, t["index"], iface, layout, n, batch, *ptrs, st)
        if rc == _capi.EUNSUPPORTED and b"LDS" in T.lib().tridiag_last_error():
            return None, None
    T.check(rc)
    torch.cuda.synchronize()
    x = ts[3].cpu().numpy()
    return check_against_oracle(arrs, x, n, batch, layout, DT[prec])


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("interleaved", [False, True])
def test_every_tile(prec, interleaved):
    import tridiagonalsolver_amd as T
    fails = []
    tiles = [t for t in T.tiles(prec) if t["interleaved"] == interleaved and not t["wide_addr"]]
    assert tiles
    for t in tiles:
        sizes = [t["cap"]] if not t["padded"] else sorted({max(1, t["cap"] - t["M"] // 2 - 1),
                                                            max(1, t["cap"] // 2 + 3)})
        if t["vec"]:                     # 16-byte row loads need n a multiple of 16 bytes
            v = 4 if prec == 0 else 2
            sizes = sorted({max(v, k - k % v) for k in sizes})
        batch = 2 * t["S"] + 1
        for n in sizes:
            announce("tile", prec, t, "n", n, "batch", batch)
            st, res = _run_tile(prec, t, n, batch, seed=t["index"] * 7 + n)
            if not (st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]]):
                fails.append((t, n, st))
    assert not fails, fails[:5]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("interleaved", [False, True])
def test_every_tile_every_interface_solver(prec, interleaved):
    """Wang-Mou half (1), PCR (2), PCR + Thomas (3) and CR (4) on every tile, at the
    tile's first test size."""
    import tridiagonalsolver_amd as T
    fails, ran = [], 0
    tiles = [t for t in T.tiles(prec) if t["interleaved"] == interleaved and not t["wide_addr"]]
    for t in tiles:
        n = t["cap"] if not t["padded"] else max(1, t["cap"] // 2 + 3)
        if t["vec"]:
            v = 4 if prec == 0 else 2
            n = max(v, n - n % v)
        batch = 2 * t["S"] + 1
        for iface in (1, 2, 3, 4):
            announce("tile iface", prec, t, "iface", iface, "n", n, "batch", batch)
            st, res = _run_tile(prec, t, n, batch, seed=t["index"] * 5 + iface, iface=iface)
            if st is None:
                continue
            ran += 1
            if not (st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]]):
                fails.append((t, iface, n, st))
    assert ran > 0
    assert not fails, fails[:5]


@pytest.mark.parametrize("prec", [0, 1])
def test_pcr_depths(prec):
    """PCR + Thomas with every PCR depth d = 1 .. log2(P) (interface-solver code
    3 | d << 4: d PCR steps, then Thomas along the 2^d remaining subsystems) on a
    contiguous and an interleaved tile of each precision; depths past log2(P) are
    refused."""
    import tridiagonalsolver_amd as T
    picked = {}
    for t in T.tiles(prec):
        if t["wide_addr"] or t["padded"] or (prec == 0 and t["vec"] and t["cap"] % 4):
            continue
        key = (t["interleaved"], t["cap"] // t["M"])
        if key[1] >= 64 and key not in picked:
            picked[key] = t
    assert picked
    fails = []
    for (il, P), t in sorted(picked.items())[:6]:
        top = P.bit_length() - 1
        for depth in range(1, top + 1):
            announce("tile pcr depth", prec, t, "P", P, "depth", depth)
            st, res = _run_tile(prec, t, t["cap"], 2 * t["S"] + 1, seed=depth, iface=3 | depth << 4)
            if st is None:
                continue
            if not (st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]]):
                fails.append((t, depth, st))
        layout = 2 if il else 0
        nul = ctypes.c_void_p(0)
        assert T.lib().tridiag_solve_tile_iface(prec, t["index"], 3 | (top + 1) << 4, layout,
                                                t["cap"], 1, nul, nul, nul, nul, nul) == _capi.EINVAL
        assert T.lib().tridiag_solve_tile_iface(prec, t["index"], 2 | 1 << 4, layout, t["cap"], 1,
                                                nul, nul, nul, nul, nul) == _capi.EINVAL
    assert not fails, fails[:5]
