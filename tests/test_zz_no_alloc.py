"""tridiag_solve_noalloc (the entry point that allocates nothing) against
tridiag_solve through the C-ABI: bit-identical results, and within 1e-5 (FP32) /
1e-12 (FP64) of the FP64 oracle, on two calls in a row (the second smaller) of
each path whose kernels use scratch memory (the cooperative row-block solve and
the chunked solve), and on on-chip tile paths."""
import ctypes

import numpy as np
import pytest

import oracle
from gpu_util import TOL, announce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tridiagonalsolver_amd import _capi
    return _capi


def _run(capi, fn, prec, algo, layout, n, batch, arrs):
    import torch
    dev = torch.device("cuda:0")
    ts = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in arrs]
    rc = fn(prec, algo, layout, n, batch, *(ctypes.c_void_p(t.data_ptr()) for t in ts),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert rc == capi.OK, capi.lib().tridiag_last_error()
    return ts[3].cpu().numpy()


@pytest.mark.parametrize("prec,layout,algo,shapes", [
    (1, 0, -1, [(24581, 40), (9000, 5)]),     # FP64 contiguous, rows beyond one workgroup
    (0, 0, -1, [(40000, 24), (16400, 3)]),    # FP32 contiguous
    (0, 2, -1, [(4096, 640), (2048, 40)]),    # FP32 interleaved, cooperative blocks
    (1, 2, -1, [(4096, 320), (1024, 9)]),     # FP64 interleaved
    (0, 2, -1, [(4096, 2048), (2048, 3008)]),  # groups of a round spread over the batch
    (0, 2, -1, [(16384, 96), (12288, 40)]),   # FP32 interleaved, 64 and 48 blocks
    (1, 2, -1, [(16384, 48), (9216, 20)]),    # FP64 interleaved, 64 and 36 blocks
    (1, 0, 5, [(24581, 40), (9000, 5)]),      # PCR interface: the chunked solve
    (1, 0, -1, [(70000, 6), (66000, 3)]),     # FP64 beyond 32 blocks: the chunked solve
])
def test_scratch_paths_two_calls(lib, prec, layout, algo, shapes):
    dtype = np.float32 if prec == 0 else np.float64
    L = lib.lib()
    for k, (n, batch) in enumerate(shapes):
        announce("noalloc scratch path", prec, layout, algo, n, batch)
        arrs = [v.astype(dtype) for v in oracle.gen_dd(n, batch, layout=layout, seed=50 + k,
                                                       blo=0.05, bhi=0.2)]
        x_na = _run(lib, L.tridiag_solve_noalloc, prec, algo, layout, n, batch, arrs)
        x_ref = _run(lib, L.tridiag_solve, prec, algo, layout, n, batch, arrs)
        assert np.array_equal(x_na, x_ref)
        xo = oracle.thomas(*(v.astype(np.float64) for v in arrs), n, batch, layout=layout)
        err = oracle.errstats(x_na.astype(np.float64), xo, n, batch, layout)["relerr"]
        assert err <= TOL[dtype], err


@pytest.mark.parametrize("prec,layout,n,batch", [(0, 2, 1000, 300), (1, 0, 4096, 64),
                                                 (0, 0, 20, 77)])
def test_tile_paths_match(lib, prec, layout, n, batch):
    dtype = np.float32 if prec == 0 else np.float64
    L = lib.lib()
    announce("noalloc tile", prec, layout, n, batch)
    arrs = [v.astype(dtype) for v in oracle.gen_dd(n, batch, layout=layout, seed=9)]
    x_na = _run(lib, L.tridiag_solve_noalloc, prec, -1, layout, n, batch, arrs)
    x_ref = _run(lib, L.tridiag_solve, prec, -1, layout, n, batch, arrs)
    assert np.array_equal(x_na, x_ref)


def test_allocating_paths_refused(lib):
    import torch
    L = lib.lib()
    t = torch.zeros(5000 * 4, device="cuda")
    p = [ctypes.c_void_p(t.data_ptr())] * 4
    # interleaved systems beyond the interleaved tiles that the cooperative solve does
    # not take (5000 rows: not a multiple of its blocks) go through a full copy
    assert L.tridiag_solve_noalloc(0, -1, 2, 5000, 4, *p, None) == lib.EUNSUPPORTED
    assert b"allocation" in L.tridiag_last_error()
    # the vendor comparator allocates its own buffer
    assert L.tridiag_solve_noalloc(0, 0, 0, 5000, 4, *p, None) == lib.EUNSUPPORTED
