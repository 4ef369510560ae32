"""GPU checks of tridiag_convert_layout: every pair of the three coefficient layouts
(0 contiguous, 1 packed, 2 interleaved), bit-exact against the same permutation
done by numpy, at ragged shapes (n and batch not multiples of the 64 x 64 tile or
of the 16-byte vector), with unaligned pointers (element-access path), skipped
coefficients, a 4096 x 4096 round trip, and a solve in the contiguous layout of
interleaved inputs compared with the FP64 oracle."""
import ctypes

import numpy as np
import pytest

from gpu_util import TOL, announce, check_against_oracle, make_case

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}
SIZES = [(1, 1), (1, 70), (70, 1), (5, 3), (5, 8), (64, 64), (100, 37), (129, 130),
         (200, 68), (4096, 3), (3, 4096), (1000, 333), (1000, 332)]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _canonical(n, batch, dtype, seed):
    """X[k, s, i]: coefficient k of system s, row i (distinct values)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((4, batch, n)).astype(dtype)


def _to_layout(X, layout):
    """numpy restatement of the three layouts (the permutation under test)."""
    if layout == 1:
        batch, n = X.shape[1], X.shape[2]
        return np.ascontiguousarray(X.transpose(1, 0, 2).reshape(batch, 4 * n))
    if layout == 2:
        return [np.ascontiguousarray(X[k].T) for k in range(4)]
    return [np.ascontiguousarray(X[k]) for k in range(4)]


def _upload(arrs, layout):
    import torch
    if layout == 1:
        return torch.from_numpy(arrs).cuda()
    return [torch.from_numpy(v).cuda() for v in arrs]


def _download(out, layout):
    import torch
    torch.cuda.synchronize()
    if layout == 1:
        return out.cpu().numpy()
    return [o.cpu().numpy() if o is not None else None for o in out]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("src", [0, 1, 2])
@pytest.mark.parametrize("dst", [0, 1, 2])
def test_every_layout_pair(prec, src, dst):
    import tridiagonalsolver_amd as T
    for n, batch in SIZES:
        announce("convert", prec, src, dst, n, batch)
        X = _canonical(n, batch, DT[prec], seed=n * 7919 + batch)
        out = T.convert_layout(_upload(_to_layout(X, src), src), src, dst, n=n, batch=batch)
        got = _download(out, dst)
        want = _to_layout(X, dst)
        if dst == 1:
            assert np.array_equal(got, want), (n, batch)
        else:
            for k in range(4):
                assert np.array_equal(got[k], want[k]), (n, batch, k)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("pair", [(0, 2), (2, 0), (0, 1), (1, 2), (2, 2)])
def test_unaligned_pointers(prec, pair):
    """Arrays that start one element past a 16-byte boundary take the element path."""
    import torch
    import tridiagonalsolver_amd as T
    src, dst = pair
    n, batch = 67, 45
    announce("convert unaligned", prec, src, dst)
    X = _canonical(n, batch, DT[prec], seed=5)
    s_arrs = _to_layout(X, src)
    w = np.dtype(DT[prec]).itemsize

    def shifted(v):
        base = torch.zeros(v.size + 1, dtype=torch.float32 if prec == 0 else torch.float64,
                           device="cuda")
        base[1:] = torch.from_numpy(v.ravel()).cuda()
        return base

    s_t = [shifted(s_arrs)] if src == 1 else [shifted(v) for v in s_arrs]
    nd = 1 if dst == 1 else 4
    per = 4 * n * batch if dst == 1 else n * batch
    d_t = [torch.full((per + 1,), -7.0, dtype=s_t[0].dtype, device="cuda") for _ in range(nd)]
    sp = [ctypes.c_void_p(t.data_ptr() + w) for t in s_t] + [None] * (4 - len(s_t))
    dp = [ctypes.c_void_p(t.data_ptr() + w) for t in d_t] + [None] * (4 - nd)
    T.check(T.lib().tridiag_convert_layout(prec, src, dst, n, batch, *sp, *dp, None))
    torch.cuda.synchronize()
    want = _to_layout(X, dst)
    if dst == 1:
        got = d_t[0].cpu().numpy()
        assert got[0] == -7.0 and np.array_equal(got[1:], want.ravel())
    else:
        for k in range(4):
            got = d_t[k].cpu().numpy()
            assert got[0] == -7.0 and np.array_equal(got[1:], want[k].ravel()), k


@pytest.mark.parametrize("prec", [0, 1])
def test_skipped_coefficients(prec):
    """Only d from interleaved to contiguous; the other destinations stay untouched."""
    import torch
    import tridiagonalsolver_amd as T
    n, batch = 300, 129
    announce("convert d only", prec)
    X = _canonical(n, batch, DT[prec], seed=11)
    il = _upload(_to_layout(X, 2), 2)
    out = T.convert_layout([None, None, None, il[3]], 2, 0)
    torch.cuda.synchronize()
    assert out[0] is None and out[1] is None and out[2] is None
    assert np.array_equal(out[3].cpu().numpy(), X[3])
    # packed source, two coefficients out
    pk = _upload(_to_layout(X, 1), 1)
    dst = [torch.full((batch, n), -1.0, dtype=pk.dtype, device="cuda") for _ in range(4)]
    P = [ctypes.c_void_p(t.data_ptr()) for t in dst]
    T.check(T.lib().tridiag_convert_layout(prec, 1, 0, n, batch, ctypes.c_void_p(pk.data_ptr()),
                                           None, None, None, None, P[1], None, P[3], None))
    torch.cuda.synchronize()
    assert np.array_equal(dst[1].cpu().numpy(), X[1])
    assert np.array_equal(dst[3].cpu().numpy(), X[3])
    assert bool((dst[0] == -1.0).all()) and bool((dst[2] == -1.0).all())


@pytest.mark.parametrize("prec", [0, 1])
def test_round_trip_4096(prec):
    import torch
    import tridiagonalsolver_amd as T
    n = batch = 4096 if prec == 0 else 2048
    announce("convert round trip", prec, n)
    dt = torch.float32 if prec == 0 else torch.float64
    g = torch.Generator(device="cuda").manual_seed(3)
    src = [torch.randn(batch, n, generator=g, device="cuda", dtype=dt) for _ in range(4)]
    il = T.convert_layout(src, 0, 2)
    pk = T.convert_layout(il, 2, 1)
    back = T.convert_layout(pk, 1, 0)
    torch.cuda.synchronize()
    for k in range(4):
        assert torch.equal(il[k], src[k].t()), k
        assert torch.equal(back[k], src[k]), k


@pytest.mark.parametrize("prec", [0, 1])
def test_solve_interleaved_inputs_in_contiguous_layout(prec):
    """Interleaved systems converted to layout 0, solved there, x converted back:
    the same bar against the FP64 oracle as a direct solve."""
    import torch
    import tridiagonalsolver_amd as T
    n, batch = 1024, 777
    announce("convert + solve", prec)
    arrs = make_case(n, batch, 2, DT[prec], seed=17)
    il = [torch.from_numpy(v).cuda() for v in arrs]
    con = T.convert_layout(il, 2, 0, n=n, batch=batch)
    T.solve(*con, layout=0)
    x_il = T.convert_layout([None, None, None, con[3]], 0, 2)[3]
    torch.cuda.synchronize()
    st, _ = check_against_oracle(arrs, x_il.cpu().numpy().ravel(), n, batch, 2, DT[prec])
    assert st["relerr"] <= TOL[DT[prec]], st
