"""GPU parity of Thomas (algo 1) in the contiguous and packed layouts, where the
lanes of a group share one system and load its rows several steps ahead: lengths
with partial last vectors and steps, odd lengths (element path), lengths at the
step and prefetch-ring boundaries, batches that leave groups idle, against the FP64
oracle: normwise relative error <= 1e-5 / 1e-12.  c holds c' afterwards, as in the
reference (Tridiagonal_CUDA/tridiagonal_basic.cu:4-30)."""
import numpy as np
import pytest

from gpu_util import TOL, announce, check_against_oracle, make_case, run_solve

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("layout", [0, 1])
def test_thomas_chunked_rows(prec, layout):
    # 16-row (FP32) / 8-row (FP64) chunks: exact multiples, partial last chunks, and
    # odd n (rows not 16-byte aligned: element accesses)
    sizes = [34, 36, 40, 48, 100, 257, 1000, 1024, 1030, 4100]
    for n in sizes:
        for batch in (1, 63, 130):
            announce("thomas rows", prec, layout, n, batch)
            arrs = make_case(n, batch, layout, DT[prec], seed=n * 31 + batch)
            x = run_solve(arrs, n, batch, layout, 1, DT[prec])
            st, res = check_against_oracle(arrs, x, n, batch, layout, DT[prec])
            assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], (n, batch, st)


@pytest.mark.parametrize("prec", [0, 1])
def test_thomas_low_margin_long(prec):
    n, batch = 8192, 300
    announce("thomas rows low margin", prec)
    arrs = make_case(n, batch, 0, DT[prec], seed=77, blo=0.05, bhi=0.2)
    x = run_solve(arrs, n, batch, 0, 1, DT[prec])
    st, res = check_against_oracle(arrs, x, n, batch, 0, DT[prec])
    assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], st


# rows per step: 64 lanes x one or two 16-byte vectors (the scan kernel) and 16 lanes x
# one vector (the team kernel); prefetch rings of 2-4 steps (forward) and 4-8 steps
# (backward)
_STEP = {0: (512, 256, 64), 1: (256, 128, 32)}


@pytest.mark.parametrize("prec", [0, 1])
def test_thomas_step_boundaries(prec):
    sizes = {17, 33}
    for sr in _STEP[prec]:
        sizes |= {sr - 1, sr, sr + 1, 2 * sr + 3, 4 * sr - 1, 4 * sr, 4 * sr + 1,
                  8 * sr - 1, 8 * sr, 8 * sr + 1, 13 * sr + 5}
    sizes = sorted(sizes)
    for n in sizes:
        for batch in (1, 9, 130):
            announce("thomas steps", prec, n, batch)
            arrs = make_case(n, batch, 0, DT[prec], seed=n * 7 + batch)
            x = run_solve(arrs, n, batch, 0, 1, DT[prec])
            st, res = check_against_oracle(arrs, x, n, batch, 0, DT[prec])
            assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], (n, batch, st)


@pytest.mark.parametrize("prec", [0, 1])
def test_thomas_leaves_c_prime_in_c(prec):
    import torch
    import tridiagonalsolver_amd as T
    n, batch = 300, 20
    announce("thomas c prime", prec, n, batch)
    a, b, c, d = make_case(n, batch, 0, DT[prec], seed=5)
    ts = [torch.from_numpy(v.copy()).cuda() for v in (a, b, c, d)]
    T.solve(*ts, algo=1, layout=0, n=n, batch=batch)
    torch.cuda.synchronize()
    cg = ts[2].cpu().numpy().reshape(batch, n).astype(np.float64)
    A, B, C = (v.reshape(batch, n).astype(np.float64) for v in (a, b, c))
    cp = np.zeros((batch, n))
    cp[:, 0] = C[:, 0] / B[:, 0]
    for i in range(1, n):
        cp[:, i] = C[:, i] / (B[:, i] - A[:, i] * cp[:, i - 1])
    err = np.max(np.abs(cg[:, :-1] - cp[:, :-1])) / np.max(np.abs(cp[:, :-1]))
    assert err <= TOL[DT[prec]], err


@pytest.mark.parametrize("prec", [0, 1])
def test_thomas_row_scaling(prec):
    # rows scaled by powers of two spread over 2^-40..2^40 (FP32) / 2^-300..2^300
    # (FP64): the solution does not change, while products of a few rows'
    # coefficients leave the format's range unless the composed maps are rescaled
    n, batch = 3000, 17
    announce("thomas row scaling", prec, n, batch)
    a, b, c, d = make_case(n, batch, 0, DT[prec], seed=41)
    rng = np.random.default_rng(9)
    span = 40 if prec == 0 else 300
    s = np.ldexp(1.0, rng.integers(-span, span + 1, size=n * batch)).astype(DT[prec])
    arrs = [v * s for v in (a, b, c, d)]
    x = run_solve(arrs, n, batch, 0, 1, DT[prec])
    xs = run_solve([a, b, c, d], n, batch, 0, 1, DT[prec])
    st, res = check_against_oracle([a, b, c, d], x, n, batch, 0, DT[prec])
    assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], st
    err = np.max(np.abs(x.astype(np.float64) - xs)) / np.max(np.abs(xs))
    assert err <= TOL[DT[prec]], err


@pytest.mark.parametrize("prec", [0, 1])
def test_thomas_both_step_widths(prec):
    # batches up to 2048 take two 16-byte vectors per lane and step, larger ones one
    for n, batch in ((1000, 2048), (1000, 2049), (4099, 2050), (257, 3000)):
        announce("thomas step width", prec, n, batch)
        arrs = make_case(n, batch, 0, DT[prec], seed=n + batch)
        x = run_solve(arrs, n, batch, 0, 1, DT[prec])
        st, res = check_against_oracle(arrs, x, n, batch, 0, DT[prec])
        assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], (n, batch, st)
