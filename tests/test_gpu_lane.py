"""GPU parity of the one-system-per-lane register Thomas kernels (short systems:
n <= 32 FP32 / 16 FP64, layouts 0 and 2; used by the Thomas id and the automatic
choice), against the FP64 oracle: normwise relative error <= 1e-5 / 1e-12."""
import numpy as np
import pytest

from gpu_util import TOL, announce, check_against_oracle, make_case, run_solve

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}
NMAX = {0: 32, 1: 16}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("layout", [0, 2])
@pytest.mark.parametrize("algo", [1, -1])
def test_every_short_size(prec, layout, algo):
    for n in range(1, NMAX[prec] + 1):
        for batch in (1, 63, 65, 300):
            announce("lane", prec, layout, algo, n, batch)
            arrs = make_case(n, batch, layout, DT[prec], seed=n * 1000 + batch)
            x = run_solve(arrs, n, batch, layout, algo, DT[prec])
            st, res = check_against_oracle(arrs, x, n, batch, layout, DT[prec])
            assert st["nonfinite"] == 0 and st["relerr"] <= TOL[DT[prec]], (n, batch, st)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("layout", [0, 2])
def test_large_batch_low_margin(prec, layout):
    n = NMAX[prec]
    arrs = make_case(n, 100000, layout, DT[prec], seed=3, blo=0.05, bhi=0.2)
    x = run_solve(arrs, n, 100000, layout, -1, DT[prec])
    st, res = check_against_oracle(arrs, x, n, 100000, layout, DT[prec])
    assert st["relerr"] <= TOL[DT[prec]], st
