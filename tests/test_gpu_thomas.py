"""GPU parity of Thomas (algo 1) in the contiguous and packed layouts, where each
lane walks its own system in 16-byte row chunks with the next chunk in flight:
lengths with partial last chunks, odd lengths (element path), batches that leave
lanes idle, against the FP64 oracle: normwise relative error <= 1e-5 / 1e-12."""
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
