"""GPU parity of the division-policy variants (include/tridiagonal.h TRIDIAG_DIV_*):
every variant library, loaded in a child process through $TRIDIAG_DIV_MODE, solves
the cases of div_mode_cases.py (lane kernels, Thomas, tiles of each interface
family, cooperative, chunked, ragged, fused forward + adjoint) within the solver's
tolerance against the FP64 oracle: normwise relative error <= 1e-5 (FP32) /
1e-12 (FP64).  The reference's division macros (Tridiagonal_CUDA/intrin.h:23-87)
have no test of their own; this holds every policy to the same bar."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TOL = {0: 1e-5, 1: 1e-12}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("mode", ["refined", "ieee", "fast"])
def test_division_policy(mode):
    from tridiagonalsolver_amd import build
    path = os.path.join(ROOT, "tridiagonalsolver_amd", "lib", build.lib_name(mode))
    assert os.path.exists(path), "%s is not built" % path
    print("[gpu-case] division policy", mode, flush=True)
    r = subprocess.run([sys.executable, os.path.join(HERE, "div_mode_cases.py")], cwd=ROOT,
                       env=dict(os.environ, TRIDIAG_DIV_MODE=mode), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    rows = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    assert len(rows) >= 25
    for row in rows:
        assert row["mode"] == mode, row
        prec = row.get("prec", 0)
        assert row.get("nonfinite", 0) == 0 and row["relerr"] <= TOL[prec], row
