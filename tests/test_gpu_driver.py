"""The command-line driver (reference test program interface) on the GPU: each run
calls a reference-compatible entry point through the solver table and appends a CSV
line whose error columns compare every system with the manufactured solution."""
import os
import subprocess

import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tridiagonalsolver_amd", "lib", "tridiag_test")
DRIVER_SRC = os.path.join(ROOT, "tridiagonalsolver_amd", "driver", "tridiag_test.cpp")
LIBRARY = os.path.join(ROOT, "tridiagonalsolver_amd", "lib", "libtridiag_amd.so")


def _driver_stale():
    """True when the built driver predates its source or the library it runs
    against (a stale binary would test an old build)."""
    t = os.path.getmtime(DRIVER)
    return any(os.path.exists(p) and os.path.getmtime(p) > t for p in (DRIVER_SRC, LIBRARY))


@pytest.mark.parametrize("prec,algo,n,batch,tol", [
    (0, 9, 1024, 512, 1e-4), (0, 1, 300, 100, 1e-4), (0, 14, 512, 64, 1e-4),
    (0, 0, 256, 64, 1e-4), (0, 11, 4096, 16, 1e-4), (1, 9, 4096, 64, 1e-10),
    (1, 3, 1000, 33, 1e-10), (1, 2, 64, 1000, 1e-10), (1, 9, 70000, 3, 1e-10)])
def test_driver_run(tmp_path, prec, algo, n, batch, tol):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    if _driver_stale():
        pytest.skip("driver binary is older than its source or the library: rebuild it")
    csv = tmp_path / "res.csv"
    announce("driver", prec, algo, n, batch)
    r = subprocess.run([DRIVER, str(prec), str(algo), str(n), str(batch), "3", str(csv)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failed" not in r.stderr, r.stderr
    f = [v.strip() for v in csv.read_text().strip().split(",")]
    assert int(f[0]) == n and int(f[1]) == batch
    mean, rmse, mx = float(f[7]), float(f[8]), float(f[9])
    assert mx <= tol and rmse <= mx and mean <= mx, f
    assert float(f[5]) >= 1.0 and float(f[6]) >= 1.0          # condition numbers
