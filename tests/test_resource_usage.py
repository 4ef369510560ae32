"""CPU check of the per-kernel resource report (the AMD counterpart of the reference's
resource-usage table), and a regression guard: the persistent ragged kernel, the
lane Thomas kernels and the layout-conversion kernels compile without scratch
(register spills), and the FP32 ragged kernel keeps two 512-lane workgroups per CU."""
import os
import shutil

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tridiagonalsolver_amd", "csrc")


@pytest.fixture(scope="module")
def ru():
    from tridiagonalsolver_amd import build, resource_usage
    if not os.path.exists(build.HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    return resource_usage


@pytest.mark.parametrize("src", ["tri_ragged.hip", "tri_lane.hip", "tri_layout.hip"])
def test_no_scratch(ru, src):
    rows = ru.report(os.path.join(CSRC, src))
    assert rows, "no kernels reported"
    for r in rows:
        assert isinstance(r.get("vgpr"), int) and isinstance(r.get("sgpr"), int), r
        assert r.get("scratch") == 0, r
        assert r.get("vspill", 0) == 0, r
    if src == "tri_ragged.hip":
        f32 = [r for r in rows if "ragged_kernel<float>" in r["name"]]
        assert f32 and f32[0]["vgpr"] <= 128, f32
    text = ru.format_rows(rows)
    assert text.splitlines()[0].startswith("VGPR")
