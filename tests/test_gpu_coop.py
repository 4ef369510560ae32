"""GPU parity of the cooperative row-block solve (K workgroups per group of
systems, block forms exchanged through global memory), through the C-ABI,
against the FP64 oracle: normwise relative error <= 1e-5 (FP32) / 1e-12 (FP64)
over every system.  Interleaved systems take it from 1024 rows when 256-row
blocks divide n (and, up to the 4096 rows the interleaved tiles hold, when rows
start on 128-byte lines), contiguous systems above one workgroup's rows in up
to 32 blocks; every case asserts the plan it tests.  A
forced zero spin limit (run in a child process) checks that a wait that gives up
ends the kernel and marks its block with NaN, never with a wrong finite value."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from gpu_util import TOL, announce, check_against_oracle, make_case, run_solve

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tridiagonalsolver_amd as T
    T.lib()


def _assert_coop(prec, layout, n, batch, K):
    import tridiagonalsolver_amd as T
    p = T.plan(prec, -1, layout, n, batch)
    assert p["kind"] == "cooperative" and p["K"] == K, p


@pytest.mark.parametrize("dtype,n,batch,K", [
    (np.float32, 1024, 64, 4), (np.float32, 2048, 96, 8), (np.float32, 2560, 160, 10),
    (np.float32, 4096, 64, 16), (np.float32, 4096, 4096, 16), (np.float32, 6144, 31, 24),
    (np.float32, 8192, 40, 32), (np.float32, 12288, 64, 48), (np.float64, 1280, 48, 5),
    (np.float64, 4096, 48, 16), (np.float64, 4096, 4096, 16), (np.float64, 5120, 40, 20),
    (np.float64, 8192, 33, 32), (np.float64, 16384, 21, 64)])
def test_interleaved(dtype, n, batch, K):
    _assert_coop(0 if dtype == np.float32 else 1, 2, n, batch, K)
    announce("coop interleaved", np.dtype(dtype).name, n, batch)
    arrs = make_case(n, batch, 2, dtype, seed=n + batch)
    x = run_solve(arrs, n, batch, 2, -1, dtype)
    st, res = check_against_oracle(arrs, x, n, batch, 2, dtype)
    assert st["relerr"] <= TOL[dtype], st
    assert res <= 10 * TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_interleaved_low_margin(dtype):
    n, batch = 4096, 48
    announce("coop interleaved low margin", np.dtype(dtype).name, n, batch)
    arrs = make_case(n, batch, 2, dtype, seed=3, blo=0.05, bhi=0.2)
    x = run_solve(arrs, n, batch, 2, -1, dtype)
    st, _ = check_against_oracle(arrs, x, n, batch, 2, dtype)
    assert st["relerr"] <= TOL[dtype], st


@pytest.mark.parametrize("dtype,n,batch,K", [
    (np.float32, 20000, 7, 5), (np.float32, 65536, 12, 16), (np.float32, 8193 * 3, 5, 7),
    (np.float32, 131072, 3, 32), (np.float32, 4096 * 31 + 17, 4, 32),
    (np.float64, 20000, 7, 10), (np.float64, 32768, 9, 16), (np.float64, 8193 * 3, 5, 13),
    (np.float64, 65536, 5, 32), (np.float64, 40000, 6, 20)])
def test_contiguous(dtype, n, batch, K):
    _assert_coop(0 if dtype == np.float32 else 1, 0, n, batch, K)
    announce("coop contiguous", np.dtype(dtype).name, n, batch)
    arrs = make_case(n, batch, 0, dtype, seed=n % 1000 + batch, blo=0.05, bhi=0.2)
    x = run_solve(arrs, n, batch, 0, -1, dtype)
    st, _ = check_against_oracle(arrs, x, n, batch, 0, dtype)
    assert st["relerr"] <= TOL[dtype], st


def test_plan_boundaries():
    """Where the planner switches to and from the cooperative solve."""
    import tridiagonalsolver_amd as T
    assert T.plan(0, -1, 2, 768, 64)["kind"] == "tile"
    assert T.plan(1, -1, 2, 768, 64)["kind"] == "tile"
    assert T.plan(0, -1, 2, 1024, 64)["kind"] == "cooperative"
    assert T.plan(1, -1, 2, 1024, 64)["kind"] == "cooperative"
    assert T.plan(1, -1, 2, 2048, 64)["kind"] == "cooperative"
    # rows off 128-byte lines keep the tile where one holds n, not beyond
    assert T.plan(0, -1, 2, 4096, 4001)["kind"] == "tile"
    assert T.plan(1, -1, 2, 2048, 17)["kind"] == "tile"
    assert T.plan(0, -1, 2, 8192, 4001)["kind"] == "cooperative"
    assert T.plan(0, -1, 2, 2304, 64)["kind"] == "cooperative"
    assert T.plan(0, -1, 2, 2100, 64)["kind"] == "tile"             # blocks must divide n
    assert T.plan(0, -1, 2, 8448, 64)["kind"] == "cooperative"      # 33 blocks
    assert T.plan(1, -1, 2, 16384, 64)["kind"] == "cooperative"     # 64 blocks
    assert T.plan(0, -1, 2, 16640, 64)["kind"] == "via_contiguous"  # 65 blocks
    assert T.plan(1, -1, 0, 65536, 8)["kind"] == "cooperative"      # 32 blocks
    assert T.plan(1, -1, 0, 65537, 8)["kind"] == "chunked"          # 33 blocks
    assert T.plan(0, -1, 0, 131072, 8)["kind"] == "cooperative"
    assert T.plan(0, -1, 0, 131073, 8)["kind"] == "chunked"
    assert T.plan(0, 5, 0, 20000, 8)["kind"] == "chunked"           # PCR interface family
    assert T.plan(0, 5, 2, 4096, 8)["kind"] == "tile"


@pytest.mark.parametrize("algo", [9, 7])
def test_wang_mou_ids_take_it(algo):
    """reg and shmem WM-pGE both use the Wang-Mou interface family."""
    n, batch = 4096, 64
    arrs = make_case(n, batch, 2, np.float32, seed=algo)
    x = run_solve(arrs, n, batch, 2, algo, np.float32)
    st, _ = check_against_oracle(arrs, x, n, batch, 2, np.float32)
    assert st["relerr"] <= TOL[np.float32], st


def test_replicated_systems_identical():
    """Every copy of one system solves to the same bits, whichever workgroups took
    its blocks."""
    n, batch = 4096, 256
    a, b, c, d = oracle.gen_dd(n, 1, layout=0, seed=17)
    arrs = [np.ascontiguousarray(np.repeat(v.reshape(n, 1), batch, axis=1).ravel())
            .astype(np.float32) for v in (a, b, c, d)]
    x = run_solve(arrs, n, batch, 2, -1, np.float32).reshape(n, batch)
    assert (x == x[:, :1]).all()


_GIVE_UP = r"""
import sys, numpy as np, torch
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(root)r + '/tests')
import oracle
from gpu_util import make_case, run_solve
n, batch = 4096, 512
arrs = make_case(n, batch, 2, np.float32, seed=5)
x = run_solve(arrs, n, batch, 2, -1, np.float32).astype(np.float64).reshape(n, batch)
xr = oracle.thomas(*(v.astype(np.float64) for v in arrs), n, batch, layout=2).reshape(n, batch)
nan = np.isnan(x)
err = np.abs(np.where(nan, 0.0, x - xr)).max() / np.abs(xr).max()
print('nan_fraction', nan.mean(), 'finite_relerr', err, flush=True)
assert err <= 1e-5, err
"""


def test_wait_that_gives_up_marks_nan():
    """TRIDIAG_COOP_SPIN=0: every wait that does not find all blocks published at
    its first poll gives up.  The kernel must end, and each row of x is either
    the right value or NaN."""
    env = dict(os.environ, TRIDIAG_COOP_SPIN="0")
    announce("coop give-up (child process)")
    r = subprocess.run([sys.executable, "-c", _GIVE_UP % {"root": ROOT}], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "nan_fraction" in r.stdout


_FORCED_KERNEL = r"""
import sys, numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(root)r + '/tests')
import tridiagonalsolver_amd as T
from gpu_util import TOL, check_against_oracle, make_case, run_solve
for dtype, layout, n, batch in ((np.float32, 2, 4096, 96), (np.float64, 2, 4096, 48),
                                (np.float32, 2, 8192, 40), (np.float64, 2, 8192, 17),
                                (np.float32, 0, 20000, 7), (np.float64, 0, 20001, 5),
                                (np.float32, 0, 65536, 3),
                                # several rounds of the persistent grid
                                (np.float32, 2, 4096, 4096), (np.float32, 0, 32768, 200),
                                (np.float64, 0, 16384, 300)):
    prec = 0 if dtype == np.float32 else 1
    assert T.plan(prec, -1, layout, n, batch)["kind"] == "cooperative"
    arrs = make_case(n, batch, layout, dtype, seed=n + batch)
    x = run_solve(arrs, n, batch, layout, -1, dtype)
    st, res = check_against_oracle(arrs, x, n, batch, layout, dtype)
    print(np.dtype(dtype).name, layout, n, batch, st["relerr"], flush=True)
    assert st["nonfinite"] == 0 and st["relerr"] <= TOL[dtype], st
print("ok", flush=True)
"""


@pytest.mark.parametrize("pipe", ["0", "1"])
def test_both_cooperative_kernels(pipe):
    """TRIDIAG_COOP_PIPE forces the kernel that waits one task later (1) or the one
    that waits before its next task (0) for every table entry, whichever each
    entry runs by default (child process: the setting is read once)."""
    env = dict(os.environ, TRIDIAG_COOP_PIPE=pipe)
    announce("coop forced kernel (child process)", pipe)
    r = subprocess.run([sys.executable, "-c", _FORCED_KERNEL % {"root": ROOT}], env=env,
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


_NO_SPREAD = r"""
import sys, numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(root)r + '/tests')
from gpu_util import make_case, run_solve
out = {}
for i, (dtype, n, batch) in enumerate(%(cases)r):
    dt = np.dtype(dtype).type
    arrs = make_case(n, batch, 2, dt, seed=i + 11)
    out["x%%d" %% i] = run_solve(arrs, n, batch, 2, -1, dt)
np.savez(%(path)r, **out)
print("ok", flush=True)
"""
_SPREAD_CASES = [("float32", 4096, 4096), ("float64", 4096, 1008), ("float32", 2048, 3008),
                 ("float64", 2048, 16), ("float32", 256, 65536), ("float64", 512, 8192)]


def test_group_spread_same_bits(tmp_path):
    """The cooperative schedule spreads the groups of one round over the whole batch
    (TRIDIAG_COOP_SPREAD), and the interleaved tiles of whole 128-byte pieces their
    resident tiles (TRIDIAG_TILE_SPREAD); both default on.  Each group's and tile's
    arithmetic is unchanged, so x must be bit-identical to the run in dispatch
    order (child process: the settings are read once).  Cases with one partial
    table column, with a single round, and two tile cases."""
    import tridiagonalsolver_amd as T
    path = str(tmp_path / "nospread.npz")
    env = dict(os.environ, TRIDIAG_COOP_SPREAD="0", TRIDIAG_TILE_SPREAD="0")
    announce("coop spread off (child process)")
    r = subprocess.run([sys.executable, "-c", _NO_SPREAD % {"root": ROOT, "cases": _SPREAD_CASES,
                                                           "path": path}],
                       env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
    ref = np.load(path)
    for i, (dtype, n, batch) in enumerate(_SPREAD_CASES):
        dt = np.dtype(dtype).type
        kind = "cooperative" if n >= 1024 else "tile"
        assert T.plan(0 if dt == np.float32 else 1, -1, 2, n, batch)["kind"] == kind
        announce("coop spread on", dtype, n, batch)
        arrs = make_case(n, batch, 2, dt, seed=i + 11)
        x = run_solve(arrs, n, batch, 2, -1, dt)
        assert np.array_equal(x, ref["x%d" % i]), (dtype, n, batch)
        st, _ = check_against_oracle(arrs, x, n, batch, 2, dt)
        assert st["relerr"] <= TOL[dt], st
