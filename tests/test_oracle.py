"""Pins the FP64 oracle (oracle/oracle.c) before any GPU result is compared to it.

Every check is independent of the oracle's own formula: closed forms, dense LU
(numpy.linalg.solve), a brute-force pure-Python SplitMix64, and planted
mistakes that the checks must catch.
"""
import numpy as np
import pytest


def dense(a, b, c, n):
    """Dense matrix of one system; a_0 and c_{n-1} are not part of A."""
    A = np.diag(np.asarray(b, dtype=np.float64))
    for i in range(1, n):
        A[i, i - 1] = a[i]
        A[i - 1, i] = c[i - 1]
    return A


def per_system(arr, n, batch, layout):
    arr = np.asarray(arr)
    if layout == 2:
        return arr.reshape(n, batch).T
    return arr.reshape(batch, n)


# ------------------------------------------------------------ closed forms

@pytest.mark.parametrize("n", [2, 3, 64, 1000])
def test_poisson_x_is_one(oracle_mod, n):
    # BPLG known-answer init (BPLGTridiagWM.hxx:28-54): tridiag(-1,2,-1), d_0=d_{n-1}=1.
    a = -np.ones(n); b = 2 * np.ones(n); c = -np.ones(n); d = np.zeros(n)
    a[0] = 0; c[-1] = 0; d[0] = 1; d[-1] = 1
    x = oracle_mod.thomas(a, b, c, d, n, 1)
    assert np.max(np.abs(x - 1.0)) < 1e-10 * n


def test_n1_is_division(oracle_mod):
    rng = np.random.default_rng(0)
    b = rng.uniform(1, 2, 7); d = rng.uniform(-1, 1, 7)
    z = np.zeros(7)
    x = oracle_mod.thomas(z + 5.0, b, z - 3.0, d, 1, 7)  # a_0, c_{n-1} ignored
    assert np.array_equal(x, d / b)


def test_n2_cramer(oracle_mod):
    rng = np.random.default_rng(1)
    for _ in range(20):
        a0, a1, c0, c1 = rng.uniform(-1, 1, 4)
        b0, b1 = rng.uniform(2, 3, 2)
        d0, d1 = rng.uniform(-1, 1, 2)
        x = oracle_mod.thomas([a0, a1], [b0, b1], [c0, c1], [d0, d1], 2, 1)
        det = b0 * b1 - c0 * a1
        assert x[0] == pytest.approx((d0 * b1 - c0 * d1) / det, rel=1e-14, abs=1e-15)
        assert x[1] == pytest.approx((b0 * d1 - a1 * d0) / det, rel=1e-14, abs=1e-15)


def test_diagonal_matrix(oracle_mod):
    rng = np.random.default_rng(2)
    n = 50
    b = rng.uniform(1, 2, n); d = rng.uniform(-1, 1, n)
    x = oracle_mod.thomas(np.zeros(n), b, np.zeros(n), d, n, 1)
    assert np.allclose(x, d / b, rtol=1e-15, atol=0)


def test_block_diagonal_splits(oracle_mod):
    rng = np.random.default_rng(3)
    n, k = 40, 17
    a, b, c, d = oracle_mod.gen_dd(n, 1, seed=9)
    c = c.copy(); a = a.copy()
    c[k - 1] = 0.0; a[k] = 0.0      # rows < k and rows >= k decouple
    x = oracle_mod.thomas(a, b, c, d, n, 1)
    a2 = a[k:].copy(); a2[0] = 0.0
    x1 = oracle_mod.thomas(a[:k], b[:k], c[:k], d[:k], k, 1)
    x2 = oracle_mod.thomas(a2, b[k:], c[k:], d[k:], n - k, 1)
    assert np.allclose(x, np.concatenate([x1, x2]), rtol=1e-14, atol=1e-15)


def test_ignores_a0_and_cn1(oracle_mod):
    n = 33
    a, b, c, d = oracle_mod.gen_dd(n, 2, seed=4)
    x = oracle_mod.thomas(a, b, c, d, n, 2)
    a2, c2 = a.copy(), c.copy()
    a2[0] = 1e30; a2[n] = -7.0; c2[n - 1] = 1e30; c2[2 * n - 1] = 3.0
    x2 = oracle_mod.thomas(a2, b, c2, d, n, 2)
    assert np.array_equal(x, x2)


# ------------------------------------------------------------ dense LU

@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("n", [1, 2, 5, 31, 64])
def test_dense_lu_all_layouts(oracle_mod, n, layout):
    batch = 3
    a, b, c, d = oracle_mod.gen_dd(n, batch, layout=layout, seed=100 + n)
    x = oracle_mod.thomas(a, b, c, d, n, batch, layout=layout)
    if layout == 1:
        blk = a.reshape(batch, 4, n)
        xs = x.reshape(batch, 4, n)[:, 3, :]
        for s in range(batch):
            A = dense(blk[s, 0], blk[s, 1], blk[s, 2], n)
            ref = np.linalg.solve(A, blk[s, 3])
            assert np.max(np.abs(xs[s] - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
        return
    A_, B_, C_, D_, X_ = (per_system(v, n, batch, layout) for v in (a, b, c, d, x))
    for s in range(batch):
        A = dense(A_[s], B_[s], C_[s], n)
        ref = np.linalg.solve(A, D_[s])
        assert np.max(np.abs(X_[s] - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_manufactured_solution_recovered(oracle_mod):
    n = 4096
    a, b, c, d, x = oracle_mod.gen_manufactured(n)
    xs = oracle_mod.thomas(a, b, c, d, n, 1)
    assert np.max(np.abs(xs - x)) < 1e-13
    # generator invariant of test.cpp:19-38: strict row and column dominance
    A = np.abs(a[1:]); C = np.abs(c[:-1])
    off_row = np.zeros(n); off_row[1:] += A; off_row[:-1] += C
    assert np.all(b > off_row)


def test_manufactured_single_recipe(oracle_mod):
    n = 257
    a, b, c, d, x = oracle_mod.gen_manufactured(n, single=True)
    for v in (a, b, c, d, x):
        assert np.array_equal(v, v.astype(np.float32).astype(np.float64))
    # d is A x formed in float: close to the FP64 product
    A = dense(a, b, c, n)
    assert np.max(np.abs(A @ x - d)) < 1e-5


# ------------------------------------------------------------ planted mistakes

def _py_thomas(a, b, c, d, mistake=None):
    """Pure-Python Thomas used only to show that the dense check catches a
    plausible mistake (the mistake is planted here, on the CPU)."""
    n = len(b)
    cp = [0.0] * n; dp = [0.0] * n
    cp[0] = c[0] / b[0]; dp[0] = d[0] / b[0]
    for i in range(1, n):
        ai = -a[i] if mistake == "sign" else a[i]
        cprev = cp[i] if mistake == "index" else cp[i - 1]
        k = b[i] - cprev * ai
        cp[i] = (a[i] if mistake == "transpose" else c[i]) / k
        dp[i] = (d[i] - dp[i - 1] * ai) / k
    x = dp[:]
    for i in range(n - 2, -1, -1):
        x[i] = dp[i] - cp[i] * x[i + 1]
    return np.array(x)


@pytest.mark.parametrize("mistake", [None, "sign", "index", "transpose"])
def test_dense_check_catches_planted_mistakes(oracle_mod, mistake):
    n = 24
    a, b, c, d = oracle_mod.gen_dd(n, 1, seed=77)
    ref = np.linalg.solve(dense(a, b, c, n), d)
    x = _py_thomas(list(a), list(b), list(c), list(d), mistake)
    ok = np.max(np.abs(x - ref)) <= 1e-13
    assert ok == (mistake is None)
    if mistake is None:
        assert np.max(np.abs(x - oracle_mod.thomas(a, b, c, d, n, 1))) <= 1e-14


# ------------------------------------------------------------ residual / errstats

def test_residual_matches_dense(oracle_mod):
    n, batch = 40, 4
    a, b, c, d = oracle_mod.gen_dd(n, batch, seed=5)
    x = oracle_mod.thomas(a, b, c, d, n, batch)
    assert oracle_mod.residual(a, b, c, d, x, n, batch) < 1e-15
    xp = x.copy(); xp[n + 7] += 1e-3       # perturb system 1
    r = oracle_mod.residual(a, b, c, d, xp, n, batch)
    A = dense(a[n:2 * n], b[n:2 * n], c[n:2 * n], n)
    xs = xp[n:2 * n]; ds = d[n:2 * n]
    expect = np.max(np.abs(A @ xs - ds)) / (np.max(np.sum(np.abs(A), axis=1)) * np.max(np.abs(xs))
                                            + np.max(np.abs(ds)))
    assert r == pytest.approx(expect, rel=1e-10)


def test_errstats_known_values(oracle_mod):
    n, batch = 4, 2
    xref = np.array([1.0, -2.0, 4.0, 1.0, 0.5, 0.5, 0.5, -1.0])
    x = xref.copy(); x[2] += 0.4; x[5] -= 0.1
    st = oracle_mod.errstats(x, xref, n, batch)
    assert st["max"] == pytest.approx(0.4)
    assert st["argmax"] == 2
    assert st["mean"] == pytest.approx(0.5 / 8)
    assert st["rmse"] == pytest.approx(np.sqrt((0.16 + 0.01) / 8))
    assert st["relerr"] == pytest.approx(max(0.4 / 4.0, 0.1 / 1.0))
    x[6] = np.nan
    st = oracle_mod.errstats(x, xref, n, batch)
    assert np.isnan(st["max"]) and st["argmax"] == 6 and st["nonfinite"] == 1
    assert np.isnan(st["relerr"])


def test_errstats_interleaved_matches_contiguous(oracle_mod):
    n, batch = 9, 5
    rng = np.random.default_rng(11)
    xr = rng.standard_normal(n * batch); x = xr + 1e-3 * rng.standard_normal(n * batch)
    s0 = oracle_mod.errstats(x, xr, n, batch, 0)
    xi = x.reshape(batch, n).T.copy().ravel(); xri = xr.reshape(batch, n).T.copy().ravel()
    s2 = oracle_mod.errstats(xi, xri, n, batch, 2)
    assert s0["max"] == s2["max"] and s0["relerr"] == pytest.approx(s2["relerr"])


# ------------------------------------------------------------ generators

M64 = (1 << 64) - 1


def _sm64(state, count):
    out = []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_bits(oracle_mod):
    # Published SplitMix64 test value: seed 0 -> first output 0xE220A8397B1DCDAF.
    assert _sm64(0, 1)[0] == 0xE220A8397B1DCDAF
    st = oracle_mod.stream_seed(1234, 5)
    assert st == (1234 ^ ((5 * 0xD1B54A32D192ED03) & M64))
    got = oracle_mod.sm64_fill(st, 16, 0.0, 1.0)
    exp = [(v >> 11) * 2.0 ** -53 for v in _sm64(st, 16)]
    assert np.array_equal(got, np.array(exp))


def test_gen_dd_draw_order(oracle_mod):
    n, batch, seed = 6, 3, 42
    a, b, c, d = oracle_mod.gen_dd(n, batch, seed=seed, blo=1.0, bhi=2.0)
    for s in range(batch):
        u = [(v >> 11) * 2.0 ** -53 for v in _sm64(oracle_mod.stream_seed(seed, s), 4 * n)]
        ea = [-1 + 2 * t for t in u[0:n]]; ec = [-1 + 2 * t for t in u[n:2 * n]]
        ea[0] = 0.0; ec[-1] = 0.0
        eb = [abs(ea[i]) + abs(ec[i]) + 1.0 + t for i, t in enumerate(u[2 * n:3 * n])]
        ed = [-1 + 2 * t for t in u[3 * n:4 * n]]
        sl = slice(s * n, (s + 1) * n)
        assert np.array_equal(a[sl], ea) and np.array_equal(c[sl], ec)
        assert np.allclose(b[sl], eb, rtol=0, atol=1e-15) and np.array_equal(d[sl], ed)


def test_gen_layouts_consistent(oracle_mod):
    n, batch = 7, 5
    g0 = oracle_mod.gen_dd(n, batch, layout=0, seed=3)
    g2 = oracle_mod.gen_dd(n, batch, layout=2, seed=3)
    g1 = oracle_mod.gen_dd(n, batch, layout=1, seed=3)[0].reshape(batch, 4, n)
    for k in range(4):
        assert np.array_equal(g0[k].reshape(batch, n), g2[k].reshape(n, batch).T)
        assert np.array_equal(g0[k].reshape(batch, n), g1[:, k, :])


# ------------------------------------------------------------ other oracle functions

def test_thomas_f32_close_to_f64(oracle_mod):
    n, batch = 512, 8
    a, b, c, d = oracle_mod.gen_dd(n, batch, seed=6)
    a32, b32, c32, d32 = (v.astype(np.float32) for v in (a, b, c, d))
    x64 = oracle_mod.thomas(a32, b32, c32, d32, n, batch)
    x32 = oracle_mod.thomas_f32(a32, b32, c32, d32, n, batch)
    assert oracle_mod.errstats(x32, x64, n, batch)["relerr"] < 1e-6


def test_ragged_matches_per_system(oracle_mod):
    sizes = oracle_mod.ragged_sizes(50, 1, 300, seed=8)
    off = oracle_mod.offsets_from_sizes(sizes)
    a, b, c, d = oracle_mod.gen_dd_ragged(off, seed=8)
    x = oracle_mod.thomas_ragged(a, b, c, d, off)
    for s in range(len(sizes)):
        sl = slice(off[s], off[s + 1])
        xs = oracle_mod.thomas(a[sl], b[sl], c[sl], d[sl], int(sizes[s]), 1)
        assert np.array_equal(x[sl], xs)
    x32 = oracle_mod.thomas_ragged_f32(a, b, c, d, off)
    assert np.max(np.abs(x32 - x)) < 1e-5


def test_ragged_skips_empty_systems(oracle_mod):
    sizes = np.array([3, 0, 0, 5, 0, 1], dtype=np.int64)
    off = oracle_mod.offsets_from_sizes(sizes)
    a, b, c, d = oracle_mod.gen_dd_ragged(off, seed=9)
    x = oracle_mod.thomas_ragged(a, b, c, d, off)
    for s in np.nonzero(sizes)[0]:
        sl = slice(off[s], off[s + 1])
        A = dense(a[sl], b[sl], c[sl], int(sizes[s]))
        assert np.allclose(x[sl], np.linalg.solve(A, d[sl]), rtol=1e-13, atol=1e-14)
    with pytest.raises(RuntimeError):
        oracle_mod.thomas_ragged(a, b, c, d, np.array([0, 4, 2, 9], dtype=np.int64))


def test_ragged_sizes_distribution(oracle_mod):
    sizes = oracle_mod.ragged_sizes(65536, 32, 8192, seed=2024)
    assert sizes.min() >= 32 and sizes.max() <= 8192
    # log-uniform: median near sqrt(32*8192) = 512
    assert 400 < np.median(sizes) < 650


@pytest.mark.parametrize("layout", [0, 2])
def test_adjoint_matches_dense(oracle_mod, layout):
    n, batch = 30, 3
    a, b, c, d = oracle_mod.gen_dd(n, batch, layout=layout, seed=12)
    g = np.random.default_rng(12).standard_normal(n * batch)
    x = oracle_mod.thomas(a, b, c, d, n, batch, layout=layout)
    ga, gb, gc, gd = oracle_mod.adjoint(a, b, c, x, g, n, batch, layout=layout)
    A_, B_, C_, X_, G_ = (per_system(v, n, batch, layout) for v in (a, b, c, x, g))
    GA, GB, GC, GD = (per_system(v, n, batch, layout) for v in (ga, gb, gc, gd))
    for s in range(batch):
        A = dense(A_[s], B_[s], C_[s], n)
        lam = np.linalg.solve(A.T, G_[s])
        assert np.allclose(GD[s], lam, rtol=1e-12, atol=1e-13)
        assert np.allclose(GB[s], -lam * X_[s], rtol=1e-12, atol=1e-13)
        assert np.allclose(GA[s][1:], -lam[1:] * X_[s][:-1], rtol=1e-12, atol=1e-13)
        assert np.allclose(GC[s][:-1], -lam[:-1] * X_[s][1:], rtol=1e-12, atol=1e-13)
        assert GA[s][0] == 0 and GC[s][-1] == 0


def test_adjoint_central_differences(oracle_mod):
    n = 16
    a, b, c, d = oracle_mod.gen_dd(n, 1, seed=13)
    w = np.random.default_rng(13).standard_normal(n)
    x = oracle_mod.thomas(a, b, c, d, n, 1)
    grads = oracle_mod.adjoint(a, b, c, x, w, n, 1)
    h = 1e-6
    for k, arr in enumerate((a, b, c, d)):
        for i in range(n):
            if (k == 0 and i == 0) or (k == 2 and i == n - 1):
                continue
            p = [v.copy() for v in (a, b, c, d)]; m = [v.copy() for v in (a, b, c, d)]
            p[k][i] += h; m[k][i] -= h
            fd = (w @ oracle_mod.thomas(*p, n, 1) - w @ oracle_mod.thomas(*m, n, 1)) / (2 * h)
            assert grads[k][i] == pytest.approx(fd, rel=1e-6, abs=1e-8)


def test_ragged_generator_stays_inside_empty_end_systems(oracle_mod):
    """Empty first and last systems: the generator writes nothing outside the rows
    (guard elements on both sides of every array stay untouched)."""
    import ctypes
    sizes = np.array([0, 4, 0, 3, 0], dtype=np.int64)
    off = oracle_mod.offsets_from_sizes(sizes)
    total = int(off[-1])
    bufs = [np.full(total + 2, 7.0) for _ in range(4)]          # guard, rows, guard
    ptrs = [b[1:].ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for b in bufs]
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    oracle_mod.lib().oracle_gen_dd_ragged(*ptrs, offp, len(sizes), 5, 1.0, 2.0, 1)
    for bf in bufs:
        assert bf[0] == 7.0 and bf[-1] == 7.0
        assert np.all(bf[1:-1] != 7.0)                           # every row was written
