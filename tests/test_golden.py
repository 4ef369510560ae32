"""Known-answer fixtures (tests/golden, written by tests/golden/make_fixtures.py) on
the CPU: the files hold what their generator makes, and the oracle recovers each
stored x (FP64 manufactured and dense-LU answers to 1e-12 relative; the FP32
manufactured system, whose d was formed in FP32 as in the reference's
single-precision run, to 1e-6; the Poisson system, whose condition number grows
like n^2, to 1e-9)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {"manufactured_f64_n1024.npz": 1e-12, "manufactured_f32_n1024.npz": 1e-6,
       "poisson_f64_n1024.npz": 1e-9, "dense_dd_f64_n64_b8.npz": 1e-12}


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(TOL))
def test_oracle_recovers_stored_x(oracle_mod, name):
    f = load(name)
    a, b, c, d, x = (np.atleast_2d(f[k]) for k in "abcdx")
    batch, n = x.shape
    xo = oracle_mod.thomas(a.ravel(), b.ravel(), c.ravel(), d.ravel(), n, batch)
    err = np.max(np.abs(xo - x.ravel())) / np.max(np.abs(x))
    assert err <= TOL[name], err


def test_fixtures_match_their_generator(oracle_mod):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(GOLDEN, "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    for name, make in mf.FIXTURES.items():
        f, g = load(name), make()
        assert set(f) == set(g) == set("abcdx")
        for k in "abcdx":
            assert np.array_equal(f[k], g[k]), (name, k)


def test_manufactured_single_is_float_exact():
    f = load("manufactured_f32_n1024.npz")
    for k in "abcdx":
        assert np.array_equal(f[k].astype(np.float32).astype(np.float64), f[k]), k
