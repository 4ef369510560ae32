"""GPU solves of the known-answer fixtures (tests/golden): every algorithm id in its
own layout (0, or 1 for the BPLG ids) and the automatic choice in layouts 0 and 2,
each fixture system replicated over a batch as the reference's driver does
(init_batch), compared with the stored x.  Bars: the FP32 / FP64 tolerances of the
other parity tests (1e-5 / 1e-12 normwise relative), 1e-9 for the Poisson system
(condition number ~ n^2)."""
import os

import numpy as np
import pytest

from gpu_util import announce, run_solve

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DT = {0: np.float32, 1: np.float64}
CASES = [  # fixture, precision, tolerance
    ("manufactured_f64_n1024.npz", 1, 1e-12),
    ("manufactured_f32_n1024.npz", 0, 1e-5),
    ("poisson_f64_n1024.npz", 1, 1e-9),
    ("dense_dd_f64_n64_b8.npz", 1, 1e-12),
    ("dense_dd_f64_n64_b8.npz", 0, 1e-5),
]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: np.atleast_2d(z[k]) for k in z.files}


def _replicate(f, reps):
    """(batch, n) arrays of the fixture systems, each repeated `reps` times."""
    return {k: np.repeat(v, reps, axis=0) for k, v in f.items()}


def _arrays(f, layout, dtype):
    a, b, c, d = (f[k].astype(dtype) for k in "abcd")
    if layout == 1:
        return (np.ascontiguousarray(np.concatenate([a, b, c, d], axis=1)).ravel(),)
    if layout == 2:
        return tuple(np.ascontiguousarray(v.T).ravel() for v in (a, b, c, d))
    return tuple(np.ascontiguousarray(v).ravel() for v in (a, b, c, d))


def _x(out, layout, batch, n):
    if layout == 1:
        return out.reshape(batch, 4 * n)[:, 3 * n:]
    if layout == 2:
        return out.reshape(n, batch).T
    return out.reshape(batch, n)


@pytest.mark.parametrize("name,prec,tol", CASES)
def test_fixture_every_algorithm(name, prec, tol):
    import tridiagonalsolver_amd as T
    L = T.lib()
    f = _load(name)
    reps = 16 if f["x"].shape[0] == 1 else 4
    g = _replicate(f, reps)
    batch, n = g["x"].shape
    runs = [(algo, L.tridiag_algo_layout(prec, algo)) for algo in range(L.tridiag_algo_count(prec))]
    runs += [(-1, 0), (-1, 2)]
    fails = []
    for algo, layout in runs:
        if layout == 1 and prec == 1:
            continue
        announce("golden", name, "prec", prec, "algo", algo, "layout", layout)
        out = run_solve(_arrays(g, layout, DT[prec]), n, batch, layout, algo, DT[prec])
        x = _x(out, layout, batch, n).astype(np.float64)
        err = np.max(np.abs(x - g["x"])) / np.max(np.abs(g["x"]))
        if not err <= tol:
            fails.append((algo, layout, float(err)))
    assert not fails, fails
