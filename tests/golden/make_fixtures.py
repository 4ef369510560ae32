"""Writes the known-answer fixtures of the parity tests (run from the repository root:
python tests/golden/make_fixtures.py).  Inputs, expected outputs, nothing else:

  manufactured_f64_n1024.npz   the reference test program's own known-answer system
                               (Tridiagonal_CUDA/test.cpp:19-38: x chosen, d = A x in
                               FP64), drawn with the oracle's seeded generator
  manufactured_f32_n1024.npz   the same recipe with values exactly representable in
                               FP32 (the single-precision run of that program)
  poisson_f64_n1024.npz        BPLG's known-answer initialisation
                               (Tridiagonal_CUDA/third_party/BPLib/alg/BPLG/BPLGTridiagWM.hxx:28-54): tridiag(-1, 2, -1),
                               d_0 = d_{n-1} = 1, x = 1 exactly
  dense_dd_f64_n64_b8.npz      8 random diagonally dominant systems of 64 rows whose x
                               comes from numpy's dense LU solve (LAPACK gesv), not from
                               the oracle

The reference itself ships no vectors; these are its two built-in answers plus one
set computed by an independent library.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def manufactured(single):
    import oracle
    a, b, c, d, x = oracle.gen_manufactured(1024, single=single)
    return dict(a=a, b=b, c=c, d=d, x=x)


def poisson(n=1024):
    a = -np.ones(n); b = 2 * np.ones(n); c = -np.ones(n); d = np.zeros(n)
    a[0] = 0.0; c[-1] = 0.0; d[0] = 1.0; d[-1] = 1.0
    return dict(a=a, b=b, c=c, d=d, x=np.ones(n))


def dense_dd(n=64, batch=8, seed=20240):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (batch, n)); c = rng.uniform(-1, 1, (batch, n))
    a[:, 0] = 0.0; c[:, -1] = 0.0
    b = np.abs(a) + np.abs(c) + rng.uniform(1, 2, (batch, n))
    d = rng.uniform(-1, 1, (batch, n))
    x = np.empty_like(d)
    for s in range(batch):
        A = np.diag(b[s]) + np.diag(a[s, 1:], -1) + np.diag(c[s, :-1], 1)
        x[s] = np.linalg.solve(A, d[s])
    return dict(a=a, b=b, c=c, d=d, x=x)


FIXTURES = {
    "manufactured_f64_n1024.npz": lambda: manufactured(False),
    "manufactured_f32_n1024.npz": lambda: manufactured(True),
    "poisson_f64_n1024.npz": poisson,
    "dense_dd_f64_n64_b8.npz": dense_dd,
}


def main():
    for name, make in FIXTURES.items():
        np.savez(os.path.join(HERE, name), **make())
        print("wrote", name)


if __name__ == "__main__":
    main()
