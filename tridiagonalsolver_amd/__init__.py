"""MI355X-native batched tridiagonal solver (HIP kernels for gfx950 behind a C-ABI).

Python mirror of the reference's solver interface (Tridiagonal_CUDA/tridiagonal.h,
helpers.h solver table): solve() runs one algorithm id in place on device
tensors, exactly like the reference's NAME_{single,double}(a, b, c, d, n, batch)
calls, plus the interleaved layout, ragged batches and the adjoint.
"""
from ._capi import (ALGO_AUTO, CONTIGUOUS, DOUBLE, INTERLEAVED, PACKED, SINGLE,  # noqa: F401
                    TridiagError, algo_names, check, div_mode, lib, plan, ragged_bucket,
                    ragged_shapes, tiles)
from .solver import (RaggedPlan, convert_layout, solve, solve_adjoint, solve_bulk,  # noqa: F401
                     solve_fwd_bwd, solve_ragged, TridiagonalSolve, tridiagonal_solve)
