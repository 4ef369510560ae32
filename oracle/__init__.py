"""Test infrastructure: the FP64 CPU oracle for the batched tridiagonal path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, and only as the checker (or, cpu_ref, as the measured CPU baseline).
The product library never imports it.
"""
from .oracle import *  # noqa: F401,F403
from . import cpu_ref  # noqa: F401
