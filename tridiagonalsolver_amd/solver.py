"""Torch-facing entry points over the C-ABI (device memory and streams come from torch;
all arithmetic happens in the HIP kernels of libtridiag_amd.so)."""
import ctypes

import torch

from ._capi import (ALGO_AUTO, CONTIGUOUS, DOUBLE, INTERLEAVED, PACKED, SINGLE, check, lib)


def _precision(t):
    if t.dtype == torch.float32:
        return SINGLE
    if t.dtype == torch.float64:
        return DOUBLE
    raise TypeError("only float32 and float64 systems are supported, got %s" % t.dtype)


def _stream_handle(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_tensors(ts):
    t0 = ts[0]
    if not t0.is_cuda:
        raise ValueError("tensors must be on a GPU device")
    for t in ts:
        if t.dtype != t0.dtype or t.device != t0.device:
            raise ValueError("all tensors must share dtype and device")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


def _shape(t, layout, n, batch):
    if n is not None and batch is not None:
        return int(n), int(batch)
    if t.dim() != 2:
        raise ValueError("pass n and batch for 1-D tensors")
    if layout == INTERLEAVED:
        return t.shape[0], t.shape[1]
    if layout == PACKED:
        return t.shape[1] // 4, t.shape[0]
    return t.shape[1], t.shape[0]


def solve(a, b, c, d, algo=ALGO_AUTO, layout=CONTIGUOUS, n=None, batch=None):
    """Solves in place: x is written to d; c may be clobbered (naive Thomas writes c').

    layout 0: (batch, n) per-system rows; layout 2: (n, batch) interleaved;
    layout 1: one (batch, 4n) packed buffer passed as `a` (b, c, d ignored).
    """
    ts = [a] if layout == PACKED else [a, b, c, d]
    _check_tensors(ts)
    n, batch = _shape(a if layout == PACKED else d, layout, n, batch)
    prec = _precision(a)
    args = [_ptr(a)] + ([_ptr(a)] * 3 if layout == PACKED else [_ptr(b), _ptr(c), _ptr(d)])
    check(lib().tridiag_solve(prec, algo, layout, n, batch, *args, _stream_handle(a.device)))
    return a if layout == PACKED else d


def solve_bulk(a, b, c, d, algo, layout=CONTIGUOUS, n=None, batch=None):
    """The reference's own form of algorithm `algo` (id 2: CR through global
    memory; ids 3-10: PCR-pThomas, PCR-half-pThomas, WM-pGE, WM-half-pGE,
    shared-memory or register exchanges; ids 11-13: CR over 4, 8, 16 rows per
    lane, n <= 1024 R), in place like solve(): x into d (and, for ids 3-10, c'
    into c).  Layouts 0 and 1."""
    ts = [a] if layout == PACKED else [a, b, c, d]
    _check_tensors(ts)
    n, batch = _shape(a if layout == PACKED else d, layout, n, batch)
    prec = _precision(a)
    args = [_ptr(a)] + ([_ptr(a)] * 3 if layout == PACKED else [_ptr(b), _ptr(c), _ptr(d)])
    check(lib().tridiag_solve_bulk(prec, algo, layout, n, batch, *args, _stream_handle(a.device)))
    return a if layout == PACKED else d


def convert_layout(arrays, src_layout, dst_layout, n=None, batch=None):
    """Copies coefficient arrays into another layout (a permutation, bit-exact).

    `arrays` is a packed (batch, 4n) tensor for layout 1, else a sequence of four
    tensors (a, b, c, d) in which None skips that coefficient.  Returns a new
    packed tensor for layout 1, else a tuple of four tensors, (batch, n) for
    layout 0 and (n, batch) for layout 2, with None where skipped.  A packed
    destination needs all four coefficients.
    """
    if src_layout == PACKED:
        src = [arrays]
        ref = arrays
    else:
        src = list(arrays)
        if len(src) != 4:
            raise ValueError("pass four arrays (a, b, c, d); None skips one")
        ref = next((t for t in src if t is not None), None)
        if ref is None:
            raise ValueError("no array to convert")
    given = [t for t in src if t is not None]
    _check_tensors(given)
    if n is None or batch is None:
        n, batch = _shape(ref, src_layout, n, batch)
    n, batch = int(n), int(batch)
    if dst_layout == PACKED:
        if src_layout != PACKED and len(given) != 4:
            raise ValueError("a packed destination needs all four coefficients")
        out = torch.empty((batch, 4 * n), dtype=ref.dtype, device=ref.device)
        dptr = [_ptr(out), None, None, None]
    else:
        shape = (n, batch) if dst_layout == INTERLEAVED else (batch, n)
        have = [True] * 4 if src_layout == PACKED else [t is not None for t in src]
        outs = [torch.empty(shape, dtype=ref.dtype, device=ref.device) if h else None
                for h in have]
        dptr = [_ptr(o) if o is not None else None for o in outs]
    sptr = [_ptr(src[0]), None, None, None] if src_layout == PACKED else \
        [_ptr(t) if t is not None else None for t in src]
    check(lib().tridiag_convert_layout(_precision(ref), src_layout, dst_layout, n, batch,
                                       *sptr, *dptr, _stream_handle(ref.device)))
    return out if dst_layout == PACKED else tuple(outs)


class RaggedPlan:
    """Schedule of a ragged batch: system s = rows [offsets[s], offsets[s+1]) of 1-D
    tensors a, b, c, d (contiguous per system, concatenated), planned once on the
    host (size buckets, per-bucket shape and interface solver) and held in a device
    work buffer.  `solve` only launches kernels: it allocates nothing and can be
    captured into a graph and replayed.  Solves on one plan must not overlap.

    shape / iface: -1 = per-bucket choice from the tuning table; >= 0 forces one
    shape of the ragged kernel (ragged_shapes()) / interface solver for all systems.
    """

    def __init__(self, offsets, dtype=torch.float32, device=None, shape=-1, iface=-1):
        self.precision = _precision(torch.empty(0, dtype=dtype))
        self.offsets_host = torch.as_tensor(offsets, dtype=torch.int64).contiguous().cpu()
        self.n_sys = self.offsets_host.numel() - 1
        if self.n_sys < 0:
            raise ValueError("offsets needs at least one entry")
        device = torch.device(device if device is not None else "cuda")
        self.device = device
        self.work = torch.empty(int(lib().tridiag_ragged_work_size(self.n_sys)),
                                dtype=torch.uint8, device=device)
        check(lib().tridiag_ragged_plan(self.precision, self.n_sys,
                                        ctypes.c_void_p(self.offsets_host.data_ptr()), shape,
                                        iface, _ptr(self.work), _stream_handle(device)))

    @property
    def rows(self):
        return int(self.offsets_host[-1]) if self.n_sys >= 0 else 0

    def solve(self, a, b, c, d):
        """Solves every system in place (x -> d) on the current stream."""
        _check_tensors([a, b, c, d])
        if _precision(a) != self.precision:
            raise ValueError("tensor dtype does not match the plan's precision")
        for t in (a, b, c, d):
            if t.numel() < self.rows:
                raise ValueError("tensors hold fewer rows than offsets[-1]")
        check(lib().tridiag_ragged_solve(self.precision, self.n_sys, _ptr(a), _ptr(b), _ptr(c),
                                         _ptr(d), _ptr(self.work), _stream_handle(a.device)))
        return d


def solve_ragged(a, b, c, d, offsets):
    """Ragged batch in one call (plan + solve): system s = rows
    [offsets[s], offsets[s+1]) of the 1-D tensors; x is written to d."""
    _check_tensors([a, b, c, d])
    RaggedPlan(offsets, dtype=a.dtype, device=a.device).solve(a, b, c, d)
    return d


def solve_adjoint(a, b, c, x, g, layout=CONTIGUOUS, n=None, batch=None):
    """Returns (ga, gb, gc, gd) = dL/d(a, b, c, d) for dL/dx = g."""
    _check_tensors([a, b, c, x, g])
    n, batch = _shape(x, layout, n, batch)
    outs = [torch.empty_like(x) for _ in range(4)]
    check(lib().tridiag_adjoint(_precision(a), layout, n, batch, _ptr(a), _ptr(b), _ptr(c),
                                _ptr(x), _ptr(g), *(_ptr(o) for o in outs),
                                _stream_handle(a.device)))
    return tuple(outs)


def solve_fwd_bwd(a, b, c, d, g, layout=CONTIGUOUS, n=None, batch=None):
    """Fused forward + adjoint for a known upstream gradient g: returns
    (x, ga, gb, gc, gd)."""
    _check_tensors([a, b, c, d, g])
    n, batch = _shape(d, layout, n, batch)
    x = torch.empty_like(d)
    outs = [torch.empty_like(d) for _ in range(4)]
    check(lib().tridiag_solve_fwd_bwd(_precision(a), layout, n, batch, _ptr(a), _ptr(b),
                                      _ptr(c), _ptr(d), _ptr(g), _ptr(x),
                                      *(_ptr(o) for o in outs), _stream_handle(a.device)))
    return (x,) + tuple(outs)


class TridiagonalSolve(torch.autograd.Function):
    """x = A^{-1} d with A = tridiag(a, b, c); a_0 and c_{n-1} are not part of A.

    Forward runs the on-chip solver on copies (inputs are not clobbered) and saves
    only a, b, c and x.  Backward solves A^T lambda = dL/dx with the adjoint
    kernel, which recomputes its own elimination, and returns
    dL/da = -lambda x_{i-1}, dL/db = -lambda x, dL/dc = -lambda x_{i+1}, dL/dd = lambda.
    """

    @staticmethod
    def forward(ctx, a, b, c, d, layout=CONTIGUOUS, algo=ALGO_AUTO):
        n, batch = _shape(d, layout, None, None)
        x = d.contiguous().clone()
        cc = c.contiguous().clone() if algo == 1 else c.contiguous()
        solve(a.contiguous(), b.contiguous(), cc, x, algo=algo, layout=layout, n=n, batch=batch)
        ctx.layout = layout
        ctx.save_for_backward(a, b, c, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        a, b, c, x = ctx.saved_tensors
        ga, gb, gc, gd = solve_adjoint(a.contiguous(), b.contiguous(), c.contiguous(), x,
                                       gx.contiguous(), layout=ctx.layout)
        return ga, gb, gc, gd, None, None


def tridiagonal_solve(a, b, c, d, layout=CONTIGUOUS, algo=ALGO_AUTO):
    """Differentiable solve: returns x, gradients flow to a, b, c, d."""
    return TridiagonalSolve.apply(a, b, c, d, layout, algo)
