"""GPU parity of the ragged persistent solve, through the C-ABI.

Systems of uneven size n_s (contiguous per system, concatenated) against the
FP64 oracle's Thomas run system by system: per-system normwise relative error
<= 1e-5 (FP32) / 1e-12 (FP64).  The ragged path is further pinned by
bit-identity with the plain tile launch of the same shape and interface solver,
and by graph replays equal to the eager solve.
"""
import ctypes

import numpy as np
import pytest

import oracle
from gpu_util import TOL, announce

pytestmark = pytest.mark.gpu

DT = {0: np.float32, 1: np.float64}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import tridiagonalsolver_amd as T
    T.lib()


def _relerr_per_system(x, xref, off):
    """max_i |x_i - xref_i| / max_i |xref_i| for every non-empty system."""
    sizes = np.diff(off)
    keep = sizes > 0
    starts = off[:-1][keep]
    err = np.maximum.reduceat(np.abs(x - xref), starts)
    ref = np.maximum.reduceat(np.abs(xref), starts)
    return err / np.where(ref > 0, ref, 1.0)


def _case(sizes, dtype, seed, blo=1.0, bhi=2.0):
    off = oracle.offsets_from_sizes(np.asarray(sizes, dtype=np.int64))
    a, b, c, d = oracle.gen_dd_ragged(off, seed=seed, blo=blo, bhi=bhi)
    return off, tuple(v.astype(dtype) for v in (a, b, c, d))


def _run(off, arrs, shape=-1, iface=-1):
    import torch
    import tridiagonalsolver_amd as T
    dev = torch.device("cuda:0")
    ts = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in arrs]
    plan = T.RaggedPlan(off, dtype=ts[0].dtype, device=dev, shape=shape, iface=iface)
    plan.solve(*ts)
    torch.cuda.synchronize()
    return ts[3].cpu().numpy()


def _check(off, arrs, x, dtype):
    xref = oracle.thomas_ragged(*arrs, off)
    assert np.isfinite(x).all()
    if len(off) > 1 and off[-1] > 0:
        rel = _relerr_per_system(x.astype(np.float64), xref, off)
        assert rel.max() <= TOL[dtype], (rel.max(), int(np.argmax(rel)))


EDGE_SIZES = [1, 2, 3, 4, 5, 31, 32, 33, 64, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1000,
              1023, 1024, 1025, 2047, 2048, 2049, 4095, 4096, 4097, 6000, 8191, 8192]


@pytest.mark.parametrize("prec", [0, 1])
def test_edge_sizes(prec):
    """Every bucket boundary, with empty systems mixed in, in a shuffled order."""
    dtype = DT[prec]
    rng = np.random.default_rng(5)
    sizes = np.array(EDGE_SIZES * 3 + [0] * 7, dtype=np.int64)
    rng.shuffle(sizes)
    announce("ragged edge sizes", prec, len(sizes))
    off, arrs = _case(sizes, dtype, seed=11)
    _check(off, arrs, _run(off, arrs), dtype)


@pytest.mark.parametrize("prec", [0, 1])
def test_log_uniform_sizes(prec):
    dtype = DT[prec]
    sizes = oracle.ragged_sizes(3000, 32, 8192, seed=3 + prec)
    announce("ragged log-uniform", prec, int(sizes.sum()))
    off, arrs = _case(sizes, dtype, seed=12)
    _check(off, arrs, _run(off, arrs), dtype)


@pytest.mark.parametrize("prec", [0, 1])
def test_low_margin(prec):
    dtype = DT[prec]
    sizes = oracle.ragged_sizes(500, 2, 8192, seed=9)
    announce("ragged low margin", prec)
    off, arrs = _case(sizes, dtype, seed=13, blo=0.05, bhi=0.2)
    _check(off, arrs, _run(off, arrs), dtype)


def _shapes():
    import tridiagonalsolver_amd as T
    return T.ragged_shapes()


@pytest.mark.parametrize("prec", [0, 1])
def test_last_system_ends_mid_chunk(prec):
    """Arrays whose last row is not the end of a 16-byte chunk: the chunk holding the
    last system's final rows runs past the arrays and is read row by row, every other
    chunk whole (also chunks past a system, which re-read its first chunk).  Every
    shape that holds the last system, and the automatic plan."""
    dtype = DT[prec]
    vv = 16 // np.dtype(dtype).itemsize
    for last in (1, 2, 3, 5, 37, 1001, 4097):
        for lead in ([], [64], [7, 300]):
            sizes = lead + [last]
            if sum(sizes) % vv == 0:
                sizes = lead + [last + 1]
            off, arrs = _case(np.array(sizes, dtype=np.int64), dtype, seed=last + len(lead))
            announce("ragged last system", prec, sizes)
            _check(off, arrs, _run(off, arrs), dtype)
            for sh in _shapes():
                if sh["cap"] >= max(sizes):
                    _check(off, arrs, _run(off, arrs, shape=sh["index"]), dtype)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("iface", [0, 1, 2, 3, 4])
def test_forced_shape_and_interface(prec, iface):
    """Every shape of the persistent kernel with every interface solver."""
    dtype = DT[prec]
    for sh in _shapes():
        cap = sh["cap"]
        rng = np.random.default_rng(100 * sh["index"] + iface)
        sizes = np.concatenate([[cap, 1, cap - 1, max(1, cap // 2 + 1)],
                                rng.integers(1, cap + 1, size=2 * sh["S"] + 1)])
        announce("ragged forced", prec, "shape", sh, "iface", iface)
        off, arrs = _case(sizes, dtype, seed=sh["index"] + 17 * iface)
        _check(off, arrs, _run(off, arrs, shape=sh["index"], iface=iface), dtype)


@pytest.mark.parametrize("prec", [0, 1])
def test_pcr_depths(prec):
    """PCR + Thomas at PCR depths 1 .. log2(lanes per system) on every shape (interface-solver
    code 3 | depth << 4), and a depth past the shape's segments refused."""
    import torch
    import tridiagonalsolver_amd as T
    dtype = DT[prec]
    for sh in _shapes():
        cap, top = sh["cap"], sh["lanes"].bit_length() - 1
        rng = np.random.default_rng(7 * sh["index"])
        sizes = np.concatenate([[cap, 1, cap - 1], rng.integers(1, cap + 1, size=2 * sh["S"])])
        off, arrs = _case(sizes, dtype, seed=sh["index"] + 3)
        for depth in sorted({1, 2, top // 2, top - 1, top} - {0}):
            announce("ragged pcr depth", prec, "shape", sh, "depth", depth)
            _check(off, arrs, _run(off, arrs, shape=sh["index"], iface=3 | depth << 4), dtype)
        with pytest.raises(T.TridiagError):
            T.RaggedPlan(off, dtype=torch.float32 if prec == 0 else torch.float64,
                         shape=sh["index"], iface=3 | (top + 1) << 4)


@pytest.mark.parametrize("prec", [0, 1])
def test_bit_identical_to_plain_tile(prec):
    """A system solved in the persistent kernel equals, bit for bit, the same
    system solved by the plain tile launch of the same lane shape (M rows per
    lane, NWS waves per system) and interface solver (Wang-Mou)."""
    import torch
    import tridiagonalsolver_amd as T
    dtype = DT[prec]
    tiles = T.tiles(prec)
    checked = 0
    for sh in _shapes():
        cands = [t for t in tiles if not t["interleaved"] and t["padded"] and not t["wide_addr"]
                 and sh["NWS"] > 0 and t["M"] == sh["M"] and t["NW"] // t["S"] == sh["NWS"]
                 and not t["vec"]]
        if not cands:
            continue
        tile = cands[0]
        for n in sorted({sh["cap"], sh["cap"] - 3, sh["cap"] // 2 + 5}):
            if n < 1:
                continue
            batch = 5
            sizes = np.full(batch, n, dtype=np.int64)
            off, arrs = _case(sizes, dtype, seed=n + 7 * sh["index"])
            announce("ragged vs tile", prec, sh, tile["index"], n)
            xr = _run(off, arrs, shape=sh["index"], iface=0)
            ts = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in arrs]
            T.check(T.lib().tridiag_solve_tile(prec, tile["index"], 0, n, batch,
                                               *(ctypes.c_void_p(t.data_ptr()) for t in ts),
                                               ctypes.c_void_p(
                                                   torch.cuda.current_stream().cuda_stream)))
            torch.cuda.synchronize()
            xt = ts[3].cpu().numpy()
            assert np.array_equal(xr, xt), (sh, n, np.max(np.abs(xr - xt)))
            checked += 1
    assert checked > 0


def test_graph_replay_matches_eager():
    """The solve captured in a graph: every replay (with eager work and syncs in
    between) equals the eager solve bit for bit."""
    import torch
    import tridiagonalsolver_amd as T
    sizes = oracle.ragged_sizes(2000, 32, 8192, seed=21)
    off, arrs = _case(sizes, np.float32, seed=22)
    dev = torch.device("cuda:0")
    src = [torch.from_numpy(v).to(dev) for v in arrs]
    work = [t.clone() for t in src]
    plan = T.RaggedPlan(off, dtype=torch.float32, device=dev)
    plan.solve(*work)
    torch.cuda.synchronize()
    eager = work[3].clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    for t, u in zip(work, src):
        t.copy_(u)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        plan.solve(*work)
    torch.cuda.synchronize()
    for rep in range(4):
        announce("ragged graph replay", rep)
        work[3].copy_(src[3])
        _ = (src[0] * 2).sum()                 # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(work[3], eager), rep
    x = eager.cpu().numpy()
    _check(off, arrs, x, np.float32)


def test_repeated_solves_reset_queue():
    """Back-to-back solves on one plan: each single launch resets its own queue."""
    import torch
    import tridiagonalsolver_amd as T
    sizes = oracle.ragged_sizes(700, 1, 4000, seed=31)
    off, arrs = _case(sizes, np.float64, seed=32)
    dev = torch.device("cuda:0")
    src = [torch.from_numpy(v).to(dev) for v in arrs]
    plan = T.RaggedPlan(off, dtype=torch.float64, device=dev)
    outs = []
    for _ in range(3):
        d = src[3].clone()
        plan.solve(src[0], src[1], src[2], d)
        outs.append(d)
        torch.cuda.synchronize()
        # the last workgroup of the launch leaves both queue counters at zero
        # (header bytes 24..27 = queue head, 32..35 = finished workgroups)
        hdr = plan.work[:64].cpu().numpy()
        assert hdr[24:28].view(np.int32)[0] == 0 and hdr[32:36].view(np.int32)[0] == 0
    assert all(torch.equal(outs[0], o) for o in outs[1:])
    _check(off, arrs, outs[0].cpu().numpy(), np.float64)


def test_empty_and_zero_sized():
    import torch
    import tridiagonalsolver_amd as T
    dev = torch.device("cuda:0")
    z = torch.zeros(4, device=dev)
    T.solve_ragged(z, z, z, z, [0])                  # no systems
    T.solve_ragged(z, z, z, z, [0, 0, 0])            # only empty systems
    torch.cuda.synchronize()


def test_plan_errors():
    import torch
    import tridiagonalsolver_amd as T
    with pytest.raises(T.TridiagError) as e:
        T.RaggedPlan([0, 8193], dtype=torch.float32)
    assert e.value.code == 2
    with pytest.raises(T.TridiagError) as e:
        T.RaggedPlan([0, 10, 5], dtype=torch.float32)
    assert e.value.code == 1
    with pytest.raises(T.TridiagError):
        T.RaggedPlan([0, 200], dtype=torch.float32, shape=0)   # capacity 128
    plan = T.RaggedPlan([0, 10], dtype=torch.float32)
    with pytest.raises(ValueError):
        x = torch.zeros(10, dtype=torch.float64, device="cuda")
        plan.solve(x, x, x, x)


def test_solve_rejects_mismatched_plan():
    """A work buffer planned for one precision or system count is refused by a
    solve of another through the C-ABI (the kernel alone would skip it silently)."""
    import torch
    import tridiagonalsolver_amd as T
    from tridiagonalsolver_amd import _capi
    L = _capi.lib()
    dev = torch.device("cuda:0")
    off = [0, 10, 30]
    plan = T.RaggedPlan(off, dtype=torch.float32, device=dev)
    x32 = torch.ones(30, dtype=torch.float32, device=dev)
    x64 = torch.ones(30, dtype=torch.float64, device=dev)
    p = [ctypes.c_void_p(t.data_ptr()) for t in (x64, x64, x64, x64)]
    w = ctypes.c_void_p(plan.work.data_ptr())
    assert L.tridiag_ragged_solve(1, 2, *p, w, None) == _capi.EINVAL
    assert b"single" in L.tridiag_last_error()
    p = [ctypes.c_void_p(t.data_ptr()) for t in (x32, x32, x32, x32)]
    assert L.tridiag_ragged_solve(0, 1, *p, w, None) == _capi.EINVAL
    assert b"2 systems" in L.tridiag_last_error()
    assert L.tridiag_ragged_solve(0, 2, *p, w, None) == _capi.OK
    torch.cuda.synchronize()


def test_full_size_workload():
    """65,536 systems, n log-uniform in [32, 8192], FP32: every system against the oracle."""
    sizes = oracle.ragged_sizes(65536, 32, 8192, seed=2024)
    announce("ragged full size", int(sizes.sum()))
    off, arrs = _case(sizes, np.float32, seed=2025)
    _check(off, arrs, _run(off, arrs), np.float32)


@pytest.mark.parametrize("prec", [0, 1])
def test_misaligned_arrays(prec):
    """Arrays that start one row past a 16-byte boundary (views into larger buffers):
    systems read on the aligned chunk grid must not reach before the array."""
    import torch
    import tridiagonalsolver_amd as T
    dtype = np.float32 if prec == 0 else np.float64
    sizes = oracle.ragged_sizes(900, 1, 3000, seed=41)
    off, arrs = _case(sizes, dtype, seed=42)
    dev = torch.device("cuda:0")
    big = [torch.zeros(len(v) + 1, dtype=torch.float32 if prec == 0 else torch.float64,
                       device=dev) for v in arrs]
    for t, v in zip(big, arrs):
        t[1:] = torch.from_numpy(v).to(dev)
    views = [t[1:] for t in big]
    plan = T.RaggedPlan(off, dtype=views[0].dtype, device=dev)
    plan.solve(*views)
    torch.cuda.synchronize()
    assert float(big[3][0]) == 0.0                       # nothing written before the array
    _check(off, arrs, views[3].cpu().numpy(), dtype)


@pytest.mark.parametrize("prec", [0, 1])
def test_sub_wave_slots_give_identical_bits(prec):
    """The sub-wave shapes hold 16-64 systems per item, several per wave, read through
    per-lane addresses.  One system copied into every slot of several items (and at
    every 4-byte misalignment of its first row) must solve to the same bits in every
    copy, for each sub-wave shape and interface solver, and match the oracle."""
    dtype = DT[prec]
    for sh in _shapes():
        if sh["NWS"] != 0:
            continue
        for n in sorted({sh["cap"], sh["cap"] - 1, max(2, sh["cap"] // 2 + 3)}):
            a, b, c, d = oracle.gen_dd(n, 1, layout=0, seed=n + sh["index"])
            reps = 3 * sh["S"] + 5
            lead = [1, 2, 3]                       # shift the copies off 16-byte boundaries
            sizes = np.array(lead + [n] * reps, dtype=np.int64)
            off = oracle.offsets_from_sizes(sizes)
            head = [oracle.gen_dd(m, 1, layout=0, seed=m) for m in lead]
            arrs = tuple(np.concatenate([h[k] for h in head] + [v] * reps).astype(dtype)
                         for k, v in enumerate((a, b, c, d)))
            for iface in (0, 1, 3):
                announce("ragged sub-wave replicas", prec, sh, n, iface)
                x = _run(off, arrs, shape=sh["index"], iface=iface)
                copies = x[sum(lead):].reshape(reps, n)
                assert (copies == copies[:1]).all(), (sh, n, iface)
                _check(off, arrs, x, dtype)
