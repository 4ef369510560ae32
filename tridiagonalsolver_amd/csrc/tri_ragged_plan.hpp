// Host planning of the ragged persistent solve: the shape table, the work-buffer
// format and the schedule builder.  Plain C++ without HIP, so that the planner
// also builds and runs under AddressSanitizer on the CPU.
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

namespace tri {

// Shapes of the ragged kernel (512 lanes): {M, NWS, S}, S systems per workgroup,
// M rows per lane.  NWS >= 1: NWS waves (64*NWS lanes) per system.  NWS = 0: a
// sub-wave shape, 512/S lanes per system and 64*S/512 systems per wave.  Capacity:
// lanes per system * M rows.
struct RaggedShape {
    int M, NWS, S;
};
constexpr int kRaggedLanes = 512;
constexpr int kRaggedShapes = 25;
constexpr RaggedShape kRaggedShapeTab[kRaggedShapes] = {
    {2, 1, 8},  {4, 1, 8},  {8, 1, 8},  {8, 2, 4},  {16, 1, 8}, {8, 4, 2},
    {16, 2, 4}, {8, 8, 1},  {16, 4, 2}, {16, 8, 1}, {4, 2, 4},  {4, 4, 2},
    // 12 rows per lane: capacities 1.5x a power of two, so that systems between
    // two powers of two leave fewer padding rows
    {12, 1, 8}, {12, 2, 4}, {12, 4, 2}, {12, 8, 1},
    // sub-wave shapes for short systems: 16 lanes (capacities 64, 128, 256) or 8
    // lanes (32, 64) per system
    {4, 0, 32}, {8, 0, 32}, {16, 0, 32}, {4, 0, 64}, {8, 0, 64},
    // 32 lanes (capacities 256, 384, 512) and 16 lanes of 12 rows (192)
    {8, 0, 16}, {12, 0, 16}, {16, 0, 16}, {12, 0, 32}};
// Lanes (segments) per system of shape k.
constexpr int ragged_lanes(int k) {
    return kRaggedShapeTab[k].NWS ? 64 * kRaggedShapeTab[k].NWS : kRaggedLanes / kRaggedShapeTab[k].S;
}
constexpr int ragged_cap(int k) { return ragged_lanes(k) * kRaggedShapeTab[k].M; }
constexpr int kRaggedWaves = kRaggedLanes / 64;
constexpr int kRaggedMaxN = 64 * 8 * 16;
constexpr int kRaggedMaxM = 16;
// Interface solver families of a bucket: 0 .. kRaggedIfaces-1 (the Iface values).
constexpr int kRaggedIfaces = 5;
// Interface-solver codes: the family in bits 0-3 and, for the PCR + Thomas family
// (3), the PCR depth d in bits 4-7 (0: the family's default).  Valid for a system
// of P lane segments when d leaves a stride of at most P/2 segments: 1 <= d <=
// log2(P).  Codes are below kIfaceCodes.
constexpr int kIfaceCodes = 256;
inline bool iface_code_valid(int code, int P) {
    if (code < 0 || code >= kIfaceCodes) return false;
    const int fam = code & 15, depth = code >> 4;
    if (fam >= kRaggedIfaces) return false;
    if (depth == 0) return true;
    return fam == 3 && (1 << (depth - 1)) <= P / 2;
}
// Lane segments per system of ragged shape k.
constexpr int ragged_segments(int k) { return ragged_lanes(k); }

// Work buffer: a 64-byte header, one 128-byte record per item, then the system
// table of the sub-wave items (16 bytes per system):
//   r[0] shape, r[1] interface solver, r[2] number of systems,
//   r[3] (sub-wave shapes) index of the item's first system-table entry, else 0,
//   r[4 + 3k], r[5 + 3k]  first row of system slot k (int64, low word first),
//   r[6 + 3k]             rows of system slot k (shapes with NWS >= 1);
//   table entry: first row (int64, low word first), rows, 0.
// Written by the host planner once (with both counters zero).  A solve leaves the
// header as it found it: the last workgroup to finish zeroes both counters.
struct RaggedHeader {
    int64_t magic;
    int64_t n_sys;
    int64_t nitems;
    int32_t counter;     // work-queue head
    int32_t precision;
    int32_t done;        // workgroups of the running solve that have left the queue
    int64_t rows;        // offsets[n_sys]: rows of the arrays that the systems cover
    int64_t ntab;        // entries of the system table (after the nitems records)
};
constexpr int64_t kRaggedMagic = 0x52414747454432LL;
constexpr int64_t kRaggedHeaderBytes = 64;
static_assert(sizeof(RaggedHeader) <= kRaggedHeaderBytes, "header exceeds its 64 bytes");
constexpr int kRaggedRecInts = 32;
constexpr int kRaggedMaxSlots = 8;
constexpr int kRaggedTabInts = 4;

// Upper bound of the work buffer: one record per system (every item holds at
// least one) and one table entry per system.
inline int64_t ragged_work_bytes(int64_t n_sys) {
    const int64_t raw = kRaggedHeaderBytes + (int64_t)(kRaggedRecInts + kRaggedTabInts) * 4 * n_sys;
    return (raw + 255) & ~(int64_t)255;
}

// Default shape of a system of n rows: the smallest capacity that holds it; among
// equal capacities the one with the fewest waves per system (no cross-wave chain).
inline int ragged_default_shape(int64_t n) {
    int best = -1;
    for (int k = 0; k < kRaggedShapes; ++k) {
        if (ragged_cap(k) < n) continue;
        if (best < 0 || ragged_cap(k) < ragged_cap(best) ||
            (ragged_cap(k) == ragged_cap(best) && kRaggedShapeTab[k].NWS < kRaggedShapeTab[best].NWS))
            best = k;
    }
    return best;
}

// Builds the work-buffer image of n_sys systems (offsets off[0..n_sys]):
// validates the offsets, buckets every non-empty system by (shape, interface
// solver code) — choice(n, shape, iface) gives a system's bucket unless
// force_shape / force_iface >= 0 fix it — orders the buckets by decreasing capacity (then
// increasing systems per item) and writes one record per item of up to S systems.
// On success `image` holds the header and records (its size is the number of
// bytes to upload) and the return value is 0; otherwise 1 (bad argument) or 2
// (a size the kernel does not take), with `err` set.
template <class Choice>
int ragged_schedule(int prec, int64_t n_sys, const int64_t *off, int force_shape, int force_iface,
                    Choice choice, std::vector<char> &image, std::string &err) {
    if (prec != 0 && prec != 1) return err = "precision must be 0 or 1", 1;
    if (n_sys < 0 || n_sys > 0x7fffffff) return err = "n_sys out of range", 1;
    if (n_sys > 0 && !off) return err = "null offsets", 1;
    if (force_shape >= kRaggedShapes || force_shape < -1) return err = "shape index out of range", 1;
    if (force_iface != -1 && !iface_code_valid(force_iface, 64 * 8))
        return err = "interface solver id out of range", 1;
    if (n_sys > 0 && off[0] < 0) return err = "offsets[0] is negative", 1;
    constexpr int NF = kIfaceCodes;
    std::vector<int32_t> key((size_t)n_sys, -1);
    std::vector<int64_t> cnt((size_t)kRaggedShapes * NF, 0);
    for (int64_t s = 0; s < n_sys; ++s) {
        const int64_t n = off[s + 1] - off[s];
        if (n < 0) return err = "offsets must be non-decreasing", 1;
        if (n == 0) continue;
        if (n > kRaggedMaxN)
            return err = "system " + std::to_string(s) + " has " + std::to_string(n) +
                         " rows; the ragged solve holds up to " + std::to_string(kRaggedMaxN),
                   2;
        int shape = -1, iface = 0;
        choice(n, shape, iface);
        if (shape < 0 || shape >= kRaggedShapes || ragged_cap(shape) < n) shape = ragged_default_shape(n);
        if (force_shape >= 0) {
            if (ragged_cap(force_shape) < n)
                return err = "system " + std::to_string(s) + " exceeds the capacity of the forced shape",
                       1;
            shape = force_shape;
        }
        if (!iface_code_valid(iface, ragged_segments(shape))) iface = 0;
        if (force_iface >= 0) {
            if (!iface_code_valid(force_iface, ragged_segments(shape)))
                return err = "PCR depth " + std::to_string(force_iface >> 4) +
                             " exceeds the segments of shape " + std::to_string(shape),
                       1;
            iface = force_iface;
        }
        key[(size_t)s] = shape * NF + iface;
        ++cnt[(size_t)key[(size_t)s]];
    }
    // buckets in order of decreasing capacity (then increasing systems per item)
    std::vector<int> order(cnt.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [](int x, int y) {
        const int sx = x / NF, sy = y / NF;
        if (ragged_cap(sx) != ragged_cap(sy)) return ragged_cap(sx) > ragged_cap(sy);
        return kRaggedShapeTab[sx].S < kRaggedShapeTab[sy].S;
    });
    // systems of each bucket, in input order
    std::vector<int64_t> first(cnt.size(), 0);
    int64_t pos = 0;
    for (int k : order) {
        first[(size_t)k] = pos;
        pos += cnt[(size_t)k];
    }
    std::vector<int32_t> perm((size_t)pos);
    {
        std::vector<int64_t> fill(first);
        for (int64_t s = 0; s < n_sys; ++s)
            if (key[(size_t)s] >= 0) perm[(size_t)fill[(size_t)key[(size_t)s]]++] = (int32_t)s;
    }
    std::vector<char> host((size_t)ragged_work_bytes(n_sys), 0);
    int32_t *recs = reinterpret_cast<int32_t *>(host.data() + kRaggedHeaderBytes);
    std::vector<int32_t> tab;
    int64_t nitems = 0;
    for (int k : order) {
        const int shape = k / NF;
        const int S = kRaggedShapeTab[shape].S;
        const bool sub = kRaggedShapeTab[shape].NWS == 0;
        for (int64_t i = 0; i < cnt[(size_t)k]; i += S) {
            int32_t *r = recs + (size_t)kRaggedRecInts * nitems++;
            const int count = (int)std::min<int64_t>(S, cnt[(size_t)k] - i);
            r[0] = shape;
            r[1] = k % NF;
            r[2] = count;
            r[3] = sub ? (int32_t)(tab.size() / kRaggedTabInts) : 0;
            for (int q = 0; q < count; ++q) {
                const int64_t s = perm[(size_t)(first[(size_t)k] + i + q)];
                const uint64_t o = (uint64_t)off[s];
                const int32_t lo = (int32_t)(uint32_t)(o & 0xffffffffu), hi = (int32_t)(uint32_t)(o >> 32);
                const int32_t rows = (int32_t)(off[s + 1] - off[s]);
                if (sub) {
                    tab.insert(tab.end(), {lo, hi, rows, 0});
                } else {
                    r[4 + 3 * q] = lo;
                    r[5 + 3 * q] = hi;
                    r[6 + 3 * q] = rows;
                }
            }
        }
    }
    const int64_t ntab = (int64_t)tab.size() / kRaggedTabInts;
    const RaggedHeader h{kRaggedMagic, n_sys, nitems, 0, prec, 0, n_sys > 0 ? off[n_sys] : 0, ntab};
    memcpy(host.data(), &h, sizeof h);
    const int64_t rec_bytes = kRaggedHeaderBytes + (int64_t)kRaggedRecInts * 4 * nitems;
    if (ntab) memcpy(host.data() + rec_bytes, tab.data(), tab.size() * sizeof(int32_t));
    host.resize((size_t)(rec_bytes + ntab * kRaggedTabInts * 4));
    image.swap(host);
    return 0;
}

// Geometry of a layout conversion (tridiag_convert_layout): element (coefficient
// k, system s, row i) sits at ptr_k[s*n + i] (layout 0), buf[s*4n + k*n + i]
// (layout 1) or ptr_k[i*batch + s] (layout 2).  Layouts 0 and 1 are row-pitched
// views of one system-major matrix and layout 2 is its transpose, so every
// conversion is a copy or a transpose of a rows x cols matrix with row pitches
// ld_src / ld_dst.
struct LayoutGeometry {
    int64_t rows, cols, ld_src, ld_dst;
    bool transpose;
};
inline LayoutGeometry layout_geometry(int src_layout, int dst_layout, int64_t n, int64_t batch) {
    const bool src_sys = src_layout != 2, dst_sys = dst_layout != 2;
    const int64_t ps = src_layout == 1 ? 4 * n : n;
    const int64_t pd = dst_layout == 1 ? 4 * n : n;
    const int64_t elems = n * batch;
    if (src_sys && dst_sys) {              // system-major to system-major: pitched copy
        if (ps == n && pd == n) return {1, elems, elems, elems, false};
        return {batch, n, ps, pd, false};
    }
    if (src_sys) return {batch, n, ps, batch, true};     // to interleaved: batch x n
    if (dst_sys) return {n, batch, batch, pd, true};     // from interleaved: n x batch
    return {1, elems, elems, elems, false};               // interleaved to interleaved
}

}  // namespace tri
