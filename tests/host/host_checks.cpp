// Host-side planning code of the library, exercised on the CPU (built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_asan.py):
// the ragged schedule builder and the layout-conversion geometry.  Every
// schedule is checked for what the kernel relies on: each non-empty system in
// exactly one record slot with its first row and size, capacities that hold the
// system, at most S systems per item, buckets in order of decreasing capacity.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "../../tridiagonalsolver_amd/csrc/tri_coop_sched.hpp"
#include "../../tridiagonalsolver_amd/csrc/tri_ragged_plan.hpp"

using namespace tri;

static int failures = 0;
#define CHECK(c, ...)                                             \
    do {                                                          \
        if (!(c)) {                                               \
            ++failures;                                           \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                         \
            fprintf(stderr, "\n");                                \
        }                                                         \
    } while (0)

static void check_schedule(const std::vector<int64_t> &off, int prec, int fs, int fi,
                           const std::vector<char> &img) {
    const int64_t n_sys = (int64_t)off.size() - 1;
    RaggedHeader h;
    CHECK(img.size() >= (size_t)kRaggedHeaderBytes, "image shorter than its header");
    memcpy(&h, img.data(), sizeof h);
    CHECK(h.magic == kRaggedMagic && h.n_sys == n_sys && h.precision == prec, "header fields");
    CHECK(h.counter == 0 && h.done == 0, "counters not zero");
    CHECK(h.rows == (n_sys > 0 ? off[n_sys] : 0), "rows");
    CHECK((int64_t)img.size() == kRaggedHeaderBytes + (int64_t)kRaggedRecInts * 4 * h.nitems +
                                     (int64_t)kRaggedTabInts * 4 * h.ntab,
          "image size %zu for %lld items, %lld table entries", img.size(), (long long)h.nitems,
          (long long)h.ntab);
    CHECK((int64_t)img.size() <= ragged_work_bytes(n_sys), "image larger than the work buffer");
    std::vector<int> seen((size_t)n_sys, 0);
    const int32_t *recs = reinterpret_cast<const int32_t *>(img.data() + kRaggedHeaderBytes);
    const int32_t *tab = recs + (size_t)kRaggedRecInts * h.nitems;
    int64_t tab_next = 0;
    int last_cap = 1 << 30;
    for (int64_t it = 0; it < h.nitems; ++it) {
        const int32_t *r = recs + (size_t)kRaggedRecInts * it;
        const int shape = r[0], iface = r[1], count = r[2];
        CHECK(shape >= 0 && shape < kRaggedShapes, "shape %d", shape);
        if (shape < 0 || shape >= kRaggedShapes) return;
        CHECK(iface_code_valid(iface, ragged_segments(shape)), "iface %d", iface);
        CHECK(count >= 1 && count <= kRaggedShapeTab[shape].S, "count %d", count);
        CHECK(ragged_cap(shape) <= last_cap, "buckets out of capacity order");
        last_cap = ragged_cap(shape);
        if (fs >= 0) CHECK(shape == fs, "forced shape ignored");
        if (fi >= 0) CHECK(iface == fi, "forced iface ignored");
        const bool sub = kRaggedShapeTab[shape].NWS == 0;
        if (sub) {
            // sub-wave items: consecutive, non-overlapping runs of the system table
            CHECK(r[3] == tab_next && tab_next + count <= h.ntab, "table run %d of %d", r[3], count);
            if (r[3] != tab_next || tab_next + count > h.ntab) return;
            tab_next += count;
        } else {
            CHECK(r[3] == 0 && count <= kRaggedMaxSlots, "record slots");
        }
        for (int q = 0; q < count; ++q) {
            const int32_t *e = sub ? tab + (size_t)kRaggedTabInts * (r[3] + q) : r + 4 + 3 * q;
            const int64_t o = (int64_t)(((uint64_t)(uint32_t)e[1] << 32) | (uint32_t)e[0]);
            const int64_t n = e[2];
            if (sub) CHECK(e[3] == 0, "table padding");
            CHECK(n >= 1 && n <= ragged_cap(shape), "system of %lld rows in capacity %d",
                  (long long)n, ragged_cap(shape));
            // which system starts at o with n rows (offsets are non-decreasing)
            const auto pos = std::lower_bound(off.begin(), off.end() - 1, o) - off.begin();
            int64_t s = pos;
            while (s < n_sys && off[s] == o && off[s + 1] - off[s] == 0) ++s;   // skip empties
            CHECK(s < n_sys && off[s] == o && off[s + 1] - off[s] == n, "record does not match a system");
            if (s < n_sys) ++seen[(size_t)s];
        }
    }
    CHECK(tab_next == h.ntab, "table entries not all used");
    for (int64_t s = 0; s < n_sys; ++s) {
        const int64_t n = off[s + 1] - off[s];
        CHECK(seen[(size_t)s] == (n > 0 ? 1 : 0), "system %lld seen %d times", (long long)s,
              seen[(size_t)s]);
    }
}

static std::vector<int64_t> offsets(const std::vector<int64_t> &sizes, int64_t start) {
    std::vector<int64_t> off(sizes.size() + 1, start);
    for (size_t i = 0; i < sizes.size(); ++i) off[i + 1] = off[i] + sizes[i];
    return off;
}

static void ragged_cases() {
    std::mt19937_64 rng(1234);
    auto dflt = [](int64_t, int &shape, int &iface) { shape = -1; iface = 0; };
    auto odd = [](int64_t n, int &shape, int &iface) {   // a tuning table choosing other shapes
        shape = (int)(n % kRaggedShapes);
        iface = (int)(n % kRaggedIfaces) | (int)((n / 7) % 11) << 4;   // some depths invalid
    };
    for (int trial = 0; trial < 300; ++trial) {
        const int64_t n_sys = (int64_t)(rng() % 3000);
        std::vector<int64_t> sizes((size_t)n_sys);
        for (auto &v : sizes) {
            const uint64_t u = rng() % 100;
            v = u < 10 ? 0 : (u < 20 ? (int64_t)(1 + rng() % 40) : (int64_t)(1 + rng() % kRaggedMaxN));
        }
        const int64_t start = (trial % 3 == 0) ? 0 : (int64_t)(rng() % (1ll << 34));
        const auto off = offsets(sizes, start);
        const int prec = trial & 1;
        std::vector<char> img;
        std::string err;
        int rc = trial % 2 ? ragged_schedule(prec, n_sys, off.data(), -1, -1, odd, img, err)
                           : ragged_schedule(prec, n_sys, off.data(), -1, -1, dflt, img, err);
        CHECK(rc == 0, "schedule failed: %s", err.c_str());
        if (rc == 0) check_schedule(off, prec, -1, -1, img);
        // every shape forced on the systems it holds
        for (int k = 0; k < kRaggedShapes; ++k) {
            std::vector<int64_t> fit;
            for (auto v : sizes) fit.push_back(std::min<int64_t>(v, ragged_cap(k)));
            const auto off2 = offsets(fit, start);
            img.clear();
            rc = ragged_schedule(prec, n_sys, off2.data(), k, -1, dflt, img, err);
            CHECK(rc == 0, "forced shape %d failed: %s", k, err.c_str());
            if (rc == 0) check_schedule(off2, prec, k, -1, img);
        }
        // forced shape of the largest capacity, forced interface solver
        int big = 0;
        for (int k = 0; k < kRaggedShapes; ++k) if (ragged_cap(k) > ragged_cap(big)) big = k;
        img.clear();
        rc = ragged_schedule(prec, n_sys, off.data(), big, trial % kRaggedIfaces, dflt, img, err);
        CHECK(rc == 0, "forced schedule failed: %s", err.c_str());
        if (rc == 0) check_schedule(off, prec, big, trial % kRaggedIfaces, img);
    }
    // argument errors
    auto dflt2 = [](int64_t, int &shape, int &iface) { shape = -1; iface = 0; };
    std::vector<char> img;
    std::string err;
    const std::vector<int64_t> bad_order = {0, 10, 5};
    CHECK(ragged_schedule(0, 2, bad_order.data(), -1, -1, dflt2, img, err) == 1, "decreasing offsets");
    const std::vector<int64_t> too_big = {0, kRaggedMaxN + 1};
    CHECK(ragged_schedule(0, 1, too_big.data(), -1, -1, dflt2, img, err) == 2, "oversized system");
    const std::vector<int64_t> neg = {-4, 10};
    CHECK(ragged_schedule(0, 1, neg.data(), -1, -1, dflt2, img, err) == 1, "negative first offset");
    const std::vector<int64_t> small = {0, 200};
    CHECK(ragged_schedule(0, 1, small.data(), 0, -1, dflt2, img, err) == 1, "forced shape too small");
    CHECK(ragged_schedule(2, 1, small.data(), -1, -1, dflt2, img, err) == 1, "precision");
    CHECK(ragged_schedule(0, 1, small.data(), kRaggedShapes, -1, dflt2, img, err) == 1, "shape id");
    CHECK(ragged_schedule(0, 1, small.data(), -1, kRaggedIfaces, dflt2, img, err) == 1, "iface id");
    // PCR depths: up to log2 of the shape's segments (the shape of 8 waves: 512, depth 9)
    CHECK(ragged_schedule(0, 1, small.data(), -1, 3 | 2 << 4, dflt2, img, err) == 0, "PCR depth 2");
    CHECK(ragged_schedule(0, 1, small.data(), -1, 3 | 10 << 4, dflt2, img, err) == 1, "PCR depth 10");
    CHECK(ragged_schedule(0, 1, small.data(), -1, 2 | 1 << 4, dflt2, img, err) == 1,
          "depth on a family without one");
    CHECK(!iface_code_valid(3 | 7 << 4, 64) && iface_code_valid(3 | 6 << 4, 64), "depth bound");
    CHECK(ragged_schedule(0, 1, nullptr, -1, -1, dflt2, img, err) == 1, "null offsets");
    img.clear();
    CHECK(ragged_schedule(0, 0, nullptr, -1, -1, dflt2, img, err) == 0 &&
              img.size() == (size_t)kRaggedHeaderBytes, "empty batch");
}

// Element (k, s, i) of a coefficient in each layout, as the conversion defines it.
static int64_t layout_index(int layout, int64_t n, int64_t batch, int k, int64_t s, int64_t i) {
    if (layout == 0) return s * n + i;
    if (layout == 1) return s * 4 * n + k * n + i;
    return i * batch + s;
}

static void layout_cases() {
    for (int sl = 0; sl < 3; ++sl)
        for (int dl = 0; dl < 3; ++dl)
            for (int64_t n : {1, 2, 3, 7, 64, 65})
                for (int64_t batch : {1, 2, 5, 64, 67}) {
                    const LayoutGeometry g = layout_geometry(sl, dl, n, batch);
                    // every element (s, i): the matrix element (row, col) the geometry
                    // reads at row*ld_src + col must be the source element, and it must
                    // land at the destination element
                    for (int64_t s = 0; s < batch; ++s)
                        for (int64_t i = 0; i < n; ++i) {
                            int64_t row, col;
                            if (sl != 2) { row = s; col = i; }
                            else { row = i; col = s; }
                            if (g.rows == 1 && !g.transpose) {       // flat copy
                                col = (sl != 2) ? s * n + i : i * batch + s;
                                row = 0;
                            }
                            CHECK(row < g.rows && col < g.cols, "element outside the matrix");
                            const int64_t src = row * g.ld_src + col;
                            const int64_t dst = g.transpose ? col * g.ld_dst + row : row * g.ld_dst + col;
                            const int k = 0;
                            CHECK(src == layout_index(sl, n, batch, k, s, i) - (sl == 1 ? k * n : 0),
                                  "source index %d->%d n=%lld b=%lld", sl, dl, (long long)n,
                                  (long long)batch);
                            CHECK(dst == layout_index(dl, n, batch, k, s, i) - (dl == 1 ? k * n : 0),
                                  "destination index %d->%d n=%lld b=%lld", sl, dl, (long long)n,
                                  (long long)batch);
                        }
                }
    // sizes whose element count needs 64 bits: no overflow in the geometry
    const LayoutGeometry g = layout_geometry(0, 2, (int64_t)1 << 22, (int64_t)1 << 20);
    CHECK(g.rows == ((int64_t)1 << 20) && g.cols == ((int64_t)1 << 22) && g.transpose, "large");
}

// The cooperative schedule's slot -> group map: a bijection on [0, G) for every
// round width W, identity for W <= 1, and within a round the groups lie at least
// floor(G/W) apart once there is more than one round.
static void coop_spread_cases() {
    for (int G = 1; G <= 300; ++G)
        for (int W = 0; W <= 70; ++W) {
            std::vector<int> seen(G, 0);
            for (int s = 0; s < G; ++s) {
                const int g = coop_group(s, G, W);
                CHECK(g >= 0 && g < G, "G=%d W=%d s=%d -> %d", G, W, s, g);
                if (g >= 0 && g < G) ++seen[g];
                if (W <= 1 || G <= W) CHECK(g == s, "identity G=%d W=%d s=%d", G, W, s);
            }
            for (int g = 0; g < G; ++g) CHECK(seen[g] == 1, "G=%d W=%d group %d x%d", G, W, g, seen[g]);
            if (W > 1 && G > W)
                for (int s = 0; s + 1 < G; ++s)
                    if ((s + 1) % W != 0) {
                        const int d = coop_group(s + 1, G, W) - coop_group(s, G, W);
                        CHECK(d >= G / W, "G=%d W=%d slots %d,%d only %d apart", G, W, s, s + 1, d);
                    }
        }
    // the measured case: 128 groups of 32 FP32 systems (4096 x 4096), 32 per round
    CHECK(coop_group(0, 128, 32) == 0 && coop_group(1, 128, 32) == 4 && coop_group(32, 128, 32) == 1,
          "4096 x 4096 spread");
}

int main() {
    ragged_cases();
    layout_cases();
    coop_spread_cases();
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("ALL OK\n");
    return 0;
}
