// C-ABI of libtridiag_amd.so: the reference-compatible symbols of
// Tridiagonal_CUDA/tridiagonal.h:8-57 plus the status-returning tridiag_* API.
#include <hip/hip_runtime.h>
#include <hipsparse/hipsparse.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <dlfcn.h>

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/tridiagonal.h"
#include "tri_adjoint.hpp"
#include "tri_coop.hpp"
#include "tri_host.hpp"
#include "tri_large.hpp"
#include "tri_ragged.hpp"

namespace tri {

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

static int hip_fail(hipError_t e, const char *where) {
    return fail(TRIDIAG_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- algorithm table
// Names and order of the reference solver table (Tridiagonal_CUDA/helpers.h:23-66).
enum Algo {
    kVendor = 0, kThomas = 1, kCR = 2, kPCR5Shmem = 3, kPCR4Shmem = 4, kPCR5Reg = 5, kPCR4Reg = 6,
    kWM5Shmem = 7, kWM4Shmem = 8, kWM5Reg = 9, kWM4Reg = 10, kCRReg4 = 11, kCRReg8 = 12,
    kCRReg16 = 13, kBPLGWM = 14, kBPLGLF = 15, kBPLGCR = 16, kBPLGPCR = 17
};

static const char *const kNamesSingle[] = {
    "hipSPARSE sgtsv", "Thomas", "CR", "shmem PCR-pThomas", "shmem PCR-half-pThomas",
    "reg PCR-pThomas", "reg PCR-half-pThomas", "shmem WM-pGE", "shmem WM-half-pGE",
    "reg WM-pGE", "reg WM-half-pGE", "reg-4 CR", "reg-8 CR", "reg-16 CR",
    "BPLG WM", "BPLG LF", "BPLG CR", "BPLG PCR"};
static const char *const kNamesDouble[] = {
    "hipSPARSE dgtsv", "Thomas", "CR", "shmem PCR-pThomas", "shmem PCR-half-pThomas",
    "reg PCR-pThomas", "reg PCR-half-pThomas", "shmem WM-pGE", "shmem WM-half-pGE",
    "reg WM-pGE", "reg WM-half-pGE", "reg-4 CR", "reg-8 CR", "reg-16 CR"};

static int algo_count(int prec) { return prec == TRIDIAG_SINGLE ? 18 : 14; }

// ---------------------------------------------------------------- tuning table
// Lines "precision layout family n_max M NW S vec lds_alpha nt xp" written by the
// autotuner (tridiagonalsolver_amd/autotune.py).  Read once, from
// $TRIDIAG_TUNING_FILE or <library dir>/../tuning/tuning_gfx950.txt.
struct TuneRow {
    int prec, layout, family;
    long long nmax;
    int M, NW, S, vec, alds, nt, xp;
};
static std::vector<TuneRow> g_tune;
// Ragged buckets: lines "R precision n_max shape iface" (shape index of the
// ragged kernel's shape table, interface solver id).
struct RaggedTuneRow {
    int prec;
    long long nmax;
    int shape, iface;
};
static std::vector<RaggedTuneRow> g_rtune;
static std::once_flag g_tune_once;

static void load_tuning() {
    std::string path;
    if (const char *e = getenv("TRIDIAG_TUNING_FILE")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr((void *)&load_tuning, &info) && info.dli_fname) {
            std::string so = info.dli_fname;
            size_t k = so.rfind('/');
            std::string dir = (k == std::string::npos) ? "." : so.substr(0, k);
            path = dir + "/../tuning/tuning_gfx950.txt";
        }
    }
    if (path.empty()) return;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        if (line[0] == 'R') {
            RaggedTuneRow q;
            if (sscanf(line + 1, "%d %lld %d %d", &q.prec, &q.nmax, &q.shape, &q.iface) == 4)
                g_rtune.push_back(q);
            continue;
        }
        TuneRow r;
        if (sscanf(line, "%d %d %d %lld %d %d %d %d %d %d %d", &r.prec, &r.layout, &r.family,
                   &r.nmax, &r.M, &r.NW, &r.S, &r.vec, &r.alds, &r.nt, &r.xp) == 11)
            g_tune.push_back(r);
    }
    fclose(f);
    std::stable_sort(g_tune.begin(), g_tune.end(),
                     [](const TuneRow &x, const TuneRow &y) { return x.nmax < y.nmax; });
    std::stable_sort(g_rtune.begin(), g_rtune.end(),
                     [](const RaggedTuneRow &x, const RaggedTuneRow &y) { return x.nmax < y.nmax; });
}

int max_lds_per_block() {
    static int v = [] {
        int dev = 0, x = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 65536;
        if (hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return 65536;
        return x;
    }();
    return v;
}

// Interface solver and exchange mechanism behind each algorithm id.
struct AlgoSpec {
    int iface;     // Iface
    int xchg_lds;  // Wang-Mou exchanges through LDS
    int lds_rows;  // prefer kernels that stage per-row coefficients in LDS
};
static AlgoSpec algo_spec(int algo) {
    switch (algo) {
    case kWM5Shmem: return {kIfaceWM, 1, 1};
    case kWM4Shmem: return {kIfaceWMHalf, 1, 1};
    case kWM4Reg: return {kIfaceWMHalf, 0, 0};
    case kPCR5Shmem: return {kIfacePCR, 0, 1};
    case kPCR5Reg: return {kIfacePCR, 0, 0};
    case kPCR4Shmem: return {kIfacePCRThomas, 0, 1};
    case kPCR4Reg: return {kIfacePCRThomas, 0, 0};
    case kCR: case kCRReg4: case kCRReg8: case kCRReg16: return {kIfaceCR, 0, 0};
    case kBPLGCR: return {kIfaceCR, 0, 0};
    case kBPLGLF: case kBPLGPCR: return {kIfacePCR, 0, 0};
    default: return {kIfaceWM, 0, 0};   // reg WM-pGE, BPLG WM, auto
    }
}

// ---------------------------------------------------------------- planning
template <typename T> static int pref_M(bool il) {
    if (il) return sizeof(T) == 4 ? 32 : 16;
    return sizeof(T) == 4 ? 4 : 8;
}
template <typename T> static int pref_S_il() { return 32 / (int)sizeof(T); }

// Rows per lane fixed by the algorithm (register-packed CR: R rows per lane).
static int forced_M(int algo) {
    switch (algo) {
    case kCRReg4: return 4;
    case kCRReg8: return 8;
    case kCRReg16: return 16;
    default: return 0;
    }
}

// LDS one workgroup of `e` needs with the interface solver of `algo`.
template <typename T> static long lds_need_iface(const TileEntry<T> &e, int iface, int xchg_lds) {
    const long w = sizeof(T), lanes = 64L * e.NW;
    const long nws = e.IL ? e.NW : e.NW / e.S;
    long st = (e.ALDS ? (long)(e.M - 2) * lanes * w : 0) + (e.XP ? lanes * e.M * w : 0) +
              (long)e.S * 2 * nws * 8 * w;
    long dyn = iface_family(iface) >= kIfacePCR ? 8 * lanes * w : (xchg_lds ? 6 * lanes * w : 0);
    return st + dyn;
}
template <typename T> static long lds_need(const TileEntry<T> &e, int algo) {
    const AlgoSpec sp = algo_spec(algo);
    return lds_need_iface(e, sp.iface, sp.xchg_lds);
}

template <typename T>
static const TileEntry<T> *find_tile(int M, int NW, int S, bool il, bool vec, bool alds, bool pad,
                                     bool big, bool nt, bool xp) {
    TileTable<T> tab = tile_table<T>();
    const TileEntry<T> *any = nullptr;
    for (int i = 0; i < tab.count; ++i) {
        const TileEntry<T> &e = tab.entries[i];
        if (e.M == M && e.NW == NW && e.S == S && e.IL == il && e.VEC == vec && e.ALDS == alds &&
            e.PAD == pad && e.BIG == big) {
            if (e.NT == nt && e.XP == xp) return &e;
            if (!e.XP) any = &e;   // padded kernels have no streaming/transposed variant
        }
    }
    return any;
}

template <typename T>
static const TileEntry<T> *choose_tile(int algo, int layout, int64_t n, bool vec_ok,
                                       bool big = false) {
    TileTable<T> tab = tile_table<T>();
    const bool il = layout == kInterleaved;
    // register-packed CR fixes R rows per lane for the reference's per-system layout;
    // interleaved batches use the tuned shape with the same CR interface
    const int fM = il ? 0 : forced_M(algo);
    if (!fM && !big) {
        std::call_once(g_tune_once, load_tuning);
        const int prec = sizeof(T) == 4 ? TRIDIAG_SINGLE : TRIDIAG_DOUBLE;
        // packed systems are solved as contiguous ones (system stride 4n): same rows
        const int tl = layout == kPacked ? kContiguous : layout;
        // rows tuned with the algorithm's own interface solver (family = its id) first,
        // then the default rows (family 0, tuned with Wang-Mou merges)
        const int fam = algo_spec(algo).iface;
        const int fams[2] = {fam, 0};
        for (int fi = 0; fi < (fam == 0 ? 1 : 2); ++fi)
        for (const TuneRow &r : g_tune) {
            if (r.prec != prec || r.layout != tl || r.family != fams[fi] || n > r.nmax) continue;
            if (r.vec && !vec_ok) continue;
            const int nws = il ? r.NW : r.NW / r.S;
            const int64_t cap = (int64_t)(il ? 64 / r.S : 64) * nws * r.M;
            if (cap < n || cap > 2 * n) continue;   // bucket too coarse for this n
            const TileEntry<T> *e = nullptr;
            if (algo_spec(algo).lds_rows)      // LDS-staged variant of the tuned shape, if built
                e = find_tile<T>(r.M, r.NW, r.S, il, r.vec, true, cap != n, false, r.nt, false);
            if (e && lds_need(*e, algo) > max_lds_per_block()) e = nullptr;
            if (!e) e = find_tile<T>(r.M, r.NW, r.S, il, r.vec, r.alds, cap != n, false, r.nt, r.xp);
            if (e && lds_need(*e, algo) <= max_lds_per_block()) return e;
        }
    }
    const TileEntry<T> *best = nullptr;
    long long best_key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < tab.count; ++i) {
        const TileEntry<T> &e = tab.entries[i];
        if (e.IL != il || e.cap < n || e.BIG != big) continue;
        if (!e.BIG && e.PAD != (e.cap != n)) continue;   // exact tiles take the unpredicated kernel
        if (e.VEC && !vec_ok) continue;
        if (fM && e.M != fM) continue;
        if (lds_need(e, algo) > max_lds_per_block()) continue;
        const int nws = il ? e.NW : e.NW / e.S;
        // Register-heavy segments keep one coefficient in LDS; so do the LDS-staged
        // ("shmem") algorithms where the shape has such a kernel.
        const bool want_alds = algo_spec(algo).lds_rows || e.M * (int)sizeof(T) >= 256 ||
                               (e.NW >= 16 && e.M * (int)sizeof(T) >= 128);
        long long key[8] = {
            il ? (e.S == pref_S_il<T>() ? 0 : 1) : 0,  // 32-byte row groups in IL
            (long long)e.cap,                          // least padding
            (long long)std::abs(e.M - pref_M<T>(il)),  // measured sweet spot per layout
            il ? 0 : -(long long)nws,                  // then more waves per system
            e.ALDS == want_alds ? 0 : 1,
            e.VEC ? 0 : 1,
            e.XP ? 0 : 1,
            e.NT ? 0 : 1};
        if (!best || std::lexicographical_compare(key, key + 8, best_key, best_key + 8)) {
            best = &e;
            std::copy(key, key + 8, best_key);
        }
    }
    return best;
}

template <typename T> static bool aligned16(const void *p) {
    return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

// Byte extent of one coefficient array if 32-bit buffer offsets can address it.
template <typename T> static int32_t buffer_extent(int64_t n, int64_t batch, int64_t sstride,
                                                   int layout) {
    const int64_t elems = (layout == kInterleaved) ? n * batch : batch * sstride;
    const int64_t bytes = elems * (int64_t)sizeof(T);
    return bytes < (int64_t)0x7fffff00 ? (int32_t)bytes : 0;
}

// Largest n any on-chip tile of the contiguous layouts holds.
template <typename T> static int64_t max_onchip_rows() {
    static const int64_t v = [] {
        TileTable<T> tab = tile_table<T>();
        int64_t m = 0;
        for (int i = 0; i < tab.count; ++i)
            if (!tab.entries[i].IL && tab.entries[i].PAD) m = std::max<int64_t>(m, tab.entries[i].cap);
        return m;
    }();
    return v;
}

// Bytes of input (four arrays) that one group of the chunked solve covers
// ($TRIDIAG_LARGE_GROUP_BYTES overrides).  Measured on MI355X at 2^20 x 16 and
// 2^16 x 256 FP64: groups small enough for pass 3 to reread its chunks from the
// 256 MB last-level cache (64-256 MB) ran 10-40% slower than one group, so the
// default covers 1 GiB per group.
static int64_t large_group_bytes() {
    static const int64_t v = [] {
        if (const char *e = getenv("TRIDIAG_LARGE_GROUP_BYTES")) {
            const long long x = atoll(e);
            if (x > 0) return (int64_t)x;
        }
        return (int64_t)1 << 30;
    }();
    return v;
}

// Contiguous systems above this many rows take the chunked solve even where an
// on-chip tile holds them ($TRIDIAG_LARGE_MIN_ROWS overrides).  Default: two
// chunks (16384 rows FP32, 8192 FP64).  Measured on MI355X at 2^24 unknowns: the
// tiles above that size hold so much state per lane that they ran at 1.3-2.4 TB/s,
// the chunked solve at 3.1-3.2 TB/s; at or below it the tiles ran 4.4-4.9 TB/s.
template <typename T> static int64_t large_min_rows() {
    static const int64_t v = [] {
        if (const char *e = getenv("TRIDIAG_LARGE_MIN_ROWS")) {
            const long long x = atoll(e);
            if (x > 0) return (int64_t)x;
        }
        return (int64_t)2 * large_chunk_rows(sizeof(T) == 4 ? TRIDIAG_SINGLE : TRIDIAG_DOUBLE, 0);
    }();
    return v;
}

// A fixed device buffer that a solve uses in place of stream-ordered
// allocations (null: allocate).
struct FixedScratch {
    void *ptr;
    size_t bytes;
};

// Systems above the on-chip capacity: the three-pass chunked solve, in groups
// of systems, with scratch for the chunk outer forms: stream-ordered, or the
// fixed buffer `fx` (groups then shrink until their scratch fits it).
template <typename T>
static int large_impl(int algo, int64_t n, int64_t batch, int64_t sstride, T *a, T *b, T *c, T *d,
                      hipStream_t st, const FixedScratch *fx = nullptr) {
    const int prec = sizeof(T) == 4 ? TRIDIAG_SINGLE : TRIDIAG_DOUBLE;
    const int shape = 0;
    const int64_t C = large_chunk_rows(prec, shape);
    const int64_t K = (n + C - 1) / C;
    if (K > 1024)
        return fail(TRIDIAG_EUNSUPPORTED, "n_eqt=" + std::to_string(n) + " needs more than 1024 "
                                          "chunks of " + std::to_string(C) + " rows");
    const int64_t per_sys = 4 * n * (int64_t)sizeof(T);
    int64_t G = std::max<int64_t>(1, std::min<int64_t>({batch, 65535,
                                                        large_group_bytes() / per_sys}));
    const int64_t sys_scratch = K * 8 * (int64_t)sizeof(T);
    void *scratch = nullptr;
    hipError_t e = hipSuccess;
    if (fx) {
        if ((int64_t)fx->bytes < sys_scratch)
            return fail(TRIDIAG_EUNSUPPORTED, "n_eqt=" + std::to_string(n) + " needs " +
                                                  std::to_string(sys_scratch) + " bytes of scratch "
                                                  "per system, more than the fixed " +
                                                  std::to_string(fx->bytes));
        G = std::min<int64_t>(G, (int64_t)fx->bytes / sys_scratch);
        scratch = fx->ptr;
    } else {
        e = hipMallocAsync(&scratch, (size_t)(G * sys_scratch), st);
        if (e != hipSuccess) return hip_fail(e, "chunked solve scratch");
    }
    const AlgoSpec sp = algo_spec(algo);
    int rc = TRIDIAG_OK;
    for (int64_t g0 = 0; g0 < batch && rc == TRIDIAG_OK; g0 += G) {
        const int64_t cnt = std::min(G, batch - g0);
        const int64_t o = g0 * sstride;
        T *red = reinterpret_cast<T *>(scratch);
        LargeParams<T> p{a + o, b + o, c + o, d + o, n, cnt, sstride, (int32_t)K, sp.iface,
                         red, red + G * K * 6};
        if constexpr (sizeof(T) == 4) e = launch_large_f32(p, shape, st);
        else e = launch_large_f64(p, shape, st);
        if (e == hipErrorInvalidConfiguration)
            rc = fail(TRIDIAG_EUNSUPPORTED, "interface solver of algorithm " + std::to_string(algo) +
                                                " needs more LDS than the chunk kernel leaves");
        else if (e != hipSuccess)
            rc = hip_fail(e, "chunked solve launch");
    }
    if (!fx) {
        hipError_t ef = hipFreeAsync(scratch, st);
        if (rc == TRIDIAG_OK && ef != hipSuccess) rc = hip_fail(ef, "chunked solve scratch");
    }
    return rc;
}

// ---------------------------------------------------------------- cooperative solve
template <typename T> static CoopTable<T> coop_table();
template <> CoopTable<float> coop_table<float>() { return coop_table_f32(); }
template <> CoopTable<double> coop_table<double>() { return coop_table_f64(); }

static int64_t env_int(const char *name, int64_t dflt) {
    const char *e = getenv(name);
    return e ? (int64_t)atoll(e) : dflt;
}

// Interleaved systems of at least this many rows take the cooperative solve
// ($TRIDIAG_COOP_MIN_ROWS overrides for both precisions; a negative value turns
// the cooperative solve off for both layouts).  Measured on MI355X (groups of a
// round spread over the batch) against the one-workgroup tiles, at about 2^24
// unknowns: FP32 768 rows 0.082 vs 0.093 ms, 1024 rows 0.082 vs 0.090 ms, 1280
// rows 0.082 vs 0.143 ms, 1536 rows 0.083 vs 0.117 ms; FP64 1024 rows 0.181 vs
// 0.218 ms, 1536 rows 0.166 vs 0.251 ms; at 512 rows the tile wins (FP32 0.084 vs
// 0.078 ms).  At 4M unknowns FP32 1024 rows is a draw (0.028 vs 0.025 ms), FP64
// 0.045 vs 0.054 ms.  Above 4096 rows (no interleaved tile) it replaces the solve
// through a contiguous copy: 8192 rows 0.123 vs 0.201 ms (FP32), 0.257 vs 0.406 ms
// (FP64); 16384 rows (64 blocks) 0.167 vs 0.213 ms (FP32), and FP64 12288 rows
// 0.266 vs 0.363 ms.
static int64_t coop_min_rows_il(int prec) {
    static const int64_t env = env_int("TRIDIAG_COOP_MIN_ROWS", 0);
    if (env != 0) return env;
    return 1024;
}

// Interleaved rows whose pitch is not a multiple of 128 bytes: the groups' row
// pieces straddle cache lines, which costs the cooperative solve more than the
// tiles (FP32 4096 x 4001 0.168 vs 0.117 ms, 2048 x 8191 0.157 vs 0.121 ms; FP64
// 0.349 vs 0.230, 0.316 vs 0.227 ms), so such batches keep the tile wherever one
// holds n.
template <typename T> static bool il_rows_on_lines(int64_t batch) {
    return (batch * (int64_t)sizeof(T)) % 128 == 0;
}
// Polls of a cooperative wait before it gives up ($TRIDIAG_COOP_SPIN).
static int32_t coop_spin_limit() {
    static const int32_t v = (int32_t)std::max<int64_t>(0, std::min<int64_t>(
        env_int("TRIDIAG_COOP_SPIN", (int64_t)1 << 22), 0x7fffffff));
    return v;
}

// Tag of the next cooperative solve: a 64-bit sequence from a random start (time,
// process and a random device), never the same for two solves of the process, and
// with a different start in every process, so that the flag words a solve finds
// in its scratch (left by earlier solves or never written) do not carry its tag.
static uint64_t coop_next_tag() {
    static std::atomic<uint64_t> next{[] {
        std::random_device rd;
        uint64_t s = ((uint64_t)rd() << 32) ^ rd();
        s ^= (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count() * 0x9E3779B97F4A7C15ull;
        s ^= (uint64_t)getpid() << 17;
        return s | 1;
    }()};
    uint64_t t = next.fetch_add(1, std::memory_order_relaxed);
    if (t == 0) t = next.fetch_add(1, std::memory_order_relaxed);   // keep 0 (fresh memory) out
    return t;
}

// Which cooperative kernel runs: the entry's choice, or $TRIDIAG_COOP_PIPE (1: the
// one that waits one task later, 0: the one that waits before its next task).
static int32_t coop_pipelined(bool entry_default) {
    static const int64_t v = env_int("TRIDIAG_COOP_PIPE", -1);
    return v < 0 ? (entry_default ? 1 : 0) : (v != 0 ? 1 : 0);
}

// Interleaved groups spread over the batch within a round (coop_group): the W
// groups a round runs, or 0 with $TRIDIAG_COOP_SPREAD=0 (neighbouring groups).
static int32_t coop_spread(int groups_per_round) {
    static const int64_t v = env_int("TRIDIAG_COOP_SPREAD", 1);
    return v != 0 ? groups_per_round : 0;
}

// The entry of the cooperative table for (layout, n): the first whose blocks
// divide n (interleaved) and whose kmax holds K = blocks per system (K >= 2);
// interleaved systems up to 64 blocks (16384 rows).  Contiguous systems take it up to
// 32 blocks (FP32 131072 rows, FP64 65536): measured on MI355X at 2^24 unknowns
// against the chunked solve, FP32 20000 rows 0.076 vs 0.113 ms and 131072 rows
// 0.088 vs 0.107 ms, FP64 20000 rows 0.147 vs 0.209 ms and 65536 rows 0.173 vs
// 0.206 ms.  Far beyond (2^20 rows, 512 blocks) every group would wait on hundreds
// of workgroups: an early version ran 1.6 ms against the chunked 0.21 ms.
template <typename T>
static const CoopEntry<T> *choose_coop(bool il, int64_t n, int &K) {
    CoopTable<T> tab = coop_table<T>();
    static const int64_t forced = env_int("TRIDIAG_COOP_ENTRY", -1);   // experiments
    for (int i = 0; i < tab.count; ++i) {
        if (forced >= 0 && i != forced) continue;
        const CoopEntry<T> &e = tab.entries[i];
        if (e.IL != il) continue;
        const int64_t k = il ? (n % e.R == 0 ? n / e.R : 0) : (n + e.R - 1) / e.R;
        if (k < 2 || k > e.kmax) continue;
        K = (int)k;
        return &e;
    }
    return nullptr;
}

// Bulk-row form of a reference id (tri_bulk.hip): form 0 PCR-pThomas, 1
// PCR-half-pThomas, 2 WM-pGE, 3 WM-half-pGE, 4 CR through global memory, 5 CR over
// R rows per lane (R = forced_M(algo)), and whether its exchanges go through LDS
// (the shared-memory ids) or ds_bpermute (the register ids).  False for ids
// without one.
static bool bulk_form(int algo, int &form, bool &lds) {
    switch (algo) {
    case kCR: form = 4; lds = false; return true;
    case kCRReg4: case kCRReg8: case kCRReg16: form = 5; lds = true; return true;
    case kPCR5Shmem: form = 0; lds = true; return true;
    case kPCR4Shmem: form = 1; lds = true; return true;
    case kPCR5Reg: form = 0; lds = false; return true;
    case kPCR4Reg: form = 1; lds = false; return true;
    case kWM5Shmem: form = 2; lds = true; return true;
    case kWM4Shmem: form = 3; lds = true; return true;
    case kWM5Reg: form = 2; lds = false; return true;
    case kWM4Reg: form = 3; lds = false; return true;
    default: return false;
    }
}
// $TRIDIAG_FORM=bulk: ids 2-13 in the contiguous and packed layouts run their
// bulk-row forms (also through the reference symbols) instead of the partition
// kernel.
static bool bulk_forms_requested() {
    static const bool v = [] {
        const char *e = getenv("TRIDIAG_FORM");
        return e && std::string(e) == "bulk";
    }();
    return v;
}
// The CR form keeps its reduced levels in a stream-ordered scratch buffer of
// 4 * n * batch words, freed after the launch.
template <typename T>
static hipError_t bulk_launch(int algo, const T *a, const T *b, T *c, T *d, int64_t n,
                              int64_t batch, int64_t sstride, int32_t bb, hipStream_t st) {
    int form = 0;
    bool lds = false;
    if (!bulk_form(algo, form, lds)) return hipErrorInvalidValue;
    if (form == 5) {
        if constexpr (sizeof(T) == 4)
            return launch_bulk_crreg_f32(forced_M(algo), (const float *)a, (const float *)b,
                                         (float *)c, (float *)d, n, batch, sstride, bb, st);
        else
            return launch_bulk_crreg_f64(forced_M(algo), (const double *)a, (const double *)b,
                                         (double *)c, (double *)d, n, batch, sstride, bb, st);
    }
    if (form == 4) {
        if (batch == 0) return hipSuccess;
        void *scratch = nullptr;
        hipError_t e = hipMallocAsync(&scratch, (size_t)(4 * n * batch) * sizeof(T), st);
        if (e != hipSuccess) return e;
        if constexpr (sizeof(T) == 4)
            e = launch_bulk_cr_f32((const float *)a, (const float *)b, (float *)c, (float *)d, n,
                                   batch, sstride, bb, (float *)scratch, st);
        else
            e = launch_bulk_cr_f64((const double *)a, (const double *)b, (double *)c,
                                   (double *)d, n, batch, sstride, bb, (double *)scratch, st);
        const hipError_t ef = hipFreeAsync(scratch, st);
        return e != hipSuccess ? e : ef;
    }
    if constexpr (sizeof(T) == 4)
        return launch_bulk_f32(form, lds, (const float *)a, (const float *)b, (float *)c,
                               (float *)d, n, batch, sstride, bb, st);
    else
        return launch_bulk_f64(form, lds, (const double *)a, (const double *)b, (double *)c,
                               (double *)d, n, batch, sstride, bb, st);
}

// How solve_impl solves one batch (decided without launching anything).
enum PlanKind {
    kPlanLane = 0,           // one system per lane, in registers (short systems)
    kPlanThomas = 1,         // the Thomas recurrence (algorithm 1)
    kPlanTile = 2,           // on chip, one workgroup per system or per S systems
    kPlanCoop = 3,           // on chip, K workgroups per group of systems
    kPlanChunked = 4,        // three passes over chunks through global memory
    kPlanViaContiguous = 5,  // interleaved systems solved through a contiguous copy
    kPlanBulk = 6,           // the reference's own form of ids 2-13 (bulk rows, CR levels)
};
template <typename T> struct SolvePlan {
    int kind = kPlanTile;
    const TileEntry<T> *tile = nullptr;
    const CoopEntry<T> *coop = nullptr;
    int K = 0;                 // cooperative: row blocks per system
    int grid = 0;              // cooperative: persistent workgroups
    size_t forms_bytes = 0;    // cooperative: scratch for the block forms, then flags
    size_t scratch_bytes = 0;
};

// Whether the cooperative row-block solve takes (algo, layout, n, batch): the
// Wang-Mou interface family, an entry whose blocks fit (interleaved: dividing n),
// 32-bit offsets, and a device that holds at least K of its workgroups at once.
// fixed_bytes > 0: the scratch must fit a fixed buffer of that size.
template <typename T>
static bool coop_plan(int algo, int layout, int64_t n, int64_t batch, size_t fixed_bytes,
                      SolvePlan<T> &pl) {
    const int prec = sizeof(T) == 4 ? TRIDIAG_SINGLE : TRIDIAG_DOUBLE;
    if (coop_min_rows_il(prec) < 0 || algo_spec(algo).iface != kIfaceWM) return false;
    const bool il = layout == kInterleaved;
    if (il && n < coop_min_rows_il(prec)) return false;
    int K = 0;
    const CoopEntry<T> *e = choose_coop<T>(il, n, K);
    if (!e) return false;
    // interleaved: 32-bit offsets into whole arrays; contiguous: into one system
    const int64_t elems = il ? n * batch : n;
    if (elems * (int64_t)sizeof(T) >= (int64_t)0x7fffff00) return false;
    const int64_t ngroups = (batch + e->S - 1) / e->S;
    const int64_t ntask = ngroups * K;
    if (ntask > 0x7fffffff) return false;
    // every task of a round must be resident: P workgroups, a multiple of K
    const int cap = e->max_grid(coop_pipelined(e->pipe) != 0);
    if (cap < K) return false;
    const int64_t P = std::min<int64_t>(ntask, (cap / K) * (int64_t)K);
    const size_t forms = (((size_t)ntask * e->S * 6 * sizeof(T)) + 255) & ~(size_t)255;
    const size_t bytes = forms + (size_t)ngroups * kCoopFlagStride * sizeof(uint64_t);
    if (fixed_bytes && bytes > fixed_bytes) return false;
    pl.kind = kPlanCoop;
    pl.coop = e;
    pl.K = K;
    pl.grid = (int)P;
    pl.forms_bytes = forms;
    pl.scratch_bytes = bytes;
    return true;
}

template <typename T>
static int coop_run(const SolvePlan<T> &pl, int layout, int64_t n, int64_t batch, int64_t sstride,
                    T *a, T *b, T *c, T *d, hipStream_t st, const FixedScratch *fx) {
    const bool il = layout == kInterleaved;
    const CoopEntry<T> &e = *pl.coop;
    const int64_t ngroups = (batch + e.S - 1) / e.S;
    void *scratch = nullptr;
    if (fx) {
        scratch = fx->ptr;
    } else {
        const hipError_t ea = hipMallocAsync(&scratch, pl.scratch_bytes, st);
        if (ea != hipSuccess) return hip_fail(ea, "cooperative solve scratch");
    }
    CoopParams<T> p{a, b, c, d, n, batch, il ? 1 : sstride, il ? batch : 1,
                    (int32_t)(il ? n * batch * (int64_t)sizeof(T) : 0), (int32_t)pl.K,
                    (int32_t)ngroups, coop_spin_limit(), coop_pipelined(e.pipe),
                    il ? coop_spread(pl.grid / pl.K) : 0, reinterpret_cast<T *>(scratch),
                    reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(scratch) +
                                                 pl.forms_bytes),
                    coop_next_tag()};
    int rc = TRIDIAG_OK;
    const hipError_t err = e.launch(p, pl.grid, st);
    if (err != hipSuccess) rc = hip_fail(err, "cooperative solve launch");
    if (!fx) {
        const hipError_t ef = hipFreeAsync(scratch, st);
        if (rc == TRIDIAG_OK && ef != hipSuccess) rc = hip_fail(ef, "cooperative solve scratch");
    }
    return rc;
}

template <typename T>
static int convert_impl(int src_layout, int dst_layout, int64_t n, int64_t batch,
                        const T *const src[4], T *const dst[4], hipStream_t st);
template <typename T>
static int solve_impl(int algo, int layout, int64_t n, int64_t batch, T *a, T *b, T *c, T *d,
                      hipStream_t st, const FixedScratch *fx = nullptr);

// Interleaved systems longer than the interleaved tiles hold: converted to the
// contiguous layout in stream-ordered scratch, solved there (the chunked solve for
// long n), and x converted back.  Moves 8w + solve + 2w bytes per unknown.
template <typename T>
static int interleaved_via_contiguous(int algo, int64_t n, int64_t batch, T *a, T *b, T *c, T *d,
                                      hipStream_t st) {
    const int64_t elems = n * batch;
    void *scratch = nullptr;
    hipError_t e = hipMallocAsync(&scratch, (size_t)(4 * elems) * sizeof(T), st);
    if (e != hipSuccess) return hip_fail(e, "interleaved solve scratch");
    T *w = reinterpret_cast<T *>(scratch);
    const T *src[4] = {a, b, c, d};
    T *dst[4] = {w, w + elems, w + 2 * elems, w + 3 * elems};
    int rc = convert_impl<T>(kInterleaved, kContiguous, n, batch, src, dst, st);
    if (rc == TRIDIAG_OK) rc = solve_impl<T>(algo, kContiguous, n, batch, dst[0], dst[1], dst[2],
                                             dst[3], st);
    if (rc == TRIDIAG_OK) {
        const T *xs[4] = {nullptr, nullptr, nullptr, dst[3]};
        T *xd[4] = {nullptr, nullptr, nullptr, d};
        rc = convert_impl<T>(kContiguous, kInterleaved, n, batch, xs, xd, st);
    }
    const hipError_t ef = hipFreeAsync(scratch, st);
    if (rc == TRIDIAG_OK && ef != hipSuccess) rc = hip_fail(ef, "interleaved solve scratch");
    return rc;
}

// Largest n for which the automatic choice takes the one-system-per-lane Thomas
// kernel instead of an on-chip tile ($TRIDIAG_LANE_ROWS overrides).
template <typename T> static int64_t lane_auto_rows(int layout) {
    static const int64_t env = [] {
        const char *e = getenv("TRIDIAG_LANE_ROWS");
        return e ? (int64_t)atoll(e) : (int64_t)-1;
    }();
    if (env >= 0) return env;
    (void)layout;
    return lane_thomas_max_rows(sizeof(T) == 4 ? TRIDIAG_SINGLE : TRIDIAG_DOUBLE);
}

// The path solve_impl takes for one batch below 2 GiB per array.  algo may be
// TRIDIAG_ALGO_AUTO.  vec_ok: rows are 16-byte aligned (contiguous layouts).
// fixed_bytes > 0: scratch must fit a fixed buffer of that many bytes.
template <typename T>
static SolvePlan<T> make_plan(int algo, int layout, int64_t n, int64_t batch, int64_t sstride,
                              bool vec_ok, size_t fixed_bytes) {
    SolvePlan<T> pl;
    const bool il = layout == kInterleaved;
    const int prec = sizeof(T) == 4 ? TRIDIAG_SINGLE : TRIDIAG_DOUBLE;
    const int32_t bb = buffer_extent<T>(n, batch, sstride, layout);
    // short systems: one system per lane in registers (Thomas), for the Thomas id and
    // for the automatic choice up to lane_auto_rows()
    if (bb != 0 && n <= lane_thomas_max_rows(prec) && (il || layout == kContiguous) &&
        (algo == kThomas || (algo == TRIDIAG_ALGO_AUTO && n <= lane_auto_rows<T>(layout)))) {
        pl.kind = kPlanLane;
        return pl;
    }
    if (algo == kThomas) {
        pl.kind = kPlanThomas;
        return pl;
    }
    int form = 0;
    bool lds = false;
    // (the register CR form holds at most 1024 R rows, as the reference's does;
    // longer systems of those ids keep the default path)
    if (!il && bb != 0 && bulk_forms_requested() && bulk_form(algo, form, lds) &&
        (form != 5 || n <= 1024 * (int64_t)forced_M(algo))) {
        pl.kind = kPlanBulk;
        return pl;
    }
    // rows beyond one workgroup: K workgroups per group of systems, on chip
    const bool il_tile_first = il && !il_rows_on_lines<T>(batch) &&
                               choose_tile<T>(algo, layout, n, vec_ok, bb == 0) != nullptr;
    if ((il || n > large_min_rows<T>()) && !il_tile_first &&
        coop_plan<T>(algo, layout, n, batch, fixed_bytes, pl))
        return pl;
    if (!il && n > large_min_rows<T>() && n > large_chunk_rows(prec, 0)) {
        pl.kind = kPlanChunked;
        return pl;
    }
    pl.tile = choose_tile<T>(algo, layout, n, vec_ok, bb == 0);
    if (!pl.tile)   // no tile of this algorithm holds n (e.g. CR_reg R, or long interleaved)
        pl.kind = il ? kPlanViaContiguous : kPlanChunked;
    return pl;
}

template <typename T>
static int solve_impl(int algo, int layout, int64_t n, int64_t batch, T *a, T *b, T *c, T *d,
                      hipStream_t st, const FixedScratch *fx) {
    if (n < 1) return fail(TRIDIAG_EINVAL, "n_eqt must be >= 1");
    if (batch < 0) return fail(TRIDIAG_EINVAL, "n_batch must be >= 0");
    if (layout < 0 || layout > 2) return fail(TRIDIAG_EINVAL, "layout must be 0, 1 or 2");
    if (!a || (layout != kPacked && (!b || !c || !d)))
        return fail(TRIDIAG_EINVAL, "null coefficient pointer");
    if (batch == 0) return TRIDIAG_OK;
    int64_t sstride = n, rstride = 1;
    if (layout == kPacked) {
        sstride = 4 * n;
    } else if (layout == kInterleaved) {
        sstride = 1;
        rstride = batch;
    }
    if (algo != kThomas && layout != kInterleaved &&
        buffer_extent<T>(n, batch, sstride, layout) == 0) {
        // Contiguous systems: solve in chunks whose arrays stay below 2 GiB so
        // the kernels keep 32-bit buffer offsets.
        const int64_t per = std::max<int64_t>(1, ((int64_t)0x7fffff00 / (int64_t)sizeof(T)) / sstride);
        if (per < batch) {
            for (int64_t s0 = 0; s0 < batch; s0 += per) {
                const int64_t cnt = std::min(per, batch - s0);
                const int64_t o = s0 * sstride;
                int rc = solve_impl<T>(algo, layout, n, cnt, a + o,
                                       layout == kPacked ? nullptr : b + o,
                                       layout == kPacked ? nullptr : c + o,
                                       layout == kPacked ? nullptr : d + o, st, fx);
                if (rc != TRIDIAG_OK) return rc;
            }
            return TRIDIAG_OK;
        }
    }
    if (layout == kPacked) {
        b = a + n; c = a + 2 * n; d = a + 3 * n;
    }
    const int V = 16 / (int)sizeof(T);
    const bool vec_ok = layout != kInterleaved && (sstride % V) == 0 && aligned16<T>(a) &&
                        aligned16<T>(b) && aligned16<T>(c) && aligned16<T>(d);
    const SolvePlan<T> pl = make_plan<T>(algo, layout, n, batch, sstride, vec_ok,
                                         fx ? fx->bytes : 0);
    const int32_t bb = buffer_extent<T>(n, batch, sstride, layout);
    const bool il = layout == kInterleaved;
    switch (pl.kind) {
    case kPlanLane: {
        const hipError_t e = (sizeof(T) == 4)
            ? launch_lane_thomas_f32((const float *)a, (const float *)b, (const float *)c, (float *)d,
                                     n, batch, sstride, rstride, il, bb, st)
            : launch_lane_thomas_f64((const double *)a, (const double *)b, (const double *)c,
                                     (double *)d, n, batch, sstride, rstride, il, bb, st);
        return e == hipSuccess ? TRIDIAG_OK : hip_fail(e, "lane Thomas launch");
    }
    case kPlanThomas: {
        if (!il && bb != 0) {
            // lane groups per system, rows as 16-byte vectors (aligned rows below 2 GiB)
            const hipError_t e = (sizeof(T) == 4)
                ? launch_thomas_rows_f32((const float *)a, (const float *)b, (float *)c, (float *)d,
                                         n, batch, sstride, bb, st)
                : launch_thomas_rows_f64((const double *)a, (const double *)b, (double *)c,
                                         (double *)d, n, batch, sstride, bb, st);
            if (e != hipErrorInvalidValue)
                return e == hipSuccess ? TRIDIAG_OK : hip_fail(e, "Thomas launch");
        }
        const hipError_t e = (sizeof(T) == 4)
            ? launch_thomas_f32((const float *)a, (const float *)b, (float *)c, (float *)d, n,
                                batch, sstride, rstride, st)
            : launch_thomas_f64((const double *)a, (const double *)b, (double *)c, (double *)d, n,
                                batch, sstride, rstride, st);
        return e == hipSuccess ? TRIDIAG_OK : hip_fail(e, "Thomas launch");
    }
    case kPlanBulk: {
        if (fx && algo == kCR)
            return fail(TRIDIAG_EUNSUPPORTED, "the global-memory CR form keeps its levels in "
                                              "an allocated buffer");
        const hipError_t e = bulk_launch<T>(algo, a, b, c, d, n, batch, sstride, bb, st);
        return e == hipSuccess ? TRIDIAG_OK : hip_fail(e, "bulk-row launch");
    }
    case kPlanCoop:
        return coop_run<T>(pl, layout, n, batch, sstride, a, b, c, d, st, fx);
    case kPlanChunked:
        return large_impl<T>(algo, n, batch, sstride, a, b, c, d, st, fx);
    case kPlanViaContiguous:
        if (fx)
            return fail(TRIDIAG_EUNSUPPORTED,
                        "interleaved n_eqt=" + std::to_string(n) + " is solved through a "
                        "contiguous copy of all four arrays, which needs an allocation");
        return interleaved_via_contiguous<T>(algo, n, batch, a, b, c, d, st);
    default:
        break;
    }
    const AlgoSpec sp = algo_spec(algo);
    TileParams<T> p{a, b, c, d, n, batch, sstride, rstride, bb, sp.iface, sp.xchg_lds};
    const hipError_t err = pl.tile->launch(p, st);
    if (err == hipErrorInvalidConfiguration)
        return fail(TRIDIAG_EUNSUPPORTED, "interface solver of algorithm " + std::to_string(algo) +
                                              " needs more LDS than the kernel for n_eqt=" +
                                              std::to_string(n) + " leaves");
    return err == hipSuccess ? TRIDIAG_OK : hip_fail(err, "tile launch");
}

// The bulk-row form of algorithm `algo` (ids 2-13) in the contiguous or packed
// layout; batches whose arrays reach 2 GiB run in sub-2-GiB launches.
template <typename T>
static int bulk_impl(int algo, int layout, int64_t n, int64_t batch, T *a, T *b, T *c, T *d,
                     hipStream_t st) {
    int form = 0;
    bool lds = false;
    if (!bulk_form(algo, form, lds))
        return fail(TRIDIAG_EUNSUPPORTED, "algorithm " + std::to_string(algo) +
                                              " has no bulk-row form (ids 2-13 do)");
    if (form == 5 && n > 1024 * (int64_t)forced_M(algo))
        return fail(TRIDIAG_EUNSUPPORTED, "register CR over " + std::to_string(forced_M(algo)) +
                                              " rows per lane holds at most " +
                                              std::to_string(1024 * forced_M(algo)) +
                                              " rows in one workgroup");
    if (n < 1) return fail(TRIDIAG_EINVAL, "n_eqt must be >= 1");
    if (batch < 0) return fail(TRIDIAG_EINVAL, "n_batch must be >= 0");
    if (layout == kInterleaved)
        return fail(TRIDIAG_EUNSUPPORTED, "the bulk-row forms stream one system per wave: "
                                          "layouts 0 and 1 only");
    if (layout != kContiguous && layout != kPacked)
        return fail(TRIDIAG_EINVAL, "layout must be 0, 1 or 2");
    if (!a || (layout == kContiguous && (!b || !c || !d)))
        return fail(TRIDIAG_EINVAL, "null coefficient pointer");
    if (batch == 0) return TRIDIAG_OK;
    const int64_t sstride = layout == kPacked ? 4 * n : n;
    const int64_t per = std::max<int64_t>(1, ((int64_t)0x7fffff00 / (int64_t)sizeof(T)) / sstride);
    if (sstride * (int64_t)sizeof(T) > (int64_t)0x7fffff00)
        return fail(TRIDIAG_EUNSUPPORTED, "one system of 2 GiB or more");
    for (int64_t s0 = 0; s0 < batch; s0 += per) {
        const int64_t cnt = std::min(per, batch - s0);
        T *a0 = a + s0 * sstride;
        T *b0 = layout == kPacked ? a0 + n : b + s0 * sstride;
        T *c0 = layout == kPacked ? a0 + 2 * n : c + s0 * sstride;
        T *d0 = layout == kPacked ? a0 + 3 * n : d + s0 * sstride;
        const int32_t bb = buffer_extent<T>(n, cnt, sstride, layout);
        const hipError_t e = bulk_launch<T>(algo, a0, b0, c0, d0, n, cnt, sstride, bb, st);
        if (e != hipSuccess) return hip_fail(e, "bulk-row launch");
    }
    return TRIDIAG_OK;
}

// ---------------------------------------------------------------- no-allocation solve
// Scratch of tridiag_solve_noalloc: the chunked solve's per-chunk outer forms and
// boundary values (8 words per chunk of every system of one launch group), or the
// cooperative solve's block forms and publication flags.  One copy per GPU, at a
// device address of its own, held until the process ends.  Every
// tridiag_solve_noalloc call on a GPU shares that GPU's copy; calls that overlap
// on one GPU race.  The chunked solve writes every word it reads; the cooperative
// solve's flags only count when they carry the solve's own tag, which no earlier
// call used, so the contents left by an earlier call need no clearing.
constexpr size_t kNoAllocScratchBytes = TRIDIAG_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];

// Device address of the current GPU's `scratch`, looked up on the first call on
// that GPU.  (The array's name is a host-side stand-in; only this address may be
// handed to kernels.)
static int noalloc_scratch(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
        return fail(TRIDIAG_EUNSUPPORTED, "device index " + std::to_string(dev) + " beyond " +
                                              std::to_string(kMaxDevices));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(TRIDIAG_EHIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_dev[dev] = ptr;
    }
    *out = g_scratch_dev[dev];
    return TRIDIAG_OK;
}

// ---------------------------------------------------------------- vendor comparator
static hipsparseHandle_t g_sparse = nullptr;
static void *g_sparse_buf = nullptr;
static size_t g_sparse_bytes = 0;

template <typename T> static int sparse_init(T *a, T *b, T *c, T *d, int n, int batch) {
    if (!g_sparse) {
        if (hipsparseCreate(&g_sparse) != HIPSPARSE_STATUS_SUCCESS)
            return fail(TRIDIAG_EHIP, "hipsparseCreate failed");
    }
    size_t bytes = 0;
    hipsparseStatus_t s;
    if (sizeof(T) == 4)
        s = hipsparseSgtsv2StridedBatch_bufferSizeExt(g_sparse, n, (float *)a, (float *)b,
                                                      (float *)c, (float *)d, batch, n, &bytes);
    else
        s = hipsparseDgtsv2StridedBatch_bufferSizeExt(g_sparse, n, (double *)a, (double *)b,
                                                      (double *)c, (double *)d, batch, n, &bytes);
    if (s != HIPSPARSE_STATUS_SUCCESS) return fail(TRIDIAG_EHIP, "gtsv2StridedBatch_bufferSizeExt");
    if (!g_sparse_buf || bytes > g_sparse_bytes) {
        if (g_sparse_buf) (void)hipFree(g_sparse_buf);
        g_sparse_buf = nullptr;
        g_sparse_bytes = 0;
        hipError_t e = hipMalloc(&g_sparse_buf, bytes ? bytes : 16);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(gtsv buffer)");
        g_sparse_bytes = bytes;
    }
    return TRIDIAG_OK;
}

template <typename T> static int sparse_solve(T *a, T *b, T *c, T *d, int n, int batch) {
    if (!g_sparse || !g_sparse_buf) return fail(TRIDIAG_EINVAL, "cuSparse_init was not called");
    hipsparseStatus_t s;
    if (sizeof(T) == 4)
        s = hipsparseSgtsv2StridedBatch(g_sparse, n, (float *)a, (float *)b, (float *)c,
                                        (float *)d, batch, n, g_sparse_buf);
    else
        s = hipsparseDgtsv2StridedBatch(g_sparse, n, (double *)a, (double *)b, (double *)c,
                                        (double *)d, batch, n, g_sparse_buf);
    if (s != HIPSPARSE_STATUS_SUCCESS) return fail(TRIDIAG_EHIP, "gtsv2StridedBatch failed");
    return TRIDIAG_OK;
}

// Synchronous reference-style call: solve on the null stream, then wait.
template <typename T> static void ref_call(int algo, int layout, T *a, T *b, T *c, T *d, int n,
                                           int batch) {
    int rc = solve_impl<T>(algo, layout, n, batch, a, b, c, d, nullptr);
    if (rc == TRIDIAG_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = hip_fail(e, "hipDeviceSynchronize");
    }
    if (rc != TRIDIAG_OK) {
        fprintf(stderr, "tridiag_amd: %s (algo %d) failed: %s\n",
                sizeof(T) == 4 ? kNamesSingle[algo] : kNamesDouble[algo], algo,
                g_last_error.c_str());
    }
}

template <typename T>
static int tile_info_impl(int index, int *M, int *NW, int *S, int *il, int *vec, int *alds,
                          int64_t *cap, int *big, int *pad, int *nt, int *xp) {
    TileTable<T> tab = tile_table<T>();
    if (index < 0 || index >= tab.count) return fail(TRIDIAG_EINVAL, "tile index out of range");
    const TileEntry<T> &e = tab.entries[index];
    if (M) *M = e.M;
    if (NW) *NW = e.NW;
    if (S) *S = e.S;
    if (il) *il = e.IL;
    if (vec) *vec = e.VEC;
    if (alds) *alds = e.ALDS;
    if (cap) *cap = e.cap;
    if (big) *big = e.BIG;
    if (pad) *pad = e.PAD;
    if (nt) *nt = e.NT;
    if (xp) *xp = e.XP;
    return TRIDIAG_OK;
}

template <typename T>
static int solve_tile_impl(int index, int iface, int layout, int64_t n, int64_t batch, T *a,
                           T *b, T *c, T *d, hipStream_t st) {
    TileTable<T> tab = tile_table<T>();
    if (index < 0 || index >= tab.count) return fail(TRIDIAG_EINVAL, "tile index out of range");
    const TileEntry<T> &e = tab.entries[index];
    if (!iface_code_valid(iface, e.cap / e.M))
        return fail(TRIDIAG_EINVAL, "interface solver code out of range (family 0-4; a PCR depth "
                                    "only for family 3, at most log2 of the tile's segments)");
    if (lds_need_iface(e, iface, 0) > max_lds_per_block())
        return fail(TRIDIAG_EUNSUPPORTED, "the interface solver's LDS does not fit beside this tile");
    if (n < 1 || batch < 0 || layout < 0 || layout > 2) return fail(TRIDIAG_EINVAL, "bad shape");
    if (!a || (layout != kPacked && (!b || !c || !d))) return fail(TRIDIAG_EINVAL, "null pointer");
    if (e.IL != (layout == kInterleaved)) return fail(TRIDIAG_EINVAL, "tile/layout mismatch");
    if (e.cap < n) return fail(TRIDIAG_EUNSUPPORTED, "n_eqt exceeds the tile capacity");
    if (!e.PAD && e.cap != n) return fail(TRIDIAG_EINVAL, "unpadded tile needs n_eqt == capacity");
    if (batch == 0) return TRIDIAG_OK;
    int64_t sstride = n, rstride = 1;
    if (layout == kPacked) {
        b = a + n; c = a + 2 * n; d = a + 3 * n;
        sstride = 4 * n;
    } else if (layout == kInterleaved) {
        sstride = 1;
        rstride = batch;
    }
    const int V = 16 / (int)sizeof(T);
    if (e.VEC && !((sstride % V) == 0 && aligned16<T>(a) && aligned16<T>(b) && aligned16<T>(c) &&
                   aligned16<T>(d)))
        return fail(TRIDIAG_EINVAL, "vector tile needs 16-byte aligned rows");
    const int32_t bb = buffer_extent<T>(n, batch, sstride, layout);
    if (bb == 0 && !e.BIG) return fail(TRIDIAG_EUNSUPPORTED, "arrays of 2 GiB or more need a BIG tile");
    TileParams<T> p{a, b, c, d, n, batch, sstride, rstride, bb, iface, 0};
    hipError_t err = e.launch(p, st);
    return err == hipSuccess ? TRIDIAG_OK : hip_fail(err, "tile launch");
}


// ---------------------------------------------------------------- ragged batches
// Bucket choice for a system of n rows: the autotuned row of its size bucket,
// else the default shape with the Wang-Mou interface solver.
static void ragged_choice(int prec, int64_t n, int &shape, int &iface) {
    std::call_once(g_tune_once, load_tuning);
    for (const RaggedTuneRow &r : g_rtune) {
        if (r.prec != prec || n > r.nmax) continue;
        if (r.shape < 0 || r.shape >= kRaggedShapes || ragged_cap(r.shape) < n) continue;
        if (!iface_code_valid(r.iface, ragged_segments(r.shape))) continue;
        shape = r.shape;
        iface = r.iface;
        return;
    }
    shape = ragged_default_shape(n);
    iface = kIfaceWM;
}

// Plans written by ragged_plan_impl, by work-buffer address: the system count
// and precision each one was made for.  The kernel ignores a work buffer whose
// header does not match its launch (it then leaves d untouched), so the host
// checks the same fields here and reports the mismatch instead.
struct RaggedPlanKey {
    int64_t n_sys;
    int prec;
};
static std::mutex g_rplan_mu;
static std::vector<std::pair<const void *, RaggedPlanKey>> g_rplans;

static void ragged_plan_record(const void *work, int64_t n_sys, int prec) {
    std::lock_guard<std::mutex> lk(g_rplan_mu);
    for (auto &e : g_rplans)
        if (e.first == work) {
            e.second = {n_sys, prec};
            return;
        }
    g_rplans.push_back({work, {n_sys, prec}});
}

static int ragged_plan_check(const void *work, int64_t n_sys, int prec) {
    std::lock_guard<std::mutex> lk(g_rplan_mu);
    for (const auto &e : g_rplans) {
        if (e.first != work) continue;
        if (e.second.prec != prec)
            return fail(TRIDIAG_EINVAL, std::string("the work buffer holds a plan for ") +
                                            (e.second.prec == TRIDIAG_SINGLE ? "single" : "double") +
                                            " precision");
        if (e.second.n_sys != n_sys)
            return fail(TRIDIAG_EINVAL, "the work buffer holds a plan for " +
                                            std::to_string(e.second.n_sys) + " systems, not " +
                                            std::to_string(n_sys));
        return TRIDIAG_OK;
    }
    return fail(TRIDIAG_EINVAL, "the work buffer holds no plan (call tridiag_ragged_plan first)");
}

// Host planner (ragged_schedule: validates the offsets, buckets the systems, one
// record per work item, largest bucket first) and the upload into `work`.
static int ragged_plan_impl(int prec, int64_t n_sys, const int64_t *off, int force_shape,
                            int force_iface, void *work, hipStream_t st) {
    if (!work) return fail(TRIDIAG_EINVAL, "null work buffer");
    std::vector<char> host;
    std::string err;
    const int rc = ragged_schedule(prec, n_sys, off, force_shape, force_iface,
                                   [prec](int64_t n, int &shape, int &iface) {
                                       ragged_choice(prec, n, shape, iface);
                                   },
                                   host, err);
    if (rc) return fail(rc == 2 ? TRIDIAG_EUNSUPPORTED : TRIDIAG_EINVAL, err);
    hipError_t e = hipMemcpyAsync(work, host.data(), host.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "ragged plan upload");
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "ragged plan upload");
    ragged_plan_record(work, n_sys, prec);
    return TRIDIAG_OK;
}

template <typename T>
static int ragged_solve_impl(int prec, int64_t n_sys, T *a, T *b, T *c, T *d, void *work,
                             hipStream_t st) {
    if (n_sys < 0 || n_sys > 0x7fffffff) return fail(TRIDIAG_EINVAL, "n_sys out of range");
    if (n_sys == 0) return TRIDIAG_OK;
    if (!work || !a || !b || !c || !d) return fail(TRIDIAG_EINVAL, "null pointer");
    if (const int rc = ragged_plan_check(work, n_sys, prec)) return rc;
    char *w = reinterpret_cast<char *>(work);
    RaggedParams<T> p{a, b, c, d, reinterpret_cast<RaggedHeader *>(w),
                      reinterpret_cast<const int32_t *>(w + kRaggedHeaderBytes), n_sys, prec};
    hipError_t e;
    if constexpr (sizeof(T) == 4) e = launch_ragged_f32(p, ragged_grid_f32(), st);
    else e = launch_ragged_f64(p, ragged_grid_f64(), st);
    return e == hipSuccess ? TRIDIAG_OK : hip_fail(e, "ragged launch");
}

// ---------------------------------------------------------------- adjoint
template <typename T>
static const AdjEntry<T> *choose_adj(int layout, int64_t n, int mode, int algo) {
    AdjTable<T> tab = adj_table<T>();
    const bool il = layout == kInterleaved;
    const AdjEntry<T> *best = nullptr;
    for (int i = 0; i < tab.count; ++i) {
        const AdjEntry<T> &e = tab.entries[i];
        if (e.IL != il || e.mode != mode || e.cap < n) continue;
        const AlgoSpec sp = algo_spec(algo);
        const long lanes = 64L * e.NW, w = sizeof(T);
        const long dyn = sp.iface >= kIfacePCR ? 8 * lanes * w : (sp.xchg_lds ? 6 * lanes * w : 0);
        if (dyn + 4 * lanes * w > max_lds_per_block()) continue;
        if (!best || e.cap < best->cap || (e.cap == best->cap && e.M > best->M)) best = &e;
    }
    return best;
}

template <typename T>
static int adjoint_impl(int mode, int algo, int layout, int64_t n, int64_t batch, const T *a,
                        const T *b, const T *c, const T *d, const T *x, const T *g, T *xo, T *ga,
                        T *gb, T *gc, T *gd, hipStream_t st) {
    if (n < 1 || batch < 0) return fail(TRIDIAG_EINVAL, "bad shape");
    if (layout != kContiguous && layout != kInterleaved)
        return fail(TRIDIAG_EINVAL, "adjoint supports layouts 0 and 2");
    if (!a || !b || !c || !g || !ga || !gb || !gc || !gd || (mode == kFused ? (!d || !xo) : !x))
        return fail(TRIDIAG_EINVAL, "null pointer");
    if (batch == 0) return TRIDIAG_OK;
    const int64_t sstride = layout == kInterleaved ? 1 : n;
    const int64_t rstride = layout == kInterleaved ? batch : 1;
    const int32_t bb = buffer_extent<T>(n, batch, sstride, layout);
    const AdjEntry<T> *e = bb ? choose_adj<T>(layout, n, mode, algo) : nullptr;
    if (!e) {
        // Beyond the fused kernels: the shifted coefficients of A^T go into ga / gc and
        // g into gd, A^T lambda = g is solved in place on gd by the plain solver (x
        // first, in place on xo, when fused), then the gradients overwrite ga, gb, gc.
        const bool il = layout == kInterleaved;
        hipError_t err = (sizeof(T) == 4)
            ? launch_adj_shift_f32((const float *)a, (const float *)c, (const float *)g,
                                   mode == kFused ? (const float *)d : nullptr, (float *)ga,
                                   (float *)gc, (float *)gd, (float *)xo, n, batch, il, st)
            : launch_adj_shift_f64((const double *)a, (const double *)c, (const double *)g,
                                   mode == kFused ? (const double *)d : nullptr, (double *)ga,
                                   (double *)gc, (double *)gd, (double *)xo, n, batch, il, st);
        if (err != hipSuccess) return hip_fail(err, "adjoint coefficients");
        const T *xs = x;
        if (mode == kFused) {
            const int rc = solve_impl<T>(algo, layout, n, batch, const_cast<T *>(a),
                                         const_cast<T *>(b), const_cast<T *>(c), xo, st);
            if (rc != TRIDIAG_OK) return rc;
            xs = xo;
        }
        const int rc = solve_impl<T>(algo, layout, n, batch, ga, const_cast<T *>(b), gc, gd, st);
        if (rc != TRIDIAG_OK) return rc;
        err = (sizeof(T) == 4)
            ? launch_adj_grad_f32((const float *)xs, (const float *)gd, (float *)ga, (float *)gb,
                                  (float *)gc, n, batch, il, st)
            : launch_adj_grad_f64((const double *)xs, (const double *)gd, (double *)ga,
                                  (double *)gb, (double *)gc, n, batch, il, st);
        return err == hipSuccess ? TRIDIAG_OK : hip_fail(err, "adjoint gradients");
    }
    const AlgoSpec sp = algo_spec(algo);
    AdjParams<T> p{a, b, c, d, x, g, xo, ga, gb, gc, gd, n, batch, sstride, rstride, bb,
                   sp.iface, sp.xchg_lds};
    hipError_t err = e->launch(p, st);
    return err == hipSuccess ? TRIDIAG_OK : hip_fail(err, "adjoint launch");
}
}  // namespace tri

using namespace tri;

// ======================================================================= Part 2
namespace tri {
// Layout conversion.  Element (coefficient k, system s, row i) sits at
// ptr_k[s*n + i] (layout 0), buf[s*4n + k*n + i] (layout 1) or ptr_k[i*batch + s]
// (layout 2); layouts 0 and 1 are row-pitched views of the same system-major
// matrix, layout 2 its transpose.
template <typename T>
static int convert_impl(int src_layout, int dst_layout, int64_t n, int64_t batch,
                        const T *const src[4], T *const dst[4], hipStream_t st) {
    if (src_layout < 0 || src_layout > 2 || dst_layout < 0 || dst_layout > 2)
        return fail(TRIDIAG_EINVAL, "layouts must be 0, 1 or 2");
    if (n < 1) return fail(TRIDIAG_EINVAL, "n_eqt must be >= 1");
    if (batch < 0) return fail(TRIDIAG_EINVAL, "n_batch must be >= 0");
    if (batch > 0 && n > INT64_MAX / 4 / batch)
        return fail(TRIDIAG_EINVAL, "n_eqt * n_batch is too large");
    if (src_layout == kPacked && !src[0])
        return fail(TRIDIAG_EINVAL, "packed source: the buffer is passed as src_a");
    if (dst_layout == kPacked && !dst[0])
        return fail(TRIDIAG_EINVAL, "packed destination: the buffer is passed as dst_a");
    const int64_t elems = n * batch;
    LayoutMove<T> p{};
    int count = 0;
    for (int k = 0; k < 4; ++k) {
        const bool has_s = src_layout == kPacked || src[k];
        const bool has_d = dst_layout == kPacked || dst[k];
        if (src_layout != kPacked && dst_layout != kPacked && has_s != has_d)
            return fail(TRIDIAG_EINVAL, "coefficient " + std::to_string(k) +
                                            ": source and destination must both be given or "
                                            "both be null");
        if (!has_s || !has_d) continue;
        const T *sp = src_layout == kPacked ? src[0] + k * n : src[k];
        T *dp = dst_layout == kPacked ? dst[0] + k * n : dst[k];
        // whole extent of each side, for the overlap check
        const T *s_lo = src_layout == kPacked ? src[0] : sp;
        const T *d_lo = dst_layout == kPacked ? dst[0] : dp;
        const int64_t s_len = src_layout == kPacked ? 4 * elems : elems;
        const int64_t d_len = dst_layout == kPacked ? 4 * elems : elems;
        const uintptr_t s0 = reinterpret_cast<uintptr_t>(s_lo), d0 = reinterpret_cast<uintptr_t>(d_lo);
        if (elems && s0 < d0 + (uintptr_t)d_len * sizeof(T) && d0 < s0 + (uintptr_t)s_len * sizeof(T))
            return fail(TRIDIAG_EINVAL, "coefficient " + std::to_string(k) +
                                            ": source and destination overlap (in-place "
                                            "conversion is not supported)");
        p.src[count] = sp;
        p.dst[count] = dp;
        ++count;
    }
    if (count == 0) return fail(TRIDIAG_EINVAL, "no coefficient array to convert");
    if (batch == 0) return TRIDIAG_OK;
    const LayoutGeometry geo = layout_geometry(src_layout, dst_layout, n, batch);
    p.rows = geo.rows; p.cols = geo.cols; p.ld_src = geo.ld_src; p.ld_dst = geo.ld_dst;
    const bool transpose = geo.transpose;
    hipError_t e;
    if constexpr (sizeof(T) == 4) e = launch_layout_move_f32(p, count, transpose, st);
    else e = launch_layout_move_f64(p, count, transpose, st);
    return e == hipSuccess ? TRIDIAG_OK : hip_fail(e, "layout conversion launch");
}
}  // namespace tri

extern "C" {

int tridiag_algo_count(int precision) {
    if (precision != TRIDIAG_SINGLE && precision != TRIDIAG_DOUBLE) return 0;
    return algo_count(precision);
}

const char *tridiag_algo_name(int precision, int algo) {
    if (algo < 0 || algo >= tridiag_algo_count(precision)) return nullptr;
    return precision == TRIDIAG_SINGLE ? kNamesSingle[algo] : kNamesDouble[algo];
}

int tridiag_algo_layout(int precision, int algo) {
    if (algo < 0 || algo >= tridiag_algo_count(precision)) return -1;
    return algo >= kBPLGWM ? TRIDIAG_LAYOUT_PACKED : TRIDIAG_LAYOUT_CONTIGUOUS;
}

int tridiag_solve(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch, void *a,
                  void *b, void *c, void *d, void *stream) {
    g_last_error.clear();
    if (precision != TRIDIAG_SINGLE && precision != TRIDIAG_DOUBLE)
        return fail(TRIDIAG_EINVAL, "precision must be 0 (single) or 1 (double)");
    if (algo != TRIDIAG_ALGO_AUTO && (algo < 0 || algo >= algo_count(precision)))
        return fail(TRIDIAG_EINVAL, "unknown algorithm id " + std::to_string(algo));
    if (algo == kVendor) {
        if (layout != kContiguous) return fail(TRIDIAG_EUNSUPPORTED, "gtsv needs layout 0");
        if (n_eqt < 3)
            return fail(TRIDIAG_EUNSUPPORTED, "the vendor gtsv2StridedBatch needs n_eqt >= 3");
        if (n_eqt > 0x7fffffff || n_batch > 0x7fffffff) return fail(TRIDIAG_EINVAL, "size");
        int rc = (precision == TRIDIAG_SINGLE)
                     ? sparse_init<float>((float *)a, (float *)b, (float *)c, (float *)d,
                                          (int)n_eqt, (int)n_batch)
                     : sparse_init<double>((double *)a, (double *)b, (double *)c, (double *)d,
                                           (int)n_eqt, (int)n_batch);
        if (rc) return rc;
        if (hipsparseSetStream(g_sparse, (hipStream_t)stream) != HIPSPARSE_STATUS_SUCCESS)
            return fail(TRIDIAG_EHIP, "hipsparseSetStream");
        return (precision == TRIDIAG_SINGLE)
                   ? sparse_solve<float>((float *)a, (float *)b, (float *)c, (float *)d,
                                         (int)n_eqt, (int)n_batch)
                   : sparse_solve<double>((double *)a, (double *)b, (double *)c, (double *)d,
                                          (int)n_eqt, (int)n_batch);
    }
    hipStream_t st = (hipStream_t)stream;
    if (precision == TRIDIAG_SINGLE)
        return solve_impl<float>(algo, layout, n_eqt, n_batch, (float *)a, (float *)b,
                                 (float *)c, (float *)d, st);
    return solve_impl<double>(algo, layout, n_eqt, n_batch, (double *)a, (double *)b,
                              (double *)c, (double *)d, st);
}

int tridiag_solve_bulk(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch,
                       void *a, void *b, void *c, void *d, void *stream) {
    g_last_error.clear();
    if (precision != TRIDIAG_SINGLE && precision != TRIDIAG_DOUBLE)
        return fail(TRIDIAG_EINVAL, "precision must be 0 (single) or 1 (double)");
    if (algo < 0 || algo >= algo_count(precision))
        return fail(TRIDIAG_EINVAL, "unknown algorithm id " + std::to_string(algo));
    hipStream_t st = (hipStream_t)stream;
    if (precision == TRIDIAG_SINGLE)
        return bulk_impl<float>(algo, layout, n_eqt, n_batch, (float *)a, (float *)b, (float *)c,
                                (float *)d, st);
    return bulk_impl<double>(algo, layout, n_eqt, n_batch, (double *)a, (double *)b, (double *)c,
                             (double *)d, st);
}

int tridiag_solve_noalloc(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch,
                          void *a, void *b, void *c, void *d, void *stream) {
    g_last_error.clear();
    if (precision != TRIDIAG_SINGLE && precision != TRIDIAG_DOUBLE)
        return fail(TRIDIAG_EINVAL, "precision must be 0 (single) or 1 (double)");
    if (algo != TRIDIAG_ALGO_AUTO && (algo < 0 || algo >= algo_count(precision)))
        return fail(TRIDIAG_EINVAL, "unknown algorithm id " + std::to_string(algo));
    if (algo == kVendor)
        return fail(TRIDIAG_EUNSUPPORTED, "the vendor gtsv comparator (algorithm 0) allocates its "
                                          "own buffer; tridiag_solve_noalloc does not run it");
    void *ptr = nullptr;
    if (const int rc = noalloc_scratch(&ptr)) return rc;
    const FixedScratch fx{ptr, kNoAllocScratchBytes};
    hipStream_t st = (hipStream_t)stream;
    if (precision == TRIDIAG_SINGLE)
        return solve_impl<float>(algo, layout, n_eqt, n_batch, (float *)a, (float *)b,
                                 (float *)c, (float *)d, st, &fx);
    return solve_impl<double>(algo, layout, n_eqt, n_batch, (double *)a, (double *)b,
                              (double *)c, (double *)d, st, &fx);
}

}  // extern "C"

namespace tri {
// Plan of tridiag_solve for arrays with 16-byte aligned rows (no launch).
template <typename T>
static int plan_impl(int algo, int layout, int64_t n, int64_t batch, SolvePlan<T> &pl) {
    if (algo != TRIDIAG_ALGO_AUTO && (algo < 0 || algo >= algo_count(sizeof(T) == 4 ? 0 : 1)))
        return fail(TRIDIAG_EINVAL, "unknown algorithm id " + std::to_string(algo));
    if (n < 1 || batch < 1 || layout < 0 || layout > 2) return fail(TRIDIAG_EINVAL, "bad shape");
    if (algo == kVendor) return fail(TRIDIAG_EUNSUPPORTED, "the vendor comparator has no plan");
    const int64_t sstride = layout == kPacked ? 4 * n : (layout == kInterleaved ? 1 : n);
    int64_t b = batch;
    if (layout != kInterleaved && algo != kThomas && buffer_extent<T>(n, b, sstride, layout) == 0)
        b = std::max<int64_t>(1, ((int64_t)0x7fffff00 / (int64_t)sizeof(T)) / sstride);
    pl = make_plan<T>(algo, layout, n, b, sstride, layout != kInterleaved, 0);
    return TRIDIAG_OK;
}
}  // namespace tri

extern "C" {

int tridiag_plan(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch, int *M,
                 int *NW, int *S) {
    g_last_error.clear();
    int m = 0, nw = 0, s = 0, kind = -1;
    if (precision == TRIDIAG_SINGLE) {
        SolvePlan<float> pl;
        if (const int rc = plan_impl<float>(algo, layout, n_eqt, n_batch, pl)) return rc;
        kind = pl.kind;
        if (pl.tile) { m = pl.tile->M; nw = pl.tile->NW; s = pl.tile->S; }
        if (pl.coop) { m = pl.coop->M; nw = pl.coop->NW; s = pl.coop->S; }
    } else if (precision == TRIDIAG_DOUBLE) {
        SolvePlan<double> pl;
        if (const int rc = plan_impl<double>(algo, layout, n_eqt, n_batch, pl)) return rc;
        kind = pl.kind;
        if (pl.tile) { m = pl.tile->M; nw = pl.tile->NW; s = pl.tile->S; }
        if (pl.coop) { m = pl.coop->M; nw = pl.coop->NW; s = pl.coop->S; }
    } else {
        return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
    }
    if (kind != kPlanTile && kind != kPlanCoop)
        return fail(TRIDIAG_EUNSUPPORTED, "this shape is not solved by an on-chip workgroup plan "
                                          "(see tridiag_plan_kind)");
    if (M) *M = m;
    if (NW) *NW = nw;
    if (S) *S = s;
    return TRIDIAG_OK;
}

int tridiag_plan_kind(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch) {
    g_last_error.clear();
    if (precision == TRIDIAG_SINGLE) {
        SolvePlan<float> pl;
        return plan_impl<float>(algo, layout, n_eqt, n_batch, pl) ? -1 : pl.kind;
    }
    if (precision == TRIDIAG_DOUBLE) {
        SolvePlan<double> pl;
        return plan_impl<double>(algo, layout, n_eqt, n_batch, pl) ? -1 : pl.kind;
    }
    fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
    return -1;
}

int tridiag_tile_count(int precision) {
    if (precision == TRIDIAG_SINGLE) return tile_table<float>().count;
    if (precision == TRIDIAG_DOUBLE) return tile_table<double>().count;
    return 0;
}

int tridiag_tile_info(int precision, int index, int *M, int *NW, int *S, int *interleaved,
                      int *vec, int *lds_alpha, int64_t *cap, int *wide_addr, int *padded,
                      int *nontemporal, int *transposed) {
    if (precision == TRIDIAG_SINGLE)
        return tile_info_impl<float>(index, M, NW, S, interleaved, vec, lds_alpha, cap, wide_addr,
                                     padded, nontemporal, transposed);
    if (precision == TRIDIAG_DOUBLE)
        return tile_info_impl<double>(index, M, NW, S, interleaved, vec, lds_alpha, cap, wide_addr,
                                      padded, nontemporal, transposed);
    return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
}

int tridiag_solve_tile_iface(int precision, int index, int iface, int layout, int64_t n_eqt,
                             int64_t n_batch, void *a, void *b, void *c, void *d, void *stream) {
    g_last_error.clear();
    if (precision == TRIDIAG_SINGLE)
        return solve_tile_impl<float>(index, iface, layout, n_eqt, n_batch, (float *)a, (float *)b,
                                      (float *)c, (float *)d, (hipStream_t)stream);
    if (precision == TRIDIAG_DOUBLE)
        return solve_tile_impl<double>(index, iface, layout, n_eqt, n_batch, (double *)a,
                                       (double *)b, (double *)c, (double *)d, (hipStream_t)stream);
    return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
}

int tridiag_solve_tile(int precision, int index, int layout, int64_t n_eqt, int64_t n_batch,
                       void *a, void *b, void *c, void *d, void *stream) {
    return tridiag_solve_tile_iface(precision, index, kIfaceWM, layout, n_eqt, n_batch, a, b, c, d,
                                    stream);
}

int64_t tridiag_ragged_work_size(int64_t n_sys) { return n_sys < 0 ? 0 : ragged_work_bytes(n_sys); }

int tridiag_ragged_shape_count(void) { return kRaggedShapes; }

int tridiag_ragged_shape_info(int index, int *M, int *NWS, int *S, int64_t *cap) {
    g_last_error.clear();
    if (index < 0 || index >= kRaggedShapes) return fail(TRIDIAG_EINVAL, "shape index out of range");
    if (M) *M = kRaggedShapeTab[index].M;
    if (NWS) *NWS = kRaggedShapeTab[index].NWS;
    if (S) *S = kRaggedShapeTab[index].S;
    if (cap) *cap = ragged_cap(index);
    return TRIDIAG_OK;
}

int tridiag_ragged_bucket(int precision, int64_t n, int *shape, int *iface) {
    g_last_error.clear();
    if (precision != TRIDIAG_SINGLE && precision != TRIDIAG_DOUBLE)
        return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
    if (n < 1 || n > kRaggedMaxN) return fail(TRIDIAG_EUNSUPPORTED, "n outside the ragged range");
    int s, f;
    ragged_choice(precision, n, s, f);
    if (shape) *shape = s;
    if (iface) *iface = f;
    return TRIDIAG_OK;
}

int tridiag_ragged_plan(int precision, int64_t n_sys, const int64_t *offsets_host, int shape,
                        int iface, void *work, void *stream) {
    g_last_error.clear();
    return ragged_plan_impl(precision, n_sys, offsets_host, shape, iface, work,
                            (hipStream_t)stream);
}

int tridiag_ragged_solve(int precision, int64_t n_sys, void *a, void *b, void *c, void *d,
                         void *work, void *stream) {
    g_last_error.clear();
    hipStream_t st = (hipStream_t)stream;
    if (precision == TRIDIAG_SINGLE)
        return ragged_solve_impl<float>(precision, n_sys, (float *)a, (float *)b, (float *)c,
                                        (float *)d, work, st);
    if (precision == TRIDIAG_DOUBLE)
        return ragged_solve_impl<double>(precision, n_sys, (double *)a, (double *)b, (double *)c,
                                         (double *)d, work, st);
    return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
}

int tridiag_solve_ragged(int precision, int64_t n_sys, const int64_t *offsets_host, void *a,
                         void *b, void *c, void *d, void *work, void *stream) {
    const int rc = tridiag_ragged_plan(precision, n_sys, offsets_host, -1, -1, work, stream);
    if (rc != TRIDIAG_OK) return rc;
    return tridiag_ragged_solve(precision, n_sys, a, b, c, d, work, stream);
}

int tridiag_adjoint(int precision, int layout, int64_t n_eqt, int64_t n_batch, const void *a,
                    const void *b, const void *c, const void *x, const void *g, void *ga,
                    void *gb, void *gc, void *gd, void *stream) {
    g_last_error.clear();
    hipStream_t st = (hipStream_t)stream;
    if (precision == TRIDIAG_SINGLE)
        return adjoint_impl<float>(kAdjoint, kWM5Reg, layout, n_eqt, n_batch, (const float *)a,
                                   (const float *)b, (const float *)c, nullptr, (const float *)x,
                                   (const float *)g, nullptr, (float *)ga, (float *)gb,
                                   (float *)gc, (float *)gd, st);
    if (precision == TRIDIAG_DOUBLE)
        return adjoint_impl<double>(kAdjoint, kWM5Reg, layout, n_eqt, n_batch, (const double *)a,
                                    (const double *)b, (const double *)c, nullptr,
                                    (const double *)x, (const double *)g, nullptr, (double *)ga,
                                    (double *)gb, (double *)gc, (double *)gd, st);
    return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
}

int tridiag_solve_fwd_bwd(int precision, int layout, int64_t n_eqt, int64_t n_batch,
                          const void *a, const void *b, const void *c, const void *d,
                          const void *g, void *x, void *ga, void *gb, void *gc, void *gd,
                          void *stream) {
    g_last_error.clear();
    hipStream_t st = (hipStream_t)stream;
    if (precision == TRIDIAG_SINGLE)
        return adjoint_impl<float>(kFused, kWM5Reg, layout, n_eqt, n_batch, (const float *)a,
                                   (const float *)b, (const float *)c, (const float *)d, nullptr,
                                   (const float *)g, (float *)x, (float *)ga, (float *)gb,
                                   (float *)gc, (float *)gd, st);
    if (precision == TRIDIAG_DOUBLE)
        return adjoint_impl<double>(kFused, kWM5Reg, layout, n_eqt, n_batch, (const double *)a,
                                    (const double *)b, (const double *)c, (const double *)d,
                                    nullptr, (const double *)g, (double *)x, (double *)ga,
                                    (double *)gb, (double *)gc, (double *)gd, st);
    return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
}

int tridiag_convert_layout(int precision, int src_layout, int dst_layout, int64_t n_eqt,
                           int64_t n_batch, const void *src_a, const void *src_b,
                           const void *src_c, const void *src_d, void *dst_a, void *dst_b,
                           void *dst_c, void *dst_d, void *stream) {
    g_last_error.clear();
    hipStream_t st = (hipStream_t)stream;
    if (precision == TRIDIAG_SINGLE) {
        const float *s[4] = {(const float *)src_a, (const float *)src_b, (const float *)src_c,
                             (const float *)src_d};
        float *d[4] = {(float *)dst_a, (float *)dst_b, (float *)dst_c, (float *)dst_d};
        return convert_impl<float>(src_layout, dst_layout, n_eqt, n_batch, s, d, st);
    }
    if (precision == TRIDIAG_DOUBLE) {
        const double *s[4] = {(const double *)src_a, (const double *)src_b,
                              (const double *)src_c, (const double *)src_d};
        double *d[4] = {(double *)dst_a, (double *)dst_b, (double *)dst_c, (double *)dst_d};
        return convert_impl<double>(src_layout, dst_layout, n_eqt, n_batch, s, d, st);
    }
    return fail(TRIDIAG_EINVAL, "precision must be 0 or 1");
}

const char *tridiag_last_error(void) { return g_last_error.c_str(); }

int tridiag_div_mode(void) { return tri::kDivMode; }

// ======================================================================= Part 1
#define REF_SYMBOL(NAME, T, ALGO, LAYOUT)                                                  \
    void NAME(T *a_gbl, T *b_gbl, T *c_gbl, T *d_gbl, int n_eqt, int n_batch) {            \
        ref_call<T>(ALGO, LAYOUT, a_gbl, b_gbl, c_gbl, d_gbl, n_eqt, n_batch);             \
    }

REF_SYMBOL(Thomas_single, float, kThomas, kContiguous)
REF_SYMBOL(Thomas_double, double, kThomas, kContiguous)
REF_SYMBOL(CR_single, float, kCR, kContiguous)
REF_SYMBOL(CR_double, double, kCR, kContiguous)
REF_SYMBOL(WM5_shmem_single, float, kWM5Shmem, kContiguous)
REF_SYMBOL(WM4_shmem_single, float, kWM4Shmem, kContiguous)
REF_SYMBOL(PCR5_shmem_single, float, kPCR5Shmem, kContiguous)
REF_SYMBOL(PCR4_shmem_single, float, kPCR4Shmem, kContiguous)
REF_SYMBOL(WM5_reg_single, float, kWM5Reg, kContiguous)
REF_SYMBOL(WM4_reg_single, float, kWM4Reg, kContiguous)
REF_SYMBOL(PCR5_reg_single, float, kPCR5Reg, kContiguous)
REF_SYMBOL(PCR4_reg_single, float, kPCR4Reg, kContiguous)
REF_SYMBOL(WM5_shmem_double, double, kWM5Shmem, kContiguous)
REF_SYMBOL(WM4_shmem_double, double, kWM4Shmem, kContiguous)
REF_SYMBOL(PCR5_shmem_double, double, kPCR5Shmem, kContiguous)
REF_SYMBOL(PCR4_shmem_double, double, kPCR4Shmem, kContiguous)
REF_SYMBOL(WM5_reg_double, double, kWM5Reg, kContiguous)
REF_SYMBOL(WM4_reg_double, double, kWM4Reg, kContiguous)
REF_SYMBOL(PCR5_reg_double, double, kPCR5Reg, kContiguous)
REF_SYMBOL(PCR4_reg_double, double, kPCR4Reg, kContiguous)
REF_SYMBOL(CR_reg4_double, double, kCRReg4, kContiguous)
REF_SYMBOL(CR_reg8_double, double, kCRReg8, kContiguous)
REF_SYMBOL(CR_reg16_double, double, kCRReg16, kContiguous)
REF_SYMBOL(CR_reg4_single, float, kCRReg4, kContiguous)
REF_SYMBOL(CR_reg8_single, float, kCRReg8, kContiguous)
REF_SYMBOL(CR_reg16_single, float, kCRReg16, kContiguous)
REF_SYMBOL(BPLG_WM, float, kBPLGWM, kPacked)
REF_SYMBOL(BPLG_LF, float, kBPLGLF, kPacked)
REF_SYMBOL(BPLG_CR, float, kBPLGCR, kPacked)
REF_SYMBOL(BPLG_PCR, float, kBPLGPCR, kPacked)

int CR_init_single(float *, float *, float *, float *, int n_eqt, int n_batch) {
    (void)n_eqt; (void)n_batch;
    return 0;
}
int CR_init_double(double *, double *, double *, double *, int n_eqt, int n_batch) {
    (void)n_eqt; (void)n_batch;
    return 0;
}
// The CR solve keeps no state between calls (its reduction runs on chip), so
// CR_init and CR_final have nothing to allocate or release.
void CR_final(void) {}

int cuSparse_init_single(float *a, float *b, float *c, float *d, int n_eqt, int n_batch) {
    return sparse_init<float>(a, b, c, d, n_eqt, n_batch);
}
int cuSparse_init_double(double *a, double *b, double *c, double *d, int n_eqt, int n_batch) {
    return sparse_init<double>(a, b, c, d, n_eqt, n_batch);
}
void cuSparse_final(void) {
    if (g_sparse_buf) (void)hipFree(g_sparse_buf);
    g_sparse_buf = nullptr;
    g_sparse_bytes = 0;
    if (g_sparse) hipsparseDestroy(g_sparse);
    g_sparse = nullptr;
}
void cuSparse_single(float *a, float *b, float *c, float *d, int n_eqt, int n_batch) {
    int rc = sparse_solve<float>(a, b, c, d, n_eqt, n_batch);
    if (rc == TRIDIAG_OK) (void)hipDeviceSynchronize();
    else fprintf(stderr, "tridiag_amd: cuSparse_single failed: %s\n", g_last_error.c_str());
}
void cuSparse_double(double *a, double *b, double *c, double *d, int n_eqt, int n_batch) {
    int rc = sparse_solve<double>(a, b, c, d, n_eqt, n_batch);
    if (rc == TRIDIAG_OK) (void)hipDeviceSynchronize();
    else fprintf(stderr, "tridiag_amd: cuSparse_double failed: %s\n", g_last_error.c_str());
}

}  // extern "C"
