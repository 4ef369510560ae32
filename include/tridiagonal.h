/*
 * tridiagonal.h -- C-ABI of the MI355X batched tridiagonal solver library
 * (libtridiag_amd.so).  Plain pointers and sizes only.
 *
 * Part 1 keeps the entry-point surface of the reference's
 * Tridiagonal_CUDA/tridiagonal.h:8-57 so that its test driver (test.cpp,
 * helpers.h solver table) and run_all.sh sweep link against this library
 * unchanged.  Semantics of every Part-1 symbol, as in the reference:
 *   - a_gbl, b_gbl, c_gbl, d_gbl are device pointers (hipMalloc), n_eqt rows
 *     per system, n_batch systems, layout 0: row i of system s at [s*n_eqt + i];
 *   - the solve is in place: x is returned in d_gbl; c_gbl may be clobbered;
 *   - a_gbl[s*n_eqt] (a_0) and c_gbl[s*n_eqt + n_eqt-1] (c_{n-1}) are ignored;
 *   - the call is synchronous (ends with a device synchronisation), runs on
 *     the null stream of the current device and returns nothing.  Unlike the
 *     reference, bad arguments or launch failures do not exit the process:
 *     they are printed to stderr and kept in tridiag_last_error().
 *   - the reference's size preconditions (n%64 for WM, n%32 for PCR, n%R for
 *     CR_reg R, batch >= 2) do not apply: every n >= 1 and batch >= 0 work.
 *
 * Part 2 adds status-returning, stream-ordered entry points with an explicit
 * layout (0 contiguous, 1 packed a|b|c|d per system, 2 interleaved
 * [i*batch + s]), ragged batches and the adjoint (backward) solve.
 */
#ifndef TRIDIAGONAL_AMD_H_INCLUDED
#define TRIDIAGONAL_AMD_H_INCLUDED 1

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================== Part 1
 * Reference-compatible symbols.  Each line cites the declaration it replaces
 * in Tridiagonal_CUDA/tridiagonal.h.
 */

/* tridiagonal.h:8-12 -- vendor library comparator (cuSPARSE gtsv2StridedBatch
 * in the reference; hipSPARSE gtsv2StridedBatch here).  init allocates the
 * library workspace and returns 0 on success. */
int cuSparse_init_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
int cuSparse_init_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void cuSparse_final(void);
void cuSparse_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void cuSparse_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:15-18 -- BPLG entry points (FP32, packed layout 1: one buffer,
 * system s at [4*n*s, 4*n*(s+1)) holding a|b|c|d; only a_gbl is read, x is
 * returned in the d slot).  Solved by this library's own kernels; the BPLG
 * restriction to power-of-two n <= 2048 does not apply. */
void BPLG_WM(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void BPLG_LF(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void BPLG_CR(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void BPLG_PCR(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:21-22 -- naive Thomas, one lane per system. */
void Thomas_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void Thomas_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:25-29 -- cyclic reduction.  init/final manage the scratch
 * buffer (4*n*batch words) used by the global-memory levels; init returns 0 on
 * success. */
int CR_init_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
int CR_init_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void CR_final(void);
void CR_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void CR_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:31-34 -- LDS-staged ("shmem") PCR-pThomas / WM-pGE, FP32. */
void WM5_shmem_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void WM4_shmem_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void PCR5_shmem_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void PCR4_shmem_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:36-39 -- register-resident ("reg") variants, FP32. */
void WM5_reg_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void WM4_reg_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void PCR5_reg_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void PCR4_reg_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:41-44 -- LDS-staged variants, FP64. */
void WM5_shmem_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void WM4_shmem_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void PCR5_shmem_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void PCR4_shmem_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:46-49 -- register-resident variants, FP64. */
void WM5_reg_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void WM4_reg_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void PCR5_reg_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void PCR4_reg_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);

/* tridiagonal.h:52-57 -- register-packed CR, R = 4, 8, 16 rows per lane. */
void CR_reg4_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void CR_reg8_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void CR_reg16_double(double *a_gbl, double *b_gbl, double *c_gbl, double *d_gbl, int n_eqt, int n_batch);
void CR_reg4_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void CR_reg8_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);
void CR_reg16_single(float *a_gbl, float *b_gbl, float *c_gbl, float *d_gbl, int n_eqt, int n_batch);

/* ===================================================================== Part 2
 * Status-returning API (no counterpart in the reference; its callers are this
 * library's test driver, the Python binding and the benchmark).
 * Return codes: */
#define TRIDIAG_OK 0
#define TRIDIAG_EINVAL 1      /* bad argument (null pointer, n < 1, bad id)   */
#define TRIDIAG_EUNSUPPORTED 2/* combination not provided by this build       */
#define TRIDIAG_EHIP 3        /* HIP runtime error (launch, memory)           */

#define TRIDIAG_SINGLE 0      /* precision ids, as the reference CLI (test.cpp:141-149) */
#define TRIDIAG_DOUBLE 1

#define TRIDIAG_LAYOUT_CONTIGUOUS 0
#define TRIDIAG_LAYOUT_PACKED 1
#define TRIDIAG_LAYOUT_INTERLEAVED 2

/* Algorithm ids = indices of the reference solver table
 * (Tridiagonal_CUDA/helpers.h:23-66): 0 vendor gtsv, 1 Thomas, 2 CR,
 * 3 shmem PCR-pThomas, 4 shmem PCR-half-pThomas, 5 reg PCR-pThomas,
 * 6 reg PCR-half-pThomas, 7 shmem WM-pGE, 8 shmem WM-half-pGE, 9 reg WM-pGE,
 * 10 reg WM-half-pGE, 11 reg-4 CR, 12 reg-8 CR, 13 reg-16 CR; FP32 only:
 * 14 BPLG WM, 15 BPLG LF, 16 BPLG CR, 17 BPLG PCR (packed layout).
 * TRIDIAG_ALGO_AUTO picks the fastest kernel for (precision, layout, n). */
#define TRIDIAG_ALGO_AUTO (-1)

int tridiag_algo_count(int precision);
const char *tridiag_algo_name(int precision, int algo);
/* Native layout of an algorithm in the reference table (helpers.h:17). */
int tridiag_algo_layout(int precision, int algo);

/* Solve n_batch systems of n_eqt rows in place (x -> d), enqueued on
 * `stream` (a hipStream_t; NULL = null stream).  Asynchronous. */
int tridiag_solve(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch,
                  void *a, void *b, void *c, void *d, void *stream);

/* The reference's own form of algorithm ids 2-13 (ids 2-10 one wave per system),
 * with the reference's memory semantics (ids 3-10: c' into c and d' into d on the
 * way down, x into d on the way up; ids 2, 11-13: a, b, c kept, x into d):
 *   2      CR (tridiagonal_basic.cu:32-164): cyclic reduction with its levels in
 *          a stream-ordered scratch buffer of 4 * n_eqt * n_batch words (the odd
 *          rows of each level kept) down to at most 64 rows, PCR across the
 *          wave's lanes for those, then the even rows of every level on the way
 *          back;
 *   3 / 5  PCR-pThomas (tridiagonal_wm_pcr.cu:886-1018, 1160-1282): 6 PCR steps
 *          over a sliding window of 64-row blocks, Thomas along the 64 stride-64
 *          subsystems;
 *   4 / 6  PCR-half-pThomas (:1020-1158, 1284-1431): 5 steps, 32 stride-32
 *          subsystems over two chained half-waves;
 *   7 / 9  WM-pGE (:17-205, 424-646): 128-row chunks, 6 Wang-Mou merge rounds to
 *          64-row blocks, partial Gaussian elimination block after block;
 *   8 / 10 WM-half-pGE (:207-422, 648-884): 5 rounds, 32-row blocks;
 *   11-13  CR_reg4/8/16 (tridiagonal_onchip.cu:170-369): one workgroup per
 *          system, R = 4, 8, 16 rows per lane reduced in registers, PCR in LDS
 *          over the lanes' last rows, then back; n_eqt <= 1024 R (longer systems
 *          return TRIDIAG_EUNSUPPORTED here and keep the default path under
 *          $TRIDIAG_FORM=bulk);
 * the shared-memory ids (3, 4, 7, 8) exchange rows between lanes through LDS,
 * the register ids (5, 6, 9, 10) through ds_bpermute.  tridiag_solve runs the
 * on-chip partition kernel for these ids unless $TRIDIAG_FORM=bulk; this entry
 * always runs the reference's form.  Layouts 0 and 1; any n_eqt >= 1 up to the
 * bound of ids 11-13 (the reference needs n % 64 for WM and n % 32 for PCR).
 * Other ids and the
 * interleaved layout return TRIDIAG_EUNSUPPORTED.  Asynchronous on `stream`.
 * tridiag_solve_noalloc refuses id 2 under $TRIDIAG_FORM=bulk (its scratch). */
int tridiag_solve_bulk(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch,
                       void *a, void *b, void *c, void *d, void *stream);

/* tridiag_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and error behaviour as tridiag_solve, except:
 *   - the paths of tridiag_solve whose kernels use scratch memory (the
 *     cooperative row-block solve and the chunked solve of long systems) keep it
 *     in a fixed device array of TRIDIAG_NOALLOC_SCRATCH_BYTES per GPU, owned by
 *     the library and held until the process ends, instead of stream-ordered
 *     allocations; the chunked solve then runs in as many launch groups as the
 *     array needs, and a cooperative solve whose scratch exceeds it takes the
 *     path tridiag_solve would take without it (its results then agree with
 *     tridiag_solve's to the solver's tolerance rather than bit for bit);
 *   - algorithm 0 (the vendor comparator, which allocates its own buffer) and
 *     interleaved systems that tridiag_solve solves through a contiguous copy of
 *     all four arrays (above 4096 rows when the cooperative solve does not take
 *     them: above 16384 rows, n not a multiple of its block rows, or its scratch
 *     beyond the array) return TRIDIAG_EUNSUPPORTED.
 * Calls on one GPU share its array: calls that overlap on one GPU (another
 * stream, another thread) race.  The first call on a GPU looks up the array's
 * device address (hipGetSymbolAddress). */
#define TRIDIAG_NOALLOC_SCRATCH_BYTES ((size_t)4 << 20)
int tridiag_solve_noalloc(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch,
                          void *a, void *b, void *c, void *d, void *stream);

/* Ragged batch: systems of uneven size, contiguous per system and concatenated.
 * System s holds rows [offsets[s], offsets[s+1]) of a, b, c, d (a_0 and c_{n-1}
 * of each system are ignored); x replaces d.  Sizes 0..8192 rows.  The systems
 * are solved by one persistent launch whose work list (size buckets, a tile
 * shape and an interface solver per bucket, largest bucket first; each item
 * records the first row and size of its systems) is planned on the host from
 * the host copy of the offsets and kept in `work`, a device buffer of
 * tridiag_ragged_work_size(n_sys) bytes.
 *
 *   tridiag_ragged_plan   builds the work list from the host copy of the
 *                         offsets and uploads it (synchronises `stream`; do not
 *                         call it during graph capture).  shape = -1 and
 *                         iface = -1 take each bucket's choice from the
 *                         autotuned table (tuning_gfx950.txt, lines
 *                         "R precision n_max shape iface"); a shape or iface
 *                         >= 0 forces it for every system (iface: an interface
 *                         solver, or TRIDIAG_IFACE_PCR_DEPTH(d) for PCR +
 *                         Thomas of depth d <= log2(lanes per system)).
 *   tridiag_ragged_solve  launches the solve for a planned work buffer (one
 *                         kernel, no host synchronisation, no allocation:
 *                         capturable into a graph and replayable).  Concurrent
 *                         solves on one work buffer race.  Returns
 *                         TRIDIAG_EINVAL when `work` was not planned by
 *                         tridiag_ragged_plan in this process, or was planned
 *                         for another precision or system count.
 *   tridiag_solve_ragged  plan + solve.
 *
 * No counterpart in the reference, whose entry points take one n_eqt for the
 * whole batch (Tridiagonal_CUDA/tridiagonal.h:8-57). */
/* Shape `index` of the persistent kernel: M rows per lane, NWS waves per system
 * (0 for the sub-wave shapes, which hold 512/S lanes per system and several
 * systems per wave), S systems per 512-lane workgroup, capacity lanes * M rows. */
int64_t tridiag_ragged_work_size(int64_t n_sys);
int tridiag_ragged_shape_count(void);
int tridiag_ragged_shape_info(int index, int *M, int *NWS, int *S, int64_t *cap);
int tridiag_ragged_bucket(int precision, int64_t n, int *shape, int *iface);
int tridiag_ragged_plan(int precision, int64_t n_sys, const int64_t *offsets_host, int shape,
                        int iface, void *work, void *stream);
int tridiag_ragged_solve(int precision, int64_t n_sys, void *a, void *b, void *c, void *d,
                         void *work, void *stream);
int tridiag_solve_ragged(int precision, int64_t n_sys, const int64_t *offsets_host, void *a,
                         void *b, void *c, void *d, void *work, void *stream);

/* Adjoint of the solve.  Given a, b, c, the forward solution x and the upstream
 * gradient g = dL/dx (all layout `layout`, layouts 0 and 2), writes
 *   gd = lambda = A^{-T} g,  ga_i = -lambda_i x_{i-1},  gb_i = -lambda_i x_i,
 *   gc_i = -lambda_i x_{i+1}  (ga_0 = gc_{n-1} = 0).
 * Outputs may not alias inputs.  Asynchronous on `stream`. */
int tridiag_adjoint(int precision, int layout, int64_t n_eqt, int64_t n_batch, const void *a,
                    const void *b, const void *c, const void *x, const void *g, void *ga,
                    void *gb, void *gc, void *gd, void *stream);

/* Fused forward + adjoint for a known upstream gradient g (e.g. a linear loss):
 * x is written to `x`, the four gradients as in tridiag_adjoint.  Reads
 * a, b, c, d, g once.  Asynchronous on `stream`. */
int tridiag_solve_fwd_bwd(int precision, int layout, int64_t n_eqt, int64_t n_batch,
                          const void *a, const void *b, const void *c, const void *d,
                          const void *g, void *x, void *ga, void *gb, void *gc, void *gd,
                          void *stream);

/* Layout conversion of the coefficient arrays.  No counterpart symbol in the
 * reference: its driver builds each solver's layout while it replicates the
 * batch (init_batch, Tridiagonal_CUDA/utils.cu:263-297); this lets callers that
 * hold one layout use a solver whose fast path is another.  Element (coefficient k, system s, row i) sits at
 *   layout 0: ptr_k[s*n_eqt + i]       (four arrays a, b, c, d)
 *   layout 1: buf[s*4*n_eqt + k*n_eqt + i]  (one buffer, passed as *_a; the
 *             *_b, *_c, *_d pointers on that side are ignored)
 *   layout 2: ptr_k[i*n_batch + s]      (four arrays)
 * Between layouts 0 and 2 a null pointer skips that coefficient (source and
 * destination must then both be null), e.g. to convert the solution d alone.
 * With a packed side, the coefficients converted are those given on the other
 * side.  Source and destination may not overlap.  Asynchronous on `stream`;
 * bit-exact (a permutation of elements). */
int tridiag_convert_layout(int precision, int src_layout, int dst_layout, int64_t n_eqt,
                           int64_t n_batch, const void *src_a, const void *src_b,
                           const void *src_c, const void *src_d, void *dst_a, void *dst_b,
                           void *dst_c, void *dst_d, void *stream);

/* Message of the last failed call on this thread ("" if none). */
const char *tridiag_last_error(void);

/* Division policy this library was built with; replaces the reference's
 * compile-time division macros (Tridiagonal_CUDA/intrin.h:23-87, selected by
 * -D flags in its Makefile).  Every variant exports the same symbols, so a caller
 * picks a policy by linking (or loading) the matching library:
 *   TRIDIAG_DIV_REFINED  libtridiag_amd.so          hardware reciprocal + Newton
 *                                                    steps (FP32 one, FP64 two)
 *   TRIDIAG_DIV_IEEE     libtridiag_amd_div_ieee.so correctly rounded reciprocal
 *   TRIDIAG_DIV_FAST     libtridiag_amd_div_fast.so hardware reciprocal alone
 *                                                    (FP32), one Newton step (FP64) */
#define TRIDIAG_DIV_REFINED 0
#define TRIDIAG_DIV_IEEE 1
#define TRIDIAG_DIV_FAST 2
int tridiag_div_mode(void);

/* Kernel configuration the dispatcher would use for arrays with 16-byte
 * aligned rows (for reports and tuning).
 *   tridiag_plan_kind  how the batch is solved: 0 one system per lane (short
 *                      systems), 1 the Thomas recurrence (algorithm 1), 2 one
 *                      workgroup per system or per S systems, on chip, 3 K
 *                      workgroups per group of S systems, each holding a block of
 *                      64/S*NW*M rows (interleaved) or 64*NW*M rows (contiguous),
 *                      on chip, 4 three passes over chunks through global
 *                      memory, 5 interleaved systems solved through a contiguous
 *                      copy; -1 on a bad argument.
 *   tridiag_plan       for kinds 2 and 3, rows-per-lane M, waves per workgroup
 *                      NW, systems per workgroup S; TRIDIAG_EUNSUPPORTED for the
 *                      other kinds. */
int tridiag_plan_kind(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch);
int tridiag_plan(int precision, int algo, int layout, int64_t n_eqt, int64_t n_batch, int *M,
                 int *NW, int *S);

/* Compiled on-chip kernel configurations, for the autotuner: count, shape of
 * entry `index` (rows per lane M, waves per workgroup NW, systems per workgroup
 * S, interleaved layout, 16-byte vector loads, one coefficient kept in LDS,
 * largest n_eqt, 64-bit addressing for arrays of 2 GiB or more, padded: accepts
 * n_eqt below the capacity; unpadded entries need n_eqt == capacity,
 * nontemporal loads/stores, rows handed to lanes through LDS after whole-line
 * loads), and a solve forced onto that entry. */
int tridiag_tile_count(int precision);
int tridiag_tile_info(int precision, int index, int *M, int *NW, int *S, int *interleaved,
                      int *vec, int *lds_alpha, int64_t *cap, int *wide_addr, int *padded,
                      int *nontemporal, int *transposed);
int tridiag_solve_tile(int precision, int index, int layout, int64_t n_eqt, int64_t n_batch,
                       void *a, void *b, void *c, void *d, void *stream);
/* The same with the interface solver chosen: 0 Wang-Mou merges (what
 * tridiag_solve_tile uses), 1 Wang-Mou half, 2 PCR, 3 PCR + Thomas, 4 CR.  The
 * autotuner times each algorithm family on its own interface solver with it.
 * For PCR + Thomas, TRIDIAG_IFACE_PCR_DEPTH(d) sets the PCR depth: d PCR steps on
 * the interface rows, then Thomas along the 2^d subsystems left (1 <= d <=
 * log2(P) for a tile of P = capacity / M lane segments per system; 0 or no depth
 * keeps the default, PCR up to the wave's segment stride).  Other codes:
 * TRIDIAG_EINVAL.  The same codes serve tridiag_ragged_plan's iface. */
#define TRIDIAG_IFACE_PCR_DEPTH(d) (3 | ((d) << 4))
int tridiag_solve_tile_iface(int precision, int index, int iface, int layout, int64_t n_eqt,
                             int64_t n_batch, void *a, void *b, void *c, void *d, void *stream);

#ifdef __cplusplus
}
#endif

#endif
