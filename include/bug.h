/*
 * bug.h -- C ABI of the MI355X (gfx950) backend for the BUG step hot path.
 *
 * Every entry point takes plain pointers and sizes; device buffers are raw HIP device pointers
 * (complex FP64 interleaved as (re, im) pairs, the memory layout of Julia's ComplexF64 and
 * numpy's complex128) and streams are hipStream_t passed as void*.  Nothing here allocates
 * caller-visible memory: outputs are caller-allocated, their block layout is described by the
 * caller, and the library checks every block against the buffer size before it launches.
 *
 * Block-sparse U(1) tensors are held as a family of "slices", each a block-diagonal matrix with
 * a charge shift (struct bug_bsm).  An MPS site tensor A(l, s, r) is the family A_s (s = local
 * state, shift = charge of s); an MPO environment E(bra, w, ket) is the family E_w (shift =
 * minus the charge of MPO channel w); a zero-site core is one slice with shift 0; the dense
 * (no-symmetry) case is the same format with one charge-0 sector per leg.
 *
 * Status codes: 0 ok, otherwise one of BUG_E*; bug_last_error() returns the message.
 * Numerical shortfall of the adaptive step is never an error code: it is reported as a verdict.
 */
#ifndef BUG_H
#define BUG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BUG_OK 0
#define BUG_EINVAL 1        /* invalid argument; no output was touched */
#define BUG_ENONFINITE 2    /* non-finite input to a small dense solver */
#define BUG_EINVARIANT 3    /* block structure inconsistent (charges, dims, offsets) */
#define BUG_EDEVICE 4       /* a HIP runtime call failed */
#define BUG_ENOMEM 5        /* device workspace could not be allocated */

/* Arithmetic of the dense block contractions. */
#define BUG_PREC_FP64 0     /* complex FP64 on FP64 MFMA, FP64 accumulate */
#define BUG_PREC_FP32 1     /* complex split into real FP32 MFMA products, FP32 accumulate */

/* Operand transforms of a block product term. */
#define BUG_OP_N 0          /* as stored */
#define BUG_OP_T 1          /* transpose */
#define BUG_OP_C 2          /* conjugate transpose */
#define BUG_OP_R 3          /* complex conjugate, no transpose */

typedef struct bug_ctx bug_ctx;

/* One leg of a block-sparse tensor: sectors sorted by charge, each with its dimension. */
typedef struct bug_leg {
    int32_t nsec;
    const int32_t* charge;   /* [nsec] U(1) charge (2 S^z for spin-1/2); 0 for dense legs */
    const int32_t* dim;      /* [nsec] sector dimension */
} bug_leg;

/*
 * A family of block-diagonal matrices with charge shifts ("block-sparse matrix", bsm).
 * Slice k has a block at row sector i iff the column leg holds charge row.charge[i] + shift[k];
 * the block is row.dim[i] x col.dim[j], column-major with leading dimension row.dim[i], at
 * element offset offset[k * row.nsec + i] of data (-1 = block absent).
 */
typedef struct bug_bsm {
    bug_leg row;
    bug_leg col;
    int32_t nslice;
    const int32_t* shift;    /* [nslice] */
    const int64_t* offset;   /* [nslice * row.nsec] element offsets, -1 = absent */
    void* data;              /* device pointer, complex FP64 */
    int64_t nelem;           /* complex elements addressable from data */
} bug_bsm;

/*
 * One site of an MPO as a sparse operator-valued matrix: W[a, b] = sum of coeff |sout><sin|
 * over the entries with that (a, b).  cl/cr are the channel charges (cumulative charge change),
 * so every entry satisfies cr[b] = cl[a] + q(sout) - q(sin).
 */
typedef struct bug_mpo_site {
    int32_t wl, wr;
    const int32_t* cl;       /* [wl] */
    const int32_t* cr;       /* [wr] */
    int32_t nent;
    const int32_t* a;        /* [nent] left channel */
    const int32_t* b;        /* [nent] right channel */
    const int32_t* sout;     /* [nent] local output state */
    const int32_t* sin;      /* [nent] local input state */
    const double* coeff;     /* [2 * nent] complex coefficients */
} bug_mpo_site;

/* One term of a generic block product: out[o] += alpha * op_a(P[p]) * op_b(Q[q]). */
typedef struct bug_term {
    int32_t o, p, q;
    int32_t op_a, op_b;
    double alpha_re, alpha_im;
} bug_term;

/*
 * One block copy dst(r, c) = op(src)(r, c) for r < rows, c < cols (dst is rows x cols).
 * mode: 0 copy, 1 transpose, 2 conjugate transpose, 3 conjugate.  scale_mode 1 multiplies row r
 * by scale[scale_off + r], 2 multiplies column c by scale[scale_off + c] (real scales).
 */
typedef struct bug_copy {
    int64_t src_off, dst_off;
    int32_t src_ld, dst_ld, rows, cols, mode, scale_mode;
    int64_t scale_off;
} bug_copy;

/* One thin QR job: A (m x n, column-major, lda) -> Q (m x k) and R (k x n), k = min(m, n). */
typedef struct bug_qr_job {
    int64_t a_off, q_off, r_off;
    int32_t m, n, lda, ldq, ldr, pad;
} bug_qr_job;

/* One SVD job, m >= n: G (m x n) -> U (m x n), S (n), Vh (n x n), descending. */
typedef struct bug_svd_job {
    int64_t g_off, u_off, s_off, vh_off;
    int32_t m, n, ldg, ldu, ldvh, phase_on_v;
} bug_svd_job;

/* ---- context ---------------------------------------------------------------------------- */
int bug_ctx_create(int device, bug_ctx** out);
int bug_ctx_destroy(bug_ctx* ctx);
const char* bug_last_error(const bug_ctx* ctx);
/* Force a GEMM tile configuration (autotuner override); -1 restores automatic choice. */
int bug_set_tile_config(bug_ctx* ctx, int config);
int bug_tile_config_count(void);
/* Grouped-GEMM schedules (tile configuration + split-k factor).  The reference has no
 * counterpart: MKL's zgemm picks its own blocking (deps/Telum/src/contract.jl:1022-1036).
 * bug_set_gemm_schedule forces both on one context (config -1 / splits 0: automatic; splits in
 * 1..8).  bug_set_autotune sets the process-wide mode: 0 heuristic only, 1 use the tuning table
 * where it holds the plan's class, 2 as 1 and time every candidate schedule on the first launch
 * of a new class (the launch then synchronizes its stream).  The table (plan class -> schedule)
 * is shared by all contexts; bug_autotune_count gives [entries, classes tuned by this process];
 * save/load write and merge a text file; clear empties it. */
int bug_set_gemm_schedule(bug_ctx* ctx, int32_t config, int32_t splits);
int bug_set_autotune(int32_t mode);
int bug_autotune_count(int64_t* out2);
int bug_autotune_clear(void);
int bug_autotune_save(const char* path);
int bug_autotune_load(const char* path);
/* Graph capture of entry-point calls (no counterpart in the reference, whose Julia calls run
 * eagerly).  Between begin and end the caller captures `stream` (hipStreamBeginCapture or a
 * framework's graph object); calls on that stream then refuse (BUG_EINVAL) anything that cannot
 * be captured: a plan upload, scratch growth, autotuner timing -- run the same call once on the
 * stream before capturing.  From begin until the matching release the context keeps every plan
 * image and scratch buffer that a replay may touch alive (buffers that must grow or reset are
 * retired, not freed).  Release (after the graph is destroyed or will not be replayed again)
 * synchronizes the device and drops the capture stream's scratch. */
int bug_graph_capture_begin(bug_ctx* ctx, void* stream);
int bug_graph_capture_end(bug_ctx* ctx);
int bug_graph_release(bug_ctx* ctx, void* stream);
/* The (tile configuration, split factor) pairs of the grouped-GEMM launches of the last
 * entry-point call, in launch order: min(cap, count) pairs into out[2 cap], the total in *count. */
int bug_last_schedules(const bug_ctx* ctx, int32_t cap, int32_t* out, int32_t* count);
/* Device workspace and plan-cache statistics: [bytes held, plans cached, plan hits, misses]. */
int bug_ctx_stats(const bug_ctx* ctx, int64_t* out4);
/* Algorithmic flops (sum of 8 m n k over the block products actually multiplied) of the last
 * contraction entry-point call on this context. */
int bug_last_op_flops(const bug_ctx* ctx, double* flops);

/* ---- block contractions ------------------------------------------------------------------- */

/*
 * Generic grouped block product: for every block of every slice of `out`,
 *   out = (accumulate ? out : 0) + sum over matching terms of alpha op_a(P) op_b(Q).
 * One launch covers every block and every term (replaces the sector-pair loop of Telum's
 * abelian contract, deps/Telum/src/contract.jl:1022-1036).
 */
int bug_bsm_product(bug_ctx* ctx, const bug_bsm* out, const bug_bsm* P, const bug_bsm* Q,
                    int32_t nterm, const bug_term* terms, int32_t accumulate, int32_t precision,
                    void* stream);

/*
 * Left environment push  Lout_b = sum_{a,sout,sin} W[a,b](sout,sin) A_sout^H L_a A_sin
 * (replaces _mpo_left_step / push_left_channels, src/BUG/Core/Environments/mpo_environments.jl:65-90,113).
 * `bra` may be NULL (bra = A); otherwise it gives the bra tensor (matrix elements between two states).
 */
int bug_env_push_left(bug_ctx* ctx, const bug_bsm* L, const bug_bsm* A, const bug_bsm* bra,
                      const bug_mpo_site* W, const bug_bsm* Lout, int32_t precision, void* stream);

/*
 * Right environment push  Rout_a = sum_{b,sout,sin} W[a,b](sout,sin) conj(B_sout) R_b B_sin^T
 * (replaces _mpo_right_step / push_right_channels, mpo_environments.jl:92-108,117).
 */
int bug_env_push_right(bug_ctx* ctx, const bug_bsm* R, const bug_bsm* B, const bug_bsm* bra,
                       const bug_mpo_site* W, const bug_bsm* Rout, int32_t precision, void* stream);

/*
 * One-site effective Hamiltonian  Y_sout = sum W[a,b](sout,sin) L_a X_sin R_b^T  - shift * X
 * (replaces apply_one_site + shift_h, src/BUG/Core/Environments/effective_operators.jl:39-64,118-129).
 */
int bug_apply_one_site(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W, const bug_bsm* R,
                       const bug_bsm* X, const bug_bsm* Y, double shift, int32_t precision,
                       void* stream);
/* The same apply with no device allocation: no counterpart in the reference, it is for callers
 * that cannot allocate (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The stage-1 intermediate, the plan images and the split-k partials go into one
 * static device workspace of 64 MiB per GPU (applies that need more are refused with BUG_EINVAL
 * before anything runs); schedules come from the tuning table or the heuristic, never from timing.
 * Calls that overlap on one GPU (different streams or threads) race on the workspace.  Results
 * equal bug_apply_one_site's. */
int bug_apply_one_site_noalloc(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W,
                               const bug_bsm* R, const bug_bsm* X, const bug_bsm* Y, double shift,
                               int32_t precision, void* stream);
/* k independent one-site applies Y_i = (H_eff,i - shifts[i]) X_i (shifts may be NULL) with one
 * grouped-GEMM launch per product stage for all of them: the instances' block products are
 * packed into one task list over the lowest of their base addresses (operands 16-byte aligned).
 * The reference applies one H_eff at a time (effective_operators.jl:39-64); this entry point is for
 * many small independent evolutions (parameter sets, serving requests) advanced together. */
int bug_apply_one_site_batch(bug_ctx* ctx, int32_t k, const bug_bsm* L, const bug_mpo_site* W,
                             const bug_bsm* R, const bug_bsm* X, const bug_bsm* Y,
                             const double* shifts, int32_t precision, void* stream);

/*
 * Zero-site (root) generator  Y = sum_w L_w C R_w^T - shift * C, with L, R the environments
 * already pushed through the root frames
 * (replaces apply_zero_site, effective_operators.jl:103-107, with zero_site_h :91-94 done by the
 * two env pushes above).
 */
int bug_apply_zero_site(bug_ctx* ctx, const bug_bsm* L, const bug_bsm* R, const bug_bsm* C,
                        const bug_bsm* Y, double shift, int32_t precision, void* stream);
/* k independent zero-site (root) applies Y_i = (H0_i - shifts[i]) C_i (shifts may be NULL) with one
 * grouped-GEMM launch per product stage for all of them, packed as in bug_apply_one_site_batch.
 * The reference solves one root at a time (effective_operators.jl:91-107, root_methods.jl:
 * 359-402); this entry point carries the lockstep root solve of bug_step_batch. */
int bug_apply_zero_site_batch(bug_ctx* ctx, int32_t k, const bug_bsm* L, const bug_bsm* R,
                              const bug_bsm* C, const bug_bsm* Y, const double* shifts,
                              int32_t precision, void* stream);

/* ---- vectors (Krylov recurrences: tensor_inner, _cgs2!, _combine; krylov.jl:296-332) ------- */

/* y += (ar + i ai) x over n complex elements. */
int bug_axpy(bug_ctx* ctx, int64_t n, double ar, double ai, const void* x, void* y, void* stream);
/* y += sum_j (alphas[2j] + i alphas[2j+1]) xs[j] for j < k <= 64, added in j order (one launch
 * for the Gram-Schmidt subtraction against a whole basis and the basis recombination,
 * _cgs2!/_combine in krylov.jl:296-332). */
int bug_axpy_many(bug_ctx* ctx, int64_t n, int32_t k, const double* alphas, const void* const* xs,
                  void* y, void* stream);
/* Segmented forms for K independent vectors packed in one buffer of n elements, segment i at
 * [segs[2i], segs[2i] + segs[2i+1]) (1 <= nseg <= 4096, no counterpart in the reference, which
 * runs one Krylov recurrence at a time): out[2(j nseg + i)], out[2(j nseg + i) + 1] = re, im of
 * <xs[j]_i | y_i> for j < k <= 64 (host output, synchronizes the stream); and
 * y_i += sum_j (alphas[2(j nseg + i)] + i alphas[2(j nseg + i) + 1]) xs[j]_i, added in j order. */
int bug_dotc_segments(bug_ctx* ctx, int64_t n, int32_t nseg, const int64_t* segs, int32_t k,
                      const void* const* xs, const void* y, double* out, void* stream);
int bug_axpy_segments(bug_ctx* ctx, int64_t n, int32_t nseg, const int64_t* segs, int32_t k,
                      const double* alphas, const void* const* xs, void* y, void* stream);
/* x *= (ar + i ai). */
int bug_scal(bug_ctx* ctx, int64_t n, double ar, double ai, void* x, void* stream);
/* out[2j], out[2j+1] = re, im of <xs[j] | y> for j < k <= 64 (host output; synchronizes the
 * stream).  Replaces tensor_inner (src/BUG/Core/TensorNetworks/symmetric_mps.jl:83-93). */
int bug_dotc(bug_ctx* ctx, int64_t n, int32_t k, const void* const* xs, const void* y,
             double* out, void* stream);

/* ---- block movement ------------------------------------------------------------------------ */

/* Execute a list of block copies in one launch; accumulate != 0 adds onto dst.  Bounds of every
 * block are checked against src_nelem / dst_nelem first (replaces Telum's permutedims, oplus and
 * getsub on block-sparse tensors). */
int bug_copy_blocks(bug_ctx* ctx, int32_t nop, const bug_copy* ops, const void* src,
                    int64_t src_nelem, void* dst, int64_t dst_nelem, const double* scale,
                    int32_t accumulate, void* stream);

/* Squared Frobenius norms of listed blocks (src_off, src_ld, rows, cols of each bug_copy) into
 * host out[nblk]; synchronizes the stream. */
int bug_block_sqnorms(bug_ctx* ctx, int32_t nblk, const bug_copy* blocks, const void* src,
                      int64_t src_nelem, double* out, void* stream);

/* ---- factorizations -------------------------------------------------------------------------- */

/*
 * Batched thin QR with the unique gauge (real non-negative R diagonal), one job per charge
 * sector; A is overwritten.  Stands in for the cutoff-0 SVDs of the exact gauge moves
 * (move_right!/move_left!/canonical!, src/BUG/Core/TensorNetworks/canonical_form.jl:10-54;
 * _frame_from, src/BUG/Core/Environments/effective_operators.jl:201-220).
 */
int bug_block_qr(bug_ctx* ctx, int32_t nmat, const bug_qr_job* jobs, void* a, int64_t a_nelem,
                 void* q, int64_t q_nelem, void* r, int64_t r_nelem, void* stream);

/*
 * Batched SVD (one-sided Jacobi), one job per charge sector; G is overwritten.  Singular values
 * descending; each singular pair's phase makes the first entry of the chosen factor's vector
 * with |x_i| >= 0.5 max|x| real positive (phase_on_v selects V instead of U).  The per-sector
 * factorizations of Telum's svd (deps/Telum/src/svd.jl:2013-2211); the global selection over
 * sectors (svd.jl:1813-1921) is made by the caller on the returned spectra.
 */
int bug_block_svd(bug_ctx* ctx, int32_t nmat, const bug_svd_job* jobs, void* g, int64_t g_nelem,
                  void* u, int64_t u_nelem, double* s, int64_t s_nelem, void* vh, int64_t vh_nelem,
                  int32_t max_sweeps, void* stream);

/* Column-pivoted thin QR of every job (the per-sector `qr(stacked, ColumnNorm())` of _thin_qr,
 * src/BUG/Randomised/dd_cbe.jl:184-259, LAPACK geqp3's greedy pivot rule with norms recomputed at
 * every step).  a is overwritten; q receives Q (m x k, ldq, k = min(m, n)) with the R diagonal
 * real non-negative; rdiag[r_off .. r_off + k) receives that diagonal.  ldr is ignored. */
int bug_block_qrp(bug_ctx* ctx, int32_t nmat, const bug_qr_job* jobs, void* a, int64_t a_nelem,
                  void* q, int64_t q_nelem, double* rdiag, int64_t rdiag_nelem, void* stream);

/* Diagnostics: force the SVD kernel (0 automatic, 1 one-sided Jacobi with one workgroup per
 * matrix, 2 QR-preconditioned one-workgroup Jacobi, 3 QR-preconditioned multi-workgroup block
 * Jacobi with the block pair staged in LDS, 4 QR-preconditioned Gram block Jacobi with the pair's
 * rows split over workgroups, 5 as 3 started from FP32 sweeps) and, with record_sweeps != 0,
 * record the (FP64) Jacobi sweeps each matrix of a batch used;
 * bug_svd_sweeps copies the counts of the last batch (synchronizes the device). */
int bug_set_svd_variant(bug_ctx* ctx, int32_t variant, int32_t record_sweeps);

/* Diagnostics: force the QR kernel (0 automatic, 1 Householder with one workgroup per matrix,
 * 2 blocked Householder: panels factored in registers, compact-WY updates on the grouped GEMM,
 * for blocks of at most 1024 rows). */
int bug_set_qr_variant(bug_ctx* ctx, int32_t variant);

/* Tuning of the multi-workgroup block Jacobi: columns staged per workgroup (0 = as many as the LDS
 * holds, at most 32; an explicit value only lowers it) and one-sided sweeps over the staged
 * columns per task (1..8). */
int bug_set_jacobi_params(bug_ctx* ctx, int32_t cols, int32_t inner_sweeps);

/* Tuning of the Gram block Jacobi: columns per block pair (2..64; up to 32 on 512-thread rotate
 * workgroups, the default 32, above 32 on 1024-thread ones), rotation sweeps over the pair's Gram
 * matrix per round (1..8), and the smallest matrix dimension of a batch from which the automatic
 * SVD choice uses it. */
int bug_set_gram_jacobi_params(bug_ctx* ctx, int32_t cols, int32_t inner_sweeps, int32_t min_n);
/* FP32-preconditioned block Jacobi (svd_variant 5; no counterpart in the reference, whose gesdd
 * runs in FP64 throughout): at most fp32_sweeps FP32 sweeps choose the starting basis of the
 * FP64 sweeps; variant 0 takes it for batches whose widest block has min_n or more columns. */
int bug_set_mixed_jacobi_params(bug_ctx* ctx, int32_t fp32_sweeps, int32_t min_n);
int bug_svd_sweeps(bug_ctx* ctx, int32_t n, int32_t* out);
/* The Jacobi noise floor (no counterpart in the reference, whose gesdd has no such knob): pairs
 * of columns holding together less than rel^2 ||X||_F^2 are not rotated, so singular values under
 * about rel ||X||_F are inexact (their sum of squares stays exact) and their vectors need not be
 * orthonormal.  Default 1e-16, below the 1e-14 sigma_max floor of every selection here; rel in
 * [0, 1e-6). */
int bug_set_svd_floor(bug_ctx* ctx, double rel);

/* ---- operator actions of the step (dd_cbe.jl:9-117, root_methods.jl:359-402) -------------- */

/* The streamed two-site sketch (apply_two_site_sketch!, dd_cbe.jl:70-89; _stream_left,
 * mpo_environments.jl:409-444): out = (H Theta) Om^dag for Theta = A V, without forming the
 * two-site tensor.  L: environment left of site A, R: right of site V, W1/W2 the MPO sites of A
 * and V, Om in V's layout (probe width on its row leg).  out (caller-allocated) has rows = L's
 * bra leg, columns = Om's row leg, slices = A's local states. */
int bug_two_site_sketch_left(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W1,
                             const bug_mpo_site* W2, const bug_bsm* R, const bug_bsm* A,
                             const bug_bsm* V, const bug_bsm* Om, const bug_bsm* out,
                             int32_t precision, void* stream);
/* The mirror (_stream_right, :446-481): out = Om^dag (H Theta) for Theta = U B, Om in U's layout
 * (probe width on its column leg); out has rows = Om's column leg, columns = R's bra leg, slices
 * = B's local states. */
int bug_two_site_sketch_right(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W1,
                              const bug_mpo_site* W2, const bug_bsm* R, const bug_bsm* U,
                              const bug_bsm* B, const bug_bsm* Om, const bug_bsm* out,
                              int32_t precision, void* stream);

/* A linear action on n-element complex device vectors: y = A x, enqueued on `stream`; returns 0
 * on success.  `user` is passed through. */
typedef int (*bug_apply_fn)(void* user, const void* x, void* y, void* stream);
typedef struct bug_expv_report {
    int32_t iterations;
    int32_t substeps;
    double err;                   /* the accumulated tail estimate (Saad), relative */
} bug_expv_report;
/* out = exp(tau A) v (expv, root_methods.jl:359-402; lanczos_expv / arnoldi_expv,
 * krylov.jl:128-263,343-442) with the basis of at most maxiter vectors per substep, the substep
 * search holding the tail estimate to conv_tol, full reorthogonalisation (reorth) for the Lanczos
 * form; hermitian selects Lanczos (A Hermitian) or Arnoldi.  v and out are device vectors of n
 * complex FP64 elements (out may not alias v). */
int bug_expv(bug_ctx* ctx, bug_apply_fn apply, void* user, int64_t n, double tau_re, double tau_im,
             const void* v, void* out, int32_t hermitian, double tol, int32_t maxiter,
             int32_t reorth, double conv_tol, bug_expv_report* report, void* stream);

/* Telum's global singular-value selection (svd.jl:1813-1921) on per-sector spectra (each sorted
 * descending; sector i has lens[i] values at spectra[sum_{k<i} lens[k]]): keep sv > max(cutoff,
 * 1e-14) sv_max over all sectors, then the largest nkeep (nkeep < 0: no limit) by a stable
 * descending sort in sector order.  keep[nsec] receives the kept count per sector, order[2 total]
 * the (sector, index) pairs of all values in that descending order (the first *nkept are the kept
 * list, the rest the truncated list), and weights[2] the kept and truncated sums of squares. */
int bug_svd_select(int32_t nsec, const int32_t* lens, const double* spectra, double cutoff,
                   int32_t nkeep, int32_t* keep, int32_t* order, int32_t* nkept, double* weights);

/* ---- the whole step ------------------------------------------------------------------------ */

/*
 * One rank-adaptive CBE-BUG step in place, the step-level drop-in for
 *   cbe_bug_step!(psi, mpo, tau; order, maxdim, hermitian, work_ceiling_rule, ...) -> CBEBugSweepInfo
 * (src/BUG/TimeEvolution/bug_step.jl:17-148, the Number overload :512-513).  The state and the MPO
 * are opaque handles; the state stays resident on the device between steps.  Options are a POD with
 * one field per keyword (bug_step_opts_default fills the reference's defaults); the report mirrors
 * CBEBugSweepInfo and the AdaptiveLedger verdict (src/BUG/Core/diagnostics.jl:209-255,
 * src/BUG/Core/RankAdaptation/basis_budget.jl:144-299).  Invalid options return BUG_EINVAL before
 * the state is touched (bug_step_stages.jl:41-111); a numerical shortfall is not an error but
 * tolerance_met = 0 with the failed tests in `unmet`.  Options this entry point does not carry
 * (the adaptive controller, exact mode, kept states, unfolded leaks, ranked Krylov selection) are
 * refused with BUG_EINVAL.
 */
typedef struct bug_state bug_state;
typedef struct bug_mpo bug_mpo;

#define BUG_CEIL_BOND 0         /* work_ceiling_rule = :bond */
#define BUG_CEIL_MAXDIM 1       /* :maxdim */
#define BUG_CEIL_BOND_STRICT 2  /* :bond_strict */
#define BUG_CEIL_FRAME_ROOT 3   /* :frame_root */

#define BUG_KRYLOV_AUTO 0       /* Lanczos for hermitian steps, Arnoldi otherwise */
#define BUG_KRYLOV_LANCZOS 1
#define BUG_KRYLOV_ARNOLDI 2

/* bug_step_opts.root_method: the root integrator (ROOT_METHODS, root_methods.jl:5-20) */
#define BUG_ROOT_EXP 0          /* the Krylov exponential with substeps (the reference's default) */
#define BUG_ROOT_TRAPEZOID 1    /* the implicit trapezoid's stability function, one stage */
#define BUG_ROOT_RADAU2 2       /* Radau IIA, two stages */
#define BUG_ROOT_RADAU3 3       /* Radau IIA, three stages */

/* bits of bug_step_info.unmet: which tests of the step's budget failed (_unmet!, basis_budget.jl) */
#define BUG_UNMET_SAMPLE 1
#define BUG_UNMET_KRYLOV 2
#define BUG_UNMET_CLOSURE 4
#define BUG_UNMET_PROJECTION 8
#define BUG_UNMET_EXPV 16
#define BUG_UNMET_TRUNC 32

typedef struct bug_step_opts {
    int32_t order;                /* >= 1 */
    int32_t maxdim;
    int32_t hermitian;            /* 1: Lanczos root solve; 0: Arnoldi */
    int32_t work_ceiling_rule;    /* BUG_CEIL_* */
    double bond_ceiling_factor;   /* < 0: the rule's default (1.2 for :bond, 0 otherwise) */
    double root_ceiling_factor;   /* :frame_root's root tier */
    double tol_basis;             /* < 0: 1e-5 */
    double tol_trunc;             /* < 0: 1e-7 */
    double tol_expv;              /* < 0: 1e-10 (tol_root) */
    double tol_krylov;            /* < 0: tol_basis */
    int32_t root;                 /* 0-based root bond; -1: L/2 - 1 */
    int32_t reorth;
    int32_t truncate;
    int32_t precision;            /* BUG_PREC_FP64 (the drop-in) or BUG_PREC_FP32 (mixed) */
    int32_t parallel;             /* the two half-sweeps (and frontier halves) concurrently */
    uint64_t seed;                /* probe stream seed (the reference's rng) */
    int32_t fixed_width;          /* DDQRExpansion(:rank_tail) probe width */
    double tail_factor;
    int32_t krylov_mode;          /* BUG_KRYLOV_* */
    int32_t mode;                 /* 0 :rank_tail; anything else is refused here */
    int32_t exact;                /* must be 0 here */
    int32_t fold_omega;           /* -1 default (folded leaks for order > 1), 1 folded; 0 refused */
    double krylov_rank_tol;       /* must be 0 here */
    int32_t root_method;          /* BUG_ROOT_* (bug_step.jl:76-81) */
    int32_t root_shift;           /* rational roots: 1 shift to the extreme Ritz value and drop the
                                     scale; 0 apply R(tau A_G) as it is (a linear PDE flow) */
    int32_t root_maxiter;         /* the root's Krylov basis cap (30) */
} bug_step_opts;

/* One half-sweep bond visit of the step (_log_bond!, basis_budget.jl:399-436). */
typedef struct bug_bond_log {
    int32_t bond;                 /* 0-based */
    int32_t side;                 /* 0 left half-sweep, 1 right */
    double sample, tail, krylov_c, krylov_d;
    int32_t n_probes, n_holdout, n_selected, n_targeted, n_krylov, n_h1;
    int32_t workcap, met, unmet, width;
    int32_t r_old, r_final;
    double discarded;             /* the bond's discarded norm relative to the state's */
} bug_bond_log;

typedef struct bug_step_info {
    int32_t tolerance_met;        /* budget_status === :met */
    int32_t unmet;                /* BUG_UNMET_* bits */
    double max_sample, max_tail, max_krylov, max_krylov_c, max_krylov_d;
    double discarded;             /* the compression's cumulative relative discarded norm */
    double worst_discarded;       /* CBEBugSweepInfo.discarded: the worst single cut */
    double expv_err;
    int32_t expv_iters;
    double proj_err;
    double defect;
    double leak[4];               /* closure leaks by degree (NaN: not measured) */
    int64_t n_probes, n_selected, n_krylov, n_h1, n_h1_closure, n_closure_sweeps;
    int32_t workcap_reached, maxdim_reached;
    int32_t r_old, r_aug, r_final;
    int32_t root_link, root_rank, root_applications;
    double seconds[6];            /* left sweep, right sweep, closure, root, truncation, step */
    int32_t* expanded;            /* caller array of expanded_cap >= L - 1 entries, or NULL */
    int32_t expanded_cap;
    bug_bond_log* log;            /* caller array of log_cap entries, or NULL */
    int32_t log_cap;
    int32_t log_count;            /* entries the step wrote (at most log_cap) */
} bug_step_info;

/* The reference's defaults (bug_step.jl:17-148, TOL_DEFAULTS basis_budget.jl:568). */
void bug_step_opts_default(bug_step_opts* opts);

/* MPO handle from L sites (copied; the caller keeps its arrays). */
int bug_mpo_create(int32_t L, const bug_mpo_site* sites, bug_mpo** out);
int bug_mpo_destroy(bug_mpo* mpo);

/* State handle from L site tensors A(l, s, r) given as families with HOST data (block offsets in
 * the bug_bsm convention); each site's slices are the d local states with shifts local_charges.
 * The step runs on `stream` (NULL: the default stream).  copy(psi) semantics: the handle owns its
 * device copy. */
int bug_state_create(bug_ctx* ctx, int32_t L, int32_t d, const int32_t* local_charges,
                     const bug_bsm* sites, int32_t center, void* stream, bug_state** out);
/* The same from DEVICE data (sites[i].data are device pointers of this context's device). */
int bug_state_create_device(bug_ctx* ctx, int32_t L, int32_t d, const int32_t* local_charges,
                            const bug_bsm* sites, int32_t center, void* stream, bug_state** out);
/* A second handle on the same state (the site tensors are shared: a step replaces tensors and
 * never writes into one, so the two evolve independently afterwards). */
int bug_state_clone(const bug_state* src, bug_state** out);
int bug_state_destroy(bug_state* state);
/* <O_j> / <psi|psi> at every site j of a diagonal local operator O = sum_s diag[s] |s><s| (S^z:
 * diag = {0.5, -0.5}); out[L] on the host (magnetisation / sz_expectation,
 * src/QMB/Observables/local_observables.jl:11-31).  The state is not changed. */
int bug_state_expect_diagonal(bug_ctx* ctx, const bug_state* state, const double* diag, double* out);
/* L, d and the canonical centre (0-based). */
int bug_state_info(const bug_state* state, int32_t* L, int32_t* d, int32_t* center);
/* Sector counts of the site's row and column legs and its element count. */
int bug_state_site_shape(const bug_state* state, int32_t site, int32_t* row_nsec, int32_t* col_nsec,
                         int64_t* nelem);
/* Download one site: leg tables, block offsets [d * row_nsec] (-1 absent) and the complex data
 * (column-major blocks, Julia's order) into host arrays sized by bug_state_site_shape. */
int bug_state_download_site(bug_ctx* ctx, const bug_state* state, int32_t site,
                            int32_t* row_charge, int32_t* row_dim, int32_t* col_charge,
                            int32_t* col_dim, int64_t* offset, void* host_data);
int bug_step(bug_ctx* ctx, bug_state* state, const bug_mpo* mpo, double tau_re, double tau_im,
             const bug_step_opts* opts, bug_step_info* info);
/* k independent steps with one option set, advanced together (cbe_bug_step! per instance,
 * bug_step.jl:17-510; the reference steps one state at a time and has no counterpart): the
 * stages up to the root solve run per instance on `workers` host threads, each with its own
 * library context and HIP stream; the k root exponentials run as ONE lockstep Krylov recurrence
 * (bug_apply_zero_site_batch for all roots per iteration, segmented Gram-Schmidt, each instance
 * with its own substep search and stopping rule); then each instance's assembly and
 * compression.  tau[2i], tau[2i+1] = re, im of instance i's step; infos (k entries) may be NULL.
 * Each instance's result equals its own bug_step to rounding (the lockstep sums its inner
 * products per segment).  On an error no state is changed. */
int bug_step_batch(bug_ctx* ctx, int32_t k, bug_state* const* states, const bug_mpo* const* mpos,
                   const double* tau, const bug_step_opts* opts, bug_step_info* infos,
                   int32_t workers);

#ifdef __cplusplus
}
#endif

#endif /* BUG_H */
