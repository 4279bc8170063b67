// Shared declarations of the gfx950 backend: device plan records, the context, error helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <list>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bug.h"

namespace bug {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            throw ::bug::Error(BUG_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// One product term of an output block: alpha * op_a(A) * op_b(B), A at a_off of the A operand,
// B at b_off of the B operand, inner dimension k.  ops = op_a | (op_b << 2).
struct GemmTerm {
    int64_t a_off;
    int64_t b_off;
    int32_t lda;
    int32_t ldb;
    int32_t k;
    int32_t ops;
    double alpha_re;
    double alpha_im;
};

// One output block: m x n at c_off (leading dimension ldc); its terms are
// terms[term_begin, term_end).  beta = 1 accumulates onto the block, 0 overwrites it.
struct GemmTask {
    int64_t c_off;
    int32_t ldc;
    int32_t m;
    int32_t n;
    int32_t term_begin;
    int32_t term_end;
    int32_t beta;
};

// One workgroup's output tile.
struct GemmTile {
    int32_t task;
    int32_t tm;
    int32_t tn;
    int32_t pad;
};

// One block copy (see bug_copy in bug.h).
struct CopyOp {
    int64_t src_off;
    int64_t dst_off;
    int32_t src_ld;
    int32_t dst_ld;
    int32_t rows;
    int32_t cols;
    int32_t mode;
    int32_t scale_mode;
    int64_t scale_off;
};

// One thin QR: A (m x n) -> Q (m x k), R (k x n); w_off = tau workspace offset.
struct MatJob {
    int64_t a_off, q_off, r_off, w_off;
    int32_t m, n, lda, ldq, ldr;
};

// One SVD by one-sided Jacobi (m >= n): G overwritten, V accumulated at v_off.
struct SvdJob {
    int64_t g_off, v_off, u_off, s_off, vh_off;
    int32_t m, n, ldg, ldv, ldu, ldvh, phase_on_v;
};

// How one grouped-GEMM launch runs: tile configuration (index into kTileConfigs) and the number
// of k windows each output tile is split into (1 = no split).
struct Schedule {
    int config = 0;
    int splits = 1;
};
constexpr int kMaxSplits = 8;

// A product plan: tasks and terms in host memory, plus their device copy.
struct Plan {
    std::vector<GemmTask> tasks;
    std::vector<GemmTerm> terms;
    double flops = 0.0;          // algorithmic 8 m n k summed over terms
    // device image: [tiles | tasks | terms] for tile configuration `config`, held in the
    // context's plan arena while `gen` equals the arena generation
    int config = -1;
    int32_t ntiles = 0;
    uint64_t gen = 0;
    size_t off_tasks = 0, off_terms = 0;
    std::vector<char> image;
    const GemmTile* d_tiles = nullptr;
    const GemmTask* d_tasks = nullptr;
    const GemmTerm* d_terms = nullptr;
    // the schedule chosen for this plan, valid while the tuning table generation, the tuning
    // mode and the precision it was chosen under are unchanged
    Schedule sched;
    uint64_t sched_gen = 0;
    int sched_mode = -1, sched_prec = -1;
};

// Tile configurations of the grouped GEMM (waves in m, waves in n, 16x16 MFMA tiles per wave).
struct TileConfig {
    int wm, wn, fm, fn;
    int tm() const { return 16 * fm * wm; }
    int tn() const { return 16 * fn * wn; }
    int threads() const { return 64 * wm * wn; }
};
extern const TileConfig kTileConfigs[];
extern const int kNumTileConfigs;

// Launch a grouped-GEMM plan cached under `key` (built by `build` on a miss) on A, B -> C.
void launch_cached_plan(bug_ctx* ctx, uint64_t key, const std::function<void(Plan&)>& build,
                        int precision, const void* A, const void* B, void* C, hipStream_t stream);
// The cached plan under `key` (built on a miss); launching a plan (cached or not) on A, B -> C.
Plan* cached_plan(bug_ctx* ctx, uint64_t key, const std::function<void(Plan&)>& build);
void launch_plan(bug_ctx* ctx, Plan& plan, int precision, const void* A, const void* B, void* C,
                 hipStream_t stream);

// FNV-1a hash of plan signatures.
struct Hasher {
    uint64_t h = 1469598103934665603ull;
    void bytes(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    }
    template <typename T>
    void pod(const T& v) { bytes(&v, sizeof(T)); }
};

// Launch the grouped GEMM of `plan` (device image already uploaded for `config`) with its k
// stages split into `splits` windows; `partial` holds split_partial_bytes() of scratch when
// splits > 1.
hipError_t launch_grouped_gemm(const Plan& plan, int config, int splits, int precision,
                               const void* A, const void* B, void* C, void* partial,
                               hipStream_t stream);
size_t split_partial_bytes(const Plan& plan, int config, int splits, int precision);

// Tile configuration by block-size statistics; device image of a plan for a configuration.
int choose_config(const Plan& plan);
void upload_plan(bug_ctx* ctx, Plan& plan, int config, hipStream_t stream);
// The host image [tiles | tasks | terms] of `plan` for tile configuration `config` (built once per
// configuration; a rebuild drops the plan's arena residency).
void build_plan_image(Plan& plan, int config);

// The schedule of a launch of `plan`: forced by the context, from the process-wide tuning table
// (timing the candidates first when tuning is on and the plan's class is new), or heuristic.
Schedule pick_schedule(bug_ctx* ctx, Plan& plan, int precision, const void* A, const void* B,
                       hipStream_t stream);
uint64_t plan_class(const Plan& plan, int precision);

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Grow-only per-stream device scratch of the context (slot selects one of several buffers).
void* scratch(bug_ctx* ctx, hipStream_t stream, size_t bytes, int slot = 0);
// Copy `bytes` of host memory to device memory `dev` in the order of `stream` without making the
// host wait for the stream: the bytes go through the context's pinned staging ring first.
void stage_upload(bug_ctx* ctx, hipStream_t stream, void* dev, const void* host, size_t bytes);

}  // namespace bug

struct bug_ctx {
    int device = 0;
    std::string last_error;
    int forced_config = -1;
    int forced_splits = 0;       // 0: the schedule's own split factor
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    hipEvent_t sync_ev[2] = {nullptr, nullptr};   // host waits of the Jacobi sweep loop
    std::mutex mu;
    // plan cache: signature hash -> plan (LRU)
    std::unordered_map<uint64_t, std::list<std::pair<uint64_t, bug::Plan*>>::iterator> plan_index;
    std::list<std::pair<uint64_t, bug::Plan*>> plan_lru;
    size_t plan_capacity = 4096;
    int64_t plan_hits = 0, plan_misses = 0;
    // device arena of plan images: bump-allocated, copied in stream order without blocking the
    // host; when full, the streams that launched from it since its last reset are synchronized
    // (not the device: other contexts' streams keep running), the arena reset and its
    // generation advanced
    void* plan_arena = nullptr;
    size_t plan_arena_bytes = 0, plan_arena_used = 0;
    uint64_t plan_arena_gen = 1;
    std::vector<hipStream_t> arena_streams;
    // pinned host staging ring of stage_upload; wrapping it first synchronizes the streams that
    // copied out of it since the previous wrap
    char* ring = nullptr;
    size_t ring_bytes = 0, ring_head = 0;
    std::vector<hipStream_t> ring_streams;
    double last_flops = 0.0;     // algorithmic flops of the last entry-point call
    std::vector<bug::Schedule> last_schedules;   // grouped-GEMM launches of that call
    // per-stream scratch for intermediates of the two-stage contractions and the solvers,
    // keyed by (stream, slot)
    std::unordered_map<void*, bug::Workspace> scratch[10];
    // SVD kernel choice (0 automatic, 1 plain one-sided Jacobi, 2 QR-preconditioned with one workgroup
    // per matrix, 3 QR-preconditioned multi-workgroup block Jacobi, 4 QR-preconditioned Gram block
    // Jacobi over row chunks) and an optional
    // device array receiving the sweeps used per matrix of the last batch (diagnostics)
    int svd_variant = 0;
    // set during an allocation-free call: schedules come from the tuning table or the heuristic,
    // never from timing candidates (which needs scratch memory)
    bool no_tune = false;
    int* svd_sweeps_dev = nullptr;
    // QR kernel choice (0 automatic, 1 one workgroup per matrix, 2 blocked panels where rows fit)
    int qr_variant = 0;
    // block-Jacobi tuning: columns staged per workgroup (0 = as many as the LDS holds, at most 32)
    // and one-sided sweeps over the staged columns per task
    int bj_cols = 0;
    int bj_inner = 1;
    // Gram block Jacobi: columns per block pair, inner sweeps per round, and the smallest matrix
    // dimension of a batch for which the automatic choice takes it over the staged kernel (128:
    // replayed on the SVD batches of an L = 128, D <= 1024 step it took the truncation cuts, 250-510
    // wide, in 0.77x the staged kernel's time and the 128-256 wide Krylov residuals in 0.95x; on
    // narrower blocks the staged kernel is faster)
    int gbj_cols = 32;
    int gbj_inner = 1;
    int gbj_min_n = 128;
    // fused Gram block Jacobi (svd_variant 6): the smallest widest-block width of a batch from
    // which the automatic choice takes it (never by default: on the same replay it took 1.4x the
    // staged kernel's time)
    int fbj_min_n = 1 << 30;
    // Jacobi pairs whose two columns hold less than svd_floor^2 of ||X||_F^2 are not rotated
    double svd_floor = 1e-16;
    // FP32-preconditioned block Jacobi (svd_variant 5): FP32 sweeps at most, and the smallest
    // matrix dimension of a batch for which the automatic choice (variant 0) takes it
    int mixed_fp32_sweeps = 12;
    int mixed_min_n = 1 << 30;
    // dynamic LDS bytes a workgroup may use on this device (queried on first use)
    int lds_limit = 0;
    // pinned host staging for small reductions
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // graph capture: the stream being captured (no allocation, upload, tuning or host wait may
    // happen on it), the number of live graphs that hold device addresses of this context's
    // plan arena and scratch (while it is non-zero those buffers are never freed or reused in
    // place: a buffer that must grow or reset is retired and freed when the count drops to 0)
    bool capture_active = false;
    hipStream_t capturing = nullptr;
    int graph_pins = 0;
    std::vector<void*> retired;
};

namespace bug {
// True while `stream` is being captured into a graph (the default stream is handle 0, so the
// capture flag, not a null handle, says whether a capture is on).
inline bool capturing_on(const bug_ctx* ctx, hipStream_t stream) {
    return ctx->capture_active && ctx->capturing == stream;
}

// Run f under the context lock, mapping exceptions to status codes and messages.
template <typename F>
int guarded(bug_ctx* ctx, F&& f) {
    if (!ctx) return BUG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->last_flops = 0.0;
        ctx->last_schedules.clear();
        f();
        return BUG_OK;
    } catch (const Error& e) {
        ctx->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return BUG_EINVAL;
    }
}
}  // namespace bug
