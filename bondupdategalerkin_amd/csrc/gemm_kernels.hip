// Grouped complex block GEMM for gfx950.
//
// One launch executes every output block of a block-sparse contraction: each output block is a
// "task" C = beta C + sum_t alpha_t op_a(A_t) op_b(B_t) over its own list of terms, and each
// workgroup owns one output tile of one task, so terms accumulate in registers with no atomics
// and the result is deterministic.  This replaces the per-sector-pair GEMM loop of Telum's
// abelian contraction (deps/Telum/src/contract.jl:1022-1036) with one launch per contraction
// stage.
//
// Complex arithmetic is split into four real MFMA products:
//   re += Br Ar - Bi Ai,  im += Br Ai + Bi Ar
// FP64 path: v_mfma_f64_16x16x4f64, FP64 accumulate (the drop-in).
// FP32 path: operands rounded to FP32 on load into LDS, v_mfma_f32_16x16x4f32, FP32 accumulate
// (the mixed-precision path); output written back as FP64.
//
// MFMA 16x16x4 operand layout (one wave, measured on gfx950): operand "a" lane l holds
// X[l % 16][l / 16], operand "b" lane l holds Y[l / 16][l % 16]; result register r of lane l is
// D[i][l % 16] with i = 4 (l / 16) + r for f32 and i = (l / 16) + 4 r for f64.
// Feeding a = B^T and b = A^T produces D = (A B)^T, so lane l % 16 walks the ROW index of the
// column-major output and 16 lanes store 16 consecutive complex elements.
#include "internal.h"

namespace bug {

const TileConfig kTileConfigs[] = {
    {2, 2, 2, 2},  // 64 x 64, 256 threads
    {1, 1, 2, 2},  // 32 x 32, 64 threads
    {2, 2, 1, 1},  // 32 x 32, 256 threads
    {2, 2, 4, 2},  // 128 x 64, 256 threads
    {1, 1, 1, 1},  // 16 x 16, 64 threads
};
const int kNumTileConfigs = sizeof(kTileConfigs) / sizeof(kTileConfigs[0]);

namespace {

// complex k per LDS stage: 32 for FP32 operands on 256-thread tiles, else 16 (the LDS of the
// widest FP64 tile, the prefetch registers of the 64-thread tiles)
template <typename T, int NT>
struct Stage {
    static constexpr int k = (sizeof(T) == 4 && NT >= 256) ? 32 : 16;
};
// LDS row padding: a row stride of 16 (mod 64) 4-byte banks keeps the four k-rows that one
// MFMA operand read touches in different bank quarters (conflict-free for FP32, the minimum
// two passes for FP64)
constexpr int kPad = 16;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef d4 acc_t;
    // row of D held by register r of lane group g = lane / 16
    static __device__ __forceinline__ int drow(int g, int r) { return g + 4 * r; }
    static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ int drow(int g, int r) { return 4 * g + r; }
    static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// Workgroups are dispatched round-robin over the 8 XCDs; map them so that consecutive logical
// tiles (tiles of one output block, which share operand rows and columns) land on one XCD and
// share its L2.
__device__ __forceinline__ int xcd_remap(int bid, int grid) {
    const int per = grid / 8, rem = grid % 8;
    const int xcd = bid % 8, local = bid / 8;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Partial tiles of a split-k launch: split s of tile t stores value e (of R = 8 FM FN per
// thread: FM x FN fragments x 4 rows x re/im) of thread tid at ((s ntiles + t) R + e) NT + tid,
// so each store instruction of a wave covers 64 consecutive elements.
template <int FM, int FN>
struct PartialLayout {
    static constexpr int R = 8 * FM * FN;
    static __device__ __forceinline__ int64_t at(int s, int ntiles, int t, int e, int nt, int tid) {
        return (((int64_t)s * ntiles + t) * R + e) * nt + tid;
    }
};

// grid = ntiles x splits.  splits = 1: each workgroup runs every k stage of its tile's terms and
// writes C.  splits > 1: the k stages of a tile (its terms' stages in order, ceil(k / kTK) each)
// are cut into `splits` contiguous windows, workgroup (tile, s) accumulates window s and stores
// the partial tile; split_reduce then sums the windows in order, so the result does not depend
// on how the windows were scheduled.
template <typename T, int WM, int WN, int FM, int FN>
__global__ __launch_bounds__(64 * WM * WN) void grouped_zgemm(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks,
    const GemmTerm* __restrict__ terms, const double2* __restrict__ A,
    const double2* __restrict__ B, double2* __restrict__ C, int ntiles, int splits,
    T* __restrict__ partial) {
    constexpr int NT = 64 * WM * WN;
    constexpr int TM = 16 * FM * WM;
    constexpr int TN = 16 * FN * WN;
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int kTK = Stage<T, 64 * WM * WN>::k;
    __shared__ T As_re[kTK][TM + kPad];
    __shared__ T As_im[kTK][TM + kPad];
    __shared__ T Bs_re[kTK][TN + kPad];
    __shared__ T Bs_im[kTK][TN + kPad];

    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    // consecutive logical ids: consecutive tiles of one window (they share operand panels)
    const int tile_id = logical % ntiles;
    const int split = logical / ntiles;
    const GemmTile tile = tiles[tile_id];
    const GemmTask task = tasks[tile.task];
    const int m0 = tile.tm * TM;
    const int n0 = tile.tn * TN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM;
    const int wn = wave / WM;
    const int lrow = lane & 15;
    const int lk = lane >> 4;

    acc_t acc_re[FM][FN];
    acc_t acc_im[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            acc_re[i][j] = acc_t{0, 0, 0, 0};
            acc_im[i][j] = acc_t{0, 0, 0, 0};
        }

    // this workgroup's window [s0, s1) of the task's k stages
    int s0 = 0, s1 = 1 << 30;
    if (splits > 1) {
        int total = 0;
        for (int t = task.term_begin; t < task.term_end; ++t) total += (terms[t].k + kTK - 1) / kTK;
        s0 = (int)((int64_t)total * split / splits);
        s1 = (int)((int64_t)total * (split + 1) / splits);
    }
    int g = 0;                                // stage index of the current term's first stage
    for (int t = task.term_begin; t < task.term_end; ++t) {
        const GemmTerm term = terms[t];
        const int nst = (term.k + kTK - 1) / kTK;
        const int lo = max(s0 - g, 0), hi = min(s1 - g, nst);
        g += nst;
        if (lo >= hi) continue;
        const int kbeg = lo * kTK, kend = min(hi * kTK, term.k);
        const int opa = term.ops & 3;
        const int opb = (term.ops >> 2) & 3;
        const bool a_plain = (opa == BUG_OP_N || opa == BUG_OP_R);
        const bool b_plain = (opb == BUG_OP_N || opb == BUG_OP_R);
        const double a_conj = (opa == BUG_OP_C || opa == BUG_OP_R) ? -1.0 : 1.0;
        const double b_conj = (opb == BUG_OP_C || opb == BUG_OP_R) ? -1.0 : 1.0;
        const double2* __restrict__ Ap = A + term.a_off;
        const double2* __restrict__ Bp = B + term.b_off;
        const int64_t lda = term.lda;
        const int64_t ldb = term.ldb;
        // software pipeline: the next stage's operands are fetched into registers while the
        // current stage's MFMAs run from LDS
        constexpr int LA = (TM * kTK) / NT;
        constexpr int LB = (kTK * TN) / NT;
        double2 ra[LA], rb[LB];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int u = 0; u < LA; ++u) {
                const int e = tid + u * NT;
                int i, kk;
                if (a_plain) {
                    i = e % TM;
                    kk = e / TM;
                } else {
                    kk = e % kTK;
                    i = e / kTK;
                }
                const int gi = m0 + i, gk = k0 + kk;
                ra[u] = (gi < task.m && gk < term.k) ? (a_plain ? Ap[gi + gk * lda] : Ap[gk + gi * lda])
                                                     : double2{0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int e = tid + u * NT;
                int kk, j;
                if (b_plain) {
                    kk = e % kTK;
                    j = e / kTK;
                } else {
                    j = e % TN;
                    kk = e / TN;
                }
                const int gj = n0 + j, gk = k0 + kk;
                rb[u] = (gj < task.n && gk < term.k) ? (b_plain ? Bp[gk + gj * ldb] : Bp[gj + gk * ldb])
                                                     : double2{0.0, 0.0};
            }
        };
        fetch(kbeg);
        for (int k0 = kbeg; k0 < kend; k0 += kTK) {
            // op_a(A) scaled by alpha and op_b(B) -> LDS (k-major)
#pragma unroll
            for (int u = 0; u < LA; ++u) {
                const int e = tid + u * NT;
                int i, kk;
                if (a_plain) {
                    i = e % TM;
                    kk = e / TM;
                } else {
                    kk = e % kTK;
                    i = e / kTK;
                }
                const double2 v = ra[u];
                const double vi = a_conj * v.y;
                As_re[kk][i] = static_cast<T>(term.alpha_re * v.x - term.alpha_im * vi);
                As_im[kk][i] = static_cast<T>(term.alpha_re * vi + term.alpha_im * v.x);
            }
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int e = tid + u * NT;
                int kk, j;
                if (b_plain) {
                    kk = e % kTK;
                    j = e / kTK;
                } else {
                    j = e % TN;
                    kk = e / TN;
                }
                Bs_re[kk][j] = static_cast<T>(rb[u].x);
                Bs_im[kk][j] = static_cast<T>(b_conj * rb[u].y);
            }
            __syncthreads();
            if (k0 + kTK < kend) fetch(k0 + kTK);
#pragma unroll
            for (int kk = 0; kk < kTK; kk += 4) {
                T a_re[FM], a_im[FM], b_re[FN], b_im[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    const int r = wm * 16 * FM + i * 16 + lrow;
                    a_re[i] = As_re[kk + lk][r];
                    a_im[i] = As_im[kk + lk][r];
                }
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const int c = wn * 16 * FN + j * 16 + lrow;
                    b_re[j] = Bs_re[kk + lk][c];
                    b_im[j] = Bs_im[kk + lk][c];
                }
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        acc_re[i][j] = Mfma<T>::step(b_re[j], a_re[i], acc_re[i][j]);
                        acc_re[i][j] = Mfma<T>::step(-b_im[j], a_im[i], acc_re[i][j]);
                        acc_im[i][j] = Mfma<T>::step(b_re[j], a_im[i], acc_im[i][j]);
                        acc_im[i][j] = Mfma<T>::step(b_im[j], a_re[i], acc_im[i][j]);
                    }
            }
            __syncthreads();
        }
    }

    if (splits > 1) {
        typedef PartialLayout<FM, FN> PL;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = ((i * FN + j) * 4 + r) * 2;
                    partial[PL::at(split, ntiles, tile_id, e, NT, tid)] = acc_re[i][j][r];
                    partial[PL::at(split, ntiles, tile_id, e + 1, NT, tid)] = acc_im[i][j][r];
                }
        return;
    }
    double2* __restrict__ Cp = C + task.c_off;
    const int64_t ldc = task.ldc;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * 16 * FM + i * 16 + lrow;
                const int gn = n0 + wn * 16 * FN + j * 16 + Mfma<T>::drow(lk, r);
                if (gm < task.m && gn < task.n) {
                    double2 out;
                    out.x = static_cast<double>(acc_re[i][j][r]);
                    out.y = static_cast<double>(acc_im[i][j][r]);
                    double2* dst = Cp + gm + gn * ldc;
                    if (task.beta) {
                        const double2 old = *dst;
                        out.x += old.x;
                        out.y += old.y;
                    }
                    *dst = out;
                }
            }
}

// Sum the split windows of each tile in window order (in FP64) and write C like the one-pass
// epilogue of grouped_zgemm (same lane -> element map, same beta rule).
template <typename T, int WM, int WN, int FM, int FN>
__global__ __launch_bounds__(64 * WM * WN) void split_reduce(
    const GemmTile* __restrict__ tiles, const GemmTask* __restrict__ tasks,
    double2* __restrict__ C, int ntiles, int splits, const T* __restrict__ partial) {
    constexpr int NT = 64 * WM * WN;
    constexpr int TM = 16 * FM * WM;
    constexpr int TN = 16 * FN * WN;
    typedef PartialLayout<FM, FN> PL;
    const int tile_id = blockIdx.x;
    const GemmTile tile = tiles[tile_id];
    const GemmTask task = tasks[tile.task];
    const int m0 = tile.tm * TM, n0 = tile.tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int lrow = lane & 15, lk = lane >> 4;
    double2* __restrict__ Cp = C + task.c_off;
    const int64_t ldc = task.ldc;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = ((i * FN + j) * 4 + r) * 2;
                double2 out{0.0, 0.0};
                for (int s = 0; s < splits; ++s) {
                    out.x += static_cast<double>(partial[PL::at(s, ntiles, tile_id, e, NT, tid)]);
                    out.y += static_cast<double>(partial[PL::at(s, ntiles, tile_id, e + 1, NT, tid)]);
                }
                const int gm = m0 + wm * 16 * FM + i * 16 + lrow;
                const int gn = n0 + wn * 16 * FN + j * 16 + Mfma<T>::drow(lk, r);
                if (gm < task.m && gn < task.n) {
                    double2* dst = Cp + gm + gn * ldc;
                    if (task.beta) {
                        const double2 old = *dst;
                        out.x += old.x;
                        out.y += old.y;
                    }
                    *dst = out;
                }
            }
}

template <typename T, int WM, int WN, int FM, int FN>
hipError_t launch_one(const Plan& plan, int splits, const void* A, const void* B, void* C,
                      void* partial, hipStream_t stream) {
    const dim3 block(64 * WM * WN);
    hipLaunchKernelGGL((grouped_zgemm<T, WM, WN, FM, FN>), dim3(plan.ntiles * splits), block, 0,
                       stream, plan.d_tiles, plan.d_tasks, plan.d_terms,
                       static_cast<const double2*>(A), static_cast<const double2*>(B),
                       static_cast<double2*>(C), plan.ntiles, splits, static_cast<T*>(partial));
    if (splits > 1)
        hipLaunchKernelGGL((split_reduce<T, WM, WN, FM, FN>), dim3(plan.ntiles), block, 0, stream,
                           plan.d_tiles, plan.d_tasks, static_cast<double2*>(C), plan.ntiles,
                           splits, static_cast<const T*>(partial));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_typed(const Plan& plan, int config, int splits, const void* A, const void* B,
                        void* C, void* partial, hipStream_t stream) {
    switch (config) {
        case 0: return launch_one<T, 2, 2, 2, 2>(plan, splits, A, B, C, partial, stream);
        case 1: return launch_one<T, 1, 1, 2, 2>(plan, splits, A, B, C, partial, stream);
        case 2: return launch_one<T, 2, 2, 1, 1>(plan, splits, A, B, C, partial, stream);
        case 3: return launch_one<T, 2, 2, 4, 2>(plan, splits, A, B, C, partial, stream);
        case 4: return launch_one<T, 1, 1, 1, 1>(plan, splits, A, B, C, partial, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

size_t split_partial_bytes(const Plan& plan, int config, int splits, int precision) {
    if (splits <= 1) return 0;
    const TileConfig& tc = kTileConfigs[config];
    const size_t elem = precision == BUG_PREC_FP32 ? sizeof(float) : sizeof(double);
    return (size_t)plan.ntiles * splits * tc.tm() * tc.tn() * 2 * elem;
}

hipError_t launch_grouped_gemm(const Plan& plan, int config, int splits, int precision,
                               const void* A, const void* B, void* C, void* partial,
                               hipStream_t stream) {
    if (plan.ntiles == 0) return hipSuccess;
    if (splits < 1 || (splits > 1 && !partial)) return hipErrorInvalidValue;
    if (precision == BUG_PREC_FP32)
        return launch_typed<float>(plan, config, splits, A, B, C, partial, stream);
    return launch_typed<double>(plan, config, splits, A, B, C, partial, stream);
}

}  // namespace bug
