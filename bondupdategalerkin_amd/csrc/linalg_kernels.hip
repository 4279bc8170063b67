// Vector, copy and factorization kernels of the BUG hot path for gfx950.
//
// * K10 vector kernels: complex axpy / scale, and a fused multi-dot <x_j | y> (j < 64) that reads
//   y once per block (Krylov inner products, CGS2 passes, norms).  HBM-bound.
// * Block copies: one launch moves any list of (sub)blocks between two buffers, optionally
//   transposed / conjugated, with an optional real per-row or per-column scale (matricisations of
//   block-sparse tensors, direct sums, trims, S Vh products).
// * Batched thin QR (Householder, one workgroup per matrix) with the unique positive-diagonal
//   gauge: the gauge moves of the step (canonical_form.jl:10-54, effective_operators.jl:201-220).
// * Batched SVD by one-sided (Hestenes) Jacobi with Brent-Luk round-robin pair ordering, one
//   workgroup per matrix, singular values sorted descending and a deterministic phase per
//   singular pair: the per-sector SVDs of Telum's svd (deps/Telum/src/svd.jl:2013-2211).
//
// Waves are 64 lanes; every reduction is a wave butterfly (__shfl_xor) so the sum lands in every
// lane.  No atomics; results are deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace bug {
namespace {

// Move a double across lanes by one DPP control (two 32-bit halves); every lane of the wave must
// be active.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes, returned in every lane: within each 16-lane row by DPP (quad swaps, then
// the half-row and row mirrors), across the four rows by lane reads.  Every lane must be active.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_mov<0xb1>(v);    // quad_perm [1,0,3,2]
    v += dpp_mov<0x4e>(v);    // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);   // row_half_mirror
    v += dpp_mov<0x140>(v);   // row_mirror
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

template <int CTRL>
__device__ __forceinline__ float dpp_movf(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ float read_lanef(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// The FP32 wave sum, with the reduction pattern of the FP64 one.
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_movf<0xb1>(v);
    v += dpp_movf<0x4e>(v);
    v += dpp_movf<0x141>(v);
    v += dpp_movf<0x140>(v);
    return (read_lanef(v, 0) + read_lanef(v, 16)) + (read_lanef(v, 32) + read_lanef(v, 48));
}

// complex element type of a real scalar type
template <typename R>
struct Cplx;
template <>
struct Cplx<double> {
    typedef double2 type;
};
template <>
struct Cplx<float> {
    typedef float2 type;
};

__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_mov<0xb1>(v));
    v = fmax(v, dpp_mov<0x4e>(v));
    v = fmax(v, dpp_mov<0x141>(v));
    v = fmax(v, dpp_mov<0x140>(v));
    return fmax(fmax(read_lane(v, 0), read_lane(v, 16)), fmax(read_lane(v, 32), read_lane(v, 48)));
}

// ---- K10: axpy / scale / multi-dot --------------------------------------------------------

__global__ void zaxpy_kernel(int64_t n, double ar, double ai, const double2* __restrict__ x,
                             double2* __restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = x[i];
        double2 w = y[i];
        w.x += ar * v.x - ai * v.y;
        w.y += ar * v.y + ai * v.x;
        y[i] = w;
    }
}

__global__ void zscal_kernel(int64_t n, double ar, double ai, double2* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = x[i];
        x[i] = double2{ar * v.x - ai * v.y, ar * v.y + ai * v.x};
    }
}

constexpr int kMaxDots = 64;
constexpr int kDotThreads = 256;
constexpr int kDotRows = 4;  // rows per thread per block

struct DotPtrs {
    const double2* p[kMaxDots];
};

struct AxpyCoeffs {
    double2 a[kMaxDots];
};

// y += sum_j a_j x_j, the terms added one after another in j order, each as zaxpy_kernel adds it
// (the same per-element arithmetic as k zaxpy launches, in one pass over y).
__global__ void zaxpy_many_kernel(int64_t n, int k, AxpyCoeffs c, DotPtrs xs,
                                  double2* __restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double2 w = y[i];
        for (int j = 0; j < k; ++j) {
            const double ar = c.a[j].x, ai = c.a[j].y;
            const double2 v = xs.p[j][i];
            w.x += ar * v.x - ai * v.y;
            w.y += ar * v.y + ai * v.x;
        }
        y[i] = w;
    }
}

// ---- segmented vector operations (K independent vectors packed in one buffer) -------------
//
// Segment i of a packed buffer is [off_i, off_i + len_i).  The multi-dot reduces each
// (vector j, segment i) over fixed chunks of kSegChunk elements and then over the chunks in
// order; the multi-axpy adds sum_j a[j][i] x_j to segment i of y, the terms in j order.

constexpr int kSegChunk = 4096;
constexpr int kSegThreads = 256;

__global__ __launch_bounds__(kSegThreads) void seg_multidot_partial(
    int nseg, const int64_t* __restrict__ segs, int k, DotPtrs xs, const double2* __restrict__ y,
    int nchunk, double2* __restrict__ partial) {
    __shared__ double red_re[kSegThreads / 64];
    __shared__ double red_im[kSegThreads / 64];
    const int i = blockIdx.x, ch = blockIdx.y;
    const int64_t off = segs[2 * i], len = segs[2 * i + 1];
    const int64_t lo = (int64_t)ch * kSegChunk, hi = min(len, lo + kSegChunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < k; ++j) {
        double re = 0.0, im = 0.0;
        for (int64_t e = lo + threadIdx.x; e < hi; e += kSegThreads) {
            const double2 xv = xs.p[j][off + e], yv = y[off + e];
            re += xv.x * yv.x + xv.y * yv.y;
            im += xv.x * yv.y - xv.y * yv.x;
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if (lane == 0) {
            red_re[wave] = re;
            red_im[wave] = im;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double sr = 0.0, si = 0.0;
            for (int w = 0; w < kSegThreads / 64; ++w) {
                sr += red_re[w];
                si += red_im[w];
            }
            partial[((int64_t)j * nseg + i) * nchunk + ch] = double2{sr, si};
        }
        __syncthreads();
    }
}

__global__ void seg_multidot_final(int nchunk, const double2* __restrict__ partial,
                                   double2* __restrict__ out) {
    const int64_t r = blockIdx.x;       // (j, i) row
    if (threadIdx.x == 0) {
        double re = 0.0, im = 0.0;
        for (int c = 0; c < nchunk; ++c) {
            re += partial[r * nchunk + c].x;
            im += partial[r * nchunk + c].y;
        }
        out[r] = double2{re, im};
    }
}

__global__ void seg_axpy_many_kernel(int nseg, const int64_t* __restrict__ segs, int k,
                                     const double2* __restrict__ a, DotPtrs xs,
                                     double2* __restrict__ y) {
    const int i = blockIdx.x;
    const int64_t off = segs[2 * i], len = segs[2 * i + 1];
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < len;
         e += (int64_t)gridDim.y * blockDim.x) {
        double2 w = y[off + e];
        for (int j = 0; j < k; ++j) {
            const double ar = a[(int64_t)j * nseg + i].x, ai = a[(int64_t)j * nseg + i].y;
            const double2 v = xs.p[j][off + e];
            w.x += ar * v.x - ai * v.y;
            w.y += ar * v.y + ai * v.x;
        }
        y[off + e] = w;
    }
}

// partial[j * nblk + b] = sum over block b's rows of conj(x_j) y  (re, im interleaved)
__global__ __launch_bounds__(kDotThreads) void multidot_partial(int64_t n, int k, DotPtrs xs,
                                                                 const double2* __restrict__ y,
                                                                 double2* __restrict__ partial) {
    __shared__ double red_re[kDotThreads / 64];
    __shared__ double red_im[kDotThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kDotThreads * kDotRows;
    double2 yv[kDotRows];
    int64_t idx[kDotRows];
#pragma unroll
    for (int r = 0; r < kDotRows; ++r) {
        idx[r] = base + threadIdx.x + (int64_t)r * kDotThreads;
        yv[r] = idx[r] < n ? y[idx[r]] : double2{0.0, 0.0};
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < k; ++j) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int r = 0; r < kDotRows; ++r) {
            if (idx[r] < n) {
                const double2 xv = xs.p[j][idx[r]];
                re += xv.x * yv[r].x + xv.y * yv[r].y;
                im += xv.x * yv[r].y - xv.y * yv[r].x;
            }
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if (lane == 0) {
            red_re[wave] = re;
            red_im[wave] = im;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double sr = 0.0, si = 0.0;
            for (int w = 0; w < kDotThreads / 64; ++w) {
                sr += red_re[w];
                si += red_im[w];
            }
            partial[(int64_t)j * gridDim.x + blockIdx.x] = double2{sr, si};
        }
        __syncthreads();
    }
}

__global__ void multidot_final(int nblk, const double2* __restrict__ partial,
                               double2* __restrict__ out) {
    const int j = blockIdx.x;
    double re = 0.0, im = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 64) {
        const double2 v = partial[(int64_t)j * nblk + b];
        re += v.x;
        im += v.y;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (threadIdx.x == 0) out[j] = double2{re, im};
}

// ---- block copies -------------------------------------------------------------------------

// dst(r, c) = scale * op(src)(r, c) for r < rows, c < cols; op by mode:
// 0 copy, 1 transpose, 2 conj-transpose, 3 conj.  Optional real scale from `scale_base`:
// scale_mode 1 multiplies row r by s[scale_off + r], 2 column c by s[scale_off + c].
__global__ void copy_blocks_kernel(const CopyOp* __restrict__ ops, const double2* __restrict__ src,
                                   double2* __restrict__ dst, const double* __restrict__ scale_base,
                                   int accumulate) {
    const CopyOp op = ops[blockIdx.x];
    const int64_t total = (int64_t)op.rows * op.cols;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.y * blockDim.x) {
        const int r = (int)(e % op.rows);
        const int c = (int)(e / op.rows);
        double2 v;
        if (op.mode == 0 || op.mode == 3)
            v = src[op.src_off + r + (int64_t)c * op.src_ld];
        else
            v = src[op.src_off + c + (int64_t)r * op.src_ld];
        if (op.mode == 2 || op.mode == 3) v.y = -v.y;
        if (op.scale_mode == 1) {
            const double s = scale_base[op.scale_off + r];
            v.x *= s;
            v.y *= s;
        } else if (op.scale_mode == 2) {
            const double s = scale_base[op.scale_off + c];
            v.x *= s;
            v.y *= s;
        }
        double2* d = dst + op.dst_off + r + (int64_t)c * op.dst_ld;
        if (accumulate) {
            const double2 o = *d;
            v.x += o.x;
            v.y += o.y;
        }
        *d = v;
    }
}

// partial squared Frobenius norms of the listed blocks (CopyOp fields: src_off, src_ld, rows,
// cols): workgroup (b, y) of a (nblk, chunks) grid sums the elements y, y + chunks, ... of block b
// in units of its 256 threads and writes out[b * chunks + y]; the caller adds the chunks in order
__global__ void block_sqnorm_kernel(const CopyOp* __restrict__ ops, const double2* __restrict__ src,
                                    double* __restrict__ out) {
    __shared__ double red[4];
    const CopyOp op = ops[blockIdx.x];
    const int64_t total = (int64_t)op.rows * op.cols;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.y * blockDim.x) {
        const int r = (int)(e % op.rows);
        const int c = (int)(e / op.rows);
        const double2 v = src[op.src_off + r + (int64_t)c * op.src_ld];
        acc += v.x * v.x + v.y * v.y;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        out[(int64_t)blockIdx.x * gridDim.y + blockIdx.y] = red[0] + red[1] + red[2] + red[3];
}

// ---- batched thin QR (Householder, unique gauge) ------------------------------------------

constexpr int kQrThreads = 1024;

// A (m x n, lda) is overwritten; Q (m x k, ldq) and R (k x n, ldr), k = min(m, n).
// R has a real non-negative diagonal (the unique thin QR of a full-rank A).
__global__ __launch_bounds__(kQrThreads) void qr_kernel(const MatJob* __restrict__ jobs,
                                                          double2* __restrict__ A,
                                                          double2* __restrict__ Q,
                                                          double2* __restrict__ R,
                                                          double2* __restrict__ tau_ws) {
    const MatJob job = jobs[blockIdx.x];
    const int m = job.m, n = job.n, k = min(m, n);
    double2* a = A + job.a_off;
    double2* q = Q + job.q_off;
    double2* rr = R + job.r_off;
    double2* tau = tau_ws + job.w_off;
    const int64_t lda = job.lda, ldq = job.ldq, ldr = job.ldr;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int nw = kQrThreads / 64;
    __shared__ double2 s_tau;

    for (int j = 0; j < k; ++j) {
        double2* col = a + (int64_t)j * lda;
        if (wave == 0) {
            double x2 = 0.0;
            for (int i = j + 1 + lane; i < m; i += 64) x2 += col[i].x * col[i].x + col[i].y * col[i].y;
            x2 = wave_sum(x2);
            const double2 alpha = col[j];
            double2 t{0.0, 0.0};
            if (!(x2 == 0.0 && alpha.y == 0.0)) {
                const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + x2);
                const double beta = alpha.x >= 0.0 ? -nrm : nrm;
                t = double2{(beta - alpha.x) / beta, -alpha.y / beta};
                // v = x / (alpha - beta)
                const double dr = alpha.x - beta, di = alpha.y;
                const double den = dr * dr + di * di;
                const double sr = dr / den, si = -di / den;
                for (int i = j + 1 + lane; i < m; i += 64) {
                    const double2 v = col[i];
                    col[i] = double2{v.x * sr - v.y * si, v.x * si + v.y * sr};
                }
                if (lane == 0) col[j] = double2{beta, 0.0};
            }
            if (lane == 0) {
                s_tau = t;
                tau[j] = t;
            }
        }
        __syncthreads();
        const double2 t = s_tau;
        // A <- (I - conj(tau) v v^H) A on columns j+1 .. n-1
        if (t.x != 0.0 || t.y != 0.0) {
            for (int c = j + 1 + wave; c < n; c += nw) {
                double2* ac = a + (int64_t)c * lda;
                double wr = 0.0, wi = 0.0;
                for (int i = j + lane; i < m; i += 64) {
                    const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                    const double2 x = ac[i];
                    wr += v.x * x.x + v.y * x.y;
                    wi += v.x * x.y - v.y * x.x;
                }
                wr = wave_sum(wr);
                wi = wave_sum(wi);
                // f = conj(tau) * w
                const double fr = t.x * wr + t.y * wi;
                const double fi = t.x * wi - t.y * wr;
                for (int i = j + lane; i < m; i += 64) {
                    const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                    double2 x = ac[i];
                    x.x -= fr * v.x - fi * v.y;
                    x.y -= fr * v.y + fi * v.x;
                    ac[i] = x;
                }
            }
        }
        __syncthreads();
    }
    // R = upper triangle (k x n)
    for (int64_t e = threadIdx.x; e < (int64_t)k * n; e += kQrThreads) {
        const int r = (int)(e % k), c = (int)(e / k);
        rr[r + c * ldr] = (r <= c) ? a[r + (int64_t)c * lda] : double2{0.0, 0.0};
    }
    // Q = H_0 ... H_{k-1} [I_k; 0]
    for (int64_t e = threadIdx.x; e < (int64_t)m * k; e += kQrThreads) {
        const int r = (int)(e % m), c = (int)(e / m);
        q[r + c * ldq] = double2{r == c ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    for (int j = k - 1; j >= 0; --j) {
        const double2 t = tau[j];
        if (t.x == 0.0 && t.y == 0.0) continue;
        const double2* col = a + (int64_t)j * lda;
        for (int c = j + wave; c < k; c += nw) {
            double2* qc = q + (int64_t)c * ldq;
            double wr = 0.0, wi = 0.0;
            for (int i = j + lane; i < m; i += 64) {
                const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                const double2 x = qc[i];
                wr += v.x * x.x + v.y * x.y;
                wi += v.x * x.y - v.y * x.x;
            }
            wr = wave_sum(wr);
            wi = wave_sum(wi);
            // f = tau * w
            const double fr = t.x * wr - t.y * wi;
            const double fi = t.x * wi + t.y * wr;
            for (int i = j + lane; i < m; i += 64) {
                const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                double2 x = qc[i];
                x.x -= fr * v.x - fi * v.y;
                x.y -= fr * v.y + fi * v.x;
                qc[i] = x;
            }
        }
        __syncthreads();
    }
    // unique gauge: make every R diagonal entry non-negative
    for (int j = wave; j < k; j += nw) {
        if (rr[j + (int64_t)j * ldr].x < 0.0) {
            for (int c = j + lane; c < n; c += 64) {
                double2 v = rr[j + (int64_t)c * ldr];
                rr[j + (int64_t)c * ldr] = double2{-v.x, -v.y};
            }
            for (int i = lane; i < m; i += 64) {
                double2 v = q[i + (int64_t)j * ldq];
                q[i + (int64_t)j * ldq] = double2{-v.x, -v.y};
            }
        }
    }
}

// Column-pivoted Householder QR of A (m x n, lda), overwritten: at step j the pivot is the
// column of largest remaining norm over rows j..m-1, the first of equals (LAPACK geqp3's greedy
// rule; the norms are recomputed at every step instead of downdated).  Outputs Q (m x k, ldq,
// k = min(m, n)) with the R diagonal made real non-negative and rdiag[r_off + j] = R_jj.
__global__ __launch_bounds__(kQrThreads) void qrp_kernel(const MatJob* __restrict__ jobs,
                                                           double2* __restrict__ A,
                                                           double2* __restrict__ Q,
                                                           double* __restrict__ rdiag,
                                                           double2* __restrict__ tau_ws) {
    extern __shared__ double cn[];            // squared norms of the remaining columns (n)
    const MatJob job = jobs[blockIdx.x];
    const int m = job.m, n = job.n, k = min(m, n);
    double2* a = A + job.a_off;
    double2* q = Q + job.q_off;
    double* rd = rdiag + job.r_off;
    double2* tau = tau_ws + job.w_off;
    const int64_t lda = job.lda, ldq = job.ldq;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int nw = kQrThreads / 64;
    __shared__ double2 s_tau;
    __shared__ int s_piv;

    for (int j = 0; j < k; ++j) {
        for (int c = j + wave; c < n; c += nw) {
            const double2* ac = a + (int64_t)c * lda;
            double s2 = 0.0;
            for (int i = j + lane; i < m; i += 64) s2 += ac[i].x * ac[i].x + ac[i].y * ac[i].y;
            s2 = wave_sum(s2);
            if (lane == 0) cn[c] = s2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int p = j;
            double best = cn[j];
            for (int c = j + 1; c < n; ++c)
                if (cn[c] > best) {
                    best = cn[c];
                    p = c;
                }
            s_piv = p;
        }
        __syncthreads();
        const int p = s_piv;
        if (p != j) {
            double2* aj = a + (int64_t)j * lda;
            double2* ap = a + (int64_t)p * lda;
            for (int i = threadIdx.x; i < m; i += kQrThreads) {
                const double2 t = aj[i];
                aj[i] = ap[i];
                ap[i] = t;
            }
        }
        __syncthreads();
        double2* col = a + (int64_t)j * lda;
        if (wave == 0) {
            double x2 = 0.0;
            for (int i = j + 1 + lane; i < m; i += 64) x2 += col[i].x * col[i].x + col[i].y * col[i].y;
            x2 = wave_sum(x2);
            const double2 alpha = col[j];
            double2 t{0.0, 0.0};
            if (!(x2 == 0.0 && alpha.y == 0.0)) {
                const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + x2);
                const double beta = alpha.x >= 0.0 ? -nrm : nrm;
                t = double2{(beta - alpha.x) / beta, -alpha.y / beta};
                const double dr = alpha.x - beta, di = alpha.y;
                const double den = dr * dr + di * di;
                const double sr = dr / den, si = -di / den;
                for (int i = j + 1 + lane; i < m; i += 64) {
                    const double2 v = col[i];
                    col[i] = double2{v.x * sr - v.y * si, v.x * si + v.y * sr};
                }
                if (lane == 0) col[j] = double2{beta, 0.0};
            }
            if (lane == 0) {
                s_tau = t;
                tau[j] = t;
            }
        }
        __syncthreads();
        const double2 t = s_tau;
        if (t.x != 0.0 || t.y != 0.0) {
            for (int c = j + 1 + wave; c < n; c += nw) {
                double2* ac = a + (int64_t)c * lda;
                double wr = 0.0, wi = 0.0;
                for (int i = j + lane; i < m; i += 64) {
                    const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                    const double2 x = ac[i];
                    wr += v.x * x.x + v.y * x.y;
                    wi += v.x * x.y - v.y * x.x;
                }
                wr = wave_sum(wr);
                wi = wave_sum(wi);
                const double fr = t.x * wr + t.y * wi;
                const double fi = t.x * wi - t.y * wr;
                for (int i = j + lane; i < m; i += 64) {
                    const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                    double2 x = ac[i];
                    x.x -= fr * v.x - fi * v.y;
                    x.y -= fr * v.y + fi * v.x;
                    ac[i] = x;
                }
            }
        }
        __syncthreads();
    }
    for (int64_t e = threadIdx.x; e < (int64_t)m * k; e += kQrThreads) {
        const int r = (int)(e % m), c = (int)(e / m);
        q[r + c * ldq] = double2{r == c ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    for (int j = k - 1; j >= 0; --j) {
        const double2 t = tau[j];
        if (t.x == 0.0 && t.y == 0.0) continue;
        const double2* col = a + (int64_t)j * lda;
        for (int c = j + wave; c < k; c += nw) {
            double2* qc = q + (int64_t)c * ldq;
            double wr = 0.0, wi = 0.0;
            for (int i = j + lane; i < m; i += 64) {
                const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                const double2 x = qc[i];
                wr += v.x * x.x + v.y * x.y;
                wi += v.x * x.y - v.y * x.x;
            }
            wr = wave_sum(wr);
            wi = wave_sum(wi);
            const double fr = t.x * wr - t.y * wi;
            const double fi = t.x * wi + t.y * wr;
            for (int i = j + lane; i < m; i += 64) {
                const double2 v = (i == j) ? double2{1.0, 0.0} : col[i];
                double2 x = qc[i];
                x.x -= fr * v.x - fi * v.y;
                x.y -= fr * v.y + fi * v.x;
                qc[i] = x;
            }
        }
        __syncthreads();
    }
    // positive-diagonal gauge on Q, and the diagonal's magnitude
    for (int j = wave; j < k; j += nw) {
        const double d = a[j + (int64_t)j * lda].x;
        if (d < 0.0)
            for (int i = lane; i < m; i += 64) {
                const double2 v = q[i + (int64_t)j * ldq];
                q[i + (int64_t)j * ldq] = double2{-v.x, -v.y};
            }
        if (lane == 0) rd[j] = fabs(d);
    }
}

// ---- blocked Householder QR (panels + grouped-GEMM updates) -------------------------------
//
// Panels of PW columns (16, or 8 for blocks of more than 512 rows so that a lane's RPT rows stay in
// registers).  The panel kernel (one workgroup per matrix, one wave per panel column, the
// column's rows held in registers, RPT per lane) factors the panel with the same
// reflectors as qr_kernel and builds the compact-WY factor T (H_p0 ... H_p0+jb-1 = I - V T V^H).
// The trailing columns are updated slab by slab, W = T^H (V^H A_trail) and A_trail -= V W;
// Q = H_0 ... H_{k-1} [I; 0] is formed backwards per panel with W = T (V^H Q_sub),
// Q_sub -= V W.  qr_finish_kernel extracts R and applies the positive-diagonal gauge.

constexpr int kQrPanel = 16;   // widest panel

struct QrPanelJob {
    int64_t a_off, v_off, pad0;    // A (lda); V: m x k column-major, leading dimension m
    int32_t m, n, lda, pad;
};

// T of one panel of one matrix: kQrPanel x kQrPanel row-major, the jb x jb upper triangle used
constexpr int kQrTSlot = kQrPanel * kQrPanel;

template <int RPT, int PW>
__global__ __launch_bounds__(64 * PW) void qr_panel_kernel(const QrPanelJob* __restrict__ jobs, int p0,
                                                           double2* __restrict__ A,
                                                           double2* __restrict__ Vb,
                                                           double2* __restrict__ Tp) {
    __shared__ double2 vbuf[2][64 * RPT];
    __shared__ double2 s_tau[PW];
    __shared__ double2 Gs[PW][PW];
    __shared__ double2 Ts[PW][PW];
    const QrPanelJob job = jobs[blockIdx.x];
    const int m = job.m, n = job.n, k = min(m, n);
    if (p0 >= k) return;
    const int jb = min(PW, k - p0);
    const int mp = m - p0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t lda = job.lda;
    const bool own = wave < jb;
    double2* col = A + job.a_off + p0 + (int64_t)(p0 + (own ? wave : 0)) * lda;
    double2 x[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
        const int r = lane + 64 * i;
        x[i] = (own && r < mp) ? col[r] : double2{0.0, 0.0};
    }
    for (int j = 0; j < jb; ++j) {
        double2* buf = vbuf[j & 1];
        if (wave == j) {
            double x2 = 0.0;
#pragma unroll
            for (int i = 0; i < RPT; ++i)
                if (lane + 64 * i > j) x2 += x[i].x * x[i].x + x[i].y * x[i].y;
            x2 = wave_sum(x2);
            const double ar = __shfl(x[0].x, j, 64), ai = __shfl(x[0].y, j, 64);  // j < 64
            const bool refl = !(x2 == 0.0 && ai == 0.0);
            double2 t{0.0, 0.0};
            double beta = 0.0, sr = 0.0, si = 0.0;
            if (refl) {
                const double nrm = sqrt(ar * ar + ai * ai + x2);
                beta = ar >= 0.0 ? -nrm : nrm;
                t = double2{(beta - ar) / beta, -ai / beta};
                const double dr = ar - beta, di = ai;
                const double den = dr * dr + di * di;
                sr = dr / den;
                si = -di / den;
            }
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int r = lane + 64 * i;
                double2 vv{0.0, 0.0};
                if (r == j) {
                    vv = double2{1.0, 0.0};
                    if (refl) x[i] = double2{beta, 0.0};
                } else if (r > j) {
                    if (refl) {
                        const double2 u = x[i];
                        x[i] = double2{u.x * sr - u.y * si, u.x * si + u.y * sr};
                        vv = x[i];
                    } else {
                        x[i] = double2{0.0, 0.0};
                    }
                }
                buf[r] = vv;
            }
            if (lane == 0) s_tau[j] = t;
        }
        __syncthreads();
        const double2 t = s_tau[j];
        if (wave > j && own && (t.x != 0.0 || t.y != 0.0)) {
            // x <- (I - conj(tau) v v^H) x
            double wr = 0.0, wi = 0.0;
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const double2 v = buf[lane + 64 * i];
                wr += v.x * x[i].x + v.y * x[i].y;
                wi += v.x * x[i].y - v.y * x[i].x;
            }
            wr = wave_sum(wr);
            wi = wave_sum(wi);
            const double fr = t.x * wr + t.y * wi;
            const double fi = t.x * wi - t.y * wr;
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const double2 v = buf[lane + 64 * i];
                x[i].x -= fr * v.x - fi * v.y;
                x[i].y -= fr * v.y + fi * v.x;
            }
        } else if (wave < j) {
            // G[wave][j] = v_wave^H v_j, v_wave = (0 above, 1 at wave, x below)
            double gr = 0.0, gi = 0.0;
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int r = lane + 64 * i;
                if (r < wave) continue;
                const double2 v = r == wave ? double2{1.0, 0.0} : x[i];
                const double2 b = buf[r];
                gr += v.x * b.x + v.y * b.y;
                gi += v.x * b.y - v.y * b.x;
            }
            gr = wave_sum(gr);
            gi = wave_sum(gi);
            if (lane == 0) Gs[wave][j] = double2{gr, gi};
        }
    }
    // the panel column back into A (R above the diagonal, v below) and explicit V
    if (own) {
        double2* vcol = Vb + job.v_off + p0 + (int64_t)(p0 + wave) * m;
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int r = lane + 64 * i;
            if (r < mp) col[r] = x[i];
            x[i] = r < wave ? double2{0.0, 0.0} : (r == wave ? double2{1.0, 0.0} : x[i]);
            if (r < mp) vcol[r] = x[i];
        }
    }
    __syncthreads();
    // T = U^-1 with U = diag(1 / tau) + the strict upper triangle of G (G[a][c] = v_a^H v_c; the
    // recurrence T[0:c, c] = -tau_c T[0:c, 0:c] G[0:c, c] says T U = I).  Column c on thread c by
    // back-substitution of U T[:, c] = e_c: T[c][c] = tau_c, T[i][c] = -tau_i sum_{a=i+1..c}
    // G[i][a] T[a][c]; the columns are independent, so no barrier separates them.
    if ((int)threadIdx.x < jb) {
        const int c = threadIdx.x;
        Ts[c][c] = s_tau[c];
        for (int i = c - 1; i >= 0; --i) {
            double sr = 0.0, si = 0.0;
            for (int a = i + 1; a <= c; ++a) {
                const double2 g = Gs[i][a], t = Ts[a][c];
                sr += g.x * t.x - g.y * t.y;
                si += g.x * t.y + g.y * t.x;
            }
            const double2 ti = s_tau[i];
            Ts[i][c] = double2{-(ti.x * sr - ti.y * si), -(ti.x * si + ti.y * sr)};
        }
        double2* tout = Tp + (int64_t)blockIdx.x * kQrTSlot;
        for (int i = 0; i < jb; ++i) tout[i * kQrPanel + c] = i <= c ? Ts[i][c] : double2{0.0, 0.0};
    }
}

// ---- fused block-reflector application: C <- C - V op(T) (V^H C) ----------------------------
//
// One workgroup per 16-column slab of C (mp rows).  W0 = V^H C_slab (jb x 16, jb <= 16) is
// accumulated by v_mfma_f64_16x16x4f64 with the row range split over the four waves and the
// partials summed through LDS; W = op(T) W0 (16 x 16 on every lane's own entries); then
// C_slab -= V W, again by MFMA, 16 rows at a time.  The trailing update of the blocked QR applies
// (I - V T V^H)^H (op = ^H, `trans` 1); the backward formation of Q applies I - V T V^H (op = I).
// MFMA f64 16x16x4 operands: lane l of "a" holds X[l % 16][l / 16], lane l of "b" holds
// Y[l / 16][l % 16]; register r of lane l of D = X Y holds D[l / 16 + 4 r][l % 16].

typedef double dvec4 __attribute__((ext_vector_type(4)));

struct ReflTask {
    int64_t c_off, p_off, t_off;   // C slab origin (ldc); V origin (leading dimension ldp); T slot
    int32_t mp, ncols, jb, ldc, ldp, trans;
};

constexpr int kReflThreads = 512;
constexpr int kReflWaves = kReflThreads / 64;   // a multiple of 4: groups of four waves share 64 rows

__global__ __launch_bounds__(kReflThreads) void apply_reflector_kernel(
    const ReflTask* __restrict__ tasks, double2* __restrict__ C, const double2* __restrict__ Vb,
    const double2* __restrict__ Tb) {
    __shared__ double2 Wp[kReflWaves][16][16];
    __shared__ double2 Ts[16][17];
    const ReflTask tk = tasks[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int mp = tk.mp, nc = tk.ncols, jb = tk.jb;
    double2* c = C + tk.c_off;
    const double2* p = Vb + tk.p_off;
    const double2* r = p;
    const int64_t ldc = tk.ldc, ldp = tk.ldp;
    static_assert(kReflThreads >= 256 && kReflWaves % 4 == 0 && kQrPanel == 16, "one T entry per thread");
    const bool entry = threadIdx.x < 256;          // the threads that own one 16 x 16 entry
    const int ea = (threadIdx.x >> 4) & 15, eb = threadIdx.x & 15;
    if (entry) Ts[ea][eb] = (ea < jb && eb < jb) ? Tb[tk.t_off + ea * kQrPanel + eb] : double2{0.0, 0.0};
    // W0 = V^H C: X = V^H (a: conj(V[k][lr])), Y = C (b: C[k][lr])
    dvec4 wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};
    // four k-steps per iteration, loads issued before the products so their latencies overlap;
    // each group of four waves covers 64 consecutive rows
    for (int k0 = 64 * (wave >> 2) + 4 * (wave & 3); k0 < mp; k0 += 16 * kReflWaves) {
        double2 xv[4], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + 16 * u + lk;
            xv[u] = double2{0.0, 0.0};
            yv[u] = double2{0.0, 0.0};
            if (k < mp) {
                if (lr < jb) xv[u] = r[k + lr * ldp];
                if (lr < nc) yv[u] = c[k + lr * ldc];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double2 x = xv[u], y = yv[u];
            // X = conj(R)^T: Xr = x.x, Xi = -x.y
            wr = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, wr, 0, 0, 0);
            wr = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.y, wr, 0, 0, 0);   // -Xi Yi = +x.y y.y
            wi = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, wi, 0, 0, 0);
            wi = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, y.x, wi, 0, 0, 0);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Wp[wave][lk + 4 * q][lr] = double2{wr[q], wi[q]};
    __syncthreads();
    // W0 = the waves' partials summed in a fixed order, then W = op(T) W0 (op(T)[i][a] =
    // conj(T[a][i]) for the trailing update, T[i][a] for the Q formation), one entry per thread
    __shared__ double2 W0s[16][17];
    if (entry) {
        double2 s{0.0, 0.0};
#pragma unroll
        for (int w = 0; w < kReflWaves; ++w) {
            s.x += Wp[w][ea][eb].x;
            s.y += Wp[w][ea][eb].y;
        }
        W0s[ea][eb] = s;
    }
    __syncthreads();
    if (entry) {
        double sr = 0.0, si = 0.0;
        for (int a = 0; a < 16; ++a) {
            const double2 t = tk.trans ? double2{Ts[a][ea].x, -Ts[a][ea].y} : Ts[ea][a];
            const double2 v = W0s[a][eb];
            sr += t.x * v.x - t.y * v.y;
            si += t.x * v.y + t.y * v.x;
        }
        Wp[0][ea][eb] = double2{sr, si};
    }
    __syncthreads();
    // every wave needs W as the b operand of V W: lane l, chunk kk: W[4 kk + lk][lr]
    double2 wb[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) wb[kk] = Wp[0][4 * kk + lk][lr];
    // C -= V W, 16 rows per tile: X = V (a: V[row0 + lr][kk + lk]), Y = W
    for (int row0 = 16 * wave; row0 < mp; row0 += 16 * kReflWaves) {
        dvec4 dr = {0.0, 0.0, 0.0, 0.0}, di = {0.0, 0.0, 0.0, 0.0};
        double2 xv[4], cv[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int col = 4 * kk + lk, row = row0 + lr;
            xv[kk] = (row < mp && col < jb) ? p[row + col * ldp] : double2{0.0, 0.0};
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = row0 + lk + 4 * q;
            cv[q] = (row < mp && lr < nc) ? c[row + lr * ldc] : double2{0.0, 0.0};
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const double2 x = xv[kk];
            const double2 y = wb[kk];
            dr = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, dr, 0, 0, 0);
            dr = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, y.y, dr, 0, 0, 0);
            di = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, di, 0, 0, 0);
            di = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.x, di, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = row0 + lk + 4 * q;
            if (row < mp && lr < nc)
                c[row + lr * ldc] = double2{cv[q].x - dr[q], cv[q].y - di[q]};
        }
    }
}

// Q = [I_k; 0] (m x k, ldq) per matrix.
__global__ void qr_qinit_kernel(const MatJob* __restrict__ jobs, double2* __restrict__ Q) {
    const MatJob job = jobs[blockIdx.x];
    const int m = job.m, k = min(job.m, job.n);
    double2* q = Q + job.q_off;
    const int64_t total = (int64_t)m * k;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.y * blockDim.x) {
        const int r = (int)(e % m), c = (int)(e / m);
        q[r + c * (int64_t)job.ldq] = double2{r == c ? 1.0 : 0.0, 0.0};
    }
}

// R = upper triangle of the factored A (k x n); rows with a negative diagonal are negated together
// with the matching columns of Q (the unique gauge).
__global__ __launch_bounds__(1024) void qr_finish_kernel(const MatJob* __restrict__ jobs,
                                                         const double2* __restrict__ A,
                                                         double2* __restrict__ Q,
                                                         double2* __restrict__ R) {
    const MatJob job = jobs[blockIdx.x];
    const int m = job.m, n = job.n, k = min(m, n);
    const double2* a = A + job.a_off;
    double2* q = Q + job.q_off;
    double2* rr = R + job.r_off;
    const int64_t lda = job.lda, ldq = job.ldq, ldr = job.ldr;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x / 64;
    // one wave per row of R (column of Q); the rows are spread over gridDim.y workgroups
    for (int j = blockIdx.y * nw + wave; j < k; j += gridDim.y * nw) {
        const bool neg = a[j + (int64_t)j * lda].x < 0.0;
        for (int c = lane; c < n; c += 64) {
            double2 v = c >= j ? a[j + (int64_t)c * lda] : double2{0.0, 0.0};
            if (neg) v = double2{-v.x, -v.y};
            rr[j + (int64_t)c * ldr] = v;
        }
        if (neg)
            for (int i = lane; i < m; i += 64) {
                const double2 v = q[i + (int64_t)j * ldq];
                q[i + (int64_t)j * ldq] = double2{-v.x, -v.y};
            }
    }
}

// ---- batched SVD: one-sided Jacobi ---------------------------------------------------------

constexpr int kSvdThreads = 1024;
// batches with a block at least this wide are QR-preconditioned
constexpr int kPrecondMinCols = 24;

__device__ void svd_finish(const SvdJob& job, const double2* __restrict__ g,
                           const double2* __restrict__ v, double2* __restrict__ U,
                           double* __restrict__ S, double2* __restrict__ Vh, double* sig, int* perm);

// G (m x n, m >= n, ldg) is the working copy (overwritten); V (n x n, ldv) accumulates the
// rotations.  Outputs: U (m x n, ldu), S (n), Vh (n x n, ldvh), sorted by descending singular
// value.  phase_on_v selects which factor carries the phase convention (first entry with
// |x_i| >= 0.5 max|x| real positive): the columns of U (0) or the columns of V (1).
__global__ __launch_bounds__(kSvdThreads) void svd_jacobi_kernel(const SvdJob* __restrict__ jobs,
                                                                   double2* __restrict__ G,
                                                                   double2* __restrict__ V,
                                                                   double2* __restrict__ U,
                                                                   double* __restrict__ S,
                                                                   double2* __restrict__ Vh,
                                                                   int max_sweeps,
                                                                   int* __restrict__ sweeps_out) {
    extern __shared__ double smem[];
    const SvdJob job = jobs[blockIdx.x];
    const int m = job.m, n = job.n;
    double2* g = G + job.g_off;
    double2* v = V + job.v_off;
    const int64_t ldg = job.ldg, ldv = job.ldv;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int nw = kSvdThreads / 64;
    double* sig = smem;                                    // n
    int* perm = reinterpret_cast<int*>(smem + n);          // n
    __shared__ int s_rotated;

    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kSvdThreads) {
        const int r = (int)(e % n), c = (int)(e / n);
        v[r + c * ldv] = double2{r == c ? 1.0 : 0.0, 0.0};
    }
    const double tol = 2.220446049250313e-16 * sqrt((double)m);
    const int N = n + (n & 1);
    __syncthreads();
    int used = 0;
    for (int sweep = 0; sweep < max_sweeps && n > 1; ++sweep) {
        ++used;
        if (threadIdx.x == 0) s_rotated = 0;
        __syncthreads();
        for (int round = 0; round < N - 1; ++round) {
            for (int pi = wave; pi < N / 2; pi += nw) {
                int p, qq;
                if (pi == 0) {
                    p = round % (N - 1);
                    qq = N - 1;
                } else {
                    p = (round + pi) % (N - 1);
                    qq = (round - pi + (N - 1)) % (N - 1);
                }
                if (p >= n || qq >= n) continue;
                if (p > qq) {
                    const int t = p;
                    p = qq;
                    qq = t;
                }
                double2* gp = g + (int64_t)p * ldg;
                double2* gq = g + (int64_t)qq * ldg;
                double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
                for (int i = lane; i < m; i += 64) {
                    const double2 a = gp[i], b = gq[i];
                    al += a.x * a.x + a.y * a.y;
                    be += b.x * b.x + b.y * b.y;
                    gr += a.x * b.x + a.y * b.y;
                    gi += a.x * b.y - a.y * b.x;
                }
                al = wave_sum(al);
                be = wave_sum(be);
                gr = wave_sum(gr);
                gi = wave_sum(gi);
                const double ag = sqrt(gr * gr + gi * gi);
                if (!(ag > tol * sqrt(al * be)) || ag == 0.0) continue;
                const double zeta = (be - al) / (2.0 * ag);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = c * t;
                const double er = gr / ag, ei = gi / ag;  // e = gamma / |gamma|
                // x_p' = c x_p - s conj(e) x_q ; x_q' = s e x_p + c x_q
                for (int i = lane; i < m; i += 64) {
                    const double2 a = gp[i], b = gq[i];
                    gp[i] = double2{c * a.x - s * (er * b.x + ei * b.y),
                                    c * a.y - s * (er * b.y - ei * b.x)};
                    gq[i] = double2{s * (er * a.x - ei * a.y) + c * b.x,
                                    s * (er * a.y + ei * a.x) + c * b.y};
                }
                double2* vp = v + (int64_t)p * ldv;
                double2* vq = v + (int64_t)qq * ldv;
                for (int i = lane; i < n; i += 64) {
                    const double2 a = vp[i], b = vq[i];
                    vp[i] = double2{c * a.x - s * (er * b.x + ei * b.y),
                                    c * a.y - s * (er * b.y - ei * b.x)};
                    vq[i] = double2{s * (er * a.x - ei * a.y) + c * b.x,
                                    s * (er * a.y + ei * a.x) + c * b.y};
                }
                if (lane == 0) s_rotated = 1;
            }
            __syncthreads();
        }
        if (!s_rotated) break;
        __syncthreads();
    }
    if (sweeps_out && threadIdx.x == 0) sweeps_out[blockIdx.x] = used;
    svd_finish(job, g, v, U, S, Vh, sig, perm);
}

// End of the SVD kernel: singular values, descending order (ties by index), phase
// convention and the U / S / Vh outputs.  Called by every thread of the workgroup.
__device__ void svd_finish(const SvdJob& job, const double2* __restrict__ g,
                           const double2* __restrict__ v, double2* __restrict__ U,
                           double* __restrict__ S, double2* __restrict__ Vh, double* sig, int* perm) {
    const int m = job.m, n = job.n;
    const int64_t ldg = job.ldg, ldv = job.ldv;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x / 64;
    __syncthreads();
    for (int j = wave; j < n; j += nw) {
        const double2* gj = g + (int64_t)j * ldg;
        double s2 = 0.0;
        for (int i = lane; i < m; i += 64) s2 += gj[i].x * gj[i].x + gj[i].y * gj[i].y;
        s2 = wave_sum(s2);
        if (lane == 0) sig[j] = sqrt(s2);
    }
    __syncthreads();
    // rank by a strict total order (descending value, ties by index, a NaN below everything) so
    // that perm is a permutation whatever the input held: a non-finite block then yields NaN
    // outputs for the caller to refuse, never an index outside the block
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        int rank = 0;
        const double sj = sig[j] == sig[j] ? sig[j] : -1.0;
        for (int i = 0; i < n; ++i) {
            const double si = sig[i] == sig[i] ? sig[i] : -1.0;
            rank += (si > sj) || (si == sj && i < j);
        }
        perm[rank] = j;
    }
    __syncthreads();
    double2* u = U + job.u_off;
    double2* vh = Vh + job.vh_off;
    double* sv = S + job.s_off;
    const int64_t ldu = job.ldu, ldvh = job.ldvh;
    for (int r = wave; r < n; r += nw) {
        const int j = perm[r];
        const double sj = sig[j];
        const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
        const double2* gj = g + (int64_t)j * ldg;
        const double2* vj = v + (int64_t)j * ldv;
        // phase of the pair from the chosen factor's column
        const double2* ref = job.phase_on_v ? vj : gj;
        const int len = job.phase_on_v ? n : m;
        double mx = 0.0;
        for (int i = lane; i < len; i += 64) mx = fmax(mx, sqrt(ref[i].x * ref[i].x + ref[i].y * ref[i].y));
        mx = wave_max(mx);
        int first = len;
        for (int base = 0; base < len && first == len; base += 64) {
            const int i = base + lane;
            bool hit = false;
            if (i < len && mx > 0.0) {
                const double ai = sqrt(ref[i].x * ref[i].x + ref[i].y * ref[i].y);
                hit = ai >= 0.5 * mx;
            }
            const unsigned long long bal = __ballot(hit);
            if (bal) first = base + __ffsll((long long)bal) - 1;
        }
        double phr = 1.0, phi = 0.0;  // ph = ref[first] / |ref[first]|
        if (first < len && (sj > 0.0 || job.phase_on_v)) {
            const double2 x = ref[first];
            const double ax = sqrt(x.x * x.x + x.y * x.y);
            if (ax > 0.0) {
                phr = x.x / ax;
                phi = x.y / ax;
            }
        }
        // U col *= conj(ph) (and / sigma), Vh row = conj(V col) * ph
        for (int i = lane; i < m; i += 64) {
            const double2 x = gj[i];
            u[i + (int64_t)r * ldu] = double2{(x.x * phr + x.y * phi) * inv,
                                              (x.y * phr - x.x * phi) * inv};
        }
        for (int i = lane; i < n; i += 64) {
            const double2 x = vj[i];  // conj(x) * ph
            vh[r + (int64_t)i * ldvh] = double2{x.x * phr + x.y * phi, x.x * phi - x.y * phr};
        }
        if (lane == 0) sv[r] = sj;
    }
}

// Phase convention on finished factors (QR-preconditioned path): for column r of U (m rows) and
// row r of Vh (n columns), with ref = the U column (or the conjugated Vh row when phase_on_v),
// make the first entry with |x_i| >= 0.5 max|x| real positive.  One wave per column.
__global__ void svd_phase_kernel(const SvdJob* __restrict__ jobs, double2* __restrict__ U,
                                 double2* __restrict__ Vh) {
    const SvdJob job = jobs[blockIdx.x];
    const int m = job.m, n = job.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x / 64;
    double2* u = U + job.u_off;
    double2* vh = Vh + job.vh_off;
    const int64_t ldu = job.ldu, ldvh = job.ldvh;
    // one wave per column; the columns of a matrix are spread over gridDim.y workgroups
    for (int r = blockIdx.y * nw + wave; r < n; r += gridDim.y * nw) {
        const bool onv = job.phase_on_v != 0;
        const int len = onv ? n : m;
        double mx = 0.0;
        for (int i = lane; i < len; i += 64) {
            const double2 x = onv ? vh[r + (int64_t)i * ldvh] : u[i + (int64_t)r * ldu];
            mx = fmax(mx, sqrt(x.x * x.x + x.y * x.y));
        }
        mx = wave_max(mx);
        int first = len;
        for (int base = 0; base < len && first == len; base += 64) {
            const int i = base + lane;
            bool hit = false;
            if (i < len && mx > 0.0) {
                const double2 x = onv ? vh[r + (int64_t)i * ldvh] : u[i + (int64_t)r * ldu];
                hit = sqrt(x.x * x.x + x.y * x.y) >= 0.5 * mx;
            }
            const unsigned long long bal = __ballot(hit);
            if (bal) first = base + __ffsll((long long)bal) - 1;
        }
        if (first >= len) continue;
        double2 x = onv ? vh[r + (int64_t)first * ldvh] : u[first + (int64_t)r * ldu];
        if (onv) x.y = -x.y;   // the V column is the conjugated Vh row
        const double ax = sqrt(x.x * x.x + x.y * x.y);
        if (!(ax > 0.0)) continue;
        const double phr = x.x / ax, phi = x.y / ax;
        for (int i = lane; i < m; i += 64) {   // U col *= conj(ph)
            const double2 y = u[i + (int64_t)r * ldu];
            u[i + (int64_t)r * ldu] = double2{y.x * phr + y.y * phi, y.y * phr - y.x * phi};
        }
        for (int i = lane; i < n; i += 64) {   // Vh row *= ph
            const double2 y = vh[r + (int64_t)i * ldvh];
            vh[r + (int64_t)i * ldvh] = double2{y.x * phr - y.y * phi, y.x * phi + y.y * phr};
        }
    }
}

// ---- multi-workgroup block one-sided Jacobi ------------------------------------------------
//
// The n columns of X (n x n) are cut into blocks of w columns.  A sweep is the Brent-Luk
// round-robin over the blocks: in every round the blocks pair up and each workgroup takes one
// pair, stages its c <= 2w columns in LDS, runs one one-sided Jacobi sweep over all column pairs
// of the stage (same rotation and threshold as the one-workgroup kernel), writes the columns
// back and applies the accumulated c x c rotation Z to the same columns of V.  One launch per
// round; a per-matrix flag records whether a sweep rotated, and a matrix whose previous sweep did
// not rotate is skipped.

constexpr int kBjThreads = 1024;
constexpr int kBjMaxCols = 32;            // columns staged per workgroup
constexpr size_t kBjStageBytes = 131072;  // LDS for the staged columns

struct BjMat {
    int64_t x_off, v_off;
    int32_t n, w, nb, N;      // rows = columns = n; block width; blocks; blocks padded to even
};

__device__ __forceinline__ void round_robin(int N, int round, int pi, int& p, int& q) {
    if (pi == 0) {
        p = round % (N - 1);
        q = N - 1;
    } else {
        p = (round + pi) % (N - 1);
        q = (round - pi + (N - 1)) % (N - 1);
    }
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
}

// R = double (the FP64 Jacobi) or float (the FP32 preconditioning sweeps of the mixed variant):
// columns, rotations and the threshold eps in R.
template <typename R>
__global__ __launch_bounds__(kBjThreads) void bj_round_kernel(const BjMat* __restrict__ mats,
                                                              const int2* __restrict__ tasks,
                                                              typename Cplx<R>::type* __restrict__ X,
                                                              typename Cplx<R>::type* __restrict__ V,
                                                              int round,
                                                              const int* __restrict__ flags_prev,
                                                              int* __restrict__ flags_cur,
                                                              R eps, int inner,
                                                              const R* __restrict__ fro2,
                                                              R floor_rel2) {
    typedef typename Cplx<R>::type C2;
    extern __shared__ __align__(16) unsigned char bj_lds[];
    C2* xs = reinterpret_cast<C2*>(bj_lds);   // n * c staged columns, then c * c Z
    __shared__ R red[kBjThreads / 64][4];
    __shared__ int s_rot;
    const int2 tk = tasks[blockIdx.x];
    const BjMat mt = mats[tk.x];
    if (round >= mt.N - 1) return;
    if (flags_prev[tk.x] == 0) return;        // converged in the previous sweep
    int bp, bq;
    round_robin(mt.N, round, tk.y, bp, bq);
    const int n = mt.n, w = mt.w;
    const int cp = bp < mt.nb ? min(w, n - bp * w) : 0;
    const int cq = bq < mt.nb ? min(w, n - bq * w) : 0;
    const int c = cp + cq;
    if (c < 2) return;
    // round 0 of a sweep rotates every pair of the staged columns (so every pair of the matrix,
    // intra-block ones included, is visited once per sweep); later rounds only the cross pairs
    const bool cross_only = round > 0 && cp > 0 && cq > 0;
    C2* x = X + mt.x_off;
    C2* v = V + mt.v_off;
    C2* Z = xs + (int64_t)n * c;
    auto gcol = [&](int j) { return j < cp ? bp * w + j : bq * w + (j - cp); };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int nw = kBjThreads / 64;
    for (int j = wave; j < c; j += nw) {
        const C2* src = x + (int64_t)gcol(j) * n;
        C2* dst = xs + (int64_t)j * n;
        for (int i = lane; i < n; i += 64) dst[i] = src[i];
    }
    for (int e = tid; e < c * c; e += kBjThreads) Z[e] = C2{(e % c) == (e / c) ? R(1) : R(0), R(0)};
    if (tid == 0) s_rot = 0;
    __syncthreads();
    const R tol2 = eps * eps * (R)n;
    // A pair whose two columns together hold less than floor_rel2 of the matrix's squared
    // Frobenius norm is left alone: rotations among such columns only reshape directions whose
    // singular values lie below floor_rel2^(1/2) ||X||_F, which no selection keeps, and a column
    // can only grow by rotating against a larger partner, which is never skipped.
    const R floor2 = floor_rel2 * fro2[tk.x];
    // pairs per inner round: round-robin over all c columns, or the cyclic shifts of the smaller
    // block against the larger
    const int big = cp >= cq ? cp : cq, small = cp >= cq ? cq : cp;
    const int big0 = cp >= cq ? 0 : cp, small0 = cp >= cq ? cp : 0;
    const int C = c + (c & 1);
    const int np = cross_only ? small : C / 2;
    const int nrounds = cross_only ? big : C - 1;
    const int wpp = max(1, nw / np);          // waves per pair
    const int grp = wave / wpp, sub = wave % wpp;
    const int gthreads = wpp * 64, gtid = sub * 64 + lane;
    for (int pass = 0; pass < inner; ++pass) {
        for (int r0 = 0; r0 < nrounds; ++r0) {
            // pair slots beyond 16 waves are taken in turns
            for (int g0 = 0; g0 < np; g0 += nw / wpp) {
                const int g = g0 + grp;
                int a = 0, b = 0;
                bool valid = grp < nw / wpp && g < np;
                if (valid) {
                    if (cross_only) {
                        const int i1 = small0 + g, i2 = big0 + (g + r0) % big;
                        a = min(i1, i2);
                        b = max(i1, i2);
                    } else {
                        round_robin(C, r0, g, a, b);
                        valid = b < c;
                    }
                }
                R al = 0, be = 0, gr = 0, gi = 0;
                if (valid) {
                    const C2* xa = xs + (int64_t)a * n;
                    const C2* xb = xs + (int64_t)b * n;
                    for (int i = gtid; i < n; i += gthreads) {
                        const C2 p = xa[i], q = xb[i];
                        al += p.x * p.x + p.y * p.y;
                        be += q.x * q.x + q.y * q.y;
                        gr += p.x * q.x + p.y * q.y;
                        gi += p.x * q.y - p.y * q.x;
                    }
                }
                al = wave_sum(al);
                be = wave_sum(be);
                gr = wave_sum(gr);
                gi = wave_sum(gi);
                if (wpp > 1) {
                    // several waves per pair: combine their partial sums through LDS
                    if (lane == 0) {
                        red[wave][0] = al;
                        red[wave][1] = be;
                        red[wave][2] = gr;
                        red[wave][3] = gi;
                    }
                    __syncthreads();
                    if (valid) {
                        al = be = gr = gi = 0;
                        for (int k = 0; k < wpp; ++k) {
                            al += red[grp * wpp + k][0];
                            be += red[grp * wpp + k][1];
                            gr += red[grp * wpp + k][2];
                            gi += red[grp * wpp + k][3];
                        }
                    }
                }
                const R ag2 = gr * gr + gi * gi;
                // rotate when |gamma| > tol sqrt(al be), tol = eps sqrt(n) (squared: no roots)
                if (valid && al + be >= floor2 && ag2 > tol2 * (al * be) && ag2 != R(0)) {
                    const R ag = sqrt(ag2);
                    const R inv = R(1) / ag;
                    const R zeta = R(0.5) * (be - al) * inv;
                    const R t = (zeta >= R(0) ? R(1) : R(-1)) / (fabs(zeta) + sqrt(R(1) + zeta * zeta));
                    const R cs = R(1) / sqrt(R(1) + t * t);
                    const R sn = cs * t;
                    const R er = gr * inv, ei = gi * inv;
                    // x_a' = c x_a - s conj(e) x_b ; x_b' = s e x_a + c x_b
                    C2* xa = xs + (int64_t)a * n;
                    C2* xb = xs + (int64_t)b * n;
                    for (int i = gtid; i < n; i += gthreads) {
                        const C2 p = xa[i], q = xb[i];
                        xa[i] = C2{cs * p.x - sn * (er * q.x + ei * q.y),
                                        cs * p.y - sn * (er * q.y - ei * q.x)};
                        xb[i] = C2{sn * (er * p.x - ei * p.y) + cs * q.x,
                                        sn * (er * p.y + ei * p.x) + cs * q.y};
                    }
                    C2* za = Z + (int64_t)a * c;
                    C2* zb = Z + (int64_t)b * c;
                    for (int i = gtid; i < c; i += gthreads) {
                        const C2 p = za[i], q = zb[i];
                        za[i] = C2{cs * p.x - sn * (er * q.x + ei * q.y),
                                        cs * p.y - sn * (er * q.y - ei * q.x)};
                        zb[i] = C2{sn * (er * p.x - ei * p.y) + cs * q.x,
                                        sn * (er * p.y + ei * p.x) + cs * q.y};
                    }
                    if (gtid == 0) s_rot = 1;
                }
                __syncthreads();
            }
        }
    }
    if (!s_rot) return;                       // uniform: read after the last barrier
    for (int j = wave; j < c; j += nw) {
        C2* dst = x + (int64_t)gcol(j) * n;
        const C2* src = xs + (int64_t)j * n;
        for (int i = lane; i < n; i += 64) dst[i] = src[i];
    }
    __syncthreads();
    for (int j = wave; j < c; j += nw) {
        const C2* src = v + (int64_t)gcol(j) * n;
        C2* dst = xs + (int64_t)j * n;
        for (int i = lane; i < n; i += 64) dst[i] = src[i];
    }
    __syncthreads();
    // V(:, cols) <- V(:, cols) Z
    for (int j = wave; j < c; j += nw) {
        C2* dst = v + (int64_t)gcol(j) * n;
        const C2* zj = Z + (int64_t)j * c;
        for (int i = lane; i < n; i += 64) {
            R sr = 0, si = 0;
            for (int k = 0; k < c; ++k) {
                const C2 p = xs[i + (int64_t)k * n], z = zj[k];
                sr += p.x * z.x - p.y * z.y;
                si += p.x * z.y + p.y * z.x;
            }
            dst[i] = C2{sr, si};
        }
    }
    if (tid == 0) flags_cur[tk.x] = 1;
}

// ---- Gram block Jacobi: the block pair's work split over row chunks --------------------------
//
// Block one-sided Jacobi where a round's pair (p, q) is not staged whole in one workgroup: the
// rows are cut into chunks of kGbjRows, each handled by its own workgroup, so a round runs
// (pairs x chunks) workgroups and the column block width no longer depends on the row count.
// Per round two launches:
//   gbj_gram_kernel   -- each chunk's partial Gram matrix G_k = X_k^H X_k of the pair's c <= 32
//                        columns, written to a per-(pair, chunk) slot;
//   gbj_rotate_kernel -- every chunk's workgroup sums the partials in chunk order (the same sum
//                        in every workgroup), runs the pair's inner sweeps of two-sided rotations
//                        on G itself (the one-sided rotations of the columns, applied to their
//                        Gram matrix: G <- J^H G J, Z <- Z J), and applies the accumulated Z to its
//                        row chunk of X and of V.
// The rotation test and the angle formula are those of the staged kernel (|g_ab|^2 > eps^2 n g_aa g_bb);
// the Gram matrix is recomputed from the columns every round, so its updates within the round
// never accumulate past one round, and convergence is judged on fresh inner products.
constexpr int kGbjCols = 32;   // widest block pair (2w) of the default kernels
constexpr int kGbjColsWide = 64;   // of the wide ones (1024 threads)
constexpr int kGbjRotThreads = 512;   // rotate kernel: every pair of an inner round at once

// 1 / sqrt(x) and 1 / x for x in [1e-280, 1e280]: the hardware estimates (about 23 correct bits)
// refined by two Newton steps, a few ulps in all; the IEEE sequences' range scaling and special
// cases are not needed in that range
__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double r = fma(-h * y, y, 0.5);
    y = fma(y, r, y);
    r = fma(-h * y, y, 0.5);
    return fma(y, r, y);
}
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// The Jacobi rotation of the Hermitian 2 x 2 [[al, g], [conj g, be]], ag2 = |g|^2 > 0:
// zeta = (be - al) / (2 |g|), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), cs = 1 / sqrt(1 + t^2),
// sn = cs t, phase e = g / |g|.  The latency of this chain sets the pace of the inner sweeps, so
// it runs on the refined estimates; |g|^2 outside their range takes the IEEE operations.
__device__ __forceinline__ void jacobi_rotation(double al, double be, double gr, double gi, double ag2,
                                                double& cs, double& sn, double& er, double& ei) {
    double inv, t;
    if (ag2 >= 1e-280 && ag2 <= 1e280) {
        inv = rsq_nr(ag2);
        const double zeta = 0.5 * (be - al) * inv;
        const double az = fabs(zeta);
        if (az < 1e100) {
            const double z1 = fma(zeta, zeta, 1.0);
            t = rcp_nr(az + z1 * rsq_nr(z1));
        } else {
            t = 0.5 / az;             // sqrt(1 + zeta^2) = |zeta| in working precision
        }
        if (zeta < 0.0) t = -t;
        cs = rsq_nr(fma(t, t, 1.0));
    } else {
        inv = 1.0 / sqrt(ag2);
        const double zeta = 0.5 * (be - al) * inv;
        t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        cs = 1.0 / sqrt(1.0 + t * t);
    }
    sn = cs * t;
    er = gr * inv;
    ei = gi * inv;
}
constexpr int kGbjRows = 32;   // rows per chunk (64 measured slower: twice the update work per workgroup)
// LDS of one instance: the gram kernel stages CC columns of a chunk, the rotate kernel holds G
// and Z (CC x (CC + 1) each) and stages the X and V rows in G's storage (and, for CC < 64, in a
// third such array)
constexpr size_t gbj_gram_lds(int cc) { return sizeof(double2) * (size_t)cc * (kGbjRows + 1); }
constexpr size_t gbj_rot_lds(int cc) {
    return (cc >= 2 * kGbjRows ? 2 : 3) * sizeof(double2) * (size_t)cc * (cc + 1);
}

// CC columns per block pair on NT = (CC / 2)^2 threads
template <int CC, int NT>
__global__ __launch_bounds__(NT) void gbj_gram_kernel(const BjMat* __restrict__ mats,
                                                      const int2* __restrict__ tasks,
                                                      const double2* __restrict__ X,
                                                      double2* __restrict__ gbuf, int round,
                                                      const int* __restrict__ flags_prev,
                                                      int maxch) {
    static_assert((CC / 2) * (CC / 2) == NT, "one 2 x 2 block of G per thread");
    extern __shared__ __align__(16) unsigned char gbj_gram_lds_raw[];
    double2 (*xs)[kGbjRows + 1] = reinterpret_cast<double2 (*)[kGbjRows + 1]>(gbj_gram_lds_raw);
    const int2 tk = tasks[blockIdx.x];
    const BjMat mt = mats[tk.x];
    if (round >= mt.N - 1 || flags_prev[tk.x] == 0) return;
    const int r0 = blockIdx.y * kGbjRows;
    if (r0 >= mt.n) return;
    int bp, bq;
    round_robin(mt.N, round, tk.y, bp, bq);
    const int n = mt.n, w = mt.w;
    const int cp = bp < mt.nb ? min(w, n - bp * w) : 0;
    const int cq = bq < mt.nb ? min(w, n - bq * w) : 0;
    const int c = cp + cq;
    if (c < 2) return;
    const int rows = min(kGbjRows, n - r0);
    const double2* x = X + mt.x_off;
    const int tid = threadIdx.x;
    // every load in flight before the first store to LDS
    constexpr int kS = (CC * kGbjRows + NT - 1) / NT;
    double2 val[kS];
#pragma unroll
    for (int t = 0; t < kS; ++t) {
        const int e = tid + t * NT;
        const int j = e / kGbjRows, i = e % kGbjRows;
        val[t] = double2{0.0, 0.0};
        if (e < CC * kGbjRows && j < c && i < rows) {
            const int col = j < cp ? bp * w + j : bq * w + (j - cp);
            val[t] = x[(int64_t)col * n + r0 + i];
        }
    }
#pragma unroll
    for (int t = 0; t < kS; ++t) {
        const int e = tid + t * NT;
        if (e < CC * kGbjRows) xs[e / kGbjRows][e % kGbjRows] = val[t];
    }
    __syncthreads();
    // G = X_k^H X_k on FP64 MFMA 16x16x4, one 16 x 16 tile of G per wave, k running over the
    // chunk's rows (zero past `rows` and outside the pair's columns, as staged):
    // re = Xr^T Xr + Xi^T Xi, im = Xr^T Xi - Xi^T Xr.  Operand a lane l holds
    // X^H[ta + l % 16][k] = conj(X[k][ta + l % 16]), operand b holds X[k][tb + l % 16] with
    // k = k0 + l / 16; register r of lane l holds G[ta + l / 16 + 4 r][tb + l % 16].
    static_assert(CC % 16 == 0 && NT == 64 * (CC / 16) * (CC / 16) && kGbjRows % 4 == 0,
                  "one 16 x 16 tile of G per wave");
    typedef double dv4 __attribute__((ext_vector_type(4)));
    const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int ta = (wave / (CC / 16)) * 16, tb = (wave % (CC / 16)) * 16;
    dv4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < kGbjRows; k0 += 4) {
        const double2 a = xs[ta + li][k0 + lk], b = xs[tb + li][k0 + lk];
        accr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, accr, 0, 0, 0);
        accr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, accr, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, acci, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.x, acci, 0, 0, 0);
    }
    double2* g = gbuf + ((int64_t)blockIdx.x * maxch + blockIdx.y) * (CC * CC);
#pragma unroll
    for (int r = 0; r < 4; ++r) g[(ta + lk + 4 * r) * CC + tb + li] = double2{accr[r], acci[r]};
}

template <int CC, int NT>
__global__ __launch_bounds__(NT) void gbj_rotate_kernel(const BjMat* __restrict__ mats,
                                                        const int2* __restrict__ tasks,
                                                        double2* __restrict__ X,
                                                        double2* __restrict__ V,
                                                        const double2* __restrict__ gbuf,
                                                        int round,
                                                        const int* __restrict__ flags_prev,
                                                        int* __restrict__ flags_cur,
                                                        double eps, int inner, int maxch) {
    extern __shared__ __align__(16) unsigned char gbj_rot_lds_raw[];
    double2 (*G)[CC + 1] = reinterpret_cast<double2 (*)[CC + 1]>(gbj_rot_lds_raw);
    double2 (*Z)[CC + 1] = G + CC;
    __shared__ int s_rot;
    const int2 tk = tasks[blockIdx.x];
    const BjMat mt = mats[tk.x];
    if (round >= mt.N - 1 || flags_prev[tk.x] == 0) return;
    const int r0 = blockIdx.y * kGbjRows;
    if (r0 >= mt.n) return;
    int bp, bq;
    round_robin(mt.N, round, tk.y, bp, bq);
    const int n = mt.n, w = mt.w;
    const int cp = bp < mt.nb ? min(w, n - bp * w) : 0;
    const int cq = bq < mt.nb ? min(w, n - bq * w) : 0;
    const int c = cp + cq;
    if (c < 2) return;
    const int rows = min(kGbjRows, n - r0);
    const int nch = (n + kGbjRows - 1) / kGbjRows;
    const int tid = threadIdx.x;
    auto gcol = [&](int j) { return j < cp ? bp * w + j : bq * w + (j - cp); };
    const double2* gp = gbuf + (int64_t)blockIdx.x * maxch * (CC * CC);
    // G and Z over the whole CC x CC: zero outside the pair's c columns, where the update's
    // products below read Z.  The chunks' partials are summed in chunk order, several chunks'
    // loads in flight at a time (a loop of single loads would wait out one memory latency per
    // chunk).
    {
        constexpr int kE = (CC * CC + NT - 1) / NT;
        constexpr int kBatch = kE <= 2 ? 8 : 16 / kE;       // loads in flight per thread: 16
        double sr[kE], si[kE];
        int ea[kE], eb[kE];
        bool live[kE];
#pragma unroll
        for (int u = 0; u < kE; ++u) {
            const int e = tid + u * NT;
            ea[u] = e / CC;
            eb[u] = e % CC;
            live[u] = e < CC * CC && ea[u] < c && eb[u] < c;
            sr[u] = 0.0;
            si[u] = 0.0;
        }
        for (int k0 = 0; k0 < nch; k0 += kBatch) {
            double2 v[kBatch][kE];
#pragma unroll
            for (int kk = 0; kk < kBatch; ++kk)
#pragma unroll
                for (int u = 0; u < kE; ++u) {
                    v[kk][u] = double2{0.0, 0.0};
                    if (live[u] && k0 + kk < nch)
                        v[kk][u] = gp[(int64_t)(k0 + kk) * (CC * CC) + ea[u] * CC + eb[u]];
                }
#pragma unroll
            for (int kk = 0; kk < kBatch; ++kk)
#pragma unroll
                for (int u = 0; u < kE; ++u) {
                    sr[u] += v[kk][u].x;
                    si[u] += v[kk][u].y;
                }
        }
#pragma unroll
        for (int u = 0; u < kE; ++u) {
            if (tid + u * NT >= CC * CC) continue;
            const int a = ea[u], b = eb[u];
            G[a][b] = double2{sr[u], si[u]};
            Z[a][b] = double2{a == b && a < c ? 1.0 : 0.0, 0.0};
        }
    }
    // The update at the end rewrites this chunk's rows of the pair's columns of X and V; no other
    // workgroup writes them in this round, so they are loaded here, after the partial Grams (whose
    // loads the sum above has waited for), and their latency runs under the inner sweep, which
    // touches only LDS, instead of after it.  Zero outside the c columns and the chunk's rows.
    static_assert((CC * kGbjRows) % NT == 0, "whole chunk rows per thread");
    constexpr int kPre = CC * kGbjRows / NT;
    double2 pre[2][kPre];
#pragma unroll
    for (int mat = 0; mat < 2; ++mat) {
        const double2* m = mat == 0 ? X + mt.x_off : V + mt.v_off;
#pragma unroll
        for (int t = 0; t < kPre; ++t) {
            const int e = tid + t * NT;
            const int j = e / kGbjRows, i = e % kGbjRows;
            pre[mat][t] = double2{0.0, 0.0};
            if (j < c && i < rows) pre[mat][t] = m[(int64_t)gcol(j) * n + r0 + i];
        }
    }
    __shared__ int s_need;
    if (tid == 0) {
        s_rot = 0;
        s_need = 0;
    }
    __syncthreads();
    const double tol2 = eps * eps * (double)n;
    const bool cross_only = round > 0 && cp > 0 && cq > 0;
    // Every pair the inner sweep visits, tested on G as summed: if none passes the rotation
    // test, the first inner round rotates nothing, G stays as it is, and so do all later ones;
    // the sweep is skipped (the result is the one the sweep would give).
    for (int e = tid; e < c * c; e += NT) {
        const int a = e / c, b = e % c;
        if (a >= b) continue;
        if (cross_only && !(a < cp && b >= cp)) continue;
        const double gr = G[a][b].x, gi = G[a][b].y;
        const double ag2 = gr * gr + gi * gi;
        if (ag2 > tol2 * (G[a][a].x * G[b][b].x) && ag2 != 0.0) s_need = 1;
    }
    __syncthreads();
    const bool needed = s_need != 0;
    const int big = cp >= cq ? cp : cq, small = cp >= cq ? cq : cp;
    const int big0 = cp >= cq ? 0 : cp, small0 = cp >= cq ? cp : 0;
    const int C = c + (c & 1);
    const int np = cross_only ? small : C / 2;
    const int nrounds = needed ? (cross_only ? big : C - 1) : 0;
    // Thread t works on pair g = t / 32, entries k = t % 32 + 32 e (e < CC / 32) of the pair's two
    // columns (then of its two rows): every pair of an inner round at once.  Every thread of a
    // pair computes the pair's rotation itself from G as the round found it: the pair's 32
    // threads sit in one wavefront and read G[a][a], G[b][b], G[a][b] before any of them writes,
    // and no other pair writes columns a or b, so a round needs only the barrier between its
    // column and row updates and the one after them.
    constexpr int kTpp = 32;
    constexpr int kPairsAtOnce = NT / kTpp;
    static_assert(kPairsAtOnce == CC / 2 && CC % kTpp == 0, "one pair per 32 threads");
    constexpr int kSlots = 1;
    constexpr int kEnt = CC / kTpp;
    const int gl = tid / kTpp, kb = tid % kTpp;
    bool any = false;
    for (int pass = 0; pass < inner; ++pass) {
        for (int r = 0; r < nrounds; ++r) {
            double cs[kSlots], sn[kSlots], er[kSlots], ei[kSlots];
            int ra[kSlots], rb[kSlots];
#pragma unroll
            for (int sl = 0; sl < kSlots; ++sl) {
                const int g = gl + sl * kPairsAtOnce;
                ra[sl] = -1;
                if (g >= np) continue;
                int a, b;
                bool valid = true;
                if (cross_only) {
                    const int i1 = small0 + g, i2 = big0 + (g + r) % big;
                    a = min(i1, i2);
                    b = max(i1, i2);
                } else {
                    round_robin(C, r, g, a, b);
                    valid = b < c;
                }
                if (!valid) continue;
                const double al = G[a][a].x, be = G[b][b].x;
                const double gr = G[a][b].x, gi = G[a][b].y;
                const double ag2 = gr * gr + gi * gi;
                if (!(ag2 > tol2 * (al * be) && ag2 != 0.0)) continue;
                jacobi_rotation(al, be, gr, gi, ag2, cs[sl], sn[sl], er[sl], ei[sl]);
                ra[sl] = a;
                rb[sl] = b;
            }
            // the reads above complete in every lane before the writes below (one wavefront
            // per pair), and the compiler keeps them in order
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
            for (int sl = 0; sl < kSlots; ++sl)
#pragma unroll
            for (int en = 0; en < kEnt; ++en) {
                const int a = ra[sl], k = kb + en * kTpp;
                if (a < 0 || k >= c) continue;
                const int b = rb[sl];
                any = true;
                // columns: M[:, a] <- c M[:, a] - s conj(e) M[:, b];  M[:, b] <- s e M[:, a] + c M[:, b]
                double2 p = G[k][a], q = G[k][b];
                G[k][a] = double2{cs[sl] * p.x - sn[sl] * (er[sl] * q.x + ei[sl] * q.y),
                                  cs[sl] * p.y - sn[sl] * (er[sl] * q.y - ei[sl] * q.x)};
                G[k][b] = double2{sn[sl] * (er[sl] * p.x - ei[sl] * p.y) + cs[sl] * q.x,
                                  sn[sl] * (er[sl] * p.y + ei[sl] * p.x) + cs[sl] * q.y};
                p = Z[k][a];
                q = Z[k][b];
                Z[k][a] = double2{cs[sl] * p.x - sn[sl] * (er[sl] * q.x + ei[sl] * q.y),
                                  cs[sl] * p.y - sn[sl] * (er[sl] * q.y - ei[sl] * q.x)};
                Z[k][b] = double2{sn[sl] * (er[sl] * p.x - ei[sl] * p.y) + cs[sl] * q.x,
                                  sn[sl] * (er[sl] * p.y + ei[sl] * p.x) + cs[sl] * q.y};
            }
            __syncthreads();
            // rows: M[a, :] <- c M[a, :] - s e M[b, :];  M[b, :] <- s conj(e) M[a, :] + c M[b, :]
#pragma unroll
            for (int sl = 0; sl < kSlots; ++sl)
#pragma unroll
            for (int en = 0; en < kEnt; ++en) {
                const int a = ra[sl], k = kb + en * kTpp;
                if (a < 0 || k >= c) continue;
                const int b = rb[sl];
                const double2 p = G[a][k], q = G[b][k];
                G[a][k] = double2{cs[sl] * p.x - sn[sl] * (er[sl] * q.x - ei[sl] * q.y),
                                  cs[sl] * p.y - sn[sl] * (er[sl] * q.y + ei[sl] * q.x)};
                G[b][k] = double2{sn[sl] * (er[sl] * p.x + ei[sl] * p.y) + cs[sl] * q.x,
                                  sn[sl] * (er[sl] * p.y - ei[sl] * p.x) + cs[sl] * q.y};
            }
            __syncthreads();
        }
    }
    if (any) s_rot = 1;
    __syncthreads();
    if (!s_rot) return;                       // uniform: read after the last barrier
    // M_chunk <- M_chunk Z for M = X and V (kGbjRows x CC complex each) on FP64 MFMA 16x16x4: the
    // rows staged in G's storage (G is not read past the barrier above; for CC < 64 the V rows in
    // a third array); wave w forms one 16 x 16 output tile of X (the first half of the waves) or
    // of V.  Fed as
    // a = Z^T and b = M^T, the product is (M Z)^T: lane l holds output row l % 16, so 16 lanes
    // store 16 consecutive complex values of one column.
    double2 (*xsx)[CC + 1] = G;
    double2 (*xsv)[CC + 1] = CC >= 2 * kGbjRows ? G + kGbjRows : G + 2 * CC;
#pragma unroll
    for (int t = 0; t < kPre; ++t) {
        const int e = tid + t * NT;
        const int j = e / kGbjRows, i = e % kGbjRows;
        xsx[i][j] = pre[0][t];
        xsv[i][j] = pre[1][t];
    }
    __syncthreads();
    {
        typedef double dv4 __attribute__((ext_vector_type(4)));
        const int wave = tid >> 6, lane = tid & 63;
        constexpr int kTpm = (kGbjRows / 16) * (CC / 16);       // output tiles per matrix
        static_assert(NT / 64 == 2 * kTpm, "one 16 x 16 output tile per wave");
        const int mat = wave / kTpm, tile = wave % kTpm;
        const int tr = (tile / (CC / 16)) * 16, tc = (tile % (CC / 16)) * 16;
        const int li = lane & 15, lk = lane >> 4;
        double2 (*xs)[CC + 1] = mat == 0 ? xsx : xsv;
        dv4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < CC; k0 += 4) {
            const double2 z = Z[k0 + lk][tc + li];
            const double2 x = xs[tr + li][k0 + lk];
            // (M Z)^T = Z^T M^T: re = Zr^T Mr^T - Zi^T Mi^T, im = Zi^T Mr^T + Zr^T Mi^T
            accr = __builtin_amdgcn_mfma_f64_16x16x4f64(z.x, x.x, accr, 0, 0, 0);
            accr = __builtin_amdgcn_mfma_f64_16x16x4f64(-z.y, x.y, accr, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f64_16x16x4f64(z.y, x.x, acci, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f64_16x16x4f64(z.x, x.y, acci, 0, 0, 0);
        }
        // register r of lane l: (M Z)[tr + l % 16][tc + l / 16 + 4 r]
        double2* m = mat == 0 ? X + mt.x_off : V + mt.v_off;
        const int i = tr + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = tc + lk + 4 * r;
            if (i < rows && j < c) m[(int64_t)gcol(j) * n + r0 + i] = double2{accr[r], acci[r]};
        }
    }
    if (tid == 0) flags_cur[tk.x] = 1;
}

// ---- Fused Gram block Jacobi: one workgroup per block pair, columns streamed --------------------
//
// The block one-sided Jacobi of the staged kernel, with the pair's columns never staged whole: a
// round's workgroup streams its c <= 32 columns of X through LDS in chunks of kFbjRows rows to form
// their Gram matrix G = X_c^H X_c (4 x 4 complex register tiles per lane, one wave per quarter of
// each chunk's rows, the four partial sums added in a fixed order), runs the pair's sweep of
// two-sided rotations on G (G <- J^H G J, Z <- Z J; the rotation test and angles of the staged
// kernel), and, when it rotated, applies Z to the columns of X and of V chunk by chunk in place.
// The block width no longer depends on the row count (the staged kernel holds n x c in LDS, so its
// blocks narrow as n grows and a sweep needs ~n^2 / 4096 rounds at n >= 256); a round costs three
// streamed passes over n x c.  The Gram is rebuilt from the columns every round, so convergence is
// judged on fresh inner products, as in the staged kernel.
constexpr int kFbjThreads = 256;
constexpr int kFbjCols = 32;                  // columns of a block pair (2w)
constexpr int kFbjRows = 64;                  // rows per streamed chunk
constexpr int kFbjLd = kFbjRows + 2;          // chunk stored column-major, padded
constexpr int kFbjGld = kFbjCols + 1;         // G and Z row stride
constexpr size_t kFbjLds = sizeof(double2) * ((size_t)kFbjCols * kFbjLd + 2 * (size_t)kFbjCols * kFbjGld);

__global__ __launch_bounds__(kFbjThreads) void fbj_round_kernel(const BjMat* __restrict__ mats,
                                                                const int2* __restrict__ tasks,
                                                                double2* __restrict__ X,
                                                                double2* __restrict__ V, int round,
                                                                const int* __restrict__ flags_prev,
                                                                int* __restrict__ flags_cur,
                                                                double eps,
                                                                const double* __restrict__ fro2,
                                                                double floor_rel2) {
    extern __shared__ __align__(16) unsigned char fbj_lds[];
    double2* xs = reinterpret_cast<double2*>(fbj_lds);              // [kFbjCols][kFbjLd]
    double2* G = xs + (size_t)kFbjCols * kFbjLd;                    // [kFbjCols][kFbjGld]
    double2* Z = G + (size_t)kFbjCols * kFbjGld;                    // [kFbjCols][kFbjGld]
    __shared__ double rot[kFbjCols / 2][4];
    __shared__ int rab[kFbjCols / 2][2];
    __shared__ int s_rot;
    const int2 tk = tasks[blockIdx.x];
    const BjMat mt = mats[tk.x];
    if (round >= mt.N - 1 || flags_prev[tk.x] == 0) return;
    int bp, bq;
    round_robin(mt.N, round, tk.y, bp, bq);
    const int n = mt.n, w = mt.w;
    const int cp = bp < mt.nb ? min(w, n - bp * w) : 0;
    const int cq = bq < mt.nb ? min(w, n - bq * w) : 0;
    const int c = cp + cq;
    if (c < 2) return;
    auto gcol = [&](int j) { return j < cp ? bp * w + j : bq * w + (j - cp); };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // ---- Gram: lane (ta, tb) owns G[4ta .. 4ta+3][4tb .. 4tb+3], wave w the rows w, w+4, ... ----
    const int ta = lane >> 3, tb = lane & 7;
    double2 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = double2{0.0, 0.0};
    const double2* x = X + mt.x_off;
    for (int r0 = 0; r0 < n; r0 += kFbjRows) {
        const int rows = min(kFbjRows, n - r0);
        __syncthreads();
        for (int e = tid; e < kFbjCols * kFbjRows; e += kFbjThreads) {
            const int j = e / kFbjRows, i = e % kFbjRows;
            double2 v = double2{0.0, 0.0};
            if (j < c && i < rows) v = x[(int64_t)gcol(j) * n + r0 + i];
            xs[j * kFbjLd + i] = v;
        }
        __syncthreads();
        for (int i = wave; i < rows; i += kFbjThreads / 64) {
            double2 av[4], bv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                av[k] = xs[(4 * ta + k) * kFbjLd + i];
                bv[k] = xs[(4 * tb + k) * kFbjLd + i];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    acc[a][b].x += av[a].x * bv[b].x + av[a].y * bv[b].y;
                    acc[a][b].y += av[a].x * bv[b].y - av[a].y * bv[b].x;
                }
        }
    }
    __syncthreads();
    // partial sums of waves 0, 1 into the chunk buffer, of waves 2, 3 into G and Z; summed in order
    {
        double2* P = wave == 0 ? xs : wave == 1 ? xs + (size_t)kFbjCols * kFbjGld : wave == 2 ? G : Z;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) P[(4 * ta + a) * kFbjGld + 4 * tb + b] = acc[a][b];
    }
    __syncthreads();
    for (int e = tid; e < kFbjCols * kFbjCols; e += kFbjThreads) {
        const int a = e / kFbjCols, b = e % kFbjCols, o = a * kFbjGld + b;
        const double2 p0 = xs[o], p1 = xs[(size_t)kFbjCols * kFbjGld + o], p2 = G[o], p3 = Z[o];
        G[o] = double2{(p0.x + p1.x) + (p2.x + p3.x), (p0.y + p1.y) + (p2.y + p3.y)};
    }
    __syncthreads();
    for (int e = tid; e < kFbjCols * kFbjCols; e += kFbjThreads) {
        const int a = e / kFbjCols, b = e % kFbjCols;
        Z[a * kFbjGld + b] = double2{a == b ? 1.0 : 0.0, 0.0};
    }
    if (tid == 0) s_rot = 0;
    __syncthreads();
    // ---- the pair's sweep of rotations on G (round 0 of a sweep: every pair of the c columns,
    // so intra-block pairs are visited once per sweep; later rounds: the cross pairs) ----
    const double tol2 = eps * eps * (double)n;
    const double floor2 = floor_rel2 * fro2[tk.x];
    const bool cross_only = round > 0 && cp > 0 && cq > 0;
    const int big = cp >= cq ? cp : cq, small = cp >= cq ? cq : cp;
    const int big0 = cp >= cq ? 0 : cp, small0 = cp >= cq ? cp : 0;
    const int C = c + (c & 1);
    const int np = cross_only ? small : C / 2;
    const int nsteps = cross_only ? big : C - 1;
    for (int r = 0; r < nsteps; ++r) {
        if (tid < np) {
            int a, b;
            bool valid = true;
            if (cross_only) {
                const int i1 = small0 + tid, i2 = big0 + (tid + r) % big;
                a = min(i1, i2);
                b = max(i1, i2);
            } else {
                round_robin(C, r, tid, a, b);
                valid = b < c;
            }
            int ra = -1;
            if (valid) {
                const double al = G[a * kFbjGld + a].x, be = G[b * kFbjGld + b].x;
                const double gr = G[a * kFbjGld + b].x, gi = G[a * kFbjGld + b].y;
                const double ag2 = gr * gr + gi * gi;
                if (al + be >= floor2 && ag2 > tol2 * (al * be) && ag2 != 0.0) {
                    const double ag = sqrt(ag2);
                    const double inv = 1.0 / ag;
                    const double zeta = 0.5 * (be - al) * inv;
                    const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double cs = 1.0 / sqrt(1.0 + t * t);
                    rot[tid][0] = cs;
                    rot[tid][1] = cs * t;
                    rot[tid][2] = gr * inv;
                    rot[tid][3] = gi * inv;
                    ra = a;
                    s_rot = 1;
                }
            }
            rab[tid][0] = ra;
            rab[tid][1] = b;
        }
        __syncthreads();
        // columns: M[:, a] <- c M[:, a] - s conj(e) M[:, b];  M[:, b] <- s e M[:, a] + c M[:, b]
        for (int e = tid; e < np * c; e += kFbjThreads) {
            const int g = e / c, k = e % c;
            const int a = rab[g][0];
            if (a < 0) continue;
            const int b = rab[g][1];
            const double cs = rot[g][0], sn = rot[g][1], er = rot[g][2], ei = rot[g][3];
            double2 p = G[k * kFbjGld + a], q = G[k * kFbjGld + b];
            G[k * kFbjGld + a] = double2{cs * p.x - sn * (er * q.x + ei * q.y), cs * p.y - sn * (er * q.y - ei * q.x)};
            G[k * kFbjGld + b] = double2{sn * (er * p.x - ei * p.y) + cs * q.x, sn * (er * p.y + ei * p.x) + cs * q.y};
            p = Z[k * kFbjGld + a];
            q = Z[k * kFbjGld + b];
            Z[k * kFbjGld + a] = double2{cs * p.x - sn * (er * q.x + ei * q.y), cs * p.y - sn * (er * q.y - ei * q.x)};
            Z[k * kFbjGld + b] = double2{sn * (er * p.x - ei * p.y) + cs * q.x, sn * (er * p.y + ei * p.x) + cs * q.y};
        }
        __syncthreads();
        // rows: M[a, :] <- c M[a, :] - s e M[b, :];  M[b, :] <- s conj(e) M[a, :] + c M[b, :]
        for (int e = tid; e < np * c; e += kFbjThreads) {
            const int g = e / c, k = e % c;
            const int a = rab[g][0];
            if (a < 0) continue;
            const int b = rab[g][1];
            const double cs = rot[g][0], sn = rot[g][1], er = rot[g][2], ei = rot[g][3];
            const double2 p = G[a * kFbjGld + k], q = G[b * kFbjGld + k];
            G[a * kFbjGld + k] = double2{cs * p.x - sn * (er * q.x - ei * q.y), cs * p.y - sn * (er * q.y + ei * q.x)};
            G[b * kFbjGld + k] = double2{sn * (er * p.x + ei * p.y) + cs * q.x, sn * (er * p.y - ei * p.x) + cs * q.y};
        }
        __syncthreads();
    }
    if (!s_rot) return;                       // uniform: read after the last barrier
    // ---- M[:, cols] <- M[:, cols] Z for M = X, then V: lane = row of the chunk, wave w the output
    // columns 8w .. 8w+7 ----
    for (int pass = 0; pass < 2; ++pass) {
        double2* m = pass == 0 ? X + mt.x_off : V + mt.v_off;
        for (int r0 = 0; r0 < n; r0 += kFbjRows) {
            const int rows = min(kFbjRows, n - r0);
            for (int e = tid; e < kFbjCols * kFbjRows; e += kFbjThreads) {
                const int j = e / kFbjRows, i = e % kFbjRows;
                if (j < c && i < rows) xs[j * kFbjLd + i] = m[(int64_t)gcol(j) * n + r0 + i];
            }
            __syncthreads();
            if (lane < rows) {
                double2 out[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) out[q] = double2{0.0, 0.0};
                for (int k = 0; k < c; ++k) {
                    const double2 p = xs[k * kFbjLd + lane];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const double2 z = Z[k * kFbjGld + 8 * wave + q];
                        out[q].x += p.x * z.x - p.y * z.y;
                        out[q].y += p.x * z.y + p.y * z.x;
                    }
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int j = 8 * wave + q;
                    if (j < c) m[(int64_t)gcol(j) * n + r0 + lane] = out[q];
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) flags_cur[tk.x] = 1;
}

// V = I per matrix; flags row 0 = 1 (every matrix runs its first sweep), later rows = 0.
template <typename C2>
__global__ void bj_setup_kernel(const BjMat* __restrict__ mats, int nmat, C2* __restrict__ V,
                                int* __restrict__ flags, int nrows) {
    const int mi = blockIdx.x;
    if (mi < nmat) {
        const BjMat mt = mats[mi];
        C2* v = V + mt.v_off;
        const int64_t nn = (int64_t)mt.n * mt.n;
        for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < nn;
             e += (int64_t)gridDim.y * blockDim.x) {
            C2 one;
            one.x = (e % mt.n) == (e / mt.n) ? 1 : 0;
            one.y = 0;
            v[e] = one;
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0)
        for (int e = threadIdx.x; e < nmat * nrows; e += blockDim.x) flags[e] = e < nmat ? 1 : 0;
}

// Squared Frobenius norm of every n x n matrix X_i (one workgroup per matrix, a fixed summation
// order, so repeated runs agree bit for bit).
template <typename R>
__global__ __launch_bounds__(256) void bj_norm_kernel(const BjMat* __restrict__ mats,
                                                      const typename Cplx<R>::type* __restrict__ X,
                                                      R* __restrict__ fro2) {
    __shared__ R part[256 / 64];
    const BjMat mt = mats[blockIdx.x];
    const typename Cplx<R>::type* x = X + mt.x_off;
    const int64_t nn = (int64_t)mt.n * mt.n;
    R acc = 0;
    for (int64_t e = threadIdx.x; e < nn; e += 256) acc += x[e].x * x[e].x + x[e].y * x[e].y;
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) fro2[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// Singular values, order, U / S / Vh of converged Jacobi factors (one workgroup per matrix).
__global__ __launch_bounds__(kSvdThreads) void svd_finish_kernel(const SvdJob* __restrict__ jobs,
                                                                   const double2* __restrict__ G,
                                                                   const double2* __restrict__ V,
                                                                   double2* __restrict__ U,
                                                                   double* __restrict__ S,
                                                                   double2* __restrict__ Vh) {
    extern __shared__ double smem[];
    const SvdJob job = jobs[blockIdx.x];
    double* sig = smem;
    int* perm = reinterpret_cast<int*>(smem + job.n);
    svd_finish(job, G + job.g_off, V + job.v_off, U, S, Vh, sig, perm);
}

}  // namespace
}  // namespace bug

using namespace bug;

namespace {
hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

unsigned grid_for(int64_t n, int threads) {
    const int64_t g = (n + threads - 1) / threads;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 65535 * 8));
}

double* pinned(bug_ctx* ctx, size_t bytes) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) HIP_CHECK(hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        HIP_CHECK(hipHostMalloc(&ctx->pinned, std::max<size_t>(bytes, 4096)));
        ctx->pinned_bytes = std::max<size_t>(bytes, 4096);
    }
    return static_cast<double*>(ctx->pinned);
}

template <typename T>
T* upload_list(bug_ctx* ctx, hipStream_t st, const T* host, size_t count, int slot) {
    void* dev = scratch(ctx, st, sizeof(T) * std::max<size_t>(count, 1), slot);
    stage_upload(ctx, st, dev, host, sizeof(T) * count);
    return static_cast<T*>(dev);
}

// The blocked QR covers batches whose blocks have at most 64 * 16 rows (a panel column per wave in
// registers) and at least one block wide enough for panels to pay off.
bool blocked_qr_fits(const std::vector<MatJob>& jobs) {
    int mmax = 0, kmax = 0;
    for (const MatJob& j : jobs) {
        mmax = std::max(mmax, j.m);
        kmax = std::max(kmax, std::min(j.m, j.n));
    }
    return mmax <= 64 * 16 && kmax >= 3 * kQrPanel;
}

bool use_blocked_qr(bug_ctx* ctx, const std::vector<MatJob>& jobs) {
    if (ctx->qr_variant == 1) return false;
    int mmax = 0;
    for (const MatJob& j : jobs) mmax = std::max(mmax, j.m);
    if (ctx->qr_variant == 2) return mmax <= 64 * 16;
    return blocked_qr_fits(jobs);
}

// Thin QR of every job (A overwritten, Q m x k, R k x n) by panels and fused block-reflector
// updates.
void blocked_qr(bug_ctx* ctx, hipStream_t st, const std::vector<MatJob>& jobs, double2* A,
                double2* Q, double2* R) {
    const int nmat = (int)jobs.size();
    std::vector<QrPanelJob> pj(nmat);
    int64_t vt = 0;
    int kmax = 0, mmax = 0;
    for (const MatJob& j : jobs) mmax = std::max(mmax, j.m);
    const int pw = mmax <= 512 ? 16 : 8;
    for (int i = 0; i < nmat; ++i) {
        const MatJob& j = jobs[i];
        const int k = std::min(j.m, j.n);
        pj[i] = QrPanelJob{j.a_off, vt, 0, j.m, j.n, j.lda, 0};
        vt += (int64_t)j.m * k;
        kmax = std::max(kmax, k);
    }
    const int npan = (kmax + pw - 1) / pw;
    // V (m x k per matrix) and the panels' T factors (kQrTSlot per panel and matrix)
    const int64_t tsz = (int64_t)npan * nmat * kQrTSlot;
    double2* ws = static_cast<double2*>(scratch(ctx, st, sizeof(double2) * std::max<int64_t>(vt + tsz, 1), 5));
    double2* Vb = ws;
    double2* Tb = ws + vt;
    // slab tasks: trailing updates of panel p at [tstart[p], tstart[p+1]), Q formation of panel p
    // at [qstart[p], qstart[p+1])
    std::vector<ReflTask> tasks;
    std::vector<size_t> tstart(npan + 1), qstart(npan + 1);
    double flops = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int p = 0; p < npan; ++p) {
            (pass == 0 ? tstart : qstart)[p] = tasks.size();
            const int p0 = p * pw;
            for (int i = 0; i < nmat; ++i) {
                const MatJob& j = jobs[i];
                const int k = std::min(j.m, j.n);
                if (k <= p0) continue;
                const int jb = std::min(pw, k - p0);
                const int mp = j.m - p0;
                const int64_t vo = pj[i].v_off + p0 + (int64_t)p0 * j.m;
                const int c0 = pass == 0 ? p0 + jb : p0;
                const int nc = pass == 0 ? j.n - p0 - jb : k - p0;
                const int64_t base = pass == 0 ? j.a_off + p0 : j.q_off + p0;
                const int64_t ld = pass == 0 ? j.lda : j.ldq;
                const int64_t to = ((int64_t)p * nmat + i) * kQrTSlot;
                for (int s0 = 0; s0 < nc; s0 += 16)
                    tasks.push_back(ReflTask{base + (int64_t)(c0 + s0) * ld, vo, to, mp,
                                             std::min(16, nc - s0), jb, (int32_t)ld, j.m,
                                             pass == 0 ? 1 : 0});
                if (nc > 0) flops += 16.0 * mp * nc * (double)jb;
            }
        }
        (pass == 0 ? tstart : qstart)[npan] = tasks.size();
    }
    ctx->last_flops += flops;
    std::vector<unsigned char> blob(sizeof(QrPanelJob) * nmat + sizeof(MatJob) * nmat +
                                    sizeof(ReflTask) * tasks.size());
    std::memcpy(blob.data(), pj.data(), sizeof(QrPanelJob) * nmat);
    std::memcpy(blob.data() + sizeof(QrPanelJob) * nmat, jobs.data(), sizeof(MatJob) * nmat);
    if (!tasks.empty())
        std::memcpy(blob.data() + (sizeof(QrPanelJob) + sizeof(MatJob)) * nmat, tasks.data(),
                    sizeof(ReflTask) * tasks.size());
    unsigned char* dblob = upload_list(ctx, st, blob.data(), blob.size(), 7);
    const QrPanelJob* dp = reinterpret_cast<const QrPanelJob*>(dblob);
    const MatJob* dj = reinterpret_cast<const MatJob*>(dblob + sizeof(QrPanelJob) * nmat);
    const ReflTask* dt = reinterpret_cast<const ReflTask*>(dblob + (sizeof(QrPanelJob) + sizeof(MatJob)) * nmat);
    hipLaunchKernelGGL(qr_qinit_kernel, dim3(nmat, 8), dim3(256), 0, st, dj, Q);
    for (int p = 0; p < npan; ++p) {
        const int p0 = p * pw;
        int mp = 0;
        for (const MatJob& j : jobs)
            if (std::min(j.m, j.n) > p0) mp = std::max(mp, j.m - p0);
        double2* Tp = Tb + (int64_t)p * nmat * kQrTSlot;
        if (pw == 8)
            hipLaunchKernelGGL((qr_panel_kernel<16, 8>), dim3(nmat), dim3(512), 0, st, dp, p0, A, Vb, Tp);
        else if (mp <= 256)
            hipLaunchKernelGGL((qr_panel_kernel<4, 16>), dim3(nmat), dim3(1024), 0, st, dp, p0, A, Vb, Tp);
        else
            hipLaunchKernelGGL((qr_panel_kernel<8, 16>), dim3(nmat), dim3(1024), 0, st, dp, p0, A, Vb, Tp);
        HIP_CHECK(hipGetLastError());
        const size_t nt = tstart[p + 1] - tstart[p];
        if (nt > 0)
            hipLaunchKernelGGL(apply_reflector_kernel, dim3((unsigned)nt), dim3(kReflThreads), 0, st,
                               dt + tstart[p], A, Vb, Tb);
    }
    for (int p = npan - 1; p >= 0; --p) {
        const size_t nq = qstart[p + 1] - qstart[p];
        if (nq > 0)
            hipLaunchKernelGGL(apply_reflector_kernel, dim3((unsigned)nq), dim3(kReflThreads), 0, st,
                               dt + qstart[p], Q, Vb, Tb);
    }
    hipLaunchKernelGGL(qr_finish_kernel, dim3(nmat, std::min(32, std::max(1, (kmax + 15) / 16))),
                       dim3(1024), 0, st, dj, A, Q, R);
    HIP_CHECK(hipGetLastError());
}

// LDS bytes the block-Jacobi stage may use: the device's per-workgroup limit less the Z matrix
// and the kernel's static arrays, at most kBjStageBytes.
size_t bj_stage_bytes(bug_ctx* ctx) {
    if (ctx->lds_limit == 0) {
        int v = 0;
        HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
        ctx->lds_limit = v > 0 ? v : 65536;
    }
    const size_t reserve = sizeof(double2) * kBjMaxCols * kBjMaxCols + 1024;
    const size_t lim = (size_t)ctx->lds_limit;
    return lim > reserve ? std::min(kBjStageBytes, lim - reserve) : 0;
}

// Columns staged per workgroup for an n-column matrix of `elem`-byte elements (even,
// 2 .. kBjMaxCols), 0 if 2 do not fit.
int bj_cols(size_t stage_bytes, int n, size_t elem = sizeof(double2)) {
    const int64_t c = std::min<int64_t>(kBjMaxCols, (int64_t)(stage_bytes / elem) / std::max(n, 1));
    return c >= 2 ? (int)(c & ~1) : 0;
}

bool block_jacobi_fits(bug_ctx* ctx, int nmax) { return bj_cols(bj_stage_bytes(ctx), nmax) >= 2; }

// Block one-sided Jacobi of the n x n matrices X_i (at the running offsets sum_{k<i} n_k^2 of X,
// leading dimension n) with V_i at list[i].v_off of V (set to I here): sweeps of round launches
// until a sweep rotates nothing in any matrix (the host reads the per-sweep flags) or max_sweeps.
template <typename R>
void block_jacobi_t(bug_ctx* ctx, hipStream_t st, const std::vector<SvdJob>& list,
                    typename Cplx<R>::type* X, typename Cplx<R>::type* V, int max_sweeps,
                    int* sweeps_dev, R eps) {
    typedef typename Cplx<R>::type C2;
    const int nmat = (int)list.size();
    const size_t stage = bj_stage_bytes(ctx);
    std::vector<BjMat> mats(nmat);
    std::vector<int2> tasks;
    int64_t xo = 0;
    int rounds = 0;
    size_t lds = 0;
    for (int i = 0; i < nmat; ++i) {
        const int n = list[i].n;
        int c = bj_cols(stage, n, sizeof(C2));
        if (ctx->bj_cols >= 2) c = std::min(c, ctx->bj_cols & ~1);
        if (c < 2) throw Error(BUG_EINVAL, "block Jacobi: matrix too wide for LDS");
        const int w = c / 2;
        const int nb = (n + w - 1) / w;
        const int N = nb + (nb & 1);
        mats[i] = BjMat{xo, list[i].v_off, n, w, nb, N};
        xo += (int64_t)n * n;
        if (n >= 2) {
            for (int p = 0; p < N / 2; ++p) tasks.push_back(make_int2(i, p));
            rounds = std::max(rounds, N - 1);
            lds = std::max(lds, sizeof(C2) * ((size_t)n * std::min(c, n) + (size_t)c * c));
        }
    }
    const int nrows = max_sweeps + 1;
    const size_t flag_bytes = (sizeof(int) * (size_t)nmat * nrows + 255) & ~size_t(255);
    char* fbase = static_cast<char*>(scratch(ctx, st, flag_bytes + sizeof(R) * (size_t)nmat, 3));
    int* flags = reinterpret_cast<int*>(fbase);
    R* fro2 = reinterpret_cast<R*>(fbase + flag_bytes);
    std::vector<unsigned char> blob(sizeof(BjMat) * nmat + sizeof(int2) * tasks.size());
    std::memcpy(blob.data(), mats.data(), sizeof(BjMat) * nmat);
    if (!tasks.empty()) std::memcpy(blob.data() + sizeof(BjMat) * nmat, tasks.data(), sizeof(int2) * tasks.size());
    unsigned char* dblob = upload_list(ctx, st, blob.data(), blob.size(), 2);
    const BjMat* dm = reinterpret_cast<const BjMat*>(dblob);
    const int2* dt = reinterpret_cast<const int2*>(dblob + sizeof(BjMat) * nmat);
    hipLaunchKernelGGL(bj_setup_kernel<C2>, dim3(nmat, 16), dim3(256), 0, st, dm, nmat, V, flags, nrows);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(bj_norm_kernel<R>, dim3(nmat), dim3(256), 0, st, dm, X, fro2);
    HIP_CHECK(hipGetLastError());
    // FP64 sweeps skip pairs below 1e-32 ||X||_F^2 (singular values under 1e-16 ||X||_F, below the
    // 1e-14 sigma_max floor of every selection for blocks up to 4096 wide); the FP32 sweeps, which
    // only precondition, rotate everything
    const R floor_rel2 = sizeof(R) == sizeof(double) ? R(ctx->svd_floor * ctx->svd_floor) : R(0);
    int used = 0;
    std::vector<int> host(nmat);
    if (!tasks.empty()) {
        if (lds > 65536)
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(bj_round_kernel<R>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        // The convergence read of sweep s overlaps sweep s + 1, which is launched before it:
        // if sweep s rotated nothing, every round of sweep s + 1 returns at once (its matrices'
        // flags from sweep s are 0) and writes no flag, so the extra sweep changes nothing.
        int* hp = reinterpret_cast<int*>(pinned(ctx, sizeof(int) * (size_t)nmat * nrows));
        for (hipEvent_t& e : ctx->sync_ev)
            if (!e) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (int s = 0; s < max_sweeps; ++s) {
            for (int r = 0; r < rounds; ++r)
                hipLaunchKernelGGL(bj_round_kernel<R>, dim3((unsigned)tasks.size()), dim3(kBjThreads), lds,
                                   st, dm, dt, X, V, r, flags + (size_t)s * nmat,
                                   flags + (size_t)(s + 1) * nmat, eps,
                                   std::max(1, ctx->bj_inner), fro2, floor_rel2);
            HIP_CHECK(hipGetLastError());
            ++used;
            HIP_CHECK(hipMemcpyAsync(hp + (size_t)s * nmat, flags + (size_t)(s + 1) * nmat,
                                     sizeof(int) * nmat, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipEventRecord(ctx->sync_ev[s & 1], st));
            if (s == 0) continue;
            HIP_CHECK(hipEventSynchronize(ctx->sync_ev[(s - 1) & 1]));
            bool any = false;
            for (int i = 0; i < nmat; ++i) any |= hp[(size_t)(s - 1) * nmat + i] != 0;
            if (!any) break;
        }
        // the last sweep's read (or the speculative sweep) must land before the flags are reused
        HIP_CHECK(hipEventSynchronize(ctx->sync_ev[(used - 1) & 1]));
    }
    if (sweeps_dev) {
        // sweeps per matrix: the sweeps that rotated plus the one that found it converged
        std::vector<int> all((size_t)nmat * nrows);
        HIP_CHECK(hipMemcpyAsync(all.data(), flags, sizeof(int) * all.size(), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (int i = 0; i < nmat; ++i) {
            int k = 0;
            for (int s = 1; s <= used; ++s) k += all[(size_t)s * nmat + i] != 0;
            host[i] = std::min(k + 1, used);
        }
        HIP_CHECK(hipMemcpyAsync(sweeps_dev, host.data(), sizeof(int) * nmat, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
}

void block_jacobi(bug_ctx* ctx, hipStream_t st, const std::vector<SvdJob>& list, double2* X,
                  double2* V, int max_sweeps, int* sweeps_dev) {
    block_jacobi_t<double>(ctx, st, list, X, V, max_sweeps, sweeps_dev, 2.220446049250313e-16);
}

// Fused Gram block Jacobi (fbj_round_kernel per round) of the n x n matrices X_i at the running
// offsets sum_{k<i} n_k^2 (leading dimension n), V_i at list[i].v_off (set to I here); blocks of
// 16 columns whatever n; sweeps and the overlapped convergence read as in block_jacobi_t.
void fused_block_jacobi(bug_ctx* ctx, hipStream_t st, const std::vector<SvdJob>& list, double2* X,
                        double2* V, int max_sweeps, int* sweeps_dev) {
    const int nmat = (int)list.size();
    std::vector<BjMat> mats(nmat);
    std::vector<int2> tasks;
    int64_t xo = 0;
    int rounds = 0;
    for (int i = 0; i < nmat; ++i) {
        const int n = list[i].n;
        const int w = std::max(1, std::min(kFbjCols / 2, (n + 1) / 2));
        const int nb = (n + w - 1) / w;
        const int N = nb + (nb & 1);
        mats[i] = BjMat{xo, list[i].v_off, n, w, nb, N};
        xo += (int64_t)n * n;
        if (n >= 2) {
            for (int p = 0; p < N / 2; ++p) tasks.push_back(make_int2(i, p));
            rounds = std::max(rounds, N - 1);
        }
    }
    const int nrows = max_sweeps + 1;
    const size_t flag_bytes = (sizeof(int) * (size_t)nmat * nrows + 255) & ~size_t(255);
    char* fbase = static_cast<char*>(scratch(ctx, st, flag_bytes + sizeof(double) * (size_t)nmat, 3));
    int* flags = reinterpret_cast<int*>(fbase);
    double* fro2 = reinterpret_cast<double*>(fbase + flag_bytes);
    std::vector<unsigned char> blob(sizeof(BjMat) * nmat + sizeof(int2) * tasks.size());
    std::memcpy(blob.data(), mats.data(), sizeof(BjMat) * nmat);
    if (!tasks.empty())
        std::memcpy(blob.data() + sizeof(BjMat) * nmat, tasks.data(), sizeof(int2) * tasks.size());
    unsigned char* dblob = upload_list(ctx, st, blob.data(), blob.size(), 2);
    const BjMat* dm = reinterpret_cast<const BjMat*>(dblob);
    const int2* dt = reinterpret_cast<const int2*>(dblob + sizeof(BjMat) * nmat);
    hipLaunchKernelGGL(bj_setup_kernel<double2>, dim3(nmat, 16), dim3(256), 0, st, dm, nmat, V, flags, nrows);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(bj_norm_kernel<double>, dim3(nmat), dim3(256), 0, st, dm, X, fro2);
    HIP_CHECK(hipGetLastError());
    int used = 0;
    if (!tasks.empty()) {
        HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fbj_round_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFbjLds));
        int* hp = reinterpret_cast<int*>(pinned(ctx, sizeof(int) * (size_t)nmat * nrows));
        for (hipEvent_t& e : ctx->sync_ev)
            if (!e) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (int s = 0; s < max_sweeps; ++s) {
            for (int r = 0; r < rounds; ++r)
                hipLaunchKernelGGL(fbj_round_kernel, dim3((unsigned)tasks.size()), dim3(kFbjThreads),
                                   kFbjLds, st, dm, dt, X, V, r, flags + (size_t)s * nmat,
                                   flags + (size_t)(s + 1) * nmat, 2.220446049250313e-16, fro2,
                                   ctx->svd_floor * ctx->svd_floor);
            HIP_CHECK(hipGetLastError());
            ++used;
            HIP_CHECK(hipMemcpyAsync(hp + (size_t)s * nmat, flags + (size_t)(s + 1) * nmat,
                                     sizeof(int) * nmat, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipEventRecord(ctx->sync_ev[s & 1], st));
            if (s == 0) continue;
            HIP_CHECK(hipEventSynchronize(ctx->sync_ev[(s - 1) & 1]));
            bool any = false;
            for (int i = 0; i < nmat; ++i) any |= hp[(size_t)(s - 1) * nmat + i] != 0;
            if (!any) break;
        }
        HIP_CHECK(hipEventSynchronize(ctx->sync_ev[(used - 1) & 1]));
    }
    if (sweeps_dev) {
        std::vector<int> all((size_t)nmat * nrows), host(nmat);
        HIP_CHECK(hipMemcpyAsync(all.data(), flags, sizeof(int) * all.size(), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (int i = 0; i < nmat; ++i) {
            int k = 0;
            for (int s = 1; s <= used; ++s) k += all[(size_t)s * nmat + i] != 0;
            host[i] = std::min(k + 1, used);
        }
        HIP_CHECK(hipMemcpyAsync(sweeps_dev, host.data(), sizeof(int) * nmat, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
}

// ---- FP32-preconditioned block Jacobi --------------------------------------------------------
//
// Sweeps of the block Jacobi in FP32 (twice the staged columns per workgroup, so about half the
// rounds per sweep) bring X's columns to orthogonality at FP32 accuracy; their rotation V32 is
// made exactly unitary in FP64 by a QR (Q_v, positive-diagonal gauge), X is rotated by it in
// FP64 (X2 = X Q_v, a grouped GEMM) and the FP64 block Jacobi finishes from there (V2); the
// accumulated rotation is V = Q_v V2.  The FP64 sweeps start from couplings of order 1e-7 and
// converge in a few sweeps; the result is an FP64 one-sided Jacobi SVD of X (the FP32 phase only
// chooses the starting basis), to rounding the same as the FP64-only path.

// One workgroup per matrix: X32 = s X with s = 2^-e, e the binary exponent of ||X||_F (so the
// FP32 column norms stay below 1 and far from overflow).
__global__ __launch_bounds__(1024) void bj_to_f32_kernel(const BjMat* __restrict__ mats,
                                                         const double2* __restrict__ X,
                                                         float2* __restrict__ X32) {
    __shared__ double red[16];
    __shared__ double s_scale;
    const BjMat mt = mats[blockIdx.x];
    const int64_t nn = (int64_t)mt.n * mt.n;
    const double2* x = X + mt.x_off;
    float2* y = X32 + mt.x_off;
    double ss = 0.0;
    for (int64_t e = threadIdx.x; e < nn; e += blockDim.x) ss += x[e].x * x[e].x + x[e].y * x[e].y;
    ss = wave_sum(ss);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        int ex = 0;
        if (t > 0.0 && isfinite(t)) frexp(sqrt(t), &ex);
        s_scale = ldexp(1.0, -ex);
    }
    __syncthreads();
    const double sc = s_scale;
    for (int64_t e = threadIdx.x; e < nn; e += blockDim.x)
        y[e] = float2{(float)(sc * x[e].x), (float)(sc * x[e].y)};
}

__global__ void f32_to_f64_kernel(int64_t n, const float2* __restrict__ in, double2* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = double2{(double)in[i].x, (double)in[i].y};
}

void mixed_block_jacobi(bug_ctx* ctx, hipStream_t st, const std::vector<SvdJob>& list, double2* X,
                        double2* V, int max_sweeps, int* sweeps_dev) {
    const int nmat = (int)list.size();
    std::vector<int64_t> off(nmat);
    int64_t ntot = 0, ksum = 0;
    for (int i = 0; i < nmat; ++i) {
        off[i] = ntot;
        ntot += (int64_t)list[i].n * list[i].n;
        ksum += list[i].n;
    }
    const size_t bytes = (size_t)ntot * (2 * sizeof(float2) + 5 * sizeof(double2)) +
                         sizeof(double2) * (size_t)std::max<int64_t>(ksum, 1) + 1024;
    char* base = static_cast<char*>(scratch(ctx, st, bytes, 9));
    auto carve = [&](size_t b) {
        char* p = base;
        base += (b + 255) & ~size_t(255);
        return p;
    };
    float2* X32 = reinterpret_cast<float2*>(carve(sizeof(float2) * ntot));
    float2* V32 = reinterpret_cast<float2*>(carve(sizeof(float2) * ntot));
    double2* V64 = reinterpret_cast<double2*>(carve(sizeof(double2) * ntot));
    double2* Qv = reinterpret_cast<double2*>(carve(sizeof(double2) * ntot));
    double2* Rv = reinterpret_cast<double2*>(carve(sizeof(double2) * ntot));
    double2* X2 = reinterpret_cast<double2*>(carve(sizeof(double2) * ntot));
    double2* V2 = reinterpret_cast<double2*>(carve(sizeof(double2) * ntot));
    double2* tau = reinterpret_cast<double2*>(carve(sizeof(double2) * std::max<int64_t>(ksum, 1)));
    (void)carve;
    // the FP32 and FP64 phases keep V at the same running offsets as X
    std::vector<SvdJob> loc(list);
    std::vector<BjMat> bm(nmat);
    for (int i = 0; i < nmat; ++i) {
        loc[i].v_off = off[i];
        bm[i] = BjMat{off[i], off[i], list[i].n, 0, 0, 0};
    }
    const BjMat* dbm = upload_list(ctx, st, bm.data(), bm.size(), 2);
    hipLaunchKernelGGL(bj_to_f32_kernel, dim3(nmat), dim3(1024), 0, st, dbm, X, X32);
    HIP_CHECK(hipGetLastError());
    block_jacobi_t<float>(ctx, st, loc, X32, V32, std::max(1, ctx->mixed_fp32_sweeps), nullptr,
                          1.1920928955078125e-07f);
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(ntot, 256)), dim3(256), 0, st, ntot, V32, V64);
    HIP_CHECK(hipGetLastError());
    // Q_v: unitary factor of V32 in FP64
    std::vector<MatJob> qj(nmat);
    int64_t to = 0;
    for (int i = 0; i < nmat; ++i) {
        const int n = list[i].n;
        qj[i] = MatJob{off[i], off[i], off[i], to, n, n, std::max(1, n), std::max(1, n), std::max(1, n)};
        to += n;
    }
    if (use_blocked_qr(ctx, qj)) {
        blocked_qr(ctx, st, qj, V64, Qv, Rv);
    } else {
        MatJob* dq = upload_list(ctx, st, qj.data(), qj.size(), 2);
        hipLaunchKernelGGL(qr_kernel, dim3(nmat), dim3(kQrThreads), 0, st, dq, V64, Qv, Rv, tau);
    }
    HIP_CHECK(hipGetLastError());
    // X2 = X Q_v
    auto square_plan = [&](uint64_t tag, bool to_v) {
        Hasher h;
        h.pod(tag);
        for (int i = 0; i < nmat; ++i) {
            h.pod(list[i].n);
            h.pod(off[i]);
            h.pod(to_v ? list[i].v_off : off[i]);
        }
        return h.h;
    };
    auto build = [&](Plan& plan, bool to_v) {
        for (int i = 0; i < nmat; ++i) {
            const int n = list[i].n;
            if (n <= 0) continue;
            plan.tasks.push_back(GemmTask{to_v ? list[i].v_off : off[i], n, n, n,
                                          (int32_t)plan.terms.size(), (int32_t)plan.terms.size() + 1, 0});
            plan.terms.push_back(GemmTerm{off[i], off[i], n, n, n, BUG_OP_N | (BUG_OP_N << 2), 1.0, 0.0});
            plan.flops += 8.0 * n * (double)n * n;
        }
    };
    launch_cached_plan(ctx, square_plan(0x5F2A, false), [&](Plan& p) { build(p, false); },
                       BUG_PREC_FP64, X, Qv, X2, st);
    block_jacobi_t<double>(ctx, st, loc, X2, V2, max_sweeps, sweeps_dev, 2.220446049250313e-16);
    // V = Q_v V2 at the caller's offsets; the rotated X back into X
    launch_cached_plan(ctx, square_plan(0x5F2B, true), [&](Plan& p) { build(p, true); },
                       BUG_PREC_FP64, Qv, V2, V, st);
    HIP_CHECK(hipMemcpyAsync(X, X2, sizeof(double2) * (size_t)ntot, hipMemcpyDeviceToDevice, st));
}


// The Gram block Jacobi (gbj_gram_kernel + gbj_rotate_kernel per round) of the n x n matrices X_i
// at the running offsets sum_{k<i} n_k^2, with V_i at list[i].v_off (set to I here); sweeps as in
// block_jacobi.
void gram_block_jacobi(bug_ctx* ctx, hipStream_t st, const std::vector<SvdJob>& list, double2* X,
                       double2* V, int max_sweeps, int* sweeps_dev) {
    const int nmat = (int)list.size();
    // block pairs of up to 32 columns on 512-thread rotate workgroups, or of up to 64 on 1024
    // threads when the context asks for wider pairs (bug_set_gram_jacobi_params)
    const bool wide = ctx->gbj_cols > kGbjCols;
    const int c = std::max(2, std::min(wide ? kGbjColsWide : kGbjCols, ctx->gbj_cols) & ~1);
    const int w = c / 2;
    std::vector<BjMat> mats(nmat);
    std::vector<int2> tasks;
    int64_t xo = 0;
    int rounds = 0, maxch = 1;
    for (int i = 0; i < nmat; ++i) {
        const int n = list[i].n;
        const int nb = (n + w - 1) / w;
        const int N = nb + (nb & 1);
        mats[i] = BjMat{xo, list[i].v_off, n, w, nb, N};
        xo += (int64_t)n * n;
        if (n >= 2) {
            for (int p = 0; p < N / 2; ++p) tasks.push_back(make_int2(i, p));
            rounds = std::max(rounds, N - 1);
            maxch = std::max(maxch, (n + kGbjRows - 1) / kGbjRows);
        }
    }
    const int nrows = max_sweeps + 1;
    int* flags = static_cast<int*>(scratch(ctx, st, sizeof(int) * (size_t)nmat * nrows, 3));
    std::vector<unsigned char> blob(sizeof(BjMat) * nmat + sizeof(int2) * tasks.size());
    std::memcpy(blob.data(), mats.data(), sizeof(BjMat) * nmat);
    if (!tasks.empty()) std::memcpy(blob.data() + sizeof(BjMat) * nmat, tasks.data(), sizeof(int2) * tasks.size());
    unsigned char* dblob = upload_list(ctx, st, blob.data(), blob.size(), 2);
    const BjMat* dm = reinterpret_cast<const BjMat*>(dblob);
    const int2* dt = reinterpret_cast<const int2*>(dblob + sizeof(BjMat) * nmat);
    hipLaunchKernelGGL(bj_setup_kernel<double2>, dim3(nmat, 16), dim3(256), 0, st, dm, nmat, V, flags, nrows);
    HIP_CHECK(hipGetLastError());
    int used = 0;
    if (!tasks.empty()) {
        const int cc = wide ? kGbjColsWide : kGbjCols;
        double2* gbuf = static_cast<double2*>(scratch(
            ctx, st, sizeof(double2) * (size_t)cc * cc * maxch * tasks.size(), 6));
        if (wide)
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(gbj_rotate_kernel<kGbjColsWide, 1024>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)gbj_rot_lds(kGbjColsWide)));
        int* hp = reinterpret_cast<int*>(pinned(ctx, sizeof(int) * (size_t)nmat * nrows));
        for (hipEvent_t& e : ctx->sync_ev)
            if (!e) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        const dim3 grid((unsigned)tasks.size(), (unsigned)maxch);
        const int inner = std::max(1, ctx->gbj_inner);
        const double eps = 2.220446049250313e-16;
        // The convergence read of sweep s overlaps sweep s + 1, launched before it (as in
        // block_jacobi_t): a sweep after one that rotated nothing returns at once in every launch.
        for (int s = 0; s < max_sweeps; ++s) {
            for (int r = 0; r < rounds; ++r) {
                const int* fp = flags + (size_t)s * nmat;
                int* fc = flags + (size_t)(s + 1) * nmat;
                if (wide) {
                    constexpr int wc = kGbjColsWide;
                    hipLaunchKernelGGL((gbj_gram_kernel<wc, 1024>), grid, dim3(1024), gbj_gram_lds(wc), st,
                                       dm, dt, X, gbuf, r, fp, maxch);
                    hipLaunchKernelGGL((gbj_rotate_kernel<wc, 1024>), grid, dim3(1024), gbj_rot_lds(wc), st,
                                       dm, dt, X, V, gbuf, r, fp, fc, eps, inner, maxch);
                } else {
                    constexpr int nc = kGbjCols;
                    hipLaunchKernelGGL((gbj_gram_kernel<nc, 256>), grid, dim3(256), gbj_gram_lds(nc), st,
                                       dm, dt, X, gbuf, r, fp, maxch);
                    hipLaunchKernelGGL((gbj_rotate_kernel<nc, kGbjRotThreads>), grid, dim3(kGbjRotThreads),
                                       gbj_rot_lds(nc), st,
                                       dm, dt, X, V, gbuf, r, fp, fc, eps, inner, maxch);
                }
            }
            HIP_CHECK(hipGetLastError());
            ++used;
            HIP_CHECK(hipMemcpyAsync(hp + (size_t)s * nmat, flags + (size_t)(s + 1) * nmat,
                                     sizeof(int) * nmat, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipEventRecord(ctx->sync_ev[s & 1], st));
            if (s == 0) continue;
            HIP_CHECK(hipEventSynchronize(ctx->sync_ev[(s - 1) & 1]));
            bool any = false;
            for (int i = 0; i < nmat; ++i) any |= hp[(size_t)(s - 1) * nmat + i] != 0;
            if (!any) break;
        }
        // the last sweep's read (or the speculative sweep) must land before the flags are reused
        HIP_CHECK(hipEventSynchronize(ctx->sync_ev[(used - 1) & 1]));
    }
    if (sweeps_dev) {
        std::vector<int> all((size_t)nmat * nrows), host(nmat);
        HIP_CHECK(hipMemcpyAsync(all.data(), flags, sizeof(int) * all.size(), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (int i = 0; i < nmat; ++i) {
            int k = 0;
            for (int s = 1; s <= used; ++s) k += all[(size_t)s * nmat + i] != 0;
            host[i] = std::min(k + 1, used);
        }
        HIP_CHECK(hipMemcpyAsync(sweeps_dev, host.data(), sizeof(int) * nmat, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
}
}  // namespace

extern "C" {

int bug_axpy(bug_ctx* ctx, int64_t n, double ar, double ai, const void* x, void* y, void* stream) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!x || !y))) throw Error(BUG_EINVAL, "bad axpy arguments");
        if (n == 0) return;
        hipLaunchKernelGGL(zaxpy_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), n,
                           ar, ai, static_cast<const double2*>(x), static_cast<double2*>(y));
        HIP_CHECK(hipGetLastError());
    });
}

int bug_scal(bug_ctx* ctx, int64_t n, double ar, double ai, void* x, void* stream) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && !x)) throw Error(BUG_EINVAL, "bad scal arguments");
        if (n == 0) return;
        hipLaunchKernelGGL(zscal_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), n,
                           ar, ai, static_cast<double2*>(x));
        HIP_CHECK(hipGetLastError());
    });
}

int bug_axpy_many(bug_ctx* ctx, int64_t n, int32_t k, const double* alphas, const void* const* xs,
                  void* y, void* stream) {
    return guarded(ctx, [&] {
        if (n < 0 || k < 0 || k > kMaxDots || (k > 0 && (!alphas || !xs)) || (n > 0 && k > 0 && !y))
            throw Error(BUG_EINVAL, "bad axpy_many arguments (k must be <= 64)");
        if (n == 0 || k == 0) return;
        AxpyCoeffs c;
        DotPtrs p;
        for (int j = 0; j < k; ++j) {
            if (!xs[j]) throw Error(BUG_EINVAL, "null vector in axpy_many");
            c.a[j] = double2{alphas[2 * j], alphas[2 * j + 1]};
            p.p[j] = static_cast<const double2*>(xs[j]);
        }
        hipLaunchKernelGGL(zaxpy_many_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                           as_stream(stream), n, k, c, p, static_cast<double2*>(y));
        HIP_CHECK(hipGetLastError());
    });
}

namespace {
// Validate segments [off, off + len) of a buffer of n elements and upload them.
const int64_t* upload_segments(bug_ctx* ctx, hipStream_t st, int32_t nseg, const int64_t* segs,
                               int64_t n, int64_t& maxlen) {
    if (nseg < 1 || nseg > 4096 || !segs) throw Error(BUG_EINVAL, "segments: 1 .. 4096 needed");
    maxlen = 0;
    for (int i = 0; i < nseg; ++i) {
        const int64_t off = segs[2 * i], len = segs[2 * i + 1];
        if (off < 0 || len < 0 || off + len > n) throw Error(BUG_EINVARIANT, "segment outside the buffer");
        maxlen = std::max(maxlen, len);
    }
    return upload_list(ctx, st, segs, 2 * (size_t)nseg, 2);
}
}  // namespace

int bug_dotc_segments(bug_ctx* ctx, int64_t n, int32_t nseg, const int64_t* segs, int32_t k,
                      const void* const* xs, const void* y, double* out, void* stream) {
    return guarded(ctx, [&] {
        if (n < 0 || k < 0 || k > kMaxDots || (k > 0 && (!xs || !out || !y)))
            throw Error(BUG_EINVAL, "bad dotc_segments arguments (k must be <= 64)");
        if (k == 0) return;
        hipStream_t st = as_stream(stream);
        int64_t maxlen = 0;
        const int64_t* dseg = upload_segments(ctx, st, nseg, segs, n, maxlen);
        DotPtrs p;
        for (int j = 0; j < k; ++j) {
            if (!xs[j]) throw Error(BUG_EINVAL, "null vector in dotc_segments");
            p.p[j] = static_cast<const double2*>(xs[j]);
        }
        const int nchunk = (int)std::max<int64_t>(1, (maxlen + kSegChunk - 1) / kSegChunk);
        const size_t rows = (size_t)k * nseg;
        double2* part = static_cast<double2*>(
            scratch(ctx, st, sizeof(double2) * (rows * nchunk + rows), 3));
        double2* res = part + rows * nchunk;
        hipLaunchKernelGGL(seg_multidot_partial, dim3(nseg, nchunk), dim3(kSegThreads), 0, st, nseg,
                           dseg, k, p, static_cast<const double2*>(y), nchunk, part);
        hipLaunchKernelGGL(seg_multidot_final, dim3((unsigned)rows), dim3(64), 0, st, nchunk, part, res);
        HIP_CHECK(hipGetLastError());
        double* h = pinned(ctx, sizeof(double2) * rows);
        HIP_CHECK(hipMemcpyAsync(h, res, sizeof(double2) * rows, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        std::memcpy(out, h, sizeof(double2) * rows);
    });
}

int bug_axpy_segments(bug_ctx* ctx, int64_t n, int32_t nseg, const int64_t* segs, int32_t k,
                      const double* alphas, const void* const* xs, void* y, void* stream) {
    return guarded(ctx, [&] {
        if (n < 0 || k < 0 || k > kMaxDots || (k > 0 && (!xs || !alphas || !y)))
            throw Error(BUG_EINVAL, "bad axpy_segments arguments (k must be <= 64)");
        if (k == 0) return;
        hipStream_t st = as_stream(stream);
        int64_t maxlen = 0;
        const int64_t* dseg = upload_segments(ctx, st, nseg, segs, n, maxlen);
        DotPtrs p;
        for (int j = 0; j < k; ++j) {
            if (!xs[j]) throw Error(BUG_EINVAL, "null vector in axpy_segments");
            p.p[j] = static_cast<const double2*>(xs[j]);
        }
        const double2* da = upload_list(ctx, st, reinterpret_cast<const double2*>(alphas),
                                        (size_t)k * nseg, 1);
        const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>((maxlen + 255) / 256, 64));
        hipLaunchKernelGGL(seg_axpy_many_kernel, dim3(nseg, gy), dim3(256), 0, st, nseg, dseg, k, da,
                           p, static_cast<double2*>(y));
        HIP_CHECK(hipGetLastError());
    });
}

int bug_dotc(bug_ctx* ctx, int64_t n, int32_t k, const void* const* xs, const void* y,
             double* out, void* stream) {
    return guarded(ctx, [&] {
        if (n < 0 || k < 0 || k > kMaxDots || (k > 0 && (!xs || !out)) || (n > 0 && !y))
            throw Error(BUG_EINVAL, "bad dotc arguments (k must be <= 64)");
        if (k == 0) return;
        hipStream_t st = as_stream(stream);
        DotPtrs p;
        for (int j = 0; j < k; ++j) {
            if (n > 0 && !xs[j]) throw Error(BUG_EINVAL, "null vector in dotc");
            p.p[j] = static_cast<const double2*>(xs[j]);
        }
        const int64_t per_block = (int64_t)kDotThreads * kDotRows;
        const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, 4096));
        double2* part = static_cast<double2*>(
            scratch(ctx, st, sizeof(double2) * ((size_t)nblk * k + kMaxDots), 3));
        double2* res = part + (size_t)nblk * k;
        if (n > 0 && nblk * per_block < n) {
            // more rows than one pass of 4096 blocks covers: fold in row chunks
            throw Error(BUG_EINVAL, "dotc vector longer than 4096 * 1024 elements");
        }
        hipLaunchKernelGGL(multidot_partial, dim3(nblk), dim3(kDotThreads), 0, st, n, k, p,
                           static_cast<const double2*>(y), part);
        hipLaunchKernelGGL(multidot_final, dim3(k), dim3(64), 0, st, nblk, part, res);
        HIP_CHECK(hipGetLastError());
        double* h = pinned(ctx, sizeof(double2) * k);
        HIP_CHECK(hipMemcpyAsync(h, res, sizeof(double2) * k, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        std::memcpy(out, h, sizeof(double2) * k);
    });
}

int bug_copy_blocks(bug_ctx* ctx, int32_t nop, const bug_copy* ops, const void* src,
                    int64_t src_nelem, void* dst, int64_t dst_nelem, const double* scale,
                    int32_t accumulate, void* stream) {
    return guarded(ctx, [&] {
        if (nop < 0 || (nop > 0 && (!ops || !src || !dst))) throw Error(BUG_EINVAL, "bad copy list");
        if (nop == 0) return;
        std::vector<CopyOp> list(nop);
        int64_t maxe = 0;
        for (int i = 0; i < nop; ++i) {
            const bug_copy& o = ops[i];
            if (o.rows < 0 || o.cols < 0 || o.mode < 0 || o.mode > 3 || o.scale_mode < 0 ||
                o.scale_mode > 2 || (o.scale_mode && !scale))
                throw Error(BUG_EINVAL, "bad copy op");
            if (o.rows == 0 || o.cols == 0) {
                list[i] = CopyOp{0, 0, 1, 1, 0, 0, 0, 0, 0};
                continue;
            }
            const bool plain = (o.mode == 0 || o.mode == 3);
            const int64_t sr = plain ? o.rows : o.cols, sc = plain ? o.cols : o.rows;
            if (o.src_off < 0 || o.dst_off < 0 || o.src_ld < sr || o.dst_ld < o.rows ||
                o.src_off + (sc - 1) * (int64_t)o.src_ld + sr > src_nelem ||
                o.dst_off + (int64_t)(o.cols - 1) * o.dst_ld + o.rows > dst_nelem)
                throw Error(BUG_EINVARIANT, "copy block outside its buffer");
            list[i] = CopyOp{o.src_off, o.dst_off, o.src_ld, o.dst_ld, o.rows, o.cols, o.mode,
                             o.scale_mode, o.scale_off};
            maxe = std::max<int64_t>(maxe, (int64_t)o.rows * o.cols);
        }
        hipStream_t st = as_stream(stream);
        CopyOp* d = upload_list(ctx, st, list.data(), list.size(), 2);
        const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>((maxe + 255) / 256, 64));
        hipLaunchKernelGGL(copy_blocks_kernel, dim3(nop, gy), dim3(256), 0, st, d,
                           static_cast<const double2*>(src), static_cast<double2*>(dst), scale,
                           accumulate);
        HIP_CHECK(hipGetLastError());
    });
}

int bug_block_sqnorms(bug_ctx* ctx, int32_t nblk, const bug_copy* blocks, const void* src,
                      int64_t src_nelem, double* out, void* stream) {
    return guarded(ctx, [&] {
        if (nblk < 0 || (nblk > 0 && (!blocks || !src || !out))) throw Error(BUG_EINVAL, "bad block list");
        if (nblk == 0) return;
        std::vector<CopyOp> list(nblk);
        for (int i = 0; i < nblk; ++i) {
            const bug_copy& o = blocks[i];
            if (o.rows < 0 || o.cols < 0) throw Error(BUG_EINVAL, "bad block dims");
            if (o.rows > 0 && o.cols > 0 &&
                (o.src_off < 0 || o.src_ld < o.rows ||
                 o.src_off + (int64_t)(o.cols - 1) * o.src_ld + o.rows > src_nelem))
                throw Error(BUG_EINVARIANT, "block outside its buffer");
            list[i] = CopyOp{o.src_off, 0, o.src_ld, 1, o.rows, o.cols, 0, 0, 0};
        }
        // enough workgroups per block to spread a large one over the device (16 elements per
        // thread and chunk at least), at most 64 chunks
        int64_t biggest = 1;
        for (const CopyOp& o : list) biggest = std::max<int64_t>(biggest, (int64_t)o.rows * o.cols);
        const int chunks = (int)std::min<int64_t>(64, std::max<int64_t>(1, biggest / (256 * 16)));
        hipStream_t st = as_stream(stream);
        CopyOp* d = upload_list(ctx, st, list.data(), list.size(), 2);
        const size_t nres = (size_t)nblk * chunks;
        double* res = static_cast<double*>(scratch(ctx, st, sizeof(double) * nres, 3));
        hipLaunchKernelGGL(block_sqnorm_kernel, dim3(nblk, chunks), dim3(256), 0, st, d,
                           static_cast<const double2*>(src), res);
        HIP_CHECK(hipGetLastError());
        double* h = pinned(ctx, sizeof(double) * nres);
        HIP_CHECK(hipMemcpyAsync(h, res, sizeof(double) * nres, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (int i = 0; i < nblk; ++i) {
            double acc = 0.0;
            for (int y = 0; y < chunks; ++y) acc += h[(size_t)i * chunks + y];
            out[i] = acc;
        }
    });
}

int bug_block_qr(bug_ctx* ctx, int32_t nmat, const bug_qr_job* jobs, void* a, int64_t a_nelem,
                 void* q, int64_t q_nelem, void* r, int64_t r_nelem, void* stream) {
    return guarded(ctx, [&] {
        if (nmat < 0 || (nmat > 0 && (!jobs || !a || !q || !r))) throw Error(BUG_EINVAL, "bad qr jobs");
        if (nmat == 0) return;
        std::vector<MatJob> list(nmat);
        int64_t wtot = 0;
        for (int i = 0; i < nmat; ++i) {
            const bug_qr_job& j = jobs[i];
            const int k = std::min(j.m, j.n);
            if (j.m < 0 || j.n < 0 || j.lda < std::max(1, j.m) || j.ldq < std::max(1, j.m) ||
                j.ldr < std::max(1, k))
                throw Error(BUG_EINVAL, "bad qr job dims");
            if (j.m > 0 && j.n > 0 &&
                (j.a_off + (int64_t)(j.n - 1) * j.lda + j.m > a_nelem ||
                 j.q_off + (int64_t)(k - 1) * j.ldq + j.m > q_nelem ||
                 j.r_off + (int64_t)(j.n - 1) * j.ldr + k > r_nelem))
                throw Error(BUG_EINVARIANT, "qr job outside its buffers");
            list[i] = MatJob{j.a_off, j.q_off, j.r_off, wtot, j.m, j.n, j.lda, j.ldq, j.ldr};
            wtot += k;
        }
        hipStream_t st = as_stream(stream);
        if (use_blocked_qr(ctx, list)) {
            blocked_qr(ctx, st, list, static_cast<double2*>(a), static_cast<double2*>(q),
                       static_cast<double2*>(r));
            return;
        }
        MatJob* d = upload_list(ctx, st, list.data(), list.size(), 2);
        double2* tau = static_cast<double2*>(scratch(ctx, st, sizeof(double2) * std::max<int64_t>(wtot, 1), 1));
        hipLaunchKernelGGL(qr_kernel, dim3(nmat), dim3(kQrThreads), 0, st, d,
                           static_cast<double2*>(a), static_cast<double2*>(q),
                           static_cast<double2*>(r), tau);
        HIP_CHECK(hipGetLastError());
    });
}

int bug_block_qrp(bug_ctx* ctx, int32_t nmat, const bug_qr_job* jobs, void* a, int64_t a_nelem,
                  void* q, int64_t q_nelem, double* rdiag, int64_t rdiag_nelem, void* stream) {
    return guarded(ctx, [&] {
        if (nmat < 0 || (nmat > 0 && (!jobs || !a || !q || !rdiag)))
            throw Error(BUG_EINVAL, "bad pivoted qr jobs");
        if (nmat == 0) return;
        std::vector<MatJob> list(nmat);
        int64_t wtot = 0;
        int nmax = 1;
        for (int i = 0; i < nmat; ++i) {
            const bug_qr_job& j = jobs[i];
            const int k = std::min(j.m, j.n);
            if (j.m < 0 || j.n < 0 || j.lda < std::max(1, j.m) || j.ldq < std::max(1, j.m))
                throw Error(BUG_EINVAL, "bad pivoted qr job dims");
            if (j.n > 16384) throw Error(BUG_EINVAL, "pivoted qr block wider than 16384");
            if (j.m > 0 && j.n > 0 &&
                (j.a_off + (int64_t)(j.n - 1) * j.lda + j.m > a_nelem ||
                 j.q_off + (int64_t)(k - 1) * j.ldq + j.m > q_nelem || j.r_off + k > rdiag_nelem))
                throw Error(BUG_EINVARIANT, "pivoted qr job outside its buffers");
            list[i] = MatJob{j.a_off, j.q_off, j.r_off, wtot, j.m, j.n, j.lda, j.ldq, 1};
            wtot += k;
            nmax = std::max(nmax, j.n);
        }
        hipStream_t st = as_stream(stream);
        MatJob* d = upload_list(ctx, st, list.data(), list.size(), 2);
        double2* tau = static_cast<double2*>(scratch(ctx, st, sizeof(double2) * std::max<int64_t>(wtot, 1), 1));
        const size_t lds = sizeof(double) * (size_t)nmax;
        if (lds > 65536)
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(qrp_kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(qrp_kernel, dim3(nmat), dim3(kQrThreads), lds, st, d,
                           static_cast<double2*>(a), static_cast<double2*>(q), rdiag, tau);
        HIP_CHECK(hipGetLastError());
    });
}

int bug_set_jacobi_params(bug_ctx* ctx, int32_t cols, int32_t inner_sweeps) {
    return guarded(ctx, [&] {
        if (cols < 0 || cols > kBjMaxCols || inner_sweeps < 1 || inner_sweeps > 8)
            throw Error(BUG_EINVAL, "jacobi params: cols in 0..32, inner sweeps in 1..8");
        ctx->bj_cols = cols;
        ctx->bj_inner = inner_sweeps;
    });
}

int bug_set_gram_jacobi_params(bug_ctx* ctx, int32_t cols, int32_t inner_sweeps, int32_t min_n) {
    return guarded(ctx, [&] {
        if (cols < 2 || cols > kGbjColsWide || inner_sweeps < 1 || inner_sweeps > 8 || min_n < 0)
            throw Error(BUG_EINVAL, "gram jacobi params: cols in 2..64, inner sweeps in 1..8, min_n >= 0");
        ctx->gbj_cols = cols & ~1;
        ctx->gbj_inner = inner_sweeps;
        ctx->gbj_min_n = min_n;
    });
}

int bug_set_mixed_jacobi_params(bug_ctx* ctx, int32_t fp32_sweeps, int32_t min_n) {
    return guarded(ctx, [&] {
        if (fp32_sweeps < 1 || fp32_sweeps > 60 || min_n < 0)
            throw Error(BUG_EINVAL, "mixed jacobi params: FP32 sweeps in 1..60, min_n >= 0");
        ctx->mixed_fp32_sweeps = fp32_sweeps;
        ctx->mixed_min_n = min_n;
    });
}

int bug_set_qr_variant(bug_ctx* ctx, int32_t variant) {
    return guarded(ctx, [&] {
        if (variant < 0 || variant > 2) throw Error(BUG_EINVAL, "qr variant must be 0, 1 or 2");
        ctx->qr_variant = variant;
    });
}

int bug_block_svd(bug_ctx* ctx, int32_t nmat, const bug_svd_job* jobs, void* g, int64_t g_nelem,
                  void* u, int64_t u_nelem, double* s, int64_t s_nelem, void* vh, int64_t vh_nelem,
                  int32_t max_sweeps, void* stream) {
    return guarded(ctx, [&] {
        if (nmat < 0 || (nmat > 0 && (!jobs || !g || !u || !s || !vh)))
            throw Error(BUG_EINVAL, "bad svd jobs");
        if (ctx->svd_sweeps_dev && nmat > 65536)
            throw Error(BUG_EINVAL, "sweep recording holds 65536 matrices; the batch has more");
        if (nmat == 0) return;
        std::vector<SvdJob> list(nmat);
        int64_t vtot = 0;
        int nmax = 1;
        for (int i = 0; i < nmat; ++i) {
            const bug_svd_job& j = jobs[i];
            if (j.m < j.n || j.n < 0 || j.ldg < std::max(1, j.m) || j.ldu < std::max(1, j.m) ||
                j.ldvh < std::max(1, j.n))
                throw Error(BUG_EINVAL, "svd job needs m >= n and valid leading dimensions");
            if (j.n > 4096) throw Error(BUG_EINVAL, "svd block wider than 4096");
            if (j.n > 0 &&
                (j.g_off + (int64_t)(j.n - 1) * j.ldg + j.m > g_nelem ||
                 j.u_off + (int64_t)(j.n - 1) * j.ldu + j.m > u_nelem ||
                 j.s_off + j.n > s_nelem || j.vh_off + (int64_t)(j.n - 1) * j.ldvh + j.n > vh_nelem))
                throw Error(BUG_EINVARIANT, "svd job outside its buffers");
            list[i] = SvdJob{j.g_off, vtot, j.u_off, j.s_off, j.vh_off, j.m, j.n, j.ldg, j.n,
                             j.ldu, j.ldvh, j.phase_on_v};
            vtot += (int64_t)j.n * j.n;
            nmax = std::max(nmax, j.n);
        }
        hipStream_t st = as_stream(stream);
        SvdJob* d = upload_list(ctx, st, list.data(), list.size(), 2);
        double2* v = static_cast<double2*>(scratch(ctx, st, sizeof(double2) * std::max<int64_t>(vtot, 1), 1));
        const int sweeps = max_sweeps > 0 ? max_sweeps : 60;
        int* sw = ctx->svd_sweeps_dev;
        bool precond = ctx->svd_variant != 1;
        if (ctx->svd_variant == 0) {
            precond = false;
            for (const SvdJob& j : list) precond |= j.n >= kPrecondMinCols;
        }
        if (!precond) {
            const size_t lds = (size_t)nmax * (sizeof(double) + sizeof(int));
            hipLaunchKernelGGL(svd_jacobi_kernel, dim3(nmat), dim3(kSvdThreads), lds, st, d,
                               static_cast<double2*>(g), v, static_cast<double2*>(u), s,
                               static_cast<double2*>(vh), sweeps, sw);
        } else {
            // QR-preconditioned one-sided Jacobi: G = Q R, Jacobi on X = R^H (n x n):
            // X V_x = U_x S  =>  G = (Q V_x) S U_x^H, so U = Q V_x and Vh = U_x^H.
            int64_t qtot = 0, ntot = 0;
            for (const SvdJob& j : list) {
                qtot += (int64_t)j.m * j.n;
                ntot += (int64_t)j.n * j.n;
            }
            double2* ws = static_cast<double2*>(
                scratch(ctx, st, sizeof(double2) * std::max<int64_t>(qtot + 4 * ntot + ntot, 1), 4));
            double2* Qw = ws;
            double2* Rw = Qw + qtot;
            double2* Xw = Rw + ntot;
            double2* Uxw = Xw + ntot;
            double2* Vxw = Uxw + ntot;
            double2* tau = Vxw + ntot;
            std::vector<MatJob> qj(nmat);
            std::vector<SvdJob> jj(nmat);
            std::vector<CopyOp> xc(nmat), vc(nmat);
            int64_t qo = 0, no = 0, to = 0;
            for (int i = 0; i < nmat; ++i) {
                const SvdJob& j = list[i];
                qj[i] = MatJob{j.g_off, qo, no, to, j.m, j.n, j.ldg, std::max(1, j.m), std::max(1, j.n)};
                xc[i] = CopyOp{no, no, std::max(1, j.n), std::max(1, j.n), j.n, j.n, 2, 0, 0};
                jj[i] = SvdJob{no, j.v_off, no, j.s_off, no, j.n, j.n, std::max(1, j.n), j.n,
                               std::max(1, j.n), std::max(1, j.n), 0};
                vc[i] = CopyOp{no, j.vh_off, std::max(1, j.n), j.ldvh, j.n, j.n, 2, 0, 0};
                qo += (int64_t)j.m * j.n;
                no += (int64_t)j.n * j.n;
                to += j.n;
            }
            if (use_blocked_qr(ctx, qj)) {
                blocked_qr(ctx, st, qj, static_cast<double2*>(g), Qw, Rw);
            } else {
                MatJob* dq = upload_list(ctx, st, qj.data(), qj.size(), 2);
                hipLaunchKernelGGL(qr_kernel, dim3(nmat), dim3(kQrThreads), 0, st, dq,
                                   static_cast<double2*>(g), Qw, Rw, tau);
            }
            HIP_CHECK(hipGetLastError());
            CopyOp* dx = upload_list(ctx, st, xc.data(), xc.size(), 2);
            hipLaunchKernelGGL(copy_blocks_kernel, dim3(nmat, 8), dim3(256), 0, st, dx, Rw, Xw,
                               nullptr, 0);
            HIP_CHECK(hipGetLastError());
            const size_t lds = (size_t)nmax * (sizeof(double) + sizeof(int));
            const bool fused = ctx->svd_variant == 6 ||
                               (ctx->svd_variant == 0 && nmax >= ctx->fbj_min_n);
            const bool gram = !fused && (ctx->svd_variant == 4 ||
                                         (ctx->svd_variant == 0 && nmax >= ctx->gbj_min_n));
            if (!fused && !gram && (ctx->svd_variant == 2 || !block_jacobi_fits(ctx, nmax))) {
                SvdJob* dj = upload_list(ctx, st, jj.data(), jj.size(), 2);
                hipLaunchKernelGGL(svd_jacobi_kernel, dim3(nmat), dim3(kSvdThreads), lds, st, dj,
                                   Xw, v, Uxw, s, Vxw, sweeps, sw);
            } else {
                const bool mixed = !fused && !gram && (ctx->svd_variant == 5 ||
                                             (ctx->svd_variant == 0 && nmax >= ctx->mixed_min_n));
                if (fused)
                    fused_block_jacobi(ctx, st, list, Xw, v, sweeps, sw);
                else if (gram)
                    gram_block_jacobi(ctx, st, list, Xw, v, sweeps, sw);
                else if (mixed)
                    mixed_block_jacobi(ctx, st, list, Xw, v, sweeps, sw);
                else
                    block_jacobi(ctx, st, list, Xw, v, sweeps, sw);
                SvdJob* dj = upload_list(ctx, st, jj.data(), jj.size(), 2);
                hipLaunchKernelGGL(svd_finish_kernel, dim3(nmat), dim3(kSvdThreads), lds, st, dj,
                                   Xw, v, Uxw, s, Vxw);
            }
            HIP_CHECK(hipGetLastError());
            // U = Q Vx  with Vx = (Vh_x)^H, one grouped-GEMM launch over the batch
            Hasher h;
            h.pod((uint64_t)0x5F0D);
            for (const SvdJob& j : list) {
                h.pod(j.m);
                h.pod(j.n);
                h.pod(j.u_off);
                h.pod(j.ldu);
            }
            launch_cached_plan(ctx, h.h, [&](Plan& plan) {
                int64_t qo2 = 0, no2 = 0;
                for (int i = 0; i < nmat; ++i) {
                    const SvdJob& j = list[i];
                    if (j.m > 0 && j.n > 0) {
                        plan.tasks.push_back(GemmTask{j.u_off, j.ldu, j.m, j.n,
                                                      (int32_t)plan.terms.size(),
                                                      (int32_t)plan.terms.size() + 1, 0});
                        plan.terms.push_back(GemmTerm{qo2, no2, std::max(1, j.m), std::max(1, j.n),
                                                      j.n, BUG_OP_N | (BUG_OP_C << 2), 1.0, 0.0});
                        plan.flops += 8.0 * j.m * j.n * (double)j.n;
                    }
                    qo2 += (int64_t)j.m * j.n;
                    no2 += (int64_t)j.n * j.n;
                }
            }, BUG_PREC_FP64, Qw, Vxw, u, st);
            CopyOp* dv = upload_list(ctx, st, vc.data(), vc.size(), 2);
            hipLaunchKernelGGL(copy_blocks_kernel, dim3(nmat, 8), dim3(256), 0, st, dv, Uxw,
                               static_cast<double2*>(vh), nullptr, 0);
            HIP_CHECK(hipGetLastError());
            SvdJob* dp = upload_list(ctx, st, list.data(), list.size(), 2);
            const int pcg = std::min(64, std::max(1, (nmax + 7) / 8));
            hipLaunchKernelGGL(svd_phase_kernel, dim3(nmat, pcg), dim3(512), 0, st, dp,
                               static_cast<double2*>(u), static_cast<double2*>(vh));
        }
        HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"

extern "C" {

int bug_set_svd_variant(bug_ctx* ctx, int32_t variant, int32_t record_sweeps) {
    return guarded(ctx, [&] {
        if (variant < 0 || variant > 6) throw Error(BUG_EINVAL, "svd variant must be 0 .. 6");
        ctx->svd_variant = variant;
        if (record_sweeps && !ctx->svd_sweeps_dev) {
            HIP_CHECK(hipMalloc(&ctx->svd_sweeps_dev, sizeof(int) * 65536));
            HIP_CHECK(hipMemset(ctx->svd_sweeps_dev, 0, sizeof(int) * 65536));
        } else if (!record_sweeps && ctx->svd_sweeps_dev) {
            HIP_CHECK(hipDeviceSynchronize());
            HIP_CHECK(hipFree(ctx->svd_sweeps_dev));
            ctx->svd_sweeps_dev = nullptr;
        }
    });
}

int bug_set_svd_floor(bug_ctx* ctx, double rel) {
    return guarded(ctx, [&] {
        if (!(rel >= 0.0 && rel < 1e-6)) throw Error(BUG_EINVAL, "svd floor must lie in [0, 1e-6)");
        ctx->svd_floor = rel;
    });
}

int bug_svd_sweeps(bug_ctx* ctx, int32_t n, int32_t* out) {
    return guarded(ctx, [&] {
        if (n < 0 || n > 65536 || (n > 0 && !out)) throw Error(BUG_EINVAL, "bad sweep buffer");
        if (!ctx->svd_sweeps_dev) throw Error(BUG_EINVAL, "sweep recording is off");
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(out, ctx->svd_sweeps_dev, sizeof(int) * n, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
