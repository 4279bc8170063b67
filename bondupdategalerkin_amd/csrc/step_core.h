// Host side of the step-level entry point: block-sparse tensors resident on the device, the tensor
// algebra the step needs (built on the library's own C entry points), the probe stream and the
// small dense solvers.  Everything here is host code; the arithmetic on tensors runs in the
// library's kernels on the step's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bug.h"

namespace bug {
namespace host {

using cplx = std::complex<double>;

// A failed entry point or an invariant the step relies on; carries the library status code.
struct StepError : std::runtime_error {
    int code;
    StepError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// a HIP runtime call whose failure ends the step with BUG_EDEVICE
#define STEP_HIP(expr)                                                                     \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            throw StepError(BUG_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)


// ---- legs and layouts (the block structure of a family of shifted block-diagonal matrices) ----

struct Leg {
    std::vector<int32_t> q, d;            // charges strictly increasing, dims > 0
    int nsec() const { return (int)q.size(); }
    int total() const;
    int index(int c) const;               // sector of charge c, -1 if absent
    int dim_of(int c) const;
    // (charge, dim) pairs: equal charges summed, empty sectors dropped, sorted by charge
    static Leg make(const std::vector<std::pair<int, int>>& pairs);
    static Leg dense(int n);
    bool operator==(const Leg& o) const { return q == o.q && d == o.d; }
    bool operator!=(const Leg& o) const { return !(*this == o); }
    std::map<int, int> as_map() const;
};

// One stored block: slice k, row charge q, column charge qc = q + shift[k], d x dc at `off`.
struct Blk {
    int k, q, qc, d, dc;
    int64_t off;
};

struct Layout {
    Leg row, col;
    std::vector<int32_t> shifts;
    std::vector<int64_t> off;             // [nslice * row.nsec], -1 = absent
    std::vector<Blk> blks;                // slice by slice, row sectors in charge order
    int64_t nelem = 0;
    Layout() = default;
    Layout(Leg r, Leg c, std::vector<int32_t> sh);
    int nslice() const { return (int)shifts.size(); }
    const Blk* find(int k, int q) const;  // the block at (slice k, row charge q) or null
    bool operator==(const Layout& o) const {
        return row == o.row && col == o.col && shifts == o.shifts;
    }
    bool operator!=(const Layout& o) const { return !(*this == o); }
};
using LayoutP = std::shared_ptr<const Layout>;

// ---- device memory: a caching allocator whose free lists are per stream ----------------------

// The stream the calling thread's step work runs on; blocks freed by this thread go back to that
// stream's free list and are handed out again only to allocations on the same stream.
void set_thread_stream(hipStream_t s);
hipStream_t thread_stream();
void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    explicit DevBuf(size_t b);
    ~DevBuf();
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// A block-sparse tensor: a layout plus a complex FP64 device buffer of at least nelem elements.
struct Tensor {
    LayoutP lay;
    std::shared_ptr<DevBuf> buf;
    bool empty() const { return !lay; }
    const Layout& L() const { return *lay; }
    double2* data() const { return static_cast<double2*>(buf->p); }
    int64_t nelem() const { return lay->nelem; }
    int64_t capacity() const { return (int64_t)(buf->bytes / sizeof(double2)); }
    bug_bsm desc() const;
};

// ---- the execution context of the step on one thread ---------------------------------------

struct Exec {
    bug_ctx* ctx;
    hipStream_t st;
    int prec;                              // contraction precision (BUG_PREC_*)
    void check(int rc) const;
};

enum Side { LEFT = 0, RIGHT = 1 };

// MPO site in the form the library consumes (bug_mpo_site with owned arrays).
struct MpoSite {
    std::vector<int32_t> cl, cr, a, b, sout, sin;
    std::vector<double> coeff;            // 2 nent
    int nent() const { return (int)a.size(); }
    bug_mpo_site desc() const;
};

// ---- tensor algebra ---------------------------------------------------------------------------

Tensor t_empty(const Exec& ex, const Layout& lay);
Tensor t_zeros(const Exec& ex, const Layout& lay);
Tensor t_clone(const Exec& ex, const Tensor& t);
Tensor t_from_host(const Exec& ex, const Layout& lay, const std::vector<cplx>& host);
std::vector<cplx> t_to_host(const Exec& ex, const Tensor& t);

struct Term {
    int o, p, q, op_a, op_b;
    cplx alpha;
};
Tensor product(const Exec& ex, const Leg& row, const Leg& col, const std::vector<int32_t>& shifts,
               const Tensor& P, const Tensor& Q, const std::vector<Term>& terms,
               int prec = BUG_PREC_FP64);
void product_into(const Exec& ex, const Tensor& out, const Tensor& P, const Tensor& Q,
                  const std::vector<Term>& terms, int prec = BUG_PREC_FP64);
Tensor core_times_site(const Exec& ex, const Tensor& C, const Tensor& A);
Tensor site_times_core(const Exec& ex, const Tensor& A, const Tensor& C);
Tensor overlap_left(const Exec& ex, const Tensor& U, const Tensor& A);
Tensor overlap_right(const Exec& ex, const Tensor& A, const Tensor& V);
Tensor core_times_core(const Exec& ex, const Tensor& A, const Tensor& B);
Tensor perp_left(const Exec& ex, const Tensor& U, const Tensor& K);
Tensor perp_right(const Exec& ex, const Tensor& V, const Tensor& K);
inline Tensor perp(const Exec& ex, Side s, const Tensor& F, const Tensor& K) {
    return s == LEFT ? perp_left(ex, F, K) : perp_right(ex, F, K);
}

std::vector<cplx> inner_many(const Exec& ex, const std::vector<const Tensor*>& xs, const Tensor& y);
cplx inner(const Exec& ex, const Tensor& a, const Tensor& b);
double norm(const Exec& ex, const Tensor& a);
void axpy_(const Exec& ex, cplx alpha, const Tensor& x, const Tensor& y);
void axpy_many_(const Exec& ex, const std::vector<cplx>& alphas,
                const std::vector<const Tensor*>& xs, const Tensor& y);
void scale_(const Exec& ex, cplx alpha, const Tensor& x);
Tensor scale(const Exec& ex, cplx alpha, const Tensor& x);

struct CopyList {
    std::vector<bug_copy> ops;
    void add(int64_t src_off, int src_ld, int64_t dst_off, int dst_ld, int rows, int cols,
             int mode = 0, int scale_mode = 0, int64_t scale_off = 0);
    void run(const Exec& ex, const void* src, int64_t src_n, void* dst, int64_t dst_n,
             const double* scale_dev = nullptr, bool accumulate = false) const;
};
// squared Frobenius norms of (off, ld, rows, cols) blocks of src
std::vector<double> block_sqnorms(const Exec& ex, const std::vector<bug_copy>& blocks,
                                  const void* src, int64_t src_n);

// matricisation sectors: left  qr -> [(s, ql, rows, off)], right  ql -> [(s, qr, cols, off)],
// both in charge order of the kept leg and slice order within
struct SecPart {
    int s, qo, n;
    int64_t off;
};
std::vector<std::pair<int, std::vector<SecPart>>> left_sectors(const Layout& lay);
std::vector<std::pair<int, std::vector<SecPart>>> right_sectors(const Layout& lay);

Tensor oplus(const Exec& ex, Side s, const Tensor& F, const Tensor& N);
// keep the leading counts[q] columns (LEFT) or rows (RIGHT) of each sector; empty when none
Tensor trim(const Exec& ex, const Tensor& T, Side s, const std::map<int, int>& counts);
// S T (rows) or T S (cols) with real diagonal scales per sector
Tensor diag_scale(const Exec& ex, const Tensor& T, const std::map<int, std::vector<double>>& S,
                  bool rows);

// unique-gauge QR (real non-negative R diagonal), roundoff-level sectors dropped
void qr_left(const Exec& ex, const Tensor& A, Tensor& Q, Tensor& R);
void qr_right(const Exec& ex, const Tensor& A, Tensor& Lc, Tensor& Q);
void qr_core(const Exec& ex, const Tensor& C, Tensor& Q, Tensor& R);

// SVD: a factorization held on the device, any number of selections read from it
struct SvEntry {
    double s;
    int q, j;
};
// Telum's global selection over per-sector spectra (sector i tagged tags[i]): `full` receives every
// value sorted descending (stable, in sector order); returns how many of them are kept
int global_selection(const std::vector<std::vector<double>>& spectra, const std::vector<int>& tags,
                     double cutoff, int nkeep, std::vector<SvEntry>& full);
struct SVDResult {
    Tensor U, Vh;
    std::map<int, std::vector<double>> S;
    std::vector<SvEntry> kept, trunc;
};
struct BlockSVD {
    int kind;                                           // 0 left, 1 right, 2 core
    LayoutP lay;
    std::vector<int> qlist;
    std::vector<std::vector<SecPart>> secs;             // per qlist entry (left/right)
    std::vector<std::pair<int, int>> mats;              // (m, n) per qlist entry
    std::vector<char> tall;
    std::vector<int64_t> uo, so, vo;
    std::shared_ptr<DevBuf> U, Vh;
    std::vector<std::vector<double>> spectra;           // per qlist entry, descending
    SVDResult select(const Exec& ex, double cutoff, int nkeep = -1) const;
};
BlockSVD factor_left(const Exec& ex, const Tensor& A);
BlockSVD factor_right(const Exec& ex, const Tensor& A);
BlockSVD factor_core(const Exec& ex, const Tensor& C);
inline BlockSVD factor(const Exec& ex, Side s, const Tensor& A) {
    return s == LEFT ? factor_left(ex, A) : factor_right(ex, A);
}
double sq_weight(const std::vector<SvEntry>& l);
double trunc_weight(const SVDResult& r);

// environments and effective operators
Layout env_layout(const Leg& bra, const Leg& ket, const std::vector<int32_t>& channel_charges);
Tensor boundary_left(const Exec& ex, const MpoSite& W0, const Leg& leg);
Tensor boundary_right(const Exec& ex, const MpoSite& Wlast, const Leg& leg);
Tensor push_left(const Exec& ex, const Tensor& L, const Tensor& A, const MpoSite& W, int prec);
Tensor push_right(const Exec& ex, const Tensor& R, const Tensor& B, const MpoSite& W, int prec);
Tensor apply_one_site(const Exec& ex, const Tensor& L, const MpoSite& W, const Tensor& R,
                      const Tensor& X, int prec);
Tensor apply_zero_site(const Exec& ex, const Tensor& L, const Tensor& R, const Tensor& C, int prec);

// ---- the probe stream (counter based: splitmix64 words, Box-Muller, E|z|^2 = 1) --------------

struct ProbeStream {
    uint64_t key = 0;
    uint64_t pos = 0;
    explicit ProbeStream(uint64_t seed);
    cplx at(uint64_t i) const;
    void skip(uint64_t n) { pos += n; }
    ProbeStream split();
};

// ---- small dense solvers on the host ----------------------------------------------------------

// eigenvalues (ascending) and orthonormal eigenvectors (column-major n x n) of a real symmetric
// tridiagonal matrix (implicit QL with Wilkinson shifts)
void tridiag_eigen(const std::vector<double>& diag, const std::vector<double>& off,
                   std::vector<double>& vals, std::vector<double>& vecs);
// exp(A) of a complex n x n matrix (row-major), scaling and squaring with the degree-13 Pade
// approximant (Higham 2005)
std::vector<cplx> expm(const std::vector<cplx>& A, int n);
// eigenvalues of a complex n x n matrix (row-major): Hessenberg reduction, shifted QR
std::vector<cplx> eigvals(std::vector<cplx> A, int n);

}  // namespace host
}  // namespace bug
