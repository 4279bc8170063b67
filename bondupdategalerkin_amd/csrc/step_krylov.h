// Krylov pieces of the step-level entry point (see step_krylov.hip).
#pragma once

#include <functional>

#include "step_core.h"

namespace bug {
namespace host {

constexpr double kBreakdown = 1e-13;      // invariant-subspace threshold of the local chains
constexpr int kMaxDepth = 64;             // local Krylov depth bound
constexpr int kRootMaxIter = 30;          // root basis per substep
constexpr double kSkewTol = 1e-12;        // |M + M^dag| below which exp(M) counts as unitary
constexpr int kMaxPanels = 4096;          // Simpson panels of one residual integral
constexpr int kTailNodes = 8;             // nodes of the real-time tail guard
constexpr double kShiftTrigger = 30.0;    // exponent past which the spectral shift switches on
constexpr double kMaxNorm = 1e150;

// smallest rank <= kmax whose scaled discarded tail fits; returns the scaled tail
double tail_rank(const std::vector<SvEntry>& kept, double scale, double allowance, int kmax,
                 double base, int& rank);
// |tau| int_0^1 |(exp(sM) e1)[index]| ds (composite Simpson in log space), M row-major n x n
double residual_integral(int index, const std::vector<cplx>& M, int n, cplx tau, bool relative);
// weight of the (k+1)-th Krylov vector in exp(tau H) v1, H row-major with leading dimension ld
double direction_weight(const std::vector<cplx>& H, int ld, int k, cplx tau, bool state_seed);

// The root integrator (root_methods.jl:5-71): the exponential, or the stability function
// R = N / D of one implicit Runge-Kutta step applied in one stage; `shift` moves a rational root's
// spectrum to the extreme Ritz value (and drops the scale) when tau has a real part.
struct RootRule {
    int method = BUG_ROOT_EXP;
    bool shift = true;
    bool rational() const { return method != BUG_ROOT_EXP; }
};
// R(z) of a rational method
cplx root_stability(int method, cplx z);
// R(M) rhs = D(M)^{-1} N(M) rhs for a small dense M (row-major n x n)
std::vector<cplx> rational_apply(int method, const std::vector<cplx>& M, int n, const std::vector<cplx>& rhs);

struct Tridiagonal {
    int n = 0;
    std::vector<double> vals, vecs;       // vecs column-major
    RootRule rule;
    Tridiagonal(const std::vector<double>& alpha, const std::vector<double>& beta, RootRule r = RootRule());
    std::vector<cplx> coeffs(cplx tau, cplx& logscale) const;
};

struct Hessenberg {
    std::vector<cplx> Hm;                 // row-major n x n
    int n = 0;
    std::vector<cplx> vals;
    RootRule rule;
    Hessenberg(const std::vector<cplx>& H, int n, RootRule r = RootRule());
    std::vector<cplx> coeffs(cplx tau, cplx& logscale) const;
};

struct ExpvReport {
    int iterations = 0;
    int substeps = 0;
    double err = 0.0;
};

// K independent vectors packed in one device buffer of n elements: instance i occupies
// [table[2i], table[2i] + table[2i+1]).
struct PackedSegs {
    std::vector<int64_t> table;
    int64_t n = 0;
    int K() const { return (int)(table.size() / 2); }
};
// y_i = A_i x_i for every instance, on packed buffers of PackedSegs::n elements
using PackedApplyFn = std::function<void(const double2* x, double2* y)>;

// exp(tau_i A_i) v0_i for K instances advanced in lockstep (lanczos_expv / arnoldi_expv per instance,
// their Krylov passes side by side): one packed apply, one segmented multi-dot and one segmented
// multi-axpy per recurrence step for all instances; each instance keeps its own substep search,
// stopping rule and scale bookkeeping, and the host-side k x k work runs on `host_threads` threads.
void expv_lockstep(const Exec& ex, const PackedSegs& sg, const PackedApplyFn& apply,
                   const std::vector<cplx>& tau, bool hermitian, const double2* v0, double tol,
                   int maxiter, bool reorth, double conv_tol, double2* out,
                   std::vector<ExpvReport>& reps, int host_threads, RootRule rule = RootRule());

// a rational `rule` takes the whole step in one stage (krylov.jl:153-160,365-370)
Tensor lanczos_expv(const Exec& ex, const std::function<Tensor(const Tensor&)>& apply, cplx tau,
                    const Tensor& v0, double tol, int maxiter, bool reorth, double conv_tol,
                    ExpvReport& rep, RootRule rule = RootRule());
Tensor arnoldi_expv(const Exec& ex, const std::function<Tensor(const Tensor&)>& apply, cplx tau,
                    const Tensor& v0, double tol, int maxiter, double conv_tol, ExpvReport& rep,
                    RootRule rule = RootRule());

}  // namespace host
}  // namespace bug
