// Krylov pieces of the step-level entry point: the host k x k side (tridiagonal and Hessenberg
// exponentials, the residual integral that weighs a local Krylov direction, the tail estimate and
// substep search) and the device recurrences (Lanczos and Arnoldi exponentials) on tensors.
//
// Reference: lanczos_expv / arnoldi_expv and their helpers, src/BUG/Core/LocalSolvers/krylov.jl:
// 44-117,128-263,296-442; the substep search and tail estimate, root_methods.jl:86-141; the
// Krylov direction weights, src/BUG/Core/RankAdaptation/krylov_enrichment.jl:58-120; Test A's tail
// rank, truncation.jl:262-278.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <limits>
#include <mutex>
#include <thread>

#include "step_krylov.h"

namespace bug {
namespace host {

double tail_rank(const std::vector<SvEntry>& kept, double scale, double allowance, int kmax,
                 double base, int& rank) {
    kmax = std::min<int>(kmax, (int)kept.size());
    std::vector<double> w(kept.size());
    for (size_t i = 0; i < kept.size(); ++i) w[i] = kept[i].s * kept[i].s;
    if (kmax <= 0) {
        double t = base;
        for (double x : w) t += x;
        rank = 0;
        return scale * std::sqrt(t);
    }
    double tail = base;
    for (size_t i = kmax; i < w.size(); ++i) tail += w[i];
    int r = kmax;
    while (r > 0 && scale * std::sqrt(tail + w[r - 1]) <= allowance) {
        tail += w[r - 1];
        --r;
    }
    rank = r;
    return scale * std::sqrt(tail);
}

double residual_integral(int index, const std::vector<cplx>& M, int n, cplx tau, bool relative) {
    double nm = 0.0, ns = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            nm += std::norm(M[(size_t)i * n + j]);
            ns += std::norm(M[(size_t)i * n + j] + std::conj(M[(size_t)j * n + i]));
        }
    const bool unitary = std::sqrt(ns) <= kSkewTol * std::sqrt(nm);
    double sigma = 0.0;
    if (!unitary) {
        const auto ev = eigvals(M, n);
        sigma = -std::numeric_limits<double>::infinity();
        for (const cplx& e : ev) sigma = std::max(sigma, e.real());
    }
    std::vector<cplx> A(M);
    for (int i = 0; i < n; ++i) A[(size_t)i * n + i] -= sigma;
    double colmax = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::abs(A[(size_t)i * n + j]);
        colmax = std::max(colmax, s);
    }
    const int panels = 2 * std::min(std::max((int)std::ceil(2.0 * colmax), 8), kMaxPanels);
    const double h = 1.0 / panels;
    std::vector<cplx> hA(A);
    for (auto& v : hA) v *= h;
    const std::vector<cplx> step = expm(hA, n);
    const int B = panels >= 64 ? 64 : 1;
    // row `index` of S^1 .. S^B and S^B by squaring
    std::vector<cplx> rows((size_t)B * n);
    for (int j = 0; j < n; ++j) rows[j] = step[(size_t)index * n + j];
    for (int b = 1; b < B; ++b)
        for (int j = 0; j < n; ++j) {
            cplx s = 0.0;
            for (int k = 0; k < n; ++k) s += rows[(size_t)(b - 1) * n + k] * step[(size_t)k * n + j];
            rows[(size_t)b * n + j] = s;
        }
    auto mm = [&](const std::vector<cplx>& X, const std::vector<cplx>& Y) {
        std::vector<cplx> Z((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                const cplx x = X[(size_t)i * n + k];
                for (int j = 0; j < n; ++j) Z[(size_t)i * n + j] += x * Y[(size_t)k * n + j];
            }
        return Z;
    };
    std::vector<cplx> big = step;
    int sq = 0;
    for (int t = B; t > 1; t >>= 1) ++sq;
    for (int t = 0; t < sq; ++t) big = mm(big, big);
    auto mv = [&](const std::vector<cplx>& X, const std::vector<cplx>& v) {
        std::vector<cplx> y(n, 0.0);
        for (int i = 0; i < n; ++i) {
            cplx s = 0.0;
            for (int k = 0; k < n; ++k) s += X[(size_t)i * n + k] * v[k];
            y[i] = s;
        }
        return y;
    };
    std::vector<double> mag(panels + 1);
    std::vector<cplx> c(n, 0.0);
    c[0] = 1.0;
    mag[0] = std::abs(c[index]);
    int j = 0;
    while (j + B <= panels) {
        for (int b = 0; b < B; ++b) {
            cplx s = 0.0;
            for (int k = 0; k < n; ++k) s += rows[(size_t)b * n + k] * c[k];
            mag[j + 1 + b] = std::abs(s);
        }
        c = mv(big, c);
        j += B;
    }
    while (j < panels) {
        c = mv(step, c);
        ++j;
        mag[j] = std::abs(c[index]);
    }
    std::vector<double> logw(panels + 1);
    double top = -std::numeric_limits<double>::infinity();
    for (int i = 0; i <= panels; ++i) {
        const double w = (i == 0 || i == panels) ? 1.0 : (i % 2 == 1 ? 4.0 : 2.0);
        logw[i] = mag[i] == 0.0 ? -std::numeric_limits<double>::infinity()
                                : std::log(w * (h / 3.0) * mag[i]) + i * h * sigma;
        top = std::max(top, logw[i]);
    }
    if (top == -std::numeric_limits<double>::infinity()) return 0.0;
    double sum = 0.0;
    for (double lw : logw) sum += std::exp(lw - top);
    const double lw = std::log(std::abs(tau)) + top + std::log(sum);
    if (unitary || !relative) return std::exp(lw);
    double cn = 0.0;
    for (const cplx& v : c) cn += std::norm(v);
    cn = std::sqrt(cn);
    return std::exp(lw - sigma - std::log(std::max(cn, std::numeric_limits<double>::min())));
}

double direction_weight(const std::vector<cplx>& H, int ld, int k, cplx tau, bool state_seed) {
    const double beta = std::abs(H[(size_t)k * ld + (k - 1)]);
    if (beta == 0.0) return 0.0;
    std::vector<cplx> M((size_t)k * k);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) M[(size_t)i * k + j] = tau * H[(size_t)i * ld + j];
    return beta * residual_integral(k - 1, M, k, tau, state_seed);
}

// ---- small propagators ----------------------------------------------------------------------------

// ---- the rational roots (root_methods.jl:5-71,143-157) --------------------------------------------

namespace {
// numerator / denominator coefficients of each method's stability function, constant term first
struct Pade {
    std::vector<double> num, den;
};
const Pade& pade_of(int method) {
    static const Pade trap{{1.0, 1.0 / 2}, {1.0, -1.0 / 2}};
    static const Pade rad2{{1.0, 1.0 / 3}, {1.0, -2.0 / 3, 1.0 / 6}};
    static const Pade rad3{{1.0, 2.0 / 5, 1.0 / 20}, {1.0, -3.0 / 5, 3.0 / 20, -1.0 / 60}};
    switch (method) {
        case BUG_ROOT_TRAPEZOID: return trap;
        case BUG_ROOT_RADAU2: return rad2;
        case BUG_ROOT_RADAU3: return rad3;
        default: throw StepError(BUG_EINVAL, "not a rational root method");
    }
}

cplx poly_at(const std::vector<double>& c, cplx z) {
    cplx out = c.back();
    for (int k = (int)c.size() - 2; k >= 0; --k) out = out * z + c[k];
    return out;
}

// the shift of a rational root: the Ritz value maximising Re(tau lambda), none without `shift`
// or in real time (tau purely imaginary)
template <typename V>
cplx rational_sigma(const std::vector<V>& vals, cplx tau, bool shift) {
    if (!shift || vals.empty() || tau.real() == 0.0) return 0.0;
    size_t best = 0;
    for (size_t i = 1; i < vals.size(); ++i)
        if ((tau * cplx(vals[i])).real() > (tau * cplx(vals[best])).real()) best = i;
    return cplx(vals[best]);
}
}  // namespace

cplx root_stability(int method, cplx z) {
    if (method == BUG_ROOT_EXP) return std::exp(z);
    const Pade& p = pade_of(method);
    return poly_at(p.num, z) / poly_at(p.den, z);
}

std::vector<cplx> rational_apply(int method, const std::vector<cplx>& M, int n, const std::vector<cplx>& rhs) {
    const Pade& p = pade_of(method);
    // P(M) by Horner, row-major
    auto poly = [&](const std::vector<double>& c) {
        std::vector<cplx> out((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) out[(size_t)i * n + i] = c.back();
        for (int k = (int)c.size() - 2; k >= 0; --k) {
            std::vector<cplx> next((size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int l = 0; l < n; ++l) {
                    const cplx a = out[(size_t)i * n + l];
                    if (a == 0.0) continue;
                    for (int j = 0; j < n; ++j) next[(size_t)i * n + j] += a * M[(size_t)l * n + j];
                }
            for (int i = 0; i < n; ++i) next[(size_t)i * n + i] += c[k];
            out.swap(next);
        }
        return out;
    };
    const std::vector<cplx> N = poly(p.num), D = poly(p.den);
    std::vector<cplx> b(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) b[i] += N[(size_t)i * n + j] * rhs[j];
    // D x = b by LU with partial pivoting
    std::vector<cplx> A = D;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(A[(size_t)r * n + c]) > std::abs(A[(size_t)piv * n + c])) piv = r;
        if (piv != c) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)c * n + j], A[(size_t)piv * n + j]);
            std::swap(b[c], b[piv]);
        }
        const cplx d = A[(size_t)c * n + c];
        if (d == 0.0) throw StepError(BUG_ENONFINITE, "singular denominator of the rational root");
        for (int r = c + 1; r < n; ++r) {
            const cplx f = A[(size_t)r * n + c] / d;
            if (f == 0.0) continue;
            for (int j = c; j < n; ++j) A[(size_t)r * n + j] -= f * A[(size_t)c * n + j];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        cplx sum = b[c];
        for (int k = c + 1; k < n; ++k) sum -= A[(size_t)c * n + k] * b[k];
        b[c] = sum / A[(size_t)c * n + c];
    }
    return b;
}

Tridiagonal::Tridiagonal(const std::vector<double>& alpha, const std::vector<double>& beta, RootRule r)
    : rule(r) {
    for (double v : alpha)
        if (!std::isfinite(v))
            throw StepError(BUG_ENONFINITE, "Lanczos tridiagonal has non-finite entries: the input "
                                            "vector or the operator is not finite");
    for (double v : beta)
        if (!std::isfinite(v))
            throw StepError(BUG_ENONFINITE, "Lanczos tridiagonal has non-finite entries: the input "
                                            "vector or the operator is not finite");
    n = (int)alpha.size();
    if (n == 1) {
        vals = alpha;
        vecs = {1.0};
    } else {
        tridiag_eigen(alpha, beta, vals, vecs);
    }
}

std::vector<cplx> Tridiagonal::coeffs(cplx tau, cplx& logscale) const {
    if (rule.rational()) {
        // R(tau (T - sigma)) e1 from the eigendecomposition; no scale returned
        const cplx sigma = rational_sigma(vals, tau, rule.shift);
        std::vector<cplx> w(n), out(n, 0.0);
        for (int k = 0; k < n; ++k)
            w[k] = root_stability(rule.method, tau * (vals[k] - sigma)) * vecs[(size_t)k * n + 0];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) out[i] += vecs[(size_t)k * n + i] * w[k];
        logscale = 0.0;
        return out;
    }
    int imax = 0;
    double exmax = -std::numeric_limits<double>::infinity();
    for (int i = 0; i < n; ++i) {
        const double e = (tau * vals[i]).real();
        if (e > exmax) {
            exmax = e;
            imax = i;
        }
    }
    const double sigma = exmax > kShiftTrigger ? vals[imax] : 0.0;
    std::vector<cplx> w(n);
    for (int k = 0; k < n; ++k) w[k] = std::exp(tau * (vals[k] - sigma)) * vecs[(size_t)k * n + 0];
    std::vector<cplx> out(n, 0.0);
    for (int i = 0; i < n; ++i) {
        cplx s = 0.0;
        for (int k = 0; k < n; ++k) s += vecs[(size_t)k * n + i] * w[k];
        out[i] = s;
    }
    logscale = tau * sigma;
    return out;
}

Hessenberg::Hessenberg(const std::vector<cplx>& H, int n_, RootRule r) : Hm(H), n(n_), rule(r) {
    for (const cplx& v : Hm)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw StepError(BUG_ENONFINITE, "Arnoldi Hessenberg matrix has non-finite entries: the "
                                            "input vector or the operator is not finite");
    if (n > 0) vals = eigvals(Hm, n);
}

std::vector<cplx> Hessenberg::coeffs(cplx tau, cplx& logscale) const {
    logscale = 0.0;
    if (n == 0) return {};
    if (rule.rational()) {
        // R(tau (Hm - sigma)) e1 by one solve with the denominator; no scale returned
        const cplx sigma = rational_sigma(vals, tau, rule.shift);
        std::vector<cplx> M((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                M[(size_t)i * n + j] = tau * (Hm[(size_t)i * n + j] - (i == j ? sigma : 0.0));
        std::vector<cplx> e1(n, 0.0);
        e1[0] = 1.0;
        std::vector<cplx> out = rational_apply(rule.method, M, n, e1);
        for (const cplx& v : out)
            if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
                throw StepError(BUG_ENONFINITE, "the rational root overflowed");
        return out;
    }
    int imax = 0;
    double exmax = -std::numeric_limits<double>::infinity();
    for (int i = 0; i < n; ++i) {
        const double e = (tau * vals[i]).real();
        if (e > exmax) {
            exmax = e;
            imax = i;
        }
    }
    const double sigma = exmax > kShiftTrigger ? vals[imax].real() : 0.0;
    std::vector<cplx> A((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            A[(size_t)i * n + j] = tau * (Hm[(size_t)i * n + j] - (i == j ? sigma : 0.0));
    const std::vector<cplx> P = expm(A, n);
    std::vector<cplx> out(n);
    for (int i = 0; i < n; ++i) {
        out[i] = P[(size_t)i * n + 0];
        if (!std::isfinite(out[i].real()) || !std::isfinite(out[i].imag()))
            throw StepError(BUG_ENONFINITE, "the Hessenberg exponential overflowed");
    }
    logscale = tau * sigma;
    return out;
}

template <typename Prop>
static double tail_sup(const Prop& p, cplx step, double betak, double above) {
    cplx ls;
    double tail = betak * std::abs(p.coeffs(step, ls).back());
    if (step.real() != 0.0) return tail;
    for (int j = 1; j < kTailNodes; ++j) {
        if (tail >= above) return tail;
        tail = std::max(tail, betak * std::abs(p.coeffs(step * ((double)j / kTailNodes), ls).back()));
    }
    return tail;
}

template <typename Prop>
static cplx pick_tau(const Prop& p, cplx tau_rem, double betak, double conv_tol) {
    auto ok = [&](cplx t) {
        const double v = tail_sup(p, t, betak, conv_tol > 0 ? conv_tol : std::numeric_limits<double>::infinity());
        return std::isfinite(v) && (conv_tol <= 0 || v <= conv_tol);
    };
    if (ok(tau_rem)) return tau_rem;
    double lo = 0.0, hi = 1.0;
    for (int i = 0; i < 40; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (ok(mid * tau_rem))
            lo = mid;
        else
            hi = mid;
    }
    return std::max(lo, 1e-6) * tau_rem;
}

// ---- the root exponentials -------------------------------------------------------------------------

static Tensor finish_scale(const Exec& ex, Tensor cur, cplx logscale) {
    const cplx sc = std::exp(logscale);
    if (std::isfinite(std::abs(sc)) && sc != 0.0) {
        const double nrm = norm(ex, cur);
        if (!(std::isfinite(nrm) && nrm > 0 && std::abs(sc) * nrm > kMaxNorm)) cur = scale(ex, sc, cur);
    }
    return cur;
}

Tensor lanczos_expv(const Exec& ex, const std::function<Tensor(const Tensor&)>& apply, cplx tau,
                    const Tensor& v0, double tol, int maxiter, bool reorth, double conv_tol,
                    ExpvReport& rep, RootRule rule) {
    rep = ExpvReport();
    const double beta0 = norm(ex, v0);
    if (beta0 == 0.0) return v0;
    Tensor cur = v0;
    cplx tau_rem = tau, logscale = 0.0;
    while (true) {
        rep.substeps += 1;
        const double cn = norm(ex, cur);
        if (cn == 0.0) break;
        std::vector<Tensor> basis{scale(ex, 1.0 / cn, cur)};
        std::vector<double> alpha, beta;
        Tensor r = apply(basis[0]);
        rep.iterations += 1;
        double a = inner(ex, basis[0], r).real();
        alpha.push_back(a);
        axpy_(ex, -a, basis[0], r);
        for (int it = 2; it <= maxiter; ++it) {
            const double b = norm(ex, r);
            if (b < tol) break;
            if (conv_tol > 0 && tail_sup(Tridiagonal(alpha, beta, rule), tau_rem, b, conv_tol) < conv_tol) break;
            beta.push_back(b);
            basis.push_back(scale(ex, 1.0 / b, r));
            r = apply(basis.back());
            rep.iterations += 1;
            a = inner(ex, basis.back(), r).real();
            alpha.push_back(a);
            axpy_(ex, -a, basis.back(), r);
            axpy_(ex, -b, basis[basis.size() - 2], r);
            if (reorth)
                for (int p = 0; p < 2; ++p) {
                    std::vector<const Tensor*> bp;
                    for (auto& t : basis) bp.push_back(&t);
                    auto cs = inner_many(ex, bp, r);
                    for (auto& c : cs) c = -c;
                    axpy_many_(ex, cs, bp, r);
                }
        }
        const Tridiagonal tri(alpha, beta, rule);
        const double betak = norm(ex, r);
        // an implicit root takes the whole step at once
        const cplx tau_s = rule.rational() ? tau_rem : pick_tau(tri, tau_rem, betak, conv_tol);
        cplx ls;
        const std::vector<cplx> coeff = tri.coeffs(tau_s, ls);
        rep.err += tail_sup(tri, tau_s, betak, std::numeric_limits<double>::infinity()) * cn / beta0;
        Tensor out = scale(ex, coeff[0] * cn, basis[0]);
        const size_t m = std::min(coeff.size(), basis.size());
        std::vector<cplx> cs;
        std::vector<const Tensor*> bp;
        for (size_t i = 1; i < m; ++i) {
            cs.push_back(coeff[i] * cn);
            bp.push_back(&basis[i]);
        }
        axpy_many_(ex, cs, bp, out);
        cur = out;
        logscale += ls;
        tau_rem -= tau_s;
        if (std::abs(tau_rem) <= 1e-14 * std::max(std::abs(tau), 1.0)) break;
        if (rep.substeps >= 10000) throw StepError(BUG_EINVARIANT, "expv: substep search failed to advance");
    }
    return finish_scale(ex, cur, logscale);
}

Tensor arnoldi_expv(const Exec& ex, const std::function<Tensor(const Tensor&)>& apply, cplx tau,
                    const Tensor& v0, double tol, int maxiter, double conv_tol, ExpvReport& rep,
                    RootRule rule) {
    rep = ExpvReport();
    const double beta0 = norm(ex, v0);
    if (beta0 == 0.0) return v0;
    Tensor cur = v0;
    cplx tau_rem = tau, logscale = 0.0;
    const int ld = maxiter;
    auto lead = [&](const std::vector<cplx>& H, int dim) {
        std::vector<cplx> S((size_t)dim * dim);
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j) S[(size_t)i * dim + j] = H[(size_t)i * ld + j];
        return Hessenberg(S, dim, rule);
    };
    while (true) {
        rep.substeps += 1;
        const double cn = norm(ex, cur);
        if (cn == 0.0) break;
        std::vector<Tensor> basis{scale(ex, 1.0 / cn, cur)};
        std::vector<cplx> H((size_t)(maxiter + 1) * ld, 0.0);
        int dim = 0;
        double betak = 0.0;
        for (int k = 1; k <= maxiter; ++k) {
            Tensor r = apply(basis[k - 1]);
            rep.iterations += 1;
            for (int p = 0; p < 2; ++p) {
                std::vector<const Tensor*> bp;
                for (auto& t : basis) bp.push_back(&t);
                auto cs = inner_many(ex, bp, r);
                for (size_t i = 0; i < cs.size(); ++i) H[i * ld + (k - 1)] += cs[i];
                for (auto& c : cs) c = -c;
                axpy_many_(ex, cs, bp, r);
            }
            dim = k;
            betak = norm(ex, r);
            if (betak < tol) break;
            if (conv_tol > 0 && tail_sup(lead(H, dim), tau_rem, betak, conv_tol) < conv_tol) break;
            H[(size_t)k * ld + (k - 1)] = betak;
            if (k == maxiter) break;
            basis.push_back(scale(ex, 1.0 / betak, r));
        }
        const Hessenberg hess = lead(H, dim);
        const cplx tau_s = rule.rational() ? tau_rem : pick_tau(hess, tau_rem, betak, conv_tol);
        cplx ls;
        const std::vector<cplx> coeff = hess.coeffs(tau_s, ls);
        rep.err += tail_sup(hess, tau_s, betak, std::numeric_limits<double>::infinity()) * cn / beta0;
        Tensor out = scale(ex, coeff[0] * cn, basis[0]);
        std::vector<cplx> cs;
        std::vector<const Tensor*> bp;
        for (int i = 1; i < dim; ++i) {
            cs.push_back(coeff[i] * cn);
            bp.push_back(&basis[i]);
        }
        axpy_many_(ex, cs, bp, out);
        cur = out;
        logscale += ls;
        tau_rem -= tau_s;
        if (std::abs(tau_rem) <= 1e-14 * std::max(std::abs(tau), 1.0)) break;
        if (rep.substeps >= 10000) throw StepError(BUG_EINVARIANT, "expv: substep search failed to advance");
    }
    return finish_scale(ex, cur, logscale);
}

// ---- many instances in lockstep ---------------------------------------------------------------------

namespace {

// A fixed set of host threads running index ranges: run(count, f) calls f(i) for i < count on the
// caller and the workers and returns when all calls are done (the first exception is rethrown).
class HostPool {
  public:
    explicit HostPool(int threads) {
        for (int t = 1; t < std::max(1, threads); ++t) workers_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& w : workers_) w.join();
    }
    template <typename F>
    void run(int count, F&& f) {
        if (count <= 0) return;
        if (workers_.empty() || count == 1) {
            for (int i = 0; i < count; ++i) f(i);
            return;
        }
        std::function<void(int)> job = f;
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &job;
            count_ = count;
            next_.store(0);
            done_ = 0;
            err_ = nullptr;
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return done_ == count_; });
        job_ = nullptr;
        if (err_) std::rethrow_exception(err_);
    }

  private:
    void work() {
        int did = 0;
        std::exception_ptr err;
        for (int i = next_.fetch_add(1); i < count_; i = next_.fetch_add(1)) {
            try {
                (*job_)(i);
            } catch (...) {
                if (!err) err = std::current_exception();
            }
            ++did;
        }
        std::lock_guard<std::mutex> lk(mu_);
        if (err && !err_) err_ = err;
        done_ += did;
        if (done_ == count_) done_cv_.notify_all();
    }
    void loop() {
        uint64_t seen = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || (gen_ != seen && job_); });
                if (stop_) return;
                seen = gen_;
            }
            work();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::function<void(int)>* job_ = nullptr;
    int count_ = 0, done_ = 0;
    std::atomic<int> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr err_;
};

// packed vectors on the step's stream
struct Packed {
    const Exec& ex;
    const PackedSegs& sg;
    int K;
    std::shared_ptr<DevBuf> fresh(bool zero) const {
        auto b = std::make_shared<DevBuf>(sizeof(double2) * (size_t)std::max<int64_t>(sg.n, 1));
        if (zero) STEP_HIP(hipMemsetAsync(b->p, 0, sizeof(double2) * (size_t)std::max<int64_t>(sg.n, 1), ex.st));
        return b;
    }
    // out[j][i] = <xs[j]_i | y_i>
    std::vector<std::vector<cplx>> dots(const std::vector<const double2*>& xs, const double2* y) const {
        std::vector<std::vector<cplx>> out(xs.size(), std::vector<cplx>(K));
        std::vector<double> res;
        for (size_t lo = 0; lo < xs.size(); lo += 64) {
            const int k = (int)std::min<size_t>(64, xs.size() - lo);
            std::vector<const void*> ptrs(xs.begin() + lo, xs.begin() + lo + k);
            res.assign((size_t)2 * k * K, 0.0);
            ex.check(bug_dotc_segments(ex.ctx, sg.n, K, sg.table.data(), k, ptrs.data(), y, res.data(), ex.st));
            for (int j = 0; j < k; ++j)
                for (int i = 0; i < K; ++i)
                    out[lo + j][i] = cplx(res[2 * ((size_t)j * K + i)], res[2 * ((size_t)j * K + i) + 1]);
        }
        return out;
    }
    std::vector<double> norms(const double2* y) const {
        const auto d = dots({y}, y);
        std::vector<double> n(K);
        for (int i = 0; i < K; ++i) n[i] = std::sqrt(std::max(d[0][i].real(), 0.0));
        return n;
    }
    // y_i += sum_j a[j][i] xs[j]_i
    void axpy(const std::vector<std::vector<cplx>>& a, const std::vector<const double2*>& xs, double2* y) const {
        std::vector<double> coef;
        for (size_t lo = 0; lo < xs.size(); lo += 64) {
            const int k = (int)std::min<size_t>(64, xs.size() - lo);
            std::vector<const void*> ptrs(xs.begin() + lo, xs.begin() + lo + k);
            coef.assign((size_t)2 * k * K, 0.0);
            for (int j = 0; j < k; ++j)
                for (int i = 0; i < K; ++i) {
                    coef[2 * ((size_t)j * K + i)] = a[lo + j][i].real();
                    coef[2 * ((size_t)j * K + i) + 1] = a[lo + j][i].imag();
                }
            ex.check(bug_axpy_segments(ex.ctx, sg.n, K, sg.table.data(), k, coef.data(), ptrs.data(), y, ex.st));
        }
    }
    // a new vector with segment i equal to s[i] x_i
    std::shared_ptr<DevBuf> scaled(const double2* x, const std::vector<cplx>& s) const {
        auto b = fresh(true);
        axpy({s}, {x}, static_cast<double2*>(b->p));
        return b;
    }
};

}  // namespace

void expv_lockstep(const Exec& ex, const PackedSegs& sg, const PackedApplyFn& apply,
                   const std::vector<cplx>& tau, bool hermitian, const double2* v0, double tol,
                   int maxiter, bool reorth, double conv_tol, double2* out,
                   std::vector<ExpvReport>& reps, int host_threads, RootRule rule) {
    const int K = sg.K();
    if ((int)tau.size() != K) throw StepError(BUG_EINVAL, "one tau per instance");
    reps.assign(K, ExpvReport());
    if (K == 0) return;
    const Packed pk{ex, sg, K};
    HostPool pool(std::min(host_threads, K));
    const std::vector<double> beta0 = pk.norms(v0);
    std::vector<char> done(K, 0);
    for (int i = 0; i < K; ++i) done[i] = beta0[i] == 0.0;
    std::vector<cplx> tau_rem(tau), logscale(K, 0.0);
    auto cur = pk.fresh(true);
    pk.axpy({std::vector<cplx>(K, 1.0)}, {v0}, static_cast<double2*>(cur->p));
    const int ld = maxiter;
    auto any = [](const std::vector<char>& m) { return std::find(m.begin(), m.end(), 1) != m.end(); };
    while (true) {
        std::vector<char> act(K, 0);
        for (int i = 0; i < K; ++i) act[i] = !done[i];
        if (!any(act)) break;
        for (int i = 0; i < K; ++i)
            if (act[i]) reps[i].substeps += 1;
        const std::vector<double> cn = pk.norms(static_cast<double2*>(cur->p));
        for (int i = 0; i < K; ++i)
            if (act[i] && cn[i] == 0.0) act[i] = 0, done[i] = 1;
        if (!any(act)) break;
        std::vector<cplx> inv(K, 0.0);
        for (int i = 0; i < K; ++i)
            if (act[i]) inv[i] = 1.0 / cn[i];
        std::vector<std::shared_ptr<DevBuf>> basis{pk.scaled(static_cast<double2*>(cur->p), inv)};
        auto bptr = [&](size_t j) { return static_cast<const double2*>(basis[j]->p); };
        auto all_basis = [&] {
            std::vector<const double2*> v;
            for (size_t j = 0; j < basis.size(); ++j) v.push_back(bptr(j));
            return v;
        };
        std::vector<char> pass(act);
        std::vector<double> betak(K, 0.0);
        std::vector<int> dim(K, 0);
        std::vector<std::vector<double>> alpha(K), beta(K);
        std::vector<std::vector<cplx>> H(hermitian ? 0 : K);
        if (!hermitian)
            for (int i = 0; i < K; ++i) H[i].assign((size_t)(maxiter + 1) * ld, 0.0);
        auto masked = [&](std::vector<cplx> c) {
            for (int i = 0; i < K; ++i)
                if (!pass[i]) c[i] = 0.0;
            return c;
        };
        auto r = pk.fresh(false);
        double2* rp = static_cast<double2*>(r->p);
        if (hermitian) {
            apply(bptr(0), rp);
            for (int i = 0; i < K; ++i)
                if (pass[i]) reps[i].iterations += 1;
            const auto a = pk.dots({bptr(0)}, rp)[0];
            std::vector<cplx> na(K);
            for (int i = 0; i < K; ++i) {
                if (pass[i]) alpha[i].push_back(a[i].real());
                na[i] = -a[i].real();
            }
            pk.axpy({masked(na)}, {bptr(0)}, rp);
            for (int it = 2; it <= maxiter; ++it) {
                const std::vector<double> b = pk.norms(rp);
                pool.run(K, [&](int i) {
                    if (!pass[i]) return;
                    if (b[i] < tol ||
                        (conv_tol > 0 && tail_sup(Tridiagonal(alpha[i], beta[i], rule), tau_rem[i], b[i], conv_tol) < conv_tol)) {
                        pass[i] = 0;
                        betak[i] = b[i];
                    }
                });
                if (!any(pass)) break;
                std::vector<cplx> ib(K, 0.0), nb(K, 0.0);
                for (int i = 0; i < K; ++i)
                    if (pass[i]) {
                        beta[i].push_back(b[i]);
                        ib[i] = 1.0 / b[i];
                        nb[i] = -b[i];
                    }
                basis.push_back(pk.scaled(rp, ib));
                const size_t last = basis.size() - 1;
                apply(bptr(last), rp);
                for (int i = 0; i < K; ++i)
                    if (pass[i]) reps[i].iterations += 1;
                const auto a2 = pk.dots({bptr(last)}, rp)[0];
                std::vector<cplx> na2(K, 0.0);
                for (int i = 0; i < K; ++i)
                    if (pass[i]) {
                        alpha[i].push_back(a2[i].real());
                        na2[i] = -a2[i].real();
                    }
                pk.axpy({masked(na2), masked(nb)}, {bptr(last), bptr(last - 1)}, rp);
                if (reorth)
                    for (int p = 0; p < 2; ++p) {
                        auto cs = pk.dots(all_basis(), rp);
                        for (auto& row : cs) row = masked(row);
                        for (auto& row : cs)
                            for (auto& c : row) c = -c;
                        pk.axpy(cs, all_basis(), rp);
                    }
            }
            // instances whose pass ran to maxiter: the tail of their last residual
            if (any(pass)) {
                const std::vector<double> b = pk.norms(rp);
                for (int i = 0; i < K; ++i)
                    if (pass[i]) betak[i] = b[i], pass[i] = 0;
            }
            for (int i = 0; i < K; ++i) dim[i] = (int)alpha[i].size();
        } else {
            for (int k = 1; k <= maxiter; ++k) {
                apply(bptr(k - 1), rp);
                for (int i = 0; i < K; ++i)
                    if (pass[i]) reps[i].iterations += 1;
                for (int p = 0; p < 2; ++p) {
                    auto cs = pk.dots(all_basis(), rp);
                    for (auto& row : cs) row = masked(row);
                    for (size_t j = 0; j < cs.size(); ++j)
                        for (int i = 0; i < K; ++i) H[i][j * ld + (k - 1)] += cs[j][i];
                    for (auto& row : cs)
                        for (auto& c : row) c = -c;
                    pk.axpy(cs, all_basis(), rp);
                }
                const std::vector<double> b = pk.norms(rp);
                pool.run(K, [&](int i) {
                    if (!pass[i]) return;
                    dim[i] = k;
                    betak[i] = b[i];
                    bool stop = b[i] < tol;
                    if (!stop && conv_tol > 0) {
                        std::vector<cplx> S((size_t)k * k);
                        for (int a = 0; a < k; ++a)
                            for (int c = 0; c < k; ++c) S[(size_t)a * k + c] = H[i][(size_t)a * ld + c];
                        stop = tail_sup(Hessenberg(S, k, rule), tau_rem[i], b[i], conv_tol) < conv_tol;
                    }
                    if (!stop) {
                        H[i][(size_t)k * ld + (k - 1)] = b[i];
                        stop = k == maxiter;
                    }
                    if (stop) pass[i] = 0;
                });
                if (!any(pass)) break;
                std::vector<cplx> ib(K, 0.0);
                for (int i = 0; i < K; ++i)
                    if (pass[i]) ib[i] = 1.0 / b[i];
                basis.push_back(pk.scaled(rp, ib));
            }
        }
        // each instance's substep: step size, coefficients, error bound, new current vector
        std::vector<std::vector<cplx>> coef(basis.size(), std::vector<cplx>(K, 0.0));
        std::vector<cplx> ls(K, 0.0), tau_s(K, 0.0);
        std::vector<double> err(K, 0.0);
        pool.run(K, [&](int i) {
            if (!act[i]) return;
            std::vector<cplx> c;
            if (hermitian) {
                const Tridiagonal tri(alpha[i], beta[i], rule);
                tau_s[i] = rule.rational() ? tau_rem[i] : pick_tau(tri, tau_rem[i], betak[i], conv_tol);
                c = tri.coeffs(tau_s[i], ls[i]);
                err[i] = tail_sup(tri, tau_s[i], betak[i], std::numeric_limits<double>::infinity());
            } else {
                const int d = dim[i];
                std::vector<cplx> S((size_t)d * d);
                for (int a = 0; a < d; ++a)
                    for (int b = 0; b < d; ++b) S[(size_t)a * d + b] = H[i][(size_t)a * ld + b];
                const Hessenberg hess(S, d, rule);
                tau_s[i] = rule.rational() ? tau_rem[i] : pick_tau(hess, tau_rem[i], betak[i], conv_tol);
                c = hess.coeffs(tau_s[i], ls[i]);
                err[i] = tail_sup(hess, tau_s[i], betak[i], std::numeric_limits<double>::infinity());
            }
            const size_t m = std::min<size_t>(std::min<size_t>(c.size(), (size_t)dim[i]), basis.size());
            for (size_t j = 0; j < m; ++j) coef[j][i] = c[j] * cn[i];
        });
        auto next = pk.fresh(true);
        pk.axpy(coef, all_basis(), static_cast<double2*>(next->p));
        // instances that took no substep keep their current vector
        std::vector<cplx> keep(K, 0.0);
        bool any_keep = false;
        for (int i = 0; i < K; ++i)
            if (!act[i]) keep[i] = 1.0, any_keep = true;
        if (any_keep) pk.axpy({keep}, {static_cast<const double2*>(cur->p)}, static_cast<double2*>(next->p));
        cur = next;
        for (int i = 0; i < K; ++i) {
            if (!act[i]) continue;
            reps[i].err += err[i] * cn[i] / beta0[i];
            logscale[i] += ls[i];
            tau_rem[i] -= tau_s[i];
            if (std::abs(tau_rem[i]) <= 1e-14 * std::max(std::abs(tau[i]), 1.0)) done[i] = 1;
            if (!done[i] && reps[i].substeps >= 10000)
                throw StepError(BUG_EINVARIANT, "expv: substep search failed to advance");
        }
    }
    // the accumulated scale, where it neither vanishes nor overflows (finish_scale per instance)
    const std::vector<double> nrm = pk.norms(static_cast<double2*>(cur->p));
    std::vector<cplx> sc(K, 1.0);
    for (int i = 0; i < K; ++i) {
        const cplx e = std::exp(logscale[i]);
        if (beta0[i] == 0.0 || !(std::isfinite(std::abs(e)) && e != 0.0)) continue;
        if (!(std::isfinite(nrm[i]) && nrm[i] > 0 && std::abs(e) * nrm[i] > kMaxNorm)) sc[i] = e;
    }
    STEP_HIP(hipMemsetAsync(out, 0, sizeof(double2) * (size_t)std::max<int64_t>(sg.n, 1), ex.st));
    pk.axpy({sc}, {static_cast<const double2*>(cur->p)}, out);
}

}  // namespace host
}  // namespace bug
