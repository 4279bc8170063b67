// Krylov pieces of the step-level entry point: the host k x k side (tridiagonal and Hessenberg
// exponentials, the residual integral that weighs a local Krylov direction, the tail estimate and
// substep search) and the device recurrences (Lanczos and Arnoldi exponentials) on tensors.
//
// Reference: lanczos_expv / arnoldi_expv and their helpers, src/BUG/Core/LocalSolvers/krylov.jl:
// 44-117,128-263,296-442; the substep search and tail estimate, root_methods.jl:86-141; the
// Krylov direction weights, src/BUG/Core/RankAdaptation/krylov_enrichment.jl:58-120; Test A's tail
// rank, truncation.jl:262-278.
#include <algorithm>
#include <cmath>
#include <limits>

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

Tridiagonal::Tridiagonal(const std::vector<double>& alpha, const std::vector<double>& beta) {
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

Hessenberg::Hessenberg(const std::vector<cplx>& H, int n_) : Hm(H), n(n_) {
    for (const cplx& v : Hm)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw StepError(BUG_ENONFINITE, "Arnoldi Hessenberg matrix has non-finite entries: the "
                                            "input vector or the operator is not finite");
    if (n > 0) vals = eigvals(Hm, n);
}

std::vector<cplx> Hessenberg::coeffs(cplx tau, cplx& logscale) const {
    logscale = 0.0;
    if (n == 0) return {};
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
                    ExpvReport& rep) {
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
            if (conv_tol > 0 && tail_sup(Tridiagonal(alpha, beta), tau_rem, b, conv_tol) < conv_tol) break;
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
        const Tridiagonal tri(alpha, beta);
        const double betak = norm(ex, r);
        const cplx tau_s = pick_tau(tri, tau_rem, betak, conv_tol);
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
                    const Tensor& v0, double tol, int maxiter, double conv_tol, ExpvReport& rep) {
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
        return Hessenberg(S, dim);
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
        const cplx tau_s = pick_tau(hess, tau_rem, betak, conv_tol);
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

}  // namespace host
}  // namespace bug
