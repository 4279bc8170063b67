// The step-level entry point: one rank-adaptive CBE-BUG step on a device-resident U(1) MPS,
// orchestrated in C++ on the library's own kernels (bug_step in bug.h).
//
// Reference (madhavmen/BondUpdateGalerkin), stage by stage:
//   options and budget         src/BUG/TimeEvolution/bug_step_stages.jl:15-142,
//                              src/BUG/Core/RankAdaptation/basis_budget.jl:75-128,454-549,568
//   half-sweeps                src/BUG/Core/Sweeps/half_sweeps.jl:218-410, bug_step.jl:222-268
//   bond frames                src/BUG/Core/Environments/effective_operators.jl:201-220
//   DD-QR CBE (:rank_tail)     src/BUG/Randomised/dd_cbe_expand.jl:19-87,252-301,383-419,
//                              dd_cbe.jl:38-89,132-161, cbe.jl:156-236, sketch.jl:141-180,
//                              the streamed sketch mpo_environments.jl:409-481
//   local Krylov (Tests C/D)   src/BUG/Core/RankAdaptation/krylov_enrichment.jl:135-292
//   order-p closure            krylov_closure_loop.jl:203-281, krylov_closure_frames.jl,
//                              src/BUG/Core/Sweeps/frontier_closure.jl:42-236, frontier_sweeps.jl
//   projection miss            krylov_closure_frames.jl:242-280
//   root solve                 effective_operators.jl:78-107, src/BUG/Core/LocalSolvers/krylov.jl
//   assembly and compression   src/BUG/Core/diagnostics.jl:269-337, truncation.jl:49-158
//
// Gauge moves are unique thin QRs (real non-negative R diagonal) where the reference takes
// cutoff-0 SVDs; both give the same isometries' spans and the same represented tensors.  The probes
// come from a counter-based stream (splitmix64 + Box-Muller) consumed in the reference's order.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <thread>

#include "internal.h"
#include "step_krylov.h"

namespace bug {
namespace host {
namespace {

constexpr double kCbeEps = 100 * 2.220446049250313e-16;   // _CBE_EPS, truncation.jl:173
constexpr double kFrontierFloor = 1e-14;                   // _FRONTIER_FLOOR
constexpr double kPolarDefect = 1e-4;

// ---- ledger ---------------------------------------------------------------------------------------

struct BondLogEntry {
    int bond, side;
    double sample, tail, krylov_c, krylov_d;
    int n_probes, n_holdout, n_selected, n_targeted, n_krylov, n_h1;
    bool workcap, met;
    int unmet, width;
    int r_old = 0, r_final = 0;
    double discarded = 0.0;
};

struct Ledger {
    double max_sample = 0, max_tail = 0, max_krylov = 0, max_krylov_c = 0, max_krylov_d = 0;
    double discarded = 0, expv_err = 0;
    int expv_iters = 0;
    int64_t n_probes = 0, n_selected = 0, n_krylov = 0, n_h1 = 0, n_holdout = 0, n_targeted = 0;
    int64_t n_h1_closure = 0, n_closure_sweeps = 0;
    bool workcap_reached = false, tolerance_met = true, maxdim_reached = false;
    double proj_err = 0.0, defect = 0.0;
    int r_old = 0, r_aug = 0, r_final = 0;
    std::vector<double> leak;
    int unmet = 0;
    std::vector<BondLogEntry> log;
    void fail(int bit) {
        tolerance_met = false;
        unmet |= bit;
    }
    void merge(const Ledger& o) {
        max_sample = std::max(max_sample, o.max_sample);
        max_tail = std::max(max_tail, o.max_tail);
        max_krylov = std::max(max_krylov, o.max_krylov);
        max_krylov_c = std::max(max_krylov_c, o.max_krylov_c);
        max_krylov_d = std::max(max_krylov_d, o.max_krylov_d);
        n_probes += o.n_probes;
        n_selected += o.n_selected;
        n_krylov += o.n_krylov;
        n_h1 += o.n_h1;
        n_holdout += o.n_holdout;
        n_targeted += o.n_targeted;
        n_h1_closure += o.n_h1_closure;
        n_closure_sweeps += o.n_closure_sweeps;
        workcap_reached |= o.workcap_reached;
        tolerance_met &= o.tolerance_met;
        unmet |= o.unmet;
        log.insert(log.end(), o.log.begin(), o.log.end());
    }
};

void log_bond(Ledger& ledger, Ledger& bl, int bond, int side, int width) {
    bl.log.push_back(BondLogEntry{bond, side, bl.max_sample, bl.max_tail, bl.max_krylov_c,
                                  bl.max_krylov_d, (int)bl.n_probes, (int)bl.n_holdout,
                                  (int)bl.n_selected, (int)bl.n_targeted, (int)bl.n_krylov,
                                  (int)bl.n_h1, bl.workcap_reached, bl.tolerance_met, bl.unmet,
                                  width});
    ledger.merge(bl);
}

struct Alloc {
    double per_bond, sample, tail, krylov, scale;
};

Alloc basis_allocation(double tol_basis, int nbonds, cplx tau, double nrm, int nfront, double tol_krylov) {
    const int nb = std::max(nbonds, 1), nf = std::max(nfront, 1);
    const double kr = tol_krylov / tol_basis;
    const double at = std::abs(tau);
    const double scale = at == 0.0 ? 1.0 : at / std::max(nrm, 2.220446049250313e-16);
    const double per_test = tol_basis / (2.0 * (nb + 1) * nf);
    return Alloc{2 * per_test, per_test, per_test, kr * per_test, scale};
}

// math.ceil(round(x, 9)) of the reference's ceilings
int ceil_round9(double x) {
    const double r = std::nearbyint(x * 1e9) / 1e9;
    return (int)std::ceil(r);
}

// ---- states ----------------------------------------------------------------------------------------

struct MPS {
    std::vector<Tensor> t;
    std::vector<int32_t> charges;
    int center = 0;
    int size() const { return (int)t.size(); }
    const Leg& bond_leg(int b) const { return b < size() ? t[b].L().row : t.back().L().col; }
    std::vector<int> bond_dims() const {
        std::vector<int> d;
        for (int i = 0; i + 1 < size(); ++i) d.push_back(t[i].L().col.total());
        return d;
    }
};

void canonical(const Exec& ex, MPS& psi, int site) {
    while (psi.center < site) {
        const int i = psi.center;
        Tensor Q, R;
        qr_left(ex, psi.t[i], Q, R);
        psi.t[i] = Q;
        psi.t[i + 1] = core_times_site(ex, R, psi.t[i + 1]);
        psi.center += 1;
    }
    while (psi.center > site) {
        const int i = psi.center;
        Tensor Lc, Q;
        qr_right(ex, psi.t[i], Lc, Q);
        psi.t[i] = Q;
        psi.t[i - 1] = site_times_core(ex, psi.t[i - 1], Lc);
        psi.center -= 1;
    }
}

// ---- probes and sketches --------------------------------------------------------------------------

std::map<int, int> fused_dims(const Layout& lay, Side side) {
    std::map<int, int> out;
    if (side == LEFT) {
        for (int i = 0; i < lay.row.nsec(); ++i)
            for (int32_t sh : lay.shifts) out[lay.row.q[i] + sh] += lay.row.d[i];
    } else {
        for (int i = 0; i < lay.col.nsec(); ++i)
            for (int32_t sh : lay.shifts) out[lay.col.q[i] - sh] += lay.col.d[i];
    }
    return out;
}

std::map<int, int> widths(const Layout& lay, Side side) {
    return side == LEFT ? lay.col.as_map() : lay.row.as_map();
}

int width_of(const Tensor& t, Side side) { return side == LEFT ? t.L().col.total() : t.L().row.total(); }

// Gaussian columns on the fused local space of `frame`: per fused charge ascending, per slice
// ascending, one (link x dfull) block column-major of which the first cols[q] columns are kept
Tensor graded_draw(const Exec& ex, const Tensor& frame, Side side, const std::map<int, int>& dfull,
                   const std::map<int, int>& cols, ProbeStream& stream) {
    const Layout& lay = frame.L();
    std::vector<std::pair<int, int>> wl;
    for (auto& c : cols)
        if (c.second > 0) wl.emplace_back(c.first, c.second);
    const Leg wleg = Leg::make(wl);
    const Leg& link = side == LEFT ? lay.row : lay.col;
    const Layout play = side == LEFT ? Layout(link, wleg, lay.shifts) : Layout(wleg, link, lay.shifts);
    std::vector<cplx> host((size_t)std::max<int64_t>(play.nelem, 1), 0.0);
    uint64_t pos = stream.pos;
    for (auto& fd : dfull) {
        const int q = fd.first, df = fd.second;
        auto ci = cols.find(q);
        const int c = ci == cols.end() ? 0 : ci->second;
        for (int s = 0; s < lay.nslice(); ++s) {
            const int sh = lay.shifts[s];
            const int lq = side == LEFT ? q - sh : q + sh;
            const int dl = link.dim_of(lq);
            if (dl == 0) continue;
            if (c > 0) {
                // numbers pos .. pos + dl * c: block (dl x c) column-major
                if (side == LEFT) {
                    const Blk* b = play.find(s, lq);
                    for (int jc = 0; jc < c; ++jc)
                        for (int r = 0; r < dl; ++r)
                            host[(size_t)(b->off + (int64_t)jc * dl + r)] = stream.at(pos + (uint64_t)jc * dl + r);
                } else {
                    // the probe block is the transpose: (c x dl), element (jc, r) = number jc*dl + r
                    const Blk* b = play.find(s, q);
                    for (int jc = 0; jc < c; ++jc)
                        for (int r = 0; r < dl; ++r)
                            host[(size_t)(b->off + (int64_t)r * c + jc)] = stream.at(pos + (uint64_t)jc * dl + r);
                }
            }
            pos += (uint64_t)dl * df;
        }
    }
    stream.pos = pos;
    return t_from_host(ex, play, host);
}

std::map<int, int> alloc_columns(const std::map<int, int>& dims, int width) {
    int tot = 0;
    for (auto& d : dims) tot += d.second;
    std::map<int, int> out;
    for (auto& d : dims)
        out[d.first] = (tot == 0 || width <= 0) ? 0
                       : std::min(d.second, (int)std::ceil((double)d.second / tot * width));
    return out;
}

// sector-graded Gaussian probe in the discarded space of `frame` (_dd_probe, dd_cbe.jl:132-161)
Tensor dd_probe(const Exec& ex, const Tensor& frame, Side side, int width, ProbeStream& stream) {
    const auto dfull = fused_dims(frame.L(), side);
    const auto dT = widths(frame.L(), side);
    std::map<int, int> dC;
    int dsum = 0;
    for (auto& f : dfull) {
        auto it = dT.find(f.first);
        dC[f.first] = std::max(f.second - (it == dT.end() ? 0 : it->second), 0);
        dsum += dC[f.first];
    }
    if (dsum == 0) return Tensor();
    const auto cols = alloc_columns(dC, width);
    Tensor Om = graded_draw(ex, frame, side, dfull, cols, stream);
    Tensor proj = perp(ex, side, frame, Om);
    return norm(ex, proj) > 0 ? proj : Tensor();
}

// sector_graded_sketch with no compression ratio (sketch.jl:45-54): columns by full fused dims
Tensor graded_probe(const Exec& ex, const Tensor& frame, Side side, int width, ProbeStream& stream) {
    const auto dfull = fused_dims(frame.L(), side);
    const auto cols = alloc_columns(dfull, width);
    bool any = false;
    for (auto& c : cols) any |= c.second > 0;
    if (!any) return Tensor();
    Tensor Om = graded_draw(ex, frame, side, dfull, cols, stream);
    return norm(ex, Om) > 0 ? Om : Tensor();
}

// (H Theta) Om^dag for Theta = A V with Om in V's layout (_stream_left, mpo_environments.jl:409-444)
Tensor stream_left(const Exec& ex, const Tensor& L, const MpoSite& W1, const MpoSite& W2,
                   const Tensor& R, const Tensor& A, const Tensor& V, const Tensor& Om, int prec) {
    const auto &Rs = R.L().shifts, &Vs = V.L().shifts, &Os = Om.L().shifts, &As = A.L().shifts;
    // F slots: (c, si) of W2's entries in first-seen order
    std::vector<std::pair<int, int>> fs;
    auto slot = [&](std::vector<std::pair<int, int>>& v, std::pair<int, int> k) {
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] == k) return (int)i;
        v.push_back(k);
        return (int)v.size() - 1;
    };
    std::map<int, int> fsh;
    std::vector<Term> tf;
    for (int e = 0; e < W2.nent(); ++e) {
        const int c = W2.a[e], b = W2.b[e], so = W2.sout[e], si = W2.sin[e];
        const int k = slot(fs, {c, si});
        fsh[k] = -Rs[b] - Os[so];
        tf.push_back({k, b, so, BUG_OP_T, BUG_OP_C, cplx(W2.coeff[2 * e], W2.coeff[2 * e + 1])});
    }
    std::vector<int32_t> fshv;
    for (size_t k = 0; k < fs.size(); ++k) fshv.push_back(fsh[(int)k]);
    Tensor F = product(ex, R.L().col, Om.L().row, fshv, R, Om, tf, prec);
    std::vector<int> cs;
    for (auto& p : fs) cs.push_back(p.first);
    std::sort(cs.begin(), cs.end());
    cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
    auto ci = [&](int c) { return (int)(std::lower_bound(cs.begin(), cs.end(), c) - cs.begin()); };
    std::map<int, int> bsh;
    std::vector<Term> tb;
    for (size_t k = 0; k < fs.size(); ++k) {
        const int c = fs[k].first, si = fs[k].second;
        tb.push_back({ci(c), si, (int)k, BUG_OP_N, BUG_OP_N, 1.0});
        bsh[ci(c)] = Vs[si] + fsh[(int)k];
    }
    std::vector<int32_t> bshv;
    for (size_t i = 0; i < cs.size(); ++i) bshv.push_back(bsh[(int)i]);
    Tensor Bm = product(ex, V.L().row, Om.L().row, bshv, V, F, tb, prec);
    std::vector<std::pair<int, int>> cslots;
    std::map<int, int> csh;
    std::vector<Term> tc;
    for (int e = 0; e < W1.nent(); ++e) {
        const int c = W1.b[e], si = W1.sin[e];
        if (!std::binary_search(cs.begin(), cs.end(), c)) continue;
        bool have = false;
        for (auto& p : cslots) have |= p == std::make_pair(c, si);
        if (have) continue;
        const int k = (int)cslots.size();
        cslots.emplace_back(c, si);
        tc.push_back({k, si, ci(c), BUG_OP_N, BUG_OP_N, 1.0});
        csh[k] = As[si] + bsh[ci(c)];
    }
    std::vector<int32_t> cshv;
    for (size_t k = 0; k < cslots.size(); ++k) cshv.push_back(csh[(int)k]);
    Tensor Cm = product(ex, A.L().row, Om.L().row, cshv, A, Bm, tc, prec);
    std::vector<Term> ty;
    for (int e = 0; e < W1.nent(); ++e) {
        const int a = W1.a[e], c = W1.b[e], so = W1.sout[e], si = W1.sin[e];
        for (size_t k = 0; k < cslots.size(); ++k)
            if (cslots[k] == std::make_pair(c, si))
                ty.push_back({so, a, (int)k, BUG_OP_N, BUG_OP_N, cplx(W1.coeff[2 * e], W1.coeff[2 * e + 1])});
    }
    return product(ex, L.L().row, Om.L().row, As, L, Cm, ty, prec);
}

// Om^dag (H Theta) for Theta = U Bv with Om in U's layout (_stream_right, mpo_environments.jl:446-481)
Tensor stream_right(const Exec& ex, const Tensor& L, const MpoSite& W1, const MpoSite& W2,
                    const Tensor& R, const Tensor& U, const Tensor& Bv, const Tensor& Om, int prec) {
    const auto &Ls = L.L().shifts, &Us = U.L().shifts, &Bs = Bv.L().shifts, &Os = Om.L().shifts;
    std::vector<std::pair<int, int>> fs;
    auto slot = [&](std::vector<std::pair<int, int>>& v, std::pair<int, int> k) {
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] == k) return (int)i;
        v.push_back(k);
        return (int)v.size() - 1;
    };
    std::map<int, int> fsh;
    std::vector<Term> tf;
    for (int e = 0; e < W1.nent(); ++e) {
        const int a = W1.a[e], c = W1.b[e], so = W1.sout[e], si = W1.sin[e];
        const int k = slot(fs, {c, si});
        fsh[k] = -Ls[a] + Os[so];
        tf.push_back({k, a, so, BUG_OP_T, BUG_OP_R, cplx(W1.coeff[2 * e], W1.coeff[2 * e + 1])});
    }
    std::vector<int32_t> fshv;
    for (size_t k = 0; k < fs.size(); ++k) fshv.push_back(fsh[(int)k]);
    Tensor F = product(ex, L.L().col, Om.L().col, fshv, L, Om, tf, prec);
    std::vector<int> cs;
    for (auto& p : fs) cs.push_back(p.first);
    std::sort(cs.begin(), cs.end());
    cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
    auto ci = [&](int c) { return (int)(std::lower_bound(cs.begin(), cs.end(), c) - cs.begin()); };
    std::map<int, int> bsh;
    std::vector<Term> tb;
    for (size_t k = 0; k < fs.size(); ++k) {
        const int c = fs[k].first, si = fs[k].second;
        tb.push_back({ci(c), si, (int)k, BUG_OP_T, BUG_OP_N, 1.0});
        bsh[ci(c)] = -Us[si] + fsh[(int)k];
    }
    std::vector<int32_t> bshv;
    for (size_t i = 0; i < cs.size(); ++i) bshv.push_back(bsh[(int)i]);
    Tensor Bm = product(ex, U.L().col, Om.L().col, bshv, U, F, tb, prec);
    std::vector<std::pair<int, int>> cslots;
    std::map<int, int> csh;
    std::vector<Term> tc;
    for (int e = 0; e < W2.nent(); ++e) {
        const int c = W2.a[e], si = W2.sin[e];
        if (!std::binary_search(cs.begin(), cs.end(), c)) continue;
        bool have = false;
        for (auto& p : cslots) have |= p == std::make_pair(c, si);
        if (have) continue;
        const int k = (int)cslots.size();
        cslots.emplace_back(c, si);
        tc.push_back({k, ci(c), si, BUG_OP_T, BUG_OP_N, 1.0});
        csh[k] = -bsh[ci(c)] + Bs[si];
    }
    std::vector<int32_t> cshv;
    for (size_t k = 0; k < cslots.size(); ++k) cshv.push_back(csh[(int)k]);
    Tensor Cm = product(ex, Om.L().col, Bv.L().col, cshv, Bm, Bv, tc, prec);
    std::vector<Term> ty;
    for (int e = 0; e < W2.nent(); ++e) {
        const int c = W2.a[e], b = W2.b[e], so = W2.sout[e], si = W2.sin[e];
        for (size_t k = 0; k < cslots.size(); ++k)
            if (cslots[k] == std::make_pair(c, si))
                ty.push_back({so, (int)k, b, BUG_OP_N, BUG_OP_T, cplx(W2.coeff[2 * e], W2.coeff[2 * e + 1])});
    }
    return product(ex, Om.L().col, R.L().row, Bs, Cm, R, ty, prec);
}

// ---- frames ---------------------------------------------------------------------------------------

struct BondFrame {
    Tensor U0, S0, V0;
    int old_rank;
};

BondFrame frame_from(const Exec& ex, const Tensor& left, const Tensor& right) {
    BondFrame f;
    Tensor Rl, Lr;
    qr_left(ex, left, f.U0, Rl);
    qr_right(ex, right, Lr, f.V0);
    f.S0 = core_times_core(ex, Rl, Lr);
    f.old_rank = f.U0.L().col.total();
    return f;
}

Tensor reortho(const Exec& ex, const Tensor& frame, const Tensor& block, Side side) {
    if (block.empty()) return Tensor();
    Tensor outside = perp(ex, side, frame, block);
    if (norm(ex, outside) <= kCbeEps * std::max(norm(ex, block), 1.0)) return Tensor();
    if (side == LEFT) {
        SVDResult r = factor_left(ex, outside).select(ex, 1e-14);
        return r.kept.empty() ? Tensor() : r.U;
    }
    SVDResult r = factor_right(ex, outside).select(ex, 1e-14);
    return r.kept.empty() ? Tensor() : r.Vh;
}

Tensor cap_to_local_room(const Exec& ex, const Tensor& frame, const Tensor& cand, Side side) {
    if (cand.empty()) return Tensor();
    const auto dfull = fused_dims(frame.L(), side);
    const auto dT = widths(frame.L(), side);
    std::map<int, int> counts;
    bool over = false;
    for (auto& w : widths(cand.L(), side)) {
        if (w.second == 0) continue;
        auto fi = dfull.find(w.first);
        auto ti = dT.find(w.first);
        const int room = (fi == dfull.end() ? 0 : fi->second) - (ti == dT.end() ? 0 : ti->second);
        counts[w.first] = std::min(w.second, room);
        over |= counts[w.first] < w.second;
    }
    return over ? trim(ex, cand, side, counts) : cand;
}

// the first sectors in ascending charge order (_trim_total, cbe.jl:219-236)
Tensor trim_total(const Exec& ex, const Tensor& t, Side side, int n) {
    if (n <= 0) return Tensor();
    const auto w = widths(t.L(), side);
    int tot = 0;
    for (auto& e : w) tot += e.second;
    if (tot <= n) return t;
    int left = n;
    std::map<int, int> keep;
    for (auto& e : w) {
        keep[e.first] = std::min(e.second, left);
        left -= keep[e.first];
    }
    return trim(ex, t, side, keep);
}

// Test B: per-width-sector normalised holdout residual (_sample_residual, sketch.jl:141-180)
double sample_residual(const Exec& ex, const Tensor& basis, Side side, const Tensor& proj_frame,
                       const Tensor& Y, const Tensor& Om) {
    const Leg& cl = side == LEFT ? Om.L().row : Om.L().col;
    std::map<int, int> cols;
    for (int i = 0; i < cl.nsec(); ++i)
        if (cl.d[i] > 0) cols[cl.q[i]] = cl.d[i];
    if (cols.empty()) return 0.0;
    Tensor res = perp(ex, side, proj_frame, Y);
    if (norm(ex, res) <= kCbeEps) return 0.0;
    if (!basis.empty()) res = perp(ex, side, proj_frame, perp(ex, side, basis, res));
    std::vector<bug_copy> blks;
    std::vector<int> keys;
    for (const Blk& b : res.L().blks) {
        blks.push_back(bug_copy{b.off, 0, std::max(b.d, 1), 1, b.d, b.dc, 0, 0, 0});
        keys.push_back(side == LEFT ? b.qc : b.q);
    }
    const std::vector<double> sq = block_sqnorms(ex, blks, res.buf->p, res.nelem());
    std::vector<std::pair<int, double>> acc;           // first-seen order of the width sectors
    for (size_t i = 0; i < keys.size(); ++i) {
        bool found = false;
        for (auto& a : acc)
            if (a.first == keys[i]) {
                a.second += sq[i];
                found = true;
            }
        if (!found) acc.emplace_back(keys[i], sq[i]);
    }
    double tot = 0.0;
    for (auto& a : acc) {
        auto it = cols.find(a.first);
        if (it == cols.end() || it->second <= 0)
            throw StepError(BUG_EINVARIANT, "Test B: residual sector without a holdout column");
        tot += a.second / it->second;
    }
    return std::sqrt(tot);
}

struct Expansion {
    Tensor U_ex, V_ex;
    int n_new;
};

// one side of one bond, :rank_tail (dd_cbe_expand.jl:19-87,252-301,383-419)
Expansion dd_cbe_expand(const Exec& ex, const BondFrame& fr, const std::vector<MpoSite>& mpo, int bond,
                        const Tensor& lenv, const Tensor& renv, Side side, const Alloc& alloc,
                        Ledger& ledger, int workcap, int fixed_width, double tail_factor,
                        ProbeStream& stream, int prec) {
    const bool left = side == LEFT;
    const Tensor& frame = left ? fr.U0 : fr.V0;
    const Tensor& opp = left ? fr.V0 : fr.U0;
    const Side oside = left ? RIGHT : LEFT;
    const Expansion none{fr.U0, fr.V0, 0};
    int full_self = 0, full_opp = 0;
    for (auto& f : fused_dims(frame.L(), side)) full_self += f.second;
    for (auto& f : fused_dims(opp.L(), oside)) full_opp += f.second;
    int cap = full_self - fr.old_rank;
    bool wbind = false;
    if (workcap > 0) {
        const int wk = std::max(workcap - fr.old_rank, 0);
        wbind = wk < cap;
        cap = std::min(cap, wk);
    }
    if (cap <= 0) return none;
    const MpoSite &W1 = mpo[bond], &W2 = mpo[bond + 1];
    const int width = std::min(fixed_width, full_opp);
    Tensor Om = dd_probe(ex, opp, oside, width, stream);
    if (Om.empty()) return none;
    ledger.n_probes += width_of(Om, oside);
    Tensor Ysk = left ? stream_left(ex, lenv, W1, W2, renv, site_times_core(ex, fr.U0, fr.S0), fr.V0, Om, prec)
                      : stream_right(ex, lenv, W1, W2, renv, fr.U0, core_times_site(ex, fr.S0, fr.V0), Om, prec);
    Tensor Y = perp(ex, side, frame, Ysk);
    const double ny = norm(ex, Y);
    if (!std::isfinite(ny))
        throw StepError(BUG_ENONFINITE, "dd_cbe_expand: non-finite sketch at bond " + std::to_string(bond));
    if (ny <= kCbeEps) return none;
    const int rmax = std::min(cap, width_of(Om, oside));
    std::map<int, Tensor> cache;
    const BlockSVD fac = factor(ex, side, Y);
    auto candidate = [&](int k) -> Tensor {
        if (k <= 0) return Tensor();
        auto it = cache.find(k);
        if (it != cache.end()) return it->second;
        SVDResult r = fac.select(ex, kCbeEps, k);
        Tensor c = r.kept.empty() ? Tensor() : reortho(ex, frame, left ? r.U : r.Vh, side);
        cache[k] = c;
        return c;
    };
    const double limit = tail_factor * alloc.sample;
    auto metric = [&](int k) { return alloc.scale * sample_residual(ex, candidate(k), side, frame, Y, Om); };
    if (metric(0) <= limit) return none;
    int hi = rmax, count;
    Tensor directions;
    if (metric(hi) > limit) {
        directions = candidate(hi);
        count = hi;
    } else {
        int lo = 0;
        while (hi - lo > 1) {
            const int k = (hi + lo) / 2;
            if (metric(k) <= limit)
                hi = k;
            else
                lo = k;
        }
        directions = candidate(hi);
        count = directions.empty() ? 0 : width_of(directions, side);
    }
    if (directions.empty()) return none;
    if (count > cap) {
        directions = trim_total(ex, directions, side, cap);
        if (directions.empty()) return none;
    }
    if (wbind && count >= cap) ledger.workcap_reached = true;
    directions = cap_to_local_room(ex, frame, directions, side);
    if (directions.empty()) return none;
    const int nq = width_of(directions, side);
    ledger.n_selected += nq;
    Tensor grown = oplus(ex, side, frame, directions);
    return left ? Expansion{grown, fr.V0, nq} : Expansion{fr.U0, grown, nq};
}

// ---- local Krylov enrichment -------------------------------------------------------------------------

struct KrylovStats {
    double eta_c = 0.0, eta_d = 0.0;
    int added = 0, n_h1 = 0;
    bool capped = false, met = true;
};

// _extend_frame (krylov_enrichment.jl:135-171) -> frame, added, fits
Tensor extend_frame(const Exec& ex, const Tensor& frame, const Tensor& residual, const Tensor& direction,
                    Side side, double weight, double allowance, int room, int& added, bool& fits) {
    added = 0;
    const double rn = norm(ex, residual);
    const double cut = kCbeEps * std::max(norm(ex, direction), rn) / rn;
    const BlockSVD fac = factor(ex, side, residual);
    SVDResult res = fac.select(ex, cut);
    const int nk = (int)res.kept.size();
    int keep;
    const double tail = tail_rank(res.kept, weight, allowance, std::min(nk, room), sq_weight(res.trunc), keep);
    fits = tail <= allowance;
    if (keep == 0) return frame;
    if (keep < nk) res = fac.select(ex, cut, keep);
    Tensor nw = reortho(ex, frame, side == LEFT ? res.U : res.Vh, side);
    nw = cap_to_local_room(ex, frame, nw, side);
    if (nw.empty()) return frame;
    added = width_of(nw, side);
    return oplus(ex, side, frame, nw);
}

Tensor krylov_enrich(const Exec& ex, const Tensor& seed, const std::function<Tensor(const Tensor&)>& applyH,
                     const Tensor& start_frame, Side side, cplx tau, const Alloc& alloc, int workcap,
                     bool lanczos, double hermitian_tol, double amp, bool state_seed, KrylovStats& st) {
    const double allow = alloc.krylov;
    Tensor frame = start_frame;
    auto score = [&](const Tensor& direction, double weight, bool first) -> bool {
        Tensor residual = perp(ex, side, frame, direction);
        const double eta = weight * norm(ex, residual);
        if (first)
            st.eta_c = eta;
        else
            st.eta_d = std::max(st.eta_d, eta);
        if (!(eta > allow)) return true;
        const int room = workcap > 0 ? workcap - width_of(frame, side) : (1 << 30);
        if (room <= 0) return false;
        int nadd;
        bool fits;
        frame = extend_frame(ex, frame, residual, direction, side, weight, allow, room, nadd, fits);
        st.added += nadd;
        return fits;
    };
    const double sn = norm(ex, seed);
    if (sn == 0.0) return frame;
    const cplx teff = tau != 0.0 ? tau : cplx(1.0);
    std::vector<Tensor> vs{scale(ex, 1.0 / sn, seed)};
    const int ld = kMaxDepth + 1;
    std::vector<cplx> H((size_t)ld * ld, 0.0);
    bool closed = false, capped = false;
    double prev_beta = 0.0;
    for (int k = 1; k <= kMaxDepth; ++k) {
        Tensor img = applyH(lanczos ? vs[k - 1] : vs.back());
        st.n_h1 += 1;
        if (lanczos) {
            if (k > 1) axpy_(ex, -prev_beta, vs[k - 2], img);
            const cplx a = inner(ex, vs[k - 1], img);
            if (!(std::abs(a.imag()) <= hermitian_tol * std::max(std::abs(a), 1.0)))
                throw StepError(BUG_EINVAL, "Lanczos needs a Hermitian one-site generator");
            H[(size_t)(k - 1) * ld + (k - 1)] = a.real();
            axpy_(ex, -a.real(), vs[k - 1], img);
            if (k >= 20 && k % 2 == 0) {
                std::vector<const Tensor*> bp;
                for (auto& t : vs) bp.push_back(&t);
                auto cs = inner_many(ex, bp, img);
                for (auto& c : cs) c = -c;
                axpy_many_(ex, cs, bp, img);
            }
        } else {
            for (int p = 0; p < 2; ++p) {
                std::vector<const Tensor*> bp;
                for (auto& t : vs) bp.push_back(&t);
                auto cs = inner_many(ex, bp, img);
                for (size_t row = 0; row < cs.size(); ++row) H[row * ld + (k - 1)] += cs[row];
                for (auto& c : cs) c = -c;
                axpy_many_(ex, cs, bp, img);
            }
        }
        const double beta = norm(ex, img);
        if (beta <= kBreakdown) {
            closed = true;
            break;
        }
        H[(size_t)k * ld + (k - 1)] = beta;
        if (lanczos) H[(size_t)(k - 1) * ld + k] = beta;
        vs.push_back(scale(ex, 1.0 / beta, img));
        const double w = amp * direction_weight(H, ld, k, teff, state_seed);
        if (!score(vs.back(), w, k == 1)) {
            capped = true;
            break;
        }
        if (w <= allow) {
            closed = true;
            break;
        }
        prev_beta = beta;
    }
    if (!closed) capped = true;
    st.capped = capped;
    st.met = closed && !capped;
    return frame;
}

void record_krylov(Ledger& ledger, const KrylovStats& st) {
    ledger.max_krylov_c = std::max(ledger.max_krylov_c, st.eta_c);
    ledger.max_krylov_d = std::max(ledger.max_krylov_d, st.eta_d);
    ledger.max_krylov = std::max({ledger.max_krylov, st.eta_c, st.eta_d});
    ledger.n_krylov += st.added;
    ledger.n_h1 += st.n_h1;
    if (st.capped) ledger.workcap_reached = true;
    if (!st.met) ledger.fail(BUG_UNMET_KRYLOV);
}

// ---- parallel halves ------------------------------------------------------------------------------

// The side thread of parallel half-sweeps: its own library context and HIP stream per main context.
struct Side2 {
    bug_ctx* ctx = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
};
std::mutex g_side_mu;
std::map<bug_ctx*, Side2> g_side;

Side2& side_of(bug_ctx* main) {
    std::lock_guard<std::mutex> lk(g_side_mu);
    Side2& s = g_side[main];
    if (!s.ctx) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) throw StepError(BUG_EDEVICE, "hipGetDevice failed");
        if (bug_ctx_create(dev, &s.ctx) != BUG_OK) throw StepError(BUG_EDEVICE, "side context creation failed");
        if (hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
            throw StepError(BUG_EDEVICE, "side stream creation failed");
    }
    return s;
}

// run `side_fn` on the side thread's stream and `main_fn` on the caller's, ordered at both ends
template <typename A, typename B>
void run_beside(const Exec& ex, A&& main_fn, B&& side_fn) {
    Side2& s = side_of(ex.ctx);
    if (hipEventRecord(s.fork, ex.st) != hipSuccess || hipStreamWaitEvent(s.st, s.fork, 0) != hipSuccess)
        throw StepError(BUG_EDEVICE, "fork of the parallel half failed");
    const Exec sex{s.ctx, s.st, ex.prec};
    int side_rc = BUG_OK;
    std::string side_msg;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::thread th([&] {
        (void)hipSetDevice(dev);
        set_thread_stream(s.st);
        try {
            side_fn(sex);
        } catch (const StepError& e) {
            side_rc = e.code;
            side_msg = e.what();
        } catch (const std::exception& e) {
            side_rc = BUG_EDEVICE;
            side_msg = e.what();
        }
        if (hipEventRecord(s.done, s.st) != hipSuccess && side_rc == BUG_OK) {
            side_rc = BUG_EDEVICE;
            side_msg = "join event of the parallel half failed";
        }
    });
    try {
        main_fn(ex);
    } catch (...) {
        th.join();
        (void)hipStreamWaitEvent(ex.st, s.done, 0);
        throw;
    }
    th.join();
    if (hipStreamWaitEvent(ex.st, s.done, 0) != hipSuccess) throw StepError(BUG_EDEVICE, "join failed");
    if (side_rc != BUG_OK) throw StepError(side_rc, side_msg);
}

// ---- order-p closure ---------------------------------------------------------------------------------

struct KLFrames {
    std::vector<Tensor> left, right;       // left[0..rl], right[rl..L-2]
    int rl;
    Tensor lenv, renv;
};

struct StepCtx {
    const std::vector<MpoSite>* mpo;
    cplx tau;
    Alloc alloc;
    std::vector<int> wcap;
    int fixed_width;
    double tail_factor;
    bool lanczos;
    double htol;
    bool parallel;
};

std::function<Tensor(const Tensor&)> root_generator(const Exec& ex, const StepCtx& cx, const KLFrames& fr,
                                                    std::vector<Tensor>& keep) {
    const auto& mpo = *cx.mpo;
    Tensor Lr = push_left(ex, fr.lenv, fr.left[fr.rl], mpo[fr.rl], ex.prec);
    Tensor Rr = push_right(ex, fr.renv, fr.right[fr.rl], mpo[fr.rl + 1], ex.prec);
    keep = {Lr, Rr};
    return [ex, Lr, Rr](const Tensor& C) { return apply_zero_site(ex, Lr, Rr, C, ex.prec); };
}

Tensor project_root(const Exec& ex, const MPS& psi, const KLFrames& fr) {
    const int L = psi.size();
    Tensor AL, AR;
    for (int i = 0; i <= fr.rl; ++i) {
        Tensor c = AL.empty() ? psi.t[i] : core_times_site(ex, AL, psi.t[i]);
        AL = overlap_left(ex, fr.left[i], c);
    }
    for (int b = L - 2; b >= fr.rl; --b) {
        Tensor c = AR.empty() ? psi.t[b + 1] : site_times_core(ex, psi.t[b + 1], AR);
        AR = overlap_right(ex, c, fr.right[b]);
    }
    return core_times_core(ex, AL, AR);
}

Tensor polar_frame(const Exec& ex, const Tensor& frame, Side side) {
    const bool left = side == LEFT;
    const int n = left ? frame.L().col.total() : frame.L().row.total();
    auto gram = [&](const Tensor& X) { return left ? overlap_left(ex, X, X) : overlap_right(ex, X, X); };
    Tensor G = gram(frame);
    const double ng = norm(ex, G), nf = norm(ex, frame);
    const double d2 = ng * ng - 2.0 * nf * nf + n;
    if (n > 0 && d2 < kPolarDefect * kPolarDefect) {
        Tensor X = frame;
        for (int it = 0; it < 2; ++it) {
            Tensor Y = left ? site_times_core(ex, X, G) : core_times_site(ex, G, X);
            scale_(ex, -0.5, Y);
            axpy_(ex, 1.5, X, Y);
            X = Y;
            if (it == 0) G = gram(X);
        }
        return X;
    }
    if (left) {
        SVDResult r = factor_left(ex, frame).select(ex, 0.0);
        return site_times_core(ex, r.U, r.Vh);
    }
    SVDResult r = factor_right(ex, frame).select(ex, 0.0);
    return core_times_site(ex, r.U, r.Vh);
}

BondFrame frame_onto_left(const Exec& ex, const Tensor& left_iso, const Tensor& carry, const Tensor& right) {
    Tensor Lr, V0;
    qr_right(ex, right, Lr, V0);
    BondFrame f{left_iso, core_times_core(ex, overlap_left(ex, left_iso, carry), Lr), V0,
                left_iso.L().col.total()};
    return f;
}

BondFrame frame_onto_right(const Exec& ex, const Tensor& left, const Tensor& carry, const Tensor& right_iso) {
    Tensor U0, Rl;
    qr_left(ex, left, U0, Rl);
    BondFrame f{U0, core_times_core(ex, Rl, overlap_right(ex, carry, right_iso)), right_iso,
                right_iso.L().row.total()};
    return f;
}

double sampled_leak(const Exec& ex, const BondFrame& fr, const StepCtx& cx, int bond, const Tensor& lenv,
                    const Tensor& renv, Side side, int count, ProbeStream& stream) {
    const auto& mpo = *cx.mpo;
    const MpoSite &W1 = mpo[bond], &W2 = mpo[bond + 1];
    if (side == LEFT) {
        Tensor Om = graded_probe(ex, fr.V0, RIGHT, count, stream);
        if (Om.empty()) return 0.0;
        Tensor Y = perp_left(ex, fr.U0, stream_left(ex, lenv, W1, W2, renv, site_times_core(ex, fr.U0, fr.S0),
                                                      fr.V0, Om, ex.prec));
        return sample_residual(ex, Tensor(), LEFT, fr.U0, Y, Om);
    }
    Tensor Om = graded_probe(ex, fr.U0, LEFT, count, stream);
    if (Om.empty()) return 0.0;
    Tensor Y = perp_right(ex, fr.V0, stream_right(ex, lenv, W1, W2, renv, fr.U0,
                                                   core_times_site(ex, fr.S0, fr.V0), Om, ex.prec));
    return sample_residual(ex, Tensor(), RIGHT, fr.V0, Y, Om);
}

Tensor enrich_bond(const Exec& ex, const BondFrame& tsf, const StepCtx& cx, int bond, const Tensor& lch,
                   const Tensor& rch, Side side, const Tensor& Fsite, const Tensor& site_core, double amp,
                   Ledger& ledger, ProbeStream& stream) {
    const auto& mpo = *cx.mpo;
    Ledger bl;
    const int cap = cx.wcap[bond];
    Expansion xp = dd_cbe_expand(ex, tsf, mpo, bond, lch, rch, side, cx.alloc, bl, cap, cx.fixed_width,
                                 cx.tail_factor, stream, ex.prec);
    Tensor seed, Hl, Hr, start;
    const MpoSite* Hw;
    if (side == LEFT) {
        seed = site_times_core(ex, site_core, overlap_right(ex, Fsite, xp.V_ex));
        Hl = lch;
        Hw = &mpo[bond];
        Hr = push_right(ex, rch, xp.V_ex, mpo[bond + 1], ex.prec);
        start = xp.U_ex;
    } else {
        seed = core_times_site(ex, overlap_left(ex, xp.U_ex, Fsite), site_core);
        Hl = push_left(ex, lch, xp.U_ex, mpo[bond], ex.prec);
        Hw = &mpo[bond + 1];
        Hr = rch;
        start = xp.V_ex;
    }
    KrylovStats st;
    Tensor frame = krylov_enrich(
        ex, seed, [&](const Tensor& v) { return apply_one_site(ex, Hl, *Hw, Hr, v, ex.prec); }, start, side,
        cx.tau, cx.alloc, cap, cx.lanczos, cx.htol, amp, false, st);
    record_krylov(bl, st);
    bl.n_h1_closure += st.n_h1;
    ledger.merge(bl);
    return frame;
}

struct FrontierHalf {
    std::vector<Tensor> frames;
    Tensor env;
    std::vector<double> leak;
    int added = 0;
    Ledger ledger;
};

void frontier_left(const Exec& ex, MPS Fm, const KLFrames& fr, const StepCtx& cx, double amp,
                   ProbeStream& stream, FrontierHalf& out) {
    const auto& mpo = *cx.mpo;
    const int L = Fm.size(), rl = fr.rl;
    canonical(ex, Fm, 0);
    std::vector<Tensor> renv(L + 1);
    renv[L] = boundary_right(ex, mpo[L - 1], Fm.bond_leg(L));
    for (int i = L - 1; i > 1; --i) renv[i] = push_right(ex, renv[i + 1], Fm.t[i], mpo[i], ex.prec);
    Tensor lch = boundary_left(ex, mpo[0], Fm.bond_leg(0));
    Tensor lenv = lch, carry, Pm;
    out.frames.assign(rl + 1, Tensor());
    out.leak.assign(rl + 1, 0.0);
    for (int b = 0; b <= rl; ++b) {
        Tensor old = Pm.empty() ? fr.left[b] : core_times_site(ex, Pm, fr.left[b]);
        Tensor site_core = carry.empty() ? Fm.t[b] : core_times_site(ex, carry, Fm.t[b]);
        BondFrame tsf = frame_onto_left(ex, old, site_core, Fm.t[b + 1]);
        const Tensor& rch = renv[b + 2];
        const int width = old.L().col.total();
        out.leak[b] = sampled_leak(ex, tsf, cx, b, lch, rch, LEFT, std::max(8, (width + 3) / 4), stream);
        Tensor Wi = cx.alloc.scale * out.leak[b] <= cx.alloc.per_bond
                        ? old
                        : enrich_bond(ex, tsf, cx, b, lch, rch, LEFT, Fm.t[b + 1], site_core, amp, out.ledger, stream);
        out.frames[b] = polar_frame(ex, Wi, LEFT);
        out.added += out.frames[b].L().col.total() - width;
        Pm = overlap_left(ex, out.frames[b], old);
        carry = overlap_left(ex, out.frames[b], site_core);
        if (b == rl)
            lenv = lch;
        else
            lch = push_left(ex, lch, out.frames[b], mpo[b], ex.prec);
    }
    out.env = lenv;
}

void frontier_right(const Exec& ex, MPS Fm, const KLFrames& fr, const StepCtx& cx, double amp,
                    ProbeStream& stream, FrontierHalf& out) {
    const auto& mpo = *cx.mpo;
    const int L = Fm.size(), rl = fr.rl;
    canonical(ex, Fm, L - 1);
    std::vector<Tensor> lenvs{boundary_left(ex, mpo[0], Fm.bond_leg(0))};
    for (int i = 0; i < L - 2; ++i) lenvs.push_back(push_left(ex, lenvs.back(), Fm.t[i], mpo[i], ex.prec));
    Tensor rch = boundary_right(ex, mpo[L - 1], Fm.bond_leg(L));
    Tensor renv = rch, carry, Pm;
    out.frames.assign(L - 1, Tensor());
    out.leak.assign(L - 1, 0.0);
    for (int b = L - 2; b >= rl; --b) {
        Tensor old = Pm.empty() ? fr.right[b] : site_times_core(ex, fr.right[b], Pm);
        Tensor site_core = carry.empty() ? Fm.t[b + 1] : site_times_core(ex, Fm.t[b + 1], carry);
        BondFrame tsf = frame_onto_right(ex, Fm.t[b], site_core, old);
        const Tensor& lch = lenvs[b];
        const int width = old.L().row.total();
        out.leak[b] = sampled_leak(ex, tsf, cx, b, lch, rch, RIGHT, std::max(8, (width + 3) / 4), stream);
        Tensor Zj = cx.alloc.scale * out.leak[b] <= cx.alloc.per_bond
                        ? old
                        : enrich_bond(ex, tsf, cx, b, lch, rch, RIGHT, Fm.t[b], site_core, amp, out.ledger, stream);
        out.frames[b] = polar_frame(ex, Zj, RIGHT);
        out.added += out.frames[b].L().row.total() - width;
        Pm = overlap_right(ex, old, out.frames[b]);
        carry = overlap_right(ex, site_core, out.frames[b]);
        if (b == rl)
            renv = rch;
        else
            rch = push_right(ex, rch, out.frames[b], mpo[b + 1], ex.prec);
    }
    out.env = renv;
}

struct SweepResult {
    int added;
    double leak;
    bool met;
};

SweepResult frontier_sweep(const Exec& ex, KLFrames& fr, const MPS& psi, const Tensor& core, const StepCtx& cx,
                           double nrm, Ledger& ledger, ProbeStream& stream) {
    const int L = psi.size(), rl = fr.rl;
    const double nF = norm(ex, core);
    const double eps = 2.220446049250313e-16;
    if (nF <= kFrontierFloor * std::max(nrm, eps)) return SweepResult{0, 0.0, true};
    MPS Fm;
    Fm.charges = psi.charges;
    for (int i = 0; i < rl; ++i) Fm.t.push_back(fr.left[i]);
    Fm.t.push_back(site_times_core(ex, fr.left[rl], core));
    for (int b = rl; b < L - 1; ++b) Fm.t.push_back(fr.right[b]);
    Fm.center = rl;
    const double amp = nF / std::max(nrm, eps);
    FrontierHalf hl, hr;
    if (cx.parallel) {
        ProbeStream sR = stream.split();
        run_beside(ex, [&](const Exec& e) { frontier_left(e, Fm, fr, cx, amp, stream, hl); },
                   [&](const Exec& e) { frontier_right(e, Fm, fr, cx, amp, sR, hr); });
    } else {
        frontier_left(ex, Fm, fr, cx, amp, stream, hl);
        frontier_right(ex, Fm, fr, cx, amp, stream, hr);
    }
    Ledger lg;
    lg.merge(hl.ledger);
    lg.merge(hr.ledger);
    lg.n_closure_sweeps = 1;
    ledger.merge(lg);
    double lL = 0.0, lR = 0.0;
    for (double x : hl.leak) lL += x * x;
    for (double x : hr.leak) lR += x * x;
    const int added = hl.added + hr.added;
    if (added > 0) {
        for (int b = 0; b <= rl; ++b) fr.left[b] = hl.frames[b];
        for (int b = rl; b < L - 1; ++b) fr.right[b] = hr.frames[b];
        fr.lenv = hl.env;
        fr.renv = hr.env;
    }
    return SweepResult{added, cx.alloc.scale * std::sqrt(lL + lR), lg.tolerance_met};
}

Tensor single_pass_closure(const Exec& ex, KLFrames& fr, const MPS& psi, const Tensor& root_start,
                           const StepCtx& cx, double nrm, Ledger& ledger, int order, ProbeStream& stream) {
    Tensor current = root_start;
    std::vector<Tensor> keep;
    auto op = root_generator(ex, cx, fr, keep);
    std::vector<double> leaks(order, std::numeric_limits<double>::quiet_NaN());
    bool met = true;
    for (int degree = 1; degree < order; ++degree) {
        // F_j = (tau / j) B F_{j-1}, j = 1 .. degree
        Tensor F = current;
        for (int j = 1; j <= degree; ++j) {
            F = op(F);
            scale_(ex, cx.tau / (double)j, F);
        }
        const SweepResult sr = frontier_sweep(ex, fr, psi, F, cx, nrm, ledger, stream);
        leaks[degree] = sr.leak;
        met &= sr.met;
        if (sr.added > 0) {
            op = root_generator(ex, cx, fr, keep);
            current = project_root(ex, psi, fr);
        }
    }
    ledger.leak = leaks;
    double defect = 0.0;
    for (int j = 0; j < order; ++j)
        if (!std::isnan(leaks[j])) defect += leaks[j] / (j + 1);
    ledger.defect = defect;
    if (!met) ledger.fail(BUG_UNMET_CLOSURE);
    return current;
}

// ---- projection miss and compression ---------------------------------------------------------------

double projection_miss(const Exec& ex, const MPS& psi, const std::vector<Tensor>& lf_frames,
                       const std::vector<Tensor>& rf_frames, int rl) {
    const int L = psi.size();
    std::vector<Tensor> rf(L + 1);
    for (int i = L - 1; i > 0; --i) {
        Tensor carried = i == L - 1 ? psi.t[L - 1] : site_times_core(ex, psi.t[i], rf[i + 1]);
        Tensor Q;
        qr_right(ex, carried, rf[i], Q);
    }
    double miss2 = 0.0;
    Tensor AL;
    for (int i = 0; i <= rl; ++i) {
        Tensor carried = AL.empty() ? psi.t[i] : core_times_site(ex, AL, psi.t[i]);
        const double m = norm(ex, site_times_core(ex, perp_left(ex, lf_frames[i], carried), rf[i + 1]));
        miss2 += m * m;
        AL = overlap_left(ex, lf_frames[i], carried);
    }
    std::vector<Tensor> lf(L - 1);
    {
        Tensor Q;
        qr_core(ex, AL, Q, lf[rl]);
    }
    for (int b = rl + 1; b < L - 1; ++b) {
        Tensor Q;
        qr_left(ex, core_times_site(ex, lf[b - 1], psi.t[b]), Q, lf[b]);
    }
    Tensor AR;
    for (int b = L - 2; b >= rl; --b) {
        Tensor carried = AR.empty() ? psi.t[b + 1] : site_times_core(ex, psi.t[b + 1], AR);
        const double m = norm(ex, core_times_site(ex, lf[b], perp_right(ex, rf_frames[b], carried)));
        miss2 += m * m;
        AR = overlap_right(ex, carried, rf_frames[b]);
    }
    return std::sqrt(miss2);
}

// budgeted root-to-leaves compression (truncate_recursive!, truncation.jl:59-158)
void truncate_recursive(const Exec& ex, MPS& psi, int root, int maxdim, double weight_tol,
                        std::vector<double>& bond_weights, double& worst, double& cumulative, bool& capped) {
    const int L = psi.size();
    const double sdiv = std::sqrt((double)root) + std::sqrt((double)(L - 1 - root));
    double nrm2 = -1.0;
    worst = 0.0;
    capped = false;
    double cum[2] = {0.0, 0.0};
    bond_weights.assign(L - 1, 0.0);
    auto cut = [&](const Tensor& t, bool left_split) {
        const BlockSVD fac = left_split ? factor_left(ex, t) : factor_right(ex, t);
        SVDResult res = fac.select(ex, 0.0, maxdim);
        const double tot = sq_weight(res.kept) + sq_weight(res.trunc);
        if (nrm2 < 0) nrm2 = tot;
        const double share = sdiv > 0 ? weight_tol * std::sqrt(std::max(nrm2, 0.0)) / sdiv : 0.0;
        const int nk = (int)res.kept.size();
        int keep;
        const double tail = tail_rank(res.kept, 1.0, share, nk, sq_weight(res.trunc), keep);
        if (tail > share && !res.trunc.empty()) capped = true;
        keep = std::max(keep, 1);
        if (keep < nk) res = fac.select(ex, 0.0, keep);
        return res;
    };
    auto account = [&](const SVDResult& res, int side, int bond) {
        worst = std::max(worst, trunc_weight(res));
        cum[side] += sq_weight(res.trunc);
        bond_weights[bond] = sq_weight(res.trunc);
    };
    for (int site = root; site > 0; --site) {
        SVDResult res = cut(psi.t[site], false);
        account(res, 0, site - 1);
        psi.t[site] = res.Vh;
        psi.t[site - 1] = site_times_core(ex, psi.t[site - 1], diag_scale(ex, res.U, res.S, false));
        psi.center = site - 1;
    }
    canonical(ex, psi, root);
    for (int site = root; site < L - 1; ++site) {
        SVDResult res = cut(psi.t[site], true);
        account(res, 1, site);
        psi.t[site] = res.U;
        psi.t[site + 1] = core_times_site(ex, diag_scale(ex, res.Vh, res.S, true), psi.t[site + 1]);
        psi.center = site + 1;
    }
    cumulative = nrm2 > 0 ? (std::sqrt(cum[0]) + std::sqrt(cum[1])) / std::sqrt(nrm2) : 0.0;
    for (double& w : bond_weights) w = nrm2 > 0 ? std::sqrt(w / nrm2) : 0.0;
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ---- the step --------------------------------------------------------------------------------------

struct StepOut {
    Ledger ledger;
    std::vector<int> expanded;
    int root_link = 0, root_rank = 0, root_applications = 0;
    double discarded = 0.0;
    double seconds[6] = {0, 0, 0, 0, 0, 0};
};

void validate(const bug_step_opts& o, int L, int& rl) {
    if (L < 2) throw StepError(BUG_EINVAL, "cbe_bug_step needs at least two sites");
    rl = o.root < 0 ? std::max(1, std::min(L - 1, L / 2)) - 1 : o.root;
    if (rl < 0 || rl > L - 2) throw StepError(BUG_EINVAL, "root must be a bond in 0:" + std::to_string(L - 2));
    if (o.order < 1) throw StepError(BUG_EINVAL, "order must be >= 1");
    if (o.maxdim < 1) throw StepError(BUG_EINVAL, "maxdim must be positive");
    if (!(o.tol_basis < 0 || o.tol_basis > 0) || !(o.tol_expv < 0 || o.tol_expv > 0))
        throw StepError(BUG_EINVAL, "tolerances must be positive (tol_trunc non-negative)");
    if (o.tol_krylov == 0.0) throw StepError(BUG_EINVAL, "tol_krylov must be positive");
    if (o.work_ceiling_rule < BUG_CEIL_BOND || o.work_ceiling_rule > BUG_CEIL_FRAME_ROOT)
        throw StepError(BUG_EINVAL, "work_ceiling_rule must be one of maxdim, bond, bond_strict, frame_root");
    const double f = o.bond_ceiling_factor;
    const bool dflt = f < 0;
    if (!dflt && !(f == 0.0 || (f > 1.0 && o.work_ceiling_rule != BUG_CEIL_MAXDIM)))
        throw StepError(BUG_EINVAL, "bond_ceiling_factor must be 0, or > 1 with a per-bond work_ceiling_rule");
    if (!(o.root_ceiling_factor > 1.0)) throw StepError(BUG_EINVAL, "root_ceiling_factor must exceed one");
    if (o.precision != BUG_PREC_FP64 && o.precision != BUG_PREC_FP32)
        throw StepError(BUG_EINVAL, "precision must be BUG_PREC_FP64 or BUG_PREC_FP32");
    if (o.mode != 0) throw StepError(BUG_EINVAL, "only the :rank_tail expansion runs through bug_step");
    if (o.fixed_width < 1) throw StepError(BUG_EINVAL, "rank_tail requires positive fixed_width");
    if (!(o.tail_factor > 0 && o.tail_factor <= 1)) throw StepError(BUG_EINVAL, "tail_factor must lie in (0, 1]");
    if (o.krylov_mode < BUG_KRYLOV_AUTO || o.krylov_mode > BUG_KRYLOV_ARNOLDI)
        throw StepError(BUG_EINVAL, "krylov_mode must be auto, lanczos or arnoldi");
    if (o.exact) throw StepError(BUG_EINVAL, "exact = true does not run through bug_step");
    if (o.fold_omega == 0) throw StepError(BUG_EINVAL, "unfolded leaks do not run through bug_step");
    if (o.krylov_rank_tol != 0.0) throw StepError(BUG_EINVAL, "ranked Krylov selection does not run through bug_step");
    if (o.root_method < BUG_ROOT_EXP || o.root_method > BUG_ROOT_RADAU3)
        throw StepError(BUG_EINVAL, "root_method must be exp, trapezoid, radau2 or radau3");
    if (o.root_maxiter < 1) throw StepError(BUG_EINVAL, "root_maxiter must be positive");
    if (o.work_ceiling_rule == BUG_CEIL_BOND_STRICT && !dflt && !(f > 1.0))
        throw StepError(BUG_EINVAL, "strict bond factor must exceed 1");
    if (o.work_ceiling_rule == BUG_CEIL_FRAME_ROOT) {
        const double ff = (dflt || f == 0.0) ? 2.0 : f;
        if (o.root_ceiling_factor > ff) throw StepError(BUG_EINVAL, "root_ceiling_factor must not exceed the frame factor");
    }
}

// What the step has built when it reaches the root solve, and what its last stage needs.
struct RootProblem {
    int rl = 0;
    double t0 = 0.0, te = 0.0, tt = 0.0;
    bool hermitian = true;
    std::vector<int> bd0;
    std::vector<Tensor> left_frames, right_frames;
    Tensor Lr, Rr, start;                  // root environments and the root start core
};

// Stages up to the root solve: options, ceilings, half-sweeps, closure, projection miss, and the
// root generator's environments (bug_step.jl:17-380).
void step_prepare(const Exec& ex, MPS& psi, const std::vector<MpoSite>& mpo, cplx tau,
                  const bug_step_opts& o, StepOut& out, RootProblem& rp) {
    const int L = psi.size();
    int rl = 0;
    validate(o, L, rl);
    if ((int)mpo.size() != L)
        throw StepError(BUG_EINVAL, "MPO has " + std::to_string(mpo.size()) + " sites, state has " + std::to_string(L));
    if (o.work_ceiling_rule == BUG_CEIL_BOND_STRICT && o.bond_ceiling_factor < 0)
        throw StepError(BUG_EINVAL, "strict bond factor must exceed 1");
    const double tb = o.tol_basis < 0 ? 1e-5 : o.tol_basis;
    const double tt = o.tol_trunc < 0 ? 1e-7 : o.tol_trunc;
    const double te = o.tol_expv < 0 ? 1e-10 : o.tol_expv;
    const double tk = o.tol_krylov < 0 ? tb : o.tol_krylov;
    const bool hermitian = o.hermitian != 0;
    const double htol = o.precision == BUG_PREC_FP32 ? 1e-5 : 1e-9;
    const bool lanczos = o.krylov_mode == BUG_KRYLOV_LANCZOS || (o.krylov_mode == BUG_KRYLOV_AUTO && hermitian);
    const int maxdim = o.maxdim;
    const int d = (int)psi.charges.size();
    const double t0 = now();
    ProbeStream stream(o.seed);
    const std::vector<int> bd0 = psi.bond_dims();
    double factor = o.bond_ceiling_factor;
    if (factor < 0) factor = o.work_ceiling_rule == BUG_CEIL_BOND ? 1.2 : 0.0;
    std::vector<int> wcap(L - 1);
    int bdmax = 0;
    for (int b : bd0) bdmax = std::max(bdmax, b);
    for (int b = 0; b < L - 1; ++b) {
        const int r = bd0[b];
        int c;
        switch (o.work_ceiling_rule) {
            case BUG_CEIL_BOND:
                c = factor > 0 ? ceil_round9(factor * d * r) : 2 * d * r;
                c = std::max(std::min(c, maxdim), 1);
                break;
            case BUG_CEIL_BOND_STRICT:
                c = std::max(1, std::min(maxdim, ceil_round9(factor * r)));
                break;
            case BUG_CEIL_FRAME_ROOT: {
                const double f = b == rl ? o.root_ceiling_factor : (factor == 0 ? 2.0 : factor);
                c = std::max(1, std::min(maxdim, ceil_round9(f * r)));
                break;
            }
            default:
                c = std::max({d * maxdim, d * bdmax, 1});
        }
        wcap[b] = c;
    }
    const double nrm = (psi.center >= 0 && psi.center < L) ? norm(ex, psi.t[psi.center]) : 1.0;
    if (!std::isfinite(nrm)) throw StepError(BUG_ENONFINITE, "cbe_bug_step: the state is not finite");
    const Alloc alloc = basis_allocation(tb, L - 1, tau, nrm, o.order, tk);
    Ledger& ledger = out.ledger;
    ledger.r_old = bdmax;
    out.expanded.assign(L - 1, 0);
    std::vector<Tensor> left_frames(rl + 1), right_frames(L - 1);
    MPS psiL = psi;
    Tensor lroot, rroot, AR;

    auto run_left = [&](const Exec& e, ProbeStream& st, Ledger& led) {
        const double ts = now();
        canonical(e, psiL, 0);
        std::vector<Tensor> renvs(L + 1);
        renvs[L] = boundary_right(e, mpo[L - 1], psiL.bond_leg(L));
        for (int i = L - 1; i > 1; --i) renvs[i] = push_right(e, renvs[i + 1], psiL.t[i], mpo[i], e.prec);
        Tensor lch = boundary_left(e, mpo[0], psiL.bond_leg(0));
        Tensor carry;
        for (int b = 0; b <= rl; ++b) {
            Tensor carried = carry.empty() ? psiL.t[b] : core_times_site(e, carry, psiL.t[b]);
            BondFrame fr = frame_from(e, carried, psiL.t[b + 1]);
            Ledger bl;
            Expansion xp = dd_cbe_expand(e, fr, mpo, b, lch, renvs[b + 2], LEFT, alloc, bl, wcap[b],
                                         o.fixed_width, o.tail_factor, st, e.prec);
            Tensor Kex = site_times_core(e, carried, overlap_right(e, psiL.t[b + 1], xp.V_ex));
            Tensor rch1 = push_right(e, renvs[b + 2], xp.V_ex, mpo[b + 1], e.prec);
            renvs[b + 2] = Tensor();
            KrylovStats ks;
            Tensor Wk = krylov_enrich(
                e, Kex, [&](const Tensor& v) { return apply_one_site(e, lch, mpo[b], rch1, v, e.prec); }, xp.U_ex,
                LEFT, tau, alloc, wcap[b], lanczos, htol, 1.0, true, ks);
            record_krylov(bl, ks);
            left_frames[b] = Wk;
            out.expanded[b] = Wk.L().col.total();
            log_bond(led, bl, b, 0, out.expanded[b]);
            carry = overlap_left(e, Wk, carried);
            if (b == rl)
                lroot = lch;
            else
                lch = push_left(e, lch, Wk, mpo[b], e.prec);
        }
        out.seconds[0] = now() - ts;
    };
    auto run_right = [&](const Exec& e, ProbeStream& st, Ledger& led) {
        const double ts = now();
        MPS& psiR = psi;
        canonical(e, psiR, L - 1);
        std::vector<Tensor> lenvs{boundary_left(e, mpo[0], psiR.bond_leg(0))};
        for (int i = 0; i < L - 2; ++i) lenvs.push_back(push_left(e, lenvs.back(), psiR.t[i], mpo[i], e.prec));
        Tensor rch = boundary_right(e, mpo[L - 1], psiR.bond_leg(L));
        Tensor carry;
        for (int b = L - 2; b >= rl; --b) {
            Tensor carried = carry.empty() ? psiR.t[b + 1] : site_times_core(e, psiR.t[b + 1], carry);
            BondFrame fr = frame_from(e, psiR.t[b], carried);
            Ledger bl;
            Expansion xp = dd_cbe_expand(e, fr, mpo, b, lenvs[b], rch, RIGHT, alloc, bl, wcap[b],
                                         o.fixed_width, o.tail_factor, st, e.prec);
            Tensor Bex = core_times_site(e, overlap_left(e, xp.U_ex, psiR.t[b]), carried);
            Tensor lch1 = push_left(e, lenvs[b], xp.U_ex, mpo[b], e.prec);
            lenvs[b] = Tensor();
            KrylovStats ks;
            Tensor Zk = krylov_enrich(
                e, Bex, [&](const Tensor& v) { return apply_one_site(e, lch1, mpo[b + 1], rch, v, e.prec); },
                xp.V_ex, RIGHT, tau, alloc, wcap[b], lanczos, htol, 1.0, true, ks);
            record_krylov(bl, ks);
            log_bond(led, bl, b, 1, Zk.L().row.total());
            right_frames[b] = Zk;
            if (b != rl) out.expanded[b] = Zk.L().row.total();
            carry = overlap_right(e, carried, Zk);
            if (b == rl)
                rroot = rch;
            else
                rch = push_right(e, rch, Zk, mpo[b + 1], e.prec);
        }
        AR = carry;
        out.seconds[1] = now() - ts;
    };

    if (o.parallel) {
        ProbeStream streamR = stream.split();
        Ledger ledL, ledR;
        const std::vector<Tensor> keep = psi.t;          // read by the side stream until the join
        run_beside(ex, [&](const Exec& e) { run_left(e, stream, ledL); },
                   [&](const Exec& e) { run_right(e, streamR, ledR); });
        ledger.merge(ledL);
        ledger.merge(ledR);
    } else {
        run_left(ex, stream, ledger);
        run_right(ex, stream, ledger);
    }

    // ---- root start ----
    Tensor AL;
    for (int i = 0; i <= rl; ++i) {
        Tensor c = AL.empty() ? psi.t[i] : core_times_site(ex, AL, psi.t[i]);
        AL = overlap_left(ex, left_frames[i], c);
    }
    Tensor root_start = core_times_core(ex, AL, AR);

    StepCtx cx{&mpo, tau, alloc, wcap, o.fixed_width, o.tail_factor, lanczos, htol, o.parallel != 0};
    if (o.order > 1) {
        const double ts = now();
        KLFrames fr{left_frames, right_frames, rl, lroot, rroot};
        root_start = single_pass_closure(ex, fr, psi, root_start, cx, nrm, ledger, o.order, stream);
        left_frames = fr.left;
        right_frames = fr.right;
        lroot = fr.lenv;
        rroot = fr.renv;
        for (int b = 0; b < rl; ++b) out.expanded[b] = left_frames[b].L().col.total();
        for (int b = rl + 1; b < L - 1; ++b) out.expanded[b] = right_frames[b].L().row.total();
        out.seconds[2] = now() - ts;
    }
    const double npsi = norm(ex, psi.t[psi.center]);
    int raug = right_frames[rl].L().row.total();
    for (int e : out.expanded) raug = std::max(raug, e);
    ledger.r_aug = raug;
    ledger.proj_err = npsi > 0 ? projection_miss(ex, psi, left_frames, right_frames, rl) / npsi : 0.0;
    if (ledger.proj_err > tb) ledger.fail(BUG_UNMET_PROJECTION);

    // ---- the root generator's environments ----
    rp.rl = rl;
    rp.t0 = t0;
    rp.te = te;
    rp.tt = tt;
    rp.hermitian = hermitian;
    rp.bd0 = bd0;
    rp.Lr = push_left(ex, lroot, left_frames[rl], mpo[rl], ex.prec);
    rp.Rr = push_right(ex, rroot, right_frames[rl], mpo[rl + 1], ex.prec);
    rp.start = root_start;
    rp.left_frames = std::move(left_frames);
    rp.right_frames = std::move(right_frames);
}

// Record the root solve's report on the ledger (bug_step.jl:383-395).
void step_root_report(const RootProblem& rp, const ExpvReport& rep, int napply, double seconds,
                      StepOut& out) {
    out.ledger.expv_err = rep.err;
    out.ledger.expv_iters = rep.iterations;
    if (rep.err > rp.te * (1 + 1e-8)) out.ledger.fail(BUG_UNMET_EXPV);
    out.root_applications = napply;
    out.seconds[3] = seconds;
}

// Assembly of the solved root into the state and the budgeted compression (bug_step.jl:396-510).
void step_finish(const Exec& ex, MPS& psi, RootProblem& rp, const Tensor& result,
                 const bug_step_opts& o, StepOut& out) {
    const int L = psi.size();
    const int rl = rp.rl;
    const int maxdim = o.maxdim;
    const double tt = rp.tt;
    Ledger& ledger = out.ledger;
    const std::vector<Tensor>& left_frames = rp.left_frames;
    const std::vector<Tensor>& right_frames = rp.right_frames;
    const std::vector<int>& bd0 = rp.bd0;
    const double ts = now();
    SVDResult res = factor_core(ex, result).select(ex, 0.0);
    out.expanded[rl] = left_frames[rl].L().col.total();
    double discarded = trunc_weight(res);
    for (int i = 0; i < rl; ++i) psi.t[i] = left_frames[i];
    psi.t[rl] = site_times_core(ex, left_frames[rl], res.U);
    psi.t[rl + 1] = core_times_site(ex, diag_scale(ex, res.Vh, res.S, true), right_frames[rl]);
    for (int b = rl + 1; b < L - 1; ++b) psi.t[b + 1] = right_frames[b];
    psi.center = rl + 1;
    out.root_rank = psi.t[rl].L().col.total();
    std::vector<double> bwts(L - 1, 0.0);
    if (o.truncate) {
        double worst, cumw;
        bool capped;
        truncate_recursive(ex, psi, rl + 1, maxdim, tt, bwts, worst, cumw, capped);
        discarded = std::max(discarded, worst);
        ledger.discarded = cumw;
        ledger.maxdim_reached = capped;
        if (cumw > tt * (1 + 1e-8)) ledger.fail(BUG_UNMET_TRUNC);
    }
    const std::vector<int> bd1 = psi.bond_dims();
    int rf = 0;
    for (int b : bd1) rf = std::max(rf, b);
    ledger.r_final = rf;
    for (BondLogEntry& e : ledger.log) {
        e.r_old = bd0[e.bond];
        e.r_final = bd1[e.bond];
        e.discarded = bwts[e.bond];
    }
    out.seconds[4] = now() - ts;
    out.seconds[5] = now() - rp.t0;
    out.root_link = rl;
    out.discarded = discarded;
}

void run_step(const Exec& ex, MPS& psi, const std::vector<MpoSite>& mpo, cplx tau, const bug_step_opts& o,
              StepOut& out) {
    RootProblem rp;
    step_prepare(ex, psi, mpo, tau, o, out, rp);
    const double ts = now();
    int napply = 0;
    auto apply_root = [&](const Tensor& C) {
        ++napply;
        return apply_zero_site(ex, rp.Lr, rp.Rr, C, ex.prec);
    };
    ExpvReport rep;
    const RootRule rule{o.root_method, o.root_shift != 0};
    Tensor result = rp.hermitian
        ? lanczos_expv(ex, apply_root, tau, rp.start, 1e-15, o.root_maxiter, o.reorth != 0, rp.te, rep, rule)
        : arnoldi_expv(ex, apply_root, tau, rp.start, 1e-15, o.root_maxiter, rp.te, rep, rule);
    step_root_report(rp, rep, napply, now() - ts, out);
    step_finish(ex, psi, rp, result, o, out);
}

}  // namespace host
}  // namespace bug

// ---- C ABI ----------------------------------------------------------------------------------------

struct bug_mpo {
    std::vector<bug::host::MpoSite> sites;
};

struct bug_state {
    bug::host::MPS psi;
    hipStream_t stream = nullptr;
};

namespace {

using namespace bug::host;

// the step's report into the caller's info POD
void fill_info(const StepOut& out, bug_step_info* info) {
    const Ledger& lg = out.ledger;
    info->tolerance_met = lg.tolerance_met ? 1 : 0;
    info->unmet = lg.unmet;
    info->max_sample = lg.max_sample;
    info->max_tail = lg.max_tail;
    info->max_krylov = lg.max_krylov;
    info->max_krylov_c = lg.max_krylov_c;
    info->max_krylov_d = lg.max_krylov_d;
    info->discarded = lg.discarded;
    info->worst_discarded = out.discarded;
    info->expv_err = lg.expv_err;
    info->expv_iters = lg.expv_iters;
    info->proj_err = lg.proj_err;
    info->defect = lg.defect;
    for (int j = 0; j < 4; ++j)
        info->leak[j] = j < (int)lg.leak.size() ? lg.leak[j] : std::numeric_limits<double>::quiet_NaN();
    info->n_probes = lg.n_probes;
    info->n_selected = lg.n_selected;
    info->n_krylov = lg.n_krylov;
    info->n_h1 = lg.n_h1;
    info->n_h1_closure = lg.n_h1_closure;
    info->n_closure_sweeps = lg.n_closure_sweeps;
    info->workcap_reached = lg.workcap_reached;
    info->maxdim_reached = lg.maxdim_reached;
    info->r_old = lg.r_old;
    info->r_aug = lg.r_aug;
    info->r_final = lg.r_final;
    info->root_link = out.root_link;
    info->root_rank = out.root_rank;
    info->root_applications = out.root_applications;
    for (int j = 0; j < 6; ++j) info->seconds[j] = out.seconds[j];
    if (info->expanded)
        for (int b = 0; b < std::min<int>(info->expanded_cap, (int)out.expanded.size()); ++b)
            info->expanded[b] = out.expanded[b];
    info->log_count = 0;
    if (info->log)
        for (const BondLogEntry& e : lg.log) {
            if (info->log_count >= info->log_cap) break;
            bug_bond_log& g = info->log[info->log_count++];
            g.bond = e.bond;
            g.side = e.side;
            g.sample = e.sample;
            g.tail = e.tail;
            g.krylov_c = e.krylov_c;
            g.krylov_d = e.krylov_d;
            g.n_probes = e.n_probes;
            g.n_holdout = e.n_holdout;
            g.n_selected = e.n_selected;
            g.n_targeted = e.n_targeted;
            g.n_krylov = e.n_krylov;
            g.n_h1 = e.n_h1;
            g.workcap = e.workcap;
            g.met = e.met;
            g.unmet = e.unmet;
            g.width = e.width;
            g.r_old = e.r_old;
            g.r_final = e.r_final;
            g.discarded = e.discarded;
        }
}

// ---- many instances stepped together ----------------------------------------------------------------

// Worker contexts and streams of a main context's batched steps (created on first use, kept).
struct BatchWorker {
    bug_ctx* ctx = nullptr;
    hipStream_t st = nullptr;
};
std::mutex g_batch_mu;
std::map<bug_ctx*, std::vector<BatchWorker>> g_batch;

std::vector<BatchWorker> batch_workers(bug_ctx* main, int n) {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    std::vector<BatchWorker>& ws = g_batch[main];
    while ((int)ws.size() < n) {
        BatchWorker w;
        if (bug_ctx_create(main->device, &w.ctx) != BUG_OK)
            throw StepError(BUG_EDEVICE, "batch worker context creation failed");
        if (hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking) != hipSuccess)
            throw StepError(BUG_EDEVICE, "batch worker stream creation failed");
        ws.push_back(w);
    }
    return std::vector<BatchWorker>(ws.begin(), ws.begin() + n);
}

// f(i, exec) for every instance i < k on `workers` threads, each with its own context and stream;
// every thread drains its stream before it takes the next instance and before it returns.
template <typename F>
void on_workers(bug_ctx* main, int k, int workers, int prec, F&& f) {
    const int n = std::max(1, std::min(workers, k));
    const std::vector<BatchWorker> ws = batch_workers(main, n);
    std::atomic<int> next{0};
    std::vector<int> rc(n, BUG_OK);
    std::vector<std::string> msg(n);
    std::vector<std::thread> th;
    for (int w = 0; w < n; ++w)
        th.emplace_back([&, w] {
            (void)hipSetDevice(main->device);
            set_thread_stream(ws[w].st);
            const Exec e{ws[w].ctx, ws[w].st, prec};
            for (int i = next.fetch_add(1); i < k && rc[w] == BUG_OK; i = next.fetch_add(1)) {
                try {
                    f(i, e);
                    STEP_HIP(hipStreamSynchronize(ws[w].st));
                } catch (const StepError& x) {
                    rc[w] = x.code;
                    msg[w] = "instance " + std::to_string(i) + ": " + x.what();
                } catch (const bug::Error& x) {
                    rc[w] = x.code;
                    msg[w] = "instance " + std::to_string(i) + ": " + x.what();
                } catch (const std::exception& x) {
                    rc[w] = BUG_EDEVICE;
                    msg[w] = "instance " + std::to_string(i) + ": " + x.what();
                }
                if (rc[w] != BUG_OK) next.store(k);     // the others stop after their instance
            }
            (void)hipStreamSynchronize(ws[w].st);
        });
    for (auto& t : th) t.join();
    for (int w = 0; w < n; ++w)
        if (rc[w] != BUG_OK) throw StepError(rc[w], msg[w]);
}

// K steps with one option set: each instance's stages up to the root solve on the worker threads,
// every root exponential in one lockstep recurrence (grouped zero-site applies, segmented
// Gram-Schmidt) on the caller's stream, then each instance's assembly and compression.
void run_step_batch(bug_ctx* ctx, hipStream_t st, std::vector<MPS>& work,
                    const std::vector<const std::vector<MpoSite>*>& mpos, const std::vector<cplx>& taus,
                    const bug_step_opts& o, std::vector<StepOut>& outs, int workers) {
    const int k = (int)work.size();
    std::vector<RootProblem> rps(k);
    outs.assign(k, StepOut());
    on_workers(ctx, k, workers, o.precision, [&](int i, const Exec& e) {
        step_prepare(e, work[i], *mpos[i], taus[i], o, outs[i], rps[i]);
    });
    // ---- the K root exponentials in lockstep ----
    set_thread_stream(st);
    const Exec ex{ctx, st, o.precision};
    const double ts = now();
    PackedSegs sg;
    for (int i = 0; i < k; ++i) {
        sg.table.push_back(sg.n);
        sg.table.push_back(rps[i].start.nelem());
        sg.n += rps[i].start.nelem();
    }
    DevBuf v0(sizeof(double2) * (size_t)std::max<int64_t>(sg.n, 1)), res(sizeof(double2) * (size_t)std::max<int64_t>(sg.n, 1));
    for (int i = 0; i < k; ++i)
        if (rps[i].start.nelem() > 0)
            STEP_HIP(hipMemcpyAsync(static_cast<double2*>(v0.p) + sg.table[2 * i], rps[i].start.data(),
                                    sizeof(double2) * (size_t)rps[i].start.nelem(), hipMemcpyDeviceToDevice, st));
    std::vector<bug_bsm> Ld(k), Rd(k), Cd(k), Yd(k);
    for (int i = 0; i < k; ++i) {
        Ld[i] = rps[i].Lr.desc();
        Rd[i] = rps[i].Rr.desc();
        Cd[i] = rps[i].start.desc();
        Yd[i] = Cd[i];
    }
    auto apply = [&](const double2* x, double2* y) {
        for (int i = 0; i < k; ++i) {
            Cd[i].data = const_cast<double2*>(x) + sg.table[2 * i];
            Yd[i].data = y + sg.table[2 * i];
        }
        ex.check(bug_apply_zero_site_batch(ctx, k, Ld.data(), Rd.data(), Cd.data(), Yd.data(), nullptr,
                                           o.precision, st));
    };
    std::vector<ExpvReport> reps;
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    expv_lockstep(ex, sg, apply, taus, o.hermitian != 0, static_cast<const double2*>(v0.p), 1e-15,
                  o.root_maxiter, o.reorth != 0, rps[0].te, static_cast<double2*>(res.p), reps, hw,
                  RootRule{o.root_method, o.root_shift != 0});
    std::vector<Tensor> results(k);
    for (int i = 0; i < k; ++i) {
        results[i] = t_empty(ex, rps[i].start.L());
        if (rps[i].start.nelem() > 0)
            STEP_HIP(hipMemcpyAsync(results[i].data(), static_cast<double2*>(res.p) + sg.table[2 * i],
                                    sizeof(double2) * (size_t)rps[i].start.nelem(), hipMemcpyDeviceToDevice, st));
    }
    STEP_HIP(hipStreamSynchronize(st));
    const double troot = now() - ts;
    for (int i = 0; i < k; ++i) step_root_report(rps[i], reps[i], reps[i].iterations, troot, outs[i]);
    on_workers(ctx, k, workers, o.precision, [&](int i, const Exec& e) {
        step_finish(e, work[i], rps[i], results[i], o, outs[i]);
    });
}

int fail_with(bug_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

template <typename F>
int step_guard(bug_ctx* ctx, F&& f) {
    if (!ctx) return BUG_EINVAL;
    try {
        f();
        return BUG_OK;
    } catch (const StepError& e) {
        return fail_with(ctx, e.code, e.what());
    } catch (const bug::Error& e) {
        return fail_with(ctx, e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail_with(ctx, BUG_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail_with(ctx, BUG_EDEVICE, e.what());
    }
}

Leg leg_from(const bug_leg& l);

// The caller's family as a tensor: a view when its blocks sit where the library's packing puts
// them, else a gathered copy.
Tensor tensor_in(const Exec& ex, const bug_bsm& b) {
    if (!b.data || b.nslice < 1) throw StepError(BUG_EINVAL, "family without data or slices");
    const Layout lay(leg_from(b.row), leg_from(b.col), std::vector<int32_t>(b.shift, b.shift + b.nslice));
    bool same = b.nelem >= lay.nelem;
    std::vector<int64_t> src(lay.blks.size(), -1);
    for (size_t i = 0; i < lay.blks.size(); ++i) {
        const Blk& blk = lay.blks[i];
        int idx = -1;
        for (int r = 0; r < b.row.nsec; ++r)
            if (b.row.charge[r] == blk.q) idx = r;
        src[i] = idx < 0 ? -1 : b.offset[(size_t)blk.k * b.row.nsec + idx];
        if (src[i] >= 0 && src[i] + (int64_t)blk.d * blk.dc > b.nelem)
            throw StepError(BUG_EINVARIANT, "block outside its family's data");
        same &= src[i] == blk.off;
    }
    if (same) {
        Tensor t;
        t.lay = std::make_shared<const Layout>(lay);
        // the view does not own the caller's memory: the deleter detaches it before the buffer
        // is destroyed, so it never reaches the pool
        t.buf = std::shared_ptr<DevBuf>(new DevBuf(0), [](DevBuf* d) {
            d->p = nullptr;
            delete d;
        });
        dev_free(t.buf->p, t.buf->bytes);
        t.buf->p = b.data;
        t.buf->bytes = sizeof(double2) * (size_t)b.nelem;
        return t;
    }
    Tensor t = t_zeros(ex, lay);
    CopyList cl;
    for (size_t i = 0; i < lay.blks.size(); ++i)
        if (src[i] >= 0) cl.add(src[i], lay.blks[i].d, lay.blks[i].off, lay.blks[i].d, lay.blks[i].d, lay.blks[i].dc);
    cl.run(ex, b.data, b.nelem, t.buf->p, t.nelem());
    return t;
}

// Scatter `t` into the caller's family (same legs and slices required).
void tensor_out(const Exec& ex, const Tensor& t, const bug_bsm& b) {
    const Layout want(leg_from(b.row), leg_from(b.col), std::vector<int32_t>(b.shift, b.shift + b.nslice));
    if (want != t.L()) throw StepError(BUG_EINVAL, "output family has the wrong legs or slices");
    CopyList cl;
    for (const Blk& blk : t.L().blks) {
        int idx = -1;
        for (int r = 0; r < b.row.nsec; ++r)
            if (b.row.charge[r] == blk.q) idx = r;
        const int64_t dst = idx < 0 ? -1 : b.offset[(size_t)blk.k * b.row.nsec + idx];
        if (dst < 0) throw StepError(BUG_EINVARIANT, "output family lacks a block of the result");
        if (dst + (int64_t)blk.d * blk.dc > b.nelem) throw StepError(BUG_EINVARIANT, "output block outside its data");
        cl.add(blk.off, blk.d, dst, blk.d, blk.d, blk.dc);
    }
    cl.run(ex, t.buf->p, t.nelem(), b.data, b.nelem);
}

MpoSite mpo_in(const bug_mpo_site& w) {
    MpoSite s;
    s.cl.assign(w.cl, w.cl + w.wl);
    s.cr.assign(w.cr, w.cr + w.wr);
    s.a.assign(w.a, w.a + w.nent);
    s.b.assign(w.b, w.b + w.nent);
    s.sout.assign(w.sout, w.sout + w.nent);
    s.sin.assign(w.sin, w.sin + w.nent);
    s.coeff.assign(w.coeff, w.coeff + 2 * w.nent);
    return s;
}

Leg leg_from(const bug_leg& l) {
    Leg out;
    for (int i = 0; i < l.nsec; ++i) {
        if (i > 0 && l.charge[i] <= l.charge[i - 1])
            throw StepError(BUG_EINVAL, "leg charges must be strictly increasing");
        if (l.dim[i] < 0) throw StepError(BUG_EINVAL, "negative sector dimension");
        if (l.dim[i] == 0) continue;
        out.q.push_back(l.charge[i]);
        out.d.push_back(l.dim[i]);
    }
    return out;
}

}  // namespace

extern "C" {

void bug_step_opts_default(bug_step_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->order = 1;
    o->maxdim = 128;
    o->hermitian = 1;
    o->work_ceiling_rule = BUG_CEIL_BOND;
    o->bond_ceiling_factor = -1.0;
    o->root_ceiling_factor = 1.2;
    o->tol_basis = -1.0;
    o->tol_trunc = -1.0;
    o->tol_expv = -1.0;
    o->tol_krylov = -1.0;
    o->root = -1;
    o->reorth = 1;
    o->truncate = 1;
    o->precision = BUG_PREC_FP64;
    o->parallel = 0;
    o->seed = 0x5EED;
    o->fixed_width = 8;
    o->tail_factor = 0.10;
    o->krylov_mode = BUG_KRYLOV_AUTO;
    o->mode = 0;
    o->exact = 0;
    o->fold_omega = -1;
    o->krylov_rank_tol = 0.0;
    o->root_method = BUG_ROOT_EXP;
    o->root_shift = 1;
    o->root_maxiter = kRootMaxIter;
}

int bug_mpo_create(int32_t L, const bug_mpo_site* sites, bug_mpo** out) {
    if (!out || L < 1 || !sites) return BUG_EINVAL;
    try {
        bug_mpo* m = new bug_mpo();
        for (int i = 0; i < L; ++i) {
            const bug_mpo_site& w = sites[i];
            if (w.wl < 1 || w.wr < 1 || w.nent < 0) {
                delete m;
                return BUG_EINVAL;
            }
            MpoSite s;
            s.cl.assign(w.cl, w.cl + w.wl);
            s.cr.assign(w.cr, w.cr + w.wr);
            s.a.assign(w.a, w.a + w.nent);
            s.b.assign(w.b, w.b + w.nent);
            s.sout.assign(w.sout, w.sout + w.nent);
            s.sin.assign(w.sin, w.sin + w.nent);
            s.coeff.assign(w.coeff, w.coeff + 2 * w.nent);
            for (int e = 0; e < w.nent; ++e)
                if (s.a[e] < 0 || s.a[e] >= w.wl || s.b[e] < 0 || s.b[e] >= w.wr) {
                    delete m;
                    return BUG_EINVAL;
                }
            m->sites.push_back(std::move(s));
        }
        *out = m;
        return BUG_OK;
    } catch (...) {
        return BUG_ENOMEM;
    }
}

int bug_mpo_destroy(bug_mpo* mpo) {
    delete mpo;
    return BUG_OK;
}

int bug_state_create(bug_ctx* ctx, int32_t L, int32_t d, const int32_t* local_charges,
                     const bug_bsm* sites, int32_t center, void* stream, bug_state** out) {
    return step_guard(ctx, [&] {
        if (!out || L < 2 || d < 1 || !local_charges || !sites || center < 0 || center >= L)
            throw StepError(BUG_EINVAL, "bad state arguments");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = static_cast<hipStream_t>(stream);
        set_thread_stream(st);
        const Exec ex{ctx, st, BUG_PREC_FP64};
        std::unique_ptr<bug_state> s(new bug_state());
        s->stream = st;
        s->psi.charges.assign(local_charges, local_charges + d);
        s->psi.center = center;
        for (int i = 0; i < L; ++i) {
            const bug_bsm& b = sites[i];
            if (b.nslice != d) throw StepError(BUG_EINVAL, "site " + std::to_string(i) + " has the wrong number of slices");
            for (int k = 0; k < d; ++k)
                if (b.shift[k] != local_charges[k]) throw StepError(BUG_EINVAL, "slice shifts differ from the local charges");
            const Layout lay(leg_from(b.row), leg_from(b.col), s->psi.charges);
            if (i > 0 && lay.row != s->psi.t[i - 1].L().col)
                throw StepError(BUG_EINVAL, "bond " + std::to_string(i) + ": legs of neighbouring sites differ");
            std::vector<cplx> host((size_t)std::max<int64_t>(lay.nelem, 1), 0.0);
            const cplx* src = static_cast<const cplx*>(b.data);
            for (const Blk& blk : lay.blks) {
                // the caller's block at (slice k, row charge q)
                int idx = -1;
                for (int r = 0; r < b.row.nsec; ++r)
                    if (b.row.charge[r] == blk.q) idx = r;
                const int64_t so = idx < 0 ? -1 : b.offset[(size_t)blk.k * b.row.nsec + idx];
                if (so < 0) continue;
                if (so + (int64_t)blk.d * blk.dc > b.nelem) throw StepError(BUG_EINVARIANT, "site block outside its data");
                std::memcpy(&host[(size_t)blk.off], src + so, sizeof(cplx) * (size_t)blk.d * blk.dc);
            }
            s->psi.t.push_back(t_from_host(ex, lay, host));
        }
        *out = s.release();
    });
}

int bug_state_create_device(bug_ctx* ctx, int32_t L, int32_t d, const int32_t* local_charges,
                            const bug_bsm* sites, int32_t center, void* stream, bug_state** out) {
    return step_guard(ctx, [&] {
        if (!out || L < 2 || d < 1 || !local_charges || !sites || center < 0 || center >= L)
            throw StepError(BUG_EINVAL, "bad state arguments");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = static_cast<hipStream_t>(stream);
        set_thread_stream(st);
        const Exec ex{ctx, st, BUG_PREC_FP64};
        std::unique_ptr<bug_state> s(new bug_state());
        s->stream = st;
        s->psi.charges.assign(local_charges, local_charges + d);
        s->psi.center = center;
        for (int i = 0; i < L; ++i) {
            const bug_bsm& b = sites[i];
            if (b.nslice != d) throw StepError(BUG_EINVAL, "site " + std::to_string(i) + " has the wrong number of slices");
            for (int k = 0; k < d; ++k)
                if (b.shift[k] != local_charges[k]) throw StepError(BUG_EINVAL, "slice shifts differ from the local charges");
            const Layout lay(leg_from(b.row), leg_from(b.col), s->psi.charges);
            if (i > 0 && lay.row != s->psi.t[i - 1].L().col)
                throw StepError(BUG_EINVAL, "bond " + std::to_string(i) + ": legs of neighbouring sites differ");
            Tensor t = t_zeros(ex, lay);
            CopyList cl;
            for (const Blk& blk : lay.blks) {
                int idx = -1;
                for (int r = 0; r < b.row.nsec; ++r)
                    if (b.row.charge[r] == blk.q) idx = r;
                const int64_t so = idx < 0 ? -1 : b.offset[(size_t)blk.k * b.row.nsec + idx];
                if (so < 0) continue;
                cl.add(so, blk.d, blk.off, blk.d, blk.d, blk.dc);
            }
            cl.run(ex, b.data, b.nelem, t.buf->p, t.nelem());
            s->psi.t.push_back(t);
        }
        *out = s.release();
    });
}

int bug_state_clone(const bug_state* src, bug_state** out) {
    if (!src || !out) return BUG_EINVAL;
    try {
        bug_state* s = new bug_state();
        s->psi = src->psi;
        s->stream = src->stream;
        *out = s;
        return BUG_OK;
    } catch (...) {
        return BUG_ENOMEM;
    }
}

int bug_state_expect_diagonal(bug_ctx* ctx, const bug_state* state, const double* diag, double* out) {
    return step_guard(ctx, [&] {
        if (!state || !diag || !out) throw StepError(BUG_EINVAL, "bad expectation arguments");
        HIP_CHECK(hipSetDevice(ctx->device));
        set_thread_stream(state->stream);
        const Exec ex{ctx, state->stream, BUG_PREC_FP64};
        MPS work = state->psi;
        canonical(ex, work, 0);
        const int d = (int)work.charges.size();
        for (int j = 0; j < work.size(); ++j) {
            if (j > 0) canonical(ex, work, j);
            const Tensor& A = work.t[j];
            std::vector<bug_copy> blks;
            std::vector<int> slices;
            for (const Blk& b : A.L().blks) {
                blks.push_back(bug_copy{b.off, 0, std::max(b.d, 1), 1, b.d, b.dc, 0, 0, 0});
                slices.push_back(b.k);
            }
            const std::vector<double> w = block_sqnorms(ex, blks, A.buf->p, A.nelem());
            std::vector<double> per(d, 0.0);
            for (size_t i = 0; i < w.size(); ++i) per[slices[i]] += w[i];
            double tot = 0.0, num = 0.0;
            for (int s = 0; s < d; ++s) {
                tot += per[s];
                num += per[s] * diag[s];
            }
            out[j] = tot > 0 ? num / tot : 0.0;
        }
    });
}

int bug_two_site_sketch_left(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W1,
                             const bug_mpo_site* W2, const bug_bsm* R, const bug_bsm* A,
                             const bug_bsm* V, const bug_bsm* Om, const bug_bsm* out,
                             int32_t precision, void* stream) {
    return step_guard(ctx, [&] {
        if (!L || !W1 || !W2 || !R || !A || !V || !Om || !out) throw StepError(BUG_EINVAL, "null argument");
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32) throw StepError(BUG_EINVAL, "unknown precision");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = static_cast<hipStream_t>(stream);
        set_thread_stream(st);
        const Exec ex{ctx, st, precision};
        Tensor y = stream_left(ex, tensor_in(ex, *L), mpo_in(*W1), mpo_in(*W2), tensor_in(ex, *R),
                               tensor_in(ex, *A), tensor_in(ex, *V), tensor_in(ex, *Om), precision);
        tensor_out(ex, y, *out);
    });
}

int bug_two_site_sketch_right(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W1,
                              const bug_mpo_site* W2, const bug_bsm* R, const bug_bsm* U,
                              const bug_bsm* B, const bug_bsm* Om, const bug_bsm* out,
                              int32_t precision, void* stream) {
    return step_guard(ctx, [&] {
        if (!L || !W1 || !W2 || !R || !U || !B || !Om || !out) throw StepError(BUG_EINVAL, "null argument");
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32) throw StepError(BUG_EINVAL, "unknown precision");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = static_cast<hipStream_t>(stream);
        set_thread_stream(st);
        const Exec ex{ctx, st, precision};
        Tensor y = stream_right(ex, tensor_in(ex, *L), mpo_in(*W1), mpo_in(*W2), tensor_in(ex, *R),
                                tensor_in(ex, *U), tensor_in(ex, *B), tensor_in(ex, *Om), precision);
        tensor_out(ex, y, *out);
    });
}

int bug_expv(bug_ctx* ctx, bug_apply_fn apply, void* user, int64_t n, double tau_re, double tau_im,
             const void* v, void* out, int32_t hermitian, double tol, int32_t maxiter, int32_t reorth,
             double conv_tol, bug_expv_report* report, void* stream) {
    return step_guard(ctx, [&] {
        if (!apply || !v || !out || n < 1 || maxiter < 1 || maxiter > 64 || !(tol >= 0))
            throw StepError(BUG_EINVAL, "bad expv arguments");
        if (v == out) throw StepError(BUG_EINVAL, "expv output may not alias its input");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = static_cast<hipStream_t>(stream);
        set_thread_stream(st);
        const Exec ex{ctx, st, BUG_PREC_FP64};
        const Layout lay(Leg::dense((int)std::min<int64_t>(n, INT32_MAX)), Leg::dense(1), {0});
        if (lay.nelem != n) throw StepError(BUG_EINVAL, "expv vectors hold at most 2^31 - 1 elements");
        Tensor v0 = t_empty(ex, lay);
        HIP_CHECK(hipMemcpyAsync(v0.buf->p, v, sizeof(double2) * (size_t)n, hipMemcpyDeviceToDevice, st));
        auto act = [&](const Tensor& x) {
            Tensor y = t_empty(ex, lay);
            const int rc = apply(user, x.buf->p, y.buf->p, stream);
            if (rc != 0) throw StepError(BUG_EINVAL, "the apply callback returned " + std::to_string(rc));
            return y;
        };
        ExpvReport rep;
        const cplx tau(tau_re, tau_im);
        Tensor r = hermitian ? lanczos_expv(ex, act, tau, v0, tol, maxiter, reorth != 0, conv_tol, rep)
                             : arnoldi_expv(ex, act, tau, v0, tol, maxiter, conv_tol, rep);
        HIP_CHECK(hipMemcpyAsync(out, r.buf->p, sizeof(double2) * (size_t)n, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (report) {
            report->iterations = rep.iterations;
            report->substeps = rep.substeps;
            report->err = rep.err;
        }
    });
}

int bug_svd_select(int32_t nsec, const int32_t* lens, const double* spectra, double cutoff,
                   int32_t nkeep, int32_t* keep, int32_t* order, int32_t* nkept, double* weights) {
    if (nsec < 0 || (nsec > 0 && (!lens || !keep)) || !nkept || !(cutoff >= 0)) return BUG_EINVAL;
    std::vector<std::vector<double>> spec(nsec);
    std::vector<int> tags(nsec);
    int64_t pos = 0;
    for (int i = 0; i < nsec; ++i) {
        if (lens[i] < 0 || (lens[i] > 0 && !spectra)) return BUG_EINVAL;
        spec[i].assign(spectra + pos, spectra + pos + lens[i]);
        for (double sv : spec[i])
            if (!std::isfinite(sv)) return BUG_ENONFINITE;
        pos += lens[i];
        tags[i] = i;
    }
    std::vector<SvEntry> full;
    const int n = global_selection(spec, tags, cutoff, nkeep, full);
    double wk = 0.0, wt = 0.0;
    for (int i = 0; i < nsec; ++i) keep[i] = 0;
    for (int e = 0; e < (int)full.size(); ++e) {
        if (e < n) {
            keep[full[e].q] += 1;
            wk += full[e].s * full[e].s;
        } else {
            wt += full[e].s * full[e].s;
        }
        if (order) {
            order[2 * e] = full[e].q;
            order[2 * e + 1] = full[e].j;
        }
    }
    *nkept = n;
    if (weights) {
        weights[0] = wk;
        weights[1] = wt;
    }
    return BUG_OK;
}

int bug_state_destroy(bug_state* state) {
    if (!state) return BUG_OK;
    set_thread_stream(state->stream);
    delete state;
    return BUG_OK;
}

int bug_state_info(const bug_state* state, int32_t* L, int32_t* d, int32_t* center) {
    if (!state) return BUG_EINVAL;
    if (L) *L = state->psi.size();
    if (d) *d = (int32_t)state->psi.charges.size();
    if (center) *center = state->psi.center;
    return BUG_OK;
}

int bug_state_site_shape(const bug_state* state, int32_t site, int32_t* row_nsec, int32_t* col_nsec,
                         int64_t* nelem) {
    if (!state || site < 0 || site >= state->psi.size()) return BUG_EINVAL;
    const Layout& l = state->psi.t[site].L();
    if (row_nsec) *row_nsec = l.row.nsec();
    if (col_nsec) *col_nsec = l.col.nsec();
    if (nelem) *nelem = l.nelem;
    return BUG_OK;
}

int bug_state_download_site(bug_ctx* ctx, const bug_state* state, int32_t site, int32_t* row_charge,
                            int32_t* row_dim, int32_t* col_charge, int32_t* col_dim, int64_t* offset,
                            void* host_data) {
    return step_guard(ctx, [&] {
        if (!state || site < 0 || site >= state->psi.size()) throw StepError(BUG_EINVAL, "bad site");
        HIP_CHECK(hipSetDevice(ctx->device));
        const Tensor& t = state->psi.t[site];
        const Layout& l = t.L();
        if (row_charge) std::memcpy(row_charge, l.row.q.data(), sizeof(int32_t) * l.row.q.size());
        if (row_dim) std::memcpy(row_dim, l.row.d.data(), sizeof(int32_t) * l.row.d.size());
        if (col_charge) std::memcpy(col_charge, l.col.q.data(), sizeof(int32_t) * l.col.q.size());
        if (col_dim) std::memcpy(col_dim, l.col.d.data(), sizeof(int32_t) * l.col.d.size());
        if (offset) std::memcpy(offset, l.off.data(), sizeof(int64_t) * l.off.size());
        if (host_data && l.nelem > 0) {
            HIP_CHECK(hipMemcpyAsync(host_data, t.buf->p, sizeof(double2) * (size_t)l.nelem,
                                     hipMemcpyDeviceToHost, state->stream));
            HIP_CHECK(hipStreamSynchronize(state->stream));
        }
    });
}

int bug_step(bug_ctx* ctx, bug_state* state, const bug_mpo* mpo, double tau_re, double tau_im,
             const bug_step_opts* opts, bug_step_info* info) {
    return step_guard(ctx, [&] {
        if (!state || !mpo || !opts) throw StepError(BUG_EINVAL, "bug_step needs a state, an MPO and options");
        HIP_CHECK(hipSetDevice(ctx->device));
        set_thread_stream(state->stream);
        const Exec ex{ctx, state->stream, opts->precision};
        int rl = 0;
        validate(*opts, state->psi.size(), rl);
        if ((int)mpo->sites.size() != state->psi.size())
            throw StepError(BUG_EINVAL, "MPO and state lengths differ");
        // the step works on a copy of the tensor list: on an error the state is left as it was
        MPS work = state->psi;
        StepOut out;
        run_step(ex, work, mpo->sites, cplx(tau_re, tau_im), *opts, out);
        state->psi = work;
        if (info) fill_info(out, info);
    });
}

int bug_step_batch(bug_ctx* ctx, int32_t k, bug_state* const* states, const bug_mpo* const* mpos,
                   const double* tau, const bug_step_opts* opts, bug_step_info* infos, int32_t workers) {
    return step_guard(ctx, [&] {
        if (k < 0 || (k > 0 && (!states || !mpos || !tau || !opts)))
            throw StepError(BUG_EINVAL, "bug_step_batch needs k states, k MPOs, 2k tau values and options");
        if (k == 0) return;
        if (workers < 1) throw StepError(BUG_EINVAL, "bug_step_batch needs at least one worker");
        HIP_CHECK(hipSetDevice(ctx->device));
        std::vector<MPS> work;
        std::vector<const std::vector<MpoSite>*> ms;
        std::vector<cplx> taus;
        for (int i = 0; i < k; ++i) {
            if (!states[i] || !mpos[i]) throw StepError(BUG_EINVAL, "null state or MPO in the batch");
            for (int j = 0; j < i; ++j)
                if (states[j] == states[i]) throw StepError(BUG_EINVAL, "a state appears twice in the batch");
            int rl = 0;
            validate(*opts, states[i]->psi.size(), rl);
            if ((int)mpos[i]->sites.size() != states[i]->psi.size())
                throw StepError(BUG_EINVAL, "MPO and state lengths differ at instance " + std::to_string(i));
            work.push_back(states[i]->psi);
            ms.push_back(&mpos[i]->sites);
            taus.emplace_back(tau[2 * i], tau[2 * i + 1]);
        }
        // every instance's pending work is ordered before the workers and the lockstep stream
        for (int i = 0; i < k; ++i) HIP_CHECK(hipStreamSynchronize(states[i]->stream));
        std::vector<StepOut> outs;
        run_step_batch(ctx, states[0]->stream, work, ms, taus, *opts, outs, workers);
        for (int i = 0; i < k; ++i) {
            states[i]->psi = work[i];
            if (infos) fill_info(outs[i], &infos[i]);
        }
    });
}

}  // extern "C"
