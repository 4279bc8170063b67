// Host tensor algebra of the step-level entry point (see step_core.h).
#include <limits>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "step_core.h"

namespace bug {
namespace host {

void Exec::check(int rc) const {
    if (rc != BUG_OK) throw StepError(rc, bug_last_error(ctx));
}

// ---- legs and layouts ---------------------------------------------------------------------------

int Leg::total() const {
    int t = 0;
    for (int x : d) t += x;
    return t;
}

int Leg::index(int c) const {
    auto it = std::lower_bound(q.begin(), q.end(), c);
    return (it != q.end() && *it == c) ? (int)(it - q.begin()) : -1;
}

int Leg::dim_of(int c) const {
    const int i = index(c);
    return i < 0 ? 0 : d[i];
}

Leg Leg::make(const std::vector<std::pair<int, int>>& pairs) {
    std::map<int, int> acc;
    for (auto& p : pairs)
        if (p.second > 0) acc[p.first] += p.second;
    Leg l;
    for (auto& e : acc) {
        l.q.push_back(e.first);
        l.d.push_back(e.second);
    }
    return l;
}

Leg Leg::dense(int n) {
    Leg l;
    l.q = {0};
    l.d = {n};
    return l;
}

std::map<int, int> Leg::as_map() const {
    std::map<int, int> m;
    for (int i = 0; i < nsec(); ++i) m[q[i]] = d[i];
    return m;
}

Layout::Layout(Leg r, Leg c, std::vector<int32_t> sh)
    : row(std::move(r)), col(std::move(c)), shifts(std::move(sh)) {
    const int ns = row.nsec();
    off.assign((size_t)shifts.size() * ns, -1);
    int64_t pos = 0;
    for (int k = 0; k < (int)shifts.size(); ++k)
        for (int i = 0; i < ns; ++i) {
            const int q = row.q[i], d = row.d[i];
            const int dc = col.dim_of(q + shifts[k]);
            if (d == 0 || dc == 0) continue;
            off[(size_t)k * ns + i] = pos;
            blks.push_back(Blk{k, q, q + shifts[k], d, dc, pos});
            pos += (int64_t)d * dc;
        }
    nelem = pos;
}

const Blk* Layout::find(int k, int q) const {
    const int i = row.index(q);
    if (i < 0 || k < 0 || k >= nslice()) return nullptr;
    const int64_t o = off[(size_t)k * row.nsec() + i];
    if (o < 0) return nullptr;
    // blocks are stored in (k, i) order: binary search by offset
    auto it = std::lower_bound(blks.begin(), blks.end(), o,
                               [](const Blk& b, int64_t v) { return b.off < v; });
    return &*it;
}

// ---- device memory --------------------------------------------------------------------------------

namespace {
thread_local hipStream_t g_thread_stream = nullptr;
std::mutex g_pool_mu;
// (stream, rounded size) -> free blocks
std::unordered_map<void*, std::unordered_map<size_t, std::vector<void*>>> g_pool;

size_t round_bytes(size_t b) {
    b = std::max<size_t>(b, 256);
    if (b <= ((size_t)1 << 20)) return (b + 4095) & ~size_t(4095);
    // above 1 MiB: steps of 1/8 of the enclosing power of two (at most 12.5 % slack)
    size_t p = 1;
    while (p < b) p <<= 1;
    const size_t step = p >> 4;
    return (b + step - 1) / step * step;
}
}  // namespace

void set_thread_stream(hipStream_t s) { g_thread_stream = s; }
hipStream_t thread_stream() { return g_thread_stream; }

void* dev_alloc(size_t bytes) {
    const size_t rb = round_bytes(bytes);
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto& lists = g_pool[(void*)g_thread_stream];
        auto it = lists.find(rb);
        if (it != lists.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            return p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, rb) != hipSuccess) {
        (void)hipGetLastError();
        // give the cached blocks of every stream back and try once more
        std::lock_guard<std::mutex> lk(g_pool_mu);
        STEP_HIP(hipDeviceSynchronize());
        for (auto& s : g_pool)
            for (auto& l : s.second)
                for (void* q : l.second) (void)hipFree(q);
        g_pool.clear();
        if (hipMalloc(&p, rb) != hipSuccess) {
            (void)hipGetLastError();
            throw StepError(BUG_ENOMEM, "device allocation of " + std::to_string(rb) + " bytes failed");
        }
    }
    return p;
}

void dev_free(void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool[(void*)g_thread_stream][round_bytes(bytes)].push_back(p);
}

DevBuf::DevBuf(size_t b) : bytes(round_bytes(b)) { p = dev_alloc(bytes); }
DevBuf::~DevBuf() { dev_free(p, bytes); }

bug_bsm Tensor::desc() const {
    const Layout& l = *lay;
    bug_bsm d;
    d.row = bug_leg{l.row.nsec(), l.row.q.data(), l.row.d.data()};
    d.col = bug_leg{l.col.nsec(), l.col.q.data(), l.col.d.data()};
    d.nslice = l.nslice();
    d.shift = l.shifts.data();
    d.offset = l.off.data();
    d.data = buf->p;
    d.nelem = l.nelem;
    return d;
}

bug_mpo_site MpoSite::desc() const {
    bug_mpo_site w;
    w.wl = (int32_t)cl.size();
    w.wr = (int32_t)cr.size();
    w.cl = cl.data();
    w.cr = cr.data();
    w.nent = (int32_t)a.size();
    w.a = a.data();
    w.b = b.data();
    w.sout = sout.data();
    w.sin = sin.data();
    w.coeff = coeff.data();
    return w;
}

// ---- tensors ---------------------------------------------------------------------------------------

Tensor t_empty(const Exec& ex, const Layout& lay) {
    (void)ex;
    Tensor t;
    t.lay = std::make_shared<const Layout>(lay);
    t.buf = std::make_shared<DevBuf>(sizeof(double2) * (size_t)std::max<int64_t>(lay.nelem, 1));
    return t;
}

Tensor t_zeros(const Exec& ex, const Layout& lay) {
    Tensor t = t_empty(ex, lay);
    STEP_HIP(hipMemsetAsync(t.buf->p, 0, sizeof(double2) * (size_t)std::max<int64_t>(lay.nelem, 1), ex.st));
    return t;
}

Tensor t_clone(const Exec& ex, const Tensor& t) {
    Tensor c = t_empty(ex, t.L());
    if (t.nelem() > 0)
        STEP_HIP(hipMemcpyAsync(c.buf->p, t.buf->p, sizeof(double2) * (size_t)t.nelem(),
                                hipMemcpyDeviceToDevice, ex.st));
    return c;
}

Tensor t_from_host(const Exec& ex, const Layout& lay, const std::vector<cplx>& host) {
    if ((int64_t)host.size() < lay.nelem) throw StepError(BUG_EINVARIANT, "host data shorter than its layout");
    Tensor t = t_empty(ex, lay);
    if (lay.nelem > 0) {
        STEP_HIP(hipMemcpyAsync(t.buf->p, host.data(), sizeof(double2) * (size_t)lay.nelem,
                                hipMemcpyHostToDevice, ex.st));
        STEP_HIP(hipStreamSynchronize(ex.st));
    }
    return t;
}

std::vector<cplx> t_to_host(const Exec& ex, const Tensor& t) {
    std::vector<cplx> h((size_t)t.nelem());
    if (t.nelem() > 0) {
        STEP_HIP(hipMemcpyAsync(h.data(), t.buf->p, sizeof(double2) * (size_t)t.nelem(),
                                hipMemcpyDeviceToHost, ex.st));
        STEP_HIP(hipStreamSynchronize(ex.st));
    }
    return h;
}

// ---- products ----------------------------------------------------------------------------------------

static std::vector<bug_term> to_terms(const std::vector<Term>& terms) {
    std::vector<bug_term> t(terms.size());
    for (size_t i = 0; i < terms.size(); ++i)
        t[i] = bug_term{terms[i].o, terms[i].p, terms[i].q, terms[i].op_a, terms[i].op_b,
                        terms[i].alpha.real(), terms[i].alpha.imag()};
    return t;
}

void product_into(const Exec& ex, const Tensor& out, const Tensor& P, const Tensor& Q,
                  const std::vector<Term>& terms, int prec) {
    const bug_bsm o = out.desc(), p = P.desc(), q = Q.desc();
    const auto t = to_terms(terms);
    ex.check(bug_bsm_product(ex.ctx, &o, &p, &q, (int32_t)t.size(), t.data(), 1, prec, ex.st));
}

Tensor product(const Exec& ex, const Leg& row, const Leg& col, const std::vector<int32_t>& shifts,
               const Tensor& P, const Tensor& Q, const std::vector<Term>& terms, int prec) {
    Tensor out = t_empty(ex, Layout(row, col, shifts));
    const bug_bsm o = out.desc(), p = P.desc(), q = Q.desc();
    const auto t = to_terms(terms);
    ex.check(bug_bsm_product(ex.ctx, &o, &p, &q, (int32_t)t.size(), t.data(), 0, prec, ex.st));
    return out;
}

Tensor core_times_site(const Exec& ex, const Tensor& C, const Tensor& A) {
    std::vector<Term> t;
    for (int s = 0; s < A.L().nslice(); ++s) t.push_back({s, 0, s, BUG_OP_N, BUG_OP_N, 1.0});
    return product(ex, C.L().row, A.L().col, A.L().shifts, C, A, t);
}

Tensor site_times_core(const Exec& ex, const Tensor& A, const Tensor& C) {
    std::vector<Term> t;
    for (int s = 0; s < A.L().nslice(); ++s) t.push_back({s, s, 0, BUG_OP_N, BUG_OP_N, 1.0});
    return product(ex, A.L().row, C.L().col, A.L().shifts, A, C, t);
}

Tensor overlap_left(const Exec& ex, const Tensor& U, const Tensor& A) {
    std::vector<Term> t;
    for (int s = 0; s < A.L().nslice(); ++s) t.push_back({0, s, s, BUG_OP_C, BUG_OP_N, 1.0});
    return product(ex, U.L().col, A.L().col, {0}, U, A, t);
}

Tensor overlap_right(const Exec& ex, const Tensor& A, const Tensor& V) {
    std::vector<Term> t;
    for (int s = 0; s < A.L().nslice(); ++s) t.push_back({0, s, s, BUG_OP_N, BUG_OP_C, 1.0});
    return product(ex, A.L().row, V.L().row, {0}, A, V, t);
}

Tensor core_times_core(const Exec& ex, const Tensor& A, const Tensor& B) {
    return product(ex, A.L().row, B.L().col, {0}, A, B, {{0, 0, 0, BUG_OP_N, BUG_OP_N, 1.0}});
}

Tensor perp_left(const Exec& ex, const Tensor& U, const Tensor& K) {
    Tensor ov = overlap_left(ex, U, K);
    Tensor out = t_clone(ex, K);
    std::vector<Term> t;
    for (int s = 0; s < K.L().nslice(); ++s) t.push_back({s, s, 0, BUG_OP_N, BUG_OP_N, -1.0});
    product_into(ex, out, U, ov, t);
    return out;
}

Tensor perp_right(const Exec& ex, const Tensor& V, const Tensor& K) {
    Tensor ov = overlap_right(ex, K, V);
    Tensor out = t_clone(ex, K);
    std::vector<Term> t;
    for (int s = 0; s < K.L().nslice(); ++s) t.push_back({s, 0, s, BUG_OP_N, BUG_OP_N, -1.0});
    product_into(ex, out, ov, V, t);
    return out;
}

// ---- vectors ------------------------------------------------------------------------------------------

static void same(const Tensor& a, const Tensor& b) {
    if (a.L() != b.L()) throw StepError(BUG_EINVARIANT, "vector operation on different block layouts");
}

std::vector<cplx> inner_many(const Exec& ex, const std::vector<const Tensor*>& xs, const Tensor& y) {
    std::vector<cplx> out;
    for (size_t c0 = 0; c0 < xs.size(); c0 += 64) {
        const size_t k = std::min<size_t>(64, xs.size() - c0);
        std::vector<const void*> ptrs(k);
        for (size_t j = 0; j < k; ++j) {
            same(*xs[c0 + j], y);
            ptrs[j] = xs[c0 + j]->buf->p;
        }
        std::vector<double> res(2 * k);
        ex.check(bug_dotc(ex.ctx, y.nelem(), (int32_t)k, ptrs.data(), y.buf->p, res.data(), ex.st));
        for (size_t j = 0; j < k; ++j) out.emplace_back(res[2 * j], res[2 * j + 1]);
    }
    return out;
}

cplx inner(const Exec& ex, const Tensor& a, const Tensor& b) { return inner_many(ex, {&a}, b)[0]; }

double norm(const Exec& ex, const Tensor& a) { return std::sqrt(std::max(inner(ex, a, a).real(), 0.0)); }

void axpy_(const Exec& ex, cplx alpha, const Tensor& x, const Tensor& y) {
    same(x, y);
    ex.check(bug_axpy(ex.ctx, y.nelem(), alpha.real(), alpha.imag(), x.buf->p, y.buf->p, ex.st));
}

void axpy_many_(const Exec& ex, const std::vector<cplx>& alphas, const std::vector<const Tensor*>& xs,
                const Tensor& y) {
    if (alphas.size() != xs.size()) throw StepError(BUG_EINVAL, "axpy_many: as many coefficients as vectors");
    for (size_t lo = 0; lo < xs.size(); lo += 64) {
        const size_t k = std::min<size_t>(64, xs.size() - lo);
        std::vector<double> coef(2 * k);
        std::vector<const void*> ptrs(k);
        for (size_t j = 0; j < k; ++j) {
            same(*xs[lo + j], y);
            coef[2 * j] = alphas[lo + j].real();
            coef[2 * j + 1] = alphas[lo + j].imag();
            ptrs[j] = xs[lo + j]->buf->p;
        }
        ex.check(bug_axpy_many(ex.ctx, y.nelem(), (int32_t)k, coef.data(), ptrs.data(), y.buf->p, ex.st));
    }
}

void scale_(const Exec& ex, cplx alpha, const Tensor& x) {
    ex.check(bug_scal(ex.ctx, x.nelem(), alpha.real(), alpha.imag(), x.buf->p, ex.st));
}

Tensor scale(const Exec& ex, cplx alpha, const Tensor& x) {
    Tensor c = t_clone(ex, x);
    scale_(ex, alpha, c);
    return c;
}

// ---- block movement ----------------------------------------------------------------------------------

void CopyList::add(int64_t src_off, int src_ld, int64_t dst_off, int dst_ld, int rows, int cols,
                   int mode, int scale_mode, int64_t scale_off) {
    if (rows > 0 && cols > 0)
        ops.push_back(bug_copy{src_off, dst_off, src_ld, dst_ld, rows, cols, mode, scale_mode, scale_off});
}

void CopyList::run(const Exec& ex, const void* src, int64_t src_n, void* dst, int64_t dst_n,
                   const double* scale_dev, bool accumulate) const {
    if (ops.empty()) return;
    ex.check(bug_copy_blocks(ex.ctx, (int32_t)ops.size(), ops.data(), src, src_n, dst, dst_n,
                             scale_dev, accumulate ? 1 : 0, ex.st));
}

std::vector<double> block_sqnorms(const Exec& ex, const std::vector<bug_copy>& blocks,
                                  const void* src, int64_t src_n) {
    std::vector<double> out(blocks.size(), 0.0);
    if (blocks.empty()) return out;
    ex.check(bug_block_sqnorms(ex.ctx, (int32_t)blocks.size(), blocks.data(), src, src_n, out.data(), ex.st));
    return out;
}

static bug_copy sqblk(int64_t off, int ld, int rows, int cols) {
    return bug_copy{off, 0, std::max(ld, 1), 1, rows, cols, 0, 0, 0};
}

std::vector<std::pair<int, std::vector<SecPart>>> left_sectors(const Layout& lay) {
    std::vector<std::pair<int, std::vector<SecPart>>> out;
    for (int qr : lay.col.q) {
        std::vector<SecPart> parts;
        for (int s = 0; s < lay.nslice(); ++s) {
            const Blk* b = lay.find(s, qr - lay.shifts[s]);
            if (b) parts.push_back({s, qr - lay.shifts[s], b->d, b->off});
        }
        if (!parts.empty()) out.emplace_back(qr, parts);
    }
    return out;
}

std::vector<std::pair<int, std::vector<SecPart>>> right_sectors(const Layout& lay) {
    std::vector<std::pair<int, std::vector<SecPart>>> out;
    for (int ql : lay.row.q) {
        std::vector<SecPart> parts;
        for (int s = 0; s < lay.nslice(); ++s) {
            const Blk* b = lay.find(s, ql);
            if (b) parts.push_back({s, ql + lay.shifts[s], b->dc, b->off});
        }
        if (!parts.empty()) out.emplace_back(ql, parts);
    }
    return out;
}

Tensor oplus(const Exec& ex, Side side, const Tensor& F, const Tensor& N) {
    const Layout &fl = F.L(), &nl = N.L();
    std::vector<std::pair<int, int>> pairs;
    const Leg& fleg = side == LEFT ? fl.col : fl.row;
    const Leg& nleg = side == LEFT ? nl.col : nl.row;
    for (int i = 0; i < fleg.nsec(); ++i) pairs.emplace_back(fleg.q[i], fleg.d[i]);
    for (int i = 0; i < nleg.nsec(); ++i) pairs.emplace_back(nleg.q[i], nleg.d[i]);
    const Leg merged = Leg::make(pairs);
    const Layout ol = side == LEFT ? Layout(fl.row, merged, fl.shifts) : Layout(merged, fl.col, fl.shifts);
    Tensor out = t_zeros(ex, ol);
    CopyList cf, cn;
    for (const Blk& b : fl.blks) {
        const Blk* o = ol.find(b.k, b.q);
        cf.add(b.off, b.d, o->off, o->d, b.d, b.dc);
    }
    for (const Blk& b : nl.blks) {
        const Blk* o = ol.find(b.k, b.q);
        if (side == LEFT)
            cn.add(b.off, b.d, o->off + (int64_t)o->d * fleg.dim_of(b.qc), o->d, b.d, b.dc);
        else
            cn.add(b.off, b.d, o->off + fleg.dim_of(b.q), o->d, b.d, b.dc);
    }
    cf.run(ex, F.buf->p, F.nelem(), out.buf->p, out.nelem());
    cn.run(ex, N.buf->p, N.nelem(), out.buf->p, out.nelem());
    return out;
}

Tensor trim(const Exec& ex, const Tensor& T, Side side, const std::map<int, int>& counts) {
    std::vector<std::pair<int, int>> pairs;
    for (auto& c : counts)
        if (c.second > 0) pairs.emplace_back(c.first, c.second);
    if (pairs.empty()) return Tensor();
    const Leg leg = Leg::make(pairs);
    const Layout lay = side == LEFT ? Layout(T.L().row, leg, T.L().shifts) : Layout(leg, T.L().col, T.L().shifts);
    Tensor out = t_zeros(ex, lay);
    CopyList cl;
    for (const Blk& b : lay.blks) {
        const Blk* e = T.L().find(b.k, b.q);
        cl.add(e->off, e->d, b.off, b.d, b.d, b.dc);
    }
    cl.run(ex, T.buf->p, T.nelem(), out.buf->p, out.nelem());
    return out;
}

Tensor diag_scale(const Exec& ex, const Tensor& T, const std::map<int, std::vector<double>>& S, bool rows) {
    std::map<int, int64_t> offs;
    std::vector<double> flat;
    for (auto& e : S) {
        offs[e.first] = (int64_t)flat.size();
        flat.insert(flat.end(), e.second.begin(), e.second.end());
    }
    if (flat.empty()) flat.push_back(0.0);
    std::shared_ptr<DevBuf> sb = std::make_shared<DevBuf>(sizeof(double) * flat.size());
    STEP_HIP(hipMemcpyAsync(sb->p, flat.data(), sizeof(double) * flat.size(), hipMemcpyHostToDevice, ex.st));
    Tensor out = t_zeros(ex, T.L());
    CopyList cl;
    for (const Blk& b : T.L().blks) {
        if (rows)
            cl.add(b.off, b.d, b.off, b.d, b.d, b.dc, 0, 1, offs.at(b.q));
        else
            cl.add(b.off, b.d, b.off, b.d, b.d, b.dc, 0, 2, offs.at(b.qc));
    }
    cl.run(ex, T.buf->p, T.nelem(), out.buf->p, out.nelem(), static_cast<const double*>(sb->p));
    STEP_HIP(hipStreamSynchronize(ex.st));      // the host scales go out of scope
    return out;
}

// ---- QR -----------------------------------------------------------------------------------------------

static constexpr double kZeroRel = 1e-14;

static std::vector<int> nonzero(const std::vector<double>& sq) {
    double tot = 0.0;
    for (double v : sq) tot += v;
    const double floor = kZeroRel * kZeroRel * tot;
    std::vector<int> keep;
    for (size_t i = 0; i < sq.size(); ++i)
        if (sq[i] > floor) keep.push_back((int)i);
    return keep;
}

struct QrOut {
    std::shared_ptr<DevBuf> Q, R;
    std::vector<int64_t> qo, ro;
    std::vector<int> ks;
};

static QrOut run_qr(const Exec& ex, const std::vector<std::pair<int, int>>& mats, void* ws,
                    int64_t ws_n, const std::vector<int64_t>& offs) {
    QrOut o;
    int64_t qp = 0, rp = 0;
    for (auto& mn : mats) {
        const int k = std::min(mn.first, mn.second);
        o.qo.push_back(qp);
        o.ro.push_back(rp);
        o.ks.push_back(k);
        qp += (int64_t)mn.first * k;
        rp += (int64_t)k * mn.second;
    }
    o.Q = std::make_shared<DevBuf>(sizeof(double2) * (size_t)std::max<int64_t>(qp, 1));
    o.R = std::make_shared<DevBuf>(sizeof(double2) * (size_t)std::max<int64_t>(rp, 1));
    if (!mats.empty()) {
        std::vector<bug_qr_job> jobs(mats.size());
        for (size_t i = 0; i < mats.size(); ++i) {
            const int m = mats[i].first, n = mats[i].second;
            jobs[i] = bug_qr_job{offs[i], o.qo[i], o.ro[i], m, n, std::max(m, 1), std::max(m, 1),
                                 std::max(o.ks[i], 1), 0};
        }
        ex.check(bug_block_qr(ex.ctx, (int32_t)jobs.size(), jobs.data(), ws, ws_n, o.Q->p,
                              std::max<int64_t>(qp, 1), o.R->p, std::max<int64_t>(rp, 1), ex.st));
    }
    return o;
}

void qr_left(const Exec& ex, const Tensor& A, Tensor& Q, Tensor& R) {
    const Layout& lay = A.L();
    const auto secs = left_sectors(lay);
    CopyList gather;
    std::vector<int64_t> offs;
    std::vector<std::pair<int, int>> mats;
    int64_t pos = 0;
    for (auto& sp : secs) {
        int m = 0;
        for (auto& p : sp.second) m += p.n;
        const int n = lay.col.dim_of(sp.first);
        int r0 = 0;
        for (auto& p : sp.second) {
            gather.add(p.off, p.n, pos + r0, m, p.n, n);
            r0 += p.n;
        }
        offs.push_back(pos);
        mats.emplace_back(m, n);
        pos += (int64_t)m * n;
    }
    DevBuf ws(sizeof(double2) * (size_t)std::max<int64_t>(pos, 1));
    gather.run(ex, A.buf->p, A.nelem(), ws.p, pos);
    std::vector<bug_copy> nb;
    for (size_t i = 0; i < mats.size(); ++i) nb.push_back(sqblk(offs[i], mats[i].first, mats[i].first, mats[i].second));
    const std::vector<int> keep = nonzero(block_sqnorms(ex, nb, ws.p, pos));
    std::vector<std::pair<int, int>> km;
    std::vector<int64_t> ko;
    for (int i : keep) {
        km.push_back(mats[i]);
        ko.push_back(offs[i]);
    }
    QrOut o = run_qr(ex, km, ws.p, std::max<int64_t>(pos, 1), ko);
    std::vector<std::pair<int, int>> bp;
    for (size_t j = 0; j < keep.size(); ++j) bp.emplace_back(secs[keep[j]].first, o.ks[j]);
    const Leg bond = Leg::make(bp);
    Q = t_zeros(ex, Layout(lay.row, bond, lay.shifts));
    R = t_zeros(ex, Layout(bond, lay.col, {0}));
    CopyList cq, cr;
    for (size_t j = 0; j < keep.size(); ++j) {
        const int i = keep[j];
        const int qr = secs[i].first;
        const int m = mats[i].first, n = mats[i].second, k = o.ks[j];
        int r0 = 0;
        for (auto& p : secs[i].second) {
            const Blk* b = Q.L().find(p.s, p.qo);
            cq.add(o.qo[j] + r0, m, b->off, b->d, p.n, k);
            r0 += p.n;
        }
        const Blk* b = R.L().find(0, qr);
        cr.add(o.ro[j], k, b->off, b->d, k, n);
    }
    cq.run(ex, o.Q->p, (int64_t)(o.Q->bytes / sizeof(double2)), Q.buf->p, Q.nelem());
    cr.run(ex, o.R->p, (int64_t)(o.R->bytes / sizeof(double2)), R.buf->p, R.nelem());
}

void qr_right(const Exec& ex, const Tensor& A, Tensor& Lc, Tensor& Q) {
    const Layout& lay = A.L();
    const auto secs = right_sectors(lay);
    CopyList gather;
    std::vector<int64_t> offs;
    std::vector<std::pair<int, int>> mats;
    int64_t pos = 0;
    for (auto& sp : secs) {
        int ncol = 0;
        for (auto& p : sp.second) ncol += p.n;
        const int rows = lay.row.dim_of(sp.first);
        int c0 = 0;
        for (auto& p : sp.second) {
            gather.add(p.off, rows, pos + c0, ncol, p.n, rows, 2);
            c0 += p.n;
        }
        offs.push_back(pos);
        mats.emplace_back(ncol, rows);
        pos += (int64_t)ncol * rows;
    }
    DevBuf ws(sizeof(double2) * (size_t)std::max<int64_t>(pos, 1));
    gather.run(ex, A.buf->p, A.nelem(), ws.p, pos);
    std::vector<bug_copy> nb;
    for (size_t i = 0; i < mats.size(); ++i) nb.push_back(sqblk(offs[i], mats[i].first, mats[i].first, mats[i].second));
    const std::vector<int> keep = nonzero(block_sqnorms(ex, nb, ws.p, pos));
    std::vector<std::pair<int, int>> km;
    std::vector<int64_t> ko;
    for (int i : keep) {
        km.push_back(mats[i]);
        ko.push_back(offs[i]);
    }
    QrOut o = run_qr(ex, km, ws.p, std::max<int64_t>(pos, 1), ko);
    std::vector<std::pair<int, int>> bp;
    for (size_t j = 0; j < keep.size(); ++j) bp.emplace_back(secs[keep[j]].first, o.ks[j]);
    const Leg bond = Leg::make(bp);
    Lc = t_zeros(ex, Layout(lay.row, bond, {0}));
    Q = t_zeros(ex, Layout(bond, lay.col, lay.shifts));
    CopyList cl, cq;
    for (size_t j = 0; j < keep.size(); ++j) {
        const int i = keep[j];
        const int ql = secs[i].first;
        const int mq = mats[i].first, nq = mats[i].second, k = o.ks[j];
        const Blk* b = Lc.L().find(0, ql);
        cl.add(o.ro[j], k, b->off, b->d, nq, k, 2);
        int c0 = 0;
        for (auto& p : secs[i].second) {
            const Blk* bq = Q.L().find(p.s, ql);
            cq.add(o.qo[j] + c0, mq, bq->off, bq->d, k, p.n, 2);
            c0 += p.n;
        }
    }
    cl.run(ex, o.R->p, (int64_t)(o.R->bytes / sizeof(double2)), Lc.buf->p, Lc.nelem());
    cq.run(ex, o.Q->p, (int64_t)(o.Q->bytes / sizeof(double2)), Q.buf->p, Q.nelem());
}

void qr_core(const Exec& ex, const Tensor& C, Tensor& Q, Tensor& R) {
    const Layout& lay = C.L();
    std::vector<bug_copy> nb;
    for (const Blk& b : lay.blks) nb.push_back(sqblk(b.off, b.d, b.d, b.dc));
    const std::vector<int> keep = nonzero(block_sqnorms(ex, nb, C.buf->p, C.nelem()));
    std::vector<std::pair<int, int>> mats;
    std::vector<int64_t> offs;
    for (int i : keep) {
        mats.emplace_back(lay.blks[i].d, lay.blks[i].dc);
        offs.push_back(lay.blks[i].off);
    }
    Tensor ws = t_clone(ex, C);
    QrOut o = run_qr(ex, mats, ws.buf->p, std::max<int64_t>(ws.nelem(), 1), offs);
    std::vector<std::pair<int, int>> bp;
    for (size_t j = 0; j < keep.size(); ++j) bp.emplace_back(lay.blks[keep[j]].q, o.ks[j]);
    const Leg bond = Leg::make(bp);
    Q = t_zeros(ex, Layout(lay.row, bond, {0}));
    R = t_zeros(ex, Layout(bond, lay.col, {0}));
    CopyList cq, cr;
    for (size_t j = 0; j < keep.size(); ++j) {
        const Blk& bl = lay.blks[keep[j]];
        const int m = mats[j].first, n = mats[j].second, k = o.ks[j];
        const Blk* bq = Q.L().find(0, bl.q);
        cq.add(o.qo[j], m, bq->off, bq->d, m, k);
        const Blk* br = R.L().find(0, bl.q);
        cr.add(o.ro[j], k, br->off, br->d, k, n);
    }
    cq.run(ex, o.Q->p, (int64_t)(o.Q->bytes / sizeof(double2)), Q.buf->p, Q.nelem());
    cr.run(ex, o.R->p, (int64_t)(o.R->bytes / sizeof(double2)), R.buf->p, R.nelem());
}

// ---- SVD ----------------------------------------------------------------------------------------------

double sq_weight(const std::vector<SvEntry>& l) {
    double s = 0.0;
    for (const SvEntry& e : l) s += e.s * e.s;
    return s;
}

double trunc_weight(const SVDResult& r) {
    const double k = sq_weight(r.kept), c = sq_weight(r.trunc), t = k + c;
    return t == 0.0 ? 0.0 : std::sqrt(c / t);
}

namespace {
// per-sector SVDs of matrices M_i (m x n) gathered by the caller at goffs into ws
struct SvdRun {
    std::vector<std::pair<int, int>> mats;
    std::vector<char> tall;
    std::vector<int64_t> goffs;
    int64_t wsize = 0;
    std::shared_ptr<DevBuf> ws;
    explicit SvdRun(const std::vector<std::pair<int, int>>& m) : mats(m) {
        for (auto& mn : mats) {
            tall.push_back(mn.first >= mn.second);
            goffs.push_back(wsize);
            wsize += (int64_t)mn.first * mn.second;
        }
        ws = std::make_shared<DevBuf>(sizeof(double2) * (size_t)std::max<int64_t>(wsize, 1));
    }
    // place block (rows x cols, ld) at (r0, c0) of matrix i (stored as M or M^H)
    void gather(CopyList& cl, int i, int64_t off, int ld, int r0, int c0, int rows, int cols) const {
        const int m = mats[i].first, n = mats[i].second;
        if (tall[i])
            cl.add(off, ld, goffs[i] + r0 + (int64_t)c0 * m, m, rows, cols);
        else
            cl.add(off, ld, goffs[i] + c0 + (int64_t)r0 * n, n, cols, rows, 2);
    }
    void run(const Exec& ex, BlockSVD& out) {
        std::vector<bug_svd_job> jobs;
        int64_t up = 0, sp = 0, vp = 0;
        for (size_t i = 0; i < mats.size(); ++i) {
            const int mj = tall[i] ? mats[i].first : mats[i].second;
            const int nj = tall[i] ? mats[i].second : mats[i].first;
            out.uo.push_back(up);
            out.so.push_back(sp);
            out.vo.push_back(vp);
            jobs.push_back(bug_svd_job{goffs[i], up, sp, vp, mj, nj, std::max(mj, 1), std::max(mj, 1),
                                       std::max(nj, 1), tall[i] ? 0 : 1});
            up += (int64_t)mj * nj;
            sp += nj;
            vp += (int64_t)nj * nj;
        }
        out.U = std::make_shared<DevBuf>(sizeof(double2) * (size_t)std::max<int64_t>(up, 1));
        out.Vh = std::make_shared<DevBuf>(sizeof(double2) * (size_t)std::max<int64_t>(vp, 1));
        DevBuf S(sizeof(double) * (size_t)std::max<int64_t>(sp, 1));
        if (!jobs.empty())
            ex.check(bug_block_svd(ex.ctx, (int32_t)jobs.size(), jobs.data(), ws->p, std::max<int64_t>(wsize, 1),
                                   out.U->p, std::max<int64_t>(up, 1), static_cast<double*>(S.p),
                                   std::max<int64_t>(sp, 1), out.Vh->p, std::max<int64_t>(vp, 1), 0, ex.st));
        std::vector<double> host((size_t)std::max<int64_t>(sp, 1));
        STEP_HIP(hipMemcpyAsync(host.data(), S.p, sizeof(double) * host.size(), hipMemcpyDeviceToHost, ex.st));
        STEP_HIP(hipStreamSynchronize(ex.st));
        for (size_t i = 0; i < mats.size(); ++i) {
            const int k = std::min(mats[i].first, mats[i].second);
            std::vector<double> sv(host.begin() + out.so[i], host.begin() + out.so[i] + k);
            for (double v : sv)
                if (!std::isfinite(v))
                    throw StepError(BUG_ENONFINITE, "SVD of a block with non-finite entries: the state or "
                                                    "the operator is not finite");
            out.spectra.push_back(std::move(sv));
        }
        out.mats = mats;
        out.tall = tall;
        ws.reset();
    }
};
}  // namespace

int global_selection(const std::vector<std::vector<double>>& spectra, const std::vector<int>& tags,
                     double cutoff, int nkeep, std::vector<SvEntry>& full) {
    // Telum's global selection: sv > max(cutoff, floor) sv_max over all sectors, then the largest
    // nkeep by a stable descending sort in sector order; the kept values lead the sorted list
    full.clear();
    double smax = 0.0;
    for (size_t i = 0; i < spectra.size(); ++i)
        for (size_t j = 0; j < spectra[i].size(); ++j) {
            smax = std::max(smax, spectra[i][j]);
            full.push_back({spectra[i][j], tags[i], (int)j});
        }
    std::stable_sort(full.begin(), full.end(), [](const SvEntry& a, const SvEntry& b) { return a.s > b.s; });
    const double thresh = std::max(cutoff, kZeroRel) * smax;
    int n = 0;
    while (n < (int)full.size() && full[n].s > thresh) ++n;
    if (nkeep >= 0) n = std::min(n, nkeep);
    return n;
}

SVDResult BlockSVD::select(const Exec& ex, double cutoff, int nkeep) const {
    std::vector<SvEntry> full;
    const int n = global_selection(spectra, qlist, cutoff, nkeep, full);
    std::map<int, int> keep;
    for (int q : qlist) keep[q] = 0;
    SVDResult r;
    r.kept.assign(full.begin(), full.begin() + n);
    r.trunc.assign(full.begin() + n, full.end());
    for (const SvEntry& e : r.kept) keep[e.q] += 1;
    std::vector<std::pair<int, int>> bp;
    for (int q : qlist)
        if (keep[q] > 0) bp.emplace_back(q, keep[q]);
    const Leg bond = Leg::make(bp);
    const Layout& l = *lay;
    if (kind == 0) {
        r.U = t_zeros(ex, Layout(l.row, bond, l.shifts));
        r.Vh = t_zeros(ex, Layout(bond, l.col, {0}));
    } else if (kind == 1) {
        r.U = t_zeros(ex, Layout(l.row, bond, {0}));
        r.Vh = t_zeros(ex, Layout(bond, l.col, l.shifts));
    } else {
        r.U = t_zeros(ex, Layout(l.row, bond, {0}));
        r.Vh = t_zeros(ex, Layout(bond, l.col, {0}));
    }
    CopyList cu_u, cu_v, cv_v, cv_u;
    auto add_u_rows = [&](CopyList& cl_tall, CopyList& cl_wide, int i, int r0, int rows, int k,
                          int64_t dst_off, int dst_ld) {
        const int m = mats[i].first;
        if (tall[i])
            cl_tall.add(uo[i] + r0, m, dst_off, dst_ld, rows, k);
        else
            cl_wide.add(vo[i] + (int64_t)r0 * m, m, dst_off, dst_ld, rows, k, 2);
    };
    auto add_vh_cols = [&](CopyList& cl_tall, CopyList& cl_wide, int i, int c0, int cols, int k,
                           int64_t dst_off, int dst_ld) {
        const int n = mats[i].second;
        if (tall[i])
            cl_tall.add(vo[i] + (int64_t)c0 * n, n, dst_off, dst_ld, k, cols);
        else
            cl_wide.add(uo[i] + c0, n, dst_off, dst_ld, k, cols, 2);
    };
    for (size_t i = 0; i < qlist.size(); ++i) {
        const int q = qlist[i];
        const int k = keep[q];
        if (k == 0) continue;
        r.S[q] = std::vector<double>(spectra[i].begin(), spectra[i].begin() + k);
        const int m = mats[i].first, n = mats[i].second;
        if (kind == 0) {
            int r0 = 0;
            for (const SecPart& p : secs[i]) {
                const Blk* b = r.U.L().find(p.s, p.qo);
                add_u_rows(cu_u, cu_v, (int)i, r0, p.n, k, b->off, b->d);
                r0 += p.n;
            }
            const Blk* b = r.Vh.L().find(0, q);
            add_vh_cols(cv_v, cv_u, (int)i, 0, n, k, b->off, b->d);
        } else if (kind == 1) {
            const Blk* b = r.U.L().find(0, q);
            add_u_rows(cu_u, cu_v, (int)i, 0, m, k, b->off, b->d);
            int c0 = 0;
            for (const SecPart& p : secs[i]) {
                const Blk* bv = r.Vh.L().find(p.s, q);
                add_vh_cols(cv_v, cv_u, (int)i, c0, p.n, k, bv->off, bv->d);
                c0 += p.n;
            }
        } else {
            const Blk* b = r.U.L().find(0, q);
            add_u_rows(cu_u, cu_v, (int)i, 0, m, k, b->off, b->d);
            const Blk* bv = r.Vh.L().find(0, q);
            add_vh_cols(cv_v, cv_u, (int)i, 0, n, k, bv->off, bv->d);
        }
    }
    const int64_t un = (int64_t)(U->bytes / sizeof(double2)), vn = (int64_t)(Vh->bytes / sizeof(double2));
    cu_u.run(ex, U->p, un, r.U.buf->p, r.U.nelem());
    cu_v.run(ex, Vh->p, vn, r.U.buf->p, r.U.nelem());
    cv_v.run(ex, Vh->p, vn, r.Vh.buf->p, r.Vh.nelem());
    cv_u.run(ex, U->p, un, r.Vh.buf->p, r.Vh.nelem());
    return r;
}

BlockSVD factor_left(const Exec& ex, const Tensor& A) {
    const Layout& lay = A.L();
    const auto secs = left_sectors(lay);
    BlockSVD f;
    f.kind = 0;
    f.lay = A.lay;
    std::vector<std::pair<int, int>> mats;
    for (auto& sp : secs) {
        int m = 0;
        for (auto& p : sp.second) m += p.n;
        mats.emplace_back(m, lay.col.dim_of(sp.first));
        f.qlist.push_back(sp.first);
        f.secs.push_back(sp.second);
    }
    SvdRun run(mats);
    CopyList g;
    for (size_t i = 0; i < secs.size(); ++i) {
        int r0 = 0;
        for (auto& p : secs[i].second) {
            run.gather(g, (int)i, p.off, p.n, r0, 0, p.n, mats[i].second);
            r0 += p.n;
        }
    }
    g.run(ex, A.buf->p, A.nelem(), run.ws->p, std::max<int64_t>(run.wsize, 1));
    run.run(ex, f);
    return f;
}

BlockSVD factor_right(const Exec& ex, const Tensor& A) {
    const Layout& lay = A.L();
    const auto secs = right_sectors(lay);
    BlockSVD f;
    f.kind = 1;
    f.lay = A.lay;
    std::vector<std::pair<int, int>> mats;
    for (auto& sp : secs) {
        int n = 0;
        for (auto& p : sp.second) n += p.n;
        mats.emplace_back(lay.row.dim_of(sp.first), n);
        f.qlist.push_back(sp.first);
        f.secs.push_back(sp.second);
    }
    SvdRun run(mats);
    CopyList g;
    for (size_t i = 0; i < secs.size(); ++i) {
        int c0 = 0;
        for (auto& p : secs[i].second) {
            run.gather(g, (int)i, p.off, mats[i].first, 0, c0, mats[i].first, p.n);
            c0 += p.n;
        }
    }
    g.run(ex, A.buf->p, A.nelem(), run.ws->p, std::max<int64_t>(run.wsize, 1));
    run.run(ex, f);
    return f;
}

BlockSVD factor_core(const Exec& ex, const Tensor& C) {
    const Layout& lay = C.L();
    BlockSVD f;
    f.kind = 2;
    f.lay = C.lay;
    std::vector<std::pair<int, int>> mats;
    for (const Blk& b : lay.blks) {
        mats.emplace_back(b.d, b.dc);
        f.qlist.push_back(b.q);
        f.secs.emplace_back();
    }
    SvdRun run(mats);
    CopyList g;
    for (size_t i = 0; i < lay.blks.size(); ++i) {
        const Blk& b = lay.blks[i];
        run.gather(g, (int)i, b.off, b.d, 0, 0, b.d, b.dc);
    }
    g.run(ex, C.buf->p, C.nelem(), run.ws->p, std::max<int64_t>(run.wsize, 1));
    run.run(ex, f);
    return f;
}

// ---- environments and effective operators --------------------------------------------------------

Layout env_layout(const Leg& bra, const Leg& ket, const std::vector<int32_t>& channel_charges) {
    std::vector<int32_t> sh;
    for (int32_t c : channel_charges) sh.push_back(-c);
    return Layout(bra, ket, sh);
}

Tensor boundary_left(const Exec& ex, const MpoSite& W0, const Leg& leg) {
    if (leg.total() != 1) throw StepError(BUG_EINVARIANT, "the left boundary link must be one-dimensional");
    const Layout lay = env_layout(leg, leg, W0.cl);
    std::vector<cplx> h((size_t)std::max<int64_t>(lay.nelem, 1), 0.0);
    const Blk* b = lay.find(0, leg.q[0]);
    if (b) h[(size_t)b->off] = 1.0;
    return t_from_host(ex, lay, h);
}

Tensor boundary_right(const Exec& ex, const MpoSite& Wlast, const Leg& leg) {
    if (leg.total() != 1) throw StepError(BUG_EINVARIANT, "the right boundary link must be one-dimensional");
    const Layout lay = env_layout(leg, leg, Wlast.cr);
    std::vector<cplx> h((size_t)std::max<int64_t>(lay.nelem, 1), 0.0);
    const Blk* b = lay.find(0, leg.q[0]);
    if (b) h[(size_t)b->off] = 1.0;
    return t_from_host(ex, lay, h);
}

Tensor push_left(const Exec& ex, const Tensor& L, const Tensor& A, const MpoSite& W, int prec) {
    Tensor out = t_empty(ex, env_layout(A.L().col, A.L().col, W.cr));
    const bug_bsm l = L.desc(), a = A.desc(), o = out.desc();
    const bug_mpo_site w = W.desc();
    ex.check(bug_env_push_left(ex.ctx, &l, &a, nullptr, &w, &o, prec, ex.st));
    return out;
}

Tensor push_right(const Exec& ex, const Tensor& R, const Tensor& B, const MpoSite& W, int prec) {
    Tensor out = t_empty(ex, env_layout(B.L().row, B.L().row, W.cl));
    const bug_bsm r = R.desc(), b = B.desc(), o = out.desc();
    const bug_mpo_site w = W.desc();
    ex.check(bug_env_push_right(ex.ctx, &r, &b, nullptr, &w, &o, prec, ex.st));
    return out;
}

Tensor apply_one_site(const Exec& ex, const Tensor& L, const MpoSite& W, const Tensor& R,
                      const Tensor& X, int prec) {
    Tensor Y = t_empty(ex, X.L());
    const bug_bsm l = L.desc(), r = R.desc(), x = X.desc(), y = Y.desc();
    const bug_mpo_site w = W.desc();
    ex.check(bug_apply_one_site(ex.ctx, &l, &w, &r, &x, &y, 0.0, prec, ex.st));
    return Y;
}

Tensor apply_zero_site(const Exec& ex, const Tensor& L, const Tensor& R, const Tensor& C, int prec) {
    Tensor Y = t_empty(ex, C.L());
    const bug_bsm l = L.desc(), r = R.desc(), c = C.desc(), y = Y.desc();
    ex.check(bug_apply_zero_site(ex.ctx, &l, &r, &c, &y, 0.0, prec, ex.st));
    return Y;
}

// ---- probe stream --------------------------------------------------------------------------------

static uint64_t mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

ProbeStream::ProbeStream(uint64_t seed) : key(mix64(seed)), pos(0) {}

cplx ProbeStream::at(uint64_t i) const {
    const uint64_t two_i = 2ull * i + key;
    const uint64_t a = mix64(two_i), b = mix64(two_i + 1ull);
    const double u1 = ((double)(a >> 11) + 0.5) / 9007199254740992.0;
    const double u2 = (double)(b >> 11) / 9007199254740992.0;
    const double rad = std::sqrt(-std::log(u1));
    const double ang = 2.0 * M_PI * u2;
    return cplx(rad * std::cos(ang), rad * std::sin(ang));
}

ProbeStream ProbeStream::split() {
    const uint64_t word = mix64(2ull * pos + key);
    pos += 1;
    return ProbeStream(word);
}

// ---- small dense solvers ---------------------------------------------------------------------------

void tridiag_eigen(const std::vector<double>& diag, const std::vector<double>& off,
                   std::vector<double>& vals, std::vector<double>& vecs) {
    const int n = (int)diag.size();
    std::vector<double> d(diag), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = off[i];
    vecs.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) vecs[(size_t)i * n + i] = 1.0;
    // implicit QL with Wilkinson-type shifts (tql2); vecs column-major: z(i, k) = vecs[k n + i]
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        while (true) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                // negligible at one ulp of its diagonal neighbours (a smaller threshold can
                // stall on entries that rounding keeps at the ulp level)
                if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (m == l) break;
            if (++iter > 200) throw StepError(BUG_ENONFINITE, "tridiagonal eigensolver did not converge");
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                double f = s * e[i];
                const double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    double* zi = &vecs[(size_t)i * n + k];
                    double* zi1 = &vecs[(size_t)(i + 1) * n + k];
                    f = *zi1;
                    *zi1 = s * *zi + c * f;
                    *zi = c * *zi - s * f;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    // sort ascending with the vectors
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    vals.resize(n);
    std::vector<double> v2((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        vals[k] = d[idx[k]];
        for (int i = 0; i < n; ++i) v2[(size_t)k * n + i] = vecs[(size_t)idx[k] * n + i];
    }
    vecs.swap(v2);
}

namespace {
using CM = std::vector<cplx>;   // row-major n x n

// C = A B with the real and imaginary parts of B and C held apart, so that the inner loop over a
// row is plain real arithmetic the compiler vectorizes (the host side of every Krylov step runs
// many of these on k x k blocks, k <= 65)
CM matmul(const CM& A, const CM& B, int n) {
    const size_t nn = (size_t)n * n;
    std::vector<double> br(nn), bi(nn), cr(nn, 0.0), ci(nn, 0.0);
    for (size_t e = 0; e < nn; ++e) {
        br[e] = B[e].real();
        bi[e] = B[e].imag();
    }
    for (int i = 0; i < n; ++i) {
        double* __restrict r = &cr[(size_t)i * n];
        double* __restrict m = &ci[(size_t)i * n];
        for (int k = 0; k < n; ++k) {
            const double ar = A[(size_t)i * n + k].real(), ai = A[(size_t)i * n + k].imag();
            if (ar == 0.0 && ai == 0.0) continue;
            const double* __restrict xr = &br[(size_t)k * n];
            const double* __restrict xi = &bi[(size_t)k * n];
            for (int j = 0; j < n; ++j) {
                r[j] += ar * xr[j] - ai * xi[j];
                m[j] += ar * xi[j] + ai * xr[j];
            }
        }
    }
    CM C(nn);
    for (size_t e = 0; e < nn; ++e) C[e] = cplx(cr[e], ci[e]);
    return C;
}

double norm1(const CM& A, int n) {
    double m = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::abs(A[(size_t)i * n + j]);
        m = std::max(m, s);
    }
    return m;
}

// solve P X = Q (n x n, row-major) by LU with partial pivoting
CM solve(CM P, CM Q, int n) {
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(P[(size_t)r * n + c]) > std::abs(P[(size_t)piv * n + c])) piv = r;
        if (piv != c)
            for (int j = 0; j < n; ++j) {
                std::swap(P[(size_t)c * n + j], P[(size_t)piv * n + j]);
                std::swap(Q[(size_t)c * n + j], Q[(size_t)piv * n + j]);
            }
        const cplx d = P[(size_t)c * n + c];
        if (d == 0.0) throw StepError(BUG_ENONFINITE, "singular Pade denominator in expm");
        for (int r = c + 1; r < n; ++r) {
            const cplx f = P[(size_t)r * n + c] / d;
            if (f == 0.0) continue;
            for (int j = c; j < n; ++j) P[(size_t)r * n + j] -= f * P[(size_t)c * n + j];
            for (int j = 0; j < n; ++j) Q[(size_t)r * n + j] -= f * Q[(size_t)c * n + j];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        const cplx d = P[(size_t)c * n + c];
        for (int j = 0; j < n; ++j) {
            cplx s = Q[(size_t)c * n + j];
            for (int k = c + 1; k < n; ++k) s -= P[(size_t)c * n + k] * Q[(size_t)k * n + j];
            Q[(size_t)c * n + j] = s / d;
        }
    }
    return Q;
}
}  // namespace

std::vector<cplx> expm(const std::vector<cplx>& A0, int n) {
    if (n == 0) return {};
    // Higham (2005), "The scaling and squaring method for the matrix exponential revisited":
    // Pade degrees 3, 5, 7, 9, 13 chosen by the 1-norm, scaling for degree 13
    static const double theta[] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1,
                                   2.097847961257068e0, 5.371920351148152e0};
    static const int degs[] = {3, 5, 7, 9, 13};
    static const double b13[] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0, 129060195264000.0, 10559470521600.0,
                                 670442572800.0, 33522128640.0, 1323241920.0, 40840800.0,
                                 960960.0, 16380.0, 182.0, 1.0};
    static const double b9[] = {17643225600.0, 8821612800.0, 2075673600.0, 302702400.0, 30270240.0,
                                2162160.0, 110880.0, 3960.0, 90.0, 1.0};
    static const double b7[] = {17297280.0, 8648640.0, 1995840.0, 277200.0, 25200.0, 1512.0, 56.0, 1.0};
    static const double b5[] = {30240.0, 15120.0, 3360.0, 420.0, 30.0, 1.0};
    static const double b3[] = {120.0, 60.0, 12.0, 1.0};
    const size_t nn = (size_t)n * n;
    CM I(nn, 0.0);
    for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
    const double nrm = norm1(A0, n);
    if (!std::isfinite(nrm)) throw StepError(BUG_ENONFINITE, "expm of a non-finite matrix");
    CM A = A0;
    int s = 0, deg = 13;
    for (int t = 0; t < 4; ++t)
        if (nrm <= theta[t]) {
            deg = degs[t];
            break;
        }
    if (deg == 13 && nrm > theta[4]) {
        s = std::max(0, (int)std::ceil(std::log2(nrm / theta[4])));
        const double f = std::ldexp(1.0, -s);
        for (auto& v : A) v *= f;
    }
    CM A2 = matmul(A, A, n);
    CM U(nn, 0.0), V(nn, 0.0);
    if (deg == 13) {
        CM A4 = matmul(A2, A2, n), A6 = matmul(A4, A2, n);
        CM W1(nn), W2(nn), Z1(nn), Z2(nn);
        for (size_t e = 0; e < nn; ++e) {
            W1[e] = b13[13] * A6[e] + b13[11] * A4[e] + b13[9] * A2[e];
            W2[e] = b13[7] * A6[e] + b13[5] * A4[e] + b13[3] * A2[e] + b13[1] * I[e];
            Z1[e] = b13[12] * A6[e] + b13[10] * A4[e] + b13[8] * A2[e];
            Z2[e] = b13[6] * A6[e] + b13[4] * A4[e] + b13[2] * A2[e] + b13[0] * I[e];
        }
        CM W = matmul(A6, W1, n);
        for (size_t e = 0; e < nn; ++e) W[e] += W2[e];
        U = matmul(A, W, n);
        V = matmul(A6, Z1, n);
        for (size_t e = 0; e < nn; ++e) V[e] += Z2[e];
    } else {
        const double* b = deg == 3 ? b3 : deg == 5 ? b5 : deg == 7 ? b7 : b9;
        std::vector<CM> pw{I, A2};
        for (int k = 2; 2 * k <= deg - 1; ++k) pw.push_back(matmul(pw.back(), A2, n));
        CM Wsum(nn, 0.0);
        for (int k = 0; 2 * k + 1 <= deg; ++k)
            for (size_t e = 0; e < nn; ++e) {
                Wsum[e] += b[2 * k + 1] * pw[k][e];
                V[e] += b[2 * k] * pw[k][e];
            }
        U = matmul(A, Wsum, n);
    }
    CM P(nn), Q(nn);
    for (size_t e = 0; e < nn; ++e) {
        P[e] = V[e] - U[e];
        Q[e] = V[e] + U[e];
    }
    CM R = solve(P, Q, n);
    for (int k = 0; k < s; ++k) R = matmul(R, R, n);
    return R;
}

std::vector<cplx> eigvals(std::vector<cplx> A, int n) {
    std::vector<cplx> w;
    if (n == 0) return w;
    auto at = [&](int i, int j) -> cplx& { return A[(size_t)i * n + j]; };
    auto cabs1 = [](cplx z) { return std::abs(z.real()) + std::abs(z.imag()); };
    // balancing (Parlett-Reinsch): a diagonal similarity by powers of two that brings each
    // row's and column's off-diagonal norms within a factor of two, so that the rounding of the
    // iteration below is relative to the balanced norm and graded matrices lose no accuracy
    for (bool again = true; again;) {
        again = false;
        for (int i = 0; i < n; ++i) {
            double c = 0.0, r = 0.0;
            for (int j = 0; j < n; ++j)
                if (j != i) {
                    c += cabs1(at(j, i));
                    r += cabs1(at(i, j));
                }
            if (c == 0.0 || r == 0.0) continue;
            double g = r / 2.0, f = 1.0;
            const double sum = c + r;
            while (c < g) {
                f *= 2.0;
                c *= 4.0;
            }
            g = r * 2.0;
            while (c > g) {
                f /= 2.0;
                c /= 4.0;
            }
            if ((c + r) / f < 0.95 * sum) {
                again = true;
                for (int j = 0; j < n; ++j) at(i, j) /= f;
                for (int j = 0; j < n; ++j) at(j, i) *= f;
            }
        }
    }
    // Householder reduction to upper Hessenberg form
    for (int k = 0; k + 2 < n; ++k) {
        double below = 0.0;                  // the column under the subdiagonal
        for (int i = k + 2; i < n; ++i) below += std::norm(at(i, k));
        if (below == 0.0) continue;          // already Hessenberg there (Arnoldi's H always is)
        double alpha = std::norm(at(k + 1, k)) + below;
        alpha = std::sqrt(alpha);
        const cplx x0 = at(k + 1, k);
        const cplx ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cplx(1.0);
        std::vector<cplx> v(n, 0.0);
        v[k + 1] = x0 + ph * alpha;
        for (int i = k + 2; i < n; ++i) v[i] = at(i, k);
        double vn = 0.0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
        if (vn == 0.0) continue;
        // A <- (I - 2 v v^H / vn) A (I - 2 v v^H / vn)
        for (int j = 0; j < n; ++j) {
            cplx s = 0.0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * at(i, j);
            s *= 2.0 / vn;
            for (int i = k + 1; i < n; ++i) at(i, j) -= v[i] * s;
        }
        for (int i = 0; i < n; ++i) {
            cplx s = 0.0;
            for (int j = k + 1; j < n; ++j) s += at(i, j) * v[j];
            s *= 2.0 / vn;
            for (int j = k + 1; j < n; ++j) at(i, j) -= s * std::conj(v[j]);
        }
    }
    // shifted QR on the Hessenberg matrix (Givens rotations, Wilkinson shifts, deflation).
    // A subdiagonal entry is negligible at one ulp of its diagonal neighbours, refined by the
    // Ahues-Tisseur test on the 2 x 2 block it couples (the criterion of LAPACK's xLAHQR): a
    // threshold below one ulp can stall on entries that rounding keeps at that level.
    const double ulp = std::numeric_limits<double>::epsilon();
    const double smlnum = std::numeric_limits<double>::min() * ((double)n / ulp);
    auto negligible = [&](int k) {
        const cplx h = at(k, k - 1);
        if (cabs1(h) <= smlnum) return true;
        double tst = cabs1(at(k - 1, k - 1)) + cabs1(at(k, k));
        if (tst == 0.0) {
            if (k - 2 >= 0) tst += std::abs(at(k - 1, k - 2).real());
            if (k + 1 < n) tst += std::abs(at(k + 1, k).real());
        }
        if (std::abs(h.real()) > ulp * tst) return false;
        const double ab = std::max(cabs1(h), cabs1(at(k - 1, k)));
        const double ba = std::min(cabs1(h), cabs1(at(k - 1, k)));
        const double aa = std::max(cabs1(at(k, k)), cabs1(at(k - 1, k - 1) - at(k, k)));
        const double bb = std::min(cabs1(at(k, k)), cabs1(at(k - 1, k - 1) - at(k, k)));
        const double s = aa + ab;
        return ba * (ab / s) <= std::max(smlnum, ulp * (bb * (aa / s)));
    };
    const int itmax = 30 * std::max(10, n);
    int hi = n - 1, iter = 0;
    while (hi >= 0) {
        if (hi == 0) {
            w.push_back(at(0, 0));
            break;
        }
        int lo = hi;
        for (; lo > 0; --lo) {
            if (negligible(lo)) {
                at(lo, lo - 1) = 0.0;
                break;
            }
        }
        if (lo == hi) {
            w.push_back(at(hi, hi));
            --hi;
            iter = 0;
            continue;
        }
        if (++iter > itmax) {
            // a block that rounding keeps from meeting the coupling test: split it at its
            // relatively smallest subdiagonal entry (already below one ulp of its neighbours
            // where the first test passed), which perturbs the eigenvalues by that entry
            int kmin = lo + 1;
            double best = std::numeric_limits<double>::infinity();
            for (int k = lo + 1; k <= hi; ++k) {
                const double tst = cabs1(at(k - 1, k - 1)) + cabs1(at(k, k));
                const double rel = cabs1(at(k, k - 1)) / (tst > 0.0 ? tst : 1.0);
                if (rel < best) {
                    best = rel;
                    kmin = k;
                }
            }
            if (!std::isfinite(best)) throw StepError(BUG_ENONFINITE, "Hessenberg QR on a non-finite matrix");
            at(kmin, kmin - 1) = 0.0;
            iter = 0;
            continue;
        }
        // shift: the eigenvalue of the trailing 2 x 2 block nearer its last diagonal entry, in the
        // cancellation-free form of xLAHQR; exceptional shifts at the 10th and 20th iteration
        // of a block (from its top, then its bottom subdiagonal)
        cplx mu;
        if (iter == 10) {
            mu = 0.75 * std::abs(at(lo + 1, lo).real()) + at(lo, lo);
        } else if (iter == 20) {
            mu = 0.75 * std::abs(at(hi, hi - 1).real()) + at(hi, hi);
        } else {
            mu = at(hi, hi);
            const cplx u = std::sqrt(at(hi - 1, hi)) * std::sqrt(at(hi, hi - 1));
            double sc = cabs1(u);
            if (sc != 0.0) {
                const cplx x = 0.5 * (at(hi - 1, hi - 1) - mu);
                const double sx = cabs1(x);
                sc = std::max(sc, sx);
                cplx y = sc * std::sqrt((x / sc) * (x / sc) + (u / sc) * (u / sc));
                if (sx > 0.0 && (x / sx).real() * y.real() + (x / sx).imag() * y.imag() < 0.0) y = -y;
                mu -= u * (u / (x + y));
            }
        }
        for (int i = lo; i <= hi; ++i) at(i, i) -= mu;
        std::vector<double> cs(hi - lo);
        std::vector<cplx> sn(hi - lo);
        for (int k = lo; k < hi; ++k) {
            const cplx x = at(k, k), y = at(k + 1, k);
            const double r = std::hypot(std::abs(x), std::abs(y));
            double cc = 1.0;
            cplx ss = 0.0;
            if (r > 0) {
                cc = std::abs(x) / r;
                const cplx ph = std::abs(x) > 0 ? x / std::abs(x) : cplx(1.0);
                ss = ph * std::conj(y) / r;
            }
            cs[k - lo] = cc;
            sn[k - lo] = ss;
            // rows k, k+1 <- [c s; -conj(s) c] rows k, k+1 within the active block (only the
            // eigenvalues are wanted, so the parts of the Schur form outside it are not kept)
            const double sr = ss.real(), si = ss.imag();
            double* __restrict rk = reinterpret_cast<double*>(&at(k, 0));
            double* __restrict rk1 = reinterpret_cast<double*>(&at(k + 1, 0));
            for (int j = k; j <= hi; ++j) {
                const double pr = rk[2 * j], pi = rk[2 * j + 1], qr = rk1[2 * j], qi = rk1[2 * j + 1];
                rk[2 * j] = cc * pr + (sr * qr - si * qi);
                rk[2 * j + 1] = cc * pi + (sr * qi + si * qr);
                rk1[2 * j] = -(sr * pr + si * pi) + cc * qr;
                rk1[2 * j + 1] = -(sr * pi - si * pr) + cc * qi;
            }
        }
        for (int k = lo; k < hi; ++k) {
            const double cc = cs[k - lo];
            const double sr = sn[k - lo].real(), si = sn[k - lo].imag();
            // columns k, k+1 <- columns k, k+1 [c -s; conj(s) c]
            for (int i = lo; i <= std::min(k + 2, hi); ++i) {
                double* a0 = reinterpret_cast<double*>(&at(i, k));
                double* a1 = reinterpret_cast<double*>(&at(i, k + 1));
                const double pr = a0[0], pi = a0[1], qr = a1[0], qi = a1[1];
                a0[0] = cc * pr + (sr * qr + si * qi);
                a0[1] = cc * pi + (sr * qi - si * qr);
                a1[0] = -(sr * pr - si * pi) + cc * qr;
                a1[1] = -(sr * pi + si * pr) + cc * qi;
            }
        }
        for (int i = lo; i <= hi; ++i) at(i, i) += mu;
    }
    return w;
}

}  // namespace host
}  // namespace bug
