// C ABI of the block-contraction engine: context, block-product planning, plan cache, the
// environment pushes and the one-/zero-site effective-Hamiltonian applies.
//
// A contraction of block-sparse U(1) tensors is planned on the host once per layout signature
// (charges, dims, offsets, terms) and cached: the plan lists every output block with the block
// products that feed it.  Execution is one grouped-GEMM launch per contraction stage.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "internal.h"

namespace bug {

// ---- hashing of layouts -----------------------------------------------------------------

struct BsmHasher : Hasher {
    void leg(const bug_leg& l) {
        pod(l.nsec);
        if (l.nsec > 0) {
            bytes(l.charge, sizeof(int32_t) * l.nsec);
            bytes(l.dim, sizeof(int32_t) * l.nsec);
        }
    }
    void bsm(const bug_bsm& m) {
        leg(m.row);
        leg(m.col);
        pod(m.nslice);
        if (m.nslice > 0) {
            bytes(m.shift, sizeof(int32_t) * m.nslice);
            bytes(m.offset, sizeof(int64_t) * (size_t)m.nslice * m.row.nsec);
        }
    }
};

// ---- host view of a block-sparse matrix family ---------------------------------------------

struct Leg {
    std::vector<int32_t> q, d;
    std::unordered_map<int32_t, int32_t> idx;
    Leg() = default;
    explicit Leg(const bug_leg& l) {
        if (l.nsec < 0 || (l.nsec > 0 && (!l.charge || !l.dim)))
            throw Error(BUG_EINVAL, "leg with negative sector count or null tables");
        q.assign(l.charge, l.charge + l.nsec);
        d.assign(l.dim, l.dim + l.nsec);
        for (int i = 0; i < l.nsec; ++i) {
            if (d[i] < 0) throw Error(BUG_EINVAL, "negative sector dimension");
            if (!idx.emplace(q[i], i).second)
                throw Error(BUG_EINVARIANT, "repeated charge on one leg");
        }
    }
    int find(int32_t charge) const {
        auto it = idx.find(charge);
        return it == idx.end() ? -1 : it->second;
    }
    bug_leg view() const {
        return bug_leg{(int32_t)q.size(), q.data(), d.data()};
    }
};

struct Block {
    int64_t off;
    int rows, cols;
    int32_t col_charge;
};

struct Bsm {
    Leg row, col;
    std::vector<int32_t> shift;
    std::vector<int64_t> offset;
    const void* data = nullptr;
    int64_t nelem = 0;

    Bsm() = default;
    explicit Bsm(const bug_bsm* m, const char* what) {
        if (!m) throw Error(BUG_EINVAL, std::string(what) + ": null descriptor");
        row = Leg(m->row);
        col = Leg(m->col);
        if (m->nslice < 0 || (m->nslice > 0 && (!m->shift || (!m->offset && m->row.nsec > 0))))
            throw Error(BUG_EINVAL, std::string(what) + ": bad slice tables");
        shift.assign(m->shift, m->shift + m->nslice);
        offset.assign(m->offset, m->offset + (size_t)m->nslice * m->row.nsec);
        data = m->data;
        nelem = m->nelem;
        validate(what);
    }
    int nslice() const { return (int)shift.size(); }
    void validate(const char* what) const {
        for (int k = 0; k < nslice(); ++k)
            for (size_t i = 0; i < row.q.size(); ++i) {
                const int64_t off = offset[k * row.q.size() + i];
                if (off < 0) continue;
                const int j = col.find(row.q[i] + shift[k]);
                if (j < 0)
                    throw Error(BUG_EINVARIANT, std::string(what) +
                                                    ": block present for a charge the column "
                                                    "leg does not hold");
                const int64_t end = off + (int64_t)row.d[i] * col.d[j];
                if (end > nelem)
                    throw Error(BUG_EINVARIANT,
                                std::string(what) + ": block extends past the buffer");
            }
        if (nelem > 0 && !data) throw Error(BUG_EINVAL, std::string(what) + ": null data");
    }
    // block of slice k at row sector i (rows x cols), false if absent
    bool block(int k, int i, Block& b) const {
        if (k < 0 || k >= nslice() || i < 0) return false;
        const int64_t off = offset[(size_t)k * row.q.size() + i];
        if (off < 0) return false;
        const int j = col.find(row.q[i] + shift[k]);
        if (j < 0) return false;
        b.off = off;
        b.rows = row.d[i];
        b.cols = col.d[j];
        b.col_charge = col.q[j];
        return true;
    }
    bug_bsm view(void* dev) const {
        bug_bsm v;
        v.row = row.view();
        v.col = col.view();
        v.nslice = nslice();
        v.shift = shift.data();
        v.offset = offset.data();
        v.data = dev;
        v.nelem = nelem;
        return v;
    }
};

// A dense layout for an intermediate family: slices with the given shifts over (row, col) legs,
// every block that the charges allow.
Bsm make_intermediate(const Leg& row, const Leg& col, const std::vector<int32_t>& shifts) {
    Bsm m;
    m.row = row;
    m.col = col;
    m.shift = shifts;
    m.offset.assign(shifts.size() * row.q.size(), -1);
    int64_t off = 0;
    for (size_t k = 0; k < shifts.size(); ++k)
        for (size_t i = 0; i < row.q.size(); ++i) {
            const int j = col.find(row.q[i] + shifts[k]);
            if (j < 0 || row.d[i] == 0 || col.d[j] == 0) continue;
            m.offset[k * row.q.size() + i] = off;
            off += (int64_t)row.d[i] * col.d[j];
        }
    m.nelem = off;
    return m;
}

bool op_plain(int op) { return op == BUG_OP_N || op == BUG_OP_R; }

// Plan out[o] (+)= sum alpha op_a(P[p]) op_b(Q[q]) over every block of out.
void build_product_plan(Plan& plan, const Bsm& out, const Bsm& P, const Bsm& Q,
                        const std::vector<bug_term>& terms, bool accumulate) {
    plan.tasks.clear();
    plan.terms.clear();
    plan.flops = 0.0;
    std::vector<std::vector<int>> by_out(out.nslice());
    for (size_t t = 0; t < terms.size(); ++t) {
        const bug_term& tm = terms[t];
        if (tm.o < 0 || tm.o >= out.nslice() || tm.p < 0 || tm.p >= P.nslice() || tm.q < 0 ||
            tm.q >= Q.nslice() || tm.op_a < 0 || tm.op_a > 3 || tm.op_b < 0 || tm.op_b > 3)
            throw Error(BUG_EINVAL, "product term out of range");
        by_out[tm.o].push_back((int)t);
    }
    for (int o = 0; o < out.nslice(); ++o)
        for (size_t i = 0; i < out.row.q.size(); ++i) {
            Block cb;
            if (!out.block(o, (int)i, cb)) continue;
            if (cb.rows == 0 || cb.cols == 0) continue;
            GemmTask task;
            task.c_off = cb.off;
            task.ldc = cb.rows;
            task.m = cb.rows;
            task.n = cb.cols;
            task.term_begin = (int32_t)plan.terms.size();
            task.beta = accumulate ? 1 : 0;
            const int32_t qi = out.row.q[i];
            const int32_t qj = cb.col_charge;
            for (int t : by_out[o]) {
                const bug_term& tm = terms[t];
                // op_a(P[p]) : rows of charge qi, inner charge qk
                Block pb;
                int32_t qk;
                int kdim;
                if (op_plain(tm.op_a)) {
                    if (!P.block(tm.p, P.row.find(qi), pb)) continue;
                    if (pb.rows != cb.rows) throw Error(BUG_EINVARIANT, "row dims disagree (P)");
                    qk = pb.col_charge;
                    kdim = pb.cols;
                } else {
                    qk = qi - P.shift[tm.p];
                    if (!P.block(tm.p, P.row.find(qk), pb)) continue;
                    if (pb.col_charge != qi) continue;
                    if (pb.cols != cb.rows) throw Error(BUG_EINVARIANT, "row dims disagree (P^T)");
                    kdim = pb.rows;
                }
                // op_b(Q[q]) : rows of charge qk, columns of charge qj
                Block qb;
                if (op_plain(tm.op_b)) {
                    if (!Q.block(tm.q, Q.row.find(qk), qb)) continue;
                    if (qb.col_charge != qj) continue;
                    if (qb.rows != kdim || qb.cols != cb.cols)
                        throw Error(BUG_EINVARIANT, "inner/column dims disagree (Q)");
                } else {
                    if (!Q.block(tm.q, Q.row.find(qj), qb)) continue;
                    if (qb.col_charge != qk) continue;
                    if (qb.cols != kdim || qb.rows != cb.cols)
                        throw Error(BUG_EINVARIANT, "inner/column dims disagree (Q^T)");
                }
                if (kdim == 0) continue;
                GemmTerm g;
                g.a_off = pb.off;
                g.b_off = qb.off;
                g.lda = pb.rows;
                g.ldb = qb.rows;
                g.k = kdim;
                g.ops = tm.op_a | (tm.op_b << 2);
                g.alpha_re = tm.alpha_re;
                g.alpha_im = tm.alpha_im;
                plan.terms.push_back(g);
                plan.flops += 8.0 * cb.rows * cb.cols * (double)kdim;
            }
            task.term_end = (int32_t)plan.terms.size();
            // An overwrite with no terms still has to write zeros; an accumulate with none is a
            // no-op and is dropped.
            if (task.term_end == task.term_begin && accumulate) continue;
            plan.tasks.push_back(task);
        }
}

// Choose a tile configuration from block-size statistics: 16 x 16 or 32 x 32 tiles of one wave
// when most of the work is in blocks 32 or fewer wide, else 32 x 32 tiles of four waves (on the
// step's contractions the autotuner picks the 64 x 64 and 128 x 64 tiles for under 2 % of the
// plan classes: their few workgroups leave most SIMDs idle).
int choose_config(const Plan& plan) {
    double work = 0.0, small = 0.0;
    for (const GemmTask& t : plan.tasks) {
        const double w = (double)t.m * t.n;
        work += w;
        if (t.m <= 32 || t.n <= 32) small += w;
    }
    if (work == 0.0) return 0;
    auto tiles_for = [&](int c) {
        const TileConfig& tc = kTileConfigs[c];
        int64_t n = 0;
        for (const GemmTask& t : plan.tasks)
            n += (int64_t)((t.m + tc.tm() - 1) / tc.tm()) * ((t.n + tc.tn() - 1) / tc.tn());
        return n;
    };
    if (small / work > 0.5) return tiles_for(4) > 4 * 256 ? 1 : 4;
    return 2;
}

// Make the device image of `plan` for tile configuration `config` resident in the plan arena:
// build it on the host once per configuration, copy it in stream order (no host wait).
void build_plan_image(Plan& plan, int config) {
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    if (plan.config != config || plan.image.empty()) {
        const TileConfig& tc = kTileConfigs[config];
        std::vector<GemmTile> tiles;
        for (size_t t = 0; t < plan.tasks.size(); ++t) {
            const GemmTask& task = plan.tasks[t];
            const int mt = (task.m + tc.tm() - 1) / tc.tm();
            const int nt = (task.n + tc.tn() - 1) / tc.tn();
            for (int a = 0; a < mt; ++a)
                for (int b = 0; b < nt; ++b) tiles.push_back(GemmTile{(int32_t)t, a, b, 0});
        }
        const size_t bt = tiles.size() * sizeof(GemmTile);
        const size_t bk = plan.tasks.size() * sizeof(GemmTask);
        const size_t bm = plan.terms.size() * sizeof(GemmTerm);
        plan.image.assign(align(bt) + align(bk) + align(bm) + 256, 0);
        if (bt) std::memcpy(plan.image.data(), tiles.data(), bt);
        if (bk) std::memcpy(plan.image.data() + align(bt), plan.tasks.data(), bk);
        if (bm) std::memcpy(plan.image.data() + align(bt) + align(bk), plan.terms.data(), bm);
        plan.ntiles = (int32_t)tiles.size();
        plan.config = config;
        plan.gen = 0;
        plan.off_tasks = align(bt);
        plan.off_terms = align(bt) + align(bk);
    }
}

void upload_plan(bug_ctx* ctx, Plan& plan, int config, hipStream_t stream) {
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    build_plan_image(plan, config);
    if (std::find(ctx->arena_streams.begin(), ctx->arena_streams.end(), stream) ==
        ctx->arena_streams.end())
        ctx->arena_streams.push_back(stream);
    if (plan.gen == ctx->plan_arena_gen) return;
    if (capturing_on(ctx, stream))
        throw Error(BUG_EINVAL, "a plan would be uploaded during graph capture: run the operation "
                                "once on the capture stream before capturing");
    const size_t total = plan.image.size();
    if (ctx->plan_arena_used + total > ctx->plan_arena_bytes) {
        const size_t prev = ctx->plan_arena_bytes;
        if (ctx->graph_pins > 0) {
            // live graphs hold addresses in this arena: keep it, start a new one
            if (ctx->plan_arena) ctx->retired.push_back(ctx->plan_arena);
            ctx->plan_arena = nullptr;
            ctx->plan_arena_bytes = 0;
        } else {
            // every launch that may read the arena went to one of these streams
            for (hipStream_t s : ctx->arena_streams) HIP_CHECK(hipStreamSynchronize(s));
        }
        ctx->arena_streams.assign(1, stream);
        if (total > ctx->plan_arena_bytes) {
            if (ctx->plan_arena) HIP_CHECK(hipFree(ctx->plan_arena));
            ctx->plan_arena = nullptr;
            const size_t want = std::max<size_t>(std::max<size_t>(total > prev ? 2 * prev : prev, 4 * total),
                                                 (size_t)64 << 20);
            if (hipMalloc(&ctx->plan_arena, want) != hipSuccess) {
                (void)hipGetLastError();
                ctx->plan_arena_bytes = 0;
                throw Error(BUG_ENOMEM, "plan arena allocation failed");
            }
            ctx->plan_arena_bytes = want;
        }
        ctx->plan_arena_used = 0;
        ++ctx->plan_arena_gen;
    }
    char* dev = static_cast<char*>(ctx->plan_arena) + ctx->plan_arena_used;
    ctx->plan_arena_used += align(total);
    stage_upload(ctx, stream, dev, plan.image.data(), total);
    plan.d_tiles = reinterpret_cast<const GemmTile*>(dev);
    plan.d_tasks = reinterpret_cast<const GemmTask*>(dev + plan.off_tasks);
    plan.d_terms = reinterpret_cast<const GemmTerm*>(dev + plan.off_terms);
    plan.gen = ctx->plan_arena_gen;
}

// Look up (or build) the plan of one product stage and launch it.
uint64_t product_key(const Bsm& out, const Bsm& P, const Bsm& Q, const std::vector<bug_term>& terms,
                     bool accumulate, uint64_t tag);

void run_product(bug_ctx* ctx, const Bsm& out, void* out_dev, const Bsm& P, const void* p_dev,
                 const Bsm& Q, const void* q_dev, const std::vector<bug_term>& terms,
                 bool accumulate, int precision, hipStream_t stream, uint64_t tag) {
    launch_cached_plan(ctx, product_key(out, P, Q, terms, accumulate, tag),
                       [&](Plan& p) { build_product_plan(p, out, P, Q, terms, accumulate); },
                       precision, p_dev, q_dev, out_dev, stream);
}

// The plan of one product stage (built on a cache miss), not launched.
Plan* cached_product_plan(bug_ctx* ctx, const Bsm& out, const Bsm& P, const Bsm& Q,
                          const std::vector<bug_term>& terms, bool accumulate, uint64_t tag) {
    return cached_plan(ctx, product_key(out, P, Q, terms, accumulate, tag),
                       [&](Plan& p) { build_product_plan(p, out, P, Q, terms, accumulate); });
}

uint64_t product_key(const Bsm& out, const Bsm& P, const Bsm& Q, const std::vector<bug_term>& terms,
                     bool accumulate, uint64_t tag) {
    BsmHasher h;
    h.pod(tag);
    {
        bug_bsm vo = out.view(nullptr), vp = P.view(nullptr), vq = Q.view(nullptr);
        vo.nelem = vp.nelem = vq.nelem = 0;
        h.bsm(vo);
        h.bsm(vp);
        h.bsm(vq);
    }
    // field by field: bug_term has padding after op_b, whose bytes are unspecified
    for (const bug_term& t : terms) {
        h.pod(t.o);
        h.pod(t.p);
        h.pod(t.q);
        h.pod(t.op_a);
        h.pod(t.op_b);
        h.pod(t.alpha_re);
        h.pod(t.alpha_im);
    }
    h.pod(accumulate);
    return h.h;
}

Plan* cached_plan(bug_ctx* ctx, uint64_t key, const std::function<void(Plan&)>& build) {
    Plan* plan = nullptr;
    auto it = ctx->plan_index.find(key);
    if (it != ctx->plan_index.end()) {
        ctx->plan_lru.splice(ctx->plan_lru.begin(), ctx->plan_lru, it->second);
        plan = it->second->second;
        ++ctx->plan_hits;
    } else {
        ++ctx->plan_misses;
        plan = new Plan();
        try {
            build(*plan);
        } catch (...) {
            delete plan;
            throw;
        }
        ctx->plan_lru.emplace_front(key, plan);
        ctx->plan_index[key] = ctx->plan_lru.begin();
        while (ctx->plan_lru.size() > ctx->plan_capacity) {
            Plan* old = ctx->plan_lru.back().second;
            ctx->plan_index.erase(ctx->plan_lru.back().first);
            ctx->plan_lru.pop_back();
            delete old;
        }
    }
    return plan;
}

void launch_plan(bug_ctx* ctx, Plan& plan, int precision, const void* A, const void* B, void* C,
                 hipStream_t stream) {
    const Schedule sc = pick_schedule(ctx, plan, precision, A, B, stream);
    upload_plan(ctx, plan, sc.config, stream);
    void* partial = nullptr;
    if (sc.splits > 1)
        partial = scratch(ctx, stream, split_partial_bytes(plan, sc.config, sc.splits, precision), 6);
    ctx->last_flops += plan.flops;
    ctx->last_schedules.push_back(sc);
    HIP_CHECK(launch_grouped_gemm(plan, sc.config, sc.splits, precision, A, B, C, partial, stream));
}

void launch_cached_plan(bug_ctx* ctx, uint64_t key, const std::function<void(Plan&)>& build,
                        int precision, const void* A, const void* B, void* C, hipStream_t stream) {
    Plan* plan = cached_plan(ctx, key, build);
    launch_plan(ctx, *plan, precision, A, B, C, stream);
}

void stage_upload(bug_ctx* ctx, hipStream_t stream, void* dev, const void* host, size_t bytes) {
    if (bytes == 0) return;
    if (capturing_on(ctx, stream))
        throw Error(BUG_EINVAL, "a host-to-device upload would be recorded during graph capture: "
                                "run the operation once on the capture stream before capturing");
    constexpr size_t kRing = (size_t)32 << 20;
    if (bytes > kRing / 4) {                 // rare large tables: a plain (pageable) copy
        HIP_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream));
        return;
    }
    if (!ctx->ring) {
        void* p = nullptr;
        HIP_CHECK(hipHostMalloc(&p, kRing));
        ctx->ring = static_cast<char*>(p);
        ctx->ring_bytes = kRing;
        ctx->ring_head = 0;
    }
    const size_t need = (bytes + 255) & ~size_t(255);
    if (ctx->ring_head + need > ctx->ring_bytes) {
        // the ring's start is about to be overwritten: the copies queued out of it must be done
        for (hipStream_t s : ctx->ring_streams) HIP_CHECK(hipStreamSynchronize(s));
        ctx->ring_streams.clear();
        ctx->ring_head = 0;
    }
    if (std::find(ctx->ring_streams.begin(), ctx->ring_streams.end(), stream) == ctx->ring_streams.end())
        ctx->ring_streams.push_back(stream);
    char* src = ctx->ring + ctx->ring_head;
    std::memcpy(src, host, bytes);
    ctx->ring_head += need;
    HIP_CHECK(hipMemcpyAsync(dev, src, bytes, hipMemcpyHostToDevice, stream));
}

void* scratch(bug_ctx* ctx, hipStream_t stream, size_t bytes, int slot) {
    Workspace& w = ctx->scratch[slot][(void*)stream];
    if (w.bytes < bytes) {
        if (capturing_on(ctx, stream))
            throw Error(BUG_EINVAL, "device scratch (slot " + std::to_string(slot) + ", " +
                                        std::to_string(w.bytes) + " -> " + std::to_string(bytes) +
                                        " bytes) would grow during graph capture: run the "
                                        "operation once on the capture stream before capturing");
        if (w.ptr) {
            if (ctx->graph_pins > 0) {
                ctx->retired.push_back(w.ptr);      // a live graph may still use it
            } else {
                HIP_CHECK(hipStreamSynchronize(stream));
                HIP_CHECK(hipFree(w.ptr));
            }
            w.ptr = nullptr;
            w.bytes = 0;
        }
        const size_t want = std::max(bytes, (size_t)(1.5 * (double)bytes));
        if (hipMalloc(&w.ptr, want) != hipSuccess) {
            (void)hipGetLastError();
            w.ptr = nullptr;
            throw Error(BUG_ENOMEM, "device scratch allocation failed");
        }
        w.bytes = want;
    }
    return w.ptr;
}

__global__ void axpy_kernel(double2* __restrict__ y, const double2* __restrict__ x, double a,
                            int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        y[i].x += a * x[i].x;
        y[i].y += a * x[i].y;
    }
}

void same_layout(const Bsm& a, const Bsm& b, const char* what) {
    if (a.row.q != b.row.q || a.row.d != b.row.d || a.col.q != b.col.q || a.col.d != b.col.d ||
        a.shift != b.shift || a.offset != b.offset)
        throw Error(BUG_EINVARIANT, std::string(what) + ": layouts differ");
}

struct MpoSite {
    int wl, wr;
    std::vector<int32_t> cl, cr, a, b, so, si;
    std::vector<double> coeff;
    explicit MpoSite(const bug_mpo_site* w, int d) {
        if (!w) throw Error(BUG_EINVAL, "null MPO site");
        wl = w->wl;
        wr = w->wr;
        if (wl <= 0 || wr <= 0 || w->nent < 0) throw Error(BUG_EINVAL, "bad MPO site dims");
        cl.assign(w->cl, w->cl + wl);
        cr.assign(w->cr, w->cr + wr);
        a.assign(w->a, w->a + w->nent);
        b.assign(w->b, w->b + w->nent);
        so.assign(w->sout, w->sout + w->nent);
        si.assign(w->sin, w->sin + w->nent);
        coeff.assign(w->coeff, w->coeff + 2 * w->nent);
        for (int e = 0; e < w->nent; ++e)
            if (a[e] < 0 || a[e] >= wl || b[e] < 0 || b[e] >= wr || so[e] < 0 || so[e] >= d ||
                si[e] < 0 || si[e] >= d)
                throw Error(BUG_EINVAL, "MPO entry out of range");
    }
    int nent() const { return (int)a.size(); }
};

// The two product stages of a one-site apply: T[(a, sin)] = L_a X_sin, then
// Y_sout = sum coeff T[(a, sin)] R_b^T (effective_operators.jl:39-64).
struct OneSiteStages {
    Bsm T;
    std::vector<bug_term> t1, t2;
};

OneSiteStages one_site_stages(const Bsm& l, const MpoSite& w, const Bsm& r, const Bsm& x,
                              const Bsm& y) {
    if (l.nslice() != w.wl || r.nslice() != w.wr)
        throw Error(BUG_EINVAL, "environment channels do not match the MPO site");
    if (y.nslice() != x.nslice()) throw Error(BUG_EINVAL, "X/Y local dims differ");
    OneSiteStages sg;
    std::vector<int> slot(w.wl * x.nslice(), -1);
    std::vector<int32_t> shifts;
    for (int e = 0; e < w.nent(); ++e) {
        int& s = slot[w.a[e] * x.nslice() + w.si[e]];
        if (s >= 0) continue;
        s = (int)shifts.size();
        shifts.push_back(l.shift[w.a[e]] + x.shift[w.si[e]]);
        sg.t1.push_back(bug_term{s, w.a[e], w.si[e], BUG_OP_N, BUG_OP_N, 1.0, 0.0});
    }
    sg.T = make_intermediate(l.row, x.col, shifts);
    for (int e = 0; e < w.nent(); ++e)
        sg.t2.push_back(bug_term{w.so[e], slot[w.a[e] * x.nslice() + w.si[e]], w.b[e], BUG_OP_N,
                                 BUG_OP_T, w.coeff[2 * e], w.coeff[2 * e + 1]});
    return sg;
}

// Y -= shift X (the H_eff - shift of the Krylov recurrences).
void shift_into(const Bsm& x, const Bsm& y, const void* xd, void* yd, double shift, hipStream_t st) {
    same_layout(x, y, "apply_one_site shift");
    const int64_t n = x.nelem;
    if (n > 0)
        hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           static_cast<double2*>(yd), static_cast<const double2*>(xd), -shift, n);
    HIP_CHECK(hipGetLastError());
}

// Append the tasks and terms of `src` to `dst` with C offsets moved by c_shift and the A and B
// offsets of its terms by a_shift and b_shift (elements): several instances' plans as one.
void append_plan(Plan& dst, const Plan& src, int64_t c_shift, int64_t a_shift, int64_t b_shift) {
    const int32_t base = (int32_t)dst.terms.size();
    for (GemmTask t : src.tasks) {
        t.c_off += c_shift;
        t.term_begin += base;
        t.term_end += base;
        dst.tasks.push_back(t);
    }
    for (GemmTerm u : src.terms) {
        u.a_off += a_shift;
        u.b_off += b_shift;
        dst.terms.push_back(u);
    }
    dst.flops += src.flops;
}

// ---- the allocation-free entry point's workspace ------------------------------------------------
//
// A file-scope device array, one copy per GPU, that bug_apply_one_site_noalloc carves into the
// stage-1 intermediate, the two plan images and the split-k partial tiles of each call.  Its device
// address is looked up once per device (with that device current) when a context is created there;
// host code uses that address, never the array's name.  Every call on one GPU shares the array and
// leaves its contents behind: the intermediate and the partials are written in full before they are
// read, so nothing needs clearing, and calls that overlap on one GPU (different streams or threads)
// race on it.
namespace noalloc {
constexpr size_t kBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kBytes];
constexpr int kMaxDevices = 64;
std::mutex mu;
void* ptr[kMaxDevices] = {};
bool looked_up[kMaxDevices] = {};

// With `device` current: the array's address on it (once per device; null if the lookup failed).
void setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(mu);
    if (looked_up[device]) return;
    looked_up[device] = true;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) {
        ptr[device] = p;
    } else {
        (void)hipGetLastError();
        ptr[device] = nullptr;
    }
}

void* address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    return ptr[device];
}
}  // namespace noalloc

}  // namespace bug

using namespace bug;

namespace {

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int bug_ctx_create(int device, bug_ctx** out) {
    if (!out) return BUG_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        return BUG_EDEVICE;
    }
    bug_ctx* ctx = new bug_ctx();
    ctx->device = device;
    {
        // the allocation-free entry point's workspace address on this device, with it current
        int prev = 0;
        if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(device) == hipSuccess) {
            noalloc::setup(device);
            (void)hipSetDevice(prev);
        } else {
            (void)hipGetLastError();
        }
    }
    *out = ctx;
    return BUG_OK;
}

int bug_ctx_destroy(bug_ctx* ctx) {
    if (!ctx) return BUG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto& e : ctx->plan_lru) delete e.second;
    if (ctx->plan_arena) (void)hipFree(ctx->plan_arena);
    for (auto& slot : ctx->scratch)
        for (auto& s : slot)
            if (s.second.ptr) (void)hipFree(s.second.ptr);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->ring) (void)hipHostFree(ctx->ring);
    if (ctx->svd_sweeps_dev) (void)hipFree(ctx->svd_sweeps_dev);
    for (hipEvent_t e : ctx->tune_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->sync_ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : ctx->retired) (void)hipFree(p);
    delete ctx;
    return BUG_OK;
}

const char* bug_last_error(const bug_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int bug_set_tile_config(bug_ctx* ctx, int config) {
    if (!ctx || config < -1 || config >= kNumTileConfigs) return BUG_EINVAL;
    ctx->forced_config = config;
    return BUG_OK;
}

int bug_tile_config_count(void) { return kNumTileConfigs; }

int bug_ctx_stats(const bug_ctx* ctx, int64_t* out4) {
    if (!ctx || !out4) return BUG_EINVAL;
    int64_t bytes = 0;
    for (auto& slot : ctx->scratch)
        for (auto& s : slot) bytes += (int64_t)s.second.bytes;
    bytes += (int64_t)ctx->plan_arena_bytes;
    out4[0] = bytes;
    out4[1] = (int64_t)ctx->plan_lru.size();
    out4[2] = ctx->plan_hits;
    out4[3] = ctx->plan_misses;
    return BUG_OK;
}

int bug_graph_capture_begin(bug_ctx* ctx, void* stream) {
    return guarded(ctx, [&] {
        if (!stream) throw Error(BUG_EINVAL, "graph capture needs a non-default stream");
        if (ctx->capture_active) throw Error(BUG_EINVAL, "a graph capture is already in progress");
        ctx->capture_active = true;
        ctx->capturing = as_stream(stream);
        ++ctx->graph_pins;
    });
}

int bug_graph_capture_end(bug_ctx* ctx) {
    return guarded(ctx, [&] {
        if (!ctx->capture_active) throw Error(BUG_EINVAL, "no graph capture in progress");
        ctx->capture_active = false;
        ctx->capturing = nullptr;
    });
}

int bug_graph_release(bug_ctx* ctx, void* stream) {
    return guarded(ctx, [&] {
        if (ctx->graph_pins <= 0) throw Error(BUG_EINVAL, "no live graph to release");
        if (ctx->capture_active) throw Error(BUG_EINVAL, "release during a graph capture");
        HIP_CHECK(hipDeviceSynchronize());
        // the capture stream's scratch belonged to the graph
        if (stream) {
            for (auto& slot : ctx->scratch) {
                auto it = slot.find(stream);
                if (it == slot.end()) continue;
                if (it->second.ptr) ctx->retired.push_back(it->second.ptr);
                slot.erase(it);
            }
        }
        if (--ctx->graph_pins == 0) {
            for (void* p : ctx->retired) HIP_CHECK(hipFree(p));
            ctx->retired.clear();
        }
    });
}

int bug_last_schedules(const bug_ctx* ctx, int32_t cap, int32_t* out, int32_t* count) {
    if (!ctx || !count || cap < 0 || (cap > 0 && !out)) return BUG_EINVAL;
    const int32_t n = (int32_t)ctx->last_schedules.size();
    for (int32_t i = 0; i < std::min(n, cap); ++i) {
        out[2 * i] = ctx->last_schedules[i].config;
        out[2 * i + 1] = ctx->last_schedules[i].splits;
    }
    *count = n;
    return BUG_OK;
}

int bug_last_op_flops(const bug_ctx* ctx, double* flops) {
    if (!ctx || !flops) return BUG_EINVAL;
    *flops = ctx->last_flops;
    return BUG_OK;
}

int bug_bsm_product(bug_ctx* ctx, const bug_bsm* out, const bug_bsm* P, const bug_bsm* Q,
                    int32_t nterm, const bug_term* terms, int32_t accumulate, int32_t precision,
                    void* stream) {
    return guarded(ctx, [&] {
        if (nterm < 0 || (nterm > 0 && !terms)) throw Error(BUG_EINVAL, "bad term list");
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        Bsm o(out, "out"), p(P, "P"), q(Q, "Q");
        std::vector<bug_term> tl(terms, terms + nterm);
        run_product(ctx, o, out->data, p, P->data, q, Q->data, tl, accumulate != 0, precision,
                    as_stream(stream), 0x01);
    });
}

int bug_env_push_left(bug_ctx* ctx, const bug_bsm* L, const bug_bsm* A, const bug_bsm* bra,
                      const bug_mpo_site* W, const bug_bsm* Lout, int32_t precision,
                      void* stream) {
    return guarded(ctx, [&] {
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        Bsm l(L, "L"), a(A, "A"), lo(Lout, "Lout");
        Bsm br = bra ? Bsm(bra, "bra") : a;
        const void* bra_dev = bra ? bra->data : A->data;
        MpoSite w(W, a.nslice());
        if (l.nslice() != w.wl || lo.nslice() != w.wr)
            throw Error(BUG_EINVAL, "environment channels do not match the MPO site");
        if (br.nslice() != a.nslice()) throw Error(BUG_EINVAL, "bra/ket local dims differ");
        // stage 1: T[(a, sin)] = L_a A_sin
        std::vector<int> slot(w.wl * a.nslice(), -1);
        std::vector<int32_t> shifts;
        std::vector<bug_term> t1;
        for (int e = 0; e < w.nent(); ++e) {
            int& s = slot[w.a[e] * a.nslice() + w.si[e]];
            if (s >= 0) continue;
            s = (int)shifts.size();
            shifts.push_back(l.shift[w.a[e]] + a.shift[w.si[e]]);
            t1.push_back(bug_term{s, w.a[e], w.si[e], BUG_OP_N, BUG_OP_N, 1.0, 0.0});
        }
        Bsm T = make_intermediate(l.row, a.col, shifts);
        hipStream_t st = as_stream(stream);
        void* tdev = scratch(ctx, st, sizeof(double2) * std::max<int64_t>(T.nelem, 1));
        run_product(ctx, T, tdev, l, L->data, a, A->data, t1, false, precision, st, 0x11);
        // stage 2: Lout_b = sum coeff bra_sout^H T[(a, sin)]
        std::vector<bug_term> t2;
        for (int e = 0; e < w.nent(); ++e)
            t2.push_back(bug_term{w.b[e], w.so[e], slot[w.a[e] * a.nslice() + w.si[e]],
                                  BUG_OP_C, BUG_OP_N, w.coeff[2 * e], w.coeff[2 * e + 1]});
        run_product(ctx, lo, Lout->data, br, bra_dev, T, tdev, t2, false, precision, st, 0x12);
    });
}

int bug_env_push_right(bug_ctx* ctx, const bug_bsm* R, const bug_bsm* B, const bug_bsm* bra,
                       const bug_mpo_site* W, const bug_bsm* Rout, int32_t precision,
                       void* stream) {
    return guarded(ctx, [&] {
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        Bsm r(R, "R"), b(B, "B"), ro(Rout, "Rout");
        Bsm br = bra ? Bsm(bra, "bra") : b;
        const void* bra_dev = bra ? bra->data : B->data;
        MpoSite w(W, b.nslice());
        if (r.nslice() != w.wr || ro.nslice() != w.wl)
            throw Error(BUG_EINVAL, "environment channels do not match the MPO site");
        if (br.nslice() != b.nslice()) throw Error(BUG_EINVAL, "bra/ket local dims differ");
        // stage 1: U[(b, sin)] = R_b B_sin^T
        std::vector<int> slot(w.wr * b.nslice(), -1);
        std::vector<int32_t> shifts;
        std::vector<bug_term> t1;
        for (int e = 0; e < w.nent(); ++e) {
            int& s = slot[w.b[e] * b.nslice() + w.si[e]];
            if (s >= 0) continue;
            s = (int)shifts.size();
            shifts.push_back(r.shift[w.b[e]] - b.shift[w.si[e]]);
            t1.push_back(bug_term{s, w.b[e], w.si[e], BUG_OP_N, BUG_OP_T, 1.0, 0.0});
        }
        Bsm U = make_intermediate(r.row, b.row, shifts);
        hipStream_t st = as_stream(stream);
        void* udev = scratch(ctx, st, sizeof(double2) * std::max<int64_t>(U.nelem, 1));
        run_product(ctx, U, udev, r, R->data, b, B->data, t1, false, precision, st, 0x21);
        // stage 2: Rout_a = sum coeff conj(bra_sout) U[(b, sin)]
        std::vector<bug_term> t2;
        for (int e = 0; e < w.nent(); ++e)
            t2.push_back(bug_term{w.a[e], w.so[e], slot[w.b[e] * b.nslice() + w.si[e]],
                                  BUG_OP_R, BUG_OP_N, w.coeff[2 * e], w.coeff[2 * e + 1]});
        run_product(ctx, ro, Rout->data, br, bra_dev, U, udev, t2, false, precision, st, 0x22);
    });
}

int bug_apply_one_site(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W, const bug_bsm* R,
                       const bug_bsm* X, const bug_bsm* Y, double shift, int32_t precision,
                       void* stream) {
    return guarded(ctx, [&] {
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        Bsm l(L, "L"), r(R, "R"), x(X, "X"), y(Y, "Y");
        MpoSite w(W, x.nslice());
        const OneSiteStages sg = one_site_stages(l, w, r, x, y);
        hipStream_t st = as_stream(stream);
        void* tdev = scratch(ctx, st, sizeof(double2) * std::max<int64_t>(sg.T.nelem, 1));
        run_product(ctx, sg.T, tdev, l, L->data, x, X->data, sg.t1, false, precision, st, 0x31);
        run_product(ctx, y, Y->data, sg.T, tdev, r, R->data, sg.t2, false, precision, st, 0x32);
        if (shift != 0.0) shift_into(x, y, X->data, Y->data, shift, st);
    });
}

int bug_apply_one_site_noalloc(bug_ctx* ctx, const bug_bsm* L, const bug_mpo_site* W,
                               const bug_bsm* R, const bug_bsm* X, const bug_bsm* Y, double shift,
                               int32_t precision, void* stream) {
    return guarded(ctx, [&] {
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        char* base = static_cast<char*>(noalloc::address(ctx->device));
        if (!base) throw Error(BUG_EDEVICE, "the static workspace has no device address on this GPU");
        Bsm l(L, "L"), r(R, "R"), x(X, "X"), y(Y, "Y");
        MpoSite w(W, x.nslice());
        const OneSiteStages sg = one_site_stages(l, w, r, x, y);
        hipStream_t st = as_stream(stream);
        struct NoTune {
            bug_ctx* c;
            ~NoTune() { c->no_tune = false; }
        } reset{ctx};
        ctx->no_tune = true;
        // the two stages' plans and schedules exactly as bug_apply_one_site takes them
        Plan* p1 = cached_product_plan(ctx, sg.T, l, x, sg.t1, false, 0x31);
        Plan* p2 = cached_product_plan(ctx, y, sg.T, r, sg.t2, false, 0x32);
        const Schedule s1 = pick_schedule(ctx, *p1, precision, L->data, X->data, st);
        const Schedule s2 = pick_schedule(ctx, *p2, precision, nullptr, R->data, st);
        build_plan_image(*p1, s1.config);
        build_plan_image(*p2, s2.config);
        auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t tb = align(sizeof(double2) * (size_t)std::max<int64_t>(sg.T.nelem, 1));
        const size_t i1 = align(p1->image.size()), i2 = align(p2->image.size());
        const size_t q1 = s1.splits > 1 ? align(split_partial_bytes(*p1, s1.config, s1.splits, precision)) : 0;
        const size_t q2 = s2.splits > 1 ? align(split_partial_bytes(*p2, s2.config, s2.splits, precision)) : 0;
        const size_t need = tb + i1 + i2 + q1 + q2;
        if (need > noalloc::kBytes)
            throw Error(BUG_EINVAL, "this apply needs " + std::to_string(need) +
                                        " bytes of workspace; the allocation-free entry point holds " +
                                        std::to_string(noalloc::kBytes));
        char* tdev = base;
        char* img1 = tdev + tb;
        char* img2 = img1 + i1;
        char* part1 = q1 ? img2 + i2 : nullptr;
        char* part2 = q2 ? img2 + i2 + q1 : nullptr;
        HIP_CHECK(hipMemcpyAsync(img1, p1->image.data(), p1->image.size(), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(img2, p2->image.data(), p2->image.size(), hipMemcpyHostToDevice, st));
        auto launch = [&](Plan& p, const Schedule& sc, char* img, const void* A, const void* B, void* C,
                          void* part) {
            const GemmTile* tiles = p.d_tiles;
            const GemmTask* tasks = p.d_tasks;
            const GemmTerm* terms = p.d_terms;
            p.d_tiles = reinterpret_cast<const GemmTile*>(img);
            p.d_tasks = reinterpret_cast<const GemmTask*>(img + p.off_tasks);
            p.d_terms = reinterpret_cast<const GemmTerm*>(img + p.off_terms);
            const hipError_t e = launch_grouped_gemm(p, sc.config, sc.splits, precision, A, B, C, part, st);
            p.d_tiles = tiles;
            p.d_tasks = tasks;
            p.d_terms = terms;
            HIP_CHECK(e);
            ctx->last_flops += p.flops;
            ctx->last_schedules.push_back(sc);
        };
        launch(*p1, s1, img1, L->data, X->data, tdev, part1);
        launch(*p2, s2, img2, tdev, R->data, Y->data, part2);
        if (shift != 0.0) shift_into(x, y, X->data, Y->data, shift, st);
    });
}

int bug_apply_one_site_batch(bug_ctx* ctx, int32_t k, const bug_bsm* L, const bug_mpo_site* W,
                             const bug_bsm* R, const bug_bsm* X, const bug_bsm* Y,
                             const double* shifts, int32_t precision, void* stream) {
    return guarded(ctx, [&] {
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        if (k < 0 || (k > 0 && (!L || !W || !R || !X || !Y)))
            throw Error(BUG_EINVAL, "bad batch arguments");
        if (k == 0) return;
        std::vector<Bsm> ls, rs, xs, ys;
        std::vector<OneSiteStages> sg;
        std::vector<int64_t> toff(k);
        int64_t ttot = 0;
        for (int i = 0; i < k; ++i) {
            ls.emplace_back(&L[i], "L");
            rs.emplace_back(&R[i], "R");
            xs.emplace_back(&X[i], "X");
            ys.emplace_back(&Y[i], "Y");
            MpoSite w(&W[i], xs[i].nslice());
            sg.push_back(one_site_stages(ls[i], w, rs[i], xs[i], ys[i]));
            toff[i] = ttot;
            ttot += std::max<int64_t>(sg[i].T.nelem, 1);
        }
        hipStream_t st = as_stream(stream);
        double2* tdev = static_cast<double2*>(scratch(ctx, st, sizeof(double2) * ttot));
        // element offset of ptr from base (both device addresses of double2 data)
        auto lowest = [&](const bug_bsm* d) {
            const char* b = static_cast<const char*>(d[0].data);
            for (int i = 1; i < k; ++i) b = std::min(b, static_cast<const char*>(d[i].data));
            return b;
        };
        auto rel = [](const void* p, const char* base) {
            const int64_t bytes = static_cast<const char*>(p) - base;
            if (bytes % (int64_t)sizeof(double2) != 0)
                throw Error(BUG_EINVAL, "batched operands must be 16-byte aligned");
            return bytes / (int64_t)sizeof(double2);
        };
        // The merged plan of a stage is cached under the instances' plan keys and their operand
        // offsets from the lowest base: callers that keep their operands in fixed buffers (a
        // packed Krylov basis) reuse it call after call.
        auto merged = [&](int stage, const char* cb, const char* ab, const char* bb) {
            std::vector<uint64_t> keys(k);
            std::vector<int64_t> co(k), ao(k), bo(k);
            Hasher h;
            h.pod((uint64_t)(0xBA7C0 + stage));
            for (int i = 0; i < k; ++i) {
                if (stage == 1) {
                    keys[i] = product_key(sg[i].T, ls[i], xs[i], sg[i].t1, false, 0x31);
                    co[i] = toff[i];
                    ao[i] = rel(L[i].data, ab);
                    bo[i] = rel(X[i].data, bb);
                } else {
                    keys[i] = product_key(ys[i], sg[i].T, rs[i], sg[i].t2, false, 0x32);
                    co[i] = rel(Y[i].data, cb);
                    ao[i] = toff[i];
                    bo[i] = rel(R[i].data, bb);
                }
                h.pod(keys[i]);
                h.pod(co[i]);
                h.pod(ao[i]);
                h.pod(bo[i]);
            }
            return cached_plan(ctx, h.h, [&](Plan& m) {
                for (int i = 0; i < k; ++i) {
                    const Plan* p = stage == 1
                        ? cached_product_plan(ctx, sg[i].T, ls[i], xs[i], sg[i].t1, false, 0x31)
                        : cached_product_plan(ctx, ys[i], sg[i].T, rs[i], sg[i].t2, false, 0x32);
                    append_plan(m, *p, co[i], ao[i], bo[i]);
                }
            });
        };
        // stage 1 of every instance in one launch: A = L_i, B = X_i, C = T_i
        const char* lb = lowest(L);
        const char* xb = lowest(X);
        Plan* m1 = merged(1, nullptr, lb, xb);
        launch_plan(ctx, *m1, precision, lb, xb, tdev, st);
        // stage 2: A = T_i, B = R_i, C = Y_i
        const char* rb = lowest(R);
        const char* yb = lowest(Y);
        Plan* m2 = merged(2, yb, nullptr, rb);
        launch_plan(ctx, *m2, precision, tdev, rb, const_cast<char*>(yb), st);
        for (int i = 0; i < k; ++i)
            if (shifts && shifts[i] != 0.0) shift_into(xs[i], ys[i], X[i].data, Y[i].data, shifts[i], st);
    });
}

int bug_apply_zero_site_batch(bug_ctx* ctx, int32_t k, const bug_bsm* L, const bug_bsm* R,
                              const bug_bsm* C, const bug_bsm* Y, const double* shifts,
                              int32_t precision, void* stream) {
    return guarded(ctx, [&] {
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        if (k < 0 || (k > 0 && (!L || !R || !C || !Y)))
            throw Error(BUG_EINVAL, "bad batch arguments");
        if (k == 0) return;
        std::vector<Bsm> ls, rs, cs, ys, ts;
        std::vector<std::vector<bug_term>> t1(k), t2(k);
        std::vector<int64_t> toff(k);
        int64_t ttot = 0;
        for (int i = 0; i < k; ++i) {
            ls.emplace_back(&L[i], "L");
            rs.emplace_back(&R[i], "R");
            cs.emplace_back(&C[i], "C");
            ys.emplace_back(&Y[i], "Y");
            if (ls[i].nslice() != rs[i].nslice())
                throw Error(BUG_EINVAL, "root environments have different channel counts");
            if (cs[i].nslice() != 1 || ys[i].nslice() != 1)
                throw Error(BUG_EINVAL, "a zero-site core has one slice");
            std::vector<int32_t> sh;
            for (int w = 0; w < ls[i].nslice(); ++w) {
                sh.push_back(ls[i].shift[w] + cs[i].shift[0]);
                t1[i].push_back(bug_term{w, w, 0, BUG_OP_N, BUG_OP_N, 1.0, 0.0});
                t2[i].push_back(bug_term{0, w, w, BUG_OP_N, BUG_OP_T, 1.0, 0.0});
            }
            ts.push_back(make_intermediate(ls[i].row, cs[i].col, sh));
            toff[i] = ttot;
            ttot += std::max<int64_t>(ts[i].nelem, 1);
        }
        hipStream_t st = as_stream(stream);
        double2* tdev = static_cast<double2*>(scratch(ctx, st, sizeof(double2) * ttot));
        auto lowest = [&](const bug_bsm* d) {
            const char* b = static_cast<const char*>(d[0].data);
            for (int i = 1; i < k; ++i) b = std::min(b, static_cast<const char*>(d[i].data));
            return b;
        };
        auto rel = [](const void* p, const char* base) {
            const int64_t bytes = static_cast<const char*>(p) - base;
            if (bytes % (int64_t)sizeof(double2) != 0)
                throw Error(BUG_EINVAL, "batched operands must be 16-byte aligned");
            return bytes / (int64_t)sizeof(double2);
        };
        // merged plans keyed by the instances' plan keys and operand offsets from the lowest base,
        // as for the one-site batch: fixed packed buffers reuse them call after call
        auto merged = [&](int stage, const char* cb, const char* ab, const char* bb) {
            std::vector<int64_t> co(k), ao(k), bo(k);
            std::vector<uint64_t> keys(k);
            Hasher h;
            h.pod((uint64_t)(0x2E50 + stage));
            for (int i = 0; i < k; ++i) {
                if (stage == 1) {
                    keys[i] = product_key(ts[i], ls[i], cs[i], t1[i], false, 0x41);
                    co[i] = toff[i];
                    ao[i] = rel(L[i].data, ab);
                    bo[i] = rel(C[i].data, bb);
                } else {
                    keys[i] = product_key(ys[i], ts[i], rs[i], t2[i], false, 0x42);
                    co[i] = rel(Y[i].data, cb);
                    ao[i] = toff[i];
                    bo[i] = rel(R[i].data, bb);
                }
                h.pod(keys[i]);
                h.pod(co[i]);
                h.pod(ao[i]);
                h.pod(bo[i]);
            }
            return cached_plan(ctx, h.h, [&](Plan& m) {
                for (int i = 0; i < k; ++i) {
                    const Plan* p = stage == 1
                        ? cached_product_plan(ctx, ts[i], ls[i], cs[i], t1[i], false, 0x41)
                        : cached_product_plan(ctx, ys[i], ts[i], rs[i], t2[i], false, 0x42);
                    append_plan(m, *p, co[i], ao[i], bo[i]);
                }
            });
        };
        // stage 1 of every instance in one launch: A = L_i, B = C_i, out T_i
        const char* lb = lowest(L);
        const char* cb = lowest(C);
        Plan* m1 = merged(1, nullptr, lb, cb);
        launch_plan(ctx, *m1, precision, lb, cb, tdev, st);
        // stage 2: A = T_i, B = R_i, out Y_i
        const char* rb = lowest(R);
        const char* yb = lowest(Y);
        Plan* m2 = merged(2, yb, nullptr, rb);
        launch_plan(ctx, *m2, precision, tdev, rb, const_cast<char*>(yb), st);
        for (int i = 0; i < k; ++i)
            if (shifts && shifts[i] != 0.0) shift_into(cs[i], ys[i], C[i].data, Y[i].data, shifts[i], st);
    });
}

int bug_apply_zero_site(bug_ctx* ctx, const bug_bsm* L, const bug_bsm* R, const bug_bsm* C,
                        const bug_bsm* Y, double shift, int32_t precision, void* stream) {
    return guarded(ctx, [&] {
        if (precision != BUG_PREC_FP64 && precision != BUG_PREC_FP32)
            throw Error(BUG_EINVAL, "unknown precision");
        Bsm l(L, "L"), r(R, "R"), c(C, "C"), y(Y, "Y");
        if (l.nslice() != r.nslice())
            throw Error(BUG_EINVAL, "root environments have different channel counts");
        if (c.nslice() != 1 || y.nslice() != 1)
            throw Error(BUG_EINVAL, "a zero-site core has one slice");
        // stage 1: T_w = L_w C
        std::vector<int32_t> shifts;
        std::vector<bug_term> t1, t2;
        for (int k = 0; k < l.nslice(); ++k) {
            shifts.push_back(l.shift[k] + c.shift[0]);
            t1.push_back(bug_term{k, k, 0, BUG_OP_N, BUG_OP_N, 1.0, 0.0});
            t2.push_back(bug_term{0, k, k, BUG_OP_N, BUG_OP_T, 1.0, 0.0});
        }
        Bsm T = make_intermediate(l.row, c.col, shifts);
        hipStream_t st = as_stream(stream);
        void* tdev = scratch(ctx, st, sizeof(double2) * std::max<int64_t>(T.nelem, 1));
        run_product(ctx, T, tdev, l, L->data, c, C->data, t1, false, precision, st, 0x41);
        // stage 2: Y = sum_w T_w R_w^T
        run_product(ctx, y, Y->data, T, tdev, r, R->data, t2, false, precision, st, 0x42);
        if (shift != 0.0) {
            same_layout(c, y, "apply_zero_site shift");
            const int64_t n = c.nelem;
            if (n > 0)
                hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                                   st, static_cast<double2*>(Y->data),
                                   static_cast<const double2*>(C->data), -shift, n);
            HIP_CHECK(hipGetLastError());
        }
    });
}

}  // extern "C"
