// Schedules of the grouped GEMM: the tile configuration and the split-k factor of a plan.
//
// A plan's schedule comes from (in order) the context's forced choice, the process-wide tuning
// table, or the heuristic below.  The table maps a plan CLASS (precision and the bucketed block
// statistics of the plan, so that a decision carries over to plans of similar shape in later
// steps and other contexts) to the schedule the autotuner measured fastest: on the first launch
// of a class in tuning mode every candidate schedule is timed with HIP events on the caller's
// stream (into a scratch copy of the output, so the caller's data are untouched) and the winner
// is recorded.  The table is shared by every context of the process, so concurrent requests
// that see the same class use the same schedule; it can be saved to and loaded from a text file
// for reproducible runs.
//
// Tile configurations give bitwise-identical results (every one accumulates each output element
// over k in the same order).  The split factor changes only the association of the k windows'
// partial sums, which split_reduce adds in window order: results stay deterministic for a given
// schedule and agree with the unsplit launch to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <map>

#include "internal.h"

namespace bug {

namespace {

// A tuned class: its schedule and, for the record, the class statistics and the measured times
// of the winner and of the heuristic's choice (zero for entries loaded from a file).
struct TuneEntry {
    Schedule s;
    int precision = -1;
    double m = 0, n = 0, k = 0, tasks = 0, flops = 0;
    float best_ms = 0.0f, heuristic_ms = 0.0f;
};

struct TuneTable {
    std::mutex mu;
    std::map<uint64_t, TuneEntry> entries;
    std::atomic<int> mode{0};
    std::atomic<uint64_t> gen{1};
    std::atomic<int64_t> tuned{0};
};

TuneTable& table() {
    static TuneTable t;
    return t;
}

int bucket(double x) { return x > 0.0 ? (int)std::lround(2.0 * std::log2(x)) : -1; }

}  // namespace

namespace {

// Flop-weighted mean m, n and total k of a plan's tasks.
void class_stats(const Plan& plan, double& m, double& n, double& k) {
    double wsum = 0.0, mm = 0.0, mn = 0.0, mk = 0.0;
    for (const GemmTask& t : plan.tasks) {
        double kt = 0.0;
        for (int i = t.term_begin; i < t.term_end; ++i) kt += plan.terms[i].k;
        const double w = std::max(1.0, (double)t.m * t.n * std::max(kt, 1.0));
        wsum += w;
        mm += w * t.m;
        mn += w * t.n;
        mk += w * kt;
    }
    m = wsum > 0 ? mm / wsum : 0.0;
    n = wsum > 0 ? mn / wsum : 0.0;
    k = wsum > 0 ? mk / wsum : 0.0;
}

}  // namespace

// The class of a plan: precision and the bucketed (half-octave) flop-weighted mean m, n and
// total k of its tasks, their count and the total flops.
uint64_t plan_class(const Plan& plan, int precision) {
    double m, n, k;
    class_stats(plan, m, n, k);
    Hasher h;
    h.pod((uint64_t)0x5C4ED);
    h.pod(precision);
    const int f[5] = {bucket(m), bucket(n), bucket(k), bucket((double)plan.tasks.size()),
                      bucket(plan.flops)};
    for (int v : f) h.pod(v);
    return h.h;
}

// The heuristic: the tile configuration by block-size statistics (choose_config); the k stages
// split in 2 or 4 windows when the tasks' total k (flop-weighted mean) reaches 512 or 1024 and
// the launch does not already hold 16 waves per CU.
Schedule heuristic_schedule(const Plan& plan) {
    Schedule s{choose_config(plan), 1};
    double m, n, k;
    class_stats(plan, m, n, k);
    const TileConfig& tc = kTileConfigs[s.config];
    int64_t tiles = 0;
    for (const GemmTask& t : plan.tasks)
        tiles += (int64_t)((t.m + tc.tm() - 1) / tc.tm()) * ((t.n + tc.tn() - 1) / tc.tn());
    const int64_t waves = tiles * (tc.threads() / 64);
    if (waves < 16 * 256) s.splits = k >= 1024.0 ? 4 : (k >= 512.0 ? 2 : 1);
    return s;
}

namespace {

// Elements of C that the plan's tasks touch (the extent a scratch stand-in needs).
int64_t output_extent(const Plan& plan) {
    int64_t e = 1;
    for (const GemmTask& t : plan.tasks)
        if (t.m > 0 && t.n > 0) e = std::max<int64_t>(e, t.c_off + (int64_t)(t.n - 1) * t.ldc + t.m);
    return e;
}

// Time every candidate schedule of `plan` on `stream` (outputs into scratch) and return the
// fastest; ties keep the earlier candidate (unsplit first).
TuneEntry tune(bug_ctx* ctx, Plan& plan, int precision, const void* A, const void* B,
               hipStream_t stream) {
    if (!ctx->tune_ev[0]) {
        HIP_CHECK(hipEventCreate(&ctx->tune_ev[0]));
        HIP_CHECK(hipEventCreate(&ctx->tune_ev[1]));
    }
    void* Cs = scratch(ctx, stream, sizeof(double2) * (size_t)output_extent(plan), 8);
    const Schedule heur = heuristic_schedule(plan);
    Schedule best = heur;
    float best_ms = 1e30f, heur_ms = 0.0f;
    static const int kSplits[] = {1, 2, 4, 8};
    for (int split : kSplits) {
        for (int c = 0; c < kNumTileConfigs; ++c) {
            upload_plan(ctx, plan, c, stream);
            void* part = nullptr;
            if (split > 1) part = scratch(ctx, stream, split_partial_bytes(plan, c, split, precision), 6);
            HIP_CHECK(launch_grouped_gemm(plan, c, split, precision, A, B, Cs, part, stream));
            HIP_CHECK(hipEventRecord(ctx->tune_ev[0], stream));
            constexpr int kReps = 3;
            for (int r = 0; r < kReps; ++r)
                HIP_CHECK(launch_grouped_gemm(plan, c, split, precision, A, B, Cs, part, stream));
            HIP_CHECK(hipEventRecord(ctx->tune_ev[1], stream));
            HIP_CHECK(hipEventSynchronize(ctx->tune_ev[1]));
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, ctx->tune_ev[0], ctx->tune_ev[1]));
            ms /= kReps;
            if (c == heur.config && split == heur.splits) heur_ms = ms;
            if (ms < best_ms) {
                best_ms = ms;
                best = Schedule{c, split};
            }
        }
    }
    TuneEntry e;
    e.s = best;
    e.precision = precision;
    class_stats(plan, e.m, e.n, e.k);
    e.tasks = (double)plan.tasks.size();
    e.flops = plan.flops;
    e.best_ms = best_ms;
    e.heuristic_ms = heur_ms;
    return e;
}

}  // namespace

Schedule pick_schedule(bug_ctx* ctx, Plan& plan, int precision, const void* A, const void* B,
                       hipStream_t stream) {
    TuneTable& tt = table();
    const uint64_t gen = tt.gen.load();
    const int mode = tt.mode.load();
    Schedule s;
    if (plan.sched_gen == gen && plan.sched_mode == mode && plan.sched_prec == precision) {
        s = plan.sched;
    } else {
        s = heuristic_schedule(plan);
        if (mode > 0 && !plan.tasks.empty()) {
            const uint64_t cls = plan_class(plan, precision);
            bool found = false;
            {
                std::lock_guard<std::mutex> lk(tt.mu);
                auto it = tt.entries.find(cls);
                if (it != tt.entries.end()) {
                    s = it->second.s;
                    found = true;
                }
            }
            // no timing (it waits on the stream) while a graph is being captured on it
            if (!found && mode == 2 && ctx->no_tune) {
                // an allocation-free call takes the heuristic for an untuned class and leaves
                // the class to be tuned by the next ordinary call
                if (ctx->forced_config >= 0) s.config = ctx->forced_config;
                if (ctx->forced_splits >= 1) s.splits = ctx->forced_splits;
                return s;
            }
            if (!found && mode == 2 && !capturing_on(ctx, stream)) {
                const TuneEntry e = tune(ctx, plan, precision, A, B, stream);
                std::lock_guard<std::mutex> lk(tt.mu);
                // the first decision recorded for a class wins (a concurrent tuner may have
                // finished first): every context then uses the same schedule
                auto ins = tt.entries.emplace(cls, e);
                s = ins.first->second.s;
                if (ins.second) ++tt.tuned;
            }
        }
        plan.sched = s;
        plan.sched_gen = tt.gen.load();
        plan.sched_mode = mode;
        plan.sched_prec = precision;
    }
    if (ctx->forced_config >= 0) s.config = ctx->forced_config;
    if (ctx->forced_splits >= 1) s.splits = ctx->forced_splits;
    return s;
}

}  // namespace bug

using namespace bug;

extern "C" {

int bug_set_autotune(int32_t mode) {
    if (mode < 0 || mode > 2) return BUG_EINVAL;
    table().mode.store(mode);
    return BUG_OK;
}

int bug_set_gemm_schedule(bug_ctx* ctx, int32_t config, int32_t splits) {
    if (!ctx || config < -1 || config >= kNumTileConfigs || splits < 0 || splits > kMaxSplits)
        return BUG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->forced_config = config;
    ctx->forced_splits = splits;
    return BUG_OK;
}

int bug_autotune_count(int64_t* out2) {
    if (!out2) return BUG_EINVAL;
    TuneTable& tt = table();
    std::lock_guard<std::mutex> lk(tt.mu);
    out2[0] = (int64_t)tt.entries.size();
    out2[1] = tt.tuned.load();
    return BUG_OK;
}

int bug_autotune_clear(void) {
    TuneTable& tt = table();
    std::lock_guard<std::mutex> lk(tt.mu);
    tt.entries.clear();
    ++tt.gen;
    return BUG_OK;
}

int bug_autotune_save(const char* path) {
    if (!path) return BUG_EINVAL;
    TuneTable& tt = table();
    std::lock_guard<std::mutex> lk(tt.mu);
    FILE* f = std::fopen(path, "w");
    if (!f) return BUG_EINVAL;
    std::fprintf(f, "# grouped-GEMM schedules: plan class, tile configuration, split factor "
                    "# precision, mean m, n, k, tasks, flops, best ms, heuristic ms\n");
    for (const auto& e : tt.entries) {
        const TuneEntry& t = e.second;
        std::fprintf(f, "%016llx %d %d # %d %.0f %.0f %.0f %.0f %.4g %.4f %.4f\n",
                     (unsigned long long)e.first, t.s.config, t.s.splits, t.precision, t.m, t.n,
                     t.k, t.tasks, t.flops, t.best_ms, t.heuristic_ms);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? BUG_OK : BUG_EINVAL;
}

int bug_autotune_load(const char* path) {
    if (!path) return BUG_EINVAL;
    FILE* f = std::fopen(path, "r");
    if (!f) return BUG_EINVAL;
    std::map<uint64_t, TuneEntry> got;
    char line[256];
    int rc = BUG_OK;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        unsigned long long cls = 0;
        int c = -1, s = 0;
        if (std::sscanf(line, "%llx %d %d", &cls, &c, &s) != 3 || c < 0 || c >= kNumTileConfigs ||
            s < 1 || s > kMaxSplits) {
            rc = BUG_EINVAL;
            break;
        }
        TuneEntry e;
        e.s = Schedule{c, s};
        got[(uint64_t)cls] = e;
    }
    std::fclose(f);
    if (rc != BUG_OK) return rc;
    TuneTable& tt = table();
    std::lock_guard<std::mutex> lk(tt.mu);
    for (const auto& e : got) {
        auto it = tt.entries.find(e.first);
        if (it == tt.entries.end())
            tt.entries.emplace(e.first, e.second);
        else
            it->second.s = e.second.s;
    }
    ++tt.gen;
    return BUG_OK;
}

}  // extern "C"
