// Settings of the passes per size bucket, and the autotuner that picks them
// by timing the real pass on device copies.  No counterpart in the reference,
// whose launch shapes and table sizes are compile-time constants.
#include "gls_internal.h"

#include <algorithm>
#include <chrono>
#include <vector>

int tuning_bucket(int64_t objects) {
    int b = 0;
    while (b < 63 && (objects >> (b + 1)) > 0) ++b;
    return b;
}

namespace {

bool tuning_valid(const gls_tuning &t) {
    return t.cut_wide >= 1 && t.rw_sweep_width >= 1 && (t.strash_table_factor == 4 || t.strash_table_factor == 8) &&
           t.strash_wave_batch >= 1 && t.strash_wave_batch <= 64 && t.balance_narrow_max >= 1 &&
           t.balance_narrow_max <= (1 << 24) && t.cut_levels >= 1 && t.cut_levels <= 6;
}

// one timed run of the command over fresh device copies of the inputs
int run_once(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t command, double &ms) {
    std::vector<gls_aig *> c(count, nullptr);
    int rc = GLS_OK;
    for (int32_t k = 0; k < count && rc == GLS_OK; k++) rc = gls_aig_copy(aigs[k], &c[k]);
    if (rc == GLS_OK) rc = hipStreamSynchronize(ctx->stream) == hipSuccess ? GLS_OK : GLS_ERR_HIP;
    const auto t0 = std::chrono::steady_clock::now();
    if (rc == GLS_OK) {
        if (count == 1) {
            rc = command == GLS_TUNE_STRASH    ? gls_strash(c[0], 0)
                 : command == GLS_TUNE_BALANCE ? gls_balance(c[0], 1, 1)
                                               : gls_rewrite(c[0], command == GLS_TUNE_RWZ ? 1 : 0, 0);
        } else {
            rc = command == GLS_TUNE_STRASH    ? gls_batch_strash(ctx, c.data(), count, 0)
                 : command == GLS_TUNE_BALANCE ? gls_batch_balance(ctx, c.data(), count, 1, 1)
                                               : gls_batch_rewrite(ctx, c.data(), count, command == GLS_TUNE_RWZ, 0);
        }
    }
    if (rc == GLS_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = GLS_ERR_HIP;
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (gls_aig *x : c)
        if (x) gls_aig_free(x);
    return rc;
}

} // namespace

extern "C" int gls_ctx_get_tuning(gls_ctx *ctx, int64_t objects, gls_tuning *out) {
    if (!ctx || !out) return GLS_ERR_ARG;
    *out = ctx->tune[tuning_bucket(objects < 1 ? 1 : objects)];
    return GLS_OK;
}

extern "C" int gls_ctx_set_tuning(gls_ctx *ctx, int64_t objects, const gls_tuning *t) {
    if (!ctx || !t) return GLS_ERR_ARG;
    if (!tuning_valid(*t)) return gls_fail(ctx, GLS_ERR_ARG, "gls_ctx_set_tuning: a setting is out of range");
    gls_tuning v = *t;
    if (objects < 0) {
        for (auto &x : ctx->tune) x = v;
    } else {
        ctx->tune[tuning_bucket(objects < 1 ? 1 : objects)] = v;
    }
    return GLS_OK;
}

extern "C" int gls_autotune(gls_ctx *ctx, gls_aig **aigs, int32_t count, int32_t command, int32_t reps,
                            gls_tuning *out, double *best_ms, double *start_ms) {
    if (!ctx || !aigs || count < 1 || reps < 1 || command < GLS_TUNE_RW || command > GLS_TUNE_BALANCE)
        return gls_fail(ctx, GLS_ERR_ARG, "gls_autotune: bad arguments");
    if (ctx->device < 0) return gls_fail(ctx, GLS_ERR_NODEV, "host-only context");
    if (ctx->active) return gls_fail(ctx, GLS_ERR_ARG, "gls_autotune: called inside a pass");
    int64_t objects = 1;
    for (int32_t k = 0; k < count; k++) {
        if (!aigs[k] || aigs[k]->ctx != ctx) return gls_fail(ctx, GLS_ERR_ARG, "gls_autotune: AIG %d", k);
        objects += aigs[k]->nObjs;
    }
    if (count == 1) objects = aigs[0]->nObjs;
    gls_tuning &slot = ctx->tune[tuning_bucket(objects < 1 ? 1 : objects)];
    const gls_tuning start = slot;
    // the fastest of reps runs under setting t
    auto measure = [&](const gls_tuning &t, double &ms) -> int {
        slot = t;
        ms = 1e300;
        for (int32_t r = 0; r < reps; r++) {
            double one = 0;
            const int rc = run_once(ctx, aigs, count, command, one);
            if (rc != GLS_OK) return rc;
            ms = std::min(ms, one);
        }
        return GLS_OK;
    };
    // the candidate values of each setting the command uses
    struct Knob {
        int32_t gls_tuning::*field;
        std::vector<int32_t> values;
    };
    std::vector<Knob> knobs;
    if (command == GLS_TUNE_RW || command == GLS_TUNE_RWZ) {
        knobs.push_back({&gls_tuning::cut_wide, {512, 1024, 2048, 4096, 16384, 1 << 30}});
        knobs.push_back({&gls_tuning::rw_sweep_width, {256, 1024, 4096, 1 << 30}});
    }
    if (command == GLS_TUNE_BALANCE) knobs.push_back({&gls_tuning::balance_narrow_max, {1024, 4096, 16384, 65536}});
    knobs.push_back({&gls_tuning::strash_table_factor, {4, 8}});
    knobs.push_back({&gls_tuning::strash_wave_batch, {1, 2, 4, 8, 16}});
    gls_tuning best = start;
    double bestMs = 0, startMs = 0;
    int rc = measure(best, bestMs);
    startMs = bestMs;
    for (size_t kn = 0; rc == GLS_OK && kn < knobs.size(); kn++) {
        for (int32_t v : knobs[kn].values) {
            if (best.*(knobs[kn].field) == v) continue;
            gls_tuning t = best;
            t.*(knobs[kn].field) = v;
            double ms = 0;
            rc = measure(t, ms);
            if (rc != GLS_OK) break;
            // a change must win by more than 1% (run-to-run noise)
            if (ms < bestMs * 0.99) {
                best = t;
                bestMs = ms;
            }
        }
    }
    if (rc != GLS_OK) {
        slot = start;
        return rc;
    }
    slot = best;
    if (out) *out = best;
    if (best_ms) *best_ms = bestMs;
    if (start_ms) *start_ms = startMs;
    return GLS_OK;
}
