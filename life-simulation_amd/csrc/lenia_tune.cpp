// Empirical tuning of the FFT form: launch shapes and radix plans.
//
// lenia_world_autotune searches the launch shape of each pass of one world:
// transforms per workgroup (row pairs / columns), workgroup size and whether
// the K inverse transforms of a tile run side by side.  Shapes only decide
// which threads run a butterfly, so results stay bit-identical; each
// candidate is timed with per-kernel GPU timestamps over a few steps and the
// world's state and step clock are restored afterwards.
//
// The radix plan of a transform length decides the arithmetic, so it is a
// property of the (axis, length) pair held in a process-wide plan cache:
// lenia_plan_autotune times every compiled plan of a world's row and column
// lengths and records the fastest; worlds created afterwards without an
// explicit plan take the cached one, so a world stepped alone and the same
// world in a batch share plan and bits.  The cache is saved to and loaded
// from a text file, one "rows|cols <length> <plan>" line per entry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "host_internal.h"
#include "launch.h"

using namespace lenia;

namespace {

struct Shape {
    int nt, T;
    bool kpar;
    bool wave;   // wave executor with T / 64 waves per workgroup (nt, kpar unused)
    bool split;  // kernel-split wave executor
};

// the candidate shapes of one pass that the plan, registers and LDS allow,
// and the wave executor at 1, 2 and 4 waves per workgroup when the plan has one
std::vector<Shape> candidates(const FftPlan &p, bool rows, int count, int K) {
    std::vector<Shape> out;
    for (int kp = 1; kp >= 0; --kp)
        for (int nt = 1; nt <= 8 && nt <= count; nt *= 2)
            for (int T = 64; T <= 1024; T *= 2)
                if (shape_ok(p, rows, nt, K, T, kp != 0)) out.push_back({nt, T, kp != 0, false, false});
    if (wave_lanes(p.code)) {
        for (int nw = 1; nw <= 4; nw *= 2) out.push_back({0, 64 * nw, false, true, false});
        out.push_back({0, 64 * kMaxKernels, false, true, true});
    }
    return out;
}

// mean [column pass, row pass] ms over nSteps after two untimed steps
int time_shape(lenia_world *w, uint32_t nSteps, double ms[2]) {
    int rc = world_upload_desc(w);
    if (rc) return rc;
    hipError_t e = enqueue_steps(w, 2);
    if (e != hipSuccess) return hip_fail(e, "autotune warm-up");
    e = time_steps(w->ctx, nSteps, 2, [&](const hipEvent_t *ev) {
        hipError_t es = enqueue_step(w, ev, 0);
        return es == hipSuccess ? world_advance_clock(w, 1) : es;
    }, ms);
    return e == hipSuccess ? LENIA_OK : hip_fail(e, "autotune timing");
}

}  // namespace

extern "C" int lenia_world_autotune(lenia_world *w, uint32_t nSteps, double *msPerStep) {
    if (!w) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_autotune: NULL world");
    if (!w->isFft) {
        if (msPerStep) *msPerStep = 0.0;
        return LENIA_OK;  // the direct forms have one launch shape
    }
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    nSteps = std::max(nSteps, 4u);
    if (w->needsInit) {
        int rc = world_apply_init(w);
        if (rc) return rc;
    }
    int rc = world_ensure_spectrum(w);
    if (rc) return rc;
    // keep the state and clock to restore them after the timed steps
    const size_t cells = (size_t)w->desc.width * w->desc.height;
    StepClock clock{};
    hipError_t e = hipMemcpyAsync(w->scratchOut, w->current_state(), cells * sizeof(float),
                                  hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&clock, w->clock, sizeof(clock), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "autotune save");
    world_drop_graph(w);
    const int K = (int)w->desc.nKernels;
    const int H = w->desc.height, Wh = w->desc.width / 2 + 1;
    double best[2] = {1e30, 1e30};
    const Shape keepRow{w->P, w->rowT, w->rowKpar, false, false}, keepCol{w->C, w->colT, w->colKpar, false, false};
    Shape bestRow = w->rowWave ? Shape{0, 64 * w->rowNW, false, true, w->rowSplit} : keepRow;
    Shape bestCol = w->colWave ? Shape{0, 64 * w->colNW, false, true, w->colSplit} : keepCol;
    // a wave candidate keeps the LDS shape fields as they are
    auto set_row = [&](const Shape &s) {
        w->rowWave = s.wave;
        w->rowSplit = s.split;
        if (s.wave) {
            w->rowNW = s.T / 64;
        } else {
            w->P = s.nt;
            w->rowT = s.T;
            w->rowKpar = s.kpar;
        }
    };
    auto set_col = [&](const Shape &s) {
        w->colWave = s.wave;
        w->colSplit = s.split;
        if (s.wave) {
            w->colNW = s.T / 64;
        } else {
            w->C = s.nt;
            w->colT = s.T;
            w->colKpar = s.kpar;
        }
    };
    for (const Shape &s : candidates(w->planW, true, H / 2, K)) {
        set_row(s);
        double ms[2];
        rc = time_shape(w, nSteps, ms);
        if (rc) break;
        if (ms[1] < best[1]) {
            best[1] = ms[1];
            bestRow = s;
        }
    }
    set_row(bestRow);
    for (const Shape &s : candidates(w->planH, false, Wh, K)) {
        if (rc) break;
        set_col(s);
        double ms[2];
        rc = time_shape(w, nSteps, ms);
        if (rc) break;
        if (ms[0] < best[0]) {
            best[0] = ms[0];
            bestCol = s;
        }
    }
    set_col(bestCol);
    w->shapeVersion += 1;
    w->desc.rowEngine = w->rowWave ? (w->rowSplit ? LENIA_ENGINE_WAVE_SPLIT : LENIA_ENGINE_WAVE) : LENIA_ENGINE_LDS;
    w->desc.colEngine = w->colWave ? (w->colSplit ? LENIA_ENGINE_WAVE_SPLIT : LENIA_ENGINE_WAVE) : LENIA_ENGINE_LDS;
    w->desc.rowPairsPerTile = w->rowWave ? 0 : w->P;
    w->desc.colsPerTile = w->colWave ? 0 : w->C;
    w->desc.rowThreads = w->rowWave ? 64 * w->rowNW : w->rowT;
    w->desc.colThreads = w->colWave ? 64 * w->colNW : w->colT;
    const int rc2 = world_upload_desc(w);
    // restore the state (its spectrum is rebuilt by the next step) and clock
    e = hipMemcpyAsync(w->current_state(), w->scratchOut, cells * sizeof(float),
                       hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_set_clock(w->clock, clock.frame, clock.sub, clock.substeps, clock.parity, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    w->specValid = false;
    if (rc) return rc;
    if (rc2) return rc2;
    if (e != hipSuccess) return hip_fail(e, "autotune restore");
    if (msPerStep) *msPerStep = best[0] + best[1];
    return LENIA_OK;
}

// ------------------------------------------------------------------ plan cache
namespace {

std::mutex g_planMu;
std::map<std::pair<uint32_t, uint32_t>, uint32_t> g_plans;  // (axis, n) -> packed plan

bool plan_valid(uint32_t n, uint32_t packed) {
    FftPlan p;
    return n >= 2 && n <= 4096 && packed != 0 && plan_unpack((int)n, packed, p);
}

}  // namespace

uint32_t lenia::plan_cache_lookup(uint32_t axis, uint32_t n) {
    std::lock_guard<std::mutex> lk(g_planMu);
    auto it = g_plans.find({axis, n});
    return it == g_plans.end() ? 0u : it->second;
}

extern "C" int lenia_plan_cache_set(uint32_t axis, uint32_t n, uint32_t plan) {
    if (axis > 1) return fail(LENIA_ERR_INVALID_ARG, "plan cache: axis must be 0 (rows) or 1 (columns)");
    if (plan != 0 && !plan_valid(n, plan))
        return fail(LENIA_ERR_INVALID_ARG, "plan cache: 0x%x is not a radix plan of length %u", plan, n);
    std::lock_guard<std::mutex> lk(g_planMu);
    if (plan == 0)
        g_plans.erase({axis, n});
    else
        g_plans[{axis, n}] = plan;
    return LENIA_OK;
}

extern "C" int lenia_plan_cache_get(uint32_t axis, uint32_t n, uint32_t *plan) {
    if (!plan) return fail(LENIA_ERR_INVALID_ARG, "lenia_plan_cache_get: NULL argument");
    if (axis > 1) return fail(LENIA_ERR_INVALID_ARG, "plan cache: axis must be 0 (rows) or 1 (columns)");
    *plan = plan_cache_lookup(axis, n);
    return LENIA_OK;
}

extern "C" void lenia_plan_cache_clear(void) {
    std::lock_guard<std::mutex> lk(g_planMu);
    g_plans.clear();
}

extern "C" int lenia_plan_cache_save(const char *path) {
    if (!path) return fail(LENIA_ERR_INVALID_ARG, "lenia_plan_cache_save: NULL path");
    std::vector<std::pair<std::pair<uint32_t, uint32_t>, uint32_t>> rows;
    {
        std::lock_guard<std::mutex> lk(g_planMu);
        rows.assign(g_plans.begin(), g_plans.end());
    }
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(LENIA_ERR_INVALID_ARG, "plan cache: cannot write %s (%s)", path, std::strerror(errno));
    std::fprintf(f, "# lenia-mi355x radix plans: axis length plan (4 bits per stage, first stage lowest)\n");
    for (const auto &r : rows)
        std::fprintf(f, "%s %u 0x%x\n", r.first.first ? "cols" : "rows", r.first.second, r.second);
    const bool ok = std::fclose(f) == 0;
    return ok ? LENIA_OK : fail(LENIA_ERR_INVALID_ARG, "plan cache: write error on %s", path);
}

extern "C" int lenia_plan_cache_load(const char *path) {
    if (!path) return fail(LENIA_ERR_INVALID_ARG, "lenia_plan_cache_load: NULL path");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(LENIA_ERR_INVALID_ARG, "plan cache: cannot read %s (%s)", path, std::strerror(errno));
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> got;
    char line[256];
    int lineNo = 0, rc = LENIA_OK;
    while (rc == LENIA_OK && std::fgets(line, sizeof line, f)) {
        ++lineNo;
        char axis[16];
        unsigned n = 0, plan = 0;
        char extra[2];
        const char *t = line;
        while (*t == ' ' || *t == '\t') ++t;
        if (*t == '#' || *t == '\n' || *t == '\r' || *t == 0) continue;
        if (std::sscanf(t, "%15s %u %x %1s", axis, &n, &plan, extra) != 3 ||
            (std::strcmp(axis, "rows") != 0 && std::strcmp(axis, "cols") != 0) || !plan_valid(n, plan))
            rc = fail(LENIA_ERR_INVALID_ARG, "plan cache: %s line %d is not \"rows|cols <length> <plan>\" "
                      "with a radix plan of that length", path, lineNo);
        else
            got[{std::strcmp(axis, "cols") == 0 ? 1u : 0u, n}] = plan;
    }
    std::fclose(f);
    if (rc) return rc;  // all or nothing
    std::lock_guard<std::mutex> lk(g_planMu);
    for (const auto &e : got) g_plans[e.first] = e.second;
    return LENIA_OK;
}

// ------------------------------------------------------------------ plan tuning
namespace {

// kernel ms per step of desc with the given plans and tuned launch shapes;
// false when no shape serves the plans
bool time_plans(lenia_ctx *ctx, lenia_world_desc d, uint32_t planW, uint32_t planH, uint32_t nSteps,
                double &ms) {
    d.fftPlanW = planW;
    d.fftPlanH = planH;
    d.rowPairsPerTile = d.colsPerTile = d.rowThreads = d.colThreads = 0;
    d.rowEngine = d.colEngine = LENIA_ENGINE_AUTO;
    lenia_world *w = nullptr;
    if (lenia_world_create(ctx, &d, &w) != LENIA_OK) return false;
    int rc = lenia_world_reset(w, 1u);
    if (!rc) rc = lenia_world_frame(w, 0, 1);
    if (!rc) rc = lenia_world_autotune(w, nSteps, &ms);
    lenia_world_destroy(w);
    return rc == LENIA_OK;
}

// the compiled plans of length n, or the built-in default when there is none
std::vector<uint32_t> plan_candidates(int n) {
    uint32_t c[16];
    const int m = compiled_plans(n, c, 16);
    std::vector<uint32_t> out(c, c + m);
    if (out.empty()) {
        FftPlan p;
        if (plan_default(n, 16, p)) out.push_back(plan_pack(p));
    }
    return out;
}

}  // namespace

extern "C" int lenia_plan_autotune(lenia_ctx *ctx, const lenia_world_desc *desc, uint32_t nSteps,
                                   uint32_t *planW, uint32_t *planH, double *msPerStep) {
    if (!ctx || !desc) return fail(LENIA_ERR_INVALID_ARG, "lenia_plan_autotune: NULL argument");
    if (desc->convMode != LENIA_CONV_FFT)
        return fail(LENIA_ERR_UNSUPPORTED, "lenia_plan_autotune: FFT form only (convMode %u)", desc->convMode);
    const int W = (int)desc->width, H = (int)desc->height;
    nSteps = std::max(nSteps, 4u);
    // rows first, with the column plan the cache or the default gives
    const uint32_t colStart = plan_cache_lookup(1u, (uint32_t)H);
    double bestMs = 1e30;
    uint32_t bestW = 0, bestH = colStart;
    for (uint32_t pw : plan_candidates(W)) {
        double ms;
        if (time_plans(ctx, *desc, pw, colStart, nSteps, ms) && ms < bestMs) {
            bestMs = ms;
            bestW = pw;
        }
    }
    if (!bestW) return fail(LENIA_ERR_UNSUPPORTED, "lenia_plan_autotune: no radix plan of width %d runs", W);
    bestMs = 1e30;
    // column plans with a wave kernel, and their times
    auto has_wave = [&](uint32_t packed) {
        FftPlan p;
        return plan_unpack(H, packed, p) && wave_lanes(p.code) > 0;
    };
    double bestWaveMs = 1e30;
    uint32_t bestWaveH = 0;
    for (uint32_t ph : plan_candidates(H)) {
        double ms;
        if (!time_plans(ctx, *desc, bestW, ph, nSteps, ms)) continue;
        if (ms < bestMs) {
            bestMs = ms;
            bestH = ph;
        }
        if (has_wave(ph) && ms < bestWaveMs) {
            bestWaveMs = ms;
            bestWaveH = ph;
        }
    }
    if (!bestH) return fail(LENIA_ERR_UNSUPPORTED, "lenia_plan_autotune: no radix plan of height %d runs", H);
    // A column plan with a wave kernel within kWaveColsSlack of the fastest is
    // taken: in a batch the worlds whose column plan has none form an LDS
    // launch of their own beside the wave class launches, which can set the
    // column pass's length (the 256-world sweep averaged ~153 us per step
    // with 384 = 8 4 4 3 against ~159 us with 16 8 3, the plan that was
    // faster for a 384-high world alone).
    constexpr double kWaveColsSlack = 1.10;
    if (bestWaveH && !has_wave(bestH) && bestWaveMs <= bestMs * kWaveColsSlack) {
        bestH = bestWaveH;
        bestMs = bestWaveMs;
    }
    int rc = lenia_plan_cache_set(0u, (uint32_t)W, bestW);
    if (!rc) rc = lenia_plan_cache_set(1u, (uint32_t)H, bestH);
    if (rc) return rc;
    if (planW) *planW = bestW;
    if (planH) *planH = bestH;
    if (msPerStep) *msPerStep = bestMs;
    return LENIA_OK;
}
