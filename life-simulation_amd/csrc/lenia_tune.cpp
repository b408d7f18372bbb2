// Empirical launch-shape tuning of one FFT-form world.
//
// The radix plan of a length is fixed (it decides the arithmetic, so it stays
// put and results stay bit-identical); only the launch shape of each pass is
// searched: transforms per workgroup (row pairs / columns), workgroup size and
// whether the K inverse transforms of a tile run side by side.  Each
// candidate is timed with per-kernel GPU timestamps over a few steps; the
// world's state and step clock are restored afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "launch.h"

using namespace lenia;

namespace {

struct Shape {
    int nt, T;
    bool kpar;
};

// the candidate shapes of one pass that the plan, registers and LDS allow
std::vector<Shape> candidates(const FftPlan &p, bool rows, int count, int K) {
    std::vector<Shape> out;
    for (int kp = 1; kp >= 0; --kp)
        for (int nt = 1; nt <= 8 && nt <= count; nt *= 2)
            for (int T = 64; T <= 1024; T *= 2)
                if (shape_ok(p, rows, nt, K, T, kp != 0)) out.push_back({nt, T, kp != 0});
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
    Shape bestRow{w->P, w->rowT, w->rowKpar}, bestCol{w->C, w->colT, w->colKpar};
    for (const Shape &s : candidates(w->planW, true, H / 2, K)) {
        w->P = s.nt;
        w->rowT = s.T;
        w->rowKpar = s.kpar;
        double ms[2];
        rc = time_shape(w, nSteps, ms);
        if (rc) break;
        if (ms[1] < best[1]) {
            best[1] = ms[1];
            bestRow = s;
        }
    }
    w->P = bestRow.nt;
    w->rowT = bestRow.T;
    w->rowKpar = bestRow.kpar;
    for (const Shape &s : candidates(w->planH, false, Wh, K)) {
        if (rc) break;
        w->C = s.nt;
        w->colT = s.T;
        w->colKpar = s.kpar;
        double ms[2];
        rc = time_shape(w, nSteps, ms);
        if (rc) break;
        if (ms[0] < best[0]) {
            best[0] = ms[0];
            bestCol = s;
        }
    }
    w->C = bestCol.nt;
    w->colT = bestCol.T;
    w->colKpar = bestCol.kpar;
    w->shapeVersion += 1;
    w->desc.rowPairsPerTile = w->P;
    w->desc.colsPerTile = w->C;
    w->desc.rowThreads = w->rowT;
    w->desc.colThreads = w->colT;
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
