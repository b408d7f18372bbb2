// Differentiable rollout: T FFT-form steps from a given state, and the
// hand-written reverse pass with sqrt(T)-style checkpointing.  No counterpart
// in the reference (its Lenia module only steps forward); the forward steps
// are exactly the library's own, so the reverse pass differentiates the
// arithmetic that produced the states.
//
// Forward: the states at t = 0, c, 2c, ... are kept as checkpoints while the
// rollout runs.  Reverse: segment by segment from the last, the segment's
// steps are recomputed bit-exactly from its checkpoint with the normal step
// kernels, which (through the save pointers of a copy of the world record)
// also leave each step's input state, potentials and 2D spectrum; the steps
// are then reversed from the segment's end to its start with two launches
// each (k_adjoint.hip).  Both sequences are captured once per (T, c) into
// hipGraphs and replayed; every kernel reads the growth parameters from the
// world's device record, so parameter updates between calls keep the graphs
// valid.  A reverse call can reuse the checkpoints of the last forward call
// (same T, c and parameters) instead of running the forward sweep again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "launch.h"

using namespace lenia;

struct AdjWorkspace {
    size_t n = 0;     // cells
    size_t spec = 0;  // floats of one complex half spectrum
    int K = 0, T = 0, c = 0, nShells = 0;
    float *ckpt = nullptr;  // [ceil(T/c)][n] states at t = 0, c, 2c, ...
    // three sets of c slots: segment sg < last recomputes into set sg % 2
    // while the reverse steps of segment sg + 1 read the other set; set 2
    // holds the last segment, saved by the forward sweep itself
    float *saveA = nullptr, *saveU = nullptr, *saveAhat = nullptr;  // [3c] slots
    FftWorldDev *recs = nullptr;  // [3c] world records with save pointers
    hipStream_t side = nullptr;   // the recompute branch of the captured graph
    hipEvent_t fork = nullptr, join = nullptr;
    float *FG = nullptr, *G1 = nullptr, *gpb = nullptr, *g = nullptr, *D = nullptr;
    double *accRows = nullptr, *accCols = nullptr, *acc = nullptr;
    int rowT = 256, colT = 256;
    bool dValid = false;  // D holds the shell spectra of the current kernels
    AdjArgs args{};
    hipGraphExec_t fwdGraph = nullptr, bwdGraph = nullptr;
    uint32_t graphShapes = 0;  // the world's shapeVersion the graphs were recorded with
    bool ckptValid = false;    // ckpt holds a forward sweep ...
    uint32_t ckptVersion = 0;  // ... run with this world descriptor version
};

static void ws_free(lenia_ctx *ctx, AdjWorkspace *ws) {
    if (ws->fwdGraph) (void)hipGraphExecDestroy(ws->fwdGraph);
    if (ws->bwdGraph) (void)hipGraphExecDestroy(ws->bwdGraph);
    float *ptrs[] = {ws->ckpt, ws->saveA, ws->saveU, ws->saveAhat, ws->FG, ws->G1, ws->gpb, ws->g, ws->D};
    for (float *p : ptrs) dev_free(ctx, p);
    dev_free(ctx, ws->recs);
    if (ws->fork) (void)hipEventDestroy(ws->fork);
    if (ws->join) (void)hipEventDestroy(ws->join);
    if (ws->side) (void)hipStreamDestroy(ws->side);
    dev_free(ctx, ws->accRows);
    dev_free(ctx, ws->accCols);
    dev_free(ctx, ws->acc);
    *ws = AdjWorkspace();
}

void lenia::world_free_adjoint(lenia_world *w) {
    if (!w->adj) return;
    ws_free(w->ctx, w->adj);
    delete w->adj;
    w->adj = nullptr;
}

// the kernel shapes changed (set_params): the shell spectra must be rebuilt;
// a different shell count needs a new workspace
void lenia::world_invalidate_adjoint(lenia_world *w) {
    if (!w->adj) return;
    int ns = 0;
    for (uint32_t k = 0; k < w->desc.nKernels; ++k)
        if (w->desc.kernels[k].shape == LENIA_SHAPE_RING) ns += (int)w->desc.kernels[k].nBetas;
    if (ns != w->adj->nShells) {
        world_free_adjoint(w);
        return;
    }
    w->adj->dValid = false;
}

static int check_differentiable(lenia_world *w) {
    if (!w->isFft)
        return fail(LENIA_ERR_UNSUPPORTED, "rollout gradients need the FFT form (convMode 1)");
    if (w->desc.growthMode != 0)
        return fail(LENIA_ERR_UNSUPPORTED, "rollout gradients need growthMode 0");
    if (w->extras())
        return fail(LENIA_ERR_UNSUPPORTED,
                    "rollout gradients need noiseAmount, muJitter and injectAmount at 0");
    return LENIA_OK;
}

// raw window of shell j of kernel k (beta_j = 1, other shells 0), double
static void shell_window(const lenia_kernel_desc &k, int j, int h, std::vector<double> &out) {
    lenia_kernel_desc one = k;
    for (int q = 0; q < 4; ++q) one.betas[q] = q == j ? 1.0f : 0.0f;
    const int n = 2 * h + 1;
    out.assign((size_t)n * n, 0.0);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) out[(size_t)y * n + x] = kernel_raw(one, x - h, y - h);
}

// shell-gradient spectra D_kj = spectrum of (c_kj - C_kj K_k) / Sraw_k, where
// c_kj is shell j's raw window, C_kj its sum and Sraw_k the raw kernel's sum
static int build_shell_spectra(lenia_world *w, AdjWorkspace *ws) {
    lenia_ctx *ctx = w->ctx;
    const int W = w->desc.width, H = w->desc.height, Wh = W / 2 + 1, K = (int)w->desc.nKernels;
    const size_t specF = (size_t)Wh * H;
    std::vector<std::vector<double>> wins;
    int hmax = 0;
    for (int k = 0; k < K; ++k) {
        const lenia_kernel_desc &kd = w->desc.kernels[k];
        hmax = std::max(hmax, kernel_half(kd));
        if (kd.shape != LENIA_SHAPE_RING) continue;
        const int h = kernel_half(kd), m = 2 * h + 1;
        std::vector<double> Kn;
        kernel_window(kd, h, Kn);
        double Sraw = 0.0;
        for (int y = 0; y < m; ++y)
            for (int x = 0; x < m; ++x) Sraw += kernel_raw(kd, x - h, y - h);
        for (int j = 0; j < (int)kd.nBetas; ++j) {
            std::vector<double> cj;
            shell_window(kd, j, h, cj);
            double Cj = 0.0;
            for (double v : cj) Cj += v;
            std::vector<double> dwin(m * m);
            for (size_t q = 0; q < dwin.size(); ++q) dwin[q] = (cj[q] - Cj * Kn[q]) / Sraw;
            wins.push_back(dwin);
        }
    }
    if (wins.empty()) {
        ws->dValid = true;
        return LENIA_OK;
    }
    const int mmax = 2 * hmax + 1;
    double *dWin = nullptr, *dCw = nullptr;
    int rc = dev_alloc(ctx, (void **)&dWin, sizeof(double) * mmax * mmax);
    if (!rc) rc = dev_alloc(ctx, (void **)&dCw, sizeof(double) * mmax * Wh);
    for (size_t i = 0; i < wins.size() && !rc; ++i) {
        const int h = kernel_half(w->desc.kernels[ws->args.shellK[i]]);
        hipError_t e = hipMemcpyAsync(dWin, wins[i].data(), sizeof(double) * wins[i].size(),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = launch_kernel_spectrum(dWin, h, W, H, dCw, ws->D + specF * i, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "shell spectra");
    }
    dev_free(ctx, dWin);
    dev_free(ctx, dCw);
    if (!rc) ws->dValid = true;
    return rc;
}

static int ws_prepare(lenia_world *w, uint32_t T, uint32_t c) {
    lenia_ctx *ctx = w->ctx;
    const int W = w->desc.width, H = w->desc.height, Wh = W / 2 + 1, K = (int)w->desc.nKernels;
    const size_t n = (size_t)W * H, spec = (size_t)Wh * H * 2;  // floats of one complex spectrum
    AdjWorkspace *ws = w->adj;
    if (ws && (ws->T != (int)T || ws->c != (int)c || ws->K != K)) {
        world_free_adjoint(w);
        ws = nullptr;
    }
    if (!ws) {
        ws = new AdjWorkspace();
        w->adj = ws;
        ws->n = n;
        ws->spec = spec;
        ws->K = K;
        ws->T = (int)T;
        ws->c = (int)c;
        const int nck = (int)((T + c - 1) / c);
        AdjArgs &a = ws->args;
        std::memset(&a, 0, sizeof(a));
        int ns = 0;
        for (int k = 0; k < K; ++k) {
            const lenia_kernel_desc &kd = w->desc.kernels[k];
            if (kd.shape != LENIA_SHAPE_RING) continue;
            for (int j = 0; j < (int)kd.nBetas; ++j) a.shellK[ns++] = k;
        }
        a.nShells = ns;
        ws->nShells = ns;
        a.W = W;
        a.H = H;
        a.K = K;
        a.planW = w->planW;
        a.planH = w->planH;
        a.gDev = &((const FftWorldDev *)w->devDesc)->g;
        a.twW = w->twW;
        a.twH = w->twH;
        a.S = w->S;
        int rc = dev_alloc(ctx, (void **)&ws->ckpt, sizeof(float) * n * nck);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->saveA, sizeof(float) * n * 3 * c);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->saveU, sizeof(float) * n * K * 3 * c);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->saveAhat, sizeof(float) * spec * 3 * c);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->recs, sizeof(FftWorldDev) * 3 * c);
        if (!rc) {
            hipError_t e = hipStreamCreateWithFlags(&ws->side, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ws->fork, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ws->join, hipEventDisableTiming);
            if (e != hipSuccess) rc = hip_fail(e, "rollout streams");
        }
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->FG, sizeof(float) * spec * K);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->G1, sizeof(float) * spec);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->gpb, sizeof(float) * n);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->g, sizeof(float) * n);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->accRows, sizeof(double) * 16 * (H / 2));
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->accCols, sizeof(double) * 16 * Wh);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->acc, sizeof(double) * 32);
        if (!rc && ns) rc = dev_alloc(ctx, (void **)&ws->D, sizeof(float) * (spec / 2) * ns);
        if (rc) {
            world_free_adjoint(w);
            return rc;
        }
        a.D = ws->D;
        // workgroup sizes: about 1 value per thread for the K transforms of
        // a row pair or column -- short stages over many threads, since one
        // reverse step has only H/2 + W/2+1 workgroups (256x256, K = 2: 512
        // threads, reverse pass 1.46 -> 1.39 ms against 256; 1024 gave the
        // same, 128 was slower) -- within <= 16 values per thread and the
        // workgroup's LDS
        const int mv = fft_max_values();
        auto pick = [&](const FftPlan &p, int nt, bool rows) {
            int want = 64;
            while (want < 512 && want < nt * p.n) want *= 2;
            for (int T2 = want; T2 <= 512; T2 *= 2)
                if (plan_fits(p, nt, T2, mv) && plan_fits(p, 1, T2, mv) &&
                    (rows ? adj_lds_rows(a, T2) : adj_lds_cols(a, T2)) <= kFftLdsMax)
                    return T2;
            for (int T2 = want / 2; T2 >= 64; T2 /= 2)
                if (plan_fits(p, nt, T2, mv) && plan_fits(p, 1, T2, mv) &&
                    (rows ? adj_lds_rows(a, T2) : adj_lds_cols(a, T2)) <= kFftLdsMax)
                    return T2;
            return 0;
        };
        ws->rowT = pick(w->planW, K, true);
        ws->colT = pick(w->planH, K, false);

        if (!ws->rowT || !ws->colT) {
            world_free_adjoint(w);
            return fail(LENIA_ERR_UNSUPPORTED, "rollout gradients: field %dx%d with %d kernels does "
                        "not fit the adjoint workgroups", W, H, K);
        }
    }
    if (!ws->dValid) {
        int rc = build_shell_spectra(w, ws);
        if (rc) {
            world_free_adjoint(w);
            return rc;
        }
    }
    return LENIA_OK;
}

static hipMemcpyKind kind_in(int onDevice) { return onDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice; }
static hipMemcpyKind kind_out(int onDevice) { return onDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost; }

static uint32_t ckpt_spacing(uint32_t T, uint32_t ckptEvery) {
    return ckptEvery ? std::min(ckptEvery, T) : (uint32_t)std::max(1.0, std::floor(std::sqrt((double)T)));
}

// state := src (device) and its row spectrum, on stream s
static hipError_t load_state(lenia_world *w, const float *src, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(w->current_state(), src, sizeof(float) * w->adj->n,
                                  hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = world_launch_rows(w, 0, (const FftWorldDev *)w->devDesc, 0, 0, nullptr, false, s);
    return e;
}

// records with save pointers for the 3c slots (refreshed from the world
// record, so parameter updates are seen)
static hipError_t enqueue_records(lenia_world *w, hipStream_t s) {
    AdjWorkspace *ws = w->adj;
    return launch_adj_records((const FftWorldDev *)w->devDesc, ws->recs, 3 * ws->c, ws->saveA, ws->saveU,
                              ws->saveAhat, ws->n, ws->spec, s);
}

// forward sweep from ckpt[0]: checkpoints at t = c, 2c, ...; the steps of the
// last segment also leave their saves (slot set 2), which the reverse pass
// reads first instead of recomputing them; the world ends at A_T with its
// clock T steps on
static hipError_t enqueue_forward(lenia_world *w) {
    AdjWorkspace *ws = w->adj;
    hipStream_t s = w->ctx->stream;
    const size_t n = ws->n;
    const uint32_t T = ws->T, c = ws->c, nck = (T + c - 1) / c;
    hipError_t e = enqueue_records(w, s);
    if (e == hipSuccess) e = load_state(w, ws->ckpt, s);
    uint32_t t = 0;
    for (uint32_t k = 1; k < nck && e == hipSuccess; ++k) {
        for (uint32_t i = 0; i < c && e == hipSuccess; ++i) e = enqueue_step(w, nullptr, t++);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ws->ckpt + n * k, w->current_state(), sizeof(float) * n,
                               hipMemcpyDeviceToDevice, s);
    }
    const FftWorldDev *last = ws->recs + (size_t)2 * c;
    for (uint32_t j = 0; t < T && e == hipSuccess; ++j, ++t) {
        e = world_launch_cols(w, last + j, s);
        if (e == hipSuccess) e = world_launch_rows(w, 1, last + j, t, 0, nullptr, false, s);
    }
    if (e == hipSuccess) e = world_advance_clock(w, T);
    return e;
}

// recompute segment sg (steps t0 .. t0+m-1) from its checkpoint on stream s;
// step t0 + j leaves A, u_k and Ahat in slot j of set sg % 2
static hipError_t recompute_segment(lenia_world *w, int sg, hipStream_t s) {
    AdjWorkspace *ws = w->adj;
    const uint32_t T = ws->T, c = ws->c, t0 = sg * c, m = std::min(T, t0 + c) - t0;
    const FftWorldDev *recs = ws->recs + (size_t)(sg % 2) * c;
    hipError_t e = load_state(w, ws->ckpt + ws->n * sg, s);
    for (uint32_t j = 0; j < m && e == hipSuccess; ++j) {
        e = world_launch_cols(w, recs + j, s);
        if (e == hipSuccess) e = world_launch_rows(w, 1, recs + j, 0, 0, nullptr, false, s);
    }
    return e;
}

// reverse the steps of segment sg from its saves, last step first.  Step t's
// row launch also finishes step t+1 (its edge: gp + row inverse of G1), so a
// reverse step is two launches; the edge of step 0 runs on its own after the
// last segment.
static hipError_t reverse_segment(lenia_world *w, int sg, hipStream_t s) {
    AdjWorkspace *ws = w->adj;
    const AdjArgs &a = ws->args;
    const size_t n = ws->n;
    const uint32_t T = ws->T, c = ws->c, t0 = sg * c, m = std::min(T, t0 + c) - t0;
    const bool last = t0 + m == T;  // saved by the forward sweep
    const size_t base = (size_t)(last ? 2 : sg % 2) * c;
    hipError_t e = hipSuccess;
    for (int j = (int)m - 1; j >= 0 && e == hipSuccess; --j) {
        const size_t slot = base + j;
        const bool first = t0 + j == T - 1;  // its incoming gradient is dL/dA_T in g
        e = launch_adj_mid(a, ws->rowT, ws->saveU + slot * a.K * n, ws->saveA + slot * n, ws->g,
                           first ? nullptr : ws->G1, ws->gpb, ws->FG, ws->accRows, s);
        if (e == hipSuccess)
            e = launch_adj_cols(a, ws->colT, ws->FG, ws->saveAhat + slot * ws->spec, ws->G1,
                                ws->accCols, s);
    }
    if (e == hipSuccess && sg == 0) e = launch_adj_edge(a, ws->rowT, ws->G1, ws->gpb, ws->g, s);
    return e;
}

// the reverse pass from the checkpoints, the last segment's saves (left by
// the forward sweep) and g = dL/dA_T: leaves the parameter sums in acc and
// dL/dA0 in g, the world at A0.  The recompute of segment sg - 1 (side
// stream) runs beside the reverse steps of segment sg.
static hipError_t enqueue_backward(lenia_world *w) {
    AdjWorkspace *ws = w->adj;
    const AdjArgs &a = ws->args;
    hipStream_t s = w->ctx->stream;
    const uint32_t T = ws->T, c = ws->c, nck = (T + c - 1) / c;
    const int H2 = a.H / 2, Wh = a.W / 2 + 1;
    // records with save pointers and zeroed FP64 slots (fill kernels: graph-safe)
    hipError_t e = enqueue_records(w, s);
    if (e == hipSuccess) e = launch_fill(reinterpret_cast<float *>(ws->accRows), (size_t)32 * H2, 0.0f, s);
    if (e == hipSuccess) e = launch_fill(reinterpret_cast<float *>(ws->accCols), (size_t)32 * Wh, 0.0f, s);
    for (int sg = (int)nck - 1; sg >= 0 && e == hipSuccess; --sg) {
        if (sg > 0) {
            e = hipEventRecord(ws->fork, s);
            if (e == hipSuccess) e = hipStreamWaitEvent(ws->side, ws->fork, 0);
            if (e == hipSuccess) e = recompute_segment(w, sg - 1, ws->side);
            if (e == hipSuccess) e = hipEventRecord(ws->join, ws->side);
        }
        if (e == hipSuccess) e = reverse_segment(w, sg, s);
        if (e == hipSuccess && sg > 0) e = hipStreamWaitEvent(s, ws->join, 0);
    }
    if (e == hipSuccess)
        e = launch_adj_final(ws->accRows, H2, 1 + 3 * a.K, ws->accCols, Wh, a.nShells, ws->acc, s);
    if (e == hipSuccess) e = load_state(w, ws->ckpt, s);
    return e;
}

static int capture(lenia_world *w, hipError_t (*body)(lenia_world *), hipGraphExec_t *out) {
    hipStream_t s = w->ctx->stream;
    if (*out) (void)hipGraphExecDestroy(*out);
    *out = nullptr;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "rollout graph capture");
    hipError_t eb = body(w);
    e = hipStreamEndCapture(s, &graph);
    if (eb != hipSuccess) e = eb;
    if (e == hipSuccess) e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        *out = nullptr;
        return hip_fail(e, "rollout graph");
    }
    return LENIA_OK;
}

// both graphs recorded with the world's current launch shapes
static int ensure_graphs(lenia_world *w) {
    AdjWorkspace *ws = w->adj;
    if (ws->fwdGraph && ws->bwdGraph && ws->graphShapes == w->shapeVersion) return LENIA_OK;
    int rc = capture(w, enqueue_forward, &ws->fwdGraph);
    if (!rc) rc = capture(w, enqueue_backward, &ws->bwdGraph);
    if (!rc) ws->graphShapes = w->shapeVersion;
    return rc;
}

extern "C" int lenia_rollout_forward(lenia_world *w, uint32_t T, uint32_t ckptEvery, const float *A0,
                                     float *AT, int onDevice) {
    if (!w || !A0 || !AT) return fail(LENIA_ERR_INVALID_ARG, "lenia_rollout_forward: NULL argument");
    if (T == 0) return fail(LENIA_ERR_INVALID_ARG, "lenia_rollout_forward: T must be at least 1");
    int rc = check_differentiable(w);
    if (rc) return rc;
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    rc = ws_prepare(w, T, ckpt_spacing(T, ckptEvery));
    if (!rc) rc = ensure_graphs(w);
    if (rc) return rc;
    AdjWorkspace *ws = w->adj;
    const size_t bytes = sizeof(float) * ws->n;
    ws->ckptValid = false;
    hipError_t e = hipMemcpyAsync(ws->ckpt, A0, bytes, kind_in(onDevice), ctx->stream);
    if (e == hipSuccess) e = hipGraphLaunch(ws->fwdGraph, ctx->stream);
    w->hostSteps += T;
    w->needsInit = false;
    w->specValid = true;
    if (e == hipSuccess) e = hipMemcpyAsync(AT, w->current_state(), bytes, kind_out(onDevice), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "rollout forward");
    ws->ckptValid = true;
    ws->ckptVersion = w->version;
    return LENIA_OK;
}

extern "C" int lenia_rollout_backward(lenia_world *w, uint32_t T, uint32_t ckptEvery,
                                      const float *A0, const float *gAT, float *gA0,
                                      lenia_param_grads *grads, int onDevice) {
    if (!w || !gAT) return fail(LENIA_ERR_INVALID_ARG, "lenia_rollout_backward: NULL argument");
    if (T == 0) return fail(LENIA_ERR_INVALID_ARG, "lenia_rollout_backward: T must be at least 1");
    int rc = check_differentiable(w);
    if (rc) return rc;
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    const uint32_t c = ckpt_spacing(T, ckptEvery);
    if (!A0) {
        const AdjWorkspace *ws0 = w->adj;
        if (!ws0 || !ws0->ckptValid || ws0->T != (int)T || ws0->c != (int)c || ws0->ckptVersion != w->version)
            return fail(LENIA_ERR_STATE, "lenia_rollout_backward: A0 is NULL but no forward rollout "
                        "with T=%u, checkpoints every %u and the current parameters precedes it", T, c);
    }
    rc = ws_prepare(w, T, c);
    if (!rc) rc = ensure_graphs(w);
    if (rc) return rc;
    AdjWorkspace *ws = w->adj;
    const size_t bytes = sizeof(float) * ws->n;
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    if (A0) {
        ws->ckptValid = false;
        e = hipMemcpyAsync(ws->ckpt, A0, bytes, kind_in(onDevice), s);
        if (e == hipSuccess) e = hipGraphLaunch(ws->fwdGraph, s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ws->g, gAT, bytes, kind_in(onDevice), s);
    if (e == hipSuccess) e = hipGraphLaunch(ws->bwdGraph, s);
    w->hostSteps += (A0 ? 2 : 1) * T;
    w->needsInit = false;
    w->specValid = true;
    if (e == hipSuccess && gA0) e = hipMemcpyAsync(gA0, ws->g, bytes, kind_out(onDevice), s);
    double acc[32];
    if (e == hipSuccess) e = hipMemcpyAsync(acc, ws->acc, sizeof(acc), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "rollout backward");
    if (A0) {
        ws->ckptValid = true;
        ws->ckptVersion = w->version;
    }
    if (grads) {
        std::memset(grads, 0, sizeof(*grads));
        grads->dDt = acc[0];
        for (int k = 0; k < ws->K; ++k) {
            grads->dMu[k] = acc[1 + 3 * k];
            grads->dSigma[k] = acc[2 + 3 * k];
            grads->dWeight[k] = acc[3 + 3 * k];
        }
        int i = 0;
        for (int k = 0; k < ws->K; ++k) {
            const lenia_kernel_desc &kd = w->desc.kernels[k];
            if (kd.shape != LENIA_SHAPE_RING) continue;
            for (int j = 0; j < (int)kd.nBetas; ++j) grads->dBetas[k][j] = acc[16 + i++];
        }
    }
    return LENIA_OK;
}
