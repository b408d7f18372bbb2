// Differentiable rollout: T FFT-form steps from a given state, and the
// hand-written reverse pass with sqrt(T)-style checkpointing.  No counterpart
// in the reference (its Lenia module only steps forward); the forward steps
// are exactly the library's own, so the reverse pass differentiates the
// arithmetic that produced the states.
//
// Reverse pass: a forward sweep stores the states at t = 0, c, 2c, ...; then,
// segment by segment from the last, the states of the segment are recomputed
// bit-exactly from its checkpoint with the normal step kernels and the steps
// are reversed from the segment's end to its start (k_adjoint.hip, four
// launches per step).  The whole sequence -- sweep, recomputes, reverse steps
// and the final sums -- is captured once per (T, c) into a hipGraph and
// replayed; every kernel reads the growth parameters from the world's device
// record, so parameter updates between calls keep the graph valid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "launch.h"

using namespace lenia;

struct AdjWorkspace {
    size_t n = 0;  // cells
    int K = 0, T = 0, c = 0, nShells = 0;
    float *ckpt = nullptr, *seg = nullptr;
    float *Ftmp = nullptr, *Ahat = nullptr, *Gk = nullptr, *FG = nullptr, *G1 = nullptr,
          *gpb = nullptr, *g = nullptr, *D = nullptr;
    double *accRows = nullptr, *accCols = nullptr, *acc = nullptr;
    int rowT = 256, colT = 256;
    bool dValid = false;  // D holds the shell spectra of the current kernels
    AdjArgs args{};
    hipGraphExec_t graph = nullptr;
    uint32_t graphShapes = 0;  // the world's shapeVersion the graph was recorded with
};

static void ws_free(lenia_ctx *ctx, AdjWorkspace *ws) {
    if (ws->graph) (void)hipGraphExecDestroy(ws->graph);
    float *ptrs[] = {ws->ckpt, ws->seg, ws->Ftmp, ws->Ahat, ws->Gk, ws->FG, ws->G1, ws->gpb, ws->g, ws->D};
    for (float *p : ptrs) dev_free(ctx, p);
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
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->seg, sizeof(float) * n * c);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->Ftmp, sizeof(float) * spec);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->Ahat, sizeof(float) * spec);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->Gk, sizeof(float) * spec * K);
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
        // workgroup sizes: about 4 values per thread for the largest batch of
        // transforms (max(K, 2) rows, K columns) -- short stages over many
        // threads, since one reverse step has only H/2 + W/2+1 workgroups --
        // within <= 16 values per thread and 64 KB of LDS
        const int mv = fft_max_values();
        const int nr = std::max(K, 2);
        auto pick = [&](const FftPlan &p, int nt, bool rows) {
            int want = 64;
            while (want < 512 && want * 4 < nt * p.n) want *= 2;
            for (int T2 = want; T2 <= 512; T2 *= 2)
                if (plan_fits(p, nt, T2, mv) && plan_fits(p, 1, T2, mv) &&
                    (rows ? adj_lds_rows(a, T2) : adj_lds_cols(a, T2)) <= 65536)
                    return T2;
            for (int T2 = want / 2; T2 >= 64; T2 /= 2)
                if (plan_fits(p, nt, T2, mv) && plan_fits(p, 1, T2, mv) &&
                    (rows ? adj_lds_rows(a, T2) : adj_lds_cols(a, T2)) <= 65536)
                    return T2;
            return 0;
        };
        ws->rowT = pick(w->planW, nr, true);
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

extern "C" int lenia_rollout_forward(lenia_world *w, uint32_t T, const float *A0, float *AT,
                                     int onDevice) {
    if (!w || !A0 || !AT) return fail(LENIA_ERR_INVALID_ARG, "lenia_rollout_forward: NULL argument");
    int rc = check_differentiable(w);
    if (rc) return rc;
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    const size_t n = (size_t)w->desc.width * w->desc.height;
    hipError_t e = hipMemcpyAsync(w->current_state(), A0, sizeof(float) * n, kind_in(onDevice), ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "rollout input");
    w->needsInit = false;
    w->specValid = false;
    rc = world_launch_steps(w, T);
    if (rc) return rc;
    e = hipMemcpyAsync(AT, w->current_state(), sizeof(float) * n, kind_out(onDevice), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? LENIA_OK : hip_fail(e, "rollout output");
}

// state := src (device) and its row spectrum, on the context stream
static hipError_t load_state(lenia_world *w, const float *src) {
    lenia_ctx *ctx = w->ctx;
    hipError_t e = hipMemcpyAsync(w->current_state(), src, sizeof(float) * w->adj->n,
                                  hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_fft_rows(0, w->rowT, false, false, (const FftWorldDev *)w->devDesc, w->rowTiles,
                            w->nRowTiles, w->nRowTiles, w->ldsRows, w->done, 1, 0, nullptr, ctx->stream);
    return e;
}

// the whole reverse pass for the workspace's (T, c), inputs already in
// ckpt[0] (A0) and g (dL/dA_T); leaves the parameter sums in acc and dL/dA0 in g
static hipError_t enqueue_backward(lenia_world *w) {
    AdjWorkspace *ws = w->adj;
    const AdjArgs &a = ws->args;
    lenia_ctx *ctx = w->ctx;
    hipStream_t s = ctx->stream;
    const size_t n = ws->n, bytes = sizeof(float) * n;
    const uint32_t T = ws->T, c = ws->c, nck = (T + c - 1) / c;
    const int H2 = a.H / 2, Wh = a.W / 2 + 1;
    // zero the per-workgroup FP64 slots (a fill kernel: graph-safe)
    hipError_t e = launch_fill(reinterpret_cast<float *>(ws->accRows), (size_t)32 * H2, 0.0f, s);
    if (e == hipSuccess) e = launch_fill(reinterpret_cast<float *>(ws->accCols), (size_t)32 * Wh, 0.0f, s);
    // forward sweep: checkpoints at t = c, 2c, ...
    if (e == hipSuccess) e = load_state(w, ws->ckpt);
    for (uint32_t k = 1; k < nck && e == hipSuccess; ++k) {
        for (uint32_t i = 0; i < c && e == hipSuccess; ++i) e = enqueue_step(w, nullptr);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ws->ckpt + n * k, w->current_state(), bytes, hipMemcpyDeviceToDevice, s);
    }
    for (int sg = (int)nck - 1; sg >= 0 && e == hipSuccess; --sg) {
        const uint32_t t0 = sg * c, t1 = std::min(T, t0 + c);
        // recompute the segment's states A_t0 .. A_t1-1 into seg
        e = load_state(w, ws->ckpt + n * sg);
        if (e == hipSuccess) e = hipMemcpyAsync(ws->seg, ws->ckpt + n * sg, bytes, hipMemcpyDeviceToDevice, s);
        for (uint32_t t = t0 + 1; t < t1 && e == hipSuccess; ++t) {
            e = enqueue_step(w, nullptr);
            if (e == hipSuccess)
                e = hipMemcpyAsync(ws->seg + n * (t - t0), w->current_state(), bytes, hipMemcpyDeviceToDevice, s);
        }
        // row spectrum of the segment's last state, then the reverse steps
        if (e == hipSuccess)
            e = launch_adj_edge(a, ws->rowT, nullptr, nullptr, nullptr, ws->seg + n * (t1 - 1 - t0), ws->Ftmp, s);
        for (int t = (int)t1 - 1; t >= (int)t0 && e == hipSuccess; --t) {
            const float *At = ws->seg + n * (t - t0);
            const float *Aprev = t > (int)t0 ? ws->seg + n * (t - 1 - t0) : nullptr;
            e = launch_adj_cols(a, ws->colT, false, ws->Ftmp, ws->Ahat, ws->Gk, nullptr, s);
            if (e == hipSuccess)
                e = launch_adj_mid(a, ws->rowT, ws->Gk, At, ws->g, ws->gpb, ws->FG, ws->accRows, s);
            if (e == hipSuccess)
                e = launch_adj_cols(a, ws->colT, true, ws->FG, ws->Ahat, ws->G1, ws->accCols, s);
            if (e == hipSuccess)
                e = launch_adj_edge(a, ws->rowT, ws->G1, ws->gpb, ws->g, Aprev, ws->Ftmp, s);
        }
    }
    if (e == hipSuccess)
        e = launch_adj_final(ws->accRows, H2, 1 + 3 * a.K, ws->accCols, Wh, a.nShells, ws->acc, s);
    // leave the world at A0 with a valid spectrum
    if (e == hipSuccess) e = load_state(w, ws->ckpt);
    return e;
}

static int build_backward_graph(lenia_world *w) {
    AdjWorkspace *ws = w->adj;
    hipStream_t s = w->ctx->stream;
    if (ws->graph) (void)hipGraphExecDestroy(ws->graph);
    ws->graph = nullptr;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "backward graph capture");
    hipError_t eb = enqueue_backward(w);
    e = hipStreamEndCapture(s, &graph);
    if (eb != hipSuccess) e = eb;
    if (e == hipSuccess) e = hipGraphInstantiate(&ws->graph, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        ws->graph = nullptr;
        return hip_fail(e, "backward graph");
    }
    return LENIA_OK;
}

extern "C" int lenia_rollout_backward(lenia_world *w, uint32_t T, uint32_t ckptEvery,
                                      const float *A0, const float *gAT, float *gA0,
                                      lenia_param_grads *grads, int onDevice) {
    if (!w || !A0 || !gAT || T == 0)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_rollout_backward: bad argument");
    int rc = check_differentiable(w);
    if (rc) return rc;
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    const uint32_t c = ckptEvery ? std::min(ckptEvery, T) : (uint32_t)std::max(1.0, std::floor(std::sqrt((double)T)));
    rc = ws_prepare(w, T, c);
    if (rc) return rc;
    AdjWorkspace *ws = w->adj;
    const size_t bytes = sizeof(float) * ws->n;
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemcpyAsync(ws->ckpt, A0, bytes, kind_in(onDevice), s);
    if (e == hipSuccess) e = hipMemcpyAsync(ws->g, gAT, bytes, kind_in(onDevice), s);
    if (e != hipSuccess) return hip_fail(e, "rollout backward inputs");
    if (!ws->graph || ws->graphShapes != w->shapeVersion) {
        rc = build_backward_graph(w);
        if (rc) return rc;
        ws->graphShapes = w->shapeVersion;
    }
    e = hipGraphLaunch(ws->graph, s);
    w->hostSteps += 2 * T;
    w->needsInit = false;
    w->specValid = true;
    if (e == hipSuccess && gA0) e = hipMemcpyAsync(gA0, ws->g, bytes, kind_out(onDevice), s);
    double acc[32];
    if (e == hipSuccess) e = hipMemcpyAsync(acc, ws->acc, sizeof(acc), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "rollout backward");
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
