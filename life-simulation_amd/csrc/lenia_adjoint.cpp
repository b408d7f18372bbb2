// Differentiable rollout: T FFT-form steps from a given state, and the
// hand-written reverse pass with sqrt(T)-style checkpointing.  No counterpart
// in the reference (its Lenia module only steps forward); the forward steps
// are exactly the library's own, so the reverse pass differentiates the
// arithmetic that produced the states.
//
// Backward over a segment [t0, t1): the states of the segment are recomputed
// bit-exactly from the checkpoint at t0 with the normal step kernels, then
// each step is reversed (k_adjoint.hip) from t1-1 down to t0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "launch.h"

using namespace lenia;

struct AdjWorkspace {
    size_t n = 0;        // cells
    int K = 0, T = 0, c = 0;
    float *ckpt = nullptr, *seg = nullptr;
    float *Ftmp = nullptr, *Ahat = nullptr, *Gk = nullptr, *U = nullptr, *GU = nullptr,
          *FG = nullptr, *G1 = nullptr, *gpb = nullptr, *g = nullptr;
    float *partC = nullptr, *partB = nullptr, *D = nullptr;
    double *acc = nullptr;
    int rowT = 256, colT = 256;
    AdjArgs args{};
};

static void ws_free(lenia_ctx *ctx, AdjWorkspace *ws) {
    float *ptrs[] = {ws->ckpt, ws->seg, ws->Ftmp, ws->Ahat, ws->Gk, ws->U, ws->GU,
                     ws->FG, ws->G1, ws->gpb, ws->g, ws->partC, ws->partB, ws->D};
    for (float *p : ptrs) dev_free(ctx, p);
    dev_free(ctx, ws->acc);
    *ws = AdjWorkspace();
}

void lenia::world_free_adjoint(lenia_world *w) {
    if (!w->adj) return;
    ws_free(w->ctx, w->adj);
    delete w->adj;
    w->adj = nullptr;
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

static int ws_prepare(lenia_world *w, uint32_t T, uint32_t c) {
    lenia_ctx *ctx = w->ctx;
    const int W = w->desc.width, H = w->desc.height, Wh = W / 2 + 1, K = (int)w->desc.nKernels;
    const size_t n = (size_t)W * H, spec = (size_t)Wh * H * 2;
    AdjWorkspace *ws = w->adj;
    const int nck = (int)((T + c - 1) / c);
    if (ws && ws->T == (int)T && ws->c == (int)c && ws->K == K) {
        ws->args.g = w->g;  // parameters may have changed
    } else {
        if (ws) world_free_adjoint(w);
        ws = new AdjWorkspace();
        w->adj = ws;
        ws->n = n;
        ws->K = K;
        ws->T = (int)T;
        ws->c = (int)c;
        int rc = dev_alloc(ctx, (void **)&ws->ckpt, sizeof(float) * n * nck);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->seg, sizeof(float) * n * c);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->Ftmp, sizeof(float) * spec);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->Ahat, sizeof(float) * spec);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->Gk, sizeof(float) * spec * K);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->U, sizeof(float) * n * K);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->GU, sizeof(float) * n * K);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->FG, sizeof(float) * spec * K);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->G1, sizeof(float) * spec);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->gpb, sizeof(float) * n);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->g, sizeof(float) * n);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->partC, sizeof(float) * 16 * ((n + 255) / 256));
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->partB, sizeof(float) * 16 * Wh);
        if (!rc) rc = dev_alloc(ctx, (void **)&ws->acc, sizeof(double) * 64);
        if (rc) {
            world_free_adjoint(w);
            return rc;
        }
        // shell-gradient spectra D_kj = spectrum of (c_kj - C_kj K_k) / S_k
        AdjArgs &a = ws->args;
        std::memset(&a, 0, sizeof(a));
        int ns = 0;
        std::vector<std::vector<double>> wins;
        for (int k = 0; k < K; ++k) {
            const lenia_kernel_desc &kd = w->desc.kernels[k];
            if (kd.shape != LENIA_SHAPE_RING) continue;
            const int h = kernel_half(kd), m = 2 * h + 1;
            std::vector<double> Kn, raw(m * m, 0.0);
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
                a.shellK[ns] = k;
                wins.push_back(dwin);
                ++ns;
            }
        }
        a.nShells = ns;
        if (ns) {
            rc = dev_alloc(ctx, (void **)&ws->D, sizeof(float) * (spec / 2) * ns);
            double *dWin = nullptr, *dCw = nullptr;
            int hmax = 0;
            for (int k = 0; k < K; ++k) hmax = std::max(hmax, kernel_half(w->desc.kernels[k]));
            const int mmax = 2 * hmax + 1;
            if (!rc) rc = dev_alloc(ctx, (void **)&dWin, sizeof(double) * mmax * mmax);
            if (!rc) rc = dev_alloc(ctx, (void **)&dCw, sizeof(double) * mmax * Wh);
            for (int i = 0; i < ns && !rc; ++i) {
                const int h = kernel_half(w->desc.kernels[a.shellK[i]]);
                hipError_t e = hipMemcpyAsync(dWin, wins[i].data(), sizeof(double) * wins[i].size(),
                                              hipMemcpyHostToDevice, ctx->stream);
                if (e == hipSuccess)
                    e = launch_kernel_spectrum(dWin, h, W, H, dCw, ws->D + (spec / 2) * i, ctx->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
                if (e != hipSuccess) rc = hip_fail(e, "shell spectra");
            }
            dev_free(ctx, dWin);
            dev_free(ctx, dCw);
            if (rc) {
                world_free_adjoint(w);
                return rc;
            }
        }
        a.W = W;
        a.H = H;
        a.K = K;
        a.planW = w->planW;
        a.planH = w->planH;
        a.g = w->g;
        a.twW = w->twW;
        a.twH = w->twH;
        a.S = w->S;
        a.D = ws->D;
        // workgroup sizes: room for K transforms at <= 16 values per thread
        const int mv = fft_max_values();
        ws->rowT = 0;
        ws->colT = 0;
        for (int T2 = 64; T2 <= 512 && !ws->rowT; T2 *= 2)
            if (plan_fits(w->planW, K, T2, mv) && plan_fits(w->planW, 1, T2, mv)) ws->rowT = T2;
        for (int T2 = 64; T2 <= 512 && !ws->colT; T2 *= 2)
            if (plan_fits(w->planH, K, T2, mv) && plan_fits(w->planH, 1, T2, mv)) ws->colT = T2;
        size_t ldsR = (size_t)(W + ((K * W + 31) & ~31)) * sizeof(float2);
        size_t ldsC = (size_t)(H + ((K * H + 31) & ~31) + ((K * H + 31) & ~31)) * sizeof(float2) +
                      512 * sizeof(float);
        if (!ws->rowT || !ws->colT || ldsR > 65536 || ldsC > 65536) {
            world_free_adjoint(w);
            return fail(LENIA_ERR_UNSUPPORTED, "rollout gradients: field %dx%d with %d kernels does "
                        "not fit the adjoint workgroups", W, H, K);
        }
    }
    return LENIA_OK;
}

static hipMemcpyKind kind_in(int onDevice) { return onDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice; }
static hipMemcpyKind kind_out(int onDevice) { return onDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost; }

// one forward step of the world from its current state (the normal kernels)
static int one_step(lenia_world *w) {
    int rc = world_launch_steps(w, 1);
    return rc;
}

// state := src (device), marks the spectrum stale
static hipError_t set_state(lenia_world *w, const float *src) {
    w->needsInit = false;
    w->specValid = false;
    return hipMemcpyAsync(w->current_state(), src, sizeof(float) * w->desc.width * w->desc.height,
                          hipMemcpyDeviceToDevice, w->ctx->stream);
}

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

static hipError_t backward_step(lenia_world *w, const float *A) {
    AdjWorkspace *ws = w->adj;
    const AdjArgs &a = ws->args;
    lenia_ctx *ctx = w->ctx;
    hipStream_t s = ctx->stream;
    const size_t n = ws->n;
    const int nBlocks = (int)((n + 255) / 256);
    const int Wh = a.W / 2 + 1;
    hipError_t e = launch_adj_rows_fwd(a, ws->rowT, A, 0, ws->Ftmp, 1, s);
    if (e == hipSuccess) e = launch_adj_cols(a, ws->colT, false, ws->Ftmp, ws->Ahat, ws->Gk, nullptr, s);
    if (e == hipSuccess) e = launch_adj_rows_inv(a, ws->rowT, ws->Gk, ws->U, a.K, nullptr, s);
    if (e == hipSuccess) e = launch_adj_cells(a, A, ws->U, ws->g, ws->gpb, ws->GU, ws->partC, s);
    if (e == hipSuccess) e = launch_adj_reduce(ws->partC, nBlocks, 1 + 3 * a.K, ws->acc, s);
    if (e == hipSuccess) e = launch_adj_rows_fwd(a, ws->rowT, ws->GU, n, ws->FG, a.K, s);
    if (e == hipSuccess) e = launch_adj_cols(a, ws->colT, true, ws->FG, ws->Ahat, ws->G1, ws->partB, s);
    if (e == hipSuccess && a.nShells) e = launch_adj_reduce(ws->partB, Wh, a.nShells, ws->acc + 16, s);
    if (e == hipSuccess) e = launch_adj_rows_inv(a, ws->rowT, ws->G1, ws->g, 1, ws->gpb, s);
    return e;
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
    const size_t n = ws->n, bytes = sizeof(float) * n;
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemsetAsync(ws->acc, 0, sizeof(double) * 64, s);
    // checkpoints: states at t = 0, c, 2c, ...
    if (e == hipSuccess) e = hipMemcpyAsync(ws->ckpt, A0, bytes, kind_in(onDevice), s);
    if (e == hipSuccess) e = set_state(w, ws->ckpt);
    const uint32_t nck = (T + c - 1) / c;
    for (uint32_t k = 1; k < nck && e == hipSuccess; ++k) {
        rc = world_launch_steps(w, c);
        if (rc) return rc;
        e = hipMemcpyAsync(ws->ckpt + n * k, w->current_state(), bytes, hipMemcpyDeviceToDevice, s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ws->g, gAT, bytes, kind_in(onDevice), s);
    for (int sg = (int)nck - 1; sg >= 0 && e == hipSuccess; --sg) {
        const uint32_t t0 = sg * c, t1 = std::min(T, t0 + c);
        e = set_state(w, ws->ckpt + n * sg);
        if (e == hipSuccess) e = hipMemcpyAsync(ws->seg, ws->ckpt + n * sg, bytes, hipMemcpyDeviceToDevice, s);
        for (uint32_t t = t0 + 1; t < t1 && e == hipSuccess; ++t) {
            rc = one_step(w);
            if (rc) return rc;
            e = hipMemcpyAsync(ws->seg + n * (t - t0), w->current_state(), bytes,
                               hipMemcpyDeviceToDevice, s);
        }
        for (int t = (int)t1 - 1; t >= (int)t0 && e == hipSuccess; --t)
            e = backward_step(w, ws->seg + n * (t - t0));
    }
    if (e == hipSuccess && gA0) e = hipMemcpyAsync(gA0, ws->g, bytes, kind_out(onDevice), s);
    double acc[64];
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
