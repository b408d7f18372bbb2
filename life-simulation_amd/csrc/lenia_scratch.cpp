// Allocation-free FFT-form stepping: lenia_fft_step_scratch.
//
// lenia_world_step runs on buffers a world allocates at create (state
// spectrum F, column-pass output G, kernel spectra S, twiddle tables, the
// device record).  Callers that cannot allocate device memory
// (code under stream capture, serving loops, harnesses that forbid hipMalloc)
// get the same step on a caller-owned device field with every one of those
// buffers carved out of one file-scope device array, `scratch`, of
// kScratchBytes per GPU.  The same plans, launch shapes, twiddles, kernel
// spectra and step kernels as a world of the same descriptor are used, so the
// results are bit-identical to lenia_world_step.
//
// The array's host-side name is not a device address: its per-device address
// comes from hipGetSymbolAddress, once per device (scratch_init_device, called
// by lenia_ctx_create with the device current), and only that address is used
// afterwards.  Every call on a GPU shares that GPU's array and leaves its
// contents behind; a call whose descriptor equals the previous call's on the
// device reuses the spectra and twiddles already there.  Calls that overlap
// on one GPU (two streams, two threads) race on the array.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "host_internal.h"
#include "launch.h"

using namespace lenia;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // per GPU, held until the process ends
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
    unsigned char *ptr = nullptr;  // this device's copy of `scratch`
    bool haveKey = false;          // the descriptor whose spectra the array holds
    lenia_world_desc key{};
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// the record of one call, written on the device from a kernel argument (no
// host buffer has to outlive the call)
__global__ void k_scratch_record(FftWorldDev rec, FftWorldDev *dst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = rec;
}

size_t up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// byte offsets of one call's buffers inside the array
struct Layout {
    size_t rec, clock, twW, twH, raw, win, cw, F, G, S, total;
    int nmax;
};

Layout layout_of(const lenia_world_desc &d) {
    const size_t W = d.width, H = d.height, Wh = W / 2 + 1, K = d.nKernels;
    int hmax = 0;
    for (uint32_t k = 0; k < d.nKernels; ++k) hmax = std::max(hmax, kernel_half(d.kernels[k]));
    Layout L;
    L.nmax = 2 * hmax + 1;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += up(bytes);
        return at;
    };
    L.rec = take(sizeof(FftWorldDev));
    L.clock = take(sizeof(StepClock));
    L.twW = take(W * sizeof(float2));
    L.twH = take(H * sizeof(float2));
    L.raw = take(std::max(W, H) * sizeof(float2));
    L.win = take(K * (size_t)L.nmax * L.nmax * sizeof(double));
    L.cw = take((size_t)L.nmax * Wh * sizeof(double));
    L.F = take(Wh * H * sizeof(float2));
    L.G = take(K * Wh * H * sizeof(float2));
    L.S = take(K * Wh * H * sizeof(float));
    L.total = o;
    return L;
}

}  // namespace

int lenia::scratch_init_device(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(LENIA_ERR_UNSUPPORTED, "scratch: device %d beyond the %d this build tracks", device,
                    kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_dev[device].ptr) return LENIA_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                               "hipGetSymbolAddress(scratch)");
    g_dev[device].ptr = static_cast<unsigned char *>(p);
    return LENIA_OK;
}

extern "C" int lenia_fft_step_scratch(lenia_ctx *ctx, const lenia_world_desc *desc,
                                      uint32_t sceneSeed, uint32_t frameIndex, float *state,
                                      uint32_t nSteps) {
    if (!ctx || !desc || !state)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_fft_step_scratch: NULL argument");
    int rc = world_check_desc(desc);
    if (rc) return rc;
    if (desc->convMode != LENIA_CONV_FFT)
        return fail(LENIA_ERR_UNSUPPORTED, "lenia_fft_step_scratch: FFT form only (convMode %u)",
                    desc->convMode);
    // the host side of a world, resolved as lenia_world_create resolves it;
    // its buffers all point into the scratch array
    lenia_world hw;
    hw.ctx = ctx;
    hw.desc = *desc;
    hw.isFft = true;
    rc = fft_sizes_and_tiles(&hw, ctx->numCUs);
    if (rc) return rc;
    const Layout L = layout_of(*desc);
    if (L.total > kScratchBytes)
        return fail(LENIA_ERR_INVALID_ARG,
                    "lenia_fft_step_scratch: a %ux%u field with %u kernels needs %zu bytes of "
                    "scratch; the maximum is %zu (use lenia_world_create for larger fields)",
                    desc->width, desc->height, desc->nKernels, L.total, kScratchBytes);
    (void)hipSetDevice(ctx->device);
    rc = scratch_init_device(ctx->device);
    if (rc) return rc;
    DeviceScratch &ds = g_dev[ctx->device];
    unsigned char *base = ds.ptr;
    hipStream_t s = ctx->stream;
    const int W = desc->width, H = desc->height, Wh = W / 2 + 1, K = (int)desc->nKernels;
    hw.moduleSeed = module_seed(sceneSeed);
    hw.g = world_growth(desc, hw.moduleSeed);
    hw.A0 = state;
    hw.F = reinterpret_cast<float *>(base + L.F);
    hw.G = reinterpret_cast<float *>(base + L.G);
    hw.S = reinterpret_cast<float *>(base + L.S);
    hw.twW = reinterpret_cast<float *>(base + L.twW);
    hw.twH = reinterpret_cast<float *>(base + L.twH);
    hw.clock = reinterpret_cast<StepClock *>(base + L.clock);
    hw.devDesc = base + L.rec;
    const FftWorldDev rec = fft_record(&hw);

    hipError_t e = hipSuccess;
    const bool same = ds.haveKey && std::memcmp(&ds.key, desc, sizeof(*desc)) == 0;
    if (!same) {
        // twiddles and kernel spectra exactly as a world of this descriptor
        // computes them (FP64 tables on the GPU, host FP64 windows)
        ds.haveKey = false;
        float *raw = reinterpret_cast<float *>(base + L.raw);
        double *win = reinterpret_cast<double *>(base + L.win);
        double *cw = reinterpret_cast<double *>(base + L.cw);
        std::vector<std::vector<double>> wins(K);
        for (int k = 0; k < K; ++k)
            if (!kernel_window(desc->kernels[k], kernel_half(desc->kernels[k]), wins[k]))
                return fail(LENIA_ERR_INVALID_ARG, "kernel %d has zero mass", k);
        e = launch_twiddles(raw, W, s);
        if (e == hipSuccess) e = launch_stage_twiddles(hw.planW, raw, hw.twW, s);
        if (e == hipSuccess) e = launch_twiddles(raw, H, s);
        if (e == hipSuccess) e = launch_stage_twiddles(hw.planH, raw, hw.twH, s);
        for (int k = 0; k < K && e == hipSuccess; ++k) {
            const int h = kernel_half(desc->kernels[k]);
            double *wk = win + (size_t)k * L.nmax * L.nmax;
            e = hipMemcpyAsync(wk, wins[k].data(), wins[k].size() * sizeof(double),
                               hipMemcpyHostToDevice, s);
            if (e == hipSuccess)
                e = launch_kernel_spectrum(wk, h, W, H, cw, hw.S + (size_t)k * Wh * H, s);
        }
        // G starts cleared, as a new world's does
        if (e == hipSuccess) e = launch_fill(hw.G, (size_t)K * Wh * H * 2, 0.0f, s);
        // the host windows must outlive their copies (also on failure)
        hipError_t es = hipStreamSynchronize(s);
        if (e == hipSuccess) e = es;
        if (e != hipSuccess) return hip_fail(e, "lenia_fft_step_scratch: set-up");
        ds.key = *desc;
        ds.haveKey = true;
    }
    hipLaunchKernelGGL(k_scratch_record, dim3(1), dim3(64), 0, s, rec,
                       reinterpret_cast<FftWorldDev *>(hw.devDesc));
    e = hipGetLastError();
    // step i runs at frame frameIndex + i, substep 0 (lenia_world_step's clock)
    if (e == hipSuccess) e = launch_set_clock(hw.clock, frameIndex, 0, 1, 0, s);
    if (e != hipSuccess) return hip_fail(e, "lenia_fft_step_scratch: record");
    hw.specValid = false;
    rc = world_ensure_spectrum(&hw);
    if (rc) return rc;
    for (uint32_t i = 0; i < nSteps && e == hipSuccess; ++i) e = enqueue_step(&hw, nullptr, i);
    if (e != hipSuccess) return hip_fail(e, "lenia_fft_step_scratch: step launch");
    return LENIA_OK;
}
