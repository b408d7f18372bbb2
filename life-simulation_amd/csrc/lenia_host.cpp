// Host side of the C ABI: contexts, worlds, batches, kernel building, FFT
// planning and launch orchestration.  Mirrors LeniaModule's lifecycle
// (Modules/FieldModules/Lenia.cpp:208-653) on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lenia_mi355x.h"
#include "host_internal.h"
#include "launch.h"

using namespace lenia;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

int lenia::fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int lenia::hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? LENIA_ERR_OUT_OF_MEMORY : LENIA_ERR_DEVICE,
                "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                       \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)


extern "C" const char *lenia_last_error(void) { return g_err.c_str(); }
extern "C" const char *lenia_version(void) { return "lenia-mi355x 0.1 (gfx950)"; }

// ------------------------------------------------------------------ device memory
int lenia::dev_alloc(lenia_ctx *ctx, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return LENIA_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    ctx->liveBytes += bytes;
    ctx->liveAllocs += 1;
    ctx->allocSizes.push_back({*p, bytes});
    return LENIA_OK;
}

void lenia::dev_free(lenia_ctx *ctx, void *p) {
    if (!p) return;
    for (size_t i = 0; i < ctx->allocSizes.size(); ++i) {
        if (ctx->allocSizes[i].first == p) {
            ctx->liveBytes -= ctx->allocSizes[i].second;
            ctx->liveAllocs -= 1;
            ctx->allocSizes.erase(ctx->allocSizes.begin() + i);
            break;
        }
    }
    (void)hipFree(p);
}

// ------------------------------------------------------------------ context
extern "C" int lenia_ctx_create(int device, lenia_ctx **out) {
    if (!out) return fail(LENIA_ERR_INVALID_ARG, "lenia_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_ctx_create: device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(LENIA_ERR_UNSUPPORTED, "lenia_ctx_create: device is %s, this build is gfx950 only",
                    prop.gcnArchName);
    // the allocation-free step's per-device scratch address, looked up once
    // per device with the device current
    int rc = scratch_init_device(device);
    if (rc) return rc;
    lenia_ctx *c = new lenia_ctx();
    c->device = device;
    c->numCUs = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    (void)hipEventCreate(&c->sp0);
    (void)hipEventCreate(&c->sp1);
    *out = c;
    return LENIA_OK;
}

extern "C" int lenia_ctx_destroy(lenia_ctx *ctx) {
    if (!ctx) return LENIA_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->sp0);
    (void)hipEventDestroy(ctx->sp1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LENIA_OK;
}

extern "C" int lenia_ctx_sync(lenia_ctx *ctx) {
    if (!ctx) return fail(LENIA_ERR_INVALID_ARG, "lenia_ctx_sync: NULL context");
    (void)hipSetDevice(ctx->device);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "stream synchronize");
    return LENIA_OK;
}

extern "C" int lenia_ctx_stream(lenia_ctx *ctx, void **stream) {
    if (!ctx || !stream) return fail(LENIA_ERR_INVALID_ARG, "lenia_ctx_stream: NULL argument");
    *stream = (void *)ctx->stream;
    return LENIA_OK;
}

extern "C" int lenia_ctx_memory(lenia_ctx *ctx, uint64_t *liveBytes, uint32_t *liveAllocs) {
    if (!ctx) return fail(LENIA_ERR_INVALID_ARG, "lenia_ctx_memory: NULL context");
    if (liveBytes) *liveBytes = ctx->liveBytes;
    if (liveAllocs) *liveAllocs = ctx->liveAllocs;
    return LENIA_OK;
}

extern "C" int lenia_ctx_last_gpu_ms(lenia_ctx *ctx, double *ms) {
    if (!ctx || !ms) return fail(LENIA_ERR_INVALID_ARG, "lenia_ctx_last_gpu_ms: NULL argument");
    if (!ctx->timed) return fail(LENIA_ERR_STATE, "no timed step has run on this context");
    (void)hipSetDevice(ctx->device);
    HIP_TRY(hipEventSynchronize(ctx->last1), "hipEventSynchronize");
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, ctx->last0, ctx->last1), "hipEventElapsedTime");
    *ms = f;
    return LENIA_OK;
}

// ------------------------------------------------------------------ kernels
// raw (unnormalised) kernel value at integer offset (dx, dy): the two reference
// builders, buildKernelTexture (Lenia.cpp:419-431) and buildRingKernelImage
// (Lenia.cpp:103-117), with an optional exponential-bump core.  Evaluated in
// double, rounded once to float after normalisation.
static double core_of(uint32_t core, double x, double centre, double sigma) {
    if (core == LENIA_CORE_EXP) {
        if (x <= 0.0 || x >= 1.0) return 0.0;
        return std::exp(4.0 - 1.0 / (x * (1.0 - x)));
    }
    double t = (x - centre) / sigma;
    return std::exp(-0.5 * t * t);
}

double lenia::kernel_raw(const lenia_kernel_desc &k, int dx, int dy) {
    double rr = std::sqrt(double(dx) * dx + double(dy) * dy);
    const double R = k.radius;
    if (k.shape == LENIA_SHAPE_SHELL) {
        double r = rr / R;
        if (!(r > 0.0 && r <= 1.0)) return 0.0;
        return core_of(k.core, r, k.shellMu, k.shellSigma);
    }
    if (!(rr > 0.0 && rr <= R)) return 0.0;
    const int B = (int)k.nBetas;
    double u = (rr / R) * B;
    int i = std::min((int)u, B - 1);
    double frac = u - i;
    if (k.core == LENIA_CORE_EXP) return k.betas[i] * core_of(LENIA_CORE_EXP, frac, 0.5, 1.0);
    double t = (frac - 0.5) / ((double)k.shellSigma * B);
    return k.betas[i] * std::exp(-0.5 * t * t);
}

int lenia::kernel_half(const lenia_kernel_desc &k) { return (int)std::ceil((double)k.radius - 1e-9); }

// (2h+1)^2 centred window normalised to sum 1 (double)
bool lenia::kernel_window(const lenia_kernel_desc &k, int h, std::vector<double> &out) {
    const int n = 2 * h + 1;
    out.assign((size_t)n * n, 0.0);
    double sum = 0.0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            double w = kernel_raw(k, x - h, y - h);
            out[(size_t)y * n + x] = w;
            sum += w;
        }
    if (!(sum > 0.0)) return false;
    for (auto &w : out) w /= sum;
    return true;
}

// ------------------------------------------------------------------ FFT plans
bool lenia::factor235(int n, int &a, int &b, int &c) {
    a = b = c = 0;
    while (n % 2 == 0) { n /= 2; ++a; }
    while (n % 3 == 0) { n /= 3; ++b; }
    while (n % 5 == 0) { n /= 5; ++c; }
    return n == 1;
}

static uint32_t div_magic(int d) {
    if (d <= 1) return 0u;
    return (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d);
}

// The radix sequences with a compile-time kernel (fft_run): the default plan
// of each common length, and for the benchmark lengths the other orders the
// radix-plan autotuner times.
namespace {
struct Fixed {
    int n, ns, r[5];
    uint32_t code;
};
const Fixed kFixedPlans[] = {
    {64, 2, {8, 8}, kPlan64},          {128, 2, {16, 8}, kPlan128},         {256, 2, {16, 16}, kPlan256},
    {512, 3, {8, 8, 8}, kPlan512},     {1024, 3, {16, 8, 8}, kPlan1024},    {2048, 3, {16, 16, 8}, kPlan2048},
    {96, 3, {8, 4, 3}, kPlan96},       {192, 3, {8, 8, 3}, kPlan192},       {384, 3, {16, 8, 3}, kPlan384},
    {720, 4, {16, 3, 3, 5}, kPlan720}, {1080, 5, {8, 3, 3, 3, 5}, kPlan1080}, {1280, 3, {16, 16, 5}, kPlan1280},
    {1920, 4, {16, 8, 3, 5}, kPlan1920}, {4096, 3, {16, 16, 16}, kPlan4096},
    {64, 2, {16, 4}, kPlan64b},        {64, 2, {4, 16}, kPlan64c},
    {128, 2, {8, 16}, kPlan128b},      {128, 3, {4, 4, 8}, kPlan128c},
    {256, 3, {8, 8, 4}, kPlan256b},    {256, 3, {4, 8, 8}, kPlan256c},
    {512, 3, {16, 8, 4}, kPlan512b},   {512, 3, {4, 8, 16}, kPlan512c},
    {1024, 3, {8, 8, 16}, kPlan1024b}, {1024, 3, {16, 16, 4}, kPlan1024c},
    {1024, 3, {4, 16, 16}, kPlan1024d}, {2048, 3, {8, 16, 16}, kPlan2048b},
    {4096, 4, {8, 8, 8, 8}, kPlan4096b}, {96, 3, {4, 8, 3}, kPlan96b},
    {192, 3, {16, 4, 3}, kPlan192b},   {384, 3, {8, 16, 3}, kPlan384b},
    {384, 4, {8, 4, 4, 3}, kPlan384c}};
}  // namespace

// fills the per-stage constants of a radix sequence; false if the product is
// not n or a division magic is not exact on the index range it serves
bool lenia::plan_make(int n, const int *radix, int ns, FftPlan &p) {
    if (ns < 1 || ns > kMaxStages) return false;
    std::memset(&p, 0, sizeof(p));
    long prod = 1;
    int Ns = 1, twOff = 0;
    for (int s = 0; s < ns; ++s) {
        int R = radix[s];
        if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8 && R != 16 && R != 32) return false;
        prod *= R;
        if (n % (Ns * R) != 0) return false;
        FftStage &st = p.st[s];
        st.radix = R;
        st.NR = n / R;
        st.Ns = Ns;
        st.twStep = n / (Ns * R);
        st.twOff = twOff;
        twOff += Ns * (R - 1);
        st.mNR = div_magic(st.NR);
        st.mNs = div_magic(Ns);
        // the kernels divide indices below 2^16 (nt * N <= 4096 in practice)
        for (uint32_t g = 0; g < 65536u; g += 1u) {
            uint32_t q1 = st.mNR ? (uint32_t)(((uint64_t)g * st.mNR) >> 32) : g;
            uint32_t q2 = st.mNs ? (uint32_t)(((uint64_t)g * st.mNs) >> 32) : g;
            if (q1 != g / (uint32_t)st.NR || q2 != g / (uint32_t)Ns) return false;
        }
        Ns *= R;
    }
    if (prod != n) return false;
    p.n = n;
    p.nStages = ns;
    p.code = kPlanGeneric;
    for (const Fixed &f : kFixedPlans)
        if (f.n == n && f.ns == ns && std::equal(radix, radix + ns, f.r)) p.code = f.code;
    return true;
}

// packed radix sequences of length n that have a compile-time kernel
int lenia::compiled_plans(int n, uint32_t *packed, int maxOut) {
    int m = 0;
    for (const Fixed &f : kFixedPlans) {
        if (f.n != n) continue;
        FftPlan p;
        if (m < maxOut && plan_make(n, f.r, f.ns, p)) packed[m++] = plan_pack(p);
    }
    return m;
}

// radix sequence for n: the power-of-two part in the fewest stages of radix
// <= maxR (balanced, e.g. 1024 = 16*8*8, 512 = 8*8*8), then 3s and 5s
bool lenia::plan_default(int n, int maxR, FftPlan &p) {
    int a, b, c;
    if (n < 2 || !factor235(n, a, b, c)) return false;
    int lmax = 1;
    while ((1 << (lmax + 1)) <= maxR && lmax < 5) ++lmax;
    std::vector<int> r;
    if (a > 0) {
        int stages = (a + lmax - 1) / lmax;
        int base = a / stages, extra = a % stages;
        for (int s = 0; s < stages; ++s) r.push_back(1 << (base + (s < extra ? 1 : 0)));
    }
    for (int i = 0; i < b; ++i) r.push_back(3);
    for (int i = 0; i < c; ++i) r.push_back(5);
    return plan_make(n, r.data(), (int)r.size(), p);
}

bool lenia::plan_unpack(int n, uint32_t packed, FftPlan &p) {
    int r[kMaxStages], s = 0;
    for (; s < kMaxStages; ++s) {
        int v = (packed >> (4 * s)) & 0xF;
        if (v == 0) break;
        r[s] = v == 0xF ? 16 : (v == 0xE ? 32 : v);  // nibbles F/E encode 16/32
    }
    return plan_make(n, r, s, p);
}

uint32_t lenia::plan_pack(const FftPlan &p) {
    uint32_t v = 0;
    for (int s = 0; s < p.nStages; ++s) {
        int R = p.st[s].radix;
        uint32_t code = R == 16 ? 0xF : (R == 32 ? 0xE : (uint32_t)R);
        v |= code << (4 * s);
    }
    return v;
}

// can one workgroup of T threads run nt transforms of this plan with at most
// maxv values per thread per stage?
bool lenia::plan_fits(const FftPlan &p, int nt, int T, int maxv) {
    if ((long)nt * p.n > (long)T * maxv) return false;
    for (int s = 0; s < p.nStages; ++s) {
        int R = p.st[s].radix;
        if (R > maxv) return false;
        long nb = (long)nt * p.n / R;
        long perThread = (nb + T - 1) / T;
        if (perThread * R > maxv) return false;
    }
    return true;
}

// ------------------------------------------------------------------ descriptor checks
static int check_desc(const lenia_world_desc *d) {
    if (!d) return fail(LENIA_ERR_INVALID_ARG, "world descriptor is NULL");
    if (d->width < 1 || d->height < 1 || d->width > 16384 || d->height > 16384)
        return fail(LENIA_ERR_INVALID_ARG, "field size %ux%u out of range", d->width, d->height);
    if (d->nKernels < 1 || d->nKernels > 4)
        return fail(LENIA_ERR_INVALID_ARG, "nKernels must be 1..4 (got %u)", d->nKernels);
    if (d->convMode > LENIA_CONV_DIRECT_BF16X3)
        return fail(LENIA_ERR_INVALID_ARG, "unknown convMode %u", d->convMode);
    if (d->growthMode > 1) return fail(LENIA_ERR_INVALID_ARG, "growthMode must be 0 or 1");
    if (!(d->dt >= 0.0f) || !std::isfinite(d->dt)) return fail(LENIA_ERR_INVALID_ARG, "dt must be finite and >= 0");
    if (d->initMode > LENIA_INIT_STAMP) return fail(LENIA_ERR_INVALID_ARG, "unknown initMode");
    for (uint32_t k = 0; k < d->nKernels; ++k) {
        const lenia_kernel_desc &kd = d->kernels[k];
        if (kd.shape > LENIA_SHAPE_RING || kd.core > LENIA_CORE_EXP)
            return fail(LENIA_ERR_INVALID_ARG, "kernel %u: unknown shape/core", k);
        if (!(kd.radius >= 1.0f) || kd.radius > 256.0f)
            return fail(LENIA_ERR_INVALID_ARG, "kernel %u: radius %g out of range [1,256]", k, kd.radius);
        if (kd.nBetas < 1 || kd.nBetas > 4)
            return fail(LENIA_ERR_INVALID_ARG, "kernel %u: nBetas must be 1..4", k);
        if (!(kd.sigma > 0.0f) || !std::isfinite(kd.mu))
            return fail(LENIA_ERR_INVALID_ARG, "kernel %u: growth sigma must be > 0", k);
        if (kd.core == LENIA_CORE_GAUSSIAN && !(kd.shellSigma > 0.0f))
            return fail(LENIA_ERR_INVALID_ARG, "kernel %u: shellSigma must be > 0", k);
        int h = kernel_half(kd);
        if (d->convMode == LENIA_CONV_DIRECT && (h > 43 || direct_lds_bytes(h, kDirectTileH, true) > kFftLdsMax))
            return fail(LENIA_ERR_UNSUPPORTED,
                        "kernel %u: the FP32 direct form supports window half-widths h <= 43 "
                        "(got %d); use the FFT form", k, h);
        if (2 * h + 1 > (int)std::min(d->width, d->height))
            return fail(LENIA_ERR_INVALID_ARG,
                        "kernel %u: window 2*%d+1 exceeds the field (%ux%u); the torus would "
                        "wrap onto itself", k, h, d->width, d->height);
    }
    return LENIA_OK;
}

static GrowthParams growth_of(const lenia_world_desc *d, uint32_t moduleSeed) {
    GrowthParams g;
    std::memset(&g, 0, sizeof(g));
    g.dt = d->dt;
    g.nKernels = d->nKernels;
    g.growthMode = d->growthMode;
    g.noiseAmount = d->noiseAmount;
    g.muJitter = d->muJitter;
    g.injectAmount = d->injectAmount;
    g.injectCount = d->injectCount;
    g.radius = d->radius;
    g.moduleSeed = moduleSeed;
    float sw = 0.0f;
    for (int k = 0; k < kMaxKernels; ++k) {
        bool on = k < (int)d->nKernels;
        g.mu[k] = on ? d->kernels[k].mu : 0.15f;
        g.sigma[k] = on ? d->kernels[k].sigma : 0.017f;
        g.bellScale[k] = (float)(std::sqrt(0.5 / std::log(2.0)) / (double)g.sigma[k]);
        g.weight[k] = on ? d->kernels[k].weight : 0.0f;
        if (on) sw += g.weight[k];  // same accumulation order as leniaGrowthMulti
    }
    g.sumWeights = sw;
    return g;
}

// ------------------------------------------------------------------ world setup
// FP32 -> BF16 bits, round to nearest even (finite values), and back
static uint16_t bf16_bits_host(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float bf16_value_host(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

// MFMA forms: Toeplitz fragments of every window row in MFMA register order.
// For kernel k, window row dy, k-step s and lane l (output column m = l % 32,
// k-group g = l / 32) the 8 values K[dy][16 s + 8 g + i - m - shift], i < 8
// (zero outside the window), as three BF16 planes: hi = bf16(v), mid =
// bf16(v - hi), lo = bf16(v - hi - mid).  The staged state tile starts shift
// columns left of x0 - h (mfma_shift), so that it loads four aligned columns
// at a time.
static int build_toeplitz(lenia_world *w, const std::vector<float> &all, int h) {
    lenia_ctx *ctx = w->ctx;
    const int K = (int)w->desc.nKernels, n = 2 * h + 1;
    const int shift = mfma_shift(h, w->desc.width);
    const int ks = mfma_ksteps(h, w->desc.width);
    w->ksteps = ks;
    const size_t nfr = (size_t)K * n * ks * 64 * 8;
    std::vector<uint16_t> hi(nfr), mid(nfr), lo(nfr);
    for (int k = 0; k < K; ++k)
        for (int dy = 0; dy < n; ++dy)
            for (int s = 0; s < ks; ++s)
                for (int l = 0; l < 64; ++l)
                    for (int i = 0; i < 8; ++i) {
                        const int t = 16 * s + 8 * (l / 32) + i - (l % 32) - shift;
                        const float v = (t >= 0 && t < n) ? all[((size_t)k * n + dy) * n + t] : 0.0f;
                        const size_t idx = ((((size_t)k * n + dy) * ks + s) * 64 + l) * 8 + i;
                        hi[idx] = bf16_bits_host(v);
                        const float r1 = v - bf16_value_host(hi[idx]);
                        mid[idx] = bf16_bits_host(r1);
                        lo[idx] = bf16_bits_host(r1 - bf16_value_host(mid[idx]));
                    }
    const size_t bytes = nfr * sizeof(uint16_t);
    if (!w->toepHi || w->toepBytes != bytes) {
        dev_free(ctx, w->toepHi);
        dev_free(ctx, w->toepMid);
        dev_free(ctx, w->toepLo);
        w->toepHi = w->toepMid = w->toepLo = nullptr;
        int rc = dev_alloc(ctx, (void **)&w->toepHi, bytes);
        if (!rc) rc = dev_alloc(ctx, (void **)&w->toepMid, bytes);
        if (!rc) rc = dev_alloc(ctx, (void **)&w->toepLo, bytes);
        if (rc) return rc;
        w->toepBytes = bytes;
    }
    HIP_TRY(hipMemcpyAsync(w->toepHi, hi.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
            "upload Toeplitz fragments");
    HIP_TRY(hipMemcpyAsync(w->toepMid, mid.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
            "upload Toeplitz fragments");
    HIP_TRY(hipMemcpyAsync(w->toepLo, lo.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
            "upload Toeplitz fragments");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "upload Toeplitz fragments");
    return LENIA_OK;
}

static int build_kernels_direct(lenia_world *w) {
    lenia_ctx *ctx = w->ctx;
    const lenia_world_desc &d = w->desc;
    int h = 0;
    for (uint32_t k = 0; k < d.nKernels; ++k) h = std::max(h, kernel_half(d.kernels[k]));
    w->half = h;
    const int n = 2 * h + 1;
    std::vector<float> all((size_t)d.nKernels * n * n, 0.0f);
    std::vector<double> win;
    for (uint32_t k = 0; k < d.nKernels; ++k) {
        if (!kernel_window(d.kernels[k], h, win))
            return fail(LENIA_ERR_INVALID_ARG, "kernel %u has zero mass", k);
        for (size_t i = 0; i < win.size(); ++i) all[(size_t)k * n * n + i] = (float)win[i];
    }
    if (w->isMfma) return build_toeplitz(w, all, h);
    // FP32 direct form: window rows zero padded to a multiple of 8 taps,
    // followed by each row's range of 8-tap chunks holding a non-zero weight
    // (the kernel skips the all-zero chunks at both ends of a row: adding
    // w = 0 times a finite state value leaves a sum unchanged, so the bits
    // do not change)
    const int np = direct_npad(h);
    const size_t nrow = (size_t)d.nKernels * n;
    std::vector<float> padded(nrow * np + 2 * nrow, 0.0f);
    int32_t *chunks = reinterpret_cast<int32_t *>(padded.data() + nrow * np);
    for (size_t r = 0; r < nrow; ++r) {
        int lo = np / 8, hi = 0;
        for (int x = 0; x < n; ++x) {
            padded[r * np + x] = all[r * n + x];
            if (all[r * n + x] != 0.0f) {
                lo = std::min(lo, x / 8);
                hi = std::max(hi, x / 8 + 1);
            }
        }
        if (hi == 0) lo = 0;  // an all-zero row: empty range
        chunks[2 * r] = lo;
        chunks[2 * r + 1] = hi;
    }
    size_t bytes = padded.size() * sizeof(float);
    if (!w->winF || w->winBytes != bytes) {
        dev_free(ctx, w->winF);
        int rc = dev_alloc(ctx, (void **)&w->winF, bytes);
        if (rc) return rc;
        w->winBytes = bytes;
    }
    HIP_TRY(hipMemcpyAsync(w->winF, padded.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
            "upload kernel windows");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "upload kernel windows");
    return LENIA_OK;
}

static int build_kernels_fft(lenia_world *w) {
    lenia_ctx *ctx = w->ctx;
    const lenia_world_desc &d = w->desc;
    const int W = d.width, H = d.height, Wh = W / 2 + 1;
    std::vector<double> win;
    int hmax = 0;
    for (uint32_t k = 0; k < d.nKernels; ++k) hmax = std::max(hmax, kernel_half(d.kernels[k]));
    const int nmax = 2 * hmax + 1;
    double *dWin = nullptr, *dCw = nullptr;
    int rc = dev_alloc(ctx, (void **)&dWin, sizeof(double) * nmax * nmax);
    if (rc) return rc;
    rc = dev_alloc(ctx, (void **)&dCw, sizeof(double) * nmax * Wh);
    if (rc) {
        dev_free(ctx, dWin);
        return rc;
    }
    for (uint32_t k = 0; k < d.nKernels && rc == LENIA_OK; ++k) {
        int h = kernel_half(d.kernels[k]);
        if (!kernel_window(d.kernels[k], h, win)) {
            rc = fail(LENIA_ERR_INVALID_ARG, "kernel %u has zero mass", k);
            break;
        }
        hipError_t e = hipMemcpyAsync(dWin, win.data(), sizeof(double) * win.size(),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = launch_kernel_spectrum(dWin, h, W, H, dCw, w->S + (size_t)k * Wh * H, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "kernel spectrum");
    }
    dev_free(ctx, dWin);
    dev_free(ctx, dCw);
    return rc;
}

// Launch shape of one FFT pass: transforms per tile (nt), threads per
// workgroup (T), whether the K inverse transforms of a tile run side by side
// (kpar, K transforms per tile slot), and the radix plan.  Explicit values
// from the descriptor win.  Otherwise: the smallest tile (nt = 1) with the
// smallest workgroup that holds <= 16 values per thread, side by side when it
// fits in the workgroup's LDS.

// Transform regions span lds_span(n) entries (one unused entry after every
// 32, the LDS placement of the kernels).
size_t lenia::lds_rows(int W, int P, int K, bool kpar, int T) {
    const int twin = rows_twin(W, P, K, kpar, T) ? 2 : 1;
    return (size_t)(W + twin * rows_region(W, P, K, kpar)) * sizeof(float2);
}
size_t lenia::lds_cols(int H, int C, int K, bool kpar) {
    return (size_t)(H + lds_span(C * H) + (kpar ? lds_span(K * C * H) : 0)) * sizeof(float2);
}

// does (nt transforms per tile, T threads, side-by-side kernels or not) fit
// this plan, the per-thread register budget and the workgroup's LDS?
bool lenia::shape_ok(const FftPlan &p, bool rows, int nt, int K, int T, bool kpar) {
    // 1024-thread workgroups: the row pass with side-by-side transforms only
    if (T > 512 && (!rows || !kpar)) return false;
    const int mv = fft_max_values();
    const size_t lds = rows ? lds_rows(p.n, nt, K, kpar, T) : lds_cols(p.n, nt, K, kpar);
    if (lds > kFftLdsMax) return false;
    if (rows && (long)(kpar ? nt * K : nt) * p.n > (long)T * fft_row_values(T)) return false;
    // side-by-side gather: each thread loads V/2 + 1 spectrum slots, and the
    // tile has K nt (W/2 + 1).  Every launched variant has K nt W <= V T, so
    // K nt <= T makes K nt (W/2 + 1) <= T (V/2 + 1): no slot is left unloaded.
    if (rows && kpar && nt * K > T) return false;
    return plan_fits(p, kpar ? nt * K : nt, T, mv) && plan_fits(p, nt, T, mv);
}

// The radix plan of a length depends only on the length (or the explicit
// override), never on the launch shape, so every shape -- and every batch a
// world joins -- computes the same bits.  Largest radices first; smaller
// ones only when no shape can run the plan.
static bool choose_plan(int N, int K, uint32_t packed, FftPlan &p, bool rows) {
    auto any_shape = [&](const FftPlan &q) {
        for (int T = 64; T <= 1024; T *= 2)
            for (int kp = 0; kp < 2; ++kp)
                if (shape_ok(q, rows, 1, K, T, kp != 0)) return true;
        return false;
    };
    if (packed) return plan_unpack(N, packed, p) && any_shape(p);
    // a plan recorded for this length and axis (lenia_plan_autotune, plan cache)
    const uint32_t cached = plan_cache_lookup(rows ? 0u : 1u, (uint32_t)N);
    if (cached && plan_unpack(N, cached, p) && any_shape(p)) return true;
    for (int maxR = 16; maxR >= 2; maxR /= 2)
        if (plan_default(N, maxR, p) && any_shape(p)) return true;
    return false;
}

// Default launch shape: transforms side by side when they fit, one transform
// (row pair / column) per tile, and the smallest workgroup holding at most
// 12 values per thread.  Explicit descriptor values restrict the choice.
bool lenia::pick_shape(const FftPlan &p, bool rows, int count, int K, uint32_t wantNt,
                       uint32_t wantT, int &ntOut, int &TOut, bool &kparOut) {
    for (int kp = 1; kp >= 0; --kp) {
        for (int nt = 1; nt <= 64; nt *= 2) {
            if (wantNt ? nt != (int)wantNt : nt > 1) continue;
            if (nt > count) break;
            int best = 0;
            for (int T = 64; T <= 1024; T *= 2) {
                if (wantT && T != (int)wantT) continue;
                if (!shape_ok(p, rows, nt, K, T, kp != 0)) continue;
                best = T;
                if ((kp ? nt * K : nt) * p.n <= 12 * T) break;
            }
            if (best) {
                ntOut = nt;
                TOut = best;
                kparOut = kp != 0;
                return true;
            }
        }
    }
    return false;
}

int lenia::wave_default_nw(int count, int L, int numCUs) {
    const int waves = (count + 64 / L - 1) / (64 / L);
    for (int nw = 4; nw > 1; nw /= 2)
        if (waves >= nw * numCUs) return nw;
    return 1;
}

// Executor of one pass.  Asked for, or the library's choice when no LDS tile
// shape is given: the kernel-split wave executor for a column pass whose plan
// has a wave kernel (measured fastest on one world: 1024x512 K=3 columns
// 8.1 us against 10.4 us on the LDS executor), the LDS executor for row
// passes (1024x512: 10.8 us against 12.9 us kernel-split, 16 us one wave per
// row pair).  lenia_world_autotune times every executor.  An explicit wave
// request fixes the waves per workgroup through the thread count.
static int choose_engine(const FftPlan &p, bool rows, int count, uint32_t engine, uint32_t wantNt,
                         uint32_t wantT, int numCUs, bool &wave, int &nw, bool &split) {
    const int L = wave_lanes(p.code);
    const char *pass = rows ? "row" : "column";
    if (engine > LENIA_ENGINE_WAVE_SPLIT)
        return fail(LENIA_ERR_INVALID_ARG, "%s pass: unknown engine %u", pass, engine);
    const bool autoSplit = engine == LENIA_ENGINE_AUTO && !rows && L > 0 && !wantNt && !wantT;
    wave = engine == LENIA_ENGINE_WAVE || engine == LENIA_ENGINE_WAVE_SPLIT || autoSplit;
    split = engine == LENIA_ENGINE_WAVE_SPLIT || autoSplit;
    if (!wave) return LENIA_OK;
    if (!L)
        return fail(LENIA_ERR_UNSUPPORTED, "%s pass: plan 0x%x of length %d has no wave executor", pass,
                    plan_pack(p), p.n);
    if (wantNt)
        return fail(LENIA_ERR_INVALID_ARG, "%s pass: the wave executor derives its tile (got %u per tile)",
                    pass, wantNt);
    if (split) {
        if (wantT && wantT != 64 * kMaxKernels)
            return fail(LENIA_ERR_UNSUPPORTED, "%s pass: the kernel-split wave executor runs %d threads (got %u)",
                        pass, 64 * kMaxKernels, wantT);
        nw = kMaxKernels;
        return LENIA_OK;
    }
    if (wantT && wantT != 64 && wantT != 128 && wantT != 256)
        return fail(LENIA_ERR_UNSUPPORTED, "%s pass: the wave executor runs 64, 128 or 256 threads (got %u)",
                    pass, wantT);
    nw = wantT ? (int)wantT / 64 : wave_default_nw(count, L, numCUs);
    return LENIA_OK;
}

static int choose_tiles_fft(lenia_world *w, int numCUs) {
    const lenia_world_desc &d = w->desc;
    const int W = d.width, H = d.height, Wh = W / 2 + 1, K = (int)d.nKernels;
    if (!choose_plan(W, K, d.fftPlanW, w->planW, true))
        return fail(LENIA_ERR_UNSUPPORTED, "no row-pass plan fits width %d (plan 0x%x)", W, d.fftPlanW);
    int rc = choose_engine(w->planW, true, H / 2, d.rowEngine, d.rowPairsPerTile, d.rowThreads, numCUs,
                           w->rowWave, w->rowNW, w->rowSplit);
    if (rc) return rc;
    // the workgroup-LDS shape (the executor of a wave world's reverse pass too)
    if (!pick_shape(w->planW, true, H / 2, K, w->rowWave ? 0 : d.rowPairsPerTile,
                    w->rowWave ? 0 : d.rowThreads, w->P, w->rowT, w->rowKpar))
        return fail(LENIA_ERR_UNSUPPORTED,
                    "no row-pass shape fits width %d (plan 0x%x, %u row pairs, %u threads)", W,
                    d.fftPlanW, d.rowPairsPerTile, d.rowThreads);
    if (!choose_plan(H, K, d.fftPlanH, w->planH, false))
        return fail(LENIA_ERR_UNSUPPORTED, "no column-pass plan fits height %d (plan 0x%x)", H, d.fftPlanH);
    rc = choose_engine(w->planH, false, Wh, d.colEngine, d.colsPerTile, d.colThreads, numCUs,
                       w->colWave, w->colNW, w->colSplit);
    if (rc) return rc;
    if (!pick_shape(w->planH, false, Wh, K, w->colWave ? 0 : d.colsPerTile,
                    w->colWave ? 0 : d.colThreads, w->C, w->colT, w->colKpar))
        return fail(LENIA_ERR_UNSUPPORTED,
                    "no column-pass shape fits height %d (plan 0x%x, %u columns, %u threads)", H,
                    d.fftPlanH, d.colsPerTile, d.colThreads);
    return LENIA_OK;
}

// Size rules and launch shapes of an FFT-form world (w->desc set): what
// lenia_world_create and the scratch-backed step both resolve.
int lenia::fft_sizes_and_tiles(lenia_world *w, int numCUs) {
    const int W = w->desc.width, H = w->desc.height;
    if (W % 2 || H % 2)
        return fail(LENIA_ERR_UNSUPPORTED, "FFT form needs even width and height (got %dx%d)", W, H);
    int a_, b_, c_;
    if (!factor235(W, a_, b_, c_) || !factor235(H, a_, b_, c_))
        return fail(LENIA_ERR_UNSUPPORTED, "FFT form needs sides of the form 2^a 3^b 5^c (got %dx%d)",
                    W, H);
    if (W > 4096 || H > 4096)
        return fail(LENIA_ERR_UNSUPPORTED, "FFT form supports sides up to 4096 (got %dx%d)", W, H);
    return choose_tiles_fft(w, numCUs);
}

int lenia::world_check_desc(const lenia_world_desc *desc) { return check_desc(desc); }

GrowthParams lenia::world_growth(const lenia_world_desc *desc, uint32_t moduleSeed) {
    return growth_of(desc, moduleSeed);
}

// The device record of an FFT-form world from its host fields (buffer
// pointers included); also sets the tile counts and LDS sizes of its passes.
FftWorldDev lenia::fft_record(lenia_world *w) {
    FftWorldDev f;
    std::memset(&f, 0, sizeof(f));
    f.W = w->desc.width;
    f.H = w->desc.height;
    f.Wh = f.W / 2 + 1;
    const int Lw = wave_lanes(w->planW.code), Lh = wave_lanes(w->planH.code);
    f.P = w->rowWave ? (w->rowSplit ? 1 : w->rowNW) * (64 / Lw) : w->P;
    f.C = w->colWave ? (w->colSplit ? 1 : w->colNW) * (64 / Lh) : w->C;
    f.rowTiles = (f.H / 2 + f.P - 1) / f.P;
    f.colTiles = (f.Wh + f.C - 1) / f.C;
    f.planW = w->planW;
    f.planH = w->planH;
    f.K = (int)w->desc.nKernels;
    f.mP = div_magic(f.P);
    f.mW = div_magic(f.W);
    f.rowThreads = w->rowWave ? 64 * w->rowNW : w->rowT;
    f.colThreads = w->colWave ? 64 * w->colNW : w->colT;
    f.g = w->g;
    f.twW = w->twW;
    f.twH = w->twH;
    f.A = w->A0;
    f.F = w->F;
    f.G = w->G;
    f.S = w->S;
    f.clock = w->clock;
    w->fdev = f;
    w->nRowTiles = f.rowTiles;
    w->nColTiles = f.colTiles;
    if (w->rowWave)
        w->ldsRows = w->rowSplit ? std::max(wave_split_lds_bytes(f.W, Lw, true), wave_lds_bytes(f.W, Lw, 1))
                                 : wave_lds_bytes(f.W, Lw, w->rowNW);
    else
        w->ldsRows = std::max(lds_rows(f.W, f.P, f.K, w->rowKpar, w->rowT), lds_rows(f.W, f.P, 1, true, w->rowT));
    if (w->colWave)
        w->ldsCols = w->colSplit ? wave_split_lds_bytes(f.H, Lh, false) : wave_lds_bytes(f.H, Lh, w->colNW);
    else
        w->ldsCols = lds_cols(f.H, f.C, f.K, w->colKpar);
    return f;
}

int lenia::world_upload_desc(lenia_world *w) {
    lenia_ctx *ctx = w->ctx;
    w->version += 1;
    if (w->isFft) {
        const FftWorldDev f = fft_record(w);
        HIP_TRY(hipMemcpyAsync(w->devDesc, &f, sizeof(f), hipMemcpyHostToDevice, ctx->stream),
                "upload world descriptor");
    } else if (w->isMfma) {
        MfmaWorldDev md;
        std::memset(&md, 0, sizeof(md));
        int tx, ty;
        mfma_tile_shape(w->split(), &tx, &ty);
        md.W = w->desc.width;
        md.H = w->desc.height;
        md.half = w->half;
        md.tilesX = (md.W + tx - 1) / tx;
        md.tilesY = (md.H + ty - 1) / ty;
        md.ksteps = w->ksteps;
        md.tileCols = mfma_tile_cols(w->ksteps);
        md.pitch = md.tileCols + 8;
        md.vec4 = md.W % 4 == 0;
        md.shift = mfma_shift(w->half, md.W);
        md.g = w->g;
        md.toepHi = w->toepHi;
        md.toepMid = w->toepMid;
        md.toepLo = w->toepLo;
        md.A0 = w->A0;
        md.A1 = w->A1;
        md.clock = w->clock;
        w->mdev = md;
        HIP_TRY(hipMemcpyAsync(w->devDesc, &md, sizeof(md), hipMemcpyHostToDevice, ctx->stream),
                "upload world descriptor");
        // no tile table: the kernels take tile j of world 0 (step_tile)
        w->nRowTiles = md.tilesX * md.tilesY;
        w->ldsRows = mfma_lds_bytes(w->half, md.W, w->split());
    } else {
        DirectWorldDev dd;
        std::memset(&dd, 0, sizeof(dd));
        int tx, ty;
        direct_tile_shape((int)w->desc.width, (int)w->desc.height, ctx->numCUs, &tx, &ty);
        dd.W = w->desc.width;
        dd.H = w->desc.height;
        dd.half = w->half;
        dd.tilesX = (dd.W + tx - 1) / tx;
        dd.tilesY = (dd.H + ty - 1) / ty;
        dd.tileH = ty;
        dd.npad = direct_npad(w->half);
        dd.colsMagic = div_magic(tx + dd.npad);
        dd.g = w->g;
        dd.win = w->winF;
        dd.rowChunks = reinterpret_cast<const int2 *>(
            w->winF + (size_t)w->desc.nKernels * (2 * w->half + 1) * dd.npad);
        dd.A0 = w->A0;
        dd.A1 = w->A1;
        dd.clock = w->clock;
        w->ddev = dd;
        HIP_TRY(hipMemcpyAsync(w->devDesc, &dd, sizeof(dd), hipMemcpyHostToDevice, ctx->stream),
                "upload world descriptor");
        w->nRowTiles = dd.tilesX * dd.tilesY;  // no tile table (step_tile)
        w->ldsRows = direct_lds_bytes(w->half, ty, direct_split(w->nRowTiles, ctx->numCUs));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream), "upload world descriptor");
    return LENIA_OK;
}

static void world_release(lenia_world *w) {
    lenia_ctx *ctx = w->ctx;
    world_drop_graph(w);
    world_free_adjoint(w);
    void *ptrs[] = {w->A0, w->A1, w->F, w->G, w->S, w->twW, w->twH, w->winF, w->toepHi, w->toepMid, w->toepLo,
                    w->clock, w->devDesc, w->scratchOut, w->org};
    for (void *p : ptrs) dev_free(ctx, p);
}

extern "C" int lenia_world_create(lenia_ctx *ctx, const lenia_world_desc *desc, lenia_world **out) {
    if (!ctx || !out) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_create: NULL argument");
    *out = nullptr;
    int rc = check_desc(desc);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    lenia_world *w = new lenia_world();
    w->ctx = ctx;
    w->desc = *desc;
    w->isFft = desc->convMode == LENIA_CONV_FFT;
    w->isMfma = desc->convMode == LENIA_CONV_DIRECT_BF16 || desc->convMode == LENIA_CONV_DIRECT_BF16X3;
    const int W = desc->width, H = desc->height;
    if (w->isFft) {
        rc = fft_sizes_and_tiles(w, ctx->numCUs);
        if (rc) {
            delete w;
            return rc;
        }
    }
    if (w->isMfma) {
        int h = 0;
        for (uint32_t k = 0; k < desc->nKernels; ++k) h = std::max(h, kernel_half(desc->kernels[k]));
        const int ks = mfma_ksteps(h, W);
        const size_t lds = mfma_lds_bytes(h, W, w->split());
        if (h > mfma_max_half(w->split()) || ks > mfma_max_ksteps() || lds > kFftLdsMax) {
            delete w;
            return fail(LENIA_ERR_UNSUPPORTED,
                        "MFMA form: window half-width %d exceeds the %s form's limit %d%s", h,
                        w->split() ? "split-BF16" : "BF16", mfma_max_half(w->split()),
                        w->split() ? "; the plain BF16 form or the FP32 forms serve it"
                                   : "; the FP32 forms serve it");
        }
    }
    w->moduleSeed = 0;
    w->g = growth_of(desc, w->moduleSeed);
    const size_t cells = (size_t)W * H;
    const int Wh = W / 2 + 1;
    const size_t K = desc->nKernels;
    rc = dev_alloc(ctx, (void **)&w->A0, cells * sizeof(float));
    if (!rc && !w->isFft) rc = dev_alloc(ctx, (void **)&w->A1, cells * sizeof(float));
    if (!rc) rc = dev_alloc(ctx, (void **)&w->clock, sizeof(StepClock));
    if (!rc)
        rc = dev_alloc(ctx, (void **)&w->devDesc,
                       std::max({sizeof(FftWorldDev), sizeof(DirectWorldDev), sizeof(MfmaWorldDev)}));
    if (!rc) rc = dev_alloc(ctx, (void **)&w->scratchOut, cells * sizeof(float));
    if (w->isFft) {
        if (!rc) rc = dev_alloc(ctx, (void **)&w->F, (size_t)Wh * H * sizeof(float2));
        if (!rc) rc = dev_alloc(ctx, (void **)&w->G, K * Wh * H * sizeof(float2));
        if (!rc) rc = dev_alloc(ctx, (void **)&w->S, K * Wh * H * sizeof(float));
        if (!rc) rc = dev_alloc(ctx, (void **)&w->twW, (size_t)W * sizeof(float2));
        if (!rc) rc = dev_alloc(ctx, (void **)&w->twH, (size_t)H * sizeof(float2));
    }
    // no tile tables: a single world's step kernels take tile j of world 0
    // from the workgroup id (step_tile)
    if (rc) {
        world_release(w);
        delete w;
        return rc;
    }
    hipError_t e = launch_fill(w->A0, cells, 0.0f, ctx->stream);
    if (e == hipSuccess && w->A1) e = launch_fill(w->A1, cells, 0.0f, ctx->stream);
    if (e == hipSuccess) e = launch_set_clock(w->clock, 0, 0, 1, 0, ctx->stream);
    if (e == hipSuccess && w->isFft) {
        // FP64-evaluated length-N tables, then the plans' per-stage tables
        float *raw = nullptr;
        int rc2 = dev_alloc(ctx, (void **)&raw, (size_t)std::max(W, H) * sizeof(float2));
        e = rc2 ? hipErrorOutOfMemory : launch_twiddles(raw, W, ctx->stream);
        if (e == hipSuccess) e = launch_stage_twiddles(w->planW, raw, w->twW, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = launch_twiddles(raw, H, ctx->stream);
        if (e == hipSuccess) e = launch_stage_twiddles(w->planH, raw, w->twH, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        dev_free(ctx, raw);
        if (e == hipSuccess) e = launch_fill(w->G, K * Wh * H * 2, 0.0f, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        rc = hip_fail(e, "world init");
        world_release(w);
        delete w;
        return rc;
    }
    rc = w->isFft ? build_kernels_fft(w) : build_kernels_direct(w);
    if (!rc) rc = world_upload_desc(w);
    if (rc) {
        world_release(w);
        delete w;
        return rc;
    }
    w->needsInit = desc->initMode != LENIA_INIT_NONE;
    w->specValid = false;
    w->frameNext = 0;
    *out = w;
    return LENIA_OK;
}

extern "C" int lenia_world_destroy(lenia_world *w) {
    if (!w) return LENIA_OK;
    (void)hipSetDevice(w->ctx->device);
    (void)hipStreamSynchronize(w->ctx->stream);
    world_release(w);
    delete w;
    return LENIA_OK;
}

extern "C" int lenia_world_set_organism(lenia_world *w, const float *cells, uint32_t orgW,
                                        uint32_t orgH, uint32_t stampCount, uint32_t stampRotate) {
    if (!w || !cells) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_set_organism: NULL argument");
    if (orgW < 1 || orgH < 1 || orgW > 4096 || orgH > 4096)
        return fail(LENIA_ERR_INVALID_ARG, "organism size %ux%u out of range", orgW, orgH);
    const size_t n = (size_t)orgW * orgH;
    for (size_t i = 0; i < n; ++i)
        if (!(cells[i] >= 0.0f && cells[i] <= 1.0f))
            return fail(LENIA_ERR_INVALID_ARG, "organism cell %zu out of [0,1]", i);
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    float *org = nullptr;
    int rc = dev_alloc(ctx, (void **)&org, n * sizeof(float));
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(org, cells, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        dev_free(ctx, org);
        return hip_fail(e, "organism upload");
    }
    dev_free(ctx, w->org);
    w->org = org;
    w->orgW = (int)orgW;
    w->orgH = (int)orgH;
    w->stampCount = std::min(stampCount, 32u);
    w->stampRotate = stampRotate ? 1u : 0u;
    return LENIA_OK;
}

extern "C" int lenia_world_reset(lenia_world *w, uint32_t sceneSeed) {
    if (!w) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_reset: NULL world");
    (void)hipSetDevice(w->ctx->device);
    w->moduleSeed = module_seed(sceneSeed);
    w->g.moduleSeed = w->moduleSeed;
    w->needsInit = w->desc.initMode != LENIA_INIT_NONE;
    return world_upload_desc(w);
}

// init pass (the needsInit_ branch of LeniaModule::encode, Lenia.cpp:468-495)
int lenia::world_apply_init(lenia_world *w) {
    lenia_ctx *ctx = w->ctx;
    const lenia_world_desc &d = w->desc;
    const int W = d.width, H = d.height;
    float *cur = w->current_state();
    hipError_t e = hipSuccess;
    if (d.initMode == LENIA_INIT_PATCH) {
        int x0 = d.initX0, y0 = d.initY0, pw = d.initW, ph = d.initH;
        if (pw == 0 || ph == 0) {
            x0 = y0 = 0;
            pw = W;
            ph = H;
        }
        e = launch_init_patch(cur, W, H, w->moduleSeed, x0, y0, pw, ph, ctx->stream);
    } else if (d.initMode == LENIA_INIT_NOISE) {
        e = launch_init_noise(cur, W, H, w->moduleSeed, d.initCoverage, d.initScale, ctx->stream);
    } else if (d.initMode == LENIA_INIT_STAMP) {
        if (!w->org) return fail(LENIA_ERR_STATE, "stamp init: no organism set (lenia_world_set_organism)");
        e = launch_init_stamp(cur, W, H, w->org, w->orgW, w->orgH, w->stampCount, w->stampRotate,
                              w->moduleSeed, ctx->stream);
    }
    if (e != hipSuccess) return hip_fail(e, "init pass");
    w->needsInit = false;
    w->specValid = false;
    return LENIA_OK;
}

hipError_t lenia::world_launch_rows(lenia_world *w, int mode, const FftWorldDev *rec, uint32_t off,
                                    int kSel, float *out, bool extras, hipStream_t s) {
    extras = extras && mode == 1;
    // no tile table: a single world's tiles are 0, 1, ... of world 0, which
    // the kernels derive from the workgroup id (step_tile)
    const int2 *tiles = nullptr;
    if (w->rowWave) {
        const int vc = wave_class(w->planW.code, w->planW.n);
        if (w->rowSplit && mode == 1)
            return launch_wave_rows_split(extras, vc, rec, tiles, w->nRowTiles, w->ldsRows, off, s);
        return launch_wave_rows(mode, w->rowSplit ? 1 : w->rowNW, vc, extras, rec, tiles,
                                w->nRowTiles, w->ldsRows, off, kSel, out, s);
    }
    // forward-only passes run the serial form, potentials the side-by-side one
    const bool kpar = mode == 1 ? w->rowKpar : mode == 2;
    return launch_fft_rows(mode, w->rowT, extras, kpar, rec, tiles, w->nRowTiles, w->nRowTiles,
                           w->ldsRows, off, kSel, out, s, mode == 1 ? w->row_tile_cells() : 0);
}

hipError_t lenia::world_launch_cols(lenia_world *w, const FftWorldDev *rec, hipStream_t s) {
    const int2 *tiles = nullptr;  // tiles 0, 1, ... of world 0 (step_tile)
    if (w->colWave) {
        const int vc = wave_class(w->planH.code, w->planH.n);
        if (w->colSplit) return launch_wave_cols_split(vc, rec, tiles, w->nColTiles, w->ldsCols, s);
        return launch_wave_cols(w->colNW, vc, rec, tiles, w->nColTiles, w->ldsCols, s);
    }
    return launch_fft_cols(w->colT, w->colKpar, rec, tiles, w->nColTiles, w->nColTiles,
                           w->ldsCols, s, w->col_tile_cells());
}

// F must hold the spectrum of the current state before a row+column step
int lenia::world_ensure_spectrum(lenia_world *w) {
    if (!w->isFft || w->specValid) return LENIA_OK;
    hipError_t e = world_launch_rows(w, 0, (const FftWorldDev *)w->devDesc, 0, 0, nullptr, false,
                                     w->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "forward row pass");
    w->specValid = true;
    return LENIA_OK;
}

static thread_local LaunchEvents tl_launch_events;
LaunchEvents &lenia::launch_events() { return tl_launch_events; }

// one step's kernels, the step `off` steps after the stored clock; with `ev`,
// kernel p of the step records ev[2p] and ev[2p+1] around itself
hipError_t lenia::enqueue_step(lenia_world *w, const hipEvent_t *ev, uint32_t off) {
    lenia_ctx *ctx = w->ctx;
    hipError_t e;
    if (w->isFft) {
        if (ev) launch_events() = {ev[0], ev[1]};
        e = world_launch_cols(w, (const FftWorldDev *)w->devDesc, ctx->stream);
        if (ev) launch_events() = {ev[2], ev[3]};
        if (e == hipSuccess)
            e = world_launch_rows(w, 1, (const FftWorldDev *)w->devDesc, off, 0, nullptr, w->extras(),
                                  ctx->stream);
    } else if (w->isMfma) {
        if (ev) launch_events() = {ev[0], ev[1]};
        e = launch_mfma(0, w->split(), w->ksteps, (const MfmaWorldDev *)w->devDesc, nullptr, w->nRowTiles,
                        w->nRowTiles, w->ldsRows, off, 0, nullptr, ctx->stream);
    } else {
        if (ev) launch_events() = {ev[0], ev[1]};
        e = launch_direct(0, (const DirectWorldDev *)w->devDesc, nullptr, w->nRowTiles,
                          w->nRowTiles, w->ldsRows, off, 0, nullptr, ctx->stream, ctx->numCUs, w->ddev.tileH);
    }
    launch_events() = LaunchEvents();
    return e;
}

hipError_t lenia::world_advance_clock(lenia_world *w, uint32_t by) {
    hipStream_t s = w->ctx->stream;
    if (w->isFft) return launch_advance_clocks((const FftWorldDev *)w->devDesc, 1, by, s);
    if (w->isMfma) return launch_advance_clocks((const MfmaWorldDev *)w->devDesc, 1, by, s);
    return launch_advance_clocks((const DirectWorldDev *)w->devDesc, 1, by, s);
}

// n steps (offsets 0 .. n-1), then the clock moves n steps on
hipError_t lenia::enqueue_steps(lenia_world *w, uint32_t n) {
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < n && e == hipSuccess; ++i) e = enqueue_step(w, nullptr, i);
    if (e == hipSuccess) e = world_advance_clock(w, n);
    return e;
}

void lenia::world_drop_graph(lenia_world *w) {
    if (w->graphExec) (void)hipGraphExecDestroy(w->graphExec);
    w->graphExec = nullptr;
}

// Steps are replayed from a hipGraph of kGraphSteps steps and one clock
// advance, captured once per world: kernels read their parameters and the
// stored clock from device memory, so one graph serves every later replay.
// Leftover steps launch directly.
static constexpr uint32_t kGraphSteps = 32;

static int world_build_graph(lenia_world *w) {
    lenia_ctx *ctx = w->ctx;
    world_drop_graph(w);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "graph capture");
    hipError_t el = enqueue_steps(w, kGraphSteps);
    e = hipStreamEndCapture(ctx->stream, &graph);
    if (el != hipSuccess) e = el;
    if (e == hipSuccess) e = hipGraphInstantiate(&w->graphExec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        w->graphExec = nullptr;
        return hip_fail(e, "graph instantiate");
    }
    w->graphExtras = w->extras();
    return LENIA_OK;
}

int lenia::world_launch_steps(lenia_world *w, uint32_t n) {
    lenia_ctx *ctx = w->ctx;
    int rc = world_ensure_spectrum(w);
    if (rc) return rc;
    if (n >= kGraphSteps) {
        if (!w->graphExec || w->graphExtras != w->extras()) {
            rc = world_build_graph(w);
            if (rc) return rc;
        }
        for (uint32_t r = 0; r < n / kGraphSteps; ++r) {
            hipError_t e = hipGraphLaunch(w->graphExec, ctx->stream);
            if (e != hipSuccess) return hip_fail(e, "graph launch");
            w->hostSteps += kGraphSteps;
        }
        n %= kGraphSteps;
    }
    if (n) {
        hipError_t e = enqueue_steps(w, n);
        if (e != hipSuccess) return hip_fail(e, "step launch");
        w->hostSteps += n;
    }
    return LENIA_OK;
}

extern "C" int lenia_world_frame(lenia_world *w, uint32_t frameIndex, uint32_t substeps) {
    if (!w) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_frame: NULL world");
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    if (w->needsInit) {
        int rc = world_apply_init(w);
        if (rc) return rc;
    }
    uint32_t steps = std::max(1u, substeps);
    hipError_t e = launch_set_clock(w->clock, frameIndex, 0, steps, w->parity(), ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "set clock");
    w->frameNext = frameIndex + 1;
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    int rc = world_launch_steps(w, steps);
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    ctx->timed = true;
    ctx->last0 = ctx->ev0;
    ctx->last1 = ctx->ev1;
    return rc;
}

extern "C" int lenia_world_step(lenia_world *w, uint32_t nSteps) {
    if (!w) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_step: NULL world");
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    if (w->needsInit) {
        int rc = world_apply_init(w);
        if (rc) return rc;
        hipError_t e = launch_set_clock(w->clock, w->frameNext, 0, 1, w->parity(), ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "set clock");
    }
    w->frameNext += nSteps;
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    int rc = world_launch_steps(w, nSteps);
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    ctx->timed = true;
    ctx->last0 = ctx->ev0;
    ctx->last1 = ctx->ev1;
    return rc;
}

extern "C" int lenia_world_set_params(lenia_world *w, const lenia_world_desc *desc) {
    if (!w) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_set_params: NULL world");
    int rc = check_desc(desc);
    if (rc) return rc;
    if (desc->width != w->desc.width || desc->height != w->desc.height ||
        desc->convMode != w->desc.convMode || desc->nKernels != w->desc.nKernels)
        return fail(LENIA_ERR_INVALID_ARG, "set_params cannot change size, convMode or kernel count");
    (void)hipSetDevice(w->ctx->device);
    bool shapes = false;
    for (uint32_t k = 0; k < desc->nKernels; ++k) {
        const lenia_kernel_desc &a = desc->kernels[k], &b = w->desc.kernels[k];
        if (a.shape != b.shape || a.core != b.core || a.radius != b.radius ||
            a.shellMu != b.shellMu || a.shellSigma != b.shellSigma || a.nBetas != b.nBetas ||
            std::memcmp(a.betas, b.betas, sizeof(a.betas)) != 0)
            shapes = true;
    }
    if (w->isFft) {
        for (uint32_t k = 0; k < desc->nKernels; ++k)
            if (kernel_half(desc->kernels[k]) > std::min(desc->width, desc->height) / 2)
                return fail(LENIA_ERR_INVALID_ARG, "kernel window too large");
    } else if (shapes) {
        int h = 0;
        for (uint32_t k = 0; k < desc->nKernels; ++k) h = std::max(h, kernel_half(desc->kernels[k]));
        if (h != w->half)
            return fail(LENIA_ERR_INVALID_ARG, "set_params cannot change the direct window size");
    }
    uint32_t keepPW = w->desc.fftPlanW, keepPH = w->desc.fftPlanH;
    uint32_t keepP = w->desc.rowPairsPerTile, keepC = w->desc.colsPerTile;
    uint32_t keepRT = w->desc.rowThreads, keepCT = w->desc.colThreads;
    uint32_t keepRE = w->desc.rowEngine, keepCE = w->desc.colEngine;
    w->desc = *desc;
    w->desc.rowEngine = keepRE;
    w->desc.colEngine = keepCE;
    w->desc.fftPlanW = keepPW;
    w->desc.fftPlanH = keepPH;
    w->desc.rowPairsPerTile = keepP;
    w->desc.colsPerTile = keepC;
    w->desc.rowThreads = keepRT;
    w->desc.colThreads = keepCT;
    w->g = growth_of(desc, w->moduleSeed);
    world_drop_graph(w);
    if (shapes) world_invalidate_adjoint(w);
    if (shapes) {
        rc = w->isFft ? build_kernels_fft(w) : build_kernels_direct(w);
        if (rc) return rc;
    }
    return world_upload_desc(w);
}

extern "C" int lenia_world_read_state(lenia_world *w, float *host, size_t count) {
    if (!w || !host) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_read_state: NULL argument");
    size_t cells = (size_t)w->desc.width * w->desc.height;
    if (count != cells)
        return fail(LENIA_ERR_INVALID_ARG, "read_state: count %zu != width*height %zu", count, cells);
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    HIP_TRY(hipMemcpyAsync(host, w->current_state(), cells * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream), "read state");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "read state");
    return LENIA_OK;
}

extern "C" int lenia_world_write_state(lenia_world *w, const float *host, size_t count) {
    if (!w || !host) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_write_state: NULL argument");
    size_t cells = (size_t)w->desc.width * w->desc.height;
    if (count != cells)
        return fail(LENIA_ERR_INVALID_ARG, "write_state: count %zu != width*height %zu", count, cells);
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    HIP_TRY(hipMemcpyAsync(w->current_state(), host, cells * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream), "write state");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "write state");
    w->needsInit = false;
    w->specValid = false;
    return LENIA_OK;
}

extern "C" int lenia_world_state_device(lenia_world *w, float **dev) {
    if (!w || !dev) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_state_device: NULL argument");
    *dev = w->current_state();
    return LENIA_OK;
}

extern "C" int lenia_world_potential(lenia_world *w, uint32_t k, float *host, size_t count) {
    if (!w || !host) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_potential: NULL argument");
    if (k >= w->desc.nKernels) return fail(LENIA_ERR_INVALID_ARG, "kernel index %u out of range", k);
    size_t cells = (size_t)w->desc.width * w->desc.height;
    if (count != cells) return fail(LENIA_ERR_INVALID_ARG, "potential: count != width*height");
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    hipError_t e;
    if (w->isFft) {
        int rc = world_ensure_spectrum(w);
        if (rc) return rc;
        e = world_launch_cols(w, (const FftWorldDev *)w->devDesc, ctx->stream);
        if (e == hipSuccess)
            e = world_launch_rows(w, 2, (const FftWorldDev *)w->devDesc, 0, (int)k, w->scratchOut, false,
                                  ctx->stream);
    } else if (w->isMfma) {
        e = launch_mfma(1, w->split(), w->ksteps, (const MfmaWorldDev *)w->devDesc, nullptr, w->nRowTiles,
                        w->nRowTiles, w->ldsRows, 0, (int)k, w->scratchOut, ctx->stream);
    } else {
        e = launch_direct(1, (const DirectWorldDev *)w->devDesc, nullptr, w->nRowTiles,
                          w->nRowTiles, w->ldsRows, 0, (int)k, w->scratchOut, ctx->stream, ctx->numCUs,
                          w->ddev.tileH);
    }
    if (e != hipSuccess) return hip_fail(e, "potential launch");
    HIP_TRY(hipMemcpyAsync(host, w->scratchOut, cells * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream), "read potential");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "read potential");
    return LENIA_OK;
}

extern "C" int lenia_world_get_desc(lenia_world *w, lenia_world_desc *out) {
    if (!w || !out) return fail(LENIA_ERR_INVALID_ARG, "lenia_world_get_desc: NULL argument");
    *out = w->desc;
    if (w->isFft) {
        out->fftPlanW = plan_pack(w->planW);
        out->fftPlanH = plan_pack(w->planH);
        out->rowPairsPerTile = w->P;
        out->colsPerTile = w->C;
        out->rowThreads = w->rowT;
        out->colThreads = w->colT;
        out->rowEngine = w->rowWave ? (w->rowSplit ? LENIA_ENGINE_WAVE_SPLIT : LENIA_ENGINE_WAVE) : LENIA_ENGINE_LDS;
        out->colEngine = w->colWave ? (w->colSplit ? LENIA_ENGINE_WAVE_SPLIT : LENIA_ENGINE_WAVE) : LENIA_ENGINE_LDS;
        if (w->rowWave) {
            out->rowPairsPerTile = 0;
            out->rowThreads = 64 * w->rowNW;
        }
        if (w->colWave) {
            out->colsPerTile = 0;
            out->colThreads = 64 * w->colNW;
        }
    }
    return LENIA_OK;
}

// pcgHash / deriveSeed (LifeCore/Math/Random.h:34-44) on the host
uint32_t lenia::pcg_hash_host(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

uint32_t lenia::module_seed(uint32_t sceneSeed) {
    return pcg_hash_host(sceneSeed ^ pcg_hash_host(0x4C454E31u));  // "LEN1"
}

// per-pass GPU timing of nSteps steps (GPUTimer per-pass timestamps,
// LifeCore/Metal/GPUTimer.mm:28-94, averaged the way LifeBench reports them,
// Apps/LifeBench/main.cpp:112-133).  Every kernel of every timed step carries
// its own start/stop timestamps (hipExtLaunchKernelGGL), so each figure is the
// mean duration of that kernel alone, launch gaps excluded.
extern "C" int lenia_world_pass_times(lenia_world *w, uint32_t nSteps, double *outMs,
                                      uint32_t maxPasses, uint32_t *nPasses) {
    if (!w || !outMs || !nPasses || nSteps == 0)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_world_pass_times: bad argument");
    lenia_ctx *ctx = w->ctx;
    (void)hipSetDevice(ctx->device);
    const uint32_t np = w->isFft ? 2u : 1u;
    if (maxPasses < np) return fail(LENIA_ERR_INVALID_ARG, "pass_times: need room for %u passes", np);
    if (w->needsInit) {
        int rc = world_apply_init(w);
        if (rc) return rc;
    }
    int rc = world_ensure_spectrum(w);
    if (rc) return rc;
    hipError_t e = time_steps(ctx, nSteps, np, [&](const hipEvent_t *ev) {
        hipError_t es = enqueue_step(w, ev, 0);
        if (es == hipSuccess) es = world_advance_clock(w, 1);
        w->hostSteps += 1;
        w->frameNext += 1;
        return es;
    }, outMs);
    if (e != hipSuccess) return hip_fail(e, "pass timing");
    *nPasses = np;
    return LENIA_OK;
}
