// Init passes and small utility kernels.
#include <hip/hip_runtime.h>

#include "lenia_device.h"
#include "launch.h"

namespace lenia {

constexpr uint32_t kTagPatch = 1009u;

// seeded uniform[0,1) noise in a rectangle, zero elsewhere: rand01 of the
// cell's global coordinates (Shaders/Common/Random.metal:16-19)
__global__ void k_init_patch(float *A, int W, int H, uint32_t seed, int x0, int y0, int pw,
                             int ph) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W || y >= H) return;
    bool in = x >= x0 && x < x0 + pw && y >= y0 && y < y0 + ph;
    A[(size_t)y * W + x] = in ? rand01(x, y, kTagPatch, seed) : 0.0f;
}

// leniaValueNoise / leniaInit (Shaders/Field/Lenia.metal:32-61), FP32 as on
// the reference GPU
__device__ float value_noise(float px, float py, uint32_t tag, uint32_t seed) {
    float fx0 = floorf(px), fy0 = floorf(py);
    float ix = fx0 + 512.0f, iy = fy0 + 512.0f;
    float fx = px - fx0, fy = py - fy0;
    fx = fx * fx * (3.0f - 2.0f * fx);
    fy = fy * fy * (3.0f - 2.0f * fy);
    uint32_t ux = (uint32_t)ix, uy = (uint32_t)iy;
    float a = rand01(ux, uy, tag, seed);
    float b = rand01(ux + 1u, uy, tag, seed);
    float c = rand01(ux, uy + 1u, tag, seed);
    float d = rand01(ux + 1u, uy + 1u, tag, seed);
    float ab = a + (b - a) * fx, cd = c + (d - c) * fx;
    return ab + (cd - ab) * fy;
}

__global__ void k_init_noise(float *A, int W, int H, uint32_t seed, float cov, float scale) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W || y >= H) return;
    float s = fmaxf(scale, 2.0f);
    float px = float(x) / s, py = float(y) / s;
    float v = 0.55f * value_noise(px, py, 11u, seed) +
              0.30f * value_noise(px * 2.13f, py * 2.13f, 23u, seed) +
              0.15f * value_noise(px * 4.31f, py * 4.31f, 37u, seed);
    v = fminf(fmaxf((v - (1.0f - cov)) / fmaxf(cov, 1e-3f), 0.0f), 1.0f);
    A[(size_t)y * W + x] = v * 0.9f;
}

// organism stamps (leniaStampInit, Shaders/Field/Lenia.metal:242-303): up to
// 32 copies of the organism's cells at hashed centres, each in one of the 8
// dihedral orientations when rotation is on; a cell takes the largest stamp
// value covering it (nearest sampling, toroidal offsets)
__device__ __forceinline__ float2 stamp_inverse(float2 d, uint32_t variant) {
    if (variant & 4u) d.x = -d.x;
    const uint32_t rot = variant & 3u;
    if (rot == 1u) d = make_float2(d.y, -d.x);
    else if (rot == 2u) d = make_float2(-d.x, -d.y);
    else if (rot == 3u) d = make_float2(-d.y, d.x);
    return d;
}

__global__ void k_init_stamp(float *A, int W, int H, const float *__restrict__ org, int ow, int oh,
                             uint32_t count, uint32_t rotate, uint32_t seed) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    float v = 0.0f;
    const float cx0 = (float)ow * 0.5f, cy0 = (float)oh * 0.5f;
    const uint32_t n = count < 32u ? count : 32u;
    for (uint32_t i = 0; i < n; ++i) {
        const float cx = rand01(i, 0u, 907u, seed) * (float)W;
        const float cy = rand01(0u, i, 911u, seed) * (float)H;
        uint32_t variant = 0u;
        if (rotate) {
            const uint32_t r = (uint32_t)(rand01(i, 1u, 919u, seed) * 8.0f);
            variant = r < 7u ? r : 7u;
        }
        float dx = (float)x + 0.5f - cx, dy = (float)y + 0.5f - cy;
        dx -= (float)W * metal_round(dx / (float)W);
        dy -= (float)H * metal_round(dy / (float)H);
        const float2 l = stamp_inverse(make_float2(dx, dy), variant);
        const int lx = (int)floorf(l.x + cx0), ly = (int)floorf(l.y + cy0);
        if (lx >= 0 && lx < ow && ly >= 0 && ly < oh) v = fmaxf(v, org[(size_t)ly * ow + lx]);
    }
    A[(size_t)y * W + x] = fminf(fmaxf(v, 0.0f), 1.0f);
}

__global__ void k_set_clock(StepClock *c, uint32_t frame, uint32_t sub, uint32_t substeps,
                            uint32_t parity) {
    StepClock v;
    v.frame = frame;
    v.sub = sub;
    v.substeps = substeps;
    v.parity = parity;
    v.steps = 0u;
    v.pad[0] = v.pad[1] = v.pad[2] = 0u;
    *c = v;
}

template <class WorldT>
__global__ void k_advance_clocks(const WorldT *worlds, int n, uint32_t by) {
    advance_clocks(worlds, n, by);
}

__global__ void k_fill(float *p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_init_patch(float *A, int W, int H, uint32_t seed, int x0, int y0, int pw,
                             int ph, hipStream_t s) {
    hipLaunchKernelGGL(k_init_patch, dim3((W + 255) / 256, H), dim3(256), 0, s, A, W, H, seed, x0,
                       y0, pw, ph);
    return hipGetLastError();
}

hipError_t launch_init_noise(float *A, int W, int H, uint32_t seed, float coverage, float scale,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_init_noise, dim3((W + 255) / 256, H), dim3(256), 0, s, A, W, H, seed,
                       coverage, scale);
    return hipGetLastError();
}

hipError_t launch_init_stamp(float *A, int W, int H, const float *org, int ow, int oh,
                             uint32_t count, uint32_t rotate, uint32_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_init_stamp, dim3((W + 255) / 256, H), dim3(256), 0, s, A, W, H, org, ow, oh,
                       count, rotate, seed);
    return hipGetLastError();
}

hipError_t launch_set_clock(StepClock *c, uint32_t frame, uint32_t sub, uint32_t substeps,
                            uint32_t parity, hipStream_t s) {
    hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(1), 0, s, c, frame, sub, substeps, parity);
    return hipGetLastError();
}

// every world's clock at the start of frame `frame` of `substeps` substeps;
// the ping-pong parity stays
template <class WorldT>
__global__ void k_frame_clocks(const WorldT *worlds, int n, uint32_t frame, uint32_t substeps) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        StepClock *c = worlds[i].clock;
        c->frame = frame;
        c->sub = 0u;
        c->substeps = substeps;
        c->steps = 0u;
    }
}

hipError_t launch_frame_clocks(const FftWorldDev *worlds, int n, uint32_t frame, uint32_t substeps,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_frame_clocks<FftWorldDev>, dim3(1), dim3(256), 0, s, worlds, n, frame, substeps);
    return hipGetLastError();
}
hipError_t launch_frame_clocks(const DirectWorldDev *worlds, int n, uint32_t frame, uint32_t substeps,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_frame_clocks<DirectWorldDev>, dim3(1), dim3(256), 0, s, worlds, n, frame,
                       substeps);
    return hipGetLastError();
}

template <class WorldT>
static hipError_t advance_launch(const WorldT *worlds, int n, uint32_t by, hipStream_t s) {
    if (n <= 0 || by == 0) return hipSuccess;
    hipLaunchKernelGGL(k_advance_clocks<WorldT>, dim3(1), dim3(256), 0, s, worlds, n, by);
    return hipGetLastError();
}
hipError_t launch_advance_clocks(const FftWorldDev *worlds, int n, uint32_t by, hipStream_t s) {
    return advance_launch(worlds, n, by, s);
}
hipError_t launch_advance_clocks(const DirectWorldDev *worlds, int n, uint32_t by, hipStream_t s) {
    return advance_launch(worlds, n, by, s);
}
hipError_t launch_advance_clocks(const MfmaWorldDev *worlds, int n, uint32_t by, hipStream_t s) {
    return advance_launch(worlds, n, by, s);
}

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace lenia
