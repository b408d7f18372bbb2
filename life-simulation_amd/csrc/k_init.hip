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

hipError_t launch_set_clock(StepClock *c, uint32_t frame, uint32_t sub, uint32_t substeps,
                            uint32_t parity, hipStream_t s) {
    hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(1), 0, s, c, frame, sub, substeps, parity);
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
