// Field -> RGBA16F colour layer with the cosine palette of the reference's
// colour-map pass (Shaders/Render/ColorMap.metal:17-35 colorMapField, :43-75
// colorMapFieldScaled; parameters LifeCore/Render/ColorMapPass.h:16-25):
//   t   = clamp(v * inputScale + inputBias, 0, 1)
//   rgb = max(a + b cos(2 pi (c t + d)), 0) * exposure,  alpha = t
// where v is the selected channel of the R32F field texel: channel 0 is the
// field, 1 and 2 read 0 and 3 reads 1 (an R32F read returns (r, 0, 0, 1)).
// When the output size differs from the field size the field is sampled
// bilinearly at the output pixel centre with wrap-around addressing, as the
// reference's linear/repeat sampler does.
//
// One thread per output pixel; a pixel is one 8-byte store, so the pass is
// bound by HBM bandwidth: 4 B read (more with bilinear taps from L2) + 8 B
// written per pixel.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace lenia {

__device__ __forceinline__ uint16_t half_bits(float x) {
    _Float16 h = (_Float16)x;  // round to nearest even
    return __builtin_bit_cast(uint16_t, h);
}

template <bool SCALED>
__global__ __launch_bounds__(256) void k_colormap(const float *__restrict__ field, int srcW,
                                                  int srcH, ColorMapParams p,
                                                  uint2 *__restrict__ out, int dstW, int dstH) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= dstW || y >= dstH) return;
    float v;
    if (!SCALED) {
        v = field[(size_t)y * srcW + x];
    } else {
        // normalized-coordinate sample at the pixel centre: texel space
        // position u*W - 0.5, two taps per axis, repeat addressing
        const float u = ((float)x + 0.5f) / (float)dstW;
        const float w = ((float)y + 0.5f) / (float)dstH;
        const float sx = u * (float)srcW - 0.5f, sy = w * (float)srcH - 0.5f;
        const float fx0 = floorf(sx), fy0 = floorf(sy);
        const float ax = sx - fx0, ay = sy - fy0;
        int x0 = (int)fx0 % srcW, y0 = (int)fy0 % srcH;
        if (x0 < 0) x0 += srcW;
        if (y0 < 0) y0 += srcH;
        const int x1 = x0 + 1 == srcW ? 0 : x0 + 1, y1 = y0 + 1 == srcH ? 0 : y0 + 1;
        const float v00 = field[(size_t)y0 * srcW + x0], v01 = field[(size_t)y0 * srcW + x1];
        const float v10 = field[(size_t)y1 * srcW + x0], v11 = field[(size_t)y1 * srcW + x1];
        const float top = v00 + ax * (v01 - v00), bot = v10 + ax * (v11 - v10);
        v = top + ay * (bot - top);
    }
    const float sel = p.channel == 0u ? v : (p.channel >= 3u ? 1.0f : 0.0f);
    const float t = fminf(fmaxf(sel * p.inputScale + p.inputBias, 0.0f), 1.0f);
    float rgb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float e = p.a[c] + p.b[c] * cosf(6.28318530718f * (p.c[c] * t + p.d[c]));
        rgb[c] = fmaxf(e, 0.0f) * p.exposure;
    }
    uint2 px;
    px.x = (uint32_t)half_bits(rgb[0]) | ((uint32_t)half_bits(rgb[1]) << 16);
    px.y = (uint32_t)half_bits(rgb[2]) | ((uint32_t)half_bits(t) << 16);
    out[(size_t)y * dstW + x] = px;
}

// Activity of a colour layer (Apps/LifePresetLab/main.cpp:67-86): Rec.709
// luma of every pixel, summed with its square in FP64.  A fixed grid and a
// fixed reduction order make the result the same on every run.
__device__ __forceinline__ double luma_of(uint2 p) {
    const float r = (float)__builtin_bit_cast(_Float16, (uint16_t)(p.x & 0xFFFFu));
    const float g = (float)__builtin_bit_cast(_Float16, (uint16_t)(p.x >> 16));
    const float b = (float)__builtin_bit_cast(_Float16, (uint16_t)(p.y & 0xFFFFu));
    return 0.2126 * (double)r + 0.7152 * (double)g + 0.0722 * (double)b;
}

constexpr int kLumaT = 256;

__device__ __forceinline__ void block_sum2(double &s, double &s2, double *red) {
    const int tid = threadIdx.x;
    red[tid] = s;
    red[kLumaT + tid] = s2;
    __syncthreads();
    for (int h = kLumaT / 2; h > 0; h >>= 1) {
        if (tid < h) {
            red[tid] += red[tid + h];
            red[kLumaT + tid] += red[kLumaT + tid + h];
        }
        __syncthreads();
    }
    s = red[0];
    s2 = red[kLumaT];
}

__global__ __launch_bounds__(kLumaT) void k_luma_partial(const uint2 *__restrict__ px, size_t n,
                                                         double *__restrict__ partial) {
    __shared__ double red[2 * kLumaT];
    double s = 0.0, s2 = 0.0;
    for (size_t i = (size_t)blockIdx.x * kLumaT + threadIdx.x; i < n; i += (size_t)gridDim.x * kLumaT) {
        const double l = luma_of(px[i]);
        s += l;
        s2 += l * l;
    }
    block_sum2(s, s2, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s;
        partial[2 * blockIdx.x + 1] = s2;
    }
}

__global__ __launch_bounds__(kLumaT) void k_luma_final(const double *__restrict__ partial, int nb,
                                                       double *__restrict__ out) {
    __shared__ double red[2 * kLumaT];
    double s = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < nb; i += kLumaT) {
        s += partial[2 * i];
        s2 += partial[2 * i + 1];
    }
    block_sum2(s, s2, red);
    if (threadIdx.x == 0) {
        out[0] = s;
        out[1] = s2;
    }
}

int luma_stats_blocks(size_t n) {
    size_t b = (n + kLumaT - 1) / kLumaT;
    return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

hipError_t launch_luma_stats(const uint16_t *rgba, size_t n, double *partial, double *out,
                             hipStream_t s) {
    const int nb = luma_stats_blocks(n);
    hipLaunchKernelGGL(k_luma_partial, dim3(nb), dim3(kLumaT), 0, s,
                       reinterpret_cast<const uint2 *>(rgba), n, partial);
    hipLaunchKernelGGL(k_luma_final, dim3(1), dim3(kLumaT), 0, s, partial, nb, out);
    return hipGetLastError();
}

hipError_t launch_colormap(const float *field, int srcW, int srcH, const ColorMapParams &p,
                           uint16_t *rgba, int dstW, int dstH, hipStream_t s) {
    dim3 grid((dstW + 255) / 256, dstH);
    uint2 *out = reinterpret_cast<uint2 *>(rgba);
    if (srcW == dstW && srcH == dstH)
        hipLaunchKernelGGL(k_colormap<false>, grid, dim3(256), 0, s, field, srcW, srcH, p, out, dstW, dstH);
    else
        hipLaunchKernelGGL(k_colormap<true>, grid, dim3(256), 0, s, field, srcW, srcH, p, out, dstW, dstH);
    return hipGetLastError();
}

}  // namespace lenia
