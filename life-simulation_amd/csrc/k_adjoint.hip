// Adjoint (reverse-mode) kernels of the FFT-form Lenia step.
//
// Forward step (growth mode 0, no noise / jitter / inject terms):
//   u_k = K_k (x) A,  b_k = exp(-(u_k - mu_k)^2 / (2 sigma_k^2)),
//   pre = A + dt sum_k w_k (2 b_k - 1),  A' = clip(pre, 0, 1)
// Reverse, for an incoming gradient g = dL/dA':
//   gp   = g [0 < pre < 1]
//   gu_k = -gp dt w_k 2 b_k (u_k - mu_k) / sigma_k^2
//   dL/dA = gp + sum_k K_k (x) gu_k          (K_k even: the adjoint of the
//                                             gather is the same operator)
//   dL/dmu_k, dL/dsigma_k, dL/dw_k, dL/ddt: sums over cells of the matching
//   partial derivatives
//   dL/dbeta_kj: a spectral dot product  sum_f D_kj(f) Re(conj(gu_k^(f)) A^(f))
//   (Parseval), with D_kj = (C_kj^ - C_kj S_k^) / S_k from the shell spectra
// In the spectral domain the adjoint convolution is the conjugate-spectrum
// product: multiply by the (real, even) kernel spectrum and transform back.
//
// These kernels reuse the LDS FFT of the forward path (same plans, same
// butterflies).  Partial sums are written per workgroup and summed in a fixed
// order, so gradients are identical run to run.
#include <hip/hip_runtime.h>

#include "fft_device.h"
#include "lenia_device.h"
#include "launch.h"

namespace lenia {

constexpr int kAV = 16;  // values per thread per FFT stage

// ------------------------------------------------------------------ rows: forward R2C
// nf real fields (field f at in + f*fieldStride) -> half spectra
// out[f][kx][y] (complex, column-major), one row pair per workgroup
template <int T>
__global__ __launch_bounds__(T) void k_adj_rows_fwd(AdjArgs a, const float *in, size_t fieldStride,
                                                    float *out, int nf) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int W = a.W, H = a.H, Wh = W / 2 + 1, H2 = H / 2;
    const int rp = blockIdx.x;  // row pair
    float2 *tw = smem;
    float2 *Z = smem + W;
    for (int t = tid; t < W; t += T) tw[t] = gld2(gp(reinterpret_cast<const float2 *>(a.twW)) + t);
    for (int q = tid; q < nf * W; q += T) {
        int f = q / W, x = q - f * W;
        const LENIA_GLOBAL float *src = gp(in) + f * fieldStride + (size_t)(2 * rp) * W;
        Z[sw(q)] = make_float2(src[x], src[x + W]);
    }
    __syncthreads();
    fft_lds<false, kAV, T>(Z, a.planW, nf, tw, tid);
    for (int m = tid; m < nf * Wh; m += T) {
        int f = m / Wh, kx = m - f * Wh;
        float2 z1 = Z[sw(f * W + kx)];
        float2 z2 = Z[sw(f * W + (kx == 0 ? 0 : W - kx))];
        LENIA_GLOBAL float4 *o = gp(reinterpret_cast<float4 *>(out)) + (size_t)f * Wh * H2;
        gst4(o + kx * H2 + rp, make_float4(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y),
                                           0.5f * (z1.y + z2.y), 0.5f * (z2.x - z1.x)));
    }
}

// ------------------------------------------------------------------ rows: inverse C2R
// ns half-spectrum slots G[s][kx][y] -> real fields out[s][y][x]; with addend,
// out = addend + field (one slot)
template <int T>
__global__ __launch_bounds__(T) void k_adj_rows_inv(AdjArgs a, const float *G, float *out, int ns,
                                                    const float *addend) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int W = a.W, H = a.H, Wh = W / 2 + 1, H2 = H / 2;
    const int rp = blockIdx.x;
    float2 *tw = smem;
    float2 *Z = smem + W;
    for (int t = tid; t < W; t += T) tw[t] = gld2(gp(reinterpret_cast<const float2 *>(a.twW)) + t);
    for (int m = tid; m < ns * Wh; m += T) {
        int s = m / Wh, kx = m - s * Wh;
        float4 g4 = gld4(gp(reinterpret_cast<const float4 *>(G)) + (size_t)s * Wh * H2 + kx * H2 + rp);
        if (kx == 0 || kx == W / 2) {
            g4.y = 0.0f;
            g4.w = 0.0f;
        }
        Z[sw(s * W + kx)] = make_float2(g4.x - g4.w, g4.y + g4.z);
        if (kx > 0 && kx < W / 2) Z[sw(s * W + W - kx)] = make_float2(g4.x + g4.w, g4.z - g4.y);
    }
    __syncthreads();
    fft_lds<true, kAV, T>(Z, a.planW, ns, tw, tid);
    for (int q = tid; q < ns * W; q += T) {
        int s = q / W, x = q - s * W;
        float2 z = Z[sw(q)];
        size_t ia = (size_t)s * W * H + (size_t)(2 * rp) * W + x;
        if (addend) {
            z.x += gp(addend)[ia];
            z.y += gp(addend)[ia + W];
        }
        gp(out)[ia] = z.x;
        gp(out)[ia + W] = z.y;
    }
}

// ------------------------------------------------------------------ columns
// POT: F -> (store the forward column spectrum Ahat) -> x S_k -> inverse -> G[k]
// ADJ: FG[k] (k < K) -> forward -> spectral partial sums for the shell
//      gradients against Ahat -> Y = sum_k S_k FG^[k] -> inverse -> G[0]
template <int T, bool ADJ>
__global__ __launch_bounds__(T) void k_adj_cols(AdjArgs a, const float *Fin, float *Ahat,
                                                float *G, float *betaPart) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int H = a.H, Wh = a.W / 2 + 1, K = a.K;
    const int c = blockIdx.x;  // spectrum column kx
    const size_t sStride = (size_t)Wh * H;
    float2 *tw = smem;
    float2 *X = smem + H;  // ADJ: K slots; POT: 1 slot + K slots
    for (int t = tid; t < H; t += T) tw[t] = gld2(gp(reinterpret_cast<const float2 *>(a.twH)) + t);
    const int nin = ADJ ? K : 1;
    for (int e = tid; e < nin * H; e += T) {
        int s = e / H, ky = e - s * H;
        X[sw(e)] = gld2(gp(reinterpret_cast<const float2 *>(Fin)) + s * sStride + (size_t)c * H + ky);
    }
    __syncthreads();
    fft_lds<false, kAV, T>(X, a.planH, nin, tw, tid);
    float2 *Y = X + ((nin * H + 31) & ~31);
    if (!ADJ) {
        for (int e = tid; e < H; e += T)
            gst2(gp(reinterpret_cast<float2 *>(Ahat)) + (size_t)c * H + e, X[sw(e)]);
        for (int e = tid; e < K * H; e += T) {
            int k = e / H, ky = e - k * H;
            Y[sw(e)] = cscale(X[sw(ky)], gp(a.S)[k * sStride + (size_t)c * H + ky]);
        }
        __syncthreads();
        fft_lds<true, kAV, T>(Y, a.planH, K, tw, tid);
        for (int e = tid; e < K * H; e += T) {
            int k = e / H, ky = e - k * H;
            gst2(gp(reinterpret_cast<float2 *>(G)) + k * sStride + (size_t)c * H + ky, Y[sw(e)]);
        }
        return;
    }
    // shell partial sums: thread-strided over ky, then a fixed-order LDS tree
    const float wt = (c == 0 || c == Wh - 1) ? 1.0f : 2.0f;
    const int nb = a.nShells;  // total (k, j) pairs, <= 16
    float acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    for (int ky = tid; ky < H; ky += T) {
        float2 ah = gld2(gp(reinterpret_cast<const float2 *>(Ahat)) + (size_t)c * H + ky);
        for (int i = 0; i < nb; ++i) {
            int k = a.shellK[i];
            float2 gu = X[sw(k * H + ky)];
            float re = gu.x * ah.x + gu.y * ah.y;  // Re(conj(gu) * A)
            acc[i] += gp(a.D)[(size_t)i * sStride + (size_t)c * H + ky] * re;
        }
    }
    // Y = sum_k S_k * gu_k^
    for (int e = tid; e < H; e += T) {
        float2 y = make_float2(0.f, 0.f);
        for (int k = 0; k < K; ++k) {
            float s = gp(a.S)[k * sStride + (size_t)c * H + e];
            float2 x = X[sw(k * H + e)];
            y.x = fmaf(s, x.x, y.x);
            y.y = fmaf(s, x.y, y.y);
        }
        Y[sw(e)] = y;
    }
    // reduce the partial sums across the workgroup (reuse the tail of LDS)
    float *red = reinterpret_cast<float *>(Y + ((H + 31) & ~31));
    for (int i = 0; i < nb; ++i) {
        red[tid] = acc[i] * wt;
        __syncthreads();
        for (int s = T / 2; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) gp(betaPart)[(size_t)c * 16 + i] = red[0];
        __syncthreads();
    }
    fft_lds<true, kAV, T>(Y, a.planH, 1, tw, tid);
    for (int e = tid; e < H; e += T) gst2(gp(reinterpret_cast<float2 *>(G)) + (size_t)c * H + e, Y[sw(e)]);
}

// ------------------------------------------------------------------ cells
// per cell: gp, gu_k and the cell terms of the parameter gradients; block
// partial sums in fixed tree order: part[block][0] = dt, then per k: mu,
// sigma, weight at 1 + 3k + {0,1,2}
__global__ __launch_bounds__(256) void k_adj_cells(AdjArgs a, const float *A, const float *U,
                                                   const float *g, float *gpOut, float *GU,
                                                   float *part) {
    __shared__ float red[256];
    const int tid = threadIdx.x;
    const size_t n = (size_t)a.W * a.H;
    const size_t i = (size_t)blockIdx.x * 256 + tid;
    float vals[13];
#pragma unroll
    for (int v = 0; v < 13; ++v) vals[v] = 0.0f;
    if (i < n) {
        float A0 = gp(A)[i];
        float b[4], u[4], sumStd = 0.0f;
        for (int k = 0; k < a.K; ++k) {
            u[k] = gp(U)[k * n + i];
            b[k] = bell_of(u[k], a.g.mu[k], a.g.bellScale[k]);
            sumStd += a.g.weight[k] * (2.0f * b[k] - 1.0f);
        }
        float pre = A0 + a.g.dt * sumStd;
        float gin = gp(g)[i];
        float gpv = (pre > 0.0f && pre < 1.0f) ? gin : 0.0f;
        gp(gpOut)[i] = gpv;
        vals[0] = gpv * sumStd;
        for (int k = 0; k < a.K; ++k) {
            float sg = a.g.sigma[k], d = u[k] - a.g.mu[k];
            float cterm = gpv * a.g.dt * a.g.weight[k] * 2.0f * b[k];
            float is2 = 1.0f / (sg * sg);
            vals[1 + 3 * k] = cterm * d * is2;
            vals[2 + 3 * k] = cterm * d * d * is2 / sg;
            vals[3 + 3 * k] = gpv * a.g.dt * (2.0f * b[k] - 1.0f);
            gp(GU)[k * n + i] = -cterm * d * is2;
        }
    }
    const int nv = 1 + 3 * a.K;
    for (int v = 0; v < nv; ++v) {
        red[tid] = vals[v];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        if (tid == 0) gp(part)[(size_t)blockIdx.x * 16 + v] = red[0];
        __syncthreads();
    }
}

// fixed-order sums of block partials, accumulated in FP64:
// acc[v] += sum_b part[b][v] for v < nv
__global__ void k_adj_reduce(const float *part, int nBlocks, int nv, double *acc) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    for (int v = 0; v < nv; ++v) {
        double s = 0.0;
        for (int b = tid; b < nBlocks; b += 256) s += (double)gp(part)[(size_t)b * 16 + v];
        red[tid] = s;
        __syncthreads();
        for (int k = 128; k > 0; k >>= 1) {
            if (tid < k) red[tid] += red[tid + k];
            __syncthreads();
        }
        if (tid == 0) gp(acc)[v] += red[0];
        __syncthreads();
    }
}

// ------------------------------------------------------------------ launchers
template <class F>
static hipError_t by_threads(int T, F f) {
    switch (T) {
        case 64: f(std::integral_constant<int, 64>()); break;
        case 128: f(std::integral_constant<int, 128>()); break;
        case 256: f(std::integral_constant<int, 256>()); break;
        case 512: f(std::integral_constant<int, 512>()); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// LDS: the twiddles, then the swizzled region rounded up to 32 entries (sw()
// permutes inside aligned blocks of 32)
static inline int r32(int n) { return (n + 31) & ~31; }

hipError_t launch_adj_rows_fwd(const AdjArgs &a, int T, const float *in, size_t fieldStride,
                               float *out, int nf, hipStream_t s) {
    size_t lds = (size_t)(a.W + r32(nf * a.W)) * sizeof(float2);
    return by_threads(T, [&](auto t) {
        constexpr int TT = decltype(t)::value;
        hipLaunchKernelGGL(k_adj_rows_fwd<TT>, dim3(a.H / 2), dim3(TT), lds, s, a, in, fieldStride,
                           out, nf);
    });
}

hipError_t launch_adj_rows_inv(const AdjArgs &a, int T, const float *G, float *out, int ns,
                               const float *addend, hipStream_t s) {
    size_t lds = (size_t)(a.W + r32(ns * a.W)) * sizeof(float2);
    return by_threads(T, [&](auto t) {
        constexpr int TT = decltype(t)::value;
        hipLaunchKernelGGL(k_adj_rows_inv<TT>, dim3(a.H / 2), dim3(TT), lds, s, a, G, out, ns,
                           addend);
    });
}

hipError_t launch_adj_cols(const AdjArgs &a, int T, bool adj, const float *Fin, float *Ahat,
                           float *G, float *betaPart, hipStream_t s) {
    const int nin = adj ? a.K : 1;
    // twiddles | X [nin][H] | POT: Y [K][H] ; ADJ: Y [H] | reduction [T <= 512]
    size_t lds = (size_t)(a.H + r32(nin * a.H) + (adj ? r32(a.H) : r32(a.K * a.H))) * sizeof(float2) +
                 (adj ? 512 * sizeof(float) : 0);
    const int Wh = a.W / 2 + 1;
    return by_threads(T, [&](auto t) {
        constexpr int TT = decltype(t)::value;
        if (adj)
            hipLaunchKernelGGL((k_adj_cols<TT, true>), dim3(Wh), dim3(TT), lds, s, a, Fin, Ahat, G,
                               betaPart);
        else
            hipLaunchKernelGGL((k_adj_cols<TT, false>), dim3(Wh), dim3(TT), lds, s, a, Fin, Ahat,
                               G, betaPart);
    });
}

hipError_t launch_adj_cells(const AdjArgs &a, const float *A, const float *U, const float *g,
                            float *gpOut, float *GU, float *part, hipStream_t s) {
    const size_t n = (size_t)a.W * a.H;
    hipLaunchKernelGGL(k_adj_cells, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, A, U,
                       g, gpOut, GU, part);
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const float *part, int nBlocks, int nv, double *acc, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_reduce, dim3(1), dim3(256), 0, s, part, nBlocks, nv, acc);
    return hipGetLastError();
}

}  // namespace lenia
