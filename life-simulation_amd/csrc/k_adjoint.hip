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
// The recompute of a segment runs the normal step kernels with their save
// pointers set (FftWorldDev::saveA/saveU/saveAhat), so each step leaves its
// input state A_t, its potentials u_k and the 2D spectrum Ahat_t.  A reverse
// step is a row launch and a column launch (row transforms along x on row
// pairs packed as one complex row, column transforms along y on the half
// spectrum):
//   k_adj_mid<EDGE>  EDGE: first finish the later step, g_t+1 = gp + row
//                    inverse of G1 (kept in registers, never stored); then
//                    per cell gp, gu_k and the parameter partials from the
//                    saved u_k and A_t; row forward of gu_k -> FG_k
//   k_adj_cols       column forward of FG_k; shell partials against Ahat_t;
//                    G1 = column inverse of sum_k S_k FG_k^
//   k_adj_edge       after the first step only: g_0 = gp + row inverse of G1
// The kernels reuse the LDS FFT of the forward path (same plans, same
// butterflies).  Each workgroup adds its partial sums to its own FP64 slots
// (one row pair or one spectrum column per workgroup, every step in order),
// and a final kernel sums the slots in a fixed order, so gradients are
// identical run to run.
#include <hip/hip_runtime.h>

#include "fft_device.h"
#include "lenia_device.h"
#include "launch.h"

namespace lenia {

constexpr int kAV = 16;  // values per thread per FFT stage

__device__ __forceinline__ void load_twiddles(float2 *tw, const float *src, int n, int tid, int T) {
    for (int t = tid; t < n; t += T) tw[t] = gld2(gp(reinterpret_cast<const float2 *>(src)) + t);
}

// half spectrum slot (float4 per kx: rows 2rp and 2rp+1) -> packed complex row
// Z[base + x] = a[x] + i b[x] (Hermitian mirror fills the negative
// frequencies; conjugated when CONJ).  LDS indices are placed as a whole:
// slot s of an FFT batch lives at sw(s W + x), not at s W + sw(x).
template <bool CONJ>
__device__ __forceinline__ void gather_row_pair(float2 *Z, int base, const float *G, int W, int H2,
                                                int rp, int tid, int T) {
    const int Wh = W / 2 + 1;
    const float sg = CONJ ? -1.0f : 1.0f;
    for (int kx = tid; kx < Wh; kx += T) {
        float4 g4 = gld4(gp(reinterpret_cast<const float4 *>(G)) + (size_t)kx * H2 + rp);
        if (kx == 0 || kx == W / 2) {
            g4.y = 0.0f;
            g4.w = 0.0f;
        }
        Z[sw(base + kx)] = make_float2(g4.x - g4.w, sg * (g4.y + g4.z));
        if (kx > 0 && kx < W / 2) Z[sw(base + W - kx)] = make_float2(g4.x + g4.w, sg * (g4.z - g4.y));
    }
}

// packed complex row spectrum at Z[base ..] (two real rows a + i b) -> their
// half spectra, stored as one float4 per kx
__device__ __forceinline__ void split_row_pair(const float2 *Z, int base, float *out, int W, int H2,
                                               int rp, int tid, int T) {
    const int Wh = W / 2 + 1;
    LENIA_GLOBAL float4 *o = gp(reinterpret_cast<float4 *>(out));
    for (int kx = tid; kx < Wh; kx += T) {
        float2 z1 = Z[sw(base + kx)];
        float2 z2 = Z[sw(base + (kx == 0 ? 0 : W - kx))];
        gst4(o + (size_t)kx * H2 + rp, make_float4(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y),
                                                   0.5f * (z1.y + z2.y), 0.5f * (z2.x - z1.x)));
    }
}

// ------------------------------------------------------------------ rows: edge
// g = gpb + row inverse of G1, one row pair per workgroup
template <int T>
__global__ __launch_bounds__(T) void k_adj_edge(AdjArgs a, const float *G1, const float *gpb,
                                                float *gOut) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int W = a.W, H2 = a.H / 2;
    const int rp = blockIdx.x;
    float2 *tw = smem;
    float2 *Z = smem + W;
    load_twiddles(tw, a.twW, W, tid, T);
    gather_row_pair<false>(Z, 0, G1, W, H2, rp, tid, T);
    __syncthreads();
    fft_run<true, kAV, T>(Z, a.planW, 1, tw, tid);
    const LENIA_GLOBAL float *ad = gp(gpb);
    LENIA_GLOBAL float *go = gp(gOut);
    for (int x = tid; x < W; x += T) {
        float2 z = Z[sw(x)];
        const size_t i0 = (size_t)(2 * rp) * W + x;
        go[i0] = ad[i0] + z.x;
        go[i0 + W] = ad[i0 + W] + z.y;
    }
}

// ------------------------------------------------------------------ rows: middle
// saved potentials U [K][H][W] and state A -> per cell gp (gpb), gu_k; row
// forward of gu_k -> FG.  Partials of dt, mu_k, sigma_k, w_k are added to
// accRows[rp][v] (FP64).
// Workgroup sums of v[i] (i < n, n <= NV) in a fixed order, added to dst[i]
// in FP64: a butterfly over the 64 lanes of each wave (partners add the same
// two values, so every lane holds the same bits), then the T/64 wave sums in
// wave order.  One barrier; `red` holds NV * T/64 floats.
template <int T, int NV>
__device__ __forceinline__ void block_sums(const float (&v)[NV], int n, float scale, float *red,
                                           LENIA_GLOBAL double *dst, int tid) {
    constexpr int NW = T / 64;
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        if (i < n) {
            float x = v[i] * scale;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
            if (lane == 0) red[i * NW + wave] = x;
        }
    }
    __syncthreads();
    if (tid < n) {
        float t = 0.0f;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += red[tid * NW + w];
        dst[tid] += (double)t;
    }
}

// EDGE: the incoming gradient of this step is gpb + row inverse of G1 (the
// later step's edge, fused here); gpb is then overwritten in place with this
// step's gp (each cell is read and written by the same thread).  Otherwise
// the incoming gradient is read from g.
template <int T, bool EDGE>
__global__ __launch_bounds__(T) void k_adj_mid(AdjArgs a, const float *U, const float *A,
                                               const float *g, const float *G1, float *gpb,
                                               float *FG, double *accRows) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int W = a.W, H = a.H, Wh = W / 2 + 1, H2 = H / 2, K = a.K;
    const int rp = blockIdx.x;
    const size_t sStride = (size_t)Wh * H;
    float2 *tw = smem;
    float2 *Z = smem + W;
    float *red = reinterpret_cast<float *>(Z + lds_span(K * W));
    const LENIA_GLOBAL float *Ag = gp(A);
    const LENIA_GLOBAL float *gg = gp(g);
    const LENIA_GLOBAL float *Ug = gp(U);
    LENIA_GLOBAL float *gpo = gp(gpb);
    const size_t n = (size_t)W * H;
    // the operands of the first loads (cells, twiddles, the later step's
    // spectrum) in scalar registers: loaded together, waited for once
    asm volatile("" ::"s"(U), "s"(A), "s"(g), "s"(G1), "s"(a.twW), "s"(W), "s"(H), "s"(K));
    // a cell pair's operands (rows 2rp and 2rp+1 at column x): potentials,
    // state, incoming gradient (EDGE: the gp part of it)
    auto load_x = [&](int x, float2 (&u2)[4], float2 &a2, float2 &g2) {
        const size_t i0 = (size_t)(2 * rp) * W + x;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            u2[k] = k < K ? make_float2(Ug[k * n + i0], Ug[k * n + i0 + W]) : make_float2(0.f, 0.f);
        a2 = make_float2(Ag[i0], Ag[i0 + W]);
        g2 = EDGE ? make_float2(gpo[i0], gpo[i0 + W]) : make_float2(gg[i0], gg[i0 + W]);
    };
    // the operands of each thread's first kPre columns are requested before
    // the transforms, which hide their latency
    constexpr int kPre = 2;
    float2 pu[kPre][4], pa[kPre], pg[kPre];
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
        const int x = tid + i * T;
        if (x < W) load_x(x, pu[i], pa[i], pg[i]);
    }
    load_twiddles(tw, a.twW, W, tid, T);
    if (EDGE) {
        // slot 0 holds the later step's row inverse; the K slots written
        // below reuse it only at each thread's own cells, after reading them
        gather_row_pair<false>(Z, 0, G1, W, H2, rp, tid, T);
        __syncthreads();
        fft_run<true, kAV, T>(Z, a.planW, 1, tw, tid);
    }
    const GrowthParams &gr = *a.gDev;
    const float dt = gr.dt;
    float mu[4], bs[4], wk[4], sg[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        mu[k] = gr.mu[k];
        bs[k] = gr.bellScale[k];
        wk[k] = gr.weight[k];
        sg[k] = gr.sigma[k];
    }
    float vals[13];
#pragma unroll
    for (int v = 0; v < 13; ++v) vals[v] = 0.0f;
    auto cell = [&](int x, const float2 (&u2)[4], float2 a2, float2 g2) {
        float2 gu2[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) gu2[k] = make_float2(0.f, 0.f);
        float gin[2] = {g2.x, g2.y};
        if (EDGE) {
            const float2 z = Z[sw(x)];
            gin[0] = g2.x + z.x;
            gin[1] = g2.y + z.y;
        }
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const size_t i = (size_t)(2 * rp + half) * W + x;
            float u[4], b[4], sumStd = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                u[k] = half ? u2[k].y : u2[k].x;
                b[k] = bell_of(u[k], mu[k], bs[k]);
                if (k < K) sumStd += wk[k] * (2.0f * b[k] - 1.0f);
            }
            const float pre = (half ? a2.y : a2.x) + dt * sumStd;
            const float gpv = (pre > 0.0f && pre < 1.0f) ? gin[half] : 0.0f;
            gpo[i] = gpv;
            vals[0] += gpv * sumStd;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k < K) {
                    const float d = u[k] - mu[k];
                    const float cterm = gpv * dt * wk[k] * 2.0f * b[k];
                    const float is2 = 1.0f / (sg[k] * sg[k]);
                    vals[1 + 3 * k] += cterm * d * is2;
                    vals[2 + 3 * k] += cterm * d * d * is2 / sg[k];
                    vals[3 + 3 * k] += gpv * dt * (2.0f * b[k] - 1.0f);
                    const float gu = -cterm * d * is2;
                    if (half) gu2[k].y = gu; else gu2[k].x = gu;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < K) Z[sw(k * W + x)] = gu2[k];
    };
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
        const int x = tid + i * T;
        if (x < W) cell(x, pu[i], pa[i], pg[i]);
    }
    for (int x = tid + kPre * T; x < W; x += T) {
        float2 u2[4], a2, g2;
        load_x(x, u2, a2, g2);
        cell(x, u2, a2, g2);
    }
    __syncthreads();
    fft_run<false, kAV, T>(Z, a.planW, K, tw, tid);
    for (int k = 0; k < K; ++k) split_row_pair(Z, k * W, FG + k * sStride * 2, W, H2, rp, tid, T);
    // fixed-order workgroup sums of the partials
    block_sums<T, 13>(vals, 1 + 3 * K, 1.0f, red, gp(accRows) + (size_t)rp * 16, tid);
}

// ------------------------------------------------------------------ columns
// FG[k] (k < K) -> column forward -> spectral partial sums for the shell
// gradients against Ahat (added to accCols[c][i]) -> Y = sum_k S_k FG^[k] ->
// column inverse -> G
template <int T>
__global__ __launch_bounds__(T) void k_adj_cols(AdjArgs a, const float *Fin, const float *Ahat,
                                                float *G, double *accCols) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int H = a.H, Wh = a.W / 2 + 1, K = a.K;
    const int c = blockIdx.x;  // spectrum column kx
    const size_t sStride = (size_t)Wh * H;
    float2 *tw = smem;
    float2 *X = smem + H;  // K slots
    const int nb = a.nShells;  // total (k, j) pairs, <= 16
    const LENIA_GLOBAL float2 *Ah = gp(reinterpret_cast<const float2 *>(Ahat)) + (size_t)c * H;
    const LENIA_GLOBAL float *Dc = gp(a.D) + (size_t)c * H;
    const LENIA_GLOBAL float *Sc = gp(a.S) + (size_t)c * H;
    // an empty statement taking the operands of the first loads in scalar
    // registers: their kernel-argument loads are issued together and waited
    // for once, before the loads they address
    asm volatile("" ::"s"(Ahat), "s"(Fin), "s"(a.D), "s"(a.S), "s"(a.twH), "s"(a.planH.code), "s"(nb), "s"(K),
                 "s"(H), "s"(Wh));
    // per spectrum row ky: the state spectrum, the shell-gradient spectra and
    // the kernel spectra at (c, ky)
    auto load_y = [&](int ky, float2 &ah, float (&d)[16], float (&sk)[4]) {
        ah = gld2(Ah + ky);
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < nb) d[i] = Dc[(size_t)i * sStride + ky];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < K) sk[k] = Sc[(size_t)k * sStride + ky];
    };
    // each thread's first kPre rows are requested before the transform
    constexpr int kPre = 2;
    float2 pah[kPre];
    float pd[kPre][16], ps[kPre][4];
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
        const int ky = tid + i * T;
        if (ky < H) load_y(ky, pah[i], pd[i], ps[i]);
    }
    load_twiddles(tw, a.twH, H, tid, T);
    for (int e = tid; e < K * H; e += T) {
        int s = e / H, ky = e - s * H;
        X[sw(e)] = gld2(gp(reinterpret_cast<const float2 *>(Fin)) + s * sStride + (size_t)c * H + ky);
    }
    __syncthreads();
    fft_run<false, kAV, T>(X, a.planH, K, tw, tid);
    float2 *Y = X + lds_span(K * H);
    // shell partial sums: thread-strided over ky, then a fixed-order
    // workgroup sum; Y = sum_k S_k * gu_k^
    const float wt = (c == 0 || c == Wh - 1) ? 1.0f : 2.0f;
    float acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    auto row = [&](int ky, float2 ah, const float (&d)[16], const float (&sk)[4]) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < nb) {
                int k = a.shellK[i];
                float2 gu = X[sw(k * H + ky)];
                float re = gu.x * ah.x + gu.y * ah.y;  // Re(conj(gu) * A)
                acc[i] += d[i] * re;
            }
        }
        float2 y = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < K) {
                float2 x = X[sw(k * H + ky)];
                y.x = fmaf(sk[k], x.x, y.x);
                y.y = fmaf(sk[k], x.y, y.y);
            }
        }
        Y[sw(ky)] = y;
    };
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
        const int ky = tid + i * T;
        if (ky < H) row(ky, pah[i], pd[i], ps[i]);
    }
    for (int ky = tid + kPre * T; ky < H; ky += T) {
        float2 ah;
        float d[16], sk[4];
        load_y(ky, ah, d, sk);
        row(ky, ah, d, sk);
    }
    // reduce the partial sums across the workgroup (the LDS after Y); the
    // barrier inside also orders the Y stores before the inverse transform
    float *red = reinterpret_cast<float *>(Y + lds_span(H));
    block_sums<T, 16>(acc, nb, wt, red, gp(accCols) + (size_t)c * 16, tid);
    fft_run<true, kAV, T>(Y, a.planH, 1, tw, tid);
    for (int e = tid; e < H; e += T) gst2(gp(reinterpret_cast<float2 *>(G)) + (size_t)c * H + e, Y[sw(e)]);
}

// ------------------------------------------------------------------ final sums
// acc[v] = sum_rp accRows[rp][v] (v < nvRows), acc[16 + i] = sum_c accCols[c][i]
// (i < nvCols), each a fixed-order FP64 tree
__global__ void k_adj_final(const double *accRows, int nRows, int nvRows, const double *accCols,
                            int nCols, int nvCols, double *acc) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    for (int v = 0; v < nvRows + nvCols; ++v) {
        const bool rows = v < nvRows;
        const double *src = rows ? accRows : accCols;
        const int n = rows ? nRows : nCols, col = rows ? v : v - nvRows;
        double s = 0.0;
        for (int b = tid; b < n; b += 256) s += gp(src)[(size_t)b * 16 + col];
        red[tid] = s;
        __syncthreads();
        for (int k = 128; k > 0; k >>= 1) {
            if (tid < k) red[tid] += red[tid + k];
            __syncthreads();
        }
        if (tid == 0) gp(acc)[rows ? v : 16 + col] = red[0];
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

// LDS: the twiddles, then the transform regions (lds_span entries each)

size_t adj_lds_rows(const AdjArgs &a, int T) {
    // mid: twiddles | Z [K][W] | reduction [T];  edge: twiddles | Z [W]
    return (size_t)(a.W + lds_span(a.K * a.W)) * sizeof(float2) + (size_t)T * sizeof(float);
}

size_t adj_lds_cols(const AdjArgs &a, int T) {
    // twiddles | X [K][H] | Y [H] | reduction [T]
    return (size_t)(a.H + lds_span(a.K * a.H) + lds_span(a.H)) * sizeof(float2) + (size_t)T * sizeof(float);
}

hipError_t launch_adj_edge(const AdjArgs &a, int T, const float *G1, const float *gpb, float *gOut,
                           hipStream_t s) {
    const size_t lds = (size_t)(a.W + lds_span(a.W)) * sizeof(float2);
    return by_threads(T, [&](auto t) {
        constexpr int TT = decltype(t)::value;
        hipLaunchKernelGGL(k_adj_edge<TT>, dim3(a.H / 2), dim3(TT), lds, s, a, G1, gpb, gOut);
    });
}

hipError_t launch_adj_mid(const AdjArgs &a, int T, const float *U, const float *A, const float *g,
                          const float *G1, float *gpb, float *FG, double *accRows, hipStream_t s) {
    const size_t lds = adj_lds_rows(a, T);
    return by_threads(T, [&](auto t) {
        constexpr int TT = decltype(t)::value;
        if (G1)
            hipLaunchKernelGGL((k_adj_mid<TT, true>), dim3(a.H / 2), dim3(TT), lds, s, a, U, A, g, G1,
                               gpb, FG, accRows);
        else
            hipLaunchKernelGGL((k_adj_mid<TT, false>), dim3(a.H / 2), dim3(TT), lds, s, a, U, A, g, G1,
                               gpb, FG, accRows);
    });
}

hipError_t launch_adj_cols(const AdjArgs &a, int T, const float *Fin, const float *Ahat, float *G,
                           double *accCols, hipStream_t s) {
    const size_t lds = adj_lds_cols(a, T);
    const int Wh = a.W / 2 + 1;
    return by_threads(T, [&](auto t) {
        constexpr int TT = decltype(t)::value;
        hipLaunchKernelGGL(k_adj_cols<TT>, dim3(Wh), dim3(TT), lds, s, a, Fin, Ahat, G, accCols);
    });
}

// copies of the world record with save pointers for c recompute steps:
// step j saves into saveA + j n, saveU + j K n, saveAhat + j spec
__global__ void k_adj_records(const FftWorldDev *src, FftWorldDev *dst, int c, float *saveA,
                              float *saveU, float *saveAhat, size_t n, size_t spec) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= c) return;
    FftWorldDev r = *src;
    r.saveA = saveA + (size_t)j * n;
    r.saveU = saveU + (size_t)j * r.K * n;
    r.saveAhat = saveAhat + (size_t)j * spec;
    dst[j] = r;
}

hipError_t launch_adj_records(const FftWorldDev *src, FftWorldDev *dst, int c, float *saveA,
                              float *saveU, float *saveAhat, size_t n, size_t spec, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_records, dim3((c + 63) / 64), dim3(64), 0, s, src, dst, c, saveA, saveU,
                       saveAhat, n, spec);
    return hipGetLastError();
}

hipError_t launch_adj_final(const double *accRows, int nRows, int nvRows, const double *accCols,
                            int nCols, int nvCols, double *acc, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_final, dim3(1), dim3(256), 0, s, accRows, nRows, nvRows, accCols, nCols,
                       nvCols, acc);
    return hipGetLastError();
}

}  // namespace lenia
