// FFT-convolution form of the Lenia step on gfx950.
//
// The reference runs, per substep and kernel count K, one dispatch per radix-2
// stage and axis plus R<->C, multiply and growth passes: 84 passes for
// 1024x512, K=3 (Modules/FieldModules/Lenia.cpp:576-643, encodeFFT2D :57-85).
// Here a step is two kernels, each touching HBM once per operand:
//
//   column pass  reads the state's half-spectrum F (after the row FFT),
//                transforms each column forward, multiplies by the K real
//                kernel spectra and transforms back: writes G_k, k < K.
//   row pass     for each row pair: K inverse complex FFTs of G_k packed as
//                one complex row (two real rows per transform), growth and
//                clip for every kernel (leniaGrowthMulti, Lenia.metal:161-232),
//                in-place state update, then the forward real FFT of the new
//                rows, which becomes the next step's F.
//
// Spectra are stored column-major ([column][row]) so the column pass reads and
// writes whole contiguous columns and the row pass moves two adjacent rows
// (16 bytes) per spectrum column.  The inverse 1/(W H) is folded into the
// kernel spectra, which are computed in FP64 from the kernel window.
#include <hip/hip_runtime.h>

#include "fft_device.h"
#include "lenia_device.h"
#include "launch.h"

namespace lenia {

constexpr int kT = 256;    // threads per workgroup
constexpr int kMaxV = 16;  // complex values per thread per transform batch

template <class WorldT>
__device__ __forceinline__ void arrive_all(uint32_t *done, const WorldT *worlds, int nWorlds) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        uint32_t prev = atomicAdd(done, 1u);
        if (prev == gridDim.x - 1) {
            __threadfence();
            for (int i = 0; i < nWorlds; ++i) {
                StepClock *cp = worlds[i].clock;
                StepClock c = *cp;
                c.sub += 1u;
                if (c.sub >= c.substeps) {
                    c.sub = 0u;
                    c.frame += 1u;
                }
                c.parity ^= 1u;
                c.steps += 1u;
                *cp = c;
            }
            __threadfence();
            atomicExch(done, 0u);
        }
    }
}

// ---------------------------------------------------------------- row pass
// mode: 0 = forward only (state -> F), 1 = step (G -> growth -> state -> F),
//       2 = potential of kernel `kSel` only (G_kSel -> out), no state change
template <int MODE>
__global__ __launch_bounds__(kT) void k_fft_rows(const FftWorldDev *__restrict__ worlds,
                                                 const int2 *__restrict__ tiles, int nTiles,
                                                 uint32_t *done, int nWorlds, int kSel,
                                                 float *__restrict__ out) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    for (int ti = blockIdx.x; ti < nTiles; ti += gridDim.x) {
        const int2 tile = tiles[ti];
        const FftWorldDev &wd = worlds[tile.x];
        const int W = wd.W, H = wd.H, Wh = wd.Wh;
        const int rp0 = tile.y * wd.P;
        const int nt = min(wd.P, H / 2 - rp0);
        float2 *Z = smem;
        const float2 *tw = reinterpret_cast<const float2 *>(wd.twW);
        float *A = wd.A;
        const int nq = nt * W;

        if (MODE == 0) {
            for (int q = tid; q < nq; q += kT) {
                int p = q / W, x = q - p * W;
                int ya = 2 * (rp0 + p);
                Z[q] = make_float2(A[(size_t)ya * W + x], A[(size_t)(ya + 1) * W + x]);
            }
            __syncthreads();
        } else {
            const GrowthParams &g = wd.g;
            const StepClock ck = *wd.clock;
            const uint32_t seed = clock_seed(ck, g.moduleSeed);
            float acc0[kMaxV], acc1[kMaxV];
#pragma unroll
            for (int i = 0; i < kMaxV; ++i) acc0[i] = acc1[i] = 0.0f;
            const int kBeg = MODE == 2 ? kSel : 0;
            const int kEnd = MODE == 2 ? kSel + 1 : (int)g.nKernels;
            for (int k = kBeg; k < kEnd; ++k) {
                const float4 *G4 = reinterpret_cast<const float4 *>(wd.G) + (size_t)k * Wh * (H / 2);
                for (int m = tid; m < nt * Wh; m += kT) {
                    int kx = m / nt, p = m - kx * nt;
                    float4 g4 = G4[(size_t)kx * (H / 2) + rp0 + p];
                    if (kx == 0 || kx == W / 2) {
                        // these columns are real for a real field: drop the
                        // rounding residue instead of leaking it across rows
                        g4.y = 0.0f;
                        g4.w = 0.0f;
                    }
                    // Z = Ga + i Gb on the stored half, Hermitian mirror on the rest
                    Z[p * W + kx] = make_float2(g4.x - g4.w, g4.y + g4.z);
                    if (kx > 0 && kx < W / 2)
                        Z[p * W + W - kx] = make_float2(g4.x + g4.w, g4.z - g4.y);
                }
                __syncthreads();
                fft_lds<true, kMaxV, kT>(Z, wd.planW, nt, tw, tid);
                if (MODE == 2) {
                    for (int q = tid; q < nq; q += kT) {
                        int p = q / W, x = q - p * W;
                        int ya = 2 * (rp0 + p);
                        float2 z = Z[q];
                        out[(size_t)ya * W + x] = z.x;
                        out[(size_t)(ya + 1) * W + x] = z.y;
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < kMaxV; ++i) {
                        int q = tid + i * kT;
                        if (q < nq) {
                            int p = q / W, x = q - p * W;
                            int ya = 2 * (rp0 + p);
                            float2 z = Z[q];
                            float ja = cell_jitter(g, x, ya, seed, ck.frame);
                            float jb = cell_jitter(g, x, ya + 1, seed, ck.frame);
                            acc0[i] += growth_term(g, k, z.x, ja);
                            acc1[i] += growth_term(g, k, z.y, jb);
                        }
                    }
                }
                __syncthreads();
            }
            if (MODE == 2) continue;
#pragma unroll
            for (int i = 0; i < kMaxV; ++i) {
                int q = tid + i * kT;
                if (q < nq) {
                    int p = q / W, x = q - p * W;
                    int ya = 2 * (rp0 + p);
                    size_t ia = (size_t)ya * W + x, ib = ia + W;
                    float a2 = growth_finish(g, A[ia], acc0[i], x, ya, W, H, seed, ck.frame);
                    float b2 = growth_finish(g, A[ib], acc1[i], x, ya + 1, W, H, seed, ck.frame);
                    A[ia] = a2;
                    A[ib] = b2;
                    Z[q] = make_float2(a2, b2);
                }
            }
            __syncthreads();
        }
        // forward FFT of the (new) row pair; split into the two real rows'
        // half-spectra: Fa = (Z[k] + conj Z[W-k]) / 2, Fb = (Z[k] - conj Z[W-k]) / 2i
        fft_lds<false, kMaxV, kT>(Z, wd.planW, nt, tw, tid);
        float4 *F4 = reinterpret_cast<float4 *>(wd.F);
        for (int m = tid; m < nt * Wh; m += kT) {
            int kx = m / nt, p = m - kx * nt;
            float2 z1 = Z[p * W + kx];
            float2 z2 = Z[p * W + (kx == 0 ? 0 : W - kx)];
            F4[(size_t)kx * (H / 2) + rp0 + p] =
                make_float4(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y), 0.5f * (z1.y + z2.y),
                            0.5f * (z2.x - z1.x));
        }
        __syncthreads();
    }
    if (MODE == 1) arrive_all(done, worlds, nWorlds);
}

// ---------------------------------------------------------------- column pass
__global__ __launch_bounds__(kT) void k_fft_cols(const FftWorldDev *__restrict__ worlds,
                                                 const int2 *__restrict__ tiles, int nTiles) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    for (int ti = blockIdx.x; ti < nTiles; ti += gridDim.x) {
        const int2 tile = tiles[ti];
        const FftWorldDev &wd = worlds[tile.x];
        const int H = wd.H, Wh = wd.Wh;
        const int c0 = tile.y * wd.C;
        const int nt = min(wd.C, Wh - c0);
        const int ne = nt * H;
        float2 *X = smem;
        const float2 *tw = reinterpret_cast<const float2 *>(wd.twH);
        const float2 *F = reinterpret_cast<const float2 *>(wd.F) + (size_t)c0 * H;
        for (int e = tid; e < ne; e += kT) X[e] = F[e];
        __syncthreads();
        fft_lds<false, kMaxV, kT>(X, wd.planH, nt, tw, tid);
        float2 xs[kMaxV];
#pragma unroll
        for (int i = 0; i < kMaxV; ++i) {
            int e = tid + i * kT;
            xs[i] = e < ne ? X[e] : make_float2(0.f, 0.f);
        }
        for (int k = 0; k < (int)wd.g.nKernels; ++k) {
            const float *S = wd.S + ((size_t)k * Wh + c0) * H;
#pragma unroll
            for (int i = 0; i < kMaxV; ++i) {
                int e = tid + i * kT;
                if (e < ne) X[e] = cscale(xs[i], S[e]);
            }
            __syncthreads();
            fft_lds<true, kMaxV, kT>(X, wd.planH, nt, tw, tid);
            float2 *G = reinterpret_cast<float2 *>(wd.G) + ((size_t)k * Wh + c0) * H;
            for (int e = tid; e < ne; e += kT) G[e] = X[e];
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- set-up kernels
// e^{-2 pi i t / N} for t < N, evaluated in FP64 and rounded once to FP32
__global__ void k_twiddles(float2 *tw, int N) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    double a = 2.0 * (double)t / (double)N;
    tw[t] = make_float2((float)cospi(a), (float)(-sinpi(a)));
}

// first half of the separable kernel spectrum, FP64:
// Cw[iy][kx] = sum_ix win[iy][ix] cos(2 pi kx (ix - h) / W)
__global__ void k_spec_rows(const double *__restrict__ win, int h, int W, int Wh,
                            double *__restrict__ Cw) {
    int kx = blockIdx.x * blockDim.x + threadIdx.x;
    int iy = blockIdx.y;
    if (kx >= Wh) return;
    const int n = 2 * h + 1;
    double acc = 0.0;
    for (int ix = 0; ix < n; ++ix) {
        double w = win[(size_t)iy * n + ix];
        if (w == 0.0) continue;
        long long m = ((long long)kx * (ix - h)) % W;
        if (m < 0) m += W;
        acc += w * cospi(2.0 * (double)m / (double)W);
    }
    Cw[(size_t)iy * Wh + kx] = acc;
}

// S[kx][ky] = scale * sum_iy Cw[iy][kx] cos(2 pi ky (iy - h) / H).  The kernel
// is even in x and y, so its DFT is this real cosine sum.
__global__ void k_spec_cols(const double *__restrict__ Cw, int h, int H, int Wh, double scale,
                            float *__restrict__ S) {
    int ky = blockIdx.x * blockDim.x + threadIdx.x;
    int kx = blockIdx.y;
    if (ky >= H) return;
    const int n = 2 * h + 1;
    double acc = 0.0;
    for (int iy = 0; iy < n; ++iy) {
        long long m = ((long long)ky * (iy - h)) % H;
        if (m < 0) m += H;
        acc += Cw[(size_t)iy * Wh + kx] * cospi(2.0 * (double)m / (double)H);
    }
    S[(size_t)kx * H + ky] = (float)(acc * scale);
}

// ---------------------------------------------------------------- launchers
hipError_t launch_fft_rows(int mode, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                           int grid, size_t lds, uint32_t *done, int nWorlds, int kSel,
                           float *out, hipStream_t s) {
    if (mode == 0)
        hipLaunchKernelGGL(k_fft_rows<0>, dim3(grid), dim3(kT), lds, s, worlds, tiles, nTiles,
                           done, nWorlds, kSel, out);
    else if (mode == 1)
        hipLaunchKernelGGL(k_fft_rows<1>, dim3(grid), dim3(kT), lds, s, worlds, tiles, nTiles,
                           done, nWorlds, kSel, out);
    else
        hipLaunchKernelGGL(k_fft_rows<2>, dim3(grid), dim3(kT), lds, s, worlds, tiles, nTiles,
                           done, nWorlds, kSel, out);
    return hipGetLastError();
}

hipError_t launch_fft_cols(const FftWorldDev *worlds, const int2 *tiles, int nTiles, int grid,
                           size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_fft_cols, dim3(grid), dim3(kT), lds, s, worlds, tiles, nTiles);
    return hipGetLastError();
}

hipError_t launch_twiddles(float *tw, int N, hipStream_t s) {
    hipLaunchKernelGGL(k_twiddles, dim3((N + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<float2 *>(tw), N);
    return hipGetLastError();
}

hipError_t launch_kernel_spectrum(const double *win, int h, int W, int H, double *Cw, float *S,
                                  hipStream_t s) {
    const int Wh = W / 2 + 1;
    hipLaunchKernelGGL(k_spec_rows, dim3((Wh + 127) / 128, 2 * h + 1), dim3(128), 0, s, win, h,
                       W, Wh, Cw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_spec_cols, dim3((H + 127) / 128, Wh), dim3(128), 0, s, Cw, h, H, Wh,
                       1.0 / ((double)W * (double)H), S);
    return hipGetLastError();
}

int fft_threads() { return kT; }
int fft_max_values() { return kMaxV; }

}  // namespace lenia
