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
// (16 bytes) per spectrum column; tiles are mapped so that neighbouring row
// pairs run on one XCD and share its L2.  Every global operand of a tile is
// requested before the FFT that precedes its use, so loads overlap the LDS
// work.  The inverse 1/(W H) is folded into the kernel spectra, which are
// computed in FP64 from the kernel window.
//
// This header holds the step kernels and their launch templates; each
// k_fft_pass<T>.hip instantiates them for one workgroup size (separate
// translation units compile in parallel), k_fft.hip holds the set-up kernels
// and the entry points that pick the instantiation.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_device.h"
#include "lenia_device.h"
#include "launch.h"

namespace lenia {

constexpr int kMaxV = 16;  // complex values per thread per stage

// Cells per thread held in registers by the row pass (state, growth sums):
// 1024-thread workgroups keep at most 12, so the values live across the
// transforms fit their 128-register budget without spilling (16 spilled);
// the host's shape check admits a launch only when (K) nt W <= row_values T.
// Tiles of at most kLeanValues cells per thread run a variant sized for
// them, which issues fewer idle loop iterations.
constexpr int row_values(int T) { return T >= 1024 ? 12 : kMaxV; }
constexpr int kLeanValues = 4;
// 256- and 512-thread step tiles of at most 12 cells per thread run a 12-cell variant
constexpr int kMidValues = 12;
constexpr int kLeanCols = 8;

__device__ __forceinline__ int tile_of(int ti, int nTiles) {
    return gridDim.x == (unsigned)nTiles ? xcd_tile(ti, nTiles) : ti;
}

// ---------------------------------------------------------------- row pass
// Twiddle-table copy global -> LDS, split in two so its loads can be in
// flight together with the other loads of a tile
template <int T>
struct TwCopy {
    static constexpr int B = 8;
    float2 v[B];
    __device__ __forceinline__ void issue(const LENIA_GLOBAL float2 *src, int n, int tid) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
            int t = tid + i * T;
            if (t < n) v[i] = gld2(src + t);
        }
    }
    __device__ __forceinline__ void store(float2 *dst, const LENIA_GLOBAL float2 *src, int n,
                                          int tid) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
            int t = tid + i * T;
            if (t < n) dst[t] = v[i];
        }
        for (int t = tid + B * T; t < n; t += T) dst[t] = gld2(src + t);
    }
};

// Stored half-spectra of a tile's row pairs (kernels kBeg .. kBeg+kn-1) ->
// full-length complex rows Z = Ga + i Gb, one LDS slot of nt rows per kernel,
// with the Hermitian symmetry of each real row's spectrum for kx > W/2.
// gather_issue puts the loads in flight; gather_store writes LDS.
// NG = row_values(T) / 2 + 1 float4 loads per thread.
// Slot geometry kept between the two: zi = LDS index of (slot s, row pair
// p, column kx), kx (-1: no load in this slot).
template <int T, int NG>
__device__ __forceinline__ void gather_issue(float4 (&gv)[NG], int (&zi)[NG], int (&kxv)[NG],
                                             const LENIA_GLOBAL float4 *G4, size_t gStride,
                                             int kBeg, int kn, int H2, int nt, int P, uint32_t mP,
                                             int nm, int W, int tid) {
    const int total = kn * nm;
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        int m = tid + i * T;
        kxv[i] = -1;
        if (m < total) {
            int s = (m >= nm) + (m >= 2 * nm) + (m >= 3 * nm);
            int mm = m - s * nm;
            int kx = (int)fast_div((uint32_t)mm, mP), p = mm - kx * P;
            if (p < nt) {
                gv[i] = gld4(G4 + (size_t)(kBeg + s) * gStride + kx * H2 + p);
                kxv[i] = kx;
                zi[i] = (s * nt + p) * W + kx;
            }
        }
    }
}

template <int NG>
__device__ __forceinline__ void gather_store(float2 *Z, const float4 (&gv)[NG], const int (&zi)[NG],
                                             const int (&kxv)[NG], int W) {
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const int kx = kxv[i];
        if (kx >= 0) {
            float4 g4 = gv[i];
            if (kx == 0 || kx == W / 2) {
                // real columns for a real field: drop the rounding
                // residue instead of leaking it across rows
                g4.y = 0.0f;
                g4.w = 0.0f;
            }
            Z[sw(zi[i])] = make_float2(g4.x - g4.w, g4.y + g4.z);
            if (kx > 0 && kx < W / 2)
                Z[sw(zi[i] + W - 2 * kx)] = make_float2(g4.x + g4.w, g4.z - g4.y);
        }
    }
}

// forward FFT of the tile's (new) row pairs, split into the two real rows'
// half-spectra Fa = (Z[k] + conj Z[W-k]) / 2, Fb = (Z[k] - conj Z[W-k]) / 2i
template <int T>
__device__ __forceinline__ void forward_rows(float2 *Z, float2 *Zt, const FftWorldDev &wd,
                                             const float2 *tw, int rp0, int nt, int tid) {
    const int W = wd.W, H2 = wd.H / 2, P = wd.P;
    const int nm = P * wd.Wh;
    // the spectra may end in either region
    const float2 *Zs = fft_run<false, kMaxV, T>(Z, wd.planW, nt, tw, tid, Zt, true);
    LENIA_GLOBAL float4 *F4 = gp(reinterpret_cast<float4 *>(wd.F)) + rp0;
    for (int m = tid; m < nm; m += T) {
        int kx = (int)fast_div((uint32_t)m, wd.mP), p = m - kx * P;
        if (p >= nt) continue;
        float2 z1 = Zs[sw(p * W + kx)];
        float2 z2 = Zs[sw(p * W + (kx == 0 ? 0 : W - kx))];
        gst4(F4 + kx * H2 + p, make_float4(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y),
                                           0.5f * (z1.y + z2.y), 0.5f * (z2.x - z1.x)));
    }
}

// MODE 0: forward only (state -> F); 1: step (G -> growth -> state -> F);
// 2: potential of kernel kSel only (G_kSel -> out), no state change.
// KPAR: the K inverse transforms of a tile run side by side (one LDS slot per
// kernel, spread over the workgroup's waves); otherwise they run one after
// the other with the growth sums kept in LDS.
// EXTRAS: the noise / muJitter / inject terms are compiled in (with all three
// amounts zero both variants compute the same bits).
// LDS: twiddles [W] | KPAR: Z [K][nt][W] ; serial: Z [nt][W] (sums in registers)
template <int MODE, int T, bool EXTRAS, bool KPAR, int V>
__device__ __forceinline__ void rows_tile(const FftWorldDev *__restrict__ worlds, int2 tile,
                                          bool newWorld, uint32_t stepOff, int kSel,
                                          float *__restrict__ out, float2 *smem, int tid) {
    constexpr int kV = V, kG = kV / 2 + 1;
    // the descriptor through the restrict table pointer: its fields may be
    // reloaded instead of held in registers across the tile
    const FftWorldDev &wd = worlds[tile.x];
    const int W = wd.W, H = wd.H, Wh = wd.Wh, P = wd.P, H2 = H / 2;
    const int rp0 = tile.y * P;
    const int nt = min(P, H2 - rp0);
    // magics read once: a field read inside an unrolled per-value loop is
    // read again in every iteration, each time after a full wait (the serial
    // forms read mW where they use it: held, it let the compiler keep every
    // value's indices live across their kernel loop, 276 -> 349 VGPRs)
    const uint32_t mP = wd.mP, mW = wd.mW;
    float2 *tw = smem;
    float2 *Z = smem + W;
    // the twin transform region, when the launch has one (rows_twin)
    float2 *Zt = rows_twin(W, P, wd.K, KPAR, T) ? Z + rows_region(W, P, wd.K, KPAR) : nullptr;
    LENIA_GLOBAL float *Arow = gp(wd.A) + (size_t)(2 * rp0) * W;  // first row of this tile
    const int nq = nt * W;
    const int nm = P * Wh;  // (kx, p) slots, p fastest; p >= nt are idle
    // the gather's operands are read here, ahead of the twiddle copy, so the
    // scalar loads of the tile's prologue go out together
    const LENIA_GLOBAL float4 *G4 = gp(reinterpret_cast<const float4 *>(wd.G)) + rp0;
    const size_t gStride = (size_t)Wh * H2;
    const int nK = (int)wd.g.nKernels;
    const LENIA_GLOBAL float2 *twg = gp(reinterpret_cast<const float2 *>(wd.twW));
    // an empty statement that takes these operands in scalar registers: their
    // scalar loads are issued together here and waited for once, where the
    // compiler would otherwise issue each next to its first use, one wait
    // after another before the gather starts (side-by-side step kernels only:
    // in the serial form it raised the registers a thread holds)
    if constexpr (KPAR && MODE == 1)
        asm volatile("" ::"s"(wd.G), "s"(wd.A), "s"(wd.twW), "s"(W), "s"(P), "s"(nK), "s"(mW), "s"(mP));
    TwCopy<T> twc;
    if (newWorld) twc.issue(twg, W, tid);

    if (MODE == 0) {
        if (newWorld) twc.store(tw, twg, W, tid);
        for (int q = tid; q < nq; q += T) {
            int p = (int)fast_div((uint32_t)q, KPAR ? mW : wd.mW), x = q - p * W;
            int ia = 2 * p * W + x;
            Z[sw(q)] = make_float2(Arow[ia], Arow[ia + W]);
        }
        __syncthreads();
        forward_rows<T>(Z, Zt, wd, tw, rp0, nt, tid);
        return;  // one tile per workgroup: no LDS reuse after it
    }
    const GrowthParams &g = wd.g;
    const GrowthRegs gr = growth_regs(g);
    LENIA_GLOBAL float *const saveU = gp(wd.saveU);
    LENIA_GLOBAL float *const saveA = gp(wd.saveA);
    const StepClock ck = clock_at(wd.clock, stepOff);
    const uint32_t seed = clock_seed(ck, g.moduleSeed);
    float ar0[kV], ar1[kV];
    if (MODE == 1) {
#pragma unroll
        for (int i = 0; i < kV; ++i) {
            int q = tid + i * T;
            if (q < nq) {
                int p = (int)fast_div((uint32_t)q, KPAR ? mW : wd.mW), x = q - p * W;
                int ia = 2 * p * W + x;
                ar0[i] = Arow[ia];
                ar1[i] = Arow[ia + W];
            }
        }
    }
    if (KPAR || MODE == 2) {
        const int kBeg = MODE == 2 ? kSel : 0;
        const int kn = MODE == 2 ? 1 : nK;
        float4 gv[kG];
        int zi[kG], kxv[kG];
        gather_issue<T, kG>(gv, zi, kxv, G4, gStride, kBeg, kn, H2, nt, P, mP, nm, W, tid);
        if (newWorld) twc.store(tw, twg, W, tid);
        gather_store<kG>(Z, gv, zi, kxv, W);
        __syncthreads();
        // the potentials may end in either region (U); the new state goes to Z
        const float2 *U = fft_run<true, kMaxV, T>(Z, wd.planW, kn * nt, tw, tid, Zt, true);
        if (MODE == 2) {
            for (int q = tid; q < nq; q += T) {
                int p = (int)fast_div((uint32_t)q, mW), x = q - p * W;
                int ia = 2 * p * W + x;
                float2 z = U[sw(q)];
                LENIA_GLOBAL float *o = gp(out) + (size_t)(2 * rp0) * W;
                o[ia] = z.x;
                o[ia + W] = z.y;
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < kV; ++i) {
            int q = tid + i * T;
            if (q < nq) {
                int p = (int)fast_div((uint32_t)q, mW), x = q - p * W;
                int ya = 2 * (rp0 + p);
                float ja = 0.0f, jb = 0.0f;
                if (EXTRAS) {
                    ja = cell_jitter(g, x, ya, seed, ck.frame);
                    jb = cell_jitter(g, x, ya + 1, seed, ck.frame);
                }
                float2 z[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (k < kn) z[k] = U[sw(k * nq + q)];
                float sa = 0.0f, sb = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (k < kn) {
                        sa += growth_term_r(gr, k, z[k].x, ja);
                        sb += growth_term_r(gr, k, z[k].y, jb);
                        if (saveU) {
                            LENIA_GLOBAL float *su = saveU + (size_t)k * W * H + (size_t)ya * W + x;
                            su[0] = z[k].x;
                            su[W] = z[k].y;
                        }
                    }
                }
                if (saveA) {
                    LENIA_GLOBAL float *sa0 = saveA + (size_t)ya * W + x;
                    sa0[0] = ar0[i];
                    sa0[W] = ar1[i];
                }
                if (EXTRAS) {
                    ar0[i] = growth_finish(g, ar0[i], sa, x, ya, W, H, seed, ck.frame);
                    ar1[i] = growth_finish(g, ar1[i], sb, x, ya + 1, W, H, seed, ck.frame);
                } else {
                    ar0[i] = growth_plain_r(gr, ar0[i], sa);
                    ar1[i] = growth_plain_r(gr, ar1[i], sb);
                }
            }
        }
        // stores after all LDS reads (no read waits behind a store)
#pragma unroll
        for (int i = 0; i < kV; ++i) {
            int q = tid + i * T;
            if (q < nq) {
                int p = (int)fast_div((uint32_t)q, mW), x = q - p * W;
                int ia = 2 * p * W + x;
                Arow[ia] = ar0[i];
                Arow[ia + W] = ar1[i];
                Z[sw(q)] = make_float2(ar0[i], ar1[i]);
            }
        }
    } else {
        float sa[kV], sb[kV];
#pragma unroll
        for (int i = 0; i < kV; ++i) sa[i] = sb[i] = 0.0f;
        float4 gv[kG];
        int zi[kG], kxv[kG];
        gather_issue<T, kG>(gv, zi, kxv, G4, gStride, 0, 1, H2, nt, P, mP, nm, W, tid);
        if (newWorld) twc.store(tw, twg, W, tid);
        for (int k = 0; k < (int)g.nKernels; ++k) {
            gather_store<kG>(Z, gv, zi, kxv, W);
            // the next kernel's spectrum loads fly during this transform
            if (k + 1 < (int)g.nKernels)
                gather_issue<T, kG>(gv, zi, kxv, G4, gStride, k + 1, 1, H2, nt, P, mP, nm, W, tid);
            __syncthreads();
            // table-driven transform inside the runtime kernel loop: the
            // compile-time one would keep its addresses live across it
            fft_lds<true, kMaxV, T>(Z, wd.planW, nt, tw, tid);
#pragma unroll
            for (int i = 0; i < kV; ++i) {
                int q = tid + i * T;
                if (q < nq) {
                    int p = (int)fast_div((uint32_t)q, wd.mW), x = q - p * W;
                    int ya = 2 * (rp0 + p);
                    float2 z = Z[sw(q)];
                    float ja = 0.0f, jb = 0.0f;
                    if (EXTRAS) {
                        ja = cell_jitter(g, x, ya, seed, ck.frame);
                        jb = cell_jitter(g, x, ya + 1, seed, ck.frame);
                    }
                    sa[i] += growth_term_r(gr, k, z.x, ja);
                    sb[i] += growth_term_r(gr, k, z.y, jb);
                    if (saveU) {
                        LENIA_GLOBAL float *su = saveU + (size_t)k * W * H + (size_t)ya * W + x;
                        su[0] = z.x;
                        su[W] = z.y;
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < kV; ++i) {
            int q = tid + i * T;
            if (q < nq) {
                int p = (int)fast_div((uint32_t)q, wd.mW), x = q - p * W;
                int ya = 2 * (rp0 + p);
                float a2, b2;
                if (saveA) {
                    LENIA_GLOBAL float *sa0 = saveA + (size_t)ya * W + x;
                    sa0[0] = ar0[i];
                    sa0[W] = ar1[i];
                }
                if (EXTRAS) {
                    a2 = growth_finish(g, ar0[i], sa[i], x, ya, W, H, seed, ck.frame);
                    b2 = growth_finish(g, ar1[i], sb[i], x, ya + 1, W, H, seed, ck.frame);
                } else {
                    a2 = growth_plain_r(gr, ar0[i], sa[i]);
                    b2 = growth_plain_r(gr, ar1[i], sb[i]);
                }
                int ia = 2 * p * W + x;
                Arow[ia] = a2;
                Arow[ia + W] = b2;
                Z[sw(q)] = make_float2(a2, b2);
            }
        }
    }
    __syncthreads();
    forward_rows<T>(Z, Zt, wd, tw, rp0, nt, tid);
}

// Waves per SIMD the register allocation must leave room for: the 256-thread
// 12-cell step kernel (the batched sweep's row pass) fits 96 VGPRs without
// spilling, which lets a CU hold five of its workgroups instead of four
// (105 VGPRs unconstrained); at 80 VGPRs (six) it spills.
constexpr int rows_waves_per_eu(int T, int V) { return T == 256 && V == kMidValues ? 5 : 1; }

// One tile per workgroup (grid = tiles)
template <int MODE, int T, bool EXTRAS, bool KPAR, int V = row_values(T)>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(rows_waves_per_eu(T, V)))) void k_fft_rows(const FftWorldDev *__restrict__ worlds,
                                                const int2 *__restrict__ tiles, int nTiles,
                                                uint32_t stepOff, int kSel,
                                                float *__restrict__ out) {
    extern __shared__ float2 smem[];
    if ((int)blockIdx.x < nTiles) {
        const int2 tile = step_tile(tiles, tile_of((int)blockIdx.x, nTiles));
        rows_tile<MODE, T, EXTRAS, KPAR, V>(worlds, tile, true, stepOff, kSel, out, smem,
                                            (int)threadIdx.x);
    }
}

// ---------------------------------------------------------------- column pass
// LDS: twiddles [H] | X [C][H] | KPAR: Y [K][C][H]
// V: column values per thread held in registers (kMaxV; a tile of at most
// kLeanCols per thread runs the variant sized for it)
template <int T, bool KPAR, int V>
__device__ __forceinline__ void cols_tile(const FftWorldDev *__restrict__ worlds, int2 tile,
                                          bool newWorld, float2 *smem, int tid) {
    const FftWorldDev &wd = worlds[tile.x];
    const int H = wd.H, Wh = wd.Wh;
    const int c0 = tile.y * wd.C;
    const int nt = min(wd.C, Wh - c0);
    const int ne = nt * H;
    const int K = (int)wd.g.nKernels;
    float2 *tw = smem;
    float2 *X = smem + H;
    const LENIA_GLOBAL float2 *twg = gp(reinterpret_cast<const float2 *>(wd.twH));
    // the operands of the column and spectrum loads, loaded together and
    // waited for once (see rows_tile)
    if constexpr (KPAR)
        asm volatile("" ::"s"(wd.F), "s"(wd.S), "s"(wd.G), "s"(wd.twH), "s"(H), "s"(Wh), "s"(wd.C), "s"(K));
    TwCopy<T> twc;
    if (newWorld) twc.issue(twg, H, tid);
    const LENIA_GLOBAL float2 *F = gp(reinterpret_cast<const float2 *>(wd.F)) + (size_t)c0 * H;
    const LENIA_GLOBAL float *S0 = gp(wd.S) + (size_t)c0 * H;
    const size_t sStride = (size_t)Wh * H;
    LENIA_GLOBAL float2 *G0 = gp(reinterpret_cast<float2 *>(wd.G)) + (size_t)c0 * H;
    {
        float2 fv[V];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            int e = tid + i * T;
            if (e < ne) fv[i] = gld2(F + e);
        }
        if (newWorld) twc.store(tw, twg, H, tid);
#pragma unroll
        for (int i = 0; i < V; ++i) {
            int e = tid + i * T;
            if (e < ne) X[sw(e)] = fv[i];
        }
    }
    if (KPAR) {
        float2 *Y = X + lds_span(ne);
        const int nk = K * ne;
        float sr[V];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            int e = tid + i * T;
            if (e < nk) {
                int k = (e >= ne) + (e >= 2 * ne) + (e >= 3 * ne), c = e - k * ne;
                // (k, c) offsets fit 32 bits (K Wh H < 2^31): no 64-bit multiply
                sr[i] = S0[k * (int)sStride + c];
            }
        }
        __syncthreads();
        // Y is free until the scaled copies: the forward transform's twin region
        fft_run<false, kMaxV, T>(X, wd.planH, nt, tw, tid, Y);
        if (wd.saveAhat) {
            LENIA_GLOBAL float2 *sv = gp(reinterpret_cast<float2 *>(wd.saveAhat)) + (size_t)c0 * H;
            for (int e = tid; e < ne; e += T) gst2(sv + e, X[sw(e)]);
        }
        // all reads of the forward spectrum first, then the scaled copies
        float2 xv[V];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            int e = tid + i * T;
            if (e < nk) {
                int k = (e >= ne) + (e >= 2 * ne) + (e >= 3 * ne), c = e - k * ne;
                xv[i] = X[sw(c)];
            }
        }
#pragma unroll
        for (int i = 0; i < V; ++i) {
            int e = tid + i * T;
            if (e < nk) Y[sw(e)] = cscale(xv[i], sr[i]);
        }
        __syncthreads();
        fft_run<true, kMaxV, T>(Y, wd.planH, K * nt, tw, tid);
        for (int e = tid; e < nk; e += T) {
            int k = (e >= ne) + (e >= 2 * ne) + (e >= 3 * ne), c = e - k * ne;
            gst2(G0 + (k * (int)sStride + c), Y[sw(e)]);
        }
        return;
    }
    float sr[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        int e = tid + i * T;
        if (e < ne) sr[i] = S0[e];
    }
    __syncthreads();
    fft_run<false, kMaxV, T>(X, wd.planH, nt, tw, tid);
    float2 xs[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
        int e = tid + i * T;
        xs[i] = e < ne ? X[sw(e)] : make_float2(0.f, 0.f);
    }
    if (wd.saveAhat) {
        LENIA_GLOBAL float2 *sv = gp(reinterpret_cast<float2 *>(wd.saveAhat)) + (size_t)c0 * H;
#pragma unroll
        for (int i = 0; i < V; ++i) {
            int e = tid + i * T;
            if (e < ne) gst2(sv + e, xs[i]);
        }
    }
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            int e = tid + i * T;
            if (e < ne) X[sw(e)] = cscale(xs[i], sr[i]);
        }
        if (k + 1 < K) {
            const LENIA_GLOBAL float *Sn = S0 + (size_t)(k + 1) * sStride;
#pragma unroll
            for (int i = 0; i < V; ++i) {
                int e = tid + i * T;
                if (e < ne) sr[i] = Sn[e];
            }
        }
        __syncthreads();
        fft_lds<true, kMaxV, T>(X, wd.planH, nt, tw, tid);
        LENIA_GLOBAL float2 *G = G0 + (size_t)k * sStride;
        for (int e = tid; e < ne; e += T) gst2(G + e, X[sw(e)]);
        if (k + 1 < K) __syncthreads();  // X is refilled for the next kernel
    }
}

// One tile per workgroup (grid = tiles)
template <int T, bool KPAR, int V = kMaxV>
__global__ __launch_bounds__(T) void k_fft_cols(const FftWorldDev *__restrict__ worlds,
                                                const int2 *__restrict__ tiles, int nTiles) {
    extern __shared__ float2 smem[];
    if ((int)blockIdx.x < nTiles) {
        const int2 tile = step_tile(tiles, tile_of((int)blockIdx.x, nTiles));
        cols_tile<T, KPAR, V>(worlds, tile, true, smem, (int)threadIdx.x);
    }
}

// ---------------------------------------------------------------- launch templates
template <int T, bool EX, bool KP, int V = row_values(T)>
void rows_launch(int mode, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                        int grid, size_t lds, uint32_t stepOff, int kSel, float *out,
                        hipStream_t s) {
    if (mode == 0)
        launch_timed(k_fft_rows<0, T, false, false>, dim3(grid), dim3(T), lds, s, worlds, tiles,
                     nTiles, stepOff, kSel, out);
    else if (mode == 1)
        launch_timed(k_fft_rows<1, T, EX, KP, V>, dim3(grid), dim3(T), lds, s, worlds, tiles, nTiles,
                     stepOff, kSel, out);
    else
        launch_timed(k_fft_rows<2, T, false, true>, dim3(grid), dim3(T), lds, s, worlds, tiles,
                     nTiles, stepOff, kSel, out);
}

template <int T>
void rows_dispatch(int mode, bool extras, bool kpar, const FftWorldDev *worlds,
                          const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff,
                          int kSel, float *out, hipStream_t s) {
    if (extras) {
        if (kpar) rows_launch<T, true, true>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
        else rows_launch<T, true, false>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
    } else {
        if (kpar) rows_launch<T, false, true>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
        else rows_launch<T, false, false>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
    }
}

template <int T>
void cols_dispatch(bool kpar, bool lean, const FftWorldDev *worlds, const int2 *tiles,
                   int nTiles, int grid, size_t lds, hipStream_t s) {
    if (kpar && lean)
        launch_timed(k_fft_cols<T, true, kLeanCols>, dim3(grid), dim3(T), lds, s, worlds, tiles,
                     nTiles);
    else if (kpar)
        launch_timed(k_fft_cols<T, true>, dim3(grid), dim3(T), lds, s, worlds, tiles, nTiles);
    else
        launch_timed(k_fft_cols<T, false>, dim3(grid), dim3(T), lds, s, worlds, tiles, nTiles);
}

// per-workgroup-size entry points (k_fft_pass<T>.hip)
#define LENIA_FFT_PASS_DECL(T)                                                                  \
    void fft_rows_pass_##T(int mode, bool extras, bool kpar, const FftWorldDev *worlds,         \
                           const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff, \
                           int kSel, float *out, hipStream_t s);                    \
    void fft_cols_pass_##T(bool kpar, bool lean, const FftWorldDev *worlds, const int2 *tiles,   \
                           int nTiles, int grid, size_t lds, hipStream_t s);
LENIA_FFT_PASS_DECL(64)
LENIA_FFT_PASS_DECL(128)
LENIA_FFT_PASS_DECL(256)
LENIA_FFT_PASS_DECL(512)
// 256- / 512-thread step passes holding kMidValues cells per thread (side-by-side transforms)
void fft_rows_step_256_mid(bool extras, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                           int grid, size_t lds, uint32_t stepOff, hipStream_t s);
void fft_rows_step_512_mid(bool extras, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                           int grid, size_t lds, uint32_t stepOff, hipStream_t s);
void fft_rows_pass_1024(int mode, bool extras, bool lean, const FftWorldDev *worlds, const int2 *tiles,
                        int nTiles, int grid, size_t lds, uint32_t stepOff, int kSel, float *out,
                        hipStream_t s);

}  // namespace lenia
