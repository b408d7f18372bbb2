// Wave-resident executor of the FFT-form step passes (gfx950, wave64).
//
// The workgroup-LDS executor (fft_pass.h) spreads the butterflies of a tile's
// transforms over every thread of a workgroup, with a workgroup barrier
// around each stage.  Here one wave owns whole transforms: the 64 lanes split
// into M = 64 / L groups of L lanes, each group holds one length-N transform
// as V = N / L values per lane (value m of lane l is entry l + L m), and the
// stages run in registers.  A stage of radix R is V / R butterflies per lane;
// its inputs are entries j + r N/R of butterfly j = l + L b, which are the
// lane's own values m = b + r V/R, so a stage needs no data movement before
// it; its outputs go through a wave-private LDS region to the next stage's
// layout (S - 1 exchanges for S stages: the last stage's outputs j + r N/R
// are again the lane's own values, and the first stage reads the values as
// loaded from HBM).  No stage waits for another wave, so waves overlap each
// other's LDS latency, and a wave loads its operands straight from HBM into
// the layout its first stage reads.
//
// The butterflies are those of the LDS executor: same radix plan, same
// inputs, twiddles (the plan's stage tables), skip of the k = 0 twiddles,
// Dft<R> and output placement, so both executors give the same bits, and so
// do the gather (Hermitian fill), growth and spectrum split around them.  A
// plan runs here when every radix divides V (wave_lanes).
#pragma once

#include <hip/hip_runtime.h>

#include "fft_device.h"
#include "lenia_device.h"

namespace lenia {

// Order a wave's LDS writes before its following reads (and reads before the
// next writes) at the compiler level; the hardware executes one wave's LDS
// instructions in order, so no wait is needed for lanes of the same wave.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Stockham stage of radix R (span NS, twiddle table at TWOFF) on the
// lane's V values; LAST: outputs stay in registers, else they go to the
// group's LDS region `xb` and come back in entry order.
template <int N, int L, int R, int NS, int TWOFF, bool INV, bool LAST>
__device__ __forceinline__ void wave_stage(float2 *v, float2 *xb, const float2 *tw, int l) {
    constexpr int V = N / L, B = V / R;
    static_assert(V % R == 0, "radix does not divide the values per lane");
    // this stage's twiddles are read first, so their LDS reads are in flight
    // together with the previous exchange's; a butterfly with k = 0 keeps its
    // inputs (a select, the same values as skipping the multiply)
    float2 w[NS > 1 ? B : 1][R - 1];
    if constexpr (NS > 1) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const float2 *t = tw + TWOFF + ((l + L * b) % NS) * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) w[b][r - 1] = t[r - 1];
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = l + L * b;
        const int k = j % NS;
        float2 u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = v[b + r * B];
        if constexpr (NS > 1) {
            const bool tw1 = k != 0;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const float2 t = ctw<INV>(u[r], w[b][r - 1]);
                u[r] = make_float2(tw1 ? t.x : u[r].x, tw1 ? t.y : u[r].y);
            }
        }
        Dft<R, INV>::run(u);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * B] = u[r];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) xb[sw(base + r * NS)] = u[r];
        }
    }
    if constexpr (!LAST) {
        wave_lds_order();
#pragma unroll
        for (int m = 0; m < V; ++m) v[m] = xb[sw(l + L * m)];
        wave_lds_order();
    }
}

template <int N, int L, bool INV, int NS, int TWOFF, int R, int... REST>
__device__ __forceinline__ void wave_fft(float2 *v, float2 *xb, const float2 *tw, int l) {
    wave_stage<N, L, R, NS, TWOFF, INV, sizeof...(REST) == 0>(v, xb, tw, l);
    if constexpr (sizeof...(REST) > 0)
        wave_fft<N, L, INV, NS * R, TWOFF + NS * (R - 1), REST...>(v, xb, tw, l);
}

// LDS entries of one wave's exchange region: M groups of N entries, placed
// i -> i + i/32 (N is a multiple of 32, so group g starts at g (N + N/32))
template <int N>
constexpr int wave_group_span() { return N + N / 32; }

// twiddle tables (n - 1 entries) of the plan -> LDS; every wave waits for them
__device__ __forceinline__ void wave_load_twiddles(float2 *tw, const float *src, int n, int nThreads) {
    const LENIA_GLOBAL float2 *g = gp(reinterpret_cast<const float2 *>(src));
    for (int t = (int)threadIdx.x; t < n - 1; t += nThreads) tw[t] = gld2(g + t);
    __syncthreads();
}

// --------------------------------------------------------------- row pass
// Where a group's row pair lives: row pair p (live: p < H/2; idle groups
// run a valid row pair, pc, and store nothing).
struct WaveRowPair {
    int l, p, pc;
    bool live;
};

// the half-spectra of row pair pc for kernel k: loads in flight into gb
// (value m of lane l is x = l + L m; x > W/2 reads column W - x)
template <int N, int L>
__device__ __forceinline__ void wave_gather_issue(const FftWorldDev &wd, const WaveRowPair &rp, int k,
                                                  float4 (&gb)[N / L]) {
    constexpr int V = N / L, W = N;
    const int H2 = wd.H / 2;
    // uniform base, 32-bit per-lane offsets ((W/2+1) H/2 < 2^31)
    const LENIA_GLOBAL float4 *G4 =
        gp(reinterpret_cast<const float4 *>(wd.G)) + (size_t)k * (W / 2 + 1) * H2;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        const int x = rp.l + L * m;
        const bool lo = m < V / 2 || (m == V / 2 && rp.l == 0);  // x <= W/2
        gb[m] = gld4(G4 + ((lo ? x : W - x) * H2 + rp.pc));
    }
}

// Z = Ga + i Gb with the Hermitian fill for x > W/2
template <int N, int L>
__device__ __forceinline__ void wave_gather_fill(const WaveRowPair &rp, const float4 (&gb)[N / L],
                                                 float2 *v) {
    constexpr int V = N / L;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        const bool lo = m < V / 2 || (m == V / 2 && rp.l == 0);
        float4 g4 = gb[m];
        if ((m == 0 || m == V / 2) && rp.l == 0) {  // kx 0 and W/2: real columns
            g4.y = 0.0f;
            g4.w = 0.0f;
        }
        v[m] = lo ? make_float2(g4.x - g4.w, g4.y + g4.z) : make_float2(g4.x + g4.w, g4.z - g4.y);
    }
}

// kernel k's growth terms of the potentials v (rows ya, ya+1) and the saved
// potentials of the reverse pass: mode 1 w b, mode 0 w (2 b - 1)
template <bool EXTRAS, int N, int L>
__device__ __forceinline__ void wave_growth_terms(const FftWorldDev &wd, const WaveRowPair &rp, int k,
                                                  const float2 *v, uint32_t seed, uint32_t frame,
                                                  float *ta, float *tb) {
    constexpr int V = N / L, W = N;
    const GrowthParams &gpar = wd.g;
    const float muk = gpar.mu[k], sck = gpar.bellScale[k], wk = gpar.weight[k];
    const bool asym = gpar.growthMode == 1u;
    const int ya = 2 * rp.pc;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        float ja = 0.0f, jb = 0.0f;
        if (EXTRAS) {
            const int x = rp.l + L * m;
            ja = cell_jitter(gpar, x, ya, seed, frame);
            jb = cell_jitter(gpar, x, ya + 1, seed, frame);
        }
        const float ba = bell_of(v[m].x, muk + ja, sck);
        const float bb = bell_of(v[m].y, muk + jb, sck);
        ta[m] = wk * (asym ? ba : 2.0f * ba - 1.0f);
        tb[m] = wk * (asym ? bb : 2.0f * bb - 1.0f);
    }
    LENIA_GLOBAL float *const saveU = gp(wd.saveU);
    if (saveU && rp.live) {
        LENIA_GLOBAL float *su = saveU + (size_t)k * W * wd.H + (size_t)ya * W + rp.l;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            su[L * m] = v[m].x;
            su[L * m + W] = v[m].y;
        }
    }
}

// growth finish from the summed terms, the state update, the forward
// transform of the two new rows and their half-spectra
// Fa = (Z[k] + conj Z[W-k]) / 2, Fb = (Z[k] - conj Z[W-k]) / 2i into F
template <bool STEP, bool EXTRAS, int N, int L, int... RS>
__device__ __forceinline__ void wave_rows_tail(const FftWorldDev &wd, const WaveRowPair &rp,
                                               const float *ar0, const float *ar1, const float *sa,
                                               const float *sb, uint32_t seed, uint32_t frame,
                                               float2 *v, float2 *xb, const float2 *tw) {
    constexpr int V = N / L, W = N;
    const int l = rp.l, H = wd.H, H2 = H / 2;
    if constexpr (STEP) {
        LENIA_GLOBAL float *Arow = gp(wd.A) + (size_t)(2 * rp.pc) * W;
        const GrowthRegs gr = growth_regs(wd.g);
        const int ya = 2 * rp.pc;
        LENIA_GLOBAL float *const saveA = gp(wd.saveA);
        if (saveA && rp.live) {
            LENIA_GLOBAL float *s0 = saveA + (size_t)ya * W + l;
#pragma unroll
            for (int m = 0; m < V; ++m) {
                s0[L * m] = ar0[m];
                s0[L * m + W] = ar1[m];
            }
        }
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const int x = l + L * m;
            float a2, b2;
            if (EXTRAS) {
                a2 = growth_finish(wd.g, ar0[m], sa[m], x, ya, W, H, seed, frame);
                b2 = growth_finish(wd.g, ar1[m], sb[m], x, ya + 1, W, H, seed, frame);
            } else {
                a2 = growth_plain_r(gr, ar0[m], sa[m]);
                b2 = growth_plain_r(gr, ar1[m], sb[m]);
            }
            if (rp.live) {
                Arow[x] = a2;
                Arow[x + W] = b2;
            }
            v[m] = make_float2(a2, b2);
        }
    }
    wave_fft<N, L, false, 1, 0, RS...>(v, xb, tw, l);
#pragma unroll
    for (int m = 0; m < V; ++m) xb[sw(l + L * m)] = v[m];
    wave_lds_order();
    LENIA_GLOBAL float4 *F4 = gp(reinterpret_cast<float4 *>(wd.F));
#pragma unroll
    for (int m = 0; m <= V / 2; ++m) {
        const int kx = l + L * m;
        if (m < V / 2 || l == 0) {
            const float2 z1 = v[m];
            const float2 z2 = xb[sw(kx == 0 ? 0 : W - kx)];
            if (rp.live)
                gst4(F4 + (kx * H2 + rp.pc), make_float4(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y),
                                                         0.5f * (z1.y + z2.y), 0.5f * (z2.x - z1.x)));
        }
    }
}

// MODE 0: forward row FFT of the state into F; 1: step (G_k -> potentials ->
// growth -> state -> F); 2: potential of kernel kSel into `out`.
// Tile: P = NW M row pairs; wave w, group g runs row pair tile.y P + w M + g,
// all kernels one after the other.
template <int MODE, bool EXTRAS, int N, int L, int... RS>
__device__ __forceinline__ void wave_rows(const FftWorldDev &wd, int2 tile, uint32_t stepOff,
                                          int kSel, float *__restrict__ out, const float2 *tw,
                                          float2 *xbw, int wave, int lane) {
    constexpr int V = N / L, M = 64 / L, W = N;
    WaveRowPair rp;
    rp.l = lane % L;
    rp.p = tile.y * wd.P + wave * M + lane / L;
    rp.live = rp.p < wd.H / 2;
    rp.pc = rp.live ? rp.p : wd.H / 2 - 1;
    float2 *xb = xbw + (lane / L) * wave_group_span<N>();
    LENIA_GLOBAL float *Arow = gp(wd.A) + (size_t)(2 * rp.pc) * W;
    float2 v[V];
    float4 gb[V];
    if constexpr (MODE == 2) {
        wave_gather_issue<N, L>(wd, rp, kSel, gb);
        wave_gather_fill<N, L>(rp, gb, v);
        wave_fft<N, L, true, 1, 0, RS...>(v, xb, tw, rp.l);
        if (rp.live) {
            LENIA_GLOBAL float *o = gp(out) + (size_t)(2 * rp.p) * W;
#pragma unroll
            for (int m = 0; m < V; ++m) {
                o[rp.l + L * m] = v[m].x;
                o[rp.l + L * m + W] = v[m].y;
            }
        }
    } else if constexpr (MODE == 0) {
#pragma unroll
        for (int m = 0; m < V; ++m) v[m] = make_float2(Arow[rp.l + L * m], Arow[rp.l + L * m + W]);
        wave_rows_tail<false, false, N, L, RS...>(wd, rp, nullptr, nullptr, nullptr, nullptr, 0u, 0u, v,
                                                  xb, tw);
    } else {
        const StepClock ck = clock_at(wd.clock, stepOff);
        const uint32_t seed = clock_seed(ck, wd.g.moduleSeed);
        // the state is read before the kernel loop when V <= 8 (its loads
        // overlap the transforms), after it otherwise (registers); V <= 16:
        // the next kernel's spectra are requested before each transform
        constexpr bool kEarlyA = V <= 8, kPrefetch = V <= 16;
        float ar0[V], ar1[V], sa[V], sb[V], ta[V], tb[V];
        if (kPrefetch) wave_gather_issue<N, L>(wd, rp, 0, gb);
#pragma unroll
        for (int m = 0; m < V; ++m) {
            if (kEarlyA) {
                ar0[m] = Arow[rp.l + L * m];
                ar1[m] = Arow[rp.l + L * m + W];
            }
            sa[m] = sb[m] = 0.0f;
        }
        const int K = (int)wd.g.nKernels;
#pragma unroll 1
        for (int k = 0; k < K; ++k) {
            if (!kPrefetch) wave_gather_issue<N, L>(wd, rp, k, gb);
            wave_gather_fill<N, L>(rp, gb, v);
            if (kPrefetch && k + 1 < K) wave_gather_issue<N, L>(wd, rp, k + 1, gb);
            wave_fft<N, L, true, 1, 0, RS...>(v, xb, tw, rp.l);
            wave_growth_terms<EXTRAS, N, L>(wd, rp, k, v, seed, ck.frame, ta, tb);
#pragma unroll
            for (int m = 0; m < V; ++m) {
                sa[m] += ta[m];
                sb[m] += tb[m];
            }
        }
        if (!kEarlyA) {
#pragma unroll
            for (int m = 0; m < V; ++m) {
                ar0[m] = Arow[rp.l + L * m];
                ar1[m] = Arow[rp.l + L * m + W];
            }
        }
        wave_rows_tail<true, EXTRAS, N, L, RS...>(wd, rp, ar0, ar1, sa, sb, seed, ck.frame, v, xb, tw);
    }
}

// Kernel-split step of M row pairs by kMaxKernels waves: wave k gathers and
// inverse-transforms kernel k (k < K) and leaves its growth terms in LDS
// (`terms`: kernels 1..3, V x 64 float2 each); after one barrier wave 0 adds
// them to its own in kernel order, finishes the growth and runs the forward
// transform and split.  The other waves end at the barrier.
template <bool EXTRAS, int N, int L, int... RS>
__device__ __forceinline__ void wave_rows_split(const FftWorldDev &wd, int2 tile, uint32_t stepOff,
                                                const float2 *tw, float2 *xbw, float2 *terms,
                                                int wave, int lane) {
    constexpr int V = N / L, W = N;
    WaveRowPair rp;
    rp.l = lane % L;
    rp.p = tile.y * wd.P + lane / L;
    rp.live = rp.p < wd.H / 2;
    rp.pc = rp.live ? rp.p : wd.H / 2 - 1;
    float2 *xb = xbw + (lane / L) * wave_group_span<N>();
    const StepClock ck = clock_at(wd.clock, stepOff);
    const uint32_t seed = clock_seed(ck, wd.g.moduleSeed);
    const int K = (int)wd.g.nKernels;
    float2 v[V];
    float ar0[V], ar1[V], ta[V], tb[V];
    if (wave == 0) {
        LENIA_GLOBAL float *Arow = gp(wd.A) + (size_t)(2 * rp.pc) * W;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            ar0[m] = Arow[rp.l + L * m];
            ar1[m] = Arow[rp.l + L * m + W];
        }
    }
    if (wave < K) {
        float4 gb[V];
        wave_gather_issue<N, L>(wd, rp, wave, gb);
        wave_gather_fill<N, L>(rp, gb, v);
        wave_fft<N, L, true, 1, 0, RS...>(v, xb, tw, rp.l);
        wave_growth_terms<EXTRAS, N, L>(wd, rp, wave, v, seed, ck.frame, ta, tb);
        if (wave > 0) {
            float2 *t = terms + (wave - 1) * V * 64 + lane;
#pragma unroll
            for (int m = 0; m < V; ++m) t[m * 64] = make_float2(ta[m], tb[m]);
        }
    }
    __syncthreads();
    if (wave != 0) return;
    float sa[V], sb[V];
#pragma unroll
    for (int m = 0; m < V; ++m) {
        sa[m] = 0.0f + ta[m];
        sb[m] = 0.0f + tb[m];
    }
    for (int k = 1; k < K; ++k) {
        const float2 *t = terms + (k - 1) * V * 64 + lane;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const float2 q = t[m * 64];
            sa[m] += q.x;
            sb[m] += q.y;
        }
    }
    wave_rows_tail<true, EXTRAS, N, L, RS...>(wd, rp, ar0, ar1, sa, sb, seed, ck.frame, v, xb, tw);
}

// --------------------------------------------------------------- column pass
// Tile: C = NW M columns; wave w, group g runs column tile.y C + w M + g.
// The record fields a column tile reads, requested together
// with the plan code before the kernel's plan switch: the column's loads then
// wait for one scalar round trip instead of two (the code, then the fields
// inside the case).
struct ColsFields {
    const float *S, *F, *twH;
    float *G, *saveAhat;
    int Wh, C, K;
};

__device__ __forceinline__ ColsFields cols_fields(const FftWorldDev &wd) {
    ColsFields f;
    f.S = wd.S;
    f.F = wd.F;
    f.twH = wd.twH;
    f.G = wd.G;
    f.saveAhat = wd.saveAhat;
    f.Wh = wd.Wh;
    f.C = wd.C;
    f.K = (int)wd.g.nKernels;
    return f;
}

template <int N, int L, int... RS>
__device__ __forceinline__ void wave_cols(const ColsFields &cf, int2 tile, float2 *tw,
                                          float2 *xbw, int wave, int lane, int nThreads) {
    constexpr int V = N / L, M = 64 / L, H = N;
    const int g = lane / L, l = lane % L;
    float2 *xb = xbw + g * wave_group_span<N>();
    const int Wh = cf.Wh;
    const int c = tile.y * cf.C + wave * M + g;
    const bool live = c < Wh;
    const int cc = live ? c : Wh - 1;
    const int off = cc * H + l;  // (k, column) offsets fit 32 bits (K Wh H < 2^31)
    const int sStride = Wh * H;
    const LENIA_GLOBAL float2 *F = gp(reinterpret_cast<const float2 *>(cf.F)) + off;
    const LENIA_GLOBAL float *S = gp(cf.S) + off;
    float2 x[V];
    float sv[V];  // the next kernel's spectrum, requested before each transform
#pragma unroll
    for (int m = 0; m < V; ++m) x[m] = gld2(F + L * m);
#pragma unroll
    for (int m = 0; m < V; ++m) sv[m] = S[L * m];
    // the twiddle table, its loads in flight together with the column's
    wave_load_twiddles(tw, cf.twH, H, nThreads);
    wave_fft<N, L, false, 1, 0, RS...>(x, xb, tw, l);
    if (cf.saveAhat && live) {
        LENIA_GLOBAL float2 *sva = gp(reinterpret_cast<float2 *>(cf.saveAhat)) + off;
#pragma unroll
        for (int m = 0; m < V; ++m) gst2(sva + L * m, x[m]);
    }
    LENIA_GLOBAL float2 *G = gp(reinterpret_cast<float2 *>(cf.G)) + off;
    const int K = cf.K;
#pragma unroll 1
    for (int k = 0; k < K; ++k) {
        float2 y[V];
#pragma unroll
        for (int m = 0; m < V; ++m) y[m] = cscale(x[m], sv[m]);
        if (k + 1 < K) {
#pragma unroll
            for (int m = 0; m < V; ++m) sv[m] = S[(k + 1) * sStride + L * m];
        }
        wave_fft<N, L, true, 1, 0, RS...>(y, xb, tw, l);
        if (live) {
#pragma unroll
            for (int m = 0; m < V; ++m) gst2(G + (k * sStride + L * m), y[m]);
        }
    }
}

// Kernel-split column pass of M columns by kMaxKernels waves: wave 0 runs
// the forward transforms and leaves the spectra in LDS (`xs`, V x 64), after
// one barrier wave k < K multiplies by kernel k's spectrum, transforms back
// and stores G_k.
template <int N, int L, int... RS>
__device__ __forceinline__ void wave_cols_split(const ColsFields &cf, int2 tile, float2 *tw,
                                                float2 *xbw, float2 *xs, int wave, int lane) {
    constexpr int V = N / L, H = N;
    const int g = lane / L, l = lane % L;
    float2 *xb = xbw + g * wave_group_span<N>();
    const int Wh = cf.Wh;
    const int c = tile.y * cf.C + g;
    const bool live = c < Wh;
    const int cc = live ? c : Wh - 1;
    const int off = cc * H + l;
    const int sStride = Wh * H;
    const int K = cf.K;
    float2 x[V];
    float sv[V];
    if (wave < K) {
        const LENIA_GLOBAL float *S = gp(cf.S) + off + wave * sStride;
#pragma unroll
        for (int m = 0; m < V; ++m) sv[m] = S[L * m];
    }
    if (wave == 0) {
        const LENIA_GLOBAL float2 *F = gp(reinterpret_cast<const float2 *>(cf.F)) + off;
#pragma unroll
        for (int m = 0; m < V; ++m) x[m] = gld2(F + L * m);
    }
    // the twiddle table, its loads in flight together with the column's
    wave_load_twiddles(tw, cf.twH, H, 64 * kMaxKernels);
    if (wave == 0) {
        wave_fft<N, L, false, 1, 0, RS...>(x, xb, tw, l);
#pragma unroll
        for (int m = 0; m < V; ++m) xs[m * 64 + lane] = x[m];
        if (cf.saveAhat && live) {
            LENIA_GLOBAL float2 *sva = gp(reinterpret_cast<float2 *>(cf.saveAhat)) + off;
#pragma unroll
            for (int m = 0; m < V; ++m) gst2(sva + L * m, x[m]);
        }
    }
    __syncthreads();
    if (wave >= K) return;
    float2 y[V];
#pragma unroll
    for (int m = 0; m < V; ++m) y[m] = cscale(xs[m * 64 + lane], sv[m]);
    wave_fft<N, L, true, 1, 0, RS...>(y, xb, tw, l);
    if (live) {
        LENIA_GLOBAL float2 *G = gp(reinterpret_cast<float2 *>(cf.G)) + off + wave * sStride;
#pragma unroll
        for (int m = 0; m < V; ++m) gst2(G + L * m, y[m]);
    }
}

// --------------------------------------------------------------- kernels
// The plans with a wave kernel: (code, N, L, radices).  wave_lanes() on the
// host lists the same table.
#define LENIA_WAVE_PLANS(X)                 \
    X(kPlan64, 64, 8, 8, 8)                 \
    X(kPlan128, 128, 8, 16, 8)              \
    X(kPlan256, 256, 16, 16, 16)            \
    X(kPlan256b, 256, 32, 8, 8, 4)          \
    X(kPlan512, 512, 64, 8, 8, 8)           \
    X(kPlan1024, 1024, 64, 16, 8, 8)        \
    X(kPlan96, 96, 4, 8, 4, 3)              \
    X(kPlan192, 192, 8, 8, 8, 3)            \
    X(kPlan384c, 384, 16, 8, 4, 4, 3)

// A kernel runs the plans of one values-per-lane class VC = N / L (8, 16 or
// 24): its registers are sized for them, so a class of small transforms keeps
// more waves per SIMD.  wave_class() on the host gives a plan's class.


// LDS: twiddles [W] | NW exchange regions of M groups
template <int MODE, bool EXTRAS, int NW, int VC>
__global__ __launch_bounds__(64 * NW) void k_wave_rows(const FftWorldDev *__restrict__ worlds,
                                                       const int2 *__restrict__ tiles, int nTiles,
                                                       uint32_t stepOff, int kSel,
                                                       float *__restrict__ out) {
    extern __shared__ float2 smem[];
    if ((int)blockIdx.x >= nTiles) return;
    const int2 tile = step_tile(tiles, xcd_tile((int)blockIdx.x, nTiles));
    const FftWorldDev &wd = worlds[tile.x];
    const int W = wd.W;
    wave_load_twiddles(smem, wd.twW, W, 64 * NW);
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    switch (wd.planW.code) {
#define LENIA_WAVE_ROWS_CASE(code, n, L, ...)                                                         \
    case code:                                                                                         \
        if constexpr (n / L == VC)                                                                     \
            wave_rows<MODE, EXTRAS, n, L, __VA_ARGS__>(wd, tile, stepOff, kSel, out, smem,            \
                                                       smem + n + wave * (64 / L) * wave_group_span<n>(), \
                                                       wave, lane);                                    \
        break;
        LENIA_WAVE_PLANS(LENIA_WAVE_ROWS_CASE)
#undef LENIA_WAVE_ROWS_CASE
        default: break;  // the host launches wave kernels for these plans only
    }
}

template <int NW, int VC>
__global__ __launch_bounds__(64 * NW) void k_wave_cols(const FftWorldDev *__restrict__ worlds,
                                                       const int2 *__restrict__ tiles, int nTiles) {
    extern __shared__ float2 smem[];
    if ((int)blockIdx.x >= nTiles) return;
    const int2 tile = step_tile(tiles, xcd_tile((int)blockIdx.x, nTiles));
    const FftWorldDev &wd = worlds[tile.x];
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    const ColsFields cf = cols_fields(wd);
    switch (wd.planH.code) {
#define LENIA_WAVE_COLS_CASE(code, n, L, ...)                                                   \
    case code:                                                                                   \
        if constexpr (n / L == VC)                                                               \
            wave_cols<n, L, __VA_ARGS__>(cf, tile, smem,                                        \
                                         smem + n + wave * (64 / L) * wave_group_span<n>(), wave, lane, \
                                         64 * NW);                                               \
        break;
        LENIA_WAVE_PLANS(LENIA_WAVE_COLS_CASE)
#undef LENIA_WAVE_COLS_CASE
        default: break;
    }
}

// kernel-split step kernels: kMaxKernels waves per workgroup, a tile is M =
// 64 / L row pairs or columns.  LDS: twiddles [N] | 4 exchange regions |
// rows: growth terms of kernels 1..3 [3][V][64]; columns: spectra [V][64]
template <bool EXTRAS, int VC>
__global__ __launch_bounds__(64 * kMaxKernels) void k_wave_rows_split(
    const FftWorldDev *__restrict__ worlds, const int2 *__restrict__ tiles, int nTiles, uint32_t stepOff) {
    extern __shared__ float2 smem[];
    if ((int)blockIdx.x >= nTiles) return;
    const int2 tile = step_tile(tiles, xcd_tile((int)blockIdx.x, nTiles));
    const FftWorldDev &wd = worlds[tile.x];
    wave_load_twiddles(smem, wd.twW, wd.W, 64 * kMaxKernels);
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    switch (wd.planW.code) {
#define LENIA_WAVE_ROWS_SPLIT_CASE(code, n, L, ...)                                                \
    case code:                                                                                      \
        if constexpr (n / L == VC) {                                                                \
            float2 *xbw = smem + n + wave * (64 / L) * wave_group_span<n>();                        \
            float2 *terms = smem + n + kMaxKernels * (64 / L) * wave_group_span<n>();               \
            wave_rows_split<EXTRAS, n, L, __VA_ARGS__>(wd, tile, stepOff, smem, xbw, terms, wave, lane); \
        }                                                                                           \
        break;
        LENIA_WAVE_PLANS(LENIA_WAVE_ROWS_SPLIT_CASE)
#undef LENIA_WAVE_ROWS_SPLIT_CASE
        default: break;
    }
}

template <int VC>
__global__ __launch_bounds__(64 * kMaxKernels) void k_wave_cols_split(
    const FftWorldDev *__restrict__ worlds, const int2 *__restrict__ tiles, int nTiles) {
    extern __shared__ float2 smem[];
    if ((int)blockIdx.x >= nTiles) return;
    const int2 tile = step_tile(tiles, xcd_tile((int)blockIdx.x, nTiles));
    const FftWorldDev &wd = worlds[tile.x];
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    const ColsFields cf = cols_fields(wd);
    switch (wd.planH.code) {
#define LENIA_WAVE_COLS_SPLIT_CASE(code, n, L, ...)                                        \
    case code:                                                                              \
        if constexpr (n / L == VC) {                                                        \
            float2 *xbw = smem + n + wave * (64 / L) * wave_group_span<n>();                \
            float2 *xs = smem + n + kMaxKernels * (64 / L) * wave_group_span<n>();          \
            wave_cols_split<n, L, __VA_ARGS__>(cf, tile, smem, xbw, xs, wave, lane);        \
        }                                                                                   \
        break;
        LENIA_WAVE_PLANS(LENIA_WAVE_COLS_SPLIT_CASE)
#undef LENIA_WAVE_COLS_SPLIT_CASE
        default: break;
    }
}

}  // namespace lenia
