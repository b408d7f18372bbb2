// LDS-resident mixed-radix Stockham FFT for gfx950.
//
// A workgroup of T threads transforms `nt` independent complex sequences of
// length N held back to back in LDS.  Each stage of radix R (2, 3, 4, 5, 8, 16
// or 32) reads the R inputs of each butterfly at stride N/R, twiddles them
// with e^{-2 pi i k r / (Ns R)} from an FP64-generated table held in LDS,
// applies the size-R DFT in registers and writes the outputs in autosort
// order, so the result lands in natural order with no bit reversal.  This is
// the Stockham formulation of the reference's radix-2 shader
// (Shaders/Common/FFT.metal:13-45, fftStageX/Y :65-121), widened to mixed and
// high radix and kept inside one workgroup: one global read and one global
// write per transform instead of one full-field round trip per stage.
//
// All arithmetic is written with explicit fmaf and the file is compiled with
// -ffp-contract=off, so a butterfly gives the same bits whichever thread runs
// it, whatever the tile shape: batched and single-world launches agree bit
// for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "lenia_types.h"

namespace lenia {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a * w, or a * conj(w) for the inverse transform
template <bool INV>
__device__ __forceinline__ float2 ctw(float2 a, float2 w) {
    if (INV) return make_float2(fmaf(a.x, w.x, a.y * w.y), fmaf(a.y, w.x, -(a.x * w.y)));
    return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.y, w.x, a.x * w.y));
}

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// W32^e = e^{-2 pi i e / 32}, e < 32 (symmetric literals: exact 0 and +-1)
__device__ __forceinline__ float2 w32(int e) {
    constexpr float c[9] = {1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f,
                            8.314696123e-01f, 7.071067812e-01f, 5.555702330e-01f,
                            3.826834324e-01f, 1.950903220e-01f, 0.0f};
    // cos(e pi/16) and sin(e pi/16) by quadrant; e is a compile-time constant
    // at every call site, so this folds to literals
    const int q = e >> 3, r = e & 7;
    float co, si;
    if (q == 0) { co = c[r]; si = c[8 - r]; }
    else if (q == 1) { co = -c[8 - r]; si = c[r]; }
    else if (q == 2) { co = -c[r]; si = -c[8 - r]; }
    else { co = c[8 - r]; si = -c[r]; }
    return make_float2(co, -si);
}

template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<4, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        float2 s13 = cadd(v[1], v[3]), d13 = rot<INV>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};

template <bool INV>
struct Dft<8, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        const float c = 0.70710678118654752440f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        // o[k] *= W8^k  (W8 = e^{-i pi/4} forward, conjugate inverse)
        float2 o1 = o[1], o3 = o[3];
        if (INV) {
            o[1] = make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y));
            o[3] = make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y));
        } else {
            o[1] = make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));
            o[3] = make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
        }
        o[2] = rot<INV>(o[2]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

// 16 = 4 x 4: x[4 n1 + n2], X[k1 + 4 k2] =
//   sum_n2 W4^{n2 k2} W16^{n2 k1} sum_n1 x[4 n1 + n2] W4^{n1 k1}
template <bool INV>
struct Dft<16, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            float2 t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
            Dft<4, INV>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
        }
#pragma unroll
        for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
            for (int k1 = 1; k1 < 4; ++k1) y[n2][k1] = ctw<INV>(y[n2][k1], w32(2 * n2 * k1));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            Dft<4, INV>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
        }
    }
};

// 32 = 8 x 4: x[4 n1 + n2] (n1 < 8, n2 < 4), X[k1 + 8 k2] =
//   sum_n2 W4^{n2 k2} W32^{n2 k1} sum_n1 x[4 n1 + n2] W8^{n1 k1}
template <bool INV>
struct Dft<32, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 y[4][8];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            float2 t[8];
#pragma unroll
            for (int n1 = 0; n1 < 8; ++n1) t[n1] = v[4 * n1 + n2];
            Dft<8, INV>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < 8; ++k1) y[n2][k1] = t[k1];
        }
#pragma unroll
        for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
            for (int k1 = 1; k1 < 8; ++k1) y[n2][k1] = ctw<INV>(y[n2][k1], w32(n2 * k1));
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) {
            float2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            Dft<4, INV>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 8 * k2] = t[k2];
        }
    }
};

template <bool INV>
struct Dft<3, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        const float s3 = 0.86602540378443864676f;
        float2 t1 = cadd(v[1], v[2]);
        float2 t2 = make_float2(fmaf(-0.5f, t1.x, v[0].x), fmaf(-0.5f, t1.y, v[0].y));
        float2 t3 = rot<INV>(cscale(csub(v[1], v[2]), s3));
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    }
};

template <bool INV>
struct Dft<5, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
        const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
        const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
        const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
        float2 a0 = v[0];
        float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        float2 b1 = make_float2(fmaf(c2, t2.x, fmaf(c1, t1.x, a0.x)),
                                fmaf(c2, t2.y, fmaf(c1, t1.y, a0.y)));
        float2 b2 = make_float2(fmaf(c1, t2.x, fmaf(c2, t1.x, a0.x)),
                                fmaf(c1, t2.y, fmaf(c2, t1.y, a0.y)));
        float2 d1 = make_float2(fmaf(s2, t4.x, s1 * t3.x), fmaf(s2, t4.y, s1 * t3.y));
        float2 d2 = make_float2(fmaf(-s1, t4.x, s2 * t3.x), fmaf(-s1, t4.y, s2 * t3.y));
        d1 = rot<INV>(d1);
        d2 = rot<INV>(d2);
        v[0] = cadd(a0, cadd(t1, t2));
        v[1] = cadd(b1, d1);
        v[4] = csub(b1, d1);
        v[2] = cadd(b2, d2);
        v[3] = csub(b2, d2);
    }
};

// floor(g / d) for g < 2^16, d < 2^16 with the host-computed magic
// m = ceil(2^32 / d) (m = 0 encodes d = 1)
__device__ __forceinline__ uint32_t fast_div(uint32_t g, uint32_t m) {
    return m ? __umulhi(g, m) : g;
}

// LDS placement of transform buffers: entry i sits at i + i/32, one unused
// entry after every 32 (256 bytes, one row of the LDS banks).  Stockham stores
// land at stride R (up to 32 x 8 bytes); unpadded, 32 of the 64 lanes hit one
// bank, padded every store and load of a stage is at most 2-way (the minimum
// for 8-byte accesses).  Unlike an XOR swizzle the placement is affine on
// each run that does not cross a 32-entry row, so the R accesses of a
// butterfly are one base address plus compile-time offsets (see stage_off).
// A region of n entries spans lds_span(n) entries.
__device__ __forceinline__ int sw(int i) { return i + (i >> 5); }

// sw(base + r s) - sw(base) for the R accesses of a Stockham butterfly, when
// that difference does not depend on base: s a multiple of 32, or s a divisor
// of 32 with every butterfly's run inside one 32-entry row (N a multiple of
// 32, and for stores the run of R s entries aligned to min(R s, 32)).
template <int N, int R, int S, bool STORE>
constexpr bool stage_fold() {
    if (S % 32 == 0) return true;
    if (N % 32 != 0 || 32 % S != 0) return false;
    if (!STORE) return true;  // loads: base = b N + j with j < S
    return (S * R) % 32 == 0 || 32 % (S * R) == 0;
}
template <int S>
__device__ __forceinline__ constexpr int stage_off(int r) { return r * S + (r * S) / 32; }

// One Stockham stage over nt transforms in LDS (padded placement), from buf
// to dst.  Every thread loads its butterflies, twiddles, transforms and
// stores.  In place (dst == buf) the group synchronises between the loads and
// the stores; between two regions (dst != buf) it does not, since no thread
// writes what another still reads: one barrier per stage instead of two.
// `tws` is this stage's twiddle table, [k][r-1] for k < Ns, 0 < r < R.
// Precondition (host): ceil(nt*N/R / T) * R <= MAXV.
template <int R, bool INV, int MAXV, int T>
__device__ __forceinline__ void fft_stage(float2 *buf, float2 *dst, const FftStage &st, int N, int nt,
                                          const float2 *tws, int tid) {
    constexpr int MB = MAXV / R;
    if constexpr (MB > 0) {
        const int NR = st.NR, Ns = st.Ns;
        const int nb = nt * NR;
        float2 v[MB][R];
        float2 w[MB][R - 1];
        // data and twiddles are read in one batch (one LDS wait per stage);
        // the twiddle tables are never written during a transform
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            int g = tid + i * T;
            if (g < nb) {
                int b = (int)fast_div((uint32_t)g, st.mNR);
                int j = g - b * NR;
                int k = j - Ns * (int)fast_div((uint32_t)j, st.mNs);
                const int base = b * N + j;
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = buf[sw(base + r * NR)];
                const float2 *t = tws + k * (R - 1);
#pragma unroll
                for (int r = 0; r < R - 1; ++r) w[i][r] = t[r];
            }
        }
        if (dst == buf) __syncthreads();
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            int g = tid + i * T;
            if (g < nb) {
                int b = (int)fast_div((uint32_t)g, st.mNR);
                int j = g - b * NR;
                int k = j - Ns * (int)fast_div((uint32_t)j, st.mNs);
                if (k != 0) {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[i][r] = ctw<INV>(v[i][r], w[i][r - 1]);
                }
                Dft<R, INV>::run(v[i]);
                const int base = b * N + (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[sw(base + r * Ns)] = v[i][r];
            }
        }
        __syncthreads();
    }
}

// Full transform; the caller synchronises after filling `buf` and the LDS
// stage-twiddle tables `tw` (plan.st[s].twOff indexes them).  Unscaled both
// ways.  The result is left in `buf`.  `alt` (NULL: none) is a second region
// of the same span: the stages then alternate between the two, ending in
// `buf` (with an odd stage count the first stage runs in place).  Placement
// never changes arithmetic, so both forms give the same bits.
template <bool INV, int MAXV, int T>
__device__ __forceinline__ void fft_lds(float2 *buf, const FftPlan &plan, int nt, const float2 *tw,
                                        int tid, float2 *alt = nullptr) {
    const int N = plan.n;
    if (!alt) alt = buf;
    float2 *src = buf;
    for (int s = 0; s < plan.nStages; ++s) {
        const FftStage &st = plan.st[s];
        const float2 *tws = tw + st.twOff;
        float2 *dst = ((plan.nStages - 1 - s) & 1) ? alt : buf;
        switch (st.radix) {
            case 2: fft_stage<2, INV, MAXV, T>(src, dst, st, N, nt, tws, tid); break;
            case 3: fft_stage<3, INV, MAXV, T>(src, dst, st, N, nt, tws, tid); break;
            case 4: fft_stage<4, INV, MAXV, T>(src, dst, st, N, nt, tws, tid); break;
            case 5: fft_stage<5, INV, MAXV, T>(src, dst, st, N, nt, tws, tid); break;
            case 8: fft_stage<8, INV, MAXV, T>(src, dst, st, N, nt, tws, tid); break;
            case 16: fft_stage<16, INV, MAXV, T>(src, dst, st, N, nt, tws, tid); break;
            default: fft_stage<32, INV, MAXV, T>(src, dst, st, N, nt, tws, tid); break;
        }
        src = dst;
    }
}

// The same Stockham stage with the plan constants known at compile time:
// index arithmetic folds to shifts and masks (multiply-high for the factors 3
// and 5), the stage tables need no loads and there is no radix switch.  The
// butterflies, twiddles and stores are those of fft_stage, so the bits are too.
template <int R, bool INV, int MAXV, int T, int N, int NS, int TWOFF>
__device__ __forceinline__ void fft_stage_c(float2 *buf, float2 *dstBuf, int nt, const float2 *tw,
                                            int tid) {
    constexpr int NR = N / R;
    constexpr int MB = MAXV / R;
    static_assert(N % (NS * R) == 0, "radix sequence does not divide the length");
    if constexpr (MB > 0) {
        constexpr bool kFoldLd = stage_fold<N, R, NR, false>();
        constexpr bool kFoldSt = stage_fold<N, R, NS, true>();
        const int nb = nt * NR;
        float2 v[MB][R];
        float2 w[MB][R - 1];
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            const int g = tid + i * T;
            if (g < nb) {
                const int b = g / NR, j = g % NR, k = j % NS;
                const int base = b * N + j;
                if constexpr (kFoldLd) {
                    const float2 *src = buf + sw(base);
#pragma unroll
                    for (int r = 0; r < R; ++r) v[i][r] = src[stage_off<NR>(r)];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[i][r] = buf[sw(base + r * NR)];
                }
                const float2 *t = tw + TWOFF + k * (R - 1);
#pragma unroll
                for (int r = 0; r < R - 1; ++r) w[i][r] = t[r];
            }
        }
        if (dstBuf == buf) __syncthreads();
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            const int g = tid + i * T;
            if (g < nb) {
                const int b = g / NR, j = g % NR, k = j % NS;
                if (k != 0) {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[i][r] = ctw<INV>(v[i][r], w[i][r - 1]);
                }
                Dft<R, INV>::run(v[i]);
                const int base = b * N + (j - k) * R + k;
                if constexpr (kFoldSt) {
                    float2 *dst = dstBuf + sw(base);
#pragma unroll
                    for (int r = 0; r < R; ++r) dst[stage_off<NS>(r)] = v[i][r];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) dstBuf[sw(base + r * NS)] = v[i][r];
                }
            }
        }
        __syncthreads();
    }
}

// stages from src, returning where the result is.  Each stage's destination
// alternates between buf and alt (alt == buf: all in place).  anyEnd false:
// the last stage ends in buf (with an odd stage count the first runs in
// place); true: every stage moves to the other region and the result ends in
// whichever the stage count gives.
template <bool INV, int MAXV, int T, int N, int NS, int TWOFF, int R, int... REST>
__device__ __forceinline__ float2 *fft_fixed_from(float2 *buf, float2 *alt, float2 *src, bool anyEnd,
                                                  int nt, const float2 *tw, int tid) {
    float2 *dst = anyEnd ? (src == buf ? alt : buf) : ((sizeof...(REST) & 1) ? alt : buf);
    fft_stage_c<R, INV, MAXV, T, N, NS, TWOFF>(src, dst, nt, tw, tid);
    if constexpr (sizeof...(REST) > 0)
        return fft_fixed_from<INV, MAXV, T, N, NS * R, TWOFF + NS * (R - 1), REST...>(buf, alt, dst, anyEnd,
                                                                                     nt, tw, tid);
    else
        return dst;
}
template <bool INV, int MAXV, int T, int N, int NS, int TWOFF, int R, int... REST>
__device__ __forceinline__ float2 *fft_fixed(float2 *buf, float2 *alt, bool anyEnd, int nt, const float2 *tw,
                                             int tid) {
    return fft_fixed_from<INV, MAXV, T, N, NS, TWOFF, R, REST...>(buf, alt ? alt : buf, buf, anyEnd, nt, tw, tid);
}

// fft_lds through the compile-time kernel of the plan when it has one
// the other radix orders of the benchmark lengths (radix-plan tuning), kept
// out of the default plans' code path
template <bool INV, int MAXV, int T>
__device__ __forceinline__ float2 *fft_run_alt(float2 *buf, uint32_t code, int nt,
                                                         const float2 *tw, int tid, float2 *alt, bool anyEnd) {
    float2 *out = buf;
    switch (code) {
        case kPlan64b: out = fft_fixed<INV, MAXV, T, 64, 1, 0, 16, 4>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan64c: out = fft_fixed<INV, MAXV, T, 64, 1, 0, 4, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan128b: out = fft_fixed<INV, MAXV, T, 128, 1, 0, 8, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan128c: out = fft_fixed<INV, MAXV, T, 128, 1, 0, 4, 4, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan256b: out = fft_fixed<INV, MAXV, T, 256, 1, 0, 8, 8, 4>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan256c: out = fft_fixed<INV, MAXV, T, 256, 1, 0, 4, 8, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan512b: out = fft_fixed<INV, MAXV, T, 512, 1, 0, 16, 8, 4>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan512c: out = fft_fixed<INV, MAXV, T, 512, 1, 0, 4, 8, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan1024b: out = fft_fixed<INV, MAXV, T, 1024, 1, 0, 8, 8, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan1024c: out = fft_fixed<INV, MAXV, T, 1024, 1, 0, 16, 16, 4>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan1024d: out = fft_fixed<INV, MAXV, T, 1024, 1, 0, 4, 16, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan2048b: out = fft_fixed<INV, MAXV, T, 2048, 1, 0, 8, 16, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan4096b: out = fft_fixed<INV, MAXV, T, 4096, 1, 0, 8, 8, 8, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan96b: out = fft_fixed<INV, MAXV, T, 96, 1, 0, 4, 8, 3>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan192b: out = fft_fixed<INV, MAXV, T, 192, 1, 0, 16, 4, 3>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan384c: out = fft_fixed<INV, MAXV, T, 384, 1, 0, 8, 4, 4, 3>(buf, alt, anyEnd, nt, tw, tid); break;
        default: out = fft_fixed<INV, MAXV, T, 384, 1, 0, 8, 16, 3>(buf, alt, anyEnd, nt, tw, tid); break;
    }
    return out;
}

// fft_lds through the compile-time kernel of the plan when it has one
// alt: an optional second LDS region of the same span (see fft_lds).  The
// result is left in buf, or with anyEnd (and alt) in whichever region the
// compile-time plans end in; the return value says where.
template <bool INV, int MAXV, int T>
__device__ __forceinline__ float2 *fft_run(float2 *buf, const FftPlan &plan, int nt, const float2 *tw,
                                           int tid, float2 *alt = nullptr, bool anyEnd = false) {
    if (__builtin_expect(plan.code >= kPlan64b, 0)) return fft_run_alt<INV, MAXV, T>(buf, plan.code, nt, tw, tid, alt, anyEnd);
    float2 *out = buf;
    switch (plan.code) {
        case kPlan64: out = fft_fixed<INV, MAXV, T, 64, 1, 0, 8, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan128: out = fft_fixed<INV, MAXV, T, 128, 1, 0, 16, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan256: out = fft_fixed<INV, MAXV, T, 256, 1, 0, 16, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan512: out = fft_fixed<INV, MAXV, T, 512, 1, 0, 8, 8, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan1024: out = fft_fixed<INV, MAXV, T, 1024, 1, 0, 16, 8, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan2048: out = fft_fixed<INV, MAXV, T, 2048, 1, 0, 16, 16, 8>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan96: out = fft_fixed<INV, MAXV, T, 96, 1, 0, 8, 4, 3>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan192: out = fft_fixed<INV, MAXV, T, 192, 1, 0, 8, 8, 3>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan384: out = fft_fixed<INV, MAXV, T, 384, 1, 0, 16, 8, 3>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan720: out = fft_fixed<INV, MAXV, T, 720, 1, 0, 16, 3, 3, 5>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan1080: out = fft_fixed<INV, MAXV, T, 1080, 1, 0, 8, 3, 3, 3, 5>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan1280: out = fft_fixed<INV, MAXV, T, 1280, 1, 0, 16, 16, 5>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan1920: out = fft_fixed<INV, MAXV, T, 1920, 1, 0, 16, 8, 3, 5>(buf, alt, anyEnd, nt, tw, tid); break;
        case kPlan4096: out = fft_fixed<INV, MAXV, T, 4096, 1, 0, 16, 16, 16>(buf, alt, anyEnd, nt, tw, tid); break;
        default: fft_lds<INV, MAXV, T>(buf, plan, nt, tw, tid, alt); break;
    }
    return out;
}

// XCD-aware tile order: consecutive workgroup ids land on different XCDs
// (round robin over 8), so give each XCD a contiguous run of tiles; tiles
// that share cache lines (neighbouring row pairs) then share one L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int x = b & 7, j = b >> 3;
    const int base = n >> 3, rem = n & 7;
    return x * base + min(x, rem) + j;
}

}  // namespace lenia
