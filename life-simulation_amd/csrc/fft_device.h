// LDS-resident mixed-radix Stockham FFT for gfx950.
//
// A workgroup transforms `nt` independent complex sequences of length N held
// back to back in LDS.  Each stage of radix R (2, 3, 4, 5 or 8) reads the R
// inputs of each butterfly at stride N/R, twiddles them with e^{-2 pi i k r /
// (Ns R)} from a table generated in FP64, applies the size-R DFT and writes
// the outputs in autosort order, so the result lands in natural order with no
// bit reversal.  This is the Stockham formulation the reference's radix-2
// shader uses (Shaders/Common/FFT.metal:13-45, fftStageX/Y :65-121), widened to
// mixed radix and kept inside one workgroup: one global read and one global
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

// One Stockham stage over nt transforms of length N in LDS.  Every thread
// first loads all its butterflies, the group synchronises, then each thread
// computes and stores, so the stage runs in place in one LDS buffer.
// Precondition (checked on the host): ceil(nt*N/R / T) <= MAXV / R.
template <int R, bool INV, int MAXV, int T>
__device__ __forceinline__ void fft_stage(float2 *buf, int N, int Ns, int nt,
                                          const float2 *__restrict__ tw, int tid) {
    constexpr int MB = MAXV / R;
    const int NR = N / R;
    const int nb = nt * NR;
    const int twStep = N / (Ns * R);
    float2 v[MB][R];
#pragma unroll
    for (int i = 0; i < MB; ++i) {
        int g = tid + i * T;
        if (g < nb) {
            int b = g / NR, j = g - b * NR;
            const float2 *src = buf + b * N + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = src[r * NR];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MB; ++i) {
        int g = tid + i * T;
        if (g < nb) {
            int b = g / NR, j = g - b * NR;
            int k = j % Ns;
            if (k != 0) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = ctw<INV>(v[i][r], tw[k * r * twStep]);
            }
            Dft<R, INV>::run(v[i]);
            float2 *dst = buf + b * N + (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * Ns] = v[i][r];
        }
    }
    __syncthreads();
}

// Full transform; the caller synchronises after filling `buf`.  Unscaled in
// both directions.
template <bool INV, int MAXV, int T>
__device__ __forceinline__ void fft_lds(float2 *buf, const FftPlan &plan, int nt,
                                        const float2 *__restrict__ tw, int tid) {
    int Ns = 1;
    const int N = plan.n;
    for (int s = 0; s < plan.nStages; ++s) {
        const int R = plan.radix[s];
        switch (R) {
            case 2: fft_stage<2, INV, MAXV, T>(buf, N, Ns, nt, tw, tid); break;
            case 3: fft_stage<3, INV, MAXV, T>(buf, N, Ns, nt, tw, tid); break;
            case 4: fft_stage<4, INV, MAXV, T>(buf, N, Ns, nt, tw, tid); break;
            case 5: fft_stage<5, INV, MAXV, T>(buf, N, Ns, nt, tw, tid); break;
            default: fft_stage<8, INV, MAXV, T>(buf, N, Ns, nt, tw, tid); break;
        }
        Ns *= R;
    }
}

}  // namespace lenia
