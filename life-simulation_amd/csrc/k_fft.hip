// FFT-convolution form of the Lenia step on gfx950: the set-up kernels
// (twiddle tables, kernel spectra) and the entry points of the two step
// passes, which pick the step kernels (fft_pass.h) instantiated for the
// requested workgroup size in k_fft_pass<T>.hip.
#include <hip/hip_runtime.h>

#include "fft_pass.h"

namespace lenia {

// ---------------------------------------------------------------- set-up kernels
// e^{-2 pi i t / N} for t < N, evaluated in FP64 and rounded once to FP32
__global__ void k_twiddles(float2 *tw, int N) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    double a = 2.0 * (double)t / (double)N;
    tw[t] = make_float2((float)cospi(a), (float)(-sinpi(a)));
}

// per-stage twiddle tables of a plan: stage s holds e^{-2 pi i k r / (Ns R)}
// at [twOff + k (R-1) + r - 1], k < Ns, 0 < r < R, copied from the length-N
// FP64-evaluated table
__global__ void k_stage_twiddles(FftPlan plan, const float2 *__restrict__ raw, float2 *out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    for (int s = 0; s < plan.nStages; ++s) {
        const FftStage &st = plan.st[s];
        const int size = st.Ns * (st.radix - 1);
        if (t >= st.twOff && t < st.twOff + size) {
            int u = t - st.twOff;
            int k = u / (st.radix - 1), r = u - k * (st.radix - 1) + 1;
            out[t] = raw[k * r * st.twStep];
        }
    }
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
hipError_t launch_fft_rows(int mode, int threads, bool extras, bool kpar,
                           const FftWorldDev *worlds, const int2 *tiles, int nTiles, int grid,
                           size_t lds, uint32_t stepOff, int kSel, float *out,
                           hipStream_t s, int tileCells) {
    // one tile per workgroup
    if (grid < 1 || grid < nTiles) return hipErrorInvalidValue;
    switch (threads) {
        case 64: fft_rows_pass_64(mode, extras, kpar, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 128: fft_rows_pass_128(mode, extras, kpar, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 256:
            if (mode == 1 && kpar && tileCells > 0 && tileCells <= kMidValues * 256)
                fft_rows_step_256_mid(extras, worlds, tiles, nTiles, grid, lds, stepOff, s);
            else
                fft_rows_pass_256(mode, extras, kpar, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
            break;
        case 512:
            if (mode == 1 && kpar && tileCells > 0 && tileCells <= kMidValues * 512)
                fft_rows_step_512_mid(extras, worlds, tiles, nTiles, grid, lds, stepOff, s);
            else
                fft_rows_pass_512(mode, extras, kpar, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
            break;
        case 1024:  // steps with side-by-side transforms only (the serial form needs more registers)
            if (mode == 1 && !kpar) return hipErrorInvalidValue;
            fft_rows_pass_1024(mode, extras, tileCells > 0 && tileCells <= kLeanValues * 1024, worlds,
                               tiles, nTiles, grid, lds, stepOff, kSel, out, s);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fft_cols(int threads, bool kpar, const FftWorldDev *worlds, const int2 *tiles,
                           int nTiles, int grid, size_t lds, hipStream_t s, int tileCells) {
    if (grid < 1 || grid < nTiles) return hipErrorInvalidValue;
    const bool lean = tileCells > 0 && tileCells <= kLeanCols * threads;
    switch (threads) {
        case 64: fft_cols_pass_64(kpar, lean, worlds, tiles, nTiles, grid, lds, s); break;
        case 128: fft_cols_pass_128(kpar, lean, worlds, tiles, nTiles, grid, lds, s); break;
        case 256: fft_cols_pass_256(kpar, lean, worlds, tiles, nTiles, grid, lds, s); break;
        case 512: fft_cols_pass_512(kpar, lean, worlds, tiles, nTiles, grid, lds, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_twiddles(float *tw, int N, hipStream_t s) {
    hipLaunchKernelGGL(k_twiddles, dim3((N + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<float2 *>(tw), N);
    return hipGetLastError();
}

hipError_t launch_stage_twiddles(const FftPlan &plan, const float *raw, float *out,
                                  hipStream_t s) {
    const int n = plan.n;  // the stage tables hold n - 1 entries
    hipLaunchKernelGGL(k_stage_twiddles, dim3((n + 255) / 256), dim3(256), 0, s, plan,
                       reinterpret_cast<const float2 *>(raw), reinterpret_cast<float2 *>(out));
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

int fft_max_values() { return kMaxV; }
int fft_row_values(int T) { return row_values(T); }

}  // namespace lenia
