// Wave-resident FFT step passes (fft_wave.h): instantiations and launchers.
#include "fft_wave.h"
#include "launch.h"

namespace lenia {

int wave_lanes(uint32_t planCode) {
    switch (planCode) {
#define LENIA_WAVE_LANES_CASE(code, n, L, ...) \
    case code: return L;
        LENIA_WAVE_PLANS(LENIA_WAVE_LANES_CASE)
#undef LENIA_WAVE_LANES_CASE
        default: return 0;
    }
}

int wave_class(uint32_t planCode, int n) {
    const int L = wave_lanes(planCode);
    return L ? n / L : 0;
}

size_t wave_lds_bytes(int N, int L, int nw) {
    const int span = N + N / 32;
    return (size_t)(N + nw * (64 / L) * span) * sizeof(float2);
}

template <int NW, int VC>
static void wave_rows_launch(int mode, bool extras, const FftWorldDev *worlds, const int2 *tiles,
                             int nTiles, size_t lds, uint32_t stepOff, int kSel, float *out,
                             hipStream_t s) {
    const dim3 grid(nTiles), block(64 * NW);
    if (mode == 0)
        launch_timed(k_wave_rows<0, false, NW, VC>, grid, block, lds, s, worlds, tiles, nTiles,
                     stepOff, kSel, out);
    else if (mode == 1 && extras)
        launch_timed(k_wave_rows<1, true, NW, VC>, grid, block, lds, s, worlds, tiles, nTiles,
                     stepOff, kSel, out);
    else if (mode == 1)
        launch_timed(k_wave_rows<1, false, NW, VC>, grid, block, lds, s, worlds, tiles, nTiles,
                     stepOff, kSel, out);
    else
        launch_timed(k_wave_rows<2, false, NW, VC>, grid, block, lds, s, worlds, tiles, nTiles,
                     stepOff, kSel, out);
}

template <int NW>
static hipError_t wave_rows_nw(int vc, int mode, bool extras, const FftWorldDev *worlds,
                               const int2 *tiles, int nTiles, size_t lds, uint32_t stepOff,
                               int kSel, float *out, hipStream_t s) {
    switch (vc) {
        case 8: wave_rows_launch<NW, 8>(mode, extras, worlds, tiles, nTiles, lds, stepOff, kSel, out, s); break;
        case 16: wave_rows_launch<NW, 16>(mode, extras, worlds, tiles, nTiles, lds, stepOff, kSel, out, s); break;
        case 24: wave_rows_launch<NW, 24>(mode, extras, worlds, tiles, nTiles, lds, stepOff, kSel, out, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

hipError_t launch_wave_rows(int mode, int nw, int vc, bool extras, const FftWorldDev *worlds,
                            const int2 *tiles, int nTiles, size_t lds, uint32_t stepOff, int kSel,
                            float *out, hipStream_t s) {
    if (nTiles < 1 || mode < 0 || mode > 2) return hipErrorInvalidValue;
    hipError_t e;
    switch (nw) {
        case 1: e = wave_rows_nw<1>(vc, mode, extras, worlds, tiles, nTiles, lds, stepOff, kSel, out, s); break;
        case 2: e = wave_rows_nw<2>(vc, mode, extras, worlds, tiles, nTiles, lds, stepOff, kSel, out, s); break;
        case 4: e = wave_rows_nw<4>(vc, mode, extras, worlds, tiles, nTiles, lds, stepOff, kSel, out, s); break;
        default: return hipErrorInvalidValue;
    }
    return e != hipSuccess ? e : hipGetLastError();
}

template <int NW>
static hipError_t wave_cols_nw(int vc, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                               size_t lds, hipStream_t s) {
    const dim3 grid(nTiles), block(64 * NW);
    switch (vc) {
        case 8: launch_timed(k_wave_cols<NW, 8>, grid, block, lds, s, worlds, tiles, nTiles); break;
        case 16: launch_timed(k_wave_cols<NW, 16>, grid, block, lds, s, worlds, tiles, nTiles); break;
        case 24: launch_timed(k_wave_cols<NW, 24>, grid, block, lds, s, worlds, tiles, nTiles); break;
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

size_t wave_split_lds_bytes(int N, int L, bool rows) {
    const int V = N / L, span = N + N / 32;
    const int extra = rows ? (kMaxKernels - 1) * V * 64 : V * 64;
    return (size_t)(N + kMaxKernels * (64 / L) * span + extra) * sizeof(float2);
}

hipError_t launch_wave_rows_split(bool extras, int vc, const FftWorldDev *worlds, const int2 *tiles,
                                  int nTiles, size_t lds, uint32_t stepOff, hipStream_t s) {
    if (nTiles < 1) return hipErrorInvalidValue;
    const dim3 grid(nTiles), block(64 * kMaxKernels);
    switch (vc * 2 + (extras ? 1 : 0)) {
        case 16: launch_timed(k_wave_rows_split<false, 8>, grid, block, lds, s, worlds, tiles, nTiles, stepOff); break;
        case 17: launch_timed(k_wave_rows_split<true, 8>, grid, block, lds, s, worlds, tiles, nTiles, stepOff); break;
        case 32: launch_timed(k_wave_rows_split<false, 16>, grid, block, lds, s, worlds, tiles, nTiles, stepOff); break;
        case 33: launch_timed(k_wave_rows_split<true, 16>, grid, block, lds, s, worlds, tiles, nTiles, stepOff); break;
        case 48: launch_timed(k_wave_rows_split<false, 24>, grid, block, lds, s, worlds, tiles, nTiles, stepOff); break;
        case 49: launch_timed(k_wave_rows_split<true, 24>, grid, block, lds, s, worlds, tiles, nTiles, stepOff); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_wave_cols_split(int vc, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                                  size_t lds, hipStream_t s) {
    if (nTiles < 1) return hipErrorInvalidValue;
    const dim3 grid(nTiles), block(64 * kMaxKernels);
    switch (vc) {
        case 8: launch_timed(k_wave_cols_split<8>, grid, block, lds, s, worlds, tiles, nTiles); break;
        case 16: launch_timed(k_wave_cols_split<16>, grid, block, lds, s, worlds, tiles, nTiles); break;
        case 24: launch_timed(k_wave_cols_split<24>, grid, block, lds, s, worlds, tiles, nTiles); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_wave_cols(int nw, int vc, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                            size_t lds, hipStream_t s) {
    if (nTiles < 1) return hipErrorInvalidValue;
    hipError_t e;
    switch (nw) {
        case 1: e = wave_cols_nw<1>(vc, worlds, tiles, nTiles, lds, s); break;
        case 2: e = wave_cols_nw<2>(vc, worlds, tiles, nTiles, lds, s); break;
        case 4: e = wave_cols_nw<4>(vc, worlds, tiles, nTiles, lds, s); break;
        default: return hipErrorInvalidValue;
    }
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace lenia
