// The FFT-form step kernels (fft_pass.h) instantiated for 512-thread workgroups.
#include "fft_pass.h"

namespace lenia {

void fft_rows_pass_512(int mode, bool extras, bool kpar, const FftWorldDev *worlds,
                       const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff,
                       int kSel, float *out, hipStream_t s) {
    rows_dispatch<512>(mode, extras, kpar, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
}

void fft_rows_step_512_mid(bool extras, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                           int grid, size_t lds, uint32_t stepOff, hipStream_t s) {
    if (extras)
        rows_launch<512, true, true, kMidValues>(1, worlds, tiles, nTiles, grid, lds, stepOff, 0, nullptr, s);
    else
        rows_launch<512, false, true, kMidValues>(1, worlds, tiles, nTiles, grid, lds, stepOff, 0, nullptr, s);
}

void fft_cols_pass_512(bool kpar, bool lean, const FftWorldDev *worlds, const int2 *tiles,
                       int nTiles, int grid, size_t lds, hipStream_t s) {
    cols_dispatch<512>(kpar, lean, worlds, tiles, nTiles, grid, lds, s);
}

}  // namespace lenia
