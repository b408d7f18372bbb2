// The FFT-form step kernels (fft_pass.h) instantiated for 256-thread workgroups.
#include "fft_pass.h"

namespace lenia {

void fft_rows_pass_256(int mode, bool extras, bool kpar, const FftWorldDev *worlds,
                       const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff,
                       int kSel, float *out, hipStream_t s) {
    rows_dispatch<256>(mode, extras, kpar, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
}

void fft_rows_step_256_mid(bool extras, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                           int grid, size_t lds, uint32_t stepOff, hipStream_t s) {
    if (extras)
        rows_launch<256, true, true, kMidValues>(1, worlds, tiles, nTiles, grid, lds, stepOff, 0, nullptr, s);
    else
        rows_launch<256, false, true, kMidValues>(1, worlds, tiles, nTiles, grid, lds, stepOff, 0, nullptr, s);
}

void fft_cols_pass_256(bool kpar, bool lean, const FftWorldDev *worlds, const int2 *tiles,
                       int nTiles, int grid, size_t lds, hipStream_t s) {
    cols_dispatch<256>(kpar, lean, worlds, tiles, nTiles, grid, lds, s);
}

}  // namespace lenia
