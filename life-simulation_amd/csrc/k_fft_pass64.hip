// The FFT-form step kernels (fft_pass.h) instantiated for 64-thread workgroups.
#include "fft_pass.h"

namespace lenia {

void fft_rows_pass_64(int mode, bool extras, bool kpar, const FftWorldDev *worlds,
                       const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff,
                       int kSel, float *out, hipStream_t s) {
    rows_dispatch<64>(mode, extras, kpar, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
}

void fft_cols_pass_64(bool kpar, bool lean, const FftWorldDev *worlds, const int2 *tiles,
                       int nTiles, int grid, size_t lds, hipStream_t s) {
    cols_dispatch<64>(kpar, lean, worlds, tiles, nTiles, grid, lds, s);
}

}  // namespace lenia
