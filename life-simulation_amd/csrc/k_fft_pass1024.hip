// The FFT-form row-pass kernels (fft_pass.h) instantiated for 1024-thread
// workgroups: side-by-side kernel transforms only (the serial form needs more
// than the 128 registers a thread has at this size).
#include "fft_pass.h"

namespace lenia {

// lean: every tile holds at most kLeanValues cells per thread
void fft_rows_pass_1024(int mode, bool extras, bool lean, const FftWorldDev *worlds,
                        const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff,
                        int kSel, float *out, hipStream_t s) {
    if (lean && mode == 1) {
        if (extras) rows_launch<1024, true, true, kLeanValues>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
        else rows_launch<1024, false, true, kLeanValues>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
        return;
    }
    if (extras) rows_launch<1024, true, true>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
    else rows_launch<1024, false, true>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
}

}  // namespace lenia
