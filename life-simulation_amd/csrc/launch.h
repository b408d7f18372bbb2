// Host-side launchers of the HIP kernels (defined in the .hip files).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lenia_types.h"

namespace lenia {

// Kernel timestamps.  When `start` is set, the next step-kernel launch on this
// thread records the two events around the kernel itself
// (hipExtLaunchKernelGGL) and clears them; otherwise launches are plain.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
LaunchEvents &launch_events();

template <class K, class... A>
inline void launch_timed(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, A... args) {
    LaunchEvents &ev = launch_events();
    if (ev.start) {
        hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)lds, s, ev.start, ev.stop, 0u, args...);
        ev = LaunchEvents();
    } else {
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
    }
}

// FFT path (k_fft.hip).  mode 0: forward row FFT of the state; 1: full step
// row pass; 2: potential of kernel kSel into `out`.  The step kernels run the
// step `stepOff` steps after each world's stored clock (mode 1 / mode 0 of
// the direct forms).  tileCells: the largest (K) x nt x W over the launch's
// tiles when the caller knows it (lets 1024-thread steps run the variant
// sized for small tiles), else 0.
hipError_t launch_fft_rows(int mode, int threads, bool extras, bool kpar,
                           const FftWorldDev *worlds, const int2 *tiles, int nTiles, int grid,
                           size_t lds, uint32_t stepOff, int kSel, float *out, hipStream_t s,
                           int tileCells = 0);
// tileCells: the largest (K) x nt x H over the launch's tiles, or 0
hipError_t launch_fft_cols(int threads, bool kpar, const FftWorldDev *worlds, const int2 *tiles,
                           int nTiles, int grid, size_t lds, hipStream_t s, int tileCells = 0);
// Wave-resident executor of the same passes (k_fft_wave.hip): nw waves per
// workgroup, each running 64 / L transforms of L lanes; a tile is nw (64 / L)
// row pairs or columns.  Only for plans with wave_lanes(code) > 0; a launch
// covers the plans of one values-per-lane class vc = N / L (wave_class).
int wave_lanes(uint32_t planCode);
int wave_class(uint32_t planCode, int n);
size_t wave_lds_bytes(int N, int L, int nw);
hipError_t launch_wave_rows(int mode, int nw, int vc, bool extras, const FftWorldDev *worlds,
                            const int2 *tiles, int nTiles, size_t lds, uint32_t stepOff, int kSel,
                            float *out, hipStream_t s);
hipError_t launch_wave_cols(int nw, int vc, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                            size_t lds, hipStream_t s);
// kernel-split step passes: kMaxKernels waves per workgroup, one per kernel
// transform (the row pass's forward transform and the column pass's forward
// transform on wave 0); a tile is 64 / L row pairs or columns
size_t wave_split_lds_bytes(int N, int L, bool rows);
hipError_t launch_wave_rows_split(bool extras, int vc, const FftWorldDev *worlds, const int2 *tiles,
                                  int nTiles, size_t lds, uint32_t stepOff, hipStream_t s);
hipError_t launch_wave_cols_split(int vc, const FftWorldDev *worlds, const int2 *tiles, int nTiles,
                                  size_t lds, hipStream_t s);
hipError_t launch_twiddles(float *tw, int N, hipStream_t s);
hipError_t launch_stage_twiddles(const FftPlan &plan, const float *raw, float *out, hipStream_t s);
hipError_t launch_kernel_spectrum(const double *win, int h, int W, int H, double *Cw, float *S,
                                  hipStream_t s);
int fft_max_values();
// cells per thread the row pass holds in registers at T threads
int fft_row_values(int T);

// direct path (k_direct.hip).  mode 0: step; 1: potential of kernel kSel.
// tileH: the output tile height the worlds' records were tiled with (32, or
// 16 for direct_tile_shape's small grids).  Grids of at most numCUs tiles run
// the split form (each output's two window-row halves on two threads; the
// same bits)
hipError_t launch_direct(int mode, const DirectWorldDev *worlds, const int2 *tiles, int nTiles,
                         int grid, size_t lds, uint32_t stepOff, int kSel, float *out,
                         hipStream_t s, int numCUs, int tileH);
// tile shape of a single W x H world on numCUs CUs (batches use 64 x 32)
void direct_tile_shape(int W, int H, int numCUs, int *tx, int *ty);
constexpr int kDirectTileW = 64, kDirectTileH = 32;
int direct_npad(int h);
// the split form runs for grids of at most one tile per CU; its LDS holds the
// second half's sums as well
bool direct_split(int nTiles, int numCUs);
size_t direct_lds_bytes(int h, int tileH, bool split);

// MFMA direct forms (k_mfma.hip).  mode 0: step; 1: potential of kernel kSel.
hipError_t launch_mfma(int mode, bool split, int ksteps, const MfmaWorldDev *worlds,
                       const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff,
                       int kSel, float *out, hipStream_t s);
void mfma_tile_shape(bool split, int *tx, int *ty);
int mfma_max_ksteps();
int mfma_max_half(bool split);
int mfma_shift(int h, int W);
int mfma_ksteps(int h, int W);
int mfma_tile_cols(int ksteps);
size_t mfma_lds_bytes(int h, int W, bool split);

// adjoint (k_adjoint.hip)
hipError_t launch_adj_edge(const AdjArgs &a, int T, const float *G1, const float *gpb, float *gOut,
                           hipStream_t s);
// G1 non-NULL: the incoming gradient is gpb + row inverse of G1 (the later
// step's edge fused in); NULL: it is read from g
hipError_t launch_adj_mid(const AdjArgs &a, int T, const float *U, const float *A, const float *g,
                          const float *G1, float *gpb, float *FG, double *accRows, hipStream_t s);
hipError_t launch_adj_cols(const AdjArgs &a, int T, const float *Fin, const float *Ahat, float *G,
                           double *accCols, hipStream_t s);
hipError_t launch_adj_final(const double *accRows, int nRows, int nvRows, const double *accCols,
                            int nCols, int nvCols, double *acc, hipStream_t s);
hipError_t launch_adj_records(const FftWorldDev *src, FftWorldDev *dst, int c, float *saveA,
                              float *saveU, float *saveAhat, size_t n, size_t spec, hipStream_t s);
size_t adj_lds_rows(const AdjArgs &a, int T);
size_t adj_lds_cols(const AdjArgs &a, int T);

// init and clock (k_init.hip)
hipError_t launch_init_patch(float *A, int W, int H, uint32_t seed, int x0, int y0, int pw,
                             int ph, hipStream_t s);
hipError_t launch_init_noise(float *A, int W, int H, uint32_t seed, float coverage, float scale,
                             hipStream_t s);
hipError_t launch_init_stamp(float *A, int W, int H, const float *org, int ow, int oh,
                             uint32_t count, uint32_t rotate, uint32_t seed, hipStream_t s);
hipError_t launch_set_clock(StepClock *c, uint32_t frame, uint32_t sub, uint32_t substeps,
                            uint32_t parity, hipStream_t s);
// move the stored clocks of n worlds `by` steps on (after the step kernels
// that read them with offsets 0 .. by-1)
hipError_t launch_advance_clocks(const FftWorldDev *worlds, int n, uint32_t by, hipStream_t s);
hipError_t launch_advance_clocks(const DirectWorldDev *worlds, int n, uint32_t by, hipStream_t s);
hipError_t launch_advance_clocks(const MfmaWorldDev *worlds, int n, uint32_t by, hipStream_t s);
// set n worlds' clocks to the start of `frame` with `substeps` substeps
hipError_t launch_frame_clocks(const FftWorldDev *worlds, int n, uint32_t frame, uint32_t substeps,
                               hipStream_t s);
hipError_t launch_frame_clocks(const DirectWorldDev *worlds, int n, uint32_t frame, uint32_t substeps,
                               hipStream_t s);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);

// colour map and activity (k_colormap.hip).  field srcW x srcH -> RGBA16F
// dstW x dstH (4 x uint16 per pixel, 8-byte aligned); bilinear when the sizes
// differ.  luma_stats: Rec.709 luma of n RGBA16F pixels -> {sum, sum of
// squares} in FP64 at out[0..1] (partial: 2 x nBlocks doubles of scratch).
hipError_t launch_colormap(const float *field, int srcW, int srcH, const ColorMapParams &p,
                           uint16_t *rgba, int dstW, int dstH, hipStream_t s);
int luma_stats_blocks(size_t n);
hipError_t launch_luma_stats(const uint16_t *rgba, size_t n, double *partial, double *out,
                             hipStream_t s);

}  // namespace lenia
