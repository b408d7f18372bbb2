// Direct-convolution form of the Lenia step (FP32 vector path) on gfx950.
//
// Same arithmetic as the reference's leniaStep (Shaders/Field/Lenia.metal:
// 63-128): u = sum_{dy,dx} K[dy][dx] A[wrap(y+dy)][wrap(x+dx)], growth, clip,
// extended to up to four kernels with the leniaGrowthMulti rule
// (Lenia.metal:161-232).  A workgroup stages a (16+2h) x (64+2h) halo tile of
// the toroidal state in LDS, each thread produces a column of M=4 outputs and
// reuses every LDS value for all M outputs it touches.  Each output's sum is
// taken per window row first, then across rows (two-level summation keeps the
// FP32 potential within ~1e-6 of the FP64 value for 3000-tap windows).
#include <hip/hip_runtime.h>

#include "lenia_device.h"
#include "launch.h"

namespace lenia {

constexpr int kDT = 256;  // threads: 64 x 4
constexpr int kDTX = 64;
constexpr int kDM = 4;    // outputs per thread (vertical)
constexpr int kDTY = 4 * kDM;


__device__ __forceinline__ int wrap_i(int p, int n) {
    int r = p % n;
    return r < 0 ? r + n : r;
}

// MODE 0: full step (src -> dst by clock parity); MODE 1: potential of kernel
// kSel of the current state into `out`
template <int MODE>
__global__ __launch_bounds__(kDT) void k_direct(const DirectWorldDev *__restrict__ worlds,
                                                const int2 *__restrict__ tiles, int nTiles,
                                                uint32_t *done, int nWorlds, int kSel,
                                                float *__restrict__ out) {
    extern __shared__ float tileS[];
    const int tid = threadIdx.x;
    const int tx = tid % kDTX, ty = tid / kDTX;
    for (int ti = blockIdx.x; ti < nTiles; ti += gridDim.x) {
        const int2 tile = tiles[ti];
        const DirectWorldDev &wd = worlds[tile.x];
        const int W = wd.W, H = wd.H, h = wd.half, n = 2 * h + 1;
        const int x0 = (tile.y % wd.tilesX) * kDTX;
        const int y0 = (tile.y / wd.tilesX) * kDTY;
        const int pitch = kDTX + 2 * h, rows = kDTY + 2 * h;
        const StepClock ck = read_clock(wd.clock);
        const LENIA_GLOBAL float *src = gp(ck.parity ? wd.A1 : wd.A0);
        LENIA_GLOBAL float *dst = gp(ck.parity ? wd.A0 : wd.A1);
        for (int i = tid; i < rows * pitch; i += kDT) {
            int ly = i / pitch, lx = i - ly * pitch;
            int gy = wrap_i(y0 - h + ly, H), gx = wrap_i(x0 - h + lx, W);
            tileS[i] = src[gy * W + gx];
        }
        __syncthreads();

        const GrowthParams &g = wd.g;
        const uint32_t seed = clock_seed(ck, g.moduleSeed);
        const int x = x0 + tx;
        float acc[kDM];
#pragma unroll
        for (int j = 0; j < kDM; ++j) acc[j] = 0.0f;
        const int kBeg = MODE == 1 ? kSel : 0;
        const int kEnd = MODE == 1 ? kSel + 1 : (int)g.nKernels;
        for (int k = kBeg; k < kEnd; ++k) {
            const LENIA_GLOBAL float *wk = gp(wd.win) + k * n * n;
            float tot[kDM];
#pragma unroll
            for (int j = 0; j < kDM; ++j) tot[j] = 0.0f;
            for (int r = 0; r < kDM + 2 * h; ++r) {
                const float *trow = tileS + (ty * kDM + r) * pitch + tx;
                float rs[kDM];
#pragma unroll
                for (int j = 0; j < kDM; ++j) rs[j] = 0.0f;
                for (int ix = 0; ix < n; ++ix) {
                    float v = trow[ix];
#pragma unroll
                    for (int j = 0; j < kDM; ++j) {
                        int iy = r - j;
                        if (iy >= 0 && iy < n) rs[j] = fmaf(wk[iy * n + ix], v, rs[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < kDM; ++j) {
                    int iy = r - j;
                    if (iy >= 0 && iy < n) tot[j] += rs[j];
                }
            }
#pragma unroll
            for (int j = 0; j < kDM; ++j) {
                int y = y0 + ty * kDM + j;
                if (MODE == 1) {
                    if (x < W && y < H) gp(out)[y * W + x] = tot[j];
                } else if (x < W && y < H) {
                    float muJ = cell_jitter(g, x, y, seed, ck.frame);
                    acc[j] += growth_term(g, k, tot[j], muJ);
                }
            }
        }
        if (MODE == 0) {
#pragma unroll
            for (int j = 0; j < kDM; ++j) {
                int y = y0 + ty * kDM + j;
                if (x < W && y < H) {
                    float a = tileS[(ty * kDM + j + h) * pitch + tx + h];
                    dst[y * W + x] = growth_finish(g, a, acc[j], x, y, W, H, seed, ck.frame);
                }
            }
        }
        __syncthreads();
    }
    if (MODE == 0) arrive_step(done, worlds, nWorlds);
}

hipError_t launch_direct(int mode, const DirectWorldDev *worlds, const int2 *tiles, int nTiles,
                         int grid, size_t lds, uint32_t *done, int nWorlds, int kSel, float *out,
                         hipStream_t s) {
    if (mode == 0)
        launch_timed(k_direct<0>, dim3(grid), dim3(kDT), lds, s, worlds, tiles, nTiles, done,
                     nWorlds, kSel, out);
    else
        launch_timed(k_direct<1>, dim3(grid), dim3(kDT), lds, s, worlds, tiles, nTiles, done,
                     nWorlds, kSel, out);
    return hipGetLastError();
}

void direct_tile_shape(int *tx, int *ty) {
    *tx = kDTX;
    *ty = kDTY;
}

}  // namespace lenia
