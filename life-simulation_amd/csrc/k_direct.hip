// Direct-convolution form of the Lenia step (FP32 vector path) on gfx950.
//
// Same arithmetic as the reference's leniaStep (Shaders/Field/Lenia.metal:
// 63-128): u = sum_{dy,dx} K[dy][dx] A[wrap(y+dy)][wrap(x+dx)], growth, clip,
// extended to up to four kernels with the leniaGrowthMulti rule
// (Lenia.metal:161-232).
//
// A workgroup owns a 64 (x) x TY (y) output tile, TY = 32, or 16 for a world
// whose 32-row tiles would leave half the CUs without one, and stages its
// (TY+2h) x (64+P) toroidal halo in LDS once (P = window width rounded up to
// 8; the window rows are zero padded to P).  Each output's potential is the sum over
// the first ceil(n/2) window rows plus the sum over the others.  In the
// 256-thread form one thread produces 8 horizontally adjacent outputs of one
// row, both halves in turn; in the split form (SPLIT, 16 TY threads, for
// grids of at most one tile per CU) a thread of the first half of the
// workgroup sums the first window-row half and its partner in the second
// half the other, and the first adds the partner's sums (through LDS).  Every
// form and tile height computes the same bits.  For every window row it slides
// a 16-value register window along the staged row, 8 taps at a time: two
// 16-byte LDS reads of state and two of weights (the kernel's window is
// staged in LDS too) feed 64 FMAs, with the next chunk's reads issued before
// the current chunk's FMAs.
// Each output's sum is taken per window row first, then across rows of each
// half, then the two halves (this summation keeps the FP32 potential within
// ~1e-6 of the FP64 value for 3000-tap windows).
#include <hip/hip_runtime.h>

#include "lenia_device.h"
#include "launch.h"

namespace lenia {

constexpr int kDTX = kDirectTileW;  // output tile width
constexpr int kDTY = kDirectTileH;  // output tile height (the 16-row form: kDTYSmall)
constexpr int kDTYSmall = 16;
constexpr int kDOX = 8;    // outputs per thread (x)
constexpr int kDStage = 16;

__device__ __forceinline__ int wrap_i(int p, int n) {
    int r = p % n;
    return r < 0 ? r + n : r;
}

// 8 consecutive LDS floats, 16-byte aligned (two ds_read_b128)
__device__ __forceinline__ void lds_read8(const float4 *p, float (&v)[8]) {
    const float4 a = p[0], b = p[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
    v[4] = b.x;
    v[5] = b.y;
    v[6] = b.z;
    v[7] = b.w;
}


// 8 taps of one window row: out[i] += w[t] * in[i + t], in = lo ++ hi
__device__ __forceinline__ void taps8(float (&rs)[8], const float (&w)[8], const float (&lo)[8],
                                      const float (&hi)[8]) {
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 8; ++i) rs[i] = fmaf(w[t], (i + t < 8) ? lo[i + t] : hi[i + t - 8], rs[i]);
}

// MODE 0: full step (src -> dst by clock parity); MODE 1: potential of kernel
// kSel of the current state into `out`.  TY: output tile height; a window-row
// half takes 8 TY threads (8 x TY outputs of 8).  SPLIT: 16 TY threads, one
// window-row half each; else 8 TY threads, both halves each.
template <int MODE, bool SPLIT, int TY>
__global__ __launch_bounds__(SPLIT ? 16 * TY : 8 * TY) void k_direct(
    const DirectWorldDev *__restrict__ worlds, const int2 *__restrict__ tiles, int nTiles, uint32_t stepOff,
    int kSel, float *__restrict__ out) {
    constexpr int kDHalf = 8 * TY;  // threads of one window-row half
    constexpr int kDT = SPLIT ? 2 * kDHalf : kDHalf;
    extern __shared__ float tileS[];
    const int tid = threadIdx.x;
    const int half = SPLIT && tid >= kDHalf, ht = tid - half * kDHalf;  // window-row half, thread in it
    const int tx = ht & 7, ty = ht >> 3;
    for (int ti = blockIdx.x; ti < nTiles; ti += gridDim.x) {
        const int2 tile = step_tile(tiles, ti);
        const DirectWorldDev &wd = worlds[tile.x];
        const int W = wd.W, H = wd.H, h = wd.half, np = wd.npad;
        const int x0 = (tile.y % wd.tilesX) * kDTX;
        const int y0 = (tile.y / wd.tilesX) * TY;
        const int rows = TY + 2 * h, cols = kDTX + np, pitch = cols + 4;
        const StepClock ck = clock_at(wd.clock, stepOff);
        const LENIA_GLOBAL float *src = gp(ck.parity ? wd.A1 : wd.A0);
        LENIA_GLOBAL float *dst = gp(ck.parity ? wd.A0 : wd.A1);
        // halo rows y0-h .. y0+TY-1+h, columns x0-h .. x0-h+cols-1 (toroidal;
        // columns past x0+63+h only meet zero weights).  The wrapped row
        // offsets and columns are tabulated in LDS first (one modulo per
        // entry), so staging a value takes a table lookup, not two modulos
        // and a division.
        const int n = 2 * h + 1;
        // (after the halo, the window and, in the split form, the second
        // half's sums)
        int *rowOff = reinterpret_cast<int *>(tileS + rows * pitch + n * np + (SPLIT ? 8 * kDHalf : 0));
        int *colIx = rowOff + rows;                                                          // [cols]
        // the current kernel's per-row chunk ranges [n], 8-byte aligned
        int2 *rcS = reinterpret_cast<int2 *>(rowOff + ((rows + cols + 1) & ~1));
        for (int i = tid; i < rows + cols; i += kDT) {
            if (i < rows)
                rowOff[i] = wrap_i(y0 - h + i, H) * W;
            else
                colIx[i - rows] = wrap_i(x0 - h + (i - rows), W);
        }
        __syncthreads();
        const int nst = rows * cols;
        const uint32_t colsMagic = wd.colsMagic;  // one scalar load, not one per staged value
        for (int base = 0; base < nst; base += kDT * kDStage) {
            float v[kDStage];
            int at[kDStage];
#pragma unroll
            for (int u = 0; u < kDStage; ++u) {
                const int i = base + u * kDT + tid;
                at[u] = -1;
                if (i < nst) {
                    // i / cols by the record's magic (exact: i < 2^16, cols < 2^8)
                    const int ly = (int)__umulhi((uint32_t)i, colsMagic), lx = i - ly * cols;
                    v[u] = src[rowOff[ly] + colIx[lx]];
                    at[u] = ly * pitch + lx;
                }
            }
#pragma unroll
            for (int u = 0; u < kDStage; ++u)
                if (at[u] >= 0) tileS[at[u]] = v[u];
        }
        __syncthreads();

        const GrowthParams &g = wd.g;
        const uint32_t seed = clock_seed(ck, g.moduleSeed);
        const int y = y0 + ty, xb = x0 + kDOX * tx;
        float gacc[kDOX];
#pragma unroll
        for (int i = 0; i < kDOX; ++i) gacc[i] = 0.0f;
        const int kBeg = MODE == 1 ? kSel : 0;
        const int kEnd = MODE == 1 ? kSel + 1 : (int)g.nKernels;
        const float4 *tile4 = reinterpret_cast<const float4 *>(tileS);
        const int pitch4 = pitch / 4;
        float *winS = tileS + rows * pitch;  // the current kernel's window, [n][np]
        float4 *part4 = reinterpret_cast<float4 *>(winS + n * np);  // SPLIT: second-half sums [2][kDHalf]
        const int nh = (n + 1) / 2;
        for (int k = kBeg; k < kEnd; ++k) {
            // stage kernel k's window in LDS: every load of the tap loop is
            // then an LDS load, and those complete in order, so the loop can
            // wait for the oldest of its loads while the newer ones fly
            __syncthreads();  // the previous kernel's window is no longer read
            {
                const LENIA_GLOBAL float4 *wg =
                    gp(reinterpret_cast<const float4 *>(wd.win)) + (size_t)k * n * (np / 4);
                float4 *ws4 = reinterpret_cast<float4 *>(winS);
                for (int i = tid; i < n * (np / 4); i += kDT) ws4[i] = gld4(wg + i);
                // its row chunk ranges too: read from LDS (in order with the
                // tap loop's other LDS reads) a row's range needs no wait for
                // every outstanding load, as a scalar load's would
                const LENIA_GLOBAL int *rcg = gp(reinterpret_cast<const int *>(wd.rowChunks)) + (size_t)2 * k * n;
                for (int i = tid; i < n; i += kDT) rcS[i] = make_int2(rcg[2 * i], rcg[2 * i + 1]);
            }
            __syncthreads();
            // the sum of window rows iy0 .. iy1-1 into tot
            auto rows_sum = [&](int iy0, int iy1, float (&tot)[kDOX]) {
#pragma unroll
                for (int i = 0; i < kDOX; ++i) tot[i] = 0.0f;
                // the next row's chunk range is read a row ahead
                int2 rNext = rcS[iy0];
                for (int iy = iy0; iy < iy1; ++iy) {
                    const float4 *trow = tile4 + (ty + iy) * pitch4 + 2 * tx;
                    const float4 *wrow = reinterpret_cast<const float4 *>(winS) + iy * (np / 4);
                    // chunks [c0, c1) of this row hold its non-zero weights
                    // (the same for every lane)
                    const int c0 = __builtin_amdgcn_readfirstlane(rNext.x);
                    const int c1 = __builtin_amdgcn_readfirstlane(rNext.y);
                    if (iy + 1 < iy1) rNext = rcS[iy + 1];
                    float rs[kDOX];
#pragma unroll
                    for (int i = 0; i < kDOX; ++i) rs[i] = 0.0f;
                    // va, vb: state chunks c, c+1; w0: weights of chunk c.  Each
                    // iteration loads chunk c+2 and its weights before the taps of
                    // chunk c, and chunk c+3 and the weights of c+2 before the
                    // taps of chunk c+1 (reads stay inside the row: chunk c+3 is
                    // loaded only when c+3 <= c1)
                    float va[8], vb[8], w0[8];
                    int c = c0;
                    lds_read8(trow + 2 * c, va);
                    lds_read8(trow + 2 * c + 2, vb);
                    lds_read8(wrow + 2 * c, w0);
                    for (; c + 1 < c1; c += 2) {
                        float vc[8], w1[8];
                        lds_read8(trow + 2 * c + 4, vc);
                        lds_read8(wrow + 2 * c + 2, w1);
                        taps8(rs, w0, va, vb);
                        if (c + 3 <= c1) lds_read8(trow + 2 * c + 6, va);
                        if (c + 2 < c1) lds_read8(wrow + 2 * c + 4, w0);
                        taps8(rs, w1, vb, vc);
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const float t = va[i];
                            va[i] = vc[i];
                            vb[i] = t;
                        }
                    }
                    if (c < c1) taps8(rs, w0, va, vb);
#pragma unroll
                    for (int i = 0; i < kDOX; ++i) tot[i] += rs[i];
                }
            };
            float lo[kDOX], hi[kDOX];
            if (SPLIT) {
                // the first half adds its partner's sums
                if (half) {
                    rows_sum(nh, n, hi);
                    part4[ht] = make_float4(hi[0], hi[1], hi[2], hi[3]);
                    part4[kDHalf + ht] = make_float4(hi[4], hi[5], hi[6], hi[7]);
                } else {
                    rows_sum(0, nh, lo);
                }
                __syncthreads();
                if (half) continue;
                const float4 a = part4[ht], b = part4[kDHalf + ht];
                hi[0] = a.x;
                hi[1] = a.y;
                hi[2] = a.z;
                hi[3] = a.w;
                hi[4] = b.x;
                hi[5] = b.y;
                hi[6] = b.z;
                hi[7] = b.w;
            } else {
                rows_sum(0, nh, lo);
                rows_sum(nh, n, hi);
            }
            float tot[kDOX];
#pragma unroll
            for (int i = 0; i < kDOX; ++i) tot[i] = lo[i] + hi[i];
            // kernel k's growth constants, read once for the 8 outputs (the
            // arithmetic of cell_jitter and growth_term)
            const float muK = g.mu[k], scK = g.bellScale[k], wK = g.weight[k], muJit = g.muJitter;
            const uint32_t gmode = g.growthMode;
#pragma unroll
            for (int i = 0; i < kDOX; ++i) {
                const int x = xb + i;
                if (MODE == 1) {
                    if (x < W && y < H) gp(out)[(size_t)y * W + x] = tot[i];
                } else if (x < W && y < H) {
                    float muJ = 0.0f;
                    if (muJit > 0.0f) muJ = (rand01(x, y, 401u + ck.frame % 613u, seed) - 0.5f) * muJit * 0.05f;
                    const float b = bell_of(tot[i], muK + muJ, scK);
                    gacc[i] += gmode == 1u ? wK * b : wK * (2.0f * b - 1.0f);
                }
            }
        }
        if (MODE == 0 && !half) {
            // the finish constants copied to registers once, and every
            // output computed before the stores (no store between the uses
            // of a constant, so none is read again)
            const GrowthParams gl = g;
            float res[kDOX];
#pragma unroll
            for (int i = 0; i < kDOX; ++i)
                res[i] = growth_finish(gl, tileS[(ty + h) * pitch + kDOX * tx + i + h], gacc[i], xb + i, y, W, H,
                                       seed, ck.frame);
#pragma unroll
            for (int i = 0; i < kDOX; ++i) {
                const int x = xb + i;
                if (x < W && y < H) dst[(size_t)y * W + x] = res[i];
            }
        }
        __syncthreads();
    }
}

template <bool SPLIT, int TY>
static void direct_launch(int mode, const DirectWorldDev *worlds, const int2 *tiles, int nTiles, int grid,
                          size_t lds, uint32_t stepOff, int kSel, float *out, hipStream_t s) {
    constexpr int T = SPLIT ? 16 * TY : 8 * TY;
    if (mode == 0)
        launch_timed(k_direct<0, SPLIT, TY>, dim3(grid), dim3(T), lds, s, worlds, tiles, nTiles, stepOff, kSel,
                     out);
    else
        launch_timed(k_direct<1, SPLIT, TY>, dim3(grid), dim3(T), lds, s, worlds, tiles, nTiles, stepOff, kSel,
                     out);
}

hipError_t launch_direct(int mode, const DirectWorldDev *worlds, const int2 *tiles, int nTiles,
                         int grid, size_t lds, uint32_t stepOff, int kSel, float *out,
                         hipStream_t s, int numCUs, int tileH) {
    // grids that leave CUs without a workgroup take the split form
    // (bit-identical); from one tile per CU on, the one-thread-per-8-outputs
    // form, which keeps two workgroups per CU (1280x720 R=32: 460 tiles, 113
    // against 124 us)
    if (tileH == kDTYSmall) {
        if (direct_split(nTiles, numCUs))
            direct_launch<true, kDTYSmall>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
        else
            direct_launch<false, kDTYSmall>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
    } else if (tileH == kDTY) {
        if (direct_split(nTiles, numCUs))
            direct_launch<true, kDTY>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
        else
            direct_launch<false, kDTY>(mode, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// 64 x 32 tiles, or 64 x 16 where 32-row tiles would number at most half the
// CUs (512x512: 128 tiles of 64 x 32 leave half the CUs idle; 256 of 64 x 16
// give every CU one)
void direct_tile_shape(int W, int H, int numCUs, int *tx, int *ty) {
    *tx = kDTX;
    const long tiles32 = (long)((W + kDTX - 1) / kDTX) * ((H + kDTY - 1) / kDTY);
    *ty = 2 * tiles32 <= numCUs ? kDTYSmall : kDTY;
}

// window row pitch (floats): 2h+1 rounded up to 8
int direct_npad(int h) { return (2 * h + 1 + 7) & ~7; }

bool direct_split(int nTiles, int numCUs) { return nTiles <= numCUs; }

size_t direct_lds_bytes(int h, int tileH, bool split) {
    const int pitch = kDTX + direct_npad(h) + 4;
    // the halo tile, one kernel window [2h+1][npad], the second half's sums,
    // the wrapped row offsets and columns of the halo
    // and the current kernel's row chunk ranges
    const int rows = tileH + 2 * h, cols = kDTX + direct_npad(h);
    return ((size_t)rows * pitch + (size_t)(2 * h + 1) * direct_npad(h)) * sizeof(float) +
           (split ? (size_t)2 * 8 * tileH * sizeof(float4) : 0) + (size_t)((rows + cols + 1) & ~1) * sizeof(int) +
           (size_t)(2 * h + 1) * sizeof(int2);
}

}  // namespace lenia
