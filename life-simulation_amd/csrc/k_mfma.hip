// Direct-convolution form of the Lenia step on the gfx950 matrix cores
// (BF16 MFMA, FP32 accumulation), plain and split-BF16.
//
// Same operator as the reference's leniaStep (Shaders/Field/Lenia.metal:
// 63-128, extended to K kernels by leniaGrowthMulti :161-232), recast as a
// banded contraction.  For one window row dy and a 32 x 32 output block at
// (y0, x0):
//
//   u[y0+n][x0+m] += sum_j T_dy[m][j] * F_dy[j][n]
//   F_dy[j][n]  = A[y0+n+dy-h][x0-h+j]                     (the state, j < J)
//   T_dy[m][j]  = K[dy][j-m] if 0 <= j-m <= 2h, else 0     (banded Toeplitz)
//
// one 32x32xJ matrix product per window row, J = 16 KS >= 32 + 2h.  T is the
// MFMA A operand (rows = output columns x), F the B operand (columns = output
// rows y): the state operand is an aligned 8-element run of one LDS row per
// lane, and the shifted Toeplitz fragments -- the same for every tile -- are
// precomputed on the host in register order (64 lanes x 8 bf16 per k-step)
// and streamed from L2.
//
// The state tile (rows y0-h .. y0+32NBY-1+h, 16 (KS+2) columns) is staged in LDS as
// BF16, four columns per 16-byte load when W is a multiple of 4: the tile then
// starts `shift` = (-h mod 4) columns left of x0-h, so every load is aligned
// and none straddles the toroidal wrap, and the fragments carry the shift
// (K[dy][j - m - shift]).
//
// Work split: a workgroup owns a 64 (x) x 32 NBY (y) output tile, NB = 2 NBY
// blocks of 32 x 32 (NBY = 2 plain, 1 split).  Its 4 waves split the 2h+1
// window rows into quarters; each wave applies every fragment it loads to the
// blocks of both columns (the right column at the fragment of k-step s - 2) and
// every state operand it reads to both columns, NB independent accumulation
// chains; the quarters' sums are reduced through LDS in a fixed order, and
// wave b finishes block b.
//
// Split-BF16 (BF16X3): the state in two terms, F = Fhi + Flo (hi = bf16(v),
// lo = bf16(v - hi), residual <= 2^-18 |v|), the kernel in three, T = Thi +
// Tmid + Tlo (residual <= 2^-27 |T|), and four MFMAs per k-step:
// Tlo*Fhi + Thi*Flo + Tmid*Fhi + Thi*Fhi.  The kernel needs the third term
// because its split residuals are systematic (equal weights of a symmetric
// window round the same way, so a 2^-18 residual would add up coherently);
// the state residuals have random signs and average out over the window.
// Accumulation is two-level: each wave sums its quarter of the window rows,
// then the four quarter sums are added.
#include <hip/hip_runtime.h>

#include "lenia_device.h"
#include "launch.h"

namespace lenia {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMW = 4;         // waves per workgroup
constexpr int kMT = 64 * kMW;  // threads
constexpr int kMaxSteps = 8;   // k-steps of 16 per window row (J <= 128, h <= 48)
constexpr int kStage = 16;     // state-tile loads in flight per thread (one column each)
constexpr int kStage4 = 8;     // the same, four columns each

// output blocks of 32 x 32 per workgroup: NBX side by side in x, NBY stacked
// in y.  The NBX blocks of a row share every state operand: block bx at k-step
// s uses the kernel fragment of k-step s - 2 bx (its columns are 32 = 2 x 16
// further right), so a state operand read from LDS feeds NBX MFMAs.
template <bool SPLIT>
constexpr int blocks_x() { return 2; }
template <bool SPLIT>
constexpr int blocks_y() { return SPLIT ? 1 : 2; }
template <bool SPLIT>
constexpr int blocks_all() { return blocks_x<SPLIT>() * blocks_y<SPLIT>(); }

// reduction buffer: the three non-owner partials of each output block
template <bool SPLIT>
constexpr int red_floats() { return blocks_all<SPLIT>() * 3 * 64 * 16; }

// round-to-nearest-even FP32 -> BF16 bits (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short b) {
    return __uint_as_float((uint32_t)b << 16);
}

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

template <bool SPLIT, int KS>
struct Frags {
    u32x4 hi[KS];
    u32x4 mid[SPLIT ? KS : 1];
    u32x4 lo[SPLIT ? KS : 1];
};

// Toeplitz fragments of window row dy of kernel k: [k][dy][s][lane] x 16 B
template <bool SPLIT, int KS>
__device__ __forceinline__ void load_toeplitz(Frags<SPLIT, KS> &f, const MfmaWorldDev &wd, int k,
                                              int dy, int lane) {
    const size_t base = (((size_t)k * (2 * wd.half + 1) + dy) * KS) * 64 + lane;
    const LENIA_GLOBAL u32x4 *hi = gp(reinterpret_cast<const u32x4 *>(wd.toepHi)) + base;
#pragma unroll
    for (int s = 0; s < KS; ++s) f.hi[s] = hi[s * 64];
    if constexpr (SPLIT) {
        const LENIA_GLOBAL u32x4 *mid = gp(reinterpret_cast<const u32x4 *>(wd.toepMid)) + base;
        const LENIA_GLOBAL u32x4 *lo = gp(reinterpret_cast<const u32x4 *>(wd.toepLo)) + base;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            f.mid[s] = mid[s * 64];
            f.lo[s] = lo[s * 64];
        }
    }
}

// one window row applied to the NBX x NBY blocks, accumulated into
// total[bx NBY + by]; rowHi/rowLo point at the row of block row 0.  All state
// operands of the row are read from LDS first and the MFMAs issued after them
// (a scheduling barrier keeps the compiler from pairing each read with its
// MFMA, which would expose the LDS latency once per MFMA with a single wave on
// the SIMD).
template <bool SPLIT, int KS, int NBX, int NBY>
__device__ __forceinline__ void rows_mma(f32x16 (&total)[NBX * NBY], const Frags<SPLIT, KS> &f,
                                         const unsigned short *rowHi, const unsigned short *rowLo,
                                         int blockStride) {
    constexpr int KSX = KS + 2 * (NBX - 1);  // state k-steps a row of blocks spans
    u32x4 bh[KSX][NBY];
    u32x4 bl[SPLIT ? KSX : 1][SPLIT ? NBY : 1];
#pragma unroll
    for (int s = 0; s < KSX; ++s)
#pragma unroll
        for (int by = 0; by < NBY; ++by) {
            bh[s][by] = *reinterpret_cast<const u32x4 *>(rowHi + by * blockStride + 16 * s);
            if constexpr (SPLIT) bl[s][by] = *reinterpret_cast<const u32x4 *>(rowLo + by * blockStride + 16 * s);
        }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < KSX; ++s) {
#pragma unroll
        for (int bx = 0; bx < NBX; ++bx) {
            const int t = s - 2 * bx;  // kernel k-step of this block
            if (t < 0 || t >= KS) continue;
            const bf16x8 th = as_bf16x8(f.hi[t]);
#pragma unroll
            for (int by = 0; by < NBY; ++by) {
                f32x16 &acc = total[bx * NBY + by];
                const bf16x8 fh = as_bf16x8(bh[s][by]);
                if constexpr (SPLIT) {
                    const bf16x8 fl = as_bf16x8(bl[s][by]);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f.lo[t]), fh, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(th, fl, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(f.mid[t]), fh, acc, 0, 0, 0);
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(th, fh, acc, 0, 0, 0);
            }
        }
    }
}

// MODE 0: full step (src -> dst by clock parity); MODE 1: potential of kernel
// kSel of the current state into `out`
template <int MODE, bool SPLIT, int KS>
__global__ __launch_bounds__(kMT) void k_mfma(const MfmaWorldDev *__restrict__ worlds,
                                              const int2 *__restrict__ tiles, int nTiles,
                                              uint32_t stepOff, int kSel,
                                              float *__restrict__ out) {
    constexpr int NBX = blocks_x<SPLIT>(), NBY = blocks_y<SPLIT>(), NB = NBX * NBY;
    constexpr int KSX = KS + 2 * (NBX - 1);
    extern __shared__ unsigned short smemMf[];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    for (int ti = blockIdx.x; ti < nTiles; ti += gridDim.x) {
        const int2 tile = step_tile(tiles, ti);
        const MfmaWorldDev &wd = worlds[tile.x];
        const int W = wd.W, H = wd.H, h = wd.half;
        const int x0 = (tile.y % wd.tilesX) * (32 * NBX);
        const int y0 = (tile.y / wd.tilesX) * (32 * NBY);
        const int rows = 32 * NBY + 2 * h, cols = 16 * KSX, pitch = wd.pitch;
        const int plane = rows * pitch;
        unsigned short *tHi = smemMf;
        unsigned short *tLo = smemMf + plane;
        float *red = reinterpret_cast<float *>(smemMf + (SPLIT ? 2 : 1) * plane);
        const StepClock ck = clock_at(wd.clock, stepOff);
        const LENIA_GLOBAL float *src = gp(ck.parity ? wd.A1 : wd.A0);
        LENIA_GLOBAL float *dst = gp(ck.parity ? wd.A0 : wd.A1);
        // state tile: rows y0-h .. y0+32NB-1+h, columns xs .. xs+cols-1 with
        // xs = x0-h-shift (toroidal); kStage4 (kStage) loads in flight per
        // thread before any is stored
        const int xs = x0 - h - wd.shift;
        if (wd.vec4) {
            // W and xs multiples of 4: aligned 16-byte loads that never
            // straddle the wrap; rows and column groups wrap by one period
            // except near a field smaller than the tile (then by modulo)
            constexpr int G4 = 4 * KSX;  // float4 groups per staged row
            const int ngr = rows * G4;
            for (int base = 0; base < ngr; base += kMT * kStage4) {
                float4 v[kStage4];
                int at[kStage4];
#pragma unroll
                for (int u = 0; u < kStage4; ++u) {
                    const int i = base + u * kMT + tid;
                    at[u] = -1;
                    if (i < ngr) {
                        const int ly = i / G4, q = i - ly * G4;
                        int gy = y0 - h + ly, gx = xs + 4 * q;
                        if (gy < 0 || gy >= H) gy = (gy % H + H) % H;
                        if (gx < 0 || gx >= W) gx = (gx % W + W) % W;
                        v[u] = gld4(reinterpret_cast<const LENIA_GLOBAL float4 *>(src + (size_t)gy * W + gx));
                        at[u] = ly * pitch + 4 * q;
                    }
                }
#pragma unroll
                for (int u = 0; u < kStage4; ++u) {
                    if (at[u] >= 0) {
                        const float4 f = v[u];
                        const unsigned short b0 = bf16_bits(f.x), b1 = bf16_bits(f.y);
                        const unsigned short b2 = bf16_bits(f.z), b3 = bf16_bits(f.w);
                        *reinterpret_cast<uint2 *>(tHi + at[u]) =
                            make_uint2(b0 | ((uint32_t)b1 << 16), b2 | ((uint32_t)b3 << 16));
                        if constexpr (SPLIT) {
                            const unsigned short l0 = bf16_bits(f.x - bf16_value(b0));
                            const unsigned short l1 = bf16_bits(f.y - bf16_value(b1));
                            const unsigned short l2 = bf16_bits(f.z - bf16_value(b2));
                            const unsigned short l3 = bf16_bits(f.w - bf16_value(b3));
                            *reinterpret_cast<uint2 *>(tLo + at[u]) =
                                make_uint2(l0 | ((uint32_t)l1 << 16), l2 | ((uint32_t)l3 << 16));
                        }
                    }
                }
            }
        } else {
            const int nst = rows * cols;
            for (int base = 0; base < nst; base += kMT * kStage) {
                float v[kStage];
                int at[kStage];
#pragma unroll
                for (int u = 0; u < kStage; ++u) {
                    const int i = base + u * kMT + tid;
                    at[u] = -1;
                    if (i < nst) {
                        const int ly = i / cols, lx = i - ly * cols;
                        int gy = (y0 - h + ly) % H, gx = (xs + lx) % W;
                        if (gy < 0) gy += H;
                        if (gx < 0) gx += W;
                        v[u] = src[(size_t)gy * W + gx];
                        at[u] = ly * pitch + lx;
                    }
                }
#pragma unroll
                for (int u = 0; u < kStage; ++u) {
                    if (at[u] >= 0) {
                        const unsigned short hb = bf16_bits(v[u]);
                        tHi[at[u]] = hb;
                        if constexpr (SPLIT) tLo[at[u]] = bf16_bits(v[u] - bf16_value(hb));
                    }
                }
            }
        }
        __syncthreads();

        const GrowthParams &g = wd.g;
        const uint32_t seed = clock_seed(ck, g.moduleSeed);
        const int n = lane & 31, grp = lane >> 5;
        const int nrows = 2 * h + 1;
        const int r0 = (nrows * wave) / kMW, r1 = (nrows * (wave + 1)) / kMW;
        const unsigned short *colHi = tHi + n * pitch + 8 * grp;
        const unsigned short *colLo = tLo + n * pitch + 8 * grp;
        // wave b < NB finishes output block b = (b / NBY, b % NBY): its rows
        // y0 + 32 (b % NBY) + n, columns xq + 8 (r/4) + 4 grp + (r%4) of lane
        // slot r with xq = x0 + 32 (b / NBY)
        const int yb = y0 + 32 * (wave % NBY) + n;
        const int xq = x0 + 32 * (wave / NBY);
        // the owners' current state values, loaded now so that they arrive
        // while the window rows multiply
        // (four consecutive columns per 16-byte load when W is a multiple of 4)
        float aCur[16];
        if (MODE == 0 && wave < NB) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int x = xq + 8 * q + 4 * grp;
                if (wd.vec4 && x < W && yb < H) {
                    const float4 a4 = gld4(reinterpret_cast<const LENIA_GLOBAL float4 *>(src + (size_t)yb * W + x));
                    aCur[4 * q] = a4.x;
                    aCur[4 * q + 1] = a4.y;
                    aCur[4 * q + 2] = a4.z;
                    aCur[4 * q + 3] = a4.w;
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        aCur[4 * q + i] = (x + i < W && yb < H) ? src[(size_t)yb * W + x + i] : 0.0f;
                }
            }
        }
        float gacc[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) gacc[r] = 0.0f;
        const int kBeg = MODE == 1 ? kSel : 0;
        const int kEnd = MODE == 1 ? kSel + 1 : (int)g.nKernels;
        for (int k = kBeg; k < kEnd; ++k) {
            f32x16 total[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) total[b] = f32x16{};
            // kD fragment sets in flight: row dy + kD loads while row dy
            // multiplies (one row's MFMAs take less time than an L2 load)
            constexpr int kD = (SPLIT || KS > 6) ? 2 : 4;
            Frags<SPLIT, KS> f[kD];
#pragma unroll
            for (int i = 0; i < kD; ++i)
                if (r0 + i < r1) load_toeplitz<SPLIT, KS>(f[i], wd, k, r0 + i, lane);
            for (int dy = r0; dy < r1; dy += kD) {
#pragma unroll
                for (int i = 0; i < kD; ++i) {
                    const int row = dy + i;
                    if (row < r1) {
                        rows_mma<SPLIT, KS, NBX, NBY>(total, f[i], colHi + row * pitch, colLo + row * pitch,
                                                32 * pitch);
                        if (row + kD < r1) load_toeplitz<SPLIT, KS>(f[i], wd, k, row + kD, lane);
                    }
                }
            }
            // fixed-order reduction of block b into wave b over the window-row
            // quarters (wave b + i holds partial p_i of block b): every wave
            // stores its partials of the blocks it does not own, one barrier,
            // then owner b adds (p_0 + p_2) + (p_1 + p_3)
            static_assert(NB * 3 * 64 * 16 <= red_floats<SPLIT>(), "reduction buffer");
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int rel = (wave - b + kMW) & (kMW - 1);  // 0 = owner
                if (rel != 0)
#pragma unroll
                    for (int r = 0; r < 16; ++r) red[((b * 3 + rel - 1) * 16 + r) * 64 + lane] = total[b][r];
            }
            __syncthreads();
            float mine[16];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b == wave) {
                    const float *p1 = red + (b * 3 + 0) * 16 * 64 + lane;
                    const float *p2 = red + (b * 3 + 1) * 16 * 64 + lane;
                    const float *p3 = red + (b * 3 + 2) * 16 * 64 + lane;
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        mine[r] = (total[b][r] + p2[r * 64]) + (p1[r * 64] + p3[r * 64]);
                }
            }
            __syncthreads();  // the next kernel's partials reuse the buffer
            if (wave < NB) {
                // kernel k's growth constants, read once for the 16 outputs
                // (the arithmetic of cell_jitter and growth_term): a field
                // read per output is read again each time, after a full wait
                const float muK = g.mu[k], scK = g.bellScale[k], wK = g.weight[k], muJit = g.muJitter;
                const uint32_t gmode = g.growthMode;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int x = xq + 8 * (r >> 2) + 4 * grp + (r & 3);
                    if (MODE == 1) {
                        if (x < W && yb < H) gp(out)[(size_t)yb * W + x] = mine[r];
                    } else if (x < W && yb < H) {
                        float muJ = 0.0f;
                        if (muJit > 0.0f)
                            muJ = (rand01(x, yb, 401u + ck.frame % 613u, seed) - 0.5f) * muJit * 0.05f;
                        const float bl = bell_of(mine[r], muK + muJ, scK);
                        gacc[r] += gmode == 1u ? wK * bl : wK * (2.0f * bl - 1.0f);
                    }
                }
            }
        }
        if (MODE == 0 && wave < NB) {
            const GrowthParams gl = g;  // the finish constants in registers, read once
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int x = xq + 8 * (r >> 2) + 4 * grp + (r & 3);
                aCur[r] = growth_finish(gl, aCur[r], gacc[r], x, yb, W, H, seed, ck.frame);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int x = xq + 8 * q + 4 * grp;
                if (wd.vec4 && x < W && yb < H) {
                    gst4(reinterpret_cast<LENIA_GLOBAL float4 *>(dst + (size_t)yb * W + x),
                         make_float4(aCur[4 * q], aCur[4 * q + 1], aCur[4 * q + 2], aCur[4 * q + 3]));
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (x + i < W && yb < H) dst[(size_t)yb * W + x + i] = aCur[4 * q + i];
                }
            }
        }
        __syncthreads();
    }
}

template <int KS>
static void mfma_launch_ks(int mode, bool split, const MfmaWorldDev *worlds, const int2 *tiles,
                           int nTiles, int grid, size_t lds, uint32_t stepOff, int kSel,
                           float *out, hipStream_t s) {
    if (mode == 0) {
        if (split)
            launch_timed(k_mfma<0, true, KS>, dim3(grid), dim3(kMT), lds, s, worlds, tiles, nTiles,
                         stepOff, kSel, out);
        else
            launch_timed(k_mfma<0, false, KS>, dim3(grid), dim3(kMT), lds, s, worlds, tiles, nTiles,
                         stepOff, kSel, out);
    } else {
        if (split)
            launch_timed(k_mfma<1, true, KS>, dim3(grid), dim3(kMT), lds, s, worlds, tiles, nTiles,
                         stepOff, kSel, out);
        else
            launch_timed(k_mfma<1, false, KS>, dim3(grid), dim3(kMT), lds, s, worlds, tiles, nTiles,
                         stepOff, kSel, out);
    }
}

// ksteps = ceil((32 + 2h) / 16): 2 (h = 0) .. 8 (h <= 48)
hipError_t launch_mfma(int mode, bool split, int ksteps, const MfmaWorldDev *worlds,
                       const int2 *tiles, int nTiles, int grid, size_t lds, uint32_t stepOff,
                       int kSel, float *out, hipStream_t s) {
    switch (ksteps) {
        case 2: mfma_launch_ks<2>(mode, split, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 3: mfma_launch_ks<3>(mode, split, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 4: mfma_launch_ks<4>(mode, split, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 5: mfma_launch_ks<5>(mode, split, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 6: mfma_launch_ks<6>(mode, split, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 7: mfma_launch_ks<7>(mode, split, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        case 8: mfma_launch_ks<8>(mode, split, worlds, tiles, nTiles, grid, lds, stepOff, kSel, out, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

void mfma_tile_shape(bool split, int *tx, int *ty) {
    *tx = 32 * (split ? blocks_x<true>() : blocks_x<false>());
    *ty = 32 * (split ? blocks_y<true>() : blocks_y<false>());
}

int mfma_max_ksteps() { return kMaxSteps; }

// largest window half-width of each form (plain BF16 41, split-BF16 32): the
// register budget of the fragment sets, and the tested range
int mfma_max_half(bool split) { return split ? 32 : 41; }

// columns between the staged tile's first column and x0 - h: the tile then
// starts on a multiple of 4 (x0 is a multiple of 32), for 16-byte loads
int mfma_shift(int h, int W) { return W % 4 == 0 ? (4 - h % 4) % 4 : 0; }

// 16-wide k-steps per window row: the staged columns j < 16 ks must cover
// 32 + 2h + shift (the shift never adds a k-step: 32 + 2h is even and
// shift <= 1 when 2h = 14 mod 16)
int mfma_ksteps(int h, int W) { return (32 + 2 * h + mfma_shift(h, W) + 15) / 16; }

// staged state columns of a tile: the kernel's k-steps plus two per extra
// block in x (the forms share blocks_x)
int mfma_tile_cols(int ksteps) { return 16 * (ksteps + 2 * (blocks_x<false>() - 1)); }
static_assert(blocks_x<false>() == blocks_x<true>(), "one column count for both forms");

size_t mfma_lds_bytes(int h, int W, bool split) {
    const int ks = mfma_ksteps(h, W);
    const int pitch = mfma_tile_cols(ks) + 8;
    const int rows = 32 * (split ? blocks_y<true>() : blocks_y<false>()) + 2 * h;
    return (size_t)rows * pitch * sizeof(unsigned short) * (split ? 2 : 1) +
           (size_t)(split ? red_floats<true>() : red_floats<false>()) * sizeof(float);
}

}  // namespace lenia
