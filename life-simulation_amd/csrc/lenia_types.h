// Plain structs shared by host code and kernels.  Device copies of these live
// in HBM (one descriptor per world) so that hipGraph replays see updates made
// between replays.
#pragma once

#include <stdint.h>

namespace lenia {

constexpr int kMaxKernels = 4;
constexpr int kMaxStages = 8;  // FFT plan: up to 8 radix stages

// growth parameters of one world (the LeniaMultiParams fields that the
// growth rule reads, Shaders/Field/Lenia.metal:138-154)
struct GrowthParams {
    float dt;
    float mu[kMaxKernels];
    float sigma[kMaxKernels];
    float weight[kMaxKernels];
    float sumWeights;
    uint32_t nKernels;
    uint32_t growthMode;
    float noiseAmount;
    float muJitter;
    float injectAmount;
    uint32_t injectCount;
    uint32_t radius;
    uint32_t moduleSeed;  // seed_ = deriveSeed(sceneSeed, 'LEN1')
};

// frame / substep position of a world, advanced on the device by the last
// kernel of every step
struct StepClock {
    uint32_t frame;
    uint32_t sub;
    uint32_t substeps;
    uint32_t parity;  // which ping-pong buffer holds the current state (direct path)
    uint32_t steps;   // steps taken since the clock was set
    uint32_t pad[3];
};

// FFT plan of one length: radices in stage order
struct FftPlan {
    int32_t n;
    int32_t nStages;
    int32_t radix[kMaxStages];
};

// device descriptor of one FFT-path world
struct FftWorldDev {
    int32_t W, H, Wh;  // Wh = W/2 + 1 half-spectrum columns
    int32_t P;         // row pairs per row-pass tile
    int32_t C;         // spectrum columns per column-pass tile
    int32_t rowTiles, colTiles;
    int32_t pad0;
    FftPlan planW, planH;
    GrowthParams g;
    const float *twW;  // interleaved complex twiddles e^{-2 pi i t / W}, t < W
    const float *twH;
    float *A;          // state, H x W row-major (updated in place)
    float *F;          // state half-spectrum after the row pass: [Wh][H] complex
    float *G;          // per-kernel column-pass output: [K][Wh][H] complex
    const float *S;    // per-kernel real spectra / (W H): [K][Wh][H]
    StepClock *clock;
};

// device descriptor of one direct-path world
struct DirectWorldDev {
    int32_t W, H;
    int32_t half;       // window half-width (max over kernels)
    int32_t tilesX, tilesY;
    int32_t pad0[3];
    GrowthParams g;
    const float *win;   // [K][(2 half+1)^2] float windows (zero padded to half)
    float *A0, *A1;     // ping-pong states
    StepClock *clock;
};

}  // namespace lenia
