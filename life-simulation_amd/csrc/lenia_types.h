// Plain structs shared by host code and kernels.  Device copies of these live
// in HBM (one descriptor per world) so that hipGraph replays see updates made
// between replays.
#pragma once

#include <stdint.h>

namespace lenia {

constexpr int kMaxKernels = 4;
constexpr int kMaxStages = 8;  // FFT plan: up to 8 radix stages

// LDS entries that a transform region of n complex values spans: the
// placement i -> i + i/32 leaves one unused entry after every 32
constexpr int lds_span(int n) { return n + (n + 31) / 32; }

// LDS bytes one workgroup may use (gfx950: 160 KiB, less a margin)
constexpr size_t kFftLdsMax = 163840 - 256;

// The row pass at 1024 threads runs its transforms between two LDS regions
// (one barrier per stage instead of two) when the second region fits: the
// twiddles [W], the transform region, then its twin.  Host and kernel both
// decide by this rule, so the launch's LDS always covers what the kernel
// uses.
constexpr int rows_region(int W, int P, int K, bool kpar) { return lds_span((kpar ? K : 1) * P * W); }
constexpr bool rows_twin(int W, int P, int K, bool kpar, int T) {
    return T >= 1024 && (size_t)(W + 2 * rows_region(W, P, K, kpar)) * 8u <= kFftLdsMax;
}

// cosine-palette colour map (ColorMapParams, LifeCore/Render/ColorMapPass.h:16-25;
// same layout as lenia_colormap in the C ABI)
struct ColorMapParams {
    float a[3], b[3], c[3], d[3];
    float inputScale, inputBias, exposure;
    uint32_t channel;
};

// growth parameters of one world (the LeniaMultiParams fields that the
// growth rule reads, Shaders/Field/Lenia.metal:138-154)
struct GrowthParams {
    float dt;
    float mu[kMaxKernels];
    float sigma[kMaxKernels];
    float weight[kMaxKernels];
    float bellScale[kMaxKernels];  // sqrt(log2(e)/2) / sigma: bell = exp2(-(t*t))
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

// frame / substep position of a world.  A sequence of steps reads it with
// per-step offsets; a one-block kernel moves it on after the sequence.
struct StepClock {
    uint32_t frame;
    uint32_t sub;
    uint32_t substeps;
    uint32_t parity;  // which ping-pong buffer holds the current state (direct path)
    uint32_t steps;   // steps taken since the clock was set
    uint32_t pad[3];
};

// one Stockham stage: radix R, butterflies per transform NR = N/R, span Ns
// (product of the earlier radices), twiddle stride N/(Ns R) into the length-N
// table, offset of this stage's [Ns][R-1] twiddle table, and division magics
// ceil(2^32/d) for d = NR and d = Ns (0 when d = 1)
struct FftStage {
    int32_t radix, NR, Ns, twStep, twOff;
    uint32_t mNR, mNs;
    int32_t pad;
};

// FFT plan of one length: stages in order
// Radix sequences that also have a compile-time kernel (the default plans of
// common lengths).  The fixed kernels run the same butterflies in
// the same order as the table-driven one, so both give the same bits.
enum : uint32_t {
    kPlanGeneric = 0,
    kPlan64 = 1,    // 8 8
    kPlan128 = 2,   // 16 8
    kPlan256 = 3,   // 16 16
    kPlan512 = 4,   // 8 8 8
    kPlan1024 = 5,  // 16 8 8
    kPlan2048 = 6,  // 16 16 8
    kPlan96 = 7,    // 8 4 3
    kPlan192 = 8,   // 8 8 3
    kPlan384 = 9,   // 16 8 3
    kPlan720 = 10,  // 16 3 3 5
    kPlan1080 = 11, // 8 3 3 3 5
    kPlan1280 = 12, // 16 16 5
    kPlan1920 = 13, // 16 8 3 5
    kPlan4096 = 14, // 16 16 16
    // other radix orders of the benchmark lengths: candidates of the radix-plan
    // autotuner (lenia_plan_autotune)
    kPlan64b = 15,    // 16 4
    kPlan64c = 16,    // 4 16
    kPlan128b = 17,   // 8 16
    kPlan128c = 18,   // 4 4 8
    kPlan256b = 19,   // 8 8 4
    kPlan256c = 20,   // 4 8 8
    kPlan512b = 21,   // 16 8 4
    kPlan512c = 22,   // 4 8 16
    kPlan1024b = 23,  // 8 8 16
    kPlan1024c = 24,  // 16 16 4
    kPlan1024d = 25,  // 4 16 16
    kPlan2048b = 26,  // 8 16 16
    kPlan4096b = 27,  // 8 8 8 8
    kPlan96b = 28,    // 4 8 3
    kPlan192b = 29,   // 16 4 3
    kPlan384b = 30,   // 8 16 3
    kPlan384c = 31,   // 8 4 4 3: every radix divides 24, so 384 runs on the wave executor
};

struct FftPlan {
    int32_t n;
    int32_t nStages;
    uint32_t code;  // kPlan*: which compile-time kernel runs this plan
    int32_t pad;
    FftStage st[kMaxStages];
};

// launch shape of one pass: threads per workgroup and values per thread
struct PassShape {
    int32_t threads;  // 64, 128 or 256
    int32_t maxv;     // 16 or 32
};

// device descriptor of one FFT-path world
struct FftWorldDev {
    int32_t W, H, Wh;  // Wh = W/2 + 1 half-spectrum columns
    int32_t P;         // row pairs per row-pass tile
    int32_t C;         // spectrum columns per column-pass tile
    int32_t rowTiles, colTiles;
    int32_t K;
    uint32_t mP;                    // division magic for P
    int32_t rowThreads, colThreads; // workgroup sizes of the two passes
    uint32_t mW;                    // division magic for W
    FftPlan planW, planH;
    GrowthParams g;
    const float *twW;  // interleaved complex twiddles e^{-2 pi i t / W}, t < W
    const float *twH;
    float *A;          // state, H x W row-major (updated in place)
    float *F;          // state half-spectrum after the row pass: [Wh][H] complex
    float *G;          // per-kernel column-pass output: [K][Wh][H] complex
    const float *S;    // per-kernel real spectra / (W H): [K][Wh][H]
    StepClock *clock;
    // optional copies of a step's intermediates for the reverse pass (null
    // normally): the state before the step (H x W), the potentials u_k
    // ([K][H][W]) and the 2D spectrum of the state ([Wh][H] complex)
    float *saveA, *saveU, *saveAhat;
};

// device descriptor of one direct-path world
struct DirectWorldDev {
    int32_t W, H;
    int32_t half;       // window half-width (max over kernels)
    int32_t tilesX, tilesY;
    int32_t npad;       // window row pitch: 2 half + 1 rounded up to 8
    int32_t tileH;      // output tile height the tiles count in (32 or 16)
    uint32_t colsMagic; // ceil(2^32 / (64 + npad)): halo index / columns by a high multiply
    GrowthParams g;
    const float *win;   // [K][2 half + 1][npad] float windows (zero padded)
    const int2 *rowChunks;  // [K][2 half + 1]: first and past-last 8-tap chunk with a non-zero weight
    float *A0, *A1;     // ping-pong states
    StepClock *clock;
};

// MFMA direct form: the direct-form world plus the precomputed Toeplitz
// fragments of every kernel row ([k][dy][kstep][lane] x 8 bf16)
struct MfmaWorldDev {
    int32_t W, H;
    int32_t half;       // window half-width (max over kernels)
    int32_t tilesX, tilesY;
    int32_t ksteps;     // 16-wide k-steps per window row: ceil((32 + 2 half) / 16)
    int32_t tileCols;   // staged tile columns: 16 (ksteps + 2), two blocks side by side
    int32_t pitch;      // LDS row pitch in bf16 (tileCols + 8)
    int32_t vec4;       // W % 4 == 0: the state tile is staged four columns per load
    int32_t shift;      // staged columns start at x0 - half - shift (a multiple of 4 when vec4)
    GrowthParams g;
    const uint16_t *toepHi, *toepMid, *toepLo;  // kernel = hi + mid + lo (bf16 planes)
    float *A0, *A1;     // ping-pong states
    StepClock *clock;
};

// arguments of the adjoint kernels (passed by value)
struct AdjArgs {
    int32_t W, H, K, nShells;
    int32_t shellK[16];  // kernel of each (k, j) shell pair
    FftPlan planW, planH;
    const GrowthParams *gDev;  // the world's growth parameters (device copy)
    const float *twW, *twH;  // stage twiddle tables of the world's plans
    const float *S;          // [K][Wh][H] kernel spectra / (W H)
    const float *D;          // [nShells][Wh][H] shell-gradient spectra
};

}  // namespace lenia
