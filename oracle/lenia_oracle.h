/*
 * lenia_oracle.h -- FP64 CPU restatement of the LifeSimSuite Lenia hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: it is the checker that the tests, the smoke entry point and
 * the CPU-baseline leg of the benchmark run beside the HIP implementation.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  All arithmetic is double precision; the
 * integer hashes are exact uint32 arithmetic.
 *
 * Pinning of each function (what independent check fails on a plausible
 * mistake) is listed in the header comment of lenia_oracle.c.
 */
#ifndef LENIA_ORACLE_H
#define LENIA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel shapes: which reference builder a kernel follows */
#define LO_SHAPE_SHELL 0 /* buildKernelTexture, Lenia.cpp:410-436 */
#define LO_SHAPE_RING 1  /* buildRingKernelImage, Lenia.cpp:95-125 */
/* kernel cores */
#define LO_CORE_GAUSSIAN 0 /* exp(-t^2/2): the reference's only core */
#define LO_CORE_EXP 1      /* exp(4 - 1/(r(1-r))): exponential bump core */

typedef struct {
    int32_t shape;
    int32_t core;
    double radius;     /* R_k in cells */
    double shellMu;    /* shell centre (SHELL shape only; RING uses 0.5) */
    double shellSigma; /* shell width */
    int32_t nBetas;
    double betas[4];
    double mu, sigma, weight; /* growth parameters of this kernel */
} lo_kernel;

typedef struct {
    int32_t width, height;
    int32_t growthMode; /* 0 = 2*bell-1, 1 = asymptotic bell-A */
    double dt;
    double noiseAmount, muJitter, injectAmount;
    int32_t injectCount;
    int32_t radius; /* module radius (only the inject blob size reads it) */
    int32_t nKernels;
    lo_kernel k[4];
} lo_world;

/* ---- integer hashes (Random.h:34-44, Random.metal:5-19) ---- */
uint32_t lo_pcg_hash(uint32_t v);
uint32_t lo_derive_seed(uint32_t globalSeed, uint32_t tag);
uint32_t lo_hash3(uint32_t x, uint32_t y, uint32_t tag, uint32_t seed);
double lo_rand01(uint32_t x, uint32_t y, uint32_t tag, uint32_t seed);
uint32_t lo_module_seed(uint32_t sceneSeed);
uint32_t lo_step_seed(uint32_t moduleSeed, uint32_t frameIndex, uint32_t substep);
/* SplitMix64 (Random.h:11-31); *state holds the raw state word */
void lo_splitmix_init(uint64_t *state, uint64_t seed);
uint64_t lo_splitmix_next(uint64_t *state);

/* ---- kernel builders ---- */
int lo_kernel_half(const lo_kernel *k);
double lo_kernel_raw(const lo_kernel *k, int dx, int dy);
/* (2*half+1)^2 window, centred, normalised to sum 1; returns 0 or -1 */
int lo_kernel_window(const lo_kernel *k, int half, double *out);
/* W*H wrap-centred image (origin at index 0), normalised to sum 1 */
void lo_kernel_image(const lo_kernel *k, int W, int H, double *out);

/* ---- init ---- */
#define LO_TAG_PATCH 1009u
void lo_init_patch(int W, int H, uint32_t seed, int x0, int y0, int pw, int ph, double *out);
void lo_init_value_noise(int W, int H, uint32_t seed, double initCoverage, double initScale,
                         double *out);

/* ---- potentials ---- */
void lo_potential_direct(int W, int H, const double *A, const double *win, int half, double *u);
void lo_fft(int N, double *re, double *im, int inverse); /* unscaled both ways */
void lo_fft2(int W, int H, double *re, double *im, int inverse);
void lo_potential_fft(int W, int H, const double *A, const double *img, double *u);

/* ---- growth and full step ---- */
void lo_growth(const lo_world *w, const double *A, const double *const *u, uint32_t stepSeed,
               uint32_t frameIndex, double *out);
/* one step; use_fft selects FFT vs direct convolution for every kernel */
int lo_step(const lo_world *w, const double *A, double *out, int use_fft, uint32_t stepSeed,
            uint32_t frameIndex);

/* ---- FP64 reference gradient of a T-step rollout (growth mode 0, no
 * noise/jitter/inject).  loss = sum(gT * A_T).  Outputs dL/dA0 (W*H),
 * and per kernel dL/dmu, dL/dsigma, dL/dweight, dL/dbetas[4]; dL/ddt. */
int lo_rollout(const lo_world *w, const double *A0, int T, int use_fft, double *AT);
int lo_rollout_grad(const lo_world *w, const double *A0, int T, const double *gT, double *gA0,
                    double *gMu, double *gSigma, double *gWeight, double *gBetas, double *gDt);
/* FFT-form rollout on one smooth piece of the clipped map (test helper): record != 0 stores
 * each step's clip state (0 inside, 1 low, 2 high) in mask[T*W*H]; record == 0 replays the
 * steps with those states imposed, so central differences never straddle a clip kink */
int lo_rollout_piece(const lo_world *w, const double *A0, int T, unsigned char *mask, int record,
                     double *AT);

/* ---- colour layer and activity ---- */
/* cosine palette of colorMapField / colorMapFieldScaled (Shaders/Render/ColorMap.metal:17-75):
 * field srcW x srcH -> rgba dstW x dstH x 4 (doubles); a[3] b[3] c[3] d[3] palette,
 * inputScale, inputBias, exposure, channel.  Sizes differ: bilinear at pixel centres, wrap. */
void lo_colormap(const double *field, int srcW, int srcH, const double *a, const double *b,
                 const double *c, const double *d, double inputScale, double inputBias,
                 double exposure, int channel, double *rgba, int dstW, int dstH);
/* organism stamps (leniaStampInit, Shaders/Field/Lenia.metal:242-303); the stamp geometry in
 * FP32 as the shader computes it */
void lo_init_stamp(int W, int H, const double *org, int ow, int oh, uint32_t count, uint32_t rotate,
                   uint32_t seed, double *out);

/* OpenMP thread count for the oracle's parallel loops (<=0: leave default) */
void lo_threads(int n);
int lo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
