/*
 * lenia_mi355x.h -- C ABI of the MI355X-native Lenia field update.
 *
 * The library replaces the GPU side of LifeSimSuite's Lenia module: the
 * module lifecycle (Modules/FieldModules/Lenia.{h,cpp}), its compute passes
 * (Shaders/Field/Lenia.metal, Shaders/Common/FFT.metal) and the allocator and
 * timing layer they sit on (LifeCore/Metal/ResourcePool.mm, GPUTimer.mm).
 * Each entry point names the reference interface it stands in for.
 *
 * Conventions: C linkage, plain pointers and sizes, no exceptions.  Every
 * function returns LENIA_OK (0) or a negative LENIA_ERR_* code and leaves a
 * message for lenia_last_error() (thread-local).  A context owns one HIP
 * stream on one device; all work of its worlds and batches is ordered on that
 * stream.  Device memory is owned by the library; host buffers by the caller.
 */
#ifndef LENIA_MI355X_H
#define LENIA_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LENIA_OK 0
#define LENIA_ERR_INVALID_ARG -1 /* bad descriptor, size or pointer */
#define LENIA_ERR_UNSUPPORTED -2 /* valid request this build cannot serve */
#define LENIA_ERR_DEVICE -3      /* a HIP call failed */
#define LENIA_ERR_OUT_OF_MEMORY -4
#define LENIA_ERR_STATE -5       /* call not valid in the object's state */

/* convolution forms ("convMode" of the scene JSON, Lenia.cpp:213-214, plus
 * the MFMA forms this library adds) */
#define LENIA_CONV_DIRECT 0        /* FP32 vector direct convolution */
#define LENIA_CONV_FFT 1           /* FP32 FFT convolution */
#define LENIA_CONV_DIRECT_BF16 2   /* BF16 MFMA banded contraction, FP32 accumulate */
#define LENIA_CONV_DIRECT_BF16X3 3 /* split-BF16 (hi+lo) compensated MFMA contraction */

/* kernel shapes: which reference builder a kernel follows */
#define LENIA_SHAPE_SHELL 0 /* buildKernelTexture (Lenia.cpp:410-436): one shell at shellMu */
#define LENIA_SHAPE_RING 1  /* buildRingKernelImage (Lenia.cpp:95-125): betas rings, centre 0.5 */

/* kernel cores */
#define LENIA_CORE_GAUSSIAN 0 /* exp(-((r-c)/s)^2/2), the reference's core */
#define LENIA_CORE_EXP 1      /* exp(4 - 1/(r(1-r))) exponential bump */

/* init modes */
#define LENIA_INIT_PATCH 0 /* seeded uniform[0,1) noise in a rectangle, 0 elsewhere */
#define LENIA_INIT_NOISE 1 /* leniaInit value-noise soup (Lenia.metal:43-61) */
#define LENIA_INIT_NONE 2  /* keep the current state (set with lenia_world_write_state) */
#define LENIA_INIT_STAMP 3 /* leniaStampInit organism stamps (Lenia.metal:242-303); needs
                              lenia_world_set_organism */

/* one entry of the scene "kernels" array (Lenia.h:99-105) with the radius
 * already resolved (radius * radiusScale) */
typedef struct lenia_kernel_desc {
    uint32_t shape;   /* LENIA_SHAPE_* */
    uint32_t core;    /* LENIA_CORE_* */
    float radius;     /* R_k in cells */
    float shellMu;    /* kernelShellMu (SHELL shape) */
    float shellSigma; /* kernelShellSigma */
    uint32_t nBetas;  /* 1..4 */
    float betas[4];   /* shell peak heights (RING shape) */
    float mu;         /* growth centre */
    float sigma;      /* growth width */
    float weight;     /* mixing weight h_k */
} lenia_kernel_desc;

/* one world: the Lenia scene-JSON surface resolved to numbers
 * (LeniaParams Lenia.h:77-94, LeniaMultiParams Lenia.h:111-129) */
typedef struct lenia_world_desc {
    uint32_t width, height; /* simWidth / simHeight */
    uint32_t convMode;      /* LENIA_CONV_* */
    uint32_t growthMode;    /* 0: A + dt*sum w(2b-1); 1: asymptotic */
    float dt;
    float noiseAmount;
    float muJitter;
    float injectAmount;
    uint32_t injectCount;
    uint32_t radius; /* module radius; sizes the inject blobs */
    uint32_t nKernels; /* 1..4 */
    lenia_kernel_desc kernels[4];
    /* init applied by the first frame after a reset (needsInit_) */
    uint32_t initMode; /* LENIA_INIT_* */
    uint32_t initX0, initY0, initW, initH; /* patch rectangle; initW or initH 0 => whole field */
    float initCoverage, initScale;         /* value-noise init */
    /* FFT launch overrides (0 = library choice).  Plans: radices packed 4
     * bits each, first stage in the lowest nibble, 0xF encodes 16 and 0xE
     * encodes 32.  Tiles: row pairs per row-pass workgroup, columns per
     * column-pass workgroup, and the two workgroup sizes (64, 128, 256 or
     * 512; the row pass also 1024 when its kernel transforms run side by
     * side and K x rowPairsPerTile x width <= 12288).  Side-by-side row
     * transforms also need K x rowPairsPerTile <= rowThreads; a shape that
     * fits no variant is rejected with LENIA_ERR_UNSUPPORTED. */
    uint32_t fftPlanW, fftPlanH;
    uint32_t rowPairsPerTile, colsPerTile;
    uint32_t rowThreads, colThreads;
    /* Transform executor of each FFT pass (LENIA_ENGINE_*; 0 = library
     * choice).  Both executors run the same radix plan with the same
     * butterflies, so the choice never changes a bit.  With the wave
     * executor the tile fields above are derived: rowThreads / colThreads
     * (64, 128 or 256) give the waves per workgroup, and a tile holds
     * 64 / L row pairs or columns per wave (L lanes per transform, set by
     * the plan); rowPairsPerTile / colsPerTile must then be 0. */
    uint32_t rowEngine, colEngine;
} lenia_world_desc;

#define LENIA_ENGINE_AUTO 0 /* no LDS shape given: kernel-split wave columns when the plan has a
                               wave kernel, LDS rows */
#define LENIA_ENGINE_LDS 1  /* a workgroup's threads share each stage through LDS barriers */
#define LENIA_ENGINE_WAVE 2 /* one wave owns whole transforms: registers + wave-private LDS */
#define LENIA_ENGINE_WAVE_SPLIT 3 /* the wave executor with one wave per kernel transform (4 waves,
                                     256 threads, per workgroup of 64 / L row pairs or columns) */

typedef struct lenia_ctx lenia_ctx;
typedef struct lenia_world lenia_world;
typedef struct lenia_batch lenia_batch;

/* message of the last failing call on this thread ("" if none) */
const char *lenia_last_error(void);
/* library version string */
const char *lenia_version(void);

/* ---- context: replaces MetalContext::create + ResourcePool + CommandGraph
 * construction (LifeCore/Sim/SceneRunner.cpp:31-35, MetalContext.mm:91-108) */
int lenia_ctx_create(int device, lenia_ctx **out);
int lenia_ctx_destroy(lenia_ctx *ctx);
/* waits for all queued work; reports a device error raised by it
 * (CommandGraph::endFrame(wait) + lastFrameHadError, CommandGraph.mm:69-90) */
int lenia_ctx_sync(lenia_ctx *ctx);
/* the context's hipStream_t, for callers that order their own work with it */
int lenia_ctx_stream(lenia_ctx *ctx, void **stream);
/* live device bytes and allocations (ResourcePool accounting, ResourcePool.h:78-81) */
int lenia_ctx_memory(lenia_ctx *ctx, uint64_t *liveBytes, uint32_t *liveAllocs);
/* GPU time in ms of the last timed call -- lenia_world_frame / _step,
 * lenia_batch_step / _frames, or lenia_scene_step (all its frames: world steps
 * and colour layers) -- (GPUTimer, LifeCore/Metal/GPUTimer.mm:9-94); waits for it */
int lenia_ctx_last_gpu_ms(lenia_ctx *ctx, double *ms);

/* ---- world: replaces LeniaModule::setup (Lenia.cpp:208-408): allocates the
 * field buffers, builds the kernels and, for the FFT form, their spectra */
int lenia_world_create(lenia_ctx *ctx, const lenia_world_desc *desc, lenia_world **out);
int lenia_world_destroy(lenia_world *w);
/* LeniaModule::reset (Lenia.cpp:448-451): seed_ = deriveSeed(sceneSeed,'LEN1');
 * the next frame re-initialises the field */
int lenia_world_reset(lenia_world *w, uint32_t sceneSeed);
/* LeniaModule::encode (Lenia.cpp:455-653): init if needed, then `substeps`
 * steps of frame `frameIndex` (step s uses seed pcg(seed_ ^ (frame*197+s))) */
int lenia_world_frame(lenia_world *w, uint32_t frameIndex, uint32_t substeps);
/* continues the frame clock: n more frames of one substep each (the
 * benchmark loop, Apps/LifeBench/main.cpp:112-133) */
int lenia_world_step(lenia_world *w, uint32_t nSteps);
/* updates the growth parameters (mu, sigma, weight, dt, growthMode, noise,
 * jitter, inject) and, if the kernel shapes changed, the kernels and spectra.
 * Sizes and convMode must be unchanged.  (The per-frame ParameterBus reads of
 * Lenia.cpp:456-464 and the init-time kernel rebuild of :497-545.) */
int lenia_world_set_params(lenia_world *w, const lenia_world_desc *desc);
/* field readback / upload, width*height floats row-major
 * (FieldModule::outputField, SimulationModule.h:132-135; ResourcePool::uploadTexture,
 * ResourcePool.mm:175-191).  Uploading clears the pending init. */
int lenia_world_read_state(lenia_world *w, float *host, size_t count);
int lenia_world_write_state(lenia_world *w, const float *host, size_t count);
/* field snapshot as a .npy file: float32 little endian, C order, shape
 * (height, width) -- the NumPy format of the reference's state dumps
 * (LifeCore/IO/NpyWriter.cpp:42-171).  Loading checks dtype, order and shape
 * and, like lenia_world_write_state, clears the pending init. */
int lenia_world_save_npy(lenia_world *w, const char *path);
int lenia_world_load_npy(lenia_world *w, const char *path);
/* device pointer of the current state buffer (valid until the next step) */
int lenia_world_state_device(lenia_world *w, float **dev);
/* potential u_k = K_k (x) A of the current state (the potential_k textures of
 * Lenia.cpp:611-616), width*height floats */
int lenia_world_potential(lenia_world *w, uint32_t k, float *host, size_t count);
/* runs nSteps steps with GPU timestamps around each kernel and writes the mean
 * duration in ms of each pass (FFT form: [column pass, row pass]; direct
 * forms: [step]) -- the per-pass GPUTimer figures LifeBench reports
 * (LifeCore/Metal/GPUTimer.mm:28-94, Apps/LifeBench/main.cpp:112-133) */
int lenia_world_pass_times(lenia_world *w, uint32_t nSteps, double *outMs, uint32_t maxPasses,
                           uint32_t *nPasses);
/* organism cells for LENIA_INIT_STAMP (row-major orgW x orgH, values in
 * [0,1]; the organism texture of Lenia.cpp:358-371) with the stamp count
 * (capped at 32) and the rotate flag (stampCount / stampRotate,
 * Lenia.cpp:283-284); copied to the device, used by every later init */
int lenia_world_set_organism(lenia_world *w, const float *cells, uint32_t orgW, uint32_t orgH,
                             uint32_t stampCount, uint32_t stampRotate);
/* resolved descriptor (tile shapes and FFT plans filled in) */
int lenia_world_get_desc(lenia_world *w, lenia_world_desc *out);
/* empirical launch-shape search for the FFT form (no counterpart in the
 * reference; the PresetLab/LifeBench callers get tuned tiles): every
 * transforms-per-workgroup, workgroup-size and side-by-side choice the plan
 * allows is timed over nSteps steps per pass, the fastest kept.  Plans are not
 * changed, so results are bit-identical before and after; the state and step
 * clock are restored.  *msPerStep (may be NULL): the tuned kernel time of one
 * step.  Direct forms: nothing to tune, *msPerStep = 0. */
int lenia_world_autotune(lenia_world *w, uint32_t nSteps, double *msPerStep);

/* ---- radix plans (no counterpart in the reference, whose FFT is radix 2
 * only; for callers that want tuned plans shared by single worlds and
 * batches).  The plan of a transform length decides the arithmetic of every
 * FFT-form step, so it belongs to the (axis, length) pair, never to a launch
 * shape or a batch: a world created with fftPlanW / fftPlanH = 0 takes the
 * plan cached for its width (axis 0, row transforms) and height (axis 1,
 * column transforms), else the built-in default, and keeps it for life.
 * Worlds created under the same cache therefore give the same bits alone and
 * batched.  The cache is process-wide and thread-safe. */
/* times every compiled radix plan of desc's width (row pass) and then of its
 * height (column pass) on the FFT form of desc, launch shapes tuned per plan,
 * nSteps steps each; records the fastest pair in the cache and returns it
 * (packed as fftPlanW / fftPlanH; any output may be NULL) with its kernel ms
 * per step.  A column plan that runs on the wave executor is preferred when
 * it is within 10% of the fastest (batches run such worlds' column passes in
 * their wave class launches).  Worlds that already exist keep their plans. */
int lenia_plan_autotune(lenia_ctx *ctx, const lenia_world_desc *desc, uint32_t nSteps,
                        uint32_t *planW, uint32_t *planH, double *msPerStep);
/* one entry: axis 0 rows, 1 columns; plan 0 removes it; a plan that is not a
 * radix sequence of length n is rejected */
int lenia_plan_cache_set(uint32_t axis, uint32_t n, uint32_t plan);
int lenia_plan_cache_get(uint32_t axis, uint32_t n, uint32_t *plan); /* 0: no entry */
void lenia_plan_cache_clear(void);
/* the cache as a text file, one "rows|cols <length> <plan hex>" line per entry
 * ('#' starts a comment line); loading is all or nothing and adds to the cache */
int lenia_plan_cache_save(const char *path);
int lenia_plan_cache_load(const char *path);

/* ---- allocation-free stepping (no counterpart in the reference): for callers
 * that cannot allocate device memory -- code under stream capture, serving
 * loops, harnesses that forbid hipMalloc.  Steps the caller's device field
 * `state` (width*height floats, row-major, updated in place) nSteps times
 * with the FFT form of `desc` on ctx's stream, step i at frame
 * frameIndex + i, substep 0, with module seed deriveSeed(sceneSeed, 'LEN1'):
 * the results equal lenia_world_step's on a world of the same descriptor,
 * reset with sceneSeed, holding the same state at that frame.  Spectra,
 * twiddles and the device record live in a library-owned device
 * array of 64 MiB per GPU (taken when the first context on a GPU is created,
 * held until the process ends); no call allocates.  Fields whose buffers do
 * not fit it (a 1920x1080 field with 3 kernels fits) are rejected with
 * LENIA_ERR_INVALID_ARG.  Every call on a GPU shares that array: calls that
 * overlap on one GPU (two streams or threads) race.  Asynchronous except
 * when the descriptor differs from the previous call's on the GPU (the
 * kernel spectra are then rebuilt, and the call waits for that). */
int lenia_fft_step_scratch(lenia_ctx *ctx, const lenia_world_desc *desc, uint32_t sceneSeed,
                           uint32_t frameIndex, float *state, uint32_t nSteps);

/* ---- differentiable rollout (no counterpart in the reference; for callers
 * that fit parameters and initial states to target fields).  FFT form,
 * growthMode 0, no noise / jitter / inject.  Pointers are device pointers
 * when onDevice is non-zero, else host pointers; width*height floats each. */
typedef struct lenia_param_grads {
    double dMu[4], dSigma[4], dWeight[4]; /* per kernel growth centre, width, mix weight */
    double dBetas[4][4];                  /* per kernel shell peaks (RING shape) */
    double dDt;
} lenia_param_grads;

/* T steps from A0; the world's state ends at AT.  The states at t = 0, c, 2c,
 * ... (c = ckptEvery, 0: floor(sqrt(T))) are kept for a following
 * lenia_rollout_backward. */
int lenia_rollout_forward(lenia_world *w, uint32_t T, uint32_t ckptEvery, const float *A0, float *AT,
                          int onDevice);
/* reverse pass of the rollout from A0 for L with dL/dA_T = gAT: writes dL/dA0
 * (may be NULL) and the parameter gradients.  The states of each segment of c
 * steps are recomputed from its checkpoint.  A0 NULL reuses the checkpoints of
 * the last lenia_rollout_forward of this world (same T and c, parameters
 * unchanged since; else LENIA_ERR_STATE); otherwise the forward sweep runs first. */
int lenia_rollout_backward(lenia_world *w, uint32_t T, uint32_t ckptEvery, const float *A0,
                           const float *gAT, float *gA0, lenia_param_grads *grads, int onDevice);

/* ---- batch: many worlds stepped together (replaces the one-variant-at-a-time
 * loop of Apps/LifePresetLab/main.cpp:718-806).  Worlds keep their own
 * buffers; a batch holds a world table and launch tiles over all of them.  All
 * worlds of a batch must use the same convMode.  Results are bit-identical to
 * stepping each world alone. */
int lenia_batch_create(lenia_ctx *ctx, lenia_world *const *worlds, uint32_t n, lenia_batch **out);
int lenia_batch_destroy(lenia_batch *b);
/* nSteps steps of every world; graphSteps > 0 captures that many steps in a
 * hipGraph once and replays it (nSteps must then be a multiple of graphSteps) */
int lenia_batch_step(lenia_batch *b, uint32_t nSteps, uint32_t graphSteps);
/* the batch's launch shapes (one workgroup size and side-by-side choice per
 * pass for all worlds) picked by timing every allowed choice over nSteps steps,
 * then the transform executors, whose finalists are timed as graph-replayed
 * steps; states and clocks restored, results unchanged bit for bit.
 * *msPerStep: the chosen configuration's graph-replayed time per step */
int lenia_batch_autotune(lenia_batch *b, uint32_t nSteps, double *msPerStep);
/* FFT-form launch grids (no counterpart in the reference).  Withdrawn: the
 * resident tile-queue grids ran 3x slower than one workgroup per tile on the
 * 256-world sweep.  set_grids accepts 0, 0 (one workgroup per tile, the only
 * launch form) and refuses anything else with LENIA_ERR_UNSUPPORTED;
 * get_grids reports 0, 0. */
int lenia_batch_set_grids(lenia_batch *b, uint32_t colPerCU, uint32_t rowPerCU);
int lenia_batch_get_grids(const lenia_batch *b, uint32_t *colPerCU, uint32_t *rowPerCU);
/* FFT-form transform executor of the batch (no counterpart in the reference):
 * with n > 0 waves per workgroup (1, 2 or 4) the worlds whose radix plan has
 * a wave kernel run the pass on the wave executor (one launch per values-per-
 * lane class, the launches of a pass side by side), the others on the
 * workgroup-LDS one; LENIA_WAVES_PER_KERNEL (3) runs the kernel-split wave
 * executor (LENIA_ENGINE_WAVE_SPLIT); 0 runs every world on the LDS
 * executor.  Default 0 rows, 4 columns; lenia_batch_autotune times 0, 1, 2,
 * 4 and 3.
 * Every executor computes the same bits. */
#define LENIA_WAVES_PER_KERNEL 3
int lenia_batch_set_waves(lenia_batch *b, uint32_t rowWaves, uint32_t colWaves);
int lenia_batch_get_waves(const lenia_batch *b, uint32_t *rowWaves, uint32_t *colWaves);
/* nFrames consecutive frames frame0, frame0+1, ... of every world, substeps
 * steps each, with the frame / substep clock of LeniaModule::encode
 * (Lenia.cpp:562-564) -- what one SceneRunner::step per frame does to each
 * variant (SceneRunner.cpp:146-246).  Pending inits run in the first frame. */
int lenia_batch_frames(lenia_batch *b, uint32_t frame0, uint32_t nFrames, uint32_t substeps,
                       uint32_t graphSteps);

/* ---- colour map: field -> RGBA16F layer (colorMapField / colorMapFieldScaled,
 * Shaders/Render/ColorMap.metal:17-75; ColorMapParams, LifeCore/Render/ColorMapPass.h:16-25).
 * rgb = max(a + b cos(2 pi (c t + d)), 0) * exposure, alpha = t, with
 * t = clamp(v * inputScale + inputBias, 0, 1). */
typedef struct lenia_colormap {
    float a[3], b[3], c[3], d[3];
    float inputScale, inputBias, exposure;
    uint32_t channel; /* 0: the field value; 1, 2: zero; 3: one (an R32F texel is (r, 0, 0, 1)) */
} lenia_colormap;

/* the palette a Lenia module starts from (ColorMapPass.h:16-25 defaults with
 * Lenia.cpp:350-354's d = (0.60, 0.45, 0.30), channel 0, inputScale 1) */
void lenia_colormap_lenia_defaults(lenia_colormap *p);
/* field (device, srcW x srcH floats) -> rgba (device, dstW x dstH pixels of 4
 * half floats), ordered on ctx's stream.  Same size: per-texel read; other
 * sizes: bilinear at pixel centres with wrap-around addressing. */
int lenia_colormap_field(lenia_ctx *ctx, const float *field, uint32_t srcW, uint32_t srcH,
                         const lenia_colormap *p, uint16_t *rgba, uint32_t dstW, uint32_t dstH);

/* ---- scene runner for Lenia scenes: the scene JSON (LifeCore/Params/Scene.cpp:24-113)
 * through LeniaModule::setup / encode (Lenia.cpp:208-653) and the per-frame
 * visibility, reseed and freeze rules of SceneRunner::step
 * (LifeCore/Sim/SceneRunner.cpp:126-215).  Every enabled module must be of
 * type "Lenia".  Compositing, post effects, audio analysis and output sinks
 * are not part of this library: a module's output is its colour layer, and
 * the values the reference's ParameterBus would hand a module each frame are
 * set with lenia_scene_set_param. */
typedef struct lenia_scene lenia_scene;
typedef struct lenia_scene_info {
    uint32_t width, height; /* "resolution": size of every module's colour layer */
    uint32_t seed;          /* scene seed */
    uint32_t substeps;
    uint32_t frameIndex;    /* frames stepped since the last reset */
    uint32_t nModules;
} lenia_scene_info;

/* what LeniaModule::setup resolves for enabled module `module` of a scene
 * (descriptor, palette, instance name), without touching a GPU.  Set-up
 * warnings of the reference (fallbacks, truncations) go to stderr as there. */
int lenia_scene_describe(const char *json, size_t len, uint32_t module, lenia_world_desc *desc,
                         lenia_colormap *cmap, char *name, size_t nameCap);
int lenia_scene_create(lenia_ctx *ctx, const char *json, size_t len, uint32_t substeps,
                       lenia_scene **out);
int lenia_scene_destroy(lenia_scene *s);
/* SceneRunner::reset: frame 0, every module reseeded from the scene seed */
int lenia_scene_reset(lenia_scene *s);
/* nFrames calls of SceneRunner::step (each: substeps steps of every visible,
 * unfrozen module, then its colour layer) */
int lenia_scene_step(lenia_scene *s, uint32_t nFrames);
/* the per-frame value of "module.key" (ParameterBus::value, base + mappings);
 * read from the next frame on */
int lenia_scene_set_param(lenia_scene *s, const char *target, float value);
int lenia_scene_get_info(lenia_scene *s, lenia_scene_info *out);
/* module's world (owned by the scene) */
int lenia_scene_world(lenia_scene *s, uint32_t module, lenia_world **world);
/* module's colour layer, width*height*4 half floats (host or device copy) */
int lenia_scene_layer(lenia_scene *s, uint32_t module, uint16_t *rgba, size_t count, int onDevice);

/* ---- parameter sweep on the batched launch (the variant loop of
 * Apps/LifePresetLab/main.cpp:156-324, 718-806): sweep JSON ("mode"
 * cartesian | per-axis, "seeds", "axes" of {target, values | range + steps})
 * expanded into variants of a one-module Lenia scene, all stepped as one
 * batch; activity (mean / stddev of Rec.709 luma of the colour layer, alive
 * when 0.01 <= mean <= 0.7 and stddev >= 0.02, main.cpp:67-90) on the GPU. */
typedef struct lenia_sweep lenia_sweep;
int lenia_sweep_create(lenia_ctx *ctx, const char *sceneJson, size_t sceneLen, const char *sweepJson,
                       size_t sweepLen, uint32_t substeps, lenia_sweep **out);
int lenia_sweep_destroy(lenia_sweep *s);
int lenia_sweep_count(lenia_sweep *s, uint32_t *nVariants);
/* frames 0 .. nFrames-1 of every variant (graphSteps as lenia_batch_step) */
int lenia_sweep_run(lenia_sweep *s, uint32_t nFrames, uint32_t graphSteps);
/* per-variant activity of the last frame run (arrays of n = count entries) */
int lenia_sweep_activity(lenia_sweep *s, double *meanLuma, double *stddevLuma, int *alive,
                         uint32_t n);
/* variant v's world (owned by the sweep) */
int lenia_sweep_world(lenia_sweep *s, uint32_t v, lenia_world **world);
/* variants.json text (id, seed, overrides, activity per variant); *needed is
 * the size including the terminating 0, buf may be NULL to query it */
int lenia_sweep_report(lenia_sweep *s, char *buf, size_t cap, size_t *needed);

#ifdef __cplusplus
}
#endif
#endif
