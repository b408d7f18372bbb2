// Internal host-side structures behind the opaque C ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "../../include/lenia_mi355x.h"
#include "lenia_types.h"

struct lenia_ctx {
    int device = 0;
    int numCUs = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around one world / batch call
    hipEvent_t sp0 = nullptr, sp1 = nullptr;  // around a scene step (which makes world calls)
    hipEvent_t last0 = nullptr, last1 = nullptr;  // the pair lenia_ctx_last_gpu_ms reads
    bool timed = false;
    uint64_t liveBytes = 0;
    uint32_t liveAllocs = 0;
    std::vector<std::pair<void *, size_t>> allocSizes;
};

struct lenia_world {
    lenia_ctx *ctx = nullptr;
    lenia_world_desc desc{};
    bool isFft = false, isMfma = false;
    lenia::GrowthParams g{};
    uint32_t moduleSeed = 0;
    bool needsInit = true;
    bool specValid = false;   // F holds the spectrum of the current state
    uint64_t hostSteps = 0;   // steps launched since creation (direct parity)
    uint32_t frameNext = 0;   // frame index the next lenia_world_step starts at
    // FFT form
    lenia::FftPlan planW{}, planH{};
    int P = 1, C = 1;       // row pairs / columns per tile
    int rowT = 256, colT = 256;  // workgroup sizes
    bool rowKpar = false, colKpar = false;  // K inverse transforms side by side
    // wave executor per pass (else the workgroup-LDS one) and its waves per workgroup
    bool rowWave = false, colWave = false;
    int rowNW = 1, colNW = 1;
    // kernel-split wave passes (one wave per kernel transform, kMaxKernels
    // waves per workgroup); the row pass's forward-only and potential passes
    // then run the one-wave kernels on the same tiles
    bool rowSplit = false, colSplit = false;
    hipGraphExec_t graphExec = nullptr;     // captured step sequence
    bool graphExtras = false;
    uint32_t version = 0;  // bumped on every descriptor upload (batches re-read)
    uint32_t shapeVersion = 0;  // bumped when launch shapes change (captured graphs re-record)
    struct AdjWorkspace *adj = nullptr;  // rollout-gradient buffers, made on first use
    float *F = nullptr, *G = nullptr, *S = nullptr, *twW = nullptr, *twH = nullptr;
    // direct forms
    int half = 0;
    float *winF = nullptr;
    size_t winBytes = 0;
    uint16_t *toepHi = nullptr, *toepMid = nullptr, *toepLo = nullptr;  // MFMA forms: Toeplitz fragments
    size_t toepBytes = 0;
    int ksteps = 0;
    bool split() const { return desc.convMode == LENIA_CONV_DIRECT_BF16X3; }
    // both
    float *A0 = nullptr, *A1 = nullptr;
    float *scratchOut = nullptr;
    lenia::StepClock *clock = nullptr;
    float *org = nullptr;  // organism cells for stamp init
    int orgW = 0, orgH = 0;
    uint32_t stampCount = 6, stampRotate = 1;
    void *devDesc = nullptr;
    lenia::FftWorldDev fdev{};
    lenia::DirectWorldDev ddev{};
    lenia::MfmaWorldDev mdev{};
    int nRowTiles = 0, nColTiles = 0;  // tiles per pass (no table: step_tile)
    size_t ldsRows = 0, ldsCols = 0;

    uint32_t parity() const { return (uint32_t)(hostSteps & 1u); }
    bool extras() const {
        return desc.noiseAmount > 0.0f || desc.muJitter > 0.0f || desc.injectAmount > 0.0f;
    }
    float *current_state() const { return (isFft || parity() == 0) ? A0 : A1; }
    // complex values one row-pass step tile holds in LDS: (K) x nt x W
    int row_tile_cells() const { return (rowKpar ? (int)desc.nKernels : 1) * P * (int)desc.width; }
    // complex values one column-pass tile holds in registers: (K) x nt x H
    int col_tile_cells() const { return (colKpar ? (int)desc.nKernels : 1) * C * (int)desc.height; }
};

namespace lenia {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
int dev_alloc(lenia_ctx *ctx, void **p, size_t bytes);
void dev_free(lenia_ctx *ctx, void *p);

double kernel_raw(const lenia_kernel_desc &k, int dx, int dy);
int kernel_half(const lenia_kernel_desc &k);
bool kernel_window(const lenia_kernel_desc &k, int h, std::vector<double> &out);

bool factor235(int n, int &a, int &b, int &c);
bool plan_make(int n, const int *radix, int ns, FftPlan &p);
bool plan_default(int n, int maxR, FftPlan &p);
bool plan_unpack(int n, uint32_t packed, FftPlan &p);
uint32_t plan_pack(const FftPlan &p);
int compiled_plans(int n, uint32_t *packed, int maxOut);
// the plan cache's entry for (axis 0 rows / 1 columns, length n), 0 if none
uint32_t plan_cache_lookup(uint32_t axis, uint32_t n);
bool plan_fits(const FftPlan &p, int nt, int T, int maxv);
// LDS of a row-pass launch of T threads (with the twin transform region of
// rows_twin when it applies); of a column-pass launch.  kFftLdsMax
// (lenia_types.h) is the limit of FFT shapes.
size_t lds_rows(int W, int P, int K, bool kpar, int T);
size_t lds_cols(int H, int C, int K, bool kpar);
bool shape_ok(const FftPlan &p, bool rows, int nt, int K, int T, bool kpar);
bool pick_shape(const FftPlan &p, bool rows, int count, int K, uint32_t wantNt, uint32_t wantT,
                int &ntOut, int &TOut, bool &kparOut);
hipError_t enqueue_step(lenia_world *w, const hipEvent_t *ev, uint32_t off);
// a world's row pass (mode 0 forward, 1 step, 2 potential of kernel kSel into
// out) and column pass on its chosen executor; `rec` is the device record to
// run (the world's own, or a copy with the reverse pass's save pointers)
hipError_t world_launch_rows(lenia_world *w, int mode, const FftWorldDev *rec, uint32_t off,
                             int kSel, float *out, bool extras, hipStream_t s);
hipError_t world_launch_cols(lenia_world *w, const FftWorldDev *rec, hipStream_t s);
// default waves per workgroup of the wave executor for `count` transforms of
// L lanes: the largest of 4, 2, 1 that still gives every CU a workgroup
int wave_default_nw(int count, int L, int numCUs);
hipError_t world_advance_clock(lenia_world *w, uint32_t by);
hipError_t enqueue_steps(lenia_world *w, uint32_t n);
int world_ensure_spectrum(lenia_world *w);
int world_apply_init(lenia_world *w);

int world_upload_desc(lenia_world *w);
int world_check_desc(const lenia_world_desc *desc);
GrowthParams world_growth(const lenia_world_desc *desc, uint32_t moduleSeed);
int fft_sizes_and_tiles(lenia_world *w, int numCUs);
FftWorldDev fft_record(lenia_world *w);
// the allocation-free step's per-device scratch address (lenia_scratch.cpp),
// looked up once per device when a context is created on it
int scratch_init_device(int device);
int world_launch_steps(lenia_world *w, uint32_t n);
void world_drop_graph(lenia_world *w);
void world_free_adjoint(lenia_world *w);
void world_invalidate_adjoint(lenia_world *w);

// Mean duration in ms of each of the np kernels of a step over nSteps steps:
// `step(ev)` enqueues one step whose kernel p records ev[2p] and ev[2p+1]
// around itself (hipExtLaunchKernelGGL timestamps), returning a hipError_t.
template <class F>
hipError_t time_steps(lenia_ctx *ctx, uint32_t nSteps, uint32_t np, F step, double *out) {
    constexpr uint32_t kChunk = 64;  // steps between event read-backs
    std::vector<hipEvent_t> ev(2 * np * kChunk, nullptr);
    hipError_t e = hipSuccess;
    for (auto &x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    std::vector<double> sum(np, 0.0);
    for (uint32_t done = 0; done < nSteps && e == hipSuccess;) {
        const uint32_t m = std::min(kChunk, nSteps - done);
        for (uint32_t i = 0; i < m && e == hipSuccess; ++i) e = step(&ev[2 * np * i]);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        for (uint32_t i = 0; i < m && e == hipSuccess; ++i)
            for (uint32_t p = 0; p < np && e == hipSuccess; ++p) {
                float f = 0.f;
                e = hipEventElapsedTime(&f, ev[2 * np * i + 2 * p], ev[2 * np * i + 2 * p + 1]);
                sum[p] += f;
            }
        done += m;
    }
    for (auto x : ev)
        if (x) (void)hipEventDestroy(x);
    for (uint32_t p = 0; p < np; ++p) out[p] = sum[p] / std::max(1u, nSteps);
    return e;
}

uint32_t pcg_hash_host(uint32_t v);
uint32_t module_seed(uint32_t sceneSeed);

}  // namespace lenia
