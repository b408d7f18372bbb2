// Internal host-side structures behind the opaque C ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/lenia_mi355x.h"
#include "lenia_types.h"

struct lenia_ctx {
    int device = 0;
    int numCUs = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
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
    hipGraphExec_t graphExec = nullptr;     // captured step sequence
    bool graphExtras = false;
    uint32_t version = 0;  // bumped on every descriptor upload (batches re-read)
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
    uint32_t *done = nullptr;
    void *devDesc = nullptr;
    lenia::FftWorldDev fdev{};
    lenia::DirectWorldDev ddev{};
    lenia::MfmaWorldDev mdev{};
    int2 *rowTiles = nullptr, *colTiles = nullptr;
    int nRowTiles = 0, nColTiles = 0;
    size_t ldsRows = 0, ldsCols = 0;

    uint32_t parity() const { return (uint32_t)(hostSteps & 1u); }
    bool extras() const {
        return desc.noiseAmount > 0.0f || desc.muJitter > 0.0f || desc.injectAmount > 0.0f;
    }
    float *current_state() const { return (isFft || parity() == 0) ? A0 : A1; }
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
bool plan_fits(const FftPlan &p, int nt, int T, int maxv);
size_t lds_rows(int W, int P, int K, bool kpar);
size_t lds_cols(int H, int C, int K, bool kpar);
bool shape_ok(const FftPlan &p, bool rows, int nt, int K, int T, bool kpar);
bool pick_shape(const FftPlan &p, bool rows, int count, int K, uint32_t wantNt, uint32_t wantT,
                int &ntOut, int &TOut, bool &kparOut);
hipError_t enqueue_step(lenia_world *w, const hipEvent_t *ev);
int world_ensure_spectrum(lenia_world *w);
int world_apply_init(lenia_world *w);

int world_upload_desc(lenia_world *w);
int world_launch_steps(lenia_world *w, uint32_t n);
void world_drop_graph(lenia_world *w);
void world_free_adjoint(lenia_world *w);

uint32_t pcg_hash_host(uint32_t v);
uint32_t module_seed(uint32_t sceneSeed);

}  // namespace lenia
