// Device-side helpers shared by every Lenia kernel: integer hashes, the
// per-cell growth rule and the step clock.  Compiled for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lenia_types.h"

namespace lenia {

// Pointers loaded from a world descriptor are generic (flat) to the compiler;
// flat accesses share a wait counter with LDS and make every LDS wait also
// wait for HBM.  Cast them to the global address space.
#define LENIA_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ LENIA_GLOBAL T *gp(T *p) {
    return (LENIA_GLOBAL T *)p;
}
// vector loads / stores through global pointers (HIP's float2/float4 cannot
// be copy-constructed from another address space; clang vectors can)
typedef float lv2 __attribute__((ext_vector_type(2)));
typedef float lv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld4(const LENIA_GLOBAL float4 *p) {
    lv4 v = *(const LENIA_GLOBAL lv4 *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst4(LENIA_GLOBAL float4 *p, float4 a) {
    *(LENIA_GLOBAL lv4 *)p = lv4{a.x, a.y, a.z, a.w};
}
__device__ __forceinline__ float2 gld2(const LENIA_GLOBAL float2 *p) {
    lv2 v = *(const LENIA_GLOBAL lv2 *)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void gst2(LENIA_GLOBAL float2 *p, float2 a) {
    *(LENIA_GLOBAL lv2 *)p = lv2{a.x, a.y};
}
// The clock `off` steps after the stored one (sub < substeps): what `off`
// single-step advances (sub + 1, wrapping into the next frame) give.
__device__ __forceinline__ StepClock clock_after(StepClock v, uint32_t off) {
    const uint32_t s = v.sub + off, n = v.substeps ? v.substeps : 1u;
    v.frame += s / n;
    v.sub = s % n;
    v.parity ^= off & 1u;
    v.steps += off;
    return v;
}

// The clock of the step `off` steps after a world's stored clock.  Steps
// launched together read the same stored clock with offsets 0, 1, 2, ... and
// one small kernel moves it on after the last of them, so no step kernel
// writes the clock that its own blocks read.
__device__ __forceinline__ StepClock clock_at(const StepClock *c, uint32_t off) {
    const LENIA_GLOBAL StepClock *g = gp(c);
    StepClock v;
    v.frame = g->frame;
    v.sub = g->sub;
    v.substeps = g->substeps;
    v.parity = g->parity;
    v.steps = g->steps;
    v.pad[0] = v.pad[1] = v.pad[2] = 0u;
    return clock_after(v, off);
}

// pcg_hash (Shaders/Common/Random.metal:5-9)
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// rand01 (Shaders/Common/Random.metal:16-19)
__device__ __forceinline__ float rand01(uint32_t x, uint32_t y, uint32_t tag, uint32_t seed) {
    uint32_t h = pcg_hash(x ^ pcg_hash(y ^ pcg_hash(tag ^ seed)));
    return float(h >> 8) * (1.0f / 16777216.0f);
}

// Metal round(): half away from zero
__device__ __forceinline__ float metal_round(float v) { return roundf(v); }

// Per-cell growth rule of leniaStep (Shaders/Field/Lenia.metal:85-127) and
// leniaGrowthMulti (Shaders/Field/Lenia.metal:172-231), split into an
// accumulate part (one call per kernel, in kernel order) and a finish part.
// With one kernel of weight 1 both reference kernels give the same value.
struct GrowthCell {
    float mu_j;  // muJitter offset of this cell
    float acc;   // mode 0: sum w(2b-1); mode 1: sum w b
};

__device__ __forceinline__ float cell_jitter(const GrowthParams &g, uint32_t x, uint32_t y,
                                             uint32_t seed, uint32_t frame) {
    if (g.muJitter > 0.0f) {
        float j = rand01(x, y, 401u + frame % 613u, seed) - 0.5f;
        return j * g.muJitter * 0.05f;
    }
    return 0.0f;
}

// bell = exp(-((u-mu)/sigma)^2 / 2) evaluated as exp2(-t*t) with
// t = (u-mu) * sqrt(log2(e)/2) / sigma (scale precomputed in FP64): one
// multiply and the native exp2 instead of an IEEE divide and expf.  Within
// 3 ulp of the FP64 bell.
__device__ __forceinline__ float bell_of(float u, float mu, float scale) {
    float t = (u - mu) * scale;
    return __builtin_amdgcn_exp2f(-(t * t));
}

// one kernel's contribution, added in kernel order
__device__ __forceinline__ float growth_term(const GrowthParams &g, int k, float u, float muJ) {
    float b = bell_of(u, g.mu[k] + muJ, g.bellScale[k]);
    if (g.growthMode == 1u) return g.weight[k] * b;
    return g.weight[k] * (2.0f * b - 1.0f);
}

__device__ __forceinline__ float growth_finish(const GrowthParams &g, float A, float acc,
                                               uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                                               uint32_t seed, uint32_t frame) {
    float A2;
    if (g.growthMode == 1u) {
        float target = (g.sumWeights > 1e-6f) ? (acc / g.sumWeights) : A;
        A2 = A + g.dt * (target - A);
    } else {
        A2 = A + g.dt * acc;
    }
    if (g.noiseAmount > 0.0f) {
        float n = rand01(x, y, 503u, seed) - 0.5f;
        A2 += n * g.noiseAmount * 0.03f;
    }
    if (g.injectAmount > 0.0f) {
        float uvx = (float(x) + 0.5f) / float(W);
        float uvy = (float(y) + 0.5f) / float(H);
        float aspect = float(W) / float(H);
        uint32_t cycle = frame / 12u;
        float blobR = float(g.radius) * 0.9f / float(H);
        for (uint32_t i = 0; i < g.injectCount; ++i) {
            float cx = rand01(i, cycle, 811u, seed);
            float cy = rand01(cycle, i, 823u, seed);
            float dx = uvx - cx, dy = uvy - cy;
            dx = dx - metal_round(dx);
            dy = dy - metal_round(dy);
            dx *= aspect;
            float r2 = (dx * dx + dy * dy) / (blobR * blobR);
            float bump = expf(-r2 * 0.5f);
            float t = fminf(fmaxf(bump * g.injectAmount, 0.0f), 1.0f);
            A2 = A2 + (0.6f - A2) * t;
        }
    }
    return fminf(fmaxf(A2, 0.0f), 1.0f);
}

// growth_finish without the noise / inject terms (identical bits when their
// amounts are zero)
__device__ __forceinline__ float growth_plain(const GrowthParams &g, float A, float acc) {
    float A2;
    if (g.growthMode == 1u) {
        float target = (g.sumWeights > 1e-6f) ? (acc / g.sumWeights) : A;
        A2 = A + g.dt * (target - A);
    } else {
        A2 = A + g.dt * acc;
    }
    return fminf(fmaxf(A2, 0.0f), 1.0f);
}

// The growth constants of a world copied into registers once per tile, so the
// per-cell loops neither reload them after each barrier nor loop over a
// runtime kernel count.  Same arithmetic as growth_term / growth_plain.
struct GrowthRegs {
    float mu[4], scale[4], weight[4];
    float dt, sumWeights;
    uint32_t mode;
    int n;
};

__device__ __forceinline__ GrowthRegs growth_regs(const GrowthParams &g) {
    GrowthRegs q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        q.mu[k] = g.mu[k];
        q.scale[k] = g.bellScale[k];
        q.weight[k] = g.weight[k];
    }
    q.dt = g.dt;
    q.sumWeights = g.sumWeights;
    q.mode = g.growthMode;
    q.n = (int)g.nKernels;
    return q;
}

__device__ __forceinline__ float growth_term_r(const GrowthRegs &q, int k, float u, float muJ) {
    float b = bell_of(u, q.mu[k] + muJ, q.scale[k]);
    if (q.mode == 1u) return q.weight[k] * b;
    return q.weight[k] * (2.0f * b - 1.0f);
}

__device__ __forceinline__ float growth_plain_r(const GrowthRegs &q, float A, float acc) {
    float A2;
    if (q.mode == 1u) {
        float target = (q.sumWeights > 1e-6f) ? (acc / q.sumWeights) : A;
        A2 = A + q.dt * (target - A);
    } else {
        A2 = A + q.dt * acc;
    }
    return fminf(fmaxf(A2, 0.0f), 1.0f);
}

// Step seed of the step a kernel is running: pcg(seed_ ^ (frame*197 + sub))
// (Modules/FieldModules/Lenia.cpp:564)
__device__ __forceinline__ uint32_t clock_seed(const StepClock &c, uint32_t moduleSeed) {
    return pcg_hash(moduleSeed ^ (c.frame * 197u + c.sub));
}

// Move every world's stored clock `by` steps on: one thread per world, run by
// a one-block kernel after the last step kernel of a launch sequence.
template <class WorldT>
__device__ __forceinline__ void advance_clocks(const WorldT *worlds, int nWorlds, uint32_t by) {
    for (int i = threadIdx.x; i < nWorlds; i += blockDim.x) {
        LENIA_GLOBAL StepClock *cp = gp(worlds[i].clock);
        StepClock v = clock_at(worlds[i].clock, by);
        cp->frame = v.frame;
        cp->sub = v.sub;
        cp->parity = v.parity;
        cp->steps = v.steps;
    }
}

// The tile a step workgroup works on: entry j of the launch's tile table, or,
// for a launch without a table (one world, whose tiles are 0, 1, ... in order),
// tile j of world 0.  Without the table the tile is known without a memory
// round trip before the descriptor loads.
__device__ __forceinline__ int2 step_tile(const int2 *__restrict__ tiles, int j) {
    return tiles ? tiles[j] : make_int2(0, j);
}

}  // namespace lenia
