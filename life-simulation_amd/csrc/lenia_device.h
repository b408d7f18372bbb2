// Device-side helpers shared by every Lenia kernel: integer hashes, the
// per-cell growth rule and the step clock.  Compiled for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lenia_types.h"

namespace lenia {

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

__device__ __forceinline__ float bell_of(float u, float mu, float sigma) {
    float t = (u - mu) / sigma;
    return expf(-0.5f * t * t);
}

// one kernel's contribution, added in kernel order
__device__ __forceinline__ float growth_term(const GrowthParams &g, int k, float u, float muJ) {
    float b = bell_of(u, g.mu[k] + muJ, g.sigma[k]);
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

// Step seed of the step a kernel is running: pcg(seed_ ^ (frame*197 + sub))
// (Modules/FieldModules/Lenia.cpp:564)
__device__ __forceinline__ uint32_t clock_seed(const StepClock &c, uint32_t moduleSeed) {
    return pcg_hash(moduleSeed ^ (c.frame * 197u + c.sub));
}

// Advance every clock once all blocks of the launch are done.  Called by
// every block at the end of the last kernel of a step; the block that
// finishes last moves the clocks on and re-arms the counter.  The next
// kernel in stream order sees the new values.
__device__ __forceinline__ void clock_arrive(uint32_t *done, StepClock *clocks, int nClocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        uint32_t prev = atomicAdd(done, 1u);
        if (prev == gridDim.x - 1) {
            __threadfence();
            for (int i = 0; i < nClocks; ++i) {
                StepClock c = clocks[i];
                c.sub += 1u;
                if (c.sub >= c.substeps) {
                    c.sub = 0u;
                    c.frame += 1u;
                }
                c.parity ^= 1u;
                c.steps += 1u;
                clocks[i] = c;
            }
            __threadfence();
            atomicExch(done, 0u);
        }
    }
}

}  // namespace lenia
