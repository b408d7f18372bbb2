// C wrapper around the reference's header-only LifeCore/Math/Random.h, built
// into oracle/_ref/ from the reference tree (test infrastructure only).  It
// lets the tests compare the oracle's integer hashes and SplitMix64 stream
// with the reference's own code, bit for bit.
#include "LifeCore/Math/Random.h"

#include <cstdint>

extern "C" {

uint32_t ref_pcg_hash(uint32_t v) { return life::pcgHash(v); }

uint32_t ref_derive_seed(uint32_t globalSeed, uint32_t tag) {
    return life::deriveSeed(globalSeed, tag);
}

// fills out[0..n) with the first n next() values of SplitMix64(seed)
void ref_splitmix_stream(uint64_t seed, int n, uint64_t *out) {
    life::SplitMix64 rng(seed);
    for (int i = 0; i < n; ++i) out[i] = rng.next();
}

// fills out[0..n) with the first n nextFloat01() values of SplitMix64(seed)
void ref_splitmix_float01(uint64_t seed, int n, float *out) {
    life::SplitMix64 rng(seed);
    for (int i = 0; i < n; ++i) out[i] = rng.nextFloat01();
}
}
