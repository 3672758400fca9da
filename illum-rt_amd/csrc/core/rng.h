// PCG32 generator with the state transition, output permutation, stream
// selection and float mapping of the reference's rng_pcg
// (libgi/random.h:59-96, itself PBRT-v3's PCG32).  Usable on host and device.
//
// The wavefront kernels derive one generator per (pixel, bounce, frame) from
// pcg_stream_id(); a ray's random numbers therefore depend only on which path
// and bounce it belongs to, never on the queue slot or lane that processes it,
// which is what lets the compacted/persistent bounce loop reproduce the plain
// per-bounce loop bit for bit.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RNG_HD __host__ __device__ __forceinline__
#else
#define RNG_HD inline
#endif

namespace rt {

struct pcg32 {
    uint64_t state, inc;

    static constexpr uint64_t kDefaultState = 0x853c49e6748fea9bULL;
    static constexpr uint64_t kMult = 0x5851f42d4c957f2dULL;

    RNG_HD explicit pcg32(uint64_t initseq) { set_sequence(initseq); }

    RNG_HD void set_sequence(uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_u32();
        state += kDefaultState;
        next_u32();
    }
    RNG_HD uint32_t next_u32() {
        uint64_t old = state;
        state = old * kMult + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    // min(1 - 2^-24, u * 2^-32), random.h:91-93.
    RNG_HD float next_float() {
        float f = (float)next_u32() * 0x1p-32f;
        const float one_minus_eps = 0x1.fffffep-1f;
        return f < one_minus_eps ? f : one_minus_eps;
    }
    // Unbiased integer in [0, b), random.h:84-90.
    RNG_HD uint32_t next_below(uint32_t b) {
        uint32_t threshold = (~b + 1u) % b;
        for (;;) {
            uint32_t r = next_u32();
            if (r >= threshold) return r % b;
        }
    }
};

// A 64-bit stream id for one (frame seed, path, bounce) triple.  The mixing is
// SplitMix64's finaliser; distinct inputs give distinct PCG sequences.  The
// path tracer's bounces use bounce = level (0..64); its sub-pixel camera
// offsets use kCameraStream.
constexpr uint64_t kCameraStream = 255;
RNG_HD uint64_t pcg_stream_id(uint64_t seed, uint64_t path, uint64_t bounce) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (path << 8) + bounce + 1;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

} // namespace rt
