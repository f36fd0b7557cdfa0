/*
 * mcpt_pcg32.h — PCG32 (O'Neill 2014, "PCG: A Family of Simple Fast
 * Space-Efficient Statistically Good Algorithms for Random Number
 * Generation"): 64-bit LCG state, multiplier 6364136223846793005, XSH-RR
 * output, seeded with pcg32_srandom_r's two-step procedure.
 *
 * Every camera path owns one stream keyed by (pixel, sample, seed):
 *   initstate = (sample << 32) ^ seed,   initseq = pixel.
 * The increment (2*pixel+1) selects the stream, so no state is carried from
 * one sample to the next and any sample range can be rendered, resumed or
 * split across devices exactly. The wavefront kernels store only `state`
 * per path; the increment is recomputed from the pixel index.
 *
 * Uniform floats are (x >> 8) * 2^-24, i.e. in [0, 1) with 24 random bits.
 */
#ifndef MCPT_PCG32_H
#define MCPT_PCG32_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MCPT_PCG_HD __host__ __device__ inline
#else
#define MCPT_PCG_HD static inline
#endif

#define MCPT_PCG_MULT 6364136223846793005ULL

typedef struct mcpt_pcg32 {
    uint64_t state;
    uint64_t inc;
} mcpt_pcg32;

MCPT_PCG_HD uint32_t mcpt_pcg32_next(mcpt_pcg32* rng) {
    uint64_t old = rng->state;
    rng->state = old * MCPT_PCG_MULT + rng->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

MCPT_PCG_HD void mcpt_pcg32_seed(mcpt_pcg32* rng, uint64_t initstate, uint64_t initseq) {
    rng->state = 0u;
    rng->inc = (initseq << 1u) | 1u;
    (void)mcpt_pcg32_next(rng);
    rng->state += initstate;
    (void)mcpt_pcg32_next(rng);
}

MCPT_PCG_HD void mcpt_pcg32_for_path(mcpt_pcg32* rng, uint32_t pixel, uint32_t sample, uint64_t seed) {
    mcpt_pcg32_seed(rng, ((uint64_t)sample << 32) ^ seed, (uint64_t)pixel);
}

/* SplitMix64's output function (Steele, Lea and Flood 2014), used to hash
 * words into a seed. */
MCPT_PCG_HD uint64_t mcpt_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* The stream a shadow ray's ratio tracking draws from, keyed by the bits
 * of the ray itself (origin, direction, length): a path's shadow ray is
 * the same bits whether it is traced inline or from a queue, so every
 * execution mode draws the same numbers without storing a key. */
MCPT_PCG_HD void mcpt_pcg32_for_shadow(mcpt_pcg32* rng, const uint32_t bits[7]) {
    uint64_t h = 0x6a09e667f3bcc909ULL;
    for (int k = 0; k < 7; k++) h = mcpt_mix64(h + 0x9e3779b97f4a7c15ULL + bits[k]);
    mcpt_pcg32_seed(rng, h, mcpt_mix64(h ^ 0x3c6ef372fe94f82bULL));
}

MCPT_PCG_HD float mcpt_pcg32_uniform(mcpt_pcg32* rng) {
    return (float)(mcpt_pcg32_next(rng) >> 8) * 5.9604644775390625e-08f;
}

#endif /* MCPT_PCG32_H */
