/* hb_rng.h -- counter-based generator of the synthetic inputs.
 *
 * The ctr-th output of a SplitMix64 sequence (Steele, Lea & Flood, "Fast
 * splittable pseudorandom number generators", OOPSLA 2014) whose state is keyed
 * by (seed, stream).  Every connectivity, delay, weight and initial-state draw
 * is a pure function of (seed, stream, counter), so host and device produce the
 * same inputs in any order and at any scale.
 *   uniform01(h)     = (h >> 11) * 2^-53                  in [0, 1)
 *   uniform_int(h,n) = ((h >> 32) * n) >> 32              in [0, n)
 */
#ifndef HB_RNG_H
#define HB_RNG_H

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define HB_RNG_FN static inline __host__ __device__
#else
#define HB_RNG_FN static inline
#include <stdint.h>
#endif

HB_RNG_FN unsigned long long hb_mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

HB_RNG_FN unsigned long long hb_rand64(unsigned long long seed, unsigned long long stream, unsigned long long ctr) {
    unsigned long long key = hb_mix64(seed ^ hb_mix64(stream + 0x9e3779b97f4a7c15ULL));
    return hb_mix64(key + (ctr + 1ULL) * 0x9e3779b97f4a7c15ULL);
}

HB_RNG_FN double hb_uniform01(unsigned long long h) { return (double)(h >> 11) * 0x1p-53; }

HB_RNG_FN unsigned int hb_uniform_int(unsigned long long h, unsigned int n) {
    return (unsigned int)(((h >> 32) * (unsigned long long)n) >> 32);
}

#endif
