/*
 * Counter-based random numbers for the phonon loop: Philox4x32-10
 * (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11), plus the mapping from counter blocks to uniform doubles.
 *
 * The reference draws every random number from a thread_local std::mt19937
 * that is seeded from std::random_device (psim/include/psim/utils.h:16-21),
 * so no reference run can be repeated.  Here every draw is a pure function of
 * (seed, stream tag, phonon or source id, measurement step, draw index), which
 * makes a phonon's trajectory independent of thread, wave, launch and pool
 * order: the HIP kernels and the CPU oracle give the same numbers.
 *
 * Counter layout (c0, c1, c2, c3), key (k0, k1) = (seed low, seed high):
 *   c0 = block index j of the stream (each block yields two doubles)
 *   c1 = measurement step for PSIM_TAG_TRANSPORT, 0 otherwise
 *   c2 = phonon id (transport, emission) or source index (budget)
 *   c3 = stream tag | stream index << 8 (the index is 0 except for batched models)
 * Every draw takes the stream's next block: psim_urand returns its first double,
 * psim_urand2 both, for two numbers used together.  (A spare second double kept
 * per lane would leave the lanes of a wave out of step, so that nearly every
 * single draw would run a Philox block for the wave anyway; fixed pairs keep the
 * blocks a wave runs to one per draw site.)  A transport stream restarts at
 * block 0 whenever the phonon's measurement step advances.
 *
 * Plain C99 so the oracle (gcc) and the HIP kernels (hipcc) share it.
 */
#ifndef PSIM_RNG_H
#define PSIM_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PSIM_QUAL __host__ __device__ static inline __attribute__((always_inline))
#else
#define PSIM_QUAL static inline
#endif

#define PSIM_TAG_TRANSPORT 0u
#define PSIM_TAG_EMIT 1u
#define PSIM_TAG_BUDGET 2u

#define PSIM_PHILOX_M0 0xD2511F53u
#define PSIM_PHILOX_M1 0xCD9E8D57u
#define PSIM_PHILOX_W0 0x9E3779B9u
#define PSIM_PHILOX_W1 0xBB67AE85u

typedef struct {
    uint32_t v[4];
} psim_u32x4;

/* One Philox4x32 round: two 32x32->64 products, then the S-box-free mixing
 * of the published algorithm. */
PSIM_QUAL psim_u32x4 psim_philox_round(psim_u32x4 c, uint32_t k0, uint32_t k1) {
    const uint64_t p0 = (uint64_t)PSIM_PHILOX_M0 * (uint64_t)c.v[0];
    const uint64_t p1 = (uint64_t)PSIM_PHILOX_M1 * (uint64_t)c.v[2];
    psim_u32x4 r;
    r.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
    r.v[1] = (uint32_t)p1;
    r.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
    r.v[3] = (uint32_t)p0;
    return r;
}

/* Philox4x32 with 10 rounds; the key is bumped by the Weyl constants between
 * rounds. */
PSIM_QUAL psim_u32x4 psim_philox4x32_10(psim_u32x4 c, uint32_t k0, uint32_t k1) {
    for (int i = 0; i < 9; ++i) {
        c = psim_philox_round(c, k0, k1);
        k0 += PSIM_PHILOX_W0;
        k1 += PSIM_PHILOX_W1;
    }
    return psim_philox_round(c, k0, k1);
}

/* 53 high bits of (hi:lo) scaled into [0, 1). */
PSIM_QUAL double psim_u01_from_pair(uint32_t hi, uint32_t lo) {
    const uint64_t bits = ((uint64_t)hi << 32) | (uint64_t)lo;
    return (double)(bits >> 11) * 0x1.0p-53;
}

/* A stream of uniform doubles in [0, 1): one Philox block per draw. */
typedef struct {
    uint32_t k0, k1; /* key = seed */
    uint32_t c1, c2, c3; /* fixed counter words */
    uint32_t block; /* next block index */
} psim_stream;

PSIM_QUAL void psim_stream_init(psim_stream* s, uint64_t seed, uint32_t tag, uint32_t id, uint32_t step) {
    s->k0 = (uint32_t)seed;
    s->k1 = (uint32_t)(seed >> 32);
    s->c1 = step;
    s->c2 = id;
    s->c3 = tag;
    s->block = 0;
}

/* Counter word c3 of a stream: its tag plus a stream index (a model's index in a batch). */
PSIM_QUAL uint32_t psim_stream_tag(uint32_t tag, uint32_t stream) { return tag | (stream << 8); }

/* Restart the stream for a new measurement step (transport streams). */
PSIM_QUAL void psim_stream_set_step(psim_stream* s, uint32_t step) {
    s->c1 = step;
    s->block = 0;
}

PSIM_QUAL psim_u32x4 psim_stream_block(psim_stream* s) {
    psim_u32x4 c;
    c.v[0] = s->block++;
    c.v[1] = s->c1;
    c.v[2] = s->c2;
    c.v[3] = s->c3;
    return psim_philox4x32_10(c, s->k0, s->k1);
}

/* One uniform double: the first of the next block. */
PSIM_QUAL double psim_urand(psim_stream* s) {
    const psim_u32x4 r = psim_stream_block(s);
    return psim_u01_from_pair(r.v[0], r.v[1]);
}

/* Two uniform doubles from the next block, in order. */
PSIM_QUAL void psim_urand2(psim_stream* s, double* a, double* b) {
    const psim_u32x4 r = psim_stream_block(s);
    *a = psim_u01_from_pair(r.v[0], r.v[1]);
    *b = psim_u01_from_pair(r.v[2], r.v[3]);
}

#endif /* PSIM_RNG_H */
