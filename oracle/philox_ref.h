/*
 * ORACLE (test infrastructure only): Philox4x32-10 counter-based generator,
 * written from Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11 (the Random123 definition).  Pinned by the published
 * known-answer vectors in tests/test_oracle_rng.py.
 */
#ifndef ORC_PHILOX_REF_H
#define ORC_PHILOX_REF_H
#include <stdint.h>

static inline void orc_philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

static inline void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        orc_philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* Per-particle stream: key = seed, counter = (id, step, tag, block).  Draw k
 * is word k&3 of block k>>2.  Uniform = (w>>9 + 1/2) / 2^23, in (0,1). */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int used;
} orc_stream;

static inline void orc_stream_init(orc_stream *s, uint64_t seed, uint32_t id, uint32_t step, uint32_t tag) {
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32);
    s->ctr[0] = id; s->ctr[1] = step; s->ctr[2] = tag; s->ctr[3] = 0;
    s->used = 4;
}

static inline double orc_uniform(orc_stream *s) {
    if (s->used == 4) {
        orc_philox4x32_10(s->ctr, s->key, s->buf);
        s->ctr[3] += 1;
        s->used = 0;
    }
    uint32_t w = s->buf[s->used++];
    return ((double)(w >> 9) + 0.5) * (1.0 / 8388608.0);
}
#endif
