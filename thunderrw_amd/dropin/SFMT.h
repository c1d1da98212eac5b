/* SFMT-19937 API names used by the reference's app functors (dependency/SFMT-src-1.5.1/SFMT.h:139-228),
 * implemented from the published algorithm (Saito & Matsumoto) for the drop-in surface.  The GPU
 * engine draws its own random numbers; these serve functor members such as PPR::sfmt_. */
#ifndef THUNDERRW_DROPIN_SFMT_H
#define THUNDERRW_DROPIN_SFMT_H

#include <stdint.h>
#include <string.h>

typedef struct {
    uint32_t state[624];
    int idx;
} sfmt_t;

static inline void sfmt_dropin_recursion(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t *c,
                                         const uint32_t *d) {
    static const uint32_t msk[4] = {0xdfffffefU, 0xddfecb7fU, 0xbffaffffU, 0xbffffff6U};
    uint64_t ahi = ((uint64_t)a[3] << 32) | a[2], alo = ((uint64_t)a[1] << 32) | a[0];
    uint64_t chi = ((uint64_t)c[3] << 32) | c[2], clo = ((uint64_t)c[1] << 32) | c[0];
    uint64_t xhi = (ahi << 8) | (alo >> 56), xlo = alo << 8;
    uint64_t ylo = (clo >> 8) | (chi << 56), yhi = chi >> 8;
    uint32_t x[4] = {(uint32_t)xlo, (uint32_t)(xlo >> 32), (uint32_t)xhi, (uint32_t)(xhi >> 32)};
    uint32_t y[4] = {(uint32_t)ylo, (uint32_t)(ylo >> 32), (uint32_t)yhi, (uint32_t)(yhi >> 32)};
    uint32_t t[4];
    for (int i = 0; i < 4; ++i) t[i] = a[i] ^ x[i] ^ ((b[i] >> 11) & msk[i]) ^ y[i] ^ (d[i] << 18);
    memcpy(r, t, sizeof(t));
}

static inline void sfmt_dropin_next_block(sfmt_t *s) {
    uint32_t *w = s->state;
    const uint32_t *r1 = w + 4 * 154, *r2 = w + 4 * 155;
    for (int i = 0; i < 156; ++i) {
        sfmt_dropin_recursion(w + 4 * i, w + 4 * i, w + 4 * ((i + 122) % 156), r1, r2);
        r1 = r2;
        r2 = w + 4 * i;
    }
}

static inline void sfmt_init_gen_rand(sfmt_t *s, uint32_t seed) {
    static const uint32_t parity[4] = {0x00000001U, 0x00000000U, 0x00000000U, 0x13c9e684U};
    s->state[0] = seed;
    for (uint32_t i = 1; i < 624; ++i) s->state[i] = 1812433253U * (s->state[i - 1] ^ (s->state[i - 1] >> 30)) + i;
    s->idx = 624;
    uint32_t inner = 0;
    for (int i = 0; i < 4; ++i) inner ^= s->state[i] & parity[i];
    for (int k = 16; k > 0; k >>= 1) inner ^= inner >> k;
    if (inner & 1U) return;
    for (int i = 0; i < 4; ++i)
        for (uint32_t bit = 1U; bit; bit <<= 1)
            if (bit & parity[i]) {
                s->state[i] ^= bit;
                return;
            }
}

static inline uint32_t sfmt_genrand_uint32(sfmt_t *s) {
    if (s->idx >= 624) {
        sfmt_dropin_next_block(s);
        s->idx = 0;
    }
    return s->state[s->idx++];
}

static inline double sfmt_to_real1(uint32_t v) { return v * (1.0 / 4294967295.0); }
static inline double sfmt_genrand_real1(sfmt_t *s) { return sfmt_to_real1(sfmt_genrand_uint32(s)); }
static inline double sfmt_to_real2(uint32_t v) { return v * (1.0 / 4294967296.0); }
static inline double sfmt_genrand_real2(sfmt_t *s) { return sfmt_to_real2(sfmt_genrand_uint32(s)); }

#endif
