/*
 * oracle_rng.c -- CPU oracle: counter-based RNGs.  TEST INFRASTRUCTURE ONLY (see mcint_oracle.h).
 *
 * Philox4x32-10 follows the published algorithm of Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11 (the Random123 family):
 * 10 rounds of two 32x32->64 multiplies with the Weyl key schedule.  The reference
 * has no Philox (its generator is SplitMix64), so this function is pinned by the
 * published known-answer vectors in tests/test_oracle_rng.py.
 *
 * SplitMix64 and u01 restate cuPythia/common/rng.cuh:47-57.
 */
#include "mcint_oracle.h"

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_fill(uint64_t key, uint64_t counter0, uint32_t stream, uint64_t n, uint32_t *out)
{
    uint32_t k[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t e = counter0 + i;
        uint32_t c[4] = { (uint32_t)e, (uint32_t)(e >> 32), stream, 0u };
        orc_philox4x32_10(c, k, out + 4 * i);
    }
}

double orc_u53(uint32_t w0, uint32_t w1)
{
    uint64_t x = ((uint64_t)w1 << 32) | (uint64_t)w0;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

float orc_u24(uint32_t w)
{
    return (float)(w >> 8) * (1.0f / 16777216.0f);
}

uint64_t orc_splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

double orc_u01(uint64_t x)
{
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}
