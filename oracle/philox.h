/*
 * TEST INFRASTRUCTURE (oracle): only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use anything under oracle/.
 *
 * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Written from the
 * published round function:
 *   (hi0, lo0) = M0 * c0 ; (hi1, lo1) = M1 * c2
 *   c' = { hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0 }
 *   k' = { k0 + W0, k1 + W1 }          (ten rounds, key bumped between rounds)
 * Pinned by the Random123 known-answer vectors and by rocRAND's independent
 * implementation (tests/golden/philox_kat.json).
 *
 * It replaces the reference's per-thread xoshiro256++ stream
 * (utils.cuh:60-72, seeded from a heap address at graph.cuh:116): every draw
 * here is a pure function of (seed, walker, step, attempt), so CPU and GPU
 * produce identical walks.
 */
#ifndef BINGO_ORACLE_PHILOX_H
#define BINGO_ORACLE_PHILOX_H
#include <stdint.h>

#define ORC_PHILOX_M0 0xD2511F53u
#define ORC_PHILOX_M1 0xCD9E8D57u
#define ORC_PHILOX_W0 0x9E3779B9u
#define ORC_PHILOX_W1 0xBB67AE85u

static inline void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)ORC_PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)ORC_PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += ORC_PHILOX_W0;
        k1 += ORC_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Convenience: block for (seed, a, b, c, d). */
static inline void orc_philox_block(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                    uint32_t out[4])
{
    uint32_t ctr[4] = {a, b, c, d};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

#endif
