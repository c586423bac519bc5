// Oracle (test infrastructure only): the seeded permeability generator of the synthetic inputs.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3",
// SC'11), keyed by (seed, problem id), counter (cell, attempt, 0x4B, 0). Two 53-bit uniforms
// v1, v2 in (-1,1) feed Marsaglia's polar method; the first normal deviate z gives
// log10 K = mu + sigma*z and K = pow(10, log10 K) with the deterministic pow of detmath.h.
// The reference problems read K from spatial parameters; this generator only fixes the
// synthetic input so that the oracle and the GPU path see identical bits.
#pragma once
#include "detmath.h"
#include <cmath>
#include <cstdint>

namespace orc {

static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double lognormalK(uint32_t seed, uint32_t pid, uint32_t cell, double mu, double sigma)
{
    const uint32_t key[2] = {seed, pid};
    for (uint32_t attempt = 0; attempt < 1000; ++attempt) {
        const uint32_t ctr[4] = {cell, attempt, 0x4Bu, 0u};
        uint32_t r[4];
        philox4x32_10(ctr, key, r);
        const uint64_t a = ((uint64_t(r[0]) << 32) | r[1]) >> 11;
        const uint64_t b = ((uint64_t(r[2]) << 32) | r[3]) >> 11;
        const double v1 = (double(a) + 0.5) * 0x1.0p-52 - 1.0;
        const double v2 = (double(b) + 0.5) * 0x1.0p-52 - 1.0;
        const double s = v1 * v1 + v2 * v2;
        if (s >= 1.0 || s == 0.0) continue;
        const double z = v1 * std::sqrt(-2.0 * dlog(s) / s);
        return dpow(10.0, mu + sigma * z);
    }
    return dpow(10.0, mu);
}

} // namespace orc
