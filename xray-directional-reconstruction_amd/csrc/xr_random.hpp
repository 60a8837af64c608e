// Counter-based random numbers and samplers shared by the synthetic-data
// kernels and the host helpers.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011): a 4x32-bit counter and 2x32-bit key, ten rounds of
// 32x32->64 multiplies with M0 = 0xD2511F53, M1 = 0xCD9E8D57 and Weyl key
// increments W0 = 0x9E3779B9, W1 = 0xBB67AE85.
// Poisson sampling: inversion by sequential multiplication for lambda < 10;
// PTRS transformed rejection (Hoermann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and Economics
// 12, 1993) for lambda >= 10.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace xr {

struct Philox {
    uint32_t c[4];
    uint32_t k[2];
};

__host__ __device__ inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0], n1 = lo1, n2 = hi0 ^ c[3] ^ k[1], n3 = lo0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
}

// Philox4x32-10 of counter (idx_lo, idx_hi, draw, 0) under key (seed_lo, seed_hi).
__host__ __device__ inline void philox4x32(uint64_t seed, uint64_t idx, uint32_t draw, uint32_t out[4]) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), draw, 0u};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int r = 0; r < 10; r++) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    for (int i = 0; i < 4; i++) out[i] = c[i];
}

// Uniform double in (0, 1) from two 32-bit words (53 random bits, never 0).
__host__ __device__ inline double u01(uint32_t a, uint32_t b) {
    const uint64_t m = (((uint64_t)a << 21) ^ (uint64_t)b) & ((1ull << 53) - 1);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Stream of uniforms for one element: two per Philox call.
struct UniformStream {
    uint64_t seed, idx;
    uint32_t draw = 0;
    uint32_t buf[4];
    int left = 0;
    __host__ __device__ UniformStream(uint64_t s, uint64_t i) : seed(s), idx(i) {}
    __host__ __device__ double next() {
        if (left == 0) {
            philox4x32(seed, idx, draw++, buf);
            left = 2;
        }
        const double u = (left == 2) ? u01(buf[0], buf[1]) : u01(buf[2], buf[3]);
        left--;
        return u;
    }
};

__host__ __device__ inline double poisson_sample(double lam, UniformStream& us) {
    if (!(lam > 0.0)) return 0.0;
    if (lam < 10.0) {
        const double L = exp(-lam);
        double p = 1.0;
        double k = 0.0;
        for (;;) {
            p *= us.next();
            if (p <= L) return k;
            k += 1.0;
        }
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = us.next() - 0.5;
        const double V = us.next();
        const double usv = 0.5 - fabs(U);
        const double k = floor((2.0 * a / usv + b) * U + lam + 0.43);
        if (usv >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (usv < 0.013 && V > usv)) continue;
        if (log(V) + log(invalpha) - log(a / (usv * usv) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
}

}  // namespace xr
