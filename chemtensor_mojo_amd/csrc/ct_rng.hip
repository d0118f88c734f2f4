// Seeded random MPS on the host: create_random_mps_c_compatible
// (reference src/tests/benchmarks/rng_c_compat.mojo:174-231), a clone of
// chemtensor's construct_random_mps with PCG32 (pcg_basic.c).  Host code: it
// only produces the input tensors that are then copied to the device.
#include <cmath>

#include "ct_internal.h"

namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 0;
    // pcg32_random_r (rng_c_compat.mojo:21-30)
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        const uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // pcg32_srandom_r (rng_c_compat.mojo:49-56)
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0;
        inc = (initseq << 1) | 1u;
        next();
        state += initstate;
        next();
    }
};

struct Rng {
    Pcg32 g0, g1;
    // seed_rng_state (rng_c_compat.mojo:129-149)
    explicit Rng(uint64_t seed) {
        Pcg32 b;
        b.seed(seed, 6364136223846793005ULL * seed + 1442695040888963407ULL);
        uint64_t v[8];
        for (auto& x : v) x = b.next();
        const uint64_t s1 = (v[0] << 32) | v[1], s2 = (v[2] << 32) | v[3];
        const uint64_t q1 = (v[4] << 32) | v[5];
        uint64_t q2 = (v[6] << 32) | v[7];
        const uint64_t mask = 0x7FFFFFFFFFFFFFFFULL;
        if ((q1 & mask) == (q2 & mask)) q2 = ~q2;
        g0.seed(s1, q1);
        g1.seed(s2, q2);
    }
    uint64_t u64() {  // pcg32x2_random_r (rng_c_compat.mojo:69-74)
        const uint64_t hi = g0.next();
        return (hi << 32) | g1.next();
    }
    uint64_t bounded(uint64_t bound) {  // pcg32x2_boundedrand_r (:76-84)
        const uint64_t thr = (~bound + 1u) % bound;
        for (;;) {
            const uint64_t r = u64();
            if (r >= thr) return r % bound;
        }
    }
    void choice_advance(int64_t bound, int64_t k) {  // rand_choice RNG use (:86-96)
        for (int64_t i = 0; i < k; ++i) (void)bounded((uint64_t)(bound - k + 1 + i));
    }
    float randnf() {  // Box-Muller in float32 (:98-108)
        const float inv = 2.32830644e-10f;
        float u1 = (float)g0.next() * inv;
        const float u2 = (float)g0.next() * inv;
        if (u1 == 0.0f) u1 = 1.0f;
        return sqrtf(-2.0f * logf(u1)) * sinf(6.2831853071795864769f * u2);
    }
};

}  // namespace

extern "C" int ct_random_mps_c_compatible(int nsites, int64_t d, int64_t chi_max, uint64_t seed, int64_t* dims,
                                          double* data, int64_t* total) {
    CT_ARG(nsites >= 1, 1, "nsites must be >= 1");
    CT_ARG(d >= 1, 2, "d must be >= 1");
    CT_ARG(chi_max >= 1, 3, "chi_max must be >= 1");
    CT_ARG(dims != nullptr, 5, "dims is NULL");
    Rng rng(seed);
    // bond dimensions (rng_c_compat.mojo:190-214)
    for (int l = 0; l <= nsites; ++l) dims[l] = 0;
    dims[0] = 1;
    dims[nsites] = 1;
    const int mid = (nsites + 1) / 2;
    for (int l = 1; l < mid; ++l) {
        const int64_t full = dims[l - 1] * d;
        dims[l] = full < chi_max ? full : chi_max;
        if (full > chi_max) rng.choice_advance(full, chi_max);
    }
    for (int l = nsites - 1; l >= mid; --l) {
        const int64_t full = dims[l + 1] * d;
        dims[l] = full < chi_max ? full : chi_max;
        if (full > chi_max) rng.choice_advance(full, chi_max);
    }
    int64_t n = 0;
    for (int i = 0; i < nsites; ++i) n += dims[i] * d * dims[i + 1];
    if (total) *total = n;
    if (!data) return 0;
    // entries randnf() * float32(1/sqrt(nelem)) (rng_c_compat.mojo:216-227)
    int64_t off = 0;
    for (int i = 0; i < nsites; ++i) {
        const int64_t ne = dims[i] * d * dims[i + 1];
        const float sc = (float)(1.0 / std::sqrt((double)ne));
        for (int64_t k = 0; k < ne; ++k) data[off++] = (double)(rng.randnf() * sc);
    }
    return 0;
}
