// nh_rng.h -- WyHash PrimeableRng, RngSampler draws and the FNV-1a dedup hash.
//
// necsim/impls/no-std/src/cogs/rng/wyhash.rs:19-119 and
// necsim/core/src/cogs/rng.rs:33-198.  All arithmetic is wrapping u64 as in
// the reference's release build.
#pragma once

#include "nh_math.h"

namespace nh {

constexpr uint64_t WY_P0 = 0xa0761d6478bd642full;
constexpr uint64_t WY_P1 = 0xe7037ed1a0b428dbull;
constexpr uint64_t WY_P2 = 0x8ebc6af09c88c6e3ull;
constexpr uint64_t WY_P5 = 0xeb44accab455d165ull;
constexpr uint64_t SEAHASH_K = 0x6eed0e9da4d94a4full;
constexpr uint64_t INV_PHI = 0x9e3779b97f4a7c15ull; // poisson.rs:10
constexpr double TWO_POW_M53 = 1.1102230246251565e-16; // 0x1.0p-53

// wyhash.rs:85-96: xor-fold of the 128-bit product.  On gfx950 the product
// is assembled from four 32x32+64 -> 64 multiply-adds (v_mad_u64_u32); no
// intermediate sum overflows 64 bits.
NH_HD uint64_t wymum(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32);
    uint32_t bl = (uint32_t)b, bh = (uint32_t)(b >> 32);
    uint64_t p0 = (uint64_t)al * bl;
    uint64_t mid = (uint64_t)al * bh + (p0 >> 32);
    uint64_t mid2 = (uint64_t)ah * bl + (uint32_t)mid;
    uint64_t lo = (mid2 << 32) | (uint32_t)p0;
    uint64_t hi = (uint64_t)ah * bh + (mid >> 32) + (mid2 >> 32);
#else
    unsigned __int128 r = (unsigned __int128)a * (unsigned __int128)b;
    uint64_t lo = (uint64_t)r;
    uint64_t hi = (uint64_t)(r >> 64);
#endif
    return (a ^ lo) ^ (b ^ hi);
}

// wyhash.rs:99-119
NH_HD uint64_t seahash_diffuse(uint64_t x) {
    x *= SEAHASH_K;
    x ^= (x >> 32) >> (x >> 60);
    x *= SEAHASH_K;
    return x;
}

NH_HD uint64_t rot32(uint64_t x) { return (x << 32) | (x >> 32); }

// PrimeableRng::prime_with, wyhash.rs:67-80.  Returns the new `state`; the
// `seed` never changes after seeding.
// prime_key is the part that depends only on the seed and the location,
// which every prime of one simulation at one location shares: the kernels
// compute it once per move, prime_keyed adds the time.
NH_HD uint64_t prime_key(uint64_t seed, uint64_t location_index) {
    return rot32(seahash_diffuse(location_index)) ^ (seed ^ WY_P0);
}
NH_HD uint64_t prime_keyed(uint64_t location_key, uint64_t time_index) {
    uint64_t t = seahash_diffuse(time_index);
    uint64_t h = wymum(location_key, rot32(t) ^ WY_P2);
    return wymum(h, 16ull ^ WY_P5);
}
NH_HD uint64_t prime(uint64_t seed, uint64_t location_index, uint64_t time_index) {
    return prime_keyed(prime_key(seed, location_index), time_index);
}

// RngCore::sample_u64, wyhash.rs:53-63
NH_HD uint64_t sample_u64(uint64_t &state) {
    state += WY_P0;
    return seahash_diffuse(wymum(state ^ WY_P1, state));
}

// rng.rs:75-81: U[0, 1)
NH_HD double uniform_closed_open(uint64_t &state) {
    return (double)(sample_u64(state) >> 11) * TWO_POW_M53;
}
// rng.rs:86-93: U(0, 1]
NH_HD double uniform_open_closed(uint64_t &state) {
    return (double)((sample_u64(state) >> 11) + 1) * TWO_POW_M53;
}

// rng.rs:180-197 (Box-Muller with sigma and mu)
// FLAT: the straight-line sin / cos (same bits), for lone lanes
template <bool FLAT = false>
NH_HD void sample_2d_normal(uint64_t &state, double mu, double sigma, double &a, double &b) {
    double u0 = uniform_open_closed(state);
    double u1 = uniform_closed_open(state);
    double r = sqrt_(-2.0 * ln(u0));
    double theta = -6.283185307179586 * u1; // -TAU
    double s, c;
    if (FLAT) sincos_flat(theta, s, c);
    else sincos(theta, s, c);
    a = (r * s) * sigma + mu;
    b = (r * c) * sigma + mu;
}

// SeedableRng::seed_from_u64 (rng.rs:33-63: rand 0.7's PCG32 fill) followed by
// WyHash::from_seed (wyhash.rs:41-49: seed = u64::from_le_bytes)
inline uint64_t seed_from_u64(uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull;
    const uint64_t INC = 11634580027462260723ull;
    uint64_t out = 0;
    for (int chunk = 0; chunk < 2; chunk++) {
        state = state * MUL + INC;
        uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        uint32_t v = (xs >> rot) | (xs << ((32u - rot) & 31u));
        out |= (uint64_t)v << (32 * chunk);
    }
    return out;
}

// FNV-1a 64 (crate fnv 1.0.7) of a SpeciationSample's Hash stream:
// x, y, index (u32 each), sample_time bits, speciation_sample bits (u64 each),
// little-endian bytes (event_sampler/tracking.rs:69-75).
NH_HD uint64_t fnv_u32(uint64_t h, uint32_t v) {
    for (int b = 0; b < 4; b++) {
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 0x100000001b3ull;
    }
    return h;
}
NH_HD uint64_t fnv_u64(uint64_t h, uint64_t v) {
    for (int b = 0; b < 8; b++) {
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 0x100000001b3ull;
    }
    return h;
}
NH_HD uint64_t speciation_sample_hash(double u, double time, uint32_t x, uint32_t y, uint32_t i) {
    uint64_t h = 0xcbf29ce484222325ull;
    h = fnv_u32(h, x);
    h = fnv_u32(h, y);
    h = fnv_u32(h, i);
    h = fnv_u64(h, bits(time));
    h = fnv_u64(h, bits(u));
    return h;
}

} // namespace nh
