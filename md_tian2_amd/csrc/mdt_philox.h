// mdt_philox.h -- counter-based normal deviates for the propagators.
//
// Generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), with its published multipliers and
// Weyl key increments.  key = (seed, trajectory id); counter = (step, slot,
// stream, component) with slot = atom*P + bead.  Every (trajectory, step,
// atom, bead, component) therefore owns one draw, independent of launch
// shape and batch composition.
//
// Deviates are Box-Muller pairs: the draw (step, slot, stream, k) gives
// u1, u2 and the two independent standard normals sqrt(-2 ln u1) cos(2 pi
// u2) (component 2k) and sqrt(-2 ln u1) sin(2 pi u2) (component 2k+1), so
// an atom-bead's three components take two draws.  The reference draws a
// fresh (u1, u2) per deviate and keeps the cosine (src/useful_things.f90:
// 121-173); the two are equal in distribution, and the reference's stream
// itself (Fortran random_number) is compiler-runtime specific and is not
// reproduced.
#pragma once
#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define MDT_HD __host__ __device__ __forceinline__
#else
#define MDT_HD inline
#endif

MDT_HD void mdt_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// 53-bit uniform in [0, 1)
MDT_HD double mdt_u53(uint32_t a, uint32_t b) {
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)m * (1.0 / 9007199254740992.0);
}

#ifdef __HIPCC__
// the Box-Muller pair k of (step, slot, stream): *zc = component 2k, *zs = 2k+1
__device__ __forceinline__ void mdt_normal_pair(uint32_t seed, uint32_t traj, uint32_t step, uint32_t stream,
                                                uint32_t slot, uint32_t k, double *zc, double *zs) {
    uint32_t c[4] = {step, slot, stream, k};
    mdt_philox4x32_10(c, seed, traj);
    const double rnd1 = 1.0 - mdt_u53(c[0], c[1]);
    const double rnd2 = mdt_u53(c[2], c[3]);
    const double rho = sqrt(-2 * log(rnd1));
    double sn, cs;
    sincospi(2 * rnd2, &sn, &cs);
    *zc = rho * cs;
    *zs = rho * sn;
}

// component comp (0, 1, 2) of (step, slot, stream)
__device__ __forceinline__ double mdt_normal(uint32_t seed, uint32_t traj, uint32_t step, uint32_t stream,
                                            uint32_t slot, uint32_t comp) {
    double zc, zs;
    mdt_normal_pair(seed, traj, step, stream, slot, comp >> 1, &zc, &zs);
    return (comp & 1) ? zs : zc;
}

// the three components of (step, slot, stream): two draws
__device__ __forceinline__ void mdt_normal3(uint32_t seed, uint32_t traj, uint32_t step, uint32_t stream,
                                           uint32_t slot, double z[3]) {
    mdt_normal_pair(seed, traj, step, stream, slot, 0u, &z[0], &z[1]);
    uint32_t c[4] = {step, slot, stream, 1u};
    mdt_philox4x32_10(c, seed, traj);
    const double rnd1 = 1.0 - mdt_u53(c[0], c[1]);
    const double rnd2 = mdt_u53(c[2], c[3]);
    z[2] = sqrt(-2 * log(rnd1)) * cospi(2 * rnd2);
}
#endif

MDT_HD double mdt_uniform(uint32_t seed, uint32_t traj, uint32_t step, uint32_t stream,
                          uint32_t slot, uint32_t comp) {
    uint32_t c[4] = {step, slot, stream, comp};
    mdt_philox4x32_10(c, seed, traj);
    return mdt_u53(c[0], c[1]);
}
