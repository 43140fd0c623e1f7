// mdt_philox.h -- counter-based normal deviates for the propagators.
//
// Generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), with its published multipliers and
// Weyl key increments.  key = (seed, trajectory id); counter = (step, slot,
// stream, component) with slot = atom*P + bead.  Every (trajectory, step,
// atom, bead, component) therefore owns one draw, independent of launch
// shape and batch composition.
//
// Deviates use the Box-Muller cosine branch in the reference's form
// sqrt(-2 ln u1) cos(2 pi u2) (src/useful_things.f90:121-173); the stream
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

MDT_HD double mdt_normal(uint32_t seed, uint32_t traj, uint32_t step, uint32_t stream,
                         uint32_t slot, uint32_t comp) {
    uint32_t c[4] = {step, slot, stream, comp};
    mdt_philox4x32_10(c, seed, traj);
    const double rnd1 = 1.0 - mdt_u53(c[0], c[1]);
    const double rnd2 = mdt_u53(c[2], c[3]);
    return sqrt(-2 * log(rnd1)) * cos(2 * 3.14159265358979323846 * rnd2);
}

MDT_HD double mdt_uniform(uint32_t seed, uint32_t traj, uint32_t step, uint32_t stream,
                          uint32_t slot, uint32_t comp) {
    uint32_t c[4] = {step, slot, stream, comp};
    mdt_philox4x32_10(c, seed, traj);
    return mdt_u53(c[0], c[1]);
}
