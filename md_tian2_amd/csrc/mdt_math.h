// mdt_math.h -- FP64 exponential for the EMT pair terms.
//
// Table-driven reduction (Tang's method): x = (64 k + j) ln2/64 + r with
// |r| <= ln2/128, exp(x) = 2^k * 2^(j/64) * exp(r).  2^(j/64) comes from a
// 64-entry table (correctly rounded doubles, generated in 60-digit decimal
// arithmetic), exp(r) - 1 from its Taylor polynomial of degree 5 (truncation
// r^6/720 below 3.5e-17 relative, a third of an ulp; degree 6 measured the
// same 1 ulp maximum for one more FMA per call).  Accuracy is within 2 ulp of the correctly rounded
// result for |x| < 700, which tests/test_math_exp.py checks against libm.
// ln2/64 is split hi/lo with hi holding 32 significant bits so that n*hi is
// exact for |n| < 2^21.
#pragma once
#include <math.h>

#ifdef __HIPCC__
#define MDT_MHD __host__ __device__ __forceinline__
#else
#define MDT_MHD inline
#endif

#define MDT_EXP_TAB_INIT                                                                     \
    {1.0, 1.0108892860517005, 1.0218971486541166, 1.0330248790212284,                         \
     1.0442737824274138, 1.0556451783605572, 1.0671404006768237, 1.0787607977571199,         \
     1.0905077326652577, 1.102382583307841, 1.1143867425958924, 1.1265216186082418,          \
     1.1387886347566916, 1.1511892299529827, 1.1637248587775775, 1.1763969916502812,         \
     1.189207115002721, 1.202156731452703, 1.215247359980469, 1.22848053610687,              \
     1.241857812073484, 1.255380757024691, 1.2690509571917332, 1.2828700160787783,           \
     1.2968395546510096, 1.3109612115247644, 1.3252366431597413, 1.339667524053303,          \
     1.3542555469368927, 1.3690024229745905, 1.383909881963832, 1.3989796725383112,          \
     1.4142135623730951, 1.42961333839197, 1.4451808069770467, 1.460917794180647,            \
     1.4768261459394993, 1.4929077282912648, 1.5091644275934228, 1.5255981507445384,         \
     1.5422108254079407, 1.559004400237837, 1.5759808451078865, 1.593142151342267,           \
     1.6104903319492543, 1.6280274218573478, 1.645755478153965, 1.6636765803267364,          \
     1.681792830507429, 1.7001063537185235, 1.718619298122478, 1.7373338352737062,           \
     1.7562521603732995, 1.7753764925265212, 1.7947090750031072, 1.8142521755003989,         \
     1.8340080864093424, 1.8539791250833855, 1.8741676341103, 1.8945759815869656,            \
     1.9152065613971474, 1.9360617934922943, 1.9571441241754002, 1.978456026387951}

#ifdef __HIPCC__
#ifndef __HIP_DEVICE_COMPILE__
// host pass of hipcc: declarations only, device code is not generated here
__device__ double mdt_rcp(double x);
__device__ double mdt_rsqrt(double x);
__device__ float mdt_rcpf(float x);
__device__ float mdt_exp2f(float x);
#else
// FP32 pair terms: single hardware instructions, each within 1 ulp
// (v_rcp_f32; v_exp_f32 = 2^x, results below 2^-126 flush to zero).
__device__ __forceinline__ float mdt_rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float mdt_exp2f(float x) { return __builtin_amdgcn_exp2f(x); }

// 1/x for finite x >= 1e-300: hardware reciprocal seed and two Newton steps
// (the refinement LLVM itself uses inside its f64 division expansion).
__device__ __forceinline__ double mdt_rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// 1/sqrt(x) for finite x > 0: hardware seed and two Newton steps.
// 1/sqrt(x): hardware seed and one Newton step.  Measured on gfx950 over 4M
// x in [0.5, 100) (squared pair distances in A^2): the seed errs by up to
// 2.5e8 ulp (2^-24.5 relative), one step by 19.5 ulp (4e-15 relative), a
// second step would reach 1 ulp for 4 more instructions per call.  A 4e-15
// relative error in r moves the pair terms by ~1e-13 relative, far below
// the 1e-9 agreement with the reference's arithmetic.
__device__ __forceinline__ double mdt_rsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double h = x * y;
    const double e = fma(-h, y, 1.0);  // 1 - x y^2
    return fma(y * 0.5, e, y);
}
#endif
#endif

// exp(x) for |x| < 700 with `tab` the 64 entries of MDT_EXP_TAB_INIT
// (held in LDS on the device).
MDT_MHD double mdt_exp(double x, const double *tab) {
    const double t = x * 92.33248261689366;  // 64 / ln 2
    const double nf = rint(t);
    const int n = (int)nf;
    double r = fma(-nf, 0.01083042469326756, x);  // ln2/64, high part
    r = fma(-nf, 2.9815858269852933e-12, r);      // ln2/64, low part
    const double p =
        r * (1.0 + r * (0.5 + r * (1.0 / 6.0 + r * (1.0 / 24.0 + r * (1.0 / 120.0)))));
    const double tj = tab[n & 63];
    return ldexp(fma(tj, p, tj), n >> 6);
}
