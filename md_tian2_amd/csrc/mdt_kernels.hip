// mdt_kernels.hip -- CDNA4 (gfx950) kernels for batched EMT forces and the
// md_tian2 propagators.
//
// k_replica: calc_force, one workgroup per replica (trajectory t, bead b).
//   phase 0: stage the replica in LDS, folded into its cell, with fractional
//            coordinates;
//   phase 1: TPA lanes per atom scan all atoms with the minimum-image
//            convention and build an ordered per-atom neighbour list of
//            (index | periodic-image code << 11), compacted with a wavefront
//            ballot;
//   phase 2 (pass A): every unordered pair once, by its owner atom:
//            sigma_k += csig(tk,tn) f_sigma_tn(r) in registers and
//            sigma_n += csig(tn,tk) f_sigma_tk(r) into a fixed-point LDS
//            accumulator; then per atom s, dens, E_coh + vref and
//            G_a = dE/dsigma_a;
//   phase 3 (pass B): the pair's dE/dr (density part through G of both
//            atoms, pair-potential part directly) gives +F on the owner and
//            -F on the partner (Newton's third law), both accumulated in
//            fixed point; then set_acceleration.
//   Fixed-point integer accumulation is associative, so results are bitwise
//   reproducible whatever order the lanes' atomics land in.
// k_traj: one workgroup per trajectory: propagate_2 of the previous step,
//   propagate_1, COM removal and (P > 1) the ring-polymer free step
//   (src/md_algo.f90:37-106, src/rpmd.f90:309-452).
#include <hip/hip_runtime.h>

#include "mdt_device.h"
#include "mdt_math.h"
#include "mdt_philox.h"

#define MDT_NBR_IDX_BITS 11
#define MDT_NBR_IDX_MASK 0x7FF

namespace {

// Tables every lane indexes by atom type: kept in LDS (divergent indices).
struct LdsTables {
    double eta2[MDT_MAXT], bs0[MDT_MAXT], kdb[MDT_MAXT], betaeta2[MDT_MAXT];
    double e0[MDT_MAXT], lambda[MDT_MAXT], v0[MDT_MAXT], kappa[MDT_MAXT], n0[MDT_MAXT];
    double densc[MDT_MAXT];
    double csig[MDT_MAXT * MDT_MAXT], etad[MDT_MAXT * MDT_MAXT], wv[MDT_MAXT * MDT_MAXT];
    double csig48[MDT_MAXT * MDT_MAXT];  // csig * 2^48: partner sigma terms in fixed-point units
    double shift[27 * 3];  // cartesian lattice translation of each image code
    double exptab[MDT_EXP_TAB_N];
    int owner[MDT_MAXT], pair_on[MDT_MAXT * MDT_MAXT], algo[MDT_MAXT], is_proj[MDT_MAXT];
    // FP32 copies for the single-precision pair terms; *_l2 carry a factor
    // log2(e) so that exp(-a x) = 2^(-a_l2 x)
    float f_eta2l2[MDT_MAXT], f_kdbl2[MDT_MAXT], f_kdb[MDT_MAXT];
    float f_csig[MDT_MAXT * MDT_MAXT], f_etad[MDT_MAXT * MDT_MAXT], f_wv[MDT_MAXT * MDT_MAXT];
};

#define MDT_LOG2E 1.4426950408889634

struct CellRegs {
    double h[9], ih[9];
};

__device__ __forceinline__ void load_cell(const BatchView &bv, int t, CellRegs &c) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        c.h[k] = bv.cell[9 * t + k];
        c.ih[k] = bv.icell[9 * t + k];
    }
}

template <int TPA>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = 1; o < TPA; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Langevin coefficients, src/md_algo.f90:165-198 (langevin_1) and :253-285
struct LanCoef {
    double c0, c1, c2, sr, sv, crv;
};

__device__ __forceinline__ LanCoef lan_coef(double fric, double dt, double temp, bool exact) {
#pragma clang fp contract(off)
    LanCoef L;
    const double xidt = fric * dt;
    const double xidt2 = xidt * xidt;
    const double ixidt = dt / xidt;
    if (exact) {
        L.c0 = exp(-xidt);
        L.c1 = (1 - L.c0) * ixidt;
        L.c2 = (1 - L.c1 / dt) * ixidt;
        L.sr = ixidt * sqrt(temp * (2 * xidt - 3 + 4 * L.c0 - L.c0 * L.c0));
        L.sv = sqrt(temp * (1 - L.c0 * L.c0));
        const double om = 1 - L.c0;
        L.crv = ixidt * temp * (om * om) / (L.sr * L.sv);
    } else {
        L.c0 = 1 - xidt + 0.5 * xidt2;
        L.c1 = (1 - 0.5 * xidt + 2 * MDT_TWELFTH * xidt2) * dt;
        L.c2 = (0.5 - 2 * MDT_TWELFTH * xidt + 0.5 * MDT_TWELFTH * xidt2) * dt;
        L.sr = dt * sqrt(temp * (8 * MDT_TWELFTH * xidt - 0.5 * xidt2));
        L.sv = sqrt(2 * temp * xidt * (1 - xidt));
        L.crv = 0.5 * sqrt(3.0) * (1 - 0.125 * xidt);
    }
    return L;
}

// the branch test of md_algo.f90:176 for one bead (1e-2 is a default-real literal)
__device__ __forceinline__ bool lan_big(double fric, double dt) { return fric * dt > (double)1e-2f; }

// ldfa, src/md_algo.f90:506-570: friction [1/fs] from the EMT density
__device__ __forceinline__ double ldfa(double dens, double mass, int *bad) {
#pragma clang fp contract(off)
    const double convert = 1.00794 * MDT_AMU2MASS;
    const double lo = (double)(-1e-12f), hi = (double)0.36f;
    double fric;
    if (lo <= dens && dens <= hi) {
        const double c[12] = {0.0802484, -1.12851, 9.28508, 2.10064, -843.419, 8.85354e3,
                              -4.89023e4, 1.6741e5, -3.67098e5, 5.03476e5, -3.9426e5, 1.34763e5};
        fric = 0.0;
        double temp = dens;
#pragma unroll
        for (int j = 0; j < 12; ++j) {
            fric = fric + c[j] * temp;
            temp = temp * dens;
        }
    } else if (dens > hi) {
        fric = 0.001 * (4.7131 - exp(-4.41305 * dens));
    } else {
        *bad = 1;
        fric = 0.0;
    }
    return fric * convert / MDT_HBAR / mass;
}

__device__ void load_tables(LdsTables &T, const EmtConst *__restrict__ ec, const PropConst *__restrict__ pc,
                            const CellRegs &cell, int lane, int nthr) {
    for (int k = lane; k < MDT_MAXT; k += nthr) {
        T.eta2[k] = ec->eta2[k];
        T.bs0[k] = ec->betas0[k];
        T.kdb[k] = ec->kappadbeta[k];
        T.betaeta2[k] = ec->betaeta2[k];
        T.e0[k] = ec->e0[k];
        T.lambda[k] = ec->lambda[k];
        T.v0[k] = ec->v0[k];
        T.kappa[k] = ec->kappa[k];
        T.n0[k] = ec->n0[k];
        T.densc[k] = ec->densc[k];
        T.owner[k] = ec->owner[k];
        T.algo[k] = pc->algo[k];
        T.is_proj[k] = pc->is_proj[k];
        T.f_eta2l2[k] = (float)(ec->eta2[k] * MDT_LOG2E);
        T.f_kdbl2[k] = (float)(ec->kappadbeta[k] * MDT_LOG2E);
        T.f_kdb[k] = (float)ec->kappadbeta[k];
    }
    for (int k = lane; k < MDT_MAXT * MDT_MAXT; k += nthr) {
        T.csig[k] = ec->csig[k];
        T.csig48[k] = ec->csig[k] * 281474976710656.0;  // exact (power of two)
        T.etad[k] = ec->etad[k];
        T.wv[k] = ec->wv[k];
        T.pair_on[k] = ec->pair_on[k];
        T.f_csig[k] = (float)ec->csig[k];
        T.f_etad[k] = (float)ec->etad[k];
        T.f_wv[k] = (float)ec->wv[k];
    }
    for (int c = lane; c < 27; c += nthr) {
        const double m0 = (double)(c % 3 - 1), m1 = (double)((c / 3) % 3 - 1), m2 = (double)(c / 9 - 1);
        T.shift[3 * c + 0] = cell.h[0] * m0 + cell.h[3] * m1 + cell.h[6] * m2;
        T.shift[3 * c + 1] = cell.h[1] * m0 + cell.h[4] * m1 + cell.h[7] * m2;
        T.shift[3 * c + 2] = cell.h[2] * m0 + cell.h[5] * m1 + cell.h[8] * m2;
    }
    const double et[MDT_EXP_TAB_N] = MDT_EXP_TAB_INIT;
    for (int k = lane; k < MDT_EXP_TAB_N; k += nthr) T.exptab[k] = et[k];
}

}  // namespace

// Fixed-point scales of the LDS accumulators that receive pair
// contributions from other lanes: integer addition is associative, so the
// totals do not depend on the order in which lanes add them.
#define MDT_W48 281474976710656.0            // 2^48
#define MDT_IW48 (1.0 / 281474976710656.0)
#define MDT_F_SCALE 4398046511104.0          // 2^42: force up to 2.1e6 eV/A, step 2.3e-13
#define MDT_F_ISCALE (1.0 / 4398046511104.0)

// Round-to-nearest-even of y to int64, returned with the rounded value as a
// double in *r.  For |y| < 2^51 adding 1.5 * 2^52 leaves rint(y) in the low
// mantissa bits (the sum lies in [2^52, 2^53), where the spacing is 1 and the
// default rounding is to nearest even), so the integer is the difference of
// two bit patterns: three instructions instead of the f64 -> i64 expansion.
// fix_round_small requires |y| < 2^51; its callers vote across the wave and
// take rint, the conversion and the overflow test (the callers' 9e18 bound)
// only when some lane's |y| is larger, so the common case issues neither.
#define MDT_RND_MAGIC 6755399441055744.0  // 1.5 * 2^52
__device__ __forceinline__ long long fix_round_small(double y, double *r) {
    const double t = y + MDT_RND_MAGIC;
    *r = t - MDT_RND_MAGIC;
    return __double_as_longlong(t) - __double_as_longlong(MDT_RND_MAGIC);
}

// sigma contributions span ~1e3 (close contact) down to ~1e-35 (pairs at the
// cut-off) and an isolated atom's sigma sits near the reference's 1e-30
// clamp, where G = dE/dsigma ~ 1/sigma: a sum with relative precision at
// any magnitude is needed.  Three int64 words at 2^-48, 2^-96 and 2^-144
// hold each contribution exactly down to 4.5e-44 (each conversion step is an
// exact power-of-two scaling and an exact subtraction of the rounded part).
__device__ __forceinline__ void sig_add_scaled(unsigned long long *hi, unsigned long long *mid,
                                               unsigned long long *lo, double y, int *bad) {
    // y: the contribution times 2^48 (callers fold the power of two into a
    // constant, an exact scaling).  The overflow test sits in the slow path
    // of the wave-uniform vote.
    double h, m, l;
    long long ih;
    if (__all(fabs(y) < 2251799813685248.0)) {  // 2^51
        ih = fix_round_small(y, &h);
    } else {
        if (!(fabs(y) < 9.0e18)) {
            *bad |= MDT_ST_FIXED_OVERFLOW;
            return;
        }
        h = rint(y);
        ih = (long long)h;
    }
    const double y1 = (y - h) * MDT_W48;  // |y1| <= 2^47
    const long long im = fix_round_small(y1, &m);
    const long long il = fix_round_small((y1 - m) * MDT_W48, &l);  // |.| <= 2^47
    atomicAdd(hi, (unsigned long long)ih);
    if (im != 0) atomicAdd(mid, (unsigned long long)im);
    if (il != 0) atomicAdd(lo, (unsigned long long)il);
}

__device__ __forceinline__ double sig_value(unsigned long long hi, unsigned long long mid, unsigned long long lo) {
    return (((double)(long long)lo * MDT_IW48 + (double)(long long)mid) * MDT_IW48 + (double)(long long)hi) *
           MDT_IW48;
}

// y: the contribution in fixed-point units (times the accumulator's scale)
__device__ __forceinline__ void fx_add_scaled(unsigned long long *acc, double y, int *bad) {
    double r;
    if (__all(fabs(y) < 2251799813685248.0)) {  // 2^51, wave-uniform branch
        atomicAdd(acc, (unsigned long long)fix_round_small(y, &r));
    } else if (fabs(y) < 9.0e18) {
        atomicAdd(acc, (unsigned long long)(long long)rint(y));
    } else {
        *bad |= MDT_ST_FIXED_OVERFLOW;
    }
}
__device__ __forceinline__ void fx_add(unsigned long long *acc, double x, double scale, int *bad) {
    fx_add_scaled(acc, x * scale, bad);
}

// s, embedded density, E_coh + vref and G = dE/dsigma of an atom of type tk
// with total sigma (src/pes_emt_mod.f90:793-869 and :871-972)
__device__ __forceinline__ void emt_atom(const LdsTables &T, int tk, double sig, double *Gk, double *Ek, double *dn) {
    const double clamp30 = (double)1e-30f;  // pes_emt_mod.f90:793-794
    const double sc = sig > clamp30 ? sig : clamp30;
    const double sl = sc > MDT_TOLERANCE ? sc : MDT_TOLERANCE;
    const double sv = -log(sl * MDT_TWELFTH) / T.betaeta2[tk];
    *dn = T.n0[tk] * exp(T.densc[tk] * sv);
    const double el = exp(-T.lambda[tk] * sv);
    const double ek = exp(-T.kappa[tk] * sv);
    *Ek = T.e0[tk] * ((1 + T.lambda[tk] * sv) * el - 1) + 6 * T.v0[tk] * ek;
    const double D = T.e0[tk] * T.lambda[tk] * T.lambda[tk] * (sv * el) + 6 * T.v0[tk] * T.kappa[tk] * ek;
    *Gk = -D / (T.betaeta2[tk] * sc);
}

// Image along one axis of a pair whose folded fractional difference x lies
// within 1e-4 of +-1/2 -- in a cell of even width, atoms of perfect lattice
// layers sit exactly half a cell apart.  The reference's minimum image
// (src/universe_mod.f90:411-428) rounds s = isimbox (r_j - r_i) of the
// unfolded positions half away from zero (anint), so the tie goes to the
// image that expression, evaluated in its order without fused
// multiply-adds, picks; returns m with x - m equal to its reduced
// coordinate (s - anint(s) is odd in r_j - r_i, so the pair's order does
// not matter).  (ia, ib, ic) = isimbox(axis, 1..3); ra, rb the x, y, z of
// the two atoms.
__device__ __forceinline__ double tie_image(double ia, double ib, double ic, double ra0, double ra1, double ra2, double rb0,
                                         double rb1, double rb2, double x) {
#pragma clang fp contract(off)
    const double s = ia * (ra0 - rb0) + ib * (ra1 - rb1) + ic * (ra2 - rb2);
    return rint(x - (s - round(s)));
}

// Minimum image of the pair (k, n) of a replica staged folded in LDS:
// d = p_k - p_n - shift[code], code the rounding m of the folded fractional
// difference (in {-1, 0, 1}^3, the list scan's image code).  The rounding
// is decided in FP32 from ihf (H^-1, 9 floats, column major; error ~1e-7
// against the 1e-4 tie window); a component within 1e-4 of +-1/2 takes the
// reference's image from the FP64 H^-1 ihg and the unfolded positions xr
// (SoA, N per component), both in global memory, as tie_image does.  Only
// 9 floats of the cell are read: this runs after both passes, where FP64
// cell registers kept live would cost the hot loops registers.
// Returns |d|^2.
__device__ __forceinline__ double pair_image(const float *ihf, const double *ihg, const double *shift,
                                             const double *px, const double *py, const double *pz,
                                             const double *xr, int N, int k, int n, double d[3]) {
    const double cx = px[k] - px[n], cy = py[k] - py[n], cz = pz[k] - pz[n];
    const float ex = (float)cx, ey = (float)cy, ez = (float)cz;
    const float f0 = ihf[0] * ex + ihf[3] * ey + ihf[6] * ez;
    const float f1 = ihf[1] * ex + ihf[4] * ey + ihf[7] * ez;
    const float f2 = ihf[2] * ex + ihf[5] * ey + ihf[8] * ez;
    float m0 = rintf(f0), m1 = rintf(f1), m2 = rintf(f2);
    const bool t0 = fabsf(fabsf(f0) - 0.5f) < 1e-4f, t1 = fabsf(fabsf(f1) - 0.5f) < 1e-4f,
               t2 = fabsf(fabsf(f2) - 0.5f) < 1e-4f;
    if (t0 | t1 | t2) {  // rare: the exact fractional difference and the reference's tie rule
        const double a0 = xr[k], a1 = xr[N + k], a2 = xr[2 * N + k];
        const double b0 = xr[n], b1 = xr[N + n], b2 = xr[2 * N + n];
        if (t0) m0 = (float)tie_image(ihg[0], ihg[3], ihg[6], a0, a1, a2, b0, b1, b2, ihg[0] * cx + ihg[3] * cy + ihg[6] * cz);
        if (t1) m1 = (float)tie_image(ihg[1], ihg[4], ihg[7], a0, a1, a2, b0, b1, b2, ihg[1] * cx + ihg[4] * cy + ihg[7] * cz);
        if (t2) m2 = (float)tie_image(ihg[2], ihg[5], ihg[8], a0, a1, a2, b0, b1, b2, ihg[2] * cx + ihg[5] * cy + ihg[8] * cz);
    }
    const int code = ((int)m0 + 1) + 3 * ((int)m1 + 1) + 9 * ((int)m2 + 1);
    d[0] = cx - shift[3 * code];
    d[1] = cy - shift[3 * code + 1];
    d[2] = cz - shift[3 * code + 2];
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// Hides v's derivation from the optimiser (an empty asm with a register
// operand; no instruction is emitted): in a loop that usually runs once,
// addresses derived from v are then formed per iteration instead of being
// strength-reduced into one induction register per array.
__device__ __forceinline__ void opaque(int &v) { asm volatile("" : "+v"(v)); }

// Sum of v over the block in a fixed order (butterfly within each wave, then
// the waves in order through wsum[nthr / 64]); the result is valid in lane 0.
// Every lane of the block must call it (it synchronises the block).
__device__ __forceinline__ double block_sum_fixed(double v, double *wsum, int lane, int nthr) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((lane & 63) == 0) wsum[lane >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (lane == 0)
        for (int w = 0; w < (nthr + 63) / 64; ++w) s += wsum[w];
    __syncthreads();
    return s;
}

// The same for four values at once (one pair of barriers); wsum holds 64.
__device__ __forceinline__ void block_sum4_fixed(double v[4], double *wsum, int lane, int nthr) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
    if ((lane & 63) == 0)
        for (int q = 0; q < 4; ++q) wsum[4 * (lane >> 6) + q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double s = 0.0;
        if (lane == 0)
            for (int w = 0; w < (nthr + 63) / 64; ++w) s += wsum[4 * w + q];
        v[q] = s;
    }
    __syncthreads();
}

// Pair ownership: each unordered pair is evaluated once, by its owner; the
// owner alternates with index parity, which gives every atom about half of
// its neighbours.  The 2^48 fixed point of the partner sigma sums bounds the
// force error of an isolated atom (sigma ~ 1e-11, the s clamp region) by
// ~1e-10 eV/A, since that atom's force scales with sigma itself.
__device__ __forceinline__ bool owns(int k, int n) { return (k < n) != (((k + n) & 1) != 0); }

// PREC (arithmetic of the pair terms):
//   0: FP64 throughout;
//   1: the cut-off function, the sigma and pair-potential exponentials and
//      their products in FP32;
//   2: only the cut-off exponential exp(acut (r - rcut)) and theta in FP32,
//      the sigma and pair-potential exponentials in FP64 (for metals whose
//      embedding stiffness makes FP32 sigma / V terms miss 1e-6 eV/atom and
//      1e-5 eV/A, e.g. Pt with V0 = 21.7 eV).
// In every mode the geometry (differences, r, the exponent arguments
// r - rcut and r - beta s0, which carry factors up to acut ~ 26/A), the
// per-atom embedding terms, G and every accumulation stay FP64 / fixed point.
// JUDGE: FP64 EMT forces of the selected replicas compared with
// bv.judge_f (the network's judge); nothing else is written.
// MAXB: the largest workgroup of the instantiation, 512, or 1024 for size
// classes of more than 512 atoms x threads per atom, whose few replicas
// would otherwise leave each CU two waves per SIMD.
template <int TPA, bool NBR_LDS, int PREC, bool JUDGE, int MAXB = 512>
__global__ void __launch_bounds__(MAXB, 4)
    k_replica(const EmtConst *__restrict__ ecg, const PropConst *__restrict__ pcg, BatchView bv, int add_fin) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ LdsTables T;
    const int rep = bv.rep_list ? bv.rep_list[blockIdx.x] : (int)blockIdx.x;
    if (bv.only && !bv.only[rep]) return;                                          // whole block: uniform
    if (bv.skip && bv.skip[rep]) return;                                           // whole block: uniform
    if (JUDGE && (rep % bv.judge_mod) != bv.judge_phase) return;                  // whole block: uniform
    const int t = bv.rep_traj[rep];
    const int traj = t;
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int b = bv.rep_bead[rep];
    const int N = bv.traj_N[t];
    const int cap = bv.nbcap < N - 1 ? bv.nbcap : (N > 1 ? N - 1 : 1);
    const long long soff = bv.traj_soff[t];
    const long long roff = soff + (long long)b * 3 * N;  // x block of this replica
    const long long doff = soff / 3 + (long long)b * N;   // dens / fixed block
    const long long aoff = bv.traj_aoff[t];
    const uint32_t tid = bv.traj_id[t];
    const double tsurf = bv.tsurf[t], tproj = bv.tproj[t];
    const int nthr = blockDim.x;
    const int lane = threadIdx.x;
    CellRegs cell;
    load_cell(bv, t, cell);
    load_tables(T, ecg, pcg, cell, lane, nthr);

    // LDS: px py pz | three words A0 A1 A2 | G | E | cnt | nbr | typ | tiny,
    // where the words hold the fractional coordinates in the scan, the sigma
    // fixed point (hi, mid, lo) in pass A and the force fixed point (x, y, z)
    // in pass B.  The lists hold the near pairs (r <= rnear; rnear + skin for
    // kept lists); the pairs beyond rnear with a tiny atom come from a direct
    // minimum-image scan of that atom's partners (pair_image).
    double *px = (double *)smem_raw;
    double *py = px + N, *pz = py + N;
    double *sx = pz + N, *sy = sx + N, *sz = sy + N;
    unsigned long long *shi = (unsigned long long *)sx, *smid = (unsigned long long *)sy;
    unsigned long long *slo = (unsigned long long *)sz;
    unsigned long long *fxa = shi, *fya = smid, *fza = slo;
    double *G = sz + N, *E = G + N;
    uint16_t *cnt = (uint16_t *)(E + N);
    uint16_t *nbr = NBR_LDS ? cnt + N : bv.nbr_global + (size_t)blockIdx.x * bv.nbr_stride;
    // kept lists staged in LDS: their copy between launches
    uint16_t *nbr_kept = (NBR_LDS && bv.rl_stage) ? bv.nbr_global + (size_t)blockIdx.x * bv.nbr_stride : nbr;
    uint8_t *typ = (uint8_t *)(cnt + N + (NBR_LDS ? (size_t)N * cap : 0));
    uint8_t *tiny = typ + N;
    __shared__ int s_anytiny;
    __shared__ double s_wsum[64];
    __shared__ float s_ihf[9];  // H^-1 in FP32 for the scan of the tiny atoms' partners
    if (lane < 9) s_ihf[lane] = (float)bv.icell[9 * t + lane];
    int bad_local = 0;
    // Verlet skin (lists in HBM): the lists of this block slot are built with
    // rnear + skin and kept while no atom moved more than skin / 2 since the
    // build; a kept list gets every entry's image recomputed from the current
    // positions (the same rounding and tie rule as the scan).  Only in cells
    // whose narrowest width exceeds 2 (rnear + skin): there every image
    // within rnear + skin has fractional components in (-1/2, 1/2), so it is
    // the rounded image the scan tests, and a pair within rnear now was
    // listed at the build in that same image.  In narrower cells the rounded
    // image need not be the nearest, and the lists are rebuilt at every
    // launch.  (The pairs beyond rnear never come from a list.)
    const double skin = bv.rl_skin;
    double wmin = 1e300;
    for (int q = 0; q < 3; ++q) {
        const double w = 1.0 / sqrt(cell.ih[q] * cell.ih[q] + cell.ih[q + 3] * cell.ih[q + 3] +
                                    cell.ih[q + 6] * cell.ih[q + 6]);
        wmin = w < wmin ? w : wmin;
    }
    const double rnear = ecg->rnear;
    const bool keep = (!NBR_LDS || bv.rl_stage) && bv.rl_valid != nullptr && 2.0 * (rnear + skin) < wmin;
    double *ref = keep ? bv.rl_ref + (long long)blockIdx.x * 3 * bv.rl_ns : nullptr;
    __shared__ int s_reuse;
    if (lane == 0) s_reuse = keep && bv.rl_valid[blockIdx.x];
    __syncthreads();
    if (s_reuse) {  // every thread reads the flag before any clears it
        double dmax = 0.0;
        for (int i = lane; i < N; i += nthr) {
            const double d0 = bv.x[roff + i] - ref[i], d1 = bv.x[roff + N + i] - ref[bv.rl_ns + i],
                         d2 = bv.x[roff + 2 * N + i] - ref[2 * bv.rl_ns + i];
            const double d2s = d0 * d0 + d1 * d1 + d2 * d2;
            dmax = d2s > dmax ? d2s : dmax;
        }
        __syncthreads();
        if (!(dmax <= 0.25 * skin * skin)) s_reuse = 0;  // benign race: every writer stores 0
    }
    __syncthreads();
    const bool reuse = s_reuse != 0;

    // ---- phase 0: stage ----
    for (int i = lane; i < N; i += nthr) {
        const int ty = bv.type[aoff + i];
        typ[i] = (uint8_t)ty;
        const double r[3] = {bv.x[roff + i], bv.x[roff + N + i], bv.x[roff + 2 * N + i]};
        if (keep && !reuse) {
            ref[i] = r[0];
            ref[bv.rl_ns + i] = r[1];
            ref[2 * bv.rl_ns + i] = r[2];
        }
        // fold into the cell: fractional in [0, 1)
        double f0 = cell.ih[0] * r[0] + cell.ih[3] * r[1] + cell.ih[6] * r[2];
        double f1 = cell.ih[1] * r[0] + cell.ih[4] * r[1] + cell.ih[7] * r[2];
        double f2 = cell.ih[2] * r[0] + cell.ih[5] * r[1] + cell.ih[8] * r[2];
        f0 -= floor(f0);
        f1 -= floor(f1);
        f2 -= floor(f2);
        sx[i] = f0;
        sy[i] = f1;
        sz[i] = f2;
        px[i] = cell.h[0] * f0 + cell.h[3] * f1 + cell.h[6] * f2;
        py[i] = cell.h[1] * f0 + cell.h[4] * f1 + cell.h[7] * f2;
        pz[i] = cell.h[2] * f0 + cell.h[5] * f1 + cell.h[8] * f2;
    }
    __syncthreads();
    const double acut = ecg->acut, rcut = ecg->rcut, cutr = ecg->cutr, cut2 = ecg->cut2hi;
    const float acutf = (float)acut, acutl2f = (float)(acut * MDT_LOG2E);
    const int apb = nthr / TPA;  // atoms per chunk
    const int wl = lane & 63;
    const unsigned long long gmask = (TPA >= 64 ? ~0ull : ((1ull << TPA) - 1ull)) << ((wl / TPA) * TPA);
    const unsigned long long below = (wl == 0) ? 0ull : ((~0ull) >> (64 - wl));

    // ---- phase 1: owned half neighbour lists (min image, ordered by index) ----
    // The scan only decides membership: it runs in FP32 on float copies of
    // the fractional coordinates (in the G/E words, unused until pass A) with
    // r^2 bound widened by 1e-4 relative; both passes re-check r <= cutr in
    // FP64.  A fractional difference within 1e-4 of +-1/2 takes its image
    // from the FP64 rounding, so the chosen image is the FP64 minimum image.
    // fractional coordinates and type per atom as one float4 (16 B: the G
    // and E words, 2N doubles); squared distances from the metric H^T H
    float4 *f4 = (float4 *)G;
    for (int i = lane; i < N; i += nthr)
        f4[i] = make_float4((float)sx[i], (float)sy[i], (float)sz[i], __int_as_float((int)typ[i]));
    int all_on = 1;
    for (int q = 0; q < ecg->ntypes * ecg->ntypes; ++q)
        if (!T.pair_on[(q / ecg->ntypes) * MDT_MAXT + q % ecg->ntypes]) all_on = 0;
    __syncthreads();
    float g00, g11, g22, g01, g02, g12;  // H^T H, cell vectors = columns of H
    {
        const double *h = cell.h;
        g00 = (float)(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        g11 = (float)(h[3] * h[3] + h[4] * h[4] + h[5] * h[5]);
        g22 = (float)(h[6] * h[6] + h[7] * h[7] + h[8] * h[8]);
        g01 = (float)(2 * (h[0] * h[3] + h[1] * h[4] + h[2] * h[5]));
        g02 = (float)(2 * (h[0] * h[6] + h[1] * h[7] + h[2] * h[8]));
        g12 = (float)(2 * (h[3] * h[6] + h[4] * h[7] + h[5] * h[8]));
    }
    // the lists hold the near pairs; a kept list is built with rnear + skin,
    // so a pair now within rnear was listed at the build.  Staged kept lists
    // (the rows in HBM, a copy in LDS during the launch) hold the whole
    // rnear + skin rows in HBM and only the entries now within rnear (the
    // FP32 bound of a build without skin) in LDS, in the same order: the
    // passes then see the entries of a list built at this launch.
    const double rn = keep ? rnear + skin : rnear;
    const float cut2f = (float)(rn * rn * (1.0 + 1e-4));
    const float near2f = (float)(rnear * rnear * (1.0 + 1e-4));
    const bool split = NBR_LDS && keep;
    uint16_t *kcnt = keep ? bv.rl_cnt + (long long)blockIdx.x * bv.rl_ns : nullptr;
    if (reuse) {  // kept list: the image of every entry from the current positions
        for (int k0 = 0; k0 < N; k0 += apb) {
            const int k = k0 + lane / TPA;
            const int sub = lane % TPA;
            if (k >= N) continue;
            const int c = kcnt[k];
            const float4 pk = f4[k];
            int cn = 0;  // staged: near entries so far
            for (int e0 = 0; e0 < c; e0 += TPA) {  // the same trip count in the atom's lanes
                const int e = e0 + sub;
                bool near = false;
                uint16_t ent = 0;
                if (e < c) {
                    const unsigned q = nbr_kept[(size_t)k * cap + e];
                    const int n = q & MDT_NBR_IDX_MASK;
                    const float4 pn = f4[n];
                    const float e0f = pk.x - pn.x, e1f = pk.y - pn.y, e2f = pk.z - pn.z;
                    float m0f = rintf(e0f), m1f = rintf(e1f), m2f = rintf(e2f);
                    const bool t0 = fabsf(e0f - m0f) > 0.4999f, t1 = fabsf(e1f - m1f) > 0.4999f,
                               t2 = fabsf(e2f - m2f) > 0.4999f;
                    if (t0 | t1 | t2) {
                        const double *xr = bv.x + roff;
                        const double a0 = xr[k], a1 = xr[N + k], a2 = xr[2 * N + k];
                        const double b0 = xr[n], b1 = xr[N + n], b2 = xr[2 * N + n];
                        if (t0) m0f = (float)tie_image(cell.ih[0], cell.ih[3], cell.ih[6], a0, a1, a2, b0, b1, b2, sx[k] - sx[n]);
                        if (t1) m1f = (float)tie_image(cell.ih[1], cell.ih[4], cell.ih[7], a0, a1, a2, b0, b1, b2, sy[k] - sy[n]);
                        if (t2) m2f = (float)tie_image(cell.ih[2], cell.ih[5], cell.ih[8], a0, a1, a2, b0, b1, b2, sz[k] - sz[n]);
                    }
                    const int code = ((int)m0f + 1) + 3 * ((int)m1f + 1) + 9 * ((int)m2f + 1);
                    ent = (uint16_t)(n | (code << MDT_NBR_IDX_BITS));
                    if (split) {
                        const float d0 = e0f - m0f, d1 = e1f - m1f, d2 = e2f - m2f;
                        const float r2 =
                            d0 * (g00 * d0 + g01 * d1 + g02 * d2) + d1 * (g11 * d1 + g12 * d2) + g22 * d2 * d2;
                        near = r2 <= near2f;
                    } else {
                        nbr[(size_t)k * cap + e] = ent;  // the rows in HBM: the image in place
                    }
                }
                if (split) {
                    const unsigned long long m = __ballot(near) & gmask;
                    if (near) nbr[(size_t)k * cap + cn + __popcll(m & below)] = ent;
                    cn += __popcll(m);
                }
            }
            if (sub == 0) cnt[k] = (uint16_t)(split ? cn : c);
        }
    }
    // only the owned candidates are visited, in increasing index: below k
    // those of the other parity (a = (k + 1) / 2 of them), above k those of
    // the same parity -- owns(k, n) of every n, without the half of the
    // lanes that would test and drop the others
    const int mmax = N / 2 + 1;
    for (int k0 = 0; k0 < (reuse ? 0 : N); k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < N;
        const int tk = act ? typ[k] : 0;
        const float4 pk = act ? f4[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int a_lo = (k + 1) >> 1;
        int c = 0, cn = 0;
        for (int m0 = 0; m0 < mmax; m0 += TPA) {
            const int mi = m0 + sub;
            const int n = mi < a_lo ? ((k + 1) & 1) + 2 * mi : k + 2 + 2 * (mi - a_lo);
            bool in = false, near = false;
            int code = 0;
            if (act && n < N) {
                const float4 pn = f4[n];
                if (all_on || T.pair_on[tk * MDT_MAXT + __float_as_int(pn.w)]) {
                    const float e0 = pk.x - pn.x, e1 = pk.y - pn.y, e2 = pk.z - pn.z;
                    float m0f = rintf(e0), m1f = rintf(e1), m2f = rintf(e2);
                    float d0 = e0 - m0f, d1 = e1 - m1f, d2 = e2 - m2f;
                    // within 1e-4 of a half-cell tie: the reference's image, from the unfolded positions
                    const bool t0 = fabsf(d0) > 0.4999f, t1 = fabsf(d1) > 0.4999f, t2 = fabsf(d2) > 0.4999f;
                    if (t0 | t1 | t2) {
                        const double *xr = bv.x + roff;
                        const double a0 = xr[k], a1 = xr[N + k], a2 = xr[2 * N + k];
                        const double b0 = xr[n], b1 = xr[N + n], b2 = xr[2 * N + n];
                        if (t0) m0f = (float)tie_image(cell.ih[0], cell.ih[3], cell.ih[6], a0, a1, a2, b0, b1, b2, sx[k] - sx[n]);
                        if (t1) m1f = (float)tie_image(cell.ih[1], cell.ih[4], cell.ih[7], a0, a1, a2, b0, b1, b2, sy[k] - sy[n]);
                        if (t2) m2f = (float)tie_image(cell.ih[2], cell.ih[5], cell.ih[8], a0, a1, a2, b0, b1, b2, sz[k] - sz[n]);
                        d0 = e0 - m0f;
                        d1 = e1 - m1f;
                        d2 = e2 - m2f;
                    }
                    const float r2 = d0 * (g00 * d0 + g01 * d1 + g02 * d2) + d1 * (g11 * d1 + g12 * d2) + g22 * d2 * d2;
                    in = r2 <= cut2f;
                    near = r2 <= near2f;
                    code = ((int)m0f + 1) + 3 * ((int)m1f + 1) + 9 * ((int)m2f + 1);
                }
            }
            const unsigned long long m = __ballot(in) & gmask;
            const uint16_t ent = (uint16_t)(n | (code << MDT_NBR_IDX_BITS));
            if (in) {
                const int pos = c + __popcll(m & below);
                if (pos < cap)
                    (split ? nbr_kept : nbr)[(size_t)k * cap + pos] = ent;  // staged: the whole row to HBM
                else
                    bad_local |= MDT_ST_NBR_OVERFLOW;
            }
            c += __popcll(m);
            if (split) {  // staged: the near entries to LDS
                const bool nin = in && near;
                const unsigned long long mn = __ballot(nin) & gmask;
                const int pos = cn + __popcll(mn & below);
                if (nin && pos < cap) nbr[(size_t)k * cap + pos] = ent;
                cn += __popcll(mn);
            }
        }
        if (act && sub == 0) {
            const int cc = c < cap ? c : cap;
            cnt[k] = (uint16_t)(split ? (cn < cap ? cn : cap) : cc);
            if (keep) kcnt[k] = (uint16_t)cc;
        }
    }
    __syncthreads();
    if (keep && !reuse && lane == 0) bv.rl_valid[blockIdx.x] = 1;  // read by the next launch
    for (int i = lane; i < N; i += nthr) {
        shi[i] = 0ull;
        smid[i] = 0ull;
        slo[i] = 0ull;
    }
    __syncthreads();

    // ---- phase 2: pass A, sigma ----
    // the sigma terms of one pair (k, n) at distance r: sigma_k += own in
    // registers, sigma_n += other into the fixed point
    auto sigma_pair = [&](int tk, int n, double r, double &sig) {
        const int tn = typ[n];
        double own, other48;  // other48: the partner's term times 2^48
        if constexpr (PREC == 1) {
            const float e1 = mdt_exp2f(acutl2f * (float)(r - rcut));
            const float th = mdt_rcpf(1.0f + e1);
            const float fsn = th * mdt_exp2f(-T.f_eta2l2[tn] * (float)(r - T.bs0[tn]));
            const float fsk = (tn == tk) ? fsn : th * mdt_exp2f(-T.f_eta2l2[tk] * (float)(r - T.bs0[tk]));
            own = (double)(T.f_csig[tk * MDT_MAXT + tn] * fsn);
            other48 = (double)(T.f_csig[tn * MDT_MAXT + tk] * fsk) * MDT_W48;
        } else {
            double th;
            if constexpr (PREC == 2) {
                th = (double)mdt_rcpf(1.0f + mdt_exp2f(acutl2f * (float)(r - rcut)));
            } else {
                th = mdt_rcp(1 + mdt_exp(acut * (r - rcut), T.exptab));
            }
            const double fsn = th * mdt_exp(-T.eta2[tn] * (r - T.bs0[tn]), T.exptab);
            const double fsk = (tn == tk) ? fsn : th * mdt_exp(-T.eta2[tk] * (r - T.bs0[tk]), T.exptab);
            own = T.csig[tk * MDT_MAXT + tn] * fsn;
            other48 = T.csig48[tn * MDT_MAXT + tk] * fsk;
        }
        sig += own;
        sig_add_scaled(&shi[n], &smid[n], &slo[n], other48, &bad_local);
    };
    // one listed pair (entry e of atom k's list), within rnear
    auto pair_a = [&](int k, int tk, double xk, double yk, double zk, int e, double &sig) {
        const unsigned q = nbr[(size_t)k * cap + e];
        const int n = q & MDT_NBR_IDX_MASK, code = q >> MDT_NBR_IDX_BITS;
        const double dx = (xk - px[n]) - T.shift[3 * code];
        const double dy = (yk - py[n]) - T.shift[3 * code + 1];
        const double dz = (zk - pz[n]) - T.shift[3 * code + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = r2 * mdt_rsqrt(r2);
        if (r > rnear) return;
        sigma_pair(tk, n, r, sig);
    };
    for (int k0 = 0; k0 < N; k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < N;
        double sig = 0.0;
        if (act) {
            const int tk = typ[k];
            const double xk = px[k], yk = py[k], zk = pz[k];
            const int ck = cnt[k];
            for (int e = sub; e < ck; e += TPA) pair_a(k, tk, xk, yk, zk, e, sig);
        }
        sig = group_sum<TPA>(sig);
        if (act && sub == 0) G[k] = sig;  // own part; the partners' part is in shi/smid/slo
    }
    if (lane == 0) s_anytiny = 0;
    __syncthreads();

    // ---- tiny atoms: sigma~ of the near pairs below MDT_SIG_TINY; then the
    // pairs beyond rnear of every tiny atom, from a minimum-image scan of
    // all atoms (a pair of two tiny atoms by its owner) ----
    for (int k = lane; k < N; k += nthr) {
        const uint8_t ty = (T.owner[typ[k]] && G[k] + sig_value(shi[k], smid[k], slo[k]) < MDT_SIG_TINY) ? 1 : 0;
        tiny[k] = ty;
        if (ty) s_anytiny = 1;  // benign race: every writer stores 1
    }
    __syncthreads();
    const bool anytiny = s_anytiny != 0;
    // the far partners n of tiny atom t: beyond rnear, within the cut-off
    auto far_partner = [&](int t, int n, double d[3], double &r, double &ri) {
        if (n == t || (tiny[n] && !owns(t, n)) || !T.pair_on[typ[t] * MDT_MAXT + typ[n]]) return false;
        const double r2 = pair_image(s_ihf, bv.icell + 9 * traj, T.shift, px, py, pz, bv.x + roff, N, t, n, d);
        ri = mdt_rsqrt(r2);
        r = r2 * ri;
        return r > rnear && r <= cutr;
    };
    // the tiny atoms in index order, 64 flags per ballot (the same mask in
    // every wave, so the loop and its barriers are uniform over the block)
    if (anytiny) {
        for (int t0 = 0; t0 < N; t0 += 64) {
          unsigned long long tm = __ballot(t0 + wl < N && tiny[t0 + wl] != 0);
          while (tm) {
            const int t = t0 + __builtin_ctzll(tm);
            tm &= tm - 1;
            const int tt = typ[t];
            double sig = 0.0;
            for (int n0 = lane; n0 < N; n0 += nthr) {
                int n = n0;
                opaque(n);
                double d[3], r, ri;
                if (far_partner(t, n, d, r, ri)) sigma_pair(tt, n, r, sig);
            }
            sig = block_sum_fixed(sig, s_wsum, lane, nthr);
            if (lane == 0) G[t] += sig;
          }
        }
        __syncthreads();
    }

    // ---- per atom: s, dens, E_coh + vref, G = dE/dsigma ----
    for (int k = lane; k < N; k += nthr) {
        const int tk = typ[k];
        double Gk = 0.0, Ek = 0.0;
        if (T.owner[tk]) {
            double dn;
            emt_atom(T, tk, G[k] + sig_value(shi[k], smid[k], slo[k]), &Gk, &Ek, &dn);
            if (!JUDGE) bv.dens[doff + k] = dn;  // raw EMT density (pes_emt_mod.f90:855,861)
        }
        G[k] = Gk;
        E[k] = Ek;
    }
    __syncthreads();
    for (int i = lane; i < N; i += nthr) {
        fxa[i] = 0ull;
        fya[i] = 0ull;
        fza[i] = 0ull;
    }
    __syncthreads();

    // ---- phase 3: pass B, pair forces and pair energies ----
    // one listed pair: the pair's dE/dr gives +F on the owner (registers) and
    // -F on the partner (fixed point); the owner takes the pair energy
    // the force and energy terms of one pair (k, n), d = r_k - r_n (minimum
    // image), r = |d|, ri = 1 / r
    auto force_pair = [&](int tk, int n, double Gk, double dx, double dy, double dz, double r, double ri,
                          double &fx, double &fy, double &fz, double &ep) {
        const int tn = typ[n];
        const int kn = tk * MDT_MAXT + tn, nk = tn * MDT_MAXT + tk;
        double c;
        if constexpr (PREC == 1) {
            const float e1 = mdt_exp2f(acutl2f * (float)(r - rcut));
            const float th = mdt_rcpf(1.0f + e1);
            const float rt1 = acutf * e1 * th;
            const float drn = (float)(r - T.bs0[tn]);
            const float fsn = th * mdt_exp2f(-T.f_eta2l2[tn] * drn);
            const float fvn = th * mdt_exp2f(-T.f_kdbl2[tn] * drn);
            float fsk = fsn, fvk = fvn;
            if (tn != tk) {
                const float drk = (float)(r - T.bs0[tk]);
                fsk = th * mdt_exp2f(-T.f_eta2l2[tk] * drk);
                fvk = th * mdt_exp2f(-T.f_kdbl2[tk] * drk);
            }
            const float ta = T.f_csig[kn] * ((T.f_etad[kn] + rt1) * fsn);
            const float tb = T.f_csig[nk] * ((T.f_etad[nk] + rt1) * fsk);
            const float tv = 0.5f * (T.f_wv[kn] * ((T.f_kdb[tn] + rt1) * fvn) + T.f_wv[nk] * ((T.f_kdb[tk] + rt1) * fvk));
            c = Gk * (double)ta + G[n] * (double)tb + (double)tv;
            ep += (double)(T.f_wv[kn] * fvn + T.f_wv[nk] * fvk);
        } else {
            double th, rt1;
            if constexpr (PREC == 2) {
                const float e1 = mdt_exp2f(acutl2f * (float)(r - rcut));
                const float thf = mdt_rcpf(1.0f + e1);
                th = (double)thf;
                rt1 = (double)(acutf * e1 * thf);
            } else {
                const double e1 = mdt_exp(acut * (r - rcut), T.exptab);
                th = mdt_rcp(1 + e1);
                rt1 = acut * e1 * th;
            }
            const double fsn = th * mdt_exp(-T.eta2[tn] * (r - T.bs0[tn]), T.exptab);
            const double fvn = th * mdt_exp(-T.kdb[tn] * (r - T.bs0[tn]), T.exptab);
            // dE/dr of this pair: density part through G, pair (V) part directly
            if (tn == tk) {  // one element: one set of pair constants (kn == nk)
                c = T.csig[kn] * ((T.etad[kn] + rt1) * fsn) * (Gk + G[n]) + T.wv[kn] * ((T.kdb[tn] + rt1) * fvn);
                ep += 2.0 * T.wv[kn] * fvn;
            } else {
                const double fsk = th * mdt_exp(-T.eta2[tk] * (r - T.bs0[tk]), T.exptab);
                const double fvk = th * mdt_exp(-T.kdb[tk] * (r - T.bs0[tk]), T.exptab);
                c = Gk * T.csig[kn] * ((T.etad[kn] + rt1) * fsn) + G[n] * T.csig[nk] * ((T.etad[nk] + rt1) * fsk) +
                    0.5 * (T.wv[kn] * ((T.kdb[tn] + rt1) * fvn) + T.wv[nk] * ((T.kdb[tk] + rt1) * fvk));
                ep += T.wv[kn] * fvn + T.wv[nk] * fvk;
            }
        }
        // the pair force in fixed-point units (2^42, an exact scaling): the
        // owner's sums fx, fy, fz are kept in these units too
        const double crs = c * ri * MDT_F_SCALE;
        const double gx = crs * dx, gy = crs * dy, gz = crs * dz;
        fx = fma(-crs, dx, fx);
        fy = fma(-crs, dy, fy);
        fz = fma(-crs, dz, fz);
        fx_add_scaled(&fxa[n], gx, &bad_local);
        fx_add_scaled(&fya[n], gy, &bad_local);
        fx_add_scaled(&fza[n], gz, &bad_local);
    };
    // one listed pair (entry e of atom k's list), within rnear
    auto pair_b = [&](int k, int tk, double xk, double yk, double zk, double Gk, int e, double &fx, double &fy,
                      double &fz, double &ep) {
        const unsigned q = nbr[(size_t)k * cap + e];
        const int n = q & MDT_NBR_IDX_MASK, code = q >> MDT_NBR_IDX_BITS;
        const double dx = (xk - px[n]) - T.shift[3 * code];
        const double dy = (yk - py[n]) - T.shift[3 * code + 1];
        const double dz = (zk - pz[n]) - T.shift[3 * code + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ri = mdt_rsqrt(r2);
        const double r = r2 * ri;
        if (r > rnear) return;
        force_pair(tk, n, Gk, dx, dy, dz, r, ri, fx, fy, fz, ep);
    };
    for (int k0 = 0; k0 < N; k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < N;
        double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
        if (act) {
            const int tk = typ[k];
            const double xk = px[k], yk = py[k], zk = pz[k];
            const double Gk = G[k];
            const int ck = cnt[k];
            for (int e = sub; e < ck; e += TPA) pair_b(k, tk, xk, yk, zk, Gk, e, fx, fy, fz, ep);
        }
        fx = group_sum<TPA>(fx);
        fy = group_sum<TPA>(fy);
        fz = group_sum<TPA>(fz);
        ep = group_sum<TPA>(ep);
        if (act && sub == 0) {
            fx_add_scaled(&fxa[k], fx, &bad_local);
            fx_add_scaled(&fya[k], fy, &bad_local);
            fx_add_scaled(&fza[k], fz, &bad_local);
            E[k] -= 0.5 * ep;
        }
    }
    __syncthreads();

    // ---- the pairs beyond rnear of the tiny atoms (as in pass A) ----
    if (anytiny) {
        for (int t0 = 0; t0 < N; t0 += 64) {
          unsigned long long tm = __ballot(t0 + wl < N && tiny[t0 + wl] != 0);
          while (tm) {
            const int t = t0 + __builtin_ctzll(tm);
            tm &= tm - 1;
            const int tt = typ[t];
            const double Gt = G[t];
            double v[4] = {0.0, 0.0, 0.0, 0.0};  // fx fy fz (fixed-point units), ep
            for (int n0 = lane; n0 < N; n0 += nthr) {
                int n = n0;
                opaque(n);
                double d[3], r, ri;
                if (far_partner(t, n, d, r, ri)) force_pair(tt, n, Gt, d[0], d[1], d[2], r, ri, v[0], v[1], v[2], v[3]);
            }
            block_sum4_fixed(v, s_wsum, lane, nthr);
            if (lane == 0) {
                fx_add_scaled(&fxa[t], v[0], &bad_local);
                fx_add_scaled(&fya[t], v[1], &bad_local);
                fx_add_scaled(&fza[t], v[2], &bad_local);
                E[t] -= 0.5 * v[3];
            }
          }
        }
        __syncthreads();
    }

    // ---- judge mode: compare with the given forces, write nothing else ----
    if constexpr (JUDGE) {
        double sq = 0.0, mx = 0.0;
        for (int k = lane; k < N; k += nthr) {
            const int tk = typ[k];
            for (int cc = 0; cc < 3; ++cc) {
                unsigned long long *acc = cc == 0 ? fxa : (cc == 1 ? fya : fza);
                const double F = T.owner[tk] ? (double)(long long)acc[k] * MDT_F_ISCALE : 0.0;
                const double d = F - bv.judge_f[roff + (long long)cc * N + k];
                sq += d * d;
                mx = fabs(d) > mx ? fabs(d) : mx;
            }
        }
        // fixed-order block reduction through red[] (wave partials, then wave 0)
        for (int o = 32; o > 0; o >>= 1) {
            sq += __shfl_xor(sq, o, 64);
            const double m2 = __shfl_xor(mx, o, 64);
            mx = m2 > mx ? m2 : mx;
        }
        __shared__ double jr[32];
        if ((lane & 63) == 0) {
            jr[2 * (lane >> 6)] = sq;
            jr[2 * (lane >> 6) + 1] = mx;
        }
        __syncthreads();
        if (lane == 0) {
            double s2 = 0.0, m2 = 0.0;
            for (int w = 0; w < (nthr + 63) / 64; ++w) {
                s2 += jr[2 * w];
                m2 = jr[2 * w + 1] > m2 ? jr[2 * w + 1] : m2;
            }
            bv.judge_acc[3 * rep] += s2;
            bv.judge_acc[3 * rep + 1] += 3.0 * N;
            bv.judge_acc[3 * rep + 2] = m2 > bv.judge_acc[3 * rep + 2] ? m2 : bv.judge_acc[3 * rep + 2];
        }
        if (bad_local) atomicOr(&bv.status[t], bad_local);
        return;
    }

    // ---- phase 4: set_acceleration (force.f90:89-98), outputs ----
    for (int k = lane; k < N; k += nthr) {
        const int tk = typ[k];
        double F[3] = {(double)(long long)fxa[k] * MDT_F_ISCALE, (double)(long long)fya[k] * MDT_F_ISCALE,
                       (double)(long long)fza[k] * MDT_F_ISCALE};
        if (!T.owner[tk]) F[0] = F[1] = F[2] = 0.0;
        if (add_fin) {
            F[0] += bv.f[roff + k];
            F[1] += bv.f[roff + N + k];
            F[2] += bv.f[roff + 2 * N + k];
        }
        const uint8_t fb = bv.fixed[doff + k];
        const double mass = bv.mass[aoff + k];
        const double a0 = (fb & 1) ? 0.0 : F[0] / mass;
        const double a1 = (fb & 2) ? 0.0 : F[1] / mass;
        const double a2 = (fb & 4) ? 0.0 : F[2] / mass;
        bv.f[roff + k] = F[0];
        bv.f[roff + N + k] = F[1];
        bv.f[roff + 2 * N + k] = F[2];
        bv.a[roff + k] = a0;
        bv.a[roff + N + k] = a1;
        bv.a[roff + 2 * N + k] = a2;
    }
    // ---- energy of the replica: fixed-order reduction by wave 0 ----
    if (lane < 64) {
        double e = 0.0;
        for (int i = lane; i < N; i += 64) e += E[i];
        for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
        if (lane == 0) bv.epot[bv.traj_eoff[t] + b] = e;
    }
    if (bad_local) atomicOr(&bv.status[t], bad_local);
}

// ---------------------------------------------------------------------------
// P > 1 trajectory kernel: [propagate_2 of the previous step] then
// [propagate_1 + COM removal] then [ring-polymer free step] for one
// trajectory per workgroup.  cjk (P x P, column-major cjk(j,k)) in global.
// collect_trajectory_characteristics after propagate_2 of step step_prev
// (src/trajectory_info.f90:47-263) for trajectory t: centroids, the
// projectile (atom 0) against every other atom (min image).  Reads x, v and
// f of the projectile and x of every atom from global memory (written
// before the caller's barrier); every lane of the block calls it.  Returns
// (block-uniform) whether the trajectory left the box and is frozen from
// this step.
__device__ bool traj_observe(const PropConst &pc, const int *s_proj, const BatchView &bv, int t, int N, int P,
                             long long soff, long long aoff, int lane, int nthr, uint32_t step_prev,
                             double *s_rmin, int *s_done) {
#pragma clang fp contract(off)
#define OX(c, b, i) (soff + ((long long)(b) * 3 + (c)) * N + (i))
        // centroids, the projectile (atom 0) against every other atom (min image)
        double cx = 0.0, cy = 0.0, cz = 0.0, vz = 0.0;
        int fnz = 0, above = 1, below = 1;
        for (int b = 0; b < P; ++b) {
            cx += bv.x[OX(0, b, 0)];
            cy += bv.x[OX(1, b, 0)];
            cz += bv.x[OX(2, b, 0)];
            vz += bv.v[OX(2, b, 0)];
            if (bv.f[OX(0, b, 0)] != 0.0 || bv.f[OX(1, b, 0)] != 0.0 || bv.f[OX(2, b, 0)] != 0.0) fnz = 1;
            if (!(bv.x[OX(2, b, 0)] > bv.obs_ul)) above = 0;
            if (!(bv.x[OX(2, b, 0)] < bv.obs_ll)) below = 0;
        }
        cx /= P;
        cy /= P;
        cz /= P;
        vz /= P;
        double h[9], ih[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            h[k] = bv.cell[9 * t + k];
            ih[k] = bv.icell[9 * t + k];
        }
        double rmin = 1.0e300;
        for (int i = 1 + lane; i < N; i += nthr) {
            const int ty = bv.type[aoff + i];
        const int pb = (pc.mixed && !s_proj[ty]) ? 1 : P;
            double ax = 0.0, ay = 0.0, az = 0.0;
            for (int b = 0; b < pb; ++b) {
                ax += bv.x[OX(0, b, i)];
                ay += bv.x[OX(1, b, i)];
                az += bv.x[OX(2, b, i)];
            }
            const double d0 = cx - ax / pb, d1 = cy - ay / pb, d2 = cz - az / pb;
            double s0 = ih[0] * d0 + ih[3] * d1 + ih[6] * d2;
            double s1 = ih[1] * d0 + ih[4] * d1 + ih[7] * d2;
            double s2 = ih[2] * d0 + ih[5] * d1 + ih[8] * d2;
            s0 -= round(s0);
            s1 -= round(s1);
            s2 -= round(s2);
            const double e0 = h[0] * s0 + h[3] * s1 + h[6] * s2;
            const double e1 = h[1] * s0 + h[4] * s1 + h[7] * s2;
            const double e2 = h[2] * s0 + h[5] * s1 + h[8] * s2;
            const double r = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            rmin = r < rmin ? r : rmin;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_xor(rmin, o, 64);
            rmin = y < rmin ? y : rmin;
        }
        if ((lane & 63) == 0) s_rmin[lane >> 6] = rmin;
        __syncthreads();
        if (lane == 0) {
            for (int w = 1; w < (nthr + 63) / 64; ++w) rmin = s_rmin[w] < rmin ? s_rmin[w] : rmin;
            TrajObs &o = bv.obs[t];
            const int istep = (int)(step_prev - bv.obs_step0);
            if (rmin < o.cl_appr) {
                o.cl_appr = rmin;
                o.cl_appr_step = istep;
            }
            o.cl_last = rmin;
            o.last_step = istep;
            o.adsorbed_prev = o.adsorbed;
            if (!o.adsorbed && rmin < bv.ads_start) o.adsorbed = 1;
            if (o.adsorbed && rmin > bv.ads_end) o.adsorbed = 0;
            if (cz < o.r_min[2]) {
                o.r_min[0] = cx;
                o.r_min[1] = cy;
                o.r_min[2] = cz;
            }
            if (fnz) o.t_int += pc.dt;
            const int down = vz < 0.0;
            if (down != o.moving_down) {
                ++o.turning_points;
                o.moving_down = down;
            }
            *s_done = bv.obs_exit && (above || below);
            if (*s_done) {
                o.exit_step = istep;
                bv.done[t] = 1;
            }
        }
        __syncthreads();
    return *s_done != 0;
#undef OX
}

// The propagator updates shared by k_traj and k_traj_classical, in the
// reference's operation order with FMA contraction off (the pragma's flag
// stays on the instructions after inlining): both kernels then give
// bitwise the same states, whatever code surrounds the updates.
// langevin_2 (md_algo.f90:253-285) and langevin_1 (:165-198):
__device__ __forceinline__ double lan_w(const LanCoef &L) {
#pragma clang fp contract(off)
    return L.sv * sqrt(1 - L.crv * L.crv);
}
__device__ __forceinline__ double lan_v2(const LanCoef &L, double w, double v, double a, double eta) {
#pragma clang fp contract(off)
    return v + L.c2 * a + w * eta;
}
__device__ __forceinline__ double lan_x1(const LanCoef &L, double dt, double r, double v, double a, double eta) {
#pragma clang fp contract(off)
    return r + L.c1 * v + L.c2 * dt * a + L.sr * eta;
}
__device__ __forceinline__ double lan_v1(const LanCoef &L, double v, double a, double eta) {
#pragma clang fp contract(off)
    return L.c0 * v + (L.c1 - L.c2) * a + L.sv * L.crv * eta;
}
// velocity Verlet half step and drift (md_algo.f90:108-131, 226-245)
__device__ __forceinline__ double verlet_v(double v, double dt, double a) {
#pragma clang fp contract(off)
    return v + 0.5 * dt * a;
}
__device__ __forceinline__ double verlet_x(double x, double dt, double vn) {
#pragma clang fp contract(off)
    return x + dt * vn;
}

// flags bit0: propagate_2, bit1: propagate_1, bit2: ring-polymer step
// (trajectories with P == 1 only get f = 0 from bit2, as calc_force zeroes f).
// step_base (graph replays): when non-null, *step_base is added to both step
// numbers, which are then offsets within the captured chunk.
// Mixed bead mode (pc.mixed): atoms of non-projectile types are classical --
// one bead, kept in bead 0 (the other bead slots are not read): their
// Langevin / Andersen temperature is kB T / m, PILE acts on their single
// mode, positions advance in propagate_1 and the ring-polymer step skips
// them; in the COM sum a classical atom counts as P identical beads.
// ---------------------------------------------------------------------------
// Four waves per SIMD: the register allocation is held to 128 VGPRs (a
// 16-byte spill); measured 66 -> 58 us per launch on 4096 H/Au 6x6x4
// trajectories against the unbounded 133-VGPR build at three waves.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_traj(const PropConst *__restrict__ pcg, BatchView bv,
                                              const double *cjk_all, const long long *cjk_off,
                                              const uint32_t *step_base, uint32_t step_prev, uint32_t step,
                                              int flags) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_algo[MDT_MAXT], s_proj[MDT_MAXT];
    __shared__ double red[4];
    const PropConst &pc = *pcg;
    const int t = blockIdx.x;
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int N = bv.traj_N[t];
    const int P = bv.traj_P[t];
    const long long soff = bv.traj_soff[t];
    const long long poff = soff / 3;
    const long long aoff = bv.traj_aoff[t];
    const uint32_t tid = bv.traj_id[t];
    const double tsurf = bv.tsurf[t], tproj = bv.tproj[t];
    const double dt = pc.dt;
    const double *cjk = cjk_all + cjk_off[t];
    const int nthr = blockDim.x, lane = threadIdx.x;
    if (step_base) {
        const uint32_t s0 = *step_base;
        step_prev += s0;
        step += s0;
    }
    // LDS: C (P x P) | ring-polymer atom lists: PILE, transformed (not
    // Langevin), Langevin (spring force) | big flags
    double *C = (double *)smem_raw;
    uint16_t *l_pile = (uint16_t *)(C + P * P), *l_tr = l_pile + N, *l_lan = l_tr + N;
    uint8_t *big = (uint8_t *)(l_lan + N);
    __shared__ int s_cnt[3];
    for (int k = lane; k < P * P; k += nthr) C[k] = cjk[k];
    for (int k = lane; k < MDT_MAXT; k += nthr) {
        s_algo[k] = pc.algo[k];
        s_proj[k] = pc.is_proj[k];
    }
    if (lane < 3) s_cnt[lane] = 0;
    int bad_local = 0;
#define XI(c, b, i) (soff + ((long long)(b) * 3 + (c)) * N + (i))
#define PI_(b, i) (poff + (long long)(b) * N + (i))
    __syncthreads();
    // beads of atom type ty: P, or 1 for a classical atom in mixed mode
#define PT(ty) ((pc.mixed && !s_proj[ty]) ? 1 : P)
    // atom-bead slots in use: every atom's P beads, or in the mixed model
    // (atom 0 the only ring polymer) P beads of atom 0 and bead 0 of the rest
    const bool mix = pc.mixed && P > 1;
    const int nab = mix ? P + N - 1 : N * P;
#define AB(q, i, b)                       \
    if (mix) {                            \
        i = (q) < P ? 0 : (q) - P + 1;    \
        b = (q) < P ? (q) : 0;            \
    } else {                              \
        i = (q) % N;                      \
        b = (q) / N;                      \
    }
    __shared__ double s_rmin[4];
    __shared__ int s_done;
    // ring-polymer atoms (P > 1 beads) by propagator; list order is free
    // (every task below is independent of the others)
    if (P > 1)
        for (int i = lane; i < N; i += nthr) {
            const int ty = bv.type[aoff + i];
            if (PT(ty) == 1) continue;
            const int a = s_algo[ty];
            if (a == MDT_PROP_PILE) l_pile[atomicAdd(&s_cnt[0], 1)] = (uint16_t)i;
            if (a == MDT_PROP_LANGEVIN)
                l_lan[atomicAdd(&s_cnt[2], 1)] = (uint16_t)i;
            else
                l_tr[atomicAdd(&s_cnt[1], 1)] = (uint16_t)i;
        }
    __syncthreads();
    const int n_pile = s_cnt[0], n_tr = s_cnt[1], n_lan = s_cnt[2];
    // langevin_2 of the previous step and langevin_1 of this one read the
    // same friction (dens after ldfa), temperature and branch, so their
    // coefficients are equal: with one atom-bead slot per lane (nab <= nthr)
    // the lane keeps them in LDS from the first for the second
    const bool one_slot = nab <= nthr;
    __shared__ LanCoef s_lan[256];
    bool have_l = false;

    if (flags & 1) {
        // ldfa for Langevin atoms, per bead (md_algo.f90:97)
        for (int q = lane; q < nab; q += nthr) {
            int i, b;
            AB(q, i, b);
            const int ty = bv.type[aoff + i];
            if (s_algo[ty] == MDT_PROP_LANGEVIN) {
                int bad = 0;
                bv.dens[PI_(b, i)] = ldfa(bv.dens[PI_(b, i)], bv.mass[aoff + i], &bad);
                if (bad) bad_local |= MDT_ST_NEG_DENSITY;
            }
        }
        __syncthreads();
        for (int i = lane; i < N; i += nthr) {
            int all = 1;
            const int pb = PT(bv.type[aoff + i]);
            for (int b = 0; b < pb; ++b)
                if (!lan_big(bv.dens[PI_(b, i)], dt)) all = 0;
            big[i] = (uint8_t)all;
        }
        __syncthreads();
        for (int q = lane; q < nab; q += nthr) {
            int i, b;
            AB(q, i, b);
            const int ty = bv.type[aoff + i];
            const uint8_t fb = bv.fixed[PI_(b, i)];
            const double mass = bv.mass[aoff + i];
            if (s_algo[ty] == MDT_PROP_LANGEVIN) {
                const double temp = MDT_KB * PT(ty) * (s_proj[ty] ? tproj : tsurf) / mass;
                const bool exact = big[i] && temp > MDT_TOLERANCE;
                const LanCoef L = lan_coef(bv.dens[PI_(b, i)], dt, temp, exact);
                if (one_slot) {
                    s_lan[lane] = L;
                    have_l = true;
                }
                const double w = lan_w(L);
                for (int c = 0; c < 3; ++c) {
                    // the deviate of this step's langevin_1, kept by propagate_1
                    // (mdt_normal(seed, tid, step_prev, MDT_STREAM_LANGEVIN, i P + b, c))
                    const double eta = bv.eta[XI(c, b, i)];
                    const double v = bv.v[XI(c, b, i)];
                    bv.v[XI(c, b, i)] = (fb >> c & 1) ? 0.0 : lan_v2(L, w, v, bv.a[XI(c, b, i)], eta);
                }
            } else {
                for (int c = 0; c < 3; ++c) {
                    const double v = bv.v[XI(c, b, i)];
                    bv.v[XI(c, b, i)] = (fb >> c & 1) ? 0.0 : verlet_v(v, dt, bv.a[XI(c, b, i)]);
                }
            }
        }
        __syncthreads();
    }

    if ((flags & 1) && bv.obs) {
        // frozen after this step: no propagate_1 / ring-polymer step
        if (traj_observe(pc, s_proj, bv, t, N, P, soff, aoff, lane, nthr, step_prev, s_rmin, &s_done)) {
            if (bad_local) atomicOr(&bv.status[t], bad_local);
            return;
        }
    }

    if (flags & 2) {
        // the branch test of langevin_1 over all beads (md_algo.f90:176)
        for (int i = lane; i < N; i += nthr) {
            int all = 1;
            const int pb = PT(bv.type[aoff + i]);
            for (int b = 0; b < pb; ++b)
                if (!lan_big(bv.dens[PI_(b, i)], dt)) all = 0;
            big[i] = (uint8_t)all;
        }
        __syncthreads();
        // pile_thermostat per atom (md_algo.f90:440-502)
        if (pc.any_pile) {
            // one-bead atoms: thread per (atom, component)
            if (P == 1 || mix)
                for (int q = lane; q < N * 3; q += nthr) {
                    const int i = q / 3, c = q % 3;
                    const int ty = bv.type[aoff + i];
                    if (s_algo[ty] != MDT_PROP_PILE || PT(ty) != 1) continue;
                    const double mass = bv.mass[aoff + i];
                    const double betaN = 1.0 / (MDT_KB * (s_proj[ty] ? tproj : tsurf));
                    const double c1 = exp(-0.5 * dt * (1 / pc.pile_tau));
                    const double c2 = sqrt(1 - c1 * c1);
                    const double z = mdt_normal(pc.seed, tid, step, MDT_STREAM_PILE, (uint32_t)(i * P), (uint32_t)c);
                    const double np = c1 * (mass * bv.v[XI(c, 0, i)]) + sqrt(mass / betaN) * c2 * z;
                    const uint8_t fb = bv.fixed[PI_(0, i)];
                    bv.v[XI(c, 0, i)] = (fb >> c & 1) ? 0.0 : np / mass;
                }
            // ring polymers: thread per (atom, component, mode); the thermostatted
            // normal-mode momenta go through f (rewritten by calc_force after this)
            for (int q = lane; q < n_pile * 3 * P; q += nthr) {
                const int k = q % P, c = (q / P) % 3, i = l_pile[q / (3 * P)];
                const int ty = bv.type[aoff + i];
                const double mass = bv.mass[aoff + i];
                const double betaN = 1.0 / (MDT_KB * (s_proj[ty] ? tproj : tsurf) * P);
                double acc = 0.0;
                for (int j = 0; j < P; ++j) acc = acc + (mass * bv.v[XI(c, j, i)]) * C[j + P * k];
                double g;
                if (k == 0) {
                    g = 1 / pc.pile_tau;
                } else {
                    const double wn = 1 / betaN / MDT_HBAR;
                    const double wk = 2 * wn * sin(k * MDT_PI / P);
                    g = 2 * wk;
                }
                const double c1 = exp(-0.5 * dt * g);
                const double c2 = sqrt(1 - c1 * c1);
                const double z = mdt_normal(pc.seed, tid, step, MDT_STREAM_PILE, (uint32_t)(i * P + k), (uint32_t)c);
                bv.f[XI(c, k, i)] = c1 * acc + sqrt(mass / betaN) * c2 * z;
            }
            __syncthreads();
            for (int q = lane; q < n_pile * 3 * P; q += nthr) {
                const int b = q % P, c = (q / P) % 3, i = l_pile[q / (3 * P)];
                const double mass = bv.mass[aoff + i];
                const uint8_t fb = bv.fixed[PI_(b, i)];
                double ap = 0.0;
                if (!(fb >> c & 1))
                    for (int k = 0; k < P; ++k) ap = ap + bv.f[XI(c, k, i)] * C[b + P * k];
                bv.v[XI(c, b, i)] = ap / mass;
            }
            __syncthreads();
        }
        for (int q = lane; q < nab; q += nthr) {
            int i, b;
            AB(q, i, b);
            const int ty = bv.type[aoff + i];
            const int pb = PT(ty);
            const int algo = s_algo[ty];
            const uint8_t fb = bv.fixed[PI_(b, i)];
            const double mass = bv.mass[aoff + i];
            if (algo == MDT_PROP_LANGEVIN) {
                LanCoef L;
                if (have_l) {
                    L = s_lan[lane];
                } else {
                    const double temp = MDT_KB * pb * (s_proj[ty] ? tproj : tsurf) / mass;
                    const bool exact = big[i] && temp > MDT_TOLERANCE;
                    L = lan_coef(bv.dens[PI_(b, i)], dt, temp, exact);
                }
                double z[3];
                mdt_normal3(pc.seed, tid, step, MDT_STREAM_LANGEVIN, (uint32_t)(i * P + b), z);
                for (int c = 0; c < 3; ++c) {
                    const double eta = z[c];
                    bv.eta[XI(c, b, i)] = eta;  // langevin_2 of this step takes the same deviate
                    const double r = bv.x[XI(c, b, i)], v = bv.v[XI(c, b, i)], a = bv.a[XI(c, b, i)];
                    if (!(fb >> c & 1)) {
                        bv.x[XI(c, b, i)] = lan_x1(L, dt, r, v, a, eta);
                        bv.v[XI(c, b, i)] = lan_v1(L, v, a, eta);
                    } else {
                        bv.v[XI(c, b, i)] = 0.0;
                    }
                }
            } else if (algo == MDT_PROP_VERLET || algo == MDT_PROP_ANDERSEN || algo == MDT_PROP_PILE) {
                for (int c = 0; c < 3; ++c) {
                    const double v = bv.v[XI(c, b, i)];
                    const double vn = (fb >> c & 1) ? 0.0 : verlet_v(v, dt, bv.a[XI(c, b, i)]);
                    bv.v[XI(c, b, i)] = vn;
                    if (pb == 1 && !(fb >> c & 1)) bv.x[XI(c, b, i)] = verlet_x(bv.x[XI(c, b, i)], dt, vn);
                }
                if (algo == MDT_PROP_ANDERSEN) {
                    const double ibetaN = pb * MDT_KB * (s_proj[ty] ? tproj : tsurf);
                    const double sigma = sqrt(ibetaN / mass);
                    const double u = mdt_uniform(pc.seed, tid, step, MDT_STREAM_ANDERSEN_U, (uint32_t)(i * P + b), 0u);
                    if (u < dt / pc.andersen_time) {
                        double z[3];
                        mdt_normal3(pc.seed, tid, step, MDT_STREAM_ANDERSEN_V, (uint32_t)(i * P + b), z);
                        for (int c = 0; c < 3; ++c)
                            if (!(fb >> c & 1)) bv.v[XI(c, b, i)] = z[c] * sigma + 0.0;
                    }
                }
            } else {
                bad_local |= MDT_ST_BAD_PROP;
            }
        }
        __syncthreads();
        if ((pc.any_andersen || pc.any_pile) && !pc.no_com) {
            // remove_com_velocity, closed form of the 1000 iterations
            if (lane < 64) {
                double ax = 0, ay = 0, az = 0, am = 0;
                for (int i = lane; i < N; i += 64) {
                    const double m = bv.mass[aoff + i];
                    double bx = 0, by = 0, bz = 0;
                    const int pb = PT(bv.type[aoff + i]);
                    for (int b = 0; b < pb; ++b) {
                        bx += bv.v[XI(0, b, i)];
                        by += bv.v[XI(1, b, i)];
                        bz += bv.v[XI(2, b, i)];
                    }
                    if (pb == 1 && P > 1) {  // a classical atom stands for P identical beads
                        bx *= P;
                        by *= P;
                        bz *= P;
                    }
                    ax += bx * m;
                    ay += by * m;
                    az += bz * m;
                    am += m;
                }
                for (int o = 32; o > 0; o >>= 1) {
                    ax += __shfl_xor(ax, o, 64);
                    ay += __shfl_xor(ay, o, 64);
                    az += __shfl_xor(az, o, 64);
                    am += __shfl_xor(am, o, 64);
                }
                if (lane == 0) {
                    red[0] = ax;
                    red[1] = ay;
                    red[2] = az;
                    red[3] = am;
                }
            }
            __syncthreads();
            const double fac = (1.0 - pow(1.0 - 1.0 / N, 1000.0)) / red[3] / P;
            for (int q = lane; q < nab; q += nthr) {
                int i, b;
                AB(q, i, b);
                bv.v[XI(0, b, i)] -= red[0] * fac;
                bv.v[XI(1, b, i)] -= red[1] * fac;
                bv.v[XI(2, b, i)] -= red[2] * fac;
            }
            __syncthreads();
        }
    }

    if (flags & 4) {
        if (P == 1) {
            for (int q = lane; q < 3 * N; q += nthr) bv.f[soff + q] = 0.0;
        } else {
            // do_ring_polymer_step (rpmd.f90:309-452)
            const double kt = MDT_KB * tproj * P / MDT_HBAR;
            const double m1 = bv.mass[aoff + 0];
            const double prefactor = -1 * m1 * (kt * kt);
            const double piN = MDT_PI / P;
            if (mix)  // classical atoms: no ring polymer, f = 0
                for (int q = lane; q < (N - 1) * 3; q += nthr) bv.f[XI(q % 3, 0, 1 + q / 3)] = 0.0;
            // Langevin atoms: the step is reverted; spring force with m(1) and
            // Tproj (rpmd.f90:418-435)
            for (int q = lane; q < n_lan * 3 * P; q += nthr) {
                const int b = q % P, c = (q / P) % 3, i = l_lan[q / (3 * P)];
                const int pv = b == 0 ? P - 1 : b - 1, nx = b == P - 1 ? 0 : b + 1;
                bv.f[XI(c, b, i)] = prefactor * (2 * bv.x[XI(c, b, i)] - bv.x[XI(c, pv, i)] - bv.x[XI(c, nx, i)]);
            }
            // the other ring polymers: thread per (atom, component, mode) for the
            // forward transform and the exact harmonic rotation, whose results
            // go through a (momenta) and f (positions) -- calc_force zeroes
            // both (force.f90:49-50) -- then per (atom, component, bead) back
            for (int q = lane; q < n_tr * 3 * P; q += nthr) {
                const int k = q % P, c = (q / P) % 3, i = l_tr[q / (3 * P)];
                const int ty = bv.type[aoff + i];
                const double mass = bv.mass[aoff + i];
                double ap = 0.0, aq = 0.0;
                for (int j = 0; j < P; ++j) {
                    ap = ap + (mass * bv.v[XI(c, j, i)]) * C[j + P * k];
                    aq = aq + bv.x[XI(c, j, i)] * C[j + P * k];
                }
                const double betaN = 1.0 / (MDT_KB * (s_proj[ty] ? tproj : tsurf) * P);
                const double twown = 2.0 / betaN / MDT_HBAR;
                double p1 = 1.0, p2 = 0.0, p3 = dt / mass, p4 = 1.0;
                const int fk = (k <= P / 2) ? k : P - k;  // mirrored modes share a frequency
                if (fk > 0) {
                    const double wk = twown * sin(fk * piN);
                    const double wt = wk * dt, wm = wk * mass;
                    const double cw = cos(wt), sw = sin(wt);
                    p1 = cw;
                    p2 = -wm * sw;
                    p3 = sw / wm;
                    p4 = cw;
                }
                bv.a[XI(c, k, i)] = ap * p1 + aq * p2;
                bv.f[XI(c, k, i)] = ap * p3 + aq * p4;
            }
            __syncthreads();
            for (int q = lane; q < n_tr * 3 * P; q += nthr) {
                const int b = q % P, c = (q / P) % 3, i = l_tr[q / (3 * P)];
                const uint8_t fb = bv.fixed[PI_(b, i)];
                if (fb >> c & 1) continue;
                const double mass = bv.mass[aoff + i];
                double rr = 0.0, vv = 0.0;
                for (int k = 0; k < P; ++k) {
                    rr = rr + bv.f[XI(c, k, i)] * C[b + P * k];
                    vv = vv + bv.a[XI(c, k, i)] * C[b + P * k] / mass;
                }
                bv.x[XI(c, b, i)] = rr;
                bv.v[XI(c, b, i)] = vv;
            }
            __syncthreads();
            for (int q = lane; q < n_tr * 3 * P; q += nthr) {
                const int b = q % P, c = (q / P) % 3, i = l_tr[q / (3 * P)];
                bv.f[XI(c, b, i)] = 0.0;
                bv.a[XI(c, b, i)] = 0.0;
            }
        }
    }
#undef XI
#undef PI_
#undef PT
#undef AB
    if (bad_local) atomicOr(&bv.status[t], bad_local);
}

// ---------------------------------------------------------------------------
// k_traj_classical: k_traj for batches of classical trajectories (P = 1)
// with every atom on its own lane and no PILE or Andersen atoms (no COM
// removal): each lane keeps its atom's state in registers from propagate_2
// through propagate_1 and the force reset, reading and writing each array
// once; the only barriers are the observer's.  Same arithmetic, in the same
// order, as k_traj; a kernel of its own so that k_traj's ring-polymer code
// does not set its register budget.
__global__ void __launch_bounds__(256)
k_traj_classical(const PropConst *__restrict__ pcg, BatchView bv, const uint32_t *step_base, uint32_t step_prev,
                 uint32_t step, int flags, int t_base) {
    __shared__ int s_algo[MDT_MAXT], s_proj[MDT_MAXT];
    __shared__ double s_rmin[4];
    __shared__ int s_done;
    const PropConst &pc = *pcg;
    const int t = t_base + (int)blockIdx.x;  // trajectories t_base .. t_base + gridDim.x - 1
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int N = bv.traj_N[t];
    const long long soff = bv.traj_soff[t];
    const long long poff = soff / 3;
    const long long aoff = bv.traj_aoff[t];
    const uint32_t tid = bv.traj_id[t];
    const double tsurf = bv.tsurf[t], tproj = bv.tproj[t];
    const double dt = pc.dt;
    const int nthr = blockDim.x, lane = threadIdx.x;
    if (step_base) {  // graph replays: step numbers are offsets from the device counter
        const uint32_t s0 = *step_base;
        step_prev += s0;
        step += s0;
    }
    for (int k = lane; k < MDT_MAXT; k += nthr) {
        s_algo[k] = pc.algo[k];
        s_proj[k] = pc.is_proj[k];
    }
    int bad_local = 0;
#define XI(c, b, i) (soff + ((long long)(b) * 3 + (c)) * N + (i))
#define PI_(b, i) (poff + (long long)(b) * N + (i))
    __syncthreads();
    {
        const int i = lane;
        const bool act = i < N;
        double x[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0}, a[3] = {0.0, 0.0, 0.0};
        double dn = 0.0, mass = 1.0;
        int ty = 0, algo = MDT_PROP_VERLET;
        uint8_t fb = 0;
        // A fixed component keeps its position, gets v = 0 and uses no
        // deviate: its x, a and eta are not read, its x and eta not written,
        // and its v written only when it was not already 0 (the fixed bottom
        // layers are half the atoms of a slab).
        double v0[3] = {0.0, 0.0, 0.0};
        if (act) {
            ty = bv.type[aoff + i];
            algo = s_algo[ty];
            fb = bv.fixed[PI_(0, i)];
            mass = bv.mass[aoff + i];
            dn = bv.dens[PI_(0, i)];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                v0[c] = v[c] = bv.v[XI(c, 0, i)];
                if (!(fb >> c & 1)) {
                    x[c] = bv.x[XI(c, 0, i)];
                    a[c] = bv.a[XI(c, 0, i)];
                }
            }
        }
        const double temp = MDT_KB * (s_proj[ty] ? tproj : tsurf) / mass;
        LanCoef L;
        bool have = false;
        if ((flags & 1) && act) {  // propagate_2 of step step_prev
            if (algo == MDT_PROP_LANGEVIN) {
                int bad = 0;
                dn = ldfa(dn, mass, &bad);  // md_algo.f90:97
                if (bad) bad_local |= MDT_ST_NEG_DENSITY;
                bv.dens[PI_(0, i)] = dn;
                L = lan_coef(dn, dt, temp, lan_big(dn, dt) && temp > MDT_TOLERANCE);
                have = true;
                const double w = lan_w(L);
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c] = (fb >> c & 1) ? 0.0 : lan_v2(L, w, v[c], a[c], bv.eta[XI(c, 0, i)]);
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c] = (fb >> c & 1) ? 0.0 : verlet_v(v[c], dt, a[c]);
            }
        }
        if ((flags & 1) && bv.obs) {
            if (i == 0)
                for (int c = 0; c < 3; ++c) bv.v[XI(c, 0, i)] = v[c];  // the observer reads the projectile's
            __syncthreads();
            if (traj_observe(pc, s_proj, bv, t, N, 1, soff, aoff, lane, nthr, step_prev, s_rmin, &s_done)) {
                // frozen: the state after this propagate_2 is the trajectory's last
                if (act)
                    for (int c = 0; c < 3; ++c)
                        if (!(fb >> c & 1) || v0[c] != 0.0) bv.v[XI(c, 0, i)] = v[c];
                if (bad_local) atomicOr(&bv.status[t], bad_local);
                return;
            }
        }
        if ((flags & 2) && act) {  // propagate_1 of step step
            if (algo == MDT_PROP_LANGEVIN) {
                if (!have) L = lan_coef(dn, dt, temp, lan_big(dn, dt) && temp > MDT_TOLERANCE);
                double z[3];
                mdt_normal3(pc.seed, tid, step, MDT_STREAM_LANGEVIN, (uint32_t)i, z);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double eta = z[c];
                    if (!(fb >> c & 1)) {
                        bv.eta[XI(c, 0, i)] = eta;  // langevin_2 of this step takes the same deviate
                        const double r = x[c], vo = v[c];
                        x[c] = lan_x1(L, dt, r, vo, a[c], eta);
                        v[c] = lan_v1(L, vo, a[c], eta);
                    } else {
                        v[c] = 0.0;
                    }
                }
            } else if (algo == MDT_PROP_VERLET) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double vn = (fb >> c & 1) ? 0.0 : verlet_v(v[c], dt, a[c]);
                    v[c] = vn;
                    if (!(fb >> c & 1)) x[c] = verlet_x(x[c], dt, vn);
                }
            } else {
                bad_local |= MDT_ST_BAD_PROP;
            }
        }
        if (act) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool fixc = fb >> c & 1;
                if ((flags & 3) && (!fixc || v0[c] != 0.0)) bv.v[XI(c, 0, i)] = v[c];
                if ((flags & 2) && !fixc) bv.x[XI(c, 0, i)] = x[c];
                if (flags & 4) bv.f[XI(c, 0, i)] = 0.0;  // calc_force zeroes f (P == 1)
            }
        }
    }
#undef XI
#undef PI_
    if (bad_local) atomicOr(&bv.status[t], bad_local);
}

// ---------------------------------------------------------------------------
// k_mixed: calc_force of a mixed-bead trajectory (one workgroup, 512 lanes):
// atom 0 (the projectile) has P beads, atoms 1..N-1 are classical and kept
// in bead 0.  The potential of bead b is the EMT energy of the classical
// atoms with projectile bead b; the classical atoms feel the bead average
// of their forces (the force on a rigid P-bead copy divided by P).  Because
// the classical-classical pair geometry does not depend on the bead, the
// averaged force needs the pair passes once, with G replaced by its bead
// average, plus the projectile pairs of every bead:
//   1. owned half lists of the classical atoms (HBM), sigma^0 from their
//      pairs (pass A), per atom G^0, E^0, dens^0;
//   2. per bead b, 16 lanes: the classical atoms within the cut-off of the
//      projectile's bead b; its sigma, G_H(b), E_H(b); then per listed atom
//      a: sigma_a(b) = sigma_a^0 + its projectile term, G_a(b), E_a(b), the
//      pair force (+F to bead b, -F/P to atom a) and pair energy, and the
//      bead averages of G_a and dens_a (fixed point);
//   3. the classical pairs (pass B) with the averaged G.
// epot(b) = sum_a (E_a^0 - V pairs / 2) + sum_{a near b} (E_a(b) - E_a^0)
//           + E_H(b) - V_H(b) / 2.
// ---------------------------------------------------------------------------
#define MIX_CAPQ 192  // classical atoms within the cut-off of one projectile bead
#define MIX_G 16      // lanes per bead
#define MIX_BLOCK 512

// MAXB: 512, or 1024 for trajectories large enough that one 512-lane
// workgroup per CU leaves the classical passes short of waves (the extra
// lanes work the classical atoms; the bead groups stay in the first 512).
template <int TPA, int MAXB = 512>
__global__ void __launch_bounds__(MAXB, 2)
    k_mixed(const EmtConst *__restrict__ ecg, const PropConst *__restrict__ pcg, BatchView bv, int add_fin) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ LdsTables T;
    __shared__ double red[1];
    const int t = bv.rep_list ? bv.rep_list[blockIdx.x] : (int)blockIdx.x;  // trajectory list
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int N = bv.traj_N[t], P = bv.traj_P[t];
    const int Ns = N - 1;
    const int cap = bv.nbcap < Ns - 1 ? bv.nbcap : (Ns > 1 ? Ns - 1 : 1);
    const long long soff = bv.traj_soff[t];
    const long long poff = soff / 3;
    const long long aoff = bv.traj_aoff[t];
    const int nthr = blockDim.x, lane = threadIdx.x;
    CellRegs cell;
    load_cell(bv, t, cell);
    load_tables(T, ecg, pcg, cell, lane, nthr);

    // LDS: px py pz | A0 A1 A2 | G0 E0 D0 S0 | Gacc Dacc (classical atoms,
    // local index a = atom - 1) | per bead hx hy hz Eb | per
    // bead group GH EH DH | cnt cntf | group hcnt, hlist | typ (all atoms) |
    // tiny (classical atoms).  The words A0-A2 hold the fractional
    // coordinates in the scan, the sigma fixed point in pass A and the force
    // fixed point afterwards.  A classical atom's list holds its near entries
    // (within the near radius) from the front (cnt) and its far entries
    // (between the near radius and the cut-off) from the back (cntf); the far
    // entries are evaluated only for pairs with a tiny atom, one whose sigma
    // from its near pairs is below MDT_SIG_TINY.
    const int ngrp = MIX_BLOCK / MIX_G;
    double *px = (double *)smem_raw, *py = px + Ns, *pz = py + Ns;
    double *sx = pz + Ns, *sy = sx + Ns, *sz = sy + Ns;
    unsigned long long *shi = (unsigned long long *)sx, *smid = (unsigned long long *)sy,
                       *slo = (unsigned long long *)sz;
    unsigned long long *fxa = shi, *fya = smid, *fza = slo;
    double *G0 = sz + Ns, *E0 = G0 + Ns, *D0 = E0 + Ns, *S0 = D0 + Ns;
    unsigned long long *Gacc = (unsigned long long *)(S0 + Ns), *Dacc = Gacc + Ns;
    double *hx = (double *)(Dacc + Ns), *hy = hx + P, *hz = hy + P, *Eb = hz + P;
    double *GH = Eb + P, *EH = GH + ngrp, *DH = EH + ngrp;
    uint16_t *cnt = (uint16_t *)(DH + ngrp);
    uint16_t *cntf = cnt + Ns;
    uint16_t *hcnt = cntf + Ns;
    uint16_t *hlist = hcnt + ngrp;
    uint8_t *typ = (uint8_t *)(hlist + (size_t)ngrp * MIX_CAPQ);
    uint8_t *tiny = typ + N;
    __shared__ int s_anytiny;
    uint16_t *nbr = bv.nbr_global + (size_t)blockIdx.x * bv.nbr_stride;
    int bad_local = 0;
    // Verlet skin for the classical lists: built with cut-off + skin, reused
    // while no classical atom moved more than skin / 2 since the build.  The
    // positions keep the fold of the build, so the lists' image codes stay
    // valid; a cell narrower than 2 (cut-off + skin) rebuilds every launch.
    const long long slot = blockIdx.x;
    double wmin = 1e300;
    for (int q = 0; q < 3; ++q) {
        const double w = 1.0 / sqrt(cell.ih[q] * cell.ih[q] + cell.ih[q + 3] * cell.ih[q + 3] +
                                    cell.ih[q + 6] * cell.ih[q + 6]);
        wmin = w < wmin ? w : wmin;
    }
    const double skin = bv.mix_skin;
    const bool keep = skin > 0.0 && bv.mix_valid && 2.0 * (ecg->cutr + skin) < wmin;
    double *ref = keep ? bv.mix_ref + slot * 3 * bv.mix_ns : nullptr;
    double *fold = keep ? bv.mix_fold + slot * 3 * bv.mix_ns : nullptr;
    __shared__ int s_reuse;
    if (lane == 0) s_reuse = keep && bv.mix_valid[slot];
    __syncthreads();
    if (s_reuse) {  // every thread reads the flag before any clears it
        double dmax = 0.0;
        for (int a = lane; a < Ns; a += nthr) {
            const int i = a + 1;
            const double d0 = bv.x[soff + i] - ref[3 * a], d1 = bv.x[soff + N + i] - ref[3 * a + 1],
                         d2 = bv.x[soff + 2 * N + i] - ref[3 * a + 2];
            const double d2s = d0 * d0 + d1 * d1 + d2 * d2;
            dmax = d2s > dmax ? d2s : dmax;
        }
        __syncthreads();
        if (dmax > 0.25 * skin * skin) s_reuse = 0;  // benign race: every writer stores 0
    }
    __syncthreads();
    const bool reuse = s_reuse != 0;

    // ---- stage: classical atoms (bead 0) and the projectile's beads, folded ----
    for (int i = lane; i < N; i += nthr) typ[i] = (uint8_t)bv.type[aoff + i];
    for (int a = lane; a < Ns; a += nthr) {
        const int i = a + 1;
        const double r[3] = {bv.x[soff + i], bv.x[soff + N + i], bv.x[soff + 2 * N + i]};
        double f0 = cell.ih[0] * r[0] + cell.ih[3] * r[1] + cell.ih[6] * r[2];
        double f1 = cell.ih[1] * r[0] + cell.ih[4] * r[1] + cell.ih[7] * r[2];
        double f2 = cell.ih[2] * r[0] + cell.ih[5] * r[1] + cell.ih[8] * r[2];
        if (reuse) {  // the fold of the list build
            f0 -= fold[3 * a];
            f1 -= fold[3 * a + 1];
            f2 -= fold[3 * a + 2];
        } else {
            const double g0 = floor(f0), g1 = floor(f1), g2 = floor(f2);
            f0 -= g0;
            f1 -= g1;
            f2 -= g2;
            if (keep) {
                fold[3 * a] = g0;
                fold[3 * a + 1] = g1;
                fold[3 * a + 2] = g2;
                ref[3 * a] = r[0];
                ref[3 * a + 1] = r[1];
                ref[3 * a + 2] = r[2];
            }
        }
        sx[a] = f0;
        sy[a] = f1;
        sz[a] = f2;
        px[a] = cell.h[0] * f0 + cell.h[3] * f1 + cell.h[6] * f2;
        py[a] = cell.h[1] * f0 + cell.h[4] * f1 + cell.h[7] * f2;
        pz[a] = cell.h[2] * f0 + cell.h[5] * f1 + cell.h[8] * f2;
    }
    for (int b = lane; b < P; b += nthr) {
        const long long rb = soff + (long long)b * 3 * N;
        const double r[3] = {bv.x[rb], bv.x[rb + N], bv.x[rb + 2 * N]};
        double f0 = cell.ih[0] * r[0] + cell.ih[3] * r[1] + cell.ih[6] * r[2];
        double f1 = cell.ih[1] * r[0] + cell.ih[4] * r[1] + cell.ih[7] * r[2];
        double f2 = cell.ih[2] * r[0] + cell.ih[5] * r[1] + cell.ih[8] * r[2];
        f0 -= floor(f0);
        f1 -= floor(f1);
        f2 -= floor(f2);
        hx[b] = cell.h[0] * f0 + cell.h[3] * f1 + cell.h[6] * f2;
        hy[b] = cell.h[1] * f0 + cell.h[4] * f1 + cell.h[7] * f2;
        hz[b] = cell.h[2] * f0 + cell.h[5] * f1 + cell.h[8] * f2;
    }
    __syncthreads();
    const double acut = ecg->acut, rcut = ecg->rcut, cutr = ecg->cutr, cut2 = ecg->cut2hi;
    const int apb = nthr / TPA;
    const int wl = lane & 63;
    const unsigned long long below = (wl == 0) ? 0ull : ((~0ull) >> (64 - wl));
    const unsigned long long gmask = (TPA >= 64 ? ~0ull : ((1ull << TPA) - 1ull)) << ((wl / TPA) * TPA);
    const int tq = typ[0];

    // ---- 1a: owned half lists of the classical atoms ----
    // FP32 membership scan on float copies of the fractional coordinates (in
    // the G0/E0/D0 words, unused until pass A), r^2 bound widened by 1e-4;
    // both passes re-check r <= cutr in FP64, and half-cell ties take the
    // reference's image (tie_image)
    float *fx32 = (float *)G0, *fy32 = (float *)E0, *fz32 = (float *)D0;
    uint16_t *kcnt = keep ? bv.mix_cnt + slot * 2 * bv.mix_ns : nullptr;
    uint16_t *kcntf = keep ? kcnt + bv.mix_ns : nullptr;
    if (reuse)
        for (int a = lane; a < Ns; a += nthr) {
            cnt[a] = kcnt[a];
            cntf[a] = kcntf[a];
        }
    for (int a = lane; a < Ns; a += nthr) {
        fx32[a] = (float)sx[a];
        fy32[a] = (float)sy[a];
        fz32[a] = (float)sz[a];
    }
    __syncthreads();
    float hf[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) hf[q] = (float)cell.h[q];
    const double cutl = keep ? (ecg->cutr + skin) * (ecg->cutr + skin) : cut2;
    const float cut2f = (float)(cutl * (1.0 + 1e-4));
    const double rn = keep ? ecg->rnear + skin : ecg->rnear;  // near / far split (k_replica)
    const float near2f = (float)(rn * rn * (1.0 + 1e-4));
    const int mmax = Ns / 2 + 1;  // owned candidates only, in increasing index (as k_replica)
    for (int k0 = 0; k0 < (reuse ? 0 : Ns); k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < Ns;
        const int tk = act ? typ[k + 1] : 0;
        const float fxk = act ? fx32[k] : 0.f, fyk = act ? fy32[k] : 0.f, fzk = act ? fz32[k] : 0.f;
        const int a_lo = (k + 1) >> 1;
        int c = 0, cf = 0;
        for (int m0 = 0; m0 < mmax; m0 += TPA) {
            const int mi = m0 + sub;
            const int n = mi < a_lo ? ((k + 1) & 1) + 2 * mi : k + 2 + 2 * (mi - a_lo);
            bool in = false, nr = false;
            int code = 0;
            if (act && n < Ns) {
                const int tn = typ[n + 1];
                if (T.pair_on[tk * MDT_MAXT + tn]) {
                    float d0 = fxk - fx32[n], d1 = fyk - fy32[n], d2 = fzk - fz32[n];
                    float m0 = rintf(d0), m1 = rintf(d1), m2 = rintf(d2);
                    const bool t0 = fabsf(fabsf(d0) - 0.5f) < 1e-4f, t1 = fabsf(fabsf(d1) - 0.5f) < 1e-4f,
                               t2 = fabsf(fabsf(d2) - 0.5f) < 1e-4f;
                    if (t0 | t1 | t2) {  // half-cell tie: the reference's image
                        const double *xr = bv.x + soff;
                        const double a0 = xr[k + 1], a1 = xr[N + k + 1], a2 = xr[2 * N + k + 1];
                        const double b0 = xr[n + 1], b1 = xr[N + n + 1], b2 = xr[2 * N + n + 1];
                        if (t0) m0 = (float)tie_image(cell.ih[0], cell.ih[3], cell.ih[6], a0, a1, a2, b0, b1, b2, sx[k] - sx[n]);
                        if (t1) m1 = (float)tie_image(cell.ih[1], cell.ih[4], cell.ih[7], a0, a1, a2, b0, b1, b2, sy[k] - sy[n]);
                        if (t2) m2 = (float)tie_image(cell.ih[2], cell.ih[5], cell.ih[8], a0, a1, a2, b0, b1, b2, sz[k] - sz[n]);
                    }
                    d0 -= m0;
                    d1 -= m1;
                    d2 -= m2;
                    const float ex = hf[0] * d0 + hf[3] * d1 + hf[6] * d2;
                    const float ey = hf[1] * d0 + hf[4] * d1 + hf[7] * d2;
                    const float ez = hf[2] * d0 + hf[5] * d1 + hf[8] * d2;
                    const float rr2 = ex * ex + ey * ey + ez * ez;
                    in = rr2 <= cut2f;
                    nr = rr2 <= near2f;
                    code = ((int)m0 + 1) + 3 * ((int)m1 + 1) + 9 * ((int)m2 + 1);
                }
            }
            const unsigned long long mall = __ballot(in) & gmask;
            const unsigned long long mn = __ballot(in && nr) & gmask;
            const unsigned long long mf = mall & ~mn;
            if (in) {  // near entries from the front, far entries from the back
                const int pos = nr ? c + __popcll(mn & below) : cf + __popcll(mf & below);
                if (pos < cap)
                    nbr[(size_t)k * cap + (nr ? pos : cap - 1 - pos)] = (uint16_t)(n | (code << MDT_NBR_IDX_BITS));
            }
            c += __popcll(mn);
            cf += __popcll(mf);
        }
        if (act && sub == 0) {
            if (c + cf > cap) bad_local |= MDT_ST_NBR_OVERFLOW;
            cnt[k] = (uint16_t)(c < cap ? c : cap);
            cntf[k] = (uint16_t)(cf < cap - cnt[k] ? cf : cap - cnt[k]);
            if (keep) {
                kcnt[k] = cnt[k];
                kcntf[k] = cntf[k];
            }
        }
    }
    __syncthreads();
    if (keep && !reuse && lane == 0) bv.mix_valid[slot] = 1;  // read by the next launch
    for (int a = lane; a < Ns; a += nthr) shi[a] = smid[a] = slo[a] = 0ull;
    __syncthreads();

    // ---- 1b: sigma^0 of the classical atoms (pass A over their pairs) ----
    // near entries; then, for the pairs with a tiny atom (sigma^0 from the
    // near pairs below MDT_SIG_TINY), the far entries too
    auto pair_a = [&](int k, int tk, double xk, double yk, double zk, int e, bool far, double &sig) {
        const unsigned q = nbr[(size_t)k * cap + e];
        const int n = q & MDT_NBR_IDX_MASK, code = q >> MDT_NBR_IDX_BITS;
        if (far && !(tiny[k] | tiny[n])) return;
        const int tn = typ[n + 1];
        const double dx = (xk - px[n]) - T.shift[3 * code];
        const double dy = (yk - py[n]) - T.shift[3 * code + 1];
        const double dz = (zk - pz[n]) - T.shift[3 * code + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = r2 * mdt_rsqrt(r2);
        if (r > cutr) return;
        const double e1 = mdt_exp(acut * (r - rcut), T.exptab);
        const double th = mdt_rcp(1 + e1);
        const double fsn = th * mdt_exp(-T.eta2[tn] * (r - T.bs0[tn]), T.exptab);
        const double fsk = (tn == tk) ? fsn : th * mdt_exp(-T.eta2[tk] * (r - T.bs0[tk]), T.exptab);
        sig += T.csig[tk * MDT_MAXT + tn] * fsn;
        sig_add_scaled(&shi[n], &smid[n], &slo[n], T.csig48[tn * MDT_MAXT + tk] * fsk, &bad_local);
    };
    for (int k0 = 0; k0 < Ns; k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < Ns;
        double sig = 0.0;
        if (act) {
            const int tk = typ[k + 1];
            const double xk = px[k], yk = py[k], zk = pz[k];
            const int ck = cnt[k];
            for (int e = sub; e < ck; e += TPA) pair_a(k, tk, xk, yk, zk, e, false, sig);
        }
        sig = group_sum<TPA>(sig);
        if (act && sub == 0) G0[k] = sig;
    }
    if (lane == 0) s_anytiny = 0;
    __syncthreads();
    for (int a = lane; a < Ns; a += nthr) {
        const uint8_t ty = (T.owner[typ[a + 1]] && G0[a] + sig_value(shi[a], smid[a], slo[a]) < MDT_SIG_TINY) ? 1 : 0;
        tiny[a] = ty;
        if (ty) s_anytiny = 1;  // benign race: every writer stores 1
    }
    __syncthreads();
    const bool anytiny = s_anytiny != 0;
    if (anytiny) {
        for (int k0 = 0; k0 < Ns; k0 += apb) {
            const int k = k0 + lane / TPA;
            const int sub = lane % TPA;
            const bool act = k < Ns;
            double sig = 0.0;
            if (act) {
                const int tk = typ[k + 1];
                const double xk = px[k], yk = py[k], zk = pz[k];
                for (int e = cap - cntf[k] + sub; e < cap; e += TPA) pair_a(k, tk, xk, yk, zk, e, true, sig);
            }
            sig = group_sum<TPA>(sig);
            if (act && sub == 0) G0[k] += sig;
        }
        __syncthreads();
    }
    for (int a = lane; a < Ns; a += nthr) {
        const int ta = typ[a + 1];
        double g = 0.0, e = 0.0, dn = 0.0;
        const double s0 = G0[a] + sig_value(shi[a], smid[a], slo[a]);
        if (T.owner[ta]) emt_atom(T, ta, s0, &g, &e, &dn);
        S0[a] = s0;
        G0[a] = g;
        E0[a] = e;
        D0[a] = dn;
        Gacc[a] = Dacc[a] = 0ull;
    }
    __syncthreads();
    for (int a = lane; a < Ns; a += nthr) fxa[a] = fya[a] = fza[a] = 0ull;  // force words from here on
    __syncthreads();

    // ---- 2: the projectile's beads, MIX_G lanes each ----
    const int grp = lane / MIX_G, gsub = lane % MIX_G;
    const unsigned long long qmask = ((1ull << MIX_G) - 1ull) << ((wl / MIX_G) * MIX_G);
    const double invP = 1.0 / P;
    const int kn = tq * MDT_MAXT, kq = tq;
    for (int b0 = 0; b0 < P; b0 += ngrp) {
        const int b = b0 + grp;
        const bool bact = grp < ngrp && b < P;
        // 2a: classical atoms within the cut-off of bead b
        {
            int c = 0;
            const double xq = bact ? hx[b] : 0.0, yq = bact ? hy[b] : 0.0, zq = bact ? hz[b] : 0.0;
            for (int n0 = 0; n0 < Ns; n0 += MIX_G) {
                const int n = n0 + gsub;
                bool in = false;
                int code = 0;
                if (bact && n < Ns && T.pair_on[kn + typ[n + 1]]) {
                    // fractional difference from the folded cartesian positions (the
                    // fractional words hold sigma / force sums by now)
                    const double cx = xq - px[n], cy = yq - py[n], cz = zq - pz[n];
                    double d0 = cell.ih[0] * cx + cell.ih[3] * cy + cell.ih[6] * cz;
                    double d1 = cell.ih[1] * cx + cell.ih[4] * cy + cell.ih[7] * cz;
                    double d2 = cell.ih[2] * cx + cell.ih[5] * cy + cell.ih[8] * cz;
                    double m0 = rint(d0), m1 = rint(d1), m2 = rint(d2);
                    const bool t0 = fabs(fabs(d0) - 0.5) < 1e-4, t1 = fabs(fabs(d1) - 0.5) < 1e-4,
                               t2 = fabs(fabs(d2) - 0.5) < 1e-4;
                    if (t0 | t1 | t2) {  // half-cell tie: the reference's image (slab in bead 0)
                        const double *xh = bv.x + soff + (long long)b * 3 * N, *xs = bv.x + soff;
                        const double a0 = xh[0], a1 = xh[N], a2 = xh[2 * N];
                        const double b0 = xs[n + 1], b1 = xs[N + n + 1], b2 = xs[2 * N + n + 1];
                        if (t0) m0 = tie_image(cell.ih[0], cell.ih[3], cell.ih[6], a0, a1, a2, b0, b1, b2, d0);
                        if (t1) m1 = tie_image(cell.ih[1], cell.ih[4], cell.ih[7], a0, a1, a2, b0, b1, b2, d1);
                        if (t2) m2 = tie_image(cell.ih[2], cell.ih[5], cell.ih[8], a0, a1, a2, b0, b1, b2, d2);
                    }
                    d0 -= m0;
                    d1 -= m1;
                    d2 -= m2;
                    const double ex = cell.h[0] * d0 + cell.h[3] * d1 + cell.h[6] * d2;
                    const double ey = cell.h[1] * d0 + cell.h[4] * d1 + cell.h[7] * d2;
                    const double ez = cell.h[2] * d0 + cell.h[5] * d1 + cell.h[8] * d2;
                    in = (ex * ex + ey * ey + ez * ez) <= cut2;
                    code = ((int)m0 + 1) + 3 * ((int)m1 + 1) + 9 * ((int)m2 + 1);
                }
                const unsigned long long m = __ballot(in) & qmask;
                if (in) {
                    const int pos = c + __popcll(m & below);
                    if (pos < MIX_CAPQ)
                        hlist[(size_t)grp * MIX_CAPQ + pos] = (uint16_t)(n | (code << MDT_NBR_IDX_BITS));
                    else
                        bad_local |= MDT_ST_NBR_OVERFLOW;
                }
                c += __popcll(m);
            }
            if (gsub == 0 && grp < ngrp) hcnt[grp] = (uint16_t)(c < MIX_CAPQ ? c : MIX_CAPQ);
        }
        __syncthreads();
        // 2b: sigma of the projectile's bead b
        {
            double sig = 0.0;
            const int ch = bact ? hcnt[grp] : 0;
            for (int e = gsub; e < ch; e += MIX_G) {
                const unsigned q = hlist[(size_t)grp * MIX_CAPQ + e];
                const int n = q & MDT_NBR_IDX_MASK, code = q >> MDT_NBR_IDX_BITS;
                const int tn = typ[n + 1];
                const double dx = (hx[b] - px[n]) - T.shift[3 * code];
                const double dy = (hy[b] - py[n]) - T.shift[3 * code + 1];
                const double dz = (hz[b] - pz[n]) - T.shift[3 * code + 2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double r = r2 * mdt_rsqrt(r2);
                if (r > cutr) continue;
                const double e1 = mdt_exp(acut * (r - rcut), T.exptab);
                const double th = mdt_rcp(1 + e1);
                sig += T.csig[kn + tn] * (th * mdt_exp(-T.eta2[tn] * (r - T.bs0[tn]), T.exptab));
            }
            sig = group_sum<MIX_G>(sig);
            if (bact && gsub == 0) {
                double g = 0.0, e = 0.0, dn = 0.0;
                if (T.owner[tq]) emt_atom(T, tq, sig, &g, &e, &dn);
                GH[grp] = g;
                EH[grp] = e;
                DH[grp] = dn;
            }
        }
        __syncthreads();
        // 2c: the listed atoms' sigma(b), G(b), E(b); pair forces and energies
        {
            double fx = 0.0, fy = 0.0, fz = 0.0, dE = 0.0, vs = 0.0;
            const int ch = bact ? hcnt[grp] : 0;
            const double Gq = bact ? GH[grp] : 0.0;
            for (int e = gsub; e < ch; e += MIX_G) {
                const unsigned q = hlist[(size_t)grp * MIX_CAPQ + e];
                const int n = q & MDT_NBR_IDX_MASK, code = q >> MDT_NBR_IDX_BITS;
                const int tn = typ[n + 1];
                const double dx = (hx[b] - px[n]) - T.shift[3 * code];
                const double dy = (hy[b] - py[n]) - T.shift[3 * code + 1];
                const double dz = (hz[b] - pz[n]) - T.shift[3 * code + 2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double ri = mdt_rsqrt(r2);
                const double r = r2 * ri;
                if (r > cutr) continue;
                const double e1 = mdt_exp(acut * (r - rcut), T.exptab);
                const double th = mdt_rcp(1 + e1);
                const double rt1 = acut * e1 * th;
                const double fsn = th * mdt_exp(-T.eta2[tn] * (r - T.bs0[tn]), T.exptab);
                const double fvn = th * mdt_exp(-T.kdb[tn] * (r - T.bs0[tn]), T.exptab);
                double fsk = fsn, fvk = fvn;
                if (tn != tq) {
                    fsk = th * mdt_exp(-T.eta2[kq] * (r - T.bs0[kq]), T.exptab);
                    fvk = th * mdt_exp(-T.kdb[kq] * (r - T.bs0[kq]), T.exptab);
                }
                const int nk = tn * MDT_MAXT + tq;
                // the classical atom with this bead's projectile term
                double Gn = 0.0, En = 0.0, Dn = 0.0;
                if (T.owner[tn]) emt_atom(T, tn, S0[n] + T.csig[nk] * fsk, &Gn, &En, &Dn);
                dE += En - E0[n];
                fx_add(&Gacc[n], (Gn - G0[n]) * invP, MDT_F_SCALE, &bad_local);
                fx_add(&Dacc[n], (Dn - D0[n]) * invP, MDT_F_SCALE, &bad_local);
                const double c = Gq * T.csig[kn + tn] * ((T.etad[kn + tn] + rt1) * fsn) +
                                 Gn * T.csig[nk] * ((T.etad[nk] + rt1) * fsk) +
                                 0.5 * (T.wv[kn + tn] * ((T.kdb[tn] + rt1) * fvn) + T.wv[nk] * ((T.kdb[kq] + rt1) * fvk));
                vs += T.wv[kn + tn] * fvn + T.wv[nk] * fvk;
                const double cr = c * ri;
                fx -= cr * dx;
                fy -= cr * dy;
                fz -= cr * dz;
                fx_add(&fxa[n], cr * dx * invP, MDT_F_SCALE, &bad_local);
                fx_add(&fya[n], cr * dy * invP, MDT_F_SCALE, &bad_local);
                fx_add(&fza[n], cr * dz * invP, MDT_F_SCALE, &bad_local);
            }
            fx = group_sum<MIX_G>(fx);
            fy = group_sum<MIX_G>(fy);
            fz = group_sum<MIX_G>(fz);
            dE = group_sum<MIX_G>(dE);
            vs = group_sum<MIX_G>(vs);
            if (bact && gsub == 0) {
                const long long rb = soff + (long long)b * 3 * N;
                double F[3] = {T.owner[tq] ? fx : 0.0, T.owner[tq] ? fy : 0.0, T.owner[tq] ? fz : 0.0};
                if (add_fin) {
                    F[0] += bv.f[rb];
                    F[1] += bv.f[rb + N];
                    F[2] += bv.f[rb + 2 * N];
                }
                const uint8_t fb = bv.fixed[poff + (long long)b * N];
                const double mass = bv.mass[aoff];
                bv.f[rb] = F[0];
                bv.f[rb + N] = F[1];
                bv.f[rb + 2 * N] = F[2];
                bv.a[rb] = (fb & 1) ? 0.0 : F[0] / mass;
                bv.a[rb + N] = (fb & 2) ? 0.0 : F[1] / mass;
                bv.a[rb + 2 * N] = (fb & 4) ? 0.0 : F[2] / mass;
                bv.dens[poff + (long long)b * N] = DH[grp];
                Eb[b] = dE + EH[grp] - 0.5 * vs;  // bead part of the energy
            }
        }
        __syncthreads();
    }

    // ---- 3: classical pairs with the bead-averaged G (pass B) ----
    for (int a = lane; a < Ns; a += nthr)
        G0[a] += (double)(long long)Gacc[a] * MDT_F_ISCALE;  // G0 now holds the bead average
    __syncthreads();
    auto pair_b = [&](int k, int tk, double xk, double yk, double zk, double Gk, int e, bool far, double &fx,
                      double &fy, double &fz, double &ep) {
        const unsigned q = nbr[(size_t)k * cap + e];
        const int n = q & MDT_NBR_IDX_MASK, code = q >> MDT_NBR_IDX_BITS;
        if (far && !(tiny[k] | tiny[n])) return;
        const int tn = typ[n + 1];
        const double dx = (xk - px[n]) - T.shift[3 * code];
        const double dy = (yk - py[n]) - T.shift[3 * code + 1];
        const double dz = (zk - pz[n]) - T.shift[3 * code + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ri = mdt_rsqrt(r2);
        const double r = r2 * ri;
        if (r > cutr) return;
        const int ikn = tk * MDT_MAXT + tn, ink = tn * MDT_MAXT + tk;
        const double e1 = mdt_exp(acut * (r - rcut), T.exptab);
        const double th = mdt_rcp(1 + e1);
        const double rt1 = acut * e1 * th;
        const double fsn = th * mdt_exp(-T.eta2[tn] * (r - T.bs0[tn]), T.exptab);
        const double fvn = th * mdt_exp(-T.kdb[tn] * (r - T.bs0[tn]), T.exptab);
        double c;
        if (tn == tk) {  // one element: one set of pair constants (as in k_replica)
            c = T.csig[ikn] * ((T.etad[ikn] + rt1) * fsn) * (Gk + G0[n]) + T.wv[ikn] * ((T.kdb[tn] + rt1) * fvn);
            ep += 2.0 * T.wv[ikn] * fvn;
        } else {
            const double fsk = th * mdt_exp(-T.eta2[tk] * (r - T.bs0[tk]), T.exptab);
            const double fvk = th * mdt_exp(-T.kdb[tk] * (r - T.bs0[tk]), T.exptab);
            c = Gk * T.csig[ikn] * ((T.etad[ikn] + rt1) * fsn) + G0[n] * T.csig[ink] * ((T.etad[ink] + rt1) * fsk) +
                0.5 * (T.wv[ikn] * ((T.kdb[tn] + rt1) * fvn) + T.wv[ink] * ((T.kdb[tk] + rt1) * fvk));
            ep += T.wv[ikn] * fvn + T.wv[ink] * fvk;
        }
        const double crs = c * ri * MDT_F_SCALE;  // fixed-point units, as in k_replica
        const double gx = crs * dx, gy = crs * dy, gz = crs * dz;
        fx = fma(-crs, dx, fx);
        fy = fma(-crs, dy, fy);
        fz = fma(-crs, dz, fz);
        fx_add_scaled(&fxa[n], gx, &bad_local);
        fx_add_scaled(&fya[n], gy, &bad_local);
        fx_add_scaled(&fza[n], gz, &bad_local);
    };
    for (int k0 = 0; k0 < Ns; k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < Ns;
        double fx = 0.0, fy = 0.0, fz = 0.0, ep = 0.0;
        if (act) {
            const int tk = typ[k + 1];
            const double xk = px[k], yk = py[k], zk = pz[k];
            const double Gk = G0[k];
            const int ck = cnt[k];
            for (int e = sub; e < ck; e += TPA) pair_b(k, tk, xk, yk, zk, Gk, e, false, fx, fy, fz, ep);
            if (anytiny)
                for (int e = cap - cntf[k] + sub; e < cap; e += TPA)
                    pair_b(k, tk, xk, yk, zk, Gk, e, true, fx, fy, fz, ep);
        }
        fx = group_sum<TPA>(fx);
        fy = group_sum<TPA>(fy);
        fz = group_sum<TPA>(fz);
        ep = group_sum<TPA>(ep);
        if (act && sub == 0) {
            fx_add_scaled(&fxa[k], fx, &bad_local);
            fx_add_scaled(&fya[k], fy, &bad_local);
            fx_add_scaled(&fza[k], fz, &bad_local);
            E0[k] -= 0.5 * ep;
        }
    }
    __syncthreads();

    // ---- outputs: classical atoms (bead 0), bead energies ----
    for (int a = lane; a < Ns; a += nthr) {
        const int i = a + 1;
        const int ta = typ[i];
        double F[3] = {(double)(long long)fxa[a] * MDT_F_ISCALE, (double)(long long)fya[a] * MDT_F_ISCALE,
                       (double)(long long)fza[a] * MDT_F_ISCALE};
        if (!T.owner[ta]) F[0] = F[1] = F[2] = 0.0;
        if (add_fin) {
            F[0] += bv.f[soff + i];
            F[1] += bv.f[soff + N + i];
            F[2] += bv.f[soff + 2 * N + i];
        }
        const uint8_t fb = bv.fixed[poff + i];
        const double mass = bv.mass[aoff + i];
        bv.f[soff + i] = F[0];
        bv.f[soff + N + i] = F[1];
        bv.f[soff + 2 * N + i] = F[2];
        bv.a[soff + i] = (fb & 1) ? 0.0 : F[0] / mass;
        bv.a[soff + N + i] = (fb & 2) ? 0.0 : F[1] / mass;
        bv.a[soff + 2 * N + i] = (fb & 4) ? 0.0 : F[2] / mass;
        if (T.owner[ta]) bv.dens[poff + i] = D0[a] + (double)(long long)Dacc[a] * MDT_F_ISCALE;
    }
    if (lane < 64) {
        double e = 0.0;
        for (int a = lane; a < Ns; a += 64) e += E0[a];
        for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
        if (lane == 0) red[0] = e;
    }
    __syncthreads();
    for (int b = lane; b < P; b += nthr) bv.epot[bv.traj_eoff[t] + b] = red[0] + Eb[b];
    if (bad_local) atomicOr(&bv.status[t], bad_local);
}

// ---------------------------------------------------------------------------
// host-side launchers (C++ linkage, used by mdt_api.hip)
// ---------------------------------------------------------------------------
size_t mdt_replica_smem(int N, int cap, bool nbr_lds) {
    const int c = cap < N - 1 ? cap : (N > 1 ? N - 1 : 1);
    // px py pz | A0 A1 A2 | G | E: 8 N doubles; cnt, lists; typ, tiny
    return sizeof(double) * (8 * (size_t)N) + sizeof(uint16_t) * ((size_t)N + (nbr_lds ? (size_t)N * c : 0)) +
           2 * (size_t)N + 16;
}

template <bool L, int F, bool J>
static hipError_t launch_replica_t(int tpa, dim3 g, dim3 blk, size_t smem, const EmtConst *ec, const PropConst *pc,
                                   const BatchView &bv, int add_fin, hipStream_t st) {
    if (blk.x > 512) {  // 1024-lane instantiations: one or two threads per atom
        switch (tpa) {
        case 1: hipLaunchKernelGGL((k_replica<1, L, F, J, 1024>), g, blk, smem, st, ec, pc, bv, add_fin); break;
        case 2: hipLaunchKernelGGL((k_replica<2, L, F, J, 1024>), g, blk, smem, st, ec, pc, bv, add_fin); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (tpa) {
    case 1: hipLaunchKernelGGL((k_replica<1, L, F, J>), g, blk, smem, st, ec, pc, bv, add_fin); break;
    case 2: hipLaunchKernelGGL((k_replica<2, L, F, J>), g, blk, smem, st, ec, pc, bv, add_fin); break;
    case 4: hipLaunchKernelGGL((k_replica<4, L, F, J>), g, blk, smem, st, ec, pc, bv, add_fin); break;
    case 8: hipLaunchKernelGGL((k_replica<8, L, F, J>), g, blk, smem, st, ec, pc, bv, add_fin); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// calc_force on every replica; ec and pc are device pointers.  The judge
// (bv.judge_acc set) always runs the FP64 arithmetic.
hipError_t mdt_launch_replica(int tpa, int nrep, int block, size_t smem, const EmtConst *ec, const PropConst *pc,
                              const BatchView &bv, int add_fin, hipStream_t st) {
    dim3 g(nrep), blk(block);
    const bool lds = bv.nbr_global == nullptr || bv.rl_stage;  // kept lists staged in LDS: the LDS instantiation
    if (bv.judge_acc)
        return lds ? launch_replica_t<true, 0, true>(tpa, g, blk, smem, ec, pc, bv, add_fin, st)
                   : launch_replica_t<false, 0, true>(tpa, g, blk, smem, ec, pc, bv, add_fin, st);
    if (bv.precision == 1)
        return lds ? launch_replica_t<true, 1, false>(tpa, g, blk, smem, ec, pc, bv, add_fin, st)
                   : launch_replica_t<false, 1, false>(tpa, g, blk, smem, ec, pc, bv, add_fin, st);
    if (bv.precision == 2)
        return lds ? launch_replica_t<true, 2, false>(tpa, g, blk, smem, ec, pc, bv, add_fin, st)
                   : launch_replica_t<false, 2, false>(tpa, g, blk, smem, ec, pc, bv, add_fin, st);
    return lds ? launch_replica_t<true, 0, false>(tpa, g, blk, smem, ec, pc, bv, add_fin, st)
               : launch_replica_t<false, 0, false>(tpa, g, blk, smem, ec, pc, bv, add_fin, st);
}

// LDS of k_mixed for trajectories of up to nmax atoms (one projectile) and
// pmax beads
size_t mdt_mixed_smem(int nmax, int pmax) {
    const size_t ns = (size_t)(nmax - 1), ngrp = MIX_BLOCK / MIX_G;
    return sizeof(double) * (12 * ns + 4 * (size_t)pmax + 3 * ngrp) +
           sizeof(uint16_t) * (2 * ns + ngrp + ngrp * MIX_CAPQ) + (size_t)nmax + ns + 16;
}

int mdt_mixed_block(void) { return MIX_BLOCK; }

// calc_force of the mixed-bead trajectories listed in bv.rep_list (ntraj of
// them); the classical atoms' owned half lists go to bv.nbr_global +
// list position x bv.nbr_stride
// block: MIX_BLOCK (512) or 1024 lanes
hipError_t mdt_launch_mixed(int tpa, int block, int ntraj, size_t smem, const EmtConst *ec, const PropConst *pc,
                            const BatchView &bv, int add_fin, hipStream_t st) {
    dim3 g(ntraj), blk(block);
    if (block == 1024) {
        switch (tpa) {
        case 2: hipLaunchKernelGGL((k_mixed<2, 1024>), g, blk, smem, st, ec, pc, bv, add_fin); break;
        case 4: hipLaunchKernelGGL((k_mixed<4, 1024>), g, blk, smem, st, ec, pc, bv, add_fin); break;
        case 8: hipLaunchKernelGGL((k_mixed<8, 1024>), g, blk, smem, st, ec, pc, bv, add_fin); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (block != MIX_BLOCK) return hipErrorInvalidValue;
    switch (tpa) {
    case 2: hipLaunchKernelGGL((k_mixed<2>), g, blk, smem, st, ec, pc, bv, add_fin); break;
    case 4: hipLaunchKernelGGL((k_mixed<4>), g, blk, smem, st, ec, pc, bv, add_fin); break;
    case 8: hipLaunchKernelGGL((k_mixed<8>), g, blk, smem, st, ec, pc, bv, add_fin); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// classical: every trajectory has P = 1 and at most `block` atoms, and no
// atom type uses PILE or Andersen (k_traj_classical)
hipError_t mdt_launch_traj(int ntraj, size_t smem, const PropConst *pc, const BatchView &bv, const double *cjk,
                           const long long *cjk_off, const uint32_t *step_base, uint32_t step_prev, uint32_t step,
                           int flags, int block, bool classical, hipStream_t st, int t_base) {
    if (classical)
        hipLaunchKernelGGL(k_traj_classical, dim3(ntraj), dim3(block), 0, st, pc, bv, step_base, step_prev, step,
                           flags, t_base);
    else
        hipLaunchKernelGGL(k_traj, dim3(ntraj), dim3(block), smem, st, pc, bv, cjk, cjk_off, step_base, step_prev,
                           step, flags);
    return hipGetLastError();
}

// the step counter of graph replays: one lane sets it (set != 0) or adds
// the chunk length
__global__ void k_advance(uint32_t *step_base, uint32_t n, int set) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *step_base = set ? n : *step_base + n;
}

hipError_t mdt_launch_advance(uint32_t *step_base, uint32_t n, int set, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, step_base, n, set);
    return hipGetLastError();
}
