// mdt_observe.hip -- per-trajectory estimators of the current batch state.
//
// k_estimators computes, for every trajectory, the quantities of the
// reference's energy file (output_nrg, src/output_mod.f90:170-229) from the
// state in HBM: atom (centroid) and bead temperatures
// (universe_mod.f90:314-327, rpmd.f90:63-77), radius of gyration
// (rpmd.f90:511-545), atom / bead / virial quantum kinetic energies
// (universe_mod.f90:433-458, rpmd.f90:547-572, 273-301), the ring-polymer
// spring energy calc_bead_epot (rpmd.f90:236-264, with the plain bead
// differences of calc_inter_bead_distances, :212-233), the bead-averaged
// potential energy, the total energy, sum(f), and the projectile centroid
// position and velocity.  One 256-thread block per trajectory; each
// thread sums its atoms in FP64, a wave-shuffle + LDS tree reduces.
// In mixed bead mode the classical atoms (bead 0 only) count as P
// identical beads, i.e. the replicated state the host interface returns.
// Bound: one read of x, v, f (3PN doubles each) and mass per trajectory --
// HBM-bound, far below a step's cost.
#include <hip/hip_runtime.h>

#include "mdt_device.h"

#define MDT_KB 8.61733238496e-5         // eV / K (constants.f90:37)
#define MDT_HBAR 0.6582119514467406     // eV fs (constants.f90:38)

namespace {

constexpr int NSUM = 14;  // partial sums reduced across the block

__device__ __forceinline__ double block_sum(double v, double *red, int lane, int nthr) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((lane & 63) == 0) red[lane >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (nthr + 63) / 64; ++w) s += red[w];
    return s;
}

}  // namespace

__global__ void __launch_bounds__(256) k_estimators(const PropConst *__restrict__ pcg, BatchView bv,
                                                    double *__restrict__ out) {
    __shared__ double red[4];
    __shared__ int s_proj[MDT_MAXT];
    const PropConst &pc = *pcg;
    const int t = blockIdx.x;
    const int N = bv.traj_N[t];
    const int P = bv.traj_P[t];
    const long long soff = bv.traj_soff[t];
    const long long poff = soff / 3;
    const long long aoff = bv.traj_aoff[t];
    const int nthr = blockDim.x, lane = threadIdx.x;
    const bool mix = pc.mixed && P > 1;
    for (int k = lane; k < MDT_MAXT; k += nthr) s_proj[k] = pc.is_proj[k];
    __syncthreads();
#define XI(c, b, i) (soff + ((long long)(b) * 3 + (c)) * N + (i))
    // s: 0 atom T, 1 bead T, 2 rgyr proj, 3 rgyr latt, 4 nproj, 5/6 atom ekin
    // p/l, 7/8 bead ekin p/l, 9/10 virial p/l, 11 spring, 12 sum f, 13 fixed
    double s[NSUM];
#pragma unroll
    for (int k = 0; k < NSUM; ++k) s[k] = 0.0;
    const double tproj = bv.tproj[t], tsurf = bv.tsurf[t];
    for (int i = lane; i < N; i += nthr) {
        const int ty = bv.type[aoff + i];
        const bool pr = s_proj[ty] != 0;
        const double m = bv.mass[aoff + i];
        const bool cl = mix && i > 0;  // classical atom: bead 0 stands for every bead
        double cx[3], vs[3];
        for (int c = 0; c < 3; ++c) {
            double xa = 0.0, va = 0.0;
            for (int b = 0; b < P; ++b) {
                const int sb = cl ? 0 : b;
                xa += bv.x[XI(c, sb, i)];
                va += bv.v[XI(c, sb, i)];
            }
            cx[c] = P > 1 ? xa / P : xa;
            vs[c] = va;
        }
        const double vc2 = vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2];
        double vb2 = 0.0, rg = 0.0, vir = 0.0, dx2 = 0.0, fsum = 0.0, nfix = 0.0;
        for (int b = 0; b < P; ++b) {
            const int sb = cl ? 0 : b, sn = cl ? 0 : (b + 1 == P ? 0 : b + 1);
            const uint8_t fb = bv.fixed[poff + (long long)sb * N + i];
            nfix += (double)((fb & 1) + (fb >> 1 & 1) + (fb >> 2 & 1));
            for (int c = 0; c < 3; ++c) {
                const double x = bv.x[XI(c, sb, i)], v = bv.v[XI(c, sb, i)], f = bv.f[XI(c, sb, i)];
                vb2 += v * v;
                rg += (x - cx[c]) * (x - cx[c]);
                vir += (cx[c] - x) * f;
                const double d = x - bv.x[XI(c, sn, i)];
                dx2 += d * d;
                fsum += f;
            }
        }
        const double T = pr ? tproj : tsurf;
        const double wn = MDT_KB * T * P / MDT_HBAR;
        s[0] += m * vc2;
        s[1] += m * vb2;
        s[pr ? 2 : 3] += rg;
        s[4] += pr ? 1.0 : 0.0;
        s[pr ? 5 : 6] += m * vc2;
        s[pr ? 7 : 8] += m * vb2;
        s[pr ? 9 : 10] += vir;
        s[11] += m * dx2 * (wn * wn);
        s[12] += fsum;
        s[13] += nfix;
    }
#pragma unroll
    for (int k = 0; k < NSUM; ++k) s[k] = block_sum(s[k], red, lane, nthr);
    if (lane == 0) {
        double *o = out + (long long)MDT_NEST * t;
        const double dof = 3.0 * N * P - s[13];
        const double nproj = s[4];
        double ep = 0.0;
        for (int b = 0; b < P; ++b) ep += bv.epot[bv.traj_eoff[t] + b];
        o[0] = s[0] / MDT_KB / dof / P;                                    // atom T
        o[1] = s[1] / MDT_KB / dof;                                        // bead T
        o[2] = P > 1 ? s[2] / P / (nproj > 1.0 ? nproj : 1.0) : 0.0;       // r_gyr^2 proj
        o[3] = P > 1 ? s[3] / P / (N - nproj > 1.0 ? N - nproj : 1.0) : 0.0;  // r_gyr^2 latt
        o[4] = 0.5 * s[5] / P / P;                                         // atom ekin proj
        o[5] = 0.5 * s[6] / P / P;                                         // atom ekin latt
        o[6] = 0.5 * s[7] / P;                                             // bead ekin proj
        o[7] = 0.5 * s[8] / P;                                             // bead ekin latt
        o[8] = P > 1 ? 0.5 * s[9] / P : 0.0;                               // quantum ekin proj
        o[9] = P > 1 ? 0.5 * s[10] / P : 0.0;                              // quantum ekin latt
        o[10] = ep / P;                                                    // atom epot
        o[11] = P > 1 ? 0.5 * s[11] / P : 0.0;                             // bead epot
        o[12] = o[6] + o[7] + o[10] + o[11];                               // e_total
        o[13] = s[12];                                                     // sum(f)
        for (int c = 0; c < 3; ++c) {                                      // projectile centroid
            double xa = 0.0, va = 0.0;
            for (int b = 0; b < P; ++b) {
                xa += bv.x[XI(c, b, 0)];
                va += bv.v[XI(c, b, 0)];
            }
            o[14 + c] = xa / P;
            o[17 + c] = va / P;
        }
    }
#undef XI
}

hipError_t mdt_launch_estimators(int ntraj, const PropConst *pc, const BatchView &bv, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_estimators, dim3(ntraj), dim3(256), 0, st, pc, bv, out);
    return hipGetLastError();
}
