// mdt_kernels.hip -- CDNA4 (gfx950) kernels for batched EMT forces and the
// md_tian2 propagators.
//
// k_replica: one workgroup per replica (trajectory t, bead b).  The replica's
//   positions are staged in LDS; TPA lanes per atom build a per-atom
//   min-image neighbour list (ordered by neighbour index, compacted with a
//   wavefront ballot), then two gather passes over it:
//     pass A: sigma_a = sum_n csig(ta,tn) f_sigma_tn(r) and the pair (V)
//             energy and force of atom a;
//     per atom: s, dens, E_coh + vref, G_a = dE/dsigma_a;
//     pass B: F_a -= sum_n (G_a c f'_tn + G_n c f'_ta) r_hat.
//   Lane partial sums are combined with wavefront shuffles in a fixed order,
//   so results are bitwise reproducible (no atomics).
//   mode MDT_MODE_FORCE: calc_force only (src/force.f90:40-98).
//   mode MDT_MODE_STEP (P == 1): propagate_1, calc_force and propagate_2
//   fused (src/md_tian2.f90:81-83) with state resident in LDS/registers.
// k_traj: one workgroup per trajectory for P > 1: propagate_2 of the
//   previous step, propagate_1, COM removal and the ring-polymer free step
//   (src/md_algo.f90:37-106, src/rpmd.f90:309-452).
#include <hip/hip_runtime.h>

#include "mdt_device.h"
#include "mdt_philox.h"

#define MDT_MODE_FORCE 0
#define MDT_MODE_STEP 1

namespace {

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

// universe_mod.f90:421-426: direct coordinates, anint, back to cartesian
__device__ __forceinline__ void minimg(const CellRegs &c, double &dx, double &dy, double &dz) {
    double s0 = c.ih[0] * dx + c.ih[3] * dy + c.ih[6] * dz;
    double s1 = c.ih[1] * dx + c.ih[4] * dy + c.ih[7] * dz;
    double s2 = c.ih[2] * dx + c.ih[5] * dy + c.ih[8] * dz;
    s0 -= round(s0);
    s1 -= round(s1);
    s2 -= round(s2);
    dx = c.h[0] * s0 + c.h[3] * s1 + c.h[6] * s2;
    dy = c.h[1] * s0 + c.h[4] * s1 + c.h[7] * s2;
    dz = c.h[2] * s0 + c.h[5] * s1 + c.h[8] * s2;
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
__device__ __forceinline__ bool lan_big(double fric, double dt) {
    return fric * dt > (double)1e-2f;
}

// ldfa, src/md_algo.f90:506-570: friction [1/fs] from the EMT density
__device__ __forceinline__ double ldfa(double dens, double mass, int *bad) {
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

struct SmemLayout {
    double *tab;  // per-type and pair tables
    double *px, *py, *pz, *fx, *fy, *fz, *G, *E;
    uint16_t *cnt, *nbr;
    uint8_t *typ;
};

// table offsets inside tab
#define TB_ETA2 0
#define TB_BS0 (MDT_MAXT)
#define TB_KDB (2 * MDT_MAXT)
#define TB_CSIG (3 * MDT_MAXT)
#define TB_ETAD (3 * MDT_MAXT + MDT_MAXT * MDT_MAXT)
#define TB_WV (3 * MDT_MAXT + 2 * MDT_MAXT * MDT_MAXT)
#define TB_SIZE (3 * MDT_MAXT + 3 * MDT_MAXT * MDT_MAXT)

__device__ __forceinline__ SmemLayout carve(unsigned char *base, int N, int cap) {
    SmemLayout s;
    double *d = (double *)base;
    s.tab = d;
    d += TB_SIZE;
    s.px = d; d += N;
    s.py = d; d += N;
    s.pz = d; d += N;
    s.fx = d; d += N;
    s.fy = d; d += N;
    s.fz = d; d += N;
    s.G = d; d += N;
    s.E = d; d += N;
    uint16_t *h = (uint16_t *)d;
    s.cnt = h; h += N;
    s.nbr = h; h += (size_t)N * cap;
    s.typ = (uint8_t *)h;
    return s;
}

}  // namespace

// Propagate_1 for one atom-bead of a P == 1 trajectory (md_algo.f90:37-70),
// positions and velocities in/out, noise drawn for this (traj, step, atom).
__device__ __forceinline__ void prop1_p1(const PropConst &pc, int algo, bool isproj, double tsurf,
                                         double tproj, double mass, uint8_t fb, double fric,
                                         uint32_t tid, uint32_t step, int i, double r[3], double v[3],
                                         const double acc[3]) {
    const double dt = pc.dt;
    if (algo == MDT_PROP_LANGEVIN) {
        const double temp = MDT_KB * 1 * (isproj ? tproj : tsurf) / mass;
        const bool exact = lan_big(fric, dt) && temp > MDT_TOLERANCE;
        const LanCoef L = lan_coef(fric, dt, temp, exact);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double eta = mdt_normal(pc.seed, tid, step, MDT_STREAM_LANGEVIN, (uint32_t)i, (uint32_t)c);
            if (!(fb >> c & 1)) {
                r[c] = r[c] + L.c1 * v[c] + L.c2 * dt * acc[c] + L.sr * eta;
                v[c] = L.c0 * v[c] + (L.c1 - L.c2) * acc[c] + L.sv * L.crv * eta;
            } else {
                v[c] = 0.0;
            }
        }
        return;
    }
    if (algo == MDT_PROP_PILE) {
        // pile_thermostat with P == 1: C = [1], centroid gamma = 1/tau (md_algo.f90:440-502)
        const double betaN = 1.0 / (MDT_KB * (isproj ? tproj : tsurf) * 1);
        const double g0 = 1 / pc.pile_tau;
        const double c1 = exp(-0.5 * dt * g0);
        const double c2 = sqrt(1 - c1 * c1);
        const double sq = sqrt(mass / betaN);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double z = mdt_normal(pc.seed, tid, step, MDT_STREAM_PILE, (uint32_t)i, (uint32_t)c);
            double np = 0.0 + (mass * v[c]) * 1.0;  // C^T p with C = sqrt(1/1)
            np = c1 * np + sq * c2 * z;
            double ap = 0.0;
            if (!(fb >> c & 1)) ap = ap + np * 1.0;
            v[c] = ap / mass;
        }
    }
    // verlet_1 (md_algo.f90:110-128), also for andersen and after pile
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (!(fb >> c & 1))
            v[c] = v[c] + 0.5 * dt * acc[c];
        else
            v[c] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (!(fb >> c & 1)) r[c] = r[c] + dt * v[c];
    if (algo == MDT_PROP_ANDERSEN) {
        // andersen (md_algo.f90:402-432)
        const double ibetaN = 1 * MDT_KB * (isproj ? tproj : tsurf);
        const double sigma = sqrt(ibetaN / mass);
        const double u = mdt_uniform(pc.seed, tid, step, MDT_STREAM_ANDERSEN_U, (uint32_t)i, 0u);
        if (u < dt / pc.andersen_time) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double z = mdt_normal(pc.seed, tid, step, MDT_STREAM_ANDERSEN_V, (uint32_t)i, (uint32_t)c);
                if (!(fb >> c & 1)) v[c] = z * sigma + 0.0;
            }
        }
    }
}

template <int TPA, bool NBR_LDS>
__global__ void __launch_bounds__(1024) k_replica(EmtConst ec, PropConst pc, BatchView bv, uint32_t step,
                                                  int mode, int add_fin) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int rep = blockIdx.x;
    const int t = bv.rep_traj[rep];
    const int b = bv.rep_bead[rep];
    const int N = bv.traj_N[t];
    const int cap = bv.nbcap < N - 1 ? bv.nbcap : (N > 1 ? N - 1 : 1);
    const long long soff = bv.traj_soff[t];
    const long long roff = soff + (long long)b * 3 * N;  // x block of this replica
    const long long doff = soff / 3 + (long long)b * N;   // dens / fixed block
    const long long aoff = bv.traj_aoff[t];
    const uint32_t tid = bv.traj_id[t];
    const double tsurf = bv.tsurf[t], tproj = bv.tproj[t];
    SmemLayout s = carve(smem_raw, N, NBR_LDS ? cap : 0);
    // neighbour lists: LDS when they fit, else a per-replica slice of global scratch
    uint16_t *nbr = NBR_LDS ? s.nbr : bv.nbr_global + (size_t)rep * bv.nbr_stride;
    CellRegs cell;
    load_cell(bv, t, cell);
    const int nthr = blockDim.x;
    const int lane = threadIdx.x;

    // tables to LDS
    for (int k = lane; k < TB_SIZE; k += nthr) {
        double v = 0.0;
        if (k < TB_BS0) v = ec.eta2[k - TB_ETA2];
        else if (k < TB_KDB) v = ec.betas0[k - TB_BS0];
        else if (k < TB_CSIG) v = ec.kappadbeta[k - TB_KDB];
        else if (k < TB_ETAD) v = ec.csig[k - TB_CSIG];
        else if (k < TB_WV) v = ec.etad[k - TB_ETAD];
        else v = ec.wv[k - TB_WV];
        s.tab[k] = v;
    }
    int bad_local = 0;

    // stage positions (after propagate_1 when fused)
    for (int i = lane; i < N; i += nthr) {
        const int ty = bv.type[aoff + i];
        s.typ[i] = (uint8_t)ty;
        double r[3] = {bv.x[roff + i], bv.x[roff + N + i], bv.x[roff + 2 * N + i]};
        if (mode == MDT_MODE_STEP) {
            double v[3] = {bv.v[roff + i], bv.v[roff + N + i], bv.v[roff + 2 * N + i]};
            const double acc[3] = {bv.a[roff + i], bv.a[roff + N + i], bv.a[roff + 2 * N + i]};
            const uint8_t fb = bv.fixed[doff + i];
            prop1_p1(pc, pc.algo[ty], pc.is_proj[ty] != 0, tsurf, tproj, bv.mass[aoff + i], fb,
                     bv.dens[doff + i], tid, step, i, r, v, acc);
            bv.v[roff + i] = v[0];
            bv.v[roff + N + i] = v[1];
            bv.v[roff + 2 * N + i] = v[2];
            bv.x[roff + i] = r[0];
            bv.x[roff + N + i] = r[1];
            bv.x[roff + 2 * N + i] = r[2];
        }
        s.px[i] = r[0];
        s.py[i] = r[1];
        s.pz[i] = r[2];
    }
    __syncthreads();
    if (mode == MDT_MODE_STEP && (pc.any_andersen || pc.any_pile)) {
        // remove_com_velocity (universe_mod.f90:222-241): 1000 subtractions of
        // S_k/(M P N) with S_k = S_0 (1-1/N)^k, summed in closed form.
        __shared__ double red[4 * 64];
        double sx = 0, sy = 0, sz = 0, sm = 0;
        if (lane < 64) {
            for (int i = lane; i < N; i += 64) {
                const double m = bv.mass[aoff + i];
                sx += bv.v[roff + i] * m;
                sy += bv.v[roff + N + i] * m;
                sz += bv.v[roff + 2 * N + i] * m;
                sm += m;
            }
            for (int o = 32; o > 0; o >>= 1) {
                sx += __shfl_xor(sx, o, 64);
                sy += __shfl_xor(sy, o, 64);
                sz += __shfl_xor(sz, o, 64);
                sm += __shfl_xor(sm, o, 64);
            }
            if (lane == 0) {
                red[0] = sx;
                red[1] = sy;
                red[2] = sz;
                red[3] = sm;
            }
        }
        __syncthreads();
        const double fac = (1.0 - pow(1.0 - 1.0 / N, 1000.0)) / red[3] / 1;
        const double cx = red[0] * fac, cy = red[1] * fac, cz = red[2] * fac;
        for (int i = lane; i < N; i += nthr) {
            bv.v[roff + i] -= cx;
            bv.v[roff + N + i] -= cy;
            bv.v[roff + 2 * N + i] -= cz;
        }
        __syncthreads();
    }

    const double *eta2 = s.tab + TB_ETA2, *bs0 = s.tab + TB_BS0, *kdb = s.tab + TB_KDB;
    const double *csig = s.tab + TB_CSIG, *etad = s.tab + TB_ETAD, *wv = s.tab + TB_WV;
    const double acut = ec.acut, rcut = ec.rcut, cutr = ec.cutr, cut2 = ec.cut2hi;
    const int apb = nthr / TPA;  // atoms per chunk
    const int wl = lane & 63;
    const unsigned long long gmask =
        (TPA >= 64 ? ~0ull : ((1ull << TPA) - 1ull)) << ((wl / TPA) * TPA);
    const unsigned long long below = (wl == 0) ? 0ull : ((~0ull) >> (64 - wl));

    // ---- neighbour list + pass A ----
    for (int k0 = 0; k0 < N; k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < N;
        const int tk = act ? s.typ[k] : 0;
        const double xk = act ? s.px[k] : 0.0, yk = act ? s.py[k] : 0.0, zk = act ? s.pz[k] : 0.0;
        int cnt = 0;
        for (int n0 = 0; n0 < N; n0 += TPA) {
            const int n = n0 + sub;
            bool in = false;
            if (act && n < N && n != k) {
                const int tn = s.typ[n];
                if (ec.pair_on[tk * MDT_MAXT + tn]) {
                    double dx = xk - s.px[n], dy = yk - s.py[n], dz = zk - s.pz[n];
                    minimg(cell, dx, dy, dz);
                    in = (dx * dx + dy * dy + dz * dz) <= cut2;
                }
            }
            const unsigned long long m = __ballot(in) & gmask;
            if (in) {
                const int pos = cnt + __popcll(m & below);
                if (pos < cap)
                    nbr[(size_t)k * cap + pos] = (uint16_t)n;
                else
                    bad_local |= MDT_ST_NBR_OVERFLOW;
            }
            cnt += __popcll(m);
        }
        if (cnt > cap) cnt = cap;
        double sig = 0.0, ev = 0.0, fvx = 0.0, fvy = 0.0, fvz = 0.0;
        if (act) {
            for (int e = sub; e < cnt; e += TPA) {
                const int n = nbr[(size_t)k * cap + e];
                const int tn = s.typ[n];
                double dx = xk - s.px[n], dy = yk - s.py[n], dz = zk - s.pz[n];
                minimg(cell, dx, dy, dz);
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r > cutr) continue;
                const double e1 = exp(acut * (r - rcut));
                const double th = 1.0 / (1 + e1);
                const double rt1 = acut * e1 * th;
                const double fs = th * exp(-eta2[tn] * (r - bs0[tn]));
                const double fv = th * exp(-kdb[tn] * (r - bs0[tn]));
                const int kn = tk * MDT_MAXT + tn;
                sig += csig[kn] * fs;
                ev += wv[kn] * fv;
                double cv;
                if (tn == tk) {
                    cv = wv[kn] * ((kdb[tn] + rt1) * fv);
                } else {
                    const double fv2 = th * exp(-kdb[tk] * (r - bs0[tk]));
                    cv = 0.5 * (wv[kn] * ((kdb[tn] + rt1) * fv) +
                                wv[tn * MDT_MAXT + tk] * ((kdb[tk] + rt1) * fv2));
                }
                const double cr = cv / r;
                fvx -= cr * dx;
                fvy -= cr * dy;
                fvz -= cr * dz;
            }
        }
        sig = group_sum<TPA>(sig);
        ev = group_sum<TPA>(ev);
        fvx = group_sum<TPA>(fvx);
        fvy = group_sum<TPA>(fvy);
        fvz = group_sum<TPA>(fvz);
        if (act && sub == 0) {
            s.cnt[k] = (uint16_t)cnt;
            double G = 0.0, E = 0.0, dn = 0.0;
            if (ec.owner[tk]) {
                const double clamp30 = (double)1e-30f;  // pes_emt_mod.f90:793-794
                const double sc = sig > clamp30 ? sig : clamp30;
                const double sl = sc > MDT_TOLERANCE ? sc : MDT_TOLERANCE;
                const double sv = -log(sl * MDT_TWELFTH) / ec.betaeta2[tk];
                dn = ec.n0[tk] * exp(ec.densc[tk] * sv);
                const double el = exp(-ec.lambda[tk] * sv);
                const double ek = exp(-ec.kappa[tk] * sv);
                E = ec.e0[tk] * ((1 + ec.lambda[tk] * sv) * el - 1) + 6 * ec.v0[tk] * ek - 0.5 * ev;
                const double D = ec.e0[tk] * ec.lambda[tk] * ec.lambda[tk] * (sv * el) +
                                 6 * ec.v0[tk] * ec.kappa[tk] * ek;
                G = -D / (ec.betaeta2[tk] * sc);
                bv.dens[doff + k] = dn;  // raw EMT density (pes_emt_mod.f90:855,861)
            }
            s.G[k] = G;
            s.E[k] = E;
            s.fx[k] = fvx;
            s.fy[k] = fvy;
            s.fz[k] = fvz;
        }
    }
    __syncthreads();

    // ---- pass B: density-derivative forces ----
    for (int k0 = 0; k0 < N; k0 += apb) {
        const int k = k0 + lane / TPA;
        const int sub = lane % TPA;
        const bool act = k < N;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        if (act) {
            const int tk = s.typ[k];
            const double xk = s.px[k], yk = s.py[k], zk = s.pz[k];
            const double Gk = s.G[k];
            const int cnt = s.cnt[k];
            for (int e = sub; e < cnt; e += TPA) {
                const int n = nbr[(size_t)k * cap + e];
                const int tn = s.typ[n];
                double dx = xk - s.px[n], dy = yk - s.py[n], dz = zk - s.pz[n];
                minimg(cell, dx, dy, dz);
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r > cutr) continue;
                const double e1 = exp(acut * (r - rcut));
                const double th = 1.0 / (1 + e1);
                const double rt1 = acut * e1 * th;
                const double fsn = th * exp(-eta2[tn] * (r - bs0[tn]));
                const int kn = tk * MDT_MAXT + tn, nk = tn * MDT_MAXT + tk;
                const double cA = Gk * csig[kn] * ((etad[kn] + rt1) * fsn);
                double cB;
                if (tn == tk) {
                    cB = s.G[n] * csig[nk] * ((etad[nk] + rt1) * fsn);
                } else {
                    const double fsk = th * exp(-eta2[tk] * (r - bs0[tk]));
                    cB = s.G[n] * csig[nk] * ((etad[nk] + rt1) * fsk);
                }
                const double cr = (cA + cB) / r;
                fx -= cr * dx;
                fy -= cr * dy;
                fz -= cr * dz;
            }
        }
        fx = group_sum<TPA>(fx);
        fy = group_sum<TPA>(fy);
        fz = group_sum<TPA>(fz);
        if (act && sub == 0) {
            const int tk = s.typ[k];
            double F[3] = {fx + s.fx[k], fy + s.fy[k], fz + s.fz[k]};
            if (!ec.owner[tk]) F[0] = F[1] = F[2] = 0.0;
            if (add_fin) {
                F[0] += bv.f[roff + k];
                F[1] += bv.f[roff + N + k];
                F[2] += bv.f[roff + 2 * N + k];
            }
            const uint8_t fb = bv.fixed[doff + k];
            const double mass = bv.mass[aoff + k];
            double acc[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] = (fb >> c & 1) ? 0.0 : F[c] / mass;
            bv.f[roff + k] = F[0];
            bv.f[roff + N + k] = F[1];
            bv.f[roff + 2 * N + k] = F[2];
            bv.a[roff + k] = acc[0];
            bv.a[roff + N + k] = acc[1];
            bv.a[roff + 2 * N + k] = acc[2];
            if (mode == MDT_MODE_STEP) {
                // propagate_2 (md_algo.f90:80-106), P == 1
                const int algo = pc.algo[tk];
                double v[3] = {bv.v[roff + k], bv.v[roff + N + k], bv.v[roff + 2 * N + k]};
                if (algo == MDT_PROP_LANGEVIN) {
                    int bad = 0;
                    const double fric = ldfa(bv.dens[doff + k], mass, &bad);
                    if (bad) bad_local |= MDT_ST_NEG_DENSITY;
                    bv.dens[doff + k] = fric;
                    const double temp = MDT_KB * 1 * (pc.is_proj[tk] ? tproj : tsurf) / mass;
                    const bool exact = lan_big(fric, pc.dt) && temp > MDT_TOLERANCE;
                    const LanCoef L = lan_coef(fric, pc.dt, temp, exact);
                    const double w = L.sv * sqrt(1 - L.crv * L.crv);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double eta =
                            mdt_normal(pc.seed, tid, step, MDT_STREAM_LANGEVIN, (uint32_t)k, (uint32_t)c);
                        v[c] = (fb >> c & 1) ? 0.0 : v[c] + L.c2 * acc[c] + w * eta;
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < 3; ++c) v[c] = (fb >> c & 1) ? 0.0 : v[c] + 0.5 * pc.dt * acc[c];
                }
                bv.v[roff + k] = v[0];
                bv.v[roff + N + k] = v[1];
                bv.v[roff + 2 * N + k] = v[2];
            }
        }
    }
    // ---- energy of the replica: fixed-order reduction by wave 0 ----
    __syncthreads();
    if (lane < 64) {
        double e = 0.0;
        for (int i = lane; i < N; i += 64) e += s.E[i];
        for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
        if (lane == 0) bv.epot[bv.traj_eoff[t] + b] = e;
    }
    if (bad_local) atomicOr(&bv.status[t], bad_local);
}

// ---------------------------------------------------------------------------
// P > 1 trajectory kernel: [propagate_2 of the previous step] then
// [propagate_1 + COM removal] then [ring-polymer free step] for one
// trajectory per workgroup.  cjk (P x P, column-major cjk(j,k)) in global.
// flags bit0: propagate_2, bit1: propagate_1, bit2: ring-polymer step
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_traj(PropConst pc, BatchView bv, const double *cjk_all,
                                              const long long *cjk_off, uint32_t step_prev,
                                              uint32_t step, int flags) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int t = blockIdx.x;
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
    double *C = (double *)smem_raw;        // P*P
    uint8_t *big = (uint8_t *)(C + P * P);  // N flags
    __shared__ double red[8];
    for (int k = lane; k < P * P; k += nthr) C[k] = cjk[k];
    int bad_local = 0;
    // index helpers: component c, bead b, atom i
#define XI(c, b, i) (soff + ((long long)(b) * 3 + (c)) * N + (i))
#define PI_(b, i) (poff + (long long)(b) * N + (i))
    __syncthreads();

    if (flags & 1) {
        // ldfa for Langevin atoms, per bead (md_algo.f90:97)
        for (int q = lane; q < N * P; q += nthr) {
            const int i = q % N, b = q / N;
            const int ty = bv.type[aoff + i];
            if (pc.algo[ty] == MDT_PROP_LANGEVIN) {
                int bad = 0;
                bv.dens[PI_(b, i)] = ldfa(bv.dens[PI_(b, i)], bv.mass[aoff + i], &bad);
                if (bad) bad_local |= MDT_ST_NEG_DENSITY;
            }
        }
        __syncthreads();
        for (int i = lane; i < N; i += nthr) {
            int all = 1;
            for (int b = 0; b < P; ++b)
                if (!lan_big(bv.dens[PI_(b, i)], dt)) all = 0;
            big[i] = (uint8_t)all;
        }
        __syncthreads();
        for (int q = lane; q < N * P; q += nthr) {
            const int i = q % N, b = q / N;
            const int ty = bv.type[aoff + i];
            const uint8_t fb = bv.fixed[PI_(b, i)];
            const double mass = bv.mass[aoff + i];
            if (pc.algo[ty] == MDT_PROP_LANGEVIN) {
                const double temp = MDT_KB * P * (pc.is_proj[ty] ? tproj : tsurf) / mass;
                const bool exact = big[i] && temp > MDT_TOLERANCE;
                const LanCoef L = lan_coef(bv.dens[PI_(b, i)], dt, temp, exact);
                const double w = L.sv * sqrt(1 - L.crv * L.crv);
                for (int c = 0; c < 3; ++c) {
                    const double eta = mdt_normal(pc.seed, tid, step_prev, MDT_STREAM_LANGEVIN,
                                                  (uint32_t)(i * P + b), (uint32_t)c);
                    const double v = bv.v[XI(c, b, i)];
                    bv.v[XI(c, b, i)] = (fb >> c & 1) ? 0.0 : v + L.c2 * bv.a[XI(c, b, i)] + w * eta;
                }
            } else {
                for (int c = 0; c < 3; ++c) {
                    const double v = bv.v[XI(c, b, i)];
                    bv.v[XI(c, b, i)] = (fb >> c & 1) ? 0.0 : v + 0.5 * dt * bv.a[XI(c, b, i)];
                }
            }
        }
        __syncthreads();
    }

    if (flags & 2) {
        // the branch test of langevin_1 over all beads (md_algo.f90:176)
        for (int i = lane; i < N; i += nthr) {
            int all = 1;
            for (int b = 0; b < P; ++b)
                if (!lan_big(bv.dens[PI_(b, i)], dt)) all = 0;
            big[i] = (uint8_t)all;
        }
        __syncthreads();
        // pile_thermostat per atom (md_algo.f90:440-502): thread per (atom, component)
        if (pc.any_pile) {
            for (int q = lane; q < N * 3; q += nthr) {
                const int i = q / 3, c = q % 3;
                const int ty = bv.type[aoff + i];
                if (pc.algo[ty] != MDT_PROP_PILE) continue;
                const double mass = bv.mass[aoff + i];
                const double betaN = 1.0 / (MDT_KB * (pc.is_proj[ty] ? tproj : tsurf) * P);
                double np[64];
                for (int k = 0; k < P; ++k) {
                    double acc = 0.0;
                    for (int j = 0; j < P; ++j) acc = acc + (mass * bv.v[XI(c, j, i)]) * C[j + P * k];
                    np[k] = acc;
                }
                for (int k = 0; k < P; ++k) {
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
                    const double z =
                        mdt_normal(pc.seed, tid, step, MDT_STREAM_PILE, (uint32_t)(i * P + k), (uint32_t)c);
                    np[k] = c1 * np[k] + sqrt(mass / betaN) * c2 * z;
                }
                for (int b = 0; b < P; ++b) {
                    const uint8_t fb = bv.fixed[PI_(b, i)];
                    double ap = 0.0;
                    if (!(fb >> c & 1))
                        for (int k = 0; k < P; ++k) ap = ap + np[k] * C[b + P * k];
                    bv.v[XI(c, b, i)] = ap / mass;
                }
            }
            __syncthreads();
        }
        for (int q = lane; q < N * P; q += nthr) {
            const int i = q % N, b = q / N;
            const int ty = bv.type[aoff + i];
            const int algo = pc.algo[ty];
            const uint8_t fb = bv.fixed[PI_(b, i)];
            const double mass = bv.mass[aoff + i];
            if (algo == MDT_PROP_LANGEVIN) {
                const double temp = MDT_KB * P * (pc.is_proj[ty] ? tproj : tsurf) / mass;
                const bool exact = big[i] && temp > MDT_TOLERANCE;
                const LanCoef L = lan_coef(bv.dens[PI_(b, i)], dt, temp, exact);
                for (int c = 0; c < 3; ++c) {
                    const double eta = mdt_normal(pc.seed, tid, step, MDT_STREAM_LANGEVIN,
                                                  (uint32_t)(i * P + b), (uint32_t)c);
                    const double r = bv.x[XI(c, b, i)], v = bv.v[XI(c, b, i)], a = bv.a[XI(c, b, i)];
                    if (!(fb >> c & 1)) {
                        bv.x[XI(c, b, i)] = r + L.c1 * v + L.c2 * dt * a + L.sr * eta;
                        bv.v[XI(c, b, i)] = L.c0 * v + (L.c1 - L.c2) * a + L.sv * L.crv * eta;
                    } else {
                        bv.v[XI(c, b, i)] = 0.0;
                    }
                }
            } else if (algo == MDT_PROP_VERLET || algo == MDT_PROP_ANDERSEN || algo == MDT_PROP_PILE) {
                for (int c = 0; c < 3; ++c) {
                    const double v = bv.v[XI(c, b, i)];
                    const double vn = (fb >> c & 1) ? 0.0 : v + 0.5 * dt * bv.a[XI(c, b, i)];
                    bv.v[XI(c, b, i)] = vn;
                    if (P == 1 && !(fb >> c & 1)) bv.x[XI(c, b, i)] += dt * vn;
                }
                if (algo == MDT_PROP_ANDERSEN) {
                    const double ibetaN = P * MDT_KB * (pc.is_proj[ty] ? tproj : tsurf);
                    const double sigma = sqrt(ibetaN / mass);
                    const double u = mdt_uniform(pc.seed, tid, step, MDT_STREAM_ANDERSEN_U,
                                                 (uint32_t)(i * P + b), 0u);
                    if (u < dt / pc.andersen_time)
                        for (int c = 0; c < 3; ++c) {
                            const double z = mdt_normal(pc.seed, tid, step, MDT_STREAM_ANDERSEN_V,
                                                        (uint32_t)(i * P + b), (uint32_t)c);
                            if (!(fb >> c & 1)) bv.v[XI(c, b, i)] = z * sigma + 0.0;
                        }
                }
            } else {
                bad_local |= MDT_ST_BAD_PROP;
            }
        }
        __syncthreads();
        if (pc.any_andersen || pc.any_pile) {
            // remove_com_velocity, closed form of the 1000 iterations
            if (lane < 64) {
                double sx = 0, sy = 0, sz = 0, sm = 0;
                for (int i = lane; i < N; i += 64) {
                    const double m = bv.mass[aoff + i];
                    double bx = 0, by = 0, bz = 0;
                    for (int b = 0; b < P; ++b) {
                        bx += bv.v[XI(0, b, i)];
                        by += bv.v[XI(1, b, i)];
                        bz += bv.v[XI(2, b, i)];
                    }
                    sx += bx * m;
                    sy += by * m;
                    sz += bz * m;
                    sm += m;
                }
                for (int o = 32; o > 0; o >>= 1) {
                    sx += __shfl_xor(sx, o, 64);
                    sy += __shfl_xor(sy, o, 64);
                    sz += __shfl_xor(sz, o, 64);
                    sm += __shfl_xor(sm, o, 64);
                }
                if (lane == 0) {
                    red[0] = sx;
                    red[1] = sy;
                    red[2] = sz;
                    red[3] = sm;
                }
            }
            __syncthreads();
            const double fac = (1.0 - pow(1.0 - 1.0 / N, 1000.0)) / red[3] / P;
            for (int q = lane; q < N * P; q += nthr) {
                const int i = q % N, b = q / N;
                bv.v[XI(0, b, i)] -= red[0] * fac;
                bv.v[XI(1, b, i)] -= red[1] * fac;
                bv.v[XI(2, b, i)] -= red[2] * fac;
            }
            __syncthreads();
        }
    }

    if (flags & 4) {
        // do_ring_polymer_step (rpmd.f90:309-452), thread per (atom, component)
        const double kt = MDT_KB * tproj * P / MDT_HBAR;
        const double m1 = bv.mass[aoff + 0];
        const double prefactor = -1 * m1 * (kt * kt);
        const double piN = MDT_PI / P;
        for (int q = lane; q < N * 3; q += nthr) {
            const int i = q / 3, c = q % 3;
            const int ty = bv.type[aoff + i];
            const double mass = bv.mass[aoff + i];
            if (pc.algo[ty] == MDT_PROP_LANGEVIN) {
                // reverted step; spring force with m(1) and Tproj (rpmd.f90:418-435)
                for (int b = 0; b < P; ++b) {
                    const int pv = b == 0 ? P - 1 : b - 1, nx = b == P - 1 ? 0 : b + 1;
                    bv.f[XI(c, b, i)] = prefactor * (2 * bv.x[XI(c, b, i)] - bv.x[XI(c, pv, i)] -
                                                     bv.x[XI(c, nx, i)]);
                }
                continue;
            }
            for (int b = 0; b < P; ++b) bv.f[XI(c, b, i)] = 0.0;
            double pq[64], qq[64];
            for (int k = 0; k < P; ++k) {
                double ap = 0.0, aq = 0.0;
                for (int j = 0; j < P; ++j) {
                    ap = ap + (mass * bv.v[XI(c, j, i)]) * C[j + P * k];
                    aq = aq + bv.x[XI(c, j, i)] * C[j + P * k];
                }
                pq[k] = ap;
                qq[k] = aq;
            }
            const double betaN = 1.0 / (MDT_KB * (pc.is_proj[ty] ? tproj : tsurf) * P);
            const double twown = 2.0 / betaN / MDT_HBAR;
            for (int k = 0; k < P; ++k) {
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
                const double pn = pq[k] * p1 + qq[k] * p2;
                qq[k] = pq[k] * p3 + qq[k] * p4;
                pq[k] = pn;
            }
            for (int b = 0; b < P; ++b) {
                const uint8_t fb = bv.fixed[PI_(b, i)];
                if (fb >> c & 1) continue;
                double rr = 0.0, vv = 0.0;
                for (int k = 0; k < P; ++k) {
                    rr = rr + qq[k] * C[b + P * k];
                    vv = vv + pq[k] * C[b + P * k] / mass;
                }
                bv.x[XI(c, b, i)] = rr;
                bv.v[XI(c, b, i)] = vv;
            }
        }
    }
#undef XI
#undef PI_
    if (bad_local) atomicOr(&bv.status[t], bad_local);
}

// ---------------------------------------------------------------------------
// host-side launchers (C++ linkage, used by mdt_api.hip)
// ---------------------------------------------------------------------------
size_t mdt_replica_smem(int N, int cap, bool nbr_lds) {
    const int c = cap < N - 1 ? cap : (N > 1 ? N - 1 : 1);
    return sizeof(double) * (TB_SIZE + 8 * (size_t)N) +
           sizeof(uint16_t) * ((size_t)N + (nbr_lds ? (size_t)N * c : 0)) + N + 16;
}

template <bool L>
static hipError_t launch_replica_t(int tpa, dim3 g, dim3 blk, size_t smem, const EmtConst &ec,
                                   const PropConst &pc, const BatchView &bv, uint32_t step, int mode,
                                   int add_fin, hipStream_t st) {
    switch (tpa) {
    case 1: hipLaunchKernelGGL((k_replica<1, L>), g, blk, smem, st, ec, pc, bv, step, mode, add_fin); break;
    case 2: hipLaunchKernelGGL((k_replica<2, L>), g, blk, smem, st, ec, pc, bv, step, mode, add_fin); break;
    case 4: hipLaunchKernelGGL((k_replica<4, L>), g, blk, smem, st, ec, pc, bv, step, mode, add_fin); break;
    case 8: hipLaunchKernelGGL((k_replica<8, L>), g, blk, smem, st, ec, pc, bv, step, mode, add_fin); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t mdt_launch_replica(int tpa, int nrep, int block, size_t smem, const EmtConst &ec,
                              const PropConst &pc, const BatchView &bv, uint32_t step, int mode,
                              int add_fin, hipStream_t st) {
    dim3 g(nrep), blk(block);
    if (bv.nbr_global == nullptr)
        return launch_replica_t<true>(tpa, g, blk, smem, ec, pc, bv, step, mode, add_fin, st);
    return launch_replica_t<false>(tpa, g, blk, smem, ec, pc, bv, step, mode, add_fin, st);
}

hipError_t mdt_launch_traj(int ntraj, size_t smem, const PropConst &pc, const BatchView &bv,
                           const double *cjk, const long long *cjk_off, uint32_t step_prev,
                           uint32_t step, int flags, hipStream_t st) {
    hipLaunchKernelGGL(k_traj, dim3(ntraj), dim3(256), smem, st, pc, bv, cjk, cjk_off, step_prev, step,
                       flags);
    return hipGetLastError();
}
