// mdt_nene.hip -- CDNA4 (gfx950) kernels of the NENE surrogate: a
// Behler-Parrinello atomic network in the form md_tian2's compute_nene
// evaluates (src/pes_nene_mod.f90:7145-7368; RuNNer getshortatomic,
// atomsymfunction2, atomsymfunction3Andi, getcutoff type 1, calconenn).
//
// k_nene: calc_force with the network, one workgroup (256 lanes) per
// replica (trajectory t, bead b):
//   stage    fractional coordinates (FP64) and element of every atom in LDS;
//   pass G   centre atoms in chunks of 16 (16 lanes each): build the chunk's
//            neighbour lists within rc over all periodic images (ordered,
//            two-pass count/write, no atomics) with d, 1/r, fc, dfc in LDS,
//            then radial sums over neighbours and angular sums over
//            neighbour pairs; raw G to the replica's scratch rows; values
//            outside [min - 1e-4, max + 1e-4] flag the replica;
//   network  per element, 16-atom tiles per wavefront on
//            v_mfma_f32_16x16x32_bf16 with operands split into bf16 hi + lo
//            parts (three products per tile, FP32 accumulation: ~16-bit
//            operand precision); forward to the atomic energy, backward to
//            c = dE/dG (scaling factor folded in), written to scratch;
//   pass F   the same chunks again: dE/dr = sum_k c_k dG_k/dr contracted
//            per neighbour / neighbour pair, +F on the centre in registers,
//            the neighbours' shares into fixed-point LDS accumulators
//            (order-independent: bitwise reproducible);
//   output   f, a = f/m (fixed components 0), epot (fixed-order sum).
// Positions arrive in Angstrom (FP64) and are converted to Bohr; energy
// and forces leave in eV and eV/Angstrom, as compute_nene adds them
// (src/pes_nene_mod.f90:7245-7250).
#include <hip/hip_runtime.h>

#include "mdt_device.h"

#define NN_TPA 16
#define NN_CA 16             // centre atoms per chunk (256 lanes / 16)
#define NN_CAP 96            // neighbours within rc per atom (fcc Au at 6.5 A: 78)
#define NN_BLOCK 256
#define NN_L2E 1.4426950408889634f
#define NN_HA2EV 27.21138602
#define NN_BOHR2ANG 0.529177211
#define NN_FSCALE 4294967296.0    // 2^32 fixed point for Hartree/Bohr
#define NN_FISCALE (1.0 / 4294967296.0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

struct NnView {
    float *G;        // nrep * maxN * NN_K1 raw symmetry functions
    float *C;        // nrep * maxN * NN_K1 dE/dG
    int *rep_flag;   // nrep: extrapolation flag of the replica (EMT fallback)
    double *fpes;    // 3PN blocks: the network's forces alone (before adding the ring-polymer forces)
    int maxN;
};

// symmetry-function tables every lane indexes by element (LDS copy)
struct NnSfTab {
    int nr[NN_MAXE], na[NN_MAXE];
    int r_el[NN_MAXE][NN_MAXR];
    float r_eta[NN_MAXE][NN_MAXR], r_rs[NN_MAXE][NN_MAXR];
    int a_cls[NN_MAXE][NN_MAXA];  // neighbour-pair class e1 * NN_MAXE + e2 (e1 <= e2)
    float a_eta[NN_MAXE][NN_MAXA], a_etal2[NN_MAXE][NN_MAXA], a_lam[NN_MAXE][NN_MAXA], a_norm[NN_MAXE][NN_MAXA];
    int a_zp[NN_MAXE][NN_MAXA];
    float glo[NN_MAXE][NN_K1], ghi[NN_MAXE][NN_K1];
};

__device__ __forceinline__ float act_f(int code, float z, float *d1) {
    if (code == 1) {
        const float h = tanhf(z);
        *d1 = 1.0f - h * h;
        return h;
    }
    if (code == 2) {
        const float h = 1.0f / (1.0f + __expf(-z));
        *d1 = h * (1.0f - h);
        return h;
    }
    *d1 = 1.0f;
    return z;
}

// cut-off type 1 of getcutoff (inner cut-off 0): fc = 0.5 (cos(pi r/rc) + 1)
__device__ __forceinline__ void cutoff(float pi_rc, float r, float *fc, float *dfc) {
    const float x = r * pi_rc;
    const float s = sinf(x), c = cosf(x);
    *fc = 0.5f * (c + 1.0f);
    *dfc = -0.5f * pi_rc * s;
}

__device__ __forceinline__ void split8(const float v[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)v[j];
        hi[j] = h;
        lo[j] = (__bf16)(v[j] - (float)h);
    }
}

// D += A B with A, B given as bf16 hi + lo: hi*hi + hi*lo + lo*hi
__device__ __forceinline__ f32x4 mfma3(const bf16x8 &ah, const bf16x8 &al, const bf16x8 &bh, const bf16x8 &bl,
                                       f32x4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
    return c;
}

__device__ __forceinline__ void fxadd(unsigned long long *acc, double x, int *bad) {
    const double y = x * NN_FSCALE;
    if (!(fabs(y) < 9.0e18)) {
        *bad |= MDT_ST_FIXED_OVERFLOW;
        return;
    }
    atomicAdd(acc, (unsigned long long)(long long)rint(y));
}

// per-chunk neighbour data (structure of arrays)
struct ChunkNbr {
    float dx[NN_CA * NN_CAP], dy[NN_CA * NN_CAP], dz[NN_CA * NN_CAP];
    float ir[NN_CA * NN_CAP], r[NN_CA * NN_CAP], fc[NN_CA * NN_CAP], dfc[NN_CA * NN_CAP];
    uint16_t j[NN_CA * NN_CAP];
    uint8_t el[NN_CA * NN_CAP];
    int cnt[NN_CA];
};

}  // namespace

// Builds the neighbour lists of centre atoms k0 .. k0+NN_CA-1 (16 lanes per
// centre; lane `sub` tests atoms n = sub, sub+16, ...; hits are counted,
// prefix-summed over the 16 lanes and written in (lane, n, image) order).
__device__ void build_chunk(float rc2, float pi_rc, ChunkNbr &L, const double *fx, const double *fy, const double *fz,
                            const uint8_t *elem, const double h[9], const int *imax, float wx, float wy, float wz,
                            int N, int k0, int lane, int *bad) {
    const int c = lane / NN_TPA, sub = lane % NN_TPA;
    const int i = k0 + c;
    int cnt = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int base = 0;
        if (pass == 1) {
            // exclusive prefix of the 16 lanes' counts
            int incl = cnt;
#pragma unroll
            for (int o = 1; o < NN_TPA; o <<= 1) {
                const int y = __shfl_up(incl, o, NN_TPA);
                if (sub >= o) incl += y;
            }
            base = incl - cnt;
            const int tot = __shfl(incl, NN_TPA - 1, NN_TPA);
            if (sub == 0) L.cnt[c] = tot < NN_CAP ? tot : NN_CAP;
            if (tot > NN_CAP) *bad |= MDT_ST_NBR_OVERFLOW;
        }
        int pos = base;
        if (i < N) {
            const double fxi = fx[i], fyi = fy[i], fzi = fz[i];
            for (int n = sub; n < N; n += NN_TPA) {
                const double f0 = fx[n] - fxi, f1 = fy[n] - fyi, f2 = fz[n] - fzi;  // in (-1, 1)
                // integer shifts s with |f + s| <= w on each axis (w: fractional half-width of the sphere)
                const int a0 = (int)ceil(-wx - f0), b0 = (int)floor(wx - f0);
                const int a1 = (int)ceil(-wy - f1), b1 = (int)floor(wy - f1);
                const int a2 = (int)ceil(-wz - f2), b2 = (int)floor(wz - f2);
                for (int s0 = max(a0, -imax[0]); s0 <= min(b0, imax[0]); ++s0)
                    for (int s1 = max(a1, -imax[1]); s1 <= min(b1, imax[1]); ++s1)
                        for (int s2 = max(a2, -imax[2]); s2 <= min(b2, imax[2]); ++s2) {
                            if (n == i && s0 == 0 && s1 == 0 && s2 == 0) continue;
                            const double g0 = f0 + s0, g1 = f1 + s1, g2 = f2 + s2;
                            const double ex = h[0] * g0 + h[3] * g1 + h[6] * g2;
                            const double ey = h[1] * g0 + h[4] * g1 + h[7] * g2;
                            const double ez = h[2] * g0 + h[5] * g1 + h[8] * g2;
                            const double r2 = ex * ex + ey * ey + ez * ez;
                            if (r2 > (double)rc2) continue;
                            if (pass == 0) {
                                ++cnt;
                            } else {
                                if (pos < NN_CAP) {
                                    const int e = c * NN_CAP + pos;
                                    const float rr = (float)sqrt(r2);
                                    float fcv, dfcv;
                                    cutoff(pi_rc, rr, &fcv, &dfcv);
                                    L.dx[e] = (float)ex;
                                    L.dy[e] = (float)ey;
                                    L.dz[e] = (float)ez;
                                    L.r[e] = rr;
                                    L.ir[e] = 1.0f / rr;
                                    L.fc[e] = fcv;
                                    L.dfc[e] = dfcv;
                                    L.j[e] = (uint16_t)n;
                                    L.el[e] = elem[n];
                                }
                                ++pos;
                            }
                        }
            }
        }
    }
    if (i >= N && sub == 0) L.cnt[c] = 0;
}

template <int WIDTH>
__device__ __forceinline__ float sum_group(float v) {
#pragma unroll
    for (int o = 1; o < WIDTH; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One 16-atom tile of element e on one wavefront: atomic energies to Ea,
// c = dE/dG to Crow (rows of the replica's C scratch).  T is the wave's
// 16 x 33 transpose buffer; every lane of the block calls this the same
// number of times (syncs inside).
__device__ void mlp_tile(const NnConst &M, int e, const int *rows, const float *Grep, float *Crep, double *Ea,
                         float *T, int lane) {
    const int l = lane & 63;
    const int rA = l & 15, kg = (l >> 4) * 8;       // A operand: row, first k
    const int nB = l & 15;                         // B operand column
    const int rC = (l >> 4) * 4;                   // C: first row of this lane
    const bool act = e >= 0;
    const int ee = act ? e : 0;
    const int nin = act ? M.nr[ee] + M.na[ee] : 0;
    const int atomA = act ? rows[rA] : -1;

    // ---- layer 1: z1 = x W1 + b1 ----
    f32x4 z1[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        float av[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = ks * 32 + kg + j;
            av[j] = (atomA >= 0 && k < nin) ? Grep[(size_t)atomA * NN_K1 + k] * M.smul[ee][k] + M.sadd[ee][k] : 0.f;
        }
        bf16x8 ah, al;
        split8(av, ah, al);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            float bv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) bv[j] = M.W1[ee][ks * 32 + kg + j][nt * 16 + nB];
            bf16x8 bh, bl;
            split8(bv, bh, bl);
            z1[nt] = mfma3(ah, al, bh, bl, z1[nt]);
        }
    }
    float h1[2][4], d1[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) h1[nt][i] = act_f(M.act1, z1[nt][i] + M.b1[ee][nt * 16 + nB], &d1[nt][i]);
    // transpose h1 (C layout) into A layout through T
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(rC + i) * 33 + nt * 16 + nB] = h1[nt][i];
    __syncthreads();
    float av2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) av2[j] = T[rA * 33 + kg + j];
    __syncthreads();
    // ---- layer 2 ----
    bf16x8 ah, al;
    split8(av2, ah, al);
    f32x4 z2[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W2[ee][kg + j][nt * 16 + nB];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        z2[nt] = mfma3(ah, al, bh, bl, z2[nt]);
    }
    float h2[2][4], d2[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) h2[nt][i] = act_f(M.act2, z2[nt][i] + M.b2[ee][nt * 16 + nB], &d2[nt][i]);
    // ---- output node and its derivative ----
    float dz3[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float p = h2[0][i] * M.w3[ee][nB] + h2[1][i] * M.w3[ee][16 + nB];
        p = sum_group<16>(p);
        float d3;
        const float E = act_f(M.act3, p + M.b3[ee], &d3);
        dz3[i] = d3;
        const int row = rC + i;
        if (act && nB == 0 && rows[row] >= 0) Ea[rows[row]] = (double)E;
    }
    // delta2 = dE/dz2 (C layout) -> A layout
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(rC + i) * 33 + nt * 16 + nB] = dz3[i] * M.w3[ee][nt * 16 + nB] * d2[nt][i];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) av2[j] = T[rA * 33 + kg + j];
    __syncthreads();
    split8(av2, ah, al);
    // dE/dh1 = delta2 W2^T
    f32x4 g1[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W2[ee][nt * 16 + nB][kg + j];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        g1[nt] = mfma3(ah, al, bh, bl, g1[nt]);
    }
    // delta1 = dE/dz1 -> A layout
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(rC + i) * 33 + nt * 16 + nB] = g1[nt][i] * d1[nt][i];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) av2[j] = T[rA * 33 + kg + j];
    __syncthreads();
    split8(av2, ah, al);
    // dE/dx = delta1 W1^T, four 16-column tiles of the 64 inputs
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W1[ee][nt * 16 + nB][kg + j];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        f32x4 dx = {0.f, 0.f, 0.f, 0.f};
        dx = mfma3(ah, al, bh, bl, dx);
        const int k = nt * 16 + nB;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = rC + i;
            if (act && rows[row] >= 0 && k < nin) Crep[(size_t)rows[row] * NN_K1 + k] = dx[i] * M.dfac[ee][k];
        }
    }
}

__global__ void __launch_bounds__(NN_BLOCK) k_nene(const NnConst *__restrict__ Mg, BatchView bv, NnView nv,
                                                   int add_fin, int skip_extrap) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ NnSfTab S;
    __shared__ ChunkNbr L;
    __shared__ float T[NN_BLOCK / 64][16 * 33];
    __shared__ float cb[NN_CA][NN_K1];
    __shared__ int rowsb[NN_BLOCK / 64][16];
    __shared__ int s_ne[NN_MAXE];
    const NnConst &M = *Mg;
    const int rep = blockIdx.x;
    const int t = bv.rep_traj[rep];
    const int b = bv.rep_bead[rep];
    const int N = bv.traj_N[t];
    const long long soff = bv.traj_soff[t];
    const long long roff = soff + (long long)b * 3 * N;
    const long long doff = soff / 3 + (long long)b * N;
    const long long aoff = bv.traj_aoff[t];
    const int lane = threadIdx.x;
    const int wave = lane >> 6;
    float *Grep = nv.G + (size_t)rep * nv.maxN * NN_K1;
    float *Crep = nv.C + (size_t)rep * nv.maxN * NN_K1;
    int bad = 0, extrap = 0;

    // dynamic LDS: fx fy fz Ea (double N each) | ax ay az (u64 N) | elist (u16 N) | elem (u8 N)
    double *fx = (double *)smem_raw, *fy = fx + N, *fz = fy + N, *Ea = fz + N;
    unsigned long long *ax = (unsigned long long *)(Ea + N), *ay = ax + N, *az = ay + N;
    uint16_t *elist = (uint16_t *)(az + N);
    uint8_t *elem = (uint8_t *)(elist + N);

    // ---- tables ----
    const int ne_model = M.nelem;
    for (int k = lane; k < NN_MAXE; k += NN_BLOCK) {
        S.nr[k] = M.nr[k];
        S.na[k] = M.na[k];
    }
    for (int k = lane; k < NN_MAXE * NN_MAXR; k += NN_BLOCK) {
        const int e = k / NN_MAXR, q = k % NN_MAXR;
        S.r_el[e][q] = M.r_el[e][q];
        S.r_eta[e][q] = M.r_eta[e][q];
        S.r_rs[e][q] = M.r_rs[e][q];
    }
    for (int k = lane; k < NN_MAXE * NN_MAXA; k += NN_BLOCK) {
        const int e = k / NN_MAXA, q = k % NN_MAXA;
        S.a_cls[e][q] = M.a_e1[e][q] * NN_MAXE + M.a_e2[e][q];
        S.a_etal2[e][q] = M.a_etal2[e][q];
        S.a_eta[e][q] = M.a_etal2[e][q] * (1.0f / NN_L2E);
        S.a_lam[e][q] = M.a_lam[e][q];
        S.a_norm[e][q] = M.a_norm[e][q];
        S.a_zp[e][q] = M.a_zp[e][q];
    }
    for (int k = lane; k < NN_MAXE * NN_K1; k += NN_BLOCK) {
        const int e = k / NN_K1, q = k % NN_K1;
        S.glo[e][q] = M.glo[e][q];
        S.ghi[e][q] = M.ghi[e][q];
    }
    // cell in Bohr (columns = lattice vectors) and the fractional half-widths
    // of the cut-off sphere, rc |row k of H^-1|
    double h[9], ih[9];
    for (int k = 0; k < 9; ++k) {
        h[k] = bv.cell[9 * t + k] * (double)M.ang2bohr;
        ih[k] = bv.icell[9 * t + k];
    }
    const double rcb = (double)M.rc;
    float w[3];
    int imax[3];
    for (int k = 0; k < 3; ++k) {
        const double bk = sqrt(ih[k] * ih[k] + ih[k + 3] * ih[k + 3] + ih[k + 6] * ih[k + 6]) / (double)M.ang2bohr;
        w[k] = (float)(rcb * bk);
        imax[k] = (int)ceil(rcb * bk) + 1;
    }

    // ---- stage ----
    for (int i = lane; i < N; i += NN_BLOCK) {
        const int ty = bv.type[aoff + i];
        const int e = M.elem_of_type[ty];
        if (e < 0 || e >= ne_model) bad |= MDT_ST_BAD_PROP;
        elem[i] = (uint8_t)(e < 0 ? 0 : e);
        const double r0 = bv.x[roff + i], r1 = bv.x[roff + N + i], r2 = bv.x[roff + 2 * N + i];
        double f0 = ih[0] * r0 + ih[3] * r1 + ih[6] * r2;
        double f1 = ih[1] * r0 + ih[4] * r1 + ih[7] * r2;
        double f2 = ih[2] * r0 + ih[5] * r1 + ih[8] * r2;
        fx[i] = f0 - floor(f0);
        fy[i] = f1 - floor(f1);
        fz[i] = f2 - floor(f2);
        Ea[i] = 0.0;
        ax[i] = ay[i] = az[i] = 0ull;
    }
    __syncthreads();
    // atoms grouped by element, ascending index within each group (wave 0)
    if (wave == 0) {
        int base = 0;
        for (int e = 0; e < ne_model; ++e) {
            const int start = base;
            for (int c0 = 0; c0 < N; c0 += 64) {
                const int i = c0 + lane;
                const bool in = i < N && elem[i] == e;
                const unsigned long long m = __ballot(in);
                if (in) elist[base + __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))))] = (uint16_t)i;
                base += __popcll(m);
            }
            if (lane == 0) s_ne[e] = base - start;
        }
    }
    __syncthreads();

    const int c = lane / NN_TPA, sub = lane % NN_TPA;
    // ---- pass G: symmetry functions ----
    for (int k0 = 0; k0 < N; k0 += NN_CA) {
        build_chunk(M.rc2, M.pi_rc, L, fx, fy, fz, elem, h, imax, w[0], w[1], w[2], N, k0, lane, &bad);
        __syncthreads();
        const int i = k0 + c;
        const int cnt = L.cnt[c];
        const int e = (i < N) ? elem[i] : 0;
        const int nr = (i < N) ? S.nr[e] : 0, na = (i < N) ? S.na[e] : 0;
        const int base = c * NN_CAP;
        float g[NN_MAXR > NN_MAXA ? NN_MAXR : NN_MAXA];
        // radial (atomsymfunction2: exp(-eta (r - rs)^2) fc(r))
#pragma unroll
        for (int k = 0; k < NN_MAXR; ++k) g[k] = 0.f;
        for (int p = sub; p < cnt; p += NN_TPA) {
            const float r = L.r[base + p], fcv = L.fc[base + p];
            const int ep = L.el[base + p];
#pragma unroll
            for (int k = 0; k < NN_MAXR; ++k)
                if (k < nr && ep == S.r_el[e][k]) {
                    const float dr = r - S.r_rs[e][k];
                    g[k] += __builtin_amdgcn_exp2f(-S.r_eta[e][k] * NN_L2E * dr * dr) * fcv;
                }
        }
#pragma unroll
        for (int k = 0; k < NN_MAXR; ++k) {
            const float v = sum_group<NN_TPA>(g[k]);
            if (i < N && k < nr && sub == (k & (NN_TPA - 1))) {
                Grep[(size_t)i * NN_K1 + k] = v;
                if (v < S.glo[e][k] || v > S.ghi[e][k]) extrap = 1;
            }
        }
        // angular (atomsymfunction3Andi): pairs p < q of the neighbour list
#pragma unroll
        for (int k = 0; k < NN_MAXA; ++k) g[k] = 0.f;
        for (int p = sub; p < cnt; p += NN_TPA) {
            const float ax_ = L.dx[base + p], ay_ = L.dy[base + p], az_ = L.dz[base + p];
            const float ra = L.r[base + p], ia = L.ir[base + p], fa = L.fc[base + p];
            const int ep = L.el[base + p];
            for (int q = p + 1; q < cnt; ++q) {
                const float bx = L.dx[base + q], by = L.dy[base + q], bz = L.dz[base + q];
                const float cx = bx - ax_, cy = by - ay_, cz = bz - az_;
                const float rc2v = cx * cx + cy * cy + cz * cz;
                if (rc2v > M.rc2) continue;
                const float rb = L.r[base + q];
                const int eq = L.el[base + q];
                const int cls = ep < eq ? ep * NN_MAXE + eq : eq * NN_MAXE + ep;
                const float cth = (ax_ * bx + ay_ * by + az_ * bz) * ia * L.ir[base + q];
                float fcc, dfcc;
                cutoff(M.pi_rc, sqrtf(rc2v), &fcc, &dfcc);
                const float Cf = fa * L.fc[base + q] * fcc;
                const float Sd = ra * ra + rb * rb + rc2v;
                float B = 0.f, etaprev = -1.f;
#pragma unroll
                for (int k = 0; k < NN_MAXA; ++k)
                    if (k < na && cls == S.a_cls[e][k]) {
                        const float x = 1.0f + S.a_lam[e][k] * cth;
                        if (x > 0.f) {
                            float pw = x;
                            for (int z = 0; z < S.a_zp[e][k]; ++z) pw *= pw;
                            if (S.a_etal2[e][k] != etaprev) {
                                etaprev = S.a_etal2[e][k];
                                B = __builtin_amdgcn_exp2f(-etaprev * Sd);
                            }
                            g[k] += S.a_norm[e][k] * pw * B * Cf;
                        }
                    }
            }
        }
#pragma unroll
        for (int k = 0; k < NN_MAXA; ++k) {
            const float v = sum_group<NN_TPA>(g[k]);
            if (i < N && k < na && sub == (k & (NN_TPA - 1))) {
                Grep[(size_t)i * NN_K1 + nr + k] = v;
                if (v < S.glo[e][nr + k] || v > S.ghi[e][nr + k]) extrap = 1;
            }
        }
        __syncthreads();
    }

    // replicas outside the training range go to the EMT fallback untouched
    extrap = __syncthreads_or(extrap);
    if (extrap && skip_extrap) {
        if (lane == 0) {
            nv.rep_flag[rep] = 1;
            atomicOr(&bv.status[t], bad | MDT_ST_NN_EXTRAPOLATION);
        } else if (bad) {
            atomicOr(&bv.status[t], bad);
        }
        return;
    }

    // ---- network: per element, 16-atom tiles, one per wavefront per round ----
    int ntiles = 0;
    for (int e = 0; e < ne_model; ++e) ntiles += (s_ne[e] + 15) / 16;
    const int nwaves = NN_BLOCK / 64;
    for (int r0 = 0; r0 < ntiles; r0 += nwaves) {
        const int tile = r0 + wave;
        int e = -1, first = 0, cntE = 0, tl = tile;
        {
            int start = 0;
            for (int q = 0; q < ne_model; ++q) {
                const int nt = (s_ne[q] + 15) / 16;
                if (e < 0 && tl < nt) {
                    e = q;
                    first = start + tl * 16;
                    cntE = s_ne[q] - tl * 16;
                } else if (e < 0) {
                    tl -= nt;
                }
                start += s_ne[q];
            }
        }
        if ((lane & 63) < 16) {
            const int rr = lane & 15;
            rowsb[wave][rr] = (e >= 0 && rr < cntE) ? (int)elist[first + rr] : -1;
        }
        __syncthreads();
        mlp_tile(M, e, rowsb[wave], Grep, Crep, Ea, T[wave], lane);
        __syncthreads();
    }

    // ---- pass F: forces ----
    for (int k0 = 0; k0 < N; k0 += NN_CA) {
        build_chunk(M.rc2, M.pi_rc, L, fx, fy, fz, elem, h, imax, w[0], w[1], w[2], N, k0, lane, &bad);
        for (int q = lane; q < NN_CA * NN_K1; q += NN_BLOCK) {
            const int ci = q / NN_K1, k = q % NN_K1;
            const int i = k0 + ci;
            cb[ci][k] = (i < N && k < S.nr[elem[i]] + S.na[elem[i]]) ? Crep[(size_t)i * NN_K1 + k] : 0.f;
        }
        __syncthreads();
        const int i = k0 + c;
        const int cnt = L.cnt[c];
        const int e = (i < N) ? elem[i] : 0;
        const int nr = (i < N) ? S.nr[e] : 0, na = (i < N) ? S.na[e] : 0;
        const int base = c * NN_CAP;
        float Fi0 = 0.f, Fi1 = 0.f, Fi2 = 0.f;  // force on the centre (Hartree/Bohr)
        // radial: dG/dr = (-2 eta (r - rs) fc + dfc) exp(-eta (r - rs)^2)
        for (int p = sub; p < cnt; p += NN_TPA) {
            const float r = L.r[base + p], fcv = L.fc[base + p], dfcv = L.dfc[base + p];
            const int ep = L.el[base + p];
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < NN_MAXR; ++k)
                if (k < nr && ep == S.r_el[e][k]) {
                    const float dr = r - S.r_rs[e][k];
                    const float ex = __builtin_amdgcn_exp2f(-S.r_eta[e][k] * NN_L2E * dr * dr);
                    s += cb[c][k] * ex * (dfcv - 2.f * S.r_eta[e][k] * dr * fcv);
                }
            const float ir = L.ir[base + p];
            const float v0 = s * L.dx[base + p] * ir, v1 = s * L.dy[base + p] * ir, v2 = s * L.dz[base + p] * ir;
            Fi0 += v0;
            Fi1 += v1;
            Fi2 += v2;
            const int j = L.j[base + p];
            fxadd(&ax[j], -(double)v0, &bad);
            fxadd(&ay[j], -(double)v1, &bad);
            fxadd(&az[j], -(double)v2, &bad);
        }
        // angular
        for (int p = sub; p < cnt; p += NN_TPA) {
            const float ax_ = L.dx[base + p], ay_ = L.dy[base + p], az_ = L.dz[base + p];
            const float ra = L.r[base + p], ia = L.ir[base + p], fa = L.fc[base + p], dfa = L.dfc[base + p];
            const int ep = L.el[base + p];
            float Fj0 = 0.f, Fj1 = 0.f, Fj2 = 0.f;
            for (int q = p + 1; q < cnt; ++q) {
                const float bx = L.dx[base + q], by = L.dy[base + q], bz = L.dz[base + q];
                const float cx = bx - ax_, cy = by - ay_, cz = bz - az_;
                const float rc2v = cx * cx + cy * cy + cz * cz;
                if (rc2v > M.rc2) continue;
                const float rb = L.r[base + q], ib = L.ir[base + q], fb = L.fc[base + q], dfb = L.dfc[base + q];
                const int eq = L.el[base + q];
                const int cls = ep < eq ? ep * NN_MAXE + eq : eq * NN_MAXE + ep;
                const float cth = (ax_ * bx + ay_ * by + az_ * bz) * ia * ib;
                const float rcc = sqrtf(rc2v);
                float fcc, dfcc;
                cutoff(M.pi_rc, rcc, &fcc, &dfcc);
                const float Cf = fa * fb * fcc;
                const float Sd = ra * ra + rb * rb + rc2v;
                float Sg = 0.f, Sc = 0.f, Ss = 0.f, B = 0.f, etaprev = -1.f;
#pragma unroll
                for (int k = 0; k < NN_MAXA; ++k)
                    if (k < na && cls == S.a_cls[e][k]) {
                        const float lam = S.a_lam[e][k];
                        const float x = 1.0f + lam * cth;
                        if (x > 0.f) {
                            float pw = x, pm = 1.f;  // x^zeta, x^(zeta-1)
                            for (int z = 0; z < S.a_zp[e][k]; ++z) {
                                pm *= pw;
                                pw *= pw;
                            }
                            if (S.a_etal2[e][k] != etaprev) {
                                etaprev = S.a_etal2[e][k];
                                B = __builtin_amdgcn_exp2f(-etaprev * Sd);
                            }
                            const float ck = cb[c][nr + k] * S.a_norm[e][k] * B;
                            const float zeta = (float)(1 << S.a_zp[e][k]);
                            Sg += ck * pw;
                            Sc += ck * zeta * lam * pm;
                            Ss += ck * S.a_eta[e][k] * pw;
                        }
                    }
                if (Sg == 0.f && Sc == 0.f) continue;
                const float iab = ia * ib;
                const float tc = Cf * Sc, ts = -2.f * Cf * Ss;
                const float ua = Sg * dfa * fb * fcc * ia, ub = Sg * fa * dfb * fcc * ib;
                const float uc = Sg * fa * fb * dfcc / rcc;
                const float ga0 = tc * (bx * iab - cth * ax_ * ia * ia) + ts * ax_ + ua * ax_;
                const float ga1 = tc * (by * iab - cth * ay_ * ia * ia) + ts * ay_ + ua * ay_;
                const float ga2 = tc * (bz * iab - cth * az_ * ia * ia) + ts * az_ + ua * az_;
                const float gb0 = tc * (ax_ * iab - cth * bx * ib * ib) + ts * bx + ub * bx;
                const float gb1 = tc * (ay_ * iab - cth * by * ib * ib) + ts * by + ub * by;
                const float gb2 = tc * (az_ * iab - cth * bz * ib * ib) + ts * bz + ub * bz;
                const float gc0 = (ts + uc) * cx, gc1 = (ts + uc) * cy, gc2 = (ts + uc) * cz;
                Fi0 += ga0 + gb0;
                Fi1 += ga1 + gb1;
                Fi2 += ga2 + gb2;
                Fj0 -= ga0 - gc0;
                Fj1 -= ga1 - gc1;
                Fj2 -= ga2 - gc2;
                const int kq = L.j[base + q];
                fxadd(&ax[kq], -(double)(gb0 + gc0), &bad);
                fxadd(&ay[kq], -(double)(gb1 + gc1), &bad);
                fxadd(&az[kq], -(double)(gb2 + gc2), &bad);
            }
            const int j = L.j[base + p];
            fxadd(&ax[j], (double)Fj0, &bad);
            fxadd(&ay[j], (double)Fj1, &bad);
            fxadd(&az[j], (double)Fj2, &bad);
        }
        const float S0 = sum_group<NN_TPA>(Fi0), S1 = sum_group<NN_TPA>(Fi1), S2 = sum_group<NN_TPA>(Fi2);
        if (i < N && sub == 0) {
            fxadd(&ax[i], (double)S0, &bad);
            fxadd(&ay[i], (double)S1, &bad);
            fxadd(&az[i], (double)S2, &bad);
        }
        __syncthreads();
    }

    // ---- outputs: f [eV/A], a, dens (no EMT density: 0), epot [eV] ----
    const double fconv = NN_HA2EV / NN_BOHR2ANG;
    for (int i = lane; i < N; i += NN_BLOCK) {
        double F[3] = {(double)(long long)ax[i] * NN_FISCALE * fconv, (double)(long long)ay[i] * NN_FISCALE * fconv,
                       (double)(long long)az[i] * NN_FISCALE * fconv};
        nv.fpes[roff + i] = F[0];
        nv.fpes[roff + N + i] = F[1];
        nv.fpes[roff + 2 * N + i] = F[2];
        if (add_fin) {
            F[0] += bv.f[roff + i];
            F[1] += bv.f[roff + N + i];
            F[2] += bv.f[roff + 2 * N + i];
        }
        const uint8_t fb = bv.fixed[doff + i];
        const double mass = bv.mass[aoff + i];
        bv.f[roff + i] = F[0];
        bv.f[roff + N + i] = F[1];
        bv.f[roff + 2 * N + i] = F[2];
        bv.a[roff + i] = (fb & 1) ? 0.0 : F[0] / mass;
        bv.a[roff + N + i] = (fb & 2) ? 0.0 : F[1] / mass;
        bv.a[roff + 2 * N + i] = (fb & 4) ? 0.0 : F[2] / mass;
        bv.dens[doff + i] = 0.0;
    }
    __syncthreads();
    if (lane < 64) {
        double e = 0.0;
        for (int i = lane; i < N; i += 64) e += Ea[i];
        for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
        if (lane == 0) bv.epot[bv.traj_eoff[t] + b] = e * NN_HA2EV;
    }
    if (lane == 0) {
        nv.rep_flag[rep] = extrap;
        if (extrap) bad |= MDT_ST_NN_EXTRAPOLATION;
    }
    if (bad) atomicOr(&bv.status[t], bad);
}

size_t mdt_nene_smem(int N) { return (size_t)N * (4 * 8 + 3 * 8 + 2 + 1) + 16; }

hipError_t mdt_nene_attributes(hipFuncAttributes *fa) { return hipFuncGetAttributes(fa, (const void *)k_nene); }

hipError_t mdt_launch_nene(int nrep, size_t smem, const NnConst *M, const BatchView &bv, float *G, float *C,
                           int *rep_flag, double *fpes, int maxN, int add_fin, int skip_extrap, hipStream_t st) {
    NnView nv{G, C, rep_flag, fpes, maxN};
    hipLaunchKernelGGL(k_nene, dim3(nrep), dim3(NN_BLOCK), smem, st, M, bv, nv, add_fin, skip_extrap);
    return hipGetLastError();
}
