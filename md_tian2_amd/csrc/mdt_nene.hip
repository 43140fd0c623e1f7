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
    // radial kernels in minimum-image cells: k_nene_rg stores its neighbour
    // lists (within rc + skin) and k_nene_rf reads them; with skin > 0 they
    // are kept across launches while every atom stays within skin/2 of its
    // position at the build.  nbr NULL: lists are built in both kernels.
    uint16_t *nbr;      // nrep * maxN * NN_CAP neighbour indices
    uint8_t *nbr_cnt;   // nrep * maxN list lengths
    uint32_t *nbr_ref;  // nrep * 3 * maxN fractional positions (frac_fix) at the build
    int *nbr_key;       // nrep: bits of (rc + skin) of valid kept lists, 0: none
    float skin;         // Bohr
};

// symmetry-function tables every lane indexes by element (LDS copy)
struct NnSfTab {
    int nr[NN_MAXE], na[NN_MAXE];
    int8_t r_el[NN_MAXE][NN_MAXR];
    float r_eta[NN_MAXE][NN_MAXR], r_rs[NN_MAXE][NN_MAXR];
    int8_t a_cls[NN_MAXE][NN_MAXA];  // neighbour-pair class e1 * NN_MAXE + e2 (e1 <= e2)
    float a_eta[NN_MAXE][NN_MAXA], a_etal2[NN_MAXE][NN_MAXA], a_lam[NN_MAXE][NN_MAXA], a_norm[NN_MAXE][NN_MAXA];
    int8_t a_zp[NN_MAXE][NN_MAXA];
    // factored angular sums: an element whose angular functions share one
    // neighbour-pair class, take at most NN_FB distinct eta and NN_FB
    // distinct (lambda, zeta), and map one-to-one onto (eta, (lambda,
    // zeta)) slots accumulates per triplet the outer product of
    // B_n = exp(-eta_n (rij^2 + rik^2 + rjk^2)) and u_m = (1 + lambda_m
    // cos)^zeta_m; function k is norm_k * slot (n_k, m_k)
    int f_ok[NN_MAXE], f_cls[NN_MAXE], f_nb[NN_MAXE], f_nu[NN_MAXE];
    float f_etal2[NN_MAXE][4], f_eta[NN_MAXE][4], f_lam[NN_MAXE][4];
    int f_zp[NN_MAXE][4];
    int8_t f_slot[NN_MAXE][NN_MAXA];  // n * NN_FB + m
};
#define NN_FB 4
static_assert(NN_FB * NN_FB == NN_TPA, "one slot of the factored angular sums per lane of a centre's group");

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
// (hardware sine / cosine of x / 2 pi revolutions; x = pi r / rc in [0, pi])
__device__ __forceinline__ void cutoff(float pi_rc, float r, float *fc, float *dfc) {
    const float rev = r * pi_rc * 0.15915494309189535f;
    const float s = __builtin_amdgcn_sinf(rev), c = __builtin_amdgcn_cosf(rev);
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

// x * 2^32 from an FP32 x as a 64-bit fixed-point addend: floor(x) and the
// 32-bit fraction (x - floor(x)) 2^32, both exact in FP32 for |x| < 2^31
// (exact when x has no bits below 2^-32, truncated below that)
__device__ __forceinline__ void fxadd32(unsigned long long *acc, float x, int *bad) {
    if (!(fabsf(x) < 2.0e9f)) {
        *bad |= MDT_ST_FIXED_OVERFLOW;
        return;
    }
    const float fl = floorf(x);
    const uint32_t lo = (uint32_t)((x - fl) * 4294967296.0f);
    atomicAdd(acc, ((unsigned long long)(uint32_t)(int)fl << 32) + lo);
}

// per-chunk neighbour data (structure of arrays)
struct ChunkNbr {
    float dx[NN_CA * NN_CAP], dy[NN_CA * NN_CAP], dz[NN_CA * NN_CAP];
    float fc[NN_CA * NN_CAP], dfc[NN_CA * NN_CAP];
    uint16_t j[NN_CA * NN_CAP];
    uint8_t el[NN_CA * NN_CAP];
    int cnt[NN_CA];
};

// r and 1/r of a neighbour entry (from its vector: one reciprocal square root)
__device__ __forceinline__ void nbr_rir(const ChunkNbr &L, int e, float *r, float *ir) {
    const float x = L.dx[e], y = L.dy[e], z = L.dz[e];
    const float r2 = x * x + y * y + z * z;
    *ir = __builtin_amdgcn_rsqf(r2);
    *r = r2 * *ir;
}

}  // namespace

// A fractional coordinate in [0, 1] as 2^-32 fixed point (1 wraps to 0,
// the same point of the cell): 8e-9 A on a 35 A cell, below the FP32
// rounding of the neighbour vectors built from it, in half the LDS of FP64.
__device__ __forceinline__ uint32_t frac_fix(double f) {
    return (uint32_t)(unsigned long long)(f * 4294967296.0);
}

// Neighbour list of centre i (16 lanes; i < 0: none) within rc over all
// periodic images: lane `sub` tests atoms n = sub, sub+16, ...; hits are
// counted, prefix-summed over the 16 lanes and passed to put(pos, n, ex, ey,
// ez, r2) in (lane, n, image) order (vector in Bohr).  Returns the number of
// hits (in every lane of the group); more than NN_CAP set the overflow flag
// and only the first NN_CAP are passed.  ux, uy, uz: the atoms' fractional
// coordinates (frac_fix).
template <class Put>
__device__ int build_images(float rc2, const uint32_t *ux, const uint32_t *uy, const uint32_t *uz, const double h[9],
                            const int *imax, float wx, float wy, float wz, int N, int i, int lane, int *bad,
                            Put put) {
    const int sub = lane % NN_TPA;
    int cnt = 0, tot = 0;
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
            tot = __shfl(incl, NN_TPA - 1, NN_TPA);
            if (tot > NN_CAP) *bad |= MDT_ST_NBR_OVERFLOW;
        }
        int pos = base;
        if (i >= 0) {
            const long long xi = ux[i], yi = uy[i], zi = uz[i];
            for (int n = sub; n < N; n += NN_TPA) {
                const double f0 = (double)((long long)ux[n] - xi) * 0x1p-32;  // in (-1, 1)
                const double f1 = (double)((long long)uy[n] - yi) * 0x1p-32;
                const double f2 = (double)((long long)uz[n] - zi) * 0x1p-32;
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
                                if (pos < NN_CAP) put(pos, n, ex, ey, ez, r2);
                                ++pos;
                            }
                        }
            }
        }
    }
    return tot;
}

// Builds the neighbour lists of centre atoms k0 .. k0+NN_CA-1 (16 lanes per
// centre) with d, fc, dfc, the neighbour and its element per entry.
__device__ void build_chunk(float rc2, float pi_rc, ChunkNbr &L, const uint32_t *ux, const uint32_t *uy,
                            const uint32_t *uz, const uint8_t *elem, const double h[9], const int *imax, float wx,
                            float wy, float wz, int N, int k0, int lane, int *bad) {
    const int c = lane / NN_TPA, sub = lane % NN_TPA;
    const int i = k0 + c;
    const int tot = build_images(rc2, ux, uy, uz, h, imax, wx, wy, wz, N, i < N ? i : -1, lane, bad,
                                 [&](int pos, int n, double ex, double ey, double ez, double r2) {
                                     const int e = c * NN_CAP + pos;
                                     const float rr = (float)sqrt(r2);
                                     float fcv, dfcv;
                                     cutoff(pi_rc, rr, &fcv, &dfcv);
                                     L.dx[e] = (float)ex;
                                     L.dy[e] = (float)ey;
                                     L.dz[e] = (float)ez;
                                     L.fc[e] = fcv;
                                     L.dfc[e] = dfcv;
                                     L.j[e] = (uint16_t)n;
                                     L.el[e] = elem[n];
                                 });
    if (sub == 0) L.cnt[c] = tot < NN_CAP ? tot : NN_CAP;
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

// Symmetry functions of the chunk's centre atoms (16 lanes each) into
// Grep rows; *extrap set when a value leaves [min - 1e-4, max + 1e-4].
__device__ void g_chunk(const NnConst &Mc, const NnSfTab &S, const ChunkNbr &L, const uint8_t *elem, int N, int k0,
                        int lane,
                        float rc2, float pi_rc, float *Grep, int *extrapp) {
    const int c = lane / NN_TPA, sub = lane % NN_TPA;
    int extrap = 0;
    struct {
        float rc2, pi_rc;
    } M{rc2, pi_rc};
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
        float r, irp;
        nbr_rir(L, base + p, &r, &irp);
        const float fcv = L.fc[base + p];
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
            if (v < Mc.glo[e][k] || v > Mc.ghi[e][k]) extrap = 1;
        }
    }
    // angular (atomsymfunction3Andi): pairs p < q of the neighbour list
    if (i < N && S.f_ok[e]) {
        // factored form: per triplet NN_FB exponentials and NN_FB powers,
        // their outer product accumulated per slot
        const int fcls = S.f_cls[e], nb = S.f_nb[e], nu = S.f_nu[e];
        float etl[NN_FB], lam[NN_FB];
        int zp[NN_FB];
#pragma unroll
        for (int q = 0; q < NN_FB; ++q) {
            etl[q] = S.f_etal2[e][q];
            lam[q] = S.f_lam[e][q];
            zp[q] = S.f_zp[e][q];
        }
        float acc[NN_FB * NN_FB];
#pragma unroll
        for (int q = 0; q < NN_FB * NN_FB; ++q) acc[q] = 0.f;
        for (int p = sub; p < cnt; p += NN_TPA) {
            const float ax_ = L.dx[base + p], ay_ = L.dy[base + p], az_ = L.dz[base + p];
            float ra, ia;
            nbr_rir(L, base + p, &ra, &ia);
            const float fa = L.fc[base + p];
            const int ep = L.el[base + p];
            for (int q = p + 1; q < cnt; ++q) {
                const float bx = L.dx[base + q], by = L.dy[base + q], bz = L.dz[base + q];
                const float cx = bx - ax_, cy = by - ay_, cz = bz - az_;
                const float rc2v = cx * cx + cy * cy + cz * cz;
                if (rc2v > M.rc2) continue;
                const int eq = L.el[base + q];
                const int cls = ep < eq ? ep * NN_MAXE + eq : eq * NN_MAXE + ep;
                if (cls != fcls) continue;
                const float rb2 = bx * bx + by * by + bz * bz;
                const float ib = __builtin_amdgcn_rsqf(rb2), rb = rb2 * ib;
                const float cth = (ax_ * bx + ay_ * by + az_ * bz) * ia * ib;
                float fcc, dfcc;
                cutoff(M.pi_rc, sqrtf(rc2v), &fcc, &dfcc);
                const float Cf = fa * L.fc[base + q] * fcc;
                const float Sd = ra * ra + rb * rb + rc2v;
                float u[NN_FB];
#pragma unroll
                for (int m = 0; m < NN_FB; ++m) {
                    const float x = 1.0f + lam[m] * cth;
                    float pw = x;
#pragma unroll
                    for (int z = 0; z < 4; ++z) pw = z < zp[m] ? pw * pw : pw;  // zeta <= 16
                    u[m] = (m < nu && x > 0.f) ? pw * Cf : 0.f;
                }
#pragma unroll
                for (int n = 0; n < NN_FB; ++n) {
                    const float B = n < nb ? __builtin_amdgcn_exp2f(-etl[n] * Sd) : 0.f;
#pragma unroll
                    for (int m = 0; m < NN_FB; ++m) acc[n * NN_FB + m] += B * u[m];
                }
            }
        }
        // slot q's sum to lane q of the group; function k reads its slot's lane
        float mine = 0.f;
#pragma unroll
        for (int q = 0; q < NN_FB * NN_FB; ++q) {
            const float v = sum_group<NN_TPA>(acc[q]);
            mine = (sub == q) ? v : mine;
        }
#pragma unroll 1
        for (int k = 0; k < NN_MAXA; ++k) {
            const float v = __shfl(mine, (lane & ~(NN_TPA - 1)) + S.f_slot[e][k < na ? k : 0], 64);
            if (k < na && sub == (k & (NN_TPA - 1))) {
                const float g4 = v * S.a_norm[e][k];
                Grep[(size_t)i * NN_K1 + nr + k] = g4;
                if (g4 < Mc.glo[e][nr + k] || g4 > Mc.ghi[e][nr + k]) extrap = 1;
            }
        }
    } else {
        // generic form: every function per triplet (no pair loop for an
        // element without angular functions)
#pragma unroll
    for (int k = 0; k < NN_MAXA; ++k) g[k] = 0.f;
    for (int p = sub; p < (na > 0 ? cnt : 0); p += NN_TPA) {
        const float ax_ = L.dx[base + p], ay_ = L.dy[base + p], az_ = L.dz[base + p];
        float ra, ia;
        nbr_rir(L, base + p, &ra, &ia);
        const float fa = L.fc[base + p];
        const int ep = L.el[base + p];
        for (int q = p + 1; q < cnt; ++q) {
            const float bx = L.dx[base + q], by = L.dy[base + q], bz = L.dz[base + q];
            const float cx = bx - ax_, cy = by - ay_, cz = bz - az_;
            const float rc2v = cx * cx + cy * cy + cz * cz;
            if (rc2v > M.rc2) continue;
            const float rb2 = bx * bx + by * by + bz * bz;
            const float ib = __builtin_amdgcn_rsqf(rb2), rb = rb2 * ib;
            const int eq = L.el[base + q];
            const int cls = ep < eq ? ep * NN_MAXE + eq : eq * NN_MAXE + ep;
            const float cth = (ax_ * bx + ay_ * by + az_ * bz) * ia * ib;
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
            if (v < Mc.glo[e][nr + k] || v > Mc.ghi[e][nr + k]) extrap = 1;
        }
    }
    }
    if (extrap) *extrapp = 1;
}

// dE/dr of the chunk's centre atoms' symmetry functions with coefficients
// cb = dE/dG (VJP): centre force in registers, neighbours' shares into the
// fixed-point accumulators ax, ay, az (Hartree/Bohr).
__device__ void force_chunk(const NnSfTab &S, const ChunkNbr &L, const float (*cb)[NN_K1], float (*Wsh)[NN_FB * NN_FB],
                            const uint8_t *elem, int N, int k0, int lane, float rc2, float pi_rc,
                            unsigned long long *ax, unsigned long long *ay, unsigned long long *az, int *badp) {
    const int c = lane / NN_TPA, sub = lane % NN_TPA;
    int bad = 0;
    struct {
        float rc2, pi_rc;
    } M{rc2, pi_rc};
    const int i = k0 + c;
    const int cnt = L.cnt[c];
    const int e = (i < N) ? elem[i] : 0;
    const int nr = (i < N) ? S.nr[e] : 0, na = (i < N) ? S.na[e] : 0;
    const int base = c * NN_CAP;
    float Fi0 = 0.f, Fi1 = 0.f, Fi2 = 0.f;  // force on the centre (Hartree/Bohr)
    // radial: dG/dr = (-2 eta (r - rs) fc + dfc) exp(-eta (r - rs)^2)
    for (int p = sub; p < cnt; p += NN_TPA) {
        float r, ir;
        nbr_rir(L, base + p, &r, &ir);
        const float fcv = L.fc[base + p], dfcv = L.dfc[base + p];
        const int ep = L.el[base + p];
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < NN_MAXR; ++k)
            if (k < nr && ep == S.r_el[e][k]) {
                const float dr = r - S.r_rs[e][k];
                const float ex = __builtin_amdgcn_exp2f(-S.r_eta[e][k] * NN_L2E * dr * dr);
                s += cb[c][k] * ex * (dfcv - 2.f * S.r_eta[e][k] * dr * fcv);
            }
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
    const bool fast = i < N && S.f_ok[e];
    float W[NN_FB * NN_FB], etl[NN_FB], eta[NN_FB], lam[NN_FB];
    int zp[NN_FB];
    int fcls = -1, nb = 0, nu = 0;
    {
        // W(n, m) = sum over the functions on slot (n, m) of dE/dG_k norm_k,
        // lane `sub` of each centre's group computing slot sub (every lane
        // of the block passes the barrier)
        float wq = 0.f;
        if (fast)
            for (int k = 0; k < na; ++k)
                if (S.f_slot[e][k] == sub) wq += cb[c][nr + k] * S.a_norm[e][k];
        Wsh[c][sub] = wq;
        __syncthreads();
    }
    if (fast) {
        fcls = S.f_cls[e];
        nb = S.f_nb[e];
        nu = S.f_nu[e];
#pragma unroll
        for (int q = 0; q < NN_FB; ++q) {
            etl[q] = S.f_etal2[e][q];
            eta[q] = S.f_eta[e][q];
            lam[q] = S.f_lam[e][q];
            zp[q] = S.f_zp[e][q];
        }
#pragma unroll
        for (int q = 0; q < NN_FB * NN_FB; ++q) W[q] = Wsh[c][q];
    }
    for (int p = sub; p < (na > 0 ? cnt : 0); p += NN_TPA) {
        const float ax_ = L.dx[base + p], ay_ = L.dy[base + p], az_ = L.dz[base + p];
        float ra, ia;
        nbr_rir(L, base + p, &ra, &ia);
        const float fa = L.fc[base + p], dfa = L.dfc[base + p];
        const int ep = L.el[base + p];
        float Fj0 = 0.f, Fj1 = 0.f, Fj2 = 0.f;
        for (int q = p + 1; q < cnt; ++q) {
            const float bx = L.dx[base + q], by = L.dy[base + q], bz = L.dz[base + q];
            const float cx = bx - ax_, cy = by - ay_, cz = bz - az_;
            const float rc2v = cx * cx + cy * cy + cz * cz;
            if (rc2v > M.rc2) continue;
            const int eq = L.el[base + q];
            const int cls = ep < eq ? ep * NN_MAXE + eq : eq * NN_MAXE + ep;
            if (fast && cls != fcls) continue;
            const float rb2 = bx * bx + by * by + bz * bz;
            const float ib = __builtin_amdgcn_rsqf(rb2), rb = rb2 * ib;
            const float fb = L.fc[base + q], dfb = L.dfc[base + q];
            const float cth = (ax_ * bx + ay_ * by + az_ * bz) * ia * ib;
            const float ircc = __builtin_amdgcn_rsqf(rc2v), rcc = rc2v * ircc;
            float fcc, dfcc;
            cutoff(M.pi_rc, rcc, &fcc, &dfcc);
            const float Cf = fa * fb * fcc;
            const float Sd = ra * ra + rb * rb + rc2v;
            float Sg = 0.f, Sc = 0.f, Ss = 0.f;
            if (fast) {
                // Sg = sum B_n W(n,m) u_m, Sc = sum B_n W(n,m) du_m/dcos,
                // Ss = sum eta_n B_n W(n,m) u_m
                float u[NN_FB], du[NN_FB];
#pragma unroll
                for (int m = 0; m < NN_FB; ++m) {
                    const float x = 1.0f + lam[m] * cth;
                    float pw = x, pm = 1.f;  // x^zeta, x^(zeta-1); zeta <= 16
#pragma unroll
                    for (int z = 0; z < 4; ++z) {
                        const bool go = z < zp[m];
                        pm = go ? pm * pw : pm;
                        pw = go ? pw * pw : pw;
                    }
                    const bool on = m < nu && x > 0.f;
                    u[m] = on ? pw : 0.f;
                    du[m] = on ? (float)(1 << zp[m]) * lam[m] * pm : 0.f;
                }
#pragma unroll
                for (int n = 0; n < NN_FB; ++n) {
                    const float B = n < nb ? __builtin_amdgcn_exp2f(-etl[n] * Sd) : 0.f;
                    float t = 0.f, td = 0.f;
#pragma unroll
                    for (int m = 0; m < NN_FB; ++m) {
                        t += W[n * NN_FB + m] * u[m];
                        td += W[n * NN_FB + m] * du[m];
                    }
                    Sg += B * t;
                    Sc += B * td;
                    Ss += eta[n] * B * t;
                }
            } else {
                float B = 0.f, etaprev = -1.f;
#pragma unroll
                for (int k = 0; k < NN_MAXA; ++k)
                    if (k < na && cls == S.a_cls[e][k]) {
                        const float lamk = S.a_lam[e][k];
                        const float x = 1.0f + lamk * cth;
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
                            Sc += ck * zeta * lamk * pm;
                            Ss += ck * S.a_eta[e][k] * pw;
                        }
                    }
            }
            if (Sg == 0.f && Sc == 0.f) continue;
            const float iab = ia * ib;
            const float tc = Cf * Sc, ts = -2.f * Cf * Ss;
            const float ua = Sg * dfa * fb * fcc * ia, ub = Sg * fa * dfb * fcc * ib;
            const float uc = Sg * fa * fb * dfcc * ircc;
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
    if (bad) *badp |= bad;
}

// u_ik = sum_a dG_ik/dr_a . g_a for the chunk's centre atoms i (the
// Jacobian-vector product of the training gradient; g per atom in LDS).
// Per neighbour pair: g_j - g_i projected on the pair direction; per
// neighbour triple: the three difference vectors dotted with the angular
// function's partial vectors, shared over all functions of the triple.
// Result scaled by -dfac (u = dLoss/d(dE/dx)) into Urep rows.
__device__ void jvp_chunk(const NnSfTab &S, const ChunkNbr &L, const uint8_t *elem, int N, int k0, int lane,
                          float rc2, float pi_rc, const float *gx, const float *gy, const float *gz,
                          const float (*dfac)[NN_K1], float *Urep) {
    const int c = lane / NN_TPA, sub = lane % NN_TPA;
    const int i = k0 + c;
    const int cnt = L.cnt[c];
    const int e = (i < N) ? elem[i] : 0;
    const int nr = (i < N) ? S.nr[e] : 0, na = (i < N) ? S.na[e] : 0;
    const int base = c * NN_CAP;
    const float gix = (i < N) ? gx[i] : 0.f, giy = (i < N) ? gy[i] : 0.f, giz = (i < N) ? gz[i] : 0.f;
    float P[NN_MAXR > NN_MAXA ? NN_MAXR : NN_MAXA];
#pragma unroll
    for (int k = 0; k < NN_MAXR; ++k) P[k] = 0.f;
    for (int p = sub; p < cnt; p += NN_TPA) {
        float r, ir;
        nbr_rir(L, base + p, &r, &ir);
        const float fcv = L.fc[base + p], dfcv = L.dfc[base + p];
        const int j = L.j[base + p];
        const float proj = (L.dx[base + p] * (gx[j] - gix) + L.dy[base + p] * (gy[j] - giy) +
                            L.dz[base + p] * (gz[j] - giz)) * ir;
        const int ep = L.el[base + p];
#pragma unroll
        for (int k = 0; k < NN_MAXR; ++k)
            if (k < nr && ep == S.r_el[e][k]) {
                const float dr = r - S.r_rs[e][k];
                const float ex = __builtin_amdgcn_exp2f(-S.r_eta[e][k] * NN_L2E * dr * dr);
                P[k] += ex * (dfcv - 2.f * S.r_eta[e][k] * dr * fcv) * proj;
            }
    }
#pragma unroll
    for (int k = 0; k < NN_MAXR; ++k) {
        const float v = sum_group<NN_TPA>(P[k]);
        if (i < N && k < nr && sub == (k & (NN_TPA - 1))) Urep[(size_t)i * NN_K1 + k] = -dfac[e][k] * v;
    }
#pragma unroll
    for (int k = 0; k < NN_MAXA; ++k) P[k] = 0.f;
    for (int p = sub; p < (na > 0 ? cnt : 0); p += NN_TPA) {
        const float ax_ = L.dx[base + p], ay_ = L.dy[base + p], az_ = L.dz[base + p];
        float ra, ia;
        nbr_rir(L, base + p, &ra, &ia);
        const float fa = L.fc[base + p], dfa = L.dfc[base + p];
        const int ep = L.el[base + p];
        const int jp = L.j[base + p];
        const float Dja0 = gx[jp] - gix, Dja1 = gy[jp] - giy, Dja2 = gz[jp] - giz;
        for (int q = p + 1; q < cnt; ++q) {
            const float bx = L.dx[base + q], by = L.dy[base + q], bz = L.dz[base + q];
            const float cx = bx - ax_, cy = by - ay_, cz = bz - az_;
            const float rc2v = cx * cx + cy * cy + cz * cz;
            if (rc2v > rc2) continue;
            const float rb2 = bx * bx + by * by + bz * bz;
            const float ib = __builtin_amdgcn_rsqf(rb2), rb = rb2 * ib;
            const float fb = L.fc[base + q], dfb = L.dfc[base + q];
            const int eq = L.el[base + q];
            const int cls = ep < eq ? ep * NN_MAXE + eq : eq * NN_MAXE + ep;
            const int kq = L.j[base + q];
            const float Dka0 = gx[kq] - gix, Dka1 = gy[kq] - giy, Dka2 = gz[kq] - giz;
            const float Dkj0 = Dka0 - Dja0, Dkj1 = Dka1 - Dja1, Dkj2 = Dka2 - Dja2;
            const float cth = (ax_ * bx + ay_ * by + az_ * bz) * ia * ib;
            const float ircc = __builtin_amdgcn_rsqf(rc2v), rcc = rc2v * ircc;
            float fcc, dfcc;
            cutoff(pi_rc, rcc, &fcc, &dfcc);
            const float Cf = fa * fb * fcc;
            const float Sd = ra * ra + rb * rb + rc2v;
            const float iab = ia * ib;
            const float aD = ax_ * Dja0 + ay_ * Dja1 + az_ * Dja2;   // a . Dja
            const float bD = bx * Dka0 + by * Dka1 + bz * Dka2;      // b . Dka
            const float cD = cx * Dkj0 + cy * Dkj1 + cz * Dkj2;      // c . Dkj
            const float bDa = bx * Dja0 + by * Dja1 + bz * Dja2;     // b . Dja
            const float aDb = ax_ * Dka0 + ay_ * Dka1 + az_ * Dka2;  // a . Dka
            const float dcosD = (bDa * iab - cth * aD * ia * ia) + (aDb * iab - cth * bD * ib * ib);
            const float SD = aD + bD + cD;
            const float FD = dfa * fb * fcc * aD * ia + fa * dfb * fcc * bD * ib + fa * fb * dfcc * cD * ircc;
            float B = 0.f, etaprev = -1.f;
#pragma unroll
            for (int k = 0; k < NN_MAXA; ++k)
                if (k < na && cls == S.a_cls[e][k]) {
                    const float lam = S.a_lam[e][k];
                    const float x = 1.0f + lam * cth;
                    if (x > 0.f) {
                        float pw = x, pm = 1.f;
                        for (int z = 0; z < S.a_zp[e][k]; ++z) {
                            pm *= pw;
                            pw *= pw;
                        }
                        if (S.a_etal2[e][k] != etaprev) {
                            etaprev = S.a_etal2[e][k];
                            B = __builtin_amdgcn_exp2f(-etaprev * Sd);
                        }
                        const float zeta = (float)(1 << S.a_zp[e][k]);
                        P[k] += S.a_norm[e][k] * B *
                                (zeta * lam * pm * Cf * dcosD - 2.f * S.a_eta[e][k] * pw * Cf * SD + pw * FD);
                    }
                }
        }
    }
#pragma unroll
    for (int k = 0; k < NN_MAXA; ++k) {
        const float v = sum_group<NN_TPA>(P[k]);
        if (i < N && k < na && sub == (k & (NN_TPA - 1))) Urep[(size_t)i * NN_K1 + nr + k] = -dfac[e][nr + k] * v;
    }
}

// LDS copy of the symmetry-function tables (the extrapolation bounds are read
// from the constants, once per atom)
__device__ void load_sftab(NnSfTab &S, const NnConst &M, int lane) {
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
    // factored angular tables, one lane per element (from M: no dependence
    // on the copies above)
    if (lane < NN_MAXE) {
        const int e = lane, na = M.na[e];
        int ok = na > 0, nb = 0, nu = 0, cls = na > 0 ? M.a_e1[e][0] * NN_MAXE + M.a_e2[e][0] : -1;
        int used = 0;  // bit per slot
        for (int k = 0; k < na && ok; ++k) {
            if (M.a_e1[e][k] * NN_MAXE + M.a_e2[e][k] != cls) ok = 0;
            int n = 0, m = 0;
            while (n < nb && S.f_etal2[e][n] != M.a_etal2[e][k]) ++n;
            if (n == nb) {
                if (nb == NN_FB) {
                    ok = 0;
                    break;
                }
                S.f_etal2[e][nb] = M.a_etal2[e][k];
                S.f_eta[e][nb] = M.a_etal2[e][k] * (1.0f / NN_L2E);
                ++nb;
            }
            while (m < nu && !(S.f_lam[e][m] == M.a_lam[e][k] && S.f_zp[e][m] == M.a_zp[e][k])) ++m;
            if (m == nu) {
                if (nu == NN_FB) {
                    ok = 0;
                    break;
                }
                S.f_lam[e][nu] = M.a_lam[e][k];
                S.f_zp[e][nu] = M.a_zp[e][k];
                ++nu;
            }
            const int slot = n * NN_FB + m;
            if (used >> slot & 1) ok = 0;  // two functions on one slot
            used |= 1 << slot;
            S.f_slot[e][k] = slot;
        }
        for (int n = nb; n < NN_FB; ++n) S.f_etal2[e][n] = S.f_eta[e][n] = 0.f;
        for (int m = nu; m < NN_FB; ++m) {
            S.f_lam[e][m] = 0.f;
            S.f_zp[e][m] = 0;
        }
        S.f_ok[e] = ok;
        S.f_cls[e] = cls;
        S.f_nb[e] = nb;
        S.f_nu[e] = nu;
    }
}

// cell in Bohr (columns = lattice vectors), its inverse in 1/Angstrom, and
// per axis the fractional half-width of the cut-off sphere rc |row k of H^-1|
// and the largest image shift to try
__device__ void cell_setup(const NnConst &M, const double *cellA, const double *icellA, double h[9], double ih[9],
                           float w[3], int imax[3]) {
    for (int k = 0; k < 9; ++k) {
        h[k] = cellA[k] * (double)M.ang2bohr;
        ih[k] = icellA[k];
    }
    const double rcb = (double)M.rc;
    for (int k = 0; k < 3; ++k) {
        const double bk = sqrt(ih[k] * ih[k] + ih[k + 3] * ih[k + 3] + ih[k + 6] * ih[k + 6]) / (double)M.ang2bohr;
        w[k] = (float)(rcb * bk);
        imax[k] = (int)ceil(rcb * bk) + 1;
    }
}

// Atoms grouped by element, ascending index within each group, into elist
// (wave 0; the caller synchronises); s_ne[e]: the group sizes.
__device__ void group_elements(const uint8_t *elem, int N, int ne_model, uint16_t *elist, int *s_ne, int lane) {
    if (lane >= 64) return;
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

// The network on every atom of the replica (per element, 16-atom tiles, one
// per wavefront per round: atomic energies to Ea, c = dE/dG rows to Crep)
// and the replica's energy [eV] to *epot (fixed-order sum).  Every lane of
// the block calls this (barriers inside).
__device__ void network_energy(const NnConst &M, const int *s_ne, const uint16_t *elist, int (*rowsb)[16],
                               float (*T)[16 * 33], const float *Grep, float *Crep, double *Ea, int N, int lane,
                               double *epot) {
    const int wave = lane >> 6;
    const int ne_model = M.nelem;
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
    if (lane < 64) {
        double e = 0.0;
        for (int i = lane; i < N; i += 64) e += Ea[i];
        for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
        if (lane == 0) *epot = e * NN_HA2EV;
    }
}

// Outputs of pass F from the fixed-point accumulators: f [eV/A] (+ the
// forces already in f when add_fin; the network's alone to nv.fpes), a = f/m
// (fixed components 0), dens = 0 (no EMT density).
__device__ void write_forces(const BatchView &bv, const NnView &nv, const unsigned long long *ax,
                             const unsigned long long *ay, const unsigned long long *az, int N, long long roff,
                             long long doff, long long aoff, int add_fin, int lane) {
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
}

// Pass G and the network: raw G rows, c = dE/dG rows, the replica's energy
// and extrapolation flag.  A replica outside the training range leaves
// with its flag set when skip_extrap (k_nene_f skips it too).
__global__ void __launch_bounds__(NN_BLOCK) k_nene_g(const NnConst *__restrict__ Mg, BatchView bv, NnView nv,
                                                     int skip_extrap) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ NnSfTab S;
    // the neighbour cache of pass G and the network's transposes share LDS
    __shared__ union {
        ChunkNbr L;
        float T[NN_BLOCK / 64][16 * 33];
    } U;
    ChunkNbr &L = U.L;
    float(*T)[16 * 33] = U.T;
    __shared__ int rowsb[NN_BLOCK / 64][16];
    __shared__ int s_ne[NN_MAXE];
    const NnConst &M = *Mg;
    const int rep = bv.rep_list ? bv.rep_list[blockIdx.x] : (int)blockIdx.x;  // size-class launches
    const int t = bv.rep_traj[rep];
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int b = bv.rep_bead[rep];
    const int N = bv.traj_N[t];
    const long long soff = bv.traj_soff[t];
    const long long roff = soff + (long long)b * 3 * N;
    const long long aoff = bv.traj_aoff[t];
    const int lane = threadIdx.x;
    const int wave = lane >> 6;
    float *Grep = nv.G + (size_t)rep * nv.maxN * NN_K1;
    float *Crep = nv.C + (size_t)rep * nv.maxN * NN_K1;
    int bad = 0, extrap = 0;

    // dynamic LDS: Ea (double N) | ux uy uz (u32 N each) | elist (u16 N) | elem (u8 N)
    double *Ea = (double *)smem_raw;
    uint32_t *ux = (uint32_t *)(Ea + N), *uy = ux + N, *uz = uy + N;
    uint16_t *elist = (uint16_t *)(uz + N);
    uint8_t *elem = (uint8_t *)(elist + N);

    // ---- tables ----
    const int ne_model = M.nelem;
    load_sftab(S, M, lane);
    double h[9], ih[9];
    float w[3];
    int imax[3];
    cell_setup(M, bv.cell + 9 * t, bv.icell + 9 * t, h, ih, w, imax);

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
        ux[i] = frac_fix(f0 - floor(f0));
        uy[i] = frac_fix(f1 - floor(f1));
        uz[i] = frac_fix(f2 - floor(f2));
        Ea[i] = 0.0;
    }
    __syncthreads();
    group_elements(elem, N, ne_model, elist, s_ne, lane);
    __syncthreads();

    // ---- pass G: symmetry functions ----
    for (int k0 = 0; k0 < N; k0 += NN_CA) {
        build_chunk(M.rc2, M.pi_rc, L, ux, uy, uz, elem, h, imax, w[0], w[1], w[2], N, k0, lane, &bad);
        __syncthreads();
        g_chunk(M, S, L, elem, N, k0, lane, M.rc2, M.pi_rc, Grep, &extrap);
        __syncthreads();
    }

    // replicas outside the training range go to the EMT fallback untouched
    extrap = __syncthreads_or(extrap);
    if (lane == 0) nv.rep_flag[rep] = extrap;
    if (extrap && skip_extrap) {
        if (lane == 0)
            atomicOr(&bv.status[t], bad | MDT_ST_NN_EXTRAPOLATION);
        else if (bad)
            atomicOr(&bv.status[t], bad);
        return;
    }

    network_energy(M, s_ne, elist, rowsb, T, Grep, Crep, Ea, N, lane, &bv.epot[bv.traj_eoff[t] + b]);
    if (extrap) bad |= MDT_ST_NN_EXTRAPOLATION;
    if (bad) atomicOr(&bv.status[t], bad);
}

// Pass F and the outputs: dE/dr from the c rows of k_nene_g, f (+ the
// forces already in f when add_fin), a = f/m, dens = 0.
__global__ void __launch_bounds__(NN_BLOCK) k_nene_f(const NnConst *__restrict__ Mg, BatchView bv, NnView nv,
                                                     int add_fin, int skip_extrap) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ NnSfTab S;
    __shared__ ChunkNbr L;
    __shared__ float cb[NN_CA][NN_K1];
    __shared__ float Wsh[NN_CA][NN_FB * NN_FB];
    const NnConst &M = *Mg;
    const int rep = bv.rep_list ? bv.rep_list[blockIdx.x] : (int)blockIdx.x;  // size-class launches
    if (skip_extrap && nv.rep_flag[rep]) return;  // whole block: uniform
    const int t = bv.rep_traj[rep];
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int b = bv.rep_bead[rep];
    const int N = bv.traj_N[t];
    const long long soff = bv.traj_soff[t];
    const long long roff = soff + (long long)b * 3 * N;
    const long long doff = soff / 3 + (long long)b * N;
    const long long aoff = bv.traj_aoff[t];
    const int lane = threadIdx.x;
    const float *Crep = nv.C + (size_t)rep * nv.maxN * NN_K1;
    int bad = 0;

    // dynamic LDS: ax ay az (u64 N each) | ux uy uz (u32 N) | elem (u8 N)
    unsigned long long *ax = (unsigned long long *)smem_raw, *ay = ax + N, *az = ay + N;
    uint32_t *ux = (uint32_t *)(az + N), *uy = ux + N, *uz = uy + N;
    uint8_t *elem = (uint8_t *)(uz + N);

    const int ne_model = M.nelem;
    load_sftab(S, M, lane);
    double h[9], ih[9];
    float w[3];
    int imax[3];
    cell_setup(M, bv.cell + 9 * t, bv.icell + 9 * t, h, ih, w, imax);
    for (int i = lane; i < N; i += NN_BLOCK) {
        const int e = M.elem_of_type[bv.type[aoff + i]];
        elem[i] = (uint8_t)(e < 0 || e >= ne_model ? 0 : e);
        const double r0 = bv.x[roff + i], r1 = bv.x[roff + N + i], r2 = bv.x[roff + 2 * N + i];
        double f0 = ih[0] * r0 + ih[3] * r1 + ih[6] * r2;
        double f1 = ih[1] * r0 + ih[4] * r1 + ih[7] * r2;
        double f2 = ih[2] * r0 + ih[5] * r1 + ih[8] * r2;
        ux[i] = frac_fix(f0 - floor(f0));
        uy[i] = frac_fix(f1 - floor(f1));
        uz[i] = frac_fix(f2 - floor(f2));
        ax[i] = ay[i] = az[i] = 0ull;
    }
    __syncthreads();

    // ---- pass F: forces ----
    for (int k0 = 0; k0 < N; k0 += NN_CA) {
        build_chunk(M.rc2, M.pi_rc, L, ux, uy, uz, elem, h, imax, w[0], w[1], w[2], N, k0, lane, &bad);
        for (int q = lane; q < NN_CA * NN_K1; q += NN_BLOCK) {
            const int ci = q / NN_K1, k = q % NN_K1;
            const int i = k0 + ci;
            cb[ci][k] = (i < N && k < S.nr[elem[i]] + S.na[elem[i]]) ? Crep[(size_t)i * NN_K1 + k] : 0.f;
        }
        __syncthreads();
        force_chunk(S, L, cb, Wsh, elem, N, k0, lane, M.rc2, M.pi_rc, ax, ay, az, &bad);
        __syncthreads();
    }

    write_forces(bv, nv, ax, ay, az, N, roff, doff, aoff, add_fin, lane);
    if (bad) atomicOr(&bv.status[t], bad);
}


// ===========================================================================
// Radial-only networks (no element has angular functions): k_nene_rg /
// k_nene_rf run the passes of k_nene_g / k_nene_f shaped for radial sums.
//  * Centres go in element order, each element's atoms padded to a multiple
//    of 4, so the 4 centres of a wavefront share one element and every
//    function parameter is wavefront-uniform (scalar loads, no LDS tables).
//  * A function term exp2(ce (r - rs)^2) fc(r) is branch-free; the
//    neighbour-element test is a select, and functions on neighbour
//    elements that none of the wavefront's current entries has are skipped
//    (ballot).
//  * List entries hold r, the neighbour element and, for pass F, the unit
//    vector and the neighbour; the cut-off is evaluated per entry.
//  * When every perpendicular width of the cell exceeds 2 rc, the only
//    image of an atom within rc is the minimum image of the fractional
//    difference: one pass over the candidates, hits compacted in ascending
//    atom order with a wavefront ballot (no count pass, no image loops).
//    Other cells take the image search of build_images.
//  * Sums over a centre's 16 lanes use DPP row permutations (no LDS).
// ===========================================================================

#define NN_WF 32  // pass F entries per centre and window (a longer list takes several windows)

// entries of one centre's list staged at a time: cnt those of this window,
// tot the centre's whole list
struct RadNbrG {  // pass G entries (the whole list)
    static constexpr int W = NN_CAP;
    float r[NN_CA * NN_CAP];
    uint8_t el[NN_CA * NN_CAP];
    int cnt[NN_CA], tot[NN_CA];
};

struct RadNbrF {  // pass F entries (a window of NN_WF)
    static constexpr int W = NN_WF;
    float r[NN_CA * NN_WF], ux[NN_CA * NN_WF], uy[NN_CA * NN_WF], uz[NN_CA * NN_WF];
    uint16_t j[NN_CA * NN_WF];
    uint8_t el[NN_CA * NN_WF];
    int cnt[NN_CA], tot[NN_CA];
};

__device__ __forceinline__ void rad_put(RadNbrG &L, int e, int n, float ex, float ey, float ez, float r2,
                                        const uint8_t *elem) {
    (void)ex;
    (void)ey;
    (void)ez;
    L.r[e] = r2 * __builtin_amdgcn_rsqf(r2);
    L.el[e] = elem[n];
}

__device__ __forceinline__ void rad_put(RadNbrF &L, int e, int n, float ex, float ey, float ez, float r2,
                                        const uint8_t *elem) {
    const float ir = __builtin_amdgcn_rsqf(r2);
    L.r[e] = r2 * ir;
    L.ux[e] = ex * ir;
    L.uy[e] = ey * ir;
    L.uz[e] = ez * ir;
    L.j[e] = (uint16_t)n;
    L.el[e] = elem[n];
}

// sum over the 16 lanes of a DPP row (row mirror, half-row mirror, quad
// swaps 2 and 1); every lane of the row receives the same sum
__device__ __forceinline__ float row_sum16(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4e, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xb1, 0xf, 0xf, false));
    return v;
}

// Vector from atom i (fractional xi, yi, zi) to atom n under the minimum
// image: the wrapped 32-bit differences are 2^32 times the fractional
// difference in [-1/2, 1/2); hs is the cell in Bohr times 2^-32 (FP32,
// columns = lattice vectors); TRI: a1 along x, a2 in the xy plane, a3 along
// z (the slab cells), whose zero entries are skipped.
template <bool TRI>
__device__ __forceinline__ void min_image_d(const float hs[9], const uint32_t *ux, const uint32_t *uy,
                                            const uint32_t *uz, int n, uint32_t xi, uint32_t yi, uint32_t zi,
                                            float *ex, float *ey, float *ez) {
    const float f0 = (float)(int)(ux[n] - xi);
    const float f1 = (float)(int)(uy[n] - yi);
    const float f2 = (float)(int)(uz[n] - zi);
    if (TRI) {
        *ex = hs[0] * f0 + hs[3] * f1;
        *ey = hs[4] * f1;
        *ez = hs[8] * f2;
    } else {
        *ex = hs[0] * f0 + hs[3] * f1 + hs[6] * f2;
        *ey = hs[1] * f0 + hs[4] * f1 + hs[7] * f2;
        *ez = hs[2] * f0 + hs[5] * f1 + hs[8] * f2;
    }
}

// Neighbour list of centre i (16 lanes; i < 0: none) within sqrt(rb2) when
// the cell admits only the minimum image: candidates n = it*16 + sub, hits
// compacted in ascending n with a wavefront ballot.  nrow non-NULL: the
// neighbour indices also go to nrow and the length to *ncnt.  Returns the
// (clamped) count in every lane of the group; more than NN_CAP hits set the
// overflow flag.
template <bool TRI, class Lt>
__device__ int rad_build_min_image(Lt &L, const uint32_t *ux, const uint32_t *uy, const uint32_t *uz,
                                   const uint8_t *elem, const float hs[9], float rb2, int N, int i, int lane,
                                   uint16_t *nrow, uint8_t *ncnt, int *bad, int w0) {
    constexpr int W = Lt::W;
    const int c = lane >> 4, sub = lane & 15, sh = lane & 48;
    const bool act = i >= 0;
    const uint32_t xi = ux[act ? i : 0], yi = uy[act ? i : 0], zi = uz[act ? i : 0];
    int cnt = 0;
    for (int n0 = 0; n0 < N; n0 += NN_TPA) {
        const int n = n0 + sub;
        bool hit = false;
        float ex = 0.f, ey = 0.f, ez = 0.f, r2 = 0.f;
        if (act && n < N && n != i) {
            min_image_d<TRI>(hs, ux, uy, uz, n, xi, yi, zi, &ex, &ey, &ez);
            r2 = ex * ex + ey * ey + ez * ez;
            hit = r2 <= rb2;
        }
        const unsigned int row = (unsigned int)(__ballot(hit) >> sh) & 0xffffu;
        const int pos = cnt + __popc(row & ((1u << sub) - 1u));
        if (hit && pos < NN_CAP) {
            if (pos >= w0 && pos < w0 + W) rad_put(L, c * W + pos - w0, n, ex, ey, ez, r2, elem);
            if (nrow) nrow[pos] = (uint16_t)n;
        }
        cnt += __popc(row);
    }
    if (cnt > NN_CAP) *bad |= MDT_ST_NBR_OVERFLOW;
    cnt = cnt < NN_CAP ? cnt : NN_CAP;
    if (nrow && act && sub == 0) ncnt[i] = (uint8_t)cnt;
    return cnt;
}

// A stored list of centre i: entries recomputed from the staged positions
// with the arithmetic of rad_build_min_image (bitwise the same vectors).
template <bool TRI, class Lt>
__device__ int rad_load(Lt &L, const uint32_t *ux, const uint32_t *uy, const uint32_t *uz, const uint8_t *elem,
                        const float hs[9], int i, int lane, const uint16_t *nrow, const uint8_t *ncnt, int w0) {
    constexpr int W = Lt::W;
    const int c = lane >> 4, sub = lane & 15;
    const int cnt = i >= 0 ? (int)ncnt[i] : 0;
    if (i >= 0) {
        const uint32_t xi = ux[i], yi = uy[i], zi = uz[i];
        const int pe = cnt < w0 + W ? cnt : w0 + W;
        for (int p = w0 + sub; p < pe; p += NN_TPA) {
            const int n = nrow[p];
            float ex, ey, ez;
            min_image_d<TRI>(hs, ux, uy, uz, n, xi, yi, zi, &ex, &ey, &ez);
            rad_put(L, c * W + p - w0, n, ex, ey, ez, ex * ex + ey * ey + ez * ez, elem);
        }
    }
    return cnt;
}

// The replica's cell for the list builds, kept in LDS so that its registers
// are live only during a build: h in Bohr (FP64 and FP32), the fractional
// half-widths of the cut-off sphere, the image bounds, and whether the
// minimum image alone suffices (every half-width below 1/2).
struct RadCell {
    double h[9];
    float hs[9], w[3];  // hs: h 2^-32 (FP32)
    int imax[3];
    int min_image, tri;
    int store;  // lists stored (minimum image within rc + skin)
};

__device__ void rad_cell(RadCell &C, const NnConst &M, const double *cellA, const double *icellA, double ih[9],
                         bool lists, float skin, int lane) {
    double h[9];
    float w[3];
    int imax[3];
    cell_setup(M, cellA, icellA, h, ih, w, imax);
    if (lane == 0) {
        for (int k = 0; k < 9; ++k) {
            C.h[k] = h[k];
            C.hs[k] = (float)h[k] * 0x1p-32f;
        }
        for (int k = 0; k < 3; ++k) {
            C.w[k] = w[k];
            C.imax[k] = imax[k];
        }
        C.min_image = w[0] < 0.4999f && w[1] < 0.4999f && w[2] < 0.4999f;
        C.tri = h[1] == 0.0 && h[2] == 0.0 && h[5] == 0.0 && h[6] == 0.0 && h[7] == 0.0;
        const float ws = 1.0f + skin / M.rc;  // half-widths of the sphere of radius rc + skin
        C.store = lists && w[0] * ws < 0.4999f && w[1] * ws < 0.4999f && w[2] * ws < 0.4999f;
    }
}

// The chunk's lists: with stored lists (C.store) read them (mode 1) or
// build them within rc + skin and store them (mode 0); otherwise the
// minimum-image pass or the image search within rc.  IMG false: the image
// search is not compiled in (its FP64 registers are then not reserved); a
// cell that needs it sets the overflow flag and gets no entries, which the
// host rules out by launching IMG for every size class with such a cell.
template <class Lt, bool IMG = true>
__device__ void rad_build(Lt &L, const RadCell &C, const NnView &nv, int rep, int mode, const uint32_t *ux,
                          const uint32_t *uy, const uint32_t *uz, const uint8_t *elem, float rc2, float rb2, int N,
                          int i, int lane, int *bad, int w0 = 0) {
    constexpr int W = Lt::W;
    const int c = lane >> 4;
    int cnt;
    if (__builtin_amdgcn_readfirstlane(C.store)) {
        float hs[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) hs[k] = C.hs[k];
        uint16_t *nrow = nv.nbr + ((size_t)rep * nv.maxN + (i >= 0 ? i : 0)) * NN_CAP;
        uint8_t *ncnt = nv.nbr_cnt + (size_t)rep * nv.maxN;
        const bool tri = __builtin_amdgcn_readfirstlane(C.tri);
        if (mode == 1)
            cnt = tri ? rad_load<true>(L, ux, uy, uz, elem, hs, i, lane, nrow, ncnt, w0)
                      : rad_load<false>(L, ux, uy, uz, elem, hs, i, lane, nrow, ncnt, w0);
        else
            cnt = tri ? rad_build_min_image<true>(L, ux, uy, uz, elem, hs, rb2, N, i, lane, nrow, ncnt, bad, w0)
                      : rad_build_min_image<false>(L, ux, uy, uz, elem, hs, rb2, N, i, lane, nrow, ncnt, bad, w0);
    } else if (__builtin_amdgcn_readfirstlane(C.min_image)) {
        float hs[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) hs[k] = C.hs[k];
        if (__builtin_amdgcn_readfirstlane(C.tri))
            cnt = rad_build_min_image<true>(L, ux, uy, uz, elem, hs, rc2, N, i, lane, nullptr, nullptr, bad, w0);
        else
            cnt = rad_build_min_image<false>(L, ux, uy, uz, elem, hs, rc2, N, i, lane, nullptr, nullptr, bad, w0);
    } else if constexpr (IMG) {
        double h[9];
        int imax[3];
#pragma unroll
        for (int k = 0; k < 9; ++k) h[k] = C.h[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) imax[k] = C.imax[k];
        cnt = build_images(rc2, ux, uy, uz, h, imax, C.w[0], C.w[1], C.w[2], N, i, lane, bad,
                           [&](int pos, int n, double ex, double ey, double ez, double r2) {
                               if (pos >= w0 && pos < w0 + W)
                                   rad_put(L, c * W + pos - w0, n, (float)ex, (float)ey, (float)ez, (float)r2,
                                           elem);
                           });
        cnt = cnt < NN_CAP ? cnt : NN_CAP;
    } else {
        *bad |= MDT_ST_NBR_OVERFLOW;
        cnt = 0;
    }
    if ((lane & 15) == 0) {
        const int wc = cnt - w0;
        L.cnt[c] = wc < 0 ? 0 : (wc < W ? wc : W);  // this window's entries
        L.tot[c] = cnt;
    }
}

// the largest list count of the wavefront's 4 centres (uniform)
__device__ __forceinline__ int wave_max_cnt(int cnt) {
    const int a = max(__builtin_amdgcn_readlane(cnt, 0), __builtin_amdgcn_readlane(cnt, 16));
    const int b = max(__builtin_amdgcn_readlane(cnt, 32), __builtin_amdgcn_readlane(cnt, 48));
    return max(a, b);
}

// Centre schedule of the radial kernels: each element's atoms (elist
// order) padded to a multiple of 4, entries (e << 16) | atom (atom 0xffff
// for padding).  All lanes of the block; returns the slot count.
__device__ int make_slots(const uint16_t *elist, const int *s_ne, int ne_model, uint32_t *slot, int lane) {
    int so = 0, start = 0;
    for (int e = 0; e < ne_model; ++e) {
        const int ne = s_ne[e], padded = (ne + 3) & ~3;
        for (int q = lane; q < padded; q += NN_BLOCK)
            slot[so + q] = ((uint32_t)e << 16) | (q < ne ? (uint32_t)elist[start + q] : 0xffffu);
        so += padded;
        start += ne;
    }
    return so;
}

#define NN_RQ (NN_CAP / NN_TPA)  // list entries per lane (6)

// The lane's list entries p = sub + 16 q into registers: r (rc past the
// count), the neighbour element (-1 past the count); *nq: the wavefront's
// largest number of entries per lane (uniform), *present: bit x set when
// some entry of the wavefront has element x (uniform).
template <class Lt>
__device__ __forceinline__ void rad_entries(const Lt &L, int lane, float rc, int nel, float rr[Lt::W / NN_TPA],
                                            int eq[Lt::W / NN_TPA], int *nq, unsigned int *present) {
    constexpr int RQ = Lt::W / NN_TPA;
    const int c = lane >> 4, sub = lane & 15;
    const int cnt = L.cnt[c], base = c * Lt::W;
    *nq = (wave_max_cnt(cnt) + NN_TPA - 1) / NN_TPA;
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        const int p = sub + NN_TPA * q;
        const float r = p < cnt ? L.r[base + p] : rc;
        const bool v = p < cnt && r <= rc;  // stored lists reach rc + skin
        rr[q] = v ? r : rc;
        eq[q] = v ? (int)L.el[base + p] : -1;
    }
    unsigned int pr = 0;
    for (int x = 0; x < nel; ++x) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < RQ; ++q) any |= eq[q] == x;
        if (__ballot(any)) pr |= 1u << x;
    }
    *present = pr;
}

// Functions lo .. hi-1 (kernel order, one neighbour element) of element e
// on the lane's first NQ entries: each function's 16-lane sum lands in the
// lane that owns it (kernel index sub -> g0, sub + 16 -> g1).  The next
// function's parameters load while the current one is evaluated.
template <int NQ>
__device__ __forceinline__ void rg_class(const NnConst &M, int e, int lo, int hi, const float rr[NN_RQ],
                                         const float fx[NN_RQ], int sub, float &g0, float &g1) {
    float rs = M.rk_rs[e][lo], ce = M.rk_ce[e][lo];
    for (int k = lo; k < hi; ++k) {
        const int kn = k + 1 < hi ? k + 1 : k;
        const float rsn = M.rk_rs[e][kn], cen = M.rk_ce[e][kn];
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const float dr = rr[q] - rs;
            acc = fmaf(__builtin_amdgcn_exp2f(ce * dr * dr), fx[q], acc);
        }
        const float v = row_sum16(acc);
        const bool mine = sub == (k & (NN_TPA - 1));
        g0 = (mine && k < NN_TPA) ? v : g0;
        g1 = (mine && k >= NN_TPA) ? v : g1;
        rs = rsn;
        ce = cen;
    }
}

// Radial symmetry functions of the chunk's centres (atom i of the lane's
// group, element e uniform over the wavefront) into Grep rows; *extrapp set
// when a value leaves [min - 1e-4, max + 1e-4].  Functions outer (their
// parameters scalar), the lane's entries inner, instantiated for the
// wavefront's number of entries per lane.
__device__ void rg_chunk(const NnConst &M, const RadNbrG &L, int e, int i, int lane, float *Grep, int *extrapp) {
    const int sub = lane & 15;
    const int nel = M.nelem, nr = M.nr[e];
    const float pi_rc = M.pi_rc;
    float rr[NN_RQ], fcq[NN_RQ];
    int eq[NN_RQ], nq;
    unsigned int present;
    rad_entries(L, lane, M.rc, nel, rr, eq, &nq, &present);
#pragma unroll
    for (int q = 0; q < NN_RQ; ++q) {
        float fc, dfc;
        cutoff(pi_rc, rr[q], &fc, &dfc);
        fcq[q] = eq[q] >= 0 ? fc : 0.f;
    }
    float g0 = 0.f, g1 = 0.f;  // functions of no present neighbour element stay 0
    for (int x = 0; x < nel; ++x) {
        const int lo = M.rk_lo[e][x], hi = M.rk_hi[e][x];
        if (lo == hi || !((present >> x) & 1u)) continue;
        float fx[NN_RQ];
#pragma unroll
        for (int q = 0; q < NN_RQ; ++q) fx[q] = eq[q] == x ? fcq[q] : 0.f;
        switch (nq) {
            case 1: rg_class<1>(M, e, lo, hi, rr, fx, sub, g0, g1); break;
            case 2: rg_class<2>(M, e, lo, hi, rr, fx, sub, g0, g1); break;
            case 3: rg_class<3>(M, e, lo, hi, rr, fx, sub, g0, g1); break;
            case 4: rg_class<4>(M, e, lo, hi, rr, fx, sub, g0, g1); break;
            case 5: rg_class<5>(M, e, lo, hi, rr, fx, sub, g0, g1); break;
            case 6: rg_class<6>(M, e, lo, hi, rr, fx, sub, g0, g1); break;
            default: break;
        }
    }
    int extrap = 0;
    if (i >= 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = sub + NN_TPA * h;
            if (k < nr) {
                const int ko = M.rk_perm[e][k];
                const float v = h ? g1 : g0;
                Grep[(size_t)i * NN_K1 + ko] = v;
                if (v < M.glo[e][ko] || v > M.ghi[e][ko]) extrap = 1;
            }
        }
    }
    if (extrap) *extrapp = 1;
}

// Functions lo .. hi-1 of element e on the lane's first NQ entries:
// sq[q] += c_k dG_k/dr at entry q, dG_k/dr = exp(-eta (r - rs)^2) (dfc - 2
// eta (r - rs) fc).  Next function's parameters prefetched.
template <int NQ>
__device__ __forceinline__ void rf_class(const NnConst &M, int e, int lo, int hi, const float *crow,
                                         const float rr[NN_RQ], const float fx[NN_RQ], const float dx[NN_RQ],
                                         float sq[NN_RQ]) {
    float rs = M.rk_rs[e][lo], ce = M.rk_ce[e][lo], bk = M.rk_b[e][lo];
    float ck = crow[M.rk_perm[e][lo]];
    for (int k = lo; k < hi; ++k) {
        const int kn = k + 1 < hi ? k + 1 : k;
        const float rsn = M.rk_rs[e][kn], cen = M.rk_ce[e][kn], bkn = M.rk_b[e][kn];
        const float ckn = crow[M.rk_perm[e][kn]];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const float dr = rr[q] - rs;
            const float t = __builtin_amdgcn_exp2f(ce * dr * dr) * ck;
            sq[q] = fmaf(t, fmaf(bk * dr, fx[q], dx[q]), sq[q]);
        }
        rs = rsn;
        ce = cen;
        bk = bkn;
        ck = ckn;
    }
}

// dE/dr of the chunk's centres' radial functions with coefficients crow =
// dE/dG of the lane's centre (input order): +F on the centre (row sum), -F
// on each neighbour into the fixed-point accumulators (Hartree/Bohr).
__device__ void rf_chunk(const NnConst &M, const RadNbrF &L, const float *crow, int e, int i, int lane,
                         unsigned long long *ax, unsigned long long *ay, unsigned long long *az, int *badp) {
    const int c = lane >> 4, sub = lane & 15;
    constexpr int RQ = RadNbrF::W / NN_TPA;  // entries per lane of a window (2)
    const int base = c * RadNbrF::W;
    const int nel = M.nelem;
    const float pi_rc = M.pi_rc;
    float rr[RQ], fcq[RQ], dfcq[RQ], sq[RQ];
    int eq[RQ], nq;
    unsigned int present;
    rad_entries(L, lane, M.rc, nel, rr, eq, &nq, &present);
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        float fc, dfc;
        cutoff(pi_rc, rr[q], &fc, &dfc);
        fcq[q] = eq[q] >= 0 ? fc : 0.f;
        dfcq[q] = eq[q] >= 0 ? dfc : 0.f;
        sq[q] = 0.f;
    }
    for (int x = 0; x < nel; ++x) {
        const int lo = M.rk_lo[e][x], hi = M.rk_hi[e][x];
        if (lo == hi || !((present >> x) & 1u)) continue;
        float fx[RQ], dx[RQ];
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
            fx[q] = eq[q] == x ? fcq[q] : 0.f;
            dx[q] = eq[q] == x ? dfcq[q] : 0.f;
        }
        switch (nq) {
            case 1: rf_class<1>(M, e, lo, hi, crow, rr, fx, dx, sq); break;
            case 2: rf_class<2>(M, e, lo, hi, crow, rr, fx, dx, sq); break;
            default: break;
        }
    }
    float F0 = 0.f, F1 = 0.f, F2 = 0.f;
    int bad = 0;
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        if (q < nq && eq[q] >= 0) {
            const int e2 = base + sub + NN_TPA * q;
            const float v0 = sq[q] * L.ux[e2], v1 = sq[q] * L.uy[e2], v2 = sq[q] * L.uz[e2];
            F0 += v0;
            F1 += v1;
            F2 += v2;
            const int j = L.j[e2];
            fxadd32(&ax[j], -v0, &bad);
            fxadd32(&ay[j], -v1, &bad);
            fxadd32(&az[j], -v2, &bad);
        }
    }
    const float S0 = row_sum16(F0), S1 = row_sum16(F1), S2 = row_sum16(F2);
    if (i >= 0 && sub == 0) {
        fxadd32(&ax[i], S0, &bad);
        fxadd32(&ay[i], S1, &bad);
        fxadd32(&az[i], S2, &bad);
    }
    if (bad) *badp |= bad;
}

// Activation and its derivative with one exponential and one reciprocal:
// tanh z = 1 - 2 / (1 + e^{2z}) (absolute error ~1e-7 near 0, exact
// saturation), sigmoid z = 1 / (1 + e^{-z}).
__device__ __forceinline__ float act_fast(int code, float z, float *d1) {
    if (code == 1) {
        const float h = 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(2.0f * NN_L2E * z));
        *d1 = 1.0f - h * h;
        return h;
    }
    if (code == 2) {
        const float h = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-NN_L2E * z));
        *d1 = h * (1.0f - h);
        return h;
    }
    *d1 = 1.0f;
    return z;
}

// orders a wavefront's LDS writes before its later LDS reads (the transpose
// buffers of mlp_tile_w are private to one wavefront)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// mlp_tile for one wavefront on its own (no block barriers): the same
// products on v_mfma_f32_16x16x32_bf16 with hi/lo split operands; the
// second 32-input block of layer 1 and the dE/dx tiles past the element's
// inputs are skipped; act_fast for the activations.
__device__ void mlp_tile_w(const NnConst &M, int e, const int *rows, const float *Grep, float *Crep, double *Ea,
                           float *T, int lane) {
    const int l = lane & 63;
    const int rA = l & 15, kg = (l >> 4) * 8;
    const int nB = l & 15;
    const int rC = (l >> 4) * 4;
    const int nin = M.nr[e] + M.na[e];
    const int nks = nin > 32 ? 2 : 1;
    const int atomA = rows[rA];

    f32x4 z1[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    for (int ks = 0; ks < nks; ++ks) {
        float av[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = ks * 32 + kg + j;
            av[j] = (atomA >= 0 && k < nin) ? Grep[(size_t)atomA * NN_K1 + k] * M.smul[e][k] + M.sadd[e][k] : 0.f;
        }
        bf16x8 ah, al;
        split8(av, ah, al);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            float bv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) bv[j] = M.W1[e][ks * 32 + kg + j][nt * 16 + nB];
            bf16x8 bh, bl;
            split8(bv, bh, bl);
            z1[nt] = mfma3(ah, al, bh, bl, z1[nt]);
        }
    }
    float h1[2][4], d1[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) h1[nt][i] = act_fast(M.act1, z1[nt][i] + M.b1[e][nt * 16 + nB], &d1[nt][i]);
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(rC + i) * 33 + nt * 16 + nB] = h1[nt][i];
    wave_lds_sync();
    float av2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) av2[j] = T[rA * 33 + kg + j];
    wave_lds_sync();
    bf16x8 ah, al;
    split8(av2, ah, al);
    f32x4 z2[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W2[e][kg + j][nt * 16 + nB];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        z2[nt] = mfma3(ah, al, bh, bl, z2[nt]);
    }
    float h2[2][4], d2[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) h2[nt][i] = act_fast(M.act2, z2[nt][i] + M.b2[e][nt * 16 + nB], &d2[nt][i]);
    float dz3[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float p = h2[0][i] * M.w3[e][nB] + h2[1][i] * M.w3[e][16 + nB];
        p = row_sum16(p);
        float d3;
        const float E = act_fast(M.act3, p + M.b3[e], &d3);
        dz3[i] = d3;
        const int row = rC + i;
        if (nB == 0 && rows[row] >= 0) Ea[rows[row]] = (double)E;
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(rC + i) * 33 + nt * 16 + nB] = dz3[i] * M.w3[e][nt * 16 + nB] * d2[nt][i];
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < 8; ++j) av2[j] = T[rA * 33 + kg + j];
    wave_lds_sync();
    split8(av2, ah, al);
    f32x4 g1[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W2[e][nt * 16 + nB][kg + j];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        g1[nt] = mfma3(ah, al, bh, bl, g1[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(rC + i) * 33 + nt * 16 + nB] = g1[nt][i] * d1[nt][i];
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < 8; ++j) av2[j] = T[rA * 33 + kg + j];
    wave_lds_sync();
    split8(av2, ah, al);
    for (int nt = 0; nt < 4; ++nt) {
        if (nt * 16 >= nin) break;
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W1[e][nt * 16 + nB][kg + j];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        f32x4 dx = {0.f, 0.f, 0.f, 0.f};
        dx = mfma3(ah, al, bh, bl, dx);
        const int k = nt * 16 + nB;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = rC + i;
            if (rows[row] >= 0 && k < nin) Crep[(size_t)rows[row] * NN_K1 + k] = dx[i] * M.dfac[e][k];
        }
    }
}

// network_energy with wavefronts working through the tiles independently
// (tile = wave, wave + 4, ...), one block barrier before the energy sum.
__device__ void network_energy_w(const NnConst &M, const int *s_ne, const uint16_t *elist, int (*rowsb)[16],
                                 float (*T)[16 * 33], const float *Grep, float *Crep, double *Ea, int N, int lane,
                                 double *epot) {
    const int wave = lane >> 6;
    const int ne_model = M.nelem;
    int ntiles = 0;
    for (int e = 0; e < ne_model; ++e) ntiles += (s_ne[e] + 15) / 16;
    for (int tile = wave; tile < ntiles; tile += NN_BLOCK / 64) {
        int e = -1, first = 0, cntE = 0, tl = tile, start = 0;
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
        e = __builtin_amdgcn_readfirstlane(e);
        if ((lane & 63) < 16) {
            const int rr = lane & 15;
            rowsb[wave][rr] = rr < cntE ? (int)elist[first + rr] : -1;
        }
        wave_lds_sync();
        mlp_tile_w(M, e, rowsb[wave], Grep, Crep, Ea, T[wave], lane);
        wave_lds_sync();
    }
    __syncthreads();
    if (lane < 64) {
        double e = 0.0;
        for (int i = lane; i < N; i += 64) e += Ea[i];
        for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
        if (lane == 0) *epot = e * NN_HA2EV;
    }
}

// Pass G and the network of a radial-only model (k_nene_g's contract).
template <bool IMG>
__global__ void __launch_bounds__(NN_BLOCK, 5) k_nene_rg(const NnConst *__restrict__ Mg, BatchView bv, NnView nv,
                                                      int skip_extrap) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ union {
        RadNbrG L;
        float T[NN_BLOCK / 64][16 * 33];
    } U;
    RadNbrG &L = U.L;
    float(*T)[16 * 33] = U.T;
    __shared__ int rowsb[NN_BLOCK / 64][16];
    __shared__ int s_ne[NN_MAXE];
    __shared__ RadCell Cl;
    const NnConst &M = *Mg;
    const int rep = bv.rep_list ? bv.rep_list[blockIdx.x] : (int)blockIdx.x;  // size-class launches
    const int t = bv.rep_traj[rep];
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int b = bv.rep_bead[rep];
    const int N = bv.traj_N[t];
    const long long soff = bv.traj_soff[t];
    const long long roff = soff + (long long)b * 3 * N;
    const long long aoff = bv.traj_aoff[t];
    const int lane = threadIdx.x;
    float *Grep = nv.G + (size_t)rep * nv.maxN * NN_K1;
    float *Crep = nv.C + (size_t)rep * nv.maxN * NN_K1;
    int bad = 0, extrap = 0;

    // dynamic LDS: Ea (double N) | ux uy uz (u32 N each) | slot (u32 N + 4 NN_MAXE) | elist (u16 N) | elem (u8 N)
    double *Ea = (double *)smem_raw;
    uint32_t *ux = (uint32_t *)(Ea + N), *uy = ux + N, *uz = uy + N;
    uint32_t *slot = uz + N;
    uint16_t *elist = (uint16_t *)(slot + N + 4 * NN_MAXE);
    uint8_t *elem = (uint8_t *)(elist + N);

    const int ne_model = M.nelem;
    double ih[9];
    rad_cell(Cl, M, bv.cell + 9 * t, bv.icell + 9 * t, ih, nv.nbr != nullptr, nv.skin, lane);

    for (int i = lane; i < N; i += NN_BLOCK) {
        const int ty = bv.type[aoff + i];
        const int e = M.elem_of_type[ty];
        if (e < 0 || e >= ne_model) bad |= MDT_ST_BAD_PROP;
        elem[i] = (uint8_t)(e < 0 ? 0 : e);
        const double r0 = bv.x[roff + i], r1 = bv.x[roff + N + i], r2 = bv.x[roff + 2 * N + i];
        double f0 = ih[0] * r0 + ih[3] * r1 + ih[6] * r2;
        double f1 = ih[1] * r0 + ih[4] * r1 + ih[7] * r2;
        double f2 = ih[2] * r0 + ih[5] * r1 + ih[8] * r2;
        ux[i] = frac_fix(f0 - floor(f0));
        uy[i] = frac_fix(f1 - floor(f1));
        uz[i] = frac_fix(f2 - floor(f2));
        Ea[i] = 0.0;
    }
    __syncthreads();
    group_elements(elem, N, ne_model, elist, s_ne, lane);
    __syncthreads();
    const int nslot = make_slots(elist, s_ne, ne_model, slot, lane);
    __syncthreads();

    // stored lists are kept while every atom stays within skin/2 of its
    // position at the build (same cut-off and skin: nbr_key)
    const bool store = __builtin_amdgcn_readfirstlane(Cl.store);
    const float rb = M.rc + nv.skin;
    const int key = __float_as_int(rb);
    int mode = 0;
    if (store && nv.skin > 0.f && nv.nbr_key[rep] == key) {
        const uint32_t *ref = nv.nbr_ref + (size_t)rep * 3 * nv.maxN;
        const float h2 = 0.25f * nv.skin * nv.skin;
        int moved = 0;
        for (int i = lane; i < N; i += NN_BLOCK) {
            const float f0 = (float)(int)(ux[i] - ref[i]);
            const float f1 = (float)(int)(uy[i] - ref[nv.maxN + i]);
            const float f2 = (float)(int)(uz[i] - ref[2 * nv.maxN + i]);
            const float dx = Cl.hs[0] * f0 + Cl.hs[3] * f1 + Cl.hs[6] * f2;
            const float dy = Cl.hs[1] * f0 + Cl.hs[4] * f1 + Cl.hs[7] * f2;
            const float dz = Cl.hs[2] * f0 + Cl.hs[5] * f1 + Cl.hs[8] * f2;
            if (dx * dx + dy * dy + dz * dz >= h2) moved = 1;
        }
        mode = __syncthreads_or(moved) ? 0 : 1;
    }

    // ---- pass G: symmetry functions ----
    for (int k0 = 0; k0 < nslot; k0 += NN_CA) {
        const int sl = k0 + (lane >> 4);
        const uint32_t sv = sl < nslot ? slot[sl] : 0xffffu;
        const int i = (sv & 0xffffu) == 0xffffu ? -1 : (int)(sv & 0xffffu);
        const int e = __builtin_amdgcn_readfirstlane((int)(sv >> 16));
        rad_build<RadNbrG, IMG>(L, Cl, nv, rep, mode, ux, uy, uz, elem, M.rc2, rb * rb, N, i, lane, &bad);
        __syncthreads();
        rg_chunk(M, L, e, i, lane, Grep, &extrap);
        __syncthreads();
    }
    if (store && mode == 0) {  // positions of this build
        uint32_t *ref = nv.nbr_ref + (size_t)rep * 3 * nv.maxN;
        for (int i = lane; i < N; i += NN_BLOCK) {
            ref[i] = ux[i];
            ref[nv.maxN + i] = uy[i];
            ref[2 * nv.maxN + i] = uz[i];
        }
        if (lane == 0) nv.nbr_key[rep] = nv.skin > 0.f ? key : 0;
    }

    // replicas outside the training range go to the EMT fallback untouched
    extrap = __syncthreads_or(extrap);
    if (lane == 0) nv.rep_flag[rep] = extrap;
    if (extrap && skip_extrap) {
        if (lane == 0)
            atomicOr(&bv.status[t], bad | MDT_ST_NN_EXTRAPOLATION);
        else if (bad)
            atomicOr(&bv.status[t], bad);
        return;
    }
    network_energy_w(M, s_ne, elist, rowsb, T, Grep, Crep, Ea, N, lane, &bv.epot[bv.traj_eoff[t] + b]);
    if (extrap) bad |= MDT_ST_NN_EXTRAPOLATION;
    if (bad) atomicOr(&bv.status[t], bad);
}

// Pass F and the outputs of a radial-only model (k_nene_f's contract).
template <bool IMG>
__global__ void __launch_bounds__(NN_BLOCK, 4) k_nene_rf(const NnConst *__restrict__ Mg, BatchView bv, NnView nv,
                                                      int add_fin, int skip_extrap) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ RadNbrF L;
    __shared__ float cb[NN_CA][NN_K1];
    __shared__ int s_ne[NN_MAXE];
    __shared__ RadCell Cl;
    const NnConst &M = *Mg;
    const int rep = bv.rep_list ? bv.rep_list[blockIdx.x] : (int)blockIdx.x;  // size-class launches
    if (skip_extrap && nv.rep_flag[rep]) return;  // whole block: uniform
    const int t = bv.rep_traj[rep];
    if (bv.done && bv.done[t]) return;  // frozen trajectory (whole block: uniform)
    const int b = bv.rep_bead[rep];
    const int N = bv.traj_N[t];
    const long long soff = bv.traj_soff[t];
    const long long roff = soff + (long long)b * 3 * N;
    const long long doff = soff / 3 + (long long)b * N;
    const long long aoff = bv.traj_aoff[t];
    const int lane = threadIdx.x;
    const float *Crep = nv.C + (size_t)rep * nv.maxN * NN_K1;
    int bad = 0;

    // dynamic LDS: ax ay az (u64 N each) | ux uy uz (u32 N) | slot (u32 N + 4 NN_MAXE) | elist (u16 N) | elem (u8 N)
    unsigned long long *ax = (unsigned long long *)smem_raw, *ay = ax + N, *az = ay + N;
    uint32_t *ux = (uint32_t *)(az + N), *uy = ux + N, *uz = uy + N;
    uint32_t *slot = uz + N;
    uint16_t *elist = (uint16_t *)(slot + N + 4 * NN_MAXE);
    uint8_t *elem = (uint8_t *)(elist + N);

    const int ne_model = M.nelem;
    double ih[9];
    rad_cell(Cl, M, bv.cell + 9 * t, bv.icell + 9 * t, ih, nv.nbr != nullptr, nv.skin, lane);
    for (int i = lane; i < N; i += NN_BLOCK) {
        const int e = M.elem_of_type[bv.type[aoff + i]];
        elem[i] = (uint8_t)(e < 0 || e >= ne_model ? 0 : e);
        const double r0 = bv.x[roff + i], r1 = bv.x[roff + N + i], r2 = bv.x[roff + 2 * N + i];
        double f0 = ih[0] * r0 + ih[3] * r1 + ih[6] * r2;
        double f1 = ih[1] * r0 + ih[4] * r1 + ih[7] * r2;
        double f2 = ih[2] * r0 + ih[5] * r1 + ih[8] * r2;
        ux[i] = frac_fix(f0 - floor(f0));
        uy[i] = frac_fix(f1 - floor(f1));
        uz[i] = frac_fix(f2 - floor(f2));
        ax[i] = ay[i] = az[i] = 0ull;
    }
    __syncthreads();
    group_elements(elem, N, ne_model, elist, s_ne, lane);
    __syncthreads();
    const int nslot = make_slots(elist, s_ne, ne_model, slot, lane);
    __syncthreads();

    // ---- pass F: forces, the chunk's lists in windows of NN_WF entries per
    // centre (one window unless some centre of the chunk has more) ----
    for (int k0 = 0; k0 < nslot; k0 += NN_CA) {
        const int sl = k0 + (lane >> 4);
        const uint32_t sv = sl < nslot ? slot[sl] : 0xffffu;
        const int i = (sv & 0xffffu) == 0xffffu ? -1 : (int)(sv & 0xffffu);
        const int e = __builtin_amdgcn_readfirstlane((int)(sv >> 16));
        // the lists k_nene_rg stored or kept for these positions
        const float rb = M.rc + nv.skin;
        int nwin = 1;  // set after the first window's barrier, the same in every lane
        for (int w = 0; w < nwin; ++w) {
            rad_build<RadNbrF, IMG>(L, Cl, nv, rep, 1, ux, uy, uz, elem, M.rc2, rb * rb, N, i, lane, &bad, w * NN_WF);
            if (w == 0)
                for (int q = lane; q < NN_CA * NN_K1; q += NN_BLOCK) {
                    const int ci = q / NN_K1, k = q % NN_K1;
                    const int s2 = k0 + ci;
                    const uint32_t v2 = s2 < nslot ? slot[s2] : 0xffffu;
                    const int ii = (int)(v2 & 0xffffu);
                    cb[ci][k] = (ii != 0xffff && k < M.nr[v2 >> 16]) ? Crep[(size_t)ii * NN_K1 + k] : 0.f;
                }
            __syncthreads();
            if (w == 0)
                for (int c2 = 0; c2 < NN_CA; ++c2) {
                    const int nw = (L.tot[c2] + NN_WF - 1) / NN_WF;
                    nwin = nw > nwin ? nw : nwin;
                }
            rf_chunk(M, L, cb[lane >> 4], e, i, lane, ax, ay, az, &bad);
            __syncthreads();
        }
    }
    write_forces(bv, nv, ax, ay, az, N, roff, doff, aoff, add_fin, lane);
    if (bad) atomicOr(&bv.status[t], bad);
}


// ===========================================================================
// Training (no reference counterpart: md_tian2 fits with RuNNer's Kalman
// filter, fit.f90).  Loss over a minibatch of Nf frames with Ntot atoms:
//   L = wE/Nf sum_f ((E_f - Eref_f)/N_f)^2 + wF/(3 Ntot) sum |F - Fref|^2
// in eV and eV/A.  One iteration: k_nn_mlp (network forward/backward on
// 16-atom tiles), k_nn_frame mode 1 (forces, residuals, the Jacobian-vector
// product u = dL/d(dE/dx)), k_nn_grad (tangent and second-order passes,
// weight gradients on MFMA, per-wave partials), k_nn_adam (fixed-order
// reduction of the partials, Adam on the FP32 master weights in NnConst).
// ===========================================================================

// k_nn_frame: mode 0 writes the raw symmetry functions of every atom of the
// frame; mode 1 computes the frame's forces from C, its energy from Eatom,
// the loss parts, cE = dL/dE and U rows.
__global__ void __launch_bounds__(NN_BLOCK) k_nn_frame(const NnConst *__restrict__ Mg, NnTrainView tv,
                                                       const int *frames, int mode, double cEw, double cFw) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ NnSfTab S;
    __shared__ ChunkNbr L;
    __shared__ float cb[NN_CA][NN_K1];
    __shared__ float Wsh[NN_CA][NN_FB * NN_FB];
    __shared__ float sdfac[NN_MAXE][NN_K1];
    __shared__ double red[NN_BLOCK / 64];
    const NnConst &M = *Mg;
    const int f = frames[blockIdx.x];
    const int N = tv.fr_natoms[f];
    const long long ao = tv.fr_aoff[f];
    const int lane = threadIdx.x;
    int bad = 0, extrap = 0;
    // dynamic LDS: ax ay az (u64 N) | ux uy uz (u32 N) | gx gy gz (float N) | elem (u8 N)
    unsigned long long *ax = (unsigned long long *)smem_raw, *ay = ax + N, *az = ay + N;
    uint32_t *ux = (uint32_t *)(az + N), *uy = ux + N, *uz = uy + N;
    float *gx = (float *)(uz + N), *gy = gx + N, *gz = gy + N;
    uint8_t *elem = (uint8_t *)(gz + N);
    load_sftab(S, M, lane);
    for (int k = lane; k < NN_MAXE * NN_K1; k += NN_BLOCK) sdfac[k / NN_K1][k % NN_K1] = M.dfac[k / NN_K1][k % NN_K1];
    double h[9], ih[9];
    float w[3];
    int imax[3];
    cell_setup(M, tv.fr_cell + 9 * f, tv.fr_icell + 9 * f, h, ih, w, imax);
    for (int i = lane; i < N; i += NN_BLOCK) {
        const int e = M.elem_of_type[tv.type[ao + i]];
        elem[i] = (uint8_t)(e < 0 ? 0 : e);
        const double *r = tv.pos + 3 * (ao + i);
        const double f0 = ih[0] * r[0] + ih[3] * r[1] + ih[6] * r[2];
        const double f1 = ih[1] * r[0] + ih[4] * r[1] + ih[7] * r[2];
        const double f2 = ih[2] * r[0] + ih[5] * r[1] + ih[8] * r[2];
        ux[i] = frac_fix(f0 - floor(f0));
        uy[i] = frac_fix(f1 - floor(f1));
        uz[i] = frac_fix(f2 - floor(f2));
        ax[i] = ay[i] = az[i] = 0ull;
    }
    __syncthreads();
    float *Grow = tv.G + (size_t)ao * NN_K1;
    if (mode == 0) {
        for (int k0 = 0; k0 < N; k0 += NN_CA) {
            build_chunk(M.rc2, M.pi_rc, L, ux, uy, uz, elem, h, imax, w[0], w[1], w[2], N, k0, lane, &bad);
            __syncthreads();
            g_chunk(M, S, L, elem, N, k0, lane, M.rc2, M.pi_rc, Grow, &extrap);
            __syncthreads();
        }
        return;
    }
    // forces with the current network (C rows from k_nn_mlp)
    const float *Crow = tv.C + (size_t)ao * NN_K1;
    for (int k0 = 0; k0 < N; k0 += NN_CA) {
        build_chunk(M.rc2, M.pi_rc, L, ux, uy, uz, elem, h, imax, w[0], w[1], w[2], N, k0, lane, &bad);
        for (int q = lane; q < NN_CA * NN_K1; q += NN_BLOCK) {
            const int ci = q / NN_K1, k = q % NN_K1;
            const int i = k0 + ci;
            cb[ci][k] = (i < N && k < S.nr[elem[i]] + S.na[elem[i]]) ? Crow[(size_t)i * NN_K1 + k] : 0.f;
        }
        __syncthreads();
        force_chunk(S, L, cb, Wsh, elem, N, k0, lane, M.rc2, M.pi_rc, ax, ay, az, &bad);
        __syncthreads();
    }
    // residuals: g = dL/dF (per Hartree/Bohr), force loss part
    const double fconv = NN_HA2EV / NN_BOHR2ANG;
    double lf = 0.0;
    for (int i = lane; i < N; i += NN_BLOCK) {
        const double F[3] = {(double)(long long)ax[i] * NN_FISCALE * fconv,
                             (double)(long long)ay[i] * NN_FISCALE * fconv,
                             (double)(long long)az[i] * NN_FISCALE * fconv};
        const double *Fr = tv.Fref + 3 * (ao + i);
        const double d0 = F[0] - Fr[0], d1 = F[1] - Fr[1], d2 = F[2] - Fr[2];
        lf += cFw * (d0 * d0 + d1 * d1 + d2 * d2);
        gx[i] = (float)(2.0 * cFw * d0 * fconv);
        gy[i] = (float)(2.0 * cFw * d1 * fconv);
        gz[i] = (float)(2.0 * cFw * d2 * fconv);
    }
    for (int o = 32; o > 0; o >>= 1) lf += __shfl_xor(lf, o, 64);
    if ((lane & 63) == 0) red[lane >> 6] = lf;
    __syncthreads();
    if (lane == 0) {
        double sF = 0.0, E = 0.0;
        for (int q = 0; q < NN_BLOCK / 64; ++q) sF += red[q];
        for (int i = 0; i < N; ++i) E += tv.Eatom[ao + i];
        const double dE = (NN_HA2EV * E - tv.Eref[f]) / N;
        tv.loss[2 * f] = cEw * dE * dE;
        tv.loss[2 * f + 1] = sF;
        tv.cE[f] = 2.0 * cEw * dE / N * NN_HA2EV;
    }
    // Jacobian-vector product
    float *Urow = tv.U + (size_t)ao * NN_K1;
    for (int k0 = 0; k0 < N; k0 += NN_CA) {
        build_chunk(M.rc2, M.pi_rc, L, ux, uy, uz, elem, h, imax, w[0], w[1], w[2], N, k0, lane, &bad);
        __syncthreads();
        jvp_chunk(S, L, elem, N, k0, lane, M.rc2, M.pi_rc, gx, gy, gz, sdfac, Urow);
        __syncthreads();
    }
}

// network forward/backward on the minibatch's tiles (rows: global atom
// indices; tiles: element, first row, row count)
__global__ void __launch_bounds__(NN_BLOCK) k_nn_mlp(const NnConst *__restrict__ Mg, NnTrainView tv,
                                                     const int4 *tiles, int ntiles, const int *rows) {
    __shared__ float T[NN_BLOCK / 64][16 * 33];
    __shared__ int rowsb[NN_BLOCK / 64][16];
    const int lane = threadIdx.x, wave = lane >> 6;
    const int tile = blockIdx.x * (NN_BLOCK / 64) + wave;
    int e = -1;
    if (tile < ntiles) e = tiles[tile].x;
    if ((lane & 63) < 16) {
        const int rr = lane & 15;
        rowsb[wave][rr] = (e >= 0 && rr < tiles[tile].z) ? rows[tiles[tile].y + rr] : -1;
    }
    __syncthreads();
    mlp_tile(*Mg, e, rowsb[wave], tv.G, tv.C, tv.Eatom, T[wave], lane);
}

namespace {
// acc += P^T Q over 16 rows: P[16][ldp] columns p0..p0+15, Q[16][ldq]
// columns q0..q0+15 (K = 32 with rows 16..31 zero)
__device__ __forceinline__ f32x4 outer16(const float *P, int ldp, int p0, const float *Q, int ldq, int q0, f32x4 acc,
                                         int l) {
    float av[8], bv[8];
    const int m = l & 15, kb = (l >> 4) * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int r = kb + j;
        av[j] = r < 16 ? P[r * ldp + p0 + m] : 0.f;
        bv[j] = r < 16 ? Q[r * ldq + q0 + m] : 0.f;
    }
    bf16x8 ah, al, bh, bl;
    split8(av, ah, al);
    split8(bv, bh, bl);
    return mfma3(ah, al, bh, bl, acc);
}

__device__ __forceinline__ float act2nd(int code, float h, float d1) {
    if (code == 1) return -2.f * h * d1;
    if (code == 2) return d1 * (1.f - 2.f * h);
    return 0.f;
}

struct GradAcc {
    f32x4 w1[4][2], w2[2][2];
    float b1[2], b2[2], w3[2], b3;
};

__device__ void acc_zero(GradAcc &a) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n) a.w1[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n) a.w2[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    a.b1[0] = a.b1[1] = a.b2[0] = a.b2[1] = a.w3[0] = a.w3[1] = a.b3 = 0.f;
}

// per-wave partial gradient of element e (padded layout, NN_NPAR floats)
__device__ void acc_flush(const GradAcc &a, float *out, int l) {
    const int nB = l & 15, rC = (l >> 4) * 4;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[NN_PW1 + (m * 16 + rC + i) * NN_H + n * 16 + nB] = a.w1[m][n][i];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[NN_PW2 + (m * 16 + rC + i) * NN_H + n * 16 + nB] = a.w2[m][n][i];
    // bias / output sums: reduce the lane partials over the 4 row groups
    float v[5] = {a.b1[0], a.b1[1], a.b2[0], a.b2[1], 0.f};
    float w3v[2] = {a.w3[0], a.w3[1]};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] += __shfl_xor(v[q], 16, 64);
        v[q] += __shfl_xor(v[q], 32, 64);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        w3v[q] += __shfl_xor(w3v[q], 16, 64);
        w3v[q] += __shfl_xor(w3v[q], 32, 64);
    }
    float b3 = a.b3;
    for (int o = 1; o < 64; o <<= 1) b3 += __shfl_xor(b3, o, 64);
    if (l < 16) {
        out[NN_PB1 + nB] = v[0];
        out[NN_PB1 + 16 + nB] = v[1];
        out[NN_PB2 + nB] = v[2];
        out[NN_PB2 + 16 + nB] = v[3];
        out[NN_PW3 + nB] = w3v[0];
        out[NN_PW3 + 16 + nB] = w3v[1];
    }
    if (l == 0) out[NN_PB3] = b3;
}
}  // namespace

// Weight gradient of one 16-row tile of element e (e < 0: padding tile; every
// lane of the block calls this the same number of times).  Stage: wave's
// LDS area of 2 x 16 x 65 + 4 x 16 x 33 floats.  Linear output layer.
__device__ void grad_tile(const NnConst &M, int e, const int *rows, const NnTrainView &tv, float *stage, GradAcc &acc,
                          int lane) {
    const int l = lane & 63;
    const int rA = l & 15, kg = (l >> 4) * 8, nB = l & 15, rC = (l >> 4) * 4;
    const bool act = e >= 0;
    const int ee = act ? e : 0;
    const int nin = act ? M.nr[ee] + M.na[ee] : 0;
    float *Xs = stage, *Us = Xs + 16 * 65, *A1 = Us + 16 * 65, *A2 = A1 + 16 * 33, *A3 = A2 + 16 * 33,
          *A4 = A3 + 16 * 33;
    const int atomA = act ? rows[rA] : -1;
    // cE of this lane's C rows
    float cEr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int a = act ? rows[rC + i] : -1;
        cEr[i] = a >= 0 ? (float)tv.cE[tv.atom_frame[a]] : 0.f;
    }
    // ---- layer 1 primal and tangent: z1 = x W1 + b1, zd1 = u W1 ----
    f32x4 z1[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}, zd1[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        float xv[8], uv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = ks * 32 + kg + j;
            const bool ok = atomA >= 0 && k < nin;
            xv[j] = ok ? tv.G[(size_t)atomA * NN_K1 + k] * M.smul[ee][k] + M.sadd[ee][k] : 0.f;
            uv[j] = ok ? tv.U[(size_t)atomA * NN_K1 + k] : 0.f;
            Xs[rA * 65 + k] = xv[j];
            Us[rA * 65 + k] = uv[j];
        }
        bf16x8 xh, xl, uh, ul;
        split8(xv, xh, xl);
        split8(uv, uh, ul);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            float bv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) bv[j] = M.W1[ee][ks * 32 + kg + j][nt * 16 + nB];
            bf16x8 bh, bl;
            split8(bv, bh, bl);
            z1[nt] = mfma3(xh, xl, bh, bl, z1[nt]);
            zd1[nt] = mfma3(uh, ul, bh, bl, zd1[nt]);
        }
    }
    float h1[2][4], d1[2][4], s1[2][4], hd1[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            h1[nt][i] = act_f(M.act1, z1[nt][i] + M.b1[ee][nt * 16 + nB], &d1[nt][i]);
            s1[nt][i] = act2nd(M.act1, h1[nt][i], d1[nt][i]);
            hd1[nt][i] = d1[nt][i] * zd1[nt][i];
            A1[(rC + i) * 33 + nt * 16 + nB] = h1[nt][i];
            A2[(rC + i) * 33 + nt * 16 + nB] = hd1[nt][i];
        }
    __syncthreads();
    // ---- layer 2 primal and tangent ----
    float hv[8], hdv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        hv[j] = A1[rA * 33 + kg + j];
        hdv[j] = A2[rA * 33 + kg + j];
    }
    bf16x8 hh, hl, dh, dl;
    split8(hv, hh, hl);
    split8(hdv, dh, dl);
    f32x4 z2[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}, zd2[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W2[ee][kg + j][nt * 16 + nB];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        z2[nt] = mfma3(hh, hl, bh, bl, z2[nt]);
        zd2[nt] = mfma3(dh, dl, bh, bl, zd2[nt]);
    }
    // rho2 = d2 w3, pi2 = a2'' zd2 w3, Y2 = cE rho2 + pi2
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float d2;
            const float h2 = act_f(M.act2, z2[nt][i] + M.b2[ee][nt * 16 + nB], &d2);
            const float s2 = act2nd(M.act2, h2, d2);
            const float w3 = M.w3[ee][nt * 16 + nB];
            const float rho2 = d2 * w3, pi2 = s2 * zd2[nt][i] * w3;
            const float y2 = cEr[i] * rho2 + pi2;
            A3[(rC + i) * 33 + nt * 16 + nB] = y2;
            A4[(rC + i) * 33 + nt * 16 + nB] = rho2;
            acc.b2[nt] += y2;
            acc.w3[nt] += cEr[i] * h2 + d2 * zd2[nt][i];
        }
    if (nB == 0)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc.b3 += cEr[i];
    __syncthreads();
    // gW2 += H1^T Y2 + Hd1^T rho2
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            acc.w2[m][n] = outer16(A1, 33, m * 16, A3, 33, n * 16, acc.w2[m][n], l);
            acc.w2[m][n] = outer16(A2, 33, m * 16, A4, 33, n * 16, acc.w2[m][n], l);
        }
    // q = rho2 W2^T, p = pi2 W2^T  (pi2 = Y2 - cE rho2 in A layout)
    float rv[8], pv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int a = act ? rows[rA] : -1;
        const float cE = a >= 0 ? (float)tv.cE[tv.atom_frame[a]] : 0.f;
        rv[j] = A4[rA * 33 + kg + j];
        pv[j] = A3[rA * 33 + kg + j] - cE * rv[j];
    }
    bf16x8 rh, rl, ph, pl;
    split8(rv, rh, rl);
    split8(pv, ph, pl);
    f32x4 q[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}, pq[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = M.W2[ee][nt * 16 + nB][kg + j];
        bf16x8 bh, bl;
        split8(bv, bh, bl);
        q[nt] = mfma3(rh, rl, bh, bl, q[nt]);
        pq[nt] = mfma3(ph, pl, bh, bl, pq[nt]);
    }
    __syncthreads();
    // rho1 = d1 q, pi1 = d1 p + a1'' zd1 q, Y1 = cE rho1 + pi1
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float rho1 = d1[nt][i] * q[nt][i];
            const float pi1 = d1[nt][i] * pq[nt][i] + s1[nt][i] * zd1[nt][i] * q[nt][i];
            const float y1 = cEr[i] * rho1 + pi1;
            A1[(rC + i) * 33 + nt * 16 + nB] = y1;
            A2[(rC + i) * 33 + nt * 16 + nB] = rho1;
            acc.b1[nt] += y1;
        }
    __syncthreads();
    // gW1 += X^T Y1 + U^T rho1
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            acc.w1[m][n] = outer16(Xs, 65, m * 16, A1, 33, n * 16, acc.w1[m][n], l);
            acc.w1[m][n] = outer16(Us, 65, m * 16, A2, 33, n * 16, acc.w1[m][n], l);
        }
    __syncthreads();
}

// per-wave partial gradients over the minibatch's tiles: wave gw takes tiles
// gw, gw + W, ... (element-grouped, so a wave meets each element in one run);
// partials[gw][e] (NN_NPAR floats each, zeroed by the host)
__global__ void __launch_bounds__(NN_BLOCK) k_nn_grad(const NnConst *__restrict__ Mg, NnTrainView tv,
                                                      const int4 *tiles, int ntiles, const int *rows,
                                                      float *partials) {
    __shared__ float stage[NN_BLOCK / 64][2 * 16 * 65 + 4 * 16 * 33];
    __shared__ int rowsb[NN_BLOCK / 64][16];
    const NnConst &M = *Mg;
    const int lane = threadIdx.x, wave = lane >> 6, l = lane & 63;
    const int W = gridDim.x * (NN_BLOCK / 64);
    const int gw = blockIdx.x * (NN_BLOCK / 64) + wave;
    const int rounds = (ntiles + W - 1) / W;
    GradAcc acc;
    acc_zero(acc);
    int cur = -1;
    for (int r = 0; r < rounds; ++r) {
        const int tile = r * W + gw;
        const int e = tile < ntiles ? tiles[tile].x : -1;
        if (e >= 0 && e != cur) {
            if (cur >= 0) acc_flush(acc, partials + ((size_t)gw * NN_MAXE + cur) * NN_NPAR, l);
            acc_zero(acc);
            cur = e;
        }
        if (l < 16) rowsb[wave][l] = (e >= 0 && l < tiles[tile].z) ? rows[tiles[tile].y + l] : -1;
        __syncthreads();
        grad_tile(M, e, rowsb[wave], tv, stage[wave], acc, lane);
    }
    if (cur >= 0) acc_flush(acc, partials + ((size_t)gw * NN_MAXE + cur) * NN_NPAR, l);
}

// fixed-order reduction of the partials; mode 0 writes the gradient to grad
// (padded layout, NN_MAXE x NN_NPAR), mode 1 also applies Adam to NnConst
__global__ void k_nn_adam(NnConst *Mg, const float *partials, int W, float *grad, float *am, float *av, float lr,
                          float beta1, float beta2, float eps, int step, int mode) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= NN_MAXE * NN_NPAR) return;
    const int e = idx / NN_NPAR, p = idx % NN_NPAR;
    if (e >= Mg->nelem) return;
    double g = 0.0;
    for (int w = 0; w < W; ++w) g += (double)partials[((size_t)w * NN_MAXE + e) * NN_NPAR + p];
    const int nin = Mg->nr[e] + Mg->na[e], n1 = Mg->nh1, n2 = Mg->nh2;
    float *par;
    bool real;
    if (p < NN_PB1) {
        const int k = p / NN_H, n = p % NN_H;
        par = &Mg->W1[e][k][n];
        real = k < nin && n < n1;
    } else if (p < NN_PW2) {
        par = &Mg->b1[e][p - NN_PB1];
        real = p - NN_PB1 < n1;
    } else if (p < NN_PB2) {
        const int k = (p - NN_PW2) / NN_H, n = (p - NN_PW2) % NN_H;
        par = &Mg->W2[e][k][n];
        real = k < n1 && n < n2;
    } else if (p < NN_PW3) {
        par = &Mg->b2[e][p - NN_PB2];
        real = p - NN_PB2 < n2;
    } else if (p < NN_PB3) {
        par = &Mg->w3[e][p - NN_PW3];
        real = p - NN_PW3 < n2;
    } else {
        par = &Mg->b3[e];
        real = true;
    }
    if (!real) g = 0.0;
    grad[idx] = (float)g;
    if (mode != 1 || !real) return;
    const float gf = (float)g;
    const float m = beta1 * am[idx] + (1.f - beta1) * gf;
    const float v = beta2 * av[idx] + (1.f - beta2) * gf * gf;
    am[idx] = m;
    av[idx] = v;
    const float mh = m / (1.f - powf(beta1, (float)step));
    const float vh = v / (1.f - powf(beta2, (float)step));
    *par -= lr * mh / (sqrtf(vh) + eps);
}

size_t mdt_nn_frame_smem(int N) { return (size_t)N * (3 * 8 + 3 * 4 + 3 * 4 + 1) + 16; }

hipError_t mdt_launch_nn_frame(int nblocks, size_t smem, const NnConst *M, const NnTrainView &tv, const int *frames,
                               int mode, double cEw, double cFw, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_frame, dim3(nblocks), dim3(NN_BLOCK), smem, st, M, tv, frames, mode, cEw, cFw);
    return hipGetLastError();
}

hipError_t mdt_launch_nn_mlp(int ntiles, const NnConst *M, const NnTrainView &tv, const int4 *tiles,
                             const int *rows, hipStream_t st) {
    const int nb = (ntiles + NN_BLOCK / 64 - 1) / (NN_BLOCK / 64);
    hipLaunchKernelGGL(k_nn_mlp, dim3(nb > 0 ? nb : 1), dim3(NN_BLOCK), 0, st, M, tv, tiles, ntiles, rows);
    return hipGetLastError();
}

hipError_t mdt_launch_nn_grad(int nblocks, const NnConst *M, const NnTrainView &tv, const int4 *tiles, int ntiles,
                              const int *rows, float *partials, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_grad, dim3(nblocks), dim3(NN_BLOCK), 0, st, M, tv, tiles, ntiles, rows, partials);
    return hipGetLastError();
}

hipError_t mdt_launch_nn_adam(NnConst *M, const float *partials, int W, float *grad, float *am, float *av, float lr,
                              float beta1, float beta2, float eps, int step, int mode, hipStream_t st) {
    const int n = NN_MAXE * NN_NPAR;
    hipLaunchKernelGGL(k_nn_adam, dim3((n + 255) / 256), dim3(256), 0, st, M, partials, W, grad, am, av, lr, beta1,
                       beta2, eps, step, mode);
    return hipGetLastError();
}

// dynamic LDS of the larger of k_nene_g (4 N doubles, u16 + u8 per atom)
// and k_nene_f (3 N doubles, 3 N u64, u8 per atom)
// the larger of k_nene_g's (23 B per atom) and k_nene_f's (37 B per atom)
size_t mdt_nene_smem(int N) { return (size_t)N * (3 * 8 + 3 * 4 + 1) + 16; }

// neighbour-list capacity per atom of the NN kernels (stored lists of the radial kernels)
int mdt_nene_cap() { return NN_CAP; }

// k_nene_rg: 27 B per atom, k_nene_rf: 43 B per atom (+ the slot padding)
size_t mdt_nene_radial_smem(int N) { return (size_t)N * (3 * 8 + 3 * 4 + 4 + 2 + 1) + 4 * 4 * NN_MAXE + 16; }

// static LDS: the larger of the two kernels'
hipError_t mdt_nene_attributes(hipFuncAttributes *fa, int radial) {
    hipFuncAttributes g, f;
    hipError_t e = hipFuncGetAttributes(&g, radial ? (const void *)k_nene_rg<true> : (const void *)k_nene_g);
    if (e != hipSuccess) return e;
    e = hipFuncGetAttributes(&f, radial ? (const void *)k_nene_rf<true> : (const void *)k_nene_f);
    if (e != hipSuccess) return e;
    *fa = g.sharedSizeBytes >= f.sharedSizeBytes ? g : f;
    return hipSuccess;
}

hipError_t mdt_launch_nene(int nrep, size_t smem, const NnConst *M, const BatchView &bv, float *G, float *C,
                           int *rep_flag, double *fpes, int maxN, int add_fin, int skip_extrap, int radial,
                           uint16_t *nbr, uint8_t *nbr_cnt, uint32_t *nbr_ref, int *nbr_key, double skin_bohr,
                           hipStream_t st, int images) {
    NnView nv{G, C, rep_flag, fpes, maxN, nbr, nbr_cnt, nbr_ref, nbr_key, (float)skin_bohr};
    if (radial)  // the image search costs k_nene_rg no occupancy: one instantiation
        hipLaunchKernelGGL(k_nene_rg<true>, dim3(nrep), dim3(NN_BLOCK), smem, st, M, bv, nv, skip_extrap);
    else
        hipLaunchKernelGGL(k_nene_g, dim3(nrep), dim3(NN_BLOCK), smem, st, M, bv, nv, skip_extrap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (radial && images)
        hipLaunchKernelGGL(k_nene_rf<true>, dim3(nrep), dim3(NN_BLOCK), smem, st, M, bv, nv, add_fin, skip_extrap);
    else if (radial)
        hipLaunchKernelGGL(k_nene_rf<false>, dim3(nrep), dim3(NN_BLOCK), smem, st, M, bv, nv, add_fin, skip_extrap);
    else
        hipLaunchKernelGGL(k_nene_f, dim3(nrep), dim3(NN_BLOCK), smem, st, M, bv, nv, add_fin, skip_extrap);
    return hipGetLastError();
}
