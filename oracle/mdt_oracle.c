/*
 * mdt_oracle.c -- TEST INFRASTRUCTURE.  Serial FP64 CPU restatement of the
 * md_tian2 reference algorithm for the EMT / propagator / RPMD hot path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, as the checker and as the CPU baseline; the product never does.
 *
 * The restatement keeps the reference's data structures and cost structure
 * on purpose (dense (3,P,N,N) derivative arrays, an N x N min-image matrix
 * rebuilt every step, per-atom propagator dispatch), so that timing it gives
 * the reference's serial cost.  Every function cites the reference lines it
 * follows.  Single-precision literals of the reference (1e-30, 1e-2, 1e-12,
 * 0.36) are reproduced as (double)<float literal>.
 *
 * Parity pins (tests/test_oracle_*.py): fcc bulk E=0/F=0 at the EMT
 * lattice constant, isolated-atom cohesive energy, energy separability,
 * central finite differences of the energy, sum F = 0, rotation /
 * translation / permutation invariance, min-image on hand-built cells,
 * Langevin against the exact Ornstein-Uhlenbeck update, Langevin(T=0,xi=0)
 * == Verlet, LDFA against numpy polyval, C^T C = I, the free ring polymer
 * against scipy expm, PILE coefficients, COM removal closed form, and the
 * Philox4x32-10 known-answer vectors.  Trajectory characteristics, the
 * energy-file estimators and the initial-condition routines
 * (tests/test_observe_cpu.py, tests/test_sampling.py): a hand-built state
 * sequence with known closest approach (through a periodic image), lowest
 * point, turning points, interaction time and adsorption; a free
 * projectile's exit step from straight flight; ring-polymer closed forms
 * (radius of gyration, spring energy, virial kinetic energy); hand-computed
 * apply_pip targets and incidence energies.
 */
#include "mdt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* constants.f90:33-39, 157-160 */
static const double ORC_PI = 3.14159265358979323846264338327950288;
#define KB 8.61733238496e-5
#define HBAR 0.6582119514467406
#define AMU2MASS 103.638239276
#define BOHR2ANG 0.529177211
#define TOLERANCE 1.0e-9
#define TWELFTH (1.0 / 12.0)

#define I3PN(c, b, i, P) ((c) + 3 * ((b) + (P) * (i)))
#define IPN(b, i, P) ((b) + (P) * (i))
#define I3PNN(c, b, i, j, P, N) ((c) + 3 * ((b) + (P) * ((i) + (N) * (j))))
#define IPNN(b, i, j, P, N) ((b) + (P) * ((i) + (N) * (j)))

static double orc_sqrt2(void) { return sqrt(2.0); }
static double orc_sqrt3(void) { return sqrt(3.0); }

/* pes_emt_mod.f90:1662-1676 select_beta('fcc') */
static double select_beta_fcc(void) {
    return pow(16.0 * ORC_PI / 3.0, 1.0 / 3.0) * (1.0 / orc_sqrt2());
}

/* universe_mod.f90:411-428: vec = r_j - r_i, direct coords, anint, back */
void orc_minimg_beads(const orc_universe *u, int i, int j, double *dist,
                      double *vec) {
    const int P = u->nbeads;
    for (int b = 0; b < P; ++b) {
        double d[3], s[3], e[3];
        for (int c = 0; c < 3; ++c)
            d[c] = u->r[I3PN(c, b, j, P)] - u->r[I3PN(c, b, i, P)];
        for (int k = 0; k < 3; ++k)
            s[k] = u->isimbox[k + 0] * d[0] + u->isimbox[k + 3] * d[1] +
                   u->isimbox[k + 6] * d[2];
        for (int k = 0; k < 3; ++k) s[k] = s[k] - round(s[k]);
        for (int k = 0; k < 3; ++k)
            e[k] = u->simbox[k + 0] * s[0] + u->simbox[k + 3] * s[1] +
                   u->simbox[k + 6] * s[2];
        for (int c = 0; c < 3; ++c) vec[c + 3 * b] = e[c];
        dist[b] = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    }
}

/* force.f90:58-74: distances(P,N,N), vectors(3,P,N,N) with
 * vectors(:,:,i,j) = r_i - r_j (min image) */
static void build_distance_matrix(const orc_universe *u, double *distances,
                                  double *vectors) {
    const int P = u->nbeads, N = u->natoms;
    double *td = (double *)malloc(sizeof(double) * P);
    double *tv = (double *)malloc(sizeof(double) * 3 * P);
    memset(distances, 0, sizeof(double) * (size_t)P * N * N);
    memset(vectors, 0, sizeof(double) * 3 * (size_t)P * N * N);
    for (int j = 0; j < N - 1; ++j)
        for (int i = j + 1; i < N; ++i) {
            orc_minimg_beads(u, i, j, td, tv);
            for (int b = 0; b < P; ++b) {
                distances[IPNN(b, i, j, P, N)] = td[b];
                distances[IPNN(b, j, i, P, N)] = td[b];
                for (int c = 0; c < 3; ++c) {
                    vectors[I3PNN(c, b, i, j, P, N)] = -tv[c + 3 * b];
                    vectors[I3PNN(c, b, j, i, P, N)] = +tv[c + 3 * b];
                }
            }
        }
    free(td);
    free(tv);
}

#define ALLOC0(n) ((double *)calloc((size_t)(n), sizeof(double)))

/* pes_emt_mod.f90:230-477 compute_emt_1species */
static int compute_emt_1species(orc_universe *u, const orc_emt_params *p,
                                const double *distances,
                                const double *vectors) {
    const int P = u->nbeads, N = u->natoms, T = u->ntypes;
    const double cutoff = 1.5;
    const double sqrt2 = orc_sqrt2(), sqrt3 = orc_sqrt3();
    int emt_idx = -1;
    for (int i = 0; i < T; ++i)
        if (u->pes[i + T * i] == ORC_PES_EMT) emt_idx = i;
    /* precondition of the reference's dense form: an EMT pair involving a
     * non-EMT species would read uninitialised exp_lambda_s (:407,:420) */
    for (int ti = 0; ti < T; ++ti)
        for (int tj = 0; tj < T; ++tj)
            if (u->pes[ti + T * tj] == ORC_PES_EMT &&
                (ti != emt_idx || tj != emt_idx))
                return 3;

    const double beta = select_beta_fcc();
    const double nneighs[3] = {12.0, 6.0, 24.0};
    const double betas0 = beta * p->s0[emt_idx];
    const double betaeta2 = beta * p->eta2[emt_idx];
    const double kappadbeta = p->kappa[emt_idx] / beta;
    double rnn[3] = {betas0, betas0 * sqrt2, betas0 * sqrt3};
    const double rcut = betas0 * sqrt3;
    const double rr = 4.0 * rcut / (sqrt3 + 2);
    const double acut = 9.21034037197618 / (rr - rcut);
    double x[3], r3temp[3];
    for (int k = 0; k < 3; ++k)
        x[k] = nneighs[k] * TWELFTH / (1 + exp(acut * (rnn[k] - rcut)));
    for (int k = 0; k < 3; ++k) r3temp[k] = rnn[k] - betas0;
    double sum1 = 0.0, sum2 = 0.0;
    for (int k = 0; k < 3; ++k) sum1 += x[k] * exp(-p->eta2[emt_idx] * r3temp[k]);
    for (int k = 0; k < 3; ++k) sum2 += x[k] * exp(-kappadbeta * r3temp[k]);
    const double igamma1 = 1.0 / sum1;
    const double igamma2 = 1.0 / sum2;

    double *V = ALLOC0(P), *Ecoh = ALLOC0(P), *vref = ALLOC0(P);
    double *s = ALLOC0(P * N), *exp_lambda_s = ALLOC0(P * N),
           *sigma = ALLOC0(P * N);
    double *dV = ALLOC0(3 * P * N), *dEcoh = ALLOC0(3 * P * N),
           *dvref = ALLOC0(3 * P * N);
    double *dsigma = ALLOC0((size_t)3 * P * N * N),
           *ds = ALLOC0((size_t)3 * P * N * N);
    double *rtemp_beads = ALLOC0(P);

    for (int i = 0; i < N - 1; ++i) {
        const int idx_i = u->idx[i];
        for (int j = i + 1; j < N; ++j) {
            const int idx_j = u->idx[j];
            if (u->pes[idx_i + T * idx_j] != ORC_PES_EMT) continue;
            for (int b = 0; b < P; ++b) {
                const double rijmag = distances[IPNN(b, i, j, P, N)];
                if (rijmag > cutoff * rcut) continue;
                double rtemp = exp(acut * (rijmag - rcut));
                const double theta = 1.0 / (1 + rtemp);
                const double rtemp1 = acut * rtemp * theta;
                rtemp = theta * exp(-p->eta2[emt_idx] * (rijmag - betas0));
                sigma[IPN(b, i, P)] += rtemp;
                sigma[IPN(b, j, P)] += rtemp;
                for (int c = 0; c < 3; ++c) {
                    const double dtheta = (p->eta2[emt_idx] + rtemp1) * rtemp *
                                          vectors[I3PNN(c, b, i, j, P, N)] /
                                          rijmag;
                    dsigma[I3PNN(c, b, i, i, P, N)] -= dtheta;
                    dsigma[I3PNN(c, b, j, j, P, N)] += dtheta;
                    dsigma[I3PNN(c, b, j, i, P, N)] = dtheta;
                    dsigma[I3PNN(c, b, i, j, P, N)] = -dtheta;
                }
                rtemp = theta * exp(-kappadbeta * (rijmag - betas0));
                V[b] += rtemp;
                for (int c = 0; c < 3; ++c) {
                    const double dtheta = (kappadbeta + rtemp1) * rtemp *
                                          vectors[I3PNN(c, b, i, j, P, N)] /
                                          rijmag;
                    dV[I3PN(c, b, i, P)] += dtheta;
                    dV[I3PN(c, b, j, P)] -= dtheta;
                }
            }
        }
    }
    for (int k = 0; k < P * N; ++k) sigma[k] *= igamma1;
    for (int b = 0; b < P; ++b) V[b] = V[b] * igamma2 * p->v0[emt_idx];
    for (size_t k = 0; k < (size_t)3 * P * N * N; ++k) dsigma[k] *= igamma1;
    for (int k = 0; k < 3 * P * N; ++k) dV[k] = dV[k] * igamma2 * p->v0[emt_idx];

    const double clamp30 = (double)1e-30f; /* :382 single-precision literal */
    for (int k = 0; k < P * N; ++k) s[k] = sigma[k] > clamp30 ? sigma[k] : clamp30;
    for (size_t k = 0; k < (size_t)3 * P * N * N; ++k) ds[k] = -dsigma[k];
    for (int i = 0; i < N; ++i) {
        const int idx_i = u->idx[i];
        for (int j = 0; j < N; ++j) {
            const int idx_j = u->idx[j];
            if (u->pes[idx_i + T * idx_j] == ORC_PES_EMT)
                for (int b = 0; b < P; ++b)
                    for (int c = 0; c < 3; ++c)
                        ds[I3PNN(c, b, i, j, P, N)] =
                            ds[I3PNN(c, b, i, j, P, N)] /
                            (betaeta2 * s[IPN(b, j, P)]);
        }
    }
    for (int k = 0; k < P * N; ++k) {
        const double t = s[k] > TOLERANCE ? s[k] : TOLERANCE;
        s[k] = -log(t * TWELFTH) / betaeta2;
    }
    const double rt1 = 0.5 / BOHR2ANG - betaeta2;
    for (int i = 0; i < N; ++i) {
        const int idx_i = u->idx[i];
        if (u->pes[idx_i + T * idx_i] == ORC_PES_EMT)
            for (int b = 0; b < P; ++b) {
                const double si = s[IPN(b, i, P)];
                u->dens[IPN(b, i, P)] = p->n0[idx_i] * exp(rt1 * si);
                exp_lambda_s[IPN(b, i, P)] = exp(-p->lambda[idx_i] * si);
                Ecoh[b] = Ecoh[b] + ((1 + p->lambda[idx_i] * si) *
                                         exp_lambda_s[IPN(b, i, P)] -
                                     1);
            }
    }
    for (int b = 0; b < P; ++b) Ecoh[b] = Ecoh[b] * p->e0[emt_idx];
    for (int i = 0; i < N; ++i) {
        const int idx_i = u->idx[i];
        for (int j = 0; j < N; ++j) {
            const int idx_j = u->idx[j];
            if (u->pes[idx_i + T * idx_j] == ORC_PES_EMT)
                for (int b = 0; b < P; ++b)
                    for (int c = 0; c < 3; ++c)
                        dEcoh[I3PN(c, b, i, P)] +=
                            s[IPN(b, j, P)] * exp_lambda_s[IPN(b, j, P)] *
                            ds[I3PNN(c, b, i, j, P, N)];
        }
    }
    for (int k = 0; k < 3 * P * N; ++k)
        dEcoh[k] = p->e0[emt_idx] * p->lambda[emt_idx] * p->lambda[emt_idx] * dEcoh[k];
    for (int i = 0; i < N; ++i) {
        if (u->idx[i] != emt_idx) continue;
        for (int b = 0; b < P; ++b) rtemp_beads[b] = exp(-p->kappa[emt_idx] * s[IPN(b, i, P)]);
        for (int b = 0; b < P; ++b) vref[b] += rtemp_beads[b];
        for (int j = 0; j < N; ++j)
            if (u->idx[j] == emt_idx)
                for (int b = 0; b < P; ++b)
                    for (int c = 0; c < 3; ++c)
                        dvref[I3PN(c, b, j, P)] +=
                            rtemp_beads[b] * ds[I3PNN(c, b, j, i, P, N)];
    }
    const double rtemp = 12 * p->v0[emt_idx];
    for (int b = 0; b < P; ++b) vref[b] = vref[b] * rtemp;
    for (int k = 0; k < 3 * P * N; ++k) dvref[k] = dvref[k] * rtemp * p->kappa[emt_idx];

    for (int b = 0; b < P; ++b) u->epot[b] = u->epot[b] + Ecoh[b] - V[b] + 0.5 * vref[b];
    for (int i = 0; i < N; ++i)
        if (u->idx[i] == emt_idx)
            for (int b = 0; b < P; ++b)
                for (int c = 0; c < 3; ++c) {
                    const int k = I3PN(c, b, i, P);
                    u->f[k] = u->f[k] + dEcoh[k] - dV[k] + 0.5 * dvref[k];
                }

    free(V); free(Ecoh); free(vref); free(s); free(exp_lambda_s); free(sigma);
    free(dV); free(dEcoh); free(dvref); free(dsigma); free(ds); free(rtemp_beads);
    return 0;
}

/* pes_emt_mod.f90:483-1027 compute_emt_2species */
static int compute_emt_2species(orc_universe *u, const orc_emt_params *p,
                                const double *distances,
                                const double *vectors) {
    const int P = u->nbeads, N = u->natoms, T = u->ntypes;
    const double cutoff = 1.5;
    const double sqrt2 = orc_sqrt2(), sqrt3 = orc_sqrt3();
    int type1 = -1, type2 = -1;
    for (int i = 0; i < T; ++i)
        if (u->pes[i + T * i] == ORC_PES_EMT) {
            if (type1 < 0) type1 = i; else type2 = i;
        }
    const double beta = select_beta_fcc();
    const double nneighs[3] = {12.0, 6.0, 24.0};
    const double betas0_1 = beta * p->s0[type1], betas0_2 = beta * p->s0[type2];
    const double betaeta2_1 = beta * p->eta2[type1], betaeta2_2 = beta * p->eta2[type2];
    const double kappadbeta_1 = p->kappa[type1] / beta, kappadbeta_2 = p->kappa[type2] / beta;
    const double chi_21 = p->n0[type1] / p->n0[type2] *
                          exp(0.5 / BOHR2ANG * (p->s0[type2] - p->s0[type1]));
    const double chi_12 = 1.0 / chi_21;
    const double rnn_1[3] = {betas0_1, betas0_1 * sqrt2, betas0_1 * sqrt3};
    const double rnn_2[3] = {betas0_2, betas0_2 * sqrt2, betas0_2 * sqrt3};
    const double rcut = (betas0_1 > betas0_2 ? betas0_1 : betas0_2) * sqrt3;
    const double rr = 4.0 * rcut / (sqrt3 + 2);
    const double acut = 9.21034037197618 / (rr - rcut);
    double x_1[3], x_2[3], r3temp[3];
    for (int k = 0; k < 3; ++k) {
        x_1[k] = nneighs[k] * TWELFTH / (1 + exp(acut * (rnn_1[k] - rcut)));
        x_2[k] = nneighs[k] * TWELFTH / (1 + exp(acut * (rnn_2[k] - rcut)));
    }
    double s11 = 0, s21 = 0, s12 = 0, s22 = 0;
    for (int k = 0; k < 3; ++k) r3temp[k] = rnn_1[k] - betas0_1;
    for (int k = 0; k < 3; ++k) s11 += x_1[k] * exp(-p->eta2[type1] * r3temp[k]);
    for (int k = 0; k < 3; ++k) s21 += x_1[k] * exp(-kappadbeta_1 * r3temp[k]);
    for (int k = 0; k < 3; ++k) r3temp[k] = rnn_2[k] - betas0_2;
    for (int k = 0; k < 3; ++k) s12 += x_2[k] * exp(-p->eta2[type2] * r3temp[k]);
    for (int k = 0; k < 3; ++k) s22 += x_2[k] * exp(-kappadbeta_2 * r3temp[k]);
    const double igamma1_1 = 1.0 / s11, igamma2_1 = 1.0 / s21;
    const double igamma1_2 = 1.0 / s12, igamma2_2 = 1.0 / s22;

    const size_t n3pn = (size_t)3 * P * N, n3pnn = (size_t)3 * P * N * N;
    double *V_12 = ALLOC0(P), *V_21 = ALLOC0(P), *V_22 = ALLOC0(P), *V_11 = ALLOC0(P);
    double *Ecoh_2 = ALLOC0(P), *Ecoh_1 = ALLOC0(P), *vref_2 = ALLOC0(P), *vref_1 = ALLOC0(P);
    double *rtemp_beads = ALLOC0(P);
    double *sigma_22 = ALLOC0(P * N), *sigma_21 = ALLOC0(P * N), *s_2 = ALLOC0(P * N);
    double *sigma_11 = ALLOC0(P * N), *sigma_12 = ALLOC0(P * N), *s_1 = ALLOC0(P * N);
    double *exp_lambda_s1 = ALLOC0(P * N), *exp_lambda_s2 = ALLOC0(P * N);
    double *dsigma_21_2 = ALLOC0(n3pn), *dsigma_12_1 = ALLOC0(n3pn);
    double *dEcoh_2_2 = ALLOC0(n3pn), *dEcoh_2_1 = ALLOC0(n3pn),
           *dEcoh_1_2 = ALLOC0(n3pn), *dEcoh_1_1 = ALLOC0(n3pn);
    double *dV_22_2 = ALLOC0(n3pn), *dV_11_1 = ALLOC0(n3pn), *dV_21_2 = ALLOC0(n3pn),
           *dV_21_1 = ALLOC0(n3pn), *dV_12_1 = ALLOC0(n3pn), *dV_12_2 = ALLOC0(n3pn);
    double *dvref_2_2 = ALLOC0(n3pn), *dvref_2_1 = ALLOC0(n3pn),
           *dvref_1_2 = ALLOC0(n3pn), *dvref_1_1 = ALLOC0(n3pn);
    double *dsigma_22 = ALLOC0(n3pnn), *dsigma_21_1 = ALLOC0(n3pnn);
    double *dsigma_11 = ALLOC0(n3pnn), *dsigma_12_2 = ALLOC0(n3pnn);
    double *ds_2_2 = ALLOC0(n3pnn), *ds_1_1 = ALLOC0(n3pnn),
           *ds_2_1 = ALLOC0(n3pnn), *ds_1_2 = ALLOC0(n3pnn);

    /* pair loop :646-763 */
    for (int i = 0; i < N - 1; ++i) {
        const int idx_i = u->idx[i];
        for (int j = i + 1; j < N; ++j) {
            const int idx_j = u->idx[j];
            if (u->pes[idx_i + T * idx_j] != ORC_PES_EMT) continue;
            for (int b = 0; b < P; ++b) {
                const double rijmag = distances[IPNN(b, i, j, P, N)];
                if (rijmag > cutoff * rcut) continue;
                double rtemp = exp(acut * (rijmag - rcut));
                const double theta = 1.0 / (1 + rtemp);
                const double rtemp1 = acut * rtemp * theta;
                double rij[3];
                for (int c = 0; c < 3; ++c) rij[c] = vectors[I3PNN(c, b, i, j, P, N)];
                if (idx_i == type2 && idx_j == type2) {
                    rtemp = theta * exp(-p->eta2[type2] * (rijmag - betas0_2));
                    sigma_22[IPN(b, i, P)] += rtemp;
                    sigma_22[IPN(b, j, P)] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        const double dt = (p->eta2[type2] + rtemp1) * rtemp * rij[c] / rijmag;
                        dsigma_22[I3PNN(c, b, i, i, P, N)] -= dt;
                        dsigma_22[I3PNN(c, b, j, j, P, N)] += dt;
                        dsigma_22[I3PNN(c, b, j, i, P, N)] = dt;
                        dsigma_22[I3PNN(c, b, i, j, P, N)] = -dt;
                    }
                    rtemp = theta * exp(-kappadbeta_2 * (rijmag - betas0_2));
                    V_22[b] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        const double dt = (kappadbeta_2 + rtemp1) * rtemp * rij[c] / rijmag;
                        dV_22_2[I3PN(c, b, i, P)] += dt;
                        dV_22_2[I3PN(c, b, j, P)] -= dt;
                    }
                } else if (idx_i == type1 && idx_j == type1) {
                    rtemp = theta * exp(-p->eta2[type1] * (rijmag - betas0_1));
                    sigma_11[IPN(b, i, P)] += rtemp;
                    sigma_11[IPN(b, j, P)] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        /* :705 uses eta2(type2) here (reference quirk) */
                        const double dt = (p->eta2[type2] + rtemp1) * rtemp * rij[c] / rijmag;
                        dsigma_11[I3PNN(c, b, i, i, P, N)] -= dt;
                        dsigma_11[I3PNN(c, b, j, j, P, N)] += dt;
                        dsigma_11[I3PNN(c, b, j, i, P, N)] = dt;
                        dsigma_11[I3PNN(c, b, i, j, P, N)] = -dt;
                    }
                    rtemp = theta * exp(-kappadbeta_1 * (rijmag - betas0_1));
                    V_11[b] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        const double dt = (kappadbeta_1 + rtemp1) * rtemp * rij[c] / rijmag;
                        dV_11_1[I3PN(c, b, i, P)] += dt;
                        dV_11_1[I3PN(c, b, j, P)] -= dt;
                    }
                } else if (idx_i != idx_j) {
                    /* :721-759 assumes i = type1 (projectile first) */
                    rtemp = theta * exp(-p->eta2[type1] * (rijmag - betas0_1));
                    sigma_21[IPN(b, j, P)] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        const double dt = (p->eta2[type1] + rtemp1) * rtemp * rij[c] / rijmag;
                        dsigma_21_2[I3PN(c, b, j, P)] += dt;
                        dsigma_21_1[I3PNN(c, b, i, j, P, N)] = -dt;
                    }
                    rtemp = theta * exp(-p->eta2[type2] * (rijmag - betas0_2));
                    sigma_12[IPN(b, i, P)] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        const double dt = (p->eta2[type2] + rtemp1) * rtemp * rij[c] / rijmag;
                        dsigma_12_1[I3PN(c, b, i, P)] -= dt;
                        dsigma_12_2[I3PNN(c, b, j, i, P, N)] = dt;
                    }
                    rtemp = theta * exp(-kappadbeta_1 * (rijmag - betas0_1));
                    V_21[b] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        const double dt = (kappadbeta_1 + rtemp1) * rtemp * rij[c] / rijmag;
                        dV_21_1[I3PN(c, b, i, P)] += dt;
                        dV_21_2[I3PN(c, b, j, P)] -= dt;
                    }
                    rtemp = theta * exp(-kappadbeta_2 * (rijmag - betas0_2));
                    V_12[b] += rtemp;
                    for (int c = 0; c < 3; ++c) {
                        const double dt = (kappadbeta_2 + rtemp1) * rtemp * rij[c] / rijmag;
                        dV_12_1[I3PN(c, b, i, P)] += dt;
                        dV_12_2[I3PN(c, b, j, P)] -= dt;
                    }
                }
            }
        }
    }

    /* :767-788 */
    for (int k = 0; k < P * N; ++k) {
        sigma_22[k] *= igamma1_2;
        sigma_11[k] *= igamma1_1;
        sigma_21[k] *= igamma1_2;
        sigma_12[k] *= igamma1_1;
    }
    for (int b = 0; b < P; ++b) {
        V_22[b] = V_22[b] * igamma2_2 * p->v0[type2];
        V_11[b] = V_11[b] * igamma2_1 * p->v0[type1];
        V_21[b] = V_21[b] * igamma2_2 * p->v0[type2] * chi_21;
        V_12[b] = V_12[b] * igamma2_1 * p->v0[type1] * chi_12;
    }
    for (size_t k = 0; k < n3pnn; ++k) {
        dsigma_22[k] *= igamma1_2;
        dsigma_11[k] *= igamma1_1;
        dsigma_21_1[k] *= igamma1_2;
        dsigma_12_2[k] *= igamma1_1;
    }
    for (size_t k = 0; k < n3pn; ++k) {
        dsigma_21_2[k] *= igamma1_2;
        dsigma_12_1[k] *= igamma1_1;
        dV_22_2[k] = dV_22_2[k] * igamma2_2 * p->v0[type2];
        dV_11_1[k] = dV_11_1[k] * igamma2_1 * p->v0[type1];
        dV_21_2[k] = dV_21_2[k] * igamma2_2 * p->v0[type2] * chi_21;
        dV_21_1[k] = dV_21_1[k] * igamma2_2 * p->v0[type2] * chi_21;
        dV_12_2[k] = dV_12_2[k] * igamma2_1 * p->v0[type1] * chi_12;
        dV_12_1[k] = dV_12_1[k] * igamma2_1 * p->v0[type1] * chi_12;
    }

    /* :793-842 */
    const double clamp30 = (double)1e-30f;
    for (int k = 0; k < P * N; ++k) {
        const double a1 = sigma_11[k] + chi_12 * sigma_12[k];
        const double a2 = sigma_22[k] + chi_21 * sigma_21[k];
        s_1[k] = a1 > clamp30 ? a1 : clamp30;
        s_2[k] = a2 > clamp30 ? a2 : clamp30;
    }
    for (size_t k = 0; k < n3pnn; ++k) {
        ds_1_1[k] = -dsigma_11[k];
        ds_2_2[k] = -dsigma_22[k];
    }
    for (int i = 0; i < N; ++i) {
        if (u->idx[i] == type2) {
            for (int b = 0; b < P; ++b)
                for (int c = 0; c < 3; ++c)
                    ds_2_2[I3PNN(c, b, i, i, P, N)] -= chi_21 * dsigma_21_2[I3PN(c, b, i, P)];
            for (int j = 0; j < N; ++j) {
                if (u->idx[j] == type2) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            ds_2_2[I3PNN(c, b, i, j, P, N)] =
                                ds_2_2[I3PNN(c, b, i, j, P, N)] / (betaeta2_2 * s_2[IPN(b, j, P)]);
                } else if (u->idx[j] == type1) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            ds_1_2[I3PNN(c, b, i, j, P, N)] =
                                -chi_12 * dsigma_12_2[I3PNN(c, b, i, j, P, N)] /
                                (betaeta2_1 * s_1[IPN(b, j, P)]);
                }
            }
        } else if (u->idx[i] == type1) {
            for (int b = 0; b < P; ++b)
                for (int c = 0; c < 3; ++c)
                    ds_1_1[I3PNN(c, b, i, i, P, N)] -= chi_12 * dsigma_12_1[I3PN(c, b, i, P)];
            for (int j = 0; j < N; ++j) {
                if (u->idx[j] == type1) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            ds_1_1[I3PNN(c, b, i, j, P, N)] =
                                ds_1_1[I3PNN(c, b, i, j, P, N)] / (betaeta2_1 * s_1[IPN(b, j, P)]);
                } else if (u->idx[j] == type2) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            ds_2_1[I3PNN(c, b, i, j, P, N)] =
                                -chi_21 * dsigma_21_1[I3PNN(c, b, i, j, P, N)] /
                                (betaeta2_2 * s_2[IPN(b, j, P)]);
                }
            }
        }
    }
    /* :844-845 */
    for (int k = 0; k < P * N; ++k) {
        const double t1 = s_1[k] > TOLERANCE ? s_1[k] : TOLERANCE;
        const double t2 = s_2[k] > TOLERANCE ? s_2[k] : TOLERANCE;
        s_1[k] = -log(t1 * TWELFTH) / betaeta2_1;
        s_2[k] = -log(t2 * TWELFTH) / betaeta2_2;
    }
    /* :849-869 */
    const double rtemp1c = 0.5 / BOHR2ANG - betaeta2_1;
    const double rtemp2c = 0.5 / BOHR2ANG - betaeta2_2;
    for (int i = 0; i < N; ++i) {
        if (u->idx[i] == type1) {
            for (int b = 0; b < P; ++b) {
                const double s = s_1[IPN(b, i, P)];
                u->dens[IPN(b, i, P)] = p->n0[type1] * exp(rtemp1c * s);
                exp_lambda_s1[IPN(b, i, P)] = exp(-p->lambda[type1] * s);
                Ecoh_1[b] = Ecoh_1[b] + ((1 + p->lambda[type1] * s) * exp_lambda_s1[IPN(b, i, P)] - 1);
            }
        } else if (u->idx[i] == type2) {
            for (int b = 0; b < P; ++b) {
                const double s = s_2[IPN(b, i, P)];
                u->dens[IPN(b, i, P)] = p->n0[type2] * exp(rtemp2c * s);
                exp_lambda_s2[IPN(b, i, P)] = exp(-p->lambda[type2] * s);
                Ecoh_2[b] = Ecoh_2[b] + ((1 + p->lambda[type2] * s) * exp_lambda_s2[IPN(b, i, P)] - 1);
            }
        }
    }
    for (int b = 0; b < P; ++b) {
        Ecoh_1[b] = Ecoh_1[b] * p->e0[type1];
        Ecoh_2[b] = Ecoh_2[b] * p->e0[type2];
    }
    /* :871-914 */
    for (int i = 0; i < N; ++i) {
        if (u->idx[i] == type2) {
            for (int j = 0; j < N; ++j) {
                if (u->idx[j] == type2) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dEcoh_2_2[I3PN(c, b, i, P)] += s_2[IPN(b, j, P)] * exp_lambda_s2[IPN(b, j, P)] *
                                                           ds_2_2[I3PNN(c, b, i, j, P, N)];
                } else if (u->idx[j] == type1) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dEcoh_1_2[I3PN(c, b, i, P)] += s_1[IPN(b, j, P)] * exp_lambda_s1[IPN(b, j, P)] *
                                                           ds_1_2[I3PNN(c, b, i, j, P, N)];
                }
            }
        } else if (u->idx[i] == type1) {
            for (int j = 0; j < N; ++j) {
                if (u->idx[j] == type2) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dEcoh_2_1[I3PN(c, b, i, P)] += s_2[IPN(b, j, P)] * exp_lambda_s2[IPN(b, j, P)] *
                                                           ds_2_1[I3PNN(c, b, i, j, P, N)];
                } else if (u->idx[j] == type1) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dEcoh_1_1[I3PN(c, b, i, P)] += s_1[IPN(b, j, P)] * exp_lambda_s1[IPN(b, j, P)] *
                                                           ds_1_1[I3PNN(c, b, i, j, P, N)];
                }
            }
        }
    }
    const double f22 = p->e0[type2] * p->lambda[type2] * p->lambda[type2];
    const double f11 = p->e0[type1] * p->lambda[type1] * p->lambda[type1];
    for (size_t k = 0; k < n3pn; ++k) {
        dEcoh_2_2[k] = f22 * dEcoh_2_2[k];
        dEcoh_1_2[k] = f11 * dEcoh_1_2[k];
        dEcoh_2_1[k] = f22 * dEcoh_2_1[k];
        dEcoh_1_1[k] = f11 * dEcoh_1_1[k];
    }
    /* :918-972 */
    for (int i = 0; i < N; ++i) {
        if (u->idx[i] == type2) {
            for (int b = 0; b < P; ++b) rtemp_beads[b] = exp(-p->kappa[type2] * s_2[IPN(b, i, P)]);
            for (int b = 0; b < P; ++b) vref_2[b] += rtemp_beads[b];
            for (int j = 0; j < N; ++j) {
                if (u->idx[j] == type2) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dvref_2_2[I3PN(c, b, j, P)] += rtemp_beads[b] * ds_2_2[I3PNN(c, b, j, i, P, N)];
                } else if (u->idx[j] == type1) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dvref_2_1[I3PN(c, b, j, P)] += rtemp_beads[b] * ds_2_1[I3PNN(c, b, j, i, P, N)];
                }
            }
        } else if (u->idx[i] == type1) {
            for (int b = 0; b < P; ++b) rtemp_beads[b] = exp(-p->kappa[type1] * s_1[IPN(b, i, P)]);
            for (int b = 0; b < P; ++b) vref_1[b] += rtemp_beads[b];
            for (int j = 0; j < N; ++j) {
                if (u->idx[j] == type1) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dvref_1_1[I3PN(c, b, j, P)] += rtemp_beads[b] * ds_1_1[I3PNN(c, b, j, i, P, N)];
                } else if (u->idx[j] == type2) {
                    for (int b = 0; b < P; ++b)
                        for (int c = 0; c < 3; ++c)
                            dvref_1_2[I3PN(c, b, j, P)] += rtemp_beads[b] * ds_1_2[I3PNN(c, b, j, i, P, N)];
                }
            }
        }
    }
    double rt = 12 * p->v0[type2];
    for (int b = 0; b < P; ++b) vref_2[b] = vref_2[b] * rt;
    for (size_t k = 0; k < n3pn; ++k) {
        dvref_2_2[k] = dvref_2_2[k] * rt * p->kappa[type2];
        dvref_2_1[k] = dvref_2_1[k] * rt * p->kappa[type2];
    }
    rt = 12 * p->v0[type1];
    for (int b = 0; b < P; ++b) vref_1[b] = vref_1[b] * rt;
    for (size_t k = 0; k < n3pn; ++k) {
        dvref_1_2[k] = dvref_1_2[k] * rt * p->kappa[type1];
        dvref_1_1[k] = dvref_1_1[k] * rt * p->kappa[type1];
    }
    /* :976 */
    for (int b = 0; b < P; ++b)
        u->epot[b] = u->epot[b] + (Ecoh_2[b] + Ecoh_1[b] - V_22[b] - V_11[b] -
                                   0.5 * (V_21[b] + V_12[b] - vref_2[b] - vref_1[b]));
    /* :1001-1011 */
    for (int i = 0; i < N; ++i)
        for (int b = 0; b < P; ++b)
            for (int c = 0; c < 3; ++c) {
                const int k = I3PN(c, b, i, P);
                if (u->idx[i] == type2)
                    u->f[k] = u->f[k] + (dEcoh_2_2[k] + dEcoh_1_2[k] - dV_22_2[k] -
                                         0.5 * (dV_21_2[k] + dV_12_2[k] - dvref_2_2[k] - dvref_1_2[k]));
                else if (u->idx[i] == type1)
                    u->f[k] = u->f[k] + (dEcoh_2_1[k] + dEcoh_1_1[k] - dV_11_1[k] -
                                         0.5 * (dV_21_1[k] + dV_12_1[k] - dvref_2_1[k] - dvref_1_1[k]));
            }

    free(V_12); free(V_21); free(V_22); free(V_11); free(Ecoh_2); free(Ecoh_1);
    free(vref_2); free(vref_1); free(rtemp_beads);
    free(sigma_22); free(sigma_21); free(s_2); free(sigma_11); free(sigma_12); free(s_1);
    free(exp_lambda_s1); free(exp_lambda_s2); free(dsigma_21_2); free(dsigma_12_1);
    free(dEcoh_2_2); free(dEcoh_2_1); free(dEcoh_1_2); free(dEcoh_1_1);
    free(dV_22_2); free(dV_11_1); free(dV_21_2); free(dV_21_1); free(dV_12_1); free(dV_12_2);
    free(dvref_2_2); free(dvref_2_1); free(dvref_1_2); free(dvref_1_1);
    free(dsigma_22); free(dsigma_21_1); free(dsigma_11); free(dsigma_12_2);
    free(ds_2_2); free(ds_1_1); free(ds_2_1); free(ds_1_2);
    return 0;
}

/* pes_emt_mod.f90:194-225 */
int orc_compute_emt(orc_universe *u, const orc_emt_params *p,
                    const double *distances, const double *vectors) {
    int nemt = 0;
    for (int i = 0; i < u->ntypes; ++i)
        if (u->pes[i + u->ntypes * i] == ORC_PES_EMT) ++nemt;
    if (nemt == 1) return compute_emt_1species(u, p, distances, vectors);
    if (nemt == 2) return compute_emt_2species(u, p, distances, vectors);
    return 2; /* 0: no EMT species (reference stops); >=3: n-species path */
}

/* rpmd.f90:36-59 */
void orc_build_cjk(int nbeads, double *cjk) {
    const int P = nbeads;
    for (int j = 1; j <= P; ++j)
        for (int k = 0; k <= P - 1; ++k) {
            double v;
            if (k == 0)
                v = sqrt(1.0 / P);
            else if (1 <= k && k <= P / 2 - 1)
                v = sqrt(2.0 / P) * cos(2.0 * ORC_PI * j * k / P);
            else if (k == P / 2)
                v = sqrt(1.0 / P) * ((j % 2) ? -1.0 : 1.0);
            else
                v = sqrt(2.0 / P) * sin(2.0 * ORC_PI * j * k / P);
            cjk[(j - 1) + P * k] = v; /* cjk(j, k+1), column-major */
        }
}

/* rpmd.f90:309-452 */
void orc_ring_polymer_step(orc_universe *u, const orc_simparams *sp,
                           const double *cjk) {
    const int P = u->nbeads, N = u->natoms;
    const size_t n = (size_t)3 * P * N;
    double *saveR = ALLOC0(n), *saveV = ALLOC0(n), *p = ALLOC0(n), *q = ALLOC0(n);
    double *newP = ALLOC0(n), *newQ = ALLOC0(n);
    double *poly = ALLOC0(4 * P);
    memcpy(saveR, u->r, n * sizeof(double));
    memcpy(saveV, u->v, n * sizeof(double));
    const double kt = KB * sp->Tproj * P / HBAR;
    const double prefactor = -1 * u->m[0] * (kt * kt);
    for (int i = 0; i < N; ++i)
        for (int b = 0; b < P; ++b)
            for (int c = 0; c < 3; ++c) p[I3PN(c, b, i, P)] = u->m[i] * u->v[I3PN(c, b, i, P)];
    memcpy(q, u->r, n * sizeof(double));
    for (int b = 0; b < P; ++b)
        for (int k = 0; k < P; ++k)
            for (int i = 0; i < N; ++i)
                for (int c = 0; c < 3; ++c) {
                    newP[I3PN(c, b, i, P)] += p[I3PN(c, k, i, P)] * cjk[k + P * b];
                    newQ[I3PN(c, b, i, P)] += q[I3PN(c, k, i, P)] * cjk[k + P * b];
                }
    memcpy(p, newP, n * sizeof(double));
    memcpy(q, newQ, n * sizeof(double));
    const double piN = ORC_PI / P;
    for (int i = 0; i < N; ++i) {
        const double mass = u->m[i];
        poly[0] = 1.0;
        poly[1] = 0.0;
        poly[2] = sp->step / mass;
        poly[3] = 1.0;
        double betaN;
        if (u->is_proj[u->idx[i]])
            betaN = 1.0 / (KB * sp->Tproj * P);
        else
            betaN = 1.0 / (KB * sp->Tsurf * P);
        if (P > 1) {
            const double twown = 2.0 / betaN / HBAR;
            for (int b = 1; b <= P / 2; ++b) {
                const double wk = twown * sin(b * piN);
                const double wt = wk * sp->step;
                const double wm = wk * mass;
                const double cos_wt = cos(wt), sin_wt = sin(wt);
                poly[0 + 4 * b] = cos_wt;
                poly[1 + 4 * b] = -wm * sin_wt;
                poly[2 + 4 * b] = sin_wt / wm;
                poly[3 + 4 * b] = cos_wt;
            }
            for (int b = 1; b <= (P - 1) / 2; ++b)
                for (int k = 0; k < 4; ++k) poly[k + 4 * (P - b)] = poly[k + 4 * b];
        }
        for (int b = 0; b < P; ++b)
            for (int c = 0; c < 3; ++c) {
                const double pp = p[I3PN(c, b, i, P)], qq = q[I3PN(c, b, i, P)];
                const double p_new = pp * poly[0 + 4 * b] + qq * poly[1 + 4 * b];
                q[I3PN(c, b, i, P)] = pp * poly[2 + 4 * b] + qq * poly[3 + 4 * b];
                p[I3PN(c, b, i, P)] = p_new;
            }
    }
    for (size_t k = 0; k < n; ++k)
        if (!u->is_fixed[k]) {
            u->r[k] = 0.0;
            u->v[k] = 0.0;
        }
    for (int i = 0; i < N; ++i)
        for (int b = 0; b < P; ++b)
            for (int k = 0; k < P; ++k)
                for (int c = 0; c < 3; ++c) {
                    const int t = I3PN(c, b, i, P);
                    if (!u->is_fixed[t]) {
                        u->r[t] = u->r[t] + q[I3PN(c, k, i, P)] * cjk[b + P * k];
                        u->v[t] = u->v[t] + p[I3PN(c, k, i, P)] * cjk[b + P * k] / u->m[i];
                    }
                }
    int any_lan = 0;
    for (int t = 0; t < u->ntypes; ++t)
        if (u->algo[t] == ORC_PROP_LANGEVIN) any_lan = 1;
    if (any_lan) {
        for (int i = 0; i < N; ++i) {
            if (u->algo[u->idx[i]] != ORC_PROP_LANGEVIN) continue;
            for (int b = 0; b < P; ++b)
                for (int c = 0; c < 3; ++c) {
                    u->r[I3PN(c, b, i, P)] = saveR[I3PN(c, b, i, P)];
                    u->v[I3PN(c, b, i, P)] = saveV[I3PN(c, b, i, P)];
                }
            for (int b = 0; b < P; ++b) {
                const int prev = b == 0 ? P - 1 : b - 1;
                const int next = b == P - 1 ? 0 : b + 1;
                for (int c = 0; c < 3; ++c)
                    u->f[I3PN(c, b, i, P)] =
                        prefactor * (2 * u->r[I3PN(c, b, i, P)] - u->r[I3PN(c, prev, i, P)] -
                                     u->r[I3PN(c, next, i, P)]);
            }
        }
    }
    free(saveR); free(saveV); free(p); free(q); free(newP); free(newQ); free(poly);
}

/* EMT energy/forces/dens only (compute_emt behind calc_force without the
 * ring-polymer step): zero f and epot, build the matrices, call EMT. */
int orc_emt_forces(orc_universe *u, const orc_emt_params *p) {
    const int P = u->nbeads, N = u->natoms;
    memset(u->f, 0, (size_t)3 * P * N * sizeof(double));
    memset(u->epot, 0, P * sizeof(double));
    double *distances = (double *)malloc(sizeof(double) * (size_t)P * N * N);
    double *vectors = (double *)malloc(sizeof(double) * 3 * (size_t)P * N * N);
    build_distance_matrix(u, distances, vectors);
    const int rc = orc_compute_emt(u, p, distances, vectors);
    free(distances);
    free(vectors);
    return rc;
}

/* force.f90:40-98; flag == energy_and_force */
int orc_calc_force(orc_universe *u, const orc_emt_params *p,
                   const orc_simparams *sp, const double *cjk) {
    const int P = u->nbeads, N = u->natoms, T = u->ntypes;
    const size_t n = (size_t)3 * P * N;
    memset(u->f, 0, n * sizeof(double));
    memset(u->a, 0, n * sizeof(double));
    memset(u->epot, 0, P * sizeof(double));
    if (P > 1) orc_ring_polymer_step(u, sp, cjk);
    double *distances = (double *)malloc(sizeof(double) * (size_t)P * N * N);
    double *vectors = (double *)malloc(sizeof(double) * 3 * (size_t)P * N * N);
    build_distance_matrix(u, distances, vectors);
    int any_emt = 0, rc = 0;
    for (int k = 0; k < T * T; ++k)
        if (u->pes[k] == ORC_PES_EMT) any_emt = 1;
    if (any_emt) rc = orc_compute_emt(u, p, distances, vectors);
    free(distances);
    free(vectors);
    /* set_acceleration :89-98 */
    for (int i = 0; i < N; ++i)
        for (int b = 0; b < P; ++b)
            for (int c = 0; c < 3; ++c) {
                const int k = I3PN(c, b, i, P);
                if (!u->is_fixed[k]) u->a[k] = u->f[k] / u->m[i];
            }
    return rc;
}

/* md_algo.f90:110-128 */
static void verlet_1(orc_universe *u, const orc_simparams *sp, int i) {
    const int P = u->nbeads;
    for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c) {
            const int k = I3PN(c, b, i, P);
            if (!u->is_fixed[k])
                u->v[k] = u->v[k] + 0.5 * sp->step * u->a[k];
            else
                u->v[k] = 0.0;
        }
    if (P == 1)
        for (int c = 0; c < 3; ++c) {
            const int k = I3PN(c, 0, i, P);
            if (!u->is_fixed[k]) u->r[k] = u->r[k] + sp->step * u->v[k];
        }
}

/* md_algo.f90:132-142 */
static void verlet_2(orc_universe *u, const orc_simparams *sp, int i) {
    const int P = u->nbeads;
    for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c) {
            const int k = I3PN(c, b, i, P);
            if (!u->is_fixed[k])
                u->v[k] = u->v[k] + 0.5 * sp->step * u->a[k];
            else
                u->v[k] = 0.0;
        }
}

/* Coefficients shared by langevin_1 (md_algo.f90:165-198) and langevin_2
 * (:253-285). Returns via arrays of length P. */
static void langevin_coefs(const orc_universe *u, const orc_simparams *sp, int i,
                           double *c0, double *c1, double *c2, double *sigma_r,
                           double *sigma_v, double *c_rv) {
    const int P = u->nbeads;
    double temp;
    if (u->is_proj[u->idx[i]])
        temp = KB * P * sp->Tproj / u->m[i];
    else
        temp = KB * P * sp->Tsurf / u->m[i];
    const double thr = (double)1e-2f; /* :176 single-precision literal */
    int all_big = 1;
    for (int b = 0; b < P; ++b)
        if (!(u->dens[IPN(b, i, P)] * sp->step > thr)) all_big = 0;
    for (int b = 0; b < P; ++b) {
        const double xidt = u->dens[IPN(b, i, P)] * sp->step;
        const double xidt2 = xidt * xidt;
        const double ixidt = sp->step / xidt;
        if (all_big && temp > TOLERANCE) {
            c0[b] = exp(-xidt);
            c1[b] = (1 - c0[b]) * ixidt;
            c2[b] = (1 - c1[b] / sp->step) * ixidt;
            sigma_r[b] = ixidt * sqrt(temp * (2 * xidt - 3 + 4 * c0[b] - c0[b] * c0[b]));
            sigma_v[b] = sqrt(temp * (1 - c0[b] * c0[b]));
            const double om = 1 - c0[b];
            c_rv[b] = ixidt * temp * (om * om) / (sigma_r[b] * sigma_v[b]);
        } else {
            c0[b] = 1 - xidt + 0.5 * xidt2;
            c1[b] = (1 - 0.5 * xidt + 2 * TWELFTH * xidt2) * sp->step;
            c2[b] = (0.5 - 2 * TWELFTH * xidt + 0.5 * TWELFTH * xidt2) * sp->step;
            sigma_r[b] = sp->step * sqrt(temp * (8 * TWELFTH * xidt - 0.5 * xidt2));
            sigma_v[b] = sqrt(2 * temp * xidt * (1 - xidt));
            c_rv[b] = 0.5 * orc_sqrt3() * (1 - 0.125 * xidt);
        }
    }
}

/* md_algo.f90:146-232 */
static void langevin_1(orc_universe *u, const orc_simparams *sp, int i,
                       const double *randy) {
    const int P = u->nbeads;
    double c0[256], c1[256], c2[256], sr[256], sv[256], crv[256];
    langevin_coefs(u, sp, i, c0, c1, c2, sr, sv, crv);
    for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c) {
            const int k = I3PN(c, b, i, P);
            if (!u->is_fixed[k]) {
                u->r[k] = u->r[k] + c1[b] * u->v[k] + c2[b] * sp->step * u->a[k] + sr[b] * randy[k];
                u->v[k] = c0[b] * u->v[k] + (c1[b] - c2[b]) * u->a[k] + sv[b] * crv[b] * randy[k];
            } else {
                u->v[k] = 0.0;
            }
        }
}

/* md_algo.f90:236-300 */
static void langevin_2(orc_universe *u, const orc_simparams *sp, int i,
                       const double *randy) {
    const int P = u->nbeads;
    double c0[256], c1[256], c2[256], sr[256], sv[256], crv[256];
    langevin_coefs(u, sp, i, c0, c1, c2, sr, sv, crv);
    for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c) {
            const int k = I3PN(c, b, i, P);
            if (!u->is_fixed[k])
                u->v[k] = u->v[k] + c2[b] * u->a[k] + sv[b] * sqrt(1 - crv[b] * crv[b]) * randy[k];
            else
                u->v[k] = 0.0;
        }
}

/* md_algo.f90:402-432; and_z unit normals scaled by sqrt(ibetaN/m) */
static void andersen(orc_universe *u, const orc_simparams *sp, int i,
                     const double *and_z, const double *and_u) {
    const int P = u->nbeads;
    const double ibetaN = u->is_proj[u->idx[i]] ? P * KB * sp->Tproj : P * KB * sp->Tsurf;
    const double sigma = sqrt(ibetaN / u->m[i]);
    const double thr = sp->step / sp->andersen_time;
    for (int b = 0; b < P; ++b)
        if (and_u[IPN(b, i, P)] < thr)
            for (int c = 0; c < 3; ++c) {
                const int k = I3PN(c, b, i, P);
                if (!u->is_fixed[k]) u->v[k] = and_z[k] * sigma + 0.0;
            }
}

/* md_algo.f90:440-502 */
static void pile_thermostat(orc_universe *u, const orc_simparams *sp, int i,
                            const double *zeta, const double *cjk) {
    const int P = u->nbeads;
    double betaN;
    if (u->is_proj[u->idx[i]])
        betaN = 1.0 / (KB * sp->Tproj * P);
    else
        betaN = 1.0 / (KB * sp->Tsurf * P);
    double *newP = ALLOC0(3 * P), *atomP = ALLOC0(3 * P), *c1 = ALLOC0(P),
           *c2 = ALLOC0(P), *gk = ALLOC0(P);
    for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c) atomP[c + 3 * b] = u->m[i] * u->v[I3PN(c, b, i, P)];
    for (int b = 0; b < P; ++b)
        for (int k = 0; k < P; ++k)
            for (int c = 0; c < 3; ++c) newP[c + 3 * b] += atomP[c + 3 * k] * cjk[k + P * b];
    for (int k = 0; k < P; ++k) {
        if (k == 0) {
            gk[k] = 1 / sp->pile_tau;
        } else {
            const double wn = 1 / betaN / HBAR;
            const double wk = 2 * wn * sin(k * ORC_PI / P);
            gk[k] = 2 * wk;
        }
    }
    for (int k = 0; k < P; ++k) {
        c1[k] = exp(-0.5 * sp->step * gk[k]);
        c2[k] = sqrt(1 - c1[k] * c1[k]);
    }
    const double sq = sqrt(u->m[i] / betaN);
    for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c)
            newP[c + 3 * b] = c1[b] * newP[c + 3 * b] + sq * c2[b] * zeta[I3PN(c, b, i, P)];
    memset(atomP, 0, sizeof(double) * 3 * P);
    for (int b = 0; b < P; ++b)
        for (int k = 0; k < P; ++k)
            for (int c = 0; c < 3; ++c)
                if (!u->is_fixed[I3PN(c, b, i, P)])
                    atomP[c + 3 * b] = atomP[c + 3 * b] + newP[c + 3 * k] * cjk[b + P * k];
    for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c) u->v[I3PN(c, b, i, P)] = atomP[c + 3 * b] / u->m[i];
    free(newP); free(atomP); free(c1); free(c2); free(gk);
}

/* universe_mod.f90:222-241 with calc_com_velocity :294-309 */
void orc_remove_com_velocity(orc_universe *u) {
    const int P = u->nbeads, N = u->natoms;
    double msum = 0.0;
    for (int i = 0; i < N; ++i) msum += u->m[i];
    for (int it = 0; it < 1000; ++it) {
        double num[3] = {0, 0, 0};
        for (int i = 0; i < N; ++i) {
            double sb[3] = {0, 0, 0};
            for (int b = 0; b < P; ++b)
                for (int c = 0; c < 3; ++c) sb[c] += u->v[I3PN(c, b, i, P)];
            for (int c = 0; c < 3; ++c) num[c] = num[c] + sb[c] * u->m[i];
        }
        double vcm[3];
        for (int c = 0; c < 3; ++c) vcm[c] = num[c] / msum / P / N;
        for (int i = 0; i < N; ++i)
            for (int b = 0; b < P; ++b)
                for (int c = 0; c < 3; ++c) u->v[I3PN(c, b, i, P)] -= vcm[c];
    }
}

/* md_algo.f90:506-570 */
int orc_ldfa(orc_universe *u, int i) {
    const int P = u->nbeads;
    const double convert = 1.00794 * AMU2MASS;
    static const double coefs[12] = {0.0802484, -1.12851, 9.28508, 2.10064,
                                     -843.419, 8.85354e3, -4.89023e4, 1.6741e5,
                                     -3.67098e5, 5.03476e5, -3.9426e5, 1.34763e5};
    const double lo = (double)(-1e-12f), hi = (double)0.36f; /* :542 */
    for (int b = 0; b < P; ++b) {
        double fric = u->dens[IPN(b, i, P)];
        if (lo <= fric && fric <= hi) {
            fric = 0.0;
            double temp = u->dens[IPN(b, i, P)];
            for (int j = 0; j < 12; ++j) {
                fric = fric + coefs[j] * temp;
                temp = temp * u->dens[IPN(b, i, P)];
            }
        } else if (fric > hi) {
            fric = 0.001 * (4.7131 - exp(-4.41305 * fric));
        } else {
            return 4; /* :551-554 negative density stops the reference */
        }
        u->dens[IPN(b, i, P)] = fric;
    }
    for (int b = 0; b < P; ++b)
        u->dens[IPN(b, i, P)] = u->dens[IPN(b, i, P)] * convert / HBAR / u->m[i];
    return 0;
}

static int any_algo(const orc_universe *u, int id) {
    for (int t = 0; t < u->ntypes; ++t)
        if (u->algo[t] == id) return 1;
    return 0;
}

/* md_algo.f90:37-76 */
int orc_propagate_1(orc_universe *u, const orc_simparams *sp,
                    const double *randy, const double *zeta,
                    const double *and_z, const double *and_u,
                    const double *cjk) {
    for (int i = 0; i < u->natoms; ++i) {
        switch (u->algo[u->idx[i]]) {
        case ORC_PROP_VERLET: verlet_1(u, sp, i); break;
        case ORC_PROP_ANDERSEN:
            verlet_1(u, sp, i);
            andersen(u, sp, i, and_z, and_u);
            break;
        case ORC_PROP_PILE:
            pile_thermostat(u, sp, i, zeta, cjk);
            verlet_1(u, sp, i);
            break;
        case ORC_PROP_LANGEVIN: langevin_1(u, sp, i, randy); break;
        default: return 5;
        }
    }
    if (any_algo(u, ORC_PROP_ANDERSEN) || any_algo(u, ORC_PROP_PILE)) orc_remove_com_velocity(u);
    return 0;
}

/* md_algo.f90:80-106 */
int orc_propagate_2(orc_universe *u, const orc_simparams *sp,
                    const double *randy) {
    for (int i = 0; i < u->natoms; ++i) {
        switch (u->algo[u->idx[i]]) {
        case ORC_PROP_VERLET:
        case ORC_PROP_ANDERSEN:
        case ORC_PROP_PILE: verlet_2(u, sp, i); break;
        case ORC_PROP_LANGEVIN: {
            const int rc = orc_ldfa(u, i);
            if (rc) return rc;
            langevin_2(u, sp, i, randy);
            break;
        }
        default: return 5;
        }
    }
    return 0;
}

/* md_tian2.f90:81-83 */
int orc_md_step(orc_universe *u, const orc_emt_params *p,
                const orc_simparams *sp, const double *cjk,
                const double *randy, const double *zeta, const double *and_z,
                const double *and_u) {
    int rc = orc_propagate_1(u, sp, randy, zeta, and_z, and_u, cjk);
    if (rc) return rc;
    rc = orc_calc_force(u, p, sp, cjk);
    if (rc) return rc;
    return orc_propagate_2(u, sp, randy);
}

/* ---- Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11.  Multipliers and Weyl key
 * increments are the published constants. ---- */
static void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words */
static double u53(uint32_t a, uint32_t b) {
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)m * (1.0 / 9007199254740992.0);
}

static void philox_draw(uint64_t seed, uint32_t traj, uint32_t step,
                        uint32_t stream, uint32_t slot, uint32_t comp,
                        uint32_t out[4]) {
    const uint32_t ctr[4] = {step, slot, stream, comp};
    const uint32_t key[2] = {(uint32_t)seed, traj};
    orc_philox4x32_10(ctr, key, out);
}

double orc_philox_uniform(uint64_t seed, uint32_t traj, uint32_t step,
                          uint32_t stream, uint32_t slot, uint32_t comp) {
    uint32_t o[4];
    philox_draw(seed, traj, step, stream, slot, comp, o);
    return u53(o[0], o[1]);
}

/* Box-Muller in the form of useful_things.f90:121-131, both branches: the
 * draw (step, slot, stream, comp / 2) gives component comp's deviate from its
 * cosine (even comp) or sine (odd comp), as the GPU path draws them; the
 * reference keeps the cosine of a fresh pair per deviate (same distribution) */
double orc_philox_normal(uint64_t seed, uint32_t traj, uint32_t step,
                         uint32_t stream, uint32_t slot, uint32_t comp) {
    uint32_t o[4];
    philox_draw(seed, traj, step, stream, slot, comp >> 1, o);
    const double rnd1 = 1.0 - u53(o[0], o[1]);
    const double rnd2 = u53(o[2], o[3]);
    const double branch = (comp & 1) ? sin(2 * ORC_PI * rnd2) : cos(2 * ORC_PI * rnd2);
    return sqrt(-2 * log(rnd1)) * branch * 1.0 + 0.0;
}

void orc_philox_fill_normals(uint64_t seed, uint32_t traj, uint32_t step,
                             uint32_t stream, int natoms, int nbeads,
                             double *out) {
    for (int i = 0; i < natoms; ++i)
        for (int b = 0; b < nbeads; ++b)
            for (int c = 0; c < 3; ++c)
                out[I3PN(c, b, i, nbeads)] =
                    orc_philox_normal(seed, traj, step, stream, (uint32_t)(i * nbeads + b), (uint32_t)c);
}

/* md_tian2.f90:78-98 without output: nsteps of propagate_1 / calc_force /
 * propagate_2 with Philox noise */
/* the step's noise for orc_md_step (counter scheme of the GPU path) */
static void fill_step_noise(const orc_universe *u, uint64_t seed, uint32_t traj, uint32_t st, int lan, int pil,
                            int and_, double *randy, double *zeta, double *and_z, double *and_u) {
    const int P = u->nbeads, N = u->natoms;
    if (lan) orc_philox_fill_normals(seed, traj, st, ORC_STREAM_LANGEVIN, N, P, randy);
    if (pil) orc_philox_fill_normals(seed, traj, st, ORC_STREAM_PILE, N, P, zeta);
    if (and_) {
        orc_philox_fill_normals(seed, traj, st, ORC_STREAM_ANDERSEN_V, N, P, and_z);
        for (int i = 0; i < N; ++i)
            for (int b = 0; b < P; ++b)
                and_u[IPN(b, i, P)] = orc_philox_uniform(seed, traj, st, ORC_STREAM_ANDERSEN_U,
                                                         (uint32_t)(i * P + b), 0u);
    }
}

int orc_run_trajectory(orc_universe *u, const orc_emt_params *p,
                       const orc_simparams *sp, uint64_t seed, uint32_t traj,
                       int step0, int nsteps) {
    const int P = u->nbeads, N = u->natoms;
    const size_t n = (size_t)3 * P * N;
    double *randy = ALLOC0(n), *zeta = ALLOC0(n), *and_z = ALLOC0(n), *and_u = ALLOC0(P * N);
    double *cjk = ALLOC0(P * P);
    orc_build_cjk(P, cjk);
    const int lan = any_algo(u, ORC_PROP_LANGEVIN), pil = any_algo(u, ORC_PROP_PILE),
              and_ = any_algo(u, ORC_PROP_ANDERSEN);
    int rc = 0;
    for (int s = 0; s < nsteps && !rc; ++s) {
        fill_step_noise(u, seed, traj, (uint32_t)(step0 + s), lan, pil, and_, randy, zeta, and_z, and_u);
        rc = orc_md_step(u, p, sp, cjk, randy, zeta, and_z, and_u);
    }
    free(randy); free(zeta); free(and_z); free(and_u); free(cjk);
    return rc;
}

/* calc_centroid_positions (rpmd.f90:81-103): bead sum / nbeads, bead order */
static void centroid(const orc_universe *u, int i, double c[3]) {
    const int P = u->nbeads;
    for (int k = 0; k < 3; ++k) {
        double a = 0.0;
        for (int b = 0; b < P; ++b) a += u->r[I3PN(k, b, i, P)];
        c[k] = P > 1 ? a / P : a;
    }
}

/* collect_trajectory_characteristics (trajectory_info.f90:47-76) with
 * check_for_lowest_position (:127-137), check_for_closest_approach
 * (:81-122), check_for_proj_surf_interaction (:142-150) and the turning
 * points of check_for_bounce (:248-255).  The bounce statistics are not
 * restated: output_scatter does not write them (output_mod.f90:699). */
void orc_collect_characteristics(const orc_universe *u, const orc_simparams *sp, double ads_start,
                                 double ads_end, int istep, orc_traj_obs *o) {
    const int P = u->nbeads, N = u->natoms;
    if (istep == 1) {
        o->closest_approach = ORC_DEFAULT_REAL;
        o->closest_approach_time = 0.0;
        for (int k = 0; k < 3; ++k) o->lowest_z[k] = ORC_DEFAULT_REAL;
        o->interaction_time = 0.0;
        o->nturning_points = 0;
        o->was_moving_down = 1;
    }
    double c1[3], ci[3], pp[3];
    centroid(u, 0, c1);
    /* check_for_lowest_position: sum(r(:,:,1), dim=2)/nbeads */
    for (int k = 0; k < 3; ++k) {
        double a = 0.0;
        for (int b = 0; b < P; ++b) a += u->r[I3PN(k, b, 0, P)];
        pp[k] = a / P;
    }
    if (pp[2] < o->lowest_z[2])
        for (int k = 0; k < 3; ++k) o->lowest_z[k] = pp[k];
    /* check_for_closest_approach */
    double here = ORC_DEFAULT_REAL;
    for (int i = 1; i < N; ++i) {
        centroid(u, i, ci);
        double d[3], s[3], e[3];
        for (int k = 0; k < 3; ++k) d[k] = c1[k] - ci[k];
        for (int k = 0; k < 3; ++k)
            s[k] = u->isimbox[k + 0] * d[0] + u->isimbox[k + 3] * d[1] + u->isimbox[k + 6] * d[2];
        for (int k = 0; k < 3; ++k) s[k] = s[k] - round(s[k]);
        for (int k = 0; k < 3; ++k)
            e[k] = u->simbox[k + 0] * s[0] + u->simbox[k + 3] * s[1] + u->simbox[k + 6] * s[2];
        const double r = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        if (r < o->closest_approach) {
            o->closest_approach = r;
            o->closest_approach_time = (double)istep;
        }
        if (r < here) here = r;
    }
    o->closest_approach_time = o->closest_approach_time * sp->step; /* every call (:111) */
    if (!o->is_adsorbed && here < ads_start) o->is_adsorbed = 1;
    if (o->is_adsorbed && here > ads_end) o->is_adsorbed = 0;
    /* check_for_proj_surf_interaction: any(f(:,:,1) /= 0) */
    int fnz = 0;
    for (int q = 0; q < 3 * P; ++q)
        if (u->f[q] != 0.0) fnz = 1;
    if (fnz) o->interaction_time = o->interaction_time + sp->step;
    /* turning points: sum(v(3,:,1))/nbeads < 0 */
    double vz = 0.0;
    for (int b = 0; b < P; ++b) vz += u->v[I3PN(2, b, 0, P)];
    const int down = vz / P < 0.0;
    if (down != o->was_moving_down) {
        o->nturning_points += 1;
        o->was_moving_down = down;
    }
}

/* One trajectory of the md loop with its exits (md_tian2.f90:75-101):
 * propagate_1, calc_force, propagate_2; with scatter output, leave when
 * every projectile bead is above proj_ul or below proj_ll; collect the
 * characteristics; after the loop collect once more with the final istep
 * (the exit step, or nsteps + 1).  *istep_end receives that istep. */
int orc_run_scatter(orc_universe *u, const orc_emt_params *p, const orc_simparams *sp, uint64_t seed,
                    uint32_t traj, int nsteps, double proj_ul, double proj_ll, int scatter, double ads_start,
                    double ads_end, orc_traj_obs *o, int *istep_end) {
    const int P = u->nbeads, N = u->natoms;
    const size_t n = (size_t)3 * P * N;
    double *randy = ALLOC0(n), *zeta = ALLOC0(n), *and_z = ALLOC0(n), *and_u = ALLOC0(P * N);
    double *cjk = ALLOC0(P * P);
    orc_build_cjk(P, cjk);
    const int lan = any_algo(u, ORC_PROP_LANGEVIN), pil = any_algo(u, ORC_PROP_PILE),
              and_ = any_algo(u, ORC_PROP_ANDERSEN);
    const int proj1 = u->is_proj[u->idx[0]];
    int rc = 0, istep = 1;
    for (; istep <= nsteps; ++istep) {
        fill_step_noise(u, seed, traj, (uint32_t)istep, lan, pil, and_, randy, zeta, and_z, and_u);
        rc = orc_md_step(u, p, sp, cjk, randy, zeta, and_z, and_u);
        if (rc) break;
        int above = 1, below = 1;
        for (int b = 0; b < P; ++b) {
            if (!(u->r[I3PN(2, b, 0, P)] > proj_ul)) above = 0;
            if (!(u->r[I3PN(2, b, 0, P)] < proj_ll)) below = 0;
        }
        if (scatter && proj1 && (above || below)) break;
        orc_collect_characteristics(u, sp, ads_start, ads_end, istep, o);
    }
    if (!rc) orc_collect_characteristics(u, sp, ads_start, ads_end, istep, o);
    *istep_end = istep;
    free(randy); free(zeta); free(and_z); free(and_u); free(cjk);
    return rc;
}

/* ---- estimators of output_nrg (output_mod.f90:170-229) ---- */

/* set_atomic_dofs (md_init.f90:882-887) */
static int dof_count(const orc_universe *u) {
    const int n = 3 * u->nbeads * u->natoms;
    int c = 0;
    for (int q = 0; q < n; ++q) c += u->is_fixed[q] != 0;
    return n - c;
}

/* calc_atom_temperature (universe_mod.f90:314-327) */
double orc_atom_temperature(const orc_universe *u) {
    const int P = u->nbeads;
    double temp = 0.0;
    for (int i = 0; i < u->natoms; ++i) {
        double s2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            double a = 0.0;
            for (int b = 0; b < P; ++b) a += u->v[I3PN(k, b, i, P)];
            s2 += a * a;
        }
        temp = temp + u->m[i] * s2;
    }
    return temp / KB / dof_count(u) / P;
}

/* calc_bead_temperature (rpmd.f90:63-77) */
double orc_bead_temperature(const orc_universe *u) {
    const int P = u->nbeads;
    double temp = 0.0;
    for (int i = 0; i < u->natoms; ++i) {
        double s2 = 0.0;
        for (int q = 0; q < 3 * P; ++q) s2 += u->v[3 * P * i + q] * u->v[3 * P * i + q];
        temp = temp + u->m[i] * s2;
    }
    return temp / KB / dof_count(u);
}

/* radius_of_gyration (rpmd.f90:511-545) */
void orc_radius_of_gyration(const orc_universe *u, double *rgyr_p, double *rgyr_l) {
    const int P = u->nbeads;
    *rgyr_p = 0.0;
    *rgyr_l = 0.0;
    if (P == 1) return;
    int nprojs = 0;
    for (int i = 0; i < u->natoms; ++i) {
        double c[3];
        centroid(u, i, c);
        double acc = 0.0;
        for (int b = 0; b < P; ++b)
            for (int k = 0; k < 3; ++k) {
                const double d = u->r[I3PN(k, b, i, P)] - c[k];
                acc += d * d;
            }
        if (u->is_proj[u->idx[i]]) {
            nprojs += 1;
            *rgyr_p += acc;
        } else {
            *rgyr_l += acc;
        }
    }
    *rgyr_p = *rgyr_p / P / (nprojs > 1 ? nprojs : 1);
    const int nl = u->natoms - nprojs;
    *rgyr_l = *rgyr_l / P / (nl > 1 ? nl : 1);
}

/* atom_ekin (universe_mod.f90:433-458) */
void orc_atom_ekin(const orc_universe *u, double *ekin_p, double *ekin_l) {
    const int P = u->nbeads;
    *ekin_p = 0.0;
    *ekin_l = 0.0;
    for (int i = 0; i < u->natoms; ++i) {
        double s2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            double a = 0.0;
            for (int b = 0; b < P; ++b) a += u->v[I3PN(k, b, i, P)];
            s2 += a * a;
        }
        const double nrg = u->m[i] * s2;
        if (u->is_proj[u->idx[i]]) *ekin_p += nrg;
        else *ekin_l += nrg;
    }
    *ekin_p = 0.5 * *ekin_p / P / P;
    *ekin_l = 0.5 * *ekin_l / P / P;
}

/* bead_ekin (rpmd.f90:547-572) */
void orc_bead_ekin(const orc_universe *u, double *ekin_p, double *ekin_l) {
    const int P = u->nbeads;
    *ekin_p = 0.0;
    *ekin_l = 0.0;
    for (int i = 0; i < u->natoms; ++i) {
        double s2 = 0.0;
        for (int q = 0; q < 3 * P; ++q) s2 += u->v[3 * P * i + q] * u->v[3 * P * i + q];
        const double nrg = u->m[i] * s2;
        if (u->is_proj[u->idx[i]]) *ekin_p += nrg;
        else *ekin_l += nrg;
    }
    *ekin_p = 0.5 * *ekin_p / P;
    *ekin_l = 0.5 * *ekin_l / P;
}

/* virial_quantum_ekin (rpmd.f90:273-301) */
void orc_virial_quantum_ekin(const orc_universe *u, double *ekin_p, double *ekin_l) {
    const int P = u->nbeads;
    *ekin_p = 0.0;
    *ekin_l = 0.0;
    if (P == 1) return;
    for (int i = 0; i < u->natoms; ++i) {
        double c[3];
        centroid(u, i, c);
        for (int b = 0; b < P; ++b) {
            double d = 0.0;
            for (int k = 0; k < 3; ++k) d += (c[k] - u->r[I3PN(k, b, i, P)]) * u->f[I3PN(k, b, i, P)];
            if (u->is_proj[u->idx[i]]) *ekin_p += d;
            else *ekin_l += d;
        }
    }
    *ekin_p = 0.5 * *ekin_p / P;
    *ekin_l = 0.5 * *ekin_l / P;
}

/* calc_bead_epot (rpmd.f90:236-264) over calc_inter_bead_distances
 * (:212-233): plain differences r(b) - r(b+1), no minimum image */
double orc_bead_epot(const orc_universe *u, const orc_simparams *sp) {
    const int P = u->nbeads;
    double epot = 0.0;
    for (int i = 0; i < u->natoms; ++i) {
        double dx2 = 0.0;
        if (P > 1)
            for (int b = 0; b < P; ++b) {
                const int k = (b + 1) % P;
                double d2 = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double d = u->r[I3PN(c, b, i, P)] - u->r[I3PN(c, k, i, P)];
                    d2 += d * d;
                }
                const double dist = sqrt(d2);
                dx2 += dist * dist;
            }
        const double T = u->is_proj[u->idx[i]] ? sp->Tproj : sp->Tsurf;
        const double w = KB * T * P / HBAR;
        epot = epot + u->m[i] * dx2 * (w * w);
    }
    return 0.5 * epot / P;
}

/* the row of output_nrg after the time column */
void orc_energy_row(const orc_universe *u, const orc_simparams *sp, double out[20]) {
    const int P = u->nbeads;
    out[0] = orc_atom_temperature(u);
    out[1] = orc_bead_temperature(u);
    orc_radius_of_gyration(u, &out[2], &out[3]);
    orc_atom_ekin(u, &out[4], &out[5]);
    orc_bead_ekin(u, &out[6], &out[7]);
    orc_virial_quantum_ekin(u, &out[8], &out[9]);
    double ep = 0.0;
    for (int b = 0; b < P; ++b) ep += u->epot[b];
    out[10] = ep / P;
    out[11] = orc_bead_epot(u, sp);
    out[12] = out[6] + out[7] + out[10] + out[11];
    double fs = 0.0;
    for (int q = 0; q < 3 * P * u->natoms; ++q) fs += u->f[q];
    out[13] = fs;
    for (int k = 0; k < 3; ++k) {
        double xa = 0.0, va = 0.0;
        for (int b = 0; b < P; ++b) {
            xa += u->r[I3PN(k, b, 0, P)];
            va += u->v[I3PN(k, b, 0, P)];
        }
        out[14 + k] = xa / P;
        out[17 + k] = va / P;
    }
}

/* ---- initial conditions with explicit uniforms ---- */

/* apply_pip (md_init.f90:498-554) on a projectile universe; U[k] is the
 * random_number drawn for a random component k */
void orc_apply_pip(orc_universe *u, const int random[3], const double value[3], const double U[3]) {
    const int P = u->nbeads, N = u->natoms;
    double *cents = (double *)malloc(sizeof(double) * 3 * N);
    for (int i = 0; i < N; ++i) centroid(u, i, cents + 3 * i);
    double target[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
        if (random[k]) {
            target[k] = U[k];
            const double cs = u->simbox[0 + 3 * k] + u->simbox[1 + 3 * k] + u->simbox[2 + 3 * k];
            for (int j = 0; j < 3; ++j) target[j] = target[j] * cs;  /* the whole vector */
        } else {
            for (int b = 0; b < P; ++b)
                for (int i = 0; i < N; ++i) u->r[I3PN(k, b, i, P)] += value[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        const double fp = u->simbox[k + 0] * target[0] + u->simbox[k + 3] * target[1] +
                          u->simbox[k + 6] * target[2];
        for (int b = 0; b < P; ++b)
            for (int i = 0; i < N; ++i) u->r[I3PN(k, b, i, P)] = u->r[I3PN(k, b, i, P)] - cents[k + 3 * i] + fp;
    }
    free(cents);
}

/* set_proj_incidence (md_init.f90:829-877) with normal_deviate
 * (useful_things.f90:121-131): rnd[0], rnd[1] for Einc; rnd[2] for a random
 * azimuth, or rnd[2], rnd[3] for a normal one */
void orc_set_proj_incidence(orc_universe *u, double einc, double sigma_einc, double polar_deg, int az_random,
                            double az_deg, double sigma_az, const double rnd[4]) {
    const int P = u->nbeads, N = u->natoms;
    const double deg2rad = ORC_PI / 180.0;
    const double drawn = sqrt(-2 * log(rnd[0])) * cos(2 * ORC_PI * rnd[1]) * sigma_einc + einc;
    double pm = 0.0;
    for (int i = 0; i < N; ++i)
        if (u->is_proj[u->idx[i]]) pm = pm + u->m[i];
    double az;
    if (az_random) {
        az = 2 * ORC_PI * rnd[2];
    } else {
        az = sqrt(-2 * log(rnd[2])) * cos(2 * ORC_PI * rnd[3]) * sigma_az + az_deg;
        az = az * deg2rad;
    }
    const double polar = polar_deg * deg2rad;
    const double vinc = sqrt(2 * drawn / pm);
    const double vx = vinc * sin(polar) * cos(az), vy = vinc * sin(polar) * sin(az), vz = -vinc * cos(polar);
    for (int i = 0; i < N; ++i) {
        if (!u->is_proj[u->idx[i]]) return;
        double vcm[3];
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (int b = 0; b < P; ++b) acc += u->v[I3PN(c, b, i, P)];
            vcm[c] = acc / P;
        }
        for (int b = 0; b < P; ++b)
            for (int c = 0; c < 3; ++c) u->v[I3PN(c, b, i, P)] = u->v[I3PN(c, b, i, P)] - vcm[c];
        for (int b = 0; b < P; ++b) {
            u->v[I3PN(0, b, i, P)] = u->v[I3PN(0, b, i, P)] + vx;
            u->v[I3PN(1, b, i, P)] = u->v[I3PN(1, b, i, P)] + vy;
            u->v[I3PN(2, b, i, P)] = u->v[I3PN(2, b, i, P)] + vz;
        }
    }
}
