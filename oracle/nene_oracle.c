/* nene_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU checker of the GPU NN
 * surrogate.  Serial FP64 restatement of md_tian2's NENE short-range
 * atomic network (RuNNer, nn_type_short 1), following:
 *   compute_nene             src/pes_nene_mod.f90:7145-7368 (units, per bead)
 *   getshortatomic           src/RuNNer_src/getshortatomic.f90:23-370
 *                            (extrapolation test :150-168, scaling :220-224,
 *                            derivative scaling :177-189, forces :294-322)
 *   atomsymfunction2         src/RuNNer_src/atomsymfunction2.f90 (value :110,
 *                            derivative :114)
 *   atomsymfunction3Andi     src/RuNNer_src/atomsymfunction3Andi.f90:24-260
 *                            (2^(1-zeta) norm :96, (1+lambda cos) <= 0 -> 0
 *                            :196-200)
 *   getcutoff (type 1)       src/RuNNer_src/getcutoff.f90:55-74
 *   calconenn                src/RuNNer_src/calconenn.f90:46-158
 *   getdnodes_values         src/RuNNer_src/getdnodes_values.f90:40-117
 * The training gradient (orc_nene_loss_grad) has no reference counterpart
 * (md_tian2 fits with RuNNer's Kalman filter in fit.f90, out of scope): it
 * is the checker of the GPU training kernels' analytic gradients and is
 * itself pinned by finite differences of the loss.
 */
#include "nene_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define HA2EV 27.21138602           /* src/constants.f90:165 */
#define BOHR2ANG 0.529177211        /* src/constants.f90:160 */
#define HABOHR2EVANG (HA2EV / BOHR2ANG)
#define EXTRAP_THRESHOLD 0.0001     /* getshortatomic.f90 threshold */

typedef struct {
    int j;
    double d[3]; /* r_j + T - r_i */
    double r;
} NbrE;

typedef struct {
    NbrE *v;
    int n, cap;
} NbrList;

static void push(NbrList *L, int j, const double *d, double r) {
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 64;
        L->v = (NbrE *)realloc(L->v, sizeof(NbrE) * L->cap);
    }
    NbrE *e = &L->v[L->n++];
    e->j = j;
    e->d[0] = d[0];
    e->d[1] = d[1];
    e->d[2] = d[2];
    e->r = r;
}

/* getcutoff type 1: fc = 0.5 (cos(pi x) + 1), x = (r - rci)/(rc - rci) */
static void fcut(double alpha, double rc, double r, double *fc, double *dfc) {
    const double rci = alpha * rc;
    if (r <= rci) {
        *fc = 1.0;
        *dfc = 0.0;
        return;
    }
    const double iw = 1.0 / (rc - rci);
    const double x = (r - rci) * iw;
    *fc = 0.5 * (cos(M_PI * x) + 1.0);
    *dfc = -sin(M_PI * x) * (M_PI / 2.0) * iw;
}

static int inv3(const double *a, double *ai) {
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (fabs(det) < 1e-300) return -1;
    const double id = 1.0 / det;
    ai[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    ai[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    ai[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    ai[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    ai[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    ai[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    ai[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    ai[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    ai[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return 0;
}

/* every periodic image (j, T) within rcmax of atom i, (j, T) != (i, 0);
 * lat rows are the lattice vectors; positions need not be folded */
static int neighbours(int nat, const double *xyz, const double *lat, int i, double rcmax, NbrList *L) {
    int M[3] = {0, 0, 0};
    if (lat) {
        double li[9];
        if (inv3(lat, li)) return -3;
        for (int k = 0; k < 3; ++k) {
            /* |column k of lat^-1| = 1 / (spacing of lattice planes k) */
            const double bk = sqrt(li[k] * li[k] + li[3 + k] * li[3 + k] + li[6 + k] * li[6 + k]);
            M[k] = (int)ceil(rcmax * bk) + 1;
        }
    }
    L->n = 0;
    for (int j = 0; j < nat; ++j)
        for (int n0 = -M[0]; n0 <= M[0]; ++n0)
            for (int n1 = -M[1]; n1 <= M[1]; ++n1)
                for (int n2 = -M[2]; n2 <= M[2]; ++n2) {
                    if (j == i && n0 == 0 && n1 == 0 && n2 == 0) continue;
                    double d[3];
                    for (int c = 0; c < 3; ++c) {
                        d[c] = xyz[3 * j + c] - xyz[3 * i + c];
                        if (lat) d[c] += n0 * lat[c] + n1 * lat[3 + c] + n2 * lat[6 + c];
                    }
                    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    if (r <= rcmax) push(L, j, d, r);
                }
    return 0;
}

long orc_nene_nweights(const NeneModel *m, int e) {
    long n = 0;
    for (int l = 1; l <= m->nlayers; ++l) n += (long)m->nodes[e][l - 1] * m->nodes[e][l] + m->nodes[e][l];
    return n;
}

int orc_nene_descriptors(const NeneModel *m, int nat, const int *elem, const double *xyz, const double *lat,
                         double *G, double *dG) {
    if (m->cutoff_type != 1) return -2;
    double rcmax = 0.0;
    for (int e = 0; e < m->nelem; ++e)
        for (int k = 0; k < m->nsf[e]; ++k) rcmax = m->rc[e][k] > rcmax ? m->rc[e][k] : rcmax;
    memset(G, 0, sizeof(double) * nat * NENE_MAXSF);
    if (dG) memset(dG, 0, sizeof(double) * (size_t)nat * NENE_MAXSF * nat * 3);
    NbrList L = {0, 0, 0};
    for (int i = 0; i < nat; ++i) {
        int rc = neighbours(nat, xyz, lat, i, rcmax, &L);
        if (rc) {
            free(L.v);
            return rc;
        }
        const int ei = elem[i];
        for (int k = 0; k < m->nsf[ei]; ++k) {
            const double rck = m->rc[ei][k], eta = m->eta[ei][k];
            double *gk = dG ? dG + ((size_t)i * NENE_MAXSF + k) * nat * 3 : NULL;
            double g = 0.0;
            if (m->type[ei][k] == 2) {
                const double rs = m->rs[ei][k];
                for (int p = 0; p < L.n; ++p) {
                    const NbrE *nb = &L.v[p];
                    if (elem[nb->j] != m->e1[ei][k] || nb->r > rck) continue;
                    double fc, dfc;
                    fcut(m->cutoff_alpha, rck, nb->r, &fc, &dfc);
                    const double ex = exp(-eta * (nb->r - rs) * (nb->r - rs));
                    g += ex * fc;
                    if (gk) {
                        const double dr = -2.0 * eta * (nb->r - rs) * ex * fc + ex * dfc; /* d/dr */
                        for (int c = 0; c < 3; ++c) {
                            const double u = dr * nb->d[c] / nb->r;
                            gk[3 * nb->j + c] += u;
                            gk[3 * i + c] -= u;
                        }
                    }
                }
            } else if (m->type[ei][k] == 3) {
                const double lam = m->lambda[ei][k], zeta = m->zeta[ei][k];
                const int a1 = m->e1[ei][k], a2 = m->e2[ei][k];
                for (int p = 0; p < L.n; ++p) {
                    const NbrE *P = &L.v[p];
                    const int ep = elem[P->j];
                    if (P->r > rck || (ep != a1 && ep != a2)) continue;
                    for (int q = p + 1; q < L.n; ++q) {
                        const NbrE *Q = &L.v[q];
                        const int eq = elem[Q->j];
                        if (!((ep == a1 && eq == a2) || (ep == a2 && eq == a1))) continue;
                        if (Q->r > rck) continue;
                        double c3[3];
                        for (int c = 0; c < 3; ++c) c3[c] = Q->d[c] - P->d[c]; /* r_k - r_j */
                        const double rjk2 = c3[0] * c3[0] + c3[1] * c3[1] + c3[2] * c3[2];
                        if (rjk2 > rck * rck) continue;
                        const double rjk = sqrt(rjk2);
                        const double rij = P->r, rik = Q->r;
                        const double cth = (P->d[0] * Q->d[0] + P->d[1] * Q->d[1] + P->d[2] * Q->d[2]) / (rij * rik);
                        const double pl = 1.0 + lam * cth;
                        if (pl <= 0.0) continue;
                        double fij, dfij, fik, dfik, fjk, dfjk;
                        fcut(m->cutoff_alpha, rck, rij, &fij, &dfij);
                        fcut(m->cutoff_alpha, rck, rik, &fik, &dfik);
                        fcut(m->cutoff_alpha, rck, rjk, &fjk, &dfjk);
                        const double A = pow(pl, zeta);
                        const double B = exp(-eta * (rij * rij + rik * rik + rjk2));
                        const double C = fij * fik * fjk;
                        g += A * B * C;
                        if (gk) {
                            /* d/da, d/db, d/dc with a = r_j - r_i, b = r_k - r_i, c = r_k - r_j */
                            const double dA = zeta * lam * pow(pl, zeta - 1.0); /* dA/dcos */
                            double ga[3], gb[3], gc[3];
                            for (int c = 0; c < 3; ++c) {
                                const double dca = Q->d[c] / (rij * rik) - cth * P->d[c] / (rij * rij);
                                const double dcb = P->d[c] / (rij * rik) - cth * Q->d[c] / (rik * rik);
                                ga[c] = dA * dca * B * C + A * B * (-2.0 * eta * P->d[c]) * C +
                                        A * B * dfij * (P->d[c] / rij) * fik * fjk;
                                gb[c] = dA * dcb * B * C + A * B * (-2.0 * eta * Q->d[c]) * C +
                                        A * B * fij * dfik * (Q->d[c] / rik) * fjk;
                                gc[c] = A * B * (-2.0 * eta * c3[c]) * C + A * B * fij * fik * dfjk * (c3[c] / rjk);
                            }
                            const double nrm = pow(2.0, 1.0 - zeta);
                            for (int c = 0; c < 3; ++c) {
                                gk[3 * i + c] -= nrm * (ga[c] + gb[c]);
                                gk[3 * P->j + c] += nrm * (ga[c] - gc[c]);
                                gk[3 * Q->j + c] += nrm * (gb[c] + gc[c]);
                            }
                        }
                    }
                }
                g *= pow(2.0, 1.0 - zeta);
            } else {
                free(L.v);
                return -2;
            }
            G[(size_t)i * NENE_MAXSF + k] = g;
        }
    }
    free(L.v);
    return 0;
}

/* ---- per-element network (calconenn, getdnodes_values) ---- */
static void act(char a, double z, double *h, double *d1, double *d2) {
    if (a == 't') {
        *h = tanh(z);
        *d1 = 1.0 - *h * *h;
        *d2 = -2.0 * *h * *d1;
    } else if (a == 's') {
        *h = 1.0 / (1.0 + exp(-z));
        *d1 = *h * (1.0 - *h);
        *d2 = *d1 * (1.0 - 2.0 * *h);
    } else { /* 'l' */
        *h = z;
        *d1 = 1.0;
        *d2 = 0.0;
    }
}

typedef struct {
    double z[NENE_MAXL + 1][NENE_MAXNODES], h[NENE_MAXL + 1][NENE_MAXNODES];
    double a1[NENE_MAXL + 1][NENE_MAXNODES], a2[NENE_MAXL + 1][NENE_MAXNODES];
    double delta[NENE_MAXL + 1][NENE_MAXNODES]; /* dE/dz */
} Net;

static const double *layer_w(const NeneModel *m, int e, int l) {
    const double *w = m->w + m->woff[e];
    for (int k = 1; k < l; ++k) w += (long)m->nodes[e][k - 1] * m->nodes[e][k] + m->nodes[e][k];
    return w;
}

static double net_forward(const NeneModel *m, int e, const double *x, Net *n) {
    for (int i = 0; i < m->nodes[e][0]; ++i) n->h[0][i] = x[i];
    for (int l = 1; l <= m->nlayers; ++l) {
        const int nin = m->nodes[e][l - 1], nout = m->nodes[e][l];
        const double *W = layer_w(m, e, l), *b = W + (long)nin * nout;
        for (int o = 0; o < nout; ++o) {
            double s = 0.0;
            for (int i = 0; i < nin; ++i) s += W[(long)i * nout + o] * n->h[l - 1][i];
            s += b[o];
            n->z[l][o] = s;
            act(m->act[e][l - 1], s, &n->h[l][o], &n->a1[l][o], &n->a2[l][o]);
        }
    }
    return n->h[m->nlayers][0];
}

/* dE/dx (x = scaled symmetry functions) and delta = dE/dz per layer */
static void net_backward(const NeneModel *m, int e, Net *n, double *dEdx) {
    const int L = m->nlayers;
    n->delta[L][0] = n->a1[L][0];
    for (int l = L; l >= 1; --l) {
        const int nin = m->nodes[e][l - 1], nout = m->nodes[e][l];
        const double *W = layer_w(m, e, l);
        for (int i = 0; i < nin; ++i) {
            double s = 0.0;
            for (int o = 0; o < nout; ++o) s += W[(long)i * nout + o] * n->delta[l][o];
            if (l > 1)
                n->delta[l - 1][i] = n->a1[l - 1][i] * s;
            else
                dEdx[i] = s;
        }
    }
}

/* gradient accumulation: g += cE dE/dW + d(u . dE/dx)/dW */
static void net_weight_grad(const NeneModel *m, int e, const Net *n, double cE, const double *u, double *g) {
    const int L = m->nlayers;
    double zd[NENE_MAXL + 1][NENE_MAXNODES], hd[NENE_MAXL + 1][NENE_MAXNODES];
    double rho[NENE_MAXL + 1][NENE_MAXNODES], pi[NENE_MAXL + 1][NENE_MAXNODES];
    memset(zd, 0, sizeof zd);
    /* tangent pass along input direction u */
    for (int i = 0; i < m->nodes[e][0]; ++i) hd[0][i] = u ? u[i] : 0.0;
    for (int l = 1; l <= L; ++l) {
        const int nin = m->nodes[e][l - 1], nout = m->nodes[e][l];
        const double *W = layer_w(m, e, l);
        for (int o = 0; o < nout; ++o) {
            double s = 0.0;
            for (int i = 0; i < nin; ++i) s += W[(long)i * nout + o] * hd[l - 1][i];
            zd[l][o] = s;
            hd[l][o] = n->a1[l][o] * s;
        }
    }
    rho[L][0] = n->a1[L][0];
    pi[L][0] = n->a2[L][0] * zd[L][0];
    double *gw = g + m->woff[e];
    long off[NENE_MAXL + 1];
    off[1] = 0;
    for (int l = 2; l <= L; ++l) off[l] = off[l - 1] + (long)m->nodes[e][l - 2] * m->nodes[e][l - 1] + m->nodes[e][l - 1];
    for (int l = L; l >= 1; --l) {
        const int nin = m->nodes[e][l - 1], nout = m->nodes[e][l];
        const double *W = layer_w(m, e, l);
        double *gW = gw + off[l], *gb = gW + (long)nin * nout;
        for (int i = 0; i < nin; ++i)
            for (int o = 0; o < nout; ++o)
                gW[(long)i * nout + o] +=
                    cE * n->h[l - 1][i] * n->delta[l][o] + n->h[l - 1][i] * pi[l][o] + hd[l - 1][i] * rho[l][o];
        for (int o = 0; o < nout; ++o) gb[o] += cE * n->delta[l][o] + pi[l][o];
        if (l > 1)
            for (int i = 0; i < nin; ++i) {
                double t = 0.0, p = 0.0;
                for (int o = 0; o < nout; ++o) {
                    t += W[(long)i * nout + o] * rho[l][o];
                    p += W[(long)i * nout + o] * pi[l][o];
                }
                rho[l - 1][i] = n->a1[l - 1][i] * t;
                pi[l - 1][i] = n->a1[l - 1][i] * p + n->a2[l - 1][i] * zd[l - 1][i] * t;
            }
    }
}

/* scaled symmetry functions (getshortatomic.f90:220-224) and the factor on
 * their derivatives (:177-189: applied whenever scaling is on) */
static int scale_sf(const NeneModel *m, int e, const double *g, double *x, double *fac, int *nextrap) {
    for (int k = 0; k < m->nsf[e]; ++k) {
        const double mn = m->gmin[e][k], mx = m->gmax[e][k];
        if ((g[k] - mx) > EXTRAP_THRESHOLD || (-g[k] + mn) > EXTRAP_THRESHOLD) ++*nextrap;
        if (m->scale && !(mx > mn)) return -4;
        if (m->center && !m->scale)
            x[k] = g[k] - m->gavg[e][k];
        else if (m->scale && !m->center)
            x[k] = (g[k] - mn) / (mx - mn) * (m->scmax - m->scmin) + m->scmin;
        else if (m->scale && m->center)
            x[k] = (g[k] - m->gavg[e][k]) / (mx - mn);
        else
            x[k] = g[k];
        fac[k] = m->scale ? (m->scmax - m->scmin) / (mx - mn) : 1.0;
    }
    return 0;
}

int orc_nene_energy_forces(const NeneModel *m, int nat, const int *elem, const double *xyz, const double *lat,
                           double *E, double *Eatom, double *F, int *nextrap) {
    double *G = (double *)malloc(sizeof(double) * nat * NENE_MAXSF);
    double *dG = (double *)malloc(sizeof(double) * (size_t)nat * NENE_MAXSF * nat * 3);
    int rc = orc_nene_descriptors(m, nat, elem, xyz, lat, G, dG);
    *E = 0.0;
    *nextrap = 0;
    memset(F, 0, sizeof(double) * nat * 3);
    Net net;
    for (int i = 0; i < nat && !rc; ++i) {
        const int e = elem[i];
        double x[NENE_MAXSF], fac[NENE_MAXSF], dx[NENE_MAXSF];
        if ((rc = scale_sf(m, e, G + (size_t)i * NENE_MAXSF, x, fac, nextrap))) break;
        const double ei = net_forward(m, e, x, &net);
        net_backward(m, e, &net, dx);
        if (Eatom) Eatom[i] = ei;
        *E += ei;
        for (int k = 0; k < m->nsf[e]; ++k) {
            const double c = dx[k] * fac[k];
            const double *gk = dG + ((size_t)i * NENE_MAXSF + k) * nat * 3;
            for (int a = 0; a < 3 * nat; ++a) F[a] -= c * gk[a];
        }
    }
    free(G);
    free(dG);
    return rc;
}

int orc_nene_loss_grad(const NeneModel *m, int nframes, const int *nat, const long *off, const int *elem,
                       const double *xyz, const double *lat, const double *Eref, const double *Fref, double wE,
                       double wF, double *L, double *grad) {
    long nw = 0, ntot = 0;
    for (int e = 0; e < m->nelem; ++e) nw = m->woff[e] + orc_nene_nweights(m, e) > nw ? m->woff[e] + orc_nene_nweights(m, e) : nw;
    for (int f = 0; f < nframes; ++f) ntot += nat[f];
    memset(grad, 0, sizeof(double) * nw);
    *L = 0.0;
    int rc = 0;
    for (int f = 0; f < nframes && !rc; ++f) {
        const int n = nat[f];
        const int *el = elem + off[f];
        const double *x0 = xyz + 3 * off[f];
        const double *lt = lat ? lat + 9 * f : NULL;
        double *G = (double *)malloc(sizeof(double) * n * NENE_MAXSF);
        double *dG = (double *)malloc(sizeof(double) * (size_t)n * NENE_MAXSF * n * 3);
        double *X = (double *)malloc(sizeof(double) * n * NENE_MAXSF);
        double *FAC = (double *)malloc(sizeof(double) * n * NENE_MAXSF);
        double *F = (double *)calloc((size_t)n * 3, sizeof(double));
        Net *nets = (Net *)malloc(sizeof(Net) * n);
        int nex = 0;
        rc = orc_nene_descriptors(m, n, el, x0, lt, G, dG);
        double E = 0.0;
        for (int i = 0; i < n && !rc; ++i) {
            const int e = el[i];
            double dx[NENE_MAXSF];
            double *xi = X + (size_t)i * NENE_MAXSF, *fi = FAC + (size_t)i * NENE_MAXSF;
            if ((rc = scale_sf(m, e, G + (size_t)i * NENE_MAXSF, xi, fi, &nex))) break;
            E += net_forward(m, e, xi, &nets[i]);
            net_backward(m, e, &nets[i], dx);
            for (int k = 0; k < m->nsf[e]; ++k) {
                const double c = dx[k] * fi[k];
                const double *gk = dG + ((size_t)i * NENE_MAXSF + k) * n * 3;
                for (int a = 0; a < 3 * n; ++a) F[a] -= c * gk[a];
            }
        }
        if (!rc) {
            const double dE = (HA2EV * E - Eref[f]) / n;
            *L += wE / nframes * dE * dE;
            const double dLdE = 2.0 * wE / nframes * dE / n * HA2EV; /* per Hartree */
            double *dLdF = (double *)malloc(sizeof(double) * n * 3);
            for (int a = 0; a < 3 * n; ++a) {
                const double r = HABOHR2EVANG * F[a] - Fref[3 * off[f] + a];
                *L += wF / (3.0 * ntot) * r * r;
                dLdF[a] = 2.0 * wF / (3.0 * ntot) * r * HABOHR2EVANG; /* per Hartree/Bohr */
            }
            for (int i = 0; i < n; ++i) {
                const int e = el[i];
                double u[NENE_MAXSF];
                for (int k = 0; k < m->nsf[e]; ++k) {
                    const double *gk = dG + ((size_t)i * NENE_MAXSF + k) * n * 3;
                    double s = 0.0;
                    for (int a = 0; a < 3 * n; ++a) s += dLdF[a] * gk[a];
                    u[k] = -FAC[(size_t)i * NENE_MAXSF + k] * s;
                }
                net_weight_grad(m, e, &nets[i], dLdE, u, grad);
            }
            free(dLdF);
        }
        free(G);
        free(dG);
        free(X);
        free(FAC);
        free(F);
        free(nets);
    }
    return rc;
}
