/* nene_oracle.h -- TEST INFRASTRUCTURE ONLY (the checker of the GPU NN
 * surrogate; never linked into the product library).
 *
 * Serial FP64 restatement of md_tian2's NENE short-range atomic neural
 * network potential (RuNNer nn_type_short 1) as compute_nene evaluates it
 * (src/pes_nene_mod.f90:7145-7368): radial (type 2) and angular (type 3,
 * "Andi" form) symmetry functions with the cosine cut-off, min/max
 * scaling, one feed-forward network per element, atomic energies summed,
 * forces from the chain rule through the symmetry functions.  Units are the
 * network's own: Bohr and Hartree.
 */
#ifndef NENE_ORACLE_H
#define NENE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define NENE_MAXE 4
#define NENE_MAXSF 64
#define NENE_MAXL 4
#define NENE_MAXNODES 64

typedef struct {
    int nelem;                          /* elements, ascending atomic number */
    int z[NENE_MAXE];                   /* atomic numbers */
    int nsf[NENE_MAXE];                 /* symmetry functions per element */
    int type[NENE_MAXE][NENE_MAXSF];    /* 2 radial, 3 angular */
    int e1[NENE_MAXE][NENE_MAXSF];      /* neighbour element index (0-based) */
    int e2[NENE_MAXE][NENE_MAXSF];      /* second neighbour element (type 3), e1 <= e2 */
    double eta[NENE_MAXE][NENE_MAXSF], rs[NENE_MAXE][NENE_MAXSF];
    double lambda[NENE_MAXE][NENE_MAXSF], zeta[NENE_MAXE][NENE_MAXSF];
    double rc[NENE_MAXE][NENE_MAXSF];   /* Bohr */
    double gmin[NENE_MAXE][NENE_MAXSF], gmax[NENE_MAXE][NENE_MAXSF], gavg[NENE_MAXE][NENE_MAXSF];
    double scmin, scmax;                /* scale_min/max_short_atomic */
    int scale, center;                  /* scale_symmetry_functions, center_symmetry_functions */
    int cutoff_type;                    /* only 1 (cosine) */
    double cutoff_alpha;                /* inner cut-off fraction */
    int nlayers;                        /* weight layers: hidden + output */
    int nodes[NENE_MAXE][NENE_MAXL + 1];/* nodes[e][0] = nsf[e] */
    char act[NENE_MAXE][NENE_MAXL];     /* 't' tanh, 'l' linear, 's' sigmoid */
    long woff[NENE_MAXE];               /* offset of element e's weights in w */
    const double *w;                    /* RuNNer order: per layer W(in, out) input-major, then biases */
} NeneModel;

/* Number of weights of element e. */
long orc_nene_nweights(const NeneModel *m, int e);

/* Raw (unscaled) symmetry functions G[i*NENE_MAXSF + k] and, when dG is
 * not NULL, their derivatives dG[((i*NENE_MAXSF + k)*nat + a)*3 + c] with
 * respect to atom a's coordinate c.  xyz in Bohr, lat the three lattice
 * vectors as rows (Bohr) or NULL for a cluster.  Returns 0 or a negative
 * error code. */
int orc_nene_descriptors(const NeneModel *m, int nat, const int *elem, const double *xyz, const double *lat,
                         double *G, double *dG);

/* Energy (Hartree), atomic energies, forces (Hartree/Bohr, nat*3) and the
 * number of symmetry-function values outside [min - 1e-4, max + 1e-4]. */
int orc_nene_energy_forces(const NeneModel *m, int nat, const int *elem, const double *xyz, const double *lat,
                           double *E, double *Eatom, double *F, int *nextrap);

/* Training loss over a set of frames and its gradient with respect to all
 * weights (same layout as m->w):
 *   L = wE/Nf sum_f ((E_f - Eref_f)/N_f)^2 + wF/(3 sum_f N_f) sum |F - Fref|^2
 * in eV and eV/Angstrom (E = ha2ev E_Ha, F = ha2ev/bohr2ang F_Ha/Bohr).
 * Frame f has nat[f] atoms starting at atom offset off[f]; lat is
 * nframes*9 (or NULL for clusters).  Eref: eV; Fref: eV/A, 3 per atom. */
int orc_nene_loss_grad(const NeneModel *m, int nframes, const int *nat, const long *off, const int *elem,
                       const double *xyz, const double *lat, const double *Eref, const double *Fref, double wE,
                       double wF, double *L, double *grad);

#ifdef __cplusplus
}
#endif
#endif
