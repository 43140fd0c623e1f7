// mdt_device.h -- shared host/device definitions for the batched EMT /
// propagator / RPMD kernels (gfx950).
//
// Device state layout in HBM (FP64, structure of arrays):
//   per trajectory t with N_t atoms and P_t beads, a block of 3*P_t*N_t
//   doubles at traj_soff[t]:  [bead b][component c][atom i]
//   so one replica (t, b) is three contiguous N_t-arrays x, y, z.
//   P*N arrays (dens, fixed bits) live at traj_soff[t]/3: [bead b][atom i].
//   Per-atom arrays (type, mass) live at traj_aoff[t]: [atom i].
#pragma once
#include <stdint.h>

#define MDT_MAXT 4

// physical constants, identical to the reference's src/constants.f90:33-39,157-160
#define MDT_PI 3.14159265358979323846
#define MDT_KB 8.61733238496e-5
#define MDT_HBAR 0.6582119514467406
#define MDT_AMU2MASS 103.638239276
#define MDT_BOHR2ANG 0.529177211
#define MDT_TOLERANCE 1.0e-9
#define MDT_TWELFTH (1.0 / 12.0)

// propagator ids, src/constants.f90:107-112
#define MDT_PROP_VERLET 1001
#define MDT_PROP_LANGEVIN 1003
#define MDT_PROP_ANDERSEN 1005
#define MDT_PROP_PILE 1006

// RNG streams (Philox counter word 2)
#define MDT_STREAM_LANGEVIN 0u
#define MDT_STREAM_PILE 1u
#define MDT_STREAM_ANDERSEN_V 2u
#define MDT_STREAM_ANDERSEN_U 3u

// per-trajectory status bits
#define MDT_ST_NBR_OVERFLOW 1
#define MDT_ST_NEG_DENSITY 2
#define MDT_ST_BAD_PROP 4
#define MDT_ST_FIXED_OVERFLOW 8

// Precision mode 1 evaluates the sigma and pair-potential exponentials in
// FP32 only when every EMT type's embedding stiffness B = 6 V0 kappa /
// (beta eta2) [eV] -- the scale of dE/d(ln sigma) at s = 0, which turns a
// relative rounding of sigma or of the pair sum into an energy error -- is
// at most this bound: H 1.25, Au 13.2 eV (FP32 terms give <= 3.3e-7
// eV/atom, 5.5e-6 eV/A); Pt has 129 eV (6e-6 eV/atom, 7.6e-5 eV/A, outside
// 1e-6 eV/atom and 1e-5 eV/A), so there only the cut-off exponential and
// theta are FP32.
#define MDT_FP32_STIFF_MAX 20.0

// An atom whose sigma~ from its near pairs is below this bound is "tiny":
// its far pairs (and those of its partners with it) are evaluated too.  For
// sigma~ >= 1e-2, |G| = |dE/dsigma~| <= 1.4e4 eV (the embedding terms give
// |G sigma~| <= 140 eV), so a far pair's force on either atom is at most
// |G| csig (eta + acut) 2^-56 < 4e-11 eV/A (csig <= 6.5, eta + acut <= 31 /A
// for H/Au and H/Pt), and its share of sigma~ at most csig 2^-56 / 1e-2 <
// 1e-14 relative.  The reference's far-H force (ds/dr divided by the
// 1e-30-clamped sigma~) depends on the relative size of those tiny terms,
// which this keeps exact.
#define MDT_SIG_TINY 1.0e-2

// EMT constants derived on the host exactly as src/pes_emt_mod.f90:533-592
// (2 species) or :267-307 (1 species), flattened into per-type and
// per-(type a, neighbour type n) tables for a gather formulation.
struct EmtConst {
    int ntypes;
    int fp32_ok;  // every EMT type's embedding stiffness 6 V0 kappa / (beta eta2) <= MDT_FP32_STIFF_MAX
    double rcut, acut, cutr;  // cutr = 1.5 * rcut (the reference's hard cut)
    double cut2hi;            // neighbour-list bound on r^2 (cutr^2 widened by 1e-12 rel)
    // Near radius: beyond it every pair weight theta(r) exp(-a (r - beta s0)),
    // a in {eta2, kappa/beta} of either type, is below 2^-56.  Pairs between
    // rnear and cutr are evaluated only when one of the two atoms has sigma~ <
    // MDT_SIG_TINY; elsewhere each changes no force by more than 4e-11 eV/A
    // and no sigma~ by more than 1e-14 relative.
    double rnear;
    double eta2[MDT_MAXT], betas0[MDT_MAXT], kappadbeta[MDT_MAXT], betaeta2[MDT_MAXT];
    double e0[MDT_MAXT], lambda[MDT_MAXT], v0[MDT_MAXT], kappa[MDT_MAXT], n0[MDT_MAXT];
    double densc[MDT_MAXT];            // 0.5/bohr2ang - beta*eta2
    double csig[MDT_MAXT * MDT_MAXT];  // sigma weight chi(ta,tn)/gamma1(ta)
    double etad[MDT_MAXT * MDT_MAXT];  // eta in the sigma derivative (type1-type1 quirk)
    double wv[MDT_MAXT * MDT_MAXT];    // pair energy weight V0(ta)chi(ta,tn)/gamma2(ta)
    int owner[MDT_MAXT];               // pes(t,t) == emt
    int pair_on[MDT_MAXT * MDT_MAXT];  // pes(ta,tn) == emt
};

struct PropConst {
    double dt, pile_tau, andersen_time;
    double com_factor_unused;
    int algo[MDT_MAXT];
    int is_proj[MDT_MAXT];
    int any_langevin, any_pile, any_andersen;
    int mixed;  // 1: atoms of non-projectile types are classical (one bead, kept in bead 0)
    uint32_t seed;
    int no_com;  // 1: no remove_com_velocity after propagate_1 (an extension; see mdt_batch_set_com_removal)
};

struct BatchView {
    int ntraj, nrep;
    int nbcap;      // neighbour-list capacity per atom
    int precision;  // pair-term arithmetic of k_replica: 0 FP64; 1 FP32 pair terms; 2 FP32 cut-off exponential only
    const int *rep_traj, *rep_bead;
    const int *traj_N, *traj_P;
    const long long *traj_soff, *traj_aoff;
    const uint8_t *type;    // per atom
    const double *mass;     // per atom
    const uint8_t *fixed;   // per atom-bead: bit c set = component c frozen
    const double *cell;     // 9 per traj, column-major (simbox)
    const double *icell;    // 9 per traj (isimbox)
    const double *tsurf, *tproj;
    const uint32_t *traj_id;
    double *x, *v, *a, *f;  // 3PN blocks
    double *dens;           // PN blocks
    double *eta;            // 3PN: langevin_1's deviates of the current step, reused by langevin_2
    double *epot;           // per replica (P per traj, at traj_soff/(3N))
    const long long *traj_eoff;  // offset of a trajectory's epot(1..P)
    int *status;            // per traj
    uint16_t *nbr_global;   // NULL: neighbour lists in LDS; else nrep * nbr_stride entries
    long long nbr_stride;   // entries per replica slice (maxN * cap)
    const int *rep_list;    // non-NULL: block k works on replica rep_list[k] (size-class launches)
    const int *only;        // non-NULL: calc_force only on replicas with only[rep] != 0
    const int *skip;        // non-NULL: no calc_force on replicas with skip[rep] != 0
    // judge mode (judge_acc non-NULL): replicas with rep % judge_mod ==
    // judge_phase compute EMT forces and accumulate, per replica, the sum of
    // squared differences to judge_f, the component count and the largest
    // difference into judge_acc[3 rep ..]; nothing else is written
    const double *judge_f;
    double *judge_acc;
    int judge_mod, judge_phase;
    // trajectory observation (collect_trajectory_characteristics,
    // src/trajectory_info.f90:47-263; the exit test of src/md_tian2.f90:88-92)
    // when obs != NULL: after each step's propagate_2 k_traj updates
    // obs[t], and a trajectory whose projectile beads are all above
    // obs_ul or all below obs_ll is frozen (done[t] set: k_traj and the
    // force kernels skip it from then on)
    // mixed bead mode, per listed trajectory (slot = position in the list):
    // the classical atoms' lists are kept across steps (Verlet skin) when
    // mix_skin > 0 and the cell's narrowest width exceeds 2 (cut-off + skin):
    // positions and fold at the list build (3 per classical atom each,
    // mix_ns atoms per slot), list counts, and a valid flag
    double *mix_ref, *mix_fold;
    uint16_t *mix_cnt;
    int *mix_valid;
    double mix_skin;
    long long mix_ns;
    // k_replica's lists kept across launches (Verlet skin), per block slot of
    // a size-class launch with lists in HBM: positions at the list build (SoA,
    // rl_ns per coordinate), list counts and a valid flag; rl_valid NULL: the
    // lists are rebuilt at every launch
    // rl_stage != 0: the kept lists live in nbr_global between launches and
    // are staged in LDS during one (the kernel's LDS-list instantiation)
    double *rl_ref;
    uint16_t *rl_cnt;
    int *rl_valid;
    double rl_skin;
    long long rl_ns;
    int rl_stage;
    struct TrajObs *obs;
    int *done;
    double obs_ul, obs_ll, ads_start, ads_end;
    int obs_exit;        // 1: the upper / lower limit freezes the trajectory
    uint32_t obs_step0;  // device step count when observation started: istep = step - obs_step0
};

// estimators per trajectory written by k_estimators (the columns 2-21 of
// the reference's energy file)
#define MDT_NEST 20

// Per trajectory, after the step istep (1-based from the start of
// observation); the layout of mdt_traj_obs in the C interface.
struct TrajObs {
    double cl_appr;        // closest centroid distance projectile - other atom [A]
    double cl_last;        // that distance at the last observed step [A]
    double r_min[3];       // projectile centroid at its lowest z
    double t_int;          // time with a nonzero force on the projectile [fs]
    int turning_points;    // sign changes of the centroid's z velocity
    int moving_down;       // the centroid's z velocity was < 0 at the last step
    int adsorbed;          // within ads_start, until farther than ads_end
    int adsorbed_prev;     // adsorbed before the last observed step
    int exit_step;         // istep after which the trajectory was frozen (0: running)
    int cl_appr_step;      // istep of the closest approach
    int last_step;         // last observed istep
    int pad_;
};

// ---------------------------------------------------------------------------
// NENE (Behler-Parrinello atomic network) surrogate: the model in the form
// compute_nene evaluates (src/pes_nene_mod.f90:7145-7368) -- radial (type 2)
// and angular (type 3) symmetry functions with one cosine cut-off, min/max
// scaling, per element a network of inputs -> 2 hidden layers -> 1.
// Lengths in Bohr, energies in Hartree (the network's units).  Inputs and
// hidden layers are zero-padded to NN_K1 and NN_H (exact: padded weights
// are zero).
// ---------------------------------------------------------------------------
#define NN_MAXE 4
#define NN_MAXR 32
#define NN_MAXA 32
#define NN_K1 64
#define NN_H 32

struct NnConst {
    int nelem;
    int act1, act2, act3;            // activation codes: 0 linear, 1 tanh, 2 sigmoid
    int elem_of_type[MDT_MAXT];      // batch atom type -> network element (-1: not in the network)
    float rc, rc2, pi_rc;            // single cut-off radius (Bohr), rc^2, pi/rc
    float ang2bohr;
    int nr[NN_MAXE], na[NN_MAXE];    // radial / angular functions per element (inputs 0..nr-1, nr..nr+na-1)
    int r_el[NN_MAXE][NN_MAXR];      // neighbour element of radial function
    float r_eta[NN_MAXE][NN_MAXR], r_rs[NN_MAXE][NN_MAXR];
    int a_e1[NN_MAXE][NN_MAXA], a_e2[NN_MAXE][NN_MAXA];
    float a_etal2[NN_MAXE][NN_MAXA];  // eta log2(e)
    float a_lam[NN_MAXE][NN_MAXA];
    float a_norm[NN_MAXE][NN_MAXA];   // 2^(1 - zeta)
    int a_zp[NN_MAXE][NN_MAXA];       // zeta = 2^a_zp (1, 2, 4, 8 or 16)
    float smul[NN_MAXE][NN_K1], sadd[NN_MAXE][NN_K1];  // network input x = G smul + sadd
    float dfac[NN_MAXE][NN_K1];                         // dE/dG = dE/dx dfac
    float glo[NN_MAXE][NN_K1], ghi[NN_MAXE][NN_K1];     // extrapolation bounds min - 1e-4, max + 1e-4
    float W1[NN_MAXE][NN_K1][NN_H], b1[NN_MAXE][NN_H];
    float W2[NN_MAXE][NN_H][NN_H], b2[NN_MAXE][NN_H];
    float w3[NN_MAXE][NN_H], b3[NN_MAXE];
    int nh1, nh2;                                       // real hidden widths (<= NN_H)
    // radial-only networks (no element has angular functions) run
    // k_nene_rg / k_nene_rf, which visit an element's radial functions
    // grouped by neighbour element: functions rk_lo[e][x] .. rk_hi[e][x]-1
    // of that order act on neighbours of element x; per function its input
    // index, shift, exponent factor -eta log2(e) and derivative factor -2 eta
    int radial_only;
    int rk_lo[NN_MAXE][NN_MAXE], rk_hi[NN_MAXE][NN_MAXE];
    int rk_perm[NN_MAXE][NN_MAXR];
    float rk_rs[NN_MAXE][NN_MAXR], rk_ce[NN_MAXE][NN_MAXR], rk_b[NN_MAXE][NN_MAXR];
};

// padded gradient / Adam layout of one element's parameters
#define NN_PW1 0
#define NN_PB1 (NN_K1 * NN_H)
#define NN_PW2 (NN_PB1 + NN_H)
#define NN_PB2 (NN_PW2 + NN_H * NN_H)
#define NN_PW3 (NN_PB2 + NN_H)
#define NN_PB3 (NN_PW3 + NN_H)
#define NN_NPAR (NN_PB3 + 1)

// training set resident on the GPU (frames of atoms, reference energies and forces)
struct NnTrainView {
    int nframes;
    const int *fr_natoms;        // per frame
    const long long *fr_aoff;    // first atom of the frame
    const double *fr_cell, *fr_icell;  // 9 per frame, Angstrom, column-major
    const uint8_t *type;         // per atom, 0-based type
    const double *pos;           // per atom x, y, z (Angstrom)
    const double *Fref;          // per atom (eV/A)
    const double *Eref;          // per frame (eV)
    const int *atom_frame;       // per atom
    float *G, *C, *U;            // per atom NN_K1: raw symmetry functions, dE/dG, dLoss/d(dE/dx)
    double *Eatom;               // per atom (Hartree)
    double *cE;                  // per frame: dLoss/dE (per Hartree)
    double *loss;                // per frame: energy part, force part
};

// per-replica NN status bit (in the trajectory status word)
#define MDT_ST_NN_EXTRAPOLATION 16
