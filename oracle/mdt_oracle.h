/*
 * mdt_oracle.h -- CPU FP64 restatement of md_tian2's force-and-propagation
 * hot path.  TEST INFRASTRUCTURE ONLY: the oracle is the checker that the
 * HIP path is compared against, and the serial CPU baseline that bench.py
 * times beside it.  Nothing in the product (the md_tian2_amd package, the
 * C-ABI library, the Fortran host) links, loads or calls it.
 *
 * Layout follows the reference exactly (src/universe_mod.f90:38-67):
 * Fortran column-major arrays, r(3,P,N) -> r[c + 3*(b + P*i)], dens(P,N) ->
 * dens[b + P*i], pes(ntypes,ntypes) -> pes[ti + ntypes*tj].  Type indices
 * (idx) are 0-based here (the reference's are 1-based).
 *
 * Randomness: the reference draws from Fortran's intrinsic random_number,
 * whose stream belongs to the compiler runtime and cannot be reproduced.
 * Every stochastic routine here therefore takes its normal deviates as an
 * explicit argument; orc_philox_* generates them with the same counter
 * scheme the GPU path uses, so GPU and oracle consume identical noise.
 */
#ifndef MDT_ORACLE_H
#define MDT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXTYPES 8

/* constants.f90:64-71, 107-115 */
#define ORC_PES_EMT 2003
#define ORC_PROP_VERLET 1001
#define ORC_PROP_LANGEVIN 1003
#define ORC_PROP_ANDERSEN 1005
#define ORC_PROP_PILE 1006

/* RNG stream identifiers (philox counter word 2) */
#define ORC_STREAM_LANGEVIN 0u
#define ORC_STREAM_PILE 1u
#define ORC_STREAM_ANDERSEN_V 2u
#define ORC_STREAM_ANDERSEN_U 3u

typedef struct {
    int ntypes;
    /* pes_emt_mod.f90:32-41, indexed by type */
    double eta2[ORC_MAXTYPES], n0[ORC_MAXTYPES], e0[ORC_MAXTYPES];
    double lambda[ORC_MAXTYPES], v0[ORC_MAXTYPES], kappa[ORC_MAXTYPES];
    double s0[ORC_MAXTYPES];
} orc_emt_params;

typedef struct {
    int natoms, nbeads, ntypes;
    double *r, *v, *f, *a; /* (3,P,N) */
    double *m;             /* (N) */
    int *is_fixed;         /* (3,P,N) 0/1 */
    int *idx;              /* (N) 0-based type */
    int *pes;              /* (ntypes,ntypes) column-major */
    int *algo;             /* (ntypes) */
    int *is_proj;          /* (ntypes) 0/1 */
    double *epot;          /* (P) */
    double *dens;          /* (P,N): pes_emt_mod.f90:45 module array */
    double simbox[9];      /* column-major 3x3 */
    double isimbox[9];
} orc_universe;

typedef struct {
    double step, Tsurf, Tproj, pile_tau, andersen_time;
} orc_simparams;

/* ---- min-image (universe_mod.f90:411-428) ---- */
void orc_minimg_beads(const orc_universe *u, int i, int j, double *dist,
                      double *vec);

/* ---- EMT (pes_emt_mod.f90:194-1027). distances (P,N,N) and vectors
 * (3,P,N,N) are the calc_force matrices (force.f90:58-74). Returns 0 on
 * success, nonzero when the reference would stop. ---- */
int orc_compute_emt(orc_universe *u, const orc_emt_params *p,
                    const double *distances, const double *vectors);

/* EMT only: zero f, epot; min-image matrices; compute_emt (no RP step) */
int orc_emt_forces(orc_universe *u, const orc_emt_params *p);

/* ---- force.f90:40-98 (calc_force incl. RP step when P>1) ---- */
int orc_calc_force(orc_universe *u, const orc_emt_params *p,
                   const orc_simparams *sp, const double *cjk);

/* ---- rpmd.f90:36-59 and 309-452 ---- */
void orc_build_cjk(int nbeads, double *cjk);
void orc_ring_polymer_step(orc_universe *u, const orc_simparams *sp,
                           const double *cjk);

/* ---- md_algo.f90 ---- */
/* randy: (3,P,N) normals shared by both Langevin halves (md_algo.f90:43-48)
 * zeta : (3,P,N) normals for PILE, one (3,P) block per atom (md_algo.f90:485)
 * and_z: (3,P,N) unit normals for Andersen (scaled inside, md_algo.f90:421)
 * and_u: (P,N) uniforms for Andersen (md_algo.f90:423)
 * Any of them may be NULL when the corresponding algorithm is absent. */
int orc_propagate_1(orc_universe *u, const orc_simparams *sp,
                    const double *randy, const double *zeta,
                    const double *and_z, const double *and_u,
                    const double *cjk);
int orc_propagate_2(orc_universe *u, const orc_simparams *sp,
                    const double *randy);
int orc_ldfa(orc_universe *u, int i);
void orc_remove_com_velocity(orc_universe *u);

/* ---- one full MD step, md_tian2.f90:81-83 ---- */
int orc_md_step(orc_universe *u, const orc_emt_params *p,
                const orc_simparams *sp, const double *cjk,
                const double *randy, const double *zeta, const double *and_z,
                const double *and_u);

/* ---- counter-based RNG: Philox4x32-10 (Salmon et al., SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
double orc_philox_normal(uint64_t seed, uint32_t traj, uint32_t step,
                         uint32_t stream, uint32_t slot, uint32_t comp);
double orc_philox_uniform(uint64_t seed, uint32_t traj, uint32_t step,
                          uint32_t stream, uint32_t slot, uint32_t comp);
/* fills out(3,P,N) for one (traj, step, stream): slot = i*P + b */
void orc_philox_fill_normals(uint64_t seed, uint32_t traj, uint32_t step,
                             uint32_t stream, int natoms, int nbeads,
                             double *out);

/* ---- serial CPU baseline driver: nsteps of the reference loop for one
 * trajectory (md_tian2.f90:78-98) with noise from orc_philox_*.
 * step0 is the index of the first step (1-based as in the reference). */
int orc_run_trajectory(orc_universe *u, const orc_emt_params *p,
                       const orc_simparams *sp, uint64_t seed, uint32_t traj,
                       int step0, int nsteps);

/* ---- trajectory characteristics (trajectory_info.f90:25-43, 47-263):
 * the module variables of one trajectory.  is_adsorbed is not reset at
 * istep 1 in the reference (it carries over from the previous trajectory);
 * the caller sets its starting value. ---- */
#define ORC_DEFAULT_REAL 1.7976931348623157e308 /* huge(0_dp), constants.f90:101 */
typedef struct {
    double closest_approach, closest_approach_time, lowest_z[3], interaction_time;
    int nturning_points, was_moving_down, is_adsorbed;
} orc_traj_obs;

void orc_collect_characteristics(const orc_universe *u, const orc_simparams *sp, double ads_start,
                                 double ads_end, int istep, orc_traj_obs *o);
/* noise of step istep from orc_philox_* with step index istep (1-based) */
int orc_run_scatter(orc_universe *u, const orc_emt_params *p, const orc_simparams *sp, uint64_t seed,
                    uint32_t traj, int nsteps, double proj_ul, double proj_ll, int scatter, double ads_start,
                    double ads_end, orc_traj_obs *o, int *istep_end);

/* ---- estimators (rpmd.f90:63-77, 236-301, 511-572; universe_mod.f90:
 * 314-327, 433-458) and the energy-file row of output_nrg
 * (output_mod.f90:170-229, columns 2-21) ---- */
double orc_atom_temperature(const orc_universe *u);
double orc_bead_temperature(const orc_universe *u);
void orc_radius_of_gyration(const orc_universe *u, double *rgyr_p, double *rgyr_l);
void orc_atom_ekin(const orc_universe *u, double *ekin_p, double *ekin_l);
void orc_bead_ekin(const orc_universe *u, double *ekin_p, double *ekin_l);
void orc_virial_quantum_ekin(const orc_universe *u, double *ekin_p, double *ekin_l);
double orc_bead_epot(const orc_universe *u, const orc_simparams *sp);
void orc_energy_row(const orc_universe *u, const orc_simparams *sp, double out[20]);

/* ---- initial conditions (md_init.f90:498-554, 829-877) with the
 * uniforms given explicitly ---- */
void orc_apply_pip(orc_universe *u, const int random[3], const double value[3], const double U[3]);
void orc_set_proj_incidence(orc_universe *u, double einc, double sigma_einc, double polar_deg, int az_random,
                            double az_deg, double sigma_az, const double rnd[4]);

#ifdef __cplusplus
}
#endif
#endif
