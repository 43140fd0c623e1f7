/*
 * mdt_gpu.h -- C-ABI of libmdt_gpu.so, the MI355X (gfx950) implementation
 * of md_tian2's force-and-propagation hot path: EMT energies/forces/density,
 * the velocity-Verlet / Langevin+LDFA / Andersen / PILE propagators and the
 * ring-polymer normal-mode step, batched over trajectories and beads.
 *
 * Conventions (chosen so a Fortran caller binds them with ISO_C_BINDING and
 * passes its arrays unchanged):
 *   - arrays are the reference's Fortran arrays in column-major order:
 *     r(3,P,N) -> r[c + 3*(b + P*i)], dens(P,N) -> dens[b + P*i],
 *     simbox(3,3) column-major, pes(ntypes,ntypes) column-major;
 *   - type indices (idx) are 1-based as in the reference;
 *   - logicals are passed as int32 (0 = .false.);
 *   - every function returns 0 on success and a negative MDT_E* code on
 *     error; mdt_last_error() returns the message of the calling thread's
 *     last failure.
 * A context or batch handle is used from one host thread at a time.
 */
#ifndef MDT_GPU_H
#define MDT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDT_OK 0
#define MDT_EINVAL -1   /* bad argument */
#define MDT_EHIP -2     /* HIP runtime failure */
#define MDT_ESTATE -3   /* call out of order (e.g. no EMT parameters) */
#define MDT_ETRAJ -4    /* a trajectory flagged a failure (see mdt_batch_status) */
#define MDT_EUNSUP -5   /* configuration outside what this library implements */

/* pes ids (src/constants.f90:64-71) and propagator ids (:107-112) */
#define MDT_PES_EMT 2003
#define MDT_PROP_VERLET 1001
#define MDT_PROP_LANGEVIN 1003
#define MDT_PROP_ANDERSEN 1005
#define MDT_PROP_PILE 1006

/* energy/force flag of calc_force (src/constants.f90:114-115) */
#define MDT_ENERGY_AND_FORCE 3001
#define MDT_ENERGY_ONLY 3002

/* precision modes */
#define MDT_PREC_FP64 0      /* everything in FP64 */
#define MDT_PREC_FP32_PAIR 1 /* FP32 pair terms, FP64 accumulation (mdt_batch_set_precision) */

/* bead modes (mdt_batch_set_bead_mode) */
#define MDT_BEADS_ALL 0   /* every atom carries the trajectory's beads (the reference's model) */
#define MDT_BEADS_MIXED 1 /* the projectile carries the beads, the other atoms are classical */

/* per-trajectory status bits (mdt_batch_status) */
#define MDT_ST_NBR_OVERFLOW 1  /* neighbour list capacity exceeded */
#define MDT_ST_NEG_DENSITY 2   /* ldfa saw a negative density (src/md_algo.f90:551-554) */
#define MDT_ST_BAD_PROP 4      /* unknown propagator (src/md_algo.f90:66-67) */
#define MDT_ST_FIXED_OVERFLOW 8 /* a pair term exceeded the fixed-point accumulator range */
#define MDT_ST_NN_EXTRAPOLATION 16 /* a network input left its scaling range (EMT fallback applied) */

typedef struct mdt_context mdt_context;
typedef struct mdt_batch mdt_batch;

/* Message of the last failed call on this thread; returns its length. */
int mdt_last_error(char *buf, int buflen);

/* Library version (major*10000 + minor*100 + patch). */
int mdt_version(void);

/* Number of visible GPUs (0 when none). */
int mdt_device_count(int *count);

/* Context on GPU `device`: owns a HIP stream and the EMT parameter table. */
int mdt_create(int device, mdt_context **out);
int mdt_destroy(mdt_context *ctx);

/*
 * EMT parameters per type and the interaction matrix.
 * Replaces read_emt (src/pes_emt_mod.f90:49-190), which fills the module
 * arrays eta2, n0, e0, lambda, v0, kappa, s0 (:32-41) per type and sets
 * pes(idx1,idx2) = pes(idx2,idx1) = 2003 (:111-112).  The number of types with
 * pes(t,t) == 2003 selects compute_emt_1species or compute_emt_2species
 * (src/pes_emt_mod.f90:194-225); other counts return MDT_EUNSUP.
 */
int mdt_emt_params_set(mdt_context *ctx, int ntypes, const double *eta2, const double *n0,
                       const double *e0, const double *lambda, const double *v0,
                       const double *kappa, const double *s0, const int32_t *pes);

/*
 * Drop-in for compute_emt(atoms, flag) (src/pes_emt_mod.f90:194) as called
 * from calc_force (src/force.f90:79) on one universe:
 *   epot(b) += E_b (src/pes_emt_mod.f90:976), f(:,b,i) += F (:1001-1011),
 *   dens(b,i) = embedded density of every EMT atom (:855,861).
 * idx(N) 1-based types; simbox/isimbox the universe's cell and its inverse;
 * r(3,P,N) positions.  `flag` is accepted and ignored, as in the reference.
 * Minimum-image distances are computed on the GPU (no distance matrix).
 */
int mdt_compute_emt(mdt_context *ctx, int natoms, int nbeads, const int32_t *idx,
                    const double *simbox, const double *isimbox, const double *r, double *epot,
                    double *f, double *dens, int32_t flag);

/*
 * mdt_compute_emt for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference: a variant of
 * mdt_compute_emt, same arithmetic and semantics (epot(b) += E_b,
 * f(:,b,i) += F, dens(b,i) = embedded density of every EMT atom), with
 * every array a DEVICE pointer in the same layout (idx(N), simbox(3,3),
 * isimbox(3,3), r(3,P,N), epot(P), f(3,P,N), dens(P,N)); status (device,
 * may be NULL) receives the call's MDT_ST_* bits.  The universe is staged
 * in a fixed scratch array of about 4.7 MB per GPU, held until the process
 * ends: natoms <= 2048, nbeads <= 64, natoms * nbeads <= 16384 (larger
 * universes return MDT_EINVAL).  Launches go to `stream` (a hipStream_t;
 * NULL: the context's stream) and are not waited for; the call allocates
 * nothing and records no memset, so it can be captured in a graph.  Calls
 * on one GPU share the scratch array: calls that overlap (other streams or
 * threads) race.
 */
int mdt_compute_emt_noalloc(mdt_context *ctx, int natoms, int nbeads, const int32_t *idx,
                            const double *simbox, const double *isimbox, const double *r, double *epot,
                            double *f, double *dens, int32_t *status, int32_t flag, void *stream);

/*
 * Batch of `ntraj` independent trajectories resident on the GPU.
 * natoms[t], nbeads[t] may differ between trajectories.
 * Replaces the trajectory loop of src/md_tian2.f90:66-115.
 */
int mdt_batch_create(mdt_context *ctx, int ntraj, const int32_t *natoms, const int32_t *nbeads,
                     mdt_batch **out);
int mdt_batch_destroy(mdt_batch *b);

/*
 * Static per-atom data, concatenated over trajectories in batch order:
 * idx (sum N, 1-based), mass (sum N, program units eV fs^2/A^2),
 * is_fixed (sum 3*P*N, as universe%is_fixed(3,P,N)).
 */
int mdt_batch_set_atoms(mdt_batch *b, const int32_t *idx, const double *mass,
                        const int32_t *is_fixed);

/* Cells: simbox and isimbox, 9 doubles each per trajectory (column-major). */
int mdt_batch_set_cells(mdt_batch *b, const double *simbox, const double *isimbox);

/*
 * Propagation setup (src/run_config.f90 simparams and universe%algo/is_proj):
 * algo(ntypes), is_proj(ntypes) per type; dt [fs]; Tsurf, Tproj [K] per
 * trajectory; pile_tau and andersen_time [fs] (defaults 200 and 30 in the
 * reference, src/run_config.f90:119-120).
 */
int mdt_batch_set_propagation(mdt_batch *b, int ntypes, const int32_t *algo,
                              const int32_t *is_proj, double dt, const double *tsurf,
                              const double *tproj, double pile_tau, double andersen_time);

/*
 * Noise: Philox4x32-10 keyed by (seed, traj_id[t]); one draw per (trajectory,
 * step, atom, bead, component).  traj_id may be NULL (ids 0..ntraj-1).
 */
int mdt_batch_seed(mdt_batch *b, uint32_t seed, const uint32_t *traj_id);

/* Positions and velocities, sum 3*P*N each, reference (3,P,N) layout per trajectory. */
int mdt_batch_set_state(mdt_batch *b, const double *r, const double *v);

/*
 * Copy state back: any pointer may be NULL.  r, v, f, a: sum 3*P*N;
 * epot: sum P; dens: sum P*N (the reference's dens module array: raw EMT
 * density after calc_force, LDFA friction for Langevin atoms after a step).
 */
int mdt_batch_get_state(mdt_batch *b, double *r, double *v, double *f, double *a, double *epot,
                        double *dens);

/*
 * calc_force(atoms, energy_and_force) on every trajectory
 * (src/force.f90:40-98, as called at src/md_tian2.f90:69), including the
 * ring-polymer free step when P > 1 (src/force.f90:54).
 */
int mdt_batch_compute_forces(mdt_batch *b);

/*
 * `nsteps` MD steps of propagate_1 / calc_force / propagate_2
 * (src/md_tian2.f90:81-83) on every trajectory.  Step numbers continue from
 * the previous call (first step is 1), which keys the noise.
 */
int mdt_batch_step(mdt_batch *b, int nsteps);

/* Per-trajectory status bits (sum over all calls since the last clear). */
int mdt_batch_status(mdt_batch *b, int32_t *status, int clear);

/*
 * Kernel timing of the last mdt_batch_step call, measured with HIP events on
 * the batch's stream: total milliseconds and number of step launches.
 */
int mdt_batch_last_timing(mdt_batch *b, double *ms_total, int *nlaunch);

/*
 * Per-launch timing of the calc_force kernel: with profiling on, every
 * mdt_batch_step records HIP events around each force launch on the batch's
 * stream; mdt_batch_force_timing returns their summed milliseconds for the
 * last call (the launch count is mdt_batch_last_timing's nlaunch).
 */
int mdt_batch_profile(mdt_batch *b, int on);
int mdt_batch_force_timing(mdt_batch *b, double *force_ms_total);

/*
 * Launch tuning: threads per atom (1, 2, 4 or 8; 0 = autotune) and
 * neighbour-list capacity per atom (0 = default).  Returns the choice made.
 */
int mdt_batch_tune(mdt_batch *b, int tpa, int nbcap, int *tpa_chosen);

/*
 * Launch shape of the k_replica size classes (replicas grouped by atom
 * count; no counterpart in the reference, which has no GPU launch): up to
 * `max` classes, 6 int32 each -- largest atom count, replicas, threads per
 * atom, lists kept across launches with the Verlet skin (1: staged in LDS
 * during a launch, 2: read from HBM) or rebuilt every launch (0), threads
 * per workgroup, list capacity per atom.
 * *nclass receives the number of classes (0 in mixed bead mode).
 */
int mdt_batch_class_shapes(mdt_batch *b, int max, int32_t *out, int *nclass);

/*
 * Workgroup width of the large size classes (no counterpart in the
 * reference): classes whose atoms x threads per atom exceed 512 (k_replica,
 * one or two threads per atom) or whose classical atoms x threads per atom
 * exceed 512 (k_mixed) may run one 1024-lane workgroup per replica instead
 * of 512 lanes looping over the atoms.  mode -1: the autotuner times both
 * (default); 0: always 512; 1: always 1024 where the class is that large.
 * Results are bitwise the same either way.  The next launch autotunes
 * again.
 */
int mdt_batch_set_wide(mdt_batch *b, int mode);

/*
 * Neighbour lists of k_replica's size classes (an extension; the reference
 * visits every pair at every call): mode -1 autotuned (default); 0 rebuilt
 * at every launch (in LDS where they fit); 1 kept across launches with the
 * Verlet skin (mdt_batch_set_list_skin) in HBM and staged in LDS during a
 * launch (read from HBM where LDS is too small); 2 kept and read from HBM.
 * Forces agree to rounding whichever mode runs.  The next launch
 * configures again.
 */
int mdt_batch_set_keep(mdt_batch *b, int mode);

/*
 * Two streams for one-bead batches (no counterpart in the reference): a
 * batch of one size class, one bead per trajectory and the EMT potential
 * steps as two independent halves of its trajectories, each on its own
 * stream, so that one half's propagator launches overlap the other half's
 * force launches.  mode -1: when the batch has at least 1024 trajectories
 * (default); 0: never; 1: whenever the batch qualifies.  Not while
 * profiling.  Results are bitwise the same either way.
 */
int mdt_batch_set_split(mdt_batch *b, int mode);

/* *halves: 2 when the last mdt_batch_step ran as two halves, else 1. */
int mdt_batch_last_split(mdt_batch *b, int *halves);

/*
 * Arithmetic of the EMT pair terms in compute_emt_2species
 * (src/pes_emt_mod.f90:646-763, 793-842, 871-972), which the reference
 * evaluates in FP64:
 *   mode 0: FP64 throughout (the reference's arithmetic; default);
 *   mode 1: FP32 pair terms where the error stays within 1e-6 eV/atom and
 *           1e-5 eV/A of FP64, with r, the exponent arguments, the per-atom
 *           embedding terms and all accumulations (sigma, energy, forces) in
 *           FP64 or exact fixed point.  When every EMT type's embedding
 *           stiffness 6 V0 kappa / (beta eta2) is at most 20 eV (H 1.25, Au
 *           13.2 eV) the cut-off function, the sigma / pair-potential
 *           exponentials and their products are FP32 (arithmetic 1); otherwise
 *           (Pt: 129 eV, where FP32 sigma / V terms err by 6e-6 eV/atom and
 *           8e-5 eV/A) only the cut-off exponential and theta are FP32
 *           (arithmetic 2).  Integration, LDFA and the normal-mode transform
 *           are FP64 in both modes.
 */
int mdt_batch_set_precision(mdt_batch *b, int mode);

/* The pair-term arithmetic the next force launch uses with the context's
 * current EMT parameters: 0 FP64, 1 FP32 pair terms, 2 FP32 cut-off
 * exponential and theta only (see mdt_batch_set_precision). */
int mdt_batch_precision_arith(mdt_batch *b, int *arith);

/*
 * Bead model of the trajectories with nbeads > 1.  The reference gives every
 * atom the same number of beads (universe%r(3,nbeads,natoms),
 * src/universe_mod.f90:38-67; calc_force evaluates the PES per bead,
 * src/force.f90:40-98).  MDT_BEADS_MIXED is an extension without a
 * counterpart there: atom 1 (a projectile type) keeps its beads and the
 * other atoms (non-projectile types only) are classical, with their state
 * in bead 1's slots (positions and velocities given for other beads are not
 * read; the state getters report bead 1 in every bead).  The potential of
 * bead b is the EMT energy of the classical atoms with the projectile's
 * bead b; the projectile's bead b feels -dV_b/dr_b and a classical atom the
 * bead average of -dV_b/dR, so that the total energy
 *   sum_b (m v_b^2 / 2 + V_b) + ring-polymer springs + P sum_classical M V^2 / 2
 * is conserved by the Verlet propagator.  epot holds V_b per bead; dens of a
 * classical atom is its bead-averaged density.  Classical atoms propagate as
 * one-bead atoms (kB T / m thermostat temperature, PILE on the single mode,
 * no ring-polymer step).  EMT only, FP64 (the precision mode applies to the
 * one-bead trajectories).  Trajectories with nbeads == 1 are unaffected.
 * Checked at the next launch: atom 1 of a projectile type, no other
 * projectile atom.
 */
int mdt_batch_set_bead_mode(mdt_batch *b, int mode);

/*
 * COM velocity removal after propagate_1 (remove_com_velocity,
 * src/universe_mod.f90:222-241, called from src/md_algo.f90:71-74 whenever an
 * Andersen or PILE atom is present): on (1, the default) or off (0).  Off is
 * an extension for thermostatting a bare slab with fixed layers (thermal
 * slab libraries): the removal runs over fixed and mobile atoms alike, damps
 * the mobile atoms' collective motion every step and lowers their
 * temperature (by about 11 % for a 3x3x4 Au(111) slab with 2 fixed layers
 * under Andersen at 300 K).
 */
int mdt_batch_set_com_removal(mdt_batch *b, int on);

/* Verlet skin (A) of the neighbour lists kept across force calls (default
 * 0.3; 0 rebuilds every list at every call): the classical atoms' lists in
 * mixed bead mode, and the lists of size classes of more than 256 atoms,
 * which then live in HBM.  A list is built with cut-off + skin and kept while
 * no atom of its replica (or trajectory) moved more than skin / 2 since the
 * build, in cells whose narrowest width exceeds 2 (cut-off + skin); narrower
 * cells rebuild at every call.  Every pair is re-checked against the
 * cut-off under the reference's minimum image, so the skin changes at most
 * the summation order.  Extension: the reference visits every pair under
 * the minimum image at every call (src/pes_emt_mod.f90:1204-1206) and keeps
 * no lists.  Switching between 0 and a positive skin re-lays the size
 * classes' lists (keeping their tuned threads per atom). */
int mdt_batch_set_list_skin(mdt_batch *b, double skin);

/*
 * Trajectory observation: collect_trajectory_characteristics
 * (src/trajectory_info.f90:47-263) after each step's propagate_2, and the
 * early end of a trajectory (src/md_tian2.f90:88-92).  The call resets the
 * observables of every trajectory; from the next step on, per trajectory:
 * the closest centroid distance of the projectile (atom 1) to any other
 * atom (minimum image) and the time [fs] it was reached, the projectile
 * centroid at its lowest z, the time with a nonzero force on the
 * projectile, the sign changes of its centroid z velocity, and the
 * adsorption status (set below ads_start, cleared above ads_end).  With
 * exit_at_limits, a trajectory whose projectile beads are all above
 * proj_ul (`pul`) or all below proj_ll (`pll`) after a step is frozen: no
 * further propagation or forces; its state stays that of the exit step.
 * Defaults of the reference: proj_ul = DBL_MAX, proj_ll = -DBL_MAX,
 * ads_start = ads_end = DBL_MAX.  The call also unfreezes every trajectory;
 * mdt_batch_set_state does not, so a batch reused for new trajectories
 * calls this again after setting their state.
 */
int mdt_batch_set_observe(mdt_batch *b, double proj_ul, double proj_ll, double ads_start, double ads_end,
                          int exit_at_limits);

/* Observables of one trajectory; istep counts the steps from the
 * mdt_batch_set_observe call, 1-based as the reference's istep. */
typedef struct mdt_traj_obs {
    double cl_appr;          /* closest_approach [A] */
    double cl_last;          /* the closest distance at the last observed step [A] */
    double r_min[3];         /* lowest_z: projectile centroid at its lowest z [A] */
    double t_int;            /* interaction_time [fs] */
    int32_t turning_points;  /* nturning_points */
    int32_t moving_down;     /* was_moving_down */
    int32_t adsorbed;        /* is_adsorbed */
    int32_t adsorbed_prev;   /* is_adsorbed before the last observed step */
    int32_t exit_step;       /* istep after which the trajectory was frozen, 0 while it runs */
    int32_t cl_appr_step;    /* istep of the closest approach (closest_approach_time / step) */
    int32_t last_step;       /* last observed istep */
    int32_t pad_;
} mdt_traj_obs;

/* Observables of every trajectory into out[ntraj]. */
int mdt_batch_get_observables(mdt_batch *b, mdt_traj_obs *out);

/*
 * Estimators of the current state, MDT_NEST doubles per trajectory into
 * out[ntraj][MDT_NEST], in the order of columns 2-21 of the reference's
 * energy file (output_nrg, src/output_mod.f90:170-229):
 *   0 atom temperature [K] (calc_atom_temperature, universe_mod.f90:314-327)
 *   1 bead temperature [K] (calc_bead_temperature, rpmd.f90:63-77)
 *   2,3 squared radius of gyration, projectile / lattice [A^2]
 *       (radius_of_gyration, rpmd.f90:511-545)
 *   4,5 centroid kinetic energy, projectile / lattice (atom_ekin,
 *       universe_mod.f90:433-458)
 *   6,7 bead kinetic energy, projectile / lattice (bead_ekin, rpmd.f90:547-572)
 *   8,9 virial quantum kinetic energy, projectile / lattice
 *       (virial_quantum_ekin, rpmd.f90:273-301)
 *   10 potential energy averaged over beads, 11 ring-polymer spring energy
 *       (calc_bead_epot, rpmd.f90:236-264), 12 total energy 6+7+10+11,
 *   13 sum of all forces, 14-16 projectile centroid [A], 17-19 its
 *       velocity [A/fs].
 * Energies in eV.  In mixed bead mode a classical atom counts as P copies
 * of its bead (the state mdt_batch_get_state returns).  The temperatures
 * use dof = 3 N P - (number of fixed components over all beads).
 */
#define MDT_NEST 20
int mdt_batch_estimators(mdt_batch *b, double *out);

/*
 * Step-loop launches through a hipGraph: with chunk > 0, mdt_batch_step
 * replays a graph of `chunk` steps (captured on first use, recaptured after
 * mdt_batch_tune or mdt_batch_set_precision) and launches the remaining
 * nsteps % chunk steps directly; results are identical to chunk = 0 (direct
 * launches, the default).  Profiling (mdt_batch_profile) forces direct
 * launches.  Replaces the per-step calls of src/md_tian2.f90:78-98.
 */
int mdt_batch_set_graph(mdt_batch *b, int chunk);

/*
 * NENE surrogate: the short-range atomic network of read_nene / compute_nene
 * (src/pes_nene_mod.f90:123-5759, :7145-7368), RuNNer nn_type_short 1 with
 * cut-off type 1.  Arrays are the contents of input.nn, scaling.data and
 * weights.XXX.data, per network element (elements in ascending atomic
 * number), symmetry functions in the reference's sorted order (radial type
 * 2 first, then angular type 3):
 *   elem_of_type[ntypes]  network element of each atom type (-1: none)
 *   nsf[nelem]            functions per element; the per-function arrays
 *                         below hold sum(nsf) entries, element-major
 *   sf_type, sf_e1, sf_e2 type 2/3 and neighbour elements (element indices)
 *   eta, rs, lambda, zeta, rc   parameters (Bohr units; one cut-off)
 *   gmin, gmax, gavg      scaling.data columns
 *   scmin, scmax, scale, center  scale_min/max_short_atomic,
 *                         scale_symmetry_functions, center_symmetry_functions
 *   nh1, nh2              nodes of the two hidden layers (<= 32)
 *   activations           three letters (t, s or l) of global_activation_short
 *   weights               per element in weights.XXX.data order, concatenated
 * Limits: 4 elements, 64 functions (32 radial, 32 angular), zeta in
 * {1, 2, 4, 8, 16}.
 */
int mdt_nn_params_set(mdt_context *ctx, int nelem, const int32_t *elem_of_type, int ntypes, const int32_t *nsf,
                      const int32_t *sf_type, const int32_t *sf_e1, const int32_t *sf_e2, const double *eta,
                      const double *rs, const double *lambda, const double *zeta, const double *rc,
                      const double *gmin, const double *gmax, const double *gavg, double scmin, double scmax,
                      int32_t scale, int32_t center, int32_t nh1, int32_t nh2, const char *activations,
                      const double *weights);

/*
 * read_nene's file part (src/pes_nene_mod.f90:123-5759): reads input.nn,
 * scaling.data and weights.XXX.data from `directory` (the reference's
 * inp_path) and uploads the network as mdt_nn_params_set does.
 * type_symbols: ntypes element symbols of 2 characters each (space padded),
 * the universe's type names in idx order.
 */
int mdt_nn_load(mdt_context *ctx, const char *directory, int ntypes, const char *type_symbols);

/*
 * compute_nene(atoms, flag) on one universe (src/pes_nene_mod.f90:7145-7368):
 * epot(b) += E_b [eV], f(:,b,i) += F [eV/A]; r, f column-major (3,P,N) in
 * Angstrom as in compute_emt.  nextrap (may be NULL) is set to 1 when a
 * symmetry function lies outside the scaling range by more than 1e-4 (the
 * reference counts these as extrapolation warnings and goes on).
 */
int mdt_compute_nene(mdt_context *ctx, int natoms, int nbeads, const int32_t *idx, const double *simbox,
                     const double *isimbox, const double *r, double *epot, double *f, int32_t *nextrap,
                     int32_t flag);

/*
 * Potential of a batch's calc_force (the dispatch of src/force.f90:77-82):
 * pes 0 = EMT, 1 = the network, with EMT (when its parameters are set) on
 * every replica whose symmetry functions leave the scaling range, and with
 * EMT as the judge of the network on a rotating 1 % of the replicas every
 * judge_interval steps (0: no judge).  Has no counterpart in the reference,
 * which runs one potential per trajectory.
 */
int mdt_batch_set_pes(mdt_batch *b, int pes, int judge_interval);

/*
 * Share of the replicas the EMT judge takes at each of its steps (no
 * counterpart in the reference): percent, a divisor of 100 (default 1).
 * Replicas with rep % (100 / percent) == phase are judged, the phase
 * rotating per judge step; the judged share of all frames is percent /
 * judge_interval %.  One judge launch takes about one replica's EMT time
 * however many replicas it judges, so 4 % every 4 steps judges the frames
 * of 1 % every step at a quarter of the launches.
 */
int mdt_batch_set_judge_percent(mdt_batch *b, int percent);

/* Judge statistics accumulated since the network first ran on the batch:
 * sum of squared force differences [(eV/A)^2], number of force components
 * judged, largest difference [eV/A]; replicas on the EMT fallback in the
 * last force evaluation. */
int mdt_batch_nn_stats(mdt_batch *b, double *judge_sq, double *judge_n, double *judge_max, int32_t *nfallback);

/*
 * Per replica (trajectories in order, beads within), 1 when the last force
 * evaluation found a symmetry function outside the scaling range and the
 * replica took the EMT fallback: the per-replica form of the extrapolation
 * warnings compute_nene prints (src/pes_nene_mod.f90:7260-7261).
 */
int mdt_batch_nn_fallback(mdt_batch *b, int32_t *flags);

/* Raw symmetry functions of the last network evaluation, float, nrep x maxN x
 * 64 (replicas in batch order, trajectory-major; unused entries
 * unspecified): the training set's descriptors and scaling statistics.
 * Has no counterpart in the reference (RuNNer writes them in mode 1). */
int mdt_batch_nn_descriptors(mdt_batch *b, float *G);

/*
 * Training of the surrogate on the GPU: the project's own forward, backward
 * and force-backward kernels (BF16 MFMA, FP32 accumulation) with Adam.  No
 * counterpart in the reference, which fits with RuNNer's Kalman filter
 * (src/fit.f90).  Loss over a minibatch of Nf frames with Ntot atoms, in eV
 * and eV/A:  L = wE/Nf sum_f ((E_f - Eref_f)/N_f)^2 + wF/(3 Ntot) sum |F - Fref|^2.
 *   mdt_nn_train_create: frames (natoms[f] atoms each; idx 1-based types,
 *     simbox 9 per frame column-major [A], r 3 per atom [A], Eref [eV],
 *     Fref 3 per atom [eV/A]) uploaded once, descriptors computed once on
 *     the GPU; fixed minibatches of batch_frames frames over a seeded
 *     permutation, the last val_fraction of them held out;
 *   mdt_nn_train_descriptor_stats: per element (nelem x 64) min, max, mean
 *     of the symmetry functions over all frames (scaling.data values; pass
 *     them to mdt_nn_params_set);
 *   mdt_nn_train_epoch: one pass over the training minibatches in a seeded
 *     random order, one Adam step each; loss = mean minibatch loss;
 *   mdt_nn_train_eval: energy RMSE per atom [eV/atom] and force-component
 *     RMSE [eV/A] on the training (set 0) or validation (set 1) frames;
 *   mdt_nn_train_loss_grad: loss and its gradient (weights.XXX.data layout,
 *     elements concatenated) over the given frames, network unchanged;
 *   mdt_nn_get_weights: the current weights in that layout.
 * The network must have a linear output layer.
 */
typedef struct mdt_nntrain mdt_nntrain;
int mdt_nn_train_create(mdt_context *ctx, int nframes, const int32_t *natoms, const int32_t *idx,
                        const double *simbox, const double *r, const double *Eref, const double *Fref,
                        int batch_frames, double val_fraction, uint32_t seed, mdt_nntrain **out);
int mdt_nn_train_destroy(mdt_nntrain *tr);
int mdt_nn_train_descriptor_stats(mdt_nntrain *tr, double *gmin, double *gmax, double *gavg);
int mdt_nn_train_epoch(mdt_nntrain *tr, double lr, double wE, double wF, double *loss);
int mdt_nn_train_eval(mdt_nntrain *tr, int set, double *rmse_e, double *rmse_f);
int mdt_nn_train_loss_grad(mdt_nntrain *tr, int nsel, const int32_t *frames, double wE, double wF, double *loss,
                           double *grad);
int mdt_nn_get_weights(mdt_context *ctx, double *weights);

/* ------------------------------------------------------------------
 * Initial conditions of a trajectory (host code, no GPU needed).  Every
 * draw is a uniform of the Philox4x32-10 stream keyed by (seed, trajectory
 * id) at its own slot, so a trajectory's initial conditions depend only on
 * its id and the seed (the reference draws them in sequence from Fortran's
 * random_number, whose stream is not reproducible across compilers).
 * Arrays are the reference's: r, v (3, nbeads, natoms) column-major, idx
 * 1-based types, is_proj per type, simbox (3,3) column-major (columns are
 * the cell vectors).
 * ------------------------------------------------------------------ */
#define MDT_SLOT_PROJ_FILE 0u
#define MDT_SLOT_LATT_FILE 1u
#define MDT_SLOT_PIP 2u       /* 2, 3, 4: x, y, z */
#define MDT_SLOT_EINC 5u      /* 5, 6 */
#define MDT_SLOT_AZIMUTH 7u   /* 7 (random) or 7, 8 (normal) */
#define MDT_SLOT_MAXWELL 16u  /* 16 + 2 (3 i + c) and the next */

/* pip x y z (src/run_config.f90: pip key): random[k] for "r" */
typedef struct mdt_pip {
    int32_t random[3];
    double value[3];
} mdt_pip;

/* Einc <e> [sigma], polar <deg>, azimuth <deg|r> [sigma]
 * (src/run_config.f90:361-400) */
typedef struct mdt_incidence {
    double einc, sigma_einc;
    double polar_deg;
    int32_t azimuth_random;
    int32_t pad_;
    double azimuth_deg, sigma_azimuth;
} mdt_incidence;

/* uniform in [0, 1) at a slot of the trajectory's sampling stream */
double mdt_sample_uniform(uint32_t seed, uint32_t traj_id, uint32_t slot);

/* the file of a conf folder: int(rnd * n) + 1 (prepare_next_traj,
 * src/md_init.f90:904-926); which = 0 projectile folder, 1 lattice folder */
int mdt_sample_conf_index(uint32_t seed, uint32_t traj_id, int which, int n);

/* apply_pip (src/md_init.f90:498-554) on the projectile universe (r
 * in-out, cartesian), with that universe's cell: each fixed component is
 * added to every bead; for each random component k the whole target vector
 * is scaled by the column sum of simbox(:,k); then every atom's beads are
 * moved so that the centroid sits at simbox . target (plus the fixed
 * components). */
int mdt_apply_pip(const mdt_pip *pip, uint32_t seed, uint32_t traj_id, int natoms, int nbeads,
                  const double *simbox, double *r);

/* set_proj_incidence (src/md_init.f90:829-877): Einc drawn from a normal of
 * width sigma_einc, azimuth uniform (random) or normal of width
 * sigma_azimuth, vinc = sqrt(2 Einc / projectile mass); each leading
 * projectile atom loses its bead-centroid velocity and gains
 * vinc (sin(polar) cos(az), sin(polar) sin(az), -cos(polar)). */
int mdt_set_proj_incidence(const mdt_incidence *inc, uint32_t seed, uint32_t traj_id, int natoms, int nbeads,
                           const int32_t *idx, const int32_t *is_proj, const double *mass, double *v);

/* Maxwell-Boltzmann velocities at `temperature` for the non-projectile
 * atoms (the same in every bead; fixed components 0).  No counterpart in
 * the reference, which reads thermalised slabs from mxt files; for
 * generated slabs. */
int mdt_sample_maxwell(double temperature, uint32_t seed, uint32_t traj_id, int natoms, int nbeads,
                       const int32_t *idx, const int32_t *is_proj, const double *mass, const int32_t *is_fixed,
                       double *v);

#ifdef __cplusplus
}
#endif
#endif /* MDT_GPU_H */
