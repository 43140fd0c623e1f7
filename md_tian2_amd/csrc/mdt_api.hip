// mdt_api.hip -- host side of libmdt_gpu.so: the C-ABI declared in
// include/mdt_gpu.h, EMT constant derivation, batch buffers, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cctype>
#include <cstddef>
#include <cstdlib>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mdt_gpu.h"
#include "mdt_device.h"

size_t mdt_replica_smem(int N, int cap, bool nbr_lds);
hipError_t mdt_launch_replica(int tpa, int nrep, int block, size_t smem, const EmtConst *ec, const PropConst *pc,
                              const BatchView &bv, int add_fin, hipStream_t st);
hipError_t mdt_launch_estimators(int ntraj, const PropConst *pc, const BatchView &bv, double *out, hipStream_t st);
hipError_t mdt_launch_traj(int ntraj, size_t smem, const PropConst *pc, const BatchView &bv, const double *cjk,
                           const long long *cjk_off, const uint32_t *step_base, uint32_t step_prev, uint32_t step,
                           int flags, int block, bool classical, hipStream_t st, int t_base = 0);
hipError_t mdt_launch_advance(uint32_t *step_base, uint32_t n, int set, hipStream_t st);
size_t mdt_mixed_smem(int nmax, int pmax);
hipError_t mdt_launch_mixed(int tpa, int block, int ntraj, size_t smem, const EmtConst *ec, const PropConst *pc,
                            const BatchView &bv, int add_fin, hipStream_t st);
size_t mdt_nene_smem(int N);
size_t mdt_nene_radial_smem(int N);
int mdt_nene_cap();
hipError_t mdt_nene_attributes(hipFuncAttributes *fa, int radial);
size_t mdt_nn_frame_smem(int N);
hipError_t mdt_launch_nn_frame(int nblocks, size_t smem, const NnConst *M, const NnTrainView &tv, const int *frames,
                               int mode, double cEw, double cFw, hipStream_t st);
hipError_t mdt_launch_nn_mlp(int ntiles, const NnConst *M, const NnTrainView &tv, const int4 *tiles,
                             const int *rows, hipStream_t st);
hipError_t mdt_launch_nn_grad(int nblocks, const NnConst *M, const NnTrainView &tv, const int4 *tiles, int ntiles,
                              const int *rows, float *partials, hipStream_t st);
hipError_t mdt_launch_nn_adam(NnConst *M, const float *partials, int W, float *grad, float *am, float *av, float lr,
                              float beta1, float beta2, float eps, int step, int mode, hipStream_t st);
hipError_t mdt_launch_nene(int nrep, size_t smem, const NnConst *M, const BatchView &bv, float *G, float *C,
                           int *rep_flag, double *fpes, int maxN, int add_fin, int skip_extrap, int radial,
                           uint16_t *nbr, uint8_t *nbr_cnt, uint32_t *nbr_ref, int *nbr_key, double skin_bohr,
                           hipStream_t st, int images);

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace

// error reporting of mdt_noalloc.hip
int mdt_noalloc_fail(int code, const char *msg) { return fail(code, "%s", msg); }
int mdt_noalloc_init(int device, void **ptr);
int mdt_noalloc_check_sizes(int natoms, int nbeads);
int mdt_noalloc_compute(void *ptr, size_t lds_limit, const EmtConst *d_ec, int ntypes, int N, int P, const int32_t *idx,
                        const double *simbox, const double *isimbox, const double *r, double *epot, double *f,
                        double *dens, int32_t *status, hipStream_t st);

namespace {

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(MDT_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                        __FILE__, __LINE__);                                                \
    } while (0)

template <typename T>
int dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(MDT_EHIP, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

const double kBeta = std::pow(16.0 * M_PI / 3.0, 1.0 / 3.0) * (1.0 / std::sqrt(2.0));
const double kNneighs[3] = {12.0, 6.0, 24.0};

}  // namespace

#define MDT_NSIDE 3  // side streams for a step's independent force launches

struct mdt_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side[MDT_NSIDE] = {};
    hipEvent_t fork_ev = nullptr, join_ev[MDT_NSIDE] = {};
    bool have_params = false;
    int ntypes = 0;
    EmtConst ec{};
    EmtConst *d_ec = nullptr;  // device copy read by the kernels
    std::vector<int32_t> pes;
    mdt_batch *scratch = nullptr;  // cached single-universe batch for mdt_compute_emt / mdt_compute_nene
    void *noalloc = nullptr;       // this device's scratch array of mdt_compute_emt_noalloc
    size_t lds_limit = 0;          // LDS per workgroup of the device
    int scratch_N = -1, scratch_P = -1;
    NnConst *nn = nullptr;    // NENE surrogate (host copy)
    NnConst *d_nn = nullptr;  // device copy read by the kernels
    bool have_nn = false;
};

// Replicas whose atom count lies in (previous bound, bound] of 64, 128, ...,
// 2048 form a size class with its own launch shape (threads per atom, block,
// LDS, neighbour buffer), so that small slabs in a mixed batch do not run
// with the largest slab's shape.
// Before autotuning, size classes whose replicas have more atoms keep their
// lists across launches (the autotuner then times both ways per class)
#define SKIN_MIN_ATOMS 256

struct SizeClass {
    int nmax = 0;              // largest replica of the class
    int nrep = 0;
    int *d_reps = nullptr;     // replica ids; NULL when the class holds every replica in order
    int tpa = 0, nbcap = 0, block = 0;
    int keep = 0;              // lists kept across launches (Verlet skin): 1 staged in LDS, 2 read from HBM
    int stage = 0;             // kept lists staged in LDS during a launch
    std::vector<int> h_mem;    // its replicas (host copy)
    int nn_img = -1;           // radial network: some cell of the class needs the image search (-1: not known)
    double nn_img_rc = 0.0;    // the cut-off (A) nn_img was decided for
    int wide = 0;              // 1: workgroups of up to 1024 lanes (k_replica's 1024-lane instantiation)
    size_t smem = 0;
    uint16_t *d_nbr = nullptr; // neighbour lists in HBM when kept or when they do not fit in LDS
    long long nbr_stride = 0;
    // Verlet skin of the kept lists: positions at the build (3 nmax per block
    // slot), counts (nmax per slot), valid flags
    double *d_ref = nullptr;
    uint16_t *d_kcnt = nullptr;
    int *d_valid = nullptr;
    float tuned_ms = 0.f;
};

struct mdt_batch {
    mdt_context *ctx = nullptr;
    std::vector<SizeClass> cls;
    int ntraj = 0, nrep = 0;
    std::vector<int> N, P;
    std::vector<int> rep_t;         // trajectory of each replica
    std::vector<long long> soff, aoff, eoff;
    long long tot3pn = 0, totpn = 0, totn = 0, totp = 0;
    int maxN = 0, maxP = 0;
    bool allP1 = true;
    // device buffers
    int *d_rep_traj = nullptr, *d_rep_bead = nullptr, *d_N = nullptr, *d_P = nullptr;
    long long *d_soff = nullptr, *d_aoff = nullptr, *d_eoff = nullptr, *d_cjkoff = nullptr;
    uint8_t *d_type = nullptr, *d_fixed = nullptr;
    double *d_mass = nullptr, *d_cell = nullptr, *d_icell = nullptr, *d_tsurf = nullptr, *d_tproj = nullptr;
    uint32_t *d_tid = nullptr;
    double *d_x = nullptr, *d_v = nullptr, *d_a = nullptr, *d_f = nullptr, *d_dens = nullptr, *d_epot = nullptr;
    double *d_eta = nullptr;  // langevin_1's deviates for langevin_2 (3PN)
    double *d_cjk = nullptr;
    int *d_status = nullptr;
    size_t lds_limit = 0;
    PropConst pc{};
    PropConst *d_pc = nullptr;  // device copy read by the kernels
    bool have_atoms = false, have_cells = false, have_prop = false, have_state = false;
    uint32_t step = 0;
    int tpa = 0, nbcap = 0, block = 0;
    int nbcap_req = 0;      // requested neighbour capacity (0 = default)
    float tuned_ms = 0.f;   // force-kernel time of the autotuned choice
    size_t smem = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int last_nlaunch = 0;
    int profile = 0;                // record events around every force launch
    std::vector<hipEvent_t> evf;    // start/stop pairs, one per step
    double last_force_ms = 0.0;     // summed force-kernel time of the last call
    int precision = 0;              // mode 0: FP64; 1: FP32 pair terms where the EMT types allow, FP64 accumulation
    int graph_arith = -1;           // pair-term arithmetic the captured graph launches
    // hipGraph replay of the step loop: a chunk of graph_chunk steps is
    // captured once per launch configuration; the step counter lives on the
    // device (d_step) and the chunk's last node advances it
    int graph_chunk = 0;
    hipGraphExec_t gexec = nullptr;
    uint32_t *d_step = nullptr;
    // two independent halves of the trajectories on two streams (one size
    // class, one bead, EMT): mode -1 auto, 0 off, 1 on; their graph chunks,
    // step counters and the identity replica list their force launches index
    int split_mode = -1;
    int last_split = 0;  // the last mdt_batch_step ran as two halves
    hipGraphExec_t gexec_h[2] = {nullptr, nullptr};
    uint32_t *d_step_h = nullptr;
    int *d_ident = nullptr;
    int graph_replays = 0;          // replays of the last call
    int pes = 0;                    // 0: EMT; 1: NENE network with EMT fallback on extrapolation
    float *d_nnG = nullptr, *d_nnC = nullptr;  // per replica maxN x NN_K1 symmetry functions, dE/dG
    int *d_nnflag = nullptr;        // per replica: network extrapolated -> EMT fallback
    // radial NN kernels' neighbour lists (minimum-image cells): per replica
    // maxN x cap indices, maxN lengths, 3 maxN positions at the build, key
    uint16_t *d_nnNbr = nullptr;
    uint8_t *d_nnCnt = nullptr;
    uint32_t *d_nnRef = nullptr;
    int *d_nnKey = nullptr;
    double *d_nnF = nullptr;        // the network's forces alone (3PN blocks)
    double *d_judge = nullptr;      // per replica: sum of squared force differences, count, max
    int judge_interval = 0;         // EMT judge every judge_interval steps (0: off)
    int judge_mod = 100;            // judged replicas: rep % judge_mod == phase (100: 1 %)
    uint32_t judge_events = 0;
    // bead mode 1 (mixed): in trajectories with P > 1 the projectile (atom 0)
    // carries the beads and the other atoms are classical, kept in bead 0;
    // those trajectories run k_mixed (one workgroup each), the rest k_replica
    int bead_mode = 0;
    std::vector<uint8_t> h_type;    // atom types (0-based), host copy for validation
    std::vector<double> h_icell;    // inverse cells (9 per trajectory), host copy
    std::vector<int> mix_traj;      // trajectories on k_mixed, by atom count
    // their size classes: a run of mix_traj (offset, count) and its LDS
    struct MixClass {
        int off = 0, n = 0, nmax = 0;
        size_t smem = 0;
        int tpa = 0, block = 512;  // launch shape (tpa 0: the batch's mix_tpa)
    };
    std::vector<MixClass> mix_cls;
    int *d_mix = nullptr;
    uint16_t *d_mix_nbr = nullptr;  // classical atoms' owned half lists, one block per listed trajectory
    long long mix_stride = 0;
    // their Verlet skin: build positions and fold, counts, valid flags
    double *d_mix_ref = nullptr, *d_mix_fold = nullptr;
    uint16_t *d_mix_cnt = nullptr;
    int *d_mix_valid = nullptr;
    long long mix_ns = 0;
    double list_skin = 0.3;  // A, mdt_batch_set_list_skin
    int wide_mode = -1;      // mdt_batch_set_wide: -1 autotuned, 0 512-lane workgroups, 1 up to 1024
    int keep_mode = -1;      // mdt_batch_set_keep: -1 autotuned, else the SizeClass::keep of every class
    int mix_tpa = 0, mix_nbcap = 0, mix_nmax = 0, mix_pmax = 0;
    size_t mix_smem = 0;
    float mix_tuned_ms = 0.f;
    // trajectory observation and the exit at the projectile limits
    TrajObs *d_obs = nullptr;
    int *d_done = nullptr;
    double obs_ul = 0, obs_ll = 0, ads_start = 0, ads_end = 0;
    int obs_exit = 0;
    uint32_t obs_step0 = 0;
    int no_com = 0;  // mdt_batch_set_com_removal(b, 0)
    double *d_est = nullptr;  // MDT_NEST per trajectory

    // k_replica's pair-term arithmetic: mode 1 is full FP32 pair terms when
    // the EMT types' stiffness allows (EmtConst::fp32_ok), else only the
    // cut-off exponential in FP32
    int arith() const { return precision == 1 ? (ctx->ec.fp32_ok ? 1 : 2) : 0; }

    BatchView view() const {
        BatchView v{};
        v.ntraj = ntraj;
        v.nrep = nrep;
        v.nbcap = 0;  // per size class (launch_replicas)
        v.precision = arith();
        v.rep_traj = d_rep_traj;
        v.rep_bead = d_rep_bead;
        v.traj_N = d_N;
        v.traj_P = d_P;
        v.traj_soff = d_soff;
        v.traj_aoff = d_aoff;
        v.type = d_type;
        v.mass = d_mass;
        v.fixed = d_fixed;
        v.cell = d_cell;
        v.icell = d_icell;
        v.tsurf = d_tsurf;
        v.tproj = d_tproj;
        v.traj_id = d_tid;
        v.x = d_x;
        v.v = d_v;
        v.a = d_a;
        v.f = d_f;
        v.dens = d_dens;
        v.eta = d_eta;
        v.epot = d_epot;
        v.traj_eoff = d_eoff;
        v.status = d_status;
        v.nbr_global = nullptr;  // per size class (launch_replicas)
        v.nbr_stride = 0;
        v.mix_ref = d_mix_ref;
        v.mix_fold = d_mix_fold;
        v.mix_cnt = d_mix_cnt;
        v.mix_valid = d_mix_valid;
        v.mix_skin = list_skin;
        v.mix_ns = mix_ns;
        v.obs = d_obs;
        v.done = d_obs ? d_done : nullptr;
        v.obs_ul = obs_ul;
        v.obs_ll = obs_ll;
        v.ads_start = ads_start;
        v.ads_end = ads_end;
        v.obs_exit = obs_exit;
        v.obs_step0 = obs_step0;
        return v;
    }
};

// ---------------------------------------------------------------------------
// EMT constants: src/pes_emt_mod.f90:533-592 (2 species), :267-307 (1 species)
// ---------------------------------------------------------------------------
static int build_emt_const(int ntypes, const double *eta2, const double *n0, const double *e0,
                           const double *lambda, const double *v0, const double *kappa,
                           const double *s0, const int32_t *pes, EmtConst &ec) {
    if (ntypes < 1 || ntypes > MDT_MAXT)
        return fail(MDT_EUNSUP, "ntypes=%d outside 1..%d", ntypes, MDT_MAXT);
    std::memset(&ec, 0, sizeof ec);
    ec.ntypes = ntypes;
    std::vector<int> emt;
    for (int t = 0; t < ntypes; ++t)
        if (pes[t + ntypes * t] == MDT_PES_EMT) emt.push_back(t);
    if (emt.empty()) return fail(MDT_EINVAL, "no type has pes(t,t) == emt (pes_emt_mod.f90:218-220)");
    if (emt.size() > 2)
        return fail(MDT_EUNSUP, "%zu EMT species: compute_emt_nspecies is not implemented", emt.size());
    for (int a = 0; a < ntypes; ++a)
        for (int n = 0; n < ntypes; ++n) {
            const bool on = pes[a + ntypes * n] == MDT_PES_EMT;
            const bool ea = std::find(emt.begin(), emt.end(), a) != emt.end();
            const bool en = std::find(emt.begin(), emt.end(), n) != emt.end();
            if (on && !(ea && en))
                return fail(MDT_EUNSUP, "EMT pair (%d,%d) involves a species without pes(t,t)=emt", a + 1,
                            n + 1);
            ec.pair_on[a * MDT_MAXT + n] = on ? 1 : 0;
        }
    const double sqrt2 = std::sqrt(2.0), sqrt3 = std::sqrt(3.0);
    auto gammas = [&](int t, double rcut, double acut, double &ig1, double &ig2) {
        const double bs0 = kBeta * s0[t];
        const double kdb = kappa[t] / kBeta;
        const double rnn[3] = {bs0, bs0 * sqrt2, bs0 * sqrt3};
        double x[3], r3[3], s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < 3; ++k) x[k] = kNneighs[k] * MDT_TWELFTH / (1 + std::exp(acut * (rnn[k] - rcut)));
        for (int k = 0; k < 3; ++k) r3[k] = rnn[k] - bs0;
        for (int k = 0; k < 3; ++k) s1 += x[k] * std::exp(-eta2[t] * r3[k]);
        for (int k = 0; k < 3; ++k) s2 += x[k] * std::exp(-kdb * r3[k]);
        ig1 = 1.0 / s1;
        ig2 = 1.0 / s2;
    };
    for (int t = 0; t < ntypes; ++t) {
        ec.eta2[t] = eta2[t];
        ec.betas0[t] = kBeta * s0[t];
        ec.kappadbeta[t] = kappa[t] / kBeta;
        ec.betaeta2[t] = kBeta * eta2[t];
        ec.e0[t] = e0[t];
        ec.lambda[t] = lambda[t];
        ec.v0[t] = v0[t];
        ec.kappa[t] = kappa[t];
        ec.n0[t] = n0[t];
        ec.densc[t] = 0.5 / MDT_BOHR2ANG - ec.betaeta2[t];
        ec.owner[t] = std::find(emt.begin(), emt.end(), t) != emt.end() ? 1 : 0;
    }
    ec.fp32_ok = 1;
    for (int t : emt)
        if (6.0 * v0[t] * kappa[t] / (kBeta * eta2[t]) > MDT_FP32_STIFF_MAX) ec.fp32_ok = 0;
    double rcut;
    if (emt.size() == 1) {
        rcut = ec.betas0[emt[0]] * sqrt3;  // :295
    } else {
        rcut = std::max(ec.betas0[emt[0]], ec.betas0[emt[1]]) * sqrt3;  // :574
    }
    const double rr = 4.0 * rcut / (sqrt3 + 2);
    const double acut = 9.21034037197618 / (rr - rcut);
    ec.rcut = rcut;
    ec.acut = acut;
    ec.cutr = 1.5 * rcut;
    ec.cut2hi = ec.cutr * ec.cutr * (1.0 + 1e-12);
    // near radius: theta(r) max_t exp(-min(eta2_t, kappa_t / beta) (r - beta s0_t)) <= 2^-56
    auto weight = [&](double r) {
        double w = 0.0;
        for (int t : emt)
            w = std::max(w, std::exp(-std::min(eta2[t], kappa[t] / kBeta) * (r - kBeta * s0[t])));
        return w / (1.0 + std::exp(acut * (r - rcut)));
    };
    const double wmax = std::ldexp(1.0, -56);
    if (weight(ec.cutr) > wmax) {
        ec.rnear = ec.cutr;  // no far pairs
    } else {
        double lo = rcut, hi = ec.cutr;
        for (int it = 0; it < 100; ++it) {
            const double mid = 0.5 * (lo + hi);
            (weight(mid) > wmax ? lo : hi) = mid;
        }
        ec.rnear = hi;
    }
    if (emt.size() == 1) {
        const int e = emt[0];
        double ig1, ig2;
        gammas(e, rcut, acut, ig1, ig2);
        ec.csig[e * MDT_MAXT + e] = ig1;
        ec.etad[e * MDT_MAXT + e] = eta2[e];
        ec.wv[e * MDT_MAXT + e] = ig2 * v0[e];
    } else {
        const int t1 = emt[0], t2 = emt[1];
        double ig1_1, ig2_1, ig1_2, ig2_2;
        gammas(t1, rcut, acut, ig1_1, ig2_1);
        gammas(t2, rcut, acut, ig1_2, ig2_2);
        const double chi_21 = n0[t1] / n0[t2] * std::exp(0.5 / MDT_BOHR2ANG * (s0[t2] - s0[t1]));
        const double chi_12 = 1.0 / chi_21;
        ec.csig[t1 * MDT_MAXT + t1] = ig1_1;
        ec.csig[t1 * MDT_MAXT + t2] = chi_12 * ig1_1;
        ec.csig[t2 * MDT_MAXT + t2] = ig1_2;
        ec.csig[t2 * MDT_MAXT + t1] = chi_21 * ig1_2;
        // the type1-type1 sigma derivative uses eta2(type2) (pes_emt_mod.f90:705)
        ec.etad[t1 * MDT_MAXT + t1] = eta2[t2];
        ec.etad[t1 * MDT_MAXT + t2] = eta2[t2];
        ec.etad[t2 * MDT_MAXT + t2] = eta2[t2];
        ec.etad[t2 * MDT_MAXT + t1] = eta2[t1];
        ec.wv[t1 * MDT_MAXT + t1] = ig2_1 * v0[t1];
        ec.wv[t2 * MDT_MAXT + t2] = ig2_2 * v0[t2];
        ec.wv[t2 * MDT_MAXT + t1] = ig2_2 * v0[t2] * chi_21;
        ec.wv[t1 * MDT_MAXT + t2] = ig2_1 * v0[t1] * chi_12;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int mdt_last_error(char *buf, int buflen) {
    if (buf && buflen > 0) {
        std::strncpy(buf, g_err.c_str(), (size_t)buflen - 1);
        buf[buflen - 1] = 0;
    }
    return (int)g_err.size();
}

int mdt_version(void) { return 100; }

int mdt_device_count(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return 0;
}

int mdt_create(int device, mdt_context **out) {
    if (!out) return fail(MDT_EINVAL, "mdt_create: out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(MDT_EHIP, "mdt_create: no GPU visible (the HIP path has no CPU fallback)");
    if (device < 0 || device >= n) return fail(MDT_EINVAL, "mdt_create: device %d of %d", device, n);
    HIPCHK(hipSetDevice(device));
    mdt_context *c = new mdt_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int i = 0; i < MDT_NSIDE && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join_ev[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        mdt_destroy(c);
        return fail(MDT_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        mdt_destroy(c);
        return fail(MDT_EHIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    }
    c->lds_limit = (size_t)prop.sharedMemPerBlock;
    // the device's scratch array of mdt_compute_emt_noalloc (its address,
    // taken once per device with the device current)
    int rc = mdt_noalloc_init(device, &c->noalloc);
    if (rc) {
        mdt_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int mdt_destroy(mdt_context *ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    if (ctx->scratch) mdt_batch_destroy(ctx->scratch);
    if (ctx->d_ec) (void)hipFree(ctx->d_ec);
    if (ctx->d_nn) (void)hipFree(ctx->d_nn);
    delete ctx->nn;
    for (int i = 0; i < MDT_NSIDE; ++i) {
        if (ctx->side[i]) (void)hipStreamDestroy(ctx->side[i]);
        if (ctx->join_ev[i]) (void)hipEventDestroy(ctx->join_ev[i]);
    }
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

int mdt_emt_params_set(mdt_context *ctx, int ntypes, const double *eta2, const double *n0,
                       const double *e0, const double *lambda, const double *v0,
                       const double *kappa, const double *s0, const int32_t *pes) {
    if (!ctx || !eta2 || !n0 || !e0 || !lambda || !v0 || !kappa || !s0 || !pes)
        return fail(MDT_EINVAL, "mdt_emt_params_set: NULL argument");
    EmtConst ec;
    int rc = build_emt_const(ntypes, eta2, n0, e0, lambda, v0, kappa, s0, pes, ec);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->d_ec) {
        int rc2 = dalloc(&ctx->d_ec, 1);
        if (rc2) return rc2;
    }
    HIPCHK(hipMemcpy(ctx->d_ec, &ec, sizeof ec, hipMemcpyHostToDevice));
    ctx->ec = ec;
    ctx->ntypes = ntypes;
    ctx->pes.assign(pes, pes + ntypes * ntypes);
    ctx->have_params = true;
    return 0;
}

static void free_classes(mdt_batch *b);

int mdt_batch_destroy(mdt_batch *b) {
    if (!b) return 0;
    (void)hipSetDevice(b->ctx->device);
    void *ptrs[] = {b->d_rep_traj, b->d_rep_bead, b->d_N, b->d_P, b->d_soff, b->d_aoff, b->d_eoff,
                    b->d_cjkoff, b->d_type, b->d_fixed, b->d_mass, b->d_cell, b->d_icell, b->d_tsurf,
                    b->d_tproj, b->d_tid, b->d_x, b->d_v, b->d_a, b->d_f, b->d_dens, b->d_epot,
                    b->d_cjk, b->d_status, b->d_pc, b->d_step, b->d_nnG, b->d_nnC,
                    b->d_nnflag, b->d_nnF, b->d_judge, b->d_mix, b->d_mix_nbr, b->d_mix_ref, b->d_mix_fold, b->d_mix_cnt, b->d_mix_valid, b->d_obs, b->d_done, b->d_est, b->d_eta,
                    b->d_nnNbr, b->d_nnCnt, b->d_nnRef, b->d_nnKey, b->d_step_h, b->d_ident};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    for (hipGraphExec_t g : b->gexec_h)
        if (g) (void)hipGraphExecDestroy(g);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (hipEvent_t e : b->evf) (void)hipEventDestroy(e);
    free_classes(b);
    delete b;
    return 0;
}

static void build_cjk_host(int P, double *cjk) {  // rpmd.f90:36-59
    for (int j = 1; j <= P; ++j)
        for (int k = 0; k <= P - 1; ++k) {
            double v;
            if (k == 0)
                v = std::sqrt(1.0 / P);
            else if (1 <= k && k <= P / 2 - 1)
                v = std::sqrt(2.0 / P) * std::cos(2.0 * M_PI * j * k / P);
            else if (k == P / 2)
                v = std::sqrt(1.0 / P) * ((j % 2) ? -1.0 : 1.0);
            else
                v = std::sqrt(2.0 / P) * std::sin(2.0 * M_PI * j * k / P);
            cjk[(j - 1) + P * k] = v;
        }
}

static void free_class_lists(SizeClass &c) {
    void *ptrs[] = {c.d_nbr, c.d_ref, c.d_kcnt, c.d_valid};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c.d_nbr = nullptr;
    c.d_ref = nullptr;
    c.d_kcnt = nullptr;
    c.d_valid = nullptr;
    c.nbr_stride = 0;
}

static void free_classes(mdt_batch *b) {
    for (SizeClass &c : b->cls) {
        if (c.d_reps) (void)hipFree(c.d_reps);
        free_class_lists(c);
    }
    b->cls.clear();
}

// new positions or skin: every kept neighbour list is rebuilt at the next
// launch (after the batch's queued launches, which may still set the flags)
static int invalidate_lists(mdt_batch *b) {
    hipStream_t st = b->ctx->stream;
    HIPCHK(hipStreamSynchronize(st));
    for (SizeClass &c : b->cls)
        if (c.d_valid) HIPCHK(hipMemsetAsync(c.d_valid, 0, sizeof(int) * c.nrep, st));
    if (b->d_mix_valid) HIPCHK(hipMemsetAsync(b->d_mix_valid, 0, sizeof(int) * b->mix_traj.size(), st));
    if (b->d_nnKey) HIPCHK(hipMemsetAsync(b->d_nnKey, 0, sizeof(int) * b->nrep, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// Size classes of the replicas that k_replica computes: every replica, or in
// mixed bead mode those of the trajectories with one bead.
static int build_classes(mdt_batch *b) {
    free_classes(b);
    const int bounds[6] = {64, 128, 256, 512, 1024, 2048};
    std::vector<std::vector<int>> mem(6);
    std::vector<int> nmax(6, 0);
    int taken = 0;
    for (int r = 0; r < b->nrep; ++r) {
        const int t = b->rep_t[r];
        if (b->bead_mode == 1 && b->P[t] > 1) continue;
        const int n = b->N[t];
        int k = 0;
        while (n > bounds[k]) ++k;
        mem[k].push_back(r);
        nmax[k] = std::max(nmax[k], n);
        ++taken;
    }
    int nonempty = 0;
    for (int k = 0; k < 6; ++k) nonempty += !mem[k].empty();
    // a list is needed unless one class holds every replica in order
    const bool lists = nonempty > 1 || taken < b->nrep;
    for (int k = 0; k < 6; ++k) {
        if (mem[k].empty()) continue;
        SizeClass c;
        c.nmax = nmax[k];
        c.nrep = (int)mem[k].size();
        c.h_mem = mem[k];
        if (lists) {
            int rc = dalloc(&c.d_reps, mem[k].size());
            if (!rc && hipMemcpy(c.d_reps, mem[k].data(), sizeof(int) * mem[k].size(), hipMemcpyHostToDevice) !=
                           hipSuccess)
                rc = fail(MDT_EHIP, "size-class replica list");
            b->cls.push_back(c);
            if (rc) return rc;
        } else {
            b->cls.push_back(c);
        }
    }
    return 0;
}

int mdt_batch_create(mdt_context *ctx, int ntraj, const int32_t *natoms, const int32_t *nbeads,
                     mdt_batch **out) {
    if (!ctx || !natoms || !nbeads || !out || ntraj <= 0)
        return fail(MDT_EINVAL, "mdt_batch_create: bad arguments");
    *out = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    mdt_batch *b = new mdt_batch();
    b->ctx = ctx;
    b->ntraj = ntraj;
    b->N.resize(ntraj);
    b->P.resize(ntraj);
    b->soff.resize(ntraj);
    b->aoff.resize(ntraj);
    b->eoff.resize(ntraj);
    std::vector<int> rep_t, rep_b;
    std::vector<long long> cjkoff(ntraj);
    std::vector<double> cjk;
    for (int t = 0; t < ntraj; ++t) {
        const int N = natoms[t], P = nbeads[t];
        if (N < 1 || N > 2048 || P < 1 || P > 64) {
            delete b;
            return fail(MDT_EUNSUP, "trajectory %d: natoms=%d nbeads=%d (supported 1..2048, 1..64)", t, N, P);
        }
        b->N[t] = N;
        b->P[t] = P;
        b->soff[t] = b->tot3pn;
        b->aoff[t] = b->totn;
        b->eoff[t] = b->totp;
        b->tot3pn += 3LL * P * N;
        b->totpn += (long long)P * N;
        b->totn += N;
        b->totp += P;
        b->maxN = std::max(b->maxN, N);
        b->maxP = std::max(b->maxP, P);
        if (P != 1) b->allP1 = false;
        for (int k = 0; k < P; ++k) {
            rep_t.push_back(t);
            rep_b.push_back(k);
        }
        cjkoff[t] = (long long)cjk.size();
        std::vector<double> c(P * P);
        build_cjk_host(P, c.data());
        cjk.insert(cjk.end(), c.begin(), c.end());
    }
    b->nrep = (int)rep_t.size();
    b->rep_t = rep_t;
    if (build_classes(b)) {
        std::string msg = g_err;
        mdt_batch_destroy(b);
        return fail(MDT_EHIP, "%s", msg.c_str());
    }
    int rc = 0;
    rc |= dalloc(&b->d_rep_traj, b->nrep);
    rc |= dalloc(&b->d_rep_bead, b->nrep);
    rc |= dalloc(&b->d_N, ntraj);
    rc |= dalloc(&b->d_P, ntraj);
    rc |= dalloc(&b->d_soff, ntraj);
    rc |= dalloc(&b->d_aoff, ntraj);
    rc |= dalloc(&b->d_eoff, ntraj);
    rc |= dalloc(&b->d_cjkoff, ntraj);
    rc |= dalloc(&b->d_type, b->totn);
    rc |= dalloc(&b->d_mass, b->totn);
    rc |= dalloc(&b->d_fixed, b->totpn);
    rc |= dalloc(&b->d_cell, 9 * ntraj);
    rc |= dalloc(&b->d_icell, 9 * ntraj);
    rc |= dalloc(&b->d_tsurf, ntraj);
    rc |= dalloc(&b->d_tproj, ntraj);
    rc |= dalloc(&b->d_tid, ntraj);
    rc |= dalloc(&b->d_x, b->tot3pn);
    rc |= dalloc(&b->d_v, b->tot3pn);
    rc |= dalloc(&b->d_a, b->tot3pn);
    rc |= dalloc(&b->d_f, b->tot3pn);
    rc |= dalloc(&b->d_dens, b->totpn);
    rc |= dalloc(&b->d_eta, b->tot3pn);
    rc |= dalloc(&b->d_epot, b->totp);
    rc |= dalloc(&b->d_cjk, cjk.size());
    rc |= dalloc(&b->d_status, ntraj);
    rc |= dalloc(&b->d_pc, 1);
    if (rc) {
        std::string msg = g_err;
        mdt_batch_destroy(b);
        return fail(MDT_EHIP, "%s", msg.c_str());
    }
    hipStream_t st = ctx->stream;
    std::vector<uint32_t> tid(ntraj);
    for (int t = 0; t < ntraj; ++t) tid[t] = (uint32_t)t;
    HIPCHK(hipMemcpyAsync(b->d_rep_traj, rep_t.data(), sizeof(int) * b->nrep, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_rep_bead, rep_b.data(), sizeof(int) * b->nrep, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_N, b->N.data(), sizeof(int) * ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_P, b->P.data(), sizeof(int) * ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_soff, b->soff.data(), sizeof(long long) * ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_aoff, b->aoff.data(), sizeof(long long) * ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_eoff, b->eoff.data(), sizeof(long long) * ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_cjkoff, cjkoff.data(), sizeof(long long) * ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_cjk, cjk.data(), sizeof(double) * cjk.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_tid, tid.data(), sizeof(uint32_t) * ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(int) * ntraj, st));
    HIPCHK(hipMemsetAsync(b->d_f, 0, sizeof(double) * b->tot3pn, st));
    HIPCHK(hipMemsetAsync(b->d_a, 0, sizeof(double) * b->tot3pn, st));
    HIPCHK(hipMemsetAsync(b->d_v, 0, sizeof(double) * b->tot3pn, st));
    HIPCHK(hipMemsetAsync(b->d_dens, 0, sizeof(double) * b->totpn, st));
    HIPCHK(hipMemsetAsync(b->d_epot, 0, sizeof(double) * b->totp, st));
    HIPCHK(hipEventCreate(&b->ev0));
    HIPCHK(hipEventCreate(&b->ev1));
    HIPCHK(hipStreamSynchronize(st));
    b->pc.seed = 0;
    *out = b;
    return 0;
}

int mdt_batch_set_atoms(mdt_batch *b, const int32_t *idx, const double *mass, const int32_t *is_fixed) {
    if (!b || !idx || !mass || !is_fixed) return fail(MDT_EINVAL, "mdt_batch_set_atoms: NULL argument");
    std::vector<uint8_t> ty(b->totn), fx(b->totpn);
    for (long long k = 0; k < b->totn; ++k) {
        if (idx[k] < 1 || idx[k] > MDT_MAXT) return fail(MDT_EINVAL, "idx[%lld]=%d outside 1..%d", k, idx[k], MDT_MAXT);
        if (!(mass[k] > 0.0)) return fail(MDT_EINVAL, "mass[%lld]=%g must be positive", k, mass[k]);
        ty[k] = (uint8_t)(idx[k] - 1);
    }
    long long q = 0;
    for (int t = 0; t < b->ntraj; ++t) {
        const int N = b->N[t], P = b->P[t];
        const long long base = b->soff[t] / 3;
        for (int i = 0; i < N; ++i)
            for (int bb = 0; bb < P; ++bb) {
                uint8_t bits = 0;
                for (int c = 0; c < 3; ++c)
                    if (is_fixed[q + c + 3 * (bb + (long long)P * i)]) bits |= (uint8_t)(1u << c);
                fx[base + (long long)bb * N + i] = bits;
            }
        q += 3LL * P * N;
    }
    hipStream_t st = b->ctx->stream;
    b->h_type = ty;
    HIPCHK(hipMemcpyAsync(b->d_type, ty.data(), b->totn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_mass, mass, sizeof(double) * b->totn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_fixed, fx.data(), b->totpn, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    b->have_atoms = true;
    return 0;
}

static void drop_graph(mdt_batch *b);

int mdt_batch_set_cells(mdt_batch *b, const double *simbox, const double *isimbox) {
    if (!b || !simbox || !isimbox) return fail(MDT_EINVAL, "mdt_batch_set_cells: NULL argument");
    hipStream_t st = b->ctx->stream;
    HIPCHK(hipMemcpyAsync(b->d_cell, simbox, sizeof(double) * 9 * b->ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_icell, isimbox, sizeof(double) * 9 * b->ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    b->h_icell.assign(isimbox, isimbox + 9 * (size_t)b->ntraj);
    for (SizeClass &k : b->cls) k.nn_img = -1;
    drop_graph(b);  // captured network launches chose their kernels for the old cells
    b->have_cells = true;
    return invalidate_lists(b);  // lists hold fractional positions of the old cells
}

int mdt_batch_set_propagation(mdt_batch *b, int ntypes, const int32_t *algo, const int32_t *is_proj,
                              double dt, const double *tsurf, const double *tproj, double pile_tau,
                              double andersen_time) {
    if (!b || !algo || !is_proj || !tsurf || !tproj) return fail(MDT_EINVAL, "mdt_batch_set_propagation: NULL argument");
    if (ntypes < 1 || ntypes > MDT_MAXT) return fail(MDT_EUNSUP, "ntypes=%d", ntypes);
    PropConst &pc = b->pc;
    const uint32_t seed = pc.seed;
    std::memset(&pc, 0, sizeof pc);
    pc.seed = seed;
    pc.dt = dt;
    pc.pile_tau = pile_tau;
    pc.andersen_time = andersen_time;
    // the free ring-polymer step's frequencies scale with kB Tproj P / hbar
    // (src/rpmd.f90:354-393): P > 1 with Tproj <= 0 has none
    for (int t = 0; t < b->ntraj; ++t)
        if (b->P[t] > 1 && !(tproj[t] > 0.0))
            return fail(MDT_EINVAL, "trajectory %d: %d beads need Tproj > 0 (got %g)", t, b->P[t], tproj[t]);
    for (int t = 0; t < ntypes; ++t) {
        const int a = algo[t];
        if (a != MDT_PROP_VERLET && a != MDT_PROP_LANGEVIN && a != MDT_PROP_ANDERSEN && a != MDT_PROP_PILE)
            return fail(MDT_EUNSUP, "type %d: propagator %d unknown (md_algo.f90:66-67)", t + 1, a);
        pc.algo[t] = a;
        pc.is_proj[t] = is_proj[t] ? 1 : 0;
        if (a == MDT_PROP_LANGEVIN) pc.any_langevin = 1;
        if (a == MDT_PROP_PILE) pc.any_pile = 1;
        if (a == MDT_PROP_ANDERSEN) pc.any_andersen = 1;
    }
    hipStream_t st = b->ctx->stream;
    HIPCHK(hipMemcpyAsync(b->d_tsurf, tsurf, sizeof(double) * b->ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_tproj, tproj, sizeof(double) * b->ntraj, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    b->have_prop = true;
    return 0;
}

int mdt_batch_seed(mdt_batch *b, uint32_t seed, const uint32_t *traj_id) {
    if (!b) return fail(MDT_EINVAL, "mdt_batch_seed: NULL batch");
    b->pc.seed = seed;
    std::vector<uint32_t> tid(b->ntraj);
    for (int t = 0; t < b->ntraj; ++t) tid[t] = traj_id ? traj_id[t] : (uint32_t)t;
    HIPCHK(hipMemcpy(b->d_tid, tid.data(), sizeof(uint32_t) * b->ntraj, hipMemcpyHostToDevice));
    return 0;
}

// reference (3,P,N) per trajectory <-> device [b][c][i]
static void to_device_layout(const mdt_batch *b, const double *src, std::vector<double> &dst) {
    dst.resize(b->tot3pn);
    for (int t = 0; t < b->ntraj; ++t) {
        const int N = b->N[t], P = b->P[t];
        const long long o = b->soff[t];
        for (int i = 0; i < N; ++i)
            for (int bb = 0; bb < P; ++bb)
                for (int c = 0; c < 3; ++c)
                    dst[o + ((long long)bb * 3 + c) * N + i] = src[o + c + 3 * (bb + (long long)P * i)];
    }
}

// bead slot holding atom i's bead bb on the device: in mixed bead mode a
// classical atom (i > 0 of a trajectory with beads) lives in bead 0 and is
// reported in every bead
static inline int dev_bead(const mdt_batch *b, int t, int i, int bb) {
    return (b->bead_mode == 1 && b->P[t] > 1 && i > 0) ? 0 : bb;
}

static void from_device_layout(const mdt_batch *b, const std::vector<double> &src, double *dst) {
    for (int t = 0; t < b->ntraj; ++t) {
        const int N = b->N[t], P = b->P[t];
        const long long o = b->soff[t];
        for (int i = 0; i < N; ++i)
            for (int bb = 0; bb < P; ++bb) {
                const int sb = dev_bead(b, t, i, bb);
                for (int c = 0; c < 3; ++c)
                    dst[o + c + 3 * (bb + (long long)P * i)] = src[o + ((long long)sb * 3 + c) * N + i];
            }
    }
}

int mdt_batch_set_state(mdt_batch *b, const double *r, const double *v) {
    if (!b || !r) return fail(MDT_EINVAL, "mdt_batch_set_state: NULL argument");
    std::vector<double> tmp;
    hipStream_t st = b->ctx->stream;
    to_device_layout(b, r, tmp);
    HIPCHK(hipMemcpyAsync(b->d_x, tmp.data(), sizeof(double) * b->tot3pn, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    if (v) {
        to_device_layout(b, v, tmp);
        HIPCHK(hipMemcpyAsync(b->d_v, tmp.data(), sizeof(double) * b->tot3pn, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    int rc = invalidate_lists(b);  // new positions: the kept lists are stale
    if (rc) return rc;
    b->have_state = true;
    return 0;
}

int mdt_batch_get_state(mdt_batch *b, double *r, double *v, double *f, double *a, double *epot, double *dens) {
    if (!b) return fail(MDT_EINVAL, "mdt_batch_get_state: NULL batch");
    hipStream_t st = b->ctx->stream;
    std::vector<double> tmp(b->tot3pn);
    double *srcs[4] = {b->d_x, b->d_v, b->d_f, b->d_a};
    double *dsts[4] = {r, v, f, a};
    for (int k = 0; k < 4; ++k) {
        if (!dsts[k]) continue;
        HIPCHK(hipMemcpyAsync(tmp.data(), srcs[k], sizeof(double) * b->tot3pn, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        from_device_layout(b, tmp, dsts[k]);
    }
    if (epot) {
        HIPCHK(hipMemcpyAsync(epot, b->d_epot, sizeof(double) * b->totp, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    if (dens) {
        std::vector<double> d(b->totpn);
        HIPCHK(hipMemcpyAsync(d.data(), b->d_dens, sizeof(double) * b->totpn, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int t = 0; t < b->ntraj; ++t) {
            const int N = b->N[t], P = b->P[t];
            const long long o = b->soff[t] / 3;
            for (int i = 0; i < N; ++i)
                for (int bb = 0; bb < P; ++bb)
                    dens[o + bb + (long long)P * i] = d[o + (long long)dev_bead(b, t, i, bb) * N + i];
        }
    }
    return 0;
}

static int push_pc(mdt_batch *b) {
    HIPCHK(hipMemcpy(b->d_pc, &b->pc, sizeof(PropConst), hipMemcpyHostToDevice));
    return 0;
}

// mixed bead mode: every trajectory with beads has one projectile, atom 0;
// its other atoms are of non-projectile types
static int check_mixed(mdt_batch *b) {
    if (b->pes == 1) return fail(MDT_EUNSUP, "mixed bead mode runs with the EMT potential only");
    if (!b->have_prop) return fail(MDT_ESTATE, "mixed bead mode: projectile types unknown (call mdt_batch_set_propagation)");
    for (int t = 0; t < b->ntraj; ++t) {
        if (b->P[t] == 1) continue;
        const int N = b->N[t];
        if (N < 2 || N > 2048) return fail(MDT_EUNSUP, "mixed bead mode: trajectory %d has %d atoms", t, N);
        const uint8_t *ty = b->h_type.data() + b->aoff[t];
        if (!b->pc.is_proj[ty[0]])
            return fail(MDT_EINVAL, "mixed bead mode: atom 1 of trajectory %d is not of a projectile type", t);
        for (int i = 1; i < N; ++i)
            if (b->pc.is_proj[ty[i]])
                return fail(MDT_EUNSUP, "mixed bead mode: trajectory %d has a second projectile atom (%d)", t, i + 1);
    }
    return 0;
}

static int ensure_ready(mdt_batch *b, bool need_prop) {
    if (b->pes == 1 && !b->ctx->have_nn) return fail(MDT_ESTATE, "no network (call mdt_nn_params_set)");
    if (b->pes == 0 && !b->ctx->have_params) return fail(MDT_ESTATE, "no EMT parameters (call mdt_emt_params_set)");
    if (b->pes == 1 && b->bead_mode == MDT_BEADS_MIXED)
        return fail(MDT_EUNSUP, "mixed bead mode runs with the EMT potential only");
    if (!b->have_atoms) return fail(MDT_ESTATE, "no atoms (call mdt_batch_set_atoms)");
    if (!b->have_cells) return fail(MDT_ESTATE, "no cells (call mdt_batch_set_cells)");
    if (!b->have_state) return fail(MDT_ESTATE, "no positions (call mdt_batch_set_state)");
    if (need_prop && !b->have_prop) return fail(MDT_ESTATE, "no propagation setup (call mdt_batch_set_propagation)");
    if (!b->have_prop) {
        // calc_force alone needs no propagator data: neutral defaults
        std::vector<double> z(b->ntraj, 0.0);
        HIPCHK(hipMemcpy(b->d_tsurf, z.data(), sizeof(double) * b->ntraj, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b->d_tproj, z.data(), sizeof(double) * b->ntraj, hipMemcpyHostToDevice));
    }
    if (b->bead_mode == 1 && b->have_atoms) {
        int rc = check_mixed(b);
        if (rc) return rc;
    }
    b->pc.mixed = b->bead_mode == 1 ? 1 : 0;
    b->pc.no_com = b->no_com;
    return push_pc(b);
}

static void drop_graph(mdt_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    for (hipGraphExec_t &g : b->gexec_h) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

static int heuristic_tpa(int maxN) {
    int tpa = 8;
    while (tpa > 1 && maxN * tpa > 512) tpa >>= 1;
    return tpa;
}

// keep: lists kept across launches when the skin is positive -- 1 kept in
// HBM between launches and staged in LDS during one (where they fit, else
// as 2), 2 kept in HBM and read from there -- or 0 rebuilt every launch (in
// LDS where they fit); -1 the default before tuning (1 above SKIN_MIN_ATOMS
// atoms).  wide: 1 lets a class of more than 512
// atoms x threads per atom (one or two threads per atom) run on one
// workgroup of up to 1024 lanes instead of looping 512 lanes over its atoms;
// -1 follows mdt_batch_set_wide (1024 lanes only for mode 1).
static int configure_class(mdt_batch *b, SizeClass &c, int tpa, int nbcap, int keep = -1, int wide = -1) {
    if (wide < 0) wide = b->wide_mode == 1 ? 1 : 0;  // untuned: 1024 lanes only when asked for
    // capacity of the owned half lists: every other atom for small replicas,
    // else 64.  The lists hold the pairs within the near radius (+ skin):
    // fcc Au or Pt within it + 1 A has 86 neighbours, of which an atom owns
    // about half; overflow is detected and reported.
    int cap = nbcap > 0 ? nbcap : 64;
    if (cap > c.nmax - 1) cap = std::max(1, c.nmax - 1);
    if (wide && tpa > 2) tpa = 2;
    const int lanes = wide ? 1024 : 512;
    while (tpa > 1 && c.nmax * tpa > lanes) tpa >>= 1;
    if (c.nmax * tpa <= 512) wide = 0;  // fits 512 lanes: the 512-lane instantiation
    // lists kept in HBM across launches (Verlet skin) unless the skin is 0
    if (b->keep_mode >= 0) keep = b->keep_mode;
    if (keep < 0) keep = c.nmax > SKIN_MIN_ATOMS ? 1 : 0;
    const bool skin = b->list_skin > 0.0 && keep >= 1;
    // dynamic LDS beside the kernel's static tables (4.4 KB, 4.7 KB in the judge's variant; 5.5 KiB allowed)
    const size_t lds_dyn = b->lds_limit > 5632 ? b->lds_limit - 5632 : 0;
    const size_t smem_lds = mdt_replica_smem(c.nmax, cap, true);
    const bool in_lds = (!skin || keep == 1) && smem_lds <= lds_dyn;
    const size_t smem = in_lds ? smem_lds : mdt_replica_smem(c.nmax, cap, false);
    if (smem > lds_dyn)
        return fail(MDT_EUNSUP, "replica of %d atoms needs %zu bytes of LDS (limit %zu)", c.nmax, smem, lds_dyn);
    free_class_lists(c);
    if (!in_lds || skin) {
        c.nbr_stride = (long long)c.nmax * cap;
        int rc = dalloc(&c.d_nbr, (size_t)c.nbr_stride * c.nrep);
        if (rc) return rc;
    }
    if (skin) {
        int rc = dalloc(&c.d_ref, 3 * (size_t)c.nmax * c.nrep) | dalloc(&c.d_kcnt, (size_t)c.nmax * c.nrep) |
                 dalloc(&c.d_valid, (size_t)c.nrep);
        if (rc) return rc;
        HIPCHK(hipMemset(c.d_valid, 0, sizeof(int) * c.nrep));
    }
    const int rounded = ((c.nmax * tpa + 63) / 64) * 64;
    c.block = std::min(wide ? 1024 : 512, std::max(64, rounded));
    c.stage = skin && in_lds ? 1 : 0;
    c.keep = skin ? (c.stage ? 1 : 2) : 0;
    c.wide = wide;
    c.tpa = tpa;
    c.nbcap = cap;
    c.smem = smem;
    return 0;
}

static int lds_limit(mdt_batch *b) {
    if (b->lds_limit == 0) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, b->ctx->device));
        b->lds_limit = (size_t)prop.sharedMemPerBlock;
    }
    return 0;
}

// the reported choice: the threads per atom of the class with most replicas
static void note_tpa(mdt_batch *b) {
    if (b->cls.empty()) {  // mixed bead mode, every trajectory on k_mixed
        b->tpa = b->mix_tpa;
        b->block = 512;
        b->smem = b->mix_smem;
        b->nbcap = b->mix_nbcap;
        return;
    }
    int best = 0;
    for (size_t k = 1; k < b->cls.size(); ++k)
        if (b->cls[k].nrep > b->cls[best].nrep) best = (int)k;
    b->tpa = b->cls[best].tpa;
    b->block = b->cls[best].block;
    b->smem = b->cls[best].smem;
    b->nbcap = b->cls[best].nbcap;
}

// tpa > 0: that many threads per atom in every class (halved where a class's
// replicas would exceed 512 threads); tpa == 0: per-class heuristic
// k_mixed's launch: trajectory list, the classical atoms' list buffer and
// LDS (tpa: lanes per classical atom in the pair passes, 2, 4 or 8)
static int configure_mixed(mdt_batch *b, int tpa, int nbcap) {
    if (b->d_mix) (void)hipFree(b->d_mix);
    if (b->d_mix_nbr) (void)hipFree(b->d_mix_nbr);
    if (b->d_mix_ref) (void)hipFree(b->d_mix_ref);
    if (b->d_mix_fold) (void)hipFree(b->d_mix_fold);
    if (b->d_mix_cnt) (void)hipFree(b->d_mix_cnt);
    if (b->d_mix_valid) (void)hipFree(b->d_mix_valid);
    b->d_mix = nullptr;
    b->d_mix_nbr = nullptr;
    b->d_mix_ref = b->d_mix_fold = nullptr;
    b->d_mix_cnt = nullptr;
    b->d_mix_valid = nullptr;
    b->mix_traj.clear();
    b->mix_cls.clear();
    b->mix_nmax = b->mix_pmax = 0;
    if (b->bead_mode != 1) return 0;
    for (int t = 0; t < b->ntraj; ++t)
        if (b->P[t] > 1) {
            b->mix_traj.push_back(t);
            b->mix_nmax = std::max(b->mix_nmax, b->N[t]);
            b->mix_pmax = std::max(b->mix_pmax, b->P[t]);
        }
    if (b->mix_traj.empty()) return 0;
    // size classes (bounds 64, 128, ..., 2048 atoms), each launched with
    // LDS for its own largest trajectory, so small slabs are not held to
    // the occupancy of the largest
    std::stable_sort(b->mix_traj.begin(), b->mix_traj.end(), [b](int x, int y) { return b->N[x] < b->N[y]; });
    const int bounds[6] = {64, 128, 256, 512, 1024, 2048};
    size_t smem = 0;
    for (int i = 0; i < (int)b->mix_traj.size();) {
        int k = 0;
        while (b->N[b->mix_traj[i]] > bounds[k]) ++k;
        mdt_batch::MixClass mc;
        mc.off = i;
        int nmax = 0, pmax = 0;
        while (i < (int)b->mix_traj.size() && b->N[b->mix_traj[i]] <= bounds[k]) {
            nmax = std::max(nmax, b->N[b->mix_traj[i]]);
            pmax = std::max(pmax, b->P[b->mix_traj[i]]);
            ++i;
        }
        mc.n = i - mc.off;
        mc.nmax = nmax;
        mc.smem = mdt_mixed_smem(nmax, pmax);
        if (mc.smem + 4096 > b->lds_limit)  // 4 KiB: the kernel's static tables
            return fail(MDT_EUNSUP, "mixed bead mode: %d atoms and %d beads need %zu bytes of LDS (limit %zu)",
                        nmax, pmax, mc.smem + 4096, b->lds_limit);
        smem = std::max(smem, mc.smem);
        b->mix_cls.push_back(mc);
    }
    const int ns = b->mix_nmax - 1;
    int cap = nbcap > 0 ? nbcap : 96;
    if (cap > ns - 1) cap = std::max(1, ns - 1);
    b->mix_stride = (long long)ns * cap;
    const size_t nm = b->mix_traj.size();
    b->mix_ns = ns;
    int rc = dalloc(&b->d_mix, nm) | dalloc(&b->d_mix_nbr, (size_t)b->mix_stride * nm) |
             dalloc(&b->d_mix_ref, 3 * (size_t)ns * nm) | dalloc(&b->d_mix_fold, 3 * (size_t)ns * nm) |
             dalloc(&b->d_mix_cnt, 2 * (size_t)ns * nm) | dalloc(&b->d_mix_valid, nm);
    if (rc) return rc;
    HIPCHK(hipMemset(b->d_mix_valid, 0, sizeof(int) * nm));
    HIPCHK(hipMemcpy(b->d_mix, b->mix_traj.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
    b->mix_tpa = (tpa == 2 || tpa == 4 || tpa == 8) ? tpa : (tpa == 1 ? 2 : 4);
    b->mix_nbcap = cap;
    b->mix_smem = smem;
    return 0;
}

static int configure(mdt_batch *b, int tpa, int nbcap) {
    drop_graph(b);  // launch shapes and the neighbour buffers may change
    int rc = lds_limit(b);
    if (rc) return rc;
    for (SizeClass &c : b->cls)
        if ((rc = configure_class(b, c, tpa > 0 ? tpa : heuristic_tpa(c.nmax), nbcap))) return rc;
    if ((rc = configure_mixed(b, tpa, nbcap))) return rc;
    note_tpa(b);
    return 0;
}

// A step's force launches over disjoint replicas (the size classes, the
// mixed-model kernel) go round-robin to the context's side streams, forked
// from the batch's stream by an event and joined back to it, so that a small
// or latency-bound launch overlaps the others.  Inside graph capture, and
// for a single launch, every launch stays on the batch's stream.
struct Fork {
    mdt_context *c = nullptr;
    hipStream_t st = nullptr;
    bool on = false;
    unsigned used = 0;
    int next = 0;
};

static int fork_open(Fork &f, mdt_context *c, hipStream_t st, int nlaunch) {
    f = Fork{};
    f.c = c;
    f.st = st;
    if (nlaunch < 2) return 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return 0;
    HIPCHK(hipEventRecord(c->fork_ev, st));
    f.on = true;
    return 0;
}

static int fork_stream(Fork &f, hipStream_t *out) {
    *out = f.st;
    if (!f.on) return 0;
    const int i = f.next++ % MDT_NSIDE;
    if (!(f.used & (1u << i))) {
        HIPCHK(hipStreamWaitEvent(f.c->side[i], f.c->fork_ev, 0));
        f.used |= 1u << i;
    }
    *out = f.c->side[i];
    return 0;
}

static int fork_close(Fork &f) {
    if (!f.on) return 0;
    for (int i = 0; i < MDT_NSIDE; ++i)
        if (f.used & (1u << i)) {
            HIPCHK(hipEventRecord(f.c->join_ev[i], f.c->side[i]));
            HIPCHK(hipStreamWaitEvent(f.st, f.c->join_ev[i], 0));
        }
    f.on = false;
    return 0;
}

// one launch per size class of the mixed-model trajectories; a class's
// block slots start at its offset in the trajectory list and in every
// per-slot buffer (lists, kept-list state).  With fk, each class takes the
// next side stream; without, all go to st.
static int launch_mixed_class(mdt_batch *b, const BatchView &base, const mdt_batch::MixClass &mc, int tpa,
                              int block, int add_fin, hipStream_t st) {
    BatchView v = base;
    const long long off = mc.off;
    v.rep_list = b->d_mix + off;
    v.nbr_global = b->d_mix_nbr + off * b->mix_stride;
    v.nbr_stride = b->mix_stride;
    v.nbcap = b->mix_nbcap;
    v.mix_ref = b->d_mix_ref + off * 3 * b->mix_ns;
    v.mix_fold = b->d_mix_fold + off * 3 * b->mix_ns;
    v.mix_cnt = b->d_mix_cnt + off * 2 * b->mix_ns;  // near and far counts per slot
    v.mix_valid = b->d_mix_valid + off;
    HIPCHK(mdt_launch_mixed(tpa, block, mc.n, mc.smem, b->ctx->d_ec, b->d_pc, v, add_fin, st));
    return 0;
}

// tpa > 0 overrides every class's tuned threads per atom
static int launch_mixed(mdt_batch *b, const BatchView &base, int tpa, int add_fin, hipStream_t st, Fork *fk = nullptr) {
    if (!b->ctx->d_ec) return fail(MDT_ESTATE, "no EMT parameters (call mdt_emt_params_set)");
    for (const mdt_batch::MixClass &mc : b->mix_cls) {
        hipStream_t s2 = st;
        if (fk) {
            int rc = fork_stream(*fk, &s2);
            if (rc) return rc;
        }
        // the class's tuned shape, else the batch's threads per atom on 512
        // lanes (1024 where mdt_batch_set_wide asks for it and the class is large)
        const bool own = tpa <= 0 && mc.tpa > 0;
        const int ct = own ? mc.tpa : (tpa > 0 ? tpa : b->mix_tpa);
        const int cb = own ? mc.block : (b->wide_mode == 1 && (mc.nmax - 1) * ct > 512 ? 1024 : 512);
        int rc = launch_mixed_class(b, base, mc, ct, cb, add_fin, s2);
        if (rc) return rc;
    }
    return 0;
}

// calc_force on every replica, one launch per size class
static int launch_replicas(mdt_batch *b, const BatchView &base, int add_fin, hipStream_t main_st) {
    mdt_context *c = b->ctx;
    if (!c->d_ec) return fail(MDT_ESTATE, "no EMT parameters (call mdt_emt_params_set)");
    Fork fk;
    int rc = fork_open(fk, c, main_st, (int)(b->cls.size() + b->mix_cls.size()));
    if (rc) return rc;
    hipStream_t st = main_st;
    for (const SizeClass &k : b->cls) {
        if ((rc = fork_stream(fk, &st))) return rc;
        BatchView v = base;
        v.rep_list = k.d_reps;
        v.nbr_global = k.d_nbr;
        v.nbr_stride = k.nbr_stride;
        v.nbcap = k.nbcap;
        v.rl_ref = k.d_ref;
        v.rl_cnt = k.d_kcnt;
        v.rl_valid = k.d_valid;
        v.rl_skin = b->list_skin;
        v.rl_ns = k.nmax;
        v.rl_stage = k.stage;
        HIPCHK(mdt_launch_replica(k.tpa, k.nrep, k.block, k.smem, c->d_ec, b->d_pc, v, add_fin, st));
    }
    if (!b->mix_traj.empty() && (rc = launch_mixed(b, base, 0, add_fin, main_st, &fk))) return rc;
    return fork_close(fk);
}

// radial-only networks run k_nene_rg / k_nene_rf
static bool nn_radial(const mdt_context *c) { return c->nn && c->nn->radial_only; }

static size_t nn_smem(const mdt_context *c, int N) {
    return nn_radial(c) ? mdt_nene_radial_smem(N) : mdt_nene_smem(N);
}

// LDS check for the context's current network (every call: the network may
// change between calls), buffers at the first call
static int ensure_nn(mdt_batch *b) {
    if (b->lds_limit == 0) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, b->ctx->device));
        b->lds_limit = (size_t)prop.sharedMemPerBlock;
    }
    hipFuncAttributes fa;
    HIPCHK(mdt_nene_attributes(&fa, nn_radial(b->ctx)));
    if (fa.sharedSizeBytes + nn_smem(b->ctx, b->maxN) > b->lds_limit)
        return fail(MDT_EUNSUP, "network on %d atoms needs %zu bytes of LDS (limit %zu)", b->maxN,
                    fa.sharedSizeBytes + nn_smem(b->ctx, b->maxN), b->lds_limit);
    if (nn_radial(b->ctx) && !b->d_nnNbr) {
        const size_t na = (size_t)b->nrep * b->maxN;
        int rc = dalloc(&b->d_nnNbr, na * mdt_nene_cap()) | dalloc(&b->d_nnCnt, na) | dalloc(&b->d_nnRef, 3 * na) |
                 dalloc(&b->d_nnKey, (size_t)b->nrep);
        if (rc) return rc;
        HIPCHK(hipMemset(b->d_nnKey, 0, sizeof(int) * b->nrep));
    }
    if (b->d_nnG) return 0;
    const size_t rows = (size_t)b->nrep * b->maxN * NN_K1;
    int rc = dalloc(&b->d_nnG, rows) | dalloc(&b->d_nnC, rows) | dalloc(&b->d_nnflag, (size_t)b->nrep) |
             dalloc(&b->d_nnF, (size_t)b->tot3pn) | dalloc(&b->d_judge, 3 * (size_t)b->nrep);
    if (rc) return rc;
    HIPCHK(hipMemset(b->d_nnflag, 0, sizeof(int) * b->nrep));
    HIPCHK(hipMemset(b->d_judge, 0, sizeof(double) * 3 * b->nrep));
    return 0;
}

// The radial network's kernels compile the image search in only for size
// classes with a cell narrower than about 2 rc (where an atom can meet
// another in more than one image): the kernels decide per replica with
// w = rc |row of H^-1| < 0.4999; the host takes 0.499 as the bound.
static int nn_images(mdt_batch *b, SizeClass &k) {
    const mdt_context *c = b->ctx;
    if (!nn_radial(c)) return 1;
    const double rc = (double)c->nn->rc / (double)c->nn->ang2bohr;
    if (k.nn_img >= 0 && k.nn_img_rc == rc) return k.nn_img;
    int img = b->h_icell.size() == 9 * (size_t)b->ntraj ? 0 : 1;
    for (int r = 0; r < (int)k.h_mem.size() && !img; ++r) {
        const double *ih = b->h_icell.data() + 9 * (size_t)b->rep_t[k.h_mem[r]];
        for (int q = 0; q < 3; ++q)
            if (rc * std::sqrt(ih[q] * ih[q] + ih[q + 3] * ih[q + 3] + ih[q + 6] * ih[q + 6]) >= 0.499) img = 1;
    }
    k.nn_img = img;
    k.nn_img_rc = rc;
    return img;
}

// calc_force's PES dispatch (src/force.f90:77-82) for every replica: EMT, or
// the network followed by EMT on the replicas whose symmetry functions left
// the training range (when EMT parameters are set).
static int launch_pes(mdt_batch *b, const BatchView &bv, int add_fin, hipStream_t st) {
    mdt_context *c = b->ctx;
    if (b->pes == 1) {
        // one launch per size class: each sizes its LDS for its largest
        // replica, so small slabs keep the occupancy of a uniform batch
        Fork fk;
        int rc = fork_open(fk, c, st, (int)b->cls.size());
        if (rc) return rc;
        for (SizeClass &k : b->cls) {
            BatchView v = bv;
            v.rep_list = k.d_reps;
            hipStream_t s2 = st;
            if ((rc = fork_stream(fk, &s2))) return rc;
            HIPCHK(mdt_launch_nene(k.nrep, nn_smem(c, k.nmax), c->d_nn, v, b->d_nnG, b->d_nnC, b->d_nnflag,
                                   b->d_nnF, b->maxN, add_fin, 1, nn_radial(c), b->d_nnNbr, b->d_nnCnt, b->d_nnRef,
                                   b->d_nnKey, b->list_skin / MDT_BOHR2ANG, s2, nn_images(b, k)));
        }
        if ((rc = fork_close(fk))) return rc;
        if (c->have_params) {
            BatchView fb = bv;
            fb.only = b->d_nnflag;
            return launch_replicas(b, fb, add_fin, st);
        }
        return 0;
    }
    return launch_replicas(b, bv, add_fin, st);
}

// EMT as the judge of the network on a rotating 100 / judge_mod % of the
// replicas
static int launch_judge(mdt_batch *b, hipStream_t st) {
    BatchView jb = b->view();
    jb.judge_f = b->d_nnF;
    jb.skip = b->d_nnflag;  // replicas on the fallback have no network forces
    jb.judge_acc = b->d_judge;
    jb.judge_mod = b->judge_mod;
    jb.judge_phase = (int)(b->judge_events % (uint32_t)b->judge_mod);
    ++b->judge_events;
    return launch_replicas(b, jb, 0, st);
}

// LDS of k_traj: the P x P normal-mode matrix, three atom lists (uint16)
// and one flag byte per atom
// k_traj's workgroup: the largest trajectory's atoms rounded up to whole
// waves, at most 256 threads (every loop of k_traj strides by blockDim.x)
static int traj_block(const mdt_batch *b) {
    const int w = ((b->maxN + 63) / 64) * 64;
    return w < 64 ? 64 : (w > 256 ? 256 : w);
}

// every trajectory classical (P = 1) with one atom per lane and no PILE or
// Andersen atoms: the propagators run in k_traj_classical
static bool traj_classical(const mdt_batch *b) {
    return b->allP1 && b->maxN <= traj_block(b) && !b->pc.any_pile && !b->pc.any_andersen;
}

static size_t traj_smem(const mdt_batch *b) {
    return sizeof(double) * b->maxP * b->maxP + 7 * (size_t)b->maxN + 16;
}

static int launch_force(mdt_batch *b) {
    mdt_context *c = b->ctx;
    const BatchView bv = b->view();
    if (!b->allP1) {
        const size_t ts = traj_smem(b);
        HIPCHK(mdt_launch_traj(b->ntraj, ts, b->d_pc, bv, b->d_cjk, b->d_cjkoff, nullptr, b->step, b->step, 4,
                               traj_block(b), false, c->stream));
    }
    return launch_pes(b, bv, b->allP1 ? 0 : 1, c->stream);
}

static int check_status(mdt_batch *b) {
    std::vector<int> s(b->ntraj);
    HIPCHK(hipMemcpy(s.data(), b->d_status, sizeof(int) * b->ntraj, hipMemcpyDeviceToHost));
    // network extrapolation is handled by the EMT fallback when EMT is set
    const int ignore = (b->pes == 1 && b->ctx->have_params) ? MDT_ST_NN_EXTRAPOLATION : 0;
    for (int t = 0; t < b->ntraj; ++t)
        if (s[t] & ~ignore)
            return fail(MDT_ETRAJ, "trajectory %d flagged status %d (1: neighbour overflow, 2: negative density, 4: bad propagator "
                        "or type, 8: fixed-point overflow, 16: network extrapolation without EMT fallback)",
                        t, s[t]);
    return 0;
}

// Times the force kernel of every size class for every threads-per-atom
// choice that fits a 512-thread block (3 launches each after one warm-up)
// and keeps each class's fastest.  f, a, epot and dens are saved and
// restored around the trial launches.
static int autotune(mdt_batch *b) {
    mdt_context *c = b->ctx;
    hipStream_t st = c->stream;
    int rc = configure(b, 0, b->nbcap_req);
    if (rc) return rc;
    // the EMT kernels read the context's EMT constants: without them (a
    // network-only potential) there is nothing to time, and the heuristic
    // launch shape of configure() stands
    if (!c->have_params || !c->d_ec) return 0;
    double *sf = nullptr, *sa = nullptr, *se = nullptr, *sd = nullptr;
    rc = dalloc(&sf, b->tot3pn) | dalloc(&sa, b->tot3pn) | dalloc(&se, b->totp) | dalloc(&sd, b->totpn);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(sf, b->d_f, sizeof(double) * b->tot3pn, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sa, b->d_a, sizeof(double) * b->tot3pn, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(se, b->d_epot, sizeof(double) * b->totp, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sd, b->d_dens, sizeof(double) * b->totpn, hipMemcpyDeviceToDevice, st));
    // per class: threads per atom x (lists rebuilt every launch, kept with
    // the Verlet skin and staged in LDS, or kept and read from HBM) x (for more than 512 atoms x threads per
    // atom: 512 lanes looping over the atoms, or up to 1024 lanes).  A kept configuration is timed over one
    // build and four reuses -- about the rebuild rate of a 5 eV H at dt 0.1
    // fs with the default skin of 0.3 A -- so its figure is not a reuse-only
    // best case.
    const int cands[4] = {1, 2, 4, 8};
    const int ntime = 5;
    float total = 0.f;
    for (SizeClass &k : b->cls) {
        int best = k.tpa, best_keep = k.keep, best_wide = k.wide;
        float best_ms = 1e30f;
        const int kfirst = b->keep_mode >= 0 ? b->keep_mode : 0;
        const int klast = b->keep_mode >= 0 ? b->keep_mode : (b->list_skin > 0.0 ? 2 : 0);
        for (int keep = kfirst; keep <= klast; ++keep)
            for (int tpa : cands)
                for (int wide = b->wide_mode == 1 ? 1 : 0; wide <= (b->wide_mode == 0 ? 0 : 1); ++wide) {
                    if (!wide && k.nmax * tpa > 512 && tpa != 1) continue;
                    if (wide && (tpa > 2 || k.nmax * tpa <= 512 || k.nmax * tpa > 1024)) continue;
                    if (!wide && b->wide_mode == 1 && tpa <= 2 && k.nmax * tpa > 512 && k.nmax * tpa <= 1024)
                        continue;
                    if ((rc = configure_class(b, k, tpa, b->nbcap_req, keep, wide))) return rc;
                    BatchView v = b->view();
                    v.rep_list = k.d_reps;
                    v.nbr_global = k.d_nbr;
                    v.nbr_stride = k.nbr_stride;
                    v.nbcap = k.nbcap;
                    v.rl_ref = k.d_ref;
                    v.rl_cnt = k.d_kcnt;
                    v.rl_valid = k.d_valid;
                    v.rl_skin = b->list_skin;
                    v.rl_ns = k.nmax;
                    v.rl_stage = k.stage;
                    HIPCHK(mdt_launch_replica(tpa, k.nrep, k.block, k.smem, c->d_ec, b->d_pc, v, 0, st));
                    if (k.keep) HIPCHK(hipMemsetAsync(k.d_valid, 0, sizeof(int) * k.nrep, st));
                    HIPCHK(hipEventRecord(b->ev0, st));
                    for (int q = 0; q < ntime; ++q)
                        HIPCHK(mdt_launch_replica(tpa, k.nrep, k.block, k.smem, c->d_ec, b->d_pc, v, 0, st));
                    HIPCHK(hipEventRecord(b->ev1, st));
                    HIPCHK(hipEventSynchronize(b->ev1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = tpa;
                        best_keep = k.keep;
                        best_wide = k.wide;
                    }
                }
        if ((rc = configure_class(b, k, best, b->nbcap_req, best_keep, best_wide))) return rc;
        k.tuned_ms = best_ms / ntime;
        total += k.tuned_ms;
    }
    if (!b->mix_traj.empty()) {
        // per mixed-model class: threads per atom x (512 lanes, or 1024 when
        // the classical atoms x threads per atom exceed 512)
        b->mix_tuned_ms = 0.f;
        int most = -1;
        for (mdt_batch::MixClass &mc : b->mix_cls) {
            int best = b->mix_tpa, best_block = 512;
            float best_ms = 1e30f;
            for (int tpa : {2, 4, 8})
                for (int block : {512, 1024}) {
                    const bool big = (mc.nmax - 1) * tpa > 512;
                    if (block == 1024 && (!big || b->wide_mode == 0)) continue;
                    if (block == 512 && big && b->wide_mode == 1) continue;
                    if ((rc = launch_mixed_class(b, b->view(), mc, tpa, block, 0, st))) return rc;
                    HIPCHK(hipEventRecord(b->ev0, st));
                    for (int q = 0; q < 3; ++q)
                        if ((rc = launch_mixed_class(b, b->view(), mc, tpa, block, 0, st))) return rc;
                    HIPCHK(hipEventRecord(b->ev1, st));
                    HIPCHK(hipEventSynchronize(b->ev1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = tpa;
                        best_block = block;
                    }
                }
            mc.tpa = best;
            mc.block = best_block;
            b->mix_tuned_ms += best_ms / 3;
            if (most < 0 || mc.n > b->mix_cls[most].n) most = (int)(&mc - b->mix_cls.data());
        }
        b->mix_tpa = b->mix_cls[most].tpa;  // the reported choice: the class with most trajectories
        total += b->mix_tuned_ms;
    }
    HIPCHK(hipMemcpyAsync(b->d_f, sf, sizeof(double) * b->tot3pn, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_a, sa, sizeof(double) * b->tot3pn, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_epot, se, sizeof(double) * b->totp, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_dens, sd, sizeof(double) * b->totpn, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(int) * b->ntraj, st));
    (void)hipFree(sf);
    (void)hipFree(sa);
    (void)hipFree(se);
    (void)hipFree(sd);
    b->tuned_ms = total;
    drop_graph(b);
    note_tpa(b);
    return 0;
}

static int autotune_if_needed(mdt_batch *b) {
    if (b->tpa) return 0;
    return configure(b, 0, b->nbcap_req);
}

int mdt_batch_compute_forces(mdt_batch *b) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    if (!b->allP1 && !b->have_prop)
        return fail(MDT_ESTATE, "beads > 1: calc_force's ring-polymer step needs Tproj (call mdt_batch_set_propagation)");
    int rc = ensure_ready(b, false);
    if (rc) return rc;
    if ((rc = autotune_if_needed(b))) return rc;
    if (b->pes == 1 && (rc = ensure_nn(b))) return rc;
    if ((rc = launch_force(b))) return rc;
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return check_status(b);
}

// Launches of `nsteps` steps on the batch's stream.  One step: [propagate_2
// of the previous step +] propagate_1 (+ ring-polymer step when P > 1) per
// trajectory, then calc_force per replica; a closing propagate_2 ends the
// sequence so that the host sees the reference's state.  step_base != NULL
// makes the step numbers offsets from the device counter (graph capture).
static int enqueue_steps(mdt_batch *b, int nsteps, const uint32_t *step_base, uint32_t first, bool events) {
    mdt_context *c = b->ctx;
    const BatchView bv = b->view();
    const size_t ts = traj_smem(b);
    const bool classical = traj_classical(b);
    const int rp = b->allP1 ? 0 : 4;
    for (int s = 0; s < nsteps; ++s) {
        const uint32_t prev = first + s, st = first + s + 1;
        const int flags = (s > 0 ? 1 : 0) | 2 | rp;
        HIPCHK(mdt_launch_traj(b->ntraj, ts, b->d_pc, bv, b->d_cjk, b->d_cjkoff, step_base, prev, st, flags,
                               traj_block(b), classical, c->stream));
        if (events) HIPCHK(hipEventRecord(b->evf[2 * s], c->stream));
        int rc = launch_pes(b, bv, b->allP1 ? 0 : 1, c->stream);
        if (rc) return rc;
        if (events) HIPCHK(hipEventRecord(b->evf[2 * s + 1], c->stream));
        if (b->pes == 1 && b->judge_interval > 0 && c->have_params && (st % (uint32_t)b->judge_interval) == 0 &&
            !step_base && (rc = launch_judge(b, c->stream)))
            return rc;
    }
    if (nsteps > 0)
        HIPCHK(mdt_launch_traj(b->ntraj, ts, b->d_pc, bv, b->d_cjk, b->d_cjkoff, step_base, first + nsteps,
                               first + nsteps, 1, traj_block(b), classical, c->stream));
    return 0;
}

// One-bead, one-class EMT batches step as two independent halves of the
// trajectories, each on its own stream (the context's and a side stream):
// one half's propagator launches overlap the other half's force launches,
// which alone fill the GPU less well.  Every trajectory runs the same
// kernels with the same step numbers, so the results are bitwise those of
// the whole batch.  Not while profiling (events around every force launch).
static bool use_split(const mdt_batch *b) {
    if (b->split_mode == 0 || b->profile) return false;
    if (!(b->allP1 && traj_classical(b) && b->pes == 0 && b->cls.size() == 1 && b->mix_traj.empty() &&
          b->cls[0].d_reps == nullptr && b->ntraj >= 2))
        return false;
    return b->split_mode == 1 || b->ntraj >= 1024;
}

// The steps of half h (trajectories [t0, t1)) on stream st: enqueue_steps
// for one size class with the replica list and the kept-list slots offset
// to the half.
static int enqueue_half(mdt_batch *b, int h, int nsteps, const uint32_t *step_base, uint32_t first, hipStream_t st) {
    mdt_context *c = b->ctx;
    if (!c->d_ec) return fail(MDT_ESTATE, "no EMT parameters (call mdt_emt_params_set)");
    const int t0 = h ? b->ntraj / 2 : 0, t1 = h ? b->ntraj : b->ntraj / 2;
    const BatchView bv = b->view();
    const SizeClass &k = b->cls[0];
    BatchView v = bv;
    v.rep_list = b->d_ident + t0;
    v.nbr_global = k.d_nbr ? k.d_nbr + (size_t)t0 * k.nbr_stride : nullptr;
    v.nbr_stride = k.nbr_stride;
    v.nbcap = k.nbcap;
    v.rl_ref = k.d_ref ? k.d_ref + (size_t)t0 * 3 * k.nmax : nullptr;
    v.rl_cnt = k.d_kcnt ? k.d_kcnt + (size_t)t0 * k.nmax : nullptr;
    v.rl_valid = k.d_valid ? k.d_valid + t0 : nullptr;
    v.rl_skin = b->list_skin;
    v.rl_ns = k.nmax;
    v.rl_stage = k.stage;
    for (int s = 0; s < nsteps; ++s) {
        const uint32_t prev = first + s, stp = first + s + 1;
        const int flags = (s > 0 ? 1 : 0) | 2;
        HIPCHK(mdt_launch_traj(t1 - t0, 0, b->d_pc, bv, b->d_cjk, b->d_cjkoff, step_base, prev, stp, flags,
                               traj_block(b), true, st, t0));
        HIPCHK(mdt_launch_replica(k.tpa, t1 - t0, k.block, k.smem, c->d_ec, b->d_pc, v, 0, st));
    }
    if (nsteps > 0)
        HIPCHK(mdt_launch_traj(t1 - t0, 0, b->d_pc, bv, b->d_cjk, b->d_cjkoff, step_base, first + nsteps,
                               first + nsteps, 1, traj_block(b), true, st, t0));
    return 0;
}

// Captures half h's chunk of graph_chunk steps (step numbers relative to
// d_step_h[h], which the chunk's last node advances) and instantiates it.
static int capture_half(mdt_batch *b, int h) {
    mdt_context *c = b->ctx;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_half(b, h, b->graph_chunk, b->d_step_h + h, 0u, c->stream);
    hipError_t e = rc ? hipSuccess : mdt_launch_advance(b->d_step_h + h, (uint32_t)b->graph_chunk, 0, c->stream);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(MDT_EHIP, "graph capture failed: %s / %s", hipGetErrorString(e), hipGetErrorString(e2));
    }
    e = hipGraphInstantiate(&b->gexec_h[h], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->gexec_h[h] = nullptr;
        return fail(MDT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return 0;
}

// nsteps steps as two halves on the context's stream and side stream 0,
// joined on the context's stream; graph chunks of S steps when S > 0
static int step_split(mdt_batch *b, int nsteps, int S, int *replays_out) {
    mdt_context *c = b->ctx;
    int rc;
    if (!b->d_ident) {
        std::vector<int> id(b->nrep);
        for (int i = 0; i < b->nrep; ++i) id[i] = i;
        if ((rc = dalloc(&b->d_ident, (size_t)b->nrep))) return rc;
        HIPCHK(hipMemcpy(b->d_ident, id.data(), sizeof(int) * b->nrep, hipMemcpyHostToDevice));
    }
    if (!b->d_step_h && (rc = dalloc(&b->d_step_h, 2))) return rc;
    int replays = 0;
    if (S > 0 && nsteps >= S) {
        if ((b->gexec_h[0] || b->gexec_h[1]) && b->graph_arith != b->arith()) drop_graph(b);
        if (!b->gexec_h[0] || !b->gexec_h[1]) {
            drop_graph(b);
            if ((rc = capture_half(b, 0)) || (rc = capture_half(b, 1))) return rc;
            b->graph_arith = b->arith();
        }
    }
    hipStream_t st[2] = {c->stream, c->side[0]};
    HIPCHK(hipEventRecord(c->fork_ev, c->stream));
    HIPCHK(hipStreamWaitEvent(st[1], c->fork_ev, 0));
    if (S > 0 && nsteps >= S) {
        for (int h = 0; h < 2; ++h) HIPCHK(mdt_launch_advance(b->d_step_h + h, b->step, 1, st[h]));
        replays = nsteps / S;
        for (int k = 0; k < replays; ++k)
            for (int h = 0; h < 2; ++h) HIPCHK(hipGraphLaunch(b->gexec_h[h], st[h]));
        b->step += (uint32_t)(replays * S);
    }
    const int rest = nsteps - replays * S;
    for (int h = 0; h < 2; ++h)
        if ((rc = enqueue_half(b, h, rest, nullptr, b->step, st[h]))) return rc;
    b->step += (uint32_t)rest;
    HIPCHK(hipEventRecord(c->join_ev[0], st[1]));
    HIPCHK(hipStreamWaitEvent(st[0], c->join_ev[0], 0));
    *replays_out = replays;
    return 0;
}

// Captures one chunk of graph_chunk steps (step numbers relative to d_step,
// which the chunk's last node advances) and instantiates it.
static int capture_chunk(mdt_batch *b) {
    mdt_context *c = b->ctx;
    if (!b->d_step) {
        int rc = dalloc(&b->d_step, 1);
        if (rc) return rc;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_steps(b, b->graph_chunk, b->d_step, 0u, false);
    hipError_t e = rc ? hipSuccess : mdt_launch_advance(b->d_step, (uint32_t)b->graph_chunk, 0, c->stream);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(MDT_EHIP, "graph capture failed: %s / %s", hipGetErrorString(e), hipGetErrorString(e2));
    }
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->gexec = nullptr;
        return fail(MDT_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return 0;
}

int mdt_batch_step(mdt_batch *b, int nsteps) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (nsteps < 0) return fail(MDT_EINVAL, "nsteps=%d", nsteps);
    HIPCHK(hipSetDevice(b->ctx->device));
    int rc = ensure_ready(b, true);
    if (rc) return rc;
    if ((rc = autotune_if_needed(b))) return rc;
    if (b->pes == 1 && (rc = ensure_nn(b))) return rc;
    mdt_context *c = b->ctx;
    if (b->profile && (int)b->evf.size() < 2 * nsteps) {
        const size_t old = b->evf.size();
        b->evf.resize(2 * nsteps);
        for (size_t k = old; k < b->evf.size(); ++k) HIPCHK(hipEventCreate(&b->evf[k]));
    }
    HIPCHK(hipEventRecord(b->ev0, c->stream));
    // profiling times every launch and the network's judge is keyed by the
    // host's step count: both launch directly
    const int S = (b->profile || (b->pes == 1 && b->judge_interval > 0)) ? 0 : b->graph_chunk;
    int replays = 0;
    b->last_split = use_split(b) ? 1 : 0;
    if (b->last_split) {
        if ((rc = step_split(b, nsteps, S, &replays))) return rc;
    } else {
        if (S > 0 && nsteps >= S) {
            if (b->gexec && b->graph_arith != b->arith()) drop_graph(b);  // EMT parameters changed the arithmetic
            if (!b->gexec) {
                if ((rc = capture_chunk(b))) return rc;
                b->graph_arith = b->arith();
            }
            HIPCHK(mdt_launch_advance(b->d_step, b->step, 1, c->stream));
            replays = nsteps / S;
            for (int k = 0; k < replays; ++k) HIPCHK(hipGraphLaunch(b->gexec, c->stream));
            b->step += (uint32_t)(replays * S);
        }
        const int rest = nsteps - replays * S;
        if ((rc = enqueue_steps(b, rest, nullptr, b->step, b->profile != 0))) return rc;
        b->step += (uint32_t)rest;
    }
    HIPCHK(hipEventRecord(b->ev1, c->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_nlaunch = nsteps;
    b->graph_replays = replays;
    b->last_force_ms = 0.0;
    if (b->profile)
        for (int s = 0; s < nsteps; ++s) {
            float fm = 0.f;
            HIPCHK(hipEventElapsedTime(&fm, b->evf[2 * s], b->evf[2 * s + 1]));
            b->last_force_ms += fm;
        }
    return check_status(b);
}

int mdt_batch_set_split(mdt_batch *b, int mode) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (mode < -1 || mode > 1) return fail(MDT_EINVAL, "mdt_batch_set_split: mode %d (want -1, 0 or 1)", mode);
    HIPCHK(hipSetDevice(b->ctx->device));
    if (mode != b->split_mode) drop_graph(b);
    b->split_mode = mode;
    return 0;
}

int mdt_batch_last_split(mdt_batch *b, int *halves) {
    if (!b || !halves) return fail(MDT_EINVAL, "mdt_batch_last_split: NULL argument");
    *halves = b->last_split ? 2 : 1;
    return 0;
}

int mdt_batch_set_graph(mdt_batch *b, int chunk) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (chunk < 0 || chunk > 4096) return fail(MDT_EINVAL, "graph chunk %d outside 0..4096", chunk);
    HIPCHK(hipSetDevice(b->ctx->device));
    if (chunk != b->graph_chunk) drop_graph(b);
    b->graph_chunk = chunk;
    return 0;
}

int mdt_batch_set_precision(mdt_batch *b, int mode) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (mode != 0 && mode != 1) return fail(MDT_EINVAL, "precision mode %d (0: FP64, 1: FP32 pair terms)", mode);
    HIPCHK(hipSetDevice(b->ctx->device));
    if (mode != b->precision) drop_graph(b);
    b->precision = mode;
    return 0;
}

int mdt_batch_set_keep(mdt_batch *b, int mode) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (mode < -1 || mode > 2) return fail(MDT_EINVAL, "mdt_batch_set_keep: mode %d (want -1, 0, 1 or 2)", mode);
    b->keep_mode = mode;
    b->tpa = 0;  // the next launch configures (and autotunes) again
    drop_graph(b);
    return 0;
}

int mdt_batch_set_wide(mdt_batch *b, int mode) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (mode < -1 || mode > 1) return fail(MDT_EINVAL, "mdt_batch_set_wide: mode %d (want -1, 0 or 1)", mode);
    b->wide_mode = mode;
    b->tpa = 0;  // the next launch autotunes again
    drop_graph(b);
    return 0;
}

int mdt_batch_class_shapes(mdt_batch *b, int max, int32_t *out, int *nclass) {
    if (!b || !nclass || (max > 0 && !out)) return fail(MDT_EINVAL, "NULL argument");
    *nclass = (int)b->cls.size();
    for (int k = 0; k < max && k < (int)b->cls.size(); ++k) {
        const SizeClass &c = b->cls[k];
        const int32_t v[6] = {c.nmax, c.nrep, c.tpa, c.keep, c.block, c.nbcap};
        for (int q = 0; q < 6; ++q) out[6 * k + q] = v[q];
    }
    return 0;
}

int mdt_batch_precision_arith(mdt_batch *b, int *arith) {
    if (!b || !arith) return fail(MDT_EINVAL, "NULL argument");
    if (!b->ctx->have_params) return fail(MDT_ESTATE, "no EMT parameters (call mdt_emt_params_set)");
    *arith = b->arith();
    return 0;
}

int mdt_batch_set_com_removal(mdt_batch *b, int on) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    const int off = on ? 0 : 1;
    if (off != b->no_com) drop_graph(b);
    b->no_com = off;
    return 0;
}

int mdt_batch_set_list_skin(mdt_batch *b, double skin) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (!(skin >= 0.0 && skin < 5.0)) return fail(MDT_EINVAL, "list skin %g A (0 <= skin < 5)", skin);
    HIPCHK(hipSetDevice(b->ctx->device));
    if (skin != b->list_skin) drop_graph(b);
    const bool relayout = (skin > 0.0) != (b->list_skin > 0.0);  // classes' lists move between LDS and HBM
    b->list_skin = skin;
    if (relayout && !b->cls.empty()) {
        int rc = lds_limit(b);
        if (rc) return rc;
        for (SizeClass &c : b->cls)
            if ((rc = configure_class(b, c, c.tpa > 0 ? c.tpa : heuristic_tpa(c.nmax), b->nbcap_req,
                                      c.tpa > 0 ? c.keep : -1, c.tpa > 0 ? c.wide : 0)))
                return rc;
        note_tpa(b);
    }
    return invalidate_lists(b);
}

int mdt_batch_set_bead_mode(mdt_batch *b, int mode) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (mode != MDT_BEADS_ALL && mode != MDT_BEADS_MIXED)
        return fail(MDT_EINVAL, "bead mode %d (0: every atom carries the beads, 1: projectile beads only)", mode);
    HIPCHK(hipSetDevice(b->ctx->device));
    if (mode == b->bead_mode) return 0;
    drop_graph(b);
    b->bead_mode = mode;
    int rc = build_classes(b);
    if (rc) return rc;
    if (b->have_atoms && b->have_prop && mode == MDT_BEADS_MIXED && (rc = check_mixed(b))) {
        b->bead_mode = MDT_BEADS_ALL;
        build_classes(b);
        b->tpa = 0;
        return rc;
    }
    return configure(b, 0, b->nbcap_req);
}

int mdt_batch_set_observe(mdt_batch *b, double proj_ul, double proj_ll, double ads_start, double ads_end,
                          int exit_at_limits) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    drop_graph(b);  // the captured launches carry the observation arguments
    if (!b->d_obs) {
        int rc = dalloc(&b->d_obs, b->ntraj) | dalloc(&b->d_done, b->ntraj);
        if (rc) return rc;
    }
    // the reference's starting values (src/trajectory_info.f90:53-67)
    std::vector<TrajObs> o(b->ntraj);
    for (TrajObs &x : o) {
        x.cl_appr = DBL_MAX;
        x.cl_last = DBL_MAX;
        x.r_min[0] = x.r_min[1] = x.r_min[2] = DBL_MAX;
        x.t_int = 0.0;
        x.turning_points = 0;
        x.moving_down = 1;
        x.adsorbed = 0;
        x.adsorbed_prev = 0;
        x.exit_step = 0;
        x.cl_appr_step = 0;
        x.last_step = 0;
        x.pad_ = 0;
    }
    HIPCHK(hipMemcpy(b->d_obs, o.data(), sizeof(TrajObs) * b->ntraj, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b->d_done, 0, sizeof(int) * b->ntraj));
    b->obs_ul = proj_ul;
    b->obs_ll = proj_ll;
    b->ads_start = ads_start;
    b->ads_end = ads_end;
    b->obs_exit = exit_at_limits ? 1 : 0;
    b->obs_step0 = b->step;
    return 0;
}

static_assert(sizeof(TrajObs) == sizeof(mdt_traj_obs) && offsetof(TrajObs, turning_points) ==
                  offsetof(mdt_traj_obs, turning_points) && offsetof(TrajObs, last_step) ==
                  offsetof(mdt_traj_obs, last_step),
              "TrajObs and mdt_traj_obs differ");

int mdt_batch_get_observables(mdt_batch *b, mdt_traj_obs *out) {
    if (!b || !out) return fail(MDT_EINVAL, "NULL argument");
    if (!b->d_obs) return fail(MDT_ESTATE, "observation is off (call mdt_batch_set_observe)");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipMemcpy(out, b->d_obs, sizeof(TrajObs) * b->ntraj, hipMemcpyDeviceToHost));
    return 0;
}

int mdt_batch_estimators(mdt_batch *b, double *out) {
    if (!b || !out) return fail(MDT_EINVAL, "NULL argument");
    HIPCHK(hipSetDevice(b->ctx->device));
    int rc = ensure_ready(b, false);
    if (rc) return rc;
    if (!b->d_est && (rc = dalloc(&b->d_est, (size_t)MDT_NEST * b->ntraj))) return rc;
    const BatchView bv = b->view();
    HIPCHK(mdt_launch_estimators(b->ntraj, b->d_pc, bv, b->d_est, b->ctx->stream));
    HIPCHK(hipMemcpyAsync(out, b->d_est, sizeof(double) * MDT_NEST * b->ntraj, hipMemcpyDeviceToHost,
                          b->ctx->stream));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return 0;
}

int mdt_batch_status(mdt_batch *b, int32_t *status, int clear) {
    if (!b || !status) return fail(MDT_EINVAL, "mdt_batch_status: NULL argument");
    HIPCHK(hipMemcpy(status, b->d_status, sizeof(int) * b->ntraj, hipMemcpyDeviceToHost));
    if (clear) HIPCHK(hipMemset(b->d_status, 0, sizeof(int) * b->ntraj));
    return 0;
}

int mdt_batch_last_timing(mdt_batch *b, double *ms_total, int *nlaunch) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (ms_total) *ms_total = b->last_ms;
    if (nlaunch) *nlaunch = b->last_nlaunch;
    return 0;
}

int mdt_batch_profile(mdt_batch *b, int on) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    b->profile = on ? 1 : 0;
    return 0;
}

int mdt_batch_force_timing(mdt_batch *b, double *force_ms_total) {
    if (!b || !force_ms_total) return fail(MDT_EINVAL, "mdt_batch_force_timing: NULL argument");
    if (!b->profile) return fail(MDT_ESTATE, "profiling is off (call mdt_batch_profile)");
    *force_ms_total = b->last_force_ms;
    return 0;
}

int mdt_batch_tune(mdt_batch *b, int tpa, int nbcap, int *tpa_chosen) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (tpa != 0 && tpa != 1 && tpa != 2 && tpa != 4 && tpa != 8)
        return fail(MDT_EINVAL, "tpa=%d (1, 2, 4, 8 or 0)", tpa);
    if (nbcap < 0 || nbcap > 65535) return fail(MDT_EINVAL, "nbcap=%d", nbcap);
    HIPCHK(hipSetDevice(b->ctx->device));
    b->nbcap_req = nbcap;
    int rc;
    if (tpa == 0) {
        if ((rc = ensure_ready(b, false))) return rc;
        rc = autotune(b);
    } else {
        rc = configure(b, tpa, nbcap);
    }
    if (rc) return rc;
    if (tpa_chosen) *tpa_chosen = b->tpa;
    return 0;
}

int mdt_compute_emt(mdt_context *ctx, int natoms, int nbeads, const int32_t *idx, const double *simbox,
                    const double *isimbox, const double *r, double *epot, double *f, double *dens,
                    int32_t flag) {
    (void)flag;  // compute_emt ignores it (forces are always computed)
    if (!ctx || !idx || !simbox || !isimbox || !r || !epot || !f || !dens)
        return fail(MDT_EINVAL, "mdt_compute_emt: NULL argument");
    if (!ctx->have_params) return fail(MDT_ESTATE, "no EMT parameters (call mdt_emt_params_set)");
    int rc;
    if (!ctx->scratch || ctx->scratch_N != natoms || ctx->scratch_P != nbeads) {
        if (ctx->scratch) mdt_batch_destroy(ctx->scratch);
        ctx->scratch = nullptr;
        int32_t n = natoms, p = nbeads;
        if ((rc = mdt_batch_create(ctx, 1, &n, &p, &ctx->scratch))) return rc;
        ctx->scratch_N = natoms;
        ctx->scratch_P = nbeads;
    }
    mdt_batch *b = ctx->scratch;
    std::vector<double> mass(natoms, 1.0);
    std::vector<int32_t> fixed(3LL * natoms * nbeads, 0);
    if ((rc = mdt_batch_set_atoms(b, idx, mass.data(), fixed.data()))) return rc;
    if ((rc = mdt_batch_set_cells(b, simbox, isimbox))) return rc;
    if ((rc = mdt_batch_set_state(b, r, nullptr))) return rc;
    // compute_emt only: one force launch per replica, no ring-polymer step
    HIPCHK(hipSetDevice(ctx->device));
    if ((rc = ensure_ready(b, false))) return rc;
    if ((rc = autotune_if_needed(b))) return rc;
    HIPCHK(hipMemsetAsync(b->d_f, 0, sizeof(double) * b->tot3pn, ctx->stream));
    if ((rc = launch_replicas(b, b->view(), 0, ctx->stream))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if ((rc = check_status(b))) return rc;
    const long long n3 = 3LL * natoms * nbeads;
    std::vector<double> ff(n3), ee(nbeads), dd((size_t)natoms * nbeads);
    if ((rc = mdt_batch_get_state(b, nullptr, nullptr, ff.data(), nullptr, ee.data(), dd.data()))) return rc;
    for (long long k = 0; k < n3; ++k) f[k] += ff[k];
    for (int k = 0; k < nbeads; ++k) epot[k] += ee[k];
    for (int i = 0; i < natoms; ++i)
        if (ctx->ec.owner[idx[i] - 1])
            for (int k = 0; k < nbeads; ++k) dens[k + (long long)nbeads * i] = dd[k + (long long)nbeads * i];
    return 0;
}

// mdt_compute_emt without allocation: device arrays in and out, the
// universe staged in the device's scratch array (mdt_noalloc.hip), every
// launch on `stream` (NULL: the context's stream), nothing waited for.
int mdt_compute_emt_noalloc(mdt_context *ctx, int natoms, int nbeads, const int32_t *idx, const double *simbox,
                            const double *isimbox, const double *r, double *epot, double *f, double *dens,
                            int32_t *status, int32_t flag, void *stream) {
    (void)flag;  // compute_emt ignores it (forces are always computed)
    if (!ctx || !idx || !simbox || !isimbox || !r || !epot || !f || !dens)
        return fail(MDT_EINVAL, "mdt_compute_emt_noalloc: NULL argument");
    if (!ctx->have_params) return fail(MDT_ESTATE, "no EMT parameters (call mdt_emt_params_set)");
    if (!ctx->noalloc) return fail(MDT_ESTATE, "mdt_compute_emt_noalloc: no scratch array on this device");
    int rc = mdt_noalloc_check_sizes(natoms, nbeads);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return mdt_noalloc_compute(ctx->noalloc, ctx->lds_limit, ctx->d_ec, ctx->ntypes, natoms, nbeads, idx, simbox,
                               isimbox, r, epot, f, dens, status, st);
}

// ---------------------------------------------------------------------------
// NENE surrogate (src/pes_nene_mod.f90: read_nene :123-5759, compute_nene
// :7145-7368)
// ---------------------------------------------------------------------------
int mdt_nn_params_set(mdt_context *ctx, int nelem, const int32_t *elem_of_type, int ntypes, const int32_t *nsf,
                      const int32_t *sf_type, const int32_t *sf_e1, const int32_t *sf_e2, const double *eta,
                      const double *rs, const double *lambda, const double *zeta, const double *rc,
                      const double *gmin, const double *gmax, const double *gavg, double scmin, double scmax,
                      int32_t scale, int32_t center, int32_t nh1, int32_t nh2, const char *activations,
                      const double *weights) {
    if (!ctx || !elem_of_type || !nsf || !sf_type || !sf_e1 || !sf_e2 || !eta || !rs || !lambda || !zeta || !rc ||
        !gmin || !gmax || !gavg || !activations || !weights)
        return fail(MDT_EINVAL, "mdt_nn_params_set: NULL argument");
    if (nelem < 1 || nelem > NN_MAXE) return fail(MDT_EUNSUP, "nelem=%d outside 1..%d", nelem, NN_MAXE);
    if (ntypes < 1 || ntypes > MDT_MAXT) return fail(MDT_EUNSUP, "ntypes=%d outside 1..%d", ntypes, MDT_MAXT);
    if (nh1 < 1 || nh1 > NN_H || nh2 < 1 || nh2 > NN_H)
        return fail(MDT_EUNSUP, "hidden layers %d, %d (1..%d nodes each)", nh1, nh2, NN_H);
    NnConst *M = new NnConst();
    std::memset(M, 0, sizeof(NnConst));
    auto bail = [&](int code, const char *msg, int a) {
        delete M;
        return fail(code, msg, a);
    };
    M->nelem = nelem;
    int acts[3];
    for (int l = 0; l < 3; ++l) {
        const char a = activations[l];
        acts[l] = a == 't' ? 1 : a == 's' ? 2 : a == 'l' ? 0 : -1;
        if (acts[l] < 0) return bail(MDT_EUNSUP, "activation %d not in t, s, l", l);
    }
    M->act1 = acts[0];
    M->act2 = acts[1];
    M->act3 = acts[2];
    for (int t = 0; t < MDT_MAXT; ++t) M->elem_of_type[t] = -1;
    for (int t = 0; t < ntypes; ++t) {
        if (elem_of_type[t] < -1 || elem_of_type[t] >= nelem) return bail(MDT_EINVAL, "elem_of_type[%d]", t);
        M->elem_of_type[t] = elem_of_type[t];
    }
    const double ang2bohr = 1.0 / 0.529177211;  // src/constants.f90:161
    double rc0 = -1.0;
    long o = 0, wo = 0;
    for (int e = 0; e < nelem; ++e) {
        const int n = nsf[e];
        if (n < 1 || n > NN_K1) return bail(MDT_EUNSUP, "element %d: number of symmetry functions outside 1..64", e);
        int nr = 0, na = 0;
        for (int k = 0; k < n; ++k) {
            const long q = o + k;
            if (rc0 < 0) rc0 = rc[q];
            if (std::fabs(rc[q] - rc0) > 1e-12 * rc0) return bail(MDT_EUNSUP, "element %d: more than one cut-off", e);
            if (sf_type[q] == 2) {
                if (na > 0) return bail(MDT_EINVAL, "element %d: radial functions must precede angular ones", e);
                if (nr >= NN_MAXR) return bail(MDT_EUNSUP, "element %d: too many radial functions", e);
                M->r_el[e][nr] = sf_e1[q];
                M->r_eta[e][nr] = (float)eta[q];
                M->r_rs[e][nr] = (float)rs[q];
                ++nr;
            } else if (sf_type[q] == 3) {
                if (na >= NN_MAXA) return bail(MDT_EUNSUP, "element %d: too many angular functions", e);
                int zp = -1;
                for (int z = 0; z <= 4; ++z)
                    if (zeta[q] == (double)(1 << z)) zp = z;
                if (zp < 0) return bail(MDT_EUNSUP, "element %d: zeta must be 1, 2, 4, 8 or 16", e);
                const int a1 = std::min(sf_e1[q], sf_e2[q]), a2 = std::max(sf_e1[q], sf_e2[q]);
                M->a_e1[e][na] = a1;
                M->a_e2[e][na] = a2;
                M->a_etal2[e][na] = (float)(eta[q] * 1.4426950408889634);
                M->a_lam[e][na] = (float)lambda[q];
                M->a_norm[e][na] = (float)std::pow(2.0, 1.0 - zeta[q]);
                M->a_zp[e][na] = zp;
                ++na;
            } else {
                return bail(MDT_EUNSUP, "symmetry function type %d (2 and 3 are supported)", sf_type[q]);
            }
            const double mn = gmin[q], mx = gmax[q], av = gavg[q];
            if (scale && !(mx > mn)) return bail(MDT_EINVAL, "element %d: scaling range max <= min", e);
            double mul = 1.0, add = 0.0, dfac = 1.0;
            if (scale && !center) {
                mul = (scmax - scmin) / (mx - mn);
                add = scmin - mn * mul;
            } else if (center && !scale) {
                add = -av;
            } else if (scale && center) {
                mul = 1.0 / (mx - mn);
                add = -av * mul;
            }
            if (scale) dfac = (scmax - scmin) / (mx - mn);  // getshortatomic.f90:177-189 (whenever scaling is on)
            M->smul[e][k] = (float)mul;
            M->sadd[e][k] = (float)add;
            M->dfac[e][k] = (float)dfac;
            M->glo[e][k] = (float)(mn - 1e-4);
            M->ghi[e][k] = (float)(mx + 1e-4);
        }
        M->nr[e] = nr;
        M->na[e] = na;
        // weights, RuNNer order: W1 (n x nh1) input-major, b1, W2 (nh1 x nh2), b2, W3 (nh2 x 1), b3
        const double *w = weights + wo;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < nh1; ++j) M->W1[e][i][j] = (float)w[i * nh1 + j];
        w += n * nh1;
        for (int j = 0; j < nh1; ++j) M->b1[e][j] = (float)w[j];
        w += nh1;
        for (int i = 0; i < nh1; ++i)
            for (int j = 0; j < nh2; ++j) M->W2[e][i][j] = (float)w[i * nh2 + j];
        w += nh1 * nh2;
        for (int j = 0; j < nh2; ++j) M->b2[e][j] = (float)w[j];
        w += nh2;
        for (int j = 0; j < nh2; ++j) M->w3[e][j] = (float)w[j];
        w += nh2;
        M->b3[e] = (float)w[0];
        wo += (long)n * nh1 + nh1 + (long)nh1 * nh2 + nh2 + nh2 + 1;
        o += n;
    }
    M->nh1 = nh1;
    M->nh2 = nh2;
    M->radial_only = 1;
    for (int e = 0; e < nelem; ++e) {
        if (M->na[e] > 0) M->radial_only = 0;
        // radial functions grouped by neighbour element (stable)
        int q = 0;
        for (int x = 0; x < NN_MAXE; ++x) {
            M->rk_lo[e][x] = q;
            for (int k = 0; k < M->nr[e]; ++k) {
                if (M->r_el[e][k] != x) continue;
                M->rk_perm[e][q] = k;
                M->rk_rs[e][q] = M->r_rs[e][k];
                M->rk_ce[e][q] = (float)(-(double)M->r_eta[e][k] * 1.4426950408889634);
                M->rk_b[e][q] = -2.0f * M->r_eta[e][k];
                ++q;
            }
            M->rk_hi[e][x] = q;
        }
        if (q != M->nr[e]) return bail(MDT_EINVAL, "element %d: radial function on an element outside the model", e);
    }
    M->rc = (float)rc0;
    M->rc2 = (float)(rc0 * rc0);
    M->pi_rc = (float)(M_PI / rc0);
    M->ang2bohr = (float)ang2bohr;
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->d_nn) {
        int rc2 = dalloc(&ctx->d_nn, 1);
        if (rc2) {
            delete M;
            return rc2;
        }
    }
    HIPCHK(hipMemcpy(ctx->d_nn, M, sizeof(NnConst), hipMemcpyHostToDevice));
    delete ctx->nn;
    ctx->nn = M;
    ctx->have_nn = true;
    return 0;
}

int mdt_batch_set_pes(mdt_batch *b, int pes, int judge_interval) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (pes != 0 && pes != 1) return fail(MDT_EINVAL, "pes %d (0: EMT, 1: network)", pes);
    if (judge_interval < 0) return fail(MDT_EINVAL, "judge_interval %d", judge_interval);
    if (pes == 1 && !b->ctx->have_nn) return fail(MDT_ESTATE, "no network (call mdt_nn_params_set)");
    if (pes == 1 && b->bead_mode == MDT_BEADS_MIXED)
        return fail(MDT_EUNSUP, "mixed bead mode runs with the EMT potential only");
    HIPCHK(hipSetDevice(b->ctx->device));
    if (pes != b->pes) drop_graph(b);
    b->pes = pes;
    b->judge_interval = judge_interval;
    if (pes == 1) return ensure_nn(b);
    return 0;
}

int mdt_batch_set_judge_percent(mdt_batch *b, int percent) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (percent < 1 || percent > 100 || 100 % percent != 0)
        return fail(MDT_EINVAL, "judge percent %d (a divisor of 100)", percent);
    b->judge_mod = 100 / percent;
    return 0;
}

int mdt_batch_nn_stats(mdt_batch *b, double *judge_sq, double *judge_n, double *judge_max, int32_t *nfallback) {
    if (!b) return fail(MDT_EINVAL, "NULL batch");
    if (!b->d_nnG) return fail(MDT_ESTATE, "the network has not run on this batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    std::vector<double> j(3 * (size_t)b->nrep);
    std::vector<int> f(b->nrep);
    HIPCHK(hipMemcpy(j.data(), b->d_judge, sizeof(double) * j.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(f.data(), b->d_nnflag, sizeof(int) * f.size(), hipMemcpyDeviceToHost));
    double sq = 0, n = 0, mx = 0;
    int nf = 0;
    for (int r = 0; r < b->nrep; ++r) {
        sq += j[3 * r];
        n += j[3 * r + 1];
        mx = std::max(mx, j[3 * r + 2]);
        nf += f[r] != 0;
    }
    if (judge_sq) *judge_sq = sq;
    if (judge_n) *judge_n = n;
    if (judge_max) *judge_max = mx;
    if (nfallback) *nfallback = nf;
    return 0;
}

int mdt_batch_nn_fallback(mdt_batch *b, int32_t *flags) {
    if (!b || !flags) return fail(MDT_EINVAL, "mdt_batch_nn_fallback: NULL argument");
    if (!b->d_nnflag) return fail(MDT_ESTATE, "the network has not run on this batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    std::vector<int> f(b->nrep);
    HIPCHK(hipMemcpy(f.data(), b->d_nnflag, sizeof(int) * f.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < b->nrep; ++r) flags[r] = f[r] != 0;
    return 0;
}

int mdt_batch_nn_descriptors(mdt_batch *b, float *G) {
    if (!b || !G) return fail(MDT_EINVAL, "mdt_batch_nn_descriptors: NULL argument");
    if (!b->d_nnG) return fail(MDT_ESTATE, "the network has not run on this batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipMemcpy(G, b->d_nnG, sizeof(float) * (size_t)b->nrep * b->maxN * NN_K1, hipMemcpyDeviceToHost));
    return 0;
}

int mdt_compute_nene(mdt_context *ctx, int natoms, int nbeads, const int32_t *idx, const double *simbox,
                     const double *isimbox, const double *r, double *epot, double *f, int32_t *nextrap,
                     int32_t flag) {
    (void)flag;  // forces are always computed
    if (!ctx || !idx || !simbox || !isimbox || !r || !epot || !f)
        return fail(MDT_EINVAL, "mdt_compute_nene: NULL argument");
    if (!ctx->have_nn) return fail(MDT_ESTATE, "no network (call mdt_nn_params_set)");
    int rc;
    if (!ctx->scratch || ctx->scratch_N != natoms || ctx->scratch_P != nbeads) {
        if (ctx->scratch) mdt_batch_destroy(ctx->scratch);
        ctx->scratch = nullptr;
        int32_t n = natoms, p = nbeads;
        if ((rc = mdt_batch_create(ctx, 1, &n, &p, &ctx->scratch))) return rc;
        ctx->scratch_N = natoms;
        ctx->scratch_P = nbeads;
    }
    mdt_batch *b = ctx->scratch;
    std::vector<double> mass(natoms, 1.0);
    std::vector<int32_t> fixed(3LL * natoms * nbeads, 0);
    if ((rc = mdt_batch_set_atoms(b, idx, mass.data(), fixed.data()))) return rc;
    if ((rc = mdt_batch_set_cells(b, simbox, isimbox))) return rc;
    if ((rc = mdt_batch_set_state(b, r, nullptr))) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    b->pes = 1;
    if ((rc = ensure_nn(b))) return rc;
    if ((rc = ensure_ready(b, false))) return rc;
    HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(int) * b->ntraj, ctx->stream));
    HIPCHK(mdt_launch_nene(b->nrep, nn_smem(ctx, b->maxN), ctx->d_nn, b->view(), b->d_nnG, b->d_nnC, b->d_nnflag,
                           b->d_nnF, b->maxN, 0, 0, nn_radial(ctx), b->d_nnNbr, b->d_nnCnt, b->d_nnRef, b->d_nnKey,
                           b->list_skin / MDT_BOHR2ANG, ctx->stream, 1));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<int> st(1);
    HIPCHK(hipMemcpy(st.data(), b->d_status, sizeof(int), hipMemcpyDeviceToHost));
    b->pes = 0;
    if (st[0] & ~MDT_ST_NN_EXTRAPOLATION)
        return fail(MDT_ETRAJ, "compute_nene: status %d (1: neighbour overflow, 4: atom type not in the network, "
                    "8: fixed-point overflow)", st[0]);
    if (nextrap) *nextrap = (st[0] & MDT_ST_NN_EXTRAPOLATION) ? 1 : 0;
    const long long n3 = 3LL * natoms * nbeads;
    std::vector<double> ff(n3), ee(nbeads);
    if ((rc = mdt_batch_get_state(b, nullptr, nullptr, ff.data(), nullptr, ee.data(), nullptr))) return rc;
    for (long long k = 0; k < n3; ++k) f[k] += ff[k];
    for (int k = 0; k < nbeads; ++k) epot[k] += ee[k];
    return 0;
}

// ---------------------------------------------------------------------------
// Training of the surrogate on the GPU (no reference counterpart: md_tian2
// fits with RuNNer's Kalman filter in fit.f90)
// ---------------------------------------------------------------------------
struct mdt_nntrain {
    mdt_context *ctx = nullptr;
    int nframes = 0, maxN = 0;
    long long natoms = 0;
    int *d_natoms = nullptr, *d_atom_frame = nullptr, *d_frames = nullptr, *d_rows = nullptr;
    long long *d_aoff = nullptr;
    double *d_cell = nullptr, *d_icell = nullptr, *d_pos = nullptr, *d_Fref = nullptr, *d_Eref = nullptr;
    uint8_t *d_type = nullptr;
    float *d_G = nullptr, *d_C = nullptr, *d_U = nullptr;
    double *d_Eatom = nullptr, *d_cE = nullptr, *d_loss = nullptr;
    int4 *d_tiles = nullptr;
    float *d_partials = nullptr, *d_grad = nullptr, *d_am = nullptr, *d_av = nullptr;
    int nblocks_grad = 256;
    int adam_step = 0;
    // minibatches: frames d_frames[fr_off[m] .. fr_off[m+1]), tiles d_tiles[tl_off[m] ..], rows d_rows[..]
    std::vector<int> fr_off, tl_off, natoms_mb;
    std::vector<int> h_natoms;
    std::vector<long long> h_aoff;
    std::vector<uint8_t> h_elem;  // network element per atom
    int ntrain_mb = 0;            // minibatches 0 .. ntrain_mb-1 train, the rest validate
    uint64_t rng = 0;

    NnTrainView view() const {
        NnTrainView v{};
        v.nframes = nframes;
        v.fr_natoms = d_natoms;
        v.fr_aoff = d_aoff;
        v.fr_cell = d_cell;
        v.fr_icell = d_icell;
        v.type = d_type;
        v.pos = d_pos;
        v.Fref = d_Fref;
        v.Eref = d_Eref;
        v.atom_frame = d_atom_frame;
        v.G = d_G;
        v.C = d_C;
        v.U = d_U;
        v.Eatom = d_Eatom;
        v.cE = d_cE;
        v.loss = d_loss;
        return v;
    }
};

static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// tiles (element, first row, count) of 16 same-element atoms over a frame list
static void build_tiles(const mdt_nntrain *tr, const int *frames, int nf, int nelem, std::vector<int4> &tiles,
                        std::vector<int> &rows, int row_base) {
    for (int e = 0; e < nelem; ++e) {
        std::vector<int> at;
        for (int q = 0; q < nf; ++q) {
            const int f = frames[q];
            for (int i = 0; i < tr->h_natoms[f]; ++i)
                if (tr->h_elem[tr->h_aoff[f] + i] == e) at.push_back((int)(tr->h_aoff[f] + i));
        }
        for (size_t k = 0; k < at.size(); k += 16) {
            const int cnt = (int)std::min<size_t>(16, at.size() - k);
            tiles.push_back(make_int4(e, row_base + (int)rows.size(), cnt, 0));
            for (int j = 0; j < cnt; ++j) rows.push_back(at[k + j]);
        }
    }
}

int mdt_nn_train_destroy(mdt_nntrain *tr) {
    if (!tr) return 0;
    (void)hipSetDevice(tr->ctx->device);
    void *ptrs[] = {tr->d_natoms, tr->d_atom_frame, tr->d_frames, tr->d_rows, tr->d_aoff, tr->d_cell, tr->d_icell,
                    tr->d_pos, tr->d_Fref, tr->d_Eref, tr->d_type, tr->d_G, tr->d_C, tr->d_U, tr->d_Eatom,
                    tr->d_cE, tr->d_loss, tr->d_tiles, tr->d_partials, tr->d_grad, tr->d_am, tr->d_av};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete tr;
    return 0;
}

int mdt_nn_train_create(mdt_context *ctx, int nframes, const int32_t *natoms, const int32_t *idx,
                        const double *simbox, const double *r, const double *Eref, const double *Fref,
                        int batch_frames, double val_fraction, uint32_t seed, mdt_nntrain **out) {
    if (!ctx || !natoms || !idx || !simbox || !r || !Eref || !Fref || !out)
        return fail(MDT_EINVAL, "mdt_nn_train_create: NULL argument");
    *out = nullptr;
    if (!ctx->have_nn) return fail(MDT_ESTATE, "no network (call mdt_nn_params_set)");
    if (ctx->nn->act3 != 0) return fail(MDT_EUNSUP, "training needs a linear output layer");
    if (nframes < 1 || batch_frames < 1) return fail(MDT_EINVAL, "nframes=%d batch_frames=%d", nframes, batch_frames);
    if (!(val_fraction >= 0.0 && val_fraction < 1.0)) return fail(MDT_EINVAL, "val_fraction %g", val_fraction);
    HIPCHK(hipSetDevice(ctx->device));
    mdt_nntrain *tr = new mdt_nntrain();
    tr->ctx = ctx;
    tr->nframes = nframes;
    tr->rng = seed;
    tr->h_natoms.assign(natoms, natoms + nframes);
    tr->h_aoff.resize(nframes);
    long long na = 0;
    for (int f = 0; f < nframes; ++f) {
        if (natoms[f] < 1 || natoms[f] > 4096) {
            delete tr;
            return fail(MDT_EINVAL, "frame %d: natoms=%d", f, natoms[f]);
        }
        tr->h_aoff[f] = na;
        na += natoms[f];
        tr->maxN = std::max(tr->maxN, natoms[f]);
    }
    tr->natoms = na;
    const NnConst &M = *ctx->nn;
    std::vector<uint8_t> type(na);
    std::vector<int> afr(na);
    tr->h_elem.resize(na);
    for (int f = 0; f < nframes; ++f)
        for (int i = 0; i < natoms[f]; ++i) {
            const long long a = tr->h_aoff[f] + i;
            const int t = idx[a] - 1;
            if (t < 0 || t >= MDT_MAXT || M.elem_of_type[t] < 0) {
                mdt_nn_train_destroy(tr);
                return fail(MDT_EINVAL, "frame %d atom %d: type %d is not in the network", f, i, idx[a]);
            }
            type[a] = (uint8_t)t;
            afr[a] = f;
            tr->h_elem[a] = (uint8_t)M.elem_of_type[t];
        }
    std::vector<double> icell(9 * (size_t)nframes);
    for (int f = 0; f < nframes; ++f) {
        const double *a = simbox + 9 * f;  // column-major: a[i + 3 j]
        double *ai = icell.data() + 9 * f;
        const double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
                           a[6] * (a[1] * a[5] - a[4] * a[2]);
        if (std::fabs(det) < 1e-12) {
            mdt_nn_train_destroy(tr);
            return fail(MDT_EINVAL, "frame %d: singular cell", f);
        }
        const double id = 1.0 / det;
        ai[0] = (a[4] * a[8] - a[7] * a[5]) * id;
        ai[3] = (a[6] * a[5] - a[3] * a[8]) * id;
        ai[6] = (a[3] * a[7] - a[6] * a[4]) * id;
        ai[1] = (a[7] * a[2] - a[1] * a[8]) * id;
        ai[4] = (a[0] * a[8] - a[6] * a[2]) * id;
        ai[7] = (a[6] * a[1] - a[0] * a[7]) * id;
        ai[2] = (a[1] * a[5] - a[4] * a[2]) * id;
        ai[5] = (a[3] * a[2] - a[0] * a[5]) * id;
        ai[8] = (a[0] * a[4] - a[3] * a[1]) * id;
    }
    // minibatches over a seeded permutation; the last val_fraction of the
    // permutation (whole minibatches) is held out for validation
    std::vector<int> perm(nframes);
    for (int f = 0; f < nframes; ++f) perm[f] = f;
    for (int f = nframes - 1; f > 0; --f) std::swap(perm[f], perm[splitmix(tr->rng) % (uint64_t)(f + 1)]);
    const int nmb = (nframes + batch_frames - 1) / batch_frames;
    int nval = (int)std::floor(val_fraction * nmb + 0.5);
    if (val_fraction > 0 && nval == 0 && nmb > 1) nval = 1;
    if (nval >= nmb) nval = nmb - 1;
    tr->ntrain_mb = nmb - nval;
    std::vector<int4> tiles;
    std::vector<int> rows;
    tr->fr_off.push_back(0);
    tr->tl_off.push_back(0);
    for (int m = 0; m < nmb; ++m) {
        const int f0 = m * batch_frames, f1 = std::min(nframes, f0 + batch_frames);
        build_tiles(tr, perm.data() + f0, f1 - f0, M.nelem, tiles, rows, 0);
        int nm = 0;
        for (int q = f0; q < f1; ++q) nm += natoms[perm[q]];
        tr->natoms_mb.push_back(nm);
        tr->fr_off.push_back(f1);
        tr->tl_off.push_back((int)tiles.size());
    }
    const size_t rowsK = (size_t)na * NN_K1;
    const int W = tr->nblocks_grad * 4;
    int rc = dalloc(&tr->d_natoms, nframes) | dalloc(&tr->d_atom_frame, na) | dalloc(&tr->d_frames, nframes) |
             dalloc(&tr->d_rows, rows.size()) | dalloc(&tr->d_aoff, nframes) | dalloc(&tr->d_cell, 9 * (size_t)nframes) |
             dalloc(&tr->d_icell, 9 * (size_t)nframes) | dalloc(&tr->d_pos, 3 * (size_t)na) |
             dalloc(&tr->d_Fref, 3 * (size_t)na) | dalloc(&tr->d_Eref, nframes) | dalloc(&tr->d_type, na) |
             dalloc(&tr->d_G, rowsK) | dalloc(&tr->d_C, rowsK) | dalloc(&tr->d_U, rowsK) |
             dalloc(&tr->d_Eatom, na) | dalloc(&tr->d_cE, nframes) | dalloc(&tr->d_loss, 2 * (size_t)nframes) |
             dalloc(&tr->d_tiles, tiles.size()) | dalloc(&tr->d_partials, (size_t)W * NN_MAXE * NN_NPAR) |
             dalloc(&tr->d_grad, (size_t)NN_MAXE * NN_NPAR) | dalloc(&tr->d_am, (size_t)NN_MAXE * NN_NPAR) |
             dalloc(&tr->d_av, (size_t)NN_MAXE * NN_NPAR);
    if (rc) {
        mdt_nn_train_destroy(tr);
        return rc;
    }
    auto up = [&](void *d, const void *h, size_t n) { return hipMemcpy(d, h, n, hipMemcpyHostToDevice); };
    HIPCHK(up(tr->d_natoms, natoms, sizeof(int) * nframes));
    HIPCHK(up(tr->d_atom_frame, afr.data(), sizeof(int) * na));
    HIPCHK(up(tr->d_frames, perm.data(), sizeof(int) * nframes));
    HIPCHK(up(tr->d_rows, rows.data(), sizeof(int) * rows.size()));
    HIPCHK(up(tr->d_aoff, tr->h_aoff.data(), sizeof(long long) * nframes));
    HIPCHK(up(tr->d_cell, simbox, sizeof(double) * 9 * nframes));
    HIPCHK(up(tr->d_icell, icell.data(), sizeof(double) * 9 * nframes));
    HIPCHK(up(tr->d_pos, r, sizeof(double) * 3 * na));
    HIPCHK(up(tr->d_Fref, Fref, sizeof(double) * 3 * na));
    HIPCHK(up(tr->d_Eref, Eref, sizeof(double) * nframes));
    HIPCHK(up(tr->d_type, type.data(), na));
    HIPCHK(up(tr->d_tiles, tiles.data(), sizeof(int4) * tiles.size()));
    HIPCHK(hipMemset(tr->d_am, 0, sizeof(float) * NN_MAXE * NN_NPAR));
    HIPCHK(hipMemset(tr->d_av, 0, sizeof(float) * NN_MAXE * NN_NPAR));
    HIPCHK(hipMemset(tr->d_G, 0, sizeof(float) * rowsK));
    // descriptors of every frame (they depend on the positions only)
    hipStream_t st = ctx->stream;
    HIPCHK(mdt_launch_nn_frame(nframes, mdt_nn_frame_smem(tr->maxN), ctx->d_nn, tr->view(), tr->d_frames, 0, 0.0,
                               0.0, st));
    HIPCHK(hipStreamSynchronize(st));
    *out = tr;
    return 0;
}

// min / max / mean of every symmetry function over the training frames'
// atoms, per element (nelem x 64 each): the scaling.data statistics
int mdt_nn_train_descriptor_stats(mdt_nntrain *tr, double *gmin, double *gmax, double *gavg) {
    if (!tr || !gmin || !gmax || !gavg) return fail(MDT_EINVAL, "mdt_nn_train_descriptor_stats: NULL argument");
    HIPCHK(hipSetDevice(tr->ctx->device));
    std::vector<float> G((size_t)tr->natoms * NN_K1);
    HIPCHK(hipMemcpy(G.data(), tr->d_G, sizeof(float) * G.size(), hipMemcpyDeviceToHost));
    const NnConst &M = *tr->ctx->nn;
    for (int e = 0; e < M.nelem; ++e) {
        const int n = M.nr[e] + M.na[e];
        std::vector<double> mn(n, 1e300), mx(n, -1e300), sm(n, 0.0);
        long long cnt = 0;
        for (long long a = 0; a < tr->natoms; ++a) {
            if (tr->h_elem[a] != e) continue;
            ++cnt;
            for (int k = 0; k < n; ++k) {
                const double g = G[(size_t)a * NN_K1 + k];
                mn[k] = std::min(mn[k], g);
                mx[k] = std::max(mx[k], g);
                sm[k] += g;
            }
        }
        for (int k = 0; k < NN_K1; ++k) {
            gmin[e * NN_K1 + k] = k < n && cnt ? mn[k] : 0.0;
            gmax[e * NN_K1 + k] = k < n && cnt ? mx[k] : 0.0;
            gavg[e * NN_K1 + k] = k < n && cnt ? sm[k] / cnt : 0.0;
        }
    }
    return 0;
}

// gradient blocks for ntiles tiles: about two tiles per wavefront, at most
// the partial buffer's nblocks_grad
static int grad_blocks(const mdt_nntrain *tr, int ntiles) {
    return std::max(1, std::min(tr->nblocks_grad, (ntiles + 7) / 8));
}

// one minibatch: forward/backward, forces, residuals, JVP, gradient partials;
// then mode 0: reduce into d_grad; mode 1: reduce + Adam
static int run_minibatch(mdt_nntrain *tr, int m, double wE, double wF, int mode, float lr) {
    mdt_context *c = tr->ctx;
    hipStream_t st = c->stream;
    const NnTrainView tv = tr->view();
    const int nf = tr->fr_off[m + 1] - tr->fr_off[m];
    const int t0 = tr->tl_off[m], nt = tr->tl_off[m + 1] - t0;
    HIPCHK(mdt_launch_nn_mlp(nt, c->d_nn, tv, tr->d_tiles + t0, tr->d_rows, st));
    HIPCHK(mdt_launch_nn_frame(nf, mdt_nn_frame_smem(tr->maxN), c->d_nn, tv, tr->d_frames + tr->fr_off[m], 1,
                               wE / nf, wF / (3.0 * tr->natoms_mb[m]), st));
    if (mode < 0) return 0;  // evaluation only
    const int nb = grad_blocks(tr, nt), W = nb * 4;
    HIPCHK(hipMemsetAsync(tr->d_partials, 0, sizeof(float) * (size_t)W * NN_MAXE * NN_NPAR, st));
    HIPCHK(mdt_launch_nn_grad(nb, c->d_nn, tv, tr->d_tiles + t0, nt, tr->d_rows, tr->d_partials, st));
    if (mode == 1) ++tr->adam_step;
    HIPCHK(mdt_launch_nn_adam(c->d_nn, tr->d_partials, W, tr->d_grad, tr->d_am, tr->d_av, lr, 0.9f, 0.999f, 1e-8f,
                              tr->adam_step > 0 ? tr->adam_step : 1, mode, st));
    return 0;
}

static int sum_loss(mdt_nntrain *tr, int m0, int m1, double *lE, double *lF) {
    std::vector<double> loss(2 * (size_t)tr->nframes);
    std::vector<int> perm(tr->nframes);
    HIPCHK(hipMemcpy(loss.data(), tr->d_loss, sizeof(double) * loss.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(perm.data(), tr->d_frames, sizeof(int) * perm.size(), hipMemcpyDeviceToHost));
    double se = 0.0, sf = 0.0;
    for (int q = tr->fr_off[m0]; q < tr->fr_off[m1]; ++q) {
        se += loss[2 * perm[q]];
        sf += loss[2 * perm[q] + 1];
    }
    *lE = se;
    *lF = sf;
    return 0;
}

int mdt_nn_train_epoch(mdt_nntrain *tr, double lr, double wE, double wF, double *loss) {
    if (!tr) return fail(MDT_EINVAL, "NULL trainer");
    HIPCHK(hipSetDevice(tr->ctx->device));
    std::vector<int> order(tr->ntrain_mb);
    for (int m = 0; m < tr->ntrain_mb; ++m) order[m] = m;
    for (int m = tr->ntrain_mb - 1; m > 0; --m) std::swap(order[m], order[splitmix(tr->rng) % (uint64_t)(m + 1)]);
    int rc;
    for (int m : order)
        if ((rc = run_minibatch(tr, m, wE, wF, 1, (float)lr))) return rc;
    HIPCHK(hipStreamSynchronize(tr->ctx->stream));
    HIPCHK(hipMemcpy(tr->ctx->nn, tr->ctx->d_nn, sizeof(NnConst), hipMemcpyDeviceToHost));
    if (loss) {
        double lE, lF;
        if ((rc = sum_loss(tr, 0, tr->ntrain_mb, &lE, &lF))) return rc;
        *loss = (lE + lF) / std::max(1, tr->ntrain_mb);
    }
    return 0;
}

// energy RMSE per atom [eV/atom] and force-component RMSE [eV/A] of the
// current network over the training (set 0) or validation (set 1) frames
int mdt_nn_train_eval(mdt_nntrain *tr, int set, double *rmse_e, double *rmse_f) {
    if (!tr || !rmse_e || !rmse_f) return fail(MDT_EINVAL, "mdt_nn_train_eval: NULL argument");
    HIPCHK(hipSetDevice(tr->ctx->device));
    const int nmb = (int)tr->fr_off.size() - 1;
    const int m0 = set == 0 ? 0 : tr->ntrain_mb, m1 = set == 0 ? tr->ntrain_mb : nmb;
    if (m1 <= m0) return fail(MDT_ESTATE, "set %d is empty", set);
    int rc;
    long long nat = 0;
    for (int m = m0; m < m1; ++m) {
        const int nf = tr->fr_off[m + 1] - tr->fr_off[m];
        // cEw = nf, cFw = 3 natoms_mb: loss parts become dE^2 per atom^2 and |dF|^2
        if ((rc = run_minibatch(tr, m, (double)nf, 3.0 * tr->natoms_mb[m], -1, 0.f))) return rc;
        nat += tr->natoms_mb[m];
    }
    HIPCHK(hipStreamSynchronize(tr->ctx->stream));
    double lE, lF;
    if ((rc = sum_loss(tr, m0, m1, &lE, &lF))) return rc;
    *rmse_e = std::sqrt(lE / (tr->fr_off[m1] - tr->fr_off[m0]));
    *rmse_f = std::sqrt(lF / (3.0 * nat));
    return 0;
}

// padded-layout index of RuNNer weight q of element e (nin x n1 x n2 x 1)
static int runner_to_padded(int q, int nin, int n1, int n2) {
    if (q < nin * n1) return NN_PW1 + (q / n1) * NN_H + q % n1;
    q -= nin * n1;
    if (q < n1) return NN_PB1 + q;
    q -= n1;
    if (q < n1 * n2) return NN_PW2 + (q / n2) * NN_H + q % n2;
    q -= n1 * n2;
    if (q < n2) return NN_PB2 + q;
    q -= n2;
    if (q < n2) return NN_PW3 + q;
    return NN_PB3;
}

// Loss and gradient (RuNNer weight layout, elements concatenated) over the
// given frames, without updating the network
int mdt_nn_train_loss_grad(mdt_nntrain *tr, int nsel, const int32_t *frames, double wE, double wF, double *loss,
                           double *grad) {
    if (!tr || !frames || !loss || !grad) return fail(MDT_EINVAL, "mdt_nn_train_loss_grad: NULL argument");
    if (nsel < 1) return fail(MDT_EINVAL, "nsel=%d", nsel);
    HIPCHK(hipSetDevice(tr->ctx->device));
    mdt_context *c = tr->ctx;
    const NnConst &M = *c->nn;
    std::vector<int4> tiles;
    std::vector<int> rows;
    int nat = 0;
    for (int q = 0; q < nsel; ++q) {
        if (frames[q] < 0 || frames[q] >= tr->nframes) return fail(MDT_EINVAL, "frame %d", frames[q]);
        nat += tr->h_natoms[frames[q]];
    }
    build_tiles(tr, frames, nsel, M.nelem, tiles, rows, 0);
    int4 *dt = nullptr;
    int *dr = nullptr, *df = nullptr;
    int rc = dalloc(&dt, tiles.size()) | dalloc(&dr, rows.size()) | dalloc(&df, nsel);
    if (rc) return rc;
    hipStream_t st = c->stream;
    const NnTrainView tv = tr->view();
    const int nbg = grad_blocks(tr, (int)tiles.size()), W = nbg * 4;
    hipError_t e1 = hipMemcpy(dt, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice);
    if (e1 == hipSuccess) e1 = hipMemcpy(dr, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice);
    if (e1 == hipSuccess) e1 = hipMemcpy(df, frames, sizeof(int) * nsel, hipMemcpyHostToDevice);
    if (e1 == hipSuccess) e1 = mdt_launch_nn_mlp((int)tiles.size(), c->d_nn, tv, dt, dr, st);
    if (e1 == hipSuccess)
        e1 = mdt_launch_nn_frame(nsel, mdt_nn_frame_smem(tr->maxN), c->d_nn, tv, df, 1, wE / nsel, wF / (3.0 * nat), st);
    if (e1 == hipSuccess)
        e1 = hipMemsetAsync(tr->d_partials, 0, sizeof(float) * (size_t)W * NN_MAXE * NN_NPAR, st);
    if (e1 == hipSuccess)
        e1 = mdt_launch_nn_grad(nbg, c->d_nn, tv, dt, (int)tiles.size(), dr, tr->d_partials, st);
    if (e1 == hipSuccess)
        e1 = mdt_launch_nn_adam(c->d_nn, tr->d_partials, W, tr->d_grad, tr->d_am, tr->d_av, 0.f, 0.9f, 0.999f,
                                1e-8f, 1, 0, st);
    if (e1 == hipSuccess) e1 = hipStreamSynchronize(st);
    (void)hipFree(dt);
    (void)hipFree(dr);
    (void)hipFree(df);
    if (e1 != hipSuccess) return fail(MDT_EHIP, "training gradient: %s", hipGetErrorString(e1));
    std::vector<float> g((size_t)NN_MAXE * NN_NPAR);
    std::vector<double> ls(2 * (size_t)tr->nframes);
    HIPCHK(hipMemcpy(g.data(), tr->d_grad, sizeof(float) * g.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ls.data(), tr->d_loss, sizeof(double) * ls.size(), hipMemcpyDeviceToHost));
    double L = 0.0;
    for (int q = 0; q < nsel; ++q) L += ls[2 * frames[q]] + ls[2 * frames[q] + 1];
    *loss = L;
    long o = 0;
    for (int e = 0; e < M.nelem; ++e) {
        const int nin = M.nr[e] + M.na[e], n1 = M.nh1, n2 = M.nh2;
        const int nw = nin * n1 + n1 + n1 * n2 + n2 + n2 + 1;
        for (int q = 0; q < nw; ++q) grad[o + q] = g[(size_t)e * NN_NPAR + runner_to_padded(q, nin, n1, n2)];
        o += nw;
    }
    return 0;
}

// current weights in the weights.XXX.data layout, elements concatenated
int mdt_nn_get_weights(mdt_context *ctx, double *weights) {
    if (!ctx || !weights) return fail(MDT_EINVAL, "mdt_nn_get_weights: NULL argument");
    if (!ctx->have_nn) return fail(MDT_ESTATE, "no network");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpy(ctx->nn, ctx->d_nn, sizeof(NnConst), hipMemcpyDeviceToHost));
    const NnConst &M = *ctx->nn;
    long o = 0;
    for (int e = 0; e < M.nelem; ++e) {
        const int nin = M.nr[e] + M.na[e], n1 = M.nh1, n2 = M.nh2;
        for (int i = 0; i < nin; ++i)
            for (int j = 0; j < n1; ++j) weights[o++] = M.W1[e][i][j];
        for (int j = 0; j < n1; ++j) weights[o++] = M.b1[e][j];
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n2; ++j) weights[o++] = M.W2[e][i][j];
        for (int j = 0; j < n2; ++j) weights[o++] = M.b2[e][j];
        for (int j = 0; j < n2; ++j) weights[o++] = M.w3[e][j];
        weights[o++] = M.b3[e];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// RuNNer model files (read_nene, src/pes_nene_mod.f90:123-5759): input.nn
// keywords, scaling.data, weights.XXX.data; symmetry functions sorted as
// sortsymfunctions does (type, cut-off, eta, zeta, lambda, rshift, element 1,
// element 2; stable), elements by atomic number.
// ---------------------------------------------------------------------------
namespace {
struct SfDef {
    int type, e1, e2;  // atomic numbers
    double eta, rs, lam, zeta, rc;
};

int atomic_number(const std::string &sym) {
    static const char *tab[] = {"H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si",
                                "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni",
                                "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo",
                                "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba",
                                "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb",
                                "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi"};
    for (int z = 0; z < (int)(sizeof(tab) / sizeof(tab[0])); ++z)
        if (sym == tab[z]) return z + 1;
    return -1;
}

std::vector<std::string> words_of(const std::string &line) {
    std::string l = line.substr(0, line.find('#'));
    std::vector<std::string> w;
    size_t i = 0;
    while (i < l.size()) {
        while (i < l.size() && std::isspace((unsigned char)l[i])) ++i;
        size_t j = i;
        while (j < l.size() && !std::isspace((unsigned char)l[j])) ++j;
        if (j > i) w.push_back(l.substr(i, j - i));
        i = j;
    }
    return w;
}
}  // namespace

int mdt_nn_load(mdt_context *ctx, const char *directory, int ntypes, const char *type_symbols) {
    if (!ctx || !directory || !type_symbols) return fail(MDT_EINVAL, "mdt_nn_load: NULL argument");
    if (ntypes < 1 || ntypes > MDT_MAXT) return fail(MDT_EINVAL, "ntypes=%d", ntypes);
    const std::string dir = std::string(directory) + "/";
    FILE *fp = std::fopen((dir + "input.nn").c_str(), "r");
    if (!fp) return fail(MDT_EINVAL, "cannot open %sinput.nn", dir.c_str());
    std::vector<int> elems;
    std::vector<int> hidden;
    std::string act;
    int nh = -1, scale = 0, center = 0, ctype = 1;
    double scmin = 0.0, scmax = 1.0, calpha = 0.0;
    std::vector<std::pair<int, SfDef>> sfs;  // (centre Z, function)
    char buf[4096];
    while (std::fgets(buf, sizeof buf, fp)) {
        const std::vector<std::string> w = words_of(buf);
        if (w.empty()) continue;
        const std::string &k = w[0];
        auto num = [&](size_t i) { return i < w.size() ? std::atof(w[i].c_str()) : 0.0; };
        if (k == "elements") {
            for (size_t i = 1; i < w.size(); ++i) elems.push_back(atomic_number(w[i]));
        } else if (k == "global_hidden_layers_short") {
            nh = (int)num(1);
        } else if (k == "global_nodes_short") {
            for (size_t i = 1; i < w.size(); ++i) hidden.push_back(std::atoi(w[i].c_str()));
        } else if (k == "global_activation_short") {
            for (size_t i = 1; i < w.size(); ++i) act += w[i];
        } else if (k == "scale_symmetry_functions") {
            scale = 1;
        } else if (k == "center_symmetry_functions") {
            center = 1;
        } else if (k == "scale_min_short_atomic") {
            scmin = num(1);
        } else if (k == "scale_max_short_atomic") {
            scmax = num(1);
        } else if (k == "cutoff_type") {
            ctype = (int)num(1);
        } else if (k == "cutoff_alpha") {
            calpha = num(1);
        } else if (k == "symfunction_short" && w.size() >= 4) {
            SfDef f{};
            const int zc = atomic_number(w[1]);
            f.type = std::atoi(w[2].c_str());
            if (f.type == 2 && w.size() >= 7) {
                f.e1 = atomic_number(w[3]);
                f.eta = num(4);
                f.rs = num(5);
                f.rc = num(6);
            } else if (f.type == 3 && w.size() >= 9) {
                const int a = atomic_number(w[3]), b2 = atomic_number(w[4]);
                f.e1 = std::min(a, b2);
                f.e2 = std::max(a, b2);
                f.eta = num(5);
                f.lam = num(6);
                f.zeta = num(7);
                f.rc = num(8);
            } else {
                std::fclose(fp);
                return fail(MDT_EUNSUP, "input.nn: symmetry function type %d (2 and 3 are supported)", f.type);
            }
            sfs.push_back({zc, f});
        }
    }
    std::fclose(fp);
    std::sort(elems.begin(), elems.end());
    if (nh >= 0 && (int)hidden.size() > nh) hidden.resize(nh);
    if (hidden.size() != 2 || act.size() != 3)
        return fail(MDT_EUNSUP, "network must have two hidden layers and three activations");
    if (ctype != 1 || calpha != 0.0) return fail(MDT_EUNSUP, "only cutoff_type 1 with cutoff_alpha 0");
    const int ne = (int)elems.size();
    if (ne < 1 || ne > NN_MAXE) return fail(MDT_EUNSUP, "%d elements", ne);
    auto eidx = [&](int z) {
        for (int i = 0; i < ne; ++i)
            if (elems[i] == z) return i;
        return -1;
    };
    std::vector<std::vector<SfDef>> per(ne);
    for (auto &p : sfs) {
        const int e = eidx(p.first);
        if (e < 0) return fail(MDT_EINVAL, "input.nn: symmetry function of an element not in `elements`");
        per[e].push_back(p.second);
    }
    auto key_less = [](const SfDef &a, const SfDef &b) {
        if (a.type != b.type) return a.type < b.type;
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.eta != b.eta) return a.eta < b.eta;
        if (a.zeta != b.zeta) return a.zeta < b.zeta;
        if (a.lam != b.lam) return a.lam < b.lam;
        if (a.rs != b.rs) return a.rs < b.rs;
        if (a.e1 != b.e1) return a.e1 < b.e1;
        return a.e2 < b.e2;
    };
    std::vector<int32_t> nsf(ne), typ, e1, e2;
    std::vector<double> eta, rs, lam, zeta, rc;
    for (int e = 0; e < ne; ++e) {
        std::stable_sort(per[e].begin(), per[e].end(), key_less);
        nsf[e] = (int)per[e].size();
        for (const SfDef &f : per[e]) {
            typ.push_back(f.type);
            e1.push_back(eidx(f.e1));
            e2.push_back(f.type == 3 ? eidx(f.e2) : 0);
            if (e1.back() < 0 || (f.type == 3 && e2.back() < 0))
                return fail(MDT_EINVAL, "input.nn: neighbour element not in `elements`");
            eta.push_back(f.eta);
            rs.push_back(f.rs);
            lam.push_back(f.lam);
            zeta.push_back(f.zeta);
            rc.push_back(f.rc);
        }
    }
    // scaling.data
    std::vector<double> gmin, gmax, gavg;
    fp = std::fopen((dir + "scaling.data").c_str(), "r");
    if (!fp) return fail(MDT_EINVAL, "cannot open %sscaling.data", dir.c_str());
    for (int e = 0; e < ne; ++e)
        for (int q = 0; q < nsf[e]; ++q) {
            int ie, iq;
            double a, b2, c;
            if (std::fscanf(fp, "%d %d %lf %lf %lf", &ie, &iq, &a, &b2, &c) != 5 || ie != e + 1 || iq != q + 1) {
                std::fclose(fp);
                return fail(MDT_EINVAL, "scaling.data: bad line for element %d function %d", e + 1, q + 1);
            }
            gmin.push_back(a);
            gmax.push_back(b2);
            gavg.push_back(c);
        }
    std::fclose(fp);
    // weights.XXX.data
    std::vector<double> wts;
    for (int e = 0; e < ne; ++e) {
        char name[32];
        std::snprintf(name, sizeof name, "weights.%03d.data", elems[e]);
        fp = std::fopen((dir + name).c_str(), "r");
        if (!fp) return fail(MDT_EINVAL, "cannot open %s%s", dir.c_str(), name);
        const long nw = (long)nsf[e] * hidden[0] + hidden[0] + (long)hidden[0] * hidden[1] + hidden[1] + hidden[1] + 1;
        for (long q = 0; q < nw; ++q) {
            double x;
            if (std::fscanf(fp, "%lf", &x) != 1) {
                std::fclose(fp);
                return fail(MDT_EINVAL, "%s: %ld weights, the network needs %ld", name, q, nw);
            }
            wts.push_back(x);
        }
        std::fclose(fp);
    }
    // atom types -> elements (type_symbols: ntypes entries of 2 characters)
    std::vector<int32_t> eot(ntypes);
    for (int t = 0; t < ntypes; ++t) {
        std::string sym(type_symbols + 2 * t, 2);
        while (!sym.empty() && (sym.back() == ' ' || sym.back() == '\0')) sym.pop_back();
        eot[t] = eidx(atomic_number(sym));
    }
    return mdt_nn_params_set(ctx, ne, eot.data(), ntypes, nsf.data(), typ.data(), e1.data(), e2.data(), eta.data(),
                             rs.data(), lam.data(), zeta.data(), rc.data(), gmin.data(), gmax.data(), gavg.data(),
                             scmin, scmax, scale, center, hidden[0], hidden[1], act.c_str(), wts.data());
}

}  // extern "C"
