// mdt_noalloc.hip -- mdt_compute_emt_noalloc: the single-universe EMT
// drop-in (mdt_compute_emt) for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  Inputs and outputs are the caller's device arrays in the
// reference's layout; everything else the force kernel needs -- the
// universe in the kernels' structure-of-arrays layout, the replica tables,
// accelerations, densities, status and neighbour lists that do not fit in
// LDS -- lives in one file-scope device array, `scratch`, of fixed size.
//
// The array exists once per GPU the library runs on and is held until the
// process ends.  Its address on each device is taken once, at context
// creation with that device current (mdt_noalloc_init), with
// hipGetSymbolAddress; host code uses that address, never the array's name.
// Every call on a GPU shares the array and leaves its contents behind: the
// pack kernel rewrites everything a call reads (including the status word),
// so nothing depends on an earlier call, and calls that overlap on one GPU
// (other streams, other threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mdt_gpu.h"
#include "mdt_device.h"

// limits of one call (also in the header's comment)
#define NOALLOC_MAX_ATOMS 2048
#define NOALLOC_MAX_BEADS 64
#define NOALLOC_MAX_AB 16384  // atoms x beads
#define NOALLOC_NBCAP 96      // owned half-list capacity per atom

// byte layout of the scratch array (every block 256-byte aligned)
#define NA_ALIGN(x) (((x) + 255) & ~(size_t)255)
#define NA_X 0
#define NA_F (NA_X + NA_ALIGN(3 * 8 * (size_t)NOALLOC_MAX_AB))
#define NA_A (NA_F + NA_ALIGN(3 * 8 * (size_t)NOALLOC_MAX_AB))
#define NA_DENS (NA_A + NA_ALIGN(3 * 8 * (size_t)NOALLOC_MAX_AB))
#define NA_FIXED (NA_DENS + NA_ALIGN(8 * (size_t)NOALLOC_MAX_AB))
#define NA_TYPE (NA_FIXED + NA_ALIGN((size_t)NOALLOC_MAX_AB))
#define NA_MASS (NA_TYPE + NA_ALIGN((size_t)NOALLOC_MAX_ATOMS))
#define NA_EPOT (NA_MASS + NA_ALIGN(8 * (size_t)NOALLOC_MAX_ATOMS))
#define NA_REPTRAJ (NA_EPOT + NA_ALIGN(8 * (size_t)NOALLOC_MAX_BEADS))
#define NA_REPBEAD (NA_REPTRAJ + NA_ALIGN(4 * (size_t)NOALLOC_MAX_BEADS))
#define NA_CELL (NA_REPBEAD + NA_ALIGN(4 * (size_t)NOALLOC_MAX_BEADS))  // cell 9, icell 9, tsurf, tproj
#define NA_TRAJ (NA_CELL + NA_ALIGN(8 * 20))  // N, P, status, traj_id | soff, aoff, eoff (8-byte aligned)
#define NA_PC (NA_TRAJ + NA_ALIGN(64))
#define NA_NBR (NA_PC + NA_ALIGN(sizeof(PropConst)))
#define NA_TOTAL (NA_NBR + NA_ALIGN(2 * (size_t)NOALLOC_MAX_AB * NOALLOC_NBCAP))

static __device__ __align__(256) unsigned char scratch[NA_TOTAL];

namespace {

std::mutex g_mu;
void *g_ptr[64] = {};  // the array's address per device

// Stage the universe: r(3,P,N) column-major -> SoA [bead][component][atom],
// types 0-based, unit masses, nothing fixed, the cell, one trajectory of P
// replicas, the status word cleared.  Any lane count; grid-stride loops.
__global__ void k_noalloc_pack(unsigned char *s, int N, int P, const int32_t *idx, const double *simbox,
                               const double *isimbox, const double *r, int ntypes) {
    double *x = (double *)(s + NA_X);
    uint8_t *fixed = s + NA_FIXED;
    uint8_t *type = s + NA_TYPE;
    double *mass = (double *)(s + NA_MASS);
    int *rep_traj = (int *)(s + NA_REPTRAJ), *rep_bead = (int *)(s + NA_REPBEAD);
    double *cell = (double *)(s + NA_CELL);
    int *ti = (int *)(s + NA_TRAJ);
    long long *tl = (long long *)(s + NA_TRAJ + 16);
    const long long n3 = 3LL * P * N;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long long k = g; k < n3; k += stride) {  // k = c + 3 (b + P i) in the reference's layout
        const int c = (int)(k % 3), b = (int)((k / 3) % P);
        const long long i = k / (3LL * P);
        x[(long long)b * 3 * N + (long long)c * N + i] = r[k];
    }
    for (long long k = g; k < (long long)P * N; k += stride) fixed[k] = 0;
    for (long long i = g; i < N; i += stride) {
        const int t = idx[i] - 1;
        type[i] = (uint8_t)(t >= 0 && t < ntypes ? t : 0);
        mass[i] = 1.0;
        if (t < 0 || t >= ntypes) atomicOr(&ti[2], MDT_ST_BAD_PROP);  // a type outside 1..ntypes
    }
    for (long long b = g; b < P; b += stride) {
        rep_traj[b] = 0;
        rep_bead[b] = (int)b;
    }
    if (g < 9) {
        cell[g] = simbox[g];
        cell[9 + g] = isimbox[g];
    }
    for (long long k = g; k < (long long)(sizeof(PropConst) / 4); k += stride)
        ((uint32_t *)(s + NA_PC))[k] = 0u;  // no propagator data: calc_force only
    if (g == 0) {
        cell[18] = 0.0;  // tsurf
        cell[19] = 0.0;  // tproj
        ti[0] = N;
        ti[1] = P;
        ti[3] = 0;  // trajectory id
        tl[0] = tl[1] = tl[2] = 0;
    }
}

// The status word is cleared by a kernel of its own before the pack kernel
// (which may set it), so no hipMemsetAsync is recorded when a caller
// captures the call in a graph.
__global__ void k_noalloc_clear(unsigned char *s) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ((int *)(s + NA_TRAJ))[2] = 0;
}

// Outputs in the reference's layout: f(:,b,i) += F, epot(b) += E_b,
// dens(b,i) = density of the EMT atoms (src/pes_emt_mod.f90:855-861, 976,
// 1001-1011); status (when given) = the call's status bits.
__global__ void k_noalloc_unpack(const unsigned char *s, const EmtConst *ec, int N, int P, double *epot, double *f,
                                 double *dens, int32_t *status) {
    const double *F = (const double *)(s + NA_F);
    const double *D = (const double *)(s + NA_DENS);
    const uint8_t *type = s + NA_TYPE;
    const double *E = (const double *)(s + NA_EPOT);
    const long long n3 = 3LL * P * N;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long long k = g; k < n3; k += stride) {
        const int c = (int)(k % 3), b = (int)((k / 3) % P);
        const long long i = k / (3LL * P);
        f[k] += F[(long long)b * 3 * N + (long long)c * N + i];
    }
    for (long long k = g; k < (long long)P * N; k += stride) {  // k = b + P i
        const int b = (int)(k % P);
        const long long i = k / P;
        if (ec->owner[type[i]]) dens[k] = D[(long long)b * N + i];
    }
    for (long long b = g; b < P; b += stride) epot[b] += E[b];
    if (g == 0 && status) *status = ((const int *)(s + NA_TRAJ))[2];
}

}  // namespace

size_t mdt_replica_smem(int N, int cap, bool nbr_lds);
hipError_t mdt_launch_replica(int tpa, int nrep, int block, size_t smem, const EmtConst *ec, const PropConst *pc,
                              const BatchView &bv, int add_fin, hipStream_t st);
int mdt_noalloc_fail(int code, const char *msg);

// The array's address on the current device (taken once per device).
int mdt_noalloc_init(int device, void **ptr) {
    if (device < 0 || device >= 64) return mdt_noalloc_fail(MDT_EINVAL, "device index outside 0..63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return mdt_noalloc_fail(MDT_EHIP, hipGetErrorString(e));
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return 0;
}

size_t mdt_noalloc_bytes(void) { return NA_TOTAL; }

// The force launch of one universe on the scratch array at `ptr` (the
// device's address from mdt_noalloc_init); lds_limit is the device's LDS per
// workgroup.  Sizes were checked by the caller.
int mdt_noalloc_compute(void *ptr, size_t lds_limit, const EmtConst *d_ec, int ntypes, int N, int P, const int32_t *idx,
                        const double *simbox, const double *isimbox, const double *r, double *epot, double *f,
                        double *dens, int32_t *status, hipStream_t st) {
    unsigned char *s = (unsigned char *)ptr;
    int cap = NOALLOC_NBCAP < N - 1 ? NOALLOC_NBCAP : (N > 1 ? N - 1 : 1);
    int tpa = 8;
    while (tpa > 1 && N * tpa > 512) tpa >>= 1;
    const size_t lds_dyn = lds_limit > 5632 ? lds_limit - 5632 : 0;  // beside k_replica's static tables (<= 4.7 KB)
    const size_t smem_lds = mdt_replica_smem(N, cap, true);
    const bool lds = smem_lds <= lds_dyn;
    const size_t smem = lds ? smem_lds : mdt_replica_smem(N, cap, false);
    if (smem > lds_dyn) return mdt_noalloc_fail(MDT_EUNSUP, "universe too large for one workgroup's LDS");
    const int block = (((N * tpa + 63) / 64) * 64) < 64 ? 64 : (((N * tpa + 63) / 64) * 64);
    const long long n3 = 3LL * P * N;
    const int grid = (int)((n3 + 255) / 256 < 1024 ? (n3 + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(64), 0, st, s);
    hipLaunchKernelGGL(k_noalloc_pack, dim3(grid), dim3(256), 0, st, s, N, P, idx, simbox, isimbox, r, ntypes);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mdt_noalloc_fail(MDT_EHIP, hipGetErrorString(e));
    BatchView v{};
    v.ntraj = 1;
    v.nrep = P;
    v.nbcap = cap;
    v.precision = 0;  // compute_emt's arithmetic: FP64
    v.rep_traj = (const int *)(s + NA_REPTRAJ);
    v.rep_bead = (const int *)(s + NA_REPBEAD);
    v.traj_N = (const int *)(s + NA_TRAJ);
    v.traj_P = (const int *)(s + NA_TRAJ + 4);
    v.status = (int *)(s + NA_TRAJ + 8);
    v.traj_id = (const uint32_t *)(s + NA_TRAJ + 12);
    v.traj_soff = (const long long *)(s + NA_TRAJ + 16);
    v.traj_aoff = (const long long *)(s + NA_TRAJ + 24);
    v.traj_eoff = (const long long *)(s + NA_TRAJ + 32);
    v.type = s + NA_TYPE;
    v.mass = (const double *)(s + NA_MASS);
    v.fixed = s + NA_FIXED;
    v.cell = (const double *)(s + NA_CELL);
    v.icell = (const double *)(s + NA_CELL) + 9;
    v.tsurf = (const double *)(s + NA_CELL) + 18;
    v.tproj = (const double *)(s + NA_CELL) + 19;
    v.x = (double *)(s + NA_X);
    v.f = (double *)(s + NA_F);
    v.a = (double *)(s + NA_A);
    v.dens = (double *)(s + NA_DENS);
    v.epot = (double *)(s + NA_EPOT);
    v.nbr_global = lds ? nullptr : (uint16_t *)(s + NA_NBR);
    v.nbr_stride = lds ? 0 : (long long)N * cap;
    e = mdt_launch_replica(tpa, P, block, smem, d_ec, (const PropConst *)(s + NA_PC), v, 0, st);
    if (e != hipSuccess) return mdt_noalloc_fail(MDT_EHIP, hipGetErrorString(e));
    hipLaunchKernelGGL(k_noalloc_unpack, dim3(grid), dim3(256), 0, st, (const unsigned char *)s, d_ec, N, P, epot, f,
                       dens, status);
    e = hipGetLastError();
    if (e != hipSuccess) return mdt_noalloc_fail(MDT_EHIP, hipGetErrorString(e));
    return 0;
}

int mdt_noalloc_check_sizes(int natoms, int nbeads) {
    if (natoms < 1 || natoms > NOALLOC_MAX_ATOMS || nbeads < 1 || nbeads > NOALLOC_MAX_BEADS ||
        (long long)natoms * nbeads > NOALLOC_MAX_AB)
        return mdt_noalloc_fail(MDT_EINVAL,
                                "mdt_compute_emt_noalloc: natoms <= 2048, nbeads <= 64 and natoms * nbeads <= 16384");
    return 0;
}
