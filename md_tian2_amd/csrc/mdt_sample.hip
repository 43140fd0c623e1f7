// mdt_sample.hip -- initial conditions of a trajectory (host code).
//
// The reference draws a trajectory's initial conditions from Fortran's
// random_number: the configuration files of `conf merge`
// (prepare_next_traj, src/md_init.f90:904-926), the projectile position of
// `pip` (apply_pip, :498-554) and its incidence velocity (set_proj_incidence,
// :829-877, with normal_deviate, src/useful_things.f90:121-131).  Here every
// draw is a Philox4x32-10 uniform keyed by (seed, trajectory id) on its own
// counter slot (MDT_SLOT_*), so a trajectory's initial conditions depend on
// nothing but its id and the seed: the Fortran and Python hosts get the same
// universes, in any batch order.  The formulas are the reference's,
// including apply_pip's scaling of the whole target vector by the column sum
// of the cell for each random component.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/mdt_gpu.h"
#include "mdt_philox.h"

namespace {

constexpr uint32_t STREAM_SAMPLE = 4u;  // beside the propagators' streams 0-3
constexpr double PI = 3.14159265358979323846;

// slot k of trajectory traj: uniform in [0, 1)
double uni(uint32_t seed, uint32_t traj, uint32_t slot) {
    return mdt_uniform(seed, traj, 0u, STREAM_SAMPLE, slot, 0u);
}

// normal_deviate (useful_things.f90:121-131): sqrt(-2 ln u1) cos(2 pi u2) sigma + mu
// with u1 in (0, 1]
double normal(uint32_t seed, uint32_t traj, uint32_t slot, double mu, double sigma) {
    const double u1 = 1.0 - uni(seed, traj, slot), u2 = uni(seed, traj, slot + 1);
    return sqrt(-2 * log(u1)) * cos(2 * PI * u2) * sigma + mu;
}

}  // namespace

extern "C" {

double mdt_sample_uniform(uint32_t seed, uint32_t traj_id, uint32_t slot) { return uni(seed, traj_id, slot); }

int mdt_sample_conf_index(uint32_t seed, uint32_t traj_id, int which, int n) {
    if (n < 1 || (which != 0 && which != 1)) return 0;
    const double rnd = uni(seed, traj_id, which == 0 ? MDT_SLOT_PROJ_FILE : MDT_SLOT_LATT_FILE);
    return (int)(rnd * n) + 1;  // int(rnd * nconfs) + 1
}

int mdt_apply_pip(const mdt_pip *pip, uint32_t seed, uint32_t traj_id, int natoms, int nbeads,
                  const double *simbox, double *r) {
    if (!pip || !simbox || !r || natoms < 1 || nbeads < 1) return MDT_EINVAL;
    // centroids before any change
    double *cent = new double[3 * natoms];
    for (int i = 0; i < natoms; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int b = 0; b < nbeads; ++b) s += r[k + 3 * (b + nbeads * i)];
            cent[k + 3 * i] = nbeads > 1 ? s / nbeads : s;
        }
    double target[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
        if (pip->random[k]) {
            target[k] = uni(seed, traj_id, MDT_SLOT_PIP + (uint32_t)k);
            const double cs = simbox[0 + 3 * k] + simbox[1 + 3 * k] + simbox[2 + 3 * k];
            for (int j = 0; j < 3; ++j) target[j] = target[j] * cs;
        } else {
            for (int i = 0; i < natoms; ++i)
                for (int b = 0; b < nbeads; ++b) r[k + 3 * (b + nbeads * i)] += pip->value[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        const double fp = simbox[k + 0] * target[0] + simbox[k + 3] * target[1] + simbox[k + 6] * target[2];
        for (int i = 0; i < natoms; ++i)
            for (int b = 0; b < nbeads; ++b) {
                double &x = r[k + 3 * (b + nbeads * i)];
                x = x - cent[k + 3 * i] + fp;
            }
    }
    delete[] cent;
    return MDT_OK;
}

int mdt_set_proj_incidence(const mdt_incidence *inc, uint32_t seed, uint32_t traj_id, int natoms, int nbeads,
                           const int32_t *idx, const int32_t *is_proj, const double *mass, double *v) {
    if (!inc || !idx || !is_proj || !mass || !v || natoms < 1 || nbeads < 1) return MDT_EINVAL;
    const double drawn = normal(seed, traj_id, MDT_SLOT_EINC, inc->einc, inc->sigma_einc);
    double pm = 0.0;
    for (int i = 0; i < natoms; ++i)
        if (is_proj[idx[i] - 1]) pm += mass[i];
    if (!(pm > 0.0)) return MDT_EINVAL;
    double az;
    if (inc->azimuth_random) {
        az = 2 * PI * uni(seed, traj_id, MDT_SLOT_AZIMUTH);
    } else {
        az = normal(seed, traj_id, MDT_SLOT_AZIMUTH, inc->azimuth_deg, inc->sigma_azimuth);
        az = az * (PI / 180.0);
    }
    const double polar = inc->polar_deg * (PI / 180.0);
    const double vinc = sqrt(2 * drawn / pm);
    const double vx = vinc * sin(polar) * cos(az), vy = vinc * sin(polar) * sin(az), vz = -vinc * cos(polar);
    for (int i = 0; i < natoms; ++i) {
        if (!is_proj[idx[i] - 1]) break;  // projectiles are always at the beginning
        double vcm[3];
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int b = 0; b < nbeads; ++b) s += v[c + 3 * (b + nbeads * i)];
            vcm[c] = s / nbeads;
        }
        for (int b = 0; b < nbeads; ++b) {
            double *p = v + 3 * (b + nbeads * i);
            p[0] = p[0] - vcm[0] + vx;
            p[1] = p[1] - vcm[1] + vy;
            p[2] = p[2] - vcm[2] + vz;
        }
    }
    return MDT_OK;
}

int mdt_sample_maxwell(double temperature, uint32_t seed, uint32_t traj_id, int natoms, int nbeads,
                       const int32_t *idx, const int32_t *is_proj, const double *mass, const int32_t *is_fixed,
                       double *v) {
    if (!idx || !is_proj || !mass || !is_fixed || !v || natoms < 1 || nbeads < 1) return MDT_EINVAL;
    for (int i = 0; i < natoms; ++i) {
        if (is_proj[idx[i] - 1]) continue;
        const double sig = temperature > 0.0 ? sqrt(8.61733238496e-5 * temperature / mass[i]) : 0.0;
        for (int c = 0; c < 3; ++c) {
            const double x = normal(seed, traj_id, MDT_SLOT_MAXWELL + 2u * (uint32_t)(3 * i + c), 0.0, sig);
            for (int b = 0; b < nbeads; ++b)
                v[c + 3 * (b + nbeads * i)] = is_fixed[c + 3 * (b + nbeads * i)] ? 0.0 : x;
        }
    }
    return MDT_OK;
}

}  // extern "C"
