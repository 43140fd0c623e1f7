"""Pins of the oracle's propagators, RPMD step and RNG by mathematics that
does not restate their code:

* Langevin (md_algo.f90:146-300): at T = 0 and zero friction it is velocity
  Verlet; with zero noise and constant acceleration the exact branch is the
  exact Ornstein-Uhlenbeck solution; its noise amplitudes equal the
  Ito-integral variances computed by numerical quadrature; the series branch
  converges to the exact one as (xi dt)^3; both halves share one deviate;
* ldfa (md_algo.f90:506-570): numpy polyval of the published fit, the
  saturation branch, the single-precision 0.36 threshold, the stop on a
  negative density;
* cjk (rpmd.f90:36-59) is orthonormal for even P (and not for odd P);
* the ring-polymer step (rpmd.f90:309-452) equals expm of the free ring
  polymer's linear equations of motion; Langevin beads are reverted and get
  the spring force -grad of the ring energy with m(1), Tproj;
* PILE (md_algo.f90:440-502): damping exp(-dt gamma_k / 2) with gamma_0 =
  1/tau, gamma_k = 2 omega_k, and noise sqrt(m/beta_P) c2;
* remove_com_velocity (universe_mod.f90:222-241) equals its geometric-series
  closed form;
* Philox4x32-10 against the Random123 known-answer vectors, and the normal
  deviates against N(0,1) (KS test).
"""
import math

import numpy as np
import pytest
from scipy import integrate, linalg, stats

import oracle
from md_tian2_amd import geometry, pesfile

KB, HBAR, AMU2MASS = 8.61733238496e-5, 0.6582119514467406, 103.638239276


def free_system(n=2, p=1, algo=geometry.PROP_LANGEVIN, mass_amu=1.00794, is_proj=False, seed=0):
    g = np.random.default_rng(seed)
    s = geometry.System([("X", is_proj)], np.zeros(n, dtype=np.int32), np.full(n, mass_amu * AMU2MASS),
                        np.zeros((n, p, 3), dtype=bool), np.array([algo], dtype=np.int32),
                        np.eye(3) * 50.0, g.normal(0, 1, (n, p, 3)), g.normal(0, 0.01, (n, p, 3)))
    return s


def state(s, **kw):
    return oracle.OracleState(s, {k: [0.0] for k in pesfile.EMT_KEYS}, [[0]], **kw)


def test_langevin_zero_T_zero_friction_is_verlet():
    g = np.random.default_rng(1)
    s = free_system()
    lan = state(s, step=0.1, tsurf=0.0)
    ver = state(free_system(algo=geometry.PROP_VERLET), step=0.1)
    a = g.normal(0, 0.3, lan.a.shape)
    for st in (lan, ver):
        st.a[:] = a
        st.dens[:] = 0.0
    noise = g.normal(0, 1, lan.r.shape)
    lan.prop1(randy=noise)
    ver.prop1()
    assert np.allclose(lan.r, ver.r, rtol=0, atol=1e-14) and np.allclose(lan.v, ver.v, rtol=0, atol=1e-15)
    lan.prop2(randy=noise)
    ver.prop2()
    assert np.allclose(lan.v, ver.v, rtol=0, atol=1e-15)


def _ldfa_xi(rho, mass):
    s = free_system(n=1)
    s.mass[:] = mass
    st = state(s)
    st.dens[:] = rho
    assert st.ldfa(0) == 0
    return float(st.dens[0, 0])


def test_langevin_exact_branch_is_ornstein_uhlenbeck():
    dt, rho = 5.0, 0.5
    s = free_system(n=1)
    m = s.mass[0]
    xi = _ldfa_xi(rho, m)
    assert xi * dt > 0.01
    st = state(s, step=dt, tsurf=300.0)
    a = np.array([0.01, -0.02, 0.005])
    st.a[0, 0] = a
    r0, v0 = st.r[0, 0].copy(), st.v[0, 0].copy()
    st.dens[:] = xi
    st.prop1()
    e = math.exp(-xi * dt)
    r_exact = r0 + v0 * (1 - e) / xi + a / xi * (dt - (1 - e) / xi)
    assert np.allclose(st.r[0, 0], r_exact, rtol=1e-12, atol=1e-13)
    st.dens[:] = rho  # propagate_2 converts it with ldfa again
    st.prop2()
    v_exact = v0 * e + a * (1 - e) / xi
    assert np.allclose(st.v[0, 0], v_exact, rtol=1e-12, atol=1e-14)


def test_langevin_noise_amplitudes_match_ito_integrals():
    dt, rho, T = 5.0, 0.5, 300.0
    s = free_system(n=1)
    m = s.mass[0]
    xi = _ldfa_xi(rho, m)
    st = state(s, step=dt, tsurf=T)
    st.r[:] = 0.0
    st.v[:] = 0.0
    st.a[:] = 0.0
    st.dens[:] = xi
    ones = np.ones_like(st.r)
    st.prop1(randy=ones)
    sr, v1 = st.r[0, 0, 0], st.v[0, 0, 0]
    temp = KB * T / m
    var_r = 2 * xi * temp / xi ** 2 * integrate.quad(lambda u: (1 - math.exp(-xi * u)) ** 2, 0, dt,
                                                       epsabs=1e-14)[0]
    var_v = 2 * xi * temp * integrate.quad(lambda u: math.exp(-2 * xi * u), 0, dt, epsabs=1e-14)[0]
    cov = 2 * xi * temp / xi * integrate.quad(lambda u: (1 - math.exp(-xi * u)) * math.exp(-xi * u), 0, dt,
                                               epsabs=1e-14)[0]
    assert abs(sr - math.sqrt(var_r)) < 1e-9 * math.sqrt(var_r)
    assert abs(v1 - cov / math.sqrt(var_r)) < 1e-9 * abs(v1)
    st.dens[:] = rho
    st.prop2(randy=ones)
    # shared deviate (md_algo.f90:43-48, 292-293): second half adds sigma_v sqrt(1 - c_rv^2)
    c_rv = cov / math.sqrt(var_r * var_v)
    assert abs((st.v[0, 0, 0] - v1) - math.sqrt(var_v) * math.sqrt(1 - c_rv ** 2)) < 1e-9 * math.sqrt(var_v)


@pytest.mark.parametrize("xidt", [0.008, 0.002])
def test_langevin_series_branch_converges(xidt):
    dt = 0.1
    s = free_system(n=1)
    st = state(s, step=dt, tsurf=0.0)
    xi = xidt / dt
    a = np.array([0.01, -0.02, 0.005])
    st.a[0, 0] = a
    r0, v0 = st.r[0, 0].copy(), st.v[0, 0].copy()
    st.dens[:] = xi
    st.prop1()
    e = math.exp(-xi * dt)
    r_exact = r0 + v0 * (1 - e) / xi + a / xi * (dt - (1 - e) / xi)
    err = np.abs(st.r[0, 0] - r_exact).max()
    assert err < 5 * xidt ** 3 * dt * (np.abs(v0).max() + dt * np.abs(a).max())


def test_ldfa_polynomial_and_branches():
    coefs = [0.0802484, -1.12851, 9.28508, 2.10064, -843.419, 8.85354e3, -4.89023e4, 1.6741e5,
             -3.67098e5, 5.03476e5, -3.9426e5, 1.34763e5]
    m = 1.00794 * AMU2MASS  # the mass that makes convert/m = 1
    for rho in (0.0, 0.01, 0.05, 0.1, 0.2, 0.3, 0.36):
        hbar_xi = np.polyval(list(reversed(coefs)) + [0.0], rho)
        assert abs(_ldfa_xi(rho, m) - hbar_xi / HBAR) < 1e-12
    for rho in (0.4, 1.0):
        assert abs(_ldfa_xi(rho, m) - 0.001 * (4.7131 - math.exp(-4.41305 * rho)) / HBAR) < 1e-15
    # 0.36 is a single-precision literal: 0.36000001 still takes the polynomial
    rho = 0.3600000100
    assert abs(_ldfa_xi(rho, m) - np.polyval(list(reversed(coefs)) + [0.0], rho) / HBAR) < 1e-12
    s = free_system(n=1)
    st = state(s)
    st.dens[:] = -1e-6
    assert st.ldfa(0) != 0  # the reference stops (md_algo.f90:551-554)


@pytest.mark.parametrize("p", [2, 4, 8, 32])
def test_cjk_orthonormal_even(p):
    st = state(free_system(n=1, p=p))
    C = st.cjk.T  # C[j, k] = cjk(j+1, k+1)
    assert np.abs(C.T @ C - np.eye(p)).max() < 1e-13


def test_cjk_not_orthonormal_odd():
    st = state(free_system(n=1, p=3))
    C = st.cjk.T
    assert np.abs(C.T @ C - np.eye(3)).max() > 1e-3


def _ring_matrix(p, m, wp, dt):
    L = 2 * np.eye(p) - np.roll(np.eye(p), 1, axis=1) - np.roll(np.eye(p), -1, axis=1)
    A = np.zeros((2 * p, 2 * p))
    A[:p, p:] = np.eye(p) / m
    A[p:, :p] = -m * wp ** 2 * L
    return linalg.expm(A * dt)


@pytest.mark.parametrize("p", [4, 8])
def test_free_ring_polymer_is_exact(p):
    T, dt = 300.0, 0.1
    s = free_system(n=1, p=p, algo=geometry.PROP_VERLET, is_proj=True, seed=3)
    st = state(s, step=dt, tproj=T, tsurf=T)
    m = s.mass[0]
    q0, v0 = st.r[0].copy(), st.v[0].copy()
    st.ring_polymer_step()
    wp = p * KB * T / HBAR
    U = _ring_matrix(p, m, wp, dt)
    for c in range(3):
        y = U @ np.concatenate([q0[:, c], m * v0[:, c]])
        assert np.allclose(st.r[0, :, c], y[:p], rtol=1e-11, atol=1e-12)
        assert np.allclose(st.v[0, :, c], y[p:] / m, rtol=1e-11, atol=1e-13)


def test_langevin_beads_reverted_with_spring_force():
    p, Tp = 4, 200.0
    s = free_system(n=2, p=p, algo=geometry.PROP_LANGEVIN, seed=4)
    s.mass[1] *= 7.0
    st = state(s, step=0.1, tproj=Tp, tsurf=50.0)
    r0, v0 = st.r.copy(), st.v.copy()
    st.ring_polymer_step()
    assert np.array_equal(st.r, r0) and np.array_equal(st.v, v0)
    m1 = s.mass[0]  # rpmd.f90:326 uses atom 1's mass for every Langevin atom
    wp = p * KB * Tp / HBAR
    for i in range(2):
        def ring_energy(q):
            return 0.5 * m1 * wp ** 2 * ((q - np.roll(q, -1, axis=0)) ** 2).sum()
        h = 1e-6
        for b in range(p):
            for c in range(3):
                qp, qm = r0[i].copy(), r0[i].copy()
                qp[b, c] += h
                qm[b, c] -= h
                fd = -(ring_energy(qp) - ring_energy(qm)) / (2 * h)
                assert abs(st.f[i, b, c] - fd) < 1e-6 * max(1.0, abs(fd))


def test_pile_damping_and_noise():
    p, T, tau, dt = 8, 100.0, 200.0, 0.1
    s = free_system(n=1, p=p, algo=geometry.PROP_PILE, is_proj=True, seed=5)
    st = state(s, step=dt, tproj=T, tsurf=T, pile_tau=tau)
    m = s.mass[0]
    C = st.cjk.T
    v0 = st.v[0].copy()
    st.a[:] = 0.0
    st.prop1(zeta=np.zeros_like(st.r))
    wp = p * KB * T / HBAR
    gam = np.array([1 / tau] + [4 * wp * math.sin(k * math.pi / p) for k in range(1, p)])
    c1 = np.exp(-0.5 * dt * gam)
    # propagate_1 also removes the COM velocity afterwards: compare relative motion
    v_exp = C @ (c1[:, None] * (C.T @ v0))
    com = lambda v: v - v.mean(axis=0)
    assert np.allclose(com(st.v[0]), com(v_exp), rtol=1e-10, atol=1e-14)
    # noise: zeta = 1 in mode k adds sqrt(m/beta_P) c2_k / m in that mode
    st2 = state(s, step=dt, tproj=T, tsurf=T, pile_tau=tau)
    st2.a[:] = 0.0
    st2.v[:] = 0.0
    z = np.zeros_like(st2.r)
    z[0, 3, 1] = 1.0
    st2.prop1(zeta=z)
    betaN = 1 / (KB * T * p)
    dv = C @ (np.eye(p)[:, 3] * math.sqrt(m / betaN) * math.sqrt(1 - c1[3] ** 2)) / m
    assert np.allclose(st2.v[0, :, 1], dv - dv.mean() + st2.v[0, :, 1].mean(), rtol=1e-10, atol=1e-16)


def test_remove_com_closed_form():
    s = free_system(n=5, p=3, seed=6)
    s.mass[:] = np.linspace(1, 5, 5) * AMU2MASS
    st = state(s)
    v0 = st.v.copy()
    st.remove_com()
    S = (s.mass[:, None] * v0.sum(axis=1)).sum(axis=0)
    shift = S / (s.mass.sum() * 3) * (1 - (1 - 1 / 5) ** 1000)
    assert np.allclose(st.v, v0 - shift, rtol=0, atol=1e-15)


def test_andersen_resample():
    p, T, dt, tau = 2, 300.0, 0.1, 0.05
    s = free_system(n=3, p=p, algo=geometry.PROP_ANDERSEN, seed=7)
    st = state(s, step=dt, tsurf=T, andersen_time=tau)
    st.a[:] = 0.0
    zs = np.random.default_rng(8).normal(0, 1, st.r.shape)
    u = np.array([[0.1, 0.9], [0.99, 0.2], [0.5, 0.5]])  # thresh = dt/tau = 2 -> all resampled
    st.prop1(and_z=zs, and_u=u)
    sig = np.sqrt(p * KB * T / s.mass)[:, None, None]
    com = lambda v: v - (s.mass[:, None, None] * v).sum(axis=(0, 1)) / (s.mass.sum() * p)
    assert np.allclose(com(st.v), com(zs * sig), atol=1e-14)


KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(oracle.philox(ctr, key)) == out


def test_normals_are_standard_normal():
    z = oracle.philox_normals(11, 3, 7, oracle.STREAM_LANGEVIN, 20000, 1).reshape(-1)
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1) < 0.02
    assert stats.kstest(z, "norm").pvalue > 1e-3


def test_zero_temperature_trajectory_is_noise_free(pes_hau):
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(0), z0=2.5, thermal=False)
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    outs = []
    for seed in (1, 2):
        st = oracle.OracleState(s, cols, pesm, step=0.1, tsurf=0.0)
        st.calc_force()
        st.run(seed, 0, 1, 20)
        outs.append(st.r.copy())
    assert np.array_equal(outs[0], outs[1])
