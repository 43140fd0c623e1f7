"""Initial-condition sampling (CPU: the sampler is host code of the library).

* The oracle's apply_pip and set_proj_incidence (restatements taking the
  uniforms explicitly) are pinned by hand-computed cases: all-fixed pip puts
  the centroid on the given point; `pip r r z` in a 10 A cubic cell with
  uniforms (0.25, 0.5) puts it at (250, 50, z) -- the reference scales the
  whole target vector by each random component's column sum -- and a fixed
  x with a random y adds both; the incidence energy of the centroid equals
  the drawn energy, polar 0 points down, bead velocity offsets survive.
* The library's sampler equals the oracle fed with the sampler's own
  uniforms, for many trajectory ids.
* The sampler's draws have the intended distributions: Einc normal with
  sigma, a uniform random azimuth, uniform configuration indices."""
import math

import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, pesfile, sampling


def _proj_universe(P=3, cell=None, rng=None):
    rng = rng or np.random.default_rng(0)
    cell = np.eye(3) * 10.0 if cell is None else cell
    r = np.array([1.0, 2.0, 3.0]) + rng.normal(0, 0.1, (1, P, 3))
    v = rng.normal(0, 0.01, (1, P, 3))
    s = geometry.System([("H", True)], np.zeros(1, dtype=np.int32), np.array([1.00794 * geometry.AMU2MASS]),
                        np.zeros((1, P, 3), dtype=bool), np.array([geometry.PROP_VERLET], dtype=np.int32), cell,
                        r, v)
    return s


def _oracle(s):
    return oracle.OracleState(s, {k: [0.0] for k in pesfile.EMT_KEYS}, [[0]])


def test_oracle_pip_closed_forms():
    s = _proj_universe()
    off = s.r[0] - s.r[0].mean(axis=0)
    st = _oracle(s)
    st.apply_pip([0, 0, 0], [4.0, 5.0, 6.0], [0, 0, 0])
    assert np.allclose(st.r[0].mean(axis=0), [4, 5, 6], atol=1e-13)
    assert np.allclose(st.r[0] - st.r[0].mean(axis=0), off, atol=1e-13)
    st = _oracle(s)
    st.apply_pip([1, 1, 0], [0, 0, 7.0], [0.25, 0.5, 0])
    assert np.allclose(st.r[0].mean(axis=0), [250.0, 50.0, 7.0], atol=1e-12)
    st = _oracle(s)
    st.apply_pip([0, 1, 0], [1.5, 0, 3.0], [0, 0.5, 0])
    assert np.allclose(st.r[0].mean(axis=0), [1.5, 50.0, 3.0], atol=1e-12)


def test_oracle_incidence_closed_forms():
    s = _proj_universe(P=4)
    m = s.mass[0]
    off = s.v[0] - s.v[0].mean(axis=0)
    st = _oracle(s)
    st.set_proj_incidence(2.0, 0.0, 0.0, False, 0.0, 0.0, [0.3, 0.6, 0.5, 0.5])
    vc = st.v[0].mean(axis=0)
    assert np.allclose(vc, [0, 0, -math.sqrt(4.0 / m)], atol=1e-14)
    assert np.allclose(st.v[0] - vc, off, atol=1e-14)
    st = _oracle(s)
    st.set_proj_incidence(1.0, 0.2, 45.0, False, 90.0, 0.0, [0.3, 0.1, 0.7, 0.2])
    e = 1.0 + 0.2 * math.sqrt(-2 * math.log(0.3)) * math.cos(2 * math.pi * 0.1)
    vc = st.v[0].mean(axis=0)
    assert 0.5 * m * (vc @ vc) == pytest.approx(e, rel=1e-13)
    assert abs(vc[0]) < 1e-15 and vc[1] == pytest.approx(-vc[2], rel=1e-13)


def test_sampler_equals_oracle_with_its_uniforms():
    cell = np.array([[8.65, 4.325, 0.0], [0.0, 7.49, 0.0], [0.0, 0.0, 40.0]])
    for tid in range(1, 40):
        for seed in (2, 777):
            s = _proj_universe(P=4, cell=cell, rng=np.random.default_rng(tid))
            pip = sampling.Pip.parse(["r", "r", "6.5"])
            r = s.r.copy()
            sampling.apply_pip(r, cell, pip, seed, tid)
            st = _oracle(s)
            U = [sampling.uniform(seed, tid, 2 + k) for k in range(3)]
            st.apply_pip([1, 1, 0], [0, 0, 6.5], U)
            assert np.abs(r - st.r).max() <= 1e-12 * max(1.0, np.abs(st.r).max())
            for az in ("r", "30"):
                inc = sampling.Incidence.make(1.92, 50.0, az, sigma_einc=0.1, sigma_azimuth=5.0)
                v = s.v.copy()
                sampling.set_proj_incidence(v, s.idx, s.is_proj, s.mass, inc, seed, tid)
                st = _oracle(s)
                rnd = [1.0 - sampling.uniform(seed, tid, 5), sampling.uniform(seed, tid, 6),
                       sampling.uniform(seed, tid, 7), sampling.uniform(seed, tid, 8)]
                if az != "r":
                    rnd[2] = 1.0 - rnd[2]
                st.set_proj_incidence(1.92, 0.1, 50.0, az == "r", 0.0 if az == "r" else 30.0, 5.0, rnd)
                assert np.abs(v - st.v).max() <= 1e-14, (tid, seed, az)


def test_sampler_distributions():
    n = 20000
    s = _proj_universe(P=1)
    m = s.mass[0]
    inc = sampling.Incidence.make(2.0, 45.0, "r", sigma_einc=0.25)
    E, phi = [], []
    for tid in range(n):
        v = np.zeros((1, 1, 3))
        sampling.set_proj_incidence(v, s.idx, s.is_proj, s.mass, inc, 9, tid)
        E.append(0.5 * m * (v[0, 0] @ v[0, 0]))
        phi.append(math.atan2(v[0, 0, 1], v[0, 0, 0]) % (2 * math.pi))
    E, phi = np.array(E), np.array(phi)
    assert abs(E.mean() - 2.0) < 4 * 0.25 / math.sqrt(n)
    assert abs(E.std() - 0.25) < 0.01
    hist, _ = np.histogram(phi, bins=8, range=(0, 2 * math.pi))
    assert (np.abs(hist - n / 8) < 5 * math.sqrt(n / 8)).all()
    k = np.array([sampling.conf_index(9, tid, 1, 7) for tid in range(n)])
    assert k.min() == 1 and k.max() == 7
    hist = np.bincount(k)[1:]
    assert (np.abs(hist - n / 7) < 5 * math.sqrt(n / 7)).all()
    assert all(sampling.conf_index(9, t, 0, 5) == sampling.conf_index(9, t, 0, 5) for t in range(10))
