"""Writes tests/golden/conf_hau: an H projectile and a 3x3x4 Au(111) lattice
(bottom 2 layers fixed, Maxwell-Boltzmann velocities at 300 K) as mxt
files, the inputs of the driver's `conf merge` test on the GPU."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from md_tian2_amd import conf, geometry  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "conf_hau")


def main():
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=2, rng=np.random.default_rng(2024), tsurf=300.0)
    latt = conf.Conf(s.r[1:].copy(), s.v[1:].copy(), True, s.fixed[1:].copy(), np.ones(s.natoms - 1, int),
                     ["Au"], [False], s.simbox.copy(), np.linalg.inv(s.simbox))
    proj = conf.Conf(s.r[:1].copy(), np.zeros((1, 1, 3)), True, np.zeros((1, 1, 3), bool), np.ones(1, int),
                     ["H"], [True], s.simbox.copy(), np.linalg.inv(s.simbox))
    for name, c in (("latt", latt), ("proj", proj)):
        os.makedirs(os.path.join(OUT, name), exist_ok=True)
        conf.write_mxt(conf.mxt_name(os.path.join(OUT, name), 1), c)


if __name__ == "__main__":
    main()
