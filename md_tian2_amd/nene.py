"""NENE (RuNNer high-dimensional neural network) model files.

Reads and writes the three files md_tian2's read_nene loads
(src/pes_nene_mod.f90:123-5759): ``input.nn`` (keywords), ``scaling.data``
(per element and symmetry function: min, max, average; last line the
short-range energy range) and ``weights.XXX.data`` (one weight per line,
XXX the atomic number, src/pes_nene_mod.f90:6930-6975).

Symmetry functions are held in the order the reference uses after
sortsymfunctions (src/RuNNer_src/sortsymfunctions.f90): ascending by
(type, cut-off, eta, zeta, lambda, rshift, element 1, element 2), stable;
``scaling.data`` and the weights index them in that order.  Elements are
ordered by atomic number (sortelements).  For type 3 the two neighbour
elements are stored with the lighter first (src/pes_nene_mod.f90:3797-3801).
Units are the network's: Bohr and Hartree.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

ATOMIC_NUMBER = {"H": 1, "He": 2, "Li": 3, "C": 6, "N": 7, "O": 8, "Al": 13, "Si": 14, "Ni": 28, "Cu": 29,
                 "Zn": 30, "Ag": 47, "Pd": 46, "Pt": 78, "Au": 79}
SYMBOL = {v: k for k, v in ATOMIC_NUMBER.items()}
BOHR2ANG = 0.529177211        # src/constants.f90:160
HA2EV = 27.21138602           # src/constants.f90:165
HABOHR2EVANG = HA2EV / BOHR2ANG


@dataclass
class SymFunc:
    type: int          # 2 radial, 3 angular
    e1: int            # neighbour element (atomic number)
    e2: int = 0        # second neighbour element (type 3), e1 <= e2
    eta: float = 0.0
    rs: float = 0.0
    lam: float = 0.0
    zeta: float = 0.0
    rc: float = 0.0    # Bohr

    def key(self):
        # sortsymfunctions.f90 passes: type, cutoff, eta, zeta, lambda, rshift, elem1, elem2
        return (self.type, self.rc, self.eta, self.zeta, self.lam, self.rs, self.e1, self.e2)

    def line(self, center):
        c = SYMBOL[center]
        if self.type == 2:
            return f"symfunction_short {c} 2 {SYMBOL[self.e1]} {self.eta:.10g} {self.rs:.10g} {self.rc:.10g}"
        return (f"symfunction_short {c} 3 {SYMBOL[self.e1]} {SYMBOL[self.e2]} {self.eta:.10g} "
                f"{self.lam:.10g} {self.zeta:.10g} {self.rc:.10g}")


@dataclass
class NeneModel:
    elements: list                      # atomic numbers, ascending
    sf: dict                            # Z -> list[SymFunc] (sorted)
    hidden: list                        # nodes per hidden layer
    activations: str                    # one letter per weight layer, e.g. "ttl"
    gmin: dict = field(default_factory=dict)   # Z -> array
    gmax: dict = field(default_factory=dict)
    gavg: dict = field(default_factory=dict)
    eshort: tuple = (0.0, 0.0)
    weights: dict = field(default_factory=dict)  # Z -> flat array, RuNNer order
    scale: bool = True
    center: bool = False
    scmin: float = 0.0
    scmax: float = 1.0
    cutoff_type: int = 1
    cutoff_alpha: float = 0.0

    def nsf(self, z):
        return len(self.sf[z])

    def nodes(self, z):
        return [self.nsf(z)] + list(self.hidden) + [1]

    def nweights(self, z):
        n = self.nodes(z)
        return sum(n[i] * n[i + 1] + n[i + 1] for i in range(len(n) - 1))

    def layers(self, z):
        """[(W (nin, nout), b (nout,)), ...] views of element z's weights."""
        n = self.nodes(z)
        w = self.weights[z]
        out, o = [], 0
        for i in range(len(n) - 1):
            W = w[o:o + n[i] * n[i + 1]].reshape(n[i], n[i + 1])
            o += n[i] * n[i + 1]
            b = w[o:o + n[i + 1]]
            o += n[i + 1]
            out.append((W, b))
        return out

    def rcmax(self):
        return max(f.rc for z in self.elements for f in self.sf[z])

    def sort(self):
        for z in self.elements:
            self.sf[z] = sorted(self.sf[z], key=SymFunc.key)


def _words(line):
    line = line.split("#", 1)[0].strip()
    return line.split() if line else []


def read_input_nn(path):
    elements, sf, hidden, act, nh = [], {}, [], "", None
    opts = dict(scale=False, center=False, scmin=0.0, scmax=1.0, cutoff_type=1, cutoff_alpha=0.0)
    for raw in open(path):
        w = _words(raw)
        if not w:
            continue
        k = w[0]
        if k == "elements":
            elements = sorted(ATOMIC_NUMBER[s] for s in w[1:])
        elif k == "global_hidden_layers_short":
            nh = int(w[1])
        elif k == "global_nodes_short":
            hidden = [int(x) for x in w[1:]]
        elif k == "global_activation_short":
            act = "".join(w[1:])
        elif k == "scale_symmetry_functions":
            opts["scale"] = True
        elif k == "center_symmetry_functions":
            opts["center"] = True
        elif k == "scale_min_short_atomic":
            opts["scmin"] = float(w[1])
        elif k == "scale_max_short_atomic":
            opts["scmax"] = float(w[1])
        elif k == "cutoff_type":
            opts["cutoff_type"] = int(w[1])
        elif k == "cutoff_alpha":
            opts["cutoff_alpha"] = float(w[1])
        elif k == "symfunction_short":
            zc = ATOMIC_NUMBER[w[1]]
            t = int(w[2])
            if t == 2:
                f = SymFunc(2, ATOMIC_NUMBER[w[3]], 0, float(w[4]), float(w[5]), 0.0, 0.0, float(w[6]))
            elif t == 3:
                a, b = sorted((ATOMIC_NUMBER[w[3]], ATOMIC_NUMBER[w[4]]))
                f = SymFunc(3, a, b, float(w[5]), 0.0, float(w[6]), float(w[7]), float(w[8]))
            else:
                raise ValueError(f"symmetry function type {t} is not supported (types 2 and 3 are)")
            sf.setdefault(zc, []).append(f)
    hidden = hidden[:nh] if nh is not None else hidden
    if len(act) != len(hidden) + 1:
        raise ValueError(f"global_activation_short {act!r} does not match {len(hidden)} hidden layers")
    for a in act:
        if a not in "tls":
            raise ValueError(f"activation {a!r} is not supported (t, l, s are)")
    if opts["cutoff_type"] != 1:
        raise ValueError("only cutoff_type 1 (cosine) is supported")
    m = NeneModel(elements, {z: sf.get(z, []) for z in elements}, hidden, act, **opts)
    m.sort()
    return m


def read_scaling(m, path):
    lines = [ln.split() for ln in open(path) if ln.strip()]
    k = 0
    for ie, z in enumerate(m.elements):
        n = m.nsf(z)
        mn, mx, av = np.zeros(n), np.zeros(n), np.zeros(n)
        for i in range(n):
            w = lines[k]
            k += 1
            if int(w[0]) != ie + 1 or int(w[1]) != i + 1:
                raise ValueError(f"scaling.data line {k}: expected element {ie + 1} function {i + 1}")
            mn[i], mx[i], av[i] = float(w[2]), float(w[3]), float(w[4])
        m.gmin[z], m.gmax[z], m.gavg[z] = mn, mx, av
    if k < len(lines) and len(lines[k]) >= 2:
        m.eshort = (float(lines[k][0]), float(lines[k][1]))


def weights_name(z):
    return f"weights.{z:03d}.data"


def read_weights(m, directory):
    for z in m.elements:
        w = np.array([float(ln.split()[0]) for ln in open(os.path.join(directory, weights_name(z))) if ln.strip()])
        if len(w) < m.nweights(z):
            raise ValueError(f"{weights_name(z)}: {len(w)} weights, the network needs {m.nweights(z)}")
        m.weights[z] = w[:m.nweights(z)].copy()


def load(directory, input_nn="input.nn"):
    m = read_input_nn(os.path.join(directory, input_nn))
    read_scaling(m, os.path.join(directory, "scaling.data"))
    read_weights(m, directory)
    return m


def write(m, directory):
    """input.nn (the keywords read_nene needs), scaling.data, weights.XXX.data."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "input.nn"), "w") as f:
        f.write("nn_type_short 1\nrunner_mode 3\n")
        f.write(f"number_of_elements {len(m.elements)}\n")
        f.write("elements " + " ".join(SYMBOL[z] for z in m.elements) + "\n")
        f.write("use_short_nn\n")
        f.write(f"global_hidden_layers_short {len(m.hidden)}\n")
        f.write("global_nodes_short " + " ".join(str(h) for h in m.hidden) + "\n")
        f.write("global_activation_short " + " ".join(m.activations) + "\n")
        f.write(f"cutoff_type {m.cutoff_type}\n")
        if m.scale:
            f.write("scale_symmetry_functions\n")
        if m.center:
            f.write("center_symmetry_functions\n")
        f.write(f"scale_min_short_atomic {m.scmin}\nscale_max_short_atomic {m.scmax}\n")
        for z in m.elements:
            for s in m.sf[z]:
                f.write(s.line(z) + "\n")
        f.write("calculate_forces\n")
    with open(os.path.join(directory, "scaling.data"), "w") as f:
        for ie, z in enumerate(m.elements):
            for i in range(m.nsf(z)):
                f.write(f"{ie + 1:4d} {i + 1:4d} {m.gmin[z][i]:18.9f} {m.gmax[z][i]:18.9f} {m.gavg[z][i]:18.9f}\n")
        f.write(f"{m.eshort[0]:20.10f} {m.eshort[1]:20.10f}\n")
    for z in m.elements:
        with open(os.path.join(directory, weights_name(z)), "w") as f:
            for x in m.weights[z]:
                f.write(f"{x:18.10f}\n")


def read_input_data(path, limit=None):
    """RuNNer input.data structures: list of dicts with lattice (3,3) rows or
    None, symbols, xyz (Bohr), forces (Ha/Bohr), energy (Ha)."""
    out, cur = [], None
    for raw in open(path):
        w = raw.split()
        if not w:
            continue
        if w[0] == "begin":
            cur = dict(lattice=[], symbols=[], xyz=[], forces=[], energy=None)
        elif w[0] == "lattice":
            cur["lattice"].append([float(x) for x in w[1:4]])
        elif w[0] == "atom":
            cur["xyz"].append([float(x) for x in w[1:4]])
            cur["symbols"].append(w[4])
            cur["forces"].append([float(x) for x in w[7:10]])
        elif w[0] == "energy":
            cur["energy"] = float(w[1])
        elif w[0] == "end":
            cur["lattice"] = np.array(cur["lattice"]) if len(cur["lattice"]) == 3 else None
            cur["xyz"] = np.array(cur["xyz"])
            cur["forces"] = np.array(cur["forces"])
            out.append(cur)
            if limit and len(out) >= limit:
                break
    return out


def hau_design(rc_ang=6.5, hidden=(32, 32)):
    """The H/Au surrogate's symmetry functions: per element 24 radial + 16
    angular, one cut-off rc_ang (6.5 A), network 40-32-32-1 tanh/tanh/linear.

    H:  24 radial on Au neighbours (8 centred, 16 shifted over 2.5-11.5 Bohr);
    Au: 18 radial on Au (6 centred, 12 shifted over 4-11.5 Bohr) and 6 centred
        radial on H neighbours;
    both: 16 angular on Au-Au pairs, eta in {0.001, 0.01, 0.03, 0.07} Bohr^-2,
        zeta in {1, 4}, lambda = +-1.
    """
    H, AU = ATOMIC_NUMBER["H"], ATOMIC_NUMBER["Au"]
    rc = rc_ang / BOHR2ANG
    ang = [SymFunc(3, AU, AU, eta, 0.0, lam, zeta, rc)
           for eta in (0.001, 0.01, 0.03, 0.07) for zeta in (1.0, 4.0) for lam in (-1.0, 1.0)]
    sf_h = [SymFunc(2, AU, 0, eta, 0.0, 0, 0, rc) for eta in np.geomspace(0.002, 0.5, 8)]
    sh = np.linspace(2.5, 11.5, 16)
    sf_h += [SymFunc(2, AU, 0, 1.0 / (2 * (sh[1] - sh[0]) ** 2), rs, 0, 0, rc) for rs in sh]
    sf_au = [SymFunc(2, AU, 0, eta, 0.0, 0, 0, rc) for eta in np.geomspace(0.002, 0.2, 6)]
    sa = np.linspace(4.0, 11.5, 12)
    sf_au += [SymFunc(2, AU, 0, 1.0 / (2 * (sa[1] - sa[0]) ** 2), rs, 0, 0, rc) for rs in sa]
    sf_au += [SymFunc(2, H, 0, eta, 0.0, 0, 0, rc) for eta in (0.003, 0.01, 0.03, 0.1, 0.3, 1.0)]
    m = NeneModel([H, AU], {H: sf_h + list(ang), AU: sf_au + [SymFunc(**vars(s)) for s in ang]},
                  list(hidden), "ttl", scale=True, center=False, scmin=0.0, scmax=1.0)
    m.sort()
    return m


def hau_radial_design(rc=4.5, hidden=(32, 32), nh=16, nau=12, nauh=4):
    """A radial-only H/Au surrogate: nh radial symmetry functions on H and
    nau + nauh on Au, cut-off rc (A), network n-32-32-1 tanh/tanh/linear.

    EMT's atomic energy is a function of distance sums only: the neighbour
    density sigma_1 and the pair sum sigma_2 are sums over neighbours of
    chi exp(-eta (r - beta s0)) theta(r) and chi exp(-kappa/beta (r - beta
    s0)) theta(r) per neighbour element (src/pes_emt_mod.f90:646-763), and
    E_i is built from them alone (:793-976).  A radial fingerprint of each
    neighbour element therefore holds everything the EMT energy depends on,
    and angular functions add nothing for an EMT surrogate.

    H:  nh on Au neighbours, a quarter centred (eta 0.002-0.5 Bohr^-2), the
        rest shifted Gaussians over 2 to 11.5 rc / 6.5 Bohr;
    Au: nau on Au, a quarter centred (eta 0.002-0.2), the rest shifted over
        4 to 11.5 rc / 6.5 Bohr; nauh centred on H (eta 0.003-1).
    The defaults (16, 12 + 4, 4.5 A) are md_tian2_amd/nn/hau_radial: on
    the 4- and 6-layer validation frames 0.44 meV/atom and 3.4 meV/A; 32 and
    24 + 8 functions at 6.5 A reach 1.0 meV/atom and 3.3 meV/A at 1.4x the
    cost (GPU training of each in about a minute).
    """
    H, AU = ATOMIC_NUMBER["H"], ATOMIC_NUMBER["Au"]
    rcb = rc / BOHR2ANG
    top = 11.5 * rc / 6.5  # Bohr

    def shifted(lo, n):
        sh = np.linspace(lo, top, n)
        return [SymFunc(2, AU, 0, 1.0 / (2 * (sh[1] - sh[0]) ** 2), rs, 0, 0, rcb) for rs in sh]

    sf_h = [SymFunc(2, AU, 0, eta, 0.0, 0, 0, rcb) for eta in np.geomspace(0.002, 0.5, nh // 4)]
    sf_h += shifted(2.0, nh - nh // 4)
    sf_au = [SymFunc(2, AU, 0, eta, 0.0, 0, 0, rcb) for eta in np.geomspace(0.002, 0.2, nau // 4)]
    sf_au += shifted(4.0, nau - nau // 4)
    sf_au += [SymFunc(2, H, 0, eta, 0.0, 0, 0, rcb) for eta in np.geomspace(0.003, 1.0, nauh)]
    m = NeneModel([H, AU], {H: sf_h, AU: sf_au}, list(hidden), "ttl", scale=True, center=False,
                  scmin=0.0, scmax=1.0)
    m.sort()
    return m


DESIGNS = {"hau": hau_design, "radial": hau_radial_design}


def init_weights(m, rng):
    """Uniform +-sqrt(6/(fan_in + fan_out)) weights, zero biases."""
    for z in m.elements:
        n = m.nodes(z)
        w = []
        for i in range(len(n) - 1):
            lim = np.sqrt(6.0 / (n[i] + n[i + 1]))
            w.append(rng.uniform(-lim, lim, n[i] * n[i + 1]))
            w.append(np.zeros(n[i + 1]))
        m.weights[z] = np.concatenate(w)


def set_scaling(m, G_by_z, margin=0.0):
    """min / max / average of each symmetry function over the training
    environments (G_by_z: Z -> array (n_env, nsf)); `margin` widens the
    range by that fraction of its width on both sides."""
    for z in m.elements:
        G = np.asarray(G_by_z[z], dtype=np.float64)[:, :m.nsf(z)]
        mn, mx = G.min(0), G.max(0)
        wd = np.maximum(mx - mn, 1e-12)
        m.gmin[z] = mn - margin * wd
        m.gmax[z] = mx + margin * wd
        m.gavg[z] = G.mean(0)
