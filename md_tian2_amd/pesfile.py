"""Reader for md_tian2 ``.pes`` files (EMT, NENE and NON blocks).

Format (src/md_init.f90:99-144, src/pes_emt_mod.f90:86-188): a block starts
with ``pes emt``, followed by ``A B proj|latt proj|latt`` and key/value lines
``eta2 n0 E0 lambda V0 kappa s0`` (case-insensitive).  In a mixed block the
first seven values belong to element A, the next seven to element B
(pes_emt_mod.f90:133-179).  A blank line ends the block.  Each block marks
the pair (A, B) as interacting through EMT (pes_emt_mod.f90:111-112).

``pes nene`` blocks (read_nene, src/pes_nene_mod.f90) mark the pair as
interacting through the network; their key/value lines name the RuNNer
directory (``inp_dir``) and the extrapolation-warning limit
(``maxnum_extrapolation_warnings``).  ``pes non`` blocks mark a pair that
does not interact (src/pes_non_mod.f90).  Any other PES (LJ, Morse, HO,
REBO) is refused: the GPU path has no kernel for it.
"""
from __future__ import annotations

from dataclasses import dataclass, field

EMT_KEYS = ("eta2", "n0", "e0", "lambda", "v0", "kappa", "s0")
PES_ID_EMT = 2003  # constants.f90:66


@dataclass
class EmtPes:
    """EMT parameters per (element, role) and the set of EMT pairs."""

    params: dict = field(default_factory=dict)  # (name, is_proj) -> {key: val}
    pairs: set = field(default_factory=set)  # {((nameA, projA), (nameB, projB))}
    nene_pairs: set = field(default_factory=set)  # pairs of `pes nene` blocks
    non_pairs: set = field(default_factory=set)   # pairs of `pes non` blocks
    nene_dir: str = None                          # inp_dir of the nene blocks
    nene_max_warnings: int = None                 # maxnum_extrapolation_warnings

    def species(self):
        return list(self.params.keys())


def parse_pes(text: str) -> EmtPes:
    """Parse every ``pes emt`` block of a .pes file.

    Errors mirror the reference's stops: wrong header length
    (pes_emt_mod.f90:90), bad role keyword (:106), redefinition (:100-103),
    non key/value lines (:124-125) and unknown keys (:181-183).
    """
    out = EmtPes()
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        words = lines[i].split()
        i += 1
        if len(words) < 2 or words[0].lower() != "pes":
            continue
        kind = words[1].lower()
        if kind not in ("emt", "nene", "non"):
            raise ValueError(f"unsupported pes block '{words[1]}' (emt, nene and non are supported)")
        if i >= len(lines):
            raise ValueError("pes block without interaction line")
        hdr = lines[i].split()
        i += 1
        if len(hdr) != 4:
            raise ValueError("need four entries in interaction-defining lines")
        if hdr[2] not in ("proj", "latt") or hdr[3] not in ("proj", "latt"):
            raise ValueError("interaction must be defined via 'proj' and 'latt' keywords")
        a = (hdr[0], hdr[2] == "proj")
        b = (hdr[1], hdr[3] == "proj")
        key = tuple(sorted([a, b]))
        if key in out.pairs or key in out.nene_pairs or key in out.non_pairs:
            raise ValueError(f"pes already defined for atoms {hdr}")
        if kind != "emt":
            (out.nene_pairs if kind == "nene" else out.non_pairs).add(key)
            while i < len(lines):
                kv = lines[i].split()
                i += 1
                if len(kv) == 0:
                    break
                if kind == "nene" and len(kv) == 2 and kv[0].lower() == "inp_dir":
                    out.nene_dir = kv[1].strip("'\"")
                elif kind == "nene" and len(kv) == 2 and kv[0].lower() == "maxnum_extrapolation_warnings":
                    out.nene_max_warnings = int(kv[1])
                else:
                    raise ValueError(f"unknown {kind} parameter line '{' '.join(kv)}'")
            continue
        out.pairs.add(key)
        counter = 1
        while i < len(lines):
            kv = lines[i].split()
            i += 1
            if len(kv) == 0:
                break
            if len(kv) != 2:
                raise ValueError("PES parameters must consist of key value pairs")
            k = kv[0].lower()
            if k not in EMT_KEYS:
                raise ValueError(f"unknown EMT parameter {kv[0]}")
            who = a if counter <= 7 else b
            out.params.setdefault(who, {})[k] = float(kv[1])
            counter += 1
    return out


def load_pes(path: str) -> EmtPes:
    with open(path) as fh:
        return parse_pes(fh.read())


def emt_tables(pes: EmtPes, type_names):
    """Arrays (eta2, n0, e0, lambda, v0, kappa, s0) and the pes matrix for
    a list of types ``[(name, is_proj), ...]`` in universe order."""
    nt = len(type_names)
    cols = {k: [0.0] * nt for k in EMT_KEYS}
    pesm = [[0] * nt for _ in range(nt)]
    for t, tn in enumerate(type_names):
        par = pes.params.get(tn)
        if par is not None:
            for k in EMT_KEYS:
                cols[k][t] = par[k]
    for ti, a in enumerate(type_names):
        for tj, b in enumerate(type_names):
            if tuple(sorted([a, b])) in pes.pairs:
                pesm[ti][tj] = PES_ID_EMT
    return cols, pesm
