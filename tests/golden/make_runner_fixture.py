"""Writes tests/golden/runner_c/ from the reference's RuNNer example
(example/md/RuNNer_input/ and example/fit/input.data of md_tian2): the
active keyword lines of input.nn (comments dropped), scaling.data,
weights.006.data, and the first 40 structures of input.data.  These are
data fixtures (a trained carbon network and DFT energies/forces in Bohr and
Hartree); run this once in a checkout that has the reference tree."""
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "runner_c")
NSTRUCT = 40


def main():
    src = os.path.join(REF, "example", "md", "RuNNer_input")
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(src, "input.nn")) as f, open(os.path.join(OUT, "input.nn"), "w") as g:
        for line in f:
            s = line.split("#", 1)[0].rstrip()
            if s.strip():
                g.write(s.strip() + "\n")
    for name in ("scaling.data", "weights.006.data"):
        with open(os.path.join(src, name)) as f, open(os.path.join(OUT, name), "w") as g:
            g.write(f.read())
    n = 0
    with open(os.path.join(REF, "example", "fit", "input.data")) as f, \
            open(os.path.join(OUT, "input.data"), "w") as g:
        for line in f:
            if line.startswith("#"):
                continue
            g.write(line)
            if line.startswith("end"):
                n += 1
                if n >= NSTRUCT:
                    break


if __name__ == "__main__":
    main()
