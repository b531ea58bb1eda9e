"""Writes the deck fixtures under tests/golden/decks from the example decks the
reference ships with their outputs (output/poro_mec/elastico/hexa).

For each case it stores
  <case>.tar.gz  the deck and every file it names (insert / setpnode), as is;
  <case>_ref.npz the reference's own results for it: the printnode
                 histories of each print node, <kind>_<node> for kind up
                 (istep, t, u, v, w, p), stress / stressE / stressB (istep,
                 t, xx yy zz xy yz xz) and flux (istep, t, qx qy qz), and the
                 nnode / neq / nad the reference logged in <case>_t_1.txt.
These are data (inputs and the reference's outputs); nothing of the
reference's code is stored.  Run from the repository root:
    python tests/golden/make_deck_fixtures.py
"""
import io
import os
import re
import tarfile

import numpy as np

SRC = "/root/reference/output/poro_mec/elastico/hexa"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")
CASES = ("exemplo1_mesh0_t2", "solo1", "solo_cilindro_hexa_1")


def named_files(deck):
    names = [os.path.basename(deck)]
    for line in open(deck, errors="replace"):
        w = line.split()
        if len(w) >= 2 and w[0] in ("insert", "setpnode"):
            names.append(w[1])
    return names


def read_hist(path):
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    return np.array([[float(v.replace("D", "E")) for v in r] for r in rows])


def main():
    os.makedirs(OUT, exist_ok=True)
    for case in CASES:
        deck = os.path.join(SRC, case + ".dat")
        names = named_files(deck)
        with tarfile.open(os.path.join(OUT, case + ".tar.gz"), "w:gz") as tar:
            for n in names:
                data = open(os.path.join(SRC, n), "rb").read()
                info = tarfile.TarInfo(n)
                info.size = len(data)
                tar.addfile(info, io.BytesIO(data))
        arrays = {}
        for fn in sorted(os.listdir(SRC)):
            m = re.fullmatch(re.escape(case) + r"_(up|stress|stressE|stressB|flux)_node_(\d+)\.txt",
                             fn)
            if m:
                arrays[f"{m.group(1)}_{m.group(2)}"] = read_hist(os.path.join(SRC, fn))
        for line in open(os.path.join(SRC, case + "_t_1.txt")):
            w = line.split()
            if len(w) == 2 and w[0] in ("nnode", "neq", "nad", "numel"):
                arrays[w[0]] = np.array(int(w[1]))
        np.savez_compressed(os.path.join(OUT, case + "_ref.npz"), **arrays)
        print(case, names, sorted(arrays))


if __name__ == "__main__":
    main()
