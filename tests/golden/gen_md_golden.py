"""Write tests/golden/md_golden.npz: ERIs and d/dr1 ERIs of selected shell
quartets of a distorted H2O/cc-pVDZ (Cartesian), by McMurchie-Davidson with an
mpmath Boys function (oracle/md_ref.py) -- an algorithm independent of Rys.

Run:  python tests/golden/gen_md_golden.py   (about 30 s)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from ccsd_gpu_amd.gto import M  # noqa: E402
from oracle import md_ref  # noqa: E402

GEOM = "O 0 0 0; H 0 0.96 0.1; H 0.93 -0.05 0.2"
QUARTETS = [(0, 0, 0, 0), (4, 2, 3, 1), (4, 4, 4, 4), (0, 5, 6, 10), (2, 3, 9, 7),
            (1, 0, 4, 0), (4, 10, 2, 0), (10, 4, 4, 10), (3, 8, 4, 2), (9, 4, 7, 4)]


def main():
    mol = M(GEOM, basis="cc-pvdz", cart=True)
    out = {"geom": np.array(GEOM), "quartets": np.array(QUARTETS)}
    for q, shls in enumerate(QUARTETS):
        out[f"int2e_{q}"] = md_ref.shell_quartet(mol, shls, 0)[0]
        out[f"ip1_{q}"] = md_ref.shell_quartet(mol, shls, 1)
    np.savez_compressed(os.path.join(HERE, "md_golden.npz"), **out)


if __name__ == "__main__":
    main()
