"""Writes grad_golden.npz: FP64 oracle CCSD gradients of seeded water systems.

Run in the build container (CPU only): python tests/golden/gen_grad_golden.py [names...]

For each system the oracle (oracle/chem.py: RHF, CCSD, lambda, RDMs, Z-vector,
gradient, restating ccsd_gpu/gpu_ccsd_lambda.py and gpu_ccsd_grad.py) runs from
scratch; the fixture keeps the geometry (Bohr), the basis, E(HF), E(CCSD corr)
and the total dE/dR (natm, 3). The GPU tests rebuild the molecule from the stored
geometry, run their whole pipeline and compare dE/dR at 1e-6 Eh/Bohr. The
oracle gradient itself is pinned by central finite differences of E(CCSD) on
small molecules (tests/test_oracle_chem.py).

Systems: "water_s0" (one H2O, seed 0, cc-pVDZ: the single-water configuration) and
"h2o4_s1" ((H2O)4, seed 1, cc-pVDZ: 96 AOs, nocc 20, nvir 76).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ccsd_gpu_amd import systems  # noqa: E402
from oracle import chem  # noqa: E402

SYSTEMS = {
    "water_s0": lambda: systems.water(seed=0, basis="cc-pvdz"),
    "h2o4_s1": lambda: systems.water_cluster(4, seed=1, basis="cc-pvdz"),
}
OUT = os.path.join(HERE, "grad_golden.npz")


def run(name):
    mol = SYSTEMS[name]()
    t0 = time.perf_counter()
    e_tot, mf, cc = chem.ccsd_total_energy(mol)
    conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, cc.ao2mo())
    assert conv, name
    de = chem.grad_elec(cc, cc.t1, cc.t2, l1, l2) + chem.grad_nuc(mol)
    print(f"{name}: nao {mol.nao_nr()} E_hf {mf.e_tot:.12f} E_corr {cc.e_corr:.12f} "
          f"|sum dE/dR| {np.abs(de.sum(0)).max():.1e} ({time.perf_counter() - t0:.0f} s)", flush=True)
    return {f"{name}/coords": mol.atom_coords(), f"{name}/symbols": np.array([mol.atom_symbol(i)
                                                                              for i in range(mol.natm)]),
            f"{name}/basis": np.array("cc-pvdz"), f"{name}/e_hf": np.array(mf.e_tot),
            f"{name}/e_corr": np.array(cc.e_corr), f"{name}/grad": de}


def main():
    names = sys.argv[1:] or list(SYSTEMS)
    data = dict(np.load(OUT)) if os.path.exists(OUT) else {}
    for n in names:
        data.update(run(n))
    np.savez(OUT, **data)
    print("wrote", OUT, sorted(data))


if __name__ == "__main__":
    main()
