"""Writes grad_golden.npz: FP64 oracle CCSD gradients of seeded water systems.

Run in the build container (CPU only): python tests/golden/gen_grad_golden.py [names...]
(GOLDEN_CHECKPOINT=<file prefix> keeps the large system's converged amplitudes and
lambdas on disk, so an interrupted run resumes after its last finished stage).

For each system the oracle (oracle/chem.py: RHF, CCSD, lambda, RDMs, Z-vector,
gradient, restating ccsd_gpu/gpu_ccsd_lambda.py and gpu_ccsd_grad.py) runs from
scratch; the fixture keeps the geometry (Bohr), the basis, E(HF), E(CCSD corr)
and the total dE/dR (natm, 3). The GPU tests rebuild the molecule from the stored
geometry, run their whole pipeline and compare dE/dR at 1e-6 Eh/Bohr. The
oracle gradient itself is pinned by central finite differences of E(CCSD) on
small molecules (tests/test_oracle_chem.py).

Systems: "water_s0" (one H2O, seed 0, cc-pVDZ: the single-water configuration),
"h2o4_s1" ((H2O)4, seed 1, cc-pVDZ: 96 AOs, nocc 20, nvir 76) and "h2o8_s1"
((H2O)8, seed 1, cc-pVDZ: 192 AOs, nocc 40, nvir 152 -- the full-path benchmark
system), the last through the memory-lean oracle (oracle/chem_large.py, about
35 GB of host memory and half an hour on 8 cores).
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
    "h2o8_s1": lambda: systems.water_cluster(8, seed=1, basis="cc-pvdz"),
}
LARGE = {"h2o8_s1"}
OUT = os.path.join(HERE, "grad_golden.npz")


def run(name):
    mol = SYSTEMS[name]()
    t0 = time.perf_counter()
    if name in LARGE:
        from oracle import chem_large
        e_hf, e_corr, de = chem_large.ccsd_gradient(mol, log=lambda m: print(f"  {name}: {m}", flush=True),
                                                    checkpoint=os.environ.get("GOLDEN_CHECKPOINT"))
    else:
        e_tot, mf, cc = chem.ccsd_total_energy(mol)
        conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, cc.ao2mo())
        assert conv, name
        de = chem.grad_elec(cc, cc.t1, cc.t2, l1, l2) + chem.grad_nuc(mol)
        e_hf, e_corr = mf.e_tot, cc.e_corr
    print(f"{name}: nao {mol.nao_nr()} E_hf {e_hf:.12f} E_corr {e_corr:.12f} "
          f"|sum dE/dR| {np.abs(de.sum(0)).max():.1e} ({time.perf_counter() - t0:.0f} s)", flush=True)
    return {f"{name}/coords": mol.atom_coords(), f"{name}/symbols": np.array([mol.atom_symbol(i)
                                                                              for i in range(mol.natm)]),
            f"{name}/basis": np.array("cc-pvdz"), f"{name}/e_hf": np.array(e_hf),
            f"{name}/e_corr": np.array(e_corr), f"{name}/grad": de}


def main():
    names = sys.argv[1:] or [n for n in SYSTEMS if n not in LARGE]
    data = dict(np.load(OUT)) if os.path.exists(OUT) else {}
    for n in names:
        data.update(run(n))
    np.savez(OUT, **data)
    print("wrote", OUT, sorted(data))


if __name__ == "__main__":
    main()
