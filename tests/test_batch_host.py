"""Host side of the batched plans: the concatenated libcint arrays."""
import numpy as np

from ccsd_gpu_amd.batch import conc_env
from ccsd_gpu_amd.systems import cnoh_batch


def test_conc_env_offsets_point_at_each_molecules_data():
    mols = cnoh_batch(4, seed=1)
    atm, bas, env, mol_nbas = conc_env(mols)
    assert list(mol_nbas) == [m.nbas for m in mols]
    a0 = s0 = 0
    for m in mols:
        for a in range(m.natm):
            p = atm[a0 + a, 1]
            assert np.array_equal(env[p:p + 3], m._env[m._atm[a, 1]:m._atm[a, 1] + 3])
        for s in range(m.nbas):
            row, ref = bas[s0 + s], m._bas[s]
            assert row[0] == ref[0] + a0
            assert np.array_equal(row[1:4], ref[1:4])
            np_, nc = ref[2], ref[3]
            assert np.array_equal(env[row[5]:row[5] + np_], m._env[ref[5]:ref[5] + np_])
            assert np.array_equal(env[row[6]:row[6] + np_ * nc], m._env[ref[6]:ref[6] + np_ * nc])
        a0 += m.natm
        s0 += m.nbas
