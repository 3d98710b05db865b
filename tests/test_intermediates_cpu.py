"""Lambda intermediates of the product (its torch code run on the CPU) against the
oracle's restatement of ccsd_gpu/gpu_ccsd_lambda.py:68-209, on the water / cc-pVDZ
CCSD amplitudes of the oracle. Pins the product's refactorings of the reference's
products (t1 t1 terms through a rank-one factor, wvvov split into its parts) against
the reference's formulas."""
import types

import numpy as np
import pytest
import torch

from oracle import chem


@pytest.fixture(scope="module")
def water():
    from ccsd_gpu_amd.gto import M
    mol = M("O 0.0 -0.0757 0.0521; H 0.7572 0.5 -0.02; H -0.76 0.47 0.03", basis="sto-3g")
    mf = chem.RHF(mol)
    mf.kernel()
    cc = chem.CCSD(mf)
    cc.kernel()
    return cc


def test_intermediates_match_oracle(water, monkeypatch):
    import ccsd_gpu_amd.gpu_ccsd_lambda as L
    monkeypatch.setattr(L, "_dev", lambda x: torch.as_tensor(np.asarray(x)) if not torch.is_tensor(x) else x)
    cc = water
    eris = cc.ao2mo()
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a))  # noqa: E731
    e = types.SimpleNamespace(fock=T(eris.fock), mo_energy=T(eris.mo_energy), ovvv=T(eris.ovvv),
                              ovoo=T(eris.ovoo), ovvo=T(eris.ovvo), oovv=T(eris.oovv), oooo=T(eris.oooo))
    ref = chem.make_intermediates(cc.t1, cc.t2, eris)
    got = L.make_intermediates_gpu(None, T(cc.t1), T(cc.t2), e)
    for name in ("w1", "w2", "w3", "w4", "woooo", "wvooo", "wVOov", "wvOOv", "wvvov"):
        r = np.asarray(getattr(ref, name))
        g = getattr(got, name).numpy()
        err = np.abs(g - r).max()
        print(name, err, flush=True)
        assert err <= 1e-12 * max(1.0, np.abs(r).max()), (name, err)
