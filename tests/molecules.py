"""Molecules of the reference's tests (tests/test_int2e_gpu.py:20-96,
tests/test_int2e_ip1_gpu.py:20-98, tests/test_gpu_ccsd_grad.py:140-157)."""
from ccsd_gpu_amd.gto import M

CH4 = """
    C  0.000  0.000  0.000
    H  0.629  0.629  0.629
    H -0.629 -0.629  0.629
    H -0.629  0.629 -0.629
    H  0.629 -0.629 -0.629
"""
H2O_REF = "O 0 0 0; H 0 0.96 0; H 0.96 0 0"


def h2_sto3g(cart=True):
    return M("H 0 0 0; H 0 0 0.74", basis="sto-3g", cart=cart)


def ch4_sto3g(cart=True):
    return M(CH4, basis="sto-3g", cart=cart)


def h2o_ccpvdz(cart=False):
    return M(H2O_REF, basis="cc-pvdz", cart=cart)


def ch4_ccpvdz(cart=False):
    return M(CH4, basis="cc-pvdz", cart=cart)


def h2o_ccpvtz(cart=False):
    return M(H2O_REF, basis="cc-pvtz", cart=cart)


def h2o_def2svp(cart=False):
    return M(H2O_REF, basis="def2-svp", cart=cart)


def h2o_ccpvqz(cart=False):
    """O carries a g shell: int2e_ip1 with nine Rys roots."""
    return M(H2O_REF, basis="cc-pvqz", cart=cart)


def hf_sto3g(cart=False):
    return M("H 0 0 0; F 0 0 1.1", basis="sto-3g", cart=cart)


def g_shell_mol(cart=False):
    """Two centres with one g shell (l=4): covers nroots up to 9 for int2e_ip1."""
    basis = {
        "O": [(0, [5.0, 1.0], [[0.4, 0.7]]), (4, [1.3], [[1.0]])],
        "H": [(0, [1.2], [[1.0]]), (1, [0.8], [[1.0]])],
    }
    return M("O 0 0 0; H 0.3 0.8 0.5", basis=basis, cart=cart)
