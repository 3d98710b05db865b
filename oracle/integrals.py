"""ctypes front end of the C oracle (rys_oracle.c): int2e / int2e_ip1 in s2kl layout.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

`intor_int2e(mol, b0, b1)` and `intor_int2e_ip1(mol, b0, b1)` return what
PySCF's `mol.intor('int2e[_ip1]', [comp=3,] aosym='s2kl',
shls_slice=(b0, b1, 0, nbas, 0, nbas, 0, nbas))` returns, reshaped as the
reference's GPU wrappers return it: (nf, nao, nao_pair) and
(3, nf, nao, nao_pair) (ccsd_gpu/cuda/int2e_gpu.py:186-215,
ccsd_gpu/cuda/int2e_ip1_gpu.py:296-326). Spherical output is the Cartesian
result transformed with mol.cart2sph_coeff() on all four indices, the same
arithmetic as int2e_ip1_gpu.py:223-289.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.oracle_int2e_s2kl_cart.restype = ctypes.c_int
        L.oracle_int2e_s2kl_cart.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.oracle_rys_roots.restype = None
        L.oracle_rys_roots.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_num_threads.restype = None
        L.oracle_set_num_threads.argtypes = [ctypes.c_int]
        L.oracle_int1e_cart.restype = ctypes.c_int
        L.oracle_int1e_cart.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _LIB = L
    return _LIB


def rys_roots(n, T):
    r = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_rys_roots(n, float(T), r.ctypes.data, w.ctypes.data)
    return r, w


def num_threads():
    return lib().oracle_num_threads()


def set_num_threads(n):
    lib().oracle_set_num_threads(int(n))


def int2e_s2kl_cart(mol, b0, b1, deriv):
    """Cartesian (ncomp, nf_cart, nao_cart, npair_cart) for i-shells [b0, b1)."""
    atm = np.ascontiguousarray(mol._atm, dtype=np.int32)
    bas = np.ascontiguousarray(mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(mol._env, dtype=np.float64)
    loc = mol.ao_loc_nr(cart=True)
    nao = int(loc[-1])
    nf = int(loc[b1] - loc[b0])
    npair = nao * (nao + 1) // 2
    ncomp = 3 if deriv else 1
    out = np.zeros((ncomp, nf, nao, npair))
    rc = lib().oracle_int2e_s2kl_cart(out.ctypes.data, int(deriv), atm.ctypes.data, mol.natm,
                                      bas.ctypes.data, mol.nbas, env.ctypes.data, int(b0), int(b1))
    if rc != 0:
        raise ValueError(f"oracle_int2e_s2kl_cart failed with code {rc}")
    return out


def _unpack_kl(x, n):
    """(..., n(n+1)/2) lower-triangle packed -> (..., n, n) symmetric."""
    k, l = np.tril_indices(n)
    full = np.zeros(x.shape[:-1] + (n, n))
    full[..., k, l] = x
    full[..., l, k] = x
    return full


def _pack_kl(x):
    n = x.shape[-1]
    k, l = np.tril_indices(n)
    return x[..., k, l]


def cart_to_sph_s2kl(out_cart, mol, b0, b1):
    """Apply cart2sph to all four indices of (ncomp, nf_c, nao_c, npair_c)."""
    C = mol.cart2sph_coeff()
    lc = mol.ao_loc_nr(cart=True)
    ls = mol.ao_loc_nr(cart=False)
    Ci = C[lc[b0]:lc[b1], ls[b0]:ls[b1]]
    nao_c = C.shape[0]
    full = _unpack_kl(out_cart, nao_c)                       # (c, i, j, k, l)
    t = np.einsum("cijkl,ia,jb,kx,ly->cabxy", full, Ci, C, C, C, optimize=True)
    return _pack_kl(t)


def _intor(mol, b0, b1, deriv):
    if b1 is None:
        b1 = mol.nbas
    out = int2e_s2kl_cart(mol, b0, b1, deriv)
    if not mol.cart:
        out = cart_to_sph_s2kl(out, mol, b0, b1)
    return out


def intor_int2e(mol, b0=0, b1=None):
    return _intor(mol, b0, b1, 0)[0]


def intor_int2e_ip1(mol, b0=0, b1=None):
    return _intor(mol, b0, b1, 1)


def int2e_full(mol):
    """(nao, nao, nao, nao) chemist's-notation ERIs in the native basis."""
    n = mol.nao_nr()
    return _unpack_kl(intor_int2e(mol), n)


def ip1_full(mol):
    """(3, nao, nao, nao, nao) d/dr1 ERIs in the native basis."""
    n = mol.nao_nr()
    return _unpack_kl(intor_int2e_ip1(mol), n)


INT1E_KINDS = {"ovlp": 0, "kin": 1, "nuc": 2, "ipovlp": 3, "ipkin": 4, "ipnuc_atom": 5}


def int1e(mol, kind):
    """One-electron integrals in the native AO basis (PySCF `int1e_*` conventions).

    ovlp / kin / nuc: (nao, nao); ipovlp / ipkin: (3, nao, nao) = <d/dr i| O |j>;
    ipnuc_atom: (natm, 3, nao, nao) = <d/dr i| -Z_C/|r-C| |j> for every nucleus C
    (its sum over C is PySCF's int1e_ipnuc; row C times -1/Z_C is int1e_iprinv at C).
    """
    k = INT1E_KINDS[kind]
    atm = np.ascontiguousarray(mol._atm, dtype=np.int32)
    bas = np.ascontiguousarray(mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(mol._env, dtype=np.float64)
    nc = mol.nao_cart()
    nmat = {0: 1, 1: 1, 2: 1, 3: 3, 4: 3, 5: 3 * mol.natm}[k]
    out = np.zeros((nmat, nc, nc))
    rc = lib().oracle_int1e_cart(out.ctypes.data, k, atm.ctypes.data, mol.natm, bas.ctypes.data, mol.nbas,
                                 env.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle_int1e_cart failed with code {rc}")
    if not mol.cart:
        C = mol.cart2sph_coeff()
        out = np.einsum("mij,ia,jb->mab", out, C, C, optimize=True)
    if k <= 2:
        return out[0]
    if k <= 4:
        return out
    n = out.shape[-1]
    return out.reshape(mol.natm, 3, n, n)
