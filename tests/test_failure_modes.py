"""Failure behaviour of the boundary, on the CPU (no kernel is launched):

  * angular momentum above 4 is rejected with CCSD_ERR_UNSUPPORTED (-3), before any
    device call (the reference gates on it in _check_mol, ccsd_gpu/cuda/int2e_ip1_gpu.py:36-40);
  * malformed shapes are rejected with CCSD_ERR_SHAPE (-2);
  * solve_lambda_gpu raises without a device, with fallback_to_cpu False as the
    reference does (ccsd_gpu/gpu_ccsd_lambda.py:427-428) and also with True (this build
    has no CPU path);
  * frozen-orbital masks are validated.
"""
import ctypes

import numpy as np
import pytest
import torch

from ccsd_gpu_amd import _lib
from ccsd_gpu_amd.gto import M

ERR_NULL, ERR_SHAPE, ERR_UNSUPPORTED = -1, -2, -3


def _raw(mol):
    atm = np.ascontiguousarray(mol._atm, dtype=np.int32)
    bas = np.ascontiguousarray(mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(mol._env, dtype=np.float64)
    return atm, bas, env


def _create(atm, bas, env):
    h = ctypes.c_void_p()
    rc = _lib.load().ccsd_plan_create(ctypes.byref(h), atm.ctypes.data, atm.shape[0], bas.ctypes.data,
                                      bas.shape[0], env.ctypes.data, env.size, None)
    return rc, _lib.load().ccsd_last_error().decode()


def test_angular_momentum_above_four_rejected():
    from ccsd_gpu_amd.hip.int2e_gpu import _check_mol
    mol = M("O 0 0 0; H 0 0.8 0.5", basis="sto-3g")
    assert _check_mol(mol)
    atm, bas, env = _raw(mol)
    bas = bas.copy()
    bas[0, 1] = 5                      # ANG_OF of the first shell: an h shell
    rc, msg = _create(atm, bas, env)
    assert rc == ERR_UNSUPPORTED and "angular momentum" in msg


@pytest.mark.parametrize("field,value", [(2, 0), (5, 10 ** 6), (0, 7)])
def test_plan_rows_outside_env_rejected(field, value):
    mol = M("O 0 0 0; H 0 0.8 0.5", basis="sto-3g")
    atm, bas, env = _raw(mol)
    bas = bas.copy()
    bas[1, field] = value              # NPRIM_OF = 0, PTR_EXP past env, ATOM_OF past natm
    rc, msg = _create(atm, bas, env)
    assert rc == ERR_SHAPE, (rc, msg)


def test_gemv_shape_errors():
    lib = _lib.load()
    buf = (ctypes.c_double * 16)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    # nslices < 1, lda shorter than a row, beta not 0/1, and a null input
    assert lib.ccsd_gemv(0, p, 4, 4, 4, p, 1, p, 1.0, 0, None, 0, None) == ERR_SHAPE
    assert lib.ccsd_gemv(0, p, 4, 4, 3, p, 1, p, 1.0, 0, None, 1, None) == ERR_SHAPE
    assert lib.ccsd_gemv(0, p, 4, 4, 4, p, 1, p, 1.0, 2, None, 1, None) == ERR_SHAPE
    assert lib.ccsd_gemv(0, None, 4, 4, 4, p, 1, p, 1.0, 0, None, 1, None) == ERR_NULL
    assert b"GEMV" in lib.ccsd_last_error() or b"null" in lib.ccsd_last_error()


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device path")
@pytest.mark.parametrize("fallback", [False, True])
def test_solve_lambda_without_device_raises(fallback):
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu

    class _CC:
        max_cycle, conv_tol_normt = 10, 1e-8

    with pytest.raises(RuntimeError, match="No usable HIP device"):
        solve_lambda_gpu(_CC(), fallback_to_cpu=fallback)


def test_frozen_mask_convention():
    from ccsd_gpu_amd.cc import frozen_mask
    assert frozen_mask(5, None).all() and frozen_mask(5, 0).all()
    assert frozen_mask(5, 2).tolist() == [False, False, True, True, True]
    assert frozen_mask(5, [0, 4]).tolist() == [False, True, True, True, False]
    with pytest.raises(ValueError):
        frozen_mask(5, [7])
    with pytest.raises(ValueError):
        frozen_mask(5, 9)


def test_pair_form_and_tuning_entries_reject_bad_arguments():
    """Argument checks of the vvvv pair-form kernels, the FP32 repack and the GEMM
    block-shape switch (all before any device work)."""
    lib = _lib.load()
    buf = (ctypes.c_double * 16)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    # np must be v (v + 1) / 2, ld >= np, npair <= o (o + 1) / 2; split needs the lo plane
    assert lib.ccsd_vvvv_pm_pack(p, 2, 3, p, p, 3, p, p, 5, 6, 0, p, None, None) == ERR_SHAPE
    assert lib.ccsd_vvvv_pm_pack(p, 2, 3, p, p, 3, p, p, 6, 5, 0, p, None, None) == ERR_SHAPE
    assert lib.ccsd_vvvv_pm_pack(p, 2, 3, p, p, 4, p, p, 6, 6, 0, p, None, None) == ERR_SHAPE
    assert lib.ccsd_vvvv_pm_pack(p, 2, 3, p, p, 3, p, p, 6, 8, 1, p, None, None) == ERR_NULL
    assert lib.ccsd_vvvv_pm_unpack(p, 2, 3, 5, 1.0, p, None) == ERR_SHAPE
    assert lib.ccsd_vvvv_pm_unpack(None, 2, 3, 6, 1.0, p, None) == ERR_NULL
    assert lib.ccsd_repack_pairs_f32(None, 2, 3, p, None) == ERR_NULL
    assert lib.ccsd_gemm_set_variant(3) == ERR_SHAPE and lib.ccsd_gemm_set_variant(-1) == ERR_SHAPE
    assert lib.ccsd_gemm_set_variant(0) == 0
