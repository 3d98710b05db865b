"""ctypes loader of the HIP library (libccsd_gpu_amd.so, C ABI in include/ccsd_gpu_amd.h).

There is no CPU fallback: if the library is missing or fails to load, every
entry point raises. Device buffers are torch tensors on the current HIP device;
calls are ordered on torch's current stream.
"""
import atexit
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccsd_gpu_amd.so")
_LIB = None

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_SIGNATURES = {
    "ccsd_last_error": (ctypes.c_char_p, []),
    "ccsd_plan_create": (c_int, [ctypes.POINTER(c_void_p), c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                                 c_void_p]),
    "ccsd_plan_destroy": (c_int, [c_void_p]),
    "ccsd_plan_set_env": (c_int, [c_void_p, c_void_p, c_int]),
    "ccsd_plan_nao": (c_int, [c_void_p, c_int]),
    "ccsd_int2e_s2kl": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "ccsd_int2e_ip1_s2kl": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "ccsd_grad_ip1_gamma": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int, c_int, c_void_p]),
    "ccsd_int1e": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "ccsd_plan_set_option": (c_int, [c_void_p, c_int, c_int]),
    "ccsd_gemm_mixed": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, ctypes.c_longlong,
                                ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                ctypes.c_longlong, c_int, c_void_p]),
    "ccsd_split_bf16": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                ctypes.c_longlong, c_void_p]),
    "ccsd_gemm_bf16x3_nt": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                    ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                    ctypes.c_longlong, ctypes.c_longlong, c_int, c_int, c_void_p]),
    "ccsd_rys_roots": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "ccsd_plan_create_batch": (c_int, [ctypes.POINTER(c_void_p), c_int, c_void_p, c_void_p, c_int, c_void_p, c_int,
                                       c_void_p, c_int]),
    "ccsd_plan_screen": (c_int, [c_void_p, c_double, c_void_p]),
    "ccsd_schwarz": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ccsd_plan_quartets": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_longlong)]),
    "ccsd_plan_last_launches": (c_int, [c_void_p, c_void_p, c_int, ctypes.POINTER(c_int)]),
    "ccsd_grad_ip1_separable": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "ccsd_grad_ip1_gamma_sep": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int, c_int,
                                        c_void_p]),
    "ccsd_sym_pairs": (c_int, [c_void_p, c_int, c_double, c_void_p]),
    "ccsd_repack_pairs": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "ccsd_repack_pairs_f32": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "ccsd_add_separable_packed": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "ccsd_pack_pairs_sym": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.c_longlong, c_void_p,
                                    c_void_p]),
    "ccsd_permute": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ccsd_combine": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int,
                             c_void_p]),
    "ccsd_gemv": (c_int, [c_int, c_void_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, c_void_p,
                          ctypes.c_longlong, c_void_p, c_double, c_int, c_void_p, c_int, c_void_p]),
    "ccsd_gemv_noalloc": (c_int, [c_int, c_void_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, c_void_p,
                                  ctypes.c_longlong, c_void_p, c_double, c_int, c_int, c_void_p]),
    "ccsd_l1_terms": (c_int, [c_void_p, c_int, c_int, c_void_p, ctypes.c_longlong, ctypes.c_longlong, c_double,
                              c_double, c_double, c_double, c_void_p, c_void_p, c_void_p]),
    "ccsd_unpack_pairs": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, ctypes.c_longlong, c_double,
                                  c_void_p, c_void_p]),
    "ccsd_pair_combine": (c_int, [c_void_p, c_void_p, c_int, c_int, c_double, c_double, c_double, c_double,
                                  c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "ccsd_gemm_set_variant": (c_int, [c_int]),
    "ccsd_vvvv_pm_pack": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, ctypes.c_longlong, c_void_p, c_void_p,
                                  ctypes.c_longlong, ctypes.c_longlong, c_int, c_void_p, c_void_p, c_void_p]),
    "ccsd_vvvv_pm_unpack": (c_int, [c_void_p, c_int, c_int, ctypes.c_longlong, ctypes.c_double, c_void_p, c_void_p]),
    "ccsd_l2_finish": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                               c_void_p, c_void_p]),
    "ccsd_diis_workspace": (ctypes.c_longlong, [ctypes.c_longlong, c_int]),
    "ccsd_diis_push": (c_int, [c_void_p, c_void_p, ctypes.c_longlong, ctypes.c_longlong, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, ctypes.c_longlong, c_int, c_int, c_void_p, c_void_p, c_void_p,
                               ctypes.c_longlong, c_void_p]),
    "ccsd_diis_combine": (c_int, [c_void_p, ctypes.c_longlong, c_int, c_void_p, c_void_p, c_void_p, ctypes.c_longlong,
                                  c_void_p]),
    "ccsd_gemm_bf16x3_nt_scatter": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                            ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                            c_void_p, c_void_p, c_void_p, c_int, c_double, c_int, c_void_p]),
    "ccsd_split_bf16_strided": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "ccsd_scatter_sum": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_double, c_int, c_void_p]),
    "ccsd_sum_slices": (c_int, [c_void_p, c_void_p, ctypes.c_longlong, c_int, c_int, c_void_p]),
}
EXPORTED = tuple(_SIGNATURES)


class HipLibraryError(RuntimeError):
    pass


def load():
    """Load the library (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise HipLibraryError(
            f"HIP library not built: {LIB_PATH}. Run `python -c 'import __graft_entry__ as g; g.build()'`.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().ccsd_last_error().decode(errors="replace")
        raise HipLibraryError(f"{what} failed ({rc}): {msg}")


def current_stream_ptr():
    import torch
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return c_void_p(t.data_ptr())


_PLANS = {}


def get_plan(mol):
    """Basis plan for `mol` on the current device (cached by basis topology)."""
    import numpy as np
    import torch
    lib = load()
    dev = torch.cuda.current_device()
    key = (mol._bas.tobytes(), mol._atm.tobytes(), len(mol._env), dev)
    env = np.ascontiguousarray(mol._env, dtype=np.float64)
    ent = _PLANS.get(key)
    if ent is None:
        atm = np.ascontiguousarray(mol._atm, dtype=np.int32)
        bas = np.ascontiguousarray(mol._bas, dtype=np.int32)
        c2s = np.ascontiguousarray(mol.cart2sph_coeff(), dtype=np.float64)
        h = c_void_p()
        check(lib.ccsd_plan_create(ctypes.byref(h), atm.ctypes.data, mol.natm, bas.ctypes.data, mol.nbas,
                                   env.ctypes.data, len(env), c2s.ctypes.data), "ccsd_plan_create")
        ent = [h, env.copy()]
        _PLANS[key] = ent
    elif not np.array_equal(ent[1], env):
        check(lib.ccsd_plan_set_env(ent[0], env.ctypes.data, len(env)), "ccsd_plan_set_env")
        ent[1] = env.copy()
    return ent[0]


@atexit.register
def _destroy_plans():
    if _LIB is None:
        return
    for h, _ in _PLANS.values():
        _LIB.ccsd_plan_destroy(h)
    _PLANS.clear()
