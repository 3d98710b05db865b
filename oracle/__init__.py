"""FP64 CPU oracle for the CCSD-gradient hot path.

TEST INFRASTRUCTURE ONLY. Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` import this package, always as the checker
(or the timed CPU baseline), never as the thing that is measured or shipped.

Contents:
  * integrals.py   ctypes wrapper of rys_oracle.c (Rys int2e / int2e_ip1,
                   s2kl layout, Cartesian and spherical) -- restates
                   ccsd_gpu/cuda/int2e_ip1_module.py and int2e_module.py;
  * md_ref.py      independent McMurchie-Davidson brute force (mpmath Boys)
                   used to pin the Rys oracle on tiny inputs.
"""
