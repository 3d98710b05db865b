// Register-resident fused int2e_ip1 x Gamma kernels of the classes with bra
// angular momenta (CCSD_REG_LI, CCSD_REG_LJ); built once per pair (0..2, 0..2).
#include "fused_ip1_reg.h"

#define CCSD_CAT2(a, b, c) a##b##c
#define CCSD_CAT(a, b, c) CCSD_CAT2(a, b, c)
CCSD_REG_TU(CCSD_REG_LI, CCSD_REG_LJ, CCSD_CAT(ccsd_launch_fused_reg_l, CCSD_REG_LI, CCSD_REG_LJ))
