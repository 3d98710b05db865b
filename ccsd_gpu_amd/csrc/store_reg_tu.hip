// Register-resident store-mode int2e / int2e_ip1 kernels of the classes with angular
// momenta (CCSD_REG_LI, CCSD_REG_LJ, CCSD_REG_LK, any l_l); built once per triple.
#include "fused_ip1_reg.h"

#define CCSD_CAT2(a, b, c, d) a##b##c##d
#define CCSD_CAT(a, b, c, d) CCSD_CAT2(a, b, c, d)
CCSD_STORE_TU(CCSD_REG_LI, CCSD_REG_LJ, CCSD_REG_LK,
              CCSD_CAT(ccsd_launch_store_reg_l, CCSD_REG_LI, CCSD_REG_LJ, CCSD_REG_LK))
