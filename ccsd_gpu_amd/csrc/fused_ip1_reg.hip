// Host side of the register-resident fused kernels: class support and dispatch
// to the per-(l_i, l_j) object files.
#include "fused_ip1_reg.h"

#define CCSD_REG_DECL(I, J) hipError_t ccsd_launch_fused_reg_l##I##J(const EriClassArgs&, int, int, hipStream_t);
CCSD_REG_DECL(0, 0) CCSD_REG_DECL(0, 1) CCSD_REG_DECL(0, 2)
CCSD_REG_DECL(1, 0) CCSD_REG_DECL(1, 1) CCSD_REG_DECL(1, 2)
CCSD_REG_DECL(2, 0) CCSD_REG_DECL(2, 1) CCSD_REG_DECL(2, 2)

// Waves per quartet ("parts") of the class's register kernel, 0 if it has none.
// The kernels assume single-column shells except s shells with up to two columns.
int ccsd_fused_reg_parts(int li, int lj, int lk, int ll, int max_ncij, int max_nckl)
{
    using namespace ccsd_reg;
    if (li > 2 || lj > 2 || lk > 2 || ll > 2) return 0;
    if (max_ncij > mcol(li) * mcol(lj) || max_nckl > mcol(lk) * mcol(ll)) return 0;
    const int gc = (max_ncij > 1 || max_nckl > 1) ? 1 : 0;
    return min_parts(li, lj, lk, ll, gc);
}

// Dynamic LDS of a register-kernel launch (the waves' bra primitive-pair tables).
size_t ccsd_fused_reg_lds(int max_nij, int wpb) { return sizeof(double) * (size_t)wpb * max_nij * ccsd_reg::kBraFields; }

hipError_t ccsd_launch_fused_reg(const EriClassArgs& a, int wpb, hipStream_t stream)
{
    const int gc = (a.max_ncij > 1 || a.max_nckl > 1) ? 1 : 0;
    switch (a.li * 3 + a.lj) {
        case 0: return ccsd_launch_fused_reg_l00(a, gc, wpb, stream);
        case 1: return ccsd_launch_fused_reg_l01(a, gc, wpb, stream);
        case 2: return ccsd_launch_fused_reg_l02(a, gc, wpb, stream);
        case 3: return ccsd_launch_fused_reg_l10(a, gc, wpb, stream);
        case 4: return ccsd_launch_fused_reg_l11(a, gc, wpb, stream);
        case 5: return ccsd_launch_fused_reg_l12(a, gc, wpb, stream);
        case 6: return ccsd_launch_fused_reg_l20(a, gc, wpb, stream);
        case 7: return ccsd_launch_fused_reg_l21(a, gc, wpb, stream);
        case 8: return ccsd_launch_fused_reg_l22(a, gc, wpb, stream);
        default: return hipErrorInvalidValue;
    }
}
