// Host side of the register-resident fused kernels: class support and dispatch
// to the per-(l_i, l_j, l_k) object files.
#include "fused_ip1_reg.h"

hipError_t ccsd_launch_fused_reg_l000(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l001(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l002(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l010(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l011(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l012(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l020(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l021(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l022(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l100(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l101(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l102(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l110(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l111(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l112(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l120(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l121(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l122(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l200(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l201(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l202(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l210(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l211(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l212(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l220(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l221(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_fused_reg_l222(const EriClassArgs&, int, int, hipStream_t);

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
    switch (a.li * 9 + a.lj * 3 + a.lk) {
        case 0: return ccsd_launch_fused_reg_l000(a, gc, wpb, stream);
        case 1: return ccsd_launch_fused_reg_l001(a, gc, wpb, stream);
        case 2: return ccsd_launch_fused_reg_l002(a, gc, wpb, stream);
        case 3: return ccsd_launch_fused_reg_l010(a, gc, wpb, stream);
        case 4: return ccsd_launch_fused_reg_l011(a, gc, wpb, stream);
        case 5: return ccsd_launch_fused_reg_l012(a, gc, wpb, stream);
        case 6: return ccsd_launch_fused_reg_l020(a, gc, wpb, stream);
        case 7: return ccsd_launch_fused_reg_l021(a, gc, wpb, stream);
        case 8: return ccsd_launch_fused_reg_l022(a, gc, wpb, stream);
        case 9: return ccsd_launch_fused_reg_l100(a, gc, wpb, stream);
        case 10: return ccsd_launch_fused_reg_l101(a, gc, wpb, stream);
        case 11: return ccsd_launch_fused_reg_l102(a, gc, wpb, stream);
        case 12: return ccsd_launch_fused_reg_l110(a, gc, wpb, stream);
        case 13: return ccsd_launch_fused_reg_l111(a, gc, wpb, stream);
        case 14: return ccsd_launch_fused_reg_l112(a, gc, wpb, stream);
        case 15: return ccsd_launch_fused_reg_l120(a, gc, wpb, stream);
        case 16: return ccsd_launch_fused_reg_l121(a, gc, wpb, stream);
        case 17: return ccsd_launch_fused_reg_l122(a, gc, wpb, stream);
        case 18: return ccsd_launch_fused_reg_l200(a, gc, wpb, stream);
        case 19: return ccsd_launch_fused_reg_l201(a, gc, wpb, stream);
        case 20: return ccsd_launch_fused_reg_l202(a, gc, wpb, stream);
        case 21: return ccsd_launch_fused_reg_l210(a, gc, wpb, stream);
        case 22: return ccsd_launch_fused_reg_l211(a, gc, wpb, stream);
        case 23: return ccsd_launch_fused_reg_l212(a, gc, wpb, stream);
        case 24: return ccsd_launch_fused_reg_l220(a, gc, wpb, stream);
        case 25: return ccsd_launch_fused_reg_l221(a, gc, wpb, stream);
        case 26: return ccsd_launch_fused_reg_l222(a, gc, wpb, stream);
        default: return hipErrorInvalidValue;
    }
}

// Store-mode kernels (int2e / int2e_ip1 into the s2kl tensor), one object per (l_i, l_j, l_k)
hipError_t ccsd_launch_store_reg_l000(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l001(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l002(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l010(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l011(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l012(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l020(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l021(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l022(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l100(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l101(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l102(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l110(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l111(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l112(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l120(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l121(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l122(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l200(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l201(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l202(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l210(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l211(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l212(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l220(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l221(const EriClassArgs&, int, int, hipStream_t);
hipError_t ccsd_launch_store_reg_l222(const EriClassArgs&, int, int, hipStream_t);

// Waves per quartet of the class's store-mode register kernel (deriv 0: int2e, 1:
// int2e_ip1), 0 if the class stays on the LDS kernel.
int ccsd_store_reg_parts(int li, int lj, int lk, int ll, int max_ncij, int max_nckl, int deriv)
{
    using namespace ccsd_reg;
    if (li > 2 || lj > 2 || lk > 2 || ll > 2) return 0;
    if (max_ncij > mcol(li) * mcol(lj) || max_nckl > mcol(lk) * mcol(ll)) return 0;
    const int gc = (max_ncij > 1 || max_nckl > 1) ? 1 : 0;
    return min_parts_store(li, lj, lk, ll, gc && has_s(li, lj, lk, ll) ? 1 : 0, deriv ? 3 : 1);
}

hipError_t ccsd_launch_store_reg(const EriClassArgs& a, int wpb, hipStream_t stream)
{
    const int gc = (a.max_ncij > 1 || a.max_nckl > 1) ? 1 : 0;
    switch (a.li * 9 + a.lj * 3 + a.lk) {
        case 0: return ccsd_launch_store_reg_l000(a, gc, wpb, stream);
        case 1: return ccsd_launch_store_reg_l001(a, gc, wpb, stream);
        case 2: return ccsd_launch_store_reg_l002(a, gc, wpb, stream);
        case 3: return ccsd_launch_store_reg_l010(a, gc, wpb, stream);
        case 4: return ccsd_launch_store_reg_l011(a, gc, wpb, stream);
        case 5: return ccsd_launch_store_reg_l012(a, gc, wpb, stream);
        case 6: return ccsd_launch_store_reg_l020(a, gc, wpb, stream);
        case 7: return ccsd_launch_store_reg_l021(a, gc, wpb, stream);
        case 8: return ccsd_launch_store_reg_l022(a, gc, wpb, stream);
        case 9: return ccsd_launch_store_reg_l100(a, gc, wpb, stream);
        case 10: return ccsd_launch_store_reg_l101(a, gc, wpb, stream);
        case 11: return ccsd_launch_store_reg_l102(a, gc, wpb, stream);
        case 12: return ccsd_launch_store_reg_l110(a, gc, wpb, stream);
        case 13: return ccsd_launch_store_reg_l111(a, gc, wpb, stream);
        case 14: return ccsd_launch_store_reg_l112(a, gc, wpb, stream);
        case 15: return ccsd_launch_store_reg_l120(a, gc, wpb, stream);
        case 16: return ccsd_launch_store_reg_l121(a, gc, wpb, stream);
        case 17: return ccsd_launch_store_reg_l122(a, gc, wpb, stream);
        case 18: return ccsd_launch_store_reg_l200(a, gc, wpb, stream);
        case 19: return ccsd_launch_store_reg_l201(a, gc, wpb, stream);
        case 20: return ccsd_launch_store_reg_l202(a, gc, wpb, stream);
        case 21: return ccsd_launch_store_reg_l210(a, gc, wpb, stream);
        case 22: return ccsd_launch_store_reg_l211(a, gc, wpb, stream);
        case 23: return ccsd_launch_store_reg_l212(a, gc, wpb, stream);
        case 24: return ccsd_launch_store_reg_l220(a, gc, wpb, stream);
        case 25: return ccsd_launch_store_reg_l221(a, gc, wpb, stream);
        case 26: return ccsd_launch_store_reg_l222(a, gc, wpb, stream);
        default: return hipErrorInvalidValue;
    }
}
