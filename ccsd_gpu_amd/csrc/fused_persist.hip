// The fused int2e_ip1 x Gamma pass as ONE persistent launch that runs the
// register-resident per-class bodies (fused_ip1_reg.h).
//
// Every class of the pass that has a register body (all (l_i, l_j, l_k, l_l) <= 2
// combinations, with or without two-column s shells, except the ten whose Gamma and
// 1-D factors do not fit the register budget in eight parts), for every molecule of a
// batch plan, is described in a device table; work items are runs of one class's
// (chunk, part) units, ordered class by class. Each one-wave block draws items from a
// device counter until the list is exhausted. The body of each class is a separate
// (non-inlined) function: the kernel dispatches on the class code of the item it drew
// (uniform over the wave), and the body keeps drawing items while they belong to its
// class, returning the first item of another class. The class table is patched with
// the call's Gamma and gradient pointers before the launch. The arithmetic is the
// per-class kernels' (same per-lane quartet evaluation, bra primitive-pair table in LDS,
// work lists and fixed-point dE/dR accumulation); where a class takes more parts here
// (below) its per-wave partial sums group differently, so the two forms agree to
// rounding, and each repeats itself bit for bit.
//
// The bodies are separate functions because one function holding all of them does not
// fit the register file: inlined together they spill even at two waves per SIMD, and
// the compilation time grows far faster than the number of bodies. As calls, each
// keeps its own allocation inside the kernel's two-waves-per-SIMD budget and saves the
// callee-saved registers it uses once per call (once per run of same-class items).
#include <utility>

#include "fused_ip1_reg.h"

namespace ccsd_reg {

// Bodies of this kernel keep a smaller register budget than the per-class kernels
// (more parts per class where needed): as functions they carry a few registers of call
// overhead, and every body must fit two waves per SIMD without spilling.
constexpr int kPersistBudget = 96;

// per-call arguments (the class table holds everything else)
struct RegPersistDyn {
    const double* gamma;
    const double* dm;
    const double* pm;
    unsigned long long* defx;
    double scale;
    int gamma_mode;
    int pad_;
};

// class code (l_i, l_j, l_k, l_l each 0..2, gc 0/1)
constexpr int kRegCodes = 162;
__host__ __device__ constexpr int reg_code(int li, int lj, int lk, int ll, int gc)
{
    return (((li * 3 + lj) * 3 + lk) * 3 + ll) * 2 + gc;
}

template <int CODE>
struct RegCode {
    static constexpr int GC = CODE & 1;
    static constexpr int LL = (CODE / 2) % 3, LK = (CODE / 6) % 3, LJ = (CODE / 18) % 3, LI = CODE / 54;
    // gc = 1 is only produced for classes with an s shell (the only multi-column shells
    // the register bodies take)
    static constexpr bool valid =
        !(GC && !has_s(LI, LJ, LK, LL)) && min_parts(LI, LJ, LK, LL, GC, kPersistBudget) > 0;
    static constexpr int P = valid ? min_parts(LI, LJ, LK, LL, GC, kPersistBudget) : 1;
};

// Next work item index of this wave (uniform): lane 0 draws it from the counter.
__device__ __forceinline__ int draw(unsigned int* counter)
{
    int t = 0;
    if (threadIdx.x == 0) t = (int)atomicAdd(counter, 1u);
    return __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
}

// A pointer made wave-uniform (its two halves through readfirstlane), so that the loads
// through it are scalar loads, as for a kernel argument.
template <typename T>
__device__ __forceinline__ const T* uniform_ptr(const T* p)
{
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<const T*>(((unsigned long long)hi << 32) | lo);
}

// Items of class `cls` from item t on, while the items drawn stay in this class; returns
// the first item index drawn that does not (>= nitems when the list is exhausted). Unit
// u of an item is part u % P of chunk u / P.
template <int CODE>
__device__ __attribute__((noinline)) int persist_class(const EriClassArgs* __restrict__ classes,
                                                       const int4* __restrict__ items, int nitems,
                                                       unsigned int* __restrict__ counter,
                                                       unsigned int* __restrict__ ticks, int t, int cls)
{
    using R = RegCode<CODE>;
    if constexpr (!R::valid) {
        return nitems;   // never dispatched (the host routes such classes elsewhere)
    } else {
        constexpr int LI = R::LI, LJ = R::LJ, LK = R::LK, LL = R::LL, GC = R::GC, P = R::P;
        constexpr int MIJ = GC ? mcol(LI) * mcol(LJ) : 1;
        t = __builtin_amdgcn_readfirstlane(t);
        cls = __builtin_amdgcn_readfirstlane(cls);
        nitems = __builtin_amdgcn_readfirstlane(nitems);
        const CCSD_AS1 int4* gitems = gld(uniform_ptr(items));
        const EriClassArgs& a = *uniform_ptr(classes + cls);
        // the kernel's dynamic LDS, named here so that the table accesses are LDS ones
        extern __shared__ double lds_tab[];
        double* tab = lds_tab;
        const int lane = threadIdx.x;
        const double norm = 34.98683665524972497 * shell_norm_dev(LI) * shell_norm_dev(LJ) * shell_norm_dev(LK) *
                            shell_norm_dev(LL);
        for (;;) {
            const unsigned long long t_start = wall_clock64();
            const int4 it = gitems[t];
            const int u0 = __builtin_amdgcn_readfirstlane(it.y), u1 = __builtin_amdgcn_readfirstlane(it.z);
            for (int u = u0; u < u1; u++) {
                const int c = u / P, part = u - c * P;
                const int4 ch = gld(a.chunks)[c];
                const int2 bp = gld(a.bra)[ch.x];
                const ShellInfo si = gld(a.shells)[bp.x], sj = gld(a.shells)[bp.y];
                double RA[3], RB[3];
                load3(a.coords + 3 * si.atom, RA);
                load3(a.coords + 3 * sj.atom, RB);
                const double AB[3] = {RB[0] - RA[0], RB[1] - RA[1], RB[2] - RA[2]};
                __syncthreads();   // one-wave block: the previous unit's table reads are done
                const int nij = fill_bra<MIJ>(a, si, sj, RA, AB, norm, lane, tab);
                __syncthreads();
                double sx = 0.0, sy = 0.0, sz = 0.0;
                run_parts<LI, LJ, LK, LL, GC, P>(a, ch, lane, part, tab, nij, sx, sy, sz,
                                                 std::make_integer_sequence<int, P>{});
                for (int off = 32; off > 0; off >>= 1) {
                    sx += __shfl_down(sx, off, 64);
                    sy += __shfl_down(sy, off, 64);
                    sz += __shfl_down(sz, off, 64);
                }
                if (lane == 0) defx_add(a.defx, si.atom, a.gamma_scale * sx, a.gamma_scale * sy, a.gamma_scale * sz);
            }
            if (ticks && lane == 0) {
                // the item's duration in constant-rate clock ticks (host calibration)
                const unsigned long long dt = wall_clock64() - t_start;
                gst(ticks)[t] = dt > 0xffffffffULL ? 0xffffffffu : (unsigned int)dt;
            }
            t = draw(counter);
            if (t >= nitems) return t;
            if (__builtin_amdgcn_readfirstlane(gitems[t].x) != cls) return t;
        }
    }
}

template <int... C>
__device__ __forceinline__ int dispatch(int code, const EriClassArgs* classes, const int4* items, int nitems,
                                        unsigned int* counter, unsigned int* ticks, int t, int cls,
                                        std::integer_sequence<int, C...>)
{
    int next = nitems;
    ((code == C ? (void)(next = persist_class<C>(classes, items, nitems, counter, ticks, t, cls)) : void()), ...);
    return next;
}

// The per-call fields of every class's launch description (before each launch).
__global__ void persist_patch_kernel(EriClassArgs* __restrict__ classes, int n, const RegPersistDyn dyn)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    classes[i].gamma = dyn.gamma;
    classes[i].dm = dyn.dm;
    classes[i].pm = dyn.pm;
    classes[i].defx = dyn.defx;
    classes[i].gamma_scale = dyn.scale;
    classes[i].gamma_mode = dyn.gamma_mode;
}

// items[t] = (class, first unit, end unit, unused); `counter` must read 0 at launch and
// every wave leaves the loop once it draws t >= nitems; the class table carries this
// call's Gamma and gradient fields (persist_patch_kernel). ticks (optional, nitems
// words): each item's duration in wall_clock64 ticks. Dynamic LDS: the bra table of the
// largest class (max_nij x kBraFields doubles). Two waves per SIMD: every body fits that
// register budget on its own.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    fused_reg_persistent_kernel(const EriClassArgs* __restrict__ classes, const int4* __restrict__ items, int nitems,
                                unsigned int* __restrict__ counter, unsigned int* __restrict__ ticks)
{
    int t = draw(counter);
    while (t < nitems) {
        const int cls = __builtin_amdgcn_readfirstlane(items[t].x);
        const EriClassArgs* ap = classes + cls;
        const int gc = (ap->max_ncij > 1 || ap->max_nckl > 1) ? 1 : 0;
        const int code = __builtin_amdgcn_readfirstlane(reg_code(ap->li, ap->lj, ap->lk, ap->ll, gc));
        t = dispatch(code, classes, items, nitems, counter, ticks, t, cls,
                     std::make_integer_sequence<int, kRegCodes>{});
    }
}

}  // namespace ccsd_reg

// Parts (units per chunk) of the class in the persistent kernel, 0 if it has no body.
int ccsd_reg_persist_parts(int li, int lj, int lk, int ll, int max_ncij, int max_nckl)
{
    using namespace ccsd_reg;
    if (li > 2 || lj > 2 || lk > 2 || ll > 2) return 0;
    if (max_ncij > mcol(li) * mcol(lj) || max_nckl > mcol(lk) * mcol(ll)) return 0;
    const int gc = (max_ncij > 1 || max_nckl > 1) ? 1 : 0;
    return min_parts(li, lj, lk, ll, gc, kPersistBudget);
}

size_t ccsd_reg_persist_lds(int max_nij) { return sizeof(double) * (size_t)max_nij * ccsd_reg::kBraFields; }

// Resident one-wave blocks per CU at `lds` bytes of dynamic LDS.
int ccsd_reg_persist_blocks_per_cu(size_t lds)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ccsd_reg::fused_reg_persistent_kernel, 64, lds) !=
        hipSuccess)
        return 0;
    return nb;
}

hipError_t ccsd_launch_reg_persistent(EriClassArgs* classes, int nclasses, const int4* items, int nitems,
                                      unsigned int* counter, unsigned int* ticks, const double* gamma, const double* dm,
                                      const double* pm, unsigned long long* defx, double scale, int gamma_mode,
                                      int nblocks, size_t lds, hipStream_t stream)
{
    if (nitems <= 0) return hipSuccess;
    if (nblocks <= 0 || nclasses <= 0 || lds > 65536) return hipErrorInvalidValue;
    ccsd_reg::RegPersistDyn dyn;
    dyn.gamma = gamma;
    dyn.dm = dm;
    dyn.pm = pm;
    dyn.defx = defx;
    dyn.scale = scale;
    dyn.gamma_mode = gamma_mode;
    dyn.pad_ = 0;
    hipLaunchKernelGGL(ccsd_reg::persist_patch_kernel, dim3((unsigned)((nclasses + 63) / 64)), dim3(64), 0, stream,
                       classes, nclasses, dyn);
    hipLaunchKernelGGL(ccsd_reg::fused_reg_persistent_kernel, dim3((unsigned)nblocks), dim3(64), lds, stream, classes,
                       items, nitems, counter, ticks);
    return hipGetLastError();
}
