// Shared definitions of the ERI library (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CCSD_BAS_SLOTS 8
#define CCSD_ATM_SLOTS 6
#define CCSD_ATOM_OF 0
#define CCSD_ANG_OF 1
#define CCSD_NPRIM_OF 2
#define CCSD_NCTR_OF 3
#define CCSD_PTR_EXP 5
#define CCSD_PTR_COEFF 6
#define CCSD_PTR_COORD 1

#define CCSD_LMAX 4
#define CCSD_NROOTS_MAX 9

// Error codes returned by every entry point.
#define CCSD_OK 0
#define CCSD_ERR_NULL -1
#define CCSD_ERR_SHAPE -2
#define CCSD_ERR_UNSUPPORTED -3
#define CCSD_ERR_HIP -4
#define CCSD_ERR_ALLOC -5

// Per-shell data in device memory (one entry per _bas row).
struct ShellInfo {
    int l, nprim, nctr, atom;
    int pexp, pcoef;     // offsets into env
    int ao_cart, ao_sph; // first AO (contraction column 0) in each basis
    int mol, pad_;       // molecule of the shell (multi-molecule plans)
};

// One unit of quartet work: bra pair x, ket pairs [y, z) of the class lists.
// A chunk never mixes molecules, so all its quartets share the bra atom.

// Launch description of one angular-momentum class of shell quartets.
struct EriClassArgs {
    const ShellInfo* shells;
    const double* env;
    const double* coords;      // natm x 3
    const int2* bra;           // (ish, jsh) pairs of this class
    const int2* ket;           // (ksh, lsh) pairs of this class, ksh >= lsh
    int nbra, nket;
    int li, lj, lk, ll;
    int deriv;                 // 0: (ij|kl); 1: d/dr_i (ij|kl)
    int nroots;
    int max_nij, max_nkl;      // max primitive pairs over this class's bra / ket pairs
    int pq_batch;              // primitive quartets per root batch (<= 64)
    int item_batch;            // (pq, root) items per recurrence sub-batch
    int groups;                // item groups G (G * O = 64)
    int units_pass;            // output units per pass
    int max_ncij, max_nckl;    // max contraction-column products per side
    // output (s2kl, Cartesian): out[comp][i - i_base][j][k(k+1)/2 + l]
    double* out;
    int i_base, nao;
    long long npair;
    long long comp_stride;
    // fused gradient (deriv == 1, out == nullptr): dE/dR[atom_i][3] += scale * sum ip1 * gamma,
    // accumulated in fixed point (defx_add below) so that the sum is independent of the
    // order in which waves finish
    const double* gamma;       // Cartesian (nao, nao, npair) packed like the output
    unsigned long long* defx;  // overflow count, then [natm][3][2] fixed-point words
    double gamma_scale;
    // work list
    const int4* chunks;        // (bra index, ket begin, ket end, unused)
    int nchunks;
    // Schwarz mode (qmax != nullptr): qmax[ish * nbas + jsh] = max sqrt((ij|ij)) as IEEE bits
    unsigned long long* qmax;
    int nbas;
    // gamma_mode 0: explicit packed gamma; 1: separable two-particle density of D and P,
    // per molecule m, Cartesian AOs local to m:
    //   Gamma_ijkl = D_ij P_kl + P_ij D_kl - (D_il P_kj + D_ik P_lj + P_il D_jk + P_ik D_jl) / 4,
    // halved on the packed diagonal k == l like the explicit form; 2: the sum of both
    int gamma_mode;
    const double* dm;
    const double* pm;
    const long long* mol_dm_off; // offset of molecule m's (nao_m x nao_m) block in dm / pm
    const int* mol_ao0;          // first Cartesian AO of molecule m
    const int* mol_nao;          // Cartesian AOs of molecule m
    // ket primitive pairs past the Gaussian-product cut (register kernels): entries
    // [ket_pp_off[kq], ket_pp_off[kq + 1]) of ket pair kq, 10 doubles each: a_kl, Q (3),
    // K_cd / a_kl, 0.5 / a_kl, c_k c_l of up to 4 contraction-column combinations
    const int* ket_pp_off;
    const double* ket_pp;
};

#ifdef __HIPCC__
// Global-memory pointer types for device code: a load through a pointer of this type
// is a global load also where the compiler cannot trace the pointer to a kernel
// argument (inside non-inlined functions, where a plain pointer would be read with
// flat instructions, which also wait for LDS). Plain pointers in the host pass, which
// type-checks the same templates.
#if defined(__HIP_DEVICE_COMPILE__)
#define CCSD_AS1 __attribute__((address_space(1)))
#else
#define CCSD_AS1
#endif
template <typename T>
__device__ __forceinline__ const CCSD_AS1 T* gld(const T* p)
{
    return (const CCSD_AS1 T*)p;
}
template <typename T>
__device__ __forceinline__ CCSD_AS1 T* gst(T* p)
{
    return (CCSD_AS1 T*)p;
}

// Three doubles (a centre's coordinates) read from global memory.
__device__ __forceinline__ void load3(const double* p, double (&v)[3])
{
    const CCSD_AS1 double* g = gld(p);
    v[0] = g[0];
    v[1] = g[1];
    v[2] = g[2];
}

// Where the two-particle density of one molecule lives: resolved once per quartet, so
// that the element reads below carry no dependent metadata loads of their own and can
// all be in flight together.
struct GammaSrc {
    const CCSD_AS1 double* gamma;
    long long nao, npair;
    int mode;
    int o, n;            // molecule's first Cartesian AO and AO count (separable part)
    const CCSD_AS1 double* D;
    const CCSD_AS1 double* P;
};

// Fixed-point gradient accumulation. A contribution x is held as two 64-bit integers,
// h = rint(x 2^20) and l = rint((x 2^20 - h) 2^40), i.e. x to 2^-60; both are added with
// integer atomics, which are exact and associative, so the accumulated sums (and the
// gradient formed from them once the pass ends) are the same bit for bit whatever the
// order of the additions -- unlike float64 atomics. Contributions must stay below 2^40
// in magnitude (gradient terms are many orders smaller); a larger or non-finite one is
// counted in the overflow word, and the conversion then returns NaN.
__device__ __forceinline__ void defx_add1(unsigned long long* w, unsigned long long* overflow, double x)
{
    const double s = x * 1048576.0;                 // x 2^20
    if (!(fabs(s) < 1.152921504606847e18)) {        // |x| < 2^40, false for NaN
        atomicAdd(overflow, 1ull);
        return;
    }
    const double h = rint(s);
    const double l = rint((s - h) * 1099511627776.0);   // the remainder (exact) times 2^40
    atomicAdd(w, (unsigned long long)(long long)h);
    atomicAdd(w + 1, (unsigned long long)(long long)l);
}

__device__ __forceinline__ void defx_add(unsigned long long* defx, int atom, double x, double y, double z)
{
    unsigned long long* w = defx + 1 + 6 * atom;
    unsigned long long* ov = defx;
    defx_add1(w, ov, x);
    defx_add1(w + 2, ov, y);
    defx_add1(w + 4, ov, z);
}

__device__ __forceinline__ GammaSrc gamma_src(const EriClassArgs& a, int mol)
{
    GammaSrc s;
    s.gamma = gld(a.gamma);
    s.nao = a.nao;
    s.npair = a.npair;
    s.mode = a.gamma_mode;
    s.o = 0;
    s.n = 0;
    s.D = nullptr;
    s.P = nullptr;
    if (s.mode != 0) {
        s.o = gld(a.mol_ao0)[mol];
        s.n = gld(a.mol_nao)[mol];
        s.D = gld(a.dm) + gld(a.mol_dm_off)[mol];
        s.P = gld(a.pm) + gld(a.mol_dm_off)[mol];
    }
    return s;
}

// Two-particle density element for Cartesian AOs (global indices) i j k l, k >= l,
// of the packed (diagonal-halved) form, in either representation.
__device__ __forceinline__ double gamma_at(const GammaSrc& s, int i, int j, int k, int l)
{
    const double ge = s.mode != 1 ? s.gamma[((long long)i * s.nao + j) * s.npair + (long long)k * (k + 1) / 2 + l]
                                  : 0.0;
    if (s.mode == 0) return ge;
    const int n = s.n;
    const CCSD_AS1 double* D = s.D;
    const CCSD_AS1 double* P = s.P;
    i -= s.o; j -= s.o; k -= s.o; l -= s.o;
    const double g = D[i * n + j] * P[k * n + l] + P[i * n + j] * D[k * n + l] -
                     0.25 * (D[i * n + l] * P[k * n + j] + D[i * n + k] * P[l * n + j] + P[i * n + l] * D[j * n + k] +
                             P[i * n + k] * D[j * n + l]);
    return ge + (k == l ? 0.5 * g : g);
}

__device__ __forceinline__ double gamma_packed(const EriClassArgs& a, int mol, int i, int j, int k, int l)
{
    return gamma_at(gamma_src(a, mol), i, j, k, l);
}
#endif
