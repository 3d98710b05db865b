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
};

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
    // fused gradient (deriv == 1, out == nullptr): de[atom_i][3] -= sum ip1 * gamma
    const double* gamma;       // Cartesian (nao, nao, npair) packed like the output
    double* de;
    double gamma_scale;
};
