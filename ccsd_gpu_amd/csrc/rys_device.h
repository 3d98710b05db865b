// Rys roots and weights on the device (Chebyshev tables of rys_table.inc).
#pragma once
#include <hip/hip_runtime.h>

// 16-byte aligned: a series' coefficients are read as pairs (every series starts at a
// multiple of RYS_DEGREE + 1 = 14 doubles)
#define RYS_STORAGE static __device__ __attribute__((aligned(16)))
#include "rys_table.inc"

namespace ccsd {

// Rys roots (u = t^2, ascending) and weights for n roots at argument T.
__device__ inline void rys_roots_dev(int n, double T, double* __restrict__ r, double* __restrict__ w)
{
    if (T >= RYS_TASYM(n)) {
        const int o = n * (n - 1) / 2;
        const double s = sqrt(0.78539816339744830962 / T);
        const double inv = 1.0 / T;
        for (int i = 0; i < n; i++) {
            r[i] = RYS_LARGE_R[o + i] * inv;
            w[i] = RYS_LARGE_W[o + i] * s;
        }
        return;
    }
    static_assert((RYS_DEGREE + 1) % 2 == 0, "coefficients are read in pairs");
    constexpr int NP = (RYS_DEGREE + 1) / 2;
    const int it = (int)(T * RYS_INV_WIDTH);
    const double u = (T - it * RYS_WIDTH) * (2.0 / RYS_WIDTH) - 1.0;
    const double u2 = 2.0 * u;
    // the lanes of a wave read different intervals: one 16-byte gather per coefficient
    // pair halves the vector memory instructions of the 8-byte form
    const double2* c = reinterpret_cast<const double2*>(RYS_TABLE + RYS_OFFSET[n - 1] +
                                                        (long)it * 2 * n * (RYS_DEGREE + 1));
    for (int s = 0; s < 2 * n; s++) {
        double cs[RYS_DEGREE + 1];
#pragma unroll
        for (int h = 0; h < NP; h++) {
            const double2 v = c[s * NP + h];
            cs[2 * h] = v.x;
            cs[2 * h + 1] = v.y;
        }
        double b1 = 0.0, b2 = 0.0;
#pragma unroll
        for (int k = RYS_DEGREE; k >= 1; k--) {
            const double b0 = cs[k] + u2 * b1 - b2;
            b2 = b1;
            b1 = b0;
        }
        const double v = cs[0] + u * b1 - b2;
        if (s < n) r[s] = v; else w[s - n] = v;
    }
}

}  // namespace ccsd
