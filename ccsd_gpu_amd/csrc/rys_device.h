// Rys roots and weights on the device (Chebyshev tables of rys_table.inc).
#pragma once
#include <hip/hip_runtime.h>

#define RYS_STORAGE static __device__
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
    const int it = (int)(T * RYS_INV_WIDTH);
    const double u = (T - it * RYS_WIDTH) * (2.0 / RYS_WIDTH) - 1.0;
    const double u2 = 2.0 * u;
    const double* c = RYS_TABLE + RYS_OFFSET[n - 1] + (long)it * 2 * n * (RYS_DEGREE + 1);
    for (int s = 0; s < 2 * n; s++) {
        const double* cs = c + s * (RYS_DEGREE + 1);
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
