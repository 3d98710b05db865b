// Two-particle density helpers (HBM-bound).
//
// sym_pairs_kernel: in place, for a dense (n, n, n, n) tensor t,
//   t[p][q][r][s] <- scale * (t[p][q][r][s] + t[q][p][r][s] + t[p][q][s][r] + t[q][p][s][r])
// for every index. One thread per orbit of the two swaps (representative
// p >= q, r >= s) reads the (up to) four members and writes the common value
// to all of them, so orbits never overlap and the update is race free.
#include <hip/hip_runtime.h>
#include <math.h>

namespace {

__device__ __forceinline__ void pair_of(long long x, int& a, int& b)
{
    // x = a (a + 1) / 2 + b, 0 <= b <= a
    long long r = (long long)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > x) r--;
    while ((r + 1) * (r + 2) / 2 <= x) r++;
    a = (int)r;
    b = (int)(x - r * (r + 1) / 2);
}

__global__ void sym_pairs_kernel(double* __restrict__ t, int n, double scale)
{
    const long long np = (long long)n * (n + 1) / 2;
    const long long total = np * np;
    const long long n2 = (long long)n * n;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < total;
         u += (long long)gridDim.x * blockDim.x) {
        const long long x = u / np, y = u - (u / np) * np;
        int p, q, r, s;
        pair_of(x, p, q);
        pair_of(y, r, s);
        const long long pq = (long long)p * n + q, qp = (long long)q * n + p;
        const long long rs = (long long)r * n + s, sr = (long long)s * n + r;
        const double v = scale * (t[pq * n2 + rs] + t[qp * n2 + rs] + t[pq * n2 + sr] + t[qp * n2 + sr]);
        t[pq * n2 + rs] = v;
        t[qp * n2 + rs] = v;
        t[pq * n2 + sr] = v;
        t[qp * n2 + sr] = v;
    }
}

}  // namespace

hipError_t ccsd_launch_sym_pairs(double* t, int n, double scale, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const long long np = (long long)n * (n + 1) / 2;
    long long nb = (np * np + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    hipLaunchKernelGGL(sym_pairs_kernel, dim3((unsigned)nb), dim3(256), 0, stream, t, n, scale);
    return hipGetLastError();
}
