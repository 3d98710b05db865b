// Two-particle density helpers (HBM-bound).
//
// pack_pairs_sym_kernel: the (ij)(kl) symmetrisation of a half-transformed 2-RDM
// t[k][l][p][q] folded into packing k >= l, the kl diagonal halved (the layout
// the fused gradient kernel reads after the remaining two index transforms).
//
// repack_pairs_kernel: z[k][l][p] (n, n, P) with p = (i, j) packed i >= j (P = m(m+1)/2)
// -> out[r][i][j] (n(n+1)/2, m, m) for the pairs r = (k, l), k >= l: the kl pair packed,
// the ij pair unpacked, with the packing weights moved from ij to kl.
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

constexpr int PT = 32;

// out[r][a][b] = w (t[k,l,a,b] + t[l,k,a,b] + t[k,l,b,a] + t[l,k,b,a]) / 2 for the pair
// r = (k, l) = (kk[r], ll[r]), k >= l, w = 1/2 on the diagonal (k == l), else 1;
// t (n, n, m, m) contiguous. One block per (pair, 32 x 32 tile of (a, b)); the
// transposed blocks come through LDS.
__global__ void __launch_bounds__(256) pack_pairs_sym_kernel(const double* __restrict__ t,
                                                             const long long* __restrict__ kk,
                                                             const long long* __restrict__ ll, int n, int m,
                                                             long long npair, double* __restrict__ out)
{
    __shared__ double tkl[PT][PT + 1];   // t[k][l][b0 + r][a0 + c]
    __shared__ double tlk[PT][PT + 1];   // t[l][k][b0 + r][a0 + c]
    const int nt = (m + PT - 1) / PT;
    const long long blk = blockIdx.x;
    const int tile = (int)(blk % ((long long)nt * nt));
    const long long r = blk / ((long long)nt * nt);
    if (r >= npair) return;
    const int k = (int)kk[r], l = (int)ll[r];
    const int a0 = (tile / nt) * PT, b0 = (tile - (tile / nt) * nt) * PT;
    const int tx = threadIdx.x & (PT - 1), ty = threadIdx.x >> 5;
    const long long mm = (long long)m * m;
    const double* Tkl = t + ((long long)k * n + l) * mm;
    const double* Tlk = t + ((long long)l * n + k) * mm;
#pragma unroll
    for (int q = 0; q < PT; q += 8) {
        const int b = b0 + ty + q, a = a0 + tx;
        const bool ok = b < m && a < m;
        tkl[ty + q][tx] = ok ? Tkl[(long long)b * m + a] : 0.0;
        tlk[ty + q][tx] = ok ? Tlk[(long long)b * m + a] : 0.0;
    }
    __syncthreads();
    const double w = (k == l) ? 0.25 : 0.5;
    double* O = out + r * mm;
#pragma unroll
    for (int q = 0; q < PT; q += 8) {
        const int a = a0 + ty + q, b = b0 + tx;
        if (a < m && b < m) {
            const long long e = (long long)a * m + b;
            O[e] = w * ((Tkl[e] + Tlk[e]) + (tkl[tx][ty + q] + tlk[tx][ty + q]));
        }
    }
}

// out[r][i][j] = (wkl / wij) z[k][l][p(i, j)] for r = k (k + 1) / 2 + l, k >= l, with
// p(i, j) = a (a + 1) / 2 + b for a = max(i, j), b = min(i, j), wkl = 1 (k == l) else 2
// and wij = 1 (i == j) else 2. One block per pair r; wave w writes rows i = w, w + 4, ...
// (coalesced along j), the gathered row of z is read from cache.
template <typename T>
__global__ void __launch_bounds__(256) repack_pairs_kernel(const T* __restrict__ z, int n, int m,
                                                          T* __restrict__ out)
{
    const long long r = blockIdx.x;
    int k, l;
    pair_of(r, k, l);
    const long long P = (long long)m * (m + 1) / 2;
    const T* Z = z + ((long long)k * n + l) * P;
    T* O = out + r * m * m;
    const double wkl = (k == l) ? 1.0 : 2.0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = wave; i < m; i += 4) {
        const long long ri = (long long)i * (i + 1) / 2;
        for (int j = lane; j < m; j += 64) {
            const long long p = (j <= i) ? ri + j : (long long)j * (j + 1) / 2 + i;
            const double f = (i == j) ? wkl : 0.5 * wkl;
            O[(long long)i * m + j] = (T)(f * (double)Z[p]);
        }
    }
}

// g[i][j][p] += w_kl * Gsep(i, j, k, l) for p = k (k + 1) / 2 + l, k >= l, w_kl = 1/2 on the
// diagonal k == l, else 1, with the separable two-particle density of D and P
//   Gsep = D_ij P_kl + P_ij D_kl - (D_il P_kj + D_ik P_lj + P_il D_jk + P_ik D_jl) / 4
// (the packed, diagonal-halved form the fused gradient kernel reads). One block per (i, j);
// D[i][:], P[i][:], D[j][:] and P[:][j] staged in LDS, D_kl / P_kl read through the cache.
__global__ void __launch_bounds__(256) add_separable_packed_kernel(double* __restrict__ g,
                                                                  const double* __restrict__ D,
                                                                  const double* __restrict__ P, int n)
{
    extern __shared__ double rows[];
    double* Di = rows;
    double* Pi = rows + n;
    double* Dj = rows + 2 * n;
    double* Pcj = rows + 3 * n;
    const int i = blockIdx.x / n, j = blockIdx.x - (blockIdx.x / n) * n;
    for (int t = threadIdx.x; t < n; t += 256) {
        Di[t] = D[(long long)i * n + t];
        Pi[t] = P[(long long)i * n + t];
        Dj[t] = D[(long long)j * n + t];
        Pcj[t] = P[(long long)t * n + j];
    }
    __syncthreads();
    const double dij = D[(long long)i * n + j], pij = P[(long long)i * n + j];
    const long long npair = (long long)n * (n + 1) / 2;
    double* G = g + ((long long)i * n + j) * npair;
    for (long long p = threadIdx.x; p < npair; p += 256) {
        int k, l;
        pair_of(p, k, l);
        const long long kl = (long long)k * n + l;
        const double s = dij * P[kl] + pij * D[kl] -
                         0.25 * (Di[l] * Pcj[k] + Di[k] * Pcj[l] + Pi[l] * Dj[k] + Pi[k] * Dj[l]);
        G[p] += (k == l) ? 0.5 * s : s;
    }
}

}  // namespace

hipError_t ccsd_launch_add_separable_packed(double* g, const double* D, const double* P, int n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if ((long long)n * n > 0x7fffffffLL || n > 2048) return hipErrorInvalidValue;
    hipLaunchKernelGGL(add_separable_packed_kernel, dim3((unsigned)((long long)n * n)), dim3(256),
                       sizeof(double) * 4 * n, stream, g, D, P, n);
    return hipGetLastError();
}

hipError_t ccsd_launch_repack_pairs(const void* z, int n, int m, void* out, int f32, hipStream_t stream)
{
    if (n <= 0 || m <= 0) return hipSuccess;
    const long long npair = (long long)n * (n + 1) / 2;
    if (npair > 0x7fffffffLL) return hipErrorInvalidValue;
    if (f32)
        hipLaunchKernelGGL(repack_pairs_kernel<float>, dim3((unsigned)npair), dim3(256), 0, stream,
                           (const float*)z, n, m, (float*)out);
    else
        hipLaunchKernelGGL(repack_pairs_kernel<double>, dim3((unsigned)npair), dim3(256), 0, stream,
                           (const double*)z, n, m, (double*)out);
    return hipGetLastError();
}

hipError_t ccsd_launch_pack_pairs_sym(const double* t, const long long* kk, const long long* ll, int n, int m,
                                      long long npair, double* out, hipStream_t stream)
{
    if (n <= 0 || m <= 0 || npair <= 0) return hipSuccess;
    const long long nt = (m + PT - 1) / PT;
    const long long nb = npair * nt * nt;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_pairs_sym_kernel, dim3((unsigned)nb), dim3(256), 0, stream, t, kk, ll, n, m, npair, out);
    return hipGetLastError();
}

hipError_t ccsd_launch_sym_pairs(double* t, int n, double scale, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const long long np = (long long)n * (n + 1) / 2;
    long long nb = (np * np + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    hipLaunchKernelGGL(sym_pairs_kernel, dim3((unsigned)nb), dim3(256), 0, stream, t, n, scale);
    return hipGetLastError();
}
