// Layout and vector kernels of the contraction planner (HBM-bound, FP64).
//
// permute_kernel / permute_tiled_kernel: a strided view (ndim <= 6, sizes in the
// destination's order, source element strides per destination dimension) copied
// into a contiguous destination, either float64 or as split-BF16 planes
// (hi = bf16(x), lo = bf16(x - hi), the operand form of the bf16x3 GEMM).
//   - when the source is contiguous along the destination's last dimension, each
//     wave copies rows: reads and writes both run along that dimension;
//   - otherwise the source's stride-1 dimension and the destination's last one
//     are tiled 32 x 32 through LDS, so reads run along the first and writes
//     along the second (the transposing copies of permuted GEMM operands).
//
// combine_rows_kernel / combine_tiled_kernel: dst (+)= sum_t coef_t src_t over up to
// four strided views of one shape (the permuted sums of the density and lambda code),
// transposed terms through the same LDS tiles.
//
// gemv_rows_kernel: y[m] (+)= alpha sum_k A[m*lda + k] x[k*incx], one wave per
// (row, K slice), coalesced along k; gemv_cols_kernel: y[n] (+)= alpha
// sum_k A[k*lda + n] x[k*incx], one thread per (column, K slice), coalesced
// along n. With several K slices the partial sums go to a workspace that the
// slice-sum kernel adds in a fixed order.
#include <hip/hip_runtime.h>

#include "bf16_split.h"
#include <stdint.h>

namespace {

constexpr int MAXD = 6;

struct PermDesc {
    int ndim;
    long long size[MAXD];      // destination order
    long long sstride[MAXD];   // source element stride of each destination dimension
    long long dstride[MAXD];   // destination (contiguous) strides
};


__device__ __forceinline__ void store(double v, long long e, double* dst, uint16_t* hi, uint16_t* lo)
{
    if (dst) {
        dst[e] = v;
    } else {
        const float f = (float)v;
        const uint16_t h = ccsd_f2bf(f);
        hi[e] = h;
        lo[e] = ccsd_f2bf(f - ccsd_bf2f(h));
    }
}

// rows of the destination's last dimension; one wave per row, grid-stride over rows
__global__ void __launch_bounds__(256) permute_rows_kernel(const double* __restrict__ src, PermDesc d,
                                                           long long nrows, double* __restrict__ dst,
                                                           uint16_t* __restrict__ hi, uint16_t* __restrict__ lo)
{
    const int lane = threadIdx.x & 63;
    const long long L = d.size[d.ndim - 1];
    const long long ls = d.sstride[d.ndim - 1];
    const long long nw = (long long)gridDim.x * 4;
    for (long long row = blockIdx.x * 4LL + (threadIdx.x >> 6); row < nrows; row += nw) {
        long long rem = row, soff = 0;
        for (int k = d.ndim - 2; k >= 0; k--) {
            const long long q = rem / d.size[k];
            soff += (rem - q * d.size[k]) * d.sstride[k];
            rem = q;
        }
        const long long doff = row * L;
        for (long long c = lane; c < L; c += 64) store(src[soff + c * ls], doff + c, dst, hi, lo);
    }
}

constexpr int TT = 32;

// dimension di (source stride 1) and the destination's last dimension dj tiled 32 x 32
__global__ void __launch_bounds__(256) permute_tiled_kernel(const double* __restrict__ src, PermDesc d, int di,
                                                            long long nouter, double* __restrict__ dst,
                                                            uint16_t* __restrict__ hi, uint16_t* __restrict__ lo)
{
    __shared__ double tile[TT][TT + 1];
    const int dj = d.ndim - 1;
    const long long Ni = d.size[di], Nj = d.size[dj];
    const long long ti = (Ni + TT - 1) / TT, tj = (Nj + TT - 1) / TT;
    const long long blk = blockIdx.x;
    const long long tile_id = blk % (ti * tj);
    const long long outer = blk / (ti * tj);
    if (outer >= nouter) return;
    const long long i0 = (tile_id / tj) * TT, j0 = (tile_id % tj) * TT;
    // source and destination offsets of the outer index (all dimensions but di, dj)
    long long rem = outer, soff = 0, doff = 0;
    for (int k = d.ndim - 1; k >= 0; k--) {
        if (k == di || k == dj) continue;
        const long long q = rem / d.size[k];
        const long long x = rem - q * d.size[k];
        soff += x * d.sstride[k];
        doff += x * d.dstride[k];
        rem = q;
    }
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x >> 5;
    const long long sj = d.sstride[dj], dsi = d.dstride[di];
#pragma unroll
    for (int k = 0; k < TT; k += 8) {
        const long long i = i0 + tx, j = j0 + ty + k;
        tile[ty + k][tx] = (i < Ni && j < Nj) ? src[soff + i + j * sj] : 0.0;   // tile[j - j0][i - i0]
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TT; k += 8) {
        const long long i = i0 + ty + k, j = j0 + tx;
        if (i < Ni && j < Nj) store(tile[tx][ty + k], doff + i * dsi + j, dst, hi, lo);
    }
}

// ---------------------------------------------------------------------------
// combine: dst (+)= sum_t coef[t] * src_t, each src_t a strided view of the
// destination's shape (ndim <= 6), dst any strided view. Terms contiguous along the
// destination's last dimension dj are read directly; terms contiguous along one
// other dimension di are staged through 32 x 32 LDS tiles (the transposed terms of
// the permuted sums in the density code); any other term is read elementwise.
constexpr int MAXT = 4;

struct CombDesc {
    int ndim, nt, di;
    long long size[MAXD];
    long long dstride[MAXD];
    long long sstride[MAXT][MAXD];
    const double* src[MAXT];
    double coef[MAXT];
    int mode[MAXT];          // 0 direct (stride 1 along dj), 1 tiled (stride 1 along di), 2 gather
};

__global__ void __launch_bounds__(256) combine_rows_kernel(CombDesc d, long long nrows, double* __restrict__ dst,
                                                           int beta)
{
    const int lane = threadIdx.x & 63;
    const int dj = d.ndim - 1;
    const long long L = d.size[dj];
    const long long nw = (long long)gridDim.x * 4;
    for (long long row = blockIdx.x * 4LL + (threadIdx.x >> 6); row < nrows; row += nw) {
        long long rem = row, doff = 0;
        long long soff[MAXT] = {0, 0, 0, 0};
        for (int k = dj - 1; k >= 0; k--) {
            const long long q = rem / d.size[k];
            const long long x = rem - q * d.size[k];
            doff += x * d.dstride[k];
#pragma unroll
            for (int t = 0; t < MAXT; t++)
                if (t < d.nt) soff[t] += x * d.sstride[t][k];
            rem = q;
        }
        for (long long c = lane; c < L; c += 64) {
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < MAXT; t++)
                if (t < d.nt) v += d.coef[t] * d.src[t][soff[t] + c * d.sstride[t][dj]];
            double* p = dst + doff + c * d.dstride[dj];
            *p = beta ? *p + v : v;
        }
    }
}

__global__ void __launch_bounds__(256) combine_tiled_kernel(CombDesc d, long long nouter, double* __restrict__ dst,
                                                            int beta)
{
    __shared__ double tile[MAXT][TT][TT + 1];
    const int dj = d.ndim - 1, di = d.di;
    const long long Ni = d.size[di], Nj = d.size[dj];
    const long long ti = (Ni + TT - 1) / TT, tj = (Nj + TT - 1) / TT;
    const long long blk = blockIdx.x;
    const long long tile_id = blk % (ti * tj);
    const long long outer = blk / (ti * tj);
    if (outer >= nouter) return;
    const long long i0 = (tile_id / tj) * TT, j0 = (tile_id % tj) * TT;
    long long rem = outer, doff = 0;
    long long soff[MAXT] = {0, 0, 0, 0};
    for (int k = d.ndim - 1; k >= 0; k--) {
        if (k == di || k == dj) continue;
        const long long q = rem / d.size[k];
        const long long x = rem - q * d.size[k];
        doff += x * d.dstride[k];
#pragma unroll
        for (int t = 0; t < MAXT; t++)
            if (t < d.nt) soff[t] += x * d.sstride[t][k];
        rem = q;
    }
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x >> 5;
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        if (t >= d.nt || d.mode[t] != 1) continue;
#pragma unroll
        for (int k = 0; k < TT; k += 8) {
            const long long i = i0 + tx, j = j0 + ty + k;
            tile[t][ty + k][tx] = (i < Ni && j < Nj) ? d.src[t][soff[t] + i + j * d.sstride[t][dj]] : 0.0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TT; k += 8) {
        const long long i = i0 + ty + k, j = j0 + tx;
        if (i < Ni && j < Nj) {
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < MAXT; t++) {
                if (t >= d.nt) continue;
                const double x = d.mode[t] == 1 ? tile[t][tx][ty + k]
                                                : d.src[t][soff[t] + i * d.sstride[t][di] + j * d.sstride[t][dj]];
                v += d.coef[t] * x;
            }
            double* p = dst + doff + i * d.dstride[di] + j * d.dstride[dj];
            *p = beta ? *p + v : v;
        }
    }
}

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// partial of slice s: sum over k in [s*kc, min(K, (s+1)*kc))
__global__ void __launch_bounds__(256) gemv_rows_kernel(const double* __restrict__ A, long long M, long long K,
                                                        long long lda, const double* __restrict__ x, long long incx,
                                                        double* __restrict__ y, double alpha, int beta,
                                                        double* __restrict__ ws, int S, long long kc)
{
    const int lane = threadIdx.x & 63;
    const long long task = blockIdx.x * 4LL + (threadIdx.x >> 6);   // (slice, row)
    if (task >= M * S) return;
    const long long s = task / M, m = task - (task / M) * M;
    const long long k0 = s * kc, k1 = (k0 + kc < K) ? k0 + kc : K;
    const double* a = A + m * lda;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
    long long k = k0 + lane;
    for (; k + 192 < k1; k += 256) {
        acc0 += a[k] * x[k * incx];
        acc1 += a[k + 64] * x[(k + 64) * incx];
        acc2 += a[k + 128] * x[(k + 128) * incx];
        acc3 += a[k + 192] * x[(k + 192) * incx];
    }
    for (; k < k1; k += 64) acc0 += a[k] * x[k * incx];
    const double v = alpha * wave_sum((acc0 + acc1) + (acc2 + acc3));
    if (lane == 0) {
        if (S == 1) y[m] = beta ? y[m] + v : v;
        else ws[s * M + m] = v;
    }
}

__global__ void __launch_bounds__(256) gemv_cols_kernel(const double* __restrict__ A, long long K, long long N,
                                                        long long lda, const double* __restrict__ x, long long incx,
                                                        double* __restrict__ y, double alpha, int beta,
                                                        double* __restrict__ ws, int S, long long kc)
{
    const long long nblk = (N + 255) / 256;
    const long long s = blockIdx.x / nblk;
    const long long n = (blockIdx.x - s * nblk) * 256 + threadIdx.x;
    if (s >= S || n >= N) return;
    const long long k0 = s * kc, k1 = (k0 + kc < K) ? k0 + kc : K;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
    long long k = k0;
    for (; k + 3 < k1; k += 4) {
        acc0 += A[k * lda + n] * x[k * incx];
        acc1 += A[(k + 1) * lda + n] * x[(k + 1) * incx];
        acc2 += A[(k + 2) * lda + n] * x[(k + 2) * incx];
        acc3 += A[(k + 3) * lda + n] * x[(k + 3) * incx];
    }
    for (; k < k1; k++) acc0 += A[k * lda + n] * x[k * incx];
    const double v = alpha * ((acc0 + acc1) + (acc2 + acc3));
    if (S == 1) y[n] = beta ? y[n] + v : v;
    else ws[s * N + n] = v;
}

}  // namespace

hipError_t ccsd_launch_permute(const double* src, int ndim, const long long* sizes, const long long* sstrides,
                               double* dst, void* hi, void* lo, hipStream_t stream)
{
    if (ndim < 1 || ndim > MAXD) return hipErrorInvalidValue;
    PermDesc d;
    d.ndim = ndim;
    long long n = 1;
    for (int k = 0; k < MAXD; k++) {
        d.size[k] = k < ndim ? sizes[k] : 1;
        d.sstride[k] = k < ndim ? sstrides[k] : 0;
        if (k < ndim) n *= sizes[k];
    }
    if (n <= 0) return hipSuccess;
    long long st = 1;
    for (int k = ndim - 1; k >= 0; k--) {
        d.dstride[k] = st;
        st *= d.size[k];
    }
    for (int k = ndim; k < MAXD; k++) d.dstride[k] = 0;
    // the destination dimension that the source walks with stride 1 (largest such)
    int di = -1;
    for (int k = 0; k < ndim; k++)
        if (d.sstride[k] == 1 && d.size[k] > 1 && (di < 0 || d.size[k] > d.size[di])) di = k;
    const int dj = ndim - 1;
    if (di < 0 || di == dj || d.sstride[dj] == 1 || d.size[dj] < 4) {
        const long long nrows = n / d.size[dj];
        long long nb = (nrows + 3) / 4;
        if (nb > 65536LL * 4) nb = 65536LL * 4;
        hipLaunchKernelGGL(permute_rows_kernel, dim3((unsigned)nb), dim3(256), 0, stream, src, d, nrows, dst,
                           (uint16_t*)hi, (uint16_t*)lo);
        return hipGetLastError();
    }
    const long long ti = (d.size[di] + TT - 1) / TT, tj = (d.size[dj] + TT - 1) / TT;
    const long long nouter = n / (d.size[di] * d.size[dj]);
    const long long nb = nouter * ti * tj;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(permute_tiled_kernel, dim3((unsigned)nb), dim3(256), 0, stream, src, d, di, nouter, dst,
                       (uint16_t*)hi, (uint16_t*)lo);
    return hipGetLastError();
}

hipError_t ccsd_launch_combine(int ndim, const long long* sizes, double* dst, const long long* dstrides, int nt,
                               const double* const* srcs, const long long* sstrides, const double* coef, int beta,
                               hipStream_t stream)
{
    if (ndim < 1 || ndim > MAXD || nt < 1 || nt > MAXT) return hipErrorInvalidValue;
    CombDesc d;
    d.ndim = ndim;
    d.nt = nt;
    long long n = 1;
    for (int k = 0; k < MAXD; k++) {
        d.size[k] = k < ndim ? sizes[k] : 1;
        d.dstride[k] = k < ndim ? dstrides[k] : 0;
        if (k < ndim) n *= sizes[k];
    }
    if (n <= 0) return hipSuccess;
    const int dj = ndim - 1;
    // the tile dimension: the (largest) dimension that most non-direct terms walk with stride 1
    int votes[MAXD] = {0, 0, 0, 0, 0, 0};
    for (int t = 0; t < MAXT; t++) {
        d.src[t] = t < nt ? srcs[t] : nullptr;
        d.coef[t] = t < nt ? coef[t] : 0.0;
        for (int k = 0; k < MAXD; k++) d.sstride[t][k] = (t < nt && k < ndim) ? sstrides[t * ndim + k] : 0;
        d.mode[t] = 2;
        if (t >= nt) continue;
        if (d.sstride[t][dj] == 1 || d.size[dj] == 1) {
            d.mode[t] = 0;
            continue;
        }
        for (int k = 0; k < dj; k++)
            if (d.sstride[t][k] == 1 && d.size[k] > 1) votes[k]++;
    }
    int di = -1;
    for (int k = 0; k < dj; k++)
        if (votes[k] > 0 && (di < 0 || votes[k] > votes[di] || (votes[k] == votes[di] && d.size[k] > d.size[di])))
            di = k;
    d.di = di;
    if (di < 0 || d.size[dj] < 4) {
        const long long nrows = n / d.size[dj];
        long long nb = (nrows + 3) / 4;
        if (nb > 65536LL * 4) nb = 65536LL * 4;
        hipLaunchKernelGGL(combine_rows_kernel, dim3((unsigned)nb), dim3(256), 0, stream, d, nrows, dst, beta);
        return hipGetLastError();
    }
    for (int t = 0; t < nt; t++)
        if (d.mode[t] == 2 && d.sstride[t][di] == 1) d.mode[t] = 1;
    const long long ti = (d.size[di] + TT - 1) / TT, tj = (d.size[dj] + TT - 1) / TT;
    const long long nouter = n / (d.size[di] * d.size[dj]);
    const long long nb = nouter * ti * tj;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(combine_tiled_kernel, dim3((unsigned)nb), dim3(256), 0, stream, d, nouter, dst, beta);
    return hipGetLastError();
}

hipError_t ccsd_launch_gemv(int cols, const double* A, long long M, long long K, long long lda, const double* x,
                            long long incx, double* y, double alpha, int beta, double* ws, int S, hipStream_t stream)
{
    // cols = 0: y[m] = sum_k A[m*lda + k] x[k]   (M rows of length K)
    // cols = 1: y[m] = sum_k A[k*lda + m] x[k]   (K rows of length M)
    if (M <= 0) return hipSuccess;
    if (S < 1) S = 1;
    const long long kc = (K + S - 1) / S;
    if (!cols) {
        const long long tasks = M * S;
        hipLaunchKernelGGL(gemv_rows_kernel, dim3((unsigned)((tasks + 3) / 4)), dim3(256), 0, stream, A, M, K, lda, x,
                           incx, y, alpha, beta, ws, S, kc);
    } else {
        const long long nblk = (M + 255) / 256;
        hipLaunchKernelGGL(gemv_cols_kernel, dim3((unsigned)(nblk * S)), dim3(256), 0, stream, A, K, M, lda, x, incx,
                           y, alpha, beta, ws, S, kc);
    }
    return hipGetLastError();
}
