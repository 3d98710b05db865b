// Split-BF16 GEMM on the CDNA4 matrix cores: C (+)= A B^T with FP32 accumulation.
//
// Operands are stored pre-split as two bfloat16 planes, x = hi + lo
// (hi = bf16(x), lo = bf16(x - hi)), both K-contiguous:
//   A{hi,lo}: [batch][M][K],  B{hi,lo}: [batch][N][K]   (the "NT" form)
// and the product keeps the three significant terms
//   C = A_hi B_hi^T + A_hi B_lo^T + A_lo B_hi^T
// on v_mfma_f32_32x32x16_bf16 (about 16 significant bits per operand, FP32
// accumulation). C is float64 [batch][M][N] (row stride ldc), overwritten
// (beta = 0) or accumulated into (beta = 1).
//
// Block: 256 threads = 4 waves, 128x128 C tile, each wave 64x64 = 2x2 MFMA
// 32x32 tiles; K staged by 64 (32 when K <= 256) through LDS, next tile
// prefetched into registers while the current one is multiplied. Tiles are handed out XCD-aware: the
// hardware deals consecutive workgroups round-robin to the 8 XCDs, so the linear
// block index is remapped to give each XCD a contiguous range of tiles, ordered
// in groups of 8 M-tiles (N-tile by N-tile inside a group) so that the blocks
// running together on one XCD share A and B tiles in that XCD's L2.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

constexpr int BM = 128, BN = 128;
constexpr int NXCD = 8, GROUP_M = 8;

// BK: K elements staged per LDS tile (64 for long K; 32 when K is short, where the
// deeper tile only adds zero padding and LDS)
template <int BK>
__global__ void __launch_bounds__(256) gemm_bf16x3_nt_kernel(
    const __bf16* __restrict__ Ahi, const __bf16* __restrict__ Alo, const __bf16* __restrict__ Bhi,
    const __bf16* __restrict__ Blo, double* __restrict__ C, int M, int N, int K, long long lda, long long ldb,
    long long sA, long long sB, long long sC, long long ldc, int beta, double alpha, const long long* __restrict__ off_b,
    const long long* __restrict__ off_r, const long long* __restrict__ off_c)
{
    constexpr int LDK = BK + 8;                 // bf16 elements per LDS row (BK + 8: conflict-free b128 reads)
    constexpr int TILE = BM * LDK;              // bf16 elements of one operand plane tile
    constexpr int CPR = BK / 8;                 // 16-byte chunks per tile row
    constexpr int HCH = BM * CPR / 256;         // chunks per thread per plane
    __shared__ __attribute__((aligned(16))) __bf16 lds[4 * TILE];   // Ahi, Alo, Bhi, Blo
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    int tm = blockIdx.x, tn = blockIdx.y;
    if (BK == 64) {
        // long K: XCD-aware grouped order (short K keeps the plain M-fastest order,
        // which measured faster when the output write dominates)
        int lin = blockIdx.x + gridDim.x * blockIdx.y;
        const int nT = gridDim.x * gridDim.y, per = nT / NXCD;
        if (lin < per * NXCD) lin = (lin % NXCD) * per + lin / NXCD;
        const int Mt = gridDim.x;
        const int gid = lin / (GROUP_M * gridDim.y), first = gid * GROUP_M;
        const int gsz = (Mt - first < GROUP_M) ? Mt - first : GROUP_M;
        tm = first + (lin % (GROUP_M * gridDim.y)) % gsz;
        tn = (lin % (GROUP_M * gridDim.y)) / gsz;
    }
    const int m0 = tm * BM, n0 = tn * BN;
    const long long bz = blockIdx.z;
    const __bf16* src[4] = {Ahi + bz * sA, Alo + bz * sA, Bhi + bz * sB, Blo + bz * sB};
    const int rows[4] = {M, M, N, N};
    const int base[4] = {m0, m0, n0, n0};
    const long long ld[4] = {lda, lda, ldb, ldb};

    // each thread moves HCH chunks of 8 bf16 (16 bytes) per plane per K tile:
    // chunk c = tid + 256*h, row = c / CPR, kchunk = c % CPR
    u32x4 stage[4][HCH];
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int h = 0; h < HCH; h++) {
                const int c = tid + 256 * h;
                const int r = c / CPR, kc = (c % CPR) * 8;
                const int gr = base[p] + r, gk = k0 + kc;
                u32x4 v = {0u, 0u, 0u, 0u};
                if (gr < rows[p]) {
                    const __bf16* s = src[p] + (long long)gr * ld[p] + gk;
                    if (gk + 8 <= K && ((reinterpret_cast<uintptr_t>(s) & 15) == 0)) {
                        v = *reinterpret_cast<const u32x4*>(s);
                    } else {
                        const uint16_t* s16 = reinterpret_cast<const uint16_t*>(s);
                        __attribute__((aligned(16))) uint16_t tmp[8];
#pragma unroll
                        for (int e = 0; e < 8; e++) tmp[e] = (gk + e < K) ? s16[e] : (uint16_t)0;
                        v = *reinterpret_cast<const u32x4*>(tmp);
                    }
                }
                stage[p][h] = v;
            }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int h = 0; h < HCH; h++) {
                const int c = tid + 256 * h;
                const int r = c / CPR, kc = (c % CPR) * 8;
                *reinterpret_cast<u32x4*>(lds + p * TILE + r * LDK + kc) = stage[p][h];
            }
    };

    v16f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;

    const int half = lane >> 5, l32 = lane & 31;
    load_tile(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        __syncthreads();           // previous tile fully consumed
        store_tile();
        __syncthreads();
        if (k0 + BK < K) load_tile(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int ko = kk + half * 8;
            v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = wm * 64 + i * 32 + l32;
                ah[i] = *reinterpret_cast<const v8bf*>(lds + 0 * TILE + r * LDK + ko);
                al[i] = *reinterpret_cast<const v8bf*>(lds + 1 * TILE + r * LDK + ko);
                const int c = wn * 64 + i * 32 + l32;
                bh[i] = *reinterpret_cast<const v8bf*>(lds + 2 * TILE + c * LDK + ko);
                bl[i] = *reinterpret_cast<const v8bf*>(lds + 3 * TILE + c * LDK + ko);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                }
        }
    }
    // lane holds C[8*(e/4) + 4*half + e%4][l32] of each 32x32 tile; the element (row, col)
    // of batch b goes to C[b*sC + row*ldc + col], or to C[off_b[b] + off_r[row] + off_c[col]]
    // when offset tables are given (a permuted destination, e.g. a term of a 4-index tensor)
    // The destination offsets of the lane's 32 rows and 2 columns are loaded first, and
    // a 32 x 32 tile's 16 previous values before any of its stores, so that the
    // epilogue's memory round trips overlap instead of running one element at a time
    // (all four tiles at once would need more than 256 registers).
    double* Cb = off_r ? C + off_b[bz] : C + bz * sC;
    long long orow[2][16];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int gr = m0 + wm * 64 + i * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
            orow[i][e] = gr < M ? (off_r ? off_r[gr] : (long long)gr * ldc) : -1;
        }
    long long ocol[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int gc = n0 + wn * 64 + j * 32 + l32;
        ocol[j] = gc < N ? (off_r ? off_c[gc] : (long long)gc) : -1;
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            if (ocol[j] < 0) continue;
            double prev[16];
#pragma unroll
            for (int e = 0; e < 16; e++) prev[e] = (beta && orow[i][e] >= 0) ? Cb[orow[i][e] + ocol[j]] : 0.0;
#pragma unroll
            for (int e = 0; e < 16; e++)
                if (orow[i][e] >= 0) Cb[orow[i][e] + ocol[j]] = prev[e] + alpha * (double)acc[i][j][e];
        }
}

// x (float64, rows x cols, row stride ld) -> hi, lo (bfloat16, rows x cols contiguous)
__global__ void split_bf16_kernel(const double* __restrict__ x, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                  long long rows, long long cols, long long ld)
{
    const long long n = rows * cols;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long r = t / cols, c = t - r * cols;
        const float f = (float)x[r * ld + c];
        const uint16_t h = f2bf(f);
        hi[t] = h;
        lo[t] = f2bf(f - bf2f(h));
    }
}

// x given as an N-d strided view (sizes, element strides, ndim <= 8) -> hi, lo planes
// contiguous in the view's row-major order: the permutation of a GEMM operand and its
// split in one pass.
struct StridedView {
    int ndim;
    long long size[8];
    long long stride[8];
};

__global__ void split_bf16_strided_kernel(const double* __restrict__ x, uint16_t* __restrict__ hi,
                                          uint16_t* __restrict__ lo, long long n, const StridedView v)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        long long rem = t, off = 0;
        for (int d = v.ndim - 1; d >= 0; d--) {
            const long long q = rem / v.size[d];
            off += (rem - q * v.size[d]) * v.stride[d];
            rem = q;
        }
        const float f = (float)x[off];
        const uint16_t h = f2bf(f);
        hi[t] = h;
        lo[t] = f2bf(f - bf2f(h));
    }
}

// C[off_b[b] + off_r[r] + off_c[c]] = beta * C[...] + alpha * sum_s ws[s][b][r][c]: the
// reduction of a K-split GEMM into a permuted destination (each element written once).
__global__ void scatter_sum_kernel(const double* __restrict__ ws, int S, int nb, int M, int N, double* __restrict__ C,
                                   const long long* __restrict__ off_b, const long long* __restrict__ off_r,
                                   const long long* __restrict__ off_c, double alpha, int beta)
{
    const long long mn = (long long)M * N, total = mn * nb;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long b = t / mn, rc = t - b * mn;
        const long long r = rc / N, c = rc - r * N;
        double s = 0.0;
        for (int q = 0; q < S; q++) s += ws[(long long)q * total + t];
        double* p = C + off_b[b] + off_r[r] + off_c[c];
        *p = beta ? *p + alpha * s : alpha * s;
    }
}

// out[i] = beta * out[i] + sum_q ws[q * rstride + i], q < S: the partial products of a
// K-split GEMM summed in a fixed order (one pass, no atomics, no memset).
__global__ void sum_slices_kernel(const double* __restrict__ ws, double* __restrict__ out, long long n, int S,
                                  long long rstride, int beta)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        double s = beta ? out[t] : 0.0;
        for (int q = 0; q < S; q++) s += ws[(long long)q * rstride + t];
        out[t] = s;
    }
}

// First level of a long slice sum: group g (blockIdx.y) adds its rows [g L, g L + L) of
// ws in order and leaves the partial in place of its first row. Many slices over few
// outputs then run on (outputs x groups) threads instead of one thread per output.
__global__ void slice_partial_kernel(double* __restrict__ ws, long long n, int rows, long long rstride, int L)
{
    const int q0 = blockIdx.y * L;
    const int q1 = q0 + L < rows ? q0 + L : rows;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = q0; q < q1; q++) s += ws[(long long)q * rstride + t];
        ws[(long long)q0 * rstride + t] = s;
    }
}

}  // namespace

hipError_t ccsd_launch_sum_slices(double* ws, double* out, long long n, int S, int beta, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    // while the slices outnumber what one thread per output should add, sum them in
    // groups of >= 8 first (partials in place), enough groups to give ~1024 workgroups
    int rows = S;
    long long rstride = n;
    while (rows > 16 && nb < 1024) {
        long long G = (1024 + nb - 1) / nb;
        const long long gmax = (rows + 7) / 8;
        if (G > gmax) G = gmax;
        if (G < 2) break;
        const int L = (int)((rows + G - 1) / G);
        G = (rows + L - 1) / L;
        hipLaunchKernelGGL(slice_partial_kernel, dim3((unsigned)nb, (unsigned)G), dim3(256), 0, stream, ws, n, rows,
                           rstride, L);
        rows = (int)G;
        rstride *= L;
    }
    hipLaunchKernelGGL(sum_slices_kernel, dim3((unsigned)nb), dim3(256), 0, stream, (const double*)ws, out, n, rows,
                       rstride, beta);
    return hipGetLastError();
}

hipError_t ccsd_launch_split_bf16_strided(const double* x, void* hi, void* lo, int ndim, const long long* sizes,
                                          const long long* strides, hipStream_t stream)
{
    if (ndim < 1 || ndim > 8) return hipErrorInvalidValue;
    StridedView v;
    v.ndim = ndim;
    long long n = 1;
    for (int d = 0; d < 8; d++) {
        v.size[d] = d < ndim ? sizes[d] : 1;
        v.stride[d] = d < ndim ? strides[d] : 0;
        if (d < ndim) n *= sizes[d];
    }
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    hipLaunchKernelGGL(split_bf16_strided_kernel, dim3((unsigned)nb), dim3(256), 0, stream, x, (uint16_t*)hi,
                       (uint16_t*)lo, n, v);
    return hipGetLastError();
}

hipError_t ccsd_launch_scatter_sum(const double* ws, int S, int nb, int M, int N, double* C, const long long* off_b,
                                   const long long* off_r, const long long* off_c, double alpha, int beta,
                                   hipStream_t stream)
{
    const long long total = (long long)nb * M * N;
    if (total <= 0) return hipSuccess;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536LL * 8) blocks = 65536LL * 8;
    hipLaunchKernelGGL(scatter_sum_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, ws, S, nb, M, N, C, off_b,
                       off_r, off_c, alpha, beta);
    return hipGetLastError();
}

hipError_t ccsd_launch_gemm_bf16x3_nt(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C,
                                      int M, int N, int K, long long lda, long long ldb, long long sA, long long sB,
                                      long long sC, long long ldc, int batch, int beta, double alpha,
                                      const long long* off_b, const long long* off_r, const long long* off_c,
                                      hipStream_t stream)
{
    if (M <= 0 || N <= 0 || batch <= 0) return hipSuccess;
    dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
    if (K > 256)
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<64>, grid, dim3(256), 0, stream, (const __bf16*)Ahi,
                           (const __bf16*)Alo, (const __bf16*)Bhi, (const __bf16*)Blo, C, M, N, K, lda, ldb, sA, sB,
                           sC, ldc, beta, alpha, off_b, off_r, off_c);
    else
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<32>, grid, dim3(256), 0, stream, (const __bf16*)Ahi,
                           (const __bf16*)Alo, (const __bf16*)Bhi, (const __bf16*)Blo, C, M, N, K, lda, ldb, sA, sB,
                           sC, ldc, beta, alpha, off_b, off_r, off_c);
    return hipGetLastError();
}

hipError_t ccsd_launch_split_bf16(const double* x, void* hi, void* lo, long long rows, long long cols, long long ld,
                                  hipStream_t stream)
{
    const long long n = rows * cols;
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)nb), dim3(256), 0, stream, x, (uint16_t*)hi, (uint16_t*)lo, rows,
                       cols, ld);
    return hipGetLastError();
}
