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
// Default block: 256 threads = 4 waves, 128x128 C tile, each wave 64x64 = 2x2 MFMA
// 32x32 tiles; K staged by 64 (32 when K <= 256) through LDS, next tile prefetched
// into registers while the current one is multiplied (other block shapes and a
// two-stage LDS pipeline are selectable, see gemm_bf16x3_nt_body). For long K, tiles
// are handed out XCD-aware: the hardware deals consecutive workgroups round-robin to
// the 8 XCDs, so the linear block index is remapped to give each XCD a contiguous
// range of tiles, ordered in groups of 8 M-tiles (N-tile by N-tile inside a group) so
// that the blocks running together on one XCD share A and B tiles in that XCD's L2.
#include <hip/hip_runtime.h>

#include "bf16_split.h"
#include <stdint.h>

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));


constexpr int NXCD = 8, GROUP_M = 8;

// One block computes a (64 WM) x (64 WN) tile of C with WM x WN waves, each wave
// 64 x 64 = 2 x 2 MFMA 32x32 tiles. BK: K elements staged per LDS tile (64 for long K;
// 32 when K is short, where the deeper tile only adds zero padding and LDS). NS: LDS
// stages -- 1: one staging buffer, two barriers per K tile; 2: the next tile is stored
// into the other buffer after the current one is multiplied, one barrier per K tile.
// The next K tile is always prefetched into registers while the current one is
// multiplied.
// AL: every row start is 16-byte aligned and K is a multiple of 8 (checked by the
// launcher), so a staged 16-byte chunk lies wholly inside or wholly outside K: full K
// tiles load through precomputed 32-bit row offsets with no per-element test, rows past
// M or N are read from the last valid row instead (they only feed outputs that are
// never stored), and only the last, partial K tile selects zeros.
template <int BK, int WM, int WN, int NS, bool AL>
__device__ __forceinline__ void gemm_bf16x3_nt_body(
    const __bf16* __restrict__ Ahi, const __bf16* __restrict__ Alo, const __bf16* __restrict__ Bhi,
    const __bf16* __restrict__ Blo, double* __restrict__ C, int M, int N, int K, long long lda, long long ldb,
    long long sA, long long sB, long long sC, long long ldc, int beta, double alpha, const long long* __restrict__ off_b,
    const long long* __restrict__ off_r, const long long* __restrict__ off_c)
{
    constexpr int NT = 64 * WM * WN;            // threads
    constexpr int BMT = 64 * WM, BNT = 64 * WN; // block tile
    constexpr int LDK = BK + 8;                 // bf16 elements per LDS row (BK + 8: conflict-free b128 reads)
    constexpr int TA = BMT * LDK, TB = BNT * LDK;   // bf16 elements of one A / B plane tile
    constexpr int STAGE = 2 * TA + 2 * TB;      // Ahi, Alo, Bhi, Blo
    constexpr int CPR = BK / 8;                 // 16-byte chunks per tile row
    constexpr int HA = BMT * CPR / NT, HB = BNT * CPR / NT;   // chunks per thread per A / B plane
    static_assert(HA * NT == BMT * CPR && HB * NT == BNT * CPR, "tile rows must split evenly over threads");
    __shared__ __attribute__((aligned(16))) __bf16 lds[NS * STAGE];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave / WN, wn = wave % WN;
    int tm = blockIdx.x, tn = blockIdx.y;
    if (K > 256) {
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
    const int m0 = tm * BMT, n0 = tn * BNT;
    const long long bz = blockIdx.z;
    const __bf16* pa[2] = {Ahi + bz * sA, Alo + bz * sA};
    const __bf16* pb[2] = {Bhi + bz * sB, Blo + bz * sB};

    // each thread moves 16-byte chunks (8 bf16): chunk c = tid + NT*h, row = c / CPR,
    // kchunk = c % CPR
    u32x4 sta[2][HA], stb[2][HB];
    auto fetch = [&](const __bf16* src, long long ld, int rows, int base, int c, int k0) {
        const int r = c / CPR, kc = (c % CPR) * 8;
        const int gr = base + r, gk = k0 + kc;
        u32x4 v = {0u, 0u, 0u, 0u};
        if (gr < rows) {
            const __bf16* sp = src + (long long)gr * ld + gk;
            if (gk + 8 <= K && ((reinterpret_cast<uintptr_t>(sp) & 15) == 0)) {
                v = *reinterpret_cast<const u32x4*>(sp);
            } else {
                const uint16_t* s16 = reinterpret_cast<const uint16_t*>(sp);
                __attribute__((aligned(16))) uint16_t tmp[8];
#pragma unroll
                for (int e = 0; e < 8; e++) tmp[e] = (gk + e < K) ? s16[e] : (uint16_t)0;
                v = *reinterpret_cast<const u32x4*>(tmp);
            }
        }
        return v;
    };
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int h = 0; h < HA; h++) sta[p][h] = fetch(pa[p], lda, M, m0, tid + NT * h, k0);
#pragma unroll
            for (int h = 0; h < HB; h++) stb[p][h] = fetch(pb[p], ldb, N, n0, tid + NT * h, k0);
        }
    };
    // aligned path: byte offset of the chunk (clamped row, column kc) from the plane
    // start, shared by the hi and lo planes (same strides); the launcher guarantees a
    // plane's last row ends below 2^31 bytes, so the offsets are 32-bit
    unsigned oa[HA], ob[HB];
    int kca[HA], kcb[HB];
    if (AL) {
#pragma unroll
        for (int h = 0; h < HA; h++) {
            const int c = tid + NT * h, r = min(m0 + c / CPR, M - 1);
            kca[h] = (c % CPR) * 8;
            oa[h] = (unsigned)(((long long)r * lda + kca[h]) * 2);
        }
#pragma unroll
        for (int h = 0; h < HB; h++) {
            const int c = tid + NT * h, r = min(n0 + c / CPR, N - 1);
            kcb[h] = (c % CPR) * 8;
            ob[h] = (unsigned)(((long long)r * ldb + kcb[h]) * 2);
        }
    }
    auto chunk = [](const __bf16* base, unsigned off) {
        return *reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(base) + off);
    };
    auto load_full = [&](int k0) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const __bf16* a = pa[p] + k0;
            const __bf16* b = pb[p] + k0;
#pragma unroll
            for (int h = 0; h < HA; h++) sta[p][h] = chunk(a, oa[h]);
#pragma unroll
            for (int h = 0; h < HB; h++) stb[p][h] = chunk(b, ob[h]);
        }
    };
    auto load_last = [&](int k0) {   // chunks at or past K read their row start and become zeros
        const u32x4 z = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int h = 0; h < HA; h++) {
                const bool in = k0 + kca[h] < K;
                const u32x4 v = chunk(pa[p], in ? oa[h] + 2u * k0 : oa[h] - 2u * kca[h]);
                sta[p][h] = in ? v : z;
            }
#pragma unroll
            for (int h = 0; h < HB; h++) {
                const bool in = k0 + kcb[h] < K;
                const u32x4 v = chunk(pb[p], in ? ob[h] + 2u * k0 : ob[h] - 2u * kcb[h]);
                stb[p][h] = in ? v : z;
            }
        }
    };
    auto load_next = [&](int k0) {
        if (!AL) load_tile(k0);
        else if (k0 + BK <= K) load_full(k0);
        else load_last(k0);
    };
    auto store_tile = [&](int stage) {
        __bf16* st = lds + stage * STAGE;
#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int h = 0; h < HA; h++) {
                const int c = tid + NT * h;
                *reinterpret_cast<u32x4*>(st + p * TA + (c / CPR) * LDK + (c % CPR) * 8) = sta[p][h];
            }
#pragma unroll
            for (int h = 0; h < HB; h++) {
                const int c = tid + NT * h;
                *reinterpret_cast<u32x4*>(st + 2 * TA + p * TB + (c / CPR) * LDK + (c % CPR) * 8) = stb[p][h];
            }
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
    // fragments of one 16-deep K step: ah[0..1], al[0..1], bh[0..1], bl[0..1]
    auto read_frags = [&](const __bf16* st, int kk, v8bf* f) {
        const int ko = kk + half * 8;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int r = wm * 64 + i * 32 + l32;
            f[i] = *reinterpret_cast<const v8bf*>(st + r * LDK + ko);
            f[2 + i] = *reinterpret_cast<const v8bf*>(st + TA + r * LDK + ko);
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int c = wn * 64 + j * 32 + l32;
            f[4 + j] = *reinterpret_cast<const v8bf*>(st + 2 * TA + c * LDK + ko);
            f[6 + j] = *reinterpret_cast<const v8bf*>(st + 2 * TA + TB + c * LDK + ko);
        }
    };
    // each accumulator takes lo.hi, hi.lo, hi.hi in that order; consecutive
    // instructions go to different accumulators (no back-to-back dependence)
    auto mfmas = [&](const v8bf* f) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[2 + i], f[4 + j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[i], f[6 + j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[i], f[4 + j], acc[i][j], 0, 0, 0);
    };
    auto multiply = [&](int stage) {
        const __bf16* st = lds + stage * STAGE;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            v8bf f[8];
            read_frags(st, kk, f);
            mfmas(f);
        }
    };

    load_next(0);
    if (NS == 1) {
        for (int k0 = 0; k0 < K; k0 += BK) {
            __syncthreads();           // previous tile fully consumed
            store_tile(0);
            __syncthreads();
            if (k0 + BK < K) load_next(k0 + BK);
            multiply(0);
        }
    } else {
        // stage s holds tile t with s = t % 2; tile t + 1 goes into the other stage once
        // tile t has been multiplied (that stage was last read for tile t - 1, before
        // the previous barrier)
        store_tile(0);
        if (BK < K) load_next(BK);
        __syncthreads();
        int stage = 0;
        for (int k0 = 0; k0 < K; k0 += BK) {
            multiply(stage);
            if (k0 + BK < K) {
                store_tile(stage ^ 1);
                if (k0 + 2 * BK < K) load_next(k0 + 2 * BK);
            }
            __syncthreads();
            stage ^= 1;
        }
    }
    // lane holds C[8*(e/4) + 4*half + e%4][l32] of each 32x32 tile; the element (row, col)
    // of batch b goes to C[b*sC + row*ldc + col], or to C[off_b[b] + off_r[row] + off_c[col]]
    // when offset tables are given (a permuted destination, e.g. a term of a 4-index tensor)
    // The destination offsets of the lane's 32 rows and 2 columns are loaded first, and
    // a 32 x 32 tile's 16 previous values before any of its stores, so that the
    // epilogue's memory round trips overlap instead of running one element at a time
    // (all tiles at once would need more than 256 registers).
    double* Cb = off_r ? C + off_b[bz] : C + bz * sC;
    long long ocol[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int gc = n0 + wn * 64 + j * 32 + l32;
        ocol[j] = gc < N ? (off_r ? off_c[gc] : (long long)gc) : -1;
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
        long long orow[16];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int gr = m0 + wm * 64 + i * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
            orow[e] = gr < M ? (off_r ? off_r[gr] : (long long)gr * ldc) : -1;
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            if (ocol[j] < 0) continue;
            double prev[16];
#pragma unroll
            for (int e = 0; e < 16; e++) prev[e] = (beta && orow[e] >= 0) ? Cb[orow[e] + ocol[j]] : 0.0;
#pragma unroll
            for (int e = 0; e < 16; e++)
                if (orow[e] >= 0) Cb[orow[e] + ocol[j]] = prev[e] + alpha * (double)acc[i][j][e];
        }
    }
}

#define GEMM_BF16X3_ARGS                                                                                              \
    const __bf16 *__restrict__ Ahi, const __bf16 *__restrict__ Alo, const __bf16 *__restrict__ Bhi,                  \
        const __bf16 *__restrict__ Blo, double *__restrict__ C, int M, int N, int K, long long lda, long long ldb,    \
        long long sA, long long sB, long long sC, long long ldc, int beta, double alpha,                              \
        const long long *__restrict__ off_b, const long long *__restrict__ off_r, const long long *__restrict__ off_c
#define GEMM_BF16X3_PASS Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, sA, sB, sC, ldc, beta, alpha, off_b, off_r, off_c

template <int BK, int WM, int WN, int NS, bool AL>
__global__ void __launch_bounds__(64 * WM * WN) gemm_bf16x3_nt_kernel(GEMM_BF16X3_ARGS)
{
    gemm_bf16x3_nt_body<BK, WM, WN, NS, AL>(GEMM_BF16X3_PASS);
}

// x (float64, rows x cols, row stride ld) -> hi, lo (bfloat16, rows x cols contiguous)
__global__ void split_bf16_kernel(const double* __restrict__ x, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                  long long rows, long long cols, long long ld)
{
    const long long n = rows * cols;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long r = t / cols, c = t - r * cols;
        const float f = (float)x[r * ld + c];
        const uint16_t h = ccsd_f2bf(f);
        hi[t] = h;
        lo[t] = ccsd_f2bf(f - ccsd_bf2f(h));
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
        const uint16_t h = ccsd_f2bf(f);
        hi[t] = h;
        lo[t] = ccsd_f2bf(f - ccsd_bf2f(h));
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

// block shape and staging for long K on aligned operands (ccsd_gemm_set_variant):
// 0: 128 x 128, BK 64, one LDS stage (the default); 1: 128 x 128, BK 32, two stages;
// 2: 128 x 256 (8 waves), BK 32, two stages. On the lambda shapes (820 x 23104 x 23104,
// 6080^3, 1600 x 23104 x 6080, 1600 x 23104 x 1600) 1 and 2 measured 0-15% slower
// than 0, except 2 on the shortest K (4% faster).
int g_gemm_variant = 0;

hipError_t ccsd_launch_gemm_bf16x3_nt(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C,
                                      int M, int N, int K, long long lda, long long ldb, long long sA, long long sB,
                                      long long sC, long long ldc, int batch, int beta, double alpha,
                                      const long long* off_b, const long long* off_r, const long long* off_c,
                                      hipStream_t stream)
{
    if (M <= 0 || N <= 0 || batch <= 0) return hipSuccess;
    // aligned path: 16-byte rows and plane offsets below 2^31 bytes
    const bool al = K % 8 == 0 && lda % 8 == 0 && ldb % 8 == 0 && sA % 8 == 0 && sB % 8 == 0 &&
                    ((uintptr_t)Ahi | (uintptr_t)Alo | (uintptr_t)Bhi | (uintptr_t)Blo) % 16 == 0 &&
                    ((long long)(M - 1) * lda + K) * 2 < (1LL << 31) && ((long long)(N - 1) * ldb + K) * 2 < (1LL << 31);
    void (*kern)(GEMM_BF16X3_ARGS) = nullptr;
    int bm = 128, bn = 128, nt = 256;
    if (K > 256 && al) {
        switch (g_gemm_variant) {
        case 1: kern = gemm_bf16x3_nt_kernel<32, 2, 2, 2, true>; break;
        case 2: kern = gemm_bf16x3_nt_kernel<32, 2, 4, 2, true>; bn = 256; nt = 512; break;
        default: kern = gemm_bf16x3_nt_kernel<64, 2, 2, 1, true>;
        }
    } else if (K > 256) {
        kern = gemm_bf16x3_nt_kernel<64, 2, 2, 1, false>;
    } else {
        kern = al ? gemm_bf16x3_nt_kernel<32, 2, 2, 1, true> : gemm_bf16x3_nt_kernel<32, 2, 2, 1, false>;
    }
    dim3 grid((M + bm - 1) / bm, (N + bn - 1) / bn, batch);
    hipLaunchKernelGGL(kern, grid, dim3(nt), 0, stream, (const __bf16*)Ahi, (const __bf16*)Alo, (const __bf16*)Bhi,
                       (const __bf16*)Blo, C, M, N, K, lda, ldb, sA, sB, sC, ldc, beta, alpha, off_b, off_r, off_c);
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
