// Batched GEMM C = A B on the CDNA4 matrix cores with FP64 operands in HBM.
//
// Used by the mixed-precision lambda / RDM contractions. Operands are read as
// float64 and converted while staging through LDS; products accumulate in
// FP32 on MFMA and are written back as float64.
//   mode 0 (fp32):   v_mfma_f32_32x32x2_f32,   operands rounded to FP32
//   mode 1 (bf16x3): v_mfma_f32_32x32x16_bf16, operands split x = hi + lo into
//                    two bfloat16 parts; C += A_hi B_hi + A_hi B_lo + A_lo B_hi
// A: [batch][M][K] (row stride lda), B: [batch][K][N] (ldb), C: [batch][M][N] (ldc).
// Block 256 threads = 4 waves, 64x64 C tile (32x32 per wave), K staged by 32.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int PADF = 1;   // fp32 A rows
constexpr int PADB = 8;   // bf16 rows (keeps 16-byte alignment)

__device__ __forceinline__ __bf16 to_bf16(float x) { return __float2bfloat16(x); }
__device__ __forceinline__ float from_bf16(__bf16 x) { return __bfloat162float(x); }

template <int MODE>
__global__ void __launch_bounds__(256) gemm_mixed_kernel(const double* __restrict__ A, const double* __restrict__ B,
                                                          double* __restrict__ C, int M, int N, int K, long long lda,
                                                          long long ldb, long long ldc, long long sA, long long sB,
                                                          long long sC)
{
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const long long bz = blockIdx.z;
    const double* Ab = A + bz * sA;
    const double* Bb = B + bz * sB;
    double* Cb = C + bz * sC;

    // fp32: As[BM][BK+PADF], Bs[BK][BN]; bf16x3: A{hi,lo}[BM][BK+PADB], B{hi,lo}[BN][BK+PADB] (transposed)
    __shared__ __attribute__((aligned(16))) float smem_f[(MODE == 0) ? (BM * (BK + PADF) + BK * BN) : 1];
    __shared__ __attribute__((aligned(16))) __bf16 smem_b[(MODE == 1) ? (4 * BM * (BK + PADB)) : 8];

    v16f acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += BK) {
        // ---- stage A and B tiles (zero padding outside the matrix) ----
        for (int e = tid; e < BM * BK; e += 256) {
            const int r = e / BK, c = e - (e / BK) * BK;
            const int gr = m0 + r, gc = k0 + c;
            const double v = (gr < M && gc < K) ? Ab[(long long)gr * lda + gc] : 0.0;
            if (MODE == 0) {
                smem_f[r * (BK + PADF) + c] = (float)v;
            } else {
                const float f = (float)v;
                const __bf16 hi = to_bf16(f);
                const __bf16 lo = to_bf16(f - from_bf16(hi));
                smem_b[r * (BK + PADB) + c] = hi;
                smem_b[BM * (BK + PADB) + r * (BK + PADB) + c] = lo;
            }
        }
        for (int e = tid; e < BK * BN; e += 256) {
            const int r = e / BN, c = e - (e / BN) * BN;   // r: k, c: n
            const int gr = k0 + r, gc = n0 + c;
            const double v = (gr < K && gc < N) ? Bb[(long long)gr * ldb + gc] : 0.0;
            if (MODE == 0) {
                smem_f[BM * (BK + PADF) + r * BN + c] = (float)v;
            } else {
                const float f = (float)v;
                const __bf16 hi = to_bf16(f);
                const __bf16 lo = to_bf16(f - from_bf16(hi));
                smem_b[2 * BM * (BK + PADB) + c * (BK + PADB) + r] = hi;
                smem_b[3 * BM * (BK + PADB) + c * (BK + PADB) + r] = lo;
            }
        }
        __syncthreads();
        const int row = wm * 32 + (lane & 31);
        const int col = wn * 32 + (lane & 31);
        const int half = lane >> 5;
        if (MODE == 0) {
            const float* As = smem_f;
            const float* Bs = smem_f + BM * (BK + PADF);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 2) {
                const float a = As[row * (BK + PADF) + kk + half];
                const float b = Bs[(kk + half) * BN + col];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
            }
        } else {
            const __bf16* Ah = smem_b;
            const __bf16* Al = smem_b + BM * (BK + PADB);
            const __bf16* Bh = smem_b + 2 * BM * (BK + PADB);
            const __bf16* Bl = smem_b + 3 * BM * (BK + PADB);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                const int ko = kk + half * 8;
                const v8bf ah = *reinterpret_cast<const v8bf*>(Ah + row * (BK + PADB) + ko);
                const v8bf al = *reinterpret_cast<const v8bf*>(Al + row * (BK + PADB) + ko);
                const v8bf bh = *reinterpret_cast<const v8bf*>(Bh + col * (BK + PADB) + ko);
                const v8bf bl = *reinterpret_cast<const v8bf*>(Bl + col * (BK + PADB) + ko);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // ---- store: lane holds C[8*(j/4) + 4*(lane/32) + j%4][lane%32] ----
    const int gc = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int gr = m0 + wm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        if (gr < M && gc < N) Cb[(long long)gr * ldc + gc] = (double)acc[j];
    }
}

}  // namespace

hipError_t ccsd_launch_gemm_mixed(int mode, const double* A, const double* B, double* C, int M, int N, int K,
                                  long long lda, long long ldb, long long ldc, long long sA, long long sB,
                                  long long sC, int batch, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || batch <= 0) return hipSuccess;
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, batch);
    if (mode == 0)
        hipLaunchKernelGGL(gemm_mixed_kernel<0>, grid, dim3(256), 0, stream, A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC);
    else
        hipLaunchKernelGGL(gemm_mixed_kernel<1>, grid, dim3(256), 0, stream, A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC);
    return hipGetLastError();
}
