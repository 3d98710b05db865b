// Lambda-update helpers over (o, o, v, v) tensors (HBM-bound, FP64).
//
// l1_block_kernel + l1_reduce_kernel: the singles terms that contract a doubles-
// shaped tensor X[i][j][a][b] with a singles-shaped vector y[j][b],
//   out[i][a] += sum_{j,b} (c1 X[i,j,a,b] + c2 X[i,j,b,a] + c3 X[j,i,a,b] + c4 X[j,i,b,a]) y[j,b],
// in one read of X. Workgroup (p, q) reads the v x v block B = X[p][q] once and
// forms B y_q and B^T y_q (the part of out[p]) and B y_p, B^T y_p (the part of
// out[q]); the per-block parts go to a workspace that a second kernel sums in a
// fixed order (no atomics: graph replays and eager calls agree bit for bit).
//
// unpack_pairs_kernel: the (i <= j) pair blocks of a doubles-shaped product (the
// symmetric half of the l2 (ac|bd) term) written to both (i, j) and, transposed,
// (j, i) of a dense (o, o, v, v) tensor.
//
// pair_combine_kernel: for every (i, j) and 32 x 32 tile of (a, b),
//   out[i,j,a,b] = beta * out + (c1 X[i,j,a,b] + c2 X[i,j,b,a] + c3 X[j,i,a,b] + c4 X[j,i,b,a]) * D[i,j,a,b]
// with the transposed blocks staged through LDS so that every global read and
// write runs along b (coalesced when X's and out's b strides are 1); X and out
// are any strided 4-index views (a permuted layout such as [i][a][j][b]), D
// optional and contiguous.
//
// vvvv_pm_pack_kernel / vvvv_pm_unpack_kernel: the l2 (ac|bd) term over index pairs.
// With P = (c, d), c >= d, and the i <= j pair blocks A = l2[i,j],
//   A+[P] = (A[c,d] + A[d,c]) (1/2 when c = d),  A-[P] = A[c,d] - A[d,c],
// the products X+- = A+- W+-^T with W+-[Q, P] = (ac|bd) +- (ad|bc) give
//   X[a,b] = (X+[Q] + s X-[Q]) / 2,  Q = (max(a,b), min(a,b)), s = +1 (a >= b) or -1,
// which the unpack writes to out[i,j] and, transposed, out[j,i] (times a scale).
#include <hip/hip_runtime.h>

#include "bf16_split.h"

namespace {

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// NC = number of 64-column chunks the kernel is compiled for (>= ceil(v / 64)); each
// wave takes 4 of its rows per step and all their chunks, so 4 NC loads are in
// flight before the row sums are reduced (one wave reduction per row and vector).
template <int NC>
__global__ void __launch_bounds__(256) l1_block_kernel(const double* __restrict__ X, int o, int v,
                                                       const double* __restrict__ y, long long sy0, long long sy1,
                                                       double c1, double c2, double c3, double c4,
                                                       double* __restrict__ ws)
{
    extern __shared__ double sh[];
    double* yq = sh;             // y[q][:]
    double* yp = yq + v;         // y[p][:]
    double* rq = yp + v;         // (B y_q)[r]
    double* rp = rq + v;         // (B y_p)[r]
    double* cq = rp + v;         // per wave: partial (B^T y_q)[c]
    double* cp = cq + 4 * v;     // per wave: partial (B^T y_p)[c]
    const int p = blockIdx.x / o, q = blockIdx.x - (blockIdx.x / o) * o;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int a = tid; a < v; a += 256) {
        yq[a] = y[q * sy0 + a * sy1];
        yp[a] = y[p * sy0 + a * sy1];
    }
    __syncthreads();
    const double* B = X + ((long long)p * o + q) * v * v;
    const bool rows = (c1 != 0.0) || (c3 != 0.0);
    const bool cols = (c2 != 0.0) || (c4 != 0.0);
    double colq[NC], colp[NC], yqc[NC], ypc[NC];
#pragma unroll
    for (int ch = 0; ch < NC; ch++) {
        const int c = ch * 64 + lane;
        colq[ch] = 0.0;
        colp[ch] = 0.0;
        yqc[ch] = c < v ? yq[c] : 0.0;
        ypc[ch] = c < v ? yp[c] : 0.0;
    }
    for (int r0 = wave; r0 < v; r0 += 16) {
        double pq[4], pp[4], yr_q[4], yr_p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = r0 + 4 * u;
            pq[u] = 0.0;
            pp[u] = 0.0;
            yr_q[u] = r < v ? yq[r] : 0.0;
            yr_p[u] = r < v ? yp[r] : 0.0;
        }
        // clamped (always in bounds) reads, so that all 4 NC of them are in flight at once:
        // a clamped column meets a zero weight (yqc, ypc) and its column sums are never
        // stored, a clamped row a zero weight (yr_q, yr_p) and its row sums are never stored
        double bv[NC][4];
#pragma unroll
        for (int ch = 0; ch < NC; ch++) {
            const int c = min(ch * 64 + lane, v - 1);
#pragma unroll
            for (int u = 0; u < 4; u++) bv[ch][u] = B[(long long)min(r0 + 4 * u, v - 1) * v + c];
        }
#pragma unroll
        for (int ch = 0; ch < NC; ch++) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double b = bv[ch][u];
                pq[u] += b * yqc[ch];
                pp[u] += b * ypc[ch];
                colq[ch] += b * yr_q[u];
                colp[ch] += b * yr_p[u];
            }
        }
        if (rows) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = r0 + 4 * u;
                const double sq = wave_sum(pq[u]);
                const double sp = wave_sum(pp[u]);
                if (lane == 0 && r < v) {
                    rq[r] = sq;
                    rp[r] = sp;
                }
            }
        }
    }
#pragma unroll
    for (int ch = 0; ch < NC; ch++) {
        const int c = ch * 64 + lane;
        if (c < v) {
            cq[wave * v + c] = colq[ch];
            cp[wave * v + c] = colp[ch];
        }
    }
    __syncthreads();
    const long long nblk = (long long)o * o * v;
    double* Wp = ws + ((long long)p * o + q) * v;          // part of out[p]
    double* Wq = ws + nblk + ((long long)p * o + q) * v;   // part of out[q]
    for (int a = tid; a < v; a += 256) {
        double wp = 0.0, wq = 0.0;
        if (rows) {
            wp += c1 * rq[a];
            wq += c3 * rp[a];
        }
        if (cols) {
            const double sq = (cq[a] + cq[v + a]) + (cq[2 * v + a] + cq[3 * v + a]);
            const double sp = (cp[a] + cp[v + a]) + (cp[2 * v + a] + cp[3 * v + a]);
            wp += c2 * sq;
            wq += c4 * sp;
        }
        Wp[a] = wp;
        Wq[a] = wq;
    }
}

__global__ void l1_reduce_kernel(const double* __restrict__ ws, int o, int v, double* __restrict__ out)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)o * v) return;
    const int i = (int)(t / v), a = (int)(t - (t / v) * v);
    const long long nblk = (long long)o * o * v;
    double s = 0.0;
    for (int q = 0; q < o; q++) s += ws[((long long)i * o + q) * v + a];
    for (int p = 0; p < o; p++) s += ws[nblk + ((long long)p * o + i) * v + a];
    out[t] += s;
}

constexpr int PT = 32;   // pair_combine tile edge

__global__ void __launch_bounds__(256) pair_combine_kernel(const double* __restrict__ X, long long x0, long long x1,
                                                           long long x2, long long x3, int o, int v, double c1,
                                                           double c2, double c3, double c4,
                                                           const double* __restrict__ D, double* __restrict__ out,
                                                           long long s0, long long s1, long long s2, long long s3,
                                                           int beta)
{
    __shared__ double tij[PT][PT + 1];   // X[i][j][b][a] for the tile, indexed [b - b0][a - a0]
    __shared__ double tji[PT][PT + 1];   // X[j][i][b][a]
    const int nt = (v + PT - 1) / PT;
    const long long blk = blockIdx.x;
    const int tile = (int)(blk % ((long long)nt * nt));
    const long long ij = blk / ((long long)nt * nt);
    const int i = (int)(ij / o), j = (int)(ij - (ij / o) * o);
    const int a0 = (tile / nt) * PT, b0 = (tile - (tile / nt) * nt) * PT;
    const int tx = threadIdx.x & (PT - 1), ty = threadIdx.x >> 5;   // 32 x 8
    const long long vv = (long long)v * v;
    const double* Xij = X + (long long)i * x0 + (long long)j * x1;
    const double* Xji = X + (long long)j * x0 + (long long)i * x1;
    // every read below uses indices clamped into range (the results of clamped lanes are
    // never stored), so that a thread's reads carry no branches and are all in flight at once
    const int vm = v - 1;
    if (c2 != 0.0 || c4 != 0.0) {
        double t1v[PT / 8], t2v[PT / 8];
#pragma unroll
        for (int k = 0; k < PT; k += 8) {
            const long long e = (long long)min(b0 + ty + k, vm) * x2 + (long long)min(a0 + tx, vm) * x3;
            t1v[k / 8] = c2 != 0.0 ? Xij[e] : 0.0;
            t2v[k / 8] = c4 != 0.0 ? Xji[e] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < PT; k += 8) {
            tij[ty + k][tx] = t1v[k / 8];
            tji[ty + k][tx] = t2v[k / 8];
        }
    }
    __syncthreads();
    double x1v[PT / 8], x3v[PT / 8], dv[PT / 8], ov[PT / 8];
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int a = min(a0 + ty + k, vm), b = min(b0 + tx, vm);
        const long long e = (long long)a * x2 + (long long)b * x3;
        x1v[k / 8] = c1 != 0.0 ? Xij[e] : 0.0;
        x3v[k / 8] = c3 != 0.0 ? Xji[e] : 0.0;
        dv[k / 8] = D ? D[((long long)i * o + j) * vv + (long long)a * v + b] : 1.0;
        ov[k / 8] = beta ? out[(long long)i * s0 + (long long)j * s1 + (long long)a * s2 + (long long)b * s3] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int a = a0 + ty + k, b = b0 + tx;
        if (a < v && b < v) {
            double val = c1 * x1v[k / 8] + c3 * x3v[k / 8];
            if (c2 != 0.0) val += c2 * tij[tx][ty + k];
            if (c4 != 0.0) val += c4 * tji[tx][ty + k];
            out[(long long)i * s0 + (long long)j * s1 + (long long)a * s2 + (long long)b * s3] =
                ov[k / 8] + val * dv[k / 8];
        }
    }
}

// out[i,j,a,b] = s X[p,a,b] and out[j,i,b,a] = s X[p,a,b] for the pair p = (iu[p], ju[p]),
// iu <= ju; on the diagonal (i == j) out[i,i,a,b] = s (X[p,a,b] + X[p,b,a]) / 2. One
// block per (pair, 32 x 32 tile); the transposed block is written from an LDS tile.
__global__ void __launch_bounds__(256) unpack_pairs_kernel(const double* __restrict__ X,
                                                           const long long* __restrict__ iu,
                                                           const long long* __restrict__ ju, int o, int v,
                                                           long long npair, double s, double* __restrict__ out)
{
    __shared__ double t1[PT][PT + 1];   // X[p][a0 + r][b0 + c]
    __shared__ double t2[PT][PT + 1];   // X[p][b0 + r][a0 + c] (diagonal pairs)
    const int nt = (v + PT - 1) / PT;
    const long long blk = blockIdx.x;
    const int tile = (int)(blk % ((long long)nt * nt));
    const long long p = blk / ((long long)nt * nt);
    if (p >= npair) return;
    const int i = (int)iu[p], j = (int)ju[p];
    const int a0 = (tile / nt) * PT, b0 = (tile - (tile / nt) * nt) * PT;
    const int tx = threadIdx.x & (PT - 1), ty = threadIdx.x >> 5;
    const long long vv = (long long)v * v;
    const double* Xp = X + p * vv;
    // clamped reads (entries outside the v x v block are never written out), all in flight at once
    const int vm = v - 1;
    double r1[PT / 8], r2[PT / 8];
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int r = ty + k;
        r1[k / 8] = Xp[(long long)min(a0 + r, vm) * v + min(b0 + tx, vm)];
        r2[k / 8] = i == j ? Xp[(long long)min(b0 + r, vm) * v + min(a0 + tx, vm)] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        t1[ty + k][tx] = r1[k / 8];
        t2[ty + k][tx] = r2[k / 8];
    }
    __syncthreads();
    double* Oij = out + ((long long)i * o + j) * vv;
    double* Oji = out + ((long long)j * o + i) * vv;
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int r = ty + k;
        if (i == j) {
            if (a0 + r < v && b0 + tx < v) Oij[(long long)(a0 + r) * v + b0 + tx] = 0.5 * s * (t1[r][tx] + t2[tx][r]);
        } else {
            if (a0 + r < v && b0 + tx < v) Oij[(long long)(a0 + r) * v + b0 + tx] = s * t1[r][tx];
            // out[j,i][b0 + r][a0 + tx] = X[p][a0 + tx][b0 + r]
            if (b0 + r < v && a0 + tx < v) Oji[(long long)(b0 + r) * v + a0 + tx] = s * t1[tx][r];
        }
    }
}

// Row r of the output is the pair (iu[r], ju[r]), column P < np the index pair
// (tc[P], td[P]), tc >= td; columns np..ld-1 are zero (K padding of a GEMM operand).
// SPLIT: write bf16 hi / lo planes (out, lo) instead of float64 (out).
template <bool SPLIT>
__global__ void __launch_bounds__(256) vvvv_pm_pack_kernel(const double* __restrict__ l2, int o, int v,
                                                            const long long* __restrict__ iu,
                                                            const long long* __restrict__ ju, long long npair,
                                                            const long long* __restrict__ tc,
                                                            const long long* __restrict__ td, long long np,
                                                            long long ld, void* __restrict__ out,
                                                            uint16_t* __restrict__ lo)
{
    const long long n = npair * ld;
    const long long plane = npair * ld;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long r = t / ld, P = t - r * ld;
        double plus = 0.0, minus = 0.0;
        if (P < np) {
            const int c = (int)tc[P], d = (int)td[P];
            const double* A = l2 + (iu[r] * o + ju[r]) * (long long)v * v;
            const double x = A[(long long)c * v + d], y = A[(long long)d * v + c];
            plus = (c == d ? 0.5 : 1.0) * (x + y);
            minus = x - y;
        }
        if (SPLIT) {
            uint16_t* hi = static_cast<uint16_t*>(out);
            ccsd_split_bf16(plus, hi[t], lo[t]);
            ccsd_split_bf16(minus, hi[plane + t], lo[plane + t]);
        } else {
            double* o64 = static_cast<double*>(out);
            o64[t] = plus;
            o64[plane + t] = minus;
        }
    }
}

// out (o, o, v, v) from X+- (2, npair, ldx), npair = o(o+1)/2 pairs i <= j in triu order.
// With P = (X+ + X-)[Q] / 2 and M = (X+ - X-)[Q] / 2 at Q = (a, b), a >= b, the pair
// (i, j), i < j, gives out[i,j,a,b] = P, out[i,j,b,a] = M, out[j,i,b,a] = P and
// out[j,i,a,b] = M (M only for a > b; X-[Q] is zero on a = b); the pair (i, i) gives
// out[i,i,a,b] = out[i,i,b,a] = X+[Q] / 2; everything times `scale`. One block per (pair,
// 32 x 32 tile (A, B) with A >= B of the v x v block): the packed rows of a tile are
// read along b, the entries a >= b written straight, the mirrored ones (b, a) from an
// LDS copy, so that every global read and write runs along the last index.
__global__ void __launch_bounds__(256) vvvv_pm_unpack_kernel(const double* __restrict__ X, int o, int v,
                                                              long long ldx, double scale, double* __restrict__ out)
{
    __shared__ double tp[PT][PT + 1];   // P (or X+ / 2 on i = j) at [a - a0][b - b0]
    __shared__ double tm[PT][PT + 1];   // M
    const int nt = (v + PT - 1) / PT;
    const int ntri = nt * (nt + 1) / 2;
    const long long r = blockIdx.x / ntri;
    const int t = (int)(blockIdx.x - r * ntri);
    int A = 0;
    while ((A + 1) * (A + 2) / 2 <= t) A++;
    const int B = t - A * (A + 1) / 2;
    // pair (i, j) of triu row r: row i starts at i o - i (i - 1) / 2
    int i = 0;
    while (i + 1 < o && (long long)(i + 1) * o - (long long)(i + 1) * i / 2 <= r) i++;
    const int j = i + (int)(r - ((long long)i * o - (long long)i * (i - 1) / 2));
    const int a0 = A * PT, b0 = B * PT;
    const int tx = threadIdx.x & (PT - 1), ty = threadIdx.x >> 5;   // 32 x 8
    const long long vv = (long long)v * v;
    const long long plane = (long long)o * (o + 1) / 2 * ldx;
    const double* Xp = X + r * ldx;
    const double* Xm = Xp + plane;
    double* Oij = out + ((long long)i * o + j) * vv;
    double* Oji = out + ((long long)j * o + i) * vv;
    const bool diag = i == j;
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int a = a0 + ty + k, b = b0 + tx;
        double P = 0.0, M = 0.0;
        if (a < v && b <= a) {
            const long long Q = (long long)a * (a + 1) / 2 + b;
            const double xp = Xp[Q];
            const double xm = diag ? 0.0 : Xm[Q];
            P = 0.5 * (xp + xm);
            M = 0.5 * (xp - xm);
            Oij[(long long)a * v + b] = scale * P;
            if (!diag) Oji[(long long)a * v + b] = scale * (a > b ? M : P);
        }
        tp[ty + k][tx] = P;
        tm[ty + k][tx] = M;
    }
    __syncthreads();
    // the mirrored entries: (row b, column a) for a > b, a = a0 + tx, b = b0 + ty + k
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int a = a0 + tx, b = b0 + ty + k;
        if (a < v && b < a) {
            const double P = tp[tx][ty + k], M = tm[tx][ty + k];
            Oij[(long long)b * v + a] = scale * (diag ? P : M);
            if (!diag) Oji[(long long)b * v + a] = scale * P;
        }
    }
}

// out[i,j,a,b] = (X[i,j,a,b] + X[j,i,b,a] + u[i,a] w[j,b] + u[j,b] w[i,a] + s S[i,j,a,b]) D[i,j,a,b]:
// the closing symmetrisation of the lambda update (divide by D_ijab) with the two terms
// that enter l2 just before it folded in (a rank-one l1 x w4 term and a fixed integral
// term weighted by s; S already symmetric under (i,a) <-> (j,b)). X, S, D and out
// contiguous (o, o, v, v), u and w contiguous (o, v); S may be NULL; s = *sp when sp is
// given (a device scalar read when the kernel runs, so one captured graph serves any
// weight), else sh. One block per (i, j, 32 x 32 tile); X[j,i] is transposed through LDS.
__global__ void __launch_bounds__(256) l2_finish_kernel(const double* __restrict__ X, int o, int v,
                                                         const double* __restrict__ u, const double* __restrict__ w,
                                                         const double* __restrict__ S, const double* __restrict__ sp,
                                                         double sh, const double* __restrict__ D,
                                                         double* __restrict__ out)
{
    __shared__ double tji[PT][PT + 1];   // X[j][i][b][a] for the tile, indexed [b - b0][a - a0]
    const int nt = (v + PT - 1) / PT;
    const long long blk = blockIdx.x;
    const int tile = (int)(blk % ((long long)nt * nt));
    const long long ij = blk / ((long long)nt * nt);
    const int i = (int)(ij / o), j = (int)(ij - (ij / o) * o);
    const int a0 = (tile / nt) * PT, b0 = (tile - (tile / nt) * nt) * PT;
    const int tx = threadIdx.x & (PT - 1), ty = threadIdx.x >> 5;   // 32 x 8
    const long long vv = (long long)v * v;
    const double* Xij = X + ((long long)i * o + j) * vv;
    const double* Xji = X + ((long long)j * o + i) * vv;
    const int vm = v - 1;
    // clamped reads (results of clamped lanes are never stored)
    double t1v[PT / 8];
#pragma unroll
    for (int k = 0; k < PT; k += 8) t1v[k / 8] = Xji[(long long)min(b0 + ty + k, vm) * v + min(a0 + tx, vm)];
#pragma unroll
    for (int k = 0; k < PT; k += 8) tji[ty + k][tx] = t1v[k / 8];
    __syncthreads();
    const double s = sp ? *sp : sh;
    const int bc = min(b0 + tx, vm);
    const double ujb = u[(long long)j * v + bc], wjb = w[(long long)j * v + bc];
    const long long base = ((long long)i * o + j) * vv;
    double xv[PT / 8], dv[PT / 8], sv[PT / 8], uw[PT / 8];
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int a = min(a0 + ty + k, vm);
        const long long e = (long long)a * v + bc;
        xv[k / 8] = Xij[e];
        dv[k / 8] = D[base + e];
        sv[k / 8] = S ? S[base + e] : 0.0;
        uw[k / 8] = u[(long long)i * v + a] * wjb + ujb * w[(long long)i * v + a];
    }
#pragma unroll
    for (int k = 0; k < PT; k += 8) {
        const int a = a0 + ty + k, b = b0 + tx;
        if (a < v && b < v)
            out[base + (long long)a * v + b] = (xv[k / 8] + tji[tx][ty + k] + uw[k / 8] + s * sv[k / 8]) * dv[k / 8];
    }
}

// DIIS with the history in FP32 (the mixed-precision lambda solve): one pass that
// stores a new vector and its error vector and forms the error's overlaps with the
// stored ones, and one pass that forms a combination of stored rows, both summing in
// FP64. The new vector x is the concatenation of two FP64 segments (s0: n0 elements,
// s1: the rest) plus an optional FP64 right-hand side.
constexpr int kDiisMaxRows = 15;

// x[k]; xs = f32(x) when given; prev_out = x when given; with prev: e = x - prev, es =
// f32(e) when given, and the per-block partials of sum_k E[r][k] ef[k] for r < nrows (the
// stored row r, or ef itself for r = slot) and of sum_k ef[k]^2 (ef = the stored e, or e
// in full when es is NULL) into part[blockIdx.x][0..nrows].
__global__ void __launch_bounds__(256) diis_push_kernel(const double* __restrict__ s0, const double* __restrict__ s1,
                                                         long long n0, long long n, const double* __restrict__ rhs,
                                                         const double* __restrict__ prev, float* __restrict__ xs,
                                                         float* __restrict__ es, const float* __restrict__ E,
                                                         long long lde, int nrows, int slot,
                                                         double* __restrict__ prev_out, double* __restrict__ part)
{
    double acc[kDiisMaxRows + 1];
#pragma unroll
    for (int r = 0; r <= kDiisMaxRows; r++) acc[r] = 0.0;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        double x = k < n0 ? s0[k] : s1[k - n0];
        if (rhs) x += rhs[k];
        if (xs) xs[k] = (float)x;
        if (prev_out) prev_out[k] = x;
        if (prev) {
            const double e = x - prev[k];
            double ef = e;
            if (es) {
                const float f = (float)e;
                es[k] = f;
                ef = (double)f;
            }
#pragma unroll
            for (int r = 0; r < kDiisMaxRows; r++)
                if (r < nrows) acc[r] += (r == slot ? ef : (double)E[(long long)r * lde + k]) * ef;
            acc[kDiisMaxRows] += ef * ef;
        }
    }
    if (!prev) return;
    __shared__ double red[4][kDiisMaxRows + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r <= kDiisMaxRows; r++) {
        if (r < nrows || r == kDiisMaxRows) {
            const double v = wave_sum(acc[r]);
            if (lane == 0) red[wave][r] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x <= nrows) {
        const int r = threadIdx.x < nrows ? threadIdx.x : kDiisMaxRows;
        part[(long long)blockIdx.x * (nrows + 1) + threadIdx.x] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
    }
}

// out[j] = sum_b part[b][j] for j = blockIdx.x <= nrows: thread t adds blocks t, t + 256,
// ..., then a fixed-order tree over the threads (deterministic).
__global__ void __launch_bounds__(256) diis_dots_kernel(const double* __restrict__ part, int nblocks, int nrows,
                                                         double* __restrict__ out)
{
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) s += part[(long long)b * (nrows + 1) + j];
    __shared__ double red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[j] = (red[0] + red[1]) + (red[2] + red[3]);
}

struct DiisCoef {
    double c[kDiisMaxRows + 1];
};

// out[k] = base[k] (0 when NULL) + sum_r c[r] X[r][k], X FP32 rows of stride ldx.
__global__ void __launch_bounds__(256) diis_combine_kernel(const float* __restrict__ X, long long ldx, int nrows,
                                                            const DiisCoef c, const double* __restrict__ base,
                                                            double* __restrict__ out, long long n)
{
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        double s = base ? base[k] : 0.0;
#pragma unroll
        for (int r = 0; r <= kDiisMaxRows; r++)
            if (r < nrows) s += c.c[r] * (double)X[(long long)r * ldx + k];
        out[k] = s;
    }
}

}  // namespace

hipError_t ccsd_launch_vvvv_pm_pack(const double* l2, int o, int v, const long long* iu, const long long* ju,
                                    long long npair, const long long* tc, const long long* td, long long np,
                                    long long ld, int split, void* out, void* lo, hipStream_t stream)
{
    const long long n = npair * ld;
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    if (split)
        hipLaunchKernelGGL(vvvv_pm_pack_kernel<true>, dim3((unsigned)nb), dim3(256), 0, stream, l2, o, v, iu, ju, npair,
                           tc, td, np, ld, out, (uint16_t*)lo);
    else
        hipLaunchKernelGGL(vvvv_pm_pack_kernel<false>, dim3((unsigned)nb), dim3(256), 0, stream, l2, o, v, iu, ju,
                           npair, tc, td, np, ld, out, (uint16_t*)nullptr);
    return hipGetLastError();
}

hipError_t ccsd_launch_vvvv_pm_unpack(const double* X, int o, int v, long long ldx, double scale, double* out,
                                      hipStream_t stream)
{
    if (o <= 0 || v <= 0) return hipSuccess;
    const long long nt = (v + PT - 1) / PT;
    const long long nb = (long long)o * (o + 1) / 2 * (nt * (nt + 1) / 2);
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(vvvv_pm_unpack_kernel, dim3((unsigned)nb), dim3(256), 0, stream, X, o, v, ldx, scale, out);
    return hipGetLastError();
}

hipError_t ccsd_launch_l2_finish(const double* X, int o, int v, const double* u, const double* w, const double* S,
                                 const double* sp, double sh, const double* D, double* out, hipStream_t stream)
{
    if (o <= 0 || v <= 0) return hipSuccess;
    const long long nt = (v + PT - 1) / PT;
    const long long nb = (long long)o * o * nt * nt;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(l2_finish_kernel, dim3((unsigned)nb), dim3(256), 0, stream, X, o, v, u, w, S, sp, sh, D, out);
    return hipGetLastError();
}

hipError_t ccsd_launch_unpack_pairs(const double* X, const long long* iu, const long long* ju, int o, int v,
                                    long long npair, double s, double* out, hipStream_t stream)
{
    if (o <= 0 || v <= 0 || npair <= 0) return hipSuccess;
    const long long nt = (v + PT - 1) / PT;
    const long long nb = npair * nt * nt;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(unpack_pairs_kernel, dim3((unsigned)nb), dim3(256), 0, stream, X, iu, ju, o, v, npair, s, out);
    return hipGetLastError();
}

size_t ccsd_l1_terms_lds(int v) { return (size_t)12 * v * sizeof(double); }

hipError_t ccsd_launch_l1_terms(const double* X, int o, int v, const double* y, long long sy0, long long sy1,
                                double c1, double c2, double c3, double c4, double* ws, double* out,
                                hipStream_t stream)
{
    if (o <= 0 || v <= 0) return hipSuccess;
    const int nch = (v + 63) / 64;
    const dim3 grid((unsigned)((long long)o * o)), block(256);
    const size_t lds = ccsd_l1_terms_lds(v);
    if (nch <= 1)
        hipLaunchKernelGGL(l1_block_kernel<1>, grid, block, lds, stream, X, o, v, y, sy0, sy1, c1, c2, c3, c4, ws);
    else if (nch <= 2)
        hipLaunchKernelGGL(l1_block_kernel<2>, grid, block, lds, stream, X, o, v, y, sy0, sy1, c1, c2, c3, c4, ws);
    else if (nch <= 3)
        hipLaunchKernelGGL(l1_block_kernel<3>, grid, block, lds, stream, X, o, v, y, sy0, sy1, c1, c2, c3, c4, ws);
    else if (nch <= 4)
        hipLaunchKernelGGL(l1_block_kernel<4>, grid, block, lds, stream, X, o, v, y, sy0, sy1, c1, c2, c3, c4, ws);
    else if (nch <= 6)
        hipLaunchKernelGGL(l1_block_kernel<6>, grid, block, lds, stream, X, o, v, y, sy0, sy1, c1, c2, c3, c4, ws);
    else if (nch <= 8)
        hipLaunchKernelGGL(l1_block_kernel<8>, grid, block, lds, stream, X, o, v, y, sy0, sy1, c1, c2, c3, c4, ws);
    else if (nch <= 11)
        hipLaunchKernelGGL(l1_block_kernel<11>, grid, block, lds, stream, X, o, v, y, sy0, sy1, c1, c2, c3, c4, ws);
    else
        return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long n = (long long)o * v;
    hipLaunchKernelGGL(l1_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ws, o, v, out);
    return hipGetLastError();
}

hipError_t ccsd_launch_pair_combine(const double* X, const long long* xs, int o, int v, double c1, double c2,
                                    double c3, double c4, const double* D, double* out, const long long* os, int beta,
                                    hipStream_t stream)
{
    if (o <= 0 || v <= 0) return hipSuccess;
    const long long nt = (v + PT - 1) / PT;
    const long long nb = (long long)o * o * nt * nt;
    hipLaunchKernelGGL(pair_combine_kernel, dim3((unsigned)nb), dim3(256), 0, stream, X, xs[0], xs[1], xs[2], xs[3], o,
                       v, c1, c2, c3, c4, D, out, os[0], os[1], os[2], os[3], beta);
    return hipGetLastError();
}

int ccsd_diis_max_rows() { return kDiisMaxRows; }

// Blocks of the push pass: its partials need ws of blocks x (nrows + 1) doubles.
long long ccsd_diis_push_blocks(long long n)
{
    long long nb = (n + 255) / 256;
    return nb < 4096 ? (nb < 1 ? 1 : nb) : 4096;
}

hipError_t ccsd_launch_diis_push(const double* s0, const double* s1, long long n0, long long n, const double* rhs,
                                 const double* prev, float* xs, float* es, const float* E, long long lde, int nrows,
                                 int slot, double* prev_out, double* dots, double* ws, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const long long nb = ccsd_diis_push_blocks(n);
    hipLaunchKernelGGL(diis_push_kernel, dim3((unsigned)nb), dim3(256), 0, stream, s0, s1, n0, n, rhs, prev, xs, es, E,
                       lde, nrows, slot, prev_out, ws);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !prev) return e;
    hipLaunchKernelGGL(diis_dots_kernel, dim3((unsigned)(nrows + 1)), dim3(256), 0, stream, ws, (int)nb, nrows, dots);
    return hipGetLastError();
}

hipError_t ccsd_launch_diis_combine(const float* X, long long ldx, int nrows, const double* coef, const double* base,
                                    double* out, long long n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    DiisCoef c;
    for (int r = 0; r <= kDiisMaxRows; r++) c.c[r] = r < nrows ? coef[r] : 0.0;
    long long nb = (n + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    hipLaunchKernelGGL(diis_combine_kernel, dim3((unsigned)nb), dim3(256), 0, stream, X, ldx, nrows, c, base, out, n);
    return hipGetLastError();
}
