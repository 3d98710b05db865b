// One-electron integrals and first derivatives on gfx950, FP64.
//
// The reference evaluates these on the CPU through PySCF (hcore_generator,
// get_ovlp, int1e_ipovlp / ipkin / ipnuc / iprinv at ccsd_gpu/gpu_ccsd_grad.py:554-583).
// One thread per (ordered shell pair, bra primitive) adds its part of the pair's
// Cartesian block for overlap and kinetic (1-D Obara-Saika tables); the
// nuclear-attraction kinds use one thread per (shell pair, bra primitive,
// nucleus), Rys quadrature with point charges (the ERI tables, n <= 5 for l <= 4
// plus one derivative). The parts of one element are added with exact fixed-point
// integer atomics (defx_add1: two 64-bit words per element, x to 2^-60, must start
// at zero) and converted to float64 once at the end, so small bases still fill the
// device and the result is the same bit for bit on every run.
//
// kinds: 0 <i|j>  1 <i|T|j>  2 <i|V|j>  3 <d/dr i|j>  4 <d/dr i|T|j>
//        5 <d/dr i| -Z_C/|r-C| |j> for every nucleus C (natm x 3 matrices)
#include "eri_common.h"
#include "rys_device.h"

namespace {

constexpr int L1 = 5;  // l + 1 derivative

__device__ __forceinline__ int ncart1(int l) { return (l + 1) * (l + 2) / 2; }

__device__ __forceinline__ void comp1(int l, int f, int* c)
{
    int n = 0;
    for (int x = l; x >= 0; x--)
        for (int y = l - x; y >= 0; y--) {
            if (n == f) { c[0] = x; c[1] = y; c[2] = l - x - y; return; }
            n++;
        }
}

__device__ __forceinline__ double snorm1(int l)
{
    return l == 0 ? 0.28209479177387814347 : (l == 1 ? 0.48860251190291992159 : 1.0);
}

// S[i][j] for i <= la, j <= lb (1-D overlap, Obara-Saika VRR then HRR).
__device__ void ovlp1d(double S[L1 + 1][L1 + 3], int la, int lb, double PA, double AB, double p)
{
    double t[2 * L1 + 4][L1 + 3];
    const int lt = la + lb;
    t[0][0] = sqrt(3.14159265358979323846 / p);
    for (int i = 0; i < lt; i++) t[i + 1][0] = PA * t[i][0] + (i > 0 ? i / (2 * p) * t[i - 1][0] : 0.0);
    for (int j = 0; j < lb; j++)
        for (int i = 0; i <= lt - j - 1; i++) t[i][j + 1] = t[i + 1][j] + AB * t[i][j];
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++) S[i][j] = t[i][j];
}

__device__ void nuc1d(double I[L1 + 1][L1 + 1], int la, int lb, double PA, double PC, double BA, double rt, double p,
                      double seed)
{
    double g[2 * L1 + 2][L1 + 1];
    const int lt = la + lb;
    const double c00 = PA - rt * PC;
    const double b10 = 0.5 / p * (1.0 - rt);
    g[0][0] = seed;
    if (lt > 0) g[1][0] = c00 * seed;
    for (int i = 1; i < lt; i++) g[i + 1][0] = c00 * g[i][0] + i * b10 * g[i - 1][0];
    for (int j = 0; j < lb; j++)
        for (int i = 0; i <= lt - j - 1; i++) g[i][j + 1] = g[i + 1][j] - BA * g[i][j];
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++) I[i][j] = g[i][j];
}

__device__ __forceinline__ void fx_add(unsigned long long* fx, size_t e, double v)
{
    defx_add1(fx + 1 + 2 * e, fx, v);
}

__global__ void int1e_kernel(int kind, unsigned long long* __restrict__ fx, const ShellInfo* __restrict__ shells,
                             const double* __restrict__ env, const double* __restrict__ coords,
                             const double* __restrict__ charges, int natm, int nbas, int nao, int maxprim)
{
    // one thread per (shell pair, bra primitive[, nucleus for the nuclear-attraction
    // kinds]); the threads of one element add atomically (fixed point)
    const bool per_nucleus = kind == 2 || kind == 5;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long npair = (long long)nbas * nbas;
    const long long nunit = npair * maxprim;
    if (t >= (per_nucleus ? nunit * natm : nunit)) return;
    const long long u = per_nucleus ? t / natm : t;
    const int pair = (int)(u / maxprim), ip_only = (int)(u - (u / maxprim) * maxprim);
    const int c_first = per_nucleus ? (int)(t - u * natm) : 0;
    const int c_last = per_nucleus ? c_first + 1 : natm;
    const int ish = pair / nbas, jsh = pair - ish * nbas;
    const ShellInfo si = shells[ish], sj = shells[jsh];
    if (ip_only >= si.nprim) return;
    const int la = si.l, lb = sj.l;
    const int deriv = kind >= 3;
    const int lae = la + deriv;
    const int nfa = ncart1(la), nfb = ncart1(lb);
    const double* A = coords + 3 * si.atom;
    const double* B = coords + 3 * sj.atom;
    const double nrm = snorm1(la) * snorm1(lb);
    const size_t nn = (size_t)nao * nao;
    double AB2 = 0.0;
    for (int x = 0; x < 3; x++) AB2 += (A[x] - B[x]) * (A[x] - B[x]);
    for (int ip = ip_only; ip <= ip_only; ip++)
        for (int jp = 0; jp < sj.nprim; jp++) {
            const double a = env[si.pexp + ip], b = env[sj.pexp + jp], p = a + b;
            const double K = exp(-a * b / p * AB2) * nrm;
            double P[3];
            for (int x = 0; x < 3; x++) P[x] = (a * A[x] + b * B[x]) / p;
            for (int ci = 0; ci < si.nctr; ci++)
                for (int cj = 0; cj < sj.nctr; cj++) {
                    const double cc = env[si.pcoef + ci * si.nprim + ip] * env[sj.pcoef + cj * sj.nprim + jp] * K;
                    const int i0 = si.ao_cart + ci * nfa, j0 = sj.ao_cart + cj * nfb;
                    if (kind <= 1 || kind == 3 || kind == 4) {
                        double S[3][L1 + 1][L1 + 3];
                        for (int x = 0; x < 3; x++) ovlp1d(S[x], lae, lb + 2, P[x] - A[x], A[x] - B[x], p);
                        for (int fa = 0; fa < nfa; fa++)
                            for (int fb = 0; fb < nfb; fb++) {
                                int ca[3], cb[3];
                                comp1(la, fa, ca);
                                comp1(lb, fb, cb);
                                const int ncomp = deriv ? 3 : 1;
                                for (int c = 0; c < ncomp; c++) {
                                    double s[3], tk[3];
                                    for (int x = 0; x < 3; x++) {
                                        const int i = ca[x], j = cb[x];
                                        auto kin = [&](int ii) {
                                            return -0.5 * ((j >= 2 ? j * (j - 1) * S[x][ii][j - 2] : 0.0) -
                                                           2 * b * (2 * j + 1) * S[x][ii][j] +
                                                           4 * b * b * S[x][ii][j + 2]);
                                        };
                                        if (deriv && x == c) {
                                            s[x] = -2 * a * S[x][i + 1][j] + (i > 0 ? i * S[x][i - 1][j] : 0.0);
                                            tk[x] = -2 * a * kin(i + 1) + (i > 0 ? i * kin(i - 1) : 0.0);
                                        } else {
                                            s[x] = S[x][i][j];
                                            tk[x] = kin(i);
                                        }
                                    }
                                    const double v = (kind == 0 || kind == 3)
                                                         ? s[0] * s[1] * s[2]
                                                         : tk[0] * s[1] * s[2] + s[0] * tk[1] * s[2] + s[0] * s[1] * tk[2];
                                    fx_add(fx, c * nn + (size_t)(i0 + fa) * nao + j0 + fb, cc * v);
                                }
                            }
                    } else {
                        const int nroots = (lae + lb) / 2 + 1;
                        for (int C = c_first; C < c_last; C++) {
                            const double* Rc = coords + 3 * C;
                            double PC2 = 0.0;
                            for (int x = 0; x < 3; x++) PC2 += (P[x] - Rc[x]) * (P[x] - Rc[x]);
                            double r[CCSD_NROOTS_MAX], w[CCSD_NROOTS_MAX];
                            ccsd::rys_roots_dev(nroots, p * PC2, r, w);
                            const double pref = -charges[C] * 2 * 3.14159265358979323846 / p * cc;
                            for (int ir = 0; ir < nroots; ir++) {
                                double I[3][L1 + 1][L1 + 1];
                                for (int x = 0; x < 3; x++)
                                    nuc1d(I[x], lae, lb, P[x] - A[x], P[x] - Rc[x], B[x] - A[x], r[ir], p,
                                          x == 2 ? w[ir] * pref : 1.0);
                                for (int fa = 0; fa < nfa; fa++)
                                    for (int fb = 0; fb < nfb; fb++) {
                                        int ca[3], cb[3];
                                        comp1(la, fa, ca);
                                        comp1(lb, fb, cb);
                                        const size_t e = (size_t)(i0 + fa) * nao + j0 + fb;
                                        if (!deriv) {
                                            fx_add(fx, e, I[0][ca[0]][cb[0]] * I[1][ca[1]][cb[1]] * I[2][ca[2]][cb[2]]);
                                        } else {
                                            for (int c = 0; c < 3; c++) {
                                                double f[3];
                                                for (int x = 0; x < 3; x++) {
                                                    const int i = ca[x], j = cb[x];
                                                    f[x] = x == c ? -2 * a * I[x][i + 1][j] + (i > 0 ? i * I[x][i - 1][j] : 0.0)
                                                                  : I[x][i][j];
                                                }
                                                fx_add(fx, (size_t)(C * 3 + c) * nn + e, f[0] * f[1] * f[2]);
                                            }
                                        }
                                    }
                            }
                        }
                    }
                }
        }
}

// out[e] = h 2^-20 + l 2^-60 of the fixed-point words (NaN everywhere after an overflow)
__global__ void int1e_convert_kernel(const unsigned long long* __restrict__ fx, long long n, double* __restrict__ out)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const long long h = (long long)fx[1 + 2 * e], l = (long long)fx[2 + 2 * e];
        out[e] = fx[0] ? __builtin_nan("") : (double)h * 9.5367431640625e-07 + (double)l * 8.673617379884035e-19;
    }
}

}  // namespace

// out (n elements) from the fixed-point accumulator fx (1 + 2 n words)
hipError_t ccsd_launch_int1e_convert(const unsigned long long* fx, long long n, double* out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536LL * 8) nb = 65536LL * 8;
    hipLaunchKernelGGL(int1e_convert_kernel, dim3((unsigned)nb), dim3(256), 0, stream, fx, n, out);
    return hipGetLastError();
}

hipError_t ccsd_launch_int1e(int kind, unsigned long long* fx, const ShellInfo* shells, const double* env,
                             const double* coords, const double* charges, int natm, int nbas, int nao, int maxprim,
                             hipStream_t stream)
{
    const long long n = (long long)nbas * nbas * maxprim * ((kind == 2 || kind == 5) ? natm : 1);
    if (n <= 0) return hipSuccess;
    if ((n + 63) / 64 > 2147483647LL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(int1e_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, stream, kind, fx, shells, env,
                       coords, charges, natm, nbas, nao, maxprim);
    return hipGetLastError();
}
