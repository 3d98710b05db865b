// Fused int2e_ip1 x Gamma contraction, register-resident, one class per template.
//
//   de[atom(i)][x] += scale * sum_{i, j, k>=l} ip1_x(ij|kl) Gamma[i][j][k(k+1)/2+l]
//
// One thread evaluates one shell quartet (ish jsh | ksh lsh): it loops over the
// primitive quartets, evaluates the Rys roots, builds the three 1-D factors
// (VRR on centre i, TRR i->k, HRR i->j and k->l; the same recurrences as the
// LDS kernel) in registers with compile-time extents, and contracts the
// derivative integrals with Gamma on the fly -- no integral is stored. A block
// shares one bra pair, so one atom receives the block's three sums.
//
// General contractions: Gamma for all contraction columns of the quartet is
// pre-contracted with the bra coefficients of each bra primitive pair
// (Gamma_bra[ckl][f]) and then with the ket coefficients of each ket primitive
// pair, so the integrals of a primitive quartet are computed once for every
// contraction column. Template bounds assume nctr <= 2 for s shells and 1 for
// l >= 1 (the host routes any other quartet class to the LDS kernel).
#include "eri_common.h"
#include "rys_device.h"

namespace {

template <int L>
struct CartTab {
    // (lx, ly, lz) of the Cartesian components, PySCF order
    static constexpr int n = (L + 1) * (L + 2) / 2;
    int x[n], y[n], z[n];
    constexpr CartTab() : x(), y(), z()
    {
        int c = 0;
        for (int lx = L; lx >= 0; lx--)
            for (int ly = L - lx; ly >= 0; ly--) {
                x[c] = lx;
                y[c] = ly;
                z[c] = L - lx - ly;
                c++;
            }
    }
};

__device__ __forceinline__ double shell_norm_dev(int l)
{
    return l == 0 ? 0.28209479177387814347 : (l == 1 ? 0.48860251190291992159 : 1.0);
}

// Registers a class needs: Gamma_bra (MCKL x NF) + Gamma of one primitive quartet (NF) + one root's
// three 1-D factors. Larger classes go to the LDS kernel.
constexpr bool fused_reg_fits(int li, int lj, int lk, int ll)
{
    const int nfi = (li + 1) * (li + 2) / 2, nfj = (lj + 1) * (lj + 2) / 2;
    const int nfk = (lk + 1) * (lk + 2) / 2, nfl = (ll + 1) * (ll + 2) / 2;
    const int NF = nfi * nfj * nfk * nfl;
    const int mckl = (lk == 0 ? 2 : 1) * (ll == 0 ? 2 : 1);
    const int gs = (li + 2) * (lj + 1) * (lk + 1) * (ll + 1);
    return li <= 2 && lj <= 2 && lk <= 2 && ll <= 2 && mckl * NF + NF + 3 * gs <= 200;
}

template <int LI, int LJ, int LK, int LL>
struct Fused {
    static constexpr int D = 1;
    static constexpr int NI = LI + 1 + D;
    static constexpr int LIJ = LI + LJ + D;
    static constexpr int LKL = LK + LL;
    static constexpr int NROOTS = (LI + LJ + LK + LL + D) / 2 + 1;
    static constexpr int NFI = (LI + 1) * (LI + 2) / 2, NFJ = (LJ + 1) * (LJ + 2) / 2;
    static constexpr int NFK = (LK + 1) * (LK + 2) / 2, NFL = (LL + 1) * (LL + 2) / 2;
    static constexpr int NF = NFI * NFJ * NFK * NFL;
    static constexpr int MCIJ = (LI == 0 ? 2 : 1) * (LJ == 0 ? 2 : 1);
    static constexpr int MCKL = (LK == 0 ? 2 : 1) * (LL == 0 ? 2 : 1);
    static constexpr int GS = NI * (LJ + 1) * (LK + 1) * (LL + 1);
};

// One direction: G[a][b][c][d], a <= LI+1, b <= LJ, c <= LK, d <= LL.
template <int LI, int LJ, int LK, int LL>
__device__ __forceinline__ void factor1d(double (&G)[LI + 2][LJ + 1][LK + 1][LL + 1], double seed, double c00, double cp,
                                         double b00, double b10, double b01, double AB, double CD)
{
    constexpr int LIJ = LI + LJ + 1, LKL = LK + LL;
    double v[LIJ + 1][LKL + 1];
    v[0][0] = seed;
#pragma unroll
    for (int i = 0; i < LIJ; i++) v[i + 1][0] = c00 * v[i][0] + (i > 0 ? i * b10 * v[i - 1][0] : 0.0);
#pragma unroll
    for (int k = 0; k < LKL; k++)
#pragma unroll
        for (int i = 0; i <= LIJ; i++) {
            double s = cp * v[i][k];
            if (k > 0) s += k * b01 * v[i][k - 1];
            if (i > 0) s += i * b00 * v[i - 1][k];
            v[i][k + 1] = s;
        }
    // HRR i -> j: h[a][b][k], (a, b+1) = (a+1, b) - AB (a, b)
    double h[LIJ + 1][LJ + 1][LKL + 1];
#pragma unroll
    for (int k = 0; k <= LKL; k++) {
#pragma unroll
        for (int a = 0; a <= LIJ; a++) h[a][0][k] = v[a][k];
#pragma unroll
        for (int b = 0; b < LJ; b++)
#pragma unroll
            for (int a = 0; a < LIJ - b; a++) h[a][b + 1][k] = h[a + 1][b][k] - AB * h[a][b][k];
    }
    // HRR k -> l
#pragma unroll
    for (int a = 0; a <= LI + 1; a++)
#pragma unroll
        for (int b = 0; b <= LJ; b++) {
            double t[LKL + 1][LL + 1];
#pragma unroll
            for (int c = 0; c <= LKL; c++) t[c][0] = h[a][b][c];
#pragma unroll
            for (int e = 0; e < LL; e++)
#pragma unroll
                for (int c = 0; c < LKL - e; c++) t[c][e + 1] = t[c + 1][e] - CD * t[c][e];
#pragma unroll
            for (int c = 0; c <= LK; c++)
#pragma unroll
                for (int e = 0; e <= LL; e++) G[a][b][c][e] = t[c][e];
        }
}

template <int LI, int LJ, int LK, int LL>
__global__ void __launch_bounds__(256) fused_ip1_reg_kernel(const EriClassArgs a)
{
    if constexpr (!fused_reg_fits(LI, LJ, LK, LL)) {
        return;
    } else {
    using F = Fused<LI, LJ, LK, LL>;
    constexpr CartTab<LI> ci_t{};
    constexpr CartTab<LJ> cj_t{};
    constexpr CartTab<LK> ck_t{};
    constexpr CartTab<LL> cl_t{};
    __shared__ double red[3][4];

    const int bq = blockIdx.y;
    const int kq = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = kq < a.nket;
    const int2 bp = a.bra[bq];
    const ShellInfo si = a.shells[bp.x], sj = a.shells[bp.y];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (active) {
        const int2 kp = a.ket[kq];
        const ShellInfo sk = a.shells[kp.x], sl = a.shells[kp.y];
        const double* RA = a.coords + 3 * si.atom;
        const double* RB = a.coords + 3 * sj.atom;
        const double* RC = a.coords + 3 * sk.atom;
        const double* RD = a.coords + 3 * sl.atom;
        const double AB[3] = {RB[0] - RA[0], RB[1] - RA[1], RB[2] - RA[2]};
        const double CD[3] = {RD[0] - RC[0], RD[1] - RC[1], RD[2] - RC[2]};
        const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        const double rcd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
        const double norm = 34.98683665524972497 * shell_norm_dev(LI) * shell_norm_dev(LJ) * shell_norm_dev(LK) *
                            shell_norm_dev(LL);
        const int ncij = si.nctr * sj.nctr, nckl = sk.nctr * sl.nctr;
        const bool same_kl = kp.x == kp.y;
        const double* ei = a.env + si.pexp;
        const double* ej = a.env + sj.pexp;
        const double* ek = a.env + sk.pexp;
        const double* el = a.env + sl.pexp;
        const double* cfi = a.env + si.pcoef;
        const double* cfj = a.env + sj.pcoef;
        const double* cfk = a.env + sk.pcoef;
        const double* cfl = a.env + sl.pcoef;

        // Gamma of every contraction column: gam[cij*MCKL + ckl][f] would need too many
        // registers, so only the bra-contracted form lives in registers.
        auto gamma_at = [&](int cij, int ckl, int f) -> double {
            const int c_i = cij / sj.nctr, c_j = cij - c_i * sj.nctr;
            const int c_k = ckl / sl.nctr, c_l = ckl - c_k * sl.nctr;
            const int fi = f % F::NFI, r1 = f / F::NFI;
            const int fj = r1 % F::NFJ, r2 = r1 / F::NFJ;
            const int fk = r2 % F::NFK, fl = r2 / F::NFK;
            const int i = si.ao_cart + c_i * F::NFI + fi;
            const int j = sj.ao_cart + c_j * F::NFJ + fj;
            const int k = sk.ao_cart + c_k * F::NFK + fk;
            const int l = sl.ao_cart + c_l * F::NFL + fl;
            if (same_kl && l > k) return 0.0;
            return a.gamma[((long long)i * a.nao + j) * a.npair + (long long)k * (k + 1) / 2 + l];
        };

        for (int ip = 0; ip < si.nprim; ip++)
            for (int jp = 0; jp < sj.nprim; jp++) {
                const double ai = ei[ip], aj = ej[jp], aij = ai + aj;
                const double Kab = exp(-ai * aj / aij * rab2);
                const double P[3] = {RA[0] + AB[0] * (aj / aij), RA[1] + AB[1] * (aj / aij), RA[2] + AB[2] * (aj / aij)};
                // Gamma contracted over the bra contraction columns of this bra primitive pair
                double gb[F::MCKL][F::NF];
#pragma unroll
                for (int c = 0; c < F::MCKL; c++)
#pragma unroll
                    for (int f = 0; f < F::NF; f++) gb[c][f] = 0.0;
                for (int cij = 0; cij < ncij; cij++) {
                    const int c_i = cij / sj.nctr, c_j = cij - c_i * sj.nctr;
                    const double cc = cfi[c_i * si.nprim + ip] * cfj[c_j * sj.nprim + jp];
#pragma unroll
                    for (int c = 0; c < F::MCKL; c++) {
                        if (c < nckl) {
#pragma unroll
                            for (int f = 0; f < F::NF; f++) gb[c][f] += cc * gamma_at(cij, c, f);
                        }
                    }
                }
                for (int kp2 = 0; kp2 < sk.nprim; kp2++)
                    for (int lp = 0; lp < sl.nprim; lp++) {
                        const double ak = ek[kp2], al = el[lp], akl = ak + al;
                        const double Kcd = exp(-ak * al / akl * rcd2);
                        const double Q[3] = {RC[0] + CD[0] * (al / akl), RC[1] + CD[1] * (al / akl),
                                             RC[2] + CD[2] * (al / akl)};
                        const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                        const double rr = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
                        const double T = aij * akl / (aij + akl) * rr;
                        // Gamma for this primitive quartet
                        double ge[F::NF];
#pragma unroll
                        for (int f = 0; f < F::NF; f++) ge[f] = 0.0;
#pragma unroll
                        for (int c = 0; c < F::MCKL; c++) {
                            if (c < nckl) {
                                const int c_k = c / sl.nctr, c_l = c - c_k * sl.nctr;
                                const double cc = cfk[c_k * sk.nprim + kp2] * cfl[c_l * sl.nprim + lp];
#pragma unroll
                                for (int f = 0; f < F::NF; f++) ge[f] += cc * gb[c][f];
                            }
                        }
                        double rt[F::NROOTS], wt[F::NROOTS];
                        ccsd::rys_roots_dev(F::NROOTS, T, rt, wt);
                        const double pref = norm * Kab * Kcd / (aij * akl * sqrt(aij + akl));
                        const double ai2 = 2.0 * ai;
#pragma unroll
                        for (int r = 0; r < F::NROOTS; r++) {
                            const double a2 = rt[r] / (aij + akl);
                            const double rt_aij = a2 * akl, rt_akl = a2 * aij;
                            const double b00 = 0.5 * a2;
                            const double b10 = 0.5 / aij * (1.0 - rt_aij);
                            const double b01 = 0.5 / akl * (1.0 - rt_akl);
                            double Gx[LI + 2][LJ + 1][LK + 1][LL + 1];
                            double Gy[LI + 2][LJ + 1][LK + 1][LL + 1];
                            double Gz[LI + 2][LJ + 1][LK + 1][LL + 1];
                            factor1d<LI, LJ, LK, LL>(Gx, 1.0, (P[0] - RA[0]) - rt_aij * PQ[0],
                                                     (Q[0] - RC[0]) + rt_akl * PQ[0], b00, b10, b01, AB[0], CD[0]);
                            factor1d<LI, LJ, LK, LL>(Gy, 1.0, (P[1] - RA[1]) - rt_aij * PQ[1],
                                                     (Q[1] - RC[1]) + rt_akl * PQ[1], b00, b10, b01, AB[1], CD[1]);
                            factor1d<LI, LJ, LK, LL>(Gz, wt[r] * pref, (P[2] - RA[2]) - rt_aij * PQ[2],
                                                     (Q[2] - RC[2]) + rt_akl * PQ[2], b00, b10, b01, AB[2], CD[2]);
#pragma unroll
                            for (int fl = 0; fl < F::NFL; fl++)
#pragma unroll
                                for (int fk = 0; fk < F::NFK; fk++)
#pragma unroll
                                    for (int fj = 0; fj < F::NFJ; fj++)
#pragma unroll
                                        for (int fi = 0; fi < F::NFI; fi++) {
                                            const int f = ((fl * F::NFK + fk) * F::NFJ + fj) * F::NFI + fi;
                                            const int ix = ci_t.x[fi], iy = ci_t.y[fi], iz = ci_t.z[fi];
                                            const int jx = cj_t.x[fj], jy = cj_t.y[fj], jz = cj_t.z[fj];
                                            const int kx = ck_t.x[fk], ky = ck_t.y[fk], kz = ck_t.z[fk];
                                            const int lx = cl_t.x[fl], ly = cl_t.y[fl], lz = cl_t.z[fl];
                                            const double gx = Gx[ix][jx][kx][lx];
                                            const double gy = Gy[iy][jy][ky][ly];
                                            const double gz = Gz[iz][jz][kz][lz];
                                            double dx = -ai2 * Gx[ix + 1][jx][kx][lx];
                                            double dy = -ai2 * Gy[iy + 1][jy][ky][ly];
                                            double dz = -ai2 * Gz[iz + 1][jz][kz][lz];
                                            if (ix > 0) dx += ix * Gx[ix > 0 ? ix - 1 : 0][jx][kx][lx];
                                            if (iy > 0) dy += iy * Gy[iy > 0 ? iy - 1 : 0][jy][ky][ly];
                                            if (iz > 0) dz += iz * Gz[iz > 0 ? iz - 1 : 0][jz][kz][lz];
                                            const double g = ge[f];
                                            sx += g * dx * gy * gz;
                                            sy += g * gx * dy * gz;
                                            sz += g * gx * gy * dz;
                                        }
                        }
                    }
            }
    }
    // block reduction (the bra pair, hence the atom of i, is shared by the block)
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_down(sx, off, 64);
        sy += __shfl_down(sy, off, 64);
        sz += __shfl_down(sz, off, 64);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[0][wave] = sx;
        red[1][wave] = sy;
        red[2][wave] = sz;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int nw = (blockDim.x + 63) >> 6;
        double s = 0.0;
        for (int w = 0; w < nw; w++) s += red[threadIdx.x][w];
        atomicAdd(a.de + 3 * si.atom + threadIdx.x, a.gamma_scale * s);
    }
    }
}

}  // namespace

// True if the class has a register kernel that accepts ket pairs with up to max_nckl
// contraction-column products.
bool ccsd_fused_reg_supported(int li, int lj, int lk, int ll, int max_nckl)
{
    return fused_reg_fits(li, lj, lk, ll) && max_nckl <= (lk == 0 ? 2 : 1) * (ll == 0 ? 2 : 1);
}

#define CCSD_FUSED_CASE(I, J, K, L)                                                                             \
    case (I * 27 + J * 9 + K * 3 + L):                                                                          \
        hipLaunchKernelGGL((fused_ip1_reg_kernel<I, J, K, L>), grid, dim3(256), 0, stream, a);                 \
        break;
#define CCSD_FUSED_L(I, J, K) CCSD_FUSED_CASE(I, J, K, 0) CCSD_FUSED_CASE(I, J, K, 1) CCSD_FUSED_CASE(I, J, K, 2)
#define CCSD_FUSED_K(I, J) CCSD_FUSED_L(I, J, 0) CCSD_FUSED_L(I, J, 1) CCSD_FUSED_L(I, J, 2)
#define CCSD_FUSED_J(I) CCSD_FUSED_K(I, 0) CCSD_FUSED_K(I, 1) CCSD_FUSED_K(I, 2)

hipError_t ccsd_launch_fused_reg(const EriClassArgs& a, hipStream_t stream)
{
    if (a.nbra <= 0 || a.nket <= 0) return hipSuccess;
    dim3 grid((a.nket + 255) / 256, a.nbra);
    switch (a.li * 27 + a.lj * 9 + a.lk * 3 + a.ll) {
        CCSD_FUSED_J(0)
        CCSD_FUSED_J(1)
        CCSD_FUSED_J(2)
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
