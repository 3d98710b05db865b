// Fused int2e_ip1 x Gamma contraction, register-resident, one class per template.
//
//   de[atom(i)][x] += scale * sum_{i, j, k>=l} ip1_x(ij|kl) Gamma[i][j][k(k+1)/2+l]
//
// One lane evaluates one shell quartet (ish jsh | ksh lsh): it loops over the
// primitive quartets, evaluates the Rys roots, builds the three 1-D factors
// (VRR on centre i, TRR i->k, HRR i->j and k->l; the same recurrences as the
// LDS kernel) in registers with compile-time extents, and contracts the
// derivative integrals with Gamma on the fly -- no integral is stored.
//
// Register budget: a lane holds Gamma for its quartet (all contraction
// columns) plus one root's three 1-D factors. Classes whose Gamma does not fit
// are split over P <= 8 waves ("parts"): part q contracts the Cartesian
// functions [q NF / P, (q+1) NF / P) and recomputes the roots and 1-D factors,
// which cost far less than the contraction; only the 1-D factor entries the
// part reads stay live. The part index is uniform over a wave, so the P
// specialised bodies never diverge.
//
// General contractions (GC = 1): Gamma of every contraction-column combination
// is loaded once per quartet and combined with the primitive coefficients of
// each primitive quartet, so each primitive quartet's integrals are evaluated
// once for all columns. GC = 0 launches hold only quartets whose shells are
// all single-column (the host splits the classes), and the coefficients then
// fold into the prefactor.
//
// A wave takes one (work chunk, part): one bra pair and up to 64 ket pairs, so
// one atom receives the wave's three sums.
#pragma once
#include <utility>

#include "eri_common.h"
#include "rys_device.h"

namespace ccsd_reg {

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

constexpr int kBudget = 110;   // doubles a lane may hold for Gamma and the 1-D factors
constexpr int kMaxParts = 8;   // beyond this the class stays on the LDS kernel

constexpr int ncart_c(int l) { return (l + 1) * (l + 2) / 2; }
// contraction columns a shell may bring into a register class: generally contracted
// s shells (two columns), single-column shells otherwise
constexpr int mcol(int l) { return l == 0 ? 2 : 1; }
constexpr bool has_s(int li, int lj, int lk, int ll) { return li == 0 || lj == 0 || lk == 0 || ll == 0; }

// exponent of direction d of Cartesian component f of angular momentum l <= 2 (PySCF order)
constexpr int kCartExp[3][6][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};
constexpr int cart_exp(int l, int f, int d) { return kCartExp[l][f][d]; }

// 1-D factor entries a part reads (the rest are dead and never held): per direction
// the (a, b, c, d) of its functions, with a + 1 and a - 1 for the derivative.
constexpr int g_live(int li, int lj, int lk, int ll, int f0, int f1)
{
    const int nfi = ncart_c(li), nfj = ncart_c(lj), nfk = ncart_c(lk);
    const int sa = li + 2, sb = lj + 1, sc = lk + 1, sd = ll + 1;
    int total = 0;
    for (int d = 0; d < 3; d++) {
        bool used[4 * 3 * 3 * 3] = {};
        for (int f = f0; f < f1; f++) {
            const int fi = f % nfi, r1 = f / nfi;
            const int fj = r1 % nfj, r2 = r1 / nfj;
            const int fk = r2 % nfk, fl = r2 / nfk;
            const int ea = cart_exp(li, fi, d), eb = cart_exp(lj, fj, d);
            const int ec = cart_exp(lk, fk, d), ed = cart_exp(ll, fl, d);
            const int base = ((eb * sc) + ec) * sd + ed;
            used[(ea + 1) * sb * sc * sd + base] = true;
            used[ea * sb * sc * sd + base] = true;
            if (ea > 0) used[(ea - 1) * sb * sc * sd + base] = true;
        }
        for (int u = 0; u < sa * sb * sc * sd; u++) total += used[u] ? 1 : 0;
    }
    return total;
}

// doubles a lane holds for part count p: Gamma of the part (with its contraction
// columns) and the live 1-D factor entries, worst part
constexpr int need(int li, int lj, int lk, int ll, int gc, int p)
{
    const int nf = ncart_c(li) * ncart_c(lj) * ncart_c(lk) * ncart_c(ll);
    const int m = gc ? mcol(li) * mcol(lj) * mcol(lk) * mcol(ll) : 1;
    int worst = 0;
    for (int q = 0; q < p; q++) {
        const int f0 = nf * q / p, f1 = nf * (q + 1) / p, nfp = f1 - f0;
        const int v = (m > 1 ? m * nfp + nfp : nfp) + g_live(li, lj, lk, ll, f0, f1);
        worst = v > worst ? v : worst;
    }
    return worst;
}

// Smallest part count that fits `budget` doubles (0: class not supported). The budget
// is an argument, not a per-file macro: this inline function must mean the same in
// every object it is compiled into.
constexpr int min_parts(int li, int lj, int lk, int ll, int gc, int budget = kBudget)
{
    if (li > 2 || lj > 2 || lk > 2 || ll > 2) return 0;
    const int nf = ncart_c(li) * ncart_c(lj) * ncart_c(lk) * ncart_c(ll);
    for (int p = 1; p <= kMaxParts && p <= nf; p *= 2)
        if (need(li, lj, lk, ll, gc, p) <= budget) return p;
    return 0;
}

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

// Bra primitive-pair table of a wave (identical for its 64 lanes), kBraFields doubles
// per pair: a_ij, P (3), norm * K_ab / a_ij, 2 a_i, 0.5 / a_ij, then the contraction-
// column coefficient products c_i c_j (up to 4).
constexpr int kBraFields = 11;

// Primitive pairs with a_1 a_2 / (a_1 + a_2) |R_12|^2 above this (Gaussian-product
// factor below e^-50 = 2e-22) contribute nothing at double precision and are skipped.
constexpr double kPrimExpCut = 50.0;

// Fills the wave's table with the bra primitive pairs that pass the cut, compacted
// (ballot prefix sums); returns their number (uniform over the wave).
template <int MIJ>
__device__ __forceinline__ int fill_bra(const EriClassArgs& a, const ShellInfo& si, const ShellInfo& sj,
                                        const double* RA, const double* AB, double norm, int lane, double* tab)
{
    const int nij = si.nprim * sj.nprim;
    const double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const CCSD_AS1 double* env = gld(a.env);
    const CCSD_AS1 double* ei = env + si.pexp;
    const CCSD_AS1 double* ej = env + sj.pexp;
    const CCSD_AS1 double* cfi = env + si.pcoef;
    const CCSD_AS1 double* cfj = env + sj.pcoef;
    const int ncij = si.nctr * sj.nctr;
    int count = 0;
    for (int p0 = 0; p0 < nij; p0 += 64) {
        const int p = p0 + lane;
        int ip = 0, jp = 0;
        double ai = 0.0, aj = 0.0, aij = 1.0, inv = 1.0, mu = 0.0;
        if (p < nij) {
            ip = p / sj.nprim;
            jp = p - ip * sj.nprim;
            ai = ei[ip];
            aj = ej[jp];
            aij = ai + aj;
            inv = 1.0 / aij;
            mu = ai * aj * inv * rab2;
        }
        const bool keep = p < nij && mu <= kPrimExpCut;
        const unsigned long long mask = __ballot(keep);
        const int slot = count + __popcll(mask & ((1ull << lane) - 1ull));
        count += __popcll(mask);
        if (!keep) continue;
        double* e = tab + slot * kBraFields;
        e[0] = aij;
        e[1] = RA[0] + AB[0] * (aj * inv);
        e[2] = RA[1] + AB[1] * (aj * inv);
        e[3] = RA[2] + AB[2] * (aj * inv);
        e[4] = norm * exp(-mu) * inv;
        e[5] = 2.0 * ai;
        e[6] = 0.5 * inv;
#pragma unroll
        for (int cij = 0; cij < MIJ; cij++) {
            const int c_i = cij / sj.nctr, c_j = cij - c_i * sj.nctr;
            e[7 + cij] = cij < ncij ? cfi[c_i * si.nprim + ip] * cfj[c_j * sj.nprim + jp] : 0.0;
        }
    }
    return count;
}

// Part Q of P of the quartet handled by this lane; adds to the lane's three sums.
template <int LI, int LJ, int LK, int LL, int GC, int P, int Q>
__device__ __forceinline__ void quartet_part(const EriClassArgs& a, const int4 ch, int lane, const double* tab,
                                             int nij, double& sx, double& sy, double& sz)
{
    constexpr int NFI = ncart_c(LI), NFJ = ncart_c(LJ), NFK = ncart_c(LK), NFL = ncart_c(LL);
    constexpr int NF = NFI * NFJ * NFK * NFL;
    constexpr int F0 = NF * Q / P, F1 = NF * (Q + 1) / P, NFP = F1 - F0;
    constexpr int NROOTS = (LI + LJ + LK + LL + 1) / 2 + 1;
    constexpr int MIJ = GC ? mcol(LI) * mcol(LJ) : 1;
    constexpr int MKL = GC ? mcol(LK) * mcol(LL) : 1;
    constexpr CartTab<LI> ci_t{};
    constexpr CartTab<LJ> cj_t{};
    constexpr CartTab<LK> ck_t{};
    constexpr CartTab<LL> cl_t{};

    const int kq = ch.y + lane;
    if (kq >= ch.z) return;
    const int2 bp = gld(a.bra)[ch.x];
    const int2 kp = gld(a.ket)[kq];
    const CCSD_AS1 ShellInfo* shells = gld(a.shells);
    const ShellInfo si = shells[bp.x], sj = shells[bp.y];
    const ShellInfo sk = shells[kp.x], sl = shells[kp.y];
    double RA[3], RB[3], RC[3], RD[3];
    load3(a.coords + 3 * si.atom, RA);
    load3(a.coords + 3 * sj.atom, RB);
    load3(a.coords + 3 * sk.atom, RC);
    load3(a.coords + 3 * sl.atom, RD);
    const double AB[3] = {RB[0] - RA[0], RB[1] - RA[1], RB[2] - RA[2]};
    const double CD[3] = {RD[0] - RC[0], RD[1] - RC[1], RD[2] - RC[2]};
    const int ncij = si.nctr * sj.nctr, nckl = sk.nctr * sl.nctr;
    const bool same_kl = kp.x == kp.y;

    // Gamma of this part for every contraction-column combination, loaded once
    const GammaSrc gs = gamma_src(a, si.mol);
    double gam[MIJ][MKL][NFP];
#pragma unroll
    for (int cij = 0; cij < MIJ; cij++)
#pragma unroll
        for (int ckl = 0; ckl < MKL; ckl++)
#pragma unroll
            for (int t = 0; t < NFP; t++) {
                const int f = F0 + t;
                const int fi = f % NFI, r1 = f / NFI;
                const int fj = r1 % NFJ, r2 = r1 / NFJ;
                const int fk = r2 % NFK, fl = r2 / NFK;
                double g = 0.0;
                if (cij < ncij && ckl < nckl) {
                    const int c_i = cij / sj.nctr, c_j = cij - c_i * sj.nctr;
                    const int c_k = ckl / sl.nctr, c_l = ckl - c_k * sl.nctr;
                    const int i = si.ao_cart + c_i * NFI + fi;
                    const int j = sj.ao_cart + c_j * NFJ + fj;
                    const int k = sk.ao_cart + c_k * NFK + fk;
                    const int l = sl.ao_cart + c_l * NFL + fl;
                    if (!(same_kl && l > k)) g = gamma_at(gs, i, j, k, l);
                }
                gam[cij][ckl][t] = g;
            }

    // ket primitive pairs past the cut, precomputed per ket pair (no divergent skips);
    // a chunk with a primitive split (w = s << 8 | S) takes pairs s, s + S, ...
    const CCSD_AS1 int* pp_off = gld(a.ket_pp_off);
    const int t1 = pp_off[kq + 1];
    const int split = ch.w & 255, sstep = split > 1 ? split : 1;
    for (int t = pp_off[kq] + (split > 1 ? (ch.w >> 8) : 0); t < t1; t += sstep) {
            const CCSD_AS1 double* kpd = gld(a.ket_pp) + (size_t)t * 10;
            const double akl = kpd[0];
            const double Qc[3] = {kpd[1], kpd[2], kpd[3]};
            const double KcdN = kpd[4];
            const double hkl = kpd[5];
            const double QC[3] = {Qc[0] - RC[0], Qc[1] - RC[1], Qc[2] - RC[2]};
            double ck[MKL];
#pragma unroll
            for (int ckl = 0; ckl < MKL; ckl++) ck[ckl] = kpd[6 + ckl];
            for (int p = 0; p < nij; p++) {
                // bra primitive pair: broadcast reads of the wave's table
                const double* e = tab + p * kBraFields;
                const double aij = e[0];
                const double Pc[3] = {e[1], e[2], e[3]};
                const double ai2 = e[5], hij = e[6];
                const double PQ[3] = {Pc[0] - Qc[0], Pc[1] - Qc[1], Pc[2] - Qc[2]};
                const double PA[3] = {Pc[0] - RA[0], Pc[1] - RA[1], Pc[2] - RA[2]};
                const double rr = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
                const double inv = 1.0 / (aij + akl);
                const double T = aij * akl * inv * rr;
                double pref = e[4] * KcdN * sqrt(inv);
                double ge[NFP];
                if constexpr (MIJ * MKL == 1) {
                    pref *= e[7] * ck[0];
#pragma unroll
                    for (int t = 0; t < NFP; t++) ge[t] = gam[0][0][t];
                } else {
#pragma unroll
                    for (int t = 0; t < NFP; t++) ge[t] = 0.0;
#pragma unroll
                    for (int cij = 0; cij < MIJ; cij++)
#pragma unroll
                        for (int ckl = 0; ckl < MKL; ckl++) {
                            const double cc = e[7 + cij] * ck[ckl];
#pragma unroll
                            for (int t = 0; t < NFP; t++) ge[t] += cc * gam[cij][ckl][t];
                        }
                }
                double rt[NROOTS], wt[NROOTS];
                ccsd::rys_roots_dev(NROOTS, T, rt, wt);
#pragma unroll
                for (int r = 0; r < NROOTS; r++) {
                    const double a2 = rt[r] * inv;
                    const double rt_aij = a2 * akl, rt_akl = a2 * aij;
                    const double b00 = 0.5 * a2;
                    const double b10 = hij * (1.0 - rt_aij);
                    const double b01 = hkl * (1.0 - rt_akl);
                    double Gx[LI + 2][LJ + 1][LK + 1][LL + 1];
                    double Gy[LI + 2][LJ + 1][LK + 1][LL + 1];
                    double Gz[LI + 2][LJ + 1][LK + 1][LL + 1];
                    factor1d<LI, LJ, LK, LL>(Gx, 1.0, PA[0] - rt_aij * PQ[0], QC[0] + rt_akl * PQ[0], b00, b10, b01,
                                             AB[0], CD[0]);
                    factor1d<LI, LJ, LK, LL>(Gy, 1.0, PA[1] - rt_aij * PQ[1], QC[1] + rt_akl * PQ[1], b00, b10, b01,
                                             AB[1], CD[1]);
                    factor1d<LI, LJ, LK, LL>(Gz, wt[r] * pref, PA[2] - rt_aij * PQ[2], QC[2] + rt_akl * PQ[2], b00,
                                             b10, b01, AB[2], CD[2]);
#pragma unroll
                    for (int t = 0; t < NFP; t++) {
                        const int f = F0 + t;
                        const int fi = f % NFI, r1 = f / NFI;
                        const int fj = r1 % NFJ, r2 = r1 / NFJ;
                        const int fk = r2 % NFK, fl = r2 / NFK;
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
                        const double g = ge[t];
                        sx += g * dx * gy * gz;
                        sy += g * gx * dy * gz;
                        sz += g * gx * gy * dz;
                    }
                }
            }
        }
}

template <int LI, int LJ, int LK, int LL, int GC, int P, int... Q>
__device__ __forceinline__ void run_parts(const EriClassArgs& a, const int4 ch, int lane, int part, const double* tab,
                                          int nij, double& sx, double& sy, double& sz, std::integer_sequence<int, Q...>)
{
    ((part == Q ? quartet_part<LI, LJ, LK, LL, GC, P, Q>(a, ch, lane, tab, nij, sx, sy, sz) : void()), ...);
}

// Dynamic LDS: one bra table of max_nij pairs per wave of the block.
template <int LI, int LJ, int LK, int LL, int GC, int P>
__global__ void __launch_bounds__(256) fused_reg_kernel(const EriClassArgs a)
{
    if constexpr (min_parts(LI, LJ, LK, LL, GC) == 0 || P < min_parts(LI, LJ, LK, LL, GC)) {
        return;
    } else {
        extern __shared__ double lds_tab[];
        constexpr int MIJ = GC ? mcol(LI) * mcol(LJ) : 1;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int gw = blockIdx.x * (blockDim.x >> 6) + wave;
        const int c = gw / P;
        const bool active = c < a.nchunks;   // every wave reaches the barrier below
        const int part = gw - c * P;
        int4 ch = make_int4(0, 0, 0, 0);
        if (active) ch = gld(a.chunks)[c];
        double* tab = lds_tab + (size_t)wave * a.max_nij * kBraFields;
        const int2 bp = gld(a.bra)[ch.x];
        const ShellInfo si = gld(a.shells)[bp.x], sj = gld(a.shells)[bp.y];
        int nij = 0;
        if (active) {
            double RA[3], RB[3];
            load3(a.coords + 3 * si.atom, RA);
            load3(a.coords + 3 * sj.atom, RB);
            const double AB[3] = {RB[0] - RA[0], RB[1] - RA[1], RB[2] - RA[2]};
            const double norm = 34.98683665524972497 * shell_norm_dev(LI) * shell_norm_dev(LJ) * shell_norm_dev(LK) *
                                shell_norm_dev(LL);
            nij = fill_bra<MIJ>(a, si, sj, RA, AB, norm, lane, tab);
        }
        __syncthreads();
        if (!active) return;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        run_parts<LI, LJ, LK, LL, GC, P>(a, ch, lane, part, tab, nij, sx, sy, sz,
                                         std::make_integer_sequence<int, P>{});
        // wave reduction (the chunk's bra pair, hence the atom of i, is shared by the wave)
        for (int off = 32; off > 0; off >>= 1) {
            sx += __shfl_down(sx, off, 64);
            sy += __shfl_down(sy, off, 64);
            sz += __shfl_down(sz, off, 64);
        }
        if (lane == 0) defx_add(a.defx, si.atom, a.gamma_scale * sx, a.gamma_scale * sy, a.gamma_scale * sz);
    }
}

template <int LI, int LJ, int LK, int LL, int GC>
hipError_t launch(const EriClassArgs& a, int wpb, hipStream_t stream)
{
    constexpr int G = (GC && has_s(LI, LJ, LK, LL)) ? 1 : 0;
    constexpr int P0 = min_parts(LI, LJ, LK, LL, G);
    constexpr int P = P0 ? P0 : 1;
    if (P0 == 0) return hipErrorInvalidValue;
    const long long waves = (long long)a.nchunks * P;
    const long long blocks = (waves + wpb - 1) / wpb;
    if (blocks > 2147483647LL) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * (size_t)wpb * a.max_nij * kBraFields;
    if (lds > 65536) return hipErrorInvalidValue;
    hipLaunchKernelGGL((fused_reg_kernel<LI, LJ, LK, LL, G, P>), dim3((unsigned)blocks), dim3(64 * wpb), lds, stream,
                       a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Store mode: the same per-lane quartet evaluation writing the integrals to the s2kl
// tensor out[comp][i - i_base][j][k(k+1)/2 + l] (the int2e / int2e_ip1 drop-ins and
// the unfused gradient). A lane accumulates its quartet's outputs of one part over
// all primitive quartets and roots (NC = 3 derivative components or 1 integral per
// Cartesian function and contraction-column combination), then writes each once:
// no ket-primitive split, every quartet of the work list is written (zeros where
// no primitive pair passes the Gaussian-product cut).
constexpr int kStoreBudget = 128;
constexpr int kStoreMaxParts = 64;

// true when every part of a p-way split fits the store budget (stops at the first part
// that does not: the compile-time evaluation stays short for the large classes)
constexpr bool fits_store(int li, int lj, int lk, int ll, int gc, int p, int nc)
{
    const int nf = ncart_c(li) * ncart_c(lj) * ncart_c(lk) * ncart_c(ll);
    const int m = gc ? mcol(li) * mcol(lj) * mcol(lk) * mcol(ll) : 1;
    if (m * ((nf + p - 1) / p) * nc > kStoreBudget) return false;
    for (int q = 0; q < p; q++) {
        const int f0 = nf * q / p, f1 = nf * (q + 1) / p, nfp = f1 - f0;
        if (m * nfp * nc + g_live(li, lj, lk, ll, f0, f1) > kStoreBudget) return false;
    }
    return true;
}

constexpr int min_parts_store(int li, int lj, int lk, int ll, int gc, int nc)
{
    if (li > 2 || lj > 2 || lk > 2 || ll > 2) return 0;
    const int nf = ncart_c(li) * ncart_c(lj) * ncart_c(lk) * ncart_c(ll);
    for (int p = 1; p <= kStoreMaxParts && p <= nf; p *= 2)
        if (fits_store(li, lj, lk, ll, gc, p, nc)) return p;
    return 0;
}

template <int LI, int LJ, int LK, int LL, int GC, int P, int Q, int DERIV>
__device__ __forceinline__ void quartet_store(const EriClassArgs& a, const int4 ch, int lane, const double* tab,
                                              int nij)
{
    constexpr int NFI = ncart_c(LI), NFJ = ncart_c(LJ), NFK = ncart_c(LK), NFL = ncart_c(LL);
    constexpr int NF = NFI * NFJ * NFK * NFL;
    constexpr int F0 = NF * Q / P, F1 = NF * (Q + 1) / P, NFP = F1 - F0;
    constexpr int NC = DERIV ? 3 : 1;
    constexpr int NROOTS = (LI + LJ + LK + LL + DERIV) / 2 + 1;
    constexpr int MIJ = GC ? mcol(LI) * mcol(LJ) : 1;
    constexpr int MKL = GC ? mcol(LK) * mcol(LL) : 1;
    constexpr CartTab<LI> ci_t{};
    constexpr CartTab<LJ> cj_t{};
    constexpr CartTab<LK> ck_t{};
    constexpr CartTab<LL> cl_t{};

    const int kq = ch.y + lane;
    if (kq >= ch.z) return;
    const int2 bp = gld(a.bra)[ch.x];
    const int2 kp = gld(a.ket)[kq];
    const CCSD_AS1 ShellInfo* shells = gld(a.shells);
    const ShellInfo si = shells[bp.x], sj = shells[bp.y];
    const ShellInfo sk = shells[kp.x], sl = shells[kp.y];
    double RA[3], RB[3], RC[3], RD[3];
    load3(a.coords + 3 * si.atom, RA);
    load3(a.coords + 3 * sj.atom, RB);
    load3(a.coords + 3 * sk.atom, RC);
    load3(a.coords + 3 * sl.atom, RD);
    const double AB[3] = {RB[0] - RA[0], RB[1] - RA[1], RB[2] - RA[2]};
    const double CD[3] = {RD[0] - RC[0], RD[1] - RC[1], RD[2] - RC[2]};
    const int ncij = si.nctr * sj.nctr, nckl = sk.nctr * sl.nctr;

    double acc[MIJ][MKL][NFP][NC];
#pragma unroll
    for (int cij = 0; cij < MIJ; cij++)
#pragma unroll
        for (int ckl = 0; ckl < MKL; ckl++)
#pragma unroll
            for (int t = 0; t < NFP; t++)
#pragma unroll
                for (int c = 0; c < NC; c++) acc[cij][ckl][t][c] = 0.0;

    const CCSD_AS1 int* pp_off = gld(a.ket_pp_off);
    const int t1 = pp_off[kq + 1];
    for (int t = pp_off[kq]; t < t1; t++) {
        const CCSD_AS1 double* kpd = gld(a.ket_pp) + (size_t)t * 10;
        const double akl = kpd[0];
        const double Qc[3] = {kpd[1], kpd[2], kpd[3]};
        const double KcdN = kpd[4];
        const double hkl = kpd[5];
        const double QC[3] = {Qc[0] - RC[0], Qc[1] - RC[1], Qc[2] - RC[2]};
        double ck[MKL];
#pragma unroll
        for (int ckl = 0; ckl < MKL; ckl++) ck[ckl] = kpd[6 + ckl];
        for (int p = 0; p < nij; p++) {
            const double* e = tab + p * kBraFields;
            const double aij = e[0];
            const double Pc[3] = {e[1], e[2], e[3]};
            const double ai2 = e[5], hij = e[6];
            const double PQ[3] = {Pc[0] - Qc[0], Pc[1] - Qc[1], Pc[2] - Qc[2]};
            const double PA[3] = {Pc[0] - RA[0], Pc[1] - RA[1], Pc[2] - RA[2]};
            const double rr = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
            const double inv = 1.0 / (aij + akl);
            const double T = aij * akl * inv * rr;
            double pref = e[4] * KcdN * sqrt(inv);
            if constexpr (MIJ * MKL == 1) pref *= e[7] * ck[0];
            double cc[MIJ][MKL];
#pragma unroll
            for (int cij = 0; cij < MIJ; cij++)
#pragma unroll
                for (int ckl = 0; ckl < MKL; ckl++) cc[cij][ckl] = (MIJ * MKL == 1) ? 1.0 : e[7 + cij] * ck[ckl];
            double rt[NROOTS], wt[NROOTS];
            ccsd::rys_roots_dev(NROOTS, T, rt, wt);
#pragma unroll
            for (int r = 0; r < NROOTS; r++) {
                const double a2 = rt[r] * inv;
                const double rt_aij = a2 * akl, rt_akl = a2 * aij;
                const double b00 = 0.5 * a2;
                const double b10 = hij * (1.0 - rt_aij);
                const double b01 = hkl * (1.0 - rt_akl);
                double Gx[LI + 2][LJ + 1][LK + 1][LL + 1];
                double Gy[LI + 2][LJ + 1][LK + 1][LL + 1];
                double Gz[LI + 2][LJ + 1][LK + 1][LL + 1];
                factor1d<LI, LJ, LK, LL>(Gx, 1.0, PA[0] - rt_aij * PQ[0], QC[0] + rt_akl * PQ[0], b00, b10, b01,
                                         AB[0], CD[0]);
                factor1d<LI, LJ, LK, LL>(Gy, 1.0, PA[1] - rt_aij * PQ[1], QC[1] + rt_akl * PQ[1], b00, b10, b01,
                                         AB[1], CD[1]);
                factor1d<LI, LJ, LK, LL>(Gz, wt[r] * pref, PA[2] - rt_aij * PQ[2], QC[2] + rt_akl * PQ[2], b00,
                                         b10, b01, AB[2], CD[2]);
#pragma unroll
                for (int t = 0; t < NFP; t++) {
                    const int f = F0 + t;
                    const int fi = f % NFI, r1 = f / NFI;
                    const int fj = r1 % NFJ, r2 = r1 / NFJ;
                    const int fk = r2 % NFK, fl = r2 / NFK;
                    const int ix = ci_t.x[fi], iy = ci_t.y[fi], iz = ci_t.z[fi];
                    const int jx = cj_t.x[fj], jy = cj_t.y[fj], jz = cj_t.z[fj];
                    const int kx = ck_t.x[fk], ky = ck_t.y[fk], kz = ck_t.z[fk];
                    const int lx = cl_t.x[fl], ly = cl_t.y[fl], lz = cl_t.z[fl];
                    const double gx = Gx[ix][jx][kx][lx];
                    const double gy = Gy[iy][jy][ky][ly];
                    const double gz = Gz[iz][jz][kz][lz];
                    double v[NC];
                    if constexpr (DERIV) {
                        double dx = -ai2 * Gx[ix + 1][jx][kx][lx];
                        double dy = -ai2 * Gy[iy + 1][jy][ky][ly];
                        double dz = -ai2 * Gz[iz + 1][jz][kz][lz];
                        if (ix > 0) dx += ix * Gx[ix > 0 ? ix - 1 : 0][jx][kx][lx];
                        if (iy > 0) dy += iy * Gy[iy > 0 ? iy - 1 : 0][jy][ky][ly];
                        if (iz > 0) dz += iz * Gz[iz > 0 ? iz - 1 : 0][jz][kz][lz];
                        v[0] = dx * gy * gz;
                        v[NC > 1 ? 1 : 0] = gx * dy * gz;
                        v[NC > 2 ? 2 : 0] = gx * gy * dz;
                    } else {
                        v[0] = gx * gy * gz;
                    }
#pragma unroll
                    for (int cij = 0; cij < MIJ; cij++)
#pragma unroll
                        for (int ckl = 0; ckl < MKL; ckl++)
#pragma unroll
                            for (int c = 0; c < NC; c++) acc[cij][ckl][t][c] += cc[cij][ckl] * v[c];
                }
            }
        }
    }
    // write every (column combination, function) of the part; when ksh == lsh only the
    // k >= l half of the shell pair belongs to the packed layout
    const bool same_kl = kp.x == kp.y;
#pragma unroll
    for (int cij = 0; cij < MIJ; cij++)
#pragma unroll
        for (int ckl = 0; ckl < MKL; ckl++) {
            if (cij >= ncij || ckl >= nckl) continue;
            const int c_i = cij / sj.nctr, c_j = cij - c_i * sj.nctr;
            const int c_k = ckl / sl.nctr, c_l = ckl - c_k * sl.nctr;
#pragma unroll
            for (int t = 0; t < NFP; t++) {
                const int f = F0 + t;
                const int fi = f % NFI, r1 = f / NFI;
                const int fj = r1 % NFJ, r2 = r1 / NFJ;
                const int fk = r2 % NFK, fl = r2 / NFK;
                const int i = si.ao_cart + c_i * NFI + fi;
                const int j = sj.ao_cart + c_j * NFJ + fj;
                const int k = sk.ao_cart + c_k * NFK + fk;
                const int l = sl.ao_cart + c_l * NFL + fl;
                if (same_kl && l > k) continue;
                const long long idx =
                    ((long long)(i - a.i_base) * a.nao + j) * a.npair + (long long)k * (k + 1) / 2 + l;
#pragma unroll
                for (int c = 0; c < NC; c++) gst(a.out)[c * a.comp_stride + idx] = acc[cij][ckl][t][c];
            }
        }
}

template <int LI, int LJ, int LK, int LL, int GC, int P, int DERIV, int... Q>
__device__ __forceinline__ void run_store_parts(const EriClassArgs& a, const int4 ch, int lane, int part,
                                                const double* tab, int nij, std::integer_sequence<int, Q...>)
{
    ((part == Q ? quartet_store<LI, LJ, LK, LL, GC, P, Q, DERIV>(a, ch, lane, tab, nij) : void()), ...);
}

template <int LI, int LJ, int LK, int LL, int GC, int P, int DERIV>
__global__ void __launch_bounds__(256) store_reg_kernel(const EriClassArgs a)
{
    constexpr int NC = DERIV ? 3 : 1;
    constexpr int P0 = min_parts_store(LI, LJ, LK, LL, GC, NC);
    if constexpr (P0 == 0 || P < P0) {
        return;
    } else {
        extern __shared__ double lds_tab[];
        constexpr int MIJ = GC ? mcol(LI) * mcol(LJ) : 1;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int gw = blockIdx.x * (blockDim.x >> 6) + wave;
        const int c = gw / P;
        const bool active = c < a.nchunks;   // every wave reaches the barrier below
        const int part = gw - c * P;
        int4 ch = make_int4(0, 0, 0, 0);
        if (active) ch = gld(a.chunks)[c];
        double* tab = lds_tab + (size_t)wave * a.max_nij * kBraFields;
        const int2 bp = gld(a.bra)[ch.x];
        const ShellInfo si = gld(a.shells)[bp.x], sj = gld(a.shells)[bp.y];
        int nij = 0;
        if (active) {
            double RA[3], RB[3];
            load3(a.coords + 3 * si.atom, RA);
            load3(a.coords + 3 * sj.atom, RB);
            const double AB[3] = {RB[0] - RA[0], RB[1] - RA[1], RB[2] - RA[2]};
            const double norm = 34.98683665524972497 * shell_norm_dev(LI) * shell_norm_dev(LJ) * shell_norm_dev(LK) *
                                shell_norm_dev(LL);
            nij = fill_bra<MIJ>(a, si, sj, RA, AB, norm, lane, tab);
        }
        __syncthreads();
        if (!active) return;
        run_store_parts<LI, LJ, LK, LL, GC, P, DERIV>(a, ch, lane, part, tab, nij,
                                                      std::make_integer_sequence<int, P>{});
    }
}

template <int LI, int LJ, int LK, int LL, int GC, int DERIV>
hipError_t launch_store(const EriClassArgs& a, int wpb, hipStream_t stream)
{
    constexpr int G = (GC && has_s(LI, LJ, LK, LL)) ? 1 : 0;
    constexpr int P0 = min_parts_store(LI, LJ, LK, LL, G, DERIV ? 3 : 1);
    constexpr int P = P0 ? P0 : 1;
    if (P0 == 0) return hipErrorInvalidValue;
    const long long waves = (long long)a.nchunks * P;
    const long long blocks = (waves + wpb - 1) / wpb;
    if (blocks > 2147483647LL) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * (size_t)wpb * a.max_nij * kBraFields;
    if (lds > 65536) return hipErrorInvalidValue;
    hipLaunchKernelGGL((store_reg_kernel<LI, LJ, LK, LL, G, P, DERIV>), dim3((unsigned)blocks), dim3(64 * wpb), lds,
                       stream, a);
    return hipGetLastError();
}

}  // namespace ccsd_reg

// Launcher of the classes with angular momenta (LI, LJ, LK, *); each triple is its own
// object file (compiled in parallel).
#define CCSD_REG_CASE(I, J, K, L)                                                                            \
    case L:                                                                                                  \
        return gc ? ccsd_reg::launch<I, J, K, L, 1>(a, wpb, stream) : ccsd_reg::launch<I, J, K, L, 0>(a, wpb, stream);
#define CCSD_REG_TU(I, J, K, NAME)                                                                           \
    hipError_t NAME(const EriClassArgs& a, int gc, int wpb, hipStream_t stream)                              \
    {                                                                                                        \
        if (a.nchunks <= 0) return hipSuccess;                                                               \
        if (wpb < 1 || wpb > 4) return hipErrorInvalidValue;                                                 \
        switch (a.ll) {                                                                                      \
            CCSD_REG_CASE(I, J, K, 0)                                                                        \
            CCSD_REG_CASE(I, J, K, 1)                                                                        \
            CCSD_REG_CASE(I, J, K, 2)                                                                        \
            default:                                                                                         \
                return hipErrorInvalidValue;                                                                 \
        }                                                                                                    \
    }

// Store-mode launcher of the classes (LI, LJ, LK, *), int2e (deriv 0) or int2e_ip1 (deriv 1).
#define CCSD_STORE_CASE(I, J, K, L)                                                                          \
    case L:                                                                                                  \
        if (a.deriv)                                                                                         \
            return gc ? ccsd_reg::launch_store<I, J, K, L, 1, 1>(a, wpb, stream)                             \
                      : ccsd_reg::launch_store<I, J, K, L, 0, 1>(a, wpb, stream);                            \
        return gc ? ccsd_reg::launch_store<I, J, K, L, 1, 0>(a, wpb, stream)                                 \
                  : ccsd_reg::launch_store<I, J, K, L, 0, 0>(a, wpb, stream);
#define CCSD_STORE_TU(I, J, K, NAME)                                                                         \
    hipError_t NAME(const EriClassArgs& a, int gc, int wpb, hipStream_t stream)                              \
    {                                                                                                        \
        if (a.nchunks <= 0) return hipSuccess;                                                               \
        if (wpb < 1 || wpb > 4) return hipErrorInvalidValue;                                                 \
        switch (a.ll) {                                                                                      \
            CCSD_STORE_CASE(I, J, K, 0)                                                                      \
            CCSD_STORE_CASE(I, J, K, 1)                                                                      \
            CCSD_STORE_CASE(I, J, K, 2)                                                                      \
            default:                                                                                         \
                return hipErrorInvalidValue;                                                                 \
        }                                                                                                    \
    }
