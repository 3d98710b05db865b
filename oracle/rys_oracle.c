/*
 * CPU FP64 oracle for the Rys-quadrature two-electron integrals.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP kernels and
 * the timed CPU baseline. Nothing in the shipped library calls it.
 *
 * Restates the reference's algorithm:
 *   - Rys roots/weights from Chebyshev tables (ccsd_gpu/cuda/int2e_ip1_module.py:192-244),
 *     here with the project's own tables (ccsd_gpu_amd/csrc/rys_table.inc);
 *   - prefactors fac_kl = 2 pi^(5/2) prod SHELL_NORM c_k c_l K_cd and
 *     cicj = c_i c_j K_ab / (a_ij a_kl sqrt(a_ij + a_kl))  (int2e_ip1_module.py:371-431);
 *   - VRR on centre i, TRR i->k, HRR i->j and k->l (int2e_ip1_module.py:440-587,
 *     ccsd_gpu/cuda/int2e_module.py:306-405);
 *   - int2e accumulates Ix Iy Iz (int2e_module.py:409-430), int2e_ip1 the
 *     d/dr1 derivative  f_x = -2 a_i g[ix+1] + ix g[ix-1]  (int2e_ip1_module.py:612-629);
 *   - output (ncomp, nf_i, nao, nao_pair) with kl packed k(k+1)/2 + l, k >= l
 *     (int2e_ip1_module.py:636-663).
 * General contractions are handled by looping over contraction columns,
 * which is the same arithmetic as the reference's decontraction
 * (ccsd_gpu/cuda/int2e_ip1_gpu.py:43-79).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RYS_STORAGE static
#include "../ccsd_gpu_amd/csrc/rys_table.inc"

#define ATM_SLOTS 6
#define PTR_COORD 1
#define BAS_SLOTS 8
#define ATOM_OF 0
#define ANG_OF 1
#define NPRIM_OF 2
#define NCTR_OF 3
#define PTR_EXP 5
#define PTR_COEFF 6

#define LMAX 4
#define MAXD (2 * LMAX + 2) /* max combined index + 1 (li+lj+1 with deriv) */

static const double TWO_PI_2_5 = 34.98683665524972497; /* 2 pi^(5/2) */

/* Chebyshev series evaluation of the Rys table; roots ascending. */
void oracle_rys_roots(int n, double T, double *r, double *w)
{
    if (T >= RYS_TASYM(n)) {
        int o = n * (n - 1) / 2;
        double s = sqrt(0.78539816339744830962 / T);
        for (int i = 0; i < n; i++) {
            r[i] = RYS_LARGE_R[o + i] / T;
            w[i] = RYS_LARGE_W[o + i] * s;
        }
        return;
    }
    int it = (int)(T * RYS_INV_WIDTH);
    double u = (T - it * RYS_WIDTH) * (2.0 / RYS_WIDTH) - 1.0;
    const double *c = RYS_TABLE + RYS_OFFSET[n - 1] + (long)it * 2 * n * (RYS_DEGREE + 1);
    for (int s = 0; s < 2 * n; s++) {
        const double *cs = c + s * (RYS_DEGREE + 1);
        /* Clenshaw: sum_k cs[k] T_k(u) */
        double b1 = 0.0, b2 = 0.0;
        for (int k = RYS_DEGREE; k >= 1; k--) {
            double b0 = cs[k] + 2.0 * u * b1 - b2;
            b2 = b1;
            b1 = b0;
        }
        double v = cs[0] + u * b1 - b2;
        if (s < n) r[s] = v; else w[s - n] = v;
    }
}

static double shell_norm(int l)
{
    if (l == 0) return 0.28209479177387814347; /* 1/(2 sqrt(pi)) */
    if (l == 1) return 0.48860251190291992159; /* sqrt(3/(4 pi)) */
    return 1.0;
}

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static void cart_comp(int l, int f, int *lx, int *ly, int *lz)
{
    int n = 0;
    for (int x = l; x >= 0; x--)
        for (int y = l - x; y >= 0; y--) {
            if (n == f) { *lx = x; *ly = y; *lz = l - x - y; return; }
            n++;
        }
}

/* One 1-D factor for a primitive quartet and one root:
 * G[i][j][k][l] for i<=li+d, j<=lj, k<=lk, l<=ll. */
typedef struct {
    double v[MAXD][LMAX + 1][MAXD][LMAX + 1];
} g4_t;

static void build_1d(g4_t *G, int li, int lj, int lk, int ll, int d,
                     double seed, double PA, double QC, double PQ, double AB, double CD,
                     double rt, double aij, double akl)
{
    int lij = li + lj + d, lkl = lk + ll;
    double a2 = rt / (aij + akl);
    double rt_aij = a2 * akl, rt_akl = a2 * aij;
    double b00 = 0.5 * a2;
    double b10 = 0.5 / aij * (1.0 - rt_aij);
    double b01 = 0.5 / akl * (1.0 - rt_akl);
    double c00 = PA - rt_aij * PQ;
    double cp = QC + rt_akl * PQ;
    double g[2 * LMAX + 2][2 * LMAX + 1];
    /* VRR (bra, centre A) */
    g[0][0] = seed;
    if (lij > 0) g[1][0] = c00 * seed;
    for (int i = 1; i < lij; i++) g[i + 1][0] = c00 * g[i][0] + i * b10 * g[i - 1][0];
    /* TRR (ket, centre C) */
    for (int k = 0; k < lkl; k++)
        for (int i = 0; i <= lij; i++) {
            double s = cp * g[i][k];
            if (k > 0) s += k * b01 * g[i][k - 1];
            if (i > 0) s += i * b00 * g[i - 1][k];
            g[i][k + 1] = s;
        }
    /* HRR bra: (a, b+1) = (a+1, b) - (B-A) (a, b), AB = B - A */
    static __thread double h[2 * LMAX + 2][LMAX + 1][2 * LMAX + 1];
    for (int k = 0; k <= lkl; k++) {
        for (int a = 0; a <= lij; a++) h[a][0][k] = g[a][k];
        for (int b = 0; b < lj; b++)
            for (int a = 0; a <= lij - b - 1; a++)
                h[a][b + 1][k] = h[a + 1][b][k] - AB * h[a][b][k];
    }
    /* HRR ket: (c, e+1) = (c+1, e) - (D-C) (c, e), CD = D - C */
    for (int a = 0; a <= li + d; a++)
        for (int b = 0; b <= lj; b++) {
            double t[2 * LMAX + 1][LMAX + 1];
            for (int c = 0; c <= lkl; c++) t[c][0] = h[a][b][c];
            for (int e = 0; e < ll; e++)
                for (int c = 0; c <= lkl - e - 1; c++)
                    t[c][e + 1] = t[c + 1][e] - CD * t[c][e];
            for (int c = 0; c <= lk; c++)
                for (int e = 0; e <= ll; e++) G->v[a][b][c][e] = t[c][e];
        }
}

typedef struct {
    int l, nprim, nctr, atom;
    const double *exp, *coef; /* coef[ctr*nprim + p] */
    const double *R;
} shell_t;

static void get_shell(shell_t *s, int ish, const int *atm, const int *bas, const double *env)
{
    const int *b = bas + ish * BAS_SLOTS;
    s->l = b[ANG_OF];
    s->nprim = b[NPRIM_OF];
    s->nctr = b[NCTR_OF];
    s->atom = b[ATOM_OF];
    s->exp = env + b[PTR_EXP];
    s->coef = env + b[PTR_COEFF];
    s->R = env + atm[s->atom * ATM_SLOTS + PTR_COORD];
}

/* Cartesian (ij|kl) or d/dr_i (ij|kl) for one shell quartet, all contraction
 * columns. out[comp][ci][cj][ck][cl][fi][fj][fk][fl] (fi fastest is NOT used:
 * layout is ((((comp*nci+ci)*ncj+cj)*nck+ck)*ncl+cl)*nf + ((fl*nfk+fk)*nfj+fj)*nfi+fi). */
static void quartet(double *out, int d, const shell_t *I, const shell_t *J,
                    const shell_t *K, const shell_t *L)
{
    int li = I->l, lj = J->l, lk = K->l, ll = L->l;
    int nfi = ncart(li), nfj = ncart(lj), nfk = ncart(lk), nfl = ncart(ll);
    int nf = nfi * nfj * nfk * nfl;
    int ncomp = d ? 3 : 1;
    int nctr = I->nctr * J->nctr * K->nctr * L->nctr;
    int nroots = (li + lj + lk + ll + d) / 2 + 1;
    memset(out, 0, sizeof(double) * (size_t)ncomp * nctr * nf);
    double AB[3], CD[3];
    for (int x = 0; x < 3; x++) { AB[x] = J->R[x] - I->R[x]; CD[x] = L->R[x] - K->R[x]; }
    double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    double rcd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
    double norm = TWO_PI_2_5 * shell_norm(li) * shell_norm(lj) * shell_norm(lk) * shell_norm(ll);
    int comps[4][15][3];
    for (int f = 0; f < nfi; f++) cart_comp(li, f, &comps[0][f][0], &comps[0][f][1], &comps[0][f][2]);
    for (int f = 0; f < nfj; f++) cart_comp(lj, f, &comps[1][f][0], &comps[1][f][1], &comps[1][f][2]);
    for (int f = 0; f < nfk; f++) cart_comp(lk, f, &comps[2][f][0], &comps[2][f][1], &comps[2][f][2]);
    for (int f = 0; f < nfl; f++) cart_comp(ll, f, &comps[3][f][0], &comps[3][f][1], &comps[3][f][2]);
    double *prim = (double *)malloc(sizeof(double) * ncomp * nf);
    g4_t *G = (g4_t *)malloc(sizeof(g4_t) * 3);
    for (int kp = 0; kp < K->nprim; kp++)
    for (int lp = 0; lp < L->nprim; lp++) {
        double ak = K->exp[kp], al = L->exp[lp], akl = ak + al;
        double Kcd = exp(-ak * al / akl * rcd2);
        double Q[3];
        for (int x = 0; x < 3; x++) Q[x] = K->R[x] + CD[x] * (al / akl);
        for (int ip = 0; ip < I->nprim; ip++)
        for (int jp = 0; jp < J->nprim; jp++) {
            double ai = I->exp[ip], aj = J->exp[jp], aij = ai + aj;
            double Kab = exp(-ai * aj / aij * rab2);
            double P[3], PQ[3];
            for (int x = 0; x < 3; x++) { P[x] = I->R[x] + AB[x] * (aj / aij); PQ[x] = P[x] - Q[x]; }
            double rr = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
            double T = aij * akl / (aij + akl) * rr;
            double rts[RYS_NMAX], wts[RYS_NMAX];
            oracle_rys_roots(nroots, T, rts, wts);
            double pref = norm * Kab * Kcd / (aij * akl * sqrt(aij + akl));
            memset(prim, 0, sizeof(double) * ncomp * nf);
            for (int ir = 0; ir < nroots; ir++) {
                for (int x = 0; x < 3; x++)
                    build_1d(&G[x], li, lj, lk, ll, d, x == 2 ? pref * wts[ir] : 1.0,
                             P[x] - I->R[x], Q[x] - K->R[x], PQ[x], AB[x], CD[x],
                             rts[ir], aij, akl);
                int f = 0;
                for (int fl = 0; fl < nfl; fl++)
                for (int fk = 0; fk < nfk; fk++)
                for (int fj = 0; fj < nfj; fj++)
                for (int fi = 0; fi < nfi; fi++, f++) {
                    int e[3][4];
                    for (int x = 0; x < 3; x++) {
                        e[x][0] = comps[0][fi][x]; e[x][1] = comps[1][fj][x];
                        e[x][2] = comps[2][fk][x]; e[x][3] = comps[3][fl][x];
                    }
                    double gx = G[0].v[e[0][0]][e[0][1]][e[0][2]][e[0][3]];
                    double gy = G[1].v[e[1][0]][e[1][1]][e[1][2]][e[1][3]];
                    double gz = G[2].v[e[2][0]][e[2][1]][e[2][2]][e[2][3]];
                    if (!d) {
                        prim[f] += gx * gy * gz;
                    } else {
                        double dg[3];
                        for (int x = 0; x < 3; x++) {
                            int ix = e[x][0];
                            double v = -2.0 * ai * G[x].v[ix + 1][e[x][1]][e[x][2]][e[x][3]];
                            if (ix > 0) v += ix * G[x].v[ix - 1][e[x][1]][e[x][2]][e[x][3]];
                            dg[x] = v;
                        }
                        prim[f] += dg[0] * gy * gz;
                        prim[nf + f] += gx * dg[1] * gz;
                        prim[2 * nf + f] += gx * gy * dg[2];
                    }
                }
            }
            /* contract into every contraction column */
            int c = 0;
            for (int ci = 0; ci < I->nctr; ci++)
            for (int cj = 0; cj < J->nctr; cj++)
            for (int ck = 0; ck < K->nctr; ck++)
            for (int cl = 0; cl < L->nctr; cl++, c++) {
                double cc = I->coef[ci * I->nprim + ip] * J->coef[cj * J->nprim + jp] *
                            K->coef[ck * K->nprim + kp] * L->coef[cl * L->nprim + lp];
                for (int comp = 0; comp < ncomp; comp++) {
                    double *o = out + ((size_t)comp * nctr + c) * nf;
                    const double *pv = prim + (size_t)comp * nf;
                    for (int q = 0; q < nf; q++) o[q] += cc * pv[q];
                }
            }
        }
    }
    free(prim);
    free(G);
}

static int ao_loc_cart(const int *bas, int nbas, int *loc)
{
    loc[0] = 0;
    for (int i = 0; i < nbas; i++)
        loc[i + 1] = loc[i] + ncart(bas[i * BAS_SLOTS + ANG_OF]) * bas[i * BAS_SLOTS + NCTR_OF];
    return loc[nbas];
}

/* Cartesian s2kl integrals for i-shells [ish0, ish1):
 * out[comp][i - ao_loc[ish0]][j][k(k+1)/2 + l], k >= l, shape
 * (ncomp, nf_i, nao_cart, nao_cart (nao_cart+1)/2). deriv = 0: int2e; 1: int2e_ip1. */
int oracle_int2e_s2kl_cart(double *out, int deriv, const int *atm, int natm,
                           const int *bas, int nbas, const double *env, int ish0, int ish1)
{
    (void)natm;
    if (deriv < 0 || deriv > 1 || ish0 < 0 || ish1 > nbas || ish0 >= ish1) return -1;
    for (int i = 0; i < nbas; i++)
        if (bas[i * BAS_SLOTS + ANG_OF] > LMAX) return -2;
    int *loc = (int *)malloc(sizeof(int) * (nbas + 1));
    int nao = ao_loc_cart(bas, nbas, loc);
    long npair = (long)nao * (nao + 1) / 2;
    int nfi_blk = loc[ish1] - loc[ish0];
    int ncomp = deriv ? 3 : 1;
    size_t comp_stride = (size_t)nfi_blk * nao * npair;
    long nkl = (long)nbas * (nbas + 1) / 2;
    long ntask = (long)(ish1 - ish0) * nbas * nkl;
    size_t maxdim = 1;
    for (int i = 0; i < nbas; i++) {
        size_t dim = (size_t)ncart(bas[i * BAS_SLOTS + ANG_OF]) * bas[i * BAS_SLOTS + NCTR_OF];
        if (dim > maxdim) maxdim = dim;
    }
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 3 * maxdim * maxdim * maxdim * maxdim);
#pragma omp for schedule(dynamic, 4)
        for (long task = 0; task < ntask; task++) {
            long klidx = task % nkl;
            long rest = task / nkl;
            int jsh = (int)(rest % nbas);
            int ish = ish0 + (int)(rest / nbas);
            int ksh = (int)((sqrt(8.0 * klidx + 1.0) - 1.0) / 2.0);
            while ((long)ksh * (ksh + 1) / 2 > klidx) ksh--;
            while ((long)(ksh + 1) * (ksh + 2) / 2 <= klidx) ksh++;
            int lsh = (int)(klidx - (long)ksh * (ksh + 1) / 2);
            shell_t I, J, K, L;
            get_shell(&I, ish, atm, bas, env);
            get_shell(&J, jsh, atm, bas, env);
            get_shell(&K, ksh, atm, bas, env);
            get_shell(&L, lsh, atm, bas, env);
            quartet(buf, deriv, &I, &J, &K, &L);
            int nfi = ncart(I.l), nfj = ncart(J.l), nfk = ncart(K.l), nfl = ncart(L.l);
            int nf = nfi * nfj * nfk * nfl;
            int nctr = I.nctr * J.nctr * K.nctr * L.nctr;
            int c = 0;
            for (int ci = 0; ci < I.nctr; ci++)
            for (int cj = 0; cj < J.nctr; cj++)
            for (int ck = 0; ck < K.nctr; ck++)
            for (int cl = 0; cl < L.nctr; cl++, c++) {
                int i0 = loc[ish] - loc[ish0] + ci * nfi;
                int j0 = loc[jsh] + cj * nfj;
                int k0 = loc[ksh] + ck * nfk;
                int l0 = loc[lsh] + cl * nfl;
                int f = 0;
                for (int fl = 0; fl < nfl; fl++)
                for (int fk = 0; fk < nfk; fk++)
                for (int fj = 0; fj < nfj; fj++)
                for (int fi = 0; fi < nfi; fi++, f++) {
                    int k = k0 + fk, l = l0 + fl;
                    if (l > k) continue;
                    size_t idx = ((size_t)(i0 + fi) * nao + (j0 + fj)) * npair + (size_t)k * (k + 1) / 2 + l;
                    for (int comp = 0; comp < ncomp; comp++)
                        out[comp * comp_stride + idx] = buf[((size_t)comp * nctr + c) * nf + f];
                }
            }
        }
        free(buf);
    }
    free(loc);
    return 0;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of the following oracle calls (the CPU baseline's single-thread sample). */
void oracle_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
