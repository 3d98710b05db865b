/*
 * CPU FP64 oracle for one-electron integrals and their first derivatives.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline; never shipped).
 *
 * The reference takes these from PySCF (mf_grad.hcore_generator, get_ovlp,
 * int1e_ipovlp / int1e_ipkin / int1e_ipnuc / int1e_iprinv, used at
 * ccsd_gpu/gpu_ccsd_grad.py:554-583). Restated here from the textbook
 * Obara-Saika recurrences (overlap, kinetic) and Rys quadrature for the
 * nuclear attraction (point charges), with the same normalization conventions
 * as the ERIs (radial norm in the coefficients, 1/(2 sqrt pi) and
 * sqrt(3/(4 pi)) factors for s and p shells).
 *
 * kinds:  0 <i|j>          1 <i|-1/2 nabla^2|j>      2 <i| sum_C -Z_C/|r-C| |j>
 *         3 <d/dr i|j>     4 <d/dr i|-1/2 nabla^2|j> 5 <d/dr i| -Z_C/|r-C| |j> per C
 * Output (Cartesian AOs, nao x nao row-major): kinds 0-2 one matrix, 3-4 three
 * (x, y, z), 5: natm x 3 matrices.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

void oracle_rys_roots(int n, double T, double *r, double *w);

#define ATM_SLOTS 6
#define PTR_COORD 1
#define BAS_SLOTS 8
#define ATOM_OF 0
#define ANG_OF 1
#define NPRIM_OF 2
#define NCTR_OF 3
#define PTR_EXP 5
#define PTR_COEFF 6
#define L1MAX 8

static int ncart1(int l) { return (l + 1) * (l + 2) / 2; }

static void comp1(int l, int f, int *c)
{
    int n = 0;
    for (int x = l; x >= 0; x--)
        for (int y = l - x; y >= 0; y--) {
            if (n == f) { c[0] = x; c[1] = y; c[2] = l - x - y; return; }
            n++;
        }
}

static double snorm(int l)
{
    if (l == 0) return 0.28209479177387814347;
    if (l == 1) return 0.48860251190291992159;
    return 1.0;
}

/* 1-D overlap table S[i][j], i <= la, j <= lb (Obara-Saika). */
static void ovlp1d(double S[L1MAX + 1][L1MAX + 3], int la, int lb, double PA, double PB, double p)
{
    double t[2 * L1MAX + 4][L1MAX + 3];
    int lt = la + lb;
    memset(t, 0, sizeof(t));
    t[0][0] = sqrt(M_PI / p);
    for (int i = 0; i < lt; i++) {
        t[i + 1][0] = PA * t[i][0] + (i > 0 ? i / (2 * p) * t[i - 1][0] : 0.0);
    }
    /* HRR: S(i, j+1) = S(i+1, j) + (A-B) S(i, j) ; A-B = PB - PA */
    double AB = PB - PA;
    for (int j = 0; j < lb; j++)
        for (int i = 0; i <= lt - j - 1; i++) t[i][j + 1] = t[i + 1][j] + AB * t[i][j];
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++) S[i][j] = t[i][j];
}

/* 1-D nuclear factors for one Rys root: I[i][j], i <= la, j <= lb. */
static void nuc1d(double I[L1MAX + 1][L1MAX + 1], int la, int lb, double PA, double PC, double AB, double rt,
                  double p, double seed)
{
    double g[2 * L1MAX + 2][L1MAX + 1];
    int lt = la + lb;
    double c00 = PA - rt * PC;
    double b10 = 0.5 / p * (1.0 - rt);
    g[0][0] = seed;
    if (lt > 0) g[1][0] = c00 * seed;
    for (int i = 1; i < lt; i++) g[i + 1][0] = c00 * g[i][0] + i * b10 * g[i - 1][0];
    for (int j = 0; j < lb; j++)
        for (int i = 0; i <= lt - j - 1; i++) g[i][j + 1] = g[i + 1][j] - AB * g[i][j];
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++) I[i][j] = g[i][j];
}

int oracle_int1e_cart(double *out, int kind, const int *atm, int natm, const int *bas, int nbas,
                      const double *env)
{
    if (kind < 0 || kind > 5) return -1;
    int *loc = (int *)malloc(sizeof(int) * (nbas + 1));
    loc[0] = 0;
    for (int i = 0; i < nbas; i++) loc[i + 1] = loc[i] + ncart1(bas[i * BAS_SLOTS + ANG_OF]) * bas[i * BAS_SLOTS + NCTR_OF];
    int nao = loc[nbas];
    int deriv = kind >= 3;
    int ncomp = deriv ? 3 : 1;
    int nmat = kind == 5 ? natm * 3 : ncomp;
    memset(out, 0, sizeof(double) * (size_t)nmat * nao * nao);
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int ish = 0; ish < nbas; ish++)
    for (int jsh = 0; jsh < nbas; jsh++) {
        const int *bi = bas + ish * BAS_SLOTS, *bj = bas + jsh * BAS_SLOTS;
        int la = bi[ANG_OF], lb = bj[ANG_OF];
        const double *A = env + atm[bi[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        const double *B = env + atm[bj[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        int nfa = ncart1(la), nfb = ncart1(lb);
        double nrm = snorm(la) * snorm(lb);
        int lae = la + deriv; /* bra extended for the derivative */
        for (int ip = 0; ip < bi[NPRIM_OF]; ip++)
        for (int jp = 0; jp < bj[NPRIM_OF]; jp++) {
            double a = env[bi[PTR_EXP] + ip], b = env[bj[PTR_EXP] + jp], p = a + b;
            double AB2 = 0, P[3];
            for (int x = 0; x < 3; x++) { AB2 += (A[x] - B[x]) * (A[x] - B[x]); P[x] = (a * A[x] + b * B[x]) / p; }
            double K = exp(-a * b / p * AB2);
            /* values per (fa, fb) for up to natm*3 matrices */
            int nv = nmat;
            double *val = (double *)calloc((size_t)nv * nfa * nfb, sizeof(double));
            if (kind <= 1 || kind == 3 || kind == 4) {
                double S[3][L1MAX + 1][L1MAX + 3];
                for (int x = 0; x < 3; x++) ovlp1d(S[x], lae, lb + 2, P[x] - A[x], P[x] - B[x], p);
                /* kinetic 1-D: T(i,j) = -1/2 [ j(j-1) S(i,j-2) - 2b(2j+1) S(i,j) + 4 b^2 S(i,j+2) ] */
                for (int fa = 0; fa < nfa; fa++)
                for (int fb = 0; fb < nfb; fb++) {
                    int ca[3], cb[3];
                    comp1(la, fa, ca);
                    comp1(lb, fb, cb);
                    for (int c = 0; c < ncomp; c++) {
                        /* bra factor per direction: plain or derivative */
                        double s[3], t[3];
                        for (int x = 0; x < 3; x++) {
                            int i = ca[x], j = cb[x];
                            double sv, tv;
                            #define TK(ii) (-0.5 * ((j >= 2 ? j * (j - 1) * S[x][ii][j - 2] : 0.0) - 2 * b * (2 * j + 1) * S[x][ii][j] + 4 * b * b * S[x][ii][j + 2]))
                            if (deriv && x == c) {
                                sv = -2 * a * S[x][i + 1][j] + (i > 0 ? i * S[x][i - 1][j] : 0.0);
                                tv = -2 * a * TK(i + 1) + (i > 0 ? i * TK(i - 1) : 0.0);
                            } else {
                                sv = S[x][i][j];
                                tv = TK(i);
                            }
                            #undef TK
                            s[x] = sv;
                            t[x] = tv;
                        }
                        double v;
                        if (kind == 0 || kind == 3) v = s[0] * s[1] * s[2];
                        else v = t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2];
                        val[(c * nfa + fa) * nfb + fb] = v * K * nrm;
                    }
                }
            } else {
                int nroots = (lae + lb) / 2 + 1;
                for (int C = 0; C < natm; C++) {
                    const double *Rc = env + atm[C * ATM_SLOTS + PTR_COORD];
                    double Z = atm[C * ATM_SLOTS + 0];
                    double PC2 = 0;
                    for (int x = 0; x < 3; x++) PC2 += (P[x] - Rc[x]) * (P[x] - Rc[x]);
                    double r[16], w[16];
                    oracle_rys_roots(nroots, p * PC2, r, w);
                    double pref = -Z * 2 * M_PI / p * K * nrm;
                    for (int ir = 0; ir < nroots; ir++) {
                        double I[3][L1MAX + 1][L1MAX + 1];
                        for (int x = 0; x < 3; x++)
                            nuc1d(I[x], lae, lb, P[x] - A[x], P[x] - Rc[x], B[x] - A[x], r[ir], p,
                                  x == 2 ? w[ir] * pref : 1.0);
                        for (int fa = 0; fa < nfa; fa++)
                        for (int fb = 0; fb < nfb; fb++) {
                            int ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
                            comp1(la, fa, ca);
                            comp1(lb, fb, cb);
                            if (!deriv) {
                                val[fa * nfb + fb] += I[0][ca[0]][cb[0]] * I[1][ca[1]][cb[1]] * I[2][ca[2]][cb[2]];
                            } else {
                                for (int c = 0; c < 3; c++) {
                                    double f[3];
                                    for (int x = 0; x < 3; x++) {
                                        int i = ca[x], j = cb[x];
                                        f[x] = (x == c) ? (-2 * a * I[x][i + 1][j] + (i > 0 ? i * I[x][i - 1][j] : 0.0))
                                                        : I[x][i][j];
                                    }
                                    val[((C * 3 + c) * nfa + fa) * nfb + fb] += f[0] * f[1] * f[2];
                                }
                            }
                        }
                    }
                }
            }
            /* contract into AO matrices */
            for (int ci = 0; ci < bi[NCTR_OF]; ci++)
            for (int cj = 0; cj < bj[NCTR_OF]; cj++) {
                double cc = env[bi[PTR_COEFF] + ci * bi[NPRIM_OF] + ip] * env[bj[PTR_COEFF] + cj * bj[NPRIM_OF] + jp];
                for (int m = 0; m < nv; m++)
                    for (int fa = 0; fa < nfa; fa++)
                        for (int fb = 0; fb < nfb; fb++) {
                            int i = loc[ish] + ci * nfa + fa, j = loc[jsh] + cj * nfb + fb;
                            out[((size_t)m * nao + i) * nao + j] += cc * val[(m * nfa + fa) * nfb + fb];
                        }
            }
            free(val);
        }
    }
    free(loc);
    return 0;
}
