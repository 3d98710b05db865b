/*
 * ccsd_gpu_amd -- C ABI of the MI355X (gfx950) CCSD-gradient hot path.
 *
 * All device pointers are HIP device memory; `stream` is a hipStream_t (NULL =
 * the null stream). Every entry returns 0 on success and a negative code on
 * failure (-1 null pointer, -2 bad shape, -3 unsupported input, -4 HIP error,
 * -5 allocation failure); ccsd_last_error() then describes the failure
 * (thread-local). Calls are stream-ordered; a plan must not be used from two
 * host threads at once.
 *
 * Integral layout (what the reference's wrappers return):
 *   out[comp][i - ao_loc[ish0]][j][k*(k+1)/2 + l],  k >= l,
 *   shape (ncomp, nf, nao, nao*(nao+1)/2), ncomp = 1 (int2e) or 3 (int2e_ip1,
 *   x, y, z of d/dr1), AO order of the native basis (Cartesian or spherical).
 */
#ifndef CCSD_GPU_AMD_H
#define CCSD_GPU_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccsd_plan ccsd_plan;

/* Text of the last error on this host thread. */
const char* ccsd_last_error(void);

/* Basis plan from libcint-layout host arrays (the `mol._atm`, `mol._bas`,
 * `mol._env` the reference hands to its kernels, ccsd_gpu/cuda/int2e_ip1_module.py:273-284,
 * uploaded per call at ccsd_gpu/cuda/int2e_ip1_gpu.py:181-182) and, optionally,
 * `mol.cart2sph_coeff()` as a row-major (nao_cart, nao_sph) host matrix
 * (ccsd_gpu/cuda/int2e_ip1_gpu.py:243). Replaces the reference's per-call
 * topology cache `_get_basis_cached` (ccsd_gpu/cuda/int2e_ip1_gpu.py:85-141). */
int ccsd_plan_create(ccsd_plan** plan, const int* atm, int natm, const int* bas, int nbas, const double* env,
                     int nenv, const double* c2s);
int ccsd_plan_destroy(ccsd_plan* plan);

/* New geometry (same basis): replaces env; coordinates are read from it. */
int ccsd_plan_set_env(ccsd_plan* plan, const double* env, int nenv);

/* Number of AOs (cart != 0: Cartesian). Returns the count or a negative code. */
int ccsd_plan_nao(const ccsd_plan* plan, int cart);

/* int2e, s2kl, i-shells [ish0, ish1): the drop-in for
 * compute_int2e_gpu(mol, b0, b1) (ccsd_gpu/cuda/int2e_gpu.py:186-215), i.e.
 * mol.intor('int2e', aosym='s2kl', shls_slice=(b0,b1,0,nbas,0,nbas,0,nbas)).
 * cart = 1: Cartesian AOs; cart = 0: spherical (needs c2s at plan creation).
 * out: device array (nf, nao, nao_pair), fully overwritten. */
int ccsd_int2e_s2kl(ccsd_plan* plan, double* out, int ish0, int ish1, int cart, void* stream);

/* int2e_ip1, s2kl: the drop-in for compute_int2e_ip1_gpu(mol, b0, b1)
 * (ccsd_gpu/cuda/int2e_ip1_gpu.py:296-326), kernel fill_int2e_ip1_kernel
 * (ccsd_gpu/cuda/int2e_ip1_module.py:273-667). out: (3, nf, nao, nao_pair). */
int ccsd_int2e_ip1_s2kl(ccsd_plan* plan, double* out, int ish0, int ish1, int cart, void* stream);

/* Fused ERI-derivative contraction for i-shells [ish0, ish1):
 *   de[atom(i)][x] += scale * sum_{i, j, k>=l} int2e_ip1[x](i j | k l) * gamma[i][j][k(k+1)/2 + l]
 * with gamma in the Cartesian s2kl layout (nao_cart, nao_cart, nao_cart*(nao_cart+1)/2).
 * With scale = -2 and gamma the packed 2-RDM rows whose kl diagonal is halved this
 * is the reference's per-block `de -= 2 * eri1 @ dm2buf` (ccsd_gpu/gpu_ccsd_grad.py:430-459)
 * without storing eri1. de: device (natm, 3), accumulated into. */
int ccsd_grad_ip1_gamma(ccsd_plan* plan, double* de, const double* gamma, double scale, int ish0, int ish1,
                        void* stream);

/* Plan options (no counterpart in the reference, which has one kernel and no
 * tuning). option 0: 1 (default) lets the fused gradient use the
 * register-resident per-class kernels where a class fits, 0 forces the LDS
 * kernel for every class (comparison and testing). option 1: 1 (default) times
 * the kernel candidates of each class once per plan and keeps the fastest;
 * 0 takes the first candidate (register kernel, 4 waves per block). option 2:
 * 1 (default) runs the class launches of a fused pass on four side streams
 * (forked from and joined back to `stream` with events), 0 on `stream` alone.
 * option 3: 1 orders the (unscreened) ket pairs of each primitive group in
 * shell order, so neighbouring lanes gather neighbouring packed Gamma; 0
 * (default) by surviving primitive pairs, descending. option 4: primitive split
 * of the register kernels: 0 (default) adds a candidate whose expensive chunks
 * are cut into up to 16 pieces along the ket primitive-pair loop (each piece at
 * most ~1/2048 of the class's work; the autotuner decides per class), 1 turns
 * the split off, S = 2..254 splits every chunk S ways (with option 1 = 0 that
 * split register kernel is the one taken). option 5: 1 runs a fused pass as ONE
 * persistent launch over every class of every molecule (one-wave blocks draw
 * work items, runs of one class's units, from a device counter until the list
 * is exhausted): with the register-resident bodies of every class in one
 * kernel when each class of the pass has one (work lists as the autotuner
 * chose them), with the LDS kernel's body otherwise; 2 the same with the LDS
 * kernel's body always; 0 (default) one launch per class as above. option 6:
 * 0 (default) times only the register-kernel candidates of classes that have
 * them (the generic kernel is timed for the other classes), 1 times the
 * generic candidates everywhere (a slower first pass of a plan: their work
 * lists are the largest). option 7: order of the register-body persistent
 * launch's work items: 0 (default) class by class (costliest class first),
 * 1 by decreasing cost only. */
int ccsd_plan_set_option(ccsd_plan* plan, int option, int value);

/* Plan over a batch of molecules: atm/bas/env concatenated (shell atom indices
 * and env offsets global, as PySCF's gto.conc_env builds them), molecule m owning
 * the next mol_nbas[m] shells. Shell pairs and quartets never mix molecules.
 * Batch counterpart of ccsd_plan_create for many molecules of uneven size; the
 * reference handles one molecule per call (ccsd_gpu/cuda/int2e_ip1_gpu.py:296). */
int ccsd_plan_create_batch(ccsd_plan** plan, int nmol, const int* mol_nbas, const int* atm, int natm, const int* bas,
                           int nbas, const double* env, int nenv);

/* Schwarz screening of the fused gradient: quartets with q_ij * q_kl < thr are
 * skipped, q_ij = max over functions a in shell i, b in shell j of sqrt((ab|ab)).
 * thr = 0 (default) evaluates every quartet. The reference screens nothing
 * (all nbas^2 * n_kl quartets, ccsd_gpu/cuda/int2e_ip1_gpu.py:199-214). */
int ccsd_plan_screen(ccsd_plan* plan, double thr, void* stream);

/* The Schwarz factors q (device, nbas x nbas, symmetric) of the plan's geometry. */
int ccsd_schwarz(ccsd_plan* plan, double* q, void* stream);

/* Shell quartets evaluated by the last integral or fused-gradient call. */
int ccsd_plan_quartets(const ccsd_plan* plan, long long* count);

/* The launches of the last integral or fused-gradient call, 10 numbers each, in
 * launch order: li, lj, lk, ll, general-contraction variant (0/1), kernel (0 LDS,
 * 1 register, 2 class inside the one persistent launch), waves per block, quartets per chunk (low 16 bits; bits 16+: the
 * register kernel's primitive split, 0 = none), waves per quartet (register
 * kernel), quartets. *nrows receives the count; up to max_rows rows are written
 * to `rows` (may be NULL). No counterpart in the reference (profiling aid). */
int ccsd_plan_last_launches(const ccsd_plan* plan, long long* rows, int max_rows, int* nrows);

/* Fused ERI-derivative contraction over all shells of every molecule of the plan
 * with the separable two-particle density of Cartesian densities D and P
 * (dm, pm: per molecule a row-major nao_m x nao_m block, blocks in molecule order):
 *   Gamma_ijkl = D_ij P_kl + P_ij D_kl - (D_il P_kj + D_ik P_lj + P_il D_jk + P_ik D_jl) / 4
 *   de[atom(i)][x] += scale * sum_{i, j, k>=l} int2e_ip1[x](i j | k l) * Gamma'_ijkl
 * with Gamma' halved on k == l (the packing of ccsd_grad_ip1_gamma). This is the
 * density-matrix part of the reference's gradient assembly (the vhf1 J/K terms
 * contracted with dm1p, ccsd_gpu/gpu_ccsd_grad.py:469-491) evaluated for many
 * molecules at once without a four-index density. de: device (natm_total, 3). */
int ccsd_grad_ip1_separable(ccsd_plan* plan, double* de, const double* dm, const double* pm, double scale,
                            void* stream);

/* Both densities in one pass (single-molecule plans): Gamma' = gamma (explicit,
 * packed as for ccsd_grad_ip1_gamma) + the separable form of dm, pm (nao_c x nao_c,
 * Cartesian). This is the reference's whole two-electron gradient, the
 * `-2 eri1 @ dm2buf` GEMV and the vhf1 J/K terms with dm1p (ccsd_gpu/gpu_ccsd_grad.py:
 * 430-491), over i-shells [ish0, ish1) without forming eri1 or a four-index
 * separable density. */
int ccsd_grad_ip1_gamma_sep(ccsd_plan* plan, double* de, const double* gamma, const double* dm, const double* pm,
                            double scale, int ish0, int ish1, void* stream);

/* One-electron integrals in the Cartesian AO basis (row-major nao_c x nao_c
 * matrices, fully overwritten): kind 0 overlap, 1 kinetic, 2 nuclear attraction
 * (all nuclei), 3 <d/dr i|j> (3 matrices), 4 <d/dr i|T|j> (3), 5 <d/dr i|-Z_C/|r-C||j>
 * for every nucleus C (natm x 3). These replace the PySCF one-electron pieces of
 * the gradient -- mf_grad.get_ovlp, hcore_generator (int1e_ipkin, int1e_ipnuc,
 * int1e_iprinv) -- called at ccsd_gpu/gpu_ccsd_grad.py:554-583. Parts of an element
 * are summed exactly in fixed point (the same bits on every run) in an accumulator
 * the plan owns: calls on one plan that overlap (other streams or threads) race. */
int ccsd_int1e(ccsd_plan* plan, int kind, double* out, void* stream);

/* Batched C = A B on the matrix cores, float64 operands in device memory:
 * A [batch][M][K] (row stride lda, batch stride strideA), B [batch][K][N],
 * C [batch][M][N] (overwritten). mode 0: FP32 MFMA (operands rounded to FP32);
 * mode 1: split-BF16 (x = hi + lo, three bf16 MFMA products); both accumulate
 * in FP32. These replace the cuBLAS / cuTENSOR GEMMs behind the CuPy einsum
 * calls of the lambda and RDM code (ccsd_gpu/gpu_ccsd_lambda.py:68-337,
 * ccsd_gpu/gpu_ccsd_grad.py:101-246) on the mixed-precision path. */
int ccsd_gemm_mixed(int mode, const double* A, const double* B, double* C, int M, int N, int K, long long lda,
                    long long ldb, long long ldc, long long strideA, long long strideB, long long strideC, int batch,
                    void* stream);

/* Split a float64 matrix (rows x cols, row stride ld) into two bfloat16 planes
 * hi = bf16(x), lo = bf16(x - hi), each rows x cols contiguous: the operand
 * form of ccsd_gemm_bf16x3_nt. No counterpart in the reference (its
 * contractions are FP64 CuPy einsums). */
int ccsd_split_bf16(const double* x, void* hi, void* lo, long long rows, long long cols, long long ld, void* stream);

/* out[i] = beta * out[i] + sum_{s < nslices} ws[s * n + i] (device arrays), in a
 * fixed order: the reduction of a K-split GEMM's partial products. ws is scratch:
 * many slices over few outputs are summed in groups first, with the group
 * partials left in ws. Plain kernels only (no memset), so the call may sit inside
 * a captured HIP graph. No counterpart in the reference (its einsums are single
 * cuBLAS/cuTENSOR calls). */
int ccsd_sum_slices(double* ws, double* out, long long n, int nslices, int beta, void* stream);

/* In place on a dense (n, n, n, n) device tensor:
 *   t[p][q][r][s] <- scale * (t[pqrs] + t[qprs] + t[pqsr] + t[qpsr]).
 * The (ij)/(kl) symmetrisation of the two-particle density, applied in the MO
 * basis before the AO transform (the reference symmetrises the AO tensor with
 * two permuted copies, ccsd_gpu/gpu_ccsd_grad.py:324-326). */
int ccsd_sym_pairs(double* t, int n, double scale, void* stream);

/* out[r][a][b] = w (t[k,l,a,b] + t[l,k,a,b] + t[k,l,b,a] + t[l,k,b,a]) / 2 for the
 * pairs r < npair with (kk[r], ll[r]) = (k, l), k >= l (device int64 arrays), w =
 * 1/2 on the diagonal k == l, else 1; t (n, n, m, m) and out (npair, m, m)
 * contiguous. The (ij)(kl) symmetrisation of the AO 2-RDM
 * (ccsd_gpu/gpu_ccsd_grad.py:324-326) and its kl packing with the diagonal
 * halved (:430-431), done once on the half-transformed density. */
int ccsd_pack_pairs_sym(const double* t, const long long* kk, const long long* ll, int n, int m, long long npair,
                        double* out, void* stream);

/* out[r][i][j] = (w_kl / w_ij) z[k][l][p(i,j)] for r = k(k+1)/2 + l, k >= l < n, with
 * p(i,j) = a(a+1)/2 + b (a = max(i,j), b = min(i,j)), w = 1 on a pair's diagonal
 * else 2; z (n, n, m(m+1)/2) and out (n(n+1)/2, m, m) contiguous device arrays.
 * With the 2-RDM symmetrised and packed over its first (MO) pair before the AO
 * transform of the second pair, this moves the packing to the second (AO) pair,
 * the layout of ccsd_gpu/gpu_ccsd_grad.py:330-333 and :430-431 for the first
 * pair's remaining transforms. */
int ccsd_repack_pairs(const double* z, int n, int m, double* out, void* stream);

/* ccsd_repack_pairs on float32 arrays (the MO -> AO transform of the 2-RDM carried in
 * FP32 storage and FP32 matrix products, the mixed gradient's default). */
int ccsd_repack_pairs_f32(const float* z, int n, int m, float* out, void* stream);

/* gamma[i][j][k(k+1)/2 + l] += w_kl * Gsep_ijkl for k >= l, w_kl = 1/2 if k == l else 1,
 *   Gsep_ijkl = D_ij P_kl + P_ij D_kl - (D_il P_kj + D_ik P_lj + P_il D_jk + P_ik D_jl) / 4,
 * gamma (n, n, n(n+1)/2), dm = D and pm = P (n, n) row-major, all device arrays, n <= 2048.
 * The HF-density terms of the gradient (the reference's vhf1 blocks contracted with dm1p,
 * ccsd_gpu/gpu_ccsd_grad.py:462-489, :570, :583) added into the packed CCSD density once,
 * so that one explicit ccsd_grad_ip1_gamma pass covers both. */
int ccsd_add_separable_packed(double* gamma, const double* dm, const double* pm, int n, void* stream);

/* Singles terms of the lambda update that contract a doubles-shaped tensor
 * x (o, o, v, v, contiguous) with a singles-shaped y (y[j][b] at j*sy0 + b*sy1):
 *   out[i][a] += sum_{j,b} (c1 x[i,j,a,b] + c2 x[i,j,b,a] + c3 x[j,i,a,b] + c4 x[j,i,b,a]) y[j,b]
 * in one read of x; ws is a device workspace of 2 o^2 v doubles, out (o, v)
 * contiguous. Replaces the einsum pairs of ccsd_gpu/gpu_ccsd_lambda.py:244-245,
 * 256-257, 288-291, 302-303, 322-323 (each of which reads its tensor twice or
 * through a permuted copy). nvir <= 682. Deterministic (fixed summation order). */
int ccsd_l1_terms(const double* x, int o, int v, const double* y, long long sy0, long long sy1, double c1, double c2,
                  double c3, double c4, double* ws, double* out, void* stream);

/* Unpack pair blocks: for p < npair with (iu[p], ju[p]) = (i, j), i <= j (device int64
 * arrays), out[i,j,a,b] = scale x[p,a,b] and out[j,i,b,a] = scale x[p,a,b]; for
 * i == j, out[i,i,a,b] = scale (x[p,a,b] + x[p,b,a]) / 2. x (npair, v, v) and out
 * (o, o, v, v) contiguous. The mirror step of the i <= j half of the l2 (ac|bd)
 * term (the reference evaluates all (i, j), ccsd_gpu/gpu_ccsd_lambda.py:238-243). */
int ccsd_unpack_pairs(const double* x, const long long* iu, const long long* ju, int o, int v, long long npair,
                      double scale, double* out, void* stream);

/* The l2 (ac|bd) term over index pairs, packing step: for the i <= j pair blocks
 * (iu[r], ju[r]) (device int64, r < npair) of l2 (o, o, v, v) and the index pairs
 * (tc[P], td[P]), tc >= td (device int64, P < np = v(v+1)/2), writes
 * A+[r,P] = (l2[i,j,c,d] + l2[i,j,d,c]) (halved when c = d) and A-[r,P] = l2[i,j,c,d] - l2[i,j,d,c]
 * as (2, npair, ld) rows, columns np..ld-1 zero: float64 into out (split 0) or
 * split-BF16 hi / lo planes into out / lo (split 1). With W+-[Q,P] = (ac|bd) +- (ad|bc)
 * the products A+- W+-^T hold the symmetric and antisymmetric halves of
 * sum_cd l2[i,j,c,d] (ac|bd) at a quarter of the o^2 v^4 work of the reference's
 * contraction (ccsd_gpu/gpu_ccsd_lambda.py:238-243). */
int ccsd_vvvv_pm_pack(const double* l2, int o, int v, const long long* iu, const long long* ju, long long npair,
                      const long long* tc, const long long* td, long long np, long long ld, int split, void* out,
                      void* lo, void* stream);

/* Unpacking step: from x = (X+, X-) (2, o(o+1)/2, ldx) float64, the products of the
 * packing step for all i <= j pairs in triu order, writes out (o, o, v, v):
 * out[i,j,a,b] = scale (X+[Q] + s X-[Q]) / 2 of the pair (i, j) for i < j and of
 * (j, i) with (a, b) swapped for i > j, Q = (max(a,b), min(a,b)) packed as
 * max(max+1)/2 + min, s = +1 when the first index is the larger; on i = j the
 * symmetrised block scale X+[Q] / 2. */
int ccsd_vvvv_pm_unpack(const double* x, int o, int v, long long ldx, double scale, double* out, void* stream);

/* out[i,j,a,b] = beta * out + (c1 x[i,j,a,b] + c2 x[i,j,b,a] + c3 x[j,i,a,b] + c4 x[j,i,b,a]) * d[i,j,a,b]
 * for x and out any strided (o, o, v, v) views (host arrays of the four element
 * strides of i, j, a, b; reads and writes are coalesced when the b strides are
 * 1; out must not overlap x), d optional (NULL = 1, else a contiguous
 * (o, o, v, v) tensor). Pair transposes go through LDS tiles. Replaces the permuted
 * elementwise forms of ccsd_gpu/gpu_ccsd_lambda.py:312 (l2 transposes),
 * :315-317 (tmp + tmp^T) and :327-334 (symmetrise and divide by D_ijab). */
int ccsd_pair_combine(const double* x, const long long* xstrides, int o, int v, double c1, double c2, double c3,
                      double c4, const double* d, double* out, const long long* ostrides, int beta, void* stream);

/* Closing step of the lambda update:
 *   out[i,j,a,b] = (x[i,j,a,b] + x[j,i,b,a] + u[i,a] w[j,b] + u[j,b] w[i,a] + sc s[i,j,a,b]) d[i,j,a,b]
 * with x, s, d, out contiguous (o, o, v, v) (out not overlapping x) and u, w contiguous
 * (o, v); s may be NULL; sc = *sdev when sdev is given (a device scalar read when the
 * kernel runs), else shost. The symmetrisation and division by D_ijab of
 * ccsd_gpu/gpu_ccsd_lambda.py:327-334 with the l1 x w4 term (:259) and the bare
 * integral term (:295, s = its (i,a) <-> (j,b) symmetrised form) folded in, so that
 * neither needs a pass of its own over l2new. */
int ccsd_l2_finish(const double* x, int o, int v, const double* u, const double* w, const double* s,
                   const double* sdev, double shost, const double* d, double* out, void* stream);

/* DIIS with its history in FP32 (the mixed-precision lambda solve; the reference keeps
 * PySCF's FP64 lib.diis on the host, ccsd_gpu/gpu_ccsd_lambda.py:376-399). The new
 * vector x = concat(s0[0:n0], s1[0:n-n0])
 * (+ rhs when given), FP64 device arrays. One pass: xs = f32(x) (when given),
 * prev_out = x (when given); with prev: e = x - prev, es = f32(e) (when given) and
 * dots[r] = sum_k E[r][k] ef[k] for r < nrows (row r of e_rows, stride lde, or ef itself
 * for r = slot), dots[nrows] = sum_k ef[k]^2, ef = the stored e (or e when es is NULL),
 * summed in FP64 in a fixed order; ws: ccsd_diis_workspace(n, nrows) doubles.
 * nrows <= 15. */
long long ccsd_diis_workspace(long long n, int nrows);
int ccsd_diis_push(const double* s0, const double* s1, long long n0, long long n, const double* rhs, const double* prev,
                   float* xs, float* es, const float* e_rows, long long lde, int nrows, int slot, double* prev_out,
                   double* dots, double* ws, long long ws_len, void* stream);

/* out[k] = base[k] (0 when NULL) + sum_{r < nrows} coef[r] rows[r][k]: FP32 rows of
 * stride ldx, host coefficients, FP64 base and out; nrows <= 16. The DIIS
 * extrapolation (base NULL) and input combination (base = the last extrapolation,
 * coefficients negated, rows = the error vectors) over an FP32 history. */
int ccsd_diis_combine(const float* rows, long long ldx, int nrows, const double* coef, const double* base, double* out,
                      long long n, void* stream);

/* Copy a strided view into a contiguous destination: ndim <= 6 dimensions in the
 * destination's order (sizes), with the source's element stride of each
 * (strides; any permutation). Writes float64 to dst, or, when dst is NULL,
 * the split-BF16 planes hi = bf16(x), lo = bf16(x - hi) (the operand form of
 * ccsd_gemm_bf16x3_nt). Transposing copies go through 32 x 32 LDS tiles. The
 * permuted copies of the reference's einsum operands (CuPy's
 * .transpose(...).copy() / cuTENSOR, e.g. ccsd_gpu/gpu_ccsd_lambda.py:120, 162, 312)
 * and the planner's operand splits. */
int ccsd_permute(const double* src, int ndim, const long long* sizes, const long long* strides, double* dst, void* hi,
                 void* lo, void* stream);

/* dst (+)= sum_t coef[t] * src_t for nterms <= 4 strided views of one shape (ndim <= 6,
 * sizes in dst's order; dstrides and sstrides[t * ndim + k] in elements; dst must
 * not overlap a source; beta 0 overwrites, 1 accumulates). Terms contiguous along
 * dst's last dimension are read directly, transposed ones through 32 x 32 LDS
 * tiles. The permuted sums of the RDM code (ccsd_gpu/gpu_ccsd_grad.py:129-246,
 * e.g. dovov = goovv.transpose(0,2,1,3)*2 - goovv.transpose(1,2,0,3)) and of the
 * 2-RDM assembly (:272-300). */
int ccsd_combine(int ndim, const long long* sizes, double* dst, const long long* dstrides, int nterms,
                 const double* const* srcs, const long long* sstrides, const double* coef, int beta, void* stream);

/* y (+)= alpha * A x in FP64: cols = 0 reads A as M rows of length K
 * (y[m] = sum_k A[m*lda + k] x[k*incx]), cols = 1 as K rows of length M
 * (y[m] = sum_k A[k*lda + m] x[k*incx]); beta 0 overwrites, 1 accumulates.
 * With nslices > 1 the K range is cut into slices whose partial sums go to ws
 * (nslices * M doubles) and are added in a fixed order. The matrix-vector
 * einsums of the lambda code (ccsd_gpu/gpu_ccsd_lambda.py:264-265, 280,
 * 283-284, 294, 298-299) and of the Z-vector's orbital Hessian. */
int ccsd_gemv(int cols, const double* A, long long M, long long K, long long lda, const double* x, long long incx,
              double* y, double alpha, int beta, double* ws, int nslices, void* stream);

/* Size (doubles) of the device scratch of ccsd_gemv_noalloc (4 MiB per device) and
 * the largest device ordinal it serves. */
#define CCSD_GEMV_SCRATCH_DOUBLES (1 << 19)
#define CCSD_MAX_DEVICES 64

/* ccsd_gemv without a workspace argument and without any allocation: the slice
 * partials go to a file-scope device array of CCSD_GEMV_SCRATCH_DOUBLES doubles
 * (one per device, held until the process ends). No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, allocation-free
 * serving loops). Same arguments and result as ccsd_gemv with the same nslices;
 * nslices * M above CCSD_GEMV_SCRATCH_DOUBLES (with nslices > 1) is rejected with
 * CCSD_ERR_SHAPE. Calls on one device share its scratch: calls that overlap on one
 * device (concurrent streams or threads) race. */
int ccsd_gemv_noalloc(int cols, const double* A, long long M, long long K, long long lda, const double* x,
                      long long incx, double* y, double alpha, int beta, int nslices, void* stream);

/* C[b] (+)= A[b] B[b]^T from split-BF16 operands (A planes [batch][M][K] with row
 * stride lda, B planes [batch][N][K] with row stride ldb, both K-contiguous;
 * strides in elements), three bf16 MFMA products per tile
 * (hi.hi + hi.lo + lo.hi), FP32 accumulation, float64 C [batch][M][N] with row
 * stride ldc; beta 0 overwrites, 1 accumulates. The mixed-precision engine for
 * the o^2 v^4 / o^3 v^3 contractions of the lambda and RDM code
 * (ccsd_gpu/gpu_ccsd_lambda.py:125, 132, 238-243; ccsd_gpu/gpu_ccsd_grad.py:222). */
int ccsd_gemm_bf16x3_nt(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C, int M, int N,
                        int K, long long lda, long long ldb, long long strideA, long long strideB, long long strideC,
                        long long ldc, int batch, int beta, void* stream);

/* ccsd_gemm_bf16x3_nt with a scattered destination: element (row, col) of batch b
 * is written to C[off_b[b] + off_r[row] + off_c[col]] (offset tables in device
 * memory, elements), as C = alpha * AB^T (beta 0) or C += alpha * AB^T (beta 1).
 * Accumulates a contraction term straight into a four-index tensor whose index
 * order differs from the GEMM's (the lambda update's `l2new += term.permute(...)`,
 * ccsd_gpu/gpu_ccsd_lambda.py:248-323), without the permuted copy. */
int ccsd_gemm_bf16x3_nt_scatter(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C, int M,
                                int N, int K, long long lda, long long ldb, long long strideA, long long strideB,
                                const long long* off_b, const long long* off_r, const long long* off_c, int batch,
                                double alpha, int beta, void* stream);

/* Block shape of ccsd_gemm_bf16x3_nt(_scatter) for K > 256 on aligned operands
 * (K and the row strides multiples of 8, 16-byte aligned planes): 0 (the default)
 * 128 x 128 blocks staging K by 64 through one LDS buffer; 1 the same blocks staging
 * K by 32 through two LDS buffers; 2 128 x 256 blocks of 8 waves, K by 32, two LDS
 * buffers. Process-wide; every variant forms the same products in the same order per
 * output element. Returns CCSD_ERR_SHAPE for another value. No counterpart in the
 * reference: a tuning switch for timing and testing the block shapes on one problem. */
int ccsd_gemm_set_variant(int v);

/* ccsd_split_bf16 from an N-d strided view (ndim <= 8; sizes, element strides as
 * host arrays): the planes are contiguous in the view's row-major order, so an
 * operand is permuted and split in one pass. */
int ccsd_split_bf16_strided(const double* x, void* hi, void* lo, int ndim, const long long* sizes,
                            const long long* strides, void* stream);

/* C[off_b[b] + off_r[r] + off_c[c]] = beta * C[...] + alpha * sum_s ws[s][b][r][c]
 * (ws: (nslices, batch, M, N) contiguous): the K-split reduction into a scattered
 * destination. */
int ccsd_scatter_sum(const double* ws, int nslices, int batch, int M, int N, double* C, const long long* off_b,
                     const long long* off_r, const long long* off_c, double alpha, int beta, void* stream);

/* Rys roots (u = t^2, ascending) and weights for `count` arguments T (device
 * arrays; r, w are (count, n)). Counterpart of rys_roots_simple
 * (ccsd_gpu/cuda/int2e_ip1_module.py:192-244), n = 1..9. */
int ccsd_rys_roots(int n, const double* T, int count, double* r, double* w, void* stream);

#ifdef __cplusplus
}
#endif
#endif
