// C-ABI of the ERI library: basis plans, class enumeration and launches.
//
// The reference launches one CUDA kernel per decontracted i-shell with the
// raw libcint arrays (ccsd_gpu/cuda/int2e_ip1_gpu.py:148-216). Here a plan
// holds the basis once in device memory; every call enumerates the bra shell
// pairs (ish in [ish0, ish1), all jsh) and ket pairs (ksh >= lsh) by angular
// momentum class and launches one quartet kernel per (bra class, ket class).
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "eri_common.h"
#include "../../include/ccsd_gpu_amd.h"

size_t ccsd_eri_lds_bytes(const EriClassArgs& a);
hipError_t ccsd_launch_eri_class(const EriClassArgs& a, hipStream_t stream);
hipError_t ccsd_launch_c2s(double* out, const double* in, const void* ti, const void* tall, int nfi_s, int nao_s,
                           long long npair_s, int nao_c, long long npair_c, int nfi_c, int ncomp, hipStream_t stream);
hipError_t ccsd_launch_rys(int n, const double* T, int count, double* r, double* w, hipStream_t stream);
size_t ccsd_c2s_term_size();
hipError_t ccsd_launch_fill(double* p, long long n, double v, hipStream_t stream);
bool ccsd_fused_reg_supported(int li, int lj, int lk, int ll, int max_nckl);
hipError_t ccsd_launch_fused_reg(const EriClassArgs& a, hipStream_t stream);
hipError_t ccsd_launch_gemm_bf16x3_nt(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C,
                                      int M, int N, int K, long long lda, long long ldb, long long sA, long long sB,
                                      long long sC, long long ldc, int batch, int beta, hipStream_t stream);
hipError_t ccsd_launch_split_bf16(const double* x, void* hi, void* lo, long long rows, long long cols, long long ld,
                                  hipStream_t stream);
hipError_t ccsd_launch_gemm_mixed(int mode, const double* A, const double* B, double* C, int M, int N, int K,
                                  long long lda, long long ldb, long long ldc, long long sA, long long sB,
                                  long long sC, int batch, hipStream_t stream);
hipError_t ccsd_launch_int1e(int kind, double* out, const ShellInfo* shells, const double* env, const double* coords,
                             const double* charges, int natm, int nbas, int nao, hipStream_t stream);

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* where)
{
    return fail(CCSD_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                      \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);  \
    } while (0)

struct C2sTermHost {
    int n;
    int idx[6];
    double c[6];
};

struct PairClass {
    int l1, l2;
    std::vector<int2> pairs;
    int max_np, max_nc;
    int2* d_pairs = nullptr;
};

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
int nsph(int l) { return 2 * l + 1; }

}  // namespace

struct ccsd_plan {
    int natm = 0, nbas = 0, nenv = 0;
    std::vector<int> atm, bas;
    std::vector<double> env;
    std::vector<ShellInfo> shells;
    int nao_cart = 0, nao_sph = 0;
    int lmax = 0;
    int device = 0;
    ShellInfo* d_shells = nullptr;
    double* d_env = nullptr;
    double* d_coords = nullptr;
    double* d_charges = nullptr;
    std::vector<PairClass> ket;                      // ksh >= lsh
    std::map<std::pair<int, int>, std::vector<PairClass>> bra_cache;
    bool has_c2s = false;
    bool fused_reg = true;                           // register kernels for the fused gradient
    std::vector<C2sTermHost> c2s;                    // per spherical AO
    void* d_c2s = nullptr;
};

namespace {

std::vector<PairClass> make_classes(const ccsd_plan* p, const std::vector<int2>& pairs)
{
    std::vector<PairClass> cls;
    std::map<std::pair<int, int>, int> where;
    for (const int2& pr : pairs) {
        const ShellInfo &a = p->shells[pr.x], &b = p->shells[pr.y];
        auto key = std::make_pair(a.l, b.l);
        auto it = where.find(key);
        if (it == where.end()) {
            where[key] = (int)cls.size();
            PairClass c;
            c.l1 = a.l;
            c.l2 = b.l;
            c.max_np = 0;
            c.max_nc = 0;
            cls.push_back(c);
            it = where.find(key);
        }
        PairClass& c = cls[it->second];
        c.pairs.push_back(pr);
        c.max_np = std::max(c.max_np, a.nprim * b.nprim);
        c.max_nc = std::max(c.max_nc, a.nctr * b.nctr);
    }
    // similar primitive counts next to each other (threads of a wave loop alike)
    for (PairClass& c : cls)
        std::stable_sort(c.pairs.begin(), c.pairs.end(), [p](const int2& x, const int2& y) {
            const ShellInfo &a0 = p->shells[x.x], &a1 = p->shells[x.y];
            const ShellInfo &b0 = p->shells[y.x], &b1 = p->shells[y.y];
            return a0.nprim * a1.nprim * a0.nctr * a1.nctr > b0.nprim * b1.nprim * b0.nctr * b1.nctr;
        });
    return cls;
}

int upload_classes(std::vector<PairClass>& cls)
{
    for (PairClass& c : cls) {
        HIP_TRY(hipMalloc(&c.d_pairs, sizeof(int2) * c.pairs.size()));
        HIP_TRY(hipMemcpy(c.d_pairs, c.pairs.data(), sizeof(int2) * c.pairs.size(), hipMemcpyHostToDevice));
    }
    return CCSD_OK;
}

void free_classes(std::vector<PairClass>& cls)
{
    for (PairClass& c : cls)
        if (c.d_pairs) (void)hipFree(c.d_pairs);
    cls.clear();
}

int get_bra(ccsd_plan* p, int ish0, int ish1, std::vector<PairClass>** out)
{
    auto key = std::make_pair(ish0, ish1);
    auto it = p->bra_cache.find(key);
    if (it == p->bra_cache.end()) {
        std::vector<int2> pairs;
        for (int i = ish0; i < ish1; i++)
            for (int j = 0; j < p->nbas; j++) pairs.push_back(make_int2(i, j));
        std::vector<PairClass> cls = make_classes(p, pairs);
        int rc = upload_classes(cls);
        if (rc) return rc;
        it = p->bra_cache.emplace(key, std::move(cls)).first;
    }
    *out = &it->second;
    return CCSD_OK;
}

int upload_geometry(ccsd_plan* p)
{
    std::vector<double> coords(3 * p->natm);
    for (int a = 0; a < p->natm; a++) {
        const int ptr = p->atm[a * CCSD_ATM_SLOTS + CCSD_PTR_COORD];
        if (ptr < 0 || ptr + 3 > p->nenv) return fail(CCSD_ERR_SHAPE, "atm PTR_COORD outside env");
        for (int x = 0; x < 3; x++) coords[3 * a + x] = p->env[ptr + x];
    }
    HIP_TRY(hipMemcpy(p->d_env, p->env.data(), sizeof(double) * p->nenv, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_coords, coords.data(), sizeof(double) * 3 * p->natm, hipMemcpyHostToDevice));
    std::vector<double> charges(p->natm);
    for (int a = 0; a < p->natm; a++) charges[a] = (double)p->atm[a * CCSD_ATM_SLOTS + 0];
    HIP_TRY(hipMemcpy(p->d_charges, charges.data(), sizeof(double) * p->natm, hipMemcpyHostToDevice));
    return CCSD_OK;
}

// Launch tuning of the generic quartet kernel for one class.
void tune_class(EriClassArgs& a)
{
    const int d = a.deriv;
    const int ncomp = d ? 3 : 1;
    const int sj = a.li + 1 + d, sk = sj * (a.lj + 1), sl = sk * (a.lk + 1);
    const int gsize = sl * (a.ll + 1);
    const int nf = ncart(a.li) * ncart(a.lj) * ncart(a.lk) * ncart(a.ll);
    const int nunits = nf * a.max_ncij * a.max_nckl;
    const int npq = a.max_nij * a.max_nkl;
    a.pq_batch = std::max(1, std::min(64, npq));
    const int g_budget = 1536;  // doubles for the 1-D factors
    a.item_batch = std::max(1, std::min(a.pq_batch * a.nroots, g_budget / (3 * gsize)));
    int O = 1;
    while (O < 64 && O < nunits) O <<= 1;
    a.groups = 64 / O;
    const int red_budget = 1024;  // doubles for the lane-owned partial sums
    a.units_pass = std::max(1, std::min(nunits, red_budget / (a.groups * ncomp)));
}

int run_classes(ccsd_plan* p, std::vector<PairClass>& bra, int deriv, double* out, int i_base,
                long long comp_stride, const double* gamma, double* de, double scale, hipStream_t stream)
{
    const long long npair = (long long)p->nao_cart * (p->nao_cart + 1) / 2;
    for (PairClass& bc : bra) {
        for (PairClass& kc : p->ket) {
            EriClassArgs a;
            memset(&a, 0, sizeof(a));
            a.shells = p->d_shells;
            a.env = p->d_env;
            a.coords = p->d_coords;
            a.bra = bc.d_pairs;
            a.ket = kc.d_pairs;
            a.nbra = (int)bc.pairs.size();
            a.nket = (int)kc.pairs.size();
            a.li = bc.l1; a.lj = bc.l2; a.lk = kc.l1; a.ll = kc.l2;
            a.deriv = deriv;
            a.nroots = (a.li + a.lj + a.lk + a.ll + deriv) / 2 + 1;
            a.max_nij = bc.max_np;
            a.max_nkl = kc.max_np;
            a.max_ncij = bc.max_nc;
            a.max_nckl = kc.max_nc;
            tune_class(a);
            a.out = out;
            a.i_base = i_base;
            a.nao = p->nao_cart;
            a.npair = npair;
            a.comp_stride = comp_stride;
            a.gamma = gamma;
            a.de = de;
            a.gamma_scale = scale;
            const long long nblk = (long long)a.nbra * a.nket;
            if (nblk > 2147483647LL) return fail(CCSD_ERR_SHAPE, "too many quartets in one class launch");
            if (!out && p->fused_reg && a.nbra <= 65535 &&
                ccsd_fused_reg_supported(a.li, a.lj, a.lk, a.ll, a.max_nckl)) {
                hipError_t e = ccsd_launch_fused_reg(a, stream);
                if (e != hipSuccess) return hip_fail(e, "fused_ip1_reg_kernel launch");
                continue;
            }
            if (ccsd_eri_lds_bytes(a) > 65536) return fail(CCSD_ERR_UNSUPPORTED, "class needs more than 64 KiB LDS");
            hipError_t e = ccsd_launch_eri_class(a, stream);
            if (e != hipSuccess) return hip_fail(e, "eri_quartet_kernel launch");
        }
    }
    return CCSD_OK;
}

int check_plan(const ccsd_plan* p)
{
    if (!p) return fail(CCSD_ERR_NULL, "null plan");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->device)
        return fail(CCSD_ERR_SHAPE, "plan was created on another device");
    return CCSD_OK;
}

int eri_s2kl(ccsd_plan* p, double* out, int ish0, int ish1, int cart, int deriv, hipStream_t stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out) return fail(CCSD_ERR_NULL, "null output pointer");
    if (ish0 < 0 || ish1 > p->nbas || ish0 >= ish1) return fail(CCSD_ERR_SHAPE, "bad shell range");
    if (!cart && !p->has_c2s) return fail(CCSD_ERR_UNSUPPORTED, "plan has no cart2sph coefficients");
    std::vector<PairClass>* bra = nullptr;
    rc = get_bra(p, ish0, ish1, &bra);
    if (rc) return rc;
    const int ncomp = deriv ? 3 : 1;
    const int i0c = p->shells[ish0].ao_cart;
    const int i1c = ish1 < p->nbas ? p->shells[ish1].ao_cart : p->nao_cart;
    const int nfi_c = i1c - i0c;
    const long long npair_c = (long long)p->nao_cart * (p->nao_cart + 1) / 2;
    const long long comp_stride_c = (long long)nfi_c * p->nao_cart * npair_c;
    double* cart_buf = out;
    if (!cart) {
        HIP_TRY(hipMallocAsync((void**)&cart_buf, sizeof(double) * ncomp * comp_stride_c, stream));
    }
    rc = run_classes(p, *bra, deriv, cart_buf, i0c, comp_stride_c, nullptr, nullptr, 0.0, stream);
    if (!rc && !cart) {
        const int i0s = p->shells[ish0].ao_sph;
        const int i1s = ish1 < p->nbas ? p->shells[ish1].ao_sph : p->nao_sph;
        const long long npair_s = (long long)p->nao_sph * (p->nao_sph + 1) / 2;
        // i-block terms re-indexed relative to the block
        std::vector<C2sTermHost> ti(p->c2s.begin() + i0s, p->c2s.begin() + i1s);
        for (auto& t : ti)
            for (int m = 0; m < t.n; m++) t.idx[m] -= i0c;
        void* d_ti = nullptr;
        hipError_t e = hipMallocAsync(&d_ti, sizeof(C2sTermHost) * ti.size(), stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_ti, ti.data(), sizeof(C2sTermHost) * ti.size(), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess)
            e = ccsd_launch_c2s(out, cart_buf, d_ti, p->d_c2s, i1s - i0s, p->nao_sph, npair_s, p->nao_cart, npair_c,
                                nfi_c, ncomp, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);  // ti is host-side and temporary
        if (d_ti) (void)hipFreeAsync(d_ti, stream);
        if (e != hipSuccess) rc = hip_fail(e, "cart2sph");
    }
    if (!cart) (void)hipFreeAsync(cart_buf, stream);
    return rc;
}

}  // namespace

extern "C" {

const char* ccsd_last_error(void) { return g_last_error.c_str(); }

int ccsd_plan_create(ccsd_plan** out, const int* atm, int natm, const int* bas, int nbas, const double* env, int nenv,
                     const double* c2s)
{
    if (!out || !atm || !bas || !env) return fail(CCSD_ERR_NULL, "null argument to ccsd_plan_create");
    if (natm <= 0 || nbas <= 0 || nenv <= 0) return fail(CCSD_ERR_SHAPE, "empty molecule");
    *out = nullptr;
    ccsd_plan* p = new ccsd_plan();
    p->natm = natm;
    p->nbas = nbas;
    p->nenv = nenv;
    p->atm.assign(atm, atm + natm * CCSD_ATM_SLOTS);
    p->bas.assign(bas, bas + nbas * CCSD_BAS_SLOTS);
    p->env.assign(env, env + nenv);
    int aoc = 0, aos = 0;
    for (int i = 0; i < nbas; i++) {
        const int* b = bas + i * CCSD_BAS_SLOTS;
        ShellInfo s;
        s.l = b[CCSD_ANG_OF];
        s.nprim = b[CCSD_NPRIM_OF];
        s.nctr = b[CCSD_NCTR_OF];
        s.atom = b[CCSD_ATOM_OF];
        s.pexp = b[CCSD_PTR_EXP];
        s.pcoef = b[CCSD_PTR_COEFF];
        s.ao_cart = aoc;
        s.ao_sph = aos;
        if (s.l < 0 || s.l > CCSD_LMAX) {
            delete p;
            return fail(CCSD_ERR_UNSUPPORTED, "angular momentum above 4");
        }
        if (s.nprim <= 0 || s.nctr <= 0 || s.atom < 0 || s.atom >= natm || s.pexp < 0 ||
            s.pexp + s.nprim > nenv || s.pcoef < 0 || s.pcoef + s.nprim * s.nctr > nenv) {
            delete p;
            return fail(CCSD_ERR_SHAPE, "bas row points outside env/atm");
        }
        if (s.nprim * s.nctr > 4096 || s.nprim > 64) {
            delete p;
            return fail(CCSD_ERR_UNSUPPORTED, "shell with more than 64 primitives");
        }
        aoc += ncart(s.l) * s.nctr;
        aos += nsph(s.l) * s.nctr;
        p->lmax = std::max(p->lmax, s.l);
        p->shells.push_back(s);
    }
    p->nao_cart = aoc;
    p->nao_sph = aos;
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return fail(CCSD_ERR_HIP, "hipGetDevice failed");
    }
    int rc = CCSD_OK;
    if (hipMalloc(&p->d_shells, sizeof(ShellInfo) * nbas) != hipSuccess ||
        hipMalloc(&p->d_env, sizeof(double) * nenv) != hipSuccess ||
        hipMalloc(&p->d_coords, sizeof(double) * 3 * natm) != hipSuccess ||
        hipMalloc(&p->d_charges, sizeof(double) * natm) != hipSuccess) {
        ccsd_plan_destroy(p);
        return fail(CCSD_ERR_ALLOC, "device allocation for the basis plan failed");
    }
    if (hipMemcpy(p->d_shells, p->shells.data(), sizeof(ShellInfo) * nbas, hipMemcpyHostToDevice) != hipSuccess) {
        ccsd_plan_destroy(p);
        return fail(CCSD_ERR_HIP, "upload of shells failed");
    }
    rc = upload_geometry(p);
    if (rc) {
        std::string msg = g_last_error;
        ccsd_plan_destroy(p);
        return fail(rc, msg);
    }
    std::vector<int2> kpairs;
    for (int k = 0; k < nbas; k++)
        for (int l = 0; l <= k; l++) kpairs.push_back(make_int2(k, l));
    p->ket = make_classes(p, kpairs);
    rc = upload_classes(p->ket);
    if (rc) {
        std::string msg = g_last_error;
        ccsd_plan_destroy(p);
        return fail(rc, msg);
    }
    if (c2s) {
        p->has_c2s = true;
        p->c2s.resize(aos);
        for (int s = 0; s < aos; s++) {
            C2sTermHost t;
            memset(&t, 0, sizeof(t));
            for (int c = 0; c < aoc; c++) {
                const double v = c2s[(size_t)c * aos + s];
                if (v != 0.0) {
                    if (t.n == 6) {
                        ccsd_plan_destroy(p);
                        return fail(CCSD_ERR_UNSUPPORTED, "cart2sph column with more than 6 terms");
                    }
                    t.idx[t.n] = c;
                    t.c[t.n] = v;
                    t.n++;
                }
            }
            p->c2s[s] = t;
        }
        if (sizeof(C2sTermHost) != ccsd_c2s_term_size() ||
            hipMalloc(&p->d_c2s, sizeof(C2sTermHost) * aos) != hipSuccess ||
            hipMemcpy(p->d_c2s, p->c2s.data(), sizeof(C2sTermHost) * aos, hipMemcpyHostToDevice) != hipSuccess) {
            ccsd_plan_destroy(p);
            return fail(CCSD_ERR_HIP, "upload of cart2sph terms failed");
        }
    }
    *out = p;
    return CCSD_OK;
}

int ccsd_plan_destroy(ccsd_plan* p)
{
    if (!p) return CCSD_OK;
    free_classes(p->ket);
    for (auto& kv : p->bra_cache) free_classes(kv.second);
    if (p->d_shells) (void)hipFree(p->d_shells);
    if (p->d_env) (void)hipFree(p->d_env);
    if (p->d_coords) (void)hipFree(p->d_coords);
    if (p->d_charges) (void)hipFree(p->d_charges);
    if (p->d_c2s) (void)hipFree(p->d_c2s);
    delete p;
    return CCSD_OK;
}

int ccsd_plan_set_env(ccsd_plan* p, const double* env, int nenv)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!env) return fail(CCSD_ERR_NULL, "null env");
    if (nenv != p->nenv) return fail(CCSD_ERR_SHAPE, "env length differs from the plan's");
    p->env.assign(env, env + nenv);
    return upload_geometry(p);
}

int ccsd_plan_nao(const ccsd_plan* p, int cart)
{
    if (!p) return fail(CCSD_ERR_NULL, "null plan");
    return cart ? p->nao_cart : p->nao_sph;
}

int ccsd_int2e_s2kl(ccsd_plan* p, double* out, int ish0, int ish1, int cart, void* stream)
{
    return eri_s2kl(p, out, ish0, ish1, cart, 0, (hipStream_t)stream);
}

int ccsd_int2e_ip1_s2kl(ccsd_plan* p, double* out, int ish0, int ish1, int cart, void* stream)
{
    return eri_s2kl(p, out, ish0, ish1, cart, 1, (hipStream_t)stream);
}

int ccsd_grad_ip1_gamma(ccsd_plan* p, double* de, const double* gamma, double scale, int ish0, int ish1,
                        void* stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!de || !gamma) return fail(CCSD_ERR_NULL, "null de or gamma");
    if (ish0 < 0 || ish1 > p->nbas || ish0 >= ish1) return fail(CCSD_ERR_SHAPE, "bad shell range");
    std::vector<PairClass>* bra = nullptr;
    rc = get_bra(p, ish0, ish1, &bra);
    if (rc) return rc;
    return run_classes(p, *bra, 1, nullptr, 0, 0, gamma, de, scale, (hipStream_t)stream);
}

int ccsd_int1e(ccsd_plan* p, int kind, double* out, void* stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out) return fail(CCSD_ERR_NULL, "null output pointer");
    if (kind < 0 || kind > 5) return fail(CCSD_ERR_SHAPE, "int1e kind must be 0..5");
    const long long nn = (long long)p->nao_cart * p->nao_cart;
    const long long nmat = kind <= 2 ? 1 : (kind <= 4 ? 3 : 3LL * p->natm);
    hipError_t e = ccsd_launch_fill(out, nmat * nn, 0.0, (hipStream_t)stream);
    if (e == hipSuccess)
        e = ccsd_launch_int1e(kind, out, p->d_shells, p->d_env, p->d_coords, p->d_charges, p->natm, p->nbas,
                              p->nao_cart, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "int1e_kernel");
    return CCSD_OK;
}

int ccsd_gemm_mixed(int mode, const double* A, const double* B, double* C, int M, int N, int K, long long lda,
                    long long ldb, long long ldc, long long strideA, long long strideB, long long strideC, int batch,
                    void* stream)
{
    if (!A || !B || !C) return fail(CCSD_ERR_NULL, "null matrix pointer");
    if (mode < 0 || mode > 1) return fail(CCSD_ERR_SHAPE, "mode must be 0 (fp32) or 1 (bf16x3)");
    if (M < 0 || N < 0 || K < 0 || batch < 0 || lda < K || ldb < N || ldc < N)
        return fail(CCSD_ERR_SHAPE, "bad GEMM dimensions");
    if ((M + 63) / 64 > 65535 || batch > 65535) return fail(CCSD_ERR_SHAPE, "M or batch too large for one launch");
    hipError_t e = ccsd_launch_gemm_mixed(mode, A, B, C, M, N, K, lda, ldb, ldc, strideA, strideB, strideC, batch,
                                          (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "gemm_mixed_kernel");
    return CCSD_OK;
}

int ccsd_split_bf16(const double* x, void* hi, void* lo, long long rows, long long cols, long long ld, void* stream)
{
    if (!x || !hi || !lo) return fail(CCSD_ERR_NULL, "null pointer");
    if (rows < 0 || cols < 0 || ld < cols) return fail(CCSD_ERR_SHAPE, "bad split shape");
    hipError_t e = ccsd_launch_split_bf16(x, hi, lo, rows, cols, ld, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "split_bf16_kernel");
    return CCSD_OK;
}

int ccsd_gemm_bf16x3_nt(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C, int M, int N,
                        int K, long long lda, long long ldb, long long strideA, long long strideB, long long strideC,
                        long long ldc, int batch, int beta, void* stream)
{
    if (!Ahi || !Alo || !Bhi || !Blo || !C) return fail(CCSD_ERR_NULL, "null pointer");
    if (M < 0 || N < 0 || K < 0 || batch < 0 || lda < K || ldb < K || ldc < N || (beta != 0 && beta != 1))
        return fail(CCSD_ERR_SHAPE, "bad GEMM arguments");
    if ((N + 127) / 128 > 65535 || batch > 65535) return fail(CCSD_ERR_SHAPE, "N or batch too large for one launch");
    hipError_t e = ccsd_launch_gemm_bf16x3_nt(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, strideA, strideB, strideC, ldc,
                                              batch, beta, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "gemm_bf16x3_nt_kernel");
    return CCSD_OK;
}

int ccsd_plan_set_option(ccsd_plan* p, int option, int value)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (option == 0) {
        p->fused_reg = value != 0;
        return CCSD_OK;
    }
    return fail(CCSD_ERR_SHAPE, "unknown option");
}

int ccsd_rys_roots(int n, const double* T, int count, double* r, double* w, void* stream)
{
    if (!T || !r || !w) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 1 || n > CCSD_NROOTS_MAX || count < 0) return fail(CCSD_ERR_SHAPE, "n must be 1..9");
    hipError_t e = ccsd_launch_rys(n, T, count, r, w, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rys_roots_kernel");
    return CCSD_OK;
}

}  // extern "C"
