// C-ABI of the ERI library: basis plans, class enumeration and launches.
//
// The reference launches one CUDA kernel per decontracted i-shell with the
// raw libcint arrays (ccsd_gpu/cuda/int2e_ip1_gpu.py:148-216). Here a plan
// holds the basis once in device memory; every call enumerates the bra shell
// pairs (ish in [ish0, ish1), all jsh) and ket pairs (ksh >= lsh) by angular
// momentum class and launches one quartet kernel per (bra class, ket class).
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "eri_common.h"
#include "../../include/ccsd_gpu_amd.h"

size_t ccsd_eri_lds_bytes(const EriClassArgs& a);
hipError_t ccsd_launch_eri_class(const EriClassArgs& a, hipStream_t stream);
int ccsd_persistent_blocks_per_cu(size_t lds);
int ccsd_reg_persist_parts(int li, int lj, int lk, int ll, int max_ncij, int max_nckl);
size_t ccsd_reg_persist_lds(int max_nij);
int ccsd_reg_persist_blocks_per_cu(size_t lds);
hipError_t ccsd_launch_reg_persistent(EriClassArgs* classes, int nclasses, const int4* items, int nitems,
                                      unsigned int* counter, unsigned int* ticks, const double* gamma, const double* dm,
                                      const double* pm, unsigned long long* defx, double scale, int gamma_mode,
                                      int nblocks, size_t lds, hipStream_t stream);
hipError_t ccsd_launch_eri_persistent(const EriClassArgs* classes, const int4* items, int nitems,
                                      unsigned int* counter, unsigned int* ticks, const double* gamma,
                                      const double* dm, const double* pm, unsigned long long* defx, double scale,
                                      int gamma_mode, int nblocks, size_t lds, hipStream_t stream);
hipError_t ccsd_launch_defx_to_de(const unsigned long long* defx, int natm, double* de, hipStream_t stream);
hipError_t ccsd_launch_c2s(double* out, const double* in, const void* ti, const void* tall, int nfi_s, int nao_s,
                           long long npair_s, int nao_c, long long npair_c, int nfi_c, int ncomp, hipStream_t stream);
hipError_t ccsd_launch_rys(int n, const double* T, int count, double* r, double* w, hipStream_t stream);
size_t ccsd_c2s_term_size();
hipError_t ccsd_launch_fill(double* p, long long n, double v, hipStream_t stream);
int ccsd_fused_reg_parts(int li, int lj, int lk, int ll, int max_ncij, int max_nckl);
int ccsd_store_reg_parts(int li, int lj, int lk, int ll, int max_ncij, int max_nckl, int deriv);
hipError_t ccsd_launch_store_reg(const EriClassArgs& a, int wpb, hipStream_t stream);
size_t ccsd_fused_reg_lds(int max_nij, int wpb);
hipError_t ccsd_launch_fused_reg(const EriClassArgs& a, int waves_per_block, hipStream_t stream);
extern int g_gemm_variant;
hipError_t ccsd_launch_gemm_bf16x3_nt(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C,
                                      int M, int N, int K, long long lda, long long ldb, long long sA, long long sB,
                                      long long sC, long long ldc, int batch, int beta, double alpha,
                                      const long long* off_b, const long long* off_r, const long long* off_c,
                                      hipStream_t stream);
hipError_t ccsd_launch_split_bf16_strided(const double* x, void* hi, void* lo, int ndim, const long long* sizes,
                                          const long long* strides, hipStream_t stream);
hipError_t ccsd_launch_scatter_sum(const double* ws, int S, int nb, int M, int N, double* C, const long long* off_b,
                                   const long long* off_r, const long long* off_c, double alpha, int beta,
                                   hipStream_t stream);
hipError_t ccsd_launch_split_bf16(const double* x, void* hi, void* lo, long long rows, long long cols, long long ld,
                                  hipStream_t stream);
hipError_t ccsd_launch_sum_slices(double* ws, double* out, long long n, int S, int beta, hipStream_t stream);
hipError_t ccsd_launch_sym_pairs(double* t, int n, double scale, hipStream_t stream);
hipError_t ccsd_launch_repack_pairs(const void* z, int n, int m, void* out, int f32, hipStream_t stream);
hipError_t ccsd_launch_add_separable_packed(double* g, const double* D, const double* P, int n, hipStream_t stream);
hipError_t ccsd_launch_pack_pairs_sym(const double* t, const long long* kk, const long long* ll, int n, int m,
                                      long long npair, double* out, hipStream_t stream);
size_t ccsd_l1_terms_lds(int v);
hipError_t ccsd_launch_vvvv_pm_pack(const double* l2, int o, int v, const long long* iu, const long long* ju,
                                    long long npair, const long long* tc, const long long* td, long long np,
                                    long long ld, int split, void* out, void* lo, hipStream_t stream);
hipError_t ccsd_launch_vvvv_pm_unpack(const double* X, int o, int v, long long ldx, double scale, double* out,
                                      hipStream_t stream);
hipError_t ccsd_launch_unpack_pairs(const double* X, const long long* iu, const long long* ju, int o, int v,
                                    long long npair, double s, double* out, hipStream_t stream);
hipError_t ccsd_launch_permute(const double* src, int ndim, const long long* sizes, const long long* sstrides,
                               double* dst, void* hi, void* lo, hipStream_t stream);
hipError_t ccsd_launch_gemv(int cols, const double* A, long long M, long long K, long long lda, const double* x,
                            long long incx, double* y, double alpha, int beta, double* ws, int S, hipStream_t stream);
hipError_t ccsd_launch_combine(int ndim, const long long* sizes, double* dst, const long long* dstrides, int nt,
                               const double* const* srcs, const long long* sstrides, const double* coef, int beta,
                               hipStream_t stream);
hipError_t ccsd_launch_l1_terms(const double* X, int o, int v, const double* y, long long sy0, long long sy1,
                                double c1, double c2, double c3, double c4, double* ws, double* out,
                                hipStream_t stream);
hipError_t ccsd_launch_l2_finish(const double* X, int o, int v, const double* u, const double* w, const double* S,
                                 const double* sp, double sh, const double* D, double* out, hipStream_t stream);
int ccsd_diis_max_rows();
long long ccsd_diis_push_blocks(long long n);
hipError_t ccsd_launch_diis_push(const double* s0, const double* s1, long long n0, long long n, const double* rhs,
                                 const double* prev, float* xs, float* es, const float* E, long long lde, int nrows,
                                 int slot, double* prev_out, double* dots, double* ws, hipStream_t stream);
hipError_t ccsd_launch_diis_combine(const float* X, long long ldx, int nrows, const double* coef, const double* base,
                                    double* out, long long n, hipStream_t stream);
hipError_t ccsd_launch_pair_combine(const double* X, const long long* xs, int o, int v, double c1, double c2,
                                    double c3, double c4, const double* D, double* out, const long long* os, int beta,
                                    hipStream_t stream);
hipError_t ccsd_launch_gemm_mixed(int mode, const double* A, const double* B, double* C, int M, int N, int K,
                                  long long lda, long long ldb, long long ldc, long long sA, long long sB,
                                  long long sC, int batch, hipStream_t stream);
hipError_t ccsd_launch_int1e(int kind, unsigned long long* fx, const ShellInfo* shells, const double* env,
                             const double* coords, const double* charges, int natm, int nbas, int nao, int maxprim,
                             hipStream_t stream);
hipError_t ccsd_launch_int1e_convert(const unsigned long long* fx, long long n, double* out, hipStream_t stream);

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

// Slice partials of ccsd_gemv_noalloc: a file-scope device array instead of an
// allocation (callers under graph capture or in allocation-free loops). Each device
// holds its own copy; its address is taken with hipGetSymbolAddress once per device,
// on that device, and only that address is handed to kernels.
static __device__ __align__(256) double scratch[CCSD_GEMV_SCRATCH_DOUBLES];
std::mutex g_scratch_mu;
double* g_scratch_dev[CCSD_MAX_DEVICES] = {};

int scratch_address(double** out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= CCSD_MAX_DEVICES) return fail(CCSD_ERR_UNSUPPORTED, "device ordinal beyond CCSD_MAX_DEVICES");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CCSD_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return CCSD_OK;
}

struct C2sTermHost {
    int n;
    int idx[6];
    double c[6];
};

// Shell pairs of one angular-momentum class, grouped into segments: a segment
// is a run of pairs of one molecule with the same primitive x contraction
// product (lanes of a wave then loop alike), ordered by decreasing Schwarz
// factor when screening is on, so that the survivors for any bra pair are a
// prefix of each segment.
struct PairClass {
    int l1, l2;
    std::vector<int2> pairs;
    std::vector<int> seg;          // segment s = pairs[seg[s], seg[s + 1])
    std::vector<int> seg_mol;
    std::vector<int> seg_eff;      // surviving primitive pairs of the segment's pairs
    int max_np = 0, max_nc = 0;
    int2* d_pairs = nullptr;
    // ket classes: the surviving primitive pairs of every pair (kKetFields doubles each),
    // pair t's entries at [pp_off[t], pp_off[t + 1])
    int* d_pp_off = nullptr;
    double* d_pp = nullptr;
};

// fields of a ket primitive-pair entry: a_kl, Q (3), K_cd / a_kl, 0.5 / a_kl, c_k c_l (4)
constexpr int kKetFields = 10;
// primitive pairs with a_1 a_2 / (a_1 + a_2) |R_12|^2 above this are not evaluated
// (Gaussian-product factor below e^-50); the same cut as the register kernels' bra side
constexpr double kPrimExpCut = 50.0;

// Device work list of one (bra class, ket class) launch.
struct ChunkList {
    int4* d = nullptr;
    int n = 0;
    long long quartets = 0;
    std::vector<double> cumq;      // quartet equivalents in the first c entries (autotuner samples)
    std::vector<float> cost;       // estimated cost of each entry (primitive products x quartets)
};

// Kernel choice of one class in the fused gradient.
struct Choice {
    int kind = -1;   // 0 generic LDS kernel, 1 register kernel
    int wpb = 4;     // waves per block (register kernel)
    int chunk = 64;  // ket pairs per chunk; bits 16+: primitive split of the register kernel (see get_chunks)
    // measured time per unit of the work list's cost (primitive products x quartets) of
    // the chosen kernel, 0 when it was not timed: converts the within-class cost model
    // into time across classes (a d-shell quartet costs far more per primitive product
    // than an s-shell one)
    double ms_per_cost = 0.0;
};

// One persistent launch: the launch description of every class it runs (work lists
// included) and its work items (runs of one class's units), drawn through a counter.
struct PersistPart {
    EriClassArgs* d_cls = nullptr;
    int4* d_items = nullptr;
    unsigned int* d_counter = nullptr;   // 8 bytes, cleared by a fill kernel before each launch
    unsigned int* d_ticks = nullptr;     // per-item durations of a calibration launch (register part)
    int ncls = 0, nitems = 0, nblocks = 0;
    size_t lds = 0;
    // what the item cut reads: each class's work list, units per chunk and weight (time
    // per cost unit), the order mode; `calibrated` once the weights come from a launch
    std::vector<const ChunkList*> lists;
    std::vector<int> parts;
    std::vector<double> wt;
    std::vector<int4> items;
    double unit_cost = 0.0;   // fixed time per unit (calibrated), in the weights' units
    double total = 0.0;       // estimated wave-time of all items, in the weights' units
    bool grouped = false;
    int launches = 0;
    bool calibrated = false;
};

// Longest run of units in one work item: every unit carries a fixed cost (its bra table,
// dependent loads and gradient atomics) that the cost model does not see, so runs of the
// cheapest chunks at the end of a class list must not grow without bound (a cut by cost
// alone put 40 000 of them into single items that ran for over a second).
constexpr int kMaxItemUnits = 256;

// The whole fused pass of a bra set as persistent launches: `reg` runs the
// register-resident bodies (plan option 5 = 1) of every class that has one, `gen` the
// LDS kernel's generic body for the others (every class when option 5 = 2).
struct PersistPlan {
    PersistPart reg, gen;
    long long quartets = 0;
    std::vector<std::array<long long, 10>> launches;
};

void free_persist_part(PersistPart& q)
{
    if (q.d_cls) (void)hipFree(q.d_cls);
    if (q.d_items) (void)hipFree(q.d_items);
    if (q.d_counter) (void)hipFree(q.d_counter);
    if (q.d_ticks) (void)hipFree(q.d_ticks);
    q = PersistPart();
}

struct BraSet {
    std::vector<PairClass> cls;
    std::map<std::tuple<int, int, int, int>, ChunkList> chunks;  // (bc, kc, chunk, screened)
    std::map<int, PersistPlan> persist;                          // by screened
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
    // molecules (one unless created by ccsd_plan_create_batch)
    int nmol = 1;
    std::vector<int> mol_shell0;                     // nmol + 1
    std::vector<int> mol_ao0, mol_nao;               // Cartesian
    int* d_mol_ao0 = nullptr;
    int* d_mol_nao = nullptr;
    long long* d_mol_dm_off = nullptr;
    std::vector<PairClass> ket;                      // ksh >= lsh, same molecule
    std::map<std::pair<int, int>, BraSet> bra_cache;
    // fused-gradient kernel choice per (ish0, ish1, bra class, ket class); class
    // numbering does not depend on the geometry or the screening threshold
    std::map<std::tuple<int, int, int, int>, Choice> choice;
    std::map<std::tuple<int, int, int, int>, Choice> choice_lds;   // generic kernel only (persistent pass)
    // fused pass as one persistent launch: 0 off, 1 register bodies (when every class of
    // the pass has one; the generic body otherwise), 2 the generic body
    int persistent = 0;
    int persist_order = 0;                           // register items: 0 class by class, 1 by cost only
    int n_cu = 0;                                    // compute units of the plan's device
    bool has_c2s = false;
    bool fused_reg = true;                           // register kernels (fused gradient and s2kl store)
    bool autotune = true;                            // time the kernel choices of each class once
    // Schwarz screening of the fused gradient
    double thr = 0.0;
    bool q_valid = false;
    std::vector<double> q;                           // nbas x nbas, symmetric
    long long last_quartets = 0;
    std::vector<std::array<long long, 10>> last_launches;   // see ccsd_plan_last_launches
    unsigned long long* d_de_scratch = nullptr;      // autotuner's throw-away gradient (fixed point)
    unsigned long long* d_defx = nullptr;            // fixed-point gradient of a fused pass (1 + 6 natm words)
    unsigned long long* d_int1e_fx = nullptr;        // fixed-point one-electron integrals of the last ccsd_int1e
    long long int1e_fx_words = 0;
    std::vector<double> coords_h;                    // natm x 3
    // side streams (default priority) that run independent class launches of a fused pass
    // concurrently; forked from and joined back to the caller's stream with events
    static constexpr int kSideStreams = 4;
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kSideStreams] = {};
    bool concurrent = true;
    bool ket_shell_order = false;                    // unscreened kets in shell order inside a primitive group
    int force_split = 0;                             // primitive split of the register kernels (0: by wave count)
    bool tune_generic = false;                       // time the generic kernel for register-capable classes too
    std::vector<int> eff;                            // surviving primitive pairs of same-molecule shell pairs
    std::vector<long long> mol_eff_off;
    std::vector<C2sTermHost> c2s;                    // per spherical AO
    void* d_c2s = nullptr;
};

namespace {

int pair_eff(const ccsd_plan* p, int i, int j)
{
    const int m = p->shells[i].mol, s0 = p->mol_shell0[m], n = p->mol_shell0[m + 1] - s0;
    return p->eff[p->mol_eff_off[m] + (long long)(i - s0) * n + (j - s0)];
}

// Primitive pairs of (i, j) past the Gaussian-product cut, for the plan's geometry.
void compute_eff(ccsd_plan* p)
{
    p->mol_eff_off.assign(p->nmol + 1, 0);
    for (int m = 0; m < p->nmol; m++) {
        const long long n = p->mol_shell0[m + 1] - p->mol_shell0[m];
        p->mol_eff_off[m + 1] = p->mol_eff_off[m] + n * n;
    }
    p->eff.assign(p->mol_eff_off[p->nmol], 0);
    for (int m = 0; m < p->nmol; m++) {
        const int s0 = p->mol_shell0[m], s1 = p->mol_shell0[m + 1], n = s1 - s0;
        for (int i = s0; i < s1; i++)
            for (int j = s0; j < s1; j++) {
                const ShellInfo &a = p->shells[i], &b = p->shells[j];
                const double* A = &p->coords_h[3 * a.atom];
                const double* B = &p->coords_h[3 * b.atom];
                const double r2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) +
                                  (A[2] - B[2]) * (A[2] - B[2]);
                int c = 0;
                for (int u = 0; u < a.nprim; u++)
                    for (int v = 0; v < b.nprim; v++) {
                        const double x = p->env[a.pexp + u], y = p->env[b.pexp + v];
                        if (x * y / (x + y) * r2 <= kPrimExpCut) c++;
                    }
                p->eff[p->mol_eff_off[m] + (long long)(i - s0) * n + (j - s0)] = c;
            }
    }
}

// Grouping key of a pair inside a class: surviving primitive pairs in power-of-two
// buckets (0, 1, 2-3, 4-7, ...) and contraction columns, so that the lanes of a wave
// loop about alike without cutting the lists into many short segments.
int prim_group(const ccsd_plan* p, const int2& pr)
{
    const ShellInfo &a = p->shells[pr.x], &b = p->shells[pr.y];
    int e = pair_eff(p, pr.x, pr.y), bucket = 0;
    while (e) {
        bucket++;
        e >>= 1;
    }
    return bucket * 16 + a.nctr * b.nctr;
}

double pair_q(const ccsd_plan* p, const int2& pr) { return p->q[(size_t)pr.x * p->nbas + pr.y]; }

// Classes by (l1, l2, several contraction columns or not); inside a class:
// molecule, then primitive group (descending), then Schwarz factor (descending)
// when `by_q`, else surviving primitive pairs (descending) or, with
// `shell_order`, the input (shell) order, which keeps the packed kl index of
// neighbouring lanes close.
std::vector<PairClass> make_classes(const ccsd_plan* p, const std::vector<int2>& pairs, bool by_q,
                                    bool shell_order = false)
{
    std::vector<PairClass> cls;
    std::map<std::tuple<int, int, int>, int> where;
    for (const int2& pr : pairs) {
        const ShellInfo &a = p->shells[pr.x], &b = p->shells[pr.y];
        auto key = std::make_tuple(a.l, b.l, a.nctr * b.nctr > 1 ? 1 : 0);
        auto it = where.find(key);
        if (it == where.end()) {
            it = where.emplace(key, (int)cls.size()).first;
            PairClass c;
            c.l1 = a.l;
            c.l2 = b.l;
            cls.push_back(c);
        }
        PairClass& c = cls[it->second];
        c.pairs.push_back(pr);
        c.max_np = std::max(c.max_np, a.nprim * b.nprim);
        c.max_nc = std::max(c.max_nc, a.nctr * b.nctr);
    }
    for (PairClass& c : cls) {
        std::stable_sort(c.pairs.begin(), c.pairs.end(), [p, by_q, shell_order](const int2& x, const int2& y) {
            const int mx = p->shells[x.x].mol, my = p->shells[y.x].mol;
            if (mx != my) return mx < my;
            const int gx = prim_group(p, x), gy = prim_group(p, y);
            if (gx != gy) return gx > gy;
            if (by_q) return pair_q(p, x) > pair_q(p, y);
            if (shell_order) return false;
            return pair_eff(p, x.x, x.y) > pair_eff(p, y.x, y.y);
        });
        for (size_t t = 0; t < c.pairs.size(); t++) {
            if (t == 0 || p->shells[c.pairs[t].x].mol != p->shells[c.pairs[t - 1].x].mol ||
                prim_group(p, c.pairs[t]) != prim_group(p, c.pairs[t - 1])) {
                c.seg.push_back((int)t);
                c.seg_mol.push_back(p->shells[c.pairs[t].x].mol);
                c.seg_eff.push_back(pair_eff(p, c.pairs[t].x, c.pairs[t].y));
            }
        }
        c.seg.push_back((int)c.pairs.size());
    }
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
    for (PairClass& c : cls) {
        if (c.d_pairs) (void)hipFree(c.d_pairs);
        if (c.d_pp_off) (void)hipFree(c.d_pp_off);
        if (c.d_pp) (void)hipFree(c.d_pp);
    }
    cls.clear();
}

// The surviving primitive pairs of every ket pair of every class, in device memory.
int upload_ket_prims(ccsd_plan* p)
{
    for (PairClass& c : p->ket) {
        std::vector<int> off(c.pairs.size() + 1, 0);
        std::vector<double> pp;
        for (size_t t = 0; t < c.pairs.size(); t++) {
            const ShellInfo &k = p->shells[c.pairs[t].x], &l = p->shells[c.pairs[t].y];
            const double* C = &p->coords_h[3 * k.atom];
            const double* D = &p->coords_h[3 * l.atom];
            const double CD[3] = {D[0] - C[0], D[1] - C[1], D[2] - C[2]};
            const double r2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
            const int nckl = k.nctr * l.nctr;
            for (int u = 0; u < k.nprim; u++)
                for (int v = 0; v < l.nprim; v++) {
                    const double ak = p->env[k.pexp + u], al = p->env[l.pexp + v], akl = ak + al;
                    const double mu = ak * al / akl * r2;
                    if (mu > kPrimExpCut) continue;
                    double e[kKetFields];
                    e[0] = akl;
                    for (int x = 0; x < 3; x++) e[1 + x] = C[x] + CD[x] * (al / akl);
                    e[4] = exp(-mu) / akl;
                    e[5] = 0.5 / akl;
                    for (int q = 0; q < 4; q++) {
                        if (q < nckl) {
                            const int ck = q / l.nctr, cl = q - ck * l.nctr;
                            e[6 + q] = p->env[k.pcoef + ck * k.nprim + u] * p->env[l.pcoef + cl * l.nprim + v];
                        } else {
                            e[6 + q] = 0.0;
                        }
                    }
                    pp.insert(pp.end(), e, e + kKetFields);
                }
            off[t + 1] = (int)(pp.size() / kKetFields);
        }
        HIP_TRY(hipMalloc(&c.d_pp_off, sizeof(int) * off.size()));
        HIP_TRY(hipMemcpy(c.d_pp_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&c.d_pp, sizeof(double) * std::max<size_t>(1, pp.size())));
        if (!pp.empty()) HIP_TRY(hipMemcpy(c.d_pp, pp.data(), sizeof(double) * pp.size(), hipMemcpyHostToDevice));
    }
    return CCSD_OK;
}

void free_braset(BraSet& b)
{
    free_classes(b.cls);
    for (auto& kv : b.chunks)
        if (kv.second.d) (void)hipFree(kv.second.d);
    b.chunks.clear();
    for (auto& kv : b.persist) {
        free_persist_part(kv.second.reg);
        free_persist_part(kv.second.gen);
    }
    b.persist.clear();
}

void clear_bra_cache(ccsd_plan* p)
{
    for (auto& kv : p->bra_cache) free_braset(kv.second);
    p->bra_cache.clear();
}

int build_ket(ccsd_plan* p)
{
    free_classes(p->ket);
    std::vector<int2> kpairs;
    for (int m = 0; m < p->nmol; m++)
        for (int k = p->mol_shell0[m]; k < p->mol_shell0[m + 1]; k++)
            for (int l = p->mol_shell0[m]; l <= k; l++) kpairs.push_back(make_int2(k, l));
    p->ket = make_classes(p, kpairs, p->thr > 0.0 && p->q_valid, p->ket_shell_order);
    int rc = upload_classes(p->ket);
    if (rc) return rc;
    return upload_ket_prims(p);
}

int get_bra(ccsd_plan* p, int ish0, int ish1, BraSet** out)
{
    auto key = std::make_pair(ish0, ish1);
    auto it = p->bra_cache.find(key);
    if (it == p->bra_cache.end()) {
        std::vector<int2> pairs;
        for (int i = ish0; i < ish1; i++) {
            const int m = p->shells[i].mol;
            for (int j = p->mol_shell0[m]; j < p->mol_shell0[m + 1]; j++) pairs.push_back(make_int2(i, j));
        }
        BraSet b;
        b.cls = make_classes(p, pairs, false);
        int rc = upload_classes(b.cls);
        if (rc) {
            free_braset(b);
            return rc;
        }
        it = p->bra_cache.emplace(key, std::move(b)).first;
    }
    *out = &it->second;
    return CCSD_OK;
}

// Work chunks of (bra class bc) x (ket class kc): per bra pair, the ket pairs of
// its molecule, per segment the Schwarz survivors (a prefix), cut into runs of at
// most `chunk`; most expensive chunks first so that the tail of the launch is short.
// Work list of one (bra class, ket class). `code` = ket pairs per chunk in its low 16
// bits; bits 16+ select a primitive split of the register kernel: S = 2..254 lists
// every chunk S times, 255 (kAdaptiveSplit) splits each chunk into S_c = ceil(cost /
// (class cost / 2048)) <= 16 pieces, so that no chunk runs much longer than a
// 2048th of the class (the expensive, heavily contracted chunks dominate a class's
// time otherwise). Piece s of S has w = (s << 8) | S and loops over the ket
// primitive pairs s, s + S, ...; the partial gradients add up in dE/dR.
constexpr int kAdaptiveSplit = 255;

int get_chunks(ccsd_plan* p, BraSet& bs, int bc, int kc, int code, bool screened, bool fused, ChunkList** out)
{
    // fused gradients also drop pairs without a primitive pair past the Gaussian-product cut
    auto key = std::make_tuple(bc, kc, code, (screened ? 1 : 0) | (fused ? 2 : 0));
    const int chunk = code & 0xffff, scode = std::min(255, code >> 16);
    auto it = bs.chunks.find(key);
    if (it == bs.chunks.end()) {
        const PairClass& B = bs.cls[bc];
        const PairClass& K = p->ket[kc];
        std::vector<std::pair<long long, int4>> work;
        long long quartets = 0;
        for (int bq = 0; bq < (int)B.pairs.size(); bq++) {
            const int2 bp = B.pairs[bq];
            const int m = p->shells[bp.x].mol;
            const int eb = pair_eff(p, bp.x, bp.y);
            if (fused && eb == 0) continue;
            const long long cb = (long long)std::max(1, eb) * p->shells[bp.x].nctr * p->shells[bp.y].nctr;
            const double qb = screened ? pair_q(p, bp) : 0.0;
            for (size_t s = 0; s + 1 < K.seg.size(); s++) {
                if (K.seg_mol[s] != m) continue;
                if (fused && K.seg_eff[s] == 0) continue;
                const int k0 = K.seg[s];
                int k1 = K.seg[s + 1];
                if (screened) {
                    // survivors: qb * q_kl >= thr, q_kl descending inside the segment
                    int lo = k0, hi = k1;
                    while (lo < hi) {
                        const int mid = (lo + hi) / 2;
                        if (qb * pair_q(p, K.pairs[mid]) >= p->thr) lo = mid + 1;
                        else hi = mid;
                    }
                    k1 = lo;
                }
                const long long ck = (long long)std::max(1, K.seg_eff[s]) * p->shells[K.pairs[k0].x].nctr *
                                     p->shells[K.pairs[k0].y].nctr;
                for (int k = k0; k < k1; k += chunk) {
                    const int ke = std::min(k1, k + chunk);
                    work.push_back(std::make_pair(cb * ck * (ke - k), make_int4(bq, k, ke, 0)));
                    quartets += ke - k;
                }
            }
        }
        if (scode > 1) {
            long long total = 0;
            for (const auto& w : work) total += w.first;
            const long long target = std::max(1LL, total / 2048);
            const int smax = std::max(1, std::min(16, K.max_np));
            std::vector<std::pair<long long, int4>> split_work;
            split_work.reserve(work.size());
            for (const auto& w : work) {
                int S = scode;
                if (scode == kAdaptiveSplit)
                    S = (int)std::min<long long>(smax, std::max(1LL, (w.first + target - 1) / target));
                if (S <= 1) {
                    split_work.push_back(w);
                    continue;
                }
                for (int sidx = 0; sidx < S; sidx++) {
                    int4 c = w.second;
                    c.w = (sidx << 8) | S;
                    split_work.push_back(std::make_pair(w.first / S, c));
                }
            }
            work.swap(split_work);
        }
        if (work.size() <= 4000000)
            std::stable_sort(work.begin(), work.end(),
                             [](const std::pair<long long, int4>& x, const std::pair<long long, int4>& y) {
                                 return x.first > y.first;
                             });
        std::vector<int4> h(work.size());
        for (size_t t = 0; t < work.size(); t++) h[t] = work[t].second;
        ChunkList cl;
        cl.cost.resize(work.size());
        for (size_t t = 0; t < work.size(); t++) cl.cost[t] = (float)work[t].first;
        cl.n = (int)h.size();
        cl.quartets = quartets;
        cl.cumq.resize(h.size() + 1);
        cl.cumq[0] = 0.0;
        // quartet equivalents: a piece of an S-way split counts 1/S of its quartets
        for (size_t t = 0; t < h.size(); t++)
            cl.cumq[t + 1] = cl.cumq[t] + (double)(h[t].z - h[t].y) / std::max(1, h[t].w & 255);
        if (cl.n) {
            HIP_TRY(hipMalloc(&cl.d, sizeof(int4) * h.size()));
            HIP_TRY(hipMemcpy(cl.d, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice));
        }
        it = bs.chunks.emplace(key, cl).first;
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
    p->coords_h = coords;
    compute_eff(p);
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

EriClassArgs class_args(const ccsd_plan* p, const PairClass& bc, const PairClass& kc, int deriv)
{
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
    a.nao = p->nao_cart;
    a.npair = (long long)p->nao_cart * (p->nao_cart + 1) / 2;
    a.nbas = p->nbas;
    a.mol_ao0 = p->d_mol_ao0;
    a.mol_nao = p->d_mol_nao;
    a.mol_dm_off = p->d_mol_dm_off;
    a.ket_pp_off = kc.d_pp_off;
    a.ket_pp = kc.d_pp;
    return a;
}

int launch_choice(const EriClassArgs& a, const Choice& c, hipStream_t stream)
{
    if (c.kind == 1) {
        hipError_t e = ccsd_launch_fused_reg(a, c.wpb, stream);
        if (e != hipSuccess) return hip_fail(e, "fused_ip1_reg_kernel launch");
        return CCSD_OK;
    }
    if (ccsd_eri_lds_bytes(a) > 65536) return fail(CCSD_ERR_UNSUPPORTED, "class needs more than 64 KiB LDS");
    hipError_t e = ccsd_launch_eri_class(a, stream);
    if (e != hipSuccess) return hip_fail(e, "eri_quartet_kernel launch");
    return CCSD_OK;
}

// Kernel choice of one fused-gradient class: every candidate is timed once on a
// sample of the class's real work (the first chunks of its list, about 2^18
// quartets, into a throw-away gradient) and the fastest per quartet is kept.
int choose(ccsd_plan* p, BraSet& bs, int ish0, int ish1, int bc, int kc, EriClassArgs a, bool screened,
           hipStream_t stream, Choice* out, bool lds_only = false)
{
    auto& cmap = lds_only ? p->choice_lds : p->choice;
    auto key = std::make_tuple(ish0, ish1, bc, kc);
    auto it = cmap.find(key);
    if (it != cmap.end()) {
        *out = it->second;
        return CCSD_OK;
    }
    std::vector<Choice> cand;
    const bool reg = !lds_only && p->fused_reg && ccsd_fused_reg_parts(a.li, a.lj, a.lk, a.ll, a.max_ncij, a.max_nckl) > 0;
    if (reg) {
        // unsplit and adaptively split work lists (a fixed split replaces both when forced)
        for (int w : {4, 2, 1}) {
            if (ccsd_fused_reg_lds(a.max_nij, w) > 65536) continue;
            Choice c;
            c.kind = 1; c.wpb = w; c.chunk = 64;
            Choice cs = c;
            cs.chunk = 64 | ((p->force_split > 1 ? p->force_split : kAdaptiveSplit) << 16);
            if (p->force_split > 1) {
                cand.push_back(cs);
            } else {
                cand.push_back(c);
                if (p->force_split == 0) cand.push_back(cs);
            }
        }
    }
    // generic kernel: runs of 1, 4 or 16 quartets per wave; work lists kept below
    // 4M entries (the bound counts every same-molecule quartet, screened or not).
    // Where register candidates exist they are the only ones timed unless
    // tune_generic is set: building the generic work lists dominated the first pass
    // of a plan, and the register kernels win those classes.
    const long long upper = (long long)a.nbra * a.nket;
    for (int ch : {1, 4, 16}) {
        if (reg && !cand.empty() && !p->tune_generic) break;
        if (upper / ch > 4000000 && ch != 16) continue;
        Choice c;
        c.kind = 0; c.wpb = 1; c.chunk = ch;
        cand.push_back(c);
    }
    Choice best = cand[0];
    if (p->autotune && cand.size() > 1) {
        if (!p->d_de_scratch) HIP_TRY(hipMalloc(&p->d_de_scratch, sizeof(unsigned long long) * (1 + 6 * p->natm)));
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        double best_t = 1e300;
        for (const Choice& c : cand) {
            ChunkList* cl = nullptr;
            int rc = get_chunks(p, bs, bc, kc, c.chunk, screened, true, &cl);
            if (rc) return rc;
            if (cl->n == 0) break;   // nothing to run in this class
            const int ns = (int)(std::lower_bound(cl->cumq.begin() + 1, cl->cumq.end(), 262144.0) - cl->cumq.begin());
            const int nsample = std::min(cl->n, std::max(1, ns));
            EriClassArgs t = a;
            t.chunks = cl->d;
            t.nchunks = nsample;
            t.defx = p->d_de_scratch;
            HIP_TRY(hipEventRecord(e0, stream));
            rc = launch_choice(t, c, stream);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(e1, stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            const double per_q = ms / cl->cumq[nsample];
            if (per_q < best_t) {
                best_t = per_q;
                best = c;
                double csum = 0.0;
                for (int t2 = 0; t2 < nsample; t2++) csum += cl->cost[t2];
                best.ms_per_cost = csum > 0.0 ? ms / csum : 0.0;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        // drop the work lists of the candidates not taken
        for (const Choice& c : cand) {
            if (c.chunk == best.chunk) continue;
            auto ck = bs.chunks.find(std::make_tuple(bc, kc, c.chunk, (screened ? 1 : 0) | 2));
            if (ck != bs.chunks.end()) {
                if (ck->second.d) (void)hipFree(ck->second.d);
                bs.chunks.erase(ck);
            }
        }
    }
    cmap[key] = best;
    *out = best;
    return CCSD_OK;
}

// Schwarz factors q[i][j] = max over functions a in i, b in j of sqrt((ab|ab)),
// from the diagonal quartets of every same-molecule pair, on the device.
int compute_q(ccsd_plan* p, hipStream_t stream)
{
    std::vector<int2> pairs;
    for (int m = 0; m < p->nmol; m++)
        for (int i = p->mol_shell0[m]; i < p->mol_shell0[m + 1]; i++)
            for (int j = p->mol_shell0[m]; j <= i; j++) pairs.push_back(make_int2(i, j));
    std::vector<PairClass> cls = make_classes(p, pairs, false);
    int rc = upload_classes(cls);
    unsigned long long* d_q = nullptr;
    int4* d_ch = nullptr;
    const size_t nq = (size_t)p->nbas * p->nbas;
    hipError_t e = hipSuccess;
    if (!rc) e = hipMalloc(&d_q, sizeof(unsigned long long) * nq);
    if (!rc && e == hipSuccess) e = ccsd_launch_fill((double*)d_q, (long long)nq, 0.0, stream);
    if (!rc && e == hipSuccess) e = hipMalloc(&d_ch, sizeof(int4) * std::max<size_t>(1, pairs.size()));
    for (size_t c = 0; !rc && e == hipSuccess && c < cls.size(); c++) {
        const PairClass& pc = cls[c];
        std::vector<int4> h(pc.pairs.size());
        for (size_t t = 0; t < h.size(); t++) h[t] = make_int4((int)t, (int)t, (int)t + 1, 0);
        e = hipMemcpyAsync(d_ch, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) break;
        EriClassArgs a = class_args(p, pc, pc, 0);
        a.chunks = d_ch;
        a.nchunks = (int)h.size();
        a.qmax = d_q;
        if (ccsd_eri_lds_bytes(a) > 65536) {
            rc = fail(CCSD_ERR_UNSUPPORTED, "class needs more than 64 KiB LDS");
            break;
        }
        e = ccsd_launch_eri_class(a, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);   // h is reused by the next class
    }
    std::vector<double> hq(nq);
    if (!rc && e == hipSuccess)
        e = hipMemcpyAsync(hq.data(), d_q, sizeof(double) * nq, hipMemcpyDeviceToHost, stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(stream);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "Schwarz factors");
    if (d_q) (void)hipFree(d_q);
    if (d_ch) (void)hipFree(d_ch);
    free_classes(cls);
    if (rc) return rc;
    p->q.assign(nq, 0.0);
    for (int i = 0; i < p->nbas; i++)
        for (int j = 0; j <= i; j++) p->q[(size_t)i * p->nbas + j] = p->q[(size_t)j * p->nbas + i] = hq[(size_t)i * p->nbas + j];
    p->q_valid = true;
    return CCSD_OK;
}

// Screening state current for the geometry: factors computed and the ket classes
// ordered by them (work lists and kernel choices rebuilt).
int ensure_screening(ccsd_plan* p, hipStream_t stream)
{
    if (p->thr <= 0.0 || p->q_valid) return CCSD_OK;
    int rc = compute_q(p, stream);
    if (rc) return rc;
    clear_bra_cache(p);
    return build_ket(p);
}

void record_launch(ccsd_plan* p, const EriClassArgs& a, const Choice& c, const ChunkList* cl)
{
    const int gc = (a.max_ncij > 1 || a.max_nckl > 1) ? 1 : 0;
    const int parts = c.kind != 1 ? 0
                      : a.out ? ccsd_store_reg_parts(a.li, a.lj, a.lk, a.ll, a.max_ncij, a.max_nckl, a.deriv)
                              : ccsd_fused_reg_parts(a.li, a.lj, a.lk, a.ll, a.max_ncij, a.max_nckl);
    p->last_launches.push_back({a.li, a.lj, a.lk, a.ll, gc, c.kind, c.kind == 1 ? c.wpb : 1, c.chunk, parts,
                                cl->quartets});
}

int run_eri(ccsd_plan* p, BraSet& bs, int deriv, double* out, int i_base, long long comp_stride, hipStream_t stream)
{
    p->last_quartets = 0;
    p->last_launches.clear();
    for (int bc = 0; bc < (int)bs.cls.size(); bc++) {
        for (int kc = 0; kc < (int)p->ket.size(); kc++) {
            EriClassArgs a = class_args(p, bs.cls[bc], p->ket[kc], deriv);
            // register-resident store kernel (one lane per quartet, 64 ket pairs per wave)
            // where the class has one, else the LDS kernel (one quartet per wave)
            Choice c;
            c.kind = (p->fused_reg &&
                      ccsd_store_reg_parts(a.li, a.lj, a.lk, a.ll, a.max_ncij, a.max_nckl, deriv) > 0 &&
                      ccsd_fused_reg_lds(a.max_nij, c.wpb) <= 65536) ? 1 : 0;
            c.chunk = c.kind == 1 ? 64 : 1;
            ChunkList* cl = nullptr;
            int rc = get_chunks(p, bs, bc, kc, c.chunk, false, false, &cl);
            if (rc) return rc;
            a.chunks = cl->d;
            a.nchunks = cl->n;
            a.out = out;
            a.i_base = i_base;
            a.comp_stride = comp_stride;
            if (c.kind == 1) {
                hipError_t e = ccsd_launch_store_reg(a, c.wpb, stream);
                if (e != hipSuccess) return hip_fail(e, "store_reg_kernel launch");
            } else {
                rc = launch_choice(a, c, stream);
                if (rc) return rc;
            }
            p->last_quartets += cl->quartets;
            record_launch(p, a, c, cl);
        }
    }
    return CCSD_OK;
}

// Fused gradient over the bra pairs of `bs`: explicit packed gamma (mode 0) or the
// separable form from dm / pm (mode 1).
int ensure_side_streams(ccsd_plan* p)
{
    if (p->fork_ev) return CCSD_OK;
    for (int i = 0; i < ccsd_plan::kSideStreams; i++) {
        HIP_TRY(hipStreamCreateWithFlags(&p->side[i], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&p->join_ev[i], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming));
    return CCSD_OK;
}

// Work items of one persistent launch over `cls` (work lists `lists`, units per chunk
// `parts`): consecutive units of one class are grouped until their estimated cost
// reaches 1/8 of a resident wave's share of the launch (a unit costlier than that is
// an item of its own). Order: class by class (`grouped`: the costliest class first,
// decreasing cost inside a class), so that the waves resident on a CU mostly run one
// body and share its instructions in cache, or by decreasing cost only.
// Cut the classes' units into work items from q.lists / q.parts / q.wt and upload them
// (replacing any earlier cut). The costs in time units are each class's cost model
// (primitive products x quartets) times its weight, so that items of different classes
// take about equally long (unweighted, a run of d-shell units cut to the same cost as an
// s-shell run took tens of times longer, and a few such items held the launch for most
// of its time).
int cut_persist_items(PersistPart& q)
{
    const size_t nc = q.lists.size();
    double total = 0.0;
    std::vector<double> cls_cost(nc, 0.0);
    for (size_t ci = 0; ci < nc; ci++) {
        for (float w : q.lists[ci]->cost) cls_cost[ci] += w;
        cls_cost[ci] = cls_cost[ci] * q.wt[ci] + q.unit_cost * (double)q.lists[ci]->n * q.parts[ci];
        total += cls_cost[ci];
    }
    q.total = total;
    std::vector<int4>& h = q.items;
    h.clear();
    if (q.grouped) {
        // classes costliest first, inside a class its units in list order (costliest chunk
        // first); items cut by guided self-scheduling: an item takes 1/(4 nblocks) of the
        // work not yet assigned, so items shrink towards the end of the list and the waves
        // drawing the last ones finish close together even where a class's weight is off
        std::vector<int> order(nc);
        for (size_t ci = 0; ci < nc; ci++) order[ci] = (int)ci;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cls_cost[x] > cls_cost[y]; });
        const double share = 1.0 / (4.0 * q.nblocks);
        double remaining = total;
        for (int ci : order) {
            const ChunkList* cl = q.lists[ci];
            const int np = q.parts[ci];
            const int nu = cl->n * np;
            int u0 = 0;
            double acc = 0.0;
            for (int u = 0; u < nu; u++) {
                acc += cl->cost[u / np] / np * q.wt[ci] + q.unit_cost;
                if (acc >= remaining * share || u + 1 == nu || u + 1 - u0 >= kMaxItemUnits) {
                    h.push_back(make_int4(ci, u0, u + 1, 0));
                    remaining -= acc;
                    u0 = u + 1;
                    acc = 0.0;
                }
            }
        }
    } else {
        // by cost only: items of 1/(32 nblocks) of the whole, costliest first
        const double target = total / ((double)q.nblocks * 32.0);
        std::vector<std::pair<double, int4>> items;
        for (size_t ci = 0; ci < nc; ci++) {
            const ChunkList* cl = q.lists[ci];
            const int np = q.parts[ci];
            const int nu = cl->n * np;
            int u0 = 0;
            double acc = 0.0;
            for (int u = 0; u < nu; u++) {
                acc += cl->cost[u / np] / np * q.wt[ci] + q.unit_cost;
                if (acc >= target || u + 1 == nu || u + 1 - u0 >= kMaxItemUnits) {
                    items.push_back(std::make_pair(acc, make_int4((int)ci, u0, u + 1, 0)));
                    u0 = u + 1;
                    acc = 0.0;
                }
            }
        }
        std::stable_sort(items.begin(), items.end(),
                         [](const std::pair<double, int4>& x, const std::pair<double, int4>& y) {
                             return x.first > y.first;
                         });
        h.resize(items.size());
        for (size_t t = 0; t < items.size(); t++) h[t] = items[t].second;
    }
    if (h.size() > 2147483647ULL) return fail(CCSD_ERR_UNSUPPORTED, "too many work items");
    if (q.d_items) (void)hipFree(q.d_items);
    if (q.d_ticks) (void)hipFree(q.d_ticks);
    q.d_items = nullptr;
    q.d_ticks = nullptr;
    q.nitems = (int)h.size();
    HIP_TRY(hipMalloc(&q.d_items, sizeof(int4) * std::max<size_t>(1, h.size())));
    HIP_TRY(hipMemcpy(q.d_items, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice));
    return CCSD_OK;
}

// One persistent launch's class table, grid and work items. wt: each class's time per
// cost unit (the autotuner's measurement, or a model where a class was not timed).
int build_persist_part(ccsd_plan* p, const std::vector<EriClassArgs>& cls, const std::vector<const ChunkList*>& lists,
                       const std::vector<int>& parts, const std::vector<double>& wt, bool reg, bool grouped,
                       PersistPart* out)
{
    PersistPart& q = *out;
    if (cls.empty()) return CCSD_OK;
    for (const EriClassArgs& a : cls) {
        const size_t lds = reg ? ccsd_reg_persist_lds(a.max_nij) : ccsd_eri_lds_bytes(a);
        if (lds > 65536) return fail(CCSD_ERR_UNSUPPORTED, "class needs more than 64 KiB LDS");
        q.lds = std::max(q.lds, lds);
    }
    if (p->n_cu == 0) HIP_TRY(hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, p->device));
    const int bpc = reg ? ccsd_reg_persist_blocks_per_cu(q.lds) : ccsd_persistent_blocks_per_cu(q.lds);
    if (bpc <= 0) return fail(CCSD_ERR_UNSUPPORTED, "persistent kernel cannot be resident at this LDS size");
    q.nblocks = bpc * p->n_cu;
    for (size_t ci = 0; ci < cls.size(); ci++)
        if ((long long)lists[ci]->n * parts[ci] > 2147483647LL) return fail(CCSD_ERR_UNSUPPORTED, "work list too long");
    q.lists = lists;
    q.parts = parts;
    q.wt = wt;
    q.grouped = grouped;
    int rc = cut_persist_items(q);
    if (rc) return rc;
    q.ncls = (int)cls.size();
    HIP_TRY(hipMalloc(&q.d_cls, sizeof(EriClassArgs) * cls.size()));
    HIP_TRY(hipMemcpy(q.d_cls, cls.data(), sizeof(EriClassArgs) * cls.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&q.d_counter, sizeof(double)));
    return CCSD_OK;
}

// Re-weight the register launch's classes from the item durations its last launch
// recorded, as this kernel runs them (its residency and part counts): item time =
// a_c x (cost of its units) + b x (number of units), a_c per class and b, the fixed
// time per unit, shared, fitted by alternating least squares; then cut the items
// again. Classes without a recorded item keep their weight scaled by the median ratio
// of the fitted ones.
int calibrate_persist_part(PersistPart& q, hipStream_t stream)
{
    std::vector<unsigned int> ticks(q.nitems);
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipMemcpy(ticks.data(), q.d_ticks, sizeof(unsigned int) * ticks.size(), hipMemcpyDeviceToHost));
    const size_t nc = q.lists.size();
    const int n = q.nitems;
    std::vector<double> C(n), N(n), T(n);
    for (int t = 0; t < n; t++) {
        const int4 it = q.items[t];
        const ChunkList* cl = q.lists[it.x];
        const int np = q.parts[it.x];
        double c = 0.0;
        for (int u = it.y; u < it.z; u++) c += cl->cost[u / np] / np;
        C[t] = c;
        N[t] = it.z - it.y;
        T[t] = ticks[t];
    }
    std::vector<double> a(nc, 0.0);
    double b = 0.0;
    for (int iter = 0; iter < 20; iter++) {
        std::vector<double> num(nc, 0.0), den(nc, 0.0);
        for (int t = 0; t < n; t++) {
            num[q.items[t].x] += (T[t] - b * N[t]) * C[t];
            den[q.items[t].x] += C[t] * C[t];
        }
        for (size_t ci = 0; ci < nc; ci++) a[ci] = den[ci] > 0.0 ? std::max(0.0, num[ci] / den[ci]) : 0.0;
        double nb = 0.0, db = 0.0;
        for (int t = 0; t < n; t++) {
            nb += (T[t] - a[q.items[t].x] * C[t]) * N[t];
            db += N[t] * N[t];
        }
        b = db > 0.0 ? std::max(0.0, nb / db) : 0.0;
    }
    std::vector<double> ratio;
    for (size_t ci = 0; ci < nc; ci++)
        if (a[ci] > 0.0 && q.wt[ci] > 0.0) ratio.push_back(a[ci] / q.wt[ci]);
    if (ratio.empty()) return CCSD_OK;
    std::nth_element(ratio.begin(), ratio.begin() + ratio.size() / 2, ratio.end());
    const double med = ratio[ratio.size() / 2];
    for (size_t ci = 0; ci < nc; ci++) q.wt[ci] = a[ci] > 0.0 ? a[ci] : q.wt[ci] * med;
    q.unit_cost = b;
    int rc = cut_persist_items(q);
    if (rc) return rc;
    q.calibrated = true;
    return CCSD_OK;
}

// The persistent pass of a bra set. Option 5 = 1: every class with a register body runs
// in one launch of those bodies (its work list as the autotuner chose it, split along
// the ket primitive pairs or not; P units per chunk), the classes without one in one
// launch of the generic body; option 5 = 2: every class in one launch of the generic
// body (one unit per chunk, chunk length from the autotuner).
int build_persist(ccsd_plan* p, BraSet& bs, int ish0, int ish1, const EriClassArgs& proto, bool screened,
                  hipStream_t stream, PersistPlan* out)
{
    const bool use_reg = p->persistent == 1 && p->fused_reg;
    // A class runs in the register-body launch when it has a body there and the
    // per-class choice is a register kernel (the same work list, hence the same
    // per-wave sums); otherwise in the generic-body launch, with the generic kernel's
    // own choice of chunk length (a register work list is not a generic one: its
    // chunks may be pieces of a primitive split).
    auto pick = [&](int bc, int kc, EriClassArgs a, Choice* c, bool* reg) {
        a.gamma_mode = proto.gamma_mode;
        a.gamma = proto.gamma;
        a.dm = proto.dm;
        a.pm = proto.pm;
        a.gamma_scale = proto.gamma_scale;
        *reg = use_reg && ccsd_reg_persist_parts(a.li, a.lj, a.lk, a.ll, a.max_ncij, a.max_nckl) > 0;
        int rc = choose(p, bs, ish0, ish1, bc, kc, a, screened, stream, c, !use_reg);
        if (rc) return rc;
        *reg = *reg && c->kind == 1;
        if (!*reg && c->kind == 1) rc = choose(p, bs, ish0, ish1, bc, kc, a, screened, stream, c, true);
        return rc;
    };
    // kernel choices (chunk length per class) first: the autotuner drops the work lists
    // of the candidates it does not take, so none is referenced before all are chosen
    for (int bc = 0; bc < (int)bs.cls.size(); bc++)
        for (int kc = 0; kc < (int)p->ket.size(); kc++) {
            Choice c;
            bool reg;
            int rc = pick(bc, kc, class_args(p, bs.cls[bc], p->ket[kc], 1), &c, &reg);
            if (rc) return rc;
        }
    // built in place in *out, so that the caller frees whatever was allocated if a step fails
    PersistPlan& pp = *out;
    std::vector<EriClassArgs> cls[2];
    std::vector<const ChunkList*> lists[2];
    std::vector<int> parts[2];
    std::vector<double> wt[2];   // measured time per cost unit of each class (0: not timed)
    std::vector<double> model[2];   // the fallback: Rys roots x Cartesian functions
    for (int bc = 0; bc < (int)bs.cls.size(); bc++)
        for (int kc = 0; kc < (int)p->ket.size(); kc++) {
            EriClassArgs a = class_args(p, bs.cls[bc], p->ket[kc], 1);
            Choice c;
            bool reg;
            int rc = pick(bc, kc, a, &c, &reg);
            if (rc) return rc;
            const int chunk = c.chunk;
            ChunkList* cl = nullptr;
            rc = get_chunks(p, bs, bc, kc, chunk, screened, true, &cl);
            if (rc) return rc;
            if (cl->n == 0) continue;
            a.chunks = cl->d;
            a.nchunks = cl->n;
            const int np = reg ? ccsd_reg_persist_parts(a.li, a.lj, a.lk, a.ll, a.max_ncij, a.max_nckl) : 1;
            pp.quartets += cl->quartets;
            pp.launches.push_back({a.li, a.lj, a.lk, a.ll, (a.max_ncij > 1 || a.max_nckl > 1) ? 1 : 0, reg ? 3 : 2, 1,
                                   chunk, reg ? np : 0, cl->quartets});
            cls[reg].push_back(a);
            lists[reg].push_back(cl);
            parts[reg].push_back(np);
            // a register body with more parts here than in its per-class kernel repeats
            // the roots and 1-D factors per part: count that as half a chunk per extra part
            const int np0 = reg ? ccsd_fused_reg_parts(a.li, a.lj, a.lk, a.ll, a.max_ncij, a.max_nckl) : 1;
            const double extra = (reg && np0 > 0 && np > np0) ? 1.0 + 0.5 * ((double)np / np0 - 1.0) : 1.0;
            wt[reg].push_back(c.ms_per_cost * extra);
            const int L = a.li + a.lj + a.lk + a.ll + 1;
            model[reg].push_back((double)(L / 2 + 1) * ncart(a.li) * ncart(a.lj) * ncart(a.lk) * ncart(a.ll) * extra);
        }
    // one scale per launch: the measured weights where every class of it was timed
    for (int r = 0; r < 2; r++)
        for (double w : wt[r])
            if (!(w > 0.0)) {
                wt[r] = model[r];
                break;
            }
    int rc = build_persist_part(p, cls[1], lists[1], parts[1], wt[1], true, p->persist_order == 0, &pp.reg);
    if (rc) return rc;
    return build_persist_part(p, cls[0], lists[0], parts[0], wt[0], false, false, &pp.gen);
}

// Zeroed fixed-point gradient words of a fused pass (allocated on first use).
int defx_begin(ccsd_plan* p, hipStream_t stream)
{
    const long long n = 1 + 6LL * p->natm;
    if (!p->d_defx) HIP_TRY(hipMalloc(&p->d_defx, sizeof(unsigned long long) * n));
    HIP_TRY(ccsd_launch_fill((double*)p->d_defx, n, 0.0, stream));   // all-zero bits
    return CCSD_OK;
}

// The fused pass as persistent launches (plan option 5): the register-body launch on
// `stream`, the generic-body launch (classes without a register body) on a side stream
// beside it; both add into the fixed-point gradient with integer atomics.
int run_fused_persistent(ccsd_plan* p, BraSet& bs, int ish0, int ish1, double* de, double scale, int gamma_mode,
                         const double* gamma, const double* dm, const double* pm, hipStream_t stream)
{
    const bool screened = p->thr > 0.0;
    auto it = bs.persist.find(screened ? 1 : 0);
    if (it == bs.persist.end()) {
        EriClassArgs proto;
        memset(&proto, 0, sizeof(proto));
        proto.gamma_mode = gamma_mode;
        proto.gamma = gamma;
        proto.dm = dm;
        proto.pm = pm;
        proto.gamma_scale = scale;
        PersistPlan pp;
        int rc = build_persist(p, bs, ish0, ish1, proto, screened, stream, &pp);
        if (rc) {
            free_persist_part(pp.reg);
            free_persist_part(pp.gen);
            return rc;
        }
        it = bs.persist.emplace(screened ? 1 : 0, pp).first;
    }
    PersistPlan& pp = it->second;
    // the second call of a launch re-cuts its items from the first one's recorded item
    // durations (not while a graph is captured: that needs a host read)
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(stream, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    bool record[2] = {false, false};
    PersistPart* parts_[2] = {&pp.reg, &pp.gen};
    for (int k = 0; k < 2; k++) {
        PersistPart& q = *parts_[k];
        if (q.nitems <= 0 || capturing) continue;
        if (q.launches > 0 && !q.calibrated && q.d_ticks) {
            HIP_TRY(hipDeviceSynchronize());   // the generic launch ran on a side stream
            int rc = calibrate_persist_part(q, stream);
            if (rc) return rc;
        }
        record[k] = !q.calibrated;
        if (record[k] && !q.d_ticks) HIP_TRY(hipMalloc(&q.d_ticks, sizeof(unsigned int) * (size_t)q.nitems));
    }
    p->last_quartets = pp.quartets;
    p->last_launches = pp.launches;
    int rc = defx_begin(p, stream);
    if (rc) return rc;
    const bool both = pp.reg.nitems > 0 && pp.gen.nitems > 0;
    if (both) {
        rc = ensure_side_streams(p);
        if (rc) return rc;
    }
    hipStream_t gst = both ? p->side[0] : stream;
    if (pp.reg.nitems > 0) HIP_TRY(ccsd_launch_fill((double*)pp.reg.d_counter, 1, 0.0, stream));
    if (pp.gen.nitems > 0) HIP_TRY(ccsd_launch_fill((double*)pp.gen.d_counter, 1, 0.0, stream));
    if (both) {
        HIP_TRY(hipEventRecord(p->fork_ev, stream));
        HIP_TRY(hipStreamWaitEvent(gst, p->fork_ev, 0));
    }
    // The generic launch goes first and its waves stay resident until its list is
    // exhausted: given every slot, it would hold the register launch back until then.
    // It gets three times the share of the wave slots its estimated work is of the
    // whole (both estimates in the same units: the autotuner's, later the recorded
    // durations): a wave's speed depends on what shares its CU, so at its bare share
    // the generic launch outlasted the register one (64 molecules: 2.81 s on 400
    // blocks against 2.47 s), while three times that ends it well before.
    int gen_blocks = pp.gen.nblocks;
    if (both && pp.gen.total > 0.0 && pp.reg.total > 0.0) {
        const double f = std::min(1.0, 3.0 * pp.gen.total / (pp.gen.total + pp.reg.total));
        gen_blocks = (int)std::min<double>(pp.gen.nblocks, std::max(1.0, std::ceil(f * pp.reg.nblocks)));
    }
    if (pp.gen.nitems > 0) {
        HIP_TRY(ccsd_launch_eri_persistent(pp.gen.d_cls, pp.gen.d_items, pp.gen.nitems, pp.gen.d_counter,
                                           record[1] ? pp.gen.d_ticks : nullptr, gamma, dm, pm, p->d_defx, scale,
                                           gamma_mode, gen_blocks, pp.gen.lds, gst));
        if (record[1]) pp.gen.launches++;
    }
    if (pp.reg.nitems > 0) {
        HIP_TRY(ccsd_launch_reg_persistent(pp.reg.d_cls, pp.reg.ncls, pp.reg.d_items, pp.reg.nitems,
                                           pp.reg.d_counter, record[0] ? pp.reg.d_ticks : nullptr, gamma, dm, pm,
                                           p->d_defx, scale, gamma_mode, pp.reg.nblocks, pp.reg.lds, stream));
        if (record[0]) pp.reg.launches++;
    }
    if (both) {
        HIP_TRY(hipEventRecord(p->join_ev[0], gst));
        HIP_TRY(hipStreamWaitEvent(stream, p->join_ev[0], 0));
    }
    HIP_TRY(ccsd_launch_defx_to_de(p->d_defx, p->natm, de, stream));
    return CCSD_OK;
}

int run_fused(ccsd_plan* p, BraSet& bs, int ish0, int ish1, double* de, double scale, int gamma_mode,
              const double* gamma, const double* dm, const double* pm, hipStream_t stream)
{
    if (p->persistent) return run_fused_persistent(p, bs, ish0, ish1, de, scale, gamma_mode, gamma, dm, pm, stream);
    const bool screened = p->thr > 0.0;
    p->last_quartets = 0;
    p->last_launches.clear();
    // kernel choices first (the autotuner synchronises), then the launches, spread over
    // the side streams: they only add into the fixed-point gradient with integer atomics,
    // so neither their order nor the result depends on the streams
    for (int bc = 0; bc < (int)bs.cls.size(); bc++)
        for (int kc = 0; kc < (int)p->ket.size(); kc++) {
            EriClassArgs a = class_args(p, bs.cls[bc], p->ket[kc], 1);
            a.gamma_mode = gamma_mode;
            a.gamma = gamma;
            a.dm = dm;
            a.pm = pm;
            a.gamma_scale = scale;
            Choice c;
            int rc = choose(p, bs, ish0, ish1, bc, kc, a, screened, stream, &c);
            if (rc) return rc;
        }
    {
        int rc = defx_begin(p, stream);
        if (rc) return rc;
    }
    const bool conc = p->concurrent;
    if (conc) {
        int rc = ensure_side_streams(p);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(p->fork_ev, stream));
        for (int i = 0; i < ccsd_plan::kSideStreams; i++) HIP_TRY(hipStreamWaitEvent(p->side[i], p->fork_ev, 0));
    }
    int nl = 0;
    for (int bc = 0; bc < (int)bs.cls.size(); bc++) {
        for (int kc = 0; kc < (int)p->ket.size(); kc++) {
            EriClassArgs a = class_args(p, bs.cls[bc], p->ket[kc], 1);
            a.gamma_mode = gamma_mode;
            a.gamma = gamma;
            a.dm = dm;
            a.pm = pm;
            a.gamma_scale = scale;
            Choice c;
            int rc = choose(p, bs, ish0, ish1, bc, kc, a, screened, stream, &c);
            if (rc) return rc;
            ChunkList* cl = nullptr;
            rc = get_chunks(p, bs, bc, kc, c.chunk, screened, true, &cl);
            if (rc) return rc;
            a.chunks = cl->d;
            a.nchunks = cl->n;
            a.defx = p->d_defx;
            if (cl->n == 0) continue;
            hipStream_t st = conc ? p->side[nl % ccsd_plan::kSideStreams] : stream;
            nl++;
            rc = launch_choice(a, c, st);
            if (rc) return rc;
            p->last_quartets += cl->quartets;
            record_launch(p, a, c, cl);
        }
    }
    if (conc) {
        for (int i = 0; i < ccsd_plan::kSideStreams; i++) {
            HIP_TRY(hipEventRecord(p->join_ev[i], p->side[i]));
            HIP_TRY(hipStreamWaitEvent(stream, p->join_ev[i], 0));
        }
    }
    HIP_TRY(ccsd_launch_defx_to_de(p->d_defx, p->natm, de, stream));
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
    if (p->nmol != 1) return fail(CCSD_ERR_UNSUPPORTED, "integral tensors need a single-molecule plan");
    if (ish0 < 0 || ish1 > p->nbas || ish0 >= ish1) return fail(CCSD_ERR_SHAPE, "bad shell range");
    if (!cart && !p->has_c2s) return fail(CCSD_ERR_UNSUPPORTED, "plan has no cart2sph coefficients");
    BraSet* bra = nullptr;
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
    rc = run_eri(p, *bra, deriv, cart_buf, i0c, comp_stride_c, stream);
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

int ccsd_plan_destroy_impl(ccsd_plan* p);

int create_plan(ccsd_plan** out, int nmol, const int* mol_nbas, const int* atm, int natm, const int* bas, int nbas,
                const double* env, int nenv, const double* c2s)
{
    if (!out || !atm || !bas || !env) return fail(CCSD_ERR_NULL, "null argument to ccsd_plan_create");
    if (natm <= 0 || nbas <= 0 || nenv <= 0) return fail(CCSD_ERR_SHAPE, "empty molecule");
    *out = nullptr;
    std::vector<int> shell0(nmol + 1, 0);
    for (int m = 0; m < nmol; m++) {
        if (mol_nbas[m] <= 0) return fail(CCSD_ERR_SHAPE, "molecule without shells");
        shell0[m + 1] = shell0[m] + mol_nbas[m];
    }
    if (shell0[nmol] != nbas) return fail(CCSD_ERR_SHAPE, "molecule shell counts do not add up to nbas");
    ccsd_plan* p = new ccsd_plan();
    p->nmol = nmol;
    p->mol_shell0 = shell0;
    p->mol_ao0.assign(nmol, 0);
    p->mol_nao.assign(nmol, 0);
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
        s.mol = (int)(std::upper_bound(shell0.begin(), shell0.end(), i) - shell0.begin()) - 1;
        s.pad_ = 0;
        if (i == shell0[s.mol]) p->mol_ao0[s.mol] = aoc;
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
    for (int m = 0; m < nmol; m++)
        p->mol_nao[m] = (m + 1 < nmol ? p->mol_ao0[m + 1] : aoc) - p->mol_ao0[m];
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return fail(CCSD_ERR_HIP, "hipGetDevice failed");
    }
    int rc = CCSD_OK;
    if (hipMalloc(&p->d_shells, sizeof(ShellInfo) * nbas) != hipSuccess ||
        hipMalloc(&p->d_env, sizeof(double) * nenv) != hipSuccess ||
        hipMalloc(&p->d_coords, sizeof(double) * 3 * natm) != hipSuccess ||
        hipMalloc(&p->d_charges, sizeof(double) * natm) != hipSuccess ||
        hipMalloc(&p->d_mol_ao0, sizeof(int) * nmol) != hipSuccess ||
        hipMalloc(&p->d_mol_nao, sizeof(int) * nmol) != hipSuccess ||
        hipMalloc(&p->d_mol_dm_off, sizeof(long long) * nmol) != hipSuccess) {
        ccsd_plan_destroy_impl(p);
        return fail(CCSD_ERR_ALLOC, "device allocation for the basis plan failed");
    }
    std::vector<long long> dm_off(nmol, 0);
    for (int m = 1; m < nmol; m++) dm_off[m] = dm_off[m - 1] + (long long)p->mol_nao[m - 1] * p->mol_nao[m - 1];
    if (hipMemcpy(p->d_shells, p->shells.data(), sizeof(ShellInfo) * nbas, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_mol_ao0, p->mol_ao0.data(), sizeof(int) * nmol, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_mol_nao, p->mol_nao.data(), sizeof(int) * nmol, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_mol_dm_off, dm_off.data(), sizeof(long long) * nmol, hipMemcpyHostToDevice) != hipSuccess) {
        ccsd_plan_destroy_impl(p);
        return fail(CCSD_ERR_HIP, "upload of shells failed");
    }
    rc = upload_geometry(p);
    if (rc) {
        std::string msg = g_last_error;
        ccsd_plan_destroy_impl(p);
        return fail(rc, msg);
    }
    rc = build_ket(p);
    if (rc) {
        std::string msg = g_last_error;
        ccsd_plan_destroy_impl(p);
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
                        ccsd_plan_destroy_impl(p);
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
            ccsd_plan_destroy_impl(p);
            return fail(CCSD_ERR_HIP, "upload of cart2sph terms failed");
        }
    }
    *out = p;
    return CCSD_OK;
}

}  // namespace

extern "C" {

const char* ccsd_last_error(void) { return g_last_error.c_str(); }

int ccsd_plan_create(ccsd_plan** out, const int* atm, int natm, const int* bas, int nbas, const double* env, int nenv,
                     const double* c2s)
{
    return create_plan(out, 1, &nbas, atm, natm, bas, nbas, env, nenv, c2s);
}

int ccsd_plan_create_batch(ccsd_plan** out, int nmol, const int* mol_nbas, const int* atm, int natm, const int* bas,
                           int nbas, const double* env, int nenv)
{
    if (!mol_nbas) return fail(CCSD_ERR_NULL, "null mol_nbas");
    if (nmol <= 0) return fail(CCSD_ERR_SHAPE, "empty batch");
    return create_plan(out, nmol, mol_nbas, atm, natm, bas, nbas, env, nenv, nullptr);
}

int ccsd_plan_destroy(ccsd_plan* p) { return ccsd_plan_destroy_impl(p); }

}  // extern "C"

namespace {

int ccsd_plan_destroy_impl(ccsd_plan* p)
{
    if (!p) return CCSD_OK;
    free_classes(p->ket);
    clear_bra_cache(p);
    if (p->d_mol_ao0) (void)hipFree(p->d_mol_ao0);
    if (p->d_mol_nao) (void)hipFree(p->d_mol_nao);
    if (p->d_mol_dm_off) (void)hipFree(p->d_mol_dm_off);
    if (p->d_de_scratch) (void)hipFree(p->d_de_scratch);
    if (p->d_defx) (void)hipFree(p->d_defx);
    if (p->d_int1e_fx) (void)hipFree(p->d_int1e_fx);
    for (int i = 0; i < ccsd_plan::kSideStreams; i++) {
        if (p->side[i]) (void)hipStreamDestroy(p->side[i]);
        if (p->join_ev[i]) (void)hipEventDestroy(p->join_ev[i]);
    }
    if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
    if (p->d_shells) (void)hipFree(p->d_shells);
    if (p->d_env) (void)hipFree(p->d_env);
    if (p->d_coords) (void)hipFree(p->d_coords);
    if (p->d_charges) (void)hipFree(p->d_charges);
    if (p->d_c2s) (void)hipFree(p->d_c2s);
    delete p;
    return CCSD_OK;
}

}  // namespace

extern "C" {

int ccsd_plan_set_env(ccsd_plan* p, const double* env, int nenv)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!env) return fail(CCSD_ERR_NULL, "null env");
    if (nenv != p->nenv) return fail(CCSD_ERR_SHAPE, "env length differs from the plan's");
    p->env.assign(env, env + nenv);
    p->q_valid = false;     // Schwarz factors follow the geometry (recomputed at the next screened call)
    rc = upload_geometry(p);
    if (rc) return rc;
    // primitive cuts, segments and work lists follow the geometry as well
    clear_bra_cache(p);
    return build_ket(p);
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
    if (p->nmol != 1) return fail(CCSD_ERR_UNSUPPORTED, "explicit gamma needs a single-molecule plan");
    if (ish0 < 0 || ish1 > p->nbas || ish0 >= ish1) return fail(CCSD_ERR_SHAPE, "bad shell range");
    rc = ensure_screening(p, (hipStream_t)stream);
    if (rc) return rc;
    BraSet* bra = nullptr;
    rc = get_bra(p, ish0, ish1, &bra);
    if (rc) return rc;
    return run_fused(p, *bra, ish0, ish1, de, scale, 0, gamma, nullptr, nullptr, (hipStream_t)stream);
}

int ccsd_grad_ip1_gamma_sep(ccsd_plan* p, double* de, const double* gamma, const double* dm, const double* pm,
                            double scale, int ish0, int ish1, void* stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!de || !gamma || !dm || !pm) return fail(CCSD_ERR_NULL, "null de, gamma, dm or pm");
    if (p->nmol != 1) return fail(CCSD_ERR_UNSUPPORTED, "explicit gamma needs a single-molecule plan");
    if (ish0 < 0 || ish1 > p->nbas || ish0 >= ish1) return fail(CCSD_ERR_SHAPE, "bad shell range");
    rc = ensure_screening(p, (hipStream_t)stream);
    if (rc) return rc;
    BraSet* bra = nullptr;
    rc = get_bra(p, ish0, ish1, &bra);
    if (rc) return rc;
    return run_fused(p, *bra, ish0, ish1, de, scale, 2, gamma, dm, pm, (hipStream_t)stream);
}

int ccsd_grad_ip1_separable(ccsd_plan* p, double* de, const double* dm, const double* pm, double scale, void* stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!de || !dm || !pm) return fail(CCSD_ERR_NULL, "null de, dm or pm");
    rc = ensure_screening(p, (hipStream_t)stream);
    if (rc) return rc;
    BraSet* bra = nullptr;
    rc = get_bra(p, 0, p->nbas, &bra);
    if (rc) return rc;
    return run_fused(p, *bra, 0, p->nbas, de, scale, 1, nullptr, dm, pm, (hipStream_t)stream);
}

int ccsd_plan_screen(ccsd_plan* p, double thr, void* stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!(thr >= 0.0)) return fail(CCSD_ERR_SHAPE, "screening threshold must be >= 0");
    if (thr == p->thr) return CCSD_OK;
    p->thr = thr;
    clear_bra_cache(p);
    if (thr > 0.0 && !p->q_valid) {
        rc = compute_q(p, (hipStream_t)stream);
        if (rc) return rc;
    }
    return build_ket(p);
}

int ccsd_schwarz(ccsd_plan* p, double* q, void* stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!q) return fail(CCSD_ERR_NULL, "null output pointer");
    if (!p->q_valid) {
        rc = compute_q(p, (hipStream_t)stream);
        if (rc) return rc;
        if (p->thr > 0.0) {
            clear_bra_cache(p);
            rc = build_ket(p);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipMemcpyAsync(q, p->q.data(), sizeof(double) * p->q.size(), hipMemcpyHostToDevice, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return CCSD_OK;
}

int ccsd_plan_last_launches(const ccsd_plan* p, long long* rows, int max_rows, int* nrows)
{
    if (!p || !nrows) return fail(CCSD_ERR_NULL, "null argument");
    *nrows = (int)p->last_launches.size();
    if (rows)
        for (int r = 0; r < std::min(max_rows, *nrows); r++)
            for (int c = 0; c < 10; c++) rows[10 * r + c] = p->last_launches[r][c];
    return CCSD_OK;
}

int ccsd_plan_quartets(const ccsd_plan* p, long long* count)
{
    if (!p || !count) return fail(CCSD_ERR_NULL, "null argument");
    *count = p->last_quartets;
    return CCSD_OK;
}

int ccsd_int1e(ccsd_plan* p, int kind, double* out, void* stream)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out) return fail(CCSD_ERR_NULL, "null output pointer");
    if (kind < 0 || kind > 5) return fail(CCSD_ERR_SHAPE, "int1e kind must be 0..5");
    const long long nn = (long long)p->nao_cart * p->nao_cart;
    const long long nmat = kind <= 2 ? 1 : (kind <= 4 ? 3 : 3LL * p->natm);
    const long long nw = 1 + 2 * nmat * nn;     // fixed-point words: overflow count, then 2 per element
    if (nw > p->int1e_fx_words) {
        if (p->d_int1e_fx) (void)hipFree(p->d_int1e_fx);
        p->d_int1e_fx = nullptr;
        p->int1e_fx_words = 0;
        HIP_TRY(hipMalloc(&p->d_int1e_fx, sizeof(unsigned long long) * nw));
        p->int1e_fx_words = nw;
    }
    hipError_t e = ccsd_launch_fill((double*)p->d_int1e_fx, nw, 0.0, (hipStream_t)stream);
    if (e == hipSuccess) {
        int maxprim = 1;
        for (const ShellInfo& sh : p->shells) maxprim = std::max(maxprim, sh.nprim);
        e = ccsd_launch_int1e(kind, p->d_int1e_fx, p->d_shells, p->d_env, p->d_coords, p->d_charges, p->natm,
                              p->nbas, p->nao_cart, maxprim, (hipStream_t)stream);
    }
    if (e == hipSuccess) e = ccsd_launch_int1e_convert(p->d_int1e_fx, nmat * nn, out, (hipStream_t)stream);
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

int ccsd_sym_pairs(double* t, int n, double scale, void* stream)
{
    if (!t) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 0 || n > 46340) return fail(CCSD_ERR_SHAPE, "bad dimension");
    hipError_t e = ccsd_launch_sym_pairs(t, n, scale, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "sym_pairs_kernel");
    return CCSD_OK;
}

int ccsd_l1_terms(const double* x, int o, int v, const double* y, long long sy0, long long sy1, double c1, double c2,
                  double c3, double c4, double* ws, double* out, void* stream)
{
    if (!x || !y || !ws || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (o < 0 || v < 0 || (long long)o * o > 0x7fffffffLL) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    if (ccsd_l1_terms_lds(v) > 65536) return fail(CCSD_ERR_UNSUPPORTED, "nvir too large for l1_block_kernel (<= 682)");
    hipError_t e = ccsd_launch_l1_terms(x, o, v, y, sy0, sy1, c1, c2, c3, c4, ws, out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "l1_block_kernel");
    return CCSD_OK;
}

int ccsd_unpack_pairs(const double* x, const long long* iu, const long long* ju, int o, int v, long long npair,
                      double scale, double* out, void* stream)
{
    if (!x || !iu || !ju || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (o < 0 || v < 0 || npair < 0 || npair > (long long)o * (o + 1) / 2) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    hipError_t e = ccsd_launch_unpack_pairs(x, iu, ju, o, v, npair, scale, out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "unpack_pairs_kernel");
    return CCSD_OK;
}

int ccsd_vvvv_pm_pack(const double* l2, int o, int v, const long long* iu, const long long* ju, long long npair,
                      const long long* tc, const long long* td, long long np, long long ld, int split, void* out,
                      void* lo, void* stream)
{
    if (!l2 || !iu || !ju || !tc || !td || !out || (split && !lo)) return fail(CCSD_ERR_NULL, "null pointer");
    if (o < 0 || v < 0 || npair < 0 || npair > (long long)o * (o + 1) / 2 || np != (long long)v * (v + 1) / 2 ||
        ld < np)
        return fail(CCSD_ERR_SHAPE, "bad dimensions");
    hipError_t e = ccsd_launch_vvvv_pm_pack(l2, o, v, iu, ju, npair, tc, td, np, ld, split, out, lo,
                                            (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "vvvv_pm_pack_kernel");
    return CCSD_OK;
}

int ccsd_vvvv_pm_unpack(const double* x, int o, int v, long long ldx, double scale, double* out, void* stream)
{
    if (!x || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (o < 0 || v < 0 || ldx < (long long)v * (v + 1) / 2) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    hipError_t e = ccsd_launch_vvvv_pm_unpack(x, o, v, ldx, scale, out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "vvvv_pm_unpack_kernel");
    return CCSD_OK;
}

int ccsd_l2_finish(const double* x, int o, int v, const double* u, const double* w, const double* s,
                   const double* sdev, double shost, const double* d, double* out, void* stream)
{
    if (!x || !u || !w || !d || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (o < 0 || v < 0) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    if (out == x) return fail(CCSD_ERR_SHAPE, "out must not overlap x");
    const long long nt = (v + 31) / 32;
    if ((long long)o * o * nt * nt > 0x7fffffffLL) return fail(CCSD_ERR_SHAPE, "too many tiles for one launch");
    hipError_t e = ccsd_launch_l2_finish(x, o, v, u, w, s, sdev, shost, d, out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "l2_finish_kernel");
    return CCSD_OK;
}

long long ccsd_diis_workspace(long long n, int nrows)
{
    if (n < 0 || nrows < 0 || nrows > ccsd_diis_max_rows()) return -1;
    return ccsd_diis_push_blocks(n) * (nrows + 1);
}

int ccsd_diis_push(const double* s0, const double* s1, long long n0, long long n, const double* rhs, const double* prev,
                   float* xs, float* es, const float* e_rows, long long lde, int nrows, int slot, double* prev_out,
                   double* dots, double* ws, long long ws_len, void* stream)
{
    if (!s0 || (n > n0 && !s1)) return fail(CCSD_ERR_NULL, "null pointer");
    if (n0 < 0 || n < n0 || nrows < 0 || nrows > ccsd_diis_max_rows() || (nrows > 0 && (slot < 0 || slot >= nrows)) ||
        (nrows > 1 && (!e_rows || lde < n)))
        return fail(CCSD_ERR_SHAPE, "bad dimensions");
    if (prev && (!dots || !ws || ws_len < ccsd_diis_workspace(n, nrows)))
        return fail(CCSD_ERR_SHAPE, "dots / workspace missing or too small");
    if (es && (nrows < 1 || !prev)) return fail(CCSD_ERR_SHAPE, "an error vector needs prev and its row");
    hipError_t e = ccsd_launch_diis_push(s0, s1, n0, n, rhs, prev, xs, es, e_rows, lde, nrows, slot, prev_out, dots,
                                         ws, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "diis_push_kernel");
    return CCSD_OK;
}

int ccsd_diis_combine(const float* rows, long long ldx, int nrows, const double* coef, const double* base, double* out,
                      long long n, void* stream)
{
    if (!out || (nrows > 0 && (!rows || !coef))) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 0 || nrows < 0 || nrows > ccsd_diis_max_rows() + 1 || (nrows > 1 && ldx < n))
        return fail(CCSD_ERR_SHAPE, "bad dimensions");
    hipError_t e = ccsd_launch_diis_combine(rows, ldx, nrows, coef, base, out, n, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "diis_combine_kernel");
    return CCSD_OK;
}

int ccsd_pair_combine(const double* x, const long long* xstrides, int o, int v, double c1, double c2, double c3,
                      double c4, const double* d, double* out, const long long* ostrides, int beta, void* stream)
{
    if (!x || !out || !xstrides || !ostrides) return fail(CCSD_ERR_NULL, "null pointer");
    if (o < 0 || v < 0 || (beta != 0 && beta != 1)) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    const long long nt = (v + 31) / 32;
    if ((long long)o * o * nt * nt > 0x7fffffffLL) return fail(CCSD_ERR_SHAPE, "too many tiles for one launch");
    hipError_t e = ccsd_launch_pair_combine(x, xstrides, o, v, c1, c2, c3, c4, d, out, ostrides, beta,
                                            (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "pair_combine_kernel");
    return CCSD_OK;
}

int ccsd_permute(const double* src, int ndim, const long long* sizes, const long long* strides, double* dst, void* hi,
                 void* lo, void* stream)
{
    if (!src || !sizes || !strides || (!dst && (!hi || !lo))) return fail(CCSD_ERR_NULL, "null pointer");
    if (ndim < 1 || ndim > 6) return fail(CCSD_ERR_SHAPE, "ndim must be 1..6");
    for (int k = 0; k < ndim; k++)
        if (sizes[k] < 0) return fail(CCSD_ERR_SHAPE, "negative size");
    hipError_t e = ccsd_launch_permute(src, ndim, sizes, strides, dst, hi, lo, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "permute_kernel");
    return CCSD_OK;
}

int ccsd_combine(int ndim, const long long* sizes, double* dst, const long long* dstrides, int nterms,
                 const double* const* srcs, const long long* sstrides, const double* coef, int beta, void* stream)
{
    if (!sizes || !dst || !dstrides || !srcs || !sstrides || !coef) return fail(CCSD_ERR_NULL, "null pointer");
    if (ndim < 1 || ndim > 6 || nterms < 1 || nterms > 4 || (beta != 0 && beta != 1))
        return fail(CCSD_ERR_SHAPE, "ndim must be 1..6 and nterms 1..4");
    for (int t = 0; t < nterms; t++)
        if (!srcs[t]) return fail(CCSD_ERR_NULL, "null source");
    for (int k = 0; k < ndim; k++)
        if (sizes[k] < 0) return fail(CCSD_ERR_SHAPE, "negative size");
    hipError_t e = ccsd_launch_combine(ndim, sizes, dst, dstrides, nterms, srcs, sstrides, coef, beta,
                                       (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "combine_kernel");
    return CCSD_OK;
}

int ccsd_gemv(int cols, const double* A, long long M, long long K, long long lda, const double* x, long long incx,
              double* y, double alpha, int beta, double* ws, int nslices, void* stream)
{
    if (!A || !x || !y) return fail(CCSD_ERR_NULL, "null pointer");
    if (M < 0 || K < 0 || (beta != 0 && beta != 1) || nslices < 1 || (nslices > 1 && !ws) ||
        lda < (cols ? M : K))
        return fail(CCSD_ERR_SHAPE, "bad GEMV arguments");
    if (K == 0) nslices = 1;   // the kernel writes beta * y (no memset: callers may capture this in a graph)
    if ((cols ? (M + 255) / 256 * nslices : (M * nslices + 3) / 4) > 0x7fffffffLL)
        return fail(CCSD_ERR_SHAPE, "GEMV too large for one launch");
    hipError_t e = ccsd_launch_gemv(cols, A, M, K, lda, x, incx, y, alpha, beta, ws, nslices, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "gemv_kernel");
    if (nslices > 1) {
        e = ccsd_launch_sum_slices(ws, y, M, nslices, beta, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "sum_slices_kernel");
    }
    return CCSD_OK;
}

int ccsd_gemv_noalloc(int cols, const double* A, long long M, long long K, long long lda, const double* x,
                      long long incx, double* y, double alpha, int beta, int nslices, void* stream)
{
    if (!A || !x || !y) return fail(CCSD_ERR_NULL, "null pointer");
    if (M < 0 || K < 0 || (beta != 0 && beta != 1) || nslices < 1 || lda < (cols ? M : K))
        return fail(CCSD_ERR_SHAPE, "bad GEMV arguments");
    if (K == 0) nslices = 1;
    if (nslices > 1 && (long long)nslices * M > CCSD_GEMV_SCRATCH_DOUBLES)
        return fail(CCSD_ERR_SHAPE, "nslices * M = " + std::to_string((long long)nslices * M) +
                                        " exceeds the GEMV scratch of " +
                                        std::to_string((long long)CCSD_GEMV_SCRATCH_DOUBLES) + " doubles");
    if ((cols ? (M + 255) / 256 * nslices : (M * nslices + 3) / 4) > 0x7fffffffLL)
        return fail(CCSD_ERR_SHAPE, "GEMV too large for one launch");
    double* ws = nullptr;
    if (nslices > 1) {
        const int rc = scratch_address(&ws);
        if (rc != CCSD_OK) return rc;
    }
    // every slice partial (nslices * M) is written by the GEMV kernel before the slice
    // sum reads it, so the scratch needs no clearing between calls
    hipError_t e = ccsd_launch_gemv(cols, A, M, K, lda, x, incx, y, alpha, beta, ws, nslices, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "gemv_kernel");
    if (nslices > 1) {
        e = ccsd_launch_sum_slices(ws, y, M, nslices, beta, (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "sum_slices_kernel");
    }
    return CCSD_OK;
}

int ccsd_pack_pairs_sym(const double* t, const long long* kk, const long long* ll, int n, int m, long long npair,
                        double* out, void* stream)
{
    if (!t || !kk || !ll || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 0 || m < 0 || npair < 0 || npair > (long long)n * (n + 1) / 2) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    hipError_t e = ccsd_launch_pack_pairs_sym(t, kk, ll, n, m, npair, out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "pack_pairs_sym_kernel");
    return CCSD_OK;
}

int ccsd_repack_pairs(const double* z, int n, int m, double* out, void* stream)
{
    if (!z || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 0 || m < 0 || (long long)n * (n + 1) / 2 > 0x7fffffffLL) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    hipError_t e = ccsd_launch_repack_pairs(z, n, m, out, 0, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "repack_pairs_kernel");
    return CCSD_OK;
}

int ccsd_repack_pairs_f32(const float* z, int n, int m, float* out, void* stream)
{
    if (!z || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 0 || m < 0 || (long long)n * (n + 1) / 2 > 0x7fffffffLL) return fail(CCSD_ERR_SHAPE, "bad dimensions");
    hipError_t e = ccsd_launch_repack_pairs(z, n, m, out, 1, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "repack_pairs_kernel");
    return CCSD_OK;
}

int ccsd_add_separable_packed(double* gamma, const double* dm, const double* pm, int n, void* stream)
{
    if (!gamma || !dm || !pm) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 0 || n > 2048) return fail(CCSD_ERR_SHAPE, "n must be 0..2048");
    hipError_t e = ccsd_launch_add_separable_packed(gamma, dm, pm, n, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "add_separable_packed_kernel");
    return CCSD_OK;
}

int ccsd_sum_slices(double* ws, double* out, long long n, int nslices, int beta, void* stream)
{
    if (!ws || !out) return fail(CCSD_ERR_NULL, "null pointer");
    if (n < 0 || nslices < 0 || (beta != 0 && beta != 1)) return fail(CCSD_ERR_SHAPE, "bad slice-sum arguments");
    hipError_t e = ccsd_launch_sum_slices(ws, out, n, nslices, beta, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "sum_slices_kernel");
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
                                              batch, beta, 1.0, nullptr, nullptr, nullptr, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "gemm_bf16x3_nt_kernel");
    return CCSD_OK;
}

int ccsd_gemm_bf16x3_nt_scatter(const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, double* C, int M,
                                int N, int K, long long lda, long long ldb, long long strideA, long long strideB,
                                const long long* off_b, const long long* off_r, const long long* off_c, int batch,
                                double alpha, int beta, void* stream)
{
    if (!Ahi || !Alo || !Bhi || !Blo || !C || !off_b || !off_r || !off_c) return fail(CCSD_ERR_NULL, "null pointer");
    if (M < 0 || N < 0 || K < 0 || batch < 0 || lda < K || ldb < K || (beta != 0 && beta != 1))
        return fail(CCSD_ERR_SHAPE, "bad GEMM arguments");
    if ((N + 127) / 128 > 65535 || batch > 65535) return fail(CCSD_ERR_SHAPE, "N or batch too large for one launch");
    hipError_t e = ccsd_launch_gemm_bf16x3_nt(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, strideA, strideB, 0, 0, batch,
                                              beta, alpha, off_b, off_r, off_c, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "gemm_bf16x3_nt_kernel");
    return CCSD_OK;
}

int ccsd_gemm_set_variant(int v)
{
    if (v < 0 || v > 2) return fail(CCSD_ERR_SHAPE, "GEMM variant must be 0, 1 or 2");
    g_gemm_variant = v;
    return CCSD_OK;
}

int ccsd_split_bf16_strided(const double* x, void* hi, void* lo, int ndim, const long long* sizes,
                            const long long* strides, void* stream)
{
    if (!x || !hi || !lo || !sizes || !strides) return fail(CCSD_ERR_NULL, "null pointer");
    if (ndim < 1 || ndim > 8) return fail(CCSD_ERR_SHAPE, "ndim must be 1..8");
    for (int d = 0; d < ndim; d++)
        if (sizes[d] < 0 || strides[d] < 0) return fail(CCSD_ERR_SHAPE, "negative size or stride");
    hipError_t e = ccsd_launch_split_bf16_strided(x, hi, lo, ndim, sizes, strides, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "split_bf16_strided_kernel");
    return CCSD_OK;
}

int ccsd_scatter_sum(const double* ws, int nslices, int batch, int M, int N, double* C, const long long* off_b,
                     const long long* off_r, const long long* off_c, double alpha, int beta, void* stream)
{
    if (!ws || !C || !off_b || !off_r || !off_c) return fail(CCSD_ERR_NULL, "null pointer");
    if (nslices < 0 || batch < 0 || M < 0 || N < 0 || (beta != 0 && beta != 1))
        return fail(CCSD_ERR_SHAPE, "bad scatter-sum arguments");
    hipError_t e = ccsd_launch_scatter_sum(ws, nslices, batch, M, N, C, off_b, off_r, off_c, alpha, beta,
                                           (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "scatter_sum_kernel");
    return CCSD_OK;
}

int ccsd_plan_set_option(ccsd_plan* p, int option, int value)
{
    int rc = check_plan(p);
    if (rc) return rc;
    if (option == 0) {
        p->fused_reg = value != 0;
        clear_bra_cache(p);
        p->choice.clear();
        p->choice_lds.clear();
        return CCSD_OK;
    }
    if (option == 1) {
        p->autotune = value != 0;
        clear_bra_cache(p);
        p->choice.clear();
        p->choice_lds.clear();
        return CCSD_OK;
    }
    if (option == 2) {
        p->concurrent = value != 0;
        return CCSD_OK;
    }
    if (option == 3) {
        p->ket_shell_order = value != 0;
        clear_bra_cache(p);
        p->choice.clear();
        p->choice_lds.clear();
        return build_ket(p);
    }
    if (option == 4) {
        if (value < 0 || value > 254) return fail(CCSD_ERR_SHAPE, "primitive split must be 0..254");
        p->force_split = value;
        clear_bra_cache(p);
        p->choice.clear();
        p->choice_lds.clear();
        return CCSD_OK;
    }
    if (option == 5) {
        if (value < 0 || value > 2) return fail(CCSD_ERR_SHAPE, "persistent mode must be 0, 1 or 2");
        p->persistent = value;
        clear_bra_cache(p);
        return CCSD_OK;
    }
    if (option == 7) {
        p->persist_order = value != 0;
        clear_bra_cache(p);
        return CCSD_OK;
    }
    if (option == 6) {
        p->tune_generic = value != 0;
        clear_bra_cache(p);
        p->choice.clear();
        p->choice_lds.clear();
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
