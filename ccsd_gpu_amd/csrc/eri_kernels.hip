// Rys-quadrature ERI kernels for gfx950 (CDNA4), FP64.
//
// One 64-lane wavefront evaluates the shell quartets (ish jsh | ksh lsh) of one
// work chunk (a bra pair and a run of ket pairs of one angular-momentum class)
// one after the other; classes are launched separately with their own LDS
// size. Work inside a quartet is spread over the wave in three phases:
//
//   A0  lanes = primitive quartets: Gaussian-product data, Boys argument T,
//       Rys roots and weights (Chebyshev tables, ccsd_gpu/cuda/int2e_ip1_module.py:192-244
//       computes the same quantities on one thread) -> LDS
//   A1  lanes = (primitive quartet, root, direction): the 1-D factor
//       I_x(i j | k l) by VRR on centre i, TRR i->k, HRR i->j, HRR k->l
//       (int2e_ip1_module.py:466-587), each lane in its own LDS slab, no barriers
//   B   lanes = (item group g, output unit o): sum over roots and primitives of
//       Ix Iy Iz (int2e, int2e_module.py:409-430) or of the d/dr_i derivative
//       f_x = -2 a_i g[ix+1] + ix g[ix-1] (int2e_ip1_module.py:612-629),
//       partial sums kept in an LDS table owned lane-by-lane
//
// then a reduction over item groups and either a Cartesian s2kl store
// (int2e_ip1_module.py:636-663 layout), a Schwarz maximum over the diagonal
// (ab|ab) of a pair with itself, or, for the fused gradient, a dot with the
// two-particle density (explicit or separable) and one atomic add per chunk
// into dE/dR.
#include "eri_common.h"

#include "rys_device.h"

namespace ccsd {

__device__ const double kShellNorm[2] = {0.28209479177387814347, 0.48860251190291992159};
// Cartesian components (lx, ly, lz) packed as lx | ly << 4 | lz << 8, l = 0..5,
// PySCF order; offset of l is l(l+1)(l+2)/6.
__device__ const int kCart[56] = {
    0x000,
    0x001, 0x010, 0x100,
    0x002, 0x011, 0x101, 0x020, 0x110, 0x200,
    0x003, 0x012, 0x102, 0x021, 0x111, 0x201, 0x030, 0x120, 0x210, 0x300,
    0x004, 0x013, 0x103, 0x022, 0x112, 0x202, 0x031, 0x121, 0x211, 0x301, 0x040, 0x130, 0x220, 0x310, 0x400,
    0x005, 0x014, 0x104, 0x023, 0x113, 0x203, 0x032, 0x122, 0x212, 0x302,
    0x041, 0x131, 0x221, 0x311, 0x401, 0x050, 0x140, 0x230, 0x320, 0x410, 0x500};

__device__ __forceinline__ int cart_comp(int l, int f) { return kCart[l * (l + 1) * (l + 2) / 6 + f]; }
__device__ __forceinline__ int ncart_dev(int l) { return (l + 1) * (l + 2) / 2; }

// LDS layout (doubles) of the generic quartet kernel; sizes from the launch.
struct LdsLayout {
    int bra_off, ket_off, cij_off, ckl_off, pq_off, g_off, red_off, total;
    int pq_stride, gsize;
};

__host__ __device__ inline LdsLayout lds_layout(const EriClassArgs& a)
{
    LdsLayout L;
    const int d = a.deriv;
    const int sj = a.li + 1 + d, sk = sj * (a.lj + 1), sl = sk * (a.lk + 1);
    L.gsize = sl * (a.ll + 1);
    L.pq_stride = 2 * a.nroots + 12;
    const int ncomp = d ? 3 : 1;
    int off = 0;
    L.bra_off = off; off += 6 * a.max_nij;           // aij, Px, Py, Pz, Kab, ai
    L.ket_off = off; off += 5 * a.max_nkl;           // akl, Qx, Qy, Qz, Kcd
    L.cij_off = off; off += a.max_ncij * a.max_nij;  // contraction-coefficient products
    L.ckl_off = off; off += a.max_nckl * a.max_nkl;
    L.pq_off = off; off += L.pq_stride * a.pq_batch;
    L.g_off = off; off += 3 * L.gsize * a.item_batch;
    L.red_off = off; off += a.groups * a.units_pass * ncomp;
    L.total = off;
    return L;
}

// One shell quartet (bra pair bq | ket pair kq); derivative sums for the fused
// gradient accumulate into gdx, gdy, gdz (lane-private).
__device__ __forceinline__ void quartet(const EriClassArgs& a, const LdsLayout& L, double* lds, int bq, int kq,
                                        double& gdx, double& gdy, double& gdz)
{
    const int lane = threadIdx.x;
    const int2 bp = a.bra[bq];
    const int2 kp = a.ket[kq];
    const ShellInfo si = a.shells[bp.x], sj = a.shells[bp.y];
    const ShellInfo sk = a.shells[kp.x], sl = a.shells[kp.y];
    const int li = a.li, lj = a.lj, lk = a.lk, ll = a.ll, d = a.deriv;
    const int n = a.nroots;
    const int ncomp = d ? 3 : 1;

    const double Ax = a.coords[3 * si.atom], Ay = a.coords[3 * si.atom + 1], Az = a.coords[3 * si.atom + 2];
    const double Bx = a.coords[3 * sj.atom], By = a.coords[3 * sj.atom + 1], Bz = a.coords[3 * sj.atom + 2];
    const double Cx = a.coords[3 * sk.atom], Cy = a.coords[3 * sk.atom + 1], Cz = a.coords[3 * sk.atom + 2];
    const double Dx = a.coords[3 * sl.atom], Dy = a.coords[3 * sl.atom + 1], Dz = a.coords[3 * sl.atom + 2];
    const double ABx = Bx - Ax, ABy = By - Ay, ABz = Bz - Az;
    const double CDx = Dx - Cx, CDy = Dy - Cy, CDz = Dz - Cz;
    const double rab2 = ABx * ABx + ABy * ABy + ABz * ABz;
    const double rcd2 = CDx * CDx + CDy * CDy + CDz * CDz;

    const int nij = si.nprim * sj.nprim, nkl = sk.nprim * sl.nprim;
    const int ncij = si.nctr * sj.nctr, nckl = sk.nctr * sl.nctr;
    const int ncombo = ncij * nckl;
    double* bra = lds + L.bra_off;
    double* ket = lds + L.ket_off;
    double* cij = lds + L.cij_off;
    double* ckl = lds + L.ckl_off;
    double* pqd = lds + L.pq_off;
    double* g = lds + L.g_off;
    double* red = lds + L.red_off;

    // ---- primitive-pair data --------------------------------------------
    const double* ei = a.env + si.pexp; const double* ej = a.env + sj.pexp;
    const double* ek = a.env + sk.pexp; const double* el = a.env + sl.pexp;
    const double* ci = a.env + si.pcoef; const double* cj = a.env + sj.pcoef;
    const double* ck = a.env + sk.pcoef; const double* cl = a.env + sl.pcoef;
    for (int p = lane; p < nij; p += 64) {
        const int ip = p / sj.nprim, jp = p - ip * sj.nprim;
        const double ai = ei[ip], aj = ej[jp], aij = ai + aj, r = aj / aij;
        double* b = bra + 6 * p;
        b[0] = aij; b[1] = Ax + ABx * r; b[2] = Ay + ABy * r; b[3] = Az + ABz * r;
        b[4] = exp(-ai * aj / aij * rab2); b[5] = ai;
        for (int c1 = 0; c1 < si.nctr; c1++)
            for (int c2 = 0; c2 < sj.nctr; c2++)
                cij[(c1 * sj.nctr + c2) * a.max_nij + p] = ci[c1 * si.nprim + ip] * cj[c2 * sj.nprim + jp];
    }
    for (int p = lane; p < nkl; p += 64) {
        const int kp2 = p / sl.nprim, lp = p - kp2 * sl.nprim;
        const double ak = ek[kp2], al = el[lp], akl = ak + al, r = al / akl;
        double* b = ket + 5 * p;
        b[0] = akl; b[1] = Cx + CDx * r; b[2] = Cy + CDy * r; b[3] = Cz + CDz * r;
        b[4] = exp(-ak * al / akl * rcd2);
        for (int c1 = 0; c1 < sk.nctr; c1++)
            for (int c2 = 0; c2 < sl.nctr; c2++)
                ckl[(c1 * sl.nctr + c2) * a.max_nkl + p] = ck[c1 * sk.nprim + kp2] * cl[c2 * sl.nprim + lp];
    }
    const double norm = 34.98683665524972497 * (li < 2 ? kShellNorm[li] : 1.0) * (lj < 2 ? kShellNorm[lj] : 1.0) *
                        (lk < 2 ? kShellNorm[lk] : 1.0) * (ll < 2 ? kShellNorm[ll] : 1.0);

    // ---- geometry of the 1-D factors -------------------------------------
    const int sJ = li + 1 + d, sK = sJ * (lj + 1), sL = sK * (lk + 1);
    const int gsize = L.gsize;
    const int lij = li + lj + d, lkl = lk + ll;
    const int nfi = ncart_dev(li), nfj = ncart_dev(lj), nfk = ncart_dev(lk), nfl = ncart_dev(ll);
    const int nf = nfi * nfj * nfk * nfl;
    const int nunits = ncombo * nf;
    const int G = a.groups, O = 64 / G;
    const int grp = lane / O, ol = lane - grp * O;
    const int npq = nij * nkl;
    const bool same_kl = (kp.x == kp.y);

    for (int u0 = 0; u0 < nunits; u0 += a.units_pass) {
        const int u1 = min(nunits, u0 + a.units_pass);
        const int nu = u1 - u0;
        for (int t = lane; t < G * a.units_pass * ncomp; t += 64) red[t] = 0.0;
        // (the barrier at the top of the pq loop orders these stores)
        for (int pq0 = 0; pq0 < npq; pq0 += a.pq_batch) {
            const int npb = min(a.pq_batch, npq - pq0);
            __syncthreads();
            // -- A0: roots and weights per primitive quartet --
            if (lane < npb) {
                const int pq = pq0 + lane;
                const int pi = pq / nkl, pk = pq - pi * nkl;
                const double* b = bra + 6 * pi;
                const double* k = ket + 5 * pk;
                const double aij = b[0], akl = k[0];
                const double PQx = b[1] - k[1], PQy = b[2] - k[2], PQz = b[3] - k[3];
                const double rr = PQx * PQx + PQy * PQy + PQz * PQz;
                const double T = aij * akl / (aij + akl) * rr;
                double* pd = pqd + lane * L.pq_stride;
                rys_roots_dev(n, T, pd, pd + n);
                double pref = norm * b[4] * k[4] / (aij * akl * sqrt(aij + akl));
                if (ncombo == 1) pref *= cij[pi] * ckl[pk];
                for (int r = 0; r < n; r++) pd[n + r] *= pref;
                double* e = pd + 2 * n;
                e[0] = aij; e[1] = akl; e[2] = b[5];
                e[3] = b[1] - Ax; e[4] = b[2] - Ay; e[5] = b[3] - Az;   // P - A
                e[6] = k[1] - Cx; e[7] = k[2] - Cy; e[8] = k[3] - Cz;   // Q - C
                e[9] = PQx; e[10] = PQy; e[11] = PQz;
            }
            const int nitems = npb * n;
            for (int it0 = 0; it0 < nitems; it0 += a.item_batch) {
                const int nib = min(a.item_batch, nitems - it0);
                __syncthreads();
                // -- A1: one 1-D factor per lane --
                for (int c = lane; c < nib * 3; c += 64) {
                    const int item = it0 + c / 3, dir = c - (c / 3) * 3;
                    const int pl = item / n, r = item - pl * n;
                    const double* pd = pqd + pl * L.pq_stride;
                    const double* e = pd + 2 * n;
                    const double rt = pd[r];
                    const double aij = e[0], akl = e[1];
                    const double a2 = rt / (aij + akl);
                    const double rt_aij = a2 * akl, rt_akl = a2 * aij;
                    const double b00 = 0.5 * a2;
                    const double b10 = 0.5 / aij * (1.0 - rt_aij);
                    const double b01 = 0.5 / akl * (1.0 - rt_akl);
                    const double PQn = e[9 + dir];
                    const double c00 = e[3 + dir] - rt_aij * PQn;
                    const double cp = e[6 + dir] + rt_akl * PQn;
                    const double ABn = dir == 0 ? ABx : (dir == 1 ? ABy : ABz);
                    const double CDn = dir == 0 ? CDx : (dir == 1 ? CDy : CDz);
                    double* gg = g + (size_t)(c) * gsize;
                    gg[0] = dir == 2 ? pd[n + r] : 1.0;
                    if (lij > 0) gg[1] = c00 * gg[0];
                    for (int i = 1; i < lij; i++) gg[i + 1] = c00 * gg[i] + i * b10 * gg[i - 1];
                    for (int kk = 0; kk < lkl; kk++) {
                        const double* g0 = gg + kk * sK;
                        double* g1 = gg + (kk + 1) * sK;
                        for (int i = 0; i <= lij; i++) {
                            double s = cp * g0[i];
                            if (kk > 0) s += kk * b01 * g0[i - sK];
                            if (i > 0) s += i * b00 * g0[i - 1];
                            g1[i] = s;
                        }
                    }
                    // HRR i -> j, in place per k row (reads precede writes)
                    for (int kk = 0; kk <= lkl; kk++) {
                        double* gk = gg + kk * sK;
                        for (int j = 0; j < lj; j++) {
                            int ij = (lij - j) + j * sJ;
                            double s1 = gk[ij];
                            for (--ij; ij >= j * sJ; --ij) {
                                const double s0 = gk[ij];
                                gk[ij + sJ] = s1 - ABn * s0;
                                s1 = s0;
                            }
                        }
                    }
                    // HRR k -> l, in place per (i, j) column
                    if (ll > 0) {
                        for (int m = 0; m < sK; m++) {
                            double* gm = gg + m;
                            for (int l2 = 0; l2 < ll; l2++) {
                                int kl = (lkl - l2) * sK + l2 * sL;
                                double s1 = gm[kl];
                                for (kl -= sK; kl >= l2 * sL; kl -= sK) {
                                    const double s0 = gm[kl];
                                    gm[kl + sL] = s1 - CDn * s0;
                                    s1 = s0;
                                }
                            }
                        }
                    }
                }
                __syncthreads();
                // -- B: accumulate owned output units over this sub-batch --
                for (int u = u0 + ol; u < u1; u += O) {
                    const int combo = u / nf, f = u - combo * nf;
                    const int fi = f % nfi, r1 = f / nfi;
                    const int fj = r1 % nfj, r2 = r1 / nfj;
                    const int fk = r2 % nfk, fl = r2 / nfk;
                    const int ci_ = cart_comp(li, fi), cj_ = cart_comp(lj, fj);
                    const int ck_ = cart_comp(lk, fk), cl_ = cart_comp(ll, fl);
                    const int ix = ci_ & 15, iy = (ci_ >> 4) & 15, iz = ci_ >> 8;
                    const int ax = ix + (cj_ & 15) * sJ + (ck_ & 15) * sK + (cl_ & 15) * sL;
                    const int ay = iy + ((cj_ >> 4) & 15) * sJ + ((ck_ >> 4) & 15) * sK + ((cl_ >> 4) & 15) * sL + gsize;
                    const int az = iz + (cj_ >> 8) * sJ + (ck_ >> 8) * sK + (cl_ >> 8) * sL + 2 * gsize;
                    const int cc_ij = combo / nckl, cc_kl = combo - cc_ij * nckl;
                    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
                    for (int item = grp; item < nib; item += G) {
                        const int gi = it0 + item;
                        const double* gg = g + (size_t)item * 3 * gsize;
                        double cf = 1.0;
                        if (ncombo > 1) {
                            const int pl = gi / n;
                            const int pq = pq0 + pl;
                            const int pi = pq / nkl, pk = pq - pi * nkl;
                            cf = cij[cc_ij * a.max_nij + pi] * ckl[cc_kl * a.max_nkl + pk];
                        }
                        if (d == 0) {
                            s0 += cf * gg[ax] * gg[ay] * gg[az];
                        } else {
                            const int pl = gi / n;
                            const double ai2 = 2.0 * pqd[pl * L.pq_stride + 2 * n + 2];
                            const double gx = gg[ax], gy = gg[ay], gz = gg[az];
                            double dx = -ai2 * gg[ax + 1];
                            double dy = -ai2 * gg[ay + 1];
                            double dz = -ai2 * gg[az + 1];
                            if (ix > 0) dx += ix * gg[ax - 1];
                            if (iy > 0) dy += iy * gg[ay - 1];
                            if (iz > 0) dz += iz * gg[az - 1];
                            s0 += cf * dx * gy * gz;
                            s1 += cf * gx * dy * gz;
                            s2 += cf * gx * gy * dz;
                        }
                    }
                    double* rp = red + ((size_t)grp * a.units_pass + (u - u0)) * ncomp;
                    rp[0] += s0;
                    if (d) { rp[1] += s1; rp[2] += s2; }
                }
            }
        }
        __syncthreads();
        // -- reduce over item groups, store or contract --
        for (int uu = lane; uu < nu; uu += 64) {
            const int u = u0 + uu;
            double v[3] = {0.0, 0.0, 0.0};
            for (int gi = 0; gi < G; gi++)
                for (int c = 0; c < ncomp; c++) v[c] += red[((size_t)gi * a.units_pass + uu) * ncomp + c];
            const int combo = u / nf, f = u - combo * nf;
            const int fi = f % nfi, r1 = f / nfi;
            const int fj = r1 % nfj, r2 = r1 / nfj;
            const int fk = r2 % nfk, fl = r2 / nfk;
            const int cc_ij = combo / nckl, cc_kl = combo - cc_ij * nckl;
            const int c_i = cc_ij / sj.nctr, c_j = cc_ij - c_i * sj.nctr;
            const int c_k = cc_kl / sl.nctr, c_l = cc_kl - c_k * sl.nctr;
            const int i = si.ao_cart + c_i * nfi + fi - a.i_base;
            const int j = sj.ao_cart + c_j * nfj + fj;
            const int k = sk.ao_cart + c_k * nfk + fk;
            const int l = sl.ao_cart + c_l * nfl + fl;
            if (same_kl && l > k) continue;
            if (a.out) {
                const long long idx = ((long long)i * a.nao + j) * a.npair + (long long)k * (k + 1) / 2 + l;
                for (int c = 0; c < ncomp; c++) a.out[c * a.comp_stride + idx] = v[c];
            } else if (a.qmax) {
                // diagonal element (ab|ab) of a bra pair with itself: Schwarz factor
                if (i == k && j == l) {
                    const double qv = sqrt(fmax(v[0], 0.0));
                    atomicMax(a.qmax + bp.x * a.nbas + bp.y, (unsigned long long)__double_as_longlong(qv));
                }
            } else {
                const double gm = gamma_packed(a, si.mol, i, j, k, l);
                gdx += v[0] * gm; gdy += v[1] * gm; gdz += v[2] * gm;
            }
        }
        __syncthreads();
    }
}

// One block (one wave) per chunk: bra pair ch.x with ket pairs [ch.y, ch.z).
__global__ void __launch_bounds__(64) eri_quartet_kernel(const EriClassArgs a)
{
    extern __shared__ double lds[];
    const LdsLayout L = lds_layout(a);
    const int4 ch = a.chunks[blockIdx.x];
    double gdx = 0.0, gdy = 0.0, gdz = 0.0;
    for (int kq = ch.y; kq < ch.z; kq++) {
        __syncthreads();   // LDS of the previous quartet fully consumed
        quartet(a, L, lds, ch.x, kq, gdx, gdy, gdz);
    }
    if (!a.out && !a.qmax) {
        // wave reduction of the three derivative sums, one atomic per component
        for (int off = 32; off > 0; off >>= 1) {
            gdx += __shfl_down(gdx, off, 64);
            gdy += __shfl_down(gdy, off, 64);
            gdz += __shfl_down(gdz, off, 64);
        }
        if (threadIdx.x == 0) {
            const int atom = a.shells[a.bra[ch.x].x].atom;
            defx_add(a.defx, atom, a.gamma_scale * gdx, a.gamma_scale * gdy, a.gamma_scale * gdz);
        }
    }
}

// Wave reduction of a chunk's three derivative sums, added (fixed point) into dE/dR of
// the bra pair's first atom (all lanes call it; lane 0 adds).
__device__ __forceinline__ void add_chunk_gradient(const EriClassArgs& a, int bq, double gdx, double gdy, double gdz)
{
    for (int off = 32; off > 0; off >>= 1) {
        gdx += __shfl_down(gdx, off, 64);
        gdy += __shfl_down(gdy, off, 64);
        gdz += __shfl_down(gdz, off, 64);
    }
    if (threadIdx.x == 0) {
        const int atom = a.shells[a.bra[bq].x].atom;
        defx_add(a.defx, atom, a.gamma_scale * gdx, a.gamma_scale * gdy, a.gamma_scale * gdz);
    }
}

// Per-call arguments of the persistent fused pass (the class table is static).
struct PersistDyn {
    const double* gamma;
    const double* dm;
    const double* pm;
    unsigned long long* defx;
    double scale;
    int gamma_mode;
};

// Fused int2e_ip1 x Gamma over every class of a pass in ONE launch. Each wave takes
// work items from a global counter until the list is exhausted; item t covers
// chunks [items[t].y, items[t].z) of class items[t].x (class table `classes`: the
// launch description of each class, work list included). Items are ordered by
// decreasing estimated cost, so the last waves to finish run short items. `counter`
// must read 0 at launch; every wave leaves the loop once it draws t >= nitems. ticks
// (optional, nitems words): each item's duration in wall_clock64 ticks.
__global__ void __launch_bounds__(64) eri_persistent_kernel(const EriClassArgs* __restrict__ classes,
                                                            const int4* __restrict__ items, int nitems,
                                                            unsigned int* __restrict__ counter,
                                                            unsigned int* __restrict__ ticks, const PersistDyn dyn)
{
    extern __shared__ double lds[];
    for (;;) {
        const unsigned long long t_start = wall_clock64();
        int t = 0;
        if (threadIdx.x == 0) t = (int)atomicAdd(counter, 1u);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
        if (t >= nitems) break;
        const int4 it = items[t];
        const int cls = __builtin_amdgcn_readfirstlane(it.x);
        const int c0 = __builtin_amdgcn_readfirstlane(it.y), c1 = __builtin_amdgcn_readfirstlane(it.z);
        EriClassArgs a = classes[cls];
        a.gamma = dyn.gamma;
        a.dm = dyn.dm;
        a.pm = dyn.pm;
        a.defx = dyn.defx;
        a.gamma_scale = dyn.scale;
        a.gamma_mode = dyn.gamma_mode;
        const LdsLayout L = lds_layout(a);
        for (int c = c0; c < c1; c++) {
            const int4 ch = a.chunks[c];
            double gdx = 0.0, gdy = 0.0, gdz = 0.0;
            for (int kq = ch.y; kq < ch.z; kq++) {
                __syncthreads();   // LDS of the previous quartet (or item) fully consumed
                quartet(a, L, lds, ch.x, kq, gdx, gdy, gdz);
            }
            add_chunk_gradient(a, ch.x, gdx, gdy, gdz);
        }
        if (ticks && threadIdx.x == 0) {
            const unsigned long long dt = wall_clock64() - t_start;
            ticks[t] = dt > 0xffffffffULL ? 0xffffffffu : (unsigned int)dt;
        }
        __syncthreads();
    }
}

// Cartesian -> spherical transform of an s2kl tensor, one thread per output.
// in[comp][i_c][j_c][pair_c], out[comp][i_s][j_s][pair_s]; per AO a list of
// (Cartesian AO, coefficient) terms, at most 6 (real solid harmonics, l <= 4).
struct C2sTerm { int n; int idx[6]; double c[6]; };

__global__ void c2s_s2kl_kernel(double* __restrict__ out, const double* __restrict__ in,
                                const C2sTerm* __restrict__ ti, const C2sTerm* __restrict__ tall,
                                int nfi_s, int nao_s, long long npair_s, int nao_c, long long npair_c,
                                int nfi_c, int ncomp)
{
    const long long total = (long long)ncomp * nfi_s * nao_s * npair_s;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long r = t;
        const long long kl = r % npair_s; r /= npair_s;
        const int j = (int)(r % nao_s); r /= nao_s;
        const int i = (int)(r % nfi_s);
        const int comp = (int)(r / nfi_s);
        int k = (int)((sqrt(8.0 * (double)kl + 1.0) - 1.0) * 0.5);
        while ((long long)k * (k + 1) / 2 > kl) k--;
        while ((long long)(k + 1) * (k + 2) / 2 <= kl) k++;
        const int l = (int)(kl - (long long)k * (k + 1) / 2);
        const C2sTerm A = ti[i], B = tall[j], C = tall[k], D = tall[l];
        const double* src = in + (long long)comp * nfi_c * nao_c * npair_c;
        double acc = 0.0;
        for (int a = 0; a < A.n; a++)
            for (int b = 0; b < B.n; b++) {
                const double cab = A.c[a] * B.c[b];
                const double* row = src + ((long long)A.idx[a] * nao_c + B.idx[b]) * npair_c;
                for (int c = 0; c < C.n; c++)
                    for (int e = 0; e < D.n; e++) {
                        const int kc = C.idx[c], lc = D.idx[e];
                        const long long p = kc >= lc ? (long long)kc * (kc + 1) / 2 + lc : (long long)lc * (lc + 1) / 2 + kc;
                        acc += cab * C.c[c] * D.c[e] * row[p];
                    }
            }
        out[t] = acc;
    }
}

__global__ void fill_kernel(double* __restrict__ p, long long n, double v)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        p[t] = v;
}

__global__ void rys_roots_kernel(int n, const double* __restrict__ T, int count, double* __restrict__ r,
                                 double* __restrict__ w)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) rys_roots_dev(n, T[t], r + (size_t)t * n, w + (size_t)t * n);
}

}  // namespace ccsd

// ---- launch helpers (host) -----------------------------------------------
size_t ccsd_eri_lds_bytes(const EriClassArgs& a) { return sizeof(double) * ccsd::lds_layout(a).total; }

hipError_t ccsd_launch_eri_class(const EriClassArgs& a, hipStream_t stream)
{
    if (a.nchunks <= 0) return hipSuccess;
    const size_t lds = ccsd_eri_lds_bytes(a);
    hipLaunchKernelGGL(ccsd::eri_quartet_kernel, dim3((unsigned)a.nchunks), dim3(64), lds, stream, a);
    return hipGetLastError();
}

// Resident one-wave blocks of the persistent kernel per CU at `lds` bytes of LDS.
int ccsd_persistent_blocks_per_cu(size_t lds)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ccsd::eri_persistent_kernel, 64, lds) != hipSuccess)
        return 0;
    return nb;
}

namespace ccsd {
// de[3 natm] += the fixed-point sums of defx (h 2^-20 + l 2^-60 per component); NaN
// everywhere when any contribution overflowed.
__global__ void defx_to_de_kernel(const unsigned long long* __restrict__ defx, int n3, double* __restrict__ de)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    const long long h = (long long)defx[1 + 2 * t], l = (long long)defx[2 + 2 * t];
    const double v = (double)h * 9.5367431640625e-07 + (double)l * 8.673617379884035e-19;   // 2^-20, 2^-60
    de[t] = defx[0] ? __builtin_nan("") : de[t] + v;
}
}  // namespace ccsd

hipError_t ccsd_launch_defx_to_de(const unsigned long long* defx, int natm, double* de, hipStream_t stream)
{
    const int n3 = 3 * natm;
    if (n3 <= 0) return hipSuccess;
    hipLaunchKernelGGL(ccsd::defx_to_de_kernel, dim3((n3 + 255) / 256), dim3(256), 0, stream, defx, n3, de);
    return hipGetLastError();
}

hipError_t ccsd_launch_eri_persistent(const EriClassArgs* classes, const int4* items, int nitems,
                                      unsigned int* counter, unsigned int* ticks, const double* gamma,
                                      const double* dm, const double* pm, unsigned long long* defx, double scale,
                                      int gamma_mode, int nblocks, size_t lds, hipStream_t stream)
{
    if (nitems <= 0 || nblocks <= 0) return hipSuccess;
    ccsd::PersistDyn dyn;
    dyn.gamma = gamma;
    dyn.dm = dm;
    dyn.pm = pm;
    dyn.defx = defx;
    dyn.scale = scale;
    dyn.gamma_mode = gamma_mode;
    hipLaunchKernelGGL(ccsd::eri_persistent_kernel, dim3((unsigned)nblocks), dim3(64), lds, stream, classes, items,
                       nitems, counter, ticks, dyn);
    return hipGetLastError();
}

hipError_t ccsd_launch_c2s(double* out, const double* in, const void* ti, const void* tall, int nfi_s, int nao_s,
                           long long npair_s, int nao_c, long long npair_c, int nfi_c, int ncomp, hipStream_t stream)
{
    const long long total = (long long)ncomp * nfi_s * nao_s * npair_s;
    if (total == 0) return hipSuccess;
    long long nb = (total + 255) / 256;
    if (nb > 65536LL * 16) nb = 65536LL * 16;
    hipLaunchKernelGGL(ccsd::c2s_s2kl_kernel, dim3((unsigned)nb), dim3(256), 0, stream, out, in,
                       (const ccsd::C2sTerm*)ti, (const ccsd::C2sTerm*)tall, nfi_s, nao_s, npair_s, nao_c,
                       npair_c, nfi_c, ncomp);
    return hipGetLastError();
}

hipError_t ccsd_launch_rys(int n, const double* T, int count, double* r, double* w, hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(ccsd::rys_roots_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, n, T, count, r, w);
    return hipGetLastError();
}

size_t ccsd_c2s_term_size() { return sizeof(ccsd::C2sTerm); }

hipError_t ccsd_launch_fill(double* p, long long n, double v, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(ccsd::fill_kernel, dim3((unsigned)nb), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}
