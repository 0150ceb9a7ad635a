// v2d_direct.hip -- exact block-tridiagonal LU of the Newton matrix on gfx950.
//
// Ordering the unknowns by radial level makes the Newton matrix block
// tridiagonal: level l (one j row, or two at the wall / far field where the
// Thom and outflow rows reach j +- 2) couples only to levels l-1 and l+1.  The
// factorisation is block Gaussian elimination from both ends at once (a
// twisted factorisation), from the wall outward and from the far field inward:
//     S_l = D_l - L_l S_{l-1}^{-1} U_{l-1}   (l < M),
//     S_l = D_l - U_l S_{l+1}^{-1} L_{l+1}   (l > M),
// meeting at level M, which takes both updates.  The two chains run in the same
// launches, which halves the sequential depth.  Every S_l^{-1} is kept
// explicitly, computed in place by blocked Gauss-Jordan without pivoting.  L and U stay sparse (at most a handful of
// entries per row / column), so the Schur update is O(n^2) per level and the
// dense inversion (2 n^3 flops, n = 2*imax) carries the work.  A solve is two
// sweeps of one dense GEMV per level (see direct_solve).
// With T = double this is the exact Newton step of the reference's MUMPS LU
// (vorti2d/petsc_solver.py:55-62); with T = float it is the mixed-precision
// preconditioner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "v2d_internal.h"

namespace v2d {

constexpr int NB = 64;   // Gauss-Jordan block width == update tile edge

// Global-address-space views of generic pointers: the job tables hand the
// kernels plain pointers, which compile to flat memory instructions, and a
// flat load is waited for together with every LDS access (lgkmcnt); global
// loads are counted in order, so later loads stay in flight.
template <typename X> using gptr = __attribute__((address_space(1))) X *;
template <typename X> __device__ __forceinline__ gptr<X> glb(X *p) { return (gptr<X>)p; }
// Native 16-byte vectors (HIP's float4 is a class type, which cannot be
// copied through an address-space-qualified pointer).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// Two consecutive elements of T in one load (8 bytes in FP32, 16 in FP64).
template <typename T> struct Pair2;
template <> struct Pair2<float> { using type = f32x2; };
template <> struct Pair2<double> { using type = f64x2; };
template <typename T> struct NVec;
template <> struct NVec<float> { using type = f32x4; };
template <> struct NVec<double> { using type = f64x2; };

// ----------------------------------------------- persistent solve sweeps ---
// Level shapes in closed form (levels 0 and nlev-1 hold two j rows, the others
// one); the pattern build checks them against its own tables.
V2D_HD inline int lvl_ld(int n) { return (n + 63) / 64 * 64; }
V2D_HD inline int64_t lvl_voff(const Levels &L, int l) { return (int64_t)2 * L.imax * (l + (l > 0)); }
V2D_HD inline int64_t lvl_soff(const Levels &L, int l)
{
    if (l == 0) return 0;
    const int n0 = L.size(0), n1 = 2 * L.imax;
    return (int64_t)n0 * lvl_ld(n0) + (int64_t)(l - 1) * n1 * lvl_ld(n1);
}

// The levels of sweep step s: forward steps s < steps (both chains toward the
// meeting level M), the middle level at s == steps, then the backward steps
// outward from M.  term: 1 = coupling to level l-1, 2 = to level l+1.
struct SwJob {
    int l, term;
};
V2D_HD inline int sweep_jobs(int s, int nlev, int M, int steps, SwJob (&jb)[2], int &phase)
{
    const int nbot = M, ntop = nlev - 1 - M;
    int nj = 0;
    if (s < steps) {
        phase = 0;
        if (s < nbot) jb[nj++] = SwJob{s, s > 0 ? 1 : 0};
        if (s < ntop) jb[nj++] = SwJob{nlev - 1 - s, s > 0 ? 2 : 0};
    } else if (s == steps) {
        phase = 1;
        jb[nj++] = SwJob{M, (M > 0 ? 1 : 0) | (M < nlev - 1 ? 2 : 0)};
    } else {
        phase = 2;
        const int st = s - steps - 1;
        if (M - 1 - st >= 0) jb[nj++] = SwJob{M - 1 - st, 2};
        if (M + 1 + st < nlev) jb[nj++] = SwJob{M + 1 + st, 1};
    }
    return nj;
}

constexpr int kSwThreads = 1024;              // persistent sweep workgroup
constexpr int kSwRows = kSwThreads / 64;      // rows per wavefront pass
constexpr int kSwRb = 3;                       // right-hand-side rows per thread in one batch

// ------------------------------------------------------------ pattern -----
int direct_build_pattern(DirectPattern &P, int imax, int jmax, const uint8_t *flags, hipStream_t s)
{
    P.L = make_levels(imax, jmax);
    const Levels &L = P.L;
    const int64_t N = (int64_t)imax * jmax;
    P.N = N;
    const int nlev = L.nlev;
    P.n.assign(nlev, 0);
    P.ld.assign(nlev, 0);
    P.ldmax = 0;
    P.voff.assign(nlev + 1, 0);
    P.soff.assign(nlev + 1, 0);
    P.nmax = 0;
    for (int l = 0; l < nlev; ++l) {
        P.n[l] = L.size(l);
        P.ld[l] = (P.n[l] + 63) / 64 * 64;
        P.voff[l + 1] = P.voff[l] + P.n[l];
        P.soff[l + 1] = P.soff[l] + (int64_t)P.n[l] * P.ld[l];
        P.nmax = std::max(P.nmax, P.n[l]);
        P.ldmax = std::max(P.ldmax, P.ld[l]);
    }
    P.sinv_elems = P.soff[nlev];
    if (P.voff[nlev] != 2 * N) {
        set_error("direct solver: level partition does not cover the unknowns");
        return -1;
    }
    const int64_t n2 = 2 * N;
    std::vector<int> cD(n2 + 1, 0), cL(n2 + 1, 0), cU(n2 + 1, 0), cUC(n2 + 1, 0), cLC(n2 + 1, 0);
    // two passes: count, then fill
    std::vector<int2> eD, eL, eU, eUC, eLC;
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int> fD, fL, fU, fUC, fLC;
        if (pass == 1) {
            for (int64_t r = 0; r < n2; ++r) {
                cD[r + 1] += cD[r]; cL[r + 1] += cL[r]; cU[r + 1] += cU[r]; cUC[r + 1] += cUC[r];
                cLC[r + 1] += cLC[r];
            }
            eD.resize(cD[n2]); eL.resize(cL[n2]); eU.resize(cU[n2]); eUC.resize(cUC[n2]);
            eLC.resize(cLC[n2]);
            fD.assign(cD.begin(), cD.end() - 1); fL.assign(cL.begin(), cL.end() - 1);
            fU.assign(cU.begin(), cU.end() - 1); fUC.assign(cUC.begin(), cUC.end() - 1);
            fLC.assign(cLC.begin(), cLC.end() - 1);
        }
        for (int64_t g = 0; g < n2; ++g) {
            const int comp = g >= N;
            const int64_t k = comp ? g - N : g;
            const int j = (int)(k / imax), i = (int)(k % imax);
            const uint8_t fl = j == jmax - 1 ? flags[i] : 0;
            int64_t cols[13];
            const int cnt = row_columns(imax, jmax, i, j, comp, fl, N, cols);
            const int lr = L.level_of_j(j);
            const int r = L.local(g, N, lr);
            const int64_t R = P.voff[lr] + r;
            for (int q = 0; q < cnt; ++q) {
                const int64_t gc = cols[q];
                const int64_t kc = gc >= N ? gc - N : gc;
                const int lc = L.level_of_j((int)(kc / imax));
                const int c = L.local(gc, N, lc);
                const int vidx = (int)(((comp ? kPsiSlots : 0) + q) * N + k);
                if (lc == lr) {
                    if (pass == 0) ++cD[R + 1]; else eD[fD[R]++] = make_int2(c, vidx);
                } else if (lc == lr - 1) {
                    const int64_t C = P.voff[lc] + c;
                    if (pass == 0) { ++cL[R + 1]; ++cLC[C + 1]; }
                    else { eL[fL[R]++] = make_int2(c, vidx); eLC[fLC[C]++] = make_int2(r, vidx); }
                } else if (lc == lr + 1) {
                    const int64_t C = P.voff[lc] + c;
                    if (pass == 0) { ++cU[R + 1]; ++cUC[C + 1]; }
                    else { eU[fU[R]++] = make_int2(c, vidx); eUC[fUC[C]++] = make_int2(r, vidx); }
                } else {
                    set_error("direct solver: a row couples levels more than one apart");
                    return -1;
                }
            }
        }
    }
    auto up = [s](int **dptr, int2 **dent, const std::vector<int> &ptr,
                 const std::vector<int2> &ent) -> int {
        V2D_HIP(hipMallocAsync(dptr, sizeof(int) * ptr.size(), s));
        V2D_HIP(hipMemcpyAsync(*dptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, s));
        V2D_HIP(hipMallocAsync(dent, sizeof(int2) * std::max<size_t>(ent.size(), 1), s));
        if (!ent.empty())
            V2D_HIP(hipMemcpyAsync(*dent, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice, s));
        return 0;
    };
    P.nent[0] = cD[n2]; P.nent[1] = cL[n2]; P.nent[2] = cU[n2]; P.nent[3] = cUC[n2]; P.nent[4] = cLC[n2];
    if (up(&P.d_Dptr, &P.d_D, cD, eD) || up(&P.d_Lptr, &P.d_Lr, cL, eL) ||
        up(&P.d_Uptr, &P.d_Ur, cU, eU) || up(&P.d_UCptr, &P.d_Uc, cUC, eUC) ||
        up(&P.d_LCptr, &P.d_Lc, cLC, eLC))
        return -1;
    auto upi = [s](int **d, const std::vector<int> &h) -> int {
        V2D_HIP(hipMallocAsync(d, sizeof(int) * std::max<size_t>(h.size(), 1), s));
        if (!h.empty())
            V2D_HIP(hipMemcpyAsync(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s));
        return 0;
    };
    // Padded coupling rows: every level row's couplings into level l-1 (L) and
    // l+1 (U), padded to kCouple entries, stored [kCouple][n_l] at
    // kCouple * voff[l] (level-ordered source index, value index or -1).
    const int64_t tot = (int64_t)kCouple * n2;
    std::vector<int> ixL(tot, 0), ixU(tot, 0), vxL(tot, -1), vxU(tot, -1);
    auto pad = [&](const std::vector<int> &ptr, const std::vector<int2> &ent, int l, int lc,
                   std::vector<int> &ix, std::vector<int> &vx) -> bool {
        const int nl = P.n[l];
        const int64_t base = (int64_t)kCouple * P.voff[l];
        for (int r = 0; r < nl; ++r) {
            const int64_t R = P.voff[l] + r;
            const int e0 = ptr[R], e1 = ptr[R + 1];
            if (e1 - e0 > kCouple) return false;
            if (e1 > e0 && (lc < 0 || lc >= nlev)) return false;
            for (int t = 0; t < kCouple; ++t) {
                const int64_t q = base + (int64_t)t * nl + r;
                if (e0 + t < e1) {
                    ix[q] = (int)(P.voff[lc] + ent[e0 + t].x);
                    vx[q] = ent[e0 + t].y;
                } else {
                    ix[q] = (lc >= 0 && lc < nlev) ? (int)P.voff[lc] : 0;   // any valid source
                }
            }
        }
        return true;
    };
    for (int l = 0; l < nlev; ++l)
        if (!pad(cL, eL, l, l - 1, ixL, vxL) || !pad(cU, eU, l, l + 1, ixU, vxU)) {
            set_error("direct solver: a row couples to more than " + std::to_string(kCouple) +
                      " unknowns of a neighbouring level");
            return -1;
        }
    // Padded coupling columns: column q of level l in U_{l-1} (side 0) and in
    // L_{l+1} (side 1), padded to the level's widest column (cx_w, 1 to 8
    // entries), stored [cx_w][n_l] at cx_off (row in the neighbouring level,
    // value index or -1).
    std::vector<int> jxC, vxC;
    P.ncx = 0;
    for (int side = 0; side < 2; ++side) {
        P.cx_off[side].assign(nlev, 0);
        P.cx_w[side].assign(nlev, 1);
    }
    for (int side = 0; side < 2; ++side)
        for (int l = 0; l < nlev; ++l) {
            const std::vector<int> &ptr = side ? cLC : cUC;
            int w = 0;
            for (int q = 0; q < P.n[l]; ++q) w = std::max(w, ptr[P.voff[l] + q + 1] - ptr[P.voff[l] + q]);
            if (w > 8) {
                set_error("direct solver: a column couples to more than 8 unknowns of a neighbouring level");
                return -1;
            }
            P.cx_w[side][l] = std::max(w, 1);
            P.cx_off[side][l] = P.ncx;
            P.ncx += (int64_t)P.cx_w[side][l] * P.n[l];
        }
    jxC.assign(P.ncx, 0);
    vxC.assign(P.ncx, -1);
    for (int side = 0; side < 2; ++side)
        for (int l = 0; l < nlev; ++l) {
            const std::vector<int> &ptr = side ? cLC : cUC;
            const std::vector<int2> &ent = side ? eLC : eUC;
            const int nl = P.n[l];
            for (int q = 0; q < nl; ++q) {
                const int e0 = ptr[P.voff[l] + q], e1 = ptr[P.voff[l] + q + 1];
                for (int t = 0; t < e1 - e0; ++t) {
                    jxC[P.cx_off[side][l] + (int64_t)t * nl + q] = ent[e0 + t].x;
                    vxC[P.cx_off[side][l] + (int64_t)t * nl + q] = ent[e0 + t].y;
                }
            }
        }
    if (upi(&P.d_ixL, ixL) || upi(&P.d_ixU, ixU) || upi(&P.d_vxL, vxL) || upi(&P.d_vxU, vxU) ||
        upi(&P.d_jxC, jxC) || upi(&P.d_vxC, vxC))
        return -1;
    P.pad_ok = true;
    // tables of the persistent solve sweeps
    P.sweep_ok = false;
    bool closed = nlev >= 2;
    for (int l = 0; l < nlev && closed; ++l)
        closed = P.ld[l] == lvl_ld(P.n[l]) && P.voff[l] == lvl_voff(L, l) && P.soff[l] == lvl_soff(L, l);
    const int M = (nlev - 1) / 2;
    const int steps = std::max(M, nlev - 1 - M);
    P.nsteps = 2 * steps + 1;
    int dev = 0, ncu = 0;
    V2D_HIP(hipGetDevice(&dev));
    V2D_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    P.ncu = ncu;
    if (closed) {
        // C workgroups per chain, one per compute unit: at most 16 rows of an
        // interior level each, at most 128 rows of any level (8 lanes per row
        // build the right-hand side)
        P.sweep_C = std::max(1, std::min(ncu / 2, (2 * imax + kSwRows - 1) / kSwRows));
        P.sweep_rows = 1;
        for (int l = 0; l < nlev; ++l) P.sweep_rows = std::max(P.sweep_rows, (P.n[l] + P.sweep_C - 1) / P.sweep_C);
        P.sweep_grid = 2 * P.sweep_C;
        P.sweep_ok = P.sweep_rows <= kSwThreads / 8 && ncu >= 2;
    }
    V2D_HIP(hipStreamSynchronize(s));   // the host lists above are locals
    return 0;
}

int direct_copy_pattern(DirectPattern &D, const DirectPattern &S, hipStream_t s)
{
    direct_free_pattern(D, s);
    D = S;
    // the device lists are D's own: allocate and copy every one
    D.d_Dptr = D.d_Lptr = D.d_Uptr = D.d_UCptr = D.d_LCptr = nullptr;
    D.d_ixL = D.d_ixU = D.d_vxL = D.d_vxU = D.d_jxC = D.d_vxC = nullptr;
    D.d_D = D.d_Lr = D.d_Ur = D.d_Uc = D.d_Lc = nullptr;
    D.sweep_ok = D.pad_ok = false;
    auto dup = [s](auto **dst, const auto *src, int64_t count) -> int {
        const size_t bytes = sizeof(*src) * (size_t)std::max<int64_t>(count, 1);
        V2D_HIP(hipMallocAsync((void **)dst, bytes, s));
        if (count > 0)
            V2D_HIP(hipMemcpyAsync(*dst, src, sizeof(*src) * (size_t)count, hipMemcpyDeviceToDevice, s));
        return 0;
    };
    const int64_t n2 = 2 * S.N, tot = (int64_t)kCouple * n2;
    if (dup(&D.d_Dptr, S.d_Dptr, n2 + 1) || dup(&D.d_Lptr, S.d_Lptr, n2 + 1) || dup(&D.d_Uptr, S.d_Uptr, n2 + 1) ||
        dup(&D.d_UCptr, S.d_UCptr, n2 + 1) || dup(&D.d_LCptr, S.d_LCptr, n2 + 1) ||
        dup(&D.d_D, S.d_D, S.nent[0]) || dup(&D.d_Lr, S.d_Lr, S.nent[1]) || dup(&D.d_Ur, S.d_Ur, S.nent[2]) ||
        dup(&D.d_Uc, S.d_Uc, S.nent[3]) || dup(&D.d_Lc, S.d_Lc, S.nent[4]) ||
        dup(&D.d_ixL, S.d_ixL, tot) || dup(&D.d_ixU, S.d_ixU, tot) || dup(&D.d_vxL, S.d_vxL, tot) ||
        dup(&D.d_vxU, S.d_vxU, tot) || dup(&D.d_jxC, S.d_jxC, S.ncx) || dup(&D.d_vxC, S.d_vxC, S.ncx))
        return -1;
    D.sweep_ok = S.sweep_ok;
    D.pad_ok = S.pad_ok;
    return 0;
}

void direct_free_pattern(DirectPattern &P, hipStream_t s)
{
    int *ptrs[] = {P.d_Dptr, P.d_Lptr, P.d_Uptr, P.d_UCptr, P.d_LCptr,
                   P.d_ixL, P.d_ixU, P.d_vxL, P.d_vxU, P.d_jxC, P.d_vxC};
    int2 *ents[] = {P.d_D, P.d_Lr, P.d_Ur, P.d_Uc, P.d_Lc};
    for (int *p : ptrs) if (p) (void)hipFreeAsync(p, s);
    for (int2 *p : ents) if (p) (void)hipFreeAsync(p, s);
    P.d_Dptr = P.d_Lptr = P.d_Uptr = P.d_UCptr = P.d_LCptr = nullptr;
    P.d_ixL = P.d_ixU = P.d_vxL = P.d_vxU = P.d_jxC = P.d_vxC = nullptr;
    P.d_D = P.d_Lr = P.d_Ur = P.d_Uc = P.d_Lc = nullptr;
    P.sweep_ok = P.pad_ok = false;
}


// ------------------------------------------------------------ batching ----
// Both elimination chains advance together: every launch below carries up to
// two level jobs and selects its job with blockIdx.z.
template <typename T>
struct Batch {
    Job<T> j[2];
    int nj;
    int kind;   // Gauss-Jordan update kernel (UpdateKind)
    __device__ __forceinline__ const Job<T> &job(int z) const { return j[z]; }
};

// Many jobs from a device table (the fronts of one elimination-tree level).
// The panels rotate over three buffers per job: step s reads panel s from
// buffer s mod 3 (rot), writes panel s + 1 into the next buffer and, in a
// heavy step, also reads panel s - 1 from the previous one (returned as
// colk_x / rowk_x).
template <typename T>
struct TableBatch {
    const Job<T> *jt;
    int nj;
    int kind;
    int rot;       // buffer of the step's read panels
    int tpj;       // grid tiles per job (the largest job's tile count)
    int compact;   // 1: the grid covers the next pivot's row and column blocks only
    __device__ __forceinline__ Job<T> job(int z) const
    {
        Job<T> J = jt[z];
        T *const c0 = J.colk, *const c1 = J.colk_w, *const c2 = J.colk_x;
        T *const r0 = J.rowk, *const r1 = J.rowk_w, *const r2 = J.rowk_x;
        J.colk = rot == 0 ? c0 : rot == 1 ? c1 : c2;
        J.rowk = rot == 0 ? r0 : rot == 1 ? r1 : r2;
        J.colk_w = rot == 0 ? c1 : rot == 1 ? c2 : c0;
        J.rowk_w = rot == 0 ? r1 : rot == 1 ? r2 : r0;
        J.colk_x = rot == 0 ? c2 : rot == 1 ? c0 : c1;
        J.rowk_x = rot == 0 ? r2 : rot == 1 ? r0 : r1;
        return J;
    }
    // block -> (job, tile): every job's first tile (its pivot tile) first,
    // then each job's remaining tiles contiguous, so that the workgroups in
    // flight together share their job's panels in the caches
    __device__ __forceinline__ void tile(int b, int &z, int &lin) const
    {
        if (b < nj) {
            z = b;
            lin = 0;
        } else {
            const int q = b - nj;
            z = q / (tpj - 1);
            lin = 1 + q % (tpj - 1);
        }
    }
};

// ------------------------------------------------------- level build ------
// One Schur source of a level:  Tm = S_p^-1 C  (np x n, row stride ld), where
// column q of the coupling C holds at most kc entries (padded columns: row of
// the neighbouring level p, coefficient).
template <typename T>
struct SchurSrc {
    const T *Sprev;        // S_p^-1 (np x ldp)
    int np, ldp;
    const int *jx;         // [kc][n] rows of level p
    const T *cf;           // [kc][n] coefficients (0 on padding)
    int kc;
    T *tm;
    int n, ld;
};

// A level's dense block before its inversion:
//   S = D - sum_k R_k Tm_k      (n x ld, zero beyond column n)
// where row r of the coupling R_k holds kCouple padded entries (level-ordered
// column, coefficient) and D is the level's own CSR row.
template <typename T>
struct LevelBuild {
    T *S;
    int n, ld;
    const int *dptr;       // CSR rows of D (this level)
    int nsrc;
    const T *tm[2];
    const int *ix[2];      // [kCouple][n] level-ordered columns
    const T *cf[2];        // [kCouple][n] coefficients
    int ixoff[2];          // level-ordered offset of the source level
};

template <typename T>
struct BuildBatch {
    SchurSrc<T> src[2];
    LevelBuild<T> lv[2];
};

constexpr int kTmRows = 4;   // rows of Tm per wavefront (Tm kernel)
constexpr int kSbRows = 4;   // rows of S per wavefront (S kernel)

// Tm[c][q] = sum_t S_p^-1[c][jx_t(q)] cf_t(q): a workgroup's four wavefronts
// each take 64 columns of kTmRows rows; the padded column lists are read once
// per lane and every row load of S_p^-1 is issued before the first product.
template <typename T, int KC>
__global__ void __launch_bounds__(256) k_schur_tm(BuildBatch<T> B)
{
    const SchurSrc<T> &A = B.src[blockIdx.z];
    const int lane = threadIdx.x & 63;
    const int c0 = (blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * kTmRows;
    const int q = blockIdx.x * 64 + lane;
    if (c0 >= A.np || q >= A.n) return;
    int p[KC];
    double u[KC];
#pragma unroll
    for (int t = 0; t < KC; ++t) {   // clamped loads, no branches: entries past kc read as 0
        const int64_t o = (int64_t)min(t, A.kc - 1) * A.n + q;
        p[t] = A.jx[o];
        const T c = A.cf[o];
        u[t] = t < A.kc ? (double)c : 0.0;
    }
    T v[kTmRows][KC];
#pragma unroll
    for (int r = 0; r < kTmRows; ++r) {
        const T *row = A.Sprev + (int64_t)min(c0 + r, A.np - 1) * A.ldp;
#pragma unroll
        for (int t = 0; t < KC; ++t) v[r][t] = row[p[t]];
    }
#pragma unroll
    for (int r = 0; r < kTmRows; ++r) {
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < KC; ++t) acc += (double)v[r][t] * u[t];
        if (c0 + r < A.np) A.tm[(int64_t)(c0 + r) * A.ld + q] = (T)acc;
    }
}

__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rdlane(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdlane(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// S[r][q] = D[r][q] - sum_k sum_t cf_k,t(r) Tm_k[ix_k,t(r)][q] for q < n and
// 0 for n <= q < ld: every entry of the block written once.  Rows are
// wavefront-uniform: one vector load brings the kSbRows x kCouple row-list
// entries of a source (lane 4 t + rr), another the D entries of the rows (at
// most 64 in all, else one row at a time), and readlane hands them out.  All
// Tm loads of the kSbRows rows are issued before the first sum.
template <typename T, int NS>
__device__ __forceinline__ void build_rows(const LevelBuild<T> &J, int r0, int q, const int2 *__restrict__ dent,
                                           const T *__restrict__ val)
{
    static_assert(kSbRows * kCouple <= 64, "row lists of a wavefront exceed its lanes");
    const int lane = threadIdx.x & 63;
    const bool in = q < J.n;
    const int qc = min(q, J.n - 1);
    int ixl[NS > 0 ? NS : 1];
    T cfl[NS > 0 ? NS : 1];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int t = (lane >> 2) & (kCouple - 1), rr = lane & 3;
        const int64_t o = (int64_t)t * J.n + min(r0 + rr, J.n - 1);
        ixl[k] = J.ix[k][o] - J.ixoff[k];
        cfl[k] = J.cf[k][o];
    }
    double tv[kSbRows][NS > 0 ? NS : 1][kCouple];
#pragma unroll
    for (int rr = 0; rr < kSbRows; ++rr)
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
            for (int t = 0; t < kCouple; ++t)
                tv[rr][k][t] = (double)J.tm[k][(int64_t)rdlane(ixl[k], 4 * t + rr) * J.ld + qc];
    int rb[kSbRows + 1];
#pragma unroll
    for (int rr = 0; rr <= kSbRows; ++rr) rb[rr] = J.dptr[min(r0 + rr, J.n)];
    double acc[kSbRows];
#pragma unroll
    for (int rr = 0; rr < kSbRows; ++rr) acc[rr] = 0.0;
    if (rb[kSbRows] - rb[0] <= 64) {
        int dc = -1;
        T dv = T(0);
        if (rb[0] + lane < rb[kSbRows]) {
            const int2 x = dent[rb[0] + lane];
            dc = x.x;
            dv = val[x.y];
        }
#pragma unroll
        for (int rr = 0; rr < kSbRows; ++rr)
            for (int i = rb[rr] - rb[0]; i < rb[rr + 1] - rb[0]; ++i)
                if (rdlane(dc, i) == q) acc[rr] = (double)rdlane(dv, i);
    } else {
#pragma unroll
        for (int rr = 0; rr < kSbRows; ++rr)
            for (int e = rb[rr]; e < rb[rr + 1]; ++e) {
                const int2 x = dent[e];
                if (x.x == q) acc[rr] = (double)val[x.y];
            }
    }
#pragma unroll
    for (int rr = 0; rr < kSbRows; ++rr) {
        if (r0 + rr >= J.n) break;
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
            for (int t = 0; t < kCouple; ++t) acc[rr] -= (double)rdlane(cfl[k], 4 * t + rr) * tv[rr][k][t];
        J.S[(int64_t)(r0 + rr) * J.ld + q] = in ? (T)acc[rr] : T(0);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_build_level(BuildBatch<T> B, const int2 *__restrict__ dent,
                                                     const T *__restrict__ val)
{
    const LevelBuild<T> J = B.lv[blockIdx.z];
    const int r0 = (blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * kSbRows;
    const int q = blockIdx.x * 64 + (threadIdx.x & 63);
    if (r0 >= J.n || blockIdx.x * 64 >= J.ld) return;
    if (J.nsrc == 2) build_rows<T, 2>(J, r0, q, dent, val);
    else if (J.nsrc == 1) build_rows<T, 1>(J, r0, q, dent, val);
    else build_rows<T, 0>(J, r0, q, dent, val);
}

// --------------------------------------------------- Gauss-Jordan ---------
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);   // ~2^-26, two Newton steps to full precision
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// Value of lane N of the lane's own 16-lane row (DPP row_newbcast:N, no LDS
// round trip), and of lane L of the wavefront (readlane, uniform).
template <int N>
__device__ __forceinline__ float row_bcast(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + N, 0xf, 0xf, false));
}
template <int N>
__device__ __forceinline__ double row_bcast(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x150 + N, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x150 + N, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
// Pivot P of the in-register 16 x 16 inverse: lane 16 g + c holds column c
// of rows 4 g + q (RG) or 4 q + g, q < 4.  The pivot row crosses 16-lane rows
// (a shuffle), the pivot value is one lane's (readlane), the pivot-column
// entries of the lane's rows sit in lane P of its own 16-lane row (DPP).
template <bool RG, int P, typename T>
__device__ __forceinline__ void gj16_pivot(T (&d)[4], T &sc, int g, int c)
{
    constexpr int pg = RG ? P >> 2 : P & 3, pq = RG ? P & 3 : P >> 2;
    const T prow = __shfl(d[pq], pg * 16 + c, 64);
    const T piv = lane_value(d[pq], pg * 16 + P);
    T pc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) pc[q] = row_bcast<P>(d[q]);
    const T ip = recip(piv);
    const bool isp = c == P;
    const T rp = isp ? T(0) : prow * ip;
    sc = isp ? -ip : sc;
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = fma(-pc[q], rp, d[q]);
    if (g == pg) d[pq] = isp ? T(-1) : rp;
}
template <bool RG, int P, typename T>
__device__ __forceinline__ void gj16_pivots(T (&d)[4], T &sc, int g, int c)
{
    if constexpr (P < 16) {
        gj16_pivot<RG, P>(d, sc, g, c);
        gj16_pivots<RG, P + 1>(d, sc, g, c);
    }
}
// In-register inverse of a 16 x 16 block held by one wavefront (layout as
// above): Gauss-Jordan without pivoting; the pivot column is scaled by
// -1 / a[p][p] once at the end, its diagonal entry set to -1 on the way so
// that it ends as 1 / a[p][p].
template <bool RG, typename T>
__device__ __forceinline__ void gj16_invert(T (&d)[4], int g, int c)
{
    T sc = T(0);
    gj16_pivots<RG, 0>(d, sc, g, c);
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] *= sc;
}

template <typename T> struct Mfma;

// Inverse of the leading kb x kb block (kb <= 64) of a 64 x 64 tile held in
// LDS (sa, row stride lda), in place, by blocked Gauss-Jordan without
// pivoting over 16-wide steps K = [16 k, 16 k + 16), the diagonal block D of
// step k already inverted in place (wavefront 0 inverts the first one in
// registers, gj16_invert):
//   1. on the matrix cores (16x16x4 blocks, operands straight from LDS):
//      wavefront w < 3 takes the row block I = ob(w) outside K: C = A[I, K] D^-1
//      into A[I, K] (negated; its own rows, so no barrier before it reads them
//      back), then its three trailing blocks A[I, J] + C A[K, J], J outside K,
//      with the pivot rows still old; the wavefront whose rows hold the next
//      diagonal block inverts that block in registers right away (the matrix
//      core output layout is the inverse's), overlapping the others' work;
//      wavefront 3 forms the pivot rows R = D^-1 A[K, J];
//   2. the trailing blocks and R are stored.
// Two barriers per step.  Padding rows and columns (kb .. 63) are zero on
// entry; they are set to the identity for the sweep (so the tile is block
// diagonal) and to zero again at the end.  256 threads.  On exit
// a[q] = inverse[16 (t / 64) + q][t % 64] and out (row stride 64) holds the
// inverse.
template <typename T>
__device__ __forceinline__ typename Mfma<T>::C mm16(const T *sa, int lda, int xr, int xc, int yr, int yc,
                                                    typename Mfma<T>::C acc)
{
    // acc += X Y, X = sa[xr.., xc..], Y = sa[yr.., yc..] (16 x 16 each);
    // element v of lane l: row drow(l / 16, v), column l % 16
    const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
#pragma unroll
    for (int h = 0; h < 4; ++h)
        acc = Mfma<T>::opd(sa[(xr + li) * lda + xc + 4 * h + lk], sa[(yr + 4 * h + lk) * lda + yc + li], acc);
    return acc;
}

template <typename T, bool kCoherentOut = false>
__device__ void gj_tile(T *sa, int lda, int kb, T (&a)[16], T *out)
{
    using M = Mfma<T>;
    using C = typename M::C;
    constexpr bool RG = M::kRowsByGroup;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, lk = lane >> 4;
    __syncthreads();
    for (int r = kb + t; r < NB; r += 256) sa[r * lda + r] = T(1);
    __syncthreads();
    const int nk = (kb + 15) / 16;
    if (wave == 0) {
        T d[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) d[q] = sa[M::drow(lk, q) * lda + li];
        gj16_invert<RG>(d, lk, li);
#pragma unroll
        for (int q = 0; q < 4; ++q) sa[M::drow(lk, q) * lda + li] = d[q];
    }
    __syncthreads();
    for (int k = 0; k < nk; ++k) {
        const int K0 = 16 * k;
        auto ob = [k](int o) { return 16 * (o < k ? o : o + 1); };   // the three blocks outside K
        C tt[3];
        if (wave < 3) {
            const int I0 = ob(wave);
            const C ct = mm16<T>(sa, lda, I0, K0, K0, K0, C{});   // A[I, K] D^-1
#pragma unroll
            for (int v = 0; v < 4; ++v) sa[(I0 + M::drow(lk, v)) * lda + K0 + li] = -ct[v];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int J0 = ob(j);
                C acc;
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] = sa[(I0 + M::drow(lk, v)) * lda + J0 + li];
                tt[j] = mm16<T>(sa, lda, I0, K0, K0, J0, acc);
            }
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (j == k && wave == k && k + 1 < nk) {   // block (k + 1, k + 1): the next diagonal
                    T d[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) d[q] = tt[j][q];
                    gj16_invert<RG>(d, lk, li);
#pragma unroll
                    for (int q = 0; q < 4; ++q) tt[j][q] = d[q];
                }
        } else {
#pragma unroll
            for (int j = 0; j < 3; ++j) tt[j] = mm16<T>(sa, lda, K0, K0, K0, ob(j), C{});   // D^-1 A[K, J]
        }
        __syncthreads();
        const int R0 = wave < 3 ? ob(wave) : K0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) sa[(R0 + M::drow(lk, v)) * lda + ob(j) + li] = tt[j][v];
        __syncthreads();
    }
    const int c = lane, rg = wave;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = rg * 16 + q;
        a[q] = (r < kb && c < kb) ? sa[r * lda + c] : T(0);
        if constexpr (kCoherentOut)
            __hip_atomic_store(out + r * NB + c, a[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            out[r * NB + c] = a[q];
    }
}

template <typename T, typename BT = Batch<T>>
__global__ void __launch_bounds__(256) k_gj_diag_b(BT B, int k0)
{
    constexpr int LT = NB + 4;   // padded LDS row stride
    __shared__ T sa[NB * LT];
    const Job<T> &J = B.job(blockIdx.z);
    const int np = J.np, ld = J.ld;
    if (k0 >= np) return;
    const int kb = min(NB, np - k0);
    const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
    T a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = rg * 16 + q;
        sa[r * LT + c] = (r < kb && c < kb) ? J.S[(int64_t)(k0 + r) * ld + k0 + c] : T(0);
    }
    gj_tile(sa, LT, kb, a, J.dinv);
}

// 16-byte vectors of T
template <typename T> struct Vec8;
template <> struct Vec8<float> { using type = float4; static constexpr int w = 4; };
template <> struct Vec8<double> { using type = double2; static constexpr int w = 2; };

// blockIdx.y == 0: pivot rows  A[K, c] <- Dinv A[K, c]  (A[K, K] <- Dinv) for a
//                  16-column tile (short per-block work: this is on the
//                  critical path of every Gauss-Jordan step)
// blockIdx.y == 1: colk[r][t] <- A[r][k0+t] for a 16-row tile of rows outside K
constexpr int RP = 16;
template <typename T, typename BT = Batch<T>>
__global__ void __launch_bounds__(256) k_gj_rowpanel_b(BT B, int k0)
{
    __shared__ T sdT[NB * NB];   // sdT[s][t] = Dinv[t][s]
    __shared__ T sp[NB * RP];    // sp[s][c]  = A[k0+s][base+c]
    const Job<T> &J = B.job(blockIdx.z);
    T *__restrict__ A = J.S;
    const int n = J.n, ld = J.ld;
    if (k0 >= J.np) return;
    const int kb = min(NB, J.np - k0);
    const int base = blockIdx.x * RP;
    if (base >= n) return;
    if (blockIdx.y == 1) {
        for (int idx = threadIdx.x; idx < RP * NB; idx += blockDim.x) {
            const int r = base + idx / NB, t = idx % NB;
            if (r < n && t < kb && (r < k0 || r >= k0 + kb))
                J.colk[(int64_t)r * NB + t] = A[(int64_t)r * ld + k0 + t];
        }
        return;
    }
    const bool cin = (base >= k0 && base < k0 + NB);
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        const int t = idx / NB, s2 = idx % NB;
        sdT[s2 * NB + t] = J.dinv[idx];
    }
    for (int idx = threadIdx.x; idx < NB * RP; idx += blockDim.x) {
        const int t = idx / RP, c = idx % RP;
        sp[idx] = (!cin && t < kb && base + c < n) ? A[(int64_t)(k0 + t) * ld + base + c] : T(0);
    }
    __syncthreads();
    const int c = threadIdx.x & (RP - 1), tr = threadIdx.x / RP;   // 16 row groups of 4
    const int col = base + c;
    T acc[4];
    if (cin) {
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = sdT[(col - k0) * NB + tr * 4 + a];
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = T(0);
        for (int s2 = 0; s2 < kb; ++s2) {
            const T bv = sp[s2 * RP + c];
#pragma unroll
            for (int a = 0; a < 4; ++a) acc[a] = fma(sdT[s2 * NB + tr * 4 + a], bv, acc[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int t = tr * 4 + a;
        if (t < kb && col < n) A[(int64_t)(k0 + t) * ld + col] = acc[a];
    }
    if (J.rowk && col < n) {   // rowk[c][t] = A[k0+t][c], zero past kb
        typename Vec8<T>::type *dst =
            reinterpret_cast<typename Vec8<T>::type *>(J.rowk + (int64_t)col * NB + tr * 4);
        constexpr int VW = Vec8<T>::w;
#pragma unroll
        for (int v = 0; v < 4 / VW; ++v) {
            typename Vec8<T>::type x;
            T *xs = reinterpret_cast<T *>(&x);
#pragma unroll
            for (int w = 0; w < VW; ++w) xs[w] = (tr * 4 + v * VW + w < kb) ? acc[v * VW + w] : T(0);
            dst[v] = x;
        }
    }
}

// A[i][c] = (c in K ? 0 : A[i][c]) - sum_t colk[i][t] A[k0+t][c]   for i outside K,
// on UT x UT output tiles (UT = 64 or 128), MT x MT accumulators per thread
// (MT = UT / 16), the K = 64 operand tiles staged in LDS.  Blocks form one 1-D
// grid with the jobs interleaved: blocks 0 and 1 are the tiles holding each
// job's next pivot block (k0+64, k0+64); they are dispatched first and also
// invert that block into dinv, so the 64-pivot chain overlaps the bulk update.

template <typename T, int UT>
__global__ void __launch_bounds__(256) k_gj_update_b(Batch<T> B, int k0)
{
    constexpr int MT = UT / 16;
    __shared__ T sc[NB * UT];   // sc[t][r]: colk rows of the tile, transposed
    __shared__ T sr[NB * UT];   // sr[t][c]: pivot-row tile
    using V = typename Vec8<T>::type;
    constexpr int VW = Vec8<T>::w;
    const int z = blockIdx.x % B.nj;
    const Job<T> &J = B.j[z];
    T *__restrict__ A = J.S;
    const int n = J.n, ld = J.ld;
    if (k0 >= n) return;
    const int kb = min(NB, n - k0);
    const int nt = (n + UT - 1) / UT;
    int lin = blockIdx.x / B.nj;
    if (lin >= nt * nt) return;
    const int k1 = k0 + NB;
    if (k1 < n) {
        const int piv = (k1 / UT) * nt + k1 / UT;
        if (lin == 0) lin = piv;
        else if (lin <= piv) lin -= 1;
    }
    const int i0 = (lin / nt) * UT, c0 = (lin % nt) * UT;
    if (UT == NB && i0 == k0) return;   // a 64-row tile inside K has nothing to update
    for (int idx = threadIdx.x; idx < UT * NB; idx += blockDim.x) {
        const int r = idx / NB, t = idx % NB;
        const int i = i0 + r;
        const bool ok = i < n && t < kb && (i < k0 || i >= k0 + kb);
        sc[t * UT + r] = ok ? J.colk[(int64_t)i * NB + t] : T(0);
    }
    for (int idx = threadIdx.x; idx < NB * UT; idx += blockDim.x) {
        const int t = idx / UT, c = idx % UT;
        sr[idx] = (t < kb && c0 + c < n) ? A[(int64_t)(k0 + t) * ld + c0 + c] : T(0);
    }
    __syncthreads();
    const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
    const int cb = c0 + tc * MT;                 // first column of this thread
    const bool cfull = cb + MT <= n;
    T acc[MT][MT];
#pragma unroll
    for (int a = 0; a < MT; ++a) {
        const int i = i0 + tr * MT + a;
        const bool rowok = i < n && (i < k0 || i >= k0 + kb);
        if (rowok && cfull) {
            const V *src = reinterpret_cast<const V *>(A + (int64_t)i * ld + cb);
#pragma unroll
            for (int v = 0; v < MT / VW; ++v) {
                const V x = src[v];
                const T *xs = reinterpret_cast<const T *>(&x);
#pragma unroll
                for (int w = 0; w < VW; ++w) acc[a][v * VW + w] = xs[w];
            }
        } else {
#pragma unroll
            for (int b = 0; b < MT; ++b) acc[a][b] = (rowok && cb + b < n) ? A[(int64_t)i * ld + cb + b] : T(0);
        }
#pragma unroll
        for (int b = 0; b < MT; ++b)
            if (cb + b >= k0 && cb + b < k0 + kb) acc[a][b] = T(0);
    }
#pragma unroll 2
    for (int t = 0; t < kb; ++t) {
        T av[MT], bv[MT];
#pragma unroll
        for (int a = 0; a < MT; ++a) av[a] = sc[t * UT + tr * MT + a];
#pragma unroll
        for (int b = 0; b < MT; ++b) bv[b] = sr[t * UT + tc * MT + b];
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int b = 0; b < MT; ++b) acc[a][b] = fma(-av[a], bv[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < MT; ++a) {
        const int i = i0 + tr * MT + a;
        if (i >= n || (i >= k0 && i < k0 + kb)) continue;
        if (cfull) {
            V *dst = reinterpret_cast<V *>(A + (int64_t)i * ld + cb);
#pragma unroll
            for (int v = 0; v < MT / VW; ++v) {
                V x;
                T *xs = reinterpret_cast<T *>(&x);
#pragma unroll
                for (int w = 0; w < VW; ++w) xs[w] = acc[a][v * VW + w];
                dst[v] = x;
            }
        } else {
#pragma unroll
            for (int b = 0; b < MT; ++b)
                if (cb + b < n) A[(int64_t)i * ld + cb + b] = acc[a][b];
        }
    }
    // fused inversion of the next pivot block
    if (k1 < n && i0 == (k1 / UT) * UT && c0 == (k1 / UT) * UT) {
        const int kb2 = min(NB, n - k1);
        const int off = k1 - i0;   // 0, or 64 inside a 128 tile
        __syncthreads();
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int b = 0; b < MT; ++b) {
                const int r = tr * MT + a - off, c = tc * MT + b - off;
                if (r >= 0 && r < NB && c >= 0 && c < NB)
                    sc[r * NB + c] = (r < kb2 && c < kb2) ? acc[a][b] : T(0);
            }
        T g[16];
        gj_tile(sc, NB, kb2, g, J.dinv);
    }
}

// MFMA form of the same update for 64 x 64 output tiles: 2 x 2 waves of
// 32 x 32, each a 2 x 2 array of 16x16x4 matrix-core blocks.  The product
// colk (rows x 64) * pivot rows (64 x cols) needs no LDS staging: lane
// (li = lane % 16, lk = lane / 16) supplies A[li][t] and B[t][li] for
// t = 16 h + 4 lk + e in instruction (h, e), h, e = 0..3, so the four lane
// groups carry four distinct t per instruction and the 16 instructions cover
// K.  Each lane reads 4-element runs colk[r][16 h + 4 lk ..] and
// rowk[c][16 h + 4 lk ..] (rowk = the pivot rows transposed, written by the
// row-panel kernel): the four lanes of a row read 4 consecutive runs, so one
// load instruction touches 16 contiguous 64-byte row segments.
// FP32 issues the product transposed (pivot-row operand first), which puts
// four consecutive columns of one row in each lane's accumulator: C moves as
// 16-byte vectors.  FP64 keeps the direct form, where accumulator element v of
// lane l holds row 4 v + l / 16, column l % 16 of its block.  The product
// starts from zero and the old tile is added at the end, so its load overlaps
// the matrix-core chain.  Rows of K form exactly the tile i0 == k0 (64-aligned),
// which is skipped, and columns of K exactly the tile c0 == k0, whose old
// values are taken as zero; operand entries with t >= kb are zeroed in
// registers; rows past n read a valid row and are not stored; columns past n
// only feed unstored outputs.
template <> struct Mfma<float> {
    using C = float __attribute__((ext_vector_type(4)));
    static constexpr bool kSwap = true;
    static __device__ __forceinline__ C op(float a, float b, C c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, c, 0, 0, 0);
    }
    // (row, column) of accumulator element v inside a 16 x 16 block
    static __device__ __forceinline__ int row(int li, int lk, int v) { return li; }
    static __device__ __forceinline__ int col(int li, int lk, int v) { return 4 * lk + v; }
    // direct form: D += A B with element v of lane l at row drow(l / 16, v), column l % 16
    static __device__ __forceinline__ C opd(float a, float b, C c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lk, int v) { return 4 * lk + v; }
    static constexpr bool kRowsByGroup = true;   // drow = 4 lk + v
};
template <> struct Mfma<double> {
    using C = double __attribute__((ext_vector_type(4)));
    static constexpr bool kSwap = false;
    static __device__ __forceinline__ C op(double a, double b, C c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int li, int lk, int v) { return 4 * v + lk; }
    static __device__ __forceinline__ int col(int li, int lk, int v) { return li; }
    static __device__ __forceinline__ C opd(double a, double b, C c) { return op(a, b, c); }
    static __device__ __forceinline__ int drow(int lk, int v) { return 4 * v + lk; }
    static constexpr bool kRowsByGroup = false;   // drow = 4 v + lk
};

// One Gauss-Jordan step of the MFMA form as a single launch (dataflow): every
// block updates its 64 x 64 tile with the step-k0 panels (read buffers), and
// the blocks that touch the next pivot block k1 = k0 + 64 also prepare step k1
// (write buffers), so no row-panel launch sits between two steps:
//   * tile (k1, k1), dispatched first: inverts its updated block (gj_tile),
//     stores the inverse into dinv, publishes flags[job] = epoch (release,
//     agent scope), then stores it as the new pivot block of A and into the
//     next transposed row panel;
//   * tiles (k1, c), c != k1: stage their updated tile in LDS, wait for the
//     flag (acquire; a bounded spin that sets *err and carries on if the flag
//     never comes, so every wave exits), and store D^-1 * tile (matrix cores,
//     direct form) as the new pivot rows of A and the next row panel;
//   * tiles (i, k1), i != k1: also store their updated tile as the next column
//     panel.
// The pivot-row work thus overlaps the bulk update of the step.
constexpr int kSpinLimit = 1 << 21;

template <typename T, typename BT = Batch<T>, int MODE = kStepFull>
__global__ void __launch_bounds__(256, sizeof(T) == 8 ? (MODE == kStepHeavy ? 2 : 1) : 4) k_gj_step_mfma_b(BT B, int k0, int epoch,
                                                        int *__restrict__ flags, int *__restrict__ err)
{
    using M = Mfma<T>;
    using C = typename M::C;
    using V = typename Vec8<T>::type;
    using NV = typename NVec<T>::type;
    constexpr int VW = Vec8<T>::w;
    constexpr int LT = NB + 4;    // padded LDS row stride
    // next pivot block, or a tile stored transposed; before that (FP32 job
    // tables) the panel blocks of the tile's product staged in LDS: a heavy
    // step's two column-panel blocks, a light or full step's column- and
    // row-panel blocks
    constexpr bool kTable = !std::is_same<BT, Batch<T>>::value;
    constexpr bool kStageP = kTable && sizeof(T) == 4 && MODE == kStepHeavy;
    constexpr bool kStage1 = kTable && sizeof(T) == 4 && (MODE == kStepLight || MODE == kStepFull);
    constexpr bool kStageD = kTable && sizeof(T) == 8 && MODE == kStepHeavy;
    __shared__ __align__(16) T sT[(kStageP || kStage1 || kStageD) ? 2 * NB * NB : NB * LT];
    static_assert(kTable || MODE == kStepFull, "the level batches take full steps only");
    int z, lin;
    if constexpr (kTable) {
        B.tile(blockIdx.x, z, lin);
    } else {
        z = blockIdx.x % B.nj;
        lin = blockIdx.x / B.nj;
    }
    const Job<T> &J = B.job(z);
    T *__restrict__ A = J.S;
    const int n = J.n, ld = J.ld, np = J.np;
    if (k0 >= np) return;
    const int kb = min(NB, np - k0);
    const int nt = (n + NB - 1) / NB;
    const int k1 = k0 + NB;
    int i0, c0;
    bool compact = false;
    if constexpr (kTable) compact = B.compact && k1 < np;
    if (compact) {
        // only the next pivot's column block (pivot tile first) and row block
        const int kt = k1 / NB;
        if (lin >= 2 * nt - 1) return;
        if (lin < nt) {
            i0 = (lin == 0 ? kt : (lin - 1 < kt ? lin - 1 : lin)) * NB;
            c0 = k1;
        } else {
            const int q = lin - nt;
            i0 = k1;
            c0 = (q < kt ? q : q + 1) * NB;
        }
    } else {
        if (lin >= nt * nt) return;
        if (k1 < np) {
            const int piv = (k1 / NB) * nt + k1 / NB;
            if (lin == 0) lin = piv;
            else if (lin <= piv) lin -= 1;
        }
        i0 = (lin / nt) * NB;
        c0 = (lin % nt) * NB;
        // a light or first step leaves the tiles outside the next pivot's row
        // and column blocks to the following heavy step
        if ((MODE == kStepLight || MODE == kStepBegin) && k1 < np && i0 != k1 && c0 != k1) return;
    }
    if (i0 == k0) {
        // the current pivot rows are not updated; their part of column tile
        // k1 is still the next column panel's rows k0 .. k0 + 63
        if (k1 < np && c0 == k1) {
            for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
                const int r = idx / NB, t = idx % NB;
                glb(J.colk_w)[(int64_t)(k0 + r) * NB + t] = (k1 + t < np) ? glb(A)[(int64_t)(k0 + r) * ld + k1 + t] : T(0);
            }
        }
        return;
    }
    const bool is_piv = k1 < np && i0 == k1 && c0 == k1;
    const bool is_prow = k1 < np && i0 == k1 && c0 != k1;
    const bool is_pcol = k1 < np && c0 == k1 && i0 != k1;
    // the pivot chain (next pivot inverse, then the next pivot rows) bounds the
    // step: its waves take issue priority over the bulk update sharing the SIMDs
    if (is_piv || is_prow) __builtin_amdgcn_s_setprio(3);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
    // The wavefront's share of the 64 x 64 tile, NBI x NBJ blocks of 16 x 16:
    // in a heavy FP32 step a 64 x 16 column strip (rwl = 0, cwl = 16 w), so no
    // row-panel operand is loaded twice in a workgroup; otherwise a 32 x 32
    // quarter.
    constexpr bool kStrip = kStageP || kStageD;
    constexpr int NBI = kStrip ? 4 : 2, NBJ = kStrip ? 1 : 2;
    const int rwl = kStrip ? 0 : (w >> 1) * 32, cwl = kStrip ? w * 16 : (w & 1) * 32;
    const int rw = i0 + rwl, cw = c0 + cwl;
    const bool colK = c0 == k0;            // the pivot columns: old values are zero
    const bool cfull = c0 + NB <= n;       // no column of the tile past n
    // A heavy step also applies the previous step's panels (colk_x, rowk_x),
    // whose update the light step deferred: every tile outside the previous
    // pivot rows (kp) and this step's pivot columns, with old values zero in
    // the previous pivot columns.
    bool oldZero = colK, useP = false;
    if constexpr (MODE == kStepHeavy) {
        const int kp = k0 - NB;
        const bool rowP = i0 == kp, colP = c0 == kp;
        useP = !rowP && !colK;
        oldZero = colK || (colP && !rowP);
    }
    // ---------------------------------------------------- tile update ----
    C acc[NBI][NBJ], old[NBI][NBJ];
#pragma unroll
    for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[bi][bj][v] = T(0);
    if constexpr (kTable && sizeof(T) == 4) {
        // Job tables (many fronts per launch, the bulk update dominates): every
        // load of a panel's product goes out before its first matrix-core
        // product -- the four 16-wide k-slices -- with the old tile (clamped
        // rows and columns, zero selected outside the matrix) behind the
        // first panel, so a tile waits for one memory round trip per panel.
        // FP32 only: in FP64 the registers this takes halve the occupancy and
        // the launch is slower.
        const bool tail = kb < NB;
        auto product = [&](const T *ck, const T *rk, bool msk, bool with_old) {
            gptr<const NV> qa[NBI], qb[NBJ];
#pragma unroll
            for (int bi = 0; bi < NBI; ++bi)
                qa[bi] = glb(reinterpret_cast<const NV *>(ck + (int64_t)min(rw + bi * 16 + li, n - 1) * NB + 4 * lk));
#pragma unroll
            for (int bj = 0; bj < NBJ; ++bj)
                qb[bj] = glb(reinterpret_cast<const NV *>(rk + (int64_t)(cw + bj * 16 + li) * NB + 4 * lk));
            NV xa[4][NBI][4 / VW], xb[4][NBJ][4 / VW];
#pragma unroll
            for (int hh = 0; hh < 4; ++hh)
#pragma unroll
                for (int u = 0; u < 4 / VW; ++u) {
#pragma unroll
                    for (int bi = 0; bi < NBI; ++bi) xa[hh][bi][u] = qa[bi][(16 * hh) / VW + u];
#pragma unroll
                    for (int bj = 0; bj < NBJ; ++bj) xb[hh][bj][u] = qb[bj][(16 * hh) / VW + u];
                }
            if (with_old) {
#pragma unroll
                for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
                    for (int bj = 0; bj < NBJ; ++bj) {
                        const int r = rw + bi * 16 + M::row(li, lk, 0), c = cw + bj * 16 + M::col(li, lk, 0);
                        const f32x4 x = *glb(reinterpret_cast<const f32x4 *>(
                            A + (int64_t)min(r, n - 1) * ld + min(c, ld - 4)));
                        const bool in = !oldZero && r < n && c < n;
                        old[bi][bj][0] = in ? x[0] : T(0);
                        old[bi][bj][1] = in ? x[1] : T(0);
                        old[bi][bj][2] = in ? x[2] : T(0);
                        old[bi][bj][3] = in ? x[3] : T(0);
                    }
            }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                T av[NBI][4], bv[NBJ][4];
#pragma unroll
                for (int u = 0; u < 4 / VW; ++u)
#pragma unroll
                    for (int e = 0; e < VW; ++e) {
#pragma unroll
                        for (int bi = 0; bi < NBI; ++bi)
                            av[bi][u * VW + e] = -reinterpret_cast<const T *>(&xa[h][bi][u])[e];
#pragma unroll
                        for (int bj = 0; bj < NBJ; ++bj)
                            bv[bj][u * VW + e] = reinterpret_cast<const T *>(&xb[h][bj][u])[e];
                    }
                if (msk) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (16 * h + 4 * lk + e >= kb)
#pragma unroll
                            for (int bi = 0; bi < NBI; ++bi) av[bi][e] = T(0);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
                        for (int bj = 0; bj < NBJ; ++bj) acc[bi][bj] = M::op(av[bi][e], bv[bj][e], acc[bi][bj]);
            }
        };
        if constexpr (MODE == kStepBegin) {
#pragma unroll
            for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
                for (int bj = 0; bj < NBJ; ++bj) {
                    const int r = rw + bi * 16 + M::row(li, lk, 0), c = cw + bj * 16 + M::col(li, lk, 0);
                    const f32x4 x =
                        *glb(reinterpret_cast<const f32x4 *>(A + (int64_t)min(r, n - 1) * ld + min(c, ld - 4)));
                    const bool in = r < n && c < n;
                    old[bi][bj][0] = in ? x[0] : T(0);
                    old[bi][bj][1] = in ? x[1] : T(0);
                    old[bi][bj][2] = in ? x[2] : T(0);
                    old[bi][bj][3] = in ? x[3] : T(0);
                }
        } else if constexpr (kStageP) {
            {
                // The two column-panel blocks (previous step's, then this
                // step's: rows of the tile, 64 pivots) go to LDS by direct
                // 16-byte loads, one round, each read once per workgroup: chunk
                // kc (4 pivots) of block row r sits at chunk kc ^ (r & 15) of
                // the row, so the sixteen rows a matrix-core operand read spans
                // fall in distinct banks; a wave's load fills 64 consecutive
                // chunks, i.e. 4 block rows.  The row-panel operands go to
                // registers in the same round (each wave the 16 columns of its
                // strip).
                const T *src[2] = {J.colk_x, J.colk};
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int q0 = (w * 4 + j) * 64;
                        const int r = (q0 + l) >> 4, kc = (l & 15) ^ (r & 15);
                        // rows past n read row n - 1 (their results are dropped)
                        const int gr = min(i0 + r, n - 1);
                        __builtin_amdgcn_global_load_lds(
                            (const void __attribute__((address_space(1))) *)(src[b] + (int64_t)gr * NB + 4 * kc),
                            (void __attribute__((address_space(3))) *)(sT + (b * 1024 + q0) * 4), 16, 0, 0);
                    }
                f32x4 xb[2][4][NBJ];   // [product][h][bj]
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const T *rk = p == 0 ? J.rowk_x : J.rowk;
#pragma unroll
                    for (int bj = 0; bj < NBJ; ++bj) {
                        const gptr<const f32x4> q =
                            glb(reinterpret_cast<const f32x4 *>(rk + (int64_t)(cw + bj * 16 + li) * NB + 4 * lk));
#pragma unroll
                        for (int h = 0; h < 4; ++h) xb[p][h][bj] = q[4 * h];
                    }
                }
#pragma unroll
                for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
                    for (int bj = 0; bj < NBJ; ++bj) {
                        const int r = rw + bi * 16 + M::row(li, lk, 0), c = cw + bj * 16 + M::col(li, lk, 0);
                        const f32x4 x = *glb(reinterpret_cast<const f32x4 *>(
                            A + (int64_t)min(r, n - 1) * ld + min(c, ld - 4)));
                        const bool in = !oldZero && r < n && c < n;
                        old[bi][bj][0] = in ? x[0] : T(0);
                        old[bi][bj][1] = in ? x[1] : T(0);
                        old[bi][bj][2] = in ? x[2] : T(0);
                        old[bi][bj][3] = in ? x[3] : T(0);
                    }
                __syncthreads();   // the panel blocks have landed
                auto lds_product = [&](int p, bool msk) {
                    const f32x4 *sa = reinterpret_cast<const f32x4 *>(sT) + p * 1024;
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        T av[NBI][4], bv[NBJ][4];
#pragma unroll
                        for (int bi = 0; bi < NBI; ++bi) {
                            const int r = rwl + bi * 16 + li;
                            const f32x4 x = sa[r * 16 + ((4 * h + lk) ^ (r & 15))];
#pragma unroll
                            for (int e = 0; e < 4; ++e) av[bi][e] = -x[e];
                        }
#pragma unroll
                        for (int bj = 0; bj < NBJ; ++bj)
#pragma unroll
                            for (int e = 0; e < 4; ++e) bv[bj][e] = xb[p][h][bj][e];
                        if (msk) {
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                if (16 * h + 4 * lk + e >= kb)
                                    for (int bi = 0; bi < NBI; ++bi) av[bi][e] = T(0);
                        }
#pragma unroll
                        for (int e = 0; e < 4; ++e)
#pragma unroll
                            for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
                                for (int bj = 0; bj < NBJ; ++bj)
                                    acc[bi][bj] = M::op(av[bi][e], bv[bj][e], acc[bi][bj]);
                    }
                };
                // the previous step's panels only where the light step deferred
                // them (the loads above are harmless elsewhere: every job holds
                // its three panel buffers)
                if (useP) lds_product(0, false);
                lds_product(1, tail);
            }
        } else if constexpr (kStage1) {
            const T *src[2] = {J.colk, J.rowk};
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int q0 = (w * 4 + j) * 64;
                    const int r = (q0 + l) >> 4, kc = (l & 15) ^ (r & 15);
                    const int gr = (b & 1) ? c0 + r : min(i0 + r, n - 1);
                    __builtin_amdgcn_global_load_lds(
                        (const void __attribute__((address_space(1))) *)(src[b] + (int64_t)gr * NB + 4 * kc),
                        (void __attribute__((address_space(3))) *)(sT + (b * 1024 + q0) * 4), 16, 0, 0);
                }
#pragma unroll
            for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
                for (int bj = 0; bj < NBJ; ++bj) {
                    const int r = rw + bi * 16 + M::row(li, lk, 0), c = cw + bj * 16 + M::col(li, lk, 0);
                    const f32x4 x = *glb(reinterpret_cast<const f32x4 *>(
                        A + (int64_t)min(r, n - 1) * ld + min(c, ld - 4)));
                    const bool in = !oldZero && r < n && c < n;
                    old[bi][bj][0] = in ? x[0] : T(0);
                    old[bi][bj][1] = in ? x[1] : T(0);
                    old[bi][bj][2] = in ? x[2] : T(0);
                    old[bi][bj][3] = in ? x[3] : T(0);
                }
            __syncthreads();   // the panel blocks have landed
            const f32x4 *sa = reinterpret_cast<const f32x4 *>(sT);
            const f32x4 *sb = reinterpret_cast<const f32x4 *>(sT) + 1024;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                T av[NBI][4], bv[NBJ][4];
#pragma unroll
                for (int bi = 0; bi < NBI; ++bi) {
                    const int r = rwl + bi * 16 + li;
                    const f32x4 x = sa[r * 16 + ((4 * h + lk) ^ (r & 15))];
#pragma unroll
                    for (int e = 0; e < 4; ++e) av[bi][e] = -x[e];
                }
#pragma unroll
                for (int bj = 0; bj < NBJ; ++bj) {
                    const int r = cwl + bj * 16 + li;
                    const f32x4 x = sb[r * 16 + ((4 * h + lk) ^ (r & 15))];
#pragma unroll
                    for (int e = 0; e < 4; ++e) bv[bj][e] = x[e];
                }
                if (tail) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (16 * h + 4 * lk + e >= kb)
                            for (int bi = 0; bi < NBI; ++bi) av[bi][e] = T(0);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
                        for (int bj = 0; bj < NBJ; ++bj) acc[bi][bj] = M::op(av[bi][e], bv[bj][e], acc[bi][bj]);
            }
        } else if (useP) {
            product(J.colk_x, J.rowk_x, false, true);
            product(J.colk, J.rowk, tail, false);
        } else {
            product(J.colk, J.rowk, tail, true);
        }
#pragma unroll
        for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
            for (int bj = 0; bj < NBJ; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[bi][bj][v] += old[bi][bj][v];
    } else {
    #pragma unroll
        for (int bi = 0; bi < NBI; ++bi)
    #pragma unroll
            for (int bj = 0; bj < NBJ; ++bj) {
                if (M::kSwap && cfull && !oldZero) {
                    const int r = rw + bi * 16 + M::row(li, lk, 0), c = cw + bj * 16 + M::col(li, lk, 0);
                    if (r < n) {
                        const f32x4 x = *glb(reinterpret_cast<const f32x4 *>(A + (int64_t)r * ld + c));
                        old[bi][bj][0] = x[0]; old[bi][bj][1] = x[1]; old[bi][bj][2] = x[2]; old[bi][bj][3] = x[3];
                    } else {
    #pragma unroll
                        for (int v = 0; v < 4; ++v) old[bi][bj][v] = T(0);
                    }
                } else {
    #pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = rw + bi * 16 + M::row(li, lk, v), c = cw + bj * 16 + M::col(li, lk, v);
                        old[bi][bj][v] = (!oldZero && r < n && c < n) ? glb(A)[(int64_t)r * ld + c] : T(0);
                    }
                }
            }
        auto product = [&](const T *ck, const T *rk, bool msk) {
            gptr<const T> pa[NBI], pb[NBJ];
    #pragma unroll
            for (int bi = 0; bi < NBI; ++bi) pa[bi] = glb(ck + (int64_t)min(rw + bi * 16 + li, n - 1) * NB + 4 * lk);
    #pragma unroll
            for (int bj = 0; bj < NBJ; ++bj) pb[bj] = glb(rk + (int64_t)(cw + bj * 16 + li) * NB + 4 * lk);
    #pragma unroll
            for (int h = 0; h < 4; ++h) {
                T av[NBI][4], bv[NBJ][4];
    #pragma unroll
                for (int bi = 0; bi < NBI; ++bi)
    #pragma unroll
                    for (int u = 0; u < 4 / VW; ++u) {
                        const NV x = reinterpret_cast<gptr<const NV>>(pa[bi] + 16 * h)[u];
    #pragma unroll
                        for (int e = 0; e < VW; ++e) av[bi][u * VW + e] = -reinterpret_cast<const T *>(&x)[e];
                    }
    #pragma unroll
                for (int bj = 0; bj < NBJ; ++bj)
    #pragma unroll
                    for (int u = 0; u < 4 / VW; ++u) {
                        const NV x = reinterpret_cast<gptr<const NV>>(pb[bj] + 16 * h)[u];
    #pragma unroll
                        for (int e = 0; e < VW; ++e) bv[bj][u * VW + e] = reinterpret_cast<const T *>(&x)[e];
                    }
                if (msk) {
    #pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (16 * h + 4 * lk + e >= kb)
                            for (int bi = 0; bi < NBI; ++bi) av[bi][e] = T(0);
                }
    #pragma unroll
                for (int e = 0; e < 4; ++e)
    #pragma unroll
                    for (int bi = 0; bi < NBI; ++bi)
    #pragma unroll
                        for (int bj = 0; bj < NBJ; ++bj) acc[bi][bj] = M::op(av[bi][e], bv[bj][e], acc[bi][bj]);
            }
        };
        if constexpr (kStageD) {
            // FP64 heavy step: the two column-panel blocks in LDS by direct
            // 16-byte loads (two pivots a chunk, 32 chunks a block row, chunk
            // kc of row r at kc ^ (r & 15); a wave's load fills 64 chunks, 2
            // block rows), the wavefront's row-panel operands (its 16 columns)
            // in registers, all in one round
            const T *src[2] = {J.colk_x, J.colk};
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int q0 = (w * 8 + j) * 64;
                    const int r = (q0 + l) >> 5, kc = ((q0 + l) & 31) ^ (r & 15);
                    const int gr = min(i0 + r, n - 1);
                    __builtin_amdgcn_global_load_lds(
                        (const void __attribute__((address_space(1))) *)(src[b] + (int64_t)gr * NB + 2 * kc),
                        (void __attribute__((address_space(3))) *)(sT + (b * 2048 + q0) * 2), 16, 0, 0);
                }
            f64x2 xb[2][4][2];   // [product][h][half]
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const gptr<const f64x2> q = glb(reinterpret_cast<const f64x2 *>(
                    (p == 0 ? J.rowk_x : J.rowk) + (int64_t)(cw + li) * NB + 4 * lk));
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    xb[p][h][0] = q[8 * h];
                    xb[p][h][1] = q[8 * h + 1];
                }
            }
            __syncthreads();   // the panel blocks have landed
            auto lds_product = [&](int p, bool msk) {
                const f64x2 *sa = reinterpret_cast<const f64x2 *>(sT) + p * 2048;
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    T av[NBI][4], bv[4];
                    const int kc = 8 * h + 2 * lk;
#pragma unroll
                    for (int bi = 0; bi < NBI; ++bi) {
                        const int r = bi * 16 + li;
                        const f64x2 x0 = sa[r * 32 + (kc ^ (r & 15))], x1 = sa[r * 32 + ((kc + 1) ^ (r & 15))];
                        av[bi][0] = -x0[0]; av[bi][1] = -x0[1]; av[bi][2] = -x1[0]; av[bi][3] = -x1[1];
                    }
                    bv[0] = xb[p][h][0][0]; bv[1] = xb[p][h][0][1]; bv[2] = xb[p][h][1][0]; bv[3] = xb[p][h][1][1];
                    if (msk) {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (16 * h + 4 * lk + e >= kb)
#pragma unroll
                                for (int bi = 0; bi < NBI; ++bi) av[bi][e] = T(0);
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int bi = 0; bi < NBI; ++bi) acc[bi][0] = M::op(av[bi][e], bv[e], acc[bi][0]);
                }
            };
            if (useP) lds_product(0, false);
            lds_product(1, kb < NB);
        } else if constexpr (MODE != kStepBegin) {
            if (useP) product(J.colk_x, J.rowk_x, false);
            product(J.colk, J.rowk, kb < NB);
        }
    #pragma unroll
        for (int bi = 0; bi < NBI; ++bi)
    #pragma unroll
            for (int bj = 0; bj < NBJ; ++bj)
    #pragma unroll
                for (int v = 0; v < 4; ++v) acc[bi][bj][v] += old[bi][bj][v];
    }
    // ------------------------------------------------ ordinary tiles ----
    if (!is_piv && !is_prow) {
#pragma unroll
        for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
            for (int bj = 0; bj < NBJ; ++bj) {
                if (M::kSwap && cfull) {
                    const int r = rw + bi * 16 + M::row(li, lk, 0), c = cw + bj * 16 + M::col(li, lk, 0);
                    if (r < n) {
                        f32x4 x;
                        x[0] = acc[bi][bj][0]; x[1] = acc[bi][bj][1]; x[2] = acc[bi][bj][2]; x[3] = acc[bi][bj][3];
                        if (MODE != kStepBegin) *glb(reinterpret_cast<f32x4 *>(A + (int64_t)r * ld + c)) = x;
                        if (is_pcol) *glb(reinterpret_cast<f32x4 *>(J.colk_w + (int64_t)r * NB + (c - k1))) = x;
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = rw + bi * 16 + M::row(li, lk, v), c = cw + bj * 16 + M::col(li, lk, v);
                        if (r < n) {
                            if (MODE != kStepBegin && c < n) glb(A)[(int64_t)r * ld + c] = acc[bi][bj][v];
                            if (is_pcol) glb(J.colk_w)[(int64_t)r * NB + (c - k1)] = acc[bi][bj][v];
                        }
                    }
                }
            }
        return;
    }
    const int kb2 = min(NB, np - k1);
    if constexpr (kStageP || kStage1 || kStageD) __syncthreads();   // every wave is past its reads of the staged panels
    if (is_piv) {
        // ------------------------------------- next pivot block inverse ----
#pragma unroll
        for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
            for (int bj = 0; bj < NBJ; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = rwl + bi * 16 + M::row(li, lk, v);
                    const int c = cwl + bj * 16 + M::col(li, lk, v);
                    sT[r * LT + c] = (r < kb2 && c < kb2) ? acc[bi][bj][v] : T(0);
                }
        const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
        T g[16];
        // publish the inverse first: the pivot-row workgroups read only dinv;
        // the pivot block of A and the next row panel are read by later
        // launches.  The inverse goes out in agent-coherent stores and the
        // flag follows their completion (s_waitcnt vmcnt(0)): a release fence
        // would write back every dirty line of this die's L2, the bulk tiles'
        // updates included.
        gj_tile<T, true>(sT, LT, kb2, g, J.dinv);
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(flags + z, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        T *rk = J.rowk_w + (int64_t)(k1 + c) * NB + rg * 16;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = rg * 16 + q;
            if (r < kb2 && c < kb2) glb(A)[(int64_t)(k1 + r) * ld + k1 + c] = g[q];
            glb(rk)[q] = g[q];
        }
        return;
    }
    // ------------------------------------------------ next pivot rows ----
#pragma unroll
    for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rwl + bi * 16 + M::row(li, lk, v);
                const int c = cwl + bj * 16 + M::col(li, lk, v);
                sT[c * LT + r] = acc[bi][bj][v];   // transposed: sT[c][t]
            }
    if (threadIdx.x == 0) {
        int spins = 0;
        while (__hip_atomic_load(flags + z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
            if (++spins > kSpinLimit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    // The published inverse is read with agent-coherent 16-byte buffer loads
    // (sc1) in place of an acquire fence, whose L2 invalidate would also drop
    // the lines the bulk tiles on this die are using.
    const __amdgpu_buffer_rsrc_t drs =
        __builtin_amdgcn_make_buffer_rsrc(J.dinv, (short)0, NB * NB * (int)sizeof(T), 0x00020000);
    const int rw2 = (w >> 1) * 32, cw2 = (w & 1) * 32;
    C d[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) d[bi][bj][v] = T(0);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        T av[2][4], bv[2][4];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int u = 0; u < 4 / VW; ++u) {
                const int off = ((rw2 + bi * 16 + li) * NB + 16 * h + 4 * lk + u * VW) * (int)sizeof(T);
                const auto raw = __builtin_amdgcn_raw_buffer_load_b128(drs, off, 0, 16);   // 16: sc1
                V x;
                __builtin_memcpy(&x, &raw, sizeof(x));
#pragma unroll
                for (int e = 0; e < VW; ++e) av[bi][u * VW + e] = reinterpret_cast<const T *>(&x)[e];
            }
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int e = 0; e < 4; ++e) bv[bj][e] = sT[(cw2 + bj * 16 + li) * LT + 16 * h + 4 * lk + e];
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) d[bi][bj] = M::opd(av[bi][e], bv[bj][e], d[bi][bj]);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int c = cw2 + bj * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rw2 + bi * 16 + M::drow(lk, v);
                if (r < kb2 && c0 + c < n) glb(A)[(int64_t)(k1 + r) * ld + c0 + c] = d[bi][bj][v];
                glb(J.rowk_w)[(int64_t)(c0 + c) * NB + r] = d[bi][bj][v];
            }
        }
}

// ------------------------------------------------ one-step big fronts ----
// Multifrontal fronts with a single 64-pivot step (nsp = 64, two children,
// nf <= kFront1MaxNf) assembled and swept by two launches per tree height
// instead of the level assembly, entry scatter and Gauss-Jordan step
// launches, with no wait on another workgroup:
// k_mf_front1_head, one workgroup per front:
//   1. the pivot rows [G W] (rows 0 .. 63, every column) and the pivot
//      columns X (rows 64 .., columns 0 .. 63) are assembled into F as the
//      level assembly does -- identity on the padding pivots ns .. 63, then
//      the children's contribution blocks in list order; only the rows and
//      columns of the ns true pivots are gathered -- and the Jacobian entries
//      the front owns are added (the analysis checks that they all lie in
//      these two panels);
//   2. G is inverted in LDS (gj_tile over its leading ns x ns block, the
//      padding pivots keeping the identity) and stored as the pivot block;
//   3. per 64-column tile: the pivot rows P = G^-1 W go to F and, transposed,
//      to the job's row panel;
// k_mf_front1_rows, one workgroup per front and 64-row tile below the pivots:
//   4. X and G^-1 are staged in LDS and -X G^-1 stored as the pivot columns; per
//      64-column tile the trailing block C is gathered from the children
//      straight into the matrix-core accumulators, takes - X P and is stored
//      (zero in the padding columns nf .. ld-1).
// KP (32 or 64, ns <= KP): the pivots the products run over.  The padding
// pivots ns .. 63 have zero rows in W and zero columns in X, so with
// ns <= 32 every product runs over 32 pivots, half the work.  The result is
// the assembly, scatter and Gauss-Jordan step launches' one,
// [G^-1, G^-1 W; -X G^-1, C - X G^-1 W], up to rounding.  256 threads; every
// global access is bounded by the front (rows < nf, columns < ld) or by a
// child's contribution block (clamped map entries).
template <typename T, int KP>
__global__ void __launch_bounds__(256) k_mf_front1_head(const MfFront *__restrict__ fr, int f0,
                                                      const Job<T> *__restrict__ jobs, const int *__restrict__ ch,
                                                      const int *__restrict__ map, const int64_t *__restrict__ aoff,
                                                      const int *__restrict__ aseg, const int *__restrict__ avid,
                                                      const T *__restrict__ val, T *__restrict__ F)
{
    static_assert(KP == 32 || KP == 64, "32 or 64 pivots");
    using M = Mfma<T>;
    using C = typename M::C;
    using NV = typename NVec<T>::type;
    using P2 = typename Pair2<T>::type;
    constexpr int VW = Vec8<T>::w;   // elements of one 16-byte access
    constexpr int LT = NB + 4;       // padded LDS row stride
    constexpr int NBJ = KP / 32;     // 16-column blocks per wave of a 64 x KP product
    __shared__ T sD[NB * LT];        // the pivot block, then its inverse
    __shared__ __align__(16) int smap[2][kFront1MaxNf];   // the children's maps, -1 past nf
    const MfFront f = fr[f0 + blockIdx.x];
    const Job<T> J = jobs[blockIdx.x];
    T *__restrict__ A = F + f.fo;
    const int nf = f.nf, ld = f.ld, ns = f.ns;
    const int nt = (nf + NB - 1) / NB;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
    const int rw = (w >> 1) * 32, cw = (w & 1) * 32;   // the wave's 32 x 32 quarter of a 64 x 64 product
    const T *cb[2];
    int cld[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const MfFront c = fr[ch[f.co + k]];
        cb[k] = F + c.fo + (int64_t)c.nsp * c.ld + c.nsp;
        cld[k] = c.ld;
    }
    for (int a = t; a < kFront1MaxNf; a += 256) {
        const int q = min(a, nf - 1);
        const int m0 = map[f.mo + q], m1 = map[f.mo + nf + q];
        smap[0][a] = a < nf ? m0 : -1;
        smap[1][a] = a < nf ? m1 : -1;
    }
    // ------------------------------------------------ 1. the panels ----
    const int ld4 = ld >> 2, nq = (ns + 3) >> 2;   // 4-column items of a row; of the true pivot columns
    {
        // no contribution: the padding pivot rows ns .. 63 (identity) and the
        // padding pivot columns of X (zero)
        const int nzq = max(NB / 4 - nq, 1);   // zero items of an X row (none when nq = 16: nz = na0)
        const int na0 = (NB - ns) * ld4, nz = na0 + (nf - NB) * (NB / 4 - nq);
        for (int e = t; e < nz; e += 256) {
            const int e2 = e - na0, q = e / ld4;
            const int r = e < na0 ? ns + q : NB + e2 / nzq;
            const int c = e < na0 ? 4 * (e - q * ld4) : 4 * (nq + e2 % nzq);
            T o[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) o[v] = c + v == r ? T(1) : T(0);
            T *d = A + (int64_t)r * ld + c;
            if constexpr (sizeof(T) == 4) {
                f32x4 x;
                x[0] = o[0]; x[1] = o[1]; x[2] = o[2]; x[3] = o[3];
                *glb(reinterpret_cast<f32x4 *>(d)) = x;
            } else {
                f64x2 x0, x1;
                x0[0] = o[0]; x0[1] = o[1]; x1[0] = o[2]; x1[1] = o[3];
                glb(reinterpret_cast<f64x2 *>(d))[0] = x0;
                glb(reinterpret_cast<f64x2 *>(d))[1] = x1;
            }
        }
    }
    __syncthreads();   // smap
    {
        // the true pivot rows (every column) and the true pivot columns of X
        const int n1 = ns * ld4, nitems = n1 + (nf - NB) * nq;
        constexpr int U = 4;
        for (int e0 = 0; e0 < nitems; e0 += 256 * U) {
            int r[U], c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = min(e0 + u * 256 + t, nitems - 1);
                const int q = e / ld4, e2 = e - n1;
                r[u] = e < n1 ? q : NB + e2 / nq;
                c[u] = e < n1 ? 4 * (e - q * ld4) : 4 * (e2 % nq);
            }
            // every map read and contribution load of the round before the
            // first sum; columns c, c + 1 (c even) are the two unknowns of one
            // node, which a child holds at an even column and the next, so
            // each pair is one load
            T x[2][U][4];
            bool in[2][U][4];
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int ia = smap[k][r[u]];
                    const int4 ib = *reinterpret_cast<const int4 *>(&smap[k][c[u]]);
                    const int ibp[2] = {ib.x, ib.z};
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const bool ok = ia >= 0 && ibp[p] >= 0;
                        const P2 y = ok ? *glb(reinterpret_cast<const P2 *>(cb[k] + ia * cld[k] + ibp[p])) : P2{T(0), T(0)};
                        in[k][u][2 * p] = in[k][u][2 * p + 1] = ok;
                        x[k][u][2 * p] = y[0];
                        x[k][u][2 * p + 1] = y[1];
                    }
                }
            T acc[U][4];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    acc[u][v] = T(0);
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        if (in[k][u][v]) acc[u][v] += x[k][u][v];
                }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                T *d = A + (int64_t)r[u] * ld + c[u];
                if constexpr (sizeof(T) == 4) {
                    f32x4 o;
                    o[0] = acc[u][0]; o[1] = acc[u][1]; o[2] = acc[u][2]; o[3] = acc[u][3];
                    *glb(reinterpret_cast<f32x4 *>(d)) = o;
                } else {
                    f64x2 o0, o1;
                    o0[0] = acc[u][0]; o0[1] = acc[u][1]; o1[0] = acc[u][2]; o1[1] = acc[u][3];
                    glb(reinterpret_cast<f64x2 *>(d))[0] = o0;
                    glb(reinterpret_cast<f64x2 *>(d))[1] = o1;
                }
            }
        }
    }
    __syncthreads();
    // the Jacobian entries the front owns (the sums of the entry scatter), U
    // positions per thread and round with every load of the round issued
    // before the first sum; every position holds at least one entry
    {
        constexpr int U = 4;
        const int64_t uend = f.ao + f.na;
        for (int64_t u0 = f.ao; u0 < uend; u0 += 256 * U) {
            int64_t q[U];
            int s0[U], s1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t uc = min(u0 + u * 256 + t, uend - 1);
                q[u] = aoff[uc];
                s0[u] = aseg[uc];
                s1[u] = aseg[uc + 1];
            }
            int v0[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v0[u] = avid[s0[u]];
            T x[U], old[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                x[u] = T(0) + val[v0[u]];
                old[u] = glb(F)[q[u]];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                for (int e = s0[u] + 1; e < s1[u]; ++e) x[u] += val[avid[e]];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (u0 + u * 256 + t < uend) glb(F)[q[u]] = old[u] + x[u];
        }
    }
    __syncthreads();
    // ---------------------------------------------- 2. pivot inverse ----
    for (int e = t; e < NB * NB; e += 256) {
        const int r = e >> 6, c = e & 63;
        sD[r * LT + c] = glb(A)[(int64_t)r * ld + c];
    }
    {
        T g[16];
        gj_tile<T>(sD, LT, ns, g, J.dinv);   // sD holds the whole inverse on return
#pragma unroll
        for (int q = 0; q < 16; ++q) glb(A)[(int64_t)(w * 16 + q) * ld + l] = sD[(w * 16 + q) * LT + l];
    }
    // ------------------------------------------------ 3. pivot rows ----
    // P = G^-1 W over the KP leading pivot rows (the others are zero in W,
    // hence in P, and already stored as such); the wave's block: the 32 rows
    // rw3 .., 16 NBJ columns cw3 .. of the tile
    const int rw3 = KP == 64 ? rw : 0, cw3 = KP == 64 ? cw : w * 16;
    for (int ct = 1; ct < nt; ++ct) {
        const int c0 = ct * NB;
        T bw[NBJ][KP / 4];
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj) {
            const int cc = c0 + cw3 + bj * 16 + li;
            const int ccl = min(cc, ld - 1);
#pragma unroll
            for (int h = 0; h < KP / 4; ++h) {
                const T x = glb(A)[(int64_t)(4 * h + lk) * ld + ccl];
                bw[bj][h] = cc < nf ? x : T(0);
            }
        }
        __syncthreads();   // every wave holds its part of W before the pivot rows overwrite it
        C d[2][NBJ];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < NBJ; ++bj) {
#pragma unroll
                for (int v = 0; v < 4; ++v) d[bi][bj][v] = T(0);
#pragma unroll
                for (int h = 0; h < KP / 4; ++h)
                    d[bi][bj] = M::opd(sD[(rw3 + bi * 16 + li) * LT + 4 * h + lk], bw[bj][h], d[bi][bj]);
            }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < NBJ; ++bj) {
                const int cc = c0 + cw3 + bj * 16 + li;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = rw3 + bi * 16 + M::drow(lk, v);
                    if (cc < nf) glb(A)[(int64_t)r * ld + cc] = d[bi][bj][v];
                    glb(J.rowk)[(int64_t)cc * NB + r] = d[bi][bj][v];
                }
            }
    }
}

template <typename T, int KP>
__global__ void __launch_bounds__(256, 2) k_mf_front1_rows(const MfFront *__restrict__ fr, int f0, int nfront,
                                                           const Job<T> *__restrict__ jobs,
                                                           const int *__restrict__ ch, const int *__restrict__ map,
                                                           T *__restrict__ F)
{
    using M = Mfma<T>;
    using C = typename M::C;
    using NV = typename NVec<T>::type;
    constexpr int VW = Vec8<T>::w;   // elements of one 16-byte access
    constexpr int LT = NB + 4;       // padded LDS row stride
    constexpr int NBJ = KP / 32;     // 16-column blocks per wave of a 64 x KP product
    __shared__ T sD[NB * LT];        // the inverse pivot block (leading KP x KP)
    __shared__ T sX[NB * LT];        // the tile's pivot columns (leading KP)
    __shared__ __align__(16) int smap[2][kFront1MaxNf];   // the children's maps, -1 past nf
    // row tile major: with a multiple of 8 fronts every row tile of a front
    // runs on its head's die (workgroups go round-robin over the 8 dies), so
    // the pivot rows are read from that die's L2
    const int fl = blockIdx.x % nfront, it = 1 + blockIdx.x / nfront;
    const MfFront f = fr[f0 + fl];
    const Job<T> J = jobs[fl];
    T *__restrict__ A = F + f.fo;
    const int nf = f.nf, ld = f.ld;
    const int nt = (nf + NB - 1) / NB;
    if (it >= nt) return;   // the front has fewer row tiles than the height's largest
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
    const int rw = (w >> 1) * 32, cw = (w & 1) * 32;   // the wave's 32 x 32 quarter of a tile
    const int r0 = it * NB;
    const T *cb[2];
    int cld[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const MfFront c = fr[ch[f.co + k]];
        cb[k] = F + c.fo + (int64_t)c.nsp * c.ld + c.nsp;
        cld[k] = c.ld;
    }
    for (int a = t; a < kFront1MaxNf; a += 256) {
        const int q = min(a, nf - 1);
        const int m0 = map[f.mo + q], m1 = map[f.mo + nf + q];
        smap[0][a] = a < nf ? m0 : -1;
        smap[1][a] = a < nf ? m1 : -1;
    }
    for (int e = t; e < KP * (KP / VW); e += 256) {
        const int r = e / (KP / VW), c = (e % (KP / VW)) * VW;
        const NV x = *glb(reinterpret_cast<const NV *>(A + (int64_t)r * ld + c));
#pragma unroll
        for (int q = 0; q < VW; ++q) sD[r * LT + c + q] = x[q];
    }
    for (int e = t; e < NB * (KP / VW); e += 256) {
        const int r = e / (KP / VW), c = (e % (KP / VW)) * VW;
        const NV x = *glb(reinterpret_cast<const NV *>(A + (int64_t)min(r0 + r, nf - 1) * ld + c));
#pragma unroll
        for (int q = 0; q < VW; ++q) sX[r * LT + c + q] = r0 + r < nf ? x[q] : T(0);
    }
    __syncthreads();
    if (r0 + rw >= nf) return;   // the wave's rows lie past the front (no barrier follows)
    // pivot columns -X G^-1 (columns KP .. 63 stay zero)
    const int cy = KP == 64 ? cw : (w & 1) * 16;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj) {
            C y;
#pragma unroll
            for (int v = 0; v < 4; ++v) y[v] = T(0);
#pragma unroll
            for (int h = 0; h < KP / 4; ++h)
                y = M::opd(sX[(rw + bi * 16 + li) * LT + 4 * h + lk],
                           sD[(4 * h + lk) * LT + cy + bj * 16 + li], y);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + rw + bi * 16 + M::drow(lk, v);
                if (r < nf) glb(A)[(int64_t)r * ld + cy + bj * 16 + li] = -y[v];
            }
        }
    // trailing blocks
    for (int ct = 1; ct < nt; ++ct) {
        const int c0 = ct * NB;
        if (c0 + cw >= ld) break;   // the wave's columns lie past the row
        // every gather of C and every load of the row panel before the first
        // sum; FP32 lanes hold four consecutive columns of a row, two node
        // pairs (even column and the next, at an even child column and the
        // next), one load each
        T x[2][2][2][4];
        bool in[2][2][2][4];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) {
                    if constexpr (M::kSwap) {
                        const int ia = smap[k][r0 + rw + bi * 16 + li];
#pragma unroll
                        for (int p = 0; p < 2; ++p) {
                            const int ib = smap[k][c0 + cw + bj * 16 + 4 * lk + 2 * p];
                            const bool ok = ia >= 0 && ib >= 0;
                            const f32x2 y =
                                ok ? *glb(reinterpret_cast<const f32x2 *>(cb[k] + ia * cld[k] + ib)) : f32x2{0.f, 0.f};
                            in[k][bi][bj][2 * p] = in[k][bi][bj][2 * p + 1] = ok;
                            x[k][bi][bj][2 * p] = y[0];
                            x[k][bi][bj][2 * p + 1] = y[1];
                        }
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int ia = smap[k][r0 + rw + bi * 16 + M::row(li, lk, v)];
                            const int ib = smap[k][c0 + cw + bj * 16 + M::col(li, lk, v)];
                            in[k][bi][bj][v] = ia >= 0 && ib >= 0;
                            x[k][bi][bj][v] = glb(cb[k])[max(ia, 0) * cld[k] + max(ib, 0)];
                        }
                    }
                }
        NV pb[KP / 16][2][4 / VW];
#pragma unroll
        for (int h = 0; h < KP / 16; ++h)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int u = 0; u < 4 / VW; ++u)
                    pb[h][bj][u] = *glb(reinterpret_cast<const NV *>(
                        J.rowk + (int64_t)(c0 + cw + bj * 16 + li) * NB + 16 * h + 4 * lk + u * VW));
        C acc[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    T s = T(0);
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        if (in[k][bi][bj][v]) s += x[k][bi][bj][v];
                    acc[bi][bj][v] = s;
                }
        // C - X P, k = 16 h + 4 lk + e
#pragma unroll
        for (int h = 0; h < KP / 16; ++h) {
            T av[2][4], bv[2][4];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) {
                const T *p = &sX[(rw + bi * 16 + li) * LT + 16 * h + 4 * lk];
#pragma unroll
                for (int e = 0; e < 4; ++e) av[bi][e] = -p[e];
            }
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int u = 0; u < 4 / VW; ++u)
#pragma unroll
                    for (int e = 0; e < VW; ++e) bv[bj][u * VW + e] = pb[h][bj][u][e];
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = M::op(av[bi][e], bv[bj][e], acc[bi][bj]);
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                if constexpr (M::kSwap) {
                    // one row, four consecutive columns per lane
                    const int r = r0 + rw + bi * 16 + li, c = c0 + cw + bj * 16 + 4 * lk;
                    if (r < nf && c < ld) {
                        f32x4 o;
#pragma unroll
                        for (int v = 0; v < 4; ++v) o[v] = c + v < nf ? acc[bi][bj][v] : T(0);
                        *glb(reinterpret_cast<f32x4 *>(A + (int64_t)r * ld + c)) = o;
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = r0 + rw + bi * 16 + M::row(li, lk, v);
                        const int c = c0 + cw + bj * 16 + M::col(li, lk, v);
                        if (r < nf && c < ld) glb(A)[(int64_t)r * ld + c] = c < nf ? acc[bi][bj][v] : T(0);
                    }
                }
            }
    }
}

// ------------------------------------------------------ launch probe ------
LaunchProbe::~LaunchProbe()
{
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
}

static int probe_flush(LaunchProbe &p)
{
    if (!p.used) return 0;
    V2D_HIP(hipEventSynchronize(p.ev[2 * p.used - 1]));
    for (int i = 0; i < p.used; ++i) {
        float ms = 0.f;
        V2D_HIP(hipEventElapsedTime(&ms, p.ev[2 * i], p.ev[2 * i + 1]));
        p.ms += ms;
    }
    p.used = 0;
    return 0;
}

// An event pair around the update launches of steps k0 = kbeg, kbeg + 64, ..
// < kend (back to back on the stream: the pair times the launches and the
// gaps between them, not a dispatch per launch).
template <typename T>
static int probe_mark(LaunchProbe *p, const Batch<T> &B, int kbeg, int kend, bool begin, hipStream_t s)
{
    if (!p) return 0;
    if (begin) {
        if (p->used == LaunchProbe::kPairs && probe_flush(*p)) return -1;
        for (int q = 0; q < 2; ++q)
            if (!p->ev[2 * p->used + q]) V2D_HIP(hipEventCreate(&p->ev[2 * p->used + q]));
        V2D_HIP(hipEventRecord(p->ev[2 * p->used], s));
        return 0;
    }
    V2D_HIP(hipEventRecord(p->ev[2 * p->used + 1], s));
    ++p->used;
    for (int k0 = kbeg; k0 < kend; k0 += NB) {
        ++p->launches;
        for (int z = 0; z < B.nj; ++z) {
            const double n = B.j[z].n;
            if (k0 >= B.j[z].n) continue;
            const double kb = std::min(NB, B.j[z].n - k0);
            p->flops += 2.0 * (n - kb) * n * kb;
            p->bytes += sizeof(T) * (2.0 * (n - kb) * n + (n - kb) * kb + kb * n);
        }
    }
    return 0;
}

// In-place Gauss-Jordan inverse of the batch's dense matrices.
template <typename T>
static int gj_invert_b(const Batch<T> &B, DirectWork<T> &W, LaunchProbe *probe, hipStream_t s)
{
    int nmax = 0;
    for (int z = 0; z < B.nj; ++z) nmax = std::max(nmax, B.j[z].n);
    const int ut = B.kind == kUpdateValu128 ? 128 : 64;
    const unsigned nt = (unsigned)((nmax + ut - 1) / ut);
    const unsigned np = (unsigned)((nmax + RP - 1) / RP);
    const unsigned nz = (unsigned)B.nj;
    hipLaunchKernelGGL(k_gj_diag_b<T>, dim3(1, 1, nz), dim3(256), 0, s, B, 0);
    if (B.kind == kUpdateMfma64) {
        // step 0's panels from the row-panel kernel, then one dataflow launch
        // per step, the panel buffers alternating
        hipLaunchKernelGGL(k_gj_rowpanel_b<T>, dim3(np, 2, nz), dim3(256), 0, s, B, 0);
        Batch<T> Bs = B;
        if (probe_mark(probe, B, 0, nmax, true, s)) return -1;
        for (int k0 = 0; k0 < nmax; k0 += NB) {
            hipLaunchKernelGGL(k_gj_step_mfma_b<T>, dim3(nt * nt * nz), dim3(256), 0, s, Bs, k0,
                               ++W.epoch, W.flags, W.err);
            for (int z = 0; z < B.nj; ++z) {
                std::swap(Bs.j[z].colk, Bs.j[z].colk_w);
                std::swap(Bs.j[z].rowk, Bs.j[z].rowk_w);
            }
        }
        if (probe_mark(probe, B, 0, nmax, false, s)) return -1;
        V2D_LAUNCH();
        return 0;
    }
    for (int k0 = 0; k0 < nmax; k0 += NB) {
        hipLaunchKernelGGL(k_gj_rowpanel_b<T>, dim3(np, 2, nz), dim3(256), 0, s, B, k0);
        if (probe_mark(probe, B, k0, k0 + 1, true, s)) return -1;
        if (ut == 128)
            hipLaunchKernelGGL((k_gj_update_b<T, 128>), dim3(nt * nt * nz), dim3(256), 0, s, B, k0);
        else
            hipLaunchKernelGGL((k_gj_update_b<T, 64>), dim3(nt * nt * nz), dim3(256), 0, s, B, k0);
        if (probe_mark(probe, B, k0, k0 + 1, false, s)) return -1;
    }
    V2D_LAUNCH();
    return 0;
}

// Gauss-Jordan sweeps of a device table of jobs (MFMA dataflow steps): every
// job's pivots [0, np) are eliminated from its n x n matrix in place, np a
// multiple of 64, n at most nmax.  flags holds one word per job.
// gj_table_begin inverts every first pivot block and forms the step-0 panels
// and pivot rows (a first step, compact grid); gj_table_step runs the step at
// pivot k0 (k0 = 0, 64, .. in order) in the given mode.
template <typename T>
int gj_table_begin(const Job<T> *jobs, int nj, int nmax, int *flags, int *err, int &epoch, hipStream_t s)
{
    if (nj > 65535) {
        set_error("Gauss-Jordan table: more than 65535 jobs in one launch");
        return -1;
    }
    const unsigned nt = (unsigned)((nmax + NB - 1) / NB);
    if (nt < 2) {
        set_error("Gauss-Jordan table: a job of one tile has no step");
        return -1;
    }
    TableBatch<T> B{jobs, nj, kUpdateMfma64, 2, (int)(2 * nt - 1), 1};
    hipLaunchKernelGGL((k_gj_step_mfma_b<T, TableBatch<T>, kStepBegin>), dim3((2 * nt - 1) * (unsigned)nj), dim3(256),
                       0, s, B, -NB, ++epoch, flags, err);
    V2D_LAUNCH();
    return 0;
}

template <typename T>
int gj_table_step(const Job<T> *jobs, int nj, int nmax, int k0, int mode, bool compact, int *flags, int *err,
                  int &epoch, hipStream_t s)
{
    const unsigned nt = (unsigned)((nmax + NB - 1) / NB);
    if (nt < 2) {
        set_error("Gauss-Jordan table: a job of one tile has no step");
        return -1;
    }
    compact = compact && mode == kStepLight;
    const unsigned tpj = compact ? 2 * nt - 1 : nt * nt;
    TableBatch<T> B{jobs, nj, kUpdateMfma64, (k0 / NB) % 3, (int)tpj, compact ? 1 : 0};
    const dim3 g(tpj * (unsigned)nj);
    switch (mode) {
    case kStepFull:
        hipLaunchKernelGGL((k_gj_step_mfma_b<T, TableBatch<T>, kStepFull>), g, dim3(256), 0, s, B, k0, ++epoch, flags,
                           err);
        break;
    case kStepLight:
        hipLaunchKernelGGL((k_gj_step_mfma_b<T, TableBatch<T>, kStepLight>), g, dim3(256), 0, s, B, k0, ++epoch,
                           flags, err);
        break;
    case kStepHeavy:
        if (k0 < NB) {
            set_error("Gauss-Jordan table: a heavy step needs a previous step");
            return -1;
        }
        hipLaunchKernelGGL((k_gj_step_mfma_b<T, TableBatch<T>, kStepHeavy>), g, dim3(256), 0, s, B, k0, ++epoch,
                           flags, err);
        break;
    default:
        set_error("Gauss-Jordan table: unknown step mode");
        return -1;
    }
    V2D_LAUNCH();
    return 0;
}

template int gj_table_begin<float>(const Job<float> *, int, int, int *, int *, int &, hipStream_t);
template int gj_table_begin<double>(const Job<double> *, int, int, int *, int *, int &, hipStream_t);
template int gj_table_step<float>(const Job<float> *, int, int, int, int, bool, int *, int *, int &, hipStream_t);
template int gj_table_step<double>(const Job<double> *, int, int, int, int, bool, int *, int *, int &, hipStream_t);

template <typename T>
int mf_front1(const MfFront *fr, int f0, int nfront, int nfmax, int kp, const Job<T> *jobs, const int *ch,
              const int *map, const int64_t *aoff, const int *aseg, const int *avid, const T *val, T *F, hipStream_t s)
{
    if (nfront <= 0) return 0;
    const int nrt = (nfmax + NB - 1) / NB - 1;   // row tiles below the pivots of the largest front
    const dim3 gh((unsigned)nfront), gr((unsigned)((int64_t)nfront * nrt));
    if (kp == 32) {
        hipLaunchKernelGGL((k_mf_front1_head<T, 32>), gh, dim3(256), 0, s, fr, f0, jobs, ch, map, aoff, aseg, avid,
                           val, F);
        hipLaunchKernelGGL((k_mf_front1_rows<T, 32>), gr, dim3(256), 0, s, fr, f0, nfront, jobs, ch, map, F);
    } else {
        hipLaunchKernelGGL((k_mf_front1_head<T, 64>), gh, dim3(256), 0, s, fr, f0, jobs, ch, map, aoff, aseg, avid,
                           val, F);
        hipLaunchKernelGGL((k_mf_front1_rows<T, 64>), gr, dim3(256), 0, s, fr, f0, nfront, jobs, ch, map, F);
    }
    V2D_LAUNCH();
    return 0;
}
template int mf_front1<float>(const MfFront *, int, int, int, int, const Job<float> *, const int *, const int *,
                              const int64_t *, const int *, const int *, const float *, float *, hipStream_t);
template int mf_front1<double>(const MfFront *, int, int, int, int, const Job<double> *, const int *, const int *,
                               const int64_t *, const int *, const int *, const double *, double *, hipStream_t);

// ------------------------------------------------------------ factor ------
// Twisted block factorisation: the bottom chain eliminates levels 0..M-1 from
// the wall outward (S_l = D_l - L_l S_{l-1}^-1 U_{l-1}), the top chain levels
// nlev-1..M+1 from the far field inward (S_l = D_l - U_l S_{l+1}^-1 L_{l+1}),
// both in the same launches; the middle level M takes both Schur updates.
template <typename T>
static Job<T> make_job(const DirectPattern &P, DirectWork<T> &W, int l, int slot)
{
    Job<T> J;
    J.S = W.sinv + P.soff[l];
    J.n = P.n[l];
    J.np = J.n;
    J.ld = P.ld[l];
    J.l = l;
    J.voff = P.voff[l];
    J.colk = W.colk + (int64_t)slot * P.nmax * NB;
    J.dinv = W.dinv + slot * NB * NB;
    J.tmat = W.tmat + (int64_t)slot * P.nmax * P.ldmax;
    const bool df = W.update_kind == kUpdateMfma64;
    J.rowk = df ? W.rowk + (int64_t)slot * P.ldmax * NB : nullptr;
    J.colk_w = df ? W.colk2 + (int64_t)slot * P.nmax * NB : nullptr;
    J.rowk_w = df ? W.rowk2 + (int64_t)slot * P.ldmax * NB : nullptr;
    J.colk_x = J.rowk_x = nullptr;
    return J;
}

// Dense block of every job of B (Schur sources: level srcp[z][k] where
// schur[z][k]), then its inverse.  A source p = l-1 couples through U_{l-1}
// (columns of side 0) and the rows L_l; a source p = l+1 through L_{l+1}
// (side 1) and the rows U_l.  At most two sources in all: the two chains take
// one each, the middle level both.
template <typename T>
static int build_and_invert(const DirectPattern &P, Batch<T> &B, const T *val, bool schur[2][2],
                            int srcp[2][2], DirectWork<T> &W, hipStream_t s)
{
    if (!P.pad_ok || !W.cfL || !W.cfU || !W.cfC) {
        set_error("block LU: padded coupling tables or their values are missing");
        return -1;
    }
    BuildBatch<T> bb{};
    int nt = 0, nmax = 0, ldmax = 0, npmax = 0, kcmax = 1;
    for (int z = 0; z < B.nj; ++z) {
        const Job<T> &J = B.j[z];
        LevelBuild<T> &lv = bb.lv[z];
        lv.S = J.S;
        lv.n = J.n;
        lv.ld = J.ld;
        lv.dptr = P.d_Dptr + J.voff;
        lv.nsrc = 0;
        nmax = std::max(nmax, J.n);
        ldmax = std::max(ldmax, J.ld);
        for (int k = 0; k < 2; ++k) {
            if (!schur[z][k]) continue;
            const int l = J.l, p = srcp[z][k], side = p == l - 1 ? 0 : 1;
            if (nt == 2 || (p != l - 1 && p != l + 1)) {
                set_error("block LU: invalid Schur source");
                return -1;
            }
            SchurSrc<T> &A = bb.src[nt];
            A.Sprev = W.sinv + P.soff[p];
            A.np = P.n[p];
            A.ldp = P.ld[p];
            A.jx = P.d_jxC + P.cx_off[side][l];
            A.cf = W.cfC + P.cx_off[side][l];
            A.kc = P.cx_w[side][l];
            A.tm = W.tmat + (int64_t)nt * P.nmax * P.ldmax;
            A.n = J.n;
            A.ld = J.ld;
            lv.tm[lv.nsrc] = A.tm;
            lv.ix[lv.nsrc] = (side == 0 ? P.d_ixL : P.d_ixU) + (int64_t)kCouple * J.voff;
            lv.cf[lv.nsrc] = (side == 0 ? W.cfL : W.cfU) + (int64_t)kCouple * J.voff;
            lv.ixoff[lv.nsrc] = (int)P.voff[p];
            ++lv.nsrc;
            npmax = std::max(npmax, A.np);
            kcmax = std::max(kcmax, A.kc);
            ++nt;
        }
    }
    if (nt) {
        const dim3 g((unsigned)((nmax + 63) / 64), (unsigned)((npmax + 4 * kTmRows - 1) / (4 * kTmRows)),
                     (unsigned)nt);
        switch (kcmax) {   // the widest column of the launch (narrower ones read as 0 past theirs)
        case 1: case 2: case 3: case 4: hipLaunchKernelGGL((k_schur_tm<T, 4>), g, dim3(256), 0, s, bb); break;
        case 5: hipLaunchKernelGGL((k_schur_tm<T, 5>), g, dim3(256), 0, s, bb); break;
        case 6: hipLaunchKernelGGL((k_schur_tm<T, 6>), g, dim3(256), 0, s, bb); break;
        default: hipLaunchKernelGGL((k_schur_tm<T, 8>), g, dim3(256), 0, s, bb); break;
        }
    }
    hipLaunchKernelGGL(k_build_level<T>,
                       dim3((unsigned)(ldmax / 64), (unsigned)((nmax + 4 * kSbRows - 1) / (4 * kSbRows)),
                            (unsigned)B.nj),
                       dim3(256), 0, s, bb, P.d_D, val);
    V2D_LAUNCH();
    return gj_invert_b(B, W, W.probe, s);
}

template <typename T>
static int direct_factor_levels(const DirectPattern &P, DirectWork<T> &W, const T *val, hipStream_t s);
template <typename T>
int direct_couplings(const DirectPattern &P, DirectWork<T> &W, const T *val, hipStream_t s);

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

template <typename T>
int direct_factor(const DirectPattern &P, DirectWork<T> &W, const T *val, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, W.err, 0);
    V2D_LAUNCH();
    if (direct_couplings(P, W, val, s)) return -1;
    if (direct_factor_levels(P, W, val, s)) return -1;
    if (W.update_kind == kUpdateMfma64) {
        int e = 0;
        V2D_HIP(hipMemcpyAsync(&e, W.err, sizeof(int), hipMemcpyDeviceToHost, s));
        V2D_HIP(hipStreamSynchronize(s));
        if (e) {
            set_error("block LU: a Gauss-Jordan step waited past its spin limit for the pivot "
                      "inverse (results invalid)");
            return -1;
        }
    }
    return 0;
}

template <typename T>
static int direct_factor_levels(const DirectPattern &P, DirectWork<T> &W, const T *val, hipStream_t s)
{
    const int kind = W.update_kind;
    const int nlev = P.L.nlev;
    const int M = (nlev - 1) / 2;            // meeting level
    const int nbot = M, ntop = nlev - 1 - M;
    const int steps = std::max(nbot, ntop);
    for (int st = 0; st < steps; ++st) {
        Batch<T> B{};
        bool schur[2][2] = {{false, false}, {false, false}};
        int srcp[2][2] = {{0, 0}, {0, 0}};
        int nj = 0;
        if (st < nbot) {
            const int l = st;
            B.j[nj] = make_job(P, W, l, nj);
            if (l > 0) { schur[nj][0] = true; srcp[nj][0] = l - 1; }
            ++nj;
        }
        if (st < ntop) {
            const int l = nlev - 1 - st;
            B.j[nj] = make_job(P, W, l, nj);
            if (st > 0) { schur[nj][0] = true; srcp[nj][0] = l + 1; }
            ++nj;
        }
        B.nj = nj;
        B.kind = kind;
        if (build_and_invert(P, B, val, schur, srcp, W, s)) return -1;
    }
    Batch<T> B{};
    bool schur[2][2] = {{M > 0, M < nlev - 1}, {false, false}};
    int srcp[2][2] = {{M - 1, M + 1}, {0, 0}};
    if (!schur[0][0]) { schur[0][0] = schur[0][1]; srcp[0][0] = srcp[0][1]; schur[0][1] = false; }
    B.nj = 1;
    B.kind = kind;
    B.j[0] = make_job(P, W, M, 0);
    if (build_and_invert(P, B, val, schur, srcp, W, s)) return -1;
    return W.probe ? probe_flush(*W.probe) : 0;
}
template int direct_factor<double>(const DirectPattern &, DirectWork<double> &, const double *, hipStream_t);
template int direct_factor<float>(const DirectPattern &, DirectWork<float> &, const float *, hipStream_t);

// ------------------------------------------------------------- solve ------
__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sparse right-hand side of a level:  y = [b_l] + sgn * sum_t list_t(r) . vec_t
// (a term's vector is level-ordered, or the global field-blocked vector
// addressed through the level map when glob is set).
template <typename T>
struct RhsTerm {
    const int *ptr;
    const int2 *ent;
    const T *vec;
    int vlev, glob;
};
template <typename T>
struct RhsJob {
    int l, n, nt;
    const T *b;        // global right-hand side or null
    RhsTerm<T> t[2];
};

// 16-byte vectors of T and their FP64 dot with an LDS vector.
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; static constexpr int w = 4; };
template <> struct Vec16<double> { using type = double2; static constexpr int w = 2; };

template <typename T>
__device__ __forceinline__ double dot16(const typename Vec16<T>::type &a, const T *b);
template <>
__device__ __forceinline__ double dot16<float>(const float4 &a, const float *b)
{
    return (double)a.x * b[0] + (double)a.y * b[1] + (double)a.z * b[2] + (double)a.w * b[3];
}
template <>
__device__ __forceinline__ double dot16<double>(const double2 &a, const double *b)
{
    return a.x * b[0] + a.y * b[1];
}

// One solve step of up to two levels, as two launches:
//   k_level_rhs_b:  y = [b_l] + sgn_rhs * sum_t list_t(r) . vec_t   (one row per thread)
//   k_level_gemv_b: out = [base] + sgn_mv * S y
// The GEMV issues the loads of y and then of its S rows (one row per
// wavefront, up to kSweepPre 16-byte vectors per lane held in registers; one
// row measured faster than two or three), stages y in LDS while the rows are
// in flight, then forms the dot products (row parts past the prefetch are
// streamed).  Output level-ordered (glob = 0) or scattered to the global
// vector.
// A 16-byte vector of a block-inverse row, loaded non-temporally: the rows
// stream through once per solve (35 MB per step at 1025^2), so they should
// not displace the right-hand sides and coupling lists from the caches.
template <typename T>
__device__ __forceinline__ typename Vec16<T>::type load_row_nt(const typename Vec16<T>::type *p)
{
    using N = T __attribute__((ext_vector_type(Vec16<T>::w)));
    const N x = __builtin_nontemporal_load(reinterpret_cast<const N *>(p));
    typename Vec16<T>::type v;
    __builtin_memcpy(&v, &x, sizeof(v));
    return v;
}

constexpr int kSweepRowsPerWave = 1;
constexpr int kSweepPre = 9;
constexpr int kGemvThreads = 256;   // four wavefronts per workgroup (measured faster than two or eight)

template <typename T>
struct SweepJob {
    RhsJob<T> rhs;
    T *y;            // level-ordered scratch for the right-hand side
    const T *S;
    int ld, glob;
    const T *base;
    T *out;
    T *lout;         // level-ordered copy of the output (or null)
    // right-hand side from the padded coupling lists (pterm != 0): bit 1 the
    // couplings to level l-1, bit 2 to level l+1; source psrc, level-ordered
    int pterm;
    const T *psrc;
    int fg;          // workgroups of the fused step (k_level_step_fused)
};
template <typename T>
struct SweepBatch {
    SweepJob<T> j[2];
};

// The same right-hand side from the padded coupling lists: one load round
// for the coefficients and source indices, one for the source entries (the
// padding adds zeros, so the sums equal the list sums above).
template <typename T>
__global__ void __launch_bounds__(256) k_level_rhs_pad(SweepBatch<T> B, Levels L, int64_t N, const T *__restrict__ cfL,
                                                       const T *__restrict__ cfU, const int *__restrict__ ixL,
                                                       const int *__restrict__ ixU, double sgn_rhs)
{
    const SweepJob<T> &J = B.j[blockIdx.y];
    const RhsJob<T> &R = J.rhs;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R.n) return;
    const int n = R.n;
    const int64_t cb = (int64_t)kCouple * lvl_voff(L, R.l);
    double sum = 0.0;
    if (J.pterm & 1)
#pragma unroll
        for (int e = 0; e < kCouple; ++e) sum += (double)cfL[cb + e * n + r] * (double)J.psrc[ixL[cb + e * n + r]];
    if (J.pterm & 2)
#pragma unroll
        for (int e = 0; e < kCouple; ++e) sum += (double)cfU[cb + e * n + r] * (double)J.psrc[ixU[cb + e * n + r]];
    J.y[r] = (T)((R.b ? (double)R.b[L.global(r, N, R.l)] : 0.0) + sgn_rhs * sum);
}

template <typename T>
__global__ void __launch_bounds__(256) k_level_rhs_b(SweepBatch<T> B, Levels L, int64_t N,
                                                     const T *__restrict__ val, double sgn_rhs)
{
    const SweepJob<T> &J = B.j[blockIdx.y];
    const RhsJob<T> &R = J.rhs;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R.n) return;
    double sum = 0.0;
    for (int k = 0; k < R.nt; ++k) {
        const RhsTerm<T> &t = R.t[k];
        for (int e = t.ptr[r]; e < t.ptr[r + 1]; ++e) {
            const int2 x = t.ent[e];
            const double v = t.glob ? (double)t.vec[L.global(x.x, N, t.vlev)] : (double)t.vec[x.x];
            sum += (double)val[x.y] * v;
        }
    }
    J.y[r] = (T)((R.b ? (double)R.b[L.global(r, N, R.l)] : 0.0) + sgn_rhs * sum);
}

template <typename T>
__global__ void __launch_bounds__(kGemvThreads) k_level_gemv_b(SweepBatch<T> B, Levels L, int64_t N,
                                                      double sgn_mv)
{
    using V = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::w;
    constexpr int RW = kSweepRowsPerWave;
    const SweepJob<T> &J = B.j[blockIdx.y];
    const int n = J.rhs.n, ld = J.ld;
    const int row0 = blockIdx.x * (kGemvThreads / 64 * RW);
    if (row0 >= n) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nv = (n + W - 1) / W;
    // y's loads go out first and the S rows' after them: vector memory
    // returns in order, so the LDS stores of y wait for y alone while the
    // rows are still in flight
    constexpr int YP = (W == 4 ? 768 : 1280) / kGemvThreads;   // y vectors per thread loaded ahead
    const V *yv = reinterpret_cast<const V *>(J.y);
    V yr[YP];
#pragma unroll
    for (int p = 0; p < YP; ++p) yr[p] = yv[min((int)threadIdx.x + kGemvThreads * p, nv - 1)];
    __builtin_amdgcn_sched_barrier(0);   // keep y's loads ahead of the rows'
    // unconditional loads of clamped columns (no branch between them); the
    // columns past the row are skipped in the sums
    V pre[RW][kSweepPre];
#pragma unroll
    for (int q = 0; q < RW; ++q) {
        const int r = min(row0 + wave * RW + q, n - 1);
        const V *row = reinterpret_cast<const V *>(J.S + (int64_t)r * ld);
#pragma unroll
        for (int u = 0; u < kSweepPre; ++u) pre[q][u] = load_row_nt<T>(row + min(lane + 64 * u, nv - 1));
    }
    // y into LDS as 16-byte vectors, entries past n zeroed
    V *vsv = reinterpret_cast<V *>(smem);
#pragma unroll
    for (int p = 0; p < YP; ++p) {
        const int c = threadIdx.x + kGemvThreads * p;
        if (c < nv) {
            V x = yr[p];
#pragma unroll
            for (int e = 0; e < W; ++e)
                if (c * W + e >= n) reinterpret_cast<T *>(&x)[e] = T(0);
            vsv[c] = x;
        }
    }
    for (int c = threadIdx.x + kGemvThreads * YP; c < nv; c += kGemvThreads) {
        V x = yv[c];
#pragma unroll
        for (int e = 0; e < W; ++e)
            if (c * W + e >= n) reinterpret_cast<T *>(&x)[e] = T(0);
        vsv[c] = x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RW; ++q) {
        const int r = row0 + wave * RW + q;
        const V *row = reinterpret_cast<const V *>(J.S + (int64_t)min(r, n - 1) * ld);
        double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
        for (int u = 0; u < kSweepPre; ++u) {
            const int c = lane + 64 * u;
            if (c < nv) {
                const V yc = vsv[c];
                if (u & 1) acc1 += dot16<T>(pre[q][u], reinterpret_cast<const T *>(&yc));
                else acc0 += dot16<T>(pre[q][u], reinterpret_cast<const T *>(&yc));
            }
        }
        for (int c = lane + 64 * kSweepPre; c < nv; c += 128) {
            const V y0 = vsv[c];
            acc0 += dot16<T>(load_row_nt<T>(row + c), reinterpret_cast<const T *>(&y0));
            if (c + 64 < nv) {
                const V y1 = vsv[c + 64];
                acc1 += dot16<T>(load_row_nt<T>(row + c + 64), reinterpret_cast<const T *>(&y1));
            }
        }
        const double acc = wsum(acc0 + acc1);
        if (lane == 0 && r < n) {
            const double b = J.base ? (double)J.base[r] : 0.0;
            const T o = (T)(b + sgn_mv * acc);
            if (J.glob) J.out[L.global(r, N, J.rhs.l)] = o;
            else J.out[r] = o;
            if (J.lout) J.lout[r] = o;
        }
    }
}

// One FP32 solve step as a single launch, for steps whose right-hand side
// takes the couplings of one neighbouring level (every forward step but the
// first, every backward step).  Every workgroup forms the whole level
// right-hand side in LDS itself, from the padded coupling lists, with the
// operations of k_level_rhs_pad, so it holds the same values bit for bit.
// 16 wavefronts per workgroup; the G workgroups of a level own 16 rows each
// (row 16 g + wave, prefetched in registers and summed as in k_level_gemv_b,
// so these rows equal the launch pair's bit for bit).  The rows past 16 G
// are split rows: workgroup g takes rows 16 G + E g .. 16 G + E g + E - 1,
// each wavefront one sixteenth of their columns, summed through LDS in
// wavefront order.  G is at most the compute units per job, so the whole
// grid is resident at once and each right-hand side is formed once per
// compute unit.  Loads go out in the order they are needed: the thread's
// coefficients, source indices and b first, then the matrix rows, then the
// source gathers, which wait for the indices only (vector memory returns in
// order).
constexpr int kFusedThreads = 1024;
constexpr int kFusedRpt = 3;                             // right-hand-side rows per thread, at most
constexpr int kFusedMaxRows = kFusedRpt * kFusedThreads;
constexpr int kFusedSplit = 2;                           // split rows per workgroup, at most

template <int RPT, int E>
__global__ void __launch_bounds__(kFusedThreads) k_level_step_fused(SweepBatch<float> B, Levels L, int64_t N,
                                                                    const float *__restrict__ cfL,
                                                                    const float *__restrict__ cfU,
                                                                    const int *__restrict__ ixL,
                                                                    const int *__restrict__ ixU, double sgn_rhs,
                                                                    double sgn_mv)
{
    using V = float4;
    constexpr int W = 4;
    const SweepJob<float> &J = B.j[blockIdx.y];
    const RhsJob<float> &R = J.rhs;
    const int n = R.n, ld = J.ld, G = J.fg, g = blockIdx.x;
    if (g >= G) return;
    __shared__ __align__(16) float ys[kFusedMaxRows];
    __shared__ double part[kFusedSplit][16];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int nv = (n + W - 1) / W;
    // 1. the thread's right-hand-side inputs (rows tid + 1024 q, clamped)
    const int64_t cb = (int64_t)kCouple * lvl_voff(L, R.l);
    const float *cf = J.pterm == 2 ? cfU : cfL;
    const int *ix = J.pterm == 2 ? ixU : ixL;
    // loaded unconditionally (a branch would wait for the load at its join):
    // without b, any vector of 2N entries stands in and its value is unused
    const float *bsrc = R.b ? R.b : J.psrc;
    float cv[RPT][kCouple], bv[RPT];
    int iv[RPT][kCouple];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = min(tid + kFusedThreads * q, n - 1);
#pragma unroll
        for (int e = 0; e < kCouple; ++e) {
            cv[q][e] = cf[cb + e * n + r];
            iv[q][e] = ix[cb + e * n + r];
        }
        bv[q] = bsrc[L.global(r, N, R.l)];
    }
    __builtin_amdgcn_sched_barrier(0);
    // 2. the wavefront's row and its slices of the split rows
    const V *row = reinterpret_cast<const V *>(J.S + (int64_t)min(16 * g + wave, n - 1) * ld);
    V pre[kSweepPre];
#pragma unroll
    for (int u = 0; u < kSweepPre; ++u) pre[u] = load_row_nt<float>(row + min(lane + 64 * u, nv - 1));
    const int sl = (nv + 15) / 16;                   // vectors per wavefront slice (at most 64)
    const int sc = min(wave * sl + min(lane, sl - 1), nv - 1);
    V sp[E > 0 ? E : 1];
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const int sr = min(16 * G + E * g + t, n - 1);
        sp[t] = load_row_nt<float>(reinterpret_cast<const V *>(J.S + (int64_t)sr * ld) + sc);
    }
    __builtin_amdgcn_sched_barrier(0);
    // 3. the source gathers; the empty asm keeps the address arithmetic on
    // the indices (and so the wait for them) behind the row loads
    float sv[RPT][kCouple];
#pragma unroll
    for (int q = 0; q < RPT; ++q)
#pragma unroll
        for (int e = 0; e < kCouple; ++e) {
            asm volatile("" : "+v"(iv[q][e]));
            sv[q][e] = J.psrc[iv[q][e]];
        }
    // 4. the right-hand side into LDS: every row r < RPT * 1024 is written,
    // so the entries from n up to the last 16-byte vector are zeros (an
    // unconditional store: a conditional one lets the compiler sink the
    // gathers behind the other loads)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = tid + kFusedThreads * q;
        double sum = 0.0;
#pragma unroll
        for (int e = 0; e < kCouple; ++e) sum += (double)cv[q][e] * (double)sv[q][e];
        const float y = (float)((R.b ? (double)bv[q] : 0.0) + sgn_rhs * sum);
        ys[r] = r < n ? y : 0.f;
    }
    __syncthreads();
    const V *vsv = reinterpret_cast<const V *>(ys);
    auto store = [&](int r, double acc) {
        const double b = J.base ? (double)J.base[r] : 0.0;
        const float o = (float)(b + sgn_mv * acc);
        if (J.glob) J.out[L.global(r, N, R.l)] = o;
        else J.out[r] = o;
        if (J.lout) J.lout[r] = o;
    };
    // 5. the wavefront's row
    {
        const int r = 16 * g + wave;
        double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
        for (int u = 0; u < kSweepPre; ++u) {
            const int c = lane + 64 * u;
            if (c < nv) {
                const V yc = vsv[c];
                if (u & 1) acc1 += dot16<float>(pre[u], reinterpret_cast<const float *>(&yc));
                else acc0 += dot16<float>(pre[u], reinterpret_cast<const float *>(&yc));
            }
        }
        for (int c = lane + 64 * kSweepPre; c < nv; c += 128) {
            const V y0 = vsv[c];
            acc0 += dot16<float>(load_row_nt<float>(row + c), reinterpret_cast<const float *>(&y0));
            if (c + 64 < nv) {
                const V y1 = vsv[c + 64];
                acc1 += dot16<float>(load_row_nt<float>(row + c + 64), reinterpret_cast<const float *>(&y1));
            }
        }
        const double acc = wsum(acc0 + acc1);
        if (lane == 0 && r < n) store(r, acc);
    }
    // 6. the split rows
    if constexpr (E > 0) {
#pragma unroll
        for (int t = 0; t < E; ++t) {
            double a = 0.0;
            if (lane < sl && wave * sl + lane < nv) {
                const V yc = vsv[wave * sl + lane];
                a = dot16<float>(sp[t], reinterpret_cast<const float *>(&yc));
            }
            a = wsum(a);
            if (lane == 0) part[t][wave] = a;
        }
        __syncthreads();
        if (wave < E && lane == 0) {
            const int sr = 16 * G + E * g + wave;
            if (sr < n) {
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < 16; ++w) t += part[wave][w];
                store(sr, t);
            }
        }
    }
}

// Grid of the fused step: per job G = min(ceil(n / 16), compute units / jobs)
// workgroups and E split rows per workgroup; false when the step does not fit
// the fused kernel (then the launch pair runs).
template <typename T>
static bool fused_plan(SweepBatch<T> &B, int nj, int nmax, const DirectPattern &P, int &grid, int &rpt, int &e)
{
    if (!std::is_same<T, float>::value || nmax > kFusedMaxRows || P.ncu < nj) return false;
    grid = 0; e = 0;
    for (int j = 0; j < nj; ++j) {
        const int pt = B.j[j].pterm;
        if (pt != 1 && pt != 2) return false;
        const int n = B.j[j].rhs.n;
        const int G = std::min((n + 15) / 16, P.ncu / nj);
        const int extra = std::max(0, n - 16 * G);
        e = std::max(e, (extra + G - 1) / G);
        B.j[j].fg = G;
        grid = std::max(grid, G);
    }
    rpt = (nmax + kFusedThreads - 1) / kFusedThreads;
    return e <= kFusedSplit;
}

template <int RPT>
static void launch_fused_e(int e, dim3 grid, const SweepBatch<float> &B, const DirectPattern &P,
                           const DirectWork<float> &W, double sgn_rhs, double sgn_mv, hipStream_t s)
{
    auto go = [&](auto k) {
        hipLaunchKernelGGL(k, grid, dim3(kFusedThreads), 0, s, B, P.L, P.N, (const float *)W.cfL,
                           (const float *)W.cfU, (const int *)P.d_ixL, (const int *)P.d_ixU, sgn_rhs, sgn_mv);
    };
    if (e == 0) go(k_level_step_fused<RPT, 0>);
    else if (e == 1) go(k_level_step_fused<RPT, 1>);
    else go(k_level_step_fused<RPT, 2>);
}

template <typename T>
static int launch_sweep(SweepBatch<T> &B, int nj, int nmax, const DirectPattern &P, const T *val,
                        double sgn_rhs, double sgn_mv, hipStream_t s, const DirectWork<T> *pad = nullptr,
                        bool fused = false)
{
    const unsigned rows_per_block = kGemvThreads / 64 * kSweepRowsPerWave;
    int fgrid = 0, rpt = 0, e = 0;
    if constexpr (std::is_same<T, float>::value) {
        if (fused && pad && fused_plan(B, nj, nmax, P, fgrid, rpt, e)) {
            const dim3 grid(fgrid, nj, 1);
            if (rpt == 1) launch_fused_e<1>(e, grid, B, P, *pad, sgn_rhs, sgn_mv, s);
            else if (rpt == 2) launch_fused_e<2>(e, grid, B, P, *pad, sgn_rhs, sgn_mv, s);
            else launch_fused_e<3>(e, grid, B, P, *pad, sgn_rhs, sgn_mv, s);
            V2D_LAUNCH();
            return 0;
        }
    }
    if (pad)
        hipLaunchKernelGGL(k_level_rhs_pad<T>, dim3(grid_for(nmax), nj, 1), dim3(256), 0, s, B, P.L, P.N,
                           (const T *)pad->cfL, (const T *)pad->cfU, P.d_ixL, P.d_ixU, sgn_rhs);
    else
        hipLaunchKernelGGL(k_level_rhs_b<T>, dim3(grid_for(nmax), nj, 1), dim3(256), 0, s, B, P.L, P.N, val,
                           sgn_rhs);
    hipLaunchKernelGGL(k_level_gemv_b<T>, dim3((nmax + rows_per_block - 1) / rows_per_block, nj, 1),
                       dim3(kGemvThreads), sizeof(T) * P.ldmax, s, B, P.L, P.N, sgn_mv);
    V2D_LAUNCH();
    return 0;
}

// ------------------------------------------ persistent sweep kernel -----
// The whole solve (every forward step, the middle level, every backward step)
// in one launch of 2 C workgroups, C = min(half the compute units, 2 imax /
// 16): workgroups [0, C) run the wall chain (and the middle level), [C, 2C)
// the far-field chain, and workgroup q of a chain owns rows [q n / C,
// (q + 1) n / C) of its chain's level at every step.  Each workgroup publishes
// its progress in its own word (2 s + 1: its part of the level right-hand side
// of step s is stored; 2 s + 2: its rows of step s are stored; release), and
// waiting workgroups poll the words of the chain in parallel, so no counter is
// contended.  Per step:
//   1. every wavefront issues the loads of its S_l^-1 row, the threads those
//      of their coupling entries and right-hand side (registers);
//   Y. wait until the producing chain has stored step s-1 (the middle level
//      waits for both chains, the first backward step for the middle level);
//      the workgroup's rows of the level right-hand side (forward: b_l -
//      couplings . z; backward: couplings . w), eight lanes per row summing
//      in list order, go to a level-ordered buffer;
//   D. wait until the chain has stored all of y_l, load it into LDS, form the
//      workgroup's rows of S_l^-1 y (FP64 accumulation) and store them
//      (forward: z_l; middle and backward: w_l and the solution).
// Every wait is bounded (error word 2 on expiry; the host then solves again
// with per-level launches): the mapping needs all 2 C workgroups resident at
// once, one per compute unit, which an idle GPU gives.
template <typename T>
struct SweepArgs {
    const T *S;
    const T *cfL, *cfU;
    const int *ixL, *ixU;
    const T *b;
    T *z, *w, *x, *y;
    int *flag, *err;
    Levels L;
    int64_t N;
    int M, steps, nsteps, C, ys_len, rows_max;
    int spin_limit;   // bound of each wait (polls)
};

// Wait until flag[first .. first + count) >= target: the first waves poll one
// word per lane; bounded, error word 2 on expiry (every later wait then gives
// up at once, so a failed launch drains quickly).  A negative spin_limit
// expires every wait without polling (the host's way to force the fallback).
// The data behind the words is read with agent-coherent loads (cload), so no
// acquire fence follows.
__device__ __forceinline__ void sweep_wait(const int *flag, int first, int count, int target, int *err,
                                           int spin_limit)
{
    const int tid = threadIdx.x, wave = tid >> 6;
    if (wave * 64 < count && spin_limit < 0) {
        __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (wave * 64 < count) {
        int spins = 0;
        for (;;) {
            const int i = tid;
            const bool mine =
                i >= count || __hip_atomic_load(flag + first + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target;
            if (__all(mine)) break;
            if (++spins > spin_limit ||
                ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

// Agent-coherent store and load of the vectors the sweep workgroups exchange
// (level right-hand sides, the chains' z and w): with them the publishing
// side waits for its stores to complete (s_waitcnt vmcnt(0)) before its
// progress word, and neither side needs a fence, whose release writes back
// and whose acquire invalidates the whole L2 of the die.
template <typename T>
__device__ __forceinline__ void cstore(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T cload(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// An empty asm that "changes" a register-held vector: the compiler cannot
// move work on it (the FP64 conversion of a prefetched FP32 row) above here.
__device__ __forceinline__ void pin(float4 &v)
{
    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
}
__device__ __forceinline__ void pin(double2 &v)
{
    asm volatile("" : "+v"(v.x), "+v"(v.y));
}

template <typename T>
__global__ void __launch_bounds__(kSwThreads) k_sweep(SweepArgs<T> A)
{
    using V = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::w;
    extern __shared__ unsigned char smem[];
    T *ys = reinterpret_cast<T *>(smem);
    T *so = ys + A.ys_len;   // the workgroup's results of a step
    const int C = A.C;
    const int chain = blockIdx.x >= C, q = blockIdx.x - chain * C;
    const int nlev = A.L.nlev, M = A.M;
    for (int s = 0; s < A.nsteps; ++s) {
        // lane-dependent offsets are cheap to form per step; an empty asm
        // keeps the compiler from hoisting (and then spilling) them
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int wave = tid >> 6, lane = tid & 63;
        // this chain's level at step s (-1: none), its coupling terms, phase
        int l = -1, term = 0, phase = 0;
        if (s < A.steps) {
            if (chain == 0 && s < M) { l = s; term = s > 0 ? 1 : 0; }
            if (chain == 1 && s < nlev - 1 - M) { l = nlev - 1 - s; term = s > 0 ? 2 : 0; }
        } else if (s == A.steps) {
            phase = 1;
            if (chain == 0) { l = M; term = (M > 0 ? 1 : 0) | (M < nlev - 1 ? 2 : 0); }
        } else {
            phase = 2;
            const int st = s - A.steps - 1;
            if (chain == 0 && M - 1 - st >= 0) { l = M - 1 - st; term = 2; }
            if (chain == 1 && M + 1 + st < nlev) { l = M + 1 + st; term = 1; }
        }
        if (l < 0) {   // nothing for this chain: the step is done at once
            if (tid == 0) __hip_atomic_store(A.flag + blockIdx.x, 2 * s + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        const int n = A.L.size(l), ld = lvl_ld(n);
        const int r0 = (int)((int64_t)q * n / C), r1 = (int)((int64_t)(q + 1) * n / C);
        const int64_t vo = lvl_voff(A.L, l);
        const T *Sl = A.S + lvl_soff(A.L, l);
        const int row = r0 + wave;
        // 1. loads that do not depend on earlier steps (indices clamped into
        //    the level: every load in bounds, none branched)
        const V *Sr = reinterpret_cast<const V *>(Sl + (int64_t)min(row, n - 1) * ld);
        const int nv = (n + W - 1) / W;
        // (vectors past the row are clamped duplicates, left out of the sums:
        // zeroing them here would wait for the loads at once)
        V pre[kSweepPre];
#pragma unroll
        for (int u = 0; u < kSweepPre; ++u) pre[u] = load_row_nt<T>(Sr + min(lane + 64 * u, nv - 1));
        // an item of 17 rows (n = 16 C + k): its 17th row is prefetched one
        // 16-byte vector per lane by wavefronts 0..8 (u = wave)
        const bool x17 = r1 - r0 == kSwRows + 1;
        V pre17 = V{};
        if (x17 && wave < kSweepPre) {
            const V *R17 = reinterpret_cast<const V *>(Sl + (int64_t)(r0 + kSwRows) * ld);
            pre17 = load_row_nt<T>(R17 + min(lane + 64 * wave, nv - 1));
        }
        // The level right-hand side, all of it, formed by every workgroup
        // itself (nothing exchanged): rows tid + 1024 k, in batches of kSwRb
        // rows per thread.  The first batch's source indices do not depend on
        // earlier steps and are loaded before the wait; the coefficients, b
        // and the source entries (agent-coherent loads) after it, all issued
        // before the first sum.  Sums in list order: the first term's entries
        // 0..3, then the second term's (the middle level's U term).
        const bool has_t = term != 0;
        const T *cft = ((term & 1) ? A.cfL : A.cfU) + (int64_t)kCouple * vo;
        const int *ixt = ((term & 1) ? A.ixL : A.ixU) + (int64_t)kCouple * vo;
        int ixp[kSwRb][kCouple];
#pragma unroll
        for (int k = 0; k < kSwRb; ++k) {
            const int r = min(tid + kSwThreads * k, n - 1);
#pragma unroll
            for (int e = 0; e < kCouple; ++e) ixp[k][e] = ixt[e * n + r];
        }
        // Y. the producing chain has stored step s - 1
        if (s > 0) {
            const int first = (s == A.steps || s == A.steps + 1) ? 0 : chain * C;
            const int count = s == A.steps ? 2 * C : C;
            sweep_wait(A.flag, first, count, 2 * s, A.err, A.spin_limit);
        }
        const T *src = phase == 2 ? A.w : A.z;
        for (int kb = 0; kb * kSwRb * kSwThreads < n; ++kb) {
            T cv[kSwRb][kCouple], sv[kSwRb][kCouple], bv[kSwRb];
#pragma unroll
            for (int k = 0; k < kSwRb; ++k) {
                const int r = min(tid + kSwThreads * (k + kSwRb * kb), n - 1);
#pragma unroll
                for (int e = 0; e < kCouple; ++e) {
                    if (kb > 0) ixp[k][e] = ixt[e * n + r];
                    cv[k][e] = cft[e * n + r];
                    sv[k][e] = cload(src + ixp[k][e]);
                }
                bv[k] = A.b[A.L.global(r, A.N, l)];
            }
#pragma unroll
            for (int k = 0; k < kSwRb; ++k) {
                const int r = tid + kSwThreads * (k + kSwRb * kb);
                if (r >= n) break;
                double acc = has_t ? (double)cv[k][0] * (double)sv[k][0] : 0.0;
#pragma unroll
                for (int e = 1; e < kCouple; ++e)
                    if (has_t) acc += (double)cv[k][e] * (double)sv[k][e];
                if (term == 3)
#pragma unroll
                    for (int e = 0; e < kCouple; ++e) {
                        const int64_t o = (int64_t)kCouple * vo + (int64_t)e * n + r;
                        acc += (double)A.cfU[o] * (double)cload(src + A.ixU[o]);
                    }
                ys[r] = phase < 2 ? (T)((double)bv[k] - acc) : (T)acc;
            }
        }
        // zero the tail the dot products read (up to the prefetch width)
        for (int r = n + tid; r < max(nv, 64 * kSweepPre) * W; r += kSwThreads) ys[r] = T(0);
        __syncthreads();
        // this wavefront's rows of S_l^-1 y: the prefetched one, then any
        // further rows of a wide level streamed, with the same split into two
        // partial sums as the per-level kernel (prefetched entries past the
        // row are zero and meet zeros in ys)
        // the 17th row: every wavefront u < 9 forms its product of that row
        // (lane l: columns l + 64 u), wavefront 0 sums them below in the
        // order of the per-level kernel
        double *x17p = reinterpret_cast<double *>(so + A.rows_max);   // [9][64]
        if (x17) {
            if (wave < kSweepPre) {
                pin(pre17);
                const double d = dot16<T>(pre17, ys + (lane + 64 * wave) * W);
                x17p[wave * 64 + lane] = lane + 64 * wave < nv ? d : 0.0;
            }
            __syncthreads();
        }
        for (int r = row; r < r1; r += kSwRows) {
            const V *Rr = reinterpret_cast<const V *>(Sl + (int64_t)r * ld);
            double acc0 = 0.0, acc1 = 0.0;
            if (x17 && r == r0 + kSwRows) {
#pragma unroll
                for (int u = 0; u < kSweepPre; ++u) {
                    if (u & 1) acc1 += x17p[u * 64 + lane];
                    else acc0 += x17p[u * 64 + lane];
                }
            } else if (r == row) {
                // three groups, so that the LDS operands of all nine are not
                // held at once beside the prefetched row
#pragma unroll
                for (int u = 0; u < kSweepPre; ++u) {
                    pin(pre[u]);
                    if (lane + 64 * u < nv) {
                        if (u & 1) acc1 += dot16<T>(pre[u], ys + (lane + 64 * u) * W);
                        else acc0 += dot16<T>(pre[u], ys + (lane + 64 * u) * W);
                    }
                    if (u % 3 == 2) __builtin_amdgcn_sched_barrier(0);
                }
            } else {
                for (int u = 0; u < kSweepPre; ++u) {
                    const int cu = lane + 64 * u;
                    if (cu < nv) {
                        if (u & 1) acc1 += dot16<T>(Rr[cu], ys + cu * W);
                        else acc0 += dot16<T>(Rr[cu], ys + cu * W);
                    }
                }
            }
            for (int c = lane + 64 * kSweepPre; c < nv; c += 128) {
                acc0 += dot16<T>(Rr[c], ys + c * W);
                if (c + 64 < nv) acc1 += dot16<T>(Rr[c + 64], ys + (c + 64) * W);
            }
            const double acc = wsum(acc0 + acc1);
            if (lane == 0) so[r - r0] = phase == 2 ? (T)((double)cload(A.z + vo + r) - acc) : (T)acc;
        }
        __syncthreads();
        if (wave == 0) {
            for (int r = r0 + lane; r < r1; r += 64) {
                const T o = so[r - r0];
                if (phase == 0) {
                    cstore(A.z + vo + r, o);
                } else {
                    cstore(A.w + vo + r, o);
                    A.x[A.L.global(r, A.N, l)] = o;   // read after the launch only
                }
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the stores are complete
            if (lane == 0) __hip_atomic_store(A.flag + blockIdx.x, 2 * s + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();   // so[] and ys[] are rewritten by the next step
    }
}

// Coefficients of the padded coupling lists: out[q] = val[vx[q]] (0 on padding).
template <typename T>
__global__ void __launch_bounds__(256) k_gather_couplings(int64_t n, const int *__restrict__ vx,
                                                          const T *__restrict__ val, T *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int a = vx[q];
    out[q] = a >= 0 ? val[a] : T(0);
}

__global__ void k_zero_ints(int *p, int n)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) p[q] = 0;
}

template <typename T>
int direct_couplings(const DirectPattern &P, DirectWork<T> &W, const T *val, hipStream_t s)
{
    if (!P.pad_ok || !W.cfL || !W.cfU || !W.cfC) {
        set_error("block LU: padded coupling tables or their values are missing");
        return -1;
    }
    const int64_t n = (int64_t)kCouple * 2 * P.N;
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_gather_couplings<T>, dim3(g), dim3(256), 0, s, n, P.d_vxL, val, W.cfL);
    hipLaunchKernelGGL(k_gather_couplings<T>, dim3(g), dim3(256), 0, s, n, P.d_vxU, val, W.cfU);
    if (P.ncx)
        hipLaunchKernelGGL(k_gather_couplings<T>, dim3((unsigned)((P.ncx + 255) / 256)), dim3(256), 0, s,
                           P.ncx, P.d_vxC, val, W.cfC);
    V2D_LAUNCH();
    return 0;
}

// dynamic LDS of k_sweep: the level right-hand side, at least as wide as the
// prefetched part of a row
template <typename T>
static int sweep_ys(const DirectPattern &P)
{
    return std::max(P.ldmax, 64 * kSweepPre * Vec16<T>::w);
}
// ys, the results of the item (rows rounded up to an even count so that the
// FP64 scratch of a 17th row after them is 8-byte aligned), that scratch
template <typename T>
static size_t sweep_lds(const DirectPattern &P)
{
    return sizeof(T) * ((size_t)sweep_ys<T>(P) + (P.sweep_rows + 1) / 2 * 2) + sizeof(double) * 64 * kSweepPre;
}

template <typename T>
static int direct_solve_levels(const DirectPattern &P, const DirectWork<T> &W, const T *val, const T *b, T *x,
                               hipStream_t s);

// One persistent launch; if one of its bounded waits expired (the workgroups
// were not all resident at once) the solve is redone with per-level launches
// and the work object keeps to them.
template <typename T>
static int direct_solve_persistent(const DirectPattern &P, DirectWork<T> &W, const T *val, const T *b, T *x,
                                   hipStream_t s)
{
    const int nlev = P.L.nlev;
    SweepArgs<T> A{};
    A.S = W.sinv;
    A.cfL = W.cfL; A.cfU = W.cfU;
    A.ixL = P.d_ixL; A.ixU = P.d_ixU;
    A.b = b; A.z = W.z; A.w = W.wv; A.x = x; A.y = W.yv;
    A.flag = W.ctr; A.err = W.err;
    A.L = P.L; A.N = P.N;
    A.M = (nlev - 1) / 2;
    A.steps = std::max(A.M, nlev - 1 - A.M);
    A.nsteps = P.nsteps;
    A.C = P.sweep_C;
    A.ys_len = sweep_ys<T>(P);
    A.rows_max = (P.sweep_rows + 1) / 2 * 2;
    A.spin_limit = W.sweep_spins > 0 ? W.sweep_spins : -1;
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(1024), 0, s, W.ctr, P.sweep_grid);
    hipLaunchKernelGGL(k_sweep<T>, dim3(P.sweep_grid), dim3(kSwThreads), sweep_lds<T>(P), s, A);
    V2D_LAUNCH();
    V2D_HIP(hipMemcpyAsync(W.err_host, W.err, sizeof(int), hipMemcpyDeviceToHost, s));
    V2D_HIP(hipStreamSynchronize(s));
    if (*W.err_host == 2) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, W.err, 0);
        V2D_LAUNCH();
        *W.err_host = 0;
        W.sweep_kind = 0;
        ++W.sweep_fallbacks;
        return direct_solve_levels(P, W, val, b, x, s);
    }
    return 0;
}

// Two sweeps over the twisted factorisation:
//   forward  bottom: z_l = S_l^-1 (b_l - L_l z_{l-1}),  top: z_l = S_l^-1 (b_l - U_l z_{l+1})
//   middle        : x_M = S_M^-1 (b_M - L_M z_{M-1} - U_M z_{M+1})
//   backward bottom: x_l = z_l - S_l^-1 U_l x_{l+1},   top: x_l = z_l - S_l^-1 L_l x_{l-1}
// sweep_kind 1 runs them as one persistent launch (k_sweep), 0 as one launch
// pair per level step.
template <typename T>
int direct_solve(const DirectPattern &P, DirectWork<T> &W, const T *val, const T *b, T *x, hipStream_t s)
{
    const bool lds_fits = sweep_lds<T>(P) <= 65536;
    if (W.sweep_kind == 1 && P.sweep_ok && lds_fits && W.cfL && W.wv && W.yv && W.ctr && W.err_host)
        return direct_solve_persistent(P, W, val, b, x, s);
    return direct_solve_levels(P, W, val, b, x, s);
}

template <typename T>
static int direct_solve_levels(const DirectPattern &P, const DirectWork<T> &W, const T *val, const T *b, T *x,
                               hipStream_t s)
{
    const int nlev = P.L.nlev;
    if (sizeof(T) * (size_t)P.ldmax > 65536) {
        set_error("direct solver: a level of " + std::to_string(P.nmax) +
                  " unknowns exceeds the 64 KiB LDS staging of the solve sweeps");
        return -1;
    }
    const int M = (nlev - 1) / 2;
    const int nbot = M, ntop = nlev - 1 - M;
    const int steps = std::max(nbot, ntop);
    auto zl = [&](int l) { return W.z + P.voff[l]; };
    auto S = [&](int l) { return (const T *)(W.sinv + P.soff[l]); };
    T *ybuf[2] = {W.colk, W.colk + (int64_t)P.nmax * NB};   // the panels are free during a solve
    // right-hand sides from the padded coupling lists when the factorisation
    // filled them (the backward steps then also keep a level-ordered copy of
    // the solution, in wv, for their sources)
    const DirectWork<T> *pad = (P.pad_ok && W.cfL && W.wv) ? &W : nullptr;
    const bool fused = W.sweep_kind == 2;
    // forward sweeps
    for (int st = 0; st < steps; ++st) {
        SweepBatch<T> G{};
        int nj = 0, nmax = 0;
        for (int side = 0; side < 2; ++side) {
            if ((side == 0 && st >= nbot) || (side == 1 && st >= ntop)) continue;
            const int l = side == 0 ? st : nlev - 1 - st;
            RhsJob<T> &r = G.j[nj].rhs;
            r.l = l; r.n = P.n[l]; r.b = b; r.nt = 0;
            if (st > 0) {
                const int p = side == 0 ? l - 1 : l + 1;
                r.t[0] = RhsTerm<T>{side == 0 ? P.d_Lptr + P.voff[l] : P.d_Uptr + P.voff[l],
                                    side == 0 ? P.d_Lr : P.d_Ur, zl(p), p, 0};
                r.nt = 1;
            }
            G.j[nj].y = ybuf[nj];
            G.j[nj].S = S(l); G.j[nj].ld = P.ld[l]; G.j[nj].glob = 0;
            G.j[nj].base = nullptr; G.j[nj].out = zl(l);
            G.j[nj].pterm = st > 0 ? (side == 0 ? 1 : 2) : 0; G.j[nj].psrc = W.z;
            nmax = std::max(nmax, P.n[l]);
            ++nj;
        }
        if (launch_sweep(G, nj, nmax, P, val, -1.0, 1.0, s, pad, fused)) return -1;
    }
    {   // middle level
        SweepBatch<T> G{};
        RhsJob<T> &r = G.j[0].rhs;
        r.l = M; r.n = P.n[M]; r.b = b; r.nt = 0;
        if (M > 0)
            r.t[r.nt++] = RhsTerm<T>{P.d_Lptr + P.voff[M], P.d_Lr, zl(M - 1), M - 1, 0};
        if (M < nlev - 1)
            r.t[r.nt++] = RhsTerm<T>{P.d_Uptr + P.voff[M], P.d_Ur, zl(M + 1), M + 1, 0};
        G.j[0].y = ybuf[0];
        G.j[0].S = S(M); G.j[0].ld = P.ld[M]; G.j[0].glob = 1;
        G.j[0].base = nullptr; G.j[0].out = x;
        G.j[0].pterm = (M > 0 ? 1 : 0) | (M < nlev - 1 ? 2 : 0); G.j[0].psrc = W.z;
        G.j[0].lout = pad ? W.wv + P.voff[M] : nullptr;
        if (launch_sweep(G, 1, P.n[M], P, val, -1.0, 1.0, s, pad)) return -1;
    }
    // backward sweeps, outward from the middle
    for (int st = 0; st < steps; ++st) {
        SweepBatch<T> G{};
        int nj = 0, nmax = 0;
        for (int side = 0; side < 2; ++side) {
            const int l = side == 0 ? M - 1 - st : M + 1 + st;
            if (l < 0 || l >= nlev) continue;
            const int p = side == 0 ? l + 1 : l - 1;
            RhsJob<T> &r = G.j[nj].rhs;
            r.l = l; r.n = P.n[l]; r.b = nullptr; r.nt = 1;
            r.t[0] = RhsTerm<T>{side == 0 ? P.d_Uptr + P.voff[l] : P.d_Lptr + P.voff[l],
                                side == 0 ? P.d_Ur : P.d_Lr, x, p, 1};
            G.j[nj].y = ybuf[nj];
            G.j[nj].S = S(l); G.j[nj].ld = P.ld[l]; G.j[nj].glob = 1;
            G.j[nj].base = zl(l); G.j[nj].out = x;
            G.j[nj].pterm = side == 0 ? 2 : 1; G.j[nj].psrc = W.wv;
            G.j[nj].lout = pad ? W.wv + P.voff[l] : nullptr;
            nmax = std::max(nmax, P.n[l]);
            ++nj;
        }
        if (!nj) break;
        if (launch_sweep(G, nj, nmax, P, val, 1.0, -1.0, s, pad, fused)) return -1;
    }
    return 0;
}
// Transposed GEMV of one solve step: out[c] = [base[c]] + sgn_mv * sum_r S[r][c] y[r].
// One workgroup per 64-column strip: lane = column, the 16 wavefronts split
// the rows (each reads 64 contiguous entries of a row), LDS reduction.
template <typename T>
__global__ void __launch_bounds__(1024) k_level_gemv_t_b(SweepBatch<T> B, Levels L, int64_t N,
                                                         double sgn_mv)
{
    __shared__ double part[16][64];
    const SweepJob<T> &J = B.j[blockIdx.y];
    const int n = J.rhs.n, ld = J.ld;
    const int c0 = blockIdx.x * 64;
    if (c0 >= n) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = c0 + lane;
    double acc = 0.0;
    if (c < n)
        for (int r = wave; r < n; r += 16) acc += (double)J.S[(int64_t)r * ld + c] * (double)J.y[r];
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && c < n) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) t += part[q][lane];
        const double b = J.base ? (double)J.base[c] : 0.0;
        const T o = (T)(b + sgn_mv * t);
        if (J.glob) J.out[L.global(c, N, J.rhs.l)] = o;
        else J.out[c] = o;
    }
}

template <typename T>
static int launch_sweep_t(const SweepBatch<T> &B, int nj, int nmax, const DirectPattern &P, const T *val,
                          double sgn_rhs, double sgn_mv, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_rhs_b<T>, dim3(grid_for(nmax), nj, 1), dim3(256), 0, s, B, P.L, P.N, val,
                       sgn_rhs);
    hipLaunchKernelGGL(k_level_gemv_t_b<T>, dim3((nmax + 63) / 64, nj, 1), dim3(1024), 0, s, B, P.L, P.N,
                       sgn_mv);
    V2D_LAUNCH();
    return 0;
}

// Solve A^T x = b with the stored twisted factorisation.  The Schur complements
// of A^T in the same elimination order are the transposes of A's, so the
// sweeps reuse the stored inverses (applied transposed) and the couplings by
// columns: forward bottom  z_l = S_l^-T (b_l - U_{l-1}^T z_{l-1}),
// forward top  z_l = S_l^-T (b_l - L_{l+1}^T z_{l+1}), the middle level both,
// backward bottom  x_l = z_l - S_l^-T L_{l+1}^T x_{l+1},
// backward top  x_l = z_l - S_l^-T U_{l-1}^T x_{l-1}.
template <typename T>
int direct_solve_t(const DirectPattern &P, const DirectWork<T> &W, const T *val, const T *b, T *x,
                   hipStream_t s)
{
    const int nlev = P.L.nlev;
    const int M = (nlev - 1) / 2;
    const int nbot = M, ntop = nlev - 1 - M;
    const int steps = std::max(nbot, ntop);
    auto zl = [&](int l) { return W.z + P.voff[l]; };
    auto S = [&](int l) { return (const T *)(W.sinv + P.soff[l]); };
    T *ybuf[2] = {W.colk, W.colk + (int64_t)P.nmax * NB};
    auto UC = [&](int l, const T *vec, int vlev, int glob) {
        return RhsTerm<T>{P.d_UCptr + P.voff[l], P.d_Uc, vec, vlev, glob};
    };
    auto LC = [&](int l, const T *vec, int vlev, int glob) {
        return RhsTerm<T>{P.d_LCptr + P.voff[l], P.d_Lc, vec, vlev, glob};
    };
    for (int st = 0; st < steps; ++st) {
        SweepBatch<T> G{};
        int nj = 0, nmax = 0;
        for (int side = 0; side < 2; ++side) {
            if ((side == 0 && st >= nbot) || (side == 1 && st >= ntop)) continue;
            const int l = side == 0 ? st : nlev - 1 - st;
            RhsJob<T> &r = G.j[nj].rhs;
            r.l = l; r.n = P.n[l]; r.b = b; r.nt = 0;
            if (st > 0) {
                const int p = side == 0 ? l - 1 : l + 1;
                r.t[0] = side == 0 ? UC(l, zl(p), p, 0) : LC(l, zl(p), p, 0);
                r.nt = 1;
            }
            G.j[nj].y = ybuf[nj];
            G.j[nj].S = S(l); G.j[nj].ld = P.ld[l]; G.j[nj].glob = 0;
            G.j[nj].base = nullptr; G.j[nj].out = zl(l);
            nmax = std::max(nmax, P.n[l]);
            ++nj;
        }
        if (launch_sweep_t(G, nj, nmax, P, val, -1.0, 1.0, s)) return -1;
    }
    {
        SweepBatch<T> G{};
        RhsJob<T> &r = G.j[0].rhs;
        r.l = M; r.n = P.n[M]; r.b = b; r.nt = 0;
        if (M > 0) r.t[r.nt++] = UC(M, zl(M - 1), M - 1, 0);
        if (M < nlev - 1) r.t[r.nt++] = LC(M, zl(M + 1), M + 1, 0);
        G.j[0].y = ybuf[0];
        G.j[0].S = S(M); G.j[0].ld = P.ld[M]; G.j[0].glob = 1;
        G.j[0].base = nullptr; G.j[0].out = x;
        if (launch_sweep_t(G, 1, P.n[M], P, val, -1.0, 1.0, s)) return -1;
    }
    for (int st = 0; st < steps; ++st) {
        SweepBatch<T> G{};
        int nj = 0, nmax = 0;
        for (int side = 0; side < 2; ++side) {
            const int l = side == 0 ? M - 1 - st : M + 1 + st;
            if (l < 0 || l >= nlev) continue;
            const int p = side == 0 ? l + 1 : l - 1;
            RhsJob<T> &r = G.j[nj].rhs;
            r.l = l; r.n = P.n[l]; r.b = nullptr; r.nt = 1;
            r.t[0] = side == 0 ? LC(l, x, p, 1) : UC(l, x, p, 1);
            G.j[nj].y = ybuf[nj];
            G.j[nj].S = S(l); G.j[nj].ld = P.ld[l]; G.j[nj].glob = 1;
            G.j[nj].base = zl(l); G.j[nj].out = x;
            nmax = std::max(nmax, P.n[l]);
            ++nj;
        }
        if (!nj) break;
        if (launch_sweep_t(G, nj, nmax, P, val, 1.0, -1.0, s)) return -1;
    }
    return 0;
}
template int direct_solve_t<double>(const DirectPattern &, const DirectWork<double> &, const double *,
                                    const double *, double *, hipStream_t);
template int direct_solve_t<float>(const DirectPattern &, const DirectWork<float> &, const float *,
                                   const float *, float *, hipStream_t);

template int direct_solve<double>(const DirectPattern &, DirectWork<double> &, const double *,
                                  const double *, double *, hipStream_t);
template int direct_solve<float>(const DirectPattern &, DirectWork<float> &, const float *,
                                 const float *, float *, hipStream_t);

}  // namespace v2d
