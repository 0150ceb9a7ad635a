// v2d_direct.hip -- exact block-tridiagonal LU of the Newton matrix on gfx950.
//
// Ordering the unknowns by radial level makes the Newton matrix block
// tridiagonal: level l (one j row, or two at the wall / far field where the
// Thom and outflow rows reach j +- 2) couples only to levels l-1 and l+1.  The
// factorisation is block Gaussian elimination from the wall outward:
//     S_0 = D_0,   S_l = D_l - L_l S_{l-1}^{-1} U_{l-1},
// and keeps every S_l^{-1} explicitly, computed in place by blocked
// Gauss-Jordan without pivoting.  L and U stay sparse (at most a handful of
// entries per row / column), so the Schur update is O(n^2) per level and the
// dense inversion (2 n^3 flops, n = 2*imax) carries the work.  A solve is two
// sweeps of one dense GEMV per level:
//     forward  z_l = S_l^{-1} (b_l - L_l z_{l-1})
//     backward x_l = z_l - S_l^{-1} U_l x_{l+1}.
// With T = double this is the exact Newton step of the reference's MUMPS LU
// (vorti2d/petsc_solver.py:55-62); with T = float it is the mixed-precision
// preconditioner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "v2d_internal.h"

namespace v2d {

constexpr int NB = 64;   // Gauss-Jordan block width == update tile edge

// ------------------------------------------------------------ pattern -----
int direct_build_pattern(DirectPattern &P, int imax, int jmax, const uint8_t *flags)
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
    std::vector<int> cD(n2 + 1, 0), cL(n2 + 1, 0), cU(n2 + 1, 0), cUC(n2 + 1, 0);
    // two passes: count, then fill
    std::vector<int2> eD, eL, eU, eUC;
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int> fD, fL, fU, fUC;
        if (pass == 1) {
            for (int64_t r = 0; r < n2; ++r) {
                cD[r + 1] += cD[r]; cL[r + 1] += cL[r]; cU[r + 1] += cU[r]; cUC[r + 1] += cUC[r];
            }
            eD.resize(cD[n2]); eL.resize(cL[n2]); eU.resize(cU[n2]); eUC.resize(cUC[n2]);
            fD.assign(cD.begin(), cD.end() - 1); fL.assign(cL.begin(), cL.end() - 1);
            fU.assign(cU.begin(), cU.end() - 1); fUC.assign(cUC.begin(), cUC.end() - 1);
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
                    if (pass == 0) ++cL[R + 1]; else eL[fL[R]++] = make_int2(c, vidx);
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
    auto up = [](int **dptr, int2 **dent, const std::vector<int> &ptr,
                 const std::vector<int2> &ent) -> int {
        V2D_HIP(hipMalloc(dptr, sizeof(int) * ptr.size()));
        V2D_HIP(hipMemcpy(*dptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
        V2D_HIP(hipMalloc(dent, sizeof(int2) * std::max<size_t>(ent.size(), 1)));
        if (!ent.empty())
            V2D_HIP(hipMemcpy(*dent, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice));
        return 0;
    };
    if (up(&P.d_Dptr, &P.d_D, cD, eD) || up(&P.d_Lptr, &P.d_Lr, cL, eL) ||
        up(&P.d_Uptr, &P.d_Ur, cU, eU) || up(&P.d_UCptr, &P.d_Uc, cUC, eUC))
        return -1;
    return 0;
}

void direct_free_pattern(DirectPattern &P)
{
    int *ptrs[] = {P.d_Dptr, P.d_Lptr, P.d_Uptr, P.d_UCptr};
    int2 *ents[] = {P.d_D, P.d_Lr, P.d_Ur, P.d_Uc};
    for (int *p : ptrs) if (p) (void)hipFree(p);
    for (int2 *p : ents) if (p) (void)hipFree(p);
    P.d_Dptr = P.d_Lptr = P.d_Uptr = P.d_UCptr = nullptr;
    P.d_D = P.d_Lr = P.d_Ur = P.d_Uc = nullptr;
}

// ------------------------------------------------------- level build ------
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *__restrict__ p, int64_t n, T v)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// S[r][c] = D entries of level rows
template <typename T>
__global__ void __launch_bounds__(256) k_scatter_D(int n, int ld, const int *__restrict__ ptr,
                                                   const int2 *__restrict__ ent,
                                                   const T *__restrict__ val, T *__restrict__ S)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int e = ptr[r]; e < ptr[r + 1]; ++e) {
        const int2 x = ent[e];
        S[(int64_t)r * ld + x.x] = val[x.y];
    }
}

// Tm[c][q] = sum_{(p,u) in Ucol(q)} Sprev[c][p] * u      (np x nl)
template <typename T>
__global__ void __launch_bounds__(256) k_schur_T(int np, int ldp, int nl, int ldl,
                                                 const int *__restrict__ ucptr,
                                                 const int2 *__restrict__ uc,
                                                 const T *__restrict__ val,
                                                 const T *__restrict__ Sprev, T *__restrict__ Tm)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (q >= nl) return;
    double acc = 0.0;
    const T *row = Sprev + (int64_t)c * ldp;
    for (int e = ucptr[q]; e < ucptr[q + 1]; ++e) {
        const int2 x = uc[e];
        acc += (double)row[x.x] * (double)val[x.y];
    }
    Tm[(int64_t)c * ldl + q] = (T)acc;
}

// S[r][q] -= sum_{(c,v) in Lrow(r)} v * Tm[c][q]
template <typename T>
__global__ void __launch_bounds__(256) k_schur_S(int nl, int ldl, const int *__restrict__ lptr,
                                                 const int2 *__restrict__ lr,
                                                 const T *__restrict__ val,
                                                 const T *__restrict__ Tm, T *__restrict__ S)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (q >= nl) return;
    const int e0 = lptr[r], e1 = lptr[r + 1];
    if (e0 == e1) return;
    double acc = (double)S[(int64_t)r * ldl + q];
    for (int e = e0; e < e1; ++e) {
        const int2 x = lr[e];
        acc -= (double)val[x.y] * (double)Tm[(int64_t)x.x * ldl + q];
    }
    S[(int64_t)r * ldl + q] = (T)acc;
}

// --------------------------------------------------- Gauss-Jordan ---------
// Inverse of a kb x kb (kb <= 64) block, unpivoted Gauss-Jordan, held in
// registers: thread t owns column c = t % 64 of the 16 consecutive rows
// r = 16*(t/64) + q.  Per pivot p the owners publish the pivot row and column
// to LDS (double-buffered: one barrier per pivot) and every entry takes one FMA
//     a[r][c] -= a[r][p] * (a[p][c] / a[p][p]).
// The pivot column is left unscaled during the sweep (later pivots act on it
// linearly) and multiplied by -1/a[p][p] once at the end, with its diagonal
// entry set to -1 so that it ends as 1/a[p][p]; only the pivot-row entry of a
// thread needs a select.  Entries outside kb x kb must be zero on entry; the
// result is written to out (row stride NB).
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);   // ~2^-26, two Newton steps to full precision
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

template <typename T>
__device__ __forceinline__ void gj_regs(T (&a)[16], int kb, T *prow, T *pcol, T *scale, T *out)
{
    const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
    if (rg == 0) scale[c] = T(0);
    for (int pw = 0; pw < 4; ++pw) {
#pragma unroll
        for (int pq = 0; pq < 16; ++pq) {
            const int p = pw * 16 + pq;
            if (p < kb) {
                T *pr = prow + (pq & 1) * NB, *pc = pcol + (pq & 1) * NB;
                if (rg == pw) pr[c] = a[pq];
                if (c == p)
#pragma unroll
                    for (int q = 0; q < 16; ++q) pc[rg * 16 + q] = a[q];
                __syncthreads();
                const T ip = recip(pr[p]);
                const T rp = (c == p) ? T(0) : pr[c] * ip;
                if (c == p && rg == 0) scale[p] = -ip;
                T cr[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) cr[q] = pc[rg * 16 + q];
#pragma unroll
                for (int q = 0; q < 16; ++q) a[q] = fma(-cr[q], rp, a[q]);
                const T rowv = (c == p) ? T(-1) : rp;
                a[pq] = (rg == pw) ? rowv : a[pq];
            }
        }
    }
    __syncthreads();
    const T sc = scale[c];
#pragma unroll
    for (int q = 0; q < 16; ++q) out[(rg * 16 + q) * NB + c] = (c < kb) ? a[q] * sc : T(0);
}

template <typename T>
__global__ void __launch_bounds__(256) k_gj_diag(const T *__restrict__ A, int n, int ld, int k0,
                                                 T *__restrict__ dinv)
{
    __shared__ T prow[2 * NB], pcol[2 * NB], scale[NB];
    const int kb = min(NB, n - k0);
    const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
    T a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = rg * 16 + q;
        a[q] = (r < kb && c < kb) ? A[(int64_t)(k0 + r) * ld + k0 + c] : T(0);
    }
    gj_regs(a, kb, prow, pcol, scale, dinv);
}

// blockIdx.y == 0: pivot rows  A[K, c] <- Dinv A[K, c]  (A[K, K] <- Dinv) for a
//                  16-column tile (short per-block work: this is on the
//                  critical path of every Gauss-Jordan step)
// blockIdx.y == 1: colk[r][t] <- A[r][k0+t] for a 16-row tile of rows outside K
constexpr int RP = 16;
template <typename T>
__global__ void __launch_bounds__(256) k_gj_rowpanel(T *__restrict__ A, int n, int ld, int k0,
                                                     const T *__restrict__ dinv,
                                                     T *__restrict__ colk)
{
    __shared__ T sdT[NB * NB];   // sdT[s][t] = Dinv[t][s]
    __shared__ T sp[NB * RP];    // sp[s][c]  = A[k0+s][base+c]
    const int kb = min(NB, n - k0);
    const int base = blockIdx.x * RP;
    if (blockIdx.y == 1) {
        for (int idx = threadIdx.x; idx < RP * NB; idx += blockDim.x) {
            const int r = base + idx / NB, t = idx % NB;
            if (r < n && t < kb && (r < k0 || r >= k0 + kb))
                colk[(int64_t)r * NB + t] = A[(int64_t)r * ld + k0 + t];
        }
        return;
    }
    const bool cin = (base >= k0 && base < k0 + NB);
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        const int t = idx / NB, s2 = idx % NB;
        sdT[s2 * NB + t] = dinv[idx];
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
}

// A[i][c] = (c in K ? 0 : A[i][c]) - sum_t colk[i][t] A[k0+t][c]   for i outside K.
// The workgroup owning the next pivot tile also inverts it into dinv.
template <typename T>
__global__ void __launch_bounds__(256) k_gj_update(T *__restrict__ A, int n, int ld, int k0,
                                                   const T *__restrict__ colk,
                                                   T *__restrict__ dinv)
{
    __shared__ T sc[NB * NB];   // sc[t][r]: colk tile, transposed
    __shared__ T sr[NB * NB];   // sr[t][c]: pivot-row tile
    const int kb = min(NB, n - k0);
    // Tile order: the next pivot tile (k0+NB, k0+NB) is block 0, so its update
    // and inversion start first and overlap the bulk of the rank-NB update.
    const int nt = (n + NB - 1) / NB;
    const int kt1 = k0 / NB + 1;
    int lin = blockIdx.y * nt + blockIdx.x;
    if (kt1 < nt) {
        const int piv = kt1 * nt + kt1;
        if (lin == 0) lin = piv;
        else if (lin <= piv) lin -= 1;
    }
    const int i0 = (lin / nt) * NB, c0 = (lin % nt) * NB;
    if (i0 == k0) return;
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        const int r = idx / NB, t = idx % NB;
        sc[t * NB + r] = (i0 + r < n && t < kb) ? colk[(int64_t)(i0 + r) * NB + t] : T(0);
        sr[idx] = (r < kb && c0 + t < n) ? A[(int64_t)(k0 + r) * ld + c0 + t] : T(0);
    }
    __syncthreads();
    const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
    T acc[4][4];
    const bool cin = (c0 == k0);   // column tile == pivot columns
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + tr * 4 + a, c = c0 + tc * 4 + b;
            acc[a][b] = (!cin && i < n && c < n) ? A[(int64_t)i * ld + c] : T(0);
        }
    for (int t = 0; t < kb; ++t) {
        T av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = sc[t * NB + tr * 4 + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = sr[t * NB + tc * 4 + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(-av[a], bv[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + tr * 4 + a, c = c0 + tc * 4 + b;
            if (i < n && c < n) A[(int64_t)i * ld + c] = acc[a][b];
        }
    // fused inversion of the next pivot block
    const int k1 = k0 + NB;
    if (k1 < n && i0 == k1 && c0 == k1) {
        const int kb2 = min(NB, n - k1);
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int r = tr * 4 + a, c = tc * 4 + b;
                sc[r * NB + c] = (r < kb2 && c < kb2) ? acc[a][b] : T(0);
            }
        __syncthreads();
        const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
        T g[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) g[q] = sc[(rg * 16 + q) * NB + c];
        gj_regs(g, kb2, sr, sr + 2 * NB, sr + 4 * NB, dinv);
    }
}

template <typename T>
int gj_invert(T *A, int n, int ld, T *colk, T *dinv, hipStream_t s)
{
    const unsigned nt = (unsigned)((n + NB - 1) / NB);
    const unsigned np = (unsigned)((n + RP - 1) / RP);
    hipLaunchKernelGGL(k_gj_diag<T>, dim3(1), dim3(256), 0, s, A, n, ld, 0, dinv);
    V2D_LAUNCH();
    for (int k0 = 0; k0 < n; k0 += NB) {
        hipLaunchKernelGGL(k_gj_rowpanel<T>, dim3(np, 2), dim3(256), 0, s, A, n, ld, k0, dinv, colk);
        hipLaunchKernelGGL(k_gj_update<T>, dim3(nt, nt), dim3(256), 0, s, A, n, ld, k0, colk, dinv);
    }
    V2D_LAUNCH();
    return 0;
}
template int gj_invert<double>(double *, int, int, double *, double *, hipStream_t);
template int gj_invert<float>(float *, int, int, float *, float *, hipStream_t);

template <typename T>
int direct_factor(const DirectPattern &P, DirectWork<T> &W, const T *val, hipStream_t s)
{
    const int nlev = P.L.nlev;
    for (int l = 0; l < nlev; ++l) {
        const int n = P.n[l], ld = P.ld[l];
        T *S = W.sinv + P.soff[l];
        const int64_t nn = (int64_t)n * ld;
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(nn)), dim3(256), 0, s, S, nn, T(0));
        hipLaunchKernelGGL(k_scatter_D<T>, dim3(grid_for(n)), dim3(256), 0, s, n, ld,
                           P.d_Dptr + P.voff[l], P.d_D, val, S);
        if (l > 0) {
            const int np = P.n[l - 1], ldp = P.ld[l - 1];
            hipLaunchKernelGGL(k_schur_T<T>, dim3(grid_for(n), np), dim3(256), 0, s, np, ldp, n, ld,
                               P.d_UCptr + P.voff[l], P.d_Uc, val, W.sinv + P.soff[l - 1], W.tmat);
            hipLaunchKernelGGL(k_schur_S<T>, dim3(grid_for(n), n), dim3(256), 0, s, n, ld,
                               P.d_Lptr + P.voff[l], P.d_Lr, val, W.tmat, S);
        }
        V2D_LAUNCH();
        if (gj_invert<T>(S, n, ld, W.colk, W.dinv, s)) return -1;
    }
    return 0;
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

constexpr int kRowsPerWave = 2;
constexpr int kRowsPerBlock = 4 * kRowsPerWave;

// forward right-hand side of level l:  y = b_l - L_l z_{l-1}
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_rhs(Levels L, int64_t N, int l, int n,
                                                 const int *__restrict__ lptr,
                                                 const int2 *__restrict__ lr,
                                                 const T *__restrict__ val, const T *__restrict__ b,
                                                 const T *__restrict__ zprev, T *__restrict__ y)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double acc = (double)b[L.global(r, N, l)];
    if (l > 0)
        for (int e = lptr[r]; e < lptr[r + 1]; ++e) {
            const int2 x = lr[e];
            acc -= (double)val[x.y] * (double)zprev[x.x];
        }
    y[r] = (T)acc;
}

// backward right-hand side of level l:  w = U_l x_{l+1}
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_rhs(Levels L, int64_t N, int l, int n,
                                                 const int *__restrict__ uptr,
                                                 const int2 *__restrict__ ur,
                                                 const T *__restrict__ val, const T *__restrict__ x,
                                                 T *__restrict__ w)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double acc = 0.0;
    for (int e = uptr[r]; e < uptr[r + 1]; ++e) {
        const int2 u = ur[e];
        acc += (double)val[u.y] * (double)x[L.global(u.x, N, l + 1)];
    }
    w[r] = (T)acc;
}

// out = base - S v  (mode 1: scatter to the global vector, base = z_l)
//        or  S v    (mode 0: level-ordered z_l)
// One wavefront per row; rows are padded to ld (a multiple of 64 elements), so
// every lane reads 16-byte vectors and the padding (zeros in S) is skipped by
// the bound on the vector length.
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

template <typename T>
__global__ void __launch_bounds__(256) k_level_gemv(Levels L, int64_t N, int l, int n, int ld,
                                                    int mode, const T *__restrict__ S,
                                                    const T *__restrict__ v,
                                                    const T *__restrict__ base,
                                                    T *__restrict__ out)
{
    using V = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::w;
    extern __shared__ unsigned char smem[];
    T *vs = reinterpret_cast<T *>(smem);
    for (int r = threadIdx.x; r < ld; r += blockDim.x) vs[r] = r < n ? v[r] : T(0);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    if (r >= n) return;
    const V *row = reinterpret_cast<const V *>(S + (int64_t)r * ld);
    const int nv = (n + W - 1) / W;
    double acc0 = 0.0, acc1 = 0.0;
    int c = lane;
    for (; c + 64 < nv; c += 128) {
        const V x0 = row[c], x1 = row[c + 64];
        acc0 += dot16<T>(x0, vs + c * W);
        acc1 += dot16<T>(x1, vs + (c + 64) * W);
    }
    if (c < nv) acc0 += dot16<T>(row[c], vs + c * W);
    const double acc = wsum(acc0 + acc1);
    if (lane == 0) {
        if (mode == 0) out[r] = (T)acc;
        else out[L.global(r, N, l)] = (T)((double)base[r] - acc);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bwd_last(Levels L, int64_t N, int l, int n,
                                                  const T *__restrict__ zl, T *__restrict__ x)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) x[L.global(r, N, l)] = zl[r];
}

template <typename T>
int direct_solve(const DirectPattern &P, const DirectWork<T> &W, const T *val, const T *b, T *x,
                 hipStream_t s)
{
    const int nlev = P.L.nlev;
    const size_t lds = sizeof(T) * (size_t)P.ldmax;
    if (lds > 65536) {
        set_error("direct solver: a level of " + std::to_string(P.nmax) +
                  " unknowns exceeds the 64 KiB LDS staging of the solve sweeps");
        return -1;
    }
    T *y = W.colk;   // level right-hand side scratch (nmax <= nmax*64)
    for (int l = 0; l < nlev; ++l) {
        const int n = P.n[l];
        const T *zprev = l > 0 ? W.z + P.voff[l - 1] : W.z;
        hipLaunchKernelGGL(k_fwd_rhs<T>, dim3(grid_for(n)), dim3(256), 0, s, P.L, P.N, l, n,
                           P.d_Lptr + P.voff[l], P.d_Lr, val, b, zprev, y);
        hipLaunchKernelGGL(k_level_gemv<T>, dim3((n + 3) / 4), dim3(256), sizeof(T) * P.ld[l], s,
                           P.L, P.N, l, n, P.ld[l], 0, W.sinv + P.soff[l], y, (const T *)nullptr,
                           W.z + P.voff[l]);
    }
    V2D_LAUNCH();
    {
        const int l = nlev - 1, n = P.n[l];
        hipLaunchKernelGGL(k_bwd_last<T>, dim3(grid_for(n)), dim3(256), 0, s, P.L, P.N, l, n,
                           W.z + P.voff[l], x);
    }
    for (int l = nlev - 2; l >= 0; --l) {
        const int n = P.n[l];
        hipLaunchKernelGGL(k_bwd_rhs<T>, dim3(grid_for(n)), dim3(256), 0, s, P.L, P.N, l, n,
                           P.d_Uptr + P.voff[l], P.d_Ur, val, x, y);
        hipLaunchKernelGGL(k_level_gemv<T>, dim3((n + 3) / 4), dim3(256), sizeof(T) * P.ld[l], s,
                           P.L, P.N, l, n, P.ld[l], 1, W.sinv + P.soff[l], y, W.z + P.voff[l], x);
    }
    V2D_LAUNCH();
    return 0;
}
template int direct_solve<double>(const DirectPattern &, const DirectWork<double> &, const double *,
                                  const double *, double *, hipStream_t);
template int direct_solve<float>(const DirectPattern &, const DirectWork<float> &, const float *,
                                 const float *, float *, hipStream_t);

}  // namespace v2d
