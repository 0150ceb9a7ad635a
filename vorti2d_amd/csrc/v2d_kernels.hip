// v2d_kernels.hip -- gfx950 kernels of the metric / assembly / Krylov hot path.
//
// All kernels run one thread per grid node (or per matrix row), i fastest, so a
// 64-lane wavefront reads 64 consecutive nodes of one radial level: every metric
// and state plane is read coalesced, and the n/s stencil rows of a wavefront are
// the rows its neighbouring wavefronts read, so they hit L2 instead of HBM.
//
// This file is compiled with -ffp-contract=off: the metric and assembly
// arithmetic uses the reference operation order (vorti2d_core.f90:104-182,
// 247-381) and IEEE division / square root, so FP64 results are bit-identical
// to the Fortran on the same inputs.
#include <hip/hip_runtime.h>

#include "v2d_common.h"
#include "v2d_internal.h"

namespace v2d {

// =============================================================== metrics ====
// d/deta at node (i, j): central inside, 3-point one-sided first derivative and
// 4-point one-sided second derivative at j = 0 and j = jmax-1
// (vorti2d_core.f90:104-152).
__device__ __forceinline__ void eta_derivs(const double *__restrict__ v, int imax,
                                           int jmax, int i, int j, double deta,
                                           double &d1, double &d2)
{
    const double *c = v + i;
    if (j == 0) {
        const double v1 = c[0], v2 = c[imax], v3 = c[2 * imax], v4 = c[3 * imax];
        d1 = (-3.0 * v1 + 4.0 * v2 - v3) / (2.0 * deta);
        d2 = (2.0 * v1 - 5.0 * v2 + 4.0 * v3 - v4) / (deta * deta);
    } else if (j == jmax - 1) {
        const int64_t t = (int64_t)imax * (jmax - 1);
        const double v1 = c[t], v2 = c[t - imax], v3 = c[t - 2 * imax], v4 = c[t - 3 * imax];
        d1 = (3.0 * v1 - 4.0 * v2 + v3) / (2.0 * deta);
        d2 = (2.0 * v1 - 5.0 * v2 + 4.0 * v3 - v4) / (deta * deta);
    } else {
        const int64_t t = (int64_t)imax * j;
        const double vn = c[t + imax], vk = c[t], vs = c[t - imax];
        d1 = (vn - vs) / (2.0 * deta);
        d2 = (vn - 2.0 * vk + vs) / (deta * deta);
    }
}

__device__ __forceinline__ double eta_d1(const double *__restrict__ v, int imax, int jmax,
                                         int i, int j, double deta)
{
    double d1, d2;
    eta_derivs(v, imax, jmax, i, j, deta, d1, d2);
    return d1;
}

// Metrics of one node from the derivative set (vorti2d_core.f90:163-182).
__device__ __forceinline__ void metric_point(double xxi, double yxi, double xeta,
                                             double yeta, double xxx, double yxx,
                                             double xee, double yee, double xxe,
                                             double yxe, double *o)
{
    const double jac = 1.0 / (xxi * yeta - xeta * yxi);
    const double kx = jac * yeta, ky = -jac * xeta;
    const double ex = -jac * yxi, ey = jac * xxi;
    const double a = kx * kx + ky * ky;
    const double b = ex * ex + ey * ey;
    const double g = kx * ex + ky * ey;
    o[0] = jac;
    o[1] = a;
    o[2] = b;
    o[3] = g;
    o[4] = -(a * (xxx * kx + yxx * ky) + 2.0 * g * (xxe * kx + yxe * ky) + b * (xee * kx + yee * ky));
    o[5] = -(a * (xxx * ex + yxx * ey) + 2.0 * g * (xxe * ex + yxe * ey) + b * (xee * ex + yee * ey));
    o[6] = ex;
    o[7] = ey;
}

// Global metrics with the seam wrap; out = 10 planes of N doubles.
__global__ void __launch_bounds__(256) k_metrics(int imax, int jmax, double dksi, double deta,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 double *__restrict__ out)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const Nbr b = neighbors(i, j, imax);
    const int ie = b.e - j * imax, iw = b.w - j * imax;   // column of e / w
    const double xe = x[b.e], xw = x[b.w], xk = x[k];
    const double ye = y[b.e], yw = y[b.w], yk = y[k];
    const double xxi = (xe - xw) / (2.0 * dksi);
    const double yxi = (ye - yw) / (2.0 * dksi);
    const double xxx = (xe - 2.0 * xk + xw) / (dksi * dksi);
    const double yxx = (ye - 2.0 * yk + yw) / (dksi * dksi);
    double xeta, yeta, xee, yee;
    eta_derivs(x, imax, jmax, i, j, deta, xeta, xee);
    eta_derivs(y, imax, jmax, i, j, deta, yeta, yee);
    const double xxe = (eta_d1(x, imax, jmax, ie, j, deta) - eta_d1(x, imax, jmax, iw, j, deta)) / (2.0 * dksi);
    const double yxe = (eta_d1(y, imax, jmax, ie, j, deta) - eta_d1(y, imax, jmax, iw, j, deta)) / (2.0 * dksi);
    double o[8];
    metric_point(xxi, yxi, xeta, yeta, xxx, yxx, xee, yee, xxe, yxe, o);
#pragma unroll
    for (int m = 0; m < 8; ++m) out[m * N + k] = o[m];
    out[8 * N + k] = xk;
    out[9 * N + k] = yk;
}

// Metrics on a ghosted (gxm x jmax) block for owned columns [il0, il1] (0-based
// here), no wrap (vorti2d_core.f90:392-495).  Unowned planes stay zero.
__global__ void __launch_bounds__(256) k_metrics_local(int gxm, int jmax, int il0, int il1,
                                                       double dksi, double deta,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ out)
{
    const int64_t N = (int64_t)gxm * jmax;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int j = (int)(k / gxm), i = (int)(k - (int64_t)j * gxm);
    out[8 * N + k] = x[k];
    out[9 * N + k] = y[k];
    if (i < il0 || i > il1) {
#pragma unroll
        for (int m = 0; m < 8; ++m) out[m * N + k] = 0.0;
        return;
    }
    const int64_t e = k + 1, w = k - 1;
    const double xxi = (x[e] - x[w]) / (2.0 * dksi);
    const double yxi = (y[e] - y[w]) / (2.0 * dksi);
    const double xxx = (x[e] - 2.0 * x[k] + x[w]) / (dksi * dksi);
    const double yxx = (y[e] - 2.0 * y[k] + y[w]) / (dksi * dksi);
    double xeta, yeta, xee, yee;
    eta_derivs(x, gxm, jmax, i, j, deta, xeta, xee);
    eta_derivs(y, gxm, jmax, i, j, deta, yeta, yee);
    const double xxe = (eta_d1(x, gxm, jmax, i + 1, j, deta) - eta_d1(x, gxm, jmax, i - 1, j, deta)) / (2.0 * dksi);
    const double yxe = (eta_d1(y, gxm, jmax, i + 1, j, deta) - eta_d1(y, gxm, jmax, i - 1, j, deta)) / (2.0 * dksi);
    double o[8];
    metric_point(xxi, yxi, xeta, yeta, xxx, yxx, xee, yee, xxe, yxe, o);
#pragma unroll
    for (int m = 0; m < 8; ++m) out[m * N + k] = o[m];
}

// ============================================================== assembly ====
// Node-local view used by the row builders.  Index arithmetic is relative to
// the node so the same code serves the global (seam-wrapped) and the local
// (ghosted) layouts.
struct NodeIx {
    int64_t k, e, w, n, s, ne, nw, se, sw;   // 9-point stencil
    int64_t n2, n3, s1, s2;                  // j+1, j+2, j-1, j-2 on the same column
};

// psi row values in emission order; returns the count, writes the RHS
// (vorti2d_core.f90:249-301).
__device__ __forceinline__ int psi_row(const AsmArgs &a, int jt, uint8_t flags,
                                       const NodeIx &s, double *v, double &b)
{
    const int64_t k = s.k;
    const double re = a.re, dk = a.dksi, de = a.deta;
    if (jt == 2) {
        if (flags & kOutPsi) {
            v[0] = 1.0; v[1] = -2.0; v[2] = 1.0;
            b = -(a.psi[k] - 2.0 * a.psi[s.s1] + a.psi[s.s2]);
            return 3;
        }
        v[0] = 1.0;
        b = (a.ca * a.m[9][k] - a.sa * a.m[8][k]) - a.psi[k];
        return 1;
    }
    if (jt == 0) {
        v[0] = 1.0;
        b = 0.0 - a.psi[k];
        return 1;
    }
    const double al = a.m[1][k], bt = a.m[2][k], g = a.m[3][k], p = a.m[4][k], q = a.m[5][k];
    const double cg = (2.0 * g) / (4.0 * re * dk * de);
    const double *f = a.psi;
    const double fk = f[k], fe = f[s.e], fw = f[s.w], fn = f[s.n], fs = f[s.s];
    v[0] = a.invdtau + ((2.0 * al) / (dk * dk) + (2.0 * bt) / (de * de)) / re;
    v[1] = -(al / (re * (dk * dk)) + p / (2.0 * re * dk));
    v[2] = (-al / (re * (dk * dk)) + p / (2.0 * re * dk));
    v[3] = -(bt / (re * (de * de)) + q / (2.0 * re * de));
    v[4] = (-bt / (re * (de * de)) + q / (2.0 * re * de));
    v[5] = -cg; v[6] = cg; v[7] = cg; v[8] = -cg;
    v[9] = -1.0 / re;
    const double t1 = (al / re) * ((fe - 2.0 * fk + fw) / (dk * dk));
    const double t2 = 2.0 * (g / re) * ((f[s.ne] - f[s.nw] - f[s.se] + f[s.sw]) / (4.0 * de * dk));
    const double t3 = (bt / re) * ((fn - 2.0 * fk + fs) / (de * de));
    const double t4 = (p / re) * ((fe - fw) / (2.0 * dk));
    const double t5 = (q / re) * ((fn - fs) / (2.0 * de));
    b = t1 + t2 + t3 + t4 + t5 + a.ome[k] / re;
    return 10;
}

// omega row values in emission order (vorti2d_core.f90:302-379).
__device__ __forceinline__ int ome_row(const AsmArgs &a, int jt, uint8_t flags,
                                       const NodeIx &s, double *v, double &b)
{
    const int64_t k = s.k;
    const double re = a.re, dk = a.dksi, de = a.deta;
    if (jt == 2) {
        if (flags & kOutOme) {
            v[0] = 1.0; v[1] = -4.0 / 3.0; v[2] = 1.0 / 3.0;
            b = (4.0 * a.ome[s.s1] - a.ome[s.s2]) / 3.0 - a.ome[k];
            return 3;
        }
        v[0] = 1.0;
        b = 0.0 - a.ome[k];
        return 1;
    }
    if (jt == 0) {
        const double bt = a.m[2][k];
        const double bw = bt / (de * de);
        v[0] = -1.0;
        v[1] = 3.5 * bw;
        v[2] = -(8.0 / 2.0) * bw;
        v[3] = 0.5 * bw;
        const double xk = a.m[8][k], yk = a.m[9][k];
        const double rad = sqrt(xk * xk + yk * yk);
        const double ct = xk / rad, st = yk / rad;
        b = a.ome[k] - 3.5 * bw * a.psi[k] + 4.0 * bw * a.psi[s.n2] - 0.5 * bw * a.psi[s.n3]
            - ((a.m[5][k] - 3.0 * (bt / de)) * (a.urot / (a.m[6][k] * ct + a.m[7][k] * st)));
        return 4;
    }
    const double *P = a.psi, *O = a.ome;
    const double al = a.m[1][k], bt = a.m[2][k], g = a.m[3][k], p = a.m[4][k], q = a.m[5][k];
    const double J = a.m[0][k];
    const double Ok = O[k], Oe = O[s.e], Ow = O[s.w], On = O[s.n], Os = O[s.s];
    const double Aa = (P[s.n] - P[s.s]) / (2.0 * de);
    const double Bb = (Oe - Ow) / (2.0 * dk);
    const double Cc = (P[s.e] - P[s.w]) / (2.0 * dk);
    const double Dd = (On - Os) / (2.0 * de);
    const double cab = al / (re * (dk * dk)), cbt = bt / (re * (de * de));
    const double cg = (2.0 * g) / (4.0 * re * dk * de);
    v[0] = a.invdtau + (2.0 * al) / (re * (dk * dk)) + (2.0 * bt) / (re * (de * de)) + 3.0 * a.inv2dt;
    v[1] = ((J * Aa) / (2.0 * dk)) - cab - p / (2.0 * re * dk);
    v[2] = (-(J * Aa) / (2.0 * dk)) - cab + p / (2.0 * re * dk);
    v[3] = (-(J * Cc) / (2.0 * de)) - cbt - q / (2.0 * re * de);
    v[4] = ((J * Cc) / (2.0 * de)) - cbt + q / (2.0 * re * de);
    v[5] = -cg; v[6] = cg; v[7] = cg; v[8] = -cg;
    v[9] = (J * Bb) / (2.0 * de);
    v[10] = -(J * Bb) / (2.0 * de);
    v[11] = -(J * Dd) / (2.0 * dk);
    v[12] = (J * Dd) / (2.0 * dk);
    const double e1 = (al / re) * ((Oe - 2.0 * Ok + Ow) / (dk * dk));
    const double e2 = 2.0 * (g / re) * ((O[s.ne] - O[s.nw] - O[s.se] + O[s.sw]) / (4.0 * de * dk));
    const double e3 = (bt / re) * ((On - 2.0 * Ok + Os) / (de * de));
    const double e4 = (p / re) * ((Oe - Ow) / (2.0 * dk));
    const double e5 = (q / re) * ((On - Os) / (2.0 * de));
    // The BDF2 history is read only for an unsteady step; with inv2dt == 0 the
    // term is the finite product 0 * (...) of the reference and contributes 0.
    double hist = 0.0;
    if (a.inv2dt != 0.0) hist = (3.0 * Ok - 4.0 * a.omeold[k] + a.omeoldold[k]) * a.inv2dt;
    b = -(J * Aa * Bb) + (J * Cc * Dd) + e1 + e2 + e3 + e4 + e5 - hist;
    return 13;
}

__device__ __forceinline__ NodeIx global_ix(int imax, int i, int j)
{
    NodeIx s;
    const Nbr b = neighbors(i, j, imax);
    s.k = b.k; s.e = b.e; s.w = b.w; s.n = b.n; s.s = b.s;
    s.ne = b.ne; s.nw = b.nw; s.se = b.se; s.sw = b.sw;
    s.n2 = b.k + imax; s.n3 = b.k + 2 * (int64_t)imax;
    s.s1 = b.k - imax; s.s2 = b.k - 2 * (int64_t)imax;
    return s;
}

// Production assembly: both rows of every node, values into the slot planes
// (cast to T), residual-form RHS in FP64 (field-blocked, 2N).  Same stencils as
// psi_row / ome_row, with every division by a launch constant replaced by a
// product with its reciprocal (AsmConsts, computed once on the host): a few
// ulp from the reference's rounding, and no FP64 divide sequences in the
// interior.  The parity COO kernels keep the reference's exact operations.
template <typename T>
__device__ __forceinline__ void assemble_node(const AsmArgs &a, const AsmConsts &c,
                                              const uint8_t *__restrict__ flags, T *__restrict__ val,
                                              double *__restrict__ rhs, int64_t k)
{
    const int imax = a.imax, jmax = a.jmax;
    const int64_t N = (int64_t)imax * jmax;
    if (k >= N) return;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const NodeIx s = global_ix(imax, i, j);
    const double *P = a.psi, *O = a.ome;
    if (j > 0 && j < jmax - 1) {
        const double J = a.m[0][k], al = a.m[1][k], bt = a.m[2][k], g = a.m[3][k];
        const double p = a.m[4][k], q = a.m[5][k];
        const double Pk = P[k], Pe = P[s.e], Pw = P[s.w], Pn = P[s.n], Ps = P[s.s];
        const double Ok = O[k], Oe = O[s.e], Ow = O[s.w], On = O[s.n], Os = O[s.s];
        const double Px = P[s.ne] - P[s.nw] - P[s.se] + P[s.sw];
        const double Ox = O[s.ne] - O[s.nw] - O[s.se] + O[s.sw];
        const double ca = al * c.c_al, cb = bt * c.c_bt, cp = p * c.c_p, cq = q * c.c_q;
        const double cg = g * c.c_g;
        // psi row (vorti2d_core.f90:249-301)
        val[0 * N + k] = (T)(a.invdtau + (2.0 * al * c.idk2 + 2.0 * bt * c.ide2) * c.ire);
        val[1 * N + k] = (T)(-(ca + cp));
        val[2 * N + k] = (T)(-ca + cp);
        val[3 * N + k] = (T)(-(cb + cq));
        val[4 * N + k] = (T)(-cb + cq);
        val[5 * N + k] = (T)(-cg);
        val[6 * N + k] = (T)cg;
        val[7 * N + k] = (T)cg;
        val[8 * N + k] = (T)(-cg);
        val[9 * N + k] = (T)(-c.ire);
        rhs[k] = c.ire * (al * ((Pe - 2.0 * Pk + Pw) * c.idk2) + 2.0 * g * (Px * c.i4dkde)
                          + bt * ((Pn - 2.0 * Pk + Ps) * c.ide2) + p * ((Pe - Pw) * c.i2dk)
                          + q * ((Pn - Ps) * c.i2de) + Ok);
        // omega row (vorti2d_core.f90:302-379)
        const double Aa = (Pn - Ps) * c.i2de, Bb = (Oe - Ow) * c.i2dk;
        const double Cc = (Pe - Pw) * c.i2dk, Dd = (On - Os) * c.i2de;
        const double ja = J * Aa * c.i2dk, jc = J * Cc * c.i2de;
        const double jb = J * Bb * c.i2de, jd = J * Dd * c.i2dk;
        T *vo = val + kPsiSlots * N;
        vo[0 * N + k] = (T)(a.invdtau + 2.0 * ca + 2.0 * cb + 3.0 * a.inv2dt);
        vo[1 * N + k] = (T)(ja - ca - cp);
        vo[2 * N + k] = (T)(-ja - ca + cp);
        vo[3 * N + k] = (T)(-jc - cb - cq);
        vo[4 * N + k] = (T)(jc - cb + cq);
        vo[5 * N + k] = (T)(-cg);
        vo[6 * N + k] = (T)cg;
        vo[7 * N + k] = (T)cg;
        vo[8 * N + k] = (T)(-cg);
        vo[9 * N + k] = (T)jb;
        vo[10 * N + k] = (T)(-jb);
        vo[11 * N + k] = (T)(-jd);
        vo[12 * N + k] = (T)jd;
        double hist = 0.0;   // the BDF2 history is read only for an unsteady step
        if (a.inv2dt != 0.0) hist = (3.0 * Ok - 4.0 * a.omeold[k] + a.omeoldold[k]) * a.inv2dt;
        rhs[N + k] = -(J * Aa * Bb) + (J * Cc * Dd)
                     + c.ire * (al * ((Oe - 2.0 * Ok + Ow) * c.idk2) + 2.0 * g * (Ox * c.i4dkde)
                                + bt * ((On - 2.0 * Ok + Os) * c.ide2) + p * ((Oe - Ow) * c.i2dk)
                                + q * ((On - Os) * c.i2de))
                     - hist;
        return;
    }
    T *vo = val + kPsiSlots * N;
    if (j == 0) {   // wall: psi Dirichlet, Thom vorticity
        val[k] = (T)1.0;
        rhs[k] = -P[k];
        const double bt = a.m[2][k];
        const double bw = bt * c.ide2;
        vo[0 * N + k] = (T)(-1.0);
        vo[1 * N + k] = (T)(3.5 * bw);
        vo[2 * N + k] = (T)(-4.0 * bw);
        vo[3 * N + k] = (T)(0.5 * bw);
        const double xk = a.m[8][k], yk = a.m[9][k];
        const double rad = sqrt(xk * xk + yk * yk);
        const double ct = xk / rad, st = yk / rad;
        rhs[N + k] = O[k] - 3.5 * bw * P[k] + 4.0 * bw * P[s.n2] - 0.5 * bw * P[s.n3]
                     - ((a.m[5][k] - 3.0 * (bt * c.ide)) * (a.urot / (a.m[6][k] * ct + a.m[7][k] * st)));
        return;
    }
    const uint8_t fl = flags[i];   // far field
    if (fl & kOutPsi) {
        val[0 * N + k] = (T)1.0;
        val[1 * N + k] = (T)(-2.0);
        val[2 * N + k] = (T)1.0;
        rhs[k] = -(P[k] - 2.0 * P[s.s1] + P[s.s2]);
    } else {
        val[k] = (T)1.0;
        rhs[k] = (a.ca * a.m[9][k] - a.sa * a.m[8][k]) - P[k];
    }
    if (fl & kOutOme) {
        vo[0 * N + k] = (T)1.0;
        vo[1 * N + k] = (T)(-4.0 / 3.0);
        vo[2 * N + k] = (T)(1.0 / 3.0);
        rhs[N + k] = (4.0 * O[s.s1] - O[s.s2]) * (1.0 / 3.0) - O[k];
    } else {
        vo[k] = (T)1.0;
        rhs[N + k] = -O[k];
    }
}
template <typename T>
__global__ void __launch_bounds__(256) k_assemble(AsmArgs a, AsmConsts c,
                                                  const uint8_t *__restrict__ flags,
                                                  T *__restrict__ val, double *__restrict__ rhs)
{
    assemble_node<T>(a, c, flags, val, rhs, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}
template __global__ void k_assemble<double>(AsmArgs, AsmConsts, const uint8_t *, double *, double *);
template __global__ void k_assemble<float>(AsmArgs, AsmConsts, const uint8_t *, float *, double *);

// Parity emission: rows [r0, r1) of the field-blocked system written in the
// reference COO order (vorti2d_core.f90:247-381) with FP64 values.
__global__ void __launch_bounds__(256) k_assemble_coo(AsmArgs a, const uint8_t *__restrict__ flags,
                                                      const int *__restrict__ pre_psi,
                                                      const int *__restrict__ pre_ome,
                                                      int npsi_out, int64_t r0, int64_t r1,
                                                      int *__restrict__ ci, int *__restrict__ cj,
                                                      double *__restrict__ cv,
                                                      double *__restrict__ bvec)
{
    const int imax = a.imax, jmax = a.jmax;
    const int64_t N = (int64_t)imax * jmax;
    const int64_t r = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    const int comp = r >= N;
    const int64_t k = comp ? r - N : r;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const int jt = j == 0 ? 0 : (j == jmax - 1 ? 2 : 1);
    const uint8_t fl = jt == 2 ? flags[i] : 0;
    const NodeIx s = global_ix(imax, i, j);
    double v[13], b;
    int64_t cols[13];
    const int n = comp ? ome_row(a, jt, fl, s, v, b) : psi_row(a, jt, fl, s, v, b);
    row_columns(imax, jmax, i, j, comp, fl, N, cols);
    const int64_t off = row_start(imax, jmax, r, pre_psi, pre_ome, npsi_out)
                        - row_start(imax, jmax, r0, pre_psi, pre_ome, npsi_out);
    for (int q = 0; q < n; ++q) {
        ci[off + q] = (int)r;
        cj[off + q] = (int)cols[q];
        cv[off + q] = v[q];
    }
    bvec[r - r0] = b;
}

// Local (ghosted, node-interleaved) parity emission (vorti2d_core.f90:510-661):
// owned node t in [0, nown) maps to (i = il0 + t % ow, j = t / ow); psi row then
// omega row; dofs through lg.  Entry offsets per owned node are precomputed.
__global__ void __launch_bounds__(256) k_assemble_coo_local(AsmArgs a, int il0, int ow,
                                                            const int *__restrict__ lg,
                                                            const int64_t *__restrict__ node_off,
                                                            int r0, int nloc,
                                                            int *__restrict__ ci, int *__restrict__ cj,
                                                            double *__restrict__ cv,
                                                            double *__restrict__ bvec)
{
    const int gxm = a.imax, jmax = a.jmax;
    const int64_t nown = (int64_t)ow * jmax;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nown) return;
    const int j = (int)(t / ow), i = il0 + (int)(t - (int64_t)j * ow);
    const int64_t k = (int64_t)gxm * j + i;
    NodeIx s;
    s.k = k; s.e = k + 1; s.w = k - 1; s.n = k + gxm; s.s = k - gxm;
    s.ne = s.n + 1; s.nw = s.n - 1; s.se = s.s + 1; s.sw = s.s - 1;
    s.n2 = k + gxm; s.n3 = k + 2 * (int64_t)gxm; s.s1 = k - gxm; s.s2 = k - 2 * (int64_t)gxm;
    const int jt = j == 0 ? 0 : (j == jmax - 1 ? 2 : 1);
    uint8_t fl = 0;
    if (jt == 2) {
        const bool out = (a.ca * a.m[6][k] + a.sa * a.m[7][k]) > 0.0;
        if (out && a.ff_bc >= 1) fl |= kOutOme;
        if (out && a.ff_bc >= 2) fl |= kOutPsi;
    }
    int64_t m = node_off[t];
    for (int comp = 0; comp < 2; ++comp) {
        double v[13], b;
        const int n = comp ? ome_row(a, jt, fl, s, v, b) : psi_row(a, jt, fl, s, v, b);
        // local column nodes in emission order (same structure as row_columns)
        int64_t cn[13];
        int cc[13];
        if (jt == 2) {
            cn[0] = k; cn[1] = s.s1; cn[2] = s.s2;
            cc[0] = cc[1] = cc[2] = comp;
        } else if (jt == 0) {
            if (comp == 0) { cn[0] = k; cc[0] = 0; }
            else {
                cn[0] = k; cc[0] = 1; cn[1] = k; cc[1] = 0;
                cn[2] = s.n2; cc[2] = 0; cn[3] = s.n3; cc[3] = 0;
            }
        } else {
            const int64_t nb9[9] = {s.k, s.e, s.w, s.n, s.s, s.ne, s.nw, s.se, s.sw};
            for (int q = 0; q < 9; ++q) { cn[q] = nb9[q]; cc[q] = comp; }
            if (comp == 0) { cn[9] = k; cc[9] = 1; }
            else {
                cn[9] = s.n; cn[10] = s.s; cn[11] = s.e; cn[12] = s.w;
                cc[9] = cc[10] = cc[11] = cc[12] = 0;
            }
        }
        const int row = lg[2 * k + comp];
        for (int q = 0; q < n; ++q) {
            ci[m + q] = row;
            cj[m + q] = lg[2 * cn[q] + cc[q]];
            cv[m + q] = v[q];
        }
        m += n;
        if (row - r0 >= 0 && row - r0 < nloc) bvec[row - r0] = b;
    }
}

// ===================================================== slot-layout exports ===
// Slot planes -> reference COO order over all 2N rows (values as FP64).
template <typename T>
__global__ void __launch_bounds__(256) k_slots_to_coo(int imax, int jmax, const uint8_t *__restrict__ flags,
                                                      const int *__restrict__ pre_psi,
                                                      const int *__restrict__ pre_ome, int npsi_out,
                                                      const T *__restrict__ val,
                                                      int *__restrict__ ci, int *__restrict__ cj,
                                                      double *__restrict__ cv)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 2 * N) return;
    const int comp = r >= N;
    const int64_t k = comp ? r - N : r;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const uint8_t fl = j == jmax - 1 ? flags[i] : 0;
    int64_t cols[13];
    const int n = row_columns(imax, jmax, i, j, comp, fl, N, cols);
    const int64_t off = row_start(imax, jmax, r, pre_psi, pre_ome, npsi_out);
    const int base = comp ? kPsiSlots : 0;
    for (int q = 0; q < n; ++q) {
        ci[off + q] = (int)r;
        cj[off + q] = (int)cols[q];
        cv[off + q] = (double)val[(base + q) * N + k];
    }
}
template __global__ void k_slots_to_coo<double>(int, int, const uint8_t *, const int *, const int *, int,
                                                const double *, int *, int *, double *);
template __global__ void k_slots_to_coo<float>(int, int, const uint8_t *, const int *, const int *, int,
                                               const float *, int *, int *, double *);

// Slot planes -> canonical CSR (columns ascending within each row; the pattern
// has no duplicate (row, col) for imax >= 4).  indptr is the COO row start.
template <typename T>
__global__ void __launch_bounds__(256) k_slots_to_csr(int imax, int jmax, const uint8_t *__restrict__ flags,
                                                      const int *__restrict__ pre_psi,
                                                      const int *__restrict__ pre_ome, int npsi_out,
                                                      const T *__restrict__ val,
                                                      int64_t *__restrict__ indptr,
                                                      int *__restrict__ indices,
                                                      double *__restrict__ data)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > 2 * N) return;
    if (r == 2 * N) {
        indptr[r] = row_start(imax, jmax, r - 1, pre_psi, pre_ome, npsi_out)
                    + row_count(jmax, jmax - 1, 1, flags[imax - 1]);
        return;
    }
    const int comp = r >= N;
    const int64_t k = comp ? r - N : r;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const uint8_t fl = j == jmax - 1 ? flags[i] : 0;
    int64_t cols[13];
    double vals[13];
    const int n = row_columns(imax, jmax, i, j, comp, fl, N, cols);
    const int base = comp ? kPsiSlots : 0;
    for (int q = 0; q < n; ++q) vals[q] = (double)val[(base + q) * N + k];
    for (int a = 1; a < n; ++a) {   // insertion sort by column
        const int64_t c = cols[a];
        const double x = vals[a];
        int b = a - 1;
        while (b >= 0 && cols[b] > c) { cols[b + 1] = cols[b]; vals[b + 1] = vals[b]; --b; }
        cols[b + 1] = c;
        vals[b + 1] = x;
    }
    const int64_t off = row_start(imax, jmax, r, pre_psi, pre_ome, npsi_out);
    indptr[r] = off;
    for (int q = 0; q < n; ++q) {
        indices[off + q] = (int)cols[q];
        data[off + q] = vals[q];
    }
}
template __global__ void k_slots_to_csr<double>(int, int, const uint8_t *, const int *, const int *, int,
                                                const double *, int64_t *, int *, double *);
template __global__ void k_slots_to_csr<float>(int, int, const uint8_t *, const int *, const int *, int,
                                               const float *, int64_t *, int *, double *);

// ================================================================== SpMV ====
// y = A x on the slot layout with implied columns; x, y field-blocked (2N).
// One thread per node computes both rows.  Accumulation type A (double for
// FP64 vectors, float for FP32 vectors).
template <typename T, typename V>
__device__ __forceinline__ void spmv_node(int imax, int jmax, const uint8_t *__restrict__ flags,
                                          const T *__restrict__ val, const V *__restrict__ x,
                                          V *__restrict__ y, int64_t k)
{
    const int64_t N = (int64_t)imax * jmax;
    if (k >= N) return;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const V *xp = x, *xo = x + N;
    const T *vp = val + k;
    V yp, yo;
    if (j == 0) {
        yp = (V)vp[0] * xp[k];
        yo = (V)vp[10 * N] * xo[k] + (V)vp[11 * N] * xp[k] + (V)vp[12 * N] * xp[k + imax]
             + (V)vp[13 * N] * xp[k + 2 * imax];
    } else if (j == jmax - 1) {
        const uint8_t fl = flags[i];
        yp = (V)vp[0] * xp[k];
        if (fl & kOutPsi) yp += (V)vp[N] * xp[k - imax] + (V)vp[2 * N] * xp[k - 2 * imax];
        yo = (V)vp[10 * N] * xo[k];
        if (fl & kOutOme) yo += (V)vp[11 * N] * xo[k - imax] + (V)vp[12 * N] * xo[k - 2 * imax];
    } else {
        const Nbr b = neighbors(i, j, imax);
        yp = (V)vp[0] * xp[b.k] + (V)vp[N] * xp[b.e] + (V)vp[2 * N] * xp[b.w] + (V)vp[3 * N] * xp[b.n]
             + (V)vp[4 * N] * xp[b.s] + (V)vp[5 * N] * xp[b.ne] + (V)vp[6 * N] * xp[b.nw]
             + (V)vp[7 * N] * xp[b.se] + (V)vp[8 * N] * xp[b.sw] + (V)vp[9 * N] * xo[b.k];
        yo = (V)vp[10 * N] * xo[b.k] + (V)vp[11 * N] * xo[b.e] + (V)vp[12 * N] * xo[b.w]
             + (V)vp[13 * N] * xo[b.n] + (V)vp[14 * N] * xo[b.s] + (V)vp[15 * N] * xo[b.ne]
             + (V)vp[16 * N] * xo[b.nw] + (V)vp[17 * N] * xo[b.se] + (V)vp[18 * N] * xo[b.sw]
             + (V)vp[19 * N] * xp[b.n] + (V)vp[20 * N] * xp[b.s] + (V)vp[21 * N] * xp[b.e]
             + (V)vp[22 * N] * xp[b.w];
    }
    y[k] = yp;
    y[N + k] = yo;
}
template <typename T, typename V>
__global__ void __launch_bounds__(256) k_spmv(int imax, int jmax, const uint8_t *__restrict__ flags,
                                              const T *__restrict__ val, const V *__restrict__ x,
                                              V *__restrict__ y)
{
    spmv_node<T, V>(imax, jmax, flags, val, x, y, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}
template __global__ void k_spmv<double, double>(int, int, const uint8_t *, const double *, const double *, double *);
template __global__ void k_spmv<float, float>(int, int, const uint8_t *, const float *, const float *, float *);

// ======================================================= vector reductions ==
// partial[b*m + c] = sum over the block's share of  V_c . w   (c < m), FP64
// accumulation.  V is column-major with leading dimension ld.  Fused
// multi-dot for classical Gram-Schmidt: w is read once per block for all c.
// Block `blk` of `nblk` (256 threads) of the fused multi-dot.
template <typename V>
__device__ __forceinline__ void multidot_block(int64_t n, int m, const V *__restrict__ Vb, int64_t ld,
                                               const V *__restrict__ w, double *__restrict__ partial,
                                               int blk, int nblk, double *sh)
{
    constexpr int kMax = 8;
    for (int c0 = 0; c0 < m; c0 += kMax) {
        const int mc = min(kMax, m - c0);
        double acc[kMax];
#pragma unroll
        for (int c = 0; c < kMax; ++c) acc[c] = 0.0;
        for (int64_t t = (int64_t)blk * blockDim.x + threadIdx.x; t < n;
             t += (int64_t)nblk * blockDim.x) {
            const double wt = (double)w[t];
#pragma unroll
            for (int c = 0; c < kMax; ++c)
                if (c < mc) acc[c] += (double)Vb[(c0 + c) * ld + t] * wt;
        }
        for (int c = 0; c < mc; ++c) {
            const double s = block_sum(acc[c], sh);
            if (threadIdx.x == 0) partial[(int64_t)blk * m + c0 + c] = s;
        }
    }
}
template <typename V>
__global__ void __launch_bounds__(256) k_multidot(int64_t n, int m, const V *__restrict__ Vb, int64_t ld,
                                                  const V *__restrict__ w, double *__restrict__ partial)
{
    __shared__ double sh[17];
    multidot_block<V>(n, m, Vb, ld, w, partial, blockIdx.x, gridDim.x, sh);
}
template __global__ void k_multidot<double>(int64_t, int, const double *, int64_t, const double *, double *);
template __global__ void k_multidot<float>(int64_t, int, const float *, int64_t, const float *, double *);

// out[c] = sum_b partial[b*m + c]
__device__ __forceinline__ void reduce_partials_block(int nb, int m, const double *__restrict__ partial,
                                                      double *__restrict__ out, double *sh)
{
    for (int c = 0; c < m; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[(int64_t)b * m + c];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) out[c] = s;
    }
}
__global__ void k_reduce_partials(int nb, int m, const double *__restrict__ partial, double *__restrict__ out)
{
    __shared__ double sh[17];
    reduce_partials_block(nb, m, partial, out, sh);
}

// w -= sum_c h[c] V_c  (h FP64 on device)
template <typename V>
__global__ void __launch_bounds__(256) k_multiaxpy(int64_t n, int m, const V *__restrict__ Vb, int64_t ld,
                                                   const double *__restrict__ h, V *__restrict__ w)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = (double)w[t];
    for (int c = 0; c < m; ++c) acc -= h[c] * (double)Vb[c * ld + t];
    w[t] = (V)acc;
}
template __global__ void k_multiaxpy<double>(int64_t, int, const double *, int64_t, const double *, double *);
template __global__ void k_multiaxpy<float>(int64_t, int, const float *, int64_t, const double *, float *);

// y = x * f with f = s[0] (inv 0), 1 / s[0] (inv 1) or, from a squared norm,
// 1 / sqrt(s[0]) (inv 2; 1 where that norm is not positive); s on device,
// FP64, with type conversion
template <typename A, typename B>
__global__ void __launch_bounds__(256) k_scale_copy(int64_t n, const A *__restrict__ x, const double *__restrict__ s,
                                                    int inv, B *__restrict__ y)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double f = s[0];
    if (inv == 1) f = 1.0 / f;
    if (inv == 2) {
        const double hn = sqrt(fmax(f, 0.0));
        f = hn > 0.0 ? 1.0 / hn : 1.0;
    }
    y[t] = (B)((double)x[t] * f);
}
template __global__ void k_scale_copy<double, double>(int64_t, const double *, const double *, int, double *);
template __global__ void k_scale_copy<double, float>(int64_t, const double *, const double *, int, float *);
template __global__ void k_scale_copy<float, float>(int64_t, const float *, const double *, int, float *);

// x += sum_c y[c] Z_c (FP64 coefficients), result accumulated in X type
template <typename Z, typename X>
__global__ void __launch_bounds__(256) k_update_sol(int64_t n, int m, const Z *__restrict__ Zb, int64_t ld,
                                                    const double *__restrict__ y, double scale,
                                                    X *__restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = 0.0;
    for (int c = 0; c < m; ++c) acc += y[c] * (double)Zb[c * ld + t];
    x[t] = (X)((double)x[t] + scale * acc);
}
template __global__ void k_update_sol<double, double>(int64_t, int, const double *, int64_t, const double *, double, double *);
template __global__ void k_update_sol<float, double>(int64_t, int, const float *, int64_t, const double *, double, double *);

// r = b - A x for an FP64 right-hand side and an FP64 iterate, via the T slot
// values: used to restart GMRES.
template <typename T>
__global__ void __launch_bounds__(256) k_residual(int imax, int jmax, const uint8_t *__restrict__ flags,
                                                  const T *__restrict__ val, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *__restrict__ r)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    // reuse the SpMV formula in FP64
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const double *xp = x, *xo = x + N;
    const T *vp = val + k;
    double yp, yo;
    if (j == 0) {
        yp = (double)vp[0] * xp[k];
        yo = (double)vp[10 * N] * xo[k] + (double)vp[11 * N] * xp[k] + (double)vp[12 * N] * xp[k + imax]
             + (double)vp[13 * N] * xp[k + 2 * imax];
    } else if (j == jmax - 1) {
        const uint8_t fl = flags[i];
        yp = (double)vp[0] * xp[k];
        if (fl & kOutPsi) yp += (double)vp[N] * xp[k - imax] + (double)vp[2 * N] * xp[k - 2 * imax];
        yo = (double)vp[10 * N] * xo[k];
        if (fl & kOutOme) yo += (double)vp[11 * N] * xo[k - imax] + (double)vp[12 * N] * xo[k - 2 * imax];
    } else {
        const Nbr nb = neighbors(i, j, imax);
        yp = (double)vp[0] * xp[nb.k] + (double)vp[N] * xp[nb.e] + (double)vp[2 * N] * xp[nb.w]
             + (double)vp[3 * N] * xp[nb.n] + (double)vp[4 * N] * xp[nb.s] + (double)vp[5 * N] * xp[nb.ne]
             + (double)vp[6 * N] * xp[nb.nw] + (double)vp[7 * N] * xp[nb.se] + (double)vp[8 * N] * xp[nb.sw]
             + (double)vp[9 * N] * xo[nb.k];
        yo = (double)vp[10 * N] * xo[nb.k] + (double)vp[11 * N] * xo[nb.e] + (double)vp[12 * N] * xo[nb.w]
             + (double)vp[13 * N] * xo[nb.n] + (double)vp[14 * N] * xo[nb.s] + (double)vp[15 * N] * xo[nb.ne]
             + (double)vp[16 * N] * xo[nb.nw] + (double)vp[17 * N] * xo[nb.se] + (double)vp[18 * N] * xo[nb.sw]
             + (double)vp[19 * N] * xp[nb.n] + (double)vp[20 * N] * xp[nb.s] + (double)vp[21 * N] * xp[nb.e]
             + (double)vp[22 * N] * xp[nb.w];
    }
    r[k] = b[k] - yp;
    r[N + k] = b[N + k] - yo;
}
template __global__ void k_residual<double>(int, int, const uint8_t *, const double *, const double *, const double *, double *);
template __global__ void k_residual<float>(int, int, const uint8_t *, const float *, const double *, const double *, double *);

__global__ void __launch_bounds__(256) k_fill_zero_d(int64_t n, double *__restrict__ p)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill_one(int64_t n, T *__restrict__ p)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = T(1);
}
template __global__ void k_fill_one<double>(int64_t, double *);
template __global__ void k_fill_one<float>(int64_t, float *);

// psi += d[0:N], ome += d[N:2N]
__global__ void __launch_bounds__(256) k_apply_update(int64_t N, const double *__restrict__ d,
                                                      double *__restrict__ psi, double *__restrict__ ome)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    psi[t] += d[t];
    ome[t] += d[N + t];
}

// ====================================================== batched launches ==
// Every case of a batched launch runs its single-case code on its own blocks
// (BMap); pointwise kernels index by the block's offset inside the case, the
// multi-dot keeps the case's own block count and stride.
template <typename T>
__global__ void __launch_bounds__(1024) k_assemble_bt(BMap bm, const BAsm *__restrict__ cs)
{
    const int c = bmap_case(bm, blockIdx.x);
    const BAsm &q = cs[c];
    assemble_node<T>(q.a, q.k, q.flags, static_cast<T *>(q.val), q.rhs,
                     (int64_t)(blockIdx.x - bm.off[c]) * blockDim.x + threadIdx.x);
}
template __global__ void k_assemble_bt<double>(BMap, const BAsm *);
template __global__ void k_assemble_bt<float>(BMap, const BAsm *);

template <typename T, typename V>
__global__ void __launch_bounds__(1024) k_spmv_bt(BMap bm, const BVec *__restrict__ v)
{
    const int c = bmap_case(bm, blockIdx.x);
    const BVec &q = v[c];
    spmv_node<T, V>(q.imax, q.jmax, q.flags, static_cast<const T *>(q.val), static_cast<const V *>(q.x),
                    static_cast<V *>(q.y), (int64_t)(blockIdx.x - bm.off[c]) * blockDim.x + threadIdx.x);
}
template __global__ void k_spmv_bt<float, float>(BMap, const BVec *);
template __global__ void k_spmv_bt<double, double>(BMap, const BVec *);

template <typename V>
__global__ void __launch_bounds__(256) k_multidot_bt(BMap bm, const BVec *__restrict__ v)
{
    __shared__ double sh[17];
    const int c = bmap_case(bm, blockIdx.x);
    const BVec &q = v[c];
    multidot_block<V>(q.n, q.m, static_cast<const V *>(q.Vb), q.ld, static_cast<const V *>(q.w), q.partial,
                      blockIdx.x - bm.off[c], q.nb, sh);
}
template __global__ void k_multidot_bt<float>(BMap, const BVec *);
template __global__ void k_multidot_bt<double>(BMap, const BVec *);

__global__ void __launch_bounds__(256) k_reduce_bt(const BVec *__restrict__ v)
{
    __shared__ double sh[17];
    const BVec &q = v[blockIdx.x];
    reduce_partials_block(q.nb, q.m, q.partial, q.out, sh);
}

template <typename V>
__global__ void __launch_bounds__(1024) k_multiaxpy_bt(BMap bm, const BVec *__restrict__ v)
{
    const int c = bmap_case(bm, blockIdx.x);
    const BVec &q = v[c];
    const int64_t t = (int64_t)(blockIdx.x - bm.off[c]) * blockDim.x + threadIdx.x;
    if (t >= q.n) return;
    const V *Vb = static_cast<const V *>(q.Vb);
    V *w = static_cast<V *>(q.y);
    double acc = (double)w[t];
    for (int k = 0; k < q.m; ++k) acc -= q.h[k] * (double)Vb[k * q.ld + t];
    w[t] = (V)acc;
}
template __global__ void k_multiaxpy_bt<float>(BMap, const BVec *);
template __global__ void k_multiaxpy_bt<double>(BMap, const BVec *);

template <typename A, typename B>
__global__ void __launch_bounds__(1024) k_scale_copy_bt(BMap bm, const BVec *__restrict__ v)
{
    const int c = bmap_case(bm, blockIdx.x);
    const BVec &q = v[c];
    const int64_t t = (int64_t)(blockIdx.x - bm.off[c]) * blockDim.x + threadIdx.x;
    if (t >= q.n) return;
    const double f = q.inv ? 1.0 / q.s[0] : q.s[0];
    static_cast<B *>(q.y)[t] = (B)((double)static_cast<const A *>(q.x)[t] * f);
}
template __global__ void k_scale_copy_bt<float, float>(BMap, const BVec *);
template __global__ void k_scale_copy_bt<double, float>(BMap, const BVec *);
template __global__ void k_scale_copy_bt<double, double>(BMap, const BVec *);

__global__ void __launch_bounds__(1024) k_apply_update_bt(BMap bm, const BVec *__restrict__ v)
{
    const int c = bmap_case(bm, blockIdx.x);
    const BVec &q = v[c];
    const int64_t t = (int64_t)(blockIdx.x - bm.off[c]) * blockDim.x + threadIdx.x;
    if (t >= q.n) return;
    const double *d = static_cast<const double *>(q.x);
    q.psi[t] += d[t];
    q.ome[t] += d[q.n + t];
}

// ================================================================ forces ====
// Wall-traction force / moment coefficients (vorti2d/forces.py:82-194), one
// workgroup, FP64.  out: cd, cl, cm, cdp, cdf, clp, clf, cmp, cmf, ref_length.
__global__ void __launch_bounds__(1024) k_forces(int imax, double re, double dksi, double deta,
                                                 const double *__restrict__ m,   // 10 planes
                                                 int64_t N, const double *__restrict__ ome,
                                                 double ref_length, double x0, double y0,
                                                 double *__restrict__ out)
{
    __shared__ double sh[17];
    const int P = imax - 1;
    const double *beta = m + 2 * N, *gama = m + 3 * N, *ex = m + 6 * N, *ey = m + 7 * N;
    const double *xp = m + 8 * N, *yp = m + 9 * N;
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int w = threadIdx.x; w < P; w += blockDim.x) {
        const int wn = (w + 1) % P, wp = (w + P - 1) % P;
        const double xw = xp[w], yw = yp[w], ow = ome[w];
        const double inv = 1.0 / sqrt(beta[w]);
        const double nx = ex[w] * inv, ny = ey[w] * inv;
        const double tx = ny, ty = -nx;
        const double oxi = (ome[wn] - ome[wp]) / (2.0 * dksi);
        const double oeta = (-3.0 * ow + 4.0 * ome[imax + w] - ome[2 * imax + w]) / (2.0 * deta);
        const double dwdn = (gama[w] * oxi + beta[w] * oeta) * inv;
        const double segf = sqrt((xp[wn] - xw) * (xp[wn] - xw) + (yp[wn] - yw) * (yp[wn] - yw));
        const double segb = sqrt((xw - xp[wp]) * (xw - xp[wp]) + (yw - yp[wp]) * (yw - yp[wp]));
        const double ds = 0.5 * (segf + segb);
        const double fric = ow / re;
        const double rx = xw - x0, ry = yw - y0;
        s[0] += yw * dwdn * ds;                    // fxp * re
        s[1] += xw * dwdn * ds;                    // -fyp * re
        s[2] += -fric * tx * ds;                   // fxf
        s[3] += -fric * ty * ds;                   // fyf
        s[4] += (rx * rx + ry * ry) * dwdn * ds;   // -2 re mzp
        s[5] += (rx * (-fric * ty) - ry * (-fric * tx)) * ds;   // mzf
        s[6] += ds;
        s[7] += xw * ds;
        s[8] += yw * ds;
    }
    double t[9];
    for (int c = 0; c < 9; ++c) t[c] = block_sum(s[c], sh);
    double d = ref_length;
    if (!(d > 0.0)) {   // body diameter from the wall: 2 x arc-weighted mean radius
        const double cx = t[7] / t[6], cy = t[8] / t[6];
        double rs = 0.0;
        for (int w = threadIdx.x; w < P; w += blockDim.x) {
            const int wn = (w + 1) % P, wp = (w + P - 1) % P;
            const double xw = xp[w], yw = yp[w];
            const double segf = sqrt((xp[wn] - xw) * (xp[wn] - xw) + (yp[wn] - yw) * (yp[wn] - yw));
            const double segb = sqrt((xw - xp[wp]) * (xw - xp[wp]) + (yw - yp[wp]) * (yw - yp[wp]));
            rs += sqrt((xw - cx) * (xw - cx) + (yw - cy) * (yw - cy)) * 0.5 * (segf + segb);
        }
        rs = block_sum(rs, sh);
        d = 2.0 * rs / t[6];
    }
    if (threadIdx.x == 0) {
        const double fxp = (1.0 / re) * t[0], fyp = -(1.0 / re) * t[1];
        const double mzp = -(1.0 / (2.0 * re)) * t[4];
        const double cdp = 2.0 * fxp / d, cdf = 2.0 * t[2] / d;
        const double clp = 2.0 * fyp / d, clf = 2.0 * t[3] / d;
        const double cmp = 2.0 * mzp / (d * d), cmf = 2.0 * t[5] / (d * d);
        out[0] = cdp + cdf; out[1] = clp + clf; out[2] = cmp + cmf;
        out[3] = cdp; out[4] = cdf; out[5] = clp; out[6] = clf; out[7] = cmp; out[8] = cmf;
        out[9] = d;
    }
}

}  // namespace v2d
