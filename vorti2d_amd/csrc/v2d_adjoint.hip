// Discrete adjoint of the steady force coefficients: vector-Jacobian products
// of the force integration, the residual and the grid metrics.
//
// With b(U; m(X), alpha, Re) the residual-form right-hand side of the
// assembly and A = -db/dU the Newton matrix, a converged state U* and an
// objective J(U*, m, Re) give
//     dJ/dtheta = dJ/dtheta|_explicit + lambda^T db/dtheta,   A^T lambda = dJ/dU,
// for theta in {alpha, Re, mesh coordinates}.  The kernels below produce the
// pieces: k_forces_vjp (dJ/dU, explicit dJ/dm, dJ/dRe for J = wd Cd + wl Cl of
// vorti2d/forces.py:82-194), k_residual_vjp (lambda^T db/dm, db/dRe,
// db/dalpha for the rows of vorti2d_core.f90:249-379) and k_metrics_vjp_* (the
// metric backward of vorti2d_core.f90:66-189 to the node coordinates).
#include <hip/hip_runtime.h>

#include "v2d_internal.h"

namespace v2d {

// --------------------------------------------------- BDF2 time coupling ---
// Right-hand side of the adjoint at time level n of an unsteady run: level
// n + 1 reads omega^n through -(3 w - 4 w^n + w^(n-1)) inv2dt and level n + 2
// through + ... w^n inv2dt, on the interior omega rows only (wall and
// far-field rows carry no BDF2 term; vorti2d_core.f90:340-379), so
//     g[N + k] += inv2dt (4 lambda_{n+1}[N + k] - lambda_{n+2}[N + k]).
__global__ void k_bdf2_adjoint_rhs(int imax, int jmax, double inv2dt, const double *__restrict__ l1,
                                   const double *__restrict__ l2, double *__restrict__ g)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int64_t j = k / imax;
    if (j == 0 || j == jmax - 1) return;
    g[N + k] += inv2dt * (4.0 * l1[N + k] - l2[N + k]);
}

// ------------------------------------------------------------ forces ------
// J = wd Cd + wl Cl with Cd = 2 Fx / d, Cl = 2 Fy / d (body axes, fixed d):
//     J = s * sum_i [ (wd y_i - wl x_i) dwdn_i ds_i - w_i (wd ny_i - wl nx_i) ds_i ],
//     s = 2 / (d Re),
// over the P = imax - 1 distinct wall nodes.  Wall node w's terms reach its own
// entries and its two neighbours' (the periodic xi difference of omega and
// the arc lengths ds), so each thread recomputes the terms of w - 1 and w + 1
// and adds the three contributions to node w in a fixed order: every entry of
// g_ome (N) and g_m (10 planes: beta, gamma, eta_x, eta_y, x, y at the wall)
// has one writer and the result does not depend on scheduling.
// out = {Cd, Cl, dJ/dRe}.  One workgroup.
struct WallTerms {
    double a;              // g_oxi / (2 dksi): + to g_ome[w + 1], - to g_ome[w - 1]
    double gome;           // to g_ome[w]
    double gome1, gome2;   // to g_ome[imax + w], g_ome[2 imax + w]
    double gb, gg, gex, gey;
    double gx, gy;         // to g_x[w], g_y[w]
    double hfx, hfy;       // to g_x[w + 1], g_y[w + 1]
    double hbx, hby;       // minus these to g_x[w - 1], g_y[w - 1]
    double fx, fy;         // the force sums' terms (times Re)
};

__device__ WallTerms wall_terms(int w, int imax, int P, double dksi, double deta, double s, double wd, double wl,
                                const double *beta, const double *gama, const double *ex, const double *ey,
                                const double *xp, const double *yp, const double *ome)
{
    WallTerms t;
    const int wn = (w + 1) % P, wp = (w + P - 1) % P;
    const double xw = xp[w], yw = yp[w], ow = ome[w];
    const double bw = beta[w], gw = gama[w];
    const double inv = 1.0 / sqrt(bw);
    const double nx = ex[w] * inv, ny = ey[w] * inv;
    const double oxi = (ome[wn] - ome[wp]) / (2.0 * dksi);
    const double oeta = (-3.0 * ow + 4.0 * ome[imax + w] - ome[2 * imax + w]) / (2.0 * deta);
    const double dwdn = (gw * oxi + bw * oeta) * inv;
    const double dxf = xp[wn] - xw, dyf = yp[wn] - yw;
    const double dxb = xw - xp[wp], dyb = yw - yp[wp];
    const double segf = sqrt(dxf * dxf + dyf * dyf), segb = sqrt(dxb * dxb + dyb * dyb);
    const double ds = 0.5 * (segf + segb);
    t.fx = yw * dwdn * ds - ow * ny * ds;
    t.fy = -xw * dwdn * ds + ow * nx * ds;
    // --- pressure term  s (wd y - wl x) dwdn ds
    const double cp = wd * yw - wl * xw;
    const double g_dwdn = s * cp * ds;
    const double g_oxi = g_dwdn * gw * inv, g_oeta = g_dwdn * bw * inv;
    t.a = g_oxi / (2.0 * dksi);
    t.gome1 = 4.0 * g_oeta / (2.0 * deta);
    t.gome2 = -g_oeta / (2.0 * deta);
    double gb = g_dwdn * (oeta * inv - 0.5 * dwdn / bw);
    t.gg = g_dwdn * oxi * inv;
    double gx = -s * wl * dwdn * ds, gy = s * wd * dwdn * ds;
    double g_ds = s * cp * dwdn;
    // --- friction term  -s w (wd ny - wl nx) ds,  n = (eta_x, eta_y) / sqrt(beta)
    const double cf = wd * ny - wl * nx;
    t.gome = -3.0 * g_oeta / (2.0 * deta) - s * cf * ds;
    t.gey = -s * ow * wd * inv * ds;
    t.gex = s * ow * wl * inv * ds;
    gb += -s * ow * (wd * ey[w] - wl * ex[w]) * ds * (-0.5 * inv / bw);
    t.gb = gb;
    g_ds += -s * ow * cf;
    // --- ds = (|X_wn - X_w| + |X_w - X_wp|) / 2
    const double hf = 0.5 * g_ds / segf, hb = 0.5 * g_ds / segb;
    t.gx = gx - hf * dxf + hb * dxb;
    t.gy = gy - hf * dyf + hb * dyb;
    t.hfx = hf * dxf;
    t.hfy = hf * dyf;
    t.hbx = hb * dxb;
    t.hby = hb * dyb;
    return t;
}

__global__ void __launch_bounds__(1024) k_forces_vjp(int imax, double re, double dksi, double deta,
                                                     const double *__restrict__ m, int64_t N,
                                                     const double *__restrict__ ome, double d,
                                                     double wd, double wl, double *__restrict__ g_ome,
                                                     double *__restrict__ g_m, double *__restrict__ out)
{
    __shared__ double sh[17];
    const int P = imax - 1;
    const double *beta = m + 2 * N, *gama = m + 3 * N, *ex = m + 6 * N, *ey = m + 7 * N;
    const double *xp = m + 8 * N, *yp = m + 9 * N;
    double *g_bt = g_m + 2 * N, *g_ga = g_m + 3 * N, *g_ex = g_m + 6 * N, *g_ey = g_m + 7 * N;
    double *g_x = g_m + 8 * N, *g_y = g_m + 9 * N;
    const double s = 2.0 / (d * re);
    double fx = 0.0, fy = 0.0;
    for (int w = threadIdx.x; w < P; w += blockDim.x) {
        const int wn = (w + 1) % P, wp = (w + P - 1) % P;
        const WallTerms t = wall_terms(w, imax, P, dksi, deta, s, wd, wl, beta, gama, ex, ey, xp, yp, ome);
        const WallTerms tp = wall_terms(wp, imax, P, dksi, deta, s, wd, wl, beta, gama, ex, ey, xp, yp, ome);
        const WallTerms tn = wall_terms(wn, imax, P, dksi, deta, s, wd, wl, beta, gama, ex, ey, xp, yp, ome);
        fx += t.fx;
        fy += t.fy;
        g_ome[w] += (t.gome + tp.a) - tn.a;
        g_ome[imax + w] += t.gome1;
        g_ome[2 * imax + w] += t.gome2;
        g_bt[w] += t.gb;
        g_ga[w] += t.gg;
        g_ex[w] += t.gex;
        g_ey[w] += t.gey;
        g_x[w] += (t.gx + tp.hfx) - tn.hbx;
        g_y[w] += (t.gy + tp.hfy) - tn.hby;
    }
    const double Fx = block_sum(fx, sh), Fy = block_sum(fy, sh);
    if (threadIdx.x == 0) {
        const double cd = 2.0 * Fx / (d * re), cl = 2.0 * Fy / (d * re);
        out[0] = cd;
        out[1] = cl;
        out[2] = -(wd * cd + wl * cl) / re;   // J is proportional to 1/Re at fixed state
    }
}

// ---------------------------------------------------------- residual ------
// Node-local: the rows of node k use the metrics of node k only, so
//     g_m[q][k] += lam_psi(k) db_psi(k)/dm_q + lam_ome(k) db_ome(k)/dm_q.
// partial[3 * block + {0, 1}] = block sums of lambda . db/dRe and
// lambda . db/dalpha (alpha in radians).
__global__ void __launch_bounds__(256) k_residual_vjp(AsmArgs a, const uint8_t *__restrict__ flags,
                                                      const double *__restrict__ lam,
                                                      double *__restrict__ g_m, double *__restrict__ partial)
{
    __shared__ double sh[17];
    const int imax = a.imax, jmax = a.jmax;
    const int64_t N = (int64_t)imax * jmax;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double g_re = 0.0, g_al = 0.0;
    if (k < N) {
        const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
        const Nbr b = neighbors(i, j, imax);
        const double lp = lam[k], lo = lam[N + k];
        const double re = a.re, dk = a.dksi, de = a.deta, ire = 1.0 / re;
        const double *P = a.psi, *O = a.ome;
        double *g[10];
        for (int q = 0; q < 10; ++q) g[q] = g_m + q * N;
        if (j > 0 && j < jmax - 1) {
            const double J = a.m[0][k], al = a.m[1][k], bt = a.m[2][k], gm = a.m[3][k];
            const double p = a.m[4][k], q = a.m[5][k];
            // psi row: b = (1/Re)(al Pxx + 2 gm Pxe + bt Pee + p Px + q Pe + omega)
            const double Pxx = (P[b.e] - 2.0 * P[k] + P[b.w]) / (dk * dk);
            const double Pxe = (P[b.ne] - P[b.nw] - P[b.se] + P[b.sw]) / (4.0 * de * dk);
            const double Pee = (P[b.n] - 2.0 * P[k] + P[b.s]) / (de * de);
            const double Px = (P[b.e] - P[b.w]) / (2.0 * dk), Pe = (P[b.n] - P[b.s]) / (2.0 * de);
            const double bpsi = ire * (al * Pxx + 2.0 * gm * Pxe + bt * Pee + p * Px + q * Pe + O[k]);
            // omega row: b = -J Aa Bb + J Cc Dd + (1/Re)(al Oxx + 2 gm Oxe + bt Oee + p Ox + q Oe) - hist
            const double Oxx = (O[b.e] - 2.0 * O[k] + O[b.w]) / (dk * dk);
            const double Oxe = (O[b.ne] - O[b.nw] - O[b.se] + O[b.sw]) / (4.0 * de * dk);
            const double Oee = (O[b.n] - 2.0 * O[k] + O[b.s]) / (de * de);
            const double Ox = (O[b.e] - O[b.w]) / (2.0 * dk), Oe = (O[b.n] - O[b.s]) / (2.0 * de);
            const double Aa = Pe, Bb = Ox, Cc = Px, Dd = Oe;
            const double visc = ire * (al * Oxx + 2.0 * gm * Oxe + bt * Oee + p * Ox + q * Oe);
            g[0][k] += lo * (-Aa * Bb + Cc * Dd);
            g[1][k] += ire * (lp * Pxx + lo * Oxx);
            g[2][k] += ire * (lp * Pee + lo * Oee);
            g[3][k] += 2.0 * ire * (lp * Pxe + lo * Oxe);
            g[4][k] += ire * (lp * Px + lo * Ox);
            g[5][k] += ire * (lp * Pe + lo * Oe);
            (void)J;
            g_re = -ire * (lp * bpsi + lo * visc);
        } else if (j == 0) {
            // omega wall row: b = w - bw(3.5 psi_k - 4 psi_n2 + 0.5 psi_n3) - K urot / Dn,
            // bw = beta / de^2, K = Q - 3 beta / de, Dn = eta_x ct + eta_y st
            const double bt = a.m[2][k], q = a.m[5][k], exk = a.m[6][k], eyk = a.m[7][k];
            const double xk = a.m[8][k], yk = a.m[9][k];
            const double rad = sqrt(xk * xk + yk * yk), ct = xk / rad, st = yk / rad;
            const double Dn = exk * ct + eyk * st;
            const double K = q - 3.0 * (bt / de);
            const double comb = -3.5 * P[k] + 4.0 * P[k + imax] - 0.5 * P[k + 2 * (int64_t)imax];
            g[2][k] += lo * (comb / (de * de) + 3.0 / de * a.urot / Dn);
            g[5][k] += lo * (-a.urot / Dn);
            const double f = lo * K * a.urot / (Dn * Dn);   // d(-K urot / Dn) / dDn * lo
            g[6][k] += f * ct;
            g[7][k] += f * st;
            const double r3 = rad * rad * rad;
            const double dct_dx = yk * yk / r3, dct_dy = -xk * yk / r3;
            const double dst_dx = -xk * yk / r3, dst_dy = xk * xk / r3;
            g[8][k] += f * (exk * dct_dx + eyk * dst_dx);
            g[9][k] += f * (exk * dct_dy + eyk * dst_dy);
        } else {
            const uint8_t fl = flags[i];
            if (!(fl & kOutPsi)) {   // psi = ca y - sa x
                g[8][k] += lp * (-a.sa);
                g[9][k] += lp * a.ca;
                g_al = lp * (-a.sa * a.m[9][k] - a.ca * a.m[8][k]);
            }
        }
    }
    const double sre = block_sum(g_re, sh), sal = block_sum(g_al, sh);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sre;
        partial[2 * blockIdx.x + 1] = sal;
    }
}

// ----------------------------------------------------------- metrics ------
// Backward of metric_point: gradient of the ten node-local derivatives
// D = (x_xi, y_xi, x_eta, y_eta, x_xixi, y_xixi, x_etaeta, y_etaeta, x_xieta,
// y_xieta) from the gradient of (jac, alpha, beta, gamma, P, Q, eta_x, eta_y).
__device__ __forceinline__ void metric_point_vjp(const double *D, const double *gm, double *gD)
{
    const double xk = D[0], yk = D[1], xe = D[2], ye = D[3];
    const double xkk = D[4], ykk = D[5], xee = D[6], yee = D[7], xke = D[8], yke = D[9];
    const double jac = 1.0 / (xk * ye - xe * yk);
    const double kx = jac * ye, ky = -jac * xe, ex = -jac * yk, ey = jac * xk;
    const double al = kx * kx + ky * ky, bt = ex * ex + ey * ey, ga = kx * ex + ky * ey;
    const double A1 = xkk * kx + ykk * ky, A2 = xke * kx + yke * ky, A3 = xee * kx + yee * ky;
    const double B1 = xkk * ex + ykk * ey, B2 = xke * ex + yke * ey, B3 = xee * ex + yee * ey;
    const double g_p = gm[4], g_q = gm[5];
    const double gal = gm[1] - g_p * A1 - g_q * B1;
    const double gga = gm[3] - 2.0 * g_p * A2 - 2.0 * g_q * B2;
    const double gbt = gm[2] - g_p * A3 - g_q * B3;
    const double gA1 = -g_p * al, gA2 = -2.0 * g_p * ga, gA3 = -g_p * bt;
    const double gB1 = -g_q * al, gB2 = -2.0 * g_q * ga, gB3 = -g_q * bt;
    gD[4] = gA1 * kx + gB1 * ex;
    gD[5] = gA1 * ky + gB1 * ey;
    gD[6] = gA3 * kx + gB3 * ex;
    gD[7] = gA3 * ky + gB3 * ey;
    gD[8] = gA2 * kx + gB2 * ex;
    gD[9] = gA2 * ky + gB2 * ey;
    const double gkx = 2.0 * gal * kx + gga * ex + gA1 * xkk + gA2 * xke + gA3 * xee;
    const double gky = 2.0 * gal * ky + gga * ey + gA1 * ykk + gA2 * yke + gA3 * yee;
    const double gex = gm[6] + 2.0 * gbt * ex + gga * kx + gB1 * xkk + gB2 * xke + gB3 * xee;
    const double gey = gm[7] + 2.0 * gbt * ey + gga * ky + gB1 * ykk + gB2 * yke + gB3 * yee;
    const double gjac = gm[0] + gkx * ye - gky * xe - gex * yk + gey * xk;
    const double gdet = -gjac * jac * jac;
    gD[0] = gey * jac + gdet * ye;     // x_xi
    gD[1] = -gex * jac - gdet * xe;    // y_xi
    gD[2] = -gky * jac - gdet * yk;    // x_eta
    gD[3] = gkx * jac + gdet * xk;     // y_eta
}

// eta-derivative stencils of column i at row j: first derivative weights w1
// and second derivative weights w2 on rows r[0..3] (vorti2d_core.f90:104-152).
__device__ __forceinline__ int eta_stencil(int j, int jmax, double deta, int *r, double *w1, double *w2)
{
    if (j == 0) {
        r[0] = 0; r[1] = 1; r[2] = 2; r[3] = 3;
        w1[0] = -3.0 / (2.0 * deta); w1[1] = 4.0 / (2.0 * deta); w1[2] = -1.0 / (2.0 * deta); w1[3] = 0.0;
    } else if (j == jmax - 1) {
        r[0] = j; r[1] = j - 1; r[2] = j - 2; r[3] = j - 3;
        w1[0] = 3.0 / (2.0 * deta); w1[1] = -4.0 / (2.0 * deta); w1[2] = 1.0 / (2.0 * deta); w1[3] = 0.0;
    } else {
        r[0] = j + 1; r[1] = j; r[2] = j - 1; r[3] = j;
        w1[0] = 1.0 / (2.0 * deta); w1[1] = 0.0; w1[2] = -1.0 / (2.0 * deta); w1[3] = 0.0;
        w2[0] = 1.0 / (deta * deta); w2[1] = -2.0 / (deta * deta); w2[2] = 1.0 / (deta * deta); w2[3] = 0.0;
        return 3;
    }
    const double h2 = deta * deta;
    w2[0] = 2.0 / h2; w2[1] = -5.0 / h2; w2[2] = 4.0 / h2; w2[3] = -1.0 / h2;
    return 4;
}

// Pass 1: per node, the local metric backward; the xi stencils and the eta
// stencils of the node scatter into g_x, g_y; the cross derivative
// x_xieta(k) = (x_eta(e) - x_eta(w)) / (2 dksi) scatters into g_eta (the
// gradient of x_eta / y_eta at the neighbours), finished by pass 2.
__global__ void __launch_bounds__(256) k_metrics_vjp_1(int imax, int jmax, double dksi, double deta,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       const double *__restrict__ g_m,
                                                       double *__restrict__ g_x, double *__restrict__ g_y,
                                                       double *__restrict__ g_eta)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    const Nbr b = neighbors(i, j, imax);
    const int ie = b.e - j * imax, iw = b.w - j * imax;
    int r[4];
    double w1[4], w2[4];
    const int ns = eta_stencil(j, jmax, deta, r, w1, w2);
    double D[10];
    D[0] = (x[b.e] - x[b.w]) / (2.0 * dksi);
    D[1] = (y[b.e] - y[b.w]) / (2.0 * dksi);
    D[4] = (x[b.e] - 2.0 * x[k] + x[b.w]) / (dksi * dksi);
    D[5] = (y[b.e] - 2.0 * y[k] + y[b.w]) / (dksi * dksi);
    D[2] = D[3] = D[6] = D[7] = 0.0;
    for (int t = 0; t < ns; ++t) {
        const int64_t c = (int64_t)r[t] * imax + i;
        D[2] += w1[t] * x[c]; D[3] += w1[t] * y[c];
        D[6] += w2[t] * x[c]; D[7] += w2[t] * y[c];
    }
    // cross derivatives from the eta derivatives at e and w
    {
        int re_[4], rw_[4];
        double a1[4], a2[4], c1[4], c2[4];
        const int ne_ = eta_stencil(j, jmax, deta, re_, a1, a2);
        const int nw_ = eta_stencil(j, jmax, deta, rw_, c1, c2);
        double xe_e = 0, ye_e = 0, xe_w = 0, ye_w = 0;
        for (int t = 0; t < ne_; ++t) {
            xe_e += a1[t] * x[(int64_t)re_[t] * imax + ie];
            ye_e += a1[t] * y[(int64_t)re_[t] * imax + ie];
        }
        for (int t = 0; t < nw_; ++t) {
            xe_w += c1[t] * x[(int64_t)rw_[t] * imax + iw];
            ye_w += c1[t] * y[(int64_t)rw_[t] * imax + iw];
        }
        D[8] = (xe_e - xe_w) / (2.0 * dksi);
        D[9] = (ye_e - ye_w) / (2.0 * dksi);
    }
    double gm[8];
    for (int q = 0; q < 8; ++q) gm[q] = g_m[q * N + k];
    double gD[10];
    metric_point_vjp(D, gm, gD);
    // xi stencils
    atomicAdd(&g_x[b.e], gD[0] / (2.0 * dksi) + gD[4] / (dksi * dksi));
    atomicAdd(&g_x[b.w], -gD[0] / (2.0 * dksi) + gD[4] / (dksi * dksi));
    atomicAdd(&g_x[k], -2.0 * gD[4] / (dksi * dksi) + g_m[8 * N + k]);
    atomicAdd(&g_y[b.e], gD[1] / (2.0 * dksi) + gD[5] / (dksi * dksi));
    atomicAdd(&g_y[b.w], -gD[1] / (2.0 * dksi) + gD[5] / (dksi * dksi));
    atomicAdd(&g_y[k], -2.0 * gD[5] / (dksi * dksi) + g_m[9 * N + k]);
    // eta stencils of this node (first and second derivative)
    for (int t = 0; t < ns; ++t) {
        const int64_t c = (int64_t)r[t] * imax + i;
        atomicAdd(&g_x[c], w1[t] * gD[2] + w2[t] * gD[6]);
        atomicAdd(&g_y[c], w1[t] * gD[3] + w2[t] * gD[7]);
    }
    // cross derivatives -> eta derivatives at e, w
    atomicAdd(&g_eta[b.e], gD[8] / (2.0 * dksi));
    atomicAdd(&g_eta[b.w], -gD[8] / (2.0 * dksi));
    atomicAdd(&g_eta[N + b.e], gD[9] / (2.0 * dksi));
    atomicAdd(&g_eta[N + b.w], -gD[9] / (2.0 * dksi));
}

// Pass 2: the gradient of the eta derivatives collected in pass 1 goes
// through each node's first-derivative eta stencil.
__global__ void __launch_bounds__(256) k_metrics_vjp_2(int imax, int jmax, double deta,
                                                       const double *__restrict__ g_eta,
                                                       double *__restrict__ g_x, double *__restrict__ g_y)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int j = (int)(k / imax), i = (int)(k - (int64_t)j * imax);
    int r[4];
    double w1[4], w2[4];
    const int ns = eta_stencil(j, jmax, deta, r, w1, w2);
    const double gx = g_eta[k], gy = g_eta[N + k];
    if (gx == 0.0 && gy == 0.0) return;
    for (int t = 0; t < ns; ++t) {
        const int64_t c = (int64_t)r[t] * imax + i;
        if (w1[t] != 0.0) {
            atomicAdd(&g_x[c], w1[t] * gx);
            atomicAdd(&g_y[c], w1[t] * gy);
        }
    }
}

}  // namespace v2d
