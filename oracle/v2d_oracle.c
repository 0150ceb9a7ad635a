/*
 * v2d_oracle.c -- CPU FP64 restatement of the vorti2d compute core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library (vorti2d_amd/csrc) never links it and has
 * no CPU fallback.
 *
 * It restates, from the equations, the algorithms of the reference Fortran
 * module src/fortran/vorti2d_core.f90:
 *   or_node_neighbors        <- node_neighbors          vorti2d_core.f90:33-61
 *   or_compute_metrics       <- compute_metrics         vorti2d_core.f90:66-189
 *   or_assemble_coo          <- assemble_coo (+ push)   vorti2d_core.f90:219-383, 664-674
 *   or_compute_metrics_local <- compute_metrics_local   vorti2d_core.f90:392-495
 *   or_assemble_coo_local    <- assemble_coo_local      vorti2d_core.f90:510-661
 *
 * Conventions (same as the reference): node pointer k = imax*(j-1)+i, 1-based,
 * i fastest; field arrays have length imax*jmax; the Newton system is
 * field-blocked (rows/cols [0,N) psi, [N,2N) omega), COO indices 0-based.
 * Mesh arrays xg/yg are (imax, jmax) in column-major order (xg[i + imax*j]).
 *
 * Pinning: see tests/test_oracle.py -- every function here is checked against
 * the reference Fortran built from its own sources (oracle/_ref), against the
 * committed golden fixtures written from it, and against closed forms
 * (annulus metrics, finite-difference Jacobian columns, pattern counts).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ a1 -- */
/* 1-based 9-point neighbour pointers with the O-grid seam wrap: columns i=1 and
 * i=imax are one physical line, so W of i=1 is i=imax-1 and E of i=imax is i=2
 * (vorti2d_core.f90:33-61).  No wrap in j. */
void or_node_neighbors(int i, int j, int imax, int out[9])
{
    int k = imax * (j - 1) + i;
    int de = 1, dw = -1;
    if (i == 1) dw = imax - 2;
    if (i == imax) de = -(imax - 2);
    int n = k + imax, s = k - imax;
    out[0] = k;
    out[1] = k + de;  out[2] = k + dw;
    out[3] = n;       out[4] = s;
    out[5] = n + de;  out[6] = n + dw;
    out[7] = s + de;  out[8] = s + dw;
}

/* ------------------------------------------------------------------ a2 -- */
/* Garmann metrics (vorti2d_core.f90:66-189).
 *   xi-derivatives : central with the seam wrap, at every j.
 *   eta-derivatives: central for 1<j<jmax; 3-point one-sided first and 4-point
 *                    one-sided second derivative at j=1 (forward) and j=jmax
 *                    (backward) (vorti2d_core.f90:118-152).
 *   cross          : xi-central difference of d(x,y)/deta (:155-161).
 *   jac = 1/(x_xi y_eta - x_eta y_xi); xi_x = jac y_eta; xi_y = -jac x_eta;
 *   eta_x = -jac y_xi; eta_y = jac x_xi; alfa=|grad xi|^2, beta=|grad eta|^2,
 *   gama = grad xi . grad eta; P and Q per :175-180.
 * out: 10 arrays of length N in order jac, alfa, beta, gama, pmet, qmet,
 * detadx, detady, xphys, yphys. */
static void eta_first(const double *v, int imax, int jmax, int i, int j,
                      double deta, double *d1, double *d2)
{
    /* v is flat k-1 indexed; i, j are 1-based */
#define V(jj) v[imax * ((jj) - 1) + (i) - 1]
    if (j == 1) {
        *d1 = (-3.0 * V(1) + 4.0 * V(2) - V(3)) / (2.0 * deta);
        *d2 = (2.0 * V(1) - 5.0 * V(2) + 4.0 * V(3) - V(4)) / (deta * deta);
    } else if (j == jmax) {
        *d1 = (3.0 * V(jmax) - 4.0 * V(jmax - 1) + V(jmax - 2)) / (2.0 * deta);
        *d2 = (2.0 * V(jmax) - 5.0 * V(jmax - 1) + 4.0 * V(jmax - 2) - V(jmax - 3))
              / (deta * deta);
    } else {
        *d1 = (V(j + 1) - V(j - 1)) / (2.0 * deta);
        *d2 = (V(j + 1) - 2.0 * V(j) + V(j - 1)) / (deta * deta);
    }
#undef V
}

static void metrics_point(double xxi, double yxi, double xeta, double yeta,
                          double xxx, double yxx, double xee, double yee,
                          double xxe, double yxe, double *o /* 8 */)
{
    double jac = 1.0 / (xxi * yeta - xeta * yxi);
    double kx = jac * yeta, ky = -jac * xeta;
    double ex = -jac * yxi, ey = jac * xxi;
    double a = kx * kx + ky * ky;
    double b = ex * ex + ey * ey;
    double g = kx * ex + ky * ey;
    double p = -(a * (xxx * kx + yxx * ky) + 2.0 * g * (xxe * kx + yxe * ky)
                 + b * (xee * kx + yee * ky));
    double q = -(a * (xxx * ex + yxx * ey) + 2.0 * g * (xxe * ex + yxe * ey)
                 + b * (xee * ex + yee * ey));
    o[0] = jac; o[1] = a; o[2] = b; o[3] = g; o[4] = p; o[5] = q; o[6] = ex; o[7] = ey;
}

void or_compute_metrics(int imax, int jmax, double dksi, double deta,
                        const double *xg, const double *yg, double *out)
{
    long N = (long)imax * jmax;
    /* xg is (imax,jmax) column-major, which is already k-1 order */
    const double *x = xg, *y = yg;
    double *xe = malloc(sizeof(double) * N), *ye = malloc(sizeof(double) * N);
    for (int j = 1; j <= jmax; ++j)
        for (int i = 1; i <= imax; ++i) {
            double dd;
            long k = (long)imax * (j - 1) + i - 1;
            eta_first(x, imax, jmax, i, j, deta, &xe[k], &dd);
            eta_first(y, imax, jmax, i, j, deta, &ye[k], &dd);
        }
    for (int j = 1; j <= jmax; ++j)
        for (int i = 1; i <= imax; ++i) {
            int nb[9];
            or_node_neighbors(i, j, imax, nb);
            long k = nb[0] - 1, e = nb[1] - 1, w = nb[2] - 1;
            double xxi = (x[e] - x[w]) / (2.0 * dksi);
            double yxi = (y[e] - y[w]) / (2.0 * dksi);
            double xxx = (x[e] - 2.0 * x[k] + x[w]) / (dksi * dksi);
            double yxx = (y[e] - 2.0 * y[k] + y[w]) / (dksi * dksi);
            double xeta, yeta, xee, yee;
            eta_first(x, imax, jmax, i, j, deta, &xeta, &xee);
            eta_first(y, imax, jmax, i, j, deta, &yeta, &yee);
            double xxe = (xe[e] - xe[w]) / (2.0 * dksi);
            double yxe = (ye[e] - ye[w]) / (2.0 * dksi);
            double o[8];
            metrics_point(xxi, yxi, xeta, yeta, xxx, yxx, xee, yee, xxe, yxe, o);
            for (int m = 0; m < 8; ++m) out[m * N + k] = o[m];
            out[8 * N + k] = x[k];
            out[9 * N + k] = y[k];
        }
    free(xe); free(ye);
}

/* ------------------------------------------------------------------ a3 -- */
/* One Newton-system row of the field-blocked system (vorti2d_core.f90:247-381
 * and Appendix-A stencils).  Returns the number of entries and fills cols[]
 * (0-based global columns) and vals[] in the reference emission order; *b gets
 * the residual-form right-hand side.  The neighbour / metric access goes through
 * a small view so the same row logic serves the replicated (seam-wrapped,
 * field-blocked) and the local (ghosted, node-interleaved) layouts. */
typedef struct {
    /* metric and state arrays, indexed by local 0-based node */
    const double *jac, *alfa, *beta, *gama, *pmet, *qmet, *detadx, *detady,
        *xphys, *yphys, *psi, *ome, *omeold, *omeoldold;
    double re, invdtau, inv2dt, urot, dksi, deta, ca, sa;
    int ff_bc;
} row_ctx;

/* node indices (0-based local) of the 9-point stencil plus the j+2 / j-2
 * nodes on the same column, and the dof mapper */
typedef struct {
    long k, e, w, n, s, ne, nw, se, sw, n2, n3, s1, s2;
} stencil_ix;

/* col(node, comp): comp 0 = psi, 1 = omega */
typedef long (*dof_fn)(const void *u, long node, int comp);

static int row_build(const row_ctx *c, int comp, int jtype /*0 wall,1 int,2 far*/,
                     const stencil_ix *s, dof_fn dof, const void *u,
                     long *cols, double *vals, double *b)
{
    const long k = s->k;
    const double re = c->re, dk = c->dksi, de = c->deta;
    int m = 0;
#define PUSH(node, comp, v) do { cols[m] = dof(u, (node), (comp)); vals[m] = (v); ++m; } while (0)
    if (comp == 0) {
        if (jtype == 2) {
            int outflow = c->ff_bc >= 2 && (c->ca * c->detadx[k] + c->sa * c->detady[k]) > 0.0;
            if (outflow) {
                PUSH(k, 0, 1.0); PUSH(s->s1, 0, -2.0); PUSH(s->s2, 0, 1.0);
                *b = -(c->psi[k] - 2.0 * c->psi[s->s1] + c->psi[s->s2]);
            } else {
                PUSH(k, 0, 1.0);
                *b = (c->ca * c->yphys[k] - c->sa * c->xphys[k]) - c->psi[k];
            }
        } else if (jtype == 0) {
            PUSH(k, 0, 1.0);
            *b = 0.0 - c->psi[k];
        } else {
            const double a = c->alfa[k], bt = c->beta[k], g = c->gama[k];
            const double p = c->pmet[k], q = c->qmet[k];
            const double cg = (2.0 * g) / (4.0 * re * dk * de);
            const double *f = c->psi;
            PUSH(k, 0, c->invdtau + ((2.0 * a) / (dk * dk) + (2.0 * bt) / (de * de)) / re);
            PUSH(s->e, 0, -(a / (re * (dk * dk)) + p / (2.0 * re * dk)));
            PUSH(s->w, 0, (-a / (re * (dk * dk)) + p / (2.0 * re * dk)));
            PUSH(s->n, 0, -(bt / (re * (de * de)) + q / (2.0 * re * de)));
            PUSH(s->s, 0, (-bt / (re * (de * de)) + q / (2.0 * re * de)));
            PUSH(s->ne, 0, -cg); PUSH(s->nw, 0, cg);
            PUSH(s->se, 0, cg);  PUSH(s->sw, 0, -cg);
            PUSH(k, 1, -1.0 / re);
            double t1 = (a / re) * ((f[s->e] - 2.0 * f[k] + f[s->w]) / (dk * dk));
            double t2 = 2.0 * (g / re) * ((f[s->ne] - f[s->nw] - f[s->se] + f[s->sw]) / (4.0 * de * dk));
            double t3 = (bt / re) * ((f[s->n] - 2.0 * f[k] + f[s->s]) / (de * de));
            double t4 = (p / re) * ((f[s->e] - f[s->w]) / (2.0 * dk));
            double t5 = (q / re) * ((f[s->n] - f[s->s]) / (2.0 * de));
            *b = t1 + t2 + t3 + t4 + t5 + c->ome[k] / re;
        }
    } else {
        if (jtype == 2) {
            int outflow = c->ff_bc >= 1 && (c->ca * c->detadx[k] + c->sa * c->detady[k]) > 0.0;
            if (outflow) {
                PUSH(k, 1, 1.0); PUSH(s->s1, 1, -4.0 / 3.0); PUSH(s->s2, 1, 1.0 / 3.0);
                *b = (4.0 * c->ome[s->s1] - c->ome[s->s2]) / 3.0 - c->ome[k];
            } else {
                PUSH(k, 1, 1.0);
                *b = 0.0 - c->ome[k];
            }
        } else if (jtype == 0) {
            const double bw = c->beta[k] / (de * de);
            PUSH(k, 1, -1.0);
            PUSH(k, 0, 3.5 * bw);
            PUSH(s->n2, 0, -4.0 * bw);
            PUSH(s->n3, 0, 0.5 * bw);
            const double r = sqrt(c->xphys[k] * c->xphys[k] + c->yphys[k] * c->yphys[k]);
            const double ct = c->xphys[k] / r, st = c->yphys[k] / r;
            *b = c->ome[k] - 3.5 * bw * c->psi[k] + 4.0 * bw * c->psi[s->n2]
                 - 0.5 * bw * c->psi[s->n3]
                 - ((c->qmet[k] - 3.0 * (c->beta[k] / de))
                    * (c->urot / (c->detadx[k] * ct + c->detady[k] * st)));
        } else {
            const double *P = c->psi, *O = c->ome;
            const double a = c->alfa[k], bt = c->beta[k], g = c->gama[k];
            const double p = c->pmet[k], q = c->qmet[k], J = c->jac[k];
            const double Aa = (P[s->n] - P[s->s]) / (2.0 * de);
            const double Bb = (O[s->e] - O[s->w]) / (2.0 * dk);
            const double Cc = (P[s->e] - P[s->w]) / (2.0 * dk);
            const double Dd = (O[s->n] - O[s->s]) / (2.0 * de);
            const double cab = a / (re * (dk * dk)), cbt = bt / (re * (de * de));
            const double cg = (2.0 * g) / (4.0 * re * dk * de);
            PUSH(k, 1, c->invdtau + (2.0 * a) / (re * (dk * dk)) + (2.0 * bt) / (re * (de * de))
                       + 3.0 * c->inv2dt);
            PUSH(s->e, 1, ((J * Aa) / (2.0 * dk)) - cab - p / (2.0 * re * dk));
            PUSH(s->w, 1, (-(J * Aa) / (2.0 * dk)) - cab + p / (2.0 * re * dk));
            PUSH(s->n, 1, (-(J * Cc) / (2.0 * de)) - cbt - q / (2.0 * re * de));
            PUSH(s->s, 1, ((J * Cc) / (2.0 * de)) - cbt + q / (2.0 * re * de));
            PUSH(s->ne, 1, -cg); PUSH(s->nw, 1, cg);
            PUSH(s->se, 1, cg);  PUSH(s->sw, 1, -cg);
            PUSH(s->n, 0, (J * Bb) / (2.0 * de));
            PUSH(s->s, 0, -(J * Bb) / (2.0 * de));
            PUSH(s->e, 0, -(J * Dd) / (2.0 * dk));
            PUSH(s->w, 0, (J * Dd) / (2.0 * dk));
            double e1 = (a / re) * ((O[s->e] - 2.0 * O[k] + O[s->w]) / (dk * dk));
            double e2 = 2.0 * (g / re) * ((O[s->ne] - O[s->nw] - O[s->se] + O[s->sw]) / (4.0 * de * dk));
            double e3 = (bt / re) * ((O[s->n] - 2.0 * O[k] + O[s->s]) / (de * de));
            double e4 = (p / re) * ((O[s->e] - O[s->w]) / (2.0 * dk));
            double e5 = (q / re) * ((O[s->n] - O[s->s]) / (2.0 * de));
            *b = -(J * Aa * Bb) + (J * Cc * Dd) + e1 + e2 + e3 + e4 + e5
                 - ((3.0 * O[k] - 4.0 * c->omeold[k] + c->omeoldold[k]) * c->inv2dt);
        }
    }
#undef PUSH
    return m;
}

typedef struct { long N; } blocked_u;
static long dof_blocked(const void *u, long node, int comp)
{
    return node + (comp ? ((const blocked_u *)u)->N : 0);
}

/* Replicated assembler (vorti2d_core.f90:219-383): rows [r0, r1) of the
 * field-blocked system, emitted row-major.  Returns 0 on success, -1 if the
 * entries would exceed maxnnz (the reference does not check; the oracle does). */
int or_assemble_coo(int imax, int jmax, double re, double invdtau, double inv2dt,
                    double urot, double dksi, double deta,
                    const double *jac, const double *alfa, const double *beta,
                    const double *gama, const double *pmet, const double *qmet,
                    const double *detadx, const double *detady,
                    const double *xphys, const double *yphys,
                    const double *psi, const double *ome, const double *omeold,
                    const double *omeoldold, int r0, int r1, int maxnnz, int ff_bc,
                    double ca, double sa, int *coo_i, int *coo_j, double *coo_v,
                    int *nnz, double *bvec)
{
    row_ctx c = {jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys, yphys,
                 psi, ome, omeold, omeoldold, re, invdtau, inv2dt, urot, dksi, deta,
                 ca, sa, ff_bc};
    const long N = (long)imax * jmax;
    blocked_u u = {N};
    long m = 0;
    for (int r = r0; r < r1; ++r) {
        int comp = r >= N;
        long node = comp ? r - N : r;
        int j = (int)(node / imax) + 1, i = (int)(node % imax) + 1;
        stencil_ix s;
        int nb[9];
        memset(&s, 0, sizeof s);
        s.k = node;
        if (j > 1 && j < jmax) {
            or_node_neighbors(i, j, imax, nb);
            s.e = nb[1] - 1; s.w = nb[2] - 1; s.n = nb[3] - 1; s.s = nb[4] - 1;
            s.ne = nb[5] - 1; s.nw = nb[6] - 1; s.se = nb[7] - 1; s.sw = nb[8] - 1;
        }
        s.n2 = node + imax; s.n3 = node + 2L * imax;
        s.s1 = node - imax; s.s2 = node - 2L * imax;
        int jt = (j == 1) ? 0 : (j == jmax ? 2 : 1);
        long cols[13];
        double vals[13], b;
        int cnt = row_build(&c, comp, jt, &s, dof_blocked, &u, cols, vals, &b);
        if (m + cnt > maxnnz) return -1;
        for (int q = 0; q < cnt; ++q) {
            coo_i[m] = r; coo_j[m] = (int)cols[q]; coo_v[m] = vals[q]; ++m;
        }
        bvec[r - r0] = b;
    }
    *nnz = (int)m;
    return 0;
}

/* ------------------------------------------------------------------ a11 - */
/* Local metrics on a ghosted (gxm x jmax) block (vorti2d_core.f90:392-495):
 * d/deta on every column (ghosts included), everything else on owned columns
 * [il0, il1] (1-based); no seam wrap (ghosts hold it).  Unowned entries are 0
 * except xphys/yphys which copy the whole block. */
void or_compute_metrics_local(int gxm, int jmax, int il0, int il1, double dksi,
                              double deta, const double *xgl, const double *ygl,
                              double *out)
{
    long N = (long)gxm * jmax;
    memset(out, 0, sizeof(double) * 8 * N);
    double *xe = malloc(sizeof(double) * N), *ye = malloc(sizeof(double) * N);
    for (int j = 1; j <= jmax; ++j)
        for (int i = 1; i <= gxm; ++i) {
            double dd;
            long k = (long)gxm * (j - 1) + i - 1;
            eta_first(xgl, gxm, jmax, i, j, deta, &xe[k], &dd);
            eta_first(ygl, gxm, jmax, i, j, deta, &ye[k], &dd);
        }
    for (int i = il0; i <= il1; ++i)
        for (int j = 1; j <= jmax; ++j) {
            long k = (long)gxm * (j - 1) + i - 1, e = k + 1, w = k - 1;
            const double *x = xgl, *y = ygl;
            double xxi = (x[e] - x[w]) / (2.0 * dksi);
            double yxi = (y[e] - y[w]) / (2.0 * dksi);
            double xxx = (x[e] - 2.0 * x[k] + x[w]) / (dksi * dksi);
            double yxx = (y[e] - 2.0 * y[k] + y[w]) / (dksi * dksi);
            double xeta, yeta, xee, yee;
            eta_first(x, gxm, jmax, i, j, deta, &xeta, &xee);
            eta_first(y, gxm, jmax, i, j, deta, &yeta, &yee);
            double xxe = (xe[e] - xe[w]) / (2.0 * dksi);
            double yxe = (ye[e] - ye[w]) / (2.0 * dksi);
            double o[8];
            metrics_point(xxi, yxi, xeta, yeta, xxx, yxx, xee, yee, xxe, yxe, o);
            for (int m = 0; m < 8; ++m) out[m * N + k] = o[m];
        }
    memcpy(out + 8 * N, xgl, sizeof(double) * N);
    memcpy(out + 9 * N, ygl, sizeof(double) * N);
    free(xe); free(ye);
}

typedef struct { const int *lg; } lg_u;
static long dof_lg(const void *u, long node, int comp)
{
    return ((const lg_u *)u)->lg[2 * node + comp];
}

/* Local assembler (vorti2d_core.f90:510-661): owned nodes of the ghosted block,
 * j outer / i inner, psi row then omega row per node; dofs through lg (0-based
 * globals, lg[2*(k-1)+c]).  bvec is indexed global_row - r0. */
int or_assemble_coo_local(int gxm, int jmax, int il0, int il1, double re,
                          double invdtau, double inv2dt, double urot, double dksi,
                          double deta, const double *jac, const double *alfa,
                          const double *beta, const double *gama, const double *pmet,
                          const double *qmet, const double *detadx,
                          const double *detady, const double *xphys,
                          const double *yphys, const double *psi, const double *ome,
                          const double *omeold, const double *omeoldold, int r0,
                          int nloc, const int *lg, int maxnnz, int ff_bc, double ca,
                          double sa, int *coo_i, int *coo_j, double *coo_v, int *nnz,
                          double *bvec)
{
    row_ctx c = {jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys, yphys,
                 psi, ome, omeold, omeoldold, re, invdtau, inv2dt, urot, dksi, deta,
                 ca, sa, ff_bc};
    lg_u u = {lg};
    long m = 0;
    for (int r = 0; r < nloc; ++r) bvec[r] = 0.0;
    for (int j = 1; j <= jmax; ++j)
        for (int i = il0; i <= il1; ++i) {
            long k = (long)gxm * (j - 1) + i - 1;
            stencil_ix s;
            memset(&s, 0, sizeof s);
            s.k = k;
            s.e = k + 1; s.w = k - 1; s.n = k + gxm; s.s = k - gxm;
            s.ne = s.n + 1; s.nw = s.n - 1; s.se = s.s + 1; s.sw = s.s - 1;
            s.n2 = k + gxm; s.n3 = k + 2L * gxm;
            s.s1 = k - gxm; s.s2 = k - 2L * gxm;
            int jt = (j == 1) ? 0 : (j == jmax ? 2 : 1);
            for (int comp = 0; comp < 2; ++comp) {
                long cols[13];
                double vals[13], b;
                int cnt = row_build(&c, comp, jt, &s, dof_lg, &u, cols, vals, &b);
                long row = lg[2 * k + comp];
                if (m + cnt > maxnnz) return -1;
                for (int q = 0; q < cnt; ++q) {
                    coo_i[m] = (int)row; coo_j[m] = (int)cols[q]; coo_v[m] = vals[q]; ++m;
                }
                if (row - r0 >= 0 && row - r0 < nloc) bvec[row - r0] = b;
            }
        }
    *nnz = (int)m;
    return 0;
}
