// v2d_common.h -- shared host/device definitions for libvorti2d.
//
// Grid conventions follow the reference core (src/fortran/vorti2d_core.f90:10-19):
// node k = imax*j + i (0-based here; the reference is 1-based), i fastest,
// i circumferential with the seam stored twice (i = 0 and i = imax-1 are one
// physical line), j radial (j = 0 wall, j = jmax-1 far field).  The Newton
// system is field-blocked: unknown / row  f*N + k  with f = 0 (psi), 1 (omega).
//
// Jacobian storage ("slot layout"): every matrix row stores its entries in the
// reference's emission order (vorti2d_core.f90:249-379), so entry q of the psi
// row of node k lives at val[q*N + k] (q < 10) and entry q of the omega row at
// val[(10+q)*N + k] (q < 13).  Column indices are implied by the row type and
// the neighbour map below and are never stored.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define V2D_HD __host__ __device__ __forceinline__
#else
#define V2D_HD inline
#endif

namespace v2d {

constexpr int kPsiSlots = 10;
constexpr int kOmeSlots = 13;
constexpr int kSlots = kPsiSlots + kOmeSlots;   // 23 value planes

// outflow flag bits for far-field nodes (j = jmax-1)
constexpr uint8_t kOutOme = 1;   // ff_bc >= 1 and ca*eta_x + sa*eta_y > 0
constexpr uint8_t kOutPsi = 2;   // ff_bc >= 2 and the same test

struct Nbr {
    int k, e, w, n, s, ne, nw, se, sw;
};

// 9-point neighbour map with the seam wrap (vorti2d_core.f90:33-61), 0-based.
// West of i=0 is i=imax-2; east of i=imax-1 is i=1.  Valid for 0<j<jmax-1;
// for boundary j the n/s entries are not used.
V2D_HD Nbr neighbors(int i, int j, int imax)
{
    Nbr b;
    const int k = imax * j + i;
    const int de = (i == imax - 1) ? -(imax - 2) : 1;
    const int dw = (i == 0) ? (imax - 2) : -1;
    b.k = k;
    b.e = k + de;
    b.w = k + dw;
    b.n = k + imax;
    b.s = k - imax;
    b.ne = b.n + de;
    b.nw = b.n + dw;
    b.se = b.s + de;
    b.sw = b.s + dw;
    return b;
}

// Row structure: number of entries and their columns (field-blocked, 0-based),
// in emission order.  comp 0 = psi row, 1 = omega row.
V2D_HD int row_columns(int imax, int jmax, int i, int j, int comp, uint8_t flags,
                       int64_t N, int64_t *cols)
{
    const int k = imax * j + i;
    if (comp == 0) {
        if (j == jmax - 1) {
            if (flags & kOutPsi) {
                cols[0] = k; cols[1] = k - imax; cols[2] = k - 2 * imax;
                return 3;
            }
            cols[0] = k;
            return 1;
        }
        if (j == 0) { cols[0] = k; return 1; }
        const Nbr b = neighbors(i, j, imax);
        cols[0] = b.k; cols[1] = b.e; cols[2] = b.w; cols[3] = b.n; cols[4] = b.s;
        cols[5] = b.ne; cols[6] = b.nw; cols[7] = b.se; cols[8] = b.sw;
        cols[9] = N + b.k;
        return 10;
    }
    if (j == jmax - 1) {
        if (flags & kOutOme) {
            cols[0] = N + k; cols[1] = N + k - imax; cols[2] = N + k - 2 * imax;
            return 3;
        }
        cols[0] = N + k;
        return 1;
    }
    if (j == 0) {
        cols[0] = N + k; cols[1] = k; cols[2] = k + imax; cols[3] = k + 2 * imax;
        return 4;
    }
    const Nbr b = neighbors(i, j, imax);
    cols[0] = N + b.k; cols[1] = N + b.e; cols[2] = N + b.w; cols[3] = N + b.n;
    cols[4] = N + b.s; cols[5] = N + b.ne; cols[6] = N + b.nw; cols[7] = N + b.se;
    cols[8] = N + b.sw;
    cols[9] = b.n; cols[10] = b.s; cols[11] = b.e; cols[12] = b.w;
    return 13;
}

// Entry count of a row, without building the columns.
V2D_HD int row_count(int jmax, int j, int comp, uint8_t flags)
{
    if (j == jmax - 1) return (flags & (comp ? kOutOme : kOutPsi)) ? 3 : 1;
    if (j == 0) return comp ? 4 : 1;
    return comp ? 13 : 10;
}

// Offset of the first entry of row r (field-blocked) in the reference's COO
// emission order, given the exclusive prefix counts of outflow nodes on the
// far-field row (pre_psi[i], pre_ome[i] = number of flagged nodes i' < i).
V2D_HD int64_t row_start(int imax, int jmax, int64_t r, const int *pre_psi,
                         const int *pre_ome, int npsi_out)
{
    const int64_t N = (int64_t)imax * jmax;
    const int64_t inner = (int64_t)imax * (jmax - 2);
    if (r < N) {
        const int j = (int)(r / imax), i = (int)(r % imax);
        if (j == 0) return r;
        if (j < jmax - 1) return imax + 10 * (r - imax);
        return imax + 10 * inner + i + 2 * (int64_t)pre_psi[i];
    }
    const int64_t psi_total = 2 * (int64_t)imax + 10 * inner + 2 * (int64_t)npsi_out;
    const int64_t rr = r - N;
    const int j = (int)(rr / imax), i = (int)(rr % imax);
    if (j == 0) return psi_total + 4 * rr;
    if (j < jmax - 1) return psi_total + 4 * (int64_t)imax + 13 * (rr - imax);
    return psi_total + 4 * (int64_t)imax + 13 * inner + i + 2 * (int64_t)pre_ome[i];
}

// Block-tridiagonal level decomposition used by the direct solver: level 0 holds
// j in {0,1} (the wall vorticity row reaches j+2), the last level holds
// j in {jmax-2, jmax-1} (outflow rows reach j-2), every other level one j.
struct Levels {
    int imax, jmax, nlev;
    V2D_HD int level_of_j(int j) const
    {
        if (j <= 1) return 0;
        if (j >= jmax - 2) return nlev - 1;
        return j - 1;
    }
    V2D_HD int first_j(int l) const { return l == 0 ? 0 : (l == nlev - 1 ? jmax - 2 : l + 1); }
    V2D_HD int nj(int l) const { return (l == 0 || l == nlev - 1) ? 2 : 1; }
    V2D_HD int size(int l) const { return 2 * imax * nj(l); }
    // level-local index of global field-blocked unknown g
    V2D_HD int local(int64_t g, int64_t N, int l) const
    {
        const int f = g >= N;
        const int64_t node = f ? g - N : g;
        const int j = (int)(node / imax), i = (int)(node % imax);
        return f * imax * nj(l) + (j - first_j(l)) * imax + i;
    }
    V2D_HD int64_t global(int loc, int64_t N, int l) const
    {
        const int half = imax * nj(l);
        const int f = loc >= half;
        const int rem = f ? loc - half : loc;
        return (int64_t)f * N + (int64_t)first_j(l) * imax + rem;
    }
};

V2D_HD Levels make_levels(int imax, int jmax)
{
    Levels L;
    L.imax = imax;
    L.jmax = jmax;
    L.nlev = jmax - 2;
    return L;
}

}  // namespace v2d
