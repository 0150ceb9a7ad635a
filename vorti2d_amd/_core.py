"""Drop-in replacement of the reference's compiled f2py module ``_core``.

``vorti2d_core`` carries the five routines of ``src/fortran/vorti2d_core.f90``
with the f2py calling conventions the reference Python uses
(``vorti2d/__init__.py:13-15``): optional trailing dimensions inferred from the
arrays, float64/int32 conversions of the inputs, and fresh NumPy outputs
returned as a tuple.  Every routine runs on the GPU through libvorti2d.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, f64c, lib, ptr, ptr_array


class vorti2d_core:  # noqa: N801  (module-like namespace, as f2py exposes it)
    """f2py-signature facade; see each method for the reference it replaces."""

    @staticmethod
    def node_neighbors(i, j, imax):
        """vorti2d_core.f90:33-61.  1-based (k, e, w, n, s, ne, nw, se, sw)."""
        out = np.zeros(9, np.int32)
        check(lib.v2d_node_neighbors(int(i), int(j), int(imax), ptr(out)))
        return tuple(int(v) for v in out)

    @staticmethod
    def compute_metrics(dksi, deta, xg, yg, imax=None, jmax=None):
        """vorti2d_core.f90:66-189 (called at vorti2d/solver.py:72-74).

        Returns (jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys, yphys),
        each of length imax*jmax in node order k = imax*j + i."""
        xg = np.asfortranarray(xg, dtype=np.float64)
        yg = np.asfortranarray(yg, dtype=np.float64)
        if xg.ndim != 2 or xg.shape != yg.shape:
            raise ValueError("xg and yg must be 2-D arrays of equal shape (imax, jmax)")
        imax = xg.shape[0] if imax is None else int(imax)
        jmax = xg.shape[1] if jmax is None else int(jmax)
        if (imax, jmax) != xg.shape:
            raise ValueError(f"imax, jmax = {imax}, {jmax} do not match xg.shape {xg.shape}")
        N = imax * jmax
        out = np.empty((10, N))
        check(lib.v2d_compute_metrics(imax, jmax, float(dksi), float(deta), ptr(xg), ptr(yg),
                                      ptr(out)))
        return tuple(np.ascontiguousarray(out[q]) for q in range(10))

    @staticmethod
    def assemble_coo(imax, jmax, re, invdtau, inv2dt, urot, dksi, deta, jac, alfa, beta,
                     gama, pmet, qmet, detadx, detady, xphys, yphys, psi, ome, omeold,
                     omeoldold, r0, r1, maxnnz, ff_bc, ca, sa, ndof=None):
        """vorti2d_core.f90:219-383 (called at vorti2d/solver.py:121-127).

        Returns (coo_i, coo_j, coo_v, nnz, bvec); the COO arrays have length
        maxnnz with the first nnz entries in the reference emission order."""
        arrs = [f64c(a) for a in (jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys,
                                  yphys, psi, ome, omeold, omeoldold)]
        ndof = arrs[0].size if ndof is None else int(ndof)
        if ndof != int(imax) * int(jmax) or any(a.size != ndof for a in arrs):
            raise ValueError("field arrays must all have length imax*jmax")
        r0, r1, maxnnz = int(r0), int(r1), int(maxnnz)
        ci = np.zeros(maxnnz, np.int32)
        cj = np.zeros(maxnnz, np.int32)
        cv = np.zeros(maxnnz, np.float64)
        bvec = np.zeros(max(r1 - r0, 0), np.float64)
        nnz = C.c_int(0)
        check(lib.v2d_assemble_coo(int(imax), int(jmax), float(re), float(invdtau), float(inv2dt),
                                   float(urot), float(dksi), float(deta), ptr_array(arrs[:10]),
                                   *[ptr(a) for a in arrs[10:]], r0, r1, maxnnz, int(ff_bc),
                                   float(ca), float(sa), ptr(ci), ptr(cj), ptr(cv), C.byref(nnz),
                                   ptr(bvec)))
        return ci, cj, cv, int(nnz.value), bvec

    @staticmethod
    def compute_metrics_local(il0, il1, dksi, deta, xgl, ygl, gxm=None, jmax=None):
        """vorti2d_core.f90:392-495 (called at vorti2d/dist_solver.py:92-93)."""
        xgl = np.asfortranarray(xgl, dtype=np.float64)
        ygl = np.asfortranarray(ygl, dtype=np.float64)
        gxm = xgl.shape[0] if gxm is None else int(gxm)
        jmax = xgl.shape[1] if jmax is None else int(jmax)
        if (gxm, jmax) != xgl.shape or xgl.shape != ygl.shape:
            raise ValueError("xgl, ygl must both have shape (gxm, jmax)")
        N = gxm * jmax
        out = np.empty((10, N))
        check(lib.v2d_compute_metrics_local(gxm, jmax, int(il0), int(il1), float(dksi),
                                            float(deta), ptr(xgl), ptr(ygl), ptr(out)))
        return tuple(np.ascontiguousarray(out[q]) for q in range(10))

    @staticmethod
    def assemble_coo_local(gxm, jmax, il0, il1, re, invdtau, inv2dt, urot, dksi, deta, jac,
                           alfa, beta, gama, pmet, qmet, detadx, detady, xphys, yphys, psi, ome,
                           omeold, omeoldold, r0, nloc, lg, maxnnz, ff_bc, ca, sa):
        """vorti2d_core.f90:510-661 (called at vorti2d/dist_solver.py:310-317)."""
        arrs = [f64c(a) for a in (jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys,
                                  yphys, psi, ome, omeold, omeoldold)]
        N = int(gxm) * int(jmax)
        if any(a.size != N for a in arrs):
            raise ValueError("field arrays must all have length gxm*jmax")
        lg = np.ascontiguousarray(lg, dtype=np.int32)
        if lg.size != 2 * N:
            raise ValueError("lg must have length 2*gxm*jmax")
        maxnnz, nloc = int(maxnnz), int(nloc)
        ci = np.zeros(maxnnz, np.int32)
        cj = np.zeros(maxnnz, np.int32)
        cv = np.zeros(maxnnz, np.float64)
        bvec = np.zeros(nloc, np.float64)
        nnz = C.c_int(0)
        check(lib.v2d_assemble_coo_local(int(gxm), int(jmax), int(il0), int(il1), float(re),
                                         float(invdtau), float(inv2dt), float(urot), float(dksi),
                                         float(deta), ptr_array(arrs[:10]),
                                         *[ptr(a) for a in arrs[10:]], int(r0), nloc, ptr(lg),
                                         maxnnz, int(ff_bc), float(ca), float(sa), ptr(ci),
                                         ptr(cj), ptr(cv), C.byref(nnz), ptr(bvec)))
        return ci, cj, cv, int(nnz.value), bvec


__all__ = ["vorti2d_core", "_lib"]
