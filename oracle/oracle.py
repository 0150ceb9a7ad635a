"""CPU FP64 oracle for the vorti2d metric / assembly / Newton path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  The product package ``vorti2d_amd`` never imports it and has no CPU
fallback.

Contents (each restates the reference file:line given):

* ``C`` / ``REF``  -- ctypes views of two implementations of the Fortran core
  (``src/fortran/vorti2d_core.f90``) behind one Python API with f2py-style
  signatures (``vorti2d/__init__.py:13-15``):
    - ``C``   : ``oracle/v2d_oracle.c`` (the restatement, built by ``make -C oracle``)
    - ``REF`` : the reference Fortran compiled from its own sources by
                ``make -C oracle ref`` into ``oracle/_ref`` (present only where
                it was built).
* ``generate_cylinder``   <- ``vorti2d/mesh.py:14-50``
* ``generate_annulus``    -- closed-form geometric-stretching / uniform-theta
                             O-grid (no reference counterpart; a benchmark mesh)
* ``smooth_state``        <- ``tests/test_assemble_local.py:23-30``
* ``force_coeffs``        <- ``vorti2d/forces.py:82-194``
* ``newton_solve``        <- ``vorti2d/solver.py:85-145`` with the MUMPS LU of
                             ``vorti2d/petsc_solver.py:55-62,97-115`` replaced by
                             scipy SuperLU (as ``tests/test_assemble_local.py:52``
                             does).
"""
from __future__ import annotations

import ctypes as _C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libvorti2d_ref.so")

_p = _C.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_p)


class _Core:
    """f2py-signature facade over a C-ABI library of the five core routines.

    ``style='c'``: the oracle restatement (by-value scalars).
    ``style='fortran'``: the reference module (by-reference scalars, mangled
    names ``_QMvorti2d_coreP<name>``).
    """

    def __init__(self, path: str, style: str):
        self.path = path
        self.style = style
        self.lib = _C.CDLL(path)

    def _fn(self, name):
        if self.style == "c":
            return getattr(self.lib, "or_" + name)
        return getattr(self.lib, "_QMvorti2d_coreP" + name)

    def _i(self, v):
        return _C.c_int(int(v)) if self.style == "c" else _C.byref(_C.c_int(int(v)))

    def _d(self, v):
        return _C.c_double(float(v)) if self.style == "c" else _C.byref(_C.c_double(float(v)))

    # vorti2d_core.f90:33-61
    def node_neighbors(self, i, j, imax):
        if self.style == "c":
            out = (_C.c_int * 9)()
            self._fn("node_neighbors")(_C.c_int(i), _C.c_int(j), _C.c_int(imax), out)
            return tuple(out)
        o = [_C.c_int(0) for _ in range(9)]
        self._fn("node_neighbors")(self._i(i), self._i(j), self._i(imax),
                                   *[_C.byref(x) for x in o])
        return tuple(x.value for x in o)

    # vorti2d_core.f90:66-189
    def compute_metrics(self, dksi, deta, xg, yg):
        xg = np.asfortranarray(xg, dtype=np.float64)
        yg = np.asfortranarray(yg, dtype=np.float64)
        imax, jmax = xg.shape
        N = imax * jmax
        if self.style == "c":
            out = np.zeros((10, N))
            self._fn("compute_metrics")(_C.c_int(imax), _C.c_int(jmax), _C.c_double(dksi),
                                        _C.c_double(deta), _ptr(xg), _ptr(yg), _ptr(out))
            return tuple(np.ascontiguousarray(out[m]) for m in range(10))
        outs = [np.zeros(N) for _ in range(10)]
        self._fn("compute_metrics")(self._i(imax), self._i(jmax), self._d(dksi),
                                    self._d(deta), _ptr(xg), _ptr(yg),
                                    *[_ptr(o) for o in outs])
        return tuple(outs)

    # vorti2d_core.f90:219-383
    def assemble_coo(self, imax, jmax, re, invdtau, inv2dt, urot, dksi, deta,
                     jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys, yphys,
                     psi, ome, omeold, omeoldold, r0, r1, maxnnz, ff_bc, ca, sa):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys, yphys,
                 psi, ome, omeold, omeoldold)]
        ndof = imax * jmax
        for a in arrs:
            if a.size != ndof:
                raise ValueError("field arrays must have length imax*jmax")
        ci = np.zeros(maxnnz, np.int32)
        cj = np.zeros(maxnnz, np.int32)
        cv = np.zeros(maxnnz, np.float64)
        bvec = np.zeros(r1 - r0, np.float64)
        nnz = _C.c_int(0)
        if self.style == "c":
            rc = self._fn("assemble_coo")(
                _C.c_int(imax), _C.c_int(jmax), *[_C.c_double(v) for v in
                                                   (re, invdtau, inv2dt, urot, dksi, deta)],
                *[_ptr(a) for a in arrs], _C.c_int(r0), _C.c_int(r1), _C.c_int(maxnnz),
                _C.c_int(ff_bc), _C.c_double(ca), _C.c_double(sa),
                _ptr(ci), _ptr(cj), _ptr(cv), _C.byref(nnz), _ptr(bvec))
            if rc != 0:
                raise RuntimeError("assemble_coo: maxnnz too small")
        else:
            self._fn("assemble_coo")(
                self._i(imax), self._i(jmax), self._i(ndof),
                *[self._d(v) for v in (re, invdtau, inv2dt, urot, dksi, deta)],
                *[_ptr(a) for a in arrs], self._i(r0), self._i(r1), self._i(maxnnz),
                self._i(ff_bc), self._d(ca), self._d(sa),
                _ptr(ci), _ptr(cj), _ptr(cv), _C.byref(nnz), _ptr(bvec))
        return ci, cj, cv, int(nnz.value), bvec

    # vorti2d_core.f90:392-495
    def compute_metrics_local(self, il0, il1, dksi, deta, xgl, ygl):
        xgl = np.asfortranarray(xgl, dtype=np.float64)
        ygl = np.asfortranarray(ygl, dtype=np.float64)
        gxm, jmax = xgl.shape
        N = gxm * jmax
        if self.style == "c":
            out = np.zeros((10, N))
            self._fn("compute_metrics_local")(
                _C.c_int(gxm), _C.c_int(jmax), _C.c_int(il0), _C.c_int(il1),
                _C.c_double(dksi), _C.c_double(deta), _ptr(xgl), _ptr(ygl), _ptr(out))
            return tuple(np.ascontiguousarray(out[m]) for m in range(10))
        outs = [np.zeros(N) for _ in range(10)]
        self._fn("compute_metrics_local")(
            self._i(gxm), self._i(jmax), self._i(il0), self._i(il1), self._d(dksi),
            self._d(deta), _ptr(xgl), _ptr(ygl), *[_ptr(o) for o in outs])
        return tuple(outs)

    # vorti2d_core.f90:510-661
    def assemble_coo_local(self, gxm, jmax, il0, il1, re, invdtau, inv2dt, urot,
                           dksi, deta, jac, alfa, beta, gama, pmet, qmet, detadx,
                           detady, xphys, yphys, psi, ome, omeold, omeoldold,
                           r0, nloc, lg, maxnnz, ff_bc, ca, sa):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (jac, alfa, beta, gama, pmet, qmet, detadx, detady, xphys, yphys,
                 psi, ome, omeold, omeoldold)]
        lg = np.ascontiguousarray(lg, dtype=np.int32)
        ci = np.zeros(maxnnz, np.int32)
        cj = np.zeros(maxnnz, np.int32)
        cv = np.zeros(maxnnz, np.float64)
        bvec = np.zeros(nloc, np.float64)
        nnz = _C.c_int(0)
        if self.style == "c":
            rc = self._fn("assemble_coo_local")(
                _C.c_int(gxm), _C.c_int(jmax), _C.c_int(il0), _C.c_int(il1),
                *[_C.c_double(v) for v in (re, invdtau, inv2dt, urot, dksi, deta)],
                *[_ptr(a) for a in arrs], _C.c_int(r0), _C.c_int(nloc), _ptr(lg),
                _C.c_int(maxnnz), _C.c_int(ff_bc), _C.c_double(ca), _C.c_double(sa),
                _ptr(ci), _ptr(cj), _ptr(cv), _C.byref(nnz), _ptr(bvec))
            if rc != 0:
                raise RuntimeError("assemble_coo_local: maxnnz too small")
        else:
            self._fn("assemble_coo_local")(
                self._i(gxm), self._i(jmax), self._i(il0), self._i(il1),
                *[self._d(v) for v in (re, invdtau, inv2dt, urot, dksi, deta)],
                *[_ptr(a) for a in arrs], self._i(r0), self._i(nloc), _ptr(lg),
                self._i(maxnnz), self._i(ff_bc), self._d(ca), self._d(sa),
                _ptr(ci), _ptr(cj), _ptr(cv), _C.byref(nnz), _ptr(bvec))
        return ci, cj, cv, int(nnz.value), bvec


def load_c() -> _Core:
    if not os.path.exists(ORACLE_SO):
        raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
    return _Core(ORACLE_SO, "c")


def load_ref() -> _Core | None:
    """The reference Fortran built from its sources, or None where not built."""
    if not os.path.exists(REF_SO):
        return None
    return _Core(REF_SO, "fortran")


# ----------------------------------------------------------------- meshes --
def generate_cylinder(imax=181, jmax=181, inner_rad=0.5, outer_rad=50.0):
    """Reference cylinder O-grid (vorti2d/mesh.py:14-50): radial spacing growing
    linearly after three unit cells, circumferential weights min(i, imax-i)^0.6
    (0-based i), theta_i = 2 pi (1 - s_i).  Returns xg, yg of shape (imax, jmax)."""
    a = np.zeros(jmax)
    for j in range(1, min(jmax, 4)):
        a[j] = float(j)
    for j in range(4, jmax):
        a[j] = a[j - 1] + (j - 2)
    a /= a[jmax - 1]
    s = np.zeros(imax)
    for i in range(1, imax):
        s[i] = s[i - 1] + min(i, imax - i) ** 0.6
    s /= s[imax - 1]
    theta = 2.0 * np.pi * (1.0 - s)
    rad = inner_rad + (outer_rad - inner_rad) * a
    xg = np.empty((imax, jmax))
    yg = np.empty((imax, jmax))
    for i in range(imax):
        xg[i, :] = rad * np.cos(theta[i])
        yg[i, :] = rad * np.sin(theta[i])
    return xg, yg


def generate_annulus(imax, jmax, inner_rad=0.5, outer_rad=50.0):
    """Geometric radial stretching, uniform theta, seam duplicated: r_j = r_in *
    q^(j-1) with q = (r_out/r_in)^(1/(jmax-1)); theta_i = 2 pi (1 - (i-1)/(imax-1)),
    same orientation as the reference generator.  Closed form for every metric."""
    q = (outer_rad / inner_rad) ** (1.0 / (jmax - 1))
    rad = inner_rad * q ** np.arange(jmax)
    rad[-1] = outer_rad
    theta = 2.0 * np.pi * (1.0 - np.arange(imax) / (imax - 1))
    return np.outer(np.cos(theta), rad), np.outer(np.sin(theta), rad)


def smooth_state(xphys, yphys):
    """Seam-consistent nonzero state of tests/test_assemble_local.py:23-30."""
    r = np.sqrt(xphys ** 2 + yphys ** 2)
    psi = np.sin(0.3 * xphys) * np.cos(0.2 * yphys) / (1.0 + 0.01 * r)
    ome = np.cos(0.25 * xphys) * np.sin(0.15 * yphys) * np.exp(-0.02 * r)
    return psi, ome, 0.9 * ome, 0.8 * ome


# ------------------------------------------------------------------ forces --
FORCE_COLUMNS = ("t", "cd", "cl", "cm", "cdp", "cdf", "clp", "clf", "cmp", "cmf")


def force_coeffs(imax, jmax, re, dksi, deta, beta, gama, detadx, detady, xphys,
                 yphys, ome, ref_length=None, moment_ref=(0.0, 0.0)):
    """Wall-traction force coefficients (vorti2d/forces.py:82-194).  Returns a
    dict with cd, cl, cm, cdp, cdf, clp, clf, cmp, cmf, ref_length."""
    P = imax - 1
    w = np.arange(P)
    xw, yw, ow = xphys[w], yphys[w], ome[w]
    bw, gw = beta[w], gama[w]
    inv = 1.0 / np.sqrt(bw)
    nx, ny = detadx[w] * inv, detady[w] * inv
    tx, ty = ny, -nx
    oxi = (np.roll(ow, -1) - np.roll(ow, 1)) / (2.0 * dksi)
    oeta = (-3.0 * ow + 4.0 * ome[imax + w] - ome[2 * imax + w]) / (2.0 * deta)
    dwdn = (gw * oxi + bw * oeta) * inv
    seg = np.hypot(np.roll(xw, -1) - xw, np.roll(yw, -1) - yw)
    ds = 0.5 * (seg + np.roll(seg, 1))
    x0, y0 = moment_ref
    fxp = (1.0 / re) * np.sum(yw * dwdn * ds)
    fyp = -(1.0 / re) * np.sum(xw * dwdn * ds)
    fric = ow / re
    fxf = np.sum(-fric * tx * ds)
    fyf = np.sum(-fric * ty * ds)
    rx, ry = xw - x0, yw - y0
    mzp = -(1.0 / (2.0 * re)) * np.sum((rx ** 2 + ry ** 2) * dwdn * ds)
    mzf = np.sum((rx * (-fric * ty) - ry * (-fric * tx)) * ds)
    if ref_length is None:
        ws = ds.sum()
        cx, cy = np.sum(xw * ds) / ws, np.sum(yw * ds) / ws
        ref_length = 2.0 * np.sum(np.hypot(xw - cx, yw - cy) * ds) / ws
    d = ref_length
    out = dict(cdp=2 * fxp / d, cdf=2 * fxf / d, clp=2 * fyp / d, clf=2 * fyf / d,
               cmp=2 * mzp / d ** 2, cmf=2 * mzf / d ** 2)
    out.update(cd=out["cdp"] + out["cdf"], cl=out["clp"] + out["clf"],
               cm=out["cmp"] + out["cmf"], ref_length=d)
    return out


# ------------------------------------------------------------ Newton loop --
@dataclass
class NewtonResult:
    psi: np.ndarray
    ome: np.ndarray
    iters: int
    res: float           # ||delta||_2 / (2N) of the last step (solver.py:134)
    res_hist: list
    rnorm_hist: list     # ||bvec||_2 before each step


def initial_state(imax, jmax, xphys, yphys, ca=1.0, sa=0.0):
    """solver.py:103-114: zeros, far-field psi = ca*y - sa*x on j=jmax."""
    N = imax * jmax
    psi = np.zeros(N)
    ome = np.zeros(N)
    ff = slice(imax * (jmax - 1), N)
    psi[ff] = ca * yphys[ff] - sa * xphys[ff]
    return psi, ome


def newton_solve(core, imax, jmax, metrics, re, *, invdtau=0.0, inv2dt=0.0, urot=0.0,
                 ff_bc=0, ca=1.0, sa=0.0, psi=None, ome=None, omeold=None,
                 omeoldold=None, tol=1e-10, max_iter=50):
    """The inner pseudo-time / Newton loop (vorti2d/solver.py:117-145), one
    physical time level, with the exact LU solve done by scipy SuperLU."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    N = imax * jmax
    dksi, deta = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    if psi is None:
        psi, ome = initial_state(imax, jmax, metrics[8], metrics[9], ca, sa)
    psi, ome = psi.copy(), ome.copy()
    omeold = np.zeros(N) if omeold is None else omeold
    omeoldold = np.zeros(N) if omeoldold is None else omeoldold
    res, it = 1.0e5, 0
    hist, rhist = [], []
    while res > tol and it < max_iter:
        ci, cj, cv, nnz, b = core.assemble_coo(
            imax, jmax, re, invdtau, inv2dt, urot, dksi, deta, *metrics,
            psi, ome, omeold, omeoldold, 0, 2 * N, 2 * N * 13, ff_bc, ca, sa)
        A = sp.csc_matrix((cv[:nnz], (ci[:nnz], cj[:nnz])), shape=(2 * N, 2 * N))
        rhist.append(float(np.linalg.norm(b)))
        d = spla.spsolve(A, b)
        psi += d[:N]
        ome += d[N:]
        res = float(np.linalg.norm(d) / (2.0 * N))
        hist.append(res)
        it += 1
    return NewtonResult(psi, ome, it, res, hist, rhist)
