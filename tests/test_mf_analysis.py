"""The multifrontal analysis (v2d_mf_analyse, host only) against the matrix it
factors.

The analysis is checked three ways on the oracle's Newton Jacobian of the
steady cylinder (vorti2d_core.f90 assemble_coo restated in oracle/), for the
three far-field conditions (Dirichlet, omega outflow, psi and omega outflow):

* structure: every unknown is the pivot of exactly one front, children sit
  lower in the tree than their parents, a child's coupling rows map onto
  parent rows holding the same unknowns, every Jacobian entry lands in one
  front position;
* numerics: a numpy restatement of what the device does with those maps --
  gather the children's contribution blocks, add the owned entries, sweep the
  pivots by Gauss-Jordan, then the forward / backward front solves -- solves
  A x = b and A^T x = b to the accuracy of scipy's sparse LU (relative
  difference 1e-10);
* cost: the dissection's factorisation flops grow like n^3 (n x n grid), not
  like the n^4 of a level-by-level block LU.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from vorti2d_amd import mf

F = {k: q for q, k in enumerate(mf.FRONT_FIELDS)}


def _system(C, oracle_mod, n, ff):
    xg, yg = oracle_mod.generate_cylinder(n, n)
    d = 1.0 / (n - 1)
    m = C.compute_metrics(d, d, xg, yg)
    psi, ome = oracle_mod.initial_state(n, n, m[8], m[9])
    r = oracle_mod.newton_solve(C, n, n, m, 40.0, tol=0.0, max_iter=1, ff_bc=ff)
    N = n * n
    z = np.zeros(N)
    ci, cj, cv, nnz, b = C.assemble_coo(n, n, 40.0, 0, 0, 0, d, d, *m, r.psi, r.ome, z, z, 0, 2 * N,
                                        26 * N, ff, 1.0, 0.0)
    ci, cj, cv = ci[:nnz], cj[:nnz], cv[:nnz]
    # slot layout: entry q of row (f, k) at value plane (10 f + q), node k;
    # the COO holds the rows in order, each row's entries in slot order
    cnt = np.bincount(ci, minlength=2 * N)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    q = np.arange(nnz) - start[ci]
    rows = ci
    plane = np.where(rows >= N, 10, 0) + q
    val = np.zeros(23 * N)
    val[plane * N + (rows % N)] = cv
    # far-field outflow bits: psi / omega rows of the last j with 3 entries
    last = N - n + np.arange(n)
    flags = ((cnt[N + last] == 3) * 1 + (cnt[last] == 3) * 2).astype(np.uint8)
    A = sp.csr_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    A.nnz_coo = nnz
    return A, np.asarray(b[:2 * N], dtype=np.float64), val, flags


def _sweep(Fm, npiv):
    """Partial Gauss-Jordan on pivots 0 .. npiv-1 (no pivoting)."""
    for p in range(npiv):
        ip = 1.0 / Fm[p, p]
        row = Fm[p, :] * ip
        col = Fm[:, p].copy()
        Fm -= np.outer(col, row)
        Fm[p, :] = row
        Fm[:, p] = -col * ip
        Fm[p, p] = ip


def _factor(a, val):
    T = a["table"]
    sums = np.add.reduceat(val[a["avid"]], a["aseg"][:-1])
    bounds = T[:, F["fo"]]
    fidx = np.searchsorted(bounds, a["aoff"], side="right") - 1
    mats = []
    for f in range(len(T)):
        fo, nf, ld, ns, nsp = (T[f, F[k]] for k in ("fo", "nf", "ld", "ns", "nsp"))
        Fm = np.zeros((nf, nf))
        Fm[np.arange(ns, nsp), np.arange(ns, nsp)] = 1.0
        for k in range(T[f, F["nch"]]):
            c = a["ch"][T[f, F["co"]] + k]
            m = a["map"][T[f, F["mo"]] + k * nf:][:nf]
            rows = np.nonzero(m >= 0)[0]
            nspc = T[c, F["nsp"]]
            Cc = mats[c][nspc:, nspc:]
            Fm[np.ix_(rows, rows)] += Cc[np.ix_(m[rows], m[rows])]
        sel = fidx == f
        off = a["aoff"][sel] - fo
        Fm[off // ld, off % ld] += sums[sel]
        _sweep(Fm, nsp)
        mats.append(Fm)
    return mats


def _solve(a, mats, b, trans=False):
    T = a["table"]
    fu = a["fu"]
    vf = []
    zf = []
    for f in range(len(T)):
        uo, nf, ns, nsp = (T[f, F[k]] for k in ("uo", "nf", "ns", "nsp"))
        u = fu[uo:uo + nf]
        v = np.zeros(nf)
        v[:ns] = b[u[:ns]]
        for k in range(T[f, F["nch"]]):
            c = a["ch"][T[f, F["co"]] + k]
            m = a["map"][T[f, F["mo"]] + k * nf:][:nf]
            rows = np.nonzero(m >= 0)[0]
            v[rows] += vf[c][T[c, F["nsp"]] + m[rows]]
        Fm = mats[f]
        if trans:
            w = Fm[:nsp, :].T @ v[:nsp]
            z = w[:nsp]
            v[nsp:] -= w[nsp:]
        else:
            w = Fm[:, :nsp] @ v[:nsp]
            z = w[:nsp]
            v[nsp:] += w[nsp:]
        vf.append(v)
        zf.append(z)
    x = np.zeros_like(b)
    for f in reversed(range(len(T))):
        uo, nf, ns, nsp = (T[f, F[k]] for k in ("uo", "nf", "ns", "nsp"))
        u = fu[uo:uo + nf]
        xb = x[u[nsp:]]
        Fm = mats[f]
        if trans:
            x[u[:ns]] = zf[f][:ns] + Fm[nsp:, :ns].T @ xb
        else:
            x[u[:ns]] = zf[f][:ns] - Fm[:ns, nsp:] @ xb
    return x


@pytest.mark.parametrize("ff", [0, 1, 2])
def test_structure(C, oracle_mod, ff):
    n = 33
    A, b, val, flags = _system(C, oracle_mod, n, ff)
    a = mf.analyse(n, n, flags)
    T = a["table"]
    N = n * n
    piv = np.concatenate([a["fu"][T[f, F["uo"]]:][:T[f, F["ns"]]] for f in range(len(T))])
    assert np.array_equal(np.sort(piv), np.arange(2 * N))
    h = T[:, F["height"]]
    assert np.all(np.diff(h) >= 0)
    for f in range(len(T)):
        nf, nsp = T[f, F["nf"]], T[f, F["nsp"]]
        u = a["fu"][T[f, F["uo"]]:][:nf]
        assert np.all(u[T[f, F["ns"]]:nsp] == -1) and np.all(u[nsp:] >= 0)
        if T[f, F["big"]]:
            assert nsp % 64 == 0
        for k in range(T[f, F["nch"]]):
            c = a["ch"][T[f, F["co"]] + k]
            assert h[c] < h[f]
            m = a["map"][T[f, F["mo"]] + k * nf:][:nf]
            uc = a["fu"][T[c, F["uo"]]:][:T[c, F["nf"]]]
            rows = np.nonzero(m >= 0)[0]
            assert len(rows) == T[c, F["nb"]]
            assert np.array_equal(uc[T[c, F["nsp"]] + m[rows]], u[rows])
    # every Jacobian entry once, each position inside one front
    assert len(a["avid"]) == A.nnz_coo and len(np.unique(a["avid"])) == A.nnz_coo
    assert np.all(np.diff(a["aoff"]) > 0)


@pytest.mark.parametrize("n,ff", [(33, 0), (33, 1), (41, 2), (65, 2)])
def test_numpy_restatement_solves(C, oracle_mod, n, ff):
    A, b, val, flags = _system(C, oracle_mod, n, ff)
    a = mf.analyse(n, n, flags)
    mats = _factor(a, val)
    lu = spla.splu(A.tocsc())
    x = _solve(a, mats, b)
    xr = lu.solve(b)
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)
    xt = _solve(a, mats, b, trans=True)
    xtr = lu.solve(b, trans="T")
    assert np.linalg.norm(xt - xtr) <= 1e-10 * np.linalg.norm(xtr)


def test_factorisation_flops_grow_like_n_cubed():
    fl = [mf.analyse(n, n)["flops"] for n in (65, 129, 257)]
    r1, r2 = fl[1] / fl[0], fl[2] / fl[1]
    print(f"multifrontal factorisation flops 65^2 {fl[0]:.3e}, 129^2 {fl[1]:.3e}, 257^2 {fl[2]:.3e}")
    assert 5.0 < r1 < 11.0 and 5.0 < r2 < 11.0   # n^3: 8x per doubling (n^4 would be 16x)
