"""Pins the CPU oracle (oracle/v2d_oracle.c + oracle/oracle.py) before it is
trusted as the checker of the HIP path.

Each oracle function is checked against the golden fixtures written from the
reference Fortran, against the reference Fortran itself where it is built
(oracle/_ref), and against what the mathematics fixes: closed-form metrics of a
log-polar grid, the Jacobian as the exact derivative of the residual (central
differences of a quadratic), closed-form pattern counts, the boundary-condition
invariants of tests/test_smoke.py:50-56 and the force known-answers of
tests/test_postproc.py:26-37.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import golden


# ----------------------------------------------------------- node_neighbors --
def test_neighbors_golden_and_bruteforce(C):
    g = golden("neighbors_7x5")
    imax, jmax = int(g["imax"]), int(g["jmax"])
    for j in range(1, jmax + 1):
        for i in range(1, imax + 1):
            got = C.node_neighbors(i, j, imax)
            assert got == tuple(g["nb"][j - 1, i - 1])
            # brute force: pointer arithmetic with the seam identification
            k = imax * (j - 1) + i
            # seam: only i=1 (west) and i=imax (east) wrap, onto columns imax-1 and 2
            e_col = i + 1 if i < imax else 2
            w_col = i - 1 if i > 1 else imax - 1
            assert got[0] == k
            assert got[1] == imax * (j - 1) + e_col
            assert got[2] == imax * (j - 1) + w_col
            assert got[3] == k + imax and got[4] == k - imax
            assert got[5] == got[1] + imax and got[6] == got[2] + imax
            assert got[7] == got[1] - imax and got[8] == got[2] - imax


# ------------------------------------------------------------------ metrics --
@pytest.mark.parametrize("name", ["cyl_25x17", "cyl_41x41", "ann_33x21"])
def test_metrics_golden_bitexact(C, name):
    g = golden("metrics_" + name)
    m = C.compute_metrics(float(g["dksi"]), float(g["deta"]), g["xg"], g["yg"])
    for q in range(10):
        assert np.array_equal(m[q], g["metrics"][q]), q


def test_metrics_vs_reference_fortran(C, REF, oracle_mod):
    for gen, imax, jmax in ((oracle_mod.generate_cylinder, 61, 37),
                            (oracle_mod.generate_annulus, 48, 29)):
        xg, yg = gen(imax, jmax)
        rng = np.random.default_rng(3)
        xg = xg + 1e-4 * rng.standard_normal(xg.shape)   # generic (non-symmetric) grid
        yg = yg + 1e-4 * rng.standard_normal(yg.shape)
        dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
        a = C.compute_metrics(dk, de, xg, yg)
        b = REF.compute_metrics(dk, de, xg, yg)
        for q in range(10):
            assert np.array_equal(a[q], b[q]), q


def annulus_closed_form(imax, jmax, r_in=0.5, r_out=50.0):
    """Exact values of the discrete metrics on the log-polar grid (interior j):
    with d = 2 pi dxi, A = (q - 1/q)/(2 deta), S = sin(d)/dxi,
    Cx = (2cos d - 2)/dxi^2, B = (q - 2 + 1/q)/deta^2:
    jac = 1/(r^2 S A), alfa = jac^2 r^2 A^2, beta = jac^2 r^2 S^2, gama = 0,
    pmet = 0, qmet = -jac r^2 S (alfa Cx + beta B)."""
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    q = (r_out / r_in) ** de
    r = r_in * q ** np.arange(jmax)
    d = 2 * np.pi * dk
    A = (q - 1 / q) / (2 * de)
    S = np.sin(d) / dk
    Cx = (2 * np.cos(d) - 2) / dk ** 2
    B = (q - 2 + 1 / q) / de ** 2
    jac = 1.0 / (r ** 2 * S * A)
    alfa = jac ** 2 * r ** 2 * A ** 2
    beta = jac ** 2 * r ** 2 * S ** 2
    qmet = -jac * r ** 2 * S * (alfa * Cx + beta * B)
    return r, jac, alfa, beta, qmet


def test_metrics_closed_form_annulus(C, oracle_mod):
    imax, jmax = 65, 33
    xg, yg = oracle_mod.generate_annulus(imax, jmax)
    m = C.compute_metrics(1.0 / (imax - 1), 1.0 / (jmax - 1), xg, yg)
    r, jac, alfa, beta, qmet = annulus_closed_form(imax, jmax)
    sl = slice(1, jmax - 1)   # interior radial levels (central differences)
    for arr, ref in ((m[0], jac), (m[1], alfa), (m[2], beta), (m[5], qmet)):
        got = arr.reshape(jmax, imax)[sl]
        want = np.broadcast_to(ref[sl, None], got.shape)
        assert np.allclose(got, want, rtol=1e-10, atol=0), np.abs(got / want - 1).max()
    scale = np.abs(m[1]).max()
    assert np.abs(m[3]).max() < 1e-12 * scale              # gama = 0 (orthogonal grid)
    assert np.abs(m[4].reshape(jmax, imax)[sl]).max() < 1e-9 * np.abs(m[5]).max()   # pmet = 0
    # eta_x, eta_y: grad eta is radial with |grad eta|^2 = beta
    ex, ey = m[6], m[7]
    assert np.allclose(ex ** 2 + ey ** 2, m[2], rtol=1e-13)
    th = np.arctan2(m[9], m[8])
    assert np.allclose(ex * np.sin(th) - ey * np.cos(th), 0, atol=1e-10 * np.sqrt(m[2]).max())
    assert (m[0] > 0).all() and np.isfinite(np.array(m)).all()   # test_smoke.py:37-41


def test_metrics_seam_copies_equal(C, oracle_mod):
    imax, jmax = 41, 41
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    m = C.compute_metrics(1 / 40, 1 / 40, xg, yg)
    for q in range(8):
        a = m[q].reshape(jmax, imax)
        assert np.allclose(a[:, 0], a[:, -1], rtol=1e-9, atol=1e-9 * np.abs(a).max())


# ----------------------------------------------------------------- assembly --
def _assemble_case(C, name):
    g = golden("assemble_" + name)
    xg, yg = g["xg"], g["yg"]
    imax, jmax = xg.shape
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    m = C.compute_metrics(dk, de, xg, yg)
    N = imax * jmax
    out = C.assemble_coo(imax, jmax, float(g["re"]), float(g["invdtau"]), float(g["inv2dt"]),
                         float(g["urot"]), dk, de, *m, *g["state"], 0, 2 * N, 26 * N,
                         int(g["ff_bc"]), float(g["ca"]), float(g["sa"]))
    return g, out


@pytest.mark.parametrize("name", ["cyl_25x17_ff1", "cyl_41x41_ff0", "ann_33x21_ff2_a12"])
def test_assemble_golden_bitexact(C, name):
    g, (ci, cj, cv, nnz, b) = _assemble_case(C, name)
    assert nnz == g["coo_i"].size
    assert np.array_equal(ci[:nnz], g["coo_i"]) and np.array_equal(cj[:nnz], g["coo_j"])
    assert np.array_equal(cv[:nnz], g["coo_v"])
    assert np.array_equal(b, g["bvec"])


def test_assemble_vs_reference_fortran_row_ranges(C, REF, oracle_mod):
    imax, jmax = 37, 23
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    m = REF.compute_metrics(dk, de, xg, yg)
    N = imax * jmax
    rng = np.random.default_rng(7)
    st = [rng.standard_normal(N) for _ in range(4)]
    for (r0, r1) in ((0, 2 * N), (0, N), (N, 2 * N), (101, 2 * N - 71)):
        for ff in (0, 1, 2):
            args = (imax, jmax, 55.0, 0.3, 1.7, -0.4, dk, de, *m, *st, r0, r1, 13 * (r1 - r0),
                    ff, 0.8, 0.6)
            a, b = C.assemble_coo(*args), REF.assemble_coo(*args)
            assert a[3] == b[3]
            n = a[3]
            for q in range(3):
                assert np.array_equal(a[q][:n], b[q][:n])
            assert np.array_equal(a[4], b[4])


def pattern_count(imax, jmax, nout_psi=0, nout_ome=0):
    """Closed-form nnz: psi rows 1 / 10 / 1(+2), omega rows 4 / 13 / 1(+2)."""
    return 23 * imax * (jmax - 2) + 7 * imax + 2 * (nout_psi + nout_ome)


def test_pattern_counts(C, oracle_mod):
    for imax, jmax in ((25, 17), (181, 181), (257, 129)):
        xg, yg = oracle_mod.generate_cylinder(imax, jmax)
        dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
        m = C.compute_metrics(dk, de, xg, yg)
        N = imax * jmax
        z = np.zeros(N)
        out_arc = int(((m[6] + 0 * m[7])[N - imax:] > 0).sum())   # ca=1, sa=0
        for ff, npsi, nome in ((0, 0, 0), (1, 0, out_arc), (2, out_arc, out_arc)):
            *_, nnz, _ = C.assemble_coo(imax, jmax, 40.0, 0, 0, 0, dk, de, *m, z, z, z, z, 0,
                                        2 * N, 26 * N, ff, 1.0, 0.0)
            assert nnz == pattern_count(imax, jmax, npsi, nome)
    # the 1025^2 figure of the benchmark workload
    assert pattern_count(1025, 1025) == 24124400


def test_jacobian_is_exact_derivative(C, oracle_mod):
    """b is at most quadratic in (psi, ome), so the central difference
    -(b(U+h e) - b(U-h e))/(2h) reproduces every column of A to round-off
    (invdtau = 0: the pseudo-time shift has no RHS counterpart)."""
    imax, jmax = 13, 9
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    m = C.compute_metrics(dk, de, xg, yg)
    N = imax * jmax
    rng = np.random.default_rng(1)
    psi, ome = rng.standard_normal(N), rng.standard_normal(N)
    old, old2 = rng.standard_normal(N), rng.standard_normal(N)
    for ff, ca, sa in ((0, 1.0, 0.0), (2, 0.6, 0.8)):
        def asm(p, o):
            return C.assemble_coo(imax, jmax, 30.0, 0.0, 1.25, 0.7, dk, de, *m, p, o, old, old2,
                                  0, 2 * N, 26 * N, ff, ca, sa)
        ci, cj, cv, nnz, b0 = asm(psi, ome)
        A = sp.csr_matrix((cv[:nnz], (ci[:nnz], cj[:nnz])), shape=(2 * N, 2 * N)).toarray()
        h = 1e-3
        J = np.empty((2 * N, 2 * N))
        for c in range(2 * N):
            e = np.zeros(2 * N)
            e[c] = h
            bp = asm(psi + e[:N], ome + e[N:])[4]
            bm = asm(psi - e[:N], ome - e[N:])[4]
            J[:, c] = -(bp - bm) / (2 * h)
        scale = np.abs(A).max(axis=1, keepdims=True)
        assert np.abs(J - A).max() / 1 <= 1e-6 * scale.max()
        assert (np.abs(J - A) <= 1e-7 * scale + 1e-9).all()


def test_wall_and_farfield_rows_closed_form(C, oracle_mod):
    imax, jmax = 21, 15
    xg, yg = oracle_mod.generate_annulus(imax, jmax)
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    m = C.compute_metrics(dk, de, xg, yg)
    N = imax * jmax
    rng = np.random.default_rng(2)
    psi, ome = rng.standard_normal(N), rng.standard_normal(N)
    z = np.zeros(N)
    ca, sa = np.cos(0.3), np.sin(0.3)
    ci, cj, cv, nnz, b = C.assemble_coo(imax, jmax, 40.0, 0, 0, 0.0, dk, de, *m, psi, ome, z, z,
                                        0, 2 * N, 26 * N, 0, ca, sa)
    A = sp.csr_matrix((cv[:nnz], (ci[:nnz], cj[:nnz])), shape=(2 * N, 2 * N))
    wall = np.arange(imax)
    far = np.arange(N - imax, N)
    assert np.allclose(b[wall], -psi[wall])                                  # psi = 0
    assert np.allclose(b[far], ca * m[9][far] - sa * m[8][far] - psi[far])   # free stream
    assert np.allclose(b[N + far], -ome[far])                                 # ome = 0
    bw = m[2][wall] / de ** 2
    want = ome[wall] - 3.5 * bw * psi[wall] + 4 * bw * psi[wall + imax] - 0.5 * bw * psi[wall + 2 * imax]
    assert np.allclose(b[N + wall], want, rtol=1e-12)                         # Thom, urot = 0
    Ad = A.toarray()
    assert np.allclose(Ad[N + wall, N + wall], -1.0)
    assert np.allclose(Ad[N + wall, wall], 3.5 * bw)


# --------------------------------------------------------- local variants --
def test_local_variants_golden(C):
    g = golden("local_cyl_25x17")
    ml = C.compute_metrics_local(int(g["il0"]), int(g["il1"]), float(g["dksi"]),
                                 float(g["deta"]), g["xgl"], g["ygl"])
    for q in range(10):
        assert np.array_equal(ml[q], g["metrics_local"][q]), q
    gxm, jmax = g["xgl"].shape
    ci, cj, cv, nnz, b = C.assemble_coo_local(
        gxm, jmax, int(g["il0"]), int(g["il1"]), float(g["re"]), float(g["invdtau"]),
        float(g["inv2dt"]), float(g["urot"]), float(g["dksi"]), float(g["deta"]),
        *g["loc_metrics"], *g["state"], 0, int(g["nloc"]), g["lg"], 2 * int(g["nloc"]) * 13,
        int(g["ff_bc"]), float(g["ca"]), float(g["sa"]))
    assert nnz == g["coo_i"].size
    assert np.array_equal(ci[:nnz], g["coo_i"]) and np.array_equal(cj[:nnz], g["coo_j"])
    assert np.array_equal(cv[:nnz], g["coo_v"]) and np.array_equal(b, g["bvec"])


def test_local_equals_global_newton_update(C):
    """tests/test_assemble_local.py:33-99: the seam-collapsed node-interleaved
    system gives the same Newton update as the field-blocked one."""
    g = golden("local_cyl_25x17")
    gxm, jmax = g["xgl"].shape
    imax = gxm - 1
    ni = imax - 1
    N = imax * jmax
    ga = golden("assemble_cyl_25x17_ff1")
    A = sp.csr_matrix((ga["coo_v"], (ga["coo_i"], ga["coo_j"])), shape=(2 * N, 2 * N))
    x_old = spla.spsolve(A.tocsc(), ga["bvec"])
    nloc = int(g["nloc"])
    Al = sp.csr_matrix((g["coo_v"], (g["coo_i"], g["coo_j"])), shape=(nloc, nloc))
    x_new = spla.spsolve(Al.tocsc(), g["bvec"]).reshape(jmax, ni, 2)
    dp = x_old[:N].reshape(jmax, imax)[:, :ni]
    do = x_old[N:].reshape(jmax, imax)[:, :ni]
    tol = 1e-9 * max(np.abs(x_new).max(), 1.0)
    assert np.abs(dp - x_new[:, :, 0]).max() < tol
    assert np.abs(do - x_new[:, :, 1]).max() < tol


# ----------------------------------------------------- Newton loop, forces --
def test_newton_reproduces_reference_state(C, oracle_mod):
    g = golden("newton_41x41_re40")
    xg, yg = g["xg"], g["yg"]
    imax, jmax = xg.shape
    m = C.compute_metrics(1.0 / (imax - 1), 1.0 / (jmax - 1), xg, yg)
    r = oracle_mod.newton_solve(C, imax, jmax, m, 40.0, tol=1e-10)
    assert r.iters == int(g["iters"])
    assert np.abs(r.psi - g["psi"]).max() < 1e-12 * np.abs(g["psi"]).max()
    assert np.abs(r.ome - g["ome"]).max() < 1e-10 * np.abs(g["ome"]).max()
    N = imax * jmax
    ff = slice(imax * (jmax - 1), N)
    wall = slice(0, imax)
    assert np.max(np.abs(r.psi[ff] - m[9][ff])) < 1e-10        # test_smoke.py:52-56
    assert np.max(np.abs(r.psi[wall])) < 1e-10
    assert np.max(np.abs(r.ome[ff])) < 1e-12
    assert abs(r.psi.max() + r.psi.min()) < 1e-6
    fc = oracle_mod.force_coeffs(imax, jmax, 40.0, 1 / 40, 1 / 40, m[2], m[3], m[6], m[7], m[8],
                                 m[9], r.ome)
    want = dict(zip(("cd", "cl", "cm", "cdp", "cdf", "clp", "clf", "cmp", "cmf", "ref_length"),
                    g["forces"]))
    for k, v in want.items():
        assert abs(fc[k] - v) <= 1e-12 * max(1.0, abs(v)), k


def test_forces_uniform_flow_and_diameter(C, oracle_mod):
    imax, jmax = 61, 31
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    m = C.compute_metrics(1 / 60, 1 / 30, xg, yg)
    fc = oracle_mod.force_coeffs(imax, jmax, 40.0, 1 / 60, 1 / 30, m[2], m[3], m[6], m[7], m[8],
                                 m[9], np.zeros(imax * jmax))
    for k in ("cd", "cl", "cm", "cdp", "cdf", "clp", "clf", "cmp", "cmf"):
        assert fc[k] == 0.0
    assert abs(fc["ref_length"] - 1.0) < 1e-3   # coarse polygon of a unit-diameter circle


@pytest.mark.parametrize("re,cd_ref", [(20.0, 1.998), (40.0, 1.50)])
def test_forces_match_ingham(C, oracle_mod, re, cd_ref):
    """tests/test_postproc.py:26-37 on the oracle Newton solve (121x121)."""
    imax = jmax = 121
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    m = C.compute_metrics(1 / 120, 1 / 120, xg, yg)
    r = oracle_mod.newton_solve(C, imax, jmax, m, re, tol=1e-11, max_iter=80)
    fc = oracle_mod.force_coeffs(imax, jmax, re, 1 / 120, 1 / 120, m[2], m[3], m[6], m[7], m[8],
                                 m[9], r.ome)
    assert abs(fc["ref_length"] - 1.0) < 1e-6
    assert abs(fc["cl"]) < 1e-3 and abs(fc["cm"]) < 1e-3
    assert abs(fc["cd"] - cd_ref) / cd_ref < 0.05
    assert fc["cdp"] > 0 and fc["cdf"] > 0


def test_planted_mistake_is_caught(C, oracle_mod):
    """A checker that cannot fail pins nothing: a sign flip planted in a copy of
    the oracle's output must break the exact-derivative identity."""
    imax, jmax = 9, 7
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    dk, de = 1 / 8, 1 / 6
    m = C.compute_metrics(dk, de, xg, yg)
    N = imax * jmax
    rng = np.random.default_rng(5)
    psi, ome = rng.standard_normal(N), rng.standard_normal(N)
    z = np.zeros(N)
    ci, cj, cv, nnz, b = C.assemble_coo(imax, jmax, 30.0, 0, 0, 0, dk, de, *m, psi, ome, z, z, 0,
                                        2 * N, 26 * N, 0, 1.0, 0.0)
    r = N + imax * 3 + 4                 # an interior omega row
    row = np.nonzero(ci[:nnz] == r)[0]
    cv_bad = cv.copy()
    cv_bad[row[1]] = -cv_bad[row[1]]     # flip omega_e
    A_bad = sp.csr_matrix((cv_bad[:nnz], (ci[:nnz], cj[:nnz])), shape=(2 * N, 2 * N)).toarray()
    c = cj[row[1]]
    e = np.zeros(2 * N)
    e[c] = 1e-3
    bp = C.assemble_coo(imax, jmax, 30.0, 0, 0, 0, dk, de, *m, psi + e[:N], ome + e[N:], z, z, 0,
                        2 * N, 26 * N, 0, 1.0, 0.0)[4]
    bm = C.assemble_coo(imax, jmax, 30.0, 0, 0, 0, dk, de, *m, psi - e[:N], ome - e[N:], z, z, 0,
                        2 * N, 26 * N, 0, 1.0, 0.0)[4]
    fd = -(bp - bm) / 2e-3
    assert abs(fd[r] - A_bad[r, c]) > 1e-3 * abs(fd[r])
