"""GPU tests of the discrete adjoint (force coefficients vs angle of attack,
Reynolds number and mesh / Hicks-Henne bump amplitudes).

The reference has no adjoint; every piece is pinned by central finite
differences of the forward function it differentiates (the residual and
metrics of the oracle / device path, both pinned to the reference Fortran),
and the whole gradient by central differences of complete steady solves.
"""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def v():
    import vorti2d_amd
    return vorti2d_amd


def _ctx(v, oracle_mod, C, imax, jmax, kind, prec, ff=0, alpha=0.0, seed=3):
    gen = oracle_mod.generate_annulus if kind == "ann" else oracle_mod.generate_cylinder
    xg, yg = gen(imax, jmax)
    a = math.radians(alpha)
    ctx = v.DeviceContext(xg, yg, ff, math.cos(a), math.sin(a), prec)
    m = C.compute_metrics(1.0 / (imax - 1), 1.0 / (jmax - 1), xg, yg)
    psi, ome, o1, o2 = oracle_mod.smooth_state(m[8], m[9])
    psi = psi + 0.01 * np.random.default_rng(seed).standard_normal(imax * jmax)
    ctx.set_state(psi, ome, o1, o2)
    return ctx, xg, yg, m, (psi, ome, o1, o2)


# ----------------------------------------------------------- A^T solve ----
@pytest.mark.parametrize("imax,jmax,kind,ff", [(41, 41, "cyl", 0), (33, 8, "ann", 2), (25, 5, "cyl", 1),
                                               (57, 33, "ann", 1)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_solve_transpose(v, C, oracle_mod, imax, jmax, kind, ff, prec):
    say(f"transpose solve {prec} {kind} {imax}x{jmax} ff={ff}")
    ctx, *_ = _ctx(v, oracle_mod, C, imax, jmax, kind, prec, ff)
    ctx.assemble(40.0, 0.0, 2.5, 0.1)
    ci, cj, cv, _ = ctx.export_coo()
    N = imax * jmax
    A = sp.csc_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    b = np.random.default_rng(5).standard_normal(2 * N)
    ctx.factor()
    x = ctx.solve_transpose(b)
    xs = spla.spsolve(A.T.tocsc(), b)
    if prec == "fp64":
        assert np.linalg.norm(A.T @ x - b) <= 1e-11 * np.linalg.norm(b)
        assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
    else:   # FP32 factorisation: error of the same size as the forward FP32 solve
        xf = ctx.solve(b)
        xfs = spla.spsolve(A, b)
        ef = np.linalg.norm(xf - xfs) / np.linalg.norm(xfs)
        et = np.linalg.norm(x - xs) / np.linalg.norm(xs)
        assert et <= max(10 * ef, 1e-4), (et, ef)


# ------------------------------------------------------- residual VJP -----
@pytest.mark.parametrize("ff,alpha,urot", [(0, 0.0, 0.0), (1, 12.0, 0.3), (2, -7.0, 0.0)])
def test_residual_vjp_matches_fd(v, C, oracle_mod, ff, alpha, urot):
    imax, jmax = 33, 21
    ctx, xg, yg, m, st = _ctx(v, oracle_mod, C, imax, jmax, "cyl", "fp64", ff, alpha)
    N = imax * jmax
    rng = np.random.default_rng(11)
    lam = rng.standard_normal(2 * N)
    re = 70.0
    gm, gre, gal = ctx.residual_vjp(lam, re, 0.0, urot)
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    a = math.radians(alpha)

    def b_of(mm, re_, a_):
        *_, b = C.assemble_coo(imax, jmax, re_, 0.0, 0.0, urot, dk, de, *mm, *st, 0, 2 * N, 26 * N, ff,
                               math.cos(a_), math.sin(a_))
        return b

    dm = [rng.standard_normal(N) * np.abs(mq).mean() for mq in m]
    h = 1e-6
    bp = b_of([mq + h * d for mq, d in zip(m, dm)], re, a)
    bm = b_of([mq - h * d for mq, d in zip(m, dm)], re, a)
    fd = lam @ (bp - bm) / (2 * h)
    an = sum(float(gm[q] @ dm[q]) for q in range(10))
    say(f"  metrics: fd {fd:.10e} adjoint {an:.10e}")
    assert abs(fd - an) <= 1e-7 * max(abs(fd), 1e-3 * np.abs(lam).sum())
    hr = 1e-4
    fdr = lam @ (b_of(m, re + hr, a) - b_of(m, re - hr, a)) / (2 * hr)
    assert abs(fdr - gre) <= 1e-6 * max(abs(fdr), 1e-8)
    ha = 1e-6
    fda = lam @ (b_of(m, re, a + ha) - b_of(m, re, a - ha)) / (2 * ha)
    assert abs(fda - gal) <= 1e-6 * max(abs(fda), 1e-8)


# -------------------------------------------------------- metric VJP ------
@pytest.mark.parametrize("grid", ["cyl", "airfoil"])
def test_metrics_vjp_matches_fd(v, grid):
    if grid == "cyl":
        xg, yg = v.generate_cylinder(25, 17, 0.5, 50.0)
    else:
        xg, yg = v.AirfoilGrid(65, 33, n_bumps=8).grid(np.full(8, 1e-3))
    imax, jmax = xg.shape
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    rng = np.random.default_rng(2)
    m0 = v.core.compute_metrics(dk, de, xg, yg)
    gm = np.stack([rng.standard_normal(imax * jmax) / (np.abs(mq).mean() + 1e-30) for mq in m0])
    gx, gy = v.metrics_vjp(dk, de, xg, yg, gm)
    dx = rng.standard_normal(xg.shape) * 1e-3
    dy = rng.standard_normal(xg.shape) * 1e-3
    dx[-1] = dx[0]                 # the seam copies move together
    dy[-1] = dy[0]
    h = 1e-5
    mp = v.core.compute_metrics(dk, de, xg + h * dx, yg + h * dy)
    mm = v.core.compute_metrics(dk, de, xg - h * dx, yg - h * dy)
    fd = sum(float(gm[q] @ (mp[q] - mm[q])) for q in range(10)) / (2 * h)
    an = float((gx * dx).sum() + (gy * dy).sum())
    say(f"  {grid}: fd {fd:.10e} adjoint {an:.10e}")
    assert abs(fd - an) <= 1e-6 * abs(fd)


# --------------------------------------------------------- force VJP ------
@pytest.mark.parametrize("wd,wl", [(1.0, 0.0), (0.0, 1.0), (0.7, -0.4)])
def test_forces_vjp_matches_fd(v, C, oracle_mod, wd, wl):
    imax, jmax = 41, 21
    ctx, xg, yg, m, st = _ctx(v, oracle_mod, C, imax, jmax, "cyl", "fp64")
    N = imax * jmax
    re, d = 60.0, 1.0
    go, gm, out = ctx.forces_vjp(re, d, wd, wl)
    ome = st[1]
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)

    def J(mm, om, re_):
        f = oracle_mod.force_coeffs(imax, jmax, re_, dk, de, mm[2], mm[3], mm[6], mm[7], mm[8], mm[9],
                                    om, ref_length=d)
        return wd * f["cd"] + wl * f["cl"]

    f0 = oracle_mod.force_coeffs(imax, jmax, re, dk, de, m[2], m[3], m[6], m[7], m[8], m[9], ome,
                                 ref_length=d)
    assert abs(out[0] - f0["cd"]) <= 1e-12 * abs(f0["cd"]) + 1e-14
    assert abs(out[1] - f0["cl"]) <= 1e-12 * max(abs(f0["cl"]), 1e-3)
    rng = np.random.default_rng(4)
    dm = [rng.standard_normal(N) * np.abs(mq).mean() * 1e-2 for mq in m]
    dw = rng.standard_normal(N) * np.abs(ome).mean()
    h = 1e-6
    fd = (J([mq + h * x for mq, x in zip(m, dm)], ome + h * dw, re)
          - J([mq - h * x for mq, x in zip(m, dm)], ome - h * dw, re)) / (2 * h)
    an = float(go @ dw) + sum(float(gm[q] @ dm[q]) for q in range(10))
    say(f"  wd {wd} wl {wl}: fd {fd:.10e} adjoint {an:.10e}")
    assert abs(fd - an) <= 1e-6 * abs(fd)
    hr = 1e-4
    fdr = (J(m, ome, re + hr) - J(m, ome, re - hr)) / (2 * hr)
    assert abs(fdr - out[2]) <= 1e-6 * abs(fdr) + 1e-12


# ------------------------------------------------- end-to-end gradient ----
def _fd_check(v, grid, amps, alpha, re, n_bumps_checked, tol, label):
    from vorti2d_amd.adjoint import steady_solve
    xg, yg = grid.grid(amps)
    r0 = steady_solve(xg, yg, re, alpha)
    say(f"  {label}: Cd {r0.cd:.8f} Cl {r0.cl:.8f} in {r0.iters} iterations (res {r0.res:.1e})")
    grads = {k: r0.ctx.adjoint_forces(re, 1.0, *w) for k, w in (("cd", (1.0, 0.0)), ("cl", (0.0, 1.0)))}

    def solve(a=amps, al=alpha, rv=re):
        r = steady_solve(*grid.grid(a), rv, al)
        return np.array([r.cd, r.cl])

    h = 1e-6
    worst = 0.0
    checks = [("alpha", lambda s: solve(al=alpha + s), "dalpha"),
              ("re", lambda s: solve(rv=re * (1 + s)), "dre")]
    for name, f, key in checks:
        fd = (f(h) - f(-h)) / (2 * h)
        if name == "re":
            fd = fd / re
        for c, q in enumerate(("cd", "cl")):
            an = grads[q][key]
            err = abs(fd[c] - an) / max(abs(fd[c]), 1e-8)
            worst = max(worst, err)
            say(f"    d{q}/d{name}: fd {fd[c]: .8e} adjoint {an: .8e} rel {err:.1e}")
    ga = {q: grid.bump_gradient(grads[q]["gx"], grads[q]["gy"]) for q in ("cd", "cl")}
    for k in range(n_bumps_checked):
        e = np.zeros_like(amps)
        e[k] = h
        fd = (solve(a=amps + e) - solve(a=amps - e)) / (2 * h)
        for c, q in enumerate(("cd", "cl")):
            an = ga[q][k]
            err = abs(fd[c] - an) / max(abs(fd[c]), 1e-6)
            worst = max(worst, err)
            say(f"    d{q}/da_{k}: fd {fd[c]: .8e} adjoint {an: .8e} rel {err:.1e}")
    assert worst <= tol, worst
    return grads, ga


def test_adjoint_small_airfoil_vs_fd(v):
    """65x33 NACA 2412 O-grid, Re 100, alpha 5 deg, 8 bumps."""
    grid = v.AirfoilGrid(65, 33, n_bumps=8)
    amps = np.random.default_rng(0).normal(0.0, 1e-3, 8)
    _fd_check(v, grid, amps, 5.0, 100.0, 8, 1e-4, "65x33 Re 100")


def test_adjoint_torch_autograd(v):
    """The torch.autograd.Function returns the same gradients as the adjoint."""
    import torch
    grid = v.AirfoilGrid(65, 33, n_bumps=8)
    amps = np.random.default_rng(0).normal(0.0, 1e-3, 8)
    a = torch.tensor(amps, dtype=torch.float64, requires_grad=True)
    al = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
    re = torch.tensor(100.0, dtype=torch.float64, requires_grad=True)
    out = v.SteadyForces.apply(a, al, re, grid)
    loss = 2.0 * out[0] - 3.0 * out[1]
    loss.backward()
    from vorti2d_amd.adjoint import steady_solve
    r = steady_solve(*grid.grid(amps), 100.0, 5.0)
    adj = r.ctx.adjoint_forces(100.0, 1.0, 2.0, -3.0)
    assert abs(al.grad.item() - adj["dalpha"]) <= 1e-9 * abs(adj["dalpha"])
    assert abs(re.grad.item() - adj["dre"]) <= 1e-9 * abs(adj["dre"])
    ga = grid.bump_gradient(adj["gx"], adj["gy"])
    assert np.allclose(a.grad.numpy(), ga, rtol=1e-9, atol=1e-12)


def test_adjoint_c4_naca_257x129_re500(v):
    """The adjoint workload: NACA 4-digit (t 0.12, m 0.02, p 0.4) O-grid
    257x129 to radius 30, Re 500, alpha 5 deg, 32 Hicks-Henne amplitudes
    ~ N(0, 1e-3) (numpy default_rng(0)); adjoint dCd, dCl w.r.t. alpha, Re and
    every amplitude against central differences with step 1e-6."""
    grid = v.AirfoilGrid(257, 129, n_bumps=32)
    amps = np.random.default_rng(0).normal(0.0, 1e-3, 32)
    _fd_check(v, grid, amps, 5.0, 500.0, 32, 1e-4, "257x129 Re 500")
