"""The device multifrontal LU (V2D_SOLVER_MULTIFRONTAL) against scipy's sparse
LU and against the block-tridiagonal LU, on the Newton Jacobian the device
assembles (the steady cylinder, Re 40, after one Newton step from the
impulsive start).

* FP64 solves A x = b and A^T x = b on the exported matrix, for the three
  far-field conditions and grids from 7 x 6 and 20 x 4 to 181^2 (small fronts
  in LDS only on the small grids; dataflow Gauss-Jordan fronts from 97^2 on):
  normwise backward error ||A x - b|| / (||A||_F ||x|| + ||b||) below 1e-12
  (both directions), solutions within 1e-8 (relative 2-norm) of SuperLU's and
  of the block LU's.  On a random right-hand side the difference to SuperLU is
  SuperLU's own error as much as this solver's: at 181^2 SuperLU's transposed
  residual is about 1000x the device solve's;
* the paired Gauss-Jordan step schedule (light / heavy steps) and the
  unpaired one give the same solution up to rounding (1e-12 in FP64);
* the fronts of one 64-pivot step assembled and swept in one launch
  (k_mf_front1, the default) and through the level assembly, entry scatter
  and step launches give the same factored fronts (every entry of every
  front, 1e-12 of the largest in FP64, 1e-5 in FP32) and solutions;
* FP32 (the mixed path's preconditioner) within 1e-3 of SuperLU;
* at 513^2 (multi-step Gauss-Jordan fronts, up to 17 pivot blocks) the FP64
  residual of the Newton right-hand side is below 1e-8 and within 4x of the
  block LU's, and the two solutions agree within 1e-8;
* a steady FP64 Newton solve with it takes the block LU's steps and ends on
  its state within 1e-10; the mixed solve converges to the same state within
  1e-8.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx(v, n, ff, prec, solver):
    imax, jmax = n if isinstance(n, tuple) else (n, n)
    xg, yg = v.generate_cylinder(imax, jmax, 0.5, 50.0)
    ctx = v.DeviceContext(xg, yg, ff, 1.0, 0.0, prec)
    ctx.set_linear_solver(solver)
    ctx.init_state()
    return ctx


def _front_mask(n):
    """Entries of the front storage that hold a front's rows and columns
    (not the row padding nf .. ld-1), from the host analysis of the grid with
    no far-field outflow coupling."""
    from vorti2d_amd import mf
    imax, jmax = n if isinstance(n, tuple) else (n, n)
    a = mf.analyse(imax, jmax)
    t = dict(zip(mf.FRONT_FIELDS, a["table"].T))
    m = np.zeros(a["front_elems"], bool)
    for fo, nf, ld in zip(t["fo"], t["nf"], t["ld"]):
        m[fo:fo + nf * ld].reshape(nf, ld)[:, :nf] = True
    return m


def _fused_vs_steps(ctx, n, b, tol):
    """Factor with and without the one-step front fusion; the largest
    difference of the fronts (relative to the largest entry; ff 0 grids only)
    and of the solutions."""
    ctx.factor()
    F, x = ctx.mf_fronts(), ctx.solve(b)
    ctx.set_front_fusion(False)
    ctx.factor()
    Fu, xu = ctx.mf_fronts(), ctx.solve(b)
    ctx.set_front_fusion(True)
    df = 0.0
    if n is not None:
        m = _front_mask(n)
        assert m.size == F.size
        df = np.abs(F[m] - Fu[m]).max() / np.abs(Fu[m]).max()
    dx = np.linalg.norm(x - xu) / np.linalg.norm(xu)
    say(f"  one-step fronts fused vs by steps: fronts {df:.2e}, solutions {dx:.2e}")
    assert df <= tol and dx <= tol


def _system_after_one_step(ctx):
    ctx.newton_step(40.0)
    ctx.assemble(40.0)
    ip, ix, dv = ctx.export_csr()
    n2 = 2 * ctx.N
    return sp.csr_matrix((dv, ix, ip), shape=(n2, n2))


@pytest.mark.parametrize("n,ff", [(33, 0), (41, 1), (97, 2), (181, 0), ((7, 6), 0), ((20, 4), 1),
                                  ((57, 33), 2), ((130, 9), 2)])
def test_fp64_solves_match_superlu(n, ff):
    import vorti2d_amd as v
    say(f"multifrontal FP64 {n} ff {ff}")
    ctx = _ctx(v, n, ff, "fp64", "multifrontal")
    A = _system_after_one_step(ctx)
    ctx.factor()
    rng = np.random.default_rng(ff + 7)
    b = rng.standard_normal(A.shape[0])
    x = ctx.solve(b)
    xt = ctx.solve_transpose(b)
    lu = spla.splu(A.tocsc())
    xr, xtr = lu.solve(b), lu.solve(b, trans="T")
    e, et = (np.linalg.norm(p - q) / np.linalg.norm(q) for p, q in ((x, xr), (xt, xtr)))
    anorm, bnorm = spla.norm(A), np.linalg.norm(b)
    eta = np.linalg.norm(A @ x - b) / (anorm * np.linalg.norm(x) + bnorm)
    etat = np.linalg.norm(A.T @ xt - b) / (anorm * np.linalg.norm(xt) + bnorm)
    say(f"  solve {e:.2e}, transposed {et:.2e} (relative to SuperLU); backward errors {eta:.2e}, {etat:.2e}")
    assert eta <= 1e-12 and etat <= 1e-12
    assert e <= 1e-8 and et <= 1e-8
    ctx.set_step_pairing(False)
    ctx.factor()
    xu, xtu = ctx.solve(b), ctx.solve_transpose(b)
    ctx.set_step_pairing(True)
    d = max(np.linalg.norm(x - xu) / np.linalg.norm(xu), np.linalg.norm(xt - xtu) / np.linalg.norm(xtu))
    say(f"  paired vs unpaired steps {d:.2e}")
    assert d <= 1e-12
    _fused_vs_steps(ctx, n if ff == 0 else None, b, 1e-12)
    blu = ctx.like("fp64")
    blu.set_linear_solver("block_lu")
    blu.assemble(40.0)
    blu.factor()
    xb = blu.solve(b)
    assert np.linalg.norm(x - xb) <= 1e-8 * np.linalg.norm(xb)
    blu.close()
    ctx.close()


def test_fp32_solve_close_to_superlu():
    import vorti2d_amd as v
    ctx = _ctx(v, 97, 0, "mixed", "multifrontal")
    ctx.set_solver(refactor_every=1)
    A = _system_after_one_step(ctx)
    ctx.factor()
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    x = ctx.solve(b)
    xr = spla.spsolve(A.tocsc(), b)
    e = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    say(f"multifrontal FP32 97x97: {e:.2e} relative to SuperLU")
    assert e <= 1e-3
    _fused_vs_steps(ctx, 97, b, 1e-5)
    ctx.close()


def test_fp64_513_residual():
    """Residual of the Newton system's own right-hand side (a random one is
    far rougher: scipy's pivoted SuperLU leaves ||A x - b|| / ||b|| ~ 2e-8 on
    it at 257^2) against the block LU's on the same system."""
    import vorti2d_amd as v
    n = 513
    say(f"multifrontal FP64 {n}x{n}")
    ctx = _ctx(v, n, 0, "fp64", "multifrontal")
    ctx.newton_step(40.0)
    ctx.assemble(40.0)
    info = ctx.mf_info()
    say(f"  {info['fronts']} fronts, {info['heights']} heights, {info['big_fronts']} big, "
        f"{info['flops']:.3e} flop")
    b = ctx.export_coo()[3]
    ctx.factor()
    x = ctx.solve(b)
    r = np.linalg.norm(ctx.spmv(x) - b) / np.linalg.norm(b)
    blu = ctx.like("fp64")
    blu.set_linear_solver("block_lu")
    blu.assemble(40.0)
    blu.factor()
    xb = blu.solve(b)
    rb = np.linalg.norm(blu.spmv(xb) - b) / np.linalg.norm(b)
    say(f"  ||A x - b|| / ||b|| = {r:.2e} (block LU {rb:.2e}); |x - x_blu| / |x_blu| "
        f"{np.linalg.norm(x - xb) / np.linalg.norm(xb):.2e}")
    assert r <= 1e-8 and r <= 4.0 * rb + 1e-12
    assert np.linalg.norm(x - xb) <= 1e-8 * np.linalg.norm(xb)
    blu.close()
    ctx.close()


def test_newton_matches_block_lu():
    import vorti2d_amd as v
    n = 181
    out = {}
    for solver in ("block_lu", "multifrontal"):
        ctx = _ctx(v, n, 0, "fp64", solver)
        r = ctx.steady_newton(40.0, tol=1e-10, rms_tol=1e-10, max_iter=30, record=True)
        out[solver] = (r, *ctx.get_state(("psi", "ome")))
        say(f"  {solver}: {r.iters} steps, res {r.res:.2e}, ||R|| {r.rnorm:.2e}")
        ctx.close()
    (rb, pb, ob), (rm, pm, om) = out["block_lu"], out["multifrontal"]
    assert rb.converged and rm.converged and rb.iters == rm.iters
    assert np.abs(pm - pb).max() <= 1e-10 * max(1.0, np.abs(pb).max())
    assert np.abs(om - ob).max() <= 1e-10 * max(1.0, np.abs(ob).max())
    ctx = _ctx(v, n, 0, "mixed", "multifrontal")
    ctx.set_solver(refactor_every=0, gmres_rtol=1e-6, gmres_max_iter=30)
    ctx.set_forcing(0.25, 1e-10 * np.sqrt(2 * ctx.N))
    r = ctx.steady_newton(40.0, tol=1e-10, rms_tol=1e-10, max_iter=30, record=True)
    p, o = ctx.get_state(("psi", "ome"))
    say(f"  mixed multifrontal: {r.iters} steps, GMRES {sum(r.lin_iters)}, res {r.res:.2e}")
    assert r.converged
    assert np.abs(p - pb).max() <= 1e-8 and np.abs(o - ob).max() <= 1e-8
    ctx.close()
