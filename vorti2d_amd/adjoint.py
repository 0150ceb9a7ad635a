"""Discrete adjoint of steady force coefficients behind torch.autograd.

`SteadyForces.apply(amplitudes, alpha_deg, re)` runs the steady FP64 Newton
solve of the device path on the airfoil grid of the given Hicks-Henne
amplitudes and returns the tensor (Cd, Cl); its backward solves one adjoint
system on the device (v2d_ctx_adjoint_forces: A^T lambda = dJ/dU with the
exact block LU, then lambda^T db/d(alpha, Re, metrics) and the metric backward
to the node coordinates) and maps the mesh gradient onto the amplitudes.

The reference computes no derivatives; its force coefficients
(vorti2d/forces.py:82-194, body axes) and its steady inner loop
(vorti2d/solver.py:136-145) are what is differentiated.  The solve is pure
Newton by default; optional switched-evolution pseudo-time steps use the
reference's dtau semantics (vorti2d/config.py:139-166: invdtau enters the
Newton matrix only).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, ptr
from .device import DeviceContext


@dataclass
class SteadyResult:
    ctx: DeviceContext
    cd: float
    cl: float
    iters: int
    res: float


def steady_solve(xg, yg, re, alpha_deg=0.0, ff_bc=0, ref_length=1.0, tol=1e-13, max_iter=200,
                 dtau0=0.0):
    """Steady FP64 solve on the device from the initial state of
    vorti2d/solver.py:103-114: pure Newton (dtau0 = 0, the reference's steady
    dtau = inf), or with switched-evolution relaxation (dtau_k = dtau0 ||R_0|| /
    ||R_k||, dropped once ||R_k|| < 1e-6 ||R_0||), until res = ||delta||_2 / (2N)
    <= tol.  Returns the context holding the converged state and its force
    coefficients."""
    a = math.radians(alpha_deg)
    ctx = DeviceContext(xg, yg, ff_bc, math.cos(a), math.sin(a), "fp64")
    ctx.init_state()
    ctx.assemble(re)
    r0 = ctx.rhs_norm()
    rn = r0
    res = 1.0
    it = 0
    while it < max_iter:
        invdtau = 0.0
        if dtau0 > 0.0 and rn > 1e-6 * r0:
            invdtau = 1.0 / (dtau0 * r0 / max(rn, 1e-300))
        res, rn, _ = ctx.newton_step(re, invdtau)
        it += 1
        if res <= tol and invdtau == 0.0:
            break
        if not math.isfinite(res):
            raise RuntimeError("steady solve diverged")
        ctx.assemble(re)
        rn = ctx.rhs_norm()
    if not (res <= tol and invdtau == 0.0):
        ctx.close()
        raise RuntimeError(f"steady solve: no convergence in {it} Newton steps (res {res:.3e} > "
                           f"{tol:g}); the adjoint needs a converged state")
    f = ctx.forces(re, ref_length)
    return SteadyResult(ctx, float(f[0]), float(f[1]), it, res)


def force_gradients(result: SteadyResult, re, ref_length=1.0, wd=1.0, wl=0.0):
    """Adjoint gradient of J = wd Cd + wl Cl at a converged steady state:
    dict(cd, cl, dalpha [per degree], dre, gx, gy [dJ/d node coordinates])."""
    return result.ctx.adjoint_forces(re, ref_length, wd, wl)


def metrics_vjp(dksi, deta, xg, yg, g_m):
    """Metric backward on the device: g_m (10 planes of N, the outputs of
    compute_metrics in order) -> (g_x, g_y) of shape (imax, jmax)."""
    xg = np.asarray(xg, dtype=np.float64)
    yg = np.asarray(yg, dtype=np.float64)
    imax, jmax = xg.shape
    N = imax * jmax
    xf = np.ascontiguousarray(xg.T)       # i fastest
    yf = np.ascontiguousarray(yg.T)
    gm = np.ascontiguousarray(np.asarray(g_m, dtype=np.float64).reshape(10, N))
    gx = np.empty(N)
    gy = np.empty(N)
    check(lib.v2d_metrics_vjp(imax, jmax, float(dksi), float(deta), ptr(xf), ptr(yf), ptr(gm),
                              ptr(gx), ptr(gy)))
    return gx.reshape(jmax, imax).T.copy(), gy.reshape(jmax, imax).T.copy()


def _torch():
    import torch
    return torch


class SteadyForces:
    """(Cd, Cl) of the steady flow past an AirfoilGrid as a differentiable
    function of (amplitudes, alpha_deg, re).  Use `SteadyForces.apply(a, alpha,
    re, grid)`; the grid is not differentiated."""

    @staticmethod
    def apply(amplitudes, alpha_deg, re, grid, ref_length=1.0, tol=1e-13):
        return _function().apply(amplitudes, alpha_deg, re, grid, ref_length, tol)


_FN = None


def _function():
    global _FN
    if _FN is not None:
        return _FN
    torch = _torch()

    class _SteadyForcesFn(torch.autograd.Function):
        @staticmethod
        def forward(fctx, amplitudes, alpha_deg, re, grid, ref_length, tol):
            a = amplitudes.detach().cpu().double().numpy()
            al = float(alpha_deg)
            rv = float(re)
            xg, yg = grid.grid(a)
            r = steady_solve(xg, yg, rv, al, ref_length=ref_length, tol=tol)
            fctx.result = r
            fctx.grid = grid
            fctx.scalars = (al, rv, ref_length)
            return torch.tensor([r.cd, r.cl], dtype=torch.float64)

        @staticmethod
        def backward(fctx, g):
            g = g.detach().cpu().double().numpy()
            al, rv, d = fctx.scalars
            adj = fctx.result.ctx.adjoint_forces(rv, d, float(g[0]), float(g[1]))
            ga = fctx.grid.bump_gradient(adj["gx"], adj["gy"])
            return (torch.from_numpy(ga), torch.tensor(adj["dalpha"], dtype=torch.float64),
                    torch.tensor(adj["dre"], dtype=torch.float64), None, None, None)

    _FN = _SteadyForcesFn
    return _FN
