"""Discrete adjoint of an unsteady BDF2 run, with checkpoint-recompute.

The forward run is the dual-time loop of vorti2d/solver.py:148-224 on one FP64
device context: at every physical level t_n = t_start + n dt (t_end
inclusive, :173-178) the inner Newton loop (:136-145) converges the residual of
the BDF2 step (vorti2d_core.f90:340-379, history omega^(n-1), omega^(n-2),
zero at the start), then the history shifts (:186-188); the rotation kick is
u_rot(t) of vorti2d/config.py:162-166.  The objective is a weighted sum of the
levels' force coefficients (vorti2d/forces.py:82-194, body axes, given
reference length): the final level's (objective="final") or the mean over all
levels (objective="mean").

Memory: the forward run writes a checkpoint every `checkpoint_every` levels in
the reference's restart format (vorti2d/restart.py:29-47: psi, omega and the
two history planes after a level's shift, t, step, imax, jmax, re) and keeps
nothing else.  The reverse sweep takes the segments between checkpoints from
the last to the first, recomputes each segment's levels from its checkpoint
(the same kernels on the same inputs, so bit for bit the forward states; the
end of every recomputed segment is checked against the next checkpoint), holds
that one segment in host memory and sweeps it backwards on the device
(v2d_ctx_adjoint_bdf2_step: per level one FP64 factorisation and one
transposed solve).  The last segment is kept from the forward run.

The reference computes no derivatives.
"""
from __future__ import annotations

import math
import os
import tempfile
from dataclasses import dataclass

import numpy as np

from . import restart as restartmod
from .config import _FF_CODES
from .device import DeviceContext


@dataclass
class BDF2Run:
    """An unsteady run's physical setting (field names as in Config)."""
    re: float
    alpha_deg: float = 0.0
    dt_phys: float = 0.2
    t_start: float = 0.0
    t_end: float = 4.0
    rot_speed: float = 0.0
    rot_until: float = 0.0
    farfield_bc: str = "dirichlet"
    pseudo_tol: float = 1e-12
    max_pseudo_iter: int = 100
    ref_length: float = 1.0

    def times(self):
        n = int(round((self.t_end - self.t_start) / self.dt_phys)) + 1
        return [self.t_start + m * self.dt_phys for m in range(max(n, 0))]

    def u_rot(self, t):
        return self.rot_speed if t <= self.rot_until else 0.0

    def inv2dt(self):
        return 1.0 / (2.0 * self.dt_phys)


def _context(xg, yg, run: BDF2Run):
    a = math.radians(run.alpha_deg)
    return DeviceContext(xg, yg, _FF_CODES[run.farfield_bc.lower()], math.cos(a), math.sin(a), "fp64")


def _solve_level(ctx, run: BDF2Run, t):
    """Newton loop of one physical level (vorti2d/solver.py:136-145)."""
    res, it = 1.0e5, 0
    while res > run.pseudo_tol and it < run.max_pseudo_iter:
        res, _, _ = ctx.newton_step(run.re, 0.0, run.inv2dt(), run.u_rot(t))
        it += 1
    if not res <= run.pseudo_tol:
        raise RuntimeError(f"level t={t:g}: no convergence in {it} Newton steps (res {res:.3e})")
    return it


def _weights(objective, nlev, wd, wl):
    if objective == "final":
        return [(0.0, 0.0)] * (nlev - 1) + [(wd, wl)]
    if objective == "mean":
        return [(wd / nlev, wl / nlev)] * nlev
    raise ValueError("objective must be 'final' or 'mean'")


def unsteady_forces(xg, yg, run: BDF2Run):
    """Forward run only: arrays t, Cd, Cl of every level and the final state
    (psi, omega)."""
    ctx = _context(xg, yg, run)
    try:
        ctx.init_state()
        ts, cd, cl = [], [], []
        for t in run.times():
            _solve_level(ctx, run, t)
            f = ctx.forces(run.re, run.ref_length)
            ts.append(t)
            cd.append(float(f[0]))
            cl.append(float(f[1]))
            ctx.bdf2_shift()
        psi, ome = ctx.get_state(("psi", "ome"))
    finally:
        ctx.close()
    return dict(t=np.array(ts), cd=np.array(cd), cl=np.array(cl), psi=psi, ome=ome)


def unsteady_forces_gradient(xg, yg, run: BDF2Run, *, wd=1.0, wl=0.0, objective="final",
                             checkpoint_every=8, checkpoint_dir=None):
    """J = sum_n w_n (wd Cd_n + wl Cl_n) of the run and its gradient with respect
    to the angle of attack (per degree), the Reynolds number and every node
    coordinate, by the discrete adjoint.  Returns dict(j, cd, cl (per level),
    dalpha, dre, gx, gy (imax x jmax), levels, checkpoints, recomputed)."""
    xg = np.asarray(xg, dtype=np.float64)
    imax, jmax = xg.shape
    times = run.times()
    nlev = len(times)
    if nlev < 1:
        raise ValueError("the run has no time level")
    K = max(1, int(checkpoint_every))
    w = _weights(objective, nlev, wd, wl)
    own_dir = checkpoint_dir is None
    cdir = tempfile.mkdtemp(prefix="v2d_ckpt_") if own_dir else checkpoint_dir
    os.makedirs(cdir, exist_ok=True)
    ckpt = lambda n: os.path.join(cdir, f"level_{n:06d}.npz")   # noqa: E731

    def save(ctx, n):
        psi, ome, o1, o2 = ctx.get_state()
        restartmod.save_restart(ckpt(n), restartmod.RestartState(
            psi, ome, o1, o2, times[n - 1] if n > 0 else run.t_start - run.dt_phys, n, imax,
            jmax, run.re))

    ctx = _context(xg, yg, run)
    try:
        # ---- forward, checkpoints before levels 0, K, 2K, ...
        ctx.init_state()
        cd, cl = np.empty(nlev), np.empty(nlev)
        starts = list(range(0, nlev, K))
        last = []
        for n, t in enumerate(times):
            if n % K == 0:
                save(ctx, n)
                last = []
            o1, o2 = ctx.get_state(("omeold", "omeoldold"))
            _solve_level(ctx, run, t)
            if n >= starts[-1]:
                psi, ome = ctx.get_state(("psi", "ome"))
                last.append((psi, ome, o1, o2))
            f = ctx.forces(run.re, run.ref_length)
            cd[n], cl[n] = f[0], f[1]
            ctx.bdf2_shift()
        # ---- reverse sweep, segment by segment
        ctx.adjoint_bdf2_begin()
        recomputed = 0
        for s in reversed(range(len(starts))):
            n0 = starts[s]
            n1 = min(n0 + K, nlev)
            if s == len(starts) - 1:
                seg = last
            else:
                st = restartmod.load_restart(ckpt(n0), imax, jmax)
                ctx.set_state(st.psi, st.ome, st.omeold, st.omeoldold)
                seg = []
                for n in range(n0, n1):
                    o1, o2 = ctx.get_state(("omeold", "omeoldold"))
                    _solve_level(ctx, run, times[n])
                    psi, ome = ctx.get_state(("psi", "ome"))
                    seg.append((psi, ome, o1, o2))
                    ctx.bdf2_shift()
                    recomputed += 1
                nxt = restartmod.load_restart(ckpt(n1), imax, jmax)
                if not (np.array_equal(seg[-1][0], nxt.psi) and np.array_equal(seg[-1][1], nxt.ome)):
                    raise RuntimeError(f"recomputed levels {n0}..{n1 - 1} differ from the forward run")
            for n in reversed(range(n0, n1)):
                ctx.set_state(*seg[n - n0])
                ctx.adjoint_bdf2_step(run.re, run.inv2dt(), run.u_rot(times[n]), run.ref_length,
                                      *w[n])
        g = ctx.adjoint_bdf2_end()
    finally:
        ctx.close()
        if own_dir:
            for f in os.listdir(cdir):
                os.remove(os.path.join(cdir, f))
            os.rmdir(cdir)
    j = float(sum(wn[0] * cd[n] + wn[1] * cl[n] for n, wn in enumerate(w)))
    return dict(j=j, cd=cd, cl=cl, levels=nlev, checkpoints=len(starts), recomputed=recomputed,
                **g)
