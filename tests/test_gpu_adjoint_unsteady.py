"""Discrete adjoint of an unsteady BDF2 run with checkpoint-recompute
(vorti2d_amd.adjoint_unsteady), against central differences of the forward run.

Run: generate_cylinder(41, 41), Re 100, dt 0.2, 20 levels (t = 0 .. 3.8),
rotation kick 0.5 until t = 2, outflow far field, alpha 3 deg (the C2 setting
of vorti2d/solver.py:148-224 on a small grid), FP64 Newton to res <= 1e-13 per
level.  Objectives: Cd at the last level and Cd averaged over all levels.
Derivatives: angle of attack, Reynolds number and four node coordinates (a
wall node and an interior node, x and y).  Bar: 1e-5 relative (the central
differences' own error is below 1e-7 at these steps).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NODES = ((7, 0), (12, 5))   # (i, j): wall node, interior node (neither on the seam)


def _run(v, **kw):
    base = dict(re=100.0, alpha_deg=3.0, dt_phys=0.2, t_start=0.0, t_end=3.8, rot_speed=0.5,
                rot_until=2.0, farfield_bc="outflow", pseudo_tol=1e-13, ref_length=1.0)
    base.update(kw)
    return v.BDF2Run(**base)


def _fd(v, xg, yg, run, field, h, node=None, comp=None):
    def f(sign):
        x, y, r = xg.copy(), yg.copy(), run
        if field == "alpha":
            r = _run(v, alpha_deg=run.alpha_deg + sign * h)
        elif field == "re":
            r = _run(v, re=run.re + sign * h)
        else:
            (x if comp == 0 else y)[node] += sign * h
        out = v.unsteady_forces(x, y, r)
        return out["cd"][-1], out["cd"].mean()
    (p_fin, p_mean), (m_fin, m_mean) = f(1.0), f(-1.0)
    return (p_fin - m_fin) / (2 * h), (p_mean - m_mean) / (2 * h)


def test_unsteady_adjoint_vs_central_differences():
    import vorti2d_amd as v
    xg, yg = v.generate_cylinder(41, 41, 0.5, 50.0)
    run = _run(v)
    print("unsteady adjoint: final-level Cd, checkpoints every 6 levels", flush=True)
    g_fin = v.unsteady_forces_gradient(xg, yg, run, objective="final", checkpoint_every=6)
    assert g_fin["levels"] == 20 and g_fin["checkpoints"] == 4 and g_fin["recomputed"] == 18
    print("unsteady adjoint: mean Cd", flush=True)
    g_mean = v.unsteady_forces_gradient(xg, yg, run, objective="mean", checkpoint_every=6)
    fwd = v.unsteady_forces(xg, yg, run)
    assert np.array_equal(fwd["cd"], g_fin["cd"])
    assert g_fin["j"] == fwd["cd"][-1] and abs(g_mean["j"] - fwd["cd"].mean()) < 1e-15
    checks = [("alpha", 1e-4, None, None, g_fin["dalpha"], g_mean["dalpha"]),
              ("re", 1e-3, None, None, g_fin["dre"], g_mean["dre"])]
    for node in NODES:
        for comp, key in ((0, "gx"), (1, "gy")):
            checks.append(("node", 1e-6, node, comp, g_fin[key][node], g_mean[key][node]))
    for field, h, node, comp, a_fin, a_mean in checks:
        fd_fin, fd_mean = _fd(v, xg, yg, run, field, h, node, comp)
        print(f"  d/d{field}{'' if node is None else (node, 'xy'[comp])}: final adjoint {a_fin:.9e} "
              f"fd {fd_fin:.9e}; mean adjoint {a_mean:.9e} fd {fd_mean:.9e}", flush=True)
        assert abs(a_fin - fd_fin) <= 1e-5 * abs(fd_fin) + 1e-9
        assert abs(a_mean - fd_mean) <= 1e-5 * abs(fd_mean) + 1e-9


def test_unsteady_adjoint_checkpointing_is_exact():
    """Recompute from checkpoints reproduces the forward levels bit for bit (the
    driver checks every recomputed segment against the next checkpoint), so
    the gradient does not depend on the checkpoint spacing beyond the order of
    the FP64 atomic sums of the force and metric backward kernels."""
    import vorti2d_amd as v
    xg, yg = v.generate_cylinder(41, 41, 0.5, 50.0)
    run = _run(v, t_end=2.2)
    a = v.unsteady_forces_gradient(xg, yg, run, checkpoint_every=3)
    b = v.unsteady_forces_gradient(xg, yg, run, checkpoint_every=100)
    assert a["recomputed"] == 9 and b["recomputed"] == 0
    for k in ("dalpha", "dre"):
        assert abs(a[k] - b[k]) <= 1e-13 * abs(b[k]), k
    for k in ("gx", "gy"):
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * np.abs(b[k]).max(), k
