"""GPU tests of the batched sweep driver: concurrent contexts reproduce single
solves bit for bit in both precisions, and the batch adjoint dCd/dalpha
matches central differences."""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cases():
    from vorti2d_amd.batch import Case
    return [Case(65, 49, 0.6, 25.0, 4.0), Case(97, 57, 1.0, 40.0, -8.0), Case(81, 65, 0.75, 33.0, 0.0),
            Case(73, 41, 0.5, 45.0, 9.5), Case(65, 61, 0.9, 20.0, -3.0), Case(89, 53, 0.55, 30.0, 6.0)]


def test_batch_fp64_equals_single_solves():
    """The FP64 path has no timing-dependent decision: concurrent contexts
    reproduce the single solves bit for bit."""
    from vorti2d_amd.batch import solve_batch, solve_case
    cases = _cases()
    res, stats = solve_batch(cases, workers=4, precision="fp64")
    say(f"  fp64 batch of {len(cases)}: {stats}")
    for c, r in zip(cases, res):
        s = solve_case(c, precision="fp64")
        assert r.newton_iters == s.newton_iters
        assert r.cd == s.cd and r.cl == s.cl
        assert r.res <= 1e-10
        if c.alpha_deg == 0.0:
            assert abs(r.cl) < 1e-8


def test_batch_mixed_equals_single_solves():
    """The mixed path's refactorisation decision depends on the iterates only
    (the default "fixed" policy): concurrent contexts reproduce the single
    solves bit for bit."""
    from vorti2d_amd.batch import solve_batch, solve_case
    cases = _cases()
    res, stats = solve_batch(cases, workers=4)
    say(f"  mixed batch of {len(cases)}: {stats}")
    for c, r in zip(cases, res):
        s = solve_case(c)
        say(f"    {c}: cd {r.cd:.17g} vs {s.cd:.17g}, iters {r.newton_iters} vs {s.newton_iters}")
        assert r.res <= 1e-10 and s.res <= 1e-10
        assert r.newton_iters == s.newton_iters and r.history == s.history
        assert r.cd == s.cd and r.cl == s.cl
        if c.alpha_deg == 0.0:
            assert abs(r.cl) < 1e-8


def test_batch_fp64_matches_mixed():
    """The mixed-precision sweep converges to the FP64 solution (Cd within 1e-6)."""
    from vorti2d_amd.batch import Case, solve_case
    for c in (Case(65, 49, 0.6, 25.0, 4.0), Case(97, 57, 1.0, 40.0, -8.0)):
        m = solve_case(c, precision="mixed")
        f = solve_case(c, precision="fp64")
        assert abs(m.cd - f.cd) < 1e-6 and abs(m.cl - f.cl) < 1e-6


def test_batch_adjoint_dcd_dalpha_vs_fd():
    from vorti2d_amd.batch import Case, solve_case
    c = Case(81, 57, 0.7, 35.0, 5.0)
    r = solve_case(c, precision="mixed", tol=1e-13, adjoint=True)
    h = 1e-6
    cp = solve_case(Case(c.imax, c.jmax, c.aspect, c.re, c.alpha_deg + h), precision="fp64", tol=1e-13)
    cm = solve_case(Case(c.imax, c.jmax, c.aspect, c.re, c.alpha_deg - h), precision="fp64", tol=1e-13)
    fd = (cp.cd - cm.cd) / (2 * h)
    say(f"  dCd/dalpha fd {fd:.8e} adjoint {r.dcd_dalpha:.8e}")
    assert abs(fd - r.dcd_dalpha) <= 1e-5 * abs(fd)


def test_create_like_equals_fresh_context():
    """A context made by v2d_ctx_create_like (mesh, state and Jacobian pattern
    copied on the device, the multifrontal analysis shared) computes what a
    fresh context on the same mesh and state computes, bit for bit: the state
    itself, the FP64 force adjoint's scalars (the force VJP gathers each wall
    node's terms in a fixed order, no atomics) and a Newton step.  The node-coordinate gradients agree to 1e-12:
    the metric backward pass sums with FP64 atomics, so they are not
    bit-reproducible even between two calls on one context.  Outflow far
    field with a nonzero angle, so the outflow flags and their pattern rows
    are exercised; the source context is destroyed before the copy is used
    again."""
    from vorti2d_amd.batch import elliptic_grid
    import vorti2d_amd as v
    a = math.radians(7.0)
    xg, yg = elliptic_grid(97, 65, 0.7)
    src = v.DeviceContext(xg, yg, 1, math.cos(a), math.sin(a), "mixed")
    src.set_solver(refactor_every=1, gmres_rtol=1e-6, gmres_max_iter=30)
    src.init_state()
    for _ in range(3):
        src.newton_step(30.0)
    st = src.get_state()
    like = src.like("fp64")
    fresh = v.DeviceContext(xg, yg, 1, math.cos(a), math.sin(a), "fp64")
    fresh.set_state(*st)
    for p, q in zip(like.get_state(), st):
        assert np.array_equal(p, q)
    for p, q in zip(like.metrics(), fresh.metrics()):
        assert np.array_equal(p, q)
    say("create_like: adjoint")
    g1 = like.adjoint_forces(30.0, 1.0, 1.0, 0.5)
    g2 = fresh.adjoint_forces(30.0, 1.0, 1.0, 0.5)
    for k in ("cd", "cl", "dalpha", "dre"):
        assert g1[k] == g2[k], k
    for k in ("gx", "gy"):
        scale = np.abs(g2[k]).max()
        assert np.abs(g1[k] - g2[k]).max() <= 1e-12 * scale, k
    src.close()
    say("create_like: Newton step after the source is gone")
    r1 = like.newton_step(30.0)
    r2 = fresh.newton_step(30.0)
    assert r1 == r2
    for p, q in zip(like.get_state(), fresh.get_state()):
        assert np.array_equal(p, q)
    like.close()
    fresh.close()
