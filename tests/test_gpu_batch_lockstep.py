"""The lockstep batch (v2d_batch_*, vorti2d_amd.batch.DeviceBatch and
solve_batch_lockstep): cases of uneven mesh sizes stepped together with shared
assembly, SpMV, multi-dot / multi-axpy and update launches reproduce their
single-context solves (solve_case) bit for bit, in both precisions, whatever
the workgroup sizes the tuner picks."""
from __future__ import annotations

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


def _same(r, s):
    assert r.newton_iters == s.newton_iters, (r.case, r.newton_iters, s.newton_iters)
    assert r.history == s.history, r.case
    assert r.cd == s.cd and r.cl == s.cl and r.cm == s.cm, r.case
    assert r.rnorm == s.rnorm, r.case


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_lockstep_equals_single_solves(prec):
    from vorti2d_amd.batch import solve_batch_lockstep, solve_case
    cases = _cases()
    say(f"lockstep {prec}: {len(cases)} cases, 4 slots")
    res, stats = solve_batch_lockstep(cases, inflight=4, precision=prec)
    say(f"  {stats}")
    for c, r in zip(cases, res):
        _same(r, solve_case(c, precision=prec))


def test_lockstep_workgroup_sizes_do_not_change_results():
    from vorti2d_amd.batch import DeviceBatch, _case_context
    cases = _cases()[:3]
    out = []
    for blocks in ((128, 128, 128), (1024, 512, 1024)):
        b = DeviceBatch("mixed", 3)
        ctxs = [_case_context(c, "mixed", 1e-10, 1e-10, 0.1) for c in cases]
        for s, ctx in enumerate(ctxs):
            b.set_case(s, ctx)
        b.set_blocks(*blocks)
        assert tuple(b.blocks().values()) == blocks
        hist = [b.newton_step([c.re for c in cases]) for _ in range(4)]
        out.append((hist, [ctx.get_state(("psi", "ome")) for ctx in ctxs]))
        b.close()
        for ctx in ctxs:
            ctx.close()
    (h1, s1), (h2, s2) = out
    for a, b in zip(h1, h2):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    for a, b in zip(s1, s2):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_lockstep_autotune_and_checks():
    import vorti2d_amd as v
    from vorti2d_amd._lib import V2DError
    from vorti2d_amd.batch import DeviceBatch, _case_context
    cases = _cases()[:2]
    ctxs = [_case_context(c, "mixed", 1e-10, 1e-10, 0.1) for c in cases]
    b = DeviceBatch("mixed", 2)
    for s, ctx in enumerate(ctxs):
        b.set_case(s, ctx)
    tuned = b.autotune(30.0)
    say(f"  tuned workgroups {tuned}")
    assert all(t in (128, 256, 512, 1024) for t in tuned.values())
    with pytest.raises(V2DError):
        b.set_case(1, ctxs[0])                      # one context in two slots
    f64 = v.DeviceContext(*_grid(cases[0]), 0, 1.0, 0.0, "fp64")
    with pytest.raises(V2DError):
        b.set_case(0, f64)                          # precision mismatch
    ctxs[0].set_refactor_policy("measured")
    with pytest.raises(V2DError):
        b.newton_step([c.re for c in cases])       # timing-based policy refused
    b.close()
    for ctx in ctxs + [f64]:
        ctx.close()


def _grid(c):
    from vorti2d_amd.batch import elliptic_grid
    return elliptic_grid(c.imax, c.jmax, c.aspect)


def test_lockstep_c5_subset_equals_single():
    """The first 16 cases of the 64-case sweep (129..257 x 129..385), 8 slots,
    with the adjoint dCd/dalpha: equal to the single solves bit for bit."""
    from vorti2d_amd.batch import solve_batch_lockstep, solve_case, sweep_cases
    cases = sweep_cases(64, seed=0)[:16]
    res, stats = solve_batch_lockstep(cases, inflight=8, adjoint=True)
    say(f"  lockstep C5 subset: {stats['solves_per_s']:.2f} solves/s, blocks {stats['blocks']}")
    for c, r in zip(cases, res):
        s = solve_case(c, adjoint=True)
        _same(r, s)
        assert abs(r.dcd_dalpha - s.dcd_dalpha) <= 1e-12 * abs(s.dcd_dalpha)
