"""The batched-sweep workload itself: 64 steady solves on elliptic-body O-grids
of uneven sizes (numpy default_rng(0): angular 129..257, radial 129..385,
aspect U[0.5, 1], Re U[20, 45], alpha U[-10, 10] deg), mixed precision, with
the batch adjoint dCd/dalpha, 8 contexts in flight.

* Every case meets res = ||delta||_2 / (2N) <= 1e-10 (vorti2d/solver.py:141-144)
  and ||R||_2 / sqrt(2N) < 1e-10.
* The smallest, the largest, the most elongated and the steepest-angle cases:
  the batch result equals the single mixed solve bit for bit, and the mixed
  state is within 1e-8 (psi, omega) and 1e-6 (Cd, Cl, Cm) of the FP64 solve.
* The smallest case (about 129 x 129 nodes) against the oracle's FP64 Newton
  solve with exact sparse LU steps: psi, omega within 1e-8, forces within 1e-6.
* dCd/dalpha of the largest case against central differences of FP64 solves.
"""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sweep():
    from vorti2d_amd.batch import solve_batch, sweep_cases
    cases = sweep_cases(64, seed=0)
    say("C5 sweep: 64 cases, 8 streams, mixed + adjoint")
    res, stats = solve_batch(cases, workers=8, adjoint=True)
    say(f"  {stats['solves_per_s']:.2f} solves/s, p50 {stats['p50_s'] * 1e3:.0f} ms, "
        f"p99 {stats['p99_s'] * 1e3:.0f} ms")
    return cases, res


def _picks(cases):
    size = [c.imax * c.jmax for c in cases]
    idx = {"smallest": int(np.argmin(size)), "largest": int(np.argmax(size)),
           "elongated": int(np.argmin([c.aspect for c in cases])),
           "steepest": int(np.argmax([abs(c.alpha_deg) for c in cases]))}
    return idx


def test_c5_every_case_converges(sweep):
    cases, res = sweep
    assert len(res) == 64
    for c, r in zip(cases, res):
        assert r.res <= 1e-10, c
        assert r.rnorm / math.sqrt(2 * c.imax * c.jmax) < 1e-10, c
        assert math.isfinite(r.dcd_dalpha), c
    assert min(c.imax for c in cases) >= 129 and max(c.imax for c in cases) <= 257
    assert min(c.jmax for c in cases) >= 129 and max(c.jmax for c in cases) <= 385


def test_c5_extreme_cases_single_and_fp64(sweep):
    from vorti2d_amd.batch import solve_case
    cases, res = sweep
    for name, i in _picks(cases).items():
        c, r = cases[i], res[i]
        say(f"  {name}: {c}")
        s = solve_case(c, precision="mixed", keep_state=True)
        assert s.cd == r.cd and s.cl == r.cl and s.newton_iters == r.newton_iters
        f = solve_case(c, precision="fp64", keep_state=True)
        dpsi = np.abs(s.psi - f.psi).max()
        dome = np.abs(s.ome - f.ome).max()
        say(f"    mixed vs fp64: max|dpsi| {dpsi:.2e} max|dome| {dome:.2e} dCd {s.cd - f.cd:.2e} "
            f"dCl {s.cl - f.cl:.2e}")
        assert dpsi <= 1e-8 and dome <= 1e-8
        assert abs(s.cd - f.cd) <= 1e-6 and abs(s.cl - f.cl) <= 1e-6 and abs(s.cm - f.cm) <= 1e-6


def test_c5_smallest_case_vs_oracle(sweep, C, oracle_mod):
    from vorti2d_amd.batch import elliptic_grid, solve_case
    cases, _ = sweep
    c = cases[_picks(cases)["smallest"]]
    say(f"  oracle solve of {c}")
    xg, yg = elliptic_grid(c.imax, c.jmax, c.aspect)
    dk, de = 1.0 / (c.imax - 1), 1.0 / (c.jmax - 1)
    m = C.compute_metrics(dk, de, xg, yg)
    a = math.radians(c.alpha_deg)
    ca, sa = math.cos(a), math.sin(a)
    r = oracle_mod.newton_solve(C, c.imax, c.jmax, m, c.re, ca=ca, sa=sa, tol=1e-12, max_iter=60)
    fo = oracle_mod.force_coeffs(c.imax, c.jmax, c.re, dk, de, m[2], m[3], m[6], m[7], m[8], m[9],
                                 r.ome, ref_length=1.0)
    for prec in ("fp64", "mixed"):
        s = solve_case(c, precision=prec, keep_state=True)
        dpsi = np.abs(s.psi - r.psi).max()
        dome = np.abs(s.ome - r.ome).max()
        say(f"    {prec} vs oracle: max|dpsi| {dpsi:.2e} max|dome| {dome:.2e} "
            f"dCd {s.cd - fo['cd']:.2e}")
        assert dpsi <= 1e-8 and dome <= 1e-8
        assert abs(s.cd - fo["cd"]) <= 1e-6 and abs(s.cl - fo["cl"]) <= 1e-6


def test_c5_adjoint_largest_case_vs_fd(sweep):
    from vorti2d_amd.batch import Case, solve_case
    cases, res = sweep
    i = _picks(cases)["largest"]
    c = cases[i]
    h = 1e-6
    cp = solve_case(Case(c.imax, c.jmax, c.aspect, c.re, c.alpha_deg + h), precision="fp64",
                    tol=1e-13)
    cm = solve_case(Case(c.imax, c.jmax, c.aspect, c.re, c.alpha_deg - h), precision="fp64",
                    tol=1e-13)
    fd = (cp.cd - cm.cd) / (2 * h)
    say(f"  {c}: dCd/dalpha fd {fd:.8e} batch adjoint {res[i].dcd_dalpha:.8e}")
    assert abs(fd - res[i].dcd_dalpha) <= 1e-5 * abs(fd) + 1e-8
