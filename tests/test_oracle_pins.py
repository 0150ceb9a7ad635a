"""Pins of the oracle's numpy restatements against the reference's own Python.

`oracle.generate_cylinder` restates vorti2d/mesh.py:14-50 and
`oracle.force_coeffs` restates vorti2d/forces.py:82-194.  Both are checked
here against fixtures that tests/golden/make_golden_forces_mesh.py wrote by
running those two reference files, and against the reference's published
steady drag coefficients on its production grid (docs/verification.rst:136-147:
Cd 2.036 at Re 20 and 1.522 at Re 40 on generate_cylinder(181, 181)).

The force fixtures use non-orthogonal grids (gamma != 0 at the wall), a moment
centre away from the origin and both reference-length modes, so a dropped
gamma * omega_xi term, a wrong moment sign or centre, or a swapped
pressure / friction split changes a pinned value; the planted-mistake test
shows that for each.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden

KEYS = ("cd", "cl", "cm", "cdp", "cdf", "clp", "clf", "cmp", "cmf", "ref_length")


def _cases():
    g = golden("forces_ref")
    out = []
    for name in g["names"]:
        name = str(name)
        m = g[f"{name}_metrics"]     # beta, gama, detadx, detady, xphys, yphys
        if name == "cyl41":
            nl = golden("newton_41x41_re40")
            imax, jmax, ome, re = 41, 41, nl["ome"], 40.0
        else:
            imax, jmax = g[f"{name}_xg"].shape
            ome, re = g[f"{name}_ome"], float(g[f"{name}_re"])
        out.append((name, imax, jmax, re, m, ome, g[f"{name}_forces"]))
    return out, g["variants_ref_length"], g["variants_moment_ref"]


def _oracle_forces(oracle_mod, imax, jmax, re, m, ome, ref_length, mref, gama=None):
    return oracle_mod.force_coeffs(imax, jmax, re, 1.0 / (imax - 1), 1.0 / (jmax - 1), m[0],
                                   m[1] if gama is None else gama, m[2], m[3], m[4], m[5], ome,
                                   ref_length=None if ref_length <= 0 else float(ref_length),
                                   moment_ref=tuple(mref))


def test_generate_cylinder_matches_reference(oracle_mod):
    import vorti2d_amd as v
    g = golden("cylinder_grids")
    for imax, jmax, rin, rout in g["cases"]:
        imax, jmax = int(imax), int(jmax)
        want_x, want_y = g[f"xg_{imax}x{jmax}"], g[f"yg_{imax}x{jmax}"]
        for gen in (oracle_mod.generate_cylinder, v.generate_cylinder):
            xg, yg = gen(imax, jmax, float(rin), float(rout))
            assert xg.shape == (imax, jmax)
            np.testing.assert_allclose(xg, want_x, rtol=0, atol=1e-14 * rout)
            np.testing.assert_allclose(yg, want_y, rtol=0, atol=1e-14 * rout)
        # radii exact at the wall and the far field
        r = np.hypot(want_x, want_y)
        assert np.allclose(r[:, 0], rin, rtol=1e-15) and np.allclose(r[:, -1], rout, rtol=1e-15)


def test_force_coeffs_match_reference(oracle_mod):
    cases, rl, mr = _cases()
    for name, imax, jmax, re, m, ome, want in cases:
        for v in range(len(rl)):
            fc = _oracle_forces(oracle_mod, imax, jmax, re, m, ome, rl[v], mr[v])
            for q, k in enumerate(KEYS):
                w = want[v, q]
                assert abs(fc[k] - w) <= 1e-14 * max(1.0, abs(w)) + 1e-15, (name, v, k, fc[k], w)


def test_force_fixture_pins_each_term(oracle_mod):
    """Plausible mistakes planted in the oracle's inputs / outputs break the fixture."""
    cases, rl, mr = _cases()
    name, imax, jmax, re, m, ome, want = cases[0]        # the bumped NACA grid
    assert name.startswith("naca")
    tol = lambda w: 1e-14 * max(1.0, abs(w)) + 1e-15   # noqa: E731
    # 1) the gamma * omega_xi term of d(omega)/dn dropped
    fc = _oracle_forces(oracle_mod, imax, jmax, re, m, ome, rl[1], mr[1], gama=np.zeros_like(m[1]))
    assert abs(fc["cdp"] - want[1, 3]) > 1e3 * tol(want[1, 3])
    # 2) the moment centre ignored (variant 1 has moment_ref = (0.1, -0.2))
    fc = _oracle_forces(oracle_mod, imax, jmax, re, m, ome, rl[1], (0.0, 0.0))
    assert abs(fc["cm"] - want[1, 2]) > 1e3 * tol(want[1, 2])
    # 3) the moment sign flipped, 4) pressure and friction drag swapped
    fc = _oracle_forces(oracle_mod, imax, jmax, re, m, ome, rl[1], mr[1])
    assert abs(-fc["cm"] - want[1, 2]) > 1e3 * tol(want[1, 2])
    assert abs(fc["cdf"] - want[1, 3]) > 1e3 * tol(want[1, 3])
    # 5) the reference length: the wall-estimated diameter differs from 2.0
    assert abs(want[0, 9] - 2.0) > 0.1


@pytest.mark.parametrize("re,cd_pub", [(40.0, 1.522), (20.0, 2.036)])
def test_published_steady_drag_181(C, oracle_mod, re, cd_pub):
    """docs/verification.rst:136-147: steady Cd on generate_cylinder(181, 181)."""
    imax = jmax = 181
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    d = 1.0 / 180
    m = C.compute_metrics(d, d, xg, yg)
    r = oracle_mod.newton_solve(C, imax, jmax, m, re, tol=1e-10, max_iter=40)
    fc = oracle_mod.force_coeffs(imax, jmax, re, d, d, m[2], m[3], m[6], m[7], m[8], m[9], r.ome)
    assert abs(fc["cd"] - cd_pub) < 1e-3, fc["cd"]
    assert abs(fc["cl"]) < 1e-6


def test_c2_fixture_first_level_and_published_values(oracle_mod, C):
    """The unsteady C2 fixture (tests/golden/make_c2_oracle.py: the reference
    Fortran assembly with SuperLU over 401 BDF2 levels on
    generate_cylinder(181, 181), Re 100) is reproduced at its first level by
    the oracle's C restatement, covers t = 0 .. 80 in steps of 0.2 with every
    level converged, and its saturated cycle lies near the reference's
    published Re 100 values (St 0.1535, docs/verification.rst:181-204; mean Cd
    about 1.31, examples/cylinder_unsteady.py:9)."""
    from lco import strouhal
    ref = golden("c2_oracle_cyl181")
    assert np.allclose(ref["t"], 0.2 * np.arange(401), rtol=0, atol=1e-12)
    assert (ref["res"] <= 1e-10).all() and (ref["newton_iters"] < 100).all()
    imax = jmax = 181
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    m = C.compute_metrics(1 / 180, 1 / 180, xg, yg)
    psi, ome = oracle_mod.initial_state(imax, jmax, m[8], m[9])
    r = oracle_mod.newton_solve(C, imax, jmax, m, 100.0, inv2dt=2.5, urot=0.5, ff_bc=1,
                                psi=psi, ome=ome, tol=1e-10)
    f = oracle_mod.force_coeffs(imax, jmax, 100.0, 1 / 180, 1 / 180, m[2], m[3], m[6], m[7],
                                m[8], m[9], r.ome)
    assert r.iters == ref["newton_iters"][0]
    for k in ("cd", "cl", "cm"):
        assert abs(f[k] - ref[k][0]) <= 1e-10 * max(1.0, abs(f[k])), k
    st = strouhal(ref["t"], ref["cl"])
    assert st["n_periods"] >= 4
    assert abs(st["st"] - 0.1535) <= 0.05 * 0.1535
    assert abs(ref["cd"][st["i0"]:].mean() - 1.31) <= 0.02 * 1.31
