"""GPU force integral (k_forces) against fixtures of the reference's own
compute_force_coeffs (vorti2d/forces.py:82-194), written by
tests/golden/make_golden_forces_mesh.py: non-orthogonal NACA and elliptic
O-grids with a seeded random vorticity and the converged 41x41 state, every
reference-length / moment-centre variant.  Bar: 1e-12 relative (the GPU sums
in a different order than numpy)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

KEYS = ("cd", "cl", "cm", "cdp", "cdf", "clp", "clf", "cmp", "cmf", "ref_length")


def _cases():
    g = golden("forces_ref")
    for name in g["names"]:
        name = str(name)
        if name == "cyl41":
            nl = golden("newton_41x41_re40")
            xg, yg, ome, re = nl["xg"], nl["yg"], nl["ome"], 40.0
        else:
            xg, yg = g[f"{name}_xg"], g[f"{name}_yg"]
            ome, re = g[f"{name}_ome"], float(g[f"{name}_re"])
        yield name, xg, yg, ome, re, g[f"{name}_metrics"], g[f"{name}_forces"], \
            g["variants_ref_length"], g["variants_moment_ref"]


def _close(got, want):
    return abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_force_host_entry_vs_reference_fixture():
    import vorti2d_amd as v
    for name, xg, yg, ome, re, m, want, rl, mr in _cases():
        imax, jmax = xg.shape
        print(f"forces host entry {name} {imax}x{jmax}", flush=True)
        for q in range(len(rl)):
            fc = v.compute_force_coeffs(imax, jmax, re, 1.0 / (imax - 1), 1.0 / (jmax - 1), *m,
                                        ome, ref_length=None if rl[q] <= 0 else float(rl[q]),
                                        moment_ref=tuple(mr[q]))
            for i, k in enumerate(KEYS):
                assert _close(getattr(fc, k), want[q, i]), (name, q, k)


def test_force_context_vs_reference_fixture():
    """Metrics computed on the device from the grid, forces of the uploaded
    state (v2d_ctx_forces)."""
    import vorti2d_amd as v
    for name, xg, yg, ome, re, _m, want, rl, mr in _cases():
        print(f"forces context {name} {xg.shape}", flush=True)
        ctx = v.DeviceContext(xg, yg, 0, 1.0, 0.0, "fp64")
        try:
            ctx.set_state(np.zeros(ome.size), ome)
            for q in range(len(rl)):
                out = ctx.forces(re, None if rl[q] <= 0 else float(rl[q]), tuple(mr[q]))
                for i, k in enumerate(KEYS):
                    assert _close(out[i], want[q, i]), (name, q, k)
        finally:
            ctx.close()
