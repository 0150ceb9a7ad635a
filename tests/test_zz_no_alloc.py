"""The no-allocation force entry (v2d_force_coeffs_noalloc) against the
allocating one (v2d_force_coeffs), through the C-ABI: the same kernel on the
same wall data, so the ten outputs agree bit for bit.  Two calls in a row, the
second on a smaller mesh (its scratch keeps the first call's larger contents),
both reference-length modes; sizes beyond the fixed scratch are rejected."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, imax, jmax, re, m, ome, ref_length, mc):
    from vorti2d_amd._lib import f64c, ptr
    arrs = [f64c(a) for a in (m[2], m[3], m[6], m[7], m[8], m[9], ome)]
    out = np.full(10, np.nan)
    rc = fn(imax, jmax, re, 1.0 / (imax - 1), 1.0 / (jmax - 1), *[ptr(a) for a in arrs],
            -1.0 if ref_length is None else ref_length, mc[0], mc[1], ptr(out))
    return rc, out


def test_noalloc_forces_equal_allocating_entry(C, oracle_mod):
    from vorti2d_amd._lib import lib
    rng = np.random.default_rng(11)
    cases = [(1025, 33, oracle_mod.generate_cylinder), (97, 41, oracle_mod.generate_annulus)]
    for imax, jmax, gen in cases:
        xg, yg = gen(imax, jmax)
        m = C.compute_metrics(1.0 / (imax - 1), 1.0 / (jmax - 1), xg, yg)
        ome = rng.standard_normal(imax * jmax)
        for ref_length, mc in ((None, (0.0, 0.0)), (2.0, (0.3, -0.1))):
            print(f"no-alloc forces {imax}x{jmax} ref_length={ref_length}", flush=True)
            rc1, a = _call(lib.v2d_force_coeffs, imax, jmax, 37.0, m, ome, ref_length, mc)
            rc2, b = _call(lib.v2d_force_coeffs_noalloc, imax, jmax, 37.0, m, ome, ref_length, mc)
            assert rc1 == 0 and rc2 == 0
            assert np.array_equal(a, b), (a, b)
            assert np.all(np.isfinite(b))


def test_noalloc_forces_rejects_oversize():
    from vorti2d_amd._lib import lib
    imax, jmax = 16386, 4
    z = np.zeros(imax * jmax)
    out = np.zeros(10)
    p = z.ctypes.data_as(ctypes.c_void_p)
    rc = lib.v2d_force_coeffs_noalloc(imax, jmax, 40.0, 1.0, 1.0, p, p, p, p, p, p, p, -1.0, 0.0,
                                      0.0, out.ctypes.data_as(ctypes.c_void_p))
    assert rc == -1
    assert b"exceeds" in lib.v2d_last_error()
