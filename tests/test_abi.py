"""CPU-side checks of the drop-in boundary: libvorti2d.so loads and exports every
entry point declared in include/vorti2d.h; the host mirror of the reference API
(Config, mesh generators, restart / CSV formats) behaves like the reference;
argument errors are reported through v2d_last_error without touching a GPU."""
from __future__ import annotations

import ctypes
import os
import re
import tempfile

import numpy as np
import pytest

from conftest import ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "vorti2d.h")).read()
    return sorted(set(re.findall(r"\b(v2d_\w+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "vorti2d_amd", "libvorti2d.so"))
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    from vorti2d_amd import _lib
    assert sorted(_lib.EXPORTED) == header_symbols()
    assert _lib.lib.v2d_version() == 100


def test_package_api_mirrors_reference():
    import vorti2d_amd as v
    for name in ("Config", "Solver", "run", "generate_cylinder", "load_mesh", "save_mesh",
                 "load_cgns_ogrid", "compute_force_coeffs", "ForceCoeffs", "mesh", "core"):
        assert hasattr(v, name), name
    for name in ("compute_metrics", "assemble_coo", "compute_metrics_local",
                 "assemble_coo_local", "node_neighbors"):
        assert callable(getattr(v.core, name)), name
    from vorti2d_amd.forces import FORCE_COLUMNS
    assert FORCE_COLUMNS == ("t", "cd", "cl", "cm", "cdp", "cdf", "clp", "clf", "cmp", "cmf")


def test_errors_reported_without_gpu():
    from vorti2d_amd import core
    from vorti2d_amd._lib import V2DError
    with pytest.raises(V2DError, match="imax >= 4"):
        core.compute_metrics(1.0, 1.0, np.zeros((3, 5)), np.zeros((3, 5)))
    with pytest.raises(V2DError, match="out of range"):
        core.node_neighbors(0, 1, 10)
    with pytest.raises(ValueError):
        core.compute_metrics(1.0, 1.0, np.zeros((5, 5)), np.zeros((5, 6)))
    with pytest.raises(ValueError):
        core.assemble_coo(5, 5, 40, 0, 0, 0, .25, .25, *[np.zeros(25)] * 13, np.zeros(24),
                          0, 50, 650, 0, 1.0, 0.0)


def test_node_neighbors_host_map(C):
    from vorti2d_amd import core
    for imax in (4, 7, 181):
        for j in (2, 3):
            for i in range(1, imax + 1):
                assert core.node_neighbors(i, j, imax) == C.node_neighbors(i, j, imax)


def test_config_semantics():
    """tests/test_postproc.py:87-142 and vorti2d/config.py:139-166."""
    from vorti2d_amd import Config
    assert Config(farfield_bc="dirichlet").ff_bc_code() == 0
    assert Config(farfield_bc="outflow").ff_bc_code() == 1
    assert Config(farfield_bc="outflow_psi").ff_bc_code() == 2
    with pytest.raises(ValueError):
        Config(farfield_bc="bogus").ff_bc_code()
    assert abs(Config(alpha_deg=0.0).sin_alpha()) < 1e-15
    assert abs(Config(alpha_deg=90.0).cos_alpha()) < 1e-12
    assert Config().inv_dtau() == 0.0 and Config(dtau=4.0).inv_dtau() == 0.25
    assert Config(steady=True).inv_2dt() == 0.0
    assert Config(steady=False, dt_phys=0.2).inv_2dt() == pytest.approx(2.5)
    c = Config(steady=False, rot_speed=0.5, rot_until=2.0)
    assert c.u_rot(2.0) == 0.5 and c.u_rot(2.2) == 0.0
    assert Config(steady=True, rot_speed=0.5).u_rot(100.0) == 0.5
    with pytest.raises(ValueError):
        Config(precision="fp16").precision_code()
    with pytest.raises(ValueError):
        Config(linsolve="lu").check_linsolve()
    Config(linsolve="gmres_asm").check_linsolve()


def test_mesh_generators_match_oracle(oracle_mod):
    import vorti2d_amd as v
    for imax, jmax in ((25, 17), (181, 181), (4, 4)):
        a, b = v.generate_cylinder(imax, jmax, 0.5, 50.0), oracle_mod.generate_cylinder(imax, jmax)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        a, b = v.generate_annulus(imax, jmax), oracle_mod.generate_annulus(imax, jmax)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    xg, yg = v.generate_cylinder(41, 31, 0.5, 50.0)
    r = np.hypot(xg, yg)
    assert np.allclose(r[:, 0], 0.5) and np.allclose(r[:, -1], 50.0)
    xa, ya = v.generate_annulus(41, 31)
    ra = np.hypot(xa, ya)
    q = v.mesh.stretch_ratio(31)
    assert np.allclose(ra[0, 1:] / ra[0, :-1], q)


def test_mesh_csv_roundtrip():
    import vorti2d_amd as v
    xg, yg = v.generate_cylinder(9, 7)
    with tempfile.TemporaryDirectory() as tmp:
        v.save_mesh(xg, yg, os.path.join(tmp, "x.csv"), os.path.join(tmp, "y.csv"))
        a, b = v.load_mesh(os.path.join(tmp, "x.csv"), os.path.join(tmp, "y.csv"))
        assert np.array_equal(a, xg) and np.array_equal(b, yg)
    with pytest.raises(NotImplementedError):
        v.load_cgns_ogrid("nothing.cgns")


def test_restart_format_roundtrip():
    """vorti2d/restart.py:29-47 file layout (npz keys and scalar types)."""
    from vorti2d_amd import restart
    rng = np.random.default_rng(0)
    st = restart.RestartState(*[rng.standard_normal(12) for _ in range(4)], 1.5, 3, 4, 3, 60.0)
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "r.npz")
        restart.save_restart(p, st)
        with np.load(p, allow_pickle=False) as d:
            assert set(d.files) == {"psi", "ome", "omeold", "omeoldold", "t", "step", "imax",
                                    "jmax", "re"}
        s2 = restart.load_restart(p, 4, 3)
        assert np.array_equal(s2.psi, st.psi) and s2.t == 1.5 and s2.step == 3
        with pytest.raises(ValueError):
            restart.load_restart(p, 5, 3)


def test_fields_io_format():
    from vorti2d_amd import fields_io
    from vorti2d_amd.forces import FORCE_COLUMNS
    with tempfile.TemporaryDirectory() as tmp:
        fields_io.ensure_dirs(tmp)
        fields_io.write_fields(tmp, 7, np.arange(4.0), -np.arange(4.0))
        assert os.path.exists(os.path.join(tmp, "psi_data", "psi_t0007.csv"))
        a = np.loadtxt(os.path.join(tmp, "omega_data", "omega_t0007.csv"), delimiter=",")
        assert np.array_equal(a, -np.arange(4.0))
        fields_io.write_residual_history(tmp, 0, [-1.0, -2.5])
        h = np.loadtxt(os.path.join(tmp, "residual_data", "residual_history_t0000.csv"),
                       delimiter=",")
        assert np.array_equal(h, [[1, -1.0], [2, -2.5]])
        fields_io.init_force_history(tmp, FORCE_COLUMNS)
        fields_io.append_force_history(tmp, FORCE_COLUMNS, {c: 1.0 for c in FORCE_COLUMNS})
        rows = open(os.path.join(tmp, "forces.csv")).read().strip().splitlines()
        assert rows[0] == "t,cd,cl,cm,cdp,cdf,clp,clf,cmp,cmf" and len(rows) == 2
        assert rows[1].split(",")[0] == "1.0000000000e+00"


def test_sweep_cases_and_elliptic_grids(C):
    """Seeded sweep ranges; elliptic O-grids have the reference layout (seam
    copies equal, wall semi-axes, far-field circle) and positive metric
    Jacobians."""
    from vorti2d_amd.batch import elliptic_grid, sweep_cases
    cs = sweep_cases(64, seed=0)
    assert len(cs) == 64 and cs == sweep_cases(64, seed=0)
    assert all(129 <= c.imax <= 257 and 129 <= c.jmax <= 385 for c in cs)
    assert all(0.5 <= c.aspect <= 1.0 and 20 <= c.re <= 45 and -10 <= c.alpha_deg <= 10 for c in cs)
    xg, yg = elliptic_grid(33, 25, 0.6)
    assert np.array_equal(xg[0], xg[-1]) and np.abs(yg[0] - yg[-1]).max() < 1e-12
    assert abs(xg[:, 0].max() - 0.5) < 1e-12 and abs(yg[:, 0].max() - 0.3) < 1e-3
    assert np.allclose(np.hypot(xg[:, -1], yg[:, -1]), 50.0)
    m = C.compute_metrics(1 / 32, 1 / 24, xg, yg)
    assert np.all(m[0] > 0) and np.all(np.isfinite(m[0]))


def test_airfoil_grid_layout(C):
    """NACA O-grid: seam copies equal, chord from -0.5 to 0.5, far field at r = 30,
    positive metric Jacobians, bumps vanish at the leading / trailing edges."""
    from vorti2d_amd.airfoil import AirfoilGrid
    g = AirfoilGrid(65, 33, n_bumps=8)
    xg, yg = g.grid(np.full(8, 2e-3))
    assert np.array_equal(xg[0], xg[-1]) and np.abs(yg[0] - yg[-1]).max() < 1e-14
    assert abs(xg[:, 0].min() + 0.5) < 1e-12 and abs(xg[:, 0].max() - 0.5) < 1e-12
    assert np.allclose(np.hypot(xg[:, -1], yg[:, -1]), 30.0)
    m = C.compute_metrics(1 / 64, 1 / 32, xg, yg)
    assert np.all(m[0] > 0)
    assert np.all(np.hypot(xg[:, 0], yg[:, 0]) > 1e-3)      # no wall node at the origin
    nh = (65 - 1) // 2
    assert np.all(g.basis[:, [0, nh, 64], 0] == 0.0)          # TE base, nose, seam
    assert np.all(g.basis[:, :, -1] == 0.0)                    # far field fixed


def test_ignored_reference_knobs_warn():
    """The reference's PETSc menu and KSP knobs are accepted (configurations
    load unchanged) but not used; each one is named in a warning."""
    from vorti2d_amd import Config
    assert Config().ignored_reference_knobs() == []
    msgs = Config(linsolve="gmres_jacobi", ksp_rtol=1e-8, ilu_fill=2,
                  distributed=True).ignored_reference_knobs()
    assert len(msgs) == 4
    assert msgs[0].startswith("linsolve='gmres_jacobi'")
    assert any(m.startswith("ksp_rtol=") for m in msgs)
    assert any(m.startswith("ilu_fill=") for m in msgs)
    assert any(m.startswith("distributed=True") for m in msgs)
