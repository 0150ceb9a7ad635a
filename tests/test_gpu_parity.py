"""GPU parity of the HIP path against the oracle and the reference fixtures.

Bars: integer outputs (COO indices, row pointers, nnz) bit-exact; FP64 metrics,
Jacobian values and RHS bit-exact (the kernels keep the reference operation
order, IEEE division and square root, no FMA contraction); FP32 Jacobian values
equal to the FP32 rounding of the FP64 values; linear solves to 1e-9 relative;
converged psi, omega within 1e-8 and force coefficients within 1e-6 of the FP64
oracle Newton solve (reference assembly + exact sparse LU).
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def v():
    import vorti2d_amd
    return vorti2d_amd


# ------------------------------------------------------------------ metrics --
@pytest.mark.parametrize("name", ["cyl_25x17", "cyl_41x41", "ann_33x21"])
def test_metrics_golden_bitexact(v, name):
    g = golden("metrics_" + name)
    m = v.core.compute_metrics(float(g["dksi"]), float(g["deta"]), g["xg"], g["yg"])
    for q in range(10):
        assert np.array_equal(m[q], g["metrics"][q]), q


@pytest.mark.parametrize("imax,jmax,kind", [(181, 181, "ann"), (181, 181, "cyl"),
                                            (257, 129, "cyl"), (1025, 1025, "ann")])
def test_metrics_vs_oracle(v, C, oracle_mod, imax, jmax, kind):
    say(f"metrics {kind} {imax}x{jmax}")
    gen = oracle_mod.generate_annulus if kind == "ann" else oracle_mod.generate_cylinder
    xg, yg = gen(imax, jmax)
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    a = v.core.compute_metrics(dk, de, xg, yg)
    b = C.compute_metrics(dk, de, xg, yg)
    for q in range(10):
        assert np.array_equal(a[q], b[q]), q


# ----------------------------------------------------------------- assembly --
@pytest.mark.parametrize("name", ["cyl_25x17_ff1", "cyl_41x41_ff0", "ann_33x21_ff2_a12"])
def test_assemble_coo_golden_bitexact(v, name):
    g = golden("assemble_" + name)
    xg, yg = g["xg"], g["yg"]
    imax, jmax = xg.shape
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    m = v.core.compute_metrics(dk, de, xg, yg)
    N = imax * jmax
    ci, cj, cv, nnz, b = v.core.assemble_coo(
        imax, jmax, float(g["re"]), float(g["invdtau"]), float(g["inv2dt"]), float(g["urot"]),
        dk, de, *m, *g["state"], 0, 2 * N, 26 * N, int(g["ff_bc"]), float(g["ca"]), float(g["sa"]))
    assert nnz == g["coo_i"].size
    assert np.array_equal(ci[:nnz], g["coo_i"]) and np.array_equal(cj[:nnz], g["coo_j"])
    assert np.array_equal(cv[:nnz], g["coo_v"]) and np.array_equal(b, g["bvec"])


def test_assemble_coo_row_ranges_vs_oracle(v, C, oracle_mod):
    imax, jmax = 61, 41
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    m = C.compute_metrics(dk, de, xg, yg)
    N = imax * jmax
    rng = np.random.default_rng(11)
    st = [rng.standard_normal(N) for _ in range(4)]
    for (r0, r1) in ((0, 2 * N), (0, N), (N, 2 * N), (37, 2 * N - 5), (N - 61, N + 61), (5, 5)):
        for ff, ca, sa in ((0, 1.0, 0.0), (1, 0.8, 0.6), (2, -0.6, 0.8)):
            args = (imax, jmax, 45.0, 0.2, 2.5, 0.3, dk, de, *m, *st, r0, r1,
                    13 * max(r1 - r0, 1), ff, ca, sa)
            a, b = v.core.assemble_coo(*args), C.assemble_coo(*args)
            assert a[3] == b[3], (r0, r1, ff)
            n = a[3]
            for q in range(3):
                assert np.array_equal(a[q][:n], b[q][:n]), (r0, r1, ff, q)
            assert np.array_equal(a[4], b[4])


def test_assemble_coo_capacity_checked(v, oracle_mod):
    from vorti2d_amd._lib import V2DError
    imax, jmax = 9, 7
    xg, yg = oracle_mod.generate_cylinder(imax, jmax)
    m = v.core.compute_metrics(1 / 8, 1 / 6, xg, yg)
    z = np.zeros(imax * jmax)
    with pytest.raises(V2DError, match="maxnnz"):
        v.core.assemble_coo(imax, jmax, 40, 0, 0, 0, 1 / 8, 1 / 6, *m, z, z, z, z, 0,
                            2 * imax * jmax, 100, 0, 1.0, 0.0)


def test_local_variants_golden(v):
    g = golden("local_cyl_25x17")
    ml = v.core.compute_metrics_local(int(g["il0"]), int(g["il1"]), float(g["dksi"]),
                                      float(g["deta"]), g["xgl"], g["ygl"])
    for q in range(10):
        assert np.array_equal(ml[q], g["metrics_local"][q]), q
    gxm, jmax = g["xgl"].shape
    ci, cj, cv, nnz, b = v.core.assemble_coo_local(
        gxm, jmax, int(g["il0"]), int(g["il1"]), float(g["re"]), float(g["invdtau"]),
        float(g["inv2dt"]), float(g["urot"]), float(g["dksi"]), float(g["deta"]),
        *g["loc_metrics"], *g["state"], 0, int(g["nloc"]), g["lg"], 2 * int(g["nloc"]) * 13,
        int(g["ff_bc"]), float(g["ca"]), float(g["sa"]))
    assert nnz == g["coo_i"].size
    assert np.array_equal(ci[:nnz], g["coo_i"]) and np.array_equal(cj[:nnz], g["coo_j"])
    assert np.array_equal(cv[:nnz], g["coo_v"]) and np.array_equal(b, g["bvec"])


# ------------------------------------------------- device context / pattern --
def _ctx_case(v, oracle_mod, C, imax, jmax, kind, prec, ff=0, ca=1.0, sa=0.0, seed=3, solver=None):
    gen = oracle_mod.generate_annulus if kind == "ann" else oracle_mod.generate_cylinder
    xg, yg = gen(imax, jmax)
    ctx = v.DeviceContext(xg, yg, ff, ca, sa, prec)
    if solver:
        ctx.set_linear_solver(solver)
    m = C.compute_metrics(1.0 / (imax - 1), 1.0 / (jmax - 1), xg, yg)
    N = imax * jmax
    psi, ome, o1, o2 = oracle_mod.smooth_state(m[8], m[9])
    rng = np.random.default_rng(seed)
    psi = psi + 0.01 * rng.standard_normal(N)
    ctx.set_state(psi, ome, o1, o2)
    return ctx, m, (psi, ome, o1, o2)


def _check_production_values(cv, ov, b, ob, prec):
    """The production assembly multiplies by reciprocals of the launch constants
    where the reference divides: entries agree to a few FP64 ulp (1e-13
    relative; the bar for Jacobian entries is 1e-12), FP32 entries are the
    rounding of those (at most one FP32 ulp from the rounding of the oracle's),
    the FP64 residual agrees to 1e-13 of its scale."""
    floor = 1e-15 * np.abs(ov).max()
    if prec == "fp64":
        assert np.all(np.abs(cv - ov) <= 1e-13 * np.abs(ov) + floor)
    else:
        assert np.all(np.abs(cv - ov) <= 2.0 ** -23 * np.abs(ov) + floor)
    assert np.abs(b - ob).max() <= 1e-13 * np.abs(ob).max()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("ff,ca,sa", [(0, 1.0, 0.0), (1, 1.0, 0.0), (2, 0.8, 0.6)])
def test_ctx_export_matches_oracle(v, C, oracle_mod, prec, ff, ca, sa):
    imax, jmax = 49, 33
    ctx, m, st = _ctx_case(v, oracle_mod, C, imax, jmax, "cyl", prec, ff, ca, sa)
    N = imax * jmax
    sc = (70.0, 0.5, 2.5, 0.25)
    ctx.assemble(*sc)
    ci, cj, cv, b = ctx.export_coo()
    oi, oj, ov, nnz, ob = C.assemble_coo(imax, jmax, sc[0], sc[1], sc[2], sc[3], 1 / 48, 1 / 32,
                                         *m, *st, 0, 2 * N, 26 * N, ff, ca, sa)
    assert ctx.nnz() == nnz
    assert np.array_equal(ci, oi[:nnz]) and np.array_equal(cj, oj[:nnz])
    _check_production_values(cv, ov[:nnz], b, ob, prec)
    # canonical CSR == scipy's COO->CSR (vorti2d/petsc_solver.py:76-95)
    ip, ix, dv = ctx.export_csr()
    A = sp.csr_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    A.sort_indices()
    assert np.array_equal(ip, A.indptr) and np.array_equal(ix, A.indices)
    assert np.array_equal(dv, A.data)
    # stencil SpMV
    x = np.random.default_rng(4).standard_normal(2 * N)
    y = ctx.spmv(x)
    ref = A @ x
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(A).dot(np.abs(x)).max()
    assert abs(ctx.rhs_norm() - np.linalg.norm(ob)) <= 1e-13 * np.linalg.norm(ob)


@pytest.mark.parametrize("imax,jmax,kind,ff", [(41, 41, "cyl", 0), (57, 33, "ann", 2), (20, 4, "cyl", 0),
                                               (25, 5, "cyl", 1), (33, 8, "ann", 2), (7, 6, "ann", 0),
                                               (181, 181, "ann", 1)])
@pytest.mark.parametrize("upd", ["mfma64", "valu64", "valu128"])
def test_direct_solve_fp64(v, C, oracle_mod, imax, jmax, kind, ff, upd):
    say(f"direct fp64 {kind} {imax}x{jmax} ff={ff} update={upd}")
    ctx, m, st = _ctx_case(v, oracle_mod, C, imax, jmax, kind, "fp64", ff, solver="block_lu")
    ctx.set_update_kernel(upd)
    ctx.assemble(40.0, 0.0, 2.5, 0.1)
    ci, cj, cv, b = ctx.export_coo()
    N = imax * jmax
    A = sp.csc_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    ctx.factor()
    x = ctx.solve(b)
    xs = spla.spsolve(A, b)
    assert np.linalg.norm(b - A @ x) <= 1e-11 * np.linalg.norm(b)
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)


@pytest.mark.parametrize("imax,jmax,kind,ff", [(41, 41, "cyl", 0), (181, 181, "ann", 1), (130, 9, "cyl", 2)])
def test_direct_solve_fp32_update_kernels(v, C, oracle_mod, imax, jmax, kind, ff):
    """FP32 block LU with each Gauss-Jordan update kernel: the solution error
    against the exact solve of the same (FP32-valued) matrix is set by FP32
    rounding, so the matrix-core and VALU kernels must agree in size."""
    say(f"direct fp32 update kernels {kind} {imax}x{jmax} ff={ff}")
    ctx, m, st = _ctx_case(v, oracle_mod, C, imax, jmax, kind, "mixed", ff, solver="block_lu")
    ctx.assemble(40.0, 0.0, 2.5, 0.1)
    ci, cj, cv, b = ctx.export_coo()
    N = imax * jmax
    A = sp.csc_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    xs = spla.spsolve(A, b)
    err = {}
    for upd in ("valu64", "valu128", "mfma64"):
        ctx.set_update_kernel(upd)
        ctx.factor()
        x = ctx.solve(b)
        assert np.all(np.isfinite(x)), upd
        err[upd] = np.linalg.norm(x - xs) / np.linalg.norm(xs)
        assert np.linalg.norm(b - A @ x) <= 5e-2 * np.linalg.norm(b), upd
    say(f"  relative errors {err}")
    for upd in ("valu128", "mfma64"):
        assert err[upd] <= max(4.0 * err["valu64"], 1e-6), (upd, err)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("imax,jmax,kind,ff", [(17, 4, "cyl", 0), (17, 5, "ann", 1), (41, 23, "cyl", 1),
                                               (73, 40, "ann", 2), (1025, 6, "ann", 0)])
def test_persistent_sweep_equals_level_sweeps(v, C, oracle_mod, prec, imax, jmax, kind, ff):
    """The one-launch persistent solve (ticketed level steps, device counters)
    forms the same sums in the same order as one launch pair per level step:
    the solutions are bit-identical.  Shapes: 2 and 3 levels, odd and even
    level counts, and 4100-unknown levels (rows past the register prefetch,
    GEMV rows past the prefetched width)."""
    say(f"persistent sweep {prec} {imax}x{jmax} {kind} ff={ff}")
    ctx, m, st = _ctx_case(v, oracle_mod, C, imax, jmax, kind, prec, ff, solver="block_lu")
    ctx.assemble(40.0, 0.0, 2.5, 0.1)
    ci, cj, cv, b = ctx.export_coo()
    ctx.factor()
    ctx.set_sweep_kernel("levels")
    x_lv = ctx.solve(b)
    ctx.set_sweep_kernel("persistent")
    x_ps = ctx.solve(b)
    x_ps2 = ctx.solve(b)
    assert np.all(np.isfinite(x_ps))
    assert np.array_equal(x_ps, x_lv)
    assert np.array_equal(x_ps, x_ps2)
    N = imax * jmax
    A = sp.csc_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    rel = np.linalg.norm(b - A @ x_ps) / np.linalg.norm(b)
    say(f"  relative residual {rel:.2e}")
    assert rel <= (1e-10 if prec == "fp64" else 5e-2)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_persistent_sweep_expired_waits_fall_back(v, C, oracle_mod, prec):
    """The persistent solve's waits are bounded.  With a bound of 0 polls
    every wait expires at once, the launch gives up and the solve is redone
    with launch pairs, which the context then keeps: the first solve falls
    back (counted), the second runs launch pairs directly, both equal the
    launch-pair solve bit for bit, and a later persistent solve with the
    default bound does too, without a fall-back."""
    say(f"persistent sweep fall-back {prec}")
    ctx, m, st = _ctx_case(v, oracle_mod, C, 73, 40, "ann", prec, 2, solver="block_lu")
    ctx.assemble(40.0, 0.0, 2.5, 0.1)
    ci, cj, cv, b = ctx.export_coo()
    ctx.factor()
    ctx.set_sweep_kernel("levels")
    x_lv = ctx.solve(b)
    ctx.set_sweep_kernel("persistent")
    assert ctx.sweep_kernel() == ("persistent", 0)
    ctx.set_sweep_spin_limit(0)
    assert np.array_equal(ctx.solve(b), x_lv)
    assert ctx.sweep_kernel() == ("levels", 1)
    assert np.array_equal(ctx.solve(b), x_lv)
    assert ctx.sweep_kernel() == ("levels", 1)
    ctx.set_sweep_spin_limit(1 << 21)
    ctx.set_sweep_kernel("persistent")
    assert np.array_equal(ctx.solve(b), x_lv)
    assert ctx.sweep_kernel() == ("persistent", 1)


@pytest.mark.parametrize("imax,jmax,kind,ff", [(17, 4, "cyl", 0), (41, 23, "cyl", 1), (73, 40, "ann", 2),
                                               (1025, 8, "ann", 0)])
def test_fused_sweep_matches_level_sweeps(v, C, oracle_mod, imax, jmax, kind, ff):
    """FP32 solve with one fused launch per level step (every workgroup forms
    the level right-hand side itself): the right-hand sides are the launch
    pair's and every row but the split rows (a level's rows past 16 per
    compute unit: the 2050-unknown levels of imax 1025 on two elimination
    chains) is summed in the same order, so the solution is bit-identical
    without split rows and agrees to FP32 rounding with them (their FP64 sums
    differ in order only, so their FP32 roundings mostly coincide).  The
    4100-unknown end levels of imax 1025 run as launch pairs."""
    say(f"fused sweep {imax}x{jmax} {kind} ff={ff}")
    ctx, m, st = _ctx_case(v, oracle_mod, C, imax, jmax, kind, "mixed", ff, solver="block_lu")
    ctx.assemble(40.0, 0.0, 2.5, 0.1)
    ci, cj, cv, b = ctx.export_coo()
    ctx.factor()
    ctx.set_sweep_kernel("levels")
    x_lv = ctx.solve(b)
    ctx.set_sweep_kernel("fused")
    x_fu = ctx.solve(b)
    x_fu2 = ctx.solve(b)
    assert np.all(np.isfinite(x_fu))
    assert np.array_equal(x_fu, x_fu2)
    N = imax * jmax
    A = sp.csc_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    if imax < 1025:
        assert np.array_equal(x_fu, x_lv)
    else:
        xs = spla.spsolve(A, b)
        err_lv = np.linalg.norm(x_lv - xs) / np.linalg.norm(xs)
        d = np.linalg.norm(x_fu - x_lv) / np.linalg.norm(x_lv)
        say(f"  levels error {err_lv:.2e}, fused - levels {d:.2e}")
        assert d <= max(4.0 * err_lv, 1e-6)
    rel = np.linalg.norm(b - A @ x_fu) / np.linalg.norm(b)
    say(f"  relative residual {rel:.2e}")
    assert rel <= 5e-2


def test_fused_sweep_newton_equals_level_sweeps(v, oracle_mod):
    """Mixed Newton-GMRES on 181^2 (refactoring every step, so no measured
    timing steers it) with the fused solve steps, captured in the HIP graph of
    the preconditioner apply: every level has 362 unknowns (no split rows), so
    the residual history and the state equal the launch-pair run bit for bit."""
    imax = jmax = 181
    xg, yg = oracle_mod.generate_annulus(imax, jmax)
    out = {}
    for kind in ("levels", "fused"):
        say(f"fused sweep newton: {kind}")
        ctx = v.DeviceContext(xg, yg, 0, 1.0, 0.0, "mixed")
        ctx.set_linear_solver("block_lu")
        ctx.set_solver(refactor_every=1, gmres_rtol=1e-6, gmres_max_iter=30)
        ctx.set_sweep_kernel(kind)
        ctx.init_state()
        hist = []
        res, k = 1.0, 0
        while res > 1e-10 and k < 20:
            res, rn, li = ctx.newton_step(40.0)
            hist.append((res, rn, li))
            k += 1
        out[kind] = (hist, ctx.get_state(("psi", "ome")))
        ctx.close()
    assert out["fused"][0] == out["levels"][0]
    assert out["fused"][0][-1][0] <= 1e-10
    for a, b in zip(out["fused"][1], out["levels"][1]):
        assert np.array_equal(a, b)


def test_forcing_residual_target(v, C, oracle_mod):
    """With a residual target (the final-state acceptance ||R||_2 / (2N) <
    1e-10) on the forcing term, the mixed solve still meets both stopping
    rules (res = ||delta||_2 / (2N) <= 1e-10 and the final residual) and
    spends no more GMRES iterations than without it (on the 1025^2 benchmark
    with a stored factorisation it saves about 25).  Refactoring every step
    keeps the iteration counts independent of measured timings."""
    imax = jmax = 181
    xg, yg = oracle_mod.generate_annulus(imax, jmax)
    n2 = 2 * imax * jmax
    tot = {}
    for tgt in (0.0, 1e-10 * n2):
        ctx = v.DeviceContext(xg, yg, 0, 1.0, 0.0, "mixed")
        ctx.set_solver(refactor_every=1, gmres_rtol=1e-6, gmres_max_iter=30)
        ctx.set_forcing(0.1, tgt)
        ctx.init_state()
        res, its, k = 1.0, [], 0
        while res > 1e-10 and k < 20:
            res, rn, li = ctx.newton_step(40.0)
            its.append(li)
            k += 1
        ctx.assemble(40.0)
        rfin = ctx.rhs_norm()
        say(f"  target {tgt:.1e}: {k} Newton steps, GMRES {its}, res {res:.2e}, ||R||/2N {rfin / n2:.2e}")
        assert res <= 1e-10 and rfin / n2 < 1e-10
        tot[tgt] = sum(its)
        ctx.close()
    assert tot[1e-10 * n2] <= tot[0.0]


@pytest.mark.parametrize("solver", ["multifrontal", "block_lu"])
def test_graph_replayed_gmres_equals_eager(v, C, oracle_mod, solver):
    """The mixed Newton step with the preconditioner apply (either linear
    solver) replayed from a captured HIP graph, and with whole GMRES
    iterations (apply, SpMV, orthogonalisation, norm, read-backs) replayed from
    graphs captured per basis index, equals the step with every kernel
    launched one by one, bit for bit, over several steps (each replay followed
    by the host's Givens update and eager work of the step), across
    refactorisations, and after switching graphs off and on again."""
    imax, jmax = 97, 49
    xg, yg = oracle_mod.generate_annulus(imax, jmax)
    hist = {}
    modes = {"eager": "off", "graph": "apply", "iteration": "iteration", "toggled": "iteration"}
    for mode in modes:
        ctx = v.DeviceContext(xg, yg, 0, 1.0, 0.0, "mixed")
        ctx.set_linear_solver(solver)
        # refactor every step: the iteration counts then do not depend on
        # measured timings, and the graph is replayed across refactorisations
        ctx.set_solver(refactor_every=1, gmres_rtol=1e-8, gmres_max_iter=30)
        ctx.set_graph(modes[mode])
        ctx.init_state()
        out = []
        for k in range(5):
            if mode == "toggled" and k == 2:
                ctx.set_graph(False)
                ctx.set_graph("iteration")
            res, rn, its = ctx.newton_step(40.0)
            out.append((res, rn, its))
        out.append(tuple(ctx.get_state(("psi", "ome"))))
        hist[mode] = out
        ctx.close()
    say(f"  eager {[h[:3] for h in hist['eager'][:5]]}")
    for mode in ("graph", "iteration", "toggled"):
        for a, b in zip(hist["eager"][:5], hist[mode][:5]):
            assert a == b, (mode, a, b)
        assert np.array_equal(hist["eager"][5][0], hist[mode][5][0])
        assert np.array_equal(hist["eager"][5][1], hist[mode][5][1])


@pytest.mark.parametrize("solver", ["multifrontal", "block_lu"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_stored_factorisation_is_self_contained(v, C, oracle_mod, prec, solver):
    """A stored factorisation keeps solving the system it was built from after
    later assemblies overwrite the Jacobian (the block LU's solve sweeps must
    not read the new inter-level couplings)."""
    imax, jmax = 73, 41
    ctx, m, st = _ctx_case(v, oracle_mod, C, imax, jmax, "cyl", prec, 1, solver=solver)
    ctx.assemble(40.0, 0.0, 2.5, 0.1)
    ci, cj, cv, b = ctx.export_coo()
    ctx.factor()
    x1 = ctx.solve(b)
    ctx.assemble(75.0, 0.3, 0.0, 0.0)          # different Jacobian, same state
    x2 = ctx.solve(b)
    assert np.array_equal(x1, x2)
    if prec == "fp64":
        N = imax * jmax
        A = sp.csc_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
        assert np.linalg.norm(b - A @ x2) <= 1e-11 * np.linalg.norm(b)


def test_mixed_newton_step_matches_fp64(v, C, oracle_mod):
    """The FP32 block LU alone is a preconditioner (FP32 accuracy); FP32 GMRES
    on it must reproduce the FP64 exact Newton update to the GMRES tolerance."""
    imax = jmax = 91
    ctx, m, st = _ctx_case(v, oracle_mod, C, imax, jmax, "cyl", "mixed")
    ctx.assemble(40.0, 0.0, 0.0, 0.0)
    ci, cj, cv, b = ctx.export_coo()
    N = imax * jmax
    A = sp.csc_matrix((cv, (ci, cj)), shape=(2 * N, 2 * N))
    ctx.factor()
    x = ctx.solve(b)
    # The stretched cylinder grid makes this matrix ill-conditioned (the FP64
    # block LU leaves a 5e-11 relative residual): the FP32 LU alone leaves
    # 3e-2 .. 6e-2 depending on the accumulation order of the update kernel.
    assert np.linalg.norm(b - A @ x) <= 1e-1 * np.linalg.norm(b)
    ref, _, _ = _ctx_case(v, oracle_mod, C, imax, jmax, "cyl", "fp64")
    p0 = ctx.get_state(("psi", "ome"))
    r64 = ref.newton_step(40.0)
    r32 = ctx.newton_step(40.0)
    d64 = np.concatenate(ref.get_state(("psi", "ome"))) - np.concatenate(p0)
    d32 = np.concatenate(ctx.get_state(("psi", "ome"))) - np.concatenate(p0)
    assert r32[2] <= 10                                   # GMRES iterations
    ref.assemble(40.0, 0.0, 0.0, 0.0)                     # FP64 matrix of the same state
    ref.set_state(*p0)
    ref.assemble(40.0, 0.0, 0.0, 0.0)
    ri, rj, rv, rb = ref.export_coo()
    A64 = sp.csc_matrix((rv, (ri, rj)), shape=(2 * N, 2 * N))
    # FP32 GMRES stops at 1e-6 relative residual of the FP32 system.  Against
    # the FP64 matrix the step is limited by the FP32 rounding of the Jacobian
    # entries (eps32 * ||A|| ||d|| / ||b|| ~ 1e-3 here); the FP64 residual of
    # the next Newton iteration corrects it (see the Newton parity tests).
    assert np.linalg.norm(rb - A64 @ d32) <= 1e-2 * np.linalg.norm(rb)
    assert np.linalg.norm(d32 - d64) <= 1e-2 * np.linalg.norm(d64)
    assert abs(r32[1] - r64[1]) <= 1e-12 * r64[1]         # same FP64 residual norm


# ------------------------------------------------------------ Newton solve --
def _run(v, xg, yg, **kw):
    with tempfile.TemporaryDirectory() as tmp:
        cfg = v.Config(out_dir=tmp, verbose=False, write_csv=False, write_xdmf=False,
                       compute_forces=False, **kw)
        return v.Solver(cfg, mesh=(xg, yg)).run()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_newton_41_matches_reference_state(v, prec):
    g = golden("newton_41x41_re40")
    s = _run(v, g["xg"], g["yg"], re=40.0, steady=True, pseudo_tol=1e-10, max_pseudo_iter=50,
             precision=prec)
    psi, ome = s.psi, s.ome
    assert np.abs(psi - g["psi"]).max() < 1e-8
    assert np.abs(ome - g["ome"]).max() < 1e-8
    if prec == "fp64":
        assert len(s.res_history[0]) == int(g["iters"])
        assert np.allclose(s.res_history[0], np.log10(g["res_hist"]), atol=1e-3)
    fc = s.force_coeffs()
    want = g["forces"]
    assert abs(fc.cd - want[0]) < 1e-6 and abs(fc.cl - want[1]) < 1e-6
    assert abs(fc.cm - want[2]) < 1e-6 and abs(fc.ref_length - want[9]) < 1e-12
    imax, jmax = g["xg"].shape
    N = imax * jmax
    ff = slice(imax * (jmax - 1), N)
    wall = slice(0, imax)
    assert np.max(np.abs(psi[ff] - s.yphys[ff])) < 1e-10     # test_smoke.py:50-56
    assert np.max(np.abs(psi[wall])) < 1e-10
    assert np.max(np.abs(ome[ff])) < 1e-12
    assert abs(psi.max() + psi.min()) < 1e-6


def test_mixed_forcing_converges(v):
    """Eisenstat-Walker GMRES tolerances (loose early, gmres_rtol near the
    solution) reach the reference converged state."""
    g = golden("newton_41x41_re40")
    s = _run(v, g["xg"], g["yg"], re=40.0, steady=True, pseudo_tol=1e-10, max_pseudo_iter=50,
             precision="mixed", refactor_every=0, gmres_max_iter=20, gmres_forcing=0.1)
    assert np.abs(s.psi - g["psi"]).max() < 1e-8
    assert np.abs(s.ome - g["ome"]).max() < 1e-8


def test_mixed_factorisation_reuse_converges(v):
    """Adaptive policy (refactor only when GMRES on the stale FP32 block LU misses
    its tolerance) reaches the same converged state as exact FP64 Newton."""
    g = golden("newton_41x41_re40")
    s = _run(v, g["xg"], g["yg"], re=40.0, steady=True, pseudo_tol=1e-10, max_pseudo_iter=50,
             precision="mixed", refactor_every=0, gmres_max_iter=12)
    assert np.abs(s.psi - g["psi"]).max() < 1e-8
    assert np.abs(s.ome - g["ome"]).max() < 1e-8
    tm = s.ctx.timers()
    assert tm["factorizations"] < tm["steps"]


@pytest.mark.parametrize("kind", ["ann", "cyl"])
def test_newton_181_matches_oracle(v, C, oracle_mod, kind):
    """Config C1: 181x181 O-grid, Re=40 steady, Dirichlet far field, alpha 0."""
    imax = jmax = 181
    gen = oracle_mod.generate_annulus if kind == "ann" else oracle_mod.generate_cylinder
    xg, yg = gen(imax, jmax)
    m = C.compute_metrics(1 / 180, 1 / 180, xg, yg)
    say(f"oracle Newton 181 {kind}")
    r = oracle_mod.newton_solve(C, imax, jmax, m, 40.0, tol=1e-10)
    fo = oracle_mod.force_coeffs(imax, jmax, 40.0, 1 / 180, 1 / 180, m[2], m[3], m[6], m[7],
                                 m[8], m[9], r.ome)
    for prec in ("fp64", "mixed"):
        say(f"device Newton 181 {kind} {prec}")
        s = _run(v, xg, yg, re=40.0, steady=True, pseudo_tol=1e-10, max_pseudo_iter=50,
                 precision=prec)
        assert np.abs(s.psi - r.psi).max() < 1e-8, prec
        assert np.abs(s.ome - r.ome).max() < 1e-8, prec
        fc = s.force_coeffs()
        for k in ("cd", "cl", "cm"):
            assert abs(getattr(fc, k) - fo[k]) < 1e-6, (prec, k)
        assert s.rnorm_history[0][-1] < 1e-6


def test_forces_host_entry_vs_oracle(v, C, oracle_mod):
    g = golden("newton_41x41_re40")
    cases = [(g["xg"], g["yg"], g["ome"])]
    xg, yg = oracle_mod.generate_cylinder(1025, 33)          # 1024 wall nodes: 16 wavefronts
    m = C.compute_metrics(1 / 1024, 1 / 32, xg, yg)
    cases.append((xg, yg, oracle_mod.smooth_state(m[8], m[9])[1]))
    for xg, yg, ome in cases:
        imax, jmax = xg.shape
        dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
        m = C.compute_metrics(dk, de, xg, yg)
        for ref_len, mc in ((None, (0.0, 0.0)), (2.0, (0.25, -0.1))):
            fc = v.compute_force_coeffs(imax, jmax, 40.0, dk, de, m[2], m[3], m[6], m[7], m[8],
                                        m[9], ome, ref_length=ref_len, moment_ref=mc)
            fo = oracle_mod.force_coeffs(imax, jmax, 40.0, dk, de, m[2], m[3], m[6], m[7], m[8],
                                         m[9], ome, ref_length=ref_len, moment_ref=mc)
            for k, val in fo.items():
                assert abs(getattr(fc, k) - val) <= 1e-12 * max(1.0, abs(val)), (imax, k)


def test_angle_of_attack_farfield(v):
    """tests/test_postproc.py:121-142 (41x41 for speed)."""
    xg, yg = v.generate_cylinder(41, 41, 0.5, 50.0)
    s = _run(v, xg, yg, re=20.0, steady=True, alpha_deg=12.0, pseudo_tol=1e-11,
             max_pseudo_iter=80)
    imax, jmax = s.imax, s.jmax
    ff = slice(imax * (jmax - 1), imax * jmax)
    ca, sa = s.cfg.cos_alpha(), s.cfg.sin_alpha()
    assert np.max(np.abs(s.psi[ff] - (ca * s.yphys[ff] - sa * s.xphys[ff]))) < 1e-9


def test_farfield_outflow_option(v):
    """tests/test_postproc.py:87-118: omega-only outflow leaves the steady Cd
    unchanged on a large domain."""
    xg, yg = v.generate_cylinder(101, 151, 0.5, 100.0)
    cd = {}
    for bc in ("dirichlet", "outflow"):
        s = _run(v, xg, yg, re=20.0, steady=True, farfield_bc=bc, pseudo_tol=1e-11,
                 max_pseudo_iter=80)
        fc = s.force_coeffs()
        assert abs(fc.cl) < 1e-3
        cd[bc] = fc.cd
    assert abs(cd["outflow"] - cd["dirichlet"]) < 1e-6


def test_unsteady_restart_equals_continuous(v):
    """tests/test_smoke.py:59-74 and tests/test_restart.py:37-54: an unsteady run
    resumed from its restart file matches the continuous run."""
    xg, yg = v.generate_cylinder(41, 41, 0.5, 50.0)
    common = dict(re=40.0, steady=False, dt_phys=0.5, rot_speed=0.5, rot_until=2.0,
                  pseudo_tol=1e-9, max_pseudo_iter=50, verbose=False, write_csv=False,
                  write_xdmf=False)
    with tempfile.TemporaryDirectory() as tmp:
        full = v.Solver(v.Config(out_dir=os.path.join(tmp, "a"), t_end=1.5, **common),
                        mesh=(xg, yg)).run()
        part = v.Solver(v.Config(out_dir=os.path.join(tmp, "b"), t_end=1.0,
                                 restart_out="r.npz", **common), mesh=(xg, yg)).run()
        rpath = os.path.join(tmp, "b", "r.npz")
        assert os.path.exists(rpath)
        res = v.Solver(v.Config(out_dir=os.path.join(tmp, "c"), t_end=1.5, restart_in=rpath,
                                **common), mesh=(xg, yg)).run()
        assert res.t == pytest.approx(1.5) and full.t == pytest.approx(1.5)
        assert np.abs(res.psi - full.psi).max() < 1e-9
        assert np.abs(res.ome - full.ome).max() < 1e-9
        rows = open(os.path.join(tmp, "a", "forces.csv")).read().strip().splitlines()
        assert rows[0].startswith("t,cd,cl,cm") and len(rows) == 1 + 4   # t = 0, .5, 1, 1.5


def test_full_size_assembly_1025(v, C, oracle_mod):
    """1025x1025 (nnz 24,124,400): the device pattern equals the oracle's, values
    and RHS agree to the production-assembly tolerance at full size."""
    imax = jmax = 1025
    xg, yg = oracle_mod.generate_annulus(imax, jmax)
    say("full-size ctx create")
    ctx = v.DeviceContext(xg, yg, 0, 1.0, 0.0, "fp64")
    m = ctx.metrics()
    N = imax * jmax
    psi, ome, o1, o2 = oracle_mod.smooth_state(m[8], m[9])
    ctx.set_state(psi, ome, o1, o2)
    ctx.assemble(40.0)
    assert ctx.nnz() == 24124400
    ci, cj, cv, b = ctx.export_coo()
    del ctx
    say("full-size oracle assembly")
    oi, oj, ov, nnz, ob = C.assemble_coo(imax, jmax, 40.0, 0, 0, 0, 1 / 1024, 1 / 1024, *m, psi,
                                         ome, o1, o2, 0, 2 * N, 23 * N, 0, 1.0, 0.0)
    assert nnz == 24124400
    assert np.array_equal(ci, oi[:nnz]) and np.array_equal(cj, oj[:nnz])
    _check_production_values(cv, ov[:nnz], b, ob, "fp64")


# ------------------------------------------------- unsteady shedding (C2) --
def _unsteady_c2(v, prec, xg, yg, tmp, t_end):
    cfg = v.Config(re=100.0, steady=False, dt_phys=0.2, t_start=0.0, t_end=t_end,
                   rot_speed=0.5, rot_until=2.0, farfield_bc="outflow", pseudo_tol=1e-10,
                   max_pseudo_iter=100, out_dir=os.path.join(tmp, prec), verbose=False,
                   write_csv=False, write_xdmf=False, compute_forces=True, precision=prec)
    s = v.Solver(cfg, mesh=(xg, yg)).run()
    d = np.genfromtxt(os.path.join(tmp, prec, "forces.csv"), delimiter=",", names=True)
    return s, d


def test_unsteady_first_steps_match_oracle(v, C, oracle_mod):
    """181x181, Re 100, BDF2 dt 0.2, rotation kick 0.5, outflow far field: the
    first three physical steps of the device dual-time loop (FP64 and mixed)
    equal the oracle's FP64 loop with exact LU solves."""
    imax = jmax = 181
    xg, yg = oracle_mod.generate_annulus(imax, jmax)
    m = C.compute_metrics(1 / 180, 1 / 180, xg, yg)
    N = imax * jmax
    psi, ome = oracle_mod.initial_state(imax, jmax, m[8], m[9])
    o1 = np.zeros(N)
    o2 = np.zeros(N)
    for step in range(3):                     # vorti2d/solver.py:169-190
        r = oracle_mod.newton_solve(C, imax, jmax, m, 100.0, inv2dt=2.5, urot=0.5, ff_bc=1,
                                    psi=psi, ome=ome, omeold=o1, omeoldold=o2, tol=1e-10)
        psi, ome = r.psi, r.ome
        o2, o1 = o1, ome.copy()
    with tempfile.TemporaryDirectory() as tmp:
        for prec in ("fp64", "mixed"):
            say(f"unsteady first steps {prec}")
            s, _ = _unsteady_c2(v, prec, xg, yg, tmp, t_end=0.4)
            assert s.step == 3
            assert np.abs(s.psi - psi).max() < 1e-8, prec
            assert np.abs(s.ome - ome).max() < 1e-8, prec


def test_unsteady_shedding_mixed_vs_fp64(v):
    """181x181 cylinder O-grid, Re 100, BDF2 dual time with dt 0.2 to t = 80,
    rotation kick 0.5 until t = 2, outflow far field.  The mixed path (FP32
    Jacobian / block LU / GMRES under FP64 Newton refinement) against the plain
    FP64 path: Cl(t), Cd(t) and the Strouhal number of the saturated cycle.
    Shedding is an instability, so round-off level differences between the two
    paths grow during the transient; the saturated cycle (amplitude, mean drag,
    frequency) must agree closely, and St must sit near the reference's own
    Re = 100 value 0.1535 (docs/verification.rst:181-204)."""
    from lco import strouhal
    xg, yg = v.generate_annulus(181, 181, 0.5, 50.0)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for prec in ("fp64", "mixed"):
            say(f"unsteady C2 {prec}")
            s, d = _unsteady_c2(v, prec, xg, yg, tmp, t_end=80.0)
            st = strouhal(d["t"], d["cl"])
            i0 = st["i0"]
            out[prec] = dict(t=d["t"], cl=d["cl"], cd=d["cd"], st=st,
                             cd_mean=float(d["cd"][i0:].mean()))
            say(f"  {prec}: St {st['st']:.4f} over {st.get('n_periods')} periods from "
                f"t={d['t'][i0]:.1f}, Cl amplitude {st['amp']:.4f}, mean Cd "
                f"{out[prec]['cd_mean']:.4f}")
    f, mx = out["fp64"], out["mixed"]
    assert len(f["t"]) == len(mx["t"]) == 401
    dcl = np.abs(f["cl"] - mx["cl"])
    dcd = np.abs(f["cd"] - mx["cd"])
    say(f"  max |dCl| {dcl.max():.2e} (t <= 20: {dcl[f['t'] <= 20].max():.2e}), "
        f"max |dCd| {dcd.max():.2e}, dSt {mx['st']['st'] - f['st']['st']:+.2e}")
    assert dcl[f["t"] <= 20].max() < 1e-6            # before the instability has grown
    assert f["st"]["n_periods"] >= 3
    assert abs(mx["st"]["st"] - f["st"]["st"]) < 2e-3
    assert abs(mx["st"]["amp"] - f["st"]["amp"]) < 0.02 * f["st"]["amp"]
    assert abs(mx["cd_mean"] - f["cd_mean"]) < 1e-3 * f["cd_mean"]
    assert 0.14 < f["st"]["st"] < 0.17


def test_unsteady_c2_reference_grid_strouhal(v):
    """C2 on the reference's own grid, generate_cylinder(181, 181): Re 100, BDF2
    dt 0.2 to t = 80, rotation kick 0.5 until t = 2, outflow far field, in FP64
    and mixed precision, against the oracle's run of the same 401 levels
    (tests/golden/c2_oracle_cyl181.npz: reference Fortran assembly and exact
    SuperLU solves, forces by the pinned force integral): Cd, Cl, Cm within 1e-6
    at every level, the final psi, omega within 1e-8 at the stored nodes, and
    the same Strouhal number.  The reference publishes St 0.1535 for Re 100
    (docs/verification.rst:181-204, run length and grid not stated) and
    "St ~ 0.16, mean Cd ~ 1.31" for this setting on a 131 x 131 grid to t = 50
    (examples/cylinder_unsteady.py:9); the saturated cycle here gives St 0.1603
    and mean Cd 1.315."""
    from lco import strouhal
    ref = golden("c2_oracle_cyl181")
    st_ref = strouhal(ref["t"], ref["cl"])
    xg, yg = v.generate_cylinder(181, 181, 0.5, 50.0)
    for prec in ("fp64", "mixed"):
        with tempfile.TemporaryDirectory() as tmp:
            say(f"unsteady C2 on generate_cylinder(181, 181), {prec}")
            s, d = _unsteady_c2(v, prec, xg, yg, tmp, t_end=80.0)
        assert len(d["t"]) == len(ref["t"]) == 401
        assert np.allclose(d["t"], ref["t"], rtol=0, atol=1e-9)   # forces.csv holds %.10e
        dev = {k: float(np.abs(d[k] - ref[k]).max()) for k in ("cd", "cl", "cm")}
        psi, ome = s.psi, s.ome
        dpsi = float(np.abs(psi[ref["idx"]] - ref["psi_s"]).max())
        dome = float(np.abs(ome[ref["idx"]] - ref["ome_s"]).max())
        st = strouhal(d["t"], d["cl"])
        cd_mean = float(d["cd"][st["i0"]:].mean())
        say(f"  St {st['st']:.4f} (oracle {st_ref['st']:.4f}; {st['n_periods']} periods from "
            f"t={d['t'][st['i0']]:.1f}), mean Cd {cd_mean:.4f}, Cl amplitude {st['amp']:.4f}; vs the "
            f"oracle over 401 levels: max |dCd| {dev['cd']:.1e}, |dCl| {dev['cl']:.1e}, "
            f"|dCm| {dev['cm']:.1e}; final |dpsi| {dpsi:.1e}, |dome| {dome:.1e}")
        assert max(dev.values()) <= 1e-6, (prec, dev)
        assert dpsi <= 1e-8 and dome <= 1e-8, (prec, dpsi, dome)
        assert abs(st["st"] - st_ref["st"]) <= 1e-6
        assert abs(np.linalg.norm(psi) - float(ref["psi_norm"])) <= 1e-8 * float(ref["psi_norm"])
        assert st["n_periods"] >= 4
        assert abs(st["st"] - 0.1535) <= 0.05 * 0.1535
        assert abs(cd_mean - 1.31) <= 0.02 * 1.31
