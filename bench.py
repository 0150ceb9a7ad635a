"""Benchmark: steady Newton solve of the 1025x1025 cylinder O-grid on one GPU.

Workload (the headline configuration): O-grid 1025x1025 around a cylinder of
radius 0.5 to radius 50, geometric radial stretching, uniform theta, seam
stored twice; Re = 40, steady, alpha = 0, Dirichlet far field, pure Newton
(dtau = inf), initial state of vorti2d/solver.py:103-114.

A "step" is one Newton solve from that state to convergence.  The stopping
rule is the reference's own (vorti2d/solver.py:134-141): iterate while
res = ||delta||_2 / (2N) > 1e-10, delta the Newton update; in addition the FP64
residual of the final state must satisfy ||R||_2 / (2N) < 1e-10 (the unscaled
||R||_2 of a converged 1025^2 solve sits at its FP64 round-off floor, about
8e-10, in both precisions).  Each Newton iteration is device assembly +
linear solve + update; the solve starts with no stored factorisation.  Mesh
upload, metrics, pattern build and kernel autotuning are one-time setup outside
the timed region.  The mixed-precision solve (FP32 Jacobian, FP32 block-LU
preconditioned FP32 GMRES, FP64 residual / state) is the reported value; the
plain FP64 exact-Newton solve is timed beside it.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Newton solve s to ||R||<1e-10 (mixed vs FP64, 1 GPU); assembly HBM GB/s"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
TOL = 1e-10


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def assembly_bytes(imax, jmax, nnz, value_bytes, n_state=2):
    """Algorithmic HBM bytes of one assembly: 8 metric planes + n_state state
    planes + 2 FP64 RHS planes per node, the Jacobian values, x/y on the two
    boundary rows."""
    N = imax * jmax
    return N * (8 * 8 + 8 * n_state + 2 * 8) + value_bytes * nnz + 16 * 2 * imax


def spmv_bytes(imax, jmax, nnz, value_bytes, vec_bytes):
    """Stencil SpMV with implied columns: values + x + y."""
    return value_bytes * nnz + 2 * (2 * imax * jmax) * vec_bytes


def newton_solve(ctx, re, max_iter=30):
    """One timed Newton solve from the initial state (no stored factorisation).
    Returns a dict with the wall seconds and per-iteration histories."""
    ctx.init_state()          # also drops any stored factorisation
    n2 = 2 * ctx.N
    rn, res_hist, lin, fac = [], [], [], []
    t0 = time.perf_counter()
    it = 0
    r_final = float("nan")
    nf = ctx.timers()["factorizations"]
    while it < max_iter:
        res, bn, li = ctx.newton_step(re)
        it += 1
        rn.append(bn)
        res_hist.append(res)
        lin.append(li)
        nf1 = ctx.timers()["factorizations"]
        fac.append(nf1 - nf)
        nf = nf1
        if res <= TOL:
            ctx.assemble(re)
            r_final = ctx.rhs_norm()
            if r_final / n2 < TOL:
                break
    else:
        ctx.assemble(re)
        r_final = ctx.rhs_norm()
    dt = time.perf_counter() - t0
    rn.append(r_final)
    return dict(seconds=dt, iters=it, rnorm=rn, res=res_hist, gmres=lin, factorizations=fac)


def traffic_table():
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE,
    scaled by the counters' reading for a copy kernel of known size (kept as a
    small JSON summary beside the profiles); None when absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def cpu_baseline(imax, jmax, n_newton, re):
    """Bounded CPU sample of the same workload, one host thread:
    the reference Fortran assemble_coo (oracle/_ref; the oracle's C restatement
    when that is absent) once at the full size, and the exact linear solve
    (COO -> CSC + scipy SuperLU factor/solve, the direct-LU class of the
    reference's MUMPS path) on the same annulus at 257^2 and 361^2, extended to
    the full size by the power law in N fitted through the two points.  The
    value is n_newton x (assembly + extrapolated solve)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    core = oracle.load_ref()
    kind = "reference"
    if core is None:
        core = oracle.load_c()
        kind = "port"

    def case(n, solve):
        xg, yg = oracle.generate_annulus(n, n)
        dk = 1.0 / (n - 1)
        m = core.compute_metrics(dk, dk, xg, yg)
        N = n * n
        psi, ome = oracle.initial_state(n, n, m[8], m[9])
        z = np.zeros(N)
        t0 = time.perf_counter()
        ci, cj, cv, nnz, b = core.assemble_coo(n, n, re, 0.0, 0.0, 0.0, dk, dk, *m, psi, ome, z, z,
                                               0, 2 * N, 13 * 2 * N, 0, 1.0, 0.0)
        t_asm = time.perf_counter() - t0
        if not solve:
            return t_asm, 0.0
        t0 = time.perf_counter()
        A = sp.csc_matrix((cv[:nnz], (ci[:nnz], cj[:nnz])), shape=(2 * N, 2 * N))
        spla.spsolve(A, b)
        return t_asm, time.perf_counter() - t0

    t_asm_full, _ = case(imax, False)
    pts = [(n, case(n, True)[1]) for n in (257, 361)]
    (n1, t1), (n2, t2) = pts
    p = np.log(t2 / t1) / np.log((n2 * n2) / (n1 * n1))
    t_solve_full = t2 * ((imax * jmax) / (n2 * n2)) ** p
    per_iter = t_asm_full + t_solve_full
    return dict(value=per_iter * n_newton, unit="s", cores=1, kind=kind,
                sample=(f"{'reference Fortran' if kind == 'reference' else 'C restatement of'} "
                        f"assemble_coo at {imax}x{jmax} once ({t_asm_full:.2f} s); COO->CSC + "
                        f"SuperLU solve at 257^2 ({t1:.2f} s) and 361^2 ({t2:.2f} s), power law "
                        f"N^{p:.2f} extended to {imax}x{jmax} ({t_solve_full:.0f} s); "
                        f"x {n_newton} Newton iterations; 1 thread"))


def adjoint_timing(v, reps=3):
    """The adjoint workload: NACA 4-digit O-grid 257x129, Re 500, alpha 5 deg,
    32 Hicks-Henne amplitudes ~ N(0, 1e-3); wall time of the FP64 steady solve
    and of one adjoint gradient (dJ/d(alpha, Re, 32 amplitudes) for one
    objective), median of `reps`."""
    from vorti2d_amd.adjoint import steady_solve
    grid = v.AirfoilGrid(257, 129, n_bumps=32)
    amps = np.random.default_rng(0).normal(0.0, 1e-3, 32)
    xg, yg = grid.grid(amps)
    fw, ad = [], []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        r = steady_solve(xg, yg, 500.0, 5.0)
        t1 = time.perf_counter()
        g = r.ctx.adjoint_forces(500.0, 1.0, 1.0, 0.0)
        grid.bump_gradient(g["gx"], g["gy"])
        t2 = time.perf_counter()
        fw.append(t1 - t0)
        ad.append(t2 - t1)
        r.ctx.close()
    fw, ad = statistics.median(fw[1:]), statistics.median(ad[1:])
    log(f"[adjoint] 257x129 Re 500: forward {fw * 1e3:.1f} ms ({r.iters} Newton iterations), "
        f"adjoint gradient {ad * 1e3:.1f} ms, Cd {r.cd:.6f} Cl {r.cl:.6f}")
    return {"workload": "NACA 4-digit (t 0.12, m 0.02, p 0.4) O-grid 257x129 to r 30, Re 500, "
                        "alpha 5 deg, 32 Hicks-Henne amplitudes",
            "forward_s": fw, "adjoint_gradient_s": ad, "params": 34, "newton_iters": r.iters,
            "central_fd_equivalent_s": 2 * 34 * fw, "cd": r.cd, "cl": r.cl}


def sweep_timing(workers=8, n=64):
    """The sweep workload: 64 elliptic O-grids from numpy default_rng(0) (angular
    129..257, radial 129..385 nodes, aspect 0.5..1, Re 20..45, alpha -10..10 deg),
    each a mixed-precision steady solve to res <= 1e-10 plus the adjoint
    dCd/dalpha, up to `workers` contexts (streams) in flight.  One untimed pass
    of the first `workers` cases warms the allocator and the code objects."""
    from vorti2d_amd.batch import solve_batch, sweep_cases
    cases = sweep_cases(n, seed=0)
    solve_batch(cases[:workers], workers=workers, adjoint=True)
    res, st = solve_batch(cases, workers=workers, adjoint=True)
    log(f"[sweep] {n} elliptic O-grids, {workers} streams: {st['solves_per_s']:.2f} solves/s, "
        f"p50 {st['p50_s'] * 1e3:.0f} ms, p99 {st['p99_s'] * 1e3:.0f} ms")
    return {"workload": f"{n} elliptic O-grids (default_rng(0): angular 129..257, radial 129..385, "
                        "aspect 0.5..1, Re 20..45, alpha -10..10 deg), mixed steady solve to "
                        "res <= 1e-10 + adjoint dCd/dalpha per case",
            "solves_per_s": st["solves_per_s"], "wall_s": st["wall_s"], "p50_s": st["p50_s"],
            "p99_s": st["p99_s"], "streams": workers,
            "newton_iters_mean": sum(r.newton_iters for r in res) / len(res),
            "max_res": max(r.res for r in res)}


def run_precision(v, xg, yg, prec, a):
    t0 = time.perf_counter()
    ctx = v.DeviceContext(xg, yg, 0, 1.0, 0.0, prec)
    if prec == "mixed":
        ctx.set_solver(refactor_every=a.refactor_every, gmres_rtol=a.gmres_rtol,
                       gmres_max_iter=a.gmres_max)
        # residual target: the final-state acceptance ||R||_2 / (2N) < TOL
        ctx.set_forcing(a.forcing, TOL * 2 * xg.size if a.forcing_target else 0.0)
        ctx.set_graph(a.graph)
    ctx.init_state()
    tune = ctx.autotune_update_kernel(a.re)
    setup = time.perf_counter() - t0
    log(f"[{prec}] setup {setup:.2f} s (update-kernel autotune "
        + ", ".join(f"{k}: {t:.3f} s" for k, t in tune.items())
        + f"), device memory {ctx.device_bytes() / 2**30:.1f} GiB")
    nsteps = a.steps if prec == "mixed" else a.fp64_steps
    for _ in range(a.warmup):
        r = newton_solve(ctx, a.re)
        log(f"[{prec}] warmup solve {r['seconds']:.3f} s, {r['iters']} iterations")
    ctx.timers(reset=True)
    runs = []
    for s in range(nsteps):
        r = newton_solve(ctx, a.re)
        runs.append(r)
        log(f"[{prec}] solve {s}: {r['seconds']:.3f} s, {r['iters']} Newton iterations; "
            f"||R||_2 " + " ".join(f"{x:.2e}" for x in r["rnorm"])
            + "; res " + " ".join(f"{x:.1e}" for x in r["res"])
            + f"; GMRES iterations {r['gmres']}; factorisations {r['factorizations']}")
    tm = ctx.timers()
    fc = ctx.forces(a.re)
    out = dict(seconds=statistics.median(r["seconds"] for r in runs),
               times=[r["seconds"] for r in runs], last=runs[-1], cd=float(fc[0]),
               cl=float(fc[1]), timers=tm, setup_s=setup, tune=tune)
    if prec == "mixed":
        out["assembly_ms"] = ctx.time_assembly(a.re, launches=50)
        out["spmv_ms"] = ctx.time_spmv(launches=50)
        out["nnz"] = ctx.nnz()
    ctx.assemble(a.re)
    out["probe"] = ctx.profile_factor()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed mixed-precision solves")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=2, help="timed FP64 solves")
    ap.add_argument("--size", type=int, default=1025)
    ap.add_argument("--re", type=float, default=40.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--refactor-every", type=int, default=0,
                    help="mixed: refactor the FP32 block LU every N Newton steps "
                         "(0 = only when GMRES on the stored one misses its tolerance)")
    ap.add_argument("--gmres-max", type=int, default=30)
    ap.add_argument("--gmres-rtol", type=float, default=1e-6)
    ap.add_argument("--forcing", type=float, default=0.25,
                    help="mixed: Eisenstat-Walker GMRES tolerance capped at this value (0: off)")
    ap.add_argument("--no-forcing-target", dest="forcing_target", action="store_false",
                    help="mixed: no residual-target safeguard on the forcing term")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="mixed: launch the block-LU apply kernels one by one instead of replaying a "
                         "captured graph (same kernels; for kernel-trace profilers)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "vorti2d_amd runs on one GPU only; the sharded design "
                          "(circumferential slabs) is not built", "n_gpus": a.gpus}))
        sys.exit(2)

    import vorti2d_amd as v

    imax = jmax = a.size
    xg, yg = v.generate_annulus(imax, jmax, 0.5, 50.0)
    out = {p: run_precision(v, xg, yg, p, a) for p in (("mixed",) if a.no_fp64 else ("mixed", "fp64"))}

    mx = out["mixed"]
    nnz = mx["nnz"]
    N = imax * jmax
    asm_b = assembly_bytes(imax, jmax, nnz, 4)
    asm_gbs = asm_b / (mx["assembly_ms"] * 1e-3) / 1e9
    spmv_b = spmv_bytes(imax, jmax, nnz, 4, 4)
    spmv_gbs = spmv_b / (mx["spmv_ms"] * 1e-3) / 1e9
    pr = mx["probe"]
    upd_avg_ms = pr["update_ms"] / max(pr["launches"], 1)
    upd_tf = pr["flops"] / (pr["update_ms"] * 1e-3) / 1e12
    traffic = traffic_table()
    kname = {"mfma64": "k_gj_step_mfma_b<float>", "valu64": "k_gj_update_b<float, 64>",
             "valu128": "k_gj_update_b<float, 128>"}[min(mx["tune"], key=mx["tune"].get)]
    alg_upd = pr["bytes"] / max(pr["launches"], 1)
    ent = traffic.get(kname) or {}
    tr_upd = ent["measured_over_algorithmic"] * alg_upd if "measured_over_algorithmic" in ent else None
    ent = traffic.get("k_assemble<float>") or {}
    tr_asm = ent.get("bytes_per_launch")
    tm = mx["timers"]
    last = mx["last"]
    res = {
        "metric": METRIC,
        "value": mx["seconds"],
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": mx["seconds"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 (f32 Jacobian, block LU, GMRES; f64 residual, state, refinement)",
        "data": "synthetic (closed-form O-grid, reference initial state)",
        "config": {"workload": f"steady cylinder {imax}x{jmax} O-grid (r 0.5..50, geometric "
                               f"radial ratio {v.mesh.stretch_ratio(jmax):.6f}), Re {a.re:g}, "
                               "pure Newton from the impulsive-start state",
                   "imax": imax, "jmax": jmax, "dof": 2 * N, "nnz": nnz,
                   "criterion": "reference res = ||delta||_2/(2N) < 1e-10 and FP64 "
                                "||R||_2/(2N) < 1e-10 at the final state",
                   "linear_solver": (f"FP32 block-LU preconditioned FP32 GMRES (rtol "
                                     f"{a.gmres_rtol:g}"
                                     + (f", forcing <= {a.forcing:g}" if a.forcing > 0 else "")
                                     + (", forcing >= 0.5 (1e-10 * 2N) / ||R_k|| (residual target)"
                                        if a.forcing > 0 and a.forcing_target else "")
                                     + ", refactor "
                                     + ("on GMRES miss" if a.refactor_every == 0 else
                                        f"every {a.refactor_every}") + ")")},
        "newton_iters": last["iters"],
        "gmres_iters": last["gmres"],
        "factorizations_per_solve": tm["factorizations"] / max(len(mx["times"]), 1),
        "rnorm_final": last["rnorm"][-1],
        "rnorm_final_scaled": last["rnorm"][-1] / (2 * N),
        "res_final": last["res"][-1],
        "cd": mx["cd"],
        "roofline": {"bound": "mfma", "achieved": upd_tf, "peak": MFMA_PEAK_TFLOPS["fp32"],
                     "unit": "TFLOP/s", "frac": upd_tf / MFMA_PEAK_TFLOPS["fp32"],
                     "traffic": tr_upd,
                     "kernel": kname,
                     "avg_launch_us": upd_avg_ms * 1e3, "launches_per_factorization": pr["launches"],
                     "flops_per_factorization": pr["flops"],
                     "algorithmic_bytes_per_launch": alg_upd,
                     "share_of_factorization": pr["update_ms"] / pr["factor_ms"]},
        "assembly": {"bound": "hbm", "achieved": asm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": asm_gbs / HBM_PEAK_GBS, "traffic": tr_asm,
                     "kernel": "k_assemble<float>",
                     "bytes_per_launch": asm_b, "avg_launch_us": mx["assembly_ms"] * 1e3},
        "spmv": {"bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": spmv_gbs / HBM_PEAK_GBS, "bytes_per_launch": spmv_b,
                 "avg_launch_us": mx["spmv_ms"] * 1e3},
        "phase_ms_per_solve": {
            k: tm[k + "_ms"] / max(len(mx["times"]), 1) for k in ("assemble", "factor", "solve")},
        "times_s": mx["times"],
        "setup_s": mx["setup_s"],
        "update_kernel_autotune_s": mx["tune"],
    }
    if "fp64" in out:
        f = out["fp64"]
        res["fp64"] = {"seconds": f["seconds"], "times_s": f["times"],
                       "newton_iters": f["last"]["iters"], "rnorm_final": f["last"]["rnorm"][-1],
                       "cd": f["cd"], "factorization_ms": f["probe"]["factor_ms"],
                       "update_kernel_tflops": f["probe"]["flops"] / (f["probe"]["update_ms"] * 1e-3)
                       / 1e12}
        res["mixed_speedup_vs_fp64"] = f["seconds"] / mx["seconds"]
        res["cd_mixed_minus_fp64"] = mx["cd"] - f["cd"]
    if not a.no_adjoint:
        res["adjoint_c4"] = adjoint_timing(v)
    if not a.no_sweep:
        res["sweep_c5"] = sweep_timing()
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(imax, jmax, last["iters"], a.re)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
