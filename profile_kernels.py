"""Fixed workloads for hardware-counter and kernel-trace runs under rocprofv3.

    python profile_kernels.py newton   [--size 1025] [--precision mixed] [--solver multifrontal]
    python profile_kernels.py factor   [--size 1025] [--precision mixed] [--solver block_lu]
    python profile_kernels.py assembly [--size 1025] [--launches 20]
    python profile_kernels.py copy     [--mib 256] [--launches 5]

newton:   the headline solve: pure Newton from the impulsive start to res <= 1e-10
          and ||R||_2 / sqrt(2N) < 1e-10 (mixed: the benchmark's GMRES budget,
          tolerance and forcing), after one warm-up solve.
factor:   one factorisation of the steady cylinder Jacobian (initial state,
          Re 40) after one warm-up factorisation (multifrontal: and one solve,
          both event-timed per kernel group, with their algorithmic counts).
assembly: repeated FP32 device assemblies of the same case.
copy:     a plain device-to-device copy of a known size (calibrates FETCH_SIZE /
          WRITE_SIZE against bytes that are certainly moved once).
"""
from __future__ import annotations

import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))   # the repository root


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["newton", "factor", "assembly", "copy"])
    ap.add_argument("--solver", default="multifrontal", choices=["multifrontal", "block_lu"])
    ap.add_argument("--graph", default="apply", choices=["off", "apply", "iteration"],
                    help="newton, mixed: GMRES launch mode (captured apply, captured iterations or eager)")
    ap.add_argument("--no-graph", dest="graph", action="store_const", const="off",
                    help="newton, mixed: launch every kernel eagerly "
                         "(kernel tracing does not follow captured-graph replays)")
    ap.add_argument("--gmres-max", type=int, default=10,
                    help="newton, mixed: GMRES budget on a stored factorisation (the benchmark's)")
    ap.add_argument("--size", type=int, default=1025)
    ap.add_argument("--jmax", type=int, default=0,
                    help="radial size (default: --size); the factor launches of a "
                         "1025 x 16 grid have the shapes of the 1025^2 ones")
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--re", type=float, default=40.0)
    a = ap.parse_args(argv)
    if a.what == "copy":
        import torch
        n = a.mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(a.launches):
            y.copy_(x)
        torch.cuda.synchronize()
        print(f"copy: {a.launches} x {a.mib} MiB", flush=True)
        return 0
    from vorti2d_amd import DeviceContext, generate_annulus
    jmax = a.jmax or a.size
    xg, yg = generate_annulus(a.size, jmax, 0.5, 50.0)
    ctx = DeviceContext(xg, yg, 0, 1.0, 0.0, a.precision)
    ctx.set_linear_solver(a.solver)
    if a.what == "newton":
        import numpy as np
        tol = 1e-10
        if a.precision == "mixed":
            ctx.set_solver(refactor_every=0, gmres_rtol=1e-6, gmres_max_iter=a.gmres_max)
            ctx.set_forcing(0.25, tol * np.sqrt(2 * ctx.N))
            ctx.set_graph(a.graph)
        for k in range(2):
            ctx.init_state()
            ctx.timers(reset=True)
            r = ctx.steady_newton(a.re, tol=tol, rms_tol=tol, max_iter=30, record=True)
            t = ctx.timers()
            print(f"newton {a.size}x{jmax} {a.precision} {a.solver} solve {k}: {r.iters} steps, "
                  f"{t['factorizations']} factorisations, {sum(r.lin_iters)} GMRES iterations, "
                  f"{t['step_ms']:.1f} ms (factor {t['factor_ms']:.1f}, solve {t['solve_ms']:.1f})",
                  flush=True)
        return 0
    ctx.init_state()
    ctx.assemble(a.re)
    if a.what == "factor" and a.solver == "multifrontal":
        # a warm-up factorisation, then one event-timed factorisation and solve:
        # the algorithmic counts of every kernel group of that pair
        ctx.factor()
        p = ctx.profile_mf()
        print(f"factor {a.size}x{jmax} {a.precision} multifrontal: factorisation {p['factor_ms']:.2f} ms, "
              f"solve {p['solve_ms']:.3f} ms, {ctx.mf_info()['flops']:.3e} flop", flush=True)
        for k, g in p["groups"].items():
            print(f"  {g['kernel']:32s} launches {g['launches']:5d}  {g['ms']:9.3f} ms  "
                  f"algorithmic_bytes={g['bytes']:.0f} algorithmic_flops={g['flops']:.0f}", flush=True)
        return 0
    if a.what == "factor":
        ctx.factor()
        p = ctx.profile_factor()
        print(f"factor {a.size}x{jmax} {a.precision}: {p['factor_ms']:.1f} ms, "
              f"{p['launches']} update launches, {p['update_ms'] / p['launches'] * 1e3:.2f} us avg, "
              f"algorithmic_bytes_per_update_launch={p['bytes'] / p['launches']:.0f}", flush=True)
    else:
        for _ in range(a.launches):
            ctx.assemble(a.re)
        ms = ctx.time_assembly(a.re, launches=a.launches)
        print(f"assembly {a.size}x{jmax} {a.precision}: {ms * 1e3:.2f} us per launch", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
