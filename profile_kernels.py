"""Fixed workloads for hardware-counter and kernel-trace runs under rocprofv3.

    python profile_kernels.py factor   [--size 1025] [--precision mixed]
    python profile_kernels.py assembly [--size 1025] [--launches 20]
    python profile_kernels.py copy     [--mib 256] [--launches 5]

factor:   one block-LU factorisation of the steady cylinder Jacobian (initial
          state, Re 40) after one warm-up factorisation.
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
    ap.add_argument("what", choices=["factor", "assembly", "copy"])
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
    ctx.init_state()
    ctx.assemble(a.re)
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
