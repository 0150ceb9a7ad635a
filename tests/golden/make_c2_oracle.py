"""Write the oracle's unsteady C2 run on the reference grid as a fixture.

Workload: generate_cylinder(181, 181) (r 0.5 -> 50), Re 100, BDF2 dual time
with dt 0.2 from t = 0 to t = 80 (401 levels), rotation kick 0.5 while
t <= 2, outflow far field, pure Newton inside each level (dtau = inf) to
res = ||delta||_2 / (2N) <= 1e-10, the time loop of vorti2d/solver.py:148-224
(level times :160-168, wall speed config.py:162-166, history shift :186-188).
Each Newton step assembles with the reference Fortran `assemble_coo`
(oracle/_ref, built from the reference sources by `make -C oracle ref`; the
oracle's C restatement, bit-identical to it, where that is absent) and solves
exactly with scipy SuperLU.  After every level the force coefficients come
from the oracle's force integral (pinned to the reference's
compute_force_coeffs by forces_ref.npz).

Stored in c2_oracle_cyl181.npz:
  t, cd, cl, cm           per level (401)
  newton_iters, res       per level
  psi_s, ome_s, idx       the final state at every 97th node
  psi_norm, ome_norm      2-norms of the final psi, omega

About 15-25 minutes on one host core.  Run from the repository root:
    python tests/golden/make_c2_oracle.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_oracle_cyl181.npz")
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402

IMAX = JMAX = 181
RE, DT, T_END, ROT, ROT_UNTIL = 100.0, 0.2, 80.0, 0.5, 2.0
FF_OUTFLOW = 1


def main():
    core = oracle.load_ref() or oracle.load_c()
    xg, yg = oracle.generate_cylinder(IMAX, JMAX, 0.5, 50.0)
    dksi, deta = 1.0 / (IMAX - 1), 1.0 / (JMAX - 1)
    m = core.compute_metrics(dksi, deta, xg, yg)
    N = IMAX * JMAX
    psi, ome = oracle.initial_state(IMAX, JMAX, m[8], m[9])
    o1, o2 = np.zeros(N), np.zeros(N)
    nsteps = int(round(T_END / DT)) + 1
    times = [m_ * DT for m_ in range(nsteps)]
    rows = {k: [] for k in ("t", "cd", "cl", "cm", "newton_iters", "res")}
    t0 = time.time()
    for k, t in enumerate(times):
        urot = ROT if t <= ROT_UNTIL else 0.0
        r = oracle.newton_solve(core, IMAX, JMAX, m, RE, inv2dt=1.0 / (2.0 * DT), urot=urot,
                                ff_bc=FF_OUTFLOW, psi=psi, ome=ome, omeold=o1, omeoldold=o2,
                                tol=1e-10, max_iter=100)
        psi, ome = r.psi, r.ome
        o2, o1 = o1, ome.copy()
        f = oracle.force_coeffs(IMAX, JMAX, RE, dksi, deta, m[2], m[3], m[6], m[7], m[8], m[9], ome)
        for key, val in (("t", t), ("cd", f["cd"]), ("cl", f["cl"]), ("cm", f["cm"]),
                         ("newton_iters", r.iters), ("res", r.res)):
            rows[key].append(val)
        if k % 20 == 0 or k == nsteps - 1:
            print(f"level {k:3d} t={t:6.2f} iters={r.iters} res={r.res:.2e} cd={f['cd']:.6f} "
                  f"cl={f['cl']:+.6f}  {time.time() - t0:.0f} s", flush=True)
    idx = np.arange(0, N, 97)
    np.savez(OUT, **{k: np.asarray(v) for k, v in rows.items()}, idx=idx, psi_s=psi[idx],
             ome_s=ome[idx], psi_norm=np.linalg.norm(psi), ome_norm=np.linalg.norm(ome),
             assembly=np.array("reference Fortran" if oracle.load_ref() else "C restatement"))
    print(f"wrote {OUT}", flush=True)


if __name__ == "__main__":
    main()
