"""Write the golden fixtures of tests/golden/ from the reference Fortran.

The reference core (src/fortran/vorti2d_core.f90) is compiled from its own
sources into oracle/_ref by ``make -C oracle ref`` and called here through
ctypes.  Each fixture stores inputs and outputs:

* neighbors_7x5.npz    node_neighbors for every (i, j) of a 7x5 grid
* metrics_*.npz        compute_metrics on reference cylinder / annulus grids
* assemble_*.npz       assemble_coo (all rows) with the canonical smooth state
                       of tests/test_assemble_local.py:23-30 and its scalars
* local_*.npz          compute_metrics_local / assemble_coo_local on a 1-rank
                       ghosted block (ghost width 1, periodic), natural lg map
* newton_41x41_re40.npz  converged steady Re=40 state from the reference
                       assembly + an exact sparse LU per Newton step, and its
                       force coefficients

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def ghost_block(field_imax, imax, jmax):
    """Global (imax-column, seam duplicated) field -> 1-rank ghosted local block
    of gxm = ni + 2 columns (ghost width 1, periodic), flat i fastest."""
    ni = imax - 1
    g2 = field_imax.reshape(jmax, imax)[:, :ni]
    gxs = -1
    L = np.empty((ni + 2, jmax))
    for il in range(ni + 2):
        L[il, :] = g2[:, (gxs + il) % ni]
    return np.ascontiguousarray(L.flatten("F"))


def natural_lg(imax, jmax):
    """Local-to-global dof map of the 1-rank block: local node (il, j) -> global
    node ((il-1) mod ni, j), dof = 2*node + c (node-interleaved, natural order)."""
    ni, gxm = imax - 1, imax + 1
    lg = np.empty(2 * gxm * jmax, np.int32)
    for j in range(jmax):
        for il in range(gxm):
            k = gxm * j + il
            g = j * ni + (il - 1) % ni
            lg[2 * k] = 2 * g
            lg[2 * k + 1] = 2 * g + 1
    return lg


def main():
    ref = O.load_ref()
    if ref is None:
        sys.exit("oracle/_ref not built: run `make -C oracle ref` first")

    nb = np.array([[ref.node_neighbors(i, j, 7) for i in range(1, 8)] for j in range(1, 6)],
                  np.int32)
    np.savez_compressed(os.path.join(OUT, "neighbors_7x5.npz"), imax=7, jmax=5, nb=nb)

    for name, gen, imax, jmax in (("cyl_25x17", O.generate_cylinder, 25, 17),
                                  ("cyl_41x41", O.generate_cylinder, 41, 41),
                                  ("ann_33x21", O.generate_annulus, 33, 21)):
        xg, yg = gen(imax, jmax)
        dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
        m = ref.compute_metrics(dk, de, xg, yg)
        np.savez_compressed(os.path.join(OUT, f"metrics_{name}.npz"), xg=xg, yg=yg,
                            dksi=dk, deta=de, metrics=np.array(m))

    cases = (("cyl_25x17_ff1", O.generate_cylinder, 25, 17,
              dict(re=80.0, invdtau=0.0, inv2dt=1.0 / (2 * 0.2), urot=0.3, ff_bc=1, ca=1.0, sa=0.0)),
             ("cyl_41x41_ff0", O.generate_cylinder, 41, 41,
              dict(re=40.0, invdtau=0.0, inv2dt=0.0, urot=0.0, ff_bc=0, ca=1.0, sa=0.0)),
             ("ann_33x21_ff2_a12", O.generate_annulus, 33, 21,
              dict(re=100.0, invdtau=1.0, inv2dt=2.5, urot=0.5, ff_bc=2,
                   ca=float(np.cos(np.radians(12.0))), sa=float(np.sin(np.radians(12.0))))))
    for name, gen, imax, jmax, sc in cases:
        xg, yg = gen(imax, jmax)
        dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
        m = ref.compute_metrics(dk, de, xg, yg)
        N = imax * jmax
        st = O.smooth_state(m[8], m[9])
        ci, cj, cv, nnz, b = ref.assemble_coo(imax, jmax, sc["re"], sc["invdtau"], sc["inv2dt"],
                                              sc["urot"], dk, de, *m, *st, 0, 2 * N, 26 * N,
                                              sc["ff_bc"], sc["ca"], sc["sa"])
        np.savez_compressed(os.path.join(OUT, f"assemble_{name}.npz"), xg=xg, yg=yg,
                            state=np.array(st), **sc, coo_i=ci[:nnz], coo_j=cj[:nnz],
                            coo_v=cv[:nnz], bvec=b)

    # local (ghosted, node-interleaved) variants on a 1-rank block
    imax, jmax = 25, 17
    xg, yg = O.generate_cylinder(imax, jmax)
    dk, de = 1.0 / (imax - 1), 1.0 / (jmax - 1)
    m = ref.compute_metrics(dk, de, xg, yg)
    gxm = imax + 1
    xgl = ghost_block(m[8], imax, jmax).reshape(jmax, gxm).T
    ygl = ghost_block(m[9], imax, jmax).reshape(jmax, gxm).T
    il0, il1 = 2, gxm - 1
    ml = ref.compute_metrics_local(il0, il1, dk, de, xgl, ygl)
    st = [ghost_block(f, imax, jmax) for f in O.smooth_state(m[8], m[9])]
    lg = natural_lg(imax, jmax)
    nloc = 2 * (imax - 1) * jmax
    sc = dict(re=80.0, invdtau=0.0, inv2dt=1.0 / (2 * 0.2), urot=0.3, ff_bc=1, ca=1.0, sa=0.0)
    loc_m = [ghost_block(f, imax, jmax) for f in m]
    ci, cj, cv, nnz, b = ref.assemble_coo_local(gxm, jmax, il0, il1, sc["re"], sc["invdtau"],
                                                sc["inv2dt"], sc["urot"], dk, de, *loc_m, *st, 0,
                                                nloc, lg, 2 * nloc * 13, sc["ff_bc"], sc["ca"],
                                                sc["sa"])
    np.savez_compressed(os.path.join(OUT, "local_cyl_25x17.npz"), xgl=xgl, ygl=ygl, il0=il0,
                        il1=il1, dksi=dk, deta=de, metrics_local=np.array(ml),
                        loc_metrics=np.array(loc_m), state=np.array(st), lg=lg, nloc=nloc,
                        **sc, coo_i=ci[:nnz], coo_j=cj[:nnz], coo_v=cv[:nnz], bvec=b)

    # converged steady state (reference assembly + exact LU per Newton step)
    imax = jmax = 41
    xg, yg = O.generate_cylinder(imax, jmax)
    m = ref.compute_metrics(1.0 / 40, 1.0 / 40, xg, yg)
    r = O.newton_solve(ref, imax, jmax, m, 40.0, tol=1e-10, max_iter=50)
    fc = O.force_coeffs(imax, jmax, 40.0, 1.0 / 40, 1.0 / 40, m[2], m[3], m[6], m[7], m[8], m[9],
                        r.ome)
    np.savez_compressed(os.path.join(OUT, "newton_41x41_re40.npz"), xg=xg, yg=yg, psi=r.psi,
                        ome=r.ome, iters=r.iters, res_hist=np.array(r.res_hist),
                        rnorm_hist=np.array(r.rnorm_hist),
                        forces=np.array([fc[k] for k in ("cd", "cl", "cm", "cdp", "cdf", "clp",
                                                         "clf", "cmp", "cmf", "ref_length")]))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
