"""Why the device solver's preconditioner is an exact block LU and not the
reference's GMRES + point-Jacobi (`linsolve="gmres_jacobi"`,
vorti2d/dist_solver.py:159-172) or GMRES + ILU (`gmres_asm`, :138-158).

The reference states that pure Newton (dtau = inf) "does not converge
iteratively" with those preconditioners (docs/parallel.rst:61-64) and that
ILU(0) "can fail" on it (vorti2d/config.py:100-104).  This test shows it on the
oracle's exact-Newton Jacobian of the steady Re 40 cylinder
(generate_cylinder(129, 129), the systems of the first and the third Newton
step): right-hand side b = the Newton residual, GMRES(50) with 10 restarts
(500 iterations, scipy), relative tolerance 1e-6.

* point-Jacobi (diagonal scaling) stalls far above 1e-6;
* ILU(0) (incomplete LU on the matrix's own pattern, natural order, no
  pivoting) stalls above 1e-6;
* an exact LU of the FP32-rounded matrix, which is what the FP32 block LU of
  the device path computes, reaches 1e-6 in at most 3 iterations.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _jacobian(C, n, m, psi, ome):
    N = n * n
    d = 1.0 / (n - 1)
    z = np.zeros(N)
    ci, cj, cv, nnz, b = C.assemble_coo(n, n, 40.0, 0, 0, 0, d, d, *m, psi, ome, z, z, 0, 2 * N,
                                        26 * N, 0, 1.0, 0.0)
    A = sp.csr_matrix((cv[:nnz], (ci[:nnz], cj[:nnz])), shape=(2 * N, 2 * N))
    A.sort_indices()
    return A, b


def ilu0(A):
    """Incomplete LU with zero fill (IKJ form on the CSR pattern of A, unit
    lower factor, no pivoting).  Returns (L, U) as CSR matrices."""
    n = A.shape[0]
    ip, ind = A.indptr, A.indices
    val = A.data.copy()
    diag = np.array([ip[i] + np.searchsorted(ind[ip[i]:ip[i + 1]], i) for i in range(n)])
    assert np.array_equal(ind[diag], np.arange(n)), "a zero diagonal in the pattern"
    pos = -np.ones(n, np.int64)
    for i in range(n):
        s, e = ip[i], ip[i + 1]
        pos[ind[s:e]] = np.arange(s, e)
        for kk in range(s, diag[i]):
            k = ind[kk]
            val[kk] /= val[diag[k]]
            ks, ke = diag[k] + 1, ip[k + 1]
            p = pos[ind[ks:ke]]
            hit = p >= 0
            val[p[hit]] -= val[kk] * val[ks:ke][hit]
        pos[ind[s:e]] = -1
    F = sp.csr_matrix((val, ind, ip), shape=A.shape)
    return sp.tril(F, -1, format="csr") + sp.eye(n, format="csr"), sp.triu(F, 0, format="csr")


def test_ilu0_is_exact_on_a_tridiagonal():
    """No fill arises on a tridiagonal matrix, so ILU(0) is its exact LU."""
    rng = np.random.default_rng(1)
    n = 40
    A = sp.diags([rng.uniform(1, 2, n - 1), rng.uniform(4, 5, n), rng.uniform(1, 2, n - 1)],
                 [-1, 0, 1], format="csr")
    L, U = ilu0(A)
    assert np.abs((L @ U - A).toarray()).max() < 1e-14


def _gmres(A, b, M):
    x, _ = spla.gmres(A, b, M=M, rtol=1e-6, atol=0.0, restart=50, maxiter=10)
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b)


def _iters_to(A, b, M):
    hist = []
    spla.gmres(A, b, M=M, rtol=1e-6, atol=0.0, restart=50, maxiter=10,
               callback=lambda r: hist.append(r), callback_type="pr_norm")
    return len(hist)


def test_jacobi_and_ilu0_fail_block_lu_converges(C, oracle_mod):
    n = 129
    xg, yg = oracle_mod.generate_cylinder(n, n)
    d = 1.0 / (n - 1)
    m = C.compute_metrics(d, d, xg, yg)
    psi, ome = oracle_mod.initial_state(n, n, m[8], m[9])
    states = [(psi, ome)]
    r = oracle_mod.newton_solve(C, n, n, m, 40.0, tol=0.0, max_iter=2)
    states.append((r.psi, r.ome))
    for step, (p, o) in zip((0, 2), states):
        A, b = _jacobian(C, n, m, p, o)
        dg = A.diagonal()
        jac = spla.LinearOperator(A.shape, matvec=lambda x, dg=dg: x / dg)
        L, U = ilu0(A)
        Ls = spla.splu(L.tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0.0)
        Us = spla.splu(U.tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0.0)
        ilu = spla.LinearOperator(A.shape, matvec=lambda x, Ls=Ls, Us=Us: Us.solve(Ls.solve(x)))
        lu = spla.splu(A.astype(np.float32).astype(np.float64).tocsc())
        exact32 = spla.LinearOperator(A.shape, matvec=lu.solve)
        r_jac, r_ilu, r_lu = _gmres(A, b, jac), _gmres(A, b, ilu), _gmres(A, b, exact32)
        print(f"Newton step {step}: relative residual after GMRES(50) x 10 restarts: "
              f"Jacobi {r_jac:.2e}, ILU(0) {r_ilu:.2e}, FP32 LU {r_lu:.2e} "
              f"({_iters_to(A, b, exact32)} iterations)")
        assert r_jac > 1e-3
        assert r_ilu > 1e-5
        assert r_lu <= 1e-6 and _iters_to(A, b, exact32) <= 3
