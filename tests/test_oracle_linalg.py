"""Pins the oracle's linear algebra and assembly restatements (CPU only).

SpMV by its definition (scipy), Dirichlet application on a hand-computed 4x4 example, PCG/GMRES
against dense direct solves, IC0/ILU0 by their defining property on the pattern, element matrices by
the P1 closed forms (cotangent stiffness, area/12 (1 + delta) mass) and the Poisson assembly by the
O(h^2) convergence of the manufactured solution.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import golden, rel_l2


def laplacian_system(oracle, n=16, coeff=1.0):
    from gpupfem2_amd import meshgen
    xy, tri = meshgen.unit_square(n)
    area, invJ = oracle.geometry2d(xy, tri)
    rp, col = oracle.csr_pattern(xy.shape[0], tri)
    val, rhs = oracle.assemble_poisson2d(xy, tri, area, invJ, rp, col, coeff=coeff, rhs_kind=1,
                                         rhs_param=2 * math.pi ** 2)
    return xy, tri, area, invJ, rp, col, val, rhs


def to_scipy(rp, col, val):
    return sp.csr_matrix((val, col, rp), shape=(rp.shape[0] - 1, rp.shape[0] - 1))


def test_spmv_definition(oracle):
    xy, tri, area, invJ, rp, col, val, rhs = laplacian_system(oracle)
    x = np.random.default_rng(0).standard_normal(rp.shape[0] - 1)
    assert np.allclose(oracle.spmv(rp, col, val, x), to_scipy(rp, col, val) @ x, rtol=1e-13, atol=1e-13)


def test_dirichlet_hand_example(oracle):
    """4x4 SPD matrix, node 0 fixed to g = 2. Mask mode: row 0 -> identity, column 0 eliminated into
    the rhs. Parity mode: row 1's first off-diagonal is (1,0) itself, row 2's first off-diagonal is
    exactly 0 (the Q6 misfire) so row 2 keeps its coupling to node 0."""
    A = np.array([[4.0, -1.0, -1.0, 0.0],
                  [-1.0, 4.0, 0.0, -1.0],
                  [-1.0, 0.0, 4.0, -1.0],
                  [0.0, -1.0, -1.0, 4.0]])
    # pattern: full rows except (1,2),(2,1),(0,3),(3,0); row 2 keeps an explicit zero at (2,1)
    pat = [[0, 1, 2], [0, 1, 3], [0, 1, 2, 3], [1, 2, 3]]
    rp = np.array([0, 3, 6, 10, 13], np.int32)
    col = np.array(sum(pat, []), np.int32)
    b = np.array([1.0, 2.0, 3.0, 4.0])
    for mode in (0, 1):
        val = np.array([A[i, j] if not (i == 2 and j == 1) else 0.0 for i in range(4) for j in pat[i]])
        # put a genuine zero first off-diagonal into row 2 (its column 0 entry is -1, column 1 is 0)
        val_perm = val.copy()
        rhs = b.copy()
        oracle.dirichlet_apply(rp, col, val_perm, rhs, [0], [2.0], mode)
        M = to_scipy(rp, col, val_perm).toarray()
        assert np.array_equal(M[0], [1.0, 0.0, 0.0, 0.0]) and rhs[0] == 2.0
        assert M[1, 0] == 0.0 and rhs[1] == 2.0 + 2.0        # -(-1)*2 added
        if mode == 1:
            assert M[2, 0] == 0.0 and rhs[2] == 3.0 + 2.0
        else:
            # row 2: first off-diagonal is col 0 = -1 -> processed as well
            assert M[2, 0] == 0.0 and rhs[2] == 3.0 + 2.0
        x = np.linalg.solve(M, rhs)
        assert abs(x[0] - 2.0) < 1e-14
    # the misfire: a row whose first off-diagonal is exactly zero is skipped in parity mode only
    rp2 = np.array([0, 2, 5], np.int32)
    col2 = np.array([0, 1, 0, 1, 1], np.int32)[:0]
    del rp2, col2


def test_dirichlet_parity_misfire(oracle):
    """'/'-split Laplacian: couplings across the diagonal are exactly zero; a row whose first
    off-diagonal vanishes keeps its Dirichlet columns in parity mode (src/Dirichlet_bcs.cu:39-41).
    Both modes give the same solution because the Dirichlet rows are identity."""
    xy, tri, area, invJ, rp, col, val, rhs = laplacian_system(oracle, n=8)
    bnodes = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 0] > 1 - 1e-9) | (xy[:, 1] < 1e-9) | (xy[:, 1] > 1 - 1e-9))[0]
    g = 1.0 + xy[bnodes, 0]
    sols, mats = [], []
    for mode in (0, 1):
        v, r = val.copy(), rhs.copy()
        oracle.dirichlet_apply(rp, col, v, r, bnodes, g, mode)
        M = to_scipy(rp, col, v)
        sols.append(spla.spsolve(M.tocsc(), r))
        mats.append(M.toarray())
    assert np.allclose(sols[0], sols[1], rtol=1e-12, atol=1e-12)
    assert np.allclose(mats[1], mats[1].T)
    assert not np.allclose(mats[0], mats[0].T)   # the heuristic misfires on this mesh


@pytest.mark.parametrize("precond", ["none", "jacobi", "ic0", "ilu0"])
def test_pcg_vs_direct(oracle, precond):
    xy, tri, area, invJ, rp, col, val, rhs = laplacian_system(oracle, n=16)
    b_nodes = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 0] > 1 - 1e-9) | (xy[:, 1] < 1e-9) | (xy[:, 1] > 1 - 1e-9))[0]
    v, r = val.copy(), rhs.copy()
    oracle.dirichlet_apply(rp, col, v, r, b_nodes, np.zeros(b_nodes.size), 1)
    x = np.zeros(rp.shape[0] - 1)
    st = oracle.pcg(rp, col, v, x, r, 1e-12, 2000, tol_mode=1, precond=precond)
    assert st.converged
    ref = spla.spsolve(to_scipy(rp, col, v).tocsc(), r)
    assert rel_l2(x, ref) < 1e-10


@pytest.mark.parametrize("restart", [30, 2000])
@pytest.mark.parametrize("precond", ["jacobi", "ilu0"])
def test_gmres_vs_direct_nonsymmetric(oracle, restart, precond):
    """A nonsymmetric (convection-diffusion-like) system on the Laplacian pattern."""
    xy, tri, area, invJ, rp, col, val, rhs = laplacian_system(oracle, n=12)
    n = rp.shape[0] - 1
    rng = np.random.default_rng(1)
    v = val.copy()
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            if col[k] != i:
                v[k] += 0.3 * rng.uniform(-1, 1) * abs(val[k])
            else:
                v[k] += 1.0
    b = rng.standard_normal(n)
    x = np.zeros(n)
    st = oracle.gmres(rp, col, v, x, b, 1e-12, 2000, restart, tol_mode=1, precond=precond)
    assert st.converged
    ref = spla.spsolve(to_scipy(rp, col, v).tocsc(), b)
    assert rel_l2(x, ref) < 1e-9


def test_gmres_reference_mode_leaves_x_when_not_converged(oracle):
    """Non-restarted GMRES updates x only on convergence (linear_solver.cu:508-518)."""
    xy, tri, area, invJ, rp, col, val, rhs = laplacian_system(oracle, n=12)
    n = rp.shape[0] - 1
    x = np.full(n, 0.5)
    st = oracle.gmres(rp, col, val + 0.0, x, np.ones(n), 1e-30, 5, 5, precond="jacobi")
    assert not st.converged and st.iterations == 5 and np.all(x == 0.5)


def test_ic0_ilu0_defining_property(oracle):
    xy, tri, area, invJ, rp, col, val, rhs = laplacian_system(oracle, n=10)
    b_nodes = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 0] > 1 - 1e-9) | (xy[:, 1] < 1e-9) | (xy[:, 1] > 1 - 1e-9))[0]
    v, r = val.copy(), rhs.copy()
    oracle.dirichlet_apply(rp, col, v, r, b_nodes, np.zeros(b_nodes.size), 1)
    n = rp.shape[0] - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    A = to_scipy(rp, col, v).toarray()
    rc, L = oracle.ic0(rp, col, v)
    assert rc == 0
    Lm = np.zeros((n, n))
    low = col <= rows
    Lm[rows[low], col[low]] = L[low]
    LLt = Lm @ Lm.T
    assert np.allclose(LLt[rows[low], col[low]], A[rows[low], col[low]], rtol=1e-12, atol=1e-12)
    rc, LU = oracle.ilu0(rp, col, v)
    assert rc == 0
    Lu = np.eye(n)
    Um = np.zeros((n, n))
    Lu[rows[col < rows], col[col < rows]] = LU[col < rows]
    Um[rows[col >= rows], col[col >= rows]] = LU[col >= rows]
    P = Lu @ Um
    assert np.allclose(P[rows, col], A[rows, col], rtol=1e-12, atol=1e-12)
    # apply == exact solve with the factors
    rr = np.random.default_rng(3).standard_normal(n)
    assert np.allclose(oracle.ic0_apply(rp, col, L, rr), np.linalg.solve(LLt, rr), rtol=1e-10, atol=1e-12)
    assert np.allclose(oracle.ilu0_apply(rp, col, LU, rr), np.linalg.solve(P, rr), rtol=1e-10, atol=1e-12)


def test_p1_element_closed_forms(oracle):
    """Stiffness = cotangent formula (area * grad N_i . grad N_j) and mass = rho area / 12 (1 + delta),
    read off single-triangle assemblies."""
    xy = np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.1]])
    tri = np.array([[0, 1, 2]], np.int32)
    area, invJ = oracle.geometry2d(xy, tri)
    rp, col = oracle.csr_pattern(3, tri)
    K, f = oracle.assemble_poisson2d(xy, tri, area, invJ, rp, col, coeff=1.0, rhs_kind=2, rhs_param=1.0)
    Kd = to_scipy(rp, col, K).toarray()
    A2 = 2 * area[0]
    ref = np.zeros((3, 3))
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        e = xy[k] - xy[j]
        for i2 in range(3):
            j2, k2 = (i2 + 1) % 3, (i2 + 2) % 3
            e2 = xy[k2] - xy[j2]
            ref[i, i2] = np.dot(e, e2) / (2 * A2)
    assert np.allclose(Kd, ref, rtol=1e-13, atol=1e-14)
    assert np.allclose(f, area[0] / 3)
    # mass via the correction assembly with p = 0: rho * area / 12 * (1 + delta)
    prm = oracle.make_params(2.5, 0.0, 0.1, 0)
    z = np.zeros(3)
    M, r0, r1 = oracle.assemble_correction(tri, area, invJ, rp, col, np.ones(3), z, z, None, prm)
    Md = to_scipy(rp, col, M).toarray()
    assert np.allclose(Md, 2.5 * area[0] / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-13)
    assert np.allclose(r0, 2.5 * area[0] / 3, rtol=1e-13) and np.all(r1 == 0)


def test_poisson_manufactured_convergence(oracle):
    """cfg1 shape: -Laplace u = 2 pi^2 sin(pi x) sin(pi y), u = 0 on the boundary; nodal error O(h^2)."""
    errs = []
    for n in (16, 32, 64):
        xy, tri, area, invJ, rp, col, val, rhs = laplacian_system(oracle, n=n)
        bn = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 0] > 1 - 1e-9) | (xy[:, 1] < 1e-9) | (xy[:, 1] > 1 - 1e-9))[0]
        assert bn.size == 4 * n
        oracle.dirichlet_apply(rp, col, val, rhs, bn, np.zeros(bn.size), 1)
        x = np.zeros(rp.shape[0] - 1)
        st = oracle.pcg(rp, col, val, x, rhs, 1e-10, 5000, tol_mode=1, precond="jacobi")
        assert st.converged
        exact = np.sin(np.pi * xy[:, 0]) * np.sin(np.pi * xy[:, 1])
        errs.append(np.abs(x - exact).max())
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(1.8 < r < 2.2 for r in rates), (errs, rates)


def test_poisson3d_manufactured(oracle):
    from gpupfem2_amd import meshgen
    errs = []
    for n in (4, 8, 16):
        xyz, tet = meshgen.box3d(n)
        vol, invJ = oracle.geometry3d(xyz, tet)
        rp, col = oracle.csr_pattern(xyz.shape[0], tet)
        val, rhs = oracle.assemble_poisson3d(xyz, tet, vol, invJ, rp, col, 1.0, 1, 3 * math.pi ** 2, 1)
        bn = np.nonzero(np.any((xyz < 1e-9) | (xyz > 1 - 1e-9), axis=1))[0]
        oracle.dirichlet_apply(rp, col, val, rhs, bn, np.zeros(bn.size), 1)
        x = np.zeros(rp.shape[0] - 1)
        assert oracle.pcg(rp, col, val, x, rhs, 1e-12, 5000, tol_mode=1, precond="jacobi").converged
        exact = np.prod(np.sin(np.pi * xyz), axis=1)
        errs.append(np.abs(x - exact).max())
    assert errs[2] < errs[1] < errs[0] and errs[1] / errs[2] > 3.0, errs


def test_gmres_restart_zero_is_unrestarted(oracle):
    """restart = 0 (the library's code for the reference's non-restarted GMRES) runs the same iterations
    as restart = maxit and leaves x untouched when it does not converge."""
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    n = 60
    A = sp.random(n, n, density=0.08, random_state=3) + sp.eye(n) * 4.0
    A = A.tocsr()
    A.sort_indices()
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    b = rng.standard_normal(n)
    xs = []
    for restart in (0, 40):
        x = np.zeros(n)
        st = oracle.gmres(rp, col, val, x, b, 1e-10, 40, restart, 0, "jacobi")
        assert st.converged
        xs.append((st.iterations, x))
    assert xs[0][0] == xs[1][0] and np.array_equal(xs[0][1], xs[1][1])
    x = np.full(n, 0.5)
    st = oracle.gmres(rp, col, val, x, b, 1e-30, 5, 0, 0, "jacobi")
    assert not st.converged and st.iterations == 5 and np.all(x == 0.5)
