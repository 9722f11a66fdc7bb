"""GPU parity: multi-right-hand-side PCG (pfem_solver_solve_multi).

Each column follows the single-RHS PCG recurrence (linear_solver.cu:301-376) with its own scalars, so it is
compared with the library's single-RHS solve and with the oracle's PCG on the same column: relative L2
<= 1e-10 at tolerance 1e-12 (the dot products are reduced in another order) and iteration counts within 3.
The 3D box of 128^3 cubes with 8 right-hand sides (2.15M unknowns) is checked through size-independent
properties: every column converges, its true residual ||b - A x|| / ||b|| meets the tolerance, and the
mixed-precision solution (PCG with reliable updates: FP32 iterations, FP64 residual replacement) agrees
with the FP64 one to 1e-8 and needs at most 25% more iterations (the recurrence is never restarted).
"""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def system2d(ctx, n=40):
    from gpupfem2_amd import capi, meshgen
    xy, tri = meshgen.unit_square(n)
    m = capi.Mesh2D(ctx, xy, tri)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, rhs_kind=1, rhs_param=2 * math.pi ** 2)
    bn = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 0] > 1 - 1e-9) | (xy[:, 1] < 1e-9) | (xy[:, 1] > 1 - 1e-9))[0]
    capi.Dirichlet(ctx, xy.shape[0], bn, np.zeros(bn.size)).apply(a, rhs, 1)
    rp, col, val = a.download()
    return m, a, rp, col, val, bn


@pytest.mark.parametrize("precond", [1, 0])
@pytest.mark.parametrize("nrhs", [1, 3, 8])
def test_multi_matches_single_and_oracle(ctx, oracle, nrhs, precond):
    from gpupfem2_amd import capi
    print("multi 2d", nrhs, precond, flush=True)
    m, a, rp, col, val, bn = system2d(ctx)
    n = rp.size - 1
    rng = np.random.default_rng(nrhs + 10 * precond)
    B = rng.standard_normal((nrhs, n))
    B[:, bn] = 0.0
    X0 = np.zeros((nrhs, n))
    if nrhs == 8:
        B[5] = 0.0                                    # a zero column: converged at once, x stays 0
        X0[2] = rng.standard_normal(n) * 0.1          # a warm start
        X0[2, bn] = 0.0
    pc = capi.Preconditioner(a, capi.PRECOND_JACOBI) if precond else None
    s = capi.Solver(a, capi.SOLVER_CG, 1e-12, 4000, 0, capi.TOL_RELATIVE, pc)
    X = ctx.to_device(X0.ravel())
    stats = s.solve_multi(a, X, ctx.to_device(B.ravel()), nrhs)
    Xh = X.download().reshape(nrhs, n)
    pname = "jacobi" if precond else "none"
    for j in range(nrhs):
        st = stats[j]
        assert st.converged and not st.nan_detected, j
        if not B[j].any():
            assert st.iterations == 0 and not Xh[j].any()
            continue
        x1 = ctx.to_device(X0[j])
        s1 = capi.Solver(a, capi.SOLVER_CG, 1e-12, 4000, 0, capi.TOL_RELATIVE,
                         capi.Preconditioner(a, capi.PRECOND_JACOBI) if precond else None)
        st1 = s1.solve(a, x1, ctx.to_device(B[j]))
        assert abs(st1.iterations - st.iterations) <= 3, (j, st1.iterations, st.iterations)
        assert rel_l2(Xh[j], x1.download()) < 1e-10, j
        xr = X0[j].copy()
        ost = oracle.pcg(rp, col, val, xr, B[j], 1e-12, 4000, 1, pname)
        assert ost.converged and abs(ost.iterations - st.iterations) <= 3, (j, ost.iterations, st.iterations)
        assert rel_l2(Xh[j], xr) < 1e-10, j


def test_multi_3d_oracle_and_mixed(ctx, oracle):
    """8 right-hand sides on a 10^3-cube box: FP64 columns vs the oracle's PCG; mixed vs FP64."""
    from gpupfem2_amd import capi, cases
    case = cases.poisson3d_multi_case(n=10, nrhs=8)
    mesh, a, B, b0 = cases.build_poisson3d_multi(ctx, case)
    rp, col, val = a.download()
    n = rp.size - 1
    Bh = B.download().reshape(8, n)
    out, its = {}, {}
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        pc = capi.Preconditioner(a, capi.PRECOND_JACOBI)
        s = capi.Solver(a, capi.SOLVER_CG, 1e-12, 4000, 0, capi.TOL_RELATIVE, pc)
        if prec:
            s.set_precision(prec, 1e-5, 30)
        X = ctx.zeros(8 * n)
        stats = s.solve_multi(a, X, B, 8)
        assert all(st.converged for st in stats)
        if prec:
            assert all(st.refinements >= 1 for st in stats)
        out[prec] = X.download().reshape(8, n)
        its[prec] = [st.iterations for st in stats]
    assert all(m <= 1.25 * f + 5 for m, f in zip(its[1], its[0])), its
    for j in range(8):
        xr = np.zeros(n)
        ost = oracle.pcg(rp, col, val, xr, Bh[j], 1e-12, 4000, 1, "jacobi")
        assert ost.converged
        assert rel_l2(out[0][j], xr) < 1e-10, j
        assert rel_l2(out[1][j], out[0][j]) < 1e-8, j


def test_multi_solver_rejects(ctx):
    from gpupfem2_amd import capi
    m, a, rp, col, val, bn = system2d(ctx, n=8)
    n = rp.size - 1
    X, B = ctx.zeros(9 * n), ctx.zeros(9 * n)
    s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 100, 0, 0, capi.Preconditioner(a, capi.PRECOND_JACOBI))
    with pytest.raises(capi.PfemError):
        s.solve_multi(a, X, B, 9)
    with pytest.raises(capi.PfemError):
        capi.Solver(a, capi.SOLVER_CG, 1e-10, 100, 0, 0, capi.Preconditioner(a, capi.PRECOND_IC0)).solve_multi(a, X, B, 2)
    with pytest.raises(capi.PfemError):
        capi.Solver(a, capi.SOLVER_GMRES, 1e-10, 100, 10, 0, None).solve_multi(a, X, B, 2)


def test_multi_3d_box128_eight_rhs(ctx):
    """The 3D box of 128^3 cubes (12.6M tetrahedra, 2.15M nodes), 8 right-hand sides, Jacobi PCG to a
    relative tolerance of 1e-12, FP64 and mixed."""
    from gpupfem2_amd import capi, cases
    print("multi 3d box128", flush=True)
    case = cases.poisson3d_multi_case(n=128, nrhs=8)
    mesh, a, B, b0 = cases.build_poisson3d_multi(ctx, case)
    n = case["xyz"].shape[0]
    Bh = B.download().reshape(8, n)
    sols = {}
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        print("precision", prec, flush=True)
        s = capi.Solver(a, capi.SOLVER_CG, 1e-12, 20000, 0, capi.TOL_RELATIVE, capi.Preconditioner(a, capi.PRECOND_JACOBI))
        if prec:
            s.set_precision(prec, 1e-5, 30)
        X = ctx.zeros(8 * n)
        stats = s.solve_multi(a, X, B, 8)
        assert all(st.converged and not st.nan_detected for st in stats)
        Xh = X.download().reshape(8, n)
        for j in range(8):
            xd, yd = ctx.to_device(Xh[j]), ctx.empty(n)
            a.spmv(xd, yd)
            res = np.linalg.norm(Bh[j] - yd.download()) / np.linalg.norm(Bh[j])
            assert res < 1e-10, (prec, j, res)
        sols[prec] = Xh
    for j in range(8):
        assert rel_l2(sols[1][j], sols[0][j]) < 1e-8, j
