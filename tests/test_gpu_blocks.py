"""GPU parity of the blocked IC0/ILU0 ordering (PFEM_ORDER_BLOCKS): blocks of up to 8 neighbouring nodes,
coloured as blocks; in the fused PCG's single-pass sweeps a block's rows sit on neighbouring lanes of one
wave and finish one after the other.

Bar: the factor is the oracle's factor of P A P^T bit for bit (P the library's permutation), and so is
the apply in the vectors' own order (per-level and generic single-pass sweeps, ascending sums). The
blocked sweeps of the fused PCG publish tagged values (one unit in the last place), finish a row by the
reciprocal of its diagonal, and in the backward sweep subtract a block's own entries after the outside
ones, so PCG iterates built on them agree with the unfused form's (per-level order of operations) to
1e-12 relative after a fixed number of iterations, and whole solves agree in iteration count within 2
and in solution to 1e-10.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def box_poisson(ctx, n):
    """P1 Poisson system of the unit box of n^3 cubes (6 tetrahedra each), boundary nodes fixed at 0."""
    from gpupfem2_amd import capi, meshgen
    xyz, tet = meshgen.box3d(n)
    nv = xyz.shape[0]
    m = capi.Mesh3D(ctx, xyz, tet)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(nv)
    capi.assemble_poisson3d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=3.0, quad=0)
    on = np.any((xyz < 1e-9) | (xyz > 1 - 1e-9), axis=1)
    bn = np.nonzero(on)[0].astype(np.int32)
    capi.Dirichlet(ctx, nv, bn, np.zeros(bn.size)).apply(a, rhs, capi.DIRICHLET_MASK)
    return m, a, rhs, nv


def cylinder_pressure(ctx, h):
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=h)
    n = case.xy.shape[0]
    m = capi.Mesh2D(ctx, case.xy, case.tri, fill_neighbors=False)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(n)
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=1.0)
    capi.Dirichlet(ctx, n, case.bcp_node, case.bcp_val).apply(a, rhs, capi.DIRICHLET_MASK)
    return m, a, rhs, n


@pytest.mark.parametrize("kind", [2, 3])
def test_block_factor_and_apply(ctx, oracle, kind):
    """The blocked factor is the oracle's IC0/ILU0 of the permuted matrix, bit for bit; its levels are
    longer than the colours (rows of a block depend on each other); the apply on natural-order vectors,
    single-pass and per level, is the oracle's sequential apply of that factor, bit for bit."""
    import scipy.sparse as sp
    from gpupfem2_amd import capi
    print("block factor", kind, flush=True)
    m, a, rhs, n = cylinder_pressure(ctx, 1.0 / 28)
    rp, col, val = a.download()
    pc = capi.Preconditioner(a, kind, capi.ORDER_BLOCKS)
    pc.initialize(a)
    fac, nl, perm = pc.download()
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert 8 <= nl <= 8 * 8
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    B = A[perm][:, perm].tocsr()
    B.sort_indices()
    rpb, colb, valb = B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.copy()
    ref = (oracle.ic0 if kind == 2 else oracle.ilu0)(rpb, colb, valb)[1]
    if kind == 2:
        low = colb <= np.repeat(np.arange(n), np.diff(rpb))
        assert np.array_equal(fac[low], ref[low])
    else:
        assert np.array_equal(fac, ref)
    r = np.random.default_rng(11).standard_normal(n)
    zp = (oracle.ic0_apply if kind == 2 else oracle.ilu0_apply)(rpb, colb, ref, r[perm])
    z = np.empty(n)
    z[perm] = zp
    dr = ctx.to_device(r)
    for single in (True, False):
        dz = ctx.empty(n)
        pc.set_sweep(single)
        pc.apply(dz, dr)
        assert np.array_equal(dz.download(), z), single


@pytest.mark.parametrize("kind,bits", [(2, 64), (2, 16), (3, 64)])
def test_block_fused_pcg_iterates(ctx, kind, bits):
    """The fused PCG on a blocked factor (three launches per iteration: SpMV, the blocked forward sweep with
    the x/r update and (y, y), the blocked backward sweep with the x/p update; ILU0: the backward sweep
    finishes (r, z)) against the unfused PCG (generic sweeps between separate vector kernels) after 1, 2,
    5 and 40 iterations, and on whole solves; FP64 factor and the mixed step's 16-bit copy."""
    from gpupfem2_amd import capi
    print("block fused pcg", kind, bits, flush=True)
    m, a, rhs, n = cylinder_pressure(ctx, 1.0 / 28)

    def solver(fuse, maxit, tol=1e-12):
        pc = capi.Preconditioner(a, kind, capi.ORDER_BLOCKS)
        if bits == 16:
            pc.set_factor_bits(16)
        s = capi.Solver(a, capi.SOLVER_CG, tol, maxit, 0, capi.TOL_RELATIVE, pc)
        if bits == 16:
            s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 20)
        s.set_option(capi.SOLVER_OPT_FUSE, fuse)
        return s

    x0 = np.linspace(0.0, 0.1, n)
    for k in (1, 2, 5, 40):
        got = []
        for fuse in (0, 3):
            s = solver(fuse, k, 1e-30)
            x = ctx.to_device(x0)
            st = s.solve(a, x, rhs)
            assert not st.converged and st.iterations == k
            got.append(x.download())
        assert rel_l2(got[1], got[0]) < 1e-12, (k, rel_l2(got[1], got[0]))
    out = []
    for fuse in (0, 3):
        s = solver(fuse, 20000)
        runs = []
        for rep in range(2):
            x = ctx.to_device(x0)
            st = s.solve(a, x, rhs)
            assert st.converged and not st.nan_detected
            runs.append((st.iterations, x.download()))
        assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1]), fuse
        out.append(runs[0])
    assert abs(out[0][0] - out[1][0]) <= 2, (out[0][0], out[1][0])
    assert rel_l2(out[1][1], out[0][1]) < 1e-10
    y = ctx.empty(n)
    a.spmv(ctx.to_device(out[1][1]), y)
    b = rhs.download()
    assert np.linalg.norm(b - y.download()) / np.linalg.norm(b) < 1e-8


@pytest.mark.parametrize("bits", [64, 16])
def test_block_pcg_full_size(ctx, bits):
    """The cylinder's pressure system (1.01M rows): the blocked IC0-PCG converges to the tolerance in fewer
    iterations than the multicolour one, with a true residual at the tolerance; repeated solves give the
    same bits (the order in which publications arrive does not matter)."""
    from gpupfem2_amd import capi
    print("block pcg full size", bits, flush=True)
    m, a, rhs, n = cylinder_pressure(ctx, 1.0 / 112)
    its = {}
    for ordering in (capi.ORDER_MULTICOLOR, capi.ORDER_BLOCKS):
        pc = capi.Preconditioner(a, capi.PRECOND_IC0, ordering)
        if bits == 16:
            pc.set_factor_bits(16)
        s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 20000, 0, capi.TOL_RELATIVE, pc)
        if bits == 16:
            s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 20)
        ref = None
        for rep in range(2 if ordering == capi.ORDER_BLOCKS else 1):
            x = ctx.zeros(n)
            st = s.solve(a, x, rhs)
            assert st.converged and not st.nan_detected
            got = (st.iterations, x.download())
            if ref is None:
                ref = got
                y = ctx.empty(n)
                a.spmv(x, y)
                b = rhs.download()
                assert np.linalg.norm(b - y.download()) / np.linalg.norm(b) < 1e-7
            else:
                assert got[0] == ref[0] and np.array_equal(got[1], ref[1]), rep
        its[ordering] = ref[0]
    print("iterations multicolour", its[capi.ORDER_MULTICOLOR], "blocks", its[capi.ORDER_BLOCKS], flush=True)
    assert its[capi.ORDER_BLOCKS] < 0.9 * its[capi.ORDER_MULTICOLOR]


def test_block_order_3d(ctx, oracle):
    """A tetrahedral box (rows of up to 14 lower entries; a block graph that is not planar, so more than 4
    block colours): the blocked IC0 factor is the oracle's factor of P A P^T bit for bit, and the fused
    blocked PCG's iterates agree with the unfused PCG's to 1e-12 after 30 iterations."""
    import scipy.sparse as sp
    from gpupfem2_amd import capi
    print("block order 3d", flush=True)
    m, a, rhs, n = box_poisson(ctx, 12)
    rp, col, val = a.download()
    pc = capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_BLOCKS)
    pc.initialize(a)
    fac, nl, perm = pc.download()
    assert np.array_equal(np.sort(perm), np.arange(n))
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    B = A[perm][:, perm].tocsr()
    B.sort_indices()
    rpb, colb = B.indptr.astype(np.int32), B.indices.astype(np.int32)
    ref = oracle.ic0(rpb, colb, B.data.copy())[1]
    low = colb <= np.repeat(np.arange(n), np.diff(rpb))
    assert np.array_equal(fac[low], ref[low])
    x0 = np.linspace(0.0, 0.1, n)
    got = []
    for fuse in (0, 3):
        s = capi.Solver(a, capi.SOLVER_CG, 1e-30, 30, 0, capi.TOL_RELATIVE,
                        capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_BLOCKS))
        s.set_option(capi.SOLVER_OPT_FUSE, fuse)
        x = ctx.to_device(x0)
        st = s.solve(a, x, rhs)
        assert st.iterations == 30
        got.append(x.download())
    assert rel_l2(got[1], got[0]) < 1e-12


def path_matrix(n, seed=0):
    """SPD tridiagonal matrix of a path of n nodes (2 on the diagonal plus a random shift, -1 off it)."""
    rng = np.random.default_rng(seed)
    rp, col, val = [0], [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                col.append(j)
                val.append(2.0 + rng.uniform(0.0, 0.5) if j == i else -1.0)
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val)


@pytest.mark.parametrize("n", [1, 3, 8, 13, 300])
@pytest.mark.parametrize("kind", [2, 3])
def test_block_order_small(ctx, oracle, n, kind):
    """Matrices smaller than a block, of one block, and of a few blocks of mixed sizes (a path graph: blocks
    of 8, 4, 2 and 1 nodes, 2 block colours): the factor is the oracle's of P A P^T bit for bit, the apply
    its sequential apply bit for bit, and the fused blocked PCG solves the system (residual 1e-10) in the
    iteration count of the unfused PCG within 1."""
    import scipy.sparse as sp
    from gpupfem2_amd import capi
    rp, col, val = path_matrix(n, seed=n)
    a = capi.CSR(ctx, rowptr=rp, col=col, val=val)
    pc = capi.Preconditioner(a, kind, capi.ORDER_BLOCKS)
    pc.initialize(a)
    fac, nl, perm = pc.download()
    assert np.array_equal(np.sort(perm), np.arange(n))
    B = sp.csr_matrix((val, col, rp), shape=(n, n))[perm][:, perm].tocsr()
    B.sort_indices()
    rpb, colb = B.indptr.astype(np.int32), B.indices.astype(np.int32)
    ref = (oracle.ic0 if kind == 2 else oracle.ilu0)(rpb, colb, B.data.copy())[1]
    if kind == 2:
        low = colb <= np.repeat(np.arange(n), np.diff(rpb))
        assert np.array_equal(fac[low], ref[low])
    else:
        assert np.array_equal(fac, ref)
    r = np.random.default_rng(1).standard_normal(n)
    zp = (oracle.ic0_apply if kind == 2 else oracle.ilu0_apply)(rpb, colb, ref, r[perm])
    z = np.empty(n)
    z[perm] = zp
    dz = ctx.empty(n)
    pc.apply(dz, ctx.to_device(r))
    assert np.array_equal(dz.download(), z)
    b = np.random.default_rng(2).standard_normal(n)
    its = []
    for fuse in (0, 3):
        s = capi.Solver(a, capi.SOLVER_CG, 1e-12, 1000, 0, capi.TOL_RELATIVE,
                        capi.Preconditioner(a, kind, capi.ORDER_BLOCKS))
        s.set_option(capi.SOLVER_OPT_FUSE, fuse)
        x = ctx.zeros(n)
        st = s.solve(a, x, ctx.to_device(b))
        assert st.converged
        xs = x.download()
        assert np.linalg.norm(b - sp.csr_matrix((val, col, rp), shape=(n, n)) @ xs) <= 1e-10 * np.linalg.norm(b)
        its.append(st.iterations)
    assert abs(its[0] - its[1]) <= 1, its


def test_block_gmres_and_fp32_refinement(ctx):
    """The other solver paths over a blocked factor (the generic single-pass sweeps, in the factor's
    ordering): GMRES(30) with ILU0 on a nonsymmetric system (the pressure matrix with its diagonal
    shifted by 1 and its off-diagonal entries perturbed by up to 10%), and the FP32 inner PCG with FP64
    refinement with IC0 (PFEM_PRECISION_MIXED) on the symmetric matrix with the shifted diagonal, each
    against the same solve over the multicolour factor and a direct solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from gpupfem2_amd import capi
    print("block gmres and fp32 refinement", flush=True)
    m, a, rhs, n = cylinder_pressure(ctx, 1.0 / 28)
    rp, col, val = a.download()
    rows = np.repeat(np.arange(n), np.diff(rp))
    rng = np.random.default_rng(9)
    v = val + np.where(col == rows, 1.0, 0.1 * rng.uniform(-1, 1, col.size) * np.abs(val))
    an = capi.CSR(ctx, like=a)
    an.upload_values(v)
    b = rng.standard_normal(n)
    exact = spla.spsolve(sp.csr_matrix((v, col, rp), shape=(n, n)).tocsc(), b)
    sols = []
    for ordering in (capi.ORDER_MULTICOLOR, capi.ORDER_BLOCKS):
        s = capi.Solver(an, capi.SOLVER_GMRES, 1e-12, 20000, 30, capi.TOL_RELATIVE,
                        capi.Preconditioner(an, capi.PRECOND_ILU0, ordering))
        x = ctx.zeros(n)
        st = s.solve(an, x, ctx.to_device(b))
        assert st.converged, ordering
        sols.append(x.download())
        assert rel_l2(sols[-1], exact) < 1e-8, ordering
    # FP32 inner iterations need a condition number FP32 can carry: the symmetric matrix with the shifted
    # diagonal (the pressure matrix itself stalls them, with either ordering)
    vs = val + np.where(col == rows, 1.0, 0.0)
    asym = capi.CSR(ctx, like=a)
    asym.upload_values(vs)
    exact_s = spla.spsolve(sp.csr_matrix((vs, col, rp), shape=(n, n)).tocsc(), b)
    for ordering in (capi.ORDER_MULTICOLOR, capi.ORDER_BLOCKS):
        s = capi.Solver(asym, capi.SOLVER_CG, 1e-12, 20000, 0, capi.TOL_RELATIVE,
                        capi.Preconditioner(asym, capi.PRECOND_IC0, ordering))
        s.set_precision(capi.PRECISION_MIXED, 1e-5, 30)
        x = ctx.zeros(n)
        st = s.solve(asym, x, ctx.to_device(b))
        assert st.converged and st.refinements >= 1, ordering
        assert rel_l2(x.download(), exact_s) < 1e-9, ordering
