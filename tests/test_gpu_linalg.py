"""GPU parity: mesh import, setup builders, SpMV, assembly, Dirichlet, preconditioners and solvers
against the CPU oracle on the same inputs.

Bar: bit-exact for integer data (cells, CSR structure, neighbour lists) and for every kernel whose
floating-point operation order matches the oracle (geometry, SpMV, assembly, Dirichlet, Jacobi, IC0/ILU0
factors); relative L2 <= 1e-10 for Krylov solutions (their dot products are reduced in a different
order, which moves the last bits and can change the iteration count by a few).
"""
import math

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu

FIX2D = ["TestProblem2.dat", "ChannelMesh.dat", "CylinderMesh.dat", "CylinderMesh3.dat"]


@pytest.mark.parametrize("fname", FIX2D)
def test_mesh2d_import_and_setup(ctx, oracle, fname):
    from gpupfem2_amd import capi
    print("mesh2d import", fname, flush=True)
    v, c, ndev = oracle.load_dat(golden(fname), 2)
    m = capi.Mesh2D(ctx, path=golden(fname))
    assert (m.nv, m.nt, m.nt_host) == (v.shape[0], ndev, c.shape[0])
    assert np.array_equal(m.download(0).reshape(-1, 2), v)
    assert np.array_equal(m.download(1).reshape(-1, 3), c[:ndev])
    area, invJ = oracle.geometry2d(v, c)
    assert np.array_equal(m.download(2), area)
    assert np.array_equal(m.download(3).reshape(-1, 4), invJ)
    if c.shape[0] < 20000:
        off, idx = oracle.neighbors_bruteforce(c)
        assert np.array_equal(m.download(4), off) and np.array_equal(m.download(5), idx)
    rp, col = oracle.csr_pattern(v.shape[0], c)
    a = capi.CSR(ctx, mesh=m)
    grp, gcol, gval = a.download()
    assert np.array_equal(grp, rp) and np.array_equal(gcol, col) and np.all(gval == 0)


def test_mesh2d_trailing_newline(ctx, oracle, tmp_path):
    from gpupfem2_amd import capi
    p = tmp_path / "nl.dat"
    p.write_text(open(golden("TestProblem2.dat")).read() + "\n")
    m = capi.Mesh2D(ctx, path=str(p))
    assert m.nt == 1711 and m.nt_host == 1712
    with pytest.raises(capi.PfemError):
        capi.Mesh2D(ctx, path=str(tmp_path / "missing.dat"))


def test_mesh3d_import(ctx, oracle):
    from gpupfem2_amd import capi
    v, c, ndev = oracle.load_dat(golden("BoxMesh.dat"), 3)
    m = capi.Mesh3D(ctx, path=golden("BoxMesh.dat"))
    assert (m.nv, m.nt) == (v.shape[0], ndev)
    assert np.array_equal(m.download(0).reshape(-1, 3), v)
    assert np.array_equal(m.download(1).reshape(-1, 4), c)
    vol, invJ = oracle.geometry3d(v, c)
    assert np.array_equal(m.download(2), vol) and np.array_equal(m.download(3).reshape(-1, 9), invJ)
    rp, col = oracle.csr_pattern(v.shape[0], c)
    grp, gcol, _ = capi.CSR(ctx, mesh=m).download()
    assert np.array_equal(grp, rp) and np.array_equal(gcol, col)


def test_mesh3d_face_neighbors(ctx, oracle):
    """Mesh3D::fillCellNeighborIndices (src/mesh_3d.cu:37-82, 198-227) built from the incidence lists:
    BoxMesh.dat, and its cells followed by cells with a repeated vertex, equal the oracle's all-pairs
    test; the box of 64^3 cubes (1.57M tetrahedra) equals the face hashing; the 128^3 box (12.6M) has
    4T - 12 n^2 entries (12 n^2 boundary triangles), at most 4 per cell, ascending, and sampled entries
    share exactly one face and are listed both ways. A second call returns the same lists."""
    from gpupfem2_amd import capi, meshgen
    from helpers import face_neighbors_by_hashing
    v, c, ndev = oracle.load_dat(golden("BoxMesh.dat"), 3)
    m = capi.Mesh3D(ctx, path=golden("BoxMesh.dat"))
    off, idx = m.fill_neighbors()
    roff, ridx = oracle.face_neighbors3d(c[:ndev])
    assert np.array_equal(off, roff) and np.array_equal(idx, ridx)
    off2, idx2 = m.fill_neighbors()
    assert np.array_equal(off2, off) and np.array_equal(idx2, idx)
    tet = c[:ndev]
    degenerate = np.array([[t[0], t[0], t[1], t[2]] for t in tet[:40:3]] +
                          [[t[1], t[1], t[2], t[2]] for t in tet[5:60:7]], np.int32)
    tet2 = np.concatenate([tet, degenerate])
    m2 = capi.Mesh3D(ctx, v, tet2, zero_small_det=True)
    off, idx = m2.fill_neighbors()
    roff, ridx = oracle.face_neighbors3d(tet2)
    assert np.array_equal(off, roff) and np.array_equal(idx, ridx)
    for n in (64, 128):
        print("face neighbours box", n, flush=True)
        xyz, tet = meshgen.box3d(n)
        m = capi.Mesh3D(ctx, xyz, tet)
        off, idx = m.fill_neighbors()
        T = tet.shape[0]
        assert off[-1] == idx.shape[0] == 4 * T - 12 * n * n
        assert np.diff(off).max() == 4
        if n == 64:
            hoff, hidx, _ = face_neighbors_by_hashing(tet)
            assert np.array_equal(off, hoff) and np.array_equal(idx, hidx)
            continue
        seg = np.repeat(np.arange(T), np.diff(off))
        assert np.all((idx[1:] > idx[:-1]) | (seg[1:] != seg[:-1]))
        rng = np.random.default_rng(7)
        for cc in rng.integers(0, T, 2000):
            for d in idx[off[cc]:off[cc + 1]]:
                assert np.sum(tet[cc][:, None] == tet[d][None, :]) == 3
                assert cc in idx[off[d]:off[d + 1]]


def test_generated_cylinder_setup(ctx, oracle):
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=1.0 / 8)
    m = capi.Mesh2D(ctx, case.xy, case.tri)
    off, idx = oracle.neighbors_bruteforce(case.tri)
    assert np.array_equal(m.download(4), off) and np.array_equal(m.download(5), idx)
    m.classify_edges(**{k: v for k, v in case.classify.items() if k != "circle_id"})
    assert np.array_equal(m.download(6).reshape(-1, 3), case.edge_ids)


def laplace(ctx, oracle, n=32):
    from gpupfem2_amd import capi, meshgen
    xy, tri = meshgen.unit_square(n)
    m = capi.Mesh2D(ctx, xy, tri)
    a = capi.CSR(ctx, mesh=m)
    area, invJ = oracle.geometry2d(xy, tri)
    rp, col = oracle.csr_pattern(xy.shape[0], tri)
    return xy, tri, m, a, area, invJ, rp, col


@pytest.mark.parametrize("rhs_kind", [2, 1, 0])
def test_poisson2d_assembly(ctx, oracle, rhs_kind):
    from gpupfem2_amd import capi
    xy, tri, m, a, area, invJ, rp, col = laplace(ctx, oracle)
    rhs = ctx.zeros(xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, coeff=0.5, rhs_kind=rhs_kind, rhs_param=2 * math.pi ** 2)
    val, r = oracle.assemble_poisson2d(xy, tri, area, invJ, rp, col, coeff=0.5, rhs_kind=rhs_kind,
                                       rhs_param=2 * math.pi ** 2)
    assert np.array_equal(a.download()[2], val)
    if rhs_kind == 2:
        assert np.array_equal(rhs.download(), r)
    else:   # sin/exp come from different math libraries: a few ulp
        assert rel_l2(rhs.download(), r) < 1e-15


@pytest.mark.parametrize("n", [6, 128])
def test_poisson3d_assembly(ctx, oracle, n):
    """Geometry, CSR pattern and the P1 Poisson system of the box of n^3 cubes (6 tetrahedra each; n = 128
    is the 12.6M-tetrahedron, 2.15M-node configuration) bit for bit against the oracle."""
    from gpupfem2_amd import capi, meshgen
    print("poisson3d assembly", n, flush=True)
    xyz, tet = meshgen.box3d(n)
    m = capi.Mesh3D(ctx, xyz, tet)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(xyz.shape[0])
    capi.assemble_poisson3d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=3.0, quad=0)
    vol, invJ = oracle.geometry3d(xyz, tet)
    rp, col = oracle.csr_pattern(xyz.shape[0], tet)
    val, r = oracle.assemble_poisson3d(xyz, tet, vol, invJ, rp, col, 1.0, 2, 3.0, 0)
    assert np.array_equal(m.download(2), vol) and np.array_equal(m.download(3).reshape(-1, 9), invJ)
    grp, gcol, gval = a.download()
    assert np.array_equal(grp, rp) and np.array_equal(gcol, col)
    assert np.array_equal(gval, val) and np.array_equal(rhs.download(), r)


def test_spmv_bitexact(ctx, oracle):
    from gpupfem2_amd import capi, cases
    rng = np.random.default_rng(0)
    for h in (1.0 / 16, 1.0 / 112):     # the second is the 2M-triangle cylinder pattern (n ~ 1.0M)
        case = cases.cylinder_case(h=h)
        print("spmv", case.name, flush=True)
        m = capi.Mesh2D(ctx, case.xy, case.tri, fill_neighbors=False)
        a = capi.CSR(ctx, mesh=m)
        rp, col, _ = a.download()
        val = rng.standard_normal(col.shape[0])
        a.upload_values(val)
        x = rng.standard_normal(rp.shape[0] - 1)
        dx, dy = ctx.to_device(x), ctx.empty(x.size)
        a.spmv(dx, dy)
        assert np.array_equal(dy.download(), oracle.spmv(rp, col, val, x))


@pytest.mark.parametrize("mode", [0, 1])
def test_dirichlet(ctx, oracle, mode):
    from gpupfem2_amd import capi
    xy, tri, m, a, area, invJ, rp, col = laplace(ctx, oracle, n=24)
    rhs = ctx.zeros(xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, rhs_kind=2, rhs_param=1.0)
    val, r = oracle.assemble_poisson2d(xy, tri, area, invJ, rp, col, rhs_kind=2, rhs_param=1.0)
    bn = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 1] < 1e-9) | (xy[:, 0] > 1 - 1e-9))[0].astype(np.int32)
    g = np.sin(3 * xy[bn, 0]) + xy[bn, 1]
    bc = capi.Dirichlet(ctx, xy.shape[0], bn, g)
    bc.apply(a, rhs, mode)
    oracle.dirichlet_apply(rp, col, val, r, bn, g, mode)
    assert np.array_equal(a.download()[2], val) and np.array_equal(rhs.download(), r)


def test_dirichlet_two_rhs(ctx, oracle):
    from gpupfem2_amd import capi
    xy, tri, m, a, area, invJ, rp, col = laplace(ctx, oracle, n=16)
    r0, r1 = ctx.zeros(xy.shape[0]), ctx.zeros(xy.shape[0])
    capi.assemble_poisson2d(m, a, r0, rhs_kind=2, rhs_param=1.0)
    r1.copy_from(r0)
    val, r = oracle.assemble_poisson2d(xy, tri, area, invJ, rp, col, rhs_kind=2, rhs_param=1.0)
    bn = np.nonzero(xy[:, 1] < 1e-9)[0].astype(np.int32)
    g0, g1 = xy[bn, 0] + 1, -xy[bn, 0]
    b0, b1 = capi.Dirichlet(ctx, xy.shape[0], bn, g0), capi.Dirichlet(ctx, xy.shape[0], bn, g1)
    b0.apply2(b1, a, r0, r1, 1)
    v0, q0 = val.copy(), r.copy()
    oracle.dirichlet_apply(rp, col, v0, q0, bn, g0, 1)
    v1, q1 = val.copy(), r.copy()
    oracle.dirichlet_apply(rp, col, v1, q1, bn, g1, 1)
    assert np.array_equal(a.download()[2], v0) and np.array_equal(v0, v1)
    assert np.array_equal(r0.download(), q0) and np.array_equal(r1.download(), q1)


def bc_system(ctx, oracle, n=40):
    from gpupfem2_amd import capi
    xy, tri, m, a, area, invJ, rp, col = laplace(ctx, oracle, n=n)
    rhs = ctx.zeros(xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, rhs_kind=1, rhs_param=2 * math.pi ** 2)
    bn = np.nonzero((xy[:, 0] < 1e-9) | (xy[:, 0] > 1 - 1e-9) | (xy[:, 1] < 1e-9) | (xy[:, 1] > 1 - 1e-9))[0]
    capi.Dirichlet(ctx, xy.shape[0], bn, np.zeros(bn.size)).apply(a, rhs, 1)
    rp_, col_, val = a.download()
    return xy, m, a, rhs, rp, col, val


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_precond_factors(ctx, oracle, kind):
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle, n=24)
    pc = capi.Preconditioner(a, kind, capi.ORDER_NATURAL)
    pc.initialize(a)
    fac, nl, perm = pc.download()
    if kind == 1:
        assert np.array_equal(fac, oracle.jacobi(rp, col, val))
        return
    ref = (oracle.ic0 if kind == 2 else oracle.ilu0)(rp, col, val)[1]
    if kind == 2:
        low = col <= np.repeat(np.arange(rp.size - 1), np.diff(rp))
        assert np.array_equal(fac[low], ref[low])
    else:
        assert np.array_equal(fac, ref)
    r = np.random.default_rng(2).standard_normal(rp.size - 1)
    dr, dz = ctx.to_device(r), ctx.empty(r.size)
    pc.apply(dz, dr)
    z_ref = (oracle.ic0_apply if kind == 2 else oracle.ilu0_apply)(rp, col, ref, r)
    assert np.array_equal(dz.download(), z_ref)


@pytest.mark.parametrize("kind", [2, 3])
def test_precond_multicolor(ctx, oracle, kind):
    """The multicolour factor is the factor of P A P^T: same bits as the oracle on the permuted matrix."""
    from gpupfem2_amd import capi
    import scipy.sparse as sp
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle, n=24)
    pc = capi.Preconditioner(a, kind, capi.ORDER_MULTICOLOR)
    pc.initialize(a)
    fac, nl, perm = pc.download()
    assert nl <= 12
    n = rp.size - 1
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
    r = np.random.default_rng(3).standard_normal(n)
    dr, dz = ctx.to_device(r), ctx.empty(n)
    pc.apply(dz, dr)
    zp = (oracle.ic0_apply if kind == 2 else oracle.ilu0_apply)(rpb, colb, ref, r[perm])
    z = np.empty(n)
    z[perm] = zp
    assert np.array_equal(dz.download(), z)


@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("ordering", [0, 1])
def test_precond_sweep_forms(ctx, oracle, kind, ordering):
    """Single-pass sweeps (one launch per triangular solve) and per-level launches give the same bits, on a
    graded cylinder mesh whose natural ordering has hundreds of levels (long dependency chains); the
    natural-order result is also the oracle's sequential solve, bit for bit."""
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=1.0 / 28)
    xy, tri = case.xy, case.tri
    m = capi.Mesh2D(ctx, xy, tri)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=1.0)
    capi.Dirichlet(ctx, xy.shape[0], case.bcp_node, case.bcp_val).apply(a, rhs, 1)
    rp, col, val = a.download()
    n = rp.size - 1
    pc = capi.Preconditioner(a, kind, ordering)
    pc.initialize(a)
    fac, nl, perm = pc.download()
    assert nl > (100 if ordering == 0 else 1)
    rng = np.random.default_rng(5)
    for trial in range(3):
        r = rng.standard_normal(n)
        dr, z1, z0 = ctx.to_device(r), ctx.empty(n), ctx.empty(n)
        pc.set_sweep(True)
        pc.apply(z1, dr)
        pc.set_sweep(False)
        pc.apply(z0, dr)
        assert np.array_equal(z1.download(), z0.download())
        if ordering == 0 and trial == 0:
            z_ref = (oracle.ic0_apply if kind == 2 else oracle.ilu0_apply)(rp, col, fac, r)
            assert np.array_equal(z1.download(), z_ref)
    pc.set_sweep(True)
    dr = ctx.to_device(r)
    pc.apply(dr, dr)   # in place, as GMRES applies it
    assert np.array_equal(dr.download(), z0.download())


@pytest.mark.parametrize("kind", [2, 3])
def test_precond_sweep_forms_full_size(ctx, kind):
    """The sliced sweep layouts at the 2M-triangle cylinder's size (1.01M rows, 4 colours), where some
    chunks' columns of one colour span more than a 16-bit code holds (32-bit fallback) and others are
    coded: single-pass sweeps against per-level launches on the combined factor, bit for bit."""
    from gpupfem2_amd import capi, cases
    print("sweep forms full size", kind, flush=True)
    case = cases.cylinder_case()
    m = capi.Mesh2D(ctx, case.xy, case.tri, fill_neighbors=False)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(case.xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=1.0)
    capi.Dirichlet(ctx, case.xy.shape[0], case.bcp_node, case.bcp_val).apply(a, rhs, 1)
    n = case.xy.shape[0]
    pc = capi.Preconditioner(a, kind, capi.ORDER_MULTICOLOR)
    pc.initialize(a)
    r = np.random.default_rng(7).standard_normal(n)
    dr, z1, z0 = ctx.to_device(r), ctx.empty(n), ctx.empty(n)
    pc.set_sweep(True)
    pc.apply(z1, dr)
    pc.set_sweep(False)
    pc.apply(z0, dr)
    assert np.array_equal(z1.download(), z0.download())


@pytest.mark.parametrize("precond,ordering", [(0, 0), (1, 0), (2, 0), (3, 0), (2, 1)])
@pytest.mark.parametrize("tol_mode", [0, 1])
def test_pcg(ctx, oracle, precond, ordering, tol_mode):
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle)
    pc = capi.Preconditioner(a, precond, ordering) if precond else None
    tol = 1e-9 if tol_mode == 0 else 1e-11
    s = capi.Solver(a, capi.SOLVER_CG, tol, 3000, 0, tol_mode, pc)
    x = ctx.zeros(xy.shape[0])
    st = s.solve(a, x, rhs)
    assert st.converged
    b = rhs.download()
    xr = np.zeros(b.size)
    if ordering == 0:
        ost = oracle.pcg(rp, col, val, xr, b, tol, 3000, tol_mode, ["none", "jacobi", "ic0", "ilu0"][precond])
        assert ost.converged and abs(ost.iterations - st.iterations) <= 3
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    exact = spla.spsolve(sp.csr_matrix((val, col, rp)).tocsc(), b)
    assert rel_l2(x.download(), exact) < 1e-7


@pytest.mark.parametrize("restart,passes", [(30, 2), (3000, 2), (30, 1)])
def test_gmres(ctx, oracle, restart, passes):
    """GMRES (classical Gram-Schmidt, passes = 2: one re-orthogonalisation, or a single pass) against the
    oracle's modified Gram-Schmidt GMRES on a nonsymmetric system."""
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle, n=24)
    n = rp.size - 1
    rng = np.random.default_rng(4)
    rows = np.repeat(np.arange(n), np.diff(rp))
    v = val + np.where(col == rows, 1.0, 0.3 * rng.uniform(-1, 1, col.size) * np.abs(val))
    a.upload_values(v)
    b = rng.standard_normal(n)
    db = ctx.to_device(b)
    pc = capi.Preconditioner(a, capi.PRECOND_JACOBI)
    s = capi.Solver(a, capi.SOLVER_GMRES, 1e-12, 3000, restart, capi.TOL_RELATIVE, pc)
    s.set_option(capi.SOLVER_OPT_GS_PASSES, passes)
    x = ctx.zeros(n)
    st = s.solve(a, x, db)
    assert st.converged
    xr = np.zeros(n)
    ost = oracle.gmres(rp, col, v, xr, b, 1e-12, 3000, restart, 1, "jacobi")
    assert ost.converged and abs(ost.iterations - st.iterations) <= 3
    assert rel_l2(x.download(), xr) < 1e-10


@pytest.mark.parametrize("prec", [0, 1])
def test_gmres_pair(ctx, oracle, prec):
    """Two nonsymmetric systems on one pattern solved together (pfem_solver_solve_pair: Arnoldi steps in
    lockstep, one SpMV launch for both) give the bits and iteration counts of two single solves, in FP64
    and in mixed precision (FP32 inner GMRES with FP64 refinement); the FP64 pair agrees with the oracle's
    modified Gram-Schmidt GMRES. A second call reuses the solvers (restarted state)."""
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle, n=24)
    n = rp.size - 1
    rng = np.random.default_rng(8)
    rows = np.repeat(np.arange(n), np.diff(rp))
    vals = [val + np.where(col == rows, 1.0, w * rng.uniform(-1, 1, col.size) * np.abs(val)) for w in (0.3, 0.1)]
    mats = [capi.CSR(ctx, like=a) for _ in vals]
    for mt, v in zip(mats, vals):
        mt.upload_values(v)
    bs = [rng.standard_normal(n), rng.standard_normal(n)]

    def solver(mt):
        s = capi.Solver(mt, capi.SOLVER_GMRES, 1e-12, 3000, 30, capi.TOL_RELATIVE, capi.Preconditioner(mt, capi.PRECOND_JACOBI))
        if prec:
            s.set_precision(capi.PRECISION_MIXED, 1e-5, 20)
        return s

    single = []
    for k in range(2):
        s = solver(mats[k])
        x = ctx.zeros(n)
        st = s.solve(mats[k], x, ctx.to_device(bs[k]))
        assert st.converged
        single.append((st.iterations, x.download()))
    s0, s1 = solver(mats[0]), solver(mats[1])
    for rep in range(2):
        x0, x1 = ctx.zeros(n), ctx.zeros(n)
        st = s0.solve_pair(mats[0], x0, ctx.to_device(bs[0]), s1, mats[1], x1, ctx.to_device(bs[1]))
        for k, (x, t) in enumerate(((x0, st[0]), (x1, st[1]))):
            assert t.converged and t.iterations == single[k][0], (rep, k, t.iterations, single[k][0])
            assert np.array_equal(x.download(), single[k][1]), (rep, k)
    if prec == 0:
        for k in range(2):
            xr = np.zeros(n)
            ost = oracle.gmres(rp, col, vals[k], xr, bs[k], 1e-12, 3000, 30, 1, "jacobi")
            assert ost.converged and abs(ost.iterations - single[k][0]) <= 3
            assert rel_l2(single[k][1], xr) < 1e-10


@pytest.mark.parametrize("kind,precond", [(0, 1), (0, 2), (1, 1)])
def test_mixed_precision_refinement(ctx, oracle, kind, precond):
    """FP32 inner Krylov + FP64 refinement reaches the FP64 tolerance."""
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle)
    pc = capi.Preconditioner(a, precond, capi.ORDER_MULTICOLOR if precond == 2 else 0)
    s = capi.Solver(a, kind, 1e-12, 5000, 30, capi.TOL_RELATIVE, pc)
    s.set_precision(capi.PRECISION_MIXED, 1e-5, 20)
    x = ctx.zeros(xy.shape[0])
    st = s.solve(a, x, rhs)
    assert st.converged and st.refinements >= 1
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    exact = spla.spsolve(sp.csr_matrix((val, col, rp)).tocsc(), rhs.download())
    assert rel_l2(x.download(), exact) < 1e-9


@pytest.mark.parametrize("kind,precond,ordering", [(0, 2, 1), (0, 2, 0), (1, 3, 1)])
def test_mixed_factor_precision(ctx, oracle, kind, precond, ordering):
    """FP64 Krylov with the IC0/ILU0 factor in FP32 reaches the FP64 tolerance in about the FP64 solver's
    iteration count."""
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle)
    its = []
    xs = []
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED_PRECOND):
        pc = capi.Preconditioner(a, precond, ordering)
        s = capi.Solver(a, kind, 1e-12, 5000, 30, capi.TOL_RELATIVE, pc)
        if prec != capi.PRECISION_FP64:
            s.set_precision(prec, 1e-5, 20)
        x = ctx.zeros(xy.shape[0])
        st = s.solve(a, x, rhs)
        assert st.converged and st.refinements == 0
        its.append(st.iterations)
        xs.append(x.download())
    assert abs(its[1] - its[0]) <= max(3, its[0] // 10)
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    exact = spla.spsolve(sp.csr_matrix((val, col, rp)).tocsc(), rhs.download())
    assert rel_l2(xs[1], exact) < 1e-9 and rel_l2(xs[0], exact) < 1e-9


@pytest.mark.parametrize("kind,precond,prec", [(0, 2, 0), (0, 2, 2), (0, 2, 1), (1, 3, 0)])
def test_factor_order_solve(ctx, oracle, kind, precond, prec):
    """A Krylov solve with a multicolour IC0/ILU0 run in the factor's ordering (P A P^T, P b, P x) gives the
    natural-order solve's solution (reduction order aside) and about its iteration count."""
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle)
    out = []
    for fo in (0, 1):
        pc = capi.Preconditioner(a, precond, capi.ORDER_MULTICOLOR)
        s = capi.Solver(a, kind, 1e-12, 5000, 30, capi.TOL_RELATIVE, pc)
        if prec:
            s.set_precision(prec, 1e-5, 20)
        s.set_option(capi.SOLVER_OPT_FACTOR_ORDER, fo)
        if kind == 0:
            s.set_option(capi.SOLVER_OPT_FUSE, 2)
        x = ctx.zeros(xy.shape[0])
        x.upload(np.linspace(0.0, 0.1, xy.shape[0]))      # a warm start exercises the permutation of x
        st = s.solve(a, x, rhs)
        assert st.converged
        out.append((st.iterations, x.download()))
    assert abs(out[0][0] - out[1][0]) <= max(3, out[0][0] // 20)
    assert rel_l2(out[1][1], out[0][1]) < 1e-10
    if kind == 0:
        # PCG fused (level 2: alpha and beta finished by the last workgroups of the SpMV and of the
        # backward sweep, the x/r update inside the forward sweep; level 1: only the x/r update) and
        # unfused: the same operations, so the same bits; two solves in a row check that the partial
        # slots are re-armed
        for fuse in (0, 1, 2):
            pc = capi.Preconditioner(a, precond, capi.ORDER_MULTICOLOR)
            s = capi.Solver(a, kind, 1e-12, 5000, 30, capi.TOL_RELATIVE, pc)
            if prec:
                s.set_precision(prec, 1e-5, 20)
            s.set_option(capi.SOLVER_OPT_FUSE, fuse)
            for rep in range(2):
                x = ctx.zeros(xy.shape[0])
                x.upload(np.linspace(0.0, 0.1, xy.shape[0]))
                st = s.solve(a, x, rhs)
                assert st.iterations == out[1][0] and np.array_equal(x.download(), out[1][1]), (fuse, rep)
        # level 3 (the default): y and z tagged in one buffer (values within one unit in the last place of
        # level 2's): the same iteration count within 1, the same solution to 1e-12, repeated solves
        # bit-identical, and the Dirichlet rows' zeros kept exactly
        pc = capi.Preconditioner(a, precond, capi.ORDER_MULTICOLOR)
        s = capi.Solver(a, kind, 1e-12, 5000, 30, capi.TOL_RELATIVE, pc)
        if prec:
            s.set_precision(prec, 1e-5, 20)
        s.set_option(capi.SOLVER_OPT_FUSE, 3)
        first = None
        for rep in range(2):
            x = ctx.zeros(xy.shape[0])
            x.upload(np.linspace(0.0, 0.1, xy.shape[0]))
            st = s.solve(a, x, rhs)
            assert st.converged and abs(st.iterations - out[1][0]) <= 1, (st.iterations, out[1][0])
            assert rel_l2(x.download(), out[1][1]) < 1e-12
            assert np.array_equal(x.download()[out[1][1] == 0.0], out[1][1][out[1][1] == 0.0])
            if first is None:
                first = x.download()
            else:
                assert np.array_equal(x.download(), first)
        # fused batches replayed from a HIP graph or launched kernel by kernel (the default), with an odd
        # poll interval too (the iteration parity is kept on the device): the same bits, and eager work
        # between the solves does not disturb the replays
        for fuse in (2, 3):
            ref = None
            for graph, poll in ((0, 8), (1, 8), (1, 3), (1, 8)):
                pc = capi.Preconditioner(a, precond, capi.ORDER_MULTICOLOR)
                s = capi.Solver(a, kind, 1e-12, 5000, 30, capi.TOL_RELATIVE, pc)
                if prec:
                    s.set_precision(prec, 1e-5, 20)
                s.set_option(capi.SOLVER_OPT_FUSE, fuse)
                s.set_option(capi.SOLVER_OPT_GRAPH, graph)
                s.set_option(capi.SOLVER_OPT_POLL, poll)
                x = ctx.zeros(xy.shape[0])
                x.upload(np.linspace(0.0, 0.1, xy.shape[0]))
                st = s.solve(a, x, rhs)
                y = ctx.empty(xy.shape[0])
                a.spmv(x, y)
                got = (st.iterations, x.download(), y.download())
                if ref is None:
                    ref = got
                else:
                    assert got[0] == ref[0] and np.array_equal(got[1], ref[1]), (fuse, graph, poll)
                    assert np.array_equal(got[2], ref[2])


def test_sweep_wait_failure_is_reported(ctx, oracle):
    """A single-pass wait that exceeds its spin limit, forced from the host with a limit of 0 polls, ends
    its launch normally (no hang), and the solve reports PFEM_ENUMERIC; with the default limit restored,
    the same solver and preconditioner solve again and reproduce the undisturbed result bit for bit."""
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle, n=160)
    n = xy.shape[0]

    def solver():
        pc = capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_MULTICOLOR)
        return pc, capi.Solver(a, capi.SOLVER_CG, 1e-10, 4000, 0, capi.TOL_RELATIVE, pc)

    pc0, s0 = solver()
    x0 = ctx.zeros(n)
    st0 = s0.solve(a, x0, rhs)
    assert st0.converged
    pc, s = solver()
    pc.set_spin_limit(0)
    x = ctx.zeros(n)
    try:
        st = s.solve(a, x, rhs)
        # no wait had to poll: then nothing failed and the result is the undisturbed one
        assert st.converged and np.array_equal(x.download(), x0.download())
        print("spin limit 0: no wait had to poll", flush=True)
    except capi.PfemError as e:
        assert e.code == -4 and "spin limit" in str(e)
        print("spin limit 0: failure reported", flush=True)
    pc.set_spin_limit(1 << 16)
    x = ctx.zeros(n)
    st = s.solve(a, x, rhs)
    assert st.converged and st.iterations == st0.iterations
    assert np.array_equal(x.download(), x0.download())


def test_solver_failure_paths(ctx, oracle):
    """maxIterations reached -> converged = 0 (the reference returns false); GMRES without restart leaves
    x untouched."""
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle, n=24)
    x = ctx.zeros(xy.shape[0])
    st = capi.Solver(a, capi.SOLVER_CG, 1e-30, 5, 0, 0, capi.Preconditioner(a)).solve(a, x, rhs)
    assert not st.converged and st.iterations == 5
    x.upload(np.full(xy.shape[0], 0.25))
    st = capi.Solver(a, capi.SOLVER_GMRES, 1e-30, 6, 0, 0, capi.Preconditioner(a)).solve(a, x, rhs)
    assert not st.converged and st.iterations == 6 and np.all(x.download() == 0.25)
    # an ordering the library does not know is refused with PFEM_EINVAL
    with pytest.raises(capi.PfemError) as e:
        capi.Preconditioner(a, capi.PRECOND_IC0, 7)
    assert e.value.code == capi.PFEM_EINVAL


def test_half_factor(ctx, oracle):
    """16-bit storage of the multicolour IC0 factor (PFEM_PRECISION_MIXED_PRECOND, factor bits 16): the
    apply is that of c L rounded to half precision, c the power of two that puts the largest entry in
    [2^13, 2^14), with c^2 b as the forward solve's input. Checked through the first PCG step, x_1 =
    alpha M^-1 b, against the oracle's apply of the factor round16(c L) / c (the same direction to 1e-13);
    then full solves converge like the 32-bit storage, also on the system scaled by 1e12 and by 1e-12,
    whose factors would overflow or vanish in half precision without the scale."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle)
    n = xy.shape[0]
    b = rhs.download()

    def solver(mat, bits, maxit, tol=1e-12):
        pc = capi.Preconditioner(mat, capi.PRECOND_IC0, capi.ORDER_MULTICOLOR)
        pc.set_factor_bits(bits)
        s = capi.Solver(mat, capi.SOLVER_CG, tol, maxit, 0, capi.TOL_RELATIVE, pc)
        s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 20)
        return pc, s

    pc, s = solver(a, 16, 1, 1e-30)
    x = ctx.zeros(n)
    s.solve(a, x, rhs)
    x1 = x.download()
    fac, nl, perm = pc.download()
    B = sp.csr_matrix((val, col, rp), shape=(n, n))[perm][:, perm].tocsr()
    B.sort_indices()
    c = 2.0 ** (14 - np.frexp(np.abs(fac).max())[1])
    f16 = (fac * c).astype(np.float16).astype(np.float64) / c
    assert not np.array_equal(f16, fac)
    zf = oracle.ic0_apply(B.indptr.astype(np.int32), B.indices.astype(np.int32), f16, b[perm])
    z = np.empty(n)
    z[perm] = zf
    assert rel_l2(x1 / np.linalg.norm(x1), z / np.linalg.norm(z)) < 1e-13
    exact = spla.spsolve(sp.csr_matrix((val, col, rp)).tocsc(), b)
    its = {}
    for scale in (1.0, 1e12, 1e-12):
        ms = capi.CSR(ctx, mesh=m)
        ms.upload_values(np.ascontiguousarray(val * scale))
        for bits in (32, 16):
            mat = a if scale == 1.0 else ms
            pc, s = solver(mat, bits, 5000)
            x = ctx.zeros(n)
            st = s.solve(mat, x, ctx.to_device(b * scale))
            assert st.converged and not st.nan_detected, (scale, bits)
            assert rel_l2(x.download(), exact) < 1e-9, (scale, bits)
            its[(scale, bits)] = st.iterations
    assert all(abs(v - its[(1.0, 32)]) <= 3 for v in its.values()), its
    # 16 bits are for IC0 factors, and are chosen before the reduced-precision copy exists
    for kind in (capi.PRECOND_JACOBI, capi.PRECOND_ILU0):
        with pytest.raises(capi.PfemError):
            capi.Preconditioner(a, kind, capi.ORDER_MULTICOLOR).set_factor_bits(16)
    pc, s = solver(a, 32, 10)
    with pytest.raises(capi.PfemError):
        pc.set_factor_bits(16)
    with pytest.raises(capi.PfemError):
        capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_MULTICOLOR).set_factor_bits(8)


@pytest.mark.parametrize("ordering", [1, 2])
def test_half_factor_wide_range_falls_back(ctx, oracle, ordering):
    """A factor whose entries span more than the 16-bit copy's range (the system D A D with D from 1 to
    1e-10 along x: IC0(D A D) = D IC0(A), so the factor's diagonal runs from about 1 to 1e-10, past the
    2^28 that half precision holds between its largest entry and its smallest normal) would lose diagonals
    to zero or subnormal halves and divide by nothing; the preconditioner then applies its 32-bit copy
    instead: the 16-bit request takes the 32-bit iterations exactly and reaches the direct solution
    (multicolour and blocked orderings)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from gpupfem2_amd import capi
    xy, m, a, rhs, rp, col, val = bc_system(ctx, oracle)
    n = xy.shape[0]
    d = 10.0 ** (-10.0 * xy[:, 0])
    rows = np.repeat(np.arange(n), np.diff(rp))
    ms = capi.CSR(ctx, mesh=m)
    vs = val * d[rows] * d[col]
    ms.upload_values(np.ascontiguousarray(vs))
    b = rhs.download() * d
    exact = spla.spsolve(sp.csr_matrix((vs, col, rp)).tocsc(), b)
    its = {}
    for bits in (32, 16):
        pc = capi.Preconditioner(ms, capi.PRECOND_IC0, ordering)
        pc.set_factor_bits(bits)
        s = capi.Solver(ms, capi.SOLVER_CG, 1e-12, 5000, 0, capi.TOL_RELATIVE, pc)
        s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 20)
        x = ctx.zeros(n)
        st = s.solve(ms, x, ctx.to_device(b))
        assert st.converged and not st.nan_detected, bits
        assert rel_l2(x.download(), exact) < 1e-8, bits
        its[bits] = st.iterations
    assert its[16] == its[32], its


def neighbors_by_incidence(nv, tri, self_included):
    """Cells sharing at least one vertex, ascending, from the pattern of C C^T (C the cell-vertex
    incidence): an independent construction for meshes too large for the brute-force oracle."""
    import scipy.sparse as sp
    nt = tri.shape[0]
    C = sp.csr_matrix((np.ones(3 * nt, np.int8), tri.ravel(), np.arange(0, 3 * nt + 1, 3)), shape=(nt, nv))
    N = (C.astype(np.int32) @ C.T.astype(np.int32)).tocsr()
    if not self_included:
        N.setdiag(0)
        N.eliminate_zeros()
    N.sort_indices()
    return N.indptr.astype(np.int64), N.indices


def test_mesh_setup_full_size(ctx, oracle):
    """Geometry, cell neighbours and the CSR pattern of the 2M-triangle cylinder mesh: geometry and CSR
    bit for bit against the oracle, neighbours against the pattern of C C^T (the construction is first
    checked against the brute-force oracle on a fixture)."""
    from gpupfem2_amd import capi, cases
    v, c, ndev = oracle.load_dat(golden("CylinderMesh.dat"), 2)
    off, idx = oracle.neighbors_bruteforce(c)
    selfin = any(i in idx[off[i]:off[i + 1]] for i in range(10))
    o2, i2 = neighbors_by_incidence(v.shape[0], c, selfin)
    assert np.array_equal(o2, off) and np.array_equal(i2, idx)
    case = cases.cylinder_case()
    xy, tri = case.xy, case.tri
    print("mesh setup full size", tri.shape[0], "triangles", flush=True)
    m = capi.Mesh2D(ctx, xy, tri)
    area, invJ = oracle.geometry2d(xy, tri)
    assert np.array_equal(m.download(2), area)
    assert np.array_equal(m.download(3).reshape(-1, 4), invJ)
    on, ni = neighbors_by_incidence(xy.shape[0], tri, selfin)
    assert np.array_equal(m.download(4), on) and np.array_equal(m.download(5), ni)
    rp, col = oracle.csr_pattern(xy.shape[0], tri)
    grp, gcol, _ = capi.CSR(ctx, mesh=m).download()
    assert np.array_equal(grp, rp) and np.array_equal(gcol, col)


@pytest.mark.parametrize("bits", [64, 16])
def test_pcg_full_size_repeatable(ctx, bits):
    """The cylinder's pressure system (1.01M rows) solved by the fused multicolour IC0-PCG three times
    over (FP64 factor, and the mixed step's 16-bit copy): identical iteration counts and bits each
    time, whatever order the workgroups' publications happen to arrive in, and a true residual at the
    requested tolerance."""
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case()
    print("pcg full size repeatable", bits, flush=True)
    m = capi.Mesh2D(ctx, case.xy, case.tri, fill_neighbors=False)
    a = capi.CSR(ctx, mesh=m)
    n = case.xy.shape[0]
    rhs = ctx.zeros(n)
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=1.0)
    capi.Dirichlet(ctx, n, case.bcp_node, case.bcp_val).apply(a, rhs, capi.DIRICHLET_MASK)
    pc = capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_MULTICOLOR)
    if bits == 16:
        pc.set_factor_bits(16)
    s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 20000, 0, capi.TOL_RELATIVE, pc)
    if bits == 16:
        s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 20)
    ref = None
    for rep in range(3):
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


def test_stream_copy_ceiling(ctx):
    """pfem_stream_copy (the roofline's measured streaming ceiling) times its copy kernel over seven grid
    sizes and two load depths and reports the best; bad sizes are rejected."""
    from gpupfem2_amd import capi
    ms, grid = capi.stream_copy(ctx, 64 << 20, 2)
    assert ms > 0.0 and grid // 100000 in (4, 8) and grid % 100000 in (256, 512, 1024, 2048, 4096, 8192, 16384)
    gbs = 2 * (64 << 20) / (ms * 1e-3) / 1e9
    print("stream copy 64 MiB: %.1f GB/s, grid %d" % (gbs, grid), flush=True)
    assert 100.0 < gbs < 20000.0
    for bad in (0, 24):
        with pytest.raises(capi.PfemError):
            capi.stream_copy(ctx, bad, 1)


def test_locality_renumbering_same_solution(ctx):
    """A mesh numbered at random and the same mesh renumbered by pfem_mesh_locality_order give the same
    Poisson solution (manufactured sin-sin right-hand side, Jacobi PCG to relative 1e-12), node for node
    through the permutation, to 1e-10."""
    from gpupfem2_amd import capi, meshgen
    xy, tri = meshgen.unit_square(48)
    rng = np.random.default_rng(12)
    new_id = rng.permutation(xy.shape[0])
    xs = np.empty_like(xy)
    xs[new_id] = xy
    ts = new_id[tri][rng.permutation(tri.shape[0])].astype(np.int32)
    xr, tr, nodes, _ = capi.renumber_mesh(xs, ts)
    sols = []
    for x, t in ((xs, ts), (xr, tr)):
        m = capi.Mesh2D(ctx, x, t)
        a = capi.CSR(ctx, mesh=m)
        rhs = ctx.zeros(x.shape[0])
        capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=1, rhs_param=2 * math.pi ** 2)
        bn = np.nonzero((x[:, 0] < 1e-9) | (x[:, 0] > 1 - 1e-9) | (x[:, 1] < 1e-9) | (x[:, 1] > 1 - 1e-9))[0]
        capi.Dirichlet(ctx, x.shape[0], bn, np.zeros(bn.size)).apply(a, rhs, capi.DIRICHLET_MASK)
        s = capi.Solver(a, capi.SOLVER_CG, 1e-12, 5000, 0, capi.TOL_RELATIVE, capi.Preconditioner(a, capi.PRECOND_JACOBI))
        u = ctx.zeros(x.shape[0])
        st = s.solve(a, u, rhs)
        assert st.converged
        sols.append(u.download())
    assert rel_l2(sols[1], sols[0][nodes]) < 1e-10
