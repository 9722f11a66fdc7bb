"""The C++ facade (include/gpupfem2.hpp) on the GPU: the reference's Poisson2D driver sequence written
against it reproduces the oracle's run of the same sequence, and the facade's own checks pass (a
Preconditioner subclass plugged into SolverCG, IC0 orderings, restarted GMRES, particles, parameters,
quadrature)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "gpupfem2_amd", "examples")


def run(args, cwd):
    print("[facade] running", " ".join(os.path.basename(a) for a in args), flush=True)
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout + p.stderr


def test_facade_checks(tmp_path):
    rc, out = run([os.path.join(EX, "facade_check"), golden("ChannelMesh.dat"), golden("BoxMesh.dat")], tmp_path)
    print(out)
    assert rc == 0 and "FAIL" not in out and out.strip().endswith("ok 0 failures")
    assert "ok mesh3d_face_neighbors 19898 5393" in out   # the oracle's count on BoxMesh.dat


def test_poisson2d_driver_matches_oracle(tmp_path, oracle):
    rc, out = run([os.path.join(EX, "poisson2d"), golden("TestProblem2.dat"), str(tmp_path)], tmp_path)
    print(out)
    assert rc == 0, out
    O = oracle
    xy, cells, ndev = O.load_dat(golden("TestProblem2.dat"))
    tri = cells[:ndev]
    area, invJ = O.geometry2d(xy, tri)
    rp, col = O.csr_pattern(xy.shape[0], tri)
    val, rhs = O.assemble_poisson2d(xy, tri, area, invJ, rp, col, coeff=0.5, rhs_kind=0, quad=1)
    eps = 2e-6
    node, gval = [], []
    for i, (x, y) in enumerate(xy):
        for cond, g in ((abs(x + 1.0) < eps, -1.0), (abs(x - 1.0) < eps, 1.0), (abs(y) < eps, 0.0),
                        (abs(y - 1.0) < eps, 2.0)):
            if cond:
                node.append(i)
                gval.append(g)
                break
    O.dirichlet_apply(rp, col, val, rhs, np.asarray(node, np.int32), np.asarray(gval), 0)
    # exported matrix and vectors carry 6 significant digits (default stream formatting, as the reference)
    m = np.loadtxt(tmp_path / "matrix.dat")
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert np.array_equal(m[:, 0].astype(int) - 1, rows) and np.array_equal(m[:, 1].astype(int) - 1, col)
    assert np.allclose(m[:, 2], val, rtol=1e-5, atol=1e-12)
    assert rel_l2(np.loadtxt(tmp_path / "rhs.dat"), rhs) < 1e-5
    x = np.zeros(rhs.size)
    st = O.pcg(rp, col, val, x, rhs, 1e-8, 1000, 0, "jacobi")
    assert st.converged
    assert rel_l2(np.loadtxt(tmp_path / "solution_cg.dat"), x) < 1e-5
    O.gmres(rp, col, val, x, rhs, 1e-8, 1000, 1000, 0, "jacobi")
    assert rel_l2(np.loadtxt(tmp_path / "solution.dat"), x) < 1e-5
    assert (tmp_path / "solution.vtu").stat().st_size > 1000


def test_cylinder2d_driver(ctx, tmp_path, oracle):
    """examples/cylinder2d.cpp, the Cylinder2D case driver (CylinderMesh3.dat, Scheme A, dt 0.001, level-2
    particles, Jacobi CG and non-restarted GMRES at the reference's absolute tolerance 1e-8), run for 4
    steps with the step graphs on (captured at steps 2 and 3, one per particle buffer, replayed at step 4):
    Forces.csv in the reference's "Time;Cx;Cy" format, VTU snapshots, and the loads of every step
    equal to those of the same case built in Python (edge IDs and Dirichlet lists from the reference's
    rules in gpupfem2_amd.cases) and stepped through the Python binding, and to the oracle's loads."""
    from gpupfem2_amd import cases
    from helpers import oracle_stepper
    steps = 4
    rc, out = run([os.path.join(EX, "cylinder2d"), golden("CylinderMesh3.dat"), str(steps), str(tmp_path), "2"],
                  tmp_path)
    print(out[-2000:])
    assert rc == 0, out[-2000:]
    lines = (tmp_path / "Forces.csv").read_text().strip().splitlines()
    assert lines[0] == "Time;Cx;Cy" and len(lines) == steps + 1
    got = np.array([[float(v) for v in ln.split(";")] for ln in lines[1:]])
    assert np.allclose(got[:, 0], 0.001 * np.arange(1, steps + 1))
    assert (tmp_path / "solution0.vtu").exists() and (tmp_path / "solution2.vtu").exists()
    v, c, ndev = oracle.load_dat(golden("CylinderMesh3.dat"), 2)
    circle = (0.2, 0.2, 0.05)
    ids = cases.classify_edges(v, c, 0.0, 2.2, 0.0, 0.41, ids=(0, 1, 2, 2), circle=circle, circle_tol=1e-3,
                               circle_id=3)
    bu, bv0, bv1, bp, bpv = cases.channel_bcs(v, 0.0, 2.2, 0.0, 0.41, umax=1.5, circle=circle)
    case = cases.Pfem2Case(
        name="Cylinder2D", xy=v, tri=c, edge_ids=ids, bcu_node=bu, bcu_val0=bv0, bcu_val1=bv1, bcp_node=bp,
        bcp_val=bpv, rho=1.0, mu=0.001, dt=0.001, scheme=0, traction_mask=0b10,
        normals=[(0.0, 0.0), (1.0, 0.0)], outer_iterations=2, substeps=3, pred_solver=cases.GMRES,
        pred_precond=cases.JACOBI, pres_solver=cases.CG, pres_precond=cases.JACOBI, corr_solver=cases.CG,
        corr_precond=cases.JACOBI, tol=1e-8, tol_mode=cases.ABS, maxit=1000, restart=0,
        dirichlet_mode=cases.PARITY, level=2, seed_mode=0, ppc=0, seed=0, body_id=3, body_center=(0.2, 0.2),
        force_coeff=2.0 / (1.0 * 1.0 * 0.1))
    mesh, ph, stepper = cases.build_gpu(ctx, case, precision=0, pres_ordering=0)
    ref = np.array([stepper.step() for _ in range(steps)])
    # Forces.csv holds 6 significant digits, as the reference writes it
    assert np.allclose(got[:, 1:], ref, rtol=1e-5, atol=1e-12), (got, ref)
    _, ost = oracle_stepper(oracle, case)
    orc = np.array([ost.step() for _ in range(steps)])
    # two reduction orders under the absolute solver tolerance 1e-8 leave load differences near 1e-6
    assert np.allclose(ref, orc, rtol=1e-5, atol=1e-8), (ref, orc)


def vtu_field(path, name):
    import re
    text = open(path).read()
    m = re.search(r'<DataArray type="Float32" Name="%s" Format="ascii">(.*?)</DataArray>' % name, text, re.S)
    return np.array(m.group(1).split(), dtype=np.float64)


def test_poiseuille2d_driver(ctx, tmp_path, oracle):
    """examples/poiseuille2d.cpp, the PoiseuilleFlow2D case driver (ChannelMesh.dat, pressure-driven: p = 10
    at x = -5 and 0 at x = 5, no slip on the walls; Scheme A, dt 0.01, level-2 particles, traction on every
    boundary edge), run for 3 steps: the exported fields equal, bit for bit, those of the same case built
    and stepped from Python."""
    from gpupfem2_amd import cases
    steps = 3
    rc, out = run([os.path.join(EX, "poiseuille2d"), golden("ChannelMesh.dat"), str(steps), str(tmp_path), "3"],
                  tmp_path)
    print(out[-1500:])
    assert rc == 0, out[-1500:]
    v, c, ndev = oracle.load_dat(golden("ChannelMesh.dat"), 2)
    ids = cases.classify_edges(v, c, -5.0, 5.0, 0.0, 1.0, ids=(0, 1, 2, 3))
    bu, bp, bpv = [], [], []
    for i, (x, y) in enumerate(v):       # the pressure-driven option, in the reference's if/else order
        if abs(x + 5.0) < cases.TOL_BC:
            bp.append(i); bpv.append(10.0)
        elif abs(x - 5.0) < cases.TOL_BC:
            bp.append(i); bpv.append(0.0)
        elif abs(y) < cases.TOL_BC or abs(y - 1.0) < cases.TOL_BC:
            bu.append(i)
    bu, bp = np.asarray(bu, np.int32), np.asarray(bp, np.int32)
    z = np.zeros(bu.size)
    case = cases.Pfem2Case(
        name="PoiseuilleFlow2D", xy=v, tri=c, edge_ids=ids, bcu_node=bu, bcu_val0=z, bcu_val1=z, bcp_node=bp,
        bcp_val=np.asarray(bpv), rho=1.0, mu=1.0, dt=0.01, scheme=0, traction_mask=0b1111,
        normals=[(-1.0, 0.0), (1.0, 0.0), (0.0, -1.0), (0.0, 1.0)], outer_iterations=1, substeps=3,
        pred_solver=cases.GMRES, pred_precond=cases.JACOBI, pres_solver=cases.CG, pres_precond=cases.JACOBI,
        corr_solver=cases.CG, corr_precond=cases.JACOBI, tol=1e-8, tol_mode=cases.ABS, maxit=1000, restart=0,
        dirichlet_mode=cases.PARITY, level=2, seed_mode=0, ppc=0, seed=0)
    mesh, ph, stepper = cases.build_gpu(ctx, case, precision=0, pres_ordering=0)
    for _ in range(steps):
        stepper.step()
    f = stepper.fields()
    for name, key in (("velX", "u0"), ("velY", "u1"), ("pressure", "p")):
        # the exporter writes 17 significant digits, which round-trip a double
        assert np.array_equal(vtu_field(tmp_path / f"solution{steps}.vtu", name), f[key]), name
    assert np.abs(f["u0"]).max() > 0.0


def test_reference_cylinder_config_vs_oracle(ctx, oracle):
    """The Cylinder2D driver's exact configuration (Scheme A, absolute tolerance 1e-8, non-restarted
    GMRES and CG with Jacobi, the reference's Dirichlet elimination heuristic, level-2 centroid seeding)
    on the smaller CylinderMesh.dat, 3 steps against the oracle: fields within 1e-6 relative L2 and loads
    within 1e-5 (the absolute solver tolerance leaves differences of that order between two reduction
    orders), particle cells identical."""
    from gpupfem2_amd import cases
    from helpers import oracle_stepper
    v, c, ndev = oracle.load_dat(golden("CylinderMesh.dat"), 2)
    circle = (0.2, 0.2, 0.05)
    ids = cases.classify_edges(v, c, 0.0, 2.2, 0.0, 0.41, ids=(0, 1, 2, 2), circle=circle, circle_tol=1e-3,
                               circle_id=3)
    bu, bv0, bv1, bp, bpv = cases.channel_bcs(v, 0.0, 2.2, 0.0, 0.41, umax=1.5, circle=circle)
    case = cases.Pfem2Case(
        name="Cylinder2D", xy=v, tri=c, edge_ids=ids, bcu_node=bu, bcu_val0=bv0, bcu_val1=bv1, bcp_node=bp,
        bcp_val=bpv, rho=1.0, mu=0.001, dt=0.001, scheme=0, traction_mask=0b10,
        normals=[(0.0, 0.0), (1.0, 0.0)], outer_iterations=2, substeps=3, pred_solver=cases.GMRES,
        pred_precond=cases.JACOBI, pres_solver=cases.CG, pres_precond=cases.JACOBI, corr_solver=cases.CG,
        corr_precond=cases.JACOBI, tol=1e-8, tol_mode=cases.ABS, maxit=1000, restart=0,
        dirichlet_mode=cases.PARITY, level=2, seed_mode=0, ppc=0, seed=0, body_id=3, body_center=(0.2, 0.2),
        force_coeff=2.0 / (1.0 * 1.0 * 0.1))
    print("reference cylinder config", c.shape[0], "cells", flush=True)
    mesh, ph, stepper = cases.build_gpu(ctx, case, precision=0, pres_ordering=0)
    oph, ost = oracle_stepper(oracle, case)
    for s in range(3):
        fg, fo = stepper.step(), ost.step()
        assert np.allclose(fg, fo, rtol=1e-5, atol=1e-8), (s, fg, fo)
    g, o = stepper.fields(), ost.fields()
    for k in ("u0", "u1", "p"):
        assert rel_l2(g[k], o[k]) < 1e-6, (k, rel_l2(g[k], o[k]))
    assert np.array_equal(np.sort(ph.get()["cell"]), np.sort(oph.get()["cell"]))


def _vtu_array(text, name):
    import re
    m = re.search(r'Name="%s" Format="ascii">\s*(.*?)\s*</DataArray>' % name, text, re.S)
    return np.array(m.group(1).split(), dtype=np.float64)


def test_poisson3d_driver_matches_oracle(tmp_path, oracle):
    """The reference's Poisson3D driver (cases/Poisson3D/main.cu:87-176) written against the facade
    (Mesh3D, SparseMatrixCSR, NumericalIntegrator3D, DirichletBCs, Jacobi PCG, DataExport3D) on BoxMesh.dat:
    the exported matrix, right-hand side and solution match the oracle's run of the same sequence, and the
    VTU holds the mesh (points, tetrahedra, offsets, type 10) and the solution at full precision."""
    rc, out = run([os.path.join(EX, "poisson3d"), golden("BoxMesh.dat"), str(tmp_path)], tmp_path)
    print(out)
    assert rc == 0, out
    O = oracle
    xyz, cells, ndev = O.load_dat(golden("BoxMesh.dat"), dim=3)
    tet = cells[:ndev]
    vol, invJ = O.geometry3d(xyz, tet)
    rp, col = O.csr_pattern(xyz.shape[0], tet)
    val, rhs = O.assemble_poisson3d(xyz, tet, vol, invJ, rp, col, coeff=0.5, rhs_kind=0, rhs_param=0.0, quad=1)
    eps = 2e-6
    node, gval = [], []
    for i, (x, y, z) in enumerate(xyz):
        for cond, g in ((abs(x + 1.0) < eps, -1.0), (abs(x - 1.0) < eps, 1.0), (abs(y) < eps, 0.0),
                        (abs(y - 1.0) < eps, 2.0), (abs(z + 0.25) < eps, 3.0), (abs(z - 0.25) < eps, -2.0)):
            if cond:
                node.append(i)
                gval.append(g)
                break
    assert len(node) > 0
    O.dirichlet_apply(rp, col, val, rhs, np.asarray(node, np.int32), np.asarray(gval), 0)
    m = np.loadtxt(tmp_path / "matrix.dat")
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert np.array_equal(m[:, 0].astype(int) - 1, rows) and np.array_equal(m[:, 1].astype(int) - 1, col)
    assert np.allclose(m[:, 2], val, rtol=1e-5, atol=1e-12)
    assert rel_l2(np.loadtxt(tmp_path / "rhs.dat"), rhs) < 1e-5
    x = np.zeros(rhs.size)
    st = O.pcg(rp, col, val, x, rhs, 1e-8, 1000, 0, "jacobi")
    assert st.converged
    assert rel_l2(np.loadtxt(tmp_path / "solution.dat"), x) < 1e-5
    text = (tmp_path / "solution.vtu").read_text()
    assert f'NumberOfPoints="{xyz.shape[0]}" NumberOfCells="{tet.shape[0]}"' in text
    assert np.array_equal(_vtu_array(text, "connectivity").astype(int), tet.ravel())
    assert np.array_equal(_vtu_array(text, "offsets").astype(int), 4 * np.arange(1, tet.shape[0] + 1))
    assert np.all(_vtu_array(text, "types") == 10)
    # both PCGs stop at an absolute (r, M^-1 r)^1/2 of 1e-8 (iterates that differ by the reduction order of
    # the dot products, and possibly one iteration): solutions agree far below the 6 digits of solution.dat
    assert rel_l2(_vtu_array(text, "solution"), x) < 1e-6
