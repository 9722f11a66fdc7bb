"""Pins of the oracle's PFEM-2 operators without the reference: on constant and linear fields the P1
operators act exactly like the continuous ones, so each integrand is checked against closed forms derived
from the vertex coordinates alone (gradients of the barycentric coordinates from edge perpendiculars, not
the oracle's inverse Jacobian). A misplaced 4/3 (x vs y), a swapped index, a sign error, a dropped term or
a wrong quadrature-weight sum fails one of them. The whole step is pinned by Poiseuille flow, whose
analytic profile and pressure gradient are the discrete steady state up to O(h^2).

Reference integrands: cases/Cylinder2D/main.cu:72-167 (prediction), :169-211 (pressure), :213-259
(correction), :261-297 (prediction BC values), :299-445 (loads); particle_handler_2d.cu:72-88 (velocity
correction); cases/PoiseuilleFlow2D/main.cu:652-736 (step)."""
import math

import numpy as np
import pytest


def dense(rp, col, val):
    n = rp.size - 1
    a = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            a[i, col[k]] += val[k]
    return a


def grads(xy):
    """Gradients of the barycentric coordinates of a triangle, and its (positive) area."""
    (x0, y0), (x1, y1), (x2, y2) = xy
    s = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    g = np.array([[y1 - y2, x2 - x1], [y2 - y0, x0 - x2], [y0 - y1, x1 - x0]]) / (2 * s)
    return g, abs(s)


@pytest.fixture
def one_triangle(oracle):
    xy = np.array([[0.13, 0.21], [1.37, 0.42], [0.55, 1.16]])
    tri = np.array([[0, 1, 2]], np.int32)
    area, invJ = oracle.geometry2d(xy, tri)
    rp, col = oracle.csr_pattern(3, tri)
    g, a = grads(xy)
    assert np.isclose(area[0], a, rtol=1e-14)
    # sanity of the closed-form gradients: sum to zero, reproduce grad x and grad y
    assert np.allclose(g.sum(0), 0.0, atol=1e-14) and np.allclose(xy.T @ g, np.eye(2), atol=1e-14)
    return xy, tri, area, invJ, rp, col, g, a


RHO, MU, DT = 1.7, 0.013, 0.05
NX, NY = 0.6, 0.8   # normal of boundary ID 1


def prediction(oracle, tri_data, u0, u1, uo0, uo1, p_old, edge_id1=True):
    xy, tri, area, invJ, rp, col, g, a = tri_data
    scheme = 0 if p_old is None else 1
    prm = oracle.make_params(RHO, MU, DT, scheme, 0b10, [(0.0, 0.0), (NX, NY)])
    edge_ids = np.array([[1 if edge_id1 else -1, -1, -1]], np.int32)   # edge 0 = vertices 0 -> 1
    v0, v1, r0, r1 = oracle.assemble_prediction(xy, tri, area, invJ, edge_ids, rp, col, u0, u1, uo0, uo1,
                                                p_old, prm)
    return dense(rp, col, v0), dense(rp, col, v1), r0, r1


def test_prediction_matrices_on_linear_fields(oracle, one_triangle):
    xy, tri, area, invJ, rp, col, g, a = one_triangle
    z = np.zeros(3)
    A0, A1, _, _ = prediction(oracle, one_triangle, z, z, z, z, None)
    x, y, one = xy[:, 0], xy[:, 1], np.ones(3)
    chi = np.array([1.0, 1.0, 0.0])                 # nodes of the traction edge
    L = np.linalg.norm(xy[1] - xy[0])
    mass = lambda f: RHO * a / 12.0 * (f + f.sum())    # rho * int N_i f, f linear
    # reference quirk: the edge integrals are scaled by the cell area with the rest (addLocalToGlobal)
    edge = MU * DT * a * (L / 2) * chi
    assert np.allclose(A0 @ one, RHO * a / 3, rtol=1e-13)
    assert np.allclose(A1 @ one, RHO * a / 3, rtol=1e-13)
    # x-system: mu dt int(dyNi dy f + 4/3 dxNi dx f) - mu dt int_edge Ni (4/3 dx f nx + dy f ny)
    assert np.allclose(A0 @ x, mass(x) + MU * DT * a * 4 / 3 * g[:, 0] - edge * 4 / 3 * NX, rtol=1e-12, atol=1e-15)
    assert np.allclose(A0 @ y, mass(y) + MU * DT * a * g[:, 1] - edge * NY, rtol=1e-12, atol=1e-15)
    # y-system: the 4/3 moves to the y derivative
    assert np.allclose(A1 @ x, mass(x) + MU * DT * a * g[:, 0] - edge * NX, rtol=1e-12, atol=1e-15)
    assert np.allclose(A1 @ y, mass(y) + MU * DT * a * 4 / 3 * g[:, 1] - edge * 4 / 3 * NY, rtol=1e-12, atol=1e-15)
    # without the traction edge the matrices are symmetric
    B0, B1, _, _ = prediction(oracle, one_triangle, z, z, z, z, None, edge_id1=False)
    assert np.allclose(B0, B0.T, rtol=1e-13) and np.allclose(B1, B1.T, rtol=1e-13)


def test_prediction_rhs_on_linear_fields(oracle, one_triangle):
    xy, tri, area, invJ, rp, col, g, a = one_triangle
    x, y, one, z = xy[:, 0], xy[:, 1], np.ones(3), np.zeros(3)
    chi = np.array([1.0, 1.0, 0.0])
    L = np.linalg.norm(xy[1] - xy[0])
    edge = MU * DT * a * (L / 2) * chi
    # uniform old velocity: rho int N_i a; uniform current velocity: no cross terms
    _, _, r0, r1 = prediction(oracle, one_triangle, 3 * one, 5 * one, 2 * one, -1 * one, None)
    assert np.allclose(r0, RHO * a / 3 * 2, rtol=1e-13) and np.allclose(r1, -RHO * a / 3, rtol=1e-13)
    # cross term of the x-equation: - mu dt int (dyNi dx u_y - 2/3 dxNi dy u_y) + edge (-2/3 dy u_y nx + dx u_y ny)
    _, _, r0, _ = prediction(oracle, one_triangle, z, x, z, z, None)
    assert np.allclose(r0, -MU * DT * a * g[:, 1] + edge * NY, rtol=1e-12, atol=1e-16)
    _, _, r0, _ = prediction(oracle, one_triangle, z, y, z, z, None)
    assert np.allclose(r0, MU * DT * a * 2 / 3 * g[:, 0] - edge * 2 / 3 * NX, rtol=1e-12, atol=1e-16)
    # y-equation mirrored: - mu dt int (dxNi dy u_x - 2/3 dyNi dx u_x) + edge (dy u_x nx - 2/3 dx u_x ny)
    _, _, _, r1 = prediction(oracle, one_triangle, y, z, z, z, None)
    assert np.allclose(r1, -MU * DT * a * g[:, 0] + edge * NX, rtol=1e-12, atol=1e-16)
    _, _, _, r1 = prediction(oracle, one_triangle, x, z, z, z, None)
    assert np.allclose(r1, MU * DT * a * 2 / 3 * g[:, 1] - edge * 2 / 3 * NY, rtol=1e-12, atol=1e-16)
    # scheme B: - dt int N_i grad p_old
    _, _, r0, r1 = prediction(oracle, one_triangle, z, z, z, z, 2 * x + 3 * y)
    assert np.allclose(r0, -DT * a / 3 * 2, rtol=1e-13) and np.allclose(r1, -DT * a / 3 * 3, rtol=1e-13)


def test_pressure_operator_on_linear_fields(oracle, one_triangle):
    xy, tri, area, invJ, rp, col, g, a = one_triangle
    x, y, one, z = xy[:, 0], xy[:, 1], np.ones(3), np.zeros(3)
    for scheme, p_old in ((0, None), (1, 4 * y - x)):
        prm = oracle.make_params(RHO, MU, DT, scheme)
        val, rhs = oracle.assemble_pressure(tri, area, invJ, rp, col, 2 * x, -0.5 * y, p_old, prm)
        K = dense(rp, col, val)
        assert np.allclose(K @ one, 0.0, atol=1e-13)
        assert np.allclose(K @ x, a * g[:, 0], rtol=1e-12) and np.allclose(K @ y, a * g[:, 1], rtol=1e-12)
        # - rho/dt int N_i div V*, div V* = 2 - 0.5; scheme B adds K p_old
        expect = -RHO / DT * 1.5 * a / 3
        if scheme:
            expect = expect + a * (4 * g[:, 1] - g[:, 0])
        assert np.allclose(rhs, expect, rtol=1e-12)


def test_correction_rhs_on_linear_fields(oracle, one_triangle):
    xy, tri, area, invJ, rp, col, g, a = one_triangle
    x, y, one, z = xy[:, 0], xy[:, 1], np.ones(3), np.zeros(3)
    prm = oracle.make_params(RHO, MU, DT, 1)
    # rho int N_i V*_k - dt int N_i d_k (p - p_old)
    val, r0, r1 = oracle.assemble_correction(tri, area, invJ, rp, col, 2 * one, -one, 3 * x + y, y, prm)
    assert np.allclose(dense(rp, col, val), RHO * a / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-13)
    assert np.allclose(r0, RHO * a / 3 * 2 - DT * a / 3 * 3, rtol=1e-13)
    assert np.allclose(r1, -RHO * a / 3 - DT * a / 3 * 0.0, rtol=1e-13)


def test_prediction_bc_values_linear_pressure(oracle):
    from gpupfem2_amd import meshgen
    xy, tri = meshgen.rectangle(0.0, 2.0, 0.0, 1.0, 8, 4, jitter=0.2, seed=7)
    area, invJ = oracle.geometry2d(xy, tri)
    prm = oracle.make_params(RHO, MU, DT, 1)
    bn = np.nonzero(xy[:, 0] < 1e-12)[0].astype(np.int32)
    g = np.linspace(0.1, 0.9, bn.size)
    p, p_old = 2 * xy[:, 0] - 5 * xy[:, 1], 0.5 * xy[:, 0]
    # g* = g + dt/rho * (area-weighted mean over the node's cells of d_k (p - p_old)): the gradient is uniform
    for k, d in ((0, 1.5), (1, -5.0)):
        out = oracle.prediction_bc_values(xy.shape[0], tri, area, invJ, p, p_old, bn, g, k, prm)
        assert np.allclose(out, g + DT / RHO * d, rtol=1e-12)
    out = oracle.prediction_bc_values(xy.shape[0], tri, area, invJ, p, p, bn, g, 0, prm)
    assert np.array_equal(out, g)


def test_body_forces_closed_forms(oracle):
    """On the cylinder channel's body ring (a regular polygon, radius R about the centre): constant p gives
    no pressure force; p = x - cx gives F_x = -pi R^2 (divergence theorem, O(h^2) polygon error); a
    velocity at rest gives no viscous force."""
    from gpupfem2_amd import cases
    R = 0.5
    errs = []
    for h in (1.0 / 8, 1.0 / 16):
        case = cases.cylinder_case(h=h)
        xy, tri = case.xy, case.tri
        area, invJ = oracle.geometry2d(xy, tri)
        cx, cy = case.body_center
        prm = oracle.make_params(1.0, 0.01, 1e-3, 1)
        ce = oracle.body_edges(case.edge_ids, 3)
        z = np.zeros(xy.shape[0])
        f = oracle.body_forces(xy, tri, invJ, ce, z, z, np.ones(xy.shape[0]), (cx, cy), prm)
        assert abs(f[0]) < 1e-12 and abs(f[1]) < 1e-12 and f[2] == 0.0 and f[3] == 0.0
        fx = oracle.body_forces(xy, tri, invJ, ce, z, z, xy[:, 0] - cx, (cx, cy), prm)
        fy = oracle.body_forces(xy, tri, invJ, ce, z, z, xy[:, 1] - cy, (cx, cy), prm)
        assert abs(fx[1]) < 1e-10 and abs(fy[0]) < 1e-10
        assert abs(fx[0] - fy[1]) < 1e-12
        errs.append(abs(fx[0] + math.pi * R * R) / (math.pi * R * R))
    assert errs[0] < 5e-3 and errs[1] < errs[0] / 3, errs


def test_particle_velocity_correction(oracle):
    from gpupfem2_amd import meshgen
    xy, tri = meshgen.rectangle(0.0, 1.0, 0.0, 1.0, 6, 6, jitter=0.2, seed=9)
    area, invJ = oracle.geometry2d(xy, tri)
    off, idx = oracle.neighbors_bruteforce(tri)
    nv = xy.shape[0]
    ph = oracle.Particles(xy, tri, invJ, off, idx, 2)
    ph.seed(1, 5, 21)
    rng = np.random.default_rng(1)
    u0, u1 = rng.standard_normal(nv), rng.standard_normal(nv)
    ph.init_velocity(u0, u1)
    v_init = ph.get()["vel"].copy()
    # u = u_old: unchanged; u = u_old + c: every particle shifts by c (the lambdas sum to one)
    ph.correct(u0, u1, u0, u1)
    assert np.array_equal(ph.get()["vel"], v_init)
    ph.correct(u0 + 0.25, u1 - 2.0, u0, u1)
    assert np.allclose(ph.get()["vel"], v_init + [0.25, -2.0], rtol=0, atol=1e-14)


def test_poiseuille_steady_state(oracle):
    """From the analytic profile u = 4 y (1 - y), v = 0 (velocity-driven channel, outlet corners free as in
    cases/PoiseuilleFlow2D/main.cu:539-548), the flow stays at it: the x-averaged velocity profile stays
    within O(h^2) + particle noise of it, and the pressure gradient averaged over steps 20-40 is the exact
    -8 mu U_max. A dropped viscous term gives a gradient near zero; a sign error diverges."""
    from gpupfem2_amd import cases
    from helpers import oracle_stepper
    case = cases.poiseuille_case(nx=24, ny=8, jitter=0.0, ppc=8, dt=1e-2)
    ph, st = oracle_stepper(oracle, case)
    x, y = case.xy[:, 0], case.xy[:, 1]
    u_exact = 4.0 * y * (1.0 - y)
    z = np.zeros_like(y)
    ph.init_velocity(u_exact, z)
    st.set_fields(u0=u_exact, u1=z, p=z, vp0=u_exact, vp1=z)
    inner = (x > 0.5) & (x < 3.5)
    slopes, prof = [], []
    levels = np.unique(np.round(y, 9))
    for s in range(40):
        st.step()
        if s >= 20:
            f = st.fields()
            slopes.append(np.polyfit(x[inner], f["p"][inner], 1)[0])
            prof.append([f["u0"][inner & (np.abs(y - yl) < 1e-9)].mean() for yl in levels])
            assert np.abs(f["u1"][x < 4.0 - 1e-9]).max() < 0.1 and np.isfinite(f["p"]).all()
    prof = np.mean(prof, axis=0)
    assert np.abs(prof - 4.0 * levels * (1.0 - levels)).max() < 0.02, prof
    assert abs(np.mean(slopes) - (-8 * case.mu * 1.0)) < 0.2 * 8 * case.mu, np.mean(slopes)


_OMP_STEP = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import oracle as O
O.use_openmp_build()
from gpupfem2_amd import cases
sys.path.insert(0, sys.argv[1] + "/tests")
from helpers import oracle_stepper
case = cases.cylinder_case(h=1.0 / 16, ppc=6, dt=2e-3)
ph, st = oracle_stepper(O, case)
st.set_coloured_pressure(True)
f = [st.step() for _ in range(2)]
np.savez(sys.argv[2], forces=np.array(f), its=st.iterations(), **st.fields())
"""


def test_openmp_build_coloured_step(oracle, tmp_path):
    """The OpenMP perf build of the oracle (bench.py's CPU baseline), with its multicolour IC0 pressure
    solve, run in a child process on the h = 1/16 cylinder: two steps agree with the parity build's
    natural-order steps to the solver tolerance (fields within relative L2 1e-8, C_D and C_L within 1e-6)."""
    import os
    import subprocess
    import sys
    from gpupfem2_amd import cases
    from helpers import oracle_stepper
    out = tmp_path / "omp.npz"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([sys.executable, "-c", _OMP_STEP, root, str(out)], check=True, timeout=600,
                   env=dict(os.environ, OMP_NUM_THREADS="4"))
    omp = np.load(out)
    case = cases.cylinder_case(h=1.0 / 16, ppc=6, dt=2e-3)
    ph, st = oracle_stepper(oracle, case)
    f = np.array([st.step() for _ in range(2)])
    ref = st.fields()
    for k in ("u0", "u1", "p"):
        err = np.linalg.norm(omp[k] - ref[k]) / np.linalg.norm(ref[k])
        assert err < 1e-8, (k, err)
    assert np.all(np.abs(omp["forces"] - f) <= 1e-6 * np.abs(f) + 1e-12), (omp["forces"], f)
    assert omp["its"][2] > 0
