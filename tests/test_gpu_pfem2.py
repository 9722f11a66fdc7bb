"""GPU parity of the PFEM-2 operators and of whole time steps against the oracle.

Element assembly, prediction BC values, particle seeding/advection/relocation/refill/ordering,
projection and correction are bit-exact (same operation order); body forces are within a few ulp
(reduction order); whole steps are within a relative L2 of 1e-8 on every field (the Krylov solves
agree to their tolerance, not to the bit) with identical particle -> cell indices.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import oracle_params, oracle_stepper

pytestmark = pytest.mark.gpu


def small_cases():
    from gpupfem2_amd import cases
    return [cases.poiseuille_case(nx=40, ny=10, ppc=9, dt=2e-3), cases.cylinder_case(h=1.0 / 8, ppc=6, dt=2e-3)]


@pytest.mark.parametrize("ci,scheme", [(0, 0), (0, 1), (1, 0), (1, 1), (2, 1)])
def test_pfem2_assembly_bitexact(ctx, oracle, ci, scheme):
    """The five PFEM-2 systems, the prediction BC values and the body forces, bit for bit against the
    oracle; ci = 2 is the full 2M-triangle cylinder mesh."""
    from gpupfem2_amd import capi, cases
    case = small_cases()[ci] if ci < 2 else cases.cylinder_case()
    case.scheme = scheme
    print("assembly", case.name, "scheme", scheme, flush=True)
    nv = case.xy.shape[0]
    m = capi.Mesh2D(ctx, case.xy, case.tri)
    m.set_edge_ids(case.edge_ids)
    rng = np.random.default_rng(ci * 10 + scheme)
    f = {k: rng.standard_normal(nv) for k in ("u0", "u1", "uo0", "uo1", "vp0", "vp1", "p", "pold")}
    d = {k: ctx.to_device(v) for k, v in f.items()}
    prm_g = capi.make_params(case.rho, case.mu, case.dt, scheme, case.traction_mask, case.normals)
    prm_o = oracle_params(oracle, case)
    pold_g = d["pold"] if scheme == 1 else None
    pold_o = f["pold"] if scheme == 1 else None
    area, invJ = oracle.geometry2d(case.xy, case.tri)
    rp, col = oracle.csr_pattern(nv, case.tri)
    a0, a1 = capi.CSR(ctx, mesh=m), None
    a1 = capi.CSR(ctx, like=a0)
    r0, r1 = ctx.empty(nv), ctx.empty(nv)
    capi.assemble_prediction(m, a0, a1, r0, r1, d["u0"], d["u1"], d["uo0"], d["uo1"], pold_g, prm_g)
    v0, v1, q0, q1 = oracle.assemble_prediction(case.xy, case.tri, area, invJ, case.edge_ids, rp, col, f["u0"],
                                                 f["u1"], f["uo0"], f["uo1"], pold_o, prm_o)
    assert np.array_equal(a0.download()[2], v0) and np.array_equal(a1.download()[2], v1)
    assert np.array_equal(r0.download(), q0) and np.array_equal(r1.download(), q1)
    capi.assemble_pressure(m, a0, r0, d["vp0"], d["vp1"], pold_g, prm_g)
    v, q = oracle.assemble_pressure(case.tri, area, invJ, rp, col, f["vp0"], f["vp1"], pold_o, prm_o)
    assert np.array_equal(a0.download()[2], v) and np.array_equal(r0.download(), q)
    capi.assemble_correction(m, a0, r0, r1, d["vp0"], d["vp1"], d["p"], pold_g, prm_g)
    v, q0, q1 = oracle.assemble_correction(case.tri, area, invJ, rp, col, f["vp0"], f["vp1"], f["p"], pold_o, prm_o)
    assert np.array_equal(a0.download()[2], v)
    assert np.array_equal(r0.download(), q0) and np.array_equal(r1.download(), q1)
    # prediction BC values
    for comp in (0, 1):
        src = capi.Dirichlet(ctx, nv, case.bcu_node, case.bcu_val0)
        tgt = capi.Dirichlet(ctx, nv, case.bcu_node, case.bcu_val0)
        capi.prediction_bc_values(m, tgt, src, d["p"], pold_g, comp, prm_g)
        ref = oracle.prediction_bc_values(nv, case.tri, area, invJ, f["p"], pold_o, case.bcu_node, case.bcu_val0,
                                          comp, prm_o)
        assert np.array_equal(tgt.values(), ref)
    if case.body_id >= 0:
        g = capi.body_forces(m, case.body_id, case.body_center, d["u0"], d["u1"], d["p"], prm_g)
        ce = oracle.body_edges(case.edge_ids, case.body_id)
        assert ce.shape[0] > 0
        o = oracle.body_forces(case.xy, case.tri, invJ, ce, f["u0"], f["u1"], f["p"], case.body_center, prm_o)
        assert np.array_equal(g, o)


@pytest.mark.parametrize("mode,ppc,level", [(0, 0, 2), (1, 20, 4)])
def test_particles_bitexact(ctx, oracle, mode, ppc, level):
    """seed, init velocity, 2 advect cycles (substeps, relocation, deletion at the outlet, refill,
    ordering), projection and correction."""
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=1.0 / 8)
    nv = case.xy.shape[0]
    m = capi.Mesh2D(ctx, case.xy, case.tri)
    ph = capi.Particles(m, level)
    ph.seed(mode, ppc, 99)
    area, invJ = oracle.geometry2d(case.xy, case.tri)
    off, idx = oracle.neighbors_bruteforce(case.tri)
    oph = oracle.Particles(case.xy, case.tri, invJ, off, idx, level)
    oph.seed(mode, ppc, 99)

    def same():
        g, o = ph.get(), oph.get()
        for k in ("cell", "id", "lam", "pos", "vel"):
            assert np.array_equal(g[k], o[k]), k

    same()
    rng = np.random.default_rng(7)
    # a velocity field that pushes particles across cells and out through the outlet (x = 20)
    u0 = 40.0 + 5.0 * rng.standard_normal(nv)
    u1 = 5.0 * rng.standard_normal(nv)
    du0, du1 = ctx.to_device(u0), ctx.to_device(u1)
    ph.init_velocity(du0, du1)
    oph.init_velocity(u0, u1)
    same()
    for cycle in range(2):
        ph.advect(du0, du1, 0.01, 3)
        oph.advect(u0, u1, 0.01, 3)
        assert ph.stats() == oph.stats()
        same()
    assert oph.stats()[0] > 0 and oph.stats()[1] > 0   # deletions and refills happened
    p0, p1 = ctx.empty(nv), ctx.empty(nv)
    ph.project(p0, p1)
    q0, q1 = oph.project(nv)
    assert np.array_equal(p0.download(), q0) and np.array_equal(p1.download(), q1)
    w0, w1 = rng.standard_normal(nv), rng.standard_normal(nv)
    ph.correct(ctx.to_device(w0), ctx.to_device(w1), du0, du1)
    oph.correct(w0, w1, u0, u1)
    same()


@pytest.mark.parametrize("ci,steps", [(0, 3), (1, 2)])
def test_step_parity(ctx, oracle, ci, steps):
    from gpupfem2_amd import cases
    case = small_cases()[ci]
    print("step parity", case.name, flush=True)
    mesh, ph, stepper = cases.build_gpu(ctx, case, pres_ordering=0)
    oph, ost = oracle_stepper(oracle, case)
    for s in range(steps):
        fg = stepper.step()
        fo = ost.step()
        g, o = stepper.fields(), ost.fields()
        for k in ("u0", "u1", "p", "vp0", "vp1"):
            assert rel_l2(g[k], o[k]) < 1e-8, (s, k, rel_l2(g[k], o[k]))
        assert np.array_equal(ph.get()["cell"], oph.get()["cell"])
        if case.body_id >= 0:
            assert np.allclose(fg, fo, rtol=1e-6, atol=1e-9), (fg, fo)
    its = stepper.iterations()
    assert its.size == 5 * case.outer_iterations and np.all(its > 0)


@pytest.mark.parametrize("precision,ordering", [(1, 0), (1, 1), (1, 2), (0, 2)])
def test_step_mixed_precision(ctx, oracle, precision, ordering):
    """Mixed precision (FP32 inner Krylov + FP64 refinement for the velocity solves; the pressure IC0 factor
    applied from an FP32 copy, or with the multicolour and blocked orderings from a 16-bit copy), and the
    FP64 step with the blocked pressure factor: fields within 1e-8 of the FP64 oracle after 2 steps."""
    from gpupfem2_amd import cases
    case = small_cases()[1]
    print("step precision", precision, "ordering", ordering, flush=True)
    mesh, ph, stepper = cases.build_gpu(ctx, case, precision=precision, inner_tol=1e-5, pres_ordering=ordering)
    oph, ost = oracle_stepper(oracle, case)
    for s in range(2):
        fg, fo = stepper.step(), ost.step()
    g, o = stepper.fields(), ost.fields()
    for k in ("u0", "u1", "p"):
        assert rel_l2(g[k], o[k]) < 1e-8, (k, rel_l2(g[k], o[k]))
    assert np.allclose(fg, fo, rtol=1e-6, atol=1e-9)


def test_particles_full_size(ctx, oracle):
    """The particle kernels at the 2M-triangle cylinder's size (40.4M particles, 20 per cell, 16
    subcells): seed, velocity initialisation, one advect cycle (3 substeps with relocation, deletion at
    the outlet, refill and reordering), projection and correction, bit for bit against the oracle. The
    oracle walks the library's neighbour lists (pinned against brute force on small meshes)."""
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case()
    xy, tri = case.xy, case.tri
    nv = xy.shape[0]
    print("particles full size", tri.shape[0], "triangles", flush=True)
    m = capi.Mesh2D(ctx, xy, tri)
    ph = capi.Particles(m, case.level)
    ph.seed(case.seed_mode, case.ppc, case.seed)
    area, invJ = oracle.geometry2d(xy, tri)
    oph = oracle.Particles(xy, tri, invJ, m.download(4), m.download(5), case.level)
    oph.seed(case.seed_mode, case.ppc, case.seed)

    def same():
        g, o = ph.get(), oph.get()
        for k in ("cell", "id", "lam", "pos", "vel"):
            assert np.array_equal(g[k], o[k]), k

    rng = np.random.default_rng(11)
    # about a cell per substep near the cylinder, and particles leaving through the outlet (x = 20)
    u0 = 1.5 + 0.5 * rng.standard_normal(nv)
    u1 = 0.5 * rng.standard_normal(nv)
    du0, du1 = ctx.to_device(u0), ctx.to_device(u1)
    ph.init_velocity(du0, du1)
    oph.init_velocity(u0, u1)
    ph.advect(du0, du1, 0.01, 3)
    oph.advect(u0, u1, 0.01, 3)
    assert ph.stats() == oph.stats()
    assert oph.stats()[0] > 0 and oph.stats()[1] > 0   # deletions and refills happened
    same()
    p0, p1 = ctx.empty(nv), ctx.empty(nv)
    ph.project(p0, p1)
    q0, q1 = oph.project(nv)
    assert np.array_equal(p0.download(), q0) and np.array_equal(p1.download(), q1)
    ph.correct(p0, p1, du0, du1)
    oph.correct(q0, q1, u0, u1)
    same()


@pytest.mark.parametrize("ci,precision", [(0, 0), (1, 0), (1, 1)])
def test_step_graphs_bitexact(ctx, ci, precision):
    """Whole-step orchestration (pfem_pfem2_set_graphs): the work between the solves captured as HIP graphs
    at the second step and replayed at the later ones gives the same bits as the eager steps -- fields,
    forces, Krylov iteration counts and every particle array -- with eager work (a field download)
    between the replays. The particle step keeps its counts on the device (no host round trip)."""
    from gpupfem2_amd import cases
    case = small_cases()[ci]
    print("step graphs", case.name, "precision", precision, flush=True)
    runs = []
    for graphs in (False, True):
        mesh, ph, stepper = cases.build_gpu(ctx, case, precision=precision, inner_tol=1e-5)
        if graphs:
            stepper.set_graphs(True)
        out = []
        for s in range(4):
            f = stepper.step()
            out.append((f.copy(), stepper.fields(), stepper.iterations().copy(), ph.get(), ph.stats(), ph.count()))
        runs.append(out)
        del stepper, ph, mesh
    for s, (a, b) in enumerate(zip(*runs)):
        assert np.array_equal(a[0], b[0]), (s, a[0], b[0])
        for k in a[1]:
            assert np.array_equal(a[1][k], b[1][k]), (s, k)
        assert np.array_equal(a[2], b[2]), s
        for k in a[3]:
            assert np.array_equal(a[3][k], b[3][k]), (s, k)
        assert a[4] == b[4] and a[5] == b[5], s


def test_step_graphs_after_growth(ctx):
    """The step graphs are captured again when the particle buffers are reallocated: after two steps (the
    graphs captured), every particle is uploaded three times over (each copy right after the particle, so
    the set stays ordered by cell), which grows the buffers (their capacity is kept at the count plus one
    particle per subcell of every cell); the graph steps that follow still give the eager steps' bits."""
    from gpupfem2_amd import cases
    case = small_cases()[1]
    print("step graphs after growth", case.name, flush=True)
    runs = []
    for graphs in (False, True):
        mesh, ph, stepper = cases.build_gpu(ctx, case)
        if graphs:
            stepper.set_graphs(True)
        out = []
        for s in range(5):
            if s == 3:
                st = ph.get()
                dup = {k: np.repeat(v, 3, axis=0) for k, v in st.items()}
                ph.set(dup, int(st["id"].max()) + 1)
                assert ph.count() == 3 * st["cell"].size
            f = stepper.step()
            out.append((f.copy(), stepper.fields(), ph.get(), ph.count()))
        runs.append(out)
        del stepper, ph, mesh
    for s, (a, b) in enumerate(zip(*runs)):
        assert np.array_equal(a[0], b[0]), (s, a[0], b[0])
        for k in a[1]:
            assert np.array_equal(a[1][k], b[1][k]), (s, k)
        for k in a[2]:
            assert np.array_equal(a[2][k], b[2][k]), (s, k)
        assert a[3] == b[3], s


def test_step_graphs_after_mesh_update(ctx):
    """The step graphs hold the mesh's edge-ID and body-cache addresses; setting the edge IDs again (which
    reallocates them and drops the body cache) between steps makes the next graph step capture afresh and
    rebuild the body cache outside the capture, and the steps keep the eager steps' bits."""
    from gpupfem2_amd import cases
    case = small_cases()[1]
    print("step graphs after mesh update", case.name, flush=True)
    runs = []
    for graphs in (False, True):
        mesh, ph, stepper = cases.build_gpu(ctx, case)
        if graphs:
            stepper.set_graphs(True)
        out = []
        for s in range(5):
            if s == 3:
                mesh.set_edge_ids(case.edge_ids)
            out.append((stepper.step().copy(), stepper.fields()))
        runs.append(out)
        del stepper, ph, mesh
    for s, (a, b) in enumerate(zip(*runs)):
        assert np.array_equal(a[0], b[0]), (s, a[0], b[0])
        for k in a[1]:
            assert np.array_equal(a[1][k], b[1][k]), (s, k)
