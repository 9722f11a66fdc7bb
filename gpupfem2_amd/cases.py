"""Case setups: geometry, edge boundary IDs, Dirichlet node lists, physical and solver parameters.

These mirror the host-side setup of the reference drivers (cases/Cylinder2D/main.cu:635-791,
cases/PoiseuilleFlow2D/main.cu:496-647, cases/Poisson2D/main.cu:87-150) for the benchmark workloads:
  * poisson   -- unit square, manufactured u = sin(pi x) sin(pi y), homogeneous Dirichlet;
  * poiseuille -- channel [0,4]x[0,1], parabolic inflow, no-slip walls, zero-pressure outlet;
  * cylinder  -- channel [0,20]x[0,4] with a cylinder of diameter 1 at (4, 2.1), Re = 100.
A case is plain data (numpy arrays and numbers) so that the same inputs can be handed to any
implementation of the step.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import meshgen

TOL_BC = 2e-6   # CONSTANTS::DOUBLE_MIN (src/common/constants.h:5)


def classify_edges(xy, tri, x0, x1, y0, y1, ids=(0, 1, 2, 3), circle=None, circle_tol=1e-3, circle_id=3):
    """Edge boundary IDs by midpoint, the if/else order of kSetEdgeBoundaryIDs
    (cases/Cylinder2D/main.cu:52-68)."""
    a = xy[tri]
    b = xy[np.roll(tri, -1, axis=1)]
    m = (a + b) * 0.5
    mx, my = m[..., 0], m[..., 1]
    out = -np.ones(tri.shape, np.int32)
    c1 = np.abs(mx - x0) < TOL_BC
    c2 = ~c1 & (np.abs(mx - x1) < TOL_BC)
    c3 = ~c1 & ~c2 & (np.abs(my - y0) < TOL_BC)
    c4 = ~c1 & ~c2 & ~c3 & (np.abs(my - y1) < TOL_BC)
    out[c1], out[c2], out[c3], out[c4] = ids[0], ids[1], ids[2], ids[3]
    if circle is not None:
        cx, cy, r = circle
        d = np.sqrt((mx - cx) ** 2 + (my - cy) ** 2)
        c5 = ~c1 & ~c2 & ~c3 & ~c4 & (np.abs(d - r) < circle_tol)
        out[c5] = circle_id
    return out


def channel_bcs(xy, x0, x1, y0, y1, umax, circle=None, circle_tol=1e-3):
    """Dirichlet lists in the if/else order of cases/Cylinder2D/main.cu:675-690: inflow profile at
    x0, zero pressure at x1 (outlet corners get the pressure condition only), no-slip on the walls and
    on the cylinder."""
    H = y1 - y0
    bu, bv0, bv1, bp = [], [], [], []
    for i in range(xy.shape[0]):
        x, y = xy[i]
        if abs(x - x0) < TOL_BC:
            yy = y - y0
            bu.append(i); bv0.append(4.0 * umax * yy * (H - yy) / (H * H)); bv1.append(0.0)
        elif abs(x - x1) < TOL_BC:
            bp.append(i)
        elif abs(y - y0) < TOL_BC or abs(y - y1) < TOL_BC:
            bu.append(i); bv0.append(0.0); bv1.append(0.0)
        elif circle is not None and abs(math.hypot(x - circle[0], y - circle[1]) - circle[2]) < circle_tol:
            bu.append(i); bv0.append(0.0); bv1.append(0.0)
    return (np.asarray(bu, np.int32), np.asarray(bv0, np.float64), np.asarray(bv1, np.float64),
            np.asarray(bp, np.int32), np.zeros(len(bp), np.float64))


@dataclass
class Pfem2Case:
    name: str
    xy: np.ndarray
    tri: np.ndarray
    edge_ids: np.ndarray
    bcu_node: np.ndarray
    bcu_val0: np.ndarray
    bcu_val1: np.ndarray
    bcp_node: np.ndarray
    bcp_val: np.ndarray
    rho: float
    mu: float
    dt: float
    scheme: int
    traction_mask: int
    normals: list
    outer_iterations: int
    substeps: int
    pred_solver: int
    pred_precond: int
    pres_solver: int
    pres_precond: int
    corr_solver: int
    corr_precond: int
    tol: float
    tol_mode: int
    maxit: int
    restart: int
    dirichlet_mode: int
    level: int
    seed_mode: int
    ppc: int
    seed: int
    body_id: int = -1
    body_center: tuple = (0.0, 0.0)
    force_coeff: float = 0.0
    classify: dict = field(default_factory=dict)
    steps: int = 1


# solver / preconditioner codes of include/pfem2.h
CG, GMRES = 0, 1
NONE, JACOBI, IC0, ILU0 = 0, 1, 2, 3
ABS, REL = 0, 1
PARITY, MASK = 0, 1


def poiseuille_case(nx=400, ny=100, jitter=0.2, mesh_seed=2, ppc=16, particle_seed=102, dt=1e-3, steps=200,
                    tol=1e-12, tol_mode=REL):
    """Velocity-driven Poiseuille flow on [0,4]x[0,1] (scheme A, one outer iteration)."""
    L, H = 4.0, 1.0
    xy, tri = meshgen.rectangle(0.0, L, 0.0, H, nx, ny, jitter=jitter, seed=mesh_seed)
    cls = dict(x0=0.0, x1=L, y0=0.0, y1=H, ids=(0, 1, 2, 3))
    ids = classify_edges(xy, tri, **cls)
    bu, bv0, bv1, bp, bpv = channel_bcs(xy, 0.0, L, 0.0, H, umax=1.0)
    return Pfem2Case(
        name=f"poiseuille_{nx}x{ny}", xy=xy, tri=tri, edge_ids=ids, bcu_node=bu, bcu_val0=bv0, bcu_val1=bv1,
        bcp_node=bp, bcp_val=bpv, rho=1.0, mu=0.01, dt=dt, scheme=0, traction_mask=0b1111,
        normals=[(-1.0, 0.0), (1.0, 0.0), (0.0, -1.0), (0.0, 1.0)], outer_iterations=1, substeps=3,
        pred_solver=GMRES, pred_precond=JACOBI, pres_solver=CG, pres_precond=JACOBI, corr_solver=GMRES,
        corr_precond=JACOBI, tol=tol, tol_mode=tol_mode, maxit=2000, restart=30, dirichlet_mode=MASK, level=4,
        seed_mode=1, ppc=ppc, seed=particle_seed, classify=cls, steps=steps)


def cylinder_case(h=1.0 / 112, ppc=20, particle_seed=103, dt=2e-4, steps=500, tol=1e-12, tol_mode=REL,
                  pres_precond=IC0, shuffle_seed=None, locality=False):
    """Flow past a cylinder (D = 1, centre (4, 2.1)) in the channel [0,20]x[0,4]; U_mean = 1 (U_max = 1.5),
    mu = 0.01 (Re = 100), scheme B, two outer iterations. shuffle_seed: number the nodes and order the
    cells at random (seeded) instead of the generator's structured order -- the same mesh with the
    numbering of an unordered unstructured mesh. locality: then renumber it with the library's locality
    order (pfem_mesh_locality_order: nodes in reverse Cuthill-McKee order, cells by their first node)."""
    xy, tri, info = meshgen.cylinder_channel(h=h)
    if shuffle_seed is not None:
        rng = np.random.default_rng(shuffle_seed)
        new_id = rng.permutation(xy.shape[0])
        xy2 = np.empty_like(xy)
        xy2[new_id] = xy
        xy, tri = xy2, new_id[tri][rng.permutation(tri.shape[0])].astype(np.int32)
    if locality:
        from . import capi
        xy, tri, _, _ = capi.renumber_mesh(xy, tri)
    L, H, R = info["L"], info["H"], info["R"]
    cx, cy = info["center"]
    cls = dict(x0=0.0, x1=L, y0=0.0, y1=H, ids=(0, 1, 2, 2), circle=(cx, cy, R), circle_tol=1e-3, circle_id=3)
    ids = classify_edges(xy, tri, **cls)
    bu, bv0, bv1, bp, bpv = channel_bcs(xy, 0.0, L, 0.0, H, umax=1.5, circle=(cx, cy, R))
    umean, rho = 1.0, 1.0
    return Pfem2Case(
        name=f"cylinder_T{tri.shape[0]}", xy=xy, tri=tri, edge_ids=ids, bcu_node=bu, bcu_val0=bv0, bcu_val1=bv1,
        bcp_node=bp, bcp_val=bpv, rho=rho, mu=0.01, dt=dt, scheme=1, traction_mask=0b10,
        normals=[(-1.0, 0.0), (1.0, 0.0), (0.0, 0.0), (0.0, 0.0)], outer_iterations=2, substeps=3,
        pred_solver=GMRES, pred_precond=JACOBI, pres_solver=CG, pres_precond=pres_precond, corr_solver=CG,
        corr_precond=JACOBI, tol=tol, tol_mode=tol_mode, maxit=5000, restart=30, dirichlet_mode=MASK, level=4,
        seed_mode=1, ppc=ppc, seed=particle_seed, body_id=3, body_center=(cx, cy),
        force_coeff=2.0 / (rho * umean * umean * info["D"]), classify=cls, steps=steps)


def poisson_case(n=64):
    """Poisson -Laplace(u) = 2 pi^2 sin(pi x) sin(pi y) on the unit square, u = 0 on the boundary."""
    xy, tri = meshgen.unit_square(n)
    on_b = (np.abs(xy[:, 0]) < TOL_BC) | (np.abs(xy[:, 0] - 1.0) < TOL_BC) | \
           (np.abs(xy[:, 1]) < TOL_BC) | (np.abs(xy[:, 1] - 1.0) < TOL_BC)
    nodes = np.nonzero(on_b)[0].astype(np.int32)
    return xy, tri, nodes, np.zeros(nodes.size)


def poisson3d_multi_case(n=128, nrhs=8, seed=5):
    """3D Poisson on the unit box: n^3 cubes, 6 Kuhn tetrahedra each, u = sin(pi x) sin(pi y) sin(pi z)
    (f = 3 pi^2 u), u = 0 on every face; nrhs right-hand sides, column j = s_j * (assembled RHS) plus
    N(0, 1e-3) noise on the interior nodes, s_j seeded in [0.5, 2]."""
    xyz, tet = meshgen.box3d(n)
    on_b = np.zeros(xyz.shape[0], bool)
    for d in range(3):
        on_b |= (np.abs(xyz[:, d]) < TOL_BC) | (np.abs(xyz[:, d] - 1.0) < TOL_BC)
    nodes = np.nonzero(on_b)[0].astype(np.int32)
    rng = np.random.default_rng(seed)
    scales = rng.uniform(0.5, 2.0, nrhs)
    noise = rng.standard_normal((nrhs, xyz.shape[0])) * 1e-3
    noise[:, on_b] = 0.0
    return dict(xyz=xyz, tet=tet, bnodes=nodes, scales=scales, noise=noise, n=n, nrhs=nrhs)


def build_poisson3d_multi(ctx, case, quad=0):
    """Device system of a poisson3d_multi_case: CSR matrix with the Dirichlet rows applied and the nrhs
    right-hand sides as one column-major device array. Returns (mesh, csr, B, base_rhs)."""
    from . import capi
    mesh = capi.Mesh3D(ctx, case["xyz"], case["tet"])
    a = capi.CSR(ctx, mesh=mesh)
    nv = case["xyz"].shape[0]
    rhs = ctx.zeros(nv)
    capi.assemble_poisson3d(mesh, a, rhs, coeff=1.0, rhs_kind=1, rhs_param=3 * math.pi ** 2, quad=quad)
    bc = capi.Dirichlet(ctx, nv, case["bnodes"], np.zeros(case["bnodes"].size))
    bc.apply(a, rhs, MASK)
    b0 = rhs.download()
    B = (case["scales"][:, None] * b0[None, :] + case["noise"]).ravel()
    return mesh, a, ctx.to_device(B), b0


def batched_poisson_case(nmesh=256, tmin=10_000, tmax=500_000, seed=11, jitter=0.2):
    """Independent pressure-Poisson systems: nmesh jittered structured triangulations of [0, 2] x [0, 1]
    with triangle counts drawn log-uniform in [tmin, tmax]; each has homogeneous Dirichlet nodes on its
    left edge and a right-hand side of i.i.d. N(0, 1) nodal values with the mean removed (zero on the
    Dirichlet nodes). The meshes are stored as one triangulation with disjoint vertex sets, so their
    matrix is block diagonal; sys_rows[k] is the first node of mesh k."""
    rng = np.random.default_rng(seed)
    target = np.exp(rng.uniform(np.log(tmin), np.log(tmax), nmesh))
    xys, tris, bnodes, rhs, rows = [], [], [], [], [0]
    for k in range(nmesh):
        nx = max(2, int(round(math.sqrt(target[k]))))
        ny = max(1, int(round(target[k] / (2 * nx))))
        xy, tri = meshgen.rectangle(0.0, 2.0, 0.0, 1.0, nx, ny, jitter=jitter, seed=seed * 1000 + k)
        off = rows[-1]
        left = np.nonzero(np.abs(xy[:, 0]) < TOL_BC)[0]
        b = rng.standard_normal(xy.shape[0])
        b -= b.mean()
        b[left] = 0.0
        xys.append(xy)
        tris.append(tri + off)
        bnodes.append(left + off)
        rhs.append(b)
        rows.append(off + xy.shape[0])
    return dict(xy=np.concatenate(xys), tri=np.concatenate(tris).astype(np.int32),
                bnodes=np.concatenate(bnodes).astype(np.int32), rhs=np.concatenate(rhs),
                sys_rows=np.asarray(rows, np.int32), ntri=np.array([t.shape[0] for t in tris]))


def build_batched_poisson(ctx, case):
    """Device system of a batched_poisson_case: the block-diagonal stiffness matrix with the Dirichlet
    rows applied and the right-hand side. Returns (mesh, csr, b)."""
    from . import capi
    mesh = capi.Mesh2D(ctx, case["xy"], case["tri"], fill_neighbors=False)
    a = capi.CSR(ctx, mesh=mesh)
    nv = case["xy"].shape[0]
    scratch = ctx.zeros(nv)
    capi.assemble_poisson2d(mesh, a, scratch, coeff=1.0, rhs_kind=2, rhs_param=0.0)
    capi.Dirichlet(ctx, nv, case["bnodes"], np.zeros(case["bnodes"].size)).apply(a, scratch, MASK)
    return mesh, a, ctx.to_device(case["rhs"])


def step_config(case, precision=0, inner_tol=1e-6, pres_ordering=1, pres_factor_bits=16, pred_gs_passes=0):
    """pfem_step_config (include/pfem2.h) for a case; in mixed precision a multicolour IC0 pressure factor
    is applied from a 16-bit copy by default."""
    from .capi import StepConfig, make_params
    cfg = StepConfig()
    cfg.prm = make_params(case.rho, case.mu, case.dt, case.scheme, case.traction_mask, case.normals)
    cfg.outer_iterations = case.outer_iterations
    cfg.substeps = case.substeps
    cfg.pred_solver, cfg.pred_precond = case.pred_solver, case.pred_precond
    cfg.pres_solver, cfg.pres_precond = case.pres_solver, case.pres_precond
    cfg.corr_solver, cfg.corr_precond = case.corr_solver, case.corr_precond
    cfg.tol, cfg.tol_mode, cfg.maxit, cfg.restart = case.tol, case.tol_mode, case.maxit, case.restart
    cfg.dirichlet_mode = case.dirichlet_mode
    cfg.body_id = case.body_id
    cfg.body_cx, cfg.body_cy = case.body_center
    cfg.force_coeff = case.force_coeff
    cfg.precision = precision
    cfg.inner_tol = inner_tol
    cfg.pres_ordering = pres_ordering
    cfg.pres_factor_bits = pres_factor_bits
    cfg.pred_gs_passes = pred_gs_passes
    return cfg


def build_gpu(ctx, case, precision=0, inner_tol=1e-6, pres_ordering=1, pred_gs_passes=0):
    """Mesh, particle handler (seeded, velocities initialised) and stepper on the GPU."""
    from . import capi
    mesh = capi.Mesh2D(ctx, case.xy, case.tri)
    mesh.set_edge_ids(case.edge_ids)
    ph = capi.Particles(mesh, case.level)
    ph.seed(case.seed_mode, case.ppc, case.seed)
    cfg = step_config(case, precision, inner_tol, pres_ordering, pred_gs_passes=pred_gs_passes)
    stepper = capi.Pfem2(mesh, ph, cfg, case.bcu_node, case.bcu_val0, case.bcu_val1, case.bcp_node, case.bcp_val)
    u0 = stepper.field_ptr("u0")
    u1 = stepper.field_ptr("u1")
    capi.lib().pfem_particles_init_velocity(ph.h, u0, u1)
    return mesh, ph, stepper
