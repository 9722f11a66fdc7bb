"""Shared test helpers: build the oracle's side of a case from the same data the GPU gets."""
import numpy as np


def oracle_case(O, case):
    area, invJ = O.geometry2d(case.xy, case.tri)
    off, idx = O.neighbors_bruteforce(case.tri) if case.tri.shape[0] <= 60000 else (None, None)
    return area, invJ, off, idx


def oracle_params(O, case):
    return O.make_params(case.rho, case.mu, case.dt, case.scheme, case.traction_mask, case.normals)


def oracle_stepper(O, case, gpu_mesh=None):
    """Oracle particles (seeded like the GPU) and stepper. Neighbour lists come from the GPU mesh for
    meshes too large for the brute-force builder (their equality is tested separately)."""
    area, invJ = O.geometry2d(case.xy, case.tri)
    if gpu_mesh is not None:
        off, idx = gpu_mesh.download(4), gpu_mesh.download(5)
    else:
        off, idx = O.neighbors_bruteforce(case.tri)
    ph = O.Particles(case.xy, case.tri, invJ, off, idx, case.level)
    ph.seed(case.seed_mode, case.ppc, case.seed)
    z = np.zeros(case.xy.shape[0])
    ph.init_velocity(z, z)
    cfg = O.StepConfig()
    cfg.prm = oracle_params(O, case)
    for k in ("outer_iterations", "substeps", "pred_solver", "pred_precond", "pres_solver", "pres_precond",
              "corr_solver", "corr_precond", "tol", "tol_mode", "maxit", "restart", "dirichlet_mode", "body_id",
              "force_coeff"):
        setattr(cfg, k, getattr(case, k))
    cfg.body_cx, cfg.body_cy = case.body_center
    st = O.Pfem2(case.xy, case.tri, case.edge_ids, cfg, ph, case.bcu_node, case.bcu_val0, case.bcu_val1,
                 case.bcp_node, case.bcp_val)
    return ph, st


def face_neighbors_by_hashing(tet):
    """Face neighbours found independently of the all-pairs test: the four sorted vertex triples of every
    tetrahedron, sorted together, pair up the two cells of each interior face."""
    T = tet.shape[0]
    faces = np.sort(tet[:, [[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]]], axis=2).reshape(-1, 3)
    owner = np.repeat(np.arange(T), 4)
    order = np.lexsort((faces[:, 2], faces[:, 1], faces[:, 0]))
    f, o = faces[order], owner[order]
    same = np.all(f[1:] == f[:-1], axis=1)
    assert not np.any(same[1:] & same[:-1]), "a face shared by three cells"
    a, b = o[:-1][same], o[1:][same]
    src, dst = np.concatenate([a, b]), np.concatenate([b, a])
    k = np.lexsort((dst, src))
    off = np.zeros(T + 1, np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), dst[k], int(same.sum())

