"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module. It loads
oracle/liboracle.so (built by oracle/Makefile) and wraps the C functions with numpy arrays.

``load_dat`` restates the reference DAT reader (src/mesh_2d.cu:36-96, src/mesh_3d.cu:84-147) in
Python, including its end-of-file quirk: a trailing newline makes the reader count one more skipped
element, so the device cell count is one short while the host cell list keeps every triangle.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_BUILD = "liboracle.so"

D = np.float64
I = np.int32


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, _BUILD)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle/{_BUILD} is not built (run `make -C oracle`)")
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def use_openmp_build():
    """Load oracle/liboracle_omp.so in this process instead of the parity build: the same sources with
    -O3 -fopenmp (per-thread-chunk reductions, atomic scatters), for bench.py's CPU baseline only. Must be
    called before the first lib()."""
    global _BUILD
    if _LIB is not None and _BUILD != "liboracle_omp.so":
        raise RuntimeError("the parity oracle is already loaded in this process")
    _BUILD = "liboracle_omp.so"


def _p(a):
    """Pointer to an array's data for the C oracle. A strided input (a column slice, say) is copied to
    contiguous memory first; the pointer keeps the copy alive. Outputs are always allocated contiguous."""
    if a is None:
        return None
    if not a.flags.c_contiguous:
        a = np.ascontiguousarray(a)
    return a.ctypes.data_as(C.c_void_p)


class SolveStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("final_residual", C.c_double),
                ("nan_detected", C.c_int)]


class Pfem2Params(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("dt", C.c_double), ("scheme", C.c_int),
                ("face_quad", C.c_int), ("edge_quad", C.c_int), ("traction_mask", C.c_int),
                ("normal", (C.c_double * 2) * 8)]


class StepConfig(C.Structure):
    _fields_ = [("prm", Pfem2Params), ("outer_iterations", C.c_int), ("substeps", C.c_int),
                ("pred_solver", C.c_int), ("pred_precond", C.c_int), ("pres_solver", C.c_int),
                ("pres_precond", C.c_int), ("corr_solver", C.c_int), ("corr_precond", C.c_int),
                ("tol", C.c_double), ("tol_mode", C.c_int), ("maxit", C.c_int), ("restart", C.c_int),
                ("dirichlet_mode", C.c_int), ("body_id", C.c_int), ("body_cx", C.c_double),
                ("body_cy", C.c_double), ("force_coeff", C.c_double)]


def _declare(L):
    vp, i, d, l, u64 = C.c_void_p, C.c_int, C.c_double, C.c_long, C.c_uint64
    sig = {
        "orc_geometry2d": (None, [i, vp, i, vp, vp, vp]),
        "orc_pcg_coloured": (i, [i, vp, vp, vp, vp, vp, d, i, i, i, vp, vp]),
        "orc_pfem2_set_coloured_pressure": (None, [vp, i]),
        "orc_geometry3d": (None, [i, vp, i, vp, vp, vp]),
        "orc_csr_count": (l, [i, i, i, vp]),
        "orc_csr_fill": (None, [i, i, i, vp, vp, vp]),
        "orc_neighbors_bruteforce_count": (l, [i, vp]),
        "orc_neighbors_bruteforce_fill": (None, [i, vp, vp, vp]),
        "orc_face_neighbors3d_count": (l, [i, vp]),
        "orc_face_neighbors3d_fill": (None, [i, vp, vp, vp]),
        "orc_quadrature": (i, [i, i, vp]),
        "orc_spmv": (None, [i, vp, vp, vp, vp, vp]),
        "orc_dirichlet_apply": (None, [i, vp, vp, vp, vp, i, vp, vp, i]),
        "orc_jacobi_setup": (None, [i, vp, vp, vp, vp]),
        "orc_ic0": (i, [i, vp, vp, vp, vp]),
        "orc_ilu0": (i, [i, vp, vp, vp, vp]),
        "orc_ic0_apply": (None, [i, vp, vp, vp, vp, vp]),
        "orc_ilu0_apply": (None, [i, vp, vp, vp, vp, vp]),
        "orc_pcg": (i, [i, vp, vp, vp, vp, vp, d, i, i, i, C.POINTER(SolveStats)]),
        "orc_gmres": (i, [i, vp, vp, vp, vp, vp, d, i, i, i, i, C.POINTER(SolveStats)]),
        "orc_assemble_poisson2d": (None, [i, vp, i, vp, vp, vp, vp, vp, vp, vp, d, i, d, i]),
        "orc_assemble_poisson3d": (None, [i, vp, i, vp, vp, vp, vp, vp, vp, vp, d, i, d, i]),
        "orc_assemble_prediction": (None, [i, vp, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                           vp, vp, vp, vp, vp, C.POINTER(Pfem2Params)]),
        "orc_assemble_pressure": (None, [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(Pfem2Params)]),
        "orc_assemble_correction": (None, [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                           C.POINTER(Pfem2Params)]),
        "orc_prediction_bc_values": (None, [i, i, vp, vp, vp, vp, vp, i, vp, vp, vp, i, C.POINTER(Pfem2Params)]),
        "orc_body_edges": (i, [i, vp, i, vp]),
        "orc_body_forces": (None, [vp, vp, vp, i, vp, vp, vp, vp, d, d, C.POINTER(Pfem2Params), vp]),
        "orc_ph_create": (vp, [i, vp, i, vp, vp, vp, vp, i]),
        "orc_ph_destroy": (None, [vp]),
        "orc_ph_seed": (None, [vp, i, i, u64]),
        "orc_ph_init_velocity": (None, [vp, vp, vp]),
        "orc_ph_advect": (None, [vp, vp, vp, d, i]),
        "orc_ph_project": (None, [vp, vp, vp]),
        "orc_ph_correct": (None, [vp, vp, vp, vp, vp]),
        "orc_ph_count": (l, [vp]),
        "orc_ph_get": (None, [vp, vp, vp, vp, vp, vp]),
        "orc_ph_set": (None, [vp, l, vp, vp, vp, vp, vp, C.c_uint32]),
        "orc_ph_stats": (None, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "orc_subcell_centers": (None, [i, vp]),
        "orc_determine_subcell": (i, [i, d, d, d]),
        "orc_philox4x32": (None, [vp, vp, vp]),
        "orc_pfem2_create": (vp, [i, vp, i, vp, vp, C.POINTER(StepConfig), vp, i, vp, vp, vp, i, vp, vp]),
        "orc_pfem2_destroy": (None, [vp]),
        "orc_pfem2_step": (None, [vp, vp]),
        "orc_pfem2_fields": (None, [vp, vp, vp, vp, vp, vp]),
        "orc_pfem2_set_fields": (None, [vp, vp, vp, vp, vp, vp]),
        "orc_pfem2_last_iterations": (None, [vp, vp]),
        "orc_threads": (i, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


# ----------------------------------------------------------------------------- mesh

def load_dat(path, dim=2, scale=1.0):
    """Reference DAT reader semantics. Returns (vertices, host_cells, device_cell_count)."""
    with open(path, "r") as f:
        text = f.read()
    tok = text.split()
    nv, ncells = int(tok[0]), int(tok[1])
    pos = 2
    verts = np.empty((nv, dim), dtype=D)
    for k in range(nv):
        vals = tok[pos + 1:pos + 4]
        verts[k, 0] = float(vals[0])
        verts[k, 1] = float(vals[1])
        if dim == 3:
            verts[k, 2] = float(vals[2])
        pos += 4
    verts *= scale
    cells = []
    want = 203 if dim == 2 else 304
    nper = 3 if dim == 2 else 4
    while pos < len(tok):
        typ = int(tok[pos + 1])
        pos += 2
        if typ == want:
            cells.append([int(t) - 1 for t in tok[pos:pos + nper]])
            pos += nper
        else:
            ncells -= 1
            pos += 2 if typ == 102 else (3 if typ == 203 else 2)
    # istream >> at EOF: when the file ends with whitespace the loop runs once more on a failed read
    if text and text[-1].isspace():
        ncells -= 1
    return verts, np.asarray(cells, dtype=I).reshape(-1, nper), ncells


def geometry2d(xy, tri):
    nt = tri.shape[0]
    area = np.empty(nt, D)
    invJ = np.empty((nt, 4), D)
    lib().orc_geometry2d(xy.shape[0], _p(np.ascontiguousarray(xy, D)), nt, _p(np.ascontiguousarray(tri, I)),
                         _p(area), _p(invJ))
    return area, invJ


def geometry3d(xyz, tet):
    nt = tet.shape[0]
    vol = np.empty(nt, D)
    invJ = np.empty((nt, 9), D)
    lib().orc_geometry3d(xyz.shape[0], _p(np.ascontiguousarray(xyz, D)), nt, _p(np.ascontiguousarray(tet, I)),
                         _p(vol), _p(invJ))
    return vol, invJ


def csr_pattern(nv, cells):
    cells = np.ascontiguousarray(cells, I)
    nt, nper = cells.shape
    nnz = lib().orc_csr_count(nv, nt, nper, _p(cells))
    rowptr = np.empty(nv + 1, I)
    col = np.empty(nnz, I)
    lib().orc_csr_fill(nv, nt, nper, _p(cells), _p(rowptr), _p(col))
    return rowptr, col


def neighbors_bruteforce(tri):
    tri = np.ascontiguousarray(tri, I)
    nt = tri.shape[0]
    total = lib().orc_neighbors_bruteforce_count(nt, _p(tri))
    off = np.empty(nt + 1, I)
    idx = np.empty(total, I)
    lib().orc_neighbors_bruteforce_fill(nt, _p(tri), _p(off), _p(idx))
    return off, idx


def face_neighbors3d(tet):
    """Face neighbours of tetrahedra by the all-pairs test of src/mesh_3d.cu:37-82 (small meshes only)."""
    tet = np.ascontiguousarray(tet, I)
    nt = tet.shape[0]
    total = lib().orc_face_neighbors3d_count(nt, _p(tet))
    off = np.empty(nt + 1, I)
    idx = np.empty(total, I)
    lib().orc_face_neighbors3d_fill(nt, _p(tet), _p(off), _p(idx))
    return off, idx


def quadrature(kind, index):
    buf = np.zeros(31 * 5, D)
    n = lib().orc_quadrature(kind, index, _p(buf))
    stride = {1: 2, 2: 4, 3: 5}[kind]
    return buf[:n * stride].reshape(n, stride).copy()


# ----------------------------------------------------------------------------- linear algebra

def spmv(rowptr, col, val, x):
    n = rowptr.shape[0] - 1
    y = np.empty(n, D)
    lib().orc_spmv(n, _p(rowptr), _p(col), _p(val), _p(np.ascontiguousarray(x, D)), _p(y))
    return y


def dirichlet_apply(rowptr, col, val, rhs, bc_node, bc_val, mode):
    """In place on val and rhs (float64 arrays)."""
    n = rowptr.shape[0] - 1
    bn = np.ascontiguousarray(bc_node, I)
    bv = np.ascontiguousarray(bc_val, D)
    lib().orc_dirichlet_apply(n, _p(rowptr), _p(col), _p(val), _p(rhs), bn.shape[0], _p(bn), _p(bv), mode)


def jacobi(rowptr, col, val):
    n = rowptr.shape[0] - 1
    d = np.zeros(n, D)
    lib().orc_jacobi_setup(n, _p(rowptr), _p(col), _p(val), _p(d))
    return d


def ic0(rowptr, col, val):
    n = rowptr.shape[0] - 1
    out = np.zeros_like(val)
    rc = lib().orc_ic0(n, _p(rowptr), _p(col), _p(val), _p(out))
    return rc, out


def ilu0(rowptr, col, val):
    n = rowptr.shape[0] - 1
    out = np.zeros_like(val)
    rc = lib().orc_ilu0(n, _p(rowptr), _p(col), _p(val), _p(out))
    return rc, out


def ic0_apply(rowptr, col, lval, r):
    z = np.empty(rowptr.shape[0] - 1, D)
    lib().orc_ic0_apply(z.shape[0], _p(rowptr), _p(col), _p(lval), _p(np.ascontiguousarray(r, D)), _p(z))
    return z


def ilu0_apply(rowptr, col, luval, r):
    z = np.empty(rowptr.shape[0] - 1, D)
    lib().orc_ilu0_apply(z.shape[0], _p(rowptr), _p(col), _p(luval), _p(np.ascontiguousarray(r, D)), _p(z))
    return z


PRECOND = {"none": 0, "jacobi": 1, "ic0": 2, "ilu0": 3}


def pcg(rowptr, col, val, x, b, tol, maxit, tol_mode=0, precond="jacobi"):
    """x is updated in place; returns SolveStats."""
    st = SolveStats()
    lib().orc_pcg(rowptr.shape[0] - 1, _p(rowptr), _p(col), _p(val), _p(x), _p(np.ascontiguousarray(b, D)),
                  tol, maxit, tol_mode, PRECOND[precond], C.byref(st))
    return st


def gmres(rowptr, col, val, x, b, tol, maxit, restart, tol_mode=0, precond="jacobi"):
    st = SolveStats()
    lib().orc_gmres(rowptr.shape[0] - 1, _p(rowptr), _p(col), _p(val), _p(x), _p(np.ascontiguousarray(b, D)),
                    tol, maxit, restart, tol_mode, PRECOND[precond], C.byref(st))
    return st


# ----------------------------------------------------------------------------- assembly

def assemble_poisson2d(xy, tri, area, invJ, rowptr, col, coeff=1.0, rhs_kind=1, rhs_param=0.0, quad=1):
    nv = xy.shape[0]
    val = np.zeros(col.shape[0], D)
    rhs = np.zeros(nv, D)
    lib().orc_assemble_poisson2d(nv, _p(xy), tri.shape[0], _p(tri), _p(area), _p(invJ), _p(rowptr), _p(col),
                                 _p(val), _p(rhs), coeff, rhs_kind, rhs_param, quad)
    return val, rhs


def assemble_poisson3d(xyz, tet, vol, invJ, rowptr, col, coeff=1.0, rhs_kind=1, rhs_param=0.0, quad=0):
    nv = xyz.shape[0]
    val = np.zeros(col.shape[0], D)
    rhs = np.zeros(nv, D)
    lib().orc_assemble_poisson3d(nv, _p(xyz), tet.shape[0], _p(tet), _p(vol), _p(invJ), _p(rowptr), _p(col),
                                 _p(val), _p(rhs), coeff, rhs_kind, rhs_param, quad)
    return val, rhs


def make_params(rho, mu, dt, scheme, traction_mask=0, normals=None, face_quad=1, edge_quad=1):
    prm = Pfem2Params()
    prm.rho, prm.mu, prm.dt, prm.scheme = rho, mu, dt, scheme
    prm.face_quad, prm.edge_quad, prm.traction_mask = face_quad, edge_quad, traction_mask
    for k in range(8):
        nx, ny = (normals[k] if normals is not None and k < len(normals) else (0.0, 0.0))
        prm.normal[k][0] = nx
        prm.normal[k][1] = ny
    return prm


def assemble_prediction(xy, tri, area, invJ, edge_ids, rowptr, col, u0, u1, uo0, uo1, p_old, prm):
    nv, nnz = xy.shape[0], col.shape[0]
    v0, v1 = np.zeros(nnz, D), np.zeros(nnz, D)
    r0, r1 = np.zeros(nv, D), np.zeros(nv, D)
    lib().orc_assemble_prediction(nv, _p(xy), tri.shape[0], _p(tri), _p(area), _p(invJ), _p(edge_ids),
                                  _p(rowptr), _p(col), _p(v0), _p(v1), _p(r0), _p(r1), _p(u0), _p(u1),
                                  _p(uo0), _p(uo1), _p(p_old), C.byref(prm))
    return v0, v1, r0, r1


def assemble_pressure(tri, area, invJ, rowptr, col, vp0, vp1, p_old, prm):
    nv, nnz = rowptr.shape[0] - 1, col.shape[0]
    val, rhs = np.zeros(nnz, D), np.zeros(nv, D)
    lib().orc_assemble_pressure(nv, tri.shape[0], _p(tri), _p(area), _p(invJ), _p(rowptr), _p(col), _p(val),
                                _p(rhs), _p(vp0), _p(vp1), _p(p_old), C.byref(prm))
    return val, rhs


def assemble_correction(tri, area, invJ, rowptr, col, vp0, vp1, p, p_old, prm):
    nv, nnz = rowptr.shape[0] - 1, col.shape[0]
    val, r0, r1 = np.zeros(nnz, D), np.zeros(nv, D), np.zeros(nv, D)
    lib().orc_assemble_correction(nv, tri.shape[0], _p(tri), _p(area), _p(invJ), _p(rowptr), _p(col), _p(val),
                                  _p(r0), _p(r1), _p(vp0), _p(vp1), _p(p), _p(p_old), C.byref(prm))
    return val, r0, r1


def prediction_bc_values(nv, tri, area, invJ, p, p_old, bc_node, bc_src, component, prm):
    bn = np.ascontiguousarray(bc_node, I)
    out = np.empty(bn.shape[0], D)
    lib().orc_prediction_bc_values(nv, tri.shape[0], _p(tri), _p(area), _p(invJ), _p(p), _p(p_old), bn.shape[0],
                                   _p(bn), _p(np.ascontiguousarray(bc_src, D)), _p(out), component, C.byref(prm))
    return out


def body_edges(edge_ids, body_id):
    e = np.ascontiguousarray(edge_ids, I)
    n = lib().orc_body_edges(e.shape[0], _p(e), body_id, None)
    out = np.empty((n, 2), I)
    lib().orc_body_edges(e.shape[0], _p(e), body_id, _p(out))
    return out


def body_forces(xy, tri, invJ, cell_edge, u0, u1, p, center, prm):
    out = np.zeros(4, D)
    ce = np.ascontiguousarray(cell_edge, I)
    lib().orc_body_forces(_p(xy), _p(tri), _p(invJ), ce.shape[0], _p(ce), _p(u0), _p(u1), _p(p),
                          center[0], center[1], C.byref(prm), _p(out))
    return out


# ----------------------------------------------------------------------------- particles

def subcell_centers(level):
    out = np.zeros(3 * level * level, D)
    lib().orc_subcell_centers(level, _p(out))
    return out.reshape(-1, 3)


def determine_subcell(level, l1, l2, l3):
    return lib().orc_determine_subcell(level, l1, l2, l3)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32(_p(c), _p(k), _p(out))
    return out


class Particles:
    """Oracle particle handler over one 2D mesh (owns its state)."""

    def __init__(self, xy, tri, invJ, nbr_off, nbr_idx, level):
        self._keep = [np.ascontiguousarray(a) for a in (xy, tri, invJ, nbr_off, nbr_idx)]
        xy, tri, invJ, nbr_off, nbr_idx = self._keep
        self.h = lib().orc_ph_create(xy.shape[0], _p(xy), tri.shape[0], _p(tri), _p(invJ), _p(nbr_off),
                                     _p(nbr_idx), level)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ph_destroy(self.h)
            self.h = None

    def seed(self, mode, ppc, seed):
        lib().orc_ph_seed(self.h, mode, ppc, seed)

    def init_velocity(self, u0, u1):
        lib().orc_ph_init_velocity(self.h, _p(u0), _p(u1))

    def advect(self, u0, u1, dt, substeps):
        lib().orc_ph_advect(self.h, _p(u0), _p(u1), dt, substeps)

    def project(self, nv):
        u0, u1 = np.empty(nv, D), np.empty(nv, D)
        lib().orc_ph_project(self.h, _p(u0), _p(u1))
        return u0, u1

    def correct(self, u0, u1, uo0, uo1):
        lib().orc_ph_correct(self.h, _p(u0), _p(u1), _p(uo0), _p(uo1))

    def count(self):
        return lib().orc_ph_count(self.h)

    def get(self):
        P = self.count()
        cell = np.empty(P, I)
        lam = np.empty((P, 3), D)
        pos = np.empty((P, 2), D)
        vel = np.empty((P, 2), D)
        pid = np.empty(P, np.uint32)
        lib().orc_ph_get(self.h, _p(cell), _p(lam), _p(pos), _p(vel), _p(pid))
        return dict(cell=cell, lam=lam, pos=pos, vel=vel, id=pid)

    def set(self, st, next_id):
        lib().orc_ph_set(self.h, st["cell"].shape[0], _p(np.ascontiguousarray(st["cell"], I)),
                         _p(np.ascontiguousarray(st["lam"], D)), _p(np.ascontiguousarray(st["pos"], D)),
                         _p(np.ascontiguousarray(st["vel"], D)), _p(np.ascontiguousarray(st["id"], np.uint32)),
                         next_id)

    def stats(self):
        a, b = C.c_long(), C.c_long()
        lib().orc_ph_stats(self.h, C.byref(a), C.byref(b))
        return a.value, b.value


class Pfem2:
    """Oracle PFEM-2 time stepper. cfg is a StepConfig; particles an oracle Particles object."""

    def __init__(self, xy, tri, edge_ids, cfg, particles, bcu_node, bcu_val0, bcu_val1, bcp_node, bcp_val):
        self.nv = xy.shape[0]
        self._keep = [np.ascontiguousarray(xy, D), np.ascontiguousarray(tri, I),
                      np.ascontiguousarray(edge_ids, I), np.ascontiguousarray(bcu_node, I),
                      np.ascontiguousarray(bcu_val0, D), np.ascontiguousarray(bcu_val1, D),
                      np.ascontiguousarray(bcp_node, I), np.ascontiguousarray(bcp_val, D)]
        k = self._keep
        self.cfg = cfg
        self.particles = particles
        self.h = lib().orc_pfem2_create(self.nv, _p(k[0]), k[1].shape[0], _p(k[1]), _p(k[2]), C.byref(cfg),
                                        particles.h, k[3].shape[0], _p(k[3]), _p(k[4]), _p(k[5]),
                                        k[6].shape[0], _p(k[6]), _p(k[7]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pfem2_destroy(self.h)
            self.h = None

    def step(self):
        f = np.zeros(2, D)
        lib().orc_pfem2_step(self.h, _p(f))
        return f

    def fields(self):
        out = [np.empty(self.nv, D) for _ in range(5)]
        lib().orc_pfem2_fields(self.h, *[_p(a) for a in out])
        return dict(zip(["u0", "u1", "p", "vp0", "vp1"], out))

    def set_coloured_pressure(self, on=True):
        lib().orc_pfem2_set_coloured_pressure(self.h, int(on))

    def set_fields(self, u0=None, u1=None, p=None, vp0=None, vp1=None):
        lib().orc_pfem2_set_fields(self.h, _p(u0), _p(u1), _p(p), _p(vp0), _p(vp1))

    def iterations(self):
        its = np.zeros(5 * self.cfg.outer_iterations, I)
        lib().orc_pfem2_last_iterations(self.h, _p(its))
        return its


def threads():
    return lib().orc_threads()
