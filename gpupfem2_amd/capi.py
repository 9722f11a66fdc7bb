"""ctypes binding of libpfem2_hip.so (the C ABI declared in include/pfem2.h).

The library is built in-tree by ``python -m gpupfem2_amd.build``. Loading fails loudly when it is
missing: there is no CPU fallback behind this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfem2_hip.so")

PFEM_OK = 0
PFEM_EINVAL, PFEM_EHIP, PFEM_EIO, PFEM_ENUMERIC, PFEM_ECAPACITY = -1, -2, -3, -4, -5
DIRICHLET_PARITY, DIRICHLET_MASK = 0, 1
PRECOND_NONE, PRECOND_JACOBI, PRECOND_IC0, PRECOND_ILU0 = 0, 1, 2, 3
ORDER_NATURAL, ORDER_MULTICOLOR, ORDER_BLOCKS = 0, 1, 2
SOLVER_CG, SOLVER_GMRES = 0, 1
TOL_ABSOLUTE, TOL_RELATIVE = 0, 1
PRECISION_FP64, PRECISION_MIXED, PRECISION_MIXED_PRECOND = 0, 1, 2
SOLVER_OPT_FACTOR_ORDER, SOLVER_OPT_POLL, SOLVER_OPT_FUSE, SOLVER_OPT_REPLACE, SOLVER_OPT_GRAPH = 1, 2, 3, 4, 5
SOLVER_OPT_GS_PASSES = 6


class SolveStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("final_residual", C.c_double),
                ("nan_detected", C.c_int), ("refinements", C.c_int)]


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
                ("body_cy", C.c_double), ("force_coeff", C.c_double), ("precision", C.c_int),
                ("inner_tol", C.c_double), ("pres_ordering", C.c_int), ("pres_factor_bits", C.c_int),
                ("pred_gs_passes", C.c_int)]


class SimulationParameters(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("tFinal", C.c_double), ("dt", C.c_double),
                ("tolerance", C.c_double), ("maxIterations", C.c_int), ("simulationScheme", C.c_int),
                ("meshFileName", C.c_char_p), ("outputFrequency", C.c_int),
                ("particleAdvectionSubsteps", C.c_int), ("exportParticles", C.c_int),
                ("calculateLoads", C.c_int), ("bodyBoundaryID", C.c_int), ("thickness", C.c_double),
                ("meanVelocity", C.c_double)]


_vp, _i, _d, _l, _sz = C.c_void_p, C.c_int, C.c_double, C.c_long, C.c_size_t
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every function of include/pfem2.h
SIGNATURES = {
    "pfem_ctx_create": (_i, [_i, _pp]),
    "pfem_ctx_destroy": (_i, [_vp]),
    "pfem_last_error": (C.c_char_p, [_vp]),
    "pfem_ctx_synchronize": (_i, [_vp]),
    "pfem_ctx_stream": (_vp, [_vp]),
    "pfem_malloc": (_i, [_vp, _sz, _pp]),
    "pfem_free": (_i, [_vp, _vp]),
    "pfem_memcpy_h2d": (_i, [_vp, _vp, _vp, _sz]),
    "pfem_memcpy_d2h": (_i, [_vp, _vp, _vp, _sz]),
    "pfem_memcpy_d2d": (_i, [_vp, _vp, _vp, _sz]),
    "pfem_memset_zero": (_i, [_vp, _vp, _sz]),
    "pfem_mesh2d_load_dat": (_i, [_vp, C.c_char_p, _i, _d, _pp]),
    "pfem_mesh2d_create": (_i, [_vp, _i, _vp, _i, _vp, _i, _pp]),
    "pfem_mesh2d_destroy": (_i, [_vp]),
    "pfem_mesh2d_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_l)]),
    "pfem_mesh2d_download": (_i, [_vp, _i, _vp]),
    "pfem_mesh2d_classify_edges": (_i, [_vp, _d, _d, _d, _d, _i, _i, _i, _i, _d, _d, _d, _d, _i]),
    "pfem_mesh2d_set_edge_ids": (_i, [_vp, _vp]),
    "pfem_mesh3d_load_dat": (_i, [_vp, C.c_char_p, _d, _i, _pp]),
    "pfem_mesh3d_create": (_i, [_vp, _i, _vp, _i, _vp, _i, _pp]),
    "pfem_mesh3d_destroy": (_i, [_vp]),
    "pfem_mesh3d_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "pfem_mesh3d_download": (_i, [_vp, _i, _vp]),
    "pfem_mesh3d_fill_neighbors": (_i, [_vp, C.POINTER(C.c_long)]),
    "pfem_mesh3d_device_ptr": (_i, [_vp, _i, C.POINTER(_vp)]),
    "pfem_csr_create_from_mesh2d": (_i, [_vp, _pp]),
    "pfem_csr_create_from_mesh3d": (_i, [_vp, _pp]),
    "pfem_csr_create": (_i, [_vp, _i, _i, _vp, _vp, _vp, _pp]),
    "pfem_csr_create_like": (_i, [_vp, _pp]),
    "pfem_csr_destroy": (_i, [_vp]),
    "pfem_csr_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "pfem_csr_pointers": (_i, [_vp, _pp, _pp, _pp]),
    "pfem_csr_clear_values": (_i, [_vp]),
    "pfem_csr_upload_values": (_i, [_vp, _vp]),
    "pfem_csr_download": (_i, [_vp, _vp, _vp, _vp]),
    "pfem_spmv": (_i, [_vp, _vp, _vp]),
    "pfem_csr_refresh_fp32": (_i, [_vp]),
    "pfem_spmv_fp32": (_i, [_vp, _vp, _vp]),
    "pfem_dot": (_i, [_vp, _i, _vp, _vp, _vp]),
    "pfem_nrm2": (_i, [_vp, _i, _vp, _vp]),
    "pfem_trsv_upper": (_i, [_vp, _i, _vp, _i, _vp]),
    "pfem_gemv": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "pfem_dirichlet_create": (_i, [_vp, _i, _i, _vp, _vp, _pp]),
    "pfem_dirichlet_destroy": (_i, [_vp]),
    "pfem_dirichlet_set_values": (_i, [_vp, _vp]),
    "pfem_dirichlet_values_device": (_i, [_vp, _pp]),
    "pfem_dirichlet_apply": (_i, [_vp, _vp, _vp, _i]),
    "pfem_dirichlet_apply2": (_i, [_vp, _vp, _vp, _vp, _vp, _i]),
    "pfem_precond_create": (_i, [_vp, _i, _i, _pp]),
    "pfem_precond_destroy": (_i, [_vp]),
    "pfem_precond_initialize": (_i, [_vp, _vp]),
    "pfem_precond_apply": (_i, [_vp, _vp, _vp]),
    "pfem_precond_download": (_i, [_vp, _vp, C.POINTER(_i), _vp]),
    "pfem_precond_set_sweep": (_i, [_vp, _i]),
    "pfem_precond_set_spin_limit": (_i, [_vp, _i]),
    "pfem_precond_set_factor_bits": (_i, [_vp, _i]),
    "pfem_solver_set_option": (_i, [_vp, _i, _i]),
    "pfem_precond_create_custom": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "pfem_mesh2d_device_ptr": (_i, [_vp, _i, _vp]),
    "pfem_solver_create": (_i, [_vp, _i, _d, _i, _i, _i, _vp, _pp]),
    "pfem_solver_set_precision": (_i, [_vp, _i, _d, _i]),
    "pfem_solver_destroy": (_i, [_vp]),
    "pfem_solver_solve": (_i, [_vp, _vp, _vp, _vp, C.POINTER(SolveStats)]),
    "pfem_solver_solve_multi": (_i, [_vp, _vp, _i, _vp, _vp, C.POINTER(SolveStats)]),
    "pfem_solver_solve_batched": (_i, [_vp, _vp, _i, _vp, _vp, _vp, C.POINTER(SolveStats), _vp]),
    "pfem_quadrature": (_i, [_i, _i, _vp]),
    "pfem_assemble_poisson2d": (_i, [_vp, _vp, _vp, _d, _i, _d, _i]),
    "pfem_assemble_poisson3d": (_i, [_vp, _vp, _vp, _d, _i, _d, _i]),
    "pfem_simulation_parameters_default": (None, [C.POINTER(SimulationParameters)]),
    "pfem_assemble_prediction": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(Pfem2Params)]),
    "pfem_assemble_pressure": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(Pfem2Params)]),
    "pfem_assemble_correction": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(Pfem2Params)]),
    "pfem_prediction_bc_values": (_i, [_vp, _vp, _vp, _vp, _vp, _i, C.POINTER(Pfem2Params)]),
    "pfem_body_forces": (_i, [_vp, _i, _d, _d, _vp, _vp, _vp, C.POINTER(Pfem2Params), _vp]),
    "pfem_body_forces_noalloc": (_i, [_vp, _i, _d, _d, _vp, _vp, _vp, C.POINTER(Pfem2Params), _vp]),
    "pfem_particles_create": (_i, [_vp, _i, _pp]),
    "pfem_particles_destroy": (_i, [_vp]),
    "pfem_particles_seed": (_i, [_vp, _i, _i, C.c_uint64]),
    "pfem_particles_init_velocity": (_i, [_vp, _vp, _vp]),
    "pfem_particles_advect": (_i, [_vp, _vp, _vp, _d, _i]),
    "pfem_particles_project": (_i, [_vp, _vp, _vp]),
    "pfem_particles_correct": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "pfem_particles_count": (_i, [_vp, C.POINTER(_l)]),
    "pfem_particles_download": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pfem_particles_upload": (_i, [_vp, _l, _vp, _vp, _vp, _vp, _vp, C.c_uint32]),
    "pfem_particles_stats": (_i, [_vp, C.POINTER(_l), C.POINTER(_l)]),
    "pfem_pfem2_create": (_i, [_vp, _vp, C.POINTER(StepConfig), _i, _vp, _vp, _vp, _i, _vp, _vp, _pp]),
    "pfem_pfem2_destroy": (_i, [_vp]),
    "pfem_pfem2_step": (_i, [_vp, _vp]),
    "pfem_pfem2_field": (_i, [_vp, _i, _pp]),
    "pfem_pfem2_last_iterations": (_i, [_vp, _vp, _i]),
    "pfem_pfem2_enable_timing": (_i, [_vp, _i]),
    "pfem_pfem2_set_graphs": (_i, [_vp, _i]),
    "pfem_pfem2_phase_times": (_i, [_vp, _vp, _i]),
    "pfem_pfem2_matrix": (_i, [_vp, _i, _pp]),
    "pfem_pfem2_precond": (_i, [_vp, _i, _pp]),
    "pfem_time_kernel": (_i, [_i, _vp, _vp, _i, C.POINTER(_d)]),
    "pfem_stream_copy": (_i, [_vp, C.c_size_t, _i, C.POINTER(_d), C.POINTER(_i)]),
    "pfem_mesh_locality_order": (_i, [_i, _i, _i, _vp, _vp, _vp]),
    "pfem_solver_solve_pair": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pfem_solver_set_batch_tiles": (_i, [_vp, _i, _vp, _vp, _vp]),
    "pfem_pfem2_sweep_fallbacks": (_i, [_vp]),
    "pfem_export_vtk": (_i, [_vp, C.c_char_p, _i, _vp, _vp]),
    "pfem_export_vtk3d": (_i, [_vp, C.c_char_p, _i, _vp, _vp]),
}

_LIB = None


class PfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pfem error {code}: {msg}")
        self.code = code


def lib():
    """Load libpfem2_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gpupfem2_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, int):
        return a
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """A device and its stream (pfem_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().pfem_ctx_create(device, C.byref(h))
        if rc != PFEM_OK:
            raise PfemError(rc, "context creation failed (is a GPU visible?)")
        self.h = h.value

    def check(self, rc):
        if rc != PFEM_OK:
            raise PfemError(rc, lib().pfem_last_error(self.h).decode())
        return rc

    def sync(self):
        self.check(lib().pfem_ctx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().pfem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- device memory
    def empty(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.size, a.dtype)
        d.upload(a)
        return d

    def zeros(self, n, dtype=np.float64):
        d = DeviceArray(self, n, dtype)
        d.zero()
        return d


class DeviceArray:
    """Device buffer owned by the context (deviceVector<T> counterpart)."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        ctx.check(lib().pfem_malloc(ctx.h, max(self.n, 1) * self.dtype.itemsize, C.byref(p)))
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        assert a.size == self.n, (a.size, self.n)
        self.ctx.check(lib().pfem_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        self.ctx.check(lib().pfem_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def zero(self):
        self.ctx.check(lib().pfem_memset_zero(self.ctx.h, self.ptr, self.nbytes))

    def copy_from(self, other):
        self.ctx.check(lib().pfem_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            lib().pfem_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Handle:
    _destroy = None

    def __del__(self):
        try:
            if getattr(self, "h", None) and self.ctx.h:
                getattr(lib(), self._destroy)(self.h)
                self.h = None
        except Exception:
            pass


class Mesh2D(_Handle):
    _destroy = "pfem_mesh2d_destroy"

    def __init__(self, ctx, xy=None, tri=None, path=None, fill_neighbors=True, scale=1.0):
        self.ctx = ctx
        h = C.c_void_p()
        if path is not None:
            ctx.check(lib().pfem_mesh2d_load_dat(ctx.h, path.encode(), int(fill_neighbors), scale, C.byref(h)))
        else:
            xy = np.ascontiguousarray(xy, np.float64)
            tri = np.ascontiguousarray(tri, np.int32)
            ctx.check(lib().pfem_mesh2d_create(ctx.h, xy.shape[0], _ptr(xy), tri.shape[0], _ptr(tri),
                                               int(fill_neighbors), C.byref(h)))
        self.h = h.value
        nv, nt, nth, nn = C.c_int(), C.c_int(), C.c_int(), C.c_long()
        lib().pfem_mesh2d_info(self.h, C.byref(nv), C.byref(nt), C.byref(nth), C.byref(nn))
        self.nv, self.nt, self.nt_host, self.n_neighbors = nv.value, nt.value, nth.value, nn.value

    def download(self, what):
        spec = {0: (2 * self.nv, np.float64), 1: (3 * self.nt, np.int32), 2: (self.nt, np.float64),
                3: (4 * self.nt, np.float64), 4: (self.nt + 1, np.int32), 5: (self.n_neighbors, np.int32),
                6: (3 * self.nt, np.int32), 7: (3 * self.nt_host, np.int32)}[what]
        out = np.empty(spec[0], spec[1])
        self.ctx.check(lib().pfem_mesh2d_download(self.h, what, _ptr(out)))
        return out

    def classify_edges(self, x0, x1, y0, y1, ids=(0, 1, 2, 3), circle=None, circle_tol=1e-3, circle_id=3):
        cx, cy, r = circle if circle is not None else (0.0, 0.0, 0.0)
        self.ctx.check(lib().pfem_mesh2d_classify_edges(self.h, x0, x1, y0, y1, ids[0], ids[1], ids[2], ids[3],
                                                        cx, cy, r, circle_tol, circle_id))

    def set_edge_ids(self, ids):
        ids = np.ascontiguousarray(ids, np.int32)
        self.ctx.check(lib().pfem_mesh2d_set_edge_ids(self.h, _ptr(ids)))


class Mesh3D(_Handle):
    _destroy = "pfem_mesh3d_destroy"

    def __init__(self, ctx, xyz=None, tet=None, path=None, scale=1.0, zero_small_det=False):
        self.ctx = ctx
        h = C.c_void_p()
        if path is not None:
            ctx.check(lib().pfem_mesh3d_load_dat(ctx.h, path.encode(), scale, int(zero_small_det), C.byref(h)))
        else:
            xyz = np.ascontiguousarray(xyz, np.float64)
            tet = np.ascontiguousarray(tet, np.int32)
            ctx.check(lib().pfem_mesh3d_create(ctx.h, xyz.shape[0], _ptr(xyz), tet.shape[0], _ptr(tet),
                                               int(zero_small_det), C.byref(h)))
        self.h = h.value
        nv, nt = C.c_int(), C.c_int()
        lib().pfem_mesh3d_info(self.h, C.byref(nv), C.byref(nt))
        self.nv, self.nt = nv.value, nt.value

    def fill_neighbors(self):
        """Face neighbour lists (Mesh3D::fillCellNeighborIndices): (offsets, indices)."""
        total = C.c_long()
        self.ctx.check(lib().pfem_mesh3d_fill_neighbors(self.h, C.byref(total)))
        self.n_neighbors = total.value
        return self.download(4), self.download(5)

    def download(self, what):
        spec = {0: (3 * self.nv, np.float64), 1: (4 * self.nt, np.int32), 2: (self.nt, np.float64),
                3: (9 * self.nt, np.float64), 4: (self.nt + 1, np.int32),
                5: (getattr(self, "n_neighbors", 0), np.int32)}[what]
        out = np.empty(spec[0], spec[1])
        self.ctx.check(lib().pfem_mesh3d_download(self.h, what, _ptr(out)))
        return out


class CSR(_Handle):
    """SparseMatrixCSR counterpart."""
    _destroy = "pfem_csr_destroy"

    def __init__(self, ctx, mesh=None, like=None, rowptr=None, col=None, val=None):
        self.ctx = ctx
        h = C.c_void_p()
        if mesh is not None:
            fn = lib().pfem_csr_create_from_mesh3d if isinstance(mesh, Mesh3D) else lib().pfem_csr_create_from_mesh2d
            ctx.check(fn(mesh.h, C.byref(h)))
        elif like is not None:
            ctx.check(lib().pfem_csr_create_like(like.h, C.byref(h)))
        else:
            rowptr = np.ascontiguousarray(rowptr, np.int32)
            col = np.ascontiguousarray(col, np.int32)
            v = None if val is None else np.ascontiguousarray(val, np.float64)
            ctx.check(lib().pfem_csr_create(ctx.h, rowptr.shape[0] - 1, col.shape[0], _ptr(rowptr), _ptr(col),
                                            _ptr(v), C.byref(h)))
        self.h = h.value
        n, nnz = C.c_int(), C.c_int()
        lib().pfem_csr_info(self.h, C.byref(n), C.byref(nnz))
        self.n, self.nnz = n.value, nnz.value

    def download(self):
        rp = np.empty(self.n + 1, np.int32)
        col = np.empty(self.nnz, np.int32)
        val = np.empty(self.nnz, np.float64)
        self.ctx.check(lib().pfem_csr_download(self.h, _ptr(rp), _ptr(col), _ptr(val)))
        return rp, col, val

    def upload_values(self, val):
        val = np.ascontiguousarray(val, np.float64)
        self.ctx.check(lib().pfem_csr_upload_values(self.h, _ptr(val)))

    def clear(self):
        self.ctx.check(lib().pfem_csr_clear_values(self.h))

    def spmv(self, x, y):
        self.ctx.check(lib().pfem_spmv(self.h, x.ptr, y.ptr))

    def refresh_fp32(self):
        self.ctx.check(lib().pfem_csr_refresh_fp32(self.h))

    def spmv_fp32(self, x, y):
        self.ctx.check(lib().pfem_spmv_fp32(self.h, x.ptr, y.ptr))


class Dirichlet(_Handle):
    _destroy = "pfem_dirichlet_destroy"

    def __init__(self, ctx, n, nodes, values):
        self.ctx = ctx
        nodes = np.ascontiguousarray(nodes, np.int32)
        values = np.ascontiguousarray(values, np.float64)
        h = C.c_void_p()
        ctx.check(lib().pfem_dirichlet_create(ctx.h, n, nodes.shape[0], _ptr(nodes), _ptr(values), C.byref(h)))
        self.h = h.value
        self.nbc = nodes.shape[0]

    def apply(self, a, rhs, mode=DIRICHLET_MASK):
        self.ctx.check(lib().pfem_dirichlet_apply(self.h, a.h, rhs.ptr, mode))

    def apply2(self, other, a, rhs0, rhs1, mode=DIRICHLET_MASK):
        self.ctx.check(lib().pfem_dirichlet_apply2(self.h, other.h, a.h, rhs0.ptr, rhs1.ptr, mode))

    def set_values(self, values):
        values = np.ascontiguousarray(values, np.float64)
        self.ctx.check(lib().pfem_dirichlet_set_values(self.h, _ptr(values)))

    def values(self):
        p = C.c_void_p()
        lib().pfem_dirichlet_values_device(self.h, C.byref(p))
        out = np.empty(self.nbc, np.float64)
        self.ctx.check(lib().pfem_memcpy_d2h(self.ctx.h, _ptr(out), p.value, 8 * self.nbc))
        return out


class Preconditioner(_Handle):
    _destroy = "pfem_precond_destroy"

    def __init__(self, a, kind=PRECOND_JACOBI, ordering=ORDER_NATURAL):
        self.ctx = a.ctx
        self.n = a.n
        self.nnz = a.nnz
        self.kind = kind
        h = C.c_void_p()
        self.ctx.check(lib().pfem_precond_create(a.h, kind, ordering, C.byref(h)))
        self.h = h.value

    def initialize(self, a):
        self.ctx.check(lib().pfem_precond_initialize(self.h, a.h))

    def apply(self, dest, src):
        self.ctx.check(lib().pfem_precond_apply(self.h, dest.ptr, src.ptr))

    def set_sweep(self, single_pass):
        """IC0/ILU0 triangular solves as one launch each (True) or one launch per level (False)."""
        self.ctx.check(lib().pfem_precond_set_sweep(self.h, 1 if single_pass else 0))

    def set_spin_limit(self, polls):
        """Polls of one single-pass wait before it is declared failed (PFEM_ENUMERIC)."""
        self.ctx.check(lib().pfem_precond_set_spin_limit(self.h, int(polls)))

    def set_factor_bits(self, bits):
        """16 or 32: storage of the factor copy PRECISION_MIXED_PRECOND applies (before set_precision)."""
        self.ctx.check(lib().pfem_precond_set_factor_bits(self.h, int(bits)))

    def download(self):
        nl = C.c_int()
        fac = np.empty(self.nnz if self.kind in (PRECOND_IC0, PRECOND_ILU0) else self.n, np.float64)
        perm = np.empty(self.n, np.int32)
        self.ctx.check(lib().pfem_precond_download(self.h, _ptr(fac), C.byref(nl), _ptr(perm)))
        return fac, nl.value, perm


class Solver(_Handle):
    _destroy = "pfem_solver_destroy"

    def __init__(self, a, kind=SOLVER_CG, tol=1e-8, maxit=1000, restart=0, tol_mode=TOL_ABSOLUTE, precond=None):
        self.ctx = a.ctx
        h = C.c_void_p()
        self.ctx.check(lib().pfem_solver_create(a.h, kind, tol, maxit, restart, tol_mode,
                                                precond.h if precond else None, C.byref(h)))
        self.h = h.value
        self.precond = precond

    def set_option(self, option, value):
        """SOLVER_OPT_FACTOR_ORDER (0/1), SOLVER_OPT_POLL (iterations between convergence polls) or
        SOLVER_OPT_FUSE (0/1/2/3, default 3), SOLVER_OPT_REPLACE, SOLVER_OPT_GRAPH (0/1, default 0) or SOLVER_OPT_GS_PASSES (GMRES: 1/2, default 2)."""
        self.ctx.check(lib().pfem_solver_set_option(self.h, option, int(value)))

    def set_precision(self, precision, inner_tol=0.0, max_refinements=0):
        self.ctx.check(lib().pfem_solver_set_precision(self.h, precision, inner_tol, max_refinements))

    def solve(self, a, x, b):
        st = SolveStats()
        self.ctx.check(lib().pfem_solver_solve(self.h, a.h, x.ptr, b.ptr, C.byref(st)))
        return st

    def solve_pair(self, a0, x0, b0, other, a1, x1, b1):
        """this solver on a0 x0 = b0 and `other` on a1 x1 = b1, together (pfem_solver_solve_pair)"""
        st = (SolveStats * 2)()
        self.ctx.check(lib().pfem_solver_solve_pair(self.h, a0.h, x0.ptr, b0.ptr, other.h, a1.h, x1.ptr, b1.ptr, st))
        return list(st)

    def set_batch_tiles(self, buckets):
        """buckets: [(max_rows, tile_rows, tile_nnz), ...] ascending in max_rows (pfem_solver_set_batch_tiles)"""
        arr = [np.ascontiguousarray([bk[i] for bk in buckets], np.int32) for i in range(3)]
        self.ctx.check(lib().pfem_solver_set_batch_tiles(self.h, len(buckets), _ptr(arr[0]), _ptr(arr[1]), _ptr(arr[2])))

    def solve_batched(self, a, sys_rows, x, b):
        """Independent systems of a block-diagonal matrix (system k: rows sys_rows[k]:sys_rows[k+1]);
        returns (stats list, per-system time to solution in ms)."""
        rows = np.ascontiguousarray(sys_rows, np.int32)
        nsys = rows.size - 1
        st = (SolveStats * nsys)()
        tms = np.zeros(nsys, np.float64)
        self.ctx.check(lib().pfem_solver_solve_batched(self.h, a.h, nsys, _ptr(rows), x.ptr, b.ptr, st, _ptr(tms)))
        return list(st), tms

    def solve_multi(self, a, X, B, nrhs):
        """A X = B for nrhs columns (device arrays of nrhs * n values, column j at offset j * n)."""
        st = (SolveStats * nrhs)()
        self.ctx.check(lib().pfem_solver_solve_multi(self.h, a.h, int(nrhs), X.ptr, B.ptr, st))
        return list(st)


class LinearAlgebra:
    """Dense BLAS-1/2 on the context's stream (LinearAlgebra, src/linear_algebra/linear_algebra.h:27-54).
    Operands are DeviceArrays or raw device addresses (ints); results stay in device memory, as with the
    reference's device pointer mode."""

    def __init__(self, ctx):
        self.ctx = ctx

    @staticmethod
    def _p(a):
        return a if isinstance(a, int) else a.ptr

    def dot(self, x, y, res, n=None):
        """res[0] = x . y (linear_algebra.h:27-38); n defaults to the arrays' common length."""
        if n is None:
            if x.n != y.n:
                raise PfemError(PFEM_EINVAL, "Inconsistent dimensions of vectors in dot product")
            n = x.n
        self.ctx.check(lib().pfem_dot(self.ctx.h, int(n), self._p(x), self._p(y), self._p(res)))

    def norm_squared(self, x, res, n=None):
        """res[0] = ||x||_2, the Euclidean norm itself (normSquared, linear_algebra.h:40-43, is cublasDnrm2)."""
        n = x.n if n is None else n
        self.ctx.check(lib().pfem_nrm2(self.ctx.h, int(n), self._p(x), self._p(res)))

    def solve_upper_triangular(self, a, x, n, lda):
        """x <- U^-1 x, U upper triangular in column-major a (solveGeneralTriangularSystem, :45-49)."""
        self.ctx.check(lib().pfem_trsv_upper(self.ctx.h, int(n), self._p(a), int(lda), self._p(x)))

    def general_mv(self, a, x, y, alpha, beta, m, n, lda):
        """y = alpha A x + beta y, A m x n column-major; alpha, beta device scalars (generalMV, :51-54)."""
        self.ctx.check(lib().pfem_gemv(self.ctx.h, int(m), int(n), self._p(alpha), self._p(a), int(lda),
                                       self._p(x), self._p(beta), self._p(y)))


def quadrature(kind, index):
    buf = np.zeros(13 * 5, np.float64)
    n = lib().pfem_quadrature(kind, index, _ptr(buf))
    if n < 0:
        raise PfemError(n, "unknown quadrature rule")
    stride = {1: 2, 2: 4, 3: 5}[kind]
    return buf[:n * stride].reshape(n, stride).copy()


def make_params(rho, mu, dt, scheme, traction_mask=0, normals=None, face_quad=1, edge_quad=1):
    prm = Pfem2Params()
    prm.rho, prm.mu, prm.dt, prm.scheme = rho, mu, dt, scheme
    prm.face_quad, prm.edge_quad, prm.traction_mask = face_quad, edge_quad, traction_mask
    for k in range(8):
        nx, ny = (normals[k] if normals is not None and k < len(normals) else (0.0, 0.0))
        prm.normal[k][0] = nx
        prm.normal[k][1] = ny
    return prm


def assemble_poisson2d(mesh, a, rhs, coeff=1.0, rhs_kind=1, rhs_param=0.0, quad=1):
    mesh.ctx.check(lib().pfem_assemble_poisson2d(mesh.h, a.h, rhs.ptr, coeff, rhs_kind, rhs_param, quad))


def assemble_poisson3d(mesh, a, rhs, coeff=1.0, rhs_kind=1, rhs_param=0.0, quad=0):
    mesh.ctx.check(lib().pfem_assemble_poisson3d(mesh.h, a.h, rhs.ptr, coeff, rhs_kind, rhs_param, quad))


def assemble_prediction(mesh, a0, a1, rhs0, rhs1, u0, u1, uo0, uo1, p_old, prm):
    mesh.ctx.check(lib().pfem_assemble_prediction(mesh.h, a0.h, a1.h, rhs0.ptr, rhs1.ptr, u0.ptr, u1.ptr, uo0.ptr,
                                                  uo1.ptr, _ptr(p_old), C.byref(prm)))


def assemble_pressure(mesh, a, rhs, vp0, vp1, p_old, prm):
    mesh.ctx.check(lib().pfem_assemble_pressure(mesh.h, a.h, rhs.ptr, vp0.ptr, vp1.ptr, _ptr(p_old), C.byref(prm)))


def assemble_correction(mesh, a, rhs0, rhs1, vp0, vp1, p, p_old, prm):
    mesh.ctx.check(lib().pfem_assemble_correction(mesh.h, a.h, rhs0.ptr, rhs1.ptr, vp0.ptr, vp1.ptr, p.ptr,
                                                  _ptr(p_old), C.byref(prm)))


def prediction_bc_values(mesh, target, source, p, p_old, component, prm):
    mesh.ctx.check(lib().pfem_prediction_bc_values(mesh.h, target.h, source.h, p.ptr, _ptr(p_old), component,
                                                   C.byref(prm)))


def body_forces(mesh, body_id, center, u0, u1, p, prm):
    out = np.zeros(4, np.float64)
    mesh.ctx.check(lib().pfem_body_forces(mesh.h, body_id, center[0], center[1], u0.ptr, u1.ptr, p.ptr,
                                          C.byref(prm), _ptr(out)))
    return out


def body_forces_noalloc(mesh, body_id, center, u0, u1, p, prm):
    """pfem_body_forces without device allocation (one static scratch array per GPU)"""
    out = np.zeros(4, np.float64)
    mesh.ctx.check(lib().pfem_body_forces_noalloc(mesh.h, body_id, center[0], center[1], u0.ptr, u1.ptr, p.ptr,
                                                  C.byref(prm), _ptr(out)))
    return out


class Particles(_Handle):
    """ParticleHandler2D counterpart."""
    _destroy = "pfem_particles_destroy"

    def __init__(self, mesh, level):
        self.ctx = mesh.ctx
        self.mesh = mesh
        h = C.c_void_p()
        self.ctx.check(lib().pfem_particles_create(mesh.h, level, C.byref(h)))
        self.h = h.value

    def seed(self, mode, ppc, seed):
        self.ctx.check(lib().pfem_particles_seed(self.h, mode, ppc, seed))

    def init_velocity(self, u0, u1):
        self.ctx.check(lib().pfem_particles_init_velocity(self.h, u0.ptr, u1.ptr))

    def advect(self, u0, u1, dt, substeps):
        self.ctx.check(lib().pfem_particles_advect(self.h, u0.ptr, u1.ptr, dt, substeps))

    def project(self, u0, u1):
        self.ctx.check(lib().pfem_particles_project(self.h, u0.ptr, u1.ptr))

    def correct(self, u0, u1, uo0, uo1):
        self.ctx.check(lib().pfem_particles_correct(self.h, u0.ptr, u1.ptr, uo0.ptr, uo1.ptr))

    def count(self):
        n = C.c_long()
        self.ctx.check(lib().pfem_particles_count(self.h, C.byref(n)))
        return n.value

    def get(self):
        P = self.count()
        cell = np.empty(P, np.int32)
        lam = np.empty((P, 3), np.float64)
        pos = np.empty((P, 2), np.float64)
        vel = np.empty((P, 2), np.float64)
        pid = np.empty(P, np.uint32)
        self.ctx.check(lib().pfem_particles_download(self.h, _ptr(cell), _ptr(lam), _ptr(pos), _ptr(vel), _ptr(pid)))
        return dict(cell=cell, lam=lam, pos=pos, vel=vel, id=pid)

    def set(self, st, next_id):
        self.ctx.check(lib().pfem_particles_upload(
            self.h, st["cell"].shape[0], _ptr(np.ascontiguousarray(st["cell"], np.int32)),
            _ptr(np.ascontiguousarray(st["lam"], np.float64)), _ptr(np.ascontiguousarray(st["pos"], np.float64)),
            _ptr(np.ascontiguousarray(st["vel"], np.float64)), _ptr(np.ascontiguousarray(st["id"], np.uint32)),
            next_id))

    def stats(self):
        a, b = C.c_long(), C.c_long()
        lib().pfem_particles_stats(self.h, C.byref(a), C.byref(b))
        return a.value, b.value


class Pfem2(_Handle):
    """One PFEM-2 problem (fields, systems, solvers) stepping over a mesh and a particle handler."""
    _destroy = "pfem_pfem2_destroy"
    FIELDS = {"u0": 0, "u1": 1, "p": 2, "vp0": 3, "vp1": 4, "uo0": 5, "uo1": 6, "pold": 7}

    def __init__(self, mesh, particles, cfg, bcu_node, bcu_val0, bcu_val1, bcp_node, bcp_val):
        self.ctx = mesh.ctx
        self.mesh = mesh
        self.particles = particles
        self.cfg = cfg
        arrs = [np.ascontiguousarray(bcu_node, np.int32), np.ascontiguousarray(bcu_val0, np.float64),
                np.ascontiguousarray(bcu_val1, np.float64), np.ascontiguousarray(bcp_node, np.int32),
                np.ascontiguousarray(bcp_val, np.float64)]
        h = C.c_void_p()
        self.ctx.check(lib().pfem_pfem2_create(mesh.h, particles.h, C.byref(cfg), arrs[0].shape[0], _ptr(arrs[0]),
                                               _ptr(arrs[1]), _ptr(arrs[2]), arrs[3].shape[0], _ptr(arrs[3]),
                                               _ptr(arrs[4]), C.byref(h)))
        self.h = h.value

    def step(self):
        f = np.zeros(2, np.float64)
        self.ctx.check(lib().pfem_pfem2_step(self.h, _ptr(f)))
        return f

    def field_ptr(self, name):
        p = C.c_void_p()
        self.ctx.check(lib().pfem_pfem2_field(self.h, self.FIELDS[name], C.byref(p)))
        return p.value

    def field(self, name):
        out = np.empty(self.mesh.nv, np.float64)
        self.ctx.check(lib().pfem_memcpy_d2h(self.ctx.h, _ptr(out), self.field_ptr(name), out.nbytes))
        return out

    def set_field(self, name, values):
        v = np.ascontiguousarray(values, np.float64)
        self.ctx.check(lib().pfem_memcpy_h2d(self.ctx.h, self.field_ptr(name), _ptr(v), v.nbytes))

    def fields(self):
        return {k: self.field(k) for k in ("u0", "u1", "p", "vp0", "vp1")}

    def iterations(self):
        its = np.zeros(64, np.int32)
        k = lib().pfem_pfem2_last_iterations(self.h, _ptr(its), 64)
        return its[:max(k, 0)]

    def enable_timing(self, on=True):
        lib().pfem_pfem2_enable_timing(self.h, int(on))

    def set_graphs(self, on=True):
        """replay the work between the solves from HIP graphs from the next step on (pfem_pfem2_set_graphs)"""
        self.ctx.check(lib().pfem_pfem2_set_graphs(self.h, int(on)))

    def phase_times(self):
        ms = np.zeros(8, np.float64)
        lib().pfem_pfem2_phase_times(self.h, _ptr(ms), 8)
        return ms

    def sweep_fallbacks(self):
        """solves re-run with per-level sweeps after a single-pass wait exceeded its spin limit"""
        return lib().pfem_pfem2_sweep_fallbacks(self.h)

    def matrix_handle(self, which):
        """raw pfem_csr of the stepper: 0 prediction x, 1 prediction y, 2 pressure, 3 correction"""
        p = C.c_void_p()
        self.ctx.check(lib().pfem_pfem2_matrix(self.h, which, C.byref(p)))
        return p.value

    def precond_handle(self, which):
        p = C.c_void_p()
        self.ctx.check(lib().pfem_pfem2_precond(self.h, which, C.byref(p)))
        return p.value


def time_kernel(ctx, op, csr_handle, precond_handle=None, launches=50):
    """average ms per launch (HIP events on the context stream): 0 SpMV FP64, 1 SpMV FP32,
    2 preconditioner apply FP64, 3 preconditioner apply FP32, 4 apply with the FP32-stored factor,
    5 / 6 ops 2 / 4 in the factor's ordering, 7 SpMV FP64 in the factor's ordering, 8 / 9 ops 6 / 5 with
    the tagged sweeps (PCG fusion level 3), 10 / 11 one fused PCG iteration with the FP64 / reduced-precision
    factor (solves of 2L and L iterations differenced)"""
    ms = C.c_double()
    ctx.check(lib().pfem_time_kernel(op, csr_handle, precond_handle, launches, C.byref(ms)))
    return ms.value


def locality_order(nv, cells):
    """(node_order, cell_order) for memory locality (pfem_mesh_locality_order, host only): node_order[new] =
    old in reverse Cuthill-McKee order, cell_order[new] = old by the cells' smallest new vertex index."""
    c = np.ascontiguousarray(cells, np.int32)
    nodes = np.empty(nv, np.int32)
    order = np.empty(max(c.shape[0], 1), np.int32)
    rc = lib().pfem_mesh_locality_order(int(nv), int(c.shape[0]), int(c.shape[1]), _ptr(c), _ptr(nodes), _ptr(order))
    if rc != 0:
        raise PfemError(rc, "pfem_mesh_locality_order failed")
    return nodes, order[:c.shape[0]]


def renumber_mesh(xy, cells):
    """The mesh in locality order: (xy, cells, node_order, cell_order) with xy_new = xy[node_order] and the
    cells' vertices mapped to the new node numbers, the cells listed in cell_order."""
    nodes, order = locality_order(xy.shape[0], cells)
    new_id = np.empty_like(nodes)
    new_id[nodes] = np.arange(nodes.size, dtype=nodes.dtype)
    return xy[nodes], new_id[np.asarray(cells)[order]].astype(np.int32), nodes, order


def stream_copy(ctx, nbytes, launches=10):
    """(average ms of one copy of nbytes, best grid): the streaming ceiling of pfem_stream_copy"""
    ms, g = C.c_double(), C.c_int()
    ctx.check(lib().pfem_stream_copy(ctx.h, int(nbytes), launches, C.byref(ms), C.byref(g)))
    return ms.value, g.value


def export_vtk(mesh, path, fields):
    names = (C.c_char_p * len(fields))(*[k.encode() for k in fields])
    ptrs = (C.c_void_p * len(fields))(*[v.ptr for v in fields.values()])
    mesh.ctx.check(lib().pfem_export_vtk(mesh.h, path.encode(), len(fields), names, ptrs))


def export_vtk3d(mesh, path, fields):
    names = (C.c_char_p * len(fields))(*[k.encode() for k in fields])
    ptrs = (C.c_void_p * len(fields))(*[v.ptr for v in fields.values()])
    mesh.ctx.check(lib().pfem_export_vtk3d(mesh.h, path.encode(), len(fields), names, ptrs))
