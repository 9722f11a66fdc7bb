"""Device memory is returned: steppers (with their captured step graphs, particle buffers and solvers) and
IC0 preconditioners and solvers of every ordering, created and destroyed over and over, leave the free
device memory where the first cycle left it (hipMemGetInfo, read through the HIP runtime the library
itself loads)."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_no_device_memory_leak(ctx):
    from gpupfem2_amd import capi, cases
    print("device memory leak check", flush=True)
    case = cases.cylinder_case(h=1.0 / 28, ppc=6, dt=2e-3)

    def cycle():
        mesh, ph, stepper = cases.build_gpu(ctx, case, precision=1, inner_tol=1e-5)
        stepper.set_graphs(True)
        for _ in range(3):
            stepper.step()
        a = capi.CSR(ctx, mesh=mesh)
        rhs = ctx.zeros(case.xy.shape[0])
        capi.assemble_poisson2d(mesh, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=1.0)
        capi.Dirichlet(ctx, case.xy.shape[0], case.bcp_node, case.bcp_val).apply(a, rhs, capi.DIRICHLET_MASK)
        for ordering in (capi.ORDER_NATURAL, capi.ORDER_MULTICOLOR, capi.ORDER_BLOCKS):
            pc = capi.Preconditioner(a, capi.PRECOND_IC0, ordering)
            if ordering != capi.ORDER_NATURAL:
                pc.set_factor_bits(16)
            s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 2000, 0, capi.TOL_RELATIVE, pc)
            if ordering != capi.ORDER_NATURAL:
                s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 20)
            x = ctx.zeros(case.xy.shape[0])
            assert s.solve(a, x, rhs).converged
            del s, pc, x
        del stepper, ph, mesh, a, rhs
        ctx.sync()

    cycle()
    base = free_bytes()
    for k in range(3):
        cycle()
        now = free_bytes()
        # allow for allocation granularity; a leak of one cycle's buffers would be tens of MB
        assert base - now < 4 << 20, (k, base - now)
