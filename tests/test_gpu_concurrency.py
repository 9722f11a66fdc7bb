"""Two contexts (two streams) on one GPU solving at the same time from two host threads: the fused
multicolour IC0-PCG (single-pass sweeps that wait on values published by earlier workgroups of their own
launch) and the blocked one, each interleaved on the device with the other's launches. Each solve must give
the bits and iteration count it gives alone: a launch only ever waits on workgroups of its own grid, which
were dispatched before it whatever else the device runs.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def pressure_system(ctx, case):
    from gpupfem2_amd import capi
    n = case.xy.shape[0]
    m = capi.Mesh2D(ctx, case.xy, case.tri, fill_neighbors=False)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(n)
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=1.0)
    capi.Dirichlet(ctx, n, case.bcp_node, case.bcp_val).apply(a, rhs, capi.DIRICHLET_MASK)
    return m, a, rhs


@pytest.mark.parametrize("orderings", [(1, 1), (1, 2)])
def test_concurrent_solves_two_streams(orderings):
    from gpupfem2_amd import capi, cases
    print("concurrent solves", orderings, flush=True)
    case = cases.cylinder_case(h=1.0 / 56)
    n = case.xy.shape[0]
    ctxs = [capi.Context(0), capi.Context(0)]
    systems = [pressure_system(c, case) for c in ctxs]

    def make(k):
        c = ctxs[k]
        m, a, rhs = systems[k]
        s = capi.Solver(a, capi.SOLVER_CG, 1e-12, 20000, 0, capi.TOL_RELATIVE,
                        capi.Preconditioner(a, capi.PRECOND_IC0, orderings[k]))
        return s

    def solve(k, s, out):
        c = ctxs[k]
        m, a, rhs = systems[k]
        x = c.zeros(n)
        st = s.solve(a, x, rhs)
        out[k] = (st.converged, st.iterations, x.download())

    solvers = [make(0), make(1)]
    alone = {}
    for k in range(2):
        solve(k, solvers[k], alone)
        assert alone[k][0]
    for rep in range(3):
        together = {}
        threads = [threading.Thread(target=solve, args=(k, solvers[k], together)) for k in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not any(t.is_alive() for t in threads)
        for k in range(2):
            assert together[k][0] and together[k][1] == alone[k][1], (rep, k, together[k][1], alone[k][1])
            assert np.array_equal(together[k][2], alone[k][2]), (rep, k)
    for c in ctxs:
        c.sync()
    # the handles before the contexts they were made on
    del solvers, systems, together, threads
    for c in ctxs:
        c.close()


def test_concurrent_steppers_with_graphs():
    """Two PFEM-2 steppers (the small cylinder and the small Poiseuille channel) on two contexts, stepped
    at the same time from two host threads with the step graphs on: one thread captures its graphs (a
    thread-local capture) while the other launches, allocates and synchronises. Each gives, step by
    step, the forces, fields, iteration counts and particle counts it gives alone."""
    from gpupfem2_amd import capi, cases
    print("concurrent steppers", flush=True)
    case_list = [cases.cylinder_case(h=1.0 / 8, ppc=6, dt=2e-3), cases.poiseuille_case(nx=40, ny=10, ppc=9, dt=2e-3)]
    ctxs = [capi.Context(0), capi.Context(0)]
    steps = 5

    def run(k, out):
        mesh, ph, stepper = cases.build_gpu(ctxs[k], case_list[k], precision=k, inner_tol=1e-5)
        stepper.set_graphs(True)
        res = []
        for s in range(steps):
            f = stepper.step()
            res.append((f.copy(), stepper.fields(), stepper.iterations().copy(), ph.count()))
        out[k] = res
        del stepper, ph, mesh

    alone = {}
    for k in range(2):
        run(k, alone)
    together = {}
    threads = [threading.Thread(target=run, args=(k, together)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=200)
    assert not any(t.is_alive() for t in threads)
    for k in range(2):
        assert k in together, k
        for s, (a, b) in enumerate(zip(alone[k], together[k])):
            assert np.array_equal(a[0], b[0]), (k, s)
            for f in a[1]:
                assert np.array_equal(a[1][f], b[1][f]), (k, s, f)
            assert np.array_equal(a[2], b[2]) and a[3] == b[3], (k, s)
    for c in ctxs:
        c.sync()
    del together, alone, threads
    for c in ctxs:
        c.close()
