"""GPU parity of the benchmark configurations at their own sizes.

* Poisson, 64 x 64 unit square (8,192 triangles), coefficient 1, Jacobi PCG to 1e-10 relative
  (cases/Poisson2D/main.cu:87-183): the GPU solution against the oracle's PCG on the same system to 1e-10,
  and the discretisation error against sin(pi x) sin(pi y) falling as h^2 from 32 x 32 to 64 x 64.
* Poiseuille channel, 400 x 100 jittered (80k triangles, 16 particles per cell = 1.28M particles,
  GMRES(30) + Jacobi for the velocity systems, Jacobi PCG for the pressure;
  cases/PoiseuilleFlow2D/main.cu:652-736): three whole steps against the oracle (every field within
  relative L2 1e-8, particle -> cell indices identical), and 200 GPU steps from the analytic profile,
  which must stay the discrete steady state (x-averaged profile within O(h^2) + particle noise, pressure
  gradient -8 mu U_max).
* Cylinder, 2,021,376 triangles, 40.4M particles, scheme B, IC0-PCG pressure: the mixed-precision pipeline
  (FP32 inner Krylov with FP64 refinement for the velocity systems, 16-bit pressure factor) against the
  FP64 pipeline over the same steps (p, u relative L2 <= 1e-8, C_D and C_L relative <= 1e-6); and the FP64
  GPU pipeline against the oracle's whole step on the same channel at h = 1/28 (126k triangles).
"""
import math

import numpy as np
import pytest

from conftest import rel_l2
from helpers import oracle_stepper

pytestmark = pytest.mark.gpu


def poisson_system(ctx, oracle, n):
    from gpupfem2_amd import capi, cases
    xy, tri, bn, bv = cases.poisson_case(n)
    m = capi.Mesh2D(ctx, xy, tri, fill_neighbors=False)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=1, rhs_param=2 * math.pi ** 2)
    capi.Dirichlet(ctx, xy.shape[0], bn, bv).apply(a, rhs, cases.MASK)
    return xy, tri, m, a, rhs, bn


def test_poisson_64x64_jacobi_pcg(ctx, oracle):
    """cfg1: 4,225 nodes, PCG + Jacobi to 1e-10 relative, against the oracle on the assembled system
    (assembly and Dirichlet rows are bit-exact, the solves agree to the tolerance's scale) and against
    the manufactured solution at O(h^2)."""
    from gpupfem2_amd import capi
    errs = {}
    for n in (32, 64):
        xy, tri, m, a, rhs, bn = poisson_system(ctx, oracle, n)
        assert tri.shape[0] == 2 * n * n and xy.shape[0] == (n + 1) ** 2
        rp, col, val = a.download()
        b = rhs.download()
        # the oracle's own assembly and Dirichlet pass on the same mesh give the same system, bit for bit
        area, invJ = oracle.geometry2d(xy, tri)
        orp, ocol = oracle.csr_pattern(xy.shape[0], tri)
        oval, ob = oracle.assemble_poisson2d(xy, tri, area, invJ, orp, ocol, 1.0, 1, 2 * math.pi ** 2)
        oracle.dirichlet_apply(orp, ocol, oval, ob, bn, np.zeros(bn.size), 1)
        assert np.array_equal(orp, rp) and np.array_equal(ocol, col) and np.array_equal(oval, val)
        assert rel_l2(b, ob) < 1e-15   # sin comes from two math libraries
        pc = capi.Preconditioner(a, capi.PRECOND_JACOBI)
        s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 5000, 0, capi.TOL_RELATIVE, pc)
        x = ctx.zeros(xy.shape[0])
        st = s.solve(a, x, rhs)
        assert st.converged and not st.nan_detected
        xo = np.zeros(xy.shape[0])
        ost = oracle.pcg(rp, col, val, xo, b, 1e-10, 5000, 1, "jacobi")
        assert ost.converged and abs(ost.iterations - st.iterations) <= 3, (st.iterations, ost.iterations)
        xg = x.download()
        assert rel_l2(xg, xo) < 1e-10, rel_l2(xg, xo)
        exact = np.sin(math.pi * xy[:, 0]) * np.sin(math.pi * xy[:, 1])
        errs[n] = np.abs(xg - exact).max()
    # nodal error of P1 on a uniform mesh: O(h^2); a dropped RHS factor or a wrong stiffness would not
    # converge at all
    assert errs[64] < 1e-3 and 3.3 < errs[32] / errs[64] < 4.7, errs


def test_poiseuille_400x100_steps_vs_oracle(ctx, oracle):
    """cfg2 at its full size: 400 x 100 jittered channel, 1.28M particles, three steps from rest."""
    from gpupfem2_amd import cases
    case = cases.poiseuille_case()
    assert case.tri.shape[0] == 80_000 and case.ppc == 16
    print("poiseuille full size", case.name, flush=True)
    mesh, ph, stepper = cases.build_gpu(ctx, case)
    assert ph.get()["cell"].size == 1_280_000
    oph, ost = oracle_stepper(oracle, case, gpu_mesh=mesh)
    for s in range(3):
        stepper.step()
        ost.step()
        g, o = stepper.fields(), ost.fields()
        for k in ("u0", "u1", "p", "vp0", "vp1"):
            assert rel_l2(g[k], o[k]) < 1e-8, (s, k, rel_l2(g[k], o[k]))
        assert np.array_equal(ph.get()["cell"], oph.get()["cell"]), s
    gi, oi = stepper.iterations(), ost.iterations()
    assert gi.size == oi.size and np.all(np.abs(gi.astype(int) - oi) <= 3), (gi, oi)


def test_poiseuille_400x100_200_steps_steady(ctx):
    """cfg2, 200 steps of dt = 1e-3 on the GPU from the analytic profile u = 4 y (1 - y), v = 0: the flow
    stays there. The x-averaged velocity profile over the interior (0.5 < x < 3.5) stays within O(h^2) +
    particle noise of the parabola, the transverse velocity stays small, and the pressure gradient is the
    exact -8 mu U_max (a dropped viscous term gives about zero, a sign error diverges)."""
    from gpupfem2_amd import capi, cases
    case = cases.poiseuille_case()
    print("poiseuille 200 steps", case.name, flush=True)
    mesh, ph, stepper = cases.build_gpu(ctx, case)
    x, y = case.xy[:, 0], case.xy[:, 1]
    u_ex = 4.0 * y * (1.0 - y)
    z = np.zeros_like(y)
    for k, v in (("u0", u_ex), ("u1", z), ("p", z), ("vp0", u_ex), ("vp1", z)):
        stepper.set_field(k, v)
    capi.lib().pfem_particles_init_velocity(ph.h, stepper.field_ptr("u0"), stepper.field_ptr("u1"))
    inner = (x > 0.5) & (x < 3.5)
    bins = np.clip((y * 100).astype(int), 0, 99)   # strips of one cell height (the nodes are jittered)
    slopes, prof = [], []
    for s in range(200):
        stepper.step()
        if s >= 150 and s % 5 == 0:
            f = stepper.fields()
            assert np.isfinite(f["p"]).all() and np.isfinite(f["u0"]).all()
            slopes.append(np.polyfit(x[inner], f["p"][inner], 1)[0])
            d = f["u0"][inner] - u_ex[inner]
            prof.append(np.bincount(bins[inner], d, 100) / np.maximum(np.bincount(bins[inner], None, 100), 1))
            assert np.abs(f["u1"][x < 4.0 - 1e-9]).max() < 0.05
    prof = np.mean(prof, axis=0)
    assert np.abs(prof).max() < 0.01, np.abs(prof).max()
    slope = np.mean(slopes)
    assert abs(slope - (-8 * case.mu)) < 0.1 * 8 * case.mu, slope


def test_cylinder_mixed_vs_fp64_full_size(ctx):
    """cfg3 at full size: the mixed pipeline against the FP64 pipeline from the same mesh, seeds and
    particles, 4 steps. Bound (stated for this configuration): pressure and velocity relative L2 <= 1e-8,
    C_D and C_L relative <= 1e-6."""
    from gpupfem2_amd import cases
    case = cases.cylinder_case()
    print("cylinder mixed vs fp64", case.name, flush=True)
    res = {}
    for prec in (0, 1):
        mesh, ph, stepper = cases.build_gpu(ctx, case, precision=prec)
        forces = [stepper.step() for _ in range(4)]
        res[prec] = (stepper.fields(), np.array(forces))
        del stepper, ph, mesh
    (g64, f64), (gmx, fmx) = res[0], res[1]
    errs = {k: rel_l2(gmx[k], g64[k]) for k in ("u0", "u1", "p", "vp0", "vp1")}
    print("mixed vs fp64 field rel L2:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-8, (k, v)
    rel_f = np.abs(fmx - f64) / np.abs(f64)
    print("mixed vs fp64 C_D, C_L rel:", rel_f[-1], flush=True)
    assert np.all(rel_f <= 1e-6), rel_f


@pytest.mark.parametrize("passes", [1, 2])
def test_cylinder_h28_fp64_vs_oracle(ctx, oracle, passes):
    """The FP64 GPU pipeline in its benchmark form (multicolour IC0 pressure factor; prediction GMRES with
    one classical Gram-Schmidt pass, the default, or two) against the oracle's whole step (natural-order
    IC0, modified Gram-Schmidt GMRES) on the cylinder channel at h = 1/28 (126k triangles, 2.5M particles),
    two steps: fields within relative L2 1e-8 (both pressure solves stop at 1e-12 relative), C_D and C_L
    within 1e-6 relative, and the same particle -> cell indices."""
    from gpupfem2_amd import cases
    case = cases.cylinder_case(h=1.0 / 28)
    print("cylinder h=1/28 vs oracle", case.name, case.tri.shape[0], "passes", passes, flush=True)
    mesh, ph, stepper = cases.build_gpu(ctx, case, pres_ordering=1, pred_gs_passes=passes)
    oph, ost = oracle_stepper(oracle, case, gpu_mesh=mesh)
    for s in range(2):
        fg = stepper.step()
        fo = ost.step()
        g, o = stepper.fields(), ost.fields()
        for k in ("u0", "u1", "p", "vp0", "vp1"):
            assert rel_l2(g[k], o[k]) < 1e-8, (s, k, rel_l2(g[k], o[k]))
        assert np.array_equal(ph.get()["cell"], oph.get()["cell"]), s
        assert np.all(np.abs(fg - fo) <= 1e-6 * np.abs(fo)), (fg, fo)
        print("passes", passes, "step", s, "rel L2 u0 %.2e p %.2e, C_D %.2e C_L %.2e" % (
            rel_l2(g["u0"], o["u0"]), rel_l2(g["p"], o["p"]), abs(fg[0] - fo[0]) / abs(fo[0]),
            abs(fg[1] - fo[1]) / abs(fo[1])), flush=True)


def test_cylinder_step_sweep_fallback(ctx):
    """A single-pass sweep wait that exceeds its spin limit inside a step (forced from the host: a limit of
    0 polls on the pressure preconditioner, at the full 2M-triangle size where some wait has to poll) does
    not fail the step: the pressure solve is repeated from the same start with per-level sweeps, the
    fallback is counted, and the fields agree with an undisturbed step to the solver tolerance's scale."""
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case()
    print("cylinder step sweep fallback", flush=True)
    _, _, ref = cases.build_gpu(ctx, case, precision=1)
    f_ref = ref.step()
    g_ref = ref.fields()
    assert ref.sweep_fallbacks() == 0
    del ref
    _, _, st = cases.build_gpu(ctx, case, precision=1)
    capi.lib().pfem_precond_set_spin_limit(st.precond_handle(1), 0)
    f = st.step()
    assert st.sweep_fallbacks() >= 1
    g = st.fields()
    for k in ("u0", "u1", "p"):
        assert rel_l2(g[k], g_ref[k]) < 1e-8, (k, rel_l2(g[k], g_ref[k]))
    assert np.all(np.abs(f - f_ref) <= 1e-6 * np.abs(f_ref) + 1e-12), (f, f_ref)
