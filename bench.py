#!/usr/bin/env python3
"""PFEM-2 steps/s on one MI355X: flow past a cylinder, ~2.02M triangles, ~40M particles (scheme B,
dt = 2e-4, two outer iterations, pressure PCG + IC0, prediction GMRES(30) + Jacobi, correction
PCG + Jacobi, relative tolerance 1e-12), in the mixed-precision pipeline and in plain FP64.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --kernels      # only the pressure-solve kernels (for counter collection)
    python bench.py --workload poisson3d   # 8-RHS 3D Poisson solves (secondary workload)
    python bench.py --workload batched     # 256 independent Poisson systems (secondary workload)

One step = advection + relocation + refill + reorder of all particles, projection, the two outer
iterations of prediction / pressure / correction (5 assemblies + 5 solves each), particle velocity
correction and the drag/lift integral. The flow starts from rest; W warm-up steps, then K timed
steps bracketed by device synchronisation. Prints one JSON line.

Mixed precision: Jacobi-preconditioned solves run FP32 inner Krylov + FP64 refinement; the IC0
pressure solve runs FP64 PCG with the IC0 factor stored in FP32 (FP32 refinement stalls on the
pressure matrix, whose condition number is ~1/eps32).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def gpu_run(ctx, case, precision, steps, warmup, inner_tol):
    from gpupfem2_amd import cases
    t0 = time.perf_counter()
    mesh, ph, st = cases.build_gpu(ctx, case, precision=precision, inner_tol=inner_tol)
    ctx.sync()
    setup = time.perf_counter() - t0
    for _ in range(warmup):
        st.step()
    ctx.sync()
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        forces = st.step()
        its.append(st.iterations().tolist())
    ctx.sync()
    dt = time.perf_counter() - t0
    out = dict(steps_per_s=steps / dt, ms_per_step=1e3 * dt / steps, setup_s=setup, iterations=its,
               particles=ph.count(), cd_cl=forces.tolist())
    return out, (mesh, ph, st)


def csr_info(a):
    from gpupfem2_amd import capi
    n, nnz = capi.C.c_int(), capi.C.c_int()
    capi.lib().pfem_csr_info(a, capi.C.byref(n), capi.C.byref(nnz))
    return n.value, nnz.value


def ic0_apply_bytes(n, nnz, factor_bytes, vec_bytes=8):
    """Algorithmic bytes of one IC0 apply (both triangular solves): the factor's two triangles with
    their diagonals and column indices, the row offsets twice, and four vector passes (r read, y
    written and read, z written)."""
    return (nnz + n) * (factor_bytes + 4) + 4 * (n + 1) * 2 + 4 * n * vec_bytes


def spmv_bytes(n, nnz, val_bytes, vec_bytes):
    return nnz * (val_bytes + 4) + 4 * (n + 1) + 2 * n * vec_bytes


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "ic0_apply_cylinder2M.md")


def measured_traffic(key):
    """HBM bytes per launch from the committed counter summary (FETCH_SIZE + WRITE_SIZE runs of
    `bench.py --kernels`, each scaled by what the counter reads for a 256 MiB copy); None if absent."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith(key + " = "):
                    return int(line.split("=")[1])
    except OSError:
        pass
    return None


def roofline_ic0(ctx, st, precision):
    """The dominant kernels of a step: the two single-pass triangular sweeps of the pressure IC0 apply
    (k_sweep_lower + k_sweep_upper, about 55% of the step), timed with HIP events on the context
    stream over 200 applies."""
    from gpupfem2_amd import capi
    a, pc = st.matrix_handle(2), st.precond_handle(1)
    n, nnz = csr_info(a)
    # the apply as the pressure PCG runs it: vectors in the multicolour factor's ordering, y and z in
    # one tagged buffer (fusion level 3); in mixed precision the step's factor copy is 16-bit
    mixed = precision == 1
    op = 8 if mixed else 9
    ms = capi.time_kernel(ctx, op, a, pc, launches=200)
    b = ic0_apply_bytes(n, nnz, 2 if mixed else 8)
    gbs = b / (ms * 1e-3) / 1e9
    traffic = measured_traffic("traffic_bytes_ic0_apply_fp16_factor" if mixed else "traffic_bytes_ic0_apply_fp64")
    return dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(gbs / HBM_PEAK_GBS, 4), traffic=traffic,
                traffic_source="profiles/ic0_apply_cylinder2M.md" if traffic else None,
                kernel=(f"IC0 apply = k_sweep_lower + k_sweep_upper (tagged), pressure matrix n={n} nnz={nnz}, "
                        f"factor {'FP16 (power-of-two scaled)' if mixed else 'FP64'}, vectors FP64 in factor order"),
                avg_us=round(ms * 1e3, 2), bytes_per_launch=b)


def pcg_iteration_bytes(n, nnz, factor_bytes):
    """Algorithmic bytes of one fused pressure PCG iteration (three launches): the SpMV (A, p gathered, q
    written), the forward sweep (the lower triangle with its diagonal; r read and written, q read, y
    written; gamma = (y, y) from registers) and the backward sweep with the x/p updates (the upper
    triangle with the diagonal; y read, z written, x and p read and written)."""
    tri = (nnz - n) // 2 * (factor_bytes + 4) + n * factor_bytes
    return spmv_bytes(n, nnz, 8, 8) + (tri + 4 * 8 * n) + (tri + 6 * 8 * n)


def roofline_pcg(ctx, st, precision, tol=1e-9):
    """One pressure solve from x = 0 (relative tolerance 1e-9, a seeded right-hand side) with the step's
    matrix and preconditioner, as the step runs it: time per iteration (the solve's wall time, which also
    holds one factorization, over its iterations) against the algorithmic bytes of an iteration."""
    import types
    import numpy as np
    from gpupfem2_amd import capi
    a = types.SimpleNamespace(h=st.matrix_handle(2), ctx=ctx)
    pc = types.SimpleNamespace(h=st.precond_handle(1))
    n, nnz = csr_info(a.h)
    s = capi.Solver(a, capi.SOLVER_CG, tol, 20000, 0, capi.TOL_RELATIVE, pc)
    if precision == 1:
        s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 0)
    b = ctx.to_device(np.random.default_rng(1).standard_normal(n))
    best = None
    for _ in range(2):
        x = ctx.zeros(n)
        ctx.sync()
        t = time.perf_counter()
        stt = s.solve(a, x, b)
        ctx.sync()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    us = best / stt.iterations * 1e6
    b_it = pcg_iteration_bytes(n, nnz, 2 if precision == 1 else 8)
    gbs = b_it / (us * 1e-6) / 1e9
    del s
    traffic = measured_traffic("traffic_bytes_pcg_iteration_fp16_factor" if precision == 1
                               else "traffic_bytes_pcg_iteration_fp64")
    return dict(iterations=stt.iterations, us_per_iteration=round(us, 2), bytes_per_iteration=b_it,
                achieved=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4), traffic_per_iteration=traffic,
                kernels="k_cg_spmv_alpha_sliced + k_sweep_lower (r update, gamma) + k_sweep_upper (x, p updates)")


def roofline_spmv(ctx, st):
    """FP64 SpMV of the pressure matrix (the SpMV of every pressure PCG iteration)."""
    from gpupfem2_amd import capi
    a = st.matrix_handle(2)
    n, nnz = csr_info(a)
    ms = capi.time_kernel(ctx, 0, a, None, launches=200)
    b = spmv_bytes(n, nnz, 8, 8)
    gbs = b / (ms * 1e-3) / 1e9
    return dict(kernel=f"k_spmv<double> pressure matrix n={n} nnz={nnz}", achieved=round(gbs, 1),
                frac=round(gbs / HBM_PEAK_GBS, 4), avg_us=round(ms * 1e3, 2), bytes_per_launch=b)


def natural_over_multicolour(ctx):
    """Iterations of IC0-PCG (oracle, relative 1e-12, random right-hand side) with the natural-order
    factor over those with the library's multicolour factor, on the cylinder channel's pressure matrix at
    h = 1/28 (65k nodes): the CPU runs the reference's natural ordering, which converges faster."""
    import numpy as np
    import scipy.sparse as sp
    from gpupfem2_amd import capi, cases
    from oracle import oracle as O
    sc = cases.cylinder_case(h=1.0 / 28)
    nv = sc.xy.shape[0]
    area, invJ = O.geometry2d(sc.xy, sc.tri)
    rp, col = O.csr_pattern(nv, sc.tri)
    val, _ = O.assemble_poisson2d(sc.xy, sc.tri, area, invJ, rp, col, 1.0, 2, 0.0, 1)
    b = np.random.default_rng(0).standard_normal(nv)
    O.dirichlet_apply(rp, col, val, b, sc.bcp_node, sc.bcp_val, sc.dirichlet_mode)
    nat = O.pcg(rp, col, val, np.zeros(nv), b, 1e-12, 20000, 1, "ic0").iterations
    m = capi.Mesh2D(ctx, sc.xy, sc.tri)
    a = capi.CSR(ctx, mesh=m)
    a.upload_values(val)
    pc = capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_MULTICOLOR)
    pc.initialize(a)
    _, _, perm = pc.download()
    B = sp.csr_matrix((val, col, rp), shape=(nv, nv))[perm][:, perm].tocsr()
    B.sort_indices()
    mc = O.pcg(B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.copy(), np.zeros(nv),
               np.ascontiguousarray(b[perm]), 1e-12, 20000, 1, "ic0").iterations
    return nat / mc


def cpu_baseline(case, st, mesh, gpu_its):
    """The oracle (1 thread, FP64) on the same full-size inputs: particle step, the 5 assemblies, and
    per-iteration Krylov costs measured with capped iteration counts, combined with the iteration
    counts of the GPU FP64 step into a time per step."""
    import numpy as np
    from oracle import oracle as O
    T0 = time.perf_counter()
    f = st.fields()
    nv = case.xy.shape[0]
    area, invJ = O.geometry2d(case.xy, case.tri)
    off, idx = mesh.download(4), mesh.download(5)
    prm = O.make_params(case.rho, case.mu, case.dt, case.scheme, case.traction_mask, case.normals)
    rp, col = O.csr_pattern(nv, case.tri)
    # particles: seed, one advect (3 substeps, relocation, refill, reorder), projection, correction
    ph = O.Particles(case.xy, case.tri, invJ, off, idx, case.level)
    ph.seed(case.seed_mode, case.ppc, case.seed)
    ph.init_velocity(f["u0"], f["u1"])
    t = time.perf_counter()
    ph.advect(f["u0"], f["u1"], case.dt, case.substeps)
    u0, u1 = ph.project(nv)
    ph.correct(f["u0"], f["u1"], u0, u1)
    t_part = time.perf_counter() - t
    del ph
    # assemblies (one of each system) + Dirichlet
    t = time.perf_counter()
    v0, v1, r0, r1 = O.assemble_prediction(case.xy, case.tri, area, invJ, case.edge_ids, rp, col, f["u0"], f["u1"],
                                           f["u0"], f["u1"], f["p"], prm)
    t_pred_asm = time.perf_counter() - t
    t = time.perf_counter()
    vp, rpv = O.assemble_pressure(case.tri, area, invJ, rp, col, f["vp0"], f["vp1"], f["p"], prm)
    O.dirichlet_apply(rp, col, vp, rpv, case.bcp_node, case.bcp_val, case.dirichlet_mode)
    t_pres_asm = time.perf_counter() - t
    t = time.perf_counter()
    vc, c0, c1 = O.assemble_correction(case.tri, area, invJ, rp, col, f["vp0"], f["vp1"], f["p"], f["p"], prm)
    t_corr_asm = time.perf_counter() - t
    O.dirichlet_apply(rp, col, v0, r0, case.bcu_node, case.bcu_val0, case.dirichlet_mode)
    O.dirichlet_apply(rp, col, vc, c0, case.bcu_node, case.bcu_val0, case.dirichlet_mode)

    def per_iter(fn, lo, hi):
        ta = time.perf_counter(); fn(lo); tb = time.perf_counter(); fn(hi); tc = time.perf_counter()
        it = ((tc - tb) - (tb - ta)) / (hi - lo)
        return max(it, 1e-9), max((tb - ta) - lo * it, 0.0)

    x = np.zeros(nv)
    pres_it, pres_setup = per_iter(lambda k: O.pcg(rp, col, vp, x * 0, rpv, 1e-300, k, 1, "ic0"), 10, 60)
    corr_it, corr_setup = per_iter(lambda k: O.pcg(rp, col, vc, x * 0, c0, 1e-300, k, 1, "jacobi"), 10, 60)
    pred_it, pred_setup = per_iter(lambda k: O.gmres(rp, col, v0, x * 0, r0, 1e-300, k, 30, 1, "jacobi"), 10, 60)
    its = np.asarray(gpu_its[-1]).reshape(case.outer_iterations, 5)
    ratio = natural_over_multicolour(mesh.ctx)
    t_solve = (its[:, 0:2].sum() * pred_it + 2 * case.outer_iterations * pred_setup +
               ratio * its[:, 2].sum() * pres_it + case.outer_iterations * pres_setup +
               its[:, 3:5].sum() * corr_it + 2 * case.outer_iterations * corr_setup)
    t_step = t_part + case.outer_iterations * (2 * t_pred_asm + t_pres_asm + t_corr_asm) + t_solve
    return dict(value=round(1.0 / t_step, 6), unit="steps/s", cores=O.threads(), kind="port",
                sample=(f"oracle (1 thread, FP64) on the full {case.tri.shape[0]}-triangle case: one particle "
                        f"advect+project+correct ({t_part:.1f} s), one assembly of each system, and 50 "
                        f"Krylov iterations per solver type (IC0-PCG {pres_it * 1e3:.1f} ms/it, Jacobi-PCG "
                        f"{corr_it * 1e3:.1f} ms/it, GMRES(30) {pred_it * 1e3:.1f} ms/it), combined with the "
                        f"GPU FP64 step's iteration counts, the pressure's scaled by {ratio:.3f} (the natural-order "
                        f"IC0 of the reference against the GPU's multicolour factor, measured on the same channel "
                        f"at h = 1/28); measured in {time.perf_counter() - T0:.0f} s"),
                seconds_per_step=round(t_step, 2))


def kernels_only(h, launches):
    """Pressure-system kernels alone: the cylinder mesh's P1 stiffness matrix (the pressure matrix:
    it depends on the geometry only) with the pressure Dirichlet rows, its multicolour IC0 factor, then
    `launches` IC0 applies on vectors in the factor's ordering (16-bit, FP32 and FP64 factor), FP64 SpMVs
    in the natural and in the factor's ordering, `launches` iterations of the fused PCG with each factor
    (16-bit and FP64), and device-to-device copies of 256 MiB (the calibration of the byte counters)."""
    import numpy as np
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=h)
    ctx = capi.Context(0)
    m = capi.Mesh2D(ctx, case.xy, case.tri)
    a = capi.CSR(ctx, mesh=m)
    rhs = ctx.zeros(case.xy.shape[0])
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=0.0)
    capi.Dirichlet(ctx, case.xy.shape[0], case.bcp_node, case.bcp_val).apply(a, rhs, capi.DIRICHLET_MASK)
    pc = capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_MULTICOLOR)
    pc.initialize(a)
    pc16 = capi.Preconditioner(a, capi.PRECOND_IC0, capi.ORDER_MULTICOLOR)
    pc16.set_factor_bits(16)
    pc16.initialize(a)
    n, nnz = a.n, a.nnz
    print(f"[kernels] n={n} nnz={nnz}", file=sys.stderr, flush=True)
    res = {}
    for op, name, p in ((8, "ic0_apply_fp16_factor", pc16), (8, "ic0_apply_fp32_factor", pc), (9, "ic0_apply_fp64", pc),
                        (6, "ic0_apply_fp32_factor_sentinel", pc), (5, "ic0_apply_fp64_sentinel", pc),
                        (0, "spmv_fp64", None), (7, "spmv_fp64_factor_order", pc)):
        ms = capi.time_kernel(ctx, op, a.h, p.h if p else None, launches=launches)
        res[name] = round(ms * 1e3, 2)
        print(f"[kernels] {name}: {ms * 1e3:.2f} us", file=sys.stderr, flush=True)
    # the fused PCG's own kernels: `launches` iterations of a pressure solve per factor (16-bit copy with
    # the mixed step's settings, FP64), on a seeded right-hand side that does not converge in so few
    b = ctx.to_device(np.random.default_rng(1).standard_normal(n))
    for name, p, mixed in (("pcg_fp16_factor", pc16, True), ("pcg_fp64", pc, False)):
        s = capi.Solver(a, capi.SOLVER_CG, 1e-30, launches, 0, capi.TOL_RELATIVE, p)
        if mixed:
            s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 0)
        x = ctx.zeros(n)
        st = s.solve(a, x, b)
        res[name + "_iterations"] = st.iterations
        print(f"[kernels] {name}: {st.iterations} iterations", file=sys.stderr, flush=True)
        del s
    res["pcg_iteration_bytes_fp16_factor"] = pcg_iteration_bytes(n, nnz, 2)
    res["pcg_iteration_bytes_fp64"] = pcg_iteration_bytes(n, nnz, 8)
    mib = 256
    src, dst = ctx.zeros(mib << 17), ctx.empty(mib << 17)
    for _ in range(launches):
        dst.copy_from(src)
    ctx.sync()
    print(f"[kernels] copied {launches} x {mib} MiB", file=sys.stderr, flush=True)
    res["copy_bytes_read"] = mib << 20
    res["ic0_bytes_fp16_factor"] = ic0_apply_bytes(n, nnz, 2)
    res["ic0_bytes_fp32_factor"] = ic0_apply_bytes(n, nnz, 4)
    res["ic0_bytes_fp64"] = ic0_apply_bytes(n, nnz, 8)
    print(json.dumps(res))


def multi_iteration_bytes(n, nnz, g, vec_bytes, val_bytes):
    """Algorithmic HBM bytes of one multi-RHS PCG iteration (Jacobi, g interleaved columns): the SpMV
    streams the matrix once (values, column indices, row offsets) and reads p, writes q; the update
    reads p, q, the diagonal and reads/writes x, r; the p update reads r, the diagonal and reads/writes p."""
    m = n * g * vec_bytes
    return nnz * (val_bytes + 4) + 4 * (n + 1) + 2 * m + 6 * m + n * vec_bytes + 3 * m + n * vec_bytes


def poisson3d_workload(args):
    """Secondary workload: 3D Poisson on the 128^3-cube box (12.6M tetrahedra, 2.15M nodes), 8 right-hand
    sides solved together by Jacobi PCG to relative 1e-10, FP64 and mixed (FP32 inner + FP64 refinement).
    A step is one 8-column solve; value = columns solved per second."""
    from gpupfem2_amd import capi, cases
    case = cases.poisson3d_multi_case(n=128, nrhs=8)
    ctx = capi.Context(0)
    mesh, a, B, b0 = cases.build_poisson3d_multi(ctx, case)
    n, nnz = csr_info(a.h)
    out = {}
    for prec in (0, 1):
        s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 20000, 0, capi.TOL_RELATIVE,
                        capi.Preconditioner(a, capi.PRECOND_JACOBI))
        if prec:
            s.set_precision(prec, 1e-5, 30)
        X = ctx.zeros(8 * n)
        times, its = [], None
        for k in range(args.warmup + args.steps):
            X.zero()
            ctx.sync()
            t = time.perf_counter()
            st = s.solve_multi(a, X, B, 8)
            ctx.sync()
            if k >= args.warmup:
                times.append(time.perf_counter() - t)
            its = [x.iterations for x in st]
            assert all(x.converged for x in st)
        dt = sum(times) / len(times)
        it = max(its)
        b = multi_iteration_bytes(n, nnz, 8, 4 if prec else 8, 4 if prec else 8)
        out[prec] = {"s_per_solve": dt, "iterations": it, "refinements": st[0].refinements,
                     "iteration_gbs": round(b * it / dt / 1e9, 1)}
        print(f"[bench] poisson3d {'mixed' if prec else 'fp64'}: {dt * 1e3:.1f} ms per 8-column solve, "
              f"{it} iterations", file=sys.stderr, flush=True)
    head = out[1]
    line = {
        "metric": "3D Poisson 8-RHS PCG+Jacobi columns solved/s (128^3 box, 12.6M tets)",
        "value": round(8 / head["s_per_solve"], 3),
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["s_per_solve"] * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (manufactured RHS scaled per column + seeded noise)",
        "config": {"workload": "poisson3d box 128^3, 8 RHS, rel tol 1e-10", "nodes": n, "nnz": nnz,
                   "parallelism": "1 GPU"},
        "fp64_solves_per_s": round(8 / out[0]["s_per_solve"], 3),
        "mixed_over_fp64": round(out[0]["s_per_solve"] / out[1]["s_per_solve"], 3),
        "iterations": {"fp64": out[0]["iterations"], "mixed_inner": out[1]["iterations"]},
        "iteration_roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                               "fp64": out[0]["iteration_gbs"], "mixed": out[1]["iteration_gbs"]},
    }
    print(json.dumps(line))


def batched_workload(args):
    """Secondary workload: 256 independent pressure-Poisson systems on jittered meshes of 10k-500k
    triangles (log-uniform), one block-diagonal matrix, Jacobi PCG per system to relative 1e-10, FP64 and
    mixed (reliable updates). A step is one batched solve; value = systems solved per second."""
    import numpy as np
    from gpupfem2_amd import capi, cases
    case = cases.batched_poisson_case()
    ctx = capi.Context(0)
    mesh, a, b = cases.build_batched_poisson(ctx, case)
    rows = case["sys_rows"]
    n = int(rows[-1])
    out = {}
    for prec in (0, 1):
        s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 20000, 0, capi.TOL_RELATIVE,
                        capi.Preconditioner(a, capi.PRECOND_JACOBI))
        if prec:
            s.set_precision(prec, 1e-5, 30)
        x = ctx.zeros(n)
        times, t2s, its = [], [], []
        for k in range(args.warmup + args.steps):
            x.zero()
            ctx.sync()
            t = time.perf_counter()
            st, tms = s.solve_batched(a, rows, x, b)
            ctx.sync()
            assert all(q.converged for q in st)
            if k >= args.warmup:
                times.append(time.perf_counter() - t)
                t2s.append(tms)
            its = [q.iterations for q in st]
        dt = sum(times) / len(times)
        t2s = np.concatenate(t2s)
        out[prec] = {"s_per_batch": dt, "p50_ms": float(np.percentile(t2s, 50)), "p99_ms": float(np.percentile(t2s, 99)),
                     "iterations": [int(min(its)), int(max(its))]}
        print(f"[bench] batched {'mixed' if prec else 'fp64'}: {dt * 1e3:.1f} ms per batch of {rows.size - 1}",
              file=sys.stderr, flush=True)
    head = out[1]
    nsys = int(rows.size - 1)
    line = {
        "metric": "batched pressure-Poisson PCG+Jacobi solves/s (256 meshes, 10k-500k triangles)",
        "value": round(nsys / head["s_per_batch"], 2),
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["s_per_batch"] * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (seeded jittered meshes, N(0,1) right-hand sides)",
        "config": {"workload": "256 independent 2D Poisson systems, rel tol 1e-10", "triangles": int(case["tri"].shape[0]),
                   "nodes": n, "parallelism": "1 GPU"},
        "fp64_solves_per_s": round(nsys / out[0]["s_per_batch"], 2),
        "time_to_solution_ms": {("mixed" if k else "fp64"): {"p50": round(v["p50_ms"], 1), "p99": round(v["p99_ms"], 1)}
                                for k, v in out.items()},
        "iterations": {("mixed" if k else "fp64"): v["iterations"] for k, v in out.items()},
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--h", type=float, default=1.0 / 112, help="background mesh spacing (D = 1)")
    ap.add_argument("--mode", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--inner-tol", type=float, default=1e-5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernels", action="store_true", help="pressure-solve kernels only (counter runs)")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--workload", choices=["cylinder", "poisson3d", "batched"], default="cylinder",
                    help="cylinder: the headline PFEM-2 step; poisson3d: 8-RHS 3D Poisson solves; "
                         "batched: 256 independent Poisson systems")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)")
    if args.workload == "poisson3d":
        poisson3d_workload(args)
        return
    if args.workload == "batched":
        batched_workload(args)
        return
    if args.kernels:
        kernels_only(args.h, args.launches)
        return
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=args.h)
    ctx = capi.Context(0)
    res = {}
    roof = roof_spmv = None
    cpu = None
    roof_pcg = {}
    for prec in ([0, 1] if args.mode == "both" else [0] if args.mode == "fp64" else [1]):
        print(f"[bench] precision {'mixed' if prec else 'fp64'}: {case.tri.shape[0]} triangles, "
              f"{args.warmup} warm-up + {args.steps} timed steps", file=sys.stderr, flush=True)
        r, (mesh, ph, st) = gpu_run(ctx, case, prec, args.steps, args.warmup, args.inner_tol)
        res[prec] = r
        print(f"[bench] {r['steps_per_s']:.3f} steps/s, iterations {r['iterations'][-1]}", file=sys.stderr,
              flush=True)
        roof_pcg["mixed" if prec else "fp64"] = roofline_pcg(ctx, st, prec)
        print(f"[bench] pressure PCG iteration {roof_pcg['mixed' if prec else 'fp64']['us_per_iteration']} us",
              file=sys.stderr, flush=True)
        if prec == (1 if args.mode != "fp64" else 0):
            roof = roofline_ic0(ctx, st, prec)
            roof_spmv = roofline_spmv(ctx, st)
            print(f"[bench] IC0 apply {roof['avg_us']} us ({roof['achieved']} GB/s), SpMV {roof_spmv['avg_us']} us",
                  file=sys.stderr, flush=True)
        if prec == 0 and not args.no_cpu:
            print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
            cpu = cpu_baseline(case, st, mesh, r["iterations"])
        del st, ph, mesh
    head = res.get(1, res.get(0))
    line = {
        "metric": "PFEM steps/s mixed vs FP64 on 2M-tri cylinder mesh; solver HBM TB/s",
        "value": round(head["steps_per_s"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if 1 in res else "f64",
        "data": "synthetic (generated mesh, Philox-seeded particles, flow from rest)",
        "config": {"workload": "cylinder Re=100, scheme B, dt=2e-4", "triangles": int(case.tri.shape[0]),
                   "nodes": int(case.xy.shape[0]), "particles_per_cell": case.ppc,
                   "particles_after_steps": head["particles"], "parallelism": "1 GPU",
                   "precision": "mixed" if 1 in res else "fp64"},
        "fp64_steps_per_s": round(res[0]["steps_per_s"], 4) if 0 in res else None,
        "mixed_steps_per_s": round(res[1]["steps_per_s"], 4) if 1 in res else None,
        "mixed_over_fp64": round(res[1]["steps_per_s"] / res[0]["steps_per_s"], 3) if len(res) == 2 else None,
        "krylov_iterations_last_step": {("mixed" if k else "fp64"): v["iterations"][-1] for k, v in res.items()},
        "cd_cl_last_step": {("mixed" if k else "fp64"): v["cd_cl"] for k, v in res.items()},
        "roofline": roof,
        "roofline_spmv": roof_spmv,
        "roofline_pcg_iteration": roof_pcg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
