#!/usr/bin/env python3
"""PFEM-2 steps/s on one MI355X: flow past a cylinder, ~2.02M triangles, ~40M particles (scheme B,
dt = 2e-4, two outer iterations, pressure PCG + IC0, prediction GMRES(30) + Jacobi, correction
PCG + Jacobi, relative tolerance 1e-12), in the mixed-precision pipeline and in plain FP64.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --kernels      # only the pressure-solve kernels (for counter collection)
    python bench.py --blocks       # multicolour against blocked IC0 ordering on the pressure system
    python bench.py --workload poisson3d   # 8-RHS 3D Poisson solves (secondary workload)
    python bench.py --workload batched     # 256 independent Poisson systems (secondary workload)
    python bench.py --workload history --steps 500   # C_D / C_L histories, FP64 and mixed

One step = advection + relocation + refill + reorder of all particles, projection, the two outer
iterations of prediction / pressure / correction (5 assemblies + 5 solves each), particle velocity
correction and the drag/lift integral. The flow starts from rest; W warm-up steps, then K timed
steps bracketed by device synchronisation. The work between the solves is replayed from HIP graphs
(pfem_pfem2_set_graphs; --no-graphs launches it kernel by kernel). Prints one JSON line, with the
mixed pipeline's error against the FP64 one after the same steps (fields, C_D, C_L).

Mixed precision: Jacobi-preconditioned solves run FP32 inner Krylov + FP64 refinement; the IC0
pressure solve runs FP64 PCG with the IC0 factor stored in 16 bits (IEEE half scaled by a power of two;
FP32 refinement stalls on the pressure matrix, whose condition number is ~1/eps32).

The line's `roofline` is the fused pressure PCG iteration (about 93% of a step), timed with HIP events;
`cpu_baseline` is one full FP64 step of the OpenMP build of the oracle on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def gpu_run(ctx, case, precision, steps, warmup, inner_tol, graphs=True):
    from gpupfem2_amd import cases
    t0 = time.perf_counter()
    mesh, ph, st = cases.build_gpu(ctx, case, precision=precision, inner_tol=inner_tol)
    if graphs:
        st.set_graphs(True)
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
               particles=ph.count(), cd_cl=forces.tolist(), fields={k: st.field(k) for k in ("u0", "u1", "p")})
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


TRAFFIC_PROFILE_REL = "profiles/ic0_apply_cylinder2M.md"
TRAFFIC_PROFILE = os.path.join(ROOT, TRAFFIC_PROFILE_REL)


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
    """Secondary: the standalone pressure IC0 apply (k_sweep_lower + k_sweep_upper, tagged, in the factor's
    ordering), timed with HIP events on the context stream over 200 applies."""
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


def roofline_pcg(ctx, st, precision, iterations=400):
    """The step's dominant work: one fused pressure PCG iteration (k_cg_spmv_alpha_sliced + k_sweep_lower
    + k_sweep_upper_xp) as the step runs it, on the step's own pressure matrix and preconditioner (16-bit
    factor copy in the mixed pipeline, FP64 factor otherwise). Timed with HIP events on the context stream
    (pfem_time_kernel ops 11 / 10: a solve of 2L iterations minus one of L, over L; the set-up cancels)
    against the algorithmic bytes of one iteration; traffic from the committed counter runs."""
    from gpupfem2_amd import capi
    a, pc = st.matrix_handle(2), st.precond_handle(1)
    n, nnz = csr_info(a)
    mixed = precision == 1
    ms = capi.time_kernel(ctx, 11 if mixed else 10, a, pc, launches=iterations)
    us = ms * 1e3
    b_it = pcg_iteration_bytes(n, nnz, 2 if mixed else 8)
    gbs = b_it / (us * 1e-6) / 1e9
    traffic = measured_traffic("traffic_bytes_pcg_iteration_fp16_factor" if mixed else "traffic_bytes_pcg_iteration_fp64")
    return dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4),
                traffic=traffic, traffic_source=TRAFFIC_PROFILE_REL if traffic else None,
                kernel=(f"fused pressure PCG iteration = k_cg_spmv_alpha_sliced + k_sweep_lower (r update, gamma) + "
                        f"k_sweep_upper_xp (x, p updates), n={n} nnz={nnz}, factor "
                        f"{'FP16 (power-of-two scaled)' if mixed else 'FP64'}, vectors FP64"),
                avg_us=round(us, 2), bytes_per_launch=b_it, timing="HIP events on the context stream, "
                f"solves of {2 * iterations} and {iterations} iterations differenced")


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


def copy_ceiling(ctx, gib=1, copies=10):
    """The device's streaming ceiling: copies of `gib` GiB (4x the 256 MB last-level cache), each reading
    and writing the buffer once -- the library's copy kernel (pfem_stream_copy: 16-byte loads, 4 or 8 in
    flight per lane, the best of 14 launch forms, HIP events) and the runtime's device-to-device copy (wall
    clock)."""
    from gpupfem2_amd import capi
    nbytes = gib << 30
    ms, grid = capi.stream_copy(ctx, nbytes, copies)
    kernel_gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    n = nbytes // 8
    src, dst = ctx.zeros(n), ctx.empty(n)
    dst.copy_from(src)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(copies):
        dst.copy_from(src)
    ctx.sync()
    dt = (time.perf_counter() - t) / copies
    memcpy_gbs = 2 * nbytes / dt / 1e9
    src.free()
    dst.free()
    best = max(kernel_gbs, memcpy_gbs)
    return dict(achieved=round(best, 1), frac_of_peak=round(best / HBM_PEAK_GBS, 4), unit="GB/s",
                copy_kernel_gbs=round(kernel_gbs, 1), copy_kernel_loads_in_flight=grid // 100000,
                copy_kernel_workgroups=grid % 100000, memcpy_d2d_gbs=round(memcpy_gbs, 1),
                method=f"copies of {gib} GiB (read + write): pfem_stream_copy (HIP events, best of 14 launch forms, "
                       f"{copies} launches each) and hipMemcpyAsync device-to-device (wall clock, {copies} copies); "
                       f"the better of the two")


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(case, st, mesh, ph):
    """One full FP64 time step of the same workload on the host cores: the oracle's OpenMP perf build
    (oracle/liboracle_omp.so, -O3 -fopenmp: the same algorithm restated in C++, particles, projection, the
    five assemblies, prediction GMRES(30) + Jacobi, pressure PCG + IC0 on a multicolour ordering (rows of
    one colour in parallel), correction PCG + Jacobi), started from the GPU FP64 run's state after its timed
    steps (fields and all particles copied to the host), timed as one step."""
    import numpy as np
    from oracle import oracle as O
    O.use_openmp_build()
    T0 = time.perf_counter()
    f = st.fields()
    off, idx = mesh.download(4), mesh.download(5)
    area, invJ = O.geometry2d(case.xy, case.tri)
    oph = O.Particles(case.xy, case.tri, invJ, off, idx, case.level)
    pst = ph.get()
    oph.set(pst, int(pst["id"].max()) + 1 if pst["id"].size else 0)
    del pst
    cfg = O.StepConfig()
    cfg.prm = O.make_params(case.rho, case.mu, case.dt, case.scheme, case.traction_mask, case.normals)
    for k in ("outer_iterations", "substeps", "pred_solver", "pred_precond", "pres_solver", "pres_precond",
              "corr_solver", "corr_precond", "tol", "tol_mode", "maxit", "restart", "dirichlet_mode", "body_id",
              "force_coeff"):
        setattr(cfg, k, getattr(case, k))
    cfg.body_cx, cfg.body_cy = case.body_center
    ost = O.Pfem2(case.xy, case.tri, case.edge_ids, cfg, oph, case.bcu_node, case.bcu_val0, case.bcu_val1,
                  case.bcp_node, case.bcp_val)
    ost.set_fields(u0=f["u0"], u1=f["u1"], p=f["p"], vp0=f["vp0"], vp1=f["vp1"])
    ost.set_coloured_pressure(True)
    npart = oph.count()
    print(f"[bench] cpu baseline: one FP64 step on {O.threads()} threads, {npart} particles", file=sys.stderr,
          flush=True)
    t = time.perf_counter()
    forces = ost.step()
    dt = time.perf_counter() - t
    its = ost.iterations().tolist()
    return dict(value=round(1.0 / dt, 6), unit="steps/s", cores=O.threads(), kind="port",
                sample=(f"one full FP64 time step of the {case.tri.shape[0]}-triangle cylinder ({npart} particles) "
                        f"by the oracle's OpenMP build (-O3 -fopenmp) on the host cores, from the GPU FP64 run's "
                        f"state: {dt:.1f} s; Krylov iterations {its}; pressure IC0 in a greedy multicolour ordering; "
                        f"set-up and transfers {t - T0:.0f} s (not timed)"),
                seconds_per_step=round(dt, 2), cd_cl=[float(v) for v in forces],
                host={"nproc": os.cpu_count(), "cpu_model": cpu_info(), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
                      "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "threads_used": O.threads()})


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
    # the bench line's roofline measurement itself (pfem_time_kernel ops 11 / 10: solves of 2L and L
    # iterations differenced), so that a kernel trace of this run holds the same launches
    for name, p, op in (("pcg_iteration_fp16_factor_us", pc16, 11), ("pcg_iteration_fp64_us", pc, 10)):
        ms = capi.time_kernel(ctx, op, a.h, p.h, launches=launches)
        res[name] = round(ms * 1e3, 2)
        print(f"[kernels] {name}: {ms * 1e3:.2f} us", file=sys.stderr, flush=True)
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


def blocks_only(h, launches):
    """The pressure system of kernels_only with the two orderings of its IC0 factor, 16-bit copy, FP64
    vectors: multicolour (the step's) and blocked (PFEM_ORDER_BLOCKS: blocks of 8 nodes coloured as
    blocks, fewer PCG iterations, longer sweeps). Per ordering: `launches` tagged applies, the fused PCG
    iteration (pfem_time_kernel op 11) and one solve to relative 1e-10 from zero on a seeded right-hand
    side (iterations and wall time). Prints one JSON line."""
    import numpy as np
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=h)
    ctx = capi.Context(0)
    m = capi.Mesh2D(ctx, case.xy, case.tri, fill_neighbors=False)
    a = capi.CSR(ctx, mesh=m)
    n = case.xy.shape[0]
    rhs = ctx.zeros(n)
    capi.assemble_poisson2d(m, a, rhs, coeff=1.0, rhs_kind=2, rhs_param=0.0)
    capi.Dirichlet(ctx, n, case.bcp_node, case.bcp_val).apply(a, rhs, capi.DIRICHLET_MASK)
    b = ctx.to_device(np.random.default_rng(1).standard_normal(n))
    res = {"n": n, "nnz": a.nnz}
    for name, ordering in (("multicolour", capi.ORDER_MULTICOLOR), ("blocks", capi.ORDER_BLOCKS)):
        t = time.perf_counter()
        pc = capi.Preconditioner(a, capi.PRECOND_IC0, ordering)
        pc.set_factor_bits(16)
        pc.initialize(a)
        ctx.sync()
        r = {"setup_s": round(time.perf_counter() - t, 2)}
        r["apply_us"] = round(capi.time_kernel(ctx, 8, a.h, pc.h, launches=launches) * 1e3, 2)
        r["pcg_iteration_us"] = round(capi.time_kernel(ctx, 11, a.h, pc.h, launches=launches) * 1e3, 2)
        s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 20000, 0, capi.TOL_RELATIVE, pc)
        s.set_precision(capi.PRECISION_MIXED_PRECOND, 1e-5, 0)
        for rep in range(2):   # the first solve also captures nothing and allocates the solver's buffers
            x = ctx.zeros(n)
            ctx.sync()
            t = time.perf_counter()
            st = s.solve(a, x, b)
            ctx.sync()
            dt = time.perf_counter() - t
        r["solve_iterations"] = st.iterations
        r["solve_ms"] = round(dt * 1e3, 1)
        res[name] = r
        print(f"[blocks] {name}: {r}", file=sys.stderr, flush=True)
        del s, pc
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
    from gpupfem2_amd import capi, cases, tune
    case = cases.batched_poisson_case()
    ctx = capi.Context(0)
    mesh, a, b = cases.build_batched_poisson(ctx, case)
    rows = case["sys_rows"]
    n = int(rows[-1])
    tuning = None if args.untuned else tune.load()
    out = {}
    for prec in (0, 1):
        s = capi.Solver(a, capi.SOLVER_CG, 1e-10, 20000, 0, capi.TOL_RELATIVE,
                        capi.Preconditioner(a, capi.PRECOND_JACOBI))
        if prec:
            s.set_precision(prec, 1e-5, 30)
        tune.apply(s, tuning)
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
        "tuning": ({"buckets": tuning["buckets"], "poll": tuning["poll"], "source": "gpupfem2_amd/tuning_batched.json"}
                   if tuning else "defaults (256-row, 2048-non-zero tiles, poll 8)"),
    }
    print(json.dumps(line))


def history_workload(args):
    """The cylinder configuration's drag and lift histories (Cylinder2D/main.cu:401-428, 872-874): `steps`
    time steps from rest in plain FP64 and in the mixed pipeline, C_D and C_L after every step, the work
    between the solves replayed from graphs as in the main workload (--no-graphs: kernel by kernel).
    Prints one JSON line with the per-mode histories (every step), the largest relative difference
    between the modes and the time per step; --out also writes the histories as CSV."""
    import numpy as np
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=args.h)
    ctx = capi.Context(0)
    hist, times = {}, {}
    for prec in (0, 1):
        mode = "mixed" if prec else "fp64"
        mesh, ph, st = cases.build_gpu(ctx, case, precision=prec, inner_tol=args.inner_tol)
        if not args.no_graphs:
            st.set_graphs(True)
        f = np.zeros((args.steps, 2))
        ctx.sync()
        t0 = time.perf_counter()
        for k in range(args.steps):
            f[k] = st.step()
            if (k + 1) % 50 == 0:
                print(f"[bench] {mode} step {k + 1}: C_D {f[k, 0]:.6f} C_L {f[k, 1]:.6e} "
                      f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        ctx.sync()
        times[mode] = (time.perf_counter() - t0) / args.steps
        hist[mode] = f
        del st, ph, mesh
    d = np.abs(hist["mixed"] - hist["fp64"])
    rel = d / np.maximum(np.abs(hist["fp64"]), 1e-300)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write("step,cd_fp64,cl_fp64,cd_mixed,cl_mixed\n")
            for k in range(args.steps):
                fh.write(f"{k + 1},{hist['fp64'][k, 0]:.17g},{hist['fp64'][k, 1]:.17g},"
                         f"{hist['mixed'][k, 0]:.17g},{hist['mixed'][k, 1]:.17g}\n")
    line = {
        "metric": "cylinder drag/lift histories, FP64 and mixed", "steps": args.steps,
        "config": {"workload": "cylinder Re=100, scheme B, dt=2e-4", "triangles": int(case.tri.shape[0])},
        "s_per_step": {k: round(v, 4) for k, v in times.items()},
        "max_rel_diff": {"cd": float(rel[:, 0].max()), "cl": float(rel[:, 1].max())},
        "max_abs_diff": {"cd": float(d[:, 0].max()), "cl": float(d[:, 1].max())},
        "final": {k: [float(v[-1, 0]), float(v[-1, 1])] for k, v in hist.items()},
        "history": {k: [[round(float(a), 9), float(b)] for a, b in v] for k, v in hist.items()},
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
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the work between the solves kernel by kernel instead of from HIP graphs")
    ap.add_argument("--kernels", action="store_true", help="pressure-solve kernels only (counter runs)")
    ap.add_argument("--blocks", action="store_true", help="multicolour against blocked IC0 on the pressure system")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--workload", choices=["cylinder", "poisson3d", "batched", "history"], default="cylinder",
                    help="cylinder: the headline PFEM-2 step; poisson3d: 8-RHS 3D Poisson solves; "
                         "batched: 256 independent Poisson systems; history: C_D/C_L over --steps steps, "
                         "FP64 and mixed")
    ap.add_argument("--out", default=None, help="history: CSV of the per-step C_D and C_L")
    ap.add_argument("--untuned", action="store_true", help="batched: ignore gpupfem2_amd/tuning_batched.json")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)")
    if args.workload == "poisson3d":
        poisson3d_workload(args)
        return
    if args.workload == "batched":
        batched_workload(args)
        return
    if args.workload == "history":
        history_workload(args)
        return
    if args.kernels:
        kernels_only(args.h, args.launches)
        return
    if args.blocks:
        blocks_only(args.h, args.launches)
        return
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=args.h)
    ctx = capi.Context(0)
    res = {}
    roof_ic0 = roof_spmv = None
    cpu = None
    roof_pcg = {}
    for prec in ([0, 1] if args.mode == "both" else [0] if args.mode == "fp64" else [1]):
        print(f"[bench] precision {'mixed' if prec else 'fp64'}: {case.tri.shape[0]} triangles, "
              f"{args.warmup} warm-up + {args.steps} timed steps", file=sys.stderr, flush=True)
        r, (mesh, ph, st) = gpu_run(ctx, case, prec, args.steps, args.warmup, args.inner_tol, not args.no_graphs)
        res[prec] = r
        print(f"[bench] {r['steps_per_s']:.3f} steps/s, iterations {r['iterations'][-1]}", file=sys.stderr,
              flush=True)
        roof_pcg["mixed" if prec else "fp64"] = roofline_pcg(ctx, st, prec)
        print(f"[bench] pressure PCG iteration {roof_pcg['mixed' if prec else 'fp64']['avg_us']} us",
              file=sys.stderr, flush=True)
        if prec == (1 if args.mode != "fp64" else 0):
            roof_ic0 = roofline_ic0(ctx, st, prec)
            roof_spmv = roofline_spmv(ctx, st)
            print(f"[bench] IC0 apply {roof_ic0['avg_us']} us ({roof_ic0['achieved']} GB/s), SpMV {roof_spmv['avg_us']} us",
                  file=sys.stderr, flush=True)
        if prec == 0 and not args.no_cpu:
            print("[bench] cpu baseline (oracle, OpenMP build)", file=sys.stderr, flush=True)
            cpu = cpu_baseline(case, st, mesh, ph)
            # the GPU's own step from the same state, for the CPU step's C_D and C_L to be compared with
            cpu["gpu_cd_cl_same_step"] = [float(v) for v in st.step()]
        del st, ph, mesh
    ceiling = copy_ceiling(ctx)
    print(f"[bench] copy ceiling {ceiling['achieved']} GB/s", file=sys.stderr, flush=True)
    for r in roof_pcg.values():
        r["frac_of_copy_ceiling"] = round(r["achieved"] / ceiling["achieved"], 4)
    head = res.get(1, res.get(0))
    err = None
    if len(res) == 2:
        # the two pipelines after the same steps from the same start (rest, the same particle seeds)
        import numpy as np
        f64, fmx = res[0]["fields"], res[1]["fields"]
        cd = [abs(a - b) / abs(a) for a, b in zip(res[0]["cd_cl"], res[1]["cd_cl"])]
        err = {"after_steps": args.warmup + args.steps,
               "rel_l2": {k: float(np.linalg.norm(fmx[k] - f64[k]) / np.linalg.norm(f64[k])) for k in f64},
               "cd_rel": cd[0], "cl_rel": cd[1],
               "bound": "fields rel L2 <= 1e-8, C_D and C_L <= 1e-6 relative against FP64"}
    for r in res.values():
        r.pop("fields", None)
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
        "dtype": ("f64+f32+f16" if 1 in res else "f64"),
        "data": "synthetic (generated mesh, Philox-seeded particles, flow from rest)",
        "config": {"workload": "cylinder Re=100, scheme B, dt=2e-4", "triangles": int(case.tri.shape[0]),
                   "nodes": int(case.xy.shape[0]), "particles_per_cell": case.ppc,
                   "particles_after_steps": head["particles"], "parallelism": "1 GPU",
                   "precision": "mixed" if 1 in res else "fp64",
                   "mixed_precision": ("velocity solves: FP32 inner Krylov with FP64 refinement; pressure: FP64 "
                                       "PCG with the IC0 factor stored in 16 bits (IEEE half, power-of-two scaled), "
                                       "FP64 sweep arithmetic; particles, assembly, loads FP64"),
                   "pres_factor_bits": 16 if 1 in res else 64},
        "fp64_steps_per_s": round(res[0]["steps_per_s"], 4) if 0 in res else None,
        "mixed_steps_per_s": round(res[1]["steps_per_s"], 4) if 1 in res else None,
        "mixed_over_fp64": round(res[1]["steps_per_s"] / res[0]["steps_per_s"], 3) if len(res) == 2 else None,
        "krylov_iterations_last_step": {("mixed" if k else "fp64"): v["iterations"][-1] for k, v in res.items()},
        "cd_cl_last_step": {("mixed" if k else "fp64"): v["cd_cl"] for k, v in res.items()},
        "mixed_vs_fp64": err,
        "step_graphs": not args.no_graphs,
        "roofline": roof_pcg.get("mixed", roof_pcg.get("fp64")),
        "roofline_pcg_iteration_fp64": roof_pcg.get("fp64"),
        "roofline_ic0_apply": roof_ic0,
        "roofline_spmv": roof_spmv,
        "copy_ceiling": ceiling,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
