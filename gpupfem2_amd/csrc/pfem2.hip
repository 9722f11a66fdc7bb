// One PFEM-2 time step: the loop body of cases/Cylinder2D/main.cu:793-880 (and of
// cases/PoiseuilleFlow2D/main.cu:652-736), with outer iterations, scheme and solvers as parameters.
// All systems share the mesh's CSR pattern; the two correction matrices of the reference are one matrix
// with two right-hand sides (they are identical, :248-251), solved together as a multi-RHS PCG in mixed
// precision when the correction solver is Jacobi PCG.
//
// With graphs on (pfem_pfem2_set_graphs), the work between the solves -- the particle step, the
// projection, the three assemblies with their Dirichlet passes, the particle correction and the loads --
// is captured as six HIP graphs on first use and replayed afterwards; only the Krylov solves, whose
// iteration counts are data-dependent, are launched kernel by kernel. The particle step has no host
// round trip (its counts stay on the device), so the graphs hold the same launches as the eager step
// and give the same bits.
#include <cstdio>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"
#include "solver.hpp"

namespace pfem {
void particles_advect(pfem_particles_t* ph, const double* u0, const double* u1, double dt, int substeps);
void particles_advect_prepare(pfem_particles_t* ph);
void particles_advect_launch(pfem_particles_t* ph, int b, const double* u0, const double* u1, double dt, int substeps);
void particles_advect_post(pfem_particles_t* ph);
long particles_generation(const pfem_particles_t* ph);
int particles_current(const pfem_particles_t* ph);
void particles_set_current(pfem_particles_t* ph, int b);
void particles_project(pfem_particles_t* ph, double* u0, double* u1);
void particles_correct(pfem_particles_t* ph, const double* u0, const double* u1, const double* uo0, const double* uo1);
void assemble_prediction(pfem_mesh2d m, pfem_csr a0, pfem_csr a1, double* rhs0, double* rhs1, const double* u0,
                         const double* u1, const double* uo0, const double* uo1, const double* p_old,
                         const pfem_pfem2_params* prm);
void assemble_pressure(pfem_mesh2d m, pfem_csr a, double* rhs, const double* vp0, const double* vp1,
                       const double* p_old, const pfem_pfem2_params* prm);
void assemble_correction(pfem_mesh2d m, pfem_csr a, double* rhs0, double* rhs1, const double* vp0, const double* vp1,
                         const double* p, const double* p_old, const pfem_pfem2_params* prm);
void prediction_bc_values(pfem_mesh2d m, pfem_dirichlet target, pfem_dirichlet source, const double* p,
                          const double* p_old, int component, const pfem_pfem2_params* prm);
void body_forces_prepare(pfem_mesh2d m, int body_id);
void body_forces(pfem_mesh2d m, int body_id, double cx, double cy, const double* u0, const double* u1,
                 const double* p, const pfem_pfem2_params* prm, double* d_out);
}  // namespace pfem

struct pfem_pfem2_t {
    pfem_ctx ctx = nullptr;
    pfem_mesh2d m = nullptr;
    pfem_particles ph = nullptr;
    pfem_step_config cfg{};
    pfem_csr A0 = nullptr, A1 = nullptr, Ap = nullptr, Ac = nullptr;
    pfem_dirichlet bcu[2] = {nullptr, nullptr}, bcpred[2] = {nullptr, nullptr}, bcp = nullptr;
    pfem_precond pc_pred = nullptr, pc_pres = nullptr, pc_corr = nullptr, pc_pred1 = nullptr;
    pfem_solver s_pred = nullptr, s_pres = nullptr, s_corr = nullptr, s_pred1 = nullptr;
    pfem::DevArray<double> f;   // u0 u1 p vp0 vp1 uo0 uo1 pold rhs0 rhs1 (n each) + 4 forces
    pfem::DevArray<double> xsave;   // a single-pass solve's start vector, kept for a per-level re-solve
    int sweep_fallbacks = 0;
    int n = 0;
    std::vector<int> its;
    bool timing = false;
    double phase_ms[8] = {0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    // graph replay of the work between the solves: segment k, buffer variant v (the particle buffers
    // alternate), captured for the particle buffers' generation `key`
    struct Segment {
        hipGraphExec_t exec = nullptr;
        long key = -1;
    };
    static constexpr int kSegments = 6;
    bool graphs = false;
    long steps = 0;   // steps taken (graphs are captured from the second one, after the lazy set-ups)
    Segment seg[kSegments][2];
    double* field(int k) { return f.p + (size_t)k * n; }
    void drop_graphs()
    {
        for (auto& row : seg)
            for (auto& g : row) {
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
                g = Segment{};
            }
    }
    ~pfem_pfem2_t()
    {
        drop_graphs();
        pfem_solver_destroy(s_pred); pfem_solver_destroy(s_pres); pfem_solver_destroy(s_corr);
        pfem_solver_destroy(s_pred1);
        pfem_precond_destroy(pc_pred); pfem_precond_destroy(pc_pres); pfem_precond_destroy(pc_corr);
        pfem_precond_destroy(pc_pred1);
        for (int k = 0; k < 2; ++k) { pfem_dirichlet_destroy(bcu[k]); pfem_dirichlet_destroy(bcpred[k]); }
        pfem_dirichlet_destroy(bcp);
        pfem_csr_destroy(A0); pfem_csr_destroy(A1); pfem_csr_destroy(Ap); pfem_csr_destroy(Ac);
        if (ev[0]) (void)hipEventDestroy(ev[0]);
        if (ev[1]) (void)hipEventDestroy(ev[1]);
    }
};

using namespace pfem;

namespace {
enum { U0, U1, P, VP0, VP1, UO0, UO1, POLD, RHS0, RHS1, NF };

void check(int rc, pfem_ctx ctx)
{
    if (rc != PFEM_OK) throw Error(rc, ctx->err);
}

struct Phase {
    pfem_pfem2_t* s;
    int k;
    Phase(pfem_pfem2_t* s_, int k_) : s(s_), k(k_)
    {
        if (s->timing) PFEM_HIP(hipEventRecord(s->ev[0], s->ctx->stream));
    }
    ~Phase()
    {
        if (!s->timing) return;
        float ms = 0.f;
        (void)hipEventRecord(s->ev[1], s->ctx->stream);
        (void)hipEventSynchronize(s->ev[1]);
        (void)hipEventElapsedTime(&ms, s->ev[0], s->ev[1]);
        s->phase_ms[k] += ms;
    }
};

// The addresses a step's launches use change only when the particle buffers or the mesh's edge IDs or
// body cache are reallocated: a graph captured under another key is captured again.
long graph_key(const pfem_pfem2_t* s)
{
    return particles_generation(s->ph) * 1000003L + s->m->generation;
}

// Runs body (launches only) directly, or from segment k's graph (variant v, captured for key on first use
// and whenever the key changes) when graphs are on and a first step has made the lazy set-ups.
template <typename F>
void run_segment(pfem_pfem2_t* s, int k, int v, long key, F&& body)
{
    if (!s->graphs || s->steps == 0) {
        body();
        return;
    }
    hipStream_t st = s->ctx->stream;
    pfem_pfem2_t::Segment& g = s->seg[k][v];
    if (!g.exec || g.key != key) {
        if (g.exec) PFEM_HIP(hipGraphExecDestroy(g.exec));
        g = pfem_pfem2_t::Segment{};
        hipGraph_t graph = nullptr;
        PFEM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        PFEM_HIP(hipStreamEndCapture(st, &graph));
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PFEM_HIP(e);
        g.key = key;
    }
    PFEM_HIP(hipGraphLaunch(g.exec, st));
}

void make_solver(pfem_pfem2_t* s, pfem_csr a, int kind, int pc_kind, int ordering, pfem_precond* pc, pfem_solver* sv,
                 int factor_bits = 32)
{
    const pfem_step_config& c = s->cfg;
    check(pfem_precond_create(a, pc_kind, ordering, pc), s->ctx);
    if (c.precision == PFEM_PRECISION_MIXED && factor_bits == 16 && pc_kind == PFEM_PRECOND_IC0 &&
        ordering != PFEM_ORDER_NATURAL)
        check(pfem_precond_set_factor_bits(*pc, 16), s->ctx);
    check(pfem_solver_create(a, kind, c.tol, c.maxit, c.restart, c.tol_mode, *pc, sv), s->ctx);
    if (c.precision == PFEM_PRECISION_MIXED) {
        const bool factor = pc_kind == PFEM_PRECOND_IC0 || pc_kind == PFEM_PRECOND_ILU0;
        check(pfem_solver_set_precision(*sv, factor ? PFEM_PRECISION_MIXED_PRECOND : PFEM_PRECISION_MIXED,
                                        c.inner_tol, 0),
              s->ctx);
    }
}

// A solve whose IC0/ILU0 preconditioner runs single-pass sweeps is solved again from the same start with
// per-level sweeps (FP64 factor) when a sweep's wait exceeded its spin limit, instead of failing the
// step; the fallback is counted and reported on stderr.
void solve(pfem_pfem2_t* s, pfem_solver sv, pfem_csr a, double* x, const double* b)
{
    pfem_solve_stats st{};
    pfem_precond_t* pc = sv->pc;
    const bool sweeps = pc && (pc->kind == PFEM_PRECOND_IC0 || pc->kind == PFEM_PRECOND_ILU0) && pc->sweep;
    if (sweeps) {
        if (s->xsave.n < (size_t)s->n) s->xsave.alloc(s->n);
        launch_copy(C(s->ctx), s->xsave.p, x, s->n);
    }
    const int failures = pc ? pc->sweep_failures : 0;
    const int rc = pfem_solver_solve(sv, a, x, b, &st);
    if (rc != PFEM_OK && sweeps && pc->sweep_failures > failures) {
        ++s->sweep_fallbacks;
        std::fprintf(stderr, "pfem_pfem2_step: a single-pass sweep wait exceeded its spin limit; solving again "
                             "with per-level sweeps (fallback %d)\n", s->sweep_fallbacks);
        launch_copy(C(s->ctx), x, s->xsave.p, s->n);
        const int precision = sv->precision;
        check(pfem_precond_set_sweep(pc, 0), s->ctx);
        if (precision == PFEM_PRECISION_MIXED_PRECOND) check(pfem_solver_set_precision(sv, PFEM_PRECISION_FP64, 0, 0), s->ctx);
        const int rc2 = pfem_solver_solve(sv, a, x, b, &st);
        check(pfem_precond_set_sweep(pc, 1), s->ctx);
        if (precision == PFEM_PRECISION_MIXED_PRECOND) check(pfem_solver_set_precision(sv, precision, 0, 0), s->ctx);
        check(rc2, s->ctx);
    } else {
        check(rc, s->ctx);
    }
    s->its.push_back(st.iterations);
}
}  // namespace

extern "C" {

int pfem_pfem2_create(pfem_mesh2d m, pfem_particles ph, const pfem_step_config* cfg, int nbcu, const int* bcu_node,
                      const double* bcu_val0, const double* bcu_val1, int nbcp, const int* bcp_node,
                      const double* bcp_val, pfem_pfem2* out)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && ph && cfg && out, "null argument");
        auto s = std::make_unique<pfem_pfem2_t>();
        s->ctx = m->ctx;
        s->m = m;
        s->ph = ph;
        s->cfg = *cfg;
        s->n = m->nv;
        Context* c = C(m->ctx);
        s->f.alloc((size_t)NF * s->n + 4);
        zero_async(c, s->f.p, s->f.bytes());
        check(pfem_csr_create_from_mesh2d(m, &s->A0), s->ctx);
        check(pfem_csr_create_like(s->A0, &s->A1), s->ctx);
        check(pfem_csr_create_like(s->A0, &s->Ap), s->ctx);
        check(pfem_csr_create_like(s->A0, &s->Ac), s->ctx);
        check(pfem_dirichlet_create(s->ctx, s->n, nbcu, bcu_node, bcu_val0, &s->bcu[0]), s->ctx);
        check(pfem_dirichlet_create(s->ctx, s->n, nbcu, bcu_node, bcu_val1, &s->bcu[1]), s->ctx);
        check(pfem_dirichlet_create(s->ctx, s->n, nbcu, bcu_node, bcu_val0, &s->bcpred[0]), s->ctx);
        check(pfem_dirichlet_create(s->ctx, s->n, nbcu, bcu_node, bcu_val1, &s->bcpred[1]), s->ctx);
        check(pfem_dirichlet_create(s->ctx, s->n, nbcp, bcp_node, bcp_val, &s->bcp), s->ctx);
        make_solver(s.get(), s->A0, cfg->pred_solver, cfg->pred_precond, PFEM_ORDER_NATURAL, &s->pc_pred, &s->s_pred);
        // the prediction systems are mass-dominated (rho M + mu dt K): one classical Gram-Schmidt pass per
        // Arnoldi step keeps their iteration counts (cylinder, 2M triangles: prediction 9.0 -> 5.9 ms per
        // step in FP64); pred_gs_passes = 2 orthogonalises twice
        PFEM_REQUIRE(cfg->pred_gs_passes >= 0 && cfg->pred_gs_passes <= 2, "pred_gs_passes must be 0, 1 or 2");
        if (cfg->pred_solver == PFEM_SOLVER_GMRES)
            check(pfem_solver_set_option(s->s_pred, PFEM_SOLVER_OPT_GS_PASSES, cfg->pred_gs_passes == 2 ? 2 : 1),
                  s->ctx);
        if (cfg->pred_solver == PFEM_SOLVER_GMRES && cfg->precision == PFEM_PRECISION_MIXED) {
            // the second component's own solver: in mixed precision the two prediction systems are solved
            // as a pair (pfem_solver_solve_pair: lockstep Arnoldi steps, one SpMV launch streaming the
            // pattern once). In FP64 they run one after the other: there the pair measured slower
            // (cylinder, 2M triangles: 6.75 against 6.20 ms per step; mixed: 5.44 against 5.77 ms).
            make_solver(s.get(), s->A1, cfg->pred_solver, cfg->pred_precond, PFEM_ORDER_NATURAL, &s->pc_pred1,
                        &s->s_pred1);
            check(pfem_solver_set_option(s->s_pred1, PFEM_SOLVER_OPT_GS_PASSES, cfg->pred_gs_passes == 2 ? 2 : 1),
                  s->ctx);
        }
        PFEM_REQUIRE(cfg->pres_factor_bits == 0 || cfg->pres_factor_bits == 16 || cfg->pres_factor_bits == 32,
                     "pres_factor_bits must be 0, 16 or 32");
        make_solver(s.get(), s->Ap, cfg->pres_solver, cfg->pres_precond, cfg->pres_ordering, &s->pc_pres, &s->s_pres,
                    cfg->pres_factor_bits ? cfg->pres_factor_bits : 32);
        make_solver(s.get(), s->Ac, cfg->corr_solver, cfg->corr_precond, PFEM_ORDER_NATURAL, &s->pc_corr, &s->s_corr);
        PFEM_HIP(hipEventCreate(&s->ev[0]));
        PFEM_HIP(hipEventCreate(&s->ev[1]));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        *out = s.release();
    });
}

int pfem_pfem2_destroy(pfem_pfem2 s)
{
    if (!s) return PFEM_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    return PFEM_OK;
}

int pfem_pfem2_step(pfem_pfem2 s, double forces[2])
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s, "null argument");
        Context* c = C(s->ctx);
        const pfem_step_config& cfg = s->cfg;
        const pfem_pfem2_params* prm = &cfg.prm;
        const int n = s->n;
        const bool B = prm->scheme == 1;
        double* pold = B ? s->field(POLD) : nullptr;
        s->its.clear();
        PFEM_REQUIRE(cfg.substeps > 0, "substeps must be positive");
        {
            Phase ph(s, 0);
            particles_advect_prepare(s->ph);
            const int b = particles_current(s->ph);
            run_segment(s, 0, b, graph_key(s), [&] {
                particles_advect_launch(s->ph, b, s->field(U0), s->field(U1), prm->dt, cfg.substeps);
            });
            particles_set_current(s->ph, 1 - b);
            particles_advect_post(s->ph);
        }
        {
            Phase ph(s, 1);
            run_segment(s, 1, particles_current(s->ph), graph_key(s), [&] {
                particles_project(s->ph, s->field(U0), s->field(U1));
                launch_copy(c, s->field(UO0), s->field(U0), n);
                launch_copy(c, s->field(UO1), s->field(U1), n);
                if (B) launch_copy(c, s->field(POLD), s->field(P), n);
            });
        }
        for (int outer = 0; outer < cfg.outer_iterations; ++outer) {
            {
                Phase ph(s, 2);
                run_segment(s, 2, 0, graph_key(s), [&] {
                    assemble_prediction(s->m, s->A0, s->A1, s->field(RHS0), s->field(RHS1), s->field(U0),
                                        s->field(U1), s->field(UO0), s->field(UO1), pold, prm);
                    for (int k = 0; k < 2; ++k) {
                        prediction_bc_values(s->m, s->bcpred[k], s->bcu[k], s->field(P), pold, k, prm);
                        dirichlet_apply(c, s->bcpred[k], nullptr, k == 0 ? s->A0 : s->A1, s->field(RHS0 + k),
                                        nullptr, cfg.dirichlet_mode);
                    }
                });
            }
            {
                Phase ph(s, 3);
                if (s->s_pred1) {
                    pfem_solve_stats st[2] = {};
                    check(pfem_solver_solve_pair(s->s_pred, s->A0, s->field(VP0), s->field(RHS0), s->s_pred1, s->A1,
                                                 s->field(VP1), s->field(RHS1), st),
                          s->ctx);
                    s->its.push_back(st[0].iterations);
                    s->its.push_back(st[1].iterations);
                } else {
                    solve(s, s->s_pred, s->A0, s->field(VP0), s->field(RHS0));
                    solve(s, s->s_pred, s->A1, s->field(VP1), s->field(RHS1));
                }
            }
            {
                Phase ph(s, 2);
                run_segment(s, 3, 0, graph_key(s), [&] {
                    assemble_pressure(s->m, s->Ap, s->field(RHS0), s->field(VP0), s->field(VP1), pold, prm);
                    dirichlet_apply(c, s->bcp, nullptr, s->Ap, s->field(RHS0), nullptr, cfg.dirichlet_mode);
                });
            }
            {
                Phase ph(s, 4);
                solve(s, s->s_pres, s->Ap, s->field(P), s->field(RHS0));
            }
            {
                Phase ph(s, 2);
                run_segment(s, 4, 0, graph_key(s), [&] {
                    assemble_correction(s->m, s->Ac, s->field(RHS0), s->field(RHS1), s->field(VP0), s->field(VP1),
                                        s->field(P), pold, prm);
                    dirichlet_apply(c, s->bcu[0], s->bcu[1], s->Ac, s->field(RHS0), s->field(RHS1),
                                    cfg.dirichlet_mode);
                });
            }
            {
                Phase ph(s, 5);
                if (cfg.precision == PFEM_PRECISION_MIXED && cfg.corr_solver == PFEM_SOLVER_CG &&
                    (cfg.corr_precond == PFEM_PRECOND_JACOBI || cfg.corr_precond == PFEM_PRECOND_NONE)) {
                    // both components at once: one matrix, two right-hand sides (U1 = U0 + n, RHS1 = RHS0 + n),
                    // each column running the PCG recurrence of a single solve; in mixed precision by reliable
                    // updates, which keep the FP64 iteration counts (the refinement of a single mixed solve
                    // needs about 30% more). In FP64 the two fused single solves are quicker for these short
                    // solves (about 20 iterations).
                    pfem_solve_stats st[2] = {};
                    check(pfem_solver_solve_multi(s->s_corr, s->Ac, 2, s->field(U0), s->field(RHS0), st), s->ctx);
                    s->its.push_back(st[0].iterations);
                    s->its.push_back(st[1].iterations);
                } else {
                    solve(s, s->s_corr, s->Ac, s->field(U0), s->field(RHS0));
                    solve(s, s->s_corr, s->Ac, s->field(U1), s->field(RHS1));
                }
            }
        }
        {
            Phase ph(s, 6);
            double* fd = s->f.p + (size_t)NF * n;
            if (cfg.body_id >= 0) body_forces_prepare(s->m, cfg.body_id);   // outside any capture
            run_segment(s, 5, particles_current(s->ph), graph_key(s), [&] {
                particles_correct(s->ph, s->field(U0), s->field(U1), s->field(UO0), s->field(UO1));
                if (cfg.body_id >= 0)
                    body_forces(s->m, cfg.body_id, cfg.body_cx, cfg.body_cy, s->field(U0), s->field(U1), s->field(P),
                                prm, fd);
            });
            if (cfg.body_id >= 0) {
                if (forces) {
                    double h[4];
                    d2h_sync(c, h, fd, sizeof(h));
                    forces[0] = (h[0] + h[2]) * cfg.force_coeff;
                    forces[1] = (h[1] + h[3]) * cfg.force_coeff;
                }
            } else if (forces) {
                forces[0] = forces[1] = 0.0;
            }
        }
        PFEM_HIP(hipGetLastError());
        ++s->steps;
    });
}

int pfem_pfem2_set_graphs(pfem_pfem2 s, int on)
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s, "null argument");
        s->graphs = on != 0;
        if (!s->graphs) s->drop_graphs();
    });
}

int pfem_pfem2_field(pfem_pfem2 s, int which, double** d_ptr)
{
    if (!s || !d_ptr) return PFEM_EINVAL;
    static const int map[8] = {U0, U1, P, VP0, VP1, UO0, UO1, POLD};
    if (which < 0 || which > 7) return PFEM_EINVAL;
    *d_ptr = s->field(map[which]);
    return PFEM_OK;
}

int pfem_pfem2_last_iterations(pfem_pfem2 s, int* its, int cap)
{
    if (!s || !its) return PFEM_EINVAL;
    const int k = std::min<int>(cap, (int)s->its.size());
    for (int i = 0; i < k; ++i) its[i] = s->its[i];
    return k;
}

int pfem_pfem2_enable_timing(pfem_pfem2 s, int on)
{
    if (!s) return PFEM_EINVAL;
    s->timing = on != 0;
    std::memset(s->phase_ms, 0, sizeof(s->phase_ms));
    return PFEM_OK;
}

int pfem_pfem2_phase_times(pfem_pfem2 s, double* ms, int cap)
{
    if (!s || !ms) return PFEM_EINVAL;
    for (int k = 0; k < std::min(cap, 8); ++k) ms[k] = s->phase_ms[k];
    return PFEM_OK;
}

int pfem_pfem2_sweep_fallbacks(pfem_pfem2 s)
{
    return s ? s->sweep_fallbacks : PFEM_EINVAL;
}

int pfem_pfem2_matrix(pfem_pfem2 s, int which, pfem_csr* out)
{
    if (!s || !out || which < 0 || which > 3) return PFEM_EINVAL;
    pfem_csr m[4] = {s->A0, s->A1, s->Ap, s->Ac};
    *out = m[which];
    return PFEM_OK;
}

int pfem_pfem2_precond(pfem_pfem2 s, int which, pfem_precond* out)
{
    if (!s || !out || which < 0 || which > 2) return PFEM_EINVAL;
    pfem_precond p[3] = {s->pc_pred, s->pc_pres, s->pc_corr};
    *out = p[which];
    return PFEM_OK;
}

}  // extern "C"
