// Oracle: one PFEM-2 time step (TEST INFRASTRUCTURE), cases/Cylinder2D/main.cu:793-880 and
// cases/PoiseuilleFlow2D/main.cu:652-736 with the outer-iteration count, scheme and solver choices as
// parameters.
#include "oracle.h"

#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_pfem2 {
    orc_step_config cfg;
    int nv, nt;
    std::vector<double> xy, area, invJ;
    std::vector<int> tri, edge_ids, rowptr, col, body;
    orc_particles* ph;
    std::vector<int> bcu_node, bcp_node;
    std::vector<double> bcu_val[2], bcp_val;
    std::vector<double> u[2], uo[2], vp[2], p, pold;
    std::vector<int> its;
    int nbody = 0;
    // multicolour pressure solve (orc_pfem2_set_coloured_pressure): new -> old rows, colour offsets, the
    // permuted pattern and the old CSR slot of each permuted slot
    bool coloured = false;
    std::vector<int> cperm, cptr, crowptr, ccol, csrc;
};

// Greedy colouring of the pattern's graph in natural order, rows ordered by (colour, index).
static void build_colouring(orc_pfem2* s)
{
    const int n = s->nv;
    std::vector<int> colour(n, -1), mark;
    int ncol = 0;
    for (int i = 0; i < n; ++i) {
        mark.assign(ncol + 1, 0);
        for (int k = s->rowptr[i]; k < s->rowptr[i + 1]; ++k)
            if (s->col[k] != i && colour[s->col[k]] >= 0) mark[colour[s->col[k]]] = 1;
        int c = 0;
        while (c < ncol && mark[c]) ++c;
        colour[i] = c;
        if (c == ncol) ++ncol;
    }
    s->cptr.assign(ncol + 1, 0);
    for (int i = 0; i < n; ++i) s->cptr[colour[i] + 1]++;
    for (int c = 0; c < ncol; ++c) s->cptr[c + 1] += s->cptr[c];
    std::vector<int> fill(s->cptr.begin(), s->cptr.end() - 1), iperm(n);
    s->cperm.resize(n);
    for (int i = 0; i < n; ++i) {
        const int q = fill[colour[i]]++;
        s->cperm[q] = i;
        iperm[i] = q;
    }
    s->crowptr.assign(n + 1, 0);
    s->ccol.clear();
    s->csrc.clear();
    std::vector<std::pair<int, int>> row;
    for (int q = 0; q < n; ++q) {
        const int o = s->cperm[q];
        row.clear();
        for (int k = s->rowptr[o]; k < s->rowptr[o + 1]; ++k) row.emplace_back(iperm[s->col[k]], k);
        std::sort(row.begin(), row.end());
        for (const auto& e : row) { s->ccol.push_back(e.first); s->csrc.push_back(e.second); }
        s->crowptr[q + 1] = (int)s->ccol.size();
    }
}

static int solve_coloured(orc_pfem2* s, const std::vector<double>& val, std::vector<double>& x,
                          const std::vector<double>& b)
{
    if (s->cperm.empty()) build_colouring(s);
    const int n = s->nv, nnz = (int)s->ccol.size();
    std::vector<double> v2(nnz), x2(n), b2(n);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nnz; ++k) v2[k] = val[s->csrc[k]];
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; ++q) { x2[q] = x[s->cperm[q]]; b2[q] = b[s->cperm[q]]; }
    orc_solve_stats st;
    const orc_step_config& c = s->cfg;
    orc_pcg_coloured(n, s->crowptr.data(), s->ccol.data(), v2.data(), x2.data(), b2.data(), c.tol, c.maxit,
                     c.tol_mode, (int)s->cptr.size() - 1, s->cptr.data(), &st);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; ++q) x[s->cperm[q]] = x2[q];
    return st.iterations;
}

static int solve(const orc_pfem2* s, int solver, int precond, const std::vector<double>& val,
                 std::vector<double>& x, const std::vector<double>& b)
{
    orc_solve_stats st;
    const orc_step_config& c = s->cfg;
    if (solver == 0)
        orc_pcg(s->nv, s->rowptr.data(), s->col.data(), val.data(), x.data(), b.data(), c.tol, c.maxit,
                c.tol_mode, precond, &st);
    else
        orc_gmres(s->nv, s->rowptr.data(), s->col.data(), val.data(), x.data(), b.data(), c.tol, c.maxit,
                  c.restart, c.tol_mode, precond, &st);
    return st.iterations;
}

extern "C" {

orc_pfem2* orc_pfem2_create(int nv, const double* xy, int nt, const int* tri, const int* edge_ids,
                            const orc_step_config* cfg, orc_particles* ph,
                            int nbcu, const int* bcu_node, const double* bcu_val0, const double* bcu_val1,
                            int nbcp, const int* bcp_node, const double* bcp_val)
{
    orc_pfem2* s = new orc_pfem2;
    s->cfg = *cfg;
    s->nv = nv; s->nt = nt;
    s->xy.assign(xy, xy + 2 * (size_t)nv);
    s->tri.assign(tri, tri + 3 * (size_t)nt);
    s->edge_ids.assign(edge_ids, edge_ids + 3 * (size_t)nt);
    s->area.resize(nt); s->invJ.resize(4 * (size_t)nt);
    orc_geometry2d(nv, xy, nt, tri, s->area.data(), s->invJ.data());
    const long nnz = orc_csr_count(nv, nt, 3, tri);
    s->rowptr.resize(nv + 1); s->col.resize(nnz);
    orc_csr_fill(nv, nt, 3, tri, s->rowptr.data(), s->col.data());
    s->ph = ph;
    s->bcu_node.assign(bcu_node, bcu_node + nbcu);
    s->bcu_val[0].assign(bcu_val0, bcu_val0 + nbcu);
    s->bcu_val[1].assign(bcu_val1, bcu_val1 + nbcu);
    s->bcp_node.assign(bcp_node, bcp_node + nbcp);
    s->bcp_val.assign(bcp_val, bcp_val + nbcp);
    for (int k = 0; k < 2; ++k) {
        s->u[k].assign(nv, 0.0); s->uo[k].assign(nv, 0.0); s->vp[k].assign(nv, 0.0);
    }
    s->p.assign(nv, 0.0); s->pold.assign(nv, 0.0);
    s->nbody = orc_body_edges(nt, edge_ids, cfg->body_id, nullptr);
    s->body.resize(2 * (size_t)s->nbody);
    orc_body_edges(nt, edge_ids, cfg->body_id, s->body.data());
    return s;
}

void orc_pfem2_destroy(orc_pfem2* s) { delete s; }

void orc_pfem2_set_coloured_pressure(orc_pfem2* s, int on) { s->coloured = on != 0; }

void orc_pfem2_step(orc_pfem2* s, double forces[2])
{
    const orc_step_config& c = s->cfg;
    const orc_pfem2_params* prm = &c.prm;
    const int nv = s->nv, nnz = s->rowptr[nv];
    const bool B = prm->scheme == 1;
    const double* pold = B ? s->pold.data() : nullptr;

    orc_ph_advect(s->ph, s->u[0].data(), s->u[1].data(), prm->dt, c.substeps);
    orc_ph_project(s->ph, s->u[0].data(), s->u[1].data());
    for (int k = 0; k < 2; ++k) s->uo[k] = s->u[k];
    if (B) s->pold = s->p;

    s->its.clear();
    std::vector<double> val0(nnz), val1(nnz), rhs0(nv), rhs1(nv), bcv(s->bcu_node.size());
    const int nbcu = (int)s->bcu_node.size();
    for (int outer = 0; outer < c.outer_iterations; ++outer) {
        // velocity prediction
        std::fill(val0.begin(), val0.end(), 0.0); std::fill(val1.begin(), val1.end(), 0.0);
        std::fill(rhs0.begin(), rhs0.end(), 0.0); std::fill(rhs1.begin(), rhs1.end(), 0.0);
        orc_assemble_prediction(nv, s->xy.data(), s->nt, s->tri.data(), s->area.data(), s->invJ.data(),
                                s->edge_ids.data(), s->rowptr.data(), s->col.data(), val0.data(), val1.data(),
                                rhs0.data(), rhs1.data(), s->u[0].data(), s->u[1].data(), s->uo[0].data(),
                                s->uo[1].data(), pold, prm);
        for (int k = 0; k < 2; ++k) {
            std::vector<double>& val = k == 0 ? val0 : val1;
            std::vector<double>& rhs = k == 0 ? rhs0 : rhs1;
            orc_prediction_bc_values(nv, s->nt, s->tri.data(), s->area.data(), s->invJ.data(), s->p.data(), pold,
                                     nbcu, s->bcu_node.data(), s->bcu_val[k].data(), bcv.data(), k, prm);
            orc_dirichlet_apply(nv, s->rowptr.data(), s->col.data(), val.data(), rhs.data(), nbcu,
                                s->bcu_node.data(), bcv.data(), c.dirichlet_mode);
            s->its.push_back(solve(s, c.pred_solver, c.pred_precond, val, s->vp[k], rhs));
        }
        // pressure
        std::fill(val0.begin(), val0.end(), 0.0); std::fill(rhs0.begin(), rhs0.end(), 0.0);
        orc_assemble_pressure(nv, s->nt, s->tri.data(), s->area.data(), s->invJ.data(), s->rowptr.data(),
                              s->col.data(), val0.data(), rhs0.data(), s->vp[0].data(), s->vp[1].data(), pold, prm);
        orc_dirichlet_apply(nv, s->rowptr.data(), s->col.data(), val0.data(), rhs0.data(),
                            (int)s->bcp_node.size(), s->bcp_node.data(), s->bcp_val.data(), c.dirichlet_mode);
        if (s->coloured && c.pres_solver == 0 && c.pres_precond == 2)
            s->its.push_back(solve_coloured(s, val0, s->p, rhs0));
        else
            s->its.push_back(solve(s, c.pres_solver, c.pres_precond, val0, s->p, rhs0));
        // correction: one matrix, two right-hand sides; each component applies its own BC values to its
        // own copy of the matrix exactly as the reference does with its two matrices
        std::fill(val0.begin(), val0.end(), 0.0);
        std::fill(rhs0.begin(), rhs0.end(), 0.0); std::fill(rhs1.begin(), rhs1.end(), 0.0);
        orc_assemble_correction(nv, s->nt, s->tri.data(), s->area.data(), s->invJ.data(), s->rowptr.data(),
                                s->col.data(), val0.data(), rhs0.data(), rhs1.data(), s->vp[0].data(),
                                s->vp[1].data(), s->p.data(), pold, prm);
        val1 = val0;
        for (int k = 0; k < 2; ++k) {
            std::vector<double>& val = k == 0 ? val0 : val1;
            std::vector<double>& rhs = k == 0 ? rhs0 : rhs1;
            orc_dirichlet_apply(nv, s->rowptr.data(), s->col.data(), val.data(), rhs.data(), nbcu,
                                s->bcu_node.data(), s->bcu_val[k].data(), c.dirichlet_mode);
            s->its.push_back(solve(s, c.corr_solver, c.corr_precond, val, s->u[k], rhs));
        }
    }
    orc_ph_correct(s->ph, s->u[0].data(), s->u[1].data(), s->uo[0].data(), s->uo[1].data());
    double f[4] = {0, 0, 0, 0};
    if (s->nbody > 0)
        orc_body_forces(s->xy.data(), s->tri.data(), s->invJ.data(), s->nbody, s->body.data(), s->u[0].data(),
                        s->u[1].data(), s->p.data(), c.body_cx, c.body_cy, prm, f);
    forces[0] = (f[0] + f[2]) * c.force_coeff;
    forces[1] = (f[1] + f[3]) * c.force_coeff;
}

void orc_pfem2_fields(const orc_pfem2* s, double* u0, double* u1, double* p, double* vp0, double* vp1)
{
    const size_t n = s->nv * sizeof(double);
    if (u0) std::memcpy(u0, s->u[0].data(), n);
    if (u1) std::memcpy(u1, s->u[1].data(), n);
    if (p) std::memcpy(p, s->p.data(), n);
    if (vp0) std::memcpy(vp0, s->vp[0].data(), n);
    if (vp1) std::memcpy(vp1, s->vp[1].data(), n);
}

void orc_pfem2_set_fields(orc_pfem2* s, const double* u0, const double* u1, const double* p,
                          const double* vp0, const double* vp1)
{
    const int n = s->nv;
    if (u0) s->u[0].assign(u0, u0 + n);
    if (u1) s->u[1].assign(u1, u1 + n);
    if (p) s->p.assign(p, p + n);
    if (vp0) s->vp[0].assign(vp0, vp0 + n);
    if (vp1) s->vp[1].assign(vp1, vp1 + n);
}

void orc_pfem2_last_iterations(const orc_pfem2* s, int* its)
{
    for (size_t i = 0; i < s->its.size(); ++i) its[i] = s->its[i];
}

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
