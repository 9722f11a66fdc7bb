/*
 * oracle.h -- CPU restatement of the gpuPfem2 PFEM-2 hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity oracle and the CPU baseline. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every function restates the reference algorithm sequentially in FP64, built with
 * -ffp-contract=off so that each expression rounds exactly as written. Each function cites the
 * reference file:line it follows. Where the reference order of floating-point accumulation is
 * undefined (device atomics), the oracle fixes one order and states it; the GPU library uses
 * the same order so the two agree bit-for-bit wherever the arithmetic is identical.
 *
 * Pinning (what checks each part besides the GPU comparison: closed forms, invariants, brute force,
 * known-answer vectors and the reference's fixtures) is done per function in tests/test_oracle_*.py;
 * every function here has such a check.
 */
#ifndef PFEM2_ORACLE_H
#define PFEM2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- mesh / geometry (src/mesh_2d.cu, src/mesh_3d.cu) ---------------- */
void orc_geometry2d(int nv, const double* xy, int nt, const int* tri, double* area, double* invJ);
void orc_geometry3d(int nv, const double* xyz, int nt, const int* tet, double* vol, double* invJ);
long orc_csr_count(int nv, int nt, int nper, const int* cells);
void orc_csr_fill(int nv, int nt, int nper, const int* cells, int* rowptr, int* col);
long orc_neighbors_bruteforce_count(int nt, const int* tri);
void orc_neighbors_bruteforce_fill(int nt, const int* tri, int* offsets, int* idx);
long orc_face_neighbors3d_count(int nt, const int* tet);
void orc_face_neighbors3d_fill(int nt, const int* tet, int* offsets, int* idx);

/* ---------------- quadrature (src/integration/quadrature_formula_*.cu) ---------------- */
/* kind 1 = edge (coord, weight), 2 = face (l1, l2, l3, weight), 3 = cell (l1..l4, weight).
 * Returns the number of points; out receives npts*(stride) doubles. */
int orc_quadrature(int kind, int index, double* out);

/* ---------------- linear algebra (src/linear_algebra, src/Dirichlet_bcs.cu) ---------------- */
void orc_spmv(int n, const int* rowptr, const int* col, const double* val, const double* x, double* y);
void orc_dirichlet_apply(int n, const int* rowptr, const int* col, double* val, double* rhs,
                         int nbc, const int* bc_node, const double* bc_val, int mode);
void orc_jacobi_setup(int n, const int* rowptr, const int* col, const double* val, double* invd);
int  orc_ic0(int n, const int* rowptr, const int* col, const double* val, double* lval);
int  orc_ilu0(int n, const int* rowptr, const int* col, const double* val, double* luval);
void orc_ic0_apply(int n, const int* rowptr, const int* col, const double* lval, const double* r, double* z);
void orc_ilu0_apply(int n, const int* rowptr, const int* col, const double* luval, const double* r, double* z);

typedef struct {
    int converged;
    int iterations;
    double final_residual;
    int nan_detected;
} orc_solve_stats;

/* precond: 0 none, 1 Jacobi, 2 IC0, 3 ILU0; tol_mode: 0 reference absolute, 1 relative */
int orc_pcg(int n, const int* rowptr, const int* col, const double* val, double* x, const double* b,
            double tol, int maxit, int tol_mode, int precond, orc_solve_stats* st);
/* PCG + IC0 on a multicolour-ordered matrix (rows of colour c: [cptr[c], cptr[c+1]), no entries between two
 * rows of one colour), colours processed in sequence and the rows of one colour in parallel (OpenMP build) */
int orc_pcg_coloured(int n, const int* rowptr, const int* col, const double* val, double* x, const double* b,
                     double tol, int maxit, int tol_mode, int ncol, const int* cptr, orc_solve_stats* st);
int orc_gmres(int n, const int* rowptr, const int* col, const double* val, double* x, const double* b,
              double tol, int maxit, int restart, int tol_mode, int precond, orc_solve_stats* st);

/* ---------------- assembly (cases/Poisson2D, cases/Poisson3D, cases/Cylinder2D, cases/PoiseuilleFlow2D) */
void orc_assemble_poisson2d(int nv, const double* xy, int nt, const int* tri, const double* area,
                            const double* invJ, const int* rowptr, const int* col, double* val, double* rhs,
                            double coeff, int rhs_kind, double rhs_param, int quad_index);
void orc_assemble_poisson3d(int nv, const double* xyz, int nt, const int* tet, const double* vol,
                            const double* invJ, const int* rowptr, const int* col, double* val, double* rhs,
                            double coeff, int rhs_kind, double rhs_param, int quad_index);

typedef struct {
    double rho, mu, dt;
    int scheme;            /* 0 = Scheme A, 1 = Scheme B */
    int face_quad, edge_quad;
    int traction_mask;     /* bit b set: boundary ID b carries the outflow traction term */
    double normal[8][2];   /* outward normal per boundary ID */
} orc_pfem2_params;

void orc_assemble_prediction(int nv, const double* xy, int nt, const int* tri, const double* area,
                             const double* invJ, const int* edge_ids, const int* rowptr, const int* col,
                             double* val0, double* val1, double* rhs0, double* rhs1,
                             const double* u0, const double* u1, const double* uo0, const double* uo1,
                             const double* p_old, const orc_pfem2_params* prm);
void orc_assemble_pressure(int nv, int nt, const int* tri, const double* area, const double* invJ,
                           const int* rowptr, const int* col, double* val, double* rhs,
                           const double* vp0, const double* vp1, const double* p_old,
                           const orc_pfem2_params* prm);
void orc_assemble_correction(int nv, int nt, const int* tri, const double* area, const double* invJ,
                             const int* rowptr, const int* col, double* val, double* rhs0, double* rhs1,
                             const double* vp0, const double* vp1, const double* p, const double* p_old,
                             const orc_pfem2_params* prm);
void orc_prediction_bc_values(int nv, int nt, const int* tri, const double* area, const double* invJ,
                              const double* p, const double* p_old, int nbc, const int* bc_node,
                              const double* bc_src, double* bc_out, int component,
                              const orc_pfem2_params* prm);
int  orc_body_edges(int nt, const int* edge_ids, int body_id, int* cell_edge /* 2 per edge, may be NULL */);
void orc_body_forces(const double* xy, const int* tri, const double* invJ, int nedges, const int* cell_edge,
                     const double* u0, const double* u1, const double* p, double cx, double cy,
                     const orc_pfem2_params* prm, double out[4]);

/* ---------------- particles (src/particles) ---------------- */
typedef struct orc_particles orc_particles;
orc_particles* orc_ph_create(int nv, const double* xy, int nt, const int* tri, const double* invJ,
                             const int* nbr_off, const int* nbr_idx, int level);
void orc_ph_destroy(orc_particles* ph);
void orc_ph_seed(orc_particles* ph, int mode, int ppc, uint64_t seed);
void orc_ph_init_velocity(orc_particles* ph, const double* u0, const double* u1);
void orc_ph_advect(orc_particles* ph, const double* u0, const double* u1, double dt, int substeps);
void orc_ph_project(orc_particles* ph, double* u0, double* u1);
void orc_ph_correct(orc_particles* ph, const double* u0, const double* u1, const double* uo0, const double* uo1);
long orc_ph_count(const orc_particles* ph);
void orc_ph_get(const orc_particles* ph, int* cell, double* lam /*3 per*/, double* pos /*2 per*/,
                double* vel /*2 per*/, uint32_t* id);
void orc_ph_set(orc_particles* ph, long count, const int* cell, const double* lam, const double* pos,
                const double* vel, const uint32_t* id, uint32_t next_id);
void orc_ph_stats(const orc_particles* ph, long* deleted_last, long* added_last);
void orc_subcell_centers(int level, double* out /* 3*level*level */);
int  orc_determine_subcell(int level, double l1, double l2, double l3);
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---------------- whole PFEM-2 time step (cases/Cylinder2D/main.cu:793-880) ---------------- */
typedef struct {
    orc_pfem2_params prm;
    int outer_iterations;
    int substeps;
    int pred_solver, pred_precond, pres_solver, pres_precond, corr_solver, corr_precond; /* solver 0 CG, 1 GMRES */
    double tol;
    int tol_mode, maxit, restart;
    int dirichlet_mode;
    int body_id;
    double body_cx, body_cy, force_coeff;
} orc_step_config;

typedef struct orc_pfem2 orc_pfem2;
orc_pfem2* orc_pfem2_create(int nv, const double* xy, int nt, const int* tri, const int* edge_ids,
                            const orc_step_config* cfg, orc_particles* ph,
                            int nbcu, const int* bcu_node, const double* bcu_val0, const double* bcu_val1,
                            int nbcp, const int* bcp_node, const double* bcp_val);
void orc_pfem2_destroy(orc_pfem2* s);
void orc_pfem2_step(orc_pfem2* s, double forces[2]);
/* on != 0: an IC0-PCG pressure solve runs on the pressure system permuted by a greedy colouring of the
 * mesh's CSR pattern (natural visiting order), with orc_pcg_coloured (the CPU baseline's form) */
void orc_pfem2_set_coloured_pressure(orc_pfem2* s, int on);
void orc_pfem2_fields(const orc_pfem2* s, double* u0, double* u1, double* p, double* vp0, double* vp1);
void orc_pfem2_set_fields(orc_pfem2* s, const double* u0, const double* u1, const double* p,
                          const double* vp0, const double* vp1);
void orc_pfem2_last_iterations(const orc_pfem2* s, int* its /* 5 per outer iteration */);

/* timing helper for the CPU baseline */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
