// Oracle: element integrands and their scatter into the CSR (TEST INFRASTRUCTURE).
//
// Each element kernel keeps the loop nest and the operand order of the reference integrand so that
// every local entry rounds identically. The scatter (src/integration/numerical_integrator_2d.cuh:22-43,
// addLocalToGlobal) adds localMatrix(i,j)*area into the CSR slot of (v_i, v_j) and localRhs(i)*area into
// rhs[v_i]. The reference does this with device atomics, so its summation order is unspecified; the
// oracle sums in ascending element order (a sequential loop), which is the order the GPU row-gather
// assembly also uses.
#include "oracle.h"

#include <cmath>
#include <vector>

namespace {

struct P2 { double x, y; };

inline P2 mv(const double* J, double gx, double gy)   // Matrix2x2::operator*(Point2) (cuda_math.cuh:148-156)
{
    return {J[0] * gx + J[1] * gy, J[2] * gx + J[3] * gy};
}

// shapeFuncGrad (numerical_integrator_2d.cuh:8-20)
const double kGrad[3][2] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, -1.0}};

int slot(const int* rowptr, const int* col, int r, int c)
{
    int lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        const int m = (lo + hi) >> 1;
        if (col[m] == c) return m;
        if (col[m] < c) lo = m + 1; else hi = m - 1;
    }
    return -1;
}

void add_local(const int* rowptr, const int* col, double* val, double* rhs, const int* v, double area,
               const double M[3][3], const double R[3])
{
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            const int s = slot(rowptr, col, v[i], v[j]);
            if (s >= 0) {
#pragma omp atomic
                val[s] += M[i][j] * area;
            }
        }
#pragma omp atomic
        rhs[v[i]] += R[i] * area;
    }
}

struct Quad2 { int n; double l[13][3]; double w[13]; };
struct Quad1 { int n; double x[5]; double w[5]; };

Quad2 face(int index)
{
    Quad2 q;
    double buf[13 * 4];
    q.n = orc_quadrature(2, index, buf);
    for (int i = 0; i < q.n; ++i) {
        q.l[i][0] = buf[4 * i]; q.l[i][1] = buf[4 * i + 1]; q.l[i][2] = buf[4 * i + 2]; q.w[i] = buf[4 * i + 3];
    }
    return q;
}

Quad1 edge(int index)
{
    Quad1 q;
    double buf[10];
    q.n = orc_quadrature(1, index, buf);
    for (int i = 0; i < q.n; ++i) { q.x[i] = buf[2 * i]; q.w[i] = buf[2 * i + 1]; }
    return q;
}

// GEOMETRY::transformGlobalToLocal (src/geometry.cuh:14-23)
inline void global_to_local(double px, double py, const double* J, double v3x, double v3y, double L[3])
{
    const double dx = px - v3x, dy = py - v3y;
    L[0] = J[0] * dx + J[2] * dy;
    L[1] = J[1] * dx + J[3] * dy;
    L[2] = 1.0 - L[0] - L[1];
}

}  // namespace

extern "C" {

// cases/Poisson2D/main.cu:28-68 (kIntegrateOverCell). coeff replaces the hard-coded lambda = 0.5 (:36).
// rhs_kind 0: exp(-(x^2 + 10 y^2)) (:24-26); 1: rhs_param * sin(pi x) sin(pi y); 2: constant rhs_param.
void orc_assemble_poisson2d(int nv, const double* xy, int nt, const int* tri, const double* area,
                            const double* invJ, const int* rowptr, const int* col, double* val, double* rhs,
                            double coeff, int rhs_kind, double rhs_param, int quad_index)
{
    (void)nv;
    const Quad2 q = face(quad_index);
    const double PI = 3.14159265358979323846;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nt; ++c) {
        const int* v = tri + 3 * c;
        const double* J = invJ + 4 * c;
        double S[6] = {0, 0, 0, 0, 0, 0};   // SymmetricMatrix3x3 storage (matrix3x3.cuh:45-74)
        double R[3] = {0, 0, 0};
        for (int k = 0; k < q.n; ++k) {
            const double* L = q.l[k];
            const double qx = L[0] * xy[2 * v[0]] + L[1] * xy[2 * v[1]] + L[2] * xy[2 * v[2]];
            const double qy = L[0] * xy[2 * v[0] + 1] + L[1] * xy[2 * v[1] + 1] + L[2] * xy[2 * v[2] + 1];
            for (int i = 0; i < 3; ++i) {
                for (int j = i; j < 3; ++j) {
                    const P2 gi = mv(J, kGrad[i][0], kGrad[i][1]);
                    const P2 gj = mv(J, kGrad[j][0], kGrad[j][1]);
                    const double aux = coeff * (gi.x * gj.x + gi.y * gj.y) * q.w[k];
                    S[i == j ? i : 2 + i + j] += aux;
                }
                double f;
                if (rhs_kind == 0) f = std::exp(-(qx * qx + 10 * qy * qy));
                else if (rhs_kind == 1) f = rhs_param * std::sin(PI * qx) * std::sin(PI * qy);
                else f = rhs_param;
                R[i] += f * L[i] * q.w[k];
            }
        }
        double M[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M[i][j] = S[i == j ? i : 2 + i + j];
        add_local(rowptr, col, val, rhs, v, area[c], M, R);
    }
}

// cases/Poisson3D/main.cu:26-67 with addLocalToGlobal3D (numerical_integrator_3d.cuh:25-46).
// rhs_kind 0: exp(-(x^2 + 10 y^2 + z^2)); 1: rhs_param * sin(pi x) sin(pi y) sin(pi z); 2: constant.
void orc_assemble_poisson3d(int nv, const double* p, int nt, const int* tet, const double* vol,
                            const double* invJ, const int* rowptr, const int* col, double* val, double* rhs,
                            double coeff, int rhs_kind, double rhs_param, int quad_index)
{
    (void)nv;
    double buf[31 * 5];
    const int nq = orc_quadrature(3, quad_index, buf);
    const double PI = 3.14159265358979323846;
    const double g3[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, -1, -1}};
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nt; ++c) {
        const int* v = tet + 4 * c;
        const double* J = invJ + 9 * c;
        double M[4][4] = {{0}};
        double R[4] = {0, 0, 0, 0};
        double G[4][3];
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d)
                G[i][d] = J[3 * d] * g3[i][0] + J[3 * d + 1] * g3[i][1] + J[3 * d + 2] * g3[i][2];
        for (int k = 0; k < nq; ++k) {
            const double* L = buf + 5 * k;
            const double w = L[4];
            double X[3];
            for (int d = 0; d < 3; ++d)
                X[d] = L[0] * p[3 * v[0] + d] + L[1] * p[3 * v[1] + d] + L[2] * p[3 * v[2] + d] + L[3] * p[3 * v[3] + d];
            for (int i = 0; i < 4; ++i) {
                for (int j = 0; j < 4; ++j) {
                    const double aux = coeff * (G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2]) * w;
                    M[i][j] += aux;
                }
                double f;
                if (rhs_kind == 0) f = std::exp(-(X[0] * X[0] + 10 * X[1] * X[1] + X[2] * X[2]));
                else if (rhs_kind == 1) f = rhs_param * std::sin(PI * X[0]) * std::sin(PI * X[1]) * std::sin(PI * X[2]);
                else f = rhs_param;
                R[i] += f * L[i] * w;
            }
        }
        for (int i = 0; i < 4; ++i) {
            for (int j = 0; j < 4; ++j) {
                const int s = slot(rowptr, col, v[i], v[j]);
                if (s >= 0) {
#pragma omp atomic
                    val[s] += M[i][j] * vol[c];
                }
            }
#pragma omp atomic
            rhs[v[i]] += R[i] * vol[c];
        }
    }
}

// cases/Cylinder2D/main.cu:72-167 (kIntegrateVelocityPrediction); the boundary term follows :130-162 for
// the IDs in prm->traction_mask with prm->normal[ID] (Cylinder2D: ID 1 with n = (1,0);
// PoiseuilleFlow2D/main.cu:139-171: every boundary ID with normalVector(ID)).
// p_old != NULL selects Scheme B (:121-125).
void orc_assemble_prediction(int nv, const double* xy, int nt, const int* tri, const double* area,
                             const double* invJ, const int* edge_ids, const int* rowptr, const int* col,
                             double* val0, double* val1, double* rhs0, double* rhs1,
                             const double* u0, const double* u1, const double* uo0, const double* uo1,
                             const double* p_old, const orc_pfem2_params* prm)
{
    (void)nv;
    const Quad2 q = face(prm->face_quad);
    const Quad1 e = edge(prm->edge_quad);
    const double rho = prm->rho, mu = prm->mu, dt = prm->dt;
    const double* uold[2] = {uo0, uo1};
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nt; ++c) {
        const int* v = tri + 3 * c;
        const double* J = invJ + 4 * c;
        double M[2][3][3] = {{{0}}};
        double R[2][3] = {{0}};
        for (int qp = 0; qp < q.n; ++qp) {
            const double* L = q.l[qp];
            const double w = q.w[qp];
            for (int i = 0; i < 3; ++i) {
                const P2 gI = mv(J, kGrad[i][0], kGrad[i][1]);
                const double Ni = L[i];
                for (int j = 0; j < 3; ++j) {
                    const P2 gJ = mv(J, kGrad[j][0], kGrad[j][1]);
                    const double Nj = L[j];
                    double aux = rho * Ni * Nj * w;
                    for (int k = 0; k < 2; ++k) {
                        M[k][i][j] += aux;
                        R[k][i] += aux * uold[k][v[j]];
                    }
                    aux = mu * dt * w;
                    M[0][i][j] += aux * (gI.y * gJ.y + 4.0 / 3.0 * gI.x * gJ.x);
                    M[1][i][j] += aux * (gI.x * gJ.x + 4.0 / 3.0 * gI.y * gJ.y);
                    R[0][i] -= aux * (gI.y * gJ.x - 2.0 / 3.0 * gI.x * gJ.y) * u1[v[j]];
                    R[1][i] -= aux * (gI.x * gJ.y - 2.0 / 3.0 * gI.y * gJ.x) * u0[v[j]];
                    if (p_old) {
                        aux = Ni * p_old[v[j]] * dt * w;
                        R[0][i] -= aux * gJ.x;
                        R[1][i] -= aux * gJ.y;
                    }
                }
            }
        }
        for (int ed = 0; ed < 3; ++ed) {
            const int id = edge_ids[3 * c + ed];
            if (id < 0 || id > 7 || !((prm->traction_mask >> id) & 1)) continue;
            const double nx = prm->normal[id][0], ny = prm->normal[id][1];
            const int a = v[ed], b = v[(ed + 1) % 3];
            const double sx = xy[2 * a], sy = xy[2 * a + 1], ex = xy[2 * b], ey = xy[2 * b + 1];
            const double halfLength = 0.5 * std::sqrt((sx - ex) * (sx - ex) + (sy - ey) * (sy - ey));
            for (int qp = 0; qp < e.n; ++qp) {
                const double cd = e.x[qp];
                const double px = 0.5 * (cd * (ex - sx) + (sx + ex));
                const double py = 0.5 * (cd * (ey - sy) + (sy + ey));
                double L[3];
                global_to_local(px, py, J, xy[2 * v[2]], xy[2 * v[2] + 1], L);
                const double aux = mu * dt * e.w[qp] * halfLength;
                for (int i = 0; i < 3; ++i) {
                    const double Ni = L[i];
                    for (int j = 0; j < 3; ++j) {
                        const P2 gJ = mv(J, kGrad[j][0], kGrad[j][1]);
                        M[0][i][j] -= aux * Ni * ((4.0 / 3.0) * gJ.x * nx + gJ.y * ny);
                        M[1][i][j] -= aux * Ni * (gJ.x * nx + (4.0 / 3.0) * gJ.y * ny);
                        R[0][i] += aux * Ni * u1[v[j]] * ((-2.0 / 3.0) * gJ.y * nx + gJ.x * ny);
                        R[1][i] += aux * Ni * u0[v[j]] * (gJ.y * nx + (-2.0 / 3.0) * gJ.x * ny);
                    }
                }
            }
        }
        add_local(rowptr, col, val0, rhs0, v, area[c], M[0], R[0]);
        add_local(rowptr, col, val1, rhs1, v, area[c], M[1], R[1]);
    }
}

// cases/Cylinder2D/main.cu:169-211 (kIntegratePressureEquation).
void orc_assemble_pressure(int nv, int nt, const int* tri, const double* area, const double* invJ,
                           const int* rowptr, const int* col, double* val, double* rhs,
                           const double* vp0, const double* vp1, const double* p_old,
                           const orc_pfem2_params* prm)
{
    (void)nv;
    const Quad2 q = face(prm->face_quad);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nt; ++c) {
        const int* v = tri + 3 * c;
        const double* J = invJ + 4 * c;
        double M[3][3] = {{0}};
        double R[3] = {0, 0, 0};
        for (int qp = 0; qp < q.n; ++qp) {
            const double* L = q.l[qp];
            const double w = q.w[qp];
            for (int i = 0; i < 3; ++i) {
                const P2 gI = mv(J, kGrad[i][0], kGrad[i][1]);
                const double Ni = L[i];
                const double aux = -prm->rho / prm->dt * Ni * w;
                for (int j = 0; j < 3; ++j) {
                    const P2 gJ = mv(J, kGrad[j][0], kGrad[j][1]);
                    const double vx = vp0[v[j]], vy = vp1[v[j]];
                    const double aux2 = (gI.x * gJ.x + gI.y * gJ.y) * w;
                    M[i][j] += aux2;
                    if (p_old) R[i] += aux2 * p_old[v[j]];
                    R[i] += aux * (gJ.x * vx + gJ.y * vy);
                }
            }
        }
        add_local(rowptr, col, val, rhs, v, area[c], M, R);
    }
}

// cases/Cylinder2D/main.cu:213-259 (kIntegrateVelocityCorrection). The two correction matrices are
// identical (:248-251), so one value array is written.
void orc_assemble_correction(int nv, int nt, const int* tri, const double* area, const double* invJ,
                             const int* rowptr, const int* col, double* val, double* rhs0, double* rhs1,
                             const double* vp0, const double* vp1, const double* p, const double* p_old,
                             const orc_pfem2_params* prm)
{
    (void)nv;
    const Quad2 q = face(prm->face_quad);
    const double* vp[2] = {vp0, vp1};
    std::vector<double> dummy(nv > 0 ? nv : 1);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nt; ++c) {
        const int* v = tri + 3 * c;
        const double* J = invJ + 4 * c;
        double M[3][3] = {{0}};
        double R[2][3] = {{0}};
        for (int qp = 0; qp < q.n; ++qp) {
            const double* L = q.l[qp];
            const double w = q.w[qp];
            for (int i = 0; i < 3; ++i) {
                const double Ni = L[i];
                const double aux2 = prm->dt * Ni * w;
                for (int j = 0; j < 3; ++j) {
                    const P2 gJ = mv(J, kGrad[j][0], kGrad[j][1]);
                    const double Nj = L[j];
                    const double aux = prm->rho * Ni * Nj * w;
                    double pv = p[v[j]];
                    if (p_old) pv -= p_old[v[j]];
                    M[i][j] += aux;   // identical for k = 0, 1
                    R[0][i] += aux * vp[0][v[j]] - aux2 * gJ.x * pv;
                    R[1][i] += aux * vp[1][v[j]] - aux2 * gJ.y * pv;
                }
            }
        }
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                const int s = slot(rowptr, col, v[i], v[j]);
                if (s >= 0) {
#pragma omp atomic
                    val[s] += M[i][j] * area[c];
                }
            }
#pragma omp atomic
            rhs0[v[i]] += R[0][i] * area[c];
#pragma omp atomic
            rhs1[v[i]] += R[1][i] * area[c];
        }
    }
}

// cases/Cylinder2D/main.cu:261-297 (kAccumulatePressureGradient, kFinalizePredictionBC) and :467-479:
// g*_b = g_b + (dt/rho) * sum_cells(area * d_k(p - p_old)) / sum_cells(area), cells in ascending order.
// p_old == NULL is Scheme A (the reference passes a null pressureSolutionOld there).
void orc_prediction_bc_values(int nv, int nt, const int* tri, const double* area, const double* invJ,
                              const double* p, const double* p_old, int nbc, const int* bc_node,
                              const double* bc_src, double* bc_out, int component,
                              const orc_pfem2_params* prm)
{
    std::vector<int> map(nv, -1);
    for (int k = 0; k < nbc; ++k) map[bc_node[k]] = k;
    std::vector<double> num(nbc, 0.0), den(nbc, 0.0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nt; ++c) {
        const int* v = tri + 3 * c;
        if (map[v[0]] == -1 && map[v[1]] == -1 && map[v[2]] == -1) continue;
        double gx = 0.0, gy = 0.0;
        for (int i = 0; i < 3; ++i) {
            const double s = p[v[i]] - (p_old ? p_old[v[i]] : 0.0);
            gx += kGrad[i][0] * s;
            gy += kGrad[i][1] * s;
        }
        const P2 g = mv(invJ + 4 * c, gx, gy);
        const double gc = component == 0 ? g.x : g.y;
        for (int i = 0; i < 3; ++i) {
            const int k = map[v[i]];
            if (k != -1) {
#pragma omp atomic
                num[k] += area[c] * gc;
#pragma omp atomic
                den[k] += area[c];
            }
        }
    }
    for (int k = 0; k < nbc; ++k) bc_out[k] = bc_src[k] + prm->dt / prm->rho * num[k] / den[k];
}

// cases/Cylinder2D/main.cu:299-320 (kCountBodyEdges): (cell, edge) pairs with the body ID, one per
// cell, ascending cell order (the reference order is that of an atomic counter).
int orc_body_edges(int nt, const int* edge_ids, int body_id, int* cell_edge)
{
    int cnt = 0;
    for (int c = 0; c < nt; ++c)
        for (int i = 0; i < 3; ++i)
            if (edge_ids[3 * c + i] == body_id) {
                if (cell_edge) { cell_edge[2 * cnt] = c; cell_edge[2 * cnt + 1] = i; }
                ++cnt;
                break;
            }
    return cnt;
}

// cases/Cylinder2D/main.cu:322-376 (kCalculateBodyForces) and :414-428 (calculateLoads); the centre
// (0.2, 0.2) is a parameter here. out = {Fp_x, Fp_y, Fv_x, Fv_y} summed over the edges in list order.
void orc_body_forces(const double* xy, const int* tri, const double* invJ, int nedges, const int* cell_edge,
                     const double* u0, const double* u1, const double* p, double cx, double cy,
                     const orc_pfem2_params* prm, double out[4])
{
    const Quad1 e = edge(prm->edge_quad);
    double tot[4] = {0, 0, 0, 0};
    for (int k = 0; k < nedges; ++k) {
        const int c = cell_edge[2 * k], ed = cell_edge[2 * k + 1];
        const int* v = tri + 3 * c;
        const double* J = invJ + 4 * c;
        const int a = v[ed], b = v[(ed + 1) % 3];
        const double sx = xy[2 * a], sy = xy[2 * a + 1], ex = xy[2 * b], ey = xy[2 * b + 1];
        const double mx = (sx + ex) * 0.5 - cx, my = (sy + ey) * 0.5 - cy;
        const double inv = 1.0 / std::sqrt(mx * mx + my * my);
        const double nx = mx * inv, ny = my * inv;
        const double halfLength = 0.5 * std::sqrt((sx - ex) * (sx - ex) + (sy - ey) * (sy - ey));
        const double tx = ny, ty = -nx;
        double l[4] = {0, 0, 0, 0};
        for (int qp = 0; qp < e.n; ++qp) {
            double pq = 0.0, dvtdn = 0.0;
            const double cd = e.x[qp];
            const double px = 0.5 * (cd * (ex - sx) + (sx + ex));
            const double py = 0.5 * (cd * (ey - sy) + (sy + ey));
            double L[3];
            global_to_local(px, py, J, xy[2 * v[2]], xy[2 * v[2] + 1], L);
            for (int i = 0; i < 3; ++i) {
                pq += p[v[i]] * L[i];
                const P2 g = mv(J, kGrad[i][0], kGrad[i][1]);
                dvtdn += (u0[v[i]] * tx + u1[v[i]] * ty) * (g.x * nx + g.y * ny);
            }
            const double weight = e.w[qp] * halfLength;
            l[0] -= pq * nx * weight;
            l[1] -= pq * ny * weight;
            l[2] += prm->mu * dvtdn * tx * weight;
            l[3] += prm->mu * dvtdn * ty * weight;
        }
        for (int d = 0; d < 4; ++d) tot[d] += l[d];
    }
    for (int d = 0; d < 4; ++d) out[d] = tot[d];
}

}  // extern "C"
