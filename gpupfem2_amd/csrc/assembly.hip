// Element assembly into the CSR by row-owner gather (no atomics, no zeroing pass).
//
// The reference integrand kernels (cases/*/main.cu) run one thread per element and scatter 9-16 local
// entries with binary searches and FP64 atomicAdd (src/integration/numerical_integrator_2d.cuh:22-43).
// Here one lane owns one CSR row r: it walks the cells incident to node r in ascending order
// (precomputed node -> (cell, local vertex) list), evaluates row `local` of that cell's local matrix
// and right-hand side with the reference's operand order, and adds them into its own row through a
// precomputed slot map. Every row is written by exactly one lane, so there are no races, the result
// is deterministic, and each entry sums its element contributions in ascending element order -- the
// order of a sequential element loop.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "solver.hpp"

namespace pfem {

// Gauss rules of src/integration/quadrature_formula_{1d,2d,3d}.cu (literals kept digit for digit).
int quadrature_table(int kind, int index, double* out)
{
    static const double e[5][10] = {
        {0.0, 2.0},
        {-0.577350269189626, 1.0, 0.577350269189626, 1.0},
        {-0.774596669241483, 0.555555555555556, 0.0, 0.888888888888889, 0.774596669241483, 0.555555555555556},
        {-0.861136311594053, 0.347854845137454, -0.339981043584856, 0.652145154862546, 0.339981043584856,
         0.652145154862546, 0.861136311594053, 0.347854845137454},
        {-0.906179845938664, 0.236926885056189, -0.538469310105683, 0.478628670499366, 0.0, 0.568888888888889,
         0.538469310105683, 0.478628670499366, 0.906179845938664, 0.236926885056189}};
    static const int ne[5] = {1, 2, 3, 4, 5};
    static const double f[8][39] = {
        {0.3333333333333333, 0.3333333333333333, 1.0},
        {0.1666666666666667, 0.1666666666666667, 0.3333333333333333, 0.6666666666666667, 0.1666666666666667,
         0.3333333333333333, 0.1666666666666667, 0.6666666666666667, 0.3333333333333333},
        {0.3333333333333333, 0.3333333333333333, -0.5625, 0.6, 0.2, 0.5208333333333333, 0.2, 0.6, 0.5208333333333333,
         0.2, 0.2, 0.5208333333333333},
        {0.816847572980459, 0.091576213509771, 0.109951743655322, 0.091576213509771, 0.816847572980459,
         0.109951743655322, 0.091576213509771, 0.091576213509771, 0.109951743655322, 0.445948490915965,
         0.108103018168070, 0.223381589678011, 0.108103018168070, 0.445948490915965, 0.223381589678011,
         0.445948490915965, 0.445948490915965, 0.223381589678011},
        {0.3333333333333333, 0.3333333333333333, 0.225, 0.101286507323456, 0.797426985353087, 0.125939180544827,
         0.797426985353087, 0.101286507323456, 0.125939180544827, 0.101286507323456, 0.101286507323456,
         0.125939180544827, 0.470142064105115, 0.059715871789770, 0.132394152788506, 0.059715871789770,
         0.470142064105115, 0.132394152788506, 0.470142064105115, 0.470142064105115, 0.132394152788506},
        {0.437525248383384, 0.124949503233232, 0.205950504760887, 0.124949503233232, 0.437525248383384,
         0.205950504760887, 0.437525248383384, 0.437525248383384, 0.205950504760887, 0.797112651860071,
         0.165409927389841, 0.063691414286223, 0.165409927389841, 0.797112651860071, 0.063691414286223,
         0.797112651860071, 0.037477420750088, 0.063691414286223, 0.037477420750088, 0.797112651860071,
         0.063691414286223, 0.165409927389841, 0.037477420750088, 0.063691414286223, 0.037477420750088,
         0.165409927389841, 0.063691414286223},
        {0.873821971016996, 0.063089014491502, 0.050844906370207, 0.063089014491502, 0.873821971016996,
         0.050844906370207, 0.063089014491502, 0.063089014491502, 0.050844906370207, 0.501426509658179,
         0.249286745170910, 0.116786275726379, 0.249286745170910, 0.501426509658179, 0.116786275726379,
         0.249286745170910, 0.249286745170910, 0.116786275726379, 0.636502499121399, 0.310352451033785,
         0.082851075618374, 0.310352451033785, 0.636502499121399, 0.082851075618374, 0.636502499121399,
         0.053145049844816, 0.082851075618374, 0.053145049844816, 0.636502499121399, 0.082851075618374,
         0.310352451033785, 0.053145049844816, 0.082851075618374, 0.053145049844816, 0.310352451033785,
         0.082851075618374},
        {0.333333333333333, 0.333333333333333, -0.149570044467670, 0.479308067841923, 0.260345966079038,
         0.175615257433204, 0.260345966079038, 0.479308067841923, 0.175615257433204, 0.260345966079038,
         0.260345966079038, 0.175615257433204, 0.869739794195598, 0.065130102902216, 0.053347235608839,
         0.065130102902216, 0.869739794195598, 0.053347235608839, 0.065130102902216, 0.065130102902216,
         0.053347235608839, 0.638444188569809, 0.312865496004875, 0.077113760890257, 0.312865496004875,
         0.638444188569809, 0.077113760890257, 0.638444188569809, 0.048690315425316, 0.077113760890257,
         0.048690315425316, 0.638444188569809, 0.077113760890257, 0.312865496004875, 0.048690315425316,
         0.077113760890257, 0.048690315425316, 0.312865496004875, 0.077113760890257}};
    static const int nf[8] = {1, 3, 4, 6, 7, 9, 12, 13};
    static const double t[2][24] = {
        {0.128515707071765, 0.139571690967945, 0.621705865573422, 0.213074006319707, 0.608054478991729,
         0.137352730077163, 0.145898694726675, 0.225533192268079, 0.161734727754746, 0.159340192246828,
         0.160612296342170, 0.342484823577903, 0.137448101688511, 0.615380511289739, 0.135300542338320,
         0.218907977834312},
        {0.0701026973651683, 0.166660699730426, 0.596575741463645, 0.180396407632724, 0.168695371477743,
         0.584652626698847, 0.168695218708539, 0.193316294293662, 0.418175789241166, 0.0803037256707293,
         0.418101348156118, 0.133594798870359, 0.596576288848820, 0.166672635105983, 0.0700786026978242,
         0.180371079146698, 0.164808547352848, 0.0851134715186107, 0.164808684755350, 0.201016720944591,
         0.0731475667668616, 0.440339136918213, 0.0731570614497775, 0.111304699111967}};
    static const int nt3[2] = {4, 6};
    if (kind == 1 && index >= 0 && index < 5) {
        for (int i = 0; i < 2 * ne[index]; ++i) out[i] = e[index][i];
        return ne[index];
    }
    if (kind == 2 && index >= 0 && index < 8) {
        for (int i = 0; i < nf[index]; ++i) {
            const double x = f[index][3 * i], y = f[index][3 * i + 1];
            out[4 * i] = x;
            out[4 * i + 1] = y;
            out[4 * i + 2] = 1.0 - x - y;
            out[4 * i + 3] = f[index][3 * i + 2];
        }
        return nf[index];
    }
    if (kind == 3 && index >= 0 && index < 2) {
        for (int i = 0; i < nt3[index]; ++i) {
            const double a = t[index][4 * i], b = t[index][4 * i + 1], c = t[index][4 * i + 2];
            out[5 * i] = a;
            out[5 * i + 1] = b;
            out[5 * i + 2] = c;
            out[5 * i + 3] = 1.0 - a - b - c;
            out[5 * i + 4] = t[index][4 * i + 3];
        }
        return nt3[index];
    }
    return 0;
}

// quadrature rules passed by value as kernel arguments
struct Quad2 {
    int n;
    double l[13][3];
    double w[13];
};
struct Quad1 {
    int n;
    double x[5];
    double w[5];
};
struct Quad3 {
    int n;
    double l[6][4];
    double w[6];
};

static Quad2 face_rule(int index)
{
    double buf[13 * 4];
    Quad2 q{};
    q.n = quadrature_table(2, index, buf);
    PFEM_REQUIRE(q.n > 0, "unknown face quadrature rule");
    for (int i = 0; i < q.n; ++i) {
        for (int d = 0; d < 3; ++d) q.l[i][d] = buf[4 * i + d];
        q.w[i] = buf[4 * i + 3];
    }
    return q;
}

static Quad1 edge_rule(int index)
{
    double buf[10];
    Quad1 q{};
    q.n = quadrature_table(1, index, buf);
    PFEM_REQUIRE(q.n > 0, "unknown edge quadrature rule");
    for (int i = 0; i < q.n; ++i) { q.x[i] = buf[2 * i]; q.w[i] = buf[2 * i + 1]; }
    return q;
}

static Quad3 cell_rule(int index)
{
    double buf[6 * 5];
    Quad3 q{};
    q.n = quadrature_table(3, index, buf);
    PFEM_REQUIRE(q.n > 0, "unknown tetrahedron quadrature rule (0: 4 points, 1: 6 points)");
    for (int i = 0; i < q.n; ++i) {
        for (int d = 0; d < 4; ++d) q.l[i][d] = buf[5 * i + d];
        q.w[i] = buf[5 * i + 4];
    }
    return q;
}

// shapeFuncGrad (numerical_integrator_2d.cuh:8-20) mapped by Matrix2x2::operator* (cuda_math.cuh:148-156)
__device__ __forceinline__ double2 grad2(const double4 J, int i)
{
    const double gx = i == 0 ? 1.0 : (i == 1 ? 0.0 : -1.0);
    const double gy = i == 0 ? 0.0 : (i == 1 ? 1.0 : -1.0);
    return make_double2(J.x * gx + J.y * gy, J.z * gx + J.w * gy);
}

struct Pfem2Prm {
    double rho, mu, dt;
    int scheme, traction_mask;
    double nrm[8][2];
};

// ------------------------------------------------------------------ Poisson 2D (cases/Poisson2D/main.cu:28-68)
__global__ __launch_bounds__(kBlock) void k_asm_poisson2d(int nv, const int* __restrict__ inc_off,
                                                          const int* __restrict__ inc_code,
                                                          const int* __restrict__ inc_slots,
                                                          const double* __restrict__ xy, const int* __restrict__ tri,
                                                          const double* __restrict__ area,
                                                          const double4* __restrict__ invJ,
                                                          const int* __restrict__ rowptr, double* __restrict__ val,
                                                          double* __restrict__ rhs, double coeff, int rhs_kind,
                                                          double rhs_param, Quad2 q)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) val[k] = 0.0;
    const double PI = 3.14159265358979323846;
    double rr = 0.0;
    for (int e = inc_off[r]; e < inc_off[r + 1]; ++e) {
        const int code = inc_code[e];
        const int c = code >> 2, li = code & 3;
        const int v0 = tri[3 * c], v1 = tri[3 * c + 1], v2 = tri[3 * c + 2];
        const double4 J = invJ[c];
        double M[3] = {0.0, 0.0, 0.0};
        double R = 0.0;
        for (int k = 0; k < q.n; ++k) {
            const double L0 = q.l[k][0], L1 = q.l[k][1], L2 = q.l[k][2];
            const double qx = L0 * xy[2 * v0] + L1 * xy[2 * v1] + L2 * xy[2 * v2];
            const double qy = L0 * xy[2 * v0 + 1] + L1 * xy[2 * v1 + 1] + L2 * xy[2 * v2 + 1];
            for (int j = 0; j < 3; ++j) {
                const int a = li < j ? li : j, b = li < j ? j : li;   // SymmetricMatrix3x3 visits (min, max)
                const double2 ga = grad2(J, a), gb = grad2(J, b);
                const double aux = coeff * (ga.x * gb.x + ga.y * gb.y) * q.w[k];
                M[j] += aux;
            }
            double f;
            if (rhs_kind == 0) f = exp(-(qx * qx + 10 * qy * qy));
            else if (rhs_kind == 1) f = rhs_param * sin(PI * qx) * sin(PI * qy);
            else f = rhs_param;
            R += f * q.l[k][li] * q.w[k];
        }
        const double ar = area[c];
        const int* s = inc_slots + 3 * (size_t)e;
        for (int j = 0; j < 3; ++j) val[s[j]] += M[j] * ar;
        rr += R * ar;
    }
    rhs[r] = rr;
}

// ------------------------------------------------------------------ Poisson 3D (cases/Poisson3D/main.cu:26-67)
__global__ __launch_bounds__(kBlock) void k_asm_poisson3d(int nv, const int* __restrict__ inc_off,
                                                          const int* __restrict__ inc_code,
                                                          const int* __restrict__ inc_slots,
                                                          const double* __restrict__ p, const int* __restrict__ tet,
                                                          const double* __restrict__ vol, const double* __restrict__ invJ,
                                                          const int* __restrict__ rowptr, double* __restrict__ val,
                                                          double* __restrict__ rhs, double coeff, int rhs_kind,
                                                          double rhs_param, Quad3 q)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) val[k] = 0.0;
    const double PI = 3.14159265358979323846;
    const double g3[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, -1, -1}};
    double rr = 0.0;
    for (int e = inc_off[r]; e < inc_off[r + 1]; ++e) {
        const int code = inc_code[e];
        const int c = code >> 2, li = code & 3;
        int v[4];
        for (int a = 0; a < 4; ++a) v[a] = tet[4 * (size_t)c + a];
        const double* J = invJ + 9 * (size_t)c;
        double G[4][3];
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d) G[i][d] = J[3 * d] * g3[i][0] + J[3 * d + 1] * g3[i][1] + J[3 * d + 2] * g3[i][2];
        double M[4] = {0.0, 0.0, 0.0, 0.0};
        double R = 0.0;
        for (int k = 0; k < q.n; ++k) {
            const double w = q.w[k];
            double X[3];
            for (int d = 0; d < 3; ++d)
                X[d] = q.l[k][0] * p[3 * v[0] + d] + q.l[k][1] * p[3 * v[1] + d] + q.l[k][2] * p[3 * v[2] + d] +
                       q.l[k][3] * p[3 * v[3] + d];
            for (int j = 0; j < 4; ++j) {
                const double aux = coeff * (G[li][0] * G[j][0] + G[li][1] * G[j][1] + G[li][2] * G[j][2]) * w;
                M[j] += aux;
            }
            double f;
            if (rhs_kind == 0) f = exp(-(X[0] * X[0] + 10 * X[1] * X[1] + X[2] * X[2]));
            else if (rhs_kind == 1) f = rhs_param * sin(PI * X[0]) * sin(PI * X[1]) * sin(PI * X[2]);
            else f = rhs_param;
            R += f * q.l[k][li] * w;
        }
        const double vc = vol[c];
        const int* s = inc_slots + 4 * (size_t)e;
        for (int j = 0; j < 4; ++j) val[s[j]] += M[j] * vc;
        rr += R * vc;
    }
    rhs[r] = rr;
}

// ------------------------------------------------------------------ PFEM-2 velocity prediction
// cases/Cylinder2D/main.cu:72-167 (and PoiseuilleFlow2D/main.cu:81-176 for the boundary term on every ID)
__global__ __launch_bounds__(kBlock) void k_asm_prediction(
    int nv, const int* __restrict__ inc_off, const int* __restrict__ inc_code, const int* __restrict__ inc_slots,
    const double* __restrict__ xy, const int* __restrict__ tri, const double* __restrict__ area,
    const double4* __restrict__ invJ, const int* __restrict__ edge_ids, const int* __restrict__ rowptr,
    double* __restrict__ val0, double* __restrict__ val1, double* __restrict__ rhs0, double* __restrict__ rhs1,
    const double* __restrict__ u0, const double* __restrict__ u1, const double* __restrict__ uo0,
    const double* __restrict__ uo1, const double* __restrict__ p_old, Pfem2Prm prm, Quad2 q, Quad1 qe)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) { val0[k] = 0.0; val1[k] = 0.0; }
    const double rho = prm.rho, mu = prm.mu, dt = prm.dt;
    double rr0 = 0.0, rr1 = 0.0;
    for (int e = inc_off[r]; e < inc_off[r + 1]; ++e) {
        const int code = inc_code[e];
        const int c = code >> 2, i = code & 3;
        int v[3];
        v[0] = tri[3 * c]; v[1] = tri[3 * c + 1]; v[2] = tri[3 * c + 2];
        const double4 J = invJ[c];
        double M0[3] = {0, 0, 0}, M1[3] = {0, 0, 0};
        double R0 = 0.0, R1 = 0.0;
        const double2 gI = grad2(J, i);
        for (int qp = 0; qp < q.n; ++qp) {
            const double w = q.w[qp];
            const double Ni = q.l[qp][i];
            for (int j = 0; j < 3; ++j) {
                const double2 gJ = grad2(J, j);
                const double Nj = q.l[qp][j];
                double aux = rho * Ni * Nj * w;
                M0[j] += aux;
                R0 += aux * uo0[v[j]];
                M1[j] += aux;
                R1 += aux * uo1[v[j]];
                aux = mu * dt * w;
                M0[j] += aux * (gI.y * gJ.y + 4.0 / 3.0 * gI.x * gJ.x);
                M1[j] += aux * (gI.x * gJ.x + 4.0 / 3.0 * gI.y * gJ.y);
                R0 -= aux * (gI.y * gJ.x - 2.0 / 3.0 * gI.x * gJ.y) * u1[v[j]];
                R1 -= aux * (gI.x * gJ.y - 2.0 / 3.0 * gI.y * gJ.x) * u0[v[j]];
                if (p_old) {
                    aux = Ni * p_old[v[j]] * dt * w;
                    R0 -= aux * gJ.x;
                    R1 -= aux * gJ.y;
                }
            }
        }
        for (int ed = 0; ed < 3; ++ed) {
            const int id = edge_ids[3 * c + ed];
            if (id < 0 || id > 7 || !((prm.traction_mask >> id) & 1)) continue;
            const double nx = prm.nrm[id][0], ny = prm.nrm[id][1];
            const int a = v[ed], b = v[(ed + 1) % 3];
            const double sx = xy[2 * a], sy = xy[2 * a + 1], ex = xy[2 * b], ey = xy[2 * b + 1];
            const double halfLength = 0.5 * sqrt((sx - ex) * (sx - ex) + (sy - ey) * (sy - ey));
            const double v3x = xy[2 * v[2]], v3y = xy[2 * v[2] + 1];
            for (int qp = 0; qp < qe.n; ++qp) {
                const double cd = qe.x[qp];
                const double px = 0.5 * (cd * (ex - sx) + (sx + ex));
                const double py = 0.5 * (cd * (ey - sy) + (sy + ey));
                const double dx = px - v3x, dy = py - v3y;
                double L[3];
                L[0] = J.x * dx + J.z * dy;
                L[1] = J.y * dx + J.w * dy;
                L[2] = 1.0 - L[0] - L[1];
                const double aux = mu * dt * qe.w[qp] * halfLength;
                const double Ni = L[i];
                for (int j = 0; j < 3; ++j) {
                    const double2 gJ = grad2(J, j);
                    M0[j] -= aux * Ni * ((4.0 / 3.0) * gJ.x * nx + gJ.y * ny);
                    M1[j] -= aux * Ni * (gJ.x * nx + (4.0 / 3.0) * gJ.y * ny);
                    R0 += aux * Ni * u1[v[j]] * ((-2.0 / 3.0) * gJ.y * nx + gJ.x * ny);
                    R1 += aux * Ni * u0[v[j]] * (gJ.y * nx + (-2.0 / 3.0) * gJ.x * ny);
                }
            }
        }
        const double ar = area[c];
        const int* s = inc_slots + 3 * (size_t)e;
        for (int j = 0; j < 3; ++j) {
            val0[s[j]] += M0[j] * ar;
            val1[s[j]] += M1[j] * ar;
        }
        rr0 += R0 * ar;
        rr1 += R1 * ar;
    }
    rhs0[r] = rr0;
    rhs1[r] = rr1;
}

// ------------------------------------------------------------------ PFEM-2 pressure (Cylinder2D/main.cu:169-211)
__global__ __launch_bounds__(kBlock) void k_asm_pressure(
    int nv, const int* __restrict__ inc_off, const int* __restrict__ inc_code, const int* __restrict__ inc_slots,
    const int* __restrict__ tri, const double* __restrict__ area, const double4* __restrict__ invJ,
    const int* __restrict__ rowptr, double* __restrict__ val, double* __restrict__ rhs, const double* __restrict__ vp0,
    const double* __restrict__ vp1, const double* __restrict__ p_old, Pfem2Prm prm, Quad2 q)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) val[k] = 0.0;
    double rr = 0.0;
    for (int e = inc_off[r]; e < inc_off[r + 1]; ++e) {
        const int code = inc_code[e];
        const int c = code >> 2, i = code & 3;
        int v[3];
        v[0] = tri[3 * c]; v[1] = tri[3 * c + 1]; v[2] = tri[3 * c + 2];
        const double4 J = invJ[c];
        double M[3] = {0, 0, 0};
        double R = 0.0;
        const double2 gI = grad2(J, i);
        for (int qp = 0; qp < q.n; ++qp) {
            const double w = q.w[qp];
            const double Ni = q.l[qp][i];
            const double aux = -prm.rho / prm.dt * Ni * w;
            for (int j = 0; j < 3; ++j) {
                const double2 gJ = grad2(J, j);
                const double vx = vp0[v[j]], vy = vp1[v[j]];
                const double aux2 = (gI.x * gJ.x + gI.y * gJ.y) * w;
                M[j] += aux2;
                if (p_old) R += aux2 * p_old[v[j]];
                R += aux * (gJ.x * vx + gJ.y * vy);
            }
        }
        const double ar = area[c];
        const int* s = inc_slots + 3 * (size_t)e;
        for (int j = 0; j < 3; ++j) val[s[j]] += M[j] * ar;
        rr += R * ar;
    }
    rhs[r] = rr;
}

// ------------------------------------------------------------------ PFEM-2 correction (Cylinder2D/main.cu:213-259)
__global__ __launch_bounds__(kBlock) void k_asm_correction(
    int nv, const int* __restrict__ inc_off, const int* __restrict__ inc_code, const int* __restrict__ inc_slots,
    const int* __restrict__ tri, const double* __restrict__ area, const double4* __restrict__ invJ,
    const int* __restrict__ rowptr, double* __restrict__ val, double* __restrict__ rhs0, double* __restrict__ rhs1,
    const double* __restrict__ vp0, const double* __restrict__ vp1, const double* __restrict__ p,
    const double* __restrict__ p_old, Pfem2Prm prm, Quad2 q)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) val[k] = 0.0;
    double rr0 = 0.0, rr1 = 0.0;
    for (int e = inc_off[r]; e < inc_off[r + 1]; ++e) {
        const int code = inc_code[e];
        const int c = code >> 2, i = code & 3;
        int v[3];
        v[0] = tri[3 * c]; v[1] = tri[3 * c + 1]; v[2] = tri[3 * c + 2];
        const double4 J = invJ[c];
        double M[3] = {0, 0, 0};
        double R0 = 0.0, R1 = 0.0;
        for (int qp = 0; qp < q.n; ++qp) {
            const double w = q.w[qp];
            const double Ni = q.l[qp][i];
            const double aux2 = prm.dt * Ni * w;
            for (int j = 0; j < 3; ++j) {
                const double2 gJ = grad2(J, j);
                const double Nj = q.l[qp][j];
                const double aux = prm.rho * Ni * Nj * w;
                double pv = p[v[j]];
                if (p_old) pv -= p_old[v[j]];
                M[j] += aux;
                R0 += aux * vp0[v[j]] - aux2 * gJ.x * pv;
                R1 += aux * vp1[v[j]] - aux2 * gJ.y * pv;
            }
        }
        const double ar = area[c];
        const int* s = inc_slots + 3 * (size_t)e;
        for (int j = 0; j < 3; ++j) val[s[j]] += M[j] * ar;
        rr0 += R0 * ar;
        rr1 += R1 * ar;
    }
    rhs0[r] = rr0;
    rhs1[r] = rr1;
}

// ------------------------------------------------------------------ prediction BC values
// kAccumulatePressureGradient + kFinalizePredictionBC (Cylinder2D/main.cu:261-297): one lane per BC node,
// its incident cells in ascending order.
__global__ void k_prediction_bc(int nbc, const int* __restrict__ node, const int* __restrict__ inc_off,
                                const int* __restrict__ inc_code, const int* __restrict__ tri,
                                const double* __restrict__ area, const double4* __restrict__ invJ,
                                const double* __restrict__ p, const double* __restrict__ p_old,
                                const double* __restrict__ src, double* __restrict__ out, int component, double dt,
                                double rho)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbc) return;
    const int b = node[k];
    double num = 0.0, den = 0.0;
    for (int e = inc_off[b]; e < inc_off[b + 1]; ++e) {
        const int c = inc_code[e] >> 2;
        double gx = 0.0, gy = 0.0;
        for (int i = 0; i < 3; ++i) {
            const int vi = tri[3 * c + i];
            const double s = p[vi] - (p_old ? p_old[vi] : 0.0);
            const double sgx = i == 0 ? 1.0 : (i == 1 ? 0.0 : -1.0);
            const double sgy = i == 0 ? 0.0 : (i == 1 ? 1.0 : -1.0);
            gx += sgx * s;
            gy += sgy * s;
        }
        const double4 J = invJ[c];
        const double g = component == 0 ? J.x * gx + J.y * gy : J.z * gx + J.w * gy;
        const double ar = area[c];
        num += ar * g;
        den += ar;
    }
    out[k] = src[k] + dt / rho * num / den;
}

// ------------------------------------------------------------------ body loads (Cylinder2D/main.cu:322-376)
// The loads {pressure x, pressure y, viscous x, viscous y} of edge ed of cell c into l.
__device__ __forceinline__ void body_edge_load(int c, int ed, const double* __restrict__ xy, const int* __restrict__ tri,
                                               const double4* __restrict__ invJ, const double* __restrict__ u0,
                                               const double* __restrict__ u1, const double* __restrict__ p, double cx,
                                               double cy, double mu, const Quad1& qe, double (&l)[4])
{
    const int v[3] = {tri[3 * c], tri[3 * c + 1], tri[3 * c + 2]};
    const double4 J = invJ[c];
    const int a = v[ed], b = v[(ed + 1) % 3];
    const double sx = xy[2 * a], sy = xy[2 * a + 1], ex = xy[2 * b], ey = xy[2 * b + 1];
    const double mx = (sx + ex) * 0.5 - cx, my = (sy + ey) * 0.5 - cy;
    const double inv = 1.0 / sqrt(mx * mx + my * my);
    const double nx = mx * inv, ny = my * inv;
    const double halfLength = 0.5 * sqrt((sx - ex) * (sx - ex) + (sy - ey) * (sy - ey));
    const double tx = ny, ty = -nx;
    const double v3x = xy[2 * v[2]], v3y = xy[2 * v[2] + 1];
    l[0] = l[1] = l[2] = l[3] = 0.0;
    for (int qp = 0; qp < qe.n; ++qp) {
        double pq = 0.0, dvtdn = 0.0;
        const double cd = qe.x[qp];
        const double px = 0.5 * (cd * (ex - sx) + (sx + ex));
        const double py = 0.5 * (cd * (ey - sy) + (sy + ey));
        const double dx = px - v3x, dy = py - v3y;
        double L[3];
        L[0] = J.x * dx + J.z * dy;
        L[1] = J.y * dx + J.w * dy;
        L[2] = 1.0 - L[0] - L[1];
        for (int i = 0; i < 3; ++i) {
            pq += p[v[i]] * L[i];
            const double2 g = grad2(J, i);
            dvtdn += (u0[v[i]] * tx + u1[v[i]] * ty) * (g.x * nx + g.y * ny);
        }
        const double weight = qe.w[qp] * halfLength;
        l[0] -= pq * nx * weight;
        l[1] -= pq * ny * weight;
        l[2] += mu * dvtdn * tx * weight;
        l[3] += mu * dvtdn * ty * weight;
    }
}

__global__ void k_body_edge_loads(int ne, const int* __restrict__ cell_edge, const double* __restrict__ xy,
                                  const int* __restrict__ tri, const double4* __restrict__ invJ,
                                  const double* __restrict__ u0, const double* __restrict__ u1,
                                  const double* __restrict__ p, double cx, double cy, double mu, Quad1 qe,
                                  double* __restrict__ loads)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ne) return;
    double l[4];
    body_edge_load(cell_edge[2 * k], cell_edge[2 * k + 1], xy, tri, invJ, u0, u1, p, cx, cy, mu, qe, l);
    for (int d = 0; d < 4; ++d) loads[4 * (size_t)k + d] = l[d];
}

// edges summed in list order by one lane (a few hundred edges)
__global__ void k_body_sum(int ne, const double* __restrict__ loads, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t[4] = {0, 0, 0, 0};
    for (int k = 0; k < ne; ++k)
        for (int d = 0; d < 4; ++d) t[d] += loads[4 * (size_t)k + d];
    for (int d = 0; d < 4; ++d) out[d] = t[d];
}

static Pfem2Prm to_prm(const pfem_pfem2_params* p)
{
    Pfem2Prm o{};
    o.rho = p->rho;
    o.mu = p->mu;
    o.dt = p->dt;
    o.scheme = p->scheme;
    o.traction_mask = p->traction_mask;
    for (int k = 0; k < 8; ++k) { o.nrm[k][0] = p->normal[k][0]; o.nrm[k][1] = p->normal[k][1]; }
    return o;
}

static void check_pattern(pfem_mesh2d m, pfem_csr a)
{
    PFEM_REQUIRE(a->pat.get() == m->pat.get(), "assembly needs a matrix created from this mesh's pattern");
}

void assemble_prediction(pfem_mesh2d m, pfem_csr a0, pfem_csr a1, double* rhs0, double* rhs1, const double* u0,
                         const double* u1, const double* uo0, const double* uo1, const double* p_old,
                         const pfem_pfem2_params* prm)
{
    check_pattern(m, a0);
    check_pattern(m, a1);
    const Quad2 q = face_rule(prm->face_quad);
    const Quad1 qe = edge_rule(prm->edge_quad);
    Context* c = C(m->ctx);
    hipLaunchKernelGGL(k_asm_prediction, dim3(blocks_for(m->nv)), dim3(kBlock), 0, c->stream, m->nv, m->inc.off.p,
                       m->inc.code.p, m->inc.slots.p, m->xy.p, m->tri.p, m->area.p, (const double4*)m->invJ.p,
                       m->edge_ids.p, m->pat->rowptr.p, a0->val.p, a1->val.p, rhs0, rhs1, u0, u1, uo0, uo1, p_old,
                       to_prm(prm), q, qe);
    PFEM_HIP(hipGetLastError());
}

void assemble_pressure(pfem_mesh2d m, pfem_csr a, double* rhs, const double* vp0, const double* vp1,
                       const double* p_old, const pfem_pfem2_params* prm)
{
    check_pattern(m, a);
    const Quad2 q = face_rule(prm->face_quad);
    Context* c = C(m->ctx);
    hipLaunchKernelGGL(k_asm_pressure, dim3(blocks_for(m->nv)), dim3(kBlock), 0, c->stream, m->nv, m->inc.off.p,
                       m->inc.code.p, m->inc.slots.p, m->tri.p, m->area.p, (const double4*)m->invJ.p,
                       m->pat->rowptr.p, a->val.p, rhs, vp0, vp1, p_old, to_prm(prm), q);
    PFEM_HIP(hipGetLastError());
}

void assemble_correction(pfem_mesh2d m, pfem_csr a, double* rhs0, double* rhs1, const double* vp0, const double* vp1,
                         const double* p, const double* p_old, const pfem_pfem2_params* prm)
{
    check_pattern(m, a);
    const Quad2 q = face_rule(prm->face_quad);
    Context* c = C(m->ctx);
    hipLaunchKernelGGL(k_asm_correction, dim3(blocks_for(m->nv)), dim3(kBlock), 0, c->stream, m->nv, m->inc.off.p,
                       m->inc.code.p, m->inc.slots.p, m->tri.p, m->area.p, (const double4*)m->invJ.p,
                       m->pat->rowptr.p, a->val.p, rhs0, rhs1, vp0, vp1, p, p_old, to_prm(prm), q);
    PFEM_HIP(hipGetLastError());
}

void prediction_bc_values(pfem_mesh2d m, pfem_dirichlet target, pfem_dirichlet source, const double* p,
                          const double* p_old, int component, const pfem_pfem2_params* prm)
{
    PFEM_REQUIRE(target->nbc == source->nbc && target->h_node == source->h_node,
                 "prediction BC target and source must have the same nodes");
    PFEM_REQUIRE(target->n == m->nv, "BC node map size differs from the mesh");
    if (target->nbc == 0) return;
    Context* c = C(m->ctx);
    hipLaunchKernelGGL(k_prediction_bc, dim3(blocks_for(target->nbc)), dim3(kBlock), 0, c->stream, target->nbc,
                       target->node.p, m->inc.off.p, m->inc.code.p, m->tri.p, m->area.p, (const double4*)m->invJ.p, p,
                       p_old, source->value.p, target->value.p, component, prm->dt, prm->rho);
    PFEM_HIP(hipGetLastError());
}

// The body's (cell, edge) list, found once per body ID (host work and an upload: never inside a captured
// launch sequence)
void body_forces_prepare(pfem_mesh2d m, int body_id)
{
    Context* c = C(m->ctx);
    if (m->body_id_cached != body_id) {
        // kCountBodyEdges (Cylinder2D/main.cu:299-320): one (cell, edge) per cell, ascending cells
        std::vector<int> ce;
        for (int cc = 0; cc < m->nt; ++cc)
            for (int i = 0; i < 3; ++i)
                if (m->h_edge_ids[3 * (size_t)cc + i] == body_id) { ce.push_back(cc); ce.push_back(i); break; }
        m->n_body = (int)(ce.size() / 2);
        upload(c, m->body_edges, ce);
        m->body_loads.alloc(4 * (size_t)std::max(m->n_body, 1) + 4);
        m->body_id_cached = body_id;
        ++m->generation;
    }
}

void body_forces(pfem_mesh2d m, int body_id, double cx, double cy, const double* u0, const double* u1,
                 const double* p, const pfem_pfem2_params* prm, double* d_out /* 4 doubles, device */)
{
    Context* c = C(m->ctx);
    body_forces_prepare(m, body_id);
    const Quad1 qe = edge_rule(prm->edge_quad);
    if (m->n_body > 0)
        hipLaunchKernelGGL(k_body_edge_loads, dim3(blocks_for(m->n_body)), dim3(kBlock), 0, c->stream, m->n_body,
                           m->body_edges.p, m->xy.p, m->tri.p, (const double4*)m->invJ.p, u0, u1, p, cx, cy, prm->mu,
                           qe, m->body_loads.p);
    hipLaunchKernelGGL(k_body_sum, dim3(1), dim3(64), 0, c->stream, m->n_body, m->body_loads.p, d_out);
    PFEM_HIP(hipGetLastError());
}

// ---- allocation-free body forces (pfem_body_forces_noalloc)
//
// The body's edges are found on the device instead of from the host copy of the edge IDs: a count of body
// cells per workgroup, an exclusive scan of the counts, and a pass in which each body cell writes its
// edge's loads at its rank (ascending cells, as the cached edge list of body_forces orders them); one lane
// then adds the loads in that order, so the totals are bit-identical to body_forces'. Everything lives in
// one static device array per GPU (about 2.4 MB): block counts, {edge count, overflow flag}, the four
// totals and the per-edge loads. Concurrent calls on one GPU share it and race.
constexpr int kNaMaxBlocks = 65536;   // cells <= 65536 * 256
constexpr int kNaMaxEdges = 65536;    // body edges
constexpr long kNaStatus = kNaMaxBlocks / 2, kNaTotals = kNaStatus + 2, kNaLoads = kNaTotals + 4;
static __device__ __align__(256) double scratch[kNaLoads + 4L * kNaMaxEdges];

int noalloc_init(Context* c)
{
    // one lookup per device; every context on that device shares the address
    static std::mutex mu;
    static std::vector<double*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= c->device) addr.resize(c->device + 1, nullptr);
    if (!addr[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        addr[c->device] = static_cast<double*>(p);
    }
    c->scratch = addr[c->device];
    return (int)hipSuccess;
}

// the edge of cell c on the body: the first of its edges with that ID (kCountBodyEdges,
// Cylinder2D/main.cu:299-320), or -1
__device__ __forceinline__ int body_edge_of(int c, const int* __restrict__ edge_ids, int body_id)
{
    for (int i = 0; i < 3; ++i)
        if (edge_ids[3 * c + i] == body_id) return i;
    return -1;
}

// body cells among the 256 cells of each workgroup
__global__ void k_na_count(int nt, const int* __restrict__ edge_ids, int body_id, int* __restrict__ counts)
{
    __shared__ int wc[kWaves];
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const bool on = c < nt && body_edge_of(c, edge_ids, body_id) >= 0;
    const unsigned long long b = __ballot(on);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// exclusive scan of the nb counts in place, one workgroup of 1024 lanes (lane t: a contiguous run of
// counts); status = {total, 0}
__global__ __launch_bounds__(1024) void k_na_scan(int nb, int* __restrict__ counts, int* __restrict__ status)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    int s = 0;
    for (int b = b0; b < b1; ++b) s += counts[b];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int b = b0; b < b1; ++b) {
        const int cnt = counts[b];
        counts[b] = run;
        run += cnt;
    }
    if (t == 1023) {
        status[0] = part[1023];
        status[1] = 0;
    }
}

// each body cell's edge loads at its rank among the body cells; ranks past the capacity raise status[1]
__global__ void k_na_loads(int nt, const int* __restrict__ edge_ids, int body_id, const int* __restrict__ offsets,
                           const double* __restrict__ xy, const int* __restrict__ tri,
                           const double4* __restrict__ invJ, const double* __restrict__ u0,
                           const double* __restrict__ u1, const double* __restrict__ p, double cx, double cy,
                           double mu, Quad1 qe, double* __restrict__ loads, int* __restrict__ status)
{
    __shared__ int wc[kWaves];
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int ed = c < nt ? body_edge_of(c, edge_ids, body_id) : -1;
    const unsigned long long b = __ballot(ed >= 0);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wc[w] = __popcll(b);
    __syncthreads();
    if (ed < 0) return;
    int rank = offsets[blockIdx.x] + __popcll(b & ((1ull << lane) - 1ull));
    for (int k = 0; k < w; ++k) rank += wc[k];
    if (rank >= kNaMaxEdges) {
        __hip_atomic_store(status + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    double l[4];
    body_edge_load(c, ed, xy, tri, invJ, u0, u1, p, cx, cy, mu, qe, l);
    for (int d = 0; d < 4; ++d) loads[4 * (size_t)rank + d] = l[d];
}

// the loads added in rank order by one lane (k_body_sum's order)
__global__ void k_na_sum(const int* __restrict__ status, const double* __restrict__ loads, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int ne = min(status[0], kNaMaxEdges);
    double t[4] = {0, 0, 0, 0};
    for (int k = 0; k < ne; ++k)
        for (int d = 0; d < 4; ++d) t[d] += loads[4 * (size_t)k + d];
    for (int d = 0; d < 4; ++d) out[d] = t[d];
}

}  // namespace pfem

using namespace pfem;

extern "C" {

int pfem_quadrature(int kind, int index, double* out)
{
    if (!out) return PFEM_EINVAL;
    const int n = quadrature_table(kind, index, out);
    return n > 0 ? n : PFEM_EINVAL;
}

int pfem_assemble_poisson2d(pfem_mesh2d m, pfem_csr a, double* d_rhs, double coeff, int rhs_kind, double rhs_param,
                            int quad_index)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && a && d_rhs, "null argument");
        check_pattern(m, a);
        const Quad2 q = face_rule(quad_index);
        Context* c = C(m->ctx);
        hipLaunchKernelGGL(k_asm_poisson2d, dim3(blocks_for(m->nv)), dim3(kBlock), 0, c->stream, m->nv,
                           m->inc.off.p, m->inc.code.p, m->inc.slots.p, m->xy.p, m->tri.p, m->area.p,
                           (const double4*)m->invJ.p, m->pat->rowptr.p, a->val.p, d_rhs, coeff, rhs_kind, rhs_param, q);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_assemble_poisson3d(pfem_mesh3d m, pfem_csr a, double* d_rhs, double coeff, int rhs_kind, double rhs_param,
                            int quad_index)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && a && d_rhs, "null argument");
        PFEM_REQUIRE(a->pat.get() == m->pat.get(), "assembly needs a matrix created from this mesh's pattern");
        const Quad3 q = cell_rule(quad_index);
        Context* c = C(m->ctx);
        hipLaunchKernelGGL(k_asm_poisson3d, dim3(blocks_for(m->nv)), dim3(kBlock), 0, c->stream, m->nv,
                           m->inc.off.p, m->inc.code.p, m->inc.slots.p, m->xyz.p, m->tet.p, m->vol.p, m->invJ.p,
                           m->pat->rowptr.p, a->val.p, d_rhs, coeff, rhs_kind, rhs_param, q);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_assemble_prediction(pfem_mesh2d m, pfem_csr a0, pfem_csr a1, double* d_rhs0, double* d_rhs1,
                             const double* d_u0, const double* d_u1, const double* d_uo0, const double* d_uo1,
                             const double* d_p_old, const pfem_pfem2_params* prm)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && a0 && a1 && d_rhs0 && d_rhs1 && d_u0 && d_u1 && d_uo0 && d_uo1 && prm, "null argument");
        assemble_prediction(m, a0, a1, d_rhs0, d_rhs1, d_u0, d_u1, d_uo0, d_uo1, d_p_old, prm);
    });
}

int pfem_assemble_pressure(pfem_mesh2d m, pfem_csr a, double* d_rhs, const double* d_vp0, const double* d_vp1,
                           const double* d_p_old, const pfem_pfem2_params* prm)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && a && d_rhs && d_vp0 && d_vp1 && prm, "null argument");
        assemble_pressure(m, a, d_rhs, d_vp0, d_vp1, d_p_old, prm);
    });
}

int pfem_assemble_correction(pfem_mesh2d m, pfem_csr a, double* d_rhs0, double* d_rhs1, const double* d_vp0,
                             const double* d_vp1, const double* d_p, const double* d_p_old,
                             const pfem_pfem2_params* prm)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && a && d_rhs0 && d_rhs1 && d_vp0 && d_vp1 && d_p && prm, "null argument");
        assemble_correction(m, a, d_rhs0, d_rhs1, d_vp0, d_vp1, d_p, d_p_old, prm);
    });
}

int pfem_prediction_bc_values(pfem_mesh2d m, pfem_dirichlet target, pfem_dirichlet source, const double* d_p,
                              const double* d_p_old, int component, const pfem_pfem2_params* prm)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && target && source && d_p && prm, "null argument");
        prediction_bc_values(m, target, source, d_p, d_p_old, component, prm);
    });
}

int pfem_body_forces(pfem_mesh2d m, int body_id, double cx, double cy, const double* d_u0, const double* d_u1,
                     const double* d_p, const pfem_pfem2_params* prm, double out[4])
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && d_u0 && d_u1 && d_p && prm && out, "null argument");
        DevArray<double> d(4);
        body_forces(m, body_id, cx, cy, d_u0, d_u1, d_p, prm, d.p);
        d2h_sync(C(m->ctx), out, d.p, 4 * sizeof(double));
    });
}

int pfem_body_forces_noalloc(pfem_mesh2d m, int body_id, double cx, double cy, const double* d_u0,
                             const double* d_u1, const double* d_p, const pfem_pfem2_params* prm, double out[4])
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && d_u0 && d_u1 && d_p && prm && out, "null argument");
        Context* c = C(m->ctx);
        PFEM_REQUIRE(c->scratch, "allocation-free scratch not resolved on this device");
        PFEM_REQUIRE(m->edge_ids.p, "edge boundary IDs are not set on the mesh");
        const long nb = blocks_for(m->nt);
        PFEM_REQUIRE(nb <= kNaMaxBlocks, "pfem_body_forces_noalloc: the mesh has more than 16777216 cells");
        double* sc = c->scratch;
        int* counts = reinterpret_cast<int*>(sc);
        int* status = reinterpret_cast<int*>(sc + kNaStatus);
        double* totals = sc + kNaTotals;
        if (nb > 0) {
            hipLaunchKernelGGL(k_na_count, dim3(nb), dim3(kBlock), 0, c->stream, m->nt, m->edge_ids.p, body_id, counts);
            hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(1024), 0, c->stream, (int)nb, counts, status);
            hipLaunchKernelGGL(k_na_loads, dim3(nb), dim3(kBlock), 0, c->stream, m->nt, m->edge_ids.p, body_id,
                               (const int*)counts, m->xy.p, m->tri.p, (const double4*)m->invJ.p, d_u0, d_u1, d_p, cx,
                               cy, prm->mu, edge_rule(prm->edge_quad), sc + kNaLoads, status);
            hipLaunchKernelGGL(k_na_sum, dim3(1), dim3(64), 0, c->stream, (const int*)status,
                               (const double*)(sc + kNaLoads), totals);
            PFEM_HIP(hipGetLastError());
            // read back through the context's pinned buffers (no allocation)
            PFEM_HIP(hipMemcpyAsync(c->pinned_d, totals, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            PFEM_HIP(hipMemcpyAsync(c->pinned_i, status, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            PFEM_HIP(hipStreamSynchronize(c->stream));
            PFEM_REQUIRE(c->pinned_i[1] == 0, "pfem_body_forces_noalloc: the body has more than 65536 edges");
            for (int d = 0; d < 4; ++d) out[d] = c->pinned_d[d];
        } else {
            for (int d = 0; d < 4; ++d) out[d] = 0.0;
        }
    });
}

}  // extern "C"
