// Oracle: mesh geometry, CSR pattern, cell neighbours, quadrature tables (TEST INFRASTRUCTURE).
// Built with -ffp-contract=off; each expression below keeps the operand order of the reference.
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <set>
#include <vector>

extern "C" {

// src/mesh_2d.cu:10-19 (kCalculateCellArea) and :21-34 (kCalculateInvJacobi) with
// src/common/cuda_math.cuh:124-141 (Matrix2x2::inverse; prints but does not zero when |det| < 2e-6).
// invJ is stored row-major (4 doubles per cell), the layout of Matrix2x2::data.
void orc_geometry2d(int nv, const double* xy, int nt, const int* tri, double* area, double* invJ)
{
    (void)nv;
    for (int c = 0; c < nt; ++c) {
        const int a = tri[3 * c], b = tri[3 * c + 1], d = tri[3 * c + 2];
        const double v12x = xy[2 * b] - xy[2 * a], v12y = xy[2 * b + 1] - xy[2 * a + 1];
        const double v13x = xy[2 * d] - xy[2 * a], v13y = xy[2 * d + 1] - xy[2 * a + 1];
        area[c] = std::fabs(v12x * v13y - v12y * v13x) * 0.5;

        const double j00 = xy[2 * a] - xy[2 * d], j01 = xy[2 * a + 1] - xy[2 * d + 1];
        const double j10 = xy[2 * b] - xy[2 * d], j11 = xy[2 * b + 1] - xy[2 * d + 1];
        const double det = j00 * j11 - j01 * j10;
        const double invdet = 1.0 / det;
        invJ[4 * c + 0] = j11 * invdet;
        invJ[4 * c + 1] = -j01 * invdet;
        invJ[4 * c + 2] = -j10 * invdet;
        invJ[4 * c + 3] = j00 * invdet;
    }
}

// src/mesh_3d.cu:12-35 with src/common/matrix3x3.cuh:127-154 (GenericMatrix3x3::det/inverse).
// Deviation (recorded): the reference returns a zero inverse when |det| < 2e-6; the oracle keeps
// the true inverse, because the 128^3 unit-box mesh has det = h^3 = 4.77e-7 for every tet.
void orc_geometry3d(int nv, const double* p, int nt, const int* tet, double* vol, double* invJ)
{
    (void)nv;
    for (int c = 0; c < nt; ++c) {
        const int i0 = tet[4 * c], i1 = tet[4 * c + 1], i2 = tet[4 * c + 2], i3 = tet[4 * c + 3];
        double r[3][3];
        const int ids[3] = {i0, i1, i2};
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) r[k][d] = p[3 * ids[k] + d] - p[3 * i3 + d];
        // volume = |v41 . (v42 x v43)| * ONE_SIXTH  (cuda_math cross/dot)
        const double cx = r[1][1] * r[2][2] - r[1][2] * r[2][1];
        const double cy = r[1][2] * r[2][0] - r[1][0] * r[2][2];
        const double cz = r[1][0] * r[2][1] - r[1][1] * r[2][0];
        vol[c] = std::fabs(r[0][0] * cx + r[0][1] * cy + r[0][2] * cz) * 0.1666666666666667;
        const double* m = &r[0][0];
        const double det = (m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                            m[2] * (m[3] * m[7] - m[4] * m[6]));
        const double id = 1.0 / det;
        double* o = invJ + 9 * c;
        o[0] = id * (m[4] * m[8] - m[5] * m[7]);
        o[1] = id * (m[2] * m[7] - m[1] * m[8]);
        o[2] = id * (m[1] * m[5] - m[2] * m[4]);
        o[3] = id * (m[5] * m[6] - m[3] * m[8]);
        o[4] = id * (m[0] * m[8] - m[2] * m[6]);
        o[5] = id * (m[2] * m[3] - m[0] * m[5]);
        o[6] = id * (m[3] * m[7] - m[4] * m[6]);
        o[7] = id * (m[1] * m[6] - m[0] * m[7]);
        o[8] = id * (m[0] * m[4] - m[1] * m[3]);
    }
}

// src/linear_algebra/sparse_matrix.cu:8-54 (2D) and :56-99 (3D): row i = {i} U {vertices sharing a cell
// with i}, columns ascending (std::set order). nper = 3 (triangles) or 4 (tets).
static void build_sets(int nv, int nt, int nper, const int* cells, std::vector<std::set<int>>& conn)
{
    conn.assign(nv, std::set<int>());
    for (int i = 0; i < nv; ++i) conn[i].insert(i);
    for (int c = 0; c < nt; ++c)
        for (int a = 0; a < nper; ++a)
            for (int b = 0; b < nper; ++b)
                if (a != b) conn[cells[nper * c + a]].insert(cells[nper * c + b]);
}

long orc_csr_count(int nv, int nt, int nper, const int* cells)
{
    std::vector<std::set<int>> conn;
    build_sets(nv, nt, nper, cells, conn);
    long nnz = 0;
    for (auto& s : conn) nnz += (long)s.size();
    return nnz;
}

void orc_csr_fill(int nv, int nt, int nper, const int* cells, int* rowptr, int* col)
{
    std::vector<std::set<int>> conn;
    build_sets(nv, nt, nper, cells, conn);
    rowptr[0] = 0;
    for (int i = 0; i < nv; ++i) {
        int k = rowptr[i];
        for (int c : conn[i]) col[k++] = c;
        rowptr[i + 1] = k;
    }
}

// src/mesh_2d.cu:107-139: all-pairs test, cells sharing >= 1 vertex, ascending (std::set). O(T^2):
// only for the small fixtures.
static void nbr_sets(int nt, const int* t, std::vector<std::set<int>>& s)
{
    s.assign(nt, std::set<int>());
    for (int i = 0; i < nt; ++i)
        for (int j = i + 1; j < nt; ++j) {
            bool share = false;
            for (int a = 0; a < 3 && !share; ++a)
                for (int b = 0; b < 3; ++b)
                    if (t[3 * i + a] == t[3 * j + b]) { share = true; break; }
            if (share) { s[i].insert(j); s[j].insert(i); }
        }
}

long orc_neighbors_bruteforce_count(int nt, const int* tri)
{
    std::vector<std::set<int>> s;
    nbr_sets(nt, tri, s);
    long total = 0;
    for (auto& x : s) total += (long)x.size();
    return total;
}

void orc_neighbors_bruteforce_fill(int nt, const int* tri, int* offsets, int* idx)
{
    std::vector<std::set<int>> s;
    nbr_sets(nt, tri, s);
    offsets[0] = 0;
    for (int i = 0; i < nt; ++i) {
        int k = offsets[i];
        for (int c : s[i]) idx[k++] = c;
        offsets[i + 1] = k;
    }
}

// src/mesh_3d.cu:37-82 (kFindNeighbors, the path loadMeshFromFile takes by default) and :149-188 (the CPU
// variant): tetrahedra whose vertex lists have exactly 3 equal pairs (a shared face), every other cell
// tested in ascending order. The GPU kernel counts all equal (i, j) pairs; the CPU variant stops at the
// first match per vertex of the first cell; the two agree on tetrahedra with distinct vertices, and the
// pair count is what is restated here. O(T^2): only for the small fixtures.
static int tet_common(const int* a, const int* b)
{
    int n = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) n += a[i] == b[j];
    return n;
}

long orc_face_neighbors3d_count(int nt, const int* tet)
{
    long total = 0;
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j)
            if (j != i && tet_common(tet + 4 * (size_t)i, tet + 4 * (size_t)j) == 3) ++total;
    return total;
}

void orc_face_neighbors3d_fill(int nt, const int* tet, int* offsets, int* idx)
{
    offsets[0] = 0;
    int k = 0;
    for (int i = 0; i < nt; ++i) {
        for (int j = 0; j < nt; ++j)
            if (j != i && tet_common(tet + 4 * (size_t)i, tet + 4 * (size_t)j) == 3) idx[k++] = j;
        offsets[i + 1] = k;
    }
}

// Quadrature tables: the decimal literals of src/integration/quadrature_formula_1d.cu:3-44,
// quadrature_formula_2d.cu:3-126 and quadrature_formula_3d.cu:3-171, kept digit for digit because the
// assembled values depend on them. The last barycentric coordinate is 1 - sum(others), as there.
int orc_quadrature(int kind, int index, double* out)
{
    if (kind == 1) {
        static const double e0[] = {0.0, 2.0};
        static const double e1[] = {-0.577350269189626, 1.0, 0.577350269189626, 1.0};
        static const double e2[] = {-0.774596669241483, 0.555555555555556, 0.0, 0.888888888888889,
                                    0.774596669241483, 0.555555555555556};
        static const double e3[] = {-0.861136311594053, 0.347854845137454, -0.339981043584856, 0.652145154862546,
                                    0.339981043584856, 0.652145154862546, 0.861136311594053, 0.347854845137454};
        static const double e4[] = {-0.906179845938664, 0.236926885056189, -0.538469310105683, 0.478628670499366,
                                    0.0, 0.568888888888889, 0.538469310105683, 0.478628670499366,
                                    0.906179845938664, 0.236926885056189};
        const double* t[] = {e0, e1, e2, e3, e4};
        const int np[] = {1, 2, 3, 4, 5};
        if (index < 0 || index > 4) return 0;
        for (int i = 0; i < 2 * np[index]; ++i) out[i] = t[index][i];
        return np[index];
    }
    if (kind == 2) {
        // (xi, eta, weight) triples
        static const double f0[] = {0.3333333333333333, 0.3333333333333333, 1.0};
        static const double f1[] = {0.1666666666666667, 0.1666666666666667, 0.3333333333333333,
                                    0.6666666666666667, 0.1666666666666667, 0.3333333333333333,
                                    0.1666666666666667, 0.6666666666666667, 0.3333333333333333};
        static const double f2[] = {0.3333333333333333, 0.3333333333333333, -0.5625, 0.6, 0.2, 0.5208333333333333,
                                    0.2, 0.6, 0.5208333333333333, 0.2, 0.2, 0.5208333333333333};
        static const double f3[] = {0.816847572980459, 0.091576213509771, 0.109951743655322,
                                    0.091576213509771, 0.816847572980459, 0.109951743655322,
                                    0.091576213509771, 0.091576213509771, 0.109951743655322,
                                    0.445948490915965, 0.108103018168070, 0.223381589678011,
                                    0.108103018168070, 0.445948490915965, 0.223381589678011,
                                    0.445948490915965, 0.445948490915965, 0.223381589678011};
        static const double f4[] = {0.3333333333333333, 0.3333333333333333, 0.225,
                                    0.101286507323456, 0.797426985353087, 0.125939180544827,
                                    0.797426985353087, 0.101286507323456, 0.125939180544827,
                                    0.101286507323456, 0.101286507323456, 0.125939180544827,
                                    0.470142064105115, 0.059715871789770, 0.132394152788506,
                                    0.059715871789770, 0.470142064105115, 0.132394152788506,
                                    0.470142064105115, 0.470142064105115, 0.132394152788506};
        static const double f5[] = {0.437525248383384, 0.124949503233232, 0.205950504760887,
                                    0.124949503233232, 0.437525248383384, 0.205950504760887,
                                    0.437525248383384, 0.437525248383384, 0.205950504760887,
                                    0.797112651860071, 0.165409927389841, 0.063691414286223,
                                    0.165409927389841, 0.797112651860071, 0.063691414286223,
                                    0.797112651860071, 0.037477420750088, 0.063691414286223,
                                    0.037477420750088, 0.797112651860071, 0.063691414286223,
                                    0.165409927389841, 0.037477420750088, 0.063691414286223,
                                    0.037477420750088, 0.165409927389841, 0.063691414286223};
        static const double f6[] = {0.873821971016996, 0.063089014491502, 0.050844906370207,
                                    0.063089014491502, 0.873821971016996, 0.050844906370207,
                                    0.063089014491502, 0.063089014491502, 0.050844906370207,
                                    0.501426509658179, 0.249286745170910, 0.116786275726379,
                                    0.249286745170910, 0.501426509658179, 0.116786275726379,
                                    0.249286745170910, 0.249286745170910, 0.116786275726379,
                                    0.636502499121399, 0.310352451033785, 0.082851075618374,
                                    0.310352451033785, 0.636502499121399, 0.082851075618374,
                                    0.636502499121399, 0.053145049844816, 0.082851075618374,
                                    0.053145049844816, 0.636502499121399, 0.082851075618374,
                                    0.310352451033785, 0.053145049844816, 0.082851075618374,
                                    0.053145049844816, 0.310352451033785, 0.082851075618374};
        static const double f7[] = {0.333333333333333, 0.333333333333333, -0.149570044467670,
                                    0.479308067841923, 0.260345966079038, 0.175615257433204,
                                    0.260345966079038, 0.479308067841923, 0.175615257433204,
                                    0.260345966079038, 0.260345966079038, 0.175615257433204,
                                    0.869739794195598, 0.065130102902216, 0.053347235608839,
                                    0.065130102902216, 0.869739794195598, 0.053347235608839,
                                    0.065130102902216, 0.065130102902216, 0.053347235608839,
                                    0.638444188569809, 0.312865496004875, 0.077113760890257,
                                    0.312865496004875, 0.638444188569809, 0.077113760890257,
                                    0.638444188569809, 0.048690315425316, 0.077113760890257,
                                    0.048690315425316, 0.638444188569809, 0.077113760890257,
                                    0.312865496004875, 0.048690315425316, 0.077113760890257,
                                    0.048690315425316, 0.312865496004875, 0.077113760890257};
        const double* t[] = {f0, f1, f2, f3, f4, f5, f6, f7};
        const int np[] = {1, 3, 4, 6, 7, 9, 12, 13};
        if (index < 0 || index > 7) return 0;
        for (int i = 0; i < np[index]; ++i) {
            const double x = t[index][3 * i], y = t[index][3 * i + 1];
            out[4 * i + 0] = x;
            out[4 * i + 1] = y;
            out[4 * i + 2] = 1.0 - x - y;
            out[4 * i + 3] = t[index][3 * i + 2];
        }
        return np[index];
    }
    if (kind == 3) {
        // (l1, l2, l3, weight); only the two lowest tet rules are used by the Poisson3D paths.
        static const double c0[] = {0.128515707071765, 0.139571690967945, 0.621705865573422, 0.213074006319707,
                                    0.608054478991729, 0.137352730077163, 0.145898694726675, 0.225533192268079,
                                    0.161734727754746, 0.159340192246828, 0.160612296342170, 0.342484823577903,
                                    0.137448101688511, 0.615380511289739, 0.135300542338320, 0.218907977834312};
        static const double c1[] = {0.0701026973651683, 0.166660699730426, 0.596575741463645, 0.180396407632724,
                                    0.168695371477743, 0.584652626698847, 0.168695218708539, 0.193316294293662,
                                    0.418175789241166, 0.0803037256707293, 0.418101348156118, 0.133594798870359,
                                    0.596576288848820, 0.166672635105983, 0.0700786026978242, 0.180371079146698,
                                    0.164808547352848, 0.0851134715186107, 0.164808684755350, 0.201016720944591,
                                    0.0731475667668616, 0.440339136918213, 0.0731570614497775, 0.111304699111967};
        const double* t[] = {c0, c1};
        const int np[] = {4, 6};
        if (index < 0 || index > 1) return 0;
        for (int i = 0; i < np[index]; ++i) {
            const double a = t[index][4 * i], b = t[index][4 * i + 1], c = t[index][4 * i + 2];
            out[5 * i + 0] = a;
            out[5 * i + 1] = b;
            out[5 * i + 2] = c;
            out[5 * i + 3] = 1.0 - a - b - c;
            out[5 * i + 4] = t[index][4 * i + 3];
        }
        return np[index];
    }
    return 0;
}

}  // extern "C"
