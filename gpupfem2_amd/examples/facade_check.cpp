// Exercises the C++ facade's plugin point and classes on one small problem and prints one line per check
// ("ok <name> <numbers>" or "FAIL <name> ..."): a user Preconditioner subclass (identity, applied with a
// device copy on the library's stream) driving SolverCG must reproduce the unpreconditioned solve
// exactly, also when it applies itself through the LinearAlgebra it was given; LinearAlgebra's dense
// wrappers (dot, norm, triangular solve, gemv, device scalars) and sparseMV must match host arithmetic;
// PreconditionerIC (natural and multicolour) and GMRES(restart) must converge; particles seed, advect and
// project; SimulationParameters defaults and quadrature tables are the reference's.
//
//   facade_check mesh.dat
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpupfem2.hpp"

namespace {
int failures = 0;
void report(bool ok, const char* name, double a = 0, double b = 0)
{
    std::printf("%s %s %.17g %.17g\n", ok ? "ok" : "FAIL", name, a, b);
    std::fflush(stdout);
    if (!ok) ++failures;
}

class IdentityPreconditioner : public Preconditioner {
public:
    IdentityPreconditioner(int n_, const LinearAlgebra* LA_) : Preconditioner(n_, LA_) {}
    int inits = 0, applies = 0;
    void initialize(const SparseMatrixCSR& csrMatrix) override
    {
        if (csrMatrix.getRows() == n) ++inits;
    }
    void applyPreconditioner(double* dest, const double* src) override
    {
        ++applies;
        if (dest != src) gpupfem2::check(pfem_memcpy_d2d(gpupfem2::context(), dest, src, sizeof(double) * n));
    }
};

// the identity again, but through the LinearAlgebra it was given, as a reference plugin would use it:
// dest = 1 * [src] (an n x 1 matrix) * [1] + 0 * dest, scalars in device memory
class GemvIdentityPreconditioner : public Preconditioner {
public:
    GemvIdentityPreconditioner(int n_, const LinearAlgebra* LA_) : Preconditioner(n_, LA_) {}
    void initialize(const SparseMatrixCSR& csrMatrix) override
    {
        (void)csrMatrix;
        if (!scalars.data) scalars = deviceVector<double>(std::vector<double>{1.0, 0.0});
    }
    void applyPreconditioner(double* dest, const double* src) override
    {
        LA->generalMV(src, scalars.data, dest, scalars.data, scalars.data + 1, n, 1, n);
    }
    deviceVector<double> scalars;
};

class NoPreconditioner : public BuiltinPreconditioner {
public:
    NoPreconditioner(int n_, const LinearAlgebra* LA_) : BuiltinPreconditioner(n_, LA_, PFEM_PRECOND_NONE, 0) {}
};

class Integrator : public NumericalIntegrator2D {
public:
    explicit Integrator(const Mesh2D& m) : NumericalIntegrator2D(m) {}
    void assemble(SparseMatrixCSR& a, deviceVector<double>& rhs) { assemblePoisson(a, rhs, 1.0, 2, 1.0, 1); }
};
}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 2) return EXIT_FAILURE;
    Mesh2D mesh;
    const bool loaded = mesh.loadMeshFromFile(argv[1]);
    report(loaded, "load");
    if (!loaded) return EXIT_FAILURE;
    report(mesh.getVertices().size > 0 && mesh.getCells().size > 0, "mesh_sizes", mesh.getVertices().size,
           mesh.getCells().size);
    report(!Mesh2D().loadMeshFromFile("/nonexistent/mesh.dat"), "load_missing_file");
    const int n = mesh.getVertices().size;

    SparseMatrixCSR a(mesh);
    deviceVector<double> rhs;
    rhs.allocate(n);
    Integrator(mesh).assemble(a, rhs);
    {
        // zero Dirichlet values on the mesh's lowest-y nodes
        const auto v = mesh.getHostVertices();
        double ymin = 1e300;
        for (const auto& p : v) ymin = std::fmin(ymin, p.y);
        std::vector<DirichletNode> bc;
        for (unsigned i = 0; i < v.size(); ++i)
            if (std::fabs(v[i].y - ymin) < 2e-6) bc.push_back({i, 0.0});
        DirichletBCs bcs;
        bcs.setEliminationMode(PFEM_DIRICHLET_MASK);
        bcs.setupDirichletBCs(bc);
        bcs.applyBCs(a, rhs);
        report(!bc.empty(), "dirichlet_nodes", (double)bc.size());
    }
    LinearAlgebra LA;
    {
        // the dense wrappers, scalars in device memory, against host arithmetic
        const auto hb = rhs.copyToHost();
        double dd = 0.0, scale = 0.0;
        for (double v : hb) {
            dd += v * v;
            scale += std::fabs(v * v);
        }
        deviceVector<double> res(std::vector<double>{0.0, 0.0});
        LA.dot(rhs, rhs, res.data);
        LA.normSquared(rhs.data, res.data + 1, n);
        const auto hr = res.copyToHost();
        report(std::fabs(hr[0] - dd) <= 1e-13 * scale && std::fabs(hr[1] - std::sqrt(dd)) <= 1e-13 * std::sqrt(dd),
               "la_dot_nrm2", hr[0], hr[1]);
        // U x = b for a 5 x 5 upper triangle stored with leading dimension 7, then y = 2 V z + 1 y
        const int k = 5, lda = 7, m = 1000;
        std::vector<double> u(lda * k, 1e300), b(k), vm((size_t)m * k), z(k), y(m);
        for (int j = 0; j < k; ++j) {
            for (int i = 0; i <= j; ++i) u[i + j * lda] = i == j ? 2.0 + j : 0.25 * (i + 1) - 0.1 * j;
            b[j] = 1.0 + 0.5 * j;
            z[j] = 0.3 * j - 0.4;
            for (int i = 0; i < m; ++i) vm[i + (size_t)j * m] = std::sin(0.01 * i + j);
        }
        for (int i = 0; i < m; ++i) y[i] = std::cos(0.02 * i);
        std::vector<double> hx = b;
        for (int j = k - 1; j >= 0; --j) {
            hx[j] /= u[j + j * lda];
            for (int i = 0; i < j; ++i) hx[i] -= u[i + j * lda] * hx[j];
        }
        deviceVector<double> du(u), dx(b), dv(vm), dz(z), dy(y), ab(std::vector<double>{2.0, 1.0});
        LA.solveGeneralTriangularSystem(du.data, dx.data, k, lda);
        LA.generalMV(dv.data, dz.data, dy.data, ab.data, ab.data + 1, m, k, m);
        const auto gx = dx.copyToHost(), gy = dy.copyToHost();
        double ex = 0.0, ey = 0.0;
        for (int j = 0; j < k; ++j) ex = std::fmax(ex, std::fabs(gx[j] - hx[j]));
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += vm[i + (size_t)j * m] * z[j];
            ey = std::fmax(ey, std::fabs(gy[i] - (2.0 * s + y[i])));
        }
        report(ex < 1e-14 && ey < 1e-13, "la_trsv_gemv", ex, ey);
    }

    deviceVector<double> x0, x1;
    x0.allocate(n);
    x1.allocate(n);
    x0.clearValues();
    x1.clearValues();
    NoPreconditioner none(n, &LA);
    SolverCG ref(1e-10, 5000, &LA, &none);
    ref.init(a);
    const bool ok0 = ref.solve(a, x0, rhs);
    IdentityPreconditioner ident(n, &LA);
    SolverCG plug(1e-10, 5000, &LA, &ident);
    plug.init(a);
    const bool ok1 = plug.solve(a, x1, rhs);
    const auto h0 = x0.copyToHost(), h1 = x1.copyToHost();
    double diff = 0.0;
    for (int i = 0; i < n; ++i) diff = std::fmax(diff, std::fabs(h0[i] - h1[i]));
    report(ok0 && ok1 && ref.lastStats().iterations == plug.lastStats().iterations && diff == 0.0, "custom_preconditioner",
           plug.lastStats().iterations, diff);
    report(ident.inits == 1 && ident.applies >= plug.lastStats().iterations, "custom_preconditioner_calls", ident.inits,
           ident.applies);
    {
        deviceVector<double> x2;
        x2.allocate(n);
        x2.clearValues();
        GemvIdentityPreconditioner gid(n, &LA);
        SolverCG cg(1e-10, 5000, &LA, &gid);
        cg.init(a);
        const bool ok2 = cg.solve(a, x2, rhs);
        const auto h2 = x2.copyToHost();
        double d2 = 0.0;
        for (int i = 0; i < n; ++i) d2 = std::fmax(d2, std::fabs(h0[i] - h2[i]));
        report(ok2 && cg.lastStats().iterations == ref.lastStats().iterations && d2 == 0.0, "plugin_using_linear_algebra",
               cg.lastStats().iterations, d2);
    }
    {
        // A x0 reproduces the right-hand side to the solve's tolerance
        deviceVector<double> ax;
        ax.allocate(n);
        LA.sparseMV(a, x0.data, ax.data);
        const auto hax = ax.copyToHost(), hb = rhs.copyToHost();
        double e = 0.0, nb = 0.0;
        for (int i = 0; i < n; ++i) {
            e += (hax[i] - hb[i]) * (hax[i] - hb[i]);
            nb += hb[i] * hb[i];
        }
        report(std::sqrt(e) <= 1e-8 + 1e-9 * std::sqrt(nb), "la_sparse_mv", std::sqrt(e), std::sqrt(nb));
    }

    for (int ordering : {PFEM_ORDER_NATURAL, PFEM_ORDER_MULTICOLOR}) {
        deviceVector<double> x;
        x.allocate(n);
        x.clearValues();
        PreconditionerIC ic(a, &LA, ordering);
        SolverCG cg(1e-10, 5000, &LA, &ic);
        cg.init(a);
        const bool ok = cg.solve(a, x, rhs);
        const auto hx = x.copyToHost();
        double d = 0.0;
        for (int i = 0; i < n; ++i) d = std::fmax(d, std::fabs(hx[i] - h0[i]));
        report(ok && d < 1e-6, ordering ? "ic0_multicolour" : "ic0_natural", cg.lastStats().iterations, d);
    }
    {
        // the multicolour IC0 factor applied from its 16-bit copy, FP64 Krylov (PFEM_PRECISION_MIXED_PRECOND)
        deviceVector<double> x;
        x.allocate(n);
        x.clearValues();
        PreconditionerIC ic(a, &LA, PFEM_ORDER_MULTICOLOR);
        ic.setFactorBits(a, 16);
        SolverCG cg(1e-10, 5000, &LA, &ic);
        cg.setPrecision(PFEM_PRECISION_MIXED_PRECOND);
        cg.init(a);
        const bool ok = cg.solve(a, x, rhs);
        const auto hx = x.copyToHost();
        double d = 0.0;
        for (int i = 0; i < n; ++i) d = std::fmax(d, std::fabs(hx[i] - h0[i]));
        report(ok && d < 1e-6, "ic0_half_factor", cg.lastStats().iterations, d);
    }
    {
        deviceVector<double> x;
        x.allocate(n);
        x.clearValues();
        PreconditionerJacobi jac(n, &LA);
        SolverGMRES gm(1e-10, 5000, &LA, &jac, 30);
        gm.init(a);
        const bool ok = gm.solve(a, x, rhs);
        report(ok, "gmres_restarted", gm.lastStats().iterations);
    }

    {
        // two right-hand sides at once (rhs and 2 rhs): each column equals the single solve's result
        const auto hb = rhs.copyToHost();
        std::vector<double> hb2(2 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            hb2[i] = hb[i];
            hb2[n + i] = 2.0 * hb[i];
        }
        deviceVector<double> B(hb2), X(std::vector<double>(2 * (size_t)n, 0.0));
        PreconditionerJacobi jac(n, &LA);
        SolverCG cg(1e-10, 5000, &LA, &jac);
        cg.init(a);
        const bool ok = cg.solveMulti(a, X, B, 2);
        const auto hx = X.copyToHost();
        double d = 0.0;
        for (int i = 0; i < n; ++i) d = std::fmax(d, std::fmax(std::fabs(hx[i] - h0[i]), std::fabs(hx[n + i] - 2.0 * h0[i])));
        report(ok && d < 1e-6 && cg.multiStats().size() == 2, "cg_multi_rhs", cg.multiStats()[0].iterations, d);
    }

    ParticleHandler2D ph(&mesh, 2);
    ph.seedParticles();
    report(ph.getParticleCount() == 4 * mesh.getCells().size, "seed_centroids", ph.getParticleCount());
    deviceVector<double> u0(std::vector<double>(n, 1.0)), u1(std::vector<double>(n, 0.5));
    deviceVector<double*> vel(std::vector<double*>{u0.data, u1.data});
    ph.initParticleVelocity(vel);
    ph.advectParticles(vel, 1e-4, 3);
    deviceVector<double> p0, p1;
    p0.allocate(n);
    p1.allocate(n);
    deviceVector<double*> proj(std::vector<double*>{p0.data, p1.data});
    ph.projectVelocityOntoGrid(proj);
    const auto hp0 = p0.copyToHost(), hp1 = p1.copyToHost();
    double err = 0.0;
    for (int i = 0; i < n; ++i)
        if (std::isfinite(hp0[i])) err = std::fmax(err, std::fmax(std::fabs(hp0[i] - 1.0), std::fabs(hp1[i] - 0.5)));
    report(err < 1e-12, "constant_velocity_projection", err);

    SimulationParameters prm;
    prm.setDefaultParameters();
    report(prm.rho == 1.0 && prm.mu == 1.0 && prm.dt == 0.01 && prm.tolerance == 1e-8 && prm.maxIterations == 1000 &&
               prm.particleAdvectionSubsteps == 3 && prm.bodyBoundaryID == -1,
           "simulation_parameters_defaults");
    const auto q = createFaceQuadratureFormula(1);
    double wsum = 0.0;
    for (const auto& g : q) wsum += g.weight;
    report(q.size() == 3 && std::fabs(wsum - 1.0) < 1e-14, "face_quadrature_1", (double)q.size(), wsum);
    const auto e = createEdgeQuadratureFormula(1);
    report(e.size() == 2 && std::fabs(e[0].weight + e[1].weight - 2.0) < 1e-14, "edge_quadrature_1", (double)e.size());

    if (argc > 2) {
        // Mesh3D with the reference's default arguments builds the face-neighbour lists: every listed pair
        // shares exactly three vertices, lists ascend, the relation is symmetric, at most 4 per cell
        Mesh3D box;
        const bool ok3 = box.loadMeshFromFile(argv[2]);
        report(ok3, "load3d");
        if (ok3) {
            std::vector<int> off, idx, tet(4 * (size_t)box.numCells());
            box.getHostCellNeighbors(off, idx);
            gpupfem2::check(pfem_mesh3d_download(box.handle(), 1, tet.data()));
            bool good = box.getCellNeighborsOffsets().size == box.numCells() + 1 &&
                        box.getCellNeighborIndices().size == (int)idx.size() && off.back() == (int)idx.size();
            for (int c = 0; c < box.numCells() && good; ++c) {
                good = off[c + 1] - off[c] <= 4;
                for (int k = off[c]; k < off[c + 1] && good; ++k) {
                    const int d = idx[k];
                    int pairs = 0;
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j) pairs += tet[4 * (size_t)c + i] == tet[4 * (size_t)d + j];
                    good = pairs == 3 && (k == off[c] || idx[k - 1] < d) &&
                           std::binary_search(idx.begin() + off[d], idx.begin() + off[d + 1], c);
                }
            }
            report(good, "mesh3d_face_neighbors", (double)idx.size(), box.numCells());
        }
    }

    std::printf("%s %d failures\n", failures ? "FAIL" : "ok", failures);
    return failures ? EXIT_FAILURE : EXIT_SUCCESS;
}
