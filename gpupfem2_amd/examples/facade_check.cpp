// Exercises the C++ facade's plugin point and classes on one small problem and prints one line per check
// ("ok <name> <numbers>" or "FAIL <name> ..."): a user Preconditioner subclass (identity, applied with a
// device copy on the library's stream) driving SolverCG must reproduce the unpreconditioned solve
// exactly; PreconditionerIC (natural and multicolour) and GMRES(restart) must converge; particles seed,
// advect and project; SimulationParameters defaults and quadrature tables are the reference's.
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
