// The Poisson3D case (cases/Poisson3D/main.cu:87-176) written against the C++ facade: the same call
// sequence -- tetrahedral mesh import, Dirichlet setup by coordinates on the six faces of the box,
// assembly with the 6-point cell rule, BCs, PCG with Jacobi, exports of the matrix, the solution, the
// right-hand side and the VTU file. The reference's integrand kernel (kIntegrateOverCell, :26-67) is the
// library operator behind NumericalIntegrator3D::assemblePoisson (coefficient 0.5,
// f = exp(-(x^2 + 10 y^2 + z^2))).
//
//   poisson3d [mesh.dat] [output directory]
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpupfem2.hpp"

class PoissonIntegrator : public NumericalIntegrator3D {
public:
    explicit PoissonIntegrator(const Mesh3D& mesh_) : NumericalIntegrator3D(mesh_) {}

    void assembleSystem(SparseMatrixCSR& csrMatrix, deviceVector<double>& rhsVector)
    {
        assemblePoisson(csrMatrix, rhsVector, 0.5, 0, 0.0, 1);
    }
};

int main(int argc, char* argv[])
{
    const std::string meshFile = argc > 1 ? argv[1] : "BoxMesh.dat";
    const std::string outDir = argc > 2 ? std::string(argv[2]) + "/" : "";
    const double eps = 2e-6;   // CONSTANTS::DOUBLE_MIN (src/common/constants.h)
    GpuTimer timer;

    timer.start();
    Mesh3D mesh;
    if (!mesh.loadMeshFromFile(meshFile)) return EXIT_FAILURE;
    timer.stop("Mesh import");

    const int problemSize = mesh.getVertices().size;

    DirichletBCs bcs;
    timer.start();
    {
        std::vector<DirichletNode> hostBcs;
        const auto vertices = mesh.getHostVertices();
        hostBcs.reserve(vertices.size() / 10);
        for (unsigned i = 0; i < vertices.size(); ++i) {
            const Point3& node = vertices[i];
            if (std::fabs(node.x - (-1.0)) < eps) hostBcs.push_back({i, -1.0});
            else if (std::fabs(node.x - 1.0) < eps) hostBcs.push_back({i, 1.0});
            else if (std::fabs(node.y) < eps) hostBcs.push_back({i, 0.0});
            else if (std::fabs(node.y - 1.0) < eps) hostBcs.push_back({i, 2.0});
            else if (std::fabs(node.z - (-0.25)) < eps) hostBcs.push_back({i, 3.0});
            else if (std::fabs(node.z - 0.25) < eps) hostBcs.push_back({i, -2.0});
        }
        bcs.setupDirichletBCs(hostBcs);
    }
    timer.stop("Boundary conditions setup");

    SparseMatrixCSR matrix(mesh);
    PoissonIntegrator integrator(mesh);

    deviceVector<double> rhsVector;
    rhsVector.allocate(problemSize);
    rhsVector.clearValues();

    timer.start();
    integrator.assembleSystem(matrix, rhsVector);
    bcs.applyBCs(matrix, rhsVector);
    timer.stop("Assembly of system and rhs");

    matrix.exportMatrix(outDir + "matrix.dat");

    deviceVector<double> solution;
    solution.allocate(problemSize);
    solution.clearValues();

    timer.start();
    LinearAlgebra LA;
    PreconditionerJacobi precond(problemSize, &LA);
    SolverCG cgSolver(1e-8, 1000, &LA, &precond);
    cgSolver.init(matrix);
    const bool ok = cgSolver.solve(matrix, solution, rhsVector);
    timer.stop("PCG solver");
    std::printf("PCG: converged %d, %d iterations\n", (int)ok, cgSolver.lastStats().iterations);

    solution.exportToFile(outDir + "solution.dat");
    rhsVector.exportToFile(outDir + "rhs.dat");

    DataExport3D dataExport(mesh);
    dataExport.addScalarDataVector(solution, "solution");
    dataExport.exportToVTK(outDir + "solution.vtu");

    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}
