// The Cylinder2D case (cases/Cylinder2D/main.cu:635-880) on the library, with the reference's inputs:
// the mesh import, the boundary IDs of kSetEdgeBoundaryIDs (:35-70), the Dirichlet lists chosen by
// coordinates in the same if/else order (:668-690), particle seeding at level 2 (one particle per subcell
// centroid), the SimulationParameters of the case (dt 0.001, mu 0.001, Scheme A, thickness 0.1, body
// boundary 3), CG with Jacobi for the pressure and the correction and non-restarted GMRES with Jacobi
// for the prediction, absolute tolerance 1e-8. The time loop body (projection, two outer iterations,
// particle correction, loads) is one pfem_pfem2_step per step. Loads go to Forces.csv as "Time;Cx;Cy"
// (boundaryLoadsCalculator, :378-445); the fields go to solution<step>.vtu every outputFrequency steps.
//
//   cylinder2d [mesh.dat] [steps] [output directory] [output frequency]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "gpupfem2.hpp"

int main(int argc, char* argv[])
{
    const std::string meshFile = argc > 1 ? argv[1] : "CylinderMesh3.dat";
    const int steps = argc > 2 ? std::atoi(argv[2]) : 7500;
    const std::string outDir = argc > 3 ? std::string(argv[3]) + "/" : "";
    const double eps = 2e-6;   // CONSTANTS::DOUBLE_MIN (src/common/constants.h)
    const double h = 0.41, l = 2.2, cx = 0.2, cy = 0.2, radius = 0.05;
    try {
        Mesh2D mesh;
        if (!mesh.loadMeshFromFile(meshFile)) return EXIT_FAILURE;
        mesh.setEdgeBoundaryIDs(0.0, l, 0.0, h, 0, 1, 2, 2, cx, cy, radius, 1e-3, 3);

        ParticleHandler2D particleHandler(&mesh, 2);
        particleHandler.seedParticles();

        // velocity: parabolic inflow (U_max = 1.5), no slip on the walls and the cylinder; pressure: zero
        // at the outlet
        std::vector<int> bcu, bcp;
        std::vector<double> bcu0, bcu1, bcpv;
        const auto vertices = mesh.getHostVertices();
        for (int i = 0; i < (int)vertices.size(); ++i) {
            const Point2& node = vertices[i];
            if (std::fabs(node.x) < eps) {
                bcu.push_back(i);
                bcu0.push_back(4 * 1.5 * node.y * (h - node.y) / (h * h));
                bcu1.push_back(0.0);
            } else if (std::fabs(node.x - l) < eps) {
                bcp.push_back(i);
                bcpv.push_back(0.0);
            } else if (std::fabs(node.y) < eps || std::fabs(node.y - h) < eps ||
                       std::fabs(std::hypot(node.x - cx, node.y - cy) - radius) < 0.001) {
                bcu.push_back(i);
                bcu0.push_back(0.0);
                bcu1.push_back(0.0);
            }
        }

        SimulationParameters params;
        params.setDefaultParameters();
        params.dt = 0.001;
        params.mu = 0.001;
        params.simulationScheme = 0;
        params.tFinal = 7.5;
        params.outputFrequency = argc > 4 ? std::atoi(argv[4]) : 500;
        params.calculateLoads = 1;
        params.bodyBoundaryID = 3;
        params.thickness = 0.1;

        pfem_step_config cfg{};
        cfg.prm.rho = params.rho;
        cfg.prm.mu = params.mu;
        cfg.prm.dt = params.dt;
        cfg.prm.scheme = params.simulationScheme;
        cfg.prm.face_quad = 1;              // createFaceQuadratureFormula(1)
        cfg.prm.edge_quad = 1;              // createEdgeQuadratureFormula(1)
        cfg.prm.traction_mask = 1 << 1;     // the viscous traction edge term on boundary 1 (the outlet)
        cfg.prm.normal[1][0] = 1.0;
        cfg.outer_iterations = 2;
        cfg.substeps = params.particleAdvectionSubsteps;
        cfg.pred_solver = PFEM_SOLVER_GMRES;
        cfg.pred_precond = PFEM_PRECOND_JACOBI;
        cfg.pres_solver = PFEM_SOLVER_CG;
        cfg.pres_precond = PFEM_PRECOND_JACOBI;
        cfg.corr_solver = PFEM_SOLVER_CG;
        cfg.corr_precond = PFEM_PRECOND_JACOBI;
        cfg.tol = params.tolerance;
        cfg.tol_mode = PFEM_TOL_ABSOLUTE;
        cfg.maxit = params.maxIterations;
        cfg.restart = 0;
        cfg.dirichlet_mode = PFEM_DIRICHLET_PARITY;
        cfg.body_id = params.bodyBoundaryID;
        cfg.body_cx = cx;
        cfg.body_cy = cy;
        cfg.force_coeff = 2.0 / (params.rho * params.meanVelocity * params.meanVelocity * params.thickness);
        cfg.precision = PFEM_PRECISION_FP64;
        cfg.pres_ordering = PFEM_ORDER_NATURAL;

        pfem_pfem2 stepper = nullptr;
        gpupfem2::check(pfem_pfem2_create(mesh.handle(), particleHandler.handle(), &cfg, (int)bcu.size(), bcu.data(),
                                          bcu0.data(), bcu1.data(), (int)bcp.size(), bcp.data(), bcpv.data(),
                                          &stepper));
        // the work between the solves replays from HIP graphs from the second step on (the same bits)
        gpupfem2::check(pfem_pfem2_set_graphs(stepper, 1));
        double* field[5];
        for (int k = 0; k < 5; ++k) gpupfem2::check(pfem_pfem2_field(stepper, k, &field[k]));
        gpupfem2::check(pfem_particles_init_velocity(particleHandler.handle(), field[0], field[1]));

        const char* names[5] = {"velX", "velY", "pressure", "velPredictionX", "velPredictionY"};
        auto exportFields = [&](int step) {
            const std::string file = outDir + "solution" + std::to_string(step) + ".vtu";
            gpupfem2::check(pfem_export_vtk(mesh.handle(), file.c_str(), 5, names, field));
        };
        exportFields(0);

        std::ofstream forces(outDir + "Forces.csv");
        forces << "Time;Cx;Cy" << std::endl;
        GpuTimer timer;
        for (int step = 1; step <= steps; ++step) {
            const double t = step * params.dt;
            timer.start();
            double f[2];
            gpupfem2::check(pfem_pfem2_step(stepper, f));
            if (step % params.outputFrequency == 0) exportFields(step);
            forces << t << ";" << f[0] << ";" << f[1] << std::endl;
            timer.stop("Time of a simulation step");
        }
        pfem_pfem2_destroy(stepper);
        std::printf("cylinder2d: %d steps, %d particles\n", steps, particleHandler.getParticleCount());
    } catch (const gpupfem2::Error& e) {
        std::fprintf(stderr, "cylinder2d: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
