// The PoiseuilleFlow2D case (cases/PoiseuilleFlow2D/main.cu:496-736) on the library, with the reference's
// inputs: the channel mesh [-5, 5] x [0, 1], the boundary IDs of kSetEdgeBoundaryIDs (:48-79) with the
// outward normals of normalVector (:32-46) and the viscous traction term on every boundary edge, the
// pressure-driven option of the Dirichlet setup (p = 10 at the inlet, 0 at the outlet, no slip on the
// walls, :536-549), particle seeding at level 2, the case's SimulationParameters (dt 0.01, Scheme A, the
// remaining defaults) and one outer iteration of GMRES (prediction), CG (pressure) and CG (correction), all
// with Jacobi. The time loop body is one pfem_pfem2_step per step; the fields go to solution<step>.vtu
// every outputFrequency steps.
//
//   poiseuille2d [mesh.dat] [steps] [output directory] [output frequency]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpupfem2.hpp"

int main(int argc, char* argv[])
{
    const std::string meshFile = argc > 1 ? argv[1] : "ChannelMesh.dat";
    const std::string outDir = argc > 3 ? std::string(argv[3]) + "/" : "";
    const double eps = 2e-6;   // CONSTANTS::DOUBLE_MIN (src/common/constants.h)
    try {
        Mesh2D mesh;
        if (!mesh.loadMeshFromFile(meshFile)) return EXIT_FAILURE;
        mesh.setEdgeBoundaryIDs(-5.0, 5.0, 0.0, 1.0, 0, 1, 2, 3);

        ParticleHandler2D particleHandler(&mesh, 2);
        particleHandler.seedParticles();

        std::vector<int> bcu, bcp;
        std::vector<double> bcu0, bcu1, bcpv;
        const auto vertices = mesh.getHostVertices();
        for (int i = 0; i < (int)vertices.size(); ++i) {
            const Point2& node = vertices[i];
            if (std::fabs(node.x - (-5.0)) < eps) {
                bcp.push_back(i);
                bcpv.push_back(10.0);
            } else if (std::fabs(node.x - 5.0) < eps) {
                bcp.push_back(i);
                bcpv.push_back(0.0);
            } else if (std::fabs(node.y) < eps || std::fabs(node.y - 1.0) < eps) {
                bcu.push_back(i);
                bcu0.push_back(0.0);
                bcu1.push_back(0.0);
            }
        }

        SimulationParameters params;
        params.setDefaultParameters();
        params.dt = 0.01;
        params.tFinal = 5.0;
        params.simulationScheme = 0;
        params.outputFrequency = argc > 4 ? std::atoi(argv[4]) : 10;
        const int steps = argc > 2 ? std::atoi(argv[2]) : (int)std::lround(params.tFinal / params.dt) - 1;

        pfem_step_config cfg{};
        cfg.prm.rho = params.rho;
        cfg.prm.mu = params.mu;
        cfg.prm.dt = params.dt;
        cfg.prm.scheme = params.simulationScheme;
        cfg.prm.face_quad = 1;
        cfg.prm.edge_quad = 1;
        cfg.prm.traction_mask = 0xF;         // every boundary ID carries the traction edge term
        const double normals[4][2] = {{-1.0, 0.0}, {1.0, 0.0}, {0.0, -1.0}, {0.0, 1.0}};
        for (int b = 0; b < 4; ++b) {
            cfg.prm.normal[b][0] = normals[b][0];
            cfg.prm.normal[b][1] = normals[b][1];
        }
        cfg.outer_iterations = 1;
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
        cfg.body_id = -1;
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
        GpuTimer timer;
        for (int step = 1; step <= steps; ++step) {
            timer.start();
            gpupfem2::check(pfem_pfem2_step(stepper, nullptr));
            if (step % params.outputFrequency == 0) exportFields(step);
            timer.stop("Time of a simulation step");
        }
        pfem_pfem2_destroy(stepper);
        std::printf("poiseuille2d: %d steps, %d particles\n", steps, particleHandler.getParticleCount());
    } catch (const gpupfem2::Error& e) {
        std::fprintf(stderr, "poiseuille2d: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
