// C ABI: CSR matrices, Dirichlet boundary conditions, VTK export, SimulationParameters defaults.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>

#include "device.hpp"
#include "internal.hpp"
#include "solver.hpp"

using namespace pfem;

namespace pfem {
__global__ void k_to32_vals(long n, float* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}
}  // namespace pfem

static pfem_csr_t* new_csr(pfem_ctx ctx, std::shared_ptr<Pattern> pat)
{
    auto a = std::make_unique<pfem_csr_t>();
    a->ctx = ctx;
    a->pat = std::move(pat);
    a->val.alloc(a->pat->nnz);
    zero_async(C(ctx), a->val.p, a->val.bytes());
    PFEM_HIP(hipStreamSynchronize(ctx->stream));
    return a.release();
}

extern "C" {

void pfem_simulation_parameters_default(pfem_simulation_parameters* p)
{
    // SimulationParameters::setDefaultParameters (src/parameters.cuh:30-50)
    p->rho = 1.0;
    p->mu = 1.0;
    p->tFinal = 10.0;
    p->dt = 0.01;
    p->particleAdvectionSubsteps = 3;
    p->tolerance = 1e-8;
    p->maxIterations = 1000;
    p->simulationScheme = 0;
    p->meshFileName = "";
    p->outputFrequency = 1;
    p->exportParticles = 0;
    p->calculateLoads = 0;
    p->bodyBoundaryID = -1;
    p->thickness = 1.0;
    p->meanVelocity = 1.0;
}

int pfem_csr_create_from_mesh2d(pfem_mesh2d m, pfem_csr* out)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && out, "null argument");
        *out = new_csr(m->ctx, m->pat);
    });
}

int pfem_csr_create_from_mesh3d(pfem_mesh3d m, pfem_csr* out)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && out, "null argument");
        *out = new_csr(m->ctx, m->pat);
    });
}

int pfem_csr_create(pfem_ctx ctx, int n, int nnz, const int* rowptr, const int* col, const double* val, pfem_csr* out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(out && rowptr && col && n > 0 && nnz >= 0, "bad CSR arguments");
        PFEM_REQUIRE(rowptr[0] == 0 && rowptr[n] == nnz, "row offsets do not match nnz");
        for (int i = 0; i < n; ++i) {
            PFEM_REQUIRE(rowptr[i] <= rowptr[i + 1], "row offsets must be non-decreasing");
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
                PFEM_REQUIRE(col[k] >= 0 && col[k] < n, "column index out of range");
                PFEM_REQUIRE(k == rowptr[i] || col[k] > col[k - 1], "columns must be strictly ascending");
            }
        }
        auto pat = make_pattern(C(ctx), n, std::vector<int>(rowptr, rowptr + n + 1), std::vector<int>(col, col + nnz));
        pfem_csr_t* a = new_csr(ctx, pat);
        if (val) h2d(C(ctx), a->val.p, val, sizeof(double) * (size_t)nnz);
        *out = a;
    });
}

int pfem_csr_create_like(pfem_csr a, pfem_csr* out)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a && out, "null argument");
        *out = new_csr(a->ctx, a->pat);
    });
}

int pfem_csr_destroy(pfem_csr a)
{
    if (!a) return PFEM_OK;
    (void)hipStreamSynchronize(a->ctx->stream);
    delete a;
    return PFEM_OK;
}

int pfem_csr_info(pfem_csr a, int* n, int* nnz)
{
    if (!a) return PFEM_EINVAL;
    if (n) *n = a->pat->n;
    if (nnz) *nnz = (int)a->pat->nnz;
    return PFEM_OK;
}

int pfem_csr_pointers(pfem_csr a, int** d_rowptr, int** d_col, double** d_val)
{
    if (!a) return PFEM_EINVAL;
    if (d_rowptr) *d_rowptr = a->pat->rowptr.p;
    if (d_col) *d_col = a->pat->col.p;
    if (d_val) *d_val = a->val.p;
    return PFEM_OK;
}

int pfem_csr_clear_values(pfem_csr a)
{
    PFEM_TRY(a ? a->ctx : nullptr, zero_async(C(a->ctx), a->val.p, a->val.bytes()));
}

int pfem_csr_upload_values(pfem_csr a, const double* val)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a && val, "null argument");
        h2d(C(a->ctx), a->val.p, val, a->val.bytes());
    });
}

int pfem_csr_download(pfem_csr a, int* rowptr, int* col, double* val)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a, "null argument");
        if (rowptr) std::memcpy(rowptr, a->pat->h_rowptr.data(), sizeof(int) * (a->pat->n + 1));
        if (col) std::memcpy(col, a->pat->h_col.data(), sizeof(int) * a->pat->nnz);
        if (val) d2h_sync(C(a->ctx), val, a->val.p, a->val.bytes());
    });
}

int pfem_spmv(pfem_csr a, const double* d_x, double* d_y)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a && d_x && d_y, "null argument");
        launch_spmv(C(a->ctx), *a->pat, a->val.p, d_x, d_y);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_dot(pfem_ctx ctx, int n, const double* d_x, const double* d_y, double* d_res)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(ctx && d_res && n >= 0 && (n == 0 || (d_x && d_y)), "invalid argument");
        launch_dot(C(ctx), n, d_x, d_y, d_res);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_nrm2(pfem_ctx ctx, int n, const double* d_x, double* d_res)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(ctx && d_res && n >= 0 && (n == 0 || d_x), "invalid argument");
        launch_nrm2(C(ctx), n, d_x, d_res);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_trsv_upper(pfem_ctx ctx, int n, const double* d_a, int lda, double* d_x)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(ctx && n >= 0 && lda >= std::max(1, n) && (n == 0 || (d_a && d_x)), "invalid argument");
        launch_trsv_upper(C(ctx), n, d_a, lda, d_x);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_gemv(pfem_ctx ctx, int m, int n, const double* d_alpha, const double* d_a, int lda, const double* d_x,
              const double* d_beta, double* d_y)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(ctx && m >= 0 && n >= 0 && lda >= std::max(1, m), "invalid argument");
        PFEM_REQUIRE(m == 0 || n == 0 || (d_alpha && d_a && d_x && d_beta && d_y), "null argument");
        launch_gemv(C(ctx), m, n, d_alpha, d_a, lda, d_x, d_beta, d_y);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_csr_refresh_fp32(pfem_csr a)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a, "null argument");
        if (!a->val32.p) a->val32.alloc(a->pat->nnz);
        hipLaunchKernelGGL(k_to32_vals, dim3(blocks_for(a->pat->nnz)), dim3(kBlock), 0, a->ctx->stream,
                           (long)a->pat->nnz, a->val32.p, a->val.p);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_spmv_fp32(pfem_csr a, const float* d_x, float* d_y)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a && d_x && d_y, "null argument");
        PFEM_REQUIRE(a->val32.p, "call pfem_csr_refresh_fp32 first");
        launch_spmv32(C(a->ctx), *a->pat, a->val32.p, d_x, d_y);
        PFEM_HIP(hipGetLastError());
    });
}

int pfem_dirichlet_create(pfem_ctx ctx, int n, int nbc, const int* node, const double* value, pfem_dirichlet* out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(out && n > 0 && nbc >= 0, "bad arguments");
        PFEM_REQUIRE(nbc == 0 || (node && value), "null BC arrays");
        auto bc = std::make_unique<pfem_dirichlet_t>();
        bc->ctx = ctx;
        bc->n = n;
        bc->nbc = nbc;
        bc->h_node.assign(node, node + nbc);
        std::vector<int> map(n, -1);
        // DirichletBCs::setupNodeMap (Dirichlet_bcs.cu:61-71): node -> BC index (the last one for a
        // repeated node)
        for (int k = 0; k < nbc; ++k) {
            PFEM_REQUIRE(node[k] >= 0 && node[k] < n, "BC node out of range");
            map[node[k]] = k;
        }
        Context* c = C(ctx);
        upload(c, bc->node, bc->h_node);
        upload(c, bc->map, map);
        std::vector<double> v(value, value + nbc);
        upload(c, bc->value, v);
        bc->num.alloc(std::max(nbc, 1));
        bc->den.alloc(std::max(nbc, 1));
        *out = bc.release();
    });
}

int pfem_dirichlet_destroy(pfem_dirichlet bc)
{
    if (!bc) return PFEM_OK;
    (void)hipStreamSynchronize(bc->ctx->stream);
    delete bc;
    return PFEM_OK;
}

int pfem_dirichlet_set_values(pfem_dirichlet bc, const double* value)
{
    PFEM_TRY(bc ? bc->ctx : nullptr, {
        PFEM_REQUIRE(bc && (value || bc->nbc == 0), "null argument");
        h2d(C(bc->ctx), bc->value.p, value, sizeof(double) * bc->nbc);
    });
}

int pfem_dirichlet_values_device(pfem_dirichlet bc, double** d_value)
{
    if (!bc || !d_value) return PFEM_EINVAL;
    *d_value = bc->value.p;
    return PFEM_OK;
}

int pfem_dirichlet_apply(pfem_dirichlet bc, pfem_csr a, double* d_rhs, int mode)
{
    PFEM_TRY(bc ? bc->ctx : nullptr, {
        PFEM_REQUIRE(bc && a && d_rhs, "null argument");
        dirichlet_apply(C(bc->ctx), bc, nullptr, a, d_rhs, nullptr, mode);
    });
}

int pfem_dirichlet_apply2(pfem_dirichlet bc0, pfem_dirichlet bc1, pfem_csr a, double* d_rhs0, double* d_rhs1, int mode)
{
    PFEM_TRY(bc0 ? bc0->ctx : nullptr, {
        PFEM_REQUIRE(bc0 && bc1 && a && d_rhs0 && d_rhs1, "null argument");
        dirichlet_apply(C(bc0->ctx), bc0, bc1, a, d_rhs0, d_rhs1, mode);
    });
}

// DataExport::exportToVTK (src/data_export.cu:17-95): ASCII VTU, points, triangles (type 5), nodal
// scalar fields. Written with 17 significant digits (the reference streams 6).
int pfem_export_vtk(pfem_mesh2d m, const char* path, int nfields, const char* const* names,
                    const double* const* d_fields)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && path, "null argument");
        std::ofstream f(path);
        if (!f.is_open()) throw Error(PFEM_EIO, std::string("cannot write ") + path);
        f << std::setprecision(17);
        f << "<?xml version=\"1.0\" ?> \n";
        f << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
        f << "  <UnstructuredGrid>\n";
        f << "    <Piece NumberOfPoints=\"" << m->nv << "\" NumberOfCells=\"" << m->nt_host << "\">\n";
        f << "      <Points>\n        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
        for (int i = 0; i < m->nv; ++i) f << "        " << m->h_xy[2 * (size_t)i] << " " << m->h_xy[2 * (size_t)i + 1] << " 0.0\n";
        f << "        </DataArray>\n      </Points>\n      <Cells>\n";
        f << "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n          ";
        for (int c = 0; c < m->nt_host; ++c)
            f << m->h_tri[3 * (size_t)c] << " " << m->h_tri[3 * (size_t)c + 1] << " " << m->h_tri[3 * (size_t)c + 2] << " ";
        f << "\n        </DataArray>\n        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n          ";
        for (int c = 0; c < m->nt_host; ++c) f << (c + 1) * 3 << " ";
        f << "\n        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" Format=\"ascii\">\n          ";
        for (int c = 0; c < m->nt_host; ++c) f << 5 << " ";
        f << "\n        </DataArray>\n      </Cells>\n";
        if (nfields > 0) {
            f << "      <PointData Scalars=\"scalars\">\n";
            std::vector<double> h(m->nv);
            for (int k = 0; k < nfields; ++k) {
                d2h_sync(C(m->ctx), h.data(), d_fields[k], sizeof(double) * m->nv);
                f << "        <DataArray type=\"Float32\" Name=\"" << names[k] << "\" Format=\"ascii\">\n        ";
                for (int i = 0; i < m->nv; ++i) f << h[i] << "  ";
                f << "\n        </DataArray>\n";
            }
            f << "      </PointData>\n";
        }
        f << "    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    });
}

// DataExport3D::exportToVTK (src/data_export_3d.cu:15-93): ASCII VTU, points, tetrahedra (type 10),
// nodal scalar fields. Written with 17 significant digits (the reference streams 6).
int pfem_export_vtk3d(pfem_mesh3d m, const char* path, int nfields, const char* const* names,
                      const double* const* d_fields)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && path && (nfields == 0 || (names && d_fields)), "null argument");
        std::ofstream f(path);
        if (!f.is_open()) throw Error(PFEM_EIO, std::string("cannot write ") + path);
        f << std::setprecision(17);
        f << "<?xml version=\"1.0\" ?> \n";
        f << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
        f << "  <UnstructuredGrid>\n";
        f << "    <Piece NumberOfPoints=\"" << m->nv << "\" NumberOfCells=\"" << m->nt << "\">\n";
        f << "      <Points>\n        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
        for (int i = 0; i < m->nv; ++i)
            f << "        " << m->h_xyz[3 * (size_t)i] << " " << m->h_xyz[3 * (size_t)i + 1] << " "
              << m->h_xyz[3 * (size_t)i + 2] << "\n";
        f << "        </DataArray>\n      </Points>\n      <Cells>\n";
        f << "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n          ";
        for (int c = 0; c < m->nt; ++c)
            for (int k = 0; k < 4; ++k) f << m->h_tet[4 * (size_t)c + k] << " ";
        f << "\n        </DataArray>\n        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n          ";
        for (int c = 0; c < m->nt; ++c) f << (c + 1) * 4 << " ";
        f << "\n        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" Format=\"ascii\">\n          ";
        for (int c = 0; c < m->nt; ++c) f << 10 << " ";
        f << "\n        </DataArray>\n      </Cells>\n";
        if (nfields > 0) {
            f << "      <PointData Scalars=\"scalars\">\n";
            std::vector<double> h(m->nv);
            for (int k = 0; k < nfields; ++k) {
                d2h_sync(C(m->ctx), h.data(), d_fields[k], sizeof(double) * m->nv);
                f << "        <DataArray type=\"Float32\" Name=\"" << names[k] << "\" Format=\"ascii\">\n        ";
                for (int i = 0; i < m->nv; ++i) f << h[i] << "  ";
                f << "\n        </DataArray>\n";
            }
            f << "      </PointData>\n";
        }
        f << "    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    });
}

}  // extern "C"
