/* Header-only C++ facade over the C ABI of libpfem2_hip.so (pfem2.h) with the class names, method
 * names and argument types of the reference library's public headers, so that a driver written against
 * the reference (cases/<name>/main.cu) keeps its call sequence:
 *
 *   Mesh2D                  src/mesh_2d.cuh:11-69
 *   SparseMatrixCSR         src/linear_algebra/sparse_matrix.cuh:9-58
 *   DirichletNode/BCs       src/Dirichlet_bcs.cuh:8-29
 *   LinearAlgebra           src/linear_algebra/linear_algebra.h:10-80 (dense wrappers, sparseMV)
 *   Preconditioner*         src/linear_algebra/preconditioners.cuh:9-90 (virtual plugin point kept)
 *   LinearSolver/SolverCG/SolverGMRES   src/linear_algebra/linear_solver.cuh:15-130
 *   ParticleHandler2D       src/particles/particle_handler_2d.cuh:9-56
 *   SimulationParameters    src/parameters.cuh:6-51
 *   GaussPoint1D/2D/3D, create*QuadratureFormula   src/integration/quadrature_formula_{1d,2d,3d}.cuh
 *   NumericalIntegrator2D   src/integration/numerical_integrator_2d.cuh:45-57 (the case integrands are
 *                           library operators here: assemblePoisson / assemblePrediction / ...)
 *   DataExport              src/data_export.cuh:10-33
 *   GpuTimer                src/common/gpu_timer.cuh:14-98
 *   deviceVector<T>         src/common/device_vector.cuh:19-146 (move-only: the reference's implicit
 *                           shallow copy double-frees)
 *
 * Errors: the reference exits on every CUDA failure and returns bool from loaders and solvers. Here
 * loaders and solvers keep their bool results; any other failure throws gpupfem2::Error (or, with
 * GPUPFEM2_EXIT_ON_ERROR defined, prints the message and exits as the reference does).
 *
 * All objects share one process-wide context (device 0, or the index in GPUPFEM2_DEVICE), whose stream
 * orders every call, as the reference's legacy default stream does. Host-only code: compile with any
 * C++17 compiler and link libpfem2_hip.so. */
#ifndef GPUPFEM2_HPP
#define GPUPFEM2_HPP

#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pfem2.h"

namespace gpupfem2 {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline pfem_ctx context()
{
    struct Holder {
        pfem_ctx ctx = nullptr;
        Holder()
        {
            const char* d = std::getenv("GPUPFEM2_DEVICE");
            const int rc = pfem_ctx_create(d ? std::atoi(d) : 0, &ctx);
            if (rc != PFEM_OK) {
                std::fprintf(stderr, "gpupfem2: no usable GPU (context creation failed, code %d)\n", rc);
                std::exit(EXIT_FAILURE);
            }
        }
        ~Holder() { pfem_ctx_destroy(ctx); }
    };
    static Holder h;
    return h.ctx;
}

inline void check(int rc)
{
    if (rc == PFEM_OK) return;
    std::string msg = pfem_last_error(context());
    if (msg.empty()) msg = "pfem error " + std::to_string(rc) + " (invalid or null handle)";
#ifdef GPUPFEM2_EXIT_ON_ERROR
    std::fprintf(stderr, "gpupfem2 error %d: %s\n", rc, msg.c_str());
    std::exit(EXIT_FAILURE);
#else
    throw Error(rc, msg);
#endif
}

}  // namespace gpupfem2

/* ------------------------------------------------------------------ small value types (common/cuda_math.cuh) */
struct Point2 {
    double x, y;
};
struct Point3 {
    double x, y, z;
};
struct Point4 {
    double x, y, z, w;
};
struct Cell3 {
    unsigned x, y, z;   // the reference's uint3 cell record
};
struct DirichletNode {
    unsigned int nodeIdx;
    double bcValue;
};

/* ------------------------------------------------------------------ deviceVector<T> */
template <class T>
struct deviceVector {
    T* data = nullptr;   //!< device pointer
    int size = 0;        //!< number of items
    int capacity = 0;    //!< allocated items

    deviceVector() = default;
    explicit deviceVector(const std::vector<T>& host)
    {
        allocate((int)host.size());
        copyFromHost(host);
    }
    deviceVector(const deviceVector&) = delete;
    deviceVector& operator=(const deviceVector&) = delete;
    deviceVector(deviceVector&& o) noexcept { swap(o); }
    deviceVector& operator=(deviceVector&& o) noexcept
    {
        if (this != &o) {
            free();
            swap(o);
        }
        return *this;
    }
    ~deviceVector() { free(); }

    void allocate(int new_size)
    {
        free();
        if (new_size <= 0) return;
        void* p = nullptr;
        gpupfem2::check(pfem_malloc(gpupfem2::context(), sizeof(T) * (size_t)new_size, &p));
        data = static_cast<T*>(p);
        size = capacity = new_size;
    }
    void free()
    {
        if (data) pfem_free(gpupfem2::context(), data);
        data = nullptr;
        size = capacity = 0;
    }
    size_t bytes() const { return sizeof(T) * (size_t)size; }
    void clearValues()
    {
        if (data) gpupfem2::check(pfem_memset_zero(gpupfem2::context(), data, bytes()));
    }
    void copyFromHost(const std::vector<T>& host)
    {
        if (host.empty()) return;
        gpupfem2::check(pfem_memcpy_h2d(gpupfem2::context(), data, host.data(), sizeof(T) * host.size()));
    }
    std::vector<T> copyToHost() const
    {
        std::vector<T> h(size);
        if (size) gpupfem2::check(pfem_memcpy_d2h(gpupfem2::context(), h.data(), data, bytes()));
        return h;
    }
    /* one value per line, default stream formatting (deviceVector::exportToFile) */
    bool exportToFile(const std::string& filename) const
    {
        std::ofstream out(filename.c_str());
        if (!out.is_open()) return false;
        for (const T& v : copyToHost()) out << v << std::endl;
        std::printf("Vector saved to %s\n", filename.c_str());
        return true;
    }

private:
    void swap(deviceVector& o)
    {
        std::swap(data, o.data);
        std::swap(size, o.size);
        std::swap(capacity, o.capacity);
    }
};

/* non-owning view with the members the reference's getters expose (.data, .size) */
template <class T>
struct deviceSpan {
    T* data = nullptr;
    int size = 0;
};

/* ------------------------------------------------------------------ meshes */
class Mesh2D {
public:
    Mesh2D() = default;
    Mesh2D(const Mesh2D&) = delete;
    Mesh2D& operator=(const Mesh2D&) = delete;
    ~Mesh2D()
    {
        if (h_) pfem_mesh2d_destroy(h_);
    }

    /* Mesh2D::loadMeshFromFile (src/mesh_2d.cu:36-96): false when the file cannot be read */
    bool loadMeshFromFile(const std::string& filename, bool fillNeighborLists = true, double scale = 1.0)
    {
        pfem_mesh2d m = nullptr;
        const int rc = pfem_mesh2d_load_dat(gpupfem2::context(), filename.c_str(), fillNeighborLists ? 1 : 0, scale, &m);
        if (rc != PFEM_OK) {
            std::printf("Error while reading mesh file %s: %s\n", filename.c_str(), pfem_last_error(gpupfem2::context()));
            return false;
        }
        reset(m);
        return true;
    }
    /* additive: a mesh from host arrays (generated meshes) */
    void createFromArrays(const std::vector<Point2>& xy, const std::vector<Cell3>& cells, bool fillNeighborLists = true)
    {
        pfem_mesh2d m = nullptr;
        gpupfem2::check(pfem_mesh2d_create(gpupfem2::context(), (int)xy.size(), &xy[0].x, (int)cells.size(),
                                           reinterpret_cast<const int*>(cells.data()), fillNeighborLists ? 1 : 0, &m));
        reset(m);
    }
    /* geometry is computed when the mesh is loaded (initMesh, src/mesh_2d.cu:98-105) */
    void initMesh() {}

    deviceSpan<Point2> getVertices() const { return {static_cast<Point2*>(ptr(0)), nv_}; }
    deviceSpan<Cell3> getCells() const { return {static_cast<Cell3*>(ptr(1)), nt_}; }
    deviceSpan<int> getEdgeBoundaryIDs() const { return {static_cast<int*>(ptr(6)), 3 * nt_}; }
    deviceSpan<double> getCellArea() const { return {static_cast<double*>(ptr(2)), nt_}; }
    deviceSpan<double> getInvJacobi() const { return {static_cast<double*>(ptr(3)), 4 * nt_}; }
    deviceSpan<int> getCellNeighborsOffsets() const { return {static_cast<int*>(ptr(4)), nt_ + 1}; }
    deviceSpan<int> getCellNeighborIndices() const { return {static_cast<int*>(ptr(5)), (int)nnb_}; }
    std::vector<Point2> getHostVertices() const
    {
        if (!h_) throw gpupfem2::Error(PFEM_EINVAL, "Mesh2D used before loadMeshFromFile");
        std::vector<Point2> v(nv_);
        gpupfem2::check(pfem_mesh2d_download(h_, 0, v.data()));
        return v;
    }
    std::vector<Cell3> getHostCells() const
    {
        if (!h_) throw gpupfem2::Error(PFEM_EINVAL, "Mesh2D used before loadMeshFromFile");
        std::vector<Cell3> c(ntHost_);
        gpupfem2::check(pfem_mesh2d_download(h_, 7, c.data()));
        return c;
    }
    /* additive: edge boundary IDs by midpoint position (kSetEdgeBoundaryIDs of the cases) */
    void setEdgeBoundaryIDs(double x0, double x1, double y0, double y1, int id_x0, int id_x1, int id_y0, int id_y1,
                            double cx = 0.0, double cy = 0.0, double r = 0.0, double tol = 1e-3, int id_circle = -1)
    {
        gpupfem2::check(pfem_mesh2d_classify_edges(h_, x0, x1, y0, y1, id_x0, id_x1, id_y0, id_y1, cx, cy, r, tol,
                                                   id_circle));
    }

    pfem_mesh2d handle() const { return h_; }

private:
    void reset(pfem_mesh2d m)
    {
        if (h_) pfem_mesh2d_destroy(h_);
        h_ = m;
        gpupfem2::check(pfem_mesh2d_info(h_, &nv_, &nt_, &ntHost_, &nnb_));
    }
    void* ptr(int what) const
    {
        if (!h_) throw gpupfem2::Error(PFEM_EINVAL, "Mesh2D used before loadMeshFromFile");
        void* p = nullptr;
        gpupfem2::check(pfem_mesh2d_device_ptr(h_, what, &p));
        return p;
    }
    pfem_mesh2d h_ = nullptr;
    int nv_ = 0, nt_ = 0, ntHost_ = 0;
    long nnb_ = 0;
};

class Mesh3D {
public:
    Mesh3D() = default;
    Mesh3D(const Mesh3D&) = delete;
    Mesh3D& operator=(const Mesh3D&) = delete;
    ~Mesh3D()
    {
        if (h_) pfem_mesh3d_destroy(h_);
    }
    /* Mesh3D::loadMeshFromFile (src/mesh_3d.cu:84-147), the reference's arguments: the face-neighbour lists
     * are built unless fillNeighborLists is false (fillCellNeighborIndices, :198-227). The additive
     * zeroSmallDet keeps the reference's zeroing of inverse Jacobians with |det| < 2e-6
     * (matrix3x3.cuh:134-154). False when the file cannot be read. */
    bool loadMeshFromFile(const std::string& filename, bool fillNeighborLists = true, double scale = 1.0,
                          bool zeroSmallDet = true)
    {
        pfem_mesh3d m = nullptr;
        if (pfem_mesh3d_load_dat(gpupfem2::context(), filename.c_str(), scale, zeroSmallDet ? 1 : 0, &m) != PFEM_OK) {
            std::printf("Error while reading mesh file %s: %s\n", filename.c_str(), pfem_last_error(gpupfem2::context()));
            return false;
        }
        if (h_) pfem_mesh3d_destroy(h_);
        h_ = m;
        nnb_ = 0;
        nbrBuilt_ = false;
        gpupfem2::check(pfem_mesh3d_info(h_, &nv_, &nt_));
        if (fillNeighborLists) fillCellNeighborIndices();
        return true;
    }
    /* Mesh3D::fillCellNeighborIndices (src/mesh_3d.cu:198-227): cells sharing a face, ascending */
    void fillCellNeighborIndices()
    {
        gpupfem2::check(pfem_mesh3d_fill_neighbors(handleChecked(), &nnb_));
        nbrBuilt_ = true;
    }
    int numVertices() const { return nv_; }
    int numCells() const { return nt_; }
    /* Mesh3D::getHostVertices (mesh_3d.cuh): the vertices as imported */
    std::vector<Point3> getHostVertices() const
    {
        std::vector<Point3> v((size_t)nv_);
        if (nv_) gpupfem2::check(pfem_mesh3d_download(handleChecked(), 0, v.data()));
        return v;
    }
    deviceSpan<Point3> getVertices() const { return {static_cast<Point3*>(ptr(0)), nv_}; }
    deviceSpan<int> getCells() const { return {static_cast<int*>(ptr(1)), 4 * nt_}; }
    deviceSpan<double> getCellVolume() const { return {static_cast<double*>(ptr(2)), nt_}; }
    deviceSpan<double> getInvJacobi() const { return {static_cast<double*>(ptr(3)), 9 * nt_}; }
    deviceSpan<int> getCellNeighborsOffsets() const { return {static_cast<int*>(ptr(4)), nbrBuilt_ ? nt_ + 1 : 0}; }
    deviceSpan<int> getCellNeighborIndices() const { return {static_cast<int*>(ptr(5)), (int)nnb_}; }
    /* additive: the face-neighbour lists on the host, as (offsets, indices) */
    void getHostCellNeighbors(std::vector<int>& offsets, std::vector<int>& indices) const
    {
        if (!nbrBuilt_) throw gpupfem2::Error(PFEM_EINVAL, "face neighbours not built");
        offsets.resize((size_t)nt_ + 1);
        indices.resize((size_t)nnb_);
        gpupfem2::check(pfem_mesh3d_download(handleChecked(), 4, offsets.data()));
        if (nnb_) gpupfem2::check(pfem_mesh3d_download(handleChecked(), 5, indices.data()));
    }
    pfem_mesh3d handle() const { return h_; }

private:
    pfem_mesh3d handleChecked() const
    {
        if (!h_) throw gpupfem2::Error(PFEM_EINVAL, "Mesh3D used before loadMeshFromFile");
        return h_;
    }
    void* ptr(int what) const
    {
        void* p = nullptr;
        gpupfem2::check(pfem_mesh3d_device_ptr(handleChecked(), what, &p));
        return p;
    }
    pfem_mesh3d h_ = nullptr;
    int nv_ = 0, nt_ = 0;
    long nnb_ = 0;
    bool nbrBuilt_ = false;
};

/* ------------------------------------------------------------------ CSR matrix */
class SparseMatrixCSR {
public:
    SparseMatrixCSR() = default;
    explicit SparseMatrixCSR(const Mesh2D& mesh) { initialize(mesh); }
    explicit SparseMatrixCSR(const Mesh3D& mesh) { initialize(mesh); }
    SparseMatrixCSR(const SparseMatrixCSR&) = delete;
    SparseMatrixCSR& operator=(const SparseMatrixCSR&) = delete;
    ~SparseMatrixCSR()
    {
        if (h_ && owned_) pfem_csr_destroy(h_);
    }

    /* SparseMatrixCSR::initialize(const Mesh2D&) (sparse_matrix.cu:8-54): pattern, values zero */
    void initialize(const Mesh2D& mesh)
    {
        pfem_csr a = nullptr;
        gpupfem2::check(pfem_csr_create_from_mesh2d(mesh.handle(), &a));
        reset(a, true);
    }
    void initialize(const Mesh3D& mesh)
    {
        pfem_csr a = nullptr;
        gpupfem2::check(pfem_csr_create_from_mesh3d(mesh.handle(), &a));
        reset(a, true);
    }
    /* non-owning view of a matrix held by the library (the argument of Preconditioner::initialize) */
    static SparseMatrixCSR view(pfem_csr a)
    {
        SparseMatrixCSR m;
        m.reset(a, false);
        return m;
    }
    SparseMatrixCSR(SparseMatrixCSR&& o) noexcept
        : h_(o.h_), owned_(o.owned_), rows_(o.rows_), nnz_(o.nnz_), rp_(o.rp_), col_(o.col_), val_(o.val_)
    {
        o.h_ = nullptr;
    }

    int* getRowOffset() const { return rp_; }
    int* getColIndices() const { return col_; }
    double* getMatrixValues() const { return val_; }
    int getRows() const { return rows_; }
    int getTotalElements() const { return nnz_; }
    void clearValues() { gpupfem2::check(pfem_csr_clear_values(h_)); }

    /* "row col value" per entry, 1-based, default stream formatting (sparse_matrix.cu:100-124) */
    bool exportMatrix(const std::string& filename) const
    {
        std::ofstream out(filename.c_str());
        if (!out.is_open()) return false;
        std::vector<int> rp(rows_ + 1), col(nnz_);
        std::vector<double> val(nnz_);
        gpupfem2::check(pfem_csr_download(h_, rp.data(), col.data(), val.data()));
        for (int i = 0; i < rows_; ++i)
            for (int j = rp[i]; j < rp[i + 1]; ++j) out << i + 1 << " " << col[j] + 1 << " " << val[j] << std::endl;
        std::printf("Matrix saved to %s\n", filename.c_str());
        return true;
    }

    pfem_csr handle() const { return h_; }

private:
    void reset(pfem_csr a, bool owned)
    {
        if (h_ && owned_) pfem_csr_destroy(h_);
        h_ = a;
        owned_ = owned;
        gpupfem2::check(pfem_csr_info(h_, &rows_, &nnz_));
        gpupfem2::check(pfem_csr_pointers(h_, &rp_, &col_, &val_));
    }
    pfem_csr h_ = nullptr;
    bool owned_ = true;
    int rows_ = 0, nnz_ = 0;
    int* rp_ = nullptr;
    int* col_ = nullptr;
    double* val_ = nullptr;
};

/* ------------------------------------------------------------------ Dirichlet boundary conditions */
class DirichletBCs {
public:
    DirichletBCs() = default;
    DirichletBCs(const DirichletBCs&) = delete;
    DirichletBCs& operator=(const DirichletBCs&) = delete;
    virtual ~DirichletBCs()
    {
        if (h_) pfem_dirichlet_destroy(h_);
    }

    /* DirichletBCs::setupDirichletBCs (Dirichlet_bcs.cu:53-59); the node map is built with the matrix size
     * at the first applyBCs, or by setupNodeMap */
    void setupDirichletBCs(const std::vector<DirichletNode>& hostBcs)
    {
        host_ = hostBcs;
        if (h_) pfem_dirichlet_destroy(h_);
        h_ = nullptr;
    }
    void setupNodeMap(int n, const std::vector<DirichletNode>& hostBCs)
    {
        host_ = hostBCs;
        create(n);
    }
    /* DirichletBCs::applyBCs (Dirichlet_bcs.cu:73-86): the reference's elimination heuristic unless
     * setEliminationMode(PFEM_DIRICHLET_MASK) */
    void applyBCs(SparseMatrixCSR& matrix, deviceVector<double>& rhs)
    {
        if (!h_ || n_ != matrix.getRows()) create(matrix.getRows());
        gpupfem2::check(pfem_dirichlet_apply(h_, matrix.handle(), rhs.data, mode_));
    }
    /* device value array (one double per BC node, in setup order): what the reference's subclasses write
     * through the protected DirichletValues (cases/Cylinder2D/main.cu:447-479) */
    double* getDeviceValues()
    {
        double* p = nullptr;
        gpupfem2::check(pfem_dirichlet_values_device(h_, &p));
        return p;
    }
    void setEliminationMode(int mode) { mode_ = mode; }
    const std::vector<DirichletNode>& getHostBCs() const { return host_; }
    pfem_dirichlet handle() const { return h_; }

protected:
    void create(int n)
    {
        std::vector<int> node(host_.size());
        std::vector<double> val(host_.size());
        for (size_t k = 0; k < host_.size(); ++k) {
            node[k] = (int)host_[k].nodeIdx;
            val[k] = host_[k].bcValue;
        }
        if (h_) pfem_dirichlet_destroy(h_);
        h_ = nullptr;
        gpupfem2::check(pfem_dirichlet_create(gpupfem2::context(), n, (int)node.size(), node.data(), val.data(), &h_));
        n_ = n;
    }
    std::vector<DirichletNode> host_;
    pfem_dirichlet h_ = nullptr;
    int n_ = -1;
    int mode_ = PFEM_DIRICHLET_PARITY;
};

/* ------------------------------------------------------------------ linear algebra */
/* LinearAlgebra (linear_algebra.h:10-80). The reference keeps cuBLAS / cuSPARSE handles here; the library's
 * context holds its stream. The dense wrappers keep the reference's device pointer mode: res, alpha and beta
 * point to device memory. The reference's sparse wrappers take cuSPARSE descriptors; here sparseMV takes the
 * matrix itself, and the incomplete factorisations and triangular solves are the Preconditioner classes. */
class LinearAlgebra {
public:
    void dot(const deviceVector<double>& v1, const deviceVector<double>& v2, double* res) const
    {
        if (v1.size == v2.size)
            dot(v1.data, v2.data, res, v1.size);
        else
            std::printf("Inconsistent dimensions of vectors in dot product\n");
    }
    void dot(const double* v1, const double* v2, double* res, int n) const
    {
        gpupfem2::check(pfem_dot(gpupfem2::context(), n, v1, v2, res));
    }
    /* the Euclidean norm itself, as the reference's cublasDnrm2 */
    void normSquared(const double* v, double* res, int n) const
    {
        gpupfem2::check(pfem_nrm2(gpupfem2::context(), n, v, res));
    }
    void solveGeneralTriangularSystem(const double* matrix, double* solution, int n, int lda) const
    {
        gpupfem2::check(pfem_trsv_upper(gpupfem2::context(), n, matrix, lda, solution));
    }
    void generalMV(const double* matrix, const double* vec, double* res, const double* alpha, const double* beta,
                   int m, int n, int lda) const
    {
        gpupfem2::check(pfem_gemv(gpupfem2::context(), m, n, alpha, matrix, lda, vec, beta, res));
    }
    /* result = A vec (sparseMV, linear_algebra.h:75-80, as the solvers call it: alpha 1, beta 0) */
    void sparseMV(const SparseMatrixCSR& matrix, const double* vec, double* result) const
    {
        gpupfem2::check(pfem_spmv(matrix.handle(), vec, result));
    }
};

/* Preconditioner (preconditioners.cuh:9-26). A subclass that overrides initialize / applyPreconditioner
 * is a plugin: the solvers call it through PFEM_PRECOND_CUSTOM on the library's stream. */
class Preconditioner {
public:
    Preconditioner(int n_, const LinearAlgebra* LA_) : n(n_), LA(LA_) {}
    virtual ~Preconditioner()
    {
        if (custom_) pfem_precond_destroy(custom_);
    }
    Preconditioner(const Preconditioner&) = delete;
    Preconditioner& operator=(const Preconditioner&) = delete;

    virtual void initialize(const SparseMatrixCSR& csrMatrix) = 0;
    virtual void applyPreconditioner(double* dest, const double* src) = 0;

    /* the C handle a solver uses (created at the first solve with the solver's matrix) */
    virtual pfem_precond handleFor(const SparseMatrixCSR& a)
    {
        (void)a;
        if (!custom_)
            gpupfem2::check(pfem_precond_create_custom(gpupfem2::context(), n, &Preconditioner::initTrampoline,
                                                       &Preconditioner::applyTrampoline, this, &custom_));
        return custom_;
    }

protected:
    int n;
    const LinearAlgebra* LA;

private:
    static void initTrampoline(void* user, pfem_csr a)
    {
        const SparseMatrixCSR m = SparseMatrixCSR::view(a);
        static_cast<Preconditioner*>(user)->initialize(m);
    }
    static void applyTrampoline(void* user, double* dest, const double* src)
    {
        static_cast<Preconditioner*>(user)->applyPreconditioner(dest, src);
    }
    pfem_precond custom_ = nullptr;
};

/* library preconditioners: kind PFEM_PRECOND_JACOBI / IC0 / ILU0 */
class BuiltinPreconditioner : public Preconditioner {
public:
    BuiltinPreconditioner(int n_, const LinearAlgebra* LA_, int kind, int ordering)
        : Preconditioner(n_, LA_), kind_(kind), ordering_(ordering)
    {
    }
    ~BuiltinPreconditioner() override
    {
        if (h_) pfem_precond_destroy(h_);
    }
    void initialize(const SparseMatrixCSR& csrMatrix) override
    {
        handleFor(csrMatrix);
        gpupfem2::check(pfem_precond_initialize(h_, csrMatrix.handle()));
    }
    void applyPreconditioner(double* dest, const double* src) override
    {
        if (!h_) throw gpupfem2::Error(PFEM_EINVAL, "preconditioner used before initialize");
        gpupfem2::check(pfem_precond_apply(h_, dest, src));
    }
    pfem_precond handleFor(const SparseMatrixCSR& a) override
    {
        if (!h_) gpupfem2::check(pfem_precond_create(a.handle(), kind_, ordering_, &h_));
        return h_;
    }

private:
    int kind_, ordering_;
    pfem_precond h_ = nullptr;
};

/* PreconditionerJacobi(n, LA) (preconditioners.cu:3-46) */
class PreconditionerJacobi : public BuiltinPreconditioner {
public:
    PreconditionerJacobi(int n_, const LinearAlgebra* LA_) : BuiltinPreconditioner(n_, LA_, PFEM_PRECOND_JACOBI, 0) {}
};

/* PreconditionerIC(matrix, LA) (preconditioners.cu:136-213); ordering PFEM_ORDER_MULTICOLOR is additive */
class PreconditionerIC : public BuiltinPreconditioner {
public:
    PreconditionerIC(const SparseMatrixCSR& matrix, const LinearAlgebra* LA_, int ordering = PFEM_ORDER_NATURAL)
        : BuiltinPreconditioner(matrix.getRows(), LA_, PFEM_PRECOND_IC0, ordering)
    {
        handleFor(matrix);
    }
    /* additive: storage of the factor copy applied in PFEM_PRECISION_MIXED_PRECOND (16 or 32 bits;
     * pfem_precond_set_factor_bits); call before the solver's precision is set */
    void setFactorBits(const SparseMatrixCSR& matrix, int bits)
    {
        gpupfem2::check(pfem_precond_set_factor_bits(handleFor(matrix), bits));
    }
};

/* PreconditionerILU(matrix, LA) (preconditioners.cu:48-134) */
class PreconditionerILU : public BuiltinPreconditioner {
public:
    PreconditionerILU(const SparseMatrixCSR& matrix, const LinearAlgebra* LA_, int ordering = PFEM_ORDER_NATURAL)
        : BuiltinPreconditioner(matrix.getRows(), LA_, PFEM_PRECOND_ILU0, ordering)
    {
        handleFor(matrix);
    }
};

/* LinearSolver (linear_solver.cuh:15-46): absolute tolerance and maxIterations as in the reference;
 * relative tolerance, GMRES restart and precision are additive options */
class LinearSolver {
public:
    LinearSolver(double tolerance_, int max_iterations, const LinearAlgebra* LA_, Preconditioner* precond_ = nullptr,
                 int kind = PFEM_SOLVER_CG, int restart_ = 0)
        : LA(LA_), precond(precond_), tolerance(tolerance_), maxIterations(max_iterations), kind_(kind), restart_(restart_)
    {
    }
    virtual ~LinearSolver()
    {
        if (h_) pfem_solver_destroy(h_);
    }
    LinearSolver(const LinearSolver&) = delete;
    LinearSolver& operator=(const LinearSolver&) = delete;

    /* LinearSolver::init (linear_solver.cu:113-171): sizes the workspace for matrices of this shape */
    virtual void init(const SparseMatrixCSR& matrix)
    {
        if (h_) pfem_solver_destroy(h_);
        h_ = nullptr;
        pfem_precond pc = precond ? precond->handleFor(matrix) : nullptr;
        gpupfem2::check(pfem_solver_create(matrix.handle(), kind_, tolerance, maxIterations, restart_, tolMode_, pc, &h_));
        if (precision_ != PFEM_PRECISION_FP64) gpupfem2::check(pfem_solver_set_precision(h_, precision_, innerTol_, 0));
        n = matrix.getRows();
        nnz = matrix.getTotalElements();
    }
    /* solve(A, x, b): x is the initial guess and the result; true when converged */
    virtual bool solve(const SparseMatrixCSR& A, deviceVector<double>& x, const deviceVector<double>& b)
    {
        if (!h_) init(A);
        gpupfem2::check(pfem_solver_solve(h_, A.handle(), x.data, b.data, &stats_));
        residual_norm = stats_.final_residual;
        return stats_.converged != 0;
    }

    /* Additions (no counterpart in the reference, which solves one system per call): nrhs <= 8 right-hand
     * sides of A at once, column j at X + j*n and B + j*n (pfem_solver_solve_multi); and independent
     * systems of a block-diagonal A, system k owning rows sysRows[k]..sysRows[k+1]
     * (pfem_solver_solve_batched). PCG with Jacobi or no preconditioner. True when every column / system
     * converged; per-column / per-system statistics in multiStats(). */
    bool solveMulti(const SparseMatrixCSR& A, deviceVector<double>& X, const deviceVector<double>& B, int nrhs)
    {
        if (!h_) init(A);
        if ((long)X.size < (long)nrhs * n || (long)B.size < (long)nrhs * n)
            throw gpupfem2::Error(PFEM_EINVAL, "solveMulti: X and B need nrhs * n entries");
        multi_.assign(nrhs, pfem_solve_stats{});
        gpupfem2::check(pfem_solver_solve_multi(h_, A.handle(), nrhs, X.data, B.data, multi_.data()));
        bool all = true;
        for (const auto& st : multi_) all = all && st.converged;
        return all;
    }
    bool solveBatched(const SparseMatrixCSR& A, const std::vector<int>& sysRows, deviceVector<double>& x,
                      const deviceVector<double>& b, std::vector<double>* timeMs = nullptr)
    {
        if (!h_) init(A);
        const int nsys = (int)sysRows.size() - 1;
        if (nsys < 1) throw gpupfem2::Error(PFEM_EINVAL, "solveBatched: no systems");
        multi_.assign(nsys, pfem_solve_stats{});
        if (timeMs) timeMs->assign(nsys, 0.0);
        gpupfem2::check(pfem_solver_solve_batched(h_, A.handle(), nsys, sysRows.data(), x.data, b.data, multi_.data(),
                                                  timeMs ? timeMs->data() : nullptr));
        bool all = true;
        for (const auto& st : multi_) all = all && st.converged;
        return all;
    }
    const std::vector<pfem_solve_stats>& multiStats() const { return multi_; }

    void setRelativeTolerance(bool relative) { tolMode_ = relative ? PFEM_TOL_RELATIVE : PFEM_TOL_ABSOLUTE; }
    void setPrecision(int precision, double innerTolerance = 1e-6)
    {
        precision_ = precision;
        innerTol_ = innerTolerance;
    }
    const pfem_solve_stats& lastStats() const { return stats_; }

protected:
    const LinearAlgebra* LA;
    Preconditioner* precond = nullptr;
    double residual_norm = 0.0;
    double tolerance;
    int maxIterations;
    int n = 0, nnz = 0;

private:
    int kind_, restart_;
    int tolMode_ = PFEM_TOL_ABSOLUTE;
    int precision_ = PFEM_PRECISION_FP64;
    double innerTol_ = 1e-6;
    pfem_solver h_ = nullptr;
    pfem_solve_stats stats_{};
    std::vector<pfem_solve_stats> multi_;
};

/* SolverCG (linear_solver.cu:173-376) */
class SolverCG : public LinearSolver {
public:
    SolverCG(double tolerance_, int max_iterations, const LinearAlgebra* LA_, Preconditioner* precond_ = nullptr)
        : LinearSolver(tolerance_, max_iterations, LA_, precond_, PFEM_SOLVER_CG)
    {
    }
};

/* SolverGMRES (linear_solver.cu:394-521): restart 0 = the reference's non-restarted GMRES */
class SolverGMRES : public LinearSolver {
public:
    SolverGMRES(double tolerance_, int max_iterations, const LinearAlgebra* LA_, Preconditioner* precond_ = nullptr,
                int restart_ = 0)
        : LinearSolver(tolerance_, max_iterations, LA_, precond_, PFEM_SOLVER_GMRES, restart_)
    {
    }
};

/* ------------------------------------------------------------------ quadrature */
struct GaussPoint1D {
    double coordinate;
    double weight;
};
struct GaussPoint2D {
    Point3 coordinates;
    double weight;
};
struct GaussPoint3D {
    Point4 coordinates;
    double weight;
};

inline std::vector<GaussPoint1D> createEdgeQuadratureFormula(int index)
{
    double buf[13 * 5];
    const int np = pfem_quadrature(1, index, buf);
    gpupfem2::check(np < 0 ? np : PFEM_OK);
    std::vector<GaussPoint1D> g(np);
    for (int k = 0; k < np; ++k) g[k] = {buf[2 * k], buf[2 * k + 1]};
    return g;
}
inline std::vector<GaussPoint2D> createFaceQuadratureFormula(int index)
{
    double buf[13 * 5];
    const int np = pfem_quadrature(2, index, buf);
    gpupfem2::check(np < 0 ? np : PFEM_OK);
    std::vector<GaussPoint2D> g(np);
    for (int k = 0; k < np; ++k) g[k] = {{buf[4 * k], buf[4 * k + 1], buf[4 * k + 2]}, buf[4 * k + 3]};
    return g;
}
inline std::vector<GaussPoint3D> createCellQuadratureFormula(int index)
{
    double buf[13 * 5];
    const int np = pfem_quadrature(3, index, buf);
    gpupfem2::check(np < 0 ? np : PFEM_OK);
    std::vector<GaussPoint3D> g(np);
    for (int k = 0; k < np; ++k)
        g[k] = {{buf[5 * k], buf[5 * k + 1], buf[5 * k + 2], buf[5 * k + 3]}, buf[5 * k + 4]};
    return g;
}

/* ------------------------------------------------------------------ simulation parameters */
struct SimulationParameters {
    double rho;
    double mu;
    double tFinal;
    double dt;
    double tolerance;
    int maxIterations;
    int simulationScheme;
    const char* meshFileName;
    int outputFrequency;
    int particleAdvectionSubsteps;
    int exportParticles;
    int calculateLoads;
    int bodyBoundaryID;
    double thickness;
    double meanVelocity;

    void setDefaultParameters()
    {
        pfem_simulation_parameters p;
        pfem_simulation_parameters_default(&p);
        rho = p.rho;
        mu = p.mu;
        tFinal = p.tFinal;
        dt = p.dt;
        particleAdvectionSubsteps = p.particleAdvectionSubsteps;
        tolerance = p.tolerance;
        maxIterations = p.maxIterations;
        simulationScheme = p.simulationScheme;
        meshFileName = p.meshFileName;
        outputFrequency = p.outputFrequency;
        exportParticles = p.exportParticles;
        calculateLoads = p.calculateLoads;
        bodyBoundaryID = p.bodyBoundaryID;
        thickness = p.thickness;
        meanVelocity = p.meanVelocity;
    }
};

/* ------------------------------------------------------------------ integration */
/* NumericalIntegrator2D (numerical_integrator_2d.cuh:45-57). The reference's cases derive from it and
 * launch their own integrand kernels; those integrands are library operators here, called from the
 * protected helpers below by the derived classes' assemble methods. */
class NumericalIntegrator2D {
public:
    explicit NumericalIntegrator2D(const Mesh2D& mesh_) : mesh(mesh_) {}
    virtual ~NumericalIntegrator2D() = default;

protected:
    /* cases/Poisson2D/main.cu:28-68 (coefficient 0.5, f = exp(-(x^2 + 10 y^2)) by default) */
    void assemblePoisson(SparseMatrixCSR& a, deviceVector<double>& rhs, double coeff = 0.5, int rhsKind = 0,
                         double rhsParam = 0.0, int faceQuadrature = 1)
    {
        gpupfem2::check(pfem_assemble_poisson2d(mesh.handle(), a.handle(), rhs.data, coeff, rhsKind, rhsParam,
                                                faceQuadrature));
    }
    void assemblePrediction(SparseMatrixCSR& a0, SparseMatrixCSR& a1, deviceVector<double>& rhs0,
                            deviceVector<double>& rhs1, const double* u0, const double* u1, const double* uo0,
                            const double* uo1, const double* pOld, const pfem_pfem2_params& prm)
    {
        gpupfem2::check(pfem_assemble_prediction(mesh.handle(), a0.handle(), a1.handle(), rhs0.data, rhs1.data, u0, u1,
                                                 uo0, uo1, pOld, &prm));
    }
    void assemblePressure(SparseMatrixCSR& a, deviceVector<double>& rhs, const double* vp0, const double* vp1,
                          const double* pOld, const pfem_pfem2_params& prm)
    {
        gpupfem2::check(pfem_assemble_pressure(mesh.handle(), a.handle(), rhs.data, vp0, vp1, pOld, &prm));
    }
    void assembleCorrection(SparseMatrixCSR& a, deviceVector<double>& rhs0, deviceVector<double>& rhs1,
                            const double* vp0, const double* vp1, const double* p, const double* pOld,
                            const pfem_pfem2_params& prm)
    {
        gpupfem2::check(pfem_assemble_correction(mesh.handle(), a.handle(), rhs0.data, rhs1.data, vp0, vp1, p, pOld,
                                                 &prm));
    }

    const Mesh2D& mesh;
};

/* NumericalIntegrator3D (numerical_integrator_3d.cuh:46-57): as NumericalIntegrator2D, the base of a
 * case's integrator; the Poisson3D integrand (cases/Poisson3D/main.cu:26-67) is the library operator
 * behind assemblePoisson. */
class NumericalIntegrator3D {
public:
    explicit NumericalIntegrator3D(const Mesh3D& mesh_) : mesh(mesh_) {}
    virtual ~NumericalIntegrator3D() = default;

protected:
    /* coefficient 0.5 and f = exp(-(x^2 + 10 y^2 + z^2)) by default (rhsKind 0; 1: rhsParam sin sin sin,
     * 2: constant rhsParam); cellQuadrature: createCellQuadratureFormula index (0 = 4 points, 1 = 6) */
    void assemblePoisson(SparseMatrixCSR& a, deviceVector<double>& rhs, double coeff = 0.5, int rhsKind = 0,
                         double rhsParam = 0.0, int cellQuadrature = 1)
    {
        gpupfem2::check(pfem_assemble_poisson3d(mesh.handle(), a.handle(), rhs.data, coeff, rhsKind, rhsParam,
                                                cellQuadrature));
    }

    const Mesh3D& mesh;
};

/* ------------------------------------------------------------------ particles */
class ParticleHandler2D {
public:
    ParticleHandler2D(const Mesh2D* mesh_, int cellDivisionLevel) : mesh(mesh_), level_(cellDivisionLevel)
    {
        gpupfem2::check(pfem_particles_create(mesh->handle(), cellDivisionLevel, &h_));
    }
    ~ParticleHandler2D() { pfem_particles_destroy(h_); }
    ParticleHandler2D(const ParticleHandler2D&) = delete;
    ParticleHandler2D& operator=(const ParticleHandler2D&) = delete;

    /* seedParticles (particle_handler_2d.cu:304-320): one particle per subcell centroid */
    void seedParticles() { gpupfem2::check(pfem_particles_seed(h_, 0, 0, 0)); }
    /* additive: ppc particles per cell at Philox-random barycentric positions */
    void seedParticlesRandom(int particlesPerCell, unsigned long long seed)
    {
        gpupfem2::check(pfem_particles_seed(h_, 1, particlesPerCell, seed));
    }
    void initParticleVelocity(const deviceVector<double*>& velocitySolution)
    {
        const auto u = velocitySolution.copyToHost();
        gpupfem2::check(pfem_particles_init_velocity(h_, u.at(0), u.at(1)));
    }
    void advectParticles(const deviceVector<double*>& velocitySolution, double timeStep, int particleSubsteps)
    {
        const auto u = velocitySolution.copyToHost();
        gpupfem2::check(pfem_particles_advect(h_, u.at(0), u.at(1), timeStep, particleSubsteps));
    }
    void correctParticleVelocity(const deviceVector<double*>& velocitySolution,
                                 const deviceVector<double*>& velocitySolutionOld)
    {
        const auto u = velocitySolution.copyToHost();
        const auto uo = velocitySolutionOld.copyToHost();
        gpupfem2::check(pfem_particles_correct(h_, u.at(0), u.at(1), uo.at(0), uo.at(1)));
    }
    void projectVelocityOntoGrid(deviceVector<double*>& velocity)
    {
        const auto u = velocity.copyToHost();
        gpupfem2::check(pfem_particles_project(h_, u.at(0), u.at(1)));
    }
    int getParticleCount() const
    {
        long c = 0;
        gpupfem2::check(pfem_particles_count(h_, &c));
        return (int)c;
    }
    /* host copies of positions and velocities (x, y pairs) */
    void downloadParticles(std::vector<double>& pos, std::vector<double>& vel) const
    {
        const int n = getParticleCount();
        pos.resize(2 * (size_t)n);
        vel.resize(2 * (size_t)n);
        gpupfem2::check(pfem_particles_download(h_, nullptr, nullptr, pos.data(), vel.data(), nullptr));
    }
    pfem_particles handle() const { return h_; }

private:
    const Mesh2D* mesh;
    int level_;
    pfem_particles h_ = nullptr;
};

/* ------------------------------------------------------------------ export */
class DataExport {
public:
    DataExport(const Mesh2D& mesh_, const ParticleHandler2D* particleHandler_ = nullptr)
        : mesh(mesh_), particleHandler(particleHandler_)
    {
    }
    void addScalarDataVector(const deviceVector<double>& dataVector, const std::string& fieldname)
    {
        scalarDataVectors[fieldname] = dataVector.data;
    }
    /* DataExport::exportToVTK (data_export.cu:17-95) */
    void exportToVTK(const std::string& filename) const
    {
        std::vector<const char*> names;
        std::vector<const double*> ptrs;
        for (const auto& kv : scalarDataVectors) {
            names.push_back(kv.first.c_str());
            ptrs.push_back(kv.second);
        }
        const int rc = pfem_export_vtk(mesh.handle(), filename.c_str(), (int)names.size(), names.data(), ptrs.data());
        if (rc == PFEM_OK) std::printf("Data saved to %s\n", filename.c_str());
        else std::printf("Error while saving data to a file\n");
    }
    /* DataExport::exportParticlesToVTK (data_export.cu:97-171): points, one vertex cell per particle,
     * velocity as point data */
    void exportParticlesToVTK(const std::string& filename)
    {
        if (!particleHandler) return;
        std::vector<double> pos, vel;
        particleHandler->downloadParticles(pos, vel);
        const size_t np = pos.size() / 2;
        std::ofstream out(filename.c_str());
        if (!out.is_open()) {
            std::printf("Error while saving particles to a file\n");
            return;
        }
        out << "<?xml version=\"1.0\" ?> \n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
        out << "  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"" << np << "\" NumberOfCells=\"" << np << "\">\n";
        out << "      <Points>\n        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
        for (size_t p = 0; p < np; ++p) out << "          " << pos[2 * p] << " " << pos[2 * p + 1] << " 0.0\n";
        out << "        </DataArray>\n      </Points>\n      <Cells>\n";
        out << "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n        ";
        for (size_t p = 0; p < np; ++p) out << "  " << p;
        out << "\n        </DataArray>\n        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n        ";
        for (size_t p = 0; p < np; ++p) out << "  " << p + 1;
        out << "\n        </DataArray>\n        <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n        ";
        for (size_t p = 0; p < np; ++p) out << "  " << 1;
        out << "\n        </DataArray>\n      </Cells>\n      <PointData Scalars=\"scalars\">\n";
        out << "        <DataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
        for (size_t p = 0; p < np; ++p) out << "          " << vel[2 * p] << " " << vel[2 * p + 1] << " 0.0\n";
        out << "        </DataArray>\n      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
        std::printf("Particles saved to %s\n", filename.c_str());
    }

private:
    const Mesh2D& mesh;
    const ParticleHandler2D* particleHandler;
    std::map<std::string, const double*> scalarDataVectors;
};

/* DataExport3D (data_export_3d.cuh:9-25): nodal scalar fields of a tetrahedral mesh to a VTU file */
class DataExport3D {
public:
    explicit DataExport3D(const Mesh3D& mesh_) : mesh(mesh_) {}
    void addScalarDataVector(const deviceVector<double>& dataVector, const std::string& fieldname)
    {
        scalarDataVectors[fieldname] = dataVector.data;
    }
    /* DataExport3D::exportToVTK (data_export_3d.cu:15-93) */
    void exportToVTK(const std::string& filename) const
    {
        std::vector<const char*> names;
        std::vector<const double*> ptrs;
        for (const auto& kv : scalarDataVectors) {
            names.push_back(kv.first.c_str());
            ptrs.push_back(kv.second);
        }
        const int rc = pfem_export_vtk3d(mesh.handle(), filename.c_str(), (int)names.size(), names.data(), ptrs.data());
        if (rc == PFEM_OK) std::printf("Mesh solution saved to %s\n", filename.c_str());
        else std::printf("Error while saving mesh solution to a file\n");
    }

private:
    const Mesh3D& mesh;
    std::map<std::string, const double*> scalarDataVectors;
};

/* ------------------------------------------------------------------ timing */
/* GpuTimer (gpu_timer.cuh:14-98): wall time between start and stop with the device synchronised */
class GpuTimer {
public:
    void start()
    {
        gpupfem2::check(pfem_ctx_synchronize(gpupfem2::context()));
        t0_ = std::chrono::steady_clock::now();
    }
    double stop(const std::string& label)
    {
        gpupfem2::check(pfem_ctx_synchronize(gpupfem2::context()));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
        std::printf("%s: %.3f ms\n", label.c_str(), ms);
        return ms;
    }

private:
    std::chrono::steady_clock::time_point t0_;
};

#endif  // GPUPFEM2_HPP
