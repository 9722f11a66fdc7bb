// Internal declarations of libpfem2_hip.so (host side). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pfem2.h"

namespace pfem {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PFEM_HIP(x)                                                                                 \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw ::pfem::Error(PFEM_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));         \
    } while (0)

#define PFEM_REQUIRE(cond, msg)                                                                     \
    do {                                                                                            \
        if (!(cond)) throw ::pfem::Error(PFEM_EINVAL, msg);                                         \
    } while (0)

// Owning device allocation (move-only; the reference deviceVector's implicit shallow copy is a
// double-free hazard, src/common/device_vector.cuh:19-146).
template <class T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    DevArray() = default;
    explicit DevArray(size_t count) { alloc(count); }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    DevArray(DevArray&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevArray& operator=(DevArray&& o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevArray() { release(); }
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        PFEM_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct Context;

// Async copies on the context stream; the *_sync variants wait for completion.
void h2d(Context* c, void* d, const void* h, size_t bytes);
void d2h_sync(Context* c, void* h, const void* d, size_t bytes);
void zero_async(Context* c, void* d, size_t bytes);

// Shared CSR structure (one per mesh; every system of a PFEM-2 step uses it).
struct Pattern {
    int n = 0;
    long nnz = 0;
    std::vector<int> h_rowptr, h_col;
    DevArray<int> rowptr, col;
    DevArray<int> tiles;   // SpMV row tiles: {first row, end row} per workgroup (2 * ntiles entries)
    int ntiles = 0;
    DevArray<int> diag;    // position of the diagonal entry of each row
    std::vector<int> h_diag;
};

std::shared_ptr<Pattern> make_pattern(Context* c, int n, std::vector<int>&& rowptr, std::vector<int>&& col);
// Reverse Cuthill-McKee order of a pattern's graph (host arrays only): order[new] = old.
std::vector<int> rcm_order(const Pattern& A);

}  // namespace pfem

struct pfem_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int* pinned_i = nullptr;      // small pinned host buffers for scalar read-backs
    double* pinned_d = nullptr;
    double* scratch = nullptr;    // device address of this device's copy of the allocation-free scratch array
    double* la_part = nullptr;    // 2 * kLaMaxBlocks partials of the dot / nrm2 reductions
};

namespace pfem {
constexpr int kLaMaxBlocks = 1024;   // largest grid of the dot / nrm2 partial kernels
}

namespace pfem {
struct Context : pfem_context {};
inline Context* C(pfem_ctx c) { return static_cast<Context*>(c); }
// Resolves the device address of the allocation-free entry points' scratch array on the context's
// device (once per device); returns a HIP error code.
int noalloc_init(Context* c);
}  // namespace pfem

struct pfem_csr_t {
    pfem_ctx ctx = nullptr;
    std::shared_ptr<pfem::Pattern> pat;
    pfem::DevArray<double> val;
    pfem::DevArray<float> val32;
};

// node -> incident (cell, local vertex) list, cells ascending; plus for each incidence the CSR
// positions of (node, cell vertex j) for every vertex j of the cell (row-gather assembly scatter map)
struct Incidence {
    int nper = 3;
    pfem::DevArray<int> off;    // nv + 1
    pfem::DevArray<int> code;   // cell * 4 + local
    pfem::DevArray<int> slots;  // nper per incidence
    std::vector<int> h_off, h_code;
};

struct pfem_mesh2d_t {
    pfem_ctx ctx = nullptr;
    int nv = 0, nt = 0, nt_host = 0;
    std::vector<double> h_xy;
    std::vector<int> h_tri;        // host cells (all parsed cells)
    std::vector<int> h_edge_ids;
    pfem::DevArray<double> xy, area, invJ;
    pfem::DevArray<int> tri, edge_ids, nbr_off, nbr_idx;
    long n_nbr = 0;
    std::shared_ptr<pfem::Pattern> pat;
    Incidence inc;
    // body edges cache
    int body_id_cached = -1000;
    pfem::DevArray<int> body_edges;
    int n_body = 0;
    pfem::DevArray<double> body_loads;
    // counts reallocations of the device arrays above after creation (edge IDs, the body cache), so that
    // captured launch sequences that hold their addresses know to be captured again
    long generation = 0;
};

struct pfem_mesh3d_t {
    pfem_ctx ctx = nullptr;
    int nv = 0, nt = 0;
    std::vector<double> h_xyz;
    std::vector<int> h_tet;
    pfem::DevArray<double> xyz, vol, invJ;
    pfem::DevArray<int> tet;
    pfem::DevArray<int> nbr_off, nbr_idx;   // face neighbours, built on request
    long n_nbr = 0;
    std::shared_ptr<pfem::Pattern> pat;
    Incidence inc;
};

struct pfem_dirichlet_t {
    pfem_ctx ctx = nullptr;
    int n = 0, nbc = 0;
    std::vector<int> h_node;
    pfem::DevArray<int> node, map;     // map: node -> BC index or -1
    pfem::DevArray<double> value;
    // rows that hold at least one Dirichlet column, per pattern (cached for the last pattern used)
    const pfem::Pattern* rows_for = nullptr;
    pfem::DevArray<int> rows;
    int nrows = 0;
    pfem::DevArray<double> num, den;   // prediction BC accumulation scratch
};

namespace pfem {
template <class T>
inline void upload(Context* c, DevArray<T>& a, const std::vector<T>& h)
{
    a.alloc(h.size());
    h2d(c, a.p, h.data(), h.size() * sizeof(T));
}

// kernels launched from more than one translation unit
void launch_spmv(Context* c, const Pattern& P, const double* val, const double* x, double* y);
void launch_spmv32(Context* c, const Pattern& P, const float* val, const float* x, float* y);
void launch_fill(Context* c, double* d, double v, long n);
void launch_copy(Context* c, double* dst, const double* src, long n);
// dense BLAS-1/2 behind the LinearAlgebra entry points; scalars (res, alpha, beta) are in device memory
void launch_dot(Context* c, long n, const double* x, const double* y, double* res);
void launch_nrm2(Context* c, long n, const double* x, double* res);
void launch_trsv_upper(Context* c, int n, const double* a, long lda, double* x);
void launch_gemv(Context* c, int m, int n, const double* alpha, const double* a, long lda, const double* x,
                 const double* beta, double* y);
void exclusive_scan_int(Context* c, const int* in, int* out, long n, int* total_dev);
void dirichlet_rows(Context* c, pfem_dirichlet_t* bc, const Pattern& P);
}  // namespace pfem

// Wraps the body of a C entry point: exceptions become status codes plus pfem_last_error() text.
#define PFEM_TRY(ctxexpr, ...)                                                                      \
    pfem_ctx ctx__ = (ctxexpr);                                                                     \
    try {                                                                                           \
        __VA_ARGS__;                                                                                \
        return PFEM_OK;                                                                             \
    } catch (const pfem::Error& e) {                                                                \
        if (ctx__) ctx__->err = e.what();                                                           \
        return e.code;                                                                              \
    } catch (const std::exception& e) {                                                             \
        if (ctx__) ctx__->err = e.what();                                                           \
        return PFEM_EINVAL;                                                                         \
    }
