// Context, device memory, mesh import and O(T) setup builders, geometry kernels.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "device.hpp"
#include "internal.hpp"

namespace pfem {

void h2d(Context* c, void* d, const void* h, size_t bytes)
{
    if (bytes == 0) return;
    PFEM_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));   // host buffer is only borrowed for the call
}

void d2h_sync(Context* c, void* h, const void* d, size_t bytes)
{
    if (bytes == 0) return;
    PFEM_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
}

void zero_async(Context* c, void* d, size_t bytes)
{
    if (bytes == 0) return;
    PFEM_HIP(hipMemsetAsync(d, 0, bytes, c->stream));
}

// ------------------------------------------------------------------ SpMV row tiles
// Rows are cut into tiles of at most kBlock rows and at most kTileNnz non-zeros (a longer single row
// gets a tile of its own); one workgroup streams one tile through LDS. The table holds {first row, end
// row} per workgroup, dealt so that the tiles of XCD x (workgroup b runs on XCD b mod 8) form the x-th
// contiguous band of rows: neighbouring rows share their gathered vector entries in one L2.
constexpr int kTileNnz = 2048;

std::shared_ptr<Pattern> make_pattern(Context* c, int n, std::vector<int>&& rowptr, std::vector<int>&& col)
{
    auto P = std::make_shared<Pattern>();
    P->n = n;
    P->nnz = rowptr[n];
    P->h_rowptr = std::move(rowptr);
    P->h_col = std::move(col);
    std::vector<int> bounds;
    bounds.push_back(0);
    int r = 0;
    while (r < n) {
        int r1 = r;
        const int base = P->h_rowptr[r];
        while (r1 < n && r1 - r < kBlock && (P->h_rowptr[r1 + 1] - base <= kTileNnz || r1 == r)) ++r1;
        bounds.push_back(r1);
        r = r1;
    }
    const int m = (int)bounds.size() - 1;
    P->ntiles = m;
    constexpr int kXcds = 8;
    std::vector<int> tiles;
    tiles.reserve(2 * (size_t)m);
    {
        std::vector<int> next(kXcds), end(kXcds);
        for (int x = 0; x < kXcds; ++x) {
            next[x] = (int)((long)m * x / kXcds);
            end[x] = (int)((long)m * (x + 1) / kXcds);
        }
        std::vector<char> used(m, 0);
        int spill = 0;
        for (int b = 0; b < m; ++b) {
            const int x = b % kXcds;
            int q;
            while (next[x] < end[x] && used[next[x]]) ++next[x];
            if (next[x] < end[x]) q = next[x]++;
            else {
                while (used[spill]) ++spill;
                q = spill;
            }
            used[q] = 1;
            tiles.push_back(bounds[q]);
            tiles.push_back(bounds[q + 1]);
        }
    }
    P->h_diag.assign(n, -1);
    for (int i = 0; i < n; ++i)
        for (int k = P->h_rowptr[i]; k < P->h_rowptr[i + 1]; ++k)
            if (P->h_col[k] == i) { P->h_diag[i] = k; break; }
    upload(c, P->rowptr, P->h_rowptr);
    upload(c, P->col, P->h_col);
    upload(c, P->tiles, tiles);
    upload(c, P->diag, P->h_diag);
    return P;
}

// CSR pattern of SparseMatrixCSR::initialize (src/linear_algebra/sparse_matrix.cu:8-99): row i holds i and
// every vertex sharing a cell with i, ascending. Built from the vertex -> cell incidence in O(nnz log).
static void build_pattern(int nv, int nt, int nper, const std::vector<int>& cells, std::vector<int>& rowptr,
                          std::vector<int>& col)
{
    std::vector<int> off(nv + 1, 0);
    for (int c = 0; c < nt; ++c)
        for (int a = 0; a < nper; ++a) off[cells[(size_t)nper * c + a] + 1]++;
    for (int v = 0; v < nv; ++v) off[v + 1] += off[v];
    std::vector<int> inc(off[nv]);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int c = 0; c < nt; ++c)
        for (int a = 0; a < nper; ++a) inc[fill[cells[(size_t)nper * c + a]]++] = c;
    rowptr.assign(nv + 1, 0);
    col.clear();
    col.reserve((size_t)nv * (nper == 3 ? 7 : 15));
    std::vector<int> buf;
    for (int v = 0; v < nv; ++v) {
        buf.clear();
        buf.push_back(v);
        for (int k = off[v]; k < off[v + 1]; ++k)
            for (int a = 0; a < nper; ++a) buf.push_back(cells[(size_t)nper * inc[k] + a]);
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        col.insert(col.end(), buf.begin(), buf.end());
        rowptr[v + 1] = (int)col.size();
    }
}

// Cell neighbour lists of Mesh2D::fillCellNeighborIndices (src/mesh_2d.cu:107-139): cells sharing at
// least one vertex, ascending -- here in O(T) through the vertex -> cell incidence.
static void build_neighbors(int nv, int nt, const std::vector<int>& tri, std::vector<int>& off_out,
                            std::vector<int>& idx_out)
{
    std::vector<int> off(nv + 1, 0);
    for (int c = 0; c < nt; ++c)
        for (int a = 0; a < 3; ++a) off[tri[3 * (size_t)c + a] + 1]++;
    for (int v = 0; v < nv; ++v) off[v + 1] += off[v];
    std::vector<int> inc(off[nv]);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int c = 0; c < nt; ++c)
        for (int a = 0; a < 3; ++a) inc[fill[tri[3 * (size_t)c + a]]++] = c;
    off_out.assign(nt + 1, 0);
    idx_out.clear();
    idx_out.reserve((size_t)nt * 13);
    std::vector<int> buf;
    for (int c = 0; c < nt; ++c) {
        buf.clear();
        for (int a = 0; a < 3; ++a) {
            const int v = tri[3 * (size_t)c + a];
            for (int k = off[v]; k < off[v + 1]; ++k)
                if (inc[k] != c) buf.push_back(inc[k]);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        idx_out.insert(idx_out.end(), buf.begin(), buf.end());
        off_out[c + 1] = (int)idx_out.size();
    }
}

// node -> (cell, local) incidence over the device cells, cells ascending, with the CSR slots of
// (node, vertex j of the cell) for the row-gather assembly.
static void build_incidence(Context* ctx, int nv, int nt, int nper, const std::vector<int>& cells,
                            const Pattern& P, Incidence& inc)
{
    inc.nper = nper;
    inc.h_off.assign(nv + 1, 0);
    for (int c = 0; c < nt; ++c)
        for (int a = 0; a < nper; ++a) inc.h_off[cells[(size_t)nper * c + a] + 1]++;
    for (int v = 0; v < nv; ++v) inc.h_off[v + 1] += inc.h_off[v];
    const size_t ninc = inc.h_off[nv];
    inc.h_code.assign(ninc, 0);
    std::vector<int> slots(ninc * nper);
    std::vector<int> fill(inc.h_off.begin(), inc.h_off.end() - 1);
    for (int c = 0; c < nt; ++c)
        for (int a = 0; a < nper; ++a) {
            const int v = cells[(size_t)nper * c + a];
            const int k = fill[v]++;
            inc.h_code[k] = c * 4 + a;
            for (int j = 0; j < nper; ++j) {
                const int w = cells[(size_t)nper * c + j];
                const int* b = P.h_col.data() + P.h_rowptr[v];
                const int* e = P.h_col.data() + P.h_rowptr[v + 1];
                const int* it = std::lower_bound(b, e, w);
                if (it == e || *it != w) throw Error(PFEM_EINVAL, "mesh cell vertex missing from CSR row");
                slots[(size_t)k * nper + j] = (int)(it - P.h_col.data());
            }
        }
    upload(ctx, inc.off, inc.h_off);
    upload(ctx, inc.code, inc.h_code);
    upload(ctx, inc.slots, slots);
}

// ------------------------------------------------------------------ geometry kernels
// kCalculateCellArea (src/mesh_2d.cu:10-19) and kCalculateInvJacobi (:21-34) with
// Matrix2x2::inverse (src/common/cuda_math.cuh:124-141); invJ row-major.
__global__ void k_geometry2d(int nt, const double* __restrict__ xy, const int* __restrict__ tri,
                             double* __restrict__ area, double* __restrict__ invJ)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nt) return;
    const int a = tri[3 * c], b = tri[3 * c + 1], d = tri[3 * c + 2];
    const double ax = xy[2 * a], ay = xy[2 * a + 1], bx = xy[2 * b], by = xy[2 * b + 1];
    const double dx = xy[2 * d], dy = xy[2 * d + 1];
    const double v12x = bx - ax, v12y = by - ay, v13x = dx - ax, v13y = dy - ay;
    area[c] = fabs(v12x * v13y - v12y * v13x) * 0.5;
    const double j00 = ax - dx, j01 = ay - dy, j10 = bx - dx, j11 = by - dy;
    const double det = j00 * j11 - j01 * j10;
    const double inv = 1.0 / det;
    double4 o;
    o.x = j11 * inv;
    o.y = -j01 * inv;
    o.z = -j10 * inv;
    o.w = j00 * inv;
    reinterpret_cast<double4*>(invJ)[c] = o;
}

// kCalculateCellVolume / kCalculateInvJacobi (src/mesh_3d.cu:12-35), GenericMatrix3x3::inverse
// (src/common/matrix3x3.cuh:127-154); zero_small reproduces the |det| < 2e-6 zero inverse.
__global__ void k_geometry3d(int nt, const double* __restrict__ p, const int* __restrict__ tet,
                             double* __restrict__ vol, double* __restrict__ invJ, int zero_small)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nt) return;
    const int i3 = tet[4 * c + 3];
    double m[9];
    for (int k = 0; k < 3; ++k) {
        const int ik = tet[4 * c + k];
        for (int d = 0; d < 3; ++d) m[3 * k + d] = p[3 * ik + d] - p[3 * i3 + d];
    }
    const double cx = m[4] * m[8] - m[5] * m[7];
    const double cy = m[5] * m[6] - m[3] * m[8];
    const double cz = m[3] * m[7] - m[4] * m[6];
    vol[c] = fabs(m[0] * cx + m[1] * cy + m[2] * cz) * 0.1666666666666667;
    const double det = (m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                        m[2] * (m[3] * m[7] - m[4] * m[6]));
    double* o = invJ + 9 * (size_t)c;
    if (zero_small && fabs(det) < 2e-6) {
        for (int k = 0; k < 9; ++k) o[k] = 0.0;
        return;
    }
    const double id = 1.0 / det;
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

// Face neighbours of Mesh3D::fillCellNeighborIndices (src/mesh_3d.cu:37-82, 198-227): the cells whose
// vertex lists have exactly 3 equal pairs with this cell's, ascending. The reference tests all T^2 pairs;
// here one lane per cell tests only the cells of its vertices' incidence lists (every cell that shares a
// pair is in one of them), each candidate once: at the first vertex of this cell it contains, and at that
// vertex's first slot in the candidate. Pass 1 counts, pass 2 writes and sorts the short segment.
template <bool FILL>
__global__ void k_face_neighbors3d(int nt, const int* __restrict__ tet, const int* __restrict__ ioff,
                                   const int* __restrict__ icode, int* __restrict__ cnt_or_off, int* __restrict__ idx)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nt) return;
    int a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = tet[4 * (size_t)c + i];
    const int base = FILL ? cnt_or_off[c] : 0;
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        bool repeat = false;
        for (int j = 0; j < i; ++j) repeat |= a[j] == a[i];
        if (repeat) continue;
        for (int e = ioff[a[i]]; e < ioff[a[i] + 1]; ++e) {
            const int code = icode[e];
            const int d = code >> 2, slot = code & 3;
            if (d == c) continue;
            int b[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = tet[4 * (size_t)d + j];
            bool skip = false;
            for (int j = 0; j < slot; ++j) skip |= b[j] == a[i];      // d listed again for a repeated vertex
            for (int j = 0; j < i; ++j)                              // d already met at an earlier vertex
                skip |= a[j] == b[0] || a[j] == b[1] || a[j] == b[2] || a[j] == b[3];
            if (skip) continue;
            int pairs = 0;
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) pairs += a[x] == b[y];
            if (pairs == 3) {
                if (FILL) idx[base + k] = d;
                ++k;
            }
        }
    }
    if (!FILL) {
        cnt_or_off[c] = k;
        return;
    }
    for (int x = 1; x < k; ++x) {   // insertion sort of the segment (at most 4 entries on a valid mesh)
        const int v = idx[base + x];
        int y = x - 1;
        while (y >= 0 && idx[base + y] > v) {
            idx[base + y + 1] = idx[base + y];
            --y;
        }
        idx[base + y + 1] = v;
    }
}

// ------------------------------------------------------------------ DAT import
// SALOME DAT reader semantics of Mesh2D/Mesh3D::loadMeshFromFile (src/mesh_2d.cu:36-96,
// src/mesh_3d.cu:84-147): header "nV nElems"; vertex records "id x y z"; element records
// "id type n1..nk" (1-based nodes); the wanted type (203 in 2D, 304 in 3D) is kept, other records are
// skipped and decrement the cell count. A file ending in whitespace makes the reference's eof loop run
// once more and decrement the count again: the device mesh then lacks the last parsed cell.
struct DatMesh {
    int nv = 0, ndev = 0;
    std::vector<double> xyz;   // 3 per vertex
    std::vector<int> cells;    // nper per parsed cell
};

static bool read_dat(const char* path, int dim, double scale, DatMesh& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::istringstream in(text);
    long nv, ncells;
    if (!(in >> nv >> ncells)) return false;
    out.nv = (int)nv;
    out.xyz.assign(3 * (size_t)nv, 0.0);
    for (long i = 0; i < nv; ++i) {
        long id;
        double x, y, z;
        if (!(in >> id >> x >> y >> z)) return false;
        out.xyz[3 * i] = scale * x;
        out.xyz[3 * i + 1] = scale * y;
        out.xyz[3 * i + 2] = dim == 3 ? scale * z : 0.0;
    }
    const int want = dim == 2 ? 203 : 304;
    const int nper = dim == 2 ? 3 : 4;
    long id, type;
    while (in >> id >> type) {
        if (type == want) {
            for (int k = 0; k < nper; ++k) {
                long v;
                if (!(in >> v)) return false;
                out.cells.push_back((int)(v - 1));
            }
        } else {
            --ncells;
            const int skip = (type == 203 && dim == 3) ? 3 : 2;
            for (int k = 0; k < skip; ++k) {
                long v;
                in >> v;
            }
        }
    }
    if (!text.empty() && std::isspace((unsigned char)text.back())) --ncells;
    out.ndev = (int)std::min<long>(ncells, (long)(out.cells.size() / nper));
    return true;
}

}  // namespace pfem

using namespace pfem;

static pfem_mesh2d_t* mesh2d_build(Context* c, int nv, std::vector<double>&& xy, std::vector<int>&& host_tri,
                                   int ndev, int fill_neighbors)
{
    auto m = std::make_unique<pfem_mesh2d_t>();
    m->ctx = c;
    m->nv = nv;
    m->nt_host = (int)(host_tri.size() / 3);
    m->nt = ndev;
    m->h_xy = std::move(xy);
    m->h_tri = std::move(host_tri);
    for (int v : m->h_tri) PFEM_REQUIRE(v >= 0 && v < nv, "cell vertex index out of range");
    std::vector<int> dtri(m->h_tri.begin(), m->h_tri.begin() + 3 * (size_t)m->nt);
    upload(c, m->xy, m->h_xy);
    upload(c, m->tri, dtri);
    m->h_edge_ids.assign(3 * (size_t)m->nt, -1);
    upload(c, m->edge_ids, m->h_edge_ids);
    m->area.alloc(m->nt);
    m->invJ.alloc(4 * (size_t)m->nt);
    if (m->nt > 0)
        hipLaunchKernelGGL(k_geometry2d, dim3(blocks_for(m->nt)), dim3(kBlock), 0, c->stream, m->nt, m->xy.p,
                           m->tri.p, m->area.p, m->invJ.p);
    PFEM_HIP(hipGetLastError());
    if (fill_neighbors) {
        std::vector<int> off, idx;
        build_neighbors(nv, m->nt, dtri, off, idx);
        m->n_nbr = idx.size();
        upload(c, m->nbr_off, off);
        upload(c, m->nbr_idx, idx);
    }
    std::vector<int> rowptr, col;
    build_pattern(nv, m->nt_host, 3, m->h_tri, rowptr, col);
    m->pat = make_pattern(c, nv, std::move(rowptr), std::move(col));
    build_incidence(c, nv, m->nt, 3, dtri, *m->pat, m->inc);
    PFEM_HIP(hipStreamSynchronize(c->stream));
    return m.release();
}

static pfem_mesh3d_t* mesh3d_build(Context* c, int nv, std::vector<double>&& xyz, std::vector<int>&& tet, int ndev,
                                   int zero_small)
{
    auto m = std::make_unique<pfem_mesh3d_t>();
    m->ctx = c;
    m->nv = nv;
    m->nt = ndev;
    m->h_xyz = std::move(xyz);
    m->h_tet = std::move(tet);
    for (int v : m->h_tet) PFEM_REQUIRE(v >= 0 && v < nv, "cell vertex index out of range");
    std::vector<int> dtet(m->h_tet.begin(), m->h_tet.begin() + 4 * (size_t)m->nt);
    upload(c, m->xyz, m->h_xyz);
    upload(c, m->tet, dtet);
    m->vol.alloc(m->nt);
    m->invJ.alloc(9 * (size_t)m->nt);
    if (m->nt > 0)
        hipLaunchKernelGGL(k_geometry3d, dim3(blocks_for(m->nt)), dim3(kBlock), 0, c->stream, m->nt, m->xyz.p,
                           m->tet.p, m->vol.p, m->invJ.p, zero_small);
    PFEM_HIP(hipGetLastError());
    std::vector<int> rowptr, col;
    build_pattern(nv, (int)(m->h_tet.size() / 4), 4, m->h_tet, rowptr, col);
    m->pat = make_pattern(c, nv, std::move(rowptr), std::move(col));
    build_incidence(c, nv, m->nt, 4, dtet, *m->pat, m->inc);
    PFEM_HIP(hipStreamSynchronize(c->stream));
    return m.release();
}

// ------------------------------------------------------------------ C ABI (mesh and memory part)
extern "C" {

int pfem_ctx_create(int device, pfem_ctx* out)
{
    if (!out) return PFEM_EINVAL;
    auto* c = new Context;
    if (hipSetDevice(device) != hipSuccess) { delete c; return PFEM_EHIP; }
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return PFEM_EHIP; }
    if (hipHostMalloc((void**)&c->pinned_i, 64 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->pinned_d, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&c->la_part, 2 * kLaMaxBlocks * sizeof(double)) != hipSuccess ||
        noalloc_init(c) != hipSuccess) {
        (void)hipFree(c->la_part);
        (void)hipHostFree(c->pinned_i);
        (void)hipHostFree(c->pinned_d);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PFEM_EHIP;
    }
    *out = c;
    return PFEM_OK;
}

int pfem_ctx_destroy(pfem_ctx ctx)
{
    if (!ctx) return PFEM_OK;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipHostFree(ctx->pinned_i);
    (void)hipHostFree(ctx->pinned_d);
    (void)hipFree(ctx->la_part);
    delete C(ctx);
    return PFEM_OK;
}

const char* pfem_last_error(pfem_ctx ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pfem_ctx_synchronize(pfem_ctx ctx)
{
    PFEM_TRY(ctx, PFEM_HIP(hipStreamSynchronize(ctx->stream)));
}

void* pfem_ctx_stream(pfem_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int pfem_malloc(pfem_ctx ctx, size_t bytes, void** d_out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(d_out, "null output");
        *d_out = nullptr;
        if (bytes) PFEM_HIP(hipMalloc(d_out, bytes));
    });
}

int pfem_free(pfem_ctx ctx, void* d_ptr)
{
    PFEM_TRY(ctx, {
        if (d_ptr) {
            PFEM_HIP(hipStreamSynchronize(ctx->stream));
            PFEM_HIP(hipFree(d_ptr));
        }
    });
}

int pfem_memcpy_h2d(pfem_ctx ctx, void* d_dst, const void* src, size_t bytes)
{
    PFEM_TRY(ctx, h2d(C(ctx), d_dst, src, bytes));
}

int pfem_memcpy_d2h(pfem_ctx ctx, void* dst, const void* d_src, size_t bytes)
{
    PFEM_TRY(ctx, d2h_sync(C(ctx), dst, d_src, bytes));
}

int pfem_memcpy_d2d(pfem_ctx ctx, void* d_dst, const void* d_src, size_t bytes)
{
    PFEM_TRY(ctx, if (bytes) PFEM_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream)));
}

int pfem_memset_zero(pfem_ctx ctx, void* d_ptr, size_t bytes)
{
    PFEM_TRY(ctx, zero_async(C(ctx), d_ptr, bytes));
}

// Node and cell order for locality (host only): the nodes in reverse Cuthill-McKee order of the vertex
// adjacency graph, the cells by the smallest new index among their vertices (ties: the next smallest,
// then the old index), so that neighbouring rows, cells and the particles sorted by cell lie close in
// memory whatever the input numbering.
int pfem_mesh_locality_order(int nv, int nt, int nvc, const int* cells, int* node_order, int* cell_order)
{
    PFEM_TRY(nullptr, {
        PFEM_REQUIRE(nv > 0 && nt >= 0 && (nvc == 3 || nvc == 4) && (cells || nt == 0) && node_order &&
                         (cell_order || nt == 0),
                     "bad arguments");
        std::vector<std::vector<int>> adj(nv);
        for (long c = 0; c < nt; ++c)
            for (int a = 0; a < nvc; ++a) {
                const int u = cells[c * nvc + a];
                PFEM_REQUIRE(u >= 0 && u < nv, "cell vertex out of range");
                for (int b = 0; b < nvc; ++b) adj[u].push_back(cells[c * nvc + b]);
            }
        Pattern P;
        P.n = nv;
        P.h_rowptr.assign(nv + 1, 0);
        for (int u = 0; u < nv; ++u) {
            auto& a = adj[u];
            a.push_back(u);
            std::sort(a.begin(), a.end());
            a.erase(std::unique(a.begin(), a.end()), a.end());
            P.h_rowptr[u + 1] = P.h_rowptr[u] + (int)a.size();
        }
        P.h_col.reserve(P.h_rowptr[nv]);
        for (auto& a : adj) {
            P.h_col.insert(P.h_col.end(), a.begin(), a.end());
            std::vector<int>().swap(a);
        }
        P.nnz = P.h_rowptr[nv];
        const std::vector<int> order = rcm_order(P);
        std::vector<int> pos(nv);
        for (int k = 0; k < nv; ++k) {
            node_order[k] = order[k];
            pos[order[k]] = k;
        }
        if (nt > 0) {
            std::vector<std::array<int, 4>> key(nt);
            for (long c = 0; c < nt; ++c) {
                std::array<int, 4> k{INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
                for (int a = 0; a < nvc; ++a) k[a] = pos[cells[c * nvc + a]];
                std::sort(k.begin(), k.begin() + nvc);
                key[c] = k;
            }
            std::vector<int> idx(nt);
            for (int c = 0; c < nt; ++c) idx[c] = c;
            std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key[x] < key[y]; });
            for (int c = 0; c < nt; ++c) cell_order[c] = idx[c];
        }
    });
}

int pfem_mesh2d_load_dat(pfem_ctx ctx, const char* path, int fill_neighbors, double scale, pfem_mesh2d* out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(path && out, "null argument");
        DatMesh d;
        if (!read_dat(path, 2, scale, d)) throw Error(PFEM_EIO, std::string("cannot read mesh file ") + path);
        std::vector<double> xy(2 * (size_t)d.nv);
        for (int i = 0; i < d.nv; ++i) { xy[2 * i] = d.xyz[3 * i]; xy[2 * i + 1] = d.xyz[3 * i + 1]; }
        *out = mesh2d_build(C(ctx), d.nv, std::move(xy), std::move(d.cells), d.ndev, fill_neighbors);
    });
}

int pfem_mesh2d_create(pfem_ctx ctx, int nv, const double* xy, int nt, const int* tri, int fill_neighbors,
                       pfem_mesh2d* out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(out && xy && tri && nv > 0 && nt > 0, "bad mesh arguments");
        std::vector<double> hxy(xy, xy + 2 * (size_t)nv);
        std::vector<int> htri(tri, tri + 3 * (size_t)nt);
        *out = mesh2d_build(C(ctx), nv, std::move(hxy), std::move(htri), nt, fill_neighbors);
    });
}

int pfem_mesh2d_destroy(pfem_mesh2d m)
{
    if (!m) return PFEM_OK;
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
    return PFEM_OK;
}

int pfem_mesh2d_info(pfem_mesh2d m, int* nv, int* nt, int* nt_host, long* n_neighbors)
{
    if (!m) return PFEM_EINVAL;
    if (nv) *nv = m->nv;
    if (nt) *nt = m->nt;
    if (nt_host) *nt_host = m->nt_host;
    if (n_neighbors) *n_neighbors = m->n_nbr;
    return PFEM_OK;
}

int pfem_mesh2d_download(pfem_mesh2d m, int what, void* dst)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && dst, "null argument");
        Context* c = C(m->ctx);
        switch (what) {
        case 0: d2h_sync(c, dst, m->xy.p, m->xy.bytes()); break;
        case 1: d2h_sync(c, dst, m->tri.p, m->tri.bytes()); break;
        case 2: d2h_sync(c, dst, m->area.p, m->area.bytes()); break;
        case 3: d2h_sync(c, dst, m->invJ.p, m->invJ.bytes()); break;
        case 4: d2h_sync(c, dst, m->nbr_off.p, m->nbr_off.bytes()); break;
        case 5: d2h_sync(c, dst, m->nbr_idx.p, m->nbr_idx.bytes()); break;
        case 6: d2h_sync(c, dst, m->edge_ids.p, m->edge_ids.bytes()); break;
        case 7: std::memcpy(dst, m->h_tri.data(), m->h_tri.size() * sizeof(int)); break;
        default: throw Error(PFEM_EINVAL, "unknown mesh array");
        }
    });
}

int pfem_mesh2d_device_ptr(pfem_mesh2d m, int what, void** d_ptr)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && d_ptr, "null argument");
        switch (what) {
        case 0: *d_ptr = m->xy.p; break;
        case 1: *d_ptr = m->tri.p; break;
        case 2: *d_ptr = m->area.p; break;
        case 3: *d_ptr = m->invJ.p; break;
        case 4: *d_ptr = m->nbr_off.p; break;
        case 5: *d_ptr = m->nbr_idx.p; break;
        case 6: *d_ptr = m->edge_ids.p; break;
        default: throw Error(PFEM_EINVAL, "unknown mesh array");
        }
    });
}

int pfem_mesh2d_set_edge_ids(pfem_mesh2d m, const int* ids)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && ids, "null argument");
        m->h_edge_ids.assign(ids, ids + 3 * (size_t)m->nt);
        upload(C(m->ctx), m->edge_ids, m->h_edge_ids);
        m->body_id_cached = -1000;
        ++m->generation;
    });
}

// Edge classification by midpoint, in the if/else order of kSetEdgeBoundaryIDs
// (cases/Cylinder2D/main.cu:52-68): |m.x - x0| < 2e-6, |m.x - x1| < 2e-6, |m.y - y0| < 2e-6 or
// |m.y - y1| < 2e-6, then the circle test. Runs on the host (setup only).
int pfem_mesh2d_classify_edges(pfem_mesh2d m, double x0, double x1, double y0, double y1, int id_x0, int id_x1,
                               int id_y0, int id_y1, double cx, double cy, double r, double tol, int id_circle)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m, "null mesh");
        std::vector<int> ids(3 * (size_t)m->nt, -1);
        for (int c = 0; c < m->nt; ++c)
            for (int i = 0; i < 3; ++i) {
                const int a = m->h_tri[3 * (size_t)c + i], b = m->h_tri[3 * (size_t)c + (i + 1) % 3];
                const double mx = (m->h_xy[2 * (size_t)a] + m->h_xy[2 * (size_t)b]) * 0.5;
                const double my = (m->h_xy[2 * (size_t)a + 1] + m->h_xy[2 * (size_t)b + 1]) * 0.5;
                int id = -1;
                if (std::fabs(mx - x0) < 2e-6) id = id_x0;
                else if (std::fabs(mx - x1) < 2e-6) id = id_x1;
                else if (std::fabs(my - y0) < 2e-6) id = id_y0;
                else if (std::fabs(my - y1) < 2e-6) id = id_y1;
                else if (r > 0.0) {
                    const double dx = mx - cx, dy = my - cy;
                    if (std::fabs(std::sqrt(dx * dx + dy * dy) - r) < tol) id = id_circle;
                }
                ids[3 * (size_t)c + i] = id;
            }
        m->h_edge_ids = ids;
        upload(C(m->ctx), m->edge_ids, m->h_edge_ids);
        m->body_id_cached = -1000;
        ++m->generation;
    });
}

int pfem_mesh3d_load_dat(pfem_ctx ctx, const char* path, double scale, int zero_small_det, pfem_mesh3d* out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(path && out, "null argument");
        DatMesh d;
        if (!read_dat(path, 3, scale, d)) throw Error(PFEM_EIO, std::string("cannot read mesh file ") + path);
        *out = mesh3d_build(C(ctx), d.nv, std::move(d.xyz), std::move(d.cells), d.ndev, zero_small_det);
    });
}

int pfem_mesh3d_create(pfem_ctx ctx, int nv, const double* xyz, int nt, const int* tet, int zero_small_det,
                       pfem_mesh3d* out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(out && xyz && tet && nv > 0 && nt > 0, "bad mesh arguments");
        std::vector<double> hx(xyz, xyz + 3 * (size_t)nv);
        std::vector<int> ht(tet, tet + 4 * (size_t)nt);
        *out = mesh3d_build(C(ctx), nv, std::move(hx), std::move(ht), nt, zero_small_det);
    });
}

int pfem_mesh3d_destroy(pfem_mesh3d m)
{
    if (!m) return PFEM_OK;
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
    return PFEM_OK;
}

int pfem_mesh3d_info(pfem_mesh3d m, int* nv, int* nt)
{
    if (!m) return PFEM_EINVAL;
    if (nv) *nv = m->nv;
    if (nt) *nt = m->nt;
    return PFEM_OK;
}

int pfem_mesh3d_fill_neighbors(pfem_mesh3d m, long* n_neighbors)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m, "null argument");
        Context* c = C(m->ctx);
        const int nt = m->nt;
        if (m->nbr_off.n != (size_t)nt + 1) {
            DevArray<int> cnt(std::max(nt, 1));
            m->nbr_off.alloc((size_t)nt + 1);
            const unsigned g = blocks_for(std::max(nt, 1));
            hipLaunchKernelGGL((k_face_neighbors3d<false>), dim3(g), dim3(kBlock), 0, c->stream, nt, m->tet.p,
                               m->inc.off.p, m->inc.code.p, cnt.p, (int*)nullptr);
            PFEM_HIP(hipGetLastError());
            exclusive_scan_int(c, cnt.p, m->nbr_off.p, nt, m->nbr_off.p + nt);
            int total = 0;
            d2h_sync(c, &total, m->nbr_off.p + nt, sizeof(int));
            m->nbr_idx.alloc(std::max(total, 1));
            m->n_nbr = total;
            hipLaunchKernelGGL((k_face_neighbors3d<true>), dim3(g), dim3(kBlock), 0, c->stream, nt, m->tet.p,
                               m->inc.off.p, m->inc.code.p, m->nbr_off.p, m->nbr_idx.p);
            PFEM_HIP(hipGetLastError());
            PFEM_HIP(hipStreamSynchronize(c->stream));
        }
        if (n_neighbors) *n_neighbors = m->n_nbr;
    });
}

int pfem_mesh3d_device_ptr(pfem_mesh3d m, int what, void** d_ptr)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && d_ptr, "null argument");
        switch (what) {
        case 0: *d_ptr = m->xyz.p; break;
        case 1: *d_ptr = m->tet.p; break;
        case 2: *d_ptr = m->vol.p; break;
        case 3: *d_ptr = m->invJ.p; break;
        case 4: *d_ptr = m->nbr_off.p; break;
        case 5: *d_ptr = m->nbr_idx.p; break;
        default: throw Error(PFEM_EINVAL, "unknown mesh array");
        }
    });
}

int pfem_mesh3d_download(pfem_mesh3d m, int what, void* dst)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && dst, "null argument");
        Context* c = C(m->ctx);
        switch (what) {
        case 0: d2h_sync(c, dst, m->xyz.p, m->xyz.bytes()); break;
        case 1: d2h_sync(c, dst, m->tet.p, m->tet.bytes()); break;
        case 2: d2h_sync(c, dst, m->vol.p, m->vol.bytes()); break;
        case 3: d2h_sync(c, dst, m->invJ.p, m->invJ.bytes()); break;
        case 4:
        case 5:
            PFEM_REQUIRE(m->nbr_off.n == (size_t)m->nt + 1, "face neighbours not built (pfem_mesh3d_fill_neighbors)");
            if (what == 4) d2h_sync(c, dst, m->nbr_off.p, m->nbr_off.bytes());
            else d2h_sync(c, dst, m->nbr_idx.p, sizeof(int) * (size_t)m->n_nbr);
            break;
        default: throw Error(PFEM_EINVAL, "unknown mesh array");
        }
    });
}

}  // extern "C"
