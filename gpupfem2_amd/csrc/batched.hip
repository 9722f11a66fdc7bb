// Batched independent PCG solves: many systems of uneven size, stored together as one block-diagonal CSR
// matrix (system s owns rows and columns [rows[s], rows[s+1])), solved by one sequence of launches. The
// reference solves one system per SolverCG::solve call (src/linear_algebra/linear_solver.cu:301-376);
// each system here runs that recurrence (Jacobi or no preconditioner, the stopping test on (r, M^-1 r))
// with its own scalars and stops on its own.
//
// Work is cut into SpMV row tiles (<= 256 rows, <= 2048 non-zeros) that never straddle two systems, and
// every vector kernel works tile by tile, so a workgroup knows its system from the tile and a finished
// system's tiles return at once. Workgroups take the tiles in XCD bands (xcd_band_tile): the tiles of one
// system's neighbouring rows, whose SpMV gathers overlap, run on one XCD and share its L2 (the FP64 SpMV
// of the 256-mesh batch fetches 1.44 instead of 1.76 GB, in the same time). Per-system dot products: one
// partial per tile, added by one workgroup per system over the system's contiguous tile range, in tile
// order (deterministic). Per-system time to
// solution is taken from the device wall clock when the system stops.
//
// PFEM_PRECISION_MIXED: PCG with reliable updates (residual replacement) per system: FP32 vectors and the
// FP32 copy of A in the iterations, the solution accumulated in FP64; after a tenfold drop of a system's
// recursive residual (by default), or when it passes the stopping test, the system spends one slot on
// x += x_s (FP64), r = b - A x with the FP64 matrix, the stopping test on that true residual, and
// continues its search direction from it. The Krylov recurrence is never restarted.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "solver.hpp"
#include "spmv.hpp"

namespace pfem {

struct BSys {
    double gam[2], alpha, beta, thr, res;
    double rz, rzmax;     // reliable updates: gamma now and at the last replacement
    int it, done, converged, nan;
    int mode, nrep;       // 0 iterating, 1 replacement due, 2 direction update after one; replacements
    unsigned long long t0, tdone;
};


// Systems are laid out by position, largest first (psys[q] = system at position q), so that the systems
// still iterating late in a batch form a short prefix and the grids can shrink to it. Tile t: rows
// [tiles[2t], tiles[2t+1]) of system tsys[t]; the tiles of position q are [stile[q], stile[q+1]).
struct BatchLayout {
    const Pattern* pat = nullptr;
    std::vector<int> rows, order, h_stile;
    std::vector<int> tiling;   // the solver's tile parameters the layout was cut with
    int nsys = 0, ntiles = 0;
    DevArray<int> tiles, tsys, stile, psys;
};

struct BatchWork {
    BatchLayout L;
    int n = 0;
    DevArray<double> r, p, q;           // FP64 PCG vectors
    DevArray<float> x32, r32, p32, q32; // reliable updates: x_s, r_s, p_s, q_s
    DevArray<double> part;              // 2 values x tiles
    DevArray<BSys> sys;
    BSys* h = nullptr;                  // pinned copy of the per-system states
    ~BatchWork()
    {
        if (h) (void)hipHostFree(h);
    }
};

// ---- kernels (one workgroup per tile unless noted)
// r = b - A x (or b), p = M^-1 r; partials (r, M^-1 r), (b, M^-1 b)
template <typename T, typename VT, bool JAC>
__global__ __launch_bounds__(kBlock) void k_b_init(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                   const int* __restrict__ col, const VT* __restrict__ val,
                                                   const T* __restrict__ x, const T* __restrict__ b,
                                                   const T* __restrict__ invd, T* __restrict__ r, T* __restrict__ p,
                                                   double* __restrict__ part, int x_zero)
{
    __shared__ T prod[kTileNnz];
    const int t = xcd_band_tile();
    int row, r1;
    T ax = 0;
    if (!x_zero) ax = tile_row_product_at<VT, T>(t, tiles, rowptr, col, val, x, prod, row, r1);
    else {
        row = tiles[2 * t] + (int)threadIdx.x;
        r1 = tiles[2 * t + 1];
    }
    double v[2] = {0.0, 0.0};
    if (row < r1) {
        const T bb = b[row];
        const T rr = bb - ax;
        const T zz = JAC ? rr * invd[row] : rr;
        const T mb = JAC ? bb * invd[row] : bb;
        r[row] = rr;
        p[row] = zz;
        v[0] = (double)(rr * zz);
        v[1] = (double)(bb * mb);
    }
    block_partials<2>(v, part, t);
}

// Sum of the tile partials of the system at position blockIdx.x (series i at part + i * ntiles), in tile
// order; valid in every thread.
template <int NV>
__device__ __forceinline__ void system_sum(const double* __restrict__ part, int ntiles, const int* __restrict__ stile,
                                           double (&tot)[NV])
{
    __shared__ double res[NV];
    const int t0 = stile[blockIdx.x], t1 = stile[blockIdx.x + 1];
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        v[i] = 0.0;
        for (int t = t0 + (int)threadIdx.x; t < t1; t += kBlock) v[i] += part[(size_t)i * ntiles + t];
    }
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) res[i] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] = res[i];
}

// gamma_0, threshold and start time per system; one workgroup per system
__global__ void k_b_begin(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                          const int* __restrict__ stile, const int* __restrict__ psys, double tol, int tol_mode)
{
    double tot[2];
    system_sum<2>(part, ntiles, stile, tot);
    if (threadIdx.x != 0) return;
    BSys& S = sys[psys[blockIdx.x]];
    const double rz = tot[0], bmb = tot[1];
    S.gam[0] = rz;
    S.thr = tol_mode == PFEM_TOL_RELATIVE && bmb > 0.0 ? tol * tol * bmb : tol * tol;
    S.it = 0;
    S.nan = rz != rz;
    S.converged = rz < S.thr ? 1 : 0;
    S.done = S.converged | S.nan;
    S.res = sqrt(fabs(rz));
    S.t0 = wall_clock64();
    S.tdone = S.t0;
}

// q = A p; partial (p, q)
template <typename T, typename VT>
__global__ __launch_bounds__(kBlock) void k_b_spmv(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                   const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const VT* __restrict__ val, const T* __restrict__ p, T* __restrict__ q,
                                                   const BSys* __restrict__ sys, double* __restrict__ part)
{
    __shared__ T prod[kTileNnz];
    const int t = xcd_band_tile();
    if (sys[tsys[t]].done) return;
    int row, r1;
    const T ap = tile_row_product_at<VT, T>(t, tiles, rowptr, col, val, p, prod, row, r1);
    double v[1] = {0.0};
    if (row < r1) {
        q[row] = ap;
        v[0] = (double)(p[row] * ap);
    }
    block_partials<1>(v, part, t);
}

// alpha per system; one workgroup per system
__global__ void k_b_alpha(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                          const int* __restrict__ stile, const int* __restrict__ psys, int k)
{
    BSys& S = sys[psys[blockIdx.x]];
    if (S.done) return;
    double tot[1];
    system_sum<1>(part, ntiles, stile, tot);
    if (threadIdx.x == 0) S.alpha = S.gam[k & 1] / tot[0];
}

// x += alpha p, r -= alpha q (updateXR, linear_solver.cu:68-77); partial (r, M^-1 r)
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_b_update(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                     T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                                                     const T* __restrict__ q, const T* __restrict__ invd,
                                                     const BSys* __restrict__ sys, double* __restrict__ part)
{
    const int t = xcd_band_tile();
    const BSys& S = sys[tsys[t]];
    if (S.done) return;
    const T alpha = (T)S.alpha;
    const int row = tiles[2 * t] + (int)threadIdx.x, r1 = tiles[2 * t + 1];
    double v[1] = {0.0};
    if (row < r1) {
        x[row] += alpha * p[row];
        const T rr = r[row] - alpha * q[row];
        r[row] = rr;
        const T zz = JAC ? rr * invd[row] : rr;
        v[0] = (double)(rr * zz);
    }
    block_partials<1>(v, part, t);
}

// gamma_{k+1}, beta, the stopping test and the time of convergence per system; one workgroup per system
__global__ void k_b_beta(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                         const int* __restrict__ stile, const int* __restrict__ psys, int k, int maxit)
{
    BSys& S = sys[psys[blockIdx.x]];
    if (S.done) return;
    double tot[1];
    system_sum<1>(part, ntiles, stile, tot);
    if (threadIdx.x != 0) return;
    const double gnew = tot[0];
    S.beta = gnew / S.gam[k & 1];
    S.gam[(k + 1) & 1] = gnew;
    S.it = k + 1;
    S.res = sqrt(fabs(gnew));
    if (gnew != gnew) {
        S.nan = 1;
        S.done = 1;
    } else if (gnew < S.thr) {
        S.converged = 1;
        S.done = 1;
    } else if (k + 1 >= maxit) {
        S.done = 1;
    }
    if (S.done) S.tdone = wall_clock64();
}

// p = M^-1 r + beta p (updateP, linear_solver.cu:79-87)
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_b_p(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                T* __restrict__ p, const T* __restrict__ r, const T* __restrict__ invd,
                                                const BSys* __restrict__ sys)
{
    const int t = xcd_band_tile();
    const BSys& S = sys[tsys[t]];
    if (S.done) return;
    const T beta = (T)S.beta;
    const int row = tiles[2 * t] + (int)threadIdx.x, r1 = tiles[2 * t + 1];
    if (row < r1) {
        const T zz = JAC ? r[row] * invd[row] : r[row];
        p[row] = zz + beta * p[row];
    }
}

// ---- reliable-update PCG (PFEM_PRECISION_MIXED)
// r = b - A x (FP64), r_s = (float) r, p_s = D_s r_s, x_s = 0; partials (r, D r), (b, D b) in FP64
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_br_init(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const double* __restrict__ val,
                                                    const double* __restrict__ x, const double* __restrict__ b,
                                                    const double* __restrict__ invd, const float* __restrict__ invd32,
                                                    float* __restrict__ rs, float* __restrict__ ps, float* __restrict__ xs,
                                                    double* __restrict__ part)
{
    __shared__ double prod[kTileNnz];
    const int t = xcd_band_tile();
    int row, r1;
    const double ax = tile_row_product_at<double, double>(t, tiles, rowptr, col, val, x, prod, row, r1);
    double v[2] = {0.0, 0.0};
    if (row < r1) {
        const double bb = b[row];
        const double rr = bb - ax;
        const float r32 = (float)rr;
        rs[row] = r32;
        ps[row] = JAC ? r32 * invd32[row] : r32;
        xs[row] = 0.0f;
        const double d = JAC ? invd[row] : 1.0;
        v[0] = rr * (rr * d);
        v[1] = bb * (bb * d);
    }
    block_partials<2>(v, part, t);
}

__global__ void k_br_begin(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                           const int* __restrict__ stile, const int* __restrict__ psys, double tol, int tol_mode)
{
    double tot[2];
    system_sum<2>(part, ntiles, stile, tot);
    if (threadIdx.x != 0) return;
    BSys& S = sys[psys[blockIdx.x]];
    const double rz = tot[0], bmb = tot[1];
    S.thr = tol_mode == PFEM_TOL_RELATIVE && bmb > 0.0 ? tol * tol * bmb : tol * tol;
    S.rz = S.rzmax = rz;
    S.res = sqrt(fabs(rz));
    S.it = 0;
    S.mode = 0;
    S.nrep = 0;
    S.nan = rz != rz;
    S.converged = rz < S.thr ? 1 : 0;
    S.done = S.converged | S.nan;
    S.t0 = wall_clock64();
    S.tdone = S.t0;
}

// tile of an iterating system: q_s = A_s p_s, partial (p_s, q_s); of a system due a replacement:
// r = b - A x with the FP64 matrix, r_s = (float) r, partial (r, D r)
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_br_spmv(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                    const int* __restrict__ rowptr, const int* __restrict__ col,
                                                    const float* __restrict__ val32, const double* __restrict__ val,
                                                    const float* __restrict__ ps, float* __restrict__ qs,
                                                    const double* __restrict__ x, const double* __restrict__ b,
                                                    const double* __restrict__ invd, float* __restrict__ rs,
                                                    const BSys* __restrict__ sys, double* __restrict__ part)
{
    __shared__ double prod[kTileNnz];
    const int t = xcd_band_tile();
    const BSys& S = sys[tsys[t]];
    if (S.done) return;
    int row, r1;
    double v[1] = {0.0};
    if (S.mode == 1) {
        const double ax = tile_row_product_at<double, double>(t, tiles, rowptr, col, val, x, prod, row, r1);
        if (row < r1) {
            const double rr = b[row] - ax;
            rs[row] = (float)rr;
            v[0] = rr * (rr * (JAC ? invd[row] : 1.0));
        }
    } else {
        const float ap = tile_row_product_at<float, float>(t, tiles, rowptr, col, val32, ps,
                                                           reinterpret_cast<float*>(prod), row, r1);
        if (row < r1) {
            qs[row] = ap;
            v[0] = (double)(ps[row] * ap);
        }
    }
    block_partials<1>(v, part, t);
}

// alpha of an iterating system; for a replaced one the true gamma: the stopping test, then beta
__global__ void k_br_alpha(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                           const int* __restrict__ stile, const int* __restrict__ psys)
{
    BSys& S = sys[psys[blockIdx.x]];
    if (S.done) return;
    double tot[1];
    system_sum<1>(part, ntiles, stile, tot);
    if (threadIdx.x != 0) return;
    const double g = tot[0];
    if (S.mode == 1) {
        S.nrep += 1;
        S.res = sqrt(fabs(g));
        if (g != g) {
            S.nan = 1;
            S.done = 1;
        } else if (g < S.thr) {
            S.converged = 1;
            S.done = 1;
        } else {
            S.beta = g / S.rz;
            S.rz = S.rzmax = g;
            S.mode = 2;
        }
        if (S.done) S.tdone = wall_clock64();
    } else {
        S.mode = 0;
        S.alpha = S.rz / g;
    }
}

// x_s += alpha p_s, r_s -= alpha q_s (FP32); partial (r_s, D_s r_s)
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_br_update(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                      float* __restrict__ xs, float* __restrict__ rs,
                                                      const float* __restrict__ ps, const float* __restrict__ qs,
                                                      const float* __restrict__ invd32, const BSys* __restrict__ sys,
                                                      double* __restrict__ part)
{
    const int t = xcd_band_tile();
    const BSys& S = sys[tsys[t]];
    if (S.done || S.mode != 0) return;
    const float alpha = (float)S.alpha;
    const int row = tiles[2 * t] + (int)threadIdx.x, r1 = tiles[2 * t + 1];
    double v[1] = {0.0};
    if (row < r1) {
        xs[row] += alpha * ps[row];
        const float rr = rs[row] - alpha * qs[row];
        rs[row] = rr;
        v[0] = (double)(rr * (JAC ? rr * invd32[row] : rr));
    }
    block_partials<1>(v, part, t);
}

// gamma of the recursive residual: the iteration count, then beta or a replacement
__global__ void k_br_beta(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                          const int* __restrict__ stile, const int* __restrict__ psys, int maxit, double drop)
{
    BSys& S = sys[psys[blockIdx.x]];
    if (S.done || S.mode != 0) return;
    double tot[1];
    system_sum<1>(part, ntiles, stile, tot);
    if (threadIdx.x != 0) return;
    const double g = tot[0];
    S.it += 1;
    if (g != g) {
        S.nan = 1;
        S.done = 1;
    } else if (S.it >= maxit) {
        S.done = 1;
    } else if (g < S.thr || g < drop * S.rzmax) {
        S.mode = 1;
    } else {
        S.beta = g / S.rz;
        S.rz = g;
    }
    if (S.done) S.tdone = wall_clock64();
}

// p_s = D_s r_s + beta p_s; a system due a replacement first folds x_s into x (FP64)
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_br_p(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                 float* __restrict__ ps, const float* __restrict__ rs,
                                                 const float* __restrict__ invd32, double* __restrict__ x,
                                                 float* __restrict__ xs, const BSys* __restrict__ sys)
{
    const int t = xcd_band_tile();
    const BSys& S = sys[tsys[t]];
    if (S.done) return;
    const int row = tiles[2 * t] + (int)threadIdx.x, r1 = tiles[2 * t + 1];
    if (row >= r1) return;
    if (S.mode == 1) {
        x[row] += (double)xs[row];
        xs[row] = 0.0f;
    } else {
        const float zz = JAC ? rs[row] * invd32[row] : rs[row];
        ps[row] = zz + (float)S.beta * ps[row];
    }
}

// x += x_s everywhere (zero for a system whose last replacement folded it in)
__global__ void k_br_fold(long n, double* __restrict__ x, const float* __restrict__ xs)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)xs[i];
}

__global__ void k_b_to32(long n, float* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

// ---- host side
// Tile limits of a system of nr rows: the first bucket whose row bound holds it (tiling = {bound, rows,
// non-zeros} per bucket); by default (no buckets) 256 rows and kTileNnz non-zeros.
static void tile_limits(const std::vector<int>& tiling, int nr, int& max_rows, int& max_nnz)
{
    max_rows = kBlock;
    max_nnz = kTileNnz;
    for (size_t k = 0; k + 2 < tiling.size(); k += 3)
        if (nr <= tiling[k]) {
            max_rows = tiling[k + 1];
            max_nnz = tiling[k + 2];
            return;
        }
}

static void build_layout(Context* c, const Pattern& P, const std::vector<int>& rows, const std::vector<int>& tiling,
                         BatchLayout& L)
{
    const int nsys = (int)rows.size() - 1;
    PFEM_REQUIRE(nsys >= 1 && rows[0] == 0 && rows[nsys] == P.n, "system row ranges must cover the matrix");
    for (int s = 0; s < nsys; ++s) {
        PFEM_REQUIRE(rows[s] < rows[s + 1], "empty system");
        for (int i = rows[s]; i < rows[s + 1]; ++i)
            for (int k = P.h_rowptr[i]; k < P.h_rowptr[i + 1]; ++k)
                PFEM_REQUIRE(P.h_col[k] >= rows[s] && P.h_col[k] < rows[s + 1],
                             "the matrix is not block diagonal over the given systems");
    }
    std::vector<int> order(nsys);
    for (int q = 0; q < nsys; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(),
                     [&](int u, int v) { return rows[u + 1] - rows[u] > rows[v + 1] - rows[v]; });
    std::vector<int> tiles, tsys, stile(nsys + 1, 0);
    for (int q = 0; q < nsys; ++q) {
        const int s = order[q];
        stile[q] = (int)tsys.size();
        int max_rows, max_nnz;
        tile_limits(tiling, rows[s + 1] - rows[s], max_rows, max_nnz);
        int r = rows[s];
        while (r < rows[s + 1]) {
            int r1 = r;
            const int base = P.h_rowptr[r];
            while (r1 < rows[s + 1] && r1 - r < max_rows && (P.h_rowptr[r1 + 1] - base <= max_nnz || r1 == r)) ++r1;
            tiles.push_back(r);
            tiles.push_back(r1);
            tsys.push_back(s);
            r = r1;
        }
    }
    stile[nsys] = (int)tsys.size();
    L.pat = &P;
    L.rows = rows;
    L.tiling = tiling;
    L.order = order;
    L.h_stile = stile;
    L.nsys = nsys;
    L.ntiles = (int)tsys.size();
    upload(c, L.tiles, tiles);
    upload(c, L.tsys, tsys);
    upload(c, L.stile, stile);
    upload(c, L.psys, order);
}

// Batched FP64 PCG over all systems (x: initial guess and solution)
template <typename T, typename VT>
static void b_cg_run(Context* c, pfem_solver_t* s, BatchWork& w, const VT* val, const T* invd, T* x, const T* b, T* r,
                     T* p, T* q, int x_zero, double tol, int tol_mode, int maxit)
{
    const BatchLayout& L = w.L;
    const Pattern& P = *L.pat;
    const bool jac = invd != nullptr;
    dim3 gt(L.ntiles), gs(L.nsys);
    const dim3 bl(kBlock);
    double* pa = w.part.p;
    if (x_zero) PFEM_HIP(hipMemsetAsync(x, 0, sizeof(T) * (size_t)w.n, c->stream));
    if (jac) hipLaunchKernelGGL((k_b_init<T, VT, true>), gt, bl, 0, c->stream, L.tiles.p, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    else hipLaunchKernelGGL((k_b_init<T, VT, false>), gt, bl, 0, c->stream, L.tiles.p, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    hipLaunchKernelGGL(k_b_begin, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, L.psys.p, tol,
                       tol_mode);
    PFEM_HIP(hipGetLastError());
    const int batch = s->poll;
    int launched = 0;
    while (true) {
        for (int j = 0; j < batch; ++j) {
            const int k = launched + j;
            hipLaunchKernelGGL((k_b_spmv<T, VT>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, P.rowptr.p, P.col.p, val, p, q, w.sys.p, pa);
            hipLaunchKernelGGL(k_b_alpha, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, L.psys.p, k);
            if (jac) hipLaunchKernelGGL((k_b_update<T, true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, x, r, p, q, invd, w.sys.p, pa);
            else hipLaunchKernelGGL((k_b_update<T, false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, x, r, p, q, invd, w.sys.p, pa);
            hipLaunchKernelGGL(k_b_beta, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, L.psys.p, k,
                               maxit);
            if (jac) hipLaunchKernelGGL((k_b_p<T, true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, p, r, invd, w.sys.p);
            else hipLaunchKernelGGL((k_b_p<T, false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, p, r, invd, w.sys.p);
        }
        launched += batch;
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(w.h, w.sys.p, sizeof(BSys) * L.nsys, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        // shrink the grids to the prefix of positions that still iterate
        int last = -1;
        for (int q2 = 0; q2 < L.nsys; ++q2)
            if (!w.h[L.order[q2]].done) last = q2;
        if (last < 0 || launched > maxit + batch) break;
        gt = dim3(L.h_stile[last + 1]);
        gs = dim3(last + 1);
    }
}

// Reliable-update PCG over all systems (x: FP64 solution, b: FP64 right-hand side)
static void b_ru_run(Context* c, pfem_solver_t* s, BatchWork& w, pfem_csr_t* a, double* x, const double* b, bool jac)
{
    const BatchLayout& L = w.L;
    const Pattern& P = *L.pat;
    const int n = w.n;
    if (!w.x32.p) {
        w.x32.alloc(n); w.r32.alloc(n); w.p32.alloc(n); w.q32.alloc(n);
    }
    if (!a->val32.p) a->val32.alloc(P.nnz);
    hipLaunchKernelGGL(k_b_to32, dim3(blocks_for(P.nnz)), dim3(kBlock), 0, c->stream, (long)P.nnz, a->val32.p, a->val.p);
    const double* invd = jac ? s->pc->invd.p : nullptr;
    const float* invd32 = jac ? s->pc->invd32.p : nullptr;
    PFEM_REQUIRE(!jac || invd32, "FP32 Jacobi diagonal missing");
    float *xs = w.x32.p, *rs = w.r32.p, *ps = w.p32.p, *qs = w.q32.p;
    const float* val32 = a->val32.p;
    const double* val = a->val.p;
    double* pa = w.part.p;
    dim3 gt(L.ntiles), gs(L.nsys);
    const dim3 bl(kBlock);
    if (jac) hipLaunchKernelGGL((k_br_init<true>), gt, bl, 0, c->stream, L.tiles.p, P.rowptr.p, P.col.p, val, (const double*)x, b, invd, invd32, rs, ps, xs, pa);
    else hipLaunchKernelGGL((k_br_init<false>), gt, bl, 0, c->stream, L.tiles.p, P.rowptr.p, P.col.p, val, (const double*)x, b, invd, invd32, rs, ps, xs, pa);
    hipLaunchKernelGGL(k_br_begin, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, L.psys.p, s->tol, s->tol_mode);
    PFEM_HIP(hipGetLastError());
    const int batch = s->poll;
    const int slots = s->maxit + s->maxit / 4 + 64;
    const double drop = std::pow(10.0, -s->replace_digits);
    int launched = 0;
    while (true) {
        for (int j = 0; j < batch; ++j) {
            if (jac) hipLaunchKernelGGL((k_br_spmv<true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, P.rowptr.p, P.col.p, val32, val, (const float*)ps, qs, (const double*)x, b, invd, rs, (const BSys*)w.sys.p, pa);
            else hipLaunchKernelGGL((k_br_spmv<false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, P.rowptr.p, P.col.p, val32, val, (const float*)ps, qs, (const double*)x, b, invd, rs, (const BSys*)w.sys.p, pa);
            hipLaunchKernelGGL(k_br_alpha, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, L.psys.p);
            if (jac) hipLaunchKernelGGL((k_br_update<true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, xs, rs, (const float*)ps, (const float*)qs, invd32, (const BSys*)w.sys.p, pa);
            else hipLaunchKernelGGL((k_br_update<false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, xs, rs, (const float*)ps, (const float*)qs, invd32, (const BSys*)w.sys.p, pa);
            hipLaunchKernelGGL(k_br_beta, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, L.psys.p, s->maxit,
                               drop);
            if (jac) hipLaunchKernelGGL((k_br_p<true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, ps, (const float*)rs, invd32, x, xs, (const BSys*)w.sys.p);
            else hipLaunchKernelGGL((k_br_p<false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, ps, (const float*)rs, invd32, x, xs, (const BSys*)w.sys.p);
        }
        launched += batch;
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(w.h, w.sys.p, sizeof(BSys) * L.nsys, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        int last = -1;
        for (int q2 = 0; q2 < L.nsys; ++q2)
            if (!w.h[L.order[q2]].done) last = q2;
        if (last < 0 || launched > slots + batch) break;
        gt = dim3(L.h_stile[last + 1]);
        gs = dim3(last + 1);
    }
    // a system stopped by the iteration limit still holds its correction in x_s
    hipLaunchKernelGGL(k_br_fold, dim3(blocks_for(n)), bl, 0, c->stream, (long)n, x, (const float*)xs);
}

static void solve_batched(Context* c, pfem_solver_t* s, pfem_csr_t* a, const std::vector<int>& rows, double* x,
                          const double* b, pfem_solve_stats* stats, double* time_ms)
{
    const Pattern& P = *a->pat;
    const int n = P.n;
    if (!s->batch || s->batch->L.pat != &P || s->batch->L.rows != rows || s->batch->L.tiling != s->batch_tiling) {
        s->batch.reset(new BatchWork());
        BatchWork& w = *s->batch;
        build_layout(c, P, rows, s->batch_tiling, w.L);
        w.n = n;
        w.r.alloc(n); w.p.alloc(n); w.q.alloc(n);
        w.part.alloc(2 * (size_t)w.L.ntiles);
        w.sys.alloc(w.L.nsys);
        PFEM_HIP(hipHostMalloc((void**)&w.h, sizeof(BSys) * w.L.nsys, hipHostMallocDefault));
    }
    BatchWork& w = *s->batch;
    const BatchLayout& L = w.L;
    const int nsys = L.nsys;
    const bool jac = s->pc && s->pc->kind == PFEM_PRECOND_JACOBI;
    if (s->pc) precond_initialize(c, s->pc, a);
    PFEM_HIP(hipMemsetAsync(w.sys.p, 0, sizeof(BSys) * nsys, c->stream));
    if (s->precision != PFEM_PRECISION_MIXED) {
        b_cg_run<double, double>(c, s, w, a->val.p, jac ? s->pc->invd.p : nullptr, x, b, w.r.p, w.p.p, w.q.p, 0, s->tol,
                                 s->tol_mode, s->maxit);
    } else {
        b_ru_run(c, s, w, a, x, b, jac);
    }
    PFEM_HIP(hipGetLastError());
    PFEM_HIP(hipMemcpyAsync(w.h, w.sys.p, sizeof(BSys) * nsys, hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
    int rate_khz = 0;
    PFEM_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device));
    const bool mixed = s->precision == PFEM_PRECISION_MIXED;
    for (int q = 0; q < nsys; ++q) {
        const BSys& S = w.h[q];
        if (stats) {
            pfem_solve_stats& o = stats[q];
            o = pfem_solve_stats{};
            o.converged = S.converged;
            o.iterations = S.it;
            o.final_residual = S.res;
            o.nan_detected = S.nan;
            o.refinements = mixed ? S.nrep : 0;
        }
        if (time_ms) time_ms[q] = rate_khz > 0 ? (double)(S.tdone - S.t0) / rate_khz : 0.0;
    }
}

void BatchWorkDeleter::operator()(BatchWork* w) const { delete w; }

}  // namespace pfem

using namespace pfem;

extern "C" {

int pfem_solver_set_batch_tiles(pfem_solver s, int nbuckets, const int* max_rows, const int* tile_rows,
                                const int* tile_nnz)
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s && nbuckets >= 0 && nbuckets <= 16, "bad arguments");
        PFEM_REQUIRE(nbuckets == 0 || (max_rows && tile_rows && tile_nnz), "null argument");
        std::vector<int> t;
        for (int k = 0; k < nbuckets; ++k) {
            PFEM_REQUIRE(k == 0 || max_rows[k] > max_rows[k - 1], "bucket row bounds must ascend");
            PFEM_REQUIRE(tile_rows[k] >= 1 && tile_rows[k] <= kBlock, "tile rows must lie in [1, 256]");
            PFEM_REQUIRE(tile_nnz[k] >= 1 && tile_nnz[k] <= kTileNnz, "tile non-zeros must lie in [1, 2048]");
            t.insert(t.end(), {max_rows[k], tile_rows[k], tile_nnz[k]});
        }
        s->batch_tiling = t;
    });
}

int pfem_solver_solve_batched(pfem_solver s, pfem_csr a, int nsys, const int* h_sys_rows, double* d_x,
                              const double* d_b, pfem_solve_stats* stats, double* h_time_ms)
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s && a && h_sys_rows && d_x && d_b, "null argument");
        PFEM_REQUIRE(nsys >= 1, "no systems");
        PFEM_REQUIRE(a->pat->n == s->n, "matrix size differs from the solver's");
        PFEM_REQUIRE(s->kind == PFEM_SOLVER_CG, "batched solves are PCG only");
        PFEM_REQUIRE(!s->pc || s->pc->kind == PFEM_PRECOND_JACOBI || s->pc->kind == PFEM_PRECOND_NONE,
                     "batched PCG takes a Jacobi preconditioner or none");
        PFEM_REQUIRE(s->precision == PFEM_PRECISION_FP64 || s->precision == PFEM_PRECISION_MIXED,
                     "batched PCG runs in FP64 or MIXED precision");
        std::vector<int> rows(h_sys_rows, h_sys_rows + nsys + 1);
        solve_batched(C(s->ctx), s, a, rows, d_x, d_b, stats, h_time_ms);
    });
}

}  // extern "C"
