// Batched independent PCG solves: many systems of uneven size, stored together as one block-diagonal CSR
// matrix (system s owns rows and columns [rows[s], rows[s+1])), solved by one sequence of launches. The
// reference solves one system per SolverCG::solve call (src/linear_algebra/linear_solver.cu:301-376);
// each system here runs that recurrence (Jacobi or no preconditioner, the stopping test on (r, M^-1 r))
// with its own scalars and stops on its own.
//
// Work is cut into SpMV row tiles (<= 256 rows, <= 2048 non-zeros) that never straddle two systems, and
// every vector kernel works tile by tile, so a workgroup knows its system from the tile and a finished
// system's tiles return at once. Per-system dot products: one partial per tile, added by one workgroup
// per system over the system's contiguous tile range, in tile order (deterministic). Per-system time to
// solution is taken from the device wall clock when the system's stopping test passes.
//
// PFEM_PRECISION_MIXED: FP32 inner batched CG on the FP32 copy of A, FP64 iterative refinement per system.
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
    double gam[2], alpha, beta, thr, res, irthr, irrz, scale;
    int it, done, converged, nan, irconv, irnan, iractive, pad;
    unsigned long long t0, tdone;
};

// Tile t: rows [tiles[2t], tiles[2t+1]) of system tsys[t]; a system's tiles are contiguous,
// [stile[s], stile[s+1]).
struct BatchLayout {
    const Pattern* pat = nullptr;
    std::vector<int> rows;
    int nsys = 0, ntiles = 0;
    DevArray<int> tiles, tsys, stile;
};

struct BatchWork {
    BatchLayout L;
    int n = 0;
    DevArray<double> r, p, q, xr;       // FP64 vectors (xr: refinement residual)
    DevArray<float> x32, b32, r32, p32, q32;
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
    int row, r1;
    T ax = 0;
    if (!x_zero) ax = tile_row_product<VT, T>(tiles, rowptr, col, val, x, prod, row, r1);
    else {
        row = tiles[2 * blockIdx.x] + (int)threadIdx.x;
        r1 = tiles[2 * blockIdx.x + 1];
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
    block_partials<2>(v, part);
}

// Sum of the tile partials of system blockIdx.x (series i at part + i * ntiles), in tile order; valid in
// every thread.
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
                          const int* __restrict__ stile, double tol, int tol_mode, int outer)
{
    double tot[2];
    system_sum<2>(part, ntiles, stile, tot);
    if (threadIdx.x != 0) return;
    BSys& S = sys[blockIdx.x];
    const double rz = tot[0], bmb = tot[1];
    S.gam[0] = rz;
    S.thr = tol_mode == PFEM_TOL_RELATIVE && bmb > 0.0 ? tol * tol * bmb : tol * tol;
    S.it = 0;
    S.nan = rz != rz;
    S.converged = rz < S.thr ? 1 : 0;
    S.done = S.converged | S.nan;
    S.res = sqrt(fabs(rz));
    if (outer) {
        // an outer (FP64) solve times itself; an inner pass leaves the refinement's clock alone
        S.t0 = wall_clock64();
        S.tdone = S.t0;
    }
}

// q = A p; partial (p, q)
template <typename T, typename VT>
__global__ __launch_bounds__(kBlock) void k_b_spmv(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                   const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const VT* __restrict__ val, const T* __restrict__ p, T* __restrict__ q,
                                                   const BSys* __restrict__ sys, double* __restrict__ part)
{
    __shared__ T prod[kTileNnz];
    if (sys[tsys[blockIdx.x]].done) return;
    int row, r1;
    const T ap = tile_row_product<VT, T>(tiles, rowptr, col, val, p, prod, row, r1);
    double v[1] = {0.0};
    if (row < r1) {
        q[row] = ap;
        v[0] = (double)(p[row] * ap);
    }
    block_partials<1>(v, part);
}

// alpha per system; one workgroup per system
__global__ void k_b_alpha(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                          const int* __restrict__ stile, int k)
{
    BSys& S = sys[blockIdx.x];
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
    const BSys& S = sys[tsys[blockIdx.x]];
    if (S.done) return;
    const T alpha = (T)S.alpha;
    const int row = tiles[2 * blockIdx.x] + (int)threadIdx.x, r1 = tiles[2 * blockIdx.x + 1];
    double v[1] = {0.0};
    if (row < r1) {
        x[row] += alpha * p[row];
        const T rr = r[row] - alpha * q[row];
        r[row] = rr;
        const T zz = JAC ? rr * invd[row] : rr;
        v[0] = (double)(rr * zz);
    }
    block_partials<1>(v, part);
}

// gamma_{k+1}, beta, the stopping test and the time of convergence per system; one workgroup per system
__global__ void k_b_beta(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                         const int* __restrict__ stile, int k, int maxit, int outer)
{
    BSys& S = sys[blockIdx.x];
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
    if (S.done && outer) S.tdone = wall_clock64();
}

// p = M^-1 r + beta p (updateP, linear_solver.cu:79-87)
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_b_p(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                T* __restrict__ p, const T* __restrict__ r, const T* __restrict__ invd,
                                                const BSys* __restrict__ sys)
{
    const BSys& S = sys[tsys[blockIdx.x]];
    if (S.done) return;
    const T beta = (T)S.beta;
    const int row = tiles[2 * blockIdx.x] + (int)threadIdx.x, r1 = tiles[2 * blockIdx.x + 1];
    if (row < r1) {
        const T zz = JAC ? r[row] * invd[row] : r[row];
        p[row] = zz + beta * p[row];
    }
}

// ---- refinement (FP64 outer loop per system)
// r = b - A x in FP64 for the systems still refining; partials (r, M^-1 r), (b, M^-1 b)
template <bool JAC>
__global__ __launch_bounds__(kBlock) void k_b_ir_residual(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                          const int* __restrict__ rowptr, const int* __restrict__ col,
                                                          const double* __restrict__ val, const double* __restrict__ x,
                                                          const double* __restrict__ b, const double* __restrict__ invd,
                                                          double* __restrict__ r, const BSys* __restrict__ sys,
                                                          double* __restrict__ part)
{
    __shared__ double prod[kTileNnz];
    if (!sys[tsys[blockIdx.x]].iractive) return;
    int row, r1;
    const double ax = tile_row_product<double, double>(tiles, rowptr, col, val, x, prod, row, r1);
    double v[2] = {0.0, 0.0};
    if (row < r1) {
        const double bb = b[row];
        const double rr = bb - ax;
        r[row] = rr;
        const double d = JAC ? invd[row] : 1.0;
        v[0] = rr * (rr * d);
        v[1] = bb * (bb * d);
    }
    block_partials<2>(v, part);
}

// per-system refinement test (the FP64 solver's criterion) and the scale of the next correction; the
// first pass sets the thresholds and the start times
__global__ void k_b_ir_test(BSys* __restrict__ sys, const double* __restrict__ part, int ntiles,
                            const int* __restrict__ stile, double tol, int tol_mode, int first)
{
    BSys& S = sys[blockIdx.x];
    if (!first && !S.iractive) return;
    double tot[2];
    system_sum<2>(part, ntiles, stile, tot);
    if (threadIdx.x != 0) return;
    if (first) {
        const bool rel = tol_mode == PFEM_TOL_RELATIVE && tot[1] > 0.0;
        S.irthr = rel ? tol * tol * tot[1] : tol * tol;
        S.t0 = wall_clock64();
    }
    S.irrz = tot[0];
    S.irnan = tot[0] != tot[0];
    S.irconv = tot[0] < S.irthr ? 1 : 0;
    S.iractive = !(S.irconv || S.irnan);
    S.scale = S.iractive ? (sqrt(tot[0]) > 0.0 ? sqrt(tot[0]) : 1.0) : 0.0;
    if (!S.iractive) S.tdone = wall_clock64();
}

__global__ __launch_bounds__(kBlock) void k_b_ir_down(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                      float* __restrict__ d, const double* __restrict__ r,
                                                      const BSys* __restrict__ sys)
{
    const double sc = sys[tsys[blockIdx.x]].scale;
    const int row = tiles[2 * blockIdx.x] + (int)threadIdx.x, r1 = tiles[2 * blockIdx.x + 1];
    if (row < r1) d[row] = sc > 0.0 ? (float)(r[row] / sc) : 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_b_ir_up(const int* __restrict__ tiles, const int* __restrict__ tsys,
                                                    double* __restrict__ x, const float* __restrict__ d,
                                                    const BSys* __restrict__ sys)
{
    const double sc = sys[tsys[blockIdx.x]].scale;
    const int row = tiles[2 * blockIdx.x] + (int)threadIdx.x, r1 = tiles[2 * blockIdx.x + 1];
    if (row < r1 && sc > 0.0) x[row] += sc * (double)d[row];
}

__global__ void k_b_to32(long n, float* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

__global__ void k_b_arm(int nsys, BSys* __restrict__ sys)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) sys[s].iractive = 1;
}

// ---- host side
static void build_layout(Context* c, const Pattern& P, const std::vector<int>& rows, BatchLayout& L)
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
    std::vector<int> tiles, tsys, stile(nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        stile[s] = (int)tsys.size();
        int r = rows[s];
        while (r < rows[s + 1]) {
            int r1 = r;
            const int base = P.h_rowptr[r];
            while (r1 < rows[s + 1] && r1 - r < kBlock && (P.h_rowptr[r1 + 1] - base <= kTileNnz || r1 == r)) ++r1;
            tiles.push_back(r);
            tiles.push_back(r1);
            tsys.push_back(s);
            r = r1;
        }
    }
    stile[nsys] = (int)tsys.size();
    L.pat = &P;
    L.rows = rows;
    L.nsys = nsys;
    L.ntiles = (int)tsys.size();
    upload(c, L.tiles, tiles);
    upload(c, L.tsys, tsys);
    upload(c, L.stile, stile);
}

// One batched CG pass (T = double: the FP64 solve; T = float: the inner solve from x = 0)
template <typename T, typename VT>
static void b_cg_run(Context* c, pfem_solver_t* s, BatchWork& w, const VT* val, const T* invd, T* x, const T* b, T* r,
                     T* p, T* q, int x_zero, double tol, int tol_mode, int maxit, bool outer)
{
    const BatchLayout& L = w.L;
    const Pattern& P = *L.pat;
    const bool jac = invd != nullptr;
    const dim3 gt(L.ntiles), gs(L.nsys), bl(kBlock);
    double* pa = w.part.p;
    if (x_zero) PFEM_HIP(hipMemsetAsync(x, 0, sizeof(T) * (size_t)w.n, c->stream));
    if (jac) hipLaunchKernelGGL((k_b_init<T, VT, true>), gt, bl, 0, c->stream, L.tiles.p, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    else hipLaunchKernelGGL((k_b_init<T, VT, false>), gt, bl, 0, c->stream, L.tiles.p, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    hipLaunchKernelGGL(k_b_begin, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, tol, tol_mode,
                       outer ? 1 : 0);
    PFEM_HIP(hipGetLastError());
    const int batch = s->poll;
    int launched = 0;
    while (true) {
        for (int j = 0; j < batch; ++j) {
            const int k = launched + j;
            hipLaunchKernelGGL((k_b_spmv<T, VT>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, P.rowptr.p, P.col.p, val, p, q, w.sys.p, pa);
            hipLaunchKernelGGL(k_b_alpha, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, k);
            if (jac) hipLaunchKernelGGL((k_b_update<T, true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, x, r, p, q, invd, w.sys.p, pa);
            else hipLaunchKernelGGL((k_b_update<T, false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, x, r, p, q, invd, w.sys.p, pa);
            hipLaunchKernelGGL(k_b_beta, gs, bl, 0, c->stream, w.sys.p, (const double*)pa, L.ntiles, L.stile.p, k, maxit,
                               outer ? 1 : 0);
            if (jac) hipLaunchKernelGGL((k_b_p<T, true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, p, r, invd, w.sys.p);
            else hipLaunchKernelGGL((k_b_p<T, false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, p, r, invd, w.sys.p);
        }
        launched += batch;
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(w.h, w.sys.p, sizeof(BSys) * L.nsys, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        bool all = true;
        for (int q2 = 0; q2 < L.nsys && all; ++q2) all = w.h[q2].done != 0;
        if (all || launched > maxit + batch) break;
    }
}

static void solve_batched(Context* c, pfem_solver_t* s, pfem_csr_t* a, const std::vector<int>& rows, double* x,
                          const double* b, pfem_solve_stats* stats, double* time_ms)
{
    const Pattern& P = *a->pat;
    const int n = P.n;
    if (!s->batch || s->batch->L.pat != &P || s->batch->L.rows != rows) {
        s->batch.reset(new BatchWork());
        BatchWork& w = *s->batch;
        build_layout(c, P, rows, w.L);
        w.n = n;
        w.r.alloc(n); w.p.alloc(n); w.q.alloc(n); w.xr.alloc(n);
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
    int refinements = 0;
    std::vector<int> inner(nsys, 0);
    if (s->precision != PFEM_PRECISION_MIXED) {
        b_cg_run<double, double>(c, s, w, a->val.p, jac ? s->pc->invd.p : nullptr, x, b, w.r.p, w.p.p, w.q.p, 0, s->tol,
                                 s->tol_mode, s->maxit, true);
    } else {
        if (!w.x32.p) {
            w.x32.alloc(n); w.b32.alloc(n); w.r32.alloc(n); w.p32.alloc(n); w.q32.alloc(n);
        }
        if (!a->val32.p) a->val32.alloc(P.nnz);
        hipLaunchKernelGGL(k_b_to32, dim3(blocks_for(P.nnz)), dim3(kBlock), 0, c->stream, (long)P.nnz, a->val32.p, a->val.p);
        const double* invd = jac ? s->pc->invd.p : nullptr;
        const float* invd32 = jac ? s->pc->invd32.p : nullptr;
        PFEM_REQUIRE(!jac || invd32, "FP32 Jacobi diagonal missing");
        const dim3 gt(L.ntiles), gs(nsys), bl(kBlock);
        hipLaunchKernelGGL(k_b_arm, dim3(blocks_for(nsys)), bl, 0, c->stream, nsys, w.sys.p);
        int total = 0;
        for (bool first = true;; first = false) {
            if (jac) hipLaunchKernelGGL((k_b_ir_residual<true>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, P.rowptr.p, P.col.p, a->val.p, (const double*)x, b, invd, w.xr.p, (const BSys*)w.sys.p, w.part.p);
            else hipLaunchKernelGGL((k_b_ir_residual<false>), gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, P.rowptr.p, P.col.p, a->val.p, (const double*)x, b, invd, w.xr.p, (const BSys*)w.sys.p, w.part.p);
            hipLaunchKernelGGL(k_b_ir_test, gs, bl, 0, c->stream, w.sys.p, (const double*)w.part.p, L.ntiles, L.stile.p,
                               s->tol, s->tol_mode, first ? 1 : 0);
            PFEM_HIP(hipMemcpyAsync(w.h, w.sys.p, sizeof(BSys) * nsys, hipMemcpyDeviceToHost, c->stream));
            PFEM_HIP(hipStreamSynchronize(c->stream));
            bool any = false;
            for (int q = 0; q < nsys; ++q) any |= w.h[q].iractive != 0;
            if (!any || refinements >= s->max_refine || total >= s->maxit) break;
            hipLaunchKernelGGL(k_b_ir_down, gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, w.b32.p, (const double*)w.xr.p, (const BSys*)w.sys.p);
            b_cg_run<float, float>(c, s, w, a->val32.p, invd32, w.x32.p, w.b32.p, w.r32.p, w.p32.p, w.q32.p, 1,
                                   s->inner_tol, PFEM_TOL_RELATIVE, std::max(1, s->maxit - total), false);
            int most = 0;
            for (int q = 0; q < nsys; ++q) {
                inner[q] += w.h[q].it;
                most = std::max(most, w.h[q].it);
            }
            total += most;
            hipLaunchKernelGGL(k_b_ir_up, gt, bl, 0, c->stream, L.tiles.p, L.tsys.p, x, (const float*)w.x32.p, (const BSys*)w.sys.p);
            ++refinements;
        }
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
            o.converged = mixed ? S.irconv : S.converged;
            o.iterations = mixed ? inner[q] : S.it;
            o.final_residual = mixed ? std::sqrt(std::fabs(S.irrz)) : S.res;
            o.nan_detected = mixed ? S.irnan : S.nan;
            o.refinements = refinements;
        }
        if (time_ms) time_ms[q] = rate_khz > 0 ? (double)(S.tdone - S.t0) / rate_khz : 0.0;
    }
}

void BatchWorkDeleter::operator()(BatchWork* w) const { delete w; }

}  // namespace pfem

using namespace pfem;

extern "C" {

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
