// Krylov solvers: PCG (SolverCG::solve, src/linear_algebra/linear_solver.cu:301-376) and GMRES(m)
// (SolverGMRES::solve, :421-521), FP64 and FP32, plus FP64 iterative refinement around an FP32 inner
// solve (the mixed-precision mode).
//
// Every scalar stays on the device. Dot products are deterministic reductions accumulated in FP64: the
// producing kernel stores per-block partials, the kernel that needs the total re-sums them in a fixed
// order (no atomics, no inter-block waits). A device "done" flag replaces the
// reference's per-iteration host read-back (linear_solver.cu:360, :499): the host polls it every few
// iterations and kernels of iterations launched past convergence return at once.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"
#include "solver.hpp"
#include "spmv.hpp"
#include "sync.hpp"

namespace pfem {

inline unsigned vec_grid(int n) { return std::min<unsigned>(blocks_for(n), (unsigned)kVecGrid); }

// ------------------------------------------------------------------ PCG kernels
// Iteration k (counted from the start of the solve, passed by the host) reads gamma_k = (r_k, z_k)
// from gam[k & 1]; its last kernel stores gamma_{k+1} in the other slot, so no kernel reads a value
// that another block of the same kernel writes.

// r = b - A x, z = invD r (Jacobi) or r, p = z; partials (r,z), (b, M^-1 b)
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_cg_init(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const T* __restrict__ val,
                                                    const T* __restrict__ x, const T* __restrict__ b,
                                                    const T* __restrict__ invd, T* __restrict__ r, T* __restrict__ z,
                                                    T* __restrict__ p, double* __restrict__ part)
{
    __shared__ T prod[kTileNnz];
    int row, r1;
    const T ax = tile_row_product<T, T>(tiles, rowptr, col, val, x, prod, row, r1);
    double v[2] = {0.0, 0.0};
    if (row < r1) {
        const T bb = b[row];
        const T rr = bb - ax;
        const T zz = JAC ? rr * invd[row] : rr;
        r[row] = rr;
        z[row] = zz;
        p[row] = zz;
        v[0] = (double)(rr * zz);
        v[1] = JAC ? (double)(bb * (bb * invd[row])) : (double)(bb * bb);
    }
    block_partials<2>(v, part);
}

// x = 0 start (inner solves of the refinement): r = b, z = M^-1 b (Jacobi), p = z
template <typename T, bool JAC>
__global__ void k_cg_init_zero(int n, const T* __restrict__ b, const T* __restrict__ invd, T* __restrict__ x,
                               T* __restrict__ r, T* __restrict__ z, T* __restrict__ p, double* __restrict__ part)
{
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T bb = b[i];
        const T zz = JAC ? bb * invd[i] : bb;
        x[i] = 0;
        r[i] = bb;
        z[i] = zz;
        p[i] = zz;
        v[0] += (double)(bb * zz);
    }
    v[1] = v[0];
    block_partials<2>(v, part);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                     const int* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ r)
{
    __shared__ T prod[kTileNnz];
    int row, r1;
    const T ax = tile_row_product<T, T>(tiles, rowptr, col, val, x, prod, row, r1);
    if (row < r1) r[row] = b[row] - ax;
}

// p = z; partials (r,z), (b,t), t = M^-1 b (t may be null)
template <typename T>
__global__ void k_cg_start(int n, const T* __restrict__ r, const T* __restrict__ z, T* __restrict__ p,
                           const T* __restrict__ b, const T* __restrict__ t, double* __restrict__ part)
{
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        p[i] = z[i];
        v[0] += (double)(r[i] * z[i]);
        if (t) v[1] += (double)(b[i] * t[i]);
    }
    block_partials<2>(v, part);
}

// gamma_0 = (r_0, z_0) and the stopping threshold: absolute (r, M^-1 r) < tol^2 (linear_solver.cu:360),
// or relative to (b, M^-1 b); a zero right-hand side falls back to the absolute test. One block.
__global__ void k_cg_begin(CgState* st, const double* __restrict__ part, int nb)
{
    double tot[2];
    sum_partials<2>(part, nb, tot);
    if (threadIdx.x != 0) return;
    const double rz = tot[0], bmb = tot[1];
    st->gam[0] = rz;
    st->thr = st->tol_mode == PFEM_TOL_RELATIVE && bmb > 0.0 ? st->tol * st->tol * bmb : st->tol * st->tol;
    st->it = 0;
    st->nan = rz != rz;
    st->converged = rz < st->thr ? 1 : 0;
    st->done = st->converged | st->nan;
    st->res = sqrt(fabs(rz));
}

// q = A p; partials (p, q)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const T* __restrict__ val,
                                                    const T* __restrict__ p, T* __restrict__ q, const CgState* st,
                                                    double* __restrict__ part)
{
    __shared__ T prod[kTileNnz];
    if (st->done) return;
    int row, r1;
    const T ap = tile_row_product<T, T>(tiles, rowptr, col, val, p, prod, row, r1);
    double v[1] = {0.0};
    if (row < r1) {
        q[row] = ap;
        v[0] = (double)(p[row] * ap);
    }
    block_partials<1>(v, part);
}

// q = A p with alpha_k = gamma_k / (p, q) finished by the launch's last workgroup from the published
// partials (the sum cg_alpha forms), for the fused PCG iteration: the same values as k_cg_spmv followed
// by k_cg_alpha, one launch fewer.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_alpha(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                          const int* __restrict__ col, const T* __restrict__ val,
                                                          const T* __restrict__ p, T* __restrict__ q, CgState* st,
                                                          double* part, int* err, int limit)
{
    __shared__ T prod[kTileNnz];
    if (st->done) {
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) st->xp_seen = st->it;
        return;
    }
    int row, r1;
    const T ap = tile_row_product<T, T>(tiles, rowptr, col, val, p, prod, row, r1);
    double v[1] = {0.0};
    if (row < r1) {
        q[row] = ap;
        v[0] = (double)(p[row] * ap);
    }
    block_sum<1>(v);
    publish_partial(v[0], part);
    if (blockIdx.x == gridDim.x - 1) {
        const double pq = collect_partials(part, gridDim.x, err, limit);
        if (threadIdx.x == 0) cg_alpha_record(st, pq);
    }
}

// As k_cg_spmv_alpha over the tile-sliced matrix (tile header {first row, rows, base, width}; entry w of
// the tile's row t at base + w * rows + t, unused slots with column -1): a lane finds its row's entries
// without loading row offsets, the loads of one entry index are unit-stride over the tile, and the row's
// products are added in CSR order, so q is bit-identical to the CSR tile kernel's. Column indices are
// 32-bit (col16 null) or 16-bit codes; the 32-bit form measured faster here (20.7 against 21.2 us).
__global__ __launch_bounds__(kBlock) void k_cg_spmv_alpha_sliced(const int* __restrict__ hdr, const int* __restrict__ col,
                                                                 const unsigned short* __restrict__ col16,
                                                                 const int* __restrict__ cbase,
                                                                 const double* __restrict__ val,
                                                                 const double* __restrict__ p, double* __restrict__ q,
                                                                 CgState* st, double* part, int* err, int limit)
{
    if (st->done) {
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) st->xp_seen = st->it;
        return;
    }
    const int* h = hdr + 4 * hdr_slot();
    const ChunkBases cb = col16 ? chunk_bases(cbase + 4 * hdr_slot()) : ChunkBases{0, 0, 0, 0, true};
    const int cnt = h[1], width = h[3];
    double v[1] = {0.0};
    if ((int)threadIdx.x < cnt) {
        const int row = h[0] + (int)threadIdx.x, e0 = h[2] + (int)threadIdx.x;
        double s = 0.0;
        constexpr int kU = 8;
        for (int w0 = 0; w0 < width; w0 += kU) {
            int cv[kU];
            double av[kU], xv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                cv[u] = -1;
                if (w0 + u < width) {
                    cv[u] = slot_col(e0 + (w0 + u) * cnt, col, col16, cb);
                    av[u] = val[e0 + (w0 + u) * cnt];
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (cv[u] >= 0) xv[u] = p[cv[u]];
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (cv[u] >= 0) s += av[u] * xv[u];
        }
        q[row] = s;
        v[0] = p[row] * s;
    }
    block_sum<1>(v);
    publish_partial(v[0], part);
    if (blockIdx.x == gridDim.x - 1) {
        const double pq = collect_partials(part, gridDim.x, err, limit);
        if (threadIdx.x == 0) cg_alpha_record(st, pq);
    }
}

// End of fused PCG iteration k = it - 1 (the backward sweep has just counted it): x += alpha_k p_k (the x
// half of updateXR, linear_solver.cu:68-77, moved here from the forward sweep so that p_k is read once for
// both updates), then, unless iteration k stopped the solve, p = z + beta p with beta_{k+1} =
// gamma_{k+1} / gamma_k already recorded by the backward sweep's last workgroup (updateP,
// linear_solver.cu:79-87). A launch queued after the solve stopped finds it == xp_seen and does nothing.
template <typename T>
__global__ void k_cg_xp_recorded(int n, T* __restrict__ x, T* __restrict__ p, const T* __restrict__ z,
                                 const CgState* st)
{
    const int it = st->it;
    if (it == st->xp_seen) return;
    const int k = it - 1;
    const bool last = st->done;
    const T al = (T)st->alpha[k & 1];
    const T beta = last ? (T)0 : (T)(st->gam[(k + 1) & 1] / st->gam[k & 1]);
    // two elements per lane (16-byte loads and stores where n allows)
    const int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i + 1 < n) {
        using V2 = HIP_vector_type<T, 2>;
        const V2 pv = *reinterpret_cast<const V2*>(p + i);
        V2 xv = *reinterpret_cast<const V2*>(x + i);
        xv.x += al * pv.x;
        xv.y += al * pv.y;
        *reinterpret_cast<V2*>(x + i) = xv;
        if (!last) {
            const V2 zv = *reinterpret_cast<const V2*>(z + i);
            V2 nv;
            nv.x = zv.x + beta * pv.x;
            nv.y = zv.y + beta * pv.y;
            *reinterpret_cast<V2*>(p + i) = nv;
        }
    } else if (i < n) {
        const T pi = p[i];
        x[i] += al * pi;
        if (!last) p[i] = z[i] + beta * pi;
    }
}

// alpha_k = gamma_k / (p, q) from the SpMV partials
__device__ __forceinline__ double cg_alpha(const CgState* st, const double* pa, int na, int k)
{
    double pq[1];
    sum_partials<1>(pa, na, pq);
    return st->gam[k & 1] / pq[0];
}

// x += alpha p, r -= alpha q (updateXR, linear_solver.cu:68-77); z = invD r; partials (r, z)
template <typename T, bool JAC>
__global__ void k_cg_update(int n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                            const T* __restrict__ q, T* __restrict__ z, const T* __restrict__ invd, const CgState* st,
                            const double* __restrict__ pa, int na, double* __restrict__ pb, int k)
{
    if (st->done) return;
    const T alpha = (T)cg_alpha(st, pa, na, k);
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const T rr = r[i] - alpha * q[i];
        r[i] = rr;
        const T zz = JAC ? rr * invd[i] : rr;
        if (JAC) z[i] = zz;
        v[0] += (double)(rr * zz);
    }
    block_partials<1>(v, pb);
}

// alpha_k alone (one block), for the iteration whose x/r update is fused into the forward sweep
__global__ void k_cg_alpha(CgState* st, const double* __restrict__ pa, int na, int k)
{
    if (st->done) return;
    const double a = cg_alpha(st, pa, na, k);
    if (threadIdx.x == 0) st->alpha[k & 1] = a;
}

template <typename T>
__global__ void k_cg_update_xr(int n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                               const T* __restrict__ q, CgState* st, const double* __restrict__ pa, int na, int k)
{
    if (st->done) return;
    const double a = cg_alpha(st, pa, na, k);
    const T alpha = (T)a;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        r[i] = r[i] - alpha * q[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha[k & 1] = a;
}

template <typename T>
__global__ void k_cg_rz(int n, const T* __restrict__ r, const T* __restrict__ z, const CgState* st,
                        double* __restrict__ pb)
{
    if (st->done) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[0] += (double)(r[i] * z[i]);
    block_partials<1>(v, pb);
}

// gamma_{k+1} from the (r, z) partials; p = z + beta p (updateP, linear_solver.cu:79-87); block 0
// records the iteration and the stopping test
template <typename T>
__global__ void k_cg_p(int n, T* __restrict__ p, const T* __restrict__ z, CgState* st,
                       const double* __restrict__ pb, int nb, int k)
{
    if (st->done) return;
    double g[1];
    sum_partials<1>(pb, nb, g);
    const double gnew = g[0];
    const T beta = (T)(gnew / st->gam[k & 1]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) cg_record(st, gnew, k);
}

// ------------------------------------------------------------------ GMRES kernels
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_gm_residual(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                        const int* __restrict__ col, const T* __restrict__ val,
                                                        const T* __restrict__ x, const T* __restrict__ b,
                                                        const T* __restrict__ invd, T* __restrict__ w,
                                                        double* __restrict__ part)
{
    __shared__ T prod[kTileNnz];
    int row, r1;
    const T ax = tile_row_product<T, T>(tiles, rowptr, col, val, x, prod, row, r1);
    double v[2] = {0.0, 0.0};
    if (row < r1) {
        const T bb = b[row];
        const T rr = bb - ax;
        const T ww = JAC ? rr * invd[row] : rr;
        const T mb = JAC ? bb * invd[row] : bb;
        w[row] = ww;
        v[0] = (double)(ww * ww);
        v[1] = (double)(mb * mb);
    }
    block_partials<2>(v, part);
}

// x = 0 start: w = M^-1 b (Jacobi or none)
template <typename T, bool JAC>
__global__ void k_gm_init_zero(int n, const T* __restrict__ b, const T* __restrict__ invd, T* __restrict__ x,
                               T* __restrict__ w, double* __restrict__ part)
{
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T bb = b[i];
        const T ww = JAC ? bb * invd[i] : bb;
        x[i] = 0;
        w[i] = ww;
        v[0] += (double)(ww * ww);
    }
    v[1] = v[0];
    block_partials<2>(v, part);
}

template <typename T>
__global__ void k_gm_norms(int n, const T* __restrict__ w, const T* __restrict__ t, double* __restrict__ part)
{
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += (double)(w[i] * w[i]);
        if (t) v[1] += (double)(t[i] * t[i]);
    }
    block_partials<2>(v, part);
}

// start of a cycle from the ||w||^2, ||M^-1 b||^2 partials (one block)
__global__ void k_gm_begin(GmState* st, const double* __restrict__ part, int nb)
{
    double tot[2];
    sum_partials<2>(part, nb, tot);
    if (threadIdx.x == 0) gm_cycle_start(st, tot[0], tot[1]);
}

// v /= ||v|| (scaleVector with divide = true, linear_solver.cu:16-26, :456, :493)
template <typename T>
__global__ void k_gm_scale(int n, T* __restrict__ v, const GmState* st)
{
    if (st->cycle_done) return;
    const T nrm = (T)st->norm;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = v[i] / nrm;
}

// w = M^-1 A v_j (Jacobi fused: applyPreconditioner(v_kp, v_kp), linear_solver.cu:473-474)
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_gm_arnoldi_spmv(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                            const int* __restrict__ col, const T* __restrict__ val,
                                                            const T* __restrict__ v, const T* __restrict__ invd,
                                                            T* __restrict__ w, const GmState* st)
{
    __shared__ T prod[kTileNnz];
    if (st->cycle_done) return;
    int row, r1;
    const T av = tile_row_product<T, T>(tiles, rowptr, col, val, v, prod, row, r1);
    if (row < r1) w[row] = JAC ? av * invd[row] : av;
}

// partials of (V_{i0+t}, w), t < cnt <= 8, stored as coefficient i0+t of part
template <typename T>
__global__ void k_gm_multidot(int n, const T* __restrict__ V, long ld, int i0, int cnt, const T* __restrict__ w,
                              const GmState* st, double* __restrict__ part)
{
    __shared__ double sh[8][kWaves];
    if (st->cycle_done) return;
    double v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T wi = w[i];
#pragma unroll
        for (int t = 0; t < 8; ++t)
            if (t < cnt) v[t] += (double)(V[(size_t)(i0 + t) * ld + i] * wi);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        if (t < cnt) {
            const double s = wave_sum(v[t]);
            if (lane == 0) sh[t][wid] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < cnt) {
        const int t = threadIdx.x;
        double s = sh[t][0];
        for (int w2 = 1; w2 < kWaves; ++w2) s += sh[t][w2];
        part[(size_t)(i0 + t) * gridDim.x + blockIdx.x] = s;
    }
}

// block i sums the partials of coefficient i: hc_i, and h_{i,j} stored (pass 0) or corrected (pass 1)
__global__ void k_gm_reduce_h(GmState* st, const double* __restrict__ part, int nb, int pass)
{
    if (st->cycle_done) return;
    double tot[1];
    sum_partials<1>(part + (size_t)blockIdx.x * nb, nb, tot);
    if (threadIdx.x == 0) {
        const int i = blockIdx.x;
        double* h = st->H + (size_t)st->k * st->m;
        st->hc[i] = tot[0];
        h[i] = pass == 0 ? tot[0] : h[i] + tot[0];
    }
}

// w -= sum_i hc_i V_i (i <= j); on the last pass also the partials of ||w||^2
template <typename T>
__global__ void k_gm_orth(int n, const T* __restrict__ V, long ld, T* __restrict__ w, const GmState* st,
                          double* __restrict__ pn, int last_pass)
{
    __shared__ double hc[kMaxRestart + 1];
    if (st->cycle_done) return;
    const int j = st->k;
    for (int i = threadIdx.x; i <= j; i += blockDim.x) hc[i] = st->hc[i];
    __syncthreads();
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T wi = w[i];
        for (int t = 0; t <= j; ++t) wi -= (T)hc[t] * V[(size_t)t * ld + i];
        w[i] = wi;
        v[0] += (double)(wi * wi);
    }
    if (last_pass) block_partials<1>(v, pn);
}

// end of Arnoldi step j from the ||w||^2 partials (one block)
__global__ void k_gm_finish(GmState* st, const double* __restrict__ pn, int nb)
{
    if (st->cycle_done) return;
    double tot[1];
    sum_partials<1>(pn, nb, tot);
    if (threadIdx.x == 0) gm_arnoldi_finish(st, sqrt(tot[0]));
}

// y = H^-1 beta (cublasDtrsv, upper, non-unit; linear_solver.cu:508-511)
__global__ void k_gm_solve_y(GmState* st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->apply_update = (st->converged || st->restarted) && st->k > 0 && !st->nan ? 1 : 0;
    if (!st->apply_update) return;
    const int k = st->k, m = st->m;
    for (int i = 0; i < k; ++i) st->y[i] = st->beta[i];
    for (int i = k - 1; i >= 0; --i) {
        double s = st->y[i];
        for (int j = i + 1; j < k; ++j) s -= st->H[(size_t)m * j + i] * st->y[j];
        st->y[i] = s / st->H[(size_t)m * i + i];
    }
}

// x += V y (cublasDgemv with alpha = beta = 1; linear_solver.cu:513-514); clears cycle_done for a restart
template <typename T>
__global__ void k_gm_update_x(int n, const T* __restrict__ V, long ld, T* __restrict__ x, const GmState* st)
{
    __shared__ double y[kMaxRestart];
    if (!st->apply_update) return;
    const int k = st->k;
    for (int i = threadIdx.x; i < k; i += blockDim.x) y[i] = st->y[i];
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T s = 0;
        for (int j = 0; j < k; ++j) s += V[(size_t)j * ld + i] * (T)y[j];
        x[i] = s + x[i];
    }
}

__global__ void k_gm_rearm(GmState* st)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) st->cycle_done = 0;
}

// ------------------------------------------------------------------ refinement kernels
// FP64 refinement state
struct IrState {
    double thr, rz, scale;
    int first, converged, nan;
};

// r = b - A x (FP64), z = invD r (Jacobi) / r; partials (r,z) and (b, M^-1 b); gmres: ||z||^2, ||M^-1 b||^2
template <bool JAC, bool GM>
__global__ __launch_bounds__(kBlock) void k_ir_residual(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                        const int* __restrict__ col, const double* __restrict__ val,
                                                        const double* __restrict__ x, const double* __restrict__ b,
                                                        const double* __restrict__ invd, double* __restrict__ r,
                                                        double* __restrict__ z, double* __restrict__ part)
{
    __shared__ double prod[kTileNnz];
    int row, r1;
    const double ax = tile_row_product<double, double>(tiles, rowptr, col, val, x, prod, row, r1);
    double v[2] = {0.0, 0.0};
    if (row < r1) {
        const double bb = b[row];
        const double rr = bb - ax;
        const double zz = JAC ? rr * invd[row] : rr;
        r[row] = rr;
        z[row] = zz;
        const double mb = JAC ? bb * invd[row] : bb;
        v[0] = GM ? zz * zz : rr * zz;
        v[1] = GM ? mb * mb : bb * mb;
    }
    block_partials<2>(v, part);
}

// general preconditioner: z = M^-1 r already applied; partials as above (t = M^-1 b, may be null)
template <bool GM>
__global__ void k_ir_norms(int n, const double* __restrict__ r, const double* __restrict__ z,
                           const double* __restrict__ b, const double* __restrict__ t, double* __restrict__ part)
{
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += GM ? z[i] * z[i] : r[i] * z[i];
        if (t) v[1] += GM ? t[i] * t[i] : b[i] * t[i];
    }
    block_partials<2>(v, part);
}

// refinement stopping test (the FP64 solver's criterion) and the scale of the next correction; one block
template <bool GM>
__global__ void k_ir_test(IrState* st, const double* __restrict__ part, int nb, double tol, int tol_mode)
{
    double tot[2];
    sum_partials<2>(part, nb, tot);
    if (threadIdx.x != 0) return;
    if (st->first) {
        const bool rel = tol_mode == PFEM_TOL_RELATIVE && tot[1] > 0.0;
        st->thr = GM ? (rel ? tol * sqrt(tot[1]) : tol) : (rel ? tol * tol * tot[1] : tol * tol);
        st->first = 0;
    }
    st->rz = tot[0];
    const double meas = GM ? sqrt(tot[0]) : tot[0];
    st->nan = tot[0] != tot[0];
    st->converged = meas < st->thr ? 1 : 0;
    st->scale = sqrt(tot[0]) > 0.0 ? sqrt(tot[0]) : 1.0;
}

__global__ void k_ir_down(int n, float* __restrict__ d, const double* __restrict__ s, const IrState* st)
{
    const double inv = 1.0 / st->scale;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = (float)(s[i] * inv);
}

__global__ void k_ir_up(int n, double* __restrict__ x, const float* __restrict__ d, const IrState* st)
{
    const double sc = st->scale;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += sc * (double)d[i];
}

__global__ void k_to32(long n, float* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

// to and from a multicolour factor's ordering: d[i] = s[perm[i]] / d[perm[i]] = s[i]
__global__ void k_perm_gather(int n, const int* __restrict__ perm, const double* __restrict__ s, double* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[perm[i]];
}

__global__ void k_perm_scatter(int n, const int* __restrict__ perm, const double* __restrict__ s, double* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[perm[i]] = s[i];
}

// ------------------------------------------------------------------ host drivers
template <typename T>
struct KArgs {
    const Pattern* P;
    const T* val;
    const T* invd;       // Jacobi diagonal (null otherwise)
    int kind;            // preconditioner kind
    pfem_precond_t* pc;
    T* x;
    const T* b;
    bool x_zero;
    double tol;
    int tol_mode, maxit;
    T *r, *z, *p, *q, *t, *V;
    bool pc_fp32;        // FP64 vectors, IC0/ILU0 factor applied in FP32
    bool factor_order;   // the system is in the multicolour factor's ordering (P A P^T, P b, P x)
};

template <typename T>
static void papply(Context* c, const KArgs<T>& A, T* z, const T* r)
{
    precond_apply_t<T>(c, A.pc, z, r, A.pc_fp32, ApplyOpts<T>{nullptr, 0, A.factor_order});
}

template <typename T>
static void cg_run(Context* c, pfem_solver_t* s, const KArgs<T>& A, CgState* st, CgState* hst)
{
    const Pattern& P = *A.P;
    const int n = P.n;
    const unsigned g = vec_grid(n);
    const bool jac = A.kind == PFEM_PRECOND_JACOBI, none = A.kind == PFEM_PRECOND_NONE;
    // x/r update fused into the forward sweep: multicolour IC0/ILU0, single-pass sweeps, factor ordering
    // (a blocked factor has the fused forms in their tagged form only)
    const bool fused = A.factor_order && A.pc && A.pc->sweep &&
                       (A.pc->contiguous || (A.pc->blocked && s->fuse >= 3)) &&
                       std::max(A.pc->nuchunks, A.pc->nlchunks) <= s->pb_cap && s->fuse >= 1;
    // level 2: four launches per iteration, the reductions finished inside the launches that produce
    // them: SpMV (+ alpha), forward sweep (+ x/r), backward sweep (+ beta and the stopping test), p update
    const bool fused2 = fused && s->fuse >= 2 && P.ntiles <= s->pa_cap;
    // level 3: the sweeps share z's storage for y and z, told apart by tags instead of re-armed sentinels
    // (two 8-byte passes less per iteration; y and z change by at most one unit in the last place)
    bool tagged = false;
    if constexpr (sizeof(T) == 8) tagged = fused2 && s->fuse >= 3 && A.pc->sell;
    // with IC0 at level 3 the forward sweep ends the iteration (gamma = (y, y)) and the backward sweep
    // applies the x and p updates row by row: three launches per iteration
    const bool xp_in_sweep = tagged && A.kind == PFEM_PRECOND_IC0;
    CgState init{};
    init.tol = A.tol;
    init.tol_mode = A.tol_mode;
    init.maxit = A.maxit;
    PFEM_HIP(hipMemcpyAsync(st, &init, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    int nb0 = (int)g;   // grid of the kernel that produced the start partials
    if (jac || none) {
        if (A.x_zero) {
            if (jac) hipLaunchKernelGGL((k_cg_init_zero<T, true>), dim3(g), dim3(kBlock), 0, c->stream, n, A.b, A.invd, A.x, A.r, A.z, A.p, s->pa);
            else hipLaunchKernelGGL((k_cg_init_zero<T, false>), dim3(g), dim3(kBlock), 0, c->stream, n, A.b, A.invd, A.x, A.r, A.z, A.p, s->pa);
        } else {
            nb0 = P.ntiles;
            if (jac) hipLaunchKernelGGL((k_cg_init<T, true>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, A.x, A.b, A.invd, A.r, A.z, A.p, s->pa);
            else hipLaunchKernelGGL((k_cg_init<T, false>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, A.x, A.b, A.invd, A.r, A.z, A.p, s->pa);
        }
    } else {
        if (A.x_zero) {
            PFEM_HIP(hipMemsetAsync(A.x, 0, sizeof(T) * n, c->stream));
            PFEM_HIP(hipMemcpyAsync(A.r, A.b, sizeof(T) * n, hipMemcpyDeviceToDevice, c->stream));
        } else {
            hipLaunchKernelGGL((k_residual<T>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, A.x, A.b, A.r);
        }
        papply<T>(c, A, A.z, A.r);
        const T* t = nullptr;
        if (A.tol_mode == PFEM_TOL_RELATIVE) {
            if (A.x_zero) t = A.z;
            else { papply<T>(c, A, A.t, A.b); t = A.t; }
        }
        hipLaunchKernelGGL((k_cg_start<T>), dim3(g), dim3(kBlock), 0, c->stream, n, A.r, A.z, A.p, A.b, t, s->pa);
    }
    hipLaunchKernelGGL(k_cg_begin, dim3(1), dim3(kBlock), 0, c->stream, st, (const double*)s->pa, nb0);
    if (fused2) {
        hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(blocks_for(P.ntiles)), dim3(kBlock), 0, c->stream,
                           (long)P.ntiles, s->pa);
        const int nslots = std::max(std::max(A.pc->nuchunks, A.pc->nlchunks), std::max(A.pc->nuslices, A.pc->nlslices));
        hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(blocks_for(nslots)), dim3(kBlock), 0, c->stream,
                           (long)nslots, s->pb);
        // z_0 has been copied into p: z's storage now becomes the sweeps' tagged buffer
        if constexpr (sizeof(T) == 8)
            if (tagged)
                hipLaunchKernelGGL((k_fill_tag0<T>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, (long)n, A.z);
    }
    PFEM_HIP(hipGetLastError());
    const int batch = s->poll;
    const int na = P.ntiles, nb = (int)g;
    int launched = 0;
    // one fused iteration (four launches, three with xp_in_sweep); the iteration index lives on the
    // device, so every iteration is the same sequence of launches
    auto fused_iteration = [&]() {
        bool sliced = false;
        if constexpr (sizeof(T) == 8) {
            sliced = A.pc->asl_val.p && A.val == A.pc->aval.p;
            if (sliced)
                hipLaunchKernelGGL(k_cg_spmv_alpha_sliced, dim3(P.ntiles), dim3(kBlock), 0, c->stream,
                                   A.pc->sl_all.hdr.p, A.pc->sl_all.col.p,
                                   (const unsigned short*)nullptr, A.pc->sl_all.cbase.p,
                                   (const double*)A.pc->asl_val.p,
                                   (const double*)A.p, A.q, st, s->pa, A.pc->sweep_err.p, A.pc->spin_limit);
        }
        if (!sliced)
            hipLaunchKernelGGL((k_cg_spmv_alpha<T>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p,
                               P.rowptr.p, P.col.p, A.val, A.p, A.q, st, s->pa, A.pc->sweep_err.p,
                               A.pc->spin_limit);
        ApplyOpts<T> o{s->pb, s->pb_cap, true};
        o.done = &st->done;
        o.q = A.q;
        o.alpha = st->alpha;
        o.alpha_cg = st;
        o.cg = st;
        o.tagged = tagged;
        if (xp_in_sweep) {
            o.xpx = A.x;
            o.xpp = A.p;
        }
        PFEM_REQUIRE(precond_apply_t<T>(c, A.pc, A.z, A.r, A.pc_fp32, o) > 0,
                     "fused PCG iteration without its partials");
        if (!xp_in_sweep)
            hipLaunchKernelGGL((k_cg_xp_recorded<T>), dim3(blocks_for((n + 1) / 2)), dim3(kBlock), 0, c->stream, n,
                               A.x, A.p, (const T*)A.z, st);
    };
    // a batch of fused iterations captured once per solve as a HIP graph and replayed
    hipGraphExec_t exec = nullptr;
    if (fused2 && s->graph) {
        hipGraph_t graph = nullptr;
        PFEM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int j = 0; j < batch; ++j) fused_iteration();
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        PFEM_HIP(hipStreamEndCapture(c->stream, &graph));
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PFEM_HIP(e);
    }
    struct ExecGuard {   // the replays queued on the stream finish before the executable graph goes
        hipGraphExec_t& e;
        hipStream_t st;
        ~ExecGuard()
        {
            if (!e) return;
            (void)hipStreamSynchronize(st);
            (void)hipGraphExecDestroy(e);
        }
    } exec_guard{exec, c->stream};
    // the state of batch b is copied to slot b & 1 and checked after batch b + 1 is queued, so the stream
    // never drains while the host waits; a converged solve runs at most one batch of no-op launches
    int pend = -1, pend_launched = 0, slot = 0;
    while (true) {
        if (exec) PFEM_HIP(hipGraphLaunch(exec, c->stream));
        for (int j = 0; j < batch && !exec; ++j) {
            const int k = launched + j;
            int nbk = nb;
            if (fused2) {
                fused_iteration();
                continue;
            }
            hipLaunchKernelGGL((k_cg_spmv<T>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, A.p, A.q, st, s->pa);
            if (jac) {
                hipLaunchKernelGGL((k_cg_update<T, true>), dim3(g), dim3(kBlock), 0, c->stream, n, A.x, A.r, A.p, A.q, A.z, A.invd, st, s->pa, na, s->pb, k);
            } else if (none) {
                hipLaunchKernelGGL((k_cg_update<T, false>), dim3(g), dim3(kBlock), 0, c->stream, n, A.x, A.r, A.p, A.q, A.r, A.invd, st, s->pa, na, s->pb, k);
            } else if (fused) {
                // alpha, then the forward sweep applies x += alpha p, r -= alpha q row by row; the backward
                // sweep produces the (r, z) partials
                hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kBlock), 0, c->stream, st, (const double*)s->pa, na, k);
                ApplyOpts<T> o{s->pb, s->pb_cap, true};
                o.done = &st->done;
                o.x = A.x;
                o.p = A.p;
                o.q = A.q;
                o.alpha = &st->alpha[k & 1];
                nbk = precond_apply_t<T>(c, A.pc, A.z, A.r, A.pc_fp32, o);
                PFEM_REQUIRE(nbk > 0, "fused PCG iteration without its partials");
            } else {
                hipLaunchKernelGGL((k_cg_update_xr<T>), dim3(g), dim3(kBlock), 0, c->stream, n, A.x, A.r, A.p, A.q, st, s->pa, na, k);
                // the triangular sweeps also produce the (r, z) partials when they can
                nbk = precond_apply_t<T>(c, A.pc, A.z, A.r, A.pc_fp32, ApplyOpts<T>{s->pb, s->pb_cap, A.factor_order});
                if (!nbk) {
                    hipLaunchKernelGGL((k_cg_rz<T>), dim3(g), dim3(kBlock), 0, c->stream, n, A.r, A.z, st, s->pb);
                    nbk = nb;
                }
            }
            hipLaunchKernelGGL((k_cg_p<T>), dim3(g), dim3(kBlock), 0, c->stream, n, A.p, none ? A.r : A.z, st, s->pb, nbk, k);
        }
        launched += batch;
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(hst + slot, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipEventRecord(s->ev[slot], c->stream));
        if (pend >= 0) {
            PFEM_HIP(hipEventSynchronize(s->ev[pend]));
            if (hst[pend].done || pend_launched > A.maxit + batch) break;
        }
        pend = slot;
        pend_launched = launched;
        slot ^= 1;
    }
    PFEM_HIP(hipMemcpyAsync(hst, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
}

// One GMRES(m) system's launch sequence (SolverGMRES::solve, linear_solver.cu:421-521), split into the
// pieces between the Arnoldi SpMVs so that gmres_run_pair can advance two systems in lockstep.
template <typename T>
struct GmresSys {
    Context* c;
    pfem_solver_t* s;
    KArgs<T> A;
    GmState* st;
    GmState* hst;
    bool first = true;
    bool done = false;

    int n() const { return A.P->n; }
    unsigned g() const { return vec_grid(A.P->n); }
    bool jac() const { return A.kind == PFEM_PRECOND_JACOBI; }
    bool none() const { return A.kind == PFEM_PRECOND_NONE; }
    T* v(int j) const { return A.V + (size_t)j * A.P->n; }

    void init()
    {
        GmState ini{};
        ini.tol = A.tol;
        ini.tol_mode = A.tol_mode;
        ini.maxit = A.maxit;
        ini.m = s->m;
        ini.restarted = s->restarted;
        ini.thr = -1.0;
        ini.H = s->H.p;
        ini.beta = s->beta.p;
        ini.cs = s->cs.p;
        ini.sn = s->sn.p;
        ini.y = s->y.p;
        ini.hc = s->hc.p;
        PFEM_HIP(hipMemcpyAsync(st, &ini, sizeof(GmState), hipMemcpyHostToDevice, c->stream));
        first = true;
        done = false;
    }
    // v_0 = M^-1 (b - A x) / ||.||, with the stopping threshold on the first cycle
    void cycle_start()
    {
        const Pattern& P = *A.P;
        const int nn = n();
        const unsigned gg = g();
        T* V = A.V;
        int nb0 = (int)gg;
        if (jac() || none()) {
            if (first && A.x_zero) {
                if (jac()) hipLaunchKernelGGL((k_gm_init_zero<T, true>), dim3(gg), dim3(kBlock), 0, c->stream, nn, A.b, A.invd, A.x, V, s->pa);
                else hipLaunchKernelGGL((k_gm_init_zero<T, false>), dim3(gg), dim3(kBlock), 0, c->stream, nn, A.b, A.invd, A.x, V, s->pa);
            } else {
                nb0 = P.ntiles;
                if (jac()) hipLaunchKernelGGL((k_gm_residual<T, true>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, A.x, A.b, A.invd, V, s->pa);
                else hipLaunchKernelGGL((k_gm_residual<T, false>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, A.x, A.b, A.invd, V, s->pa);
            }
        } else {
            if (first && A.x_zero) {
                PFEM_HIP(hipMemsetAsync(A.x, 0, sizeof(T) * nn, c->stream));
                PFEM_HIP(hipMemcpyAsync(V, A.b, sizeof(T) * nn, hipMemcpyDeviceToDevice, c->stream));
            } else {
                hipLaunchKernelGGL((k_residual<T>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, A.x, A.b, V);
            }
            papply<T>(c, A, V, V);
            const T* t = nullptr;
            if (first && A.tol_mode == PFEM_TOL_RELATIVE) {
                if (A.x_zero) t = V;
                else { papply<T>(c, A, A.t, A.b); t = A.t; }
            }
            hipLaunchKernelGGL((k_gm_norms<T>), dim3(gg), dim3(kBlock), 0, c->stream, nn, V, t, s->pa);
        }
        hipLaunchKernelGGL(k_gm_begin, dim3(1), dim3(kBlock), 0, c->stream, st, (const double*)s->pa, nb0);
        first = false;
        hipLaunchKernelGGL((k_gm_scale<T>), dim3(gg), dim3(kBlock), 0, c->stream, nn, V, st);
    }
    // w = M^-1 A v_j into v_{j+1}
    void spmv(int j)
    {
        const Pattern& P = *A.P;
        if (jac()) hipLaunchKernelGGL((k_gm_arnoldi_spmv<T, true>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, v(j), A.invd, v(j + 1), st);
        else hipLaunchKernelGGL((k_gm_arnoldi_spmv<T, false>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, A.val, v(j), A.invd, v(j + 1), st);
    }
    // the rest of Arnoldi step j: preconditioner (other than Jacobi), Gram-Schmidt, Givens, normalisation
    void arnoldi_rest(int j)
    {
        const int nn = n();
        const unsigned gg = g();
        const long ld = nn;
        T* w = v(j + 1);
        if (!jac() && !none()) papply<T>(c, A, w, w);
        const int passes = s->reorth ? 2 : 1;
        for (int pass = 0; pass < passes; ++pass) {
            for (int i0 = 0; i0 <= j; i0 += 8)
                hipLaunchKernelGGL((k_gm_multidot<T>), dim3(gg), dim3(kBlock), 0, c->stream, nn, A.V, ld, i0, std::min(8, j + 1 - i0), w, st, s->pb);
            hipLaunchKernelGGL(k_gm_reduce_h, dim3(j + 1), dim3(kBlock), 0, c->stream, st, (const double*)s->pb, (int)gg, pass);
            hipLaunchKernelGGL((k_gm_orth<T>), dim3(gg), dim3(kBlock), 0, c->stream, nn, A.V, ld, w, st, s->pn, pass == passes - 1 ? 1 : 0);
        }
        hipLaunchKernelGGL(k_gm_finish, dim3(1), dim3(kBlock), 0, c->stream, st, (const double*)s->pn, (int)gg);
        hipLaunchKernelGGL((k_gm_scale<T>), dim3(gg), dim3(kBlock), 0, c->stream, nn, w, st);
    }
    void read_state() { PFEM_HIP(hipMemcpyAsync(hst, st, sizeof(GmState), hipMemcpyDeviceToHost, c->stream)); }
    // y = H^-1 beta, x += V y (when the cycle converged, or every cycle of GMRES(m))
    void finish_cycle()
    {
        hipLaunchKernelGGL(k_gm_solve_y, dim3(1), dim3(64), 0, c->stream, st);
        hipLaunchKernelGGL((k_gm_update_x<T>), dim3(g()), dim3(kBlock), 0, c->stream, n(), A.V, (long)n(), A.x, st);
    }
    // after finish_cycle and a state read: true when the solve is over, else the next cycle is armed
    bool next_cycle()
    {
        if (hst->converged || hst->nan || hst->it >= A.maxit || !s->restarted) {
            done = true;
            return true;
        }
        hipLaunchKernelGGL(k_gm_rearm, dim3(1), dim3(64), 0, c->stream, st);
        return false;
    }
};

template <typename T>
static void gmres_run(Context* c, pfem_solver_t* s, const KArgs<T>& A, GmState* st, GmState* hst)
{
    GmresSys<T> g{c, s, A, st, hst};
    g.init();
    while (true) {
        g.cycle_start();
        for (int j0 = 0; j0 < s->m; j0 += s->poll) {
            const int j1 = std::min(s->m, j0 + s->poll);
            for (int j = j0; j < j1; ++j) {
                g.spmv(j);
                g.arnoldi_rest(j);
            }
            PFEM_HIP(hipGetLastError());
            g.read_state();
            PFEM_HIP(hipStreamSynchronize(c->stream));
            if (hst->cycle_done) break;
        }
        g.finish_cycle();
        PFEM_HIP(hipGetLastError());
        g.read_state();
        PFEM_HIP(hipStreamSynchronize(c->stream));
        if (g.next_cycle()) break;
    }
}

// w_k = M_k^-1 A_k v_k for two systems on one pattern (Jacobi or no preconditioner) in one launch: each
// tile stages the first system's products in LDS, sums its rows, then does the same for the second, so
// the tile's column indices come from HBM once (the second pass finds them in L2) with the LDS of a
// single-system tile; a system whose cycle has ended is skipped. Each row is summed in the order of
// k_gm_arnoldi_spmv (the same bits as two single launches).
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_gm_arnoldi_spmv_pair(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                                 const int* __restrict__ col, const T* __restrict__ val0,
                                                                 const T* __restrict__ val1, const T* __restrict__ v0,
                                                                 const T* __restrict__ v1, const T* __restrict__ invd0,
                                                                 const T* __restrict__ invd1, T* __restrict__ w0,
                                                                 T* __restrict__ w1, const GmState* st0, const GmState* st1)
{
    __shared__ T prod[kTileNnz];
    const bool on[2] = {!st0->cycle_done, !st1->cycle_done};
    if (!on[0] && !on[1]) return;
    const T* val[2] = {val0, val1};
    const T* v[2] = {v0, v1};
    const T* invd[2] = {invd0, invd1};
    T* w[2] = {w0, w1};
    for (int q = 0; q < 2; ++q) {
        if (!on[q]) continue;
        int row, r1;
        const T av = tile_row_product<T, T>(tiles, rowptr, col, val[q], v[q], prod, row, r1);
        if (row < r1) w[q][row] = JAC ? av * invd[q][row] : av;
        __syncthreads();   // the LDS tile is reused by the second system
    }
}

// Two GMRES(m) solves with one pattern (the velocity prediction's two components, whose matrices share
// the mesh's CSR structure; the reference solves them one after the other, Cylinder2D/main.cu:820-824):
// both systems advance in lockstep, their Arnoldi SpMVs in one launch that streams the pattern once, and
// one host poll serves both. Each system keeps its own recurrence, scalars and stopping test, so its
// iterates are those of a single solve (the SpMV sums each row in the same order).
template <typename T>
static void gmres_run_pair(Context* c, GmresSys<T>& g0, GmresSys<T>& g1)
{
    GmresSys<T>* g[2] = {&g0, &g1};
    const bool fused = g0.A.P == g1.A.P && g0.jac() == g1.jac() && (g0.jac() || g0.none()) && g0.s->m == g1.s->m;
    for (auto* x : g) x->init();
    while (!g0.done || !g1.done) {
        for (auto* x : g)
            if (!x->done) x->cycle_start();
        const int m = std::max(g0.s->m, g1.s->m), poll = std::min(g0.s->poll, g1.s->poll);
        for (int j0 = 0; j0 < m; j0 += poll) {
            const int j1 = std::min(m, j0 + poll);
            for (int j = j0; j < j1; ++j) {
                // a finished system's state stays cycle_done, so its launches return at once
                if (fused && !g0.done && !g1.done) {
                    const Pattern& P = *g0.A.P;
                    if (g0.jac())
                        hipLaunchKernelGGL((k_gm_arnoldi_spmv_pair<T, true>), dim3(P.ntiles), dim3(kBlock), 0, c->stream,
                                           P.tiles.p, P.rowptr.p, P.col.p, g0.A.val, g1.A.val, g0.v(j), g1.v(j),
                                           g0.A.invd, g1.A.invd, g0.v(j + 1), g1.v(j + 1), g0.st, g1.st);
                    else
                        hipLaunchKernelGGL((k_gm_arnoldi_spmv_pair<T, false>), dim3(P.ntiles), dim3(kBlock), 0, c->stream,
                                           P.tiles.p, P.rowptr.p, P.col.p, g0.A.val, g1.A.val, g0.v(j), g1.v(j),
                                           g0.A.invd, g1.A.invd, g0.v(j + 1), g1.v(j + 1), g0.st, g1.st);
                } else {
                    for (auto* x : g)
                        if (!x->done && j < x->s->m) x->spmv(j);
                }
                for (auto* x : g)
                    if (!x->done && j < x->s->m) x->arnoldi_rest(j);
            }
            PFEM_HIP(hipGetLastError());
            for (auto* x : g)
                if (!x->done) x->read_state();
            PFEM_HIP(hipStreamSynchronize(c->stream));
            if ((g0.done || g0.hst->cycle_done) && (g1.done || g1.hst->cycle_done)) break;
        }
        for (auto* x : g)
            if (!x->done) x->finish_cycle();
        PFEM_HIP(hipGetLastError());
        for (auto* x : g)
            if (!x->done) x->read_state();
        PFEM_HIP(hipStreamSynchronize(c->stream));
        for (auto* x : g)
            if (!x->done) x->next_cycle();
    }
}

static KArgs<double> args64(pfem_solver_t* s, pfem_csr_t* a, double* x, const double* b)
{
    KArgs<double> A{};
    A.pc_fp32 = s->precision == PFEM_PRECISION_MIXED_PRECOND && s->pc &&
                (s->pc->kind == PFEM_PRECOND_IC0 || s->pc->kind == PFEM_PRECOND_ILU0);
    A.P = a->pat.get();
    A.val = a->val.p;
    A.kind = s->pc ? s->pc->kind : PFEM_PRECOND_NONE;
    A.invd = A.kind == PFEM_PRECOND_JACOBI ? s->pc->invd.p : nullptr;
    A.pc = s->pc;
    A.x = x;
    A.b = b;
    A.x_zero = false;
    A.tol = s->tol;
    A.tol_mode = s->tol_mode;
    A.maxit = s->maxit;
    A.r = s->r.p; A.z = s->z.p; A.p = s->p.p; A.q = s->q.p; A.t = s->t.p; A.V = s->V.p;
    return A;
}

// A multicolour IC0/ILU0 with single-pass sweeps: the whole Krylov solve runs on P A P^T, P b, P x (the
// factor's ordering), so the sweeps read and write contiguously and nothing is permuted per iteration.
static bool factor_order_eligible(const pfem_solver_t* s)
{
    const pfem_precond_t* pc = s->pc;
    return pc && (pc->kind == PFEM_PRECOND_IC0 || pc->kind == PFEM_PRECOND_ILU0) &&
           pc->ordering != PFEM_ORDER_NATURAL && pc->sweep && pc->aval.p && s->factor_order;
}

static KArgs<double> args_factor_order(Context* c, pfem_solver_t* s, double* x, const double* b)
{
    pfem_precond_t* pc = s->pc;
    const int n = pc->n;
    if (!s->xp.p) {
        s->xp.alloc(n);
        s->bp.alloc(n);
    }
    hipLaunchKernelGGL(k_perm_gather, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->perm.p, b, s->bp.p);
    hipLaunchKernelGGL(k_perm_gather, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->perm.p,
                       (const double*)x, s->xp.p);
    PFEM_HIP(hipGetLastError());
    KArgs<double> A{};
    A.P = pc->fpat.get();
    A.val = pc->aval.p;
    A.kind = pc->kind;
    A.pc = pc;
    A.x = s->xp.p;
    A.b = s->bp.p;
    A.tol = s->tol;
    A.tol_mode = s->tol_mode;
    A.maxit = s->maxit;
    A.r = s->r.p; A.z = s->z.p; A.p = s->p.p; A.q = s->q.p; A.t = s->t.p; A.V = s->V.p;
    A.pc_fp32 = s->precision == PFEM_PRECISION_MIXED_PRECOND;
    A.factor_order = true;
    PFEM_REQUIRE(2 * (size_t)A.P->ntiles <= (size_t)s->pa_cap, "partials too small for the permuted pattern");
    return A;
}

static void scatter_back(Context* c, pfem_solver_t* s, double* x)
{
    const int n = s->pc->n;
    hipLaunchKernelGGL(k_perm_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, s->pc->perm.p,
                       (const double*)s->xp.p, x);
    PFEM_HIP(hipGetLastError());
}

void cg_solve_fp64(Context* c, pfem_solver_t* s, pfem_csr_t* a, double* x, const double* b, pfem_solve_stats* out)
{
    if (s->pc) precond_initialize(c, s->pc, a);
    if (factor_order_eligible(s)) {
        cg_run<double>(c, s, args_factor_order(c, s, x, b), s->cg.p, s->h_cg);
        scatter_back(c, s, x);
    } else {
        cg_run<double>(c, s, args64(s, a, x, b), s->cg.p, s->h_cg);
    }
    precond_check(c, s->pc);
    if (out) {
        out->converged = s->h_cg->converged;
        out->iterations = s->h_cg->it;
        out->final_residual = s->h_cg->res;
        out->nan_detected = s->h_cg->nan;
        out->refinements = 0;
    }
}

void gmres_solve_fp64(Context* c, pfem_solver_t* s, pfem_csr_t* a, double* x, const double* b, pfem_solve_stats* out)
{
    if (s->pc) precond_initialize(c, s->pc, a);
    if (factor_order_eligible(s)) {
        gmres_run<double>(c, s, args_factor_order(c, s, x, b), s->gm.p, s->h_gm);
        scatter_back(c, s, x);
    } else {
        gmres_run<double>(c, s, args64(s, a, x, b), s->gm.p, s->h_gm);
    }
    precond_check(c, s->pc);
    if (out) {
        out->converged = s->h_gm->converged;
        out->iterations = s->h_gm->it;
        out->final_residual = s->h_gm->res;
        out->nan_detected = s->h_gm->nan;
        out->refinements = 0;
    }
}

// FP64 iterative refinement around FP32 inner solves: r = b - A x and the stopping test in FP64
// (the same criterion as the FP64 solver), the correction A d = r / ||r|| solved in FP32 to the
// relative tolerance inner_tol, x += ||r|| d in FP64.
// FP64 iterative refinement around FP32 inner solves of one system: r = b - A x and the stopping test in
// FP64 (the same criterion as the FP64 solver), the correction A d = r / ||r|| solved in FP32 to the
// relative tolerance inner_tol, x += ||r|| d in FP64. Split into pieces so that two systems can refine in
// lockstep (solve_mixed_pair).
struct MixedSys {
    Context* c;
    pfem_solver_t* s;
    pfem_csr_t* a;
    double* x;
    const double* b;
    IrState* ist = nullptr;
    KArgs<float> A{};
    int total_inner = 0, refinements = 0, conv = 0, nan = 0;
    double res = 0.0;
    bool active = true;

    bool gm() const { return s->kind == PFEM_SOLVER_GMRES; }
    void start()
    {
        const Pattern& P = *a->pat;
        if (!a->val32.p) a->val32.alloc(P.nnz);
        hipLaunchKernelGGL(k_to32, dim3(blocks_for(P.nnz)), dim3(kBlock), 0, c->stream, (long)P.nnz, a->val32.p, a->val.p);
        if (s->pc) precond_initialize(c, s->pc, a);   // also refreshes the FP32 copies of the factor
        ist = reinterpret_cast<IrState*>(s->xr.p + 2 * (size_t)P.n);   // state lives after the two n-vectors
        IrState init{};
        init.first = 1;
        PFEM_HIP(hipMemcpyAsync(ist, &init, sizeof(IrState), hipMemcpyHostToDevice, c->stream));
        const int kind = s->pc ? s->pc->kind : PFEM_PRECOND_NONE;
        A.P = &P;
        A.val = a->val32.p;
        A.kind = kind;
        A.invd = kind == PFEM_PRECOND_JACOBI ? s->pc->invd32.p : nullptr;
        A.pc = s->pc;
        A.x = s->d32.p;
        A.b = s->r32.p;
        A.x_zero = true;
        A.tol = s->inner_tol;
        A.tol_mode = PFEM_TOL_RELATIVE;
        A.r = s->z32.p; A.z = s->p32.p; A.p = s->q32.p; A.q = s->t32.p; A.t = nullptr; A.V = s->V32.p;
    }
    // r = b - A x, its norms and the stopping test (launched; read with read_test)
    void residual_test()
    {
        const Pattern& P = *a->pat;
        const int n = P.n;
        const unsigned g = vec_grid(n);
        const int kind = A.kind;
        const bool jac = kind == PFEM_PRECOND_JACOBI, none = kind == PFEM_PRECOND_NONE;
        double* r = s->xr.p;
        double* z = s->xr.p + n;
        int nb0 = (int)g;
        if (jac || none) {
            nb0 = P.ntiles;
            const double* invd = jac ? s->pc->invd.p : nullptr;
            if (gm()) {
                if (jac) hipLaunchKernelGGL((k_ir_residual<true, true>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, a->val.p, x, b, invd, r, z, s->pa);
                else hipLaunchKernelGGL((k_ir_residual<false, true>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, a->val.p, x, b, invd, r, z, s->pa);
            } else {
                if (jac) hipLaunchKernelGGL((k_ir_residual<true, false>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, a->val.p, x, b, invd, r, z, s->pa);
                else hipLaunchKernelGGL((k_ir_residual<false, false>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, a->val.p, x, b, invd, r, z, s->pa);
            }
        } else {
            hipLaunchKernelGGL((k_residual<double>), dim3(P.ntiles), dim3(kBlock), 0, c->stream, P.tiles.p, P.rowptr.p, P.col.p, a->val.p, x, b, r);
            precond_apply(c, s->pc, z, r);
            const double* t = nullptr;
            if (refinements == 0 && s->tol_mode == PFEM_TOL_RELATIVE) {
                precond_apply(c, s->pc, s->t.p, b);
                t = s->t.p;
            }
            if (gm()) hipLaunchKernelGGL((k_ir_norms<true>), dim3(g), dim3(kBlock), 0, c->stream, n, r, z, b, t, s->pa);
            else hipLaunchKernelGGL((k_ir_norms<false>), dim3(g), dim3(kBlock), 0, c->stream, n, r, z, b, t, s->pa);
        }
        if (gm()) hipLaunchKernelGGL(k_ir_test<true>, dim3(1), dim3(kBlock), 0, c->stream, ist, (const double*)s->pa, nb0, s->tol, s->tol_mode);
        else hipLaunchKernelGGL(k_ir_test<false>, dim3(1), dim3(kBlock), 0, c->stream, ist, (const double*)s->pa, nb0, s->tol, s->tol_mode);
        PFEM_HIP(hipMemcpyAsync(s->h_cg, ist, sizeof(IrState), hipMemcpyDeviceToHost, c->stream));
    }
    // after the stream is synchronised: false when the refinement is over
    bool read_test()
    {
        IrState h{};
        std::memcpy(&h, s->h_cg, sizeof(IrState));
        res = gm() ? std::sqrt(h.rz) : std::sqrt(std::fabs(h.rz));
        if (h.converged) conv = 1;
        if (h.nan) nan = 1;
        if (conv || nan || refinements >= s->max_refine || total_inner >= s->maxit) active = false;
        return active;
    }
    // d's right-hand side r / ||r|| in FP32
    void down()
    {
        const int n = a->pat->n;
        hipLaunchKernelGGL(k_ir_down, dim3(vec_grid(n)), dim3(kBlock), 0, c->stream, n, s->r32.p, s->xr.p, ist);
        A.maxit = std::max(1, s->maxit - total_inner);
    }
    void up(int inner_its)
    {
        const int n = a->pat->n;
        total_inner += inner_its;
        hipLaunchKernelGGL(k_ir_up, dim3(vec_grid(n)), dim3(kBlock), 0, c->stream, n, x, s->d32.p, ist);
        ++refinements;
    }
    void finish(pfem_solve_stats* out)
    {
        precond_check(c, s->pc);
        if (out) {
            out->converged = conv;
            out->iterations = total_inner;
            out->final_residual = res;
            out->nan_detected = nan;
            out->refinements = refinements;
        }
    }
};

void solve_mixed(Context* c, pfem_solver_t* s, pfem_csr_t* a, double* x, const double* b, pfem_solve_stats* out)
{
    MixedSys m{c, s, a, x, b};
    m.start();
    while (true) {
        m.residual_test();
        PFEM_HIP(hipStreamSynchronize(c->stream));
        if (!m.read_test()) break;
        m.down();
        if (m.gm()) {
            gmres_run<float>(c, s, m.A, s->gm.p, s->h_gm);
            m.up(s->h_gm->it);
        } else {
            cg_run<float>(c, s, m.A, s->cg.p, s->h_cg);
            m.up(s->h_cg->it);
        }
    }
    m.finish(out);
}

// Two GMRES systems on one pattern (the two velocity-prediction components): FP64 (the Arnoldi steps in
// lockstep, gmres_run_pair), or mixed precision with both refinements in lockstep and their FP32 inner
// GMRES solves run as a pair.
static void gmres_solve_pair(Context* c, pfem_solver_t* s0, pfem_csr_t* a0, double* x0, const double* b0,
                             pfem_solver_t* s1, pfem_csr_t* a1, double* x1, const double* b1, pfem_solve_stats* out)
{
    if (s0->precision == PFEM_PRECISION_MIXED) {
        MixedSys m[2] = {{c, s0, a0, x0, b0}, {c, s1, a1, x1, b1}};
        for (auto& q : m) q.start();
        while (m[0].active || m[1].active) {
            for (auto& q : m)
                if (q.active) q.residual_test();
            PFEM_HIP(hipStreamSynchronize(c->stream));
            bool on[2];
            for (int k = 0; k < 2; ++k) on[k] = m[k].active && m[k].read_test();
            if (!on[0] && !on[1]) break;
            for (int k = 0; k < 2; ++k)
                if (on[k]) m[k].down();
            if (on[0] && on[1]) {
                GmresSys<float> g0{c, s0, m[0].A, s0->gm.p, s0->h_gm}, g1{c, s1, m[1].A, s1->gm.p, s1->h_gm};
                gmres_run_pair<float>(c, g0, g1);
            } else {
                const int k = on[0] ? 0 : 1;
                gmres_run<float>(c, m[k].s, m[k].A, m[k].s->gm.p, m[k].s->h_gm);
            }
            for (int k = 0; k < 2; ++k)
                if (on[k]) m[k].up(m[k].s->h_gm->it);
        }
        m[0].finish(out);
        m[1].finish(out ? out + 1 : nullptr);
        return;
    }
    for (pfem_solver_t* s : {s0, s1})
        if (s->pc) precond_initialize(c, s->pc, s == s0 ? a0 : a1);
    GmresSys<double> g0{c, s0, args64(s0, a0, x0, b0), s0->gm.p, s0->h_gm};
    GmresSys<double> g1{c, s1, args64(s1, a1, x1, b1), s1->gm.p, s1->h_gm};
    gmres_run_pair<double>(c, g0, g1);
    for (int k = 0; k < 2; ++k) {
        pfem_solver_t* s = k ? s1 : s0;
        precond_check(c, s->pc);
        if (out) {
            out[k].converged = s->h_gm->converged;
            out[k].iterations = s->h_gm->it;
            out[k].final_residual = s->h_gm->res;
            out[k].nan_detected = s->h_gm->nan;
            out[k].refinements = 0;
        }
    }
}

void precond_enable_fp32(Context* c, pfem_precond_t* pc);

static void alloc_mixed(pfem_solver_t* s)
{
    const int n = s->n;
    if (s->r32.p) return;
    s->r32.alloc(n); s->z32.alloc(n); s->p32.alloc(n); s->q32.alloc(n); s->d32.alloc(n); s->t32.alloc(n);
    if (s->kind == PFEM_SOLVER_GMRES) s->V32.alloc((size_t)n * (s->m + 1));
    // two FP64 n-vectors followed by the refinement state
    s->xr.alloc(2 * (size_t)n + (sizeof(IrState) + sizeof(double) - 1) / sizeof(double) + 1);
    PFEM_HIP(hipMemset(s->xr.p, 0, s->xr.bytes()));
}

}  // namespace pfem

using namespace pfem;

extern "C" {

int pfem_solver_create(pfem_csr a, int kind, double tol, int max_iterations, int restart, int tol_mode,
                       pfem_precond precond, pfem_solver* out)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a && out, "null argument");
        PFEM_REQUIRE(kind == PFEM_SOLVER_CG || kind == PFEM_SOLVER_GMRES, "unknown solver kind");
        PFEM_REQUIRE(max_iterations > 0, "max_iterations must be positive");
        auto s = std::make_unique<pfem_solver_t>();
        s->ctx = a->ctx;
        s->kind = kind;
        s->tol = tol;
        s->maxit = max_iterations;
        s->tol_mode = tol_mode;
        s->pc = precond;
        const int n = a->pat->n;
        s->n = n;
        if (precond) PFEM_REQUIRE(precond->n == n, "preconditioner size differs from the matrix");
        if (kind == PFEM_SOLVER_GMRES) {
            // SolverGMRES::init clips maxIterations to the matrix size (linear_solver.cu:398-401)
            if (s->maxit > n) s->maxit = n;
            s->restart = restart <= 0 ? s->maxit : restart;
            s->m = std::min(s->restart, s->maxit);
            s->restarted = s->restart < s->maxit ? 1 : 0;
            PFEM_REQUIRE(s->m <= kMaxRestart, "GMRES cycle longer than 1024 vectors");
            s->V.alloc((size_t)n * (s->m + 1));
            s->H.alloc((size_t)s->m * s->m);
            s->beta.alloc(s->m + 1);
            s->cs.alloc(s->m);
            s->sn.alloc(s->m);
            s->y.alloc(s->m);
            s->hc.alloc(s->m + 1);
            s->gm.alloc(1);
            PFEM_HIP(hipMemset(s->gm.p, 0, sizeof(GmState)));
            PFEM_HIP(hipHostMalloc((void**)&s->h_gm, sizeof(GmState), hipHostMallocDefault));
            s->poll = 4;
        } else {
            s->r.alloc(n); s->z.alloc(n); s->p.alloc(n); s->q.alloc(n);
            s->poll = 8;
        }
        s->t.alloc(n);
        s->cg.alloc(1);
        PFEM_HIP(hipMemset(s->cg.p, 0, sizeof(CgState)));
        PFEM_HIP(hipHostMalloc((void**)&s->h_cg, 2 * std::max(sizeof(CgState), (size_t)128), hipHostMallocDefault));
        for (hipEvent_t& e : s->ev) PFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        int tiles = a->pat->ntiles;
        if (precond && precond->fpat) tiles = std::max(tiles, precond->fpat->ntiles);
        const size_t na = 2 * (size_t)std::max(tiles, kVecGrid);
        // pb also takes the per-chunk (r, z) partials of the single-pass sweeps (n/256 + one per level)
        const size_t nb = std::max<size_t>((size_t)std::max(2, s->m + 1) * kVecGrid, blocks_for(n) + 4096);
        s->partials.alloc(na + nb + kVecGrid);
        s->pa = s->partials.p;
        s->pb = s->pa + na;
        s->pn = s->pb + nb;
        s->pb_cap = (int)nb;
        s->pa_cap = (int)na;
        *out = s.release();
    });
}

int pfem_solver_set_precision(pfem_solver s, int precision, double inner_tol, int max_refinements)
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s, "null solver");
        PFEM_REQUIRE(precision == PFEM_PRECISION_FP64 || precision == PFEM_PRECISION_MIXED ||
                         precision == PFEM_PRECISION_MIXED_PRECOND,
                     "unknown precision");
        if (precision == PFEM_PRECISION_MIXED_PRECOND)
            PFEM_REQUIRE(s->pc && (s->pc->kind == PFEM_PRECOND_IC0 || s->pc->kind == PFEM_PRECOND_ILU0),
                         "PFEM_PRECISION_MIXED_PRECOND needs an IC0 or ILU0 preconditioner");
        s->precision = precision;
        if (inner_tol > 0) s->inner_tol = inner_tol;
        if (max_refinements > 0) s->max_refine = max_refinements;
        if (precision == PFEM_PRECISION_MIXED_PRECOND) {
            precond_enable_fp32(C(s->ctx), s->pc);
            s->pc->sweep = 1;
        } else if (precision == PFEM_PRECISION_MIXED) {
            PFEM_REQUIRE(!s->pc || s->pc->kind != PFEM_PRECOND_CUSTOM,
                         "a custom preconditioner applies to FP64 vectors only");
            if (s->kind == PFEM_SOLVER_GMRES && !s->gm.p) {
                s->gm.alloc(1);
                PFEM_HIP(hipMemset(s->gm.p, 0, sizeof(GmState)));
            }
            alloc_mixed(s);
            if (s->pc) precond_enable_fp32(C(s->ctx), s->pc);
        }
    });
}

int pfem_solver_set_option(pfem_solver s, int option, int value)
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s, "null solver");
        switch (option) {
        case PFEM_SOLVER_OPT_FACTOR_ORDER: s->factor_order = value ? 1 : 0; break;
        case PFEM_SOLVER_OPT_FUSE:
            PFEM_REQUIRE(value >= 0 && value <= 3, "fusion level out of range");
            s->fuse = value;
            break;
        case PFEM_SOLVER_OPT_REPLACE:
            PFEM_REQUIRE(value >= 1 && value <= 8, "replacement interval out of range");
            s->replace_digits = value;
            break;
        case PFEM_SOLVER_OPT_GRAPH: s->graph = value ? 1 : 0; break;
        case PFEM_SOLVER_OPT_GS_PASSES:
            PFEM_REQUIRE(value == 1 || value == 2, "Gram-Schmidt passes must be 1 or 2");
            s->reorth = value == 2 ? 1 : 0;
            break;
        case PFEM_SOLVER_OPT_POLL:
            PFEM_REQUIRE(value >= 1 && value <= 4096, "poll interval out of range");
            s->poll = value;
            break;
        default: throw Error(PFEM_EINVAL, "unknown solver option");
        }
    });
}

// ops 10 and 11: the fused PCG iteration as a solve runs it (FP64 factor, or the reduced-precision factor
// copy of PFEM_PRECISION_MIXED_PRECOND), from x = 0 with b = 1 and a zero tolerance, so that every
// iteration runs; the device time of a solve of 2L iterations minus that of L iterations, over L, on the
// context stream (the solve's set-up, factorization included, cancels).
static double time_pcg_iteration(Context* c, pfem_csr a, pfem_precond pc, bool reduced, int launches)
{
    const int n = a->pat->n;
    pfem_solver s = nullptr;
    if (pfem_solver_create(a, PFEM_SOLVER_CG, 0.0, launches, 0, PFEM_TOL_ABSOLUTE, pc, &s) != PFEM_OK)
        throw Error(PFEM_EINVAL, c->err);
    std::unique_ptr<pfem_solver_t, int (*)(pfem_solver)> guard(s, pfem_solver_destroy);
    if (reduced && pfem_solver_set_precision(s, PFEM_PRECISION_MIXED_PRECOND, 0.0, 0) != PFEM_OK)
        throw Error(PFEM_EINVAL, c->err);
    DevArray<double> x(n), b(n);
    launch_fill(c, b.p, 1.0, n);
    hipEvent_t e0, e1;
    PFEM_HIP(hipEventCreate(&e0));
    PFEM_HIP(hipEventCreate(&e1));
    float ms[2] = {0.f, 0.f};
    for (int k = 0; k < 2; ++k) {
        s->maxit = launches * (k + 1);
        PFEM_HIP(hipEventRecord(e0, c->stream));
        const int rc = pfem_solver_solve(s, a, x.p, b.p, nullptr);
        PFEM_HIP(hipEventRecord(e1, c->stream));
        PFEM_HIP(hipEventSynchronize(e1));
        if (rc != PFEM_OK) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw Error(rc, c->err);
        }
        PFEM_HIP(hipEventElapsedTime(&ms[k], e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (ms[1] - ms[0]) / launches;
}

int pfem_time_kernel(int op, pfem_csr a, pfem_precond pc, int launches, double* avg_ms)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a && avg_ms && launches > 0 && op >= 0 && op <= 11, "bad arguments");
        if (op >= 10) {
            PFEM_REQUIRE(pc && (pc->kind == PFEM_PRECOND_IC0 || pc->kind == PFEM_PRECOND_ILU0),
                         "ops 10 and 11 need an IC0 or ILU0 preconditioner");
            *avg_ms = time_pcg_iteration(C(a->ctx), a, pc, op == 11, launches);
            return PFEM_OK;
        }
        PFEM_REQUIRE(op < 5 || (pc && pc->ordering != PFEM_ORDER_NATURAL && pc->sweep),
                     "ops 5 to 9 need a multicolour preconditioner with single-pass sweeps");
        PFEM_REQUIRE(op < 8 || pc->sell, "ops 8 and 9 need the sliced sweep layout");
        PFEM_REQUIRE(op < 2 || pc, "ops 2 and 3 need a preconditioner");
        Context* c = C(a->ctx);
        const Pattern& P = *a->pat;
        const int n = P.n;
        DevArray<double> x(n), y(n);
        DevArray<float> x32(n), y32(n);
        launch_fill(c, x.p, 1.0, n);
        hipLaunchKernelGGL(k_to32, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, (long)n, x32.p, x.p);
        if (op == 1) {
            if (!a->val32.p) a->val32.alloc(P.nnz);
            hipLaunchKernelGGL(k_to32, dim3(blocks_for(P.nnz)), dim3(kBlock), 0, c->stream, (long)P.nnz, a->val32.p, a->val.p);
        }
        if (op >= 2) {
            if (op == 3 || op == 4 || op == 6 || op == 8) precond_enable_fp32(c, pc);
            precond_initialize(c, pc, a);
        }
        ApplyOpts<double> tag_opts{nullptr, 0, true};
        tag_opts.tagged = true;
        if (op >= 8)   // the tagged form's output buffer carries tag 0 between applies
            hipLaunchKernelGGL((k_fill_tag0<double>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, (long)n, y.p);
        auto launch = [&]() {
            switch (op) {
            case 0: launch_spmv(c, P, a->val.p, x.p, y.p); break;
            case 1: launch_spmv32(c, P, a->val32.p, x32.p, y32.p); break;
            case 2: precond_apply(c, pc, y.p, x.p); break;
            case 3: precond_apply32(c, pc, y32.p, x32.p); break;
            case 4: precond_apply_t<double>(c, pc, y.p, x.p, true, ApplyOpts<double>{}); break;
            case 5: precond_apply_t<double>(c, pc, y.p, x.p, false, ApplyOpts<double>{nullptr, 0, true}); break;
            case 6: precond_apply_t<double>(c, pc, y.p, x.p, true, ApplyOpts<double>{nullptr, 0, true}); break;
            case 7: launch_spmv(c, *pc->fpat, pc->aval.p, x.p, y.p); break;
            default: precond_apply_t<double>(c, pc, y.p, x.p, op == 8, tag_opts); break;
            }
        };
        launch();   // warm-up
        hipEvent_t e0, e1;
        PFEM_HIP(hipEventCreate(&e0));
        PFEM_HIP(hipEventCreate(&e1));
        PFEM_HIP(hipEventRecord(e0, c->stream));
        for (int k = 0; k < launches; ++k) launch();
        PFEM_HIP(hipEventRecord(e1, c->stream));
        PFEM_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_ms = ms / launches;
    });
}

int pfem_solver_destroy(pfem_solver s)
{
    if (!s) return PFEM_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->h_cg) (void)hipHostFree(s->h_cg);
    if (s->h_gm) (void)hipHostFree(s->h_gm);
    for (hipEvent_t e : s->ev)
        if (e) (void)hipEventDestroy(e);
    delete s;
    return PFEM_OK;
}

int pfem_solver_solve_pair(pfem_solver s0, pfem_csr a0, double* d_x0, const double* d_b0, pfem_solver s1,
                           pfem_csr a1, double* d_x1, const double* d_b1, pfem_solve_stats stats[2])
{
    PFEM_TRY(s0 ? s0->ctx : nullptr, {
        PFEM_REQUIRE(s0 && s1 && a0 && a1 && d_x0 && d_b0 && d_x1 && d_b1, "null argument");
        PFEM_REQUIRE(s0 != s1, "the pair needs two solver objects (each keeps its own Krylov state)");
        PFEM_REQUIRE(s0->ctx == s1->ctx && a0->ctx == s0->ctx && a1->ctx == s0->ctx, "objects of different contexts");
        PFEM_REQUIRE(a0->pat->n == s0->n && a1->pat->n == s1->n, "matrix size differs from the solver's");
        Context* c = C(s0->ctx);
        pfem_solve_stats local[2] = {};
        const bool pairable = s0->kind == PFEM_SOLVER_GMRES && s1->kind == PFEM_SOLVER_GMRES &&
                              s0->precision == s1->precision && s0->precision != PFEM_PRECISION_MIXED_PRECOND &&
                              s0->m == s1->m && a0->pat.get() == a1->pat.get() &&
                              (!s0->pc || s0->pc->kind == PFEM_PRECOND_JACOBI || s0->pc->kind == PFEM_PRECOND_NONE) &&
                              (!s1->pc || s1->pc->kind == PFEM_PRECOND_JACOBI || s1->pc->kind == PFEM_PRECOND_NONE);
        if (pairable) {
            gmres_solve_pair(c, s0, a0, d_x0, d_b0, s1, a1, d_x1, d_b1, local);
        } else {
            // other combinations: the two solves one after the other
            const int rc0 = pfem_solver_solve(s0, a0, d_x0, d_b0, &local[0]);
            if (rc0 != PFEM_OK) throw Error(rc0, c->err);
            const int rc1 = pfem_solver_solve(s1, a1, d_x1, d_b1, &local[1]);
            if (rc1 != PFEM_OK) throw Error(rc1, c->err);
        }
        if (stats) {
            stats[0] = local[0];
            stats[1] = local[1];
        }
    });
}

int pfem_solver_solve(pfem_solver s, pfem_csr a, double* d_x, const double* d_b, pfem_solve_stats* stats)
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s && a && d_x && d_b, "null argument");
        PFEM_REQUIRE(a->pat->n == s->n, "matrix size differs from the solver's");
        Context* c = C(s->ctx);
        pfem_solve_stats local{};
        if (s->precision == PFEM_PRECISION_MIXED) solve_mixed(c, s, a, d_x, d_b, &local);
        else if (s->kind == PFEM_SOLVER_CG) cg_solve_fp64(c, s, a, d_x, d_b, &local);
        else gmres_solve_fp64(c, s, a, d_x, d_b, &local);
        if (stats) *stats = local;
    });
}

}  // extern "C"
