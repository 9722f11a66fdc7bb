// Multi-right-hand-side PCG: A X = B for up to 8 right-hand sides sharing one matrix, in one sequence of
// launches. The reference solves one system per call (SolverCG::solve, src/linear_algebra/
// linear_solver.cu:301-376); each column here follows that recurrence (Jacobi or no preconditioner,
// the same stopping test on (r, M^-1 r)) with its own scalars and its own convergence, so a column's
// iterates equal a single-RHS solve's up to the order of the dot-product reductions.
//
// Layout: the solver's work vectors are interleaved, v[i * G + j] = component i of column j, with G the
// column count rounded up to 1, 2, 4 or 8. A row is handled by a group of G consecutive lanes, lane j of
// the group taking column j: the group reads the row's values and column indices once (one broadcast
// load per entry for all columns) and gathers G contiguous values per column index, so the matrix stream
// is shared by the columns and the gathers are G-wide. Dot products are reduced per column in a fixed
// order (deterministic). The caller's X and B are column-major (column j at X + j * n).
//
// PFEM_PRECISION_MIXED: PCG with reliable updates (residual replacement). The iterations run on FP32
// vectors and the FP32 copy of A (half the bytes); the solution accumulates in FP64. When a column's
// recursive residual has fallen by a set factor (10x by default) since the last replacement, or passes
// the stopping test, the
// column spends one iteration slot on a replacement: x += x_s in FP64, r = b - A x with the FP64 matrix,
// the stopping test on that true residual, and the search direction continued from it (beta = new
// gamma / old gamma). Unlike restarted refinement, the Krylov recurrence is never restarted, so the
// iteration count stays close to the FP64 solver's, and convergence is decided on FP64 residuals.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "solver.hpp"

namespace pfem {

constexpr int kMaxRhs = 8;
constexpr int kMGrid = 2048;   // grid of the multi-RHS kernels (grid-stride over row groups)
static_assert(kMGrid % kBlock == 0, "partial series are read in whole blocks");

struct MState {
    double tol;
    int tol_mode, maxit, ncol, all_done;
    double gam[2][kMaxRhs], alpha[kMaxRhs], beta[kMaxRhs], thr[kMaxRhs], res[kMaxRhs];
    int it[kMaxRhs], done[kMaxRhs], converged[kMaxRhs], nan[kMaxRhs];
    // reliable updates: gamma of the current residual, gamma at the last replacement, the slot state
    // (0 iterating, 1 replacement due, 2 direction update after a replacement) and the replacements made
    double rz[kMaxRhs], rzmax[kMaxRhs], drop;
    int mode[kMaxRhs], nrep[kMaxRhs];
};


// Per-column block sums of NV values: lane l holds column l % G. Reduced over the wave with shuffles that
// keep the column (offsets 32 .. G), then over the four waves in order; part[(i * G + j) * gridDim.x +
// blockIdx.x] = value i of column j.
template <int G, int NV>
__device__ __forceinline__ void group_partials(double (&v)[NV], double* __restrict__ part)
{
    __shared__ double sh[NV][kWaves][G];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
#pragma unroll
        for (int off = 32; off >= G; off >>= 1) v[i] += __shfl_down(v[i], off, 64);
        if (lane < G) sh[i][wid][lane] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < G) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = sh[i][0][threadIdx.x];
            for (int w = 1; w < kWaves; ++w) s += sh[i][w][threadIdx.x];
            part[((size_t)i * G + threadIdx.x) * gridDim.x + blockIdx.x] = s;
        }
    }
}

// Totals of NJ partial series of nb <= kMGrid terms each (series j at part + j * nb: the grid of the
// kernel that wrote them), valid in thread 0. Every load of a lane is issued before the first addition
// (one memory round trip); lane t adds terms t, t + 256, ... of each series in order, then block_sum.
template <int NJ>
__device__ __forceinline__ void series_totals(const double* __restrict__ part, int nb, double (&tot)[NJ])
{
    constexpr int kPer = kMGrid / kBlock;
    double t[NJ][kPer];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = threadIdx.x + u * kBlock;
            t[j][u] = k < nb ? part[(size_t)j * nb + k] : 0.0;
        }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) s += t[j][u];
        tot[j] = s;
    }
    block_sum<NJ>(tot);
}

// ---- FP64 / FP32 multi-RHS CG kernels (T = vector type, VT = matrix value type)
constexpr int kMTileNnz = 1024;   // non-zeros of a row tile staged in LDS
constexpr int kMRows = 2;         // rows per lane in a row tile

// Row tiles of the multi-RHS products: tile t = rows [tiles[2t], tiles[2t+1]), at most kMRows * kBlock / G
// rows and at most kMTileNnz non-zeros unless a single row is longer. Lane (q, j) of a workgroup takes
// column j of rows q and q + kBlock / G of the tile. Workgroup b takes a contiguous range of tiles; the
// ranges of XCD x (workgroups b = x mod 8) form the x-th band of rows, so neighbouring rows share their
// gathered vector entries in one L2. A tile's values, column indices and row offsets are staged in LDS
// with unit-stride loads while the next tile's bounds are fetched; each lane then adds val * x[col * G + j]
// over its rows left to right (the sequential order), with up to 8 gathers per row in flight, and hands
// each sum to f(row, j, s). The G lanes of a row gather G contiguous values.
template <int G, typename T, typename VT, typename F>
__device__ __forceinline__ void for_tile_rows(int ntiles, const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                              const int* __restrict__ col, const VT* __restrict__ val,
                                              const T* __restrict__ x, F f)
{
    constexpr int kLanesRows = kBlock / G;
    constexpr int kTileRows = kMRows * kLanesRows;
    constexpr int kU = 8;
    __shared__ VT sval[kMTileNnz];
    __shared__ int scol[kMTileNnz];
    __shared__ int rp[kTileRows + 1];
    const int nb = gridDim.x, b = blockIdx.x;
    const int rb = (b % 8) * (nb / 8) + b / 8;
    const int t0 = (int)((long)ntiles * rb / nb), t1 = (int)((long)ntiles * (rb + 1) / nb);
    const int j = threadIdx.x % G, q = threadIdx.x / G;
    int r0 = 0, r1 = 0, k0 = 0, k1 = 0;
    if (t0 < t1) {
        r0 = tiles[2 * t0];
        r1 = tiles[2 * t0 + 1];
        k0 = rowptr[r0];
        k1 = rowptr[r1];
    }
    for (int t = t0; t < t1; ++t) {
        const int nk = k1 - k0;
        const bool staged = nk <= kMTileNnz;
        __syncthreads();   // the previous tile's readers are done with the LDS arrays
        for (int i = threadIdx.x; i <= r1 - r0; i += kBlock) rp[i] = rowptr[r0 + i] - k0;
        if (staged)
            for (int k = threadIdx.x; k < nk; k += kBlock) {
                sval[k] = val[k0 + k];
                scol[k] = col[k0 + k];
            }
        // the next tile's bounds, fetched while this one is computed
        int n0 = 0, n1 = 0, m0 = 0, m1 = 0;
        if (t + 1 < t1) {
            n0 = tiles[2 * t + 2];
            n1 = tiles[2 * t + 3];
            m0 = rowptr[n0];
            m1 = rowptr[n1];
        }
        __syncthreads();
        T s[kMRows];
#pragma unroll
        for (int u = 0; u < kMRows; ++u) s[u] = 0;
        if (staged) {
            int ka[kMRows], ke[kMRows];
#pragma unroll
            for (int u = 0; u < kMRows; ++u) {
                const int lr = q + u * kLanesRows;
                ka[u] = lr < r1 - r0 ? rp[lr] : 0;
                ke[u] = lr < r1 - r0 ? rp[lr + 1] : 0;
            }
            int len = 0;
#pragma unroll
            for (int u = 0; u < kMRows; ++u) len = max(len, ke[u] - ka[u]);
            for (int d = 0; d < len; d += kU) {
                T xv[kMRows][kU], vv[kMRows][kU];
#pragma unroll
                for (int u = 0; u < kMRows; ++u)
#pragma unroll
                    for (int w = 0; w < kU; ++w) {
                        const int k = ka[u] + d + w;
                        if (k < ke[u]) {
                            vv[u][w] = (T)sval[k];
                            xv[u][w] = x[(size_t)scol[k] * G + j];
                        }
                    }
#pragma unroll
                for (int u = 0; u < kMRows; ++u)
#pragma unroll
                    for (int w = 0; w < kU; ++w)
                        if (ka[u] + d + w < ke[u]) s[u] += vv[u][w] * xv[u][w];
            }
        } else {
#pragma unroll
            for (int u = 0; u < kMRows; ++u) {
                const int row = r0 + q + u * kLanesRows;
                if (row < r1)
                    for (int k = rowptr[row]; k < rowptr[row + 1]; ++k) s[u] += (T)val[k] * x[(size_t)col[k] * G + j];
            }
        }
#pragma unroll
        for (int u = 0; u < kMRows; ++u) {
            const int row = r0 + q + u * kLanesRows;
            if (row < r1) f(row, j, s[u]);
        }
        r0 = n0;
        r1 = n1;
        k0 = m0;
        k1 = m1;
    }
}

// r = b - A x (or b when x = 0), p = z = M^-1 r; partials (r, z) and (b, M^-1 b) per column
template <int G, typename T, typename VT, bool JAC>
__global__ __launch_bounds__(kBlock) void k_m_init(int n, int ntiles, const int* __restrict__ tiles,
                                                   const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const VT* __restrict__ val, const T* __restrict__ x,
                                                   const T* __restrict__ b, const T* __restrict__ invd, T* __restrict__ r,
                                                   T* __restrict__ p, double* __restrict__ part, int x_zero)
{
    double v[2] = {0.0, 0.0};
    auto row_init = [&](int i, int j, T ax) {
        const size_t o = (size_t)i * G + j;
        const T bb = b[o];
        const T rr = bb - ax;
        const T zz = JAC ? rr * invd[i] : rr;
        const T mb = JAC ? bb * invd[i] : bb;
        r[o] = rr;
        p[o] = zz;
        v[0] += (double)(rr * zz);
        v[1] += (double)(bb * mb);
    };
    if (x_zero) {
        const int j = threadIdx.x % G;
        for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n; i += gridDim.x * (kBlock / G))
            row_init(i, j, (T)0);
    } else {
        for_tile_rows<G, T, VT>(ntiles, tiles, rowptr, col, val, x, row_init);
    }
    group_partials<G, 2>(v, part);
}

// gamma_0, the per-column thresholds (linear_solver.cu:360; relative to (b, M^-1 b) on request, absolute
// for a zero column) and the columns already converged; one block
template <int G>
__global__ void k_m_begin(MState* st, const double* __restrict__ part, int nb)
{
    double tot[2 * G];
    series_totals<2 * G>(part, nb, tot);
    if (threadIdx.x != 0) return;
    int all = 1;
    for (int j = 0; j < G; ++j) {
        const double rz = tot[j], bmb = tot[G + j];
        st->gam[0][j] = rz;
        st->thr[j] = st->tol_mode == PFEM_TOL_RELATIVE && bmb > 0.0 ? st->tol * st->tol * bmb : st->tol * st->tol;
        st->it[j] = 0;
        st->nan[j] = rz != rz;
        st->converged[j] = rz < st->thr[j] ? 1 : 0;
        st->done[j] = j >= st->ncol || st->converged[j] || st->nan[j];
        st->res[j] = sqrt(fabs(rz));
        all &= st->done[j];
    }
    st->all_done = all;
}

// q = A p; partials (p, q) per column
template <int G, typename T, typename VT>
__global__ __launch_bounds__(kBlock) void k_m_spmv(int ntiles, const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                   const int* __restrict__ col, const VT* __restrict__ val,
                                                   const T* __restrict__ p, T* __restrict__ q, const MState* st,
                                                   double* __restrict__ part)
{
    if (st->all_done) return;
    double v[1] = {0.0};
    for_tile_rows<G, T, VT>(ntiles, tiles, rowptr, col, val, p, [&](int i, int j, T ap) {
        const size_t o = (size_t)i * G + j;
        q[o] = ap;
        v[0] += (double)(p[o] * ap);
    });
    group_partials<G, 1>(v, part);
}

// alpha_j = gamma_j / (p_j, q_j) for the active columns; one block
template <int G>
__global__ void k_m_alpha(MState* st, const double* __restrict__ part, int nb, int k)
{
    if (st->all_done) return;
    double tot[G];
    series_totals<G>(part, nb, tot);
    if (threadIdx.x != 0) return;
    for (int j = 0; j < G; ++j) st->alpha[j] = st->done[j] ? 0.0 : st->gam[k & 1][j] / tot[j];
}

// x += alpha p, r -= alpha q (updateXR, linear_solver.cu:68-77) for the active columns; partials (r, M^-1 r)
template <int G, typename T, bool JAC>
__global__ void k_m_update(int n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ q,
                           const T* __restrict__ invd, const MState* st, double* __restrict__ part)
{
    if (st->all_done) return;
    const int j = threadIdx.x % G;
    const bool active = !st->done[j];
    const T alpha = (T)st->alpha[j];
    double v[1] = {0.0};
    for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n; i += gridDim.x * (kBlock / G)) {
        if (!active) continue;
        const size_t o = (size_t)i * G + j;
        x[o] += alpha * p[o];
        const T rr = r[o] - alpha * q[o];
        r[o] = rr;
        const T zz = JAC ? rr * invd[i] : rr;
        v[0] += (double)(rr * zz);
    }
    group_partials<G, 1>(v, part);
}

// gamma_{k+1}, beta and the stopping test per active column; one block
template <int G>
__global__ void k_m_beta(MState* st, const double* __restrict__ part, int nb, int k)
{
    if (st->all_done) return;
    double tot[G];
    series_totals<G>(part, nb, tot);
    if (threadIdx.x != 0) return;
    int all = 1;
    for (int j = 0; j < G; ++j) {
        if (!st->done[j]) {
            const double gnew = tot[j];
            st->beta[j] = gnew / st->gam[k & 1][j];
            st->gam[(k + 1) & 1][j] = gnew;
            st->it[j] = k + 1;
            st->res[j] = sqrt(fabs(gnew));
            if (gnew != gnew) {
                st->nan[j] = 1;
                st->done[j] = 1;
            } else if (gnew < st->thr[j]) {
                st->converged[j] = 1;
                st->done[j] = 1;
            } else if (k + 1 >= st->maxit) {
                st->done[j] = 1;
            }
            // a column that stops here keeps its p: it is not used again
            if (st->done[j]) st->beta[j] = 0.0;
        }
        all &= st->done[j];
    }
    st->all_done = all;
}

// p = M^-1 r + beta p (updateP, linear_solver.cu:79-87) for the columns still active
template <int G, typename T, bool JAC>
__global__ void k_m_p(int n, T* __restrict__ p, const T* __restrict__ r, const T* __restrict__ invd, const MState* st)
{
    if (st->all_done) return;
    const int j = threadIdx.x % G;
    if (st->done[j]) return;
    const T beta = (T)st->beta[j];
    for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n; i += gridDim.x * (kBlock / G)) {
        const size_t o = (size_t)i * G + j;
        const T zz = JAC ? r[o] * invd[i] : r[o];
        p[o] = zz + beta * p[o];
    }
}

// ---- layout changes and refinement kernels
// column-major (column j at s + j * n, j < ncol) -> interleaved (columns ncol..G-1 zero)
template <int G, typename T>
__global__ void k_m_pack(int n, int ncol, const double* __restrict__ s, T* __restrict__ d)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)n * G; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t / G), j = (int)(t % G);
        d[t] = j < ncol ? (T)s[(size_t)j * n + i] : (T)0;
    }
}

template <int G>
__global__ void k_m_unpack(int n, int ncol, const double* __restrict__ s, double* __restrict__ d)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)n * G; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t / G), j = (int)(t % G);
        if (j < ncol) d[(size_t)j * n + i] = s[t];
    }
}

// ---- reliable-update PCG (PFEM_PRECISION_MIXED)
// r = b - A x (FP64), r_s = (float) r, p_s = D_s r_s, x_s = 0; partials (r, D r) and (b, D b) in FP64
template <int G, bool JAC>
__global__ __launch_bounds__(kBlock) void k_mr_init(int ntiles, const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const double* __restrict__ val,
                                                    const double* __restrict__ x, const double* __restrict__ b,
                                                    const double* __restrict__ invd, const float* __restrict__ invd32,
                                                    float* __restrict__ rs, float* __restrict__ ps, float* __restrict__ xs,
                                                    double* __restrict__ part)
{
    double v[2] = {0.0, 0.0};
    for_tile_rows<G, double, double>(ntiles, tiles, rowptr, col, val, x, [&](int i, int j, double ax) {
        const size_t o = (size_t)i * G + j;
        const double bb = b[o];
        const double rr = bb - ax;
        const float r32 = (float)rr;
        rs[o] = r32;
        ps[o] = JAC ? r32 * invd32[i] : r32;
        xs[o] = 0.0f;
        const double d = JAC ? invd[i] : 1.0;
        v[0] += rr * (rr * d);
        v[1] += bb * (bb * d);
    });
    group_partials<G, 2>(v, part);
}

template <int G>
__global__ void k_mr_begin(MState* st, const double* __restrict__ part, int nb)
{
    double tot[2 * G];
    series_totals<2 * G>(part, nb, tot);
    if (threadIdx.x != 0) return;
    int all = 1;
    for (int j = 0; j < G; ++j) {
        const double rz = tot[j], bmb = tot[G + j];
        st->thr[j] = st->tol_mode == PFEM_TOL_RELATIVE && bmb > 0.0 ? st->tol * st->tol * bmb : st->tol * st->tol;
        st->rz[j] = st->rzmax[j] = rz;
        st->res[j] = sqrt(fabs(rz));
        st->it[j] = 0;
        st->mode[j] = 0;
        st->nrep[j] = 0;
        st->nan[j] = rz != rz;
        st->converged[j] = rz < st->thr[j] ? 1 : 0;
        st->done[j] = j >= st->ncol || st->converged[j] || st->nan[j];
        all &= st->done[j];
    }
    st->all_done = all;
}

// Iterating columns: q_s = A_s p_s, partial (p_s, q_s). Columns due a replacement: r = b - A x with the
// FP64 matrix, r_s = (float) r, partial (r, D r). The tile stages FP64 values only while some column is
// being replaced ((float) of the FP64 value is the FP32 copy's value, so the FP32 products do not change).
template <int G, bool JAC>
__global__ __launch_bounds__(kBlock) void k_mr_spmv(int ntiles, const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const float* __restrict__ val32,
                                                    const double* __restrict__ val, const float* __restrict__ ps,
                                                    float* __restrict__ qs, const double* __restrict__ x,
                                                    const double* __restrict__ b, const double* __restrict__ invd,
                                                    float* __restrict__ rs, const MState* st, double* __restrict__ part)
{
    if (st->all_done) return;
    int anyrep = 0;
    for (int c = 0; c < G; ++c) anyrep |= st->mode[c] == 1 && !st->done[c];
    double v[1] = {0.0};
    if (!anyrep) {
        for_tile_rows<G, float, float>(ntiles, tiles, rowptr, col, val32, ps, [&](int i, int j, float ap) {
            const size_t o = (size_t)i * G + j;
            qs[o] = ap;
            v[0] += (double)(ps[o] * ap);
        });
    } else {
        const int j = threadIdx.x % G;
        const bool rep = st->mode[j] == 1;
        __shared__ double sval[kMTileNnz];
        __shared__ int scol[kMTileNnz];
        const int nb = gridDim.x, bk = blockIdx.x;
        const int rb = (bk % 8) * (nb / 8) + bk / 8;
        const int t0 = (int)((long)ntiles * rb / nb), t1 = (int)((long)ntiles * (rb + 1) / nb);
        for (int t = t0; t < t1; ++t) {
            const int r0 = tiles[2 * t], r1 = tiles[2 * t + 1];
            const int k0 = rowptr[r0], nk = rowptr[r1] - k0;
            const bool staged = nk <= kMTileNnz;
            if (staged) {
                __syncthreads();
                for (int k = threadIdx.x; k < nk; k += kBlock) {
                    sval[k] = val[k0 + k];
                    scol[k] = col[k0 + k];
                }
                __syncthreads();
            }
            for (int u = 0; u < kMRows; ++u) {
                const int row = r0 + (int)threadIdx.x / G + u * (kBlock / G);
                if (row >= r1) continue;
                const size_t o = (size_t)row * G + j;
                const int ka = rowptr[row], ke = rowptr[row + 1];
                if (rep) {
                    double ax = 0.0;
                    for (int k = ka; k < ke; ++k) {
                        const double a = staged ? sval[k - k0] : val[k];
                        const int cc = staged ? scol[k - k0] : col[k];
                        ax += a * x[(size_t)cc * G + j];
                    }
                    const double rr = b[o] - ax;
                    rs[o] = (float)rr;
                    v[0] += rr * (rr * (JAC ? invd[row] : 1.0));
                } else {
                    float ap = 0.0f;
                    for (int k = ka; k < ke; ++k) {
                        const float a = staged ? (float)sval[k - k0] : val32[k];
                        const int cc = staged ? scol[k - k0] : col[k];
                        ap += a * ps[(size_t)cc * G + j];
                    }
                    qs[o] = ap;
                    v[0] += (double)(ps[o] * ap);
                }
            }
        }
    }
    group_partials<G, 1>(v, part);
}

// alpha = gamma / (p_s, q_s) for iterating columns; for a replaced column the true gamma: the stopping
// test, then beta for the direction update; one block
template <int G>
__global__ void k_mr_alpha(MState* st, const double* __restrict__ part, int nb)
{
    if (st->all_done) return;
    double tot[G];
    series_totals<G>(part, nb, tot);
    if (threadIdx.x != 0) return;
    int all = 1;
    for (int j = 0; j < G; ++j) {
        if (!st->done[j]) {
            if (st->mode[j] == 1) {
                const double g = tot[j];
                st->nrep[j] += 1;
                st->res[j] = sqrt(fabs(g));
                if (g != g) {
                    st->nan[j] = 1;
                    st->done[j] = 1;
                } else if (g < st->thr[j]) {
                    st->converged[j] = 1;
                    st->done[j] = 1;
                } else {
                    st->beta[j] = g / st->rz[j];
                    st->rz[j] = st->rzmax[j] = g;
                    st->mode[j] = 2;
                }
            } else {
                st->mode[j] = 0;
                st->alpha[j] = st->rz[j] / tot[j];
            }
        }
        all &= st->done[j];
    }
    st->all_done = all;
}

// x_s += alpha p_s, r_s -= alpha q_s for iterating columns (FP32); partial (r_s, D_s r_s)
template <int G, bool JAC>
__global__ void k_mr_update(int n, float* __restrict__ xs, float* __restrict__ rs, const float* __restrict__ ps,
                            const float* __restrict__ qs, const float* __restrict__ invd32, const MState* st,
                            double* __restrict__ part)
{
    if (st->all_done) return;
    const int j = threadIdx.x % G;
    const bool active = !st->done[j] && st->mode[j] == 0;
    const float alpha = (float)st->alpha[j];
    double v[1] = {0.0};
    for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n && active; i += gridDim.x * (kBlock / G)) {
        const size_t o = (size_t)i * G + j;
        xs[o] += alpha * ps[o];
        const float rr = rs[o] - alpha * qs[o];
        rs[o] = rr;
        v[0] += (double)(rr * (JAC ? rr * invd32[i] : rr));
    }
    group_partials<G, 1>(v, part);
}

// gamma of the recursive residual: the iteration count, then either beta or a replacement (after a 10x
// drop since the last one, or when the recursive residual passes the stopping test); one block
template <int G>
__global__ void k_mr_beta(MState* st, const double* __restrict__ part, int nb)
{
    if (st->all_done) return;
    double tot[G];
    series_totals<G>(part, nb, tot);
    if (threadIdx.x != 0) return;
    int all = 1;
    for (int j = 0; j < G; ++j) {
        if (!st->done[j] && st->mode[j] == 0) {
            const double g = tot[j];
            st->it[j] += 1;
            if (g != g) {
                st->nan[j] = 1;
                st->done[j] = 1;
            } else if (st->it[j] >= st->maxit) {
                st->done[j] = 1;
            } else if (g < st->thr[j] || g < st->drop * st->rzmax[j]) {
                st->mode[j] = 1;
            } else {
                st->beta[j] = g / st->rz[j];
                st->rz[j] = g;
            }
        }
        all &= st->done[j];
    }
    st->all_done = all;
}

// p_s = D_s r_s + beta p_s (iterating columns and after a replacement); a column due a replacement
// first moves its correction into the FP64 solution: x += x_s, x_s = 0
template <int G, bool JAC>
__global__ void k_mr_p(int n, float* __restrict__ ps, const float* __restrict__ rs, const float* __restrict__ invd32,
                       double* __restrict__ x, float* __restrict__ xs, const MState* st)
{
    if (st->all_done) return;
    const int j = threadIdx.x % G;
    if (st->done[j]) return;
    const int mode = st->mode[j];
    const float beta = (float)st->beta[j];
    for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n; i += gridDim.x * (kBlock / G)) {
        const size_t o = (size_t)i * G + j;
        if (mode == 1) {
            x[o] += (double)xs[o];
            xs[o] = 0.0f;
        } else {
            const float zz = JAC ? rs[o] * invd32[i] : rs[o];
            ps[o] = zz + beta * ps[o];
        }
    }
}

// x += x_s over every column (a converged column's x_s is zero: its last replacement folded it in)
__global__ void k_mr_fold(long m, double* __restrict__ x, const float* __restrict__ xs)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) x[t] += (double)xs[t];
}

__global__ void k_m_to32(long n, float* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

// ---- host side
struct MultiWork {
    int n = 0, G = 0, ncol = 0;
    const Pattern* pat = nullptr;
    DevArray<int> tiles;                // row tiles of the products (for_tile_rows)
    int ntiles = 0;
    DevArray<double> x, b, r, p, q;    // FP64 interleaved (x, b also the refinement's solution and RHS)
    DevArray<float> x32, r32, p32, q32;  // reliable updates: x_s, r_s, p_s, q_s
    DevArray<double> part;
    DevArray<MState> st;
    MState* h = nullptr;                // pinned, two slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~MultiWork()
    {
        if (h) (void)hipHostFree(h);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Grid of a row-tile kernel: as many workgroups as are resident at once (occupancy per CU times the
// CUs), at most kMGrid. Each workgroup takes a contiguous range of tiles, so with the whole grid resident
// from the start no workgroup waits for a second round (3D 8-RHS solve: FP64 298 -> 275 ms at 5
// workgroups per CU, FP32 iterations 264 -> 205 ms at 7, against a fixed 2048).
template <typename K>
static int resident_grid(K kernel)
{
    int per_cu = 0, dev = 0, cus = 0;
    PFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), kBlock, 0));
    PFEM_HIP(hipGetDevice(&dev));
    PFEM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max(kBlock, std::min(kMGrid, per_cu * cus));
}

// One CG pass over the interleaved system (T = double: FP64 solve of b; T = float: inner FP32 solve from
// x = 0). Ends with st copied to h[0].
template <int G, typename T, typename VT>
static void m_cg_run(Context* c, pfem_solver_t* s, MultiWork& w, const Pattern& P, const VT* val, const T* invd,
                     T* x, const T* b, T* r, T* p, T* q, int x_zero, double tol, int tol_mode, int maxit, int ncol)
{
    const int n = P.n;
    const bool jac = invd != nullptr;
    MState init{};
    init.tol = tol;
    init.tol_mode = tol_mode;
    init.maxit = maxit;
    init.ncol = ncol;
    PFEM_HIP(hipMemcpyAsync(w.st.p, &init, sizeof(MState), hipMemcpyHostToDevice, c->stream));
    if (x_zero) PFEM_HIP(hipMemsetAsync(x, 0, sizeof(T) * (size_t)n * G, c->stream));
    double* pa = w.part.p;
    double* pb = w.part.p + (size_t)2 * G * kMGrid;
    const dim3 gr(kMGrid), bl(kBlock);
    const int ns = resident_grid(k_m_spmv<G, T, VT>);   // the tile kernels' grid
    const dim3 gs(ns);
    if (jac) hipLaunchKernelGGL((k_m_init<G, T, VT, true>), gs, bl, 0, c->stream, n, w.ntiles, w.tiles.p, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    else hipLaunchKernelGGL((k_m_init<G, T, VT, false>), gs, bl, 0, c->stream, n, w.ntiles, w.tiles.p, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    hipLaunchKernelGGL((k_m_begin<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pa, ns);
    PFEM_HIP(hipGetLastError());
    const int batch = s->poll;
    int launched = 0, pend = -1, pend_launched = 0, slot = 0;
    while (true) {
        for (int jb = 0; jb < batch; ++jb) {
            const int k = launched + jb;
            hipLaunchKernelGGL((k_m_spmv<G, T, VT>), gs, bl, 0, c->stream, w.ntiles, w.tiles.p, P.rowptr.p, P.col.p, val, p, q, w.st.p, pa);
            hipLaunchKernelGGL((k_m_alpha<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pa, ns, k);
            if (jac) hipLaunchKernelGGL((k_m_update<G, T, true>), gr, bl, 0, c->stream, n, x, r, p, q, invd, w.st.p, pb);
            else hipLaunchKernelGGL((k_m_update<G, T, false>), gr, bl, 0, c->stream, n, x, r, p, q, invd, w.st.p, pb);
            hipLaunchKernelGGL((k_m_beta<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pb, kMGrid, k);
            if (jac) hipLaunchKernelGGL((k_m_p<G, T, true>), gr, bl, 0, c->stream, n, p, r, invd, w.st.p);
            else hipLaunchKernelGGL((k_m_p<G, T, false>), gr, bl, 0, c->stream, n, p, r, invd, w.st.p);
        }
        launched += batch;
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(w.h + slot, w.st.p, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipEventRecord(w.ev[slot], c->stream));
        if (pend >= 0) {
            PFEM_HIP(hipEventSynchronize(w.ev[pend]));
            if (w.h[pend].all_done || pend_launched > maxit + batch) break;
        }
        pend = slot;
        pend_launched = launched;
        slot ^= 1;
    }
    PFEM_HIP(hipMemcpyAsync(w.h, w.st.p, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
}

// Reliable-update PCG over the interleaved system (x, b in w.x, w.b); ends with the state in h[0]
template <int G>
static void m_ru_run(Context* c, pfem_solver_t* s, MultiWork& w, const Pattern& P, pfem_csr_t* a, bool jac)
{
    const int n = P.n;
    const size_t m = (size_t)n * G;
    if (!w.x32.p) {
        w.x32.alloc(m); w.r32.alloc(m); w.p32.alloc(m); w.q32.alloc(m);
    }
    if (!a->val32.p) a->val32.alloc(P.nnz);
    hipLaunchKernelGGL(k_m_to32, dim3(blocks_for(P.nnz)), dim3(kBlock), 0, c->stream, (long)P.nnz, a->val32.p, a->val.p);
    const double* invd = jac ? s->pc->invd.p : nullptr;
    const float* invd32 = jac ? s->pc->invd32.p : nullptr;
    PFEM_REQUIRE(!jac || invd32, "FP32 Jacobi diagonal missing");
    MState init{};
    init.tol = s->tol;
    init.tol_mode = s->tol_mode;
    init.maxit = s->maxit;
    init.ncol = w.ncol;
    init.drop = std::pow(10.0, -s->replace_digits);
    PFEM_HIP(hipMemcpyAsync(w.st.p, &init, sizeof(MState), hipMemcpyHostToDevice, c->stream));
    double* pa = w.part.p;
    double* pb = w.part.p + (size_t)2 * G * kMGrid;
    float *xs = w.x32.p, *rs = w.r32.p, *ps = w.p32.p, *qs = w.q32.p;
    const dim3 gr(kMGrid), bl(kBlock);
    const int ns = jac ? resident_grid(k_mr_spmv<G, true>) : resident_grid(k_mr_spmv<G, false>);
    const dim3 gs(ns);
    if (jac) hipLaunchKernelGGL((k_mr_init<G, true>), gs, bl, 0, c->stream, w.ntiles, w.tiles.p, P.rowptr.p, P.col.p, a->val.p, (const double*)w.x.p, (const double*)w.b.p, invd, invd32, rs, ps, xs, pa);
    else hipLaunchKernelGGL((k_mr_init<G, false>), gs, bl, 0, c->stream, w.ntiles, w.tiles.p, P.rowptr.p, P.col.p, a->val.p, (const double*)w.x.p, (const double*)w.b.p, invd, invd32, rs, ps, xs, pa);
    hipLaunchKernelGGL((k_mr_begin<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pa, ns);
    PFEM_HIP(hipGetLastError());
    const int batch = s->poll;
    int launched = 0, pend = -1, slot = 0;
    // a replacement takes one extra slot per tenfold residual drop; the slot budget allows for them
    const int slots = s->maxit + s->maxit / 4 + 64;
    while (true) {
        for (int jb = 0; jb < batch; ++jb) {
            if (jac) hipLaunchKernelGGL((k_mr_spmv<G, true>), gs, bl, 0, c->stream, w.ntiles, w.tiles.p, P.rowptr.p, P.col.p, (const float*)a->val32.p, (const double*)a->val.p, (const float*)ps, qs, (const double*)w.x.p, (const double*)w.b.p, invd, rs, (const MState*)w.st.p, pa);
            else hipLaunchKernelGGL((k_mr_spmv<G, false>), gs, bl, 0, c->stream, w.ntiles, w.tiles.p, P.rowptr.p, P.col.p, (const float*)a->val32.p, (const double*)a->val.p, (const float*)ps, qs, (const double*)w.x.p, (const double*)w.b.p, invd, rs, (const MState*)w.st.p, pa);
            hipLaunchKernelGGL((k_mr_alpha<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pa, ns);
            if (jac) hipLaunchKernelGGL((k_mr_update<G, true>), gr, bl, 0, c->stream, n, xs, rs, (const float*)ps, (const float*)qs, invd32, (const MState*)w.st.p, pb);
            else hipLaunchKernelGGL((k_mr_update<G, false>), gr, bl, 0, c->stream, n, xs, rs, (const float*)ps, (const float*)qs, invd32, (const MState*)w.st.p, pb);
            hipLaunchKernelGGL((k_mr_beta<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pb, kMGrid);
            if (jac) hipLaunchKernelGGL((k_mr_p<G, true>), gr, bl, 0, c->stream, n, ps, (const float*)rs, invd32, w.x.p, xs, (const MState*)w.st.p);
            else hipLaunchKernelGGL((k_mr_p<G, false>), gr, bl, 0, c->stream, n, ps, (const float*)rs, invd32, w.x.p, xs, (const MState*)w.st.p);
        }
        launched += batch;
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(w.h + slot, w.st.p, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipEventRecord(w.ev[slot], c->stream));
        if (pend >= 0) {
            PFEM_HIP(hipEventSynchronize(w.ev[pend]));
            if (w.h[pend].all_done || launched > slots + 2 * batch) break;
        }
        pend = slot;
        slot ^= 1;
    }
    // columns still iterating at the slot limit keep their FP64 part; their FP32 correction is folded in
    hipLaunchKernelGGL(k_mr_fold, dim3(blocks_for((long)m)), bl, 0, c->stream, (long)m, w.x.p, (const float*)xs);
    PFEM_HIP(hipMemcpyAsync(w.h, w.st.p, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
}

template <int G>
static void solve_multi_g(Context* c, pfem_solver_t* s, pfem_csr_t* a, int ncol, double* X, const double* B,
                          pfem_solve_stats* stats)
{
    const Pattern& P = *a->pat;
    const int n = P.n;
    if (!s->multi || s->multi->G != G || s->multi->n != n || s->multi->pat != &P) {
        s->multi.reset(new MultiWork());
        MultiWork& w = *s->multi;
        w.n = n;
        w.G = G;
        w.pat = &P;
        std::vector<int> tiles;
        for (int r = 0; r < n;) {
            int r1 = r;
            const int base = P.h_rowptr[r];
            while (r1 < n && r1 - r < kMRows * kBlock / G && (P.h_rowptr[r1 + 1] - base <= kMTileNnz || r1 == r)) ++r1;
            tiles.push_back(r);
            tiles.push_back(r1);
            r = r1;
        }
        w.ntiles = (int)tiles.size() / 2;
        upload(c, w.tiles, tiles);
        const size_t m = (size_t)n * G;
        w.x.alloc(m); w.b.alloc(m); w.r.alloc(m); w.p.alloc(m); w.q.alloc(m);
        w.part.alloc((size_t)4 * G * kMGrid);
        w.st.alloc(1);
        PFEM_HIP(hipHostMalloc((void**)&w.h, 2 * sizeof(MState), hipHostMallocDefault));
        for (hipEvent_t& e : w.ev) PFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    MultiWork& w = *s->multi;
    w.ncol = ncol;
    const size_t m = (size_t)n * G;
    const unsigned gp = (unsigned)std::min<size_t>(blocks_for((long)m), 8192);
    const bool jac = s->pc && s->pc->kind == PFEM_PRECOND_JACOBI;
    if (s->pc) precond_initialize(c, s->pc, a);
    hipLaunchKernelGGL((k_m_pack<G, double>), dim3(gp), dim3(kBlock), 0, c->stream, n, ncol, (const double*)X, w.x.p);
    hipLaunchKernelGGL((k_m_pack<G, double>), dim3(gp), dim3(kBlock), 0, c->stream, n, ncol, B, w.b.p);
    std::vector<pfem_solve_stats> st(ncol);
    if (s->precision != PFEM_PRECISION_MIXED) {
        m_cg_run<G, double, double>(c, s, w, P, a->val.p, jac ? s->pc->invd.p : nullptr, w.x.p, w.b.p, w.r.p, w.p.p,
                                    w.q.p, 0, s->tol, s->tol_mode, s->maxit, ncol);
        for (int j = 0; j < ncol; ++j) {
            st[j].converged = w.h->converged[j];
            st[j].iterations = w.h->it[j];
            st[j].final_residual = w.h->res[j];
            st[j].nan_detected = w.h->nan[j];
        }
    } else {
        m_ru_run<G>(c, s, w, P, a, jac);
        for (int j = 0; j < ncol; ++j) {
            st[j].converged = w.h->converged[j];
            st[j].iterations = w.h->it[j];
            st[j].final_residual = w.h->res[j];
            st[j].nan_detected = w.h->nan[j];
            st[j].refinements = w.h->nrep[j];
        }
    }
    hipLaunchKernelGGL((k_m_unpack<G>), dim3(gp), dim3(kBlock), 0, c->stream, n, ncol, (const double*)w.x.p, X);
    PFEM_HIP(hipGetLastError());
    PFEM_HIP(hipStreamSynchronize(c->stream));
    if (stats) std::memcpy(stats, st.data(), sizeof(pfem_solve_stats) * ncol);
}

void MultiWorkDeleter::operator()(MultiWork* w) const { delete w; }

}  // namespace pfem

using namespace pfem;

extern "C" {

int pfem_solver_solve_multi(pfem_solver s, pfem_csr a, int nrhs, double* d_X, const double* d_B,
                            pfem_solve_stats* stats)
{
    PFEM_TRY(s ? s->ctx : nullptr, {
        PFEM_REQUIRE(s && a && d_X && d_B, "null argument");
        PFEM_REQUIRE(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs must be 1..8");
        PFEM_REQUIRE(a->pat->n == s->n, "matrix size differs from the solver's");
        PFEM_REQUIRE(s->kind == PFEM_SOLVER_CG, "multi-RHS solves are PCG only");
        PFEM_REQUIRE(!s->pc || s->pc->kind == PFEM_PRECOND_JACOBI || s->pc->kind == PFEM_PRECOND_NONE,
                     "multi-RHS PCG takes a Jacobi preconditioner or none");
        PFEM_REQUIRE(s->precision == PFEM_PRECISION_FP64 || s->precision == PFEM_PRECISION_MIXED,
                     "multi-RHS PCG runs in FP64 or MIXED precision");
        Context* c = C(s->ctx);
        if (nrhs == 1) solve_multi_g<1>(c, s, a, nrhs, d_X, d_B, stats);
        else if (nrhs == 2) solve_multi_g<2>(c, s, a, nrhs, d_X, d_B, stats);
        else if (nrhs <= 4) solve_multi_g<4>(c, s, a, nrhs, d_X, d_B, stats);
        else solve_multi_g<8>(c, s, a, nrhs, d_X, d_B, stats);
    });
}

}  // extern "C"
