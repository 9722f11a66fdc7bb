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
// PFEM_PRECISION_MIXED: FP32 inner multi-RHS CG on the FP32 copy of A with FP64 iterative refinement of
// every column (residuals, the stopping test and the solution update in FP64), as the single-RHS mixed
// solver does.
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
    // refinement
    double irthr[kMaxRhs], irrz[kMaxRhs], scale[kMaxRhs];
    int irfirst, irconv[kMaxRhs], irnan[kMaxRhs], irdone;
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

// Totals of NJ partial series of kMGrid terms each (series j at part + j * kMGrid), valid in thread 0.
// Every load of a lane is issued before the first addition (one memory round trip); lane t adds terms
// t, t + 256, ... of each series in order, then block_sum.
template <int NJ>
__device__ __forceinline__ void series_totals(const double* __restrict__ part, double (&tot)[NJ])
{
    constexpr int kPer = kMGrid / kBlock;
    double t[NJ][kPer];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int u = 0; u < kPer; ++u) t[j][u] = part[(size_t)j * kMGrid + threadIdx.x + u * kBlock];
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
// sum_k A_ik x_kj for the calling lane's row i and column j, left to right
template <int G, typename T, typename VT>
__device__ __forceinline__ T group_row_product(int i, int j, const int* __restrict__ rowptr, const int* __restrict__ col,
                                               const VT* __restrict__ val, const T* __restrict__ x)
{
    T s = 0;
    const int k1 = rowptr[i + 1];
    for (int k = rowptr[i]; k < k1; ++k) s += (T)val[k] * x[(size_t)col[k] * G + j];
    return s;
}

// r = b - A x (or b when x = 0), p = z = M^-1 r; partials (r, z) and (b, M^-1 b) per column
template <int G, typename T, typename VT, bool JAC>
__global__ __launch_bounds__(kBlock) void k_m_init(int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const VT* __restrict__ val, const T* __restrict__ x,
                                                   const T* __restrict__ b, const T* __restrict__ invd, T* __restrict__ r,
                                                   T* __restrict__ p, double* __restrict__ part, int x_zero)
{
    const int j = threadIdx.x % G;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n; i += gridDim.x * (kBlock / G)) {
        const size_t o = (size_t)i * G + j;
        const T bb = b[o];
        const T rr = x_zero ? bb : bb - group_row_product<G, T, VT>(i, j, rowptr, col, val, x);
        const T zz = JAC ? rr * invd[i] : rr;
        const T mb = JAC ? bb * invd[i] : bb;
        r[o] = rr;
        p[o] = zz;
        v[0] += (double)(rr * zz);
        v[1] += (double)(bb * mb);
    }
    group_partials<G, 2>(v, part);
}

// gamma_0, the per-column thresholds (linear_solver.cu:360; relative to (b, M^-1 b) on request, absolute
// for a zero column) and the columns already converged; one block
template <int G>
__global__ void k_m_begin(MState* st, const double* __restrict__ part)
{
    double tot[2 * G];
    series_totals<2 * G>(part, tot);
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
__global__ __launch_bounds__(kBlock) void k_m_spmv(int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const VT* __restrict__ val, const T* __restrict__ p, T* __restrict__ q,
                                                   const MState* st, double* __restrict__ part)
{
    if (st->all_done) return;
    const int j = threadIdx.x % G;
    double v[1] = {0.0};
    for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n; i += gridDim.x * (kBlock / G)) {
        const size_t o = (size_t)i * G + j;
        const T ap = group_row_product<G, T, VT>(i, j, rowptr, col, val, p);
        q[o] = ap;
        v[0] += (double)(p[o] * ap);
    }
    group_partials<G, 1>(v, part);
}

// alpha_j = gamma_j / (p_j, q_j) for the active columns; one block
template <int G>
__global__ void k_m_alpha(MState* st, const double* __restrict__ part, int k)
{
    if (st->all_done) return;
    double tot[G];
    series_totals<G>(part, tot);
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
__global__ void k_m_beta(MState* st, const double* __restrict__ part, int k)
{
    if (st->all_done) return;
    double tot[G];
    series_totals<G>(part, tot);
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

// FP64 refinement: r = b - A x (FP64); partials (r, M^-1 r) and (b, M^-1 b) per column
template <int G, bool JAC>
__global__ __launch_bounds__(kBlock) void k_m_ir_residual(int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                          const double* __restrict__ val, const double* __restrict__ x,
                                                          const double* __restrict__ b, const double* __restrict__ invd,
                                                          double* __restrict__ r, double* __restrict__ part)
{
    const int j = threadIdx.x % G;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * (kBlock / G) + threadIdx.x / G; i < n; i += gridDim.x * (kBlock / G)) {
        const size_t o = (size_t)i * G + j;
        const double bb = b[o];
        const double rr = bb - group_row_product<G, double, double>(i, j, rowptr, col, val, x);
        r[o] = rr;
        const double d = JAC ? invd[i] : 1.0;
        v[0] += rr * (rr * d);
        v[1] += bb * (bb * d);
    }
    group_partials<G, 2>(v, part);
}

// refinement test per column (the FP64 solver's criterion) and the scale of its next correction
template <int G>
__global__ void k_m_ir_test(MState* st, const double* __restrict__ part)
{
    double tot[2 * G];
    series_totals<2 * G>(part, tot);
    if (threadIdx.x != 0) return;
    int all = 1;
    for (int j = 0; j < G; ++j) {
        if (st->irfirst) {
            const bool rel = st->tol_mode == PFEM_TOL_RELATIVE && tot[G + j] > 0.0;
            st->irthr[j] = rel ? st->tol * st->tol * tot[G + j] : st->tol * st->tol;
        }
        st->irrz[j] = tot[j];
        st->irnan[j] = tot[j] != tot[j];
        st->irconv[j] = tot[j] < st->irthr[j] ? 1 : 0;
        const bool stop = j >= st->ncol || st->irconv[j] || st->irnan[j];
        // a finished column gets a zero inner right-hand side (scale 0): its inner solve stops at once
        st->scale[j] = stop ? 0.0 : (sqrt(tot[j]) > 0.0 ? sqrt(tot[j]) : 1.0);
        all &= stop ? 1 : 0;
    }
    st->irfirst = 0;
    st->irdone = all;
}

template <int G>
__global__ void k_m_ir_down(int n, float* __restrict__ d, const double* __restrict__ s, const MState* st)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)n * G; t += (long)gridDim.x * blockDim.x) {
        const double sc = st->scale[t % G];
        d[t] = sc > 0.0 ? (float)(s[t] / sc) : 0.0f;
    }
}

template <int G>
__global__ void k_m_ir_up(int n, double* __restrict__ x, const float* __restrict__ d, const MState* st)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)n * G; t += (long)gridDim.x * blockDim.x) {
        const double sc = st->scale[t % G];
        if (sc > 0.0) x[t] += sc * (double)d[t];
    }
}

__global__ void k_m_to32(long n, float* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

// ---- host side
struct MultiWork {
    int n = 0, G = 0;
    DevArray<double> x, b, r, p, q;    // FP64 interleaved (x, b also the refinement's solution and RHS)
    DevArray<float> x32, b32, r32, p32, q32;
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
    // refinement fields are kept across inner passes
    PFEM_HIP(hipMemcpyAsync(w.st.p, &init, offsetof(MState, irthr), hipMemcpyHostToDevice, c->stream));
    if (x_zero) PFEM_HIP(hipMemsetAsync(x, 0, sizeof(T) * (size_t)n * G, c->stream));
    double* pa = w.part.p;
    double* pb = w.part.p + (size_t)2 * G * kMGrid;
    const dim3 gr(kMGrid), bl(kBlock);
    if (jac) hipLaunchKernelGGL((k_m_init<G, T, VT, true>), gr, bl, 0, c->stream, n, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    else hipLaunchKernelGGL((k_m_init<G, T, VT, false>), gr, bl, 0, c->stream, n, P.rowptr.p, P.col.p, val, x, b, invd, r, p, pa, x_zero);
    hipLaunchKernelGGL((k_m_begin<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pa);
    PFEM_HIP(hipGetLastError());
    const int batch = s->poll;
    int launched = 0, pend = -1, pend_launched = 0, slot = 0;
    while (true) {
        for (int jb = 0; jb < batch; ++jb) {
            const int k = launched + jb;
            hipLaunchKernelGGL((k_m_spmv<G, T, VT>), gr, bl, 0, c->stream, n, P.rowptr.p, P.col.p, val, p, q, w.st.p, pa);
            hipLaunchKernelGGL((k_m_alpha<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pa, k);
            if (jac) hipLaunchKernelGGL((k_m_update<G, T, true>), gr, bl, 0, c->stream, n, x, r, p, q, invd, w.st.p, pb);
            else hipLaunchKernelGGL((k_m_update<G, T, false>), gr, bl, 0, c->stream, n, x, r, p, q, invd, w.st.p, pb);
            hipLaunchKernelGGL((k_m_beta<G>), dim3(1), bl, 0, c->stream, w.st.p, (const double*)pb, k);
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

template <int G>
static void solve_multi_g(Context* c, pfem_solver_t* s, pfem_csr_t* a, int ncol, double* X, const double* B,
                          pfem_solve_stats* stats)
{
    const Pattern& P = *a->pat;
    const int n = P.n;
    if (!s->multi || s->multi->G != G || s->multi->n != n) {
        s->multi.reset(new MultiWork());
        MultiWork& w = *s->multi;
        w.n = n;
        w.G = G;
        const size_t m = (size_t)n * G;
        w.x.alloc(m); w.b.alloc(m); w.r.alloc(m); w.p.alloc(m); w.q.alloc(m);
        w.part.alloc((size_t)4 * G * kMGrid);
        w.st.alloc(1);
        PFEM_HIP(hipHostMalloc((void**)&w.h, 2 * sizeof(MState), hipHostMallocDefault));
        for (hipEvent_t& e : w.ev) PFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    MultiWork& w = *s->multi;
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
        if (!w.x32.p) {
            w.x32.alloc(m); w.b32.alloc(m); w.r32.alloc(m); w.p32.alloc(m); w.q32.alloc(m);
        }
        if (!a->val32.p) a->val32.alloc(P.nnz);
        hipLaunchKernelGGL(k_m_to32, dim3(blocks_for(P.nnz)), dim3(kBlock), 0, c->stream, (long)P.nnz, a->val32.p, a->val.p);
        MState init{};
        init.tol = s->tol;
        init.tol_mode = s->tol_mode;
        init.ncol = ncol;
        init.irfirst = 1;
        PFEM_HIP(hipMemcpyAsync(w.st.p, &init, sizeof(MState), hipMemcpyHostToDevice, c->stream));
        const double* invd = jac ? s->pc->invd.p : nullptr;
        const float* invd32 = jac ? s->pc->invd32.p : nullptr;
        PFEM_REQUIRE(!jac || invd32, "FP32 Jacobi diagonal missing");
        std::vector<int> inner(ncol, 0);
        int refinements = 0, total = 0;
        while (true) {
            if (jac) hipLaunchKernelGGL((k_m_ir_residual<G, true>), dim3(kMGrid), dim3(kBlock), 0, c->stream, n, P.rowptr.p, P.col.p, a->val.p, (const double*)w.x.p, (const double*)w.b.p, invd, w.r.p, w.part.p);
            else hipLaunchKernelGGL((k_m_ir_residual<G, false>), dim3(kMGrid), dim3(kBlock), 0, c->stream, n, P.rowptr.p, P.col.p, a->val.p, (const double*)w.x.p, (const double*)w.b.p, invd, w.r.p, w.part.p);
            hipLaunchKernelGGL((k_m_ir_test<G>), dim3(1), dim3(kBlock), 0, c->stream, w.st.p, (const double*)w.part.p);
            PFEM_HIP(hipMemcpyAsync(w.h, w.st.p, sizeof(MState), hipMemcpyDeviceToHost, c->stream));
            PFEM_HIP(hipStreamSynchronize(c->stream));
            if (w.h->irdone || refinements >= s->max_refine || total >= s->maxit) break;
            hipLaunchKernelGGL((k_m_ir_down<G>), dim3(gp), dim3(kBlock), 0, c->stream, n, w.b32.p, (const double*)w.r.p, (const MState*)w.st.p);
            const int maxit = std::max(1, s->maxit - total);
            m_cg_run<G, float, float>(c, s, w, P, a->val32.p, invd32, w.x32.p, w.b32.p, w.r32.p, w.p32.p, w.q32.p, 1,
                                      s->inner_tol, PFEM_TOL_RELATIVE, maxit, ncol);
            int most = 0;
            for (int j = 0; j < ncol; ++j) {
                inner[j] += w.h->it[j];
                most = std::max(most, w.h->it[j]);
            }
            total += most;
            hipLaunchKernelGGL((k_m_ir_up<G>), dim3(gp), dim3(kBlock), 0, c->stream, n, w.x.p, (const float*)w.x32.p, (const MState*)w.st.p);
            ++refinements;
        }
        for (int j = 0; j < ncol; ++j) {
            st[j].converged = w.h->irconv[j];
            st[j].iterations = inner[j];
            st[j].final_residual = std::sqrt(std::fabs(w.h->irrz[j]));
            st[j].nan_detected = w.h->irnan[j];
            st[j].refinements = refinements;
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
