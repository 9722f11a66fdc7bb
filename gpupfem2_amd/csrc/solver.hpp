// Krylov solver and preconditioner state (internal).
#pragma once

#include <cmath>
#include <cstddef>

#include "internal.hpp"

namespace pfem {

// A matrix or triangle sliced by row chunks, column-major within a chunk: entry w of the chunk's row t
// in slot base + w * rows + t; a row's unused slots follow its entries and hold column -1. hdr: {first
// row, rows, base, width} per chunk; col: 32-bit column of each slot; col16: its 16-bit code (colour << 14
// | offset from cbase[4 * chunk + colour], 0xFFFF unused), valid where cbase[4 * chunk] >= 0; pos: the
// CSR slot each entry's value comes from (-1 unused).
struct SlicedCols {
    DevArray<int> hdr, cbase, col, pos;
    DevArray<unsigned short> col16;
};

// Blocked layout (PFEM_ORDER_BLOCKS): a chunk holds blocks of one colour and one size tc (a power of two,
// at most kBlockRows), a block's rows consecutive, one row per lane, so a block lies in one wave; the
// triangle is chunk-sliced as for the multicolour order, except that a column inside the row's own block
// is coded -2 - t' (t' the column's position in the block); cbase per chunk: {tc, 0, 0, 0}.
constexpr int kBlockRows = 8;

// A chunk's four colour bases, loaded once (uniform values) so that decoding a column is register work
struct ChunkBases {
    int b0, b1, b2, b3;
    bool wide;
};

// Position of this workgroup's entry in a sliced layout's header and base tables. Workgroup b runs on XCD
// b mod 8; the tables hold the entries of XCD x's workgroups as the x-th contiguous band (the order of
// xcd_band_tile), so the workgroups that one XCD runs one after the other read neighbouring entries, four to
// a cache line of its L2, rather than one line each. A bijection for every grid size; the tables are
// stored in this order by the layout builder.
__device__ __forceinline__ int hdr_slot()
{
    const int b = (int)blockIdx.x, g = (int)gridDim.x;
    const int x = b & 7, q = g >> 3, rem = g & 7;
    return x * q + min(x, rem) + (b >> 3);
}

__device__ __forceinline__ ChunkBases chunk_bases(const int* __restrict__ cb)
{
    return ChunkBases{cb[0], cb[1], cb[2], cb[3], cb[0] < 0};
}

// Column of sliced slot e (-1 for an unused slot)
__device__ __forceinline__ int slot_col(int e, const int* __restrict__ col32, const unsigned short* __restrict__ col16,
                                        const ChunkBases& cb)
{
    if (cb.wide) return col32[e];
    const unsigned v = col16[e];
    const unsigned c = v >> 14;
    const int base = c == 0 ? cb.b0 : c == 1 ? cb.b1 : c == 2 ? cb.b2 : cb.b3;
    return v == 0xFFFFu ? -1 : base + (int)(v & 0x3FFFu);
}

struct MultiWork;   // multi-RHS PCG work space (multi.hip)
struct MultiWorkDeleter {
    void operator()(MultiWork* w) const;
};
struct BatchWork;   // batched-systems PCG work space (batched.hip)
struct BatchWorkDeleter {
    void operator()(BatchWork* w) const;
};

constexpr int kMaxRestart = 1024;   // largest GMRES cycle (the reference's maxIterations default is 1000)
constexpr int kVecGrid = 1024;      // grid of the vector kernels (grid-stride loops)

// Device-resident PCG scalars, (re)written by the host at the start of each solve. gam[k & 1] holds
// (r_k, z_k): iteration k reads one slot while its last kernel writes the other; alpha[k & 1] is alpha_k
// (recorded for diagnostics).
struct CgState {
    double tol;
    int tol_mode, maxit;
    double gam[2], alpha[2], thr, res;
    int it, done, converged, nan;
    // fused iterations: the value of it when the current iteration's SpMV ran; the x/p update of an
    // iteration applies only while it differs (the iteration ended since), so launches queued past
    // the end of the solve change nothing
    int xp_seen;
};

// End of PCG iteration k, given gnew = (r_{k+1}, z_{k+1}): gamma_{k+1}, the iteration count and the
// stopping test (linear_solver.cu:360); one thread.
__device__ __forceinline__ void cg_record(CgState* st, double gnew, int k)
{
    st->gam[(k + 1) & 1] = gnew;
    st->it = k + 1;
    st->res = sqrt(fabs(gnew));
    if (gnew != gnew) {
        st->nan = 1;
        st->done = 1;
    } else if (gnew < st->thr) {
        st->converged = 1;
        st->done = 1;
    } else if (k + 1 >= st->maxit) {
        st->done = 1;
    }
}

// alpha_k = gamma_k / (p_k, A p_k) of the iteration the state is in (k = it), recorded by the fused SpMV's
// last workgroup, which also marks the iteration's x/p update as due (xp_seen = k < it once the
// iteration ends); one thread.
__device__ __forceinline__ void cg_alpha_record(CgState* st, double pq)
{
    const int k = st->it;
    st->alpha[k & 1] = st->gam[k & 1] / pq;
    st->xp_seen = k;
}

// Device-resident GMRES(m) scalars and small arrays (H column-major with leading dimension m).
struct GmState {
    double tol;
    int tol_mode, maxit, m, restarted;
    double thr;
    double *H, *beta, *cs, *sn, *y, *hc;
    double res, norm;
    int k, it, done, converged, cycle_done, nan, apply_update;
};

#ifdef __HIPCC__
// Start of a cycle (thread 0 of the last block): beta_0 = ||w||, convergence test
// (linear_solver.cu:446-453; tolerance relative to ||M^-1 b|| when tol_mode is relative).
__device__ inline void gm_cycle_start(GmState* st, double w2, double mb2)
{
    // relative to ||M^-1 b||; a zero right-hand side falls back to the absolute test
    if (st->thr < 0.0) st->thr = st->tol_mode == PFEM_TOL_RELATIVE && mb2 > 0.0 ? st->tol * sqrt(mb2) : st->tol;
    const double b0 = sqrt(w2);
    st->beta[0] = b0;
    st->res = b0;
    st->norm = b0;
    st->k = 0;
    st->cycle_done = 0;
    if (b0 != b0) {
        st->nan = 1; st->done = 1; st->cycle_done = 1;
    } else if (b0 < st->thr) {
        st->converged = 1; st->done = 1; st->cycle_done = 1;
    }
}

// End of Arnoldi step j = st->k (thread 0 of the last block): updateHcolumn (linear_solver.cu:40-66)
// with h_{j+1,j} = ||w||. GivensRotation (cuda_math.cuh:105-113) uses 1/sqrt here.
__device__ inline void gm_arnoldi_finish(GmState* st, double hn)
{
    const int j = st->k, m = st->m;
    double* h = st->H + (size_t)j * m;
    double* cs = st->cs;
    double* sn = st->sn;
    for (int i = 0; i < j; ++i) {
        const double t = cs[i] * h[i] + sn[i] * h[i + 1];
        h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
        h[i] = t;
    }
    const double coeff = 1.0 / sqrt(h[j] * h[j] + hn * hn);
    cs[j] = h[j] * coeff;
    sn[j] = hn * coeff;
    h[j] = cs[j] * h[j] + sn[j] * hn;
    st->beta[j + 1] = -sn[j] * st->beta[j];
    st->beta[j] *= cs[j];
    const double res = fabs(st->beta[j + 1]);
    st->res = res;
    st->norm = hn;
    st->it += 1;
    st->k = j + 1;
    if (res != res) {
        st->nan = 1; st->done = 1; st->cycle_done = 1;
    } else if (res < st->thr) {
        st->converged = 1; st->done = 1; st->cycle_done = 1;
    } else if (st->it >= st->maxit) {
        st->done = 1; st->cycle_done = 1;
    } else if (st->k >= m) {
        st->cycle_done = 1;
    }
}
#endif

}  // namespace pfem

struct pfem_precond_t {
    pfem_ctx ctx = nullptr;
    int kind = PFEM_PRECOND_NONE;
    // PFEM_PRECOND_CUSTOM: host callbacks that enqueue their work on the context stream
    pfem_precond_init_fn custom_init = nullptr;
    pfem_precond_apply_fn custom_apply = nullptr;
    void* custom_user = nullptr;
    int ordering = PFEM_ORDER_NATURAL;
    int n = 0;
    std::shared_ptr<pfem::Pattern> pat;   // pattern of the matrices it is initialised from
    pfem::DevArray<double> invd;
    pfem::DevArray<float> invd32;
    // IC0 / ILU0: factor of (permuted) A, stored on the permuted pattern
    std::shared_ptr<pfem::Pattern> fpat;  // factor pattern (== pat for natural ordering)
    pfem::DevArray<int> src_pos;          // factor slot k takes A's value at src_pos[k]
    pfem::DevArray<int> perm, iperm;      // new -> old, old -> new (multicolour)
    std::vector<int> h_perm;
    pfem::DevArray<int> mirror;           // for each slot (i,k) with k > i: position of (k,i)
    pfem::DevArray<int> lrows, urows;     // rows grouped by forward / backward level
    std::vector<int> h_lptr, h_uptr;      // level offsets into lrows / urows
    pfem::DevArray<double> fac;
    pfem::DevArray<float> fac32;
    pfem::DevArray<double> aval;          // A's values on the permuted pattern (Krylov in factor order)
    pfem::DevArray<double> tmp, tmp2;
    pfem::DevArray<float> tmp32, tmp32b;
    pfem::DevArray<int> status;           // first failing row (+1), 0 if none
    int nlevels = 0;
    // single-pass sweeps: rows in chunks of <= 256 of one level, numbered level by level
    // ({offset into lrows/urows, count} pairs); sy / sw hold the forward / backward results in the
    // factor's ordering and read as a sentinel NaN until a row has published its value
    int sweep = 1;                        // 1: one launch per triangular solve, 0: one launch per level
    int spin_limit = 1 << 16;             // polls of one single-pass wait before it is declared failed
    pfem::DevArray<int> lchunks, uchunks;
    int nlchunks = 0, nuchunks = 0;
    int nlslices = 0, nuslices = 0;       // chunks of the sliced layouts (up to 512 rows each)
    bool contiguous = false;              // levels are row ranges (multicolour): chunks hold row ranges
    pfem::DevArray<double> sy, sw;
    pfem::DevArray<float> sy32, sw32;
    pfem::DevArray<int> sweep_err;        // set when a wait exceeded its spin limit
    int sweep_failures = 0;               // solves reported failed by a wait past its spin limit
    // the sweeps' copy of the factor: all strictly lower entries (row by row), then all strictly upper
    // entries, then the n diagonal values, each row's entries in the factor's order; split[k] takes
    // fac[split_pos[k]]. split_ptr holds the lower row offsets (n + 1), then the upper ones (n + 1).
    pfem::DevArray<int> split_pos, split_ptr, split_col;
    pfem::DevArray<double> split;
    pfem::DevArray<float> split32;
    long split_diag = 0;                  // offset of the diagonal values in split
    // colour-range chunks (contiguous levels): the split factor's triangles sliced by chunk (sl_lower,
    // sl_upper), and A in the factor's ordering sliced over the SpMV tiles of fpat (sl_all, values in
    // asl_val, for the fused PCG's SpMV)
    pfem::SlicedCols sl_lower, sl_upper, sl_all;
    pfem::DevArray<double> sell_l, sell_u, asl_val;
    pfem::DevArray<float> sell_l32, sell_u32;
    // the reduced-precision copy used with FP64 vectors (PFEM_PRECISION_MIXED_PRECOND): 32 or 16 bits per
    // value (16: the split and sliced factor only, for the single-pass sweeps)
    int factor_bits = 32;
    bool half_ok = true;                  // the last 16-bit copy kept every diagonal (else the 32-bit copy is used)
    pfem::DevArray<_Float16> split16, sell_l16, sell_u16;   // c L, c = scale16[0] (a power of two)
    pfem::DevArray<double> scale16;       // c, c^2, and the largest |L| found (as bits)
    bool sell = false;
    // PFEM_ORDER_BLOCKS: sl_lower / sl_upper hold the blocked layout (one block per lane; the sliced slots
    // take the dependencies outside the block, a dense triangle per block the ones inside); the per-level
    // tables (lrows, lchunks, ...) hold the longest-path levels of the same factor
    bool blocked = false;
};

struct pfem_solver_t {
    pfem_ctx ctx = nullptr;
    int kind = PFEM_SOLVER_CG;
    double tol = 1e-8;
    int maxit = 1000, restart = 1000, m = 0, restarted = 0, tol_mode = PFEM_TOL_ABSOLUTE;
    pfem_precond_t* pc = nullptr;
    int poll = 8;       // iterations launched between two host polls of the done flag
    int reorth = 1;     // GMRES: classical Gram-Schmidt with one re-orthogonalisation pass
    int precision = PFEM_PRECISION_FP64;
    double inner_tol = 1e-6;
    int max_refine = 50;
    int n = 0;
    pfem::DevArray<double> r, z, p, q, t, V, H, beta, cs, sn, y, hc;
    pfem::DevArray<double> xp, bp;   // x and b in a multicolour factor's ordering
    // per-block partial sums of the reductions: pa (SpMV-shaped producers, 2 values x tiles), pb (vector
    // producers: 2 values, or one per GMRES coefficient, x kVecGrid), pn (||w||^2 of an Arnoldi step)
    pfem::DevArray<double> partials;
    double *pa = nullptr, *pb = nullptr, *pn = nullptr;
    int pb_cap = 0, pa_cap = 0;
    int factor_order = 1;   // run multicolour IC0/ILU0 solves in the factor's ordering
    // PCG kernel fusion in the factor's ordering: 0 none; 1 x/r update inside the forward sweep; 2 also
    // alpha and beta finished by the last workgroups of the SpMV and of the backward sweep; 3 tagged
    // sweeps (IC0: three launches per iteration)
    int fuse = 3;
    int graph = 0;                        // fused PCG: replay a captured batch of iterations
    int replace_digits = 2;   // reliable updates (multi/batched MIXED): replace after gamma fell 10^-digits
    hipEvent_t ev[2] = {nullptr, nullptr};   // pipelined polls of the device state
    std::unique_ptr<pfem::MultiWork, pfem::MultiWorkDeleter> multi;   // pfem_solver_solve_multi
    std::unique_ptr<pfem::BatchWork, pfem::BatchWorkDeleter> batch;   // pfem_solver_solve_batched
    std::vector<int> batch_tiling;   // {row bound, tile rows, tile non-zeros} per size bucket (empty: defaults)
    pfem::DevArray<pfem::CgState> cg;
    pfem::DevArray<pfem::GmState> gm;
    pfem::CgState* h_cg = nullptr;   // pinned, two slots
    pfem::GmState* h_gm = nullptr;   // pinned
    // mixed precision (FP32 inner CG, FP64 refinement)
    pfem::DevArray<float> r32, z32, p32, q32, d32, t32, V32;
    pfem::DevArray<double> xr;       // refinement residual, its preconditioned copy, refinement state
};

namespace pfem {
void precond_initialize(Context* c, pfem_precond_t* pc, pfem_csr_t* a);
void precond_apply(Context* c, pfem_precond_t* pc, double* z, const double* r);
void precond_apply32(Context* c, pfem_precond_t* pc, float* z, const float* r);
void precond_check(Context* c, pfem_precond_t* pc);
// Options of an apply. part: the dot product an IC0/ILU0 apply can produce on the way, per-chunk
// partials of (r, z) written to part[0..m) (m = returned count, at most cap). factor_order: r and z are
// already in the factor's (multicolour) ordering, so the sweeps do not permute.
template <typename TV>
struct ApplyOpts {
    double* part = nullptr;
    int cap = 0;
    bool factor_order = false;
    // PCG fusion (single-pass sweeps in factor order): skip when *done; with alpha, r -= alpha q (r = the
    // apply's input) before the forward sweep uses r, and x += alpha p as well when x is given
    const int* done = nullptr;
    TV* x = nullptr;
    const TV* p = nullptr;
    const TV* q = nullptr;
    const double* alpha = nullptr;
    // with part: the backward sweep publishes its partials (sentinel-armed slots) and its last workgroup
    // adds them and ends the PCG iteration cg is in (iteration cg->it, cg_record); nothing further is
    // needed for beta. With an IC0 factor and tagged sweeps the forward sweep does this from (y, y).
    CgState* cg = nullptr;
    // with alpha pointing at alpha_cg->alpha: the forward sweep takes alpha[it & 1] of the iteration the
    // state is in (the iteration index lives on the device, so a batch of iterations is always the same
    // sequence of launches)
    const CgState* alpha_cg = nullptr;
    // single-pass sliced sweeps in factor order with FP64 values: y and z share z's storage and are told
    // apart by tags (no sentinel re-arming); z must carry tag 0 on entry (+0.0 does) and leaves with tag 0
    bool tagged = false;
    // tagged, IC0, with cg: the backward sweep also applies x += alpha p and p = z + beta p (the fused
    // PCG then launches no separate vector kernel)
    TV* xpx = nullptr;
    TV* xpp = nullptr;
};
// z = M^-1 r with vectors of type TV. fp32 (TV = double): IC0/ILU0 with the factor stored in FP32 and the
// solves computed in FP64. Returns
// the number of fused partials written (0 if none: single-pass sweeps, z != r and a large enough cap are
// needed).
template <typename TV>
int precond_apply_t(Context* c, pfem_precond_t* pc, TV* z, const TV* r, bool fp32, const ApplyOpts<TV>& opt);
void jacobi_setup(Context* c, const Pattern& P, const double* val, double* invd, float* invd32);
void jacobi_apply(Context* c, int n, double* dest, const double* src, const double* invd);
void jacobi_apply32(Context* c, int n, float* dest, const float* src, const float* invd);
void cg_solve_fp64(Context* c, pfem_solver_t* s, pfem_csr_t* a, double* x, const double* b, pfem_solve_stats* out);
void gmres_solve_fp64(Context* c, pfem_solver_t* s, pfem_csr_t* a, double* x, const double* b,
                      pfem_solve_stats* out);
void solve_mixed(Context* c, pfem_solver_t* s, pfem_csr_t* a, double* x, const double* b, pfem_solve_stats* out);
void dirichlet_apply(Context* c, pfem_dirichlet_t* bc0, pfem_dirichlet_t* bc1, pfem_csr_t* a, double* rhs0,
                     double* rhs1, int mode);
}  // namespace pfem
