// Preconditioners: Jacobi, IC0 and ILU0 (PreconditionerJacobi / PreconditionerIC / PreconditionerILU,
// src/linear_algebra/preconditioners.cu). The reference calls cusparseDcsric02 / csrilu02 and SpSV; here
// the zero-fill factorizations and the triangular solves are level-scheduled HIP kernels: rows whose
// dependencies are complete form a level, one lane computes one row, exactly the row recurrence of the
// sequential algorithm (so the factor is bit-identical to a sequential CPU factorization of the same
// matrix). With PFEM_ORDER_MULTICOLOR the matrix is first permuted symmetrically by a greedy colouring,
// so a level is a colour.
//
// Applies run each triangular solve as ONE launch (single-pass sweep): one workgroup per chunk of <= 256
// rows of one level, chunks numbered level by level. A row publishes its result with an agent-scope
// atomic store into a buffer that holds a sentinel NaN until then; a row that needs it reads it with an
// ordinary load and, only while it still sees the sentinel, polls with agent-scope atomic loads. No fence
// is involved (the value is its own flag), and a workgroup waits only on chunks with smaller indices, so
// the grid drains in dispatch order. Every wait has a spin limit; exceeding it sets an error flag that
// the host turns into PFEM_ENUMERIC instead of waiting for ever. The arithmetic per row is that of the
// per-level kernels, so both forms give bit-identical results.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>

#include "device.hpp"
#include "internal.hpp"
#include "solver.hpp"
#include "sync.hpp"

namespace pfem {

// ------------------------------------------------------------------ kernels
__global__ void k_gather_values(long nnz, const int* __restrict__ src_pos, const double* __restrict__ a,
                                double* __restrict__ f)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) f[k] = a[src_pos[k]];
}

// IC0, row i (left-looking): L_ik = (a_ik - sum_{j<k} L_ij L_kj) / L_kk for k < i, then
// L_ii = sqrt(a_ii - sum L_ik^2). f holds A's values on entry; lower entries become L.
__global__ void k_ic0_level(int cnt, const int* __restrict__ rows, const int* __restrict__ rowptr,
                            const int* __restrict__ col, const int* __restrict__ diag, double* __restrict__ f,
                            int* status)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    const int si = rowptr[i], di = diag[i];
    for (int p = si; p < di; ++p) {
        const int k = col[p];
        double s = f[p];
        int q = rowptr[k];
        const int dk = diag[k];
        for (int pp = si; pp < p; ++pp) {
            const int j = col[pp];
            while (q < dk && col[q] < j) ++q;
            if (q < dk && col[q] == j) s -= f[pp] * f[q];
        }
        f[p] = s / f[dk];
    }
    double d = f[di];
    for (int p = si; p < di; ++p) d -= f[p] * f[p];
    if (!(d > 0.0)) atomicCAS(status, 0, i + 1);
    f[di] = sqrt(d);
}

// ILU0 (IKJ), row i: for k < i: a_ik /= a_kk; a_ij -= a_ik a_kj for j > k in both patterns.
__global__ void k_ilu0_level(int cnt, const int* __restrict__ rows, const int* __restrict__ rowptr,
                             const int* __restrict__ col, const int* __restrict__ diag, double* __restrict__ f,
                             int* status)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    const int si = rowptr[i], ei = rowptr[i + 1];
    for (int p = si; p < ei; ++p) {
        const int k = col[p];
        if (k >= i) break;
        const int dk = diag[k];
        const double lik = f[p] / f[dk];
        f[p] = lik;
        int q = dk + 1;
        const int ek = rowptr[k + 1];
        for (int pp = p + 1; pp < ei; ++pp) {
            const int j = col[pp];
            while (q < ek && col[q] < j) ++q;
            if (q < ek && col[q] == j) f[pp] -= lik * f[q];
        }
    }
    if (f[diag[i]] == 0.0) atomicCAS(status, 0, i + 1);
}

// IC0: copy L_ki into the upper slot (i,k) so that the backward sweep with L^T reads its row
// contiguously instead of gathering through the mirror map
__global__ void k_ic0_transpose(int n, const int* __restrict__ rowptr, const int* __restrict__ diag,
                                const int* __restrict__ mirror, double* __restrict__ f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int p = diag[i] + 1; p < rowptr[i + 1]; ++p) f[p] = f[mirror[p]];
}

// forward solve of one level: y_i = (b_i - sum_{k<i} L_ik y_k) [/ L_ii for IC0], b_i = src[perm[i]]
// (perm null: src[i]); y may alias src when perm is null
template <typename T, bool UNIT>
__global__ void k_lower_level(int cnt, const int* __restrict__ rows, const int* __restrict__ rowptr,
                              const int* __restrict__ col, const int* __restrict__ diag, const T* __restrict__ f,
                              const int* __restrict__ perm, const T* src, T* y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    const int di = diag[i];
    T s = src[perm ? perm[i] : i];
    for (int p = rowptr[i]; p < di; ++p) s -= f[p] * y[col[p]];
    y[i] = UNIT ? s : s / f[di];
}

// backward solve of one level, in place on z: z_i = (z_i - sum_{k>i} U_ik z_k) / U_ii (IC0: U = L^T,
// stored in the upper slots); with a permutation the result is also scattered to out[perm[i]]
template <typename T>
__global__ void k_upper_level(int cnt, const int* __restrict__ rows, const int* __restrict__ rowptr,
                              const int* __restrict__ col, const int* __restrict__ diag, const T* __restrict__ f,
                              const int* __restrict__ perm, T* z, T* out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    const int di = diag[i];
    T s = z[i];
    for (int p = di + 1; p < rowptr[i + 1]; ++p) s -= f[p] * z[col[p]];
    s = s / f[di];
    z[i] = s;
    if (perm) out[perm[i]] = s;
}

// ---- single-pass sweeps
// s - sum_{p0 <= p < p1} f[p] * buf[col[p]], subtracted in ascending p (the sequential order); the
// column indices, factor values and first reads of up to kDeps dependencies are issued together, and
// only dependencies that still read as the sentinel are polled
constexpr int kDeps = 8;
// Sliced layouts: each lane takes kLaneRows rows of its chunk (t, t + 256, ...), a chunk up to kSliceRows
// rows of one level, the dependency loads of the lane's rows issued together. Two rows per lane (the
// 1M-row pressure matrix's sweeps then fit the 256 CUs in one round of about 2,000 workgroups) measured
// slower: the tagged 16-bit apply 22.7 -> 40.7 us, a PCG iteration 52.7 -> 69.1 us.
constexpr int kLaneRows = 1;
constexpr int kSliceRows = kLaneRows * kBlock;

template <typename T, typename TS>
__device__ __forceinline__ T dep_sum(T s, int p0, int p1, const int* __restrict__ col, const TS* __restrict__ f,
                                     T* buf, int* err, int limit)
{
    using U = typename Sentinel<T>::U;
    constexpr U sent = Sentinel<T>::bits;
    for (int base = p0; base < p1; base += kDeps) {
        const int m = min(kDeps, p1 - base);
        int cv[kDeps];
        T fv[kDeps];
        U yv[kDeps];
#pragma unroll
        for (int k = 0; k < kDeps; ++k)
            if (k < m) {
                cv[k] = col[base + k];
                fv[k] = (T)f[base + k];
            }
#pragma unroll
        for (int k = 0; k < kDeps; ++k)
            if (k < m) yv[k] = __hip_atomic_load(reinterpret_cast<U*>(buf + cv[k]), __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int k = 0; k < kDeps; ++k)
            if (k < m) {
                if (yv[k] == sent) yv[k] = __builtin_bit_cast(U, await_value(buf + cv[k], err, limit));
                s -= fv[k] * __builtin_bit_cast(T, yv[k]);
            }
    }
    return s;
}

// Arguments of the single-pass sweeps (passed by value).
template <typename TV, typename TS, typename TF>
struct SweepArgs {
    const int* chunks;   // {first row (or position in rows), count} per workgroup
    const int* rows;     // rows by level (unused when the chunks are row ranges)
    const int* ptr;      // row offsets of this sweep's triangle in the split factor
    const int* col;
    const unsigned short* col16;   // sliced layout: 16-bit column codes, decoded with cbase (4 per chunk)
    const int* cbase;
    const TS* f;         // split factor values (off-diagonal entries)
    const TS* d;         // diagonal values
    const int* perm;     // factor row -> vector index (null: vectors in the factor's ordering)
    TF* sy;              // forward results (sentinel until published)
    TF* sw;              // backward results
    int* err;            // set when a wait exceeds its spin limit
    int limit;
    const int* done;     // the Krylov solve's done flag: both sweeps skip once it is set
    const TV* src;       // forward: right-hand side b
    // forward, fused PCG update (x != null, factor ordering): x += alpha p, r -= alpha q, then b = r
    TV* x;
    TV* r;
    const TV* p;
    const TV* q;
    const double* alpha;
    // backward: output and the optional (r, z) partials; with cg, the partials are published into
    // sentinel-armed slots and the last workgroup ends PCG iteration k (cg_record)
    TV* out;
    const TV* rdot;
    double* part;
    CgState* cg;
    const CgState* acg;   // non-null: alpha = acg->alpha[acg->it & 1] (the device's iteration index)
    // backward, tagged, fused IC0-PCG: the x and p updates of the iteration (null: a separate kernel)
    TV* xu;
    TV* pu;
    // 16-bit factors are stored as c L (c a power of two that puts the largest entry near 2^14); the
    // forward sweep then solves for c^2 b, which leaves z unchanged. in_scale[1] = c^2.
    const double* in_scale;
};

// forward solve, all levels: y_i = (b_i - sum_{k<i} L_ik y_k) [/ L_ii for IC0], b_i = src[perm[i]];
// also re-arms sw for the backward solve.
// TV is the type of the vectors, TS that of the stored factor, TF that of the arithmetic and of the
// published values (FP64 vectors over an FP32-stored factor: the mixed-precision preconditioner).
// Chunk layouts: kRows (positions in the level's row list, CSR triangles), kRange (a row range, CSR
// triangles), kSliced (a row range, the chunk-sliced triangle: header {first row, rows, base, width}).
constexpr int kRows = 0, kRange = 1, kSliced = 2;

// The calling lane's rows: one row (CSR layouts), or kLaneRows rows t, t + 256, ... of a sliced chunk
// (header {first row, rows, base, width}); e0, stride and width locate a sliced row's entries.
template <int LAYOUT, int R, typename TV, typename TS, typename TF>
__device__ __forceinline__ bool sweep_rows(const SweepArgs<TV, TS, TF>& a, int (&i)[R], int (&e0)[R], bool (&has)[R],
                                           int& stride, int& width)
{
    bool any = false;
    if constexpr (LAYOUT == kSliced) {
        const int* h = a.chunks + 4 * hdr_slot();
        const int cnt = h[1];
        stride = cnt;
        width = h[3];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int t = (int)threadIdx.x + r * kBlock;
            i[r] = h[0] + t;
            e0[r] = h[2] + t;
            has[r] = t < cnt;
            any |= has[r];
        }
    } else {
        static_assert(R == 1, "CSR layouts take one row per lane");
        const int c0 = a.chunks[2 * blockIdx.x], cnt = a.chunks[2 * blockIdx.x + 1];
        has[0] = (int)threadIdx.x < cnt;
        i[0] = LAYOUT == kRange ? c0 + (int)threadIdx.x : (has[0] ? a.rows[c0 + threadIdx.x] : 0);
        e0[0] = stride = width = 0;
        any = has[0];
    }
    return any;
}

// s[r] - sum_k T_ik buf[k] for each of the lane's rows, subtracted in ascending k (the sequential order).
// Sliced rows: the column codes, factor values and first reads of the dependencies of all the lane's rows
// (kDeps / R entries each) are issued together before any is consumed. WAIT < 0: buf holds the sentinel
// until published; WAIT = 0 or 1: buf is a tagged buffer (FP64), a dependency is ready once it carries
// tag WAIT.
template <int LAYOUT, int WAIT, int R, typename TV, typename TS, typename TF>
__device__ __forceinline__ void sweep_sums(const SweepArgs<TV, TS, TF>& a, TF (&s)[R], const int (&i)[R],
                                          const int (&e0)[R], const bool (&has)[R], int stride, int width, TF* buf)
{
    if constexpr (LAYOUT == kSliced) {
        using U = typename Sentinel<TF>::U;
        constexpr U sent = Sentinel<TF>::bits;
        constexpr int K = kDeps / R;
        const ChunkBases cb = a.col16 ? chunk_bases(a.cbase + 4 * hdr_slot()) : ChunkBases{0, 0, 0, 0, true};
        for (int w0 = 0; w0 < width; w0 += K) {
            int cv[R][K];
            TF fv[R][K];
            U yv[R][K];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    cv[r][k] = -1;
                    if (has[r] && w0 + k < width) {
                        const int e = e0[r] + (w0 + k) * stride;
                        cv[r][k] = slot_col(e, a.col, a.col16, cb);
                        fv[r][k] = (TF)a.f[e];
                    }
                }
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (cv[r][k] >= 0)
                        yv[r][k] = __hip_atomic_load(reinterpret_cast<U*>(buf + cv[r][k]), __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (cv[r][k] >= 0) {
                        if constexpr (WAIT < 0) {
                            if (yv[r][k] == sent) yv[r][k] = __builtin_bit_cast(U, await_value(buf + cv[r][k], a.err, a.limit));
                        } else {
                            static_assert(sizeof(TF) == 8, "tagged buffers hold FP64 values");
                            if ((yv[r][k] & 1ull) != (U)WAIT)
                                yv[r][k] = __builtin_bit_cast(U, await_tagged<WAIT>(buf + cv[r][k], a.err, a.limit));
                        }
                        s[r] -= fv[r][k] * __builtin_bit_cast(TF, yv[r][k]);
                    }
        }
    } else {
        if (has[0]) s[0] = dep_sum(s[0], a.ptr[i[0]], a.ptr[i[0] + 1], a.col, a.f, buf, a.err, a.limit);
    }
}

template <int LAYOUT>
constexpr int lane_rows()
{
    return LAYOUT == kSliced ? kLaneRows : 1;
}

// TAGGED (sliced layout, FP64 published values): sy and sw are one buffer; the forward sweep publishes y
// with tag 1, the backward sweep replaces it by z with tag 0, and neither re-arms anything. z keeps tag 0
// so that an exact zero (a Dirichlet row's) stays exactly zero.
//
// In the fused PCG with an IC0 factor (TAGGED, XR, not UNIT, a.cg set) this sweep also ends the iteration:
// with M = L L^T, gamma_{k+1} = (r_{k+1}, M^-1 r_{k+1}) = (y, y) for y = L^-1 r_{k+1}, so the chunks
// publish their partials of (y, y) and the last workgroup adds them and records the iteration
// (cg_record), which the backward sweep would otherwise do from (r, z) after reading r once more. A
// 16-bit factor stored as c L gives c y here: the sum is divided by c^2 (a power of two, exact).
template <typename TV, typename TS, typename TF, bool UNIT, int LAYOUT, bool XR, bool TAGGED = false>
__global__ __launch_bounds__(kBlock) void k_sweep_lower(SweepArgs<TV, TS, TF> a)
{
    constexpr bool kYY = TAGGED && XR && !UNIT;
    constexpr int R = lane_rows<LAYOUT>();
    __shared__ double wsum[kWaves];
    __shared__ unsigned arrived;
    if (a.done && *a.done) return;
    const bool yy = kYY && a.cg;
    if (yy) {
        if (threadIdx.x == 0) arrived = 0u;
        __syncthreads();
    }
    int i[R], e0[R], stride, width;
    bool has[R];
    const bool any = sweep_rows<LAYOUT, R>(a, i, e0, has, stride, width);
    if (!yy && !any) return;
    double v = 0.0;
    if (any) {
        TF s[R];
        TV al = 0;
        if constexpr (XR) al = (TV)(a.acg ? a.acg->alpha[a.acg->it & 1] : *a.alpha);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!has[r]) continue;
            TV b;
            if constexpr (XR) {
                // updateXR (linear_solver.cu:68-77) for this row, as the separate vector kernel computes it
                // (the x half is left to the p-update kernel when a.x is null)
                if (a.x) a.x[i[r]] += al * a.p[i[r]];
                b = a.r[i[r]] - al * a.q[i[r]];
                a.r[i[r]] = b;
            } else {
                b = a.src[a.perm ? a.perm[i[r]] : i[r]];
            }
            TF b0 = (TF)b;
            if constexpr (sizeof(TS) == 2) b0 *= (TF)a.in_scale[1];
            s[r] = b0;
        }
        if constexpr (TAGGED) {
            static_assert(LAYOUT == kSliced && sizeof(TF) == 8, "tagged sweeps: sliced layout, FP64 values");
            sweep_sums<LAYOUT, 1>(a, s, i, e0, has, stride, width, a.sy);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!has[r]) continue;
                const TF y = UNIT ? s[r] : s[r] / (TF)a.d[i[r]];
                publish_tagged<1>(a.sy + i[r], y);
                v += (double)y * (double)y;
            }
        } else {
            sweep_sums<LAYOUT, -1>(a, s, i, e0, has, stride, width, a.sy);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!has[r]) continue;
                publish(a.sy + i[r], UNIT ? s[r] : s[r] / (TF)a.d[i[r]]);
                a.sw[i[r]] = sentinel_value<TF>();
            }
        }
    }
    if constexpr (kYY) {
        if (!yy) return;
        // the backward sweep's reduction: per-wave sums in LDS, the last wave to arrive adds them in wave
        // order and publishes the chunk's partial into its sentinel-armed slot
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) {
            wsum[threadIdx.x >> 6] = v;
            __threadfence_block();
            if (atomicAdd(&arrived, 1u) == kWaves - 1) {
                __threadfence_block();
                double t = wsum[0];
                for (int w = 1; w < kWaves; ++w) t += wsum[w];
                publish(a.part + blockIdx.x, t);
            }
        }
        if (blockIdx.x == gridDim.x - 1) {
            __syncthreads();
            double g = collect_partials(a.part, gridDim.x, a.err, a.limit);
            if constexpr (sizeof(TS) == 2) g /= a.in_scale[1];
            // iteration k = it (it counts the iterations ended so far)
            if (threadIdx.x == 0) cg_record(a.cg, g, a.cg->it);
        }
    }
}

// backward solve, all levels: z_i = (y_i - sum_{k>i} U_ik z_k) / U_ii, out[perm[i]] = z_i; re-arms sy.
// part != null: the chunk's partial of (r, z) (r = rdot, in the vectors' order) goes to part[chunk].
//
template <typename TV, typename TS, typename TF, int LAYOUT, bool TAGGED = false>
__global__ __launch_bounds__(kBlock) void k_sweep_upper(SweepArgs<TV, TS, TF> a)
{
    constexpr int R = lane_rows<LAYOUT>();
    // (r, z) partial without a workgroup barrier after the sweep: each wave leaves its sum in LDS and the
    // last wave to arrive adds the four in wave order (the order block_sum uses)
    __shared__ double wsum[kWaves];
    __shared__ unsigned arrived;
    int i[R], e0[R], stride, width;
    bool has[R];
    if (a.done && *a.done) return;
    if (a.part) {
        if (threadIdx.x == 0) arrived = 0u;
        __syncthreads();
    }
    double v = 0.0;
    if (sweep_rows<LAYOUT, R>(a, i, e0, has, stride, width)) {
        TF s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!has[r]) continue;
            s[r] = a.sy[i[r]];
            if constexpr (!TAGGED) a.sy[i[r]] = sentinel_value<TF>();
        }
        sweep_sums<LAYOUT, TAGGED ? 0 : -1>(a, s, i, e0, has, stride, width, a.sw);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!has[r]) continue;
            TF zr;
            if constexpr (TAGGED) {
                zr = tagged<0>(s[r] / (TF)a.d[i[r]]);
                publish_tagged<0>(a.sw + i[r], zr);
            } else {
                zr = s[r] / (TF)a.d[i[r]];
                publish(a.sw + i[r], zr);
            }
            const int o = a.perm ? a.perm[i[r]] : i[r];
            const TV z = (TV)zr;
            if ((const void*)a.out != (const void*)a.sw) a.out[o] = z;
            if (a.part) v += (double)(a.rdot[o] * z);
        }
    }
    if (!a.part) return;
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) {
        wsum[threadIdx.x >> 6] = v;
        __threadfence_block();
        if (atomicAdd(&arrived, 1u) == kWaves - 1) {
            __threadfence_block();
            double t = wsum[0];
            for (int w = 1; w < kWaves; ++w) t += wsum[w];
            if (a.cg) publish(a.part + blockIdx.x, t);
            else a.part[blockIdx.x] = t;
        }
    }
    if (a.cg && blockIdx.x == gridDim.x - 1) {
        // gamma_{k+1} = (r, z) and the stopping test, which the separate p-update kernel does otherwise
        __syncthreads();
        const double g = collect_partials(a.part, gridDim.x, a.err, a.limit);
        // iteration k = it (it counts the iterations ended so far)
        if (threadIdx.x == 0) cg_record(a.cg, g, a.cg->it);
    }
}

// The fused IC0-PCG's backward sweep (tagged, sliced): z = L^-T y published with tag 0, and for each row
// as soon as its z is known the end-of-iteration vector updates of iteration k = it - 1 (the forward sweep
// has recorded gamma and beta): x += alpha_k p_k (the x half of updateXR, linear_solver.cu:68-77) and
// p = z + beta p (updateP, :79-87), with the operations of k_cg_xp_recorded, so no separate vector pass
// runs. When iteration k stopped the solve only the x update runs, and a launch finding it equal to the
// xp_seen its SpMV left does nothing.
template <typename TV, typename TS, typename TF>
__global__ __launch_bounds__(kBlock) void k_sweep_upper_xp(SweepArgs<TV, TS, TF> a)
{
    constexpr int R = kLaneRows;
    int i[R], e0[R], stride, width;
    bool has[R];
    const CgState* st = a.acg;
    const int it = st->it;
    if (it == st->xp_seen) return;
    const int k = it - 1;
    const bool last = st->done;
    const TV al = (TV)st->alpha[k & 1];
    if (!sweep_rows<kSliced, R>(a, i, e0, has, stride, width)) return;
    if (last) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (has[r]) a.xu[i[r]] += al * a.pu[i[r]];
        return;
    }
    const TV beta = (TV)(st->gam[(k + 1) & 1] / st->gam[k & 1]);
    TF z[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (has[r]) z[r] = a.sy[i[r]];
    sweep_sums<kSliced, 0>(a, z, i, e0, has, stride, width, a.sw);
    // z is published first (the rows of lower levels wait on it); p_k is read after the sweep
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!has[r]) continue;
        z[r] = tagged<0>(z[r] / (TF)a.d[i[r]]);
        publish_tagged<0>(a.sw + i[r], z[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!has[r]) continue;
        const TV pi = a.pu[i[r]];
        a.xu[i[r]] += al * pi;
        a.pu[i[r]] = (TV)z[r] + beta * pi;
    }
}

// ------------------------------------------------------------------ blocked sweeps (PFEM_ORDER_BLOCKS)
// A lane per row, as the multicolour sweeps: each row first subtracts its entries outside its block
// (other colours, ascending, waiting on tags), keeping the coefficients of its block's own entries; then
// the block's rows finish one after the other inside the wave, each value passed to the block's other
// rows by a shuffle. A row is finished by the reciprocal of its diagonal (one division per row rather
// than one per step of the block), so values differ from a division by rounding. Forward, the terms of
// every row come in the sequential order (the block's columns are the row's largest); backward, the
// block's entries (the smallest columns) come after the outside ones instead of before them.

// entries of a row loaded together by the blocked sweeps (fewer than kDeps: the block's coefficients take
// registers of their own)
constexpr int kBlockDeps = 4;

// Outside entries of the calling lane's row subtracted from s (tagged buffer, ready once a value carries
// tag WAIT); the coefficients of the row's in-block entries go to cf[t'] with bit t' of mask.
template <int WAIT, typename TV, typename TS, typename TF>
__device__ __forceinline__ void block_row_sums(const SweepArgs<TV, TS, TF>& a, TF& s, TF (&cf)[kBlockRows],
                                               unsigned& mask, int e0, int stride, int width, TF* buf)
{
    using U = unsigned long long;
    constexpr int K = kBlockDeps;
    for (int w0 = 0; w0 < width; w0 += K) {
        int cv[K];
        TF fv[K];
        U yv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            cv[k] = -1;
            if (w0 + k < width) {
                const int e = e0 + (w0 + k) * stride;
                cv[k] = a.col[e];
                fv[k] = (TF)a.f[e];
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (cv[k] >= 0)
                yv[k] = __hip_atomic_load(reinterpret_cast<U*>(buf + cv[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (cv[k] >= 0) {
                if ((yv[k] & 1ull) != (U)WAIT) yv[k] = __builtin_bit_cast(U, await_tagged<WAIT>(buf + cv[k], a.err, a.limit));
                s -= fv[k] * __builtin_bit_cast(TF, yv[k]);
            } else if (cv[k] <= -2) {
                const int tp = -2 - cv[k];
#pragma unroll
                for (int q = 0; q < kBlockRows; ++q)
                    if (q == tp) cf[q] = fv[k];
                mask |= 1u << tp;
            }
        }
    }
}

__device__ __forceinline__ double shfl_f64(double v, int src)
{
    return __shfl(v, src, 64);
}

// Forward sweep, tagged (y published with tag 1). XR: the fused PCG's x/r update of the row first (as
// k_sweep_lower); with an IC0 factor and a.cg set the chunks also publish their partials of (y, y) and the
// last workgroup records the iteration (as k_sweep_lower).
template <typename TV, typename TS, typename TF, bool UNIT, bool XR>
__global__ __launch_bounds__(kBlock) void k_bsweep_lower(SweepArgs<TV, TS, TF> a)
{
    static_assert(sizeof(TF) == 8 && sizeof(TV) == 8, "tagged sweeps: FP64 vectors and values");
    constexpr bool kYY = XR && !UNIT;
    __shared__ double wsum[kWaves];
    __shared__ unsigned arrived;
    if (a.done && *a.done) return;
    const bool yy = kYY && a.cg;
    if (yy) {
        if (threadIdx.x == 0) arrived = 0u;
        __syncthreads();
    }
    const int slot = hdr_slot();
    const int* h = a.chunks + 4 * slot;
    const int first = h[0], cnt = h[1], base = h[2], width = h[3];
    const int tc = a.cbase[4 * slot];
    const int u = (int)threadIdx.x;
    const bool has = u < cnt;
    const int t = u & (tc - 1);
    const int i = first + u;
    TF s = 0, d = 1;
    TF cf[kBlockRows];
#pragma unroll
    for (int q = 0; q < kBlockRows; ++q) cf[q] = 0;
    unsigned mask = 0;
    if (has) {
        TV b;
        if constexpr (XR) {
            const TV al = (TV)(a.acg ? a.acg->alpha[a.acg->it & 1] : *a.alpha);
            if (a.x) a.x[i] += al * a.p[i];
            b = a.r[i] - al * a.q[i];
            a.r[i] = b;
        } else {
            b = a.src[i];
        }
        s = (TF)b;
        if constexpr (sizeof(TS) == 2) s *= (TF)a.in_scale[1];
        if constexpr (!UNIT) d = (TF)a.d[i];
        block_row_sums<1>(a, s, cf, mask, base + u, cnt, width, a.sy);
    }
    // the block's rows in order: at step tau row tau finishes (times the reciprocal of its diagonal,
    // computed once) and the block's later rows subtract its term; lb: the block's first lane in the wave
    const int lb = (u & 63) - t;
    const TF dinv = UNIT ? (TF)1 : (TF)1 / d;
    TF yv = 0;
#pragma unroll
    for (int tau = 0; tau < kBlockRows; ++tau) {
        if (tau >= tc) break;
        if (t == tau) yv = UNIT ? s : s * dinv;
        const TF yt = shfl_f64(tagged<1>(yv), lb + tau);   // what a dependent row reads from the buffer
        if (mask & (1u << tau)) s -= cf[tau] * yt;
    }
    double v = 0.0;
    if (has) {
        publish_tagged<1>(a.sy + i, yv);
        v = (double)yv * (double)yv;
    }
    if constexpr (kYY) {
        if (!yy) return;
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) {
            wsum[threadIdx.x >> 6] = v;
            __threadfence_block();
            if (atomicAdd(&arrived, 1u) == kWaves - 1) {
                __threadfence_block();
                double tot = wsum[0];
                for (int w = 1; w < kWaves; ++w) tot += wsum[w];
                publish(a.part + blockIdx.x, tot);
            }
        }
        if (blockIdx.x == gridDim.x - 1) {
            __syncthreads();
            double g = collect_partials(a.part, gridDim.x, a.err, a.limit);
            if constexpr (sizeof(TS) == 2) g /= a.in_scale[1];
            if (threadIdx.x == 0) cg_record(a.cg, g, a.cg->it);
        }
    }
}

// Backward sweep, tagged (z published with tag 0 in place of y). XP: the fused IC0-PCG's end-of-iteration
// x and p updates of the row (as k_sweep_upper_xp); otherwise out (when distinct) and the optional (r, z)
// partials with the last workgroup's cg_record (as k_sweep_upper).
template <typename TV, typename TS, typename TF, bool XP>
__global__ __launch_bounds__(kBlock) void k_bsweep_upper(SweepArgs<TV, TS, TF> a)
{
    static_assert(sizeof(TF) == 8 && sizeof(TV) == 8, "tagged sweeps: FP64 vectors and values");
    __shared__ double wsum[kWaves];
    __shared__ unsigned arrived;
    TV al = 0, beta = 0;
    bool last = false;
    if constexpr (XP) {
        const CgState* st = a.acg;
        const int it = st->it;
        if (it == st->xp_seen) return;
        const int kk = it - 1;
        last = st->done;
        al = (TV)st->alpha[kk & 1];
        if (!last) beta = (TV)(st->gam[(kk + 1) & 1] / st->gam[kk & 1]);
    } else {
        if (a.done && *a.done) return;
        if (a.part) {
            if (threadIdx.x == 0) arrived = 0u;
            __syncthreads();
        }
    }
    const int slot = hdr_slot();
    const int* h = a.chunks + 4 * slot;
    const int first = h[0], cnt = h[1], base = h[2], width = h[3];
    const int tc = a.cbase[4 * slot];
    const int u = (int)threadIdx.x;
    const bool has = u < cnt;
    const int t = u & (tc - 1);
    const int i = first + u;
    if (XP && last) {
        if (has) a.xu[i] += al * a.pu[i];
        return;
    }
    TF s = 0, d = 1;
    TF cf[kBlockRows];
#pragma unroll
    for (int q = 0; q < kBlockRows; ++q) cf[q] = 0;
    unsigned mask = 0;
    if (has) {
        s = a.sy[i];
        d = (TF)a.d[i];
        block_row_sums<0>(a, s, cf, mask, base + u, cnt, width, a.sw);
    }
    const int lb = (u & 63) - t;
    const TF dinv = (TF)1 / d;
    TF zmine = 0;
#pragma unroll
    for (int tau = kBlockRows - 1; tau >= 0; --tau) {
        if (tau >= tc) continue;
        if (t == tau) zmine = tagged<0>(s * dinv);
        const TF zt = shfl_f64(zmine, lb + tau);
        if (mask & (1u << tau)) s -= cf[tau] * zt;
    }
    double v = 0.0;
    if (has) {
        publish_tagged<0>(a.sw + i, zmine);
        if constexpr (XP) {
            const TV pi = a.pu[i];
            a.xu[i] += al * pi;
            a.pu[i] = (TV)zmine + beta * pi;
        } else {
            if ((const void*)a.out != (const void*)a.sw) a.out[i] = (TV)zmine;
            if (a.part) v = (double)(a.rdot[i] * (TV)zmine);
        }
    }
    if constexpr (!XP) {
        if (!a.part) return;
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) {
            wsum[threadIdx.x >> 6] = v;
            __threadfence_block();
            if (atomicAdd(&arrived, 1u) == kWaves - 1) {
                __threadfence_block();
                double tot = wsum[0];
                for (int w = 1; w < kWaves; ++w) tot += wsum[w];
                if (a.cg) publish(a.part + blockIdx.x, tot);
                else a.part[blockIdx.x] = tot;
            }
        }
        if (a.cg && blockIdx.x == gridDim.x - 1) {
            __syncthreads();
            const double g = collect_partials(a.part, gridDim.x, a.err, a.limit);
            if (threadIdx.x == 0) cg_record(a.cg, g, a.cg->it);
        }
    }
}

template <typename T>
__global__ void k_copyT(long n, T* __restrict__ d, const T* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

__global__ void k_gather_or_zero(long n, const int* __restrict__ pos, const double* __restrict__ a,
                                 double* __restrict__ f)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) f[k] = pos[k] >= 0 ? a[pos[k]] : 0.0;
}

__global__ void k_f64_to_f32(long n, float* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

// Largest |s_i| into *mx (bit patterns of non-negative doubles order as the values do; *mx starts at 0).
__global__ void k_absmax(long n, const double* __restrict__ s, unsigned long long* mx)
{
    double m = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = fmax(m, fabs(s[i]));
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(mx, __builtin_bit_cast(unsigned long long, m));
}

// sc[0] = c = 2^(14 - e) with max = f 2^e, 0.5 <= f < 1 (so max c lies in [2^13, 2^14)), sc[1] = c^2;
// c = 1 for an all-zero or non-finite factor. One thread.
__global__ void k_f16_scale(const unsigned long long* mx, double* sc)
{
    const double m = __builtin_bit_cast(double, *mx);
    int e = 0;
    if (m > 0.0 && m < INFINITY) (void)frexp(m, &e);
    else e = 14;
    const double c = ldexp(1.0, 14 - e);
    sc[0] = c;
    sc[1] = c * c;
}

__global__ void k_f64_to_f16(long n, _Float16* __restrict__ d, const double* __restrict__ s, const double* __restrict__ sc)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (_Float16)(s[i] * sc[0]);
}

__global__ void k_zero_u64(unsigned long long* p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

// Diagonal entries of the 16-bit copy that came out zero or subnormal (|c L_ii| < 2^-14): those rows would
// divide by (nearly) nothing. Counted into *lost (zeroed before).
__global__ void k_f16_lost_diag(int n, const _Float16* __restrict__ d, int* lost)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(fabs((double)d[i]) >= 6.103515625e-05)) atomicAdd(lost, 1);
}

// ------------------------------------------------------------------ host setup
// Reverse Cuthill-McKee order of the matrix graph: breadth-first from a pseudo-peripheral node of each
// connected component (the last node reached by a BFS from the component's minimum-degree node),
// neighbours visited by ascending degree (ties by index), the whole sequence reversed.
std::vector<int> rcm_order(const Pattern& A)
{
    const int n = A.n;
    std::vector<int> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = A.h_rowptr[i + 1] - A.h_rowptr[i] - 1;
    std::vector<int> order;
    order.reserve(n);
    std::vector<char> seen(n, 0);
    std::vector<int> q, nb, last(n, -1);
    auto bfs = [&](int root, std::vector<int>& out, int stamp) {
        out.clear();
        out.push_back(root);
        last[root] = stamp;
        for (size_t h = 0; h < out.size(); ++h) {
            const int u = out[h];
            nb.clear();
            for (int k = A.h_rowptr[u]; k < A.h_rowptr[u + 1]; ++k) {
                const int v = A.h_col[k];
                if (v != u && last[v] != stamp && !seen[v]) {
                    last[v] = stamp;
                    nb.push_back(v);
                }
            }
            std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
            out.insert(out.end(), nb.begin(), nb.end());
        }
    };
    std::vector<int> by_degree(n);
    std::iota(by_degree.begin(), by_degree.end(), 0);
    std::stable_sort(by_degree.begin(), by_degree.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    int stamp = 0;
    for (int s0 : by_degree) {
        if (seen[s0]) continue;
        bfs(s0, q, stamp++);
        const int root = q.back();   // pseudo-peripheral
        bfs(root, q, stamp++);
        for (int v : q) seen[v] = 1;
        order.insert(order.end(), q.begin(), q.end());
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Reduces a proper colouring to K colours where a few stragglers exceed it (planar node graphs of
// triangle meshes are 4-colourable, and greedy leaves a handful of nodes at colour 4). A straggler v
// takes colour a after a Kempe interchange: a <-> b swapped in the union of the {a, b}-components of
// v's a-coloured neighbours, allowed when that union holds none of v's b-coloured neighbours (every
// such swap keeps the colouring proper). When no pair works, the chain through one neighbour is swapped
// and the search repeats, a bounded number of times. Deterministic; stragglers that resist keep their
// colour. More than max_bad stragglers are left as they are. Returns the colour count.
static int reduce_colours(const Pattern& A, std::vector<int>& colour, int ncol, int K, size_t max_bad)
{
    const int n = A.n;
    std::vector<int> bad;
    for (int i = 0; i < n; ++i)
        if (colour[i] >= K) bad.push_back(i);
    if (bad.empty() || bad.size() > max_bad) return ncol;
    constexpr size_t kCap = 1u << 18;   // largest chain considered
    std::vector<int> seen(n, -1), comp, nb;
    std::vector<char> forbid(n, 0);
    int stamp = 0;
    // union of the {a, b}-components of the start nodes; false if it reaches a forbidden node or the cap
    auto chain = [&](const std::vector<int>& starts, int a, int b) {
        ++stamp;
        comp.clear();
        for (int u : starts)
            if (seen[u] != stamp) {
                seen[u] = stamp;
                comp.push_back(u);
            }
        for (size_t h = 0; h < comp.size(); ++h) {
            const int u = comp[h];
            for (int k = A.h_rowptr[u]; k < A.h_rowptr[u + 1]; ++k) {
                const int w = A.h_col[k];
                if (w == u || seen[w] == stamp || (colour[w] != a && colour[w] != b)) continue;
                if (forbid[w] || comp.size() >= kCap) return false;
                seen[w] = stamp;
                comp.push_back(w);
            }
        }
        return true;
    };
    auto swap_chain = [&](int a, int b) {
        for (int u : comp) colour[u] = colour[u] == a ? b : a;
    };
    auto neighbours = [&](int v) {
        nb.clear();
        for (int k = A.h_rowptr[v]; k < A.h_rowptr[v + 1]; ++k)
            if (A.h_col[k] != v) nb.push_back(A.h_col[k]);
    };
    auto fix = [&](int v) {
        neighbours(v);
        for (int a = 0; a < K; ++a) {
            std::vector<int> na;
            for (int u : nb)
                if (colour[u] == a) na.push_back(u);
            if (na.empty()) {
                colour[v] = a;
                return true;
            }
            for (int b = 0; b < K; ++b) {
                if (b == a) continue;
                for (int u : nb)
                    if (colour[u] == b) forbid[u] = 1;
                const bool ok = chain(na, a, b);
                for (int u : nb) forbid[u] = 0;
                if (ok) {
                    swap_chain(a, b);
                    colour[v] = a;
                    return true;
                }
            }
        }
        return false;
    };
    for (int v : bad) {
        if (colour[v] < K || fix(v)) continue;
        neighbours(v);
        const std::vector<int> ring = nb;
        for (int t = 0; t < 64 && colour[v] >= K; ++t) {
            const int u = ring[t % ring.size()];
            const int a = colour[u];
            if (a >= K) continue;
            const int b = (a + 1 + (t / (int)ring.size()) % (K - 1)) % K;
            if (chain(std::vector<int>{u}, a, b)) swap_chain(a, b);
            fix(v);
        }
    }
    int m = 0;
    for (int i = 0; i < n; ++i) m = std::max(m, colour[i] + 1);
    return m;
}

// PFEM_ORDER_BLOCKS. Blocks of at most kBlockRows nodes, each grown breadth-first over unassigned nodes
// from a seed taken in reverse Cuthill-McKee order (on a triangulation nearly every block is full and
// compact), a block of another size split into power-of-two pieces in its own node order; the block
// graph coloured greedily in block order, reduced to 4 colours where Kempe interchanges allow. New order:
// colour by colour; within a colour, chunks of up to kBlock rows made of blocks of one size tc (full
// blocks first, in block order), a block's nodes consecutive, so a block's rows are tc neighbouring
// lanes of one wave. A zero-fill factor depends only on which of two coupled nodes comes first, which is
// what the order (colour, block, node within block) fixes. chunks[colour] lists {first row, rows | tc << 16}.
static void block_order(const Pattern& A, std::vector<int>& perm, std::vector<int>& cptr,
                        std::vector<std::vector<std::pair<int, int>>>& chunks)
{
    const int n = A.n;
    const std::vector<int> order = rcm_order(A);
    std::vector<int> blk(n, -1), bptr{0}, bnode, grown;
    bnode.reserve(n);
    for (int s : order) {
        if (blk[s] >= 0) continue;
        grown.assign(1, s);
        blk[s] = 0;
        for (size_t h = 0; h < grown.size() && grown.size() < (size_t)kBlockRows; ++h) {
            const int u = grown[h];
            for (int k = A.h_rowptr[u]; k < A.h_rowptr[u + 1] && grown.size() < (size_t)kBlockRows; ++k) {
                const int v = A.h_col[k];
                if (blk[v] < 0) {
                    blk[v] = 0;
                    grown.push_back(v);
                }
            }
        }
        // power-of-two pieces, largest first
        size_t o = 0;
        for (int piece = kBlockRows; piece >= 1; piece >>= 1)
            while (grown.size() - o >= (size_t)piece) {
                const int b = (int)bptr.size() - 1;
                for (int q = 0; q < piece; ++q) {
                    blk[grown[o + q]] = b;
                    bnode.push_back(grown[o + q]);
                }
                o += piece;
                bptr.push_back((int)bnode.size());
            }
    }
    const int nb = (int)bptr.size() - 1;
    // block graph (host CSR only; each row lists the block itself too)
    Pattern B;
    B.n = nb;
    B.h_rowptr.assign(nb + 1, 0);
    {
        std::vector<int> seen(nb, -1), row;
        for (int b = 0; b < nb; ++b) {
            row.clear();
            for (int q = bptr[b]; q < bptr[b + 1]; ++q) {
                const int u = bnode[q];
                for (int k = A.h_rowptr[u]; k < A.h_rowptr[u + 1]; ++k) {
                    const int cb = blk[A.h_col[k]];
                    if (seen[cb] != b) {
                        seen[cb] = b;
                        row.push_back(cb);
                    }
                }
            }
            std::sort(row.begin(), row.end());
            B.h_col.insert(B.h_col.end(), row.begin(), row.end());
            B.h_rowptr[b + 1] = (int)B.h_col.size();
        }
    }
    std::vector<int> colour(nb, -1), mark;
    int ncol = 0;
    for (int b = 0; b < nb; ++b) {
        mark.assign(ncol + 1, 0);
        for (int k = B.h_rowptr[b]; k < B.h_rowptr[b + 1]; ++k) {
            const int j = B.h_col[k];
            if (j != b && colour[j] >= 0) mark[colour[j]] = 1;
        }
        int cc = 0;
        while (cc < ncol && mark[cc]) ++cc;
        colour[b] = cc;
        if (cc == ncol) ++ncol;
    }
    // a colour is a dependency level of the sweeps: worth a long search where the block graph is planar
    // (a few per cent of stragglers on triangulations); a graph with many is left as coloured (3D meshes)
    ncol = reduce_colours(B, colour, ncol, 4, (size_t)nb / 10 + 16);
    for (int b = 0; b < nb; ++b)
        for (int k = B.h_rowptr[b]; k < B.h_rowptr[b + 1]; ++k)
            PFEM_REQUIRE(B.h_col[k] == b || colour[B.h_col[k]] != colour[b], "block colouring is not proper");
    perm.clear();
    perm.reserve(n);
    cptr.assign(1, 0);
    chunks.assign(ncol, {});
    std::vector<std::vector<int>> by_size(kBlockRows + 1);
    for (int cc = 0; cc < ncol; ++cc) {
        for (auto& v : by_size) v.clear();
        for (int b = 0; b < nb; ++b)
            if (colour[b] == cc) by_size[bptr[b + 1] - bptr[b]].push_back(b);
        for (int size = kBlockRows; size >= 1; size >>= 1) {
            const std::vector<int>& bl = by_size[size];
            const size_t per = (size_t)(kBlock / size);
            for (size_t q = 0; q < bl.size(); q += per) {
                const int m = (int)std::min(per, bl.size() - q);
                chunks[cc].emplace_back((int)perm.size(), (m * size) | (size << 16));
                for (int l = 0; l < m; ++l)
                    for (int t = 0; t < size; ++t) perm.push_back(bnode[bptr[bl[q + l]] + t]);
            }
        }
        cptr.push_back((int)perm.size());
    }
    PFEM_REQUIRE((int)perm.size() == n, "block ordering lost a node");
}

// Chunks of each level, dealt to workgroup slots XCD by XCD: workgroup b runs on XCD b mod 8, and each XCD
// receives one contiguous band of the level's chunks (in the level's own order, which follows the mesh).
// The reordering stays inside the level, so every workgroup still waits only on lower-numbered ones.
// Returns the flat {first, second} pairs in workgroup order.
static std::vector<int> deal_chunks(const std::vector<std::vector<std::pair<int, int>>>& levels)
{
    constexpr int kXcds = 8;
    std::vector<int> ch;
    for (const auto& level : levels) {
        const int m = (int)level.size();
        const int base = (int)ch.size() / 2;   // workgroup index of the level's first chunk
        std::vector<int> next(kXcds), end(kXcds);
        for (int x = 0; x < kXcds; ++x) {
            next[x] = (int)((long)m * x / kXcds);
            end[x] = (int)((long)m * (x + 1) / kXcds);
        }
        std::vector<char> used(m, 0);
        int spill = 0;   // lowest chunk not yet dealt, for slots whose band is exhausted
        for (int b = 0; b < m; ++b) {
            const int x = (base + b) % kXcds;
            int q = -1;
            while (next[x] < end[x] && used[next[x]]) ++next[x];
            if (next[x] < end[x]) q = next[x]++;
            else {
                while (used[spill]) ++spill;
                q = spill;
            }
            used[q] = 1;
            ch.push_back(level[q].first);
            ch.push_back(level[q].second);
        }
    }
    return ch;
}

// Chunk q's header and base entries (4 ints each) moved to hdr_slot's position for workgroup q (grid =
// number of chunks).
static void to_slot_order(std::vector<int>& hdr, std::vector<int>& cbase)
{
    const int G = (int)(hdr.size() / 4), q8 = G >> 3, rem = G & 7;
    std::vector<int> h2(hdr.size()), c2(cbase.size());
    for (int b = 0; b < G; ++b) {
        const int x = b & 7, sl = x * q8 + std::min(x, rem) + (b >> 3);
        for (int k = 0; k < 4; ++k) {
            h2[4 * sl + k] = hdr[4 * b + k];
            c2[4 * sl + k] = cbase[4 * b + k];
        }
    }
    hdr.swap(h2);
    cbase.swap(c2);
}

static void build_factor_structure(Context* c, pfem_precond_t* pc)
{
    const Pattern& A = *pc->pat;
    const int n = A.n;
    std::vector<int> perm(n), iperm(n);
    std::vector<int> cptr;   // multicolour, blocks: colour c occupies rows [cptr[c], cptr[c+1]) of the new order
    std::vector<std::vector<std::pair<int, int>>> bchunks;   // blocks: the chunks of each colour
    const bool blocked = pc->ordering == PFEM_ORDER_BLOCKS;
    if (pc->ordering == PFEM_ORDER_MULTICOLOR) {
        // greedy colouring visiting the nodes in reverse Cuthill-McKee order, then Kempe-chain reduction of
        // the few nodes above colour 3; new order = (colour, RCM position). On the cylinder channel this
        // gives 4 colours where greedy in natural order gives 6: the colours are the sweeps' dependency
        // levels, so two fewer waits per triangular solve.
        const std::vector<int> order = rcm_order(A);
        std::vector<int> colour(n, -1);
        std::vector<int> mark;
        int ncol = 0;
        for (int oi = 0; oi < n; ++oi) {
            const int i = order[oi];
            mark.assign(ncol + 1, 0);
            for (int k = A.h_rowptr[i]; k < A.h_rowptr[i + 1]; ++k) {
                const int j = A.h_col[k];
                if (j != i && colour[j] >= 0) mark[colour[j]] = 1;
            }
            int cc = 0;
            while (cc < ncol && mark[cc]) ++cc;
            colour[i] = cc;
            if (cc == ncol) ++ncol;
        }
        ncol = reduce_colours(A, colour, ncol, 4, (size_t)n / 100 + 16);
        for (int i = 0; i < n; ++i)
            for (int k = A.h_rowptr[i]; k < A.h_rowptr[i + 1]; ++k)
                PFEM_REQUIRE(A.h_col[k] == i || colour[A.h_col[k]] != colour[i], "colouring is not proper");
        perm = order;
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return colour[a] < colour[b]; });
        cptr.assign(ncol + 1, 0);
        for (int i = 0; i < n; ++i) cptr[colour[i] + 1]++;
        for (int k = 0; k < ncol; ++k) cptr[k + 1] += cptr[k];
    } else if (pc->ordering == PFEM_ORDER_BLOCKS) {
        block_order(A, perm, cptr, bchunks);
    } else {
        std::iota(perm.begin(), perm.end(), 0);
    }
    for (int i = 0; i < n; ++i) iperm[perm[i]] = i;
    pc->h_perm = perm;
    // permuted pattern and the source positions of its values
    std::vector<int> rowptr(n + 1, 0), col, src;
    col.reserve(A.nnz);
    src.reserve(A.nnz);
    std::vector<std::pair<int, int>> buf;
    for (int i = 0; i < n; ++i) {
        const int o = perm[i];
        buf.clear();
        for (int k = A.h_rowptr[o]; k < A.h_rowptr[o + 1]; ++k) buf.emplace_back(iperm[A.h_col[k]], k);
        std::sort(buf.begin(), buf.end());
        for (auto& e : buf) { col.push_back(e.first); src.push_back(e.second); }
        rowptr[i + 1] = (int)col.size();
    }
    if (pc->ordering == PFEM_ORDER_NATURAL) pc->fpat = pc->pat;
    else pc->fpat = make_pattern(c, n, std::vector<int>(rowptr), std::vector<int>(col));
    const Pattern& F = *pc->fpat;
    for (int i = 0; i < n; ++i) PFEM_REQUIRE(F.h_diag[i] >= 0, "IC0/ILU0 needs every diagonal entry in the pattern");
    // forward levels: longest-path levels of the lower triangle; in the multicolour order a colour is an
    // independent set whose lower neighbours all have smaller colours, so the colours are the levels
    // (contiguous row ranges, forward in colour order and backward in reverse colour order). The blocked
    // order keeps the longest-path levels (a colour holds several: the rows of a block depend on each other)
    const bool colour_levels = !cptr.empty() && !blocked;
    std::vector<int> lev(n, 0);
    int nl = 0;
    for (int i = 0; i < n; ++i) {
        int l = 0;
        for (int k = F.h_rowptr[i]; k < F.h_diag[i]; ++k) l = std::max(l, lev[F.h_col[k]] + 1);
        lev[i] = l;
        nl = std::max(nl, l + 1);
    }
    if (colour_levels) {
        nl = (int)cptr.size() - 1;
        for (int k = 0; k < nl; ++k)
            for (int i = cptr[k]; i < cptr[k + 1]; ++i) lev[i] = k;
    }
    std::vector<int> lptr(nl + 1, 0), lrows(n);
    for (int i = 0; i < n; ++i) lptr[lev[i] + 1]++;
    for (int l = 0; l < nl; ++l) lptr[l + 1] += lptr[l];
    {
        std::vector<int> fill(lptr.begin(), lptr.end() - 1);
        for (int i = 0; i < n; ++i) lrows[fill[lev[i]]++] = i;
    }
    // backward levels
    std::vector<int> ulev(n, 0);
    int nu = 0;
    for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        for (int k = F.h_diag[i] + 1; k < F.h_rowptr[i + 1]; ++k) l = std::max(l, ulev[F.h_col[k]] + 1);
        ulev[i] = l;
        nu = std::max(nu, l + 1);
    }
    if (colour_levels) {
        nu = nl;
        for (int k = 0; k < nl; ++k)
            for (int i = cptr[k]; i < cptr[k + 1]; ++i) ulev[i] = nl - 1 - k;
    }
    std::vector<int> uptr(nu + 1, 0), urows(n);
    for (int i = 0; i < n; ++i) uptr[ulev[i] + 1]++;
    for (int l = 0; l < nu; ++l) uptr[l + 1] += uptr[l];
    {
        std::vector<int> fill(uptr.begin(), uptr.end() - 1);
        for (int i = 0; i < n; ++i) urows[fill[ulev[i]]++] = i;
    }
    // mirror positions: slot (i,k) -> slot (k,i) (-1 when the pattern is not symmetric there)
    std::vector<int> mirror(F.nnz, -1);
    for (int i = 0; i < n; ++i)
        for (int k = F.h_rowptr[i]; k < F.h_rowptr[i + 1]; ++k) {
            const int j = F.h_col[k];
            const int* b = F.h_col.data() + F.h_rowptr[j];
            const int* e = F.h_col.data() + F.h_rowptr[j + 1];
            const int* it = std::lower_bound(b, e, i);
            if (it != e && *it == i) mirror[k] = (int)(it - F.h_col.data());
        }
    if (pc->kind == PFEM_PRECOND_IC0)
        for (int i = 0; i < n; ++i)
            for (int k = F.h_diag[i] + 1; k < F.h_rowptr[i + 1]; ++k)
                PFEM_REQUIRE(mirror[k] >= 0, "IC0 needs a structurally symmetric pattern");
    pc->nlevels = std::max(nl, nu);
    pc->h_lptr = lptr;
    pc->h_uptr = uptr;
    upload(c, pc->lrows, lrows);
    upload(c, pc->urows, urows);
    upload(c, pc->mirror, mirror);
    upload(c, pc->src_pos, src);
    upload(c, pc->perm, perm);
    upload(c, pc->iperm, iperm);
    // chunk = {first position in rows, count}; with colour levels every level is a contiguous row range
    // and a chunk stores {first row, count} instead
    //
    // Within a level the chunks are dealt to workgroup slots XCD by XCD: workgroup b runs on XCD b mod 8,
    // and each XCD receives one contiguous band of the level's chunks. The rows of every colour are in
    // RCM order, so band x of one colour and band x of the next cover the same region of the mesh, and a
    // row's dependencies were published on its own XCD (its L2) rather than on any of the eight. The
    // reordering stays inside the level, so every workgroup still waits only on lower-numbered ones.
    auto chunk_table = [&](const std::vector<int>& ptr, const std::vector<int>& rows, int rc) {
        std::vector<std::vector<std::pair<int, int>>> levels(ptr.size() - 1);
        for (size_t l = 0; l + 1 < ptr.size(); ++l)
            for (int o = ptr[l]; o < ptr[l + 1]; o += rc)
                levels[l].emplace_back(colour_levels ? rows[o] : o, std::min(rc, ptr[l + 1] - o));
        return deal_chunks(levels);
    };
    const std::vector<int> lch = chunk_table(lptr, lrows, kBlock), uch = chunk_table(uptr, urows, kBlock);
    pc->contiguous = colour_levels;
    pc->nlchunks = (int)lch.size() / 2;
    pc->nuchunks = (int)uch.size() / 2;
    upload(c, pc->lchunks, lch);
    upload(c, pc->uchunks, uch);
    {
        // split copy for the sweeps: a forward (backward) sweep then fetches only the lower (upper) triangle
        std::vector<int> sptr(2 * ((size_t)n + 1)), scol, spos;
        scol.reserve(F.nnz);
        spos.reserve(F.nnz);
        for (int t = 0; t < 2; ++t)
            for (int i = 0; i < n; ++i) {
                sptr[(size_t)t * (n + 1) + i] = (int)scol.size();
                const int k0 = t == 0 ? F.h_rowptr[i] : F.h_diag[i] + 1;
                const int k1 = t == 0 ? F.h_diag[i] : F.h_rowptr[i + 1];
                for (int k = k0; k < k1; ++k) {
                    scol.push_back(F.h_col[k]);
                    spos.push_back(k);
                }
                if (i == n - 1) sptr[(size_t)t * (n + 1) + n] = (int)scol.size();
            }
        pc->split_diag = (long)scol.size();
        for (int i = 0; i < n; ++i) spos.push_back(F.h_diag[i]);
        PFEM_REQUIRE((long)spos.size() == F.nnz, "factor pattern with a repeated diagonal entry");
        if (scol.empty()) scol.push_back(0);   // keep the device array non-null for a diagonal-only factor
        upload(c, pc->split_ptr, sptr);
        upload(c, pc->split_col, scol);
        upload(c, pc->split_pos, spos);
        pc->split.alloc(F.nnz);
    }
    if (colour_levels || blocked) {
        // Sliced layouts: a lane finds its row's entries from its chunk's header alone (no row offsets to
        // load first), and the loads of one entry index are unit-stride over the chunk. Column indices
        // are stored in 16 bits where a chunk allows it: 2 bits of colour and a 14-bit offset from the
        // chunk's smallest column in that colour (0xFFFF marks an unused slot); a chunk whose columns
        // of one colour span more than 16383 rows keeps 32-bit indices (cbase[0] = -1).
        const int ncol = (int)cptr.size() - 1;
        auto slice = [&](const std::vector<int>& ranges, int part, SlicedCols& out, DevArray<double>& val_d) {
            std::vector<int> hdr, col, pos, cbase;
            std::vector<unsigned short> c16;
            for (size_t q = 0; q < ranges.size(); q += 2) {
                const int r0 = ranges[q], cnt = ranges[q + 1];
                auto k0 = [&](int i) { return part == 1 ? F.h_diag[i] + 1 : F.h_rowptr[i]; };
                auto k1 = [&](int i) { return part == 0 ? F.h_diag[i] : F.h_rowptr[i + 1]; };
                int width = 0;
                for (int i = r0; i < r0 + cnt; ++i) width = std::max(width, k1(i) - k0(i));
                const int base = (int)col.size();
                col.resize(col.size() + (size_t)width * cnt, -1);
                pos.resize(col.size(), -1);
                c16.resize(col.size(), 0xFFFF);
                int lo[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX}, hi[4] = {-1, -1, -1, -1};
                bool narrow = ncol <= 4;
                for (int t = 0; t < cnt; ++t) {
                    const int i = r0 + t;
                    for (int k = k0(i); k < k1(i); ++k) {
                        const size_t e = base + (size_t)(k - k0(i)) * cnt + t;
                        col[e] = F.h_col[k];
                        pos[e] = k;
                        if (narrow) {
                            const int cc = (int)(std::upper_bound(cptr.begin(), cptr.end(), F.h_col[k]) - cptr.begin()) - 1;
                            lo[cc] = std::min(lo[cc], F.h_col[k]);
                            hi[cc] = std::max(hi[cc], F.h_col[k]);
                        }
                    }
                }
                for (int cc = 0; cc < 4 && narrow; ++cc)
                    if (hi[cc] >= 0 && hi[cc] - lo[cc] > 0x3FFE) narrow = false;
                if (narrow) {
                    for (int t = 0; t < cnt; ++t) {
                        const int i = r0 + t;
                        for (int k = k0(i); k < k1(i); ++k) {
                            const size_t e = base + (size_t)(k - k0(i)) * cnt + t;
                            const int cc = (int)(std::upper_bound(cptr.begin(), cptr.end(), F.h_col[k]) - cptr.begin()) - 1;
                            c16[e] = (unsigned short)((cc << 14) | (F.h_col[k] - lo[cc]));
                        }
                    }
                    for (int cc = 0; cc < 4; ++cc) cbase.push_back(hi[cc] >= 0 ? lo[cc] : 0);
                } else {
                    cbase.insert(cbase.end(), {-1, 0, 0, 0});
                }
                hdr.insert(hdr.end(), {r0, cnt, base, width});
            }
            if (col.empty()) {
                col.push_back(-1);
                pos.push_back(-1);
                c16.push_back(0xFFFF);
            }
            to_slot_order(hdr, cbase);
            upload(c, out.hdr, hdr);
            upload(c, out.cbase, cbase);
            upload(c, out.col, col);
            upload(c, out.col16, c16);
            upload(c, out.pos, pos);
            val_d.alloc(col.size());
        };
        // Blocked layouts: the chunk-sliced triangle of each chunk (as above, a lane per row), with the
        // columns inside the row's own block coded -2 - t' (t' the column's position in the block);
        // cbase per chunk: {tc, 0, 0, 0}. A column inside the chunk's rows must be in the row's block.
        auto slice_blocks = [&](const std::vector<int>& ch, int part, SlicedCols& out, DevArray<double>& val_d) {
            std::vector<int> hdr, col, pos, cbase;
            for (size_t q = 0; q < ch.size(); q += 2) {
                const int r0 = ch[q], cnt = ch[q + 1] & 0xFFFF, tc = ch[q + 1] >> 16;
                PFEM_REQUIRE(cnt >= 1 && cnt <= kBlock && tc >= 1 && tc <= kBlockRows && (tc & (tc - 1)) == 0 &&
                                 cnt % tc == 0,
                             "bad block chunk");
                auto k0 = [&](int i) { return part == 1 ? F.h_diag[i] + 1 : F.h_rowptr[i]; };
                auto k1 = [&](int i) { return part == 0 ? F.h_diag[i] : F.h_rowptr[i + 1]; };
                int width = 0;
                for (int i = r0; i < r0 + cnt; ++i) width = std::max(width, k1(i) - k0(i));
                const int base = (int)col.size();
                col.resize(col.size() + (size_t)width * cnt, -1);
                pos.resize(col.size(), -1);
                for (int u = 0; u < cnt; ++u) {
                    const int i = r0 + u, b0 = i - u % tc;
                    for (int k = k0(i); k < k1(i); ++k) {
                        const int j = F.h_col[k];
                        const size_t e = base + (size_t)(k - k0(i)) * cnt + u;
                        if (j >= r0 && j < r0 + cnt) {
                            PFEM_REQUIRE(j >= b0 && j < b0 + tc, "block ordering: a coupling across blocks of one colour");
                            col[e] = -2 - (j - b0);
                        } else {
                            col[e] = j;
                        }
                        pos[e] = k;
                    }
                }
                hdr.insert(hdr.end(), {r0, cnt, base, width});
                cbase.insert(cbase.end(), {tc, 0, 0, 0});
            }
            if (col.empty()) {
                col.push_back(-1);
                pos.push_back(-1);
            }
            to_slot_order(hdr, cbase);
            std::vector<unsigned short> c16(1, 0xFFFF);
            upload(c, out.hdr, hdr);
            upload(c, out.cbase, cbase);
            upload(c, out.col, col);
            upload(c, out.col16, c16);
            upload(c, out.pos, pos);
            val_d.alloc(col.size());
        };
        if (blocked) {
            // the forward sweep takes the colours in order, the backward sweep in reverse order
            const std::vector<int> lch2 = deal_chunks(bchunks);
            const std::vector<int> uch2 = deal_chunks(std::vector<std::vector<std::pair<int, int>>>(bchunks.rbegin(), bchunks.rend()));
            pc->nlslices = (int)lch2.size() / 2;
            pc->nuslices = (int)uch2.size() / 2;
            slice_blocks(lch2, 0, pc->sl_lower, pc->sell_l);
            slice_blocks(uch2, 1, pc->sl_upper, pc->sell_u);
            pc->blocked = true;
        } else {
            // the sliced sweeps take kSliceRows rows per chunk (kLaneRows per lane)
            const std::vector<int> lch2 = chunk_table(lptr, lrows, kSliceRows), uch2 = chunk_table(uptr, urows, kSliceRows);
            pc->nlslices = (int)lch2.size() / 2;
            pc->nuslices = (int)uch2.size() / 2;
            slice(lch2, 0, pc->sl_lower, pc->sell_l);
            slice(uch2, 1, pc->sl_upper, pc->sell_u);
        }
        pc->sell = true;
        // The whole matrix over the SpMV tiles, dealt by region across colours: XCD x (workgroups b = x mod
        // 8) takes the tiles of the x-th eighth of every colour's range. A row's neighbours lie in the other
        // colours at about the same relative position, so they are gathered from the L2 of the XCD that
        // also computes them, where a band of the factor order would span one colour only.
        std::vector<int> ftiles(2 * (size_t)F.ntiles);
        PFEM_HIP(hipMemcpy(ftiles.data(), F.tiles.p, ftiles.size() * sizeof(int), hipMemcpyDeviceToHost));
        {
            constexpr int kXcds = 8;
            std::vector<std::pair<int, int>> t2;
            for (int t = 0; t < F.ntiles; ++t) t2.emplace_back(ftiles[2 * t], ftiles[2 * t + 1]);
            std::sort(t2.begin(), t2.end());
            std::vector<std::vector<std::pair<int, int>>> region(kXcds);
            for (const auto& tr : t2) {
                const int cc = (int)(std::upper_bound(cptr.begin(), cptr.end(), tr.first) - cptr.begin()) - 1;
                const long size = std::max(1, cptr[cc + 1] - cptr[cc]);
                const int x = (int)std::min<long>(kXcds - 1, (long)(tr.first - cptr[cc]) * kXcds / size);
                region[x].push_back(tr);
            }
            std::vector<size_t> next(kXcds, 0);
            size_t spill_x = 0;
            for (int b = 0; b < F.ntiles; ++b) {
                int x = b % kXcds;
                if (next[x] == region[x].size()) {   // this region is used up: take from the next non-empty one
                    while (next[spill_x] == region[spill_x].size()) ++spill_x;
                    x = (int)spill_x;
                }
                const auto tr = region[x][next[x]++];
                ftiles[2 * b] = tr.first;
                ftiles[2 * b + 1] = tr.second;
            }
            // the CSR tile kernels of the factor's pattern use the same order (the same partials, so every
            // form of the fused PCG gives the same bits)
            upload(c, pc->fpat->tiles, ftiles);
            for (int b = 0; b < F.ntiles; ++b) ftiles[2 * b + 1] -= ftiles[2 * b];
        }
        slice(ftiles, 2, pc->sl_all, pc->asl_val);
    }
    pc->fac.alloc(F.nnz);
    pc->tmp.alloc(n);
    pc->sy.alloc(n);
    pc->sw.alloc(n);
    hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, (long)n, pc->sy.p);
    hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, (long)n, pc->sw.p);
    pc->status.alloc(2);   // first failing row (+1); diagonals lost to the 16-bit copy
    zero_async(c, pc->status.p, 2 * sizeof(int));
    pc->sweep_err.alloc(1);
    zero_async(c, pc->sweep_err.p, sizeof(int));
}

void precond_initialize(Context* c, pfem_precond_t* pc, pfem_csr_t* a)
{
    PFEM_REQUIRE(a->pat->n == pc->n, "preconditioner and matrix sizes differ");
    if (pc->kind == PFEM_PRECOND_NONE) return;
    if (pc->kind == PFEM_PRECOND_CUSTOM) {
        if (pc->custom_init) pc->custom_init(pc->custom_user, a);
        return;
    }
    if (pc->kind == PFEM_PRECOND_JACOBI) {
        jacobi_setup(c, *a->pat, a->val.p, pc->invd.p, pc->invd32.p);
        return;
    }
    PFEM_REQUIRE(a->pat.get() == pc->pat.get() ||
                     (a->pat->h_rowptr == pc->pat->h_rowptr && a->pat->h_col == pc->pat->h_col),
                 "IC0/ILU0 initialised from a matrix with another pattern");
    const Pattern& F = *pc->fpat;
    hipLaunchKernelGGL(k_gather_values, dim3(blocks_for(F.nnz)), dim3(kBlock), 0, c->stream, (long)F.nnz,
                       pc->src_pos.p, a->val.p, pc->fac.p);
    if (pc->ordering != PFEM_ORDER_NATURAL) {
        // A on the permuted pattern, for Krylov iterations run in the factor's ordering
        if (!pc->aval.p) pc->aval.alloc(F.nnz);
        hipLaunchKernelGGL((k_copyT<double>), dim3(blocks_for(F.nnz)), dim3(kBlock), 0, c->stream, (long)F.nnz,
                           pc->aval.p, (const double*)pc->fac.p);
        if (pc->asl_val.p) {
            const long m = (long)pc->asl_val.n;
            hipLaunchKernelGGL(k_gather_or_zero, dim3(blocks_for(m)), dim3(kBlock), 0, c->stream, m, pc->sl_all.pos.p,
                               (const double*)pc->aval.p, pc->asl_val.p);
        }
    }
    for (size_t l = 0; l + 1 < pc->h_lptr.size(); ++l) {
        const int cnt = pc->h_lptr[l + 1] - pc->h_lptr[l];
        if (pc->kind == PFEM_PRECOND_IC0)
            hipLaunchKernelGGL(k_ic0_level, dim3(blocks_for(cnt)), dim3(kBlock), 0, c->stream, cnt,
                               pc->lrows.p + pc->h_lptr[l], F.rowptr.p, F.col.p, F.diag.p, pc->fac.p, pc->status.p);
        else
            hipLaunchKernelGGL(k_ilu0_level, dim3(blocks_for(cnt)), dim3(kBlock), 0, c->stream, cnt,
                               pc->lrows.p + pc->h_lptr[l], F.rowptr.p, F.col.p, F.diag.p, pc->fac.p, pc->status.p);
    }
    if (pc->kind == PFEM_PRECOND_IC0)
        hipLaunchKernelGGL(k_ic0_transpose, dim3(blocks_for(F.n)), dim3(kBlock), 0, c->stream, F.n, F.rowptr.p,
                           F.diag.p, pc->mirror.p, pc->fac.p);
    hipLaunchKernelGGL(k_gather_values, dim3(blocks_for(F.nnz)), dim3(kBlock), 0, c->stream, (long)F.nnz,
                       pc->split_pos.p, pc->fac.p, pc->split.p);
    if (pc->sell) {
        const long nl = (long)pc->sell_l.n, nu = (long)pc->sell_u.n;
        hipLaunchKernelGGL(k_gather_or_zero, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream, nl, pc->sl_lower.pos.p,
                           pc->fac.p, pc->sell_l.p);
        hipLaunchKernelGGL(k_gather_or_zero, dim3(blocks_for(nu)), dim3(kBlock), 0, c->stream, nu, pc->sl_upper.pos.p,
                           pc->fac.p, pc->sell_u.p);
    }
    if (pc->fac32.p) {
        hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(F.nnz)), dim3(kBlock), 0, c->stream, (long)F.nnz,
                           pc->fac32.p, pc->fac.p);
        hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(F.nnz)), dim3(kBlock), 0, c->stream, (long)F.nnz,
                           pc->split32.p, pc->split.p);
        if (pc->sell) {
            const long nl = (long)pc->sell_l.n, nu = (long)pc->sell_u.n;
            hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream, nl, pc->sell_l32.p,
                               pc->sell_l.p);
            hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(nu)), dim3(kBlock), 0, c->stream, nu, pc->sell_u32.p,
                               pc->sell_u.p);
        }
    }
    if (pc->split16.p) {
        // the scale from the largest factor entry, then c L in 16 bits
        const long nl = (long)pc->sell_l.n, nu = (long)pc->sell_u.n;
        auto* mx = reinterpret_cast<unsigned long long*>(pc->scale16.p + 2);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, mx);
        hipLaunchKernelGGL(k_absmax, dim3(std::min<long>(blocks_for(F.nnz), 1024)), dim3(kBlock), 0, c->stream,
                           (long)F.nnz, (const double*)pc->split.p, mx);
        hipLaunchKernelGGL(k_f16_scale, dim3(1), dim3(1), 0, c->stream, mx, pc->scale16.p);
        hipLaunchKernelGGL(k_f64_to_f16, dim3(blocks_for(F.nnz)), dim3(kBlock), 0, c->stream, (long)F.nnz,
                           pc->split16.p, pc->split.p, pc->scale16.p);
        hipLaunchKernelGGL(k_f64_to_f16, dim3(blocks_for(nl)), dim3(kBlock), 0, c->stream, nl, pc->sell_l16.p,
                           pc->sell_l.p, pc->scale16.p);
        hipLaunchKernelGGL(k_f64_to_f16, dim3(blocks_for(nu)), dim3(kBlock), 0, c->stream, nu, pc->sell_u16.p,
                           pc->sell_u.p, pc->scale16.p);
        // a diagonal lost to the 16-bit range (the scale follows the largest entry, so this happens when
        // the factor's entries span more than about 2^27): this factor is applied from its 32-bit copy
        int* lost = pc->status.p + 1;
        zero_async(c, lost, sizeof(int));
        hipLaunchKernelGGL(k_f16_lost_diag, dim3(blocks_for(F.n)), dim3(kBlock), 0, c->stream, F.n,
                           (const _Float16*)(pc->split16.p + pc->split_diag), lost);
        int h = 0;
        d2h_sync(c, &h, lost, sizeof(int));
        pc->half_ok = h == 0;
    }
    PFEM_HIP(hipGetLastError());
}

// Per-level launches (TV == TF only). Natural ordering: in place on out (src may equal out). Multicolour:
// the forward pass gathers src through the permutation into work, the backward pass scatters into out.
template <typename T>
static void level_solves(Context* c, pfem_precond_t* pc, const T* src, T* work, T* out, const T* f)
{
    const Pattern& F = *pc->fpat;
    const bool ic = pc->kind == PFEM_PRECOND_IC0;
    const int* perm = pc->ordering == PFEM_ORDER_NATURAL ? nullptr : pc->perm.p;
    T* y = perm ? work : out;
    for (size_t l = 0; l + 1 < pc->h_lptr.size(); ++l) {
        const int cnt = pc->h_lptr[l + 1] - pc->h_lptr[l];
        const int* rows = pc->lrows.p + pc->h_lptr[l];
        if (ic)
            hipLaunchKernelGGL((k_lower_level<T, false>), dim3(blocks_for(cnt)), dim3(kBlock), 0, c->stream, cnt, rows,
                               F.rowptr.p, F.col.p, F.diag.p, f, perm, src, y);
        else
            hipLaunchKernelGGL((k_lower_level<T, true>), dim3(blocks_for(cnt)), dim3(kBlock), 0, c->stream, cnt, rows,
                               F.rowptr.p, F.col.p, F.diag.p, f, perm, src, y);
    }
    for (size_t l = 0; l + 1 < pc->h_uptr.size(); ++l) {
        const int cnt = pc->h_uptr[l + 1] - pc->h_uptr[l];
        const int* rows = pc->urows.p + pc->h_uptr[l];
        hipLaunchKernelGGL((k_upper_level<T>), dim3(blocks_for(cnt)), dim3(kBlock), 0, c->stream, cnt, rows,
                           F.rowptr.p, F.col.p, F.diag.p, f, perm, y, out);
    }
}

// The two single-pass sweeps; o.part also produces the (r, z) partials (one per backward chunk), o.x
// fuses the PCG x/r update into the forward sweep.
template <typename TV, typename TS, typename TF>
static void sweeps(Context* c, pfem_precond_t* pc, const TV* src, TV* out, const TS* f, TF* sy, TF* sw, double* part,
                   const ApplyOpts<TV>& o)
{
    const Pattern& F = *pc->fpat;
    const bool ic = pc->kind == PFEM_PRECOND_IC0;
    SweepArgs<TV, TS, TF> a{};
    a.rows = nullptr;
    a.col = pc->split_col.p;
    a.f = f;
    a.d = f + pc->split_diag;
    a.perm = pc->ordering == PFEM_ORDER_NATURAL || o.factor_order ? nullptr : pc->perm.p;
    a.sy = sy;
    a.sw = sw;
    // in the factor's ordering with z and r distinct and z of the arithmetic's type, the backward sweep
    // publishes straight into z (the forward sweep re-arms z instead of sw): one 8-byte pass less
    if (o.factor_order && pc->contiguous && sizeof(TV) == sizeof(TF) && (const void*)out != (const void*)src)
        a.sw = reinterpret_cast<TF*>(out);
    a.err = pc->sweep_err.p;
    a.limit = pc->spin_limit;
    a.done = o.done;
    a.src = src;
    a.x = o.x;
    a.r = const_cast<TV*>(src);
    a.p = o.p;
    a.q = o.q;
    a.alpha = o.alpha;
    a.out = out;
    a.rdot = src;
    a.part = part;
    a.cg = part ? o.cg : nullptr;
    a.acg = o.alpha_cg;
    const bool xr = o.alpha != nullptr;
    PFEM_REQUIRE(!xr || (o.factor_order && (pc->contiguous || (pc->blocked && o.tagged)) && o.q && (o.p || !o.x)),
                 "the fused x/r update needs a multicolour or blocked factor and vectors in its ordering");
    dim3 gl(pc->nlchunks), gu(pc->nuchunks);
    const dim3 bl(kBlock);
    SweepArgs<TV, TS, TF> lo = a, up = a;
    lo.chunks = pc->lchunks.p;
    up.chunks = pc->uchunks.p;
    lo.ptr = pc->split_ptr.p;
    up.ptr = pc->split_ptr.p + (F.n + 1);
    if constexpr (sizeof(TS) == 2) lo.in_scale = up.in_scale = pc->scale16.p;
    if constexpr (sizeof(TF) == 8 && sizeof(TV) == 8) {
        if (pc->blocked && o.tagged) {
            // y and z share out: y with tag 1, then z with tag 0 (out must carry tag 0 on entry)
            PFEM_REQUIRE(o.factor_order && (const void*)out != (const void*)src,
                         "tagged sweeps need vectors in the factor's ordering and z distinct from r");
            lo.chunks = pc->sl_lower.hdr.p;
            up.chunks = pc->sl_upper.hdr.p;
            lo.col = pc->sl_lower.col.p;
            up.col = pc->sl_upper.col.p;
            lo.col16 = up.col16 = nullptr;
            lo.cbase = pc->sl_lower.cbase.p;
            up.cbase = pc->sl_upper.cbase.p;
            if constexpr (sizeof(TS) == 2) {
                lo.f = pc->sell_l16.p;
                up.f = pc->sell_u16.p;
            } else if constexpr (sizeof(TS) == 4) {
                lo.f = pc->sell_l32.p;
                up.f = pc->sell_u32.p;
            } else {
                lo.f = pc->sell_l.p;
                up.f = pc->sell_u.p;
            }
            lo.sy = lo.sw = up.sy = up.sw = reinterpret_cast<TF*>(out);
            gl = dim3(pc->nlslices);
            gu = dim3(pc->nuslices);
            bool xp = false;
            if (xr && ic && up.cg) {
                // as in the multicolour form: the forward sweep ends the iteration, the backward sweep
                // applies the x and p updates given o.xpx
                up.part = nullptr;
                up.cg = nullptr;
                up.xu = o.xpx;
                up.pu = o.xpp;
                xp = o.xpx != nullptr;
            } else {
                lo.cg = nullptr;
            }
            if (xr) {
                if (ic) hipLaunchKernelGGL((k_bsweep_lower<TV, TS, TF, false, true>), gl, bl, 0, c->stream, lo);
                else hipLaunchKernelGGL((k_bsweep_lower<TV, TS, TF, true, true>), gl, bl, 0, c->stream, lo);
            } else {
                if (ic) hipLaunchKernelGGL((k_bsweep_lower<TV, TS, TF, false, false>), gl, bl, 0, c->stream, lo);
                else hipLaunchKernelGGL((k_bsweep_lower<TV, TS, TF, true, false>), gl, bl, 0, c->stream, lo);
            }
            if (xp) hipLaunchKernelGGL((k_bsweep_upper<TV, TS, TF, true>), gu, bl, 0, c->stream, up);
            else hipLaunchKernelGGL((k_bsweep_upper<TV, TS, TF, false>), gu, bl, 0, c->stream, up);
            return;
        }
    }
    if (pc->contiguous && pc->sell) {
        gl = dim3(pc->nlslices);
        gu = dim3(pc->nuslices);
        lo.chunks = pc->sl_lower.hdr.p;
        up.chunks = pc->sl_upper.hdr.p;
        lo.col = pc->sl_lower.col.p;
        up.col = pc->sl_upper.col.p;
        // 16-bit column codes pay with the FP64 factor (apply 37.6 -> 28.5 us on the cylinder's pressure
        // matrix) and cost with the FP32 one (27.1 -> 30.1 us), whose entries are already narrow
        lo.col16 = sizeof(TS) == 8 ? pc->sl_lower.col16.p : nullptr;
        up.col16 = sizeof(TS) == 8 ? pc->sl_upper.col16.p : nullptr;
        lo.cbase = pc->sl_lower.cbase.p;
        up.cbase = pc->sl_upper.cbase.p;
        if constexpr (sizeof(TS) == 2) {
            // (16-bit column codes measured slower here too: apply 22.1 -> 26.6 us)
            lo.f = pc->sell_l16.p;
            up.f = pc->sell_u16.p;
            lo.in_scale = up.in_scale = pc->scale16.p;
        } else if constexpr (sizeof(TS) == 4) {
            lo.f = pc->sell_l32.p;
            up.f = pc->sell_u32.p;
        } else {
            lo.f = pc->sell_l.p;
            up.f = pc->sell_u.p;
        }
        if constexpr (sizeof(TF) == 8) {
            if (o.tagged) {
                // y and z share out: y with tag 1, then z with tag 0 (out must carry tag 0 on entry)
                PFEM_REQUIRE(o.factor_order && (const void*)out != (const void*)src && sizeof(TV) == 8,
                             "tagged sweeps need vectors in the factor's ordering and z distinct from r");
                lo.sy = lo.sw = up.sy = up.sw = reinterpret_cast<TF*>(out);
                if (xr && ic && up.cg) {
                    // IC0: the forward sweep forms gamma = (y, y) and ends the iteration; the backward
                    // sweep solves and, given o.xpx, applies the x and p updates (its partial slots and r
                    // are not touched)
                    up.part = nullptr;
                    up.cg = nullptr;
                    up.xu = o.xpx;
                    up.pu = o.xpp;
                } else {
                    lo.cg = nullptr;
                }
                if (xr) {
                    if (ic) hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, false, kSliced, true, true>), gl, bl, 0, c->stream, lo);
                    else hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, true, kSliced, true, true>), gl, bl, 0, c->stream, lo);
                } else {
                    if (ic) hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, false, kSliced, false, true>), gl, bl, 0, c->stream, lo);
                    else hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, true, kSliced, false, true>), gl, bl, 0, c->stream, lo);
                }
                if (up.xu) hipLaunchKernelGGL((k_sweep_upper_xp<TV, TS, TF>), gu, bl, 0, c->stream, up);
                else hipLaunchKernelGGL((k_sweep_upper<TV, TS, TF, kSliced, true>), gu, bl, 0, c->stream, up);
                return;
            }
        }
        PFEM_REQUIRE(!o.tagged, "tagged sweeps need the sliced layout and FP64 values");
        if (xr) {
            if (ic) hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, false, kSliced, true>), gl, bl, 0, c->stream, lo);
            else hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, true, kSliced, true>), gl, bl, 0, c->stream, lo);
        } else {
            if (ic) hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, false, kSliced, false>), gl, bl, 0, c->stream, lo);
            else hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, true, kSliced, false>), gl, bl, 0, c->stream, lo);
        }
        hipLaunchKernelGGL((k_sweep_upper<TV, TS, TF, kSliced>), gu, bl, 0, c->stream, up);
    } else if (pc->contiguous) {
        PFEM_REQUIRE(!o.tagged, "tagged sweeps need the sliced layout");
        if (xr) {
            if (ic) hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, false, kRange, true>), gl, bl, 0, c->stream, lo);
            else hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, true, kRange, true>), gl, bl, 0, c->stream, lo);
        } else {
            if (ic) hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, false, kRange, false>), gl, bl, 0, c->stream, lo);
            else hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, true, kRange, false>), gl, bl, 0, c->stream, lo);
        }
        hipLaunchKernelGGL((k_sweep_upper<TV, TS, TF, kRange>), gu, bl, 0, c->stream, up);
    } else {
        PFEM_REQUIRE(!o.tagged, "tagged sweeps need the sliced layout");
        lo.rows = pc->lrows.p;
        up.rows = pc->urows.p;
        if (ic) hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, false, kRows, false>), gl, bl, 0, c->stream, lo);
        else hipLaunchKernelGGL((k_sweep_lower<TV, TS, TF, true, kRows, false>), gl, bl, 0, c->stream, lo);
        hipLaunchKernelGGL((k_sweep_upper<TV, TS, TF, kRows>), gu, bl, 0, c->stream, up);
    }
}

template <typename TV>
int precond_apply_t(Context* c, pfem_precond_t* pc, TV* z, const TV* r, bool fp32, const ApplyOpts<TV>& opt)
{
    constexpr bool f64 = sizeof(TV) == 8;
    const int n = pc->n;
    // the partials need r intact, so an in-place apply (z == r) does not fuse them
    // (the slots serve one partial per backward chunk, or per forward chunk when the forward sweep ends a
    // PCG iteration over an IC0 factor)
    const bool fuse_ok = opt.part && z != r && pc->sweep && std::max(pc->nuchunks, pc->nlchunks) <= opt.cap;
    double* part = fuse_ok ? opt.part : nullptr;
    PFEM_REQUIRE(!opt.factor_order || pc->sweep, "factor-order apply needs the single-pass sweeps");
    switch (pc->kind) {
    case PFEM_PRECOND_NONE:
        if (z != r) hipLaunchKernelGGL((k_copyT<TV>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, (long)n, z, r);
        break;
    case PFEM_PRECOND_JACOBI:
        if constexpr (f64) jacobi_apply(c, n, z, r, pc->invd.p);
        else jacobi_apply32(c, n, z, r, pc->invd32.p);
        break;
    case PFEM_PRECOND_CUSTOM:
        if constexpr (f64) pc->custom_apply(pc->custom_user, z, r);
        else PFEM_REQUIRE(false, "a custom preconditioner applies to FP64 vectors only");
        break;
    default:
        if constexpr (f64) {
            if (fp32) {
                // factor stored in FP32, solves in FP64: an exact apply of the rounded factor
                PFEM_REQUIRE(pc->fac32.p && pc->sweep, "FP32 factor not enabled (or per-level apply selected)");
                if (pc->factor_bits == 16 && pc->half_ok) {
                    PFEM_REQUIRE(pc->split16.p && pc->sell && (pc->contiguous || pc->blocked),
                                 "a 16-bit factor needs a multicolour or blocked factor with the sliced sweep layout");
                    sweeps<double, _Float16, double>(c, pc, r, z, pc->split16.p, pc->sy.p, pc->sw.p, part, opt);
                } else {
                    sweeps<double, float, double>(c, pc, r, z, pc->split32.p, pc->sy.p, pc->sw.p, part, opt);
                }
            } else if (pc->sweep) {
                sweeps<double, double, double>(c, pc, r, z, pc->split.p, pc->sy.p, pc->sw.p, part, opt);
            } else {
                level_solves<double>(c, pc, r, pc->tmp.p, z, pc->fac.p);
            }
        } else {
            PFEM_REQUIRE(pc->fac32.p, "FP32 factor not enabled");
            if (pc->sweep) sweeps<float, float, float>(c, pc, r, z, pc->split32.p, pc->sy32.p, pc->sw32.p, part, opt);
            else level_solves<float>(c, pc, r, pc->tmp32.p, z, pc->fac32.p);
        }
    }
    PFEM_HIP(hipGetLastError());
    // partials written: one per backward launch workgroup
    const int nbw = (pc->contiguous && pc->sell) || (pc->blocked && opt.tagged) ? pc->nuslices : pc->nuchunks;
    return part && (pc->kind == PFEM_PRECOND_IC0 || pc->kind == PFEM_PRECOND_ILU0) ? nbw : 0;
}

template int precond_apply_t<double>(Context*, pfem_precond_t*, double*, const double*, bool, const ApplyOpts<double>&);
template int precond_apply_t<float>(Context*, pfem_precond_t*, float*, const float*, bool, const ApplyOpts<float>&);

void precond_apply(Context* c, pfem_precond_t* pc, double* z, const double* r)
{
    precond_apply_t<double>(c, pc, z, r, false, ApplyOpts<double>{});
}

void precond_apply32(Context* c, pfem_precond_t* pc, float* z, const float* r)
{
    precond_apply_t<float>(c, pc, z, r, true, ApplyOpts<float>{});
}

void precond_enable_fp32(Context* c, pfem_precond_t* pc)
{
    if (pc->kind == PFEM_PRECOND_JACOBI && !pc->invd32.p) {
        pc->invd32.alloc(pc->n);
    } else if ((pc->kind == PFEM_PRECOND_IC0 || pc->kind == PFEM_PRECOND_ILU0) && !pc->fac32.p) {
        pc->fac32.alloc(pc->fpat->nnz);
        pc->split32.alloc(pc->fpat->nnz);
        if (pc->sell) {
            pc->sell_l32.alloc(pc->sell_l.n);
            pc->sell_u32.alloc(pc->sell_u.n);
        }
        if (pc->factor_bits == 16 && pc->sell) {
            pc->scale16.alloc(3);
            pc->split16.alloc(pc->fpat->nnz);
            pc->sell_l16.alloc(pc->sell_l.n);
            pc->sell_u16.alloc(pc->sell_u.n);
        }
        pc->tmp32.alloc(pc->n);
        pc->sy32.alloc(pc->n);
        pc->sw32.alloc(pc->n);
        const long n = pc->n;
        hipLaunchKernelGGL((k_fill_sentinel<float>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->sy32.p);
        hipLaunchKernelGGL((k_fill_sentinel<float>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->sw32.p);
        PFEM_HIP(hipGetLastError());
    }
}

// Reports a single-pass sweep whose wait exceeded its spin limit (and re-arms the sweep buffers).
void precond_check(Context* c, pfem_precond_t* pc)
{
    if (!pc || !pc->sweep_err.p) return;
    int e = 0;
    d2h_sync(c, &e, pc->sweep_err.p, sizeof(int));
    if (!e) return;
    const long n = pc->n;
    zero_async(c, pc->sweep_err.p, sizeof(int));
    hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->sy.p);
    hipLaunchKernelGGL((k_fill_sentinel<double>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->sw.p);
    if (pc->sy32.p) {
        hipLaunchKernelGGL((k_fill_sentinel<float>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->sy32.p);
        hipLaunchKernelGGL((k_fill_sentinel<float>), dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, n, pc->sw32.p);
    }
    PFEM_HIP(hipStreamSynchronize(c->stream));
    ++pc->sweep_failures;
    throw Error(PFEM_ENUMERIC, "triangular sweep: a wait exceeded its spin limit");
}

}  // namespace pfem

using namespace pfem;

extern "C" {

int pfem_precond_create(pfem_csr a, int kind, int ordering, pfem_precond* out)
{
    PFEM_TRY(a ? a->ctx : nullptr, {
        PFEM_REQUIRE(a && out, "null argument");
        PFEM_REQUIRE(kind >= PFEM_PRECOND_NONE && kind <= PFEM_PRECOND_ILU0, "unknown preconditioner kind");
        PFEM_REQUIRE(ordering >= PFEM_ORDER_NATURAL && ordering <= PFEM_ORDER_BLOCKS, "unknown ordering");
        auto pc = std::make_unique<pfem_precond_t>();
        pc->ctx = a->ctx;
        pc->kind = kind;
        pc->ordering = ordering;
        pc->n = a->pat->n;
        pc->pat = a->pat;
        Context* c = C(a->ctx);
        if (kind == PFEM_PRECOND_JACOBI) {
            pc->invd.alloc(pc->n);
        } else if (kind == PFEM_PRECOND_IC0 || kind == PFEM_PRECOND_ILU0) {
            build_factor_structure(c, pc.get());
        }
        PFEM_HIP(hipStreamSynchronize(c->stream));
        *out = pc.release();
    });
}

int pfem_precond_create_custom(pfem_ctx ctx, int n, pfem_precond_init_fn init, pfem_precond_apply_fn apply,
                               void* user, pfem_precond* out)
{
    PFEM_TRY(ctx, {
        PFEM_REQUIRE(ctx && apply && out && n > 0, "bad arguments");
        auto pc = std::make_unique<pfem_precond_t>();
        pc->ctx = ctx;
        pc->kind = PFEM_PRECOND_CUSTOM;
        pc->n = n;
        pc->custom_init = init;
        pc->custom_apply = apply;
        pc->custom_user = user;
        *out = pc.release();
    });
}

int pfem_precond_destroy(pfem_precond pc)
{
    if (!pc) return PFEM_OK;
    (void)hipStreamSynchronize(pc->ctx->stream);
    delete pc;
    return PFEM_OK;
}

int pfem_precond_initialize(pfem_precond pc, pfem_csr a)
{
    PFEM_TRY(pc ? pc->ctx : nullptr, {
        PFEM_REQUIRE(pc && a, "null argument");
        precond_initialize(C(pc->ctx), pc, a);
        if (pc->kind == PFEM_PRECOND_IC0 || pc->kind == PFEM_PRECOND_ILU0) {
            int st = 0;
            d2h_sync(C(pc->ctx), &st, pc->status.p, sizeof(int));
            if (st) {
                zero_async(C(pc->ctx), pc->status.p, sizeof(int));
                throw Error(PFEM_ENUMERIC, "zero or negative pivot in row " + std::to_string(st - 1) +
                                               " of the (permuted) matrix");
            }
        }
    });
}

int pfem_precond_apply(pfem_precond pc, double* d_dest, const double* d_src)
{
    PFEM_TRY(pc ? pc->ctx : nullptr, {
        PFEM_REQUIRE(pc && d_dest && d_src, "null argument");
        precond_apply(C(pc->ctx), pc, d_dest, d_src);
        precond_check(C(pc->ctx), pc);
    });
}

int pfem_precond_set_sweep(pfem_precond pc, int single_pass)
{
    PFEM_TRY(pc ? pc->ctx : nullptr, {
        PFEM_REQUIRE(pc, "null argument");
        pc->sweep = single_pass ? 1 : 0;
    });
}

int pfem_precond_set_factor_bits(pfem_precond pc, int bits)
{
    PFEM_TRY(pc ? pc->ctx : nullptr, {
        PFEM_REQUIRE(pc && (bits == 16 || bits == 32), "bits must be 16 or 32");
        PFEM_REQUIRE(bits == 32 || pc->kind == PFEM_PRECOND_IC0, "16-bit storage is for IC0 factors");
        PFEM_REQUIRE(!pc->fac32.p, "set the factor's storage before the reduced-precision copy is enabled");
        pc->factor_bits = bits;
    });
}

int pfem_precond_set_spin_limit(pfem_precond pc, int polls)
{
    PFEM_TRY(pc ? pc->ctx : nullptr, {
        PFEM_REQUIRE(pc && polls >= 0, "bad arguments");
        pc->spin_limit = polls;
    });
}

int pfem_precond_download(pfem_precond pc, double* factor, int* n_levels, int* perm)
{
    PFEM_TRY(pc ? pc->ctx : nullptr, {
        PFEM_REQUIRE(pc, "null argument");
        Context* c = C(pc->ctx);
        if (n_levels) *n_levels = pc->nlevels;
        if (perm && !pc->h_perm.empty()) std::memcpy(perm, pc->h_perm.data(), pc->h_perm.size() * sizeof(int));
        if (factor) {
            if (pc->kind == PFEM_PRECOND_JACOBI) d2h_sync(c, factor, pc->invd.p, pc->invd.bytes());
            else if (pc->fac.p) d2h_sync(c, factor, pc->fac.p, pc->fac.bytes());
        }
    });
}

}  // extern "C"
