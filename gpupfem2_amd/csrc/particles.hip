// ParticleHandler2D on gfx950 (src/particles/particle_handler_2d.cu).
//
// Layout: structure of arrays (cell, lambda_1..2 as one 16-byte load, lambda_3, position, velocity, ID),
// double buffered for the reorder. The reference's 96-byte AoS record is dragged through every kernel;
// here each kernel touches only the fields it needs.
//
// Per step (advect): one kernel advects and relocates every particle through all substeps (the
// substeps only depend on the particle itself and the fixed velocity field); a particle outside its
// cell's 1-ring is marked dead. A census kernel ORs each live particle's subcell bit into a per-cell
// mask, empty subcells are refilled (scan over cells, so new particles are appended in (cell, subcell)
// order), and a stable LSD radix sort by cell (dead particles last) reorders the state. The sorted order
// is what makes the projection a deterministic segmented reduction: per cell over its contiguous
// particle range, then per node over its incident cells -- no atomics (the reference uses 9 FP64
// atomicAdd per particle, :90-107).
//
// The particle count lives on the device ({count, next ID, deleted, added} in pst): every kernel of a step
// reads it there and is launched over the buffers' capacity, which the host keeps at least the count plus
// the largest possible refill (one particle per subcell of every cell), so a step never waits for the
// device to report how many particles it added or deleted. The host learns the new count from an
// asynchronous copy read at its next use (the next step's capacity check, pfem_particles_count); by then
// the stream has long passed it. With fixed capacity and no host-side count, a step's particle kernels
// are the same launch sequence every step (two sequences: the buffers alternate), which the stepper can
// replay from a HIP graph.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"

struct pfem_particles_t {
    pfem_ctx ctx = nullptr;
    pfem_mesh2d_t* mesh = nullptr;
    int level = 2, ns = 4;
    double centers[16][3];
    long count = 0, capacity = 0;
    uint32_t next_id = 0;
    int cur = 0;
    pfem::DevArray<int> cell[2];
    pfem::DevArray<double2> lam12[2];
    pfem::DevArray<double> lam3[2];
    pfem::DevArray<double2> pos[2], vel[2];
    pfem::DevArray<uint32_t> id[2];
    pfem::DevArray<uint32_t> mask;
    pfem::DevArray<int> add, add_off, counters, cell_start;
    pfem::DevArray<uint32_t> keys[2], vals[2];
    pfem::DevArray<int> hist, scan_tmp;
    pfem::DevArray<double> part;
    long deleted_last = 0, added_last = 0;
    // device state {count, next ID, deleted, added} of the last step, its pinned host copy and the event
    // that completes the copy; generation counts reallocations of the particle buffers
    pfem::DevArray<long> pst;
    long* h_pst = nullptr;
    hipEvent_t pst_ev = nullptr;
    bool pst_pending = false;
    long generation = 0;
    ~pfem_particles_t()
    {
        if (pst_ev) (void)hipEventDestroy(pst_ev);
        if (h_pst) (void)hipHostFree(h_pst);
    }
};

namespace pfem {

struct Centers {
    double c[16][3];
};

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11)
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ double u01(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// determineSubcell (particle_handler_2d.cu:10-33), row/column clamped to the triangle (the reference
// indexes past the cell's counters for particles inside the 2e-6 tolerance band below/right of it)
__device__ __forceinline__ int subcell(int L, double step, double l1, double l2, double l3)
{
    int i = (int)((1.0 - l2) * L);
    i = i > L - 1 ? L - 1 : (i < 0 ? 0 : i);
    int j = (int)(l1 * L);
    j = j > i ? i : (j < 0 ? 0 : j);
    int res = i * i + 2 * j;
    if (j != i) {
        const double x = step * (j + 1);
        const double y = 1.0 - step * (i + 1);
        const double z = 1.0 - x - y;
        if (l3 < z) ++res;
    }
    return res;
}

// ------------------------------------------------------------------ kernels
// kSeedParticlesIntoCell (particle_handler_2d.cu:35-52), deterministic: particle k of cell c at c*ppc + k
__global__ void k_seed(int nt, int ppc, int mode, const double* __restrict__ xy, const int* __restrict__ tri,
                       Centers cen, uint2 key, int* __restrict__ cell, double2* __restrict__ lam12,
                       double* __restrict__ lam3, double2* __restrict__ pos, double2* __restrict__ vel,
                       uint32_t* __restrict__ id)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)nt * ppc) return;
    const int c = (int)(p / ppc), k = (int)(p % ppc);
    double L0, L1, L2;
    if (mode == 0) {
        L0 = cen.c[k][0]; L1 = cen.c[k][1]; L2 = cen.c[k][2];
    } else {
        const uint4 r = philox4x32_10(make_uint4((uint32_t)c, (uint32_t)k, 0u, 0u), key);
        const double s = sqrt(u01(r.x, r.y));
        const double r2 = u01(r.z, r.w);
        L0 = 1.0 - s; L1 = s * (1.0 - r2); L2 = s * r2;
    }
    const int a = tri[3 * c], b = tri[3 * c + 1], d = tri[3 * c + 2];
    cell[p] = c;
    lam12[p] = make_double2(L0, L1);
    lam3[p] = L2;
    pos[p] = make_double2(L0 * xy[2 * a] + L1 * xy[2 * b] + L2 * xy[2 * d],
                          L0 * xy[2 * a + 1] + L1 * xy[2 * b + 1] + L2 * xy[2 * d + 1]);
    vel[p] = make_double2(0.0, 0.0);
    id[p] = (uint32_t)p;
}

// kCorrectParticleVelocity (particle_handler_2d.cu:72-88); uo == nullptr is initParticleVelocity
__global__ void k_correct(const long* __restrict__ pst, const int* __restrict__ cell, const double2* __restrict__ lam12,
                          const double* __restrict__ lam3, double2* __restrict__ vel, const int* __restrict__ tri,
                          const double* __restrict__ u0, const double* __restrict__ u1,
                          const double* __restrict__ uo0, const double* __restrict__ uo1)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pst[0]) return;
    const int c = cell[p];
    const double2 l12 = lam12[p];
    const double l[3] = {l12.x, l12.y, lam3[p]};
    double ix = 0.0, iy = 0.0;
    for (int i = 0; i < 3; ++i) {
        const int v = tri[3 * c + i];
        ix += l[i] * (u0[v] - (uo0 ? uo0[v] : 0.0));
        iy += l[i] * (u1[v] - (uo1 ? uo1[v] : 0.0));
    }
    double2 vv = vel[p];
    vv.x = vv.x + ix;
    vv.y = vv.y + iy;
    vel[p] = vv;
}

// Particle2D::isInsideCell (particle_2d.cu:14-22), GEOMETRY::transformGlobalToLocal and
// isPointInsideUnitTriangle (geometry.cuh:14-46)
__device__ __forceinline__ bool inside(const double4 J, double v3x, double v3y, double px, double py, double* L)
{
    const double dx = px - v3x, dy = py - v3y;
    L[0] = J.x * dx + J.z * dy;
    L[1] = J.y * dx + J.w * dy;
    L[2] = 1.0 - L[0] - L[1];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (L[d] > 1.0 + 2e-6 || L[d] < -2e-6) return false;
    return true;
}

// All substeps of advectParticles (particle_handler_2d.cu:328-338): kAdvectParticles (:54-70) then the
// relocation of sortParticlesInCells (:363-395): own cell, then the vertex neighbours ascending, first
// hit wins; a particle found nowhere is marked dead (cell = -1).
__global__ __launch_bounds__(kBlock) void k_advect_relocate(
    const long* __restrict__ pst, int substeps, double dts, int* __restrict__ cell, double2* __restrict__ lam12,
    double* __restrict__ lam3, double2* __restrict__ pos, const double* __restrict__ xy, const int* __restrict__ tri,
    const double4* __restrict__ invJ, const int* __restrict__ nbr_off, const int* __restrict__ nbr_idx,
    const double* __restrict__ u0, const double* __restrict__ u1, int* __restrict__ dead)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pst[0]) return;
    int c = cell[p];
    if (c < 0) return;
    const double2 l12 = lam12[p];
    double l[3] = {l12.x, l12.y, lam3[p]};
    double2 x = pos[p];
    for (int s = 0; s < substeps; ++s) {
        const int t0 = tri[3 * c], t1 = tri[3 * c + 1], t2 = tri[3 * c + 2];
        double ax = 0.0, ay = 0.0;
        ax += l[0] * u0[t0]; ay += l[0] * u1[t0];
        ax += l[1] * u0[t1]; ay += l[1] * u1[t1];
        ax += l[2] * u0[t2]; ay += l[2] * u1[t2];
        x.x = x.x + dts * ax;
        x.y = x.y + dts * ay;
        double L[3];
        if (inside(invJ[c], xy[2 * t2], xy[2 * t2 + 1], x.x, x.y, L)) {
            l[0] = L[0]; l[1] = L[1]; l[2] = L[2];
            continue;
        }
        int found = -1;
        for (int k = nbr_off[c]; k < nbr_off[c + 1]; ++k) {
            const int nb = nbr_idx[k];
            const int v3 = tri[3 * nb + 2];
            if (inside(invJ[nb], xy[2 * v3], xy[2 * v3 + 1], x.x, x.y, L)) {
                found = nb;
                break;
            }
        }
        if (found < 0) {
            c = -1;
            atomicAdd(dead, 1);
            break;
        }
        c = found;
        l[0] = L[0]; l[1] = L[1]; l[2] = L[2];
    }
    cell[p] = c;
    lam12[p] = make_double2(l[0], l[1]);
    lam3[p] = l[2];
    pos[p] = x;
}

// kCountParticlesInSubcells (particle_handler_2d.cu:173-181): only emptiness matters, so one bit per
// subcell (atomicOr into a 16-bit mask per cell) replaces the reference's int counter per subcell
__global__ void k_census(const long* __restrict__ pst, const int* __restrict__ cell, const double2* __restrict__ lam12,
                         const double* __restrict__ lam3, int L, double step, uint32_t* __restrict__ mask)
{
    // Particles arrive grouped by cell (the array was sorted by cell at the end of the previous step and
    // advection moves few of them), so the lanes of a wave mostly share two or three cells. The subcell
    // bits are OR-ed across lanes with the same cell first (a suffix scan keyed on the cell), and only the
    // first lane of each run of equal cells issues the atomic: one atomic per run instead of per particle.
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long P = pst[0];
    if ((long)blockIdx.x * blockDim.x >= P) return;   // a whole workgroup past the count (the grid covers the capacity)
    const int lane = threadIdx.x & 63;
    int c = -1;
    uint32_t bits = 0u;
    if (p < P) {
        c = cell[p];
        if (c >= 0) {
            const double2 l12 = lam12[p];
            bits = 1u << subcell(L, step, l12.x, l12.y, lam3[p]);
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ob = __shfl_down(bits, off, 64);
        const int oc = __shfl_down(c, off, 64);
        if (lane + off < 64 && oc == c) bits |= ob;
    }
    const int prev = __shfl_up(c, 1, 64);
    if (c >= 0 && (lane == 0 || prev != c)) atomicOr(&mask[c], bits);
}

// kCountParticlesToBeAdded (:183-195)
__global__ void k_refill_count(int nt, int ns, const uint32_t* __restrict__ mask, int* __restrict__ add)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nt) return;
    const uint32_t full = ns >= 32 ? 0xffffffffu : ((1u << ns) - 1u);
    add[c] = ns - __popc(mask[c] & full);
}

// kAddParticlesToCell (:197-236): new particles at the centroids of the empty subcells, velocity
// interpolated from the grid; appended after the current count in (cell, subcell) order with consecutive
// IDs from the next free one
__global__ void k_refill_write(int nt, int ns, const long* __restrict__ pst, const uint32_t* __restrict__ mask,
                               const int* __restrict__ add, const int* __restrict__ add_off,
                               const double* __restrict__ xy, const int* __restrict__ tri, Centers cen,
                               const double* __restrict__ u0, const double* __restrict__ u1, int* __restrict__ cell,
                               double2* __restrict__ lam12, double* __restrict__ lam3, double2* __restrict__ pos,
                               double2* __restrict__ vel, uint32_t* __restrict__ id)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nt || add[c] == 0) return;
    const long base = pst[0];
    const uint32_t id0 = (uint32_t)pst[1];
    const uint32_t m = mask[c];
    const int a = tri[3 * c], b = tri[3 * c + 1], d = tri[3 * c + 2];
    int k = 0;
    for (int s = 0; s < ns; ++s) {
        if (m & (1u << s)) continue;
        const long p = base + add_off[c] + k;
        const double L0 = cen.c[s][0], L1 = cen.c[s][1], L2 = cen.c[s][2];
        cell[p] = c;
        lam12[p] = make_double2(L0, L1);
        lam3[p] = L2;
        pos[p] = make_double2(L0 * xy[2 * a] + L1 * xy[2 * b] + L2 * xy[2 * d],
                              L0 * xy[2 * a + 1] + L1 * xy[2 * b + 1] + L2 * xy[2 * d + 1]);
        double vx = 0.0, vy = 0.0;
        vx += L0 * u0[a]; vy += L0 * u1[a];
        vx += L1 * u0[b]; vy += L1 * u1[b];
        vx += L2 * u0[d]; vy += L2 * u1[d];
        vel[p] = make_double2(vx, vy);
        id[p] = id0 + (uint32_t)(add_off[c] + k);
        ++k;
    }
}

// ------------------------------------------------------------------ exclusive scan (int)
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__device__ int block_exclusive_scan(int v, int* total)
{
    __shared__ int wsum[kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
    for (int w = 0; w < kWaves; ++w) {
        if (w < wid) pre += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

__global__ void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out, long n, int* __restrict__ sums)
{
    __shared__ int buf[kScanTile];
    const long base = (long)blockIdx.x * kScanTile;
    for (int i = threadIdx.x; i < kScanTile; i += kBlock) buf[i] = base + i < n ? in[base + i] : 0;
    __syncthreads();
    int run = 0;
    for (int i = 0; i < kScanItems; ++i) run += buf[threadIdx.x * kScanItems + i];
    int tot;
    int pre = block_exclusive_scan(run, &tot);
    for (int i = 0; i < kScanItems; ++i) {
        const int v = buf[threadIdx.x * kScanItems + i];
        buf[threadIdx.x * kScanItems + i] = pre;
        pre += v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kScanTile; i += kBlock)
        if (base + i < n) out[base + i] = buf[i];
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = tot;
}

__global__ void k_scan_add(int* __restrict__ out, long n, const int* __restrict__ sums_scanned)
{
    const long base = (long)blockIdx.x * kScanTile;
    const int add = sums_scanned[blockIdx.x];
    for (int i = threadIdx.x; i < kScanTile; i += kBlock)
        if (base + i < n) out[base + i] += add;
}

// total = sum(in) written to total_dev (if not null); tmp holds at least ntiles + ntiles2 + 1 ints
static void scan_with_tmp(Context* c, const int* in, int* out, long n, int* tmp, int* total_dev);

__global__ void k_scan_total(const int* __restrict__ in, const int* __restrict__ out, long n, int* __restrict__ total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = n > 0 ? out[n - 1] + in[n - 1] : 0;
}

static void scan_with_tmp(Context* c, const int* in, int* out, long n, int* tmp, int* total_dev)
{
    if (n <= 0) return;
    const long nt = (n + kScanTile - 1) / kScanTile;
    if (nt == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, c->stream, in, out, n, (int*)nullptr);
    } else {
        int* sums = tmp;
        int* sums_scan = tmp + nt;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kBlock), 0, c->stream, in, out, n, sums);
        PFEM_REQUIRE(nt <= kScanTile * (long)kScanTile, "scan too long");
        scan_with_tmp(c, sums, sums_scan, nt, tmp + 2 * nt, nullptr);
        hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kBlock), 0, c->stream, out, n, sums_scan);
    }
    if (total_dev) hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, c->stream, in, out, n, total_dev);
}

static long scan_tmp_size(long n)
{
    long s = 0;
    while (n > kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        s += 2 * n;
    }
    return s + 1;
}

void exclusive_scan_int(Context* c, const int* in, int* out, long n, int* total_dev)
{
    DevArray<int> tmp(scan_tmp_size(n));
    scan_with_tmp(c, in, out, n, tmp.p, total_dev);
    PFEM_HIP(hipStreamSynchronize(c->stream));
}

// ------------------------------------------------------------------ stable LSD radix sort (8-bit digits)
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kBlock * kRadixItems;

// Particles to sort after a refill: the count before it plus the particles it appended
__device__ __forceinline__ long sort_count(const long* __restrict__ pst, const int* __restrict__ counters)
{
    return pst[0] + counters[1];
}

__global__ __launch_bounds__(kBlock) void k_radix_hist(const long* __restrict__ pst, const int* __restrict__ counters,
                                                       const uint32_t* __restrict__ keys, int shift,
                                                       int* __restrict__ hist, int nblocks)
{
    const long n = sort_count(pst, counters);
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * kRadixTile;
    for (int i = 0; i < kRadixItems; ++i) {
        const long idx = base + (long)i * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Scatter with stable ranks: the tile is processed in 256-key rounds in index order; inside a round
// each wave finds the lanes sharing its digit with 8 ballots, and waves are ordered through LDS counts.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const long* __restrict__ pst, const int* __restrict__ counters,
                                                          const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                          uint32_t* __restrict__ vout, int shift,
                                                          const int* __restrict__ offs, int nblocks)
{
    const long n = sort_count(pst, counters);
    if ((long)blockIdx.x * kRadixTile >= n) return;   // a tile past the count: no keys, nothing to place
    __shared__ int base[256];
    __shared__ int wcnt[kWaves][256];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    base[threadIdx.x] = offs[(size_t)threadIdx.x * nblocks + blockIdx.x];
    for (int w = 0; w < kWaves; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const long tbase = (long)blockIdx.x * kRadixTile;
    for (int i = 0; i < kRadixItems; ++i) {
        const long idx = tbase + (long)i * kBlock + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t key = valid ? kin[idx] : 0u;
        const uint32_t val = valid ? (vin ? vin[idx] : (uint32_t)idx) : 0u;
        const uint32_t d = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wid][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int off = base[d] + rank;
            for (int w = 0; w < wid; ++w) off += wcnt[w][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        {
            const int dd = threadIdx.x;
            int s = 0;
            for (int w = 0; w < kWaves; ++w) {
                s += wcnt[w][dd];
                wcnt[w][dd] = 0;
            }
            base[dd] += s;
        }
        __syncthreads();
    }
}

__global__ void k_make_keys(const long* __restrict__ pst, const int* __restrict__ counters, const int* __restrict__ cell,
                            int nt, uint32_t* __restrict__ keys)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < sort_count(pst, counters)) {
        const int c = cell[p];
        keys[p] = c < 0 ? (uint32_t)nt : (uint32_t)c;
    }
}

__global__ void k_gather(const long* __restrict__ pst, const int* __restrict__ counters,
                         const uint32_t* __restrict__ perm, const int* __restrict__ cell0,
                         const double2* __restrict__ l120, const double* __restrict__ l30,
                         const double2* __restrict__ pos0, const double2* __restrict__ vel0,
                         const uint32_t* __restrict__ id0, int* __restrict__ cell1, double2* __restrict__ l121,
                         double* __restrict__ l31, double2* __restrict__ pos1, double2* __restrict__ vel1,
                         uint32_t* __restrict__ id1)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sort_count(pst, counters)) return;
    const uint32_t s = perm[p];
    cell1[p] = cell0[s];
    l121[p] = l120[s];
    l31[p] = l30[s];
    pos1[p] = pos0[s];
    vel1[p] = vel0[s];
    id1[p] = id0[s];
}

// The step's new device state: count = count + added - deleted (the deleted particles sorted last), the
// next free ID, and the step's deleted and added counts
__global__ void k_pstate_finish(long* __restrict__ pst, const int* __restrict__ counters)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long dead = counters[0], added = counters[1];
    pst[0] = pst[0] + added - dead;
    pst[1] = pst[1] + added;
    pst[2] = dead;
    pst[3] = added;
}

__global__ void k_zero_u32(long n, uint32_t* __restrict__ d)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0u;
}

// first particle of every cell in the cell-sorted array (cell_start[nt] = number of particles)
__global__ void k_cell_start(const long* __restrict__ pst, const int* __restrict__ cell, int nt, int* __restrict__ start)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = pst[0];
    if (p > n) return;
    const int prev = p == 0 ? -1 : cell[p - 1];
    const int curc = p == n ? nt : cell[p];
    for (int c = prev + 1; c <= curc; ++c) start[c] = (int)p;
}

// ------------------------------------------------------------------ projection
// kProjectParticleVelocityOntoGrid / kFinalizeVelocityProjection (:90-115) as a segmented reduction
__global__ void k_proj_cells(int nt, const int* __restrict__ start, const double2* __restrict__ lam12,
                             const double* __restrict__ lam3, const double2* __restrict__ vel,
                             double* __restrict__ part)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nt) return;
    double s[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) s[k] = 0.0;
    // the cell's particles in array order; the loads of up to 8 particles are issued together
    constexpr int kU = 8;
    const int p1 = start[c + 1];
    for (int p0 = start[c]; p0 < p1; p0 += kU) {
        double2 l12[kU], v[kU];
        double l3[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (p0 + u < p1) {
                l12[u] = lam12[p0 + u];
                l3[u] = lam3[p0 + u];
                v[u] = vel[p0 + u];
            }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (p0 + u < p1) {
                const double l[3] = {l12[u].x, l12[u].y, l3[u]};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    s[3 * i + 0] += l[i] * v[u].x;
                    s[3 * i + 1] += l[i] * v[u].y;
                    s[3 * i + 2] += l[i];
                }
            }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) part[9 * (size_t)c + k] = s[k];
}

__global__ void k_proj_nodes(int nv, const int* __restrict__ inc_off, const int* __restrict__ inc_code,
                             const double* __restrict__ part, double* __restrict__ u0, double* __restrict__ u1)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double sx = 0.0, sy = 0.0, w = 0.0;
    for (int e = inc_off[v]; e < inc_off[v + 1]; ++e) {
        const int code = inc_code[e];
        const double* s = part + 9 * (size_t)(code >> 2) + 3 * (code & 3);
        sx += s[0];
        sy += s[1];
        w += s[2];
    }
    u0[v] = sx / w;
    u1[v] = sy / w;
}

// ------------------------------------------------------------------ host side
static void alloc_state(pfem_particles_t* ph, long cap)
{
    for (int b = 0; b < 2; ++b) {
        ph->cell[b].alloc(cap);
        ph->lam12[b].alloc(cap);
        ph->lam3[b].alloc(cap);
        ph->pos[b].alloc(cap);
        ph->vel[b].alloc(cap);
        ph->id[b].alloc(cap);
        ph->keys[b].alloc(cap);
        ph->vals[b].alloc(cap);
    }
    const long nblk = (cap + kRadixTile - 1) / kRadixTile;
    ph->hist.alloc(256 * (size_t)nblk);
    ph->scan_tmp.alloc(scan_tmp_size(std::max<long>(256 * nblk, ph->mesh->nt + 1)) + 8);
    ph->capacity = cap;
    ++ph->generation;
}

// The host copy of the device state {count, next ID, deleted, added}, once its pending copy is complete
// (normally long before: the stream has passed it by the next call that needs it).
static void sync_host_state(pfem_particles_t* ph)
{
    if (!ph->pst_pending) return;
    PFEM_HIP(hipEventSynchronize(ph->pst_ev));
    ph->count = ph->h_pst[0];
    ph->next_id = (uint32_t)ph->h_pst[1];
    ph->deleted_last = ph->h_pst[2];
    ph->added_last = ph->h_pst[3];
    ph->pst_pending = false;
}

// The device state from the host's (after a seed or an upload)
static void put_device_state(Context* c, pfem_particles_t* ph)
{
    ph->h_pst[0] = ph->count;
    ph->h_pst[1] = ph->next_id;
    ph->h_pst[2] = ph->deleted_last;
    ph->h_pst[3] = ph->added_last;
    h2d(c, ph->pst.p, ph->h_pst, 4 * sizeof(long));
    ph->pst_pending = false;
}

// capacity for `need` particles; the first ph->count particles of the current buffers are kept (the host
// state must be current)
static void grow(Context* c, pfem_particles_t* ph, long need)
{
    if (need <= ph->capacity) return;
    // deviceVector::resize grows by 1.1 (src/common/device_vector.cuh:96-109); keep more headroom
    const long cap = std::max<long>(need + need / 4, 1024);
    const int b = ph->cur;
    DevArray<int> cell(cap);
    DevArray<double2> l12(cap), pos(cap), vel(cap);
    DevArray<double> l3(cap);
    DevArray<uint32_t> id(cap);
    const size_t n = ph->count;
    PFEM_HIP(hipMemcpyAsync(cell.p, ph->cell[b].p, n * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    PFEM_HIP(hipMemcpyAsync(l12.p, ph->lam12[b].p, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
    PFEM_HIP(hipMemcpyAsync(l3.p, ph->lam3[b].p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    PFEM_HIP(hipMemcpyAsync(pos.p, ph->pos[b].p, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
    PFEM_HIP(hipMemcpyAsync(vel.p, ph->vel[b].p, n * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
    PFEM_HIP(hipMemcpyAsync(id.p, ph->id[b].p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
    alloc_state(ph, cap);
    ph->cell[b] = std::move(cell);
    ph->lam12[b] = std::move(l12);
    ph->lam3[b] = std::move(l3);
    ph->pos[b] = std::move(pos);
    ph->vel[b] = std::move(vel);
    ph->id[b] = std::move(id);
}

// cell_start of buffer b for the device count (grid over the capacity)
static void compute_cell_start(Context* c, pfem_particles_t* ph, int b)
{
    hipLaunchKernelGGL(k_cell_start, dim3(blocks_for(ph->capacity + 1)), dim3(kBlock), 0, c->stream,
                       (const long*)ph->pst.p, ph->cell[b].p, ph->mesh->nt, ph->cell_start.p);
}

// stable reorder of buffer b's particles (count + added, from the device) by cell into buffer 1 - b, the
// deleted ones (cell = -1) last
static void sort_by_cell(Context* c, pfem_particles_t* ph, int b)
{
    const int nt = ph->mesh->nt;
    const int o = 1 - b;
    const long cap = ph->capacity;
    const long* pst = ph->pst.p;
    const int* cnt = ph->counters.p;
    int bits = 1;
    while ((1u << bits) <= (uint32_t)nt) ++bits;   // keys are in [0, nt]
    const int passes = (bits + 7) / 8;
    hipLaunchKernelGGL(k_make_keys, dim3(blocks_for(cap)), dim3(kBlock), 0, c->stream, pst, cnt, ph->cell[b].p, nt,
                       ph->keys[0].p);
    const int nblk = (int)((cap + kRadixTile - 1) / kRadixTile);
    int kb = 0;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(kBlock), 0, c->stream, pst, cnt, ph->keys[kb].p, shift,
                           ph->hist.p, nblk);
        scan_with_tmp(c, ph->hist.p, ph->hist.p, 256L * nblk, ph->scan_tmp.p, nullptr);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(kBlock), 0, c->stream, pst, cnt, ph->keys[kb].p,
                           pass == 0 ? (const uint32_t*)nullptr : ph->vals[kb].p, ph->keys[1 - kb].p,
                           ph->vals[1 - kb].p, shift, ph->hist.p, nblk);
        kb = 1 - kb;
    }
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(cap)), dim3(kBlock), 0, c->stream, pst, cnt, ph->vals[kb].p,
                       ph->cell[b].p, ph->lam12[b].p, ph->lam3[b].p, ph->pos[b].p, ph->vel[b].p, ph->id[b].p,
                       ph->cell[o].p, ph->lam12[o].p, ph->lam3[o].p, ph->pos[o].p, ph->vel[o].p, ph->id[o].p);
    PFEM_HIP(hipGetLastError());
}

static Centers centers_of(const pfem_particles_t* ph)
{
    Centers cen{};
    std::memcpy(cen.c, ph->centers, sizeof(cen.c));
    return cen;
}

// Host part of an advection step, before any launch: the count of the previous step, and capacity for
// the largest refill this step can make (every subcell of every cell). Never inside a captured region.
void particles_advect_prepare(pfem_particles_t* ph)
{
    Context* c = C(ph->ctx);
    pfem_mesh2d_t* m = ph->mesh;
    PFEM_REQUIRE(m->nbr_off.p, "the mesh was created without neighbour lists");
    sync_host_state(ph);
    grow(c, ph, ph->count + (long)m->nt * ph->ns);
}

// The launches of an advection step on buffer b (the particles end in buffer 1 - b): advection and
// relocation, census, refill, reorder by cell, the new device state and cell_start. Launches only (no
// host waits, allocations or host-side state), so that it can be captured; the caller sets ph->cur.
void particles_advect_launch(pfem_particles_t* ph, int b, const double* u0, const double* u1, double dt, int substeps)
{
    Context* c = C(ph->ctx);
    pfem_mesh2d_t* m = ph->mesh;
    const long cap = ph->capacity;
    const long* pst = ph->pst.p;
    const double dts = dt / substeps;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, 4L, (uint32_t*)ph->counters.p);
    hipLaunchKernelGGL(k_advect_relocate, dim3(blocks_for(cap)), dim3(kBlock), 0, c->stream, pst, substeps, dts,
                       ph->cell[b].p, ph->lam12[b].p, ph->lam3[b].p, ph->pos[b].p, m->xy.p, m->tri.p,
                       (const double4*)m->invJ.p, m->nbr_off.p, m->nbr_idx.p, u0, u1, ph->counters.p);
    // census
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(m->nt)), dim3(kBlock), 0, c->stream, (long)m->nt, ph->mask.p);
    hipLaunchKernelGGL(k_census, dim3(blocks_for(cap)), dim3(kBlock), 0, c->stream, pst, ph->cell[b].p,
                       ph->lam12[b].p, ph->lam3[b].p, ph->level, 1.0 / ph->level, ph->mask.p);
    hipLaunchKernelGGL(k_refill_count, dim3(blocks_for(m->nt)), dim3(kBlock), 0, c->stream, m->nt, ph->ns,
                       ph->mask.p, ph->add.p);
    scan_with_tmp(c, ph->add.p, ph->add_off.p, m->nt, ph->scan_tmp.p, ph->counters.p + 1);
    hipLaunchKernelGGL(k_refill_write, dim3(blocks_for(m->nt)), dim3(kBlock), 0, c->stream, m->nt, ph->ns, pst,
                       ph->mask.p, ph->add.p, ph->add_off.p, m->xy.p, m->tri.p, centers_of(ph), u0, u1,
                       ph->cell[b].p, ph->lam12[b].p, ph->lam3[b].p, ph->pos[b].p, ph->vel[b].p, ph->id[b].p);
    sort_by_cell(c, ph, b);
    hipLaunchKernelGGL(k_pstate_finish, dim3(1), dim3(64), 0, c->stream, ph->pst.p, (const int*)ph->counters.p);
    compute_cell_start(c, ph, 1 - b);
    PFEM_HIP(hipGetLastError());
}

// After the launches of a step: the device state's copy to the host, read at its next use
void particles_advect_post(pfem_particles_t* ph)
{
    Context* c = C(ph->ctx);
    PFEM_HIP(hipMemcpyAsync(ph->h_pst, ph->pst.p, 4 * sizeof(long), hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipEventRecord(ph->pst_ev, c->stream));
    ph->pst_pending = true;
}

void particles_advect(pfem_particles_t* ph, const double* u0, const double* u1, double dt, int substeps)
{
    PFEM_REQUIRE(substeps > 0, "substeps must be positive");
    particles_advect_prepare(ph);
    const int b = ph->cur;
    particles_advect_launch(ph, b, u0, u1, dt, substeps);
    ph->cur = 1 - b;
    particles_advect_post(ph);
}

void particles_project(pfem_particles_t* ph, double* u0, double* u1)
{
    Context* c = C(ph->ctx);
    pfem_mesh2d_t* m = ph->mesh;
    const int b = ph->cur;
    hipLaunchKernelGGL(k_proj_cells, dim3(blocks_for(m->nt)), dim3(kBlock), 0, c->stream, m->nt, ph->cell_start.p,
                       ph->lam12[b].p, ph->lam3[b].p, ph->vel[b].p, ph->part.p);
    hipLaunchKernelGGL(k_proj_nodes, dim3(blocks_for(m->nv)), dim3(kBlock), 0, c->stream, m->nv, m->inc.off.p,
                       m->inc.code.p, ph->part.p, u0, u1);
    PFEM_HIP(hipGetLastError());
}

void particles_correct(pfem_particles_t* ph, const double* u0, const double* u1, const double* uo0, const double* uo1)
{
    Context* c = C(ph->ctx);
    const int b = ph->cur;
    hipLaunchKernelGGL(k_correct, dim3(blocks_for(ph->capacity)), dim3(kBlock), 0, c->stream, (const long*)ph->pst.p,
                       ph->cell[b].p, ph->lam12[b].p, ph->lam3[b].p, ph->vel[b].p, ph->mesh->tri.p, u0, u1, uo0,
                       uo1);
    PFEM_HIP(hipGetLastError());
}

// The particle set's buffer generation (changes when the buffers are reallocated) and current buffer
long particles_generation(const pfem_particles_t* ph) { return ph->generation; }
int particles_current(const pfem_particles_t* ph) { return ph->cur; }
void particles_set_current(pfem_particles_t* ph, int b) { ph->cur = b; }

}  // namespace pfem

using namespace pfem;

extern "C" {

// ParticleHandler2D constructor (particle_handler_2d.cu:238-294): level clamped to [1, 4], L^2 subcell
// centroids, row i from the top holding 2i+1 sub-triangles
int pfem_particles_create(pfem_mesh2d m, int level, pfem_particles* out)
{
    PFEM_TRY(m ? m->ctx : nullptr, {
        PFEM_REQUIRE(m && out, "null argument");
        auto ph = std::make_unique<pfem_particles_t>();
        ph->ctx = m->ctx;
        ph->mesh = m;
        const int L = std::max(std::min(level, 4), 1);
        ph->level = L;
        ph->ns = L * L;
        const double dx = 1.0 / L;
        int k = -1;
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < 2 * i + 1; ++j) {
                const double xmin = (j / 2) * dx, xmax = xmin + dx;
                const double ymin = (L - 1 - i) * dx, ymax = ymin + dx;
                const double v0[3] = {xmin, ymax, 1.0 - xmin - ymax};
                double v1[3];
                v1[0] = (j % 2 == 0) ? xmin : xmax;
                v1[1] = (j % 2 == 0) ? ymin : ymax;
                v1[2] = 1.0 - v1[0] - v1[1];
                const double v2[3] = {xmax, ymin, 1.0 - xmax - ymin};
                ++k;
                for (int d = 0; d < 3; ++d) ph->centers[k][d] = (v0[d] + v1[d] + v2[d]) * 0.3333333333333333;
            }
        Context* c = C(m->ctx);
        ph->mask.alloc(m->nt);
        ph->add.alloc(m->nt + 1);
        ph->add_off.alloc(m->nt + 1);
        ph->counters.alloc(4);
        ph->cell_start.alloc(m->nt + 1);
        ph->part.alloc(9 * (size_t)m->nt);
        ph->pst.alloc(4);
        PFEM_HIP(hipHostMalloc((void**)&ph->h_pst, 4 * sizeof(long), hipHostMallocDefault));
        PFEM_HIP(hipEventCreateWithFlags(&ph->pst_ev, hipEventDisableTiming));
        zero_async(c, ph->counters.p, 4 * sizeof(int));
        alloc_state(ph.get(), 1024);
        put_device_state(c, ph.get());
        zero_async(c, ph->cell_start.p, ph->cell_start.bytes());   // no particles: every cell's range is empty
        PFEM_HIP(hipStreamSynchronize(c->stream));
        *out = ph.release();
    });
}

int pfem_particles_destroy(pfem_particles ph)
{
    if (!ph) return PFEM_OK;
    (void)hipStreamSynchronize(ph->ctx->stream);
    delete ph;
    return PFEM_OK;
}

int pfem_particles_seed(pfem_particles ph, int mode, int ppc, uint64_t seed)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph, "null argument");
        PFEM_REQUIRE(mode == 0 || mode == 1, "seed mode must be 0 (subcell centroids) or 1 (random)");
        if (mode == 0) ppc = ph->ns;
        PFEM_REQUIRE(ppc > 0, "particles per cell must be positive");
        Context* c = C(ph->ctx);
        const long P = (long)ph->mesh->nt * ppc;
        sync_host_state(ph);
        ph->count = 0;
        grow(c, ph, P + P / 5);
        const int b = ph->cur;
        const uint2 key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
        hipLaunchKernelGGL(k_seed, dim3(blocks_for(P)), dim3(kBlock), 0, c->stream, ph->mesh->nt, ppc, mode,
                           ph->mesh->xy.p, ph->mesh->tri.p, centers_of(ph), key, ph->cell[b].p, ph->lam12[b].p,
                           ph->lam3[b].p, ph->pos[b].p, ph->vel[b].p, ph->id[b].p);
        PFEM_HIP(hipGetLastError());
        ph->count = P;
        ph->next_id = (uint32_t)P;
        ph->deleted_last = ph->added_last = 0;
        put_device_state(c, ph);
        compute_cell_start(c, ph, b);
        PFEM_HIP(hipStreamSynchronize(c->stream));
    });
}

int pfem_particles_init_velocity(pfem_particles ph, const double* d_u0, const double* d_u1)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph && d_u0 && d_u1, "null argument");
        particles_correct(ph, d_u0, d_u1, nullptr, nullptr);
    });
}

int pfem_particles_advect(pfem_particles ph, const double* d_u0, const double* d_u1, double dt, int substeps)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph && d_u0 && d_u1, "null argument");
        particles_advect(ph, d_u0, d_u1, dt, substeps);
    });
}

int pfem_particles_project(pfem_particles ph, double* d_u0, double* d_u1)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph && d_u0 && d_u1, "null argument");
        particles_project(ph, d_u0, d_u1);
    });
}

int pfem_particles_correct(pfem_particles ph, const double* d_u0, const double* d_u1, const double* d_uo0,
                           const double* d_uo1)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph && d_u0 && d_u1 && d_uo0 && d_uo1, "null argument");
        particles_correct(ph, d_u0, d_u1, d_uo0, d_uo1);
    });
}

int pfem_particles_count(pfem_particles ph, long* count)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph && count, "null argument");
        sync_host_state(ph);
        *count = ph->count;
    });
}

int pfem_particles_download(pfem_particles ph, int* cell, double* lam, double* pos, double* vel, uint32_t* id)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph, "null argument");
        Context* c = C(ph->ctx);
        sync_host_state(ph);
        const size_t n = ph->count;
        const int b = ph->cur;
        if (cell) d2h_sync(c, cell, ph->cell[b].p, n * sizeof(int));
        if (lam) {
            std::vector<double2> l12(n);
            std::vector<double> l3(n);
            d2h_sync(c, l12.data(), ph->lam12[b].p, n * sizeof(double2));
            d2h_sync(c, l3.data(), ph->lam3[b].p, n * sizeof(double));
            for (size_t p = 0; p < n; ++p) {
                lam[3 * p] = l12[p].x;
                lam[3 * p + 1] = l12[p].y;
                lam[3 * p + 2] = l3[p];
            }
        }
        if (pos) d2h_sync(c, pos, ph->pos[b].p, n * sizeof(double2));
        if (vel) d2h_sync(c, vel, ph->vel[b].p, n * sizeof(double2));
        if (id) d2h_sync(c, id, ph->id[b].p, n * sizeof(uint32_t));
    });
}

int pfem_particles_upload(pfem_particles ph, long count, const int* cell, const double* lam, const double* pos,
                          const double* vel, const uint32_t* id, uint32_t next_id)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph && count >= 0 && cell && lam && pos && vel && id, "null argument");
        for (long p = 0; p < count; ++p) {
            PFEM_REQUIRE(cell[p] >= 0 && cell[p] < ph->mesh->nt, "particle cell out of range");
            PFEM_REQUIRE(p == 0 || cell[p] >= cell[p - 1], "uploaded particles must be ordered by cell");
        }
        Context* c = C(ph->ctx);
        sync_host_state(ph);
        ph->count = 0;
        grow(c, ph, count + count / 5 + 16);
        const int b = ph->cur;
        std::vector<double2> l12(count);
        std::vector<double> l3(count);
        for (long p = 0; p < count; ++p) {
            l12[p] = make_double2(lam[3 * p], lam[3 * p + 1]);
            l3[p] = lam[3 * p + 2];
        }
        h2d(c, ph->cell[b].p, cell, count * sizeof(int));
        h2d(c, ph->lam12[b].p, l12.data(), count * sizeof(double2));
        h2d(c, ph->lam3[b].p, l3.data(), count * sizeof(double));
        h2d(c, ph->pos[b].p, pos, count * sizeof(double2));
        h2d(c, ph->vel[b].p, vel, count * sizeof(double2));
        h2d(c, ph->id[b].p, id, count * sizeof(uint32_t));
        ph->count = count;
        ph->next_id = next_id;
        ph->deleted_last = ph->added_last = 0;
        put_device_state(c, ph);
        compute_cell_start(c, ph, b);
        PFEM_HIP(hipStreamSynchronize(c->stream));
    });
}

int pfem_particles_stats(pfem_particles ph, long* deleted_last, long* added_last)
{
    PFEM_TRY(ph ? ph->ctx : nullptr, {
        PFEM_REQUIRE(ph, "null argument");
        sync_host_state(ph);
        if (deleted_last) *deleted_last = ph->deleted_last;
        if (added_last) *added_last = ph->added_last;
    });
}

}  // extern "C"
