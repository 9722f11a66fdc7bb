// Device-side helpers shared by the kernels (wave64 reductions, deterministic grid reductions).
#pragma once

#include <hip/hip_runtime.h>

namespace pfem {

constexpr int kBlock = 256;          // 4 wavefronts of 64 lanes
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_sumf(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum of NV values over the block; the result is valid in thread 0. The order is fixed
// (shuffle tree within each wave, then waves 0..3 in sequence), so results are reproducible.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double sh[NV][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        v[i] = wave_sum(v[i]);
        if (lane == 0) sh[i][wid] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = sh[i][0];
            for (int w = 1; w < kWaves; ++w) s += sh[i][w];
            v[i] = s;
        }
    }
    __syncthreads();
}

// Deterministic grid-wide reduction of NV doubles: every block deposits its partial sums, the last
// block to arrive adds all partials in block order and returns true (in all of its threads) with
// the totals in tot[] (valid in thread 0). partials has NV * gridDim.x entries; counter must be zero
// before the launch and is reset to zero by the last block.
template <int NV>
__device__ bool grid_sum(double (&v)[NV], double* partials, unsigned* counter, double (&tot)[NV])
{
    __shared__ bool last;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; ++i) partials[(size_t)i * gridDim.x + blockIdx.x] = v[i];
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        s[i] = 0.0;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
            s[i] += __hip_atomic_load(&partials[(size_t)i * gridDim.x + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(s);
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; ++i) tot[i] = s[i];
        *counter = 0u;
    }
    return true;
}

// Same for a runtime number of values (m <= 64), used by the GMRES multi-dot.
__device__ inline bool grid_sum_dyn(double* v /* m per thread, in registers or LDS */, int m, double* partials,
                                    unsigned* counter, double* tot /* m */, double* scratch /* LDS m*kWaves */)
{
    __shared__ bool last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = 0; i < m; ++i) {
        const double w = wave_sum(v[i]);
        if (lane == 0) scratch[i * kWaves + wid] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < m; ++i) {
            double s = scratch[i * kWaves];
            for (int w = 1; w < kWaves; ++w) s += scratch[i * kWaves + w];
            partials[(size_t)i * gridDim.x + blockIdx.x] = s;
        }
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
            s += __hip_atomic_load(&partials[(size_t)i * gridDim.x + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s = wave_sum(s);
        if (lane == 0) scratch[i * kWaves + wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < m; ++i) {
            double s = scratch[i * kWaves];
            for (int w = 1; w < kWaves; ++w) s += scratch[i * kWaves + w];
            tot[i] = s;
        }
        *counter = 0u;
    }
    __syncthreads();
    return true;
}

inline unsigned blocks_for(long n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace pfem
