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

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
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

// Deterministic grid reductions carried across a kernel boundary: the producing kernel stores one
// partial per block, the consuming kernel re-sums the partials in a fixed order (lane-strided, then
// block_sum), so the total depends only on the producer's grid size. There is no inter-block
// synchronisation inside a kernel: on gfx950 an agent-scope fence writes back and invalidates the L2 of
// the issuing XCD, and a last-block-done reduction (one fence and one counter update per block) costs an
// SpMV several times its own run time by evicting the gathered vector.
//
// part[i * nb + b] holds value i of block b (nb = producer's grid size).
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ part, int slot)
{
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) part[(size_t)i * gridDim.x + slot] = v[i];
    }
}

template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ part)
{
    block_partials<NV>(v, part, (int)blockIdx.x);
}

// Totals of the nb partials of each of the NV values; valid in every thread. Lane t adds partials
// t, t + 256, t + 512, ... in that order; the loads of 8 consecutive terms are issued together.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int nb, double (&tot)[NV])
{
    __shared__ double res[NV];
    constexpr int kU = 8;
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = 0.0;
    for (int k0 = threadIdx.x; k0 < nb; k0 += kU * kBlock) {
        double t[NV][kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = k0 + u * kBlock;
#pragma unroll
            for (int i = 0; i < NV; ++i) t[i][u] = k < nb ? part[(size_t)i * nb + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
#pragma unroll
            for (int i = 0; i < NV; ++i) v[i] += t[i][u];
        }
    }
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) res[i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] = res[i];
}

inline unsigned blocks_for(long n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace pfem
