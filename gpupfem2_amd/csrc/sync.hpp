// Publication of values between workgroups of one launch without fences, shared by the single-pass
// triangular sweeps and the fused Krylov iterations.
//
// A value is written once per launch into a location that reads as a sentinel NaN until then, with an
// agent-scope atomic store; the value is its own flag. A reader takes an ordinary load and, only while it
// still sees the sentinel, polls with agent-scope atomic loads. A workgroup only ever waits on
// workgroups dispatched before it (smaller indices), which are resident or finished, so the grid drains
// in dispatch order. Every wait has a spin limit; exceeding it sets an error flag that the host reports
// instead of waiting for ever.
#pragma once

#include "device.hpp"

namespace pfem {

template <typename T>
struct Sentinel;
template <>
struct Sentinel<double> {
    using U = unsigned long long;
    static constexpr U bits = 0x7FFC0DEC0DEC0DECull;   // quiet NaN with a payload no arithmetic produces
    static constexpr U nan = 0x7FF8000000000000ull;
};
template <>
struct Sentinel<float> {
    using U = unsigned int;
    static constexpr U bits = 0x7FE0DEC5u;
    static constexpr U nan = 0x7FC00000u;
};


template <typename T>
__device__ __forceinline__ T await_value(T* p, int* err, int limit)
{
    using U = typename Sentinel<T>::U;
    constexpr U sent = Sentinel<T>::bits;
    U* q = reinterpret_cast<U*>(p);
    // an ordinary load sees either the sentinel or the published value: the location is written once
    // per launch, and the previous launch's contents are not visible past the kernel boundary
    U v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (v == sent) {
        int spins = 0;
        do {
            __builtin_amdgcn_s_sleep(1);
            v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } while (v == sent && ++spins < limit);
        if (v == sent) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __builtin_bit_cast(T, v);
}

template <typename T>
__device__ __forceinline__ void publish(T* p, T v)
{
    using U = typename Sentinel<T>::U;
    constexpr U sent = Sentinel<T>::bits, canonical = Sentinel<T>::nan;
    U b = __builtin_bit_cast(U, v);
    if (b == sent) b = canonical;
    __hip_atomic_store(reinterpret_cast<U*>(p), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ T sentinel_value()
{
    constexpr typename Sentinel<T>::U sent = Sentinel<T>::bits;
    return __builtin_bit_cast(T, sent);
}

template <typename T>
__global__ void k_fill_sentinel(long n, T* __restrict__ d)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = sentinel_value<T>();
}

// Tagged publication: two launches that alternate on one FP64 buffer tell their values apart by the
// lowest mantissa bit, the tag (0 in one launch, 1 in the other), so no launch has to re-arm a sentinel
// before the next one. A published value has its lowest bit forced to the launch's tag (a change of at
// most one unit in the last place, 2^-52 relative); a location reads with the other tag until it is
// published in this launch, which plays the part of the sentinel. Quiet NaNs stay NaNs (their top
// mantissa bit is set).
template <int TAG>
__device__ __forceinline__ double tagged(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    return __builtin_bit_cast(double, TAG ? (b | 1ull) : (b & ~1ull));
}

template <int TAG>
__device__ __forceinline__ void publish_tagged(double* p, double v)
{
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __builtin_bit_cast(unsigned long long, tagged<TAG>(v)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Polls a location that did not yet carry tag TAG (the first ordinary load saw the other tag).
template <int TAG>
__device__ __forceinline__ double await_tagged(double* p, int* err, int limit)
{
    using U = unsigned long long;
    U* q = reinterpret_cast<U*>(p);
    U v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if ((v & 1ull) != (U)TAG) {
        int spins = 0;
        do {
            __builtin_amdgcn_s_sleep(1);
            v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } while ((v & 1ull) != (U)TAG && ++spins < limit);
        if ((v & 1ull) != (U)TAG) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __builtin_bit_cast(double, v);
}

// Every element of d reads +0.0, which carries tag 0 (the state a launch publishing with tag 1 expects).
template <typename T>
__global__ void k_fill_tag0(long n, T* __restrict__ d)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (T)0;
}


// Grid total finished inside the launch by its last workgroup. Every workgroup publishes its partial
// (one double, after block_sum) into part[blockIdx.x], which reads as the sentinel until then. The last
// workgroup is dispatched after all the others, so it can wait for every slot: it adds them in the order
// sum_partials uses (lane t: slots t, t + 256, ...; loads of 8 terms issued together; then block_sum), so
// the total is bit-identical to a consumer kernel's sum_partials, and it re-arms each slot for the next
// launch. A slot still unpublished after the lane's spin budget sets *err (the total is then a NaN).
// Returns the total in every thread of the last workgroup; all its threads must call it.
__device__ __forceinline__ double collect_partials(double* part, int nb, int* err, int limit)
{
    using U = Sentinel<double>::U;
    constexpr U sent = Sentinel<double>::bits;
    constexpr int kU = 8;
    __shared__ double res;
    double v[1] = {0.0};
    int budget = limit;   // polls left to this lane over all its slots, so a failure costs one spin limit
    for (int k0 = threadIdx.x; k0 < nb; k0 += kU * kBlock) {
        U t[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = k0 + u * kBlock;
            t[u] = k < nb ? __hip_atomic_load(reinterpret_cast<U*>(part + k), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT)
                          : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = k0 + u * kBlock;
            if (k < nb) {
                while (t[u] == sent && budget > 0) {
                    __builtin_amdgcn_s_sleep(1);
                    t[u] = __hip_atomic_load(reinterpret_cast<U*>(part + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    --budget;
                }
                if (t[u] == sent) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(reinterpret_cast<U*>(part + k), sent, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            v[0] += __builtin_bit_cast(double, t[u]);
        }
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) res = v[0];
    __syncthreads();
    return res;
}

// The calling workgroup's partial (valid in thread 0 after block_sum) into its slot.
__device__ __forceinline__ void publish_partial(double v, double* part)
{
    if (threadIdx.x == 0) publish(part + blockIdx.x, v);
}

}  // namespace pfem
