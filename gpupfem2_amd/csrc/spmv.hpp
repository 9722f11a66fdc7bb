// CSR SpMV tile routine shared by the SpMV-based kernels.
//
// One 256-lane workgroup per row tile (<= 256 rows, <= kTileNnz non-zeros). The tile's val/col stream is
// read with unit stride into LDS as products val*x[col]; each lane then sums its row's products left to
// right. Loads are fully coalesced, the only irregular access is the x gather, and the per-row order is
// fixed (bit-identical to a sequential CPU loop without FMA contraction).
#pragma once

#include "device.hpp"

namespace pfem {

constexpr int kTileNnz = 2048;

// Product of the calling lane's row with x (tile t: rows [tiles[2t], tiles[2t+1])). Returns the row index
// in r (r >= r1 means the lane owns no row). prod is an LDS buffer of kTileNnz entries.
template <typename VT, typename XT>
__device__ __forceinline__ XT tile_row_product_at(int t, const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col, const VT* __restrict__ val,
                                                  const XT* __restrict__ x, XT* prod, int& r, int& r1)
{
    const int r0 = tiles[2 * t];
    r1 = tiles[2 * t + 1];
    const int k0 = rowptr[r0];
    const int nk = rowptr[r1] - k0;
    r = r0 + (int)threadIdx.x;
    XT sum = 0;
    if (nk <= kTileNnz) {
        for (int k = threadIdx.x; k < nk; k += kBlock) prod[k] = (XT)val[k0 + k] * x[col[k0 + k]];
        __syncthreads();
        if (r < r1) {
            const int a = rowptr[r] - k0, b = rowptr[r + 1] - k0;
            for (int k = a; k < b; ++k) sum += prod[k];
        }
    } else {
        // a single row longer than the LDS tile: streamed in chunks, summed in order by lane 0
        for (int base = 0; base < nk; base += kTileNnz) {
            const int m = min(kTileNnz, nk - base);
            for (int k = threadIdx.x; k < m; k += kBlock) prod[k] = (XT)val[k0 + base + k] * x[col[k0 + base + k]];
            __syncthreads();
            if (threadIdx.x == 0)
                for (int k = 0; k < m; ++k) sum += prod[k];
            __syncthreads();
        }
    }
    return sum;
}

// The same for the tile of this workgroup's own index (tile t = blockIdx.x).
template <typename VT, typename XT>
__device__ __forceinline__ XT tile_row_product(const int* __restrict__ tiles, const int* __restrict__ rowptr,
                                               const int* __restrict__ col, const VT* __restrict__ val,
                                               const XT* __restrict__ x, XT* prod, int& r, int& r1)
{
    return tile_row_product_at<VT, XT>((int)blockIdx.x, tiles, rowptr, col, val, x, prod, r, r1);
}

// Tile of this workgroup when a grid of gridDim.x workgroups covers tiles [0, gridDim.x): workgroup b runs
// on XCD b mod 8, and XCD x takes the x-th contiguous band of tiles, so neighbouring row tiles (whose
// gathers overlap) meet in one L2. A bijection for every grid size.
__device__ __forceinline__ int xcd_band_tile()
{
    const int b = (int)blockIdx.x, g = (int)gridDim.x;
    const int x = b & 7, q = g >> 3, rem = g & 7;
    return x * q + min(x, rem) + (b >> 3);
}

}  // namespace pfem
