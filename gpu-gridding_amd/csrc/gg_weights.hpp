// gg_weights.hpp -- kernels of the uniform-weighting pre-pass
// (oskar_grid_weights.c): per-cell sums of the input weights ("write") and
// the input weights divided by their cell's sum ("read").
//
// Write is an ordered floating-point sum per cell, so it runs as
//   k_weight_cells  key = cell index (or `outside`), value = point index
//   radix sort      stable, so each cell's points stay in input order
//   k_weight_sum    one thread per cell adds the cell's weights, in input
//                   order, to the cell's existing value
// which performs, per cell, exactly the additions of the serial loop in the
// same order: the grid is bit-identical to the serial one.  Read is one
// independent division per point.
#pragma once

#include "gg_device.hpp"

namespace gg {

// Grid cell of one point: oskar_grid_weights.c:117-128 (float) and :48-59
// (double), with the x86-64 conversion of out-of-range values and the
// wrapping add of the compiled reference.  True when the cell lies inside
// the grid.
template <typename T>
__device__ __forceinline__ bool weights_cell(T u, T v, T grid_scale, int grid_size,
        int& gu, int& gv)
{
#pragma clang fp contract(off)
    const int centre = grid_size / 2;
    gu = wrap_add(x86_int(gg_round(-u * grid_scale)), centre);
    gv = wrap_add(x86_int(gg_round(v * grid_scale)), centre);
    return gu < grid_size && gu >= 0 && gv < grid_size && gv >= 0;
}

// Sets the sort count, zeroes the skipped counter and resets the bounding
// box {min u, min v, max u, max v} (box may be null).
__global__ void k_weight_init(unsigned long long* __restrict__ d_n, unsigned long long n,
        unsigned long long* __restrict__ skipped, int* __restrict__ box)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (d_n) *d_n = n;
    *skipped = 0;
    if (box) { box[0] = INT_MAX; box[1] = INT_MAX; box[2] = INT_MIN; box[3] = INT_MIN; }
}

__device__ __forceinline__ int wave_min_i(int x)
{
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ int wave_max_i(int x)
{
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
    return x;
}

// Adds a block's count of points off the grid to *skipped and merges its
// bounding box into box (when given).  Every thread of the block calls it.
__device__ __forceinline__ void weight_block_flush(unsigned cnt, int u0, int v0, int u1, int v1,
        unsigned long long* __restrict__ skipped, int* __restrict__ box)
{
    __shared__ unsigned s_cnt[4];
    __shared__ int s_box[4][4];
    const int wave = threadIdx.x >> 6;
    cnt = wave_sum(cnt);
    u0 = wave_min_i(u0); v0 = wave_min_i(v0);
    u1 = wave_max_i(u1); v1 = wave_max_i(v1);
    if (lane_id() == 0)
    {
        s_cnt[wave] = cnt;
        s_box[wave][0] = u0; s_box[wave][1] = v0; s_box[wave][2] = u1; s_box[wave][3] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const int nw = (int)(blockDim.x >> 6);
        unsigned long long c = 0;
        for (int w = 0; w < nw; ++w)
        {
            c += s_cnt[w];
            u0 = min(u0, s_box[w][0]); v0 = min(v0, s_box[w][1]);
            u1 = max(u1, s_box[w][2]); v1 = max(v1, s_box[w][3]);
        }
        if (c) atomicAdd(skipped, c);
        if (box && u0 <= u1)
        {
            atomicMin(&box[0], u0); atomicMin(&box[1], v0);
            atomicMax(&box[2], u1); atomicMax(&box[3], v1);
        }
    }
}

// Per point: when keys is given, key = v * G + u for points inside the grid
// and `outside` (> every cell index) for the others, value = point index.
// Counts the points off the grid and, when box is given, the bounding box of
// the cells in use.  Block size 256.
template <typename T>
__global__ __launch_bounds__(256)
void k_weight_cells(const T* __restrict__ uu, const T* __restrict__ vv, unsigned long long n,
        T grid_scale, int grid_size, unsigned outside, unsigned* __restrict__ keys,
        unsigned* __restrict__ vals, unsigned long long* __restrict__ skipped,
        int* __restrict__ box)
{
    unsigned cnt = 0;
    int u0 = INT_MAX, v0 = INT_MAX, u1 = INT_MIN, v1 = INT_MIN;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            i < n; i += (unsigned long long)gridDim.x * blockDim.x)
    {
        int gu, gv;
        const bool in = weights_cell(uu[i], vv[i], grid_scale, grid_size, gu, gv);
        if (keys)
        {
            keys[i] = in ? (unsigned)gv * (unsigned)grid_size + (unsigned)gu : outside;
            vals[i] = (unsigned)i;
        }
        if (in)
        {
            u0 = min(u0, gu); v0 = min(v0, gv);
            u1 = max(u1, gu); v1 = max(v1, gv);
        }
        else
            ++cnt;
    }
    weight_block_flush(cnt, u0, v0, u1, v1, skipped, box);
}

// Over the stably sorted (key, point) pairs: the first position of each cell
// adds that cell's weights, in point order, to the cell's value.  The cell
// (u, v) lives at grid[(v - v0) * pitch + (u - u0)] (a window of the grid
// starting at cell (u0, v0), or the whole grid with u0 = v0 = 0, pitch = G).
template <typename T>
__global__ __launch_bounds__(256)
void k_weight_sum(const unsigned* __restrict__ keys, const unsigned* __restrict__ vals,
        const T* __restrict__ weight, unsigned long long n, unsigned outside, int grid_size,
        T* __restrict__ grid, int u0, int v0, unsigned long long pitch)
{
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            p < n; p += (unsigned long long)gridDim.x * blockDim.x)
    {
        const unsigned k = keys[p];
        if (k == outside || (p > 0 && keys[p - 1] == k)) continue;
        unsigned long long e = p + 1;
        while (e < n && keys[e] == k) ++e;
        const unsigned cv = k / (unsigned)grid_size;
        const unsigned cu = k - cv * (unsigned)grid_size;
        T* cell = grid + (unsigned long long)(cv - (unsigned)v0) * pitch + (cu - (unsigned)u0);
        T acc = *cell;
        #pragma unroll 4
        for (unsigned long long q = p; q < e; ++q) acc += weight[vals[q]];
        *cell = acc;
    }
}

// Per point inside the grid: weight_out = weight_in / (cell sum), or 0 where
// the cell sum is 0 (oskar_grid_weights.c:162-163 / :93-94); points off the
// grid leave weight_out untouched.  The cell addressing is k_weight_sum's.
// Counts the points off the grid when skipped is given.
template <typename T>
__global__ __launch_bounds__(256)
void k_weight_read(const T* __restrict__ uu, const T* __restrict__ vv, const T* weight_in,
        T* weight_out, unsigned long long n, T grid_scale, int grid_size,
        const T* __restrict__ grid, int u0, int v0, unsigned long long pitch,
        unsigned long long* __restrict__ skipped)
{
    unsigned cnt = 0;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            i < n; i += (unsigned long long)gridDim.x * blockDim.x)
    {
        int gu, gv;
        if (!weights_cell(uu[i], vv[i], grid_scale, grid_size, gu, gv)) { ++cnt; continue; }
        const T g = grid[(unsigned long long)(gv - v0) * pitch + (unsigned)(gu - u0)];
        weight_out[i] = g != (T)0 ? weight_in[i] / g : (T)0;
    }
    if (skipped) weight_block_flush(cnt, 0, 0, -1, -1, skipped, nullptr);
}

} // namespace gg
