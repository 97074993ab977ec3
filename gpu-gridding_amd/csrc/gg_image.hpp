// gg_image.hpp -- grid -> image: the 2-D FFT of the grid that follows
// gridding (README:24-26).  For an even grid size G, c = G/2,
//   image[y][x] = Re( sum_{v,u} grid[v][u]
//                     * exp(+2 pi i ((u - c)(x - c) + (v - c)(y - c)) / G) ) * scale
// which equals (-1)^(x+y) * BACKWARD_DFT[ (-1)^(u+v) grid ][y][x] (the
// constant phase exp(2 pi i c^2 / G) squares to 1 over the two axes).  So:
//   k_image_pre    work = (-1)^(u+v) grid          (one pass, HBM-bound)
//   hipFFT         unnormalised backward C2C / Z2Z, in place on work
//   k_image_post   image = (-1)^(x+y) Re(work) * scale
// Each thread handles two horizontally adjacent cells (G is even, so a pair
// never straddles a row and the two signs are opposite).
#pragma once

#include "gg_device.hpp"

namespace gg {

template <typename T>
__global__ __launch_bounds__(256)
void k_image_pre(const typename V2<T>::t* __restrict__ grid, typename V2<T>::t* __restrict__ work,
        unsigned G, unsigned long long npairs)
{
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            p < npairs; p += (unsigned long long)gridDim.x * blockDim.x)
    {
        const unsigned long long i = 2 * p;
        const unsigned v = (unsigned)(i / G);          // row; column i - v*G is even
        const T s = (v & 1u) ? (T)-1 : (T)1;
        const typename V2<T>::t a = grid[i], b = grid[i + 1];
        work[i] = {s * a.x, s * a.y};
        work[i + 1] = {-s * b.x, -s * b.y};
    }
}

template <typename T>
__global__ __launch_bounds__(256)
void k_image_post(const typename V2<T>::t* __restrict__ work, T* __restrict__ image, unsigned G,
        unsigned long long npairs, T scale)
{
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            p < npairs; p += (unsigned long long)gridDim.x * blockDim.x)
    {
        const unsigned long long i = 2 * p;
        const unsigned y = (unsigned)(i / G);
        const T s = (y & 1u) ? -scale : scale;
        image[i] = s * work[i].x;
        image[i + 1] = -s * work[i + 1].x;
    }
}

} // namespace gg
