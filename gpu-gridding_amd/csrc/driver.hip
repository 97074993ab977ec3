// driver.hip -- the reference harness's flow (main.cpp:55-453) over its input
// files, on the GPU: uniform weighting over every block, then weights read
// and gridding block by block, with every array the passes share (kernel
// stack, weights grid, visibility grid, norm) resident on the device.
// Built on the library's public C ABI (mi355_gridder.h, mi355_io.h).
#include "gg_status.hpp"
#include "mi355_gridder.h"
#include "mi355_io.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using ggs::fail;
using ggs::ok;

namespace {

#define DRV_HIP(expr)                                                        \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            return fail(MI355_ERR_HIP, "%s failed: %s", #expr,               \
                    hipGetErrorString(e_));                                  \
    } while (0)
#define DRV_TRY(expr)                                                        \
    do { int s_ = (expr); if (s_ != MI355_OK) return s_; } while (0)

// Device allocations of one run, released on every return path.
struct DevArrays
{
    std::vector<void*> ptrs;
    ~DevArrays() { for (void* p : ptrs) (void)hipFree(p); }
    int alloc(void** p, size_t bytes)
    {
        *p = nullptr;
        hipError_t e = hipMalloc(p, bytes ? bytes : 16);
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            return fail(MI355_ERR_OUT_OF_MEMORY, "device allocation of %zu bytes failed", bytes);
        }
        ptrs.push_back(*p);
        return MI355_OK;
    }
};

struct Files
{
    mi355_bin* vis = nullptr;
    void* kre = nullptr; void* kim = nullptr; int* support = nullptr;
    ~Files()
    {
        mi355_bin_close(vis);
        mi355_io_free(kre); mi355_io_free(kim); mi355_io_free(support);
    }
};

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Norm-wise error of a grid against a reference grid (both G*G complex):
// per block, max |a - b|, max |b|, sum |a - b|^2, sum |b|^2.
template <typename TA, typename TB>
__global__ void k_grid_error(const TA* __restrict__ a, const TB* __restrict__ b,
        unsigned long long n, double* __restrict__ part)
{
    __shared__ double sh[4][256];
    double md = 0, mb = 0, sd = 0, sb = 0;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
            i += (unsigned long long)gridDim.x * blockDim.x)
    {
        const double d = (double)a[i] - (double)b[i], r = (double)b[i];
        md = fmax(md, fabs(d)); mb = fmax(mb, fabs(r));
        sd += d * d; sb += r * r;
    }
    sh[0][threadIdx.x] = md; sh[1][threadIdx.x] = mb; sh[2][threadIdx.x] = sd; sh[3][threadIdx.x] = sb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if ((int)threadIdx.x < s)
        {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
            sh[3][threadIdx.x] += sh[3][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = sh[k][0];
}

__global__ void k_fill_ones(void* p, unsigned long long n, int dbl)
{
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
            i += (unsigned long long)gridDim.x * blockDim.x)
    {
        if (dbl) static_cast<double*>(p)[i] = 1.0;
        else static_cast<float*>(p)[i] = 1.0f;
    }
}

// Precision-specific entry points of the public ABI.
struct ApiF
{
    typedef float T;
    static int wwrite(mi355_ctx* c, size_t n, const T* u, const T* v, const T* w, T cell, int G,
            unsigned long long* sk, T* g)
    { return mi355_grid_weights_write_f_dev(c, n, u, v, w, cell, G, sk, g); }
    static int wread(mi355_ctx* c, size_t n, const T* u, const T* v, const T* wi, T* wo, T cell,
            int G, unsigned long long* sk, const T* g)
    { return mi355_grid_weights_read_f_dev(c, n, u, v, wi, wo, cell, G, sk, g); }
    static int grid(mi355_ctx* c, size_t P, const int* sup, int ov, int csh, const T* conv,
            size_t n, const T* u, const T* v, const T* w, const T* vis, const T* wt, T cell,
            T ws, int G, unsigned long long* sk, double* nrm, T* g)
    { return mi355_grid_wproj_f_dev(c, P, sup, ov, csh, conv, n, u, v, w, vis, wt, cell, ws, G, sk, nrm, g); }
    static int grid64(mi355_ctx* c, size_t P, const int* sup, int ov, int csh, const T* conv,
            size_t n, const T* u, const T* v, const T* w, const T* vis, const T* wt, T cell,
            T ws, int G, unsigned long long* sk, double* nrm, double* g)
    { return mi355_grid_wproj_fsel_d_dev(c, P, sup, ov, csh, conv, n, u, v, w, vis, wt, cell, ws, G, sk, nrm, g); }
    static int image(mi355_ctx* c, int G, const T* g, double nrm, T* img)
    { return mi355_grid_to_image_f_dev(c, G, g, nrm, img); }
};
struct ApiD
{
    typedef double T;
    static int wwrite(mi355_ctx* c, size_t n, const T* u, const T* v, const T* w, T cell, int G,
            unsigned long long* sk, T* g)
    { return mi355_grid_weights_write_d_dev(c, n, u, v, w, cell, G, sk, g); }
    static int wread(mi355_ctx* c, size_t n, const T* u, const T* v, const T* wi, T* wo, T cell,
            int G, unsigned long long* sk, const T* g)
    { return mi355_grid_weights_read_d_dev(c, n, u, v, wi, wo, cell, G, sk, g); }
    static int grid(mi355_ctx* c, size_t P, const int* sup, int ov, int csh, const T* conv,
            size_t n, const T* u, const T* v, const T* w, const T* vis, const T* wt, T cell,
            T ws, int G, unsigned long long* sk, double* nrm, T* g)
    { return mi355_grid_wproj_d_dev(c, P, sup, ov, csh, conv, n, u, v, w, vis, wt, cell, ws, G, sk, nrm, g); }
    static int grid64(mi355_ctx* c, size_t P, const int* sup, int ov, int csh, const T* conv,
            size_t n, const T* u, const T* v, const T* w, const T* vis, const T* wt, T cell,
            T ws, int G, unsigned long long* sk, double* nrm, double* g)
    { return grid(c, P, sup, ov, csh, conv, n, u, v, w, vis, wt, cell, ws, G, sk, nrm, g); }
    static int image(mi355_ctx* c, int G, const T* g, double nrm, T* img)
    { return mi355_grid_to_image_d_dev(c, G, g, nrm, img); }
};

// Pinned host allocations of one run, released on every return path.
struct HostArrays
{
    std::vector<void*> ptrs;
    ~HostArrays() { for (void* p : ptrs) (void)hipHostFree(p); }
    int alloc(void** p, size_t bytes)
    {
        *p = nullptr;
        hipError_t e = hipHostMalloc(p, bytes ? bytes : 16, hipHostMallocDefault);
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            return fail(MI355_ERR_OUT_OF_MEMORY, "pinned host allocation of %zu bytes failed", bytes);
        }
        ptrs.push_back(*p);
        return MI355_OK;
    }
};

// Events of one run, destroyed on every return path.
struct Events
{
    std::vector<hipEvent_t> ev;
    ~Events() { for (hipEvent_t e : ev) (void)hipEventDestroy(e); }
    int make(hipEvent_t* e, bool timing)
    {
        hipError_t r = hipEventCreateWithFlags(e, timing ? hipEventDefault : hipEventDisableTiming);
        if (r != hipSuccess)
            return fail(MI355_ERR_HIP, "hipEventCreateWithFlags failed: %s", hipGetErrorString(r));
        ev.push_back(*e);
        return MI355_OK;
    }
};

// Waits for the run's stream when it goes out of scope, so that no copy or
// kernel still uses the buffers that are released after it (error paths).
struct StreamDrain
{
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

// The run, ordered on the context's stream.  Blocks go through two sets of
// pinned host and device input buffers: block b+1 is read from the file
// while block b's copies and passes run; a set is refilled once the event
// recorded after its copies has passed.  Every block's num_skipped goes to
// a slot of its own on the device, read back once at the end, so nothing in
// the block loops waits for the device (the gridding call itself waits for
// its selection to size its buffers).
template <typename Api>
int run(mi355_ctx* ctx, Files& F, const mi355_vis_header& hd, const mi355_kernel_info& ki,
        int check, void* grid_host, size_t grid_capacity, const char* image_path,
        mi355_driver_result* res)
{
    typedef typename Api::T T;
    const int G = ki.grid_size, P = ki.num_w_planes, csh = ki.conv_size_half;
    // main.cpp casts cellsize_rad and w_scale to the precision of the data
    // (:329-330, :374-375)
    const T cell = (T)ki.cellsize_rad, w_scale = (T)ki.w_scale;
    const size_t cap = (size_t)hd.max_times_per_block * (size_t)hd.num_baselines;
    const size_t cells = (size_t)G * G;
    const int nb = hd.num_blocks;
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t st = nullptr;
    DRV_TRY(mi355_ctx_get_stream(ctx, reinterpret_cast<void**>(&st)));

    DevArrays D;
    HostArrays H;
    Events E;
    void *d_conv, *d_w1, *d_w, *d_wgrid, *d_grid, *d_norm, *d_sk;
    void *d_uu[2], *d_vv[2], *d_ww[2], *d_vis[2];
    void *h_conv, *h_uu[2], *h_vv[2], *h_ww[2], *h_vis[2], *h_sk, *h_norm;
    // kernel stack: real and imaginary cubes interleaved (main.cpp:186-199)
    const size_t nk = (size_t)csh * csh * P;
    DRV_TRY(H.alloc(&h_conv, 2 * nk * sizeof(T)));
    for (size_t j = 0; j < nk; ++j)
    {
        static_cast<T*>(h_conv)[2 * j] = static_cast<const T*>(F.kre)[j];
        static_cast<T*>(h_conv)[2 * j + 1] = static_cast<const T*>(F.kim)[j];
    }
    // num_skipped slots: pass 1 per block, pass 2 per block, the FP64 check
    // per block, one for the weights read
    const size_t n_sk = 3 * (size_t)nb + 1;
    DRV_TRY(H.alloc(&h_sk, n_sk * sizeof(unsigned long long)));
    DRV_TRY(H.alloc(&h_norm, 2 * sizeof(double)));
    DRV_TRY(D.alloc(&d_conv, 2 * nk * sizeof(T)));
    for (int k = 0; k < 2; ++k)
    {
        DRV_TRY(D.alloc(&d_uu[k], cap * sizeof(T)));
        DRV_TRY(D.alloc(&d_vv[k], cap * sizeof(T)));
        DRV_TRY(D.alloc(&d_ww[k], cap * sizeof(T)));
        DRV_TRY(D.alloc(&d_vis[k], 2 * cap * sizeof(T)));
        DRV_TRY(H.alloc(&h_uu[k], cap * sizeof(T)));
        DRV_TRY(H.alloc(&h_vv[k], cap * sizeof(T)));
        DRV_TRY(H.alloc(&h_ww[k], cap * sizeof(T)));
        DRV_TRY(H.alloc(&h_vis[k], 2 * cap * sizeof(T)));
    }
    DRV_TRY(D.alloc(&d_w1, cap * sizeof(T)));
    DRV_TRY(D.alloc(&d_w, cap * sizeof(T)));
    DRV_TRY(D.alloc(&d_wgrid, cells * sizeof(T)));
    DRV_TRY(D.alloc(&d_grid, 2 * cells * sizeof(T)));
    DRV_TRY(D.alloc(&d_norm, 2 * sizeof(double)));
    DRV_TRY(D.alloc(&d_sk, n_sk * sizeof(unsigned long long)));
    void* d_grid64 = nullptr;
    if (check) DRV_TRY(D.alloc(&d_grid64, 2 * cells * sizeof(double)));
    hipEvent_t ev_free[2];
    bool pending[2] = {false, false};
    std::vector<hipEvent_t> ev_t(4 * (size_t)nb);   // per block and pass: start, end
    for (int k = 0; k < 2; ++k) DRV_TRY(E.make(&ev_free[k], false));
    for (auto& e : ev_t) DRV_TRY(E.make(&e, true));
    StreamDrain drain{st};

    DRV_HIP(hipMemcpyAsync(d_conv, h_conv, 2 * nk * sizeof(T), hipMemcpyHostToDevice, st));
    DRV_HIP(hipMemsetAsync(d_wgrid, 0, cells * sizeof(T), st));
    DRV_HIP(hipMemsetAsync(d_grid, 0, 2 * cells * sizeof(T), st));
    DRV_HIP(hipMemsetAsync(d_norm, 0, 2 * sizeof(double), st));
    DRV_HIP(hipMemsetAsync(d_sk, 0, n_sk * sizeof(unsigned long long), st));
    DRV_HIP(hipMemsetAsync(d_w, 0, cap * sizeof(T), st));     // (main.cpp callocs it, :185)
    if (check) DRV_HIP(hipMemsetAsync(d_grid64, 0, 2 * cells * sizeof(double), st));
    // unit input weights (main.cpp:195-205)
    hipLaunchKernelGGL(k_fill_ones, dim3(256), dim3(256), 0, st, d_w1, (unsigned long long)cap,
            (int)(sizeof(T) == 8));
    DRV_HIP(hipGetLastError());

    double t_read = 0;
    unsigned long long points = 0;
    size_t n_in[2] = {0, 0};
    // block b into host set k (once the set's previous copies are done)
    auto read = [&](int b, int k, bool amps) -> int {
        if (pending[k])
        {
            DRV_HIP(hipEventSynchronize(ev_free[k]));
            pending[k] = false;
        }
        const auto t0 = std::chrono::steady_clock::now();
        DRV_TRY(mi355_vis_read_block(F.vis, &hd, b, cap, &n_in[k], (T*)h_uu[k], (T*)h_vv[k],
                (T*)h_ww[k], amps ? (T*)h_vis[k] : nullptr));
        t_read += seconds_since(t0);
        return MI355_OK;
    };
    auto upload = [&](int k, bool amps) -> int {
        const size_t n = n_in[k];
        DRV_HIP(hipMemcpyAsync(d_uu[k], h_uu[k], n * sizeof(T), hipMemcpyHostToDevice, st));
        DRV_HIP(hipMemcpyAsync(d_vv[k], h_vv[k], n * sizeof(T), hipMemcpyHostToDevice, st));
        if (amps)
        {
            DRV_HIP(hipMemcpyAsync(d_ww[k], h_ww[k], n * sizeof(T), hipMemcpyHostToDevice, st));
            DRV_HIP(hipMemcpyAsync(d_vis[k], h_vis[k], 2 * n * sizeof(T), hipMemcpyHostToDevice, st));
        }
        DRV_HIP(hipEventRecord(ev_free[k], st));
        pending[k] = true;
        return MI355_OK;
    };
    unsigned long long* sk = static_cast<unsigned long long*>(d_sk);

    // Pass 1: the uniform-weighting grid over every block (main.cpp:202-239).
    if (nb > 0) DRV_TRY(read(0, 0, false));
    for (int b = 0; b < nb; ++b)
    {
        const int k = b & 1;
        const size_t n = n_in[k];
        DRV_HIP(hipEventRecord(ev_t[4 * b], st));
        DRV_TRY(upload(k, false));
        DRV_TRY(Api::wwrite(ctx, n, (const T*)d_uu[k], (const T*)d_vv[k], (const T*)d_w1, cell, G,
                sk + b, (T*)d_wgrid));
        DRV_HIP(hipEventRecord(ev_t[4 * b + 1], st));
        if (b + 1 < nb) DRV_TRY(read(b + 1, k ^ 1, false));
    }

    // Pass 2: weights read and gridding, block by block (main.cpp:242-381).
    if (nb > 0) DRV_TRY(read(0, nb & 1, true));
    for (int b = 0; b < nb; ++b)
    {
        const int k = (b + nb) & 1;
        const size_t n = n_in[k];
        points += n;
        DRV_HIP(hipEventRecord(ev_t[4 * b + 2], st));
        DRV_TRY(upload(k, true));
        DRV_TRY(Api::wread(ctx, n, (const T*)d_uu[k], (const T*)d_vv[k], (const T*)d_w1, (T*)d_w,
                cell, G, sk + 3 * nb, (const T*)d_wgrid));
        DRV_TRY(Api::grid(ctx, (size_t)P, F.support, ki.oversample, csh, (const T*)d_conv, n,
                (const T*)d_uu[k], (const T*)d_vv[k], (const T*)d_ww[k], (const T*)d_vis[k],
                (const T*)d_w, cell, w_scale, G, sk + nb + b, (double*)d_norm, (T*)d_grid));
        if (check)
            DRV_TRY(Api::grid64(ctx, (size_t)P, F.support, ki.oversample, csh, (const T*)d_conv, n,
                    (const T*)d_uu[k], (const T*)d_vv[k], (const T*)d_ww[k], (const T*)d_vis[k],
                    (const T*)d_w, cell, w_scale, G, sk + 2 * nb + b, (double*)d_norm + 1,
                    (double*)d_grid64));
        DRV_HIP(hipEventRecord(ev_t[4 * b + 3], st));
        if (b + 1 < nb) DRV_TRY(read(b + 1, k ^ 1, true));
    }
    DRV_HIP(hipMemcpyAsync(h_sk, d_sk, n_sk * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    DRV_HIP(hipMemcpyAsync(h_norm, d_norm, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    DRV_HIP(hipStreamSynchronize(st));
    const double t_wall = seconds_since(t_start);
    double t_weights = 0, t_grid = 0;
    unsigned long long skipped_total = 0;
    for (int b = 0; b < nb; ++b)
    {
        float ms1 = 0, ms2 = 0;
        DRV_HIP(hipEventElapsedTime(&ms1, ev_t[4 * b], ev_t[4 * b + 1]));
        DRV_HIP(hipEventElapsedTime(&ms2, ev_t[4 * b + 2], ev_t[4 * b + 3]));
        t_weights += 1e-3 * ms1;
        t_grid += 1e-3 * ms2;
        skipped_total += static_cast<unsigned long long*>(h_sk)[nb + b];
    }
    const double* norm = static_cast<const double*>(h_norm);

    res->num_blocks = nb;
    res->grid_size = G;
    res->num_w_planes = P;
    res->precision = sizeof(T) == 8 ? MI355_BIN_DOUBLE : MI355_BIN_SINGLE;
    res->num_points = points;
    res->num_skipped_total = skipped_total;
    res->num_skipped_last = nb > 0 ? static_cast<unsigned long long*>(h_sk)[2 * nb - 1] : 0;
    res->norm = norm[0];
    res->seconds_read = t_read;
    res->seconds_weights = t_weights;
    res->seconds_grid = t_grid;
    res->seconds_wall = t_wall;
    res->check_max_rel = res->check_l2_rel = -1.0;
    if (check)
    {
        const unsigned blocks = 1024;
        void *d_part, *h_part;
        DRV_TRY(D.alloc(&d_part, 4 * blocks * sizeof(double)));
        DRV_TRY(H.alloc(&h_part, 4 * blocks * sizeof(double)));
        hipLaunchKernelGGL((k_grid_error<T, double>), dim3(blocks), dim3(256), 0, st,
                (const T*)d_grid, (const double*)d_grid64, (unsigned long long)(2 * cells),
                (double*)d_part);
        DRV_HIP(hipGetLastError());
        DRV_HIP(hipMemcpyAsync(h_part, d_part, 4 * blocks * sizeof(double), hipMemcpyDeviceToHost, st));
        DRV_HIP(hipStreamSynchronize(st));
        const double* part = static_cast<const double*>(h_part);
        double md = 0, mb = 0, sd = 0, sb = 0;
        for (unsigned k = 0; k < blocks; ++k)
        {
            md = std::fmax(md, part[4 * k]); mb = std::fmax(mb, part[4 * k + 1]);
            sd += part[4 * k + 2]; sb += part[4 * k + 3];
        }
        res->check_max_rel = mb > 0 ? md / mb : md;
        res->check_l2_rel = sb > 0 ? std::sqrt(sd / sb) : std::sqrt(sd);
    }
    if (grid_host)
    {
        if (grid_capacity < cells)
            return fail(MI355_ERR_INVALID_ARG, "grid holds %zu cells, need %zu", grid_capacity, cells);
        DRV_HIP(hipMemcpyAsync(grid_host, d_grid, 2 * cells * sizeof(T), hipMemcpyDeviceToHost, st));
        DRV_HIP(hipStreamSynchronize(st));
    }
    if (image_path)
    {
        void* d_img;
        DRV_TRY(D.alloc(&d_img, cells * sizeof(T)));
        DRV_TRY(Api::image(ctx, G, (const T*)d_grid, norm[0], (T*)d_img));
        std::vector<T> img(cells);
        DRV_HIP(hipMemcpyAsync(img.data(), d_img, cells * sizeof(T), hipMemcpyDeviceToHost, st));
        DRV_HIP(hipStreamSynchronize(st));
        DRV_TRY(mi355_fits_write_image(image_path, res->precision, G, G, img.data()));
    }
    return ok();
}

} // namespace

extern "C" {

MI355_EXPORT int mi355_grid_vis_file(mi355_ctx* ctx, const char* root, int check, void* grid,
        size_t grid_capacity, const char* image_path, mi355_driver_result* res)
{
    if (!ctx || !root || !res) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    memset(res, 0, sizeof(*res));
    Files F;
    const std::string r = root;
    DRV_TRY(mi355_bin_open((r + ".vis").c_str(), 'r', &F.vis));
    mi355_vis_header hd;
    DRV_TRY(mi355_vis_read_header(F.vis, &hd));
    // single channel, single polarisation (main.cpp:162-171)
    if (hd.num_channels_total != 1 || (hd.amp_type & MI355_BIN_MATRIX))
        return fail(MI355_ERR_INVALID_ARG, "the gridder needs single-channel, "
                "single-polarisation data (%d channels, amplitude type 0x%02x)",
                hd.num_channels_total, hd.amp_type);
    const int prec = hd.amp_type & 0x0F;
    if (prec != MI355_BIN_SINGLE && prec != MI355_BIN_DOUBLE)
        return fail(MI355_ERR_FORMAT, "amplitude type 0x%02x", hd.amp_type);
    // the harness hands the coordinate arrays to the gridder as arrays of the
    // amplitude precision (main.cpp:297-332), so the two must agree
    if (hd.coord_precision != prec)
        return fail(MI355_ERR_FORMAT, "coordinate precision 0x%02x differs from the amplitudes' "
                "0x%02x", hd.coord_precision, prec);
    mi355_kernel_info ki, ki2;
    DRV_TRY(mi355_fits_read_kernel((r + "_KERNELS_REAL.fits").c_str(), prec, &ki, &F.kre, &F.support));
    DRV_TRY(mi355_fits_read_kernel((r + "_KERNELS_IMAG.fits").c_str(), prec, &ki2, &F.kim, nullptr));
    if (ki.conv_size_half != ki2.conv_size_half || ki.num_w_planes != ki2.num_w_planes)
        return fail(MI355_ERR_FORMAT, "real and imaginary kernel cubes differ in shape");
    return prec == MI355_BIN_DOUBLE
            ? run<ApiD>(ctx, F, hd, ki, check, grid, grid_capacity, image_path, res)
            : run<ApiF>(ctx, F, hd, ki, check, grid, grid_capacity, image_path, res);
}

} // extern "C"
