// synth.cpp -- seeded synthetic inputs for the W-projection gridder.
//
// The reference's data sets (an OSKAR .vis file and two FITS kernel cubes,
// read at main.cpp:92-189) are a 3.2 GB download that is not available
// offline, so every workload here is synthetic, with the shapes of the
// reference's data sets (README:126-134).  Everything is a deterministic
// function of (seed, index): draws use a counter-based hash, so the output
// does not depend on the number of threads.
//
// Kernels: plane k has support s_k = ceil(4 + (smax-4)*sqrt(k/(P-1))); its
// first quadrant (conv_size_half^2 samples at `oversample` samples per cell,
// the layout read by main.cpp:158-189) is a Gaussian taper times a chirp,
//   c(r) = exp(-r^2 / (2 sigma_k^2)) * exp(i*pi*a_k*r^2),  r^2 = (ix^2+iy^2)/ov^2,
// with sigma_k = s_k / sqrt(2 ln 1000) (taper 1e-3 at r = s_k) and
// a_k = 4 (k/(P-1)) / s_k^2 (two turns of phase at r = s_k on the last plane).
//
// Visibilities: radius r from an exponential (scale 0.15 umax) or a
// log-normal (median 0.05 umax, sigma_ln 1.2, clipped at umax) law, angle
// uniform, ww ~ N(0, (0.3 r)^2).  With num_times > 1 each of num_baselines
// baselines draws (r, theta, w) once and rotates through num_times one-minute
// steps of hour angle; visibilities are stored time-major (all baselines of
// time 0, then time 1, ...), as a .vis block holds them.  Amplitudes are
// complex normal (re, im ~ N(0, 1/2)); weights are 1.
//
// The geometry is fixed so that footprints just fit: umax = 1e4 wavelengths,
// cell_size_rad = (G/2 - smax - 2) / (umax * G), and
// w_scale = (P-1)^2 / w_ref with w_ref the 99.9th percentile of |ww|, so that
// the plane index sqrt(|w| * w_scale) (oskar_grid_wproj.c:162) runs from 0 to
// P-1 as |w| runs from 0 to w_ref and the top 0.1% exercise the clamp to the
// last plane (:176-179).  (A scale of (P-1)^2 / w_ref^2 would put almost every
// visibility on the first few planes.)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#define SYNTH_API extern "C" __attribute__((visibility("default")))

namespace {

inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in (0, 1): 53 random bits, never exactly 0.
inline double unit(uint64_t seed, uint64_t stream, uint64_t index, uint64_t k)
{
    const uint64_t h = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull ^
            mix64(index * 0xABC98388FB8FAC03ull + k)));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

inline double normal(uint64_t seed, uint64_t stream, uint64_t index, uint64_t k)
{
    const double u1 = unit(seed, stream, index, 2 * k);
    const double u2 = unit(seed, stream, index, 2 * k + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

const double kUmax = 1.0e4;

} // namespace

// Support schedule s_k = ceil(4 + (smax-4) sqrt(k/(P-1))).
SYNTH_API int mi355_synth_support(int num_w_planes, int max_support, int* support)
{
    if (num_w_planes < 1 || max_support < 1 || !support) return 1;
    for (int k = 0; k < num_w_planes; ++k)
    {
        const double f = num_w_planes > 1 ? (double)k / (num_w_planes - 1) : 1.0;
        int s = (int)std::ceil(4.0 + (max_support - 4) * std::sqrt(f));
        if (s > max_support) s = max_support;
        if (s < 1) s = 1;
        support[k] = s;
    }
    return 0;
}

// Kernel stack [P][csh][csh] complex float (re, im interleaved), the layout of
// the `kernels` array built at main.cpp:167-189.
SYNTH_API int mi355_synth_kernels(int num_w_planes, int max_support,
        int oversample, int conv_size_half, int* support, float* conv)
{
    if (mi355_synth_support(num_w_planes, max_support, support)) return 1;
    if (oversample < 1 || conv_size_half < 1 || !conv) return 1;
    const size_t plane_len = (size_t)conv_size_half * conv_size_half;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < num_w_planes; ++k)
    {
        const double s = support[k];
        const double sigma = s / std::sqrt(2.0 * std::log(1000.0));
        const double f = num_w_planes > 1 ? (double)k / (num_w_planes - 1) : 1.0;
        const double a = 4.0 * f / (s * s);
        float* p = conv + 2 * plane_len * (size_t)k;
        for (int iy = 0; iy < conv_size_half; ++iy)
            for (int ix = 0; ix < conv_size_half; ++ix)
            {
                const double r2 = ((double)ix * ix + (double)iy * iy) /
                        ((double)oversample * oversample);
                const double amp = std::exp(-r2 / (2.0 * sigma * sigma));
                const double ph = M_PI * a * r2;
                float* c = p + 2 * ((size_t)iy * conv_size_half + ix);
                c[0] = (float)(amp * std::cos(ph));
                c[1] = (float)(amp * std::sin(ph));
            }
    }
    return 0;
}

// Cell size that makes a footprint at radius umax just fit the grid.
SYNTH_API double mi355_synth_cell_size(int grid_size, int max_support)
{
    return (grid_size / 2 - max_support - 2) / (kUmax * grid_size);
}

// Visibilities.  dist: 0 = exponential radius (scale 0.15 umax),
// 1 = log-normal radius (median 0.05 umax, sigma_ln 1.2, clipped at umax).
// num_points = num_baselines * num_times when num_times > 1 (tracks,
// time-major); with num_times <= 1, num_points independent draws.
// Returns w_scale = (P-1)^2 / w_ref through *w_scale_out.
SYNTH_API int mi355_synth_vis(unsigned long long seed, int dist,
        size_t num_points, size_t num_baselines, size_t num_times,
        int num_w_planes, float* uu, float* vv, float* ww, float* vis,
        float* weight, double* w_scale_out)
{
    if (!uu || !vv || !ww || !vis || !weight || num_w_planes < 1) return 1;
    const bool tracks = num_times > 1;
    if (tracks && num_baselines * num_times != num_points) return 1;
    const size_t n_src = tracks ? num_baselines : num_points;
    const double step = 2.0 * M_PI * 60.0 / 86400.0;  // one minute of hour angle
    #pragma omp parallel for schedule(static)
    for (long long bl = 0; bl < (long long)n_src; ++bl)
    {
        const uint64_t b = (uint64_t)bl;
        double r;
        if (dist == 0)
            r = -std::log(unit(seed, 1, b, 0)) * 0.15 * kUmax;
        else
        {
            r = 0.05 * kUmax * std::exp(1.2 * normal(seed, 1, b, 1));
            if (r > kUmax) r = kUmax;
        }
        const double theta = 2.0 * M_PI * unit(seed, 2, b, 0);
        const double w0 = 0.3 * r * normal(seed, 3, b, 0);
        const size_t nt = tracks ? num_times : 1;
        for (size_t t = 0; t < nt; ++t)
        {
            const size_t i = tracks ? t * num_baselines + b : b;
            const double ha = tracks ? ((double)t - 0.5 * (nt - 1)) * step : 0.0;
            uu[i] = (float)(r * std::cos(theta + ha));
            vv[i] = (float)(r * std::sin(theta + ha));
            ww[i] = (float)(w0 + 0.05 * r * std::sin(ha));
        }
    }
    #pragma omp parallel for schedule(static)
    for (long long il = 0; il < (long long)num_points; ++il)
    {
        const uint64_t i = (uint64_t)il;
        vis[2 * i]     = (float)(std::sqrt(0.5) * normal(seed, 4, i, 0));
        vis[2 * i + 1] = (float)(std::sqrt(0.5) * normal(seed, 4, i, 1));
        weight[i] = 1.0f;
    }
    if (w_scale_out)
    {
        std::vector<float> aw(num_points);
        for (size_t i = 0; i < num_points; ++i) aw[i] = std::fabs(ww[i]);
        double w_ref = 1.0;
        if (num_points)
        {
            const size_t q = (size_t)(0.999 * (double)(num_points - 1));
            std::nth_element(aw.begin(), aw.begin() + q, aw.end());
            w_ref = aw[q] > 0.0f ? aw[q] : 1.0;
        }
        *w_scale_out = (double)(num_w_planes - 1) * (num_w_planes - 1) / w_ref;
    }
    return 0;
}
