// gg_device.hpp -- device-side types and helpers shared by the gridder kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

namespace gg {

// ---- Tile geometry --------------------------------------------------------
// A grid tile is 64 columns x 64 rows.  One workgroup owns one tile; its four
// wave64s own 16 rows each, lane x owning column x of every row of its wave.
constexpr int kTileW = 64;
constexpr int kLogTileW = 6;
constexpr int kWaves = 4;
constexpr int kRowsPerWave = 16;
constexpr int kTileH = kWaves * kRowsPerWave;
constexpr int kLogTileH = 6;
constexpr int kTileCells = kTileW * kTileH;
constexpr int kBlock = kWaves * 64;

// ---- Per-visibility record ------------------------------------------------
// meta  = int4 {grid_u, grid_v, packed, table_base}
//   packed bits 0-4 flags, 5-16 support, 17-31 kernel-table id (for
//   ordering; shifted right as needed to fit 15 bits)
//   table_base = element offset of the expanded kernel table in use
// amp   = T4 {a, b, c, d}: the four FMA coefficients of the complex MAC
//   re += c_re*a + c_im*b ;  im += c_re*c + c_im*d
constexpr int kFlagNegV = 1;   // off_v < 0: table rows read mirrored
constexpr int kFlagNegU = 2;   // off_u < 0: table columns read mirrored
constexpr int kFlagConj = 4;   // ww > 0: Im(kernel) negated
constexpr int kFlagSkip = 8;   // footprint leaves the grid
constexpr int kFlagNonFinite = 16;  // weight*vis not finite: kept out of the
                                    // tiled path and gridded by k_grid_nonfinite
constexpr int kMaxSupport = 4095;
constexpr int kSupportShift = 5;
constexpr int kTableIdShift = 17;
constexpr int kTableIdBits = 15;
constexpr int kOrientMask = kFlagNegV | kFlagNegU;   // mirroring of a table read
// In the sorted visit records of the strip kernel (never skipped, never
// non-finite) the skip bit marks a visit whose run is staged in LDS and the
// non-finite bit the first visit of a run.
constexpr int kVisitStaged = kFlagSkip;
constexpr int kVisitRunStart = kFlagNonFinite;
// ... and the table-id bits (not needed past the sort) hold the table's
// symmetry: bits 17-18 its columns, 19-20 its rows, 1 = symmetric about the
// centre tap (offset 0: tap index |ov k|), 2 = about the half-tap before it
// (offset ov/2, ov even: |ov/2 + ov k| = |ov/2 + ov (-1-k)|), 0 = neither.
constexpr int kVisitFoldShift = kTableIdShift;
// Expanded kernel tables carry this many zero rows above and below each
// (2s+1)^2 table, so that reads of up to that many rows past a footprint
// edge return zero taps.
constexpr int kTablePadRows = 7;
// The table buffer starts and ends with this many rows of the widest table
// (zeros), so that reads running up to that many rows past any table stay
// inside the buffer.
constexpr int kTableGuardRows = 16;

__device__ __forceinline__ int meta_support(int packed) { return (packed >> kSupportShift) & 0xFFF; }
__device__ __forceinline__ unsigned meta_table_id(int packed) { return (unsigned)packed >> kTableIdShift; }

template <typename T> struct V2;
template <> struct V2<float> { typedef float2 t; };
template <> struct V2<double> { typedef double2 t; };
template <typename T> struct V4;
template <> struct V4<float> { typedef float4 t; };
template <> struct V4<double> { typedef double4 t; };

// Per-job parameters (a single call is one job).  T is the accumulation
// type; the selection constants are kept in the input precision, which is
// float for the FP32 entry points and double for the FP64 ones.
template <typename T>
struct JobDev
{
    typename V2<T>::t* grid;   // G*G complex, row-major
    double* norm;              // accumulated
    unsigned long long* skipped;
    float grid_scale_f;        // grid_size * cell_size_rad in float (oskar_grid_wproj.c:148)
    float w_scale_f;
    double grid_scale_d;       // ... in double (oskar_grid_wproj.c:59)
    double w_scale_d;
    int grid_size;
    int ntx, nty;              // tiles across / down
    unsigned tile_base;        // first global tile id of this job
    unsigned long long vis_begin, vis_end;
    unsigned chunk_base;       // first selection chunk of this job
};

// Selection constants of a job in the selection precision.
template <typename JD>
__device__ __forceinline__ void job_scales(const JD& j, float& gs, float& ws)
{
    gs = j.grid_scale_f; ws = j.w_scale_f;
}
template <typename JD>
__device__ __forceinline__ void job_scales(const JD& j, double& gs, double& ws)
{
    gs = j.grid_scale_d; ws = j.w_scale_d;
}

// ---- x86-64 conversion semantics ------------------------------------------
// The serial reference is C compiled for x86-64, where (int)x of a float or
// double out of int range or NaN yields INT_MIN (cvtts?2si), and (size_t)x
// is emitted as a compare with 2^63 followed by cvtts?2si: NaN gives 2^63,
// values >= 2^64 (and +Inf) give 0.  The GPU's own conversions saturate
// instead, so selection mirrors the x86 results explicitly.
__device__ __forceinline__ int x86_int(float x)
{
    return (x >= -2147483648.0f && x < 2147483648.0f) ? (int)x : INT_MIN;
}
__device__ __forceinline__ int x86_int(double x)
{
    return (x > -2147483649.0 && x < 2147483648.0) ? (int)x : INT_MIN;
}
template <typename T>
__device__ __forceinline__ unsigned long long x86_size(T x)
{
    if (x != x) return 0x8000000000000000ull;
    if (x >= (T)18446744073709551616.0) return 0ull;
    return (unsigned long long)x;
}
// Two's-complement wrap, as the x86 add/sub the reference compiles to.
__device__ __forceinline__ int wrap_add(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
__device__ __forceinline__ int wrap_sub(int a, int b) { return (int)((unsigned)a - (unsigned)b); }

__device__ __forceinline__ float gg_round(float x) { return roundf(x); }
__device__ __forceinline__ double gg_round(double x) { return round(x); }
// llvm.sqrt: correctly rounded in both precisions under HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt (the __fsqrt_rn header wrapper may
// map to the approximate native instruction, so it is not used).
__device__ __forceinline__ float gg_sqrt_rn(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double gg_sqrt_rn(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float gg_abs(float x) { return fabsf(x); }
__device__ __forceinline__ double gg_abs(double x) { return fabs(x); }

struct Selection
{
    int grid_u, grid_v, off_u, off_v, plane, support, conj, skip;
};

// Cell, oversampled offset and w-plane of one visibility: the selection of
// oskar_grid_wproj.c:158-187 (T = float) and :69-98 (T = double), operation
// for operation.  Contraction is off here so that (round(p) - p) with
// p = -u*gs is not fused into an FMA, which would change off_u.
template <typename T>
__device__ __forceinline__ Selection select_vis(T u, T v, T w, T grid_scale,
        T w_scale, int grid_size, int oversample, int num_planes,
        const int* __restrict__ support)
{
#pragma clang fp contract(off)
    Selection o;
    const int centre = grid_size / 2;
    const T pos_u = -u * grid_scale;
    const T pos_v = v * grid_scale;
    const unsigned long long gw = x86_size(gg_round(gg_sqrt_rn(gg_abs(w * w_scale))));
    const T ru = gg_round(pos_u);
    const T rv = gg_round(pos_v);
    o.conj = (w > (T)0);
    o.grid_u = wrap_add(x86_int(ru), centre);
    o.grid_v = wrap_add(x86_int(rv), centre);
    o.off_u = x86_int(gg_round((ru - pos_u) * (T)oversample));
    o.off_v = x86_int(gg_round((rv - pos_v) * (T)oversample));
    o.plane = gw < (unsigned long long)num_planes ? (int)gw : num_planes - 1;
    o.support = support[o.plane];
    o.skip = (wrap_add(o.grid_u, o.support) >= grid_size ||
              wrap_sub(o.grid_u, o.support) < 0 ||
              wrap_add(o.grid_v, o.support) >= grid_size ||
              wrap_sub(o.grid_v, o.support) < 0);
    return o;
}

// ---- Scalar (constant address space) loads --------------------------------
// Data written by an earlier kernel and only read by the current one may be
// read through address space 4: a wave-uniform address then compiles to an
// s_load into scalar registers (one load per wave, values uniform by
// construction) instead of a per-lane vector load.
template <typename X>
__device__ __forceinline__ X ld_const(const X* p, unsigned long long i)
{
    static_assert(sizeof(X) % 4 == 0, "scalar loads move whole dwords");
    typedef __attribute__((address_space(4))) const unsigned CU;
    const CU* q = (const CU*)(p + i);
    X out;
    unsigned* o = reinterpret_cast<unsigned*>(&out);
    #pragma unroll
    for (unsigned k = 0; k < sizeof(X) / 4; ++k) o[k] = q[k];
    return out;
}

// ---- wave64 helpers -------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ unsigned long long lanemask_lt()
{
    return (1ull << lane_id()) - 1ull;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x)
{
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Workgroup b of a launch padded to a multiple of 8 runs on XCD b % 8; give
// each XCD a contiguous range of work so that neighbouring tiles share an L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned padded)
{
    const unsigned per = padded >> 3;
    return (b & 7u) * per + (b >> 3);
}

} // namespace gg
