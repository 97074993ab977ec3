// gg_grid_staged.hpp -- tile gridding with kernel tables staged in LDS.
//
// Visits of a tile are sorted by kernel table, so a tile's list is a
// sequence of runs that all read the same (2s+1)^2 table.  A workgroup of
// eight waves owns one 64x64 tile: four 16-row strips, two waves per strip
// splitting the strip's visits (even / odd positions).  At the start of a run
// of two or more visits whose table fits, the workgroup copies the table
// into LDS once (rows at a fixed stride, three zero rows above and below);
// every visit of the run then reads its taps from LDS, four rows per group
// with immediate offsets.  Rows of a group outside the footprint land on the
// zero rows, so no per-row branches are needed.  Single visits, tables wider
// than the stride and non-finite amplitudes use the direct path (buffer
// loads from the table in global memory).
//
// Per-visit records (meta, amp) are fetched 64 visits at a time per wave --
// one vector load per lane -- and broadcast with readlane, because scalar
// loads complete out of order and cannot be prefetched deeply.
//
// Accumulators stay in registers (16 rows x re/im per lane).  At the end of
// a work item the two waves of a strip add their partials in a fixed order
// through LDS, and the strip is written to the grid (single-item tiles) or
// to the item's partial slot (merged later in item order, FP64).
#pragma once

#include "gg_kernels.hpp"

namespace gg {

template <typename T> struct Stage;
template <> struct Stage<float>  { static constexpr int kStride = 96; };  // tables up to 95^2
template <> struct Stage<double> { static constexpr int kStride = 80; };  // tables up to 79^2
constexpr int kPadRows = 3;
constexpr int kStageLanes = 2;                        // waves per strip
constexpr int kStageWaves = kWaves * kStageLanes;     // 8
constexpr int kStageBlock = kStageWaves * 64;         // 512 threads
constexpr int kBatch = 64;                            // visit records per wave fetch

template <typename T>
__host__ __device__ constexpr int stage_lds_elems()
{
    return (Stage<T>::kStride + 2 * kPadRows) * Stage<T>::kStride;
}

__device__ __forceinline__ int bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float bcast(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double bcast(double v, int l)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, l);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <typename T>
__device__ __forceinline__ void mac4(const typename V2<T>::t (&c)[4], const typename V4<T>::t& a,
        T* re, T* im)
{
    #pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        re[r] = fma(c[r].x, a.x, re[r]);
        re[r] = fma(c[r].y, a.y, re[r]);
        im[r] = fma(c[r].x, a.z, im[r]);
        im[r] = fma(c[r].y, a.w, im[r]);
    }
}

// The visit functions have no divergent control flow (only wave-uniform
// skips): lanes outside the footprint's columns read zeros, and multiplying
// the (finite) coefficients by zero leaves their accumulators unchanged.
// (Amplitudes are finite here: k_select keeps non-finite ones out of the
// tile lists.)

// One visit, taps from the staged table in LDS.  Inactive lanes read column
// S-1, which no staged table uses and staging fills with zeros.
template <typename T>
__device__ __forceinline__ void visit_lds(const int4 mt, const typename V4<T>::t a, int x0,
        int ys, int lane, const typename V2<T>::t* tab, T (&re)[kRowsPerWave],
        T (&im)[kRowsPerWave])
{
    typedef typename V2<T>::t T2;
    constexpr int S = Stage<T>::kStride;
    const int s = meta_support(mt.z);
    const int j0 = ys - mt.y;
    if (j0 > s || j0 + kRowsPerWave - 1 < -s) return;
    const int k = x0 + lane - mt.x;
    const bool act = k >= -s && k <= s;
    const int kcol = act ? ((mt.z & kFlagNegU) ? s - k : s + k) : S - 1;
    const bool neg_v = mt.z & kFlagNegV;
    // Two groups of four rows are read before their multiplies (four: no
    // faster; one: 1.5% slower for the FP64 gridder).
    constexpr int GL = 2;
    #pragma unroll
    for (int g0 = 0; g0 < kRowsPerWave; g0 += 4 * GL)
    {
        T2 c[GL][4];
        #pragma unroll
        for (int u = 0; u < GL; ++u)
        {
            const int g = g0 + 4 * u;
            const int jg = j0 + g;
            if (g >= kRowsPerWave || jg > s || jg + 3 < -s) continue;
            if (!neg_v)
            {
                const T2* p = tab + (kPadRows + s + jg) * S + kcol;     // rows increase with r
                c[u][0] = p[0]; c[u][1] = p[S]; c[u][2] = p[2 * S]; c[u][3] = p[3 * S];
            }
            else
            {
                const T2* p = tab + (kPadRows + s - jg - 3) * S + kcol; // row of r = 3 first
                c[u][3] = p[0]; c[u][2] = p[S]; c[u][1] = p[2 * S]; c[u][0] = p[3 * S];
            }
        }
        #pragma unroll
        for (int u = 0; u < GL; ++u)
        {
            const int g = g0 + 4 * u;
            const int jg = j0 + g;
            if (g >= kRowsPerWave || jg > s || jg + 3 < -s) continue;
            mac4<T>(c[u], a, re + g, im + g);
        }
    }
}

// One visit, taps straight from the table in global memory (buffer loads).
// Rows outside the footprint read the zero row; inactive lanes use an
// offset past the buffer's end, for which raw buffer loads return zero.
template <typename T>
__device__ __forceinline__ void visit_direct(const int4 mt, const typename V4<T>::t a, int x0,
        int ys, int lane, __amdgpu_buffer_rsrc_t rs, unsigned zero_row_byte,
        T (&re)[kRowsPerWave], T (&im)[kRowsPerWave])
{
    typedef typename V2<T>::t T2;
    const int s = meta_support(mt.z);
    const int j0 = ys - mt.y;
    if (j0 > s || j0 + kRowsPerWave - 1 < -s) return;
    const int W = 2 * s + 1;
    const int k = x0 + lane - mt.x;
    const bool act = k >= -s && k <= s;
    const unsigned kbyte = act
            ? (unsigned)((mt.z & kFlagNegU) ? s - k : s + k) * (unsigned)sizeof(T2)
            : 0x80000000u;
    const bool neg_v = mt.z & kFlagNegV;
    // The loads of every touched group of four rows are issued before any
    // multiply, so that a wave has all of a visit's taps in flight at once
    // (the pass waits on memory latency: one group at a time took 27% longer
    // for the FP64 gridder at support 72, two groups 9%).
    constexpr int GD = kRowsPerWave / 4;
    #pragma unroll
    for (int g0 = 0; g0 < kRowsPerWave; g0 += 4 * GD)
    {
        T2 c[GD][4];
        #pragma unroll
        for (int u = 0; u < GD; ++u)
        {
            const int g = g0 + 4 * u;
            const int jg = j0 + g;
            if (g >= kRowsPerWave || jg > s || jg + 3 < -s) continue;
            #pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                const int jr = jg + r;
                const unsigned soff = (jr >= -s && jr <= s)
                        ? ((unsigned)mt.w + (unsigned)((neg_v ? s - jr : s + jr) * W)) *
                          (unsigned)sizeof(T2)
                        : zero_row_byte;
                c[u][r] = load_tap(rs, kbyte, soff, (T2*)nullptr);
            }
        }
        #pragma unroll
        for (int u = 0; u < GD; ++u)
        {
            const int g = g0 + 4 * u;
            const int jg = j0 + g;
            if (g >= kRowsPerWave || jg > s || jg + 3 < -s) continue;
            mac4<T>(c[u], a, re + g, im + g);
        }
    }
}

// Copy a W x W table (row-major in global memory) into LDS rows
// kPadRows .. kPadRows+W-1 at stride S, zero rows around it.  Wave w copies
// rows w, w+8, ...; lane l copies columns l and l+64 of a row.  Rows are
// taken four at a time so that a wave's loads are in flight together.
template <typename T>
__device__ __forceinline__ void stage_table(typename V2<T>::t* tab, const typename V2<T>::t* src,
        int W, int wave, int lane)
{
    typedef typename V2<T>::t T2;
    constexpr int S = Stage<T>::kStride;
    const int rows = W + 2 * kPadRows;
    for (int r0 = wave; r0 < rows; r0 += 4 * kStageWaves)
    {
        T2 v[4][2];
        #pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int jj = r0 + u * kStageWaves - kPadRows;
            #pragma unroll
            for (int h = 0; h < 2; ++h)
            {
                const int c = lane + 64 * h;
                v[u][h].x = (T)0; v[u][h].y = (T)0;
                if (jj >= 0 && jj < W && c < W) v[u][h] = src[jj * W + c];
            }
        }
        #pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int r = r0 + u * kStageWaves;
            #pragma unroll
            for (int h = 0; h < 2; ++h)
            {
                const int c = lane + 64 * h;
                if (r < rows && c < S) tab[r * S + c] = v[u][h];
            }
        }
    }
}

// Plain FP32: two workgroups per CU (LDS) -> four waves per SIMD, 128 VGPRs
// each.  The compensated and FP64 variants carry twice the accumulator
// state and get 256 VGPRs (two waves per SIMD) instead of spilling.
template <typename T, bool COMP>
constexpr int stage_waves_per_eu() { return (COMP || sizeof(T) == 8) ? 2 : 4; }

template <typename T, bool COMP>
__global__ __launch_bounds__(kStageBlock)
__attribute__((amdgpu_waves_per_eu(stage_waves_per_eu<T, COMP>())))
void k_grid_staged(const GridArgs<T> A)
{
    typedef typename V2<T>::t T2;
    typedef typename V4<T>::t T4;
    constexpr int S = Stage<T>::kStride;
    __shared__ T2 tab[stage_lds_elems<T>()];
    static_assert(stage_lds_elems<T>() >= kWaves * kRowsPerWave * 64 * (COMP ? 2 : 1),
                  "combine buffer must fit in the table buffer");
    const unsigned num_items = ld_const(A.d_num_items, 0);
    const unsigned num_runs = (unsigned)ld_const(A.d_num_runs, 0);
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int strip = wave & (kWaves - 1);
    const int q = wave >> 2;                       // 0 or 1: visit parity
    const __amdgpu_buffer_rsrc_t rs = table_rsrc(A.tables, A.table_bytes);

    for (unsigned item = blockIdx.x; item < num_items; item += gridDim.x)
    {
        const unsigned t = ld_const(A.item_tile, item);
        const unsigned ioff = ld_const(A.item_off, t);
        const unsigned m = item - ioff;
        const unsigned nit = ld_const(A.item_off, t + 1) - ioff;
        const unsigned v0 = ld_const(A.tile_start, t) + m * A.visits_per_item;
        const unsigned v1 = min(v0 + A.visits_per_item, ld_const(A.tile_end, t));
        int j = 0;
        {
            int lo = 0, hi = A.num_jobs - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (ld_const(A.job_tile_base, mid) <= t) lo = mid; else hi = mid - 1;
            }
            j = lo;
        }
        const JobDev<T> J = ld_const(A.jobs, j);
        const int G = J.grid_size;
        const unsigned local = t - J.tile_base;
        const int ty = (int)(local / (unsigned)J.ntx);
        const int tx = (int)(local - (unsigned)ty * J.ntx);
        const int x0 = tx * kTileW;
        const int ys = ty * kTileH + strip * kRowsPerWave;

        T re[kRowsPerWave], im[kRowsPerWave];
        T sre[kRowsPerWave], sim[kRowsPerWave], cre[kRowsPerWave], cim[kRowsPerWave];
        #pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r)
        {
            re[r] = im[r] = (T)0;
            if (COMP) { sre[r] = sim[r] = cre[r] = cim[r] = (T)0; }
        }
        int since = 0;
        int staged_base = -1;          // table base currently in LDS (per item)

        // Runs of this item: the run of v0 onwards, their records fetched 64
        // at a time (lane l holds run rbase + l) and broadcast.
        unsigned rbase = ld_const(A.run_excl, v0) + ld_const(A.run_flag, v0) - 1u;
        unsigned rv_start = 0, rv_next = 0;
        int rv_z = 0, rv_tb = 0;
        auto load_runs = [&]() {
            const unsigned r = rbase + lane;
            if (r < num_runs)
            {
                rv_start = A.run_start[r];
                rv_next = A.run_start[r + 1];
                rv_z = A.run_z[r];
                rv_tb = A.run_tb[r];
            }
        };
        load_runs();
        int rl = 0;

        for (unsigned b = v0; b < v1; b += kStageLanes * kBatch)
        {
            // Records of this wave's visits b + q + 2l, l < 64.
            const unsigned pl = b + q + kStageLanes * lane;
            const unsigned iv = pl < v1 ? A.visit_vis[pl] : 0u;
            int4 mv = make_int4(0, 0, 0, 0);
            T4 av;
            av.x = av.y = av.z = av.w = (T)0;
            if (pl < v1) { mv = A.meta[iv]; av = A.amp[iv]; }
            const unsigned bend = min(b + kStageLanes * kBatch, v1);
            for (unsigned p = b; p < bend;)
            {
                if (rl == 64)
                {
                    rbase += 64;
                    load_runs();
                    rl = 0;
                }
                const unsigned rstart = (unsigned)bcast((int)rv_start, rl);
                const unsigned rnext = (unsigned)bcast((int)rv_next, rl);
                const int rtb = bcast(rv_tb, rl);
                const int W0 = 2 * meta_support(bcast(rv_z, rl)) + 1;
                // (max: every step advances, whatever the run table holds)
                const unsigned e = max(min(rnext, bend), p + 1);
                bool staged = rtb == staged_base;
                if (!staged && W0 <= S && rnext - rstart >= A.min_stage_run)
                {
                    __syncthreads();
                    stage_table<T>(tab, A.tables + (unsigned)rtb, W0, wave, lane);
                    __syncthreads();
                    staged_base = rtb;
                    staged = true;
                }
                if (e >= rnext) ++rl;      // the run ends in this batch
                // this wave's visits in [p, e): l with p <= b + q + 2l < e
                const int lbeg = p - b <= (unsigned)q ? 0 : (int)((p - b - q + 1) / kStageLanes);
                const int lend = e - b <= (unsigned)q ? 0 : (int)((e - b - q + 1) / kStageLanes);
                auto flush = [&]() {
                    if (COMP && ++since == kFlushEvery)
                    {
                        since = 0;
                        #pragma unroll
                        for (int r = 0; r < kRowsPerWave; ++r)
                        {
                            T s0, e0, s1, e1;
                            two_sum(sre[r], re[r], s0, e0);
                            two_sum(sim[r], im[r], s1, e1);
                            sre[r] = s0; sim[r] = s1; cre[r] += e0; cim[r] += e1;
                            re[r] = im[r] = (T)0;
                        }
                    }
                };
                if (staged)
                {
                    for (int l = lbeg; l < lend; ++l)
                    {
                        int4 mt;
                        mt.x = bcast(mv.x, l); mt.y = bcast(mv.y, l);
                        mt.z = bcast(mv.z, l); mt.w = bcast(mv.w, l);
                        T4 a;
                        a.x = bcast(av.x, l); a.y = bcast(av.y, l);
                        a.z = bcast(av.z, l); a.w = bcast(av.w, l);
                        visit_lds<T>(mt, a, x0, ys, lane, tab, re, im);
                        flush();
                    }
                }
                else
                {
                    for (int l = lbeg; l < lend; ++l)
                    {
                        int4 mt;
                        mt.x = bcast(mv.x, l); mt.y = bcast(mv.y, l);
                        mt.z = bcast(mv.z, l); mt.w = bcast(mv.w, l);
                        T4 a;
                        a.x = bcast(av.x, l); a.y = bcast(av.y, l);
                        a.z = bcast(av.z, l); a.w = bcast(av.w, l);
                        visit_direct<T>(mt, a, x0, ys, lane, rs, A.zero_row_byte, re, im);
                        flush();
                    }
                }
                p = e;
            }
        }

        // Combine the two waves of each strip in a fixed order (q=0 + q=1).
        __syncthreads();
        T2* cb = tab + strip * (kRowsPerWave * 64) * (COMP ? 2 : 1);
        if (q == 1)
        {
            #pragma unroll
            for (int r = 0; r < kRowsPerWave; ++r)
            {
                if (COMP)
                {
                    T s0, e0, s1, e1;
                    two_sum(sre[r], re[r], s0, e0);
                    two_sum(sim[r], im[r], s1, e1);
                    T2 sv; sv.x = s0; sv.y = s1;
                    T2 cv; cv.x = cre[r] + e0; cv.y = cim[r] + e1;
                    cb[(2 * r) * 64 + lane] = sv;
                    cb[(2 * r + 1) * 64 + lane] = cv;
                }
                else
                {
                    T2 v; v.x = re[r]; v.y = im[r];
                    cb[r * 64 + lane] = v;
                }
            }
        }
        __syncthreads();
        if (q == 0)
        {
            const int x = x0 + lane;
            #pragma unroll
            for (int r = 0; r < kRowsPerWave; ++r)
            {
                double dre, dim;
                T pre, pim;
                if (COMP)
                {
                    const T2 sv = cb[(2 * r) * 64 + lane];
                    const T2 cv = cb[(2 * r + 1) * 64 + lane];
                    dre = (double)sre[r] + (double)cre[r] + (double)re[r] + (double)sv.x + (double)cv.x;
                    dim = (double)sim[r] + (double)cim[r] + (double)im[r] + (double)sv.y + (double)cv.y;
                    pre = (T)dre; pim = (T)dim;
                }
                else
                {
                    const T2 v = cb[r * 64 + lane];
                    pre = re[r] + v.x;
                    pim = im[r] + v.y;
                    dre = pre; dim = pim;
                }
                const int y = ys + r;
                if (nit == 1)
                {
                    if (x < G && y < G)
                    {
                        T2* g = J.grid + (size_t)y * G + x;
                        T2 gv = *g;
                        if (COMP) { gv.x = (T)((double)gv.x + dre); gv.y = (T)((double)gv.y + dim); }
                        else { gv.x += pre; gv.y += pim; }
                        *g = gv;
                    }
                }
                else
                {
                    const unsigned slot = ld_const(A.slot_off, t) + m;
                    const int c = (strip * kRowsPerWave + r) * kTileW + lane;
                    if (COMP)
                        reinterpret_cast<double2*>(A.scratch)[(size_t)slot * kTileCells + c] =
                                make_double2(dre, dim);
                    else
                    {
                        T2 v; v.x = pre; v.y = pim;
                        reinterpret_cast<T2*>(A.scratch)[(size_t)slot * kTileCells + c] = v;
                    }
                }
            }
        }
        __syncthreads();
    }
}

} // namespace gg
