// gg_kernels.hpp -- HIP kernels of the W-projection gridding pipeline (gfx950).
//
// Pipeline, in launch order:
//   k_expand_tables  kernel stack -> per-(plane, |off_u|, |off_v|) tables
//   k_norm_table     sum of Re(tap) over each footprint (FP64, serial order)
//   k_select         per-visibility selection record + tile-visit count
//   k_job_reduce     per-job norm (FP64), num_skipped, tap count, bbox
//   scan             exclusive prefix sum of visit counts
//   k_emit           one (tile|kernel-table key, visibility) pair per visit
//   radix sort       stable LSD sort of the pairs by key
//   k_tile_bounds    first/last visit of each tile in sorted order
//   k_items / scan / k_item_map   split tiles into work items
//   k_grid_tiles     one workgroup per work item, tile in registers (or LDS)
//   k_merge          ordered FP64 sum of partial tiles into the grid
//   k_degrid         adjoint: one wave per visibility, taps dealt to all lanes
#pragma once

#include "gg_device.hpp"

namespace gg {

// ===========================================================================
// Kernel-table preparation
// ===========================================================================
// The reference reads tap (j, k) of plane p at conv[p][|off_v + ov*j|]
// [|off_u + ov*k|] (oskar_grid_wproj.c:192-203): stride ov in both
// directions, so a wave reading 64 neighbouring columns would touch 64
// scattered samples.  The table for (p, a, b), a = |off_u|, b = |off_v|,
// stores that tap at row j + s, column k + s, a dense (2s+1)^2 complex
// array: a wave reads one contiguous 512-byte row segment.  Negative offsets
// use the same table mirrored (|-a + ov*k| = |a + ov*(-k)|).  Each table
// occupies (W + 2*kTablePadRows) rows of W taps, the pad rows zero; the
// table's row 0 sits kTablePadRows rows into its block.
// TI: precision of the caller's kernel stack; TA: precision the gridder
// accumulates in (the tables are stored in TA; float -> double is exact).
// One wave per table row (pad rows included): the row's table and row index
// come from one division per row, and its lanes write the row's W taps in
// order.  Workgroups stride over a plane's rows (blockIdx.y: plane), so a
// launch needs few workgroups per plane whatever the planes' sizes.
template <typename TI, typename TA>
__global__ __launch_bounds__(256)
void k_expand_tables(const typename V2<TI>::t* __restrict__ conv,
        int num_planes, int conv_size_half, int oversample, int half_ov,
        const int* __restrict__ support, const unsigned* __restrict__ plane_off,
        typename V2<TA>::t* __restrict__ tables)
{
    const size_t plane_len = (size_t)conv_size_half * conv_size_half;
    const int noff = half_ov + 1;
    const int lane = lane_id();
    const int wave = (int)(threadIdx.x >> 6), waves = (int)(blockDim.x >> 6);
    for (int p = blockIdx.y; p < num_planes; p += gridDim.y)
    {
        const int s = support[p];
        const int W = 2 * s + 1;
        const int rows = W + 2 * kTablePadRows;          // rows of one padded table
        const int total_rows = rows * noff * noff;
        const typename V2<TI>::t* src = conv + plane_len * p;
        typename V2<TA>::t* dst = tables + plane_off[p];
        for (int rr = blockIdx.x * waves + wave; rr < total_rows; rr += gridDim.x * waves)
        {
            const int tab = rr / rows;
            const int jj = rr - tab * rows - kTablePadRows;
            const int a = tab / noff, b = tab - a * noff;
            typename V2<TA>::t* out = dst + (size_t)rr * W;
            const bool in = jj >= 0 && jj < W;
            const typename V2<TI>::t* srow =
                    src + (in ? (size_t)abs(b + oversample * (jj - s)) * conv_size_half : 0);
            for (int kk = lane; kk < W; kk += 64)
            {
                typename V2<TA>::t o;
                o.x = (TA)0;
                o.y = (TA)0;
                if (in)
                {
                    const typename V2<TI>::t v = srow[abs(a + oversample * (kk - s))];
                    o.x = (TA)v.x;
                    o.y = (TA)v.y;
                }
                out[kk] = o;
            }
        }
    }
}

// norm_table[p][off_v + H][off_u + H] = sum over the footprint of Re(tap) in
// double: what one visibility with these (plane, off_u, off_v) adds to the
// norm per unit weight (the `sum` of oskar_grid_wproj.c:154, :207, :210).
// The footprint's taps for -off are those for +off in reverse order (k ->
// -k), so one wave sums the entry for (|off_u|, |off_v|) -- lanes take
// columns, rows in order, then a fixed-order wave reduction (deterministic;
// agrees with the serial order's double sum to a few ulps) -- and writes it
// to the four entries with those magnitudes.
template <typename T>
__global__ __launch_bounds__(256)
void k_norm_table(const typename V2<T>::t* __restrict__ conv,
        int num_planes, int conv_size_half, int oversample, int half_ov,
        const int* __restrict__ support, double* __restrict__ norm_table)
{
    const int side = 2 * half_ov + 1, noff = half_ov + 1;
    const int total = num_planes * noff * noff;
    const int lane = lane_id();
    const int nw = gridDim.x * (blockDim.x >> 6);
    for (int e = blockIdx.x * (blockDim.x >> 6) + (int)(threadIdx.x >> 6); e < total; e += nw)
    {
        const int p = e / (noff * noff);
        const int r = e - p * noff * noff;
        const int ov = r / noff, ou = r % noff;           // magnitudes
        const int s = support[p];
        const typename V2<T>::t* src = conv + (size_t)conv_size_half * conv_size_half * p;
        double sum = 0.0;
        for (int j = -s; j <= s; ++j)
        {
            const typename V2<T>::t* row = src + (size_t)abs(ov + oversample * j) * conv_size_half;
            for (int k = -s + lane; k <= s; k += 64)
                sum += (double)row[abs(ou + oversample * k)].x;
        }
        sum = wave_sum(sum);
        if (lane < 4)
        {
            const int sv = (lane & 2) ? -ov : ov, su = (lane & 1) ? -ou : ou;
            norm_table[(size_t)p * side * side + (size_t)(sv + half_ov) * side + (su + half_ov)] = sum;
        }
    }
}

// ===========================================================================
// Selection
// ===========================================================================
constexpr int kSelThreads = 256;
constexpr int kSelPerThread = 4;
constexpr int kSelChunk = kSelThreads * kSelPerThread;   // visibilities per chunk

// A visibility's record and weighted amplitude side by side (32 bytes), for
// the strip kernel's gather, which reads them in sorted-visit order, that is
// at random: one 32-byte read instead of two 16-byte ones.
struct __align__(32) VisRec
{
    int4 meta;
    float2 amp;                      // weight * vis (re, im)
    float2 pad;
};

template <typename TI, typename TA>
struct SelectArgs
{
    const TI* uu; const TI* vv; const TI* ww; const TI* vis; const TI* weight;
    const JobDev<TA>* jobs; int num_jobs;
    const unsigned* job_chunk_off;   // [num_jobs + 1]
    int num_planes; const int* support; int oversample; int half_ov;
    const double* norm_table; const unsigned* plane_off;
    int table_id_shift;              // table id >> shift fits kTableIdBits
    int tw, th;                      // bucketing tile width / height (cells)
    // outputs
    int4* meta; typename V4<TA>::t* amp; unsigned* num_visits;
    VisRec* vrec;                    // FP32 strip path: meta and (re, im) in one record, or null
    double* part_norm; unsigned long long* part_skipped;
    unsigned long long* part_taps; unsigned long long* part_nonfinite; int4* part_bbox;
};

__device__ __forceinline__ int find_job(const unsigned* off, int n, unsigned x)
{
    int lo = 0, hi = n - 1;                 // largest j with off[j] <= x
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Block-wide sum in a fixed order (deterministic): 256 threads.
template <typename V>
__device__ __forceinline__ V block_sum(V x, V* sh)
{
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int s = kSelThreads / 2; s > 0; s >>= 1)
    {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const V r = sh[0];
    __syncthreads();
    return r;
}
__device__ __forceinline__ int block_min(int x, int* sh)
{
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int s = kSelThreads / 2; s > 0; s >>= 1)
    {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = min(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const int r = sh[0];
    __syncthreads();
    return r;
}

// Selection in TI (bit-identical to the serial code of that precision),
// amplitudes and table addressing for accumulation in TA.
template <typename TI, typename TA>
__global__ __launch_bounds__(kSelThreads)
void k_select(const SelectArgs<TI, TA> A, unsigned num_chunks)
{
    // No FMA contraction anywhere in selection: the per-visibility norm term
    // sum * weight (oskar_grid_wproj.c:210) is rounded before it is added.
#pragma clang fp contract(off)
    __shared__ double sh_d[kSelThreads];
    __shared__ unsigned long long sh_u[kSelThreads];
    __shared__ int sh_i[kSelThreads];
    const int side = 2 * A.half_ov + 1;
    const int noff = A.half_ov + 1;
    for (unsigned c = blockIdx.x; c < num_chunks; c += gridDim.x)
    {
        const int j = find_job(A.job_chunk_off, A.num_jobs, c);
        const JobDev<TA> J = A.jobs[j];
        TI gscale, wscale;
        job_scales(J, gscale, wscale);
        const unsigned long long first = J.vis_begin +
                (unsigned long long)(c - A.job_chunk_off[j]) * kSelChunk;
        double nsum = 0.0;
        unsigned long long nskip = 0, ntaps = 0, nnonfin = 0;
        int bx0 = INT_MAX, by0 = INT_MAX, bx1 = INT_MAX, by1 = INT_MAX; // max as -min
        for (int q = 0; q < kSelPerThread; ++q)
        {
            const unsigned long long i = first + q * kSelThreads + threadIdx.x;
            if (i >= J.vis_end) break;
            const Selection s = select_vis<TI>(A.uu[i], A.vv[i], A.ww[i],
                    gscale, wscale, J.grid_size, A.oversample,
                    A.num_planes, A.support);
            int4 m;
            m.x = s.grid_u; m.y = s.grid_v;
            if (s.skip)
            {
                m.z = kFlagSkip; m.w = 0;
                A.meta[i] = m;
                A.num_visits[i] = 0;
                ++nskip;
                continue;
            }
            const int a = abs(s.off_u), b = abs(s.off_v);
            const int W = 2 * s.support + 1;
            const unsigned tab = (unsigned)(a * noff + b);
            const unsigned table_id = (unsigned)s.plane * noff * noff + tab;
            int flags = (s.off_v < 0 ? kFlagNegV : 0) |
                    (s.off_u < 0 ? kFlagNegU : 0) | (s.conj ? kFlagConj : 0);
            const TI w_in = A.weight[i];
            TA amp_re = (TA)0, amp_im = (TA)0;
            if (A.vis)
            {
                // Amplitude: weight * vis (oskar_grid_wproj.c:166-169) in the
                // accumulation precision, the conjugation folded into the FMA
                // coefficients.
                const TA w_i = (TA)w_in;
                const TA v_re = w_i * (TA)A.vis[2 * i];
                const TA v_im = w_i * (TA)A.vis[2 * i + 1];
                typename V4<TA>::t am;
                am.x = v_re;
                am.y = s.conj ? v_im : -v_im;
                am.z = v_im;
                am.w = s.conj ? -v_re : v_re;
                amp_re = v_re;
                amp_im = v_im;
                const bool nonfinite = !isfinite(v_re) || !isfinite(v_im);
                // (the strip path reads amplitudes from vrec; only the
                // non-finite pass reads amp there)
                if (!A.vrec || nonfinite) A.amp[i] = am;
                if (nonfinite)
                {
                    flags |= kFlagNonFinite;
                    ++nnonfin;
                }
            }
            m.z = (int)(((table_id >> A.table_id_shift) << kTableIdShift) |
                    ((unsigned)s.support << kSupportShift) | (unsigned)flags);
            m.w = (int)(A.plane_off[s.plane] + tab * (unsigned)(W + 2 * kTablePadRows) * (unsigned)W +
                    (unsigned)(kTablePadRows * W));
            A.meta[i] = m;
            if constexpr (std::is_same<TA, float>::value)
            {
                if (A.vrec)
                {
                    VisRec rr;
                    rr.meta = m;
                    rr.amp = make_float2(amp_re, amp_im);
                    rr.pad = make_float2(0.f, 0.f);
                    A.vrec[i] = rr;
                }
            }
            // Tile visits (none for non-finite amplitudes: k_grid_nonfinite
            // grids those after the tiles are done).
            // (footprints of gridded visibilities lie inside the grid: the
            // operands are non-negative)
            const int tx0 = (s.grid_u - s.support) / A.tw;
            const int tx1 = (s.grid_u + s.support) / A.tw;
            const int ty0 = (s.grid_v - s.support) / A.th;
            const int ty1 = (s.grid_v + s.support) / A.th;
            A.num_visits[i] = (flags & kFlagNonFinite) ? 0u
                    : (unsigned)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
            // Norm term (oskar_grid_wproj.c:210), taps, bounding box.
            nsum += A.norm_table[(s.plane * side + s.off_v + A.half_ov) * side +
                    s.off_u + A.half_ov] * (double)w_in;
            ntaps += (unsigned long long)W * W;
            bx0 = min(bx0, s.grid_u - s.support);
            by0 = min(by0, s.grid_v - s.support);
            bx1 = min(bx1, -(s.grid_u + s.support));
            by1 = min(by1, -(s.grid_v + s.support));
        }
        nsum = block_sum(nsum, sh_d);
        nskip = block_sum(nskip, sh_u);
        ntaps = block_sum(ntaps, sh_u);
        nnonfin = block_sum(nnonfin, sh_u);
        bx0 = block_min(bx0, sh_i);
        by0 = block_min(by0, sh_i);
        bx1 = block_min(bx1, sh_i);
        by1 = block_min(by1, sh_i);
        if (threadIdx.x == 0)
        {
            A.part_norm[c] = nsum;
            A.part_skipped[c] = nskip;
            A.part_taps[c] = ntaps;
            A.part_nonfinite[c] = nnonfin;
            A.part_bbox[c] = make_int4(bx0, by0, bx1, by1);
        }
    }
}

// Selection only, for callers that need the per-visibility cells and kernel
// indices: out[f*n + i] for the fields grid_u, grid_v, off_u, off_v, plane,
// support, conj, skip (f = 0..7).  One job.
template <typename TI>
__global__ __launch_bounds__(256)
void k_select_out(const TI* __restrict__ uu, const TI* __restrict__ vv,
        const TI* __restrict__ ww, unsigned long long n, TI grid_scale, TI w_scale,
        int grid_size, int oversample, int num_planes, const int* __restrict__ support,
        int* __restrict__ out)
{
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            i < n; i += (unsigned long long)gridDim.x * blockDim.x)
    {
        const Selection s = select_vis<TI>(uu[i], vv[i], ww[i], grid_scale, w_scale,
                grid_size, oversample, num_planes, support);
        out[i] = s.grid_u;
        out[n + i] = s.grid_v;
        out[2 * n + i] = s.off_u;
        out[3 * n + i] = s.off_v;
        out[4 * n + i] = s.plane;
        out[5 * n + i] = s.support;
        out[6 * n + i] = s.conj;
        out[7 * n + i] = s.skip;
    }
}

// Per-job totals over the job's chunks, in chunk order.  job_out receives
// {skipped, taps, non-finite} and job_bbox the {x0, y0, x1, y1} cell box
// (x1 < x0 if nothing is gridded), job_norm the norm term sum.  The
// caller's outputs are written by k_job_commit at the end of the pass.
template <typename T>
__global__ __launch_bounds__(kSelThreads)
void k_job_reduce(const JobDev<T>* __restrict__ jobs,
        const unsigned* __restrict__ job_chunk_off,
        const double* __restrict__ part_norm,
        const unsigned long long* __restrict__ part_skipped,
        const unsigned long long* __restrict__ part_taps,
        const unsigned long long* __restrict__ part_nonfinite,
        const int4* __restrict__ part_bbox,
        unsigned long long* __restrict__ job_out, int4* __restrict__ job_bbox,
        double* __restrict__ job_norm)
{
    __shared__ double sh_d[kSelThreads];
    __shared__ unsigned long long sh_u[kSelThreads];
    __shared__ int sh_i[kSelThreads];
    const int j = blockIdx.x;
    const unsigned c0 = job_chunk_off[j], c1 = job_chunk_off[j + 1];
    double nsum = 0.0;
    unsigned long long nskip = 0, ntaps = 0, nnf = 0;
    int b0 = INT_MAX, b1 = INT_MAX, b2 = INT_MAX, b3 = INT_MAX;
    for (unsigned c = c0 + threadIdx.x; c < c1; c += kSelThreads)
    {
        nsum += part_norm[c];
        nskip += part_skipped[c];
        ntaps += part_taps[c];
        nnf += part_nonfinite[c];
        const int4 b = part_bbox[c];
        b0 = min(b0, b.x); b1 = min(b1, b.y); b2 = min(b2, b.z); b3 = min(b3, b.w);
    }
    nsum = block_sum(nsum, sh_d);
    nskip = block_sum(nskip, sh_u);
    ntaps = block_sum(ntaps, sh_u);
    nnf = block_sum(nnf, sh_u);
    b0 = block_min(b0, sh_i); b1 = block_min(b1, sh_i);
    b2 = block_min(b2, sh_i); b3 = block_min(b3, sh_i);
    if (threadIdx.x == 0)
    {
        job_norm[j] = nsum;
        job_out[3 * j] = nskip;
        job_out[3 * j + 1] = ntaps;
        job_out[3 * j + 2] = nnf;
        job_bbox[j] = make_int4(b0, b1, b2 == INT_MAX ? INT_MIN : -b2,
                b3 == INT_MAX ? INT_MIN : -b3);
    }
}

// The pass's outputs besides the grids: *norm += the job's norm sum (the
// `*norm += sum * weight` of oskar_grid_wproj.c:210, summed per job),
// *num_skipped = its count (set per call, :151).  Launched after the last
// kernel that can fail to be launched, so that a call refused on the way
// (workspace) leaves the caller's outputs as they were.
template <typename T>
__global__ void k_job_commit(const JobDev<T>* __restrict__ jobs, int num_jobs,
        const double* __restrict__ job_norm, const unsigned long long* __restrict__ job_out)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < num_jobs; j += gridDim.x * blockDim.x)
    {
        *jobs[j].norm += job_norm[j];
        *jobs[j].skipped = job_out[3 * j];
    }
}

// ===========================================================================
// Exclusive scan of unsigned counts (three passes, deterministic)
// ===========================================================================
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;
constexpr int kScanChunk = kScanThreads * kScanPerThread;

// Count of an array given on the device: n = mul * ceil(*src / div) + add.
struct DevCount
{
    const unsigned long long* src;
    unsigned long long div, mul, add;
    __device__ __forceinline__ unsigned long long get() const
    {
        return src ? mul * ((*src + div - 1) / div) + add : add;
    }
};

__device__ __forceinline__ unsigned long long block_excl_scan_u64(
        unsigned long long x, unsigned long long* sh, unsigned long long* total)
{
    // inclusive wave scan
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    unsigned long long v = x;
    #pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const unsigned long long y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) sh[wave] = v;
    __syncthreads();
    unsigned long long base = 0, tot = 0;
    for (int w = 0; w < kScanThreads / 64; ++w)
    {
        if (w < wave) base += sh[w];
        tot += sh[w];
    }
    __syncthreads();
    *total = tot;
    return base + v - x;
}

__global__ __launch_bounds__(kScanThreads)
void k_scan_reduce(const unsigned* __restrict__ in, DevCount cnt,
        unsigned long long* __restrict__ partial)
{
    __shared__ unsigned long long sh[kScanThreads / 64];
    const unsigned long long n = cnt.get();
    const unsigned long long nchunk = (n + kScanChunk - 1) / kScanChunk;
    for (unsigned long long c = blockIdx.x; c < nchunk; c += gridDim.x)
    {
        const unsigned long long base = c * kScanChunk + (unsigned long long)threadIdx.x * kScanPerThread;
        unsigned long long s = 0;
        #pragma unroll
        for (int q = 0; q < kScanPerThread; ++q)
            if (base + q < n) s += in[base + q];
        unsigned long long tot;
        block_excl_scan_u64(s, sh, &tot);
        if (threadIdx.x == 0) partial[c] = tot;
    }
}

// Single workgroup: exclusive scan of the chunk totals; *total = grand total,
// *overflow set when it exceeds 32 bits; out[n] = total when write_end.
__global__ __launch_bounds__(1024)
void k_scan_spine(DevCount cnt, unsigned long long* __restrict__ partial,
        unsigned long long* __restrict__ total, unsigned* __restrict__ overflow,
        unsigned* __restrict__ out, int write_end)
{
    __shared__ unsigned long long sh[16];
    const unsigned long long n = cnt.get();
    const unsigned long long nchunk = (n + kScanChunk - 1) / kScanChunk;
    const unsigned long long per = (nchunk + 1023) / 1024;
    const unsigned long long b0 = threadIdx.x * per;
    unsigned long long s = 0;
    for (unsigned long long c = b0; c < b0 + per && c < nchunk; ++c) s += partial[c];
    // block exclusive scan over 1024 threads
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    unsigned long long v = s;
    #pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const unsigned long long y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) sh[wave] = v;
    __syncthreads();
    unsigned long long base = 0, tot = 0;
    for (int w = 0; w < 16; ++w) { if (w < wave) base += sh[w]; tot += sh[w]; }
    unsigned long long run = base + v - s;
    for (unsigned long long c = b0; c < b0 + per && c < nchunk; ++c)
    {
        const unsigned long long x = partial[c];
        partial[c] = run;
        run += x;
    }
    if (threadIdx.x == 0)
    {
        *total = tot;
        if (tot > 0xFFFFFFFFull) *overflow = 1u;
        if (write_end) out[n] = (unsigned)tot;
    }
}

__global__ __launch_bounds__(kScanThreads)
void k_scan_down(const unsigned* __restrict__ in, DevCount cnt,
        const unsigned long long* __restrict__ partial, unsigned* __restrict__ out)
{
    __shared__ unsigned long long sh[kScanThreads / 64];
    const unsigned long long n = cnt.get();
    const unsigned long long nchunk = (n + kScanChunk - 1) / kScanChunk;
    for (unsigned long long c = blockIdx.x; c < nchunk; c += gridDim.x)
    {
        const unsigned long long base = c * kScanChunk + (unsigned long long)threadIdx.x * kScanPerThread;
        unsigned vals[kScanPerThread];
        unsigned long long s = 0;
        #pragma unroll
        for (int q = 0; q < kScanPerThread; ++q)
        {
            vals[q] = base + q < n ? in[base + q] : 0u;
            s += vals[q];
        }
        unsigned long long tot;
        unsigned long long run = partial[c] + block_excl_scan_u64(s, sh, &tot);
        #pragma unroll
        for (int q = 0; q < kScanPerThread; ++q)
        {
            if (base + q < n) out[base + q] = (unsigned)run;
            run += vals[q];
        }
    }
}

// ===========================================================================
// Visit emission and radix sort
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256)
void k_emit(const int4* __restrict__ meta, const unsigned* __restrict__ num_visits,
        const unsigned* __restrict__ visit_off, const JobDev<T>* __restrict__ jobs,
        const unsigned long long* __restrict__ job_vis_off, int num_jobs,
        unsigned long long n, int tw, int th, int key_table_shift, int key_table_bits,
        int orient_bits, unsigned* __restrict__ keys, unsigned* __restrict__ vals)
{
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            i < n; i += (unsigned long long)gridDim.x * blockDim.x)
    {
        const unsigned nv = num_visits[i];
        if (!nv) continue;
        int lo = 0, hi = num_jobs - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (job_vis_off[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const int ntx = jobs[lo].ntx;
        const unsigned tile_base = jobs[lo].tile_base;
        const int4 m = meta[i];
        const int s = meta_support(m.z);
        // kernel table, then (orient_bits = 2) its orientation, the two mirror
        // flags; a kernel that mirrors when reading takes runs of all
        // orientations (orient_bits = 0)
        const unsigned kappa = ((meta_table_id(m.z) << orient_bits) |
                (orient_bits ? (unsigned)(m.z & kOrientMask) : 0u)) >> key_table_shift;
        const int tx0 = (m.x - s) / tw, tx1 = (m.x + s) / tw;
        const int ty0 = (m.y - s) / th, ty1 = (m.y + s) / th;
        unsigned o = visit_off[i];
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx, ++o)
            {
                const unsigned tile = tile_base + (unsigned)(ty * ntx + tx);
                keys[o] = (tile << key_table_bits) | kappa;
                vals[o] = (unsigned)i;
            }
    }
}

constexpr int kSortThreads = 256;
constexpr int kSortPerThread = 16;
constexpr int kSortChunk = kSortThreads * kSortPerThread;

// Histogram of one BITS-bit digit (8 or 9) per chunk, stored digit-major:
// hist[d * nchunk + c].
template <int BITS>
__global__ __launch_bounds__(kSortThreads)
void k_sort_hist(const unsigned* __restrict__ keys,
        const unsigned long long* __restrict__ d_n, int shift,
        unsigned* __restrict__ hist)
{
    constexpr int R = 1 << BITS;
    constexpr unsigned M = R - 1;
    __shared__ unsigned h[R];
    const unsigned long long n = *d_n;
    const unsigned long long nchunk = (n + kSortChunk - 1) / kSortChunk;
    for (unsigned long long c = blockIdx.x; c < nchunk; c += gridDim.x)
    {
        for (int d = threadIdx.x; d < R; d += kSortThreads) h[d] = 0;
        __syncthreads();
        const unsigned long long base = c * kSortChunk;
        #pragma unroll 4
        for (int r = 0; r < kSortPerThread; ++r)
        {
            const unsigned long long idx = base + r * kSortThreads + threadIdx.x;
            if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & M], 1u);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < R; d += kSortThreads) hist[d * nchunk + c] = h[d];
        __syncthreads();
    }
}

// Stable scatter.  Wave w of a chunk takes its 1024 consecutive elements
// (w * 1024 ..), 64 at a time: lanes whose keys share a digit are matched
// with BITS ballots, and each element's rank among the wave's earlier
// elements of its digit comes from a running per-wave count in LDS (the
// lowest lane of each digit group advances it).  One prefix over the waves
// per digit then gives each wave's base, so that ranks follow index order
// and the chunk's base for each digit comes from the digit-major scan.  The
// keys, values and ranks stay in registers meanwhile: two barriers per
// chunk.
template <int BITS>
__global__ __launch_bounds__(kSortThreads)
void k_sort_scatter(const unsigned* __restrict__ keys_in,
        const unsigned* __restrict__ vals_in,
        const unsigned long long* __restrict__ d_n, int shift,
        const unsigned* __restrict__ hist_off,
        unsigned* __restrict__ keys_out, unsigned* __restrict__ vals_out)
{
    constexpr int R = 1 << BITS;
    constexpr unsigned M = R - 1;
    constexpr int NW = kSortThreads / 64;
    constexpr int SUB = kSortChunk / NW / 64;          // 64-element rows per wave
    __shared__ unsigned goff[R];
    __shared__ unsigned wcnt[NW][R];
    const unsigned long long n = *d_n;
    const unsigned long long nchunk = (n + kSortChunk - 1) / kSortChunk;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = lanemask_lt();
    for (unsigned long long c = blockIdx.x; c < nchunk; c += gridDim.x)
    {
        for (int d = threadIdx.x; d < R; d += kSortThreads)
        {
            goff[d] = hist_off[d * nchunk + c];
            #pragma unroll
            for (int w = 0; w < NW; ++w) wcnt[w][d] = 0;
        }
        __syncthreads();
        const unsigned long long base = c * kSortChunk + (unsigned long long)wave * (SUB * 64);
        unsigned key[SUB], val[SUB], rank[SUB];
        #pragma unroll
        for (int r = 0; r < SUB; ++r)
        {
            const unsigned long long idx = base + r * 64 + lane;
            const bool valid = idx < n;
            key[r] = valid ? keys_in[idx] : 0u;
            val[r] = valid ? vals_in[idx] : 0u;
        }
        #pragma unroll
        for (int r = 0; r < SUB; ++r)
        {
            const bool valid = base + r * 64 + lane < n;
            const unsigned d = (key[r] >> shift) & M;
            unsigned long long peers = __ballot(valid);
            #pragma unroll
            for (int b = 0; b < BITS; ++b)
            {
                const bool bit = (d >> b) & 1u;
                const unsigned long long bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            const unsigned before = (unsigned)__popcll(peers & lt);
            const unsigned prior = valid ? wcnt[wave][d] : 0u;
            rank[r] = prior + before;
            if (valid && before == 0) wcnt[wave][d] = prior + (unsigned)__popcll(peers);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < R; d += kSortThreads)
        {
            unsigned acc = goff[d];
            #pragma unroll
            for (int w = 0; w < NW; ++w)
            {
                const unsigned t = wcnt[w][d];
                wcnt[w][d] = acc;
                acc += t;
            }
        }
        __syncthreads();
        #pragma unroll
        for (int r = 0; r < SUB; ++r)
        {
            if (base + r * 64 + lane < n)
            {
                const unsigned dst = wcnt[wave][(key[r] >> shift) & M] + rank[r];
                keys_out[dst] = key[r];
                vals_out[dst] = val[r];
            }
        }
        __syncthreads();
    }
}

// ===========================================================================
// Tiles and work items
// ===========================================================================
__global__ void k_fill_u32(unsigned* __restrict__ p, unsigned long long n, unsigned v)
{
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            i < n; i += (unsigned long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_tile_bounds(const unsigned* __restrict__ keys,
        const unsigned long long* __restrict__ d_n, int key_table_bits,
        unsigned* __restrict__ tile_start, unsigned* __restrict__ tile_end)
{
    const unsigned long long n = *d_n;
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            p < n; p += (unsigned long long)gridDim.x * blockDim.x)
    {
        const unsigned t = keys[p] >> key_table_bits;
        if (p == 0 || (keys[p - 1] >> key_table_bits) != t) tile_start[t] = (unsigned)p;
        if (p == n - 1 || (keys[p + 1] >> key_table_bits) != t) tile_end[t] = (unsigned)(p + 1);
    }
}

// Runs: maximal stretches of sorted visits with the same tile, the same
// kernel table and the same mirror flags of orient_mask.
// flag[p] = 1 where a run starts.  When the sort key holds the full table id
// (and orientation), key equality decides; otherwise the visit records are
// compared.
__global__ void k_run_flags(const unsigned* __restrict__ keys,
        const unsigned* __restrict__ vals, const int4* __restrict__ meta,
        const unsigned long long* __restrict__ d_n, int exact_key, int orient_mask,
        unsigned* __restrict__ flag)
{
    const unsigned long long n = *d_n;
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            p < n; p += (unsigned long long)gridDim.x * blockDim.x)
    {
        unsigned f = p == 0 || keys[p] != keys[p - 1];
        if (!f && !exact_key)
        {
            const int4 a = meta[vals[p]], b = meta[vals[p - 1]];
            f = a.w != b.w || ((a.z ^ b.z) & orient_mask);
        }
        flag[p] = f;
    }
}

// Per run r (excl = exclusive scan of flag): run_start[r] = its first visit,
// run_z[r] / run_tb[r] = packed flags-support-table word and table base of
// its visits (unless run_z is null: the strip kernel reads the gathered
// records instead); run_start[R] = n.  The run of visit p is excl[p] +
// flag[p] - 1.
__global__ void k_run_starts(const unsigned* __restrict__ flag,
        const unsigned* __restrict__ excl, const unsigned* __restrict__ vals,
        const int4* __restrict__ meta, const unsigned long long* __restrict__ d_n,
        const unsigned long long* __restrict__ d_runs, unsigned* __restrict__ run_start,
        int* __restrict__ run_z, int* __restrict__ run_tb)
{
    const unsigned long long n = *d_n;
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            p < n; p += (unsigned long long)gridDim.x * blockDim.x)
    {
        if (flag[p])
        {
            const unsigned r = excl[p];
            run_start[r] = (unsigned)p;
            if (run_z)
            {
                const int4 m = meta[vals[p]];
                run_z[r] = m.z;
                run_tb[r] = m.w;
            }
        }
        if (p == 0) run_start[*d_runs] = (unsigned)n;
    }
}

// Records of the sorted visits, in visit order (one load per batch of
// visits in the strip kernel instead of an index load and a dependent one),
// with kVisitStaged set where the visit's run is staged in LDS (a table at
// most max_w wide, and at least the threshold of its support class of
// visits: min_run.x .. .w for tables at most 33, 65, 97 and max_w wide) and
// kVisitRunStart where
// the visit starts a run; y holds the visit's placement in its region (see
// below), the table-id bits of z the table's symmetry (kVisitFoldShift), and
// vlds the visit's part of its staged LDS offsets.
template <typename T>
__global__ void k_gather_visits(const unsigned* __restrict__ vals,
        const VisRec* __restrict__ vrec,
        const unsigned* __restrict__ run_flag, const unsigned* __restrict__ run_excl,
        const unsigned* __restrict__ run_start, uint4 min_run, int max_w, int lds_stride,
        const unsigned* __restrict__ keys, int key_table_bits,
        const JobDev<T>* __restrict__ jobs, const unsigned* __restrict__ job_tile_base,
        int num_jobs, int region_h,
        int fold_noff, int fold_half,
        const unsigned long long* __restrict__ d_n, int4* __restrict__ vmeta,
        typename V2<T>::t* __restrict__ vamp, int* __restrict__ vlds)
{
    const unsigned long long n = *d_n;
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            p < n; p += (unsigned long long)gridDim.x * blockDim.x)
    {
        const unsigned i = vals[p];
        const VisRec rec = vrec[i];
        int4 m = rec.meta;
        const unsigned r = run_excl[p] + run_flag[p] - 1u;
        const int W = 2 * meta_support(m.z) + 1;
        const unsigned mr = W <= 33 ? min_run.x : W <= 65 ? min_run.y : W <= 97 ? min_run.z
                                                                           : min_run.w;
        if (run_start[r + 1] - run_start[r] >= mr && W <= max_w) m.z |= kVisitStaged;
        if (run_flag[p]) m.z |= kVisitRunStart;
        vamp[p] = {rec.amp.x, rec.amp.y};  // the weighted amplitude (re, im)
        // The visit's part of its staged-tap LDS byte offsets (the strip
        // kernel adds the lane's part): footprint column 0 is table column
        // s - x + (strip column), or x + s - (strip column) mirrored in u;
        // footprint row 0 is staged row -(y - s) + (strip row), or the rows
        // run up from y + s when mirrored in v.
        const int sp = meta_support(m.z);
        const int cu = (m.z & kFlagNegU) ? m.x + sp : sp - m.x;
        const int rv = (m.z & kFlagNegV) ? m.y + sp : sp - m.y;
        vlds[p] = 8 * (cu + rv * lds_stride);
        // The visit's rows in its region (sorted key -> tile -> job -> row
        // of the region): footprint row 0 at region row top, and the mask
        // of the 8-row chunks the footprint touches, packed into y as
        // (chunk mask << 16) | (top & 0xffff).
        const unsigned t = keys[p] >> key_table_bits;
        const int j = find_job(job_tile_base, num_jobs, t);
        const unsigned local = t - jobs[j].tile_base;
        const int ys = (int)(local / (unsigned)jobs[j].ntx) * region_h;
        const int top = m.y - sp - ys;
        const int q0 = max(top, 0) >> 3;
        const int q1 = min(top + 2 * sp, region_h - 1) >> 3;
        const unsigned cm = (2u << q1) - (1u << q0);
        m.y = (int)((cm << 16) | ((unsigned)top & 0xffffu));
        // The table's symmetry from its id (table = (plane noff + |off_u|)
        // noff + |off_v|; fold_noff = 0 when the ids were shortened).
        int fold = 0;
        if (fold_noff > 0)
        {
            const unsigned tab = meta_table_id(m.z) % (unsigned)(fold_noff * fold_noff);
            const int a = (int)tab / fold_noff, b = (int)tab % fold_noff;
            fold = (a == 0 ? 1 : a == fold_half ? 2 : 0) | ((b == 0 ? 1 : b == fold_half ? 2 : 0) << 2);
        }
        m.z = (m.z & ((1 << kTableIdShift) - 1)) | (fold << kVisitFoldShift);
        vmeta[p] = m;
    }
}

// Work items per tile: ceil(visits / visits_per_item); tiles with two or more
// items also take that many partial-tile slots.
__global__ void k_items(const unsigned* __restrict__ tile_start,
        const unsigned* __restrict__ tile_end, unsigned num_tiles,
        unsigned visits_per_item, unsigned* __restrict__ n_items,
        unsigned* __restrict__ n_slots)
{
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < num_tiles;
            t += gridDim.x * blockDim.x)
    {
        const unsigned cnt = tile_end[t] - tile_start[t];
        const unsigned ni = (cnt + visits_per_item - 1) / visits_per_item;
        n_items[t] = ni;
        n_slots[t] = ni >= 2 ? ni : 0u;
    }
}

__global__ void k_item_map(const unsigned* __restrict__ item_off, unsigned num_tiles,
        unsigned* __restrict__ item_tile)
{
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < num_tiles;
            t += gridDim.x * blockDim.x)
        for (unsigned it = item_off[t]; it < item_off[t + 1]; ++it) item_tile[it] = t;
}

// ===========================================================================
// Gridding
// ===========================================================================
template <typename T>
struct GridArgs
{
    const typename V2<T>::t* tables;     // followed by 2*kMaxSupport+1 zero taps
    unsigned table_bytes;                // tables + zero row, < 2^32
    unsigned zero_row_byte;              // byte offset of the zero row
    const int4* meta;
    const typename V4<T>::t* amp;
    const unsigned* visit_vis;     // visibility index per sorted visit
    const unsigned* tile_start; const unsigned* tile_end;
    const unsigned* item_off;      // [num_tiles + 1]
    const unsigned* slot_off;      // [num_tiles]
    const unsigned* item_tile;
    const unsigned* d_num_items;   // = item_off[num_tiles]
    const JobDev<T>* jobs; const unsigned* job_tile_base; int num_jobs;
    unsigned visits_per_item;
    void* scratch;                 // partial tiles of multi-item tiles
    // runs (staged kernel): see k_run_starts
    const unsigned* run_flag; const unsigned* run_excl;
    const unsigned* run_start; const int* run_z; const int* run_tb;
    const unsigned long long* d_num_runs;
    unsigned min_stage_run;        // shortest run staged through LDS
    const int4* vmeta;             // meta / amp of each sorted visit (strip kernel)
    const typename V2<T>::t* vamp;        // weighted amplitude per sorted visit
    const int* vlds;               // per sorted visit: its part of the staged LDS offsets
    unsigned item_group;           // strip kernel: consecutive items per XCD (>= 1)
    unsigned* item_ctr;            // strip kernel: 8 zeroed counters (positions
                                   // taken per XCD)
};

// TwoSum (Knuth): s + e == a + b exactly.
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e)
{
#pragma clang fp contract(off)
    s = a + b;
    const T bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

template <typename T>
__device__ __forceinline__ bool all_finite(const typename V4<T>::t& a)
{
    return isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(a.w);
}

// Raw buffer loads of one complex tap: scalar resource and row offset,
// per-lane column offset -> one buffer_load per row, no address VALU work.
// (0x00020000: the gfx9-family resource word 3 for a raw dword buffer.)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const void* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 load_tap(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff, float2*)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ double2 load_tap(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff, double2*)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, (int)soff, 0));
}

// One workgroup per work item: tile (tx, ty) of job j, visits [v0, v1) of
// that tile's sorted list.  Wave w owns rows y0+16w .. y0+16w+15, lane x
// owns column x0+x; the 16 complex accumulators per lane live in registers.
// For each visit the wave reads the kernel-table row segment under its
// lanes (one coalesced 512-byte load per row) and does 4 FMAs per tap.  The
// visit record is read with scalar loads one visit ahead, so the
// coefficients sit in scalar registers and all row addressing is scalar.
// Rows are taken four at a time; rows of a group that fall outside the
// footprint read a zero row instead (multiplying finite coefficients by
// zero leaves the accumulator unchanged), so the four loads issue together.
// No atomics: the tile is owned.  COMP: every kFlushEvery visits the
// register partials are added into a TwoSum-compensated (sum, error) pair
// per cell in LDS.
constexpr int kFlushEvery = 32;

template <typename T, bool COMP>
__global__ __launch_bounds__(kBlock)
void k_grid_tiles(const GridArgs<T> A, unsigned grid_padded)
{
    typedef typename V2<T>::t T2;
    typedef typename V4<T>::t T4;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const unsigned num_items = ld_const(A.d_num_items, 0);
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const __amdgpu_buffer_rsrc_t rs = table_rsrc(A.tables, A.table_bytes);
    for (unsigned b = blockIdx.x; b < grid_padded; b += gridDim.x)
    {
        const unsigned item = xcd_remap(b, grid_padded);
        if (item >= num_items) continue;
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
        const int ys = ty * kTileH + wave * kRowsPerWave;

        T acc_re[kRowsPerWave], acc_im[kRowsPerWave];
        #pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) { acc_re[r] = (T)0; acc_im[r] = (T)0; }

        T* lds = reinterpret_cast<T*>(lds_raw) + (size_t)wave * kRowsPerWave * 64 * 4;
        if (COMP)
        {
            #pragma unroll
            for (int r = 0; r < kRowsPerWave; ++r)
            {
                T* c = lds + (r * 64 + lane) * 4;
                c[0] = (T)0; c[1] = (T)0; c[2] = (T)0; c[3] = (T)0;
            }
        }
        int since_flush = 0;

        unsigned i_next = ld_const(A.visit_vis, v0);
        int4 m_next = ld_const(A.meta, i_next);
        T4 a_next = ld_const(A.amp, i_next);
        for (unsigned p = v0; p < v1; ++p)
        {
            const int4 mt = m_next;
            const T4 a = a_next;
            if (p + 1 < v1)
            {
                i_next = ld_const(A.visit_vis, p + 1);
                m_next = ld_const(A.meta, i_next);
                a_next = ld_const(A.amp, i_next);
            }
            const int s = meta_support(mt.z);
            const int j0 = ys - mt.y;                  // row offset of r = 0
            if (j0 > s || j0 + kRowsPerWave - 1 < -s) continue;
            const int W = 2 * s + 1;
            const int k = x0 + lane - mt.x;
            const bool neg_u = mt.z & kFlagNegU, neg_v = mt.z & kFlagNegV;
            if (k >= -s && k <= s)
            {
                const unsigned kcol = (unsigned)(neg_u ? s - k : s + k);
                const unsigned kbyte = kcol * (unsigned)sizeof(T2);
                // (amplitudes are finite: non-finite ones never reach a tile)
                #pragma unroll
                for (int g = 0; g < kRowsPerWave; g += 4)
                {
                    const int jg = j0 + g;
                    if (jg > s || jg + 3 < -s) continue;
                    T2 c[4];
                    #pragma unroll
                    for (int r = 0; r < 4; ++r)
                    {
                        const int jr = jg + r;
                        const unsigned soff = (jr >= -s && jr <= s)
                                ? ((unsigned)mt.w + (unsigned)((neg_v ? s - jr : s + jr) * W)) *
                                  (unsigned)sizeof(T2)
                                : A.zero_row_byte;
                        c[r] = load_tap(rs, kbyte, soff, (T2*)nullptr);
                    }
                    #pragma unroll
                    for (int r = 0; r < 4; ++r)
                    {
                        acc_re[g + r] = fma(c[r].x, a.x, acc_re[g + r]);
                        acc_re[g + r] = fma(c[r].y, a.y, acc_re[g + r]);
                        acc_im[g + r] = fma(c[r].x, a.z, acc_im[g + r]);
                        acc_im[g + r] = fma(c[r].y, a.w, acc_im[g + r]);
                    }
                }
            }
            if (COMP && ++since_flush == kFlushEvery)
            {
                since_flush = 0;
                #pragma unroll
                for (int r = 0; r < kRowsPerWave; ++r)
                {
                    T* c = lds + (r * 64 + lane) * 4;
                    T s0, e0, s1, e1;
                    two_sum(c[0], acc_re[r], s0, e0);
                    two_sum(c[1], acc_im[r], s1, e1);
                    c[0] = s0; c[1] = s1; c[2] += e0; c[3] += e1;
                    acc_re[r] = (T)0; acc_im[r] = (T)0;
                }
            }
        }

        // Write-back.  A tile with one item adds straight into the grid; a
        // tile split over several items writes its partial to a slot that
        // k_merge sums in item order.
        const int x = x0 + lane;
        if (nit == 1)
        {
            #pragma unroll
            for (int r = 0; r < kRowsPerWave; ++r)
            {
                const int y = ys + r;
                if (x < G && y < G)
                {
                    T2* g = J.grid + (size_t)y * G + x;
                    T2 v = *g;
                    if (COMP)
                    {
                        const T* c = lds + (r * 64 + lane) * 4;
                        const double re = (double)v.x + ((double)c[0] + (double)c[2]) + (double)acc_re[r];
                        const double im = (double)v.y + ((double)c[1] + (double)c[3]) + (double)acc_im[r];
                        v.x = (T)re; v.y = (T)im;
                    }
                    else
                    {
                        v.x += acc_re[r]; v.y += acc_im[r];
                    }
                    *g = v;
                }
            }
        }
        else
        {
            const unsigned slot = A.slot_off[t] + m;
            if (COMP)
            {
                double2* out = reinterpret_cast<double2*>(A.scratch) + (size_t)slot * kTileCells;
                #pragma unroll
                for (int r = 0; r < kRowsPerWave; ++r)
                {
                    const T* c = lds + (r * 64 + lane) * 4;
                    out[(wave * kRowsPerWave + r) * kTileW + lane] = make_double2(
                            ((double)c[0] + (double)c[2]) + (double)acc_re[r],
                            ((double)c[1] + (double)c[3]) + (double)acc_im[r]);
                }
            }
            else
            {
                T2* out = reinterpret_cast<T2*>(A.scratch) + (size_t)slot * kTileCells;
                #pragma unroll
                for (int r = 0; r < kRowsPerWave; ++r)
                {
                    T2 v; v.x = acc_re[r]; v.y = acc_im[r];
                    out[(wave * kRowsPerWave + r) * kTileW + lane] = v;
                }
            }
        }
        if (COMP) __syncthreads();
    }
}

// Ordered merge: for each tile split over n >= 2 items, grid += sum of the
// item partials, summed in item order in double.  Block (c, t): tile t's
// cells 1024c .. 1024c+1023, four per thread; the partial loads of up to
// four items are in flight together, the sums stay in item order.
constexpr int kMergeCells = 1024;

template <typename T, typename P2>
__global__ __launch_bounds__(256)
void k_merge(const unsigned* __restrict__ item_off, const unsigned* __restrict__ slot_off,
        unsigned num_tiles, const JobDev<T>* __restrict__ jobs,
        const unsigned* __restrict__ job_tile_base, int num_jobs, int tw, int th,
        const P2* __restrict__ scratch)
{
    typedef typename V2<T>::t T2;
    const int cells = tw * th;   // a multiple of kMergeCells
    const unsigned chunks = (unsigned)(cells / kMergeCells);
    for (unsigned long long w = blockIdx.x; w < (unsigned long long)num_tiles * chunks; w += gridDim.x)
    {
        const unsigned t = (unsigned)(w / chunks);
        const unsigned ch = (unsigned)(w - (unsigned long long)t * chunks);
        const unsigned nit = item_off[t + 1] - item_off[t];
        if (nit < 2) continue;
        int lo = 0, hi = num_jobs - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (job_tile_base[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const JobDev<T>& J = jobs[lo];
        const int G = J.grid_size;
        const unsigned local = t - J.tile_base;
        const int ty = (int)(local / (unsigned)J.ntx);
        const int tx = (int)(local - (unsigned)ty * J.ntx);
        const P2* part = scratch + (size_t)slot_off[t] * cells;
        #pragma unroll
        for (int k = 0; k < kMergeCells / 256; ++k)
        {
            const int c = (int)ch * kMergeCells + k * 256 + (int)threadIdx.x;
            double re = 0.0, im = 0.0;
            unsigned q = 0;
            for (; q + 4 <= nit; q += 4)
            {
                P2 v[4];
                #pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = part[(size_t)(q + u) * cells + c];
                #pragma unroll
                for (int u = 0; u < 4; ++u) { re += (double)v[u].x; im += (double)v[u].y; }
            }
            for (; q < nit; ++q)
            {
                const P2 v = part[(size_t)q * cells + c];
                re += (double)v.x; im += (double)v.y;
            }
            const int y = c / tw;
            const int x = tx * tw + (c - y * tw);
            const int yg = ty * th + y;
            if (x < G && yg < G)
            {
                T2* g = J.grid + (size_t)yg * G + x;
                T2 gv = *g;
                gv.x = (T)((double)gv.x + re);
                gv.y = (T)((double)gv.y + im);
                *g = gv;
            }
        }
    }
}

// Visibilities whose weighted amplitude is not finite: one wave per
// visibility adds its footprint with float atomics, after the tiled path
// has finished.  The cells it touches end up NaN or +-Inf whatever the
// order of the additions, exactly as the serial code leaves them; keeping
// these few visibilities out of the tiled kernels lets those read zero
// rows for out-of-footprint taps (0 * Inf would be NaN).  A pass whose jobs
// counted no such visibility (job_out[3j+2], from k_job_reduce) returns at
// once: the host-sync-free pipeline launches this kernel without knowing.
template <typename T>
__global__ __launch_bounds__(256)
void k_grid_nonfinite(const typename V2<T>::t* __restrict__ tables,
        const int4* __restrict__ meta, const typename V4<T>::t* __restrict__ amp,
        const JobDev<T>* __restrict__ jobs, const unsigned long long* __restrict__ job_vis_off,
        int num_jobs, unsigned long long n, const unsigned long long* __restrict__ job_out)
{
    typedef typename V2<T>::t T2;
    const int lane = lane_id();
    {
        unsigned long long cnt = 0;
        for (int j = lane; j < num_jobs; j += 64) cnt += job_out[3 * j + 2];
        if (!__ballot(cnt != 0)) return;
    }
    const unsigned long long nw = (unsigned long long)gridDim.x * (blockDim.x >> 6);
    const unsigned long long w0 = (unsigned long long)blockIdx.x * (blockDim.x >> 6) +
            (unsigned)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (unsigned long long i = w0; i < n; i += nw)
    {
        const int4 mt = ld_const(meta, i);
        if ((mt.z & (kFlagNonFinite | kFlagSkip)) != kFlagNonFinite) continue;
        int lo = 0, hi = num_jobs - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (ld_const(job_vis_off, mid) <= i) lo = mid; else hi = mid - 1;
        }
        const JobDev<T> J = ld_const(jobs, lo);
        const typename V4<T>::t a = ld_const(amp, i);
        const int s = meta_support(mt.z);
        const int W = 2 * s + 1;
        const bool neg_u = mt.z & kFlagNegU, neg_v = mt.z & kFlagNegV;
        for (int jr = -s; jr <= s; ++jr)
        {
            const T2* krow = tables + (unsigned)mt.w + (neg_v ? s - jr : s + jr) * W;
            T* grow = reinterpret_cast<T*>(J.grid + (size_t)(mt.y + jr) * J.grid_size + mt.x);
            for (int k = -s + lane; k <= s; k += 64)
            {
                const T2 c = krow[neg_u ? s - k : s + k];
                atomicAdd(grow + 2 * k, fma(c.y, a.y, c.x * a.x));
                atomicAdd(grow + 2 * k + 1, fma(c.y, a.w, c.x * a.z));
            }
        }
    }
}

// ===========================================================================
// Degridding (adjoint)
// ===========================================================================
// One wave per visibility.  The footprint's W*W taps are dealt to the 64
// lanes in row-major order (lane l takes taps l, l+64, ...: row t / W,
// column t % W, kept up incrementally), so every lane works whatever W is
// and a wave's 64 table reads are one contiguous run; per-lane partials are
// reduced across the wave at the end.  The visibilities go in the order
// `order` gives (grid tiles of their centre cells: waves that run together
// read neighbouring cells), or in index order when it is null.  Output per
// visibility: sum over taps of conj-folded tap * grid cell, times the
// weight when apply_weight (the adjoint of oskar_grid_wproj.c:190-209).
template <typename T>
__global__ __launch_bounds__(256)
void k_degrid(const typename V2<T>::t* __restrict__ tables,
        const int4* __restrict__ meta, const unsigned* __restrict__ order,
        const T* __restrict__ weight, unsigned long long n,
        const typename V2<T>::t* __restrict__ grid, int grid_size, int apply_weight,
        typename V2<T>::t* __restrict__ out)
{
    typedef typename V2<T>::t T2;
    const int lane = lane_id();
    const unsigned long long nw = (unsigned long long)gridDim.x * (blockDim.x >> 6);
    const unsigned long long w0 = (unsigned long long)blockIdx.x * (blockDim.x >> 6) +
            (unsigned)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (unsigned long long q = w0; q < n; q += nw)
    {
        const unsigned long long i = order ? (unsigned long long)ld_const(order, q) : q;
        const int4 mt = ld_const(meta, i);
        T re = (T)0, im = (T)0;
        if (!(mt.z & kFlagSkip))
        {
            const int s = meta_support(mt.z);
            const int W = 2 * s + 1;
            const int n2 = W * W;
            const bool neg_u = mt.z & kFlagNegU, neg_v = mt.z & kFlagNegV;
            const T cj = (mt.z & kFlagConj) ? (T)-1 : (T)1;
            // 64 taps further on: dr rows and dc columns (one more row when
            // the column passes W)
            const int dr = 64 / W, dc = 64 - dr * W;
            int r = lane / W, c = lane - (lane / W) * W;
            const T2* tab = tables + (unsigned)mt.w;                 // table row 0
            const T2* g0 = grid + (size_t)(mt.y - s) * grid_size + (mt.x - s);
            for (int t = lane; t < n2; t += 64)
            {
                const int tr = neg_v ? W - 1 - r : r;
                const int tc = neg_u ? W - 1 - c : c;
                const T2 cv = tab[tr * W + tc];
                const T2 g = g0[r * grid_size + c];
                const T ci = cv.y * cj;
                re = fma(cv.x, g.x, re); re = fma(ci, g.y, re);
                im = fma(cv.x, g.y, im); im = fma(-ci, g.x, im);
                r += dr;
                c += dc;
                if (c >= W) { c -= W; ++r; }
            }
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if (lane == 0)
        {
            const T w = apply_weight ? weight[i] : (T)1;
            T2 v; v.x = re * w; v.y = im * w;
            out[i] = v;
        }
    }
}

// Degridding order keys: the 64 x 64 grid tile of each visibility's centre
// cell (tiles_x per tile row), skipped visibilities under skip_key.
__global__ void k_degrid_keys(const int4* __restrict__ meta, unsigned long long n, int tiles_x,
        unsigned skip_key, unsigned* __restrict__ keys, unsigned* __restrict__ vals)
{
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
            i < n; i += (unsigned long long)gridDim.x * blockDim.x)
    {
        const int4 m = meta[i];
        keys[i] = (m.z & kFlagSkip) ? skip_key
                                     : (unsigned)((m.y >> 6) * tiles_x + (m.x >> 6));
        vals[i] = (unsigned)i;
    }
}

} // namespace gg
