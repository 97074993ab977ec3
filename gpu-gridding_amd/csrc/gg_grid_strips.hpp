// gg_grid_strips.hpp -- FP32 region gridding, one wave per 16-column strip.
//
// A workgroup of sixteen waves owns a 256 x 128 region of the grid; wave w
// owns the strip of columns 16w .. 16w+15 of it, all 128 rows.  A lane owns
// one column of four rows of each of 32 slots (lane = 16 * row-in-slot +
// column), so the strip's 2048 cells are 32 complex accumulators per lane,
// kept in registers.  A footprint of width F covers about F/16 + 1 strip
// columns-blocks of 16, so few lanes idle at its left and right edges.
//
// The region's visits are sorted by kernel table and orientation (mirror
// flags), so its list is a sequence of runs that read one table in one
// orientation.  A run of at least min_stage_run visits whose table is at
// most kStripMaxW wide is staged in LDS once, already mirrored, at a fixed
// row stride with zero rows and columns around it; every visit of the run
// then reads a lane's tap of slot i at a constant offset from one per-visit
// base address.  Other visits read taps straight from the padded tables in
// global memory with raw buffer loads.
//
// Rows go in chunks of eight (two slots).  A visit touches chunks q0..q1 of
// the strip; a chunk is skipped with a wave-uniform branch.  From LDS the
// taps of chunk q+1 are read while chunk q is multiplied; from global memory
// the touched chunks are requested eight at a time before their multiplies.  Rows of a touched chunk
// outside the footprint read zero rows, columns outside it zero columns
// (staged) or nothing (direct: the buffer offset is out of range).
//
// Each wave walks all visits of a run segment 64 at a time: one vector load
// per lane of the records, a ballot of the visits whose footprint meets its
// strip, then those visits one by one with their records broadcast by
// readlane and the per-visit addressing done in scalar registers.  Every
// load in the loop nest is a raw buffer load (scalar base, 32-bit offset),
// which keeps per-lane address registers to one per stream.  At the end of a work item the strip is added to the grid
// (single-item regions) or stored in the item's partial slot (merged in item
// order, FP64, afterwards).
#pragma once

#include "gg_kernels.hpp"

#include <type_traits>

namespace gg {

constexpr int kStripW = 16;                        // columns per wave
constexpr int kStripRows = 128;                    // rows per wave
constexpr int kStripWaves = 16;
constexpr int kStripBlock = kStripWaves * 64;      // 1024 threads
constexpr int kRegionW = kStripWaves * kStripW;    // 256 columns
constexpr int kRegionH = kStripRows;               // 128 rows
static_assert(kRegionW * kRegionH % kMergeCells == 0, "merge blocks tile a region");
constexpr int kStripSlots = kStripRows / 4;        // 32 slots of 16 columns x 4 rows
constexpr int kStripChunks = kStripSlots / 2;      // 16 chunks of 8 rows
constexpr int kStripS = 144;      // LDS row stride in taps; 144 = 16 (mod 32): the two
                                  // 128-byte row segments of a half-wave use disjoint banks
constexpr int kStripPad = 7;      // zero rows above and below a staged table
constexpr int kStripGuard = 16;   // zero taps before the first staged row
constexpr int kStripLds = 20480;  // float2 taps: 160 KiB
constexpr int kStripMaxW = 119;   // widest table staged
// Staged reads stay inside rows 0 .. W+21 (the chunk after the last one
// touched is read ahead) and columns -15 .. W+14 of the staged layout.
static_assert(kStripGuard + (kStripMaxW + 21) * kStripS + kStripMaxW + 15 < kStripLds,
              "read-ahead of the widest staged table must stay inside the LDS array");
static_assert(kStripS - 15 >= kStripMaxW, "columns -15..-1 must land on zero columns");
static_assert(kTablePadRows >= 7, "direct reads of a touched chunk go 7 rows past a table");
// Direct read-ahead goes up to 15 rows past a table: the table buffer has
// kTableGuardRows rows of the widest table before the first and after the
// last table.
static_assert(kTableGuardRows >= 15, "direct read-ahead must stay inside the table buffer");

// Raw buffer loads (scalar resource, 32-bit byte offset; out of range -> 0).
__device__ __forceinline__ unsigned rb_u32(__amdgpu_buffer_rsrc_t rs, unsigned off)
{
    return (unsigned)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0);
}
template <typename X>
__device__ __forceinline__ X rb_x4(__amdgpu_buffer_rsrc_t rs, unsigned off)
{
    static_assert(sizeof(X) == 16, "16-byte record");
    return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0));
}

// Copy a table into LDS, mirrored as orient says, at stride kStripS: staged
// row r + kStripPad, column c holds the tap a visit of this orientation reads
// at footprint offset (r, c).  Staged rows 0 .. W+13 come from table rows
// -7 .. W+6 of the padded block (the pad rows are zero); columns W ..
// kStripS-1 are zero (out-of-range loads).  Wave w copies rows w, w+8, ...
// kStageRowsInFlight at a time; lane l columns l, l+64 and l+128.
constexpr int kStageRowsInFlight = 4;

__device__ __forceinline__ void stage_strip_table(float2* tab, __amdgpu_buffer_rsrc_t rs,
        int tb, int W, int orient, int wave, int lane)
{
    const int rows = W + 2 * kStripPad;
    const bool neg_v = orient & kFlagNegV, neg_u = orient & kFlagNegU;
    unsigned vo[3];
    #pragma unroll
    for (int h = 0; h < 3; ++h)
    {
        const int c = lane + 64 * h;
        vo[h] = c < W ? (unsigned)(neg_u ? W - 1 - c : c) * 8u : 0x80000000u;
    }
    float2* dst = tab + kStripGuard + lane;
    for (int r0 = wave; r0 < rows; r0 += kStageRowsInFlight * kStripWaves)
    {
        float2 v[kStageRowsInFlight][3];
        #pragma unroll
        for (int u = 0; u < kStageRowsInFlight; ++u)
        {
            const int r = r0 + u * kStripWaves;
            const int jr = r - kStripPad;
            const unsigned so = (unsigned)((tb + (neg_v ? W - 1 - jr : jr) * W) * 8);
            #pragma unroll
            for (int h = 0; h < 3; ++h)
            {
                v[u][h] = make_float2(0.f, 0.f);
                if (r < rows) v[u][h] = load_tap(rs, vo[h], so, (float2*)nullptr);
            }
        }
        #pragma unroll
        for (int u = 0; u < kStageRowsInFlight; ++u)
        {
            const int r = r0 + u * kStripWaves;
            if (r < rows)
            {
                dst[r * kStripS] = v[u][0];
                dst[r * kStripS + 64] = v[u][1];
                if (lane + 128 < kStripS) dst[r * kStripS + 128] = v[u][2];
            }
        }
    }
}

__device__ __forceinline__ void mac_strip(float& re, float& im, const float2 c, const float4 a)
{
    re = fmaf(c.x, a.x, re);
    re = fmaf(c.y, a.y, re);
    im = fmaf(c.x, a.z, im);
    im = fmaf(c.y, a.w, im);
}

// Chunk loop over LDS taps: LOAD(q, dst0, dst1) fetches the taps of chunk q
// (slots 2q, 2q+1).  Even chunks' taps live in (e0, e1), odd chunks' in
// (o0, o1), so the read-ahead of chunk q+1 needs no copies.
#define GG_STRIP_CHUNKS(cm, LOAD)                                              \
    {                                                                         \
        const int qa_ = __builtin_ctz((unsigned)(cm));                        \
        float2 e0, e1, o0, o1;                                                \
        if (qa_ & 1) { LOAD(qa_, o0, o1); } else { LOAD(qa_, e0, e1); }       \
        _Pragma("unroll")                                                     \
        for (int q = 0; q < kStripChunks; ++q)                                \
        {                                                                     \
            if ((cm) & (1 << q))                                              \
            {                                                                 \
                if (q + 1 < kStripChunks)                                     \
                {                                                             \
                    if ((q + 1) & 1) { LOAD(q + 1, o0, o1); }                 \
                    else { LOAD(q + 1, e0, e1); }                             \
                }                                                             \
                if (q & 1)                                                    \
                {                                                             \
                    mac_strip(re[2 * q], im[2 * q], o0, a);                   \
                    mac_strip(re[2 * q + 1], im[2 * q + 1], o1, a);           \
                }                                                             \
                else                                                          \
                {                                                             \
                    mac_strip(re[2 * q], im[2 * q], e0, a);                   \
                    mac_strip(re[2 * q + 1], im[2 * q + 1], e1, a);           \
                }                                                             \
            }                                                                 \
        }                                                                     \
    }

__global__ __launch_bounds__(kStripBlock) __attribute__((amdgpu_waves_per_eu(3)))
void k_grid_strips(const GridArgs<float> A)
{
    __shared__ float2 tab[kStripLds];
    const unsigned num_items = ld_const(A.d_num_items, 0);
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lr = lane >> 4, lc = lane & 15;
    // LDS tap of this lane for footprint offset (row lr, column lc) when the
    // visit's footprint starts at strip row 0, column 0.
    const int lane_lds = kStripGuard + (kStripPad + lr) * kStripS + lc;
    const __amdgpu_buffer_rsrc_t rs_tab = table_rsrc(A.tables, A.table_bytes);
    const __amdgpu_buffer_rsrc_t rs_meta = table_rsrc(A.meta, A.meta_bytes);
    const __amdgpu_buffer_rsrc_t rs_amp = table_rsrc(A.amp, A.meta_bytes);
    const __amdgpu_buffer_rsrc_t rs_vis = table_rsrc(A.visit_vis, A.visit_bytes);
    const __amdgpu_buffer_rsrc_t rs_rstart = table_rsrc(A.run_start, A.run_bytes);
    const __amdgpu_buffer_rsrc_t rs_rz = table_rsrc(A.run_z, A.run_bytes);
    const __amdgpu_buffer_rsrc_t rs_rtb = table_rsrc(A.run_tb, A.run_bytes);
    if (threadIdx.x < kStripGuard) tab[threadIdx.x] = make_float2(0.f, 0.f);
    int staged_base = -1, staged_orient = -1;      // table in LDS (uniform)

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
        const JobDev<float> J = ld_const(A.jobs, j);
        const int G = J.grid_size;
        const unsigned local = t - J.tile_base;
        const int ty = (int)(local / (unsigned)J.ntx);
        const int tx = (int)(local - (unsigned)ty * J.ntx);
        const int xs = tx * kRegionW + wave * kStripW;   // strip's first column
        const int ys = ty * kRegionH;                    // strip's first row

        float re[kStripSlots], im[kStripSlots];
        #pragma unroll
        for (int i = 0; i < kStripSlots; ++i) re[i] = im[i] = 0.f;

        // Two passes over the item's visits: first the visits of runs read
        // from global memory (no barriers: waves run at their own pace), then
        // the runs staged in LDS.  Runs' records are fetched 64 at a time
        // (lane l: run rbase + l).
        auto walk = [&](auto staged_pass) {
            constexpr bool STAGED = decltype(staged_pass)::value;
            unsigned rbase = ld_const(A.run_excl, v0) + ld_const(A.run_flag, v0) - 1u;
            unsigned rv_start, rv_next;
            int rv_z, rv_tb;
            auto load_runs = [&]() {
                const unsigned o = (rbase + (unsigned)lane) * 4u;
                rv_start = rb_u32(rs_rstart, o);
                rv_next = rb_u32(rs_rstart, o + 4u);
                rv_z = (int)rb_u32(rs_rz, o);
                rv_tb = (int)rb_u32(rs_rtb, o);
            };
            load_runs();
            int rl = 0;

            for (unsigned b = v0; b < v1; b += 64)
            {
                // Records of visits b + l; which of them meet this strip.
                const unsigned pl = b + (unsigned)lane;
                const unsigned iv = rb_u32(rs_vis, pl * 4u);
                const int4 mv = rb_x4<int4>(rs_meta, iv * 16u);
                const float4 av = rb_x4<float4>(rs_amp, iv * 16u);
                const unsigned long long meets_mask = __ballot(pl < v1 &&
                        mv.x - meta_support(mv.z) <= xs + kStripW - 1 &&
                        mv.x + meta_support(mv.z) >= xs);
                const unsigned bend = min(b + 64u, v1);

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
                    const int rz = bcast(rv_z, rl);
                    const int W0 = 2 * meta_support(rz) + 1;
                    const int orient = rz & kOrientMask;
                    // (max: every step advances, whatever the run table holds)
                    const unsigned e = max(min(rnext, bend), p + 1);
                    const bool stage_run = W0 <= kStripMaxW && rnext - rstart >= A.min_stage_run;
                    if (e >= rnext) ++rl;      // the run ends in this batch
                    const int lb = (int)(p - b), le = (int)(e - b);
                    p = e;
                    if (stage_run != STAGED) continue;
                    if (STAGED && (rtb != staged_base || orient != staged_orient))
                    {
                        __syncthreads();
                        stage_strip_table(tab, rs_tab, rtb, W0, orient, wave, lane);
                        __syncthreads();
                        staged_base = rtb;
                        staged_orient = orient;
                    }
                    const unsigned long long seg = (le == 64 ? ~0ull : ((1ull << le) - 1ull)) &
                            ~((1ull << lb) - 1ull);
                    unsigned long long todo = meets_mask & seg;
                    while (todo)
                    {
                        // Per visit (scalar): support, footprint origin relative
                        // to the strip (kc: footprint column of strip column 0;
                        // top: footprint row 0), chunks q0..q1 touched.
                        const int l = __builtin_ctzll(todo);
                        todo &= todo - 1ull;
                        const int x = bcast(mv.x, l), y = bcast(mv.y, l), z = bcast(mv.z, l);
                        float4 a;
                        a.x = bcast(av.x, l); a.y = bcast(av.y, l);
                        a.z = bcast(av.z, l); a.w = bcast(av.w, l);
                        const int s = meta_support(z);
                        const int kc = xs - (x - s);
                        const int top = y - s - ys;
                        // (a visit of the region's list meets its rows:
                        // top <= 127 and top + 2s >= 0)
                        const int q0 = max(top, 0) >> 3;
                        const int q1 = min(top + 2 * s, kStripRows - 1) >> 3;
                        const int cm = (int)((2u << q1) - (1u << q0));
                        if constexpr (STAGED)
                        {
                            const float2* t0 = tab + (lane_lds + kc - top * kStripS);
                            const float2* t1 = t0 + 64 * kStripS;
#define GG_LDS_LOAD(qq, d0, d1)                                                 \
                            {                                                   \
                                const float2* pq = (qq) < 8 ? t0 + (qq) * 8 * kStripS \
                                                            : t1 + ((qq) - 8) * 8 * kStripS; \
                                d0 = pq[0];                                     \
                                d1 = pq[4 * kStripS];                           \
                            }
                            GG_STRIP_CHUNKS(cm, GG_LDS_LOAD)
#undef GG_LDS_LOAD
                        }
                        else
                        {
                            const int tb = bcast(mv.w, l);
                            const int W = 2 * s + 1;
                            const bool neg_v = z & kFlagNegV, neg_u = z & kFlagNegU;
                            // Lane's footprint column, its table column, and its
                            // row within a slot (rows run backwards when mirrored).
                            const int k = kc + lc;
                            const int kt = neg_u ? W - 1 - k : k;
                            const int rrow = neg_v ? 3 - lr : lr;
                            const unsigned voff = (k >= 0 && k < W) ? (unsigned)(rrow * W + kt) * 8u
                                                                    : 0x80000000u;
                            // Byte offset of chunk 0's first table row (the row
                            // of lane row 0, or row 3 when mirrored), steps.
                            const int row0 = neg_v ? (W - 1) - 3 + top : -top;
                            const int base = (tb + row0 * W) * 8;
                            const int step = (neg_v ? -64 : 64) * W;
                            const int half = step >> 1;
                            // Touched chunks' taps are requested eight chunks
                            // at a time before their multiplies: up to 16
                            // loads in flight.
                            #pragma unroll
                            for (int h = 0; h < kStripChunks; h += 8)
                            {
                                if (!((cm >> h) & 0xFF)) continue;
                                float2 tp[8][2];
                                #pragma unroll
                                for (int u = 0; u < 8; ++u)
                                {
                                    if (cm & (1 << (h + u)))
                                    {
                                        const unsigned so = (unsigned)(base + (h + u) * step);
                                        tp[u][0] = load_tap(rs_tab, voff, so, (float2*)nullptr);
                                        tp[u][1] = load_tap(rs_tab, voff, so + (unsigned)half,
                                                            (float2*)nullptr);
                                    }
                                }
                                #pragma unroll
                                for (int u = 0; u < 8; ++u)
                                {
                                    const int q = h + u;
                                    if (cm & (1 << q))
                                    {
                                        mac_strip(re[2 * q], im[2 * q], tp[u][0], a);
                                        mac_strip(re[2 * q + 1], im[2 * q + 1], tp[u][1], a);
                                    }
                                }
                            }
                        }
                    }
                }
            }
        };
        walk(std::false_type());
        walk(std::true_type());

        // Strip out: lane (lr, lc) of slot i is cell (xs + lc, ys + 4i + lr).
        const int xo = xs + lc;
        if (nit == 1)
        {
            #pragma unroll
            for (int i = 0; i < kStripSlots; ++i)
            {
                const int yo = ys + 4 * i + lr;
                if (xo < G && yo < G)
                {
                    float2* g = J.grid + (size_t)yo * G + xo;
                    float2 gv = *g;
                    gv.x += re[i];
                    gv.y += im[i];
                    *g = gv;
                }
            }
        }
        else
        {
            const unsigned slot = ld_const(A.slot_off, t) + m;
            float2* out = reinterpret_cast<float2*>(A.scratch) +
                    (size_t)slot * (kRegionW * kRegionH);
            #pragma unroll
            for (int i = 0; i < kStripSlots; ++i)
                out[(4 * i + lr) * kRegionW + wave * kStripW + lc] = make_float2(re[i], im[i]);
        }
        // (the staged table stays valid for the next item: only staging
        // writes LDS, after a barrier)
    }
}

#undef GG_STRIP_CHUNKS

} // namespace gg
