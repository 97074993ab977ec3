// gg_grid_strips.hpp -- FP32 region gridding, one wave per 16-column strip.
//
// A workgroup of 16 waves owns a 256 x 128 region of the grid (12 waves and
// 192 x 64 when compensated); wave w owns the strip of columns 16w .. 16w+15
// of it, all its rows.  The strip's rows go in slots of four; a lane owns
// one cell of each slot (lane = 16 * row-in-slot + column), so the strip's
// cells are 32 (compensated: 16) complex accumulators per lane, kept in
// registers.  A footprint of width F covers about F/16 + 1 strips, so few
// lanes idle at its left and right edges.
//
// The region's visits are sorted by kernel table, so its list is a sequence
// of runs that read one table (in any of the four mirror orientations).  A
// run at least as long as its table-width class's threshold, whose table is
// at most kStripMaxW wide, is staged in LDS once, unmirrored, at a fixed row
// stride with zero rows and columns around it; every visit of the run then
// reads a lane's tap of slot i at a constant offset from one per-visit base
// address.  Other visits read taps straight from the padded tables in global
// memory with raw buffer loads.
//
// Rows go in chunks of eight (two slots).  A visit touches chunks q0..q1 of
// the strip.  From LDS it takes the pairs of chunks (16 rows, four slots)
// that hold them, the four slots of a pair read together and all four
// multiplied; from global memory the touched chunks' footprint slots are
// requested kDirectGroup chunks at a time, others skipped with wave-uniform
// branches.  Rows outside the footprint read zero rows, columns outside it
// zero columns (staged) or nothing (direct: the buffer offset is out of
// range).
//
// A work item's direct visits go first (pass 1; barriers keep the waves on
// the same visits where they read wide tables, so that neighbouring strips
// share the lines they fetch), then its staged runs (pass 2).  Each wave
// walks all visits of a run segment 64 at a time: one vector load per lane
// of the records, a ballot of the visits whose footprint meets its strip,
// then those visits one by one with their records broadcast by readlane and
// the per-visit addressing done in scalar registers.  Every load in the loop
// nest is a raw buffer load (scalar base, 32-bit offset), which keeps
// per-lane address registers to one per stream.  At the end of a work item
// the strip is added to the grid (single-item regions) or stored in the
// item's partial slot (merged in item order, FP64, afterwards).
#pragma once

#include "gg_kernels.hpp"

namespace gg {

constexpr int kStripW = 16;                        // columns per wave
// Geometry of a strip kernel: WAVES waves per workgroup (one workgroup per
// CU: the staged table takes the LDS), each owning 16 columns x ROWS rows,
// and COMP for compensated accumulation.
//  - Plain FP32: 16 waves (four per SIMD, up to 128 VGPRs, 64 of them
//    accumulators) on 256 x 128 regions, 8 chunks of global taps loaded
//    together (2, 3, 4: +4.9%, +3.2%, +0.7%; 16 spills).  Four waves per
//    SIMD hide more of the staged pass's LDS
//    latency than three at 12 waves (8 chunks, up to 168 VGPRs): the staged
//    pass alone took 53.7 ms against 60.5 ms, for a slightly slower direct
//    pass.
//  - Compensated FP32: 12 waves on 192 x 64 regions (16 slots per lane:
//    32 running sums, and 64 TwoSum sums and error terms), 8 chunks (2
//    chunks: +5%, 4: +1%).
#ifndef GG_STRIP_WAVES
#define GG_STRIP_WAVES 16
#endif
template <int WAVES, int ROWS, bool COMP>
struct StripCfg
{
    static constexpr int kWaves = WAVES;
    static constexpr int kBlock = WAVES * 64;
    static constexpr int kRegionW = WAVES * kStripW;
    static constexpr int kRegionH = ROWS;
    static constexpr int kSlots = ROWS / 4;          // slots of 16 columns x 4 rows
    static constexpr int kChunks = ROWS / 8;         // chunks of 8 rows
#ifndef GG_DIRECT_GROUP
#define GG_DIRECT_GROUP 8
#endif
#ifndef GG_DIRECT_GROUP_COMP
#define GG_DIRECT_GROUP_COMP 8
#endif
    static constexpr int kDirectGroup = COMP ? GG_DIRECT_GROUP_COMP : GG_DIRECT_GROUP;
    static constexpr bool kComp = COMP;
    static_assert(ROWS % 16 == 0 && ROWS <= 128, "blocks of 16 rows, two LDS bases of 64 rows");
    static_assert(kRegionW * kRegionH % kMergeCells == 0, "merge blocks tile a region");
};
using StripPlain = StripCfg<GG_STRIP_WAVES, 128, false>;
using StripComp = StripCfg<12, 64, true>;
constexpr int kStripS = 144;      // LDS row stride in taps; 144 = 16 (mod 32): the two
                                  // 128-byte row segments of a half-wave use disjoint banks
constexpr int kStripPad = 7;      // zero rows above and below a staged table
// Staged row 0 starts at element kStripGuard.  Reads of a touched pair of
// chunks cover rows -8 .. W+21 and columns -15 .. W+14 of the staged
// layout: the eight rows before row 0 are inside the array (their taps are
// never multiplied), the 16 taps right before row 0 are zero (columns -15..-1
// of row 0 are multiplied).
constexpr int kStripGuard = 8 * kStripS + 16;
constexpr int kStripLds = 20446;  // float2 taps: 160 KiB less the prefetch area and a counter
constexpr int kStripMaxW = 111;   // widest table staged
#ifndef GG_STRIP_PREFETCH
#define GG_STRIP_PREFETCH 1
#endif
static_assert(kStripGuard - 8 * kStripS - 15 >= 0, "rows -8.. must stay inside the LDS array");
static_assert(kStripGuard + (kStripMaxW + 2 * kStripPad + 7) * kStripS + 128 <= kStripLds,
              "the zero rows after the widest staged table must stay inside the LDS array");
static_assert(kStripGuard + (kStripMaxW + 21) * kStripS + kStripMaxW + 15 < kStripLds,
              "reads of the widest staged table must stay inside the LDS array");
static_assert(kStripS - 15 >= kStripMaxW, "columns -15..-1 must land on zero columns");
static_assert(kStripGuard + (kStripMaxW + 2 * kStripPad - 1) * kStripS + 128 <= kStripLds,
              "staging of the widest table must stay inside the LDS array");
static_assert(kTablePadRows >= kStripPad && kTablePadRows >= 3,
              "staging copies the kStripPad zero rows on either side of a table from its padded block; "
              "direct reads of a touched slot go up to 3 rows past it");
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

// Copy a table into LDS at stride kStripS: staged row t + kStripPad,
// column c holds table row t, column c, for t = -7 .. W+6 (the pad rows of
// the padded block are zero), and staged rows -8 .. -1 and W+14 .. W+21
// (table rows -15 .. -8 and W+7 .. W+14, which the pair of chunks around a
// footprint's first or last row reaches) are cleared; columns W .. 127 are
// zero and columns 128 .. kStripS-1 stay zero from the kernel's start (no
// staged table reaches them).  Visits of all four orientations read this
// one copy: mirrored columns and rows are in their LDS addresses.
//
// The copy goes straight from memory to LDS (buffer loads with the LDS
// flag): no registers hold the data, so every row's loads are issued at
// once and a table costs one round trip to memory.  Wave w copies staged
// rows w-8, w-8 + WAVES, ...; one 16-byte load per lane and row moves
// columns 2l, 2l+1 to LDS (columns 0 .. 127 per row).  Lanes with 2l >= W,
// and every lane of the zero rows, load nothing (the offset is out of the
// buffer's range) and write zeros; lane s (2s = W-1) also brings the tap
// after the row's last (the next row's first), which one zero store per
// row then clears.  The caller's barrier after the copy orders it before
// any read.
typedef __attribute__((address_space(3))) void lds_void_t;
// s_waitcnt immediate (gfx9 encoding) for vmcnt(0) with expcnt and lgkmcnt
// left at their maxima: wait until this wave's vector memory loads,
// direct-to-LDS copies included, have completed.
constexpr int kWaitVmcnt0 = (7 << 4) | (15 << 8);
static_assert(kStripMaxW < 128, "staged columns 128 .. kStripS-1 are always zero");
static_assert((kStripGuard * 8) % 16 == 0 && (kStripS * 8) % 16 == 0,
              "staged rows must start on 16 bytes");

template <int WAVES>
__device__ __forceinline__ void stage_strip_table(float2* tab, __amdgpu_buffer_rsrc_t rs,
        int tb, int W, int wave, int lane)
{
    const int rows = W + 2 * kStripPad;
    const unsigned vo = 2 * lane < W ? (unsigned)lane * 16u : 0x80000000u;
    // staged rows -8 .. rows+7: the table's rows, and eight zero rows on
    // either side (an out-of-range offset: the copy writes zeros)
    for (int r = wave - 8; r < rows + 8; r += WAVES)
    {
        const bool in = r >= 0 && r < rows;
        const unsigned so = in ? (unsigned)((tb + (r - kStripPad) * W) * 8) : 0u;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void_t*)(tab + kStripGuard + r * kStripS),
                16, in ? vo : 0x80000000u, so, 0, 0);
    }
    // column W of the rows this wave copied (lane k: row wave-8 + k WAVES),
    // after its copies have landed in LDS (they write that column too).  The
    // wait is explicit: the compiler does not reliably order an LDS store
    // after direct-to-LDS loads of the same array.
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    const int r = wave - 8 + lane * WAVES;
    if (r >= 0 && r < rows) tab[kStripGuard + r * kStripS + W] = make_float2(0.f, 0.f);
}

// Pull a table's block (pad rows included) towards the compute units ahead
// of its staging: one dword per 64 bytes, so that a wave's load touches 4 KiB
// and a 70-wide table takes 10 loads shared by the waves.  The loads go
// through the LDS path into a 256-byte array of their own, so they hold no
// registers, and reads of the staged table need not wait for them (it is a
// different LDS object); the next staging's waits cover them.
template <int WAVES>
__device__ __forceinline__ void prefetch_strip_table(float* sink, __amdgpu_buffer_rsrc_t rs,
        int tb, int W, int wave, int lane)
{
    const unsigned b0 = (unsigned)((tb - kTablePadRows * W) * 8);
    const unsigned len = (unsigned)((W + 2 * kTablePadRows) * W * 8);
    for (unsigned o = (unsigned)wave * 4096u; o < len; o += WAVES * 4096u)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void_t*)sink, 4,
                o + (unsigned)lane * 64u, b0, 0, 0);
}

__device__ __forceinline__ void mac_strip(float& re, float& im, const float2 c, const float4 a)
{
    re = fmaf(c.x, a.x, re);
    re = fmaf(c.y, a.y, re);
    im = fmaf(c.x, a.z, im);
    im = fmaf(c.y, a.w, im);
}

// Chunk loop over LDS taps, two chunks (16 rows, four slots) per block:
// the four slots of a pair that holds a touched chunk are read together,
// then all four multiplied (rows outside the footprint are zero rows).
// t0 / t1: this lane's taps of chunk 0 / chunk 8.
#define GG_STRIP_PAIRS(cm)                                                     \
    {                                                                         \
        _Pragma("unroll")                                                     \
        for (int pq = 0; pq < kStripChunks / 2; ++pq)                         \
        {                                                                     \
            if (((cm) >> (2 * pq)) & 3)                                       \
            {                                                                 \
                const float2* pp = pq < 4 ? t0 + pq * 16 * kStripS            \
                                          : t1 + (pq - 4) * 16 * kStripS;     \
                const float2 c0 = pp[0], c1 = pp[4 * kStripS];                \
                const float2 c2 = pp[8 * kStripS], c3 = pp[12 * kStripS];     \
                {                                                             \
                    mac_strip(re[4 * pq], im[4 * pq], c0, a);                 \
                    mac_strip(re[4 * pq + 1], im[4 * pq + 1], c1, a);         \
                }                                                             \
                {                                                             \
                    mac_strip(re[4 * pq + 2], im[4 * pq + 2], c2, a);         \
                    mac_strip(re[4 * pq + 3], im[4 * pq + 3], c3, a);         \
                }                                                             \
            }                                                                 \
        }                                                                     \
    }

// The same for a visit whose table rows run backwards (off_v < 0): slot i
// lies 4i rows further up the staged table, so a block's four slots are
// read from its lowest address (that of its last slot) and assigned in
// reverse.  t0 / t1: the lowest addresses of blocks 3 and 7 (slots 15 and
// 31), so that every offset is positive.
#define GG_STRIP_PAIRS_VMIRROR(cm)                                             \
    {                                                                         \
        _Pragma("unroll")                                                     \
        for (int pq = 0; pq < kStripChunks / 2; ++pq)                         \
        {                                                                     \
            if (((cm) >> (2 * pq)) & 3)                                       \
            {                                                                 \
                const float2* pp = pq < 4 ? t0 + (3 - pq) * 16 * kStripS      \
                                          : t1 + (7 - pq) * 16 * kStripS;     \
                const float2 c0 = pp[0], c1 = pp[4 * kStripS];                \
                const float2 c2 = pp[8 * kStripS], c3 = pp[12 * kStripS];     \
                {                                                             \
                    mac_strip(re[4 * pq], im[4 * pq], c3, a);                 \
                    mac_strip(re[4 * pq + 1], im[4 * pq + 1], c2, a);         \
                }                                                             \
                {                                                             \
                    mac_strip(re[4 * pq + 2], im[4 * pq + 2], c1, a);         \
                    mac_strip(re[4 * pq + 3], im[4 * pq + 3], c0, a);         \
                }                                                             \
            }                                                                 \
        }                                                                     \
    }

template <class Cfg>
__global__ __launch_bounds__(Cfg::kBlock) __attribute__((amdgpu_waves_per_eu(Cfg::kWaves / 4)))
void k_grid_strips(const GridArgs<float> A)
{
    constexpr int kStripWaves = Cfg::kWaves;
    constexpr int kStripBlock = Cfg::kBlock;
    constexpr int kRegionW = Cfg::kRegionW;
    constexpr int kRegionH = Cfg::kRegionH;
    constexpr int kStripRows = Cfg::kRegionH;
    constexpr int kStripSlots = Cfg::kSlots;
    constexpr int kStripChunks = Cfg::kChunks;
    constexpr int kDirectGroup = Cfg::kDirectGroup;
    constexpr bool kComp = Cfg::kComp;
    __shared__ float2 tab[kStripLds];
    __shared__ float pf_sink[64];                 // prefetch loads land here
    const unsigned num_items = ld_const(A.d_num_items, 0);
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // Each pass and the write-out take the lane's row and column in its slots
    // (lr, lc) from their own copy of the lane id, hidden from the compiler,
    // so that nothing derived from it is kept live across the item loop.
#define GG_LANE_RC                                                              \
        int ln_ = lane;                                                         \
        __asm__ volatile("" : "+v"(ln_));                                       \
        const int lr = ln_ >> 4, lc = ln_ & 15;
    const __amdgpu_buffer_rsrc_t rs_tab = table_rsrc(A.tables, A.table_bytes);
    // all of LDS zero: the 16 taps before staged row 0 and columns 128 ..
    // kStripS-1 of every row are read as zeros and never written again
    for (int i = threadIdx.x; i < kStripLds; i += kStripBlock) tab[i] = make_float2(0.f, 0.f);
    __syncthreads();
    int staged_base = -1;                          // table in LDS (uniform)

    // A persistent grid (a multiple of 8 workgroups, dealt out to the XCDs
    // round-robin): workgroup b takes the next free position of its XCD,
    // x = b mod 8, from that XCD's counter in item_ctr -- positions x, x+8,
    // x+16, ... in turn -- so that uneven items balance over the CUs and a
    // staged table stays in LDS for the next item.  Within each block of 8g
    // positions XCD x takes the g consecutive items x*g .. x*g+g-1 (g =
    // item_group), so neighbouring regions -- which share the tables of the
    // visibilities whose footprints straddle them -- run on one L2, while all
    // XCDs still advance through the item list together.
    const unsigned g = A.item_group, blk = 8u * g;
    const unsigned npos = (num_items + blk - 1) / blk * blk;
    __shared__ unsigned s_next;
    for (;;)
    {
        {
            const unsigned x = blockIdx.x & 7u;
            if (threadIdx.x == 0) s_next = atomicAdd(A.item_ctr + x, 1u) * 8u + x;
        }
        __syncthreads();
        const unsigned p = __builtin_amdgcn_readfirstlane(s_next);
        __syncthreads();                   // (read by every wave before the next write)
        if (p >= npos) break;
        const unsigned item = p / blk * blk + (p & 7u) * g + (p >> 3) % g;
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
        // Compensated: every kFlushEvery visits of the strip the FP32 sums
        // re / im fold into TwoSum running sums (sre, sim) and error terms
        // (cre, cim), which the write-out combines in FP64.
        constexpr int kC = kComp ? kStripSlots : 1;
        float sre[kC], sim[kC], cre[kC], cim[kC];
        #pragma unroll
        for (int i = 0; i < kC; ++i) sre[i] = sim[i] = cre[i] = cim[i] = 0.f;
        int since = 0;                                 // visits since the last fold
        auto visit_done = [&]() {
            if constexpr (kComp)
            {
                if (++since == kFlushEvery)
                {
                    since = 0;
                    #pragma unroll
                    for (int i = 0; i < kStripSlots; ++i)
                    {
                        float s0, e0, s1, e1;
                        two_sum(sre[i], re[i], s0, e0);
                        two_sum(sim[i], im[i], s1, e1);
                        sre[i] = s0; sim[i] = s1; cre[i] += e0; cim[i] += e1;
                        re[i] = im[i] = 0.f;
                    }
                }
            }
        };

        // (resources based at the item's records keep buffer offsets small;
        // reads past the item's last visit return zero)
        const __amdgpu_buffer_rsrc_t rs_vmeta = table_rsrc(A.vmeta + v0, (v1 - v0) * 16u);
        const __amdgpu_buffer_rsrc_t rs_vamp = table_rsrc(A.vamp + v0, (v1 - v0) * 8u);

        // The visits of a batch of 64 (lane l: visit b + l) that this strip
        // processes: those that meet its columns, of the wanted kind.
        auto batch_mask = [&](unsigned b, unsigned e, int4& mv, float2& av, bool staged_kind) {
            const unsigned pl = b + (unsigned)lane;
            mv = rb_x4<int4>(rs_vmeta, (pl - v0) * 16u);
            av = load_tap(rs_vamp, (pl - v0) * 8u, 0u, (float2*)nullptr);
            const int s = meta_support(mv.z);
            return __ballot(pl < e && ((mv.z & kVisitStaged) != 0) == staged_kind &&
                            mv.x - s <= xs + kStripW - 1 && mv.x + s >= xs);
        };
        // Per visit (scalar, from lane l of the batch): support, footprint
        // origin relative to the strip (kc: footprint column of strip column
        // 0; top: footprint row 0, region row), the mask cm of the 8-row
        // chunks touched (both precomputed in the record: y = (cm << 16) |
        // (top & 0xffff)), coefficients.  (A visit of the region's list meets
        // its rows: top <= region height - 1, top + 2s >= 0.)
// The FMA coefficients (re += c.x a.x + c.y a.y, im += c.x a.z + c.y a.w)
// come from the weighted amplitude (vr, vi) and the conjugation flag:
// (vr, -vi, vi, vr), or (vr, vi, vi, -vr) when conjugated.
#define GG_VISIT_HEAD                                                           \
        const int x = bcast(mv.x, l), py = bcast(mv.y, l), z = bcast(mv.z, l);  \
        float4 a;                                                               \
        {                                                                       \
            const float vr = bcast(av.x, l), vi = bcast(av.y, l);               \
            const bool cj = z & kFlagConj;                                      \
            a.x = vr; a.y = cj ? vi : -vi; a.z = vi; a.w = cj ? -vr : vr;       \
        }                                                                       \
        const int s = meta_support(z);                                          \
        const int kc = xs - (x - s);                                            \
        const int top = (py << 16) >> 16;                                       \
        const int cm = (int)((unsigned)py >> 16);

        {
        GG_LANE_RC
        // Pass 1: visits of runs read straight from the tables in global
        // memory.  A barrier per batch that holds a direct visit of support
        // 33 or more keeps the waves on the same visits, so that the cache
        // lines that neighbouring strips' row segments share are fetched
        // once (about a third less traffic than free-running waves; without
        // it the kernel took 8% longer at support 72).  A batch that holds
        // one of support 49 or more is taken in quarters of 16 visits with a
        // barrier before each (plain kernel: 1.9% faster than one barrier per
        // batch at support 72, halves 1.7%).  Smaller tables stay in the caches, and
        // there the barrier only costs (8% at support 36).  (Every wave holds
        // the batch's records: the tests are uniform.)
        for (unsigned b = v0; b < v1; b += 64)
        {
            int4 mv;
            float2 av;
            const unsigned long long todo_batch = batch_mask(b, v1, mv, av, false);
            const bool direct = b + (unsigned)lane < v1 && !(mv.z & kVisitStaged);
            const bool sync = __ballot(direct && meta_support(mv.z) >= 33) != 0ull;
            const int parts = !kComp && __ballot(direct && meta_support(mv.z) >= 49) ? 4 : 1;
            for (int part = 0; part < parts; ++part)
            {
                if (part > 0 || sync) __syncthreads();
                unsigned long long todo = parts == 1 ? todo_batch : todo_batch & (0xffffull << (16 * part));
                while (todo)
                {
                    const int l = __builtin_ctzll(todo);
                    todo &= todo - 1ull;
                    GG_VISIT_HEAD
                    const int tb = bcast(mv.w, l);
                    const int W = 2 * s + 1;
                    const bool neg_v = z & kFlagNegV, neg_u = z & kFlagNegU;
                    const int fu = (z >> kVisitFoldShift) & 3, fv = (z >> (kVisitFoldShift + 2)) & 3;
                    // Lane's footprint column and its table column.  A table
                    // symmetric in its columns is read in its right half only:
                    // column c < s reads column 2s - c (fold 1: symmetric about
                    // the centre) or 2s - 1 - c (fold 2: about the half-column
                    // before it), which holds the same tap, so that a visit
                    // fetches fewer bytes and the strips either side of the
                    // centre share lines.
                    const int k = kc + lc;
                    int kt = neg_u ? W - 1 - k : k;
                    if (fu && kt < s) kt = (2 * s + 1 - fu) - kt;
                    const unsigned vcol = (k >= 0 && k < W) ? (unsigned)kt * 8u : 0x80000000u;
                    // Rows: the lane's table row in slot i is t = t0 + i dt (rows
                    // run backwards when mirrored).  A table symmetric about its
                    // centre row (fold 1) reads row max(t, 2s - t), the same tap;
                    // rows above and below the footprint (t in -3..-1 and W..W+2:
                    // slots end at most 3 rows past it) stay zero pad rows.  (A
                    // table symmetric about the half-row before its centre pairs
                    // rows t and 2s-1-t but leaves row 2s unpaired: folding it
                    // would cost more per slot than it saves.)  The table's byte
                    // offsets ascend with t, so the row read is the larger of two
                    // offsets, one add each per slot: oa + i da (row t) and
                    // ob + i db (row 2s - t, or row t again when the rows do not
                    // fold).  Offsets with vcol out of range are at least 2^31,
                    // negative as ints, and the signed max of two such is one of
                    // them: still out of range.
                    const int t0 = neg_v ? W - 1 + top - lr : lr - top;
                    const unsigned vtb = vcol + (unsigned)tb * 8u;
                    const int da = (neg_v ? -32 : 32) * W;
                    const unsigned oa = vtb + (unsigned)(t0 * 8 * W);
                    const unsigned ob = fv == 1 ? vtb + (unsigned)((2 * s - t0) * 8 * W) : oa;
                    const int db = fv == 1 ? -da : da;
                    // slots (4 rows) that hold footprint rows: only those are read
                    const int i0 = max(top, 0) >> 2;
                    const int i1 = min(top + 2 * s, kStripRows - 1) >> 2;
                    const unsigned sm = (unsigned)((2ull << i1) - (1ull << i0));
                    // Touched chunks' taps are requested kDirectGroup chunks at a
                    // time before their multiplies: up to 2 * kDirectGroup loads in
                    // flight.  slot_off(i): the byte offset of the lane's tap in
                    // slot i.
                    auto chunks = [&](auto slot_off) {
                        #pragma unroll
                        for (int h = 0; h < kStripChunks; h += kDirectGroup)
                        {
                            if (!((cm >> h) & ((1 << kDirectGroup) - 1))) continue;
                            float2 tp[kDirectGroup][2];
                            #pragma unroll
                            for (int u = 0; u < kDirectGroup; ++u)
                            {
                                const int q = h + u;
                                if (q >= kStripChunks) continue;
                                if (sm & (1u << (2 * q)))
                                    tp[u][0] = load_tap(rs_tab, slot_off(2 * q), 0u, (float2*)nullptr);
                                if (sm & (2u << (2 * q)))
                                    tp[u][1] = load_tap(rs_tab, slot_off(2 * q + 1), 0u, (float2*)nullptr);
                            }
                            #pragma unroll
                            for (int u = 0; u < kDirectGroup; ++u)
                            {
                                const int q = h + u;
                                if (q >= kStripChunks) continue;
                                if (sm & (1u << (2 * q))) mac_strip(re[2 * q], im[2 * q], tp[u][0], a);
                                if (sm & (2u << (2 * q)))
                                    mac_strip(re[2 * q + 1], im[2 * q + 1], tp[u][1], a);
                            }
                        }
                    };
                    chunks([&](int i) {
                        return (unsigned)max((int)(oa + (unsigned)(i * da)), (int)(ob + (unsigned)(i * db)));
                    });
                    visit_done();
                }
            }
        }
        }
        {
        GG_LANE_RC
        // LDS row of this lane's table row when the visit's footprint starts at
        // strip row 0: staged row kStripPad + lr, or kStripPad - lr (plus the
        // visit's offset) when its rows run backwards; the column (lc, or
        // backwards when mirrored in u) is added per visit.
        const int lane_lds = kStripGuard + (kStripPad + lr) * kStripS;
        const int lane_lds_m = kStripGuard + (kStripPad - lr) * kStripS;
        // The lane's part of a visit's LDS byte offsets (the visit's part,
        // from its record, is added per visit): forwards / backwards in u,
        // rows running forwards (f) or backwards (r); the strip's origin
        // (xs, ys) enters with the sign of the column and row directions.
        const int bpf = 8 * (lane_lds + lc + xs + ys * kStripS);
        const int bnf = 8 * (lane_lds - lc - xs + ys * kStripS);
        const int bpr = 8 * (lane_lds_m + lc + xs - (ys + 60) * kStripS);
        const int bnr = 8 * (lane_lds_m - lc - xs - (ys + 60) * kStripS);
        const __amdgpu_buffer_rsrc_t rs_vlds = table_rsrc(A.vlds + v0, (v1 - v0) * 4u);
        // Pass 2: visits of staged runs, 64 at a time.  The batch's staged
        // visits split into segments at run starts (the first segment may
        // continue a run of the previous batch); each segment's table is
        // copied to LDS unless it is already there, then this strip's visits
        // of the segment are processed.  Every wave sees the same segments,
        // so the barriers around staging are uniform.
        for (unsigned b = v0; b < v1; b += 64)
        {
            const unsigned pl = b + (unsigned)lane;
            const int4 mv = rb_x4<int4>(rs_vmeta, (pl - v0) * 16u);
            const float2 av = load_tap(rs_vamp, (pl - v0) * 8u, 0u, (float2*)nullptr);
            const int vl = (int)rb_u32(rs_vlds, (pl - v0) * 4u);
            const bool stg = pl < v1 && (mv.z & kVisitStaged);
            const unsigned long long stg_mask = __ballot(stg);
            if (!stg_mask) continue;
            const int sv = meta_support(mv.z);
            const unsigned long long meets = __ballot(stg && mv.x - sv <= xs + kStripW - 1 &&
                                                      mv.x + sv >= xs);
            unsigned long long starts = __ballot(stg && (mv.z & kVisitRunStart)) |
                    (stg_mask & (~stg_mask + 1ull));
            const unsigned long long vneg = __ballot(mv.z & kFlagNegV);
            while (starts)
            {
                const int l0 = __builtin_ctzll(starts);
                starts &= starts - 1ull;
                const unsigned long long upto = starts ? (starts & (~starts + 1ull)) - 1ull : ~0ull;
                const int rtb = bcast(mv.w, l0);
                const int rz = bcast(mv.z, l0);
                if (rtb != staged_base)
                {
                    __syncthreads();
                    stage_strip_table<kStripWaves>(tab, rs_tab, rtb, 2 * meta_support(rz) + 1, wave, lane);
                    __syncthreads();
                    staged_base = rtb;
                }
#if GG_STRIP_PREFETCH
                if (starts)           // the batch's next run: its table heads for L2
                {
                    const int l1 = __builtin_ctzll(starts);
                    prefetch_strip_table<kStripWaves>(pf_sink, rs_tab, bcast(mv.w, l1),
                                         2 * meta_support(bcast(mv.z, l1)) + 1, wave, lane);
                }
#endif
                const unsigned long long seg = meets & upto & ~((1ull << l0) - 1ull);
                // visits whose rows run forwards, then those running backwards
                unsigned long long todo = seg & ~vneg;
                while (todo)
                {
                    const int l = __builtin_ctzll(todo);
                    todo &= todo - 1ull;
                    GG_VISIT_HEAD
                    // tap of chunk 0: table column kc + lc (or W-1-kc-lc
                    // mirrored), row -top + lr.  (Two independent bases:
                    // every block's reads then fit the positive offset fields
                    // of ds_read2st64.)
                    const int u = bcast(vl, l);
                    const int bs = (z & kFlagNegU) ? bnf : bpf;
                    int o0 = bs + u, o1 = bs + (u + 8 * 64 * kStripS);
                    __asm__ volatile("" : "+v"(o0), "+v"(o1));
                    const float2* t0 = (const float2*)((const char*)tab + o0);
                    const float2* t1 = (const float2*)((const char*)tab + o1);
                    GG_STRIP_PAIRS(cm)
                    visit_done();
                }
                todo = seg & vneg;
                while (todo)
                {
                    const int l = __builtin_ctzll(todo);
                    todo &= todo - 1ull;
                    GG_VISIT_HEAD
                    // table row of strip row 0: W-1+top, slot i 4i rows up;
                    // t0 / t1 at block 3 / 7 (60 rows up from strip row 0)
                    const int u = bcast(vl, l);
                    const int bs = (z & kFlagNegU) ? bnr : bpr;
                    int o0 = bs + u, o1 = bs + (u - 8 * 64 * kStripS);
                    __asm__ volatile("" : "+v"(o0), "+v"(o1));
                    const float2* t0 = (const float2*)((const char*)tab + o0);
                    const float2* t1 = (const float2*)((const char*)tab + o1);
                    GG_STRIP_PAIRS_VMIRROR(cm)
                    visit_done();
                }
            }
        }
        }
#undef GG_VISIT_HEAD

        {
        GG_LANE_RC
        // Strip out: lane (lr, lc) of slot i is cell (xs + lc, ys + 4i + lr).
        // The lane's base offset goes through an empty asm so that the 32
        // slot addresses are formed here rather than kept live (or spilled)
        // from the start of the kernel.
        const int xo = xs + lc;
        if (nit == 1)
        {
            unsigned long long lb = (unsigned long long)(ys + lr) * (unsigned)G + (unsigned)xo;
            __asm__ volatile("" : "+v"(lb));
            float2* g0 = J.grid + lb;
            const size_t step4 = (size_t)4 * (unsigned)G;
            #pragma unroll
            for (int i = 0; i < kStripSlots; ++i)
            {
                const int yo = ys + 4 * i + lr;
                if (xo < G && yo < G)
                {
                    float2* g = g0 + (size_t)i * step4;
                    float2 gv = *g;
                    if constexpr (kComp)
                    {
                        gv.x = (float)((double)gv.x + ((double)sre[i] + (double)cre[i] + (double)re[i]));
                        gv.y = (float)((double)gv.y + ((double)sim[i] + (double)cim[i] + (double)im[i]));
                    }
                    else
                    {
                        gv.x += re[i];
                        gv.y += im[i];
                    }
                    *g = gv;
                }
            }
        }
        else
        {
            const unsigned slot = ld_const(A.slot_off, t) + m;
            unsigned lo = (unsigned)(lr * kRegionW + wave * kStripW + lc);
            __asm__ volatile("" : "+v"(lo));
            if constexpr (kComp)
            {
                // (FP64 partials: the merge sums them in item order)
                double2* out = reinterpret_cast<double2*>(A.scratch) +
                        (size_t)slot * (kRegionW * kRegionH) + lo;
                #pragma unroll
                for (int i = 0; i < kStripSlots; ++i)
                    out[i * 4 * kRegionW] = make_double2(
                            (double)sre[i] + (double)cre[i] + (double)re[i],
                            (double)sim[i] + (double)cim[i] + (double)im[i]);
            }
            else
            {
                float2* out = reinterpret_cast<float2*>(A.scratch) +
                        (size_t)slot * (kRegionW * kRegionH) + lo;
                #pragma unroll
                for (int i = 0; i < kStripSlots; ++i)
                    out[i * 4 * kRegionW] = make_float2(re[i], im[i]);
            }
        }
        }
        // (the staged table stays valid for the next item: only staging
        // writes LDS, after a barrier)
    }
}

#undef GG_LANE_RC
#undef GG_STRIP_PAIRS
#undef GG_STRIP_PAIRS_VMIRROR

} // namespace gg
