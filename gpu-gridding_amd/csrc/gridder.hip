// gridder.hip -- host orchestration and C ABI of the MI355X W-projection gridder.
//
// Replaces oskar_grid_wproj_f/_d (oskar_grid_wproj.h:68-131) behind the same
// signature.  One call = one pass of the pipeline in gg_kernels.hpp over all
// of the call's visibilities, on the context's stream, with two small
// device-to-host reads (visit total, work-item total) to size buffers.
#include "gg_kernels.hpp"
#include "gg_grid_staged.hpp"
#include "gg_grid_strips.hpp"
#include "gg_weights.hpp"
#include "gg_image.hpp"
#include "gg_status.hpp"
#include "mi355_gridder.h"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

using namespace gg;

// ---------------------------------------------------------------------------
// Status
// ---------------------------------------------------------------------------
namespace ggs {

thread_local int t_last_status = MI355_OK;
thread_local char t_last_msg[512] = "";

int fail(int status, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_last_msg, sizeof(t_last_msg), fmt, ap);
    va_end(ap);
    t_last_status = status;
    return status;
}

int ok()
{
    t_last_status = MI355_OK;
    t_last_msg[0] = 0;
    return MI355_OK;
}

} // namespace ggs

// Static device scratch of the allocation-free entry point
// (mi355_grid_wproj_f_dev_noalloc): every device the library runs on holds
// one copy until the process ends.  Host code never uses the name itself
// (on the host it is a stand-in, not a device address); init_ctx resolves
// each device's address once with hipGetSymbolAddress and the pipeline
// carves its buffers out of that address.
// Degridding calls of at least this many visibilities go in grid-tile order.
constexpr size_t kDegridSortMin = 4096;
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

using ggs::fail;
using ggs::ok;

std::mutex g_scratch_lock;
void* g_scratch_addr[64] = {};       // per device: the address of `scratch`

#define HIP_TRY(expr)                                                        \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            return fail(e_ == hipErrorOutOfMemory ? MI355_ERR_OUT_OF_MEMORY  \
                                                  : MI355_ERR_HIP,           \
                    "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                    __FILE__, __LINE__);                                     \
    } while (0)

#define TRY(expr)                                                            \
    do { int s_ = (expr); if (s_ != MI355_OK) return s_; } while (0)

// Grow-only device allocation.
struct DevBuf
{
    void* p = nullptr;
    size_t cap = 0;
    template <typename X> X* as() const { return static_cast<X*>(p); }
};


// Pinned host memory (hipHostMalloc) that the host-pointer drop-ins stage
// through: the DMA engines read and write it at full PCIe rate, which
// pageable memory does not reach.
struct HostBuf
{
    void* p = nullptr;
    size_t cap = 0;
};

int ensure_host(HostBuf& b, size_t bytes)
{
    if (bytes <= b.cap) return MI355_OK;
    if (b.p) { (void)hipHostFree(b.p); b.p = nullptr; b.cap = 0; }
    if (hipHostMalloc(&b.p, bytes, hipHostMallocDefault) != hipSuccess)
    {
        (void)hipGetLastError();
        b.p = nullptr;
        return MI355_ERR_OUT_OF_MEMORY;
    }
    b.cap = bytes;
    return MI355_OK;
}

void release_host(HostBuf& b)
{
    if (b.p) (void)hipHostFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// rows x row_bytes from src (row pitch sp) to dst (row pitch dp), the rows
// split over up to 16 host threads (one per 4 MiB)
void copy_rows(void* dst, size_t dp, const void* src, size_t sp, size_t row_bytes, size_t rows)
{
    const size_t total = row_bytes * rows;
    unsigned nt = std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, total >> 22));
    nt = (unsigned)std::min<size_t>(nt, rows);
    auto part = [&](size_t r0, size_t r1) {
        for (size_t r = r0; r < r1; ++r)
            std::memcpy(static_cast<char*>(dst) + r * dp,
                        static_cast<const char*>(src) + r * sp, row_bytes);
    };
    if (nt <= 1) { part(0, rows); return; }
    std::vector<std::thread> th;
    const size_t per = (rows + nt - 1) / nt;
    for (unsigned t = 1; t < nt; ++t)
    {
        const size_t r0 = std::min(rows, t * per), r1 = std::min(rows, r0 + per);
        if (r0 < r1) th.emplace_back(part, r0, r1);
    }
    part(0, std::min(rows, per));
    for (auto& t : th) t.join();
}

int bits_for(unsigned long long n)   // bits needed to hold values 0..n-1
{
    int b = 0;
    while (b < 64 && (1ull << b) < n) ++b;
    return b;
}

unsigned grid_blocks(unsigned long long work, unsigned per_block, unsigned cap)
{
    unsigned long long b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

enum Phase { EV_START, EV_PREP, EV_SELECT, EV_BUCKET, EV_GRID, EV_MERGE, EV_COUNT };

} // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct mi355_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned visits_per_item = 0;   // 0: automatic (auto_visits_per_item)
    bool compensated = false;
    bool timing = false;
    bool cache_kernels = false;
    int grid_kernel = 0;            // 0 auto (strips for plain FP32), 1 direct 64x64, 2 staged 64x64
    unsigned min_stage_run = 0;     // 0: automatic (stage_run_default)
    // per support class (tables at most 33, 65, 97, 111 wide): the strip
    // kernel's staged-run threshold, 0 = min_stage_run
    unsigned msr_class[4] = {0, 0, 0, 0};
    // the autotuner's choices per workload shape (signature -> settings)
    struct Tuned { std::vector<unsigned char> sig; int kernel; unsigned vpi, msr, group;
                   unsigned cls[4]; double ms; };
    std::vector<Tuned> tuned;
    bool tune_cache = true;
    unsigned item_group = 1;        // strip kernel: consecutive work items per XCD
    int num_cu = 256;

    // Host-sync-free mode (buffers sized from upper bounds, counts read on
    // the device) and stream capture in progress; set by the graph entry.
    bool sync_free = false;
    bool capturing = false;
    bool stats_deferred = false;    // stats to be read from the device
    unsigned long long deferred_points = 0;
    int deferred_jobs = 0;
    // captured batched pipeline
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    std::vector<unsigned char> graph_key;
    bool stats_untimed = false;      // last call was a graph launch (no phase events)
    // The batched graph runs on a context of its own, whose buffers, job
    // table and uploads no other call touches; the last call on this
    // context was a graph call when stats_in_graph is set.
    mi355_ctx* graph_ctx = nullptr;
    bool stats_in_graph = false;
    // The allocation-free entry point runs on a context of its own whose
    // buffers are carved, call by call, out of the device's static scratch
    // array (arena); the last call on this context was one of those when
    // stats_in_noalloc is set.
    mi355_ctx* noalloc_ctx = nullptr;
    bool stats_in_noalloc = false;
    // the last strip-kernel pass's sorted tile lists (mi355_ctx_get_buckets)
    bool bk_valid = false;
    int bk_cur = 0, bk_key_table_bits = 0;
    unsigned long long bk_count = 0;
    unsigned char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;

    // prepared kernel tables
    const void* cache_conv = nullptr;
    int cache_prec = 0, cache_ov = 0, cache_csh = 0;
    std::vector<int> cache_support;
    bool tables_valid = false;
    std::vector<unsigned> plane_off;
    unsigned long long table_elems = 0;   // guard + tables (the zero tail follows)
    unsigned long long table_tail = 0;
    int max_support = 0;

    DevBuf tables, norm_table, support_d, plane_off_d;
    DevBuf meta, amp, num_visits, visit_off;
    DevBuf part_norm, part_skipped, part_taps, part_nonfinite, part_bbox;
    DevBuf scan_partial, scalars;
    DevBuf keys[2], vals[2], hist, hist_off;
    DevBuf tile_start, tile_end, n_items, n_slots, item_off, slot_off, item_tile;
    DevBuf run_start, run_z, run_tb;
    DevBuf vmeta, vamp, vlds;       // sorted visit records (strip kernel)
    DevBuf vrec;                    // per-visibility record + amplitude (strip kernel's gather)
    DevBuf scratch;
    DevBuf jobs_d, job_chunk_off_d, job_vis_off_d, job_tile_base_d, job_out_d, job_bbox_d, job_norm_d;
    // drop-in staging
    DevBuf in_conv, in_uu, in_vv, in_ww, in_vis, in_weight, in_grid, in_norm, in_skipped;
    DevBuf in_weight_out;
    // pinned staging of the host-pointer drop-ins: two bands in turn, each
    // with the event of the copy that last used it
    HostBuf h_stage[2];
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // grid -> image
    DevBuf fft_work, in_image;
    hipfftHandle fft_plan = 0;
    bool fft_valid = false;
    int fft_G = 0, fft_double = 0;

    // host mirrors kept alive across asynchronous uploads
    std::vector<unsigned char> h_jobs;
    std::vector<unsigned> h_job_chunk_off, h_job_tile_base;
    std::vector<unsigned long long> h_job_vis_off;
    std::vector<unsigned long long> h_job_out;
    std::vector<int4> h_job_bbox;

    hipEvent_t ev[EV_COUNT] = {};
    mi355_stats stats = {};

    // device workspace held by the context's buffers, and its cap (0: none)
    size_t ws_bytes = 0;
    unsigned long long max_workspace = 0;
    // set by mi355_ctx_set_stream when the stream changes: switch_ev marks
    // the end of the work queued on the previous stream
    hipEvent_t switch_ev = nullptr;
    bool switch_pending = false;
};

namespace {

std::mutex g_default_lock;
mi355_ctx* g_default[64] = {};

// The strip kernel's staged-run thresholds per support class (tables at
// most 33, 65, 97 and 111 wide).
// Automatic visits per work item: the largest of 8192, 4096, 2048, 1024
// for which the call has at least 6 item sizes of visibilities per CU
// (about 8 items per CU at 1.5 tile visits per visibility), else 1024:
// fewer, larger items mean fewer partial regions to merge (31.4M
// visibilities: 8192 took 78.1 ms per pass against 78.8 with 4096), and
// small calls need small items to spread their heaviest regions over the
// CUs (at 1M and 2.15M visibilities the strip kernel took 28% and 15% less
// time with 1024 than with 4096, the FP64 64x64 kernel 17% and 4%).  The
// choice depends on the visibility count only, so the host-sync-free form
// (which knows the tile-visit count only as a bound) splits the work
// exactly as the synchronous one does.
unsigned auto_visits_per_item(unsigned long long num_points, int num_cu)
{
    unsigned v = 8192;
    while (v > 1024 && num_points < 6ull * v * (unsigned long long)std::max(num_cu, 1)) v >>= 1;
    return v;
}

// Automatic shortest staged run: 3 visits; 8 for the FP64 64x64 kernel,
// whose direct reads keep a whole visit's taps in flight, so that staging a
// table pays only over longer runs (support-72 shape: 255.4 ms at 3, 241.8
// ms at 8; 2.15M-visibility shape 17.0 -> 15.3 ms).
unsigned stage_run_default(bool fp64_tiles) { return fp64_tiles ? 8u : 3u; }

// The strip kernel's thresholds per support class (tables at most 33, 65,
// 97, 111 wide): the class option, else the option for all classes, else
// automatic: 6 for tables up to 65 wide, whose direct reads mostly hit the
// caches, and 3 for wider ones (2.15M-visibility shape 4.30 -> 4.13 ms, 1M
// 1.59 -> 1.51 ms, support-72 shape unchanged).
uint4 stage_thresholds(const mi355_ctx* C)
{
    const unsigned* c = C->msr_class;
    const unsigned b = C->min_stage_run;
    const unsigned d[4] = {b ? b : 6u, b ? b : 6u, b ? b : 3u, b ? b : 3u};
    return make_uint4(c[0] ? c[0] : d[0], c[1] ? c[1] : d[1], c[2] ? c[2] : d[2],
                      c[3] ? c[3] : d[3]);
}

// Work this context has queued that may still read its buffers: everything
// on its current stream, and on the stream it used before the last
// mi355_ctx_set_stream (an event recorded there at the switch).
int wait_ctx_work(mi355_ctx* C)
{
    HIP_TRY(hipStreamSynchronize(C->stream));
    if (C->switch_pending)
    {
        HIP_TRY(hipEventSynchronize(C->switch_ev));
        C->switch_pending = false;
    }
    return MI355_OK;
}

// Grow-only device allocation of a context's buffer b to at least `bytes`.
// A buffer that grows is freed only after the context's queued work is done
// (not the whole device's).  The context's workspace cap
// (MI355_OPT_MAX_WORKSPACE) is checked before any allocation.
int ensure(mi355_ctx* C, DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (bytes <= b.cap) return MI355_OK;
    if (C->arena)
    {
        // allocation-free context: the next 256-byte-aligned piece of the
        // static scratch array (pieces are handed out afresh on every call)
        const size_t at = (C->arena_used + 255) & ~(size_t)255;
        if (at > C->arena_cap || bytes > C->arena_cap - at)
            return fail(MI355_ERR_TOO_LARGE,
                    "this call needs more than the %zu bytes of static scratch the "
                    "allocation-free entry point has (at least %zu)", C->arena_cap, at + bytes);
        b.p = C->arena + at;
        b.cap = bytes;
        C->arena_used = at + bytes;
        return MI355_OK;
    }
    const size_t others = C->ws_bytes - b.cap;
    if (C->max_workspace && others + bytes > C->max_workspace)
        return fail(MI355_ERR_OUT_OF_MEMORY,
                "workspace of %zu bytes would exceed the context's cap of %zu bytes",
                others + bytes, (size_t)C->max_workspace);
    if (b.p)
    {
        TRY(wait_ctx_work(C));
        (void)hipFree(b.p);
        b.p = nullptr;
        C->ws_bytes -= b.cap;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8;      // headroom for slowly growing calls
    if (C->max_workspace && others + want > C->max_workspace) want = bytes;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        want = bytes;
        e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            b.p = nullptr;
            return fail(MI355_ERR_OUT_OF_MEMORY, "device allocation of %zu bytes failed", bytes);
        }
    }
    b.cap = want;
    C->ws_bytes += want;
    return MI355_OK;
}

void release(mi355_ctx* C, DevBuf& b)
{
    if (C->arena)        // (a view into the static scratch: nothing to free)
    {
        b.p = nullptr;
        b.cap = 0;
        return;
    }
    if (b.p) (void)hipFree(b.p);
    C->ws_bytes -= b.cap;
    b.p = nullptr;
    b.cap = 0;
}

void record(mi355_ctx* C, Phase p)
{
    if (C->timing && !C->capturing) (void)hipEventRecord(C->ev[p], C->stream);
}

int init_ctx(mi355_ctx* C, int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    {
        (void)hipGetLastError();
        return fail(MI355_ERR_NO_DEVICE, "no HIP device visible");
    }
    if (device < 0 || device >= n)
        return fail(MI355_ERR_INVALID_ARG, "device %d out of range (have %d)", device, n);
    C->device = device;
    HIP_TRY(hipSetDevice(device));
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
            ncu > 0)
        C->num_cu = ncu;
    for (int i = 0; i < EV_COUNT; ++i) HIP_TRY(hipEventCreate(&C->ev[i]));
    HIP_TRY(hipEventCreateWithFlags(&C->switch_ev, hipEventDisableTiming));
    {
        // the device's static scratch, resolved once per device (device current)
        std::lock_guard<std::mutex> lock(g_scratch_lock);
        if (device < 64 && !g_scratch_addr[device])
        {
            void* ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr)
            {
                (void)hipGetLastError();
                return fail(MI355_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
            }
            g_scratch_addr[device] = ptr;
        }
    }
    return MI355_OK;
}

// Every device buffer a context holds.
std::vector<DevBuf*> device_buffers(mi355_ctx* C)
{
    return {&C->tables, &C->norm_table, &C->support_d, &C->plane_off_d, &C->meta,
        &C->amp, &C->num_visits, &C->visit_off, &C->part_norm, &C->part_skipped,
        &C->part_taps, &C->part_nonfinite, &C->part_bbox, &C->scan_partial, &C->scalars,
        &C->keys[0], &C->keys[1],
        &C->vals[0], &C->vals[1], &C->hist, &C->hist_off, &C->tile_start, &C->tile_end,
        &C->n_items, &C->n_slots, &C->item_off, &C->slot_off, &C->item_tile, &C->run_start,
        &C->run_z, &C->run_tb, &C->scratch,
        &C->jobs_d, &C->job_chunk_off_d, &C->job_vis_off_d, &C->job_tile_base_d,
        &C->job_out_d, &C->job_bbox_d, &C->job_norm_d, &C->in_conv, &C->in_uu, &C->in_vv, &C->in_ww,
        &C->in_vis, &C->in_weight, &C->in_grid, &C->in_norm, &C->in_skipped, &C->in_weight_out,
        &C->fft_work, &C->in_image, &C->vmeta, &C->vamp, &C->vlds, &C->vrec};
}

int get_default_ctx(mi355_ctx** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
    {
        (void)hipGetLastError();
        return fail(MI355_ERR_NO_DEVICE, "no HIP device visible");
    }
    if (dev < 0 || dev >= 64) return fail(MI355_ERR_INVALID_ARG, "device id %d", dev);
    if (!g_default[dev])
    {
        mi355_ctx* C = new mi355_ctx();
        const int s = init_ctx(C, dev);
        if (s != MI355_OK) { delete C; return s; }
        g_default[dev] = C;
    }
    *out = g_default[dev];
    return MI355_OK;
}

// ---------------------------------------------------------------------------
// Argument checks shared by every entry point (host only, before any HIP call)
// ---------------------------------------------------------------------------
int check_kernel_args(size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const void* conv_func)
{
    if (num_w_planes < 1) return fail(MI355_ERR_INVALID_ARG, "num_w_planes must be >= 1");
    if (num_w_planes > (1u << 24)) return fail(MI355_ERR_TOO_LARGE, "num_w_planes > 2^24");
    if (!support) return fail(MI355_ERR_INVALID_ARG, "support is NULL");
    if (!conv_func) return fail(MI355_ERR_INVALID_ARG, "conv_func is NULL");
    if (oversample < 1) return fail(MI355_ERR_INVALID_ARG, "oversample must be >= 1");
    if (conv_size_half < 1) return fail(MI355_ERR_INVALID_ARG, "conv_size_half must be >= 1");
    const int half = (oversample + 1) / 2;
    for (size_t p = 0; p < num_w_planes; ++p)
    {
        const int s = support[p];
        if (s < 0 || s > kMaxSupport)
            return fail(MI355_ERR_INVALID_ARG, "support[%zu] = %d outside [0, %d]", p, s, kMaxSupport);
        // Largest kernel index read by the serial code: |off + ov*s| <= ov/2 + ov*s.
        if ((long long)oversample * s + half >= conv_size_half)
            return fail(MI355_ERR_KERNEL_TOO_SMALL,
                    "plane %zu: support %d needs conv_size_half > %lld, have %d",
                    p, s, (long long)oversample * s + half, conv_size_half);
    }
    return MI355_OK;
}

int check_grid_args(int grid_size)
{
    if (grid_size < 1) return fail(MI355_ERR_INVALID_ARG, "grid_size must be >= 1");
    if (grid_size > (1 << 20)) return fail(MI355_ERR_TOO_LARGE, "grid_size > 2^20");
    return MI355_OK;
}

// ---------------------------------------------------------------------------
// Kernel tables
// ---------------------------------------------------------------------------
template <typename TI, typename TA>
int prepare_tables(mi355_ctx* C, size_t P, const int* support, int ov, int csh,
        const typename V2<TI>::t* conv_dev)
{
    const int prec = (int)(16 * sizeof(TI) + sizeof(TA));
    if (!C->sync_free && C->cache_kernels && C->tables_valid && C->cache_conv == conv_dev &&
            C->cache_prec == prec && C->cache_ov == ov && C->cache_csh == csh &&
            C->cache_support.size() == P &&
            std::equal(support, support + P, C->cache_support.begin()))
        return MI355_OK;
    C->tables_valid = false;
    const int half = (ov + 1) / 2, noff = half + 1;
    int smax = 0;
    for (size_t p = 0; p < P; ++p) smax = std::max(smax, support[p]);
    // Layout: guard rows, then per plane noff^2 padded tables (see
    // k_expand_tables), then the zero tail (guard rows and the zero row the
    // 64-column kernels read for rows outside a footprint).  32-bit byte
    // offsets, below 2^31 so that an offset of 2^31 is out of range.
    const unsigned long long Wmax = 2ull * smax + 1;
    const unsigned long long guard = (unsigned long long)kTableGuardRows * Wmax;
    C->plane_off.resize(P);
    unsigned long long off = guard;
    for (size_t p = 0; p < P; ++p)
    {
        C->plane_off[p] = (unsigned)off;
        const unsigned long long W = 2ull * support[p] + 1;
        off += (W + 2 * kTablePadRows) * W * noff * noff;
        if (off > 0x7FFFFFFFull)
            return fail(MI355_ERR_TOO_LARGE, "expanded kernel tables exceed 2^31 elements");
    }
    C->table_elems = off;
    C->max_support = smax;
    typedef typename V2<TA>::t T2;
    const unsigned long long tail = std::max(guard, 2ull * kMaxSupport + 1);
    if ((off + tail) * sizeof(T2) > 0x7FFFFFFFull)
        return fail(MI355_ERR_TOO_LARGE, "expanded kernel tables exceed 2 GiB");
    TRY(ensure(C, C->tables, (off + tail) * sizeof(T2)));
    C->table_tail = tail;
    {
        // zero the guard and the tail (fill kernels: no memset nodes)
        const unsigned wpe = sizeof(T2) / 4;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_blocks(guard * wpe, 256, 1024)), dim3(256), 0,
                C->stream, C->tables.as<unsigned>(), guard * wpe, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_blocks(tail * wpe, 256, 1024)), dim3(256), 0,
                C->stream, reinterpret_cast<unsigned*>(C->tables.as<T2>() + off), tail * wpe, 0u);
        HIP_TRY(hipGetLastError());
    }
    TRY(ensure(C, C->norm_table, P * (2 * half + 1) * (2 * half + 1) * sizeof(double)));
    TRY(ensure(C, C->support_d, P * sizeof(int)));
    TRY(ensure(C, C->plane_off_d, P * sizeof(unsigned)));
    C->cache_support.assign(support, support + P);
    if (!C->capturing)   // (a captured pipeline reuses what the eager run uploaded)
    {
        HIP_TRY(hipMemcpyAsync(C->support_d.p, C->cache_support.data(), P * sizeof(int),
                hipMemcpyHostToDevice, C->stream));
        HIP_TRY(hipMemcpyAsync(C->plane_off_d.p, C->plane_off.data(), P * sizeof(unsigned),
                hipMemcpyHostToDevice, C->stream));
    }
    {
        // four rows (waves) per workgroup, at most 16 workgroups per plane
        const unsigned rows_max = (unsigned)((Wmax + 2 * kTablePadRows) * noff * noff);
        dim3 g(grid_blocks(rows_max, 4, 16), (unsigned)std::min<size_t>(P, 65535));
        hipLaunchKernelGGL((k_expand_tables<TI, TA>), g, dim3(256), 0, C->stream,
                conv_dev, (int)P, csh, ov, half, C->support_d.as<int>(),
                C->plane_off_d.as<unsigned>(), C->tables.as<T2>());
        HIP_TRY(hipGetLastError());
        const unsigned entries = (unsigned)(P * noff * noff);   // (|off_u|, |off_v|) per plane
        hipLaunchKernelGGL(k_norm_table<TI>, dim3(grid_blocks(entries, 4, 65535)), dim3(256), 0,
                C->stream, conv_dev, (int)P, csh, ov, half, C->support_d.as<int>(),
                C->norm_table.as<double>());
        HIP_TRY(hipGetLastError());
    }
    C->cache_conv = conv_dev;
    C->cache_prec = prec;
    C->cache_ov = ov;
    C->cache_csh = csh;
    C->tables_valid = true;
    return MI355_OK;
}

// ---------------------------------------------------------------------------
// Exclusive scan of n unsigned values (n given on the host or the device)
// ---------------------------------------------------------------------------
int scan_u32(mi355_ctx* C, const unsigned* in, unsigned* out, DevCount cnt,
        unsigned long long n_bound, unsigned long long* d_total, unsigned* d_overflow,
        int write_end)
{
    const unsigned long long nchunk = (n_bound + kScanChunk - 1) / kScanChunk;
    TRY(ensure(C, C->scan_partial, (nchunk + 1) * sizeof(unsigned long long)));
    unsigned long long* part = C->scan_partial.as<unsigned long long>();
    const unsigned blocks = grid_blocks(nchunk, 1, 16384);
    hipLaunchKernelGGL(k_scan_reduce, dim3(blocks), dim3(kScanThreads), 0, C->stream, in, cnt, part);
    hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(1024), 0, C->stream, cnt, part, d_total,
            d_overflow, out, write_end);
    hipLaunchKernelGGL(k_scan_down, dim3(blocks), dim3(kScanThreads), 0, C->stream, in, cnt, part, out);
    HIP_TRY(hipGetLastError());
    return MI355_OK;
}

// Device scalars (u64 slots in C->scalars).  The weighting entry points use
// their own slots (count and bounding box), so that they leave the counts of
// a gridding call, read later by mi355_ctx_get_stats, in place.
enum Scalar { SC_VISITS, SC_OVERFLOW, SC_ITEMS, SC_SLOTS, SC_SCRATCH0, SC_RUNS = 6,
              SC_W_COUNT = 8, SC_W_BOX = 10, SC_ITEM_CTR = 12, SC_COUNT = 16 };

// Stable LSD radix sort of the pairs (C->keys[0], C->vals[0]) on the low
// key_bits bits of the keys: *d_n pairs, n_bound an upper bound of *d_n
// known on the host.  Digits are 8 bits, or 9 where that saves a pass (27
// bits in 3 passes instead of 4).  *cur receives the index of the
// ping-pong buffers holding the result.
int radix_sort(mi355_ctx* C, unsigned long long n_bound, const unsigned long long* d_n,
        int key_bits, int* cur_out)
{
    C->bk_valid = false;              // (the ping-pong buffers are rewritten)
    unsigned long long* sc = C->scalars.as<unsigned long long>();
    const int bits = (key_bits + 8) / 9 < (key_bits + 7) / 8 ? 9 : 8;
    const unsigned radix = 1u << bits;
    const unsigned long long nchunk = (n_bound + kSortChunk - 1) / kSortChunk;
    TRY(ensure(C, C->hist, radix * nchunk * sizeof(unsigned)));
    TRY(ensure(C, C->hist_off, (radix * nchunk + 1) * sizeof(unsigned)));
    int cur = 0;
    for (int shift = 0; shift < key_bits; shift += bits)
    {
        const unsigned blocks = grid_blocks(nchunk, 1, 16384);
        if (bits == 9)
            hipLaunchKernelGGL(k_sort_hist<9>, dim3(blocks), dim3(kSortThreads), 0, C->stream,
                    C->keys[cur].as<unsigned>(), d_n, shift, C->hist.as<unsigned>());
        else
            hipLaunchKernelGGL(k_sort_hist<8>, dim3(blocks), dim3(kSortThreads), 0, C->stream,
                    C->keys[cur].as<unsigned>(), d_n, shift, C->hist.as<unsigned>());
        DevCount hc{d_n, (unsigned long long)kSortChunk, radix, 0};
        TRY(scan_u32(C, C->hist.as<unsigned>(), C->hist_off.as<unsigned>(), hc, radix * nchunk,
                sc + SC_SCRATCH0, reinterpret_cast<unsigned*>(sc + SC_SCRATCH0 + 1), 0));
        if (bits == 9)
            hipLaunchKernelGGL(k_sort_scatter<9>, dim3(blocks), dim3(kSortThreads), 0, C->stream,
                    C->keys[cur].as<unsigned>(), C->vals[cur].as<unsigned>(), d_n, shift,
                    C->hist_off.as<unsigned>(), C->keys[cur ^ 1].as<unsigned>(),
                    C->vals[cur ^ 1].as<unsigned>());
        else
            hipLaunchKernelGGL(k_sort_scatter<8>, dim3(blocks), dim3(kSortThreads), 0, C->stream,
                    C->keys[cur].as<unsigned>(), C->vals[cur].as<unsigned>(), d_n, shift,
                    C->hist_off.as<unsigned>(), C->keys[cur ^ 1].as<unsigned>(),
                    C->vals[cur ^ 1].as<unsigned>());
        HIP_TRY(hipGetLastError());
        cur ^= 1;
    }
    *cur_out = cur;
    return MI355_OK;
}

// ---------------------------------------------------------------------------
// One pass of the pipeline over num_jobs jobs.
// ---------------------------------------------------------------------------
// TI: precision of the caller's inputs and of selection; TA: accumulation.
template <typename TI, typename TA>
struct JobHost
{
    unsigned long long vis_offset, num_points;
    TI cell_size_rad, w_scale;
    int grid_size;
    typename V2<TA>::t* grid;
    double* norm;
    unsigned long long* skipped;
};

// Visibilities with non-finite weighted amplitudes (counted by selection),
// gridded after the tiled path on the same stream.
template <typename T>
int grid_nonfinite(mi355_ctx* C, unsigned long long count, int num_jobs,
        unsigned long long n_total)
{
    if (!count) return MI355_OK;
    typedef typename V2<T>::t T2;
    typedef typename V4<T>::t T4;
    hipLaunchKernelGGL(k_grid_nonfinite<T>, dim3(grid_blocks(n_total, 4, 65535)), dim3(256), 0,
            C->stream, C->tables.as<T2>(), C->meta.as<int4>(), C->amp.as<T4>(),
            C->jobs_d.as<JobDev<T>>(), C->job_vis_off_d.as<unsigned long long>(), num_jobs,
            n_total, C->job_out_d.as<unsigned long long>());
    HIP_TRY(hipGetLastError());
    return MI355_OK;
}

// The outputs besides the grids (norm, num_skipped), after every other
// kernel of the pass has been launched.
template <typename T>
int commit_jobs(mi355_ctx* C, int num_jobs)
{
    hipLaunchKernelGGL(k_job_commit<T>, dim3(grid_blocks(num_jobs, 64, 1024)), dim3(64), 0,
            C->stream, C->jobs_d.as<JobDev<T>>(), num_jobs, C->job_norm_d.as<double>(),
            C->job_out_d.as<unsigned long long>());
    HIP_TRY(hipGetLastError());
    return MI355_OK;
}

// Called once the gridded bounding box of each job is known (before the
// grid is touched); the host drop-in uploads the caller's grid window here.
struct BeforeGrid
{
    virtual int operator()(mi355_ctx* C, const int4* bbox) = 0;
    virtual ~BeforeGrid() {}
};

template <typename TI, typename TA>
int run_pipeline(mi355_ctx* C, size_t P, const int* support, int ov, int csh,
        const typename V2<TI>::t* conv, int num_jobs, const JobHost<TI, TA>* jobs,
        const TI* uu, const TI* vv, const TI* ww, const TI* vis, const TI* weight,
        BeforeGrid* before_grid)
{
    typedef TA T;
    typedef typename V2<T>::t T2;
    typedef typename V4<T>::t T4;
    C->stats = mi355_stats();
    C->stats_in_graph = false;
    C->stats_in_noalloc = false;
    C->bk_valid = false;
    record(C, EV_START);
    TRY((prepare_tables<TI, TA>(C, P, support, ov, csh, conv)));
    record(C, EV_PREP);
    const int half = (ov + 1) / 2, noff = half + 1;
    // Gridding kernel and bucketing geometry: FP32 uses the strip kernel on
    // 256 x 128 regions (compensated: 192 x 64) unless another kernel is
    // asked for; the 64 x 64 kernels take every other case.
    const bool comp = C->compensated;
    const bool strips = std::is_same<T, float>::value && C->grid_kernel == 0;
    // (the direct 64 x 64 kernel launches one workgroup per item: not in the
    // host-sync-free mode, which takes the staged one instead)
    const bool direct64 = !strips && C->grid_kernel == 1 && !C->sync_free;
    const int tw = strips ? (comp ? StripComp::kRegionW : StripPlain::kRegionW) : kTileW;
    const int th = strips ? (comp ? StripComp::kRegionH : StripPlain::kRegionH) : kTileH;

    // Jobs.
    C->h_jobs.resize(sizeof(JobDev<T>) * num_jobs);
    JobDev<T>* hj = reinterpret_cast<JobDev<T>*>(C->h_jobs.data());
    C->h_job_chunk_off.assign(num_jobs + 1, 0);
    C->h_job_tile_base.assign(num_jobs + 1, 0);
    C->h_job_vis_off.assign(num_jobs + 1, 0);
    unsigned long long chunks = 0, tiles = 0, n_total = 0;
    for (int j = 0; j < num_jobs; ++j)
    {
        const JobHost<TI, TA>& J = jobs[j];
        TRY(check_grid_args(J.grid_size));
        if (J.vis_offset != n_total)
            return fail(MI355_ERR_INVALID_ARG, "job %d: visibilities must be contiguous and in job order", j);
        JobDev<T> d;
        d.grid = J.grid;
        d.norm = J.norm;
        d.skipped = J.skipped;
        // grid_scale = grid_size * cell_size_rad in the input precision
        // (oskar_grid_wproj.c:148 for float, :59 for double).
        const TI gs = J.grid_size * J.cell_size_rad;
        d.grid_scale_f = (float)gs;
        d.w_scale_f = (float)J.w_scale;
        d.grid_scale_d = (double)gs;
        d.w_scale_d = (double)J.w_scale;
        d.grid_size = J.grid_size;
        d.ntx = (J.grid_size + tw - 1) / tw;
        d.nty = (J.grid_size + th - 1) / th;
        d.tile_base = (unsigned)tiles;
        d.vis_begin = n_total;
        d.vis_end = n_total + J.num_points;
        d.chunk_base = (unsigned)chunks;
        hj[j] = d;
        C->h_job_chunk_off[j] = (unsigned)chunks;
        C->h_job_tile_base[j] = (unsigned)tiles;
        C->h_job_vis_off[j] = n_total;
        chunks += (J.num_points + kSelChunk - 1) / kSelChunk;
        tiles += (unsigned long long)d.ntx * d.nty;
        n_total += J.num_points;
        if (tiles > 0xFFFFFFFFull || chunks > 0xFFFFFFFFull)
            return fail(MI355_ERR_TOO_LARGE, "too many tiles or visibilities in one call");
    }
    C->h_job_chunk_off[num_jobs] = (unsigned)chunks;
    C->h_job_tile_base[num_jobs] = (unsigned)tiles;
    C->h_job_vis_off[num_jobs] = n_total;
    if (n_total >= 0xFFFFFFFFull)
        return fail(MI355_ERR_TOO_LARGE, "more than 2^32-1 visibilities in one call");
    const unsigned num_tiles = (unsigned)tiles;
    const int tile_bits = std::max(1, bits_for(tiles));
    if (tile_bits > 32) return fail(MI355_ERR_TOO_LARGE, "tile id exceeds 32 bits");
    const unsigned long long num_tables = (unsigned long long)P * noff * noff;
    // Sort key: tile, then the table id as the visit records hold it, then
    // (64x64 staged kernels, which stage a table already mirrored) the two
    // orientation bits; the strip kernel mirrors when reading LDS, so its
    // runs span orientations.  Low bits are dropped if the key exceeds 32
    // bits.
    const int orient_bits = strips ? 0 : 2;
    const int table_bits_full = bits_for(num_tables);
    const int table_id_shift = std::max(0, table_bits_full - kTableIdBits);
    const int kappa_bits = table_bits_full - table_id_shift + orient_bits;
    const int key_table_bits = std::min(kappa_bits, 32 - tile_bits);
    const int key_table_shift = kappa_bits - key_table_bits;

    TRY(ensure(C, C->jobs_d, C->h_jobs.size()));
    TRY(ensure(C, C->job_chunk_off_d, (num_jobs + 1) * sizeof(unsigned)));
    TRY(ensure(C, C->job_tile_base_d, (num_jobs + 1) * sizeof(unsigned)));
    TRY(ensure(C, C->job_vis_off_d, (num_jobs + 1) * sizeof(unsigned long long)));
    TRY(ensure(C, C->job_out_d, 3 * num_jobs * sizeof(unsigned long long)));
    TRY(ensure(C, C->job_bbox_d, num_jobs * sizeof(int4)));
    TRY(ensure(C, C->job_norm_d, num_jobs * sizeof(double)));
    if (!C->capturing)
    {
        HIP_TRY(hipMemcpyAsync(C->jobs_d.p, C->h_jobs.data(), C->h_jobs.size(), hipMemcpyHostToDevice, C->stream));
        HIP_TRY(hipMemcpyAsync(C->job_chunk_off_d.p, C->h_job_chunk_off.data(), (num_jobs + 1) * sizeof(unsigned), hipMemcpyHostToDevice, C->stream));
        HIP_TRY(hipMemcpyAsync(C->job_tile_base_d.p, C->h_job_tile_base.data(), (num_jobs + 1) * sizeof(unsigned), hipMemcpyHostToDevice, C->stream));
        HIP_TRY(hipMemcpyAsync(C->job_vis_off_d.p, C->h_job_vis_off.data(), (num_jobs + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, C->stream));
    }

    // Selection.
    TRY(ensure(C, C->meta, n_total * sizeof(int4)));
    TRY(ensure(C, C->amp, n_total * sizeof(T4)));
    if (strips) TRY(ensure(C, C->vrec, n_total * sizeof(VisRec)));
    TRY(ensure(C, C->num_visits, n_total * sizeof(unsigned)));
    TRY(ensure(C, C->visit_off, (n_total + 1) * sizeof(unsigned)));
    TRY(ensure(C, C->part_norm, chunks * sizeof(double)));
    TRY(ensure(C, C->part_skipped, chunks * sizeof(unsigned long long)));
    TRY(ensure(C, C->part_taps, chunks * sizeof(unsigned long long)));
    TRY(ensure(C, C->part_nonfinite, chunks * sizeof(unsigned long long)));
    TRY(ensure(C, C->part_bbox, chunks * sizeof(int4)));
    TRY(ensure(C, C->scalars, SC_COUNT * sizeof(unsigned long long)));
    unsigned long long* sc = C->scalars.as<unsigned long long>();
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, C->stream,
            reinterpret_cast<unsigned*>(sc), (unsigned long long)(2 * SC_COUNT), 0u);

    SelectArgs<TI, TA> sa{};
    sa.uu = uu; sa.vv = vv; sa.ww = ww; sa.vis = vis; sa.weight = weight;
    sa.jobs = C->jobs_d.as<JobDev<T>>(); sa.num_jobs = num_jobs;
    sa.job_chunk_off = C->job_chunk_off_d.as<unsigned>();
    sa.num_planes = (int)P; sa.support = C->support_d.as<int>();
    sa.oversample = ov; sa.half_ov = half;
    sa.norm_table = C->norm_table.as<double>();
    sa.plane_off = C->plane_off_d.as<unsigned>();
    sa.table_id_shift = table_id_shift;
    sa.tw = tw; sa.th = th;
    sa.meta = C->meta.as<int4>(); sa.amp = C->amp.as<T4>();
    sa.vrec = strips ? C->vrec.as<VisRec>() : nullptr;
    sa.num_visits = C->num_visits.as<unsigned>();
    sa.part_norm = C->part_norm.as<double>();
    sa.part_skipped = C->part_skipped.as<unsigned long long>();
    sa.part_taps = C->part_taps.as<unsigned long long>();
    sa.part_nonfinite = C->part_nonfinite.as<unsigned long long>();
    sa.part_bbox = C->part_bbox.as<int4>();
    if (chunks)
    {
        hipLaunchKernelGGL((k_select<TI, TA>), dim3(grid_blocks(chunks, 1, 65535)), dim3(kSelThreads), 0,
                C->stream, sa, (unsigned)chunks);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_job_reduce<T>, dim3(num_jobs), dim3(kSelThreads), 0, C->stream,
            C->jobs_d.as<JobDev<T>>(), C->job_chunk_off_d.as<unsigned>(),
            C->part_norm.as<double>(), C->part_skipped.as<unsigned long long>(),
            C->part_taps.as<unsigned long long>(), C->part_nonfinite.as<unsigned long long>(),
            C->part_bbox.as<int4>(), C->job_out_d.as<unsigned long long>(),
            C->job_bbox_d.as<int4>(), C->job_norm_d.as<double>());
    HIP_TRY(hipGetLastError());
    record(C, EV_SELECT);

    // Visit offsets and total.
    {
        DevCount cnt{nullptr, 1, 0, n_total};
        TRY(scan_u32(C, C->num_visits.as<unsigned>(), C->visit_off.as<unsigned>(), cnt,
                n_total, sc + SC_VISITS, reinterpret_cast<unsigned*>(sc + SC_OVERFLOW), 1));
    }
    unsigned long long host_sc[SC_COUNT];
    C->h_job_out.resize(3 * num_jobs);
    C->h_job_bbox.resize(num_jobs);
    // T_vis sizes the buffers from here on: the exact count of tile visits,
    // or (host-sync-free) an upper bound from the largest support, the
    // kernels reading the exact count on the device.
    unsigned long long T_vis = 0, num_nonfinite = 0;
    C->stats.num_points = n_total;
    C->stats_deferred = C->sync_free;
    if (C->sync_free)
    {
        if (before_grid) return fail(MI355_ERR_INVALID_ARG, "host window upload needs host syncs");
        int smax = 0;
        for (size_t p = 0; p < P; ++p) smax = std::max(smax, support[p]);
        const unsigned long long per = (unsigned long long)(2 * smax / tw + 2) * (2 * smax / th + 2);
        T_vis = std::max(1ull, n_total * per);
        if (T_vis > 0xFFFFFFFFull)
            return fail(MI355_ERR_TOO_LARGE, "tile-visit bound exceeds 2^32-1 (split the call)");
        num_nonfinite = 1;            // (not known: the pass always runs)
        C->deferred_points = n_total;
        C->deferred_jobs = num_jobs;
    }
    else
    {
        HIP_TRY(hipMemcpyAsync(host_sc, sc, sizeof(host_sc), hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipMemcpyAsync(C->h_job_out.data(), C->job_out_d.p, 3 * num_jobs * sizeof(unsigned long long), hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipMemcpyAsync(C->h_job_bbox.data(), C->job_bbox_d.p, num_jobs * sizeof(int4), hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipStreamSynchronize(C->stream));
        if (host_sc[SC_OVERFLOW]) return fail(MI355_ERR_TOO_LARGE, "more than 2^32-1 tile visits in one call");
        T_vis = host_sc[SC_VISITS];
        for (int j = 0; j < num_jobs; ++j)
        {
            C->stats.num_skipped += C->h_job_out[3 * j];
            C->stats.num_taps += C->h_job_out[3 * j + 1];
            num_nonfinite += C->h_job_out[3 * j + 2];
        }
        C->stats.num_visits = T_vis;
        if (before_grid) TRY((*before_grid)(C, C->h_job_bbox.data()));
        if (T_vis == 0)
        {
            record(C, EV_BUCKET); record(C, EV_GRID); record(C, EV_MERGE);
            TRY(grid_nonfinite<T>(C, num_nonfinite, num_jobs, n_total));
            return commit_jobs<T>(C, num_jobs);
        }
    }

    // Visits per work item: the option, or automatic.
    const unsigned vpi = C->visits_per_item ? C->visits_per_item
                                            : auto_visits_per_item(n_total, C->num_cu);
    C->stats.visits_per_item = vpi;

    // Emit (tile | table) keys and sort them.
    for (int b = 0; b < 2; ++b)
    {
        TRY(ensure(C, C->keys[b], T_vis * sizeof(unsigned)));
        TRY(ensure(C, C->vals[b], T_vis * sizeof(unsigned)));
    }
    hipLaunchKernelGGL(k_emit<T>, dim3(grid_blocks(n_total, 256, 16384)), dim3(256), 0, C->stream,
            C->meta.as<int4>(), C->num_visits.as<unsigned>(), C->visit_off.as<unsigned>(),
            C->jobs_d.as<JobDev<T>>(), C->job_vis_off_d.as<unsigned long long>(), num_jobs,
            n_total, tw, th, key_table_shift, key_table_bits, orient_bits, C->keys[0].as<unsigned>(),
            C->vals[0].as<unsigned>());
    HIP_TRY(hipGetLastError());
    const int key_bits = tile_bits + key_table_bits;
    int cur = 0;
    TRY(radix_sort(C, T_vis, sc + SC_VISITS, key_bits, &cur));
    const unsigned* skeys = C->keys[cur].as<unsigned>();
    // (for mi355_ctx_get_buckets: where this pass's sorted lists are)
    C->bk_valid = strips && !C->sync_free;
    C->bk_cur = cur;
    C->bk_count = T_vis;
    C->bk_key_table_bits = key_table_bits;
    const unsigned* svals = C->vals[cur].as<unsigned>();

    // Runs of visits sharing a tile, a kernel table and an orientation (the
    // LDS-staging kernels): flags -> exclusive scan -> per-run start, packed
    // word and table base.  Flags and scan live in the sort's free ping-pong
    // buffers.
    unsigned* run_flag = C->keys[cur ^ 1].as<unsigned>();
    unsigned* run_excl = C->vals[cur ^ 1].as<unsigned>();
    if (!direct64)
    {
        const int exact_key = table_id_shift == 0 && key_table_shift == 0;
        const unsigned vb = grid_blocks(T_vis, 256, 65535);
        hipLaunchKernelGGL(k_run_flags, dim3(vb), dim3(256), 0, C->stream, skeys, svals,
                C->meta.as<int4>(), sc + SC_VISITS, exact_key, orient_bits ? kOrientMask : 0,
                run_flag);
        DevCount rc{sc + SC_VISITS, 1, 1, 0};
        TRY(scan_u32(C, run_flag, run_excl, rc, T_vis, sc + SC_RUNS,
                reinterpret_cast<unsigned*>(sc + SC_OVERFLOW), 0));
        TRY(ensure(C, C->run_start, (T_vis + 1) * sizeof(unsigned)));
        if (!strips)   // (the per-run words are for the 64x64 staged kernel)
        {
            TRY(ensure(C, C->run_z, (T_vis + 1) * sizeof(int)));
            TRY(ensure(C, C->run_tb, (T_vis + 1) * sizeof(int)));
        }
        hipLaunchKernelGGL(k_run_starts, dim3(vb), dim3(256), 0, C->stream, run_flag, run_excl,
                svals, C->meta.as<int4>(), sc + SC_VISITS, sc + SC_RUNS,
                C->run_start.as<unsigned>(), strips ? nullptr : C->run_z.as<int>(),
                strips ? nullptr : C->run_tb.as<int>());
        HIP_TRY(hipGetLastError());
    }
    if (strips)
    {
        TRY(ensure(C, C->vmeta, T_vis * sizeof(int4)));
        TRY(ensure(C, C->vamp, T_vis * sizeof(T2)));
        TRY(ensure(C, C->vlds, T_vis * sizeof(int)));
        hipLaunchKernelGGL(k_gather_visits<T>, dim3(grid_blocks(T_vis, 256, 65535)), dim3(256), 0,
                C->stream, svals, C->vrec.as<VisRec>(), run_flag, run_excl,
                C->run_start.as<unsigned>(), stage_thresholds(C), kStripMaxW, kStripS, skeys,
                key_table_bits, C->jobs_d.as<JobDev<T>>(), C->job_tile_base_d.as<unsigned>(),
                num_jobs, th, table_id_shift == 0 ? noff : 0, ov % 2 == 0 ? ov / 2 : -1,
                sc + SC_VISITS, C->vmeta.as<int4>(), C->vamp.as<T2>(), C->vlds.as<int>());
        HIP_TRY(hipGetLastError());
    }

    // Tiles -> work items.
    TRY(ensure(C, C->tile_start, num_tiles * sizeof(unsigned)));
    TRY(ensure(C, C->tile_end, num_tiles * sizeof(unsigned)));
    TRY(ensure(C, C->n_items, num_tiles * sizeof(unsigned)));
    TRY(ensure(C, C->n_slots, num_tiles * sizeof(unsigned)));
    TRY(ensure(C, C->item_off, (num_tiles + 1) * sizeof(unsigned)));
    TRY(ensure(C, C->slot_off, (num_tiles + 1) * sizeof(unsigned)));
    const unsigned tb = grid_blocks(num_tiles, 256, 16384);
    hipLaunchKernelGGL(k_fill_u32, dim3(tb), dim3(256), 0, C->stream, C->tile_start.as<unsigned>(), (unsigned long long)num_tiles, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(tb), dim3(256), 0, C->stream, C->tile_end.as<unsigned>(), (unsigned long long)num_tiles, 0u);
    hipLaunchKernelGGL(k_tile_bounds, dim3(grid_blocks(T_vis, 256, 65535)), dim3(256), 0, C->stream,
            skeys, sc + SC_VISITS, key_table_bits, C->tile_start.as<unsigned>(), C->tile_end.as<unsigned>());
    hipLaunchKernelGGL(k_items, dim3(tb), dim3(256), 0, C->stream, C->tile_start.as<unsigned>(),
            C->tile_end.as<unsigned>(), num_tiles, vpi, C->n_items.as<unsigned>(),
            C->n_slots.as<unsigned>());
    HIP_TRY(hipGetLastError());
    {
        DevCount cnt{nullptr, 1, 0, num_tiles};
        TRY(scan_u32(C, C->n_items.as<unsigned>(), C->item_off.as<unsigned>(), cnt, num_tiles,
                sc + SC_ITEMS, reinterpret_cast<unsigned*>(sc + SC_OVERFLOW), 1));
        TRY(scan_u32(C, C->n_slots.as<unsigned>(), C->slot_off.as<unsigned>(), cnt, num_tiles,
                sc + SC_SLOTS, reinterpret_cast<unsigned*>(sc + SC_OVERFLOW), 1));
    }
    unsigned long long num_items, num_slots;
    if (C->sync_free)
    {
        // a tile of c visits has ceil(c / V) items, and slots only when that
        // is 2 or more, i.e. fewer than 2c / V
        const unsigned long long V = vpi;
        num_items = std::min<unsigned long long>(num_tiles, T_vis) + T_vis / V + 1;
        num_slots = 2 * T_vis / V + 1;
    }
    else
    {
        HIP_TRY(hipMemcpyAsync(host_sc, sc, sizeof(host_sc), hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipStreamSynchronize(C->stream));
        num_items = host_sc[SC_ITEMS];
        num_slots = host_sc[SC_SLOTS];
    }
    TRY(ensure(C, C->item_tile, num_items * sizeof(unsigned)));
    hipLaunchKernelGGL(k_item_map, dim3(tb), dim3(256), 0, C->stream, C->item_off.as<unsigned>(),
            num_tiles, C->item_tile.as<unsigned>());
    HIP_TRY(hipGetLastError());
    record(C, EV_BUCKET);
    C->stats.num_items = num_items;

    // Grid.
    const size_t part_bytes = comp ? sizeof(double2) : sizeof(T2);
    TRY(ensure(C, C->scratch, num_slots * (unsigned long long)(tw * th) * part_bytes));
    GridArgs<T> ga;
    ga.tables = C->tables.as<T2>();
    ga.table_bytes = (unsigned)((C->table_elems + C->table_tail) * sizeof(T2));
    ga.zero_row_byte = (unsigned)(C->table_elems * sizeof(T2));
    ga.meta = C->meta.as<int4>();
    ga.amp = C->amp.as<T4>();
    ga.visit_vis = svals;
    ga.tile_start = C->tile_start.as<unsigned>();
    ga.tile_end = C->tile_end.as<unsigned>();
    ga.item_off = C->item_off.as<unsigned>();
    ga.slot_off = C->slot_off.as<unsigned>();
    ga.item_tile = C->item_tile.as<unsigned>();
    ga.d_num_items = reinterpret_cast<const unsigned*>(C->item_off.as<unsigned>() + num_tiles);
    ga.jobs = C->jobs_d.as<JobDev<T>>();
    ga.job_tile_base = C->job_tile_base_d.as<unsigned>();
    ga.num_jobs = num_jobs;
    ga.visits_per_item = vpi;
    ga.scratch = C->scratch.p;
    ga.run_flag = run_flag;
    ga.run_excl = run_excl;
    ga.run_start = C->run_start.as<unsigned>();
    ga.run_z = C->run_z.as<int>();
    ga.run_tb = C->run_tb.as<int>();
    ga.d_num_runs = sc + SC_RUNS;
    ga.min_stage_run = C->min_stage_run ? C->min_stage_run
                                        : stage_run_default(std::is_same<T, double>::value);
    if (strips && (T_vis + 1) * 4ull > 0xFFFFFFFFull)
        return fail(MI355_ERR_TOO_LARGE, "more than 2^30 tile visits in one call");
    ga.vmeta = C->vmeta.as<int4>();
    ga.vamp = C->vamp.as<T2>();
    ga.vlds = C->vlds.as<int>();
    ga.item_group = C->item_group;
    // the strip kernel's persistent grid takes positions from per-XCD
    // counters (zeroed with the other scalars at the start of the pass)
    ga.item_ctr = reinterpret_cast<unsigned*>(sc + SC_ITEM_CTR);
    const unsigned padded = (unsigned)((num_items + 7) & ~7ull);
    // The strip kernel: a persistent grid of one workgroup per CU that takes
    // item positions from per-XCD counters, so that uneven items balance and
    // a staged table stays in LDS for the next item.  The 64x64 kernels: one
    // workgroup per work item, in item order (consecutive items of a heavy
    // tile land on different XCDs); host-sync-free, a persistent grid of two
    // workgroups per CU walks the items the device counted.
    const unsigned wg_items = C->sync_free
            ? (unsigned)std::min<unsigned long long>(num_items, 2ull * C->num_cu)
            : (unsigned)num_items;
    if (num_items && strips)
    {
        if constexpr (std::is_same<T, float>::value)
        {
            // (a multiple of 8 workgroups: the kernel maps positions to XCDs)
            const unsigned long long blk = 8ull * C->item_group;
            const unsigned wg8 = (unsigned)std::min<unsigned long long>(
                    std::max(8u, (unsigned)C->num_cu & ~7u), (num_items + blk - 1) / blk * blk);
            if (comp)
                hipLaunchKernelGGL(k_grid_strips<StripComp>, dim3(wg8), dim3(StripComp::kBlock),
                        0, C->stream, ga);
            else
                hipLaunchKernelGGL(k_grid_strips<StripPlain>, dim3(wg8), dim3(StripPlain::kBlock),
                        0, C->stream, ga);
            HIP_TRY(hipGetLastError());
        }
    }
    else if (num_items && !direct64)
    {
        if constexpr (std::is_same<T, float>::value)
        {
            if (comp)
                hipLaunchKernelGGL((k_grid_staged<T, true>), dim3(wg_items),
                        dim3(kStageBlock), 0, C->stream, ga);
            else
                hipLaunchKernelGGL((k_grid_staged<T, false>), dim3(wg_items),
                        dim3(kStageBlock), 0, C->stream, ga);
        }
        else
            hipLaunchKernelGGL((k_grid_staged<T, false>), dim3(wg_items),
                    dim3(kStageBlock), 0, C->stream, ga);
        HIP_TRY(hipGetLastError());
    }
    else if (num_items)
    {
        if (comp)
        {
            const size_t lds = (size_t)kWaves * kRowsPerWave * 64 * 4 * sizeof(T);
            HIP_TRY(hipFuncSetAttribute((const void*)k_grid_tiles<T, true>,
                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_grid_tiles<T, true>), dim3(padded), dim3(kBlock), lds, C->stream, ga, padded);
        }
        else
            hipLaunchKernelGGL((k_grid_tiles<T, false>), dim3(padded), dim3(kBlock), 0, C->stream, ga, padded);
        HIP_TRY(hipGetLastError());
    }
    record(C, EV_GRID);
    if (num_slots)
    {
        const unsigned merge_blocks = grid_blocks(
                (unsigned long long)num_tiles * (unsigned)(tw * th / kMergeCells), 1, 1u << 20);
        if (comp)
            hipLaunchKernelGGL((k_merge<T, double2>), dim3(merge_blocks), dim3(256), 0, C->stream,
                    C->item_off.as<unsigned>(), C->slot_off.as<unsigned>(), num_tiles,
                    C->jobs_d.as<JobDev<T>>(), C->job_tile_base_d.as<unsigned>(), num_jobs,
                    tw, th, C->scratch.as<double2>());
        else
            hipLaunchKernelGGL((k_merge<T, T2>), dim3(merge_blocks), dim3(256), 0, C->stream,
                    C->item_off.as<unsigned>(), C->slot_off.as<unsigned>(), num_tiles,
                    C->jobs_d.as<JobDev<T>>(), C->job_tile_base_d.as<unsigned>(), num_jobs,
                    tw, th, C->scratch.as<T2>());
        HIP_TRY(hipGetLastError());
    }
    record(C, EV_MERGE);
    TRY(grid_nonfinite<T>(C, num_nonfinite, num_jobs, n_total));
    return commit_jobs<T>(C, num_jobs);
}

void collect_timing(mi355_ctx* C)
{
    if (!C->timing) return;
    if (hipEventSynchronize(C->ev[EV_MERGE]) != hipSuccess) { (void)hipGetLastError(); return; }
    float ms[EV_COUNT] = {};
    for (int i = 1; i < EV_COUNT; ++i)
        if (hipEventElapsedTime(&ms[i], C->ev[i - 1], C->ev[i]) != hipSuccess)
        { (void)hipGetLastError(); ms[i] = 0; }
    C->stats.ms_prepare = ms[EV_PREP];
    C->stats.ms_select = ms[EV_SELECT];
    C->stats.ms_bucket = ms[EV_BUCKET];
    C->stats.ms_grid = ms[EV_GRID];
    C->stats.ms_merge = ms[EV_MERGE];
    float tot = 0;
    if (hipEventElapsedTime(&tot, C->ev[EV_START], C->ev[EV_MERGE]) != hipSuccess)
    { (void)hipGetLastError(); tot = 0; }
    C->stats.ms_total = tot;
}

// ---------------------------------------------------------------------------
// Device-pointer gridding, one job.
// ---------------------------------------------------------------------------
template <typename TI, typename TA>
int grid_dev(mi355_ctx* C, size_t P, const int* support, int ov, int csh,
        const TI* conv, size_t N, const TI* uu, const TI* vv, const TI* ww,
        const TI* vis, const TI* weight, TI cell, TI w_scale, int G,
        unsigned long long* skipped, double* norm, TA* grid, BeforeGrid* hook)
{
    TRY(check_kernel_args(P, support, ov, csh, conv));
    TRY(check_grid_args(G));
    if (!skipped || !norm || !grid) return fail(MI355_ERR_INVALID_ARG, "NULL output pointer");
    if (N && (!uu || !vv || !ww || !vis || !weight))
        return fail(MI355_ERR_INVALID_ARG, "NULL visibility array");
    HIP_TRY(hipSetDevice(C->device));
    JobHost<TI, TA> J;
    J.vis_offset = 0; J.num_points = N;
    J.cell_size_rad = cell; J.w_scale = w_scale; J.grid_size = G;
    J.grid = reinterpret_cast<typename V2<TA>::t*>(grid);
    J.norm = norm; J.skipped = skipped;
    TRY((run_pipeline<TI, TA>(C, P, support, ov, csh,
            reinterpret_cast<const typename V2<TI>::t*>(conv), 1, &J, uu, vv, ww, vis, weight,
            hook)));
    collect_timing(C);
    return ok();
}

// Per-visibility selection only (one job): out is int32 [8][N].
template <typename TI>
int select_dev(mi355_ctx* C, size_t P, const int* support, int ov, size_t N,
        const TI* uu, const TI* vv, const TI* ww, TI cell, TI w_scale, int G, int* out)
{
    if (P < 1 || !support || ov < 1) return fail(MI355_ERR_INVALID_ARG, "bad kernel arguments");
    TRY(check_grid_args(G));
    if (N && (!uu || !vv || !ww || !out)) return fail(MI355_ERR_INVALID_ARG, "NULL pointer argument");
    if (N == 0) return ok();
    HIP_TRY(hipSetDevice(C->device));
    TRY(ensure(C, C->support_d, P * sizeof(int)));
    C->cache_support.assign(support, support + P);
    C->tables_valid = false;
    HIP_TRY(hipMemcpyAsync(C->support_d.p, C->cache_support.data(), P * sizeof(int),
            hipMemcpyHostToDevice, C->stream));
    const TI gs = G * cell;
    hipLaunchKernelGGL(k_select_out<TI>, dim3(grid_blocks(N, 256, 65535)), dim3(256), 0,
            C->stream, uu, vv, ww, (unsigned long long)N, gs, w_scale, G, ov, (int)P,
            C->support_d.as<int>(), out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(C->stream));
    return ok();
}

// ---------------------------------------------------------------------------
// Host-pointer drop-in.
// ---------------------------------------------------------------------------
// Transfers between the caller's (pageable) arrays and the device go
// through the context's two pinned bands of kStageBand bytes: the host
// threads copy band k while the DMA engine moves band k-1.
constexpr size_t kStageBand = 64u << 20;

int stage_slots(mi355_ctx* C, size_t bytes)
{
    for (int i = 0; i < 2; ++i)
    {
        if (C->stage_ev[i]) HIP_TRY(hipEventSynchronize(C->stage_ev[i]));
        else HIP_TRY(hipEventCreateWithFlags(&C->stage_ev[i], hipEventDisableTiming));
        TRY(ensure_host(C->h_stage[i], bytes));
    }
    return MI355_OK;
}

// rows x row_bytes, host (pitch hp) -> device (pitch dp), on C->stream
int stage_upload(mi355_ctx* C, void* dev, size_t dp, const void* host, size_t hp,
        size_t row_bytes, size_t rows)
{
    if (!rows || !row_bytes) return MI355_OK;
    const size_t band = std::max<size_t>(1, kStageBand / row_bytes);
    TRY(stage_slots(C, std::min(rows, band) * row_bytes));
    for (size_t k = 0, r0 = 0; r0 < rows; ++k, r0 += band)
    {
        const int sl = (int)(k & 1);
        const size_t n = std::min(band, rows - r0);
        HIP_TRY(hipEventSynchronize(C->stage_ev[sl]));     // its last DMA is done
        copy_rows(C->h_stage[sl].p, row_bytes, static_cast<const char*>(host) + r0 * hp, hp,
                  row_bytes, n);
        HIP_TRY(hipMemcpy2DAsync(static_cast<char*>(dev) + r0 * dp, dp, C->h_stage[sl].p,
                row_bytes, row_bytes, n, hipMemcpyHostToDevice, C->stream));
        HIP_TRY(hipEventRecord(C->stage_ev[sl], C->stream));
    }
    return MI355_OK;
}

int stage_upload_1d(mi355_ctx* C, void* dev, const void* host, size_t bytes)
{
    const size_t full = bytes / kStageBand, tail = bytes - full * kStageBand;
    TRY(stage_upload(C, dev, kStageBand, host, kStageBand, kStageBand, full));
    return stage_upload(C, static_cast<char*>(dev) + full * kStageBand, tail,
                        static_cast<const char*>(host) + full * kStageBand, tail, tail, 1);
}

// rows x row_bytes, device (pitch dp) -> host (pitch hp); returns when the
// host has them.  (Before band k is read into slot k&1, the host has copied
// band k-2 out of it.)
int stage_download(mi355_ctx* C, void* host, size_t hp, const void* dev, size_t dp,
        size_t row_bytes, size_t rows)
{
    if (!rows || !row_bytes) return MI355_OK;
    const size_t band = std::max<size_t>(1, kStageBand / row_bytes);
    const size_t nb = (rows + band - 1) / band;
    TRY(stage_slots(C, std::min(rows, band) * row_bytes));
    for (size_t k = 0; k <= nb; ++k)
    {
        if (k < nb)
        {
            const int sl = (int)(k & 1);
            const size_t r0 = k * band, n = std::min(band, rows - r0);
            HIP_TRY(hipMemcpy2DAsync(C->h_stage[sl].p, row_bytes,
                    static_cast<const char*>(dev) + r0 * dp, dp, row_bytes, n,
                    hipMemcpyDeviceToHost, C->stream));
            HIP_TRY(hipEventRecord(C->stage_ev[sl], C->stream));
        }
        if (k >= 1)
        {
            const size_t j = k - 1;
            const int sl = (int)(j & 1);
            const size_t r0 = j * band, n = std::min(band, rows - r0);
            HIP_TRY(hipEventSynchronize(C->stage_ev[sl]));
            copy_rows(static_cast<char*>(host) + r0 * hp, hp, C->h_stage[sl].p, row_bytes,
                      row_bytes, n);
        }
    }
    return MI355_OK;
}

int stage_download_1d(mi355_ctx* C, void* host, const void* dev, size_t bytes)
{
    const size_t full = bytes / kStageBand, tail = bytes - full * kStageBand;
    TRY(stage_download(C, host, kStageBand, dev, kStageBand, kStageBand, full));
    return stage_download(C, static_cast<char*>(host) + full * kStageBand, tail,
                          static_cast<const char*>(dev) + full * kStageBand, tail, tail, 1);
}

template <typename T>
struct UploadWindow : BeforeGrid
{
    const T* host_grid; int G; int4 box; bool any = false;
    int operator()(mi355_ctx* C, const int4* bbox) override
    {
        box = bbox[0];
        any = box.x <= box.z && box.y <= box.w;
        if (!any) return MI355_OK;
        typedef typename V2<T>::t T2;
        const size_t pitch = (size_t)G * sizeof(T2);
        const size_t off = (size_t)box.y * G + box.x;
        return stage_upload(C, C->in_grid.as<T2>() + off, pitch,
                reinterpret_cast<const T2*>(host_grid) + off, pitch,
                (size_t)(box.z - box.x + 1) * sizeof(T2), (size_t)(box.w - box.y + 1));
    }
};

template <typename T>
int grid_host(size_t P, const int* support, int ov, int csh, const T* conv,
        size_t N, const T* uu, const T* vv, const T* ww, const T* vis,
        const T* weight, T cell, T w_scale, int G, size_t* num_skipped,
        double* norm, T* grid)
{
    TRY(check_kernel_args(P, support, ov, csh, conv));
    TRY(check_grid_args(G));
    if (!num_skipped || !norm || !grid) return fail(MI355_ERR_INVALID_ARG, "NULL output pointer");
    if (N && (!uu || !vv || !ww || !vis || !weight))
        return fail(MI355_ERR_INVALID_ARG, "NULL visibility array");
    if (N == 0) { *num_skipped = 0; return ok(); }   // oskar_grid_wproj.c:151
    std::lock_guard<std::mutex> lock(g_default_lock);
    mi355_ctx* C = nullptr;
    TRY(get_default_ctx(&C));
    HIP_TRY(hipSetDevice(C->device));
    typedef typename V2<T>::t T2;
    const size_t conv_n = (size_t)P * csh * csh * 2;
    TRY(ensure(C, C->in_conv, conv_n * sizeof(T)));
    TRY(ensure(C, C->in_uu, N * sizeof(T)));
    TRY(ensure(C, C->in_vv, N * sizeof(T)));
    TRY(ensure(C, C->in_ww, N * sizeof(T)));
    TRY(ensure(C, C->in_vis, 2 * N * sizeof(T)));
    TRY(ensure(C, C->in_weight, N * sizeof(T)));
    TRY(ensure(C, C->in_grid, (size_t)G * G * sizeof(T2)));
    TRY(ensure(C, C->in_norm, sizeof(double)));
    TRY(ensure(C, C->in_skipped, sizeof(unsigned long long)));
    TRY(stage_upload_1d(C, C->in_conv.p, conv, conv_n * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_uu.p, uu, N * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_vv.p, vv, N * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_ww.p, ww, N * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_vis.p, vis, 2 * N * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_weight.p, weight, N * sizeof(T)));
    const double zero = 0.0;
    HIP_TRY(hipMemcpyAsync(C->in_norm.p, &zero, sizeof(double), hipMemcpyHostToDevice, C->stream));
    // The host conv_func may change between calls at the same address.
    C->tables_valid = false;
    UploadWindow<T> up;
    up.host_grid = grid;
    up.G = G;
    TRY((grid_dev<T, T>(C, P, support, ov, csh, C->in_conv.as<T>(), N, C->in_uu.as<T>(),
            C->in_vv.as<T>(), C->in_ww.as<T>(), C->in_vis.as<T>(), C->in_weight.as<T>(),
            cell, w_scale, G, C->in_skipped.as<unsigned long long>(), C->in_norm.as<double>(),
            C->in_grid.as<T>(), &up)));
    if (up.any)
    {
        const size_t pitch = (size_t)G * sizeof(T2);
        const size_t off = (size_t)up.box.y * G + up.box.x;
        TRY(stage_download(C, reinterpret_cast<T2*>(grid) + off, pitch,
                C->in_grid.as<T2>() + off, pitch, (size_t)(up.box.z - up.box.x + 1) * sizeof(T2),
                (size_t)(up.box.w - up.box.y + 1)));
    }
    double dnorm = 0.0;
    unsigned long long dskip = 0;
    HIP_TRY(hipMemcpyAsync(&dnorm, C->in_norm.p, sizeof(double), hipMemcpyDeviceToHost, C->stream));
    HIP_TRY(hipMemcpyAsync(&dskip, C->in_skipped.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, C->stream));
    HIP_TRY(hipStreamSynchronize(C->stream));
    *norm += dnorm;
    *num_skipped = (size_t)dskip;
    return ok();
}

// ---------------------------------------------------------------------------
// Degridding (adjoint), device pointers.
// ---------------------------------------------------------------------------
template <typename T>
int degrid_dev(mi355_ctx* C, size_t P, const int* support, int ov, int csh,
        const T* conv, size_t N, const T* uu, const T* vv, const T* ww,
        const T* weight, T cell, T w_scale, int G, const T* grid,
        int apply_weight, T* out)
{
    TRY(check_kernel_args(P, support, ov, csh, conv));
    TRY(check_grid_args(G));
    if (!grid || (N && (!out || !uu || !vv || !ww || !weight)))
        return fail(MI355_ERR_INVALID_ARG, "NULL pointer argument");
    if (N >= 0xFFFFFFFFull) return fail(MI355_ERR_TOO_LARGE, "too many visibilities");
    if (N == 0) return ok();
    HIP_TRY(hipSetDevice(C->device));
    typedef typename V2<T>::t T2;
    typedef typename V4<T>::t T4;
    TRY((prepare_tables<T, T>(C, P, support, ov, csh, reinterpret_cast<const T2*>(conv))));
    const int half = (ov + 1) / 2, noff = half + 1;
    // One job; its norm/skipped totals go to library scratch (not reported).
    TRY(ensure(C, C->in_norm, sizeof(double)));
    TRY(ensure(C, C->in_skipped, sizeof(unsigned long long)));
    C->h_jobs.resize(sizeof(JobDev<T>));
    JobDev<T>* hj = reinterpret_cast<JobDev<T>*>(C->h_jobs.data());
    hj->grid = const_cast<T2*>(reinterpret_cast<const T2*>(grid));
    hj->norm = C->in_norm.as<double>();
    hj->skipped = C->in_skipped.as<unsigned long long>();
    const T gs = G * cell;
    hj->grid_scale_f = (float)gs;
    hj->w_scale_f = (float)w_scale;
    hj->grid_scale_d = (double)gs;
    hj->w_scale_d = (double)w_scale;
    hj->grid_size = G;
    hj->ntx = (G + kTileW - 1) / kTileW;
    hj->nty = (G + kTileH - 1) / kTileH;
    hj->tile_base = 0;
    hj->vis_begin = 0;
    hj->vis_end = N;
    hj->chunk_base = 0;
    const unsigned long long chunks = (N + kSelChunk - 1) / kSelChunk;
    C->h_job_chunk_off.assign(2, 0);
    C->h_job_chunk_off[1] = (unsigned)chunks;
    TRY(ensure(C, C->jobs_d, sizeof(JobDev<T>)));
    TRY(ensure(C, C->job_chunk_off_d, 2 * sizeof(unsigned)));
    HIP_TRY(hipMemcpyAsync(C->jobs_d.p, hj, sizeof(JobDev<T>), hipMemcpyHostToDevice, C->stream));
    HIP_TRY(hipMemcpyAsync(C->job_chunk_off_d.p, C->h_job_chunk_off.data(), 2 * sizeof(unsigned), hipMemcpyHostToDevice, C->stream));
    TRY(ensure(C, C->meta, N * sizeof(int4)));
    TRY(ensure(C, C->amp, N * sizeof(T4)));
    TRY(ensure(C, C->num_visits, N * sizeof(unsigned)));
    TRY(ensure(C, C->part_norm, chunks * sizeof(double)));
    TRY(ensure(C, C->part_skipped, chunks * sizeof(unsigned long long)));
    TRY(ensure(C, C->part_taps, chunks * sizeof(unsigned long long)));
    TRY(ensure(C, C->part_nonfinite, chunks * sizeof(unsigned long long)));
    TRY(ensure(C, C->part_bbox, chunks * sizeof(int4)));
    SelectArgs<T, T> sa{};
    sa.uu = uu; sa.vv = vv; sa.ww = ww; sa.weight = weight;
    sa.vis = nullptr;      // the adjoint needs the selection only, no amplitudes
    sa.tw = kTileW; sa.th = kTileH;   // (tile-visit counts are computed, not used)
    sa.jobs = C->jobs_d.as<JobDev<T>>(); sa.num_jobs = 1;
    sa.job_chunk_off = C->job_chunk_off_d.as<unsigned>();
    sa.num_planes = (int)P; sa.support = C->support_d.as<int>();
    sa.oversample = ov; sa.half_ov = half;
    sa.norm_table = C->norm_table.as<double>();
    sa.plane_off = C->plane_off_d.as<unsigned>();
    sa.table_id_shift = std::max(0, bits_for((unsigned long long)P * noff * noff) - kTableIdBits);
    sa.meta = C->meta.as<int4>(); sa.amp = C->amp.as<T4>();
    sa.num_visits = C->num_visits.as<unsigned>();
    sa.part_norm = C->part_norm.as<double>();
    sa.part_skipped = C->part_skipped.as<unsigned long long>();
    sa.part_taps = C->part_taps.as<unsigned long long>();
    sa.part_nonfinite = C->part_nonfinite.as<unsigned long long>();
    sa.part_bbox = C->part_bbox.as<int4>();
    hipLaunchKernelGGL((k_select<T, T>), dim3(grid_blocks(chunks, 1, 65535)), dim3(kSelThreads), 0,
            C->stream, sa, (unsigned)chunks);
    HIP_TRY(hipGetLastError());
    // Order by the 64 x 64 grid tile of the centre cell (a stable radix
    // sort of N keys), so that the waves running together read neighbouring
    // grid cells; small calls keep index order.
    const unsigned* order = nullptr;
    if (N >= kDegridSortMin)
    {
        const int tiles_x = (G + 63) / 64;
        const unsigned long long ntile = (unsigned long long)tiles_x * tiles_x;
        TRY(ensure(C, C->scalars, SC_COUNT * sizeof(unsigned long long)));
        unsigned long long* sc = C->scalars.as<unsigned long long>();
        for (int b = 0; b < 2; ++b)
        {
            TRY(ensure(C, C->keys[b], N * sizeof(unsigned)));
            TRY(ensure(C, C->vals[b], N * sizeof(unsigned)));
        }
        // the sort's count: N in sc[SC_VISITS] (two 32-bit words)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, C->stream,
                reinterpret_cast<unsigned*>(sc + SC_VISITS), 1ull, (unsigned)N);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, C->stream,
                reinterpret_cast<unsigned*>(sc + SC_VISITS) + 1, 1ull, 0u);
        hipLaunchKernelGGL(k_degrid_keys, dim3(grid_blocks(N, 256, 16384)), dim3(256), 0, C->stream,
                C->meta.as<int4>(), (unsigned long long)N, tiles_x, (unsigned)ntile,
                C->keys[0].as<unsigned>(), C->vals[0].as<unsigned>());
        HIP_TRY(hipGetLastError());
        int cur = 0;
        TRY(radix_sort(C, N, sc + SC_VISITS, bits_for(ntile + 1), &cur));
        order = C->vals[cur].as<unsigned>();
    }
    hipLaunchKernelGGL(k_degrid<T>, dim3(grid_blocks(N, 4, 65535)), dim3(256), 0, C->stream,
            C->tables.as<T2>(), C->meta.as<int4>(), order, weight, (unsigned long long)N,
            reinterpret_cast<const T2*>(grid), G, apply_weight, reinterpret_cast<T2*>(out));
    HIP_TRY(hipGetLastError());
    return ok();
}

// Job table of the batched entry points -> per-job host descriptions.
int batched_jobs(size_t P, const int* support, int ov, int csh, const float* conv,
        int num_jobs, const mi355_job_f* jobs, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight,
        std::vector<JobHost<float, float>>& J)
{
    TRY(check_kernel_args(P, support, ov, csh, conv));
    if (num_jobs < 1 || !jobs) return fail(MI355_ERR_INVALID_ARG, "no jobs");
    J.assign(num_jobs, JobHost<float, float>());
    unsigned long long n = 0;
    for (int j = 0; j < num_jobs; ++j)
    {
        if (!jobs[j].grid || !jobs[j].norm || !jobs[j].num_skipped)
            return fail(MI355_ERR_INVALID_ARG, "job %d: NULL output pointer", j);
        J[j].vis_offset = jobs[j].vis_offset;
        J[j].num_points = jobs[j].num_points;
        J[j].cell_size_rad = jobs[j].cell_size_rad;
        J[j].w_scale = jobs[j].w_scale;
        J[j].grid_size = jobs[j].grid_size;
        J[j].grid = reinterpret_cast<float2*>(jobs[j].grid);
        J[j].norm = jobs[j].norm;
        J[j].skipped = jobs[j].num_skipped;
        n += jobs[j].num_points;
    }
    if (n && (!uu || !vv || !ww || !vis || !weight))
        return fail(MI355_ERR_INVALID_ARG, "NULL visibility array");
    // Each job's outputs are written by plain read-modify-writes from many
    // workgroups at once, so no two jobs may share a grid cell, a norm or a
    // count: sort the output ranges by address and check neighbours.
    struct Range { uintptr_t lo, hi; int job; };
    for (int kind = 0; kind < 3; ++kind)
    {
        std::vector<Range> r((size_t)num_jobs);
        for (int j = 0; j < num_jobs; ++j)
        {
            const uintptr_t p = kind == 0 ? (uintptr_t)jobs[j].grid
                    : kind == 1 ? (uintptr_t)jobs[j].norm : (uintptr_t)jobs[j].num_skipped;
            const size_t bytes = kind == 0 ? (size_t)std::max(jobs[j].grid_size, 0) *
                    (size_t)std::max(jobs[j].grid_size, 0) * sizeof(float2)
                    : kind == 1 ? sizeof(double) : sizeof(unsigned long long);
            r[j] = {p, p + std::max<size_t>(bytes, 1), j};
        }
        std::sort(r.begin(), r.end(), [](const Range& a, const Range& b) { return a.lo < b.lo; });
        for (size_t k = 1; k < r.size(); ++k)
            if (r[k].lo < r[k - 1].hi)
                return fail(MI355_ERR_INVALID_ARG, "jobs %d and %d share their %s", r[k - 1].job,
                        r[k].job, kind == 0 ? "grid" : kind == 1 ? "norm" : "num_skipped");
    }
    return MI355_OK;
}

// ---------------------------------------------------------------------------
// Uniform-weighting pre-pass (oskar_grid_weights.c)
// ---------------------------------------------------------------------------
int check_weights_args(size_t N, int G)
{
    if (G < 1) return fail(MI355_ERR_INVALID_ARG, "grid_size must be >= 1");
    // cell indices v * G + u and the off-grid key G^2 fit 32 bits
    if (G > 65535) return fail(MI355_ERR_TOO_LARGE, "grid_size > 65535 for the weighting pre-pass");
    if (N > 0xFFFFFFFFull) return fail(MI355_ERR_TOO_LARGE, "more than 2^32-1 points in one call");
    return MI355_OK;
}

// Cell keys of the N points, stably sorted (point order kept within a cell);
// counts points off the grid into *skipped and the cells' bounding box into
// box (when given).  Expects k_weight_init to have set SC_W_COUNT to N.
template <typename T>
int weights_sorted_cells(mi355_ctx* C, size_t N, const T* uu, const T* vv, T gs, int G,
        unsigned long long* skipped, int* box, int* cur)
{
    for (int b = 0; b < 2; ++b)
    {
        TRY(ensure(C, C->keys[b], N * sizeof(unsigned)));
        TRY(ensure(C, C->vals[b], N * sizeof(unsigned)));
    }
    const unsigned long long cells = (unsigned long long)G * G;
    hipLaunchKernelGGL(k_weight_cells<T>, dim3(grid_blocks(N, 256, 16384)), dim3(256), 0,
            C->stream, uu, vv, (unsigned long long)N, gs, G, (unsigned)cells,
            C->keys[0].as<unsigned>(), C->vals[0].as<unsigned>(), skipped, box);
    HIP_TRY(hipGetLastError());
    return radix_sort(C, N, C->scalars.as<unsigned long long>() + SC_W_COUNT,
            bits_for(cells + 1), cur);
}

template <typename T>
int weights_write_dev(mi355_ctx* C, size_t N, const T* uu, const T* vv, const T* weight,
        T cell, int G, unsigned long long* skipped, T* grid)
{
    TRY(check_weights_args(N, G));
    if (!skipped || !grid) return fail(MI355_ERR_INVALID_ARG, "NULL output pointer");
    if (N && (!uu || !vv || !weight)) return fail(MI355_ERR_INVALID_ARG, "NULL input array");
    HIP_TRY(hipSetDevice(C->device));
    TRY(ensure(C, C->scalars, SC_COUNT * sizeof(unsigned long long)));
    unsigned long long* sc = C->scalars.as<unsigned long long>();
    hipLaunchKernelGGL(k_weight_init, dim3(1), dim3(64), 0, C->stream, sc + SC_W_COUNT,
            (unsigned long long)N, skipped, (int*)nullptr);
    HIP_TRY(hipGetLastError());
    if (N == 0) return ok();
    const T gs = (T)G * cell;     // oskar_grid_weights.c:113 / :44
    int cur = 0;
    TRY(weights_sorted_cells<T>(C, N, uu, vv, gs, G, skipped, nullptr, &cur));
    hipLaunchKernelGGL(k_weight_sum<T>, dim3(grid_blocks(N, 256, 65535)), dim3(256), 0,
            C->stream, C->keys[cur].as<unsigned>(), C->vals[cur].as<unsigned>(), weight,
            (unsigned long long)N, (unsigned)((unsigned long long)G * G), G, grid, 0, 0,
            (unsigned long long)G);
    HIP_TRY(hipGetLastError());
    return ok();
}

template <typename T>
int weights_read_dev(mi355_ctx* C, size_t N, const T* uu, const T* vv, const T* weight_in,
        T* weight_out, T cell, int G, unsigned long long* skipped, const T* grid)
{
    TRY(check_weights_args(N, G));
    if (!skipped || !grid) return fail(MI355_ERR_INVALID_ARG, "NULL pointer argument");
    if (N && (!uu || !vv || !weight_in || !weight_out))
        return fail(MI355_ERR_INVALID_ARG, "NULL point array");
    HIP_TRY(hipSetDevice(C->device));
    hipLaunchKernelGGL(k_weight_init, dim3(1), dim3(64), 0, C->stream,
            (unsigned long long*)nullptr, 0ull, skipped, (int*)nullptr);
    if (N)
    {
        const T gs = (T)G * cell;     // oskar_grid_weights.c:146 / :77
        hipLaunchKernelGGL(k_weight_read<T>, dim3(grid_blocks(N, 256, 16384)), dim3(256), 0,
                C->stream, uu, vv, weight_in, weight_out, (unsigned long long)N, gs, G, grid, 0, 0,
                (unsigned long long)G, skipped);
    }
    HIP_TRY(hipGetLastError());
    return ok();
}

// Host-pointer drop-ins: the points go to the device, and only the window of
// the weights grid that the points' cells span travels each way.
template <typename T>
int upload_weight_window(mi355_ctx* C, const T* grid, int G, const int* box)
{
    const size_t ww = (size_t)(box[2] - box[0] + 1), wh = (size_t)(box[3] - box[1] + 1);
    TRY(ensure(C, C->in_grid, ww * wh * sizeof(T)));
    return stage_upload(C, C->in_grid.p, ww * sizeof(T), grid + (size_t)box[1] * G + box[0],
            (size_t)G * sizeof(T), ww * sizeof(T), wh);
}

template <typename T>
int weights_write_host(size_t N, const T* uu, const T* vv, const T* weight, T cell, int G,
        size_t* num_skipped, T* grid)
{
    TRY(check_weights_args(N, G));
    if (!num_skipped || !grid) return fail(MI355_ERR_INVALID_ARG, "NULL output pointer");
    if (N && (!uu || !vv || !weight)) return fail(MI355_ERR_INVALID_ARG, "NULL input array");
    if (N == 0) { *num_skipped = 0; return ok(); }   // oskar_grid_weights.c:116
    std::lock_guard<std::mutex> lock(g_default_lock);
    mi355_ctx* C = nullptr;
    TRY(get_default_ctx(&C));
    HIP_TRY(hipSetDevice(C->device));
    TRY(ensure(C, C->in_uu, N * sizeof(T)));
    TRY(ensure(C, C->in_vv, N * sizeof(T)));
    TRY(ensure(C, C->in_weight, N * sizeof(T)));
    TRY(ensure(C, C->in_skipped, sizeof(unsigned long long)));
    TRY(ensure(C, C->scalars, SC_COUNT * sizeof(unsigned long long)));
    TRY(stage_upload_1d(C, C->in_uu.p, uu, N * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_vv.p, vv, N * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_weight.p, weight, N * sizeof(T)));
    unsigned long long* sc = C->scalars.as<unsigned long long>();
    int* box = reinterpret_cast<int*>(sc + SC_W_BOX);
    unsigned long long* skipped = C->in_skipped.as<unsigned long long>();
    hipLaunchKernelGGL(k_weight_init, dim3(1), dim3(64), 0, C->stream, sc + SC_W_COUNT,
            (unsigned long long)N, skipped, box);
    const T gs = (T)G * cell;
    int cur = 0;
    TRY(weights_sorted_cells<T>(C, N, C->in_uu.as<T>(), C->in_vv.as<T>(), gs, G, skipped, box,
            &cur));
    int hbox[4];
    unsigned long long hskip = 0;
    HIP_TRY(hipMemcpyAsync(hbox, box, sizeof(hbox), hipMemcpyDeviceToHost, C->stream));
    HIP_TRY(hipMemcpyAsync(&hskip, skipped, sizeof(hskip), hipMemcpyDeviceToHost, C->stream));
    HIP_TRY(hipStreamSynchronize(C->stream));
    if (hbox[0] <= hbox[2])
    {
        TRY(upload_weight_window<T>(C, grid, G, hbox));
        const size_t ww = (size_t)(hbox[2] - hbox[0] + 1), wh = (size_t)(hbox[3] - hbox[1] + 1);
        hipLaunchKernelGGL(k_weight_sum<T>, dim3(grid_blocks(N, 256, 65535)), dim3(256), 0,
                C->stream, C->keys[cur].as<unsigned>(), C->vals[cur].as<unsigned>(),
                C->in_weight.as<T>(), (unsigned long long)N,
                (unsigned)((unsigned long long)G * G), G, C->in_grid.as<T>(), hbox[0], hbox[1],
                (unsigned long long)ww);
        HIP_TRY(hipGetLastError());
        TRY(stage_download(C, grid + (size_t)hbox[1] * G + hbox[0], (size_t)G * sizeof(T),
                C->in_grid.p, ww * sizeof(T), ww * sizeof(T), wh));
    }
    *num_skipped = (size_t)hskip;
    return ok();
}

template <typename T>
int weights_read_host(size_t N, const T* uu, const T* vv, const T* weight_in, T* weight_out,
        T cell, int G, size_t* num_skipped, const T* grid)
{
    TRY(check_weights_args(N, G));
    if (!num_skipped || !grid) return fail(MI355_ERR_INVALID_ARG, "NULL pointer argument");
    if (N && (!uu || !vv || !weight_in || !weight_out))
        return fail(MI355_ERR_INVALID_ARG, "NULL point array");
    if (N == 0) { *num_skipped = 0; return ok(); }   // oskar_grid_weights.c:149
    std::lock_guard<std::mutex> lock(g_default_lock);
    mi355_ctx* C = nullptr;
    TRY(get_default_ctx(&C));
    HIP_TRY(hipSetDevice(C->device));
    TRY(ensure(C, C->in_uu, N * sizeof(T)));
    TRY(ensure(C, C->in_vv, N * sizeof(T)));
    TRY(ensure(C, C->in_weight, N * sizeof(T)));
    TRY(ensure(C, C->in_weight_out, N * sizeof(T)));
    TRY(ensure(C, C->in_skipped, sizeof(unsigned long long)));
    TRY(ensure(C, C->scalars, SC_COUNT * sizeof(unsigned long long)));
    TRY(stage_upload_1d(C, C->in_uu.p, uu, N * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_vv.p, vv, N * sizeof(T)));
    unsigned long long* sc = C->scalars.as<unsigned long long>();
    int* box = reinterpret_cast<int*>(sc + SC_W_BOX);
    unsigned long long* skipped = C->in_skipped.as<unsigned long long>();
    hipLaunchKernelGGL(k_weight_init, dim3(1), dim3(64), 0, C->stream,
            (unsigned long long*)nullptr, 0ull, skipped, box);
    const T gs = (T)G * cell;
    hipLaunchKernelGGL(k_weight_cells<T>, dim3(grid_blocks(N, 256, 16384)), dim3(256), 0,
            C->stream, C->in_uu.as<T>(), C->in_vv.as<T>(), (unsigned long long)N, gs, G, 0u,
            (unsigned*)nullptr, (unsigned*)nullptr, skipped, box);
    HIP_TRY(hipGetLastError());
    int hbox[4];
    unsigned long long hskip = 0;
    HIP_TRY(hipMemcpyAsync(hbox, box, sizeof(hbox), hipMemcpyDeviceToHost, C->stream));
    HIP_TRY(hipMemcpyAsync(&hskip, skipped, sizeof(hskip), hipMemcpyDeviceToHost, C->stream));
    HIP_TRY(hipStreamSynchronize(C->stream));
    if (hbox[0] <= hbox[2])
    {
        TRY(stage_upload_1d(C, C->in_weight.p, weight_in, N * sizeof(T)));
        // points off the grid leave their weight_out entries as they were
        if (hskip) TRY(stage_upload_1d(C, C->in_weight_out.p, weight_out, N * sizeof(T)));
        TRY(upload_weight_window<T>(C, grid, G, hbox));
        const size_t ww = (size_t)(hbox[2] - hbox[0] + 1);
        hipLaunchKernelGGL(k_weight_read<T>, dim3(grid_blocks(N, 256, 16384)), dim3(256), 0,
                C->stream, C->in_uu.as<T>(), C->in_vv.as<T>(), C->in_weight.as<T>(),
                C->in_weight_out.as<T>(), (unsigned long long)N, gs, G, C->in_grid.as<T>(),
                hbox[0], hbox[1], (unsigned long long)ww, (unsigned long long*)nullptr);
        HIP_TRY(hipGetLastError());
        TRY(stage_download_1d(C, weight_out, C->in_weight_out.p, N * sizeof(T)));
    }
    *num_skipped = (size_t)hskip;
    return ok();
}

// ---------------------------------------------------------------------------
// Grid -> image (centred inverse 2-D DFT, real part, scaled by 1/norm)
// ---------------------------------------------------------------------------
#define FFT_TRY(expr)                                                        \
    do {                                                                     \
        hipfftResult r_ = (expr);                                            \
        if (r_ != HIPFFT_SUCCESS)                                            \
            return fail(MI355_ERR_HIP, "%s failed: hipFFT status %d (%s:%d)", \
                    #expr, (int)r_, __FILE__, __LINE__);                     \
    } while (0)

int check_image_args(int G)
{
    if (G < 2 || (G & 1)) return fail(MI355_ERR_INVALID_ARG, "grid_size must be even and >= 2");
    if (G > 65536) return fail(MI355_ERR_TOO_LARGE, "grid_size > 65536 for the image transform");
    return MI355_OK;
}

template <typename T>
int image_dev(mi355_ctx* C, int G, const T* grid, double norm, T* image)
{
    TRY(check_image_args(G));
    if (!grid || !image) return fail(MI355_ERR_INVALID_ARG, "NULL pointer argument");
    HIP_TRY(hipSetDevice(C->device));
    typedef typename V2<T>::t T2;
    const unsigned long long n = (unsigned long long)G * G;
    TRY(ensure(C, C->fft_work, n * sizeof(T2)));
    const int dbl = sizeof(T) == 8;
    if (!C->fft_valid || C->fft_G != G || C->fft_double != dbl)
    {
        if (C->fft_valid) (void)hipfftDestroy(C->fft_plan);
        C->fft_valid = false;
        FFT_TRY(hipfftPlan2d(&C->fft_plan, G, G, dbl ? HIPFFT_Z2Z : HIPFFT_C2C));
        C->fft_valid = true;
        C->fft_G = G;
        C->fft_double = dbl;
    }
    FFT_TRY(hipfftSetStream(C->fft_plan, C->stream));
    const unsigned blocks = grid_blocks(n / 2, 256, 65535);
    hipLaunchKernelGGL(k_image_pre<T>, dim3(blocks), dim3(256), 0, C->stream,
            reinterpret_cast<const T2*>(grid), C->fft_work.as<T2>(), (unsigned)G, n / 2);
    HIP_TRY(hipGetLastError());
    if (dbl)
        FFT_TRY(hipfftExecZ2Z(C->fft_plan, C->fft_work.as<hipfftDoubleComplex>(),
                C->fft_work.as<hipfftDoubleComplex>(), HIPFFT_BACKWARD));
    else
        FFT_TRY(hipfftExecC2C(C->fft_plan, C->fft_work.as<hipfftComplex>(),
                C->fft_work.as<hipfftComplex>(), HIPFFT_BACKWARD));
    const T scale = (T)(norm != 0.0 ? 1.0 / norm : 1.0);
    hipLaunchKernelGGL(k_image_post<T>, dim3(blocks), dim3(256), 0, C->stream,
            C->fft_work.as<T2>(), image, (unsigned)G, n / 2, scale);
    HIP_TRY(hipGetLastError());
    return ok();
}

template <typename T>
int image_host(int G, const T* grid, double norm, T* image)
{
    TRY(check_image_args(G));
    if (!grid || !image) return fail(MI355_ERR_INVALID_ARG, "NULL pointer argument");
    std::lock_guard<std::mutex> lock(g_default_lock);
    mi355_ctx* C = nullptr;
    TRY(get_default_ctx(&C));
    HIP_TRY(hipSetDevice(C->device));
    const size_t n = (size_t)G * G;
    TRY(ensure(C, C->in_grid, 2 * n * sizeof(T)));
    TRY(ensure(C, C->in_image, n * sizeof(T)));
    TRY(stage_upload_1d(C, C->in_grid.p, grid, 2 * n * sizeof(T)));
    TRY(image_dev<T>(C, G, C->in_grid.as<T>(), norm, C->in_image.as<T>()));
    TRY(stage_download_1d(C, image, C->in_image.p, n * sizeof(T)));
    return ok();
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

MI355_EXPORT int mi355_grid_last_error(void) { return ggs::t_last_status; }
MI355_EXPORT const char* mi355_grid_last_error_message(void) { return ggs::t_last_msg; }

MI355_EXPORT const char* mi355_grid_error_string(int status)
{
    switch (status)
    {
    case MI355_OK: return "success";
    case MI355_ERR_INVALID_ARG: return "invalid argument";
    case MI355_ERR_KERNEL_TOO_SMALL: return "conv_size_half too small for the support table";
    case MI355_ERR_NO_DEVICE: return "no HIP device";
    case MI355_ERR_HIP: return "HIP runtime error";
    case MI355_ERR_OUT_OF_MEMORY: return "device out of memory";
    case MI355_ERR_TOO_LARGE: return "problem exceeds 32-bit indexing";
    case MI355_ERR_IO: return "file I/O error";
    case MI355_ERR_FORMAT: return "file format error";
    default: return "unknown status";
    }
}

MI355_EXPORT void mi355_grid_wproj_d(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half, const double* conv_func,
        const size_t num_points, const double* uu, const double* vv,
        const double* ww, const double* vis, const double* weight,
        const double cell_size_rad, const double w_scale, const int grid_size,
        size_t* num_skipped, double* norm, double* grid)
{
    (void)grid_host<double>(num_w_planes, support, oversample, conv_size_half, conv_func,
            num_points, uu, vv, ww, vis, weight, cell_size_rad, w_scale, grid_size,
            num_skipped, norm, grid);
}

MI355_EXPORT void mi355_grid_wproj_f(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half, const float* conv_func,
        const size_t num_points, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight,
        const float cell_size_rad, const float w_scale, const int grid_size,
        size_t* num_skipped, double* norm, float* grid)
{
    (void)grid_host<float>(num_w_planes, support, oversample, conv_size_half, conv_func,
            num_points, uu, vv, ww, vis, weight, cell_size_rad, w_scale, grid_size,
            num_skipped, norm, grid);
}

MI355_EXPORT int mi355_ctx_create(int device, mi355_ctx** ctx)
{
    if (!ctx) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    *ctx = nullptr;
    mi355_ctx* C = new mi355_ctx();
    const int s = init_ctx(C, device);
    if (s != MI355_OK) { delete C; return s; }
    *ctx = C;
    return ok();
}

MI355_EXPORT int mi355_ctx_destroy(mi355_ctx* C)
{
    if (!C) return ok();
    if (C->graph_ctx) (void)mi355_ctx_destroy(C->graph_ctx);
    C->graph_ctx = nullptr;
    if (C->noalloc_ctx) (void)mi355_ctx_destroy(C->noalloc_ctx);
    C->noalloc_ctx = nullptr;
    if (C->graph_exec) (void)hipGraphExecDestroy(C->graph_exec);
    if (C->graph) (void)hipGraphDestroy(C->graph);
    if (C->capture_stream) (void)hipStreamDestroy(C->capture_stream);
    (void)hipSetDevice(C->device);
    (void)wait_ctx_work(C);
    std::vector<DevBuf*> all = device_buffers(C);
    if (C->fft_valid) (void)hipfftDestroy(C->fft_plan);
    for (DevBuf* b : all) release(C, *b);
    for (int i = 0; i < 2; ++i)
    {
        release_host(C->h_stage[i]);
        if (C->stage_ev[i]) (void)hipEventDestroy(C->stage_ev[i]);
    }
    for (int i = 0; i < EV_COUNT; ++i) if (C->ev[i]) (void)hipEventDestroy(C->ev[i]);
    if (C->switch_ev) (void)hipEventDestroy(C->switch_ev);
    delete C;
    return ok();
}

MI355_EXPORT int mi355_ctx_set_stream(mi355_ctx* C, void* stream)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (s != C->stream)
    {
        // mark the end of the work queued so far: a buffer that grows later
        // is freed only once that work is done
        HIP_TRY(hipSetDevice(C->device));
        if (C->switch_pending) HIP_TRY(hipEventSynchronize(C->switch_ev));
        HIP_TRY(hipEventRecord(C->switch_ev, C->stream));
        C->switch_pending = true;
    }
    C->stream = s;
    return ok();
}

MI355_EXPORT int mi355_ctx_get_stream(mi355_ctx* C, void** stream)
{
    if (!C || !stream) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    *stream = reinterpret_cast<void*>(C->stream);
    return ok();
}

MI355_EXPORT int mi355_ctx_set_option(mi355_ctx* C, int key, long long value)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    switch (key)
    {
    case MI355_OPT_VISITS_PER_ITEM:
        if (value < 0 || value > (1 << 24)) return fail(MI355_ERR_INVALID_ARG, "visits per item %lld", value);
        C->visits_per_item = (unsigned)value;
        break;
    case MI355_OPT_COMPENSATED: C->compensated = value != 0; break;
    case MI355_OPT_PHASE_TIMING: C->timing = value != 0; break;
    case MI355_OPT_CACHE_KERNELS: C->cache_kernels = value != 0; C->tables_valid = false; break;
    case MI355_OPT_GRID_KERNEL:
        if (value < 0 || value > 2) return fail(MI355_ERR_INVALID_ARG, "grid kernel %lld", value);
        C->grid_kernel = (int)value;
        break;
    case MI355_OPT_MIN_STAGE_RUN:
        if (value < 0) return fail(MI355_ERR_INVALID_ARG, "min stage run %lld", value);
        C->min_stage_run = (unsigned)std::min<long long>(value, 0xFFFFFFFFll);
        break;
    case MI355_OPT_ITEM_GROUP:
        if (value < 1 || value > 4096) return fail(MI355_ERR_INVALID_ARG, "item group %lld", value);
        C->item_group = (unsigned)value;
        break;
    case MI355_OPT_MIN_STAGE_RUN_W33:
    case MI355_OPT_MIN_STAGE_RUN_W65:
    case MI355_OPT_MIN_STAGE_RUN_W97:
    case MI355_OPT_MIN_STAGE_RUN_W111:
        if (value < 0) return fail(MI355_ERR_INVALID_ARG, "min stage run %lld", value);
        C->msr_class[key - MI355_OPT_MIN_STAGE_RUN_W33] =
                (unsigned)std::min<long long>(value, 0xFFFFFFFFll);
        break;
    case MI355_OPT_AUTOTUNE_CACHE: C->tune_cache = value != 0; break;
    case MI355_OPT_MAX_WORKSPACE:
        if (value < 0) return fail(MI355_ERR_INVALID_ARG, "workspace cap %lld", value);
        C->max_workspace = (unsigned long long)value;
        break;
    default: return fail(MI355_ERR_INVALID_ARG, "unknown option %d", key);
    }
    return ok();
}

MI355_EXPORT int mi355_ctx_get_option(const mi355_ctx* C, int key, long long* value)
{
    if (!C || !value) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    switch (key)
    {
    case MI355_OPT_VISITS_PER_ITEM: *value = C->visits_per_item; break;
    case MI355_OPT_COMPENSATED: *value = C->compensated; break;
    case MI355_OPT_PHASE_TIMING: *value = C->timing; break;
    case MI355_OPT_CACHE_KERNELS: *value = C->cache_kernels; break;
    case MI355_OPT_GRID_KERNEL: *value = C->grid_kernel; break;
    case MI355_OPT_MIN_STAGE_RUN: *value = C->min_stage_run; break;
    case MI355_OPT_ITEM_GROUP: *value = C->item_group; break;
    case MI355_OPT_MIN_STAGE_RUN_W33:
    case MI355_OPT_MIN_STAGE_RUN_W65:
    case MI355_OPT_MIN_STAGE_RUN_W97:
    case MI355_OPT_MIN_STAGE_RUN_W111:
        *value = C->msr_class[key - MI355_OPT_MIN_STAGE_RUN_W33];
        break;
    case MI355_OPT_AUTOTUNE_CACHE: *value = C->tune_cache; break;
    case MI355_OPT_MAX_WORKSPACE: *value = (long long)C->max_workspace; break;
    default: return fail(MI355_ERR_INVALID_ARG, "unknown option %d", key);
    }
    return ok();
}

MI355_EXPORT int mi355_ctx_get_stats(const mi355_ctx* Cc, mi355_stats* stats)
{
    if (!Cc || !stats) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    mi355_ctx* C = const_cast<mi355_ctx*>(Cc);
    if (C->stats_in_graph && C->graph_ctx) C = C->graph_ctx;
    else if (C->stats_in_noalloc && C->noalloc_ctx) C = C->noalloc_ctx;
    if (C->stats_deferred)
    {
        // host-sync-free call: read the counts the device kept
        HIP_TRY(hipSetDevice(C->device));
        unsigned long long sc[SC_COUNT];
        std::vector<unsigned long long> jo(3 * (size_t)C->deferred_jobs);
        HIP_TRY(hipMemcpyAsync(sc, C->scalars.p, sizeof(sc), hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipMemcpyAsync(jo.data(), C->job_out_d.p, jo.size() * sizeof(unsigned long long),
                hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipStreamSynchronize(C->stream));
        C->stats.num_points = C->deferred_points;
        C->stats.num_visits = sc[SC_VISITS];
        C->stats.num_items = sc[SC_ITEMS];
        C->stats.num_skipped = C->stats.num_taps = 0;
        for (int j = 0; j < C->deferred_jobs; ++j)
        {
            C->stats.num_skipped += jo[3 * j];
            C->stats.num_taps += jo[3 * j + 1];
        }
        if (!C->stats_untimed) collect_timing(C);
        C->stats_deferred = false;
    }
    *stats = C->stats;
    return ok();
}

MI355_EXPORT int mi355_ctx_get_buckets(mi355_ctx* C, unsigned long long* count,
        int* key_table_bits, unsigned* keys, unsigned* vis, int* flags, unsigned long long cap)
{
    if (!C || !count) return fail(MI355_ERR_INVALID_ARG, "NULL argument");
    if (!C->bk_valid) return fail(MI355_ERR_INVALID_ARG, "no strip-kernel gridding pass to report");
    *count = C->bk_count;
    if (key_table_bits) *key_table_bits = C->bk_key_table_bits;
    if (!keys && !vis && !flags) return ok();
    if (cap < C->bk_count) return fail(MI355_ERR_INVALID_ARG, "cap %llu < %llu visits", cap, C->bk_count);
    HIP_TRY(hipSetDevice(C->device));
    const size_t n = (size_t)C->bk_count;
    if (keys)
        HIP_TRY(hipMemcpyAsync(keys, C->keys[C->bk_cur].p, n * sizeof(unsigned),
                hipMemcpyDeviceToHost, C->stream));
    if (vis)
        HIP_TRY(hipMemcpyAsync(vis, C->vals[C->bk_cur].p, n * sizeof(unsigned),
                hipMemcpyDeviceToHost, C->stream));
    std::vector<int4> vm;
    if (flags)
    {
        vm.resize(n);
        HIP_TRY(hipMemcpyAsync(vm.data(), C->vmeta.p, n * sizeof(int4), hipMemcpyDeviceToHost,
                C->stream));
    }
    HIP_TRY(hipStreamSynchronize(C->stream));
    for (size_t i = 0; i < vm.size(); ++i)
        flags[i] = ((vm[i].z & kVisitStaged) ? 1 : 0) | ((vm[i].z & kVisitRunStart) ? 2 : 0);
    return ok();
}

MI355_EXPORT int mi355_default_ctx(mi355_ctx** ctx)
{
    if (!ctx) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lock(g_default_lock);
    TRY(get_default_ctx(ctx));
    return ok();
}

MI355_EXPORT int mi355_grid_wproj_f_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const float* conv, size_t N, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight, float cell, float w_scale,
        int G, unsigned long long* skipped, double* norm, float* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return grid_dev<float, float>(C, P, support, ov, csh, conv, N, uu, vv, ww, vis, weight, cell,
            w_scale, G, skipped, norm, grid, nullptr);
}

MI355_EXPORT int mi355_grid_wproj_f_dev_noalloc(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const float* conv, size_t N, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight, float cell, float w_scale,
        int G, unsigned long long* skipped, double* norm, float* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    HIP_TRY(hipSetDevice(C->device));
    // Set-up on the first call: the context's allocation-free twin, whose
    // buffers are pieces of this device's static scratch (its address was
    // resolved when the context was created).
    if (!C->noalloc_ctx)
    {
        if (C->device >= 64 || !g_scratch_addr[C->device])
            return fail(MI355_ERR_HIP, "no static scratch address for device %d", C->device);
        mi355_ctx* Z = new mi355_ctx();
        const int st = init_ctx(Z, C->device);
        if (st != MI355_OK) { delete Z; return st; }
        Z->arena = static_cast<unsigned char*>(g_scratch_addr[C->device]);
        Z->arena_cap = kNoAllocBytes;
        C->noalloc_ctx = Z;
    }
    mi355_ctx* Z = C->noalloc_ctx;
    Z->visits_per_item = C->visits_per_item;
    Z->compensated = C->compensated;
    Z->grid_kernel = C->grid_kernel;
    Z->min_stage_run = C->min_stage_run;
    std::copy(C->msr_class, C->msr_class + 4, Z->msr_class);
    Z->item_group = C->item_group;
    Z->timing = C->timing;
    Z->cache_kernels = false;          // (the scratch is shared: tables are rebuilt per call)
    Z->tables_valid = false;
    if (Z->stream != C->stream) TRY(mi355_ctx_set_stream(Z, C->stream));
    // every buffer is handed out afresh from the start of the scratch
    Z->arena_used = 0;
    for (DevBuf* b : device_buffers(Z)) { b->p = nullptr; b->cap = 0; }
    C->stats_in_noalloc = true;
    C->stats_in_graph = false;
    return grid_dev<float, float>(Z, P, support, ov, csh, conv, N, uu, vv, ww, vis, weight, cell,
            w_scale, G, skipped, norm, grid, nullptr);
}

MI355_EXPORT int mi355_grid_wproj_fsel_d_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const float* conv, size_t N, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight, float cell, float w_scale,
        int G, unsigned long long* skipped, double* norm, double* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return grid_dev<float, double>(C, P, support, ov, csh, conv, N, uu, vv, ww, vis, weight,
            cell, w_scale, G, skipped, norm, grid, nullptr);
}

MI355_EXPORT int mi355_select_wproj_f_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, size_t N, const float* uu, const float* vv, const float* ww, float cell,
        float w_scale, int G, int* out)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return select_dev<float>(C, P, support, ov, N, uu, vv, ww, cell, w_scale, G, out);
}

MI355_EXPORT int mi355_select_wproj_d_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, size_t N, const double* uu, const double* vv, const double* ww, double cell,
        double w_scale, int G, int* out)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return select_dev<double>(C, P, support, ov, N, uu, vv, ww, cell, w_scale, G, out);
}

MI355_EXPORT int mi355_grid_wproj_d_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const double* conv, size_t N, const double* uu, const double* vv,
        const double* ww, const double* vis, const double* weight, double cell,
        double w_scale, int G, unsigned long long* skipped, double* norm, double* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return grid_dev<double, double>(C, P, support, ov, csh, conv, N, uu, vv, ww, vis, weight,
            cell, w_scale, G, skipped, norm, grid, nullptr);
}

MI355_EXPORT int mi355_degrid_wproj_f_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const float* conv, size_t N, const float* uu, const float* vv,
        const float* ww, const float* weight, float cell, float w_scale, int G,
        const float* grid, int apply_weight, float* out)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return degrid_dev<float>(C, P, support, ov, csh, conv, N, uu, vv, ww, weight, cell,
            w_scale, G, grid, apply_weight, out);
}

MI355_EXPORT int mi355_degrid_wproj_d_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const double* conv, size_t N, const double* uu, const double* vv,
        const double* ww, const double* weight, double cell, double w_scale, int G,
        const double* grid, int apply_weight, double* out)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return degrid_dev<double>(C, P, support, ov, csh, conv, N, uu, vv, ww, weight, cell,
            w_scale, G, grid, apply_weight, out);
}

MI355_EXPORT int mi355_grid_wproj_f_batched_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const float* conv, int num_jobs, const mi355_job_f* jobs,
        const float* uu, const float* vv, const float* ww, const float* vis,
        const float* weight)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    std::vector<JobHost<float, float>> J;
    TRY(batched_jobs(P, support, ov, csh, conv, num_jobs, jobs, uu, vv, ww, vis, weight, J));
    HIP_TRY(hipSetDevice(C->device));
    C->stats_untimed = false;
    TRY((run_pipeline<float, float>(C, P, support, ov, csh, reinterpret_cast<const float2*>(conv),
            num_jobs, J.data(), uu, vv, ww, vis, weight, nullptr)));
    collect_timing(C);
    return ok();
}

MI355_EXPORT int mi355_grid_wproj_f_batched_graph_dev(mi355_ctx* C, size_t P,
        const int* support, int ov, int csh, const float* conv, int num_jobs,
        const mi355_job_f* jobs, const float* uu, const float* vv, const float* ww,
        const float* vis, const float* weight)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    std::vector<JobHost<float, float>> J;
    TRY(batched_jobs(P, support, ov, csh, conv, num_jobs, jobs, uu, vv, ww, vis, weight, J));
    HIP_TRY(hipSetDevice(C->device));
    // The graph's own context: its buffers, device job table, support and
    // plane offsets are written only by the graph's eager build run, so
    // calls of any other entry point on C leave a captured graph intact.
    if (!C->graph_ctx)
    {
        mi355_ctx* Gc = new mi355_ctx();
        const int s = init_ctx(Gc, C->device);
        if (s != MI355_OK) { delete Gc; return s; }
        C->graph_ctx = Gc;
    }
    mi355_ctx* const parent = C;
    C = parent->graph_ctx;
    C->visits_per_item = parent->visits_per_item;
    C->compensated = parent->compensated;
    C->grid_kernel = parent->grid_kernel;
    C->min_stage_run = parent->min_stage_run;
    std::copy(parent->msr_class, parent->msr_class + 4, C->msr_class);
    C->item_group = parent->item_group;
    C->max_workspace = parent->max_workspace;
    C->timing = parent->timing;
    C->cache_kernels = false;
    if (C->stream != parent->stream) TRY(mi355_ctx_set_stream(C, parent->stream));
    parent->stats_in_graph = true;
    // What the captured launch depends on: shapes, pointers, job table,
    // options.  The input values themselves may change between launches.
    std::vector<unsigned char> key;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        key.insert(key.end(), b, b + n);
    };
    const void* ptrs[6] = {conv, uu, vv, ww, vis, weight};
    const long long opts[11] = {(long long)C->visits_per_item, (long long)C->min_stage_run,
                                C->grid_kernel, C->compensated, (long long)C->item_group,
                                (long long)C->max_workspace, (long long)(intptr_t)C->stream,
                                C->msr_class[0], C->msr_class[1], C->msr_class[2],
                                C->msr_class[3]};
    put(&P, sizeof(P));
    put(support, P * sizeof(int));
    put(&ov, sizeof(ov));
    put(&csh, sizeof(csh));
    put(&num_jobs, sizeof(num_jobs));
    put(jobs, (size_t)num_jobs * sizeof(mi355_job_f));
    put(ptrs, sizeof(ptrs));
    put(opts, sizeof(opts));
    if (C->graph_exec && key == C->graph_key)
    {
        HIP_TRY(hipGraphLaunch(C->graph_exec, C->stream));
        C->stats = mi355_stats();
        C->stats_deferred = true;
        C->stats_untimed = true;
        return ok();
    }
    // Build: one eager host-sync-free run (which sizes every buffer and
    // produces this call's result), then the same sequence captured on the
    // library's own stream for the following calls.
    if (C->graph_exec) { (void)hipGraphExecDestroy(C->graph_exec); C->graph_exec = nullptr; }
    if (C->graph) { (void)hipGraphDestroy(C->graph); C->graph = nullptr; }
    C->graph_key.clear();
    C->stats_untimed = false;
    C->sync_free = true;
    int st = run_pipeline<float, float>(C, P, support, ov, csh,
            reinterpret_cast<const float2*>(conv), num_jobs, J.data(), uu, vv, ww, vis, weight,
            nullptr);
    const mi355_stats eager_stats = C->stats;
    const bool eager_deferred = C->stats_deferred;
    if (st == MI355_OK)
    {
        if (!C->capture_stream &&
                hipStreamCreateWithFlags(&C->capture_stream, hipStreamNonBlocking) != hipSuccess)
            st = fail(MI355_ERR_HIP, "capture stream");
        // the capture stream must see the eager run's uploads and buffers
        else if (hipStreamSynchronize(C->stream) != hipSuccess)
            st = fail(MI355_ERR_HIP, "stream synchronize");
    }
    if (st == MI355_OK)
    {
        hipStream_t user = C->stream;
        C->stream = C->capture_stream;
        C->capturing = true;
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(C->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            st = fail(MI355_ERR_HIP, "begin capture");
        else
        {
            st = run_pipeline<float, float>(C, P, support, ov, csh,
                    reinterpret_cast<const float2*>(conv), num_jobs, J.data(), uu, vv, ww, vis,
                    weight, nullptr);
            const hipError_t e = hipStreamEndCapture(C->stream, &g);
            if (st == MI355_OK && (e != hipSuccess || !g))
                st = fail(MI355_ERR_HIP, "end capture: %s", hipGetErrorString(e));
        }
        C->capturing = false;
        C->stream = user;
        if (st == MI355_OK)
        {
            if (hipGraphInstantiate(&C->graph_exec, g, nullptr, nullptr, 0) != hipSuccess)
            {
                C->graph_exec = nullptr;
                st = fail(MI355_ERR_HIP, "graph instantiate");
            }
            else
            {
                C->graph = g;
                C->graph_key = key;
            }
        }
        if (st != MI355_OK && g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
    }
    C->sync_free = false;
    C->stats = eager_stats;
    C->stats_deferred = eager_deferred;
    return st == MI355_OK ? ok() : st;
}

MI355_EXPORT int mi355_autotune_f_batched_dev(mi355_ctx* C, size_t P, const int* support,
        int ov, int csh, const float* conv, int num_jobs, const mi355_job_f* jobs,
        const float* uu, const float* vv, const float* ww, const float* vis,
        const float* weight, double* best_ms)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    std::vector<JobHost<float, float>> J;
    TRY(batched_jobs(P, support, ov, csh, conv, num_jobs, jobs, uu, vv, ww, vis, weight, J));
    HIP_TRY(hipSetDevice(C->device));
    // The workload's shape: what the best settings depend on (not the
    // pointers or the values).
    std::vector<unsigned char> sig;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        sig.insert(sig.end(), b, b + n);
    };
    put(&P, sizeof(P));
    put(support, P * sizeof(int));
    put(&ov, sizeof(ov));
    put(&csh, sizeof(csh));
    put(&num_jobs, sizeof(num_jobs));
    for (int j = 0; j < num_jobs; ++j)
    {
        put(&J[j].grid_size, sizeof(int));
        put(&J[j].num_points, sizeof(J[j].num_points));
    }
    auto apply = [&](int kernel, unsigned vpi, unsigned msr, unsigned group, const unsigned* cls) {
        C->grid_kernel = kernel;
        C->visits_per_item = vpi;
        C->min_stage_run = msr;
        C->item_group = group;
        std::copy(cls, cls + 4, C->msr_class);
    };
    if (C->tune_cache)
        for (const auto& t : C->tuned)
            if (t.sig == sig)
            {
                apply(t.kernel, t.vpi, t.msr, t.group, t.cls);
                if (best_ms) *best_ms = t.ms;
                return ok();
            }
    // Trial outputs: a private scratch grid per job (sized for it), norm and
    // count per job, so the caller's outputs are not touched and the trial
    // jobs write as many grids as the real call does.
    size_t cells = 0;
    for (int j = 0; j < num_jobs; ++j) cells += (size_t)J[j].grid_size * J[j].grid_size;
    DevBuf tgrid, tnorm, tskip;
    TRY(ensure(C, tgrid, cells * sizeof(float2)));
    TRY(ensure(C, tnorm, (size_t)num_jobs * sizeof(double)));
    TRY(ensure(C, tskip, (size_t)num_jobs * sizeof(unsigned long long)));
    for (int j = 0, off = 0; j < num_jobs; ++j)
    {
        J[j].grid = tgrid.as<float2>() + off;      // (results are discarded)
        J[j].norm = tnorm.as<double>() + j;
        J[j].skipped = tskip.as<unsigned long long>() + j;
        off += J[j].grid_size * J[j].grid_size;
    }
    const bool timing0 = C->timing;
    const unsigned vpi0 = C->visits_per_item, msr0 = C->min_stage_run, ig0 = C->item_group;
    const int gk0 = C->grid_kernel;
    unsigned cls0[4];
    std::copy(C->msr_class, C->msr_class + 4, cls0);
    C->timing = true;
    int st = MI355_OK;
    // one timed pass with the context's current settings
    auto trial = [&](double& ms) -> int {
        const int rs = run_pipeline<float, float>(C, P, support, ov, csh,
                reinterpret_cast<const float2*>(conv), num_jobs, J.data(), uu, vv, ww, vis,
                weight, nullptr);
        if (rs != MI355_OK) return rs;
        collect_timing(C);
        ms = C->stats.ms_total;
        return MI355_OK;
    };
    struct Setting { int kernel; unsigned vpi, msr, group; unsigned cls[4]; };
    Setting best{gk0, vpi0, msr0, ig0, {cls0[0], cls0[1], cls0[2], cls0[3]}};
    double best_t = 1e300, t = 0;
    // warm-up (buffers grow to their sizes, tables are expanded)
    apply(best.kernel, best.vpi, best.msr, best.group, best.cls);
    st = trial(t);
    // Stage 1: kernel / geometry, work-item size, item grouping, with one
    // staged-run threshold for every class.
    const unsigned zero4[4] = {0, 0, 0, 0};
    for (int kern : {0, 2})
        for (unsigned vpi : {1024u, 2048u, 4096u, 8192u})
            for (unsigned grp : {1u, 8u})
            {
                if (st != MI355_OK) break;
                if (kern == 2 && grp != 1u) continue;
                apply(kern, vpi, kern == 2 ? 4u : 3u, grp, zero4);
                st = trial(t);
                if (st == MI355_OK && t < best_t)
                {
                    best_t = t;
                    best = Setting{kern, vpi, C->min_stage_run, grp, {0, 0, 0, 0}};
                }
            }
    // Stage 2 (strip kernel): each support class's threshold in turn, the
    // others held at their best so far; two passes per value, the faster
    // counts.
    if (st == MI355_OK && best.kernel == 0)
    {
        for (int c = 0; c < 4 && st == MI355_OK; ++c)
        {
            unsigned cls[4];
            std::copy(best.cls, best.cls + 4, cls);
            for (unsigned v : {1u, 2u, 3u, 4u, 6u, 8u, 12u})
            {
                cls[c] = v;
                apply(best.kernel, best.vpi, best.msr, best.group, cls);
                double t2 = 0;
                st = trial(t);
                if (st == MI355_OK) st = trial(t2);
                if (st != MI355_OK) break;
                t = std::min(t, t2);
                if (t < best_t)
                {
                    best_t = t;
                    std::copy(cls, cls + 4, best.cls);
                }
            }
        }
    }
    C->timing = timing0;
    release(C, tgrid); release(C, tnorm); release(C, tskip);
    if (st != MI355_OK)
    {
        apply(gk0, vpi0, msr0, ig0, cls0);
        return st;
    }
    apply(best.kernel, best.vpi, best.msr, best.group, best.cls);
    mi355_ctx::Tuned rec{sig, best.kernel, best.vpi, best.msr, best.group,
              {best.cls[0], best.cls[1], best.cls[2], best.cls[3]}, best_t};
    bool stored = false;
    for (auto& t0 : C->tuned)
        if (t0.sig == sig) { t0 = rec; stored = true; }
    if (!stored) C->tuned.push_back(rec);
    if (best_ms) *best_ms = best_t;
    return ok();
}

// ---- uniform-weighting pre-pass ------------------------------------------
MI355_EXPORT void mi355_grid_weights_write_d(const size_t num_points, const double* uu,
        const double* vv, const double* weight, const double cell_size_rad,
        const int grid_size, size_t* num_skipped, double* grid)
{
    (void)weights_write_host<double>(num_points, uu, vv, weight, cell_size_rad, grid_size,
            num_skipped, grid);
}

MI355_EXPORT void mi355_grid_weights_write_f(const size_t num_points, const float* uu,
        const float* vv, const float* weight, const float cell_size_rad,
        const int grid_size, size_t* num_skipped, float* grid)
{
    (void)weights_write_host<float>(num_points, uu, vv, weight, cell_size_rad, grid_size,
            num_skipped, grid);
}

MI355_EXPORT void mi355_grid_weights_read_d(const size_t num_points, const double* uu,
        const double* vv, const double* weight_in, double* weight_out,
        const double cell_size_rad, const int grid_size, size_t* num_skipped,
        const double* grid)
{
    (void)weights_read_host<double>(num_points, uu, vv, weight_in, weight_out, cell_size_rad,
            grid_size, num_skipped, grid);
}

MI355_EXPORT void mi355_grid_weights_read_f(const size_t num_points, const float* uu,
        const float* vv, const float* weight_in, float* weight_out,
        const float cell_size_rad, const int grid_size, size_t* num_skipped,
        const float* grid)
{
    (void)weights_read_host<float>(num_points, uu, vv, weight_in, weight_out, cell_size_rad,
            grid_size, num_skipped, grid);
}

MI355_EXPORT int mi355_grid_weights_write_f_dev(mi355_ctx* C, size_t N, const float* uu,
        const float* vv, const float* weight, float cell, int G,
        unsigned long long* num_skipped, float* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return weights_write_dev<float>(C, N, uu, vv, weight, cell, G, num_skipped, grid);
}

MI355_EXPORT int mi355_grid_weights_write_d_dev(mi355_ctx* C, size_t N, const double* uu,
        const double* vv, const double* weight, double cell, int G,
        unsigned long long* num_skipped, double* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return weights_write_dev<double>(C, N, uu, vv, weight, cell, G, num_skipped, grid);
}

MI355_EXPORT int mi355_grid_weights_read_f_dev(mi355_ctx* C, size_t N, const float* uu,
        const float* vv, const float* weight_in, float* weight_out, float cell, int G,
        unsigned long long* num_skipped, const float* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return weights_read_dev<float>(C, N, uu, vv, weight_in, weight_out, cell, G, num_skipped,
            grid);
}

MI355_EXPORT int mi355_grid_weights_read_d_dev(mi355_ctx* C, size_t N, const double* uu,
        const double* vv, const double* weight_in, double* weight_out, double cell, int G,
        unsigned long long* num_skipped, const double* grid)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return weights_read_dev<double>(C, N, uu, vv, weight_in, weight_out, cell, G, num_skipped,
            grid);
}

// ---- grid -> image -------------------------------------------------------
MI355_EXPORT void mi355_grid_to_image_f(int grid_size, const float* grid, double norm,
        float* image)
{
    (void)image_host<float>(grid_size, grid, norm, image);
}

MI355_EXPORT void mi355_grid_to_image_d(int grid_size, const double* grid, double norm,
        double* image)
{
    (void)image_host<double>(grid_size, grid, norm, image);
}

MI355_EXPORT int mi355_grid_to_image_f_dev(mi355_ctx* C, int grid_size, const float* grid,
        double norm, float* image)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return image_dev<float>(C, grid_size, grid, norm, image);
}

MI355_EXPORT int mi355_grid_to_image_d_dev(mi355_ctx* C, int grid_size, const double* grid,
        double norm, double* image)
{
    if (!C) return fail(MI355_ERR_INVALID_ARG, "ctx is NULL");
    return image_dev<double>(C, grid_size, grid, norm, image);
}

} // extern "C"
