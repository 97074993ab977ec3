/*
 * mi355_gridder.h -- C ABI of the MI355X (gfx950) W-projection gridder.
 *
 * The library replaces the reference's serial convolutional gridder
 * (oskar_grid_wproj.c) behind the same C signature.  All entry points are
 * plain C: pointers, sizes and status codes, no C++ or framework types.
 *
 * Semantics kept from the serial gridder (oskar_grid_wproj.c:126-212):
 *   - num_skipped is SET (zeroed, then counted) on every call (:151, :185);
 *   - norm and grid are ACCUMULATED across calls (:205-206, :210);
 *   - cell, oversampled offset and w-plane selection are bit-identical to
 *     the serial code's float (for _f) or double (for _d) arithmetic;
 *   - the grid is G*G complex values, row-major, re/im interleaved.
 * The grid itself is summed in a different order (per-tile register
 * accumulation, then an FP64 merge of tile partials), so FP32 results agree
 * with the serial code to rounding, not bit for bit.
 *
 * Errors: the drop-ins return void, like the functions they replace; the
 * status of the most recent call on the calling thread is available from
 * mi355_grid_last_error().  All other entry points return that status.
 */
#ifndef MI355_GRIDDER_H_
#define MI355_GRIDDER_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI355_EXPORT __attribute__((visibility("default")))

/* Status codes. */
#define MI355_OK                   0
#define MI355_ERR_INVALID_ARG      1  /* null pointer, bad size, oversample < 1 ... */
#define MI355_ERR_KERNEL_TOO_SMALL 2  /* conv_size_half too small for a plane's support */
#define MI355_ERR_NO_DEVICE        3  /* no HIP device visible */
#define MI355_ERR_HIP              4  /* a HIP runtime call failed */
#define MI355_ERR_OUT_OF_MEMORY    5  /* device allocation failed */
#define MI355_ERR_TOO_LARGE        6  /* a count exceeds the library's 32-bit indexing */
#define MI355_ERR_IO               7  /* file open / read / write failed (mi355_io.h) */
#define MI355_ERR_FORMAT           8  /* file is not in the expected format (mi355_io.h) */

/* ------------------------------------------------------------------------
 * Drop-in replacements (host pointers, synchronous).
 * ---------------------------------------------------------------------- */

/* Replaces oskar_grid_wproj_d (oskar_grid_wproj.h:68-86); same arguments,
 * same meaning.  FP64 throughout, like the function it replaces. */
MI355_EXPORT void mi355_grid_wproj_d(
        const size_t num_w_planes, const int* support, const int oversample,
        const int conv_size_half, const double* conv_func,
        const size_t num_points, const double* uu, const double* vv,
        const double* ww, const double* vis, const double* weight,
        const double cell_size_rad, const double w_scale, const int grid_size,
        size_t* num_skipped, double* norm, double* grid);

/* Replaces oskar_grid_wproj_f (oskar_grid_wproj.h:113-131), the function
 * main.cpp's replacement hook calls (main.cpp:288-338); same arguments, same
 * meaning.  The reference's own GPU entry oskar_grid_wproj_f_gpu
 * (oskar_grid_wproj_gpu.hpp:92-99) narrows num_points to int, takes double
 * cell and w_scale, zeroes *norm and assigns the grid; this one does not. */
MI355_EXPORT void mi355_grid_wproj_f(
        const size_t num_w_planes, const int* support, const int oversample,
        const int conv_size_half, const float* conv_func,
        const size_t num_points, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight,
        const float cell_size_rad, const float w_scale, const int grid_size,
        size_t* num_skipped, double* norm, float* grid);

/* Status of the last call made on this thread (MI355_OK when it succeeded),
 * for the void drop-ins.  The reference's GPU path aborts instead
 * (gpu_support.hpp:143-151). */
MI355_EXPORT int mi355_grid_last_error(void);
MI355_EXPORT const char* mi355_grid_last_error_message(void);
MI355_EXPORT const char* mi355_grid_error_string(int status);

/* ------------------------------------------------------------------------
 * Contexts: one per device and caller thread; they own the persistent
 * device workspace and the stream.  The drop-ins use a library-owned
 * context on the current device, guarded by a mutex.  No reference
 * counterpart: the reference allocates and frees per call
 * (oskar_grid_wproj_gpu.cu:1759-2121).
 * ---------------------------------------------------------------------- */
typedef struct mi355_ctx mi355_ctx;

MI355_EXPORT int mi355_ctx_create(int device, mi355_ctx** ctx);
MI355_EXPORT int mi355_ctx_destroy(mi355_ctx* ctx);
/* The library-owned context the drop-ins use on the current device (for
 * setting its options); do not destroy it. */
MI355_EXPORT int mi355_default_ctx(mi355_ctx** ctx);
/* Work is enqueued on this HIP stream (a hipStream_t; 0 = null stream). */
MI355_EXPORT int mi355_ctx_set_stream(mi355_ctx* ctx, void* hip_stream);
/* The stream the context's work is enqueued on (as set above; 0 = null). */
MI355_EXPORT int mi355_ctx_get_stream(mi355_ctx* ctx, void** hip_stream);

/* Tuning and behaviour options (mi355_ctx_set_option). */
#define MI355_OPT_VISITS_PER_ITEM 1  /* tile visits per work item; 0: automatic
                                        (default): the largest of 8192, 4096,
                                        2048 for which the call has at least 6 x
                                        that many visibilities per CU, else 1024 */
#define MI355_OPT_COMPENSATED     2  /* 1: TwoSum-compensated LDS accumulation */
#define MI355_OPT_PHASE_TIMING    3  /* 1: record hipEvents around each phase */
#define MI355_OPT_CACHE_KERNELS   4  /* 1: reuse prepared kernel tables while the
                                        conv_func pointer and shape are unchanged */
#define MI355_OPT_GRID_KERNEL     5  /* 0: automatic (default): 16-column strips on
                                        256x128 regions for FP32 (192x64 when
                                        compensated), 2 for FP64;
                                        1: 64x64 tiles, tables read directly;
                                        2: 64x64 tiles, tables staged in LDS */
#define MI355_OPT_MIN_STAGE_RUN   6  /* shortest same-table run staged in LDS;
                                        0: automatic (default): for the strip
                                        kernel 6 for tables up to 65 wide and 3
                                        for wider ones; 3 for the 64x64 kernels,
                                        8 for the FP64 one */
#define MI355_OPT_ITEM_GROUP      7  /* strip kernel: work items are dealt to the
                                        8 XCDs in groups of this many consecutive
                                        items (default 1: round-robin) */
#define MI355_OPT_MAX_WORKSPACE   8  /* cap in bytes on the device workspace the
                                        context may hold (0: no cap, the default);
                                        a call that would need more fails with
                                        MI355_ERR_OUT_OF_MEMORY before allocating */
/* Strip kernel (FP32): shortest same-table run staged in LDS per support
 * class, tables at most 33, 65, 97 and 111 wide (0, the default: the value
 * of MI355_OPT_MIN_STAGE_RUN). */
#define MI355_OPT_MIN_STAGE_RUN_W33  9
#define MI355_OPT_MIN_STAGE_RUN_W65  10
#define MI355_OPT_MIN_STAGE_RUN_W97  11
#define MI355_OPT_MIN_STAGE_RUN_W111 12
#define MI355_OPT_AUTOTUNE_CACHE  13 /* 1 (default): the autotuner answers a
                                        workload shape it has tuned on this
                                        context from its cache; 0: it always
                                        runs its trials (and updates the cache) */
MI355_EXPORT int mi355_ctx_set_option(mi355_ctx* ctx, int key, long long value);
MI355_EXPORT int mi355_ctx_get_option(const mi355_ctx* ctx, int key, long long* value);

/* Figures of the last gridding call on this context. */
typedef struct
{
    double ms_prepare;   /* kernel-table expansion and norm table */
    double ms_select;    /* per-visibility selection (index stage) */
    double ms_bucket;    /* visit scan, emit and radix sort by tile */
    double ms_grid;      /* tile gridding kernel */
    double ms_merge;     /* ordered FP64 merge of multi-item tiles */
    double ms_total;     /* all of the above, first event to last */
    unsigned long long num_points;
    unsigned long long num_skipped;
    unsigned long long num_visits;  /* (visibility, tile) pairs */
    unsigned long long num_items;   /* work items launched */
    unsigned long long num_taps;    /* sum over gridded visibilities of (2s+1)^2 */
    unsigned long long num_merge_tiles;
    unsigned long long visits_per_item;  /* work-item size used (the option, or
                                            the automatic choice) */
} mi355_stats;

MI355_EXPORT int mi355_ctx_get_stats(const mi355_ctx* ctx, mi355_stats* stats);

/* Introspection of the last FP32 strip-kernel gridding pass on the context
 * (not a host-synchronisation-free one): its sorted tile lists, for checking
 * the bucketing against a restatement of the selection.  *count receives
 * the number of (region, visibility) visits and *key_table_bits the number
 * of low key bits below the region id; with cap >= *count, keys[i] and
 * vis[i] receive the i-th sorted visit's key ((region << key_table_bits) |
 * kernel-table id) and visibility index, and flags[i] bit 0 = the visit's
 * run is staged in LDS, bit 1 = the visit starts a run (any output may be
 * NULL).  No reference counterpart: it stands for the per-tile count and
 * scatter of oskar_grid_wproj_gpu.cu:359-649. */
MI355_EXPORT int mi355_ctx_get_buckets(mi355_ctx* ctx, unsigned long long* count,
        int* key_table_bits, unsigned* keys, unsigned* vis, int* flags, unsigned long long cap);

/* ------------------------------------------------------------------------
 * Device-pointer gridding: every array is in device memory except
 * `support` (host, num_w_planes ints).  num_skipped (one unsigned long long)
 * and norm (one double) are device pointers; the call sets *num_skipped and
 * adds to *norm and to grid.  Same arithmetic as the drop-ins.
 * ---------------------------------------------------------------------- */
MI355_EXPORT int mi355_grid_wproj_f_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const float* conv_func, size_t num_points,
        const float* uu, const float* vv, const float* ww, const float* vis,
        const float* weight, float cell_size_rad, float w_scale,
        int grid_size, unsigned long long* num_skipped, double* norm,
        float* grid);

/* The same gridding pass without any device allocation.  No reference
 * counterpart: for callers that cannot allocate device memory during the
 * call (work captured into a HIP graph, tight serving loops, harnesses that
 * forbid hipMalloc).  Every buffer the pass needs beyond the caller's inputs
 * and outputs is a piece of a static device array of 64 MiB that each device
 * the library runs on holds until the process ends (its address is resolved
 * when a context is created; the first call on a context also creates its
 * events).  A call that needs more fails with MI355_ERR_TOO_LARGE before its
 * outputs are touched.  All calls on one device share that array, and the
 * array keeps the last call's contents: calls that overlap on one device
 * (on any contexts or streams) race.  The kernel tables are rebuilt on every
 * call (MI355_OPT_CACHE_KERNELS does not apply); the other options are the
 * context's. */
MI355_EXPORT int mi355_grid_wproj_f_dev_noalloc(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const float* conv_func, size_t num_points,
        const float* uu, const float* vv, const float* ww, const float* vis,
        const float* weight, float cell_size_rad, float w_scale,
        int grid_size, unsigned long long* num_skipped, double* norm,
        float* grid);

MI355_EXPORT int mi355_grid_wproj_d_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const double* conv_func, size_t num_points,
        const double* uu, const double* vv, const double* ww,
        const double* vis, const double* weight, double cell_size_rad,
        double w_scale, int grid_size, unsigned long long* num_skipped,
        double* norm, double* grid);

/* FP32 inputs and FP32 selection (bit-identical cells, offsets and planes to
 * mi355_grid_wproj_f_dev and oskar_grid_wproj_f), accumulated in FP64 into a
 * complex128 grid.  No reference counterpart: it is the FP64 truth against
 * which the FP32 grid's error is stated without counting the cell or plane
 * flips a pure-double selection (oskar_grid_wproj_d) would make. */
MI355_EXPORT int mi355_grid_wproj_fsel_d_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const float* conv_func, size_t num_points,
        const float* uu, const float* vv, const float* ww, const float* vis,
        const float* weight, float cell_size_rad, float w_scale,
        int grid_size, unsigned long long* num_skipped, double* norm,
        double* grid);

/* Selection only: for each visibility, the cell, oversampled offsets,
 * clamped plane, support, conjugation flag (ww > 0) and skip flag that the
 * gridder uses -- the per-visibility part of oskar_grid_wproj.c:158-187
 * (float) / :69-98 (double).  out is int32 [8][num_points] in that field
 * order (grid_u, grid_v, off_u, off_v, plane, support, conj, skip). */
MI355_EXPORT int mi355_select_wproj_f_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        size_t num_points, const float* uu, const float* vv, const float* ww,
        float cell_size_rad, float w_scale, int grid_size, int* out);
MI355_EXPORT int mi355_select_wproj_d_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        size_t num_points, const double* uu, const double* vv, const double* ww,
        double cell_size_rad, double w_scale, int grid_size, int* out);

/* ------------------------------------------------------------------------
 * Degridding: the exact adjoint of gridding with respect to the complex
 * amplitudes (no reference counterpart).  For every visibility i,
 *   vis_out[i] = w_i * sum over its footprint of conj(c'_i) * grid[cell],
 * c' being the kernel tap after the conjugation flip for ww > 0; with
 * apply_weight = 0 the factor w_i is omitted.  Skipped visibilities get 0.
 * Device pointers except `support`.
 * ---------------------------------------------------------------------- */
MI355_EXPORT int mi355_degrid_wproj_f_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const float* conv_func, size_t num_points,
        const float* uu, const float* vv, const float* ww,
        const float* weight, float cell_size_rad, float w_scale,
        int grid_size, const float* grid, int apply_weight, float* vis_out);

MI355_EXPORT int mi355_degrid_wproj_d_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const double* conv_func, size_t num_points,
        const double* uu, const double* vv, const double* ww,
        const double* weight, double cell_size_rad, double w_scale,
        int grid_size, const double* grid, int apply_weight,
        double* vis_out);

/* ------------------------------------------------------------------------
 * Batched gridding: many independent grids (snapshots / channels) sharing
 * one kernel stack, gridded in one pass of the pipeline.  Job j grids
 * visibilities [vis_offset, vis_offset + num_points) of the concatenated
 * arrays onto its own grid.  The reference grids one block per call
 * (main.cpp:243-381); this is the same operation for many blocks at once.
 * Every job needs outputs of its own: grids, norms and counts that overlap
 * another job's are refused (MI355_ERR_INVALID_ARG), since the jobs'
 * outputs are written concurrently.  (main.cpp's blocks, which all add to
 * one grid, are gridded by calling a single-job entry point per block.)
 * ---------------------------------------------------------------------- */
typedef struct
{
    size_t vis_offset;
    size_t num_points;
    float cell_size_rad;
    float w_scale;
    int grid_size;
    float* grid;                      /* device, grid_size^2 complex */
    double* norm;                     /* device, accumulated */
    unsigned long long* num_skipped;  /* device, set */
} mi355_job_f;

MI355_EXPORT int mi355_grid_wproj_f_batched_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const float* conv_func, int num_jobs,
        const mi355_job_f* jobs, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight);

/* The same batched pass as one HIP graph launched on the context's stream.
 * The first call with a given job table, set of pointers and options runs
 * the pipeline without host synchronisation (buffers sized from upper
 * bounds, counts kept on the device), then captures that sequence; later
 * calls with the same table, pointers and options replay the graph, reading
 * whatever the input arrays hold at the time.  Results arrive through the
 * job table's device pointers when the stream reaches them;
 * mi355_ctx_get_stats synchronises the stream and reads the counts back.
 * No reference counterpart (the reference grids one block per call). */
MI355_EXPORT int mi355_grid_wproj_f_batched_graph_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const float* conv_func, int num_jobs,
        const mi355_job_f* jobs, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight);

/* Autotuner for a batched workload: times the pipeline into private
 * outputs (one scratch grid per job; the job table's grids are not
 * touched), first over the gridding kernel / tile geometry (16-column
 * strips on 256x128 regions, or 64x64 tiles), work-item size and item
 * grouping, then, for the strip kernel, over the shortest staged run of
 * each support class in turn (tables at most 33, 65, 97, 111 wide), and
 * leaves the fastest settings in the context's options
 * (MI355_OPT_GRID_KERNEL, _VISITS_PER_ITEM, _ITEM_GROUP, _MIN_STAGE_RUN,
 * _MIN_STAGE_RUN_W33 .. _W111).  The choice is cached on the context per
 * workload shape (job count, visibility counts, grid sizes, w-planes and
 * supports, oversampling, kernel-stack size): a later call with the same
 * shape applies it without trials (see MI355_OPT_AUTOTUNE_CACHE).
 * *best_ms receives the chosen settings' time per pass.  Counterpart:
 * OskarTuningParams (oskar_grid_wproj_gpu.cu:1646-1655, set at
 * :1683-1688), fixed there. */
MI355_EXPORT int mi355_autotune_f_batched_dev(mi355_ctx* ctx,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, const float* conv_func, int num_jobs,
        const mi355_job_f* jobs, const float* uu, const float* vv,
        const float* ww, const float* vis, const float* weight, double* best_ms);

/* ------------------------------------------------------------------------
 * Uniform-weighting pre-pass.
 *
 * Write adds each point's weight to the cell of the weights grid (a G x G
 * real array, row v, column u) that the point falls in; read divides each
 * point's weight by its cell's total (0 where the total is 0).  Points whose
 * cell lies off the grid are counted in num_skipped: write ignores them and
 * read leaves their weight_out entries untouched.  The grid written is
 * bit-identical to the serial one (per cell, the same additions in point
 * order).  grid_size is limited to 65535 and num_points to 2^32-1 per call
 * (MI355_ERR_TOO_LARGE).
 * ---------------------------------------------------------------------- */

/* Replace oskar_grid_weights_write_d / _f (oskar_grid_weights.h:60-64 /
 * :107-111; called at main.cpp:229-237); same arguments, same meaning.  Only
 * the window of the grid spanned by the points' cells is copied to the
 * device and back.  Errors: mi355_grid_last_error. */
MI355_EXPORT void mi355_grid_weights_write_d(const size_t num_points,
        const double* uu, const double* vv, const double* weight,
        const double cell_size_rad, const int grid_size,
        size_t* num_skipped, double* grid);
MI355_EXPORT void mi355_grid_weights_write_f(const size_t num_points,
        const float* uu, const float* vv, const float* weight,
        const float cell_size_rad, const int grid_size,
        size_t* num_skipped, float* grid);

/* Replace oskar_grid_weights_read_d / _f (oskar_grid_weights.h:84-88 /
 * :131-135; called at main.cpp:273-285); same arguments, same meaning. */
MI355_EXPORT void mi355_grid_weights_read_d(const size_t num_points,
        const double* uu, const double* vv, const double* weight_in,
        double* weight_out, const double cell_size_rad, const int grid_size,
        size_t* num_skipped, const double* grid);
MI355_EXPORT void mi355_grid_weights_read_f(const size_t num_points,
        const float* uu, const float* vv, const float* weight_in,
        float* weight_out, const float cell_size_rad, const int grid_size,
        size_t* num_skipped, const float* grid);

/* Device-pointer forms of the four calls above, asynchronous on the
 * context's stream and free of host synchronisation.  *num_skipped (a
 * device u64) is set, not accumulated, like the reference's; the grid
 * accumulates across write calls. */
MI355_EXPORT int mi355_grid_weights_write_f_dev(mi355_ctx* ctx, size_t num_points,
        const float* uu, const float* vv, const float* weight, float cell_size_rad,
        int grid_size, unsigned long long* num_skipped, float* grid);
MI355_EXPORT int mi355_grid_weights_write_d_dev(mi355_ctx* ctx, size_t num_points,
        const double* uu, const double* vv, const double* weight, double cell_size_rad,
        int grid_size, unsigned long long* num_skipped, double* grid);
MI355_EXPORT int mi355_grid_weights_read_f_dev(mi355_ctx* ctx, size_t num_points,
        const float* uu, const float* vv, const float* weight_in, float* weight_out,
        float cell_size_rad, int grid_size, unsigned long long* num_skipped,
        const float* grid);
MI355_EXPORT int mi355_grid_weights_read_d_dev(mi355_ctx* ctx, size_t num_points,
        const double* uu, const double* vv, const double* weight_in, double* weight_out,
        double cell_size_rad, int grid_size, unsigned long long* num_skipped,
        const double* grid);

/* ------------------------------------------------------------------------
 * Grid -> image: the 2-D FFT of the grid that README:24-26 names as the step
 * after gridding (the reference has no code for it).  For an even
 * grid_size G and c = G/2,
 *   image[y][x] = Re( sum_{v,u} grid[v][u]
 *                     * exp(+2 pi i ((u-c)(x-c) + (v-c)(y-c)) / G) ) / norm
 * (1/norm taken as 1 when norm is 0): the centred inverse transform, so a
 * grid holding one value at cell (c, c) gives a constant image.  grid is G*G
 * complex values (interleaved re, im, row-major), image G*G reals.  hipFFT
 * does the transform; two HIP passes apply the centring signs and the
 * scale.  Odd grid sizes: MI355_ERR_INVALID_ARG; above 65536:
 * MI355_ERR_TOO_LARGE.  The device forms use a work buffer of G*G complex
 * values kept in the context.
 * ---------------------------------------------------------------------- */
MI355_EXPORT void mi355_grid_to_image_f(int grid_size, const float* grid, double norm,
        float* image);
MI355_EXPORT void mi355_grid_to_image_d(int grid_size, const double* grid, double norm,
        double* image);
MI355_EXPORT int mi355_grid_to_image_f_dev(mi355_ctx* ctx, int grid_size, const float* grid,
        double norm, float* image);
MI355_EXPORT int mi355_grid_to_image_d_dev(mi355_ctx* ctx, int grid_size, const double* grid,
        double norm, double* image);

#ifdef __cplusplus
}
#endif

#endif /* MI355_GRIDDER_H_ */
