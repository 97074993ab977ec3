/*
 * grid_oracle.c -- CPU restatement of the serial W-projection gridder.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker.
 *
 * It restates, function by function, the reference's single-threaded
 * gridders:
 *   oskar_grid_wproj_f   oskar_grid_wproj.c:126-212  (FP32)
 *   oskar_grid_wproj_d   oskar_grid_wproj.c:37-123   (FP64)
 * keeping every floating-point operation, its precision and its order, so
 * that on x86-64 (gcc, -O2 -ffp-contract=off, no FMA) the grid, norm and
 * num_skipped are bit-identical to the compiled reference.  That claim is
 * pinned by tests/test_oracle.py against (1) the committed golden fixtures in
 * tests/golden/, written by the compiled reference, (2) oracle/_ref/ (the
 * reference compiled from its own sources) when present, and (3) closed forms
 * (delta and constant kernels, conjugation symmetry, linearity, Hermitian
 * mirror, norm table, brute-force cell selection on tiny grids).
 *
 * Additional checkers with no reference counterpart (pinned by the same
 * closed forms and by agreement with the functions above where they coincide):
 *   oracle_select_f / _d       per-visibility cell / offset / plane selection
 *                              (the index part of oskar_grid_wproj.c:158-187)
 *   oracle_grid_wproj_fsel_d   FP32 selection, FP64 accumulation: the FP64
 *                              truth that the FP32 GPU grid is measured
 *                              against without counting selection flips.
 *   oracle_degrid_wproj_fsel_d exact adjoint of the above (FP64).
 *   oracle_grid_wproj_fsel_d_tiles  the same FP64 truth restricted to a list
 *                              of 64x64 grid tiles, multi-threaded, for
 *                              spot checks of full-size grids.
 *   oracle_norm_f              norm and num_skipped of oracle_grid_wproj_f
 *                              without the grid (bit-identical to it).
 *
 * Uniform weighting (the pre-pass upstream of the gridder), restated from
 * the reference's serial functions, same precision and order:
 *   oracle_grid_weights_write_f / _d   oskar_grid_weights.c:106-134 / :37-65
 *   oracle_grid_weights_read_f / _d    oskar_grid_weights.c:136-165 / :67-96
 * pinned against the same functions compiled from the reference (oracle/_ref)
 * and by closed forms (unit weights give per-cell counts; read after write
 * gives weight / cell total).
 *
 * Grid -> image (the "2D FFT of the grid" that README:24-26 names as the
 * step after gridding; the reference has no code for it):
 *   oracle_grid_to_image_d   centred inverse DFT, real part, scaled by
 *                            1/norm, by direct summation in FP64
 * pinned by closed forms (point sources at known cells give cosine images)
 * and by numpy's FFT on the same grid.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORACLE_API __attribute__((visibility("default")))

/* Selection of one visibility: cell, oversampled offsets, plane and support.
 * Mirrors oskar_grid_wproj.c:158-187 (float) and :69-98 (double). */
typedef struct
{
    int grid_u, grid_v;     /* nearest cell (oskar_grid_wproj.c:163-164) */
    int off_u, off_v;       /* oversampled offsets (:172-173) */
    size_t plane;           /* plane actually used, after the clamp (:176-179) */
    int support;            /* w_support (:176-177) */
    int conj_neg;           /* 1 when ww > 0, i.e. Im(kernel) is negated (:161) */
    int skip;               /* footprint leaves the grid (:182-187) */
} oracle_sel;

static void select_f(float u, float v, float w, float grid_scale,
        float w_scale, int grid_size, int oversample, size_t num_w_planes,
        const int* support, oracle_sel* s)
{
    const int centre = grid_size / 2;
    const float pos_u = -u * grid_scale;
    const float pos_v = v * grid_scale;
    const size_t wi = (size_t)roundf(sqrtf(fabsf(w * w_scale)));
    const float ru = roundf(pos_u), rv = roundf(pos_v);
    s->conj_neg = (w > 0.0f);
    s->grid_u = (int)ru + centre;
    s->grid_v = (int)rv + centre;
    s->off_u = (int)roundf((ru - pos_u) * oversample);
    s->off_v = (int)roundf((rv - pos_v) * oversample);
    s->plane = wi < num_w_planes ? wi : num_w_planes - 1;
    s->support = support[s->plane];
    s->skip = (s->grid_u + s->support >= grid_size ||
            s->grid_u - s->support < 0 ||
            s->grid_v + s->support >= grid_size ||
            s->grid_v - s->support < 0);
}

static void select_d(double u, double v, double w, double grid_scale,
        double w_scale, int grid_size, int oversample, size_t num_w_planes,
        const int* support, oracle_sel* s)
{
    const int centre = grid_size / 2;
    const double pos_u = -u * grid_scale;
    const double pos_v = v * grid_scale;
    const size_t wi = (size_t)round(sqrt(fabs(w * w_scale)));
    const double ru = round(pos_u), rv = round(pos_v);
    s->conj_neg = (w > 0.0);
    s->grid_u = (int)ru + centre;
    s->grid_v = (int)rv + centre;
    s->off_u = (int)round((ru - pos_u) * oversample);
    s->off_v = (int)round((rv - pos_v) * oversample);
    s->plane = wi < num_w_planes ? wi : num_w_planes - 1;
    s->support = support[s->plane];
    s->skip = (s->grid_u + s->support >= grid_size ||
            s->grid_u - s->support < 0 ||
            s->grid_v + s->support >= grid_size ||
            s->grid_v - s->support < 0);
}

static void unpack(const oracle_sel* s, size_t i, int* gu, int* gv, int* ou,
        int* ov, int* plane, int* sup, int* conj_neg, int* skip)
{
    if (gu) gu[i] = s->grid_u;
    if (gv) gv[i] = s->grid_v;
    if (ou) ou[i] = s->off_u;
    if (ov) ov[i] = s->off_v;
    if (plane) plane[i] = (int)s->plane;
    if (sup) sup[i] = s->support;
    if (conj_neg) conj_neg[i] = s->conj_neg;
    if (skip) skip[i] = s->skip;
}

ORACLE_API void oracle_select_f(size_t n, const float* uu, const float* vv,
        const float* ww, size_t num_w_planes, const int* support,
        int oversample, float cell_size_rad, float w_scale, int grid_size,
        int* gu, int* gv, int* ou, int* ov, int* plane, int* sup,
        int* conj_neg, int* skip)
{
    const float grid_scale = grid_size * cell_size_rad; /* :148 */
    size_t i;
    for (i = 0; i < n; ++i)
    {
        oracle_sel s;
        select_f(uu[i], vv[i], ww[i], grid_scale, w_scale, grid_size,
                oversample, num_w_planes, support, &s);
        unpack(&s, i, gu, gv, ou, ov, plane, sup, conj_neg, skip);
    }
}

ORACLE_API void oracle_select_d(size_t n, const double* uu, const double* vv,
        const double* ww, size_t num_w_planes, const int* support,
        int oversample, double cell_size_rad, double w_scale, int grid_size,
        int* gu, int* gv, int* ou, int* ov, int* plane, int* sup,
        int* conj_neg, int* skip)
{
    const double grid_scale = grid_size * cell_size_rad; /* :59 */
    size_t i;
    for (i = 0; i < n; ++i)
    {
        oracle_sel s;
        select_d(uu[i], vv[i], ww[i], grid_scale, w_scale, grid_size,
                oversample, num_w_planes, support, &s);
        unpack(&s, i, gu, gv, ou, ov, plane, sup, conj_neg, skip);
    }
}

/* FP32 gridder, same signature and semantics as oskar_grid_wproj_f
 * (oskar_grid_wproj.h:113-131): num_skipped is set, norm and grid are
 * accumulated.  The footprint walk follows oskar_grid_wproj.c:190-210: rows
 * j outer, columns k inner, kernel tap at |off_v+j*ov| x |off_u+k*ov|,
 * complex MAC in float, sum of Re(tap) in double. */
ORACLE_API void oracle_grid_wproj_f(size_t num_w_planes, const int* support,
        int oversample, int conv_size_half, const float* conv_func,
        size_t num_points, const float* uu, const float* vv, const float* ww,
        const float* vis, const float* weight, float cell_size_rad,
        float w_scale, int grid_size, size_t* num_skipped, double* norm,
        float* grid)
{
    const size_t plane_len = (size_t)conv_size_half * conv_size_half;
    const float grid_scale = grid_size * cell_size_rad;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        oracle_sel s;
        int j, k;
        double tap_sum = 0.0;
        select_f(uu[i], vv[i], ww[i], grid_scale, w_scale, grid_size,
                oversample, num_w_planes, support, &s);
        if (s.skip) { *num_skipped += 1; continue; }
        {
            const float w_i = weight[i];
            const float a_re = w_i * vis[2 * i];
            const float a_im = w_i * vis[2 * i + 1];
            const float cj = s.conj_neg ? -1.0f : 1.0f;
            const float* plane = conv_func + 2 * s.plane * plane_len;
            for (j = -s.support; j <= s.support; ++j)
            {
                const float* krow = plane +
                        2 * (size_t)abs(s.off_v + j * oversample) * conv_size_half;
                float* grow = grid + 2 * ((size_t)(s.grid_v + j) * grid_size
                        + s.grid_u);
                for (k = -s.support; k <= s.support; ++k)
                {
                    const float* c = krow + 2 * abs(s.off_u + k * oversample);
                    const float c_re = c[0];
                    const float c_im = c[1] * cj;
                    grow[2 * k]     += (a_re * c_re - a_im * c_im);
                    grow[2 * k + 1] += (a_im * c_re + a_re * c_im);
                    tap_sum += c_re;
                }
            }
            *norm += tap_sum * w_i;
        }
    }
}

/* Norm and num_skipped of oracle_grid_wproj_f without the grid: the same
 * per-visibility double sum of Re(tap) over the footprint, rows j outer and
 * columns k inner (oskar_grid_wproj.c:190-207), added as sum * weight in
 * visibility order (:210).  The sum depends only on (plane, off_u, off_v),
 * so each distinct one is computed once, by the same loop, and reused: the
 * result is bit-identical to oracle_grid_wproj_f's norm.  For checking the
 * norm of workloads too large for the full serial gridder. */
ORACLE_API void oracle_norm_f(size_t num_w_planes, const int* support,
        int oversample, int conv_size_half, const float* conv_func,
        size_t num_points, const float* uu, const float* vv, const float* ww,
        const float* weight, float cell_size_rad, float w_scale, int grid_size,
        size_t* num_skipped, double* norm)
{
    const size_t plane_len = (size_t)conv_size_half * conv_size_half;
    const float grid_scale = grid_size * cell_size_rad;
    const int side = 2 * oversample + 1;          /* offsets in [-ov, ov] */
    double* memo = (double*)malloc(num_w_planes * side * side * sizeof(double));
    unsigned char* have = (unsigned char*)calloc(num_w_planes * side * side, 1);
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        oracle_sel s;
        size_t e;
        select_f(uu[i], vv[i], ww[i], grid_scale, w_scale, grid_size,
                oversample, num_w_planes, support, &s);
        if (s.skip) { *num_skipped += 1; continue; }
        e = (s.plane * side + (size_t)(s.off_v + oversample)) * side + (size_t)(s.off_u + oversample);
        if (!have[e])
        {
            const float* plane = conv_func + 2 * s.plane * plane_len;
            double tap_sum = 0.0;
            int j, k;
            for (j = -s.support; j <= s.support; ++j)
            {
                const float* krow = plane +
                        2 * (size_t)abs(s.off_v + j * oversample) * conv_size_half;
                for (k = -s.support; k <= s.support; ++k)
                    tap_sum += krow[2 * abs(s.off_u + k * oversample)];
            }
            memo[e] = tap_sum;
            have[e] = 1;
        }
        *norm += memo[e] * weight[i];
    }
    free(memo);
    free(have);
}

/* FP64 gridder, oskar_grid_wproj_d (oskar_grid_wproj.h:68-86,
 * oskar_grid_wproj.c:37-123). */
ORACLE_API void oracle_grid_wproj_d(size_t num_w_planes, const int* support,
        int oversample, int conv_size_half, const double* conv_func,
        size_t num_points, const double* uu, const double* vv,
        const double* ww, const double* vis, const double* weight,
        double cell_size_rad, double w_scale, int grid_size,
        size_t* num_skipped, double* norm, double* grid)
{
    const size_t plane_len = (size_t)conv_size_half * conv_size_half;
    const double grid_scale = grid_size * cell_size_rad;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        oracle_sel s;
        int j, k;
        double tap_sum = 0.0;
        select_d(uu[i], vv[i], ww[i], grid_scale, w_scale, grid_size,
                oversample, num_w_planes, support, &s);
        if (s.skip) { *num_skipped += 1; continue; }
        {
            const double w_i = weight[i];
            const double a_re = w_i * vis[2 * i];
            const double a_im = w_i * vis[2 * i + 1];
            const double cj = s.conj_neg ? -1.0 : 1.0;
            const double* plane = conv_func + 2 * s.plane * plane_len;
            for (j = -s.support; j <= s.support; ++j)
            {
                const double* krow = plane +
                        2 * (size_t)abs(s.off_v + j * oversample) * conv_size_half;
                double* grow = grid + 2 * ((size_t)(s.grid_v + j) * grid_size
                        + s.grid_u);
                for (k = -s.support; k <= s.support; ++k)
                {
                    const double* c = krow + 2 * abs(s.off_u + k * oversample);
                    const double c_re = c[0];
                    const double c_im = c[1] * cj;
                    grow[2 * k]     += (a_re * c_re - a_im * c_im);
                    grow[2 * k + 1] += (a_im * c_re + a_re * c_im);
                    tap_sum += c_re;
                }
            }
            *norm += tap_sum * w_i;
        }
    }
}

/* FP64 truth for the FP32 gridder: selection exactly as oskar_grid_wproj_f,
 * accumulation of the same (float) inputs in double.  No reference
 * counterpart; it coincides with oskar_grid_wproj_d wherever the float and
 * double selections agree. */
ORACLE_API void oracle_grid_wproj_fsel_d(size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const float* conv_func, size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, float cell_size_rad, float w_scale,
        int grid_size, size_t* num_skipped, double* norm, double* grid)
{
    const size_t plane_len = (size_t)conv_size_half * conv_size_half;
    const float grid_scale = grid_size * cell_size_rad;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        oracle_sel s;
        int j, k;
        double tap_sum = 0.0;
        select_f(uu[i], vv[i], ww[i], grid_scale, w_scale, grid_size,
                oversample, num_w_planes, support, &s);
        if (s.skip) { *num_skipped += 1; continue; }
        {
            const double w_i = weight[i];
            const double a_re = w_i * (double)vis[2 * i];
            const double a_im = w_i * (double)vis[2 * i + 1];
            const double cj = s.conj_neg ? -1.0 : 1.0;
            const float* plane = conv_func + 2 * s.plane * plane_len;
            for (j = -s.support; j <= s.support; ++j)
            {
                const float* krow = plane +
                        2 * (size_t)abs(s.off_v + j * oversample) * conv_size_half;
                double* grow = grid + 2 * ((size_t)(s.grid_v + j) * grid_size
                        + s.grid_u);
                for (k = -s.support; k <= s.support; ++k)
                {
                    const float* c = krow + 2 * abs(s.off_u + k * oversample);
                    const double c_re = c[0];
                    const double c_im = (double)c[1] * cj;
                    grow[2 * k]     += (a_re * c_re - a_im * c_im);
                    grow[2 * k + 1] += (a_im * c_re + a_re * c_im);
                    tap_sum += c_re;
                }
            }
            *norm += tap_sum * w_i;
        }
    }
}

/* Exact adjoint (FP64) of oracle_grid_wproj_fsel_d with respect to the
 * complex amplitudes:  out[i] = w_i * sum_taps conj(c'_i) * grid[cell],
 * where c' is the tap after the conjugation flip.  Skipped visibilities get 0.
 * With apply_weight = 0 the factor w_i is left out. */
ORACLE_API void oracle_degrid_wproj_fsel_d(size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const float* conv_func, size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* weight,
        float cell_size_rad, float w_scale, int grid_size,
        const double* grid, int apply_weight, double* out)
{
    const size_t plane_len = (size_t)conv_size_half * conv_size_half;
    const float grid_scale = grid_size * cell_size_rad;
    size_t i;
    for (i = 0; i < num_points; ++i)
    {
        oracle_sel s;
        int j, k;
        double acc_re = 0.0, acc_im = 0.0;
        select_f(uu[i], vv[i], ww[i], grid_scale, w_scale, grid_size,
                oversample, num_w_planes, support, &s);
        out[2 * i] = 0.0;
        out[2 * i + 1] = 0.0;
        if (s.skip) continue;
        {
            const double cj = s.conj_neg ? -1.0 : 1.0;
            const float* plane = conv_func + 2 * s.plane * plane_len;
            const double w_i = apply_weight ? (double)weight[i] : 1.0;
            for (j = -s.support; j <= s.support; ++j)
            {
                const float* krow = plane +
                        2 * (size_t)abs(s.off_v + j * oversample) * conv_size_half;
                const double* grow = grid + 2 * ((size_t)(s.grid_v + j) *
                        grid_size + s.grid_u);
                for (k = -s.support; k <= s.support; ++k)
                {
                    const float* c = krow + 2 * abs(s.off_u + k * oversample);
                    const double c_re = c[0];
                    const double c_im = (double)c[1] * cj;
                    const double g_re = grow[2 * k], g_im = grow[2 * k + 1];
                    acc_re += c_re * g_re + c_im * g_im;
                    acc_im += c_re * g_im - c_im * g_re;
                }
            }
            out[2 * i] = w_i * acc_re;
            out[2 * i + 1] = w_i * acc_im;
        }
    }
}

/* ---- Tile-restricted FP64 truth for spot checks of large grids. ---- */

typedef struct
{
    size_t num_w_planes; const int* support; int oversample;
    int conv_size_half; const float* conv_func; size_t num_points;
    const float* uu; const float* vv; const float* ww; const float* vis;
    const float* weight; float cell_size_rad; float w_scale; int grid_size;
    int tile; const int* tiles_xy; size_t num_tiles; double* out;
    size_t next; pthread_mutex_t lock;
} tile_job;

static void* tile_worker(void* arg)
{
    tile_job* J = (tile_job*)arg;
    const size_t plane_len = (size_t)J->conv_size_half * J->conv_size_half;
    const float grid_scale = J->grid_size * J->cell_size_rad;
    const int T = J->tile;
    for (;;)
    {
        size_t t, i;
        pthread_mutex_lock(&J->lock);
        t = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (t >= J->num_tiles) break;
        {
            const int x0 = J->tiles_xy[2 * t] * T, y0 = J->tiles_xy[2 * t + 1] * T;
            double* o = J->out + 2 * (size_t)T * T * t;
            memset(o, 0, sizeof(double) * 2 * (size_t)T * T);
            for (i = 0; i < J->num_points; ++i)
            {
                oracle_sel s;
                int j, k, jlo, jhi, klo, khi;
                select_f(J->uu[i], J->vv[i], J->ww[i], grid_scale, J->w_scale,
                        J->grid_size, J->oversample, J->num_w_planes,
                        J->support, &s);
                if (s.skip) continue;
                jlo = y0 - s.grid_v; if (jlo < -s.support) jlo = -s.support;
                jhi = y0 + T - 1 - s.grid_v; if (jhi > s.support) jhi = s.support;
                klo = x0 - s.grid_u; if (klo < -s.support) klo = -s.support;
                khi = x0 + T - 1 - s.grid_u; if (khi > s.support) khi = s.support;
                if (jlo > jhi || klo > khi) continue;
                {
                    const double w_i = J->weight[i];
                    const double a_re = w_i * (double)J->vis[2 * i];
                    const double a_im = w_i * (double)J->vis[2 * i + 1];
                    const double cj = s.conj_neg ? -1.0 : 1.0;
                    const float* plane = J->conv_func + 2 * s.plane * plane_len;
                    for (j = jlo; j <= jhi; ++j)
                    {
                        const float* krow = plane + 2 * (size_t)abs(
                                s.off_v + j * J->oversample) * J->conv_size_half;
                        double* orow = o + 2 * (size_t)(s.grid_v + j - y0) * T;
                        for (k = klo; k <= khi; ++k)
                        {
                            const float* c = krow + 2 * abs(s.off_u + k * J->oversample);
                            const double c_re = c[0];
                            const double c_im = (double)c[1] * cj;
                            const int x = s.grid_u + k - x0;
                            orow[2 * x]     += (a_re * c_re - a_im * c_im);
                            orow[2 * x + 1] += (a_im * c_re + a_re * c_im);
                        }
                    }
                }
            }
        }
    }
    return NULL;
}

/* FP64 truth (as oracle_grid_wproj_fsel_d) for the listed tiles only:
 * tiles_xy holds (tile_x, tile_y) pairs of tile x tile cells; out receives
 * num_tiles blocks of tile*tile complex doubles, row-major.  Contributions
 * only (the block starts at zero).  Uses num_threads pthreads. */
ORACLE_API void oracle_grid_wproj_fsel_d_tiles(size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const float* conv_func, size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, float cell_size_rad, float w_scale,
        int grid_size, int tile, const int* tiles_xy, size_t num_tiles,
        int num_threads, double* out)
{
    tile_job J;
    pthread_t th[256];
    int t;
    J.num_w_planes = num_w_planes; J.support = support;
    J.oversample = oversample; J.conv_size_half = conv_size_half;
    J.conv_func = conv_func; J.num_points = num_points;
    J.uu = uu; J.vv = vv; J.ww = ww; J.vis = vis; J.weight = weight;
    J.cell_size_rad = cell_size_rad; J.w_scale = w_scale;
    J.grid_size = grid_size; J.tile = tile; J.tiles_xy = tiles_xy;
    J.num_tiles = num_tiles; J.out = out; J.next = 0;
    pthread_mutex_init(&J.lock, NULL);
    if (num_threads < 1) num_threads = 1;
    if (num_threads > 256) num_threads = 256;
    for (t = 0; t < num_threads; ++t)
        pthread_create(&th[t], NULL, tile_worker, &J);
    for (t = 0; t < num_threads; ++t)
        pthread_join(th[t], NULL);
    pthread_mutex_destroy(&J.lock);
}


/* ------------------------------------------------------------------------
 * Uniform weighting: cell selection of oskar_grid_weights.c (:117-128 for
 * float, :48-59 for double), then the per-cell sum in visibility order
 * (write) or the division by the cell total (read).
 * ---------------------------------------------------------------------- */
static int weights_cell_f(float u, float v, float grid_scale, int grid_size, size_t* t)
{
    const int centre = grid_size / 2;
    const int grid_u = (int)roundf(-u * grid_scale) + centre;
    const int grid_v = (int)roundf(v * grid_scale) + centre;
    *t = (size_t)grid_v * (size_t)grid_size + (size_t)grid_u;
    return !(grid_u >= grid_size || grid_u < 0 || grid_v >= grid_size || grid_v < 0);
}

static int weights_cell_d(double u, double v, double grid_scale, int grid_size, size_t* t)
{
    const int centre = grid_size / 2;
    const int grid_u = (int)round(-u * grid_scale) + centre;
    const int grid_v = (int)round(v * grid_scale) + centre;
    *t = (size_t)grid_v * (size_t)grid_size + (size_t)grid_u;
    return !(grid_u >= grid_size || grid_u < 0 || grid_v >= grid_size || grid_v < 0);
}

ORACLE_API void oracle_grid_weights_write_f(size_t n, const float* uu, const float* vv,
        const float* weight, float cell_size_rad, int grid_size, size_t* num_skipped,
        float* grid)
{
    const float grid_scale = grid_size * cell_size_rad;
    size_t t;
    *num_skipped = 0;
    for (size_t i = 0; i < n; ++i)
    {
        if (!weights_cell_f(uu[i], vv[i], grid_scale, grid_size, &t)) { ++*num_skipped; continue; }
        grid[t] += weight[i];
    }
}

ORACLE_API void oracle_grid_weights_write_d(size_t n, const double* uu, const double* vv,
        const double* weight, double cell_size_rad, int grid_size, size_t* num_skipped,
        double* grid)
{
    const double grid_scale = grid_size * cell_size_rad;
    size_t t;
    *num_skipped = 0;
    for (size_t i = 0; i < n; ++i)
    {
        if (!weights_cell_d(uu[i], vv[i], grid_scale, grid_size, &t)) { ++*num_skipped; continue; }
        grid[t] += weight[i];
    }
}

ORACLE_API void oracle_grid_weights_read_f(size_t n, const float* uu, const float* vv,
        const float* weight_in, float* weight_out, float cell_size_rad, int grid_size,
        size_t* num_skipped, const float* grid)
{
    const float grid_scale = grid_size * cell_size_rad;
    size_t t;
    *num_skipped = 0;
    for (size_t i = 0; i < n; ++i)
    {
        if (!weights_cell_f(uu[i], vv[i], grid_scale, grid_size, &t)) { ++*num_skipped; continue; }
        weight_out[i] = (grid[t] != 0.0) ? weight_in[i] / grid[t] : 0.0f;
    }
}

ORACLE_API void oracle_grid_weights_read_d(size_t n, const double* uu, const double* vv,
        const double* weight_in, double* weight_out, double cell_size_rad, int grid_size,
        size_t* num_skipped, const double* grid)
{
    const double grid_scale = grid_size * cell_size_rad;
    size_t t;
    *num_skipped = 0;
    for (size_t i = 0; i < n; ++i)
    {
        if (!weights_cell_d(uu[i], vv[i], grid_scale, grid_size, &t)) { ++*num_skipped; continue; }
        weight_out[i] = (grid[t] != 0.0) ? weight_in[i] / grid[t] : 0.0;
    }
}

/* ------------------------------------------------------------------------
 * Grid -> image.  For an even grid size G, c = G/2:
 *   image[y][x] = Re( sum_{v,u} grid[v][u]
 *                     * exp(+2 pi i ((u - c)(x - c) + (v - c)(y - c)) / G) )
 *                 * scale,   scale = 1/norm (1 when norm is 0).
 * Separable direct summation (rows, then columns) in FP64, with the phase
 * index ((u - c)(x - c)) mod G taken exactly in integers.
 * grid: G*G complex values, interleaved (re, im), row v, column u.
 * ---------------------------------------------------------------------- */
ORACLE_API int oracle_grid_to_image_d(int G, const double* grid, double norm, double* image)
{
    if (G < 2 || (G & 1)) return 1;
    const long long c = G / 2;
    const double scale = norm != 0.0 ? 1.0 / norm : 1.0;
    double* cs = (double*) malloc(2 * (size_t)G * sizeof(double));
    double* tmp = (double*) malloc(2 * (size_t)G * G * sizeof(double));
    if (!cs || !tmp) { free(cs); free(tmp); return 2; }
    for (int k = 0; k < G; ++k)
    {
        const double a = 2.0 * M_PI * (double)k / (double)G;
        cs[2 * k] = cos(a);
        cs[2 * k + 1] = sin(a);
    }
    /* rows: tmp[v][x] = sum_u grid[v][u] e^{2 pi i (u-c)(x-c)/G} */
    for (int v = 0; v < G; ++v)
        for (int x = 0; x < G; ++x)
        {
            double re = 0.0, im = 0.0;
            for (int u = 0; u < G; ++u)
            {
                long long k = ((long long)(u - c) * (long long)(x - c)) % G;
                if (k < 0) k += G;
                const double gr = grid[2 * ((size_t)v * G + u)];
                const double gi = grid[2 * ((size_t)v * G + u) + 1];
                re += gr * cs[2 * k] - gi * cs[2 * k + 1];
                im += gr * cs[2 * k + 1] + gi * cs[2 * k];
            }
            tmp[2 * ((size_t)v * G + x)] = re;
            tmp[2 * ((size_t)v * G + x) + 1] = im;
        }
    /* columns, real part only */
    for (int y = 0; y < G; ++y)
        for (int x = 0; x < G; ++x)
        {
            double re = 0.0;
            for (int v = 0; v < G; ++v)
            {
                long long k = ((long long)(v - c) * (long long)(y - c)) % G;
                if (k < 0) k += G;
                re += tmp[2 * ((size_t)v * G + x)] * cs[2 * k] -
                      tmp[2 * ((size_t)v * G + x) + 1] * cs[2 * k + 1];
            }
            image[(size_t)y * G + x] = re * scale;
        }
    free(cs);
    free(tmp);
    return 0;
}
