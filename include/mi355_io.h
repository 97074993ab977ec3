/*
 * mi355_io.h -- reading and writing the reference's input files, and the
 * end-to-end gridding driver over them (part of libmi355_gridder.so).
 *
 * The reference's harness (main.cpp:55-453) reads
 *   <root>.vis                  an OSKAR binary file (tagged chunks, CRC-32C)
 *                               of visibility blocks,
 *   <root>_KERNELS_REAL.fits    the real and imaginary parts of the W-kernel
 *   <root>_KERNELS_IMAG.fits    stack (3-D FITS cubes, header keywords and a
 *                               W_KERNELS binary table of supports),
 * through the OSKAR binary library (utils/binary) and cfitsio.  This file
 * declares the library's own readers and writers for those formats, with
 * the reference functions' argument meaning and error behaviour, and the
 * driver that runs main.cpp's whole flow (uniform weighting, then gridding
 * block by block) on the GPU.
 *
 * Status codes are those of mi355_gridder.h; I/O failures return
 * MI355_ERR_IO (open / read / write / seek) or MI355_ERR_FORMAT (not the
 * expected format, tag missing, CRC mismatch), with a message in
 * mi355_grid_last_error_message().
 */
#ifndef MI355_IO_H_
#define MI355_IO_H_

#include <stddef.h>

#include "mi355_gridder.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * OSKAR binary files.
 *
 * A file is a 64-byte header ("OSKARBIN", format version) followed by
 * chunks: a 20-byte tag (type, group and tag ids or extended names, user
 * index, block size), the payload and an optional CRC-32C.  Opening for
 * reading indexes every chunk; queries search the index from the "query
 * search start" onwards (oskar_binary_query.c:48-80), as main.cpp relies on
 * to read block i's tags (main.cpp:214, :253).
 * ---------------------------------------------------------------------- */
typedef struct mi355_bin mi355_bin;

/* Data type codes of a chunk's payload (oskar_binary_data_types.h:43-74). */
#define MI355_BIN_CHAR    0x01
#define MI355_BIN_INT     0x02
#define MI355_BIN_SINGLE  0x04
#define MI355_BIN_DOUBLE  0x08
#define MI355_BIN_COMPLEX 0x20
#define MI355_BIN_MATRIX  0x40

/* Replaces oskar_binary_create (oskar_binary_create.h; create.c:49-290):
 * mode 'r' reads the header and indexes the chunks, 'w' creates the file
 * and writes the header (format version 2). */
MI355_EXPORT int mi355_bin_open(const char* path, char mode, mi355_bin** handle);
/* Replaces oskar_binary_free. */
MI355_EXPORT int mi355_bin_close(mi355_bin* handle);
/* Number of chunks indexed (oskar_binary_num_tags). */
MI355_EXPORT int mi355_bin_num_tags(const mi355_bin* handle);
/* Replaces oskar_binary_set_query_search_start (query.c:137-146). */
MI355_EXPORT int mi355_bin_set_query_search_start(mi355_bin* handle, int start);
/* Replaces oskar_binary_query (query.c:48-80): index of the first standard
 * chunk at or after the search start with this group, tag and user index
 * (and data type, unless data_type is 0). */
MI355_EXPORT int mi355_bin_query(const mi355_bin* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t* payload_size, int* chunk_index);
/* Replaces oskar_binary_read (read.c:110-127): the payload of that chunk
 * into data (data_size bytes available), CRC-checked when the chunk has a
 * CRC. */
MI355_EXPORT int mi355_bin_read(mi355_bin* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, void* data);
/* Replaces oskar_binary_write (write.c:38-124): appends a standard chunk
 * with a CRC-32C. */
MI355_EXPORT int mi355_bin_write(mi355_bin* handle, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, const void* data);

/* Tag groups and tags of a visibility file that main.cpp reads
 * (oskar_binary.h:57-71; read_vis.h:39-59). */
#define MI355_VIS_GROUP_HEADER                 11
#define MI355_VIS_GROUP_BLOCK                  12
#define MI355_VIS_HEADER_NUM_TAGS_PER_BLOCK     2
#define MI355_VIS_HEADER_AMP_TYPE               5
#define MI355_VIS_HEADER_COORD_PRECISION        6
#define MI355_VIS_HEADER_MAX_TIMES_PER_BLOCK    7
#define MI355_VIS_HEADER_NUM_TIMES_TOTAL        8
#define MI355_VIS_HEADER_NUM_CHANNELS_TOTAL    10
#define MI355_VIS_HEADER_NUM_STATIONS          11
#define MI355_VIS_HEADER_FREQ_START_HZ         23
#define MI355_VIS_BLOCK_DIM_START_AND_SIZE      1
#define MI355_VIS_BLOCK_CROSS_CORRELATIONS      3
#define MI355_VIS_BLOCK_BASELINE_UU             4
#define MI355_VIS_BLOCK_BASELINE_VV             5
#define MI355_VIS_BLOCK_BASELINE_WW             6

/* The visibility header fields main.cpp reads (main.cpp:120-160), and what
 * it derives from them. */
typedef struct mi355_vis_header
{
    int num_tags_per_block;
    int amp_type;               /* MI355_BIN_* type of the amplitudes */
    int coord_precision;        /* MI355_BIN_SINGLE or MI355_BIN_DOUBLE */
    int max_times_per_block;
    int num_times_total;
    int num_channels_total;
    int num_stations;
    double freq_start_hz;
    int num_baselines;          /* num_stations * (num_stations - 1) / 2 */
    int num_blocks;             /* ceil(num_times_total / max_times_per_block) */
} mi355_vis_header;

MI355_EXPORT int mi355_vis_read_header(mi355_bin* handle, mi355_vis_header* header);

/* Block i of the file, as main.cpp reads it (main.cpp:213-227, :252-271):
 * the block's time count (DIM_START_AND_SIZE[2]) times num_baselines
 * points; uu, vv, ww in the coordinate precision, scaled from metres to
 * wavelengths as scale_coords does (read_vis.c:49-73: freq / c in double,
 * then cast to float for single precision); vis (may be NULL) receives the
 * complex amplitudes in the amplitude precision.  capacity = points the
 * arrays hold. */
MI355_EXPORT int mi355_vis_read_block(mi355_bin* handle, const mi355_vis_header* header,
        int block, size_t capacity, size_t* block_size, void* uu, void* vv, void* ww,
        void* vis);

/* ------------------------------------------------------------------------
 * W-kernel FITS cubes.
 * ---------------------------------------------------------------------- */
typedef struct mi355_kernel_info
{
    int conv_size_half;         /* NAXIS1 (= NAXIS2) */
    int num_w_planes;           /* NAXIS3 */
    int oversample;             /* OVERSAMP */
    int grid_size;              /* GRIDSIZE */
    double cellsize_rad;        /* CELLSIZE */
    double w_scale;             /* W_SCALE */
} mi355_kernel_info;

/* Replaces read_kernel (read_kernel.c:35-118): a 3-D primary image of
 * conv_size_half^2 * num_w_planes values, converted to precision
 * (MI355_BIN_SINGLE or MI355_BIN_DOUBLE) after BSCALE / BZERO, into a new
 * array *data; when support is not NULL, the first column of the binary
 * table extension named W_KERNELS (num_w_planes values) into a new int
 * array *support.  Arrays are released with mi355_io_free. */
MI355_EXPORT int mi355_fits_read_kernel(const char* path, int precision,
        mi355_kernel_info* info, void** data, int** support);
/* Writes such a file: a FLOAT_IMG (single) or DOUBLE_IMG (double) cube with
 * the four keywords and, when support is not NULL, the W_KERNELS table with
 * one 32-bit integer column SUPPORT. */
MI355_EXPORT int mi355_fits_write_kernel(const char* path, int precision,
        const mi355_kernel_info* info, const void* data, const int* support);
/* Writes a 2-D FITS image (FLOAT_IMG or DOUBLE_IMG), e.g. the output of
 * mi355_grid_to_image_*. */
MI355_EXPORT int mi355_fits_write_image(const char* path, int precision, int width,
        int height, const void* data);
MI355_EXPORT void mi355_io_free(void* p);

/* ------------------------------------------------------------------------
 * Driver: main.cpp's flow on the GPU.
 *
 * Opens <root>.vis and <root>_KERNELS_REAL/_IMAG.fits, checks for single-
 * channel, single-polarisation data (main.cpp:162-171), interleaves the
 * kernel parts and sets unit input weights (:186-206), then
 *   pass 1: every block's coordinates -> uniform-weighting grid
 *           (mi355_grid_weights_write_*_dev, :202-239);
 *   pass 2: every block's coordinates and amplitudes -> weights read
 *           (:273-286), then gridding (mi355_grid_wproj_*_dev, :288-338),
 * with the weights grid, the visibility grid and norm resident on the
 * device.  The amplitude precision picks FP32 or FP64 throughout, as in
 * main.cpp.  With check set, every block is also gridded with FP64
 * accumulation (FP32 data: mi355_grid_wproj_fsel_d_dev) and the grid's
 * norm-wise error against it is reported.  Everything runs in order on
 * the context's stream (mi355_ctx_set_stream): blocks are read into two
 * alternating pinned buffers, block b+1 from the file while block b's
 * copies and passes run, and the per-block num_skipped stay on the device
 * until the end, so the loops do not wait for the device between blocks.
 * The context's options (kernel choice, item size ...) apply.
 * ---------------------------------------------------------------------- */
typedef struct mi355_driver_result
{
    int num_blocks, grid_size, num_w_planes, precision;   /* MI355_BIN_SINGLE / _DOUBLE */
    unsigned long long num_points;
    unsigned long long num_skipped_total;                 /* over all blocks */
    unsigned long long num_skipped_last;                  /* last block's, as main.cpp reports */
    double norm;
    double seconds_read;                                  /* host file reads */
    double seconds_weights, seconds_grid;                 /* per pass: sum over blocks of the
                                                           * device time from a block's
                                                           * copies to its last pass (events) */
    double check_max_rel, check_l2_rel;                   /* check only (else -1) */
    double seconds_wall;                                  /* both passes, reads included */
} mi355_driver_result;

/* grid: host array of grid_size^2 complex values in the amplitude precision
 * that receives the grid (may be NULL); image_path: when not NULL, the
 * image of the grid (mi355_grid_to_image_*, scaled by 1/norm) is written
 * there as FITS. */
MI355_EXPORT int mi355_grid_vis_file(mi355_ctx* ctx, const char* input_root, int check,
        void* grid, size_t grid_capacity, const char* image_path,
        mi355_driver_result* result);

#ifdef __cplusplus
}
#endif

#endif /* MI355_IO_H_ */
