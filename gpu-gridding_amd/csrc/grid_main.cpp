// grid_main.cpp -- command-line driver with the reference harness's usage
// (main.cpp:80-88): `mi355_grid_main <test_data_root_name> [--check]
// [--image FILE.fits] [--device N]` grids <root>.vis with the kernels in
// <root>_KERNELS_REAL/_IMAG.fits on the GPU (uniform weighting, as main.cpp)
// and reports what main.cpp reports.
#include "mi355_io.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv)
{
    const char* root = nullptr;
    const char* image = nullptr;
    int check = 0, device = 0;
    for (int i = 1; i < argc; ++i)
    {
        if (!strcmp(argv[i], "--check")) check = 1;
        else if (!strcmp(argv[i], "--image") && i + 1 < argc) image = argv[++i];
        else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (!root) root = argv[i];
        else root = nullptr, i = argc;
    }
    if (!root)
    {
        fprintf(stderr, "Usage: %s <test_data_root_name> [--check] [--image FILE.fits] "
                "[--device N]\n", argv[0]);
        return EXIT_FAILURE;
    }
    mi355_ctx* ctx = nullptr;
    if (mi355_ctx_create(device, &ctx) != MI355_OK)
    {
        fprintf(stderr, "Error: %s\n", mi355_grid_last_error_message());
        return EXIT_FAILURE;
    }
    mi355_driver_result r;
    printf("Gridding '%s.vis' on device %d...\n", root, device);
    const int st = mi355_grid_vis_file(ctx, root, check, nullptr, 0, image, &r);
    if (st != MI355_OK)
    {
        fprintf(stderr, "Error (code %d): %s\n", st, mi355_grid_last_error_message());
        mi355_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    printf("%llu visibilities in %d blocks, %d w-planes, grid %d x %d (%s)\n",
            r.num_points, r.num_blocks, r.num_w_planes, r.grid_size, r.grid_size,
            r.precision == MI355_BIN_DOUBLE ? "double" : "single");
    printf("num_skipped (last block) = %llu, total %llu; norm = %.17g\n", r.num_skipped_last,
            r.num_skipped_total, r.norm);
    if (check)
        printf("Checking grids: max|G - G64| / max|G64| = %.3e, L2 = %.3e (%s)\n",
                r.check_max_rel, r.check_l2_rel, r.check_max_rel <= 1e-5 ? "ok" : "FAILED");
    printf("Reading the file took %.3f seconds\n", r.seconds_read);
    printf("Generating weights took %.3f seconds (GPU, copies included)\n", r.seconds_weights);
    printf("Gridding visibilities took %.3f seconds (GPU, copies included)\n", r.seconds_grid);
    printf("Both passes took %.3f seconds (wall clock, file reads included)\n", r.seconds_wall);
    if (image) printf("Image written to '%s'\n", image);
    mi355_ctx_destroy(ctx);
    return check && !(r.check_max_rel <= 1e-5) ? EXIT_FAILURE : EXIT_SUCCESS;
}
