/* Prints the C ABI as a C compiler sees include/mdzg.h: struct sizes and
   field offsets, enum values, and the results of the calls that need no
   device.  The CPU tests compare it with the ctypes mirror of mdz_amd. */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <mdzg.h>

#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))

int main(int argc, char** argv)
{
    printf("sizeof.mdzg_view %zu\n", sizeof(mdzg_view));
    F(mdzg_view, precision); F(mdzg_view, depth); F(mdzg_view, xmin); F(mdzg_view, julia_im);
    printf("sizeof.mdzg_tile %zu\n", sizeof(mdzg_tile));
    F(mdzg_tile, line0); F(mdzg_tile, counts);
    printf("sizeof.mdzg_stats %zu\n", sizeof(mdzg_stats));
    printf("sizeof.mdzg_palette %zu\n", sizeof(mdzg_palette));
    F(mdzg_palette, n); F(mdzg_palette, offset); F(mdzg_palette, colour_scale);
    F(mdzg_palette, interpolate);
    printf("sizeof.mdzg_scene %zu\n", sizeof(mdzg_scene));
    F(mdzg_scene, depth); F(mdzg_scene, aspect); F(mdzg_scene, precision); F(mdzg_scene, c0);
    F(mdzg_scene, julia_im); F(mdzg_scene, palette_offset); F(mdzg_scene, palette_n);
    F(mdzg_scene, palette);
    printf("sizeof.mdzg_server_cfg %zu\n", sizeof(mdzg_server_cfg));
    printf("sizeof.mdzg_req_info %zu\n", sizeof(mdzg_req_info));
    printf("sizeof.mdzg_serve_stats %zu\n", sizeof(mdzg_serve_stats));
    printf("enum.path %d %d %d %d\n", MDZG_PATH_MPFR, MDZG_PATH_LD, MDZG_PATH_DOUBLE, MDZG_PATH_GMP);
    printf("enum.fractal %d %d %d %d\n", MDZG_MANDELBROT, MDZG_BURNING_SHIP,
           MDZG_GENERALIZED_CELTIC, MDZG_VARIANT);
    printf("abi %d\n", mdzg_abi_version());
    char a[4096], b[4096], c[4096], d[4096];
    int rc = mdzg_rect_from_center("-0.75", "0", "3.0", 128, 240, 180, a, b, c, d, 4096);
    printf("rect %d %s %s %s %s\n", rc, a, b, c, d);
    if (argc > 1) {
        mdzg_scene sc;
        mdzg_palette pal;
        int w = 0, h = 0;
        rc = mdzg_scene_load(argv[1], &sc);
        printf("scene %d %d %d %lld %ld %d %s\n", rc, sc.fractal, sc.path, sc.depth, sc.precision,
               sc.centre, sc.c0);
        rc = mdzg_scene_dims(&sc, &w, &h);
        printf("dims %d %d %d\n", rc, w, h);
        rc = mdzg_scene_rect(&sc, w, h, 0, a, b, c, d, 4096);
        printf("srect %d %s %s %s %s\n", rc, a, b, c, d);
        rc = mdzg_scene_palette(&sc, &pal);
        printf("palette %d %d %d %.17g %d %u %u\n", rc, pal.n, pal.offset, pal.colour_scale,
               pal.interpolate, pal.colours[0], pal.colours[pal.n - 1]);
        mdzg_scene_free(&sc);
    }
    if (argc > 2) {
        uint32_t px[12];
        for (int i = 0; i < 12; ++i)
            px[i] = 0x010203u * (uint32_t)i;
        printf("png %d\n", mdzg_png_write(argv[2], px, 4, 3));
    }
    return 0;
}
