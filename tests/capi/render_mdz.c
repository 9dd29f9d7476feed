/* render_mdz FILE.mdz WIDTH OUT.png [AA]: a settings file rendered, coloured
   and written as a PNG through the C ABI alone -- the standalone form of
   `mdz -s FILE.mdz -w WIDTH -a AA -r OUT.png` (height from the file's aspect). */
#include <stdio.h>
#include <stdlib.h>
#include <mdzg.h>

static int render_mdz(const char* file, int w, int h, int aa, const char* png)
{
    mdzg_scene sc;
    mdzg_palette pal;
    char xmin[4096], xmax[4096], ymax[4096], width[4096];
    int rc = mdzg_scene_load(file, &sc);
    if (rc) return rc;
    if ((rc = mdzg_scene_palette(&sc, &pal)) || (rc = mdzg_scene_dims(&sc, &w, &h)) ||
        (rc = mdzg_scene_rect(&sc, w, h, 0, xmin, xmax, ymax, width, sizeof xmin))) {
        mdzg_scene_free(&sc);
        return rc;
    }
    mdzg_view v = { .fractal = sc.fractal, .family = sc.family, .path = sc.path,
                    .real_width = w * aa, .real_height = h * aa,
                    .precision = sc.precision, .depth = sc.depth,
                    .xmin = xmin, .xmax = xmax, .ymax = ymax, .width = width,
                    .julia_re = sc.julia_re, .julia_im = sc.julia_im };
    uint32_t* rgb = malloc(sizeof(uint32_t) * (size_t)w * h);
    mdzg_ctx* ctx;
    if (!rgb) rc = MDZG_ERR_NOMEM;
    else if (!(rc = mdzg_create(0, &ctx))) {
        rc = mdzg_render_image(ctx, &v, aa, &pal, rgb, NULL, NULL);   /* counts stay on the GPU */
        mdzg_destroy(ctx);
        if (!rc) rc = mdzg_png_write(png, rgb, w, h);
    }
    free(rgb);
    mdzg_scene_free(&sc);
    return rc;
}

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: render_mdz FILE.mdz WIDTH OUT.png [AA]\n");
        return 2;
    }
    const int rc = render_mdz(argv[1], atoi(argv[2]), 0, argc > 4 ? atoi(argv[4]) : 1, argv[3]);
    if (rc)
        fprintf(stderr, "render_mdz: %s (%d)\n", mdzg_strerror(rc), rc);
    return rc ? 1 : 0;
}
