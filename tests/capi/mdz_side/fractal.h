/*
 * Test-side declarations for building mdz_amd/shim/fractal_gpu.c without
 * MDZ: only the image_info, coords and Julia fields the shim reads, under
 * the member names it uses (MDZ's own image_info holds many more, and its
 * header needs GTK).  Not MDZ's header.
 */
#ifndef MDZ_SIDE_FRACTAL_H
#define MDZ_SIDE_FRACTAL_H

#include <stdint.h>

#include <gmp.h>
#include <mpfr.h>

typedef int64_t depth_t;

enum { FAMILY_MANDEL = 0, FAMILY_JULIA = 1 };

typedef struct coords {
    mpfr_t xmin, xmax, ymax, width;
} coords;

typedef struct julia_info {
    mpfr_t c_re, c_im;
} julia_info;

typedef struct image_info {
    mpfr_t xmin, xmax, ymax, width;
    coords* pcoords;
    depth_t depth;
    int* raw_data;
    int real_width, real_height;
    int family, fractal;
    union {
        julia_info julia;
    } u;
    mpfr_prec_t precision;
    void* rth_ptr;
} image_info;

int fractal_gpu_mpfr_calculate_line(image_info* img, int line);
int fractal_gpu_ld_calculate_line(image_info* img, int line);
int fractal_gpu_gmp_calculate_line(image_info* img, int line);
void fractal_gpu_shutdown(void);

#endif
