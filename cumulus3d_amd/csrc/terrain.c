/*
 * terrain.c — seeded synthetic aerial scenes for the PatchStereo engine's tests and benchmarks,
 * rendered on the host (OpenMP over rows). The scene model is terrain_core.h, which the device
 * renderer (terrain.hip) shares.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define TERRAIN_FN static inline
#include "terrain_core.h"

/* Renders `ncam` cameras. img/depth are ncam*h*w floats (h, w of each camera, packed one after
 * the other), normal 3 floats per pixel. Returns 0. */
int c3d_render_terrain(const c3d_terrain_params* p, const c3d_terrain_camera* cams, int ncam, float* img,
                       float* depth, float* normal) {
    size_t off = 0;
    for (int c = 0; c < ncam; ++c) {
        const c3d_terrain_camera* cm = &cams[c];
        double C[3];
        terrain_center(cm, C);
        const int w = cm->width, h = cm->height;
#pragma omp parallel for schedule(dynamic, 2)
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                size_t pi = off + (size_t)y * w + x;
                terrain_pixel(p, cm, C, x, y, &img[pi], &depth[pi], &normal[3 * pi]);
            }
        }
        off += (size_t)w * h;
    }
    return 0;
}
