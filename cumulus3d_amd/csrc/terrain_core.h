/*
 * terrain_core.h — the per-pixel model of the seeded synthetic aerial scenes, shared by the host
 * renderer (terrain.c, OpenMP) and the device renderer (terrain.hip, one thread per pixel).
 *
 * A procedural heightfield (fBm of value noise) over a square patch, with a procedural albedo and
 * Lambertian shading (view-independent, so the views are photo-consistent), ray-marched from
 * pinhole cameras. Per pixel: a float image value in [0,1] (2x2 supersampled), the ground-truth
 * camera-z depth of the pixel-centre ray (0 where the ray leaves the patch or sees the sky) and
 * the ground-truth normal in camera coordinates.
 *
 * The including file defines TERRAIN_FN, the qualifiers of every function here (plain static
 * inline for the host build, static __device__ inline for the device build). Everything is double
 * precision, in the same operation order for both builds.
 *
 * The reference has no terrain generator; its synthetic scene is a textured floor with boxes
 * (crates/dense/src/synthetic.rs:108-292), which cumulus3d_amd/scenes.py restates. Pixel
 * convention: integer pixel = pixel centre (crates/dense/src/scene.rs:34-36).
 */
#ifndef C3D_TERRAIN_CORE_H
#define C3D_TERRAIN_CORE_H

#include <stdint.h>

typedef struct c3d_terrain_params {
    uint64_t seed;
    double extent;          /* side of the square patch, metres (centred on the origin) */
    double relief;          /* maximum |height|, metres */
    double base_wavelength; /* wavelength of the first fBm octave, metres */
    int32_t octaves;
    int32_t pad0;
    double albedo_scale;    /* multiplies every albedo wavelength (1 = 0.3 m finest) */
    double sky;             /* image value of the sky and of the ground outside the patch */
} c3d_terrain_params;

typedef struct c3d_terrain_camera {
    double k[4]; /* fx, fy, cx, cy (integer-centre) */
    double r[9]; /* world -> camera rotation, row-major */
    double t[3];
    int32_t width;
    int32_t height;
} c3d_terrain_camera;

TERRAIN_FN uint64_t terrain_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

TERRAIN_FN double terrain_hash2(int64_t ix, int64_t iy, uint64_t seed) {
    uint64_t z = ((uint64_t)ix * 0x9e3779b97f4a7c15ULL) ^ ((uint64_t)iy * 0xc2b2ae3d27d4eb4fULL) ^ seed;
    return (double)(terrain_mix64(z) >> 11) / 9007199254740992.0;
}

TERRAIN_FN double value_noise(double x, double y, uint64_t seed) {
    double x0 = floor(x), y0 = floor(y);
    double fx = x - x0, fy = y - y0;
    double sx = fx * fx * (3.0 - 2.0 * fx), sy = fy * fy * (3.0 - 2.0 * fy);
    int64_t ix = (int64_t)x0, iy = (int64_t)y0;
    double a = terrain_hash2(ix, iy, seed), b = terrain_hash2(ix + 1, iy, seed);
    double c = terrain_hash2(ix, iy + 1, seed), d = terrain_hash2(ix + 1, iy + 1, seed);
    double top = a + (b - a) * sx, bot = c + (d - c) * sx;
    return top + (bot - top) * sy;
}

TERRAIN_FN double terrain_height(const c3d_terrain_params* p, double x, double y) {
    double amp = 1.0, wl = p->base_wavelength, s = 0.0, norm = 0.0;
    for (int o = 0; o < p->octaves; ++o) {
        s += amp * (2.0 * value_noise(x / wl, y / wl, p->seed + 101 * (uint64_t)o) - 1.0);
        norm += amp;
        amp *= 0.5;
        wl *= 0.5;
    }
    return p->relief * s / norm;
}

TERRAIN_FN double terrain_albedo(const c3d_terrain_params* p, double x, double y) {
    double sc = p->albedo_scale;
    uint64_t s = p->seed ^ 0xa1bed0ULL;
    double blocks = terrain_hash2((int64_t)floor(x / (0.6 * sc)), (int64_t)floor(y / (0.6 * sc)), s ^ 0x55);
    double fine = value_noise(x / (0.3 * sc), y / (0.3 * sc), s + 1);
    double mid = value_noise(x / (2.0 * sc), y / (2.0 * sc), s + 2);
    double coarse = value_noise(x / (8.0 * sc), y / (8.0 * sc), s + 3);
    double a = 0.35 * blocks + 0.25 * fine + 0.25 * mid + 0.15 * coarse;
    return a < 0 ? 0 : (a > 1 ? 1 : a);
}

TERRAIN_FN void terrain_normal(const c3d_terrain_params* p, double x, double y, double n[3]) {
    const double e = 0.005;
    double gx = (terrain_height(p, x + e, y) - terrain_height(p, x - e, y)) / (2 * e);
    double gy = (terrain_height(p, x, y + e) - terrain_height(p, x, y - e)) / (2 * e);
    double l = sqrt(gx * gx + gy * gy + 1.0);
    n[0] = -gx / l;
    n[1] = -gy / l;
    n[2] = 1.0 / l;
}

/* Marches the ray o + t d (|d| = 1). Returns 1 and the hit distance when it meets the terrain
 * inside the patch. */
TERRAIN_FN int terrain_march(const c3d_terrain_params* p, const double o[3], const double d[3], double* t_hit) {
    const double L = 1.3; /* bound on |grad h| of the fBm */
    double dxy = sqrt(d[0] * d[0] + d[1] * d[1]);
    double rate = fabs(d[2]) + L * dxy;
    double half = 0.5 * p->extent;
    double t = 0.0;
    for (int it = 0; it < 4000; ++it) {
        double x = o[0] + t * d[0], y = o[1] + t * d[1], z = o[2] + t * d[2];
        /* the patch is convex and every camera sits above it: a ray that leaves it never returns */
        if (fabs(x) > half || fabs(y) > half) return 0;
        if (z > p->relief + 1.0 && d[2] >= 0.0) return 0;
        double gap = z - terrain_height(p, x, y);
        if (gap < 1e-5) {
            *t_hit = t;
            return 1;
        }
        double step = gap / rate;
        t += step > 1e-5 ? step : 1e-5;
        if (t > 5000.0) return 0;
    }
    return 0;
}

TERRAIN_FN double terrain_shade(const c3d_terrain_params* p, double x, double y) {
    double n[3];
    terrain_normal(p, x, y, n);
    const double Lx = 0.3, Ly = 0.2, Lz = 1.0;
    double ll = sqrt(Lx * Lx + Ly * Ly + Lz * Lz);
    double nl = (n[0] * Lx + n[1] * Ly + n[2] * Lz) / ll;
    if (nl < 0) nl = 0;
    double a = terrain_albedo(p, x, y);
    return 0.05 + 0.9 * a * (0.75 + 0.25 * nl);
}

/* Renders pixel (x, y) of a camera whose centre is C: four sub-pixel rays for the image value and
 * the pixel-centre ray for depth and normal. */
TERRAIN_FN void terrain_pixel(const c3d_terrain_params* p, const c3d_terrain_camera* cm, const double C[3], int x, int y,
                              float* img, float* depth, float* normal) {
    const double* R = cm->r;
    double acc = 0.0;
    *depth = 0.0f;
    normal[0] = normal[1] = normal[2] = 0.0f;
    for (int s = 0; s < 5; ++s) {
        /* sub-pixel offsets (-0.25, -0.25), (0.25, -0.25), (-0.25, 0.25), (0.25, 0.25), then the centre */
        const double ox = s == 4 ? 0.0 : ((s & 1) ? 0.25 : -0.25);
        const double oy = s == 4 ? 0.0 : ((s & 2) ? 0.25 : -0.25);
        double rc[3] = {((double)x + ox - cm->k[2]) / cm->k[0], ((double)y + oy - cm->k[3]) / cm->k[1], 1.0};
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = R[0 * 3 + i] * rc[0] + R[1 * 3 + i] * rc[1] + R[2 * 3 + i] * rc[2];
        double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int i = 0; i < 3; ++i) d[i] /= l;
        double th;
        int hit = terrain_march(p, C, d, &th);
        if (s < 4) {
            acc += hit ? terrain_shade(p, C[0] + th * d[0], C[1] + th * d[1]) : p->sky;
        } else if (hit) {
            double P[3] = {C[0] + th * d[0], C[1] + th * d[1], C[2] + th * d[2]};
            double zc = R[6] * P[0] + R[7] * P[1] + R[8] * P[2] + cm->t[2];
            double nw[3], nc[3];
            terrain_normal(p, P[0], P[1], nw);
            for (int i = 0; i < 3; ++i) nc[i] = R[i * 3 + 0] * nw[0] + R[i * 3 + 1] * nw[1] + R[i * 3 + 2] * nw[2];
            *depth = (float)zc;
            normal[0] = (float)nc[0];
            normal[1] = (float)nc[1];
            normal[2] = (float)nc[2];
        }
    }
    *img = (float)(acc / 4.0);
}

/* Camera centre C = -R^T t. */
TERRAIN_FN void terrain_center(const c3d_terrain_camera* cm, double C[3]) {
    const double* R = cm->r;
    for (int i = 0; i < 3; ++i) C[i] = -(R[0 * 3 + i] * cm->t[0] + R[1 * 3 + i] * cm->t[1] + R[2 * 3 + i] * cm->t[2]);
}

#endif
