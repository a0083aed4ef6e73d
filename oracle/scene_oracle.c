/* scene_oracle.c — CPU restatement of the scene preparation's statistics (PairStats::new,
 * crates/dense/src/neighbors.rs:47-80; depth_ranges / depth_range_of, :162-189) — TEST
 * INFRASTRUCTURE ONLY: the checker of cumulus3d_amd/csrc/neighbors.hip. The selections
 * (select, :83-103) are restated in oracle/__init__.py; select_diverse is checked against a
 * brute force in tests/test_scene_oracle.py.
 *
 * As the reference: per pair of distinct views a list of f32 angles (appended track by track),
 * sorted in total order when its 75th percentile is asked for; per view a list of positive depths,
 * sorted. libm acos, double precision.
 *
 * Pinning: the reference's tests host_parts.rs:61-89 (selection orders [1,3], [1], [2,1,3], [0,3]
 * and the range 7.5..12.5) and neighbors.rs:194-198 (range 1.5..125 of 1..100), and brute-force
 * recomputation in numpy (tests/test_scene_oracle.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/c3d_scene.h"
#include "pm_oracle.h"

typedef struct {
    size_t n, cap;
    float* v;
} flist;

static void fpush(flist* l, float x) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 8;
        l->v = (float*)realloc(l->v, l->cap * sizeof(float));
    }
    l->v[l->n++] = x;
}

static void cam_centre(const double* r, const double* t, double c[3]) {
    c[0] = -(r[0] * t[0] + r[3] * t[1] + r[6] * t[2]);
    c[1] = -(r[1] * t[0] + r[4] * t[1] + r[7] * t[2]);
    c[2] = -(r[2] * t[0] + r[5] * t[1] + r[8] * t[2]);
}

static int cmp_d(const void* a, const void* b) {
    /* f64::total_cmp on the (non-negative or NaN) values here */
    const double x = *(const double*)a, y = *(const double*)b;
    const int nx = x != x, ny = y != y;
    if (nx || ny) return nx - ny;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int pmo_pair_stats(const c3d_sparse_scene* s, uint32_t* count, double* q75) {
    const uint32_t n = s->num_views;
    double* c = (double*)malloc(3 * (size_t)n * sizeof(double));
    for (uint32_t v = 0; v < n; ++v) cam_centre(s->rotation + 9 * (size_t)v, s->translation + 3 * (size_t)v, c + 3 * (size_t)v);
    flist* L = (flist*)calloc((size_t)n * n, sizeof(flist));
    for (uint64_t p = 0; p < s->num_points; ++p) {
        const double* x = s->xyz + 3 * p;
        const uint64_t a0 = s->track_offset[p], a1 = s->track_offset[p + 1];
        for (uint64_t i = a0; i < a1; ++i)
            for (uint64_t j = a0; j < i; ++j) {
                const uint32_t a = s->track_views[i], b = s->track_views[j];
                if (a == b) continue;
                const double* ca = c + 3 * (size_t)a;
                const double* cb = c + 3 * (size_t)b;
                const double u[3] = {ca[0] - x[0], ca[1] - x[1], ca[2] - x[2]};
                const double w[3] = {cb[0] - x[0], cb[1] - x[1], cb[2] - x[2]};
                const double nu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                const double nv = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                double ang;
                if (nu == 0.0 || nv == 0.0) {
                    ang = 0.0;
                } else {
                    double cc = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / (nu * nv);
                    if (cc < -1.0) cc = -1.0;
                    if (cc > 1.0) cc = 1.0;
                    const double t = acos(cc);
                    ang = fmin(t, 3.141592653589793 - t);
                }
                const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
                fpush(&L[(size_t)lo * n + hi], (float)ang);
            }
    }
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t b = 0; b < n; ++b) {
            const flist* l = &L[(size_t)(a < b ? a : b) * n + (a < b ? b : a)];
            count[(size_t)a * n + b] = a == b ? 0 : (uint32_t)l->n;
            double q = 0.0;
            if (a != b && l->n) {
                double* d = (double*)malloc(l->n * sizeof(double));
                for (size_t k = 0; k < l->n; ++k) d[k] = l->v[k];
                qsort(d, l->n, sizeof(double), cmp_d);
                q = pmo_quantile_sorted(d, l->n, 0.75);
                free(d);
            }
            q75[(size_t)a * n + b] = q;
        }
    for (size_t k = 0; k < (size_t)n * n; ++k) free(L[k].v);
    free(L);
    free(c);
    return 0;
}

int pmo_depth_ranges(const c3d_sparse_scene* s, double* dmin, double* dmax, uint8_t* has) {
    const uint32_t n = s->num_views;
    size_t* cnt = (size_t*)calloc(n, sizeof(size_t));
    double** l = (double**)calloc(n, sizeof(double*));
    size_t* cap = (size_t*)calloc(n, sizeof(size_t));
    for (uint64_t p = 0; p < s->num_points; ++p) {
        const double* x = s->xyz + 3 * p;
        for (uint64_t i = s->track_offset[p]; i < s->track_offset[p + 1]; ++i) {
            const uint32_t v = s->track_views[i];
            const double* r = s->rotation + 9 * (size_t)v + 6;
            const double d = r[0] * x[0] + r[1] * x[1] + r[2] * x[2] + s->translation[3 * (size_t)v + 2];
            if (d > 0.0) {
                if (cnt[v] == cap[v]) {
                    cap[v] = cap[v] ? 2 * cap[v] : 16;
                    l[v] = (double*)realloc(l[v], cap[v] * sizeof(double));
                }
                l[v][cnt[v]++] = d;
            }
        }
    }
    for (uint32_t v = 0; v < n; ++v) {
        const size_t m = cnt[v];
        if (!m) {
            dmin[v] = dmax[v] = 0.0;
            has[v] = 0;
            continue;
        }
        qsort(l[v], m, sizeof(double), cmp_d);
        size_t lo = (size_t)floor((double)m * 0.01), hi = (size_t)floor((double)m * 0.99);
        if (lo > m - 1) lo = m - 1;
        if (hi > m - 1) hi = m - 1;
        dmin[v] = 0.75 * l[v][lo];
        dmax[v] = 1.25 * l[v][hi];
        has[v] = 1;
        free(l[v]);
    }
    free(cnt);
    free(l);
    free(cap);
    return 0;
}
