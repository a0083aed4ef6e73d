/*
 * pm_oracle.c — CPU restatement of the reference PatchStereo backend. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * product path.
 *
 * Sources restated (all paths relative to the reference repository):
 *   crates/cuda/src/kernels/patchmatch.cu   kernels pm_half, pm_eval, pm_filter, pm_upsample,
 *                                           pm_median and their device helpers
 *   crates/cuda/src/patchmatch.rs           constants (496-521), level_pair (131-151), run
 *                                           schedule (752-799), evaluate (801-824), filter (842-868)
 *   crates/dense/src/kernel.rs              normative rules (32-74), rng_state/rng_next (248-259)
 *   crates/dense/src/math.rs                hash_bytes, mix64, erf, emission, priors (31-154)
 *   crates/dense/src/densify.rs             single-level schedule and threshold (350-454)
 *   crates/dense/src/upsample.rs            num_levels, build_pyramid (35-102)
 *   crates/dense/src/image.rs               bilinear_clamped (218-236)
 *
 * Arithmetic (FP32 mode) follows the device implementation exactly: explicit fmaf where the device
 * uses FMA, plain IEEE operations elsewhere (compiled with -ffp-contract=off), IEEE division and
 * square root, the same polynomial exp (pmo_expf) and sin/cos (pmo_sincosf), source samples kept
 * in gray levels (0..255) and shifted by the reference centre value before the NCC sums, a
 * two-pass weighted reference variance, and the covariance as sum(w s' I)/W - mu mean(s'). These are exact algebraic restatements of the reference's
 * formulas (NCC is invariant to the 1/255 scale and to the shift); they change only rounding.
 * The per-sample reciprocal of the projective divide is the device's seeded reciprocal (one
 * third-order and one Newton step; mode 1, the default); mode 0 (IEEE division) is kept to measure what that choice changes.
 *
 * Pinning (what fixes each part independently of this file):
 *   hash_bytes, mix64, rng_*     published FNV-1a-64 / splitmix64 vectors (tests/test_oracle_kats.py)
 *   erf, emission, priors,       reference unit tests math.rs:164-212 (known-answer values)
 *   plane_transfer, quantile
 *   bilinear_clamped             image.rs:316-326 (ramp value 26.25)
 *   num_levels, build_pyramid    upsample.rs:266-276 (levels 3/2/1, cx = 1.75, constant image)
 *   pmo_expf / pmo_sincosf       libm to 2 ulp over their domains
 *   rcp_seeded                   IEEE division to 1 ulp over every float of [1, 2) and 10^6
 *                                values over [2^-40, 2^40]
 *   level_pair + homography      brute-force projection of the plane's 3-D point (1e-4 px)
 *   ncc_cost (FP32)              FP64 cost oracle below to 1e-4; closed forms: identical
 *                                textures under the true plane give cost 0, constant patch 2,
 *                                projection outside 2
 *   ncc_cost (FP64)              closed forms above and a brute-force double evaluation
 *   pmo_region_pick,             known answers derived by hand from the rule text (kernel.rs:50-66,
 *   pmo_transfer_plane,          patchmatch.cu:327-345,425-436,714-733, densify.rs:446-454):
 *   pmo_select_views,            vote edges, previous-view weight, prior-weighted fallback, ties
 *   pmo_refine_accept,           to the earlier index, acceptance order, perturbation retries,
 *   pmo_perturb_normal_det,      top-K, read-out tests (tests/test_oracle_rules.py), and planted
 *   pmo_topk_mean,               slips in them rejected (tests/test_oracle_mutations.py)
 *   pmo_readout_source
 *   pm_half/pm_eval/pm_filter,   the rules above, plus synthetic ground truth: depth median rel. error < 0.005,
 *   schedule                     normal median < 3 deg, valid > 0.5 (patchmatch_gpu.rs:96-98
 *                                accuracy bars, applied to this oracle on the floor+boxes scene)
 *   pm_upsample, pm_median       upsample.rs:279-307 (plane reproduced to 1e-5)
 * Parity with an execution of the reference itself is unpinned: nothing in the reference can be
 * built or run here (Rust + CUDA/NVRTC), and it ships no golden depth maps.
 */
#include "pm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXSRC C3D_MAX_SOURCES

/* PMO_MUT selects a deliberately planted mistake for tests/test_oracle_mutations.py (0 = none).
 * Only test builds set it; the shipped oracle is compiled without it. */
#ifndef PMO_MUT
#define PMO_MUT 0
#endif
#define INF_ERR 1e30f
#define PI_F 3.14159265358979323846f

typedef struct { float x, y, z, w; } f4;
typedef struct { float x, y, z; } f3;

static __thread char g_err[512];
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char* pmo_last_error(void) { return g_err; }

static int g_rcp_mode = 1; /* 0: IEEE 1/z, 1: seeded reciprocal (the device's) */
void pmo_set_rcp_mode(int mode) { g_rcp_mode = mode; }

/* ================================================================ integer primitives */

/* FNV-1a 64 (math.rs:31-39) */
uint64_t pmo_hash_bytes(const uint8_t* b, size_t n) {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < n; ++i) {
        h ^= (uint64_t)b[i];
        h *= PMO_MUT == 10 ? 0x00000100000001b5ULL : 0x00000100000001b3ULL;
    }
    return h;
}

/* splitmix64 finaliser (math.rs:41-47, patchmatch.cu:112-116) */
uint64_t pmo_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* per-pixel start state (kernel.rs:248-252, patchmatch.cu:122-125) */
uint64_t pmo_rng_state(uint64_t seed, uint64_t key, uint32_t level, uint32_t run, uint32_t t,
                       uint32_t half, uint64_t pixel) {
    uint64_t tag = ((uint64_t)level << 48) ^ ((uint64_t)run << 32) ^ ((uint64_t)t << (PMO_MUT == 1 ? 15 : 16)) ^ (uint64_t)half;
    return pmo_mix64(seed ^ pmo_mix64(key ^ pmo_mix64(tag)) ^ pixel);
}

/* one step, uniform (0,1] (kernel.rs:254-259, patchmatch.cu:117-121) */
float pmo_rng_next(uint64_t* s) {
    *s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = pmo_mix64(*s);
    return ((float)(z >> 40) + 1.0f) * (1.0f / 16777216.0f);
}

/* ================================================================ math.rs twins (double) */

/* Taylor-series erf (math.rs:49-66) */
double pmo_erf(double x) {
    if (fabs(x) > 5.0) return x > 0 ? 1.0 : -1.0;
    double sum = 0.0, term = x, x2 = x * x;
    for (int n = 0; n < 200; ++n) {
        double t = term / (double)(2 * n + 1);
        sum += t;
        double as = fabs(sum) > 1e-300 ? fabs(sum) : 1e-300;
        if (fabs(t) < 1e-17 * as) break;
        term *= -x2 / (double)(n + 1);
    }
    return sum * 2.0 / sqrt(M_PI);
}

/* A(sigma) of the half-normal truncated to [0,2] (math.rs:68-71) */
double pmo_emission_norm(double sigma) {
    return 2.0 / (sqrt(2.0 * M_PI) * sigma * pmo_erf(2.0 / (sigma * M_SQRT2)));
}
/* E(c) (math.rs:73-76) */
double pmo_emission(double cost, double sigma) {
    return pmo_emission_norm(sigma) * exp(-cost * cost / (2.0 * sigma * (PMO_MUT == 2 ? 1.0 : sigma)));
}
/* E/(E+0.5) (math.rs:78-82) */
double pmo_visibility_probability(double cost, double sigma) {
    double e = pmo_emission(cost, sigma);
    return e / (e + 0.5);
}
/* math.rs:84-93 */
double pmo_triangulation_prior(double cos_theta, double min_angle_rad) {
    double cmin = cos(min_angle_rad);
    if (cos_theta > cmin) {
        double s = 1.0 - (1.0 - cos_theta) / (1.0 - cmin);
        double v = PMO_MUT == 9 ? 1.0 + s * s : 1.0 - s * s;
        return v < 0 ? 0 : (v > 1 ? 1 : v);
    }
    return 1.0;
}
/* math.rs:95-99 */
double pmo_incident_prior(double cos_phi, double sigma) {
    double x = 1.0 - (cos_phi > 0 ? cos_phi : 0.0);
    return exp(-x * x / (2.0 * sigma * sigma));
}
/* math.rs:101-123 */
double pmo_resolution_prior(const double h[9], double x, double y, double radius) {
    double cx[4] = {x - radius, x - radius, x + radius, x + radius};
    double cy[4] = {y - radius, y + radius, y + radius, y - radius};
    double px[4], py[4];
    for (int i = 0; i < 4; ++i) {
        double z = h[6] * cx[i] + h[7] * cy[i] + h[8];
        px[i] = (h[0] * cx[i] + h[1] * cy[i] + h[2]) / z;
        py[i] = (h[3] * cx[i] + h[4] * cy[i] + h[5]) / z;
    }
    double area = 0.0;
    for (int i = 0; i < 4; ++i) area += px[i] * py[(i + 1) % 4] - px[(i + 1) % 4] * py[i];
    double a_src = 0.5 * fabs(area);
    double a_ref = (2.0 * radius + 1.0) * (2.0 * radius + 1.0);
    if (a_src <= 0.0 || !isfinite(a_src)) return 0.0;
    return a_src / a_ref < a_ref / a_src ? a_src / a_ref : a_ref / a_src;
}
/* math.rs:143-154; returns 1 and writes *d_out on success */
int pmo_plane_transfer(double dq, const double n[3], const double rq[3], const double rp[3],
                       double* d_out) {
    double num = n[0] * rq[0] + n[1] * rq[1] + n[2] * rq[2];
    double den = n[0] * rp[0] + n[1] * rp[1] + n[2] * rp[2];
    if (den > -1e-6) return 0;
    double d = dq * num / den;
    if (!(d > 0.0 && isfinite(d))) return 0;
    *d_out = d;
    return 1;
}
/* math.rs:125-135 */
double pmo_quantile_sorted(const double* v, size_t n, double q) {
    if (n == 0) return 0.0;
    double pos = q * (double)(n - 1);
    size_t lo = (size_t)floor(pos);
    size_t hi = lo + 1 < n ? lo + 1 : n - 1;
    double f = pos - (double)lo;
    return v[lo] + (v[hi] - v[lo]) * f;
}
/* image.rs:218-236 */
float pmo_bilinear_clamped(const float* d, size_t w, size_t h, float x, float y) {
    float xm = (float)(w - 1), ym = (float)(h - 1);
    x = x < 0 ? 0 : (x > xm ? xm : x);
    y = y < 0 ? 0 : (y > ym ? ym : y);
    size_t x0 = (size_t)x, y0 = (size_t)y;
    size_t x1 = x0 + 1 < w ? x0 + 1 : w - 1, y1 = y0 + 1 < h ? y0 + 1 : h - 1;
    float fx = x - (float)x0, fy = y - (float)y0;
    float a = d[y0 * w + x0] + fx * (d[y0 * w + x1] - d[y0 * w + x0]);
    float b = d[y1 * w + x0] + fx * (d[y1 * w + x1] - d[y1 * w + x0]);
    return a + fy * (b - a);
}

/* ================================================================ elementary functions */

static inline float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* exp(x): Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, degree-7 Taylor polynomial for e^r,
 * scale by 2^n. Flushes to 0 below -86 (no subnormal results). Same steps as the device. */
float pmo_expf(float x) {
    if (x != x) return x;
    if (!(x > -86.0f)) return 0.0f;
    if (x > 88.0f) return INFINITY;
    float n = rintf(x * 1.44269504088896341f);
    float r = fmaf(-n, 0.693145751953125f, x);
    r = fmaf(-n, 1.42860677e-06f, r);
    float p = PMO_MUT == 7 ? 0.0f : 1.98412698e-04f; /* 1/5040 */
    p = fmaf(p, r, 1.38888889e-03f);     /* 1/720 */
    p = fmaf(p, r, 8.33333333e-03f);     /* 1/120 */
    p = fmaf(p, r, 4.16666667e-02f);     /* 1/24 */
    p = fmaf(p, r, 1.66666667e-01f);     /* 1/6 */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int ni = (int)n;
    return p * bits_f((uint32_t)(ni + 127) << 23);
}

/* sin and cos: k = rint(x 2/pi), three-part Cody-Waite reduction, Taylor polynomials on
 * |r| <= pi/4 (degree 9 / 10), quadrant swap. Same steps as the device. */
void pmo_sincosf(float x, float* so, float* co) {
    float k = rintf(x * 0.636619772367581343f);
    float r = fmaf(-k, 1.57079625129699707f, x);
    r = fmaf(-k, 7.54978941586159635e-08f, r);
    r = fmaf(-k, 5.39030252995776476e-15f, r);
    float r2 = r * r;
    float ps = 2.75573192e-06f;          /* 1/9! */
    ps = fmaf(r2, ps, -1.98412698e-04f); /* -1/7! */
    ps = fmaf(r2, ps, 8.33333333e-03f);  /* 1/5! */
    ps = fmaf(r2, ps, -1.66666667e-01f); /* -1/3! */
    float s = fmaf(r * r2, ps, r);
    float pc = -2.75573192e-07f;         /* -1/10! */
    pc = fmaf(r2, pc, 2.48015873e-05f);  /* 1/8! */
    pc = fmaf(r2, pc, -1.38888889e-03f); /* -1/6! */
    pc = fmaf(r2, pc, 4.16666667e-02f);  /* 1/4! */
    pc = fmaf(r2, pc, -0.5f);
    float c = fmaf(r2, pc, 1.0f);
    int q = ((int)k) & 3;
    if (q == 0) { *so = s; *co = c; }
    else if (q == 1) { *so = c; *co = -s; }
    else if (q == 2) { *so = -s; *co = -c; }
    else { *so = -c; *co = s; }
}

/* reciprocal of a positive normal float: bit-pattern seed, one third-order step r(1 + e + e^2)
 * and one Newton step r(1 + e) with e = 1 - z r (mode 1), or IEEE division (mode 0) */
static inline float rcp_seeded(float z) {
    float r = bits_f(0x7EF311C3u - f_bits(z));
    float e = fmaf(-z, r, 1.0f);
    r = fmaf(r, fmaf(e, e, e), r);
    e = fmaf(-z, r, 1.0f);
    r = fmaf(r, e, r);
    return r;
}

void pmo_rcp_sample(const float* z, float* r, size_t n) {
    for (size_t i = 0; i < n; ++i) r[i] = rcp_seeded(z[i]);
}

static inline float rcp_sample(float z) {
    if (g_rcp_mode == 0) return 1.0f / z;
    return rcp_seeded(z);
}

/* ================================================================ geometry */

/* A = K_j R K^-1, b = K_j t in double, stored as float (patchmatch.rs:131-151) */
static void level_pair_d(const float kr[4], const float kj[4], const float r[9], const float t[3],
                         double a[9], double b[3]) {
    double fx = kr[0], fy = kr[1], cx = kr[2], cy = kr[3];
    double kinv[3][3] = {{1.0 / fx, 0.0, -cx / fx}, {0.0, 1.0 / fy, -cy / fy}, {0.0, 0.0, 1.0}};
    double kjm[3][3] = {{kj[0], 0.0, kj[2]}, {0.0, kj[1], kj[3]}, {0.0, 0.0, 1.0}};
    double kr_[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s += kjm[i][m] * (double)(PMO_MUT == 3 ? r[k * 3 + m] : r[m * 3 + k]);
            kr_[i][k] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s += kr_[i][m] * kinv[m][k];
            a[i * 3 + k] = s;
        }
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int m = 0; m < 3; ++m) s += kjm[i][m] * (double)t[m];
        b[i] = s;
    }
}
void pmo_level_pair(const float kr[4], const float kj[4], const float r[9], const float t[3],
                    float a_out[9], float b_out[3]) {
    double a[9], b[3];
    level_pair_d(kr, kj, r, t, a, b);
    for (int i = 0; i < 9; ++i) a_out[i] = (float)a[i];
    for (int i = 0; i < 3; ++i) b_out[i] = (float)b[i];
}

/* ================================================================ pyramid (upsample.rs:35-102) */

uint32_t pmo_num_levels(uint32_t min_side, uint32_t max_levels, uint32_t min_size) {
    uint32_t l = 1, s = min_side;
    while (l < max_levels && s / 2 >= min_size) {
        s /= 2;
        l += 1;
    }
    return l;
}

static void blur3(const float* src, uint32_t w, uint32_t h, float* out, float* tmp) {
    float g0 = expf(-1.0f / (2.0f * 0.25f)), g1 = 1.0f, g2 = g0;
    float s = g0 + g1 + g2;
    float g[3] = {g0 / s, g1 / s, g2 / s};
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            float a = 0.0f;
            for (int k = 0; k < 3; ++k) {
                long xx = (long)x + k - 1;
                if (xx < 0) xx = 0;
                if (xx > (long)w - 1) xx = (long)w - 1;
                a += g[k] * src[(size_t)y * w + (size_t)xx];
            }
            tmp[(size_t)y * w + x] = a;
        }
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            float a = 0.0f;
            for (int k = 0; k < 3; ++k) {
                long yy = (long)y + k - 1;
                if (yy < 0) yy = 0;
                if (yy > (long)h - 1) yy = (long)h - 1;
                a += g[k] * tmp[(size_t)yy * w + x];
            }
            out[(size_t)y * w + x] = a;
        }
}

void pmo_build_pyramid(const uint8_t* gray, uint32_t w, uint32_t h, const double k[4],
                       uint32_t levels, uint8_t* const* out, float* k_out) {
    double kk[4] = {k[0], k[1], k[2], k[3]};
    for (int i = 0; i < 4; ++i) k_out[i] = (float)kk[i];
    float* cur = (float*)malloc(sizeof(float) * (size_t)w * h);
    float* b = (float*)malloc(sizeof(float) * (size_t)w * h);
    float* tmp = (float*)malloc(sizeof(float) * (size_t)w * h);
    for (size_t i = 0; i < (size_t)w * h; ++i) cur[i] = (float)gray[i];
    for (uint32_t l = 1; l < levels; ++l) {
        blur3(cur, w, h, b, tmp);
        uint32_t nw = w / 2, nh = h / 2;
        uint8_t* q = out[l - 1];
        for (uint32_t y = 0; y < nh; ++y)
            for (uint32_t x = 0; x < nw; ++x) {
                size_t i = 2 * (size_t)y * w + 2 * (size_t)x;
                float v = 0.25f * (b[i] + b[i + 1] + b[i + w] + b[i + w + 1]);
                float r = roundf(v);
                q[(size_t)y * nw + x] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
            }
        for (size_t i = 0; i < (size_t)nw * nh; ++i) cur[i] = (float)q[i];
        kk[0] *= 0.5;
        kk[1] *= 0.5;
        kk[2] = (kk[2] + 0.5) * 0.5 - 0.5;
        kk[3] = (kk[3] + 0.5) * 0.5 - 0.5;
        w = nw;
        h = nh;
        for (int i = 0; i < 4; ++i) k_out[4 * l + i] = (float)kk[i];
    }
    free(cur);
    free(b);
    free(tmp);
}

/* ================================================================ context */

typedef struct {
    int w, h;
    const uint8_t* img;
    float fx, fy, cx, cy;
} lvl_t;

typedef struct {
    int src;
    float r[9], t[3], c[3];
} pair_t;

typedef struct {
    float A[9], b[3];
} lpair_t;

struct pmo_ctx {
    const c3d_kernel_input* in;
    int nthreads;
    int nv, nl;
    lvl_t* lv;       /* nv * nl */
    pair_t* pairs;   /* nv * MAXSRC */
    lpair_t* lpairs; /* nv * nl * MAXSRC */
    int wr, wstep, wn;
    float inv2s, inv2c255, tau_bad, inv2beta2, prev_w;
    int n1, n2, topk;
    float cos_tri_min, inv2inc, lambda, gmax;
    float cos_filter_tri, q_min, ncc_norm, inv2ncc, filter_gmax, weak_var255, vmin255;
    double inv2s_d, inv2c_d;
    /* snapshot */
    const float** snap;
    int snap_level;
};

typedef struct {
    int level, geometric, run_id, t, color, random_init, use_prior;
    float tau_good, eps, phi;
} launch_t;

typedef struct {
    int w, h;
    f4* plane;
    float* cost;
    uint8_t* bview;
    const f4* prior;
} vstate_t;

static inline float deg2rad(float d) { return d * (PI_F / 180.0f); }

pmo_ctx* pmo_create(const c3d_kernel_input* in, int num_threads) {
    const c3d_pm_params* p = &in->pm;
    if (p->window_radius < 1 || p->window_radius > 20 || p->window_step < 1 || p->window_step > 2) {
        set_err("window radius 1..=20 and step 1..=2 required");
        return NULL;
    }
    pmo_ctx* c = (pmo_ctx*)calloc(1, sizeof(pmo_ctx));
    c->in = in;
    c->nthreads = num_threads > 0 ? num_threads : 1;
    c->nv = (int)in->num_views;
    c->nl = (int)in->num_levels;
    c->lv = (lvl_t*)calloc((size_t)c->nv * c->nl, sizeof(lvl_t));
    c->pairs = (pair_t*)calloc((size_t)c->nv * MAXSRC, sizeof(pair_t));
    c->lpairs = (lpair_t*)calloc((size_t)c->nv * c->nl * MAXSRC, sizeof(lpair_t));
    c->snap = (const float**)calloc((size_t)(c->nv > 0 ? c->nv : 1), sizeof(float*));
    for (int v = 0; v < c->nv; ++v) {
        const c3d_kernel_view* kv = &in->views[v];
        if (kv->num_sources > MAXSRC) {
            set_err("more than 32 sources");
            pmo_destroy(c);
            return NULL;
        }
        for (int l = 0; l < c->nl; ++l) {
            const c3d_level_image* li = &kv->levels[l];
            lvl_t* L = &c->lv[v * c->nl + l];
            L->w = (int)li->width;
            L->h = (int)li->height;
            L->img = li->gray;
            L->fx = li->k[0];
            L->fy = li->k[1];
            L->cx = li->k[2];
            L->cy = li->k[3];
        }
        for (uint32_t j = 0; j < kv->num_sources; ++j) {
            pair_t* pr = &c->pairs[v * MAXSRC + j];
            pr->src = (int)kv->sources[j];
            memcpy(pr->r, kv->pairs[j].r, sizeof pr->r);
            memcpy(pr->t, kv->pairs[j].t, sizeof pr->t);
            memcpy(pr->c, kv->pairs[j].center, sizeof pr->c);
            if (pr->src < 0 || pr->src >= c->nv) {
                set_err("source index out of range");
                pmo_destroy(c);
                return NULL;
            }
            for (int l = 0; l < c->nl; ++l) {
                lpair_t* lp = &c->lpairs[((size_t)v * c->nl + l) * MAXSRC + j];
                pmo_level_pair(kv->levels[l].k, in->views[pr->src].levels[l].k, pr->r, pr->t, lp->A, lp->b);
            }
        }
    }
    /* constants (patchmatch.rs:496-521); f32 arithmetic as in the reference */
    float sc = p->sigma_spatial > 0.0f ? p->sigma_spatial : (float)p->window_radius;
    c->wr = p->window_radius;
    c->wstep = p->window_step;
    c->wn = (2 * c->wr) / c->wstep + 1;
    c->inv2s = 1.0f / (2.0f * sc * sc);
    float inv2c = 1.0f / (2.0f * p->sigma_color * p->sigma_color);
    c->inv2c255 = inv2c / 65025.0f;
    c->inv2s_d = 1.0 / (2.0 * (double)sc * (double)sc);
    c->inv2c_d = 1.0 / (2.0 * (double)p->sigma_color * (double)p->sigma_color);
    c->tau_bad = p->tau1;
    c->inv2beta2 = 1.0f / (2.0f * p->beta * p->beta);
    c->prev_w = p->prev_view_weight;
    c->n1 = (int)p->n1;
    c->n2 = (int)p->n2;
    c->topk = (int)p->top_k;
    c->cos_tri_min = cosf(deg2rad(p->min_triangulation_angle_deg));
    c->inv2inc = 1.0f / (2.0f * p->incident_angle_sigma * p->incident_angle_sigma);
    c->lambda = p->geom_lambda;
    c->gmax = p->geom_max_cost;
    const c3d_filter_params* f = &in->filter;
    c->cos_filter_tri = cosf(deg2rad(f->min_triangulation_angle_deg));
    c->q_min = (float)pmo_emission(1.0 - (double)f->min_ncc, (double)f->ncc_sigma);
    c->ncc_norm = (float)pmo_emission_norm((double)f->ncc_sigma);
    c->inv2ncc = 1.0f / (2.0f * f->ncc_sigma * f->ncc_sigma);
    c->filter_gmax = f->geom_max_cost;
    c->weak_var255 = p->weak_texture_var * 65025.0f;
    c->vmin255 = 1e-5f * 65025.0f;
    c->snap_level = -1;
    return c;
}

void pmo_destroy(pmo_ctx* c) {
    if (!c) return;
    free(c->lv);
    free(c->pairs);
    free(c->lpairs);
    free(c->snap);
    free(c);
}

/* ================================================================ device helpers, restated */

static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 mk3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 nrm3(f3 a) {
    float d = dot3(a, a);
    float l = 1.0f / sqrtf(d > 1e-30f ? d : 1e-30f);
    return mk3(a.x * l, a.y * l, a.z * l);
}
static inline f3 xyz(f4 h) { return mk3(h.x, h.y, h.z); }
static inline f4 mk4(f3 n, float d) { f4 r = {n.x, n.y, n.z, d}; return r; }

/* patchmatch.cu:386 */
static inline f3 pixel_ray(const lvl_t* R, int x, int y) {
    return mk3(((float)x - R->cx) / R->fx, ((float)y - R->cy) / R->fy, 1.0f);
}

static inline float ref_at(const lvl_t* R, int x, int y) {
    if (x < 0 || y < 0 || x >= R->w || y >= R->h) return 0.0f;
    return (float)R->img[(size_t)y * R->w + x];
}

#define MAXWN2 (41 * 41)

/* reference patch statistics (patchmatch.cu:199-223), gray-level units, two-pass variance; the
 * window values ik are kept for the covariance sum */
typedef struct {
    float ic, W, mu, var;
    float w[MAXWN2];
    float ik[MAXWN2];
} rstats_t;

static void ref_stats(const pmo_ctx* c, const lvl_t* R, int x, int y, rstats_t* s) {
    const int WN = c->wn;
    float* ik = s->ik;
    s->ic = ref_at(R, x, y);
    float W = 0.0f, S = 0.0f;
    for (int a = 0; a < WN; ++a) {
        int dy = -c->wr + a * c->wstep;
        for (int b = 0; b < WN; ++b) {
            int dx = -c->wr + b * c->wstep;
            int k = a * WN + b;
            float v = ref_at(R, x + dx, y + dy);
            float d = s->ic - v;
            float sp = (float)(dx * dx + dy * dy) * c->inv2s;
            float w = pmo_expf(-fmaf(d * d, c->inv2c255, sp));
            ik[k] = v;
            s->w[k] = w;
            W += w;
            S = fmaf(w, v, S);
        }
    }
    float mu = S / W;
    float V = 0.0f;
    for (int k = 0; k < WN * WN; ++k) {
        float dk = ik[k] - mu;
        float wd = s->w[k] * dk;
        V = fmaf(wd, dk, V);
    }
    s->W = W;
    s->mu = mu;
    s->var = V / W;
}

/* plane-induced homography (patchmatch.cu:225-242), translated to the pixel: H maps the window
 * offset (dx, dy, 1) of pixel (x, y), i.e. H = (A + b n^T K^-1 / delta) T(x, y). Its third column
 * is A (x, y, 1)^T + b / d exactly (n^T K^-1 (x, y, 1)^T / delta = 1 / d), which avoids the
 * cancellation of the untranslated form for strongly slanted planes. */
static int homography(const lvl_t* R, const lpair_t* L, float x, float y, float d, f3 n, float* H) {
    float ifx = 1.0f / R->fx, ify = 1.0f / R->fy;
    f3 ray = mk3((x - R->cx) * ifx, (y - R->cy) * ify, 1.0f);
    float delta = d * dot3(n, ray);
    if (!(delta < -1e-12f)) return 0;
    float id = 1.0f / delta, invd = 1.0f / d;
    float u0 = n.x * ifx, u1 = n.y * ify;
    if (PMO_MUT == 8) {
        float t_ = u0;
        u0 = u1;
        u1 = t_;
    }
    for (int r = 0; r < 3; ++r) {
        float br = L->b[r] * id;
        H[r * 3 + 0] = L->A[r * 3 + 0] + br * u0;
        H[r * 3 + 1] = L->A[r * 3 + 1] + br * u1;
        H[r * 3 + 2] = fmaf(L->A[r * 3 + 0], x, fmaf(L->A[r * 3 + 1], y, L->A[r * 3 + 2])) + L->b[r] * invd;
    }
    return 1;
}

/* software bilinear source sample in gray levels, 0 outside (patchmatch.cu:169-192, HW_INTERP 0) */
static inline float src_sample(const lvl_t* S, float u, float v) {
    float xf = floorf(u), yf = floorf(v);
    float ax = u - xf, ay = v - yf;
    float p00 = 0.f, p10 = 0.f, p01 = 0.f, p11 = 0.f;
    if (xf >= -1.0f && xf <= (float)(S->w - 1) && yf >= -1.0f && yf <= (float)(S->h - 1)) {
        int x0 = (int)xf, y0 = (int)yf;
        int xin0 = x0 >= 0, xin1 = x0 + 1 < S->w;
        if (y0 >= 0) {
            const uint8_t* r = S->img + (size_t)y0 * S->w;
            if (xin0) p00 = r[x0];
            if (xin1) p10 = r[x0 + 1];
        }
        if (y0 + 1 < S->h) {
            const uint8_t* r = S->img + (size_t)(y0 + 1) * S->w;
            if (xin0) p01 = r[x0];
            if (xin1) p11 = r[x0 + 1];
        }
    }
    float top = fmaf(ax, p10 - p00, p00);
    float bot = fmaf(ax, p11 - p01, p01);
    return fmaf(ay, bot - top, top);
}

/* bilateral-weighted NCC cost 1 - rho, 2 on failure (patchmatch.cu:244-290); H is translated to
 * the pixel, samples are at window offsets (dx, dy) */
static float ncc_cost(const pmo_ctx* c, const rstats_t* rs, const lvl_t* S, const float* H, int x, int y) {
    const int WN = c->wn;
    (void)x;
    (void)y;
    float z = H[8];
    if (!(z > 1e-8f)) return 2.0f;
    float iz = 1.0f / z;
    float u = H[2] * iz;
    float v = H[5] * iz;
    if (!(u >= 0.f && v >= 0.f && u <= (float)(S->w - 1) && v <= (float)(S->h - 1))) return 2.0f;
    if (rs->var < c->vmin255) return 2.0f;
    float Ss = 0.f, SSs = 0.f, IS = 0.f;
    for (int a = 0; a < WN; ++a) {
        float oy = (float)(-c->wr + a * c->wstep);
        float rz = fmaf(H[7], oy, H[8]);
        float ru = fmaf(H[1], oy, H[2]);
        float rv = fmaf(H[4], oy, H[5]);
        for (int b = 0; b < WN; ++b) {
            float ox = (float)(-c->wr + b * c->wstep);
            int k = a * WN + b;
            float zz = fmaf(H[6], ox, rz);
            float sv = 0.0f;
            if (zz > 1e-8f) {
                float r = rcp_sample(zz);
                sv = src_sample(S, fmaf(H[0], ox, ru) * r, fmaf(H[3], ox, rv) * r);
            }
            float sp = sv - rs->ic;
            float ws = rs->w[k] * sp;
            Ss += ws;
            SSs = fmaf(ws, sp, SSs);
            IS = fmaf(ws, rs->ik[k], IS);
        }
    }
    float iW = 1.0f / rs->W;
    float mus = Ss * iW;
    float vars = PMO_MUT == 5 ? SSs * iW : fmaf(-mus, mus, SSs * iW);
    if (vars < c->vmin255) return 2.0f;
    float cov = fmaf(-rs->mu, mus, IS * iW);
    float cc = 1.0f - cov / sqrtf(rs->var * vars);
    return fminf(fmaxf(cc, 0.0f), 2.0f);
}

/* forward-backward reprojection error (patchmatch.cu:292-314) */
static float geo_error(const pmo_ctx* c, const lvl_t* R, const lvl_t* S, const pair_t* pr, int x, int y, float d) {
    const float* ds = c->snap[pr->src];
    if (!ds) return INF_ERR;
    f3 X = mk3(((float)x - R->cx) / R->fx * d, ((float)y - R->cy) / R->fy * d, d);
    float xj = pr->r[0] * X.x + pr->r[1] * X.y + pr->r[2] * X.z + pr->t[0];
    float yj = pr->r[3] * X.x + pr->r[4] * X.y + pr->r[5] * X.z + pr->t[1];
    float zj = pr->r[6] * X.x + pr->r[7] * X.y + pr->r[8] * X.z + pr->t[2];
    if (!(zj > 1e-8f)) return INF_ERR;
    float u = S->fx * xj / zj + S->cx, v = S->fy * yj / zj + S->cy;
    float uf = floorf(u + 0.5f), vf = floorf(v + 0.5f);
    if (!(uf >= 0.f && vf >= 0.f && uf < (float)S->w && vf < (float)S->h)) return INF_ERR;
    float dj = ds[(size_t)vf * S->w + (size_t)uf];
    if (!(dj > 0.f)) return INF_ERR;
    f3 Yj = mk3((u - S->cx) / S->fx * dj, (v - S->cy) / S->fy * dj, dj);
    f3 q = mk3(Yj.x - pr->t[0], Yj.y - pr->t[1], Yj.z - pr->t[2]);
    float Yx = pr->r[0] * q.x + pr->r[3] * q.y + pr->r[6] * q.z;
    float Yy = pr->r[1] * q.x + pr->r[4] * q.y + pr->r[7] * q.z;
    float Yz = pr->r[2] * q.x + pr->r[5] * q.y + pr->r[8] * q.z;
    if (!(Yz > 1e-8f)) return INF_ERR;
    float ex = R->fx * Yx / Yz + R->cx - (float)x, ey = R->fy * Yy / Yz + R->cy - (float)y;
    return sqrtf(ex * ex + ey * ey);
}

/* (photometric cost, truncated geometric error) of one hypothesis x one source
 * (patchmatch.cu:316-324) */
static inline float pair_cost(const pmo_ctx* c, const launch_t* P, const rstats_t* rs, const lvl_t* R,
                              const lvl_t* S, const pair_t* pr, const lpair_t* lp, int x, int y, f4 h,
                              float* geo) {
    float H[9];
    float cc = 2.0f;
    if (homography(R, lp, (float)x, (float)y, h.w, xyz(h), H)) cc = ncc_cost(c, rs, S, H, x, y);
    if (P->geometric) *geo = fminf(geo_error(c, R, S, pr, x, y, h.w), c->gmax);
    return cc;
}

/* top-K mean (patchmatch.cu:326-345): the per-source costs in source order, then pmo_topk_mean */
static float topk_cost(const pmo_ctx* c, const launch_t* P, const rstats_t* rs, const lvl_t* R, int view,
                       int nsrc, int x, int y, f4 h) {
    float cs[MAXSRC];
    for (int j = 0; j < nsrc; ++j) {
        const pair_t* pr = &c->pairs[view * MAXSRC + j];
        const lvl_t* S = &c->lv[pr->src * c->nl + P->level];
        const lpair_t* lp = &c->lpairs[((size_t)view * c->nl + P->level) * MAXSRC + j];
        float g = 0.f;
        float cc = pair_cost(c, P, rs, R, S, pr, lp, x, y, h, &g);
        if (P->geometric) cc = cc + c->lambda * g;
        cs[j] = cc;
    }
    return pmo_topk_mean(cs, nsrc, c->topk);
}

/* uniform normal on the sphere facing the camera (patchmatch.cu:347-361) */
static f3 random_normal(uint64_t* st, f3 ray) {
    f3 n = mk3(0.f, 0.f, -1.f);
    for (int k = 0; k < 64; ++k) {
        float v1 = 2.f * pmo_rng_next(st) - 1.f;
        float v2 = 2.f * pmo_rng_next(st) - 1.f;
        float s = v1 * v1 + v2 * v2;
        if (s < 1.f && s > 0.f) {
            float r = 2.f * sqrtf(1.f - s);
            n = mk3(v1 * r, v2 * r, 1.f - 2.f * s);
            break;
        }
    }
    if (dot3(n, ray) > 0.f) n = mk3(-n.x, -n.y, -n.z);
    return n;
}

/* uniform in inverse depth (patchmatch.cu:363-366) */
static float random_depth(uint64_t* st, float dmin, float dmax) {
    float a = 1.f / dmax, b = 1.f / dmin;
    return 1.f / (a + pmo_rng_next(st) * (b - a));
}

/* rotate by u*phi about a random perpendicular axis, up to 4 halvings (patchmatch.cu:368-384); the
 * draws are taken here in the device's order, the rotation is pmo_perturb_normal_det */
static f3 perturb_normal(uint64_t* st, f3 n, f3 ray, float phi) {
    f3 a = random_normal(st, mk3(0.f, 0.f, 0.f));
    float u = pmo_rng_next(st);
    float av[3] = {a.x, a.y, a.z}, nv[3] = {n.x, n.y, n.z}, rv[3] = {ray.x, ray.y, ray.z}, o[3];
    pmo_perturb_normal_det(av, u, nv, rv, phi, o);
    return mk3(o[0], o[1], o[2]);
}

/* triangulation x incidence x resolution prior (patchmatch.cu:388-422) */
static float view_prior(const pmo_ctx* c, const lvl_t* R, const pair_t* pr, const float* H, int hok, int x,
                        int y, f4 h) {
    if (!hok) return 0.f;
    f3 ray = pixel_ray(R, x, y);
    f3 X = mk3(ray.x * h.w, ray.y * h.w, h.w);
    f3 S = mk3(pr->c[0] - X.x, pr->c[1] - X.y, pr->c[2] - X.z);
    float ls = sqrtf(dot3(S, S)), lx = sqrtf(dot3(X, X));
    if (ls <= 0.f || lx <= 0.f) return 0.f;
    float cost = -dot3(S, X) / (ls * lx);
    float tri = 1.f;
    if (cost > c->cos_tri_min) {
        float s = 1.f - (1.f - cost) / (1.f - c->cos_tri_min);
        tri = fminf(fmaxf(1.f - s * s, 0.f), 1.f);
    }
    float cphi = dot3(S, xyz(h)) / ls;
    float xi = 1.f - fmaxf(0.f, cphi);
    float inc = pmo_expf(-xi * xi * c->inv2inc);
    float r = (float)c->wr;
    float cxs[4] = {-r, -r, r, r};
    float cys[4] = {-r, r, r, -r};
    float px[4], py[4];
    for (int k = 0; k < 4; ++k) {
        float z = H[6] * cxs[k] + H[7] * cys[k] + H[8];
        if (!(z > 1e-8f)) return 0.f;
        px[k] = (H[0] * cxs[k] + H[1] * cys[k] + H[2]) / z;
        py[k] = (H[3] * cxs[k] + H[4] * cys[k] + H[5]) / z;
    }
    float area = 0.f;
    for (int k = 0; k < 4; ++k) area += px[k] * py[(k + 1) & 3] - px[(k + 1) & 3] * py[k];
    area = 0.5f * fabsf(area);
    float aref = (2.f * r + 1.f) * (2.f * r + 1.f);
    if (!(area > 0.f)) return 0.f;
    float res = fminf(area / aref, aref / area);
    return tri * inc * res;
}

/* ================================================================ decision rules over tables
 *
 * The selection, acceptance and read-out rules of the half step, run start and read-out as pure
 * functions of cost tables, so that tests/test_oracle_rules.py can pin each against cases derived
 * by hand from the normative text (kernel.rs:50-66) and the kernel's read-out and top-K code. The
 * per-pixel kernels below compute the tables and call these; the float operations and their order
 * are the device's. */

/* Region pick of the propagation step (kernel.rs:50-52, patchmatch.cu:480-498): among the
 * region's offsets inside the image, the neighbour with the smallest stored cost, the first one in
 * offset order on ties. Regions 0-3 are the V shapes up/right/down/left, 4-7 the bands. Returns 1
 * with the neighbour in (*qx, *qy), 0 when no offset lies inside the image. */
int pmo_region_pick(const float* cost, int w, int h, int x, int y, int region, int* qx_out, int* qy_out) {
        /* region offsets for the up direction (patchmatch.cu:424-436) */
    static const int V_OFF[7][2] = {{0, -1}, {-1, -2}, {1, -2}, {-2, -3}, {2, -3}, {-3, -4}, {3, -4}};
    const int dir = region & 3;
    const int strip = region >= 4;
    const int cnt = strip ? 11 : 7;
    float best = 1e30f;
    int found = 0;
    for (int k = 0; k < cnt; ++k) {
        int dx = strip ? 0 : V_OFF[k][0], dy = strip ? -(3 + 2 * k) : V_OFF[k][1];
        for (int r = 0; r < dir; ++r) { /* up -> right -> down -> left: (dx, dy) -> (-dy, dx) */
            int t = dx;
            dx = -dy;
            dy = t;
        }
        const int qx = x + dx, qy = y + dy;
        if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
        const float cq = cost[qy * w + qx];
#if PMO_MUT == 19
        if (cq <= best) {
#else
        if (cq < best) {
#endif
            best = cq;
            *qx_out = qx;
            *qy_out = qy;
            found = 1;
        }
    }
    return found;
}

/* Moves the neighbour's plane q = (n, d_q) seen along ray_q onto the pixel's ray (kernel.rs:51-52,
 * patchmatch.cu:499-510): d = d_q (n . ray_q) / (n . ray), kept when n . ray < -1e-6, d_q > 0 and
 * d is positive and finite. Returns 1 with the moved plane in out. */
int pmo_transfer_plane(const float q[4], const float ray_q[3], const float ray[3], float out[4]) {
    const float den = q[0] * ray[0] + q[1] * ray[1] + q[2] * ray[2];
    if (!(den < -1e-6f && q[3] > 0.f)) return 0;
    const float d = q[3] * (q[0] * ray_q[0] + q[1] * ray_q[1] + q[2] * ray_q[2]) / den;
    if (!(d > 0.f && isfinite(d))) return 0;
    out[0] = q[0];
    out[1] = q[1];
    out[2] = q[2];
    out[3] = d;
    return 1;
}

/* View selection and aggregation (kernel.rs:53-58, patchmatch.cu:513-594) over the cost table of
 * one pixel: prior[j] of each source, m[9 j + i] the photometric cost and g[9 j + i] the truncated
 * geometric error of hypothesis i (0 = current, 1-8 = propagated) against source j (g NULL for a
 * photometric run). Sources with prior <= 0 take no part (their m and g are not read). Writes the
 * refinement weights wv[j] (> 0 for the sources that take part), their sum *den_out, the aggregated
 * costs agg[9]; *best_out is the adopted hypothesis (the earlier index on ties) and *view_out the
 * most important view (the earlier source on ties, 255 when no view is adopted). */
void pmo_select_views(const pmo_select_params* p, int nsrc, int vprev, const float* prior, const float* m, const float* g,
                      float* wv, float* den_out, float* agg, int* best_out, int* view_out) {
    float num[9], numf[9];
    for (int i = 0; i < 9; ++i) num[i] = numf[i] = 0.f;
    float den = 0.f, wbest = 0.f, denf = 0.f;
    int vbest = 255;
    for (int j = 0; j < nsrc; ++j) {
        wv[j] = 0.f;
        if (!(prior[j] > 0.f)) continue;
        const float* mj = m + 9 * j;
        const float* gj = g ? g + 9 * j : NULL;
        int good = 0, bad = 0;
        float conf = 0.f;
        for (int i = 1; i < 9; ++i) {
            if (mj[i] < p->tau_good) {
                good++;
                conf += pmo_expf(-mj[i] * mj[i] * p->inv2beta2);
            }
            if (mj[i] > p->tau_bad) bad++;
        }
#if PMO_MUT == 11
        int adopt = good >= p->n1 && bad < p->n2;
#elif PMO_MUT == 12
        int adopt = good > p->n1 && bad <= p->n2;
#else
        int adopt = good > p->n1 && bad < p->n2;
#endif
        int isprev = j == vprev;
#if PMO_MUT == 13
        float wp = adopt ? 1.f * conf / (float)good : (isprev ? p->prev_w : 0.f);
#elif PMO_MUT == 14
        float wp = adopt ? (isprev ? 2.f : 1.f) * conf / (float)good : 0.f;
#else
        float wp = adopt ? (isprev ? 2.f : 1.f) * conf / (float)good : (isprev ? p->prev_w : 0.f);
#endif
        float w2 = wp * prior[j];
        wv[j] = w2 > 0.f ? w2 : -prior[j];
        for (int i = 0; i < 9; ++i) numf[i] += prior[j] * (mj[i] + p->lambda * (gj ? gj[i] : 0.f));
        denf += prior[j];
        if (w2 > 0.f) {
            for (int i = 0; i < 9; ++i) num[i] += w2 * (mj[i] + p->lambda * (gj ? gj[i] : 0.f));
            den += w2;
#if PMO_MUT == 16
            if (w2 >= wbest) {
#else
            if (w2 > wbest) {
#endif
                wbest = w2;
                vbest = j;
            }
        }
    }
    int bi9 = 0;
    if (den > 0.f) {
        for (int i = 0; i < 9; ++i) agg[i] = num[i] / den;
        for (int i = 1; i < 9; ++i)
#if PMO_MUT == 15
            if (agg[i] <= agg[bi9]) bi9 = i;
#else
            if (agg[i] < agg[bi9]) bi9 = i;
#endif
    } else if (denf > 0.f) {
        /* no adopted view: prior-weighted average, and refinement is compared with the priors */
        for (int i = 0; i < 9; ++i) agg[i] = numf[i] / denf;
        for (int i = 1; i < 9; ++i)
            if (agg[i] < agg[bi9]) bi9 = i;
#if PMO_MUT != 17
        for (int j = 0; j < nsrc; ++j) wv[j] = -wv[j];
#endif
        den = denf;
    } else {
        for (int i = 0; i < 9; ++i) agg[i] = 2.f;
    }
    *den_out = den;
    *best_out = bi9;
    *view_out = vbest;
}

/* Refinement acceptance (kernel.rs:59-62, patchmatch.cu:608-653): candidate c's cost is the
 * wv-weighted mean over the sources with wv[j] > 0 of m[c nsrc + j] + lambda g[c nsrc + j] (g NULL
 * for a photometric run), divided by den, or 2 when den <= 0; candidates are tried in order and one
 * is adopted when its cost is strictly below the running best *bcost (updated). Returns the index
 * of the last adopted candidate, or -1. */
int pmo_refine_accept(const pmo_select_params* p, int ncand, int nsrc, const float* wv, float den, const float* m, const float* g,
                      float* bcost) {
    int taken = -1;
    for (int ci = 0; ci < ncand; ++ci) {
        float cc;
        if (den > 0.f) {
            float nm = 0.f;
            for (int j = 0; j < nsrc; ++j) {
                if (!(wv[j] > 0.f)) continue;
                nm += wv[j] * (m[ci * nsrc + j] + p->lambda * (g ? g[ci * nsrc + j] : 0.f));
            }
            cc = nm / den;
        } else {
            cc = 2.f;
        }
#if PMO_MUT == 20
        if (cc <= *bcost) {
#else
        if (cc < *bcost) {
#endif
            *bcost = cc;
            taken = ci;
        }
    }
    return taken;
}

/* Normal perturbation given its random draws (kernel.rs:60-61, patchmatch.cu:368-384): `a` is the
 * random unit vector that fixes the rotation axis a x n and u in (0, 1] the angle fraction. Rotates n
 * by u phi about that axis toward b = axis x n; when the result does not face the camera (m . ray >= 0)
 * the angle is halved, up to 4 tries in all, after which n is kept. n is also kept when a is
 * parallel to n (|a x n|^2 < 1e-12). */
void pmo_perturb_normal_det(const float a[3], float u, const float n[3], const float ray[3], float phi, float out[3]) {
    f3 nn = mk3(n[0], n[1], n[2]);
    f3 ax = cross3(mk3(a[0], a[1], a[2]), nn);
    float l2 = dot3(ax, ax);
    float ang = u * phi;
    out[0] = n[0];
    out[1] = n[1];
    out[2] = n[2];
    if (l2 < 1e-12f) return;
    ax = nrm3(ax);
    f3 bn = cross3(ax, nn);
#if PMO_MUT == 18
    for (int k = 0; k < 1; ++k) {
#else
    for (int k = 0; k < 4; ++k) {
#endif
        float sa, ca;
        pmo_sincosf(ang, &sa, &ca);
        f3 mm = mk3(nn.x * ca + bn.x * sa, nn.y * ca + bn.y * sa, nn.z * ca + bn.z * sa);
        if (dot3(mm, mk3(ray[0], ray[1], ray[2])) < 0.f) {
            mm = nrm3(mm);
            out[0] = mm.x;
            out[1] = mm.y;
            out[2] = mm.z;
            return;
        }
        ang *= 0.5f;
    }
}

/* Top-K mean (kernel.rs:64-65 run start, patchmatch.cu:326-345): the mean of the min(k, n) smallest
 * of the n costs, 2 when there is none. */
float pmo_topk_mean(const float* cost, int n, int k) {
    float b0 = 1e30f, b1 = 1e30f, b2 = 1e30f;
    for (int j = 0; j < n; ++j) {
        const float cc = cost[j];
        if (cc < b0) { b2 = b1; b1 = b0; b0 = cc; }
#if PMO_MUT == 21
        else if (cc < b1) { b1 = cc; }
#else
        else if (cc < b1) { b2 = b1; b1 = cc; }
#endif
        else if (cc < b2) { b2 = cc; }
    }
#if PMO_MUT == 22
    int kk = k;
#else
    int kk = k < n ? k : n;
#endif
    float s = b0;
    if (kk >= 2) s += b1;
    if (kk >= 3) s += b2;
    return kk > 0 ? s / (float)kk : 2.0f;
}

/* One source's read-out test at a pixel (patchmatch.cu:714-733): X the pixel's 3-D point, lx = |X|,
 * c the source centre and n the normal. The source counts when the triangulation angle at X is at
 * least the minimum (cos <= cos_filter_tri), the surface faces the source (n . (c - X) > 0), the
 * homography exists (hok), the visibility probability of the cost m, E/(E + 0.5) with
 * E = ncc_norm exp(-m^2 inv2ncc), is not below q_min, and the geometric error is at most
 * filter_gmax. Returns 1 when it counts. */
static int readout_angles(const pmo_readout_params* k, const float X[3], float lx, const float c[3], const float n[3]) {
    f3 Sv = mk3(c[0] - X[0], c[1] - X[1], c[2] - X[2]);
    float ls = sqrtf(dot3(Sv, Sv));
    if (!(ls > 0.f)) return 0;
    float cost = -dot3(Sv, mk3(X[0], X[1], X[2])) / (ls * lx);
#if PMO_MUT == 23
    if (cost >= k->cos_filter_tri) return 0;
#else
    if (cost > k->cos_filter_tri) return 0;
#endif
    float cphi = dot3(Sv, mk3(n[0], n[1], n[2])) / ls;
#if PMO_MUT == 24
    if (cphi < 0.f) return 0;
#else
    if (cphi <= 0.f) return 0;
#endif
    return 1;
}

int pmo_readout_source(const pmo_readout_params* k, const float X[3], float lx, const float c[3], const float n[3], int hok, float m,
                       float geo) {
    if (!readout_angles(k, X, lx, c, n)) return 0;
    if (!hok) return 0;
    float e = k->ncc_norm * pmo_expf(-m * m * k->inv2ncc);
    if (e / (e + 0.5f) < k->q_min) return 0;
    if (geo > k->filter_gmax) return 0;
    return 1;
}

/* Count threshold of the read-out (densify.rs:446-454): max(1, min(min_num_consistent, nsrc)),
 * at most 255 (the count is a u8). */
uint32_t pmo_count_need(uint32_t min_num_consistent, uint32_t nsrc) {
#if PMO_MUT == 25
    uint32_t need = min_num_consistent;
#else
    uint32_t need = min_num_consistent < nsrc ? min_num_consistent : nsrc;
#endif
    if (need < 1) need = 1;
    if (need > 255) need = 255;
    return need;
}

/* ================================================================ kernels, per pixel */

/* one red-black half step at an active pixel (patchmatch.cu:439-657) */
static void half_pixel(const pmo_ctx* c, const launch_t* P, int view, const vstate_t* vs, int x, int y) {
    const lvl_t* R = &c->lv[view * c->nl + P->level];
    const c3d_kernel_view* kv = &c->in->views[view];
    const int w = vs->w, h = vs->h;
    const int nsrc = (int)kv->num_sources;
    const int pix = y * w + x;
    f4* pl = vs->plane;
    const f3 ray = pixel_ray(R, x, y);
    const float rayv[3] = {ray.x, ray.y, ray.z};

    /* 1. candidates: the current plane and one per region (patchmatch.cu:480-511) */
    f4 hyp[9];
    hyp[0] = pl[pix];
    for (int r = 0; r < 8; ++r) {
        hyp[1 + r] = hyp[0];
        int bx, by;
        if (pmo_region_pick(vs->cost, w, h, x, y, r, &bx, &by)) {
            const f4 q = pl[by * w + bx];
            const f3 rq = pixel_ray(R, bx, by);
            const float qv[4] = {q.x, q.y, q.z, q.w}, rqv[3] = {rq.x, rq.y, rq.z};
            float o[4];
            if (pmo_transfer_plane(qv, rqv, rayv, o)) hyp[1 + r] = (f4){o[0], o[1], o[2], o[3]};
        }
    }

    /* 2. cost table over the sources, then view selection and aggregation (patchmatch.cu:513-594) */
    rstats_t rs;
    ref_stats(c, R, x, y, &rs);
    const int vprev = vs->bview[pix];
    float prior[MAXSRC], mt[MAXSRC * 9], gt[MAXSRC * 9], wv[MAXSRC];
    for (int j = 0; j < nsrc; ++j) {
        const pair_t* pr = &c->pairs[view * MAXSRC + j];
        const lvl_t* S = &c->lv[pr->src * c->nl + P->level];
        const lpair_t* lp = &c->lpairs[((size_t)view * c->nl + P->level) * MAXSRC + j];
        float* m = mt + 9 * j;
        float* gg = gt + 9 * j;
        {
            float H[9];
            int ok = homography(R, lp, (float)x, (float)y, hyp[0].w, xyz(hyp[0]), H);
            prior[j] = view_prior(c, R, pr, H, ok, x, y, hyp[0]);
            if (!(prior[j] > 0.f)) continue; /* not evaluated */
            m[0] = ok ? ncc_cost(c, &rs, S, H, x, y) : 2.f;
            gg[0] = P->geometric ? fminf(geo_error(c, R, S, pr, x, y, hyp[0].w), c->gmax) : 0.f;
        }
        for (int i = 1; i < 9; ++i) {
            float H[9];
            int ok = homography(R, lp, (float)x, (float)y, hyp[i].w, xyz(hyp[i]), H);
            m[i] = ok ? ncc_cost(c, &rs, S, H, x, y) : 2.f;
            gg[i] = P->geometric ? fminf(geo_error(c, R, S, pr, x, y, hyp[i].w), c->gmax) : 0.f;
        }
    }
    const pmo_select_params sp = {P->tau_good, c->tau_bad, c->inv2beta2, c->prev_w, c->lambda, c->n1, c->n2};
    float den, agg[9];
    int bi9, vbest;
    pmo_select_views(&sp, nsrc, vprev, prior, mt, P->geometric ? gt : NULL, wv, &den, agg, &bi9, &vbest);
    f4 best = hyp[bi9];
    float bcost = agg[bi9];

    /* 3. refinement (patchmatch.cu:596-628): the six candidates from the pixel's random stream,
     * their costs over the sources with wv > 0, then acceptance */
    uint64_t st = pmo_rng_state(c->in->seed, kv->key, (uint32_t)P->level, (uint32_t)P->run_id, (uint32_t)P->t,
                                (uint32_t)P->color, (uint64_t)pix);
    float d_r = random_depth(&st, kv->depth_min, kv->depth_max);
    f3 n_r = random_normal(&st, ray);
    float u = 2.f * pmo_rng_next(&st) - 1.f;
    float inv = best.w > 0.f ? (1.f / best.w) * (1.f + u * P->eps) : -1.f;
    float d_p = inv > 0.f ? 1.f / inv : d_r;
    if (!(best.w > 0.f)) best.w = d_r;
    f3 n0 = xyz(best);
    f3 n_p = perturb_normal(&st, n0, ray, P->phi);
    f4 cand[7] = {mk4(n0, d_r), mk4(n_r, best.w), mk4(n_r, d_r), mk4(n0, d_p), mk4(n_p, best.w), mk4(n_p, d_p)};
    int ncand = 6;
    /* 4. coarse-scale candidate at weakly textured pixels (patchmatch.cu:629-653) */
    if (P->use_prior && rs.var < c->weak_var255) {
        f4 q = vs->prior[pix];
        if (q.w > 0.f && dot3(xyz(q), ray) < 0.f) cand[ncand++] = q;
    }
    float mc[7 * MAXSRC], gc[7 * MAXSRC];
    if (den > 0.f)
        for (int ci = 0; ci < ncand; ++ci)
            for (int j = 0; j < nsrc; ++j) {
                if (!(wv[j] > 0.f)) continue;
                const pair_t* pr = &c->pairs[view * MAXSRC + j];
                const lvl_t* S = &c->lv[pr->src * c->nl + P->level];
                const lpair_t* lp = &c->lpairs[((size_t)view * c->nl + P->level) * MAXSRC + j];
                float g = 0.f;
                mc[ci * nsrc + j] = pair_cost(c, P, &rs, R, S, pr, lp, x, y, cand[ci], &g);
                gc[ci * nsrc + j] = g;
            }
    int taken = pmo_refine_accept(&sp, ncand, nsrc, wv, den, mc, P->geometric ? gc : NULL, &bcost);
    if (taken >= 0) best = cand[taken];
    pl[pix] = best;
    vs->cost[pix] = bcost;
    vs->bview[pix] = (uint8_t)vbest;
}

/* run start / evaluation at one pixel (patchmatch.cu:659-692) */
static void eval_pixel(const pmo_ctx* c, const launch_t* P, int view, const vstate_t* vs, int x, int y, int mode) {
    const lvl_t* R = &c->lv[view * c->nl + P->level];
    const c3d_kernel_view* kv = &c->in->views[view];
    const int pix = y * vs->w + x;
    f3 ray = pixel_ray(R, x, y);
    f4 h = vs->plane[pix];
    /* random_init 1: every pixel (patchmatch.cu:668-676); 2: pixels without a positive depth */
    if (mode == 1 && (P->random_init == 1 || (P->random_init == 2 && !(h.w > 0.f)))) {
        uint64_t st = pmo_rng_state(c->in->seed, kv->key, (uint32_t)P->level, (uint32_t)P->run_id, 0xFFFFu, 0u,
                                    (uint64_t)pix);
        float d = random_depth(&st, kv->depth_min, kv->depth_max);
        f3 n = random_normal(&st, ray);
        h = mk4(n, d);
        vs->plane[pix] = h;
    }
    float cc = 2.f;
    if (h.w > 0.f) {
        rstats_t rs;
        ref_stats(c, R, x, y, &rs);
        cc = topk_cost(c, P, &rs, R, view, (int)kv->num_sources, x, y, h);
    }
    vs->cost[pix] = cc;
    if (mode == 1) vs->bview[pix] = 255;
}

/* final read-out count at one pixel (patchmatch.cu:694-737): each source's test is
 * pmo_readout_source */
static uint8_t filter_pixel(const pmo_ctx* c, int level, int view, const f4* plane, int w, int x, int y) {
    const lvl_t* R = &c->lv[view * c->nl + level];
    const c3d_kernel_view* kv = &c->in->views[view];
    const int pix = y * w + x;
    f4 h = plane[pix];
    int count = 0;
    if (h.w > 0.f) {
        rstats_t rs;
        ref_stats(c, R, x, y, &rs);
        f3 ray = pixel_ray(R, x, y);
        const float X[3] = {ray.x * h.w, ray.y * h.w, h.w}, nv[3] = {h.x, h.y, h.z};
        float lx = sqrtf(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
        const pmo_readout_params k = {c->cos_filter_tri, c->ncc_norm, c->inv2ncc, c->q_min, c->filter_gmax};
        for (int j = 0; j < (int)kv->num_sources; ++j) {
            const pair_t* pr = &c->pairs[view * MAXSRC + j];
            const lvl_t* S = &c->lv[pr->src * c->nl + level];
            const lpair_t* lp = &c->lpairs[((size_t)view * c->nl + level) * MAXSRC + j];
            /* the photometric and geometric terms only where the angle tests pass (they decide
             * first); pmo_readout_source repeats those tests */
            if (!readout_angles(&k, X, lx, pr->c, nv)) continue;
            float H[9];
            int hok = homography(R, lp, (float)x, (float)y, h.w, xyz(h), H);
            float m = hok ? ncc_cost(c, &rs, S, H, x, y) : 2.f;
            float ge = hok ? geo_error(c, R, S, pr, x, y, h.w) : 0.f;
            count += pmo_readout_source(&k, X, lx, pr->c, nv, hok, m, ge);
        }
    }
    return (uint8_t)(count < 255 ? count : 255);
}

/* ================================================================ schedule (patchmatch.rs:752-868) */

static void states_to_planes(const c3d_view_state* s, f4* pl) {
    size_t n = (size_t)s->width * s->height;
    for (size_t i = 0; i < n; ++i) {
        pl[i].x = s->normal[3 * i + 0];
        pl[i].y = s->normal[3 * i + 1];
        pl[i].z = s->normal[3 * i + 2];
        pl[i].w = s->depth[i];
    }
}
static void planes_to_states(const f4* pl, c3d_view_state* s) {
    size_t n = (size_t)s->width * s->height;
    for (size_t i = 0; i < n; ++i) {
        s->normal[3 * i + 0] = pl[i].x;
        s->normal[3 * i + 1] = pl[i].y;
        s->normal[3 * i + 2] = pl[i].z;
        s->depth[i] = pl[i].w;
    }
}

static int set_snapshot(pmo_ctx* c, const c3d_depth_snapshot* snap) {
    for (int v = 0; v < c->nv; ++v) c->snap[v] = NULL;
    if (!snap) return 0;
    if ((int)snap->level >= c->nl) {
        set_err("snapshot level out of range");
        return C3D_INVALID_ARGUMENT;
    }
    for (uint32_t v = 0; v < snap->num_maps && (int)v < c->nv; ++v) c->snap[v] = snap->maps[v];
    c->snap_level = (int)snap->level;
    return 0;
}

static int check_states(const pmo_ctx* c, uint32_t level, const uint32_t* views, uint32_t n,
                        const c3d_view_state* states) {
    if ((int)level >= c->nl) {
        set_err("level out of range");
        return C3D_INVALID_ARGUMENT;
    }
    for (uint32_t i = 0; i < n; ++i) {
        if ((int)views[i] >= c->nv) {
            set_err("view index out of range");
            return C3D_INVALID_ARGUMENT;
        }
        const lvl_t* L = &c->lv[views[i] * c->nl + level];
        if ((int)states[i].width != L->w || (int)states[i].height != L->h) {
            set_err("state size does not match the level size");
            return C3D_INVALID_ARGUMENT;
        }
    }
    return 0;
}

static void launch_constants(const pmo_ctx* c, launch_t* k, uint32_t t) {
    const c3d_pm_params* pm = &c->in->pm;
    float tf = (float)t;
    float eps0 = k->geometric ? pm->eps0_geometric : pm->eps0_photometric;
    float phi0 = k->geometric ? pm->phi0_geometric_deg : pm->phi0_photometric_deg;
    float p = 1.0f;
    for (uint32_t i = 0; i < t && p > 0.0f; ++i) p *= 0.5f;
    k->t = (int)t;
    k->tau_good = pm->tau0 * expf(-tf * tf / pm->tau_alpha);
    k->eps = fmaxf(eps0 * p, pm->eps_min);
    k->phi = deg2rad(fmaxf(phi0 * p, pm->phi_min_deg));
}

static void run_view(pmo_ctx* c, const launch_t* base, int view, vstate_t* vs, uint32_t iters, int row0, int row1) {
    const int w = vs->w, h = vs->h;
    launch_t k = *base;
    /* pm_eval mode 1 over all pixels (over the band and its 24-row propagation reach for a band run) */
    int e0 = row0 - 24 > 0 ? row0 - 24 : 0, e1 = row1 + 24 < h ? row1 + 24 : h;
    if (row0 == 0 && row1 == h) e0 = 0, e1 = h;
#pragma omp parallel for schedule(dynamic, 1) num_threads(c->nthreads)
    for (int y = e0; y < e1; ++y)
        for (int x = 0; x < w; ++x) eval_pixel(c, &k, view, vs, x, y, 1);
    for (uint32_t t = 0; t < iters; ++t) {
        launch_constants(c, &k, t);
        for (int color = 0; color < 2; ++color) {
            k.color = color;
#pragma omp parallel for schedule(dynamic, 1) num_threads(c->nthreads)
            for (int y = row0; y < row1; ++y)
                for (int x = ((y + color) & 1); x < w; x += 2) half_pixel(c, &k, view, vs, x, y);
        }
    }
}

int pmo_run(pmo_ctx* c, const uint32_t* views, uint32_t n, c3d_view_state* states, const c3d_run_params* p,
            const c3d_depth_snapshot* snap) {
    int rc = check_states(c, p->level, views, n, states);
    if (rc) return rc;
    if (p->geometric) {
        if (!snap) {
            set_err("a geometric run needs a snapshot");
            return C3D_INVALID_ARGUMENT;
        }
        rc = set_snapshot(c, snap);
        if (rc) return rc;
    }
    int use_prior = p->use_prior;
    for (uint32_t i = 0; i < n; ++i)
        if (!states[i].prior) use_prior = 0;
    for (uint32_t i = 0; i < n; ++i) {
        c3d_view_state* s = &states[i];
        size_t np = (size_t)s->width * s->height;
        vstate_t vs;
        vs.w = (int)s->width;
        vs.h = (int)s->height;
        vs.plane = (f4*)calloc(np, sizeof(f4));
        vs.cost = (float*)calloc(np, sizeof(float));
        vs.bview = (uint8_t*)calloc(np, 1);
        vs.prior = use_prior ? (const f4*)s->prior : NULL;
        if (p->random_init != 1) states_to_planes(s, vs.plane);
        launch_t k;
        memset(&k, 0, sizeof k);
        k.level = (int)p->level;
        k.geometric = p->geometric;
        k.run_id = (int)p->run_id;
        k.random_init = p->random_init;
        k.use_prior = use_prior;
        run_view(c, &k, (int)views[i], &vs, p->iterations, 0, vs.h);
        planes_to_states(vs.plane, s);
        memcpy(s->cost, vs.cost, np * sizeof(float));
        free(vs.plane);
        free(vs.cost);
        free(vs.bview);
    }
    return 0;
}

int pmo_run_rows(pmo_ctx* c, uint32_t view, c3d_view_state* s, const c3d_run_params* p, uint32_t row0, uint32_t row1) {
    int rc = check_states(c, p->level, &view, 1, s);
    if (rc) return rc;
    if (p->geometric) {
        set_err("pmo_run_rows: photometric runs only");
        return C3D_INVALID_ARGUMENT;
    }
    size_t np = (size_t)s->width * s->height;
    vstate_t vs;
    vs.w = (int)s->width;
    vs.h = (int)s->height;
    vs.plane = (f4*)calloc(np, sizeof(f4));
    vs.cost = (float*)calloc(np, sizeof(float));
    vs.bview = (uint8_t*)calloc(np, 1);
    vs.prior = NULL;
    if (p->random_init != 1) states_to_planes(s, vs.plane);
    launch_t k;
    memset(&k, 0, sizeof k);
    k.level = (int)p->level;
    k.run_id = (int)p->run_id;
    k.random_init = p->random_init;
    if (row1 > (uint32_t)vs.h) row1 = (uint32_t)vs.h;
    run_view(c, &k, (int)view, &vs, p->iterations, (int)row0, (int)row1);
    planes_to_states(vs.plane, s);
    memcpy(s->cost, vs.cost, np * sizeof(float));
    free(vs.plane);
    free(vs.cost);
    free(vs.bview);
    return 0;
}

int pmo_evaluate(pmo_ctx* c, uint32_t level, const uint32_t* views, uint32_t n, const c3d_view_state* states,
                 int geometric, const c3d_depth_snapshot* snap, float* const* costs) {
    int rc = check_states(c, level, views, n, states);
    if (rc) return rc;
    if (geometric) {
        if (!snap) {
            set_err("a geometric evaluation needs a snapshot");
            return C3D_INVALID_ARGUMENT;
        }
        rc = set_snapshot(c, snap);
        if (rc) return rc;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const c3d_view_state* s = &states[i];
        size_t np = (size_t)s->width * s->height;
        vstate_t vs;
        vs.w = (int)s->width;
        vs.h = (int)s->height;
        vs.plane = (f4*)calloc(np, sizeof(f4));
        vs.cost = costs[i];
        vs.bview = (uint8_t*)calloc(np, 1);
        vs.prior = NULL;
        states_to_planes(s, vs.plane);
        launch_t k;
        memset(&k, 0, sizeof k);
        k.level = (int)level;
        k.geometric = geometric;
#pragma omp parallel for schedule(dynamic, 1) num_threads(c->nthreads)
        for (int y = 0; y < vs.h; ++y)
            for (int x = 0; x < vs.w; ++x) eval_pixel(c, &k, (int)views[i], &vs, x, y, 0);
        free(vs.plane);
        free(vs.bview);
    }
    return 0;
}

int pmo_filter(pmo_ctx* c, const uint32_t* views, uint32_t n, const c3d_view_state* states,
               const c3d_depth_snapshot* snap, uint8_t* const* counts) {
    if (!snap) {
        set_err("filter needs a snapshot");
        return C3D_INVALID_ARGUMENT;
    }
    int rc = check_states(c, snap->level, views, n, states);
    if (rc) return rc;
    rc = set_snapshot(c, snap);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        const c3d_view_state* s = &states[i];
        size_t np = (size_t)s->width * s->height;
        f4* pl = (f4*)calloc(np, sizeof(f4));
        states_to_planes(s, pl);
        int w = (int)s->width, h = (int)s->height;
        uint8_t* out = counts[i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(c->nthreads)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) out[y * w + x] = filter_pixel(c, (int)snap->level, (int)views[i], pl, w, x, y);
        free(pl);
    }
    return 0;
}

/* joint bilateral upsampling (patchmatch.cu:739-801) */
int pmo_upsample(pmo_ctx* c, uint32_t level, const uint32_t* views, uint32_t n, const c3d_view_state* low,
                 float sigma_s, float sigma_c, c3d_view_state* out) {
    if (level == 0 || (int)level >= c->nl) {
        set_err("upsample level must be 1..num_levels-1");
        return C3D_INVALID_ARGUMENT;
    }
    int rc = check_states(c, level - 1, views, n, low);
    if (rc) return rc;
    rc = check_states(c, level, views, n, out);
    if (rc) return rc;
    float inv2s = 1.0f / (2.0f * sigma_s * sigma_s), inv2c = 1.0f / (2.0f * sigma_c * sigma_c);
    for (uint32_t i = 0; i < n; ++i) {
        const lvl_t* R = &c->lv[views[i] * c->nl + level];
        const lvl_t* L = &c->lv[views[i] * c->nl + level - 1];
        size_t nl_ = (size_t)L->w * L->h;
        f4* lo = (f4*)calloc(nl_, sizeof(f4));
        states_to_planes(&low[i], lo);
        c3d_view_state* so = &out[i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(c->nthreads)
        for (int y = 0; y < R->h; ++y)
            for (int x = 0; x < R->w; ++x) {
                f3 rp = pixel_ray(R, x, y);
                float ip = (float)R->img[(size_t)y * R->w + x] * (1.0f / 255.0f);
                float px = ((float)x - 0.5f) * 0.5f, py = ((float)y - 0.5f) * 0.5f;
                int qx0 = (int)roundf(px), qy0 = (int)roundf(py);
                float sw = 0.f, sd = 0.f;
                f3 sn = mk3(0.f, 0.f, 0.f);
                float nd = 1e30f, ndep = 0.f;
                f3 nn = mk3(0.f, 0.f, 0.f);
                for (int qy = qy0 - 2; qy <= qy0 + 2; ++qy) {
                    if (qy < 0 || qy >= L->h) continue;
                    for (int qx = qx0 - 2; qx <= qx0 + 2; ++qx) {
                        if (qx < 0 || qx >= L->w) continue;
                        f4 q = lo[qy * L->w + qx];
                        if (!(q.w > 0.f)) continue;
                        f3 nq = xyz(q);
                        float den = dot3(nq, rp);
                        if (!(den < -1e-6f)) continue;
                        float d = q.w * dot3(nq, pixel_ray(L, qx, qy)) / den;
                        if (!(d > 0.f) || !isfinite(d)) continue;
                        float ds = ((float)qx - px) * ((float)qx - px) + ((float)qy - py) * ((float)qy - py);
                        if (ds < nd) {
                            nd = ds;
                            ndep = d;
                            nn = nq;
                        }
                        float iq = (float)L->img[(size_t)qy * L->w + qx] * (1.0f / 255.0f);
                        float wgt = pmo_expf(-ds * inv2s - (ip - iq) * (ip - iq) * inv2c);
                        sw += wgt;
                        sd += wgt * d;
                        sn.x += wgt * nq.x;
                        sn.y += wgt * nq.y;
                        sn.z += wgt * nq.z;
                    }
                }
                f4 o = {0.f, 0.f, 0.f, 0.f};
                float d = 0.f;
                f3 nv = mk3(0.f, 0.f, 0.f);
                int ok = 0;
                if (sw > 1e-12f) {
                    d = sd / sw;
                    nv = sn;
                    ok = 1;
                } else if (nd < 1e29f) {
                    d = ndep;
                    nv = nn;
                    ok = 1;
                }
                if (ok && dot3(nv, nv) > 1e-24f) {
                    nv = nrm3(nv);
                    if (dot3(nv, rp) >= 0.f) nv = mk3(-nv.x, -nv.y, -nv.z);
                    o = mk4(nv, d);
                }
                size_t pi = (size_t)y * R->w + x;
                so->depth[pi] = o.w;
                so->normal[3 * pi + 0] = o.x;
                so->normal[3 * pi + 1] = o.y;
                so->normal[3 * pi + 2] = o.z;
            }
        free(lo);
    }
    return 0;
}

/* 5x5 median plane filter (patchmatch.cu:803-849) */
int pmo_median(pmo_ctx* c, uint32_t level, const uint32_t* views, uint32_t n, const c3d_view_state* in,
               c3d_view_state* out) {
    int rc = check_states(c, level, views, n, in);
    if (rc) return rc;
    rc = check_states(c, level, views, n, out);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        const lvl_t* R = &c->lv[views[i] * c->nl + level];
        const int w = R->w, h = R->h;
        size_t np = (size_t)w * h;
        f4* src = (f4*)calloc(np, sizeof(f4));
        states_to_planes(&in[i], src);
        c3d_view_state* so = &out[i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(c->nthreads)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                f4 self = src[y * w + x];
                f4 o = self;
                if (self.w > 0.f) {
                    f3 rp = pixel_ray(R, x, y);
                    float ds[25];
                    int qs[25];
                    int cnt = 0;
                    int y0 = y - 2 > 0 ? y - 2 : 0, y1 = y + 3 < h ? y + 3 : h;
                    int x0 = x - 2 > 0 ? x - 2 : 0, x1 = x + 3 < w ? x + 3 : w;
                    for (int qy = y0; qy < y1; ++qy)
                        for (int qx = x0; qx < x1; ++qx) {
                            f4 q = src[qy * w + qx];
                            if (!(q.w > 0.f)) continue;
                            f3 nq = xyz(q);
                            float den = dot3(nq, rp);
                            if (!(den < -1e-6f)) continue;
                            float d = q.w * dot3(nq, pixel_ray(R, qx, qy)) / den;
                            if (!(d > 0.f) || !isfinite(d)) continue;
                            int qi = qy * w + qx;
                            int k = cnt++;
                            while (k > 0 && (ds[k - 1] > d || (ds[k - 1] == d && qs[k - 1] > qi))) {
                                ds[k] = ds[k - 1];
                                qs[k] = qs[k - 1];
                                --k;
                            }
                            ds[k] = d;
                            qs[k] = qi;
                        }
                    if (cnt > 0) {
                        int m = (cnt - 1) / 2;
                        f4 q = src[qs[m]];
                        o = q;
                        o.w = ds[m];
                    }
                }
                size_t pi = (size_t)y * w + x;
                so->depth[pi] = o.w;
                so->normal[3 * pi + 0] = o.x;
                so->normal[3 * pi + 1] = o.y;
                so->normal[3 * pi + 2] = o.z;
                if (so->cost && in[i].cost && so->cost != in[i].cost) so->cost[pi] = in[i].cost[pi];
            }
        free(src);
    }
    return 0;
}

/* ---------------------------------------------------------------- compute_depth_maps schedule */

/* ViewState::new at `level` (kernel.rs:154-159): depth 0, normal 0, cost 2; prior zeroed when asked */
static c3d_view_state* new_states(const pmo_ctx* c, uint32_t level, const uint32_t* views, uint32_t n, int prior) {
    c3d_view_state* s = (c3d_view_state*)calloc(n, sizeof(c3d_view_state));
    for (uint32_t i = 0; i < n; ++i) {
        const lvl_t* L = &c->lv[views[i] * c->nl + level];
        size_t np = (size_t)L->w * L->h;
        s[i].width = (uint32_t)L->w;
        s[i].height = (uint32_t)L->h;
        s[i].depth = (float*)calloc(np, sizeof(float));
        s[i].normal = (float*)calloc(3 * np, sizeof(float));
        s[i].cost = (float*)malloc(np * sizeof(float));
        for (size_t p = 0; p < np; ++p) s[i].cost[p] = 2.f;
        s[i].prior = prior ? (float*)calloc(4 * np, sizeof(float)) : NULL;
    }
    return s;
}

static void free_states(c3d_view_state* s, uint32_t n) {
    if (!s) return;
    for (uint32_t i = 0; i < n; ++i) {
        free(s[i].depth);
        free(s[i].normal);
        free(s[i].cost);
        free(s[i].prior);
    }
    free(s);
}

static float** new_maps(const c3d_view_state* s, uint32_t n) {
    float** m = (float**)calloc(n, sizeof(float*));
    for (uint32_t i = 0; i < n; ++i) m[i] = (float*)malloc((size_t)s[i].width * s[i].height * sizeof(float));
    return m;
}

static void free_maps(float** m, uint32_t n) {
    if (!m) return;
    for (uint32_t i = 0; i < n; ++i) free(m[i]);
    free(m);
}

/* The snapshot closure of densify.rs:335-342: every active view's current depth at `level`. */
static c3d_depth_snapshot take_snapshot(const pmo_ctx* c, const uint32_t* views, uint32_t n, const c3d_view_state* s,
                                        uint32_t level, uint64_t id, float** copies, const float** maps) {
    for (int v = 0; v < c->nv; ++v) maps[v] = NULL;
    for (uint32_t i = 0; i < n; ++i) {
        memcpy(copies[i], s[i].depth, (size_t)s[i].width * s[i].height * sizeof(float));
        maps[views[i]] = copies[i];
    }
    c3d_depth_snapshot snap = {id, level, (uint32_t)c->nv, maps};
    return snap;
}

/* compute_depth_maps over all scales (densify.rs:350-454). `out`, `counts`, `raw_depth` are at the
 * top scale. */
int pmo_densify(pmo_ctx* c, const uint32_t* views, uint32_t n, const c3d_schedule* sch, c3d_view_state* out,
                uint8_t* const* counts, float* const* raw_depth) {
    const uint32_t top = (uint32_t)c->nl - 1;
    int rc = check_states(c, top, views, n, out);
    if (rc) return rc;
    uint32_t run_id = sch->first_run_id;
    uint64_t snap_id = 0;
    const float** maps = (const float**)calloc((size_t)c->nv, sizeof(float*));
    c3d_view_state* st = NULL;
    for (uint32_t level = 0; !rc && level <= top; ++level) {
        /* params.rs:476-506 via the schedule struct; use_prior densify.rs:366 */
        const int use_prior = level > 0 && c->in->pm.weak_texture_var > 0.f;
        if (level == 0) {
            st = new_states(c, 0, views, n, 0);
            c3d_run_params p = {0, 0, 1, sch->photometric_iters, run_id++, 0};
            rc = pmo_run(c, views, n, st, &p, NULL);
        } else {
            /* detail restorer (densify.rs:366-409) */
            c3d_view_state* up = new_states(c, level, views, n, use_prior);
            rc = pmo_upsample(c, level, views, n, st, sch->jbu_sigma_spatial, sch->jbu_sigma_color, up);
            free_states(st, n);
            st = up;
            if (rc) break;
            if (use_prior)
                for (uint32_t i = 0; i < n; ++i) {
                    size_t np = (size_t)up[i].width * up[i].height;
                    for (size_t q = 0; q < np; ++q) {
                        up[i].prior[4 * q + 0] = up[i].normal[3 * q + 0];
                        up[i].prior[4 * q + 1] = up[i].normal[3 * q + 1];
                        up[i].prior[4 * q + 2] = up[i].normal[3 * q + 2];
                        up[i].prior[4 * q + 3] = up[i].depth[q];
                    }
                }
            float** c_init = new_maps(up, n);
            float** c_photo = new_maps(up, n);
            c3d_view_state* photo = new_states(c, level, views, n, use_prior);
            for (uint32_t i = 0; i < n; ++i) {
                size_t np = (size_t)up[i].width * up[i].height;
                if (use_prior) memcpy(photo[i].prior, up[i].prior, 4 * np * sizeof(float));
                if (!sch->restorer_random_init) {
                    memcpy(photo[i].depth, up[i].depth, np * sizeof(float));
                    memcpy(photo[i].normal, up[i].normal, 3 * np * sizeof(float));
                    memcpy(photo[i].cost, up[i].cost, np * sizeof(float));
                }
            }
            rc = pmo_evaluate(c, level, views, n, up, 0, NULL, c_init);
            c3d_run_params p = {level, 0, sch->restorer_random_init ? 1 : 0, sch->photometric_iters_upper, run_id++,
                                use_prior};
            if (!rc) rc = pmo_run(c, views, n, photo, &p, NULL);
            if (!rc) rc = pmo_evaluate(c, level, views, n, photo, 0, NULL, c_photo);
            const float xi = sch->restorer_threshold;
            for (uint32_t i = 0; !rc && i < n; ++i) {
                size_t np = (size_t)up[i].width * up[i].height;
                for (size_t q = 0; q < np; ++q)
                    if (up[i].depth[q] <= 0.f || c_init[i][q] - c_photo[i][q] > xi) {
                        up[i].depth[q] = photo[i].depth[q];
                        up[i].normal[3 * q + 0] = photo[i].normal[3 * q + 0];
                        up[i].normal[3 * q + 1] = photo[i].normal[3 * q + 1];
                        up[i].normal[3 * q + 2] = photo[i].normal[3 * q + 2];
                    }
            }
            free_states(photo, n);
            free_maps(c_init, n);
            free_maps(c_photo, n);
        }
        float** copies = new_maps(st, n);
        for (uint32_t r = 0; !rc && r < sch->geometric_rounds; ++r) {
            c3d_depth_snapshot snap = take_snapshot(c, views, n, st, level, ++snap_id, copies, maps);
            c3d_run_params g = {level, 1, 0, sch->geometric_iters, run_id++, use_prior};
            rc = pmo_run(c, views, n, st, &g, &snap);
        }
        free_maps(copies, n);
    }
    /* priors dropped, median, read-out and threshold at the top scale (densify.rs:428-454) */
    for (uint32_t i = 0; st && i < n; ++i) {
        free(st[i].prior);
        st[i].prior = NULL;
    }
    if (!rc && c->in->filter.median_filter) {
        c3d_view_state* med = new_states(c, top, views, n, 0);
        rc = pmo_median(c, top, views, n, st, med);
        free_states(st, n);
        st = med;
    }
    uint8_t** cnt = (uint8_t**)calloc(n, sizeof(uint8_t*));
    float** copies = rc ? NULL : new_maps(st, n);
    if (!rc) {
        for (uint32_t i = 0; i < n; ++i)
            cnt[i] = counts ? counts[i] : (uint8_t*)malloc((size_t)st[i].width * st[i].height);
        c3d_depth_snapshot snap = take_snapshot(c, views, n, st, top, ++snap_id, copies, maps);
        rc = pmo_filter(c, views, n, st, &snap, cnt);
    }
    for (uint32_t i = 0; !rc && i < n; ++i) {
        const c3d_kernel_view* kv = &c->in->views[views[i]];
        const uint32_t need = pmo_count_need(c->in->filter.min_num_consistent, kv->num_sources);
        size_t np = (size_t)st[i].width * st[i].height;
        if (raw_depth && raw_depth[i]) memcpy(raw_depth[i], st[i].depth, np * sizeof(float));
        for (size_t q = 0; q < np; ++q) {
            int keep = st[i].depth[q] > 0.f && cnt[i][q] >= need;
            out[i].depth[q] = keep ? st[i].depth[q] : 0.f;
            out[i].normal[3 * q + 0] = keep ? st[i].normal[3 * q + 0] : 0.f;
            out[i].normal[3 * q + 1] = keep ? st[i].normal[3 * q + 1] : 0.f;
            out[i].normal[3 * q + 2] = keep ? st[i].normal[3 * q + 2] : 0.f;
            if (out[i].cost) out[i].cost[q] = st[i].cost[q];
        }
        /* read-out post-processing (densify.rs:455-462): speckles judged on the raw cost, gaps
         * filled from the raw cost with the top-scale gray image */
        float* cw = out[i].cost ? out[i].cost : (float*)malloc(np * sizeof(float));
        if (!out[i].cost) memcpy(cw, st[i].cost, np * sizeof(float));
        if (sch->post.remove_speckles)
            pmo_remove_speckles(out[i].depth, out[i].normal, st[i].cost, st[i].width, st[i].height, &sch->post, NULL);
        if (sch->post.fill_holes) {
            if (!sch->post_k) rc = C3D_INVALID_ARGUMENT;
            else
                pmo_fill_holes(out[i].depth, out[i].normal, cw, st[i].cost, c->in->views[views[i]].levels[top].gray,
                               sch->post_k + 4 * i, st[i].width, st[i].height, &sch->post, NULL);
        }
        if (!out[i].cost) free(cw);
    }
    for (uint32_t i = 0; i < n; ++i)
        if (!counts && cnt[i]) free(cnt[i]);
    free(cnt);
    free_maps(copies, n);
    free_states(st, n);
    free(maps);
    return rc;
}

/* ================================================================ single-hypothesis probes */

float pmo_ncc_cost_f32(pmo_ctx* c, uint32_t level, uint32_t view, uint32_t slot, int x, int y, const float plane[4]) {
    const lvl_t* R = &c->lv[view * c->nl + level];
    const pair_t* pr = &c->pairs[view * MAXSRC + slot];
    const lvl_t* S = &c->lv[pr->src * c->nl + level];
    const lpair_t* lp = &c->lpairs[((size_t)view * c->nl + level) * MAXSRC + slot];
    rstats_t rs;
    ref_stats(c, R, x, y, &rs);
    float H[9];
    f3 n = mk3(plane[0], plane[1], plane[2]);
    if (!homography(R, lp, (float)x, (float)y, plane[3], n, H)) return 2.0f;
    return ncc_cost(c, &rs, S, H, x, y);
}

/* FP64 evaluation of the reference formula (patchmatch.cu:199-290) in [0,1] intensity units */
static double ref_at_d(const lvl_t* R, int x, int y) {
    if (x < 0 || y < 0 || x >= R->w || y >= R->h) return 0.0;
    return (double)R->img[(size_t)y * R->w + x] / 255.0;
}
static double src_sample_d(const lvl_t* S, double u, double v) {
    double xf = floor(u), yf = floor(v);
    double ax = u - xf, ay = v - yf;
    double p00 = 0, p10 = 0, p01 = 0, p11 = 0;
    if (xf >= -1.0 && xf <= (double)(S->w - 1) && yf >= -1.0 && yf <= (double)(S->h - 1)) {
        int x0 = (int)xf, y0 = (int)yf;
        int xin0 = x0 >= 0, xin1 = x0 + 1 < S->w;
        if (y0 >= 0) {
            if (xin0) p00 = S->img[(size_t)y0 * S->w + x0];
            if (xin1) p10 = S->img[(size_t)y0 * S->w + x0 + 1];
        }
        if (y0 + 1 < S->h) {
            if (xin0) p01 = S->img[(size_t)(y0 + 1) * S->w + x0];
            if (xin1) p11 = S->img[(size_t)(y0 + 1) * S->w + x0 + 1];
        }
    }
    if (PMO_MUT == 4) {
        double t_ = ax;
        ax = ay;
        ay = t_;
    }
    double top = p00 + ax * (p10 - p00), bot = p01 + ax * (p11 - p01);
    return (top + ay * (bot - top)) / 255.0;
}

double pmo_ncc_cost_f64(pmo_ctx* c, uint32_t level, uint32_t view, uint32_t slot, int x, int y, const double pl[4]) {
    const lvl_t* R = &c->lv[view * c->nl + level];
    const pair_t* pr = &c->pairs[view * MAXSRC + slot];
    const lvl_t* S = &c->lv[pr->src * c->nl + level];
    const c3d_kernel_view* kv = &c->in->views[view];
    double A[9], b[3];
    level_pair_d(kv->levels[level].k, c->in->views[pr->src].levels[level].k, pr->r, pr->t, A, b);
    const int WN = c->wn;
    /* reference statistics */
    double ic = ref_at_d(R, x, y), W = 0, Sm = 0, SS = 0;
    double wk[MAXWN2], ik[MAXWN2];
    for (int a = 0; a < WN; ++a) {
        int dy = -c->wr + a * c->wstep;
        for (int bb = 0; bb < WN; ++bb) {
            int dx = -c->wr + bb * c->wstep;
            double v = ref_at_d(R, x + dx, y + dy);
            double d = ic - v;
            double w = exp(-(double)(dx * dx + dy * dy) * c->inv2s_d - d * d * c->inv2c_d);
            wk[a * WN + bb] = w;
            ik[a * WN + bb] = v;
            W += w;
            Sm += w * v;
            SS += w * v * v;
        }
    }
    double mu = Sm / W, var = SS / W - mu * mu;
    /* homography */
    double fx = R->fx, fy = R->fy, cx = R->cx, cy = R->cy;
    double ray[3] = {((double)x - cx) / fx, ((double)y - cy) / fy, 1.0};
    double delta = pl[3] * (pl[0] * ray[0] + pl[1] * ray[1] + pl[2] * ray[2]);
    if (!(delta < -1e-12)) return 2.0;
    double u0 = pl[0] / fx, u1 = pl[1] / fy, u2 = pl[2] - u0 * cx - u1 * cy;
    double H[9];
    for (int r = 0; r < 3; ++r) {
        double br = b[r] / delta;
        H[3 * r + 0] = A[3 * r + 0] + br * u0;
        H[3 * r + 1] = A[3 * r + 1] + br * u1;
        H[3 * r + 2] = A[3 * r + 2] + br * u2;
    }
    double z = H[6] * x + H[7] * y + H[8];
    if (!(z > 1e-8)) return 2.0;
    double u = (H[0] * x + H[1] * y + H[2]) / z, v = (H[3] * x + H[4] * y + H[5]) / z;
    if (!(u >= 0 && v >= 0 && u <= S->w - 1 && v <= S->h - 1)) return 2.0;
    if (var < 1e-5) return 2.0;
    double Ss = 0, SSs = 0, IS = 0;
    for (int a = 0; a < WN; ++a) {
        int dy = -c->wr + a * c->wstep;
        for (int bb = 0; bb < WN; ++bb) {
            int dx = -c->wr + bb * c->wstep;
            double sx = x + dx, sy = y + dy;
            double zz = H[6] * sx + H[7] * sy + H[8];
            double sv = 0;
            if (zz > 1e-8) sv = src_sample_d(S, (H[0] * sx + H[1] * sy + H[2]) / zz, (H[3] * sx + H[4] * sy + H[5]) / zz);
            double w = wk[a * WN + bb];
            Ss += w * sv;
            SSs += w * sv * sv;
            IS += w * ik[a * WN + bb] * sv;
        }
    }
    double mus = Ss / W, vars = SSs / W - mus * mus;
    if (vars < 1e-5) return 2.0;
    double cov = IS / W - (PMO_MUT == 6 ? 0.0 : mu * mus);
    double cc = 1.0 - cov / sqrt(var * vars);
    return cc < 0 ? 0 : (cc > 2 ? 2 : cc);
}
