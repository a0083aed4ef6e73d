/* fusion_oracle.c — CPU restatement of the agreement fusion (fuse_consistency,
 * crates/dense/src/fusion.rs:202-350 with ConsCtx::evaluate :420-490, mark_hits :360-383, inv_var
 * :78-89, point_key :54-56, Cam::project/backproject/normal_world :146-176) — TEST INFRASTRUCTURE
 * ONLY: the checker of cumulus3d_amd/csrc/fusion.hip.
 *
 * Views in index order; within a view, pixels in row-major order. (The reference processes the
 * rows of a view in parallel; a pixel only marks pixels of other views, so the order inside a view
 * does not change the result.) Double precision throughout, as in the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/c3d_fusion.h"

typedef struct {
    const c3d_fusion_view* v;
} cam_t;

/* Cam::project (fusion.rs:146-155): returns 0 when the point is not in front */
static int project(const c3d_fusion_view* c, const double x[3], double* u, double* v, double* z) {
    const double* r = c->r;
    double cz = r[6] * x[0] + r[7] * x[1] + r[8] * x[2] + c->t[2];
    if (cz <= 0.0) return 0;
    double cx = r[0] * x[0] + r[1] * x[1] + r[2] * x[2] + c->t[0];
    double cy = r[3] * x[0] + r[4] * x[1] + r[5] * x[2] + c->t[1];
    *u = c->k[0] * cx / cz + c->k[2];
    *v = c->k[1] * cy / cz + c->k[3];
    *z = cz;
    return 1;
}

/* Cam::backproject (fusion.rs:157-166) */
static void backproject(const c3d_fusion_view* c, double col, double row, double d, double out[3]) {
    double p[3] = {(col - c->k[2]) / c->k[0] * d - c->t[0], (row - c->k[3]) / c->k[1] * d - c->t[1], d - c->t[2]};
    const double* r = c->r;
    out[0] = r[0] * p[0] + r[3] * p[1] + r[6] * p[2];
    out[1] = r[1] * p[0] + r[4] * p[1] + r[7] * p[2];
    out[2] = r[2] * p[0] + r[5] * p[1] + r[8] * p[2];
}

/* Cam::normal_world (fusion.rs:168-176) */
static void normal_world(const c3d_fusion_view* c, const float* n, double out[3]) {
    const double* r = c->r;
    double a = n[0], b = n[1], e = n[2];
    out[0] = r[0] * a + r[3] * b + r[6] * e;
    out[1] = r[1] * a + r[4] * b + r[7] * e;
    out[2] = r[2] * a + r[5] * b + r[8] * e;
}

/* inv_var (fusion.rs:78-89) */
static double inv_var(const c3d_fusion_params* p, const c3d_fusion_view* c, size_t pix, double d) {
    if (p->inverse_variance && c->cost && c->baseline > 0.0) {
        double cc = c->cost[pix];
        cc = cc < 0.0 ? 0.0 : (cc > 2.0 ? 2.0 : cc);
        double spx = p->sigma_px0 + p->sigma_px_slope * cc;
        double sd = d * d * spx / (c->k[0] * c->baseline);
        double v = sd * sd;
        return 1.0 / (v > 1e-300 ? v : 1e-300);
    }
    return 1.0;
}

static double gray_at(const c3d_fusion_view* c, size_t col, size_t row) {
    if (!c->gray) return 0.0;
    if (col > c->width - 1) col = c->width - 1;
    if (row > c->height - 1) row = c->height - 1;
    return (double)c->gray[row * c->width + col];
}

static uint8_t to_u8(double v) {
    double r = round(v);
    return (uint8_t)(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
}

int pmo_fuse_consistency(const c3d_fusion_view* views, uint32_t n, const c3d_fusion_params* p, c3d_point_cloud* out) {
    uint8_t** marks = (uint8_t**)calloc(n, sizeof(uint8_t*));
    for (uint32_t v = 0; v < n; ++v) marks[v] = (uint8_t*)calloc((size_t)views[v].width * views[v].height, 1);
    const double tol_depth = p->max_depth_error;
    const double cos_n = cos(p->max_normal_error_deg * (3.14159265358979323846 / 180.0));
    const double r2 = p->max_reproj_error * p->max_reproj_error;
    uint32_t* hit_v = (uint32_t*)malloc(sizeof(uint32_t) * (p->consistency_num_images + 1));
    size_t* hit_p = (size_t*)malloc(sizeof(size_t) * (p->consistency_num_images + 1));
    uint64_t count = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const c3d_fusion_view* cam = &views[v];
        const size_t w = cam->width, h = cam->height;
        uint32_t nov = 0;
        for (uint32_t j = 0; j < cam->num_overlap && nov < p->consistency_num_images; ++j)
            if (cam->overlap[j] < n) nov++;
        for (size_t row = 0; row < h; ++row)
            for (size_t col = 0; col < w; ++col) {
                const size_t pix = row * w + col;
                if (cam->depth[pix] <= 0.0f || (p->mark_used && marks[v][pix])) continue;
                const double d = cam->depth[pix];
                double x[3], nw[3];
                backproject(cam, (double)col, (double)row, d, x);
                normal_world(cam, cam->normal + 3 * pix, nw);
                const double w0 = inv_var(p, cam, pix, d);
                double sw = w0;
                double sx[3] = {x[0] * w0, x[1] * w0, x[2] * w0};
                double sn[3] = {nw[0] * w0, nw[1] * w0, nw[2] * w0};
                const double c0 = gray_at(cam, col, row);
                double sc[3] = {c0 * w0, c0 * w0, c0 * w0};
                uint32_t nh = 0, taken = 0;
                for (uint32_t j = 0; j < cam->num_overlap && taken < nov; ++j) {
                    const uint32_t o = cam->overlap[j];
                    if (o >= n) continue;
                    taken++;
                    const c3d_fusion_view* oc = &views[o];
                    double u, vv, z;
                    if (!project(oc, x, &u, &vv, &z)) continue;
                    const double cu = floor(u + 0.5), rv = floor(vv + 0.5);
                    if (cu < 0.0 || rv < 0.0 || cu >= (double)oc->width || rv >= (double)oc->height) continue;
                    const size_t opix = (size_t)rv * oc->width + (size_t)cu;
                    const double od = oc->depth[opix];
                    if (od <= 0.0 || fabs(z - od) / od > tol_depth) continue;
                    double xo[3];
                    backproject(oc, cu, rv, od, xo);
                    double bu, bv, bz;
                    if (!project(cam, xo, &bu, &bv, &bz)) continue;
                    if ((bu - (double)col) * (bu - (double)col) + (bv - (double)row) * (bv - (double)row) > r2) continue;
                    double no[3];
                    normal_world(oc, oc->normal + 3 * opix, no);
                    if (no[0] * nw[0] + no[1] * nw[1] + no[2] * nw[2] < cos_n) continue;
                    const double co = gray_at(oc, (size_t)cu, (size_t)rv);
                    const double wo = inv_var(p, oc, opix, od);
                    sw += wo;
                    for (int k = 0; k < 3; ++k) {
                        sx[k] += wo * xo[k];
                        sn[k] += wo * no[k];
                        sc[k] += wo * co;
                    }
                    hit_v[nh] = o;
                    hit_p[nh] = opix;
                    nh++;
                }
                const double l = sqrt(sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2]);
                if (nh < p->min_consistent_views || l < 1.1920928955078125e-07) continue;
                if (p->mark_used) {
                    for (uint32_t q = 0; q < nh; ++q) {
                        const c3d_fusion_view* oc = &views[hit_v[q]];
                        if (p->mark_radius == 0) {
                            marks[hit_v[q]][hit_p[q]] = 1;
                            continue;
                        }
                        const long rr = (long)p->mark_radius;
                        const long cx0 = (long)(hit_p[q] % oc->width), cy0 = (long)(hit_p[q] / oc->width);
                        const double d0 = oc->depth[hit_p[q]];
                        for (long yy = cy0 - rr < 0 ? 0 : cy0 - rr; yy <= (cy0 + rr < (long)oc->height - 1 ? cy0 + rr : (long)oc->height - 1); ++yy)
                            for (long xx = cx0 - rr < 0 ? 0 : cx0 - rr; xx <= (cx0 + rr < (long)oc->width - 1 ? cx0 + rr : (long)oc->width - 1);
                                 ++xx) {
                                const size_t qq = (size_t)yy * oc->width + (size_t)xx;
                                const double dq = oc->depth[qq];
                                if (dq > 0.0 && fabs(dq - d0) <= p->max_depth_error * d0) marks[hit_v[q]][qq] = 1;
                            }
                    }
                }
                if (count < out->capacity) {
                    for (int k = 0; k < 3; ++k) {
                        out->positions[3 * count + k] = (float)(sx[k] / sw);
                        out->normals[3 * count + k] = (float)(sn[k] / l);
                        out->colors[3 * count + k] = to_u8(sc[k] / sw);
                    }
                    out->keys[count] = ((uint64_t)cam->image_id << 32) | (((uint64_t)row & 0xffff) << 16) | ((uint64_t)col & 0xffff);
                }
                count++;
            }
    }
    out->count = count;
    for (uint32_t v = 0; v < n; ++v) free(marks[v]);
    free(marks);
    free(hit_v);
    free(hit_p);
    return count > out->capacity ? C3D_INVALID_ARGUMENT : 0;
}
