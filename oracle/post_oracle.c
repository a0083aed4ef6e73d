/* post_oracle.c — CPU restatement of the read-out post-processing (remove_speckles,
 * crates/dense/src/postproc.rs:77-133; fill_holes, postproc.rs:135-204) — TEST INFRASTRUCTURE
 * ONLY: the checker of cumulus3d_amd/csrc/postproc.hip.
 *
 * Same algorithm and order as the reference: components flooded from each unlabelled pixel in
 * raster order with an explicit stack (neighbours left, right, up, down), their costs sorted in
 * IEEE total order, element len/2 compared with speckle_keep_cost; the fill walks the 8 directions
 * in the reference's order and applies all fills after the scan. Depth rule and component join in
 * f32, fill geometry in double (plane_transfer, pm_oracle.c).
 *
 * Pinning: the reference's own test postproc.rs:208-231 (4 speckle pixels erased, 13 filled, the
 * hole at depth 10) in tests/test_post_oracle.py, components against scipy's connected components
 * of the join graph, medians against numpy, fills of a tilted plane against its closed-form depth.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pm_oracle.h"

/* f32::total_cmp as an unsigned key */
static uint32_t total_key(float c) {
    uint32_t u;
    memcpy(&u, &c, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

static int cmp_total(const void* a, const void* b) {
    const uint32_t x = total_key(*(const float*)a), y = total_key(*(const float*)b);
    return x < y ? -1 : (x > y ? 1 : 0);
}

int pmo_remove_speckles(float* depth, float* normal, const float* cost, uint32_t w, uint32_t h, const c3d_post_params* p,
                        uint64_t* removed_out) {
    const size_t np = (size_t)w * h;
    const size_t min_area = (size_t)ceil((double)np * p->speckle_min_area_frac);
    if (removed_out) *removed_out = 0;
    if (min_area <= 1) return 0;
    uint32_t* label = (uint32_t*)malloc(np * sizeof(uint32_t));
    size_t* stack = (size_t*)malloc(np * sizeof(size_t));
    size_t* comp = (size_t*)malloc(np * sizeof(size_t));
    float* costs = (float*)malloc(np * sizeof(float));
    for (size_t i = 0; i < np; ++i) label[i] = 0xFFFFFFFFu;
    uint64_t removed = 0;
    for (size_t start = 0; start < np; ++start) {
        if (depth[start] <= 0.0f || label[start] != 0xFFFFFFFFu) continue;
        size_t ns = 0, nc = 0;
        stack[ns++] = start;
        label[start] = (uint32_t)start;
        while (ns) {
            const size_t i = stack[--ns];
            comp[nc++] = i;
            const size_t x = i % w, y = i / w;
            const float di = depth[i];
            size_t nb[4] = {(size_t)-1, (size_t)-1, (size_t)-1, (size_t)-1};
            if (x > 0) nb[0] = i - 1;
            if (x + 1 < w) nb[1] = i + 1;
            if (y > 0) nb[2] = i - w;
            if (y + 1 < h) nb[3] = i + w;
            for (int k = 0; k < 4; ++k) {
                const size_t j = nb[k];
                if (j == (size_t)-1 || label[j] != 0xFFFFFFFFu || depth[j] <= 0.0f) continue;
                const float dj = depth[j];
                if (fabsf(dj - di) <= p->speckle_rel_depth * fminf(di, dj)) {
                    label[j] = (uint32_t)start;
                    stack[ns++] = j;
                }
            }
        }
        if (nc < min_area) {
            for (size_t q = 0; q < nc; ++q) costs[q] = cost[comp[q]];
            qsort(costs, nc, sizeof(float), cmp_total);
            if (costs[nc / 2] > p->speckle_keep_cost) {
                for (size_t q = 0; q < nc; ++q) {
                    depth[comp[q]] = 0.0f;
                    normal[3 * comp[q] + 0] = normal[3 * comp[q] + 1] = normal[3 * comp[q] + 2] = 0.0f;
                }
                removed += nc;
            }
        }
    }
    free(label);
    free(stack);
    free(comp);
    free(costs);
    if (removed_out) *removed_out = removed;
    return 0;
}

typedef struct {
    size_t i, q;
    float d, n[3], c;
} fill_t;

int pmo_fill_holes(float* depth, float* normal, float* cost, const float* raw_cost, const uint8_t* gray, const double k[4], uint32_t w,
                   uint32_t h, const c3d_post_params* p, uint64_t* filled_out) {
    const size_t np = (size_t)w * h;
    long long maxd = (long long)round(sqrt((double)((unsigned long long)w * w + (unsigned long long)h * h)) * p->fill_max_frac);
    if (maxd < 1) maxd = 1;
    uint8_t* valid = (uint8_t*)malloc(np);
    for (size_t i = 0; i < np; ++i) valid[i] = depth[i] > 0.0f;
    const float sf = p->fill_edge_step * 255.0f;
    const int step = sf != sf ? 0 : (sf >= 2147483647.0f ? 2147483647 : (sf <= -2147483648.0f ? (-2147483647 - 1) : (int)sf));
    static const int DX[8] = {1, -1, 0, 0, 1, -1, 1, -1};
    static const int DY[8] = {0, 0, 1, -1, 1, -1, -1, 1};
    fill_t* fills = (fill_t*)malloc(np * sizeof(fill_t) + 1);
    size_t nf = 0;
    for (long long y = 0; y < (long long)h; ++y)
        for (long long x = 0; x < (long long)w; ++x) {
            const size_t i = (size_t)y * w + (size_t)x;
            if (valid[i] || raw_cost[i] > p->fill_max_cost) continue;
            const int ip = gray[i];
            int found = 0, bd = 0;
            long long bs = 0;
            size_t bq = 0;
            for (int dir = 0; dir < 8; ++dir) {
                int prev = ip;
                for (long long s = 1; s <= maxd; ++s) {
                    const long long qx = x + DX[dir] * s, qy = y + DY[dir] * s;
                    if (qx < 0 || qy < 0 || qx >= (long long)w || qy >= (long long)h) break;
                    const size_t q = (size_t)qy * w + (size_t)qx;
                    const int iq = gray[q];
                    if (abs(iq - prev) > step) break;
                    prev = iq;
                    if (valid[q]) {
                        const int dd = abs(iq - ip);
                        if (!found || dd < bd || (dd == bd && s < bs)) {
                            found = 1;
                            bd = dd;
                            bs = s;
                            bq = q;
                        }
                        break;
                    }
                }
            }
            if (!found) continue;
            const double nd[3] = {normal[3 * bq], normal[3 * bq + 1], normal[3 * bq + 2]};
            const double qx = (double)(bq % w), qy = (double)(bq / w);
            const double rq[3] = {(qx - k[2]) / k[0], (qy - k[3]) / k[1], 1.0};
            const double rp[3] = {((double)x - k[2]) / k[0], ((double)y - k[3]) / k[1], 1.0};
            double d;
            if (pmo_plane_transfer((double)depth[bq], nd, rq, rp, &d)) {
                fill_t* f = &fills[nf++];
                f->i = i;
                f->q = bq;
                f->d = (float)d;
                f->n[0] = normal[3 * bq];
                f->n[1] = normal[3 * bq + 1];
                f->n[2] = normal[3 * bq + 2];
                f->c = cost[bq];
            }
        }
    for (size_t j = 0; j < nf; ++j) {
        depth[fills[j].i] = fills[j].d;
        normal[3 * fills[j].i + 0] = fills[j].n[0];
        normal[3 * fills[j].i + 1] = fills[j].n[1];
        normal[3 * fills[j].i + 2] = fills[j].n[2];
        cost[fills[j].i] = fills[j].c;
    }
    free(valid);
    free(fills);
    if (filled_out) *filled_out = nf;
    return 0;
}
