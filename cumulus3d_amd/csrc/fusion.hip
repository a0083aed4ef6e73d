// fusion.hip — agreement fusion of the final depth maps on the MI355X (include/c3d_fusion.h;
// fuse_consistency, crates/dense/src/fusion.rs:202-490).
//
// Views are fused one after another (later views read the "used" marks that earlier views left):
// one thread per reference pixel evaluates the agreement in double precision, as the reference
// does, writes its point into a per-pixel slot and marks the agreeing pixels of the other views;
// an exclusive scan of the adoption flags (hipcub) then compacts the points in row-major order,
// so the cloud comes out in the reference's order without a sort.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/c3d_fusion.h"
#include "backend.hpp"

namespace c3d {
namespace {

struct FView {
    int w, h;
    long long off;  // first pixel in the concatenated per-pixel arrays
    double k[4], r[9], t[3], baseline;
    unsigned long long key_hi;  // image_id << 32
    int ov_off, ov_n;           // overlap views (already filtered and truncated)
    int has_cost, has_gray;
};

struct FArgs {
    const FView* views;
    const int* overlap;
    const float* depth;
    const float* normal;
    const float* cost;
    const uint8_t* gray;
    uint8_t* marks;
    int v;
    unsigned min_views, mark_radius;
    int mark_used, inverse_variance;
    double sigma_px0, sigma_px_slope, tol_depth, cos_n, r2;
    // per-pixel outputs of view v
    unsigned* flag;
    float* pos;
    float* nrm;
    uint8_t* col;
};

__device__ bool project(const FView& c, const double x[3], double& u, double& v, double& z) {
    const double cz = c.r[6] * x[0] + c.r[7] * x[1] + c.r[8] * x[2] + c.t[2];
    if (cz <= 0.0) return false;
    const double cx = c.r[0] * x[0] + c.r[1] * x[1] + c.r[2] * x[2] + c.t[0];
    const double cy = c.r[3] * x[0] + c.r[4] * x[1] + c.r[5] * x[2] + c.t[1];
    u = c.k[0] * cx / cz + c.k[2];
    v = c.k[1] * cy / cz + c.k[3];
    z = cz;
    return true;
}

__device__ void backproject(const FView& c, double col, double row, double d, double out[3]) {
    const double p0 = (col - c.k[2]) / c.k[0] * d - c.t[0], p1 = (row - c.k[3]) / c.k[1] * d - c.t[1], p2 = d - c.t[2];
    out[0] = c.r[0] * p0 + c.r[3] * p1 + c.r[6] * p2;
    out[1] = c.r[1] * p0 + c.r[4] * p1 + c.r[7] * p2;
    out[2] = c.r[2] * p0 + c.r[5] * p1 + c.r[8] * p2;
}

__device__ void normal_world(const FView& c, const float* n, double out[3]) {
    const double a = n[0], b = n[1], e = n[2];
    out[0] = c.r[0] * a + c.r[3] * b + c.r[6] * e;
    out[1] = c.r[1] * a + c.r[4] * b + c.r[7] * e;
    out[2] = c.r[2] * a + c.r[5] * b + c.r[8] * e;
}

__device__ double inv_var(const FArgs& a, const FView& c, long long pix, double d) {
    if (a.inverse_variance && c.has_cost && c.baseline > 0.0) {
        double cc = a.cost[c.off + pix];
        cc = cc < 0.0 ? 0.0 : (cc > 2.0 ? 2.0 : cc);
        const double spx = a.sigma_px0 + a.sigma_px_slope * cc;
        const double sd = d * d * spx / (c.k[0] * c.baseline);
        const double v = sd * sd;
        return 1.0 / (v > 1e-300 ? v : 1e-300);
    }
    return 1.0;
}

__device__ double gray_at(const FArgs& a, const FView& c, int col, int row) {
    if (!c.has_gray) return 0.0;
    col = col > c.w - 1 ? c.w - 1 : col;
    row = row > c.h - 1 ? c.h - 1 : row;
    return (double)a.gray[c.off + (long long)row * c.w + col];
}

__device__ uint8_t to_u8(double v) {
    const double r = round(v);
    return (uint8_t)(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
}

constexpr int MAX_OV = 32;

__global__ void k_fuse_eval(FArgs a) {
    const FView cam = a.views[a.v];
    const long long pix = (long long)blockIdx.x * 256 + threadIdx.x;
    if (pix >= (long long)cam.w * cam.h) return;
    a.flag[pix] = 0u;
    const float dd = a.depth[cam.off + pix];
    if (dd <= 0.0f || (a.mark_used && a.marks[cam.off + pix])) return;
    const int row = (int)(pix / cam.w), col = (int)(pix - (long long)row * cam.w);
    const double d = dd;
    double x[3], nw[3];
    backproject(cam, (double)col, (double)row, d, x);
    normal_world(cam, a.normal + 3 * (cam.off + pix), nw);
    const double w0 = inv_var(a, cam, pix, d);
    double sw = w0;
    double sx[3] = {x[0] * w0, x[1] * w0, x[2] * w0};
    double sn[3] = {nw[0] * w0, nw[1] * w0, nw[2] * w0};
    const double c0 = gray_at(a, cam, col, row);
    double sc = c0 * w0;
    int hv[MAX_OV];
    long long hp[MAX_OV];
    unsigned nh = 0;
    for (int j = 0; j < cam.ov_n; ++j) {
        const int o = a.overlap[cam.ov_off + j];
        const FView oc = a.views[o];
        double u, vv, z;
        if (!project(oc, x, u, vv, z)) continue;
        const double cu = floor(u + 0.5), rv = floor(vv + 0.5);
        if (cu < 0.0 || rv < 0.0 || cu >= (double)oc.w || rv >= (double)oc.h) continue;
        const long long opix = (long long)rv * oc.w + (long long)cu;
        const double od = a.depth[oc.off + opix];
        if (od <= 0.0 || fabs(z - od) / od > a.tol_depth) continue;
        double xo[3];
        backproject(oc, cu, rv, od, xo);
        double bu, bv, bz;
        if (!project(cam, xo, bu, bv, bz)) continue;
        if ((bu - (double)col) * (bu - (double)col) + (bv - (double)row) * (bv - (double)row) > a.r2) continue;
        double no[3];
        normal_world(oc, a.normal + 3 * (oc.off + opix), no);
        if (no[0] * nw[0] + no[1] * nw[1] + no[2] * nw[2] < a.cos_n) continue;
        const double co = gray_at(a, oc, (int)cu, (int)rv);
        const double wo = inv_var(a, oc, opix, od);
        sw += wo;
        sx[0] += wo * xo[0];
        sx[1] += wo * xo[1];
        sx[2] += wo * xo[2];
        sn[0] += wo * no[0];
        sn[1] += wo * no[1];
        sn[2] += wo * no[2];
        sc += wo * co;
        hv[nh] = o;
        hp[nh] = opix;
        nh++;
    }
    const double l = sqrt(sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2]);
    if (nh < a.min_views || l < 1.1920928955078125e-07) return;
    if (a.mark_used) {
        for (unsigned q = 0; q < nh; ++q) {
            const FView oc = a.views[hv[q]];
            if (a.mark_radius == 0) {
                a.marks[oc.off + hp[q]] = 1;
                continue;
            }
            const int rr = (int)a.mark_radius;
            const int cx0 = (int)(hp[q] % oc.w), cy0 = (int)(hp[q] / oc.w);
            const double d0 = a.depth[oc.off + hp[q]];
            for (int yy = max(cy0 - rr, 0); yy <= min(cy0 + rr, oc.h - 1); ++yy)
                for (int xx = max(cx0 - rr, 0); xx <= min(cx0 + rr, oc.w - 1); ++xx) {
                    const long long qq = (long long)yy * oc.w + xx;
                    const double dq = a.depth[oc.off + qq];
                    if (dq > 0.0 && fabs(dq - d0) <= a.tol_depth * d0) a.marks[oc.off + qq] = 1;
                }
        }
    }
    a.flag[pix] = 1u;
    const uint8_t cc = to_u8(sc / sw);
    for (int k = 0; k < 3; ++k) {
        a.pos[3 * pix + k] = (float)(sx[k] / sw);
        a.nrm[3 * pix + k] = (float)(sn[k] / l);
        a.col[3 * pix + k] = cc;
    }
}

__global__ void k_fuse_compact(const unsigned* flag, const unsigned* idx, long long npix, int w, unsigned long long key_hi,
                               unsigned long long base, unsigned long long cap, const float* pos, const float* nrm, const uint8_t* col,
                               float* opos, float* onrm, uint8_t* ocol, unsigned long long* okey) {
    const long long pix = (long long)blockIdx.x * 256 + threadIdx.x;
    if (pix >= npix || !flag[pix]) return;
    const unsigned long long q = base + idx[pix];
    if (q >= cap) return;
    for (int k = 0; k < 3; ++k) {
        opos[3 * q + k] = pos[3 * pix + k];
        onrm[3 * q + k] = nrm[3 * pix + k];
        ocol[3 * q + k] = col[3 * pix + k];
    }
    const unsigned long long row = (unsigned long long)(pix / w), c = (unsigned long long)(pix % w);
    okey[q] = key_hi | ((row & 0xffffull) << 16) | (c & 0xffffull);
}

struct Bufs {
    std::vector<void*> p;
    ~Bufs() {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T>
    hipError_t get(T** out, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) {
            p.push_back(q);
            *out = (T*)q;
        }
        return e;
    }
};

}  // namespace
}  // namespace c3d

using namespace c3d;

#define F_TRY(expr)                                                                                                     \
    do {                                                                                                                \
        hipError_t e_ = (expr);                                                                                         \
        if (e_ != hipSuccess) return set_error(C3D_DEVICE, std::string("HIP: ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int c3d_fuse_consistency(c3d_pm_backend* be, const c3d_fusion_view* views, uint32_t n, const c3d_fusion_params* p,
                                    c3d_point_cloud* out) {
    if (!be || !views || !p || !out || n == 0) return set_error(C3D_INVALID_ARGUMENT, "null argument");
    if (out->capacity && (!out->positions || !out->normals || !out->colors || !out->keys))
        return set_error(C3D_INVALID_ARGUMENT, "null output arrays");
    F_TRY(hipSetDevice(be->device));
    hipStream_t st = be->stream;
    std::vector<FView> fv(n);
    std::vector<int> ov;
    long long tot = 0;
    long long maxpix = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const c3d_fusion_view& c = views[v];
        if (!c.depth || !c.normal || c.width < 1 || c.height < 1 || c.width > 65536 || c.height > 65536 ||
            (long long)c.width * c.height >= (1LL << 31))
            return set_error(C3D_INVALID_ARGUMENT, "fusion view without depth/normal or of bad size");
        FView& f = fv[v];
        f.w = (int)c.width;
        f.h = (int)c.height;
        f.off = tot;
        tot += (long long)c.width * c.height;
        maxpix = std::max(maxpix, (long long)c.width * c.height);
        std::memcpy(f.k, c.k, sizeof f.k);
        std::memcpy(f.r, c.r, sizeof f.r);
        std::memcpy(f.t, c.t, sizeof f.t);
        f.baseline = c.baseline;
        f.key_hi = (unsigned long long)c.image_id << 32;
        f.has_cost = c.cost != nullptr;
        f.has_gray = c.gray != nullptr;
        f.ov_off = (int)ov.size();
        for (uint32_t j = 0; j < c.num_overlap && (uint32_t)(ov.size() - f.ov_off) < p->consistency_num_images; ++j)
            if (c.overlap[j] < n) ov.push_back((int)c.overlap[j]);
        f.ov_n = (int)ov.size() - f.ov_off;
        if (f.ov_n > MAX_OV) return set_error(C3D_INVALID_ARGUMENT, "more than 32 overlap views");
    }
    Bufs b;
    FView* d_views;
    int* d_ov;
    float *d_depth, *d_normal, *d_cost, *d_pos, *d_nrm;
    uint8_t *d_gray, *d_marks, *d_col;
    unsigned *d_flag, *d_idx;
    F_TRY(b.get(&d_views, n));
    F_TRY(b.get(&d_ov, ov.size()));
    F_TRY(b.get(&d_depth, tot));
    F_TRY(b.get(&d_normal, 3 * tot));
    F_TRY(b.get(&d_cost, tot));
    F_TRY(b.get(&d_gray, tot));
    F_TRY(b.get(&d_marks, tot));
    F_TRY(b.get(&d_flag, maxpix));
    F_TRY(b.get(&d_idx, maxpix));
    F_TRY(b.get(&d_pos, 3 * maxpix));
    F_TRY(b.get(&d_nrm, 3 * maxpix));
    F_TRY(b.get(&d_col, 3 * maxpix));
    const unsigned long long cap = out->capacity;
    float *o_pos, *o_nrm;
    uint8_t* o_col;
    unsigned long long* o_key;
    F_TRY(b.get(&o_pos, 3 * cap));
    F_TRY(b.get(&o_nrm, 3 * cap));
    F_TRY(b.get(&o_col, 3 * cap));
    F_TRY(b.get(&o_key, cap));
    F_TRY(hipMemcpyAsync(d_views, fv.data(), n * sizeof(FView), hipMemcpyHostToDevice, st));
    if (!ov.empty()) F_TRY(hipMemcpyAsync(d_ov, ov.data(), ov.size() * sizeof(int), hipMemcpyHostToDevice, st));
    for (uint32_t v = 0; v < n; ++v) {
        const c3d_fusion_view& c = views[v];
        const size_t np = (size_t)c.width * c.height;
        F_TRY(hipMemcpyAsync(d_depth + fv[v].off, c.depth, np * sizeof(float), hipMemcpyHostToDevice, st));
        F_TRY(hipMemcpyAsync(d_normal + 3 * fv[v].off, c.normal, 3 * np * sizeof(float), hipMemcpyHostToDevice, st));
        if (c.cost) F_TRY(hipMemcpyAsync(d_cost + fv[v].off, c.cost, np * sizeof(float), hipMemcpyHostToDevice, st));
        if (c.gray) F_TRY(hipMemcpyAsync(d_gray + fv[v].off, c.gray, np, hipMemcpyHostToDevice, st));
    }
    F_TRY(hipMemsetAsync(d_marks, 0, tot, st));
    size_t tmp_bytes = 0;
    F_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_flag, d_idx, (int)maxpix, st));
    void* d_tmp;
    F_TRY(b.get((uint8_t**)&d_tmp, tmp_bytes));
    FArgs a{};
    a.views = d_views;
    a.overlap = d_ov;
    a.depth = d_depth;
    a.normal = d_normal;
    a.cost = d_cost;
    a.gray = d_gray;
    a.marks = d_marks;
    a.min_views = p->min_consistent_views;
    a.mark_radius = p->mark_radius;
    a.mark_used = p->mark_used;
    a.inverse_variance = p->inverse_variance;
    a.sigma_px0 = p->sigma_px0;
    a.sigma_px_slope = p->sigma_px_slope;
    a.tol_depth = p->max_depth_error;
    a.cos_n = std::cos(p->max_normal_error_deg * (3.14159265358979323846 / 180.0));
    a.r2 = p->max_reproj_error * p->max_reproj_error;
    a.flag = d_flag;
    a.pos = d_pos;
    a.nrm = d_nrm;
    a.col = d_col;
    unsigned long long count = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const long long np = (long long)fv[v].w * fv[v].h;
        a.v = (int)v;
        hipLaunchKernelGGL(k_fuse_eval, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, a);
        F_TRY(hipGetLastError());
        F_TRY(hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_flag, d_idx, (int)np, st));
        hipLaunchKernelGGL(k_fuse_compact, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, d_flag, d_idx, np, fv[v].w,
                           fv[v].key_hi, count, cap, d_pos, d_nrm, d_col, o_pos, o_nrm, o_col, o_key);
        F_TRY(hipGetLastError());
        unsigned last_idx = 0, last_flag = 0;
        F_TRY(hipMemcpyAsync(&last_idx, d_idx + np - 1, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        F_TRY(hipMemcpyAsync(&last_flag, d_flag + np - 1, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        F_TRY(hipStreamSynchronize(st));
        count += (unsigned long long)last_idx + last_flag;
    }
    const unsigned long long m = std::min(count, cap);
    if (m) {
        F_TRY(hipMemcpyAsync(out->positions, o_pos, 3 * m * sizeof(float), hipMemcpyDeviceToHost, st));
        F_TRY(hipMemcpyAsync(out->normals, o_nrm, 3 * m * sizeof(float), hipMemcpyDeviceToHost, st));
        F_TRY(hipMemcpyAsync(out->colors, o_col, 3 * m, hipMemcpyDeviceToHost, st));
        F_TRY(hipMemcpyAsync(out->keys, o_key, m * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    }
    F_TRY(hipStreamSynchronize(st));
    out->count = count;
    if (count > cap) return set_error(C3D_INVALID_ARGUMENT, "point cloud capacity " + std::to_string(cap) + " < " + std::to_string(count));
    return C3D_OK;
}
