// neighbors.hip — scene preparation on the MI355X (include/c3d_scene.h): pair statistics of the
// sparse tracks (PairStats::new, crates/dense/src/neighbors.rs:47-80), source selection
// (select / select_diverse, :83-151, host) and per-view depth ranges (depth_ranges, :162-189).
//
// Every track pair becomes one 64-bit key (pair index above, the f32 angle's bits below: angles
// are in [0, pi/2], so their bits sort as their values) and every observation in front of its
// camera one (view, depth) record; a radix sort groups the pairs and orders each pair's angles,
// after which counts, 75th percentiles and 1 % / 99 % depths are positions in the sorted runs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/c3d_scene.h"
#include "backend.hpp"

namespace c3d {
namespace {

struct Cam {
    double r[9], t[3], c[3];
};

// Camera centre -(R^T t) (scene.rs:82-89), in the reference's operation order
void centre(const double* r, const double* t, double c[3]) {
    c[0] = -(r[0] * t[0] + r[3] * t[1] + r[6] * t[2]);
    c[1] = -(r[1] * t[0] + r[4] * t[1] + r[7] * t[2]);
    c[2] = -(r[2] * t[0] + r[5] * t[1] + r[8] * t[2]);
}

__global__ __launch_bounds__(256) void k_pairs_per_point(const unsigned long long* off, const unsigned* tv, unsigned long long np,
                                                         unsigned long long* npairs) {
    const unsigned long long p = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const unsigned long long a0 = off[p], a1 = off[p + 1];
    unsigned long long c = 0;
    for (unsigned long long i = a0; i < a1; ++i)
        for (unsigned long long j = a0; j < i; ++j) c += tv[i] != tv[j];
    npairs[p] = c;
}

// the angle of neighbors.rs:57-70 and the pair key
__global__ __launch_bounds__(256) void k_pair_keys(const unsigned long long* off, const unsigned* tv, unsigned long long np,
                                                   const double* xyz, const Cam* cams, unsigned n, const unsigned long long* base,
                                                   unsigned long long* keys) {
    const unsigned long long p = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const unsigned long long a0 = off[p], a1 = off[p + 1];
    const double x0 = xyz[3 * p], x1 = xyz[3 * p + 1], x2 = xyz[3 * p + 2];
    unsigned long long q = base[p];
    for (unsigned long long i = a0; i < a1; ++i)
        for (unsigned long long j = a0; j < i; ++j) {
            const unsigned a = tv[i], b = tv[j];
            if (a == b) continue;
            const double* ca = cams[a].c;
            const double* cb = cams[b].c;
            const double u0 = ca[0] - x0, u1 = ca[1] - x1, u2 = ca[2] - x2;
            const double v0 = cb[0] - x0, v1 = cb[1] - x1, v2 = cb[2] - x2;
            const double nu = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
            const double nv = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
            double ang = 0.0;
            if (!(nu == 0.0 || nv == 0.0)) {
                double c = (u0 * v0 + u1 * v1 + u2 * v2) / (nu * nv);
                c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);  // f64::clamp keeps NaN
                const double t = acos(c);
                const double s = 3.141592653589793 - t;
                ang = t < s ? t : s;                          // f64::min (NaN only with NaN input)
            }
            const unsigned lo = a < b ? a : b, hi = a < b ? b : a;
            const unsigned long long pair = (unsigned long long)lo * n + hi;
            keys[q++] = (pair << 32) | (unsigned long long)__float_as_uint((float)ang);
        }
}

__global__ __launch_bounds__(256) void k_widen(const unsigned* v, unsigned long long n, unsigned long long* out) {
    const unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = v[i];
}

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* p, unsigned long long n, unsigned v) {
    const unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// first and one-past-last position of every run of equal key >> shift
__global__ __launch_bounds__(256) void k_runs(const unsigned long long* keys, unsigned long long n, int shift, unsigned* start,
                                              unsigned* end) {
    const unsigned long long p = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const unsigned long long g = keys[p] >> shift;
    if (p == 0 || (keys[p - 1] >> shift) != g) start[g] = (unsigned)p;
    if (p + 1 == n || (keys[p + 1] >> shift) != g) end[g] = (unsigned)(p + 1);
}

// count and quantile_sorted(angles, 0.75) (math.rs:126-135) of pair (a, b), mirrored to (b, a)
__global__ __launch_bounds__(256) void k_pair_quantile(const unsigned long long* keys, const unsigned* start, const unsigned* end,
                                                       unsigned n, unsigned* count, double* q75) {
    const unsigned long long g = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= (unsigned long long)n * n) return;
    const unsigned a = (unsigned)(g / n), b = (unsigned)(g % n);
    if (a >= b) return;
    const unsigned s = start[g], c = end[g] - s;
    double q = 0.0;
    if (c > 0) {
        const double pos = 0.75 * (double)(c - 1);
        const unsigned lo = (unsigned)floor(pos);
        const unsigned hi = lo + 1 < c - 1 ? lo + 1 : c - 1;
        const double f = pos - (double)lo;
        const double vl = (double)__uint_as_float((unsigned)(keys[s + lo] & 0xFFFFFFFFull));
        const double vh = (double)__uint_as_float((unsigned)(keys[s + hi] & 0xFFFFFFFFull));
        q = vl + (vh - vl) * f;
    }
    count[g] = c;
    count[(unsigned long long)b * n + a] = c;
    q75[g] = q;
    q75[(unsigned long long)b * n + a] = q;
}

__global__ __launch_bounds__(256) void k_depth_per_point(const unsigned long long* off, const unsigned* tv, unsigned long long np,
                                                         const double* xyz, const Cam* cams, unsigned long long* nobs) {
    const unsigned long long p = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const double x0 = xyz[3 * p], x1 = xyz[3 * p + 1], x2 = xyz[3 * p + 2];
    unsigned long long c = 0;
    for (unsigned long long i = off[p]; i < off[p + 1]; ++i) {
        const double* r = cams[tv[i]].r;
        const double d = r[6] * x0 + r[7] * x1 + r[8] * x2 + cams[tv[i]].t[2];
        c += d > 0.0;
    }
    nobs[p] = c;
}

__global__ __launch_bounds__(256) void k_depth_write(const unsigned long long* off, const unsigned* tv, unsigned long long np,
                                                     const double* xyz, const Cam* cams, const unsigned long long* base,
                                                     unsigned long long* dbits, unsigned* view) {
    const unsigned long long p = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const double x0 = xyz[3 * p], x1 = xyz[3 * p + 1], x2 = xyz[3 * p + 2];
    unsigned long long q = base[p];
    for (unsigned long long i = off[p]; i < off[p + 1]; ++i) {
        const double* r = cams[tv[i]].r;
        const double d = r[6] * x0 + r[7] * x1 + r[8] * x2 + cams[tv[i]].t[2];
        if (d > 0.0) {  // positive doubles: the bits sort as the values
            dbits[q] = (unsigned long long)__double_as_longlong(d);
            view[q] = tv[i];
            ++q;
        }
    }
}

// depth_range_of (neighbors.rs:178-189) on the view's sorted run
__global__ void k_ranges(const unsigned long long* dbits, const unsigned* start, const unsigned* end, unsigned n, double* dmin,
                         double* dmax, uint8_t* has) {
    const unsigned v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const unsigned s = start[v], c = end[v] - s;
    if (c == 0) {
        dmin[v] = dmax[v] = 0.0;
        has[v] = 0;
        return;
    }
    unsigned lo = (unsigned)floor((double)c * 0.01), hi = (unsigned)floor((double)c * 0.99);
    lo = lo < c - 1 ? lo : c - 1;
    hi = hi < c - 1 ? hi : c - 1;
    dmin[v] = 0.75 * __longlong_as_double((long long)dbits[s + lo]);
    dmax[v] = 1.25 * __longlong_as_double((long long)dbits[s + hi]);
    has[v] = 1;
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

#define N_TRY(expr)                                                                                                     \
    do {                                                                                                                \
        hipError_t e_ = (expr);                                                                                         \
        if (e_ != hipSuccess) return set_error(C3D_DEVICE, std::string("HIP: ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

int check_scene(const c3d_sparse_scene* s) {
    if (!s || !s->rotation || !s->translation || (s->num_points && (!s->xyz || !s->track_offset)))
        return set_error(C3D_INVALID_ARGUMENT, "null scene array");
    if (s->num_views == 0 || s->num_views > 65535) return set_error(C3D_INVALID_ARGUMENT, "1..65535 views");
    if (s->num_points && s->track_offset[0] != 0) return set_error(C3D_INVALID_ARGUMENT, "track_offset[0] must be 0");
    for (uint64_t p = 0; p < s->num_points; ++p)
        if (s->track_offset[p + 1] < s->track_offset[p]) return set_error(C3D_INVALID_ARGUMENT, "decreasing track_offset");
    const uint64_t nobs = s->num_points ? s->track_offset[s->num_points] : 0;
    if (nobs && !s->track_views) return set_error(C3D_INVALID_ARGUMENT, "null track_views");
    for (uint64_t i = 0; i < nobs; ++i)
        if (s->track_views[i] >= s->num_views) return set_error(C3D_INVALID_ARGUMENT, "track view index out of range");
    return C3D_OK;
}

// uploads cameras, points and tracks
struct SceneDev {
    Cam* cams = nullptr;
    double* xyz = nullptr;
    unsigned long long* off = nullptr;
    unsigned* tv = nullptr;
    unsigned long long nobs = 0;
};

int upload_scene(const c3d_sparse_scene* s, Bufs& b, SceneDev& d, hipStream_t st, std::vector<Cam>& hc) {
    const unsigned n = s->num_views;
    hc.resize(n);
    for (unsigned v = 0; v < n; ++v) {
        std::memcpy(hc[v].r, s->rotation + 9 * (size_t)v, sizeof hc[v].r);
        std::memcpy(hc[v].t, s->translation + 3 * (size_t)v, sizeof hc[v].t);
        centre(hc[v].r, hc[v].t, hc[v].c);
    }
    const unsigned long long np = s->num_points;
    d.nobs = np ? s->track_offset[np] : 0;
    N_TRY(b.get(&d.cams, n));
    N_TRY(b.get(&d.xyz, 3 * np));
    N_TRY(b.get(&d.off, np + 1));
    N_TRY(b.get(&d.tv, d.nobs));
    N_TRY(hipMemcpyAsync(d.cams, hc.data(), n * sizeof(Cam), hipMemcpyHostToDevice, st));
    if (np) {
        N_TRY(hipMemcpyAsync(d.xyz, s->xyz, 3 * np * sizeof(double), hipMemcpyHostToDevice, st));
        N_TRY(hipMemcpyAsync(d.off, s->track_offset, (np + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, st));
    }
    if (d.nobs) N_TRY(hipMemcpyAsync(d.tv, s->track_views, d.nobs * sizeof(unsigned), hipMemcpyHostToDevice, st));
    return C3D_OK;
}

// exclusive scan of per-point counts; returns the total
int scan_counts(Bufs& b, const unsigned long long* cnt, unsigned long long* base, unsigned long long np, hipStream_t st,
                unsigned long long& total) {
    total = 0;
    if (!np) return C3D_OK;
    if (np > 0x7FFFFFFFull) return set_error(C3D_INVALID_ARGUMENT, "more than 2^31 - 1 points");
    size_t bytes = 0;
    N_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, base, (int)np, st));
    uint8_t* tmp;
    N_TRY(b.get(&tmp, bytes));
    N_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, base, (int)np, st));
    unsigned long long last_base = 0, last_cnt = 0;
    N_TRY(hipMemcpyAsync(&last_base, base + np - 1, sizeof last_base, hipMemcpyDeviceToHost, st));
    N_TRY(hipMemcpyAsync(&last_cnt, cnt + np - 1, sizeof last_cnt, hipMemcpyDeviceToHost, st));
    N_TRY(hipStreamSynchronize(st));
    total = last_base + last_cnt;
    return C3D_OK;
}

unsigned grid_of(unsigned long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace c3d

using namespace c3d;

extern "C" int c3d_pair_stats(c3d_pm_backend* be, const c3d_sparse_scene* s, uint32_t* count, double* q75) {
    if (!be || !count || !q75) return set_error(C3D_INVALID_ARGUMENT, "null argument");
    int rc = check_scene(s);
    if (rc) return rc;
    if (s->num_views > 8192) return set_error(C3D_INVALID_ARGUMENT, "pair statistics of at most 8192 views");
    N_TRY(hipSetDevice(be->device));
    hipStream_t st = be->stream;
    const unsigned n = s->num_views;
    const unsigned long long nn = (unsigned long long)n * n, np = s->num_points;
    Bufs b;
    SceneDev d;
    std::vector<Cam> hc;
    rc = upload_scene(s, b, d, st, hc);
    if (rc) return rc;
    unsigned long long *cnt, *base;
    N_TRY(b.get(&cnt, np));
    N_TRY(b.get(&base, np));
    if (np) hipLaunchKernelGGL(k_pairs_per_point, dim3(grid_of(np)), dim3(256), 0, st, d.off, d.tv, np, cnt);
    N_TRY(hipGetLastError());
    unsigned long long total = 0;
    rc = scan_counts(b, cnt, base, np, st, total);
    if (rc) return rc;
    if (total > 0x7FFFFFFFull) return set_error(C3D_INVALID_ARGUMENT, "more than 2^31 - 1 track pairs");
    unsigned *d_count, *start, *end;
    double* d_q;
    N_TRY(b.get(&d_count, nn));
    N_TRY(b.get(&d_q, nn));
    N_TRY(b.get(&start, nn));
    N_TRY(b.get(&end, nn));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(nn)), dim3(256), 0, st, start, nn, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(nn)), dim3(256), 0, st, end, nn, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(nn)), dim3(256), 0, st, d_count, nn, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(2 * nn)), dim3(256), 0, st, (unsigned*)d_q, 2 * nn, 0u);
    N_TRY(hipGetLastError());
    if (total) {
        unsigned long long *keys, *sorted;
        N_TRY(b.get(&keys, total));
        N_TRY(b.get(&sorted, total));
        hipLaunchKernelGGL(k_pair_keys, dim3(grid_of(np)), dim3(256), 0, st, d.off, d.tv, np, d.xyz, d.cams, n, base, keys);
        N_TRY(hipGetLastError());
        size_t bytes = 0;
        N_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys, sorted, (int)total, 0, 64, st));
        uint8_t* tmp;
        N_TRY(b.get(&tmp, bytes));
        N_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, bytes, keys, sorted, (int)total, 0, 64, st));
        hipLaunchKernelGGL(k_runs, dim3(grid_of(total)), dim3(256), 0, st, sorted, total, 32, start, end);
        hipLaunchKernelGGL(k_pair_quantile, dim3(grid_of(nn)), dim3(256), 0, st, sorted, start, end, n, d_count, d_q);
        N_TRY(hipGetLastError());
    }
    N_TRY(hipMemcpyAsync(count, d_count, nn * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    N_TRY(hipMemcpyAsync(q75, d_q, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    N_TRY(hipStreamSynchronize(st));
    return C3D_OK;
}

extern "C" int c3d_depth_ranges(c3d_pm_backend* be, const c3d_sparse_scene* s, double* depth_min, double* depth_max,
                                uint8_t* has_range) {
    if (!be || !depth_min || !depth_max || !has_range) return set_error(C3D_INVALID_ARGUMENT, "null argument");
    int rc = check_scene(s);
    if (rc) return rc;
    N_TRY(hipSetDevice(be->device));
    hipStream_t st = be->stream;
    const unsigned n = s->num_views;
    const unsigned long long np = s->num_points;
    Bufs b;
    SceneDev d;
    std::vector<Cam> hc;
    rc = upload_scene(s, b, d, st, hc);
    if (rc) return rc;
    unsigned long long *cnt, *base;
    N_TRY(b.get(&cnt, np));
    N_TRY(b.get(&base, np));
    if (np) hipLaunchKernelGGL(k_depth_per_point, dim3(grid_of(np)), dim3(256), 0, st, d.off, d.tv, np, d.xyz, d.cams, cnt);
    N_TRY(hipGetLastError());
    unsigned long long total = 0;
    rc = scan_counts(b, cnt, base, np, st, total);
    if (rc) return rc;
    if (total > 0x7FFFFFFFull) return set_error(C3D_INVALID_ARGUMENT, "more than 2^31 - 1 observations");
    unsigned *start, *end;
    double *dmin, *dmax;
    uint8_t* has;
    N_TRY(b.get(&start, n));
    N_TRY(b.get(&end, n));
    N_TRY(b.get(&dmin, n));
    N_TRY(b.get(&dmax, n));
    N_TRY(b.get(&has, n));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(n)), dim3(256), 0, st, start, (unsigned long long)n, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(n)), dim3(256), 0, st, end, (unsigned long long)n, 0u);
    N_TRY(hipGetLastError());
    unsigned long long* dsorted = nullptr;
    if (total) {
        unsigned long long *dbits, *dbits2, *vkey;
        unsigned *view, *view2;
        N_TRY(b.get(&dbits, total));
        N_TRY(b.get(&dbits2, total));
        N_TRY(b.get(&view, total));
        N_TRY(b.get(&view2, total));
        N_TRY(b.get(&vkey, total));
        hipLaunchKernelGGL(k_depth_write, dim3(grid_of(np)), dim3(256), 0, st, d.off, d.tv, np, d.xyz, d.cams, base, dbits, view);
        N_TRY(hipGetLastError());
        // by depth, then stably by view: each view's depths in increasing order
        size_t b1 = 0, b2 = 0;
        N_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, dbits, dbits2, view, view2, (int)total, 0, 64, st));
        N_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b2, view2, view, dbits2, dbits, (int)total, 0, 32, st));
        uint8_t* tmp;
        N_TRY(b.get(&tmp, std::max(b1, b2)));
        N_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, b1, dbits, dbits2, view, view2, (int)total, 0, 64, st));
        N_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, b2, view2, view, dbits2, dbits, (int)total, 0, 32, st));
        // runs of equal view
        hipLaunchKernelGGL(k_widen, dim3(grid_of(total)), dim3(256), 0, st, view, total, vkey);
        hipLaunchKernelGGL(k_runs, dim3(grid_of(total)), dim3(256), 0, st, vkey, total, 0, start, end);
        N_TRY(hipGetLastError());
        dsorted = dbits;
    }
    hipLaunchKernelGGL(k_ranges, dim3((n + 255) / 256), dim3(256), 0, st, dsorted, start, end, n, dmin, dmax, has);
    N_TRY(hipGetLastError());
    N_TRY(hipMemcpyAsync(depth_min, dmin, n * sizeof(double), hipMemcpyDeviceToHost, st));
    N_TRY(hipMemcpyAsync(depth_max, dmax, n * sizeof(double), hipMemcpyDeviceToHost, st));
    N_TRY(hipMemcpyAsync(has_range, has, n, hipMemcpyDeviceToHost, st));
    N_TRY(hipStreamSynchronize(st));
    return C3D_OK;
}

extern "C" int c3d_select_views(uint32_t n, const uint32_t* count, const double* q75, uint32_t r, uint32_t max_views,
                                double min_angle_rad, uint32_t* out, uint32_t* num_out) {
    if (!count || !q75 || !num_out || (max_views && !out)) return set_error(C3D_INVALID_ARGUMENT, "null argument");
    if (r >= n) return set_error(C3D_INVALID_ARGUMENT, "view index out of range");
    std::vector<std::pair<uint32_t, uint32_t>> c;  // (count, view)
    for (uint32_t o = 0; o < n; ++o) {
        if (o == r) continue;
        const size_t g = (size_t)r * n + o;
        if (count[g] == 0) continue;
        if (min_angle_rad > 0.0 && q75[g] < min_angle_rad) continue;
        c.emplace_back(count[g], o);
    }
    std::sort(c.begin(), c.end(), [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    const uint32_t m = (uint32_t)std::min<size_t>(c.size(), max_views);
    for (uint32_t i = 0; i < m; ++i) out[i] = c[i].second;
    *num_out = m;
    return C3D_OK;
}

namespace {
// integer power by repeated squaring (the lowering of f64::powi)
double powi(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}
}  // namespace

extern "C" int c3d_select_diverse(const c3d_sparse_scene* s, const uint32_t* count, const double* q75, uint32_t r, uint32_t max_views,
                                  double min_angle_rad, uint32_t bins, double decay, uint32_t* out, uint32_t* num_out) {
    if (!s || !s->rotation || !s->translation || !count || !q75 || !num_out || (max_views && !out))
        return set_error(C3D_INVALID_ARGUMENT, "null argument");
    const uint32_t n = s->num_views;
    if (r >= n) return set_error(C3D_INVALID_ARGUMENT, "view index out of range");
    std::vector<uint32_t> cands(n);
    uint32_t nc = 0;
    int rc = c3d_select_views(n, count, q75, r, n, min_angle_rad, cands.data(), &nc);
    if (rc) return rc;
    bins = std::max<uint32_t>(bins, 1);
    const double* rr = s->rotation + 9 * (size_t)r;
    double cr[3];
    centre(rr, s->translation + 3 * (size_t)r, cr);
    struct Info {
        uint32_t o;
        double sc;
        uint32_t bin;
    };
    std::vector<Info> info(nc);
    for (uint32_t i = 0; i < nc; ++i) {
        const uint32_t o = cands[i];
        double co[3];
        centre(s->rotation + 9 * (size_t)o, s->translation + 3 * (size_t)o, co);
        const double dd[3] = {co[0] - cr[0], co[1] - cr[1], co[2] - cr[2]};
        double bb[3];
        for (int k = 0; k < 3; ++k) bb[k] = rr[3 * k] * dd[0] + rr[3 * k + 1] * dd[1] + rr[3 * k + 2] * dd[2];  // dir_to_cam
        const double lat = std::sqrt(bb[0] * bb[0] + bb[1] * bb[1]);
        uint32_t bin;
        if (lat < 0.3 * std::fabs(bb[2])) {
            bin = bins;
        } else {
            const double a = std::atan2(bb[1], bb[0]) + 3.141592653589793;
            const double f = a / (2.0 * 3.141592653589793) * (double)bins;
            bin = std::min<uint32_t>(f >= 4294967295.0 ? 4294967295u : (uint32_t)f, bins - 1);
        }
        info[i] = {o, (double)count[(size_t)r * n + o], bin};
    }
    std::vector<int> used(bins + 1, 0);
    std::vector<char> taken(nc, 0);
    uint32_t m = 0;
    while (m < max_views) {
        bool found = false;
        double bs = 0.0;
        uint32_t bi = 0;
        for (uint32_t i = 0; i < nc; ++i) {
            if (taken[i]) continue;
            const double sc = info[i].sc * powi(decay, used[info[i].bin]);
            if (!found || sc > bs) {
                found = true;
                bs = sc;
                bi = i;
            }
        }
        if (!found) break;
        taken[bi] = 1;
        used[info[bi].bin] += 1;
        out[m++] = info[bi].o;
    }
    *num_out = m;
    return C3D_OK;
}
