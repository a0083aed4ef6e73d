// postproc.hip — read-out post-processing of depth maps on the MI355X (include/c3d_pm.h,
// c3d_remove_speckles / c3d_fill_holes and the post stage of c3d_pm_densify):
// small-component removal (remove_speckles, crates/dense/src/postproc.rs:77-133) and
// boundary-aware gap filling (fill_holes, postproc.rs:135-204), applied after the threshold as in
// densify.rs:455-462.
//
// Small components. The reference floods each component from its first pixel in raster order;
// the joining rule |d_j - d_i| <= rel * min(d_i, d_j) is symmetric, so the components are the
// connected components of the 4-neighbour graph of joining pairs, whatever the visiting order,
// and erasing one component never changes another. Here they are found by concurrent union-find
// (atomicMin links, every label points to a smaller index of its component: the root is the
// component's first pixel, as in the reference), first inside 16 x 16 tiles in LDS, then across
// the tile edges in global memory, counted with wave-aggregated atomics, and the
// median cost of every component below the area bound comes from one radix sort of
// (root, cost in total order) keys: the key at position start + size/2 of a root's run is the
// element costs[len / 2] of the reference's sorted list.
//
// Gap filling reads only pixels that were valid before it (a mask taken first) and writes only
// pixels that were not, so every pixel is one independent thread, in double precision like the
// reference.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/c3d_pm.h"
#include "backend.hpp"
#include "postproc.hpp"

namespace c3d {
namespace {

constexpr unsigned NONE = 0xFFFFFFFFu;

struct PItem {
    long long off;   // first pixel in the concatenated arrays
    long long goff;  // first byte of the view's gray image
    int w, h;
    unsigned min_area;  // 0 or 1: no removal (postproc.rs:78-81)
    int maxd;
    double k[4];
};

__device__ __forceinline__ bool item_pixel(const PItem* items, PItem& it, long long& i, int& x, int& y) {
    it = items[blockIdx.y];
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long long)it.w * it.h) return false;
    x = (int)(q % it.w);
    y = (int)(q / it.w);
    i = it.off + q;
    return true;
}

// f32 -> u32 key in IEEE total order (f32::total_cmp) and back
__device__ __forceinline__ unsigned order_key(float c) {
    const unsigned u = __float_as_uint(c);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float order_value(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__device__ __forceinline__ unsigned load_label(const unsigned* L, unsigned a) {
    return __hip_atomic_load(L + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned find_root(const unsigned* L, unsigned a) {
    unsigned p = load_label(L, a);
    while (p != a) {  // labels only point to smaller indices: the walk ends
        a = p;
        p = load_label(L, a);
    }
    return a;
}

// Links the trees of a and b: the larger root is pointed at the smaller (Playne and Hawick,
// "A new algorithm for parallel connected-component labelling on GPUs", IEEE TPDS 2018).
// Every retry starts from a strictly smaller label, so the loop ends.
__device__ void unite(unsigned* L, unsigned a, unsigned b) {
    for (;;) {
        a = find_root(L, a);
        b = find_root(L, b);
        if (a == b) return;
        if (a < b) {
            const unsigned old = atomicMin(L + b, a);
            if (old == b) return;
            b = old;
        } else {
            const unsigned old = atomicMin(L + a, b);
            if (old == a) return;
            a = old;
        }
    }
}

// remove_speckles' joining rule (postproc.rs:108-110), in f32 as there
__device__ __forceinline__ bool joins(float di, float dj, float rel) { return fabsf(dj - di) <= rel * fminf(di, dj); }

// adds `v` for every lane with `on` to *dst with one atomic per wave
__device__ __forceinline__ void wave_count(bool on, unsigned long long* dst) {
    const unsigned long long m = __ballot(on);
    if (m && (threadIdx.x & 63) == (unsigned)__ffsll((long long)m) - 1) atomicAdd(dst, (unsigned long long)__popcll(m));
}

__global__ void k_clear(unsigned long long* c, int n) {
    for (int j = (int)threadIdx.x; j < n; j += (int)blockDim.x) c[j] = 0ull;
}

// Components inside 16 x 16 tiles first, with the labels and the union-find in LDS; each tile's
// labels then point to the tile-local root, which is the smallest global index of its piece (the
// tile's row-major order is the image's). Only the pairs across tile edges are united in global
// memory (k_cc_border).
constexpr int CC_T = 16;

__device__ __forceinline__ unsigned find_lds(const unsigned* L, unsigned a) {
    unsigned p = L[a];
    while (p != a) {
        a = p;
        p = __hip_atomic_load(L + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return a;
}

__device__ void unite_lds(unsigned* L, unsigned a, unsigned b) {
    for (;;) {
        a = find_lds(L, a);
        b = find_lds(L, b);
        if (a == b) return;
        if (a < b) {
            const unsigned old = atomicMin(L + b, a);
            if (old == b) return;
            b = old;
        } else {
            const unsigned old = atomicMin(L + a, b);
            if (old == a) return;
            a = old;
        }
    }
}

__device__ __forceinline__ bool cc_tile(const PItem& it, int& x0, int& y0) {
    const int tx = (it.w + CC_T - 1) / CC_T, ty = (it.h + CC_T - 1) / CC_T;
    if ((long long)blockIdx.x >= (long long)tx * ty) return false;
    x0 = (int)(blockIdx.x % tx) * CC_T;
    y0 = (int)(blockIdx.x / tx) * CC_T;
    return true;
}

__global__ __launch_bounds__(256) void k_cc_local(const PItem* items, const float4* plane, unsigned* label, unsigned* size, float rel) {
    __shared__ unsigned L[CC_T * CC_T];
    __shared__ float D[CC_T * CC_T];
    const PItem it = items[blockIdx.y];
    int x0, y0;
    if (!cc_tile(it, x0, y0)) return;  // uniform over the workgroup
    const int t = threadIdx.x, lx = t % CC_T, ly = t / CC_T;
    const int x = x0 + lx, y = y0 + ly;
    const bool in = x < it.w && y < it.h;
    const long long i = it.off + (long long)y * it.w + x;
    const float d = in ? plane[i].w : 0.0f;
    const bool active = in && !(d <= 0.0f);  // postproc.rs:90: depth <= 0 is no pixel
    D[t] = active ? d : 0.0f;
    L[t] = active ? (unsigned)t : NONE;
    if (in) size[i] = 0u;
    __syncthreads();
    if (active) {
        if (lx + 1 < CC_T && x + 1 < it.w) {
            const float dj = D[t + 1];
            if (!(dj <= 0.0f) && joins(d, dj, rel)) unite_lds(L, (unsigned)t, (unsigned)(t + 1));
        }
        if (ly + 1 < CC_T && y + 1 < it.h) {
            const float dj = D[t + CC_T];
            if (!(dj <= 0.0f) && joins(d, dj, rel)) unite_lds(L, (unsigned)t, (unsigned)(t + CC_T));
        }
    }
    __syncthreads();
    if (in) {
        unsigned g = NONE;
        if (active) {
            const unsigned r = find_lds(L, (unsigned)t);
            g = (unsigned)(it.off + (long long)(y0 + (int)(r / CC_T)) * it.w + x0 + (int)(r % CC_T));
        }
        label[i] = g;
    }
}

// unions across the right and bottom edges of every tile
__global__ __launch_bounds__(256) void k_cc_border(const PItem* items, const float4* plane, unsigned* label, float rel) {
    const PItem it = items[blockIdx.y];
    int x0, y0;
    if (!cc_tile(it, x0, y0)) return;
    const int t = threadIdx.x, lx = t % CC_T, ly = t / CC_T;
    if (lx != CC_T - 1 && ly != CC_T - 1) return;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= it.w || y >= it.h) return;
    const long long i = it.off + (long long)y * it.w + x;
    const float di = plane[i].w;
    if (di <= 0.0f) return;
    if (lx == CC_T - 1 && x + 1 < it.w) {
        const float dj = plane[i + 1].w;
        if (!(dj <= 0.0f) && joins(di, dj, rel)) unite(label, (unsigned)i, (unsigned)(i + 1));
    }
    if (ly == CC_T - 1 && y + 1 < it.h) {
        const float dj = plane[i + it.w].w;
        if (!(dj <= 0.0f) && joins(di, dj, rel)) unite(label, (unsigned)i, (unsigned)(i + it.w));
    }
}

__global__ __launch_bounds__(256) void k_cc_flatten(const PItem* items, unsigned* label, unsigned* size) {
    PItem it;
    long long i;
    int x, y;
    const bool in = item_pixel(items, it, i, x, y);
    unsigned r = NONE;
    if (in && load_label(label, (unsigned)i) != NONE) {
        r = find_root(label, (unsigned)i);
        __hip_atomic_store(label + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // one atomic per distinct root of the wave
    bool pend = r != NONE;
    for (;;) {
        const unsigned long long m = __ballot(pend);
        if (!m) break;
        const int lead = __ffsll((long long)m) - 1;
        const unsigned rl = __shfl(r, lead);
        const bool same = pend && r == rl;
        const unsigned long long ms = __ballot(same);
        if ((int)(threadIdx.x & 63) == lead) atomicAdd(size + rl, (unsigned)__popcll(ms));
        pend = pend && !same;
    }
}

__global__ __launch_bounds__(256) void k_small_keys(const PItem* items, const unsigned* label, const unsigned* size,
                                                    const float* cost, unsigned long long* keys, unsigned long long* counter) {
    PItem it;
    long long i;
    int x, y;
    const bool in = item_pixel(items, it, i, x, y);
    bool small = false;
    unsigned r = NONE;
    if (in && it.min_area > 1u) {
        r = label[i];
        small = r != NONE && size[r] < it.min_area;
    }
    const unsigned long long m = __ballot(small);
    if (!m) return;
    const int lane = (int)(threadIdx.x & 63);
    const int lead = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == lead) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, lead);
    if (small) {
        const unsigned long long below = m & ((1ull << lane) - 1ull);
        keys[base + __popcll(below)] = ((unsigned long long)r << 32) | order_key(cost[i]);
    }
}

__global__ __launch_bounds__(256) void k_seg_median(const unsigned long long* keys, long long n, const unsigned* size, float keep,
                                                    uint8_t* erase) {
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const unsigned r = (unsigned)(keys[p] >> 32);
    if (p > 0 && (unsigned)(keys[p - 1] >> 32) == r) return;
    // costs.sort_by(total_cmp); costs[len / 2] > speckle_keep_cost (postproc.rs:119-121)
    const long long m = p + size[r] / 2;
    if (m >= n) return;  // the run of r holds size[r] keys
    const float med = order_value((unsigned)(keys[m] & 0xFFFFFFFFull));
    erase[r] = med > keep ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_erase(const PItem* items, float4* plane, const unsigned* label, const unsigned* size,
                                               const uint8_t* erase, uint8_t* mask, unsigned long long* removed) {
    PItem it;
    long long i;
    int x, y;
    const bool in = item_pixel(items, it, i, x, y);
    bool gone = false;
    if (in) {
        const unsigned r = label[i];
        gone = it.min_area > 1u && r != NONE && size[r] < it.min_area && erase[r];
        if (gone) plane[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (mask) mask[i] = !gone && plane[i].w > 0.0f;
    }
    wave_count(gone, removed + blockIdx.y);
}

__global__ __launch_bounds__(256) void k_mask(const PItem* items, const float4* plane, uint8_t* mask) {
    PItem it;
    long long i;
    int x, y;
    if (!item_pixel(items, it, i, x, y)) return;
    mask[i] = plane[i].w > 0.0f;  // postproc.rs:148
}

// fill_holes (postproc.rs:135-204) for one pixel
__global__ __launch_bounds__(256) void k_fill(const PItem* items, float4* plane, float* cost, const float* raw_cost, const uint8_t* gray,
                                              const uint8_t* mask, int step, float max_cost, unsigned long long* filled) {
    PItem it;
    long long i;
    int x, y;
    const bool in = item_pixel(items, it, i, x, y);
    bool done = false;
    if (in && !mask[i] && !(raw_cost[i] > max_cost)) {
        const uint8_t* g = gray + it.goff;
        const long long q0 = i - it.off;
        const int ip = g[q0];
        bool found = false;
        int bd = 0, bs = 0;
        long long bq = 0;
        const int DX[8] = {1, -1, 0, 0, 1, -1, 1, -1};
        const int DY[8] = {0, 0, 1, -1, 1, -1, -1, 1};
#pragma unroll 1
        for (int dir = 0; dir < 8; ++dir) {
            int prev = ip;
            for (int s = 1; s <= it.maxd; ++s) {
                const int qx = x + DX[dir] * s, qy = y + DY[dir] * s;
                if (qx < 0 || qy < 0 || qx >= it.w || qy >= it.h) break;
                const long long q = (long long)qy * it.w + qx;
                const int iq = g[q];
                if (abs(iq - prev) > step) break;
                prev = iq;
                if (mask[it.off + q]) {
                    const int dd = abs(iq - ip);
                    if (!found || dd < bd || (dd == bd && s < bs)) {
                        found = true;
                        bd = dd;
                        bs = s;
                        bq = q;
                    }
                    break;
                }
            }
        }
        if (found) {
            const float4 pq = plane[it.off + bq];
            const double n0 = pq.x, n1 = pq.y, n2 = pq.z;
            const double qx = (double)(bq % it.w), qy = (double)(bq / it.w);
            const double rq0 = (qx - it.k[2]) / it.k[0], rq1 = (qy - it.k[3]) / it.k[1];
            const double rp0 = ((double)x - it.k[2]) / it.k[0], rp1 = ((double)y - it.k[3]) / it.k[1];
            // plane_transfer (crates/dense/src/math.rs:146-154)
            const double num = n0 * rq0 + n1 * rq1 + n2 * 1.0;
            const double den = n0 * rp0 + n1 * rp1 + n2 * 1.0;
            if (!(den > -1e-6)) {
                const double d = (double)pq.w * num / den;
                if (d > 0.0 && isfinite(d)) {
                    plane[i] = make_float4(pq.x, pq.y, pq.z, (float)d);
                    cost[i] = cost[it.off + bq];
                    done = true;
                }
            }
        }
    }
    wave_count(done, filled + blockIdx.y);
}

// planes (nx, ny, nz, d) from separate depth and 3-float normal arrays, and back
__global__ __launch_bounds__(256) void k_pack_planes(const float* depth, const float* normal, float4* plane, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) plane[i] = make_float4(normal[3 * i], normal[3 * i + 1], normal[3 * i + 2], depth[i]);
}

__global__ __launch_bounds__(256) void k_unpack_planes(const float4* plane, float* depth, float* normal, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const float4 q = plane[i];
        depth[i] = q.w;
        normal[3 * i] = q.x;
        normal[3 * i + 1] = q.y;
        normal[3 * i + 2] = q.z;
    }
}

template <class T>
hipError_t grow(T*& p, size_t& have, size_t want) {
    if (want <= have && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    have = 0;
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T));
    if (e == hipSuccess) have = std::max<size_t>(want, 1);
    return e;
}

}  // namespace

struct PostWork {
    unsigned* label = nullptr;
    unsigned* size = nullptr;
    unsigned long long* keys = nullptr;
    unsigned long long* keys2 = nullptr;
    uint8_t* erase = nullptr;
    uint8_t* mask = nullptr;
    unsigned long long* counters = nullptr;  // [0] small keys, then removed[n], filled[n]
    PItem* items = nullptr;
    void* tmp = nullptr;
    size_t n_label = 0, n_size = 0, n_keys = 0, n_keys2 = 0, n_erase = 0, n_mask = 0, n_counters = 0, n_items = 0, n_tmp = 0;
    ~PostWork() {
        void* all[] = {label, size, keys, keys2, erase, mask, counters, items, tmp};
        for (void* p : all)
            if (p) (void)hipFree(p);
    }
};

PostWork* post_work_new() { return new PostWork(); }
void post_work_free(PostWork* w) { delete w; }

#define P_TRY(expr)                                                                                                     \
    do {                                                                                                                \
        hipError_t e_ = (expr);                                                                                         \
        if (e_ != hipSuccess) return set_error(C3D_DEVICE, std::string("HIP: ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

int post_check(const c3d_post_params& p) {
    if (p.remove_speckles && !(p.speckle_min_area_frac >= 0.0 && p.speckle_min_area_frac <= 1.0))
        return set_error(C3D_INVALID_ARGUMENT, "speckle_min_area_frac outside [0, 1]");
    if (p.fill_holes && !(p.fill_max_frac >= 0.0 && p.fill_max_frac <= 1.0))
        return set_error(C3D_INVALID_ARGUMENT, "fill_max_frac outside [0, 1]");
    return C3D_OK;
}

int post_run(PostWork* w, const PostView* views, int n, float4* plane, float* cost, const float* raw_cost, const uint8_t* gray,
             const c3d_post_params& p, hipStream_t st, unsigned long long* removed, unsigned long long* filled) {
    if (n < 1 || n > 65535) return set_error(C3D_INVALID_ARGUMENT, "post-processing of 1..65535 views at a time");
    int rc = post_check(p);
    if (rc) return rc;
    if (!p.remove_speckles && !p.fill_holes) {
        for (int v = 0; v < n; ++v) {
            if (removed) removed[v] = 0;
            if (filled) filled[v] = 0;
        }
        return C3D_OK;
    }
    std::vector<PItem> it((size_t)n);
    long long tot = 0, maxpix = 0;
    for (int v = 0; v < n; ++v) {
        const PostView& pv = views[v];
        if (pv.w < 1 || pv.h < 1 || pv.w > 65536 || pv.h > 65536)
            return set_error(C3D_INVALID_ARGUMENT, "post-processing of an image outside 1..65536 pixels a side");
        PItem& q = it[(size_t)v];
        q.off = pv.off;
        q.goff = pv.goff;
        q.w = pv.w;
        q.h = pv.h;
        const long long np = (long long)pv.w * pv.h;
        tot = std::max(tot, pv.off + np);
        maxpix = std::max(maxpix, np);
        // postproc.rs:78: ceil(w h frac); 0 or 1 = nothing to remove
        const double ma = std::ceil((double)np * p.speckle_min_area_frac);
        q.min_area = p.remove_speckles ? (unsigned)std::min(ma, 4294967295.0) : 0u;
        // postproc.rs:146: max(1, round(sqrt(w^2 + h^2) frac)), the walk cannot leave the image anyway
        const unsigned long long d2 = (unsigned long long)pv.w * pv.w + (unsigned long long)pv.h * pv.h;
        const double md = std::round(std::sqrt((double)d2) * p.fill_max_frac);
        q.maxd = (int)std::max(1.0, std::min(md, 131072.0));
        std::memcpy(q.k, pv.k, sizeof q.k);
    }
    if (tot >= (long long)NONE) return set_error(C3D_INVALID_ARGUMENT, "more than 2^32 - 2 pixels in one post-processing call");
    P_TRY(grow(w->items, w->n_items, (size_t)n));
    P_TRY(grow(w->counters, w->n_counters, (size_t)(1 + 2 * n)));
    P_TRY(grow(w->mask, w->n_mask, (size_t)tot));
    P_TRY(hipMemcpyAsync(w->items, it.data(), it.size() * sizeof(PItem), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(1024), 0, st, w->counters, 1 + 2 * n);
    P_TRY(hipGetLastError());
    unsigned long long* d_removed = w->counters + 1;
    unsigned long long* d_filled = w->counters + 1 + n;
    const dim3 grid((unsigned)((maxpix + 255) / 256), (unsigned)n);
    long long maxtiles = 0;
    for (const PItem& q : it) maxtiles = std::max(maxtiles, (long long)((q.w + CC_T - 1) / CC_T) * ((q.h + CC_T - 1) / CC_T));
    const dim3 cc_grid((unsigned)maxtiles, (unsigned)n);
    bool any_small = false;
    for (const PItem& q : it) any_small = any_small || q.min_area > 1u;
    if (p.remove_speckles && any_small) {
        P_TRY(grow(w->label, w->n_label, (size_t)tot));
        P_TRY(grow(w->size, w->n_size, (size_t)tot));
        P_TRY(grow(w->keys, w->n_keys, (size_t)tot));
        P_TRY(grow(w->erase, w->n_erase, (size_t)tot));
        hipLaunchKernelGGL(k_cc_local, cc_grid, dim3(256), 0, st, w->items, plane, w->label, w->size, p.speckle_rel_depth);
        hipLaunchKernelGGL(k_cc_border, cc_grid, dim3(256), 0, st, w->items, plane, w->label, p.speckle_rel_depth);
        hipLaunchKernelGGL(k_cc_flatten, grid, dim3(256), 0, st, w->items, w->label, w->size);
        hipLaunchKernelGGL(k_small_keys, grid, dim3(256), 0, st, w->items, w->label, w->size, cost, w->keys, w->counters);
        P_TRY(hipGetLastError());
        unsigned long long nsmall = 0;
        P_TRY(hipMemcpyAsync(&nsmall, w->counters, sizeof nsmall, hipMemcpyDeviceToHost, st));
        P_TRY(hipStreamSynchronize(st));
        if (nsmall > (unsigned long long)tot || nsmall > 0x7FFFFFFFull)
            return set_error(C3D_DEVICE, "post-processing: more small-component pixels than one sort takes");
        if (nsmall) {
            P_TRY(grow(w->keys2, w->n_keys2, (size_t)nsmall));
            size_t bytes = 0;
            P_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, w->keys, w->keys2, (int)nsmall, 0, 64, st));
            P_TRY(grow((uint8_t*&)w->tmp, w->n_tmp, bytes));
            P_TRY(hipcub::DeviceRadixSort::SortKeys(w->tmp, bytes, w->keys, w->keys2, (int)nsmall, 0, 64, st));
            hipLaunchKernelGGL(k_seg_median, dim3((unsigned)((nsmall + 255) / 256)), dim3(256), 0, st, w->keys2, (long long)nsmall,
                               w->size, p.speckle_keep_cost, w->erase);
            P_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(k_erase, grid, dim3(256), 0, st, w->items, plane, w->label, w->size, w->erase,
                           p.fill_holes ? w->mask : nullptr, d_removed);
        P_TRY(hipGetLastError());
    } else if (p.fill_holes) {
        hipLaunchKernelGGL(k_mask, grid, dim3(256), 0, st, w->items, plane, w->mask);
        P_TRY(hipGetLastError());
    }
    if (p.fill_holes) {
        // postproc.rs:150: (fill_edge_step * 255) as i32, f32 product, truncation (saturating)
        const float sf = p.fill_edge_step * 255.0f;
        const int step = sf != sf ? 0 : (sf >= 2147483647.0f ? 2147483647 : (sf <= -2147483648.0f ? (-2147483647 - 1) : (int)sf));
        hipLaunchKernelGGL(k_fill, grid, dim3(256), 0, st, w->items, plane, cost, raw_cost, gray, w->mask, step, p.fill_max_cost,
                           d_filled);
        P_TRY(hipGetLastError());
    }
    std::vector<unsigned long long> h((size_t)(1 + 2 * n));
    P_TRY(hipMemcpyAsync(h.data(), w->counters, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    P_TRY(hipStreamSynchronize(st));
    for (int v = 0; v < n; ++v) {
        if (removed) removed[v] = h[(size_t)(1 + v)];
        if (filled) filled[v] = h[(size_t)(1 + n + v)];
    }
    return C3D_OK;
}

}  // namespace c3d

using namespace c3d;

namespace {

// one image through the device post stage: planes packed (nx, ny, nz, d) as in the session
int post_one(c3d_pm_backend* be, float* depth, float* normal, float* cost, const float* raw_cost, const uint8_t* gray,
             const double* k, uint32_t w, uint32_t h, const c3d_post_params& p, uint64_t* count) {
    if (w < 1 || h < 1 || (unsigned long long)w * h >= 0xFFFFFFFFull) return set_error(C3D_INVALID_ARGUMENT, "bad image size");
    P_TRY(hipSetDevice(be->device));
    hipStream_t st = be->stream;
    const size_t np = (size_t)w * h;
    std::vector<float4> pl(np);
    for (size_t i = 0; i < np; ++i) pl[i] = make_float4(normal[3 * i], normal[3 * i + 1], normal[3 * i + 2], depth[i]);
    float4* d_plane = nullptr;
    float *d_cost = nullptr, *d_raw = nullptr;
    uint8_t* d_gray = nullptr;
    struct Free {
        std::vector<void*> p;
        ~Free() {
            for (void* q : p)
                if (q) (void)hipFree(q);
        }
    } fr;
    P_TRY(hipMalloc((void**)&d_plane, np * sizeof(float4)));
    fr.p.push_back(d_plane);
    P_TRY(hipMalloc((void**)&d_cost, np * sizeof(float)));
    fr.p.push_back(d_cost);
    if (raw_cost) {
        P_TRY(hipMalloc((void**)&d_raw, np * sizeof(float)));
        fr.p.push_back(d_raw);
        P_TRY(hipMemcpyAsync(d_raw, raw_cost, np * sizeof(float), hipMemcpyHostToDevice, st));
    }
    if (gray) {
        P_TRY(hipMalloc((void**)&d_gray, np));
        fr.p.push_back(d_gray);
        P_TRY(hipMemcpyAsync(d_gray, gray, np, hipMemcpyHostToDevice, st));
    }
    P_TRY(hipMemcpyAsync(d_plane, pl.data(), np * sizeof(float4), hipMemcpyHostToDevice, st));
    P_TRY(hipMemcpyAsync(d_cost, cost, np * sizeof(float), hipMemcpyHostToDevice, st));
    PostView v{};
    v.off = 0;
    v.goff = 0;
    v.w = (int)w;
    v.h = (int)h;
    if (k) std::memcpy(v.k, k, sizeof v.k);
    PostWork work;
    unsigned long long removed = 0, filled = 0;
    int rc = post_run(&work, &v, 1, d_plane, d_cost, d_raw, d_gray, p, st, &removed, &filled);
    if (rc) return rc;
    P_TRY(hipMemcpyAsync(pl.data(), d_plane, np * sizeof(float4), hipMemcpyDeviceToHost, st));
    if (p.fill_holes) P_TRY(hipMemcpyAsync(cost, d_cost, np * sizeof(float), hipMemcpyDeviceToHost, st));
    P_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < np; ++i) {
        depth[i] = pl[i].w;
        normal[3 * i] = pl[i].x;
        normal[3 * i + 1] = pl[i].y;
        normal[3 * i + 2] = pl[i].z;
    }
    if (count) *count = p.remove_speckles ? removed : filled;
    return C3D_OK;
}

}  // namespace

extern "C" int c3d_remove_speckles(c3d_pm_backend* be, float* depth, float* normal, const float* cost, uint32_t width,
                                   uint32_t height, const c3d_post_params* params, uint64_t* removed) {
    if (!be || !depth || !normal || !cost || !params) return set_error(C3D_INVALID_ARGUMENT, "null argument");
    c3d_post_params p = *params;
    p.remove_speckles = 1;
    p.fill_holes = 0;
    // the cost is only read: a private copy keeps the caller's buffer const
    std::vector<float> c(cost, cost + (size_t)width * height);
    return post_one(be, depth, normal, c.data(), nullptr, nullptr, nullptr, width, height, p, removed);
}

extern "C" int c3d_fill_holes(c3d_pm_backend* be, float* depth, float* normal, float* cost, const float* raw_cost, const uint8_t* gray,
                              const double k[4], uint32_t width, uint32_t height, const c3d_post_params* params, uint64_t* filled) {
    if (!be || !depth || !normal || !cost || !raw_cost || !gray || !k || !params)
        return set_error(C3D_INVALID_ARGUMENT, "null argument");
    c3d_post_params p = *params;
    p.remove_speckles = 0;
    p.fill_holes = 1;
    return post_one(be, depth, normal, cost, raw_cost, gray, k, width, height, p, filled);
}

// ---------------------------------------------------------------- allocation-free gap filling
//
// c3d_fill_holes_noalloc is c3d_fill_holes with every device buffer it needs (the staged inputs,
// the packed planes, the validity mask, the item record and the counters) inside `scratch`, a
// fixed array of device memory that each GPU holds from the library's load until the process
// ends. It calls no allocating routine, so callers that cannot allocate (code under HIP graph
// capture, tight serving loops) can use it. Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU race. The per-device address of the array comes from
// hipGetSymbolAddress on the first call on that device; host code uses only that address.
namespace {

constexpr size_t NA_MAX_PIXELS = (size_t)1 << 21;  // 2,097,152 pixels: one 1920 x 1080 view fits
constexpr size_t NA_PLANE = 0;                                     // float4 per pixel
constexpr size_t NA_DEPTH = NA_PLANE + 16 * NA_MAX_PIXELS;         // float per pixel
constexpr size_t NA_NORMAL = NA_DEPTH + 4 * NA_MAX_PIXELS;         // 3 floats per pixel
constexpr size_t NA_COST = NA_NORMAL + 12 * NA_MAX_PIXELS;         // float per pixel
constexpr size_t NA_RAW = NA_COST + 4 * NA_MAX_PIXELS;             // float per pixel
constexpr size_t NA_GRAY = NA_RAW + 4 * NA_MAX_PIXELS;             // byte per pixel
constexpr size_t NA_MASK = NA_GRAY + NA_MAX_PIXELS;                // byte per pixel
constexpr size_t NA_ITEM = NA_MASK + NA_MAX_PIXELS;                // one PItem
constexpr size_t NA_COUNTERS = NA_ITEM + 256;                      // 3 counters (post_run's layout)
constexpr size_t NA_BYTES = NA_COUNTERS + 256;                     // 88,081,408 bytes
static_assert(sizeof(c3d::PItem) <= 256, "item record larger than its slot");
static_assert(NA_ITEM % 256 == 0 && NA_DEPTH % 256 == 0 && NA_COST % 256 == 0, "slots stay 256-byte aligned");

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
unsigned char* na_base[NA_MAX_DEVICES] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

// The device address of `scratch` on `device` (current), looked up once per device.
int na_scratch(int device, unsigned char** out) {
    if (device < 0 || device >= NA_MAX_DEVICES) return set_error(C3D_INVALID_ARGUMENT, "device ordinal beyond 64");
    std::lock_guard<std::mutex> g(na_mutex);
    if (!na_base[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(C3D_DEVICE, std::string("HIP: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        na_base[device] = static_cast<unsigned char*>(p);
    }
    *out = na_base[device];
    return C3D_OK;
}

}  // namespace

extern "C" int c3d_fill_holes_noalloc(c3d_pm_backend* be, float* depth, float* normal, float* cost, const float* raw_cost,
                                      const uint8_t* gray, const double k[4], uint32_t width, uint32_t height,
                                      const c3d_post_params* params, uint64_t* filled) {
    if (!be || !depth || !normal || !cost || !raw_cost || !gray || !k || !params)
        return set_error(C3D_INVALID_ARGUMENT, "null argument");
    if (width < 1 || height < 1) return set_error(C3D_INVALID_ARGUMENT, "bad image size");
    const size_t np = (size_t)width * height;
    if (np > NA_MAX_PIXELS)
        return set_error(C3D_INVALID_ARGUMENT, "c3d_fill_holes_noalloc takes at most 2^21 pixels per call (its device scratch is "
                                               "fixed); use c3d_fill_holes for larger images");
    c3d_post_params p = *params;
    p.remove_speckles = 0;
    p.fill_holes = 1;
    int rc = post_check(p);
    if (rc) return rc;
    P_TRY(hipSetDevice(be->device));
    unsigned char* base = nullptr;
    rc = na_scratch(be->device, &base);
    if (rc) return rc;
    hipStream_t st = be->stream;
    float4* d_plane = reinterpret_cast<float4*>(base + NA_PLANE);
    float* d_depth = reinterpret_cast<float*>(base + NA_DEPTH);
    float* d_normal = reinterpret_cast<float*>(base + NA_NORMAL);
    float* d_cost = reinterpret_cast<float*>(base + NA_COST);
    float* d_raw = reinterpret_cast<float*>(base + NA_RAW);
    uint8_t* d_gray = base + NA_GRAY;
    uint8_t* d_mask = base + NA_MASK;
    PItem* d_item = reinterpret_cast<PItem*>(base + NA_ITEM);
    unsigned long long* d_counters = reinterpret_cast<unsigned long long*>(base + NA_COUNTERS);
    // the one view as post_run describes it (postproc.rs:146 walk length)
    PItem it{};
    it.off = 0;
    it.goff = 0;
    it.w = (int)width;
    it.h = (int)height;
    it.min_area = 0u;
    const unsigned long long d2 = (unsigned long long)width * width + (unsigned long long)height * height;
    const double md = std::round(std::sqrt((double)d2) * p.fill_max_frac);
    it.maxd = (int)std::max(1.0, std::min(md, 131072.0));
    std::memcpy(it.k, k, sizeof it.k);
    const float sf = p.fill_edge_step * 255.0f;  // postproc.rs:150, as in post_run
    const int step = sf != sf ? 0 : (sf >= 2147483647.0f ? 2147483647 : (sf <= -2147483648.0f ? (-2147483647 - 1) : (int)sf));
    P_TRY(hipMemcpyAsync(d_depth, depth, np * sizeof(float), hipMemcpyHostToDevice, st));
    P_TRY(hipMemcpyAsync(d_normal, normal, 3 * np * sizeof(float), hipMemcpyHostToDevice, st));
    P_TRY(hipMemcpyAsync(d_cost, cost, np * sizeof(float), hipMemcpyHostToDevice, st));
    P_TRY(hipMemcpyAsync(d_raw, raw_cost, np * sizeof(float), hipMemcpyHostToDevice, st));
    P_TRY(hipMemcpyAsync(d_gray, gray, np, hipMemcpyHostToDevice, st));
    P_TRY(hipMemcpyAsync(d_item, &it, sizeof it, hipMemcpyHostToDevice, st));
    const unsigned blocks = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_pack_planes, dim3(blocks), dim3(256), 0, st, d_depth, d_normal, d_plane, (long long)np);
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(1024), 0, st, d_counters, 3);
    hipLaunchKernelGGL(k_mask, dim3(blocks, 1), dim3(256), 0, st, d_item, d_plane, d_mask);
    hipLaunchKernelGGL(k_fill, dim3(blocks, 1), dim3(256), 0, st, d_item, d_plane, d_cost, d_raw, d_gray, d_mask, step,
                       p.fill_max_cost, d_counters + 2);
    hipLaunchKernelGGL(k_unpack_planes, dim3(blocks), dim3(256), 0, st, d_plane, d_depth, d_normal, (long long)np);
    P_TRY(hipGetLastError());
    unsigned long long nfilled = 0;
    P_TRY(hipMemcpyAsync(depth, d_depth, np * sizeof(float), hipMemcpyDeviceToHost, st));
    P_TRY(hipMemcpyAsync(normal, d_normal, 3 * np * sizeof(float), hipMemcpyDeviceToHost, st));
    P_TRY(hipMemcpyAsync(cost, d_cost, np * sizeof(float), hipMemcpyDeviceToHost, st));
    P_TRY(hipMemcpyAsync(&nfilled, d_counters + 2, sizeof nfilled, hipMemcpyDeviceToHost, st));
    P_TRY(hipStreamSynchronize(st));
    if (filled) *filled = nfilled;
    return C3D_OK;
}
