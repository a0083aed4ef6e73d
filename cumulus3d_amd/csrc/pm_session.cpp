// pm_session.cpp — the C ABI of include/c3d_pm.h: backend (device) and session (device-resident
// images, geometry, states) for the gfx950 PatchStereo kernels.
//
// Host-side behaviour follows the reference's CUDA backend host code:
//   constants          crates/cuda/src/patchmatch.rs:496-521
//   run schedule       crates/cuda/src/patchmatch.rs:752-799 (eval mode 1, then t x {colour 0, 1})
//   per-iteration      crates/cuda/src/patchmatch.rs:767-779 (tau(t), eps_t, phi_t)
//   evaluate / filter  crates/cuda/src/patchmatch.rs:801-824, 842-868
//   snapshot           crates/cuda/src/patchmatch.rs:573-604 (re-upload only on id change)
//   batching           crates/cuda/src/patchmatch.rs:606-625 (pixel and view budgets)
//   errors             crates/cuda/src/patchmatch.rs:292-294, 366-371, 638-642, 753-756
// Differences by design: every image is uploaded once as u8 plus a packed quad table (no texture
// objects); all views of a batch run in one flattened grid; c3d_pm_densify keeps states on the
// device for the whole multi-scale schedule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/c3d_init.h"
#include "../../include/c3d_pm.h"
#include "backend.hpp"
#include "nn_kernels.hpp"
#include "nn_net.hpp"
#include "pm_kernels.hpp"
#include "postproc.hpp"

using namespace c3d;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(C3D_DEVICE, std::string("HIP: ") + #expr + ": " + \
                                                          hipGetErrorString(e_));           \
    } while (0)

const float kPi = 3.14159265358979323846f;
inline float deg2rad(float d) { return d * (kPi / 180.0f); }

// math.rs:49-76 (double)
double erf_series(double x) {
    if (std::fabs(x) > 5.0) return x > 0 ? 1.0 : -1.0;
    double sum = 0.0, term = x, x2 = x * x;
    for (int n = 0; n < 200; ++n) {
        double t = term / (double)(2 * n + 1);
        sum += t;
        double as = std::fabs(sum) > 1e-300 ? std::fabs(sum) : 1e-300;
        if (std::fabs(t) < 1e-17 * as) break;
        term *= -x2 / (double)(n + 1);
    }
    return sum * 2.0 / std::sqrt(M_PI);
}
double emission_norm(double s) { return 2.0 / (std::sqrt(2.0 * M_PI) * s * erf_series(2.0 / (s * M_SQRT2))); }
double emission(double c, double s) { return emission_norm(s) * std::exp(-c * c / (2.0 * s * s)); }

// A = K_j R K^-1, b = K_j t (patchmatch.rs:131-151)
LevelPair level_pair(const float kr[4], const float kj[4], const float r[9], const float t[3]) {
    double fx = kr[0], fy = kr[1], cx = kr[2], cy = kr[3];
    double kinv[3][3] = {{1.0 / fx, 0.0, -cx / fx}, {0.0, 1.0 / fy, -cy / fy}, {0.0, 0.0, 1.0}};
    double kjm[3][3] = {{kj[0], 0.0, kj[2]}, {0.0, kj[1], kj[3]}, {0.0, 0.0, 1.0}};
    double kr_[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s += kjm[i][m] * (double)r[m * 3 + k];
            kr_[i][k] = s;
        }
    LevelPair L;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s += kr_[i][m] * kinv[m][k];
            L.A[i * 3 + k] = (float)s;
        }
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int m = 0; m < 3; ++m) s += kjm[i][m] * (double)t[m];
        L.b[i] = (float)s;
    }
    return L;
}

// A device buffer that grows on demand and is freed with its owner (not copyable).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    hipError_t reserve(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = std::max<size_t>(want, 1);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};


}  // namespace

namespace {

void free_backend(c3d_pm_backend* be) {
    (void)hipSetDevice(be->device);
    if (be->stream) (void)hipStreamDestroy(be->stream);
    delete be;
}

}  // namespace

namespace c3d {
int set_error(int code, const std::string& msg) { return fail(code, msg); }

int backend_hold(c3d_pm_backend* be) {
    std::lock_guard<std::mutex> g(be->life);
    if (be->destroy_pending) return fail(C3D_INVALID_ARGUMENT, "the backend was destroyed");
    ++be->holders;
    return C3D_OK;
}

void backend_release(c3d_pm_backend* be) {
    bool last;
    {
        std::lock_guard<std::mutex> g(be->life);
        last = --be->holders == 0 && be->destroy_pending;
    }
    if (last) free_backend(be);
}
}  // namespace c3d

struct c3d_pm_session {
    c3d_pm_backend* be = nullptr;
    std::unique_lock<std::mutex> guard;
    const KernelSet* ks = nullptr;
    int nv = 0, nl = 0;
    std::vector<LevelDesc> lv;
    std::vector<ViewDesc> vd;
    std::vector<int> h_pair_src;        // [view][MAXSRC] source view index
    std::vector<PairRec> h_pair_rt;     // [view][MAXSRC] reference -> source geometry
    // device tables
    DevBuf<uint8_t> gray;
    DevBuf<uint32_t> quad;
    DevBuf<LevelDesc> d_lv;
    DevBuf<PairRec> d_pairs;
    DevBuf<LevelPair> d_lpairs;
    DevBuf<ViewDesc> d_views;
    DevBuf<BatchItem> d_batch;
    // snapshot
    bool snap_valid = false;
    uint64_t snap_id = 0;
    int snap_level = -1;
    DevBuf<float> snap;
    DevBuf<long long> snap_off;
    // states
    DevBuf<float4> plane, prior, plane2;
    DevBuf<float> cost;
    // device-resident schedule: lower-scale planes, upsampled planes, restorer costs
    DevBuf<float4> low, up;
    DevBuf<float> cinit;
    DevBuf<uint8_t> bview, count, need;
    // downloads: the final planes split into depth and 3-float normal arrays (the page-locked
    // staging buffers belong to the backend)
    DevBuf<float> dl_depth, dl_normal;
    // first half-step tile of each batch item (launch_half_tile_prefix), rebuilt by every run
    DevBuf<int> tile_pre;
    // probes
    DevBuf<int> q_xy;
    DevBuf<float4> q_planes;
    DevBuf<float> q_out;
    KParams base{};
    c3d_pm_params pm_copy{};
    c3d_filter_params filter_copy{};
    std::vector<float4> host4;
    hipEvent_t ev[8] = {};
    // per-launch timing of the half-step kernels: pairs of events, read after the final sync
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    // stage boundaries of c3d_pm_densify: an event opens a segment of a stage kind
    std::vector<hipEvent_t> marks;
    std::vector<int> mark_kind;
    size_t marks_used = 0;
    // c3d_pm_densify: per-scale batch tables and snapshot offsets, uploaded before the device work
    DevBuf<BatchItem> d_batch_lv;
    DevBuf<long long> d_snap_lv;
    std::vector<BatchItem> h_batch_lv;
    std::vector<long long> h_snap_lv;
    std::vector<uint8_t> h_need;
    // half-step workgroup shape: rows of active pixels and lanes per pixel (options.tile_rows and
    // options.threads_per_pixel, or c3d_pm_autotune)
    int half_rows = 1;
    int half_tpp = 1;
    // the captured schedule (options.use_graph) and the key it was captured for
    hipGraphExec_t gexec = nullptr;
    std::vector<uint8_t> gkey;
    c3d::PostWork* post = nullptr;  // scratch of the read-out post-processing
};

namespace {

int set_device(c3d_pm_backend* be) {
    HIP_TRY(hipSetDevice(be->device));
    return C3D_OK;
}

// States of `views` at `level` must match the level size (patchmatch.rs:638-642).
int check_views(c3d_pm_session* s, uint32_t level, const uint32_t* views, uint32_t n, const c3d_view_state* st) {
    if ((int)level >= s->nl) return fail(C3D_INVALID_ARGUMENT, "level " + std::to_string(level) + " out of range");
    for (uint32_t i = 0; i < n; ++i) {
        if ((int)views[i] >= s->nv) return fail(C3D_INVALID_ARGUMENT, "view index out of range");
        if (st) {
            const LevelDesc& L = s->lv[views[i] * s->nl + level];
            if ((int)st[i].width != L.w || (int)st[i].height != L.h)
                return fail(C3D_INVALID_ARGUMENT, "state size " + std::to_string(st[i].width) + "x" + std::to_string(st[i].height) +
                                                      " != level " + std::to_string(L.w) + "x" + std::to_string(L.h));
            if (!st[i].depth || !st[i].normal) return fail(C3D_INVALID_ARGUMENT, "state without depth/normal buffers");
        }
    }
    return C3D_OK;
}

// Bundles views by the pixel and view budgets (patchmatch.rs:606-625): [start, end) pairs.
std::vector<std::pair<uint32_t, uint32_t>> batches(c3d_pm_session* s, const uint32_t* views, uint32_t n, uint32_t level) {
    std::vector<std::pair<uint32_t, uint32_t>> out;
    long long budget = s->be->opts.batch_pixels > 0 ? s->be->opts.batch_pixels : (1LL << 40);
    uint32_t maxb = s->be->opts.max_batch > 0 ? (uint32_t)s->be->opts.max_batch : 64u;
    uint32_t start = 0;
    long long pix = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const LevelDesc& L = s->lv[views[i] * s->nl + level];
        long long p = (long long)L.w * L.h;
        if (i > start && (pix + p > budget || i - start >= maxb)) {
            out.push_back({start, i});
            start = i;
            pix = 0;
        }
        pix += p;
    }
    if (start < n) out.push_back({start, n});
    return out;
}

// Writes batch items for views[a..b) at `level` to the device; returns the pixel count.
int set_batch(c3d_pm_session* s, const uint32_t* views, uint32_t a, uint32_t b, uint32_t level, std::vector<BatchItem>& items,
              long long& npix, int& max_nsrc) {
    items.clear();
    npix = 0;
    max_nsrc = 1;
    for (uint32_t i = a; i < b; ++i) {
        const LevelDesc& L = s->lv[views[i] * s->nl + level];
        BatchItem it{};
        it.off = npix;
        it.view = (int)views[i];
        it.w = L.w;
        it.h = L.h;
        items.push_back(it);
        npix += (long long)L.w * L.h;
        max_nsrc = std::max(max_nsrc, s->vd[views[i]].nsrc);
    }
    HIP_TRY(s->d_batch.reserve(items.size()));
    HIP_TRY(hipMemcpyAsync(s->d_batch.p, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice, s->be->stream));
    HIP_TRY(s->plane.reserve(npix));
    HIP_TRY(s->cost.reserve(npix));
    HIP_TRY(s->bview.reserve(npix));
    HIP_TRY(s->count.reserve(npix));
    HIP_TRY(s->tile_pre.reserve(items.size() + 1));
    return C3D_OK;
}

KParams params_for(c3d_pm_session* s, uint32_t level, bool geometric, const std::vector<BatchItem>& items, int max_nsrc) {
    KParams k = s->base;
    k.gray = s->gray.p;
    k.quad = s->quad.p;
    k.lv = s->d_lv.p;
    k.pairs = s->d_pairs.p;
    k.lpairs = s->d_lpairs.p;
    k.views = s->d_views.p;
    k.batch = s->d_batch.p;
    k.snap = s->snap.p;
    k.snap_off = s->snap_off.p;
    k.plane = s->plane.p;
    k.cost = s->cost.p;
    k.bview = s->bview.p;
    k.count = s->count.p;
    k.prior = s->prior.p;
    k.tile_pre = nullptr;
    k.nitems = (int)items.size();
    k.level = (int)level;
    k.geometric = geometric ? 1 : 0;
    k.wv_stride = 256;
    k.max_nsrc = max_nsrc;
    return k;
}

// The top-scale results of a densify call copied straight into the caller's arrays: the planes
// are split into depth and 3-float normal arrays on the device, then each view's depth, normal,
// cost, raw depth and count come back with one transfer each, in stream order, and one final
// synchronisation (no host staging buffer, no host-side unpacking).
int download_results(c3d_pm_session* s, const BatchItem* items, uint32_t n, long long npix, c3d_view_state* out,
                     uint8_t* const* counts, float* const* raw_depth);
int download_states(c3d_pm_session* s, const float4* plane, const float* cost, const BatchItem* items, uint32_t n, long long npix,
                    c3d_view_state* out, uint8_t* const* counts, float* const* raw_depth);
int upload_states(c3d_pm_session* s, const c3d_view_state* st, const std::vector<BatchItem>& items, long long npix, bool prior,
                  float4* dst);

// Snapshot upload, only when the id changes (patchmatch.rs:573-604).
// The snapshot's maps are laid out with the sizes of snap->level and read with the sizes of the
// run's level, so the two must be the same scale.
int ensure_snapshot(c3d_pm_session* s, const c3d_depth_snapshot* snap, uint32_t level) {
    if (!snap) return fail(C3D_INVALID_ARGUMENT, "a geometric run needs a depth snapshot");
    if ((int)snap->level >= s->nl) return fail(C3D_INVALID_ARGUMENT, "snapshot level out of range");
    if (snap->level != level)
        return fail(C3D_INVALID_ARGUMENT, "snapshot of level " + std::to_string(snap->level) + " for a run at level " + std::to_string(level));
    if (s->snap_valid && s->snap_id == snap->id && s->snap_level == (int)snap->level) return C3D_OK;
    std::vector<long long> offs((size_t)s->nv, -1);
    size_t total = 0;
    for (int v = 0; v < s->nv && (uint32_t)v < snap->num_maps; ++v) {
        if (snap->maps[v]) {
            const LevelDesc& L = s->lv[v * s->nl + snap->level];
            offs[v] = (long long)total;
            total += (size_t)L.w * L.h;
        }
    }
    HIP_TRY(s->snap.reserve(std::max<size_t>(total, 1)));
    HIP_TRY(s->snap_off.reserve((size_t)s->nv));
    for (int v = 0; v < s->nv; ++v) {
        if (offs[v] >= 0) {
            const LevelDesc& L = s->lv[v * s->nl + snap->level];
            HIP_TRY(hipMemcpyAsync(s->snap.p + offs[v], snap->maps[v], (size_t)L.w * L.h * sizeof(float), hipMemcpyHostToDevice,
                                   s->be->stream));
        }
    }
    HIP_TRY(hipMemcpyAsync(s->snap_off.p, offs.data(), offs.size() * sizeof(long long), hipMemcpyHostToDevice, s->be->stream));
    HIP_TRY(hipStreamSynchronize(s->be->stream));
    s->snap_valid = true;
    s->snap_id = snap->id;
    s->snap_level = (int)snap->level;
    return C3D_OK;
}

// Per-iteration constants (patchmatch.rs:767-779).
void iteration_constants(const c3d_pm_params& pm, bool geometric, uint32_t t, KParams& k) {
    float tf = (float)t;
    float eps0 = geometric ? pm.eps0_geometric : pm.eps0_photometric;
    float phi0 = geometric ? pm.phi0_geometric_deg : pm.phi0_photometric_deg;
    float p = 1.0f;
    for (uint32_t i = 0; i < t && p > 0.0f; ++i) p *= 0.5f;
    k.t = (int)t;
    k.tau_good = pm.tau0 * std::exp(-tf * tf / pm.tau_alpha);
    k.eps = std::fmax(eps0 * p, pm.eps_min);
    k.phi = deg2rad(std::fmax(phi0 * p, pm.phi_min_deg));
}

// One run over the current batch: eval mode 1, then iterations x colours (patchmatch.rs:766-779).
// With `timed`, every half-step launch is bracketed by a pair of pool events (read later).
int launch_run(c3d_pm_session* s, KParams k, const std::vector<BatchItem>& items, const c3d_pm_params& pm, bool geometric,
               uint32_t iterations, bool timed) {
    hipStream_t st = s->be->stream;
    k.ntiles = pix_tiles(items.data(), (int)items.size());
    if (iterations == 0) {
        HIP_TRY(launch_pix_tile_prefix(k.batch, k.nitems, s->tile_pre.p, st));
        KParams e = k;
        e.tile_pre = s->tile_pre.p;
        HIP_TRY(s->ks->eval(e, 1, st));
    } else {
        // the colour-0 costs of the run start are overwritten unread by the first half step: only
        // colour 1 is costed (k_eval_init initialises every pixel and resets the best view)
        const long long npix = items.empty() ? 0 : items.back().off + (long long)items.back().w * items.back().h;
        HIP_TRY(launch_eval_init(k, npix, st));
        KParams e = k;
        e.color = 1;
        e.ntiles = half_tiles(items.data(), (int)items.size(), 4, 1);
        HIP_TRY(launch_span_tile_prefix(e.batch, e.nitems, s->tile_pre.p, st));
        e.tile_pre = s->tile_pre.p;
        HIP_TRY(s->ks->eval_color(e, st));
    }
    k.ntiles = half_tiles(items.data(), (int)items.size(), s->half_rows, s->half_tpp);
    if (iterations > 0) {
        HIP_TRY(launch_half_tile_prefix(k.batch, k.nitems, s->half_rows, s->half_tpp, s->tile_pre.p, st));
        k.tile_pre = s->tile_pre.p;
    }
    for (uint32_t t = 0; t < iterations; ++t) {
        iteration_constants(pm, geometric, t, k);
        for (int color = 0; color < 2; ++color) {
            k.color = color;
            if (timed) {
                while (s->pool.size() < s->pool_used + 2) {
                    hipEvent_t e;
                    HIP_TRY(hipEventCreate(&e));
                    s->pool.push_back(e);
                }
                HIP_TRY(hipEventRecord(s->pool[s->pool_used], st));
            }
            HIP_TRY(s->ks->half(k, s->half_rows, s->half_tpp, st));
            if (timed) {
                HIP_TRY(hipEventRecord(s->pool[s->pool_used + 1], st));
                s->pool_used += 2;
            }
        }
    }
    return C3D_OK;
}

// The half step's rows per workgroup for options.tile_rows = 0: one row (one-wave workgroups), or
// four with two lanes per pixel, the only two-lane shape.
int default_rows(int rows, int tpp) { return rows != 0 ? rows : (tpp == 2 ? 4 : 1); }

// Sum of the recorded half-step launch times (after the stream has drained).
int pool_sum(c3d_pm_session* s, double* ms_out, uint32_t* n_out) {
    double tot = 0.0;
    for (size_t i = 0; i + 1 < s->pool_used; i += 2) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->pool[i], s->pool[i + 1]));
        tot += ms;
    }
    *ms_out = tot;
    *n_out = (uint32_t)(s->pool_used / 2);
    s->pool_used = 0;
    return C3D_OK;
}

// Opens a timed segment of stage `kind` (0 photometric, 1 geometric, 2 read-out; -1 closes).
int mark(c3d_pm_session* s, int kind) {
    if (s->marks.size() <= s->marks_used) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        s->marks.push_back(e);
        s->mark_kind.push_back(0);
    }
    HIP_TRY(hipEventRecord(s->marks[s->marks_used], s->be->stream));
    s->mark_kind[s->marks_used] = kind;
    s->marks_used++;
    return C3D_OK;
}

#define MARK(kind)              \
    do {                        \
        int m_ = mark(s, kind); \
        if (m_) return m_;      \
    } while (0)

}  // namespace

// ================================================================ C ABI

extern "C" {

const char* c3d_pm_last_error(void) { return g_error.c_str(); }

int c3d_pm_is_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, 0) != hipSuccess) return 0;
    return std::strncmp(pr.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int c3d_pm_create(int ordinal, const c3d_pm_options* opts, c3d_pm_backend** out) {
    if (!out) return fail(C3D_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(C3D_NOT_AVAILABLE, "no HIP device");
    if (ordinal < 0 || ordinal >= n) return fail(C3D_NOT_AVAILABLE, "device ordinal out of range");
    hipDeviceProp_t pr;
    HIP_TRY(hipGetDeviceProperties(&pr, ordinal));
    if (std::strncmp(pr.gcnArchName, "gfx950", 6) != 0)
        return fail(C3D_NOT_AVAILABLE, std::string("device is ") + pr.gcnArchName + ", this library is built for gfx950");
    c3d_pm_backend* be = new c3d_pm_backend();
    be->device = ordinal;
    be->name = std::string("hip(") + (pr.name[0] ? pr.name : "AMD Instinct") + ", " + pr.gcnArchName + ")";
    if (opts) be->opts = *opts;
    {
        const int rows = be->opts.tile_rows, tpp = be->opts.threads_per_pixel;
        if ((rows != 0 && rows != 1 && rows != 2 && rows != 4) || (tpp != 0 && tpp != 1 && tpp != 2) ||
            !half_shape_ok(default_rows(rows, tpp), tpp == 2 ? 2 : 1)) {
            delete be;
            return fail(C3D_INVALID_ARGUMENT,
                        "half-step shapes are (tile_rows, threads_per_pixel) = (1, 1), (2, 1), (4, 1), (4, 2)");
        }
    }
    if (be->opts.batch_pixels <= 0) be->opts.batch_pixels = 1LL << 40;
    if (be->opts.max_batch <= 0) be->opts.max_batch = 64;
    if (hipSetDevice(ordinal) != hipSuccess || hipStreamCreateWithFlags(&be->stream, hipStreamNonBlocking) != hipSuccess) {
        delete be;
        return fail(C3D_DEVICE, "HIP: cannot create a stream");
    }
    *out = be;
    return C3D_OK;
}

const char* c3d_pm_name(const c3d_pm_backend* be) { return be ? be->name.c_str() : ""; }

int c3d_pm_destroy(c3d_pm_backend* be) {
    if (!be) return 0;
    {
        std::lock_guard<std::mutex> g(be->life);
        if (be->holders > 0) {
            be->destroy_pending = true;
            return be->holders;
        }
    }
    free_backend(be);
    return 0;
}

int c3d_pm_begin(c3d_pm_backend* be, const c3d_kernel_input* in, c3d_pm_session** out) {
    if (!be || !in || !out) return fail(C3D_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const c3d_pm_params& pm = in->pm;
    if (pm.window_radius < 1 || pm.window_radius > 20 || pm.window_step < 1 || pm.window_step > 2)
        return fail(C3D_INVALID_ARGUMENT, "window radius " + std::to_string(pm.window_radius) + " (1..=20), step " +
                                              std::to_string(pm.window_step) + " (1|2)");
    const KernelSet* ks = kernel_set(pm.window_radius, pm.window_step, be->opts.any_radius != 0);
    if (!ks)
        return fail(C3D_INVALID_ARGUMENT, "window radius " + std::to_string(pm.window_radius) + " step " + std::to_string(pm.window_step) +
                                              " has no compiled kernels in this build");
    if (in->num_levels < 1) return fail(C3D_INVALID_ARGUMENT, "num_levels must be >= 1");
    int rc = set_device(be);
    if (rc) return rc;
    // the hold refuses a destroyed backend before waiting for its session lock (an open session
    // keeps that lock); a failed begin gives the hold back after unlocking
    rc = c3d::backend_hold(be);
    if (rc) return rc;
    struct Hold {
        c3d_pm_backend* be;
        bool kept = false;
        ~Hold() {
            if (!kept) c3d::backend_release(be);
        }
    } hold{be};
    std::unique_lock<std::mutex> guard(be->lock);
    const int nv = (int)in->num_views, nl = (int)in->num_levels;
    // validate
    for (int v = 0; v < nv; ++v) {
        const c3d_kernel_view& kv = in->views[v];
        if (kv.num_sources > C3D_MAX_SOURCES)
            return fail(C3D_INVALID_ARGUMENT, "source views " + std::to_string(kv.num_sources) + " > 32");
        for (uint32_t j = 0; j < kv.num_sources; ++j)
            if ((int)kv.sources[j] >= nv || (int)kv.sources[j] == v)
                return fail(C3D_INVALID_ARGUMENT, "source index out of range");
        for (int l = 0; l < nl; ++l) {
            if (!kv.levels[l].gray || kv.levels[l].width < 1 || kv.levels[l].height < 1)
                return fail(C3D_INVALID_ARGUMENT, "empty level image");
            // source-sample table byte offsets are formed exactly in FP32 (multiples of 4 below 2^26)
            if (((uint64_t)kv.levels[l].width + 3) * ((uint64_t)kv.levels[l].height + 3) > (1ULL << 24))
                return fail(C3D_INVALID_ARGUMENT, "image larger than (w + 3)(h + 3) <= 2^24 pixels");
        }
    }
    c3d_pm_session* s = new c3d_pm_session();
    s->be = be;
    hold.kept = true;  // the session owns the hold from here; c3d_pm_end gives it back
    s->ks = ks;
    s->nv = nv;
    s->nl = nl;
    s->lv.resize((size_t)nv * nl);
    s->vd.resize((size_t)nv);
    // image layout
    size_t gtot = 0, qtot = 0;
    for (int v = 0; v < nv; ++v)
        for (int l = 0; l < nl; ++l) {
            const c3d_level_image& li = in->views[v].levels[l];
            LevelDesc& L = s->lv[(size_t)v * nl + l];
            L.w = (int)li.width;
            L.h = (int)li.height;
            L.qpitch = L.w + 3;
            L.goff = (long long)gtot;
            gtot += (size_t)L.w * L.h;
            gtot = (gtot + 255) & ~(size_t)255;
            L.qoff = (long long)qtot + 2LL * L.qpitch + 2;
            qtot += (size_t)L.qpitch * (L.h + 3);
            qtot = (qtot + 63) & ~(size_t)63;
            L.fx = li.k[0];
            L.fy = li.k[1];
            L.cx = li.k[2];
            L.cy = li.k[3];
        }
    auto bail = [&](int code) {
        c3d_pm_end(s);
        return code;
    };
    s->guard = std::move(guard);
    if (hipEventCreate(&s->ev[0]) != hipSuccess) return bail(fail(C3D_DEVICE, "HIP: event"));
    for (int i = 1; i < 8; ++i) (void)hipEventCreate(&s->ev[i]);
    hipStream_t st = be->stream;
    if (s->gray.reserve(gtot) != hipSuccess || s->quad.reserve(qtot) != hipSuccess)
        return bail(fail(C3D_DEVICE, "HIP: out of device memory for images"));
    for (int v = 0; v < nv; ++v)
        for (int l = 0; l < nl; ++l) {
            const c3d_level_image& li = in->views[v].levels[l];
            const LevelDesc& L = s->lv[(size_t)v * nl + l];
            if (hipMemcpyAsync(s->gray.p + L.goff, li.gray, (size_t)L.w * L.h, hipMemcpyHostToDevice, st) != hipSuccess)
                return bail(fail(C3D_DEVICE, "HIP: image upload"));
            if (launch_quad_build(s->gray.p + L.goff, L.w, L.h, s->quad.p + (L.qoff - 2LL * L.qpitch - 2), st) != hipSuccess)
                return bail(fail(C3D_DEVICE, "HIP: quad build"));
        }
    std::vector<PairRec> prs((size_t)nv * MAXSRC);
    std::vector<LevelPair> lps((size_t)nv * nl * MAXSRC);
    std::memset(prs.data(), 0, prs.size() * sizeof(PairRec));
    std::memset(lps.data(), 0, lps.size() * sizeof(LevelPair));
    for (int v = 0; v < nv; ++v) {
        const c3d_kernel_view& kv = in->views[v];
        for (uint32_t j = 0; j < kv.num_sources; ++j) {
            PairRec& p = prs[(size_t)v * MAXSRC + j];
            p.src = (int)kv.sources[j];
            std::memcpy(p.r, kv.pairs[j].r, sizeof p.r);
            std::memcpy(p.t, kv.pairs[j].t, sizeof p.t);
            std::memcpy(p.c, kv.pairs[j].center, sizeof p.c);
            for (int l = 0; l < nl; ++l)
                lps[((size_t)v * nl + l) * MAXSRC + j] = level_pair(kv.levels[l].k, in->views[p.src].levels[l].k, p.r, p.t);
        }
        ViewDesc& d = s->vd[v];
        d.key = kv.key;
        d.dmin = kv.depth_min;
        d.dmax = kv.depth_max;
        d.nsrc = (int)kv.num_sources;
    }
    s->h_pair_rt = prs;
    s->h_pair_src.assign((size_t)nv * MAXSRC, 0);
    for (size_t i = 0; i < prs.size(); ++i) s->h_pair_src[i] = prs[i].src;
    if (s->d_lv.reserve(s->lv.size()) != hipSuccess || s->d_pairs.reserve(prs.size()) != hipSuccess ||
        s->d_lpairs.reserve(lps.size()) != hipSuccess || s->d_views.reserve(s->vd.size()) != hipSuccess ||
        s->snap_off.reserve((size_t)nv) != hipSuccess || s->snap.reserve(1) != hipSuccess)
        return bail(fail(C3D_DEVICE, "HIP: out of device memory for tables"));
    std::vector<long long> none((size_t)nv, -1);
    if (hipMemcpyAsync(s->d_lv.p, s->lv.data(), s->lv.size() * sizeof(LevelDesc), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->d_pairs.p, prs.data(), prs.size() * sizeof(PairRec), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->d_lpairs.p, lps.data(), lps.size() * sizeof(LevelPair), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->d_views.p, s->vd.data(), s->vd.size() * sizeof(ViewDesc), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->snap_off.p, none.data(), none.size() * sizeof(long long), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return bail(fail(C3D_DEVICE, "HIP: table upload"));
    // constants (patchmatch.rs:496-521)
    KParams& k = s->base;
    std::memset(&k, 0, sizeof k);
    float sc = pm.sigma_spatial > 0.0f ? pm.sigma_spatial : (float)pm.window_radius;
    const c3d_filter_params& fl = in->filter;
    k.seed = in->seed;
    k.wr = pm.window_radius;
    k.nlevels = nl;
    k.topk = (int)pm.top_k;
    k.tau_bad = pm.tau1;
    k.inv2beta2 = 1.0f / (2.0f * pm.beta * pm.beta);
    k.prev_w = pm.prev_view_weight;
    k.n1 = (int)pm.n1;
    k.n2 = (int)pm.n2;
    k.inv2s = 1.0f / (2.0f * sc * sc);
    float inv2c = 1.0f / (2.0f * pm.sigma_color * pm.sigma_color);
    k.inv2c255 = inv2c / 65025.0f;
    k.vmin255 = 1e-5f * 65025.0f;
    k.weak_var255 = pm.weak_texture_var * 65025.0f;
    k.cos_tri_min = std::cos(deg2rad(pm.min_triangulation_angle_deg));
    k.inv2inc = 1.0f / (2.0f * pm.incident_angle_sigma * pm.incident_angle_sigma);
    k.lambda = pm.geom_lambda;
    k.gmax = pm.geom_max_cost;
    k.cos_filter_tri = std::cos(deg2rad(fl.min_triangulation_angle_deg));
    k.q_min = (float)emission(1.0 - (double)fl.min_ncc, (double)fl.ncc_sigma);
    k.ncc_norm = (float)emission_norm((double)fl.ncc_sigma);
    k.inv2ncc = 1.0f / (2.0f * fl.ncc_sigma * fl.ncc_sigma);
    k.filter_gmax = fl.geom_max_cost;
    s->pm_copy = pm;
    s->filter_copy = fl;
    s->half_tpp = be->opts.threads_per_pixel == 2 ? 2 : 1;
    s->half_rows = default_rows(be->opts.tile_rows, s->half_tpp);
    *out = s;
    return C3D_OK;
}

void c3d_pm_end(c3d_pm_session* s) {
    if (!s) return;
    (void)hipSetDevice(s->be->device);
    (void)hipStreamSynchronize(s->be->stream);
    s->gray.release();
    s->quad.release();
    s->d_lv.release();
    s->d_pairs.release();
    s->d_lpairs.release();
    s->d_views.release();
    s->d_batch.release();
    s->snap.release();
    s->snap_off.release();
    s->plane.release();
    s->prior.release();
    s->plane2.release();
    s->cost.release();
    s->low.release();
    s->up.release();
    s->cinit.release();
    s->bview.release();
    s->count.release();
    s->need.release();
    s->q_xy.release();
    s->q_planes.release();
    s->q_out.release();
    for (auto& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : s->pool) (void)hipEventDestroy(e);
    for (auto& e : s->marks) (void)hipEventDestroy(e);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->post) c3d::post_work_free(s->post);
    s->post = nullptr;
    s->d_batch_lv.release();
    s->d_snap_lv.release();
    s->dl_depth.release();
    s->dl_normal.release();
    s->tile_pre.release();
    c3d_pm_backend* be = s->be;
    delete s;  // unlocks the backend's session mutex
    c3d::backend_release(be);
}

int c3d_pm_run(c3d_pm_session* s, const uint32_t* views, uint32_t n, c3d_view_state* states, const c3d_run_params* p,
               const c3d_depth_snapshot* snap) {
    if (!s || !p || (n && (!views || !states))) return fail(C3D_INVALID_ARGUMENT, "null argument");
    int rc = set_device(s->be);
    if (rc) return rc;
    rc = check_views(s, p->level, views, n, states);
    if (rc) return rc;
    if (p->geometric) {
        rc = ensure_snapshot(s, snap, p->level);
        if (rc) return rc;
    }
    hipStream_t st = s->be->stream;
    for (auto [a, b] : batches(s, views, n, p->level)) {
        std::vector<BatchItem> items;
        long long npix;
        int max_nsrc;
        rc = set_batch(s, views, a, b, p->level, items, npix, max_nsrc);
        if (rc) return rc;
        bool use_prior = p->use_prior != 0;
        for (uint32_t i = a; i < b; ++i)
            if (!states[i].prior) use_prior = false;
        if (p->random_init != 1) {
            rc = upload_states(s, states + a, items, npix, false, s->plane.p);
            if (rc) return rc;
        }
        if (use_prior) {
            HIP_TRY(s->prior.reserve(npix));
            rc = upload_states(s, states + a, items, npix, true, s->prior.p);
            if (rc) return rc;
        }
        KParams k = params_for(s, p->level, p->geometric != 0, items, max_nsrc);
        k.run_id = (int)p->run_id;
        k.random_init = p->random_init == 1 ? 1 : (p->random_init == 2 ? 2 : 0);
        k.use_prior = use_prior ? 1 : 0;
        rc = launch_run(s, k, items, s->pm_copy, p->geometric != 0, p->iterations, false);
        if (rc) return rc;
        rc = download_states(s, s->plane.p, s->cost.p, items.data(), (uint32_t)items.size(), npix, states + a, nullptr, nullptr);
        if (rc) return rc;
    }
    return C3D_OK;
}

int c3d_pm_evaluate(c3d_pm_session* s, uint32_t level, const uint32_t* views, uint32_t n, const c3d_view_state* states, int geometric,
                    const c3d_depth_snapshot* snap, float* const* costs) {
    if (!s || (n && (!views || !states || !costs))) return fail(C3D_INVALID_ARGUMENT, "null argument");
    int rc = set_device(s->be);
    if (rc) return rc;
    rc = check_views(s, level, views, n, states);
    if (rc) return rc;
    if (geometric) {
        rc = ensure_snapshot(s, snap, level);
        if (rc) return rc;
    }
    hipStream_t st = s->be->stream;
    for (auto [a, b] : batches(s, views, n, level)) {
        std::vector<BatchItem> items;
        long long npix;
        int max_nsrc;
        rc = set_batch(s, views, a, b, level, items, npix, max_nsrc);
        if (rc) return rc;
        rc = upload_states(s, states + a, items, npix, false, s->plane.p);
        if (rc) return rc;
        KParams k = params_for(s, level, geometric != 0, items, max_nsrc);
        k.ntiles = pix_tiles(items.data(), (int)items.size());
        HIP_TRY(s->ks->eval(k, 0, st));
        for (size_t i = 0; i < items.size(); ++i)
            HIP_TRY(hipMemcpyAsync(costs[a + i], s->cost.p + items[i].off, (size_t)items[i].w * items[i].h * sizeof(float),
                                   hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return C3D_OK;
}

int c3d_pm_filter(c3d_pm_session* s, const uint32_t* views, uint32_t n, const c3d_view_state* states, const c3d_depth_snapshot* snap,
                  uint8_t* const* counts) {
    if (!s || !snap || (n && (!views || !states || !counts))) return fail(C3D_INVALID_ARGUMENT, "null argument");
    int rc = set_device(s->be);
    if (rc) return rc;
    rc = ensure_snapshot(s, snap, snap->level);
    if (rc) return rc;
    const uint32_t level = snap->level;
    rc = check_views(s, level, views, n, states);
    if (rc) return rc;
    hipStream_t st = s->be->stream;
    for (auto [a, b] : batches(s, views, n, level)) {
        std::vector<BatchItem> items;
        long long npix;
        int max_nsrc;
        rc = set_batch(s, views, a, b, level, items, npix, max_nsrc);
        if (rc) return rc;
        rc = upload_states(s, states + a, items, npix, false, s->plane.p);
        if (rc) return rc;
        KParams k = params_for(s, level, true, items, max_nsrc);
        k.ntiles = pix_tiles(items.data(), (int)items.size());
        HIP_TRY(s->ks->filter(k, st));
        for (size_t i = 0; i < items.size(); ++i)
            HIP_TRY(hipMemcpyAsync(counts[a + i], s->count.p + items[i].off, (size_t)items[i].w * items[i].h, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return C3D_OK;
}

int c3d_pm_upsample(c3d_pm_session* s, uint32_t level, const uint32_t* views, uint32_t n, const c3d_view_state* low, float sigma_s,
                    float sigma_c, c3d_view_state* out) {
    if (!s || (n && (!views || !low || !out))) return fail(C3D_INVALID_ARGUMENT, "null argument");
    if (level == 0) return fail(C3D_INVALID_ARGUMENT, "upsampling needs level >= 1");
    int rc = set_device(s->be);
    if (rc) return rc;
    rc = check_views(s, level - 1, views, n, low);
    if (rc) return rc;
    rc = check_views(s, level, views, n, out);
    if (rc) return rc;
    hipStream_t st = s->be->stream;
    for (auto [a, b] : batches(s, views, n, level)) {
        std::vector<BatchItem> items;
        long long npix;
        int max_nsrc;
        rc = set_batch(s, views, a, b, level, items, npix, max_nsrc);
        if (rc) return rc;
        // low-resolution planes into the prior buffer
        std::vector<BatchItem> low_items;
        long long lpix = 0;
        for (uint32_t i = a; i < b; ++i) {
            BatchItem li{};
            li.off = lpix;
            li.w = (int)low[i].width;
            li.h = (int)low[i].height;
            low_items.push_back(li);
            items[i - a].low_off = (int)lpix;
            lpix += (long long)li.w * li.h;
        }
        HIP_TRY(hipMemcpyAsync(s->d_batch.p, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice, st));
        HIP_TRY(s->prior.reserve(lpix));
        rc = upload_states(s, low + a, low_items, lpix, false, s->prior.p);
        if (rc) return rc;
        KParams k = params_for(s, level, false, items, max_nsrc);
        k.up_inv2s = 1.0f / (2.0f * sigma_s * sigma_s);
        k.up_inv2c = 1.0f / (2.0f * sigma_c * sigma_c);
        k.ntiles = pix_tiles(items.data(), (int)items.size());
        HIP_TRY(launch_upsample(k, st));
        rc = download_states(s, s->plane.p, nullptr, items.data(), (uint32_t)items.size(), npix, out + a, nullptr, nullptr);
        if (rc) return rc;
    }
    return C3D_OK;
}

int c3d_pm_median(c3d_pm_session* s, uint32_t level, const uint32_t* views, uint32_t n, const c3d_view_state* in,
                  c3d_view_state* out) {
    if (!s || (n && (!views || !in || !out))) return fail(C3D_INVALID_ARGUMENT, "null argument");
    int rc = set_device(s->be);
    if (rc) return rc;
    rc = check_views(s, level, views, n, in);
    if (rc) return rc;
    rc = check_views(s, level, views, n, out);
    if (rc) return rc;
    hipStream_t st = s->be->stream;
    for (auto [a, b] : batches(s, views, n, level)) {
        std::vector<BatchItem> items;
        long long npix;
        int max_nsrc;
        rc = set_batch(s, views, a, b, level, items, npix, max_nsrc);
        if (rc) return rc;
        HIP_TRY(s->prior.reserve(npix));
        rc = upload_states(s, in + a, items, npix, false, s->prior.p);
        if (rc) return rc;
        KParams k = params_for(s, level, false, items, max_nsrc);
        k.ntiles = pix_tiles(items.data(), (int)items.size());
        HIP_TRY(launch_median(k, st));
        rc = download_states(s, s->plane.p, nullptr, items.data(), (uint32_t)items.size(), npix, out + a, nullptr, nullptr);
        if (rc) return rc;
        for (size_t i = 0; i < items.size(); ++i)
            if (out[a + i].cost && in[a + i].cost && out[a + i].cost != in[a + i].cost)
                std::memcpy(out[a + i].cost, in[a + i].cost, (size_t)items[i].w * items[i].h * sizeof(float));
    }
    return C3D_OK;
}

}  // extern "C"

namespace {

// Everything the device-resident schedule needs, laid out before any device work is queued.
struct DensifyPlan {
    uint32_t top = 0, n = 0;
    std::vector<long long> npix;  // per scale
    long long npix_max = 0;
    int max_nsrc = 1;
    bool any_prior = false;
    uint32_t half_launches = 0;
};

// Per-scale batch tables (views back to back, low_off = the view's offset one scale down) and
// snapshot offset tables, written to the device in one upload.
int plan_densify(c3d_pm_session* s, const uint32_t* views, uint32_t n, const c3d_schedule* sch, DensifyPlan& pl) {
    pl.top = (uint32_t)s->nl - 1;
    pl.n = n;
    pl.npix.assign(pl.top + 1, 0);
    s->h_batch_lv.assign((size_t)(pl.top + 1) * n, BatchItem{});
    s->h_snap_lv.assign((size_t)(pl.top + 1) * s->nv, -1);
    for (uint32_t l = 0; l <= pl.top; ++l) {
        long long off = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const LevelDesc& L = s->lv[views[i] * s->nl + l];
            BatchItem& it = s->h_batch_lv[(size_t)l * n + i];
            it.off = off;
            it.view = (int)views[i];
            it.w = L.w;
            it.h = L.h;
            it.low_off = l > 0 ? (int)s->h_batch_lv[(size_t)(l - 1) * n + i].off : 0;
            s->h_snap_lv[(size_t)l * s->nv + views[i]] = off;
            off += (long long)L.w * L.h;
            pl.max_nsrc = std::max(pl.max_nsrc, s->vd[views[i]].nsrc);
        }
        pl.npix[l] = off;
        pl.npix_max = std::max(pl.npix_max, off);
    }
    pl.any_prior = s->nl > 1 && s->pm_copy.weak_texture_var > 0.f;
    pl.half_launches = 2 * (sch->photometric_iters + pl.top * sch->photometric_iters_upper +
                            (pl.top + 1) * sch->geometric_rounds * sch->geometric_iters);
    s->h_need.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t nd = std::min<uint32_t>(s->filter_copy.min_num_consistent, (uint32_t)s->vd[views[i]].nsrc);
        s->h_need[i] = (uint8_t)std::min<uint32_t>(std::max<uint32_t>(nd, 1u), 255u);
    }
    const long long m = pl.npix_max;
    HIP_TRY(s->plane.reserve(m));
    HIP_TRY(s->cost.reserve(m));
    HIP_TRY(s->bview.reserve(m));
    HIP_TRY(s->count.reserve(m));
    HIP_TRY(s->snap.reserve(m));
    HIP_TRY(s->plane2.reserve(m));
    HIP_TRY(s->tile_pre.reserve((size_t)n + 1));
    HIP_TRY(s->need.reserve(n));
    if (s->nl > 1) {
        HIP_TRY(s->low.reserve(m));
        HIP_TRY(s->up.reserve(m));
        HIP_TRY(s->cinit.reserve(m));
    }
    if (pl.any_prior || s->filter_copy.median_filter) HIP_TRY(s->prior.reserve(m));
    HIP_TRY(s->d_batch_lv.reserve(s->h_batch_lv.size()));
    HIP_TRY(s->d_snap_lv.reserve(s->h_snap_lv.size()));
    hipStream_t st = s->be->stream;
    HIP_TRY(hipMemcpyAsync(s->d_batch_lv.p, s->h_batch_lv.data(), s->h_batch_lv.size() * sizeof(BatchItem), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(s->d_snap_lv.p, s->h_snap_lv.data(), s->h_snap_lv.size() * sizeof(long long), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(s->need.p, s->h_need.data(), n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return C3D_OK;
}

// Kernel parameters of scale `level` over the planned batch.
KParams plan_params(c3d_pm_session* s, const DensifyPlan& pl, uint32_t level, bool geometric) {
    std::vector<BatchItem> none;
    KParams k = params_for(s, level, geometric, none, pl.max_nsrc);
    k.nitems = (int)pl.n;
    k.batch = s->d_batch_lv.p + (size_t)level * pl.n;
    k.snap_off = s->d_snap_lv.p + (size_t)level * s->nv;
    return k;
}

// Queues the whole schedule (densify.rs:350-454) on the session stream: kernels and
// device-to-device copies only, so that it can be captured into a graph. With `timed`, stage
// boundaries and every half-step launch are bracketed by events.
int enqueue_densify(c3d_pm_session* s, const DensifyPlan& pl, const c3d_schedule* sch, bool timed) {
    hipStream_t st = s->be->stream;
    const BatchItem* items_all = s->h_batch_lv.data();
    uint32_t run_id = sch->first_run_id;
    for (uint32_t level = 0; level <= pl.top; ++level) {
        // params.rs:476-506 through the schedule struct; use_prior densify.rs:366
        const bool use_prior = level > 0 && pl.any_prior;
        const BatchItem* items = items_all + (size_t)level * pl.n;
        const std::vector<BatchItem> iv(items, items + pl.n);
        const long long npix = pl.npix[level];
        if (timed) MARK(0);
        if (level == 0) {
            KParams k = plan_params(s, pl, 0, false);
            k.run_id = (int)run_id++;
            k.random_init = 1;
            int rc = launch_run(s, k, iv, s->pm_copy, false, sch->photometric_iters, timed);
            if (rc) return rc;
        } else {
            // joint bilateral upsampling of the lower scale (patchmatch.cu:741-800)
            HIP_TRY(hipMemcpyAsync(s->low.p, s->plane.p, (size_t)pl.npix[level - 1] * sizeof(float4), hipMemcpyDeviceToDevice, st));
            KParams u = plan_params(s, pl, level, false);
            u.prior = s->low.p;
            u.up_inv2s = 1.0f / (2.0f * sch->jbu_sigma_spatial * sch->jbu_sigma_spatial);
            u.up_inv2c = 1.0f / (2.0f * sch->jbu_sigma_color * sch->jbu_sigma_color);
            u.ntiles = pix_tiles(items, (int)pl.n);
            HIP_TRY(launch_pix_tile_prefix(u.batch, u.nitems, s->tile_pre.p, st));
            u.tile_pre = s->tile_pre.p;
            HIP_TRY(launch_upsample(u, st));
            if (use_prior)
                HIP_TRY(hipMemcpyAsync(s->prior.p, s->plane.p, (size_t)npix * sizeof(float4), hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(s->up.p, s->plane.p, (size_t)npix * sizeof(float4), hipMemcpyDeviceToDevice, st));
            // detail restorer (densify.rs:373-409): c_init, photometric run, c_photo, merge
            KParams e = plan_params(s, pl, level, false);
            e.ntiles = pix_tiles(items, (int)pl.n);
            e.tile_pre = s->tile_pre.p;  // the 64 x 4 table built for the upsampling
            HIP_TRY(s->ks->eval(e, 0, st));
            HIP_TRY(hipMemcpyAsync(s->cinit.p, s->cost.p, (size_t)npix * sizeof(float), hipMemcpyDeviceToDevice, st));
            KParams k = plan_params(s, pl, level, false);
            k.run_id = (int)run_id++;
            k.random_init = sch->restorer_random_init ? 1 : 0;
            k.use_prior = use_prior ? 1 : 0;
            int rc = launch_run(s, k, iv, s->pm_copy, false, sch->photometric_iters_upper, timed);
            if (rc) return rc;
            HIP_TRY(launch_pix_tile_prefix(e.batch, e.nitems, s->tile_pre.p, st));  // the run rebuilt it
            HIP_TRY(s->ks->eval(e, 0, st));
            HIP_TRY(launch_restore(s->plane.p, s->up.p, s->cinit.p, s->cost.p, sch->restorer_threshold, npix, st));
        }
        if (timed) MARK(1);
        for (uint32_t r = 0; r < sch->geometric_rounds; ++r) {
            HIP_TRY(launch_extract_depth(s->plane.p, s->snap.p, npix, st));
            KParams g = plan_params(s, pl, level, true);
            g.run_id = (int)run_id++;
            g.random_init = 0;
            g.use_prior = use_prior ? 1 : 0;
            int rc = launch_run(s, g, iv, s->pm_copy, true, sch->geometric_iters, timed);
            if (rc) return rc;
        }
    }
    // read-out at the top scale (densify.rs:428-454)
    if (timed) MARK(2);
    const BatchItem* items = items_all + (size_t)pl.top * pl.n;
    const long long npix = pl.npix[pl.top];
    if (s->filter_copy.median_filter) {
        HIP_TRY(hipMemcpyAsync(s->prior.p, s->plane.p, (size_t)npix * sizeof(float4), hipMemcpyDeviceToDevice, st));
        KParams m = plan_params(s, pl, pl.top, false);
        m.ntiles = pix_tiles(items, (int)pl.n);
        HIP_TRY(launch_pix_tile_prefix(m.batch, m.nitems, s->tile_pre.p, st));
        m.tile_pre = s->tile_pre.p;
        HIP_TRY(launch_median(m, st));
    }
    HIP_TRY(launch_extract_depth(s->plane.p, s->snap.p, npix, st));
    KParams f = plan_params(s, pl, pl.top, true);
    f.ntiles = pix_tiles(items, (int)pl.n);
    HIP_TRY(launch_pix_tile_prefix(f.batch, f.nitems, s->tile_pre.p, st));
    f.tile_pre = s->tile_pre.p;
    HIP_TRY(s->ks->filter(f, st));
    HIP_TRY(launch_threshold(f, s->need.p, s->plane2.p, npix, st));
    if (timed) MARK(-1);
    return C3D_OK;
}

// Parameters of the read-out post-processing, checked before any device work.
int check_post(const c3d_post_params& p, const double* post_k) {
    int rc = c3d::post_check(p);
    if (rc) return rc;
    if (p.fill_holes && !post_k) return fail(C3D_INVALID_ARGUMENT, "fill_holes needs the views' intrinsics (post_k)");
    return C3D_OK;
}

// remove_speckles / fill_holes on the thresholded top-scale maps (densify.rs:455-462): planes in
// plane2, the raw cost in `cost` (read by both, overwritten where a pixel is filled), the views'
// gray images at the top scale.
int run_post(c3d_pm_session* s, const c3d_post_params& p, const double* post_k, const BatchItem* items, uint32_t n, c3d_timing* timing) {
    if (timing) {
        timing->post_ms = 0.0;
        timing->post_removed = timing->post_filled = 0;
    }
    if (!p.remove_speckles && !p.fill_holes) return C3D_OK;
    if (!s->post) s->post = c3d::post_work_new();
    const int top = s->nl - 1;
    std::vector<c3d::PostView> pv(n);
    for (uint32_t i = 0; i < n; ++i) {
        pv[i].off = items[i].off;
        pv[i].goff = s->lv[(size_t)items[i].view * s->nl + top].goff;
        pv[i].w = items[i].w;
        pv[i].h = items[i].h;
        for (int q = 0; q < 4; ++q) pv[i].k[q] = post_k ? post_k[4 * i + q] : 0.0;
    }
    std::vector<unsigned long long> rem(n), fil(n);
    const auto t0 = std::chrono::steady_clock::now();
    int rc = c3d::post_run(s->post, pv.data(), (int)n, s->plane2.p, s->cost.p, s->cost.p, s->gray.p, p, s->be->stream, rem.data(),
                           fil.data());
    if (rc) return rc;
    if (timing) {
        timing->post_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (uint32_t i = 0; i < n; ++i) {
            timing->post_removed += rem[i];
            timing->post_filled += fil[i];
        }
    }
    return C3D_OK;
}

// Identity of a captured schedule: the views, the schedule and every device address it reads.
std::vector<uint8_t> graph_key(c3d_pm_session* s, const uint32_t* views, uint32_t n, const c3d_schedule* sch) {
    std::vector<uint8_t> k;
    auto put = [&](const void* p, size_t b) { k.insert(k.end(), (const uint8_t*)p, (const uint8_t*)p + b); };
    put(&n, sizeof n);
    put(views, n * sizeof(uint32_t));
    put(sch, sizeof *sch);
    put(&s->half_rows, sizeof s->half_rows);
    put(&s->half_tpp, sizeof s->half_tpp);
    const void* ptrs[] = {s->plane.p, s->cost.p, s->bview.p, s->count.p, s->snap.p, s->plane2.p, s->need.p, s->low.p,
                          s->up.p, s->cinit.p, s->prior.p, s->d_batch_lv.p, s->d_snap_lv.p, s->tile_pre.p};
    put(ptrs, sizeof ptrs);
    return k;
}

}  // namespace

extern "C" {

int c3d_pm_densify(c3d_pm_session* s, const uint32_t* views, uint32_t n, const c3d_schedule* sch, c3d_view_state* out,
                   uint8_t* const* counts, float* const* raw_depth, c3d_timing* timing) {
    if (!s || !sch || (n && !views)) return fail(C3D_INVALID_ARGUMENT, "null argument");
    if (n == 0) {  // nothing to estimate (densify.rs:290-425 loops over no views)
        if (timing) *timing = c3d_timing{};
        return C3D_OK;
    }
    int rc = set_device(s->be);
    if (rc) return rc;
    const uint32_t top = (uint32_t)s->nl - 1;
    rc = check_views(s, top, views, n, out);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < i; ++j)
            if (views[i] == views[j]) return fail(C3D_INVALID_ARGUMENT, "duplicate view in c3d_pm_densify");
    if (s->nl > 1 && !(sch->jbu_sigma_spatial > 0.f && sch->jbu_sigma_color > 0.f))
        return fail(C3D_INVALID_ARGUMENT, "joint bilateral upsampling needs positive sigmas");
    rc = check_post(sch->post, sch->post_k);
    if (rc) return rc;
    hipStream_t st = s->be->stream;
    DensifyPlan pl;
    rc = plan_densify(s, views, n, sch, pl);
    if (rc) return rc;
    s->snap_valid = false;  // the snapshot buffer now holds device-made maps
    s->pool_used = 0;
    s->marks_used = 0;
    if (s->be->opts.use_graph) {
        // the whole schedule as one graph, captured once per (views, schedule, buffers) and replayed
        std::vector<uint8_t> key = graph_key(s, views, n, sch);
        if (!s->gexec || key != s->gkey) {
            if (s->gexec) HIP_TRY(hipGraphExecDestroy(s->gexec));
            s->gexec = nullptr;
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = enqueue_densify(s, pl, sch, false);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(C3D_DEVICE, std::string("HIP: graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                s->gexec = nullptr;
                return fail(C3D_DEVICE, std::string("HIP: graph instantiate: ") + hipGetErrorString(e));
            }
            s->gkey = key;
        }
        HIP_TRY(hipEventRecord(s->ev[0], st));
        HIP_TRY(hipGraphLaunch(s->gexec, st));
        HIP_TRY(hipEventRecord(s->ev[1], st));
        HIP_TRY(hipEventSynchronize(s->ev[1]));
        if (timing) {
            float tot = 0.f;
            HIP_TRY(hipEventElapsedTime(&tot, s->ev[0], s->ev[1]));
            *timing = c3d_timing{};
            timing->total_ms = tot;
            timing->half_launches = pl.half_launches;
        }
    } else {
        rc = enqueue_densify(s, pl, sch, timing != nullptr);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(st));
        if (timing) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (size_t i = 0; i + 1 < s->marks_used; ++i) {
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, s->marks[i], s->marks[i + 1]));
                if (s->mark_kind[i] >= 0) acc[s->mark_kind[i]] += ms;
            }
            float tot = 0.f;
            HIP_TRY(hipEventElapsedTime(&tot, s->marks[0], s->marks[s->marks_used - 1]));
            timing->photometric_ms = acc[0];
            timing->geometric_ms = acc[1];
            timing->filter_ms = acc[2];
            timing->total_ms = tot;
            rc = pool_sum(s, &timing->half_kernel_ms, &timing->half_launches);
            if (rc) return rc;
        }
    }
    const BatchItem* items = s->h_batch_lv.data() + (size_t)top * n;
    const long long npix = pl.npix[top];
    rc = run_post(s, sch->post, sch->post_k, items, n, timing);
    if (rc) return rc;
    rc = download_results(s, items, n, npix, out, counts, raw_depth);
    if (rc) return rc;
    return C3D_OK;
}

int c3d_pm_densify_learned(c3d_pm_session* s, c3d_init_net* net, const uint32_t* views, uint32_t n, const c3d_learned_schedule* ls,
                           c3d_view_state* out, uint8_t* const* counts, float* const* raw_depth, float* const* init_depth,
                           float* const* init_normal, c3d_timing* timing) {
    if (!s || !net || !ls || (n && !views)) return fail(C3D_INVALID_ARGUMENT, "null argument");
    if (n == 0) {
        if (timing) *timing = c3d_timing{};
        return C3D_OK;
    }
    const nn::NetView nv = nn::net_view(net);
    if (nv.be != s->be) return fail(C3D_INVALID_ARGUMENT, "the network belongs to another backend");
    int rc = check_post(ls->post, ls->post_k);
    if (rc) return rc;
    rc = set_device(s->be);
    if (rc) return rc;
    const uint32_t top = (uint32_t)s->nl - 1;
    rc = check_views(s, top, views, n, out);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < i; ++j)
            if (views[i] == views[j]) return fail(C3D_INVALID_ARGUMENT, "duplicate view in c3d_pm_densify_learned");
    // quarter-resolution images of the views and of their first two sources
    std::vector<int> need_q((size_t)s->nv, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const int v = (int)views[i];
        const LevelDesc& L = s->lv[(size_t)v * s->nl + top];
        if (L.w / 4 > nv.W || L.h / 4 > nv.H || L.w < 4 || L.h < 4)
            return fail(C3D_INVALID_ARGUMENT, "the network input is smaller than a quarter-resolution view");
        if (s->vd[v].nsrc < 1) return fail(C3D_INVALID_ARGUMENT, "a view without sources");
        need_q[v] = 1;
        for (int j = 0; j < std::min(2, s->vd[v].nsrc); ++j) {
            // source index from the pair table
            need_q[(size_t)s->h_pair_src[(size_t)v * MAXSRC + j]] = 1;
        }
    }
    hipStream_t st = s->be->stream;
    std::vector<size_t> qoff((size_t)s->nv, 0), hoff((size_t)s->nv, 0);
    size_t qtot = 0, htot = 0;
    for (int v = 0; v < s->nv; ++v)
        if (need_q[v]) {
            const LevelDesc& L = s->lv[(size_t)v * s->nl + top];
            hoff[v] = htot;
            htot += (((size_t)(L.w / 2) * (L.h / 2)) + 255) & ~(size_t)255;
            qoff[v] = qtot;
            qtot += (((size_t)(L.w / 4) * (L.h / 4)) + 255) & ~(size_t)255;
        }
    DevBuf<uint8_t> half, quarter;
    HIP_TRY(half.reserve(htot));
    HIP_TRY(quarter.reserve(qtot));
    // every buffer of the schedule at the top scale
    std::vector<BatchItem> items;
    long long npix;
    int max_nsrc;
    rc = set_batch(s, views, 0, n, top, items, npix, max_nsrc);
    if (rc) return rc;
    HIP_TRY(s->snap.reserve(npix));
    HIP_TRY(s->plane2.reserve(npix));
    HIP_TRY(s->need.reserve(n));
    std::vector<uint8_t> need(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t nd = std::min<uint32_t>(s->filter_copy.min_num_consistent, (uint32_t)s->vd[views[i]].nsrc);
        need[i] = (uint8_t)std::min<uint32_t>(std::max<uint32_t>(nd, 1u), 255u);
    }
    HIP_TRY(hipMemcpyAsync(s->need.p, need.data(), n, hipMemcpyHostToDevice, st));
    std::vector<long long> offs((size_t)s->nv, -1);
    for (auto& it : items) offs[it.view] = it.off;
    HIP_TRY(hipMemcpyAsync(s->snap_off.p, offs.data(), offs.size() * sizeof(long long), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    s->snap_valid = false;
    s->pool_used = 0;
    s->marks_used = 0;
    MARK(0);
    // blur3 weights of the pyramid (upsample.rs:47-50), as in c3d_pm_build_pyramid
    const float e2 = 0x1.152aaap-3f;
    const float gsum = e2 + 1.0f + e2;
    const float gw[3] = {e2 / gsum, 1.0f / gsum, e2 / gsum};
    for (int v = 0; v < s->nv; ++v)
        if (need_q[v]) {
            const LevelDesc& L = s->lv[(size_t)v * s->nl + top];
            HIP_TRY(launch_pyramid_step(s->gray.p + L.goff, L.w, L.h, half.p + hoff[v], gw, st));
            HIP_TRY(launch_pyramid_step(half.p + hoff[v], L.w / 2, L.h / 2, quarter.p + qoff[v], gw, st));
        }
    auto qk = [&](int v, float k[4]) {
        const LevelDesc& L = s->lv[(size_t)v * s->nl + top];
        double kk[4] = {L.fx, L.fy, L.cx, L.cy};
        for (int r = 0; r < 2; ++r) {
            kk[0] *= 0.5;
            kk[1] *= 0.5;
            kk[2] = (kk[2] + 0.5) * 0.5 - 0.5;
            kk[3] = (kk[3] + 0.5) * 0.5 - 0.5;
        }
        for (int j = 0; j < 4; ++j) k[j] = (float)kk[j];
    };
    for (uint32_t c0 = 0; c0 < n; c0 += (uint32_t)nv.B) {
        const uint32_t c1 = std::min<uint32_t>(n, c0 + (uint32_t)nv.B);
        for (uint32_t i = c0; i < c1; ++i) {
            const int v = (int)views[i];
            const LevelDesc& L = s->lv[(size_t)v * s->nl + top];
            nn::SweepArgs a{};
            a.ref = quarter.p + qoff[v];
            a.rw = L.w / 4;
            a.rh = L.h / 4;
            qk(v, a.rk);
            a.nsrc = std::min(2, s->vd[v].nsrc);
            for (int j = 0; j < a.nsrc; ++j) {
                const int sv = s->h_pair_src[(size_t)v * MAXSRC + j];
                const LevelDesc& S = s->lv[(size_t)sv * s->nl + top];
                a.src[j] = quarter.p + qoff[sv];
                a.sw[j] = S.w / 4;
                a.sh[j] = S.h / 4;
                qk(sv, a.sk[j]);
                std::memcpy(a.r[j], s->h_pair_rt[(size_t)v * MAXSRC + j].r, sizeof a.r[j]);
                std::memcpy(a.t[j], s->h_pair_rt[(size_t)v * MAXSRC + j].t, sizeof a.t[j]);
            }
            a.dmin = s->vd[v].dmin;
            a.dmax = s->vd[v].dmax;
            a.feat = nn::net_features(net, (int)(i - c0));
            a.H = nv.H;
            a.W = nv.W;
            HIP_TRY(nn::sweep(a, st));
            nn::net_set_range(net, (int)(i - c0), a.dmin, a.dmax);
        }
        rc = nn::net_forward(net, (int)(c1 - c0));
        if (rc) return rc;
        for (uint32_t i = c0; i < c1; ++i) {
            const int v = (int)views[i];
            const LevelDesc& L = s->lv[(size_t)v * s->nl + top];
            nn::InitPlanesArgs a{};
            a.out = nn::net_output(net) + (size_t)(i - c0) * nv.H * nv.W * 32;
            a.H = nv.H;
            a.W = nv.W;
            a.qw = L.w / 4;
            a.qh = L.h / 4;
            a.dmin = s->vd[v].dmin;
            a.dmax = s->vd[v].dmax;
            a.w = L.w;
            a.h = L.h;
            a.k[0] = L.fx;
            a.k[1] = L.fy;
            a.k[2] = L.cx;
            a.k[3] = L.cy;
            a.conf_min = ls->conf_min;
            a.plane = s->plane.p + items[i].off;
            HIP_TRY(nn::init_planes(a, st));
        }
    }
    MARK(0);
    if (init_depth || init_normal) {
        s->host4.resize((size_t)npix);
        HIP_TRY(hipMemcpyAsync(s->host4.data(), s->plane.p, (size_t)npix * sizeof(float4), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < n; ++i) {
            const float4* p = s->host4.data() + items[i].off;
            const size_t np = (size_t)items[i].w * items[i].h;
            for (size_t q = 0; q < np; ++q) {
                if (init_depth && init_depth[i]) init_depth[i][q] = p[q].w;
                if (init_normal && init_normal[i]) {
                    init_normal[i][3 * q] = p[q].x;
                    init_normal[i][3 * q + 1] = p[q].y;
                    init_normal[i][3 * q + 2] = p[q].z;
                }
            }
        }
    }
    uint32_t run_id = ls->first_run_id;
    KParams k = params_for(s, top, false, items, max_nsrc);
    k.run_id = (int)run_id++;
    k.random_init = 2;
    rc = launch_run(s, k, items, s->pm_copy, false, ls->refine_iters, timing != nullptr);
    if (rc) return rc;
    MARK(1);
    for (uint32_t r = 0; r < ls->geometric_rounds; ++r) {
        HIP_TRY(launch_extract_depth(s->plane.p, s->snap.p, npix, st));
        KParams g = params_for(s, top, true, items, max_nsrc);
        g.run_id = (int)run_id++;
        g.random_init = 0;
        rc = launch_run(s, g, items, s->pm_copy, true, ls->geometric_iters, timing != nullptr);
        if (rc) return rc;
    }
    MARK(2);
    HIP_TRY(launch_extract_depth(s->plane.p, s->snap.p, npix, st));
    KParams f = params_for(s, top, true, items, max_nsrc);
    f.ntiles = pix_tiles(items.data(), (int)items.size());
    HIP_TRY(launch_pix_tile_prefix(f.batch, f.nitems, s->tile_pre.p, st));
    f.tile_pre = s->tile_pre.p;
    HIP_TRY(s->ks->filter(f, st));
    HIP_TRY(launch_threshold(f, s->need.p, s->plane2.p, npix, st));
    MARK(-1);
    HIP_TRY(hipEventSynchronize(s->marks[s->marks_used - 1]));
    if (timing) {
        // marks: 0 init start, 0 refinement start, 1 geometric, 2 read-out, end
        float a = 0.f, b = 0.f, c = 0.f, d = 0.f, tot = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, s->marks[0], s->marks[1]));
        HIP_TRY(hipEventElapsedTime(&b, s->marks[1], s->marks[2]));
        HIP_TRY(hipEventElapsedTime(&c, s->marks[2], s->marks[3]));
        HIP_TRY(hipEventElapsedTime(&d, s->marks[3], s->marks[4]));
        HIP_TRY(hipEventElapsedTime(&tot, s->marks[0], s->marks[4]));
        timing->init_ms = a;
        timing->photometric_ms = b;
        timing->geometric_ms = c;
        timing->filter_ms = d;
        timing->total_ms = tot;
        rc = pool_sum(s, &timing->half_kernel_ms, &timing->half_launches);
        if (rc) return rc;
    }
    half.release();
    quarter.release();
    rc = run_post(s, ls->post, ls->post_k, items.data(), n, timing);
    if (rc) return rc;
    rc = download_results(s, items.data(), n, npix, out, counts, raw_depth);
    if (rc) return rc;
    return C3D_OK;
}

}  // extern "C"

namespace {
// The depth and normal temporaries of state transfers for `npix` pixels; growing them frees the old
// ones, which queued kernels of the same call may still read, so the stream drains first.
hipError_t grow_temp(c3d_pm_session* s, long long npix) {
    if ((size_t)npix <= s->dl_depth.n && 3 * (size_t)npix <= s->dl_normal.n && s->dl_depth.p && s->dl_normal.p) return hipSuccess;
    hipError_t e = hipStreamSynchronize(s->be->stream);
    if (e != hipSuccess) return e;
    e = s->dl_depth.reserve((size_t)npix);
    if (e != hipSuccess) return e;
    return s->dl_normal.reserve(3 * (size_t)npix);
}

// One result array's device source and host destination.
struct Download {
    const unsigned char* src;
    unsigned char* dst;
    size_t bytes;
};

// A piece of a result array in a staging buffer: its offset there, its destination, its size.
struct Staged {
    size_t at;
    unsigned char* dst;
    size_t bytes;
};

// Copies staged pieces to their destinations on up to 16 host threads, each taking an equal share
// of the bytes.
void copy_out(const unsigned char* stage, const std::vector<Staged>& pieces) {
    size_t total = 0;
    for (const Staged& d : pieces) total += d.bytes;
    if (total == 0) return;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<size_t>(hw, (total + (4u << 20) - 1) / (4u << 20));
    auto work = [&](size_t lo, size_t hi) {  // bytes [lo, hi) of the concatenated pieces
        size_t at = 0;
        for (const Staged& d : pieces) {
            const size_t a = std::max(lo, at), b = std::min(hi, at + d.bytes);
            if (a < b) std::memcpy(d.dst + (a - at), stage + d.at + (a - at), b - a);
            at += d.bytes;
            if (at >= hi) break;
        }
    };
    if (nt <= 1) {
        work(0, total);
        return;
    }
    const size_t share = (total + nt - 1) / nt;
    std::vector<std::thread> th;
    unsigned started = 1;  // shares [1, started) run on their own threads
    try {
        th.reserve(nt - 1);
        for (unsigned t = 1; t < nt; ++t) {
            th.emplace_back(work, std::min(total, t * share), std::min(total, (t + 1) * share));
            started = t + 1;
        }
    } catch (...) {
        // no thread (or no memory for one): the calling thread takes the shares that did not start
    }
    for (unsigned t = started; t < nt; ++t) work(std::min(total, t * share), std::min(total, (t + 1) * share));
    work(0, std::min(total, share));
    for (std::thread& x : th) x.join();
}

// The top-scale results of a densify call copied into the caller's arrays: the planes are split
// into depth and 3-float normal arrays on the device; then every array goes through two
// page-locked staging buffers of DL_STAGE bytes: while the device copies into one, host threads
// copy the other out to the caller's (pageable) arrays.
constexpr size_t DL_STAGE = 256u << 20;
int download_states(c3d_pm_session* s, const float4* plane, const float* cost, const BatchItem* items, uint32_t n, long long npix,
                    c3d_view_state* out, uint8_t* const* counts, float* const* raw_depth) {
    hipStream_t st = s->be->stream;
    if (out) {
        HIP_TRY(grow_temp(s, npix));
        HIP_TRY(launch_split_planes(plane, s->dl_depth.p, s->dl_normal.p, npix, st));
    }
    std::vector<Download> xs;
    for (uint32_t i = 0; i < n; ++i) {
        const size_t np = (size_t)items[i].w * items[i].h;
        const long long off = items[i].off;
        auto add = [&](const void* src, void* dst, size_t bytes) {
            if (dst && bytes) xs.push_back({(const unsigned char*)src, (unsigned char*)dst, bytes});
        };
        if (out) {
            add(s->dl_depth.p + off, out[i].depth, np * sizeof(float));
            add(s->dl_normal.p + 3 * off, out[i].normal, 3 * np * sizeof(float));
            if (cost) add(cost + off, out[i].cost, np * sizeof(float));
        }
        if (raw_depth) add(s->snap.p + off, raw_depth[i], np * sizeof(float));
        if (counts) add(s->count.p + off, counts[i], np);
    }
    if (xs.empty()) {
        HIP_TRY(hipStreamSynchronize(st));
        return C3D_OK;
    }
    // $C3D_PM_STAGE_BYTES (64 bytes or more) shrinks the staging buffers, so that tests can switch
    // between them many times on small jobs
    size_t stage_bytes = DL_STAGE;
    if (const char* e = std::getenv("C3D_PM_STAGE_BYTES")) {
        const long long v = std::atoll(e);
        if (v >= 64 && (size_t)v < DL_STAGE) stage_bytes = (size_t)v;
    }
    PinnedBuf* stage = s->be->dl_stage;
    HIP_TRY(stage[0].reserve(stage_bytes));
    HIP_TRY(stage[1].reserve(stage_bytes));
    std::vector<Staged> staged[2];
    bool busy[2] = {false, false};
    int cur = 0;
    size_t fill = 0;
    auto drain = [&](int b) -> hipError_t {
        if (!busy[b]) return hipSuccess;
        hipError_t e = hipEventSynchronize(s->ev[6 + b]);
        if (e != hipSuccess) return e;
        copy_out(stage[b].p, staged[b]);
        staged[b].clear();
        busy[b] = false;
        return hipSuccess;
    };
    auto close_buffer = [&]() -> hipError_t {  // the device copies into `cur` are queued: switch
        hipError_t e = hipEventRecord(s->ev[6 + cur], st);
        if (e != hipSuccess) return e;
        busy[cur] = true;
        cur = 1 - cur;
        fill = 0;
        return drain(cur);  // the other buffer's pieces out while the device fills this one
    };
    for (const Download& x : xs) {
        size_t done = 0;
        while (done < x.bytes) {
            const size_t take = std::min(x.bytes - done, stage_bytes - fill);
            HIP_TRY(hipMemcpyAsync(stage[cur].p + fill, x.src + done, take, hipMemcpyDeviceToHost, st));
            staged[cur].push_back({fill, x.dst + done, take});
            fill += take;
            done += take;
            if (fill == stage_bytes) HIP_TRY(close_buffer());
        }
    }
    if (fill > 0) HIP_TRY(close_buffer());
    HIP_TRY(drain(0));
    HIP_TRY(drain(1));
    return C3D_OK;
}

int download_results(c3d_pm_session* s, const BatchItem* items, uint32_t n, long long npix, c3d_view_state* out,
                     uint8_t* const* counts, float* const* raw_depth) {
    return download_states(s, s->plane2.p, s->cost.p, items, n, npix, out, counts, raw_depth);
}

// Caller states into device planes at the item offsets (patchmatch.rs:327-336): each view's depth
// and normals (or, with `prior`, its float4 prior planes) go to the device with one transfer each,
// and the planes are assembled there.
int upload_states(c3d_pm_session* s, const c3d_view_state* st, const std::vector<BatchItem>& items, long long npix, bool prior,
                  float4* dst) {
    hipStream_t sm = s->be->stream;
    if (prior) {
        for (size_t i = 0; i < items.size(); ++i)
            HIP_TRY(hipMemcpyAsync(dst + items[i].off, st[i].prior, (size_t)items[i].w * items[i].h * sizeof(float4),
                                   hipMemcpyHostToDevice, sm));
        return C3D_OK;
    }
    HIP_TRY(grow_temp(s, npix));
    for (size_t i = 0; i < items.size(); ++i) {
        const size_t np = (size_t)items[i].w * items[i].h;
        HIP_TRY(hipMemcpyAsync(s->dl_depth.p + items[i].off, st[i].depth, np * sizeof(float), hipMemcpyHostToDevice, sm));
        HIP_TRY(hipMemcpyAsync(s->dl_normal.p + 3 * items[i].off, st[i].normal, 3 * np * sizeof(float), hipMemcpyHostToDevice, sm));
    }
    HIP_TRY(launch_merge_planes(s->dl_depth.p, s->dl_normal.p, dst, npix, sm));
    return C3D_OK;
}
}  // namespace

extern "C" {

int c3d_pm_autotune(c3d_pm_session* s, const uint32_t* views, uint32_t n, uint32_t level, uint32_t* rows_out, uint32_t* tpp_out,
                    double* ms_out) {
    if (!s || !views) return fail(C3D_INVALID_ARGUMENT, "null argument");
    if (n == 0) return fail(C3D_INVALID_ARGUMENT, "no views to time");
    int rc = set_device(s->be);
    if (rc) return rc;
    rc = check_views(s, level, views, n, nullptr);
    if (rc) return rc;
    hipStream_t st = s->be->stream;
    std::vector<BatchItem> items;
    long long npix;
    int max_nsrc;
    rc = set_batch(s, views, 0, n, level, items, npix, max_nsrc);
    if (rc) return rc;
    // scratch states: a random-init evaluation, then two timed iterations per shape after one untimed
    KParams k = params_for(s, level, false, items, max_nsrc);
    k.run_id = 0x7FFFFFFF;
    k.random_init = 1;
    k.ntiles = pix_tiles(items.data(), (int)items.size());
    HIP_TRY(s->ks->eval(k, 1, st));
    const int cand[C3D_TILE_SHAPES][2] = {{1, 1}, {2, 1}, {4, 1}, {4, 2}};
    double ms[C3D_TILE_SHAPES] = {};
    int bestc = 0;
    for (int c = 0; c < C3D_TILE_SHAPES; ++c) {
        KParams h = k;
        h.random_init = 0;
        h.ntiles = half_tiles(items.data(), (int)items.size(), cand[c][0], cand[c][1]);
        HIP_TRY(launch_half_tile_prefix(h.batch, h.nitems, cand[c][0], cand[c][1], s->tile_pre.p, st));
        h.tile_pre = s->tile_pre.p;
        for (int it = 0; it < 3; ++it) {
            iteration_constants(s->pm_copy, false, (uint32_t)it, h);
            if (it == 1) HIP_TRY(hipEventRecord(s->ev[0], st));
            for (int color = 0; color < 2; ++color) {
                h.color = color;
                HIP_TRY(s->ks->half(h, cand[c][0], cand[c][1], st));
            }
        }
        HIP_TRY(hipEventRecord(s->ev[1], st));
        HIP_TRY(hipEventSynchronize(s->ev[1]));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, s->ev[0], s->ev[1]));
        ms[c] = t;
        if (ms[c] < ms[bestc]) bestc = c;
    }
    s->half_rows = cand[bestc][0];
    s->half_tpp = cand[bestc][1];
    if (rows_out) *rows_out = (uint32_t)s->half_rows;
    if (tpp_out) *tpp_out = (uint32_t)s->half_tpp;
    if (ms_out)
        for (int c = 0; c < C3D_TILE_SHAPES; ++c) ms_out[c] = ms[c];
    return C3D_OK;
}

uint32_t c3d_pm_num_levels(uint32_t min_side, uint32_t max_levels, uint32_t min_size) {
    uint32_t l = 1, sz = min_side;
    while (l < max_levels && sz / 2 >= min_size) {
        sz /= 2;
        l += 1;
    }
    return l;
}

int c3d_pm_build_pyramid(c3d_pm_backend* be, const uint8_t* gray, uint32_t width, uint32_t height, const double k[4],
                         uint32_t levels, uint8_t* const* out, float* k_out) {
    if (!be || !gray || !k || !out || !k_out) return fail(C3D_INVALID_ARGUMENT, "null argument");
    if (levels < 1 || levels > 16) return fail(C3D_INVALID_ARGUMENT, "levels must be 1..16");
    if (width < 1 || height < 1) return fail(C3D_INVALID_ARGUMENT, "empty image");
    if ((width >> (levels - 1)) < 1 || (height >> (levels - 1)) < 1)
        return fail(C3D_INVALID_ARGUMENT, "too many levels for the image size");
    for (uint32_t l = 0; l < levels; ++l)
        if (!out[l]) return fail(C3D_INVALID_ARGUMENT, "null level buffer");
    int rc = set_device(be);
    if (rc) return rc;
    // sizes and byte offsets, finest first (reference order is coarsest first)
    std::vector<uint32_t> ws(levels), hs(levels);
    std::vector<size_t> off(levels);
    size_t tot = 0;
    for (uint32_t i = 0; i < levels; ++i) {
        ws[i] = i == 0 ? width : ws[i - 1] / 2;
        hs[i] = i == 0 ? height : hs[i - 1] / 2;
        off[i] = tot;
        tot += ((size_t)ws[i] * hs[i] + 255) & ~(size_t)255;
    }
    // blur3 weights (upsample.rs:47-50): exp(-1 / (2 * 0.25)) = 0x1.152aaap-3 (correctly rounded), normalised
    const float e2 = 0x1.152aaap-3f;
    const float sum = e2 + 1.0f + e2;
    const float g[3] = {e2 / sum, 1.0f / sum, e2 / sum};
    hipStream_t st = be->stream;
    DevBuf<uint8_t> buf;
    HIP_TRY(buf.reserve(tot));
    hipError_t e = hipMemcpyAsync(buf.p, gray, (size_t)width * height, hipMemcpyHostToDevice, st);
    for (uint32_t i = 1; e == hipSuccess && i < levels; ++i)
        e = launch_pyramid_step(buf.p + off[i - 1], (int)ws[i - 1], (int)hs[i - 1], buf.p + off[i], g, st);
    for (uint32_t i = 0; e == hipSuccess && i < levels; ++i)
        e = hipMemcpyAsync(out[levels - 1 - i], buf.p + off[i], (size_t)ws[i] * hs[i], hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    buf.release();
    if (e != hipSuccess) return fail(C3D_DEVICE, std::string("HIP: pyramid: ") + hipGetErrorString(e));
    double kk[4] = {k[0], k[1], k[2], k[3]};
    for (uint32_t i = 0; i < levels; ++i) {
        if (i > 0) {
            kk[0] *= 0.5;
            kk[1] *= 0.5;
            kk[2] = (kk[2] + 0.5) * 0.5 - 0.5;
            kk[3] = (kk[3] + 0.5) * 0.5 - 0.5;
        }
        for (int j = 0; j < 4; ++j) k_out[4 * (levels - 1 - i) + j] = (float)kk[j];
    }
    return C3D_OK;
}

int c3d_pm_pair_costs(c3d_pm_session* s, uint32_t level, uint32_t view, uint32_t slot, uint32_t nq, const int32_t* xy,
                      const float* planes, float* costs) {
    if (!s || (nq && (!xy || !planes || !costs))) return fail(C3D_INVALID_ARGUMENT, "null argument");
    int rc = set_device(s->be);
    if (rc) return rc;
    if ((int)level >= s->nl || (int)view >= s->nv) return fail(C3D_INVALID_ARGUMENT, "level or view out of range");
    if ((int)slot >= s->vd[view].nsrc) return fail(C3D_INVALID_ARGUMENT, "source slot out of range");
    const LevelDesc& R = s->lv[view * s->nl + level];
    for (uint32_t i = 0; i < nq; ++i)
        if (xy[2 * i] < 0 || xy[2 * i + 1] < 0 || xy[2 * i] >= R.w || xy[2 * i + 1] >= R.h)
            return fail(C3D_INVALID_ARGUMENT, "query pixel outside the image");
    if (nq == 0) return C3D_OK;
    hipStream_t st = s->be->stream;
    HIP_TRY(s->q_xy.reserve(2 * (size_t)nq));
    HIP_TRY(s->q_planes.reserve(nq));
    HIP_TRY(s->q_out.reserve(nq));
    HIP_TRY(hipMemcpyAsync(s->q_xy.p, xy, 2 * (size_t)nq * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(s->q_planes.p, planes, (size_t)nq * sizeof(float4), hipMemcpyHostToDevice, st));
    std::vector<BatchItem> none;
    KParams k = params_for(s, level, false, none, 1);
    HIP_TRY(s->ks->pair_costs(k, (int)view, (int)slot, (int)nq, s->q_xy.p, s->q_planes.p, s->q_out.p, st));
    HIP_TRY(hipMemcpyAsync(costs, s->q_out.p, (size_t)nq * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return C3D_OK;
}

int c3d_pm_sample_reciprocal(c3d_pm_backend* be, const float* z, float* r, uint64_t n) {
    if (!be || (n && (!z || !r))) return fail(C3D_INVALID_ARGUMENT, "null argument");
    if (n == 0) return C3D_OK;
    if (n > (1ULL << 32)) return fail(C3D_INVALID_ARGUMENT, "more than 2^32 values in one call");
    int rc = set_device(be);
    if (rc) return rc;
    hipStream_t st = be->stream;
    DevBuf<float> dz, dr;
    HIP_TRY(dz.reserve((size_t)n));
    HIP_TRY(dr.reserve((size_t)n));
    HIP_TRY(hipMemcpyAsync(dz.p, z, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_sample_reciprocal(dz.p, dr.p, (long long)n, st));
    HIP_TRY(hipMemcpyAsync(r, dr.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return C3D_OK;
}

}  // extern "C"
