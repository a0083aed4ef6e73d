// pm_device.hpp — device-side data layout and arithmetic of the gfx950 PatchStereo kernels.
//
// Layout in HBM (one session):
//   gray   all view-level images, u8 row-major, no padding (reference patch loads)
//   quad   per view-level, a (w+3) x (h+3) table of u32: entry (x0, y0) for x0 in [-2, w],
//          y0 in [-2, h] packs the bilinear 2x2 neighbourhood p(x0,y0) | p(x0+1,y0) << 8 |
//          p(x0,y0+1) << 16 | p(x0+1,y0+1) << 24 with 0 outside the image, so that one dword load
//          gives all four texels of a source sample and clamping the integer sample position into
//          the table reproduces "0 outside" exactly (patchmatch.cu:169-192 software bilinear).
//   plane  float4 (nx, ny, nz, d) per pixel, row-major per view, views of a batch back to back
//          (ViewState packed as in patchmatch.rs:327-348); cost f32, best view u8, count u8 alike.
//
// Arithmetic: every operation is written out (explicit fmaf where an FMA is intended; the library
// is compiled with -ffp-contract=off) in the order the CPU oracle uses, with IEEE division and
// square root, the polynomial exp / sin-cos below and a seeded reciprocal for the per-sample
// projective divide, so that the device and the CPU oracle agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace c3d {

constexpr int MAXSRC = 32;
constexpr int MAX_WR = 20;  // largest window radius (patchmatch.rs:292-294)
constexpr float INF_ERR = 1e30f;

struct LevelDesc {
    int w, h;
    int qpitch;      // quad row pitch (w + 3)
    int pad0;
    long long qoff;  // dword offset of quad entry (0, 0) (the 2-texel border precedes it)
    long long goff;  // byte offset of the u8 image
    float fx, fy, cx, cy;
};

struct PairRec {
    int src;
    float r[9], t[3], c[3];
};

struct LevelPair {
    float A[9];
    float b[3];
};

struct ViewDesc {
    unsigned long long key;
    float dmin, dmax;
    int nsrc, pad0;
};

struct BatchItem {
    long long off;  // first pixel of this view in the batch state buffers
    int view, w, h;
    int low_off;    // upsampling: first pixel of the low-resolution input in the prior buffer
};

struct KParams {
    unsigned long long seed;
    const uint8_t* gray;
    const uint32_t* quad;
    const LevelDesc* lv;
    const PairRec* pairs;
    const LevelPair* lpairs;
    const ViewDesc* views;
    const BatchItem* batch;
    const float* snap;
    const long long* snap_off;
    float4* plane;
    float* cost;
    uint8_t* bview;
    uint8_t* count;
    const float4* prior;
    int nitems, nlevels, level, geometric;
    int run_id, t, color, random_init;
    int topk, n1, n2, use_prior;
    float tau_good, tau_bad, inv2beta2, prev_w;
    float inv2s, inv2c255, vmin255, weak_var255;
    float cos_tri_min, inv2inc, lambda, gmax;
    float eps, phi, up_inv2s, up_inv2c;
    float cos_filter_tri, q_min, ncc_norm, inv2ncc;
    float filter_gmax;
    // first flattened tile of each batch item for the launch's tile shape, tile_pre[nitems] = all
    // tiles (null: locate_tile searches the items linearly)
    const int* tile_pre;
    int wv_stride;  // LDS floats per source in the per-thread view-weight table
    int ntiles;     // total workgroups of the launch
    int max_nsrc;   // largest source count of the batch (sizes the LDS view-weight table)
    int wr;         // window radius (read by the kernels compiled for any radius)
};

// ---------------------------------------------------------------- work distribution

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2). Map consecutive logical
// tiles onto one XCD so that neighbouring tiles share their source footprints in that L2.
__device__ __forceinline__ int xcd_tile(int bid, int n) {
    const int nx = 8;
    int q = n / nx, r = n % nx;
    int x = bid % nx, k = bid / nx;
    return x * q + (x < r ? x : r) + k;
}

// Finds the batch item of flattened tile `t` for tiles of tw x th pixels; returns the item and
// the tile's pixel origin, or false when t is past the last item.
__device__ __forceinline__ bool locate_tile(const KParams& P, int t, int tw, int th, BatchItem& it, int& x0, int& y0,
                                            int* item = nullptr) {
    for (int i = 0; i < P.nitems; ++i) {
        BatchItem b = P.batch[i];
        int tx = (b.w + tw - 1) / tw, ty = (b.h + th - 1) / th;
        int n = tx * ty;
        if (t < n) {
            it = b;
            if (item) *item = i;
            y0 = (t / tx) * th;
            x0 = (t % tx) * tw;
            return true;
        }
        t -= n;
    }
    return false;
}

// locate_tile through the launch's tile offsets (P.tile_pre, built for the same tw x th): a binary
// search over the items instead of a linear one. The last item whose first tile is at or before t
// holds it (items without tiles share their successor's offset and are passed over).
__device__ __forceinline__ bool locate_tile_pre(const KParams& P, int t, int tw, int th, BatchItem& it, int& x0, int& y0) {
    const int* pre = P.tile_pre;
    if (t >= pre[P.nitems]) return false;
    int lo = 0, hi = P.nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    it = P.batch[lo];
    const int tl = t - pre[lo];
    const int tx = (it.w + tw - 1) / tw;
    y0 = (tl / tx) * th;
    x0 = (tl % tx) * tw;
    return true;
}

// The tile's item and pixel origin through P.tile_pre when the launch provides it (built for the
// same tw x th), else by the linear search.
__device__ __forceinline__ bool find_tile(const KParams& P, int t, int tw, int th, BatchItem& it, int& x0, int& y0) {
    return P.tile_pre ? locate_tile_pre(P, t, tw, th, it, x0, y0) : locate_tile(P, t, tw, th, it, x0, y0);
}

// ---------------------------------------------------------------- random numbers (kernel.rs:248-259)

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float rnd(unsigned long long& s) {
    s += 0x9e3779b97f4a7c15ULL;
    unsigned long long z = mix64(s);
    return ((float)(z >> 40) + 1.0f) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ unsigned long long rng_state(unsigned long long seed, unsigned long long key, int level, int run,
                                                        int t, int half, unsigned long long pixel) {
    unsigned long long tag = ((unsigned long long)level << 48) ^ ((unsigned long long)run << 32) ^
                             ((unsigned long long)t << 16) ^ (unsigned long long)half;
    return mix64(seed ^ mix64(key ^ mix64(tag)) ^ pixel);
}

// ---------------------------------------------------------------- elementary functions

// exp: Cody-Waite reduction, degree-7 Taylor polynomial, scale by 2^n; 0 below -86, +inf above
// 88, NaN kept. Branch-free: the polynomial runs on the argument clamped to [-86, 88] and the
// special values are selected afterwards (same results as testing first).
__device__ __forceinline__ float c3d_expf(float x) {
    const float xc = fminf(fmaxf(x, -86.0f), 88.0f);
    float n = rintf(xc * 1.44269504088896341f);
    float r = fmaf(-n, 0.693145751953125f, xc);
    r = fmaf(-n, 1.42860677e-06f, r);
    float p = 1.98412698e-04f;
    p = fmaf(p, r, 1.38888889e-03f);
    p = fmaf(p, r, 8.33333333e-03f);
    p = fmaf(p, r, 4.16666667e-02f);
    p = fmaf(p, r, 1.66666667e-01f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int ni = (int)n;
    float e = p * __int_as_float((ni + 127) << 23);
    e = x > -86.0f ? e : 0.0f;
    e = x > 88.0f ? __builtin_inff() : e;
    return x != x ? x : e;
}

// sin and cos: quadrant reduction, Taylor polynomials on |r| <= pi/4.
__device__ __forceinline__ void c3d_sincosf(float x, float& so, float& co) {
    float k = rintf(x * 0.636619772367581343f);
    float r = fmaf(-k, 1.57079625129699707f, x);
    r = fmaf(-k, 7.54978941586159635e-08f, r);
    r = fmaf(-k, 5.39030252995776476e-15f, r);
    float r2 = r * r;
    float ps = 2.75573192e-06f;
    ps = fmaf(r2, ps, -1.98412698e-04f);
    ps = fmaf(r2, ps, 8.33333333e-03f);
    ps = fmaf(r2, ps, -1.66666667e-01f);
    float s = fmaf(r * r2, ps, r);
    float pc = -2.75573192e-07f;
    pc = fmaf(r2, pc, 2.48015873e-05f);
    pc = fmaf(r2, pc, -1.38888889e-03f);
    pc = fmaf(r2, pc, 4.16666667e-02f);
    pc = fmaf(r2, pc, -0.5f);
    float c = fmaf(r2, pc, 1.0f);
    int q = ((int)k) & 3;
    if (q == 0) { so = s; co = c; }
    else if (q == 1) { so = c; co = -s; }
    else if (q == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
}

// Reciprocal of the projective divide of one source sample: bit-pattern seed (relative error
// below 2^-4.3), one third-order step r(1 + e + e^2) and one Newton step r(1 + e), e = 1 - z r,
// leaving 2^-25.8 before the last rounding: within 1 ulp of 1/z for positive normal z (0.78 ulp
// at most over a binade), reproducible on any IEEE machine, 6 full-rate VALU operations. The
// hardware estimate v_rcp_f32 plus one Newton step would be the correctly rounded quotient, but
// v_rcp_f32 issues at a quarter of the rate and measured slower (26.65 against 27.21 ms per half
// step on 7 x 1920x1080 with 6 sources).
__device__ __forceinline__ float rcp_sample(float z) {
    float r = __uint_as_float(0x7EF311C3u - __float_as_uint(z));
    float e = fmaf(-z, r, 1.0f);
    r = fmaf(r, fmaf(e, e, e), r);
    e = fmaf(-z, r, 1.0f);
    r = fmaf(r, e, r);
    return r;
}

// ---------------------------------------------------------------- vectors

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 nrm3(f3 a) {
    float d = dot3(a, a);
    float l = 1.0f / sqrtf(d > 1e-30f ? d : 1e-30f);
    return mk3(a.x * l, a.y * l, a.z * l);
}
__device__ __forceinline__ f3 xyz(float4 h) { return mk3(h.x, h.y, h.z); }
__device__ __forceinline__ float4 mk4(f3 n, float d) { return make_float4(n.x, n.y, n.z, d); }

// patchmatch.cu:386
__device__ __forceinline__ f3 pixel_ray(const LevelDesc& R, int x, int y) {
    return mk3(((float)x - R.cx) / R.fx, ((float)y - R.cy) / R.fy, 1.0f);
}

// ---------------------------------------------------------------- reference window

template <int WR, int WSTEP>
struct Win {
    static constexpr int WN = (2 * WR) / WSTEP + 1;
    static constexpr int WN2 = WN * WN;
    static constexpr int LAST = -WR + (WN - 1) * WSTEP;  // last sample offset
};

// Weighted statistics of the reference patch (patchmatch.cu:199-223): bilateral weights w_k in
// registers, weighted mean and two-pass variance in gray levels. The window values I_k themselves
// are re-read from the reference accessor (LDS) where the covariance needs them.
template <int WR, int WSTEP>
struct RefWin {
    float ic, mu, W, iW, var;
    float w[Win<WR, WSTEP>::WN2];
};

// Reference image accessors: 0 outside the image.
struct RefGlobal {
    const uint8_t* p;
    int w, h;
    __device__ __forceinline__ float at(int x, int y) const {
        if (x < 0 || y < 0 || x >= w || y >= h) return 0.0f;
        return (float)p[(size_t)y * w + x];
    }
};

template <int TW>
struct RefTile {
    const float* s;
    int x0, y0;
    __device__ __forceinline__ float at(int x, int y) const { return s[(y - y0) * TW + (x - x0)]; }
};

template <int WR, int WSTEP, class REF>
__device__ __forceinline__ void ref_stats(const KParams& P, const REF& ref, int x, int y, RefWin<WR, WSTEP>& rw) {
    constexpr int WN = Win<WR, WSTEP>::WN;
    rw.ic = ref.at(x, y);
    float W = 0.0f, S = 0.0f;
#pragma unroll
    for (int a = 0; a < WN; ++a) {
        const int dy = -WR + a * WSTEP;
#pragma unroll
        for (int b = 0; b < WN; ++b) {
            const int dx = -WR + b * WSTEP;
            const int k = a * WN + b;
            float v = ref.at(x + dx, y + dy);
            float d = rw.ic - v;
            float sp = (float)(dx * dx + dy * dy) * P.inv2s;
            float w = c3d_expf(-fmaf(d * d, P.inv2c255, sp));
            rw.w[k] = w;
            W += w;
            S = fmaf(w, v, S);
        }
    }
    float mu = S / W;
    float V = 0.0f;
#pragma unroll
    for (int a = 0; a < WN; ++a) {
        const int dy = -WR + a * WSTEP;
#pragma unroll
        for (int b = 0; b < WN; ++b) {
            const int dx = -WR + b * WSTEP;
            const int k = a * WN + b;
            float dk = ref.at(x + dx, y + dy) - mu;
            float wd = rw.w[k] * dk;
            V = fmaf(wd, dk, V);
        }
    }
    rw.mu = mu;
    rw.W = W;
    rw.iW = 1.0f / W;
    rw.var = V / W;
}

// ---------------------------------------------------------------- homography and cost

// Plane-induced homography (patchmatch.cu:225-242) translated to the pixel: H maps the window offset
// (dx, dy, 1) of pixel (x, y), H = (A + b n^T K^-1 / delta) T(x, y). Its third column is
// A (x, y, 1)^T + b / d exactly (n^T K^-1 (x, y, 1)^T / delta = 1 / d), free of the cancellation of
// the untranslated form for strongly slanted planes.
__device__ __forceinline__ bool homography(const LevelDesc& R, const LevelPair& L, float ifx, float ify, float x, float y, float d,
                                           f3 n, float* H) {
    f3 ray = mk3((x - R.cx) * ifx, (y - R.cy) * ify, 1.0f);
    float delta = d * dot3(n, ray);
    if (!(delta < -1e-12f)) return false;
    float id = 1.0f / delta, invd = 1.0f / d;
    float u0 = n.x * ifx, u1 = n.y * ify;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float br = L.b[r] * id;
        H[r * 3 + 0] = L.A[r * 3 + 0] + br * u0;
        H[r * 3 + 1] = L.A[r * 3 + 1] + br * u1;
        H[r * 3 + 2] = fmaf(L.A[r * 3 + 0], x, fmaf(L.A[r * 3 + 1], y, L.A[r * 3 + 2])) + L.b[r] * invd;
    }
    return true;
}

// Bilinear interpolation of one quad entry q = p00 | p10 << 8 | p01 << 16 | p11 << 24 at the
// fractions (ax, ay): top = p00 + ax (p10 - p00), bot = p01 + ax (p11 - p01), top + ay (bot - top),
// each an FP32 FMA. The two row starts and the two row differences are formed as packed half
// floats: a byte permute sets the exponent byte 0x64 above each texel, which reads 1024 + p
// exactly, and one packed subtraction each gives (p00, p01) and (p10 - p00, p11 - p01), all
// integers of magnitude below 2^11 and so exact. Each row is then one mixed FMA on the exact
// half-float operands: the same value as the FP32 formula, in 8 instructions instead of 10.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float quad_lerp(uint32_t q, float ax, float ay) {
    const uint32_t lo = __builtin_amdgcn_perm(0x64646464u, q, 0x04020400u);  // (1024 + p00, 1024 + p01)
    const uint32_t hi = __builtin_amdgcn_perm(0x64646464u, q, 0x04030401u);  // (1024 + p10, 1024 + p11)
    const half2_t l = __builtin_bit_cast(half2_t, lo), h = __builtin_bit_cast(half2_t, hi);
    const half2_t d = h - l;
    const half2_t p = l - (half2_t){(_Float16)1024.0f, (_Float16)1024.0f};
    const float top = fmaf(ax, (float)d.x, (float)p.x);
    const float bot = fmaf(ax, (float)d.y, (float)p.y);
    return fmaf(ay, bot - top, top);
}

// One source sample in gray levels through the quad table (see the layout note above). `qbase` is
// the table's first entry (x0, y0) = (-2, -2); the byte offset 4 ((y0 + 2) qpitch + (x0 + 2)) is formed
// exactly in FP32 (multiples of 4 below 2^26) from the clamped floor of the sample position, and the
// load uses a uniform base plus a 32-bit offset.
__device__ __forceinline__ float quad_sample(const uint32_t* qbase, float qp4, float koff4, float w1, float h1, float u, float v) {
    float xf = floorf(u), yf = floorf(v);
    float ax = u - xf, ay = v - yf;
    float xc = __builtin_amdgcn_fmed3f(xf, -2.0f, w1);
    float yc = __builtin_amdgcn_fmed3f(yf, -2.0f, h1);
    uint32_t off = (uint32_t)fmaf(yc, qp4, fmaf(xc, 4.0f, koff4));
    uint32_t q = *(const uint32_t*)((const char*)qbase + off);
    return quad_lerp(q, ax, ay);
}

// Gathers kept in flight by the window loop: this default serves the cost-probe kernel; the
// evaluation and read-out keep five (three waves per SIMD) and the half step picks its own per
// build. Measured on 7 x 1920x1080, 6 sources, with 4-row half-step workgroups: 4 -> 502.8 ms,
// 6 -> 500.8 ms, 8 -> 500.8 ms per job (photometric half steps 2.9 ms faster, geometric ones
// 2.8 ms slower), 12 spills; the compiler's own batches of eight gathers drained to zero before
// the next batch gave 512.9 ms.
constexpr int WINDOW_PIPE = 8;

// Window sums of one hypothesis x source (H translated to the pixel). A sample whose projective
// depth z is not above 1e-8 contributes the value 0 (patchmatch.cu:271-276). CHECK = false when all
// four window corners have z > 1e-8: z is monotone in each offset (every step is a correctly
// rounded affine operation), so then every sample does. With CHECK the sample is computed anyway
// (any position is clamped into the quad table) and replaced by a select.
// AHEAD: the reference value of a sample is read from LDS with the sample's gather, PIPE samples
// before it is used, instead of at its use. The half step's windows take it (one fewer wait on
// LDS per sample; it fits their registers); the evaluation and read-out, held to three waves
// per SIMD, would spill with it.
template <int WR, int WSTEP, bool CHECK, int PIPE, bool AHEAD, class REF>
__device__ __forceinline__ void window_sums(const RefWin<WR, WSTEP>& rw, const REF& ref, int x, int y, const uint32_t* qbase, float qp4,
                                            float koff4, float w1, float h1, const float* H, float& Ss, float& SSs, float& IS) {
    constexpr int WN = Win<WR, WSTEP>::WN;
    Ss = 0.f;
    SSs = 0.f;
    IS = 0.f;
    // A rolling pipeline over the row-major samples: the gather of sample k + PIPE is issued
    // before sample k is consumed, so PIPE gathers stay in flight. The sums take the samples in
    // row-major order as before.
    constexpr int N = WN * WN;
    float fx[N], fy[N], rv[N];
    uint32_t qv[N];
    auto issue = [&](int k) {
        const int a = k / WN, b = k % WN;
        if constexpr (AHEAD) rv[k] = ref.at(x + (-WR + b * WSTEP), y + (-WR + a * WSTEP));
        const float oy = (float)(-WR + a * WSTEP);
        const float ox = (float)(-WR + b * WSTEP);
        const float zz = fmaf(H[6], ox, fmaf(H[7], oy, H[8]));
        const float r = rcp_sample(zz);
        const float u = fmaf(H[0], ox, fmaf(H[1], oy, H[2])) * r, v = fmaf(H[3], ox, fmaf(H[4], oy, H[5])) * r;
        const float xf = floorf(u), yf = floorf(v);
        fx[k] = u - xf;
        fy[k] = v - yf;
        const float xc = __builtin_amdgcn_fmed3f(xf, -2.0f, w1);
        const float yc = __builtin_amdgcn_fmed3f(yf, -2.0f, h1);
        const uint32_t off = (uint32_t)fmaf(yc, qp4, fmaf(xc, 4.0f, koff4));
        qv[k] = *(const uint32_t*)((const char*)qbase + off);
    };
#pragma unroll
    for (int k = 0; k < PIPE && k < N; ++k) issue(k);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k + PIPE < N) issue(k + PIPE);
        const int a = k / WN, b = k % WN;
        const uint32_t q = qv[k];
        const float ax = fx[k], ay = fy[k];
        float sv = quad_lerp(q, ax, ay);
        if (CHECK) sv = fmaf(H[6], (float)(-WR + b * WSTEP), fmaf(H[7], (float)(-WR + a * WSTEP), H[8])) > 1e-8f ? sv : 0.0f;
        float sp = sv - rw.ic;
        float ws = rw.w[k] * sp;
        Ss += ws;
        SSs = fmaf(ws, sp, SSs);
        IS = fmaf(ws, AHEAD ? rv[k] : ref.at(x + (-WR + b * WSTEP), y + (-WR + a * WSTEP)), IS);
    }
}

// Bilateral-weighted NCC cost 1 - rho, 2 on failure (patchmatch.cu:244-290); the covariance is
// sum(w s' I)/W - mu mean(s') with s' the shifted source samples.
template <int WR, int WSTEP, int PIPE = WINDOW_PIPE, bool AHEAD = false, class REF>
__device__ __forceinline__ float ncc_cost(const KParams& P, const RefWin<WR, WSTEP>& rw, const REF& ref, const LevelDesc& S,
                                          const float* H, int x, int y) {
    float z = H[8];
    if (!(z > 1e-8f)) return 2.0f;
    float iz = 1.0f / z;
    float u = H[2] * iz;
    float v = H[5] * iz;
    if (!(u >= 0.f && v >= 0.f && u <= (float)(S.w - 1) && v <= (float)(S.h - 1))) return 2.0f;
    if (rw.var < P.vmin255) return 2.0f;
    const uint32_t* qbase = P.quad + (S.qoff - 2LL * S.qpitch - 2);
    const float qp4 = 4.0f * (float)S.qpitch, koff4 = 2.0f * qp4 + 8.0f;
    const float w1 = (float)S.w, h1 = (float)S.h;
    constexpr float lo = (float)(-WR), hi = (float)Win<WR, WSTEP>::LAST;
    const float za = fmaf(H[7], lo, H[8]), zb = fmaf(H[7], hi, H[8]);
    const bool all_front = fmaf(H[6], lo, za) > 1e-8f && fmaf(H[6], hi, za) > 1e-8f && fmaf(H[6], lo, zb) > 1e-8f &&
                           fmaf(H[6], hi, zb) > 1e-8f;
    float Ss, SSs, IS;
    if (all_front) window_sums<WR, WSTEP, false, PIPE, AHEAD>(rw, ref, x, y, qbase, qp4, koff4, w1, h1, H, Ss, SSs, IS);
    else window_sums<WR, WSTEP, true, PIPE, AHEAD>(rw, ref, x, y, qbase, qp4, koff4, w1, h1, H, Ss, SSs, IS);
    const float iW = rw.iW;
    float mus = Ss * iW;
    float vars = fmaf(-mus, mus, SSs * iW);
    if (vars < P.vmin255) return 2.0f;
    float cov = fmaf(-rw.mu, mus, IS * iW);
    float cc = 1.0f - cov / sqrtf(rw.var * vars);
    return fminf(fmaxf(cc, 0.0f), 2.0f);
}

// ---------------------------------------------------------------- windows of any radius

// Windows too large for register-resident weights (any radius up to 20, step 1 or 2; the kernels
// instantiated with WR = 0 read the radius from KParams::wr): the bilateral weight of a sample is
// recomputed where it is used, as the reference does for windows above 36 samples
// (patchmatch.cu:47,153-165). Every operation and its order match the register-resident path, so
// both give identical results for the same radius.
struct RefWinDyn {
    float ic, mu, W, iW, var;
    int wr;
};

struct RefTileDyn {
    const float* s;
    int x0, y0, pitch;
    __device__ __forceinline__ float at(int x, int y) const { return s[(y - y0) * pitch + (x - x0)]; }
};

template <int WR, int WSTEP>
using RefWinT = typename std::conditional<(WR > 0), RefWin<WR, WSTEP>, RefWinDyn>::type;

__device__ __forceinline__ float bilateral_weight(const KParams& P, float ic, float v, int dx, int dy) {
    float d = ic - v;
    float sp = (float)(dx * dx + dy * dy) * P.inv2s;
    return c3d_expf(-fmaf(d * d, P.inv2c255, sp));
}

template <int WSTEP, class REF>
__device__ void ref_stats_dyn(const KParams& P, const REF& ref, int x, int y, int wr, RefWinDyn& rw) {
    const int wn = (2 * wr) / WSTEP + 1;
    rw.ic = ref.at(x, y);
    rw.wr = wr;
    float W = 0.0f, S = 0.0f;
    for (int a = 0; a < wn; ++a) {
        const int dy = -wr + a * WSTEP;
        for (int b = 0; b < wn; ++b) {
            const int dx = -wr + b * WSTEP;
            float v = ref.at(x + dx, y + dy);
            float w = bilateral_weight(P, rw.ic, v, dx, dy);
            W += w;
            S = fmaf(w, v, S);
        }
    }
    float mu = S / W;
    float V = 0.0f;
    for (int a = 0; a < wn; ++a) {
        const int dy = -wr + a * WSTEP;
        for (int b = 0; b < wn; ++b) {
            const int dx = -wr + b * WSTEP;
            float v = ref.at(x + dx, y + dy);
            float dk = v - mu;
            float wd = bilateral_weight(P, rw.ic, v, dx, dy) * dk;
            V = fmaf(wd, dk, V);
        }
    }
    rw.mu = mu;
    rw.W = W;
    rw.iW = 1.0f / W;
    rw.var = V / W;
}

template <int WSTEP, bool CHUNK, class REF>
__device__ float ncc_cost_dyn(const KParams& P, const RefWinDyn& rw, const REF& ref, const LevelDesc& S, const float* H, int x,
                              int y) {
    float z = H[8];
    if (!(z > 1e-8f)) return 2.0f;
    float iz = 1.0f / z;
    float u = H[2] * iz;
    float v = H[5] * iz;
    if (!(u >= 0.f && v >= 0.f && u <= (float)(S.w - 1) && v <= (float)(S.h - 1))) return 2.0f;
    if (rw.var < P.vmin255) return 2.0f;
    const uint32_t* qbase = P.quad + (S.qoff - 2LL * S.qpitch - 2);
    const float qp4 = 4.0f * (float)S.qpitch, koff4 = 2.0f * qp4 + 8.0f;
    const float w1 = (float)S.w, h1 = (float)S.h;
    const int wr = rw.wr, wn = (2 * wr) / WSTEP + 1;
    float Ss = 0.f, SSs = 0.f, IS = 0.f;
    for (int a = 0; a < wn; ++a) {
        const int dyi = -wr + a * WSTEP;
        const float oy = (float)dyi;
        const float rz = fmaf(H[7], oy, H[8]);
        const float ru = fmaf(H[1], oy, H[2]);
        const float rv = fmaf(H[4], oy, H[5]);
        if constexpr (!CHUNK) {
#pragma unroll 2
            for (int b = 0; b < wn; ++b) {
                const int dxi = -wr + b * WSTEP;
                const float ox = (float)dxi;
                const float zz = fmaf(H[6], ox, rz);
                const float r = rcp_sample(zz);
                float sv = quad_sample(qbase, qp4, koff4, w1, h1, fmaf(H[0], ox, ru) * r, fmaf(H[3], ox, rv) * r);
                sv = zz > 1e-8f ? sv : 0.0f;
                const float iv = ref.at(x + dxi, y + dyi);
                const float sp = sv - rw.ic;
                const float ws = bilateral_weight(P, rw.ic, iv, dxi, dyi) * sp;
                Ss += ws;
                SSs = fmaf(ws, sp, SSs);
                IS = fmaf(ws, iv, IS);
            }
            continue;
        }
        // chunks of four samples: the four gathers are issued before the first is consumed (the
        // samples past the row's end in the last chunk repeat its last sample and are not summed);
        // faster in the geometric builds, slower in the photometric one
        for (int b0 = 0; b0 < wn; b0 += 4) {
            uint32_t qs[4];
            float fxs[4], fys[4], zs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int b = b0 + u < wn ? b0 + u : wn - 1;
                const float ox = (float)(-wr + b * WSTEP);
                const float zz = fmaf(H[6], ox, rz);
                const float r = rcp_sample(zz);
                const float su = fmaf(H[0], ox, ru) * r, sw = fmaf(H[3], ox, rv) * r;
                const float xf = floorf(su), yf = floorf(sw);
                fxs[u] = su - xf;
                fys[u] = sw - yf;
                zs[u] = zz;
                const float xc = __builtin_amdgcn_fmed3f(xf, -2.0f, w1);
                const float yc = __builtin_amdgcn_fmed3f(yf, -2.0f, h1);
                qs[u] = *(const uint32_t*)((const char*)qbase + (uint32_t)fmaf(yc, qp4, fmaf(xc, 4.0f, koff4)));
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int b = b0 + u;
                if (b >= wn) break;
                const int dxi = -wr + b * WSTEP;
                const uint32_t q = qs[u];
                const float ax = fxs[u], ay = fys[u];
                float sv = quad_lerp(q, ax, ay);
                sv = zs[u] > 1e-8f ? sv : 0.0f;
                const float iv = ref.at(x + dxi, y + dyi);
                const float sp = sv - rw.ic;
                const float ws = bilateral_weight(P, rw.ic, iv, dxi, dyi) * sp;
                Ss += ws;
                SSs = fmaf(ws, sp, SSs);
                IS = fmaf(ws, iv, IS);
            }
        }
    }
    const float iW = rw.iW;
    float mus = Ss * iW;
    float vars = fmaf(-mus, mus, SSs * iW);
    if (vars < P.vmin255) return 2.0f;
    float cov = fmaf(-rw.mu, mus, IS * iW);
    float cc = 1.0f - cov / sqrtf(rw.var * vars);
    return fminf(fmaxf(cc, 0.0f), 2.0f);
}

// Dispatch between the register-resident window (WR > 0) and the window of any radius (WR = 0).
template <int WR, int WSTEP, class REF>
__device__ __forceinline__ void ref_stats_any(const KParams& P, const REF& ref, int x, int y, int wr, RefWinT<WR, WSTEP>& rw) {
    if constexpr (WR > 0) ref_stats<WR, WSTEP>(P, ref, x, y, rw);
    else ref_stats_dyn<WSTEP>(P, ref, x, y, wr, rw);
}

template <int WR, int WSTEP, int PIPE = WINDOW_PIPE, bool CHUNK = false, bool AHEAD = false, class REF>
__device__ __forceinline__ float ncc_cost_any(const KParams& P, const RefWinT<WR, WSTEP>& rw, const REF& ref, const LevelDesc& S,
                                              const float* H, int x, int y) {
    if constexpr (WR > 0) return ncc_cost<WR, WSTEP, PIPE, AHEAD>(P, rw, ref, S, H, x, y);
    else return ncc_cost_dyn<WSTEP, CHUNK>(P, rw, ref, S, H, x, y);
}

// Forward-backward reprojection error through the source depth snapshot (patchmatch.cu:292-314).
// geo_error in two halves: geo_fetch projects the pixel at depth d into the source and issues the
// load of the source's snapshot depth at the nearest pixel; geo_finish back-projects it and
// returns the reprojection distance. A caller can put other work between the two, so that the
// load's latency is hidden; geo_finish(geo_fetch(...)) is geo_error exactly.
struct GeoFetch {
    float u, v, dj;
    bool ok;  // false: the result is INF_ERR
};

__device__ __forceinline__ GeoFetch geo_fetch(const KParams& P, const LevelDesc& R, const LevelDesc& S, const PairRec& pr, int x,
                                              int y, float d) {
    GeoFetch g{0.f, 0.f, 0.f, false};
    long long off = P.snap_off[pr.src];
    if (off < 0) return g;
    f3 X = mk3(((float)x - R.cx) / R.fx * d, ((float)y - R.cy) / R.fy * d, d);
    float xj = pr.r[0] * X.x + pr.r[1] * X.y + pr.r[2] * X.z + pr.t[0];
    float yj = pr.r[3] * X.x + pr.r[4] * X.y + pr.r[5] * X.z + pr.t[1];
    float zj = pr.r[6] * X.x + pr.r[7] * X.y + pr.r[8] * X.z + pr.t[2];
    if (!(zj > 1e-8f)) return g;
    float u = S.fx * xj / zj + S.cx, v = S.fy * yj / zj + S.cy;
    float uf = floorf(u + 0.5f), vf = floorf(v + 0.5f);
    if (!(uf >= 0.f && vf >= 0.f && uf < (float)S.w && vf < (float)S.h)) return g;
    g.dj = P.snap[off + (long long)vf * S.w + (long long)uf];
    g.u = u;
    g.v = v;
    g.ok = true;
    return g;
}

__device__ __forceinline__ float geo_finish(const LevelDesc& R, const LevelDesc& S, const PairRec& pr, int x, int y, const GeoFetch& g) {
    if (!g.ok) return INF_ERR;
    const float dj = g.dj, u = g.u, v = g.v;
    if (!(dj > 0.f)) return INF_ERR;
    f3 Yj = mk3((u - S.cx) / S.fx * dj, (v - S.cy) / S.fy * dj, dj);
    f3 q = mk3(Yj.x - pr.t[0], Yj.y - pr.t[1], Yj.z - pr.t[2]);
    float Yx = pr.r[0] * q.x + pr.r[3] * q.y + pr.r[6] * q.z;
    float Yy = pr.r[1] * q.x + pr.r[4] * q.y + pr.r[7] * q.z;
    float Yz = pr.r[2] * q.x + pr.r[5] * q.y + pr.r[8] * q.z;
    if (!(Yz > 1e-8f)) return INF_ERR;
    float ex = R.fx * Yx / Yz + R.cx - (float)x, ey = R.fy * Yy / Yz + R.cy - (float)y;
    return sqrtf(ex * ex + ey * ey);
}

__device__ __forceinline__ float geo_error(const KParams& P, const LevelDesc& R, const LevelDesc& S, const PairRec& pr, int x, int y,
                                           float d) {
    return geo_finish(R, S, pr, x, y, geo_fetch(P, R, S, pr, x, y, d));
}

// (photometric cost, truncated geometric error) of one hypothesis x one source (patchmatch.cu:316-324)
// In geometric runs the snapshot depth is fetched before the NCC window and used after it.
template <int WR, int WSTEP, int PIPE = WINDOW_PIPE, class REF>
__device__ __forceinline__ float pair_cost(const KParams& P, const RefWinT<WR, WSTEP>& rw, const REF& ref, const LevelDesc& R,
                                           const LevelDesc& S, const PairRec& pr, const LevelPair& lp, int x, int y, float4 h,
                                           float* geo) {
    float H[9];
    float c = 2.0f;
    GeoFetch gf{0.f, 0.f, 0.f, false};
    if (P.geometric) gf = geo_fetch(P, R, S, pr, x, y, h.w);
    if (homography(R, lp, 1.0f / R.fx, 1.0f / R.fy, (float)x, (float)y, h.w, xyz(h), H))
        c = ncc_cost_any<WR, WSTEP, PIPE>(P, rw, ref, S, H, x, y);
    if (P.geometric) *geo = fminf(geo_finish(R, S, pr, x, y, gf), P.gmax);
    return c;
}

// ---------------------------------------------------------------- random hypotheses

// patchmatch.cu:347-361
__device__ __forceinline__ f3 random_normal(unsigned long long& st, f3 ray) {
    f3 n = mk3(0.f, 0.f, -1.f);
    for (int k = 0; k < 64; ++k) {
        float v1 = 2.f * rnd(st) - 1.f;
        float v2 = 2.f * rnd(st) - 1.f;
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

// patchmatch.cu:363-366
__device__ __forceinline__ float random_depth(unsigned long long& st, float dmin, float dmax) {
    float a = 1.f / dmax, b = 1.f / dmin;
    return 1.f / (a + rnd(st) * (b - a));
}

// patchmatch.cu:368-384
__device__ __forceinline__ f3 perturb_normal(unsigned long long& st, f3 n, f3 ray, float phi) {
    f3 a = random_normal(st, mk3(0.f, 0.f, 0.f));
    f3 ax = cross3(a, n);
    float l2 = dot3(ax, ax);
    float ang = rnd(st) * phi;
    if (l2 < 1e-12f) return n;
    ax = nrm3(ax);
    f3 bn = cross3(ax, n);
    for (int k = 0; k < 4; ++k) {
        float sa, ca;
        c3d_sincosf(ang, sa, ca);
        f3 m = mk3(n.x * ca + bn.x * sa, n.y * ca + bn.y * sa, n.z * ca + bn.z * sa);
        if (dot3(m, ray) < 0.f) return nrm3(m);
        ang *= 0.5f;
    }
    return n;
}

// triangulation x incidence x resolution prior (patchmatch.cu:388-422); r = window radius
__device__ __forceinline__ float view_prior(const KParams& P, const LevelDesc& R, const PairRec& pr, const float* H, bool hok, int x,
                                            int y, float4 h, float r) {
    if (!hok) return 0.f;
    f3 ray = pixel_ray(R, x, y);
    f3 X = mk3(ray.x * h.w, ray.y * h.w, h.w);
    f3 S = mk3(pr.c[0] - X.x, pr.c[1] - X.y, pr.c[2] - X.z);
    float ls = sqrtf(dot3(S, S)), lx = sqrtf(dot3(X, X));
    if (ls <= 0.f || lx <= 0.f) return 0.f;
    float cost = -dot3(S, X) / (ls * lx);
    float tri = 1.f;
    if (cost > P.cos_tri_min) {
        float s = 1.f - (1.f - cost) / (1.f - P.cos_tri_min);
        tri = fminf(fmaxf(1.f - s * s, 0.f), 1.f);
    }
    float cphi = dot3(S, xyz(h)) / ls;
    float xi = 1.f - fmaxf(0.f, cphi);
    float inc = c3d_expf(-xi * xi * P.inv2inc);
    const float cxs[4] = {-r, -r, r, r};
    const float cys[4] = {-r, r, r, -r};
    float px[4], py[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float z = H[6] * cxs[k] + H[7] * cys[k] + H[8];
        if (!(z > 1e-8f)) return 0.f;
        px[k] = (H[0] * cxs[k] + H[1] * cys[k] + H[2]) / z;
        py[k] = (H[3] * cxs[k] + H[4] * cys[k] + H[5]) / z;
    }
    float area = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) area += px[k] * py[(k + 1) & 3] - px[(k + 1) & 3] * py[k];
    area = 0.5f * fabsf(area);
    float aref = (2.f * r + 1.f) * (2.f * r + 1.f);
    if (!(area > 0.f)) return 0.f;
    float res = fminf(area / aref, aref / area);
    return tri * inc * res;
}

}  // namespace c3d
