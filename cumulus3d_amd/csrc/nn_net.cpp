// nn_net.cpp — the learned initialiser's U-Net on the device: parameters, activations, the
// forward and backward schedules over the kernels of nn_kernels.hip, Adam, and the C ABI of
// include/c3d_init.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/c3d_init.h"
#include "backend.hpp"
#include "nn_kernels.hpp"
#include "nn_net.hpp"

using namespace c3d;
using namespace c3d::nn;

namespace {

#define NN_TRY(expr)                                                                                              \
    do {                                                                                                          \
        hipError_t e_ = (expr);                                                                                   \
        if (e_ != hipSuccess) return set_error(C3D_DEVICE, std::string("HIP: ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Layer {
    int cin, cout, k, relu;
    int cin_real;  // input channels that carry data (fan-in of the initialisation)
    size_t woff, boff;
    uint16_t* w = nullptr;   // [cout][k*k][cin] bf16
    uint16_t* wt = nullptr;  // [cin][k*k][cout] bf16, taps flipped (data gradient)
};

enum { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2, D1, D0, HEAD, NLAYER };

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double gauss(uint64_t& s) {
    double u1 = ((splitmix(s) >> 11) + 1) * (1.0 / 9007199254740993.0);
    double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace

struct c3d_init_net {
    c3d_pm_backend* be = nullptr;
    int B = 0, H = 0, W = 0;
    Layer L[NLAYER];
    size_t nparam = 0;
    float *w32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;
    // activations (bf16) and their gradients (fp32)
    uint16_t *feat, *e0a, *e0b, *p1, *e1a, *e1b, *p2, *e2a, *e2b, *p3, *e3a, *e3b, *cat2, *d2, *cat1, *d1, *cat0, *d0, *dout, *dz;
    float* out32;
    float *g_d0, *g_cat0, *g_d1, *g_cat1, *g_d2, *g_cat2, *g_e3b, *g_e3a, *g_p3, *g_e2b, *g_e2a, *g_p2, *g_e1b, *g_e1a, *g_p1, *g_e0b,
        *g_e0a;
    float *label, *range, *sums;
    std::vector<float> h_range;
    std::vector<int> h_valid;  // per slot: valid width, height
    int adam_t = 0;
    std::vector<void*> allocs;
};

namespace {

template <class T>
hipError_t alloc(c3d_init_net* n, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
        n->allocs.push_back(q);
        *p = (T*)q;
    }
    return e;
}

int upload_weights(c3d_init_net* n) {
    hipStream_t s = n->be->stream;
    for (int i = 0; i < NLAYER; ++i) {
        const Layer& l = n->L[i];
        NN_TRY(weights_bf16(n->w32 + l.woff, l.cout, l.k * l.k, l.cin, l.w, l.wt, s));
    }
    return C3D_OK;
}

int conv_fwd(c3d_init_net* n, int li, const uint16_t* x, int nimg, int h, int w, uint16_t* y, float* y32) {
    const Layer& l = n->L[li];
    ConvArgs a{};
    a.x = x;
    a.xs = l.cin;
    a.cin = l.cin;
    a.nimg = nimg;
    a.H = h;
    a.W = w;
    a.ksize = l.k;
    a.w = l.w;
    a.bias = n->w32 + l.boff;
    a.cout = l.cout;
    a.y = y;
    a.ys = l.cout;
    a.yoff = 0;
    a.y32 = y32;
    a.y32s = l.cout;
    a.mode = y32 ? 1 : 0;
    a.relu = l.relu;
    NN_TRY(conv(a, n->be->stream));
    return C3D_OK;
}

int forward(c3d_init_net* n, int b) {
    hipStream_t s = n->be->stream;
    const int H = n->H, W = n->W;
    int rc;
    if ((rc = conv_fwd(n, E0A, n->feat, b, H, W, n->e0a, nullptr))) return rc;
    if ((rc = conv_fwd(n, E0B, n->e0a, b, H, W, n->e0b, nullptr))) return rc;
    NN_TRY(pool(n->e0b, 32, 32, b, H, W, n->p1, 32, s));
    if ((rc = conv_fwd(n, E1A, n->p1, b, H / 2, W / 2, n->e1a, nullptr))) return rc;
    if ((rc = conv_fwd(n, E1B, n->e1a, b, H / 2, W / 2, n->e1b, nullptr))) return rc;
    NN_TRY(pool(n->e1b, 64, 64, b, H / 2, W / 2, n->p2, 64, s));
    if ((rc = conv_fwd(n, E2A, n->p2, b, H / 4, W / 4, n->e2a, nullptr))) return rc;
    if ((rc = conv_fwd(n, E2B, n->e2a, b, H / 4, W / 4, n->e2b, nullptr))) return rc;
    NN_TRY(pool(n->e2b, 128, 128, b, H / 4, W / 4, n->p3, 128, s));
    if ((rc = conv_fwd(n, E3A, n->p3, b, H / 8, W / 8, n->e3a, nullptr))) return rc;
    if ((rc = conv_fwd(n, E3B, n->e3a, b, H / 8, W / 8, n->e3b, nullptr))) return rc;
    const long long P0 = (long long)b * H * W, P1 = P0 / 4, P2 = P0 / 16;
    NN_TRY(upsample(n->e3b, 128, 128, b, H / 8, W / 8, n->cat2, 256, 0, s));
    NN_TRY(copy_channels(n->e2b, 128, 128, P2, n->cat2, 256, 128, s));
    if ((rc = conv_fwd(n, D2, n->cat2, b, H / 4, W / 4, n->d2, nullptr))) return rc;
    NN_TRY(upsample(n->d2, 128, 128, b, H / 4, W / 4, n->cat1, 192, 0, s));
    NN_TRY(copy_channels(n->e1b, 64, 64, P1, n->cat1, 192, 128, s));
    if ((rc = conv_fwd(n, D1, n->cat1, b, H / 2, W / 2, n->d1, nullptr))) return rc;
    NN_TRY(upsample(n->d1, 64, 64, b, H / 2, W / 2, n->cat0, 96, 0, s));
    NN_TRY(copy_channels(n->e0b, 32, 32, P0, n->cat0, 96, 64, s));
    if ((rc = conv_fwd(n, D0, n->cat0, b, H, W, n->d0, nullptr))) return rc;
    if ((rc = conv_fwd(n, HEAD, n->d0, b, H, W, nullptr, n->out32))) return rc;
    return C3D_OK;
}

// Gradient of a layer from the gradient at its (pre-activation) output dz: weights, bias, and
// (when dx is given) the input gradient, stored.
int layer_bwd(c3d_init_net* n, int li, const uint16_t* x, const uint16_t* dz, int nimg, int h, int w, float* dx) {
    const Layer& l = n->L[li];
    hipStream_t s = n->be->stream;
    WgradArgs g{};
    g.x = x;
    g.xs = l.cin;
    g.cin = l.cin;
    g.dz = dz;
    g.dzs = l.cout;
    g.cout = l.cout;
    g.nimg = nimg;
    g.H = h;
    g.W = w;
    g.ksize = l.k;
    g.dw = n->g32 + l.woff;
    NN_TRY(wgrad(g, s));
    const long long np = (long long)nimg * h * w;
    NN_TRY(bias_grad(dz, l.cout, l.cout, np, n->g32 + l.boff, s));
    if (dx) {
        ConvArgs a{};
        a.x = dz;
        a.xs = l.cout;
        a.cin = l.cout;
        a.nimg = nimg;
        a.H = h;
        a.W = w;
        a.ksize = l.k;
        a.w = l.wt;
        a.bias = nullptr;
        a.cout = l.cin;
        a.y32 = dx;
        a.y32s = l.cin;
        a.mode = 1;
        NN_TRY(conv(a, s));
    }
    return C3D_OK;
}

// dz = relu'(a) * da, then the layer's gradients
int relu_layer_bwd(c3d_init_net* n, int li, const uint16_t* x, const uint16_t* a, const float* da, int nimg, int h, int w, float* dx) {
    const Layer& l = n->L[li];
    const long long np = (long long)nimg * h * w;
    NN_TRY(relu_grad(da, l.cout, a, l.cout, n->dz, l.cout, l.cout, np, n->be->stream));
    return layer_bwd(n, li, x, n->dz, nimg, h, w, dx);
}

int backward(c3d_init_net* n, int b) {
    hipStream_t s = n->be->stream;
    const int H = n->H, W = n->W;
    const long long P0 = (long long)b * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    int rc;
    NN_TRY(fill(n->g32, 0.f, (long long)n->nparam, s));
    NN_TRY(fill(n->g_d1, 0.f, P1 * 64, s));
    NN_TRY(fill(n->g_d2, 0.f, P2 * 128, s));
    NN_TRY(fill(n->g_e3b, 0.f, P3 * 128, s));
    NN_TRY(fill(n->g_e2b, 0.f, P2 * 128, s));
    NN_TRY(fill(n->g_e1b, 0.f, P1 * 64, s));
    NN_TRY(fill(n->g_e0b, 0.f, P0 * 32, s));
    // head (no activation): dz = dout
    if ((rc = layer_bwd(n, HEAD, n->d0, n->dout, b, H, W, n->g_d0))) return rc;
    if ((rc = relu_layer_bwd(n, D0, n->cat0, n->d0, n->g_d0, b, H, W, n->g_cat0))) return rc;
    NN_TRY(upsample_grad(n->g_cat0, 96, 0, 64, b, H / 2, W / 2, n->g_d1, 64, s));
    NN_TRY(add_channels(n->g_cat0, 96, 64, 32, P0, n->g_e0b, 32, s));
    if ((rc = relu_layer_bwd(n, D1, n->cat1, n->d1, n->g_d1, b, H / 2, W / 2, n->g_cat1))) return rc;
    NN_TRY(upsample_grad(n->g_cat1, 192, 0, 128, b, H / 4, W / 4, n->g_d2, 128, s));
    NN_TRY(add_channels(n->g_cat1, 192, 128, 64, P1, n->g_e1b, 64, s));
    if ((rc = relu_layer_bwd(n, D2, n->cat2, n->d2, n->g_d2, b, H / 4, W / 4, n->g_cat2))) return rc;
    NN_TRY(upsample_grad(n->g_cat2, 256, 0, 128, b, H / 8, W / 8, n->g_e3b, 128, s));
    NN_TRY(add_channels(n->g_cat2, 256, 128, 128, P2, n->g_e2b, 128, s));
    if ((rc = relu_layer_bwd(n, E3B, n->e3a, n->e3b, n->g_e3b, b, H / 8, W / 8, n->g_e3a))) return rc;
    if ((rc = relu_layer_bwd(n, E3A, n->p3, n->e3a, n->g_e3a, b, H / 8, W / 8, n->g_p3))) return rc;
    NN_TRY(pool_grad(n->g_p3, 128, 128, b, H / 4, W / 4, n->g_e2b, 128, s));
    if ((rc = relu_layer_bwd(n, E2B, n->e2a, n->e2b, n->g_e2b, b, H / 4, W / 4, n->g_e2a))) return rc;
    if ((rc = relu_layer_bwd(n, E2A, n->p2, n->e2a, n->g_e2a, b, H / 4, W / 4, n->g_p2))) return rc;
    NN_TRY(pool_grad(n->g_p2, 64, 64, b, H / 2, W / 2, n->g_e1b, 64, s));
    if ((rc = relu_layer_bwd(n, E1B, n->e1a, n->e1b, n->g_e1b, b, H / 2, W / 2, n->g_e1a))) return rc;
    if ((rc = relu_layer_bwd(n, E1A, n->p1, n->e1a, n->g_e1a, b, H / 2, W / 2, n->g_p1))) return rc;
    NN_TRY(pool_grad(n->g_p1, 32, 32, b, H, W, n->g_e0b, 32, s));
    if ((rc = relu_layer_bwd(n, E0B, n->e0a, n->e0b, n->g_e0b, b, H, W, n->g_e0a))) return rc;
    if ((rc = relu_layer_bwd(n, E0A, n->feat, n->e0a, n->g_e0a, b, H, W, nullptr))) return rc;
    return C3D_OK;
}

int check_net(c3d_init_net* n) {
    if (!n) return set_error(C3D_INVALID_ARGUMENT, "null network");
    NN_TRY(hipSetDevice(n->be->device));
    return C3D_OK;
}

}  // namespace

namespace c3d {
namespace nn {

NetView net_view(const c3d_init_net* n) { return NetView{n->be, n->B, n->H, n->W}; }
uint16_t* net_features(c3d_init_net* n, int slot) { return n->feat + (size_t)slot * n->H * n->W * 64; }
void net_set_range(c3d_init_net* n, int slot, float dmin, float dmax) {
    n->h_range[2 * slot] = dmin;
    n->h_range[2 * slot + 1] = dmax;
}
int net_forward(c3d_init_net* n, int batch) { return forward(n, batch); }
const float* net_output(const c3d_init_net* n) { return n->out32; }

}  // namespace nn
}  // namespace c3d

extern "C" {

int c3d_init_create(c3d_pm_backend* be, uint32_t max_batch, uint32_t height, uint32_t width, uint64_t seed, c3d_init_net** out) {
    if (!be || !out) return set_error(C3D_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (max_batch < 1 || height < 8 || width < 8 || height % 8 || width % 8)
        return set_error(C3D_INVALID_ARGUMENT, "batch >= 1 and height, width positive multiples of 8 required");
    NN_TRY(hipSetDevice(be->device));
    const int hold = c3d::backend_hold(be);
    if (hold) return hold;
    c3d_init_net* n = new c3d_init_net();
    n->be = be;
    n->B = (int)max_batch;
    n->H = (int)height;
    n->W = (int)width;
    // (cin, cout, k, relu, cin_real)
    const int spec[NLAYER][5] = {{64, 32, 3, 1, 33},   {32, 32, 3, 1, 32},   {32, 64, 3, 1, 32},   {64, 64, 3, 1, 64},
                                 {64, 128, 3, 1, 64},  {128, 128, 3, 1, 128}, {128, 128, 3, 1, 128}, {128, 128, 3, 1, 128},
                                 {256, 128, 3, 1, 256}, {192, 64, 3, 1, 192}, {96, 32, 3, 1, 96},    {32, 32, 1, 0, 32}};
    size_t off = 0;
    for (int i = 0; i < NLAYER; ++i) {
        Layer& l = n->L[i];
        l.cin = spec[i][0];
        l.cout = spec[i][1];
        l.k = spec[i][2];
        l.relu = spec[i][3];
        l.cin_real = spec[i][4];
        l.woff = off;
        off += (size_t)l.cout * l.k * l.k * l.cin;
        l.boff = off;
        off += (size_t)l.cout;
    }
    n->nparam = off;
    auto bail = [&](hipError_t e) {
        c3d_init_destroy(n);
        return set_error(C3D_DEVICE, std::string("HIP: out of device memory for the network: ") + hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
#define A(ptr, cnt)                                   \
    if ((e = alloc(n, &(ptr), (size_t)(cnt))) != hipSuccess) return bail(e);
    A(n->w32, off);
    A(n->g32, off);
    A(n->m32, off);
    A(n->v32, off);
    for (int i = 0; i < NLAYER; ++i) {
        size_t c = (size_t)n->L[i].cout * n->L[i].k * n->L[i].k * n->L[i].cin;
        A(n->L[i].w, c);
        A(n->L[i].wt, c);
    }
    const size_t P0 = (size_t)n->B * n->H * n->W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    A(n->feat, P0 * 64);
    A(n->e0a, P0 * 32);
    A(n->e0b, P0 * 32);
    A(n->p1, P1 * 32);
    A(n->e1a, P1 * 64);
    A(n->e1b, P1 * 64);
    A(n->p2, P2 * 64);
    A(n->e2a, P2 * 128);
    A(n->e2b, P2 * 128);
    A(n->p3, P3 * 128);
    A(n->e3a, P3 * 128);
    A(n->e3b, P3 * 128);
    A(n->cat2, P2 * 256);
    A(n->d2, P2 * 128);
    A(n->cat1, P1 * 192);
    A(n->d1, P1 * 64);
    A(n->cat0, P0 * 96);
    A(n->d0, P0 * 32);
    A(n->dout, P0 * 32);
    A(n->dz, P0 * 32);
    A(n->out32, P0 * 32);
    A(n->g_d0, P0 * 32);
    A(n->g_cat0, P0 * 96);
    A(n->g_d1, P1 * 64);
    A(n->g_cat1, P1 * 192);
    A(n->g_d2, P2 * 128);
    A(n->g_cat2, P2 * 256);
    A(n->g_e3b, P3 * 128);
    A(n->g_e3a, P3 * 128);
    A(n->g_p3, P3 * 128);
    A(n->g_e2b, P2 * 128);
    A(n->g_e2a, P2 * 128);
    A(n->g_p2, P2 * 64);
    A(n->g_e1b, P1 * 64);
    A(n->g_e1a, P1 * 64);
    A(n->g_p1, P1 * 32);
    A(n->g_e0b, P0 * 32);
    A(n->g_e0a, P0 * 32);
    A(n->label, P0 * 4);
    A(n->range, (size_t)n->B * 2);
    A(n->sums, 8);
#undef A
    n->h_range.assign((size_t)n->B * 2, 1.0f);
    n->h_valid.assign((size_t)n->B * 2, 0);
    // He-normal initialisation over the data-carrying fan-in; the head starts small and its
    // unused outputs 5..31 at zero
    std::vector<float> w(off, 0.0f);
    uint64_t st = seed;
    for (int i = 0; i < NLAYER; ++i) {
        const Layer& l = n->L[i];
        const int taps = l.k * l.k;
        const double sd = (i == HEAD ? 0.1 : 1.0) * std::sqrt(2.0 / (double)(taps * l.cin_real));
        for (int co = 0; co < l.cout; ++co)
            for (int t = 0; t < taps; ++t)
                for (int ci = 0; ci < l.cin; ++ci) {
                    const double g = gauss(st);
                    const bool live = ci < l.cin_real && (i != HEAD || co < 5);
                    w[l.woff + ((size_t)co * taps + t) * l.cin + ci] = live ? (float)(g * sd) : 0.0f;
                }
    }
    hipStream_t s = be->stream;
    if ((e = hipMemcpyAsync(n->w32, w.data(), off * sizeof(float), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(e);
    if ((e = fill(n->m32, 0.f, (long long)off, s)) != hipSuccess) return bail(e);
    if ((e = fill(n->v32, 0.f, (long long)off, s)) != hipSuccess) return bail(e);
    if ((e = fill(n->label, 0.f, (long long)P0 * 4, s)) != hipSuccess) return bail(e);
    if ((e = hipMemsetAsync(n->feat, 0, P0 * 64 * sizeof(uint16_t), s)) != hipSuccess) return bail(e);
    if ((e = hipMemsetAsync(n->dout, 0, P0 * 32 * sizeof(uint16_t), s)) != hipSuccess) return bail(e);
    int rc = upload_weights(n);
    if (rc) {
        c3d_init_destroy(n);
        return rc;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return bail(e);
    *out = n;
    return C3D_OK;
}

void c3d_init_destroy(c3d_init_net* n) {
    if (!n) return;
    c3d_pm_backend* be = n->be;
    (void)hipSetDevice(be->device);
    (void)hipStreamSynchronize(be->stream);
    for (void* p : n->allocs) (void)hipFree(p);
    delete n;
    c3d::backend_release(be);
}

uint64_t c3d_init_num_params(const c3d_init_net* n) { return n ? (uint64_t)n->nparam : 0; }

int c3d_init_get_params(c3d_init_net* n, float* out) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!out) return set_error(C3D_INVALID_ARGUMENT, "null output");
    NN_TRY(hipMemcpyAsync(out, n->w32, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->be->stream));
    NN_TRY(hipStreamSynchronize(n->be->stream));
    return C3D_OK;
}

int c3d_init_set_params(c3d_init_net* n, const float* in) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!in) return set_error(C3D_INVALID_ARGUMENT, "null input");
    NN_TRY(hipMemcpyAsync(n->w32, in, n->nparam * sizeof(float), hipMemcpyHostToDevice, n->be->stream));
    NN_TRY(fill(n->m32, 0.f, (long long)n->nparam, n->be->stream));
    NN_TRY(fill(n->v32, 0.f, (long long)n->nparam, n->be->stream));
    n->adam_t = 0;
    rc = upload_weights(n);
    if (rc) return rc;
    NN_TRY(hipStreamSynchronize(n->be->stream));
    return C3D_OK;
}

int c3d_init_features(c3d_init_net* n, uint32_t slot0, uint32_t num, const c3d_sweep_view* views) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!views || slot0 + num > (uint32_t)n->B) return set_error(C3D_INVALID_ARGUMENT, "slots out of range");
    hipStream_t s = n->be->stream;
    for (uint32_t i = 0; i < num; ++i) {
        const c3d_sweep_view& v = views[i];
        if (!v.gray || v.width > (uint32_t)n->W || v.height > (uint32_t)n->H || v.num_sources < 1 || v.num_sources > 2)
            return set_error(C3D_INVALID_ARGUMENT, "sweep view larger than the network input or without 1..2 sources");
        if (!(v.depth_min > 0.f && v.depth_max > v.depth_min)) return set_error(C3D_INVALID_ARGUMENT, "bad depth range");
        std::vector<size_t> sz = {(size_t)v.width * v.height};
        for (uint32_t j = 0; j < v.num_sources; ++j) sz.push_back((size_t)v.sources[j].width * v.sources[j].height);
        size_t tot = 0;
        for (size_t z : sz) tot += (z + 255) & ~(size_t)255;
        uint8_t* img = nullptr;
        NN_TRY(hipMalloc(&img, tot));
        size_t o = 0;
        SweepArgs a{};
        hipError_t e = hipMemcpyAsync(img, v.gray, sz[0], hipMemcpyHostToDevice, s);
        a.ref = img;
        o += (sz[0] + 255) & ~(size_t)255;
        for (uint32_t j = 0; e == hipSuccess && j < v.num_sources; ++j) {
            const c3d_sweep_source& src = v.sources[j];
            e = hipMemcpyAsync(img + o, src.gray, sz[j + 1], hipMemcpyHostToDevice, s);
            a.src[j] = img + o;
            o += (sz[j + 1] + 255) & ~(size_t)255;
            a.sw[j] = (int)src.width;
            a.sh[j] = (int)src.height;
            std::memcpy(a.sk[j], src.k, sizeof a.sk[j]);
            std::memcpy(a.r[j], src.r, sizeof a.r[j]);
            std::memcpy(a.t[j], src.t, sizeof a.t[j]);
        }
        a.rw = (int)v.width;
        a.rh = (int)v.height;
        std::memcpy(a.rk, v.k, sizeof a.rk);
        a.nsrc = (int)v.num_sources;
        a.dmin = v.depth_min;
        a.dmax = v.depth_max;
        a.feat = n->feat + (size_t)(slot0 + i) * n->H * n->W * 64;
        a.H = n->H;
        a.W = n->W;
        if (e == hipSuccess) e = sweep(a, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(img);
        if (e != hipSuccess) return set_error(C3D_DEVICE, std::string("HIP: sweep: ") + hipGetErrorString(e));
        n->h_range[2 * (slot0 + i)] = v.depth_min;
        n->h_range[2 * (slot0 + i) + 1] = v.depth_max;
        n->h_valid[2 * (slot0 + i)] = (int)v.width;
        n->h_valid[2 * (slot0 + i) + 1] = (int)v.height;
    }
    return C3D_OK;
}

int c3d_init_set_features(c3d_init_net* n, uint32_t num, const uint16_t* feat) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!feat || num > (uint32_t)n->B) return set_error(C3D_INVALID_ARGUMENT, "slots out of range");
    NN_TRY(hipMemcpyAsync(n->feat, feat, (size_t)num * n->H * n->W * 64 * sizeof(uint16_t), hipMemcpyHostToDevice, n->be->stream));
    NN_TRY(hipStreamSynchronize(n->be->stream));
    return C3D_OK;
}

int c3d_init_get_features(c3d_init_net* n, uint32_t slot, uint16_t* feat) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!feat || slot >= (uint32_t)n->B) return set_error(C3D_INVALID_ARGUMENT, "slot out of range");
    const size_t cnt = (size_t)n->H * n->W * 64;
    NN_TRY(hipMemcpyAsync(feat, n->feat + slot * cnt, cnt * sizeof(uint16_t), hipMemcpyDeviceToHost, n->be->stream));
    NN_TRY(hipStreamSynchronize(n->be->stream));
    return C3D_OK;
}

int c3d_init_set_labels(c3d_init_net* n, uint32_t slot, const float* depth, const float* normal, uint32_t vw, uint32_t vh, float dmin,
                        float dmax) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!depth || !normal || slot >= (uint32_t)n->B || vw > (uint32_t)n->W || vh > (uint32_t)n->H)
        return set_error(C3D_INVALID_ARGUMENT, "bad label arguments");
    if (!(dmin > 0.f && dmax > dmin)) return set_error(C3D_INVALID_ARGUMENT, "bad depth range");
    const size_t np = (size_t)n->H * n->W;
    std::vector<float> lb(np * 4);
    for (size_t p = 0; p < np; ++p) {
        lb[4 * p] = depth[p];
        lb[4 * p + 1] = normal[3 * p];
        lb[4 * p + 2] = normal[3 * p + 1];
        lb[4 * p + 3] = normal[3 * p + 2];
    }
    NN_TRY(hipMemcpyAsync(n->label + slot * np * 4, lb.data(), np * 4 * sizeof(float), hipMemcpyHostToDevice, n->be->stream));
    NN_TRY(hipStreamSynchronize(n->be->stream));
    n->h_range[2 * slot] = dmin;
    n->h_range[2 * slot + 1] = dmax;
    n->h_valid[2 * slot] = (int)vw;
    n->h_valid[2 * slot + 1] = (int)vh;
    return C3D_OK;
}

int c3d_init_train_step(c3d_init_net* n, uint32_t batch, const c3d_init_train_params* tp, int step, float* losses, float* grads) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!tp || batch < 1 || batch > (uint32_t)n->B) return set_error(C3D_INVALID_ARGUMENT, "batch out of range");
    for (uint32_t i = 1; i < batch; ++i)
        if (n->h_valid[2 * i] != n->h_valid[0] || n->h_valid[2 * i + 1] != n->h_valid[1])
            return set_error(C3D_INVALID_ARGUMENT, "every slot of a batch needs the same valid size");
    hipStream_t s = n->be->stream;
    NN_TRY(hipMemcpyAsync(n->range, n->h_range.data(), (size_t)batch * 2 * sizeof(float), hipMemcpyHostToDevice, s));
    rc = forward(n, (int)batch);
    if (rc) return rc;
    NN_TRY(fill(n->sums, 0.f, 8, s));
    LossArgs la{};
    la.out = n->out32;
    la.label = n->label;
    la.range = n->range;
    la.nimg = (int)batch;
    la.H = n->H;
    la.W = n->W;
    la.Hv = n->h_valid[1];
    la.Wv = n->h_valid[0];
    la.w_normal = tp->w_normal;
    la.w_conf = tp->w_conf;
    la.conf_tol = tp->conf_tol;
    la.dout = n->dout;
    la.sums = n->sums;
    NN_TRY(loss(la, s));  // channels 5..31 of dout stay zero from c3d_init_create
    rc = backward(n, (int)batch);
    if (rc) return rc;
    if (step) {
        n->adam_t++;
        const float b1t = 1.f - std::pow(tp->beta1, (float)n->adam_t), b2t = 1.f - std::pow(tp->beta2, (float)n->adam_t);
        NN_TRY(adam(n->w32, n->g32, n->m32, n->v32, (long long)n->nparam, tp->lr, tp->beta1, tp->beta2, tp->eps, b1t, b2t, n->sums + 3,
                    s));
        rc = upload_weights(n);
        if (rc) return rc;
    }
    if (losses) NN_TRY(hipMemcpyAsync(losses, n->sums, 4 * sizeof(float), hipMemcpyDeviceToHost, s));
    if (grads) NN_TRY(hipMemcpyAsync(grads, n->g32, n->nparam * sizeof(float), hipMemcpyDeviceToHost, s));
    NN_TRY(hipStreamSynchronize(s));
    return C3D_OK;
}

int c3d_init_predict(c3d_init_net* n, uint32_t batch, float* depth, float* normal, float* conf, float* head) {
    int rc = check_net(n);
    if (rc) return rc;
    if (batch < 1 || batch > (uint32_t)n->B) return set_error(C3D_INVALID_ARGUMENT, "batch out of range");
    rc = forward(n, (int)batch);
    if (rc) return rc;
    const size_t np = (size_t)batch * n->H * n->W;
    std::vector<float> o(np * 32);
    NN_TRY(hipMemcpyAsync(o.data(), n->out32, np * 32 * sizeof(float), hipMemcpyDeviceToHost, n->be->stream));
    NN_TRY(hipStreamSynchronize(n->be->stream));
    const size_t per = (size_t)n->H * n->W;
    for (size_t p = 0; p < np; ++p) {
        const float* q = &o[p * 32];
        const size_t b = p / per;
        const float ia = 1.f / n->h_range[2 * b + 1], ib = 1.f / n->h_range[2 * b];
        const float s0 = 1.f / (1.f + std::exp(-q[0]));
        if (depth) depth[p] = 1.f / (ia + s0 * (ib - ia));
        if (normal) {
            const float nv = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + 1e-6f);
            normal[3 * p] = q[1] / nv;
            normal[3 * p + 1] = q[2] / nv;
            normal[3 * p + 2] = q[3] / nv;
        }
        if (conf) conf[p] = 1.f / (1.f + std::exp(-q[4]));
        if (head)
            for (int k = 0; k < 5; ++k) head[p * 5 + k] = q[k];
    }
    return C3D_OK;
}

}  // extern "C"
