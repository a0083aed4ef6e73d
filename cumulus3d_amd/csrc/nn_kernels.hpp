// nn_kernels.hpp — host-side interface of the learned-initialiser kernels (nn_kernels.hip).
//
// Tensors are NHWC: element (image b, row y, column x, channel c) of a tensor with channel stride
// cs lives at ((b * H + y) * W + x) * cs + c. Activations and weights are bf16 (stored as uint16),
// gradients of activations and every parameter-side buffer are fp32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace c3d {
namespace nn {

// Convolution (3x3 or 1x1, stride 1, zero padding) as an implicit GEMM on MFMA:
// y[p][co] = sum_{tap, ci} x[p + off(tap)][ci] * w[co][tap][ci].
struct ConvArgs {
    const uint16_t* x;
    int xs, cin;  // channel stride and count of the input (cin a multiple of 32)
    int nimg, H, W, ksize;
    const uint16_t* w;  // [cout][ksize * ksize][cin]
    const float* bias;  // [cout] or null
    int cout;           // a multiple of 16
    uint16_t* y;        // mode 0: bf16 output at channel offset yoff of a tensor with stride ys
    int ys, yoff;
    float* y32;  // modes 1 and 2: fp32 output with stride y32s (1: store, 2: accumulate)
    int y32s;
    int mode, relu;
};

// Weight gradient dw[co][tap][ci] += sum_p dz[p][co] * x[p + off(tap)][ci] (fp32 atomics).
struct WgradArgs {
    const uint16_t* x;
    int xs, cin;
    const uint16_t* dz;
    int dzs, cout;
    int nimg, H, W, ksize;
    float* dw;
};

hipError_t conv(const ConvArgs& a, hipStream_t s);
hipError_t wgrad(const WgradArgs& a, hipStream_t s);
// db[c] += sum_p dz[p][c]
hipError_t bias_grad(const uint16_t* dz, int dzs, int c, long long npix, float* db, hipStream_t s);
// dz = bf16(da * (a > 0)) over c channels
hipError_t relu_grad(const float* da, int das, const uint16_t* a, int as, uint16_t* dz, int dzs, int c, long long npix,
                     hipStream_t s);
// 2x2 mean pooling (H, W even) and its gradient (accumulated)
hipError_t pool(const uint16_t* x, int xs, int c, int nimg, int H, int W, uint16_t* y, int ys, hipStream_t s);
hipError_t pool_grad(const float* dy, int dys, int c, int nimg, int H, int W, float* dx, int dxs, hipStream_t s);
// nearest 2x upsampling of x (nimg, H, W) into channels [yoff, yoff + c) of y (nimg, 2H, 2W)
hipError_t upsample(const uint16_t* x, int xs, int c, int nimg, int H, int W, uint16_t* y, int ys, int yoff, hipStream_t s);
hipError_t upsample_grad(const float* dy, int dys, int yoff, int c, int nimg, int H, int W, float* dx, int dxs, hipStream_t s);
// y[:, yoff : yoff + c] = x ; dx += dy[:, off : off + c]
hipError_t copy_channels(const uint16_t* x, int xs, int c, long long npix, uint16_t* y, int ys, int yoff, hipStream_t s);
hipError_t add_channels(const float* dy, int dys, int off, int c, long long npix, float* dx, int dxs, hipStream_t s);
// fp32 fill
hipError_t fill(float* p, float v, long long n, hipStream_t s);
// weight layout transforms: bf16 forward copy and the flipped, transposed data-gradient copy
hipError_t weights_bf16(const float* w32, int cout, int taps, int cin, uint16_t* w, uint16_t* wt, hipStream_t s);

// Loss of the 5 head outputs against plain-PatchStereo labels, and its gradient (bf16, stride 32).
struct LossArgs {
    const float* out;  // [p][32] head output (channels 0..4 used)
    const float* label;  // [p][4] (depth, nx, ny, nz); depth 0 = no label
    const float* range;  // [image][2] (depth_min, depth_max)
    int nimg, H, W, Hv, Wv;  // padded size and the valid (unpadded) image size
    float w_normal, w_conf, conf_tol;
    uint16_t* dout;  // [p][32]
    float* sums;     // [5]: depth, normal, conf, count, (unused)
};
hipError_t loss(const LossArgs& a, hipStream_t s);

// Adam over a flat parameter vector (lr, beta1, beta2, eps, bias corrections b1t = 1 - beta1^t,
// b2t = 1 - beta2^t); the gradient is divided by max(1, *count) (read on the device).
hipError_t adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps, float b1t,
                float b2t, const float* count, hipStream_t s);

// Plane-sweep features of one reference view at the network resolution.
struct SweepArgs {
    const uint8_t* ref;
    int rw, rh;  // valid reference size (the feature tensor is W x H, padded with zeros)
    float rk[4];
    const uint8_t* src[2];
    int sw[2], sh[2];
    float sk[2][4];
    float r[2][9], t[2][3];  // reference -> source
    int nsrc;
    float dmin, dmax;
    uint16_t* feat;  // [H][W][64]
    int H, W;
};
hipError_t sweep(const SweepArgs& a, hipStream_t s);

// Planes of a full-resolution view from the network output of its quarter-resolution image:
// inverse depth bilinear, normal and confidence nearest, normal turned towards the camera; depth
// 0 where the confidence is below `conf_min` (random initialisation there).
struct InitPlanesArgs {
    const float* out;  // head output of the view's slot, [H][W][32]
    int H, W;          // network size
    int qw, qh;        // quarter-resolution image size (valid part of the network input)
    float dmin, dmax;
    int w, h;          // full resolution
    float k[4];
    float conf_min;
    float4* plane;     // full-resolution planes of the view
};
hipError_t init_planes(const InitPlanesArgs& a, hipStream_t s);

}  // namespace nn
}  // namespace c3d
