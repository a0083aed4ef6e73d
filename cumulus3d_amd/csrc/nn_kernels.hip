// nn_kernels.hip — kernels of the learned initialiser: a small U-Net in bf16 with fp32
// accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), its backward pass, Adam,
// the training loss against plain-PatchStereo labels and the plane-sweep input features.
//
// Convolutions are implicit GEMMs over NHWC tensors: the 64 lanes of a wave load 16 pixels x 32
// channels (one 16-byte load per lane) as an A operand and 32 channels x 16 output channels of
// the weights as the B operand, which serves the wave's 1 or 4 groups of 16 pixels; a 256-thread
// workgroup computes 64 or 256 pixels x 64 output channels.
// The weight gradient stages 32 pixels x 32 channels of both operands transposed in LDS so that
// the reduction over pixels is the MFMA's K dimension.
#include "nn_kernels.hpp"

namespace c3d {
namespace nn {

typedef short s16x8 __attribute__((__vector_size__(16)));
typedef float f32x4 __attribute__((__vector_size__(16)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma(s16x8 a, s16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// ---------------------------------------------------------------- convolution (forward / data gradient)

// G pixel groups of 16 per wave: the wave loads each B operand (32 input x 16 output channels of
// one tap) once for its G groups. Every output still accumulates tap by tap and 32 channels at a
// time in the same order, so G changes no result.
template <int G>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kg = lane >> 4;
    const long long HW = (long long)a.H * a.W;
    const long long P = (long long)a.nimg * HW;
    const long long p0 = ((long long)blockIdx.x * 4 + wave) * (16 * G);
    const int co0 = blockIdx.y * 64;
    const int ntile = min(4, (a.cout - co0) / 16);
    int img[G], py[G], px[G];
    bool pin[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const long long pa = p0 + 16 * g + r16;
        pin[g] = pa < P;
        img[g] = py[g] = px[g] = 0;
        if (pin[g]) {
            img[g] = (int)(pa / HW);
            const int rem = (int)(pa - (long long)img[g] * HW);
            py[g] = rem / a.W;
            px[g] = rem - py[g] * a.W;
        }
    }
    f32x4 acc[G][4];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[g][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int taps = a.ksize * a.ksize, rad = a.ksize / 2;
    for (int t = 0; t < taps; ++t) {
        const int dy = t / a.ksize - rad, dx = t % a.ksize - rad;
        bool ok[G];
        long long xoff[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int qy = py[g] + dy, qx = px[g] + dx;
            ok[g] = pin[g] && qy >= 0 && qy < a.H && qx >= 0 && qx < a.W;
            xoff[g] = ok[g] ? (((long long)img[g] * HW + (long long)qy * a.W + qx) * a.xs + kg * 8) : 0;
        }
        for (int c0 = 0; c0 < a.cin; c0 += 32) {
            s16x8 A[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                A[g] = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
                if (ok[g]) A[g] = *(const s16x8*)(a.x + xoff[g] + c0);
            }
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) {
                if (tt < ntile) {
                    const int co = co0 + tt * 16 + r16;
                    const s16x8 B = *(const s16x8*)(a.w + ((long long)co * taps + t) * a.cin + c0 + kg * 8);
#pragma unroll
                    for (int g = 0; g < G; ++g) acc[g][tt] = mfma(A[g], B, acc[g][tt]);
                }
            }
        }
    }
    // lane holds rows (pixels) 4 kg + j and column (output channel) r16 of each 16 x 16 tile
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            if (tt >= ntile) continue;
            const int co = co0 + tt * 16 + r16;
            const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long p = p0 + 16 * g + 4 * kg + j;
                if (p >= P) continue;
                float v = acc[g][tt][j] + b;
                if (a.mode == 0) {
                    if (a.relu) v = fmaxf(v, 0.f);
                    a.y[p * a.ys + a.yoff + co] = f2bf(v);
                } else if (a.mode == 1) {
                    a.y32[p * a.y32s + co] = v;
                } else {
                    a.y32[p * a.y32s + co] += v;
                }
            }
        }
    }
}

// Four pixel groups per wave where the launch still has at least 512 workgroups (two per CU),
// one otherwise (the small deep layers keep their parallelism).
hipError_t conv(const ConvArgs& a, hipStream_t s) {
    const long long P = (long long)a.nimg * a.H * a.W;
    const unsigned ny = (unsigned)((a.cout + 63) / 64);
    const long long n4 = (P + 255) / 256;
    if (n4 * ny >= 512) hipLaunchKernelGGL(k_conv<4>, dim3((unsigned)n4, ny), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_conv<1>, dim3((unsigned)((P + 63) / 64), ny), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- weight gradient

constexpr int WG_CHUNK = 4096;  // pixels reduced by one workgroup

__global__ __launch_bounds__(64) void k_wgrad(WgradArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t sdz[32][40];  // [output channel][pixel]
    __shared__ __attribute__((aligned(16))) uint16_t sx[32][40];   // [input channel][pixel]
    const int lane = threadIdx.x;
    const int r16 = lane & 15, kg = lane >> 4;
    const int ci0 = blockIdx.x * 32, co0 = blockIdx.y * 32;
    const int taps = a.ksize * a.ksize, rad = a.ksize / 2;
    const int t = blockIdx.z % taps;
    const long long chunk = blockIdx.z / taps;
    const int dy = t / a.ksize - rad, dx = t % a.ksize - rad;
    const long long HW = (long long)a.H * a.W;
    const long long P = (long long)a.nimg * HW;
    const long long pbeg = chunk * WG_CHUNK;
    const long long pend = min(P, pbeg + WG_CHUNK);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long pb = pbeg; pb < pend; pb += 32) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = lane * 2 + h;  // 0..127: pixel e / 4, channel group (e % 4) * 8
            const int pxl = e >> 2, cg = (e & 3) * 8;
            const long long p = pb + pxl;
            s16x8 vz = {0, 0, 0, 0, 0, 0, 0, 0}, vx = {0, 0, 0, 0, 0, 0, 0, 0};
            if (p < pend) {
                vz = *(const s16x8*)(a.dz + p * a.dzs + co0 + cg);
                const int img = (int)(p / HW);
                const int rem = (int)(p - (long long)img * HW);
                const int qy = rem / a.W + dy, qx = rem % a.W + dx;
                if (qy >= 0 && qy < a.H && qx >= 0 && qx < a.W)
                    vx = *(const s16x8*)(a.x + ((long long)img * HW + (long long)qy * a.W + qx) * a.xs + ci0 + cg);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                sdz[cg + j][pxl] = (uint16_t)vz[j];
                sx[cg + j][pxl] = (uint16_t)vx[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const s16x8 A = *(const s16x8*)&sdz[i * 16 + r16][kg * 8];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const s16x8 B = *(const s16x8*)&sx[j * 16 + r16][kg * 8];
                acc[i][j] = mfma(A, B, acc[i][j]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int co = co0 + i * 16 + 4 * kg + q, ci = ci0 + j * 16 + r16;
                atomicAdd(&a.dw[((long long)co * taps + t) * a.cin + ci], acc[i][j][q]);
            }
}

hipError_t wgrad(const WgradArgs& a, hipStream_t s) {
    const long long P = (long long)a.nimg * a.H * a.W;
    const long long chunks = (P + WG_CHUNK - 1) / WG_CHUNK;
    dim3 grid((unsigned)(a.cin / 32), (unsigned)(a.cout / 32), (unsigned)(chunks * a.ksize * a.ksize));
    hipLaunchKernelGGL(k_wgrad, grid, dim3(64), 0, s, a);
    return hipGetLastError();
}

__global__ void k_bias_grad(const uint16_t* dz, int dzs, int c, long long npix, float* db) {
    const int ch = blockIdx.y * 64 + (threadIdx.x & 63);
    if (ch >= c) return;
    const long long p0 = (long long)blockIdx.x * WG_CHUNK;
    const long long p1 = min(npix, p0 + WG_CHUNK);
    float acc = 0.f;
    for (long long p = p0 + (threadIdx.x >> 6); p < p1; p += 4) acc += bf2f(dz[p * dzs + ch]);
    atomicAdd(&db[ch], acc);
}

hipError_t bias_grad(const uint16_t* dz, int dzs, int c, long long npix, float* db, hipStream_t s) {
    dim3 grid((unsigned)((npix + WG_CHUNK - 1) / WG_CHUNK), (unsigned)((c + 63) / 64));
    hipLaunchKernelGGL(k_bias_grad, grid, dim3(256), 0, s, dz, dzs, c, npix, db);
    return hipGetLastError();
}

// ---------------------------------------------------------------- elementwise and resampling

__global__ void k_relu_grad(const float* da, int das, const uint16_t* a, int as, uint16_t* dz, int dzs, int c, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n * c) return;
    const long long p = i / c;
    const int ch = (int)(i - p * c);
    const float g = bf2f(a[p * as + ch]) > 0.f ? da[p * das + ch] : 0.f;
    dz[p * dzs + ch] = f2bf(g);
}

hipError_t relu_grad(const float* da, int das, const uint16_t* a, int as, uint16_t* dz, int dzs, int c, long long npix,
                     hipStream_t s) {
    const long long n = npix * c;
    hipLaunchKernelGGL(k_relu_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, da, das, a, as, dz, dzs, c, npix);
    return hipGetLastError();
}

__global__ void k_pool(const uint16_t* x, int xs, int c, int nimg, int H, int W, uint16_t* y, int ys) {
    const int Ho = H / 2, Wo = W / 2;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)nimg * Ho * Wo * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    const long long q = i / c;
    const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), b = (int)(q / ((long long)Wo * Ho));
    const long long base = ((long long)b * H + 2 * yo) * W + 2 * xo;
    const float v = bf2f(x[base * xs + ch]) + bf2f(x[(base + 1) * xs + ch]) + bf2f(x[(base + W) * xs + ch]) +
                    bf2f(x[(base + W + 1) * xs + ch]);
    y[q * ys + ch] = f2bf(0.25f * v);
}

hipError_t pool(const uint16_t* x, int xs, int c, int nimg, int H, int W, uint16_t* y, int ys, hipStream_t s) {
    const long long n = (long long)nimg * (H / 2) * (W / 2) * c;
    hipLaunchKernelGGL(k_pool, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, xs, c, nimg, H, W, y, ys);
    return hipGetLastError();
}

__global__ void k_pool_grad(const float* dy, int dys, int c, int nimg, int H, int W, float* dx, int dxs) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)nimg * H * W * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    const long long p = i / c;
    const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long long)W * H));
    const int Ho = H / 2, Wo = W / 2;
    const long long q = ((long long)b * Ho + y / 2) * Wo + x / 2;
    dx[p * dxs + ch] += 0.25f * dy[q * dys + ch];
}

hipError_t pool_grad(const float* dy, int dys, int c, int nimg, int H, int W, float* dx, int dxs, hipStream_t s) {
    const long long n = (long long)nimg * H * W * c;
    hipLaunchKernelGGL(k_pool_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dy, dys, c, nimg, H, W, dx, dxs);
    return hipGetLastError();
}

__global__ void k_upsample(const uint16_t* x, int xs, int c, int nimg, int H, int W, uint16_t* y, int ys, int yoff) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const int H2 = 2 * H, W2 = 2 * W;
    const long long n = (long long)nimg * H2 * W2 * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    const long long p = i / c;
    const int xx = (int)(p % W2), yy = (int)((p / W2) % H2), b = (int)(p / ((long long)W2 * H2));
    const long long q = ((long long)b * H + yy / 2) * W + xx / 2;
    y[p * ys + yoff + ch] = x[q * xs + ch];
}

hipError_t upsample(const uint16_t* x, int xs, int c, int nimg, int H, int W, uint16_t* y, int ys, int yoff, hipStream_t s) {
    const long long n = (long long)nimg * 4 * H * W * c;
    hipLaunchKernelGGL(k_upsample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, xs, c, nimg, H, W, y, ys, yoff);
    return hipGetLastError();
}

__global__ void k_upsample_grad(const float* dy, int dys, int yoff, int c, int nimg, int H, int W, float* dx, int dxs) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)nimg * H * W * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    const long long q = i / c;
    const int x = (int)(q % W), y = (int)((q / W) % H), b = (int)(q / ((long long)W * H));
    const int W2 = 2 * W;
    const long long p = ((long long)b * 2 * H + 2 * y) * W2 + 2 * x;
    const float g = dy[p * dys + yoff + ch] + dy[(p + 1) * dys + yoff + ch] + dy[(p + W2) * dys + yoff + ch] +
                    dy[(p + W2 + 1) * dys + yoff + ch];
    dx[q * dxs + ch] += g;
}

hipError_t upsample_grad(const float* dy, int dys, int yoff, int c, int nimg, int H, int W, float* dx, int dxs, hipStream_t s) {
    const long long n = (long long)nimg * H * W * c;
    hipLaunchKernelGGL(k_upsample_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dy, dys, yoff, c, nimg, H, W, dx, dxs);
    return hipGetLastError();
}

__global__ void k_copy_channels(const uint16_t* x, int xs, int c, long long npix, uint16_t* y, int ys, int yoff) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= npix * c) return;
    const long long p = i / c;
    const int ch = (int)(i - p * c);
    y[p * ys + yoff + ch] = x[p * xs + ch];
}

hipError_t copy_channels(const uint16_t* x, int xs, int c, long long npix, uint16_t* y, int ys, int yoff, hipStream_t s) {
    const long long n = npix * c;
    hipLaunchKernelGGL(k_copy_channels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, xs, c, npix, y, ys, yoff);
    return hipGetLastError();
}

__global__ void k_add_channels(const float* dy, int dys, int off, int c, long long npix, float* dx, int dxs) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= npix * c) return;
    const long long p = i / c;
    const int ch = (int)(i - p * c);
    dx[p * dxs + ch] += dy[p * dys + off + ch];
}

hipError_t add_channels(const float* dy, int dys, int off, int c, long long npix, float* dx, int dxs, hipStream_t s) {
    const long long n = npix * c;
    hipLaunchKernelGGL(k_add_channels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dy, dys, off, c, npix, dx, dxs);
    return hipGetLastError();
}

__global__ void k_fill(float* p, float v, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t fill(float* p, float v, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

__global__ void k_weights_bf16(const float* w32, int cout, int taps, int cin, uint16_t* w, uint16_t* wt) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)cout * taps * cin;
    if (i >= n) return;
    const int ci = (int)(i % cin);
    const int t = (int)((i / cin) % taps);
    const int co = (int)(i / ((long long)cin * taps));
    const uint16_t b = f2bf(w32[i]);
    w[i] = b;
    // data gradient: wt[ci][taps - 1 - t][co]
    wt[((long long)ci * taps + (taps - 1 - t)) * cout + co] = b;
}

hipError_t weights_bf16(const float* w32, int cout, int taps, int cin, uint16_t* w, uint16_t* wt, hipStream_t s) {
    const long long n = (long long)cout * taps * cin;
    hipLaunchKernelGGL(k_weights_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w32, cout, taps, cin, w, wt);
    return hipGetLastError();
}

// ---------------------------------------------------------------- loss

__device__ __forceinline__ float sigm(float x) { return 1.f / (1.f + __expf(-x)); }

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__global__ void k_loss(LossArgs a) {
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long HW = (long long)a.H * a.W;
    const long long P = a.nimg * HW;
    float ld = 0.f, ln = 0.f, lc = 0.f, cnt = 0.f;
    if (p < P) {
        float g[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        const int b = (int)(p / HW);
        const int rem = (int)(p - (long long)b * HW);
        const int y = rem / a.W, x = rem - y * a.W;
        const float* o = a.out + p * 32;
        const float* lb = a.label + p * 4;
        if (y < a.Hv && x < a.Wv && lb[0] > 0.f) {
            const float ia = 1.f / a.range[2 * b + 1], ib = 1.f / a.range[2 * b];
            const float s0 = sigm(o[0]);
            const float id = ia + s0 * (ib - ia);
            const float idt = 1.f / lb[0];
            const float e = id - idt;
            ld = fabsf(e) / (ib - ia);
            g[0] = (e > 0.f ? 1.f : (e < 0.f ? -1.f : 0.f)) * s0 * (1.f - s0);
            const float vx = o[1], vy = o[2], vz = o[3];
            const float nv = sqrtf(vx * vx + vy * vy + vz * vz + 1e-6f);
            const float nx = vx / nv, ny = vy / nv, nz = vz / nv;
            const float dot = nx * lb[1] + ny * lb[2] + nz * lb[3];
            ln = 1.f - dot;
            g[1] = -a.w_normal * (lb[1] - dot * nx) / nv;
            g[2] = -a.w_normal * (lb[2] - dot * ny) / nv;
            g[3] = -a.w_normal * (lb[3] - dot * nz) / nv;
            const float c = sigm(o[4]);
            const float tgt = fabsf(1.f / id - lb[0]) <= a.conf_tol * lb[0] ? 1.f : 0.f;
            lc = -(tgt * __logf(fmaxf(c, 1e-7f)) + (1.f - tgt) * __logf(fmaxf(1.f - c, 1e-7f)));
            g[4] = a.w_conf * (c - tgt);
            cnt = 1.f;
        }
        uint16_t* d = a.dout + p * 32;
#pragma unroll
        for (int k = 0; k < 5; ++k) d[k] = f2bf(g[k]);
    }
    ld = wave_sum(ld);
    ln = wave_sum(ln);
    lc = wave_sum(lc);
    cnt = wave_sum(cnt);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&a.sums[0], ld);
        atomicAdd(&a.sums[1], ln);
        atomicAdd(&a.sums[2], lc);
        atomicAdd(&a.sums[3], cnt);
    }
}

hipError_t loss(const LossArgs& a, hipStream_t s) {
    const long long P = (long long)a.nimg * a.H * a.W;
    hipLaunchKernelGGL(k_loss, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- Adam

__global__ void k_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps, float b1t,
                       float b2t, const float* count) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] / fmaxf(count[0], 1.f);
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / b1t) / (sqrtf(vi / b2t) + eps);
}

hipError_t adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps, float b1t,
                float b2t, const float* count, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, b1t, b2t, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------- plane-sweep features

// Channels 0..31: mean over the sources of |I_ref - I_src(H_k x)| / 255 for fronto-parallel planes
// at 32 inverse depths evenly spaced (bin centres) between 1/dmax and 1/dmin; 1 where no source
// sees the pixel. Channel 32: I_ref / 255. Channels 33..63 and the padding pixels: 0.
__device__ __forceinline__ float bilin_u8(const uint8_t* g, int w, int h, float u, float v, bool& inside) {
    const float xf = floorf(u), yf = floorf(v);
    const int x0 = (int)xf, y0 = (int)yf;
    inside = u >= 0.f && v >= 0.f && u <= (float)(w - 1) && v <= (float)(h - 1);
    if (!inside) return 0.f;
    const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    const float ax = u - xf, ay = v - yf;
    const float p00 = g[y0 * w + x0], p10 = g[y0 * w + x1], p01 = g[y1 * w + x0], p11 = g[y1 * w + x1];
    const float top = p00 + ax * (p10 - p00), bot = p01 + ax * (p11 - p01);
    return top + ay * (bot - top);
}

__global__ void k_sweep(SweepArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.H * a.W) return;
    const int y = i / a.W, x = i - y * a.W;
    uint16_t* f = a.feat + (long long)i * 64;
    if (x >= a.rw || y >= a.rh) {
        for (int c = 0; c < 64; ++c) f[c] = 0;
        return;
    }
    const float ir = a.ref[y * a.rw + x];
    const float rx = (x - a.rk[2]) / a.rk[0], ry = (y - a.rk[3]) / a.rk[1];
    const float ia = 1.f / a.dmax, ib = 1.f / a.dmin;
    for (int k = 0; k < 32; ++k) {
        const float d = 1.f / (ia + (k + 0.5f) / 32.f * (ib - ia));
        const float X = rx * d, Y = ry * d, Z = d;
        float cost = 0.f;
        int n = 0;
        for (int s = 0; s < a.nsrc; ++s) {
            const float* R = a.r[s];
            const float xs = R[0] * X + R[1] * Y + R[2] * Z + a.t[s][0];
            const float ys = R[3] * X + R[4] * Y + R[5] * Z + a.t[s][1];
            const float zs = R[6] * X + R[7] * Y + R[8] * Z + a.t[s][2];
            if (!(zs > 1e-6f)) continue;
            const float u = a.sk[s][0] * xs / zs + a.sk[s][2], v = a.sk[s][1] * ys / zs + a.sk[s][3];
            bool inside;
            const float is = bilin_u8(a.src[s], a.sw[s], a.sh[s], u, v, inside);
            if (!inside) continue;
            cost += fabsf(ir - is);
            n++;
        }
        f[k] = f2bf(n > 0 ? cost / (255.f * n) : 1.f);
    }
    f[32] = f2bf(ir / 255.f);
    for (int c = 33; c < 64; ++c) f[c] = 0;
}

hipError_t sweep(const SweepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_sweep, dim3((unsigned)((a.H * a.W + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- network output -> PatchStereo planes

__global__ void k_init_planes(InitPlanesArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.w * a.h) return;
    const int y = i / a.w, x = i - y * a.w;
    // the full-resolution pixel centre in quarter-resolution pixels (two halvings of the pyramid,
    // c <- (c + 0.5) / 2 - 0.5)
    const float u = (x + 0.5f) * 0.25f - 0.5f, v = (y + 0.5f) * 0.25f - 0.5f;
    const float uc = fminf(fmaxf(u, 0.f), (float)(a.qw - 1)), vc = fminf(fmaxf(v, 0.f), (float)(a.qh - 1));
    const int x0 = (int)floorf(uc), y0 = (int)floorf(vc);
    const int x1 = min(x0 + 1, a.qw - 1), y1 = min(y0 + 1, a.qh - 1);
    const float ax = uc - x0, ay = vc - y0;
    const float ia = 1.f / a.dmax, ib = 1.f / a.dmin;
    auto inv = [&](int qx, int qy) { return ia + sigm(a.out[((long long)qy * a.W + qx) * 32]) * (ib - ia); };
    const float top = inv(x0, y0) + ax * (inv(x1, y0) - inv(x0, y0));
    const float bot = inv(x0, y1) + ax * (inv(x1, y1) - inv(x0, y1));
    const float id = top + ay * (bot - top);
    const int qx = min(max((int)roundf(uc), 0), a.qw - 1), qy = min(max((int)roundf(vc), 0), a.qh - 1);
    const float* o = a.out + ((long long)qy * a.W + qx) * 32;
    float4 pl = make_float4(0.f, 0.f, 0.f, 0.f);
    if (sigm(o[4]) >= a.conf_min) {
        const float nv = sqrtf(o[1] * o[1] + o[2] * o[2] + o[3] * o[3] + 1e-6f);
        float nx = o[1] / nv, ny = o[2] / nv, nz = o[3] / nv;
        const float rx = (x - a.k[2]) / a.k[0], ry = (y - a.k[3]) / a.k[1];
        if (nx * rx + ny * ry + nz > 0.f) {
            nx = -nx;
            ny = -ny;
            nz = -nz;
        }
        pl = make_float4(nx, ny, nz, 1.f / id);
    }
    a.plane[i] = pl;
}

hipError_t init_planes(const InitPlanesArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_init_planes, dim3((unsigned)((a.w * a.h + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace c3d
