// pm_kernels.hpp — host-side interface of the kernels in pm_kernels.hip.
#pragma once

#include "pm_device.hpp"

namespace c3d {

// One compiled window variant (radius, step) of the window-templated kernels.
struct KernelSet {
    int wr, wstep;
    hipError_t (*half)(const KParams&, int rows, int tpp, hipStream_t);
    hipError_t (*eval)(const KParams&, int mode, hipStream_t);
    // cost of the pixels of colour P.color only (P.ntiles = half_tiles(items, 4, 1)), no initialisation
    hipError_t (*eval_color)(const KParams&, hipStream_t);
    hipError_t (*filter)(const KParams&, hipStream_t);
    hipError_t (*pair_costs)(const KParams&, int view, int slot, int n, const int* xy, const float4* planes, float* out,
                             hipStream_t);
    const void* (*half_symbol)();
};

// Kernels of a window (radius 1..MAX_WR, step 1 or 2): the fixed-radius build when compiled and
// `any_radius` is false, else the any-radius kernels of the step (KernelSet::wr == 0); nullptr
// outside those bounds.
const KernelSet* kernel_set(int wr, int wstep, bool any_radius);

// half-step workgroup shapes compiled: (rows of a workgroup, lanes per pixel) = (1, 1), (2, 1), (4, 1), (4, 2)
bool half_shape_ok(int rows, int tpp);

// workgroup counts of the two tilings for a batch
int half_tiles(const BatchItem* items, int n, int rows, int tpp);
int pix_tiles(const BatchItem* items, int n);

hipError_t launch_sample_reciprocal(const float* z, float* r, long long n, hipStream_t s);
// random initialisation by P.random_init and best-view reset of every pixel (P.ntiles = pix_tiles)
hipError_t launch_eval_init(const KParams& P, long long npix, hipStream_t s);
hipError_t launch_split_planes(const float4* plane, float* depth, float* normal, long long n, hipStream_t s);
hipError_t launch_merge_planes(const float* depth, const float* normal, float4* plane, long long n, hipStream_t s);
hipError_t launch_upsample(const KParams& P, hipStream_t s);
hipError_t launch_median(const KParams& P, hipStream_t s);
hipError_t launch_quad_build(const uint8_t* g, int w, int h, uint32_t* q, hipStream_t s);
hipError_t launch_extract_depth(const float4* plane, float* out, long long n, hipStream_t s);
hipError_t launch_restore(float4* plane, const float4* up, const float* c_init, float* cost, float xi, long long n,
                          hipStream_t s);
hipError_t launch_pyramid_step(const uint8_t* src, int w, int h, uint8_t* dst, const float g[3], hipStream_t s);
hipError_t launch_threshold(const KParams& P, const uint8_t* need, float4* out, long long npix, hipStream_t s);
hipError_t launch_half_tile_prefix(const BatchItem* batch, int n, int rows, int tpp, int* pre, hipStream_t s);
hipError_t launch_pix_tile_prefix(const BatchItem* batch, int n, int* pre, hipStream_t s);
hipError_t launch_span_tile_prefix(const BatchItem* batch, int n, int* pre, hipStream_t s);

}  // namespace c3d
