// terrain.hip — the synthetic aerial scenes of terrain.c rendered on the device: one thread per
// pixel runs the shared per-pixel model (terrain_core.h: fBm heightfield, ray march, shading) in
// double precision. Used to render large workloads (a 108-view 1080p zone is 224 Mpx, five
// marched rays per pixel) in seconds instead of minutes on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define TERRAIN_FN static __device__ inline
#include "terrain_core.h"

namespace {

__global__ void __launch_bounds__(256) k_render(c3d_terrain_params p, c3d_terrain_camera cm, float* img, float* depth, float* normal) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= cm.width || y >= cm.height) return;
    double C[3];
    terrain_center(&cm, C);
    const size_t pi = (size_t)y * cm.width + x;
    terrain_pixel(&p, &cm, C, x, y, &img[pi], &depth[pi], &normal[3 * pi]);
}

}  // namespace

// Renders `ncam` cameras on device 0 into host arrays laid out as c3d_render_terrain's (img/depth
// ncam*h*w floats packed camera after camera, normal 3 floats per pixel). Returns 0 or the HIP
// error code.
extern "C" int c3d_render_terrain_device(const c3d_terrain_params* p, const c3d_terrain_camera* cams, int ncam, float* img,
                                         float* depth, float* normal) {
    size_t maxpix = 0;
    for (int c = 0; c < ncam; ++c) {
        if (cams[c].width < 1 || cams[c].height < 1) return (int)hipErrorInvalidValue;
        const size_t n = (size_t)cams[c].width * cams[c].height;
        if (n > maxpix) maxpix = n;
    }
    if (ncam <= 0) return 0;
    float *d_img = nullptr, *d_dep = nullptr, *d_nor = nullptr;
    hipError_t e = hipMalloc((void**)&d_img, maxpix * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&d_dep, maxpix * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&d_nor, 3 * maxpix * sizeof(float));
    size_t off = 0;
    for (int c = 0; c < ncam && e == hipSuccess; ++c) {
        const c3d_terrain_camera& cm = cams[c];
        const size_t n = (size_t)cm.width * cm.height;
        dim3 grid((unsigned)((cm.width + 15) / 16), (unsigned)((cm.height + 15) / 16));
        hipLaunchKernelGGL(k_render, grid, dim3(256), 0, 0, *p, cm, d_img, d_dep, d_nor);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(img + off, d_img, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(depth + off, d_dep, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(normal + 3 * off, d_nor, 3 * n * sizeof(float), hipMemcpyDeviceToHost);
        off += n;
    }
    if (d_img) (void)hipFree(d_img);
    if (d_dep) (void)hipFree(d_dep);
    if (d_nor) (void)hipFree(d_nor);
    return (int)e;
}
