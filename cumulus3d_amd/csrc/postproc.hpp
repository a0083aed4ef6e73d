// postproc.hpp — the device post stage (postproc.hip) as the PatchStereo session calls it.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/c3d_pm.h"

namespace c3d {

// One view of a post-processing batch: its pixels at `off` of the concatenated plane/cost arrays,
// its gray image at byte `goff`, its top-scale intrinsics in double (fill_holes reads view.k).
struct PostView {
    long long off;
    long long goff;
    int w, h;
    double k[4];
};

struct PostWork;  // device scratch, grown on demand, kept between calls
PostWork* post_work_new();
void post_work_free(PostWork* w);

// Validates the parameters (C3D_INVALID_ARGUMENT).
int post_check(const c3d_post_params& p);

// remove_speckles then fill_holes (densify.rs:455-462) on planes (nx, ny, nz, d), in place, on
// stream `st`; `raw_cost` may alias `cost`. Synchronises the stream (once to size the sort, once at
// the end). removed/filled (per view, may be null) receive the counts.
int post_run(PostWork* w, const PostView* views, int n, float4* plane, float* cost, const float* raw_cost, const uint8_t* gray,
             const c3d_post_params& p, hipStream_t st, unsigned long long* removed, unsigned long long* filled);

}  // namespace c3d
