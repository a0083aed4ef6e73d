/*
 * c3d_fusion.h — C ABI of the depth-map fusion on the MI355X: agreement ("consistency") fusion of
 * the final depth maps into a point cloud, the reference's default fusion mode
 * (fuse_consistency, crates/dense/src/fusion.rs:202-350; FusionMode::Consistency, params.rs).
 *
 * Views are fused in index order. A reference pixel with a positive depth that no earlier view
 * has marked as used becomes a point when at least min_consistent_views of its first
 * consistency_num_images overlap views agree: the nearest pixel of the projection has a depth
 * within max_depth_error (relative), its point reprojects within max_reproj_error pixels, and the
 * normals differ by at most max_normal_error_deg. The point is the inverse-variance weighted
 * mean (sigma_d = d^2 sigma_px / (f b), sigma_px = sigma_px0 + sigma_px_slope * cost,
 * fusion.rs:78-89) of the reference and agreeing points, normals and colours; with mark_used the
 * agreeing pixels (and, with mark_radius, their neighbours of matching depth) are not used as
 * reference pixels later (fusion.rs:360-383). All geometry is double precision, as in the
 * reference. Points are returned in the reference's order (views, rows, columns).
 *
 * Not here: traversal fusion, the second pass of leftover pixels, the visibility filter and the
 * normal-direction smoothing of the cloud (fusion.rs:219-460 options that the fast profile does
 * not need for the comparison of clouds).
 */
#ifndef C3D_FUSION_H
#define C3D_FUSION_H

#include <stdint.h>

#include "c3d_pm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One fused view (FusionInput, fusion.rs:58-76, plus its camera). */
typedef struct c3d_fusion_view {
    uint32_t width;
    uint32_t height;
    const float* depth;   /* width*height, 0 = invalid */
    const float* normal;  /* 3 per pixel, reference camera coordinates */
    const float* cost;    /* final cost per pixel, or NULL (no inverse-variance weight) */
    const uint8_t* gray;  /* colour source (gray replicated to RGB), or NULL (black) */
    double k[4];          /* fx, fy, cx, cy */
    double r[9];          /* world -> camera rotation, row-major */
    double t[3];          /* world -> camera translation */
    double baseline;      /* median centre distance to the sources (densify.rs:260-276); 0 = no weight */
    uint32_t image_id;    /* point key (fusion.rs:54-56) */
    uint32_t num_overlap;
    const uint32_t* overlap; /* overlap views in decreasing shared-point order (PairStats::select) */
} c3d_fusion_view;

/* FusionParams (params.rs:236-290); the fast profile sets min_consistent_views 3,
 * max_depth_error 0.005, max_normal_error_deg 23, max_reproj_error 100 (params.rs:456-458). */
typedef struct c3d_fusion_params {
    uint32_t min_consistent_views;
    uint32_t consistency_num_images;
    int32_t mark_used;
    uint32_t mark_radius;
    int32_t inverse_variance;
    uint32_t pad0;
    double sigma_px0;
    double sigma_px_slope;
    double max_reproj_error;
    double max_depth_error;
    double max_normal_error_deg;
} c3d_fusion_params;

/* Output cloud (PointCloud): caller-owned arrays of `capacity` points. */
typedef struct c3d_point_cloud {
    uint64_t capacity;
    uint64_t count;      /* points fused; when above capacity only capacity are written (INVALID_ARGUMENT) */
    float* positions;    /* 3 per point, world */
    float* normals;      /* 3 per point, world */
    uint8_t* colors;     /* 3 per point */
    uint64_t* keys;      /* (image_id << 32) | (row << 16) | col */
} c3d_point_cloud;

/* Agreement fusion on the device of `backend`. */
int c3d_fuse_consistency(c3d_pm_backend* backend, const c3d_fusion_view* views, uint32_t num_views,
                         const c3d_fusion_params* params, c3d_point_cloud* out);

#ifdef __cplusplus
}
#endif

#endif /* C3D_FUSION_H */
