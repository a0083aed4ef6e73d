/*
 * c3d_scene.h — C ABI of the scene preparation that feeds the PatchStereo engine: source-view
 * selection from sparse point tracks and per-view depth ranges (crates/dense/src/neighbors.rs,
 * called by compute_depth_maps, densify.rs:224-241, and for the fusion overlap lists,
 * densify.rs:513-519).
 *
 * The pair statistics and the depth ranges run on the device of a backend (every track pair and
 * every observation is one key of a radix sort; the per-pair counts, 75th-percentile angles and
 * per-view 1 % / 99 % depths are read off the sorted runs). The selection itself (a sort of at
 * most num_views candidates per view) runs on the host.
 */
#ifndef C3D_SCENE_H
#define C3D_SCENE_H

#include <stdint.h>

#include "c3d_pm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The cameras and sparse points of a DenseScene (scene.rs): the views of point p are
 * track_views[track_offset[p] .. track_offset[p + 1]), in the track's order. */
typedef struct c3d_sparse_scene {
    uint32_t num_views;
    uint32_t pad0;
    const double* rotation;        /* 9 per view: world -> camera, row-major */
    const double* translation;     /* 3 per view */
    uint64_t num_points;
    const double* xyz;             /* 3 per point, world */
    const uint64_t* track_offset;  /* num_points + 1, non-decreasing, track_offset[0] = 0 */
    const uint32_t* track_views;   /* view indices < num_views */
} c3d_sparse_scene;

/* PairStats::new (neighbors.rs:47-80) on the device: for every pair of entries (i, j < i) of a
 * track whose views a, b differ, the pair (min, max) counts once and records the angle
 * min(t, pi - t) at the point between the rays to the two camera centres (double, stored as
 * float; 0 when a centre coincides with the point). Outputs are num_views x num_views, symmetric,
 * row-major: count[a n + b] = shared count, q75[a n + b] = quantile_sorted(sorted angles, 0.75)
 * (math.rs:126-135, in double), 0 where the count is 0. */
int c3d_pair_stats(c3d_pm_backend* backend, const c3d_sparse_scene* scene, uint32_t* count, double* q75);

/* PairStats::select (neighbors.rs:83-103), host only: the views o != r with count > 0 (and, when
 * min_angle_rad > 0, q75 >= min_angle_rad), in decreasing count then increasing index, at most
 * max_views of them into out; *num_out receives how many. */
int c3d_select_views(uint32_t num_views, const uint32_t* count, const double* q75, uint32_t r, uint32_t max_views,
                     double min_angle_rad, uint32_t* out, uint32_t* num_out);

/* PairStats::select_diverse (neighbors.rs:110-151), host only: candidates of select(r, all,
 * min_angle_rad) scored by shared count times decay^k, k = views already taken in the same sector
 * of the baseline's azimuth in camera r (`bins` sectors, one more for baselines within
 * atan(0.3) of the optical axis: lateral part below 0.3 times the axial part), greedily. */
int c3d_select_diverse(const c3d_sparse_scene* scene, const uint32_t* count, const double* q75, uint32_t r, uint32_t max_views,
                       double min_angle_rad, uint32_t bins, double decay, uint32_t* out, uint32_t* num_out);

/* depth_ranges and depth_range_of (neighbors.rs:162-189) on the device: per view, the depths r_3 . x + t_3 > 0 of
 * every observation of a point (a view listed twice counts twice), sorted; lo = d[floor(0.01 n)],
 * hi = d[floor(0.99 n)] (both at most n - 1); range = (0.75 lo, 1.25 hi). has_range[v] = 0 and
 * the range 0 when the view observes no point in front. */
int c3d_depth_ranges(c3d_pm_backend* backend, const c3d_sparse_scene* scene, double* depth_min, double* depth_max,
                     uint8_t* has_range);

#ifdef __cplusplus
}
#endif

#endif /* C3D_SCENE_H */
