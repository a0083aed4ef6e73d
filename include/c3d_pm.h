/*
 * c3d_pm.h — C ABI of the MI355X (gfx950) PatchStereo depth-map engine.
 *
 * This is the drop-in boundary that replaces the Rust trait pair of cumulus3d-dense
 *   PatchMatchBackend  (crates/dense/src/kernel.rs:239-245)
 *   PatchMatchSession  (crates/dense/src/kernel.rs:191-237)
 * and the CUDA implementation behind it (crates/cuda/src/patchmatch.rs:285-869).
 * Every entry point below names the reference item it stands in for.
 *
 * Conventions (crates/dense/src/kernel.rs:36-74): integer pixel = pixel centre, ray(x,y) =
 * ((x-cx)/fx, (y-cy)/fy, 1), depth = camera z, normals in reference-camera coordinates with
 * n.ray < 0. Images are 8-bit gray, row-major, no padding (LevelImage, kernel.rs:81-91).
 *
 * Ownership: c3d_pm_begin copies every image and geometry table to the device; the caller may
 * free its input afterwards (the reference session borrows it instead, kernel.rs:244).
 * View states are caller-owned host buffers updated in place (kernel.rs:193-194).
 *
 * Threading: a backend handle may be shared by threads; c3d_pm_begin blocks while another
 * session of the same backend is open (one session at a time per device, patchmatch.rs:96-97,
 * 290). A session is used by one thread at a time.
 *
 * Errors: every int-returning call returns a c3d_status; the message of the last failure on the
 * calling thread is returned by c3d_pm_last_error().
 */
#ifndef C3D_PM_H
#define C3D_PM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define C3D_PM_ABI_VERSION 2
/* Maximum number of source views per reference view (MAXSRC, patchmatch.rs:52). */
#define C3D_MAX_SOURCES 32

/* Status codes. The Rust shim maps them to cumulus3d_core::Error:
 * INVALID_ARGUMENT -> Error::InvalidArgument, DEVICE -> Error::Invariant("HIP: ..."),
 * NOT_AVAILABLE / COMPILE -> GpuError::{NotAvailable, Compile} (crates/cuda/src/device.rs:66-78). */
typedef enum c3d_status {
    C3D_OK = 0,
    C3D_INVALID_ARGUMENT = 1,
    C3D_DEVICE = 2,
    C3D_NOT_AVAILABLE = 3,
    C3D_COMPILE = 4
} c3d_status;

typedef struct c3d_pm_backend c3d_pm_backend;
typedef struct c3d_pm_session c3d_pm_session;

/* Backend options; stands in for CudaPatchMatchOptions (patchmatch.rs:56-81).
 * Zero-initialised fields select the defaults given in brackets. */
typedef struct c3d_pm_options {
    int64_t batch_pixels;     /* pixel budget of one launch [2,200,000] (batch_pixels) */
    int32_t max_batch;        /* views per launch [64] (max_batch) */
    int32_t tile_rows;        /* rows of active pixels per half-step workgroup (1, 2 or 4); 0 = 1 (4 with
                                 threads_per_pixel 2) until c3d_pm_autotune */
    int32_t threads_per_pixel;/* lanes per pixel in the half step: 1, or 2 with tile_rows 4 (the two lanes
                                 split the pixel's NCC windows); 0 = 1 until c3d_pm_autotune */
    int32_t use_graph;        /* 1 = capture each run's iteration loop in a hipGraph; 0 = plain launches */
    int32_t any_radius;       /* 1 = always the any-radius kernels (weights recomputed per sample), even
                                 for the windows with a register-resident build; 0 = fastest available */
    int32_t reserved[5];
} c3d_pm_options;

/* One view image at one scale (LevelImage, kernel.rs:81-91). k = fx, fy, cx, cy. */
typedef struct c3d_level_image {
    uint32_t width;
    uint32_t height;
    const uint8_t* gray;
    float k[4];
} c3d_level_image;

/* Reference -> source geometry (PairGeometry, kernel.rs:93-103): X_src = r X_ref + t;
 * center = source camera centre in reference-camera coordinates. */
typedef struct c3d_pair_geometry {
    float r[9];
    float t[3];
    float center[3];
} c3d_pair_geometry;

/* One kernel input view (KernelView, kernel.rs:105-120). levels has num_levels entries,
 * 0 = lowest resolution. */
typedef struct c3d_kernel_view {
    uint64_t key;
    const c3d_level_image* levels;
    uint32_t num_sources;
    uint32_t pad0;
    const uint32_t* sources;
    const c3d_pair_geometry* pairs;
    float depth_min;
    float depth_max;
} c3d_kernel_view;

/* PatchMatch constants (PmParams, crates/dense/src/params.rs:76-124); defaults params.rs:126-154. */
typedef struct c3d_pm_params {
    int32_t window_radius;
    int32_t window_step;
    float sigma_spatial; /* <= 0: window_radius (PmParams::sigma_s, params.rs:158-164) */
    float sigma_color;
    float tau0;
    float tau_alpha;
    float tau1;
    float beta;
    uint32_t n1;
    uint32_t n2;
    float prev_view_weight;
    float min_triangulation_angle_deg;
    float incident_angle_sigma;
    float geom_lambda;
    float geom_max_cost;
    uint32_t top_k;
    float eps0_photometric;
    float eps0_geometric;
    float phi0_photometric_deg;
    float phi0_geometric_deg;
    float eps_min;
    float phi_min_deg;
    float weak_texture_var;
    int32_t pad0;
} c3d_pm_params;

/* Read-out tolerances (FilterParams, params.rs:179-207). */
typedef struct c3d_filter_params {
    float min_ncc;
    float min_triangulation_angle_deg;
    uint32_t min_num_consistent;
    float geom_max_cost;
    float ncc_sigma;
    int32_t median_filter;
} c3d_filter_params;

/* Whole kernel input (KernelInput, kernel.rs:122-135). */
typedef struct c3d_kernel_input {
    const c3d_kernel_view* views;
    uint32_t num_views;
    uint32_t num_levels;
    c3d_pm_params pm;
    c3d_filter_params filter;
    uint64_t seed;
} c3d_kernel_input;

/* Pixel state of one reference view at one scale (ViewState, kernel.rs:137-160), structure of
 * arrays: depth[n], normal[3n], cost[n], prior[4n] (nx, ny, nz, d) or NULL when absent. */
typedef struct c3d_view_state {
    uint32_t width;
    uint32_t height;
    float* depth;
    float* normal;
    float* cost;
    float* prior;
} c3d_view_state;

/* Depth snapshot for geometric runs (DepthSnapshot, kernel.rs:162-171): maps[v] is the depth map
 * of view v at `level` (width*height floats) or NULL. Equal id implies equal content; the session
 * re-uploads only when the id changes (patchmatch.rs:573-576). */
typedef struct c3d_depth_snapshot {
    uint64_t id;
    uint32_t level;
    uint32_t num_maps;
    const float* const* maps;
} c3d_depth_snapshot;

/* Parameters of one run (RunParams, kernel.rs:173-189). random_init: 0 = start from the given
 * planes, 1 = random initialisation of every pixel (the reference's bool), 2 = random
 * initialisation only where the given depth is not positive (no reference counterpart: the
 * learned initialiser marks its low-confidence pixels that way). */
typedef struct c3d_run_params {
    uint32_t level;
    int32_t geometric;
    int32_t random_init;
    uint32_t iterations;
    uint32_t run_id;
    int32_t use_prior;
} c3d_run_params;

/* ---------------------------------------------------------------- backend */

/* Opens device `ordinal`. Stands in for CudaPatchMatch::new / try_default
 * (patchmatch.rs:239-246) and make_pm_backend("cuda") (crates/cli/src/densewrap.rs:42-49).
 * opts may be NULL (defaults). Returns C3D_NOT_AVAILABLE without a usable gfx950 device. */
int c3d_pm_create(int ordinal, const c3d_pm_options* opts, c3d_pm_backend** out);

/* Backend name for logs, "hip(<device name>)"; PatchMatchBackend::name (kernel.rs:241-242,
 * patchmatch.rs:286-288). Valid until c3d_pm_destroy. */
const char* c3d_pm_name(const c3d_pm_backend* backend);

/* Releases the backend (drop of CudaPatchMatch, patchmatch.rs:285-298). Open sessions and
 * learned-initialiser networks (c3d_init.h) hold the backend, as a Session borrows its backend for
 * its lifetime in the reference (kernel.rs:244). Returns 0 when the backend was freed now; otherwise
 * the number of sessions and networks still open: the backend is then freed when the last of them
 * is closed (c3d_pm_end, c3d_init_destroy), and every further c3d_pm_begin or c3d_init_create on
 * it fails with C3D_INVALID_ARGUMENT. The handle must not be used by the caller after this call.
 * A backend keeps two 256 MB page-locked host buffers, through which sessions download their
 * results, from the first download until it is freed. */
int c3d_pm_destroy(c3d_pm_backend* backend);

/* 1 when a gfx950 device can be opened; is_available (crates/cuda/src/device.rs:80-89). */
int c3d_pm_is_available(void);

/* ---------------------------------------------------------------- session */

/* Uploads images and geometry and opens a session; PatchMatchBackend::begin (kernel.rs:243-244,
 * patchmatch.rs:289-297, Session::new patchmatch.rs:355-547). Fails with INVALID_ARGUMENT for a
 * window radius outside 1..=20, step outside 1..=2 (patchmatch.rs:292-294), more than 32 sources
 * (patchmatch.rs:369-371), a level count mismatch (patchmatch.rs:366-368) or a source index out
 * of range. Blocks while another session of this backend is open. */
int c3d_pm_begin(c3d_pm_backend* backend, const c3d_kernel_input* input, c3d_pm_session** out);

/* Red-black iteration run on states[i] of views[i], updated in place (depth, normal, cost);
 * PatchMatchSession::run (kernel.rs:193-194, patchmatch.rs:752-799). A geometric run needs a
 * snapshot (patchmatch.rs:753-756). States must match the level size (patchmatch.rs:638-642). */
int c3d_pm_run(c3d_pm_session* session, const uint32_t* views, uint32_t num_views,
               c3d_view_state* states, const c3d_run_params* params,
               const c3d_depth_snapshot* snapshot);

/* Top-K mean cost of each state's current hypothesis into costs[i] (width*height floats);
 * PatchMatchSession::evaluate (kernel.rs:195-203, patchmatch.rs:801-824). */
int c3d_pm_evaluate(c3d_pm_session* session, uint32_t level, const uint32_t* views,
                    uint32_t num_views, const c3d_view_state* states, int geometric,
                    const c3d_depth_snapshot* snapshot, float* const* costs);

/* Final read-out: per pixel the number of sources passing every photometric and geometric
 * check, into counts[i]; PatchMatchSession::filter (kernel.rs:204-207, patchmatch.rs:842-868). */
int c3d_pm_filter(c3d_pm_session* session, const uint32_t* views, uint32_t num_views,
                  const c3d_view_state* states, const c3d_depth_snapshot* snapshot,
                  uint8_t* const* counts);

/* Joint bilateral upsampling of level-1 states `low` to `level` into `out` (caller-allocated at
 * the level size; prior and cost untouched); PatchMatchSession::upsample (kernel.rs:208-232,
 * patchmatch.rs:826-836, 931-960). */
int c3d_pm_upsample(c3d_pm_session* session, uint32_t level, const uint32_t* views,
                    uint32_t num_views, const c3d_view_state* low, float sigma_s, float sigma_c,
                    c3d_view_state* out);

/* 5x5 median plane filter of `in` into `out` (same size; cost copied);
 * PatchMatchSession::median_filter (kernel.rs:233-236, patchmatch.rs:838-840, 962-989). */
int c3d_pm_median(c3d_pm_session* session, uint32_t level, const uint32_t* views,
                  uint32_t num_views, const c3d_view_state* in, c3d_view_state* out);

/* Closes the session (drop of Session, patchmatch.rs:992-1009). */
void c3d_pm_end(c3d_pm_session* session);

/* Message of the last failure on this thread ("" if none). */
const char* c3d_pm_last_error(void);

/* ---------------------------------------------------------------- device-resident schedule */

/* Schedule of compute_depth_maps (crates/dense/src/densify.rs:350-454) executed with every state
 * resident on the device, over all `num_levels` scales of the input (0 = coarsest):
 *   scale 0: random-init photometric run of `photometric_iters` iterations;
 *   scale l > 0: joint bilateral upsampling of scale l-1 (`jbu_sigma_spatial`, `jbu_sigma_color`),
 *     the detail restorer (densify.rs:366-409): evaluate the upsampled planes (c_init), run
 *     `photometric_iters_upper` photometric iterations from them (or from a random init when
 *     `restorer_random_init`), evaluate that result (c_photo), and keep the photometric plane where
 *     the upsampled depth is not positive or c_init - c_photo > `restorer_threshold`;
 *   every scale: `geometric_rounds` rounds that each snapshot all views' depth at that scale on the
 *     device and run `geometric_iters` geometric iterations;
 *   top scale: optional 5x5 median (filter.median_filter), the filter read-out with a final
 *     snapshot, and the count threshold max(1, min(min_num_consistent, nsrc)) (densify.rs:446-454).
 * Runs at scales > 0 use the upsampled plane as the weak-texture prior when pm.weak_texture_var > 0
 * (densify.rs:366-372). LevelSchedule (params.rs:476-506) gives: fast profile 6 / 3 iterations, no
 * restorer random init, 2 geometric rounds of 2; quality profile 7 / 6, random init, 2 rounds of 6.
 * There is no host round trip between runs (the reference copies states through the host after
 * every run, patchmatch.rs:785).
 * After the threshold, `post` applies the read-out post-processing of densify.rs:455-462 (see
 * c3d_post_params) to the thresholded maps; fill_holes needs `post_k`. */

/* PostParams (crates/dense/src/postproc.rs:40-73). Reference defaults: remove_speckles 1,
 * speckle_rel_depth 0.02, speckle_min_area_frac 2e-4, speckle_keep_cost 0.2, fill_holes 0
 * (1 in the fast profile, params.rs:456), fill_max_frac 0.01, fill_edge_step 0.06,
 * fill_max_cost 1.0. */
typedef struct c3d_post_params {
    int32_t remove_speckles;
    float speckle_rel_depth;      /* neighbours join when |d_j - d_i| <= rel * min(d_i, d_j) */
    double speckle_min_area_frac; /* components below ceil(w h frac) pixels ... */
    float speckle_keep_cost;      /* ... are erased unless their median cost is <= this */
    int32_t fill_holes;
    double fill_max_frac;         /* walk at most max(1, round(diagonal * frac)) pixels */
    float fill_edge_step;         /* stop at a gray step above (int)(step * 255) */
    float fill_max_cost;          /* fill only pixels whose own raw cost is <= this */
} c3d_post_params;

typedef struct c3d_schedule {
    uint32_t photometric_iters;       /* scale 0 */
    uint32_t photometric_iters_upper; /* scales > 0 */
    uint32_t geometric_rounds;        /* per scale */
    uint32_t geometric_iters;
    uint32_t first_run_id; /* run ids are first_run_id, first_run_id+1, ... (densify.rs:345-348) */
    int32_t restorer_random_init;
    float jbu_sigma_spatial;  /* DensifyOptions::jbu_sigma_spatial [1.0] (params.rs:423) */
    float jbu_sigma_color;    /* DensifyOptions::jbu_sigma_color [0.04] */
    float restorer_threshold; /* DensifyOptions::restorer_threshold [0.1] */
    uint32_t pad0;
    c3d_post_params post;     /* both flags 0: the thresholded maps are the result */
    const double* post_k;     /* with post.fill_holes: fx, fy, cx, cy (double) of each view's top
                                 scale, 4 per entry of `views`, as the scene holds them */
} c3d_schedule;

/* Per-stage device times of the last c3d_pm_densify call, milliseconds (hipEvent on the session
 * stream). */
typedef struct c3d_timing {
    double photometric_ms;
    double geometric_ms;
    double filter_ms;
    double total_ms;
    double half_kernel_ms; /* sum over all red-black half-step kernels */
    uint32_t half_launches;
    uint32_t pad0;
    double init_ms;        /* c3d_pm_densify_learned: features, network and initial planes */
    double post_ms;        /* post-processing (c3d_post_params), host wall time of the stage */
    uint64_t post_removed; /* pixels erased by remove_speckles, all views */
    uint64_t post_filled;  /* pixels filled by fill_holes, all views */
} c3d_timing;

/* Runs the schedule for `views`; `out`, `counts` and `raw_depth` are at the top scale
 * (num_levels - 1). When `out` is not NULL the final states are written: depth/normal zeroed where the threshold fails, raw cost kept; when
 * `counts` is not NULL the filter counts are written. With both NULL the results stay on the
 * device (for throughput measurement). `timing` may be NULL. No views (num_views 0, `views` may
 * then be NULL) is a no-op. */
int c3d_pm_densify(c3d_pm_session* session, const uint32_t* views, uint32_t num_views,
                   const c3d_schedule* schedule, c3d_view_state* out, uint8_t* const* counts,
                   float* const* raw_depth, c3d_timing* timing);

/* ---------------------------------------------------------------- post-processing */

/* Small-component removal of one depth map on the device of `backend` (remove_speckles,
 * crates/dense/src/postproc.rs:77-133): 4-neighbour components of pixels with depth > 0 joined
 * when |d_j - d_i| <= speckle_rel_depth * min(d_i, d_j) (f32); a component of fewer than
 * ceil(width * height * speckle_min_area_frac) pixels whose median cost (element len/2 of its costs
 * sorted in total order) is above speckle_keep_cost gets depth 0 and normal 0. Nothing happens when
 * that area bound is 0 or 1. The remove_speckles and fill_holes flags of `params` are not read.
 * `removed` (may be NULL) receives the number of erased pixels. depth: width*height, normal:
 * 3 per pixel, cost: width*height. */
int c3d_remove_speckles(c3d_pm_backend* backend, float* depth, float* normal, const float* cost, uint32_t width,
                        uint32_t height, const c3d_post_params* params, uint64_t* removed);

/* Boundary-aware gap filling of one depth map on the device of `backend` (fill_holes,
 * postproc.rs:135-204): every pixel with depth <= 0 (or NaN) and raw_cost <= fill_max_cost walks
 * the 8 directions (+x, -x, +y, -y, +x+y, -x-y, +x-y, -x+y) for up to
 * max(1, round(sqrt(w^2 + h^2) * fill_max_frac)) steps, stopping at the border or where the gray
 * value changes by more than (int)(fill_edge_step * 255) from the previous step; the first pixel
 * with depth > 0 met in a direction is a candidate, and the candidate of least gray difference to
 * the pixel (then least distance, then earliest direction) lends its plane: the depth moves onto
 * the pixel's ray (plane_transfer, math.rs:146-154, double precision with `k` = fx, fy, cx, cy),
 * the normal and the cost are copied. Only pixels valid before the call are read. `filled` (may be
 * NULL) receives the number of filled pixels. The flags of `params` are not read. */
int c3d_fill_holes(c3d_pm_backend* backend, float* depth, float* normal, float* cost, const float* raw_cost,
                   const uint8_t* gray, const double k[4], uint32_t width, uint32_t height,
                   const c3d_post_params* params, uint64_t* filled);

/* c3d_fill_holes without any allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops). Same
 * arguments, rules and results as c3d_fill_holes, but every device buffer it needs lives in one
 * fixed device array of 88,081,408 bytes that each GPU holds from the library's load until the
 * process ends. Images above 2^21 pixels (2,097,152; one 1920 x 1080 view fits) are refused with
 * C3D_INVALID_ARGUMENT. Calls that overlap on one GPU share that array and race: serialise them. */
int c3d_fill_holes_noalloc(c3d_pm_backend* backend, float* depth, float* normal, float* cost, const float* raw_cost,
                           const uint8_t* gray, const double k[4], uint32_t width, uint32_t height,
                           const c3d_post_params* params, uint64_t* filled);

/* ---------------------------------------------------------------- autotuner (no reference counterpart) */

/* Tile autotuner: times the red-black half step with each of the C3D_TILE_SHAPES compiled workgroup
 * shapes — (rows of active pixels, lanes per pixel) = (1, 1), (2, 1), (4, 1), (4, 2); 64 / lanes
 * active pixels per wave row — on `views` at `level` (a random initialisation and two timed
 * photometric iterations on the session's scratch states, after one untimed) and keeps the fastest
 * for the session's later runs. Every shape gives identical results. rows_out / tpp_out (optional)
 * receive the choice, ms_out[C3D_TILE_SHAPES] (optional) the times of the two iterations per shape
 * in that order. */
#define C3D_TILE_SHAPES 4
int c3d_pm_autotune(c3d_pm_session* session, const uint32_t* views, uint32_t num_views, uint32_t level,
                    uint32_t* rows_out, uint32_t* tpp_out, double* ms_out);

/* ---------------------------------------------------------------- scene preparation */

/* Number of scales (num_levels, crates/dense/src/upsample.rs:36-44): halve the short side while it
 * stays >= min_size, at most max_levels, at least 1. Host only. */
uint32_t c3d_pm_num_levels(uint32_t min_side, uint32_t max_levels, uint32_t min_size);

/* Gray pyramid of one image, computed on the device of `backend` (build_pyramid,
 * crates/dense/src/upsample.rs:78-102, called by densify.rs:293): each level is the previous one
 * smoothed by the 3-tap Gaussian of blur3 (upsample.rs:45-75, edge clamp, rows then columns),
 * 2x2-averaged and rounded to 8 bits; intrinsics f <- f/2, c <- (c + 0.5)/2 - 0.5 in double.
 * Levels are in the reference's order, 0 = coarsest: out[levels-1] receives `gray` itself,
 * level l is (width >> (levels-1-l)) x (height >> (levels-1-l)) pixels (successive floor
 * halving), k_out[4l..4l+3] = fx, fy, cx, cy of level l. */
int c3d_pm_build_pyramid(c3d_pm_backend* backend, const uint8_t* gray, uint32_t width, uint32_t height,
                         const double k[4], uint32_t levels, uint8_t* const* out, float* k_out);

/* ---------------------------------------------------------------- probes (no reference counterpart) */

/* Photometric cost 1 - rho of explicit hypotheses, one per query: pixel (xy[2i], xy[2i+1]) of
 * `view` at `level` against source slot `src_slot` with plane (nx, ny, nz, d) = planes[4i..4i+3].
 * The arithmetic is the one inside c3d_pm_run (ncc_cost, patchmatch.cu:245-290); this entry
 * exists so that per-hypothesis costs can be checked against an FP64 host evaluation. */
int c3d_pm_pair_costs(c3d_pm_session* session, uint32_t level, uint32_t view, uint32_t src_slot,
                      uint32_t num_queries, const int32_t* xy, const float* planes, float* costs);

/* The per-sample reciprocal of the projective divide inside every window loop (in place of the
 * reference's `1.0f / zz`, patchmatch.cu:273), applied on the device of `backend` to n values (at
 * most 2^32): r[i] = the kernels' 1/z[i]. It exists so that this arithmetic can be checked against
 * the host restatement and IEEE division over whole binades. */
int c3d_pm_sample_reciprocal(c3d_pm_backend* backend, const float* z, float* r, uint64_t n);

#ifdef __cplusplus
}
#endif

#endif /* C3D_PM_H */
