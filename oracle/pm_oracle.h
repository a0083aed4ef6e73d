/*
 * pm_oracle.h — CPU oracle for the PatchStereo hot path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only as the checker (or as the timed CPU baseline); the product path never calls it.
 *
 * It restates, function by function, the rules of the reference's CUDA PatchStereo backend
 * (crates/cuda/src/kernels/patchmatch.cu, crates/cuda/src/patchmatch.rs) and of the trait's
 * normative text (crates/dense/src/kernel.rs:32-74), with the host-side formulas of
 * crates/dense/src/{math,upsample,densify,params}.rs. Each function cites the lines it follows.
 *
 * Two arithmetic modes:
 *  - FP32: the device's operation order (explicit fmaf where the device uses FMA, IEEE division
 *    and square root, the same polynomial exp and sin/cos, source samples summed in gray-level
 *    units shifted by the reference centre value). Built with -ffp-contract=off.
 *  - FP64: the reference's NCC formula in double with libm exp, used as the cost oracle
 *    (pmo_ncc_cost_f64).
 * The input structures are the POD types of include/c3d_pm.h (data only).
 *
 * Pinning (what fixes each part independently of this restatement) is listed per function in
 * pm_oracle.c and summarised at the top of that file.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#include "../include/c3d_pm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmo_ctx pmo_ctx;

/* integer primitives (math.rs:31-47, kernel.rs:248-259) */
uint64_t pmo_hash_bytes(const uint8_t* b, size_t n);
uint64_t pmo_mix64(uint64_t z);
uint64_t pmo_rng_state(uint64_t seed, uint64_t key, uint32_t level, uint32_t run, uint32_t t,
                       uint32_t half, uint64_t pixel);
float pmo_rng_next(uint64_t* state);

/* math.rs twins (double) */
double pmo_erf(double x);
double pmo_emission_norm(double sigma);
double pmo_emission(double cost, double sigma);
double pmo_visibility_probability(double cost, double sigma);
double pmo_triangulation_prior(double cos_theta, double min_angle_rad);
double pmo_incident_prior(double cos_phi, double sigma);
double pmo_resolution_prior(const double h[9], double x, double y, double radius);
int pmo_plane_transfer(double dq, const double n[3], const double ray_q[3],
                       const double ray_p[3], double* d_out);
double pmo_quantile_sorted(const double* v, size_t n, double q);
float pmo_bilinear_clamped(const float* data, size_t w, size_t h, float x, float y);

/* the device's elementary functions, restated */
/* per-sample reciprocal of the projective divide: 1 = the device's reciprocal from a bit-pattern
 * seed, one third-order and one Newton step (default), 0 = IEEE division */
void pmo_set_rcp_mode(int mode);
/* the device's reciprocal (mode 1) of n floats */
void pmo_rcp_sample(const float* z, float* r, size_t n);
float pmo_expf(float x);
void pmo_sincosf(float x, float* s, float* c);

/* homography constants A = K_j R K^-1, b = K_j t (patchmatch.rs:131-151) */
void pmo_level_pair(const float kr[4], const float kj[4], const float r[9], const float t[3],
                    float a_out[9], float b_out[3]);

/* pyramid helpers (upsample.rs:35-102, 240-260) */
uint32_t pmo_num_levels(uint32_t min_side, uint32_t max_levels, uint32_t min_size);
/* Writes levels-1 coarser images of `gray` (w x h) into out (caller-sized, coarsest last);
 * returns the intrinsics of each level in k_out[4*l] for l = 0 (input) .. levels-1 (coarsest). */
void pmo_build_pyramid(const uint8_t* gray, uint32_t w, uint32_t h, const double k[4],
                       uint32_t levels, uint8_t* const* out, float* k_out);

/* decision rules of the half step, run start and read-out over tables (pm_oracle.c); the
 * per-pixel kernels call these with the costs they compute */
typedef struct pmo_select_params {
    float tau_good, tau_bad, inv2beta2, prev_w, lambda;
    int32_t n1, n2;
} pmo_select_params;
typedef struct pmo_readout_params {
    float cos_filter_tri, ncc_norm, inv2ncc, q_min, filter_gmax;
} pmo_readout_params;
int pmo_region_pick(const float* cost, int w, int h, int x, int y, int region, int* qx, int* qy);
int pmo_transfer_plane(const float q[4], const float ray_q[3], const float ray[3], float out[4]);
void pmo_select_views(const pmo_select_params* p, int nsrc, int vprev, const float* prior, const float* m,
                      const float* g, float* wv, float* den, float* agg, int* best, int* view);
int pmo_refine_accept(const pmo_select_params* p, int ncand, int nsrc, const float* wv, float den,
                      const float* m, const float* g, float* bcost);
void pmo_perturb_normal_det(const float a[3], float u, const float n[3], const float ray[3], float phi,
                            float out[3]);
float pmo_topk_mean(const float* cost, int n, int k);
int pmo_readout_source(const pmo_readout_params* k, const float X[3], float lx, const float c[3],
                       const float n[3], int hok, float m, float geo);
uint32_t pmo_count_need(uint32_t min_num_consistent, uint32_t nsrc);

/* context over a kernel input (borrowed; must outlive the context) */
pmo_ctx* pmo_create(const c3d_kernel_input* input, int num_threads);
void pmo_destroy(pmo_ctx* ctx);
const char* pmo_last_error(void);

/* trait operations, same semantics as the C ABI entry points */
int pmo_run(pmo_ctx* ctx, const uint32_t* views, uint32_t n, c3d_view_state* states,
            const c3d_run_params* p, const c3d_depth_snapshot* snap);
int pmo_evaluate(pmo_ctx* ctx, uint32_t level, const uint32_t* views, uint32_t n,
                 const c3d_view_state* states, int geometric, const c3d_depth_snapshot* snap,
                 float* const* costs);
int pmo_filter(pmo_ctx* ctx, const uint32_t* views, uint32_t n, const c3d_view_state* states,
               const c3d_depth_snapshot* snap, uint8_t* const* counts);
int pmo_upsample(pmo_ctx* ctx, uint32_t level, const uint32_t* views, uint32_t n,
                 const c3d_view_state* low, float sigma_s, float sigma_c, c3d_view_state* out);
int pmo_median(pmo_ctx* ctx, uint32_t level, const uint32_t* views, uint32_t n,
               const c3d_view_state* in, c3d_view_state* out);
/* scene preparation (scene_oracle.c; neighbors.rs:47-189); the struct is include/c3d_scene.h's */
struct c3d_sparse_scene;
int pmo_pair_stats(const struct c3d_sparse_scene* s, uint32_t* count, double* q75);
int pmo_depth_ranges(const struct c3d_sparse_scene* s, double* dmin, double* dmax, uint8_t* has);

/* read-out post-processing (post_oracle.c; postproc.rs:77-204) */
int pmo_remove_speckles(float* depth, float* normal, const float* cost, uint32_t w, uint32_t h,
                        const c3d_post_params* p, uint64_t* removed);
int pmo_fill_holes(float* depth, float* normal, float* cost, const float* raw_cost, const uint8_t* gray,
                   const double k[4], uint32_t w, uint32_t h, const c3d_post_params* p, uint64_t* filled);

int pmo_densify(pmo_ctx* ctx, const uint32_t* views, uint32_t n, const c3d_schedule* sch,
                c3d_view_state* out, uint8_t* const* counts, float* const* raw_depth);

/* Restricted run for timing samples: like pmo_run with random_init on one view, but only rows
 * [row0, row1) are updated (other rows keep their initial state). */
int pmo_run_rows(pmo_ctx* ctx, uint32_t view, c3d_view_state* state, const c3d_run_params* p,
                 uint32_t row0, uint32_t row1);

/* single-hypothesis costs: FP32 (device arithmetic) and FP64 (reference formula in double) */
float pmo_ncc_cost_f32(pmo_ctx* ctx, uint32_t level, uint32_t view, uint32_t src_slot, int x,
                       int y, const float plane[4]);
double pmo_ncc_cost_f64(pmo_ctx* ctx, uint32_t level, uint32_t view, uint32_t src_slot, int x,
                        int y, const double plane[4]);

#ifdef __cplusplus
}
#endif

#endif
