/*
 * c3d_init.h — C ABI of the learned initialiser of the MI355X PatchStereo engine.
 *
 * No counterpart in the reference (cumulus3d initialises PatchStereo randomly,
 * crates/cuda/src/kernels/patchmatch.cu:661-692). A small U-Net predicts per-pixel depth, normal
 * and confidence at a quarter of the image resolution from plane-sweep features of the reference
 * view and two sources; PatchStereo starts from these planes where the confidence is at least
 * 0.5 and from its random initialisation elsewhere, then refines.
 *
 * Network (about 1.03 M parameters): 4 scales, 32/64/128/128 channels, two 3x3 convolutions per
 * encoder scale, 2x2 mean pooling, nearest 2x upsampling with skip concatenation and one 3x3
 * convolution per decoder scale, ReLU, a 1x1 head with 5 outputs (inverse depth, normal (3),
 * confidence). Weights and activations are bf16 with fp32 accumulation on the matrix cores;
 * master weights, gradients and the Adam moments are fp32. Forward, backward and Adam are
 * hand-written HIP kernels.
 *
 * Inputs (64 channels, H x W with H and W multiples of 8): channels 0..31 are the mean absolute
 * difference |I_ref - I_src(H_k x)| / 255 over the sources for fronto-parallel planes at 32 inverse
 * depths evenly spaced between 1/depth_max and 1/depth_min (bin centres), 1 where no source sees
 * the pixel; channel 32 is I_ref / 255; the rest are 0.
 *
 * Training loss per labelled pixel: |1/d - 1/d*| / (1/dmin - 1/dmax) + w_normal (1 - n . n*) +
 * w_conf BCE(c, |d - d*| <= conf_tol d*), averaged over the labelled pixels of the batch.
 */
#ifndef C3D_INIT_H
#define C3D_INIT_H

#include <stddef.h>
#include <stdint.h>

#include "c3d_pm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct c3d_init_net c3d_init_net;

/* Geometry of one source for the plane-sweep features: x_src = r x_ref + t; k = fx, fy, cx, cy. */
typedef struct c3d_sweep_source {
    const uint8_t* gray;
    uint32_t width;
    uint32_t height;
    float k[4];
    float r[9];
    float t[3];
} c3d_sweep_source;

/* One reference view for the features: its image and intrinsics, up to 2 sources, depth range. */
typedef struct c3d_sweep_view {
    const uint8_t* gray;
    uint32_t width;
    uint32_t height;
    float k[4];
    uint32_t num_sources;
    c3d_sweep_source sources[2];
    float depth_min;
    float depth_max;
} c3d_sweep_view;

/* Training hyper-parameters. */
typedef struct c3d_init_train_params {
    float lr;
    float beta1;
    float beta2;
    float eps;
    float w_normal;
    float w_conf;
    float conf_tol;
    uint32_t pad0;
} c3d_init_train_params;

/* Allocates the network on the device of `backend` for batches of up to max_batch images of
 * height x width (multiples of 8), with weights drawn from a He-normal distribution seeded by
 * `seed`. The network holds the backend until c3d_init_destroy: a c3d_pm_destroy in between
 * defers the backend's release to that call. */
int c3d_init_create(c3d_pm_backend* backend, uint32_t max_batch, uint32_t height, uint32_t width, uint64_t seed,
                    c3d_init_net** out);
/* Frees the network and drops its hold on the backend (freeing the backend if c3d_pm_destroy
 * was called while the network was open). */
void c3d_init_destroy(c3d_init_net* net);
/* Number of fp32 parameters (weights then bias of each layer, in layer order; weights [cout][k*k][cin]). */
uint64_t c3d_init_num_params(const c3d_init_net* net);
int c3d_init_get_params(c3d_init_net* net, float* out);
int c3d_init_set_params(c3d_init_net* net, const float* in);
/* Plane-sweep features of views[i] into batch slot slot0 + i (computed on the device). */
int c3d_init_features(c3d_init_net* net, uint32_t slot0, uint32_t num_views, const c3d_sweep_view* views);
/* Explicit features (bf16 bit patterns, [slot][H][W][64]) for num_slots slots from slot 0. */
int c3d_init_set_features(c3d_init_net* net, uint32_t num_slots, const uint16_t* feat);
/* The features of `slot` (bf16 bit patterns, [H][W][64]). */
int c3d_init_get_features(c3d_init_net* net, uint32_t slot, uint16_t* feat);
/* Labels of slot0 + i: depth[H*W] (0 = none) and normal[3*H*W] at the network size, the valid image
 * size (the rest is padding) and the depth range of the slot. */
int c3d_init_set_labels(c3d_init_net* net, uint32_t slot, const float* depth, const float* normal, uint32_t valid_width,
                        uint32_t valid_height, float depth_min, float depth_max);
/* Forward, loss, backward over slots [0, batch). losses[4] = summed depth, normal and confidence
 * terms and the labelled pixel count; grads (optional) receives the fp32 parameter gradient of the
 * summed loss; with `step` nonzero an Adam step is taken (gradient averaged over labelled pixels). */
int c3d_init_train_step(c3d_init_net* net, uint32_t batch, const c3d_init_train_params* params, int step, float* losses,
                        float* grads);
/* Forward over slots [0, batch): depth, normal (3 per pixel, facing the camera ray is not enforced)
 * and confidence at the network size; `head` (optional) receives the raw head outputs [p][5]. */
int c3d_init_predict(c3d_init_net* net, uint32_t batch, float* depth, float* normal, float* conf, float* head);

/* ---------------------------------------------------------------- learned-init PatchStereo */

/* Schedule after the learned initialisation (top scale of the session input): `refine_iters`
 * photometric iterations from the predicted planes, with the reference's random initialisation
 * (patchmatch.cu:668-676) where the confidence is below conf_min; then `geometric_rounds` rounds
 * of `geometric_iters` geometric iterations; the read-out filter and count threshold
 * (densify.rs:428-454). */
typedef struct c3d_learned_schedule {
    uint32_t refine_iters;
    uint32_t geometric_rounds;
    uint32_t geometric_iters;
    uint32_t first_run_id;
    float conf_min;
    uint32_t pad0;
    c3d_post_params post;  /* read-out post-processing, as in c3d_schedule */
    const double* post_k;
} c3d_learned_schedule;

/* c3d_pm_densify with the learned initialiser in place of the random initialisation, all on the
 * device: quarter-resolution images (two gray-pyramid steps) of the views and of their first two
 * sources, plane-sweep features, the network over batches of `net`'s size, planes at full
 * resolution, then the schedule above. The network's input size must hold the quarter-resolution
 * images. Outputs as c3d_pm_densify; init_depth / init_normal (optional, per view) receive the
 * initial planes (depth 0 = random initialisation). No views is a no-op. */
int c3d_pm_densify_learned(c3d_pm_session* session, c3d_init_net* net, const uint32_t* views, uint32_t num_views,
                           const c3d_learned_schedule* schedule, c3d_view_state* out, uint8_t* const* counts,
                           float* const* raw_depth, float* const* init_depth, float* const* init_normal,
                           c3d_timing* timing);

#ifdef __cplusplus
}
#endif

#endif /* C3D_INIT_H */
