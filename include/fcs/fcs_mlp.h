/*
 * Learned brick prefetch: a small FP32 MLP trained on the marcher's own
 * visit logs, used only to order and issue brick decodes ahead of the
 * marcher. It has no counterpart in the reference (the reference decodes the
 * visible set on demand each frame, web/scripts/cacheProcessing.js:99-232);
 * on-demand decode stays the fallback, so rendered frames never depend on
 * the network's output.
 *
 * Network: x (n_in) -> relu(x W1 + b1) (n_hidden) -> relu(. W2 + b2)
 * (n_hidden) -> . W3 + b3 (n_out logits). Loss: binary cross-entropy with
 * logits, averaged over batch x n_out. Parameters are one flat FP32 array:
 * W1 (n_in x n_hidden, row-major), b1, W2 (n_hidden x n_hidden), b2,
 * W3 (n_hidden x n_out), b3. Every product runs in hand-written HIP kernels
 * (MFMA FP32 GEMM, fused epilogues, BCE, Adam) on the context's stream.
 */
#ifndef FCS_MLP_H
#define FCS_MLP_H

#include <stddef.h>
#include <stdint.h>

#include "fcs/fcs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fcs_mlp fcs_mlp;

/* Create with He-uniform weights and zero biases drawn from `seed`
 * (SplitMix64). `max_batch` bounds every later batch. */
int fcs_mlp_create(fcs_ctx* ctx, uint32_t n_in, uint32_t n_hidden, uint32_t n_out, uint32_t max_batch,
                   uint64_t seed, fcs_mlp** out);
void fcs_mlp_destroy(fcs_mlp* m);
uint64_t fcs_mlp_num_params(const fcs_mlp* m);
int fcs_mlp_get_params(fcs_mlp* m, float* host);
int fcs_mlp_set_params(fcs_mlp* m, const float* host);
/* Gradients of the last fcs_mlp_backward, same layout as the parameters. */
int fcs_mlp_get_grads(fcs_mlp* m, float* host);

/* Forward pass on a dense host batch x (batch x n_in); logits copied to
 * `logits` (batch x n_out) when not NULL. */
int fcs_mlp_forward(fcs_mlp* m, const float* x, uint32_t batch, float* logits);
/* Forward + loss + backward on a dense host batch (y: targets in [0,1]);
 * gradients stay on the device; the loss is returned in *loss. */
int fcs_mlp_backward(fcs_mlp* m, const float* x, const float* y, uint32_t batch, float* loss);
/* One Adam step with the current gradients (bias-corrected, step count kept
 * inside the network). */
int fcs_mlp_adam(fcs_mlp* m, float lr, float beta1, float beta2, float eps);

/* Train on bit-packed samples held on the device for the call. Sample i has
 * n_pose floats poses[i*n_pose..], an input mask of (n_in - n_pose) bits and
 * a target mask of n_out bits, each packed LSB-first into 32-bit words, one
 * row of ceil(bits/32) words per sample. Mini-batches of `batch` samples in
 * an order shuffled by `seed` every epoch; Adam with `lr`, betas 0.9/0.999.
 * The mean loss of each epoch goes to epoch_loss[e] when not NULL. */
int fcs_mlp_train_masks(fcs_mlp* m, const float* poses, uint32_t n_pose, const uint32_t* in_bits,
                        const uint32_t* tgt_bits, uint32_t n_samples, uint32_t epochs, uint32_t batch, float lr,
                        uint64_t seed, float* epoch_loss);

/* Predict for one sample given as pose floats and an input mask read from a
 * device array of one u32 per mask position (nonzero = set), e.g. a frame's
 * visited-brick array. Probabilities (sigmoid of the logits) are copied to
 * `probs` (n_out floats) when not NULL. */
int fcs_mlp_predict_device_mask(fcs_mlp* m, const float* pose, uint32_t n_pose, const uint32_t* d_mask_u32,
                                float* probs);

/* One frame with learned prefetch: fcs_render_visible(cam) (on-demand decode
 * of the bricks the rays enter remains the fallback, so the image is the same
 * as without prefetch), then the
 * network reads the camera features and the frame's visited-brick mask on the
 * device (n_in = 12 + bricks, n_out = bricks) and every non-empty brick with
 * probability >= threshold that is not resident at its LOD under `cam_ahead`
 * (NULL: `cam`) is decoded ahead with fcs_prefetch_bricks, most probable
 * first, at most max_prefetch. *out_prefetched (nullable) = decodes issued;
 * out_bricks (nullable, max_prefetch entries) lists the bricks issued, one
 * entry per candidate, 0xFFFFFFFF where the pool had no room. */
int fcs_render_prefetch(fcs_volume* vol, fcs_mlp* m, const fcs_camera* cam, const fcs_camera* cam_ahead,
                        uint32_t w, uint32_t h, float threshold, uint32_t max_prefetch, fcs_frame* out,
                        uint32_t* out_prefetched, uint32_t* out_bricks);

#ifdef __cplusplus
}
#endif

#endif /* FCS_MLP_H */
