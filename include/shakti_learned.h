/*
 * shakti_learned.h -- C ABI of the learned part of libshakti_hip.so: a 4-level
 * convolutional U-Net (16/32/64/128 channels, 3x3 convolutions, LeakyReLU,
 * BF16 weights and activations, FP32 accumulation on the gfx950 matrix cores)
 * with hand-written forward, backward and Adam kernels, and the hooks that
 * let the head solver record training pairs and use the network as a warm
 * start (shk_solver_cfg.form = SHK_FORM_LEARNED).
 *
 * The reference has no learned component; nothing here replaces a reference
 * interface.  The solver-side hooks sit behind the same plug-in seam as the
 * rest of the library (solve_elliptic_linear_system!,
 * src/linear_solver.jl:915-925).
 *
 * What the network computes.  For the head system A e = r of one Picard
 * iteration (A in the symmetric face layout aP, aE, aN; r = rhs - A h_prev)
 * the input image holds, per cell, the Jacobi-scaled residual and couplings
 *   ch0  r / (sqrt(aP) s),  s = rms of r / sqrt(aP) over the grid
 *   ch1  2 aE / sqrt(aP aP_east),   ch2  the same for the west face
 *   ch3  2 aN / sqrt(aP aP_north),  ch4  the same for the south face
 *   ch5  c / aP (c: reaction plus Dirichlet-face conductance, so 1 on identity rows)
 *   ch6, ch7  zero
 * and the network output N approximates, up to a scale, sqrt(aP) e / s; the
 * warm start is h_prev + gamma e_N with e_N = s N / sqrt(aP) and gamma the
 * energy-minimising step gamma = (e_N . r) / (e_N . A e_N), computed in FP64.
 *
 * Layout.  Images are NHWC: element (b, y, x, c) at ((b*H + y)*W + x)*C + c,
 * with y the grid's iy and x its ix.  H must be a multiple of 32 and W of 64
 * (three 2x2 poolings, and rows of 8 pixels at the coarsest level for the
 * weight-gradient kernel); the solver pads its nx x ny grid with zero cells
 * up to those multiples.  Network input: BF16, 8 channels.  Output: FP32, one
 * channel.  Device pointers are plain device addresses (hipMalloc'd or from
 * any allocator on the same device).
 *
 * Parameters.  One flat FP32 vector, layer after layer in the order of
 * enum shk_unet_layer; a 3x3 layer stores W[cout][tap][cin] (tap = 3*(dy+1) +
 * (dx+1)) followed by bias[cout]; the 1x1 head stores w[16] then its bias.
 *
 * Conventions as in shakti_hip.h: SHK_OK or a negative SHK_ERR_* code, the
 * message from shk_unet_last_error; calls return after the network's stream
 * has synchronised; one network is used by one host thread at a time.
 */
#ifndef SHAKTI_LEARNED_H
#define SHAKTI_LEARNED_H

#include <stddef.h>
#include <stdint.h>

#include "shakti_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shk_unet shk_unet;

/* Capacity and activation slope.  Buffers for every activation and gradient
 * are sized for max_batch images of max_h x max_w at creation. */
typedef struct {
    int32_t max_batch;    /* 1 .. 256 */
    int32_t max_h, max_w; /* multiples of 32 and 64 */
    float leaky_slope;    /* LeakyReLU slope for negative inputs (0.1) */
    int32_t loss_kind;    /* SHK_LOSS_L2 or SHK_LOSS_ENERGY (shk_unet_loss) */
} shk_unet_cfg;

#define SHK_LOSS_L2 0     /* 1 - (N.t)^2 / (|N|^2 |t|^2) */
#define SHK_LOSS_ENERGY 1 /* 1 - (N.r)^2 / (E(N) E(t)), E the energy of the input's operator */

/* Layers, in parameter order. */
enum shk_unet_layer {
    SHK_L_ENC0A = 0, SHK_L_ENC0B, SHK_L_ENC1A, SHK_L_ENC1B, SHK_L_ENC2A, SHK_L_ENC2B, SHK_L_ENC3A,
    SHK_L_ENC3B, SHK_L_DEC2A, SHK_L_DEC2B, SHK_L_DEC1A, SHK_L_DEC1B, SHK_L_DEC0A, SHK_L_DEC0B, SHK_L_HEAD,
    SHK_L_COUNT
};

/* Internal tensors that shk_unet_download_tensor can copy out (as FP32),
 * for checks of each layer.  Level l images are (H >> l) x (W >> l).
 * Forward: x (input), a/s encoder outputs, p pooled, t3 bottleneck, u/v
 * decoder outputs, out.  Backward: after shk_unet_backward, da*, ds*, dt3,
 * du*, dv* hold the gradient of the loss with respect to the pre-activation
 * of the layer that produced the tensor of the same name (the activation's
 * derivative is applied in place); dp*, dcat* and dout are gradients of those
 * tensors themselves (dcat*: the concatenated decoder inputs [upsampled
 * deeper tensor, skip]). */
enum shk_unet_tensor {
    SHK_T_X = 0, SHK_T_A0, SHK_T_S0, SHK_T_P0, SHK_T_A1, SHK_T_S1, SHK_T_P1, SHK_T_A2, SHK_T_S2, SHK_T_P2,
    SHK_T_A3, SHK_T_T3, SHK_T_U2, SHK_T_V2, SHK_T_U1, SHK_T_V1, SHK_T_U0, SHK_T_V0, SHK_T_OUT,
    SHK_T_DOUT, SHK_T_DV0, SHK_T_DU0, SHK_T_DCAT0, SHK_T_DV1, SHK_T_DU1, SHK_T_DCAT1, SHK_T_DV2, SHK_T_DU2,
    SHK_T_DCAT2, SHK_T_DT3, SHK_T_DA3, SHK_T_DP2, SHK_T_DS2, SHK_T_DA2, SHK_T_DP1, SHK_T_DS1, SHK_T_DA1,
    SHK_T_DP0, SHK_T_DS0, SHK_T_DA0,
    SHK_T_COUNT
};

int shk_unet_create(const shk_unet_cfg* cfg, shk_unet** out);
int shk_unet_destroy(shk_unet* net);
const char* shk_unet_last_error(const shk_unet* net);

/* Number of FP32 parameters (the same for every network). */
int64_t shk_unet_num_params(void);
/* Offset of a layer's weights in the flat vector and its channel counts
 * (cin counts both concatenated inputs of a decoder layer). */
int shk_unet_layer_info(int32_t layer, int64_t* offset, int32_t* cin, int32_t* cout);

/* He-normal initialisation from a seed (biases zero), or copy in / out. */
int shk_unet_init_params(shk_unet* net, uint64_t seed);
int shk_unet_set_params(shk_unet* net, const float* host, int64_t n);
int shk_unet_get_params(shk_unet* net, float* host, int64_t n);
/* Gradient of the last shk_unet_backward / shk_unet_train_step. */
int shk_unet_get_grads(shk_unet* net, float* host, int64_t n);

/* Forward pass: x_dev BF16 [B][H][W][8] -> y_dev FP32 [B][H][W] (y_dev may be
 * NULL; the output stays in the network's SHK_T_OUT tensor). */
int shk_unet_forward(shk_unet* net, const void* x_dev, int32_t B, int32_t H, int32_t W, float* y_dev);

/* Scale-free loss of the last forward output N against target t (FP32
 * [B][H][W]), averaged over images.  SHK_LOSS_L2: 1 - (N.t)^2 / (|N|^2 |t|^2),
 * the relative squared error left after the best scaling of N.
 * SHK_LOSS_ENERGY: the same in the energy norm of the Jacobi-scaled operator
 * that the last forward's input carries, evaluated in flux form,
 * E(u) = sum_cells c u^2 + sum_faces a_f (u_i - u_j)^2, so
 * 1 - (N.r)^2 / (E(N) E(t)) with r = input channel 0 (t solves A t = r).
 * Writes the loss to *loss and, when with_grad, its gradient to the
 * SHK_T_DOUT tensor. */
int shk_unet_loss(shk_unet* net, const float* t_dev, int32_t with_grad, double* loss);

/* Backward pass of the last forward from dout_dev (FP32 [B][H][W]; NULL: use
 * the SHK_T_DOUT tensor written by shk_unet_loss): parameter gradient. */
int shk_unet_backward(shk_unet* net, const float* dout_dev);

/* One Adam step (FP32 master weights and moments) on the last gradient,
 * then the BF16 copies used by the kernels are refreshed. */
int shk_unet_adam(shk_unet* net, float lr, float beta1, float beta2, float eps);

/* forward + loss + backward + Adam on one batch; returns the batch loss. */
int shk_unet_train_step(shk_unet* net, const void* x_dev, const float* t_dev, int32_t B, int32_t H, int32_t W,
                        float lr, double* loss);

/* Copy an internal tensor of the last pass to host, as FP32; count must be
 * its element count for the last pass's B, H, W. */
int shk_unet_download_tensor(shk_unet* net, int32_t tensor, float* host, int64_t count);

/* ---- device buffers for inputs, targets and recorded pairs ---------------- */

/* Plain device memory on the current HIP device, for callers that have no
 * allocator of their own on this HIP runtime.  kind: 0 host to device, 1
 * device to host, 2 device to device.  Synchronous. */
int shk_dev_alloc(size_t bytes, void** out);
int shk_dev_free(void* ptr);
int shk_dev_copy(void* dst, const void* src, size_t bytes, int32_t kind);

/* ---- solver hooks (shk_ctx from shakti_hip.h) ---------------------------- */

/* Attach a network to a context (NULL detaches).  The network then runs on
 * the context's stream; it must be large enough for one image of the
 * context's grid padded to multiples of 32 (ny) and 64 (nx), and must not be
 * used elsewhere while attached. */
int shk_attach_unet(shk_ctx* ctx, shk_unet* net);

/* Learned preconditioner inside flexible CG (form SHK_FORM_LEARNED with the
 * multigrid or Chebyshev preconditioner; no counterpart in the reference,
 * whose CG takes a fixed preconditioner).  every = 0: the network gives the
 * warm start only (default).  every >= 2 (rounded up to even): the solve
 * queues batches of `every` CG iterations; the first iteration of each batch
 * preconditions with z = g e + P(r - g A e), e the network's correction for
 * r, g = e'r / e'Ae and P the fixed preconditioner, and that iteration and
 * the next take the flexible beta = -z'Ap / p'Ap.  A learned batch whose
 * final preconditioned residual norm is not below stall_ratio times the one
 * at its start ends the learned batches: the solve goes on as plain PCG and
 * shk_solve_stats.fallback gets bit 1 (value 2).  The stopping test is the
 * fixed preconditioner's, checked on the iterations that use it.
 * every in {0} U [2, 4096], stall_ratio in [0, 1000]; 0 makes every learned
 * batch count as stalled (the first one ends them). */
int shk_set_learned_precond(shk_ctx* ctx, int32_t every, double stall_ratio);

/* Record training pairs: from now on every elliptic solve of the context
 * writes, into slot k = (number recorded so far), the network input of its
 * Picard iteration (BF16 [H8][W8][8] at x_dev + k*H8*W8*8) and the target
 * sqrt(aP) (h_new - h_prev) / s (FP32 [H8][W8] at t_dev + k*H8*W8), until
 * `capacity` pairs are held.  H8, W8: ny and nx rounded up to multiples of 32
 * and 64.  x_dev == NULL stops recording; shk_recorded_pairs returns the
 * number of pairs written since recording started. */
int shk_record_pairs(shk_ctx* ctx, void* x_dev, float* t_dev, int32_t capacity);

/* Record training pairs for the learned preconditioner: from now on every
 * learned CG iteration (form SHK_FORM_LEARNED with shk_set_learned_precond
 * every >= 2) writes its network input (the Jacobi-scaled residual r_k /
 * (sqrt(aP) s_k) with the coupling channels) into the next slot, and when the
 * solve has converged the target sqrt(aP) (x_final - x_k) / s_k, the
 * remaining error of that iteration.  Same layout and capacity rules as
 * shk_record_pairs; the library holds capacity x n doubles of iterates until
 * recording stops.  Solves run uncaptured while recording.  Only one kind of
 * recording is active at a time; x_dev == NULL stops either. */
int shk_record_cg_pairs(shk_ctx* ctx, void* x_dev, float* t_dev, int32_t capacity);
int shk_recorded_pairs(shk_ctx* ctx, int32_t* recorded);

#ifdef __cplusplus
}
#endif
#endif /* SHAKTI_LEARNED_H */
