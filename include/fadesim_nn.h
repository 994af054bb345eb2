/*
 * fadesim_nn.h -- C ABI of the learned warm start's GPU kernels.
 *
 * The warm start has no counterpart in the reference: it feeds the batched
 * solver (fadesim_solve_batch's d_warm) an initial drop field in place of the
 * reference's initial guess u = vin, d = vout
 * (crossbarHspice/V2/concrete_command.h:1171-1178).  Its training data are
 * solutions of fadesim_solve_batch only.  The related reference component is
 * the SCN conductance predictor (pimtorch/pimfixedpoint/SCN.py:127-153),
 * trained on solver output as well (SCN_data_gen.py:134-179).
 *
 * Tensors are device pointers; activations are NHWC bf16 ("void *": the
 * bf16 storage of __bf16 / torch.bfloat16), weights bf16 [Cout][Kp] with
 * k = (3*ky + kx)*Cin + ci, the bias in column 9*Cin and zero columns up to
 * Kp = roundup(9*Cin + 1, 32).  All calls are stream-ordered.
 */
#ifndef FADESIM_NN_H
#define FADESIM_NN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physics features of B crossbars (uniform wire r, grounded bitlines):
 * first- and second-order superposition drop fields.  feat [B][M][N][8]
 * bf16, base [B][2][M][N] FP32 (second-order p, q), scale [B] FP32. */
int fadesim_nn_features(int B, int M, int N, const float *d_g, const double *d_r, const double *d_vin,
                        void *d_feat, float *d_base, float *d_scale, void *stream);

/* The same features averaged over 2^level x 2^level cells (level 2,
 * 128x128 crossbars): feat [B][M>>2][N>>2][8] bf16, base [B][2][M][N] FP32
 * (or NULL: not written -- fadesim_solve_batch forms it itself when given
 * d_warm_coarse without d_warm), scale [B] -- the input of a network whose
 * first layer works at level 2. */
int fadesim_nn_features_pooled(int B, int M, int N, int level, const float *d_g, const double *d_r,
                               const double *d_vin, void *d_feat, float *d_base, float *d_scale, void *stream);

/* warm = base + scale * out[..., 0:2]  ->  [B][2][M][N] FP32 */
int fadesim_nn_combine(int B, int M, int N, const float *d_base, const float *d_scale, const void *d_out, int cout,
                       float *d_warm, void *stream);

/* warm = base + scale * bilinear(out[..., 0:2]) for a network output 2^level
 * times coarser than the crossbar (out [B][M>>level][N>>level][cout]) ->
 * [B][2][M][N] FP32; sample points at the fine cells' centres. */
int fadesim_nn_combine_up(int B, int M, int N, const float *d_base, const float *d_scale, const void *d_out,
                          int cout, int level, float *d_warm, void *stream);

/* The coarse32f (widths 32 / 64) or coarse16f (16 / 32) network of B
 * crossbars in one launch, one workgroup per crossbar with every activation
 * in LDS: x [B][32][32][8] bf16 (pooled features) -> out [B][32][32][16]
 * bf16.  d_w: 7 layer weight pointers (e2a, e2, e3a, e3, t3, d2, out) in
 * fragment-major order [Kp/32][Cout][32] (w'[kb][co][kk] = w[co][32 kb + kk]
 * of the [Cout][Kp] layout), kp: their padded depths (the width is read from
 * kp[1]). */
int fadesim_nn_coarse_forward(int B, const void *d_x, const void *const *d_w, const int *kp, void *d_out,
                              void *stream);

/* fadesim_solve_batch (include/fadesim.h) with the coarse16f warm start formed
 * inside the solve: per crossbar, the pooled features and base field of
 * fadesim_nn_features_pooled, the network of fadesim_nn_coarse_forward and the
 * combine at load of d_warm_coarse (level 2, 16 channels), all in the
 * guess-first kernel with the activations in LDS -- the same outputs, bit for
 * bit, as those three steps followed by fadesim_solve_batch with d_warm = base,
 * d_warm_coarse = out, d_warm_scale = scale, and no base field, features or
 * network output written to HBM.  Like the reference's initial guess
 * (crossbarHspice/V2/concrete_command.h:1171-1178), which it replaces, it feeds
 * the solve of NodebasedGSMethod::nodeBaseGS named at fadesim_solve_batch.
 * Needs rows = cols = 128, FADESIM_F_TEST_FIRST (not FIXED_SWEEPS or
 * NO_FASTPATH), and d_warm, d_warm_coarse, d_vout NULL (grounded bitlines);
 * d_w, kp: the 7 fragment-major layer weights and depths as for
 * fadesim_nn_coarse_forward, width 16.  The library keeps a workspace of
 * batch x 128 KB for the starts of crossbars that need sweeps. */
struct fadesim_batch_desc;
int fadesim_nn_solve_batch(const struct fadesim_batch_desc *desc, const void *const *d_w, const int *kp,
                           void *stream);

/* 3x3 same convolution (MFMA), optional ReLU: y [B][H][W][Cout] */
int fadesim_nn_conv3x3_fwd(int B, int H, int W, int Cin, int Cout, int Kp, const void *d_x, const void *d_w, void *d_y,
                           int relu, void *stream);

/* Flipped, transposed weights for the data gradient: wT [Cin][KpT] */
int fadesim_nn_flip_weights(int Cin, int Cout, int Kp, int KpT, const void *d_w, void *d_wT, void *stream);

/* Weight (and bias) gradient, accumulated: dw [Cout][Kp] FP32 += col^T dz */
int fadesim_nn_conv3x3_wgrad(int B, int H, int W, int Cin, int Cout, int Kp, const void *d_x, const void *d_dz,
                             float *d_dw, void *stream);

int fadesim_nn_pool2(int B, int H, int W, int C, const void *d_x, void *d_y, void *stream);
int fadesim_nn_pool2_bwd(int B, int H, int W, int C, const void *d_dy, int accumulate, void *d_dx, void *stream);
/* u [B][2H][2W][C] = nearest-upsample(t [B][H][W][C]) + e */
int fadesim_nn_up2_add(int B, int H, int W, int C, const void *d_t, const void *d_e, void *d_u, void *stream);
/* dt [B][H][W][C] = 2x2 block sums of du */
int fadesim_nn_up2_bwd(int B, int H, int W, int C, const void *d_du, void *d_dt, void *stream);
/* dy = (dy + add) * (y > 0); add may be NULL */
int fadesim_nn_relu_mask(size_t n, void *d_dy, const void *d_add, const void *d_y, void *stream);
/* c = a + b (bf16, FP32 arithmetic) */
int fadesim_nn_add(size_t n, const void *d_a, const void *d_b, void *d_c, void *stream);
/* loss += mean squared error of out[..., 0:2] vs target [B][2][M][N] / 2; dout */
int fadesim_nn_mse(int B, int M, int N, const void *d_out, int cout, const float *d_target, void *d_dout,
                   float *d_loss, void *stream);
/* Adam step on FP32 master weights; refreshes the bf16 copy wb */
int fadesim_nn_adam(size_t n, float *d_w, float *d_m, float *d_v, const float *d_grad, void *d_wb, float lr,
                    float beta1, float beta2, float eps, int step, float weight_decay, float grad_scale,
                    void *stream);
/* src_bf16 ? bf16 -> f32 : f32 -> bf16 */
int fadesim_nn_cast(size_t n, const void *d_src, int src_bf16, void *d_dst, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FADESIM_NN_H */
