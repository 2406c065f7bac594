/*
 * isac_posenet.h -- kernels of the learned pose-proposal network.
 *
 * No counterpart in the reference: ucdavis/GPU-ISAC searches every reference,
 * shift and psi exhaustively (cuda/gpu_aln_noref.cu:326-554).  These entry
 * points are the forward, backward and optimiser kernels of a small CNN that
 * proposes (reference, psi bin, shift) candidates for a narrowed search
 * (isac_mref_proposal_device in gpu_aln_pack.h).  They serve the Python host
 * layer that trains the network and runs it (isac_amd/posenet.py).
 *
 * Conventions: every pointer is a device pointer; activations are bf16
 * (stored as uint16_t) in NHWC order; accumulation is fp32.  Master parameters
 * of a layer are fp32 [Cout][K+1] with K = ks*ks*Cin: row co holds the weights
 * in (tap, ci) order, tap = ky*ks + kx, followed by the bias.  `stream` is a
 * hipStream_t (NULL = the null stream).  Every call returns 0, or an error
 * code (see gpu_last_error) for bad arguments or a failed launch; no call
 * synchronises.
 */
#ifndef ISAC_AMD_POSENET_H
#define ISAC_AMD_POSENET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* x0[i][y][x] = bf16((img - mean_i) / (std_i + 1e-6)), per image over its D*D
 * pixels (sums in FP64). */
int isac_nn_standardize(const float *img, int n, int D, uint16_t *x0, void *stream);

/* First layer: 3x3 convolution of a one-channel bf16 image, zero padding 1,
 * bias and ReLU: y [n][H][W][Cout] bf16.  Cout must be 32. */
int isac_nn_conv1_fwd(const uint16_t *x0, int n, int H, int W, const float *master, int Cout,
                      uint16_t *y, void *stream);

/* ks x ks convolution (ks = 1 or 3, zero padding ks/2) on MFMA: x [n][H][W][Cin]
 * bf16, w [Cout][ks*ks][Cin] bf16; y = x * w (+ bias[co * bias_stride]) with an
 * optional ReLU, stored bf16 (out_f32 = 0) or fp32 [n][H][W][Cout].  Cin and Cout
 * must be multiples of 16.  A ks = 1 layer on a 1x1 image is a fully connected
 * layer.  With wflip from isac_nn_pack_weights and no bias / ReLU it computes
 * the input gradient of a layer. */
int isac_nn_conv_fwd(const uint16_t *x, int n, int H, int W, int Cin, const uint16_t *w,
                     const float *bias, int bias_stride, int Cout, int ks, int relu, int out_f32,
                     void *y, void *stream);

/* Inference forms with the 2x2 max pool fused (results equal the unfused
 * convolution followed by isac_nn_maxpool2): conv1 (Cout 32) and the 3x3 MFMA
 * convolution with bias and ReLU; a [n][H/2][W/2][Cout] bf16. */
int isac_nn_conv1_pool_fwd(const uint16_t *x0, int n, int H, int W, const float *master, int Cout,
                           uint16_t *a, void *stream);
int isac_nn_conv_pool_fwd(const uint16_t *x, int n, int H, int W, int Cin, const uint16_t *w,
                          const float *bias, int bias_stride, int Cout, uint16_t *a, void *stream);

/* 2x2 max pooling, stride 2, floor: a [n][H/2][W/2][C]; C a multiple of 8. */
int isac_nn_maxpool2(const uint16_t *y, int n, int H, int W, int C, uint16_t *a, void *stream);

/* Gradient through the pooling and the ReLU before it: dz [n][H][W][C] holds
 * da at the first maximum of each window (when that maximum is > 0), else 0. */
int isac_nn_maxpool2_bwd(const uint16_t *y, const uint16_t *da, int n, int H, int W, int C,
                         uint16_t *dz, void *stream);

/* dz = dy where y > 0, else 0 (count elements, bf16). */
int isac_nn_relu_bwd(const uint16_t *y, const uint16_t *dy, size_t count, uint16_t *dz,
                     void *stream);

/* Weight and bias gradient of a ks x ks layer (Cin >= 1, any Cout):
 * g [Cout][K+1] fp32 = sum over pixels of dz[p][co] * x-patch[p][k] (the last
 * column sums dz: the bias gradient).  part is scratch of part_bytes; the
 * pixel range is split into as many partial sums as fit (fixed order, so the
 * result is deterministic). */
int isac_nn_conv_wgrad(const uint16_t *x, const uint16_t *dz, int n, int H, int W, int Cin,
                       int Cout, int ks, float *g, float *part, size_t part_bytes, void *stream);

/* Softmax cross-entropy over three heads of logits [n][ld] (fp32): columns
 * [0, h0) references, [h0, h0+h1) psi classes (h1 = 2 n_psi, mirror-major),
 * [h0+h1, h0+h1+h2) shifts (x-major grid, shift_ny per row); labels [n][3].
 * Targets are one-hot, or with sigma_psi > 0 a circular Gaussian (sigma in
 * bins) over the psi bins of the labelled mirror and with sigma_shift > 0 a
 * 2-D Gaussian (sigma in grid steps) over the shifts, normalised per head.
 * loss[i] = sum of the three head cross-entropies; dlogits [n][ld] bf16 =
 * (softmax - target) / n, zero in the columns after the heads. */
int isac_nn_softmax_xent(const float *logits, int n, int ld, int h0, int h1, int h2,
                         const int *labels, float sigma_psi, float sigma_shift, int shift_ny,
                         float *loss, uint16_t *dlogits, void *stream);

/* AdamW step t (1-based) over count parameters: p -= lr * (mhat / (sqrt(vhat) +
 * eps) + wd * p); mask (or NULL) zeroes the update where mask[i] == 0. */
int isac_nn_adamw(float *p, const float *g, float *m, float *v, const uint8_t *mask,
                  size_t count, float lr, float beta1, float beta2, float eps, float wd, int t,
                  void *stream);

/* Master [Cout][K+1] -> w [Cout][ks*ks][Cin] bf16 and (if wflip != NULL) the
 * input-gradient weights wflip [Cin][ks*ks][Cout] bf16, wflip[ci][t][co] =
 * w[co][ks*ks-1-t][ci]. */
int isac_nn_pack_weights(const float *master, int Cout, int Cin, int ks, uint16_t *w,
                         uint16_t *wflip, void *stream);

/* Top-k (k <= 8) proposals per particle from logits [n][ld]: heads of n_ref
 * references, 2*n_psi (mirror-major) psi classes and n_shift shifts.  The
 * candidates are the triples of the top-8 references, top-16 psi classes and
 * top-16 shifts, scored by the sum of their three logits and taken greedily,
 * best first (ties to the lower ranks).  With psi_r >= 0 each choice removes
 * the candidates of the same reference and mirror within psi_r bins and
 * shift_r grid steps (x-major grid of shift_ny per row), so that the narrowed
 * windows spread out.  The first min_refs choices take the best remaining
 * triple of the 1st, 2nd, ... ranked reference, so that many references are
 * proposed.  prop[i][j] = {ref, mirror, psi bin, shift index}; {-1, 0, 0, 0}
 * when fewer than k candidates remain.  conf (or NULL) [n]: the softmax mass of
 * the reference head on the distinct proposed references, the network's own
 * estimate that the true reference is among them. */
int isac_nn_proposals(const float *logits, int n, int ld, int n_ref, int n_psi, int n_shift,
                      int k, int psi_r, int shift_r, int shift_ny, int min_refs, int *prop,
                      float *conf, void *stream);

#ifdef __cplusplus
}
#endif
#endif
