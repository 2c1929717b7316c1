/* C ABI of the learned preconditioner's kernels (libefwi.so, gfx950).
 *
 * None of these has a counterpart in the reference (EFWI-3D has no learned
 * part); they serve the 3-D U-Net that maps the raw FWI gradient plus the
 * source illumination to a model update (efwi/unet.py).
 *
 * Tensors are device pointers to channels-last volumes of nz x nx x ny voxels
 * (voxel v = (z * nx + x) * ny + y): BF16 values (uint16_t bit patterns) at
 * p[v * ld + c], with the channel stride ld and the channel counts multiples
 * of 16 where stated, so a tensor may be a channel slice of a wider one.
 * Weights of a 3x3x3 convolution: w[t][cout][cin], tap t = 9 (dz+1) + 3 (dx+1)
 * + (dy+1). Every call is asynchronous on `stream` (a hipStream_t, NULL for
 * the null stream) and returns 0 or an error code below.
 */
#ifndef EFWI_NN_H
#define EFWI_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EFWI_NN_OK 0
#define EFWI_NN_ERR_ARG 1 /* invalid sizes, strides or pointers */
#define EFWI_NN_ERR_HIP 2 /* a kernel launch failed */
#define EFWI_NN_ERR_SCRATCH 3 /* efwi_nn_wgrad_noalloc: partial sums beyond EFWI_NN_WGRAD_STATIC_FLOATS */

/* 3x3x3 convolution with zero padding, BF16 in and out, FP32 accumulation on
 * the matrix cores (v_mfma_f32_16x16x32_bf16, the input staged through LDS).
 *   mode 0 (forward): dst[v][n] = relu?(sum_t,k src[v + d_t][k] w[t][n][k] + bias[n]),
 *                     kc = cin, nc = cout, bias may be NULL;
 *   mode 1 (input gradient): dst[v][n] = sum_t,k src[v - d_t][k] w[t][k][n],
 *                     src = output gradient (kc = cout), nc = cin.
 * Mode 1 gathers its weights with 2-byte loads; mode 0 with the weights of
 * efwi_nn_flip_weights computes the same input gradient at the forward rate.
 * When mask is not NULL, dst[v][n] is zeroed where mask[v * ldm + n] <= 0 (the
 * ReLU derivative of a layer's activation). kc, nc, lds, ldd, ldm: multiples
 * of 16, lds >= kc, ldd >= nc; src and w 16-byte aligned (a channel slice
 * starts at a multiple of 8 channels). */
int efwi_nn_conv3(const uint16_t* src, int lds, int kc, const uint16_t* w, const float* bias, uint16_t* dst,
                  int ldd, int nc, const uint16_t* mask, int ldm, int nz, int nx, int ny, int mode, int relu,
                  void* stream);

/* efwi_nn_conv3 on an input given as two channel ranges and an output split
 * in two: input channels [0, kc1) from src (stride lds), [kc1, kc1 + kc2)
 * from src2 (stride lds2); output channels [0, nc1) to dst (stride ldd),
 * [nc1, nc1 + nc2) to dst2 (stride ldd2). src2 or dst2 may be NULL (one
 * range). With up1, src is a (nz/2, nx/2, ny/2) volume (even sizes) whose
 * nearest upsampling is read: voxel (z, x, y) takes src's (z/2, x/2, y/2).
 * Weights, bias and mask (stride ldm >= nc1 + nc2) address the joined
 * channels. Results equal efwi_nn_conv3 on the joined (and upsampled) tensors
 * bit for bit. For a skip concatenation kept as two separate tensors. */
int efwi_nn_conv3_cat(const uint16_t* src, int lds, int kc1, int up1, const uint16_t* src2, int lds2, int kc2,
                      const uint16_t* w, const float* bias, uint16_t* dst, int ldd, int nc1, uint16_t* dst2, int ldd2,
                      int nc2, const uint16_t* mask, int ldm, int nz, int nx, int ny, int mode, int relu,
                      void* stream);

/* wt[t][k][n] = w[26 - t][n][k] for a layer's weights w[27][nc][kc]: the
 * input gradient of that layer is the mode-0 convolution of the output
 * gradient (kc' = nc input channels, nc' = kc outputs) with wt. wt must not
 * alias w. */
int efwi_nn_flip_weights(const uint16_t* w, int kc, int nc, uint16_t* wt, void* stream);

/* efwi_nn_flip_weights for n layers (1 <= n <= 16) in one launch: host arrays
 * of n weight pointers, channel counts and output pointers. */
int efwi_nn_flip_weights_n(int n, const uint16_t* const* w, const int* kc, const int* nc, uint16_t* const* wt,
                           void* stream);

/* Weight gradient of a 3x3x3 convolution, dw[t][m][n] = sum_v dpre[v][m]
 * src[v + d_t][n] (FP32 result), and, when db is not NULL, its bias gradient
 * db[m] = sum_v dpre[v][m]; dpre = gradient of the layer's pre-activation (mc
 * channels), src = its input (kc channels); dpre and src 16-byte aligned (a
 * channel slice starts at a multiple of 8 channels). The volume is cut into
 * bricks of 2 x 2 rows of 64 voxels along y (32 where ny <= 32); each
 * workgroup reduces `per` bricks for one 16 x 16 block of (m, n) on
 * v_mfma_f32_16x16x32_bf16 into FP32 partials in `scratch`
 * (efwi_nn_wgrad_scratch floats), which are then summed. */
long efwi_nn_wgrad_scratch(int mc, int kc, int nz, int nx, int ny, int per);
int efwi_nn_wgrad(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* scratch,
                  long scratch_floats, float* dw, float* db, int nz, int nx, int ny, int per, void* stream);

/* efwi_nn_wgrad with src given as two channel ranges: [0, kc1) from src
 * (stride lds; with up1 a half-size volume read through nearest upsampling,
 * as in efwi_nn_conv3_cat), [kc1, kc1 + kc2) from src2 (stride lds2);
 * dw[t][m][n] over the joined kc1 + kc2 channels, equal to efwi_nn_wgrad on
 * the joined tensor. scratch as efwi_nn_wgrad_scratch(mc, kc1 + kc2, ...). */
int efwi_nn_wgrad_cat(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc1, int up1,
                      const uint16_t* src2, int lds2, int kc2, float* scratch, long scratch_floats, float* dw,
                      float* db, int nz, int nx, int ny, int per, void* stream);

/* efwi_nn_wgrad without a workspace argument and without any allocation:
 * the partial sums live in a static device array of
 * EFWI_NN_WGRAD_STATIC_FLOATS floats (16 MiB) that each GPU holds for the
 * process's lifetime. Calls whose efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny,
 * per) exceeds it return EFWI_NN_ERR_SCRATCH (raise `per`). Every call on a
 * GPU uses that GPU's one array: calls that overlap on one GPU (concurrent
 * streams or threads) race. Same kernels as efwi_nn_wgrad: results equal its
 * results up to the order of the FP32 atomic additions that combine groups
 * of 32 chunks (which varies between calls once there are three or more). No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under graph capture, training loops that forbid it). */
#define EFWI_NN_WGRAD_STATIC_FLOATS (4L << 20)
int efwi_nn_wgrad_noalloc(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* dw,
                          float* db, int nz, int nx, int ny, int per, void* stream);

/* out[c] = sum_v x[v][c] (a bias gradient on its own), c divides 256, ld >= c. */
int efwi_nn_chan_sum(const uint16_t* x, int ld, int c, long nvox, float* out, void* stream);

/* 2x2x2 average pooling of a (nz, nx, ny) volume (even sizes) and its
 * gradient (accumulate: add into df instead of overwriting; zeroed where
 * mask <= 0 when mask is given, after the BF16 rounding). */
int efwi_nn_pool2(const uint16_t* src, int lds, uint16_t* dst, int ldd, int c, int nz, int nx, int ny, void* stream);
int efwi_nn_pool2_bwd(const uint16_t* dc, int ldc, uint16_t* df, int ldf, int c, int nz, int nx, int ny, int accumulate,
                      const uint16_t* mask, int ldm, void* stream);

/* Nearest-neighbour upsampling x2 to a (nz, nx, ny) volume and its gradient
 * (sum over the 8 children, zeroed where mask <= 0 when mask is given). */
int efwi_nn_up2(const uint16_t* cs, int ldc, uint16_t* fd, int ldf, int c, int nz, int nx, int ny, void* stream);
int efwi_nn_up2_bwd(const uint16_t* df, int ldf, uint16_t* dc, int ldc, int c, int nz, int nx, int ny,
                    const uint16_t* mask, int ldm, void* stream);

/* a[v][c] = (a + b)[v][c] (b may be NULL: a alone), zeroed where mask <= 0
 * when mask is given. */
int efwi_nn_add(uint16_t* a, int lda, const uint16_t* b, int ldb, int c, long nvox, const uint16_t* mask, int ldm,
                void* stream);

/* 1x1x1 output head in FP32: out[v] = sum_c w[c] act[v][c] + b[0] (c <= 16),
 * and its gradient: dact = w dout where act > 0, dw[c] = sum_v act dout,
 * db = sum_v dout. */
int efwi_nn_head(const uint16_t* act, int ld, int c, const float* w, const float* b, float* out, long nvox,
                 void* stream);
int efwi_nn_head_bwd(const uint16_t* act, int ld, int c, const float* w, const float* dout, uint16_t* dact, int ldd,
                     float* dw, float* db, long nvox, void* stream);

/* Multiplicative output: u[v] = g[v] exp(clamp(o[v], -4, 6)) with g =
 * channel 0 of the network input feat (BF16, stride ld) and o the head's FP32
 * output; its gradient dout[v] = du[v] u[v] where -4 < o[v] < 6, else 0. */
int efwi_nn_gate(const uint16_t* feat, int ld, const float* o, float* u, long nvox, void* stream);
int efwi_nn_gate_bwd(const float* o, const float* u, const float* du, float* dout, long nvox, void* stream);

/* Masked mean-square loss over `count` voxels: *loss = sum m (out - t)^2 /
 * count (FP64, device), dout = 2 m (out - t) / count. */
/* efwi_nn_mse, efwi_nn_gate_bwd and efwi_nn_head_bwd (16 channels) in one
 * pass: with u (the gated output) the loss of u and its gradient through the
 * gate, with u NULL the loss of the head output o itself; dact, dw, db and
 * *loss as those calls give them (dact bit for bit; the sums in another
 * order). act and dact 16-byte aligned, ld and ldd multiples of 8. */
int efwi_nn_loss_head_bwd(const uint16_t* act, int ld, const float* w, const float* o, const float* u,
                          const float* target, const float* mask, long nvox, double count, uint16_t* dact, int ldd,
                          float* dw, float* db, double* loss, void* stream);
int efwi_nn_mse(const float* out, const float* target, const float* mask, long nvox, double count, float* dout,
                double* loss, void* stream);

/* Adam step on FP32 master weights (bias-corrected, step >= 1); w16, when
 * not NULL, receives the BF16 copy. */
int efwi_nn_adam(float* w, const float* g, float* m1, float* m2, long n, float lr, float b1, float b2, float eps,
                 int step, uint16_t* w16, void* stream);

/* y = bf16(x), round to nearest even. */
int efwi_nn_to_bf16(const float* x, uint16_t* y, long n, void* stream);

/* *out = max |x| (FP32, device). */
int efwi_nn_absmax(const float* x, long n, float* out, void* stream);

/* Network input from padded FWI-grid volumes (npml cells per face), written
 * for the interior only (16 channels): channel 0 = g * inv_g, channel 1 =
 * log10(illum * inv_i + 1e-6) / 6 + 1, the other channels 0. */
int efwi_nn_pack(const float* g, const float* illum, int nz, int nx, int ny, int npml, float inv_g, float inv_i,
                 uint16_t* feat, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EFWI_NN_H */
