/* tiger_mlp: the learned part of the chemistry step (C ABI).
 *
 * A multilayer perceptron nin -> hidden (GELU) x n_hidden_layers -> nout
 * trained and evaluated on the GPU with BF16 MFMA products accumulated in
 * FP32 (FP32 master weights, Adam).  It maps a scaled chemistry state and
 * time step to the scaled increment of the state over that step; the
 * surrogate-guided step (tgr_chem_source_chemistry_surrogate_device in
 * tiger_chem.h) uses it as the starting point of its FP64 Newton iteration.
 * No counterpart in the reference (which integrates every cell with CVODE,
 * combustion/finitechem.f90:987-1068).
 *
 * Parameter vector (FP32, host order of tgr_mlp_set_params/get_params): for
 * each linear layer l = 0..n_hidden_layers, W_l[out_pad_l][in_pad_l] row-major
 * then b_l[out_pad_l]; widths padded to multiples of 64 (in_pad_0 = nin
 * rounded up, out_pad_last = nout rounded up, hidden layers `hidden`), padded
 * entries zero.  Inputs x: BF16 bit patterns [B][in_pad] row-major, padded
 * columns zero; outputs and targets FP32, outputs [B][out_pad], targets
 * [B][nout].  B must be a positive multiple of 64.  All arrays are device
 * memory; every call is enqueued on `stream` (a hipStream_t, NULL = default).
 * Return 0 or a negative code; tgr_mlp_last_error() gives the message. */
#ifndef TIGER_MLP_H
#define TIGER_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TGR_MLP_E_ARG (-1)
#define TGR_MLP_E_HIP (-2)
#define TGR_MLP_E_NOMEM (-4)

typedef struct tgr_mlp tgr_mlp;

/* A network with all parameters zero, on the current HIP device. */
int tgr_mlp_create(int nin, int nout, int hidden, int n_hidden_layers, tgr_mlp **out);
int tgr_mlp_destroy(tgr_mlp *h);

/* Length of the parameter vector and padded input / output widths. */
int tgr_mlp_layout(const tgr_mlp *h, int64_t *nparam, int *in_pad, int *out_pad);

/* The shape given to tgr_mlp_create. */
int tgr_mlp_shape(const tgr_mlp *h, int *nin, int *nout, int *hidden, int *n_hidden_layers);

/* Copy the FP32 parameters from / to host memory (set also resets Adam). */
int tgr_mlp_set_params(tgr_mlp *h, const float *host);
int tgr_mlp_get_params(const tgr_mlp *h, float *host);

/* out[B][out_pad] = network(x[B][in_pad]). */
int tgr_mlp_forward(tgr_mlp *h, const uint16_t *x, int64_t B, float *out, void *stream);

/* Mean-squared-error loss against tgt[B][nout] and its gradient with respect
 * to every parameter (host vector of nparam floats, may be NULL); synchronous. */
int tgr_mlp_gradients(tgr_mlp *h, const uint16_t *x, const float *tgt, int64_t B, float *grad_host,
                      float *loss_host, void *stream);

/* One Adam step on the loss of the batch (bias-corrected, no weight decay).
 * loss_host != NULL: waits and returns the batch loss before the update. */
int tgr_mlp_train_step(tgr_mlp *h, const uint16_t *x, const float *tgt, int64_t B, float lr, float beta1,
                       float beta2, float eps, float *loss_host, void *stream);

/* The product kernel alone: C[M][N] (FP32, row stride ldc) = sum_k A(m,k) B(k,n)
 * with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (BF16); one of
 * A's strides and one of B's must be 1, the other a multiple of 8; M, N
 * multiples of 64, K of 32; 16-byte aligned operands. */
int tgr_mlp_gemm_bf16(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk, int64_t sbn,
                      int M, int N, int K, float *C, int64_t ldc, void *stream);

const char *tgr_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TIGER_MLP_H */
