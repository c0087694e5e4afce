/* mfa_mlp.h -- C ABI of the profile surrogate (libmfa_hip.so).
 *
 * A small multilayer perceptron, x -> [Linear -> SiLU]^(L-1) -> Linear -> y,
 * trained and evaluated on the GPU with hand-written BF16 MFMA kernels (FP32
 * accumulation, FP32 master weights, Adam).  It maps a channel's
 * (grafting density, chain length, brush-solvent epsilon, z) to the brush and
 * solvent densities at z; the warm start built from its output seeds initial
 * configurations in place of a plain lattice start.
 *
 * No counterpart in the reference: mfa_brush generates starts with
 * gen_brush.f90 / conf_default.f90 and has no learned component.
 *
 * Parameter layout (FP32, one flat array): W_0 (n_0 x n_1, row-major), b_0
 * (n_1), W_1, b_1, ...  Arrays of rows are row-major, rows x width.
 * Entry points return 0, or a negative code with mfa_mlp_last_error().
 */
#ifndef MFA_MLP_H
#define MFA_MLP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define MFA_MLP_E_ARG (-1)
#define MFA_MLP_E_HIP (-2)

typedef struct mfa_mlp mfa_mlp;

/* widths[0] = inputs, widths[n_widths-1] = outputs; He-uniform initial
 * weights from splitmix64(seed), zero biases */
int  mfa_mlp_create(const int32_t *widths, int32_t n_widths, uint64_t seed, mfa_mlp **out);
int  mfa_mlp_destroy(mfa_mlp *m);
int  mfa_mlp_num_params(mfa_mlp *m, int64_t *n);
int  mfa_mlp_get_params(mfa_mlp *m, float *params, int64_t n);
int  mfa_mlp_set_params(mfa_mlp *m, const float *params, int64_t n);   /* also resets the Adam moments */
/* y (rows x n_out) = f(x (rows x n_in)) */
int  mfa_mlp_forward(mfa_mlp *m, const float *x, int32_t rows, float *y);
/* mean squared error over rows x n_out and its gradient w.r.t. the flat parameters */
int  mfa_mlp_gradient(mfa_mlp *m, const float *x, const float *t, int32_t rows, float *loss, float *grad);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) on minibatches of `batch`
 * rows drawn without replacement per epoch (splitmix64(seed) shuffles);
 * loss_history[steps] (optional) receives each minibatch MSE.  One step is
 * captured as a hipGraph and replayed `steps` times (see mfa_mlp_set_graph) */
int  mfa_mlp_train(mfa_mlp *m, const float *x, const float *t, int64_t n_rows, int32_t steps, int32_t batch,
                   float lr, uint64_t seed, float *loss_history);
/* learning-rate schedule of later train calls: a cosine from lr at a call's
 * first step to lr * lr_final at its last; lr_final = 1 (the default) keeps
 * it constant.  No reference counterpart */
int  mfa_mlp_set_lr_schedule(mfa_mlp *m, float lr_final);
/* training by graph replay (on, the default) or by launching each step's
 * kernels (off); both give the same bits.  No reference counterpart */
int  mfa_mlp_set_graph(mfa_mlp *m, int32_t on);
/* C (M x N) = A (M x K) B (K x N) through the BF16 MFMA GEMM (operands
 * rounded to BF16, FP32 accumulation); exposes the tile for tests */
int  mfa_mlp_gemm_check(const float *a, const float *b, int32_t M, int32_t N, int32_t K, float *c);
const char *mfa_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
