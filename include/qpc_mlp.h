/*
 * qpc_mlp.h -- C ABI of the learned Krylov controller's network.
 *
 * No counterpart in the reference (its Krylov propagator, propagator_arnoldi
 * in propagator/la_1e.hpp:90-335, uses a fixed subspace size and step).  A
 * 7 -> 64 -> 64 -> 2 perceptron with GELU activations (exact erf form) and
 * FP32 parameters, trained and evaluated on the device by hand-written
 * kernels.  It predicts, per propagation step, the Krylov subspace size to
 * start from and a step-size multiplier; the propagator's a-posteriori error
 * estimate still decides (the subspace is extended whenever the estimate
 * exceeds the tolerance), so the plain method is always the fallback.
 *
 * Parameters, NPARAM floats, row-major (out x in):
 *   W1[64][7], b1[64], W2[64][64], b2[64], W3[2][64], b3[2].
 * Inputs X[n][7], targets / outputs [n][2]; the caller normalises them.
 * Loss (qpc_mlp_loss_grad, qpc_mlp_train): L = 1/(2n) sum_s sum_o (y - t)^2.
 * Every entry point returns 0 or a negative QPC_E_* status (qpc_gpu.h).
 */
#ifndef QPC_MLP_H
#define QPC_MLP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define QPC_MLP_IN      7
#define QPC_MLP_HID     64
#define QPC_MLP_OUT     2
#define QPC_MLP_NPARAM  4802

typedef struct qpc_mlp qpc_mlp;

/* Creates a network on the current HIP device with Glorot-uniform weights
 * and zero biases drawn from a splitmix64 stream seeded by `seed`. */
int  qpc_mlp_create(uint64_t seed, qpc_mlp** out);
int  qpc_mlp_destroy(qpc_mlp* net);
const char* qpc_mlp_last_error(const qpc_mlp* net);
int  qpc_mlp_get_params(qpc_mlp* net, float* params);
int  qpc_mlp_set_params(qpc_mlp* net, const float* params);
/* Y[n][2] = f(X[n][7]) (synchronous). */
int  qpc_mlp_forward(qpc_mlp* net, const float* X, int64_t n, float* Y);
/* Loss and its gradient with respect to every parameter on (X, T). */
int  qpc_mlp_loss_grad(qpc_mlp* net, const float* X, const float* T, int64_t n, float* loss, float* grad);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of `batch`
 * samples drawn without replacement each epoch (splitmix64 permutation from
 * `seed`); *loss_out receives the loss over all n samples afterwards. */
int  qpc_mlp_train(qpc_mlp* net, const float* X, const float* T, int64_t n, int64_t epochs, int64_t batch,
                   float lr, uint64_t seed, float* loss_out);

#ifdef __cplusplus
}
#endif
#endif
