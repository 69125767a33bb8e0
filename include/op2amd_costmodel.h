/*
 * A small MLP cost model for plan selection, trained and evaluated on the GPU
 * with this library's own HIP kernels (FP32).
 *
 * Architecture: nin -> H -> H -> 1, ReLU after both hidden layers, linear
 * output (predicted log kernel time).  Loss: mean squared error.  Optimiser:
 * Adam (beta1 0.9, beta2 0.999, eps 1e-8).  Parameters are stored flat in the
 * order W1[nin][H], b1[H], W2[H][H], b2[H], W3[H], b3.  All reductions over
 * samples are sequential per output element, so a gradient is the same from
 * run to run.  Optional decoupled weight decay (AdamW).
 */
#ifndef OP2AMD_COSTMODEL_H
#define OP2AMD_COSTMODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct op2amd_mlp op2amd_mlp;

/* He-initialised weights from `seed`, zero biases. */
op2amd_mlp *op2amd_mlp_create(int nin, int hidden, uint64_t seed);
void op2amd_mlp_destroy(op2amd_mlp *m);
int op2amd_mlp_num_params(const op2amd_mlp *m);
int op2amd_mlp_get_params(const op2amd_mlp *m, float *out);
int op2amd_mlp_set_params(op2amd_mlp *m, const float *in);

/* Forward pass on n samples x[n][nin] (host arrays); y[n] out. */
int op2amd_mlp_forward(op2amd_mlp *m, const float *x, int n, float *y);

/* Loss and its gradient with respect to every parameter (flat layout) at
 * the current parameters, on samples (x, t). */
int op2amd_mlp_loss_grad(op2amd_mlp *m, const float *x, const float *t, int n,
                         float *loss, float *grad);

/* Decoupled weight decay for the following training steps (AdamW:
 * p -= lr * wd * p after each Adam update; default 0).  Returns 0. */
int op2amd_mlp_set_weight_decay(op2amd_mlp *m, float wd);

/* `epochs` full-batch Adam steps with learning rate `lr`; final loss out.
 * Training continues from the current parameters and optimiser state. */
int op2amd_mlp_train(op2amd_mlp *m, const float *x, const float *t, int n,
                     int epochs, float lr, float *final_loss);

#ifdef __cplusplus
}
#endif

#endif
