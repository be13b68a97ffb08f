/*
 * psim_surrogate.h -- C ABI of the steady-state surrogate used for warm starts.
 *
 * No counterpart in the reference: psim starts every sensor at its configured initial
 * temperature (SensorController::t_steady_ -> getInitTemp,
 * psim/include/psim/sensorController.h:63-67).  The surrogate is an MLP
 * n_in -> hidden (x layers, ReLU) -> 1 whose weights and activations are bfloat16 and whose
 * products accumulate in FP32 on the MFMA units; Adam keeps FP32 master weights.  It proposes
 * initial sensor temperatures only; results always come from the Monte Carlo.
 *
 * Parameters are one flat FP32 vector: for each layer l (inputs dims[l], outputs dims[l + 1],
 * dims = n_in, hidden x layers, 16), the weights [dims[l]][dims[l + 1]] row-major, then the
 * biases [dims[l + 1]].  Only column 0 of the last layer is used; its other 15 columns stay 0.
 * Every function returns 0 or a negative PSIM_E_* code (psim_gpu.h).
 */
#ifndef PSIM_SURROGATE_H
#define PSIM_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psim_mlp psim_mlp;

int psim_mlp_create(psim_mlp** mlp, int device, int32_t n_in, int32_t hidden, int32_t layers);
void psim_mlp_destroy(psim_mlp* mlp);
const char* psim_mlp_last_error(const psim_mlp* mlp);
int64_t psim_mlp_num_params(const psim_mlp* mlp);

/* Sets the FP32 master parameters (and their bf16 copy); resets the optimiser state. */
int psim_mlp_set_params(psim_mlp* mlp, const float* params);
int psim_mlp_get_params(psim_mlp* mlp, float* params);

/* y[r] = network(x[r][0 .. n_in)) for r < rows.  Calls taking rows accept 1 .. 4194240 (65535
 * blocks of 64 rows); more is rejected with PSIM_E_INVALID and a message. */
int psim_mlp_forward(psim_mlp* mlp, const float* x, int32_t rows, float* y);

/* psim_mlp_forward without any device allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels and launch shapes as psim_mlp_forward, bit-identical results, with every buffer beyond
 * the caller's x and y carved from a 64 MiB device array that the library holds on each GPU it runs
 * on until the process ends.  Takes at most as many rows as that array holds for this network
 * (29056 for 32 -> 512 x 4); more is refused with PSIM_E_INVALID and a message.  Every call on one GPU
 * uses that GPU's one array: calls that overlap on one GPU (two threads, or two networks on
 * different streams) race. */
int psim_mlp_forward_noalloc(psim_mlp* mlp, const float* x, int32_t rows, float* y);

/* Mean squared error against t and its gradient with respect to every parameter (FP32). */
int psim_mlp_gradients(psim_mlp* mlp, const float* x, const float* t, int32_t rows, float* grad, float* loss);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the mean squared error of one batch.  The first
 * step at a batch size runs eagerly (and tunes the products' tiles); the next one captures the
 * step's kernels (forward, backward, Adam) as a HIP graph, which later steps of that size replay;
 * the batch's copy to the device and the loss read-back stay outside the graph.  Replayed and
 * eager steps give the same parameters bit for bit. */
int psim_mlp_train_step(psim_mlp* mlp, const float* x, const float* t, int32_t rows, float lr, float* loss);

/* Training steps replayed from a graph (1, the default) or all eager (0). */
int psim_mlp_set_graph(psim_mlp* mlp, int32_t enable);

/* C[M][N] = op(A)[M][K] x op(B)[K][N] on the MFMA GEMM used by the network, with the operands
 * rounded to bf16 (A stored [K][M] when ta, B stored [N][K] when tb): a check of the kernel. */
int psim_mlp_gemm_check(int device, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float* A,
    const float* B, float* C);

/* Average duration (ms) of one launch of that GEMM kernel on op(A)[M][K] x op(B)[K][N], from HIP
 * events around `reps` launches on a stream of its own: the measurement behind its roofline. */
int psim_mlp_time_gemm(int device, int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, int32_t reps,
    float* ms_per_launch);

/* Average duration (ms) of one of the network's own fused products of its first hidden x hidden
 * layer at `rows` rows, on the network's buffers and stream (HIP events around `reps` launches):
 *   kind 0  forward: bias, ReLU and bf16 rounding fused, bf16 activations out (inference);
 *   kind 1  forward in training: the same plus the transposed activation copy;
 *   kind 2  weight gradient: transposed activations x transposed gradient, split over the rows;
 *   kind 3  input gradient through the ReLU: bf16 gradient out, row-major and transposed.
 * *tile (optional) receives the tile the autotuner chose, as 1000 x rows + columns (128128 ...),
 * negated for the kernel that keeps its 128-column weight slice in LDS, with its band's rows:
 * -256128 (4 waves, 256-row bands) or -512128 (8 waves, 512-row bands).
 * Needs at least two hidden layers. */
int psim_mlp_time_product(psim_mlp* mlp, int32_t kind, int32_t rows, int32_t reps, float* ms_per_launch,
    int32_t* tile);

#ifdef __cplusplus
}
#endif

#endif /* PSIM_SURROGATE_H */
