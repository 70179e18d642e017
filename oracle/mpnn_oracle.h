/*
 * CPU oracle of the learned warm start's message-passing network (test
 * infrastructure only: tests/ use it as the checker of the device forward
 * and backward kernels; the product library never links or calls it).
 *
 * The network has no counterpart in the reference (TAPSim relaxes from the
 * previous potentials only); its definition is this build's own:
 *
 *   h0   = relu(X W_in + b_in)                       X: n x 32, W_in: 32 x 64
 *   for l = 1..3:
 *     m  = A h_{l-1}                                 A: symmetric edge weights
 *     h_l = relu([h_{l-1} | m] W_l + b_l) + h_{l-1}  W_l: 128 x 64
 *   y    = h_3 w_out + b_out                         w_out: 64
 *   loss = sum_i mask_i (y_i - t_i)^2 / max(1, sum_i mask_i)
 *
 * Parameters are one fp32 array laid out as TSG_MPNN_* in
 * include/tapsim_gpu.h.  With emulate = 1 every quantity the device rounds
 * to bfloat16 (features, weights, the GEMM operands and the back-propagated
 * pre-activation gradients) is rounded here at the same point; everything
 * else is carried in double.  With emulate = 0 the network is evaluated in
 * double throughout, which is what the finite-difference checks of the
 * gradient use.
 */
#ifndef MPNN_ORACLE_H
#define MPNN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_MPNN_IN = 32, ORA_MPNN_HID = 64, ORA_MPNN_LAYERS = 3 };
enum {
	ORA_MPNN_W_IN = 0,
	ORA_MPNN_B_IN = ORA_MPNN_IN * ORA_MPNN_HID,
	ORA_MPNN_LAYER0 = ORA_MPNN_B_IN + ORA_MPNN_HID,
	ORA_MPNN_LAYER_STRIDE = 2 * ORA_MPNN_HID * ORA_MPNN_HID + ORA_MPNN_HID,
	ORA_MPNN_W_OUT = ORA_MPNN_LAYER0 + ORA_MPNN_LAYERS * ORA_MPNN_LAYER_STRIDE,
	ORA_MPNN_B_OUT = ORA_MPNN_W_OUT + ORA_MPNN_HID,
	ORA_MPNN_NPARAM = ORA_MPNN_B_OUT + 1
};

typedef struct
{
	int n;              /* nodes */
	const int* ptr;     /* n + 1 row offsets */
	const int* col;     /* neighbour (local index) per edge */
	const float* a;     /* edge weight per edge; a_ij == a_ji */
	const float* x;     /* n * ORA_MPNN_IN features */
	const float* t;     /* n targets */
	const float* mask;  /* n loss weights */
} ora_mpnn_graph;

/* forward pass; y (n values) and loss may be NULL.  grad (ORA_MPNN_NPARAM
 * values) receives d loss / d params when not NULL. */
int ora_mpnn(const ora_mpnn_graph* g, const float* params, int emulate, double* loss, double* y, double* grad);

/* round-to-nearest-even to bfloat16 (returned as float) */
float ora_bf16(float v);

#ifdef __cplusplus
}
#endif

#endif
