// Warm-start network: sizes, the flat parameter layout and the launchers of
// its device kernels.
//
//   h0   = relu(X W_in + b_in)                       X: n x 32 (bf16), W_in: 32 x 64
//   for l = 1..3:
//     m   = A h_{l-1}                                A: symmetric edge weights
//     h_l = relu([h_{l-1} | m] W_l + b_l) + h_{l-1}  W_l: 128 x 64
//   y    = h_3 w_out + b_out
//
// Parameters are one fp32 array (row-major W[k][o]); the GEMMs multiply
// bfloat16 copies of the weights and of their operands with fp32
// accumulation on the MFMA units.
#pragma once

#include <hip/hip_runtime.h>

namespace tsg {

constexpr int kMpnnIn = 32, kMpnnHid = 64, kMpnnLayers = 3, kMpnnCat = 2 * kMpnnHid;
constexpr int kMpnnWIn = 0;
constexpr int kMpnnBIn = kMpnnIn * kMpnnHid;
constexpr int kMpnnLayer0 = kMpnnBIn + kMpnnHid;
constexpr int kMpnnLayerStride = kMpnnCat * kMpnnHid + kMpnnHid;
constexpr int kMpnnWOut = kMpnnLayer0 + kMpnnLayers * kMpnnLayerStride;
constexpr int kMpnnBOut = kMpnnWOut + kMpnnHid;
constexpr int kMpnnNParam = kMpnnBOut + 1;

typedef unsigned short bf16_t;  // bfloat16 bits

// bfloat16 copies of the weights in the layouts the kernels read
struct MpnnWeights {
	bf16_t* win_t;   // [64][32]   W_in transposed (forward B operand)
	bf16_t* wl_t;    // [3][64][128] W_l transposed (forward)
	bf16_t* wl;      // [3][128][64] W_l (input-gradient GEMM)
	bf16_t* wout;    // [64]
};

// C[n x N] = A[n x K] * Bt[N x K]^T (bf16 operands, fp32 accumulation),
// N a multiple of 64, K a multiple of 32.  With `bias` set the epilogue of a
// network layer is applied: z = C + bias -> Z, h = relu(z) (+ res) -> H, and
// bf16(h) -> Hb[row * hb_stride + col] when Hb is set; otherwise C is stored.
int launch_mpnn_gemm(hipStream_t, int n, int K, int N, const bf16_t* A, const bf16_t* Bt, float* C, const float* bias,
                     const float* res, float* Z, float* H, bf16_t* Hb, int hb_stride);
// m = A h over rows [0, n) of a batch whose global rows start at row0:
// cat[i][64 + o] = bf16(sum_e a_e h[col_e - row0][o])
int launch_mpnn_aggregate(hipStream_t, int n, int row0, const int* ptr, const int* col, const float* a, const float* h,
                          bf16_t* cat);
// dh_prev[i][o] = dh[i][o] + dcat[i][o] + sum_e a_e dcat[col_e - row0][64 + o]
int launch_mpnn_aggregate_bwd(hipStream_t, int n, int row0, const int* ptr, const int* col, const float* a,
                              const float* dh, const float* dcat, float* dh_prev);
// y[i] = sum_o bf16(h3[i][o]) * wout[o] + b_out
int launch_mpnn_out(hipStream_t, int n, const float* h3, const bf16_t* wout, const float* b_out, float* y);
// dy[i] = 2 mask_i (y_i - t_i) / msum ; loss partials (sum of mask (y-t)^2 / msum) per block
int launch_mpnn_loss(hipStream_t, int n, const float* y, const float* t, const float* mask, float msum, float* dy,
                     float* loss_part, int* nblocks);
// dh3[i][o] = dy[i] * wout[o]
int launch_mpnn_out_bwd(hipStream_t, int n, const float* dy, const bf16_t* wout, float* dh3);
// dz = dh * (z > 0) -> dZ (fp32) and dZb (bf16)
int launch_mpnn_relu_bwd(hipStream_t, int n, const float* dh, const float* z, float* dz, bf16_t* dzb);
// partial column sums over blocks of 256 rows: part[b][c] = sum_rows X[r][c] * (scale ? scale[r] : 1),
// with X rounded to bf16 first when round_bf16 is set; 64 columns
int launch_mpnn_colsum(hipStream_t, int n, const float* X, const float* scale, int round_bf16, float* part, int* nblocks);
// partial W gradients over blocks of 256 rows: part[b][k][o] = sum_rows A[r][k] * D[r][o]  (A: n x K bf16, D: n x 64 bf16)
int launch_mpnn_wgrad(hipStream_t, int n, int K, const bf16_t* A, const bf16_t* D, float* part, int* nblocks);
// out[c] (+)= sum_b part[b][c] in block order, c < width
int launch_mpnn_reduce(hipStream_t, int nblocks, int width, const float* part, float* out, int accumulate);
// Adam on the fp32 parameters, then the bf16 weight copies
int launch_mpnn_adam(hipStream_t, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                     int step);
int launch_mpnn_pack(hipStream_t, const float* p, const MpnnWeights& w);
// fp32 -> bf16 copy
int launch_to_bf16(hipStream_t, long long n, const float* x, bf16_t* out);

}  // namespace tsg
