// Warm-start network kernels: bf16 MFMA GEMMs (v_mfma_f32_16x16x32_bf16,
// fp32 accumulation) for the dense layers and their gradients, message
// aggregation over the shell graph, the output layer, the loss, and Adam.
//
// MFMA operand layout (16x16x32, one wavefront): lane l holds
//   A[row = l % 16][k = 8 * (l / 16) + i]   i = 0..7   (16 x 32 tile)
//   B[k = 8 * (l / 16) + i][col = l % 16]               (32 x 16 tile)
//   C[row = 4 * (l / 16) + i][col = l % 16] i = 0..3   (16 x 16 result)
// The GEMMs read B as Bt (N x K row-major), so both operands of a lane are
// eight consecutive bf16 values (one 16-byte load).
#include "kernels.hpp"
#include "mpnn.hpp"

namespace tsg {
namespace dev {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

// round to nearest even; NaN stays NaN
__device__ __forceinline__ bf16_t f2bf(float f)
{
	unsigned u = __float_as_uint(f);
	if ((u & 0x7f800000u) == 0x7f800000u) {
		if (u & 0x007fffffu) u |= 0x00400000u;
		return (bf16_t) (u >> 16);
	}
	u += 0x7fffu + ((u >> 16) & 1u);
	return (bf16_t) (u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(((unsigned) b) << 16); }

__device__ __forceinline__ bf16x8_t load8(const bf16_t* p)
{
	const u32x4_t v = *(const u32x4_t*) p;
	return __builtin_bit_cast(bf16x8_t, v);
}

// one workgroup = 4 wavefronts x 16 rows; blockIdx.y selects 64 output columns
__global__ __launch_bounds__(256) void k_gemm(int n, int K, int N, const bf16_t* __restrict__ A,
                                              const bf16_t* __restrict__ Bt, float* __restrict__ C,
                                              const float* __restrict__ bias, const float* __restrict__ res,
                                              float* __restrict__ Z, float* __restrict__ H, bf16_t* __restrict__ Hb,
                                              int hb_stride)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int r0 = blockIdx.x * 64 + wave * 16;
	const int c0 = blockIdx.y * 64;
	if (r0 >= n) return;  // wave-uniform
	const int ar = min(r0 + (lane & 15), n - 1);
	const int kq = 8 * (lane >> 4);
	f32x4_t acc[4];
	for (int t = 0; t < 4; t++) acc[t] = (f32x4_t){0.0f, 0.0f, 0.0f, 0.0f};
	for (int k0 = 0; k0 < K; k0 += 32) {
		const bf16x8_t a = load8(A + (size_t) ar * K + k0 + kq);
		bf16x8_t b[4];
#pragma unroll
		for (int t = 0; t < 4; t++) b[t] = load8(Bt + (size_t) (c0 + t * 16 + (lane & 15)) * K + k0 + kq);
#pragma unroll
		for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], acc[t], 0, 0, 0);
	}
#pragma unroll
	for (int t = 0; t < 4; t++) {
		const int col = c0 + t * 16 + (lane & 15);
#pragma unroll
		for (int i = 0; i < 4; i++) {
			const int row = r0 + 4 * (lane >> 4) + i;
			if (row >= n) continue;
			const float v = acc[t][i];
			const size_t o = (size_t) row * N + col;
			if (bias) {
				const float z = v + bias[col];
				Z[o] = z;
				float h = z > 0.0f ? z : 0.0f;
				if (res) h += res[o];
				H[o] = h;
				if (Hb) Hb[(size_t) row * hb_stride + col] = f2bf(h);
			} else
				C[o] = v;
		}
	}
}

// one wavefront per row, one lane per channel; edges in CSR order
// sum over the row's edges e (in CSR order) of a[e] x[col[e] - row0][o]:
// the wavefront loads 64 edges' columns and weights at once (lane k holds
// edge c + k) and gathers the neighbours' rows eight at a time, so the
// gathers do not wait on one column load each; the additions stay in edge
// order
__device__ __forceinline__ float edge_sum(int e0, int e1, int row0, const int* __restrict__ col,
                                          const float* __restrict__ a, const float* x, int stride, int o, float m)
{
	for (int c = e0; c < e1; c += 64) {
		const int cnt = min(64, e1 - c);
		int cj = 0;
		float aj = 0.0f;
		if (o < cnt) {
			cj = col[c + o] - row0;
			aj = a[c + o];
		}
		for (int k0 = 0; k0 < cnt; k0 += 8) {
			float v[8];
#pragma unroll
			for (int k = 0; k < 8; k++) {
				const int j = __shfl(cj, k0 + k, 64);
				v[k] = k0 + k < cnt ? x[(size_t) j * stride + o] : 0.0f;
			}
#pragma unroll
			for (int k = 0; k < 8; k++)
				if (k0 + k < cnt) m += __shfl(aj, k0 + k, 64) * v[k];
		}
	}
	return m;
}

__global__ __launch_bounds__(256) void k_aggregate(int n, int row0, const int* __restrict__ ptr, const int* __restrict__ col,
                                                   const float* __restrict__ a, const float* __restrict__ h,
                                                   bf16_t* __restrict__ cat)
{
	const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int o = threadIdx.x & 63;
	if (i >= n) return;
	const int e0 = ptr[row0 + i], e1 = ptr[row0 + i + 1];
	const float m = edge_sum(e0, e1, row0, col, a, h, kMpnnHid, o, 0.0f);
	cat[(size_t) i * kMpnnCat + kMpnnHid + o] = f2bf(m);
}

__global__ __launch_bounds__(256) void k_aggregate_bwd(int n, int row0, const int* __restrict__ ptr,
                                                       const int* __restrict__ col, const float* __restrict__ a,
                                                       const float* dh, const float* __restrict__ dcat, float* dh_prev)
{
	const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int o = threadIdx.x & 63;
	if (i >= n) return;
	const int e0 = ptr[row0 + i], e1 = ptr[row0 + i + 1];
	const float d0 = dh[(size_t) i * kMpnnHid + o] + dcat[(size_t) i * kMpnnCat + o];
	dh_prev[(size_t) i * kMpnnHid + o] = edge_sum(e0, e1, row0, col, a, dcat + kMpnnHid, kMpnnCat, o, d0);
}

__global__ __launch_bounds__(256) void k_out(int n, const float* __restrict__ h3, const bf16_t* __restrict__ wout,
                                             const float* __restrict__ b_out, float* __restrict__ y)
{
	const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int o = threadIdx.x & 63;
	if (i >= n) return;  // wave-uniform
	float v = bf2f(f2bf(h3[(size_t) i * kMpnnHid + o])) * bf2f(wout[o]);
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	if (o == 0) y[i] = v + b_out[0];
}

// 256 rows per block; loss_part[2b] = sum mask (y-t)^2 / msum, loss_part[2b+1] = sum dy
__global__ __launch_bounds__(256) void k_loss(int n, const float* __restrict__ y, const float* __restrict__ t,
                                              const float* __restrict__ mask, float msum, float* __restrict__ dy,
                                              float* __restrict__ loss_part)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	float l = 0.0f, g = 0.0f;
	if (i < n) {
		const float d = y[i] - t[i];
		g = 2.0f * mask[i] * d / msum;
		l = mask[i] * d * d / msum;
		dy[i] = g;
	}
	__shared__ float sl[256], sg[256];
	sl[threadIdx.x] = l;
	sg[threadIdx.x] = g;
	__syncthreads();
	for (int h = 128; h > 0; h >>= 1) {
		if ((int) threadIdx.x < h) {
			sl[threadIdx.x] += sl[threadIdx.x + h];
			sg[threadIdx.x] += sg[threadIdx.x + h];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		loss_part[2 * blockIdx.x] = sl[0];
		loss_part[2 * blockIdx.x + 1] = sg[0];
	}
}

__global__ void k_out_bwd(int n, const float* __restrict__ dy, const bf16_t* __restrict__ wout, float* __restrict__ dh3)
{
	const long long k = (long long) blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= (long long) n * kMpnnHid) return;
	dh3[k] = dy[k / kMpnnHid] * bf2f(wout[k % kMpnnHid]);
}

__global__ void k_relu_bwd(int n, const float* __restrict__ dh, const float* __restrict__ z, float* __restrict__ dz,
                           bf16_t* __restrict__ dzb)
{
	const long long k = (long long) blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= (long long) n * kMpnnHid) return;
	const float d = z[k] > 0.0f ? dh[k] : 0.0f;
	dz[k] = d;
	dzb[k] = f2bf(d);
}

__global__ __launch_bounds__(256) void k_colsum(int n, const float* __restrict__ X, const float* __restrict__ scale,
                                                int rnd, float* __restrict__ part)
{
	const int c = threadIdx.x & 63, sub = threadIdx.x >> 6;
	const int rb = blockIdx.x * 256;
	float s = 0.0f;
	for (int k = 0; k < 64; k++) {
		const int r = rb + sub + 4 * k;
		if (r >= n) break;
		float x = X[(size_t) r * kMpnnHid + c];
		if (rnd) x = bf2f(f2bf(x));
		if (scale) x *= scale[r];
		s += x;
	}
	__shared__ float sp[4][64];
	sp[sub][c] = s;
	__syncthreads();
	if (sub == 0) part[(size_t) blockIdx.x * 64 + c] = ((sp[0][c] + sp[1][c]) + sp[2][c]) + sp[3][c];
}

// partial dW = A^T D over 256 rows per block, staged through LDS 64 rows at a time
__global__ __launch_bounds__(256) void k_wgrad(int n, int K, const bf16_t* __restrict__ A, const bf16_t* __restrict__ D,
                                               float* __restrict__ part)
{
	__shared__ __attribute__((aligned(16))) bf16_t As[64 * kMpnnCat];
	__shared__ __attribute__((aligned(16))) bf16_t Ds[64 * kMpnnHid];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int ntiles = (K / 16) * 4;
	f32x4_t acc[8];
	for (int j = 0; j < 8; j++) acc[j] = (f32x4_t){0.0f, 0.0f, 0.0f, 0.0f};
	const int kc = K / 8;  // 16-byte chunks per row of A
	for (int sub = 0; sub < 4; sub++) {
		const int rb = blockIdx.x * 256 + sub * 64;
		for (int idx = threadIdx.x; idx < 64 * kc; idx += 256) {
			const int r = idx / kc, cc = idx % kc;
			u32x4_t v = {0u, 0u, 0u, 0u};
			if (rb + r < n) v = *(const u32x4_t*) (A + (size_t) (rb + r) * K + cc * 8);
			*(u32x4_t*) (As + r * K + cc * 8) = v;
		}
		for (int idx = threadIdx.x; idx < 64 * 8; idx += 256) {
			const int r = idx / 8, cc = idx % 8;
			u32x4_t v = {0u, 0u, 0u, 0u};
			if (rb + r < n) v = *(const u32x4_t*) (D + (size_t) (rb + r) * kMpnnHid + cc * 8);
			*(u32x4_t*) (Ds + r * kMpnnHid + cc * 8) = v;
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < 8; j++) {
			const int tile = wave + 4 * j;
			if (tile < ntiles) {  // wave-uniform
				const int m0 = (tile >> 2) * 16, c0 = (tile & 3) * 16;
#pragma unroll
				for (int rr = 0; rr < 64; rr += 32) {
					u16x8_t ta, tb;
#pragma unroll
					for (int i = 0; i < 8; i++) {
						const int r = rr + 8 * (lane >> 4) + i;
						ta[i] = As[r * K + m0 + (lane & 15)];
						tb[i] = Ds[r * kMpnnHid + c0 + (lane & 15)];
					}
					acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, ta),
					                                                 __builtin_bit_cast(bf16x8_t, tb), acc[j], 0, 0, 0);
				}
			}
		}
		__syncthreads();
	}
	float* out = part + (size_t) blockIdx.x * K * kMpnnHid;
#pragma unroll
	for (int j = 0; j < 8; j++) {
		const int tile = wave + 4 * j;
		if (tile >= ntiles) continue;
		const int m0 = (tile >> 2) * 16, c0 = (tile & 3) * 16;
#pragma unroll
		for (int i = 0; i < 4; i++) out[(m0 + 4 * (lane >> 4) + i) * kMpnnHid + c0 + (lane & 15)] = acc[j][i];
	}
}

__global__ void k_reduce(int nblocks, int width, const float* __restrict__ part, float* __restrict__ out, int accumulate)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= width) return;
	float s = 0.0f;
	for (int b = 0; b < nblocks; b++) s += part[(size_t) b * width + c];
	out[c] = accumulate ? out[c] + s : s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= kMpnnNParam) return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.0f - b1) * gi;
	const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_pack(const float* __restrict__ p, MpnnWeights w)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < kMpnnBIn) {  // W_in[k][o]
		const int k = i / kMpnnHid, o = i % kMpnnHid;
		w.win_t[o * kMpnnIn + k] = f2bf(p[i]);
		return;
	}
	if (i >= kMpnnWOut && i < kMpnnBOut) {
		w.wout[i - kMpnnWOut] = f2bf(p[i]);
		return;
	}
	if (i < kMpnnLayer0 || i >= kMpnnWOut) return;
	const int l = (i - kMpnnLayer0) / kMpnnLayerStride;
	const int r = (i - kMpnnLayer0) % kMpnnLayerStride;
	if (r >= kMpnnCat * kMpnnHid) return;  // bias
	const int k = r / kMpnnHid, o = r % kMpnnHid;
	const bf16_t b = f2bf(p[i]);
	w.wl_t[(size_t) l * kMpnnCat * kMpnnHid + o * kMpnnCat + k] = b;
	w.wl[(size_t) l * kMpnnCat * kMpnnHid + k * kMpnnHid + o] = b;
}

__global__ void k_to_bf16(long long n, const float* __restrict__ x, bf16_t* __restrict__ out)
{
	const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = f2bf(x[i]);
}

}  // namespace dev

static inline int blocks_for(long long n, int per) { return (int) ((n + per - 1) / per); }

int launch_mpnn_gemm(hipStream_t s, int n, int K, int N, const bf16_t* A, const bf16_t* Bt, float* C, const float* bias,
                     const float* res, float* Z, float* H, bf16_t* Hb, int hb_stride)
{
	if (n <= 0) return 0;
	if (K % 32 || N % 64) return -1;
	hipLaunchKernelGGL(dev::k_gemm, dim3(blocks_for(n, 64), N / 64), dim3(256), 0, s, n, K, N, A, Bt, C, bias, res, Z, H,
	                   Hb, hb_stride);
	return (int) hipGetLastError();
}

int launch_mpnn_aggregate(hipStream_t s, int n, int row0, const int* ptr, const int* col, const float* a, const float* h,
                          bf16_t* cat)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_aggregate, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, row0, ptr, col, a, h, cat);
	return (int) hipGetLastError();
}

int launch_mpnn_aggregate_bwd(hipStream_t s, int n, int row0, const int* ptr, const int* col, const float* a,
                              const float* dh, const float* dcat, float* dh_prev)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_aggregate_bwd, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, row0, ptr, col, a, dh, dcat,
	                   dh_prev);
	return (int) hipGetLastError();
}

int launch_mpnn_out(hipStream_t s, int n, const float* h3, const bf16_t* wout, const float* b_out, float* y)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_out, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, h3, wout, b_out, y);
	return (int) hipGetLastError();
}

int launch_mpnn_loss(hipStream_t s, int n, const float* y, const float* t, const float* mask, float msum, float* dy,
                     float* loss_part, int* nblocks)
{
	*nblocks = blocks_for(n, 256);
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_loss, dim3(*nblocks), dim3(256), 0, s, n, y, t, mask, msum, dy, loss_part);
	return (int) hipGetLastError();
}

int launch_mpnn_out_bwd(hipStream_t s, int n, const float* dy, const bf16_t* wout, float* dh3)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_out_bwd, dim3(blocks_for((long long) n * kMpnnHid, 256)), dim3(256), 0, s, n, dy, wout, dh3);
	return (int) hipGetLastError();
}

int launch_mpnn_relu_bwd(hipStream_t s, int n, const float* dh, const float* z, float* dz, bf16_t* dzb)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_relu_bwd, dim3(blocks_for((long long) n * kMpnnHid, 256)), dim3(256), 0, s, n, dh, z, dz, dzb);
	return (int) hipGetLastError();
}

int launch_mpnn_colsum(hipStream_t s, int n, const float* X, const float* scale, int round_bf16, float* part, int* nblocks)
{
	*nblocks = blocks_for(n, 256);
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_colsum, dim3(*nblocks), dim3(256), 0, s, n, X, scale, round_bf16, part);
	return (int) hipGetLastError();
}

int launch_mpnn_wgrad(hipStream_t s, int n, int K, const bf16_t* A, const bf16_t* D, float* part, int* nblocks)
{
	*nblocks = blocks_for(n, 256);
	if (n <= 0) return 0;
	if (K % 32 || K > kMpnnCat) return -1;
	hipLaunchKernelGGL(dev::k_wgrad, dim3(*nblocks), dim3(256), 0, s, n, K, A, D, part);
	return (int) hipGetLastError();
}

int launch_mpnn_reduce(hipStream_t s, int nblocks, int width, const float* part, float* out, int accumulate)
{
	if (width <= 0) return 0;
	hipLaunchKernelGGL(dev::k_reduce, dim3(blocks_for(width, 256)), dim3(256), 0, s, nblocks, width, part, out, accumulate);
	return (int) hipGetLastError();
}

int launch_mpnn_adam(hipStream_t s, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                     int step)
{
	const float c1 = 1.0f - powf(b1, (float) step), c2 = 1.0f - powf(b2, (float) step);
	hipLaunchKernelGGL(dev::k_adam, dim3(blocks_for(kMpnnNParam, 256)), dim3(256), 0, s, p, g, m, v, lr, b1, b2, eps, c1, c2);
	return (int) hipGetLastError();
}

int launch_mpnn_pack(hipStream_t s, const float* p, const MpnnWeights& w)
{
	hipLaunchKernelGGL(dev::k_pack, dim3(blocks_for(kMpnnNParam, 256)), dim3(256), 0, s, p, w);
	return (int) hipGetLastError();
}

int launch_to_bf16(hipStream_t s, long long n, const float* x, bf16_t* out)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_to_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, out);
	return (int) hipGetLastError();
}

}  // namespace tsg
