// Warm start of the relaxation after an evaporation event: the network's
// input features over the shell of the evaporated atom, the application of
// its predicted potential change (and its undoing), and the training
// targets recorded once the relaxation has converged.
#include "kernels.hpp"
#include "mpnn.hpp"

namespace tsg {
namespace dev {

__device__ __forceinline__ bf16_t to_bf16(float f)
{
	unsigned u = __float_as_uint(f);
	if ((u & 0x7f800000u) == 0x7f800000u) {
		if (u & 0x007fffffu) u |= 0x00400000u;
		return (bf16_t) (u >> 16);
	}
	u += 0x7fffu + ((u >> 16) & 1u);
	return (bf16_t) (u >> 16);
}

// the scaled FP64 residual J(phi)_i - phi_i of every shell row, one
// wavefront per row (lane l takes entries l, l+64, ...; a fixed butterfly
// sums the lanes), and phi_pre; Dirichlet shell nodes get 0
__global__ __launch_bounds__(256) void k_warm_resid(WarmArgs w, SweepArgs a)
{
	const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= w.n) return;
	const int node = w.nodes[i];
	const int slot = w.slots[i];
	double acc = 0.0;
	if (slot >= 0) {
		const int deg = a.slot_deg[slot];
		const long long base = a.slice_off[slot >> 6] + (slot & 63);
		const double* coup = (const double*) a.coup;
		for (int k = lane; k < deg; k += 64) acc += w.phi[a.col[base + 64ll * k]] * coup[base + 64ll * k];
	}
	for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
	if (lane != 0) return;
	const double p = w.phi[node];
	w.phi_pre[i] = p;
	w.resid[i] = slot >= 0 ? (acc - a.q[slot]) / a.ws[slot] - p : 0.0;
}

// one workgroup: the residuals' maximum s over the shell, the features
__global__ __launch_bounds__(1024) void k_warm_features(WarmArgs w)
{
	__shared__ double sred[16];
	double rmax = 0.0;
	for (int i = threadIdx.x; i < w.n; i += blockDim.x) rmax = fmax(rmax, fabs(w.resid[i]));
	for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_xor(rmax, off, 64));
	if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = rmax;
	__syncthreads();
	double s = 0.0;
	for (int k = 0; k < (int) (blockDim.x >> 6); k++) s = fmax(s, sred[k]);
	if (threadIdx.x == 0) *w.scale = s;
	const double inv = s > 0.0 ? 1.0 / s : 0.0;
	for (int i = threadIdx.x; i < w.n; i += blockDim.x) {
		const int node = w.nodes[i];
		const bool unknown = w.slots[i] >= 0;
		float f[kMpnnIn];
		for (int k = 0; k < kMpnnIn; k++) f[k] = 0.0f;
		f[0] = (float) (w.resid[i] * inv);
		const int hop = w.hop[i];
		if (hop >= 0 && hop < 4) f[1 + hop] = 1.0f;
		f[5] = unknown ? 1.0f : 0.0f;
		f[6] = unknown ? 0.0f : 1.0f;
		float d2 = 0.0f;
		for (int c = 0; c < 3; c++) {
			const float d = (w.xyz[3 * (size_t) node + c] - w.v[c]) / w.ell;
			f[7 + c] = d;
			d2 += d * d;
		}
		f[10] = sqrtf(d2);
		f[11] = w.wv[i];
		f[12] = 1.0f;
		f[13] = (float) ((w.phi_pre[i] - w.phi_ref) * inv);
		for (int k = 0; k < kMpnnIn; k++) w.x[(size_t) i * kMpnnIn + k] = to_bf16(f[k]);
		w.mask[i] = unknown ? 1.0f : 0.0f;
	}
}

// phi_i = phi_pre_i + s * y_i on the shell's unknown rows (y null: restore phi_pre)
__global__ void k_warm_apply(WarmArgs w, const float* __restrict__ y)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= w.n || w.slots[i] < 0) return;
	double v = w.phi_pre[i];
	if (y) v += *w.scale * (double) y[i];
	w.phi[w.nodes[i]] = v;
}

// t_i = (phi_i - phi_pre_i) / s after the relaxation
__global__ void k_warm_targets(WarmArgs w)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= w.n) return;
	const double s = *w.scale;
	float t = 0.0f;
	if (w.slots[i] >= 0 && s > 0.0) t = (float) ((w.phi[w.nodes[i]] - w.phi_pre[i]) / s);
	w.t[i] = t;
}

}  // namespace dev

int launch_warm_features(hipStream_t s, const WarmArgs& w, const SweepArgs& a)
{
	if (w.n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_warm_resid, dim3((w.n + 3) / 4), dim3(256), 0, s, w, a);
	hipLaunchKernelGGL(dev::k_warm_features, dim3(1), dim3(1024), 0, s, w);
	return (int) hipGetLastError();
}

int launch_warm_apply(hipStream_t s, const WarmArgs& w, const float* y)
{
	if (w.n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_warm_apply, dim3((w.n + 255) / 256), dim3(256), 0, s, w, y);
	return (int) hipGetLastError();
}

int launch_warm_targets(hipStream_t s, const WarmArgs& w)
{
	if (w.n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_warm_targets, dim3((w.n + 255) / 256), dim3(256), 0, s, w);
	return (int) hipGetLastError();
}

}  // namespace tsg
