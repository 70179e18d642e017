// Accelerated relaxation (the TSG_MIXED_F32_REFINE method): iterative
// refinement whose residual and update run in FP64 and whose correction
// solve runs as Chebyshev-accelerated Jacobi sweeps in FP32.
//
// Scaled residual (the method-independent stopping rule; for a Jacobi sweep
// it is exactly the reference's deviation, grid_3d.cpp:597-605):
//     r_i = J(phi)_i - phi_i,   J(phi)_i = (sum_j c_ij phi_j - q_i) / sum_j c_ij
// Correction equation on the unknown rows (Dirichlet neighbours drop out):
//     delta = G delta + r,      G_ij = c_ij / sum_k c_ik   (j unknown)
// solved with the Chebyshev semi-iteration on the Jacobi sweep
//     delta_{k+1} = w_{k+1} (G delta_k + r - delta_{k-1}) + delta_{k-1}
// for a spectral radius bound rho of G (Varga, Matrix Iterative Analysis, 5.1).
#include "device_math.hpp"
#include "kernels.hpp"
#include "rows.hpp"

namespace tsg {
namespace dev {

__device__ __forceinline__ double wave_maxd(double v)
{
	for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
	return v;
}

// FP64 residual of the current potentials; writes r (float) per row slot and
// the per-block max |r|; slots without a node are skipped
__global__ __launch_bounds__(256) void k_resid64(SweepArgs a, const double* __restrict__ phi, float* __restrict__ r32)
{
	const int main_blocks = (a.nslots + 255) >> 8;
	const RowBlock rb = row_block(a.nlong);
	const int lb = rb.is_long ? -1 : rb.index;
	int slot = -1, node = -1, deg = 0;
	long long soff = 0;
	double q = 0.0, ws = 1.0;
	if (!rb.is_long) {
		if (lb < main_blocks) slot = lb * 256 + threadIdx.x;
		if (slot >= 0 && slot < a.nslots) {
			node = a.slot_node[slot];
			deg = a.slot_deg[slot];
			soff = a.slice_off[slot >> 6];
			q = a.q[slot];
			ws = a.ws[slot];
		}
	} else {
		const int li = long_row_index(rb);
		if (li < a.nlong) {
			slot = a.long_slots[li];
			node = a.slot_node[slot];
			deg = a.slot_deg[slot];
			soff = a.long_off[li];  // the contiguous copy
			q = a.q[slot];
			ws = a.ws[slot];
		}
	}
	const long long base = soff + (slot & 63);
	double lv = 0.0, old = 0.0;
	if (!rb.is_long) {
		const bool act = node >= 0 && deg <= kLongRow;
		if (act) old = phi[node];
		else node = -1;
		lv = short_row_dot_exact<double>(a.col, (const double*) a.coup, phi, act, deg, base);
	} else if (node >= 0) {
		old = phi[node];
		lv = long_row_dot_exact<double>(a.lcol, (const double*) a.lcoup, phi, deg, soff);
		if (lane_id() != 0) node = -1;
	}
	double rmax = 0.0;
	int nan = 0;
	if (node >= 0) {
		lv -= q;
		lv /= ws;
		const double r = lv - old;
		if (r != r) nan = 1;
		r32[slot] = (float) r;
		rmax = fabs(r);
	}
	__shared__ double smax[4];
	__shared__ int snan[4];
	rmax = wave_maxd(rmax);
	const int anynan = __any(nan);
	if ((threadIdx.x & 63) == 0) { smax[threadIdx.x >> 6] = rmax; snan[threadIdx.x >> 6] = anynan; }
	__syncthreads();
	if (threadIdx.x == 0) {
		a.block_max[blockIdx.x] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
		a.block_flag[blockIdx.x] = snan[0] | snan[1] | snan[2] | snan[3];
	}
}

// one-block reduction of the per-block maxima into out[0] (and NaN into out[1])
__global__ __launch_bounds__(256) void k_reduce_max(int nblocks, const double* block_max, const int* block_flag, double* out)
{
	double m = 0.0;
	int nan = 0;
	for (int b = threadIdx.x; b < nblocks; b += 256) {
		m = fmax(m, block_max[b]);
		nan |= block_flag[b];
	}
	__shared__ double smax[4];
	__shared__ int snan[4];
	m = wave_maxd(m);
	const int anynan = __any(nan);
	if ((threadIdx.x & 63) == 0) { smax[threadIdx.x >> 6] = m; snan[threadIdx.x >> 6] = anynan; }
	__syncthreads();
	if (threadIdx.x == 0) {
		out[0] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
		out[1] = (snan[0] | snan[1] | snan[2] | snan[3]) ? 1.0 : 0.0;
	}
}

// delta_prev = 0, delta_cur = r on the rows (the first Jacobi step from 0)
__global__ void k_cheb_init(AccelArgs x, const float* __restrict__ r32, float* __restrict__ dprev, float* __restrict__ dcur)
{
	const int slot = blockIdx.x * blockDim.x + threadIdx.x;
	if (slot >= x.nslots || x.slot_node[slot] < 0) return;
	dprev[slot] = 0.0f;
	dcur[slot] = r32[slot];
}

// one Chebyshev step in FP32: dnext overwrites dprev slot by slot
__global__ __launch_bounds__(256) void k_cheb32(AccelArgs x, const float* __restrict__ r32, const float* __restrict__ dcur,
                                              float* __restrict__ dprev_next, float omega)
{
	float s, r = 0.0f, prev = 0.0f;
	const int i = accel_row(x, dcur, &s, r32, &r, dprev_next, &prev);
	if (i < 0) return;
	dprev_next[i] = fmaf(omega, s + r - prev, prev);
}

// phi += delta on the rows, in FP64 (delta by slot, phi by node); both
// potential buffers take the result
__global__ void k_apply64(AccelArgs x, const float* __restrict__ delta, double* __restrict__ phi, double* __restrict__ phi_other)
{
	const int slot = blockIdx.x * blockDim.x + threadIdx.x;
	if (slot >= x.nslots) return;
	const int node = x.slot_node[slot];
	if (node < 0) return;
	const double v = phi[node] + (double) delta[slot];
	phi[node] = v;
	if (phi_other) phi_other[node] = v;
}

// y = G x (power iteration for the spectral radius)
__global__ __launch_bounds__(256) void k_apply_g(AccelArgs x, const float* __restrict__ in, float* __restrict__ out)
{
	float s;
	const int i = accel_row(x, in, &s, nullptr, nullptr, nullptr, nullptr);
	if (i >= 0) out[i] = s;
}

}  // namespace dev

int launch_resid64(hipStream_t s, const SweepArgs& a, const double* phi, float* r32, double* out)
{
	const int blocks = row_launch_blocks(a.nslots, a.nlong);
	if (blocks == 0) return 0;
	hipLaunchKernelGGL(dev::k_resid64, dim3(blocks), dim3(256), 0, s, a, phi, r32);
	hipLaunchKernelGGL(dev::k_reduce_max, dim3(1), dim3(256), 0, s, blocks, (const double*) a.block_max,
	                   (const int*) a.block_flag, out);
	return (int) hipGetLastError();
}

int launch_cheb_init(hipStream_t s, const AccelArgs& x, const float* r32, float* dprev, float* dcur)
{
	hipLaunchKernelGGL(dev::k_cheb_init, dim3((x.nslots + 255) / 256), dim3(256), 0, s, x, r32, dprev, dcur);
	return (int) hipGetLastError();
}

int launch_cheb_step(hipStream_t s, const AccelArgs& x, const float* r32, const float* dcur, float* dprev_next, float omega)
{
	hipLaunchKernelGGL(dev::k_cheb32, dim3(row_launch_blocks(x.nslots, x.nlong)), dim3(256), 0, s, x, r32, dcur, dprev_next, omega);
	return (int) hipGetLastError();
}

int launch_apply64(hipStream_t s, const AccelArgs& x, const float* delta, double* phi, double* phi_other)
{
	hipLaunchKernelGGL(dev::k_apply64, dim3((x.nslots + 255) / 256), dim3(256), 0, s, x, delta, phi, phi_other);
	return (int) hipGetLastError();
}

int launch_apply_g(hipStream_t s, const AccelArgs& x, const float* in, float* out)
{
	hipLaunchKernelGGL(dev::k_apply_g, dim3(row_launch_blocks(x.nslots, x.nlong)), dim3(256), 0, s, x, in, out);
	return (int) hipGetLastError();
}

}  // namespace tsg
