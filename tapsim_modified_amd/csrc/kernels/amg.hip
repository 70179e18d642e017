// Multigrid-preconditioned conjugate gradients for the correction equation of
// the accelerated relaxation (the inner solve of TSG_MIXED_F32_REFINE when the
// multigrid solver is selected).
//
// The correction equation delta = G delta + r (G_ij = c_ij / D_i over the
// unknown neighbours, D_i the row's coupling sum) is solved in its unscaled
// form A delta = D r with A = D (I - G): symmetric positive definite on the
// rows that are not Neumann rows (a Neumann row also sees the non-Neumann
// rows that do not see it; the outer FP64 refinement absorbs that slight
// asymmetry).  The preconditioner is one V-cycle of aggregation multigrid:
// piecewise-constant prolongation over aggregates of strongly coupled rows,
// Galerkin coarse operators, one damped-Jacobi step before and after each
// coarse correction, and the coarsest level solved with its dense inverse.
//
// Level 0 is the grid's own FP32 correction rows (SELL-64, node-indexed
// vectors); the coarse levels are CSR.  Every reduction is summed in a fixed
// order (block partials, then one block), so the solve is deterministic.
#include "kernels.hpp"
#include "rows.hpp"

namespace tsg {
namespace dev {

__device__ __forceinline__ double wave_sumd(double v)
{
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	return v;
}

// sum of v over the 256-thread block into part[blockIdx.x] (fixed order)
__device__ __forceinline__ void block_sum(double v, double* part)
{
	__shared__ double s[4];
	v = wave_sumd(v);
	if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

// the row operand of this thread on level 0 (G in over the correction rows),
// as the Chebyshev kernels read it
__device__ __forceinline__ int fine_row(const AccelArgs& x, const float* __restrict__ in, float* gin)
{
	return accel_row(x, in, gin, nullptr, nullptr, nullptr, nullptr);
}

// q = A p = D (p - G p); part += p.q
__global__ __launch_bounds__(256) void k_amg_fine_matvec(AccelArgs x, const float* __restrict__ dn,
                                                         const float* __restrict__ p, float* __restrict__ q,
                                                         double* __restrict__ part)
{
	float g;
	const int node = fine_row(x, p, &g);
	double c = 0.0;
	if (node >= 0) {
		const float pi = p[node];
		const float v = dn[node] * (pi - g);
		q[node] = v;
		c = (double) pi * (double) v;
	}
	block_sum(c, part);
}

// level-0 slot loop: node of slot i of this thread (or -1)
__device__ __forceinline__ int slot_node_of(const AccelArgs& x, int* slot_out)
{
	const int slot = blockIdx.x * blockDim.x + threadIdx.x;
	*slot_out = slot;
	if (slot >= x.nslots) return -1;
	return x.slot_node[slot];
}

// delta += alpha p, r -= alpha q; part += r.r
__global__ __launch_bounds__(256) void k_amg_fine_update(AccelArgs x, const double* __restrict__ st,
                                                         const float* __restrict__ p, const float* __restrict__ q,
                                                         float* __restrict__ delta, float* __restrict__ r,
                                                         double* __restrict__ part)
{
	int slot;
	const int node = slot_node_of(x, &slot);
	double c = 0.0;
	if (node >= 0) {
		const float alpha = (float) st[kAmgAlpha];
		delta[node] += alpha * p[node];
		const float rv = r[node] - alpha * q[node];
		r[node] = rv;
		c = (double) rv * (double) rv;
	}
	block_sum(c, part);
}

// r = D r32, delta = 0; part += r.r
__global__ __launch_bounds__(256) void k_amg_fine_start(AccelArgs x, const float* __restrict__ dn,
                                                        const float* __restrict__ r32, float* __restrict__ r,
                                                        float* __restrict__ delta, double* __restrict__ part)
{
	int slot;
	const int node = slot_node_of(x, &slot);
	double c = 0.0;
	if (node >= 0) {
		const float v = dn[node] * r32[node];
		r[node] = v;
		delta[node] = 0.0f;
		c = (double) v * (double) v;
	}
	block_sum(c, part);
}

// V-cycle on level 0: pre-smoothing from zero, x0 = w b / D
__global__ __launch_bounds__(256) void k_amg_fine_pre(AccelArgs x, const float* __restrict__ dn, const float* __restrict__ b,
                                                      float* __restrict__ x0, float w)
{
	int slot;
	const int node = slot_node_of(x, &slot);
	if (node >= 0) x0[node] = w * b[node] / dn[node];
}

// res = b - D (x0 - G x0)
__global__ __launch_bounds__(256) void k_amg_fine_resid(AccelArgs x, const float* __restrict__ dn,
                                                        const float* __restrict__ b, const float* __restrict__ x0,
                                                        float* __restrict__ res)
{
	float g;
	const int node = fine_row(x, x0, &g);
	if (node >= 0) res[node] = b[node] - dn[node] * (x0[node] - g);
}

// x += e[agg] (prolongation from level 1; rows outside every aggregate keep x)
__global__ __launch_bounds__(256) void k_amg_fine_prolong(AccelArgs x, const int* __restrict__ agg,
                                                          const float* __restrict__ e, float* __restrict__ xv)
{
	int slot;
	const int node = slot_node_of(x, &slot);
	if (node < 0) return;
	const int a = agg[node];
	if (a >= 0) xv[node] += e[a];
}

// z = xin + w (b / D - xin + G xin); part += b.z (b is the PCG residual)
__global__ __launch_bounds__(256) void k_amg_fine_post(AccelArgs x, const float* __restrict__ dn,
                                                       const float* __restrict__ b, const float* __restrict__ xin,
                                                       float* __restrict__ z, float w, double* __restrict__ part)
{
	float g;
	const int node = fine_row(x, xin, &g);
	double c = 0.0;
	if (node >= 0) {
		const float bi = b[node];
		const float xi = xin[node];
		const float v = xi + w * (bi / dn[node] - xi + g);
		z[node] = v;
		c = (double) bi * (double) v;
	}
	block_sum(c, part);
}

// p = z + beta p (beta = 0 on the first iteration)
__global__ __launch_bounds__(256) void k_amg_fine_dir(AccelArgs x, const double* __restrict__ st,
                                                      const float* __restrict__ z, float* __restrict__ p, int first)
{
	int slot;
	const int node = slot_node_of(x, &slot);
	if (node < 0) return;
	p[node] = first ? z[node] : z[node] + (float) st[kAmgBeta] * p[node];
}

// ---- coarse levels (CSR)

__global__ __launch_bounds__(256) void k_amg_pre(int n, const float* __restrict__ dinv, const float* __restrict__ b,
                                                 float* __restrict__ x0, float w)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) x0[i] = w * dinv[i] * b[i];
}

__global__ __launch_bounds__(256) void k_amg_resid(AmgLevel L, const float* __restrict__ b, const float* __restrict__ x0,
                                                   float* __restrict__ res)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= L.n) return;
	float s = 0.0f;
	for (int k = L.ptr[i]; k < L.ptr[i + 1]; k++) s += L.val[k] * x0[L.col[k]];
	res[i] = b[i] - s;
}

// bc[I] = sum of res over the members of aggregate I (fixed member order)
__global__ __launch_bounds__(256) void k_amg_restrict(int nc, const int* __restrict__ rptr, const int* __restrict__ ridx,
                                                      const float* __restrict__ res, float* __restrict__ bc)
{
	const int I = blockIdx.x * blockDim.x + threadIdx.x;
	if (I >= nc) return;
	float s = 0.0f;
	for (int k = rptr[I]; k < rptr[I + 1]; k++) s += res[ridx[k]];
	bc[I] = s;
}

__global__ __launch_bounds__(256) void k_amg_prolong(int n, const int* __restrict__ agg, const float* __restrict__ e,
                                                     float* __restrict__ xv)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int a = agg[i];
	if (a >= 0) xv[i] += e[a];
}

// xout = xin + w dinv (b - A xin)
__global__ __launch_bounds__(256) void k_amg_post(AmgLevel L, const float* __restrict__ b, const float* __restrict__ xin,
                                                  float* __restrict__ xout, float w)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= L.n) return;
	float s = 0.0f;
	for (int k = L.ptr[i]; k < L.ptr[i + 1]; k++) s += L.val[k] * xin[L.col[k]];
	xout[i] = xin[i] + w * L.dinv[i] * (b[i] - s);
}

// x = Minv b on the coarsest level (one wavefront per row, fixed-order sum)
__global__ __launch_bounds__(256) void k_amg_dense(int n, const float* __restrict__ minv, const float* __restrict__ b,
                                                   float* __restrict__ xv)
{
	const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (row >= n) return;
	float s = 0.0f;
	for (int k = lane; k < n; k += 64) s += minv[(size_t) row * n + k] * b[k];
	for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
	if (lane == 0) xv[row] = s;
}

// the PCG scalars from block partials (one block): mode 0 p.q -> alpha;
// 1 r.r -> rr (2: also rr0); 3 r.z -> beta and rz (4: first rz)
__global__ __launch_bounds__(256) void k_amg_scalars(int nblocks, const double* __restrict__ part, double* __restrict__ st,
                                                     int mode)
{
	double v = 0.0;
	for (int b = threadIdx.x; b < nblocks; b += 256) v += part[b];
	__shared__ double s[4];
	v = wave_sumd(v);
	if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x != 0) return;
	const double t = (s[0] + s[1]) + (s[2] + s[3]);
	if (mode == 0) {
		st[kAmgAlpha] = t != 0.0 ? st[kAmgRz] / t : 0.0;
	} else if (mode == 1 || mode == 2) {
		st[kAmgRr] = t;
		if (mode == 2) st[kAmgRr0] = t;
	} else {
		if (mode == 3) st[kAmgBeta] = st[kAmgRz] != 0.0 ? t / st[kAmgRz] : 0.0;
		st[kAmgRz] = t;
	}
}

}  // namespace dev

static inline int blocks_of(long long n) { return (int) ((n + 255) / 256); }

int amg_fine_blocks(const AccelArgs& x) { return row_launch_blocks(x.nslots, x.nlong); }
int amg_slot_blocks(const AccelArgs& x) { return blocks_of(x.nslots); }

int launch_amg_start(hipStream_t s, const AccelArgs& x, const float* dn, const float* r32, float* r, float* delta,
                     double* part, double* st)
{
	hipLaunchKernelGGL(dev::k_amg_fine_start, dim3(amg_slot_blocks(x)), dim3(256), 0, s, x, dn, r32, r, delta, part);
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(256), 0, s, amg_slot_blocks(x), (const double*) part, st, 2);
	return (int) hipGetLastError();
}

// z = V(b): one V-cycle over the levels; lv[0] is the first coarse level,
// lv[nlev-1] the coarsest (solved with minv); bl, xl, rl are per-level
// right-hand side, iterate and scratch vectors
int launch_amg_vcycle(hipStream_t s, const AccelArgs& x, const float* dn, const int* agg0, float w0, const AmgLevel* lv,
                      const float* w, int nlev, const float* b, float* x0, float* res, float* z, float* const* bl,
                      float* const* xl, float* const* rl, const float* minv, double* part)
{
	const int fb = amg_fine_blocks(x), sb = amg_slot_blocks(x);
	hipLaunchKernelGGL(dev::k_amg_fine_pre, dim3(sb), dim3(256), 0, s, x, dn, b, x0, w0);
	hipLaunchKernelGGL(dev::k_amg_fine_resid, dim3(fb), dim3(256), 0, s, x, dn, b, (const float*) x0, res);
	const float* rfine = res;
	for (int l = 0; l < nlev; l++) {
		const AmgLevel& L = lv[l];
		hipLaunchKernelGGL(dev::k_amg_restrict, dim3(blocks_of(L.n)), dim3(256), 0, s, L.n, L.rptr, L.ridx, rfine, bl[l]);
		if (l == nlev - 1) {
			hipLaunchKernelGGL(dev::k_amg_dense, dim3((L.n + 3) / 4), dim3(256), 0, s, L.n, minv, (const float*) bl[l],
			                   xl[l]);
			break;
		}
		hipLaunchKernelGGL(dev::k_amg_pre, dim3(blocks_of(L.n)), dim3(256), 0, s, L.n, L.dinv, (const float*) bl[l], xl[l],
		                   w[l]);
		hipLaunchKernelGGL(dev::k_amg_resid, dim3(blocks_of(L.n)), dim3(256), 0, s, L, (const float*) bl[l],
		                   (const float*) xl[l], rl[l]);
		rfine = rl[l];
	}
	// up: the final iterate of level l+1 corrects level l, then one smoothing
	// step writes level l's final iterate into its scratch vector
	const float* fin = xl[nlev - 1];
	for (int l = nlev - 2; l >= 0; l--) {
		const AmgLevel& L = lv[l];
		const AmgLevel& C = lv[l + 1];
		hipLaunchKernelGGL(dev::k_amg_prolong, dim3(blocks_of(L.n)), dim3(256), 0, s, L.n, C.agg, fin, xl[l]);
		hipLaunchKernelGGL(dev::k_amg_post, dim3(blocks_of(L.n)), dim3(256), 0, s, L, (const float*) bl[l],
		                   (const float*) xl[l], rl[l], w[l]);
		fin = rl[l];
	}
	hipLaunchKernelGGL(dev::k_amg_fine_prolong, dim3(sb), dim3(256), 0, s, x, agg0, fin, x0);
	hipLaunchKernelGGL(dev::k_amg_fine_post, dim3(fb), dim3(256), 0, s, x, dn, b, (const float*) x0, z, w0, part);
	return (int) hipGetLastError();
}

int launch_amg_scalars(hipStream_t s, int nblocks, const double* part, double* st, int mode)
{
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(256), 0, s, nblocks, part, st, mode);
	return (int) hipGetLastError();
}

int launch_amg_dir(hipStream_t s, const AccelArgs& x, const double* st, const float* z, float* p, int first)
{
	hipLaunchKernelGGL(dev::k_amg_fine_dir, dim3(amg_slot_blocks(x)), dim3(256), 0, s, x, st, z, p, first);
	return (int) hipGetLastError();
}

int launch_amg_matvec_update(hipStream_t s, const AccelArgs& x, const float* dn, const float* p, float* q, float* delta,
                             float* r, double* part, double* st)
{
	const int fb = amg_fine_blocks(x), sb = amg_slot_blocks(x);
	hipLaunchKernelGGL(dev::k_amg_fine_matvec, dim3(fb), dim3(256), 0, s, x, dn, p, q, part);
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(256), 0, s, fb, (const double*) part, st, 0);
	hipLaunchKernelGGL(dev::k_amg_fine_update, dim3(sb), dim3(256), 0, s, x, (const double*) st, p, (const float*) q, delta,
	                   r, part);
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(256), 0, s, sb, (const double*) part, st, 1);
	return (int) hipGetLastError();
}

}  // namespace tsg
