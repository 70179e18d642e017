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
// Level 0 is the grid's own FP32 correction rows (SELL-64; the vectors are
// indexed by row slot, dense over the unknown rows, and the rows' columns are
// slots); the coarse levels are CSR.  Every reduction is summed in a fixed
// order (block partials, then one block sums them in block order), so the
// solve is deterministic.  Smoothing steps are fused into the
// neighbouring kernels: the pre-smoothing from zero (x0 = w b / D) is formed
// where b is formed and recomputed on the fly at the neighbours, the
// prolongation is applied on the fly inside the post-smoothing.
#include "kernels.hpp"
#include "rows.hpp"

namespace tsg {
namespace dev {

__device__ __forceinline__ double wave_sumd(double v)
{
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	return v;
}

// the sum of v over the 256-thread block goes to part[blockIdx.x] (fixed order)
__device__ __forceinline__ void block_sum(double v, double* part)
{
	__shared__ double s[4];
	v = wave_sumd(v);
	if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

// one block sums the partials in block order and applies `mode` to the
// scalars: 0 p.q -> alpha = rz / p.q; 1 r.r -> rr, one more iteration,
// done once rr <= tol2 rr0; 2 r.r -> rr, rr0 (the start: tol2 stored, done
// and the iteration count cleared); 3 r.z -> beta = r.z / rz, rz; 4 r.z -> rz.
// (A last-block-done fusion into the producing kernel needs a device-scope
// fence per block, which writes back the L2 and costs more than this launch.)
__global__ __launch_bounds__(1024) void k_amg_scalars(int nblocks, const double* __restrict__ part, double* __restrict__ st,
                                                      int mode, double tol2)
{
	if (mode != 2 && st[kAmgDone] != 0.0) return;
	double t = 0.0;
	for (int b = threadIdx.x; b < nblocks; b += 1024) t += part[b];
	__shared__ double s[16];
	t = wave_sumd(t);
	if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = t;
	__syncthreads();
	if (threadIdx.x != 0) return;
	t = 0.0;
	for (int k = 0; k < 16; k++) t += s[k];
	if (mode == 0) {
		st[kAmgAlpha] = t != 0.0 ? st[kAmgRz] / t : 0.0;
	} else if (mode == 1) {
		st[kAmgRr] = t;
		st[kAmgIters] += 1.0;
		if (t <= st[kAmgTol2] * st[kAmgRr0] || t != t) st[kAmgDone] = 1.0;
	} else if (mode == 2) {
		st[kAmgRr] = t;
		st[kAmgRr0] = t;
		st[kAmgTol2] = tol2;
		st[kAmgIters] = 0.0;
		st[kAmgDone] = 0.0;
	} else {
		if (mode == 3) st[kAmgBeta] = st[kAmgRz] != 0.0 ? t / st[kAmgRz] : 0.0;
		st[kAmgRz] = t;
	}
}

// the solve has met its target: the remaining launches return at once
__device__ __forceinline__ bool amg_done(const double* st) { return st[kAmgDone] != 0.0; }

// level-0 slot loop: this thread's slot when it holds a row (or -1)
__device__ __forceinline__ int active_slot(const AccelArgs& x)
{
	const int slot = blockIdx.x * blockDim.x + threadIdx.x;
	if (slot >= x.nslots || x.slot_node[slot] < 0) return -1;
	return slot;
}

// the row operand of this thread on level 0 (G in over the correction rows);
// returns the slot
__device__ __forceinline__ int fine_row(const AccelArgs& x, const float* __restrict__ in, float* gin)
{
	return accel_row(x, in, gin, nullptr, nullptr, nullptr, nullptr);
}

// r = D r32, delta = 0, x0 = w0 r / D; partials of r.r
__global__ __launch_bounds__(256) void k_amg_fine_start(AmgSolve S, const float* __restrict__ r32)
{
	const int i = active_slot(S.x);
	double c = 0.0;
	if (i >= 0) {
		const float d = S.dn[i];
		const float v = d * r32[i];
		S.r[i] = v;
		S.delta[i] = 0.0f;
		S.x0[i] = S.w0 * v / d;
		c = (double) v * (double) v;
	}
	block_sum(c, S.part);
}

// q = A p = D (p - G p); partials of p.q
__global__ __launch_bounds__(256) void k_amg_fine_matvec(AmgSolve S)
{
	if (amg_done(S.st)) return;
	float g;
	const int i = fine_row(S.x, S.p, &g);
	double c = 0.0;
	if (i >= 0) {
		const float pi = S.p[i];
		const float v = S.dn[i] * (pi - g);
		S.q[i] = v;
		c = (double) pi * (double) v;
	}
	block_sum(c, S.part);
}

// delta += alpha p, r -= alpha q, x0 = w0 r / D; partials of r.r
__global__ __launch_bounds__(256) void k_amg_fine_update(AmgSolve S)
{
	if (amg_done(S.st)) return;
	const int i = active_slot(S.x);
	double c = 0.0;
	if (i >= 0) {
		const float alpha = (float) S.st[kAmgAlpha];
		S.delta[i] += alpha * S.p[i];
		const float rv = S.r[i] - alpha * S.q[i];
		S.r[i] = rv;
		S.x0[i] = S.w0 * rv / S.dn[i];
		c = (double) rv * (double) rv;
	}
	block_sum(c, S.part);
}

// res = r - D (x0 - G x0)
__global__ __launch_bounds__(256) void k_amg_fine_resid(AmgSolve S)
{
	if (amg_done(S.st)) return;
	float g;
	const int i = fine_row(S.x, S.x0, &g);
	if (i >= 0) S.res[i] = S.r[i] - S.dn[i] * (S.x0[i] - g);
}

// x0 += wc e[agg0] (the coarse correction; rows outside every aggregate keep x0)
__global__ __launch_bounds__(256) void k_amg_fine_prolong(AmgSolve S, const float* __restrict__ e)
{
	if (amg_done(S.st)) return;
	const int i = active_slot(S.x);
	if (i < 0) return;
	const int a = S.agg0[i];
	if (a >= 0) S.x0[i] += S.wc * e[a];
}

// z = x0 + w0 (r / D - x0 + G x0); partials of r.z
__global__ __launch_bounds__(256) void k_amg_fine_post(AmgSolve S)
{
	if (amg_done(S.st)) return;
	float g;
	const int i = fine_row(S.x, S.x0, &g);
	double c = 0.0;
	if (i >= 0) {
		const float bi = S.r[i];
		const float xi = S.x0[i];
		const float v = xi + S.w0 * (bi / S.dn[i] - xi + g);
		S.z[i] = v;
		c = (double) bi * (double) v;
	}
	block_sum(c, S.part);
}

// p = z + beta p (beta = 0 on the first iteration)
__global__ __launch_bounds__(256) void k_amg_fine_dir(AmgSolve S, int first)
{
	if (amg_done(S.st)) return;
	const int i = active_slot(S.x);
	if (i < 0) return;
	S.p[i] = first ? S.z[i] : S.z[i] + (float) S.st[kAmgBeta] * S.p[i];
}

// ---- coarse levels (CSR)

// Coarse rows hold ~20-40 entries: eight lanes share a row (entries k, k+8,
// ... per lane, then a fixed butterfly over the eight), so the row's column
// and value loads are contiguous across the lanes.
constexpr int kCsrLanes = 8;

__device__ __forceinline__ float row_group_sum(float v)
{
	v += __shfl_xor(v, 4, 64);
	v += __shfl_xor(v, 2, 64);
	v += __shfl_xor(v, 1, 64);
	return v;
}

// pre-smoothing from zero and its residual in one pass: x0 = w Dinv b,
// res = b - A x0 (x0 at the neighbours recomputed from b)
__global__ __launch_bounds__(256) void k_amg_pre_resid(const double* __restrict__ st, AmgLevel L, float w,
                                                       const float* __restrict__ b, float* __restrict__ x0,
                                                       float* __restrict__ res)
{
	if (amg_done(st)) return;
	const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kCsrLanes;
	const int sub = threadIdx.x % kCsrLanes;
	float s = 0.0f;
	if (i < L.n)
		for (int k = L.ptr[i] + sub; k < L.ptr[i + 1]; k += kCsrLanes) {
			const int j = L.col[k];
			s += L.val[k] * (w * L.dinv[j] * b[j]);
		}
	s = row_group_sum(s);
	if (i >= L.n || sub != 0) return;
	const float bi = b[i];
	x0[i] = w * L.dinv[i] * bi;
	res[i] = bi - s;
}

// bc[I] = sum of res over the members of aggregate I (fixed member order per
// lane, then the butterfly)
__global__ __launch_bounds__(256) void k_amg_restrict(const double* __restrict__ st, int nc, const int* __restrict__ rptr,
                                                      const int* __restrict__ ridx, const float* __restrict__ res,
                                                      float* __restrict__ bc)
{
	if (amg_done(st)) return;
	const int I = (blockIdx.x * blockDim.x + threadIdx.x) / kCsrLanes;
	const int sub = threadIdx.x % kCsrLanes;
	float s = 0.0f;
	if (I < nc)
		for (int k = rptr[I] + sub; k < rptr[I + 1]; k += kCsrLanes) s += res[ridx[k]];
	s = row_group_sum(s);
	if (I < nc && sub == 0) bc[I] = s;
}

// the coarse correction and the post-smoothing in one pass:
// xin = x0 + wc e[agg], xout = xin + w Dinv (b - A xin) (xin at the
// neighbours recomputed from x0 and e)
__global__ __launch_bounds__(256) void k_amg_prolong_post(const double* __restrict__ st, AmgLevel L, float w, float wc,
                                                          const int* __restrict__ agg, const float* __restrict__ e,
                                                          const float* __restrict__ b, const float* __restrict__ x0,
                                                          float* __restrict__ xout)
{
	if (amg_done(st)) return;
	const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kCsrLanes;
	const int sub = threadIdx.x % kCsrLanes;
	float s = 0.0f;
	if (i < L.n)
		for (int k = L.ptr[i] + sub; k < L.ptr[i + 1]; k += kCsrLanes) {
			const int j = L.col[k];
			const int a = agg[j];
			s += L.val[k] * (a >= 0 ? x0[j] + wc * e[a] : x0[j]);
		}
	s = row_group_sum(s);
	if (i >= L.n || sub != 0) return;
	const int a = agg[i];
	const float xi = a >= 0 ? x0[i] + wc * e[a] : x0[i];
	xout[i] = xi + w * L.dinv[i] * (b[i] - s);
}

// x = Minv b on the coarsest level (one wavefront per row, fixed-order sum)
__global__ __launch_bounds__(256) void k_amg_dense(const double* __restrict__ st, int n, const float* __restrict__ minv,
                                                   const float* __restrict__ b, float* __restrict__ xv)
{
	if (amg_done(st)) return;
	const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (row >= n) return;
	float s = 0.0f;
	for (int k = lane; k < n; k += 64) s += minv[(size_t) row * n + k] * b[k];
	for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
	if (lane == 0) xv[row] = s;
}

}  // namespace dev

static inline int blocks_of(long long n) { return (int) ((n + 255) / 256); }
static inline int csr_blocks(long long rows) { return blocks_of(rows * dev::kCsrLanes); }

int amg_fine_blocks(const AccelArgs& x) { return row_launch_blocks(x.nslots, x.nlong); }
int amg_slot_blocks(const AccelArgs& x) { return blocks_of(x.nslots); }

int launch_amg_start(hipStream_t s, const AmgSolve& S, const float* r32, double tol2)
{
	hipLaunchKernelGGL(dev::k_amg_fine_start, dim3(amg_slot_blocks(S.x)), dim3(256), 0, s, S, r32);
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(1024), 0, s, amg_slot_blocks(S.x), (const double*) S.part, S.st, 2,
	                   tol2);
	return (int) hipGetLastError();
}

// z = V(r): S.x0 holds the level-0 pre-smoothing (formed with r); lv[0] is
// the first coarse level, lv[nlev-1] the coarsest (solved with minv)
int launch_amg_vcycle(hipStream_t s, const AmgSolve& S, int mode)
{
	const int fb = amg_fine_blocks(S.x), sb = amg_slot_blocks(S.x);
	hipLaunchKernelGGL(dev::k_amg_fine_resid, dim3(fb), dim3(256), 0, s, S);
	const float* rfine = S.res;
	for (int l = 0; l < S.nlev; l++) {
		const AmgLevel& L = S.lv[l];
		hipLaunchKernelGGL(dev::k_amg_restrict, dim3(csr_blocks(L.n)), dim3(256), 0, s, (const double*) S.st, L.n, L.rptr, L.ridx, rfine,
		                   S.bl[l]);
		if (l == S.nlev - 1) {
			hipLaunchKernelGGL(dev::k_amg_dense, dim3((L.n + 3) / 4), dim3(256), 0, s, (const double*) S.st, L.n, S.minv,
			                   (const float*) S.bl[l], S.xl[l]);
			break;
		}
		hipLaunchKernelGGL(dev::k_amg_pre_resid, dim3(csr_blocks(L.n)), dim3(256), 0, s, (const double*) S.st, L, S.w[l],
		                   (const float*) S.bl[l], S.xl[l], S.rl[l]);
		rfine = S.rl[l];
	}
	// up: level l+1's final iterate corrects level l inside its
	// post-smoothing, whose result goes to level l's scratch vector
	const float* fin = S.xl[S.nlev - 1];
	for (int l = S.nlev - 2; l >= 0; l--) {
		const AmgLevel& L = S.lv[l];
		hipLaunchKernelGGL(dev::k_amg_prolong_post, dim3(csr_blocks(L.n)), dim3(256), 0, s, (const double*) S.st, L, S.w[l], S.wc, S.lv[l + 1].agg,
		                   fin, (const float*) S.bl[l], (const float*) S.xl[l], S.rl[l]);
		fin = S.rl[l];
	}
	hipLaunchKernelGGL(dev::k_amg_fine_prolong, dim3(sb), dim3(256), 0, s, S, fin);
	hipLaunchKernelGGL(dev::k_amg_fine_post, dim3(fb), dim3(256), 0, s, S);
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(1024), 0, s, fb, (const double*) S.part, S.st, mode, 0.0);
	return (int) hipGetLastError();
}

int launch_amg_dir(hipStream_t s, const AmgSolve& S, int first)
{
	hipLaunchKernelGGL(dev::k_amg_fine_dir, dim3(amg_slot_blocks(S.x)), dim3(256), 0, s, S, first);
	return (int) hipGetLastError();
}

int launch_amg_fine_matvec(hipStream_t s, const AmgSolve& S)
{
	hipLaunchKernelGGL(dev::k_amg_fine_matvec, dim3(amg_fine_blocks(S.x)), dim3(256), 0, s, S);
	return (int) hipGetLastError();
}

int launch_amg_matvec_update(hipStream_t s, const AmgSolve& S)
{
	const int fb = amg_fine_blocks(S.x), sb = amg_slot_blocks(S.x);
	hipLaunchKernelGGL(dev::k_amg_fine_matvec, dim3(fb), dim3(256), 0, s, S);
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(1024), 0, s, fb, (const double*) S.part, S.st, 0, 0.0);
	hipLaunchKernelGGL(dev::k_amg_fine_update, dim3(sb), dim3(256), 0, s, S);
	hipLaunchKernelGGL(dev::k_amg_scalars, dim3(1), dim3(1024), 0, s, sb, (const double*) S.part, S.st, 1, 0.0);
	return (int) hipGetLastError();
}

}  // namespace tsg
