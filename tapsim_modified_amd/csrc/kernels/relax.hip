// Relaxation kernels: one Jacobi cycle over the unknown rows (global
// relaxation) and the whole local relaxation of a node shell.
//
// Storage: the unknown rows (non-Dirichlet nodes) live in row slots laid
// out as SELL-64: slice s holds slots 64s..64s+63, entry k of the slot in
// lane l sits at slice_off[s] + 64k + l, so a wavefront reads 64 consecutive
// column indices and couplings per step.  A row lists the neighbours that
// take part in Grid_3d::Table::computePotential (grid_3d.cpp:1317-1371) in
// ascending node index -- all neighbours for a Neumann row, the non-Neumann
// ones otherwise -- and its coupling sum is precomputed on the host in that
// same order, so the value is bit-identical to the reference's
//     lv = sum_k (double) phi[j_k] * c_k ;  lv -= charge ;  lv /= sum_k c_k
// (double accumulation for both float and double storage; the build uses
// -ffp-contract=off so no product is fused into an add).
#include <algorithm>

#include "device_math.hpp"
#include "kernels.hpp"
#include "rows.hpp"

namespace tsg {
namespace dev {

__device__ __forceinline__ double wave_max(double v)
{
	for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
	return v;
}

// the deviation queue and least-squares slope test of
// Grid_3d::Table::relax(threshold,cycleSize,queueSize), grid_3d.cpp:621-654;
// returns true when the slope is exactly zero (RelaxationFault)
__device__ bool push_and_test(double* queue, int* qhead, int* qsize, int qcap, int queueSize, int cycle, double dev,
                              double* slope_out)
{
	if (queueSize <= 0) return false;
	if (*qsize < qcap) {
		queue[(*qhead + *qsize) % qcap] = dev;
		(*qsize)++;
	} else {
		queue[*qhead] = dev;
		*qhead = (*qhead + 1) % qcap;
	}
	while (*qsize > queueSize) { *qhead = (*qhead + 1) % qcap; (*qsize)--; }
	if (*qsize != queueSize) return false;
	const double xMean = (double) ((float) (queueSize - 1) / 2.0f);
	double yMean = (double) 0.0f;
	for (int i = 0; i < *qsize; i++) yMean += queue[(*qhead + i) % qcap];
	yMean /= (double) (unsigned) queueSize;
	double nom = (double) 0.0f;
	for (int x = 0; x < *qsize; x++) nom += (x - xMean) * (queue[(*qhead + x) % qcap] - yMean);
	double den = (double) (unsigned) queueSize;
	den *= (double) (unsigned) queueSize;
	den -= 1;
	den *= (double) (unsigned) queueSize;
	den *= (double) (unsigned) cycle;
	den /= 12.0f;
	const double slope = nom / den;
	*slope_out = slope;
	return slope == 0.0f;
}

// value of one row (computePotential) from the given read buffer.  Loads
// are issued eight entries at a time (column indices, then couplings and the
// gathered potentials) so that a wavefront keeps many requests in flight;
// the products are then added strictly in entry order, which keeps the sum
// bit-identical to the reference's sequential loop.
template <typename R, bool kAtomicRead>
__device__ __forceinline__ double row_value(const SweepArgs& a, int slot, const R* __restrict__ pin)
{
	constexpr int B = 8;
	const int deg = a.slot_deg[slot];
	const long long base = a.slice_off[slot >> 6] + (slot & 63);
	const R* __restrict__ coup = (const R*) a.coup;
	const int* __restrict__ col = a.col;
	double lv = 0.0;
	for (int k = 0; k < deg; k += B) {
		int j[B];
		R c[B];
		R p[B];
#pragma unroll
		for (int u = 0; u < B; u++) j[u] = (k + u < deg) ? col[base + 64ll * (k + u)] : 0;
#pragma unroll
		for (int u = 0; u < B; u++) c[u] = (k + u < deg) ? coup[base + 64ll * (k + u)] : (R) 0;
#pragma unroll
		for (int u = 0; u < B; u++) {
			if (k + u < deg) {
				if constexpr (kAtomicRead)
					p[u] = __hip_atomic_load(pin + j[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				else
					p[u] = pin[j[u]];
			} else
				p[u] = (R) 0;
		}
#pragma unroll
		for (int u = 0; u < B; u++) {
			if (k + u < deg) {
				double adj = (double) p[u];
				adj *= (double) c[u];
				lv += adj;
			}
		}
	}
	lv -= a.q[slot];
	lv /= a.ws[slot];
	return lv;
}

// One Jacobi cycle: each block writes its max |dphi| (and a NaN flag) to its
// own slot; the cycle's decision is taken by k_finalize.  No fence or atomic
// per block: on a multi-XCD device an agent-scope fence writes back the L2,
// and same-address atomics from thousands of blocks serialise.
template <typename R>
__global__ __launch_bounds__(256) void k_jacobi(SweepArgs a, int read)
{
	const R* pin = (const R*) (read ? a.phi1 : a.phi0);
	R* pout = (R*) (read ? a.phi0 : a.phi1);
	const int main_blocks = (a.nslots + 255) >> 8;
	const RowBlock rb = row_block(a.nlong);
	const int lb = rb.is_long ? -1 : rb.index;
	// the per-slot record is loaded in one round before any branch on it
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
	if (a.ctl->done) return;  // uniform per launch: converged earlier in this batch
	const long long base = soff + (slot & 63);
	double lv = 0.0;
	R old = (R) 0;
	if (!rb.is_long) {  // block-uniform
		const bool act = node >= 0 && deg <= kLongRow;  // else empty slot, or on the long-row list
		if (act) old = pin[node];
		else node = -1;
		lv = short_row_dot_exact<R>(a.col, (const R*) a.coup, pin, act, deg, base);
	} else if (node >= 0) {  // wave-uniform
		old = pin[node];
		lv = long_row_dot_exact<R>(a.lcol, (const R*) a.lcoup, pin, deg, soff);
		if (lane_id() != 0) node = -1;
	}
	double dev = 0.0;
	int nan = 0;
	if (node >= 0) {
		lv -= q;
		lv /= ws;
		if (lv != lv) nan = 1;
		const R nv = (R) lv;
		pout[node] = nv;
		// relax(cycleSize): |phi(0) - phi(1)| in double (grid_3d.cpp:597-605)
		double ld = (double) nv;
		ld -= (double) old;
		dev = fabs(ld);
	}
	__shared__ double smax[4];
	__shared__ int snan[4];
	dev = wave_max(dev);
	const int anynan = __any(nan);
	const int w = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0) { smax[w] = dev; snan[w] = anynan; }
	__syncthreads();
	if (threadIdx.x == 0) {
		a.block_max[blockIdx.x] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
		const int f = snan[0] | snan[1] | snan[2] | snan[3];
		a.block_flag[blockIdx.x] = f;
		// a NaN in a cycle before the last of a relax(cycleSize) call would be
		// overwritten by the next cycle's flags: it also sticks in the control
		// block (rare path, so the atomic costs nothing in a normal sweep)
		if (f) atomicOr(&a.ctl->nan, 1);
	}
}

// One in-place cycle of relax(cycleSize) in one workgroup of 256 threads.
// Rows go in ascending node order in chunks of at most 256 rows and
// kInplaceCap staged entries.  For a row of node i in a chunk starting at
// node lo, the reference's ascending sum (grid_3d.cpp:1324-1366) reads
//   j < lo          : values already written in this cycle (earlier chunks),
//   lo <= j < i     : values of this chunk written before i (or Dirichlet),
//   j > i           : values not yet written in this cycle.
// Phase A (one thread per row) sums the j < lo part in order and stages the
// remaining entries: the product for a known value, the coupling and the
// chunk position for a value written earlier in the chunk.  Phase B (one
// thread) finishes the rows in order, adding the staged terms in entry
// order, so every sum is the reference's one-thread sum bit for bit.
// Phase C writes the chunk back.  Loads and stores of the swept buffer are
// agent-scope atomics so that later chunks see the earlier chunks' values.
template <typename R>
__global__ __launch_bounds__(256) void k_inplace(SweepArgs a, InplaceArgs P)
{
	constexpr int CH = 256;
	R* pw = (R*) (P.buf ? a.phi1 : a.phi0);
	const R* po = (const R*) (P.buf ? a.phi0 : a.phi1);
	const R* __restrict__ coup = (const R*) a.coup;
	__shared__ int s_node[CH];
	__shared__ int s_e0[CH + 1];
	__shared__ double s_prefix[CH], s_q[CH], s_ws[CH];
	__shared__ R s_val[CH];
	__shared__ double s_term[kInplaceCap];
	__shared__ short s_pos[kInplaceCap];
	__shared__ int s_cnt[CH];
	__shared__ int s_rows, s_nan, s_bad;
	__shared__ double s_max[4];
	const int t = threadIdx.x;
	if (a.ctl->done) return;  // uniform: converged earlier in this batch
	if (t == 0) { s_nan = 0; s_bad = 0; }
	double dmax = 0.0;
	for (int c0 = 0; c0 < P.count;) {
		const int p = c0 + t;
		int slot = -1, node = 0x7fffffff, deg = 0;
		long long base = 0;
		if (p < P.count) {
			slot = P.ord_slot[p];
			node = a.slot_node[slot];
			deg = a.slot_deg[slot];
			base = a.slice_off[slot >> 6] + (slot & 63);
		}
		s_node[t] = node;
		__syncthreads();
		const int lo = s_node[0];
		// the part before the chunk, in entry order
		double prefix = 0.0;
		int k = 0;
		for (; k < deg; k++) {
			const int j = a.col[base + 64ll * k];
			if (j >= lo) break;
			double adj = (double) __hip_atomic_load(pw + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			adj *= (double) coup[base + 64ll * k];
			prefix += adj;
		}
		const int cnt = deg - k;
		s_cnt[t] = cnt;
		__syncthreads();
		if (t == 0) {
			// leading rows whose staged entries fit
			int off = 0, r = 0;
			for (; r < CH && c0 + r < P.count; r++) {
				if (off + s_cnt[r] > kInplaceCap) break;
				s_e0[r] = off;
				off += s_cnt[r];
			}
			s_e0[r] = off;
			s_rows = r;
			if (r == 0) s_bad = 1;
		}
		__syncthreads();
		const int rows = s_rows;
		if (rows == 0) break;  // one row longer than the staging area (reported below)
		if (t < rows) {
			s_prefix[t] = prefix;
			s_q[t] = a.q[slot];
			s_ws[t] = a.ws[slot];
			int e = s_e0[t];
			for (int kk = k; kk < deg; kk++, e++) {
				const int j = a.col[base + 64ll * kk];
				const R c = coup[base + 64ll * kk];
				int pos = -1;
				if (j < node) {  // an earlier row of this chunk, or a Dirichlet node
					int l = 0, h = t;
					while (l < h) {
						const int m = (l + h) >> 1;
						if (s_node[m] < j) l = m + 1;
						else h = m;
					}
					if (l < t && s_node[l] == j) pos = l;
				}
				if (pos >= 0) {
					s_term[e] = (double) c;
				} else {
					double adj = (double) __hip_atomic_load(pw + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
					adj *= (double) c;
					s_term[e] = adj;
				}
				s_pos[e] = (short) pos;
			}
		}
		__syncthreads();
		if (t == 0) {
			for (int r = 0; r < rows; r++) {
				double lv = s_prefix[r];
				const int e1 = s_e0[r + 1];
				for (int e = s_e0[r]; e < e1; e++) {
					const int pos = s_pos[e];
					double term = s_term[e];
					if (pos >= 0) {
						double adj = (double) s_val[pos];
						adj *= term;
						term = adj;
					}
					lv += term;
				}
				lv -= s_q[r];
				lv /= s_ws[r];
				if (lv != lv) s_nan = 1;
				s_val[r] = (R) lv;
			}
		}
		__syncthreads();
		if (t < rows) {
			const R v = s_val[t];
			__hip_atomic_store(pw + node, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (P.last) {
				double ld = (double) v;
				ld -= (double) po[node];
				dmax = fmax(dmax, fabs(ld));
			}
		}
		__threadfence_block();
		__syncthreads();
		c0 += rows;
	}
	dmax = wave_max(dmax);
	if ((t & 63) == 0) s_max[t >> 6] = dmax;
	__syncthreads();
	if (t == 0) {
		if (s_bad) atomicOr(&a.ctl->nan, 2);
		if (s_nan) atomicOr(&a.ctl->nan, 1);
		if (P.last) {
			a.block_max[0] = fmax(fmax(s_max[0], s_max[1]), fmax(s_max[2], s_max[3]));
			a.block_flag[0] = s_nan;
		}
	}
}

// the decision of one relax(1) call: the deviation over all blocks, the
// deviation queue / zero-slope test and the threshold (grid_3d.cpp:611-671)
__global__ __launch_bounds__(256) void k_finalize(int nblocks, RelaxCtl* ctl, const double* block_max, const int* block_flag)
{
	// the partials are loaded four per thread per round (clamped indices,
	// issued together) before the control block is consulted
	double m = 0.0;
	int nan = 0;
	for (int b0 = 0; b0 < nblocks; b0 += 1024) {
		double v[4];
		int f[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int b = min(b0 + u * 256 + (int) threadIdx.x, nblocks - 1);
			v[u] = block_max[b];
			f[u] = block_flag[b];
		}
#pragma unroll
		for (int u = 0; u < 4; u++) {
			m = fmax(m, v[u]);
			nan |= f[u];
		}
	}
	if (ctl->done) return;
	__shared__ double smax[4];
	__shared__ int snan[4];
	m = wave_max(m);
	const int anynan = __any(nan);
	if ((threadIdx.x & 63) == 0) { smax[threadIdx.x >> 6] = m; snan[threadIdx.x >> 6] = anynan; }
	__syncthreads();
	if (threadIdx.x != 0) return;
	const double d = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
	ctl->sweeps += 1;
	ctl->last_dev = d;
	if (ctl->nan & 2) { ctl->done = 4; return; }  // a row too long for the in-place staging
	if ((snan[0] | snan[1] | snan[2] | snan[3]) || ctl->nan) { ctl->done = 3; return; }
	double slope = 0.0;
	if (push_and_test(ctl->queue, &ctl->qhead, &ctl->qsize, 1024, ctl->queue_size, ctl->cycle, d, &slope)) {
		ctl->fault_dev = d;
		ctl->fault_slope = slope;
		ctl->done = 2;
		return;
	}
	ctl->iteration += ctl->cycle;
	if (!(d > ctl->threshold)) ctl->done = 1;
}

// Whole localRelax(index, order, threshold, cycleSize, queueSize) in one
// workgroup.  doLocalRelax (grid_3d.cpp:1260-1282, 1516-1575) reads buffer
// _phiSwitch and writes the buffer after _phiLocalSwitch, which toggles per
// cycle: a cycle whose write buffer differs from the read buffer is Jacobi,
// otherwise it updates in place in ascending node order (one reference
// thread).  The in-place order is reproduced exactly with a level schedule:
// a node's level exceeds the levels of all its lower-index shell neighbours,
// so every node reads exactly the values the sequential sweep would read.
template <typename R>
__global__ __launch_bounds__(1024) void k_local(SweepArgs a, LocalArgs L)
{
	R* phi[2] = {(R*) a.phi0, (R*) a.phi1};
	const R* pr = phi[L.read_buf];
	__shared__ double sred[16];
	__shared__ int sflag;
	__shared__ double squeue[1024];
	__shared__ int qhead, qsize;
	__shared__ long long iteration;
	__shared__ int ps, status;
	if (threadIdx.x == 0) {
		qhead = 0; qsize = 0; iteration = 0; ps = L.local_switch; status = 0;
	}
	__syncthreads();
	for (;;) {
		for (int c = 0; c < L.cycle; c++) {
			const int wbuf = (ps + 1) % 2;
			R* pw = phi[wbuf];
			if (wbuf != L.read_buf) {
				for (int p = threadIdx.x; p < L.count; p += blockDim.x) {
					const double lv = row_value<R, true>(a, L.slots[p], pr);
					if (lv != lv) atomicOr(&status, 3);
					__hip_atomic_store(pw + L.nodes[p], (R) lv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				}
				__syncthreads();
			} else {
				for (int l = 0; l < L.nlevels; l++) {
					for (int q = L.level_ptr[l] + threadIdx.x; q < L.level_ptr[l + 1]; q += blockDim.x) {
						const int p = L.order[q];
						const double lv = row_value<R, true>(a, L.slots[p], pr);
						if (lv != lv) atomicOr(&status, 3);
						__hip_atomic_store(pw + L.nodes[p], (R) lv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
					}
					__threadfence_block();
					__syncthreads();
				}
			}
			if (threadIdx.x == 0) ps = (ps + 1) % 2;
			__syncthreads();
		}
		// deviation over the shell (grid_3d.cpp:748-757)
		double dev = 0.0;
		for (int p = threadIdx.x; p < L.count; p += blockDim.x) {
			const int node = L.nodes[p];
			double ld = (double) __hip_atomic_load(phi[0] + node, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			ld -= (double) __hip_atomic_load(phi[1] + node, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			ld = fabs(ld);
			if (ld > dev) dev = ld;
		}
		dev = wave_max(dev);
		if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = dev;
		__syncthreads();
		if (threadIdx.x == 0) {
			double m = 0.0;
			for (int w = 0; w < (int) (blockDim.x >> 6); w++) m = fmax(m, sred[w]);
			double slope;
			int stop = 0;
			if (status) stop = 1;
			else if (push_and_test(squeue, &qhead, &qsize, 1024, L.queue, L.cycle, m, &slope)) {
				iteration = -1 - iteration;
				stop = 1;
			} else {
				iteration += L.cycle;
				if (!(m > L.threshold)) stop = 1;
				else if (iteration >= L.max_iterations) { status = 4; stop = 1; }
			}
			sflag = stop;
		}
		__syncthreads();
		if (sflag) break;
	}
	if (threadIdx.x == 0) {
		L.out[0] = iteration;
		L.out[1] = ps;
		L.out[2] = status;
	}
}

__global__ void k_fill_i32(int n, int* dst, int value)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) dst[i] = value;
}

// row y < count: job y's scatter (the values moved as bits); row count: the
// node's species values
__global__ __launch_bounds__(256) void k_patch_batch(PatchBatch b)
{
	const int j = blockIdx.y;
	if (j == b.count) {
		if (blockIdx.x == 0 && threadIdx.x == 0) {
			const NodeSpecies& v = b.node;
			v.d_flags[v.node] = v.flags;
			v.d_evap_field[v.node] = v.evap_field;
			v.d_evap_energy[v.node] = v.evap_energy;
			v.d_epsilon[v.node] = v.epsilon;
		}
		return;
	}
	const PatchJob p = b.job[j];
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += gridDim.x * blockDim.x) {
		const long long a = p.addr[i];
		if (p.width == 8) ((unsigned long long*) p.dst)[a] = ((const unsigned long long*) p.val)[i];
		else ((unsigned int*) p.dst)[a] = ((const unsigned int*) p.val)[i];
	}
}

}  // namespace dev

// long-row workgroups (one wavefront per long row) then slot workgroups,
// each part padded to a multiple of the XCD count (see row_block)
int row_launch_blocks(int nslots, int nlong)
{
	const int lg = ((nlong + 3) / 4 + 7) / 8 * 8;
	const int mg = ((nslots + 255) / 256 + 7) / 8 * 8;
	return lg + mg;
}

int launch_jacobi_sweep(hipStream_t s, const SweepArgs& a, int read, bool f64)
{
	const int blocks = row_launch_blocks(a.nslots, a.nlong);
	if (blocks == 0) return 0;
	if (f64) hipLaunchKernelGGL(dev::k_jacobi<double>, dim3(blocks), dim3(256), 0, s, a, read);
	else hipLaunchKernelGGL(dev::k_jacobi<float>, dim3(blocks), dim3(256), 0, s, a, read);
	hipLaunchKernelGGL(dev::k_finalize, dim3(1), dim3(256), 0, s, blocks, a.ctl, (const double*) a.block_max,
	                   (const int*) a.block_flag);
	return (int) hipGetLastError();
}

int launch_jacobi_cycle(hipStream_t s, const SweepArgs& a, int read, bool f64, bool decide)
{
	if (decide) return launch_jacobi_sweep(s, a, read, f64);
	const int blocks = row_launch_blocks(a.nslots, a.nlong);
	if (blocks == 0) return 0;
	if (f64) hipLaunchKernelGGL(dev::k_jacobi<double>, dim3(blocks), dim3(256), 0, s, a, read);
	else hipLaunchKernelGGL(dev::k_jacobi<float>, dim3(blocks), dim3(256), 0, s, a, read);
	return (int) hipGetLastError();
}

int launch_inplace_cycle(hipStream_t s, const SweepArgs& a, const InplaceArgs& P, bool f64)
{
	if (f64) hipLaunchKernelGGL(dev::k_inplace<double>, dim3(1), dim3(256), 0, s, a, P);
	else hipLaunchKernelGGL(dev::k_inplace<float>, dim3(1), dim3(256), 0, s, a, P);
	if (P.last)
		hipLaunchKernelGGL(dev::k_finalize, dim3(1), dim3(256), 0, s, 1, a.ctl, (const double*) a.block_max,
		                   (const int*) a.block_flag);
	return (int) hipGetLastError();
}

int launch_local_relax(hipStream_t s, const SweepArgs& a, const LocalArgs& L, bool f64)
{
	if (f64) hipLaunchKernelGGL(dev::k_local<double>, dim3(1), dim3(1024), 0, s, a, L);
	else hipLaunchKernelGGL(dev::k_local<float>, dim3(1), dim3(1024), 0, s, a, L);
	return (int) hipGetLastError();
}

int launch_fill_i32(hipStream_t s, int n, int* dst, int value)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, n, dst, value);
	return (int) hipGetLastError();
}
int launch_patch_batch(hipStream_t s, const PatchBatch& b)
{
	if (b.count < 0 || b.count > kPatchJobs) return (int) hipErrorInvalidValue;
	int blocks = 1;
	for (int j = 0; j < b.count; j++) {
		if (b.job[j].n < 0 || (b.job[j].width != 4 && b.job[j].width != 8)) return (int) hipErrorInvalidValue;
		blocks = std::max(blocks, std::min(1024, (b.job[j].n + 255) / 256));
	}
	hipLaunchKernelGGL(dev::k_patch_batch, dim3(blocks, b.count + 1), dim3(256), 0, s, b);
	return (int) hipGetLastError();
}

}  // namespace tsg
