// Surface evaluation: node fields by least squares (Grid_3d::Table::field_o1,
// grid_3d.cpp:909-991), the evaporation score |E| / F_evap of every surface
// site (Surface_3d::evap_compute_specificFields / probLinear_field,
// surface_3d.cpp:194-219, 266-297), the first-maximum choice
// (evapMaximum, surface_3d.cpp:423-431) with the apex, the sequential
// normaliser, the surface normal, and the device copy of the surface set.
#include "device_math.hpp"
#include "kernels.hpp"
#include "field.hpp"
#include "plane.hpp"

namespace tsg {
namespace dev {

__device__ __forceinline__ float exp_r(float x) { return expf(x); }
__device__ __forceinline__ double exp_r(double x) { return exp(x); }

// probVoronoiFluxForce (surface_3d.cpp:337-400): Maxwell stress over the
// node's Voronoi faces between media of different permittivity
template <typename R>
__device__ R flux_force(const FieldArgs& f, int i)
{
	const R* coup = (const R*) f.coup;
	const R* phi = (const R*) f.phi;
	const R eps0 = (R) 8.85418781762e-12;
	R area = 0;
	V3<R> force = v3<R>(0, 0, 0);
	const R pi = phi[i];
	const V3<R> ci = load_coord<R>(f.xyz, i);
	for (int kk = f.ptr[i]; kk < f.ptr[i + 1]; kk++) {
		const int j = f.idx[kk];
		R v = phi[j];
		v -= pi;
		v *= v;
		v = (R) (v * (-1.0 * eps0));
		v *= coup[kk];
		const R e1 = (R) (1.0 / f.epsilon[i]);
		const R e2 = (R) (1.0 / f.epsilon[j]);
		v *= e2 - e1;
		const V3<R> cj = load_coord<R>(f.xyz, j);
		R corr = coup[kk];
		corr *= vlen(vsub(ci, cj));
		corr = (R) (corr / 2.0);
		corr *= e1 + e2;
		area += corr;
		V3<R> lf = vsub(cj, ci);
		const R dd = vdot(lf, lf);
		for (int c = 0; c < 3; c++) lf.x[c] /= dd;
		for (int c = 0; c < 3; c++) lf.x[c] *= v;
		for (int c = 0; c < 3; c++) force.x[c] += lf.x[c];
	}
	R value = vlen(force);
	value /= area;
	return value;
}

template <typename R>
__global__ __launch_bounds__(256) void k_surface_fields(FieldArgs f)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= f.count) return;
	const int node = f.nodes[s];
	R E[3];
	if (!field_o1<R>(node, f.xyz, f.ptr, f.idx, (const R*) f.coup, f.flags, (const R*) f.phi, E)) {
		atomicOr(f.status, 1);
		E[0] = E[1] = E[2] = (R) 0;
	}
	if (E[0] != E[0] || E[1] != E[1] || E[2] != E[2]) atomicOr(f.status, 2);
	const R flen = vlen(v3<R>(E[0], E[1], E[2]));
	R ef = flen;
	ef /= (R) f.evap_field[node];
	R value = ef;
	if (f.mode == 1) {  // probBoltzmann (surface_3d.cpp:223-262)
		const R kB = (R) 8.6173431e-5f;
		R pr = (R) -1.0;
		pr *= flen;
		pr /= (R) f.evap_field[node];
		pr = (R) (pr + 1.0);
		pr *= (R) f.evap_energy[node];
		pr /= (R) f.temperature;
		pr /= kB;
		pr = (R) (pr / -2.0);
		value = exp_r(pr);
	} else if (f.mode == 2) {  // probLinear_force (surface_3d.cpp:300-333)
		R e2[3], q;
		const int rc = field_o2<R>(node, f.xyz, f.ptr, f.idx, f.flags, (const R*) f.phi, e2, &q);
		if (rc) atomicOr(f.status, rc == 1 ? 4 : 2);
		const R fs = vlen(v3<R>(e2[0], e2[1], e2[2]));
		const R F = (R) f.evap_field[node];
		value = q;
		value *= fs;
		value /= F * F;
	} else if (f.mode == 3) {
		const R F = (R) f.evap_field[node];
		value = flux_force<R>(f, node);
		value /= F * F;
	}
	((R*) f.value)[s] = value;
	if (f.value_ef) ((R*) f.value_ef)[s] = ef;
}

template <typename R>
__global__ void k_node_fields(int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                              const R* coup, const unsigned char* flags, const R* phi, double* out, int* status)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= count) return;
	R E[3];
	if (!field_o1<R>(nodes[s], xyz, ptr, idx, coup, flags, phi, E)) atomicOr(status, 1);
	for (int i = 0; i < 3; i++) out[3 * (size_t) s + i] = (double) E[i];
}

// field_o2 (quadratic least squares) of a list of nodes: -grad phi and the
// charge -eps0 laplacian; status bit 2 for a singular system, bit 1 for NaN
template <typename R>
__global__ void k_node_fields_o2(int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                                 const unsigned char* flags, const R* phi, double* out, double* charge, int* status)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= count) return;
	R E[3], q;
	const int rc = field_o2<R>(nodes[s], xyz, ptr, idx, flags, phi, E, &q);
	if (rc == 1) atomicOr(status, 4);
	if (rc == 2) atomicOr(status, 2);
	for (int i = 0; i < 3; i++) out[3 * (size_t) s + i] = (double) E[i];
	charge[s] = (double) q;
}

// ---- the device surface set

__global__ __launch_bounds__(256) void k_sstate_scatter(int count, const int* __restrict__ nodes,
                                                        const unsigned char* __restrict__ vals,
                                                        unsigned char* __restrict__ sstate)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < count) sstate[nodes[i]] = vals[i];
}

// the compaction: 256 threads x 16 nodes per block
constexpr int kCompactPerThread = 16;
constexpr int kCompactBlock = 256 * kCompactPerThread;

// surface sites among the thread's 16 nodes
__device__ __forceinline__ int sites_of_thread(int n, const unsigned char* __restrict__ sstate, int first)
{
	int c = 0;
	if (first + kCompactPerThread <= n) {
		const uint4 w = *(const uint4*) (sstate + first);
		c = __popc(w.x & 0x01010101u) + __popc(w.y & 0x01010101u) + __popc(w.z & 0x01010101u) +
		    __popc(w.w & 0x01010101u);
	} else {
		for (int v = first; v < n; v++) c += sstate[v] & kSurfSite;
	}
	return c;
}

// exclusive prefix of c over the 256 threads of the block, and the block total
__device__ __forceinline__ int block_exclusive_scan(int c, int* total)
{
	__shared__ int wsum[4];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int inc = c;
	for (int off = 1; off < 64; off <<= 1) {
		const int t = __shfl_up(inc, off, 64);
		if (lane >= off) inc += t;
	}
	if (lane == 63) wsum[wave] = inc;
	__syncthreads();
	int before = 0;
	for (int w = 0; w < wave; w++) before += wsum[w];
	*total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
	__syncthreads();
	return before + inc - c;
}

__global__ __launch_bounds__(256) void k_surf_count(int n, const unsigned char* __restrict__ sstate, int* __restrict__ blk)
{
	const int first = blockIdx.x * kCompactBlock + threadIdx.x * kCompactPerThread;
	const int c = first < n ? sites_of_thread(n, sstate, first) : 0;
	int total;
	block_exclusive_scan(c, &total);
	if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_surf_write(int n, const unsigned char* __restrict__ sstate,
                                                    const int* __restrict__ blk, int* __restrict__ list,
                                                    int* __restrict__ count)
{
	// sites in the blocks before this one
	int pre = 0;
	for (int b = threadIdx.x; b < (int) blockIdx.x; b += blockDim.x) pre += blk[b];
	int ptotal;
	block_exclusive_scan(pre, &ptotal);
	const int first = blockIdx.x * kCompactBlock + threadIdx.x * kCompactPerThread;
	const int c = first < n ? sites_of_thread(n, sstate, first) : 0;
	int total;
	int pos = ptotal + block_exclusive_scan(c, &total);
	if (c)
		for (int v = first; v < first + kCompactPerThread && v < n; v++)
			if (sstate[v] & kSurfSite) list[pos++] = v;
	if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *count = ptotal + total;
}

// ---- the choice

// orderable key of a coordinate (-0 counts as +0, as the comparison of the
// reference's apex scan does) and the slot, larger for a smaller slot
__device__ __forceinline__ unsigned long long apex_key(float z, int s)
{
	const unsigned int u = __float_as_uint(z + 0.0f);
	const unsigned int k = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
	return ((unsigned long long) k << 32) | (unsigned long long) (0xffffffffu - (unsigned int) s);
}

// evapMaximum (surface_3d.cpp:423-431), the apex (surface_3d.cpp:171-190),
// the scaling reference (surface_3d.cpp:194-219) and the chosen site's
// potential and field in one pass over the site list.  The first maximum in
// slot order (slots ascend with the node index) decides; a slot below the
// winner whose value lies within a few ulps of the maximum could tie with it
// after the reference's float normalisation (v / sum), and a negative or NaN
// value breaks the order of v / sum: both set the near-tie flag for the host
// to resolve with the exact normaliser.
template <typename R>
__global__ __launch_bounds__(1024) void k_choose(ChooseArgs a)
{
	__shared__ R sv[1024], se[1024];
	__shared__ int ss[1024];
	__shared__ unsigned long long sa[1024];
	__shared__ int sneg, snear, sstat;
	const R* v = (const R*) a.value;
	const R* ve = a.value_ef ? (const R*) a.value_ef : v;
	const R ninf = (R) -INFINITY;
	R best = ninf, emax = ninf;
	int bs = 0x7fffffff, neg = 0;
	unsigned long long ak = 0;
	for (int s = threadIdx.x; s < a.count; s += blockDim.x) {
		const R x = v[s];
		if (x > best) { best = x; bs = s; }  // increasing s: ties keep the first
		if (!(x >= (R) 0)) neg = 1;
		const R e = ve[s];
		if (e > emax) emax = e;
		const unsigned long long k = apex_key(a.xyz[3 * (size_t) a.list[s] + 2], s);
		if (k > ak) ak = k;
	}
	if (threadIdx.x == 0) { sneg = 0; snear = 0; sstat = 0; }
	sv[threadIdx.x] = best;
	ss[threadIdx.x] = bs;
	se[threadIdx.x] = emax;
	sa[threadIdx.x] = ak;
	__syncthreads();
	if (neg) sneg = 1;
	for (int h = blockDim.x / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h) {
			const R o = sv[threadIdx.x + h];
			const int os = ss[threadIdx.x + h];
			if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && os < ss[threadIdx.x])) {
				sv[threadIdx.x] = o;
				ss[threadIdx.x] = os;
			}
			if (se[threadIdx.x + h] > se[threadIdx.x]) se[threadIdx.x] = se[threadIdx.x + h];
			if (sa[threadIdx.x + h] > sa[threadIdx.x]) sa[threadIdx.x] = sa[threadIdx.x + h];
		}
		__syncthreads();
	}
	const R vmax = sv[0];
	const int w = ss[0] == 0x7fffffff ? 0 : ss[0];
	const R eps = (sizeof(R) == 4) ? (R) (8.0 / 16777216.0) : (R) (8.0 / 9007199254740992.0);
	const R lim = vmax - vmax * eps;
	for (int s = threadIdx.x; s < w; s += blockDim.x)
		if (v[s] >= lim) snear = 1;
	const R* phi = (const R*) a.phi;
	ChoiceRec* rec = a.rec;
	if (threadIdx.x == 0 && a.count > 0) {
		const int node = a.list[w];
		R E[3];
		if (!field_o1<R>(node, a.xyz, a.ptr, a.idx, (const R*) a.coup, a.flags, phi, E)) {
			atomicOr(&sstat, 1);
			E[0] = E[1] = E[2] = (R) 0;
		}
		rec->node = node;
		rec->slot = w;
		rec->value = (double) v[w];
		rec->pot = (double) phi[node];
		for (int c = 0; c < 3; c++) rec->field[c] = (double) E[c];
	}
	if (threadIdx.x == 64 && a.prev >= 0) {
		R E[3];
		if (!field_o1<R>(a.prev, a.xyz, a.ptr, a.idx, (const R*) a.coup, a.flags, phi, E)) {
			atomicOr(&sstat, 16);
			E[0] = E[1] = E[2] = (R) 0;
		}
		rec->pot_prev = (double) phi[a.prev];
		for (int c = 0; c < 3; c++) rec->field_prev[c] = (double) E[c];
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		if (a.count == 0) {
			rec->node = -1;
			rec->slot = 0;
		}
		const int as = (int) (0xffffffffu - (unsigned int) (sa[0] & 0xffffffffu));
		const int an = a.count > 0 ? a.list[as] : -1;
		for (int c = 0; c < 3; c++) rec->apex[c] = an >= 0 ? a.xyz[3 * (size_t) an + c] : 0.0f;
		rec->scaling_ref = (double) se[0];
		rec->near_tie = (snear || sneg) ? 1 : 0;
		rec->count = *a.dev_count;
		rec->status = *a.status | sstat;
	}
}

// ---- the surface normal (Surface_3d::Table::normal, surface_3d.cpp:123-169)

constexpr int kNormalChunk = 1024;  // shell entries staged per pass of the sums
constexpr int kNormalExpandBlocks = 256;

__global__ __launch_bounds__(64) void k_nrm_init(NormalArgs a)
{
	if (threadIdx.x != 0) return;
	NormalCtl* c = a.ctl;
	c->cnt = 0;
	c->over = 0;
	c->lo = a.n;
	c->hi = -1;
	for (int l = 0; l < 4; l++) c->end[l] = 0;
	c->sorted = 0;
	c->skip = (a.sstate[a.node] & kSurfSite) ? 0 : 1;
	a.marks[a.node] = a.stamp;
}

// one level of adjacentNodes(node, set, 3) (geometry_3d.cpp:1112-1145): the
// neighbours of the previous level's entries (level 0: of the node) not
// reached yet are appended; one wavefront per frontier entry, its lanes over
// the neighbour list, one append per wavefront
__global__ __launch_bounds__(256) void k_nrm_expand(NormalArgs a, int level)
{
	const NormalCtl* c = a.ctl;
	if (c->skip || c->over) return;
	const int lb = level == 0 ? 0 : c->end[level - 1];
	const int le = level == 0 ? 1 : c->end[level];
	const int lane = threadIdx.x & 63;
	const int waves = gridDim.x * (blockDim.x >> 6);
	for (int f = lb + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); f < le; f += waves) {
		const int u = level == 0 ? a.node : a.shell[f];
		const int k0 = a.ptr[u], k1 = a.ptr[u + 1];
		for (int kb = k0; kb < k1; kb += 64) {
			const int k = kb + lane;
			int v = -1;
			bool claimed = false;
			if (k < k1) {
				v = a.idx[k];
				claimed = atomicExch(&a.marks[v], a.stamp) != a.stamp;
			}
			const unsigned long long m = __ballot(claimed);
			if (m == 0) continue;
			int base = 0;
			if (lane == 0) base = atomicAdd(&a.ctl->cnt, __popcll(m));
			base = __shfl(base, 0, 64);
			const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
			int vmin = claimed ? v : 0x7fffffff, vmax = claimed ? v : -1;
			for (int off = 32; off > 0; off >>= 1) {
				vmin = min(vmin, __shfl_xor(vmin, off, 64));
				vmax = max(vmax, __shfl_xor(vmax, off, 64));
			}
			if (claimed) {
				if (pos < kNormalShellCap) a.shell[pos] = v;
				else a.ctl->over = 1;
			}
			if (lane == 0) {
				atomicMin(&a.ctl->lo, vmin);
				atomicMax(&a.ctl->hi, vmax);
			}
		}
	}
}

__global__ __launch_bounds__(64) void k_nrm_close(NormalArgs a, int level)
{
	if (threadIdx.x == 0) a.ctl->end[level + 1] = min(a.ctl->cnt, kNormalShellCap);
}

// the shell's nodes among the thread's 16: marked by this call, not the node
__device__ __forceinline__ int shell_bits(const NormalArgs& a, int first, int lo, int hi)
{
	int bits = 0;
	for (int t = 0; t < kCompactPerThread; t++) {
		const int v = first + t;
		if (v >= lo && v <= hi && a.marks[v] == a.stamp && v != a.node) bits |= 1 << t;
	}
	return bits;
}

__global__ __launch_bounds__(256) void k_nrm_count(NormalArgs a)
{
	const NormalCtl* c = a.ctl;
	const int b0 = blockIdx.x * kCompactBlock;
	int n = 0;
	if (!c->skip && !c->over && b0 <= c->hi && b0 + kCompactBlock > c->lo)
		n = __popc(shell_bits(a, b0 + threadIdx.x * kCompactPerThread, c->lo, c->hi));
	int total;
	block_exclusive_scan(n, &total);
	if (threadIdx.x == 0) a.blk[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_nrm_write(NormalArgs a)
{
	const NormalCtl* c = a.ctl;
	int pre = 0;
	for (int b = threadIdx.x; b < (int) blockIdx.x; b += blockDim.x) pre += a.blk[b];
	int ptotal;
	block_exclusive_scan(pre, &ptotal);
	const int b0 = blockIdx.x * kCompactBlock;
	const int first = b0 + threadIdx.x * kCompactPerThread;
	int bits = 0;
	if (!c->skip && !c->over && b0 <= c->hi && b0 + kCompactBlock > c->lo) bits = shell_bits(a, first, c->lo, c->hi);
	int total;
	int pos = ptotal + block_exclusive_scan(__popc(bits), &total);
	for (int t = 0; t < kCompactPerThread; t++)
		if (bits & (1 << t)) a.sorted[pos++] = first + t;
	if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) a.ctl->sorted = ptotal + total;
}

// the sums in ascending shell order: threads 0-5 the moments (xx, xy, xz,
// yy, yz, zz) of the offsets x0 - x over the surface sites, threads 6-8 the
// orientation x - x0 over the vacuum nodes; then the plane on one thread
__global__ __launch_bounds__(256) void k_nrm_finish(NormalArgs a)
{
	__shared__ float cx[kNormalChunk], cy[kNormalChunk], cz[kNormalChunk];
	__shared__ unsigned char cst[kNormalChunk];
	__shared__ float s_acc[9];
	const NormalCtl* c = a.ctl;
	ChoiceRec* rec = a.rec;
	const int tid = threadIdx.x;
	if (c->skip || c->over) {
		if (tid == 0) {
			for (int k = 0; k < 3; k++) rec->normal[k] = 0.0f;
			rec->normal_status = c->over ? 2 : 0;
		}
		return;
	}
	const int cnt = c->sorted;
	const float x0[3] = {a.xyz[3 * (size_t) a.node], a.xyz[3 * (size_t) a.node + 1], a.xyz[3 * (size_t) a.node + 2]};
	const int P0 = tid < 3 ? 0 : (tid < 5 ? 1 : 2);
	const int Q0 = tid == 0 ? 0 : (tid == 1 || tid == 3 ? 1 : 2);
	float acc = 0.0f;
	for (int c0 = 0; c0 < cnt; c0 += kNormalChunk) {
		const int m = min(kNormalChunk, cnt - c0);
		for (int i = tid; i < m; i += blockDim.x) {
			const int v = a.sorted[c0 + i];
			cx[i] = a.xyz[3 * (size_t) v];
			cy[i] = a.xyz[3 * (size_t) v + 1];
			cz[i] = a.xyz[3 * (size_t) v + 2];
			cst[i] = a.sstate[v];
		}
		__syncthreads();
		if (tid < 6) {
			for (int i = 0; i < m; i++)
				if (cst[i] & kSurfSite) {
					const float x[3] = {cx[i], cy[i], cz[i]};
					const float op = x0[P0] - x[P0], oq = x0[Q0] - x[Q0];
					acc += op * oq;
				}
		} else if (tid < 9) {
			const int k = tid - 6;
			for (int i = 0; i < m; i++)
				if (cst[i] & kSurfVacuum) {
					const float x[3] = {cx[i], cy[i], cz[i]};
					acc += x[k] - x0[k];
				}
		}
		__syncthreads();
	}
	if (tid < 9) s_acc[tid] = acc;
	__syncthreads();
	if (tid == 0) {
		float mom[6], nv[3];
		for (int t = 0; t < 6; t++) mom[t] = s_acc[t];
		if (!plane::from_moments(mom, nv)) {
			rec->normal_status = 1;
			return;
		}
		float dot = 0.0f;
		for (int k = 0; k < 3; k++) dot += nv[k] * s_acc[6 + k];
		if (dot < 0.0f)
			for (int k = 0; k < 3; k++) nv[k] *= -1.0f;
		for (int k = 0; k < 3; k++) rec->normal[k] = nv[k];
		rec->normal_status = 0;
	}
}

// sequential sum in slot order, reproducing the reference's running float
// (or double) sum exactly: one thread, loads issued 32 at a time ahead of the
// ordered additions
template <typename R>
__global__ __launch_bounds__(64) void k_seq_sum(int count, const R* __restrict__ v, double* out)
{
	if (threadIdx.x != 0) return;
	R sum = (R) 0;
	int s = 0;
	for (; s + 32 <= count; s += 32) {
		R x[32];
#pragma unroll
		for (int u = 0; u < 32; u++) x[u] = v[s + u];
#pragma unroll
		for (int u = 0; u < 32; u++) sum += x[u];
	}
	for (; s < count; s++) sum += v[s];
	*out = (double) sum;
}

}  // namespace dev

int launch_surface_fields(hipStream_t st, const FieldArgs& f, bool f64)
{
	if (f.count <= 0) return 0;
	if (f64) hipLaunchKernelGGL(dev::k_surface_fields<double>, dim3((f.count + 255) / 256), dim3(256), 0, st, f);
	else hipLaunchKernelGGL(dev::k_surface_fields<float>, dim3((f.count + 255) / 256), dim3(256), 0, st, f);
	return (int) hipGetLastError();
}

int launch_sstate_scatter(hipStream_t st, int count, const int* nodes, const unsigned char* vals, unsigned char* sstate)
{
	if (count <= 0) return 0;
	hipLaunchKernelGGL(dev::k_sstate_scatter, dim3((count + 255) / 256), dim3(256), 0, st, count, nodes, vals, sstate);
	return (int) hipGetLastError();
}

int surface_compact_blocks(int n) { return (n + dev::kCompactBlock - 1) / dev::kCompactBlock; }

int launch_surface_compact(hipStream_t st, int n, const unsigned char* sstate, int* blk, int* list, int* count)
{
	const int nb = surface_compact_blocks(n);
	if (nb <= 0) return -1;
	hipLaunchKernelGGL(dev::k_surf_count, dim3(nb), dim3(256), 0, st, n, sstate, blk);
	hipLaunchKernelGGL(dev::k_surf_write, dim3(nb), dim3(256), 0, st, n, sstate, (const int*) blk, list, count);
	return (int) hipGetLastError();
}

int launch_surface_choose(hipStream_t st, const ChooseArgs& a, bool f64)
{
	if (f64) hipLaunchKernelGGL(dev::k_choose<double>, dim3(1), dim3(1024), 0, st, a);
	else hipLaunchKernelGGL(dev::k_choose<float>, dim3(1), dim3(1024), 0, st, a);
	return (int) hipGetLastError();
}

size_t surface_normal_workspace(int n)
{
	return sizeof(NormalCtl) + 64 + 2 * sizeof(int) * (size_t) kNormalShellCap + sizeof(int) * (size_t) surface_compact_blocks(n);
}

int launch_surface_normal(hipStream_t st, const NormalArgs& a)
{
	const int nb = surface_compact_blocks(a.n);
	if (nb <= 0) return -1;
	hipLaunchKernelGGL(dev::k_nrm_init, dim3(1), dim3(64), 0, st, a);
	for (int l = 0; l < 3; l++) {
		hipLaunchKernelGGL(dev::k_nrm_expand, dim3(dev::kNormalExpandBlocks), dim3(256), 0, st, a, l);
		hipLaunchKernelGGL(dev::k_nrm_close, dim3(1), dim3(64), 0, st, a, l);
	}
	hipLaunchKernelGGL(dev::k_nrm_count, dim3(nb), dim3(256), 0, st, a);
	hipLaunchKernelGGL(dev::k_nrm_write, dim3(nb), dim3(256), 0, st, a);
	hipLaunchKernelGGL(dev::k_nrm_finish, dim3(1), dim3(256), 0, st, a);
	return (int) hipGetLastError();
}

int launch_surface_sum(hipStream_t st, int count, const void* value, bool f64, double* out)
{
	if (f64) hipLaunchKernelGGL(dev::k_seq_sum<double>, dim3(1), dim3(64), 0, st, count, (const double*) value, out);
	else hipLaunchKernelGGL(dev::k_seq_sum<float>, dim3(1), dim3(64), 0, st, count, (const float*) value, out);
	return (int) hipGetLastError();
}

int launch_node_fields_o2(hipStream_t st, int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                          const unsigned char* flags, const void* phi, bool f64, double* out, double* charge, int* status)
{
	if (count <= 0) return 0;
	if (f64)
		hipLaunchKernelGGL(dev::k_node_fields_o2<double>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz,
		                   ptr, idx, flags, (const double*) phi, out, charge, status);
	else
		hipLaunchKernelGGL(dev::k_node_fields_o2<float>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz,
		                   ptr, idx, flags, (const float*) phi, out, charge, status);
	return (int) hipGetLastError();
}

int launch_node_fields(hipStream_t st, int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                       const void* coup, const unsigned char* flags, const void* phi, bool f64, double* out, int* status)
{
	if (count <= 0) return 0;
	if (f64)
		hipLaunchKernelGGL(dev::k_node_fields<double>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz, ptr,
		                   idx, (const double*) coup, flags, (const double*) phi, out, status);
	else
		hipLaunchKernelGGL(dev::k_node_fields<float>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz, ptr,
		                   idx, (const float*) coup, flags, (const float*) phi, out, status);
	return (int) hipGetLastError();
}

}  // namespace tsg
