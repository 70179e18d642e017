// Geometry precompute on the GPU: tetrahedron circumcentres and, for every
// directed Delaunay edge (i -> j), the Voronoi face area with the three
// closing factors the reference uses:
//   inf = 0.5 * |p_i - p_j|  -> Grid_3d::Table::computeCoupling (grid_3d.cpp:1392-1409)
//   inf = 1.0                -> Surface_3d::Table::init/update test (surface_3d.cpp:73,109)
//   inf = meanDist(i)/2      -> computeCharge's cell volume (grid_3d.cpp:1373-1390)
// The reference recomputes these with std::list walks on the host per call
// (hours at the largest specimens); here one thread owns one directed edge.
//
// Face construction restates Geometry_3d::Table::get_voronoiFace
// (geometry_3d.cpp:1211-1382) and VoronoiFace::finiteCenter/finiteArea
// (:244-360) including the start tet, the walk direction given by the
// permutation parity, and the float operation order.
#include <cstring>

#include "device_math.hpp"
#include "kernels.hpp"

namespace tsg {
namespace dev {

// Geometry_3d::Table::tetCenter, geometry_3d.cpp:830-870: double LU solve
// relative to vertex 3; the float result is (float)(x3 + rhs)
__global__ void k_circumcentres(int m, const float* __restrict__ xyz, const int* __restrict__ tv,
                                double* __restrict__ cc64, float* __restrict__ cc32, int* __restrict__ status)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= m) return;
	const int v[4] = {tv[4 * (size_t) t], tv[4 * (size_t) t + 1], tv[4 * (size_t) t + 2], tv[4 * (size_t) t + 3]};
	double lumat[9], rhs[3];
	for (int i = 0; i < 3; i++) {
		double tmp = 0.0;
		for (int j = 0; j < 3; j++) {
			double diff = xyz[3 * (size_t) v[i] + j];
			diff -= xyz[3 * (size_t) v[3] + j];
			tmp += diff * diff;
			lumat[i * 3 + j] = diff;
		}
		rhs[i] = 0.5 * tmp;
	}
	double d;
	int index[3];
	if (!lu_decmp<double, 3>(lumat, index, &d)) {
		atomicOr(status, 1);
		for (int k = 0; k < 3; k++) { cc64[3 * (size_t) t + k] = 0.0; cc32[3 * (size_t) t + k] = 0.0f; }
		return;
	}
	lu_solve<double, 3>(lumat, index, rhs);
	for (int k = 0; k < 3; k++) {
		const float x3 = xyz[3 * (size_t) v[3] + k];
		cc64[3 * (size_t) t + k] = (double) x3 + rhs[k];
		cc32[3 * (size_t) t + k] = (float) ((double) x3 + rhs[k]);
	}
}

// slot indices of n1, n2 in tet t, the remaining two slots ordered so that an
// even permutation swaps them (Geometry_3d::Table::sign, geometry_3d.cpp:1690)
__device__ __forceinline__ bool local_slots(const int* __restrict__ tv, int t, int n1, int n2, int li[4])
{
	const int* v = tv + 4 * (size_t) t;
	li[0] = 0;
	while (li[0] < 4 && v[li[0]] != n1) li[0]++;
	li[1] = 0;
	while (li[1] < 4 && v[li[1]] != n2) li[1]++;
	if (li[0] == 4 || li[1] == 4) return false;
	int j = 2;
	for (int i = 0; i < 4; i++)
		if (i != li[0] && i != li[1]) li[j++] = i;
	int inv = 0;
	for (int a = 0; a < 4; a++)
		for (int b = a + 1; b < 4; b++)
			if (li[a] > li[b]) inv++;
	if ((inv & 1) == 0) { const int x = li[2]; li[2] = li[3]; li[3] = x; }
	return true;
}

// The ring of tets around the edge (n1, n2), walked without storing it:
// `front` is its first tet, `len` the number of tets in ring order (a
// closed ring repeats its start at the end, as the reference's list does),
// and the ring is re-walked from `front` by ring_next.  The order is the
// reference's: from the start tet forward (the neighbour opposite slot
// li[2]) until the ring closes or opens; an open ring then extends backwards
// from the start (opposite li[3]) to its front.  Any number of tets: the
// fans of the boundary-plane nodes under a wide tip base exceed any fixed
// buffer.
struct Ring {
	int front;
	int len;
	bool closed;
	int front_open_vertex;  // vertex defining d1 (open rings)
	int back_open_vertex;   // vertex defining d2
	int n1, n2;
};

constexpr int kMaxRingWalk = 1 << 20;  // bounds a walk on a broken mesh

// the tet after t in ring order (-1 at an open end, -2 on a broken mesh)
__device__ __forceinline__ int ring_next(const int* __restrict__ tv, const int* __restrict__ tn, const Ring& r, int t)
{
	int li[4];
	if (!local_slots(tv, t, r.n1, r.n2, li)) return -2;
	return tn[4 * (size_t) t + li[2]];
}

// returns false on a broken mesh
__device__ bool build_ring(const int* __restrict__ tv, const int* __restrict__ tn, const int* __restrict__ assoc,
                           const int* __restrict__ nt_ptr, const int* __restrict__ nt_idx, int n1, int n2, Ring& r)
{
	int li[4];
	int start = assoc[n1];
	if (!local_slots(tv, start, n1, n2, li)) {
		start = -1;
		for (int k = nt_ptr[n1]; k < nt_ptr[n1 + 1]; k++)
			if (local_slots(tv, nt_idx[k], n1, n2, li)) { start = nt_idx[k]; break; }
		if (start < 0) return false;
	}
	r.n1 = n1;
	r.n2 = n2;
	r.front = start;
	r.len = 1;
	bool finite = true;
	{
		const int nb = tn[4 * (size_t) start + li[2]];
		if (nb != -1) {
			r.len++;
			int t = nb;
			while (t != start) {
				if (!local_slots(tv, t, n1, n2, li) || r.len > kMaxRingWalk) return false;
				const int nx = tn[4 * (size_t) t + li[2]];
				if (nx == -1) {
					r.back_open_vertex = tv[4 * (size_t) t + li[3]];
					finite = false;
					break;
				}
				r.len++;
				t = nx;
			}
		} else {
			r.back_open_vertex = tv[4 * (size_t) start + li[3]];
			finite = false;
		}
	}
	if (!finite) {
		// extend backwards from the start to the front
		int t = start;
		for (;;) {
			if (!local_slots(tv, t, n1, n2, li) || r.len > kMaxRingWalk) return false;
			const int nb = tn[4 * (size_t) t + li[3]];
			if (nb == -1) {
				r.front_open_vertex = tv[4 * (size_t) t + li[2]];
				r.front = t;
				break;
			}
			r.len++;
			t = nb;
		}
	}
	r.closed = finite && r.len > 1;
	return true;
}

template <typename R>
__device__ __forceinline__ V3<R> centre_of(const double* cc64, const float* cc32, int t)
{
	if constexpr (sizeof(R) == 4)
		return v3<R>(cc32[3 * (size_t) t], cc32[3 * (size_t) t + 1], cc32[3 * (size_t) t + 2]);
	else
		return v3<R>(cc64[3 * (size_t) t], cc64[3 * (size_t) t + 1], cc64[3 * (size_t) t + 2]);
}

// VoronoiFace::finiteArea(inf) on a ring (inf > 0 or a closed ring): the
// centre of the ring's circumcentres (with the two closing points of an
// open ring), then the fan of triangles about it, both in ring order
template <typename R>
__device__ R ring_area(const Ring& r, const float* __restrict__ xyz, const int* __restrict__ tv,
                       const int* __restrict__ tn, const double* cc64, const float* cc32, R inf)
{
	const V3<R> p1 = load_coord<R>(xyz, r.n1), p2 = load_coord<R>(xyz, r.n2);
	V3<R> d1 = v3<R>(0, 0, 0), d2 = v3<R>(0, 0, 0);
	if (!r.closed) {
		const V3<R> qb = load_coord<R>(xyz, r.back_open_vertex);
		d2 = vcross(vsub(qb, p1), vsub(p2, p1));
		if (vlen(d2) > (R) 0) { const R l = vlen(d2); for (int k = 0; k < 3; k++) d2.x[k] /= l; }
		const V3<R> qf = load_coord<R>(xyz, r.front_open_vertex);
		d1 = vcross(vsub(p2, p1), vsub(qf, p1));
		if (vlen(d1) > (R) 0) { const R l = vlen(d1); for (int k = 0; k < 3; k++) d1.x[k] /= l; }
	}
	V3<R> c = v3<R>(0, 0, 0);
	int t = r.front, last = r.front;
	for (int k = 0; k < r.len; k++) {
		c = vadd(c, centre_of<R>(cc64, cc32, t));
		last = t;
		if (k + 1 < r.len) t = ring_next(tv, tn, r, t);
	}
	int cnt = r.len;
	const V3<R> vf = centre_of<R>(cc64, cc32, r.front);
	const V3<R> vb = centre_of<R>(cc64, cc32, last);
	if (!r.closed) {
		c = vadd(c, vf);
		c = vadd(c, vscale(inf, d1));
		c = vadd(c, vb);
		c = vadd(c, vscale(inf, d2));
		cnt += 2;
	}
	for (int k = 0; k < 3; k++) c.x[k] /= (R) cnt;
	R value = (R) 0;
	V3<R> cur = vf;
	t = r.front;
	for (int k = 0; k + 1 < r.len; k++) {
		t = ring_next(tv, tn, r, t);
		const V3<R> nxt = centre_of<R>(cc64, cc32, t);
		value += vlen(vcross(vsub(cur, c), vsub(nxt, c))) / (R) 2.0;
		cur = nxt;
	}
	if (!r.closed) {
		V3<R> e1 = vf;
		e1 = vadd(e1, vscale(inf, d1));
		e1 = vsub(e1, c);
		V3<R> e2 = vsub(vf, c);
		value += vlen(vcross(e1, e2)) / (R) 2.0;
		e1 = vsub(vb, c);
		e2 = vb;
		e2 = vadd(e2, vscale(inf, d2));
		e2 = vsub(e2, c);
		value += vlen(vcross(e1, e2)) / (R) 2.0;
	}
	return value;
}

// One thread per directed edge k = (row i, column ptr[i]..ptr[i+1]).
// Outputs per directed edge: area with inf = |pi-pj|/2 in float and double,
// area with inf = 1 (float) for the surface test, |pi-pj| (float, double).
__global__ void k_edge_areas(int n, long long nnz, const float* __restrict__ xyz, const int* __restrict__ tv,
                             const int* __restrict__ tn, const int* __restrict__ assoc, const int* __restrict__ nt_ptr,
                             const int* __restrict__ nt_idx, const int* __restrict__ ptr, const int* __restrict__ idx,
                             const int* __restrict__ row_of, const double* __restrict__ cc64, const float* __restrict__ cc32,
                             float* __restrict__ a_half32, double* __restrict__ a_half64, float* __restrict__ a_one32,
                             float* __restrict__ dist32, double* __restrict__ dist64, int* __restrict__ status)
{
	const long long k = (long long) blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nnz) return;
	const int i = row_of[k];
	const int j = idx[k];
	Ring r;
	if (!build_ring(tv, tn, assoc, nt_ptr, nt_idx, i, j, r)) {
		atomicOr(status, 2);
		a_half32[k] = 0.0f; a_half64[k] = 0.0; a_one32[k] = 0.0f;
		return;
	}
	// Geometry_3d::Table::distance, geometry_3d.cpp:1203-1209
	const float d32 = vlen(vsub(load_coord<float>(xyz, i), load_coord<float>(xyz, j)));
	const double d64 = vlen(vsub(load_coord<double>(xyz, i), load_coord<double>(xyz, j)));
	dist32[k] = d32;
	dist64[k] = d64;
	// computeCoupling: distance = 0.5 * dist (double product stored to float)
	const float half32 = (float) (0.5 * (double) d32);
	const double half64 = 0.5 * d64;
	a_half32[k] = ring_area<float>(r, xyz, tv, tn, cc64, cc32, half32);
	a_half64[k] = ring_area<double>(r, xyz, tv, tn, cc64, cc32, half64);
	a_one32[k] = ring_area<float>(r, xyz, tv, tn, cc64, cc32, 1.0f);
}

// Geometry_3d::Table::voronoiVolume(i, inf) with computeCharge's closing
// factor (grid_3d.cpp:1373-1390): inf = half the mean distance to the
// neighbours; VoronoiCell::finiteVolume (geometry_3d.cpp:520-535) sums
// finiteArea(inf) * |pi - pj| / 6 over the faces in ascending neighbour order.
// One thread per node, float and double (the plain FP64 grid) variants.
__global__ void k_cell_volumes(int n, const float* __restrict__ xyz, const int* __restrict__ tv,
                               const int* __restrict__ tn, const int* __restrict__ assoc,
                               const int* __restrict__ nt_ptr, const int* __restrict__ nt_idx,
                               const int* __restrict__ ptr, const int* __restrict__ idx,
                               const double* __restrict__ cc64, const float* __restrict__ cc32,
                               const float* __restrict__ dist32, const double* __restrict__ dist64,
                               float* __restrict__ vol32, double* __restrict__ vol64, int* __restrict__ status)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int k0 = ptr[i], k1 = ptr[i + 1];
	if (k1 == k0) {
		vol32[i] = 0.0f;
		vol64[i] = 0.0;
		return;
	}
	float m32 = 0.0f;
	double m64 = 0.0;
	for (int k = k0; k < k1; k++) {
		m32 += dist32[k];
		m64 += dist64[k];
	}
	m32 /= (float) (k1 - k0);
	m32 /= 2.0f;
	m64 /= (double) (k1 - k0);
	m64 /= 2.0;
	float v32 = 0.0f;
	double v64 = 0.0;
	for (int k = k0; k < k1; k++) {
		Ring r;
		if (!build_ring(tv, tn, assoc, nt_ptr, nt_idx, i, idx[k], r)) {
			atomicOr(status, 2);
			vol32[i] = 0.0f;
			vol64[i] = 0.0;
			return;
		}
		float t32 = ring_area<float>(r, xyz, tv, tn, cc64, cc32, m32);
		t32 *= dist32[k];
		t32 /= 3.0f * 2.0f;
		v32 += t32;
		double t64 = ring_area<double>(r, xyz, tv, tn, cc64, cc32, m64);
		t64 *= dist64[k];
		t64 /= 3.0 * 2.0;
		v64 += t64;
	}
	vol32[i] = v32;
	vol64[i] = v64;
}

}  // namespace dev

int launch_cell_volumes(hipStream_t s, int n, const float* xyz, const int* tv, const int* tn, const int* assoc,
                        const int* nt_ptr, const int* nt_idx, const int* ptr, const int* idx, const double* cc64,
                        const float* cc32, const float* dist32, const double* dist64, float* vol32, double* vol64,
                        int* status)
{
	if (n <= 0) return 0;
	hipLaunchKernelGGL(dev::k_cell_volumes, dim3((n + 127) / 128), dim3(128), 0, s, n, xyz, tv, tn, assoc, nt_ptr, nt_idx,
	                   ptr, idx, cc64, cc32, dist32, dist64, vol32, vol64, status);
	return (int) hipGetLastError();
}

int launch_circumcentres(hipStream_t s, int m, const float* xyz, const int* tv, double* cc64, float* cc32, int* status)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(dev::k_circumcentres, dim3((m + 255) / 256), dim3(256), 0, s, m, xyz, tv, cc64, cc32, status);
	return (int) hipGetLastError();
}

namespace dev {

// order-preserving integer key of a float (no NaN): negative values have
// their magnitude bits flipped
__device__ __forceinline__ int key_of(float f)
{
	const int b = __float_as_int(f);
	return b >= 0 ? b : (b ^ 0x7fffffff);
}

__global__ void k_hull_extents(int m, const int* __restrict__ tn, const float* __restrict__ cc32, int* __restrict__ ext)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= m) return;
	const int* nb = tn + 4 * (size_t) t;
	if (nb[0] >= 0 && nb[1] >= 0 && nb[2] >= 0 && nb[3] >= 0) return;
	for (int k = 0; k < 3; k++) {
		const int key = key_of(cc32[3 * (size_t) t + k]);
		atomicMin(ext + k, key);
		atomicMax(ext + 3 + k, key);
	}
}

}  // namespace dev

float float_of_key(int k)
{
	const int b = k >= 0 ? k : (k ^ 0x7fffffff);
	float f;
	std::memcpy(&f, &b, sizeof(f));
	return f;
}

int launch_hull_extents(hipStream_t s, int m, const int* tn, const float* cc32, int* ext)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(dev::k_hull_extents, dim3((m + 255) / 256), dim3(256), 0, s, m, tn, cc32, ext);
	return (int) hipGetLastError();
}

int launch_edge_areas(hipStream_t s, int n, long long nnz, const float* xyz, const int* tv, const int* tn,
                      const int* assoc, const int* nt_ptr, const int* nt_idx, const int* ptr, const int* idx,
                      const int* row_of, const double* cc64, const float* cc32, float* a_half32, double* a_half64,
                      float* a_one32, float* dist32, double* dist64, int* status)
{
	if (nnz <= 0) return 0;
	const long long blocks = (nnz + 127) / 128;
	hipLaunchKernelGGL(dev::k_edge_areas, dim3((unsigned) blocks), dim3(128), 0, s, n, nnz, xyz, tv, tn, assoc, nt_ptr,
	                   nt_idx, ptr, idx, row_of, cc64, cc32, a_half32, a_half64, a_one32, dist32, dist64, status);
	return (int) hipGetLastError();
}

}  // namespace tsg
