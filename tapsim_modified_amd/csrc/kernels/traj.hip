// Ion trajectories through the tetrahedral mesh (Trajectory_3d,
// tapsim/trajectory_3d.cpp:99-696) with the point location and field
// interpolation they use (Geometry_3d::Table::findTetrahedron / tetWeights,
// geometry_3d.cpp:872-1044; Grid_3d::FastField::compute, grid_3d.cpp:1669-1721).
//
// One thread integrates one ion, in float as the reference does.  The node
// fields come from a snapshot (field_o1 of every node, taken after the
// relaxation), the counterpart of FastField's per-node cache.  Point
// location is the visibility walk with an exact orientation sign.  The
// reference draws each walk step's first face from rand(); when the caller
// hands over the process's rand() stream (one ion, as in the event loop) the
// walk draws from it in the same order, otherwise it starts at face 0, which
// changes which tet is returned only for a point exactly on a face.
#include "device_math.hpp"
#include "field.hpp"
#include "kernels.hpp"

namespace tsg {
namespace dev {

// ---------------------------------------------------------------- exact orientation sign

constexpr int kBn = 16;  // 1024-bit two's complement
struct Bn {
	unsigned long long w[kBn];
};

__device__ void bn_zero(Bn* r)
{
	for (int i = 0; i < kBn; i++) r->w[i] = 0ull;
}

__device__ void bn_neg(Bn* r, const Bn& a)
{
	unsigned long long carry = 1ull;
	for (int i = 0; i < kBn; i++) {
		const unsigned long long x = ~a.w[i] + carry;
		carry = (carry && x == 0ull) ? 1ull : 0ull;
		r->w[i] = x;
	}
}

__device__ void bn_add(Bn* r, const Bn& a, const Bn& b)
{
	unsigned long long carry = 0ull;
	for (int i = 0; i < kBn; i++) {
		const unsigned long long s1 = a.w[i] + b.w[i];
		const unsigned long long c1 = s1 < a.w[i] ? 1ull : 0ull;
		const unsigned long long s2 = s1 + carry;
		const unsigned long long c2 = s2 < s1 ? 1ull : 0ull;
		r->w[i] = s2;
		carry = c1 + c2;
	}
}

__device__ void bn_sub(Bn* r, const Bn& a, const Bn& b)
{
	Bn nb;
	bn_neg(&nb, b);
	bn_add(r, a, nb);
}

__device__ int bn_sign(const Bn& a)
{
	if (a.w[kBn - 1] >> 63) return -1;
	for (int i = 0; i < kBn; i++)
		if (a.w[i]) return 1;
	return 0;
}

__device__ void bn_mul(Bn* r, const Bn& a, const Bn& b)
{
	Bn x = a, y = b;
	const int sa = bn_sign(x), sb = bn_sign(y);
	if (sa < 0) bn_neg(&x, x);
	if (sb < 0) bn_neg(&y, y);
	unsigned long long out[kBn];
	for (int i = 0; i < kBn; i++) out[i] = 0ull;
	for (int i = 0; i < kBn; i++) {
		if (x.w[i] == 0ull) continue;
		unsigned long long carry = 0ull;
		for (int j = 0; i + j < kBn; j++) {
			const unsigned long long lo = x.w[i] * y.w[j];
			const unsigned long long hi = __umul64hi(x.w[i], y.w[j]);
			unsigned long long s = out[i + j] + lo;
			unsigned long long c = s < lo ? 1ull : 0ull;
			const unsigned long long s2 = s + carry;
			c += s2 < s ? 1ull : 0ull;
			out[i + j] = s2;
			carry = hi + c;
		}
	}
	for (int i = 0; i < kBn; i++) r->w[i] = out[i];
	if ((sa < 0) != (sb < 0)) bn_neg(r, *r);
}

// v = m * 2^(e-24) with a 24-bit integer m, scaled to 2^emin
__device__ void bn_from_float(Bn* r, float v, int emin)
{
	bn_zero(r);
	if (v == 0.0f) return;
	int e;
	const float fm = frexpf(v, &e);
	const long long m = (long long) ldexpf(fm, 24);
	const int shift = (e - 24) - emin;
	const unsigned long long um = (unsigned long long) (m < 0 ? -m : m);
	const int limb = shift / 64, bit = shift % 64;
	if (limb < kBn) r->w[limb] |= um << bit;
	if (bit != 0 && limb + 1 < kBn) r->w[limb + 1] |= um >> (64 - bit);
	if (m < 0) bn_neg(r, *r);
}

__device__ __noinline__ int orient_exact(const float* a, const float* b, const float* c, const float* d)
{
	const float* p[4] = {a, b, c, d};
	int emin = 1 << 30;
	for (int i = 0; i < 4; i++)
		for (int k = 0; k < 3; k++)
			if (p[i][k] != 0.0f) {
				int e;
				frexpf(p[i][k], &e);
				if (e - 24 < emin) emin = e - 24;
			}
	Bn m[3][3];
	for (int i = 0; i < 3; i++)
		for (int k = 0; k < 3; k++) {
			Bn x, y;
			bn_from_float(&x, p[i][k], emin);
			bn_from_float(&y, p[3][k], emin);
			bn_sub(&m[i][k], x, y);
		}
	Bn t1, t2, minor, acc, term;
	bn_zero(&acc);
	bn_mul(&t1, m[1][1], m[2][2]); bn_mul(&t2, m[1][2], m[2][1]); bn_sub(&minor, t1, t2);
	bn_mul(&term, m[0][0], minor); bn_add(&acc, acc, term);
	bn_mul(&t1, m[1][0], m[2][2]); bn_mul(&t2, m[1][2], m[2][0]); bn_sub(&minor, t1, t2);
	bn_mul(&term, m[0][1], minor); bn_sub(&acc, acc, term);
	bn_mul(&t1, m[1][0], m[2][1]); bn_mul(&t2, m[1][1], m[2][0]); bn_sub(&minor, t1, t2);
	bn_mul(&term, m[0][2], minor); bn_add(&acc, acc, term);
	return bn_sign(acc);
}

// sign of det[a-d; b-d; c-d] (the orient3d convention of the reference's
// predicates, used through Geometry_3d::Table::tetOrient): a double
// evaluation with a forward error bound, exact big integers otherwise
__device__ int orient_sign_f(const float* a, const float* b, const float* c, const float* d)
{
	const double adx = (double) a[0] - d[0], ady = (double) a[1] - d[1], adz = (double) a[2] - d[2];
	const double bdx = (double) b[0] - d[0], bdy = (double) b[1] - d[1], bdz = (double) b[2] - d[2];
	const double cdx = (double) c[0] - d[0], cdy = (double) c[1] - d[1], cdz = (double) c[2] - d[2];
	const double det = adx * (bdy * cdz - bdz * cdy) - ady * (bdx * cdz - bdz * cdx) + adz * (bdx * cdy - bdy * cdx);
	const double perm = fabs(adx) * (fabs(bdy * cdz) + fabs(bdz * cdy)) + fabs(ady) * (fabs(bdx * cdz) + fabs(bdz * cdx)) +
	                    fabs(adz) * (fabs(bdx * cdy) + fabs(bdy * cdx));
	const double bound = 1e-14 * perm;
	if (det > bound) return 1;
	if (det < -bound) return -1;
	if (perm == 0.0) return 0;
	return orient_exact(a, b, c, d);
}

// ---------------------------------------------------------------- point location and weights

struct Mesh3 {
	const float* xyz;
	const int* tv;
	const int* tn;
	int m;
	// one wavefront runs the (uniform) calculation together: the four face
	// tests of a walk step and the five determinants of the weights go to
	// lanes 0-4; every lane holds the same results
	bool wave = false;
	__device__ const float* P(int i) const { return xyz + 3 * (size_t) i; }
};

// the process's rand() stream continued on the device: glibc's 31-word
// additive table (host/prand.hpp), staged in LDS by the one thread using it
struct DevRand {
	int* s;
	int r;
	__device__ int next()
	{
		const int f = r + 3 < 31 ? r + 3 : r + 3 - 31;
		const unsigned v = (unsigned) s[f] + (unsigned) s[r];
		s[f] = (int) v;
		r = r + 1 < 31 ? r + 1 : 0;
		return (int) (v >> 1);
	}
};

__device__ void rand_stage(DevRand* d, int* lds, const int* src)
{
	for (int i = 0; i < 31; i++) lds[i] = src[i];
	d->s = lds;
	d->r = src[31];
}

__device__ void rand_unstage(const DevRand& d, int* dst)
{
	for (int i = 0; i < 31; i++) dst[i] = d.s[i];
	dst[31] = d.r;
}

// findTetrahedron walk from t (geometry_3d.cpp:988-1044); -1 outside the
// mesh, -2 when the walk does not end.  Each step starts at face
// rand() % 4 when a stream is given, else at face 0.
// the walk with the four face tests of a step on lanes 0-3 (face lo + lane):
// the first negative face in the order lo, lo+1, ... is the lowest set lane
// of the ballot, so the path and the draws are those of the one-lane walk.
// Lane q (0-3) loads vertex q of the tet and the neighbour across face q
// before the step's draw; each test gathers its three vertices by shuffles.
__device__ int find_tet_wave(const Mesh3& g, const float* pos, int t, DevRand* rng)
{
	const int lane = threadIdx.x & 63;
	const int limit = min(g.m, 1 << 20);
	for (int loop = 0; loop < limit; loop++) {
		if (t < 0) return -1;
		float px = 0.0f, py = 0.0f, pz = 0.0f;
		int across = -1;
		if (lane < 4) {
			const float* P = g.P(g.tv[4 * (size_t) t + lane]);
			px = P[0];
			py = P[1];
			pz = P[2];
			across = g.tn[4 * (size_t) t + lane];
		}
		const int lo = rng ? rng->next() % 4 : 0;
		// tetOrient(n1, n2, n3, pos) = orient3d(p2, p1, p3, pos) with
		// n1 = v[index % 4], n2 = v[3 - (2 + index) % 4], n3 = v[(2 + index) % 4]
		const int index = lo + (lane & 3);
		const int i1 = index % 4, i2 = 3 - (2 + index) % 4, i3 = (2 + index) % 4;
		const float p2[3] = {__shfl(px, i2, 64), __shfl(py, i2, 64), __shfl(pz, i2, 64)};
		const float p1[3] = {__shfl(px, i1, 64), __shfl(py, i1, 64), __shfl(pz, i1, 64)};
		const float p3[3] = {__shfl(px, i3, 64), __shfl(py, i3, 64), __shfl(pz, i3, 64)};
		bool neg = false;
		if (lane < 4) neg = orient_sign_f(p2, p1, p3, pos) < 0;
		const unsigned long long mask = __ballot(neg);
		if (mask == 0) return t;
		const int k = __ffsll((unsigned long long) mask) - 1;
		t = __shfl(across, 3 - (lo + k) % 4, 64);
	}
	return -2;
}

__device__ int find_tet(const Mesh3& g, const float* pos, int t, DevRand* rng = nullptr)
{
	if (g.wave) return find_tet_wave(g, pos, t, rng);
	const int limit = min(g.m, 1 << 20);
	for (int loop = 0; loop < limit; loop++) {
		if (t < 0) return -1;
		const int* v = g.tv + 4 * (size_t) t;
		const int lo = rng ? rng->next() % 4 : 0, hi = lo + 4;
		int index;
		for (index = lo; index < hi; index++) {
			// tetOrient(n1, n2, n3, pos) = orient3d(p2, p1, p3, pos)
			const int n1 = v[index % 4], n2 = v[3 - (2 + index) % 4], n3 = v[(2 + index) % 4];
			if (orient_sign_f(g.P(n2), g.P(n1), g.P(n3), pos) < 0) {
				t = g.tn[4 * (size_t) t + 3 - index % 4];
				break;
			}
		}
		if (index == hi) return t;
	}
	return -2;
}

// tapsim/utils.h:212-226 (float): determinant from the LU diagonal by
// mantissa products and an exponent sum
__device__ float lu_det_f(const float* lu, float d)
{
	float x = d;
	int y = 0;
	for (int i = 0; i < 3; i++) {
		int tmp;
		x *= frexpf(lu[i * 3 + i], &tmp);
		y += tmp;
	}
	return ldexpf(x, y);
}

__device__ float tet_volume(const Mesh3& g, int t)
{
	float lumat[9];
	const int* v = g.tv + 4 * (size_t) t;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) lumat[i * 3 + j] = g.P(v[i])[j] - g.P(v[3])[j];
	float d;
	int idx[3];
	if (!lu_decmp<float, 3>(lumat, idx, &d)) return 0.0f;
	float value = lu_det_f(lumat, d);
	value /= 6.0f;
	return fabsf(value);
}

// the weights with the determinant of vertex i's opposite sub-tet on lane i
// (0-3) and the volume on lane 4, gathered to every lane
__device__ void tet_weights_wave(const Mesh3& g, int t, const float* p, float* w)
{
	const int lane = threadIdx.x & 63;
	const int* v = g.tv + 4 * (size_t) t;
	float mine = 0.0f;
	bool ok = true;
	if (lane < 4) {
		const int i = lane;
		float lumat[9];
		for (int j = 0; j < 3; j++)
			for (int k = 0; k < 3; k++) lumat[j * 3 + k] = g.P(v[(i + j) % 4])[k] - p[k];
		float d;
		int idx[3];
		if (!lu_decmp<float, 3>(lumat, idx, &d)) {
			ok = false;
		} else {
			float x = lu_det_f(lumat, d);
			x /= 6.0f;
			mine = fabsf(x);
		}
	} else if (lane == 4) {
		mine = tet_volume(g, t);
	}
	const float volume = __shfl(mine, 4, 64);
	if (volume == 0.0f) {
		for (int i = 0; i < 4; i++) w[i] = 0.25f;
		return;
	}
	float wi = 0.0f;
	if (ok) {
		wi = mine;
		wi /= volume;
	}
	for (int i = 0; i < 4; i++) w[(i + 3) % 4] = __shfl(wi, i, 64);
}

__device__ void tet_weights(const Mesh3& g, int t, const float* p, float* w)
{
	if (g.wave) {
		tet_weights_wave(g, t, p, w);
		return;
	}
	const float volume = tet_volume(g, t);
	if (volume == 0.0f) {
		for (int i = 0; i < 4; i++) w[i] = 0.25f;
		return;
	}
	const int* v = g.tv + 4 * (size_t) t;
	for (int i = 0; i < 4; i++) {
		float lumat[9];
		for (int j = 0; j < 3; j++)
			for (int k = 0; k < 3; k++) lumat[j * 3 + k] = g.P(v[(i + j) % 4])[k] - p[k];
		float d;
		int idx[3];
		if (!lu_decmp<float, 3>(lumat, idx, &d)) {
			w[(i + 3) % 4] = 0.0f;
			continue;
		}
		float x = lu_det_f(lumat, d);
		x /= 6.0f;
		x = fabsf(x);
		x /= volume;
		w[(i + 3) % 4] = x;
	}
}

// FastField::compute: field at pos from the node snapshot (zero outside the mesh)
__device__ void fast_field(const Mesh3& g, const float* nfield, const float* pos, int guess, float* out,
                           DevRand* rng = nullptr)
{
	out[0] = out[1] = out[2] = 0.0f;
	const int t = find_tet(g, pos, guess, rng);
	if (t < 0) return;
	float w[4];
	tet_weights(g, t, pos, w);
	const int* v = g.tv + 4 * (size_t) t;
	for (int k = 0; k < 4; k++)
		for (int c = 0; c < 3; c++) out[c] += w[k] * nfield[3 * (size_t) v[k] + c];
}

// ---------------------------------------------------------------- the flight's tet cache

// The tets a one-wavefront flight's walks visit, kept in the wavefront's
// registers: slot s (0-15) lives on lanes 4s..4s+3, lane 4s+q holding vertex
// q of the slot's tet (its index, coordinates and field-snapshot value) and
// the neighbour across face q.  The walks of an integration step all start
// at the step's tet and mostly end in it or a neighbour, so most walk steps
// find their tet here instead of behind two dependent loads (tet -> vertices
// -> coordinates).  The cached values are the loaded ones, so the walks, the
// draws and the weights are those of the uncached walk.  Slots are refilled
// round-robin.
struct TetCache {
	int tet = -1;
	int across = -1;
	float x = 0.0f, y = 0.0f, z = 0.0f;
	float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
	float vol = -1.0f;  // the tet's volume once computed (-1 before)
	int next = 0;  // the slot the next miss fills (uniform)
};

// the first lane of t's slot, filled from memory on a miss (t >= 0)
__device__ int cache_slot(TetCache& c, const Mesh3& g, const float* nfield, int t)
{
	const unsigned long long hit = __ballot(c.tet == t);
	if (hit != 0) return __ffsll(hit) - 1;
	const int s = c.next;
	c.next = (s + 1) & 15;
	const int lane = threadIdx.x & 63;
	if ((lane >> 2) == s) {
		const int q = lane & 3;
		const int v = g.tv[4 * (size_t) t + q];
		c.tet = t;
		c.vol = -1.0f;
		c.across = g.tn[4 * (size_t) t + q];
		const float* P = g.P(v);
		const float* F = nfield + 3 * (size_t) v;
		c.x = P[0];
		c.y = P[1];
		c.z = P[2];
		c.f0 = F[0];
		c.f1 = F[1];
		c.f2 = F[2];
	}
	return 4 * s;
}

// find_tet_wave with the tets taken from the cache; *slot is the found
// tet's first lane
__device__ int find_tet_cached(TetCache& c, const Mesh3& g, const float* nfield, const float* pos, int t, DevRand* rng,
                               int* slot)
{
	const int lane = threadIdx.x & 63;
	const int limit = min(g.m, 1 << 20);
	for (int loop = 0; loop < limit; loop++) {
		if (t < 0) return -1;
		const int base = cache_slot(c, g, nfield, t);
		const int lo = rng ? rng->next() % 4 : 0;
		const int index = lo + (lane & 3);
		const int i1 = base + index % 4, i2 = base + 3 - (2 + index) % 4, i3 = base + (2 + index) % 4;
		const float p2[3] = {__shfl(c.x, i2, 64), __shfl(c.y, i2, 64), __shfl(c.z, i2, 64)};
		const float p1[3] = {__shfl(c.x, i1, 64), __shfl(c.y, i1, 64), __shfl(c.z, i1, 64)};
		const float p3[3] = {__shfl(c.x, i3, 64), __shfl(c.y, i3, 64), __shfl(c.z, i3, 64)};
		bool neg = false;
		if (lane < 4) neg = orient_sign_f(p2, p1, p3, pos) < 0;
		const unsigned long long mask = __ballot(neg);
		if (mask == 0) {
			*slot = base;
			return t;
		}
		const int k = __ffsll(mask) - 1;
		t = __shfl(c.across, base + 3 - (lo + k) % 4, 64);
	}
	return -2;
}

// tet_weights_wave from the cached slot: lane i (0-3) takes vertices i, i+1,
// i+2 (mod 4); the volume is computed once per cached tet (lane 4, from the
// four vertices) and kept in the slot
__device__ void tet_weights_cached(TetCache& c, int base, const float* p, float* w)
{
	const int lane = threadIdx.x & 63;
	float volume = __shfl(c.vol, base, 64);
	if (volume < 0.0f) {
		float X[4][3];
		for (int j = 0; j < 4; j++) {
			X[j][0] = __shfl(c.x, base + j, 64);
			X[j][1] = __shfl(c.y, base + j, 64);
			X[j][2] = __shfl(c.z, base + j, 64);
		}
		float v = 0.0f;
		if (lane == 4) {
			float lumat[9];
			for (int i = 0; i < 3; i++)
				for (int j = 0; j < 3; j++) lumat[i * 3 + j] = X[i][j] - X[3][j];
			float d;
			int idx[3];
			if (lu_decmp<float, 3>(lumat, idx, &d)) {
				float value = lu_det_f(lumat, d);
				value /= 6.0f;
				v = fabsf(value);
			}
		}
		volume = __shfl(v, 4, 64);
		if ((lane >> 2) == (base >> 2)) c.vol = volume;
	}
	if (volume == 0.0f) {
		for (int i = 0; i < 4; i++) w[i] = 0.25f;
		return;
	}
	float X[3][3];
	for (int j = 0; j < 3; j++) {
		const int src = base + ((lane & 3) + j) % 4;
		X[j][0] = __shfl(c.x, src, 64);
		X[j][1] = __shfl(c.y, src, 64);
		X[j][2] = __shfl(c.z, src, 64);
	}
	float wi = 0.0f;
	if (lane < 4) {
		float lumat[9];
		for (int j = 0; j < 3; j++)
			for (int k = 0; k < 3; k++) lumat[j * 3 + k] = X[j][k] - p[k];
		float d;
		int idx[3];
		if (lu_decmp<float, 3>(lumat, idx, &d)) {
			float x = lu_det_f(lumat, d);
			x /= 6.0f;
			wi = fabsf(x);
			wi /= volume;
		}
	}
	for (int i = 0; i < 4; i++) w[(i + 3) % 4] = __shfl(wi, i, 64);
}

// fast_field with the walk and the weights on the cache
__device__ void fast_field_cached(TetCache& c, const Mesh3& g, const float* nfield, const float* pos, int guess,
                                  float* out, DevRand* rng)
{
	out[0] = out[1] = out[2] = 0.0f;
	int base = 0;
	const int t = find_tet_cached(c, g, nfield, pos, guess, rng, &base);
	if (t < 0) return;
	float w[4];
	tet_weights_cached(c, base, pos, w);
	for (int k = 0; k < 4; k++) {
		const float f[3] = {__shfl(c.f0, base + k, 64), __shfl(c.f1, base + k, 64), __shfl(c.f2, base + k, 64)};
		for (int i = 0; i < 3; i++) out[i] += w[k] * f[i];
	}
}

// ---------------------------------------------------------------- integration

struct Phase {
	float t;
	float p[3], v[3];
	int tet;
};

struct Tr {
	Mesh3 g;
	const float* nfield;
	float charge, mass;
	int integrator, stepper;
	float dt_min;
	bool geom_error;       // the work limit ended the integration
	long long calls, max_calls;
	DevRand* rng;          // the process's rand() stream, or null
	float ep[3], ev[3];    // the cumulative error estimate (position, velocity)
	TetCache cache;        // the walks' tets (one-wavefront flights)
};

// the field at pos and the tet of pos, walked from guess
__device__ void tr_field(Tr& tr, const float* pos, int guess, float* out)
{
	if (tr.g.wave) fast_field_cached(tr.cache, tr.g, tr.nfield, pos, guess, out, tr.rng);
	else fast_field(tr.g, tr.nfield, pos, guess, out, tr.rng);
}

__device__ int tr_locate(Tr& tr, const float* pos, int guess)
{
	if (!tr.g.wave) return find_tet(tr.g, pos, guess, tr.rng);
	int base;
	return find_tet_cached(tr.cache, tr.g, tr.nfield, pos, guess, tr.rng, &base);
}

// Trajectory_3d::cumulative_error (trajectory_3d.cpp:680-696): per component
// sqrt(e^2 + step^2), the squares as the reference's powf(x, 2.0f)
__device__ __forceinline__ void cumulate_error(Tr& tr, const Phase& e)
{
	for (int i = 0; i < 3; i++) {
		float a = tr.ep[i] * tr.ep[i];
		float b = e.p[i] * e.p[i];
		a += b;
		tr.ep[i] = sqrtf(a);
		a = tr.ev[i] * tr.ev[i];
		b = e.v[i] * e.v[i];
		a += b;
		tr.ev[i] = sqrtf(a);
	}
}

// Trajectory_3d::o5_integrator (trajectory_3d.cpp:583-676): Cash-Karp
// embedded Runge-Kutta 4(5), tableau of Cash & Karp (1990)
__device__ Phase o5_step(Tr& tr, const Phase& back, float delta, Phase* error)
{
	const float b[6][5] = {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
	                       {1.0f / 5.0f, 0.0f, 0.0f, 0.0f, 0.0f},
	                       {3.0f / 40.0f, 9.0f / 40.0f, 0.0f, 0.0f, 0.0f},
	                       {3.0f / 10.0f, -9.0f / 10.0f, 6.0f / 5.0f, 0.0f, 0.0f},
	                       {-11.0f / 54.0f, 5.0f / 2.0f, -70.0f / 27.0f, 35.0f / 27.0f, 0.0f},
	                       {1631.0f / 55296.0f, 175.0f / 512.0f, 575.0f / 13824.0f, 44275.0f / 110592.0f, 253.0f / 4096.0f}};
	const float c1[6] = {37.0f / 378.0f, 0.0f, 250.0f / 621.0f, 125.0f / 594.0f, 0.0f, 512.0f / 1771.0f};
	const float c2[6] = {2825.0f / 27648.0f, 0.0f, 18575.0f / 48384.0f, 13525.0f / 55296.0f, 277.0f / 14336.0f, 1.0f / 4.0f};
	const float qm = tr.charge / tr.mass;
	float dv[6][3], dp[6][3];
	for (int i = 0; i < 6; i++) {
		float vel[3], pos[3];
		for (int l = 0; l < 3; l++) {
			vel[l] = back.v[l];
			pos[l] = back.p[l];
		}
		for (int j = 0; j < i; j++)
			for (int l = 0; l < 3; l++) {
				vel[l] += b[i][j] * dv[j][l];
				pos[l] += b[i][j] * dp[j][l];
			}
		tr_field(tr, pos, back.tet, dv[i]);
		for (int l = 0; l < 3; l++) dv[i][l] *= qm;
		for (int l = 0; l < 3; l++) dv[i][l] *= delta;
		for (int l = 0; l < 3; l++) dp[i][l] = vel[l] * delta;
	}
	Phase next = back;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 6; j++) {
			next.p[i] += c1[j] * dp[j][i];
			next.v[i] += c1[j] * dv[j][i];
		}
	next.t += delta;
	error->t = 0.0f;
	for (int i = 0; i < 3; i++) {
		error->p[i] = 0.0f;
		error->v[i] = 0.0f;
		for (int j = 0; j < 6; j++) {
			error->p[i] += (c1[j] - c2[j]) * dp[j][i];
			error->v[i] += (c1[j] - c2[j]) * dv[j][i];
		}
	}
	next.tet = tr_locate(tr, next.p, back.tet);
	return next;
}

// Trajectory_3d::o1_integrator (trajectory_3d.cpp:549-581)
__device__ Phase o1_step(Tr& tr, const Phase& back, float delta, Phase* error)
{
	float f[3];
	tr_field(tr, back.p, back.tet, f);
	for (int i = 0; i < 3; i++) f[i] *= tr.charge;
	for (int i = 0; i < 3; i++) f[i] /= tr.mass;
	Phase next = back;
	for (int i = 0; i < 3; i++) {
		next.v[i] += delta * f[i];
		next.p[i] += delta * back.v[i];
	}
	next.t += delta;
	error->t = 0.0f;
	const float d2 = delta * delta;
	for (int i = 0; i < 3; i++) {
		error->p[i] = next.p[i] * d2;
		error->v[i] = next.v[i] * d2;
	}
	next.tet = tr_locate(tr, next.p, back.tet);
	return next;
}

// The classical fourth-order Runge-Kutta step (integrator 2, an addition:
// the reference offers the Euler and Cash-Karp integrators).  Slopes at 0,
// h/2, h/2 and h -- from the start, the start plus half of the first slope,
// plus half of the second, plus the third -- and the solution
// y + (k1 + 2 k2 + 2 k3 + k4) / 6.  The error estimate is the difference to
// the embedded third-order solution y + (k1 + 2 k2 + 2 k3 + k5) / 6, where
// k5 is the slope at the new point: (k4 - k5) / 6.  Five field evaluations
// per attempt (Cash-Karp: six).
__device__ Phase o4_step(Tr& tr, const Phase& back, float delta, Phase* error)
{
	const float a[4] = {0.0f, 0.5f, 0.5f, 1.0f};
	const float c[4] = {1.0f / 6.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 6.0f};
	const float qm = tr.charge / tr.mass;
	float dv[5][3], dp[5][3];
	for (int i = 0; i < 4; i++) {
		float vel[3], pos[3];
		for (int l = 0; l < 3; l++) {
			vel[l] = back.v[l];
			pos[l] = back.p[l];
		}
		if (i > 0)
			for (int l = 0; l < 3; l++) {
				vel[l] += a[i] * dv[i - 1][l];
				pos[l] += a[i] * dp[i - 1][l];
			}
		tr_field(tr, pos, back.tet, dv[i]);
		for (int l = 0; l < 3; l++) dv[i][l] *= qm;
		for (int l = 0; l < 3; l++) dv[i][l] *= delta;
		for (int l = 0; l < 3; l++) dp[i][l] = vel[l] * delta;
	}
	Phase next = back;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 4; j++) {
			next.p[i] += c[j] * dp[j][i];
			next.v[i] += c[j] * dv[j][i];
		}
	next.t += delta;
	tr_field(tr, next.p, back.tet, dv[4]);
	for (int l = 0; l < 3; l++) dv[4][l] *= qm;
	for (int l = 0; l < 3; l++) dv[4][l] *= delta;
	for (int l = 0; l < 3; l++) dp[4][l] = next.v[l] * delta;
	error->t = 0.0f;
	for (int i = 0; i < 3; i++) {
		float e = dp[3][i];
		e -= dp[4][i];
		e /= 6.0f;
		error->p[i] = e;
		e = dv[3][i];
		e -= dv[4][i];
		e /= 6.0f;
		error->v[i] = e;
	}
	next.tet = tr_locate(tr, next.p, back.tet);
	return next;
}

// (tol/err)^(1/4) of the step-size control: the reference calls glibc's
// powf; the correctly rounded fourth root (two double square roots, one
// rounding to float) agrees with it on all but about one argument in 1500
__device__ __forceinline__ float fourth_root(float x) { return (float) sqrt(sqrt((double) x)); }

__device__ float rel_error(const Phase& value, const Phase& error)
{
	float rel = 0.0f;
	for (int i = 0; i < 3; i++) {
		float tmp;
		if (value.p[i] != 0.0f) {
			tmp = error.p[i];
			tmp /= value.p[i];
			tmp = fabsf(tmp);
			if (rel < tmp) rel = tmp;
		}
		if (value.v[i] != 0.0f) {
			tmp = error.v[i];
			tmp /= value.v[i];
			tmp = fabsf(tmp);
			if (rel < tmp) rel = tmp;
		}
	}
	return rel;
}

// a tet of adjacentTetrahedra(t) (geometry_3d.cpp:1046-1096): another tet
// sharing a vertex with t
__device__ bool adjacent(const Mesh3& g, int t, int u)
{
	if (u < 0 || u == t) return false;
	const int* a = g.tv + 4 * (size_t) t;
	const int* b = g.tv + 4 * (size_t) u;
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			if (a[i] == b[j]) return true;
	return false;
}

__device__ bool boundary_tet(const Mesh3& g, const unsigned char* gbnd, int t)
{
	for (int q = 0; q < 4; q++)
		if (gbnd[g.tv[4 * (size_t) t + q]]) return true;
	return false;
}

// the steppers (trajectory_3d.cpp:358-532): 0 none, 1 error-restricted,
// 2 geometry-restricted, 3 adaptive
__device__ Phase stepper_step(Tr& tr, const unsigned char* gbnd, const Phase& back, float* delta, float err_limit,
                              Phase* error)
{
	const float safety = 0.9f;
	for (;;) {
		if (++tr.calls > tr.max_calls) {  // bounds the work of one ion
			tr.geom_error = true;
			return back;
		}
		// the NONE stepper integrates with the step as given (trajectory_3d.cpp:155-158);
		// the other three raise it to the step limit first (:362,432,458)
		if (tr.stepper != 0 && *delta < tr.dt_min) *delta = tr.dt_min;
		const Phase value = tr.integrator == 0   ? o1_step(tr, back, *delta, error)
		                    : tr.integrator == 2 ? o4_step(tr, back, *delta, error)
		                                         : o5_step(tr, back, *delta, error);
		if (tr.stepper == 0) return value;
		if (tr.stepper == 2) {
			if (value.tet < 0 && boundary_tet(tr.g, gbnd, back.tet)) return value;
			if (adjacent(tr.g, back.tet, value.tet)) return value;
			if (value.tet == back.tet)
				*delta *= 1.1f;
			else {
				if (*delta == tr.dt_min) return value;
				*delta *= 0.9f;
			}
			continue;
		}
		if (tr.stepper == 3) {
			if (value.tet >= 0 && value.tet != back.tet && !adjacent(tr.g, back.tet, value.tet)) {
				*delta *= 0.1f;
				continue;
			}
		}
		const float rel = rel_error(value, *error);
		if (rel < err_limit || *delta == tr.dt_min) {
			if (rel != 0.0f) {
				float tmp = *delta;
				tmp *= fourth_root(err_limit / rel);
				tmp *= safety;
				if (tmp < *delta) *delta *= 1.1f;
				else *delta = tmp;
			}
			return value;
		}
		float tmp = *delta;
		tmp *= fourth_root(err_limit / rel);
		tmp *= safety;
		if (tmp > *delta) *delta *= 0.9f;
		else *delta = tmp;
	}
}

__device__ void put(const TrajArgs& a, int ion, int k, const Phase& q)
{
	if (a.summary) {
		float* sm = a.summary + 12 * (size_t) ion;
		if (k == 0)
			for (int c = 0; c < 3; c++) sm[c] = q.p[c];
		for (int c = 0; c < 3; c++) sm[3 + c] = q.p[c];
		sm[6] = q.t;
		if (q.tet >= 0) {
			for (int c = 0; c < 3; c++) sm[7 + c] = q.p[c];
			sm[10] = (float) q.tet;
			sm[11] = 1.0f;
		}
	}
	if (k >= a.cap) return;
	float* o = a.out + ((size_t) ion * a.cap + k) * 8;
	o[0] = q.t;
	for (int c = 0; c < 3; c++) {
		o[1 + c] = q.p[c];
		o[4 + c] = q.v[c];
	}
	o[7] = (float) q.tet;
}

// Trajectory_3d::init / integrate(delta, errorLimit) / extrapolate(Z, zmax)
// (trajectory_3d.cpp:99-126, 130-205, 303-332) for ion `ion` of the batch
__device__ void run_ion(const TrajArgs& a, int ion, DevRand* rng, bool wave = false)
{
	Tr tr;
	tr.g.wave = wave;
	tr.g.xyz = a.xyz;
	tr.g.tv = a.tv;
	tr.g.tn = a.tn;
	tr.g.m = a.m;
	tr.nfield = a.nfield + (size_t) a.snapshot[ion] * 3 * (size_t) a.n;
	tr.charge = a.charge[ion];
	tr.mass = a.mass[ion];
	tr.integrator = a.integrator;
	tr.stepper = a.stepper;
	tr.dt_min = a.dt_min;
	tr.geom_error = false;
	tr.calls = 0;
	tr.max_calls = a.max_calls;
	tr.rng = rng;
	for (int k = 0; k < 3; k++) tr.ep[k] = tr.ev[k] = 0.0f;
	const int node = a.nodes[ion];
	Phase back;
	back.t = 0.0f;
	for (int k = 0; k < 3; k++) {
		back.p[k] = a.xyz[3 * (size_t) node + k];
		back.v[k] = a.v0 ? a.v0[3 * (size_t) ion + k] : 0.0f;
	}
	back.tet = tr_locate(tr, back.p, a.assoc[node]);
	int status;
	int len = 0;
	if (a.summary)
		for (int c = 0; c < 12; c++) a.summary[12 * (size_t) ion + c] = 0.0f;
	if (a.err_out)
		for (int c = 0; c < 6; c++) a.err_out[6 * (size_t) ion + c] = 0.0f;
	if (back.tet < 0) {
		a.len[ion] = 0;
		a.status[ion] = -10;  // initial point outside the mesh
		return;
	}
	put(a, ion, len++, back);
	float f[3];
	tr_field(tr, back.p, back.tet, f);
	if (f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f)
		status = 1;  // NO_FIELD
	else {
		float delta = a.dt0;
		for (;;) {
			Phase error;
			const Phase next = stepper_step(tr, a.gbnd, back, &delta, a.err_limit, &error);
			if (tr.geom_error) break;
			back = next;
			put(a, ion, len++, back);
			cumulate_error(tr, error);
			if (back.tet < 0) {
				status = 5;  // SYSTEM_LIMIT: the ion left the mesh
				break;
			}
			if (len >= a.max_steps) {
				status = 3;  // ITERATION_LIMIT
				break;
			}
		}
	}
	if (tr.geom_error) {
		a.len[ion] = len;
		a.status[ion] = -4;  // work limit
		if (a.err_out)
			for (int c = 0; c < 3; c++) {
				a.err_out[6 * (size_t) ion + c] = tr.ep[c];
				a.err_out[6 * (size_t) ion + 3 + c] = tr.ev[c];
			}
		return;
	}
	// extrapolate(Z, zmax)
	{
		float dt = a.zmax;
		dt -= back.p[2];
		dt /= back.v[2];
		Phase item;
		item.t = back.t;
		item.t += dt;
		for (int k = 0; k < 3; k++) item.p[k] = back.p[k];
		for (int k = 0; k < 3; k++) item.p[k] += dt * back.v[k];
		for (int k = 0; k < 3; k++) item.v[k] = back.v[k];
		item.tet = tr_locate(tr, item.p, back.tet);
		put(a, ion, len++, item);
		// the error estimate follows the extrapolation (trajectory_3d.cpp:327)
		if (dt > 0.0f)
			for (int k = 0; k < 3; k++) {
				float d = dt;
				d *= tr.ev[k];
				tr.ep[k] += d;
			}
	}
	a.len[ion] = len;
	a.status[ion] = status;
	if (a.err_out)
		for (int c = 0; c < 3; c++) {
			a.err_out[6 * (size_t) ion + c] = tr.ep[c];
			a.err_out[6 * (size_t) ion + 3 + c] = tr.ev[c];
		}
}

__global__ __launch_bounds__(64) void k_trajectory(TrajArgs a)
{
	const int ion = blockIdx.x * blockDim.x + threadIdx.x;
	if (ion >= a.count) return;
	if (a.rng) {  // one ion (checked on the host): the stream stays in order
		__shared__ int table[31];
		DevRand rng;
		rand_stage(&rng, table, a.rng);
		run_ion(a, ion, &rng);
		rand_unstage(rng, a.rng);
	} else {
		run_ion(a, ion, nullptr);
	}
}

// ---------------------------------------------------------------- field lines

// Geometry_3d::Table::tetGuess (geometry_3d.cpp:1712-1748): the sampled tet
// with the vertex nearest to pos (samples from the stream, or evenly spaced)
__device__ int tet_guess(const Mesh3& g, const float* pos, int cnt, DevRand* rng)
{
	int ref = -1;
	float refd = 3.402823466e+38f;
	for (int k = 0; k < cnt; k++) {
		const int t = rng ? rng->next() % g.m : (int) (((long long) k * g.m) / cnt);
		for (int v = 0; v < 4; v++) {
			const float* c = g.P(g.tv[4 * (size_t) t + v]);
			float d = 0.0f;
			for (int i = 0; i < 3; i++) {
				float tmp = c[i] - pos[i];
				tmp *= tmp;
				d += tmp;
			}
			if (d < refd) {
				refd = d;
				ref = t;
			}
		}
	}
	return ref;
}

// FastField::compute(pos) without a guess (grid_3d.cpp:1663-1667): locate
// from tetGuess, then the guessed evaluation (a second walk)
__device__ void field_unguessed(const Mesh3& g, const float* nf, const float* p, int cnt, DevRand* rng, float* out)
{
	const int t = find_tet(g, p, tet_guess(g, p, cnt, rng), rng);
	fast_field(g, nf, p, t, out, rng);
}

__device__ void put_point(const FieldlineArgs& a, int line, int k, const float* p, const float* f, int tet)
{
	if (k >= a.cap) return;
	float* o = a.out + ((size_t) line * a.cap + k) * 7;
	for (int c = 0; c < 3; c++) {
		o[c] = p[c];
		o[3 + c] = f[c];
	}
	o[6] = (float) tet;
}

__device__ void run_fieldline(const FieldlineArgs& a, int line, DevRand* rng)
{
	Mesh3 g;
	g.xyz = a.xyz;
	g.tv = a.tv;
	g.tn = a.tn;
	g.m = a.m;
	float p0[3], bp[3], bf[3];
	for (int c = 0; c < 3; c++) p0[c] = a.start[3 * (size_t) line + c];
	field_unguessed(g, a.nfield, p0, a.guess_samples, rng, bf);
	int bt = find_tet(g, p0, tet_guess(g, p0, a.guess_samples, rng), rng);
	if (bt < 0) {
		a.len[line] = 0;
		a.status[line] = -10;  // init() throws: the start is outside the mesh
		return;
	}
	for (int c = 0; c < 3; c++) bp[c] = p0[c];
	int n = 0;
	put_point(a, line, n++, bp, bf, bt);
	// the first step: half the edge of a cube of the start tet's volume
	float delta = tet_volume(g, bt);
	delta = (float) pow((double) delta, (double) (1.0f / 3.0f));
	delta /= 2.0f;
	int status = 3;  // ITERATION_LIMIT
	long long calls = 0;
	for (long long it = 0; it < a.iterations; it++) {
		if (bf[0] == 0.0f && bf[1] == 0.0f && bf[2] == 0.0f) {
			status = 1;  // NO_FIELD
			break;
		}
		float len2 = 0.0f;
		for (int c = 0; c < 3; c++) len2 += bf[c] * bf[c];
		const float flen = sqrtf(len2);
		float dir[3];
		for (int c = 0; c < 3; c++) dir[c] = bf[c] / flen;
		bool found = false;
		float vp[3], vf[3];
		int vt = -1;
		for (long long st = 0; st < a.stepper_limit; st++) {
			if (++calls > a.max_calls) break;
			for (int c = 0; c < 3; c++) vp[c] = bp[c] + delta * dir[c];
			vt = find_tet(g, vp, bt, rng);
			// adjacentTetrahedra(back) excludes the back tet itself
			if ((vt < 0 && boundary_tet(g, a.gbnd, bt)) || adjacent(g, bt, vt)) {
				field_unguessed(g, a.nfield, vp, a.guess_samples, rng, vf);
				found = true;
				break;
			}
			if (vt == bt) delta *= 1.1f;
			else delta *= 0.9f;
		}
		if (calls > a.max_calls) {
			status = -4;  // work limit
			break;
		}
		if (!found) {
			status = 2;  // STEPPER_LIMIT
			break;
		}
		if (vt < 0) {
			status = 5;  // SYSTEM_LIMIT: the line left the mesh
			break;
		}
		if (a.distance >= 0.0f) {
			float q2 = 0.0f;
			for (int c = 0; c < 3; c++) {
				const float e = p0[c] - vp[c];
				q2 += e * e;
			}
			if (sqrtf(q2) > a.distance) {
				status = 4;  // GEOMETRY_LIMIT
				break;
			}
		}
		for (int c = 0; c < 3; c++) {
			bp[c] = vp[c];
			bf[c] = vf[c];
		}
		bt = vt;
		put_point(a, line, n++, bp, bf, bt);
	}
	a.len[line] = n;
	a.status[line] = status;
}

__global__ __launch_bounds__(64) void k_fieldlines(FieldlineArgs a)
{
	const int tid = blockIdx.x * blockDim.x + threadIdx.x;
	if (a.rng) {  // the lines in order on one stream
		if (tid != 0) return;
		__shared__ int table[31];
		DevRand rng;
		rand_stage(&rng, table, a.rng);
		for (int line = 0; line < a.count; line++) run_fieldline(a, line, &rng);
		rand_unstage(rng, a.rng);
	} else if (tid < a.count) {
		run_fieldline(a, tid, nullptr);
	}
}

// evap_initialVelocity (surface_3d.cpp:472-502) on the stream: 2kT/m split
// by two uniform draws, then a random sign per axis (five draws)
__device__ void thermal_velocity(DevRand* rng, float temperature, float mass_amu, float* v)
{
	double emax = 2.0;
	emax *= 1.3806504e-23;
	emax *= (double) temperature;
	emax /= (double) mass_amu;
	emax /= 1.660538782e-27;
	const float rmax = (float) 2147483647;
	float x = (float) rng->next();
	x /= rmax;
	x = (float) ((double) x * emax);
	float y = (float) rng->next();
	y /= rmax;
	y = (float) ((double) y * (emax - (double) x));
	float z = (float) emax;
	z -= x;
	z -= y;
	v[0] = x;
	v[1] = y;
	v[2] = z;
	for (int i = 0; i < 3; i++) {
		v[i] = sqrtf(v[i]);
		if (rng->next() % 2) v[i] *= -1.0f;
	}
}

// Grid_3d::potential(pos, guess) (grid_3d.cpp:1585-1602) after `walks`
// locations of the point, each drawing from the stream
template <typename R>
__device__ float exit_potential(const Mesh3& g, const float* p, int guess, int walks, DevRand* rng, const R* phi)
{
	int t = -1;
	for (int w = 0; w < walks; w++) t = find_tet(g, p, guess, rng);
	if (t < 0) return 0.0f;
	float wt[4];
	tet_weights(g, t, p, wt);
	float value = 0.0f;
	for (int k = 0; k < 4; k++) value += wt[k] * (float) phi[g.tv[4 * (size_t) t + k]];
	return value;
}

// one wavefront: every lane runs the flight's uniform calculation (the same
// values, the stream's table in LDS updated with the same words), lanes 0-4
// share the face tests and the weights' determinants, the walks' tets are
// cached in the lanes' registers
__global__ __launch_bounds__(64) void k_flight(FlightArgs f)
{
	if (blockIdx.x != 0) return;
	__shared__ int table[31];
	DevRand rng;
	rand_stage(&rng, table, f.t.rng);
	TrajArgs a = f.t;
	if (f.thermal) {
		float v[3];
		thermal_velocity(&rng, f.temperature, f.mass_amu, v);
		for (int c = 0; c < 3; c++) f.v0[c] = v[c];
		a.v0 = f.v0;
	}
	run_ion(a, 0, &rng, true);
	f.exit[0] = 0.0f;
	f.exit[1] = 0.0f;
	if (f.exit_walks > 0 && a.summary && a.summary[11] != 0.0f) {
		const float p[3] = {a.summary[7], a.summary[8], a.summary[9]};
		const int guess = (int) a.summary[10];
		Mesh3 g;
		g.wave = true;
		g.xyz = a.xyz;
		g.tv = a.tv;
		g.tn = a.tn;
		g.m = a.m;
		f.exit[0] = f.f64 ? exit_potential<double>(g, p, guess, f.exit_walks, &rng, (const double*) f.phi)
		                  : exit_potential<float>(g, p, guess, f.exit_walks, &rng, (const float*) f.phi);
		f.exit[1] = 1.0f;
	}
	rand_unstage(rng, f.t.rng);
}

template <typename R>
__global__ __launch_bounds__(64) void k_point_potential(int count, const float* pos, const int* guess, Mesh3 g, const R* phi, float* out,
                                  int* rng_state, int walks)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count) return;
	const float* p = pos + 3 * (size_t) i;
	int t;
	if (rng_state) {  // one point (checked on the host)
		__shared__ int table[31];
		DevRand rng;
		rand_stage(&rng, table, rng_state);
		t = find_tet(g, p, guess[i], &rng);
		for (int w = 1; w < walks; w++) t = find_tet(g, p, guess[i], &rng);
		rand_unstage(rng, rng_state);
	} else {
		t = find_tet(g, p, guess[i]);
	}
	if (t < 0) {
		out[i] = 0.0f;
		return;
	}
	float w[4];
	tet_weights(g, t, p, w);
	float value = 0.0f;
	for (int k = 0; k < 4; k++) value += w[k] * (float) phi[g.tv[4 * (size_t) t + k]];
	out[i] = value;
}

// field_o1 of every node into a float snapshot (the FastField cache)
template <typename R>
__global__ void k_field_snapshot(int n, const float* xyz, const int* ptr, const int* idx, const R* coup,
                                 const unsigned char* flags, const R* phi, float* out, int* status)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	R E[3];
	if (!field_o1<R>(s, xyz, ptr, idx, coup, flags, phi, E)) {
		atomicOr(status, 1);
		E[0] = E[1] = E[2] = (R) 0;
	}
	for (int i = 0; i < 3; i++) out[3 * (size_t) s + i] = (float) E[i];
}

}  // namespace dev

int launch_trajectories(hipStream_t s, const TrajArgs& a)
{
	if (a.count <= 0) return 0;
	if (a.rng && a.count != 1) return (int) hipErrorInvalidValue;
	hipLaunchKernelGGL(dev::k_trajectory, dim3((a.count + 63) / 64), dim3(64), 0, s, a);
	return (int) hipGetLastError();
}

int launch_fieldlines(hipStream_t s, const FieldlineArgs& a)
{
	if (a.count <= 0) return 0;
	if (a.m <= 0 || a.guess_samples < 1 || a.cap < 0) return (int) hipErrorInvalidValue;
	const int blocks = a.rng ? 1 : (a.count + 63) / 64;
	hipLaunchKernelGGL(dev::k_fieldlines, dim3(blocks), dim3(64), 0, s, a);
	return (int) hipGetLastError();
}

int launch_flight(hipStream_t s, const FlightArgs& f)
{
	if (f.t.count != 1 || !f.t.rng || !f.exit || (f.thermal && !f.v0)) return (int) hipErrorInvalidValue;
	hipLaunchKernelGGL(dev::k_flight, dim3(1), dim3(64), 0, s, f);
	return (int) hipGetLastError();
}

int launch_point_potential(hipStream_t s, int count, const float* pos, const int* guess, const float* xyz, const int* tv,
                           const int* tn, int m, const void* phi, bool f64, float* out, int* rng, int walks)
{
	if (count <= 0) return 0;
	if (rng && (count != 1 || walks < 1)) return (int) hipErrorInvalidValue;
	dev::Mesh3 g;
	g.xyz = xyz;
	g.tv = tv;
	g.tn = tn;
	g.m = m;
	if (f64)
		hipLaunchKernelGGL(dev::k_point_potential<double>, dim3((count + 63) / 64), dim3(64), 0, s, count, pos, guess, g,
		                   (const double*) phi, out, rng, walks);
	else
		hipLaunchKernelGGL(dev::k_point_potential<float>, dim3((count + 63) / 64), dim3(64), 0, s, count, pos, guess, g,
		                   (const float*) phi, out, rng, walks);
	return (int) hipGetLastError();
}

int launch_field_snapshot(hipStream_t s, int n, const float* xyz, const int* ptr, const int* idx, const void* coup,
                          const unsigned char* flags, const void* phi, bool f64, float* out, int* status)
{
	if (n <= 0) return 0;
	if (f64)
		hipLaunchKernelGGL(dev::k_field_snapshot<double>, dim3((n + 255) / 256), dim3(256), 0, s, n, xyz, ptr, idx,
		                   (const double*) coup, flags, (const double*) phi, out, status);
	else
		hipLaunchKernelGGL(dev::k_field_snapshot<float>, dim3((n + 255) / 256), dim3(256), 0, s, n, xyz, ptr, idx,
		                   (const float*) coup, flags, (const float*) phi, out, status);
	return (int) hipGetLastError();
}

}  // namespace tsg
