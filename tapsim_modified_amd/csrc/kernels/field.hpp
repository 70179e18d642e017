// Node fields by least squares in element type R: linear
// (Grid_3d::Table::field_o1, grid_3d.cpp:909-991), shared by the surface
// evaluation and the trajectory field snapshot, and quadratic with the
// charge estimate (Grid_3d::Table::field_o2, grid_3d.cpp:993-1199).
#pragma once

#include "device_math.hpp"

namespace tsg {
namespace dev {

// field_o1 in element type R; returns false on a singular system
template <typename R>
__device__ bool field_o1(int node, const float* __restrict__ xyz, const int* __restrict__ ptr,
                         const int* __restrict__ idx, const R* __restrict__ coup, const unsigned char* __restrict__ flags,
                         const R* __restrict__ phi, R E[3])
{
	E[0] = E[1] = E[2] = (R) 0;
	if (flags[node] & 0x01) return true;  // geometric boundary: zero field
	const V3<R> pi = load_coord<R>(xyz, node);
	const R p0 = phi[node];
	R a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0;
	R b[3] = {0, 0, 0};
	// the reference forms each of the nine sums in its own loop over the same
	// ascending neighbour list; the sums are independent, so one pass with
	// nine accumulators performs exactly the same roundings.  The loads of
	// eight neighbours are issued before their (ordered) additions.
	constexpr int U = 8;
	const int k0 = ptr[node], k1 = ptr[node + 1];
	for (int kb = k0; kb < k1; kb += U) {
		int j[U];
		R cp[U], pj[U];
		V3<R> xj[U];
#pragma unroll
		for (int u = 0; u < U; u++) {
			const int k = min(kb + u, k1 - 1);
			j[u] = idx[k];
			cp[u] = coup[k];
		}
#pragma unroll
		for (int u = 0; u < U; u++) {
			xj[u] = load_coord<R>(xyz, j[u]);
			pj[u] = phi[j[u]];
		}
#pragma unroll
		for (int u = 0; u < U; u++) {
			if (kb + u >= k1 || cp[u] == (R) 0) continue;
			const V3<R> c = vsub(xj[u], pi);
			a00 += c.x[0] * c.x[0];
			a01 += c.x[0] * c.x[1];
			a02 += c.x[0] * c.x[2];
			a11 += c.x[1] * c.x[1];
			a12 += c.x[1] * c.x[2];
			a22 += c.x[2] * c.x[2];
			const R dp = pj[u] - p0;
			for (int i = 0; i < 3; i++) b[i] += dp * c.x[i];
		}
	}
	R a[9] = {a00, a01, a02, a01, a11, a12, a02, a12, a22};
	R d;
	int perm[3];
	if (!lu_decmp<R, 3>(a, perm, &d)) return false;
	lu_solve<R, 3>(a, perm, b);
	for (int i = 0; i < 3; i++) E[i] = (R) (-1.0f * b[i]);
	return true;
}

// field_o2: the 9-parameter quadratic fit over every neighbour (no
// coupling test), normal equations in double in the reference's term order,
// NR LU in double; E = -grad phi and charge = -eps0 (phi_xx + phi_yy + phi_zz).
// Returns 0, 1 for a singular system, 2 for a NaN field.
template <typename R>
__device__ int field_o2(int node, const float* __restrict__ xyz, const int* __restrict__ ptr, const int* __restrict__ idx,
                        const unsigned char* __restrict__ flags, const R* __restrict__ phi, R E[3], R* charge)
{
	E[0] = E[1] = E[2] = (R) 0;
	*charge = (R) 0;
	if (flags[node] & 0x01) return 0;
	const V3<R> pi = load_coord<R>(xyz, node);
	const R pidx = phi[node];
	const int k0 = ptr[node], k1 = ptr[node + 1];
	double a[81];
	for (int i = 0; i < 81; i++) a[i] = 0.0;
	for (int k = k0; k < k1; k++) {
		const V3<R> c = vsub(load_coord<R>(xyz, idx[k]), pi);
		const R x = c.x[0], y = c.x[1], z = c.x[2];
		// each product left to right in R, promoted by the 0.5 double literal
		// or by the double accumulator
		a[0 * 9 + 0] += x * x;
		a[0 * 9 + 1] += x * y;
		a[0 * 9 + 2] += x * z;
		a[0 * 9 + 3] += x * x * x * 0.5;
		a[0 * 9 + 4] += x * x * y;
		a[0 * 9 + 5] += x * x * z;
		a[0 * 9 + 6] += x * y * y * 0.5;
		a[0 * 9 + 7] += x * y * z;
		a[0 * 9 + 8] += x * z * z * 0.5;
		a[1 * 9 + 1] += y * y;
		a[1 * 9 + 2] += y * z;
		a[1 * 9 + 3] += y * x * x * 0.5;
		a[1 * 9 + 4] += y * x * y;
		a[1 * 9 + 5] += y * x * z;
		a[1 * 9 + 6] += y * y * y * 0.5;
		a[1 * 9 + 7] += y * y * z;
		a[1 * 9 + 8] += y * z * z * 0.5;
		a[2 * 9 + 2] += z * z;
		a[2 * 9 + 3] += z * x * x * 0.5;
		a[2 * 9 + 4] += z * x * y;
		a[2 * 9 + 5] += z * x * z;
		a[2 * 9 + 6] += z * y * y * 0.5;
		a[2 * 9 + 7] += z * y * z;
		a[2 * 9 + 8] += z * z * z * 0.5;
		a[3 * 9 + 3] += x * x * 0.5 * x * x * 0.5;
		a[3 * 9 + 4] += x * x * 0.5 * x * y;
		a[3 * 9 + 5] += x * x * 0.5 * x * z;
		a[3 * 9 + 6] += x * x * 0.5 * y * y * 0.5;
		a[3 * 9 + 7] += x * x * 0.5 * y * z;
		a[3 * 9 + 8] += x * x * 0.5 * z * z * 0.5;
		a[4 * 9 + 4] += x * y * x * y;
		a[4 * 9 + 5] += x * y * x * z;
		a[4 * 9 + 6] += x * y * y * y * 0.5;
		a[4 * 9 + 7] += x * y * y * z;
		a[4 * 9 + 8] += x * y * z * z * 0.5;
		a[5 * 9 + 5] += x * z * x * z;
		a[5 * 9 + 6] += x * z * y * y * 0.5;
		a[5 * 9 + 7] += x * z * y * z;
		a[5 * 9 + 8] += x * z * z * z * 0.5;
		a[6 * 9 + 6] += y * y * 0.5 * y * y * 0.5;
		a[6 * 9 + 7] += y * y * 0.5 * y * z;
		a[6 * 9 + 8] += y * y * 0.5 * z * z * 0.5;
		a[7 * 9 + 7] += y * z * y * z;
		a[7 * 9 + 8] += y * z * z * z * 0.5;
		a[8 * 9 + 8] += z * z * 0.5 * z * z * 0.5;
	}
	for (int i = 0; i < 9; i++)
		for (int j = 0; j < i; j++) a[i * 9 + j] = a[j * 9 + i];
	double b[9];
	for (int i = 0; i < 9; i++) b[i] = 0.0;
	for (int k = k0; k < k1; k++) {
		const V3<R> c = vsub(load_coord<R>(xyz, idx[k]), pi);
		const R pot = phi[idx[k]] - pidx;
		double t1[3];
		for (int i = 0; i < 3; i++) {
			t1[i] = pot;
			t1[i] *= c.x[i];
			b[i] += t1[i];
		}
		int row = 3;
		for (int i = 0; i < 3; i++)
			for (int j = i; j < 3; j++) {
				R t2 = (R) t1[i];
				t2 *= c.x[j];
				if (i == j) t2 *= (R) 0.5;
				b[row++] += t2;
			}
	}
	int perm[9];
	double d;
	if (!lu_decmp<double, 9>(a, perm, &d)) return 1;
	lu_solve<double, 9>(a, perm, b);
	for (int i = 0; i < 3; i++) {
		if (b[i] != b[i]) return 2;
		E[i] = (R) (-1.0f * b[i]);
	}
	// a float target: each compound step with a double operand is evaluated
	// in double and rounded back (grid_3d.cpp:1187-1191)
	R q = (R) b[3];
	q = (R) ((double) q + b[6]);
	q = (R) ((double) q + b[8]);
	q *= (R) -1.0f;
	q = (R) ((double) q * 8.85418781762e-12);
	*charge = q;
	return 0;
}

}  // namespace dev
}  // namespace tsg
