// Best-fit plane of a 3x3 scatter matrix, shared by the host (the
// tsg_surface_normal / tsg_best_plane entry points) and the device (the event
// loop's surface normal): Geometry_3d::find_best_plane
// (tapsim/geometry_3d.cpp:1806-1870) with the Numerical Recipes tridiagonal
// reduction and implicit-shift QL of tapsim/utils.h:245-458 (Press et al.,
// Numerical Recipes in C, 2nd ed., tred2 and tqli).
//
// Arithmetic follows the reference's float code, with its float/double
// promotions kept where they occur; both compilations are built without
// multiply-add contraction and with IEEE division and square root, so the
// host and device results are the same bits.
#pragma once

#include <math.h>

#ifdef __HIP__
#define TSG_HD __host__ __device__
#else
#define TSG_HD
#endif

namespace tsg {
namespace plane {

// hypotenuse as the reference's inline pythag: the square root is taken in
// double of a float ratio
TSG_HD inline float pythag(float a, float b)
{
	const float aa = fabsf(a), ab = fabsf(b);
	if (aa > ab) {
		const float q = ab / aa;
		return (float) ((double) aa * sqrt(1.0 + (double) (q * q)));
	}
	if (ab == 0.0f) return 0.0f;
	const float q = aa / ab;
	return (float) ((double) ab * sqrt(1.0 + (double) (q * q)));
}

// tred2 (utils.h:245-330): symmetric a (row-major 3x3) -> tridiagonal d, e;
// a is replaced by the orthogonal transformation
TSG_HD inline void tridiagonalise(float* a, float* d, float* e)
{
	const int n = 3;
	for (int i = n - 1; i > 0; i--) {
		const int l = i - 1;
		float h = 0.0f, scale = 0.0f;
		if (l > 0) {
			for (int k = 0; k <= l; k++) scale += fabsf(a[i * n + k]);
			if (scale == 0.0f) {
				e[i] = a[i * n + l];
			} else {
				for (int k = 0; k <= l; k++) {
					a[i * n + k] /= scale;
					h += a[i * n + k] * a[i * n + k];
				}
				float f = a[i * n + l];
				float g = f >= 0.0f ? -sqrtf(h) : sqrtf(h);
				e[i] = scale * g;
				h -= f * g;
				a[i * n + l] = f - g;
				f = 0.0f;
				for (int j = 0; j <= l; j++) {
					a[j * n + i] = a[i * n + j] / h;
					g = 0.0f;
					for (int k = 0; k <= j; k++) g += a[j * n + k] * a[i * n + k];
					for (int k = j + 1; k <= l; k++) g += a[k * n + j] * a[i * n + k];
					e[j] = g / h;
					f += e[j] * a[i * n + j];
				}
				const float hh = f / (h + h);
				for (int j = 0; j <= l; j++) {
					f = a[i * n + j];
					g = e[j] - hh * f;
					e[j] = g;
					for (int k = 0; k <= j; k++) a[j * n + k] -= (f * e[k] + g * a[i * n + k]);
				}
			}
		} else {
			e[i] = a[i * n + l];
		}
		d[i] = h;
	}
	d[0] = 0.0f;
	e[0] = 0.0f;
	for (int i = 0; i < n; i++) {
		const int l = i;
		if (d[i] != 0.0f) {
			for (int j = 0; j < l; j++) {
				float g = 0.0f;
				for (int k = 0; k < l; k++) g += a[i * n + k] * a[k * n + j];
				for (int k = 0; k < l; k++) a[k * n + j] -= g * a[k * n + i];
			}
		}
		d[i] = a[i * n + i];
		a[i * n + i] = 1.0f;
		for (int j = 0; j < l; j++) a[j * n + i] = a[i * n + j] = 0.0f;
	}
}

// tqli (utils.h:358-456): eigenvalues into d, eigenvectors as the columns of
// z.  The reference constructs but never throws its iteration-limit error, so
// it iterates until convergence; a NaN matrix would loop for ever there.  Here
// the loop stops after 10^4 sweeps of one eigenvalue and reports failure.
TSG_HD inline bool ql_implicit(float* d, float* e, float* z)
{
	const int n = 3;
	for (int i = 1; i < n; i++) e[i - 1] = e[i];
	e[n - 1] = 0.0f;
	for (int l = 0; l < n; l++) {
		int iter = 0, m;
		do {
			for (m = l; m < n - 1; m++) {
				const float dd = fabsf(d[m]) + fabsf(d[m + 1]);
				if (fabsf(e[m]) + dd == dd) break;
			}
			if (m != l) {
				if (++iter > 10000) return false;
				float g = (float) ((double) (d[l + 1] - d[l]) / (2.0 * (double) e[l]));
				float r = pythag(g, 1.0f);
				g = d[m] - d[l] + e[l] / (g + (g >= 0.0f ? fabsf(r) : -fabsf(r)));
				float s = 1.0f, c = 1.0f, p = 0.0f;
				int i;
				for (i = m - 1; i >= l; i--) {
					float f = s * e[i];
					const float b = c * e[i];
					r = pythag(f, g);
					e[i + 1] = r;
					if (r == 0.0f) {
						d[i + 1] -= p;
						e[m] = 0.0f;
						break;
					}
					s = f / r;
					c = g / r;
					g = d[i + 1] - p;
					r = (float) ((double) ((d[i] - g) * s) + 2.0 * (double) c * (double) b);
					d[i + 1] = g + (s * r);
					p = s * r;
					g = c * r - b;
					for (int k = 0; k < n; k++) {
						f = z[k * n + i + 1];
						z[k * n + i + 1] = s * z[k * n + i] + c * f;
						z[k * n + i] = c * z[k * n + i] - s * f;
					}
				}
				if (r == 0.0f && i >= l) continue;
				d[l] -= p;
				e[l] = g;
				e[m] = 0.0f;
			}
		} while (m != l);
	}
	return true;
}

// find_best_plane from the scatter moments m = (xx, xy, xz, yy, yz, zz)
// (float sums over the offsets in offset order): the eigenvector of the
// smallest eigenvalue, normalised.  Returns false for a NaN component (the
// reference throws there) or a QL iteration that does not converge.
TSG_HD inline bool from_moments(const float m[6], float out[3])
{
	float a[9];
	a[0] = m[0];
	a[1] = m[1];
	a[2] = m[2];
	a[4] = m[3];
	a[5] = m[4];
	a[8] = m[5];
	a[3] = a[1];
	a[6] = a[2];
	a[7] = a[5];
	float d[3] = {0.0f, 0.0f, 0.0f}, e[3] = {0.0f, 0.0f, 0.0f};
	tridiagonalise(a, d, e);
	if (!ql_implicit(d, e, a)) return false;
	int index = 0;
	for (int i = 1; i < 3; i++)
		if (d[i] < d[index]) index = i;
	for (int i = 0; i < 3; i++) {
		const float v = a[i * 3 + index];
		if (v != v) return false;
		out[i] = v;
	}
	if (out[0] != 0.0f || out[1] != 0.0f || out[2] != 0.0f) {
		float len2 = 0.0f;
		for (int i = 0; i < 3; i++) len2 += out[i] * out[i];
		const float len = sqrtf(len2);
		for (int i = 0; i < 3; i++) out[i] /= len;
	}
	return true;
}

// the moment index of (p, q), p <= q
TSG_HD inline int moment_index(int p, int q) { return p == 0 ? q : (p == 1 ? 2 + q : 5); }

}  // namespace plane
}  // namespace tsg
