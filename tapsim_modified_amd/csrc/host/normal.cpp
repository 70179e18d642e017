// Surface normal of an evaporation site: Surface_3d::Table::normal
// (tapsim/surface_3d.cpp:123-169) with Geometry_3d::find_best_plane
// (tapsim/geometry_3d.cpp:1806-1870) and the Numerical Recipes tridiagonal
// reduction / implicit-shift QL of tapsim/utils.h:245-458 (Press et al.,
// Numerical Recipes in C, 2nd ed., tred2 and tqli).
//
// The result is a per-event record quantity (it does not steer the
// evaporation sequence) over a few hundred nodes, so it runs on the host next
// to the surface set.  Arithmetic follows the reference's float code, with
// its float/double promotions kept where they occur, so that the vector is
// bit-identical: the library is built without FMA contraction.
#include <algorithm>
#include <cmath>
#include <set>
#include <vector>

#include "mesh.hpp"

namespace tsg {

namespace {

// hypotenuse as the reference's inline pythag: the square root is taken in
// double of a float ratio
float pythag(float a, float b) {
	const float aa = std::fabs(a), ab = std::fabs(b);
	if (aa > ab) {
		const float q = ab / aa;
		return (float) ((double) aa * std::sqrt(1.0 + (double) (q * q)));
	}
	if (ab == 0.0f) return 0.0f;
	const float q = aa / ab;
	return (float) ((double) ab * std::sqrt(1.0 + (double) (q * q)));
}

// tred2 (utils.h:245-330): symmetric a (row-major 3x3) -> tridiagonal d, e;
// a is replaced by the orthogonal transformation
void tridiagonalise(float* a, float* d, float* e) {
	const int n = 3;
	for (int i = n - 1; i > 0; i--) {
		const int l = i - 1;
		float h = 0.0f, scale = 0.0f;
		if (l > 0) {
			for (int k = 0; k <= l; k++) scale += std::fabs(a[i * n + k]);
			if (scale == 0.0f) {
				e[i] = a[i * n + l];
			} else {
				for (int k = 0; k <= l; k++) {
					a[i * n + k] /= scale;
					h += a[i * n + k] * a[i * n + k];
				}
				float f = a[i * n + l];
				float g = f >= 0.0f ? -std::sqrt(h) : std::sqrt(h);
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
bool ql_implicit(float* d, float* e, float* z) {
	const int n = 3;
	for (int i = 1; i < n; i++) e[i - 1] = e[i];
	e[n - 1] = 0.0f;
	for (int l = 0; l < n; l++) {
		int iter = 0, m;
		do {
			for (m = l; m < n - 1; m++) {
				const float dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
				if (std::fabs(e[m]) + dd == dd) break;
			}
			if (m != l) {
				if (++iter > 10000) return false;
				float g = (float) ((double) (d[l + 1] - d[l]) / (2.0 * (double) e[l]));
				float r = pythag(g, 1.0f);
				g = d[m] - d[l] + e[l] / (g + (g >= 0.0f ? std::fabs(r) : -std::fabs(r)));
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

}  // namespace

// find_best_plane: the eigenvector of the smallest eigenvalue of the 3x3
// scatter matrix of the offsets (float sums in offset order), normalised.
// Returns false for a NaN component (the reference throws there).
bool best_plane(const std::vector<float>& off, float out[3]) {
	const size_t c = off.size() / 3;
	float a[9];
	auto moment = [&](int p, int q) {
		float s = 0.0f;
		for (size_t i = 0; i < c; i++) s += off[3 * i + p] * off[3 * i + q];
		return s;
	};
	a[0] = moment(0, 0);
	a[1] = moment(0, 1);
	a[2] = moment(0, 2);
	a[4] = moment(1, 1);
	a[5] = moment(1, 2);
	a[8] = moment(2, 2);
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
		if (std::isnan(a[i * 3 + index])) return false;
		out[i] = a[i * 3 + index];
	}
	if (out[0] != 0.0f || out[1] != 0.0f || out[2] != 0.0f) {
		float len2 = 0.0f;
		for (int i = 0; i < 3; i++) len2 += out[i] * out[i];
		const float len = std::sqrt(len2);
		for (int i = 0; i < 3; i++) out[i] /= len;
	}
	return true;
}

// Surface_3d::Table::normal: the best plane through the surface sites within
// three Delaunay hops of `node` (offsets node - site, ascending site index),
// oriented so that it points toward the vacuum nodes of the same shell.  A
// node that is not a surface site gets the zero vector.
bool surface_normal(const Mesh& mesh, const std::set<int>& surface, const int16_t* ids, int vacuum_id, int node,
                    float out[3]) {
	out[0] = out[1] = out[2] = 0.0f;
	if (!surface.count(node)) return true;
	// adjacentNodes(node, set, 3), geometry_3d.cpp:1112-1145: every node
	// within three hops, the node itself excluded
	static thread_local std::vector<uint8_t> mark;
	if ((int) mark.size() < mesh.n) mark.assign(mesh.n, 0);
	std::vector<int> shell, cur(1, node), next;
	mark[node] = 1;
	for (int lvl = 0; lvl < 3; lvl++) {
		next.clear();
		for (int u : cur)
			for (int k = mesh.ptr[u]; k < mesh.ptr[u + 1]; k++) {
				const int v = mesh.idx[k];
				if (!mark[v]) { mark[v] = 1; next.push_back(v); shell.push_back(v); }
			}
		cur.swap(next);
	}
	mark[node] = 0;
	for (int v : shell) mark[v] = 0;
	std::sort(shell.begin(), shell.end());
	const float* x0 = &mesh.xyz[3 * (size_t) node];
	std::vector<float> off;
	float toward[3] = {0.0f, 0.0f, 0.0f};
	for (int v : shell) {
		const float* x = &mesh.xyz[3 * (size_t) v];
		if (ids[v] == vacuum_id)
			for (int c = 0; c < 3; c++) toward[c] += x[c] - x0[c];
		if (surface.count(v))
			for (int c = 0; c < 3; c++) off.push_back(x0[c] - x[c]);
	}
	float n[3];
	if (!best_plane(off, n)) return false;
	float dot = 0.0f;
	for (int c = 0; c < 3; c++) dot += n[c] * toward[c];
	if (dot < 0.0f)
		for (int c = 0; c < 3; c++) n[c] *= -1.0f;
	for (int c = 0; c < 3; c++) out[c] = n[c];
	return true;
}

}  // namespace tsg
