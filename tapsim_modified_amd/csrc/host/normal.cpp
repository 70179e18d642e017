// Surface normal of an evaporation site on the host: Surface_3d::Table::normal
// (tapsim/surface_3d.cpp:123-169) with Geometry_3d::find_best_plane
// (tapsim/geometry_3d.cpp:1806-1870).  The eigen-solution is the shared
// float code of kernels/plane.hpp, which the event loop's device normal
// (kernels/surface.hip) runs too; this host version serves the
// tsg_surface_normal and tsg_best_plane entry points.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../kernels/plane.hpp"
#include "mesh.hpp"

namespace tsg {

// find_best_plane: the scatter moments as float sums in offset order
bool best_plane(const std::vector<float>& off, float out[3]) {
	const size_t c = off.size() / 3;
	float m[6];
	for (int p = 0; p < 3; p++)
		for (int q = p; q < 3; q++) {
			float s = 0.0f;
			for (size_t i = 0; i < c; i++) s += off[3 * i + p] * off[3 * i + q];
			m[plane::moment_index(p, q)] = s;
		}
	return plane::from_moments(m, out);
}

// Surface_3d::Table::normal: the best plane through the surface sites within
// three Delaunay hops of `node` (offsets node - site, ascending site index),
// oriented so that it points toward the vacuum nodes of the same shell.  A
// node that is not a surface site gets the zero vector.
bool surface_normal(const Mesh& mesh, const uint8_t* sstate, int node, float out[3]) {
	out[0] = out[1] = out[2] = 0.0f;
	if (!(sstate[node] & 1)) return true;
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
		if (sstate[v] & 2)
			for (int c = 0; c < 3; c++) toward[c] += x[c] - x0[c];
		if (sstate[v] & 1)
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
