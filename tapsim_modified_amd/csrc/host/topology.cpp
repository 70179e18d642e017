// Node/tet topology derived from a tetrahedralisation, as the reference
// derives it in Geometry_3d::Table (setBoundary, setAssociatedTetrahedra,
// geometry_3d.cpp:1475-1507; adjacentNodes, :1098-1110).
#include <algorithm>
#include <thread>
#include <vector>

#include "mesh.hpp"

namespace tsg {

void build_topology(Mesh* m, int threads)
{
	const int n = m->n, M = m->m();
	m->assoc.assign(n, -1);
	m->boundary.assign(n, 0);
	m->nt_ptr.assign((size_t) n + 1, 0);
	for (int t = 0; t < M; t++)
		for (int k = 0; k < 4; k++) {
			const int v = m->tet_v[4 * (size_t) t + k];
			m->assoc[v] = t;  // the last tet naming the node
			m->nt_ptr[v + 1]++;
			if (m->tet_n[4 * (size_t) t + k] < 0)
				for (int q = 0; q < 4; q++)
					if (q != k) m->boundary[m->tet_v[4 * (size_t) t + q]] = 1;
		}
	for (int i = 0; i < n; i++) m->nt_ptr[i + 1] += m->nt_ptr[i];
	m->nt_idx.resize((size_t) m->nt_ptr[n]);
	{
		std::vector<int> fill(m->nt_ptr.begin(), m->nt_ptr.end() - 1);
		for (int t = 0; t < M; t++)
			for (int k = 0; k < 4; k++) m->nt_idx[(size_t) fill[m->tet_v[4 * (size_t) t + k]]++] = t;
	}
	// neighbours: vertices of incident tets, ascending and unique
	if (threads <= 0) threads = (int) std::max(1u, std::thread::hardware_concurrency());
	threads = std::min(threads, 64);
	std::vector<int> deg(n, 0);
	std::vector<std::vector<int>> part(threads);
	auto work = [&](int tid) {
		const int lo = (int) ((long long) n * tid / threads), hi = (int) ((long long) n * (tid + 1) / threads);
		std::vector<int> tmp;
		std::vector<int>& out = part[tid];
		for (int i = lo; i < hi; i++) {
			tmp.clear();
			for (int k = m->nt_ptr[i]; k < m->nt_ptr[i + 1]; k++)
				for (int q = 0; q < 4; q++) {
					const int v = m->tet_v[4 * (size_t) m->nt_idx[k] + q];
					if (v != i) tmp.push_back(v);
				}
			std::sort(tmp.begin(), tmp.end());
			tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
			deg[i] = (int) tmp.size();
			out.insert(out.end(), tmp.begin(), tmp.end());
		}
	};
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; t++) pool.emplace_back(work, t);
	for (auto& th : pool) th.join();
	m->ptr.assign((size_t) n + 1, 0);
	for (int i = 0; i < n; i++) m->ptr[i + 1] = m->ptr[i] + deg[i];
	m->idx.resize((size_t) m->ptr[n]);
	size_t off = 0;
	for (int t = 0; t < threads; t++) {
		std::copy(part[t].begin(), part[t].end(), m->idx.begin() + off);
		off += part[t].size();
	}
}

}  // namespace tsg
