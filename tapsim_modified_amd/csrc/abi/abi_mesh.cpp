// C ABI over the host mesh pipeline (specimen generation, Delaunay,
// topology).  See include/tapsim_gpu.h for the reference interfaces.
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../../include/tapsim_gpu.h"
#include "../host/mesh.hpp"
#include "errors.hpp"

namespace tsg {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) { g_error = msg; return code; }
}  // namespace tsg

extern "C" {

const char* tsg_last_error(void) { return tsg::g_error.c_str(); }
int tsg_version(void) { return 1; }
void tsg_free(void* p) { std::free(p); }

int tsg_specimen_generate(const tsg_specimen_params* p, int* n, float** xyz, short** ids, long long counts[5])
{
	if (!p || !n || !xyz || !ids) return tsg::fail(TSG_E_ARG, "tsg_specimen_generate: null argument");
	try {
		tsg::SpecimenParams q;
		q.a = p->a; q.radius = p->radius; q.shank = p->shank;
		for (int k = 0; k < 3; k++) q.box[k] = p->box[k];
		q.shell = p->shell; q.vacuum_target = p->vacuum_target; q.decay = p->decay;
		q.face_spacing = p->face_spacing; q.cluster_density = p->cluster_density;
		q.cluster_radius = p->cluster_radius; q.seed = p->seed;
		q.ids.vacuum = p->id_vacuum; q.ids.top = p->id_top; q.ids.bottom = p->id_bottom;
		q.ids.sides = p->id_sides; q.ids.matrix = p->id_matrix; q.ids.solute = p->id_solute;
		tsg::Specimen s;
		tsg::generate_specimen(q, &s);
		const size_t N = s.ids.size();
		*n = (int) N;
		*xyz = (float*) std::malloc(sizeof(float) * 3 * (N ? N : 1));
		*ids = (short*) std::malloc(sizeof(short) * (N ? N : 1));
		std::memcpy(*xyz, s.xyz.data(), sizeof(float) * 3 * N);
		std::memcpy(*ids, s.ids.data(), sizeof(short) * N);
		if (counts) {
			counts[0] = s.n_atoms; counts[1] = s.n_solute; counts[2] = s.n_vacuum_lattice;
			counts[3] = s.n_vacuum_random; counts[4] = s.n_box;
		}
		return TSG_OK;
	} catch (const std::exception& e) {
		return tsg::fail(TSG_E_ARG, e.what());
	}
}

int tsg_mesh_create(int n, const float* xyz, tsg_mesh** out)
{
	if (n < 4 || !xyz || !out) return tsg::fail(TSG_E_ARG, "tsg_mesh_create: bad arguments");
	try {
		tsg_mesh* h = new tsg_mesh();
		std::vector<float> pts(xyz, xyz + 3 * (size_t) n);
		std::string err;
		if (!tsg::delaunay(pts, &h->m, &err)) {
			delete h;
			return tsg::fail(TSG_E_GEOM, "tsg_mesh_create: " + err);
		}
		tsg::build_topology(&h->m);
		*out = h;
		return TSG_OK;
	} catch (const std::exception& e) {
		return tsg::fail(TSG_E_GEOM, e.what());
	}
}

int tsg_mesh_create_from_tets(int n, const float* xyz, int m, const int* tet_v, const int* tet_n, tsg_mesh** out)
{
	if (n < 4 || m < 1 || !xyz || !tet_v || !tet_n || !out) return tsg::fail(TSG_E_ARG, "tsg_mesh_create_from_tets: bad arguments");
	tsg_mesh* h = new tsg_mesh();
	h->m.n = n;
	h->m.xyz.assign(xyz, xyz + 3 * (size_t) n);
	h->m.tet_v.assign(tet_v, tet_v + 4 * (size_t) m);
	h->m.tet_n.assign(tet_n, tet_n + 4 * (size_t) m);
	for (size_t k = 0; k < h->m.tet_v.size(); k++)
		if (h->m.tet_v[k] < 0 || h->m.tet_v[k] >= n || h->m.tet_n[k] < -1 || h->m.tet_n[k] >= m) {
			delete h;
			return tsg::fail(TSG_E_ARG, "tsg_mesh_create_from_tets: index out of range");
		}
	tsg::build_topology(&h->m);
	*out = h;
	return TSG_OK;
}

int tsg_mesh_info(const tsg_mesh* h, int* n, int* m, long long* nnz)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null mesh");
	if (n) *n = h->m.n;
	if (m) *m = h->m.m();
	if (nnz) *nnz = (long long) h->m.idx.size();
	return TSG_OK;
}

int tsg_mesh_get_tets(const tsg_mesh* h, int* tv, int* tn)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null mesh");
	if (tv) std::memcpy(tv, h->m.tet_v.data(), sizeof(int) * h->m.tet_v.size());
	if (tn) std::memcpy(tn, h->m.tet_n.data(), sizeof(int) * h->m.tet_n.size());
	return TSG_OK;
}

int tsg_mesh_get_csr(const tsg_mesh* h, int* ptr, int* idx)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null mesh");
	if (ptr) std::memcpy(ptr, h->m.ptr.data(), sizeof(int) * h->m.ptr.size());
	if (idx) std::memcpy(idx, h->m.idx.data(), sizeof(int) * h->m.idx.size());
	return TSG_OK;
}

int tsg_mesh_get_nodeinfo(const tsg_mesh* h, int* assoc, unsigned char* boundary)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null mesh");
	if (assoc) std::memcpy(assoc, h->m.assoc.data(), sizeof(int) * h->m.assoc.size());
	if (boundary) std::memcpy(boundary, h->m.boundary.data(), h->m.boundary.size());
	return TSG_OK;
}

void tsg_mesh_destroy(tsg_mesh* h) { delete h; }

int tsg_best_plane(const float* offsets, int count, float out[3])
{
	if (count < 0 || (count > 0 && !offsets) || !out) return tsg::fail(TSG_E_ARG, "bad arguments");
	std::vector<float> off(offsets, offsets + 3 * (size_t) count);
	if (!tsg::best_plane(off, out)) return tsg::fail(TSG_E_NAN, "Geometry_3d::find_best_plane(): NaN");
	return TSG_OK;
}

}  // extern "C"
