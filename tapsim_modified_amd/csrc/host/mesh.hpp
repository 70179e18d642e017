// Host-side mesh pipeline: specimen node generation, Delaunay
// tetrahedralisation and node/tet topology.  These produce the inputs of the
// per-event path (the reference gets them from its meshgen tool and TetGen,
// tapsim/geometry_3d.cpp:759-810); the layout is the one the device path
// consumes directly.
#pragma once

#include <cstdint>
#include <string>
#include <set>
#include <vector>

namespace tsg {

// Species ids of a generated specimen (the reference's Configuration ids,
// tapsim/configuration.h:62; ids are a char in the reference).
struct SpecimenIds {
	int vacuum = 0, top = 1, bottom = 2, sides = 3, matrix = 10, solute = 11;
};

struct SpecimenParams {
	double a = 0.287e-9;          // BCC lattice constant (m)
	double radius = 5e-9;         // tip radius (m)
	double shank = 10e-9;         // cylinder height below the cap centre (m)
	double box[3] = {60e-9, 60e-9, 120e-9};  // box extent (m); x,y centred on the axis
	double shell = 2.0;           // vacuum lattice shell thickness (lattice constants)
	long long vacuum_target = 150000;  // graded random vacuum nodes
	double decay = 3.0;           // density ~ distance^-decay
	double face_spacing = 6e-9;   // spacing of box surface nodes (m)
	double cluster_density = 0.0; // solute cluster centres per m^3 (0: one phase)
	double cluster_radius = 2e-9;
	uint64_t seed = 20261203ull;
	SpecimenIds ids;
};

struct Specimen {
	std::vector<float> xyz;       // 3N metres
	std::vector<int16_t> ids;     // N
	long long n_atoms = 0, n_solute = 0, n_vacuum_lattice = 0, n_vacuum_random = 0, n_box = 0;
};

void generate_specimen(const SpecimenParams&, Specimen*);

// key of a point on the 3-D Hilbert curve, coordinates quantised to `bits` bits
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits);

// Tetrahedralisation result, in the reference's convention
// (Geometry_3d::Table::TetrahedronData, geometry_3d.h:164-176): neighbours[i]
// is the tet across the face opposite vertices[i] (-1 on the hull), every
// tet positively oriented (Geometry_3d::Table::setOrientation,
// geometry_3d.cpp:1584-1594).
struct Mesh {
	int n = 0;
	std::vector<float> xyz;
	std::vector<int> tet_v;       // 4M
	std::vector<int> tet_n;       // 4M
	// derived topology
	std::vector<int> assoc;       // last tet naming the node (setAssociatedTetrahedra)
	std::vector<uint8_t> boundary;// vertex of a hull face (setBoundary)
	std::vector<int> nt_ptr, nt_idx;   // node -> incident tets, ascending
	std::vector<int> ptr, idx;    // node -> Delaunay neighbours, ascending (adjacentNodes)
	int m() const { return (int) (tet_v.size() / 4); }
};

// Exact-predicate Bowyer-Watson.  Points must be distinct; returns false and
// fills err on failure.
bool delaunay(const std::vector<float>& xyz, Mesh* out, std::string* err, int threads = 0);

// Builds assoc, boundary, node->tet and node->node tables from tet_v/tet_n.
void build_topology(Mesh* m, int threads = 0);

// Geometry_3d::find_best_plane over offsets (x,y,z triples) and
// Surface_3d::Table::normal; both false on a NaN result (normal.cpp).
// sstate holds per node bit 0 (surface site) and bit 1 (vacuum).
bool best_plane(const std::vector<float>& off, float out[3]);
bool surface_normal(const Mesh& mesh, const uint8_t* sstate, int node, float out[3]);

// Orientation sign of det[b-a, c-a, d-a] with exact fallback.
int orient_sign(const float* a, const float* b, const float* c, const float* d);

}  // namespace tsg
