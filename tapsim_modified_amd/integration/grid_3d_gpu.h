// Drop-in for the hot methods of the reference's Grid_3d::Table
// (tapsim/grid_3d.h:100-295) on the MI355X library: a C++98 wrapper over the
// C ABI (include/tapsim_gpu.h) that a maintainer of the reference compiles
// into tapsim/ so that Process::evaporation keeps its calls.  The mesh comes
// from the reference's Geometry_3d::Table, the species from its
// Configuration::Table, the ids and potentials from its Grid_3d::Table; the
// reference's errors are rethrown with their own types and fields.
//
// Built with the reference's flags plus
//   -I<repo>/include -L<repo>/tapsim_modified_amd -ltapsim_gpu
#ifndef TAPSIM_GRID_3D_GPU_H
#define TAPSIM_GRID_3D_GPU_H

#include <stdexcept>
#include <vector>

#include "configuration.h"
#include "geometry_3d.h"
#include "grid_3d.h"
#include "mathVector_3d.h"
#include "tapsim_gpu.h"

class GpuGrid
{
public:
	// the geometry after tetrahedralize() / set_metaInformation(), the grid
	// after fastSync() (ids and both potential buffers are taken from it)
	GpuGrid(const Geometry_3d::Table& geom, const Configuration::Table& cfg, const Grid_3d::Table& grid, int mode)
	    : mesh_(0), grid_(0)
	{
		const int n = geom.numNodes(), m = geom.numTetrahedra();
		std::vector<float> xyz(3 * (size_t) n);
		std::vector<short> ids(n);
		std::vector<double> phi0(n), phi1(n);
		for (int i = 0; i < n; i++)
		{
			for (int c = 0; c < 3; c++) xyz[3 * (size_t) i + c] = geom.nodeCoords(i)[c];
			ids[i] = grid.id(i).toValue();
			phi0[i] = grid.node(i).phi(0);
			phi1[i] = grid.node(i).phi(1);
		}
		std::vector<int> tv(4 * (size_t) m), tn(4 * (size_t) m);
		for (int t = 0; t < m; t++)
			for (int k = 0; k < 4; k++)
			{
				tv[4 * (size_t) t + k] = geom.tetrahedron(t).vertices[k];
				tn[4 * (size_t) t + k] = geom.tetrahedron(t).neighbours[k];
			}
		check(tsg_mesh_create_from_tets(n, &xyz[0], m, &tv[0], &tn[0], &mesh_));
		tsg_species sp[128];
		for (int s = 0; s < 128; s++) sp[s] = tsg_species();
		const std::list<Configuration::NodeId> idList = cfg.ids();
		for (std::list<Configuration::NodeId>::const_iterator i = idList.begin(); i != idList.end(); i++)
		{
			const int s = i->toValue();
			if (s < 0 || s > 127) throw std::runtime_error("GpuGrid: species id out of range");
			const Configuration::NodeData& d = cfg[*i];
			sp[s].defined = 1;
			sp[s].charge_density = d.chargeDensity();
			sp[s].epsilon = d.epsilon();
			sp[s].removable = d.removable() ? 1 : 0;
			sp[s].neumann = d.neumannBoundary() ? 1 : 0;
			sp[s].dirichlet = d.dirichletBoundary() ? 1 : 0;
			sp[s].phi = d.phi();
			sp[s].mass = d.mass();
			sp[s].evap_charge = d.evapCharge();
			sp[s].evap_field = d.evapField();
			sp[s].evap_energy = d.evapEnergy();
		}
		check(tsg_grid_create(mesh_, &ids[0], sp, cfg.temperature(), mode, &grid_));
		check(tsg_set_phi(grid_, &phi0[0], &phi1[0], 0, 0));
	}

	~GpuGrid()
	{
		if (grid_) tsg_grid_destroy(grid_);
		if (mesh_) tsg_mesh_destroy(mesh_);
	}

	// Grid_3d::Table::relax(cycleSize) (grid_3d.cpp:591-609)
	double relax(const unsigned int cycleSize)
	{
		double dev;
		check(tsg_relax_cycles(grid_, (int) cycleSize, &dev));
		return dev;
	}

	// Grid_3d::Table::relax(threshold, cycleSize, queueSize) (grid_3d.cpp:611-671)
	signed long relax(const double threshold, const unsigned int cycleSize, const unsigned int queueSize)
	{
		long long it;
		double fdev, fslope, last;
		check(tsg_relax_threshold(grid_, threshold, (int) cycleSize, (int) queueSize, &it, &fdev, &fslope, &last));
		if (it < 0) throw Grid_3d::Table::RelaxationFault("Grid_3d::Table::relax()", (signed long) (-1 - it), fdev, fslope);
		return (signed long) it;
	}

	// Grid_3d::Table::localRelax(index, order, threshold, cycleSize, queueSize) (grid_3d.cpp:707-795)
	signed long localRelax(const int index, const unsigned int order, const double threshold,
	                       const unsigned int cycleSize, const unsigned int queueSize)
	{
		long long it;
		check(tsg_local_relax(grid_, index, (int) order, threshold, (int) cycleSize, (int) queueSize, &it));
		if (it < 0) throw Grid_3d::Table::RelaxationFault("Grid_3d::Table::localRelax()", (signed long) (-1 - it));
		return (signed long) it;
	}

	// Grid_3d::Table::potential(i) (grid_3d.h:184)
	float potential(const int i)
	{
		double v;
		check(tsg_potential(grid_, i, &v));
		return (float) v;
	}

	// Grid_3d::Table::field_o1(i, geometry) (grid_3d.cpp:909-991)
	MathVector3d<float> field_o1(const int i)
	{
		double e[3];
		check(tsg_field_o1(grid_, 1, &i, e));
		return MathVector3d<float>((float) e[0], (float) e[1], (float) e[2]);
	}

	// Grid_3d::Table::field_o2(i, geometry, &charge) (grid_3d.cpp:993-1199)
	MathVector3d<float> field_o2(const int i, float* charge)
	{
		double e[3], q;
		check(tsg_field_o2(grid_, 1, &i, e, &q));
		if (charge) *charge = (float) q;
		return MathVector3d<float>((float) e[0], (float) e[1], (float) e[2]);
	}

	// Node::setId + Grid_3d::Table::fastResync (grid_3d.cpp:555-589)
	void fastResync(const int i, const Configuration::NodeId id) { check(tsg_set_id(grid_, i, id.toValue())); }

	// Process::evaporation steps I-VI (process.cpp:680-1037) for k events
	std::vector<tsg_event> evaporate(const tsg_evap_options& o, const int k)
	{
		std::vector<tsg_event> out(k > 0 ? k : 1);
		int done = 0;
		check(tsg_evaporate(grid_, &o, k, &out[0], &done));
		out.resize(done);
		return out;
	}

	tsg_grid* handle() { return grid_; }

private:
	GpuGrid(const GpuGrid&);
	GpuGrid& operator=(const GpuGrid&);
	static void check(int rc)
	{
		if (rc) throw std::runtime_error(tsg_last_error());
	}
	tsg_mesh* mesh_;
	tsg_grid* grid_;
};

#endif
