/*
 * Fixture driver for the reference TAPSim build (test infrastructure only).
 *
 * Links the reference's own objects (built by oracle/Makefile.ref from the
 * reference sources) and runs their Grid_3d / Surface_3d / Process code on a
 * mesh given to it, writing raw little-endian arrays that tests/golden/
 * turns into committed fixtures.  Nothing in the product links this.
 *
 * usage: ref_driver IN.bin CONFIG.aly OUTDIR STEP [STEP ...]
 *
 * IN.bin (all little endian):
 *   char[4] "TSGM", int32 version (1), int32 N, int32 M,
 *   float32 coords[3N] (metres), int16 ids[N], int32 tets[4M]
 *   M == 0 -> the reference meshes the nodes itself with TetGen
 *             (Geometry_3d::Table::tetrahedralize, tapsim/geometry_3d.cpp:759)
 *   M  > 0 -> the given tets are used (positively oriented) and the
 *             reference derives neighbours/boundary/associated tets
 *             (Geometry_3d::Table::set_metaInformation, geometry_3d.cpp:695)
 *
 * STEPs are applied in order to one System_3d:
 *   mesh                 build the tetrahedralisation
 *   sync                 Grid_3d::Table::fastSync      (grid_3d.cpp:491)
 *   dumpmesh             tets, neighbours, associated tets, boundary flags
 *   dumpgrid             CSR neighbours, couplings, charges, flags
 *   dumpphi:TAG          both potential buffers + potential() buffer index
 *   relax:K:C            K x Grid_3d::Table::relax(C)   (grid_3d.cpp:591)
 *   relaxthr:T:C:Q       Grid_3d::Table::relax(T,C,Q)   (grid_3d.cpp:611)
 *   local:NODE:ORD:T:C:Q Grid_3d::Table::localRelax     (grid_3d.cpp:707)
 *   evapnode:NODE        setId(vacuum) + fastResync     (surface_3d.cpp:513)
 *   field:TAG            field_o1 / field_o2 of every node (grid_3d.cpp:909,993)
 *   vorarea:TAG          voronoiArea(i,j,1.0f) both directions for every edge
 *   evap:...             the steps I-VI of Process::evaporation (process.cpp:680-1037)
 *   traj:TAG:NODE:INT:STEP  one Trajectory_3d from a node (trajectory_3d.cpp:99); its phases,
 *                        status and final error estimate
 *   writetraj:IN:OUT:MODE:CHUNK  File_Io::LogTrajectory of the datasets in IN
 *   writedump:NAME:INDEX  File_Io::writeSystem (configuration, geometry, grid)
 *   randreset:SEED       std::srand(SEED), then Grid_3d::Table::randomReset (grid_3d.cpp:811)
 *   threads:T            Grid_3d::Table::setThreadNum(T)  (grid_3d.cpp:194; the driver starts with 1)
 *
 * The wall time of every step (steady clock, seconds) goes to step_seconds.f64.
 */

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <set>
#include <fstream>
#include <stdexcept>
#include <queue>
#include <limits>
#include <cmath>
#include <ctime>
#include <iostream>
#include <list>
#include <map>
#include <ostream>
#include <sstream>
#include <typeinfo>
#include <pthread.h>
#include <unistd.h>

// The reference's classes are opened (private -> public; the layout is
// unchanged) for two things they have no setters for: Fieldline_3d's
// stepper and iteration limits, which its constructor leaves uninitialised,
// and the status and error estimate of a Trajectory_3d handed to
// File_Io::LogTrajectory.  Every standard header they use is included above.
#define private public
#include "system_3d.h"
#include "file_io.h"
#include "logResults.h"
#include "logTrajectory.h"
#include "process.h"
#include "surface_3d.h"
#include "trajectory_3d.h"
#include "fieldline_3d.h"
#undef private

namespace
{
	std::string g_out;

	template<class T>
	void writeArray(const std::string& name, const std::vector<T>& v)
	{
		const std::string path = g_out + "/" + name;
		FILE* f = std::fopen(path.c_str(), "wb");
		if (f == 0) throw std::runtime_error("cannot write " + path);
		if (!v.empty()) std::fwrite(&v[0], sizeof(T), v.size(), f);
		std::fclose(f);
	}

	std::vector<float> readFloats(const std::string& name)
	{
		const std::string path = g_out + "/" + name;
		FILE* f = std::fopen(path.c_str(), "rb");
		if (f == 0) throw std::runtime_error("cannot read " + path);
		std::vector<float> v;
		float x;
		while (std::fread(&x, sizeof(float), 1, f) == 1) v.push_back(x);
		std::fclose(f);
		return v;
	}

	std::vector<std::string> split(const std::string& s, char c)
	{
		std::vector<std::string> out;
		std::string cur;
		for (size_t i = 0; i < s.size(); i++)
		{
			if (s[i] == c) { out.push_back(cur); cur.clear(); }
			else cur += s[i];
		}
		out.push_back(cur);
		return out;
	}

	struct Input
	{
		int n, m;
		std::vector<float> coords;
		std::vector<short> ids;
		std::vector<int> tets;
	};

	Input readInput(const char* path)
	{
		Input in;
		FILE* f = std::fopen(path, "rb");
		if (f == 0) throw std::runtime_error("cannot open input");
		char magic[4];
		int version;
		if (std::fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "TSGM", 4) != 0) throw std::runtime_error("bad magic");
		if (std::fread(&version, 4, 1, f) != 1 || version != 1) throw std::runtime_error("bad version");
		if (std::fread(&in.n, 4, 1, f) != 1 || std::fread(&in.m, 4, 1, f) != 1) throw std::runtime_error("bad header");
		in.coords.resize(3 * (size_t)in.n);
		in.ids.resize(in.n);
		in.tets.resize(4 * (size_t)in.m);
		if (std::fread(&in.coords[0], 4, in.coords.size(), f) != in.coords.size()) throw std::runtime_error("short coords");
		if (std::fread(&in.ids[0], 2, in.ids.size(), f) != in.ids.size()) throw std::runtime_error("short ids");
		if (in.m > 0 && std::fread(&in.tets[0], 4, in.tets.size(), f) != in.tets.size()) throw std::runtime_error("short tets");
		std::fclose(f);
		return in;
	}

	void dumpPhi(const System_3d& s, const std::string& tag)
	{
		const int n = s.gridTable.numNodes();
		std::vector<float> p0(n), p1(n), pot(n);
		for (int i = 0; i < n; i++)
		{
			p0[i] = s.gridTable.node(i).phi(0);
			p1[i] = s.gridTable.node(i).phi(1);
			pot[i] = s.gridTable.potential(i);
		}
		writeArray("phi0_" + tag + ".f32", p0);
		writeArray("phi1_" + tag + ".f32", p1);
		writeArray("potential_" + tag + ".f32", pot);
	}
}

int main(int argc, char** argv)
{
	if (argc < 5)
	{
		std::fprintf(stderr, "usage: ref_driver IN.bin CONFIG.aly OUTDIR STEP...\n");
		return 2;
	}

	std::srand(99); // as tapsim/tapsim.cpp:235

	try
	{
		const Input in = readInput(argv[1]);
		g_out = argv[3];

		System_3d system;
		system.gridTable.setThreadNum(1);
		File_Io::readConfig(argv[2], &system.configTable);

		system.geomTable.allocate(in.n);
		system.gridTable.allocate(in.n);
		for (int i = 0; i < in.n; i++)
		{
			system.geomTable.node(i).coords = Geometry_3d::Point(in.coords[3*i], in.coords[3*i+1], in.coords[3*i+2]);
			system.geomTable.node(i).boundary = false;
			system.gridTable.node(i).setId(Configuration::NodeId::fromValue(in.ids[i]));
			system.gridTable.node(i).setNumber(Configuration::NodeNumber::fromValue(i + 1));
			const float phi = system.configTable[system.gridTable.id(i)].phi();
			system.gridTable.node(i).setPhi(0, phi);
			system.gridTable.node(i).setPhi(1, phi);
		}

		std::vector<double> relaxDeviations;
		std::vector<long long> relaxIterations;
		std::vector<double> stepSeconds;

		for (int a = 4; a < argc; a++)
		{
			const std::vector<std::string> tok = split(argv[a], ':');
			const std::string& cmd = tok[0];
			std::fprintf(stderr, "[ref_driver] step %s\n", argv[a]);
			struct timespec stepStart;
			clock_gettime(CLOCK_MONOTONIC, &stepStart);

			if (cmd == "mesh")
			{
				if (in.m == 0)
					system.geomTable.tetrahedralize();
				else
				{
					system.geomTable.allocateTets(in.m);
					for (int t = 0; t < in.m; t++)
						for (int k = 0; k < 4; k++)
							system.geomTable.tetrahedron(t).vertices[k] = in.tets[4*t+k];
					system.geomTable.set_metaInformation();
				}
			}
			else if (cmd == "sync")
			{
				system.gridTable.fastSync(system.geomTable, system.configTable);
			}
			else if (cmd == "dumpmesh")
			{
				const int m = system.geomTable.numTetrahedra();
				std::vector<int> tets(4 * (size_t)m), neigh(4 * (size_t)m), assoc(in.n);
				std::vector<unsigned char> boundary(in.n);
				for (int t = 0; t < m; t++)
					for (int k = 0; k < 4; k++)
					{
						tets[4*t+k] = system.geomTable.tetrahedron(t).vertices[k];
						neigh[4*t+k] = system.geomTable.tetrahedron(t).neighbours[k];
					}
				for (int i = 0; i < in.n; i++)
				{
					assoc[i] = system.geomTable.node(i).associatedTetrahedron;
					boundary[i] = system.geomTable.node(i).boundary ? 1 : 0;
				}
				writeArray("tets.i32", tets);
				writeArray("neigh.i32", neigh);
				writeArray("assoc.i32", assoc);
				writeArray("boundary.u8", boundary);
			}
			else if (cmd == "dumpgrid")
			{
				const int n = system.gridTable.numNodes();
				std::vector<int> ptr(n + 1, 0), idx;
				std::vector<float> coup, charge(n);
				std::vector<unsigned char> flags(n);
				std::vector<short> ids(n);
				for (int i = 0; i < n; i++)
				{
					const Grid_3d::Node& node = system.gridTable.node(i);
					for (int j = 0; j < node.numNeighbours(); j++)
					{
						idx.push_back(node.neighbour(j));
						coup.push_back(node.coupling(j));
					}
					ptr[i+1] = (int) idx.size();
					charge[i] = node.charge();
					flags[i] = (node.boundary() ? 1 : 0) | (node.dirichletBoundary() ? 2 : 0) | (node.neumannBoundary() ? 4 : 0);
					ids[i] = node.id().toValue();
				}
				writeArray("nbr_ptr.i32", ptr);
				writeArray("nbr_idx.i32", idx);
				writeArray("coupling.f32", coup);
				writeArray("charge.f32", charge);
				writeArray("flags.u8", flags);
				writeArray("ids.i16", ids);
			}
			else if (cmd == "dumpphi")
			{
				dumpPhi(system, tok.at(1));
			}
			else if (cmd == "relax")
			{
				const int k = std::atoi(tok.at(1).c_str());
				const unsigned int c = std::atoi(tok.at(2).c_str());
				for (int i = 0; i < k; i++)
					relaxDeviations.push_back(system.gridTable.relax(c));
				writeArray("relax_dev.f64", relaxDeviations);
			}
			else if (cmd == "relaxthr")
			{
				const double thr = std::atof(tok.at(1).c_str());
				const unsigned int c = std::atoi(tok.at(2).c_str());
				const unsigned int q = std::atoi(tok.at(3).c_str());
				long long it;
				double faultDev = -1.0, faultSlope = 0.0;
				try { it = system.gridTable.relax(thr, c, q); }
				catch (Grid_3d::Table::RelaxationFault& e) { it = -1 - e.iteration(); faultDev = e.deviation(); faultSlope = e.slope(); }
				relaxIterations.push_back(it);
				std::vector<double> fault(2);
				fault[0] = faultDev; fault[1] = faultSlope;
				writeArray("relaxthr_iter.i64", relaxIterations);
				writeArray("relaxthr_fault_" + tok.at(1) + ".f64", fault);
			}
			else if (cmd == "local")
			{
				const int node = std::atoi(tok.at(1).c_str());
				const unsigned int order = std::atoi(tok.at(2).c_str());
				const double thr = std::atof(tok.at(3).c_str());
				const unsigned int c = std::atoi(tok.at(4).c_str());
				const unsigned int q = std::atoi(tok.at(5).c_str());
				long long it;
				try { it = system.gridTable.localRelax(node, order, thr, c, q); }
				catch (Grid_3d::Table::RelaxationFault& e) { it = -1 - e.iteration(); }
				relaxIterations.push_back(it);
				writeArray("local_iter.i64", relaxIterations);
			}
			else if (cmd == "threads")
			{
				system.gridTable.setThreadNum((unsigned int) std::atoi(tok.at(1).c_str()));
			}
			else if (cmd == "randreset")
			{
				std::srand((unsigned int) std::atoi(tok.at(1).c_str()));
				system.gridTable.randomReset(system.configTable);
			}
			else if (cmd == "evapnode")
			{
				const int node = std::atoi(tok.at(1).c_str());
				system.gridTable.node(node).setId(system.configTable["Vacuum"].id());
				system.gridTable.fastResync(node, system.geomTable, system.configTable);
			}
			else if (cmd == "field")
			{
				const int n = system.gridTable.numNodes();
				std::vector<float> e1(3 * (size_t)n), e2(3 * (size_t)n), q2(n);
				for (int i = 0; i < n; i++)
				{
					const MathVector3d<float> f1 = system.gridTable.field_o1(i, system.geomTable);
					float q = 0.0f;
					const MathVector3d<float> f2 = system.gridTable.field_o2(i, system.geomTable, &q);
					for (int k = 0; k < 3; k++) { e1[3*i+k] = f1[k]; e2[3*i+k] = f2[k]; }
					q2[i] = q;
				}
				writeArray("field_o1_" + tok.at(1) + ".f32", e1);
				writeArray("field_o2_" + tok.at(1) + ".f32", e2);
				writeArray("charge_o2_" + tok.at(1) + ".f32", q2);
			}
			else if (cmd == "vorarea")
			{
				const int n = system.gridTable.numNodes();
				std::vector<float> a;
				for (int i = 0; i < n; i++)
				{
					const Grid_3d::Node& node = system.gridTable.node(i);
					for (int j = 0; j < node.numNeighbours(); j++)
						a.push_back(system.geomTable.voronoiArea(i, node.neighbour(j), 1.0f));
				}
				writeArray("vorarea_" + tok.at(1) + ".f32", a);
			}
			else if (cmd == "evap")
			{
				// evap:EVENTS:PROB:CHOICE:LTHR:SHELL:LCYC:LQUE:GLOBAL:RINT:RTHR:RCYC:RQUE:INTEG:STEPPER[:THERMAL[:COORD]]
				// The steps I-VI of Process::evaporation (process.cpp:680-1037)
				// composed of the reference's own functions.  The fork's
				// update_coordination (process.cpp:647,770) is left out: it adds
				// neighbour counts to species ids every event and the
				// configuration lookup then throws for the undefined ids
				// (configuration.cpp:214-232), so upstream TAPSim semantics are
				// what can be run and compared.
				const int events = std::atoi(tok.at(1).c_str());
				const int probMode = std::atoi(tok.at(2).c_str());
				const int evapMode = std::atoi(tok.at(3).c_str());
				const float lthr = (float) std::atof(tok.at(4).c_str());
				const unsigned int shell = std::atoi(tok.at(5).c_str());
				const unsigned int lcyc = std::atoi(tok.at(6).c_str());
				const unsigned int lque = std::atoi(tok.at(7).c_str());
				const unsigned int gcyc = std::atoi(tok.at(8).c_str());
				const int rint = std::atoi(tok.at(9).c_str());
				const float rthr = (float) std::atof(tok.at(10).c_str());
				const unsigned int rcyc = std::atoi(tok.at(11).c_str());
				const unsigned int rque = std::atoi(tok.at(12).c_str());
				const int integ = std::atoi(tok.at(13).c_str());
				const int stepper = std::atoi(tok.at(14).c_str());
				const int thermal = tok.size() > 15 ? std::atoi(tok.at(15).c_str()) : 0;
				// COORD 1: the fork's update_coordination where its Process calls it
				// (process.cpp:647 after the initial surface evaluation, :770 after
				// every take-away)
				const int coord = tok.size() > 16 ? std::atoi(tok.at(16).c_str()) : 0;

				Surface_3d::Table surf;
				surf.setVacuumId(system.configTable["Vacuum"].id());
				surf.init(system);
				Surface_3d::evap_compute_specificFields(&surf, system);
				Surface_3d::evap_compute_probabilities(probMode, &surf, system);
				if (coord) update_coordination(&system);

				std::vector<float> rec, recNormal;
				std::vector<int> recNode;
				std::vector<long long> recIter;
				std::queue<float> voltageQueue;
				float voltageQueueSum = 0.0f;
				for (int ev = 1; ev <= events && surf.nodes().size() > 0; ev++)
				{
					const float scalingReference = surf.scalingReference();
					const Surface_3d::Nodeset::const_iterator it = Surface_3d::evap_findCandidate(evapMode, surf);
					const int node = it->index();
					const float probability = it->probability();
					const float potBefore = system.gridTable.potential(node);
					const MathVector3d<float> fieldBefore = system.gridTable.field_o1(node, system.geomTable);
					const Geometry_3d::Point apex = system.geomTable.nodeCoords(surf.apex(system)->index());
					const Geometry_3d::Point normal = surf.normal(node, system);
					const Configuration::NodeId nodeId = system.gridTable.id(node);

					evap_takeAway(&system, &surf, node);
					if (coord) update_coordination(&system);

					long long localIter = 0, refreshIter = 0;
					try { localIter = system.gridTable.localRelax(node, shell, lthr, lcyc, lque); }
					catch (Grid_3d::Table::RelaxationFault& e) { localIter = -1 - e.iteration(); }
					try { system.gridTable.relax(gcyc); }
					catch (Grid_3d::Table::RelaxationFault&) {}
					if (rint > 0 && ev % rint == 0)
					{
						try { refreshIter = system.gridTable.relax(rthr, rcyc, rque); }
						catch (Grid_3d::Table::RelaxationFault& e) { refreshIter = -1 - e.iteration(); }
					}

					Surface_3d::evap_compute_specificFields(&surf, system);
					Surface_3d::evap_compute_probabilities(probMode, &surf, system);

					float charge = system.configTable[nodeId].evapCharge();
					charge *= Trajectory_3d::eCharge;
					float mass = system.configTable[nodeId].mass();
					mass *= Trajectory_3d::amu2kg;
					Trajectory_3d traj;
					traj.setGeometry(&system.geomTable);
					traj.setGrid(&system.gridTable);
					traj.setIntegratorType(integ);
					traj.setStepperType(stepper);
					traj.setTimeStepLimit(1e-16f);
					traj.setStepperLimit(std::numeric_limits<unsigned long>::max());
					traj.setIterationLimit(std::numeric_limits<unsigned long>::max());
					MathVector3d<float> velocity(0.0f);
					if (thermal) velocity = evap_initialVelocity(nodeId, system.configTable);
					traj.init(system.geomTable.nodeCoords(node), velocity, charge, mass,
					          system.geomTable.node(node).associatedTetrahedron);
					traj.integrate(1e-14f, 1e-6f);
					traj.extrapolate(Geometry_3d::Point::Z, system.geomTable.max().z());

					const float potAfter = system.gridTable.potential(node);
					const MathVector3d<float> fieldAfter = system.gridTable.field_o1(node, system.geomTable);
					float voltage, timeScale;
					std::vector<Trajectory_3d::phaseVector>::const_reverse_iterator back = traj.data().rbegin();
					while (traj.data().rend() != back && back->tetIndex() < 0) back++;
					if (traj.data().rend() != back && fieldBefore.length() != 0.0f)
					{
						float v = potBefore;
						v -= Grid_3d::potential(system.geomTable, system.gridTable, back->position(), back->tetIndex());
						v *= system.configTable[nodeId].evapField();
						v /= fieldBefore.length();
						v /= scalingReference;
						voltageQueue.push(v);
						voltageQueueSum += v;
						while (voltageQueue.size() > 250) { voltageQueueSum -= voltageQueue.front(); voltageQueue.pop(); }
						voltage = voltageQueueSum;
						voltage /= voltageQueue.size();
						timeScale = potAfter;
						timeScale -= Grid_3d::potential(system.geomTable, system.gridTable, back->position(), back->tetIndex());
						timeScale /= voltage;
						timeScale = std::sqrt(timeScale);
					}
					else if (fieldBefore.length() == 0.0f) { voltage = 1.0f; timeScale = 1.0f; }
					else { voltage = 0.0f; timeScale = 1.0f; }

					recNode.push_back(node);
					recNode.push_back(nodeId.toValue());
					recNode.push_back((int) traj.data().size());
					recNode.push_back((int) surf.nodes().size());
					recIter.push_back(localIter);
					recIter.push_back(refreshIter);
					rec.push_back(probability);
					rec.push_back(potBefore);
					for (int c = 0; c < 3; c++) rec.push_back(fieldBefore[c]);
					rec.push_back(potAfter);
					for (int c = 0; c < 3; c++) rec.push_back(fieldAfter[c]);
					rec.push_back(voltage);
					rec.push_back(timeScale);
					for (int c = 0; c < 3; c++) rec.push_back(traj.data().front().position()[c]);
					for (int c = 0; c < 3; c++) rec.push_back(traj.data().back().position()[c]);
					rec.push_back(traj.data().back().time() * timeScale);
					for (int c = 0; c < 3; c++) rec.push_back(apex[c]);
					rec.push_back(surf.scalingReference());
					for (int c = 0; c < 3; c++) recNormal.push_back(normal[c]);
				}
				writeArray("evap_normal.f32", recNormal);
				// where the events leave the process's rand() stream
				std::vector<int> nextRand(1, std::rand());
				writeArray("evap_next_rand.i32", nextRand);
				writeArray("evap_int.i32", recNode);
				writeArray("evap_iter.i64", recIter);
				writeArray("evap_f.f32", rec);
			}
			else if (cmd == "traj")
			{
				// traj:TAG:NODE:INTEGRATOR:STEPPER  (charge 2e, mass 55.845 u)
				const int node = std::atoi(tok.at(2).c_str());
				Trajectory_3d t;
				t.setGeometry(&system.geomTable);
				t.setGrid(&system.gridTable);
				t.setIntegratorType(std::atoi(tok.at(3).c_str()));
				t.setStepperType(std::atoi(tok.at(4).c_str()));
				t.setTimeStepLimit(1e-16f);
				t.setStepperLimit(100000);
				t.setIterationLimit(1000000);
				const float charge = 2.0f * Trajectory_3d::eCharge;
				const float mass = 55.845f * Trajectory_3d::amu2kg;
				t.init(system.geomTable.nodeCoords(node), MathVector3d<float>(0.0f), charge, mass,
				       system.geomTable.node(node).associatedTetrahedron);
				const int status = t.integrate(1e-14f, 1e-6f);
				t.extrapolate(Geometry_3d::Point::Z, system.geomTable.max().z());
				std::vector<float> data;
				for (size_t k = 0; k < t.data().size(); k++)
				{
					const Trajectory_3d::phaseVector& p = t.data()[k];
					data.push_back(p.time());
					for (int c = 0; c < 3; c++) data.push_back(p.position(c));
					for (int c = 0; c < 3; c++) data.push_back(p.velocity(c));
					data.push_back((float) p.tetIndex());
				}
				std::vector<int> st(1, status);
				writeArray("traj_" + tok.at(1) + ".f32", data);
				writeArray("traj_status_" + tok.at(1) + ".i32", st);
				std::vector<float> err;
				for (int c = 0; c < 3; c++) err.push_back(t.error_estimate().position(c));
				for (int c = 0; c < 3; c++) err.push_back(t.error_estimate().velocity(c));
				writeArray("traj_err_" + tok.at(1) + ".f32", err);
			}
			else if (cmd == "writegeom")
			{
				// writegeom:NAME:MODE:NUMBERS:POTENTIALS -- File_Io::writeGeomGrid
				// of the current grid into NAME in the work directory (MODE 0
				// ASCII, 1 BINARY)
				File_Io::writeGeomGrid((g_out + "/" + tok.at(1)).c_str(), system.geomTable, system.gridTable,
				                       std::atoi(tok.at(2).c_str()), std::atoi(tok.at(3).c_str()),
				                       std::atoi(tok.at(4).c_str()));
			}
			else if (cmd == "readgeom")
			{
				// readgeom:NAME:TAG -- File_Io::readGeomGrid of NAME in the work
				// directory into fresh tables; TAG_xyz.f32, TAG_ids.i16,
				// TAG_numbers.u32, TAG_phi.f32 and TAG_flags.i32 (the reported
				// withNumbers, withPotentials) hold what it read
				Geometry_3d::Table geom;
				Grid_3d::Table grid;
				int wn = -1, wp = -1;
				File_Io::readGeomGrid((g_out + "/" + tok.at(1)).c_str(), &geom, &grid, &wn, &wp);
				const std::string tag = tok.at(2);
				std::vector<float> xyz, phi;
				std::vector<short> ids;
				std::vector<unsigned int> numbers;
				for (int i = 0; i < grid.numNodes(); i++)
				{
					for (int c = 0; c < 3; c++) xyz.push_back(geom.node(i).coords[c]);
					ids.push_back(grid.node(i).id().toValue());
					numbers.push_back(grid.node(i).number().toValue());
					phi.push_back((float) grid.node(i).phi(0));
				}
				std::vector<int> flags;
				flags.push_back(wn);
				flags.push_back(wp);
				writeArray(tag + "_xyz.f32", xyz);
				writeArray(tag + "_ids.i16", ids);
				writeArray(tag + "_numbers.u32", numbers);
				writeArray(tag + "_phi.f32", phi);
				writeArray(tag + "_flags.i32", flags);
			}
			else if (cmd == "writeresults")
			{
				// writeresults:IN:OUT:MODE -- File_Io::LogResults (MODE 0 ASCII,
				// 1 BINARY) writes the events of IN (27 floats each: index, id,
				// number, voltage, start xyz, stop xyz, stop time, time scale,
				// probability, potential and field before, potential and field
				// after, normal, apex) into OUT of the work directory
				const std::vector<float> rec = readFloats(tok.at(1));
				const int K = (int) (rec.size() / 27);
				{
					File_Io::LogResults log((g_out + "/" + tok.at(2)).c_str(), std::atoi(tok.at(3).c_str()));
					for (int e = 0; e < K; e++)
					{
						const float* r = &rec[27 * e];
						File_Io::LogResults::Dataset d;
						d.simTime = 0.0f;
						d.eventIndex = (unsigned int) r[0];
						d.nodeIndex = 0;
						d.nodeId = Configuration::NodeId::fromValue((short) r[1]);
						d.nodeNumber = Configuration::NodeNumber::fromValue((unsigned int) r[2]);
						d.voltage = r[3];
						Trajectory_3d::phaseVector a(0.0f, MathVector3d<float>(r[4], r[5], r[6]), MathVector3d<float>(0.0f, 0.0f, 0.0f));
						Trajectory_3d::phaseVector b(r[10], MathVector3d<float>(r[7], r[8], r[9]), MathVector3d<float>(0.0f, 0.0f, 0.0f));
						const_cast<std::vector<Trajectory_3d::phaseVector>&>(d.trajectory.data()).push_back(a);
						const_cast<std::vector<Trajectory_3d::phaseVector>&>(d.trajectory.data()).push_back(b);
						d.timeScale = r[11];
						d.probability = r[12];
						d.potentialBefore = r[13];
						d.fieldBefore = MathVector3d<float>(r[14], r[15], r[16]);
						d.potentialAfter = r[17];
						d.fieldAfter = MathVector3d<float>(r[18], r[19], r[20]);
						d.normal = Geometry_3d::Point(r[21], r[22], r[23]);
						d.apex = Geometry_3d::Point(r[24], r[25], r[26]);
						log.push(d);
					}
					// the writer thread empties its queue only while it runs; give
					// it time before the destructor stops it
					usleep(300000);
				}
			}
			else if (cmd == "writetraj")
			{
				// writetraj:IN:OUT:MODE:CHUNK -- File_Io::LogTrajectory (MODE 0
				// ASCII, 1 BINARY; CHUNK events per file, 0 one file) writes the
				// datasets of IN into OUT of the work directory.  A dataset in IN
				// (floats): event index, node number, time scale, status, charge
				// (C), mass (kg), phase count P, error estimate (position xyz,
				// velocity xyz), then P phases of t, position xyz, velocity xyz,
				// tet
				const std::vector<float> rec = readFloats(tok.at(1));
				{
					// (name, mode, delay 0, chunk size)
					File_Io::LogTrajectory log((g_out + "/" + tok.at(2)).c_str(), std::atoi(tok.at(3).c_str()), 0,
					                           (unsigned int) std::atoi(tok.at(4).c_str()));
					size_t at = 0;
					while (at + 13 <= rec.size())
					{
						const float* h = &rec[at];
						Trajectory_3d t;
						t._charge = h[4];
						t._mass = h[5];
						t._status = (int) h[3];
						const int P = (int) h[6];
						t._error_estimate = Trajectory_3d::phaseVector(0.0f, MathVector3d<float>(h[7], h[8], h[9]),
						                                              MathVector3d<float>(h[10], h[11], h[12]));
						at += 13;
						t._data.clear();
						for (int k = 0; k < P; k++, at += 8)
						{
							const float* q = &rec[at];
							t._data.push_back(Trajectory_3d::phaseVector(q[0], MathVector3d<float>(q[1], q[2], q[3]),
							                                             MathVector3d<float>(q[4], q[5], q[6]), (int) q[7]));
						}
						log.push((unsigned int) h[0], (unsigned int) h[1], h[2], t);
					}
					// the writer thread empties its queue only while it runs; give
					// it time before the destructor stops it
					usleep(300000);
				}
			}
			else if (cmd == "writedump")
			{
				// writedump:NAME:INDEX -- File_Io::writeSystem of the current
				// system into NAME in the work directory
				File_Io::writeSystem((g_out + "/" + tok.at(1)).c_str(), system,
				                     (unsigned int) std::atoi(tok.at(2).c_str()));
			}
			else if (cmd == "writeconfig")
			{
				// writeconfig:NAME -- File_Io::writeConfig (ASCII) of the
				// configuration table the driver read
				File_Io::writeConfig((g_out + "/" + tok.at(1)).c_str(), system.configTable, File_Io::ASCII);
			}
			else if (cmd == "coord")
			{
				// the fork's update_coordination (surface_3d.cpp:568-593); it
				// prints every new id, which goes to the driver's stdout
				update_coordination(&system);
			}
			else if (cmd == "fieldlines")
			{
				// fieldlines:TAG:DISTANCE:ITERATIONS:STEPPER with the start points
				// (3 floats each) in TAG_start.f32 of the work directory; each line
				// is init() + compute(distance) in turn (one rand() stream)
				const std::string tag = tok.at(1);
				const float distance = (float) std::atof(tok.at(2).c_str());
				const unsigned long iters = std::strtoul(tok.at(3).c_str(), 0, 10);
				const unsigned long steps = std::strtoul(tok.at(4).c_str(), 0, 10);
				std::vector<float> start = readFloats(tag + "_start.f32");
				std::vector<float> data;
				std::vector<int> info;
				for (size_t l = 0; l + 2 < start.size(); l += 3)
				{
					Fieldline_3d f(&system.geomTable, &system.gridTable);
					f._stepperLimit = steps;
					f._iterationLimit = iters;
					f.init(MathVector3d<float>(start[l], start[l + 1], start[l + 2]));
					const int status = f.compute(distance);
					info.push_back(status);
					info.push_back((int) f.data().size());
					for (size_t k = 0; k < f.data().size(); k++)
					{
						const Fieldline_3d::phaseVector& p = f.data()[k];
						for (int c = 0; c < 3; c++) data.push_back(p.position()[c]);
						for (int c = 0; c < 3; c++) data.push_back(p.field()[c]);
						data.push_back((float) p.tetIndex());
					}
				}
				std::vector<int> nextRand(1, std::rand());
				writeArray("fieldlines_" + tag + ".f32", data);
				writeArray("fieldlines_" + tag + ".i32", info);
				writeArray("fieldlines_next_rand_" + tag + ".i32", nextRand);
			}
			else
				throw std::runtime_error("unknown step " + cmd);
			struct timespec stepEnd;
			clock_gettime(CLOCK_MONOTONIC, &stepEnd);
			stepSeconds.push_back((double) (stepEnd.tv_sec - stepStart.tv_sec) +
			                      1e-9 * (double) (stepEnd.tv_nsec - stepStart.tv_nsec));
			writeArray("step_seconds.f64", stepSeconds);
		}
	}
	catch (std::exception& e)
	{
		std::fprintf(stderr, "ref_driver: %s\n", e.what());
		return 1;
	}
	return 0;
}
