// C ABI over the device grid (include/tapsim_gpu.h).
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "../../../include/tapsim_gpu.h"
#include "../host/grid.hpp"
#include "../host/prand.hpp"
#include "errors.hpp"

struct tsg_grid {
	std::unique_ptr<tsg::Grid> g;
};

tsg::Grid* tsg_grid_unwrap(tsg_grid* h) { return h->g.get(); }

namespace {
static_assert(sizeof(tsg_species) == sizeof(tsg::Species), "species layout");
static_assert(sizeof(tsg_event) == sizeof(tsg::EventRecord), "event layout");

int check(tsg_grid* h, int rc)
{
	if (rc) tsg::set_error(h->g->error());
	return rc;
}
}  // namespace

extern "C" {

int tsg_grid_create(const tsg_mesh* mesh, const short* ids, const tsg_species* species, float temperature, int mode,
                    tsg_grid** out)
{
	if (!mesh || !ids || !species || !out) return tsg::fail(TSG_E_ARG, "tsg_grid_create: null argument");
	if (mode != TSG_PLAIN_F64 && mode != TSG_FAITHFUL_F32) return tsg::fail(TSG_E_ARG, "tsg_grid_create: unknown mode");
	int dev = 0;
	if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) return tsg::fail(TSG_E_HIP, "tsg_grid_create: no HIP device");
	try {
		tsg::Species sp[128];
		std::memcpy(sp, species, sizeof(sp));
		auto* h = new tsg_grid();
		h->g.reset(new tsg::Grid(mode, mesh->m, ids, sp, temperature));
		if (h->g->status()) {
			const int rc = h->g->status();
			tsg::set_error(h->g->error());
			delete h;
			return rc;
		}
		*out = h;
		return TSG_OK;
	} catch (const std::exception& e) {
		return tsg::fail(TSG_E_HIP, std::string("tsg_grid_create: ") + e.what());
	}
}

void tsg_grid_destroy(tsg_grid* h) { delete h; }

int tsg_grid_info(const tsg_grid* h, int* n, long long* nnz, int* nrows)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	if (n) *n = h->g->n();
	if (nnz) *nnz = h->g->nnz();
	if (nrows) *nrows = h->g->nrows();
	return TSG_OK;
}

int tsg_relax_cycles(tsg_grid* h, int cycles, double* deviation)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->relax_cycles(cycles, deviation));
}

int tsg_relax_threshold(tsg_grid* h, double threshold, int cycles, int queue, long long* iterations, double* fault_dev,
                        double* fault_slope, double* last_dev)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->relax_threshold(threshold, cycles, queue, iterations, fault_dev, fault_slope, last_dev));
}

int tsg_local_relax(tsg_grid* h, int node, int order, double threshold, int cycles, int queue, long long* iterations)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->local_relax(node, order, threshold, cycles, queue, iterations));
}

int tsg_get_phi(tsg_grid* h, int buffer, double* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->get_phi(buffer, out));
}

int tsg_set_phi(tsg_grid* h, const double* p0, const double* p1, int sw, int lsw)
{
	if (!h || !p0 || !p1) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->set_phi(p0, p1, sw, lsw));
}

int tsg_phi_switches(const tsg_grid* h, int* sw, int* lsw)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	if (sw) *sw = h->g->phi_switch();
	if (lsw) *lsw = h->g->local_switch();
	return TSG_OK;
}

int tsg_potential(tsg_grid* h, int node, double* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->potential(node, out));
}

int tsg_field_o1(tsg_grid* h, int count, const int* nodes, double* out)
{
	if (!h || (count > 0 && (!nodes || !out))) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->field_o1(count, nodes, out));
}

int tsg_field_o2(tsg_grid* h, int count, const int* nodes, double* out, double* charge)
{
	if (!h || (count > 0 && (!nodes || !out || !charge))) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->field_o2(count, nodes, out, charge));
}

int tsg_get_couplings(const tsg_grid* h, double* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return h->g->get_couplings(out);
}

int tsg_get_charges(const tsg_grid* h, double* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return h->g->get_charges(out);
}

int tsg_get_flags(const tsg_grid* h, unsigned char* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return h->g->get_flags(out);
}

int tsg_get_ids(const tsg_grid* h, short* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return h->g->get_ids((int16_t*) out);
}

int tsg_get_areas(const tsg_grid* h, float* half32, double* half64, float* one32)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null argument");
	return h->g->get_areas(half32, half64, one32);
}

int tsg_set_id(tsg_grid* h, int node, int id)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->set_id_and_resync(node, id));
}

int tsg_surface_init(tsg_grid* h, int vacuum_id)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->surface_init(vacuum_id));
}

int tsg_surface_size(const tsg_grid* h, int* count)
{
	if (!h || !count) return tsg::fail(TSG_E_ARG, "null argument");
	*count = h->g->surface_size();
	return TSG_OK;
}

int tsg_surface_nodes(const tsg_grid* h, int* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	const int rc = h->g->surface_nodes(out);
	if (rc) tsg::set_error(h->g->error());
	return rc;
}

int tsg_fieldlines(tsg_grid* h, int count, const float* start, float distance, long long iterations,
                   long long stepper_limit, int process_rand, int cap, float* out, int* len, int* status)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->fieldlines(count, start, distance, iterations, stepper_limit, process_rand != 0, cap, out,
	                                 len, status));
}

int tsg_update_coordination(tsg_grid* h)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->update_coordination());
}

int tsg_surface_normal(const tsg_grid* h, int node, float out[3])
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	const int rc = h->g->surface_normal(node, out);
	if (rc) tsg::set_error(h->g->error());
	return rc;
}

int tsg_surface_evaluate_model(tsg_grid* h, int mode, int exact_sum, double* scaling_ref)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->surface_evaluate(exact_sum != 0, scaling_ref, mode));
}

void tsg_srand(unsigned int seed) { tsg::rand_seed(&tsg::process_rand(), seed); }

int tsg_rand(int count, int* out)
{
	if (count < 0 || (count > 0 && !out)) return tsg::fail(TSG_E_ARG, "bad arguments");
	for (int i = 0; i < count; i++) out[i] = tsg::prand();
	return TSG_OK;
}

int tsg_surface_evaluate(tsg_grid* h, int exact_sum, double* scaling_ref)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->surface_evaluate(exact_sum != 0, scaling_ref));
}

int tsg_surface_values(tsg_grid* h, double* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->surface_values(out));
}

int tsg_surface_probabilities(tsg_grid* h, float* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->surface_probabilities(out));
}

int tsg_surface_candidate(tsg_grid* h, int* node, int* near_tie)
{
	if (!h || !node) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->surface_candidate(node, near_tie));
}

int tsg_take_away(tsg_grid* h, int node)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->take_away(node));
}

namespace {

tsg::EvapOptions options_of(const tsg_evap_options* o)
{
	tsg::EvapOptions e;
	e.prob_mode = o->prob_mode; e.evap_mode = o->evap_mode; e.vacuum_id = o->vacuum_id;
	e.relax_threshold = o->relax_threshold; e.relax_shell = o->relax_shell; e.relax_cycle = o->relax_cycle;
	e.relax_queue = o->relax_queue; e.relax_global_cycles = o->relax_global_cycles;
	e.refresh_interval = o->refresh_interval; e.refresh_threshold = o->refresh_threshold;
	e.refresh_cycle = o->refresh_cycle; e.refresh_queue = o->refresh_queue;
	e.exact_probability = o->exact_probability;
	e.method = o->method;
	e.accel_tol = o->accel_tol;
	e.trajectory = o->trajectory;
	e.integrator = o->integrator;
	e.stepper = o->stepper;
	e.dt0 = o->dt0;
	e.dt_min = o->dt_min;
	e.err_limit = o->err_limit;
	e.voltage_queue = o->voltage_queue;
	e.shrink_limit = o->shrink_limit;
	e.thermal_velocity = o->thermal_velocity;
	e.update_coordination = o->update_coordination;
	e.accel_solver = o->accel_solver;
	e.init_event_cnt = o->init_event_cnt;
	e.detector_z = o->detector_z;
	e.ion_table = o->ion_table;
	e.ion_seed = o->ion_seed;
	return e;
}

}  // namespace

int tsg_evaporate(tsg_grid* h, const tsg_evap_options* o, int k, tsg_event* out, int* done)
{
	if (!h || !o || !out || !done) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->evaporate(options_of(o), k, reinterpret_cast<tsg::EventRecord*>(out), done));
}

int tsg_evaporate_traj(tsg_grid* h, const tsg_evap_options* o, int k, tsg_event* out, int* done, int cap, float* phases,
                       int* phase_len)
{
	if (!h || !o || !out || !done || cap < 0 || (cap > 0 && (!phases || !phase_len)))
		return tsg::fail(TSG_E_ARG, "tsg_evaporate_traj: bad arguments");
	return check(h, h->g->evaporate(options_of(o), k, reinterpret_cast<tsg::EventRecord*>(out), done, cap, phases,
	                                phase_len));
}

int tsg_grid_extents(tsg_grid* h, float* out)
{
	if (!h || !out) return tsg::fail(TSG_E_ARG, "null argument");
	for (int k = 0; k < 12; k++) out[k] = h->g->extents()[k];
	return 0;
}

int tsg_set_ion_table(tsg_grid* h, int species, int count, const float* mass_amu, const int* charge,
                      const float* weight)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->set_ion_table(species, count, mass_amu, charge, weight));
}

int tsg_trajectories(tsg_grid* h, int count, const int* nodes, const float* charge, const float* mass, int integrator,
                     int stepper, float dt0, float dt_min, float err_limit, int cap, float* out, int* len, int* status,
                     float* error_estimate)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return check(h, h->g->trajectories(count, nodes, charge, mass, integrator, stepper, dt0, dt_min, err_limit, cap, out,
	                                   len, status, nullptr, nullptr, false, error_estimate));
}

int tsg_accel_setup(tsg_grid* h, double safety, double* rho)
{
	if (!h || !(safety > 0.0 && safety <= 1.0)) return tsg::fail(TSG_E_ARG, "tsg_accel_setup: bad arguments");
	const int rc = check(h, h->g->accel_setup(safety));
	if (!rc && rho) *rho = h->g->accel_rho();
	return rc;
}

int tsg_accel_relax(tsg_grid* h, double tol, int* outer, long long* inner, double* final_residual)
{
	if (!h || !(tol > 0.0)) return tsg::fail(TSG_E_ARG, "tsg_accel_relax: bad arguments");
	return check(h, h->g->accel_relax(tol, outer, inner, final_residual));
}

int tsg_phase_times(tsg_grid* h, double* ms, int n, long long* events)
{
	if (!h || !ms || n < 0) return tsg::fail(TSG_E_ARG, "tsg_phase_times: bad arguments");
	return h->g->phase_times(ms, n, events);
}

int tsg_accel_multigrid(tsg_grid* h, int on, int* levels)
{
	if (!h || on < 0 || on > 2) return tsg::fail(TSG_E_ARG, "tsg_accel_multigrid: bad arguments");
	h->g->amg_select(on >= 1, on == 2);
	if (levels) *levels = h->g->amg_levels();
	return 0;
}

int tsg_reset_potentials(tsg_grid* h, int kind, const float* center, float pmin, float pmax)
{
	if (!h || kind < 0 || kind > 2 || (kind == 2 && !center)) return tsg::fail(TSG_E_ARG, "tsg_reset_potentials: bad arguments");
	if (kind == 0) return check(h, h->g->reset_potentials());
	if (kind == 1) return check(h, h->g->random_potentials());
	return check(h, h->g->spherical_potentials(center, pmin, pmax));
}

int tsg_copy_probe(long long bytes, int reps, double* avg_ms)
{
	if (bytes <= 0 || bytes % 16 || reps <= 0 || !avg_ms) return tsg::fail(TSG_E_ARG, "tsg_copy_probe: bad arguments");
	void *a = nullptr, *b = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	int rc = TSG_OK;
	if (hipMalloc(&a, (size_t) bytes) != hipSuccess || hipMalloc(&b, (size_t) bytes) != hipSuccess ||
	    hipMemset(a, 0, (size_t) bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
	    hipEventCreate(&e1) != hipSuccess)
		rc = tsg::fail(TSG_E_HIP, "tsg_copy_probe: allocation failed");
	float ms = 0.0f;
	if (!rc) {
		hipEventRecord(e0, nullptr);
		for (int r = 0; r < reps && !rc; r++)
			if (tsg::launch_copy_probe(nullptr, bytes, a, b)) rc = tsg::fail(TSG_E_HIP, "tsg_copy_probe: launch failed");
		hipEventRecord(e1, nullptr);
		if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
			rc = tsg::fail(TSG_E_HIP, "tsg_copy_probe: timing failed");
	}
	if (e0) hipEventDestroy(e0);
	if (e1) hipEventDestroy(e1);
	if (a) hipFree(a);
	if (b) hipFree(b);
	if (!rc) *avg_ms = ms / reps;
	return rc;
}

int tsg_time_kernel(tsg_grid* h, int kind, int reps, double* avg_ms, double* bytes)
{
	if (!h || !avg_ms) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->g->time_kernel(kind, reps, avg_ms, bytes));
}

int tsg_sweep_stats(const tsg_grid* h, double* bytes, long long* rows, long long* entries, long long* referenced,
                    double* ms, long long* sweeps, long long* launched)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null grid");
	return h->g->sweep_stats(bytes, rows, entries, referenced, ms, sweeps, launched);
}

}  // extern "C"
