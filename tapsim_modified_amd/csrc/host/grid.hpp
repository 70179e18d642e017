// Host side of the device grid: the counterpart of Grid_3d::Table
// (tapsim/grid_3d.h:100-295) together with the surface table and the
// per-event loop of Process::evaporation that drive it.  Hot data lives on
// the device; the host keeps the small mirrors needed to patch it when an
// atom is removed (flags, ids, couplings, row layout, surface set).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <set>
#include <string>
#include <memory>
#include <vector>

#include "../kernels/kernels.hpp"
#include "mesh.hpp"
#include "prand.hpp"

namespace tsg {

class MpnnNet;

struct Species {
	int defined = 0;
	float charge_density = 0, epsilon = 1;
	int removable = 0, neumann = 0, dirichlet = 0;
	float phi = 0, mass = 0;
	int evap_charge = 0;
	float evap_field = 0, evap_energy = 0;
};

// one evaporation event as recorded by File_Io::LogResults::Dataset
// (tapsim/logResults.h:41-66) plus the relaxation counters
struct EventRecord {
	int node, id;
	float probability;
	double pot_before, field_before[3];
	double pot_after, field_after[3];
	long long local_iter, global_iter;
	int surface_size;
	double scaling_ref;
	int near_tie;
	// learned warm start: 0 none, 1 prediction applied, 2 prediction rejected
	// (it raised the residual), 3 graph recorded only; the max |residual|
	// over the shell before and (when predicted) over the grid after it
	int warm_start;
	double residual_cold, residual_warm;
	// steps V and VI (process.cpp:940-1037): the ion's flight and the
	// reconstruction quantities; apex = highest surface site before the event,
	// normal = the surface normal at the evaporated site (Surface_3d::normal)
	float start[3], stop[3], apex[3], normal[3];
	float voltage, time_scale, tof;
	int traj_len, traj_status;
	long long event_index;  // LogResults::Dataset::eventIndex (process.cpp:696)
	float ion_mass;         // the ion's mass (amu) and charge (e) in its flight
	int ion_charge;
	float err_pos[3], err_vel[3];  // the flight's final error estimate (Trajectory_3d::error_estimate)
};

struct EvapOptions {
	int prob_mode = 0, evap_mode = 0, vacuum_id = 0;
	double relax_threshold = 1e-2;
	int relax_shell = 3, relax_cycle = 5, relax_queue = 5;
	int relax_global_cycles = 1;
	int refresh_interval = 0;
	double refresh_threshold = 0.0;
	int refresh_cycle = 1, refresh_queue = 0;
	int exact_probability = 1;  // compute the sequential normaliser every event
	int method = 0;             // 0: the reference's relaxation, 1: accelerated (mixed precision),
	                            // 2: accelerated from the learned warm start
	double accel_tol = 1e-8;    // scaled-residual target of the accelerated relaxation
	int accel_solver = 0;       // correction solve: 0 Chebyshev-Jacobi, 1 multigrid-preconditioned CG
	// step V: the ion trajectory (EvaporationOptions, process.h:39-122)
	int trajectory = 0, integrator = 1, stepper = 1;
	float dt0 = 1e-14f, dt_min = 1e-16f, err_limit = 1e-6f;
	int voltage_queue = 250;
	float shrink_limit = 1.0f;  // stop once the apex has receded this far (m)
	int thermal_velocity = 0;   // 1: evap_initialVelocity from the process's rand()
	int update_coordination = 0;  // 1: the fork's update_coordination (surface_3d.cpp:568-593)
	int init_event_cnt = 0;       // EvaporationOptions::initEventCnt (process.h:98)
	// additions without a reference counterpart (integrator 2 is the classical
	// fourth-order Runge-Kutta step)
	float detector_z = 0.0f;      // > 0: extrapolate each flight to this plane (m), else to the mesh's top
	int ion_table = 0;            // 1: each ion's mass and charge drawn from its species' ion table
	unsigned int ion_seed = 0;    // seed of those draws (one per event, keyed by the event index)
};

// a species' ion table: (mass amu, charge e, weight) entries
constexpr int kIonTableMax = 8;
struct IonTable {
	int count = 0;
	float mass[kIonTableMax];
	int charge[kIonTableMax];
	float weight[kIonTableMax];
};
// the entry drawn for an event: u in [0, 1) from splitmix64 of (seed, event
// index), then the first entry whose cumulative weight exceeds u times the
// total
int ion_pick(unsigned int seed, long long event_index, const IonTable& t);

class DevBuf {
public:
	DevBuf() = default;
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	~DevBuf();
	void alloc(size_t bytes);
	void free();
	template <typename T> T* as() const { return (T*) p_; }
	void* ptr() const { return p_; }
	size_t bytes() const { return n_; }

private:
	void* p_ = nullptr;
	size_t n_ = 0;
};

class Grid {
public:
	Grid(int mode, const Mesh& mesh, const int16_t* ids, const Species* species, float temperature);
	~Grid();

	int status() const { return status_; }
	const std::string& error() const { return err_; }
	bool f64() const { return f64_; }
	int n() const { return mesh_.n; }
	long long nnz() const { return (long long) mesh_.idx.size(); }
	int nrows() const { return nrows_; }

	// Grid_3d::Table::fastSync (grid_3d.cpp:491-553); runs at construction
	int sync();
	// Grid_3d::Table::fastResync after setId (grid_3d.cpp:555-589)
	int set_id_and_resync(int node, int id);

	// Grid_3d::Table::relax(cycleSize) (grid_3d.cpp:591-609)
	int relax_cycles(int cycles, double* deviation);
	// Grid_3d::Table::relax(threshold, cycleSize, queueSize) (grid_3d.cpp:611-671)
	int relax_threshold(double threshold, int cycles, int queue, long long* iterations, double* fault_dev,
	                    double* fault_slope, double* last_dev);
	// Grid_3d::Table::localRelax(index, order, threshold, cycleSize, queueSize) (grid_3d.cpp:707-795)
	int local_relax(int node, int order, double threshold, int cycles, int queue, long long* iterations);

	// initial potentials: species presets (Grid_3d::Table::reset), random
	// (randomReset, from the process's rand()), radial ramp
	// (preset_sphericalPotential); both buffers, switches kept
	int reset_potentials();
	int random_potentials();
	int spherical_potentials(const float center[3], float pmin, float pmax);
	int get_phi(int buffer, double* out);
	int set_phi(const double* phi0, const double* phi1, int phi_switch, int local_switch);
	int phi_switch() const { return sw_; }
	int local_switch() const { return lsw_; }
	int potential(int node, double* out);

	int field_o1(int count, const int* nodes, double* out);
	int field_o2(int count, const int* nodes, double* out, double* charge);
	// field_o1 through caller-provided device staging (count node indices,
	// 3 count doubles, one status int) instead of the grid's own buffers
	int field_o1_staged(int count, const int* nodes, double* out, int* d_nodes, double* d_fields, int* d_status);
	// the device the grid was created on (its streams and buffers live there)
	int device() const { return device_; }
	int get_couplings(double* out) const;
	int get_charges(double* out) const;
	// the fork's update_coordination (surface_3d.cpp:568-593) on the ids
	int update_coordination();
	// Geometry_3d::Table::min() / max() and the constrained extents
	const float* extents() const { return extents_; }
	// the ion table of a species for the flights (no reference counterpart)
	int set_ion_table(int species, int count, const float* mass, const int* charge, const float* weight);
	// the node's id names a defined species (Configuration::Table::operator[]
	// throws otherwise)
	bool species_ok(int node) const
	{
		const int id = ids_[node];
		return id >= 0 && id < 128 && species_[id].defined;
	}
	int get_flags(unsigned char* out) const;
	int get_ids(int16_t* out) const;
	int get_areas(float* a_half32, double* a_half64, float* a_one32) const;

	// surface (Surface_3d::Table, surface_3d.cpp:50-119) and selection
	int surface_init(int vacuum_id);
	int surface_size() const { return surf_count_; }
	int surface_nodes(int* out);
	int surface_normal(int node, float out[3]);  // Surface_3d::Table::normal (host)
	// evap_compute_specificFields + evap_compute_probabilities (mode 0 linear
	// field, 1 Boltzmann, 2 linear force, 3 Voronoi flux force), the choice,
	// the apex and, for prev >= 0, potential() and field_o1 of node prev
	int surface_evaluate(bool exact_sum, double* scaling_ref, int mode = 0, int prev = -1);
	int surface_values(double* out);                  // |E|/F per site, site order
	int surface_probabilities(float* out);            // normalised as the reference
	// evapMaximum: node, and its position in the site list
	int surface_candidate(int* node, int* near_tie, int* slot = nullptr);
	int mc_candidate(int* node, int* slot = nullptr);  // evapMonteCarlo
	// evap_takeAway (surface_3d.cpp:507-564)
	int take_away(int node);

	// accelerated relaxation: FP32 Chebyshev-Jacobi correction solves inside
	// FP64 residual refinement, to max|J(phi)-phi| <= tol
	int accel_setup(double safety);
	int accel_relax(double tol, int* outer, long long* inner, double* final_residual);
	double accel_rho() const { return rho_; }
	// aggregation multigrid for the correction solve (built on first use;
	// rows created later are smoothed but not aggregated)
	int amg_setup();
	// (eager: every kernel launched on its own instead of replaying the
	// captured iteration graph; the same kernels in the same order)
	void amg_select(bool on, bool eager = false) { amg_want_ = on; amg_eager_ = eager; }
	int amg_levels() const { return (int) amg_levels_.size(); }
	int time_kernel(int kind, int reps, double* avg_ms, double* bytes);
	// host wall time of the event loop's phases, accumulated over all events
	// (ms): choice (with the record's fields before), take-away, warm start,
	// relaxation, surface evaluation (with the fields after), flight hand-off
	// and collection
	enum { kPhaseChoice, kPhaseTakeAway, kPhaseWarm, kPhaseRelax, kPhaseSurface, kPhaseFlight, kPhases };
	int phase_times(double* ms, int n, long long* events) const
	{
		for (int p = 0; p < n && p < kPhases; p++) ms[p] = stat_phase_ms_[p];
		if (events) *events = stat_phase_events_;
		return 0;
	}

	// Process::evaporation steps I-IV and VI (process.cpp:680-1037) for k events
	int evaporate(const EvapOptions& o, int k, EventRecord* out, int* done, int phase_cap = 0, float* phases = nullptr,
	              int* phase_len = nullptr);

	// ion trajectories (Trajectory_3d, trajectory_3d.cpp:99-332) of a batch of
	// ions from their nodes in the current potentials; out holds cap phases
	// (t, x, y, z, vx, vy, vz, tet) per ion
	// v0: optional start velocities (3 per ion); process_rand: the walks
	// continue the process's rand() stream (one ion only)
	int trajectories(int count, const int* nodes, const float* charge, const float* mass, int integrator, int stepper,
	                 float dt0, float dt_min, float err_limit, int cap, float* out, int* len, int* status,
	                 float* summary = nullptr, const float* v0 = nullptr, bool process_rand = false,
	                 float* error_estimate = nullptr);
	// Grid_3d::potential(pos, guess): barycentric potential of a point; with
	// process_rand the point is located `walks` times drawing from the
	// process's rand() stream, as that many calls of the reference would
	int point_potential(const float pos[3], int guess, float* out, bool process_rand = false, int walks = 1);
	// Fieldline_3d::init + compute(distance) (fieldline_3d.cpp:55-190) for
	// `count` start points in the current potentials; out holds cap points of
	// (x, y, z, Ex, Ey, Ez, tet) per line
	int fieldlines(int count, const float* start, float distance, long long iterations, long long stepper_limit,
	               bool process_rand, int cap, float* out, int* len, int* status);
	// evap_initialVelocity (surface_3d.cpp:472-502) from the process's rand()
	void thermal_velocity(const Species& sp, float v[3]) const;

	// learned warm start: events of methods 1 and 2 record their shell graph
	// and relaxation target into the network's store (record), method 2
	// starts the relaxation from the network's prediction (predict)
	int attach_mpnn(MpnnNet* net, bool record, bool predict);

	hipStream_t stream() const { return stream_; }
	// algorithmic bytes per sweep and the accumulated sweep timing (HIP events
	// around the threshold relaxations on the grid's stream)
	int sweep_stats(double* bytes, long long* rows, long long* entries, long long* referenced, double* ms,
	                long long* sweeps, long long* launched) const;
	// device pointers for the benchmark's own timing of the sweep kernel
	SweepArgs sweep_args();

private:
	int fail(int code, const std::string& msg);
	int hip_check(hipError_t e, const char* where);
	double compute_coupling(int n1, int n2, long long entry) const;
	bool participates(int row, int col) const;
	void row_entries(int node, std::vector<int>* cols, std::vector<double>* coups, double* ws) const;
	int build_rows();
	int write_slot(int slot, int node, std::vector<long long>* ai, std::vector<int>* vi, std::vector<long long>* ar,
	               std::vector<double>* vr, std::vector<long long>* aw, std::vector<double>* vw);
	long long entry_index(int row, int col) const;
	int arm_ctl(double threshold, int cycles, int queue);
	// relax(cycleSize > 1): the in-place cycles' row order and one call's launches
	DevBuf d_ord_slot_;
	int ord_count_ = 0;
	bool ord_dirty_ = true;
	int ensure_inplace_order();
	int launch_relax_call(const SweepArgs& a, int s, int cycles);
	int resolve_candidate(int* node, int* near_tie, int* slot);

	int mode_;
	int device_ = 0;
	bool f64_;
	int status_ = 0;
	std::string err_;
	Mesh mesh_;
	std::vector<int16_t> ids_;
	std::vector<uint8_t> flags_;
	Species species_[128];
	IonTable ions_[128];
	float temperature_;
	std::vector<long long> row_of_entry_;  // row index of every directed entry (int64 for >2^31)
	std::vector<long long> mirror_;        // entry of (j,i) for entry (i,j)
	std::vector<float> a_half32_, a_one32_, dist32_;
	std::vector<double> a_half64_, dist64_;
	std::vector<double> coup_;             // host mirror (values exact in R)
	std::vector<double> charge_;
	bool coord_done_ = false;         // the run-start update_coordination has been applied
	// steps V-VI pipelined on a second stream: the flight of event e runs
	// while event e+1 relaxes; two slots, each with the field snapshot and
	// the potentials of its event, its inputs and outputs and two events
	struct FlightSlot {
		bool active = false;
		long long seq = 0;
		int ev = 0;
		float flen = 0.0f, pot_before = 0.0f, pot_after = 0.0f, evap_field = 0.0f, scaling = 0.0f;
		DevBuf nfield, phi, io;
		float* host = nullptr;  // pinned: inputs [0, 4), outputs [4, 30)
		hipEvent_t ready = nullptr, done = nullptr;
		// the flight's phases when the caller records trajectories
		DevBuf phases;
		float* phases_host = nullptr;  // pinned, phases_cap phases of 8 floats
		int phases_cap = 0;
	};
	FlightSlot fslot_[2];
	hipStream_t fstream_ = nullptr;
	bool rng_on_device_ = false;      // the flight stream holds the process's rand() stream
	long long flight_seq_ = 0;
	EventRecord* flight_out_ = nullptr;  // the records of the evaporate call in progress
	// trajectory recording of the evaporate call in progress (cap 0: off):
	// event e's phases go to rec_phases_ + 8 cap e, their count to rec_len_[e]
	float* rec_phases_ = nullptr;
	int* rec_len_ = nullptr;
	int rec_cap_ = 0;
	int flight_start(const EvapOptions& o, int ev, int node, const Species& sp, float flen, float pot_before,
	                 float pot_after, float scaling);
	int flight_finish(FlightSlot& f, EventRecord* out, const EvapOptions& o);
	int flight_drain(EventRecord* out, const EvapOptions& o);
	int evaporate_events(const EvapOptions& o, int k, EventRecord* out, int* done);
	std::vector<float> cell_vol32_;   // Voronoi cell volume closed at half the mean neighbour distance
	std::vector<double> cell_vol64_;
	double charge_of(int node) const;
	int sw_ = 0, lsw_ = 0;

	// rows (SELL-64)
	int nrows_ = 0, nslots_ = 0, cap_slots_ = 0, next_spare_ = 0, spare_width_ = 0;
	std::vector<int> slot_of_, slot_node_, slot_deg_, slice_width_;
	std::vector<long long> slice_off_;
	// slots whose row is longer than kLongRow (one wavefront each in the row kernels)
	std::set<int> long_set_;
	bool long_dirty_ = false;
	int long_cap_ = 0;
	DevBuf d_long_;
	void set_long(int slot, bool on);
	int upload_long();
	// contiguous copies of the long rows (FP64 / faithful rows and FP32
	// correction rows share the offsets): per slot its offset (-1: none) and
	// capacity, bump-allocated from a store with spare room for new long rows
	std::vector<long long> loff_;
	std::vector<int> lcap_;
	long long lnext_ = 0, lstore_cap_ = 0;
	DevBuf d_long_off_, d_lcol_, d_lcoup_, d_licol_, d_liw_;
	// patches of the contiguous copies collected while rows are rewritten:
	// int (target 0 columns, 1 correction columns), real (couplings), float
	// (correction weights)
	std::vector<long long> lpi_addr_, lpr_addr_, lpf_addr_;
	std::vector<int> lpi_val_;
	std::vector<double> lpr_val_;
	std::vector<float> lpf_val_;
	int long_storage(int slot, int deg);

	// the surface set (Surface_3d::Table's nodes): the host mirror of the
	// device state bytes (bit 0 site, bit 1 vacuum), the site count, the
	// nodes whose byte changed since the last upload; the ascending site list
	// itself is rebuilt on the device and copied to surface_vec_ only where
	// the host needs it (Monte Carlo choice, exact tie resolution)
	int vacuum_id_ = -1;
	std::vector<uint8_t> sstate_;
	int surf_count_ = 0;
	std::vector<int> sst_pending_;
	bool list_dirty_ = false;        // a site bit changed: the list is rebuilt at the next flush
	bool host_list_stale_ = true;    // surface_vec_ no longer matches the device list
	bool eval_stale_ = true;         // the sites changed since the last evaluation
	std::vector<int> surface_vec_;
	double scaling_ref_ = 1.0;
	DevBuf d_sstate_, d_surf_blk_, d_sst_upd_, d_marks_, d_nshell_, d_choice_;
	unsigned char* sst_host_ = nullptr;  // pinned staging of the state updates
	size_t sst_host_bytes_ = 0;
	hipEvent_t sst_ev_ = nullptr;        // the last upload from that staging
	ChoiceRec* choice_host_ = nullptr;   // pinned readback of the evaluation's choice record
	bool choice_valid_ = false;          // choice_host_ describes the current sites and potentials
	int normal_stamp_ = 0;
	hipEvent_t norm_src_ = nullptr, norm_done_ = nullptr;
	bool normal_pending_ = false;         // a normal is in flight on the side stream
	int ensure_side_stream();
	void set_site(int node, bool on);
	void refresh_vacuum(int node);
	int surface_flush();
	int host_list();
	int launch_normal(int node);

	hipStream_t stream_ = nullptr;
	DevBuf d_xyz_, d_ptr_, d_idx_, d_coupfull_, d_flags_, d_evapF_, d_evapE_, d_eps_, d_phi0_, d_phi1_, d_surf_ef_;
	int prob_mode_ = 0;
	DevBuf d_slot_node_, d_slot_deg_, d_slice_off_, d_col_, d_coup_, d_ws_, d_q_, d_ctl_, d_block_max_, d_block_flag_;
	DevBuf d_surf_, d_surf_val_, d_sum_, d_status_;
	DevBuf d_shell_nodes_, d_shell_slots_, d_level_ptr_, d_order_, d_local_out_;
	DevBuf d_patch_addr_, d_patch_i_, d_patch_r_, d_nodes_, d_fields_;
	void* h_pinned_ = nullptr;
	size_t h_pinned_bytes_ = 0;
	hipEvent_t patch_ev_ = nullptr;       // the last upload from h_pinned_
	hipEvent_t ev_start_ = nullptr, ev_stop_ = nullptr;
	// accelerated relaxation state
	bool accel_ = false;
	double rho_ = 0.0;
	std::vector<int> inner_deg_;
	DevBuf d_inner_deg_, d_icol_, d_iw_, d_d0_, d_d1_, d_r32_, d_redout_;
	int write_inner(int slot, int node, std::vector<long long>* ai, std::vector<int>* vi, std::vector<long long>* af,
	                std::vector<float>* vf);
	AccelArgs accel_args() const;
	// multigrid state: coarse levels (CSR), level-0 row sums, aggregate of
	// every level-0 node, the coarsest level's dense inverse, PCG vectors
	struct AmgLevelHost {
		int n = 0;
		float w = 0.0f;
		DevBuf ptr, col, val, dinv, rptr, ridx, agg, b, x, r;
	};
	static constexpr int kAmgCoarsest = 1024;
	bool amg_want_ = false, amg_on_ = false;
	float amg_w0_ = 0.0f;
	std::vector<std::unique_ptr<AmgLevelHost>> amg_levels_;
	DevBuf amg_dn_, amg_agg0_, amg_minv_, amg_part_, amg_st_;
	DevBuf amg_r_, amg_z_, amg_p_, amg_q_, amg_x0_, amg_res_, amg_delta_;
	bool amg_eager_ = false;
	// the captured PCG iterations (kAmgBatch per replay) for the current rows
	static constexpr int kAmgBatch = 2;
	static constexpr int kAmgMaxIters = 160;
	hipGraph_t amg_graph_ = nullptr;
	hipGraphExec_t amg_exec_ = nullptr;
	AccelArgs amg_graph_args_{};  // the row arguments the graph was captured with
	double* amg_host_ = nullptr;  // pinned readback of the scalars, two batches
	hipEvent_t amg_ev_[2] = {nullptr, nullptr};
	void amg_graph_drop();
	// the level-0 row sums of the touched rows (slot, value) and the slots
	// whose residual entry is cleared, shipped with the row patches
	void amg_touch(const std::vector<int>& nodes, const std::vector<int>& off_slots, std::vector<long long>* dn_addr,
	               std::vector<float>* dn_val, std::vector<long long>* res_addr);
	int amg_solve(const AccelArgs& x, float f, int* iters, float** delta);
	// trajectories: mesh arrays for the point location, node-field snapshot
	DevBuf d_tv_, d_tn_, d_assoc_, d_gbnd_, d_nfield_, d_ppos_, d_traj_in_, d_traj_res_, d_traj_out_, d_rng_;
	float zmax_ = 0.0f;
	float extents_[12] = {};  // Geometry_3d::Table min, max, constrained min, constrained max
	int ensure_point_location();
	// Process::evaporation's voltage queue (process.cpp:983-1037) and initial apex
	std::vector<float> vq_;
	size_t vq_head_ = 0;
	float vq_sum_ = 0.0f;
	float apex0_ = 0.0f;
	bool apex0_set_ = false;
	// Process::evaporation's eventCnt (process.cpp:676-696), kept across
	// evaporate calls; -1 until the first call sets it from initEventCnt
	long long event_cnt_ = -1;
	// warm start state of the current event
	MpnnNet* net_ = nullptr;
	bool net_record_ = false, net_predict_ = false;
	DevBuf d_w_nodes_, d_w_slots_, d_w_hop_, d_w_wv_, d_w_pre_, d_w_res_, d_w_scale_;
	int warm_cap_ = 0;
	std::vector<int> warm_local_;  // node -> shell index while a shell is built, else -1
	void* warm_host_ = nullptr;    // pinned staging of the shell's uploads and readbacks
	size_t warm_host_bytes_ = 0;
	// the FP64 residual of the current potentials is in d_r32_ (max r32_rmax_):
	// the warm start's check formed it, the next relaxation starts from it
	// the surface normaliser on its side stream (surface_evaluate)
	hipStream_t sstream_ = nullptr;
	hipEvent_t sum_src_ = nullptr, sum_done_ = nullptr;
	double* sum_host_ = nullptr;
	bool sum_pending_ = false;
	bool r32_ready_ = false;
	double r32_rmax_ = 0.0;
	WarmArgs warm_{};
	long long warm_row0_ = 0, warm_edges_ = 0;
	int warm_unknown_ = 0;
	int warm_begin(int node, double phi_ref, bool predict, EventRecord* rec);
	int warm_end();
	double stat_phase_ms_[kPhases] = {};
	long long stat_phase_events_ = 0;
	double stat_accel_ms_ = 0.0;
	long long stat_cheb_ = 0, stat_resid_ = 0, stat_rho_raised_ = 0;
	double stat_sweep_ms_ = 0.0;
	long long stat_sweeps_ = 0, stat_launched_ = 0, last_sweeps_ = 0;
};

}  // namespace tsg
