// Host-visible launchers of the device kernels (one translation unit per
// kernel family; the host Grid calls only these).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "mpnn.hpp"

namespace tsg {

// ---------------------------------------------------------------- geometry
int launch_circumcentres(hipStream_t, int m, const float* xyz, const int* tv, double* cc64, float* cc32, int* status);
// Geometry_3d::Table::setConstrainedExtents (geometry_3d.cpp:1575-1610): the
// float circumcentres of the tets with a hull face are the first and last
// vertices of the open Voronoi faces; their per-axis min and max go to
// ext[0..2] and ext[3..5] as order-preserving integer keys (set ext to
// INT_MAX x3, INT_MIN x3 before)
int launch_hull_extents(hipStream_t, int m, const int* tn, const float* cc32, int* ext);
// the float of an order-preserving key
float float_of_key(int k);
// per node: the Voronoi cell volume closed at half the mean neighbour
// distance (computeCharge), float and double; needs dist32/dist64 of
// launch_edge_areas
int launch_cell_volumes(hipStream_t, int n, const float* xyz, const int* tv, const int* tn, const int* assoc,
                        const int* nt_ptr, const int* nt_idx, const int* ptr, const int* idx, const double* cc64,
                        const float* cc32, const float* dist32, const double* dist64, float* vol32, double* vol64,
                        int* status);
int launch_edge_areas(hipStream_t, int n, long long nnz, const float* xyz, const int* tv, const int* tn,
                      const int* assoc, const int* nt_ptr, const int* nt_idx, const int* ptr, const int* idx,
                      const int* row_of, const double* cc64, const float* cc32, float* a_half32, double* a_half64,
                      float* a_one32, float* dist32, double* dist64, int* status);

// ---------------------------------------------------------------- relaxation

// longest row taken by one lane in the row kernels; longer rows are listed
// separately and each is taken by a whole wavefront
constexpr int kLongRow = 24;

// Device-resident control block of a threshold relaxation
// (Grid_3d::Table::relax(threshold, cycleSize, queueSize), grid_3d.cpp:611-671).
struct RelaxCtl {
	int done;                 // 0 running, 1 converged, 2 RelaxationFault, 3 NaN
	int nblocks;              // blocks per sweep launch
	unsigned int counter;     // blocks finished in the current sweep
	int queue_size;
	unsigned long long devbits;  // running max |dphi| of the current sweep (double bits)
	long long iteration;      // iterations counted as the reference does
	long long sweeps;         // sweeps executed since the control block was armed
	double threshold;
	double last_dev;
	double fault_dev, fault_slope;
	int cycle;                // cycle size (1: every sweep is one relax(1) call)
	int qhead, qsize;
	int nan;                  // sticky: a cycle of the current relax(cycleSize) call produced a NaN
	double queue[1024];
};

struct SweepArgs {
	int nslots;
	const int* slot_node;
	const int* slot_deg;
	const long long* slice_off;
	const int* col;
	const void* coup;         // R per entry
	const double* ws;         // per slot
	const double* q;          // per slot
	void* phi0;               // R per node
	void* phi1;
	RelaxCtl* ctl;
	double* block_max;        // per-block partial max |dphi| of the current cycle
	int* block_flag;          // per-block NaN flag
	int threshold_mode;       // 1: use ctl convergence/early exit; 0: plain fixed sweeps
	const int* long_slots;    // slots whose row has more than kLongRow entries (one wavefront each), ascending
	int nlong;
	// the long rows' entries again, contiguous: long row li (long_slots[li])
	// has its columns and couplings at long_off[li] + k
	const long long* long_off;
	const int* lcol;
	const void* lcoup;        // R per entry
};
// workgroups of a row launch: 256 slots each, then one wavefront per long row
int row_launch_blocks(int nslots, int nlong);

// One Jacobi cycle (read buffer `read`, write the other) followed by the
// one-block decision kernel that reduces max|phi_new - phi_old| into ctl.
int launch_jacobi_sweep(hipStream_t, const SweepArgs&, int read, bool f64);
// A Jacobi cycle inside a relax(cycleSize) call: with `decide` the decision
// kernel follows (the call's last cycle), otherwise only the cycle runs.
int launch_jacobi_cycle(hipStream_t, const SweepArgs&, int read, bool f64, bool decide);

// The in-place cycles of relax(cycleSize > 1) (doRelax with the table's
// _phiSwitch held for the whole call, grid_3d.cpp:1238-1258,1458-1514):
// every even cycle updates buffer `buf` in place in ascending node order, as
// the reference's one-thread sweep.  One workgroup: the rows go in chunks
// whose known terms are staged in parallel, then the chain runs in order.
struct InplaceArgs {
	int count;                // unknown rows
	const int* ord_slot;      // their slots in ascending node order
	int buf;                  // the buffer updated in place (_phiSwitch of the call)
	int last;                 // 1: the call's last cycle: max |phi0 - phi1| into block_max[0], then the decision
};
// entries of one chunk staged in LDS; a row longer than this cannot be swept
constexpr int kInplaceCap = 4096;
int launch_inplace_cycle(hipStream_t, const SweepArgs&, const InplaceArgs&, bool f64);

// Local relaxation (Grid_3d::Table::localRelax(index, order, threshold,
// cycleSize, queueSize), grid_3d.cpp:707-795) of a node shell in one
// workgroup: the shell (ascending node index, Dirichlet removed) with its
// level schedule for the in-place cycles.
struct LocalArgs {
	int count;
	const int* nodes;         // shell nodes (ascending)
	const int* slots;         // their row slots
	const int* level_ptr;     // in-place levels: positions level_ptr[l]..level_ptr[l+1] of order[]
	const int* order;         // shell positions grouped by level
	int nlevels;
	int read_buf;             // _phiSwitch
	int local_switch;         // _phiLocalSwitch on entry
	int cycle, queue;
	double threshold;
	long long max_iterations;
	long long* out;           // [0] iterations (or -1-iteration on fault), [1] local switch after, [2] status
};
int launch_local_relax(hipStream_t, const SweepArgs&, const LocalArgs&, bool f64);

// dst[0..n) = value
int launch_fill_i32(hipStream_t, int n, int* dst, int value);
// one node's flags and species data (values passed with the launch)
struct NodeSpecies {
	int node;
	unsigned char flags;
	float evap_field, evap_energy, epsilon;
	unsigned char* d_flags;
	float *d_evap_field, *d_evap_energy, *d_epsilon;
};

// the patches of one take-away in one launch: job j writes dst_j[addr_j[k]]
// = val_j[k] for k < n_j, elements of `width` bytes (4 or 8), and the node's
// species values are written beside them; the jobs write distinct arrays,
// so their order does not matter
constexpr int kPatchJobs = 16;
struct PatchJob {
	int n = 0, width = 4;
	const long long* addr = nullptr;
	const void* val = nullptr;
	void* dst = nullptr;
};
struct PatchBatch {
	int count = 0;
	PatchJob job[kPatchJobs];
	NodeSpecies node;
};
int launch_patch_batch(hipStream_t, const PatchBatch&);

// ---------------------------------------------------------------- accelerated relaxation
// FP32 correction rows: the unknown neighbours of every row slot with the
// weights c_ij / sum_k c_ik, in the slot layout of the FP64 rows
struct AccelArgs {
	int nslots;
	const int* slot_node;
	const int* inner_deg;
	const long long* slice_off;
	const int* col;
	const float* w;
	const int* slot_deg;      // full row length: the long-row test
	const int* long_slots;
	int nlong;
	// the long rows' correction entries, contiguous at long_off[li] + k
	const long long* long_off;
	const int* licol;
	const float* liw;
};
// FP64 scaled residual r = J(phi) - phi into r32, max |r| into out[0], NaN flag into out[1]
int launch_resid64(hipStream_t, const SweepArgs&, const double* phi, float* r32, double* out);
int launch_cheb_init(hipStream_t, const AccelArgs&, const float* r32, float* dprev, float* dcur);
int launch_cheb_step(hipStream_t, const AccelArgs&, const float* r32, const float* dcur, float* dprev_next, float omega);
int launch_apply64(hipStream_t, const AccelArgs&, const float* delta, double* phi, double* phi_other);
int launch_apply_g(hipStream_t, const AccelArgs&, const float* in, float* out);

// ---------------------------------------------------------------- multigrid correction solve
// A coarse level of the aggregation multigrid (CSR, float): the operator, its
// inverse diagonal, the members of each row among the finer level's indices
// (restriction) and the map of the finer level's indices onto this level
// (prolongation; -1 outside every aggregate)
struct AmgLevel {
	int n;
	const int* ptr;
	const int* col;
	const float* val;
	const float* dinv;
	const int* rptr;
	const int* ridx;
	const int* agg;
};
// slots of the PCG scalars in device memory
// (kAmgDone: set once r.r <= kAmgTol2 * rr0; every later kernel of the
// solve returns at once; kAmgIters counts the completed iterations)
enum { kAmgAlpha = 0, kAmgBeta = 1, kAmgRz = 2, kAmgRr = 3, kAmgRr0 = 4, kAmgDone = 5, kAmgTol2 = 6, kAmgIters = 7,
       kAmgScalars = 8 };
// one PCG solve's device state: level 0 (the grid's FP32 correction rows,
// slot-indexed vectors), the coarse levels lv[0..nlev) with their Jacobi
// weights w and per-level vectors (right-hand side, iterate, scratch), the
// coarsest level's dense inverse, the block partials of the reductions, the
// scalars
struct AmgSolve {
	AccelArgs x;
	const float* dn;   // row coupling sums D
	const int* agg0;   // level-0 slot -> aggregate of lv[0]
	float w0;
	float wc;          // scale of every coarse correction (over-correction)
	const AmgLevel* lv;
	const float* w;
	int nlev;
	float *r, *z, *p, *q, *x0, *res, *delta;
	float* const* bl;
	float* const* xl;
	float* const* rl;
	const float* minv;
	double* part;
	double* st;
};
// r = D r32, delta = 0, x0 = w0 r / D (the V-cycle's pre-smoothing), rr0 = rr = r.r,
// the stopping target tol2 (on r.r / rr0), done = 0, iterations = 0
int launch_amg_start(hipStream_t, const AmgSolve&, const float* r32, double tol2);
// z = V(r); mode 3: beta = r.z / rz, rz = r.z; mode 4: rz = r.z
int launch_amg_vcycle(hipStream_t, const AmgSolve&, int mode);
// p = z + beta p (first: p = z)
int launch_amg_dir(hipStream_t, const AmgSolve&, int first);
// q = A p, alpha = rz / p.q, delta += alpha p, r -= alpha q, x0 = w0 r / D, rr = r.r,
// done once rr <= tol2 rr0 (every launch here is a no-op once done is set)
int launch_amg_matvec_update(hipStream_t, const AmgSolve&);
// q = A p and the block partials of p.q alone (the kernel's timing)
int launch_amg_fine_matvec(hipStream_t, const AmgSolve&);

// ---------------------------------------------------------------- warm start
// the shell of an evaporated atom as the network's graph (rows of the
// network's store) and the quantities that tie it to the grid
struct WarmArgs {
	int n;                    // shell nodes (ascending node index)
	const int* nodes;
	const int* slots;         // row slot, or -1 for a Dirichlet node
	const signed char* hop;   // BFS distance from the evaporated atom
	const float* wv;          // edge weight to the evaporated atom (0 if not adjacent)
	const float* xyz;
	float v[3];               // position of the evaporated atom
	float ell;                // length scale (mean distance to its neighbours)
	double phi_ref;           // its potential before the removal
	double* phi;              // the potential buffer the relaxation works on
	double* phi_pre;          // per shell node: potential before the warm start
	double* resid;            // per shell node: residual before the warm start
	double* scale;            // device scalar: max |residual| over the shell
	bf16_t* x;                // features (network store rows)
	float* t;                 // targets (network store rows)
	float* mask;              // loss weights: 1 on unknown rows
};
int launch_warm_features(hipStream_t, const WarmArgs&, const SweepArgs&);
int launch_warm_apply(hipStream_t, const WarmArgs&, const float* y);
int launch_warm_targets(hipStream_t, const WarmArgs&);

// ---------------------------------------------------------------- trajectories
// a batch of ions, one thread each; node fields from per-ion snapshots
struct TrajArgs {
	int count;
	const int* nodes;         // start node of each ion
	const int* snapshot;      // index of its field snapshot
	const float* charge;      // C
	const float* mass;        // kg
	int integrator, stepper;
	float dt0, dt_min, err_limit, zmax;
	int max_steps;            // accepted steps per ion
	long long max_calls;      // integrator calls per ion
	int n, m;
	const float* xyz;
	const int* tv;
	const int* tn;
	const int* assoc;
	const unsigned char* gbnd;
	const float* nfield;      // snapshots: 3 floats per node each
	float* out;               // per ion: cap phases of (t, p[3], v[3], tet)
	int cap;
	int* len;                 // phases produced (the output holds the first cap)
	int* status;              // Trajectory_3d status (trajectory_3d.h:89), negative on failure
	float* summary;           // optional, 12 floats per ion: start xyz, stop xyz, t of the last
	                          // phase, last phase inside the mesh (xyz, tet), 1 if there is one
	const float* v0;          // optional start velocities, 3 per ion (m/s); at rest when null
	int* rng;                 // optional rand() stream (31-word table + index, see host/prand.hpp):
	                          // the walks draw their first face from it; count must be 1
	float* err_out;           // optional, 6 floats per ion: the final error estimate
	                          // (Trajectory_3d::error_estimate: position xyz, velocity xyz)
};
// One event's flight (steps V-VI of Process::evaporation, process.cpp:940-1037)
// in one thread, on the process's rand() stream: the thermal start velocity
// (evap_initialVelocity, surface_3d.cpp:472-502) when asked, the trajectory,
// and the potential at its last point inside the mesh, located `exit_walks`
// times as the reference's repeated Grid_3d::potential calls do.
struct FlightArgs {
	TrajArgs t;               // count 1; t.rng required; t.v0 is set by the kernel when thermal
	int thermal;
	float temperature, mass_amu;
	float* v0;                // 3 floats: the start velocity (written when thermal)
	int exit_walks;           // 0: no exit potential
	const void* phi;          // potentials the exit point is interpolated from
	bool f64;
	float* exit;              // [0] the exit potential, [1] 1 when it was formed
};
int launch_flight(hipStream_t, const FlightArgs&);

// Grid_3d::potential(pos, guess) (grid_3d.cpp:1585-1602) of `count` points:
// barycentric interpolation of the potentials in float; 0 outside the mesh.
// With rng (count 1) the point is located `walks` times in a row, each walk
// drawing from the stream as the reference's repeated calls do.
int launch_point_potential(hipStream_t, int count, const float* pos, const int* guess, const float* xyz, const int* tv,
                           const int* tn, int m, const void* phi, bool f64, float* out, int* rng = nullptr,
                           int walks = 1);
int launch_trajectories(hipStream_t, const TrajArgs&);

// Fieldline_3d::init + compute(distance) (fieldline_3d.cpp:55-190) of
// `count` field lines through the node-field snapshot.  With rng the lines
// run one after another in one thread, continuing the process's rand()
// stream (tetGuess samples and walk faces) in the reference's order; without
// it one thread per line, tetGuess sampling evenly spaced tets and the walks
// starting at face 0 (the located tet differs only on a face).
struct FieldlineArgs {
	int count;
	const float* start;       // 3 floats per line
	float distance;           // < 0: no distance limit
	long long iterations, stepper_limit, max_calls;
	int guess_samples;        // tetGuess's sample count, (int) pow(m, 0.25)
	int m;
	const float* xyz;
	const int* tv;
	const int* tn;
	const unsigned char* gbnd;
	const float* nfield;      // field_o1 snapshot, 3 floats per node
	int cap;
	float* out;               // per line: cap points of (x, y, z, Ex, Ey, Ez, tet)
	int* len;                 // points of each line (the output holds the first cap)
	int* status;              // Fieldline_3d status; -10 start outside the mesh, -4 work limit
	int* rng;
};
int launch_fieldlines(hipStream_t, const FieldlineArgs&);
int launch_field_snapshot(hipStream_t, int n, const float* xyz, const int* ptr, const int* idx, const void* coup,
                          const unsigned char* flags, const void* phi, bool f64, float* out, int* status);

// ---------------------------------------------------------------- measurement
// dst = src over `bytes` (a multiple of 16): the counter calibration copy
int launch_copy_probe(hipStream_t, long long bytes, const void* src, void* dst);

// ---------------------------------------------------------------- surface
struct FieldArgs {
	int count;
	const int* nodes;         // surface nodes, ascending
	const float* xyz;
	const int* ptr;           // full node CSR (ascending neighbours)
	const int* idx;
	const void* coup;         // R per full-CSR entry
	const unsigned char* flags;
	const void* phi;          // potential() buffer
	const float* evap_field;  // per node
	void* value;              // R per surface slot: the probability model's unnormalised value
	int* status;
	// probability model (Surface_3d::evap_compute_probabilities, surface_3d.cpp:223-417):
	// 0 linear field |E|/F, 1 Boltzmann, 2 linear force (field_o2), 3 Voronoi flux force
	int mode;
	float temperature;        // K (mode 1)
	const float* evap_energy; // per node, eV (mode 1)
	const float* epsilon;     // per node (mode 3)
	void* value_ef;           // R per slot: |E|/F (specific fields) when mode != 0, else null
};
int launch_surface_fields(hipStream_t, const FieldArgs&, bool f64);

// ---------------------------------------------------------------- surface set
// The surface (Surface_3d::Table's node set) lives on the device as one state
// byte per node, bit 0 a surface site and bit 1 a vacuum node, and as the
// ascending list of the sites that a compaction rebuilds from the bytes.
enum { kSurfSite = 1, kSurfVacuum = 2 };
// sstate[nodes[i]] = vals[i] (distinct nodes)
int launch_sstate_scatter(hipStream_t, int count, const int* nodes, const unsigned char* vals, unsigned char* sstate);
// list = the surface sites in ascending order, *count their number; blk
// holds one int per 4096 nodes
int surface_compact_blocks(int n);
int launch_surface_compact(hipStream_t, int n, const unsigned char* sstate, int* blk, int* list, int* count);

// what the event loop reads of one surface evaluation, in one readback
struct ChoiceRec {
	double value;          // the first maximum's unnormalised value
	double pot;            // potential() of the chosen site
	double field[3];       // field_o1 of the chosen site
	double pot_prev;       // potential() and field_o1 of `prev` (the last evaporated atom)
	double field_prev[3];
	double scaling_ref;    // max |E|/F (evap_compute_specificFields)
	float apex[3];         // Surface_3d::Table::apex: the first highest site
	int node;              // the first maximum's node (-1: empty surface)
	int slot;              // and its position in the site list
	int near_tie;          // 1: a value below it may tie after normalisation, or a value is negative / NaN
	int count;             // sites in the device list
	int status;            // field status bits (1, 2, 4 as the fields kernel; 16: singular field_o1 of prev)
	// Surface_3d::Table::normal of the event's site (k_normal, written apart
	// from the choice): status 0 done, 1 NaN (find_best_plane), 2 shell too large
	float normal[3];
	int normal_status;
};
struct ChooseArgs {
	int count;
	const int* list;
	const void* value;     // R per slot
	const void* value_ef;  // R per slot, or null (then value)
	const float* xyz;
	const int* ptr;
	const int* idx;
	const void* coup;
	const unsigned char* flags;
	const void* phi;
	const int* status;     // the fields kernel's status word
	const int* dev_count;  // the compaction's count
	int prev;              // node whose potential and field follow the evaluation, or -1
	ChoiceRec* rec;
};
int launch_surface_choose(hipStream_t, const ChooseArgs&, bool f64);

// Surface_3d::Table::normal of one node on the device: the three-hop shell
// gathered level by level (marks[] holds the stamp of the call that last
// reached a node), put in ascending order by a compaction of the marks over
// the shell's node range, then the moments and the shared plane solution
// (plane.hpp) in one workgroup; kernels on one stream, one call after another
constexpr int kNormalShellCap = 1 << 20;
struct NormalCtl {
	int cnt;      // shell entries gathered
	int over;     // the shell outgrew kNormalShellCap
	int lo, hi;   // node range of the shell
	int end[4];   // end of each level's entries (end[0] = 0)
	int sorted;   // entries of the ascending list
	int skip;     // the node is not a surface site: zero normal
};
struct NormalArgs {
	int node, n;
	const int* ptr;
	const int* idx;
	const float* xyz;
	const unsigned char* sstate;
	int* marks;
	int stamp;             // != 0, new per call
	NormalCtl* ctl;
	int* shell;            // kNormalShellCap ints, gathered order
	int* sorted;           // kNormalShellCap ints, ascending
	int* blk;              // surface_compact_blocks(n) ints
	ChoiceRec* rec;        // normal, normal_status
};
// bytes of the device workspace of launch_surface_normal for n nodes
size_t surface_normal_workspace(int n);
int launch_surface_normal(hipStream_t, const NormalArgs&);

// sequential sum in slot order (the reference's float/double normaliser)
int launch_surface_sum(hipStream_t, int count, const void* value, bool f64, double* out);

// field_o1 of arbitrary nodes (one thread each)
int launch_node_fields(hipStream_t, int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                       const void* coup, const unsigned char* flags, const void* phi, bool f64, double* out, int* status);
// field_o2 of a node list: fields (3 per node) and charges
int launch_node_fields_o2(hipStream_t, int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                          const unsigned char* flags, const void* phi, bool f64, double* out, double* charge,
                          int* status);

}  // namespace tsg
