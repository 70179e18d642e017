/*
 * tapsim_gpu.h -- C ABI of the MI355X (gfx950) per-event electrostatics of
 * TAPSim (ashtonmv/tapsim_modified).
 *
 * Plain pointers and sizes only; every call returns an int status
 * (TSG_OK or a negative TSG_E_*), with tsg_last_error() describing the last
 * failure of the calling thread.  No exception crosses this boundary.  One
 * HIP stream per grid handle; calls on one handle are serialised by the
 * caller (the reference's Grid_3d::Table is not re-entrant either,
 * tapsim/grid_3d.cpp:1238-1282).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef TAPSIM_GPU_H
#define TAPSIM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define TSG_OK 0
#define TSG_E_ARG (-1)
#define TSG_E_HIP (-2)
#define TSG_E_NAN (-3)          /* std::runtime_error of computePotential & co. */
#define TSG_E_GEOM (-4)
#define TSG_E_RELAX_FAULT (-5)  /* Grid_3d::Table::RelaxationFault */
#define TSG_E_SURFACE (-6)
#define TSG_E_STATE (-7)
#define TSG_E_CONFIG (-8)       /* Configuration::Table::operator[]: an id without a species */

/* precision / method of a grid (selected at creation) */
#define TSG_PLAIN_F64 0         /* reference algorithm promoted to FP64 (drop-in, bit-exact to the FP64 oracle) */
#define TSG_FAITHFUL_F32 1      /* the reference's own float storage, double accumulation */
#define TSG_MIXED_F32_REFINE 2  /* FP32 sweeps on the correction, FP64 residual and refinement */

/* node flags, tapsim/grid_3d.h:84 */
#define TSG_FLAG_BOUNDARY 0x01
#define TSG_FLAG_DIRICHLET 0x02
#define TSG_FLAG_NEUMANN 0x04

const char* tsg_last_error(void);
int tsg_version(void);
void tsg_free(void* p);

/* ------------------------------------------------------------------ mesh */

typedef struct tsg_specimen_params
{
	double a;                 /* BCC lattice constant, m */
	double radius;            /* tip radius, m */
	double shank;             /* cylinder height below the cap centre, m */
	double box[3];            /* box extent, m (x, y centred on the axis; z from the base) */
	double shell;             /* vacuum lattice shell, in lattice constants */
	long long vacuum_target;  /* graded random vacuum nodes (approximate) */
	double decay;             /* vacuum density ~ distance^-decay */
	double face_spacing;      /* box surface node spacing, m */
	double cluster_density;   /* solute clusters per m^3 (0: one phase) */
	double cluster_radius;    /* m */
	unsigned long long seed;
	int id_vacuum, id_top, id_bottom, id_sides, id_matrix, id_solute;
} tsg_specimen_params;

typedef struct tsg_mesh tsg_mesh;

/* Synthetic BCC [110] tip node set (replaces the node file that the
   reference reads with File_Io::readGeomGrid, tapsim/file_io.cpp:822).
   *xyz (3n floats) and *ids (n shorts) are released with tsg_free. */
int tsg_specimen_generate(const tsg_specimen_params* p, int* n, float** xyz, short** ids, long long counts[5]);

/* Delaunay tetrahedralisation + topology (replaces
   Geometry_3d::Table::tetrahedralize, tapsim/geometry_3d.cpp:759-810). */
int tsg_mesh_create(int n, const float* xyz, tsg_mesh** out);
/* Adopt a given tetrahedralisation (neighbours[i] opposite vertices[i], -1
   on the hull; Geometry_3d::Table::set_metaInformation, geometry_3d.cpp:695) */
int tsg_mesh_create_from_tets(int n, const float* xyz, int m, const int* tet_v, const int* tet_n, tsg_mesh** out);
int tsg_mesh_info(const tsg_mesh*, int* n, int* m, long long* nnz);
int tsg_mesh_get_tets(const tsg_mesh*, int* tet_v, int* tet_n);
int tsg_mesh_get_csr(const tsg_mesh*, int* ptr, int* idx);          /* adjacentNodes, ascending */
int tsg_mesh_get_nodeinfo(const tsg_mesh*, int* assoc, unsigned char* boundary);
void tsg_mesh_destroy(tsg_mesh*);


/* ------------------------------------------------------------------ grid */

/* One species of the configuration table (Configuration::NodeData,
   tapsim/configuration.h:100-150); the table is indexed by NodeId (0..127). */
typedef struct tsg_species
{
	int defined;
	float charge_density;     /* C/m^3 */
	float epsilon;
	int removable, neumann, dirichlet;
	float phi;                /* preset potential, V */
	float mass;               /* amu */
	int evap_charge;
	float evap_field;         /* V/m */
	float evap_energy;        /* eV */
} tsg_species;

typedef struct tsg_grid tsg_grid;

/* Builds the device grid on a mesh: Voronoi face areas on the GPU, then
   Grid_3d::Table::fastSync (tapsim/grid_3d.cpp:491-553); potentials start at
   the species presets (Process::initialization, process.cpp:65-80).
   mode: TSG_PLAIN_F64 or TSG_FAITHFUL_F32. */
int tsg_grid_create(const tsg_mesh* mesh, const short* ids, const tsg_species* species /* [128] */,
                    float temperature, int mode, tsg_grid** out);
void tsg_grid_destroy(tsg_grid*);
int tsg_grid_info(const tsg_grid*, int* n, long long* nnz, int* nrows);

/* Grid_3d::Table::relax(cycleSize) (grid_3d.cpp:591-609); returns the deviation */
int tsg_relax_cycles(tsg_grid*, int cycles, double* deviation);
/* Grid_3d::Table::relax(threshold, cycleSize, queueSize) (grid_3d.cpp:611-671);
   *iterations = -1 - iteration when a RelaxationFault (zero slope) ends it */
int tsg_relax_threshold(tsg_grid*, double threshold, int cycles, int queue, long long* iterations,
                        double* fault_dev, double* fault_slope, double* last_dev);
/* Grid_3d::Table::localRelax(index, order, threshold, cycleSize, queueSize)
   (grid_3d.cpp:707-795) */
int tsg_local_relax(tsg_grid*, int node, int order, double threshold, int cycles, int queue, long long* iterations);

/* potential buffers (Node::phi, grid_3d.h:57-58) and the buffer switches */
int tsg_get_phi(tsg_grid*, int buffer, double* out);
int tsg_set_phi(tsg_grid*, const double* phi0, const double* phi1, int phi_switch, int local_switch);
int tsg_phi_switches(const tsg_grid*, int* phi_switch, int* local_switch);
/* initial potentials, both buffers: kind 0 the species presets
   (Grid_3d::Table::reset, grid_3d.cpp:797-809), 1 random for the
   non-Dirichlet nodes from the process's rand() (randomReset, :811-848),
   2 the radial ramp of preset_sphericalPotential (system_3d.cpp:24-57) from
   pmax at center */
int tsg_reset_potentials(tsg_grid*, int kind, const float* center, float pmin, float pmax);
/* Grid_3d::Table::potential(i) (grid_3d.h:184) */
int tsg_potential(tsg_grid*, int node, double* out);
/* Grid_3d::Table::field_o1(i) (grid_3d.cpp:909-991), 3 values per node */
int tsg_field_o1(tsg_grid*, int count, const int* nodes, double* out);
/* field_o1 of a node list, as tsg_field_o1, allocating no device memory.
   No counterpart in the reference: for callers that cannot allocate (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The node list, fields and status word are staged in a
   file-scope device array of the library (1.8 MB per GPU, held until the
   process ends); count above TSG_NOALLOC_MAX_NODES fails with TSG_E_ARG.
   Every call on one GPU shares that array: calls that overlap on one GPU
   race and must be serialised by the caller. */
#define TSG_NOALLOC_MAX_NODES 65536
int tsg_field_o1_noalloc(tsg_grid*, int count, const int* nodes, double* out);
/* Grid_3d::Table::field_o2(i, geometry, &charge) (grid_3d.cpp:993-1199): the
   quadratic least-squares field (3 values per node into out) and the charge
   -eps0 (phi_xx + phi_yy + phi_zz) of the fit (one per node into charge) */
int tsg_field_o2(tsg_grid*, int count, const int* nodes, double* out, double* charge);
/* the host mirror of Node::coupling for every directed entry of the CSR */
int tsg_get_couplings(const tsg_grid*, double* out);
/* Node::charge of every node (computeCharge, grid_3d.cpp:1373-1390: Voronoi
   cell volume closed at half the mean neighbour distance x charge density /
   epsilon0), as the grid stores it */
int tsg_get_charges(const tsg_grid*, double* out);
int tsg_get_flags(const tsg_grid*, unsigned char* out);
int tsg_get_ids(const tsg_grid*, short* out);
/* per directed entry: Voronoi face area closed at |pi-pj|/2 (float and
   double) and closed at 1 m (float) -- Geometry_3d::Table::voronoiArea */
int tsg_get_areas(const tsg_grid*, float* half32, double* half64, float* one32);
/* Node::setId + Grid_3d::Table::fastResync (grid_3d.cpp:555-589) */
int tsg_set_id(tsg_grid*, int node, int id);

/* Surface_3d::Table::init (surface_3d.cpp:50-81) */
int tsg_surface_init(tsg_grid*, int vacuum_id);
int tsg_surface_size(const tsg_grid*, int* count);
int tsg_surface_nodes(const tsg_grid*, int* out);
/* evap_compute_specificFields + evap_compute_probabilities (linear field),
   surface_3d.cpp:194-219,266-297; exact_sum also forms the normaliser */
int tsg_surface_evaluate(tsg_grid*, int exact_sum, double* scaling_ref);
/* the same with the probability model of evap_compute_probabilities
   (surface_3d.cpp:223-417): 0 probLinear_field, 1 probBoltzmann (needs the
   grid's temperature), 2 probLinear_force (field_o2), 3 probVoronoiFluxForce */
int tsg_surface_evaluate_model(tsg_grid*, int mode, int exact_sum, double* scaling_ref);
/* seeds the process's rand() stream (the reference seeds the C library's
   with 99, tapsim.cpp:235).  The library keeps this stream itself, so that
   the point-location walks on the GPU can continue it: evapMonteCarlo, the
   thermal start velocity, randomReset and every walk draw from it in the
   reference's order.  For equal seeds it is glibc's rand() sequence. */
void tsg_srand(unsigned int seed);
/* the next count draws of that stream (std::rand() in the reference) */
int tsg_rand(int count, int* out);
int tsg_surface_values(tsg_grid*, double* out);
int tsg_surface_probabilities(tsg_grid*, float* out);
/* evapMaximum (surface_3d.cpp:423-431) */
int tsg_surface_candidate(tsg_grid*, int* node, int* near_tie);
/* evap_takeAway (surface_3d.cpp:507-564) */
int tsg_take_away(tsg_grid*, int node);
/* the fork's update_coordination (surface_3d.cpp:568-593): every id >= 10
   grows by its number of neighbours with id >= 10, in node order, wrapping
   as a signed char; only the ids change */
int tsg_update_coordination(tsg_grid*);
/* Surface_3d::Table::normal (surface_3d.cpp:123-169): best-plane normal of
   the surface sites within three Delaunay hops, oriented toward the vacuum
   nodes of that shell; zero for a node that is not a surface site */
int tsg_surface_normal(const tsg_grid*, int node, float out[3]);
/* Geometry_3d::find_best_plane (geometry_3d.cpp:1806-1870): unit
   eigenvector of the smallest eigenvalue of the scatter matrix of count
   offsets (x,y,z floats); host only, no device needed */
int tsg_best_plane(const float* offsets, int count, float out[3]);

typedef struct tsg_evap_options
{
	int prob_mode, evap_mode, vacuum_id;
	double relax_threshold;
	int relax_shell, relax_cycle, relax_queue;
	int relax_global_cycles;
	int refresh_interval;
	double refresh_threshold;
	int refresh_cycle, refresh_queue;
	int exact_probability;
	int method;               /* 0: the reference's relaxation (steps III as process.cpp:824-900), 1: accelerated,
	                             2: accelerated from the learned warm start (tsg_grid_attach_mpnn) */
	double accel_tol;         /* accelerated: max |J(phi) - phi| target, V */
	/* step V, the ion trajectory (Process::EvaporationOptions, process.h:39-122) */
	int trajectory;           /* 0 off, 1 on */
	int integrator, stepper;  /* see tsg_trajectories */
	float dt0, dt_min, err_limit;
	int voltage_queue;        /* running mean length of the voltage (process.cpp:983-1010) */
	float shrink_limit;       /* stop once the apex has receded this far, m */
	int update_coordination;  /* 1: the fork's update_coordination (surface_3d.cpp:568-593) after the initial
	                             surface evaluation and after every take-away (process.cpp:647,770); a later
	                             lookup of an id without a species fails with TSG_E_CONFIG, as the reference throws */
	int thermal_velocity;     /* 1: start velocity of evap_initialVelocity (surface_3d.cpp:472-502) from the
	                             process's rand() and the grid's temperature; 0: at rest (the reference's
	                             TRAJECTORY_NO_INITIAL_VELOCITY = 1) */
	int accel_solver;         /* accelerated methods' FP32 correction solve (no reference counterpart):
	                             0 Chebyshev-Jacobi, 1 conjugate gradients preconditioned by one V-cycle of
	                             aggregation multigrid (iterations replayed from a captured HIP graph), 2 the
	                             same launched kernel by kernel; the FP64 refinement and its target are the same */
	int init_event_cnt;       /* EvaporationOptions::initEventCnt (process.h:98): the event index the run
	                             starts after; the grid keeps counting across tsg_evaporate calls, so events
	                             evaporated in chunks get the indices (and refresh relaxations,
	                             process.cpp:876) of one uninterrupted run */
	/* additions without a reference counterpart */
	float detector_z;         /* > 0: each flight is extrapolated to the plane z = detector_z (m) instead of the
	                             mesh's top (process.cpp:960 extrapolates to Geometry_3d::Table::max().z()) */
	int ion_table;            /* 1: each ion's mass and charge are drawn from its species' ion table
	                             (tsg_set_ion_table); 0 or an empty table: the species' mass and charge */
	unsigned int ion_seed;    /* seed of the ion draws: one counter-based draw (splitmix64 of seed and event
	                             index) per event, independent of the process's rand() stream */
} tsg_evap_options;

typedef struct tsg_event
{
	int node, id;
	float probability;
	double pot_before, field_before[3];
	double pot_after, field_after[3];
	long long local_iter, global_iter;
	int surface_size;
	double scaling_ref;
	int near_tie;
	int warm_start;           /* 0 none, 1 prediction applied, 2 prediction rejected, 3 graph recorded only */
	double residual_cold;     /* max |J(phi)-phi| over the shell after the removal */
	double residual_warm;     /* max |J(phi)-phi| over the grid after the prediction */
	/* steps V and VI (process.cpp:940-1037) */
	float start[3], stop[3];  /* first and last (extrapolated) trajectory point */
	float apex[3];            /* highest surface site before the event */
	float normal[3];          /* surface normal at the site, surface_3d.cpp:123-169 */
	float voltage;            /* running mean of (phi_before - phi(exit)) F / |E_before| / scaling */
	float time_scale;         /* sqrt((phi_after - phi(exit)) / voltage) */
	float tof;                /* time of the last trajectory point times time_scale */
	int traj_len, traj_status;
	long long event_index;    /* LogResults::Dataset::eventIndex (process.cpp:696): initEventCnt + events so far */
	float ion_mass;           /* the ion's mass (amu) and charge (e) in its flight: the species' own, or the
	                             entry of its ion table drawn for this event */
	int ion_charge;
	float err_pos[3];         /* the flight's final error estimate, Trajectory_3d::error_estimate (position,
	                             velocity; trajectory_3d.cpp:199,327,680-696), which LogTrajectory writes */
	float err_vel[3];
} tsg_event;

/* Process::evaporation steps I-IV and VI (process.cpp:680-1037) for up to k
   events; *done is the number recorded in out */
int tsg_evaporate(tsg_grid*, const tsg_evap_options*, int k, tsg_event* out, int* done);
/* tsg_evaporate recording each event's trajectory for File_Io::LogTrajectory
   (process.cpp:976-979 pushes every event's Trajectory_3d): the first cap
   phases (t, x, y, z, vx, vy, vz, tet; t unscaled, as Trajectory_3d holds
   it) of event e go to phases + 8 cap e and their full count to
   phase_len[e]. */
int tsg_evaporate_traj(tsg_grid*, const tsg_evap_options*, int k, tsg_event* out, int* done, int cap, float* phases,
                       int* phase_len);

/* Accelerated relaxation (no reference counterpart; replaces step III's
   relaxations when selected): FP32 Chebyshev-accelerated Jacobi correction
   solves inside FP64 residual refinement until max|J(phi)-phi| <= tol, where
   J is one Jacobi sweep (the reference's deviation for that sweep).  Both
   potential buffers then hold the result.  setup estimates the spectral
   radius (power iteration) with the given safety factor (0 < safety <= 1). */
int tsg_accel_setup(tsg_grid*, double safety, double* rho);
int tsg_accel_relax(tsg_grid*, double tol, int* outer, long long* inner, double* final_residual);
/* Selects the correction solve of tsg_accel_relax (no reference counterpart):
   on = 1 conjugate gradients preconditioned by one V-cycle of aggregation
   multigrid (hierarchy built from the current rows on the first solve; inner
   then counts CG iterations; the iterations replay a captured HIP graph),
   on = 2 the same with every kernel launched on its own (identical results),
   on = 0 the Chebyshev-Jacobi steps.  *levels receives the number of coarse
   levels built so far (0 before the first solve). */
int tsg_accel_multigrid(tsg_grid*, int on, int* levels);

/* Measurement (no reference counterpart): host wall time of tsg_evaporate's
   phases, accumulated over every event so far, in ms: ms[0] the choice (with
   the fields before), [1] the take-away, [2] the warm start, [3] the
   relaxation, [4] the surface evaluation (with the fields after), [5] the
   flight hand-off and collection; n <= 6 entries are written; *events the
   number of events timed. */
int tsg_phase_times(tsg_grid*, double* ms, int n, long long* events);

/* Measurement: average device time of one launch (HIP events on the grid's
   stream around reps launches) of kind 0 = the Jacobi cycle, kind 1 = the
   FP32 Chebyshev correction step, kind 2 = the multigrid solve's level-0
   operator product q = A p (hierarchy built if needed), and that kernel's
   algorithmic bytes per launch.  Potentials are restored afterwards. */
int tsg_time_kernel(tsg_grid*, int kind, int reps, double* avg_ms, double* bytes);

/* Measurement: algorithmic bytes of one sweep (column index + coupling per
   entry; node, degree, coupling sum, charge and written potential per row;
   one read of every referenced potential) and the accumulated device time of
   the threshold relaxations (HIP events on the grid's stream).  No reference
   counterpart: the reference has no instrumentation of its sweep. */
int tsg_sweep_stats(const tsg_grid*, double* bytes, long long* rows, long long* entries, long long* referenced,
                    double* ms, long long* sweeps, long long* launched);
/* a plain device-to-device copy of `bytes` (multiple of 16), launched reps
   times: the known-size kernel that calibrates the HBM traffic counters;
   no counterpart in the reference (measurement only) */
int tsg_copy_probe(long long bytes, int reps, double* avg_ms);

/* ------------------------------------------------------------------ trajectories
 * Trajectory_3d::init + integrate(delta, errorLimit) + extrapolate(Z, zmax)
 * (tapsim/trajectory_3d.cpp:99-332, with Grid_3d::FastField::compute,
 * grid_3d.cpp:1669-1721, and Geometry_3d::Table::findTetrahedron,
 * geometry_3d.cpp:988-1044) for `count` ions, each starting at rest on
 * nodes[i] with charge[i] (C) and mass[i] (kg), in the field of the current
 * potentials.  integrator: 0 first order, 1 Cash-Karp (the reference's
 * default), 2 the classical fourth-order Runge-Kutta step with the embedded
 * third-order error estimate (k4 - k5) / 6, k5 the slope at the new point
 * (an addition, no reference counterpart); stepper: 0 none, 1 error-restricted (default), 2
 * geometry-restricted, 3 adaptive.  out receives, per ion, the first cap
 * phases (t, x, y, z, vx, vy, vz, tet) of its trajectory, the last one the
 * extrapolation to the mesh's largest z; len[i] the number of phases and
 * status[i] the Trajectory_3d status (trajectory_3d.h:89; 5 = left the
 * mesh, 1 = no field), negative on failure (-4: work limit);
 * error_estimate (may be null) 6 floats per ion, Trajectory_3d::error_estimate
 * (trajectory_3d.cpp:199,327,680-696): the root-sum-square of the accepted
 * steps' error estimates, position xyz then velocity xyz, the position part
 * moved along by the extrapolation. */
int tsg_trajectories(tsg_grid*, int count, const int* nodes, const float* charge, const float* mass, int integrator,
                     int stepper, float dt0, float dt_min, float err_limit, int cap, float* out, int* len, int* status,
                     float* error_estimate);

/* Geometry_3d::Table's extents (geometry_3d.cpp:1555-1610), as the dump of
   File_Io::writeSystem holds them: out[0..2] min, [3..5] max of the node
   coordinates, [6..8] and [9..11] the same widened by the first and last
   vertices of every open Voronoi face (setConstrainedExtents) */
int tsg_grid_extents(tsg_grid*, float* out);

/* The ion table of a species for the flights of tsg_evaporate (no reference
 * counterpart; the reference flies every ion of a species with the species'
 * mass and charge, process.cpp:930-935): up to 8 entries of mass (amu),
 * charge (e) and weight >= 0; with tsg_evap_options.ion_table = 1 each event
 * draws one entry with probability weight / sum of weights.  count 0 clears
 * the table. */
int tsg_set_ion_table(tsg_grid*, int species, int count, const float* mass_amu, const int* charge, const float* weight);

/* Fieldline_3d::init(start) + compute(distance) (tapsim/fieldline_3d.cpp:55-190)
 * for `count` start points (3 floats each) in the field of the current
 * potentials: steps of (start tet volume)^(1/3)/2 along the unit field,
 * accepted when they land in a tet sharing a vertex with the current one,
 * the step growing by 1.1 while it stays in the current tet and shrinking by
 * 0.9 otherwise.  distance < 0: no distance limit.  The reference leaves its
 * iteration and stepper limits uninitialised; here they are arguments.
 * process_rand 1: the lines run in order and continue the process's rand()
 * stream (tetGuess samples, walk faces) as the reference; 0: one device
 * thread per line.  out: per line the first cap points (x, y, z, Ex, Ey, Ez,
 * tet); status[i] the Fieldline_3d status (fieldline_3d.h:74; 5 = left the
 * mesh, 4 = distance reached, 1 = no field), -10 for a start outside the mesh. */
int tsg_fieldlines(tsg_grid*, int count, const float* start, float distance, long long iterations,
                   long long stepper_limit, int process_rand, int cap, float* out, int* len, int* status);

/* ------------------------------------------------------------------ learned warm start
 * No counterpart in the reference (TAPSim starts every relaxation from the
 * previous potentials).  A message-passing network over the 3-hop shell of
 * the evaporated atom predicts the change of the potentials that the
 * relaxation will produce; the prediction is added before the relaxation,
 * which still runs to its residual target (the network never decides the
 * result, only where the relaxation starts).
 *
 *   h0   = relu(X W_in + b_in)                        X: n x 32, W_in: 32 x 64
 *   h_l  = relu([h_{l-1} | A h_{l-1}] W_l + b_l) + h_{l-1},  l = 1..3, W_l: 128 x 64
 *   y    = h_3 w_out + b_out
 * A holds symmetric edge weights.  The GEMMs run on bfloat16 MFMA with fp32
 * accumulation; parameters are one fp32 array in this layout (row-major W[k][o]):
 */
#define TSG_MPNN_IN 32
#define TSG_MPNN_HID 64
#define TSG_MPNN_LAYERS 3
#define TSG_MPNN_W_IN 0
#define TSG_MPNN_B_IN (TSG_MPNN_IN * TSG_MPNN_HID)
#define TSG_MPNN_LAYER0 (TSG_MPNN_B_IN + TSG_MPNN_HID)
#define TSG_MPNN_LAYER_STRIDE (2 * TSG_MPNN_HID * TSG_MPNN_HID + TSG_MPNN_HID)
#define TSG_MPNN_W_OUT (TSG_MPNN_LAYER0 + TSG_MPNN_LAYERS * TSG_MPNN_LAYER_STRIDE)
#define TSG_MPNN_B_OUT (TSG_MPNN_W_OUT + TSG_MPNN_HID)
#define TSG_MPNN_NPARAM (TSG_MPNN_B_OUT + 1)

typedef struct tsg_mpnn tsg_mpnn;

/* a network with its graph store in device memory: batches of up to
 * max_batch_rows nodes; the store holds data_rows nodes and data_edges edges */
int tsg_mpnn_create(int max_batch_rows, long long data_rows, long long data_edges, unsigned seed, tsg_mpnn** out);
void tsg_mpnn_destroy(tsg_mpnn*);
int tsg_mpnn_set_params(tsg_mpnn*, const float* params);
int tsg_mpnn_get_params(tsg_mpnn*, float* params);
/* one graph from the host (local CSR ptr[n+1]/col/a, x: n*32 features,
 * targets t and loss weights mask): loss = sum mask (y-t)^2 / max(1, sum mask),
 * y (n) and d loss / d params (TSG_MPNN_NPARAM); each output may be NULL */
int tsg_mpnn_eval(tsg_mpnn*, int n, const int* ptr, const int* col, const float* a, const float* x, const float* t,
                  const float* mask, double* loss, float* y, float* grad);
/* the training store: graphs recorded by evaporation events (see
 * tsg_grid_attach_mpnn) or added from the host */
int tsg_mpnn_add_graph(tsg_mpnn*, int n, const int* ptr, const int* col, const float* a, const float* x,
                       const float* t, const float* mask);
int tsg_mpnn_dataset(const tsg_mpnn*, int* graphs, long long* rows, long long* edges);
int tsg_mpnn_clear(tsg_mpnn*);
/* Adam over graphs [g_begin, g_end), batches of up to batch_graphs
 * consecutive graphs in shuffled order, the learning rate lr decayed over
 * the epochs on a cosine (epoch e of E uses lr (1 + cos(pi e / E)) / 2);
 * mean batch loss of the first and last epoch */
int tsg_mpnn_train(tsg_mpnn*, int g_begin, int g_end, int epochs, int batch_graphs, float lr, unsigned seed,
                   double* first_loss, double* last_loss);
int tsg_mpnn_evaluate(tsg_mpnn*, int g_begin, int g_end, double* loss);
/* events of methods 1 and 2 on this grid record their 3-hop shell graph and
 * relaxation target into the network's store (record != 0); method 2 starts
 * each relaxation from the network's prediction (predict != 0).  net may be
 * NULL to detach; the network must outlive the attachment. */
int tsg_grid_attach_mpnn(tsg_grid*, tsg_mpnn* net, int record, int predict);

#ifdef __cplusplus
}
#endif

#endif
