/*
 * CPU oracle: a restatement of TAPSim's per-event electrostatics
 * (ashtonmv/tapsim_modified, TAPSim 1.0b r3225) used ONLY as the checker in
 * tests/, in __graft_entry__.smoke() and as bench.py's cpu_baseline leg.
 * The product library never links or calls it.
 *
 * Two precisions are restated:
 *   ORA_F32  "faithful": float storage of potentials, couplings and
 *            geometry exactly as the reference (tapsim/grid_3d.h:89-97),
 *            double accumulation in computePotential (grid_3d.cpp:1317).
 *   ORA_F64  the same algorithm with every float promoted to double; this
 *            is the plain-FP64 oracle the GPU path must match bit for bit.
 *
 * Parity pinning: every function is checked in tests/ against fixtures
 * produced by the reference itself (oracle/_ref, built from the reference
 * sources by oracle/Makefile.ref) and against closed forms (linear fixed
 * points, BCC couplings, exact linear fields, uniform-field flights).
 */
#ifndef TAPSIM_ORACLE_H
#define TAPSIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_F32 = 0, ORA_F64 = 1 };

enum { ORA_OK = 0, ORA_E_ARG = -1, ORA_E_NAN = -2, ORA_E_GEOM = -3, ORA_E_SURF = -4, ORA_E_CONFIG = -5 };

typedef struct
{
	int defined;
	float charge_density;  /* C/m^3 */
	float epsilon;
	int removable;
	int neumann;
	int dirichlet;
	float phi;
	float mass;            /* amu */
	int evap_charge;
	float evap_field;      /* V/m */
	float evap_energy;     /* eV */
} ora_species;

typedef struct
{
	int prob_mode;         /* 0 linear field, 1 boltzmann, 2 linear force, 3 voronoi flux */
	int evap_mode;         /* 0 maximum, 1 monte carlo */
	int vacuum_id;
	float relax_threshold; int relax_shell; int relax_cycle; int relax_queue;
	int relax_global_cycles;
	int refresh_interval; float refresh_threshold; int refresh_cycle; int refresh_queue;
	int trajectory;        /* 0: skip step V, 1: integrate */
	int integrator; int stepper;
	float dt0; float dt_min; float err_limit;
	unsigned int voltage_queue;
	float shrink_limit;
	int inplace_order;     /* 0 ascending sequential (one thread), 1 colour classes */
	int thermal_velocity;  /* 1: evap_initialVelocity from rand(), 0: at rest */
	int update_coordination; /* 1: the fork's update_coordination at the run's start and after each take-away */
	int init_event_cnt;    /* EvaporationOptions::initEventCnt (process.h:98) */
	/* additions without a reference counterpart (integrator 2 is the classical
	   fourth-order Runge-Kutta step) */
	float detector_z;      /* > 0: extrapolate the flight to this plane (m) instead of the mesh's top */
	int ion_table;         /* 1: each ion's mass and charge drawn from its species' ion table */
	unsigned int ion_seed; /* seed of those draws (one per event, keyed by the event index) */
} ora_evap_opts;

typedef struct
{
	int node;
	int id;
	float probability;
	float pot_before; float field_before[3];
	float pot_after;  float field_after[3];
	float voltage; float time_scale;
	float start[3]; float stop[3]; float tof;
	int traj_len; int traj_status;
	long long local_iter; long long refresh_iter;
	int surface_size;
	float scaling_ref;
	float apex[3];
	float normal[3];
	int traj_root_diff;    /* flight steps where powf(x, 0.25f) is not the correctly rounded root */
	long long event_index; /* LogResults::Dataset::eventIndex (process.cpp:696) */
	float ion_mass;        /* the ion's mass (amu) and charge (e) in its flight */
	int ion_charge;
	float err_pos[3];      /* the flight's final error estimate (Trajectory_3d::error_estimate) */
	float err_vel[3];
} ora_event;

typedef struct ora_sys ora_sys;

ora_sys* ora_create(int mode, int n, const float* xyz, const short* ids,
                    int m, const int* tet_v, const int* tet_n,
                    const ora_species* species /* 128 */, float temperature);
void ora_destroy(ora_sys*);
const char* ora_last_error(void);

int ora_sync(ora_sys*);                                   /* fastSync */
int ora_set_couplings(ora_sys*, const double* coupling);
int ora_resync(ora_sys*, int node);                       /* fastResync */
int ora_set_id(ora_sys*, int node, int id);
void ora_update_coordination(ora_sys*);                  /* the fork's patch, surface_3d.cpp:568-593 */

int ora_nnz(const ora_sys*);
void ora_get_csr(const ora_sys*, int* ptr, int* idx);
void ora_get_coupling(const ora_sys*, double* out);
void ora_get_charge(const ora_sys*, double* out);
void ora_get_flags(const ora_sys*, unsigned char* out);
void ora_get_ids(const ora_sys*, short* out);
void ora_get_assoc(const ora_sys*, int* out);
void ora_get_boundary(const ora_sys*, unsigned char* out);

void ora_get_phi(const ora_sys*, int buffer, double* out);
void ora_set_phi(ora_sys*, const double* phi0, const double* phi1, int phi_switch, int local_switch);
int ora_phi_switch(const ora_sys*);
int ora_local_switch(const ora_sys*);
double ora_potential(const ora_sys*, int node);
void ora_set_inplace_order(ora_sys*, int kind, const int* colour, int ncolours);

/* Grid_3d::Table::randomReset and preset_sphericalPotential (initial potentials) */
void ora_random_reset(ora_sys* s);
void ora_spherical_preset(ora_sys* s, const float center[3], float pmin, float pmax);
double ora_relax(ora_sys*, int cycles, int* err);
/* `sweeps` relax(1) calls with `threads` workers (FP64 mode; timing of the
   CPU baseline); wall time in *seconds, the last deviation returned */
double ora_relax_threaded(ora_sys* s, int sweeps, int threads, double* seconds, int* err);
/* worker threads of the threshold relaxations with cycleSize 1 (default 1) */
void ora_set_threads(ora_sys* s, int threads);
long long ora_relax_threshold(ora_sys*, double threshold, int cycles, int queue,
                              double* fault_dev, double* fault_slope, int* err);
int ora_find_local_nodes(const ora_sys*, int node, int order, int* out, int cap);
double ora_local_relax(ora_sys*, int node, int order, int cycles, int* err);
long long ora_local_relax_threshold(ora_sys*, int node, int order, double threshold,
                                    int cycles, int queue, int* err);

int ora_field_o1(const ora_sys*, int node, double out[3]);
int ora_field_o2(const ora_sys*, int node, double out[3], double* charge);
double ora_voronoi_area(const ora_sys*, int n1, int n2, double inf);
double ora_voronoi_volume(const ora_sys*, int node, double inf);
double ora_distance(const ora_sys*, int n1, int n2);

/* surface / selection */
int ora_surface_init(ora_sys*, int vacuum_id);
int ora_surface_size(const ora_sys*);
void ora_surface_nodes(const ora_sys*, int* out);
void ora_surface_probabilities(const ora_sys*, float* out);
int ora_compute_probabilities(ora_sys*, int mode);
float ora_specific_fields(ora_sys*);
int ora_find_candidate(const ora_sys*, int mode);
int ora_take_away(ora_sys*, int node);
/* Surface_3d::Table::normal / Geometry_3d::find_best_plane */
int ora_surface_normal(const ora_sys*, int node, float out[3]);
int ora_best_plane(const float* offsets, int count, float out[3]);

/* trajectory (float, as trajectory_3d.cpp) */
void ora_srand(unsigned int seed);
int ora_find_tetrahedron(const ora_sys*, const float pos[3], int guess);  /* guess < 0: -1 */
int ora_locate(const ora_sys*, const float pos[3]);                       /* tetGuess + walk */
/* the trajectories' (tol/err)^(1/4): 0 glibc's powf as the reference, 1 the
   correctly rounded root (the device's arithmetic) */
void ora_set_root_mode(int mode);
/* Fieldline_3d::init(start) + compute(distance) with the given stepper and
   iteration limits; out: 7 floats per point (x, y, z, Ex, Ey, Ez, tet) */
int ora_fieldline(const ora_sys*, const float start[3], float distance, unsigned long iterations,
                  unsigned long stepper_limit, float* out, int cap, int* len);
void ora_tet_weights(const ora_sys*, int tet, const float pos[3], float w[4]);
float ora_point_potential(const ora_sys*, const float pos[3], int guess);
int ora_point_field(const ora_sys*, const float pos[3], int guess, float out[3]);
int ora_trajectory(const ora_sys*, int node, float charge, float mass, int integrator,
                   int stepper, float dt0, float dt_min, float err_limit,
                   float* out, int cap, int* len, float* err);

int ora_evaporate(ora_sys*, const ora_evap_opts*, int nevents, ora_event* out, int* done);

/* the ion table of a species (no reference counterpart): up to 8 (mass amu,
   charge e, weight) entries; count 0 clears it */
int ora_set_ion_table(ora_sys*, int species, int count, const float* mass, const int* charge, const float* weight);
/* the ion drawn for an event: index into the species' table */
int ora_ion_pick(unsigned int seed, long long event_index, int count, const float* weight);
/* the fourth-order Runge-Kutta step in an analytic field (its pinning):
   kind 0 uniform E = g, kind 1 E = -(g0 x, g1 y, g2 z); steps of delta from
   (p0, v0); out = t, p[3], v[3]; err = the last step's estimate (p[3], v[3]) */
int ora_rk4_analytic(int kind, const float g[3], float qm, const float p0[3], const float v0[3], float delta,
                     int steps, float* out, float* err);

#ifdef __cplusplus
}
#endif

#endif
