/*
 * CPU oracle for TAPSim's per-event electrostatics -- TEST INFRASTRUCTURE.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, always as the checker; the product (tapsim_modified_amd/) never does.
 *
 * Restates, one function at a time, the reference's
 *   tapsim/grid_3d.cpp       relaxation, couplings, fields
 *   tapsim/geometry_3d.cpp   Voronoi faces/cells, point location, weights
 *   tapsim/surface_3d.cpp    surface set, probabilities, selection, take-away
 *   tapsim/trajectory_3d.cpp Cash-Karp / Euler flights and step control
 *   tapsim/process.cpp       the per-event loop (steps I-VI)
 * in the reference's float storage (ORA_F32) and promoted to double
 * (ORA_F64).  Multi-threaded reference relaxation is chaotic in its in-place
 * cycles; the oracle fixes the one-thread order (ascending node index) or a
 * colour-class order.
 */
#include "tapsim_oracle.h"
#include "oracle_exact.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <pthread.h>
#include <time.h>

#define ORA_EPSILON0 8.85418781762e-12 /* grid_3d.h:285 */
#define ORA_BND 0x01                   /* grid_3d.h:84 */
#define ORA_DIR 0x02
#define ORA_NEU 0x04

static __thread char g_err[256];
const char* ora_last_error(void) { return g_err; }
static int fail(int code, const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); return code; }

struct ora_sys
{
	int mode;
	int n, m;
	float* xyz;
	short* ids;
	int* tv;
	int* tn;
	int* assoc;
	unsigned char* gbnd;      /* geometric boundary (Geometry_3d NodeData::boundary) */
	int* nt_ptr;              /* node -> incident tets, ascending tet index */
	int* nt_idx;
	int* ptr;                 /* node -> Delaunay neighbours, ascending */
	int* idx;
	int nnz;
	void* coup;
	void* charge;
	unsigned char* flags;
	void* phi0;
	void* phi1;
	int sw, lsw;
	ora_species species[128];
	float temperature;
	int inplace_kind;
	int* colour;
	int ncolours;
	/* surface table (Surface_3d::Table), ascending node index */
	int* surf;
	double* prob;
	int nsurf, capsurf;
	int coord_done;        /* the run-start update_coordination has been applied */
	int vacuum_id;
	double scaling_ref;
	/* Process::evaporation state kept across ora_evaporate calls, so that a
	   run evaporated in chunks equals one uninterrupted run: the event counter
	   (process.cpp:676-696), the voltage queue (:983-1010), the initial apex */
	long long event_cnt;   /* -1 until the first call sets it from initEventCnt */
	int threads;           /* workers of the cycleSize-1 threshold relaxations (1: calling thread) */
	float vq[4096];
	int vq_head, vq_size;
	float vq_sum;
	float apex0;
	int apex0_set;
	/* per species: the ion table of the flights (no reference counterpart) */
	int ion_n[128];
	float ion_mass[128][8];
	int ion_q[128][8];
	float ion_w[128][8];
};

/* Configuration::Table::operator[] (configuration.cpp:213-231) throws for an
   id without a species; ids can leave the table through the fork's
   update_coordination */
static int sp_bad(const ora_sys* s, int id) { return id < 0 || id > 127 || !s->species[id].defined; }

static void ora_flags_from_config(ora_sys* s, int i)
{
	const ora_species* sp = &s->species[s->ids[i]];
	unsigned char f = 0;
	if (s->gbnd[i]) f |= ORA_BND;
	if (sp->dirichlet) f |= ORA_DIR;
	if (sp->neumann) f |= ORA_NEU;
	s->flags[i] = f;
}

static int ora_find_slot(const ora_sys* s, int row, int col)
{
	int k = s->ptr[row];
	while (s->idx[k] != col) k++;
	return k;
}

/* tapsim/utils.h:52-140 (Numerical Recipes ludcmp with the reference's
   singular checks); T is the matrix element type */
#define LU_DECMP(T, NAME)                                                              \
static int NAME(T* lu, int n, int* index, T* d)                                        \
{                                                                                      \
	T parity = (T) 1.0;                                                                \
	T ii[9], vv[9];                                                                    \
	for (int i = 0; i < n; i++)                                                        \
	{                                                                                  \
		T big = (T) 0.0;                                                               \
		for (int j = 0; j < n; j++)                                                    \
		{                                                                              \
			const T tmp = (T) fabs(lu[i * n + j]);                                     \
			if (tmp > big) big = tmp;                                                  \
		}                                                                              \
		if (big == 0.0) return 0;                                                      \
		vv[i] = (T) 1.0;                                                               \
		vv[i] /= big;                                                                  \
	}                                                                                  \
	int imax = 0;                                                                      \
	for (int j = 0; j < n; j++)                                                        \
	{                                                                                  \
		for (int i = 0; i < j; i++)                                                    \
		{                                                                              \
			T sum = lu[i * n + j];                                                     \
			for (int k = 0; k < i; k++) sum -= lu[i * n + k] * lu[k * n + j];          \
			lu[i * n + j] = sum;                                                       \
		}                                                                              \
		T big = (T) 0.0;                                                               \
		for (int i = j; i < n; i++)                                                    \
		{                                                                              \
			T sum = lu[i * n + j];                                                     \
			for (int k = 0; k < j; k++) sum -= lu[i * n + k] * lu[k * n + j];          \
			lu[i * n + j] = sum;                                                       \
			const T dum = vv[i] * (T) fabs(sum);                                       \
			if (dum >= big) { big = dum; imax = i; }                                   \
		}                                                                              \
		if (j != imax)                                                                 \
		{                                                                              \
			for (int k = 0; k < n; k++)                                                \
			{ T t = lu[imax * n + k]; lu[imax * n + k] = lu[j * n + k]; lu[j * n + k] = t; } \
			parity *= (T) -1.0;                                                        \
			vv[imax] = vv[j];                                                          \
		}                                                                              \
		ii[j] = (T) imax;                                                              \
		if (lu[j * n + j] == 0.0) return 0;                                            \
		{                                                                              \
			/* 1.0 is a double literal in the reference: the reciprocal is */        \
			/* formed in double and rounded to T                           */        \
			const T dum = (T) (1.0 / lu[j * n + j]);                                   \
			for (int i = j + 1; i < n; i++) lu[i * n + j] *= dum;                      \
		}                                                                              \
	}                                                                                  \
	if (index) for (int i = 0; i < n; i++) index[i] = (int) ii[i];                     \
	if (d) *d = parity;                                                                \
	return 1;                                                                          \
}

/* tapsim/utils.h:163-195 */
#define LU_SOLVE(T, NAME)                                                              \
static void NAME(const T* lu, int n, const int* index, T* b)                           \
{                                                                                      \
	int ii = 0;                                                                        \
	for (int i = 0; i < n; i++)                                                        \
	{                                                                                  \
		const int ip = index[i];                                                       \
		T sum = b[ip];                                                                 \
		b[ip] = b[i];                                                                  \
		if (ii) { for (int j = ii - 1; j < i; j++) sum -= lu[i * n + j] * b[j]; }      \
		else if (sum) ii = i + 1;                                                      \
		b[i] = sum;                                                                    \
	}                                                                                  \
	for (int i = n - 1; i >= 0; i--)                                                   \
	{                                                                                  \
		T sum = b[i];                                                                  \
		for (int j = i + 1; j < n; j++) sum -= lu[i * n + j] * b[j];                   \
		b[i] = sum / lu[i * n + i];                                                    \
	}                                                                                  \
}

LU_DECMP(double, ora_lu_decmp_d)
LU_SOLVE(double, ora_lu_solve_d)
LU_DECMP(float, ora_lu_decmp_fl)

/* tapsim/utils.h:212-226, float instance */
static float ora_lu_det_fl(const float* lu, int n, float d)
{
	float x = d;
	int y = 0;
	for (int i = 0; i < n; i++)
	{
		int tmp;
		x *= frexpf(lu[i * n + i], &tmp);
		y += tmp;
	}
	return ldexpf(x, y);
}

#define R float
#define SQRT sqrtf
#define FN(x) x##_f32
#include "oracle_impl.inc"
#undef R
#undef SQRT
#undef FN

#define R double
#define SQRT sqrt
#define FN(x) x##_f64
#include "oracle_impl.inc"
#undef R
#undef SQRT
#undef FN

#define DISPATCH(s, name, ...) ((s)->mode == ORA_F32 ? name##_f32(__VA_ARGS__) : name##_f64(__VA_ARGS__))

/* ------------------------------------------------------------ construction */

static int cmp_int(const void* a, const void* b)
{
	const int x = *(const int*) a, y = *(const int*) b;
	return (x > y) - (x < y);
}

ora_sys* ora_create(int mode, int n, const float* xyz, const short* ids,
                    int m, const int* tet_v, const int* tet_n,
                    const ora_species* species, float temperature)
{
	if (n <= 0 || m <= 0 || (mode != ORA_F32 && mode != ORA_F64)) { fail(ORA_E_ARG, "bad arguments"); return 0; }
	ora_sys* s = (ora_sys*) calloc(1, sizeof(ora_sys));
	if (!s) return NULL;
	s->event_cnt = -1;
	s->threads = 1;
	s->mode = mode;
	s->n = n;
	s->m = m;
	s->xyz = (float*) malloc(sizeof(float) * 3 * (size_t) n);
	memcpy(s->xyz, xyz, sizeof(float) * 3 * (size_t) n);
	s->ids = (short*) malloc(sizeof(short) * (size_t) n);
	memcpy(s->ids, ids, sizeof(short) * (size_t) n);
	s->tv = (int*) malloc(sizeof(int) * 4 * (size_t) m);
	memcpy(s->tv, tet_v, sizeof(int) * 4 * (size_t) m);
	s->tn = (int*) malloc(sizeof(int) * 4 * (size_t) m);
	memcpy(s->tn, tet_n, sizeof(int) * 4 * (size_t) m);
	memcpy(s->species, species, sizeof(ora_species) * 128);
	s->temperature = temperature;

	/* Geometry_3d::Table::setAssociatedTetrahedra (geometry_3d.cpp:1497-1507):
	   the last tet naming the node; setBoundary (:1475-1495): vertices of
	   faces without neighbour */
	s->assoc = (int*) malloc(sizeof(int) * (size_t) n);
	s->gbnd = (unsigned char*) calloc((size_t) n, 1);
	for (int i = 0; i < n; i++) s->assoc[i] = -1;
	s->nt_ptr = (int*) calloc((size_t) n + 1, sizeof(int));
	for (int t = 0; t < m; t++)
		for (int k = 0; k < 4; k++)
		{
			const int v = s->tv[4 * (size_t) t + k];
			s->assoc[v] = t;
			s->nt_ptr[v + 1]++;
			if (s->tn[4 * (size_t) t + k] < 0)
				for (int q = 0; q < 4; q++)
					if (q != k) s->gbnd[s->tv[4 * (size_t) t + q]] = 1;
		}
	for (int i = 0; i < n; i++) s->nt_ptr[i + 1] += s->nt_ptr[i];
	s->nt_idx = (int*) malloc(sizeof(int) * (size_t) s->nt_ptr[n]);
	{
		int* fill = (int*) malloc(sizeof(int) * (size_t) n);
		memcpy(fill, s->nt_ptr, sizeof(int) * (size_t) n);
		for (int t = 0; t < m; t++)
			for (int k = 0; k < 4; k++)
				s->nt_idx[fill[s->tv[4 * (size_t) t + k]]++] = t;
		free(fill);
	}
	/* Geometry_3d::Table::adjacentNodes (geometry_3d.cpp:1098-1110): all
	   vertices of the tets around the node, as an ascending std::set */
	s->ptr = (int*) calloc((size_t) n + 1, sizeof(int));
	int maxinc = 0;
	for (int i = 0; i < n; i++)
		if (s->nt_ptr[i + 1] - s->nt_ptr[i] > maxinc) maxinc = s->nt_ptr[i + 1] - s->nt_ptr[i];
	int* tmp = (int*) malloc(sizeof(int) * (3 * (size_t) maxinc + 1));
	int cap = 1 << 20;
	s->idx = (int*) malloc(sizeof(int) * (size_t) cap);
	for (int i = 0; i < n; i++)
	{
		int c = 0;
		for (int k = s->nt_ptr[i]; k < s->nt_ptr[i + 1]; k++)
			for (int q = 0; q < 4; q++)
			{
				const int v = s->tv[4 * (size_t) s->nt_idx[k] + q];
				if (v != i) tmp[c++] = v;
			}
		qsort(tmp, (size_t) c, sizeof(int), cmp_int);
		int u = 0;
		for (int k = 0; k < c; k++)
			if (u == 0 || tmp[u - 1] != tmp[k]) tmp[u++] = tmp[k];
		if (s->ptr[i] + u > cap)
		{
			while (s->ptr[i] + u > cap) cap *= 2;
			s->idx = (int*) realloc(s->idx, sizeof(int) * (size_t) cap);
		}
		memcpy(s->idx + s->ptr[i], tmp, sizeof(int) * (size_t) u);
		s->ptr[i + 1] = s->ptr[i] + u;
	}
	free(tmp);
	s->nnz = s->ptr[n];
	const size_t rs = mode == ORA_F32 ? sizeof(float) : sizeof(double);
	s->coup = calloc((size_t) s->nnz, rs);
	s->charge = calloc((size_t) n, rs);
	s->flags = (unsigned char*) calloc((size_t) n, 1);
	s->phi0 = calloc((size_t) n, rs);
	s->phi1 = calloc((size_t) n, rs);
	for (int i = 0; i < n; i++)
	{
		ora_flags_from_config(s, i);
		const float p = s->species[s->ids[i]].phi;
		if (mode == ORA_F32) { ((float*) s->phi0)[i] = p; ((float*) s->phi1)[i] = p; }
		else { ((double*) s->phi0)[i] = p; ((double*) s->phi1)[i] = p; }
	}
	s->colour = (int*) calloc((size_t) n, sizeof(int));
	s->ncolours = 1;
	s->vacuum_id = -1;
	s->scaling_ref = 1.0;
	return s;
}

void ora_destroy(ora_sys* s)
{
	if (!s) return;
	free(s->xyz); free(s->ids); free(s->tv); free(s->tn); free(s->assoc); free(s->gbnd);
	free(s->nt_ptr); free(s->nt_idx); free(s->ptr); free(s->idx); free(s->coup); free(s->charge);
	free(s->flags); free(s->phi0); free(s->phi1); free(s->colour); free(s->surf); free(s->prob);
	free(s);
}

int ora_sync(ora_sys* s) { return DISPATCH(s, fast_sync, s); }

/* adopt couplings computed elsewhere (used to time the per-event loop without
   the setup, as the reference's own timer excludes it) */
int ora_set_couplings(ora_sys* s, const double* coup)
{
	for (int i = 0; i < s->n; i++) ora_flags_from_config(s, i);
	for (int k = 0; k < s->nnz; k++)
	{
		if (s->mode == ORA_F32) ((float*) s->coup)[k] = (float) coup[k];
		else ((double*) s->coup)[k] = coup[k];
	}
	return ORA_OK;
}
int ora_resync(ora_sys* s, int node) { return DISPATCH(s, fast_resync, s, node); }
int ora_set_id(ora_sys* s, int node, int id) { s->ids[node] = (short) id; return ORA_OK; }

int ora_nnz(const ora_sys* s) { return s->nnz; }
void ora_get_csr(const ora_sys* s, int* ptr, int* idx)
{
	memcpy(ptr, s->ptr, sizeof(int) * ((size_t) s->n + 1));
	memcpy(idx, s->idx, sizeof(int) * (size_t) s->nnz);
}
static void get_real(const ora_sys* s, const void* src, size_t count, double* out)
{
	for (size_t i = 0; i < count; i++)
		out[i] = s->mode == ORA_F32 ? (double) ((const float*) src)[i] : ((const double*) src)[i];
}
void ora_get_coupling(const ora_sys* s, double* out) { get_real(s, s->coup, (size_t) s->nnz, out); }
void ora_get_charge(const ora_sys* s, double* out) { get_real(s, s->charge, (size_t) s->n, out); }
void ora_get_flags(const ora_sys* s, unsigned char* out) { memcpy(out, s->flags, (size_t) s->n); }
void ora_get_ids(const ora_sys* s, short* out) { memcpy(out, s->ids, sizeof(short) * (size_t) s->n); }
void ora_get_assoc(const ora_sys* s, int* out) { memcpy(out, s->assoc, sizeof(int) * (size_t) s->n); }
void ora_get_boundary(const ora_sys* s, unsigned char* out) { memcpy(out, s->gbnd, (size_t) s->n); }

void ora_get_phi(const ora_sys* s, int b, double* out) { get_real(s, b ? s->phi1 : s->phi0, (size_t) s->n, out); }
void ora_set_phi(ora_sys* s, const double* p0, const double* p1, int sw, int lsw)
{
	for (int i = 0; i < s->n; i++)
	{
		if (s->mode == ORA_F32) { ((float*) s->phi0)[i] = (float) p0[i]; ((float*) s->phi1)[i] = (float) p1[i]; }
		else { ((double*) s->phi0)[i] = p0[i]; ((double*) s->phi1)[i] = p1[i]; }
	}
	s->sw = sw;
	s->lsw = lsw;
}
int ora_phi_switch(const ora_sys* s) { return s->sw; }
void ora_random_reset(ora_sys* s) { if (s->mode == ORA_F32) random_reset_f32(s); else random_reset_f64(s); }
void ora_spherical_preset(ora_sys* s, const float center[3], float pmin, float pmax)
{
	if (s->mode == ORA_F32) spherical_preset_f32(s, center, pmin, pmax);
	else spherical_preset_f64(s, center, pmin, pmax);
}
int ora_local_switch(const ora_sys* s) { return s->lsw; }
double ora_potential(const ora_sys* s, int i) { return DISPATCH(s, potential, s, i); }
void ora_set_inplace_order(ora_sys* s, int kind, const int* colour, int ncolours)
{
	s->inplace_kind = kind;
	if (kind == 1 && colour)
	{
		memcpy(s->colour, colour, sizeof(int) * (size_t) s->n);
		s->ncolours = ncolours;
	}
}

/* ------------------------------------------------------------ relaxation */

double ora_relax(ora_sys* s, int cycles, int* err)
{
	int e = 0;
	const double d = DISPATCH(s, relax, s, cycles, &e);
	if (err) *err = e;
	return d;
}

/* relax(1) calls with the reference's threading of a cycle
   (cwrapper_threadedRelax, grid_3d.cpp:1458-1514): T workers each take a
   contiguous chunk of about N/T nodes, a barrier ends the cycle, the first
   worker forms the deviation and flips the switch.  A Jacobi cycle's values
   do not depend on who computes them, so the potentials and deviations equal
   ora_relax(s, 1) repeated.  With `ctl` the first worker also applies the
   threshold relaxation's queue test and threshold after every cycle. */
typedef struct
{
	double threshold;
	unsigned int queue;
	long long iteration;
	int fault;
	double fault_dev, fault_slope;
} ora_thr_ctl;

typedef struct ora_queue_s ora_queue;
static int queue_slope_zero(ora_queue* q, double dev, unsigned int qsize, unsigned int cycles, double* slope_out);

typedef struct
{
	ora_sys* s;
	long long sweeps;
	int threads, id, err;
	double dev;
	pthread_barrier_t* bar;
	double* devs;
	int* stop;
	ora_thr_ctl* ctl;
	ora_queue* q;
} ora_relax_job;

static void* relax_worker(void* arg)
{
	ora_relax_job* j = (ora_relax_job*) arg;
	ora_sys* s = j->s;
	const int n = s->n;
	const int lo = (int) ((long long) n * j->id / j->threads), hi = (int) ((long long) n * (j->id + 1) / j->threads);
	for (long long it = 0; j->ctl || it < j->sweeps; it++)
	{
		double dev = 0.0;
		const int e = s->mode == ORA_F32 ? jacobi_range_f32(s, lo, hi, &dev) : jacobi_range_f64(s, lo, hi, &dev);
		if (e) j->err = e;
		j->devs[j->id] = e ? -1.0 : dev;
		pthread_barrier_wait(j->bar);
		if (j->id == 0)
		{
			double d = 0.0;
			int bad = 0;
			for (int t = 0; t < j->threads; t++)
			{
				if (j->devs[t] < 0.0) bad = 1;
				if (j->devs[t] > d) d = j->devs[t];
			}
			j->dev = d;
			s->sw = (s->sw + 1) % 2;
			if (bad) *j->stop = 1;
			else if (j->ctl)
			{
				double slope = 0.0;
				if (queue_slope_zero(j->q, d, j->ctl->queue, 1u, &slope))
				{
					j->ctl->fault = 1;
					j->ctl->fault_dev = d;
					j->ctl->fault_slope = slope;
					*j->stop = 1;
				}
				else
				{
					j->ctl->iteration += 1;
					if (!(d > j->ctl->threshold)) *j->stop = 1;
				}
			}
		}
		pthread_barrier_wait(j->bar);
		if (*j->stop) break;
	}
	return NULL;
}

/* runs the workers; returns the last deviation */
static double run_relax_workers(ora_sys* s, long long sweeps, int threads, ora_thr_ctl* ctl, ora_queue* q,
                                double* seconds, int* err)
{
	pthread_barrier_t bar;
	pthread_barrier_init(&bar, NULL, (unsigned) threads);
	ora_relax_job jobs[256];
	double devs[256];
	pthread_t tid[256];
	int stop = 0;
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	for (int t = 0; t < threads; t++)
	{
		jobs[t].s = s; jobs[t].sweeps = sweeps; jobs[t].threads = threads; jobs[t].id = t; jobs[t].err = 0;
		jobs[t].dev = 0.0; jobs[t].bar = &bar; jobs[t].devs = devs; jobs[t].stop = &stop; jobs[t].ctl = ctl;
		jobs[t].q = q;
		if (t > 0) pthread_create(&tid[t], NULL, relax_worker, &jobs[t]);
	}
	relax_worker(&jobs[0]);
	for (int t = 1; t < threads; t++) pthread_join(tid[t], NULL);
	clock_gettime(CLOCK_MONOTONIC, &t1);
	pthread_barrier_destroy(&bar);
	if (seconds) *seconds = (double) (t1.tv_sec - t0.tv_sec) + 1e-9 * (double) (t1.tv_nsec - t0.tv_nsec);
	int e = 0;
	for (int t = 0; t < threads; t++) if (jobs[t].err) e = jobs[t].err;
	if (err) *err = e;
	return jobs[0].dev;
}

double ora_relax_threaded(ora_sys* s, int sweeps, int threads, double* seconds, int* err)
{
	if (threads < 1 || threads > 256 || sweeps < 0) { if (err) *err = ORA_E_ARG; return 0.0; }
	if (sweeps == 0) { if (seconds) *seconds = 0.0; if (err) *err = 0; return 0.0; }
	return run_relax_workers(s, sweeps, threads, NULL, NULL, seconds, err);
}

/* worker threads of the threshold relaxations with cycleSize 1 (the event
   loop's refresh relaxation); 1 runs them on the calling thread */
void ora_set_threads(ora_sys* s, int threads) { s->threads = threads < 1 ? 1 : (threads > 256 ? 256 : threads); }

/* the deviation queue and least-squares slope test of
   Grid_3d::Table::relax(threshold,cycleSize,queueSize), grid_3d.cpp:611-671 */
struct ora_queue_s { double v[4096]; int head, size; };

static int queue_slope_zero(ora_queue* q, double dev, unsigned int qsize, unsigned int cycles, double* slope_out)
{
	if (qsize == 0) return 0;
	/* std::deque push_back then pop_front while longer than queueSize */
	q->v[(q->head + q->size) % 4096] = dev;
	q->size++;
	while (q->size > (int) qsize) { q->head = (q->head + 1) % 4096; q->size--; }
	if (q->size != (int) qsize) return 0;
	const double xMean = (double) ((float) (qsize - 1) / 2.0f);
	double yMean = (double) 0.0f;
	for (int i = 0; i < q->size; i++) yMean += q->v[(q->head + i) % 4096];
	yMean /= qsize;
	double nom = (double) 0.0f;
	for (int x = 0; x < q->size; x++) nom += (x - xMean) * (q->v[(q->head + x) % 4096] - yMean);
	double den = qsize;
	den *= qsize;
	den -= 1;
	den *= qsize;
	den *= cycles;
	den /= 12.0f;
	const double slope = nom / den;
	*slope_out = slope;
	return slope == 0.0f;
}

long long ora_relax_threshold(ora_sys* s, double threshold, int cycles, int queue,
                              double* fault_dev, double* fault_slope, int* err)
{
	ora_queue q;
	q.head = 0; q.size = 0;
	long long iteration = 0;
	double dev;
	if (queue > 4096) { if (err) *err = ORA_E_ARG; return 0; }
	if (cycles == 1 && s->threads > 1)
	{
		ora_thr_ctl c;
		c.threshold = threshold; c.queue = (unsigned) queue; c.iteration = 0; c.fault = 0;
		c.fault_dev = 0.0; c.fault_slope = 0.0;
		int e = 0;
		run_relax_workers(s, 0, s->threads, &c, &q, NULL, &e);
		if (err) *err = e;
		if (e) return c.iteration;
		if (c.fault)
		{
			if (fault_dev) *fault_dev = c.fault_dev;
			if (fault_slope) *fault_slope = c.fault_slope;
			return -1 - c.iteration;
		}
		return c.iteration;
	}
	do
	{
		int e = 0;
		dev = DISPATCH(s, relax, s, cycles, &e);
		if (e) { if (err) *err = e; return iteration; }
		double slope = 0.0;
		if (queue_slope_zero(&q, dev, (unsigned) queue, (unsigned) cycles, &slope))
		{
			if (fault_dev) *fault_dev = dev;
			if (fault_slope) *fault_slope = slope;
			if (err) *err = 0;
			return -1 - iteration;
		}
		iteration += cycles;
	}
	while (dev > threshold);
	if (err) *err = 0;
	return iteration;
}

/* Grid_3d::Table::find_localNodes, grid_3d.cpp:1284-1315 (BFS levels into an
   ascending set; the primary node enters only as a neighbour's neighbour) */
int ora_find_local_nodes(const ora_sys* s, int node, int order, int* out, int cap)
{
	unsigned char* in = (unsigned char*) calloc((size_t) s->n, 1);
	int* cur = (int*) malloc(sizeof(int) * (size_t) s->n);
	int* next = (int*) malloc(sizeof(int) * (size_t) s->n);
	int ncur = 1, count = 0;
	cur[0] = node;
	for (int lvl = 0; lvl < order; lvl++)
	{
		int nn = 0;
		for (int u = 0; u < ncur; u++)
			for (int k = s->ptr[cur[u]]; k < s->ptr[cur[u] + 1]; k++)
			{
				const int v = s->idx[k];
				if (!in[v]) { in[v] = 1; next[nn++] = v; count++; }
			}
		int* t = cur; cur = next; next = t;
		ncur = nn;
	}
	int c = 0;
	for (int i = 0; i < s->n; i++)
		if (in[i]) { if (c < cap) out[c] = i; c++; }
	free(in); free(cur); free(next);
	return c;
}

static int local_nodes_nodir(const ora_sys* s, int node, int order, int** out)
{
	int* all = (int*) malloc(sizeof(int) * (size_t) s->n);
	const int c = ora_find_local_nodes(s, node, order, all, s->n);
	int u = 0;
	for (int k = 0; k < c; k++)
		if (!(s->flags[all[k]] & ORA_DIR)) all[u++] = all[k];
	*out = all;
	return u;
}

double ora_local_relax(ora_sys* s, int node, int order, int cycles, int* err)
{
	int* nodes;
	const int c = local_nodes_nodir(s, node, order, &nodes);
	int e = DISPATCH(s, do_local_relax, s, cycles, nodes, c);
	const double dev = DISPATCH(s, max_dev, s, nodes, c);
	free(nodes);
	if (err) *err = e;
	return dev;
}

/* Grid_3d::Table::localRelax(index,order,threshold,cycle,queue), grid_3d.cpp:707-795 */
long long ora_local_relax_threshold(ora_sys* s, int node, int order, double threshold,
                                    int cycles, int queue, int* err)
{
	int* nodes;
	const int c = local_nodes_nodir(s, node, order, &nodes);
	ora_queue q;
	q.head = 0; q.size = 0;
	long long iteration = 0;
	double dev;
	do
	{
		int e = DISPATCH(s, do_local_relax, s, cycles, nodes, c);
		if (e) { free(nodes); if (err) *err = e; return iteration; }
		dev = DISPATCH(s, max_dev, s, nodes, c);
		double slope;
		if (queue_slope_zero(&q, dev, (unsigned) queue, (unsigned) cycles, &slope))
		{
			free(nodes);
			if (err) *err = 0;
			return -1 - iteration;
		}
		iteration += cycles;
	}
	while (dev > threshold);
	free(nodes);
	if (err) *err = 0;
	return iteration;
}

/* ------------------------------------------------------------ fields */

int ora_field_o1(const ora_sys* s, int node, double out[3])
{
	if (s->mode == ORA_F32)
	{
		float v[3];
		const int e = field_o1_f32(s, node, v);
		for (int k = 0; k < 3; k++) out[k] = v[k];
		return e;
	}
	return field_o1_f64(s, node, out);
}

int ora_field_o2(const ora_sys* s, int node, double out[3], double* charge)
{
	if (s->mode == ORA_F32)
	{
		float v[3], q;
		const int e = field_o2_f32(s, node, v, &q);
		for (int k = 0; k < 3; k++) out[k] = v[k];
		if (charge) *charge = q;
		return e;
	}
	return field_o2_f64(s, node, out, charge);
}

double ora_voronoi_area(const ora_sys* s, int n1, int n2, double inf)
{
	int e = 0;
	return s->mode == ORA_F32 ? voronoi_area_f32(s, n1, n2, (float) inf, &e) : voronoi_area_f64(s, n1, n2, inf, &e);
}

double ora_voronoi_volume(const ora_sys* s, int node, double inf)
{
	int e = 0;
	return s->mode == ORA_F32 ? voronoi_volume_f32(s, node, (float) inf, &e) : voronoi_volume_f64(s, node, inf, &e);
}

double ora_distance(const ora_sys* s, int n1, int n2) { return DISPATCH(s, distance, s, n1, n2); }

/* ------------------------------------------------------------ surface */

#define R float
#define SQRT sqrtf
#define EXP expf
#define POW powf
#define FN(x) x##_f32
#include "oracle_surface.inc"
#undef R
#undef SQRT
#undef EXP
#undef POW
#undef FN
#define R double
#define SQRT sqrt
#define EXP exp
#define POW pow
#define FN(x) x##_f64
#include "oracle_surface.inc"
#undef R
#undef SQRT
#undef EXP
#undef POW
#undef FN

static int surf_find(const ora_sys* s, int node)
{
	int lo = 0, hi = s->nsurf;
	while (lo < hi)
	{
		const int mid = (lo + hi) / 2;
		if (s->surf[mid] < node) lo = mid + 1; else hi = mid;
	}
	return lo;
}

static int surf_has(const ora_sys* s, int node)
{
	const int pos = surf_find(s, node);
	return pos < s->nsurf && s->surf[pos] == node;
}

static void surf_insert(ora_sys* s, int node)
{
	const int pos = surf_find(s, node);
	if (pos < s->nsurf && s->surf[pos] == node) return; /* std::set::insert keeps the old entry */
	if (s->nsurf == s->capsurf)
	{
		s->capsurf = s->capsurf ? 2 * s->capsurf : 1024;
		s->surf = (int*) realloc(s->surf, sizeof(int) * (size_t) s->capsurf);
		s->prob = (double*) realloc(s->prob, sizeof(double) * (size_t) s->capsurf);
	}
	memmove(s->surf + pos + 1, s->surf + pos, sizeof(int) * (size_t) (s->nsurf - pos));
	memmove(s->prob + pos + 1, s->prob + pos, sizeof(double) * (size_t) (s->nsurf - pos));
	s->surf[pos] = node;
	s->prob[pos] = 0.0;
	s->nsurf++;
}

static int surf_erase(ora_sys* s, int node)
{
	const int pos = surf_find(s, node);
	if (pos >= s->nsurf || s->surf[pos] != node) return 0;
	memmove(s->surf + pos, s->surf + pos + 1, sizeof(int) * (size_t) (s->nsurf - pos - 1));
	memmove(s->prob + pos, s->prob + pos + 1, sizeof(double) * (size_t) (s->nsurf - pos - 1));
	s->nsurf--;
	return 1;
}

static int has_vacuum_face(const ora_sys* s, int i)
{
	for (int k = s->ptr[i]; k < s->ptr[i + 1]; k++)
	{
		const int j = s->idx[k];
		if (s->ids[j] == s->vacuum_id && ora_voronoi_area(s, i, j, 1.0) > 0.0) return 1;
	}
	return 0;
}

/* Surface_3d::Table::init, surface_3d.cpp:50-81 */
int ora_surface_init(ora_sys* s, int vacuum_id)
{
	s->vacuum_id = vacuum_id;
	s->nsurf = 0;
	for (int i = 0; i < s->n; i++)
	{
		if (s->ids[i] == vacuum_id) continue;
		if (!s->species[s->ids[i]].removable) continue;
		if (has_vacuum_face(s, i)) surf_insert(s, i);
	}
	return ORA_OK;
}

/* Surface_3d::Table::update, surface_3d.cpp:83-119 */
static int surface_update(ora_sys* s, int node)
{
	if (!surf_erase(s, node)) return fail(ORA_E_SURF, "Surface_3d::Table::update()");
	for (int k = s->ptr[node]; k < s->ptr[node + 1]; k++)
	{
		const int j = s->idx[k];
		if (s->ids[j] == s->vacuum_id) continue;
		if (sp_bad(s, s->ids[j])) return fail(ORA_E_CONFIG, "Configuration::Table::operator[]()");
		if (!s->species[s->ids[j]].removable) continue;
		if (has_vacuum_face(s, j)) surf_insert(s, j);
	}
	return ORA_OK;
}

int ora_surface_size(const ora_sys* s) { return s->nsurf; }
void ora_surface_nodes(const ora_sys* s, int* out) { memcpy(out, s->surf, sizeof(int) * (size_t) s->nsurf); }
void ora_surface_probabilities(const ora_sys* s, float* out)
{
	for (int k = 0; k < s->nsurf; k++) out[k] = (float) s->prob[k];
}

int ora_compute_probabilities(ora_sys* s, int mode) { return DISPATCH(s, probabilities, s, mode); }

float ora_specific_fields(ora_sys* s)
{
	int e = 0;
	return (float) DISPATCH(s, specific_fields, s, &e);
}

/* evapMaximum / evapMonteCarlo, surface_3d.cpp:423-450 */
int ora_find_candidate(const ora_sys* s, int mode)
{
	if (s->nsurf == 0) return -1;
	if (mode == 0)
	{
		int pick = 0;
		for (int k = 0; k < s->nsurf; k++)
			if (s->prob[pick] < s->prob[k]) pick = k;
		return s->surf[pick];
	}
	double thr = rand();
	thr /= RAND_MAX;
	int k = 0;
	double acc = (float) s->prob[0];
	while (thr > acc && k + 1 < s->nsurf)
	{
		k++;
		acc += (float) s->prob[k];
	}
	return s->surf[k];
}

/* evap_takeAway, surface_3d.cpp:507-564 (detached neighbours purged) */
int ora_take_away(ora_sys* s, int node)
{
	s->ids[node] = (short) s->vacuum_id;
	int e = ora_resync(s, node);
	if (e) return e;
	e = surface_update(s, node);
	if (e) return e;
	int lonely[256], nl = 0;
	for (int k = s->ptr[node]; k < s->ptr[node + 1]; k++)
	{
		const int j = s->idx[k];
		if (s->ids[j] == s->vacuum_id) continue;
		int flag = 1;
		for (int kk = s->ptr[j]; kk < s->ptr[j + 1]; kk++)
			if (s->ids[s->idx[kk]] != s->vacuum_id) { flag = 0; break; }
		if (flag && nl < 256) lonely[nl++] = j;
	}
	for (int q = 0; q < nl; q++)
	{
		s->ids[lonely[q]] = (short) s->vacuum_id;
		e = ora_resync(s, lonely[q]);
		if (e) return e;
		e = surface_update(s, lonely[q]);
		if (e) return e;
	}
	return ORA_OK;
}

/* ------------------------------------------------------------ point location */

void ora_srand(unsigned int seed) { srand(seed); }

/* update_coordination, the fork's patch (surface_3d.cpp:568-593): in node
   order, every id >= 10 grows by the number of neighbours whose id (already
   updated for lower indices) is >= 10; ids are a signed char and wrap */
void ora_update_coordination(ora_sys* s)
{
	for (int i = 0; i < s->n; i++)
	{
		if (s->ids[i] < 10) continue;
		int count = 0;
		for (int k = s->ptr[i]; k < s->ptr[i + 1]; k++)
			if (s->ids[s->idx[k]] >= 10) count++;
		s->ids[i] = (short) (signed char) (s->ids[i] + count);
	}
}

static const float* nc(const ora_sys* s, int i) { return s->xyz + 3 * (size_t) i; }

/* Geometry_3d::Table::tetOrient(n1,n2,n3,pos) = orient3d(p2,p1,p3,pos),
   geometry_3d.cpp:1750-1758; only its sign is used */
static int tet_orient_sign(const ora_sys* s, int n1, int n2, int n3, const float* pos)
{
	return ora_orient3d_sign(nc(s, n2), nc(s, n1), nc(s, n3), pos);
}

/* Geometry_3d::Table::tetGuess, geometry_3d.cpp:1712-1748 */
static int tet_guess(const ora_sys* s, const float* pos)
{
	int ref = -1;
	float refd = FLT_MAX;
	int cnt = (int) pow((double) s->m, 0.25);
	while (cnt-- > 0)
	{
		int t = rand();
		t %= s->m;
		for (int v = 0; v < 4; v++)
		{
			const float* c = nc(s, s->tv[4 * (size_t) t + v]);
			float d = 0.0f;
			for (int i = 0; i < 3; i++) { float tmp = c[i] - pos[i]; tmp *= tmp; d += tmp; }
			if (d < refd) { refd = d; ref = t; }
		}
	}
	return ref;
}

/* Geometry_3d::Table::findTetrahedron(pos, tetIndex), geometry_3d.cpp:988-1044:
   stochastic visibility walk, random first face from rand() % 4; a negative
   start tet returns -1 at once, without a draw */
int ora_find_tetrahedron(const ora_sys* s, const float pos[3], int t)
{
	for (int loop = 0; loop < s->m; loop++)
	{
		if (t < 0) return -1;
		const int minI = rand() % 4, maxI = minI + 4;
		int index;
		for (index = minI; index < maxI; index++)
		{
			const int* v = s->tv + 4 * (size_t) t;
			if (tet_orient_sign(s, v[index % 4], v[3 - (2 + index) % 4], v[(2 + index) % 4], pos) < 0)
			{
				t = s->tn[4 * (size_t) t + 3 - index % 4];
				break;
			}
		}
		if (index == maxI) return t;
	}
	return -2;
}

/* Geometry_3d::Table::findTetrahedron(pos), geometry_3d.cpp:982-986: the
   walk from tetGuess's sample */
int ora_locate(const ora_sys* s, const float pos[3])
{
	return ora_find_tetrahedron(s, pos, tet_guess(s, pos));
}

/* Geometry_3d::Table::tetVolume / tetWeights, geometry_3d.cpp:872-980 */
static float tet_volume(const ora_sys* s, int t)
{
	float lumat[9];
	const int* v = s->tv + 4 * (size_t) t;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++)
			lumat[i * 3 + j] = nc(s, v[i])[j] - nc(s, v[3])[j];
	float d;
	if (!ora_lu_decmp_fl(lumat, 3, 0, &d)) return 0.0f;
	float value = ora_lu_det_fl(lumat, 3, d);
	value /= 6.0f;
	return fabsf(value);
}

void ora_tet_weights(const ora_sys* s, int t, const float p[3], float w[4])
{
	const float volume = tet_volume(s, t);
	if (volume == 0.0f) { for (int i = 0; i < 4; i++) w[i] = 0.25f; return; }
	const int* v = s->tv + 4 * (size_t) t;
	for (int i = 0; i < 4; i++)
	{
		float lumat[9];
		for (int j = 0; j < 3; j++)
			for (int k = 0; k < 3; k++)
				lumat[j * 3 + k] = nc(s, v[(i + j) % 4])[k] - p[k];
		float d;
		if (!ora_lu_decmp_fl(lumat, 3, 0, &d)) { w[(i + 3) % 4] = 0.0f; continue; }
		float x = ora_lu_det_fl(lumat, 3, d);
		x /= 6.0f;
		x = fabsf(x);
		x /= volume;
		w[(i + 3) % 4] = x;
	}
}

static void node_field_f(const ora_sys* s, int node, float out[3])
{
	double e[3];
	ora_field_o1(s, node, e);
	for (int k = 0; k < 3; k++) out[k] = (float) e[k];
}

/* Grid_3d::potential(geom, grid, pos, guess), grid_3d.cpp:1585-1602 */
float ora_point_potential(const ora_sys* s, const float pos[3], int guess)
{
	const int t = ora_find_tetrahedron(s, pos, guess);
	if (t < 0) return 0.0f;
	float w[4];
	ora_tet_weights(s, t, pos, w);
	float value = 0.0f;
	for (int k = 0; k < 4; k++) value += w[k] * (float) ora_potential(s, s->tv[4 * (size_t) t + k]);
	return value;
}

/* Grid_3d::FastField::compute(pos, guess), grid_3d.cpp:1669-1721 (the
   node-field cache changes no value and is not restated) */
static int fast_field(const ora_sys* s, const float pos[3], int guess, float out[3])
{
	out[0] = out[1] = out[2] = 0.0f;
	const int t = ora_find_tetrahedron(s, pos, guess);
	if (t < 0) return t == -2 ? ORA_E_GEOM : ORA_OK;
	float f[4][3];
	for (int k = 0; k < 4; k++) node_field_f(s, s->tv[4 * (size_t) t + k], f[k]);
	float w[4];
	ora_tet_weights(s, t, pos, w);
	for (int k = 0; k < 4; k++)
		for (int c = 0; c < 3; c++)
			out[c] += w[k] * f[k][c];
	return ORA_OK;
}

int ora_point_field(const ora_sys* s, const float pos[3], int guess, float out[3])
{
	return fast_field(s, pos, guess, out);
}

/* ------------------------------------------------------------ trajectory */

typedef struct { float t; float p[3]; float v[3]; int tet; } phase;

enum { T_INVALID = -1, T_NO_INTEGRATION, T_NO_FIELD, T_STEPPER_LIMIT, T_ITERATION_LIMIT, T_GEOMETRY_LIMIT, T_SYSTEM_LIMIT };

typedef struct
{
	const ora_sys* s;
	float charge, mass;
	int integrator, stepper;
	float dt_min;
	phase* data;
	int len, cap;
	phase err_est;
	int root_diff;   /* steps whose glibc powf(x, 0.25f) differs from the correctly rounded fourth root */
} traj;

/* the step-size control's powf(x, 0.25f): glibc's, as the reference calls
   it (mode 0), or the correctly rounded fourth root that the device computes
   (mode 1); steps where the two differ are counted */
static int g_root_mode = 0;
void ora_set_root_mode(int mode) { g_root_mode = mode; }

static float quarter_pow(traj* tr, float x)
{
	const float p = powf(x, 0.25f);
	const float c = (float) sqrt(sqrt((double) x));
	if (p != c) tr->root_diff++;
	return g_root_mode ? c : p;
}

/* the error estimate's powf(x, 2.0f): glibc's, as the reference calls it
   (mode 0), or the product x * x that the device forms (mode 1); glibc's
   powf rounds x^2 correctly except, possibly, at exact ties */
static float square(float x)
{
	return g_root_mode ? x * x : powf(x, 2.0f);
}

/* Trajectory_3d::o5_integrator, trajectory_3d.cpp:583-676 (Cash-Karp tableau
   in float, Numerical Recipes 16.2) */
static phase o5_step(const traj* tr, float delta, phase* error)
{
	static const float b[6][5] =
	{
		{ 0.0f, 0.0f, 0.0f, 0.0f, 0.0f },
		{ 1.0f / 5.0f, 0.0f, 0.0f, 0.0f, 0.0f },
		{ 3.0f / 40.0f, 9.0f / 40.0f, 0.0f, 0.0f, 0.0f },
		{ 3.0f / 10.0f, -9.0f / 10.0f, 6.0f / 5.0f, 0.0f, 0.0f },
		{ -11.0f / 54.0f, 5.0f / 2.0f, -70.0f / 27.0f, 35.0f / 27.0f, 0.0f },
		{ 1631.0f / 55296.0f, 175.0f / 512.0f, 575.0f / 13824.0f, 44275.0f / 110592.0f, 253.0f / 4096.0f }
	};
	static const float c1[6] = { 37.0f / 378.0f, 0.0f, 250.0f / 621.0f, 125.0f / 594.0f, 0.0f, 512.0f / 1771.0f };
	static const float c2[6] = { 2825.0f / 27648.0f, 0.0f, 18575.0f / 48384.0f, 13525.0f / 55296.0f, 277.0f / 14336.0f, 1.0f / 4.0f };
	const phase* back = &tr->data[tr->len - 1];
	const float qm = tr->charge / tr->mass;
	float dv[6][3], dp[6][3];
	for (int i = 0; i < 6; i++)
	{
		float vel[3], pos[3];
		for (int l = 0; l < 3; l++) { vel[l] = back->v[l]; pos[l] = back->p[l]; }
		for (int j = 0; j < i; j++)
			for (int l = 0; l < 3; l++)
			{
				vel[l] += b[i][j] * dv[j][l];
				pos[l] += b[i][j] * dp[j][l];
			}
		fast_field(tr->s, pos, back->tet, dv[i]);
		for (int l = 0; l < 3; l++) dv[i][l] *= qm;
		for (int l = 0; l < 3; l++) dv[i][l] *= delta;
		for (int l = 0; l < 3; l++) dp[i][l] = vel[l] * delta;
	}
	phase next = *back;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 6; j++)
		{
			next.p[i] += c1[j] * dp[j][i];
			next.v[i] += c1[j] * dv[j][i];
		}
	next.t += delta;
	if (error)
	{
		error->t = 0.0f;
		for (int i = 0; i < 3; i++)
		{
			error->p[i] = 0.0f;
			error->v[i] = 0.0f;
			for (int j = 0; j < 6; j++)
			{
				error->p[i] += (c1[j] - c2[j]) * dp[j][i];
				error->v[i] += (c1[j] - c2[j]) * dv[j][i];
			}
		}
	}
	next.tet = ora_find_tetrahedron(tr->s, next.p, back->tet);
	return next;
}

/* Trajectory_3d::o1_integrator, trajectory_3d.cpp:549-581 */
static phase o1_step(const traj* tr, float delta, phase* error)
{
	const phase* back = &tr->data[tr->len - 1];
	float f[3];
	fast_field(tr->s, back->p, back->tet, f);
	for (int i = 0; i < 3; i++) f[i] *= tr->charge;
	for (int i = 0; i < 3; i++) f[i] /= tr->mass;
	phase next = *back;
	for (int i = 0; i < 3; i++)
	{
		next.v[i] += delta * f[i];
		next.p[i] += delta * back->v[i];
	}
	next.t += delta;
	if (error)
	{
		error->t = 0.0f;
		const float d2 = powf(delta, 2.0f);
		for (int i = 0; i < 3; i++) { error->p[i] = next.p[i] * d2; error->v[i] = next.v[i] * d2; }
	}
	next.tet = ora_find_tetrahedron(tr->s, next.p, back->tet);
	return next;
}

/* The classical fourth-order Runge-Kutta step (integrator 2; not in the
   reference, which offers the Euler and Cash-Karp integrators): slopes at 0,
   h/2, h/2 and h from the start, from the start plus half of the first and
   of the second slope, and from the start plus the third; the solution adds
   (k1 + 2 k2 + 2 k3 + k4) / 6.  The error estimate is the difference to the
   embedded third-order solution y + (k1 + 2 k2 + 2 k3 + k5) / 6, k5 the
   slope at the new point: (k4 - k5) / 6.  The field comes from a callback so
   that the pinning tests can drive the same step with an analytic field. */
typedef void (*field_fn)(const void* ctx, const float pos[3], int guess, float out[3]);

static phase rk4_step(field_fn field, const void* ctx, const phase* back, float qm, float delta, phase* error)
{
	static const float a[4] = { 0.0f, 0.5f, 0.5f, 1.0f };
	static const float c[4] = { 1.0f / 6.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 6.0f };
	float dv[5][3], dp[5][3];
	for (int i = 0; i < 4; i++)
	{
		float vel[3], pos[3];
		for (int l = 0; l < 3; l++) { vel[l] = back->v[l]; pos[l] = back->p[l]; }
		if (i > 0)
			for (int l = 0; l < 3; l++)
			{
				vel[l] += a[i] * dv[i - 1][l];
				pos[l] += a[i] * dp[i - 1][l];
			}
		field(ctx, pos, back->tet, dv[i]);
		for (int l = 0; l < 3; l++) dv[i][l] *= qm;
		for (int l = 0; l < 3; l++) dv[i][l] *= delta;
		for (int l = 0; l < 3; l++) dp[i][l] = vel[l] * delta;
	}
	phase next = *back;
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 4; j++)
		{
			next.p[i] += c[j] * dp[j][i];
			next.v[i] += c[j] * dv[j][i];
		}
	next.t += delta;
	field(ctx, next.p, back->tet, dv[4]);
	for (int l = 0; l < 3; l++) dv[4][l] *= qm;
	for (int l = 0; l < 3; l++) dv[4][l] *= delta;
	for (int l = 0; l < 3; l++) dp[4][l] = next.v[l] * delta;
	if (error)
	{
		error->t = 0.0f;
		for (int i = 0; i < 3; i++)
		{
			float e = dp[3][i];
			e -= dp[4][i];
			e /= 6.0f;
			error->p[i] = e;
			e = dv[3][i];
			e -= dv[4][i];
			e /= 6.0f;
			error->v[i] = e;
		}
	}
	return next;
}

static void mesh_field(const void* ctx, const float pos[3], int guess, float out[3])
{
	fast_field((const ora_sys*) ctx, pos, guess, out);
}

static phase o4_step(const traj* tr, float delta, phase* error)
{
	const phase* back = &tr->data[tr->len - 1];
	const float qm = tr->charge / tr->mass;
	phase next = rk4_step(mesh_field, tr->s, back, qm, delta, error);
	next.tet = ora_find_tetrahedron(tr->s, next.p, back->tet);
	return next;
}

/* analytic fields of the integrator's pinning tests: kind 0 uniform,
   E = (g0, g1, g2); kind 1 linear restoring, E = -(g0 x, g1 y, g2 z) */
typedef struct { int kind; float g[3]; } analytic_field;

static void analytic_eval(const void* ctx, const float pos[3], int guess, float out[3])
{
	(void) guess;
	const analytic_field* f = (const analytic_field*) ctx;
	for (int l = 0; l < 3; l++)
	{
		if (f->kind == 0) out[l] = f->g[l];
		else
		{
			float v = f->g[l];
			v *= pos[l];
			out[l] = -v;
		}
	}
}

int ora_rk4_analytic(int kind, const float g[3], float qm, const float p0[3], const float v0[3], float delta,
                     int steps, float* out, float* err)
{
	analytic_field f;
	f.kind = kind;
	for (int l = 0; l < 3; l++) f.g[l] = g[l];
	phase y;
	memset(&y, 0, sizeof(y));
	for (int l = 0; l < 3; l++) { y.p[l] = p0[l]; y.v[l] = v0[l]; }
	phase e;
	memset(&e, 0, sizeof(e));
	for (int s = 0; s < steps; s++) y = rk4_step(analytic_eval, &f, &y, qm, delta, &e);
	out[0] = y.t;
	for (int l = 0; l < 3; l++) { out[1 + l] = y.p[l]; out[4 + l] = y.v[l]; }
	for (int l = 0; l < 3; l++) { err[l] = e.p[l]; err[3 + l] = e.v[l]; }
	return ORA_OK;
}

static phase do_integration(const traj* tr, float delta, phase* error)
{
	if (tr->integrator == 2) return o4_step(tr, delta, error);
	return tr->integrator == 0 ? o1_step(tr, delta, error) : o5_step(tr, delta, error);
}

static float rel_error(const phase* value, const phase* error)
{
	float rel = 0.0f;
	for (int i = 0; i < 3; i++)
	{
		float tmp;
		if (value->p[i] != 0.0f) { tmp = error->p[i]; tmp /= value->p[i]; tmp = fabsf(tmp); if (rel < tmp) rel = tmp; }
		if (value->v[i] != 0.0f) { tmp = error->v[i]; tmp /= value->v[i]; tmp = fabsf(tmp); if (rel < tmp) rel = tmp; }
	}
	return rel;
}

/* adjacentTetrahedra, geometry_3d.cpp:1046-1096: tets sharing a vertex with t */
static int adjacent_tets(const ora_sys* s, int t, int* out, int cap)
{
	int c = 0;
	for (int q = 0; q < 4; q++)
	{
		const int v = s->tv[4 * (size_t) t + q];
		for (int k = s->nt_ptr[v]; k < s->nt_ptr[v + 1]; k++)
		{
			const int u = s->nt_idx[k];
			if (u == t) continue;
			int seen = 0;
			for (int z = 0; z < c; z++) if (out[z] == u) { seen = 1; break; }
			if (!seen && c < cap) out[c++] = u;
		}
	}
	return c;
}

static int in_list(const int* l, int n, int x) { for (int i = 0; i < n; i++) if (l[i] == x) return 1; return 0; }

static int is_boundary_tet(const ora_sys* s, int t)
{
	for (int q = 0; q < 4; q++) if (s->gbnd[s->tv[4 * (size_t) t + q]]) return 1;
	return 0;
}

/* the three steppers, trajectory_3d.cpp:358-532 */
static int stepper_step(traj* tr, float* delta, float err_limit, phase* error, phase* out)
{
	const float safety = 0.9f;
	int allowed[2048];
	int na = 0;
	if (tr->stepper == 2 || tr->stepper == 3)
		na = adjacent_tets(tr->s, tr->data[tr->len - 1].tet, allowed, 2048);
	for (;;)
	{
		/* NONE integrates with the step as given (trajectory_3d.cpp:155-158);
		   the other steppers raise it to the step limit first (:362,432,458) */
		if (tr->stepper != 0 && *delta < tr->dt_min) *delta = tr->dt_min;
		phase value = do_integration(tr, *delta, error);
		if (tr->stepper == 0) { *out = value; return 0; }
		if (tr->stepper == 2)
		{
			const int back = tr->data[tr->len - 1].tet;
			if (value.tet < 0 && is_boundary_tet(tr->s, back)) { *out = value; return 0; }
			if (in_list(allowed, na, value.tet)) { *out = value; return 0; }
			if (value.tet == back) *delta *= 1.1f;
			else
			{
				if (*delta == tr->dt_min) { *out = value; return 0; }
				*delta *= 0.9f;
			}
			continue;
		}
		if (tr->stepper == 3)
		{
			if (value.tet >= 0 && value.tet != tr->data[tr->len - 1].tet && !in_list(allowed, na, value.tet))
			{
				*delta *= 0.1f;
				continue;
			}
		}
		const float rel = rel_error(&value, error);
		if (rel < err_limit || *delta == tr->dt_min)
		{
			if (rel != 0.0f)
			{
				float tmp = *delta;
				tmp *= quarter_pow(tr, err_limit / rel);
				tmp *= safety;
				if (tmp < *delta) *delta *= 1.1f; else *delta = tmp;
			}
			*out = value;
			return 0;
		}
		float tmp = *delta;
		tmp *= quarter_pow(tr, err_limit / rel);
		tmp *= safety;
		if (tmp > *delta) *delta *= 0.9f; else *delta = tmp;
	}
}

static void traj_push(traj* tr, phase p)
{
	if (tr->len == tr->cap)
	{
		tr->cap = tr->cap ? tr->cap * 2 : 256;
		tr->data = (phase*) realloc(tr->data, sizeof(phase) * (size_t) tr->cap);
	}
	tr->data[tr->len++] = p;
}

/* Trajectory_3d::init / integrate(delta, errorLimit) / extrapolate,
   trajectory_3d.cpp:99-126,130-205,303-332 */
static int traj_run(traj* tr, const float start[3], const float* v0, int guess, float dt0, float err_limit, float zmax)
{
	phase p0;
	memset(&p0, 0, sizeof(p0));
	for (int k = 0; k < 3; k++) p0.p[k] = start[k];
	if (v0) for (int k = 0; k < 3; k++) p0.v[k] = v0[k];
	p0.tet = ora_find_tetrahedron(tr->s, start, guess);
	if (p0.tet < 0) return fail(ORA_E_GEOM, "Trajectory_3d::init(): initial point is outside");
	tr->len = 0;
	traj_push(tr, p0);
	memset(&tr->err_est, 0, sizeof(tr->err_est));
	int status;
	float f[3];
	fast_field(tr->s, tr->data[0].p, tr->data[0].tet, f);
	if (f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f) status = T_NO_FIELD;
	else
	{
		float delta = dt0;
		for (;;)
		{
			phase error, next;
			memset(&error, 0, sizeof(error));
			stepper_step(tr, &delta, err_limit, &error, &next);
			traj_push(tr, next);
			/* Trajectory_3d::cumulative_error, trajectory_3d.cpp:680-696 */
			for (int i = 0; i < 3; i++)
			{
				float a = square(tr->err_est.p[i]);
				a += square(error.p[i]);
				tr->err_est.p[i] = sqrtf(a);
				float b = square(tr->err_est.v[i]);
				b += square(error.v[i]);
				tr->err_est.v[i] = sqrtf(b);
			}
			if (tr->data[tr->len - 1].tet < 0) { status = T_SYSTEM_LIMIT; break; }
		}
	}
	/* extrapolate(Z, zmax) */
	{
		const phase* back = &tr->data[tr->len - 1];
		float dt = zmax;
		dt -= back->p[2];
		dt /= back->v[2];
		phase item;
		item.t = back->t;
		item.t += dt;
		for (int k = 0; k < 3; k++) item.p[k] = back->p[k];
		for (int k = 0; k < 3; k++) item.p[k] += dt * back->v[k];
		for (int k = 0; k < 3; k++) item.v[k] = back->v[k];
		item.tet = ora_find_tetrahedron(tr->s, item.p, back->tet);
		traj_push(tr, item);
		/* the error estimate follows the extrapolation (trajectory_3d.cpp:327) */
		if (dt > 0.0f)
			for (int k = 0; k < 3; k++)
			{
				float d = dt;
				d *= tr->err_est.v[k];
				tr->err_est.p[k] += d;
			}
	}
	return status;
}

/* evap_initialVelocity, surface_3d.cpp:472-502 (two uniform splits of
   2kT/m, then a random sign per axis; five rand() draws) */
static void initial_velocity(const ora_sys* s, const ora_species* sp, float v[3])
{
	double emax = 2.0 * 1.3806504e-23;
	emax *= s->temperature;
	emax /= sp->mass;
	emax /= 1.660538782e-27;
	float x = (float) rand();
	x /= (float) RAND_MAX;
	x = (float) (x * emax);
	float y = (float) rand();
	y /= (float) RAND_MAX;
	y = (float) (y * (emax - x));
	float z = (float) emax;
	z -= x;
	z -= y;
	v[0] = x; v[1] = y; v[2] = z;
	for (int i = 0; i < 3; i++)
	{
		v[i] = sqrtf(v[i]);
		if (rand() % 2) v[i] = -v[i];
	}
}

static float zmax_of(const ora_sys* s)
{
	/* Geometry_3d::Table::setExtents, geometry_3d.cpp:1509-1528 */
	float mx = s->xyz[2];
	for (int i = 1; i < s->n; i++)
		if (s->xyz[3 * (size_t) i + 2] > mx) mx = s->xyz[3 * (size_t) i + 2];
	return mx;
}

int ora_trajectory(const ora_sys* s, int node, float charge, float mass, int integrator,
                   int stepper, float dt0, float dt_min, float err_limit,
                   float* out, int cap, int* len, float* err)
{
	traj tr;
	memset(&tr, 0, sizeof(tr));
	tr.s = s; tr.charge = charge; tr.mass = mass; tr.integrator = integrator; tr.stepper = stepper;
	tr.dt_min = dt_min;
	const int status = traj_run(&tr, nc(s, node), NULL, s->assoc[node], dt0, err_limit, zmax_of(s));
	for (int k = 0; k < tr.len && k < cap; k++)
	{
		float* o = out + 8 * (size_t) k;
		o[0] = tr.data[k].t;
		for (int c = 0; c < 3; c++) { o[1 + c] = tr.data[k].p[c]; o[4 + c] = tr.data[k].v[c]; }
		o[7] = (float) tr.data[k].tet;
	}
	*len = tr.len;
	if (err)
		for (int k = 0; k < 3; k++) { err[k] = tr.err_est.p[k]; err[3 + k] = tr.err_est.v[k]; }
	free(tr.data);
	return status;
}

/* ------------------------------------------------------------ ion table */

/* splitmix64 (Steele, Lea and Flood 2014): the mixing function of one
   counter-based draw */
static unsigned long long splitmix64(unsigned long long x)
{
	unsigned long long z = x + 0x9e3779b97f4a7c15ull;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

/* u in [0, 1) from (seed, event index) with 53 random bits; the first entry
   whose cumulative weight (double, in table order) exceeds u times the
   total */
int ora_ion_pick(unsigned int seed, long long event_index, int count, const float* weight)
{
	const unsigned long long key = ((unsigned long long) seed << 32) ^ (unsigned long long) event_index;
	const double u = (double) (splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
	double total = 0.0;
	for (int k = 0; k < count; k++) total += (double) weight[k];
	const double target = u * total;
	double cum = 0.0;
	for (int k = 0; k < count; k++)
	{
		cum += (double) weight[k];
		if (target < cum) return k;
	}
	return count - 1;
}

int ora_set_ion_table(ora_sys* s, int species, int count, const float* mass, const int* charge, const float* weight)
{
	if (species < 0 || species >= 128 || count < 0 || count > 8) return fail(ORA_E_ARG, "ion table: species or count");
	for (int k = 0; k < count; k++)
		if (!(weight[k] >= 0.0f) || !(mass[k] > 0.0f)) return fail(ORA_E_ARG, "ion table: mass or weight");
	s->ion_n[species] = count;
	for (int k = 0; k < count; k++)
	{
		s->ion_mass[species][k] = mass[k];
		s->ion_q[species][k] = charge[k];
		s->ion_w[species][k] = weight[k];
	}
	return ORA_OK;
}

/* ------------------------------------------------------------ event loop */

static float flen3(const double e[3])
{
	float v = 0.0f;
	for (int k = 0; k < 3; k++) v += (float) e[k] * (float) e[k];
	return sqrtf(v);
}

/* ------------------------------------------------------------ field lines */

/* the field line's first step: powf(volume, 1/3) as the reference (mode 0)
   or correctly rounded (mode 1, the device's) */
static float cube_root(float x)
{
	return g_root_mode ? (float) pow((double) x, (double) (1.0f / 3.0f)) : powf(x, 1.0f / 3.0f);
}

/* Grid_3d::FastField::compute(pos) without a guess (grid_3d.cpp:1663-1667) */
static void field_unguessed(const ora_sys* s, const float p[3], float out[3])
{
	const int g = ora_locate(s, p);
	fast_field(s, p, g, out);
}

/* Fieldline_3d::init + compute(distance) (fieldline_3d.cpp:55-130) with the
   stepper (:155-190); out holds 7 floats per point (position, field, tet).
   Returns the Fieldline_3d status, or ORA_E_GEOM when the start lies
   outside the mesh (init() throws there). */
int ora_fieldline(const ora_sys* s, const float start[3], float distance, unsigned long iterations,
                  unsigned long stepper_limit, float* out, int cap, int* len)
{
	enum { NO_FIELD = 1, STEPPER_LIMIT = 2, ITERATION_LIMIT = 3, GEOMETRY_LIMIT = 4, SYSTEM_LIMIT = 5 };
	typedef struct { float p[3], f[3]; int tet; } pt;
	int n = 0, ncap = 64;
	pt* d = (pt*) malloc(sizeof(pt) * (size_t) ncap);
	pt p0;
	for (int c = 0; c < 3; c++) p0.p[c] = start[c];
	field_unguessed(s, p0.p, p0.f);
	p0.tet = ora_locate(s, p0.p);
	if (p0.tet < 0) { free(d); *len = 0; return fail(ORA_E_GEOM, "Fieldline_3d::init()"); }
	d[n++] = p0;
	float delta = tet_volume(s, p0.tet);
	delta = cube_root(delta);
	delta /= 2.0f;
	int status = ITERATION_LIMIT;
	int* allowed = (int*) malloc(sizeof(int) * 4096);
	unsigned long it = iterations;
	while (it--)
	{
		const pt back = d[n - 1];
		if (back.f[0] == 0.0f && back.f[1] == 0.0f && back.f[2] == 0.0f) { status = NO_FIELD; break; }
		const int na = adjacent_tets(s, back.tet, allowed, 4096);
		float len2 = 0.0f;
		for (int c = 0; c < 3; c++) len2 += back.f[c] * back.f[c];
		const float flen = sqrtf(len2);
		float dir[3];
		for (int c = 0; c < 3; c++) dir[c] = back.f[c] / flen;
		int found = 0;
		pt v;
		unsigned long st = stepper_limit;
		while (st--)
		{
			for (int c = 0; c < 3; c++) v.p[c] = back.p[c] + delta * dir[c];
			v.tet = ora_find_tetrahedron(s, v.p, back.tet);
			if ((v.tet < 0 && is_boundary_tet(s, back.tet)) || (v.tet >= 0 && in_list(allowed, na, v.tet)))
			{
				field_unguessed(s, v.p, v.f);
				found = 1;
				break;
			}
			if (v.tet == back.tet) delta *= 1.1f;
			else delta *= 0.9f;
		}
		if (!found) { status = STEPPER_LIMIT; break; }
		if (v.tet < 0) { status = SYSTEM_LIMIT; break; }
		if (distance >= 0.0f)
		{
			float q2 = 0.0f;
			for (int c = 0; c < 3; c++) { const float e = d[0].p[c] - v.p[c]; q2 += e * e; }
			if (sqrtf(q2) > distance) { status = GEOMETRY_LIMIT; break; }
		}
		if (n == ncap) { ncap *= 2; d = (pt*) realloc(d, sizeof(pt) * (size_t) ncap); }
		d[n++] = v;
	}
	for (int k = 0; k < n && k < cap; k++)
	{
		for (int c = 0; c < 3; c++) { out[7 * k + c] = d[k].p[c]; out[7 * k + 3 + c] = d[k].f[c]; }
		out[7 * k + 6] = (float) d[k].tet;
	}
	*len = n;
	free(allowed);
	free(d);
	return status;
}

/* ---- surface normal: Surface_3d::Table::normal (surface_3d.cpp:123-169),
   Geometry_3d::find_best_plane (geometry_3d.cpp:1806-1870), and utils.h's
   Numerical Recipes householder (tred2, utils.h:245-330) and tri_eigen_qli
   (tqli, utils.h:358-456) for n = 3, float with the reference's promotions */

static float nr_hyp(float a, float b)
{
	float x = fabsf(a), y = fabsf(b);
	if (x > y) { float t = y / x; return (float) (x * sqrt(1.0 + t * t)); }
	if (y == 0.0f) return 0.0f;
	float t = x / y;
	return (float) (y * sqrt(1.0 + t * t));
}

static void nr_tred2(float a[3][3], float d[3], float e[3])
{
	for (int i = 2; i >= 1; i--)
	{
		int l = i - 1;
		float h = 0.0f, scale = 0.0f;
		if (l >= 1)
		{
			for (int k = 0; k <= l; k++) scale += fabsf(a[i][k]);
			if (scale == 0.0f) e[i] = a[i][l];
			else
			{
				for (int k = 0; k <= l; k++) { a[i][k] /= scale; h += a[i][k] * a[i][k]; }
				float f = a[i][l];
				float g = (f >= 0.0f) ? -sqrtf(h) : sqrtf(h);
				e[i] = scale * g;
				h -= f * g;
				a[i][l] = f - g;
				f = 0.0f;
				for (int j = 0; j <= l; j++)
				{
					a[j][i] = a[i][j] / h;
					g = 0.0f;
					for (int k = 0; k <= j; k++) g += a[j][k] * a[i][k];
					for (int k = j + 1; k <= l; k++) g += a[k][j] * a[i][k];
					e[j] = g / h;
					f += e[j] * a[i][j];
				}
				float hh = f / (h + h);
				for (int j = 0; j <= l; j++)
				{
					f = a[i][j];
					e[j] = g = e[j] - hh * f;
					for (int k = 0; k <= j; k++) a[j][k] -= (f * e[k] + g * a[i][k]);
				}
			}
		}
		else e[i] = a[i][l];
		d[i] = h;
	}
	d[0] = 0.0f;
	e[0] = 0.0f;
	for (int i = 0; i < 3; i++)
	{
		if (d[i] != 0.0f)
			for (int j = 0; j < i; j++)
			{
				float g = 0.0f;
				for (int k = 0; k < i; k++) g += a[i][k] * a[k][j];
				for (int k = 0; k < i; k++) a[k][j] -= g * a[k][i];
			}
		d[i] = a[i][i];
		a[i][i] = 1.0f;
		for (int j = 0; j < i; j++) { a[j][i] = 0.0f; a[i][j] = 0.0f; }
	}
}

/* returns 0, or ORA_E_NAN when an eigenvalue does not settle in 10^4 sweeps
   (the reference never throws there and would keep iterating) */
static int nr_tqli(float d[3], float e[3], float z[3][3])
{
	e[0] = e[1]; e[1] = e[2]; e[2] = 0.0f;
	for (int l = 0; l < 3; l++)
	{
		int iter = 0, m;
		for (;;)
		{
			for (m = l; m < 2; m++)
			{
				float dd = fabsf(d[m]) + fabsf(d[m + 1]);
				if (fabsf(e[m]) + dd == dd) break;
			}
			if (m == l) break;
			if (++iter > 10000) return ORA_E_NAN;
			float g = (float) ((d[l + 1] - d[l]) / (2.0 * e[l]));
			float r = nr_hyp(g, 1.0f);
			g = d[m] - d[l] + e[l] / (g + (g >= 0.0f ? fabsf(r) : -fabsf(r)));
			float s = 1.0f, c = 1.0f, p = 0.0f;
			int i = m - 1, early = 0;
			for (; i >= l; i--)
			{
				float f = s * e[i], b = c * e[i];
				e[i + 1] = r = nr_hyp(f, g);
				if (r == 0.0f) { d[i + 1] -= p; e[m] = 0.0f; early = 1; break; }
				s = f / r;
				c = g / r;
				g = d[i + 1] - p;
				r = (float) ((d[i] - g) * s + 2.0 * c * b);
				d[i + 1] = g + (s * r);
				p = s * r;
				g = c * r - b;
				for (int k = 0; k < 3; k++)
				{
					f = z[k][i + 1];
					z[k][i + 1] = s * z[k][i] + c * f;
					z[k][i] = c * z[k][i] - s * f;
				}
			}
			if (early) continue;
			d[l] -= p;
			e[l] = g;
			e[m] = 0.0f;
		}
	}
	return ORA_OK;
}

int ora_best_plane(const float* off, int count, float out[3])
{
	float a[3][3], d[3] = {0, 0, 0}, e[3] = {0, 0, 0};
	for (int p = 0; p < 3; p++)
		for (int q = p; q < 3; q++)
		{
			float acc = 0.0f;
			for (int i = 0; i < count; i++) acc += off[3 * i + p] * off[3 * i + q];
			a[p][q] = acc;
			a[q][p] = acc;
		}
	nr_tred2(a, d, e);
	if (nr_tqli(d, e, a)) return fail(ORA_E_NAN, "find_best_plane: no convergence");
	int col = 0;
	if (d[1] < d[col]) col = 1;
	if (d[2] < d[col]) col = 2;
	for (int i = 0; i < 3; i++)
	{
		out[i] = a[i][col];
		if (out[i] != out[i]) return fail(ORA_E_NAN, "Geometry_3d::find_best_plane()");
	}
	if (out[0] != 0.0f || out[1] != 0.0f || out[2] != 0.0f)
	{
		float len = 0.0f;
		for (int i = 0; i < 3; i++) len += out[i] * out[i];
		len = sqrtf(len);
		for (int i = 0; i < 3; i++) out[i] /= len;
	}
	return ORA_OK;
}

int ora_surface_normal(const ora_sys* s, int node, float out[3])
{
	out[0] = out[1] = out[2] = 0.0f;
	if (!surf_has(s, node)) return ORA_OK;
	int* shell = (int*) malloc(sizeof(int) * (size_t) s->n);
	int c = ora_find_local_nodes(s, node, 3, shell, s->n);
	float* off = (float*) malloc(sizeof(float) * 3 * (size_t) (c + 1));
	float vac[3] = {0.0f, 0.0f, 0.0f};
	int noff = 0;
	const float* p0 = s->xyz + 3 * (size_t) node;
	for (int k = 0; k < c; k++)
	{
		const int v = shell[k];
		if (v == node) continue;  /* adjacentNodes(.., limit) erases the primary node */
		const float* p = s->xyz + 3 * (size_t) v;
		if (s->ids[v] == s->vacuum_id)
			for (int q = 0; q < 3; q++) vac[q] += p[q] - p0[q];
		if (surf_has(s, v))
		{
			for (int q = 0; q < 3; q++) off[3 * noff + q] = p0[q] - p[q];
			noff++;
		}
	}
	int e = ora_best_plane(off, noff, out);
	free(off); free(shell);
	if (e) return e;
	float dot = 0.0f;
	for (int q = 0; q < 3; q++) dot += out[q] * vac[q];
	if (dot < 0.0f)
		for (int q = 0; q < 3; q++) out[q] *= -1.0f;
	return ORA_OK;
}

static int surface_species_bad(const ora_sys* s)
{
	for (int k = 0; k < s->nsurf; k++)
		if (sp_bad(s, s->ids[s->surf[k]])) return 1;
	return 0;
}

/* Process::evaporation, process.cpp:441-1124, steps I-VI (logging and dumps
   are not restated; update_coordination is off) */
int ora_evaporate(ora_sys* s, const ora_evap_opts* o, int nevents, ora_event* out, int* done)
{
	*done = 0;
	int e;
	if (s->vacuum_id != o->vacuum_id || s->nsurf == 0) ora_surface_init(s, o->vacuum_id);
	if (surface_species_bad(s)) return fail(ORA_E_CONFIG, "Configuration::Table::operator[]()");
	ora_specific_fields(s);
	e = ora_compute_probabilities(s, o->prob_mode);
	if (e) return e;
	float* vq = s->vq;
	const float zmax = o->detector_z > 0.0f ? o->detector_z : zmax_of(s);
	if (!s->apex0_set)
	{
		float apex0 = -FLT_MAX;
		for (int k = 0; k < s->nsurf; k++)
			if (s->xyz[3 * (size_t) s->surf[k] + 2] > apex0) apex0 = s->xyz[3 * (size_t) s->surf[k] + 2];
		s->apex0 = apex0;
		s->apex0_set = 1;
	}
	const float apex0 = s->apex0;
	if (o->voltage_queue > 4096) return fail(ORA_E_ARG, "voltage queue too long");
	if (o->init_event_cnt < 0) return fail(ORA_E_ARG, "negative initial event count");
	if (s->event_cnt < 0) s->event_cnt = o->init_event_cnt;
	/* the fork applies update_coordination once after the initial surface
	   evaluation (process.cpp:645-647) and after every take-away (:770) */
	if (o->update_coordination && !s->coord_done) { ora_update_coordination(s); s->coord_done = 1; }
	for (int ev = 0; ev < nevents && s->nsurf > 0; ev++)
	{
		ora_event* r = &out[ev];
		memset(r, 0, sizeof(*r));
		r->event_index = ++s->event_cnt;
		const float scalingReference = (float) s->scaling_ref;
		/* I) candidate */
		const int node = ora_find_candidate(s, o->evap_mode);
		const int pos = surf_find(s, node);
		r->node = node;
		r->id = s->ids[node];
		r->probability = (float) s->prob[pos];
		r->pot_before = (float) ora_potential(s, node);
		double fb[3];
		ora_field_o1(s, node, fb);
		for (int k = 0; k < 3; k++) r->field_before[k] = (float) fb[k];
		{
			int top = s->surf[0];
			for (int k = 0; k < s->nsurf; k++)
				if (s->xyz[3 * (size_t) s->surf[k] + 2] > s->xyz[3 * (size_t) top + 2]) top = s->surf[k];
			for (int k = 0; k < 3; k++) r->apex[k] = s->xyz[3 * (size_t) top + k];
		}
		e = ora_surface_normal(s, node, r->normal);
		if (e) return e;
		const int nodeId = s->ids[node];
		/* II) take away */
		e = ora_take_away(s, node);
		if (e) return e;
		if (o->update_coordination) ora_update_coordination(s);
		/* III) relaxation */
		r->local_iter = ora_local_relax_threshold(s, node, o->relax_shell, (double) o->relax_threshold,
		                                          o->relax_cycle, o->relax_queue, &e);
		if (e) return e;
		if (o->relax_global_cycles > 0) { ora_relax(s, o->relax_global_cycles, &e); if (e) return e; }
		/* process.cpp:876 adds initEventCnt to the already offset eventIndex,
		   in unsigned int */
		if (o->refresh_interval > 0 &&
		    ((unsigned int) o->init_event_cnt + (unsigned int) r->event_index) % (unsigned int) o->refresh_interval == 0)
		{
			double fd, fs;
			r->refresh_iter = ora_relax_threshold(s, (double) o->refresh_threshold, o->refresh_cycle,
			                                      o->refresh_queue, &fd, &fs, &e);
			if (e) return e;
		}
		/* IV) fields and probabilities */
		if (surface_species_bad(s)) return fail(ORA_E_CONFIG, "Configuration::Table::operator[]()");
		ora_specific_fields(s);
		e = ora_compute_probabilities(s, o->prob_mode);
		if (e) return e;
		r->surface_size = s->nsurf;
		r->scaling_ref = (float) s->scaling_ref;
		/* V) trajectory */
		traj tr;
		memset(&tr, 0, sizeof(tr));
		int have_back = 0;
		float backp[3] = { 0, 0, 0 };
		int backt = -1;
		if (o->trajectory)
		{
			if (sp_bad(s, nodeId)) return fail(ORA_E_CONFIG, "Configuration::Table::operator[]()");
			const ora_species* sp = &s->species[nodeId];
			float ion_m = sp->mass;
			int ion_q = sp->evap_charge;
			if (o->ion_table && s->ion_n[nodeId] > 0)
			{
				const int k = ora_ion_pick(o->ion_seed, r->event_index, s->ion_n[nodeId], s->ion_w[nodeId]);
				ion_m = s->ion_mass[nodeId][k];
				ion_q = s->ion_q[nodeId][k];
			}
			r->ion_mass = ion_m;
			r->ion_charge = ion_q;
			float charge = (float) ion_q;
			charge *= 1.60217733e-19f;
			float mass = ion_m;
			mass *= 1.660538782e-27f;
			tr.s = s; tr.charge = charge; tr.mass = mass;
			tr.integrator = o->integrator; tr.stepper = o->stepper; tr.dt_min = o->dt_min;
			float v0[3] = { 0.0f, 0.0f, 0.0f };
			if (o->thermal_velocity) initial_velocity(s, sp, v0);
			r->traj_status = traj_run(&tr, nc(s, node), v0, s->assoc[node], o->dt0, o->err_limit, zmax);
			r->traj_len = tr.len;
			r->traj_root_diff = tr.root_diff;
			for (int k = 0; k < 3; k++) { r->err_pos[k] = tr.err_est.p[k]; r->err_vel[k] = tr.err_est.v[k]; }
			for (int k = 0; k < 3; k++) { r->start[k] = tr.data[0].p[k]; r->stop[k] = tr.data[tr.len - 1].p[k]; }
			for (int k = tr.len - 1; k >= 0; k--)
				if (tr.data[k].tet >= 0) { have_back = 1; backt = tr.data[k].tet; for (int c = 0; c < 3; c++) backp[c] = tr.data[k].p[c]; break; }
		}
		/* VI) voltage and time scale */
		r->pot_after = (float) ora_potential(s, node);
		double fa[3];
		ora_field_o1(s, node, fa);
		for (int k = 0; k < 3; k++) r->field_after[k] = (float) fa[k];
		const float flen = flen3(fb);
		if (have_back && flen != 0.0f)
		{
			float v = r->pot_before;
			v -= ora_point_potential(s, backp, backt);
			v *= s->species[nodeId].evap_field;
			v /= flen;
			v /= scalingReference;
			vq[(s->vq_head + s->vq_size) % 4096] = v;
			s->vq_size++;
			s->vq_sum += v;
			while (s->vq_size > (int) o->voltage_queue) { s->vq_sum -= vq[s->vq_head]; s->vq_head = (s->vq_head + 1) % 4096; s->vq_size--; }
			r->voltage = s->vq_sum;
			r->voltage /= (float) s->vq_size;
			float ts = r->pot_after;
			ts -= ora_point_potential(s, backp, backt);
			ts /= r->voltage;
			r->time_scale = sqrtf(ts);
		}
		else if (flen == 0.0f) { r->voltage = 1.0f; r->time_scale = 1.0f; }
		else { r->voltage = 0.0f; r->time_scale = 1.0f; }
		r->tof = o->trajectory ? tr.data[tr.len - 1].t * r->time_scale : 0.0f;
		free(tr.data);
		*done = ev + 1;
		if (apex0 - r->apex[2] >= o->shrink_limit) break;
	}
	return ORA_OK;
}
