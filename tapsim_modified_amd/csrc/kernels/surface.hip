// Surface evaluation: node fields by least squares (Grid_3d::Table::field_o1,
// grid_3d.cpp:909-991), the evaporation score |E| / F_evap of every surface
// site (Surface_3d::evap_compute_specificFields / probLinear_field,
// surface_3d.cpp:194-219, 266-297), the first-maximum choice
// (evapMaximum, surface_3d.cpp:423-431) and the sequential normaliser.
#include "device_math.hpp"
#include "kernels.hpp"
#include "field.hpp"

namespace tsg {
namespace dev {

__device__ __forceinline__ float exp_r(float x) { return expf(x); }
__device__ __forceinline__ double exp_r(double x) { return exp(x); }

// probVoronoiFluxForce (surface_3d.cpp:337-400): Maxwell stress over the
// node's Voronoi faces between media of different permittivity
template <typename R>
__device__ R flux_force(const FieldArgs& f, int i)
{
	const R* coup = (const R*) f.coup;
	const R* phi = (const R*) f.phi;
	const R eps0 = (R) 8.85418781762e-12;
	R area = 0;
	V3<R> force = v3<R>(0, 0, 0);
	const R pi = phi[i];
	const V3<R> ci = load_coord<R>(f.xyz, i);
	for (int kk = f.ptr[i]; kk < f.ptr[i + 1]; kk++) {
		const int j = f.idx[kk];
		R v = phi[j];
		v -= pi;
		v *= v;
		v = (R) (v * (-1.0 * eps0));
		v *= coup[kk];
		const R e1 = (R) (1.0 / f.epsilon[i]);
		const R e2 = (R) (1.0 / f.epsilon[j]);
		v *= e2 - e1;
		const V3<R> cj = load_coord<R>(f.xyz, j);
		R corr = coup[kk];
		corr *= vlen(vsub(ci, cj));
		corr = (R) (corr / 2.0);
		corr *= e1 + e2;
		area += corr;
		V3<R> lf = vsub(cj, ci);
		const R dd = vdot(lf, lf);
		for (int c = 0; c < 3; c++) lf.x[c] /= dd;
		for (int c = 0; c < 3; c++) lf.x[c] *= v;
		for (int c = 0; c < 3; c++) force.x[c] += lf.x[c];
	}
	R value = vlen(force);
	value /= area;
	return value;
}

template <typename R>
__global__ __launch_bounds__(256) void k_surface_fields(FieldArgs f)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= f.count) return;
	const int node = f.nodes[s];
	R E[3];
	if (!field_o1<R>(node, f.xyz, f.ptr, f.idx, (const R*) f.coup, f.flags, (const R*) f.phi, E)) {
		atomicOr(f.status, 1);
		E[0] = E[1] = E[2] = (R) 0;
	}
	if (E[0] != E[0] || E[1] != E[1] || E[2] != E[2]) atomicOr(f.status, 2);
	const R flen = vlen(v3<R>(E[0], E[1], E[2]));
	R ef = flen;
	ef /= (R) f.evap_field[node];
	R value = ef;
	if (f.mode == 1) {  // probBoltzmann (surface_3d.cpp:223-262)
		const R kB = (R) 8.6173431e-5f;
		R pr = (R) -1.0;
		pr *= flen;
		pr /= (R) f.evap_field[node];
		pr = (R) (pr + 1.0);
		pr *= (R) f.evap_energy[node];
		pr /= (R) f.temperature;
		pr /= kB;
		pr = (R) (pr / -2.0);
		value = exp_r(pr);
	} else if (f.mode == 2) {  // probLinear_force (surface_3d.cpp:300-333)
		R e2[3], q;
		const int rc = field_o2<R>(node, f.xyz, f.ptr, f.idx, f.flags, (const R*) f.phi, e2, &q);
		if (rc) atomicOr(f.status, rc == 1 ? 4 : 2);
		const R fs = vlen(v3<R>(e2[0], e2[1], e2[2]));
		const R F = (R) f.evap_field[node];
		value = q;
		value *= fs;
		value /= F * F;
	} else if (f.mode == 3) {
		const R F = (R) f.evap_field[node];
		value = flux_force<R>(f, node);
		value /= F * F;
	}
	((R*) f.value)[s] = value;
	if (f.value_ef) ((R*) f.value_ef)[s] = ef;
	if (f.field)
		for (int i = 0; i < 3; i++) f.field[3 * (size_t) s + i] = (float) E[i];
}

template <typename R>
__global__ void k_node_fields(int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                              const R* coup, const unsigned char* flags, const R* phi, double* out, int* status)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= count) return;
	R E[3];
	if (!field_o1<R>(nodes[s], xyz, ptr, idx, coup, flags, phi, E)) atomicOr(status, 1);
	for (int i = 0; i < 3; i++) out[3 * (size_t) s + i] = (double) E[i];
}

// field_o2 (quadratic least squares) of a list of nodes: -grad phi and the
// charge -eps0 laplacian; status bit 2 for a singular system, bit 1 for NaN
template <typename R>
__global__ void k_node_fields_o2(int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                                 const unsigned char* flags, const R* phi, double* out, double* charge, int* status)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= count) return;
	R E[3], q;
	const int rc = field_o2<R>(nodes[s], xyz, ptr, idx, flags, phi, E, &q);
	if (rc == 1) atomicOr(status, 4);
	if (rc == 2) atomicOr(status, 2);
	for (int i = 0; i < 3; i++) out[3 * (size_t) s + i] = (double) E[i];
	charge[s] = (double) q;
}

// first maximum in slot order (slots are in ascending node order); a slot
// below the winner whose value lies within a few ulps of the maximum could
// tie with it after the reference's float normalisation (v / sum), which
// sets the near-tie flag for the host to resolve with the exact normaliser
template <typename R>
__global__ __launch_bounds__(1024) void k_argmax(int count, const R* __restrict__ v, long long* out, double* maxval)
{
	__shared__ R sv[1024];
	__shared__ int ss[1024];
	R best = (R) -1;
	int bs = 0x7fffffff;
	for (int s = threadIdx.x; s < count; s += blockDim.x) {
		const R x = v[s];
		if (x > best) { best = x; bs = s; }  // increasing s: ties keep the first
	}
	sv[threadIdx.x] = best;
	ss[threadIdx.x] = bs;
	__syncthreads();
	for (int h = blockDim.x / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h) {
			const R o = sv[threadIdx.x + h];
			const int os = ss[threadIdx.x + h];
			if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && os < ss[threadIdx.x])) {
				sv[threadIdx.x] = o;
				ss[threadIdx.x] = os;
			}
		}
		__syncthreads();
	}
	const R vmax = sv[0];
	const int w = ss[0];
	__shared__ int near;
	if (threadIdx.x == 0) near = 0;
	__syncthreads();
	const R eps = (sizeof(R) == 4) ? (R) (8.0 / 16777216.0) : (R) (8.0 / 9007199254740992.0);
	const R lim = vmax - vmax * eps;
	for (int s = threadIdx.x; s < w; s += blockDim.x)
		if (v[s] >= lim) near = 1;
	__syncthreads();
	if (threadIdx.x == 0) {
		out[0] = w;
		out[1] = near;
		*maxval = (double) vmax;
	}
}

// sequential sum in slot order, reproducing the reference's running float
// (or double) sum exactly: one thread, loads issued 32 at a time ahead of the
// ordered additions
template <typename R>
__global__ __launch_bounds__(64) void k_seq_sum(int count, const R* __restrict__ v, double* out)
{
	if (threadIdx.x != 0) return;
	R sum = (R) 0;
	int s = 0;
	for (; s + 32 <= count; s += 32) {
		R x[32];
#pragma unroll
		for (int u = 0; u < 32; u++) x[u] = v[s + u];
#pragma unroll
		for (int u = 0; u < 32; u++) sum += x[u];
	}
	for (; s < count; s++) sum += v[s];
	*out = (double) sum;
}

}  // namespace dev

int launch_surface_fields(hipStream_t st, const FieldArgs& f, bool f64)
{
	if (f.count <= 0) return 0;
	if (f64) hipLaunchKernelGGL(dev::k_surface_fields<double>, dim3((f.count + 255) / 256), dim3(256), 0, st, f);
	else hipLaunchKernelGGL(dev::k_surface_fields<float>, dim3((f.count + 255) / 256), dim3(256), 0, st, f);
	return (int) hipGetLastError();
}

int launch_surface_argmax(hipStream_t st, int count, const void* value, bool f64, long long* out, double* maxval)
{
	if (f64) hipLaunchKernelGGL(dev::k_argmax<double>, dim3(1), dim3(1024), 0, st, count, (const double*) value, out, maxval);
	else hipLaunchKernelGGL(dev::k_argmax<float>, dim3(1), dim3(1024), 0, st, count, (const float*) value, out, maxval);
	return (int) hipGetLastError();
}

int launch_surface_sum(hipStream_t st, int count, const void* value, bool f64, double* out)
{
	if (f64) hipLaunchKernelGGL(dev::k_seq_sum<double>, dim3(1), dim3(64), 0, st, count, (const double*) value, out);
	else hipLaunchKernelGGL(dev::k_seq_sum<float>, dim3(1), dim3(64), 0, st, count, (const float*) value, out);
	return (int) hipGetLastError();
}

int launch_node_fields_o2(hipStream_t st, int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                          const unsigned char* flags, const void* phi, bool f64, double* out, double* charge, int* status)
{
	if (count <= 0) return 0;
	if (f64)
		hipLaunchKernelGGL(dev::k_node_fields_o2<double>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz,
		                   ptr, idx, flags, (const double*) phi, out, charge, status);
	else
		hipLaunchKernelGGL(dev::k_node_fields_o2<float>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz,
		                   ptr, idx, flags, (const float*) phi, out, charge, status);
	return (int) hipGetLastError();
}

int launch_node_fields(hipStream_t st, int count, const int* nodes, const float* xyz, const int* ptr, const int* idx,
                       const void* coup, const unsigned char* flags, const void* phi, bool f64, double* out, int* status)
{
	if (count <= 0) return 0;
	if (f64)
		hipLaunchKernelGGL(dev::k_node_fields<double>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz, ptr,
		                   idx, (const double*) coup, flags, (const double*) phi, out, status);
	else
		hipLaunchKernelGGL(dev::k_node_fields<float>, dim3((count + 63) / 64), dim3(64), 0, st, count, nodes, xyz, ptr,
		                   idx, (const float*) coup, flags, (const float*) phi, out, status);
	return (int) hipGetLastError();
}

}  // namespace tsg
