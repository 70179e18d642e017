// Device-side restatements of the reference's small numerical kernels.
// Every routine follows the reference's operation order exactly (the library
// is compiled with -ffp-contract=off and IEEE division/sqrt), so float
// instances reproduce the reference's float results and double instances the
// FP64 promotion bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace tsg {
namespace dev {

template <typename R>
struct V3 {
	R x[3];
};

template <typename R>
__device__ __forceinline__ V3<R> v3(R a, R b, R c)
{
	V3<R> r;
	r.x[0] = a; r.x[1] = b; r.x[2] = c;
	return r;
}

template <typename R>
__device__ __forceinline__ V3<R> load_coord(const float* __restrict__ xyz, int i)
{
	return v3<R>((R) xyz[3 * (size_t) i], (R) xyz[3 * (size_t) i + 1], (R) xyz[3 * (size_t) i + 2]);
}

template <typename R>
__device__ __forceinline__ V3<R> vsub(const V3<R>& a, const V3<R>& b)
{
	V3<R> r;
	for (int k = 0; k < 3; k++) r.x[k] = a.x[k] - b.x[k];
	return r;
}

template <typename R>
__device__ __forceinline__ V3<R> vadd(const V3<R>& a, const V3<R>& b)
{
	V3<R> r;
	for (int k = 0; k < 3; k++) r.x[k] = a.x[k] + b.x[k];
	return r;
}

// scalar * vector, vector/mathVector.h:125-139
template <typename R>
__device__ __forceinline__ V3<R> vscale(R f, const V3<R>& a)
{
	V3<R> r;
	for (int k = 0; k < 3; k++) r.x[k] = a.x[k] * f;
	return r;
}

// vector/mathVector.h:114-122: ((0 + a0 b0) + a1 b1) + a2 b2
template <typename R>
__device__ __forceinline__ R vdot(const V3<R>& a, const V3<R>& b)
{
	R v = (R) 0;
	for (int k = 0; k < 3; k++) v += a.x[k] * b.x[k];
	return v;
}

template <typename R>
__device__ __forceinline__ R vlen(const V3<R>& a)
{
	return sqrt(vdot(a, a));
}

// vector/mathVector_3d.h:63-77
template <typename R>
__device__ __forceinline__ V3<R> vcross(const V3<R>& a, const V3<R>& b)
{
	V3<R> r;
	for (int i = 0; i < 3; i++) {
		const int k = (1 + i) % 3, j = (2 + i) % 3;
		r.x[i] = a.x[k] * b.x[j];
		r.x[i] -= a.x[j] * b.x[k];
	}
	return r;
}

// Numerical Recipes LU decomposition with partial pivoting as restated in
// tapsim/utils.h:52-140 (including the double reciprocal of the pivot).
// Every loop is unrolled and the row exchange is written as conditional
// swaps with fixed indices, so that the matrix stays in registers (an
// exchange through a computed row index would place it in scratch memory);
// the arithmetic and its order are those of the reference.
template <typename T, int n>
__device__ __forceinline__ bool lu_decmp(T* lu, int* index, T* d)
{
	T parity = (T) 1.0;
	T vv[n];
	int ii[n];
#pragma unroll
	for (int i = 0; i < n; i++) {
		T big = (T) 0.0;
#pragma unroll
		for (int j = 0; j < n; j++) {
			const T tmp = fabs(lu[i * n + j]);
			if (tmp > big) big = tmp;
		}
		if (big == (T) 0.0) return false;
		vv[i] = (T) 1.0;
		vv[i] /= big;
	}
	int imax = 0;
#pragma unroll
	for (int j = 0; j < n; j++) {
#pragma unroll
		for (int i = 0; i < j; i++) {
			T sum = lu[i * n + j];
#pragma unroll
			for (int k = 0; k < i; k++) sum -= lu[i * n + k] * lu[k * n + j];
			lu[i * n + j] = sum;
		}
		T big = (T) 0.0;
#pragma unroll
		for (int i = j; i < n; i++) {
			T sum = lu[i * n + j];
#pragma unroll
			for (int k = 0; k < j; k++) sum -= lu[i * n + k] * lu[k * n + j];
			lu[i * n + j] = sum;
			const T dum = vv[i] * fabs(sum);
			if (dum >= big) { big = dum; imax = i; }
		}
#pragma unroll
		for (int r = j + 1; r < n; r++) {
			if (r == imax) {
#pragma unroll
				for (int k = 0; k < n; k++) {
					const T t = lu[r * n + k];
					lu[r * n + k] = lu[j * n + k];
					lu[j * n + k] = t;
				}
				parity *= (T) -1.0;
				vv[r] = vv[j];
			}
		}
		ii[j] = imax;
		if (lu[j * n + j] == (T) 0.0) return false;
		const T dum = (T) (1.0 / (double) lu[j * n + j]);
#pragma unroll
		for (int i = j + 1; i < n; i++) lu[i * n + j] *= dum;
	}
	if (index)
#pragma unroll
		for (int i = 0; i < n; i++) index[i] = ii[i];
	if (d) *d = parity;
	return true;
}

// tapsim/utils.h:163-195, with the permuted load and store of b and the
// first-nonzero start of the forward sum as fixed-index selects
template <typename T, int n>
__device__ __forceinline__ void lu_solve(const T* lu, const int* index, T* b)
{
	int ii = 0;
#pragma unroll
	for (int i = 0; i < n; i++) {
		const int ip = index[i];
		T sum = b[0];
#pragma unroll
		for (int r = 1; r < n; r++)
			if (r == ip) sum = b[r];
#pragma unroll
		for (int r = 0; r < n; r++)
			if (r == ip) b[r] = b[i];
		if (ii) {
#pragma unroll
			for (int j = 0; j < i; j++)
				if (j >= ii - 1) sum -= lu[i * n + j] * b[j];
		} else if (sum != (T) 0)
			ii = i + 1;
		b[i] = sum;
	}
#pragma unroll
	for (int i = n - 1; i >= 0; i--) {
		T sum = b[i];
#pragma unroll
		for (int j = i + 1; j < n; j++) sum -= lu[i * n + j] * b[j];
		b[i] = sum / lu[i * n + i];
	}
}

}  // namespace dev
}  // namespace tsg
