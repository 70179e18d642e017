// Row gathers over the SELL-64 row layout shared by the relaxation kernels.
//
// A row's entries are summed strictly in entry order (ascending neighbour
// index), so the value does not depend on how the loads are scheduled.
//
// Short rows (at most kLongRow entries) are taken by their own lane, with
// the loads of the whole row issued before the ordered additions (the batch
// is 8, 16 or kLongRow entries, chosen per wavefront from its longest row;
// rows are sorted by length within small windows so that a wavefront's
// rows have similar lengths).  Rows
// longer than that (the fan of a boundary-plane node under the tip base
// reaches several thousand entries) are kept on a separate list and each is
// taken by a whole wavefront in extra workgroups of the same launch: the
// lanes load 256 entries at a time and every lane adds the products in entry
// order through v_readlane.  A long row on the per-lane path would keep one
// wavefront looping for the length of the row after the rest of the launch
// has finished.  The bulk sweeps read a long row from a contiguous copy of
// its entries (in the SELL layout consecutive entries of one row are 64
// elements apart, so every entry would cost a cache line of its own); the
// SELL copy stays for the kernels that take rows one by one.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace tsg {
namespace dev {

constexpr int kLongWindows = 4;  // 64-entry windows loaded per round on the long path

__device__ __forceinline__ int lane_id() { return (int) (threadIdx.x & 63); }

__device__ __forceinline__ int wave_max_int(int v)
{
	for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
	return v;
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
	const unsigned long long b = (unsigned long long) __double_as_longlong(v);
	const unsigned lo = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) b, l);
	const unsigned hi = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) (b >> 32), l);
	return __longlong_as_double((long long) (((unsigned long long) hi << 32) | lo));
}

__device__ __forceinline__ float readlane_f32(float v, int l)
{
	return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Workgroups of a row launch.  The long rows come first, one wavefront
// each, in a multiple of eight workgroups: they are the slowest wavefronts,
// so they start at once and spread over all XCDs.  The slot workgroups
// follow.  The hardware deals workgroups to the eight XCDs in turn
// (workgroup b runs on XCD b % 8), each XCD with its own L2; the slot
// workgroups are padded to a multiple of eight and XCD x takes the
// contiguous range [x q, (x+1) q) of them, so each L2 holds one contiguous
// share of the rows and of the potentials they gather.
constexpr int kXcds = 8;
struct RowBlock {
	bool is_long;
	int index;  // long: first long row of the workgroup; slot: slot workgroup
};
__device__ __forceinline__ int long_groups(int nlong) { return ((nlong + 3) / 4 + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ RowBlock row_block(int nlong)
{
	const int lg = long_groups(nlong);
	RowBlock r;
	const int b = (int) blockIdx.x;
	if (b < lg) {
		r.is_long = true;
		r.index = b * 4;
	} else {
		const int bm = b - lg;
		const int q = ((int) gridDim.x - lg) / kXcds;
		r.is_long = false;
		r.index = (bm % kXcds) * q + bm / kXcds;
	}
	return r;
}

// the long row handled by this wavefront (wave-uniform)
__device__ __forceinline__ int long_row_index(const RowBlock& rb)
{
	return __builtin_amdgcn_readfirstlane(rb.index + (int) (threadIdx.x >> 6));
}

// sum_k (double) p[col_k] * (double) c_k, added in entry order starting from
// 0.0, each product rounded to double on its own (no fused multiply-add):
// the reference's computePotential accumulation.  Per-lane path, deg <= kLongRow.
template <typename R, int B>
__device__ __forceinline__ double short_row_dot_exact_b(const int* __restrict__ col, const R* __restrict__ coup,
                                                        const R* __restrict__ p, int deg, long long base, int k0,
                                                        double lv)
{
	// entries k0 .. k0+B-1 continuing the running sum lv; loads are
	// unconditional (entries past the end re-read the last one) so that the
	// compiler issues the whole batch before the first wait
	int j[B];
	R c[B], v[B];
	const int last = max(deg - 1, 0);
#pragma unroll
	for (int u = 0; u < B; u++) j[u] = col[base + 64ll * min(k0 + u, last)];
#pragma unroll
	for (int u = 0; u < B; u++) c[u] = coup[base + 64ll * min(k0 + u, last)];
#pragma unroll
	for (int u = 0; u < B; u++) v[u] = p[j[u]];
#pragma unroll
	for (int u = 0; u < B; u++) {
		double adj = (double) v[u];
		adj *= (double) c[u];
		const double t = lv + adj;
		lv = (k0 + u < deg) ? t : lv;
	}
	return lv;
}

// deg <= kLongRow on every active lane; `active` false lanes load nothing.
// The batch is chosen per wavefront from its longest row (wave-uniform);
// double rows of more than 16 entries take a 16-entry batch and an 8-entry
// one (a single 24-entry batch of doubles needs more registers than the
// compiler keeps for one round of loads)
template <typename R>
__device__ __forceinline__ double short_row_dot_exact(const int* __restrict__ col, const R* __restrict__ coup,
                                                      const R* __restrict__ p, bool active, int deg, long long base)
{
	const int wmax = wave_max_int(active ? deg : 0);
	if (!active || wmax == 0) return 0.0;
	if (wmax <= 8) return short_row_dot_exact_b<R, 8>(col, coup, p, deg, base, 0, 0.0);
	const double lv = short_row_dot_exact_b<R, 16>(col, coup, p, deg, base, 0, 0.0);
	if (wmax <= 16) return lv;
	return short_row_dot_exact_b<R, kLongRow - 16>(col, coup, p, deg, base, 16, lv);
}

// the same sum for one long row taken by the whole wavefront from its
// contiguous entries off .. off+deg-1 (deg and off wave-uniform; blocks of
// 256 threads); every lane returns the sum.  The lanes load and multiply
// 64 W entries at a time and stage the products in LDS; lane 0 then adds
// them in entry order, 16 per round with the loads of a round issued
// first.  A round past the row's end adds -0.0, the exact identity of the
// IEEE addition (s + -0.0 == s for every s, signed zeros included), so the
// sum is the sequential one bit for bit while the loop count stays a
// multiple of 16.  (A sum through v_readlane costs about three times as
// many cycles per entry, and the longest row -- a few thousand entries --
// bounds the whole sweep.)
template <typename R>
__device__ __forceinline__ double long_row_dot_exact(const int* __restrict__ col, const R* __restrict__ coup,
                                                     const R* __restrict__ p, int deg, long long off)
{
	constexpr int W = kLongWindows;
	constexpr int U = 16;
	static_assert((64 * W) % U == 0, "the staged window holds whole rounds");
	__shared__ double s_prod[4][64 * W];
	double* sp = s_prod[(threadIdx.x >> 6) & 3];
	deg = __builtin_amdgcn_readfirstlane(deg);
	double acc = 0.0;
	for (int k0 = 0; k0 < deg; k0 += 64 * W) {
		int j[W];
		R c[W], v[W];
#pragma unroll
		for (int w = 0; w < W; w++) {
			const long long e = off + min(k0 + 64 * w + lane_id(), deg - 1);
			j[w] = col[e];
			c[w] = coup[e];
		}
#pragma unroll
		for (int w = 0; w < W; w++) v[w] = p[j[w]];
#pragma unroll
		for (int w = 0; w < W; w++) {
			double prod = (double) v[w];
			prod *= (double) c[w];
			sp[64 * w + lane_id()] = (k0 + 64 * w + lane_id() < deg) ? prod : -0.0;
		}
		// the window's products are visible to lane 0 (LDS operations of one
		// wavefront complete in order; the fence keeps the compiler's order)
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		__builtin_amdgcn_wave_barrier();
		if (lane_id() == 0) {
			const int rounds = (min(64 * W, deg - k0) + U - 1) / U;
			for (int r = 0; r < rounds; r++) {
				double t[U];
#pragma unroll
				for (int u = 0; u < U; u++) t[u] = sp[r * U + u];
#pragma unroll
				for (int u = 0; u < U; u++) acc += t[u];
			}
		}
		// the next window overwrites the staging only after lane 0 read it
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		__builtin_amdgcn_wave_barrier();
	}
	return readlane_f64(acc, 0);
}

// sum_k w_k * x[col_k] in FP32 as a chain of fused multiply-adds in entry
// order; per-lane path, deg <= kLongRow
// entries k0 .. k0+B-1 continuing the running sum s (loads of entries past
// the end re-read the last one, so that the batch is issued before the
// first wait)
template <int B>
__device__ __forceinline__ float short_row_fma_f32_b(const int* __restrict__ col, const float* __restrict__ w,
                                                     const float* __restrict__ x, int deg, long long base, int k0,
                                                     float s)
{
	int j[B];
	float c[B], v[B];
	const int last = max(deg - 1, 0);
#pragma unroll
	for (int u = 0; u < B; u++) j[u] = col[base + 64ll * min(k0 + u, last)];
#pragma unroll
	for (int u = 0; u < B; u++) c[u] = w[base + 64ll * min(k0 + u, last)];
#pragma unroll
	for (int u = 0; u < B; u++) v[u] = x[j[u]];
#pragma unroll
	for (int u = 0; u < B; u++) {
		const float t = fmaf(c[u], v[u], s);
		s = (k0 + u < deg) ? t : s;
	}
	return s;
}

// rows of more than 16 entries take a 16-entry batch and an 8-entry one
// (one 24-entry batch holds registers that cost a third of the occupancy)
__device__ __forceinline__ float short_row_fma_f32(const int* __restrict__ col, const float* __restrict__ w,
                                                   const float* __restrict__ x, bool active, int deg, long long base)
{
	const int wmax = wave_max_int(active ? deg : 0);
	if (!active || wmax == 0) return 0.0f;
	if (wmax <= 8) return short_row_fma_f32_b<8>(col, w, x, deg, base, 0, 0.0f);
	const float s = short_row_fma_f32_b<16>(col, w, x, deg, base, 0, 0.0f);
	if (wmax <= 16) return s;
	return short_row_fma_f32_b<kLongRow - 16>(col, w, x, deg, base, 16, s);
}

// the long row's contiguous entries off .. off+deg-1: the FP32 sum of a
// correction row (one wavefront; every lane returns the sum)
__device__ __forceinline__ float long_row_fma_f32(const int* __restrict__ col, const float* __restrict__ w,
                                                  const float* __restrict__ x, int deg, long long off)
{
	constexpr int W = kLongWindows;
	float acc = 0.0f;
	for (int k0 = 0; k0 < deg; k0 += 64 * W) {
		int j[W];
		float c[W], v[W];
#pragma unroll
		for (int ww = 0; ww < W; ww++) {
			const long long e = off + min(k0 + 64 * ww + lane_id(), deg - 1);
			j[ww] = col[e];
			c[ww] = w[e];
		}
#pragma unroll
		for (int ww = 0; ww < W; ww++) v[ww] = x[j[ww]];
#pragma unroll
		for (int ww = 0; ww < W; ww++) {
			const int cnt = min(64, deg - (k0 + 64 * ww));
			for (int u = 0; u < cnt; u++) acc = fmaf(readlane_f32(c[ww], u), readlane_f32(v[ww], u), acc);
		}
	}
	return acc;
}

// G x over the FP32 correction rows for this thread's slot (or long row):
// returns the slot (-1 when none) and the sum.  The correction vectors are
// slot-indexed (dense over the unknown rows) and the rows' columns are the
// neighbours' slots; the per-slot record and the slot's extra operands
// (e1[slot], e2[slot], either may be null) are loaded in one round before
// the row
__device__ __forceinline__ int accel_row(const AccelArgs& x, const float* __restrict__ in, float* sum,
                                         const float* __restrict__ e1, float* v1, const float* __restrict__ e2, float* v2)
{
	const int main_blocks = (x.nslots + 255) >> 8;
	const RowBlock rb = row_block(x.nlong);
	const int lb = rb.is_long ? -1 : rb.index;
	int slot = -1, node = -1, deg = 0, ideg = 0;
	long long soff = 0, loff = 0;
	if (!rb.is_long) {
		if (lb < main_blocks) slot = lb * 256 + threadIdx.x;
		if (slot >= 0 && slot < x.nslots) {
			node = x.slot_node[slot];
			deg = x.slot_deg[slot];
			ideg = x.inner_deg[slot];
			soff = x.slice_off[slot >> 6];
		}
	} else {
		const int li = long_row_index(rb);
		if (li < x.nlong) {
			slot = x.long_slots[li];
			node = x.slot_node[slot];
			ideg = x.inner_deg[slot];
			loff = x.long_off[li];
		}
	}
	const long long base = soff + (slot & 63);
	*sum = 0.0f;
	if (!rb.is_long) {  // block-uniform; the row sum needs the whole wavefront
		const bool act = node >= 0 && deg <= kLongRow;
		if (act) {
			if (e1) *v1 = e1[slot];
			if (e2) *v2 = e2[slot];
		}
		*sum = short_row_fma_f32(x.col, x.w, in, act, ideg, base);
		return act ? slot : -1;
	}
	if (node < 0) return -1;
	if (e1) *v1 = e1[slot];
	if (e2) *v2 = e2[slot];
	*sum = long_row_fma_f32(x.licol, x.liw, in, ideg, loff);
	return lane_id() == 0 ? slot : -1;
}

}  // namespace dev
}  // namespace tsg
