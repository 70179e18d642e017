// Exact geometric predicates on float coordinates.
//
// Both predicates are evaluated first in double with a forward error bound
// and, only when the bound cannot decide the sign, exactly: every float is
// m * 2^e with a 24-bit integer m, so after scaling all inputs of one call to
// the smallest exponent they are integers and the determinants are formed in
// fixed-width two's complement big integers (1024 bits covers coordinates
// spanning more than 2^90 in magnitude).  Written from the definition of the
// determinants (cofactor expansion); the reference uses Shewchuk's adaptive
// predicates (predicates/predicates.c) for the same signs.
#include "predicates.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace tsg {
namespace {

constexpr int kMaxLimbs = 16;
// active limb count of the current exact evaluation (set from the exponent
// range of its inputs; sums and products stay below 2^(64*nl-1))
thread_local int kLimbs = kMaxLimbs;

struct Big {
	uint64_t w[kMaxLimbs];
};

inline void big_zero(Big* r) { std::memset(r->w, 0, sizeof(uint64_t) * kLimbs); }

inline void big_neg(Big* r, const Big& a)
{
	uint64_t carry = 1;
	for (int i = 0; i < kLimbs; i++) {
		const uint64_t x = ~a.w[i] + carry;
		carry = (carry && x == 0) ? 1 : 0;
		r->w[i] = x;
	}
}

inline void big_add(Big* r, const Big& a, const Big& b)
{
	unsigned __int128 c = 0;
	for (int i = 0; i < kLimbs; i++) {
		c += (unsigned __int128) a.w[i] + b.w[i];
		r->w[i] = (uint64_t) c;
		c >>= 64;
	}
}

inline void big_sub(Big* r, const Big& a, const Big& b)
{
	Big nb;
	big_neg(&nb, b);
	big_add(r, a, nb);
}

inline int big_sign(const Big& a)
{
	if (a.w[kLimbs - 1] >> 63) return -1;
	for (int i = 0; i < kLimbs; i++)
		if (a.w[i]) return 1;
	return 0;
}

void big_mul(Big* r, const Big& a, const Big& b)
{
	Big x = a, y = b;
	const int sa = big_sign(x), sb = big_sign(y);
	if (sa < 0) big_neg(&x, x);
	if (sb < 0) big_neg(&y, y);
	uint64_t out[kMaxLimbs] = {0};
	for (int i = 0; i < kLimbs; i++) {
		if (x.w[i] == 0) continue;
		unsigned __int128 carry = 0;
		for (int j = 0; i + j < kLimbs; j++) {
			const unsigned __int128 t = (unsigned __int128) x.w[i] * y.w[j] + out[i + j] + carry;
			out[i + j] = (uint64_t) t;
			carry = t >> 64;
		}
	}
	std::memcpy(r->w, out, sizeof(uint64_t) * kLimbs);
	if ((sa < 0) != (sb < 0)) big_neg(r, *r);
}

void big_from_float(Big* r, float v, int emin)
{
	big_zero(r);
	if (v == 0.0f) return;
	int e;
	const float fm = std::frexp(v, &e);
	const int64_t m = (int64_t) std::ldexp(fm, 24);
	const int shift = (e - 24) - emin;
	const uint64_t um = (uint64_t) (m < 0 ? -m : m);
	const int limb = shift / 64, bit = shift % 64;
	if (limb < kLimbs) r->w[limb] |= um << bit;
	if (bit != 0 && limb + 1 < kLimbs) r->w[limb + 1] |= um >> (64 - bit);
	if (m < 0) big_neg(r, *r);
}

// smallest ulp exponent of the inputs; sets the active limb count so that a
// product of `degree` differences (plus carries) fits
int min_exponent(const float* const* p, int np, int degree)
{
	int emin = 1 << 30, emax = -(1 << 30);
	for (int i = 0; i < np; i++)
		for (int k = 0; k < 3; k++)
			if (p[i][k] != 0.0f) {
				int e;
				std::frexp(p[i][k], &e);
				if (e - 24 < emin) emin = e - 24;
				if (e > emax) emax = e;
			}
	if (emin == (1 << 30)) { kLimbs = 2; return 0; }
	const int bits = (emax - emin) + 2;          // magnitude bits of one difference
	const int total = degree * bits + 8;         // products, squares and sums
	kLimbs = std::min(kMaxLimbs, total / 64 + 2);
	return emin;
}

// det of rows (a, b, c) in big integers
void big_det3(Big* r, const Big a[3], const Big b[3], const Big c[3])
{
	Big t1, t2, minor, term;
	big_zero(r);
	big_mul(&t1, b[1], c[2]); big_mul(&t2, b[2], c[1]); big_sub(&minor, t1, t2);
	big_mul(&term, a[0], minor); big_add(r, *r, term);
	big_mul(&t1, b[0], c[2]); big_mul(&t2, b[2], c[0]); big_sub(&minor, t1, t2);
	big_mul(&term, a[1], minor); big_sub(r, *r, term);
	big_mul(&t1, b[0], c[1]); big_mul(&t2, b[1], c[0]); big_sub(&minor, t1, t2);
	big_mul(&term, a[2], minor); big_add(r, *r, term);
}

}  // namespace

int orient_sign(const float* a, const float* b, const float* c, const float* d)
{
	const double bx = (double) b[0] - a[0], by = (double) b[1] - a[1], bz = (double) b[2] - a[2];
	const double cx = (double) c[0] - a[0], cy = (double) c[1] - a[1], cz = (double) c[2] - a[2];
	const double dx = (double) d[0] - a[0], dy = (double) d[1] - a[1], dz = (double) d[2] - a[2];
	const double m1 = cy * dz - cz * dy, m2 = cx * dz - cz * dx, m3 = cx * dy - cy * dx;
	const double det = bx * m1 - by * m2 + bz * m3;
	const double perm = std::fabs(bx) * (std::fabs(cy * dz) + std::fabs(cz * dy)) +
	                    std::fabs(by) * (std::fabs(cx * dz) + std::fabs(cz * dx)) +
	                    std::fabs(bz) * (std::fabs(cx * dy) + std::fabs(cy * dx));
	const double bound = 4e-15 * perm;
	if (det > bound) return 1;
	if (det < -bound) return -1;
	if (perm == 0.0) return 0;
	const float* pts[4] = {a, b, c, d};
	const int emin = min_exponent(pts, 4, 3);
	Big v[4][3];
	for (int i = 0; i < 4; i++)
		for (int k = 0; k < 3; k++) big_from_float(&v[i][k], pts[i][k], emin);
	Big rows[3][3];
	for (int i = 0; i < 3; i++)
		for (int k = 0; k < 3; k++) big_sub(&rows[i][k], v[i + 1][k], v[0][k]);
	Big det3;
	big_det3(&det3, rows[0], rows[1], rows[2]);
	return big_sign(det3);
}

// > 0 iff e lies strictly inside the circumsphere of the positively oriented
// (orient_sign(a,b,c,d) > 0) tetrahedron abcd.
int insphere_sign(const float* a, const float* b, const float* c, const float* d, const float* e)
{
	const float* p[4] = {a, b, c, d};
	double r[4][3], w[4], pw[4];
	for (int i = 0; i < 4; i++) {
		for (int k = 0; k < 3; k++) r[i][k] = (double) p[i][k] - e[k];
		w[i] = r[i][0] * r[i][0] + r[i][1] * r[i][1] + r[i][2] * r[i][2];
		pw[i] = w[i];
	}
	auto det3 = [&](int i, int j, int k, double* perm) {
		const double m1 = r[j][1] * r[k][2] - r[j][2] * r[k][1];
		const double m2 = r[j][0] * r[k][2] - r[j][2] * r[k][0];
		const double m3 = r[j][0] * r[k][1] - r[j][1] * r[k][0];
		*perm = std::fabs(r[i][0]) * (std::fabs(r[j][1] * r[k][2]) + std::fabs(r[j][2] * r[k][1])) +
		        std::fabs(r[i][1]) * (std::fabs(r[j][0] * r[k][2]) + std::fabs(r[j][2] * r[k][0])) +
		        std::fabs(r[i][2]) * (std::fabs(r[j][0] * r[k][1]) + std::fabs(r[j][1] * r[k][0]));
		return r[i][0] * m1 - r[i][1] * m2 + r[i][2] * m3;
	};
	double p0, p1, p2, p3;
	const double M0 = det3(1, 2, 3, &p0), M1 = det3(0, 2, 3, &p1), M2 = det3(0, 1, 3, &p2), M3 = det3(0, 1, 2, &p3);
	// det4 of rows [r_i, w_i] expanded along the last column
	const double det = -w[0] * M0 + w[1] * M1 - w[2] * M2 + w[3] * M3;
	const double perm = pw[0] * p0 + pw[1] * p1 + pw[2] * p2 + pw[3] * p3;
	const double bound = 1e-14 * perm;
	// det4 < 0 <=> inside for a positive tet in this orientation convention
	if (det < -bound) return 1;
	if (det > bound) return -1;
	if (perm == 0.0) return 0;
	const float* pts[5] = {a, b, c, d, e};
	const int emin = min_exponent(pts, 5, 5);
	Big v[5][3];
	for (int i = 0; i < 5; i++)
		for (int k = 0; k < 3; k++) big_from_float(&v[i][k], pts[i][k], emin);
	Big rr[4][3], ww[4];
	for (int i = 0; i < 4; i++) {
		for (int k = 0; k < 3; k++) big_sub(&rr[i][k], v[i][k], v[4][k]);
		Big t;
		big_zero(&ww[i]);
		for (int k = 0; k < 3; k++) {
			big_mul(&t, rr[i][k], rr[i][k]);
			big_add(&ww[i], ww[i], t);
		}
	}
	Big M[4];
	big_det3(&M[0], rr[1], rr[2], rr[3]);
	big_det3(&M[1], rr[0], rr[2], rr[3]);
	big_det3(&M[2], rr[0], rr[1], rr[3]);
	big_det3(&M[3], rr[0], rr[1], rr[2]);
	Big acc, t;
	big_zero(&acc);
	big_mul(&t, ww[0], M[0]); big_sub(&acc, acc, t);
	big_mul(&t, ww[1], M[1]); big_add(&acc, acc, t);
	big_mul(&t, ww[2], M[2]); big_sub(&acc, acc, t);
	big_mul(&t, ww[3], M[3]); big_add(&acc, acc, t);
	return -big_sign(acc);
}

}  // namespace tsg
