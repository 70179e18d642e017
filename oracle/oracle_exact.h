/*
 * Exact orientation sign for float coordinates (oracle side).
 *
 * The reference decides point location with Shewchuk's adaptive exact
 * orient3d evaluated on float input (predicates/predicates.c:1740-2393,
 * called through Geometry_3d::Table::tetOrient, tapsim/geometry_3d.cpp:1757).
 * Only the SIGN is used by the reference, so the oracle computes the sign of
 *     det[a-d; b-d; c-d]
 * with a double evaluation guarded by an error bound and, when the bound is
 * not conclusive, an exact evaluation in fixed-width big integers: every
 * float is m*2^e with a 24-bit m, so scaling all twelve inputs to the
 * smallest exponent gives integers below 2^280, whose 3x3 determinant fits in
 * 1024 bits.  Written from the definition of the determinant; no code from
 * the reference's predicates file.
 */
#ifndef ORACLE_EXACT_H
#define ORACLE_EXACT_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#define ORA_BN_LIMBS 16 /* 1024-bit two's complement */

typedef struct { uint64_t w[ORA_BN_LIMBS]; } ora_bn;

static void ora_bn_from_float_scaled(ora_bn* r, float v, int emin)
{
	memset(r, 0, sizeof(*r));
	if (v == 0.0f) return;
	int e;
	float fm = frexpf(v, &e);                 /* v = fm * 2^e, 0.5 <= |fm| < 1 */
	int64_t m = (int64_t) ldexpf(fm, 24);     /* exact 24-bit integer mantissa */
	int shift = (e - 24) - emin;              /* v = m * 2^(e-24) = m*2^shift*2^emin */
	int neg = m < 0;
	uint64_t um = (uint64_t) (neg ? -m : m);
	int limb = shift / 64, bit = shift % 64;
	r->w[limb] |= um << bit;
	if (bit != 0 && limb + 1 < ORA_BN_LIMBS) r->w[limb + 1] |= um >> (64 - bit);
	if (neg)
	{
		uint64_t carry = 1;
		for (int i = 0; i < ORA_BN_LIMBS; i++)
		{
			uint64_t x = ~r->w[i] + carry;
			carry = (carry && x == 0) ? 1 : 0;
			r->w[i] = x;
		}
	}
}

static void ora_bn_add(ora_bn* r, const ora_bn* a, const ora_bn* b)
{
	unsigned __int128 c = 0;
	for (int i = 0; i < ORA_BN_LIMBS; i++)
	{
		c += (unsigned __int128) a->w[i] + b->w[i];
		r->w[i] = (uint64_t) c;
		c >>= 64;
	}
}

static void ora_bn_neg(ora_bn* r, const ora_bn* a)
{
	uint64_t carry = 1;
	for (int i = 0; i < ORA_BN_LIMBS; i++)
	{
		uint64_t x = ~a->w[i] + carry;
		carry = (carry && x == 0) ? 1 : 0;
		r->w[i] = x;
	}
}

static void ora_bn_sub(ora_bn* r, const ora_bn* a, const ora_bn* b)
{
	ora_bn nb;
	ora_bn_neg(&nb, b);
	ora_bn_add(r, a, &nb);
}

static int ora_bn_sign(const ora_bn* a)
{
	if (a->w[ORA_BN_LIMBS - 1] >> 63) return -1;
	for (int i = 0; i < ORA_BN_LIMBS; i++)
		if (a->w[i]) return 1;
	return 0;
}

static void ora_bn_mul(ora_bn* r, const ora_bn* a, const ora_bn* b)
{
	ora_bn x = *a, y = *b;
	int sa = ora_bn_sign(&x), sb = ora_bn_sign(&y);
	if (sa < 0) ora_bn_neg(&x, &x);
	if (sb < 0) ora_bn_neg(&y, &y);
	uint64_t out[ORA_BN_LIMBS];
	memset(out, 0, sizeof(out));
	for (int i = 0; i < ORA_BN_LIMBS; i++)
	{
		if (x.w[i] == 0) continue;
		unsigned __int128 carry = 0;
		for (int j = 0; i + j < ORA_BN_LIMBS; j++)
		{
			unsigned __int128 t = (unsigned __int128) x.w[i] * y.w[j] + out[i + j] + carry;
			out[i + j] = (uint64_t) t;
			carry = t >> 64;
		}
	}
	memcpy(r->w, out, sizeof(out));
	if ((sa < 0) != (sb < 0)) ora_bn_neg(r, r);
}

/* sign of det[a-d; b-d; c-d] for float points */
static int ora_orient3d_sign(const float* a, const float* b, const float* c, const float* d)
{
	const double adx = (double) a[0] - d[0], ady = (double) a[1] - d[1], adz = (double) a[2] - d[2];
	const double bdx = (double) b[0] - d[0], bdy = (double) b[1] - d[1], bdz = (double) b[2] - d[2];
	const double cdx = (double) c[0] - d[0], cdy = (double) c[1] - d[1], cdz = (double) c[2] - d[2];
	const double det = adx * (bdy * cdz - bdz * cdy) - ady * (bdx * cdz - bdz * cdx) + adz * (bdx * cdy - bdy * cdx);
	const double perm = fabs(adx) * (fabs(bdy * cdz) + fabs(bdz * cdy)) + fabs(ady) * (fabs(bdx * cdz) + fabs(bdz * cdx))
	                  + fabs(adz) * (fabs(bdx * cdy) + fabs(bdy * cdx));
	/* the float differences carry at most one rounding each; 1e-14 covers
	   every rounding of the double evaluation with a wide margin */
	const double bound = 1e-14 * perm;
	if (det > bound) return 1;
	if (det < -bound) return -1;
	if (perm == 0.0) return 0;

	/* exact path */
	const float* p[4] = { a, b, c, d };
	int emin = 1 << 30;
	for (int i = 0; i < 4; i++)
		for (int k = 0; k < 3; k++)
			if (p[i][k] != 0.0f)
			{
				int e;
				frexpf(p[i][k], &e);
				if (e - 24 < emin) emin = e - 24;
			}
	ora_bn v[4][3];
	for (int i = 0; i < 4; i++)
		for (int k = 0; k < 3; k++)
			ora_bn_from_float_scaled(&v[i][k], p[i][k], emin);
	ora_bn m[3][3];
	for (int i = 0; i < 3; i++)
		for (int k = 0; k < 3; k++)
			ora_bn_sub(&m[i][k], &v[i][k], &v[3][k]);
	ora_bn t1, t2, minor, acc, term;
	memset(&acc, 0, sizeof(acc));
	/* + m00 (m11 m22 - m12 m21) */
	ora_bn_mul(&t1, &m[1][1], &m[2][2]); ora_bn_mul(&t2, &m[1][2], &m[2][1]); ora_bn_sub(&minor, &t1, &t2);
	ora_bn_mul(&term, &m[0][0], &minor); ora_bn_add(&acc, &acc, &term);
	/* - m01 (m10 m22 - m12 m20) */
	ora_bn_mul(&t1, &m[1][0], &m[2][2]); ora_bn_mul(&t2, &m[1][2], &m[2][0]); ora_bn_sub(&minor, &t1, &t2);
	ora_bn_mul(&term, &m[0][1], &minor); ora_bn_sub(&acc, &acc, &term);
	/* + m02 (m10 m21 - m11 m20) */
	ora_bn_mul(&t1, &m[1][0], &m[2][1]); ora_bn_mul(&t2, &m[1][1], &m[2][0]); ora_bn_sub(&minor, &t1, &t2);
	ora_bn_mul(&term, &m[0][2], &minor); ora_bn_add(&acc, &acc, &term);
	return ora_bn_sign(&acc);
}

#endif
