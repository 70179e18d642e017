/*
 * CPU oracle of the warm-start network (see mpnn_oracle.h).  Test
 * infrastructure only.  Straightforward loops in double; the gradient is
 * the textbook reverse-mode derivative of the forward pass written out
 * layer by layer.
 */
#include "mpnn_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

float ora_bf16(float v)
{
	uint32_t u;
	memcpy(&u, &v, 4);
	if ((u & 0x7f800000u) == 0x7f800000u) {  /* inf / nan: truncate, keep nan quiet */
		if (u & 0x007fffffu) u |= 0x00400000u;
		u &= 0xffff0000u;
	} else {
		u += 0x7fffu + ((u >> 16) & 1u);
		u &= 0xffff0000u;
	}
	memcpy(&v, &u, 4);
	return v;
}

/* value as the device holds it: an fp32 quantity rounded to bfloat16 when emulating */
static double rb(int emulate, double v) { return emulate ? (double) ora_bf16((float) v) : v; }

enum { IN = ORA_MPNN_IN, H = ORA_MPNN_HID, L = ORA_MPNN_LAYERS, H2 = 2 * ORA_MPNN_HID };

int ora_mpnn(const ora_mpnn_graph* g, const float* params, int emulate, double* loss_out, double* y_out, double* grad)
{
	const int n = g->n;
	if (n <= 0) return -1;
	/* weights as the device multiplies them */
	double* P = (double*) malloc(sizeof(double) * ORA_MPNN_NPARAM);
	for (int i = 0; i < ORA_MPNN_NPARAM; i++) P[i] = params[i];
	if (emulate) {
		for (int i = 0; i < IN * H; i++) P[ORA_MPNN_W_IN + i] = ora_bf16(params[ORA_MPNN_W_IN + i]);
		for (int l = 0; l < L; l++)
			for (int i = 0; i < H2 * H; i++) {
				const int o = ORA_MPNN_LAYER0 + l * ORA_MPNN_LAYER_STRIDE + i;
				P[o] = ora_bf16(params[o]);
			}
		for (int i = 0; i < H; i++) P[ORA_MPNN_W_OUT + i] = ora_bf16(params[ORA_MPNN_W_OUT + i]);
	}
	double* X = (double*) malloc(sizeof(double) * (size_t) n * IN);
	for (size_t i = 0; i < (size_t) n * IN; i++) X[i] = rb(emulate, g->x[i]);

	/* activations: h[l] (l = 0..L), pre-activations z[l], the GEMM operands cat[l] (l = 1..L) */
	double* h = (double*) calloc((size_t) (L + 1) * n * H, sizeof(double));
	double* z = (double*) calloc((size_t) (L + 1) * n * H, sizeof(double));
	double* cat = (double*) calloc((size_t) (L + 1) * n * H2, sizeof(double));
	double* y = (double*) calloc((size_t) n, sizeof(double));
#define Hh(l, i, o) h[((size_t) (l) * n + (i)) * H + (o)]
#define Zz(l, i, o) z[((size_t) (l) * n + (i)) * H + (o)]
#define Cc(l, i, k) cat[((size_t) (l) * n + (i)) * H2 + (k)]

	for (int i = 0; i < n; i++)
		for (int o = 0; o < H; o++) {
			double s = 0.0;
			for (int k = 0; k < IN; k++) s += X[(size_t) i * IN + k] * P[ORA_MPNN_W_IN + k * H + o];
			s += P[ORA_MPNN_B_IN + o];
			Zz(0, i, o) = s;
			Hh(0, i, o) = s > 0.0 ? s : 0.0;
		}
	for (int l = 1; l <= L; l++) {
		const double* W = P + ORA_MPNN_LAYER0 + (l - 1) * ORA_MPNN_LAYER_STRIDE;
		const double* b = W + H2 * H;
		for (int i = 0; i < n; i++)
			for (int o = 0; o < H; o++) {
				double m = 0.0;
				for (int e = g->ptr[i]; e < g->ptr[i + 1]; e++) m += (double) g->a[e] * Hh(l - 1, g->col[e], o);
				Cc(l, i, o) = rb(emulate, Hh(l - 1, i, o));
				Cc(l, i, H + o) = rb(emulate, m);
			}
		for (int i = 0; i < n; i++)
			for (int o = 0; o < H; o++) {
				double s = 0.0;
				for (int k = 0; k < H2; k++) s += Cc(l, i, k) * W[k * H + o];
				s += b[o];
				Zz(l, i, o) = s;
				Hh(l, i, o) = (s > 0.0 ? s : 0.0) + Hh(l - 1, i, o);
			}
	}
	double msum = 0.0;
	for (int i = 0; i < n; i++) msum += g->mask[i];
	const double M = msum > 1.0 ? msum : 1.0;
	double loss = 0.0;
	for (int i = 0; i < n; i++) {
		double s = 0.0;
		for (int o = 0; o < H; o++) s += rb(emulate, Hh(L, i, o)) * P[ORA_MPNN_W_OUT + o];
		s += P[ORA_MPNN_B_OUT];
		y[i] = s;
		const double d = s - g->t[i];
		loss += g->mask[i] * d * d;
	}
	loss /= M;
	if (loss_out) *loss_out = loss;
	if (y_out) memcpy(y_out, y, sizeof(double) * n);

	if (grad) {
		memset(grad, 0, sizeof(double) * ORA_MPNN_NPARAM);
		double* dh = (double*) calloc((size_t) n * H, sizeof(double));
		double* dz = (double*) calloc((size_t) n * H, sizeof(double));
		double* dcat = (double*) calloc((size_t) n * H2, sizeof(double));
		for (int i = 0; i < n; i++) {
			const double dy = 2.0 * g->mask[i] * (y[i] - g->t[i]) / M;
			grad[ORA_MPNN_B_OUT] += dy;
			for (int o = 0; o < H; o++) {
				grad[ORA_MPNN_W_OUT + o] += dy * rb(emulate, Hh(L, i, o));
				dh[(size_t) i * H + o] = dy * P[ORA_MPNN_W_OUT + o];
			}
		}
		for (int l = L; l >= 1; l--) {
			const double* W = P + ORA_MPNN_LAYER0 + (l - 1) * ORA_MPNN_LAYER_STRIDE;
			double* gW = grad + ORA_MPNN_LAYER0 + (l - 1) * ORA_MPNN_LAYER_STRIDE;
			double* gb = gW + H2 * H;
			for (int i = 0; i < n; i++)
				for (int o = 0; o < H; o++) {
					const double d = Zz(l, i, o) > 0.0 ? dh[(size_t) i * H + o] : 0.0;
					gb[o] += d;
					dz[(size_t) i * H + o] = rb(emulate, d);
				}
			for (int k = 0; k < H2; k++)
				for (int o = 0; o < H; o++) {
					double s = 0.0;
					for (int i = 0; i < n; i++) s += Cc(l, i, k) * dz[(size_t) i * H + o];
					gW[k * H + o] += s;
				}
			for (int i = 0; i < n; i++)
				for (int k = 0; k < H2; k++) {
					double s = 0.0;
					for (int o = 0; o < H; o++) s += dz[(size_t) i * H + o] * W[k * H + o];
					dcat[(size_t) i * H2 + k] = s;
				}
			/* h_{l-1} feeds the residual path, the left half of cat and, through
			 * m = A h_{l-1}, the right half: dh_{l-1} += A^T dm */
			for (int i = 0; i < n; i++)
				for (int o = 0; o < H; o++) dh[(size_t) i * H + o] += dcat[(size_t) i * H2 + o];
			for (int i = 0; i < n; i++)
				for (int e = g->ptr[i]; e < g->ptr[i + 1]; e++) {
					const int j = g->col[e];
					for (int o = 0; o < H; o++) dh[(size_t) j * H + o] += (double) g->a[e] * dcat[(size_t) i * H2 + H + o];
				}
		}
		for (int i = 0; i < n; i++)
			for (int o = 0; o < H; o++) {
				const double d = Zz(0, i, o) > 0.0 ? dh[(size_t) i * H + o] : 0.0;
				grad[ORA_MPNN_B_IN + o] += d;
				dz[(size_t) i * H + o] = rb(emulate, d);
			}
		for (int k = 0; k < IN; k++)
			for (int o = 0; o < H; o++) {
				double s = 0.0;
				for (int i = 0; i < n; i++) s += X[(size_t) i * IN + k] * dz[(size_t) i * H + o];
				grad[ORA_MPNN_W_IN + k * H + o] += s;
			}
		free(dh);
		free(dz);
		free(dcat);
	}
#undef Hh
#undef Zz
#undef Cc
	free(P);
	free(X);
	free(h);
	free(z);
	free(cat);
	free(y);
	return 0;
}
