// Synthetic specimen: a BCC tip with its [110] axis along z (cylindrical shank
// plus hemispherical cap), optional solute clusters, a few lattice layers of
// vacuum sites around it, graded random vacuum nodes whose density falls as
// distance^-decay from the tip surface, and a box of boundary nodes
// (Top / Sides Dirichlet, Bottom Neumann as in examples/example.aly).  The
// reference reads such node sets from a file produced by its meshgen tool
// (meshgen/meshgen.cpp); this generator is seeded and deterministic.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "mesh.hpp"

namespace tsg {
namespace {

struct Rng {  // splitmix64
	uint64_t s;
	explicit Rng(uint64_t seed) : s(seed) {}
	uint64_t next()
	{
		uint64_t z = (s += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// signed distance to the tip surface (negative inside)
double tip_distance(double x, double y, double z, double R, double H)
{
	if (z <= H) return std::sqrt(x * x + y * y) - R;
	return std::sqrt(x * x + y * y + (z - H) * (z - H)) - R;
}

}  // namespace

void generate_specimen(const SpecimenParams& P, Specimen* out)
{
	out->xyz.clear();
	out->ids.clear();
	const double a = P.a, R = P.radius, H = P.shank;
	const double bx = P.box[0], by = P.box[1], bz = P.box[2];
	const double zb = -0.25 * a;            // bottom face just below the lowest atom plane
	const double shell = P.shell * a;
	Rng rng(P.seed);

	// solute cluster centres, uniform in the tip volume
	std::vector<double> centres;
	if (P.cluster_density > 0) {
		const double vol = M_PI * R * R * H + 2.0 / 3.0 * M_PI * R * R * R;
		const double bbox = (2 * R) * (2 * R) * (H + R);
		const long long want = (long long) std::llround(P.cluster_density * vol);
		long long got = 0;
		while (got < want) {
			const double x = (rng.uniform() * 2 - 1) * R, y = (rng.uniform() * 2 - 1) * R, z = rng.uniform() * (H + R);
			if (tip_distance(x, y, z, R, H) > 0) continue;
			centres.push_back(x); centres.push_back(y); centres.push_back(z);
			got++;
		}
		(void) bbox;
	}

	auto push = [&](double x, double y, double z, int id) {
		out->xyz.push_back((float) x);
		out->xyz.push_back((float) y);
		out->xyz.push_back((float) z);
		out->ids.push_back((int16_t) id);
	};

	// BCC sites: cubic (i,j,k)a and (i+1/2,j+1/2,k+1/2)a, rotated so that
	// x = [001], y = [1-10]/sqrt2, z = [110]/sqrt2
	const double s2 = std::sqrt(0.5);
	const double zmaxAtoms = H + R + shell;
	const double rmax = R + shell;
	const double lim = (std::max(rmax, zmaxAtoms) * 2.0) / a + 4;
	const int L = (int) std::ceil(lim);
	const double cr2 = P.cluster_radius * P.cluster_radius;
	for (int i = -L; i <= L; i++)
		for (int j = -L; j <= L; j++) {
			// z = (i + j) a / sqrt2 must fall in [0, zmaxAtoms]
			const double zc = (i + j) * a * s2;
			if (zc < -a || zc > zmaxAtoms + a) continue;
			const double yc = (i - j) * a * s2;
			if (std::fabs(yc) > rmax + a) continue;
			for (int k = -L; k <= L; k++)
				for (int h = 0; h < 2; h++) {
					const double cx = (i + 0.5 * h) * a, cy = (j + 0.5 * h) * a, cz = (k + 0.5 * h) * a;
					const double x = cz, y = (cx - cy) * s2, z = (cx + cy) * s2;
					if (z < 0 || z > zmaxAtoms) continue;
					if (std::fabs(x) > bx / 2 || std::fabs(y) > by / 2 || z >= bz) continue;
					const double d = tip_distance(x, y, z, R, H);
					if (d > shell) continue;
					if (d <= 0) {
						int id = P.ids.matrix;
						for (size_t c = 0; c < centres.size(); c += 3) {
							const double dx = x - centres[c], dy = y - centres[c + 1], dz = z - centres[c + 2];
							if (dx * dx + dy * dy + dz * dz <= cr2) { id = P.ids.solute; break; }
						}
						push(x, y, z, id);
						out->n_atoms++;
						if (id == P.ids.solute) out->n_solute++;
					} else {
						push(x, y, z, P.ids.vacuum);
						out->n_vacuum_lattice++;
					}
				}
		}

	// graded random vacuum: node density rho(d) = rho_near * min(1, (dc/d)^decay)
	// at distance d from the tip surface, with rho_near one node per (2a)^3
	// and dc chosen (bisection on a pilot sample) so that about vacuum_target
	// nodes are expected; proposals are uniform in the box and accepted with
	// probability min(1, (dc/d)^decay)
	if (P.vacuum_target > 0) {
		const double dmin = shell + 0.5 * a;
		auto dist = [&](const double* p) { return tip_distance(p[0], p[1], p[2], R, H); };
		// proposals keep a margin from the box faces so that no vacuum node
		// lands on (or rounds onto) a face
		const double mx = 0.01 * P.face_spacing;
		auto propose = [&](Rng& r, double* p) {
			p[0] = (r.uniform() - 0.5) * (bx - 2 * mx);
			p[1] = (r.uniform() - 0.5) * (by - 2 * mx);
			p[2] = zb + mx + r.uniform() * (bz - zb - 2 * mx);
		};
		const double V = bx * by * (bz - zb);
		const double rho_near = 1.0 / (8.0 * a * a * a);
		const double nprop = rho_near * V;
		Rng pilot(P.seed ^ 0x5bd1e995ull);
		const int np = 400000;
		std::vector<double> dd;
		dd.reserve(np);
		for (int q = 0; q < np; q++) {
			double p[3];
			propose(pilot, p);
			const double d = dist(p);
			if (d >= dmin) dd.push_back(d);
		}
		auto expected = [&](double dc) {
			double acc = 0;
			for (size_t q = 0; q < dd.size(); q++) acc += std::min(1.0, std::pow(dc / dd[q], P.decay));
			return acc / np * nprop;
		};
		double lo = 1e-3 * a, hi = std::max(bx, std::max(by, bz));
		for (int it = 0; it < 60; it++) {
			const double mid = std::sqrt(lo * hi);
			if (expected(mid) < (double) P.vacuum_target) lo = mid; else hi = mid;
		}
		const double dc = std::sqrt(lo * hi);
		const long long total = (long long) nprop;
		for (long long q = 0; q < total; q++) {
			double p[3];
			propose(rng, p);
			const double d = dist(p);
			if (d < dmin) { rng.next(); continue; }
			const double w = std::min(1.0, std::pow(dc / d, P.decay));
			if (rng.uniform() >= w) continue;
			push(p[0], p[1], p[2], P.ids.vacuum);
			out->n_vacuum_random++;
		}
	}

	// box surface nodes: corners, edges and faces on a grid of face_spacing
	const double xs[2] = {-bx / 2, bx / 2}, ys[2] = {-by / 2, by / 2}, zs[2] = {zb, bz};
	const int nx = std::max(1, (int) std::ceil(bx / P.face_spacing));
	const int ny = std::max(1, (int) std::ceil(by / P.face_spacing));
	const int nz = std::max(1, (int) std::ceil((bz - zb) / P.face_spacing));
	for (int i = 0; i <= nx; i++)
		for (int j = 0; j <= ny; j++)
			for (int k = 0; k <= nz; k++) {
				const bool ex = (i == 0 || i == nx), ey = (j == 0 || j == ny), ez = (k == 0 || k == nz);
				if (!(ex || ey || ez)) continue;
				const double x = xs[0] + (xs[1] - xs[0]) * i / nx;
				const double y = ys[0] + (ys[1] - ys[0]) * j / ny;
				const double z = zs[0] + (zs[1] - zs[0]) * k / nz;
				int id;
				if (k == nz) id = P.ids.top;
				else if (ex || ey) id = P.ids.sides;
				else id = P.ids.bottom;
				push(x, y, z, id);
				out->n_box++;
			}

	// node numbering along a 3-D Hilbert curve: nodes close in space get
	// close numbers, so a wavefront of consecutive rows gathers the
	// potentials of a compact range of nodes (the device sweeps' locality)
	const size_t n = out->ids.size();
	float lo[3], hi[3];
	for (int k = 0; k < 3; k++) lo[k] = hi[k] = out->xyz[k];
	for (size_t i = 0; i < n; i++)
		for (int k = 0; k < 3; k++) {
			lo[k] = std::min(lo[k], out->xyz[3 * i + k]);
			hi[k] = std::max(hi[k], out->xyz[3 * i + k]);
		}
	double ext = 0.0;
	for (int k = 0; k < 3; k++) ext = std::max(ext, (double) hi[k] - lo[k]);
	const int bits = 21;
	const double scale = ext > 0.0 ? ((1u << bits) - 1) / ext : 0.0;
	std::vector<uint64_t> key(n);
	for (size_t i = 0; i < n; i++) {
		uint32_t q[3];
		for (int k = 0; k < 3; k++)
			q[k] = (uint32_t) std::min<double>((1u << bits) - 1, std::max(0.0, (out->xyz[3 * i + k] - lo[k]) * scale));
		key[i] = hilbert3(q[0], q[1], q[2], bits);
	}
	std::vector<size_t> order(n);
	for (size_t i = 0; i < n; i++) order[i] = i;
	std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
	std::vector<float> xyz(3 * n);
	std::vector<int16_t> ids(n);
	for (size_t i = 0; i < n; i++) {
		for (int k = 0; k < 3; k++) xyz[3 * i + k] = out->xyz[3 * order[i] + k];
		ids[i] = out->ids[order[i]];
	}
	out->xyz.swap(xyz);
	out->ids.swap(ids);
}

}  // namespace tsg
