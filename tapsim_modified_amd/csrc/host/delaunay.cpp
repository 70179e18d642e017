// Incremental Delaunay tetrahedralisation (Bowyer-Watson) with exact
// predicates.
//
// Points are inserted in 3-D Hilbert order; each insertion locates the point
// by a visibility walk from the last new tet, grows the cavity of tets whose
// circumsphere strictly contains it, enlarges the cavity until it is strictly
// star-shaped from the point (which also resolves cospherical ties), and
// re-fans the cavity boundary.  A large enclosing tetrahedron is removed at
// the end; with the box corners and faces among the input, the remaining
// hull is the box.  The reference obtains the same tetrahedralisation from
// TetGen with switches "znNFCQ" (tapsim/geometry_3d.cpp:759-810).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "predicates.hpp"

namespace tsg {

uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits)
{
	// Skilling's transform of the coordinate bits into Hilbert order
	uint32_t X[3] = {x, y, z};
	const uint32_t M = 1u << (bits - 1);
	for (uint32_t Q = M; Q > 1; Q >>= 1) {
		const uint32_t P = Q - 1;
		for (int i = 0; i < 3; i++) {
			if (X[i] & Q) X[0] ^= P;
			else {
				const uint32_t t = (X[0] ^ X[i]) & P;
				X[0] ^= t;
				X[i] ^= t;
			}
		}
	}
	for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
	uint32_t t = 0;
	for (uint32_t Q = M; Q > 1; Q >>= 1)
		if (X[2] & Q) t ^= Q - 1;
	for (int i = 0; i < 3; i++) X[i] ^= t;
	uint64_t key = 0;
	for (int b = bits - 1; b >= 0; b--)
		for (int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> b) & 1u);
	return key;
}

namespace {

struct Tet {
	int v[4];
	int n[4];  // (tet << 2) | face of the neighbour, -1 outside
};

class Builder {
public:
	Builder(const std::vector<float>& xyz) : n_(xyz.size() / 3)
	{
		pts_.resize((n_ + 4) * 3);
		std::copy(xyz.begin(), xyz.end(), pts_.begin());
	}

	bool run(Mesh* out, std::string* err);

private:
	const float* P(int i) const { return &pts_[3 * (size_t) i]; }

	int new_tet()
	{
		if (!free_.empty()) {
			const int t = free_.back();
			free_.pop_back();
			return t;
		}
		tets_.push_back(Tet());
		stamp_.push_back(0);
		return (int) tets_.size() - 1;
	}

	// orientation of tet t with vertex slot i replaced by p
	int orient_replaced(int t, int i, const float* p) const
	{
		const float* q[4];
		for (int k = 0; k < 4; k++) q[k] = k == i ? p : P(tets_[t].v[k]);
		return orient_sign(q[0], q[1], q[2], q[3]);
	}

	bool in_conflict(int t, const float* p) const
	{
		const Tet& T = tets_[t];
		return insphere_sign(P(T.v[0]), P(T.v[1]), P(T.v[2]), P(T.v[3]), p) > 0;
	}

	int locate(const float* p, int start, bool* dup, int pi);
	bool insert(int pi, std::string* err);

	size_t n_;
	std::vector<float> pts_;
	std::vector<Tet> tets_;
	std::vector<uint32_t> stamp_;
	uint32_t cur_stamp_ = 0;
	std::vector<int> free_;
	int last_ = 0;
	// scratch
	std::vector<int> cavity_, stack_;
	std::vector<std::pair<int, int>> bfaces_;  // (tet, face)
	std::vector<uint64_t> hkey_;
	std::vector<int> hval_;
	std::vector<int> hused_;
	std::vector<int> created_;
	std::vector<Tet> old_;
};

int Builder::locate(const float* p, int t, bool* dup, int pi)
{
	*dup = false;
	const int maxSteps = (int) tets_.size() * 2 + 16;
	int rot = 0;
	for (int step = 0; step < maxSteps; step++) {
		const Tet& T = tets_[t];
		int next = -1;
		rot = (rot + 1) & 3;
		for (int kk = 0; kk < 4; kk++) {
			const int k = (kk + rot) & 3;
			if (orient_replaced(t, k, p) < 0) {
				next = T.n[k];
				break;
			}
		}
		if (next < 0) {
			for (int k = 0; k < 4; k++) {
				const float* q = P(T.v[k]);
				if (q[0] == p[0] && q[1] == p[1] && q[2] == p[2] && T.v[k] != pi) *dup = true;
			}
			return t;
		}
		t = next >> 2;
	}
	return -1;
}

bool Builder::insert(int pi, std::string* err)
{
	const float* p = P(pi);
	bool dup = false;
	const int t0 = locate(p, last_, &dup, pi);
	if (t0 < 0) { *err = "point location failed"; return false; }
	if (dup) { *err = "duplicate point " + std::to_string(pi); return false; }

	// cavity of tets whose circumsphere strictly contains p
	cur_stamp_++;
	cavity_.clear();
	stack_.clear();
	stamp_[t0] = cur_stamp_;
	cavity_.push_back(t0);
	stack_.push_back(t0);
	while (!stack_.empty()) {
		const int t = stack_.back();
		stack_.pop_back();
		for (int k = 0; k < 4; k++) {
			const int nb = tets_[t].n[k];
			if (nb < 0) continue;
			const int u = nb >> 2;
			if (stamp_[u] == cur_stamp_) continue;
			if (in_conflict(u, p)) {
				stamp_[u] = cur_stamp_;
				cavity_.push_back(u);
				stack_.push_back(u);
			}
		}
	}
	// enlarge until every boundary face sees p strictly on its inner side
	for (;;) {
		bool grown = false;
		bfaces_.clear();
		for (size_t c = 0; c < cavity_.size(); c++) {
			const int t = cavity_[c];
			for (int k = 0; k < 4; k++) {
				const int nb = tets_[t].n[k];
				if (nb >= 0 && stamp_[nb >> 2] == cur_stamp_) continue;
				if (orient_replaced(t, k, p) <= 0) {
					if (nb < 0) { *err = "point outside the enclosing tetrahedron"; return false; }
					stamp_[nb >> 2] = cur_stamp_;
					cavity_.push_back(nb >> 2);
					grown = true;
				} else
					bfaces_.push_back(std::make_pair(t, k));
			}
		}
		if (!grown) break;
	}

	// hash of the new internal faces keyed by their edge opposite p
	size_t hs = 64;
	while (hs < bfaces_.size() * 4) hs <<= 1;
	if (hkey_.size() < hs) { hkey_.assign(hs, ~0ull); hval_.assign(hs, -1); }
	hused_.clear();
	const size_t mask = hs - 1;

	std::vector<int>& created = created_;
	created.clear();
	std::vector<Tet>& old = old_;
	old.resize(bfaces_.size());
	for (size_t f = 0; f < bfaces_.size(); f++) old[f] = tets_[bfaces_[f].first];
	// the cavity tets become free slots, reused first
	for (size_t c = 0; c < cavity_.size(); c++) free_.push_back(cavity_[c]);

	for (size_t f = 0; f < bfaces_.size(); f++) {
		const Tet& src = old[f];
		const int k = bfaces_[f].second;
		const int nt = new_tet();
		stamp_[nt] = 0;
		Tet& T = tets_[nt];
		for (int q = 0; q < 4; q++) T.v[q] = src.v[q];
		T.v[k] = pi;
		T.n[k] = src.n[k];
		if (src.n[k] >= 0) tets_[src.n[k] >> 2].n[src.n[k] & 3] = (nt << 2) | k;
		for (int j = 0; j < 4; j++) {
			if (j == k) continue;
			int a = -1, b = -1;
			for (int q = 0; q < 4; q++) {
				if (q == j || q == k) continue;
				if (a < 0) a = T.v[q]; else b = T.v[q];
			}
			const uint64_t key = a < b ? ((uint64_t) (uint32_t) a << 32) | (uint32_t) b
			                           : ((uint64_t) (uint32_t) b << 32) | (uint32_t) a;
			size_t h = (key * 0x9E3779B97F4A7C15ull) >> 20 & mask;
			for (;;) {
				if (hkey_[h] == ~0ull) {
					hkey_[h] = key;
					hval_[h] = (nt << 2) | j;
					hused_.push_back((int) h);
					break;
				}
				if (hkey_[h] == key) {
					const int other = hval_[h];
					T.n[j] = other;
					tets_[other >> 2].n[other & 3] = (nt << 2) | j;
					hkey_[h] = ~1ull;  // consumed; keeps the probe chain intact
					break;
				}
				h = (h + 1) & mask;
			}
		}
		created.push_back(nt);
	}
	for (size_t u = 0; u < hused_.size(); u++) {
		if (hkey_[hused_[u]] != ~1ull) { *err = "unmatched cavity face"; return false; }
		hkey_[hused_[u]] = ~0ull;
		hval_[hused_[u]] = -1;
	}
	// mark freed slots that were not reused as dead
	for (size_t i = 0; i < free_.size(); i++) tets_[free_[i]].v[0] = -1;
	last_ = created.back();
	return true;
}

bool Builder::run(Mesh* out, std::string* err)
{
	if (n_ < 4) { *err = "need at least 4 points"; return false; }
	float lo[3], hi[3];
	for (int k = 0; k < 3; k++) { lo[k] = pts_[k]; hi[k] = pts_[k]; }
	for (size_t i = 0; i < n_; i++)
		for (int k = 0; k < 3; k++) {
			lo[k] = std::min(lo[k], pts_[3 * i + k]);
			hi[k] = std::max(hi[k], pts_[3 * i + k]);
		}
	double L = 0, c[3];
	for (int k = 0; k < 3; k++) { L = std::max(L, (double) hi[k] - lo[k]); c[k] = 0.5 * ((double) hi[k] + lo[k]); }
	if (L <= 0) { *err = "degenerate point set"; return false; }
	// enclosing tetrahedron, ~1000 box sizes away
	const double S = 1000.0 * L;
	const double sv[4][3] = {{S, S, S}, {-S, -S, S}, {-S, S, -S}, {S, -S, -S}};
	for (int q = 0; q < 4; q++)
		for (int k = 0; k < 3; k++) pts_[3 * (n_ + q) + k] = (float) (c[k] + sv[q][k]);
	Tet T0;
	for (int q = 0; q < 4; q++) { T0.v[q] = (int) n_ + q; T0.n[q] = -1; }
	if (orient_sign(P(T0.v[0]), P(T0.v[1]), P(T0.v[2]), P(T0.v[3])) < 0) std::swap(T0.v[0], T0.v[1]);
	tets_.reserve(n_ * 7 + 16);
	stamp_.reserve(n_ * 7 + 16);
	tets_.push_back(T0);
	stamp_.push_back(0);
	last_ = 0;

	// Hilbert order
	std::vector<uint64_t> key(n_);
	const int bits = 21;
	const double scale = ((1u << bits) - 1) / L;
	for (size_t i = 0; i < n_; i++) {
		uint32_t q[3];
		for (int k = 0; k < 3; k++) q[k] = (uint32_t) std::min<double>((1u << bits) - 1, std::max(0.0, (pts_[3 * i + k] - lo[k]) * scale));
		key[i] = hilbert3(q[0], q[1], q[2], bits);
	}
	std::vector<int> order(n_);
	std::iota(order.begin(), order.end(), 0);
	std::sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });

	// a build of millions of points takes minutes: with TSG_VERBOSE set,
	// every 2 M insertions are reported on stderr
	const bool verbose = std::getenv("TSG_VERBOSE") != nullptr;
	for (size_t i = 0; i < n_; i++) {
		if (!insert(order[i], err)) return false;
		if (verbose && (i + 1) % 2000000 == 0) {
			std::fprintf(stderr, "[tsg mesh] %zu of %zu points inserted, %zu tets\n", i + 1, n_, tets_.size());
			std::fflush(stderr);
		}
	}

	// drop tets touching the enclosing vertices, compact, re-encode neighbours
	std::vector<int> remap(tets_.size(), -1);
	int m = 0;
	for (size_t t = 0; t < tets_.size(); t++) {
		const Tet& T = tets_[t];
		if (T.v[0] < 0) continue;
		bool keep = true;
		for (int q = 0; q < 4; q++)
			if (T.v[q] >= (int) n_) keep = false;
		if (keep) remap[t] = m++;
	}
	out->n = (int) n_;
	out->xyz.assign(pts_.begin(), pts_.begin() + 3 * n_);
	out->tet_v.resize(4 * (size_t) m);
	out->tet_n.resize(4 * (size_t) m);
	for (size_t t = 0; t < tets_.size(); t++) {
		const int r = remap[t];
		if (r < 0) continue;
		for (int q = 0; q < 4; q++) {
			out->tet_v[4 * (size_t) r + q] = tets_[t].v[q];
			const int nb = tets_[t].n[q];
			out->tet_n[4 * (size_t) r + q] = nb < 0 ? -1 : remap[nb >> 2];
		}
	}
	return true;
}

}  // namespace

bool delaunay(const std::vector<float>& xyz, Mesh* out, std::string* err, int)
{
	Builder b(xyz);
	return b.run(out, err);
}

}  // namespace tsg
