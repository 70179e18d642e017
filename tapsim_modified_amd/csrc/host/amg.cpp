// Aggregation multigrid for the correction equation of the accelerated
// relaxation: the hierarchy (host) and the preconditioned conjugate gradient
// driver (device kernels of kernels/amg.hip).  No reference counterpart: it
// replaces the Chebyshev iteration of the FP32 correction solve when
// selected; the FP64 residual refinement around it, and so the stopping rule,
// are the same.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "grid.hpp"

#define HIPCHK(x)                                      \
	do {                                               \
		if (int rc__ = hip_check((x), #x)) return rc__; \
	} while (0)

namespace tsg {

namespace {

// every coarse correction of the V-cycle is scaled by this factor: piecewise
// constant prolongation underestimates smooth errors, and over-correcting
// (Vanek, Brezina and Mandel's remedy for unsmoothed aggregation) takes PCG
// iterations away on tip meshes at no cost per iteration; the preconditioner
// stays symmetric and, below 2, positive definite
constexpr float kAmgCoarseScale = 1.6f;

// the damped-Jacobi weight of every level is this factor times 4 / (3 lmax),
// lmax the power iteration's estimate of the largest eigenvalue of D^-1 A,
// and at most 1.  With the over-correction above it takes a fifth of the
// PCG iterations away on tip meshes (factor 1 and scale 1.4 before).  The
// bound keeps the smoother convergent whatever the estimate: D^-1 A = I - G
// with G >= 0 and row sums of G at most 1 on every level (the correction rows
// and their Galerkin products are weakly diagonally dominant M-matrices), so
// lmax <= 2 and any weight up to 1 contracts.
constexpr double kAmgSmoothFactor = 1.4;

struct Csr {
	int n = 0;
	std::vector<int> ptr, col;
	std::vector<double> val;
};

// greedy aggregation over strong couplings |a_ij| >= theta sqrt(a_ii a_jj):
// seeds whose strong neighbours are all free take them; the rest join the
// aggregate of a strong neighbour, or stay alone
int aggregate(const Csr& A, double theta, std::vector<int>* agg)
{
	const int n = A.n;
	std::vector<double> d(n, 0.0);
	for (int i = 0; i < n; i++)
		for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++)
			if (A.col[k] == i) d[i] = A.val[k];
	auto strong = [&](int i, int k) {
		const int j = A.col[k];
		return j != i && std::fabs(A.val[k]) >= theta * std::sqrt(std::fabs(d[i] * d[j]));
	};
	agg->assign(n, -1);
	int na = 0;
	for (int i = 0; i < n; i++) {
		if ((*agg)[i] >= 0) continue;
		bool free = true;
		for (int k = A.ptr[i]; k < A.ptr[i + 1] && free; k++)
			if (strong(i, k) && (*agg)[A.col[k]] >= 0) free = false;
		if (!free) continue;
		(*agg)[i] = na;
		for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++)
			if (strong(i, k)) (*agg)[A.col[k]] = na;
		na++;
	}
	for (int i = 0; i < n; i++) {
		if ((*agg)[i] >= 0) continue;
		int best = -1;
		double bv = 0.0;
		for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) {
			const int j = A.col[k];
			if (j == i || (*agg)[j] < 0) continue;
			if (std::fabs(A.val[k]) > bv) { bv = std::fabs(A.val[k]); best = (*agg)[j]; }
		}
		(*agg)[i] = best >= 0 ? best : na++;
	}
	return na;
}

// Galerkin operator P^T A P for piecewise-constant P
Csr galerkin(const Csr& A, const std::vector<int>& agg, int nc)
{
	std::vector<std::vector<std::pair<int, double>>> rows(nc);
	for (int i = 0; i < A.n; i++) {
		const int I = agg[i];
		for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) rows[I].push_back({agg[A.col[k]], A.val[k]});
	}
	Csr C;
	C.n = nc;
	C.ptr.assign(nc + 1, 0);
	for (int I = 0; I < nc; I++) {
		auto& r = rows[I];
		std::sort(r.begin(), r.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
			return a.first < b.first;
		});
		for (size_t k = 0; k < r.size();) {
			size_t e = k;
			double s = 0.0;
			while (e < r.size() && r[e].first == r[k].first) s += r[e++].second;
			C.col.push_back(r[k].first);
			C.val.push_back(s);
			k = e;
		}
		C.ptr[I + 1] = (int) C.col.size();
		std::vector<std::pair<int, double>>().swap(r);
	}
	return C;
}

// largest eigenvalue of D^-1 A by power iteration (the damping of Jacobi)
double jacobi_lmax(const Csr& A)
{
	std::vector<double> d(A.n, 1.0), x(A.n), y(A.n);
	for (int i = 0; i < A.n; i++) {
		for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++)
			if (A.col[k] == i) d[i] = A.val[k];
		x[i] = 1.0 + 0.5 * std::sin(0.7 * i);
	}
	double lam = 1.0;
	for (int it = 0; it < 40; it++) {
		double ny = 0.0, nx = 0.0;
		for (int i = 0; i < A.n; i++) {
			double s = 0.0;
			for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) s += A.val[k] * x[A.col[k]];
			y[i] = s / d[i];
			ny += y[i] * y[i];
			nx += x[i] * x[i];
		}
		lam = std::sqrt(ny / nx);
		const double inv = 1.0 / std::sqrt(ny);
		for (int i = 0; i < A.n; i++) x[i] = y[i] * inv;
	}
	return lam;
}

double smoother_weight(const Csr& A) { return std::min(1.0, kAmgSmoothFactor * 4.0 / (3.0 * jacobi_lmax(A))); }

// dense inverse by Gauss-Jordan with partial pivoting (the coarsest level)
bool dense_inverse(const Csr& A, std::vector<float>* out)
{
	const int n = A.n;
	std::vector<double> a((size_t) n * n, 0.0), inv((size_t) n * n, 0.0);
	for (int i = 0; i < n; i++) {
		for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) a[(size_t) i * n + A.col[k]] += A.val[k];
		inv[(size_t) i * n + i] = 1.0;
	}
	for (int c = 0; c < n; c++) {
		int p = c;
		for (int r = c + 1; r < n; r++)
			if (std::fabs(a[(size_t) r * n + c]) > std::fabs(a[(size_t) p * n + c])) p = r;
		if (a[(size_t) p * n + c] == 0.0) return false;
		if (p != c)
			for (int k = 0; k < n; k++) {
				std::swap(a[(size_t) p * n + k], a[(size_t) c * n + k]);
				std::swap(inv[(size_t) p * n + k], inv[(size_t) c * n + k]);
			}
		const double piv = 1.0 / a[(size_t) c * n + c];
		for (int k = 0; k < n; k++) {
			a[(size_t) c * n + k] *= piv;
			inv[(size_t) c * n + k] *= piv;
		}
		for (int r = 0; r < n; r++) {
			if (r == c) continue;
			const double f = a[(size_t) r * n + c];
			if (f == 0.0) continue;
			for (int k = 0; k < n; k++) {
				a[(size_t) r * n + k] -= f * a[(size_t) c * n + k];
				inv[(size_t) r * n + k] -= f * inv[(size_t) c * n + k];
			}
		}
	}
	out->resize((size_t) n * n);
	for (size_t k = 0; k < out->size(); k++) (*out)[k] = (float) inv[k];
	return true;
}

// the row arguments a captured graph holds by value
bool same_rows(const AccelArgs& a, const AccelArgs& b)
{
	return a.nslots == b.nslots && a.nlong == b.nlong && a.slot_node == b.slot_node && a.inner_deg == b.inner_deg &&
	       a.slice_off == b.slice_off && a.col == b.col && a.w == b.w && a.slot_deg == b.slot_deg &&
	       a.long_slots == b.long_slots && a.long_off == b.long_off && a.licol == b.licol && a.liw == b.liw;
}

template <typename T>
int upload(DevBuf* b, const std::vector<T>& v)
{
	b->alloc(sizeof(T) * std::max<size_t>(v.size(), 1));
	if (!b->ptr()) return -1;
	if (!v.empty() && hipMemcpy(b->ptr(), v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess) return -1;
	return 0;
}

}  // namespace

// Builds the hierarchy from the current correction rows: level 0 is the rows
// themselves (A = D - C over the unknown neighbours that enter each row's
// computePotential), coarser levels by aggregation until at most
// kAmgCoarsest rows remain.
int Grid::amg_setup()
{
	if (!accel_)
		if (int rc = accel_setup(0.9)) return rc;
	const int n = mesh_.n;
	// level 0 in a compact index u over the row slots in slot order; on the
	// device the level-0 vectors are slot-indexed
	std::vector<int> unode, uslot;
	std::vector<int> upos(n, -1);
	for (int slot = 0; slot < next_spare_; slot++) {
		const int node = slot_node_[slot];
		if (node < 0) continue;
		upos[node] = (int) unode.size();
		unode.push_back(node);
		uslot.push_back(slot);
	}
	Csr A;
	A.n = (int) unode.size();
	A.ptr.assign(A.n + 1, 0);
	const int nv = cap_slots_;
	std::vector<float> dn(nv, 1.0f);
	std::vector<int> cols;
	std::vector<double> coups;
	for (int u = 0; u < A.n; u++) {
		const int i = unode[u];
		double ws;
		row_entries(i, &cols, &coups, &ws);
		dn[uslot[u]] = (float) ws;
		std::vector<std::pair<int, double>> row;
		row.push_back({u, ws});
		for (size_t k = 0; k < cols.size(); k++)
			if (!(flags_[cols[k]] & 2) && upos[cols[k]] >= 0) row.push_back({upos[cols[k]], -coups[k]});
		std::sort(row.begin(), row.end());
		for (auto& e : row) {
			A.col.push_back(e.first);
			A.val.push_back(e.second);
		}
		A.ptr[u + 1] = (int) A.col.size();
	}
	amg_w0_ = (float) smoother_weight(A);
	// coarse levels (a captured iteration graph holds the old ones' pointers)
	amg_graph_drop();
	amg_levels_.clear();
	std::vector<int> agg0_slot(nv, -1);
	Csr fine = A;
	for (int l = 0; l < 12 && (l == 0 || fine.n > kAmgCoarsest); l++) {
		std::vector<int> agg;
		const int nc = aggregate(fine, 0.08, &agg);
		if (l > 0 && nc >= fine.n * 0.8) break;  // no useful coarsening left
		Csr C = galerkin(fine, agg, nc);
		amg_levels_.push_back(std::unique_ptr<AmgLevelHost>(new AmgLevelHost));
		AmgLevelHost& L = *amg_levels_.back();
		L.n = nc;
		// restriction members (finer-level indices; level 0 by slot)
		std::vector<int> rptr(nc + 1, 0), ridx(fine.n);
		for (int i = 0; i < fine.n; i++) rptr[agg[i] + 1]++;
		for (int I = 0; I < nc; I++) rptr[I + 1] += rptr[I];
		std::vector<int> fill(rptr.begin(), rptr.end() - 1);
		for (int i = 0; i < fine.n; i++) ridx[fill[agg[i]]++] = l == 0 ? uslot[i] : i;
		std::vector<float> val(C.val.begin(), C.val.end()), dinv(nc);
		for (int I = 0; I < nc; I++) {
			double d = 1.0;
			for (int k = C.ptr[I]; k < C.ptr[I + 1]; k++)
				if (C.col[k] == I) d = C.val[k];
			dinv[I] = (float) (1.0 / d);
		}
		L.w = (float) smoother_weight(C);
		if (l == 0)
			for (int u = 0; u < fine.n; u++) agg0_slot[uslot[u]] = agg[u];
		if (upload(&L.ptr, C.ptr) || upload(&L.col, C.col) || upload(&L.val, val) || upload(&L.dinv, dinv) ||
		    upload(&L.rptr, rptr) || upload(&L.ridx, ridx) || (l > 0 && upload(&L.agg, agg)))
			return fail(-2, "device allocation failed (multigrid levels)");
		L.b.alloc(sizeof(float) * nc);
		L.x.alloc(sizeof(float) * nc);
		L.r.alloc(sizeof(float) * nc);
		if (!L.b.ptr() || !L.x.ptr() || !L.r.ptr()) return fail(-2, "device allocation failed (multigrid vectors)");
		fine = std::move(C);
		if (fine.n <= kAmgCoarsest) {
			std::vector<float> minv;
			if (!dense_inverse(fine, &minv)) return fail(-3, "multigrid: singular coarsest operator");
			if (upload(&amg_minv_, minv)) return fail(-2, "device allocation failed (multigrid coarsest)");
		}
	}
	if (amg_levels_.empty() || fine.n > kAmgCoarsest) {
		amg_levels_.clear();
		return fail(-7, "multigrid: the rows do not coarsen");
	}
	if (upload(&amg_dn_, dn) || upload(&amg_agg0_, agg0_slot)) return fail(-2, "device allocation failed (multigrid)");
	for (DevBuf* b : {&amg_r_, &amg_z_, &amg_p_, &amg_q_, &amg_x0_, &amg_res_, &amg_delta_}) {
		b->alloc(sizeof(float) * nv);
		if (!b->ptr()) return fail(-2, "device allocation failed (multigrid vectors)");
		HIPCHK(hipMemset(b->ptr(), 0, sizeof(float) * nv));
	}
	const int nb = std::max(row_launch_blocks(cap_slots_, long_cap_), (cap_slots_ + 255) / 256) + 8;
	amg_part_.alloc(sizeof(double) * nb);
	amg_st_.alloc(sizeof(double) * kAmgScalars);
	if (!amg_part_.ptr() || !amg_st_.ptr()) return fail(-2, "device allocation failed (multigrid)");
	HIPCHK(hipMemset(amg_st_.ptr(), 0, sizeof(double) * kAmgScalars));
	amg_on_ = true;
	return 0;
}

// rows rewritten by a resync: the new row sums (by slot) of rows that remain
// (or became) unknown; a slot that lost its row gets a zero residual entry,
// which its aggregate's restriction keeps reading
void Grid::amg_touch(const std::vector<int>& nodes, const std::vector<int>& off_slots, std::vector<long long>* dn_addr,
                     std::vector<float>* dn_val, std::vector<long long>* res_addr)
{
	std::vector<int> cols;
	std::vector<double> coups;
	for (const int node : nodes) {
		const int slot = slot_of_[node];
		if (slot < 0) continue;
		double ws;
		row_entries(node, &cols, &coups, &ws);
		dn_addr->push_back(slot);
		dn_val->push_back((float) ws);
	}
	for (int slot : off_slots) res_addr->push_back(slot);
}

// PCG with one V-cycle as preconditioner on A delta = D r32, until the
// residual's 2-norm has fallen by f; delta (node-indexed) receives the solution
int Grid::amg_solve(const AccelArgs& x, float f, int* iters, float** delta)
{
	const int nlev = (int) amg_levels_.size();
	std::vector<AmgLevel> lv(nlev);
	std::vector<float> w(nlev);
	std::vector<float*> bl(nlev), xl(nlev), rl(nlev);
	for (int l = 0; l < nlev; l++) {
		AmgLevelHost& H = *amg_levels_[l];
		lv[l].n = H.n;
		lv[l].ptr = H.ptr.as<int>();
		lv[l].col = H.col.as<int>();
		lv[l].val = H.val.as<float>();
		lv[l].dinv = H.dinv.as<float>();
		lv[l].rptr = H.rptr.as<int>();
		lv[l].ridx = H.ridx.as<int>();
		lv[l].agg = l > 0 ? H.agg.as<int>() : nullptr;
		w[l] = H.w;
		bl[l] = H.b.as<float>();
		xl[l] = H.x.as<float>();
		rl[l] = H.r.as<float>();
	}
	AmgSolve S;
	S.x = x;
	S.dn = amg_dn_.as<float>();
	S.agg0 = amg_agg0_.as<int>();
	S.w0 = amg_w0_;
	S.wc = kAmgCoarseScale;
	S.lv = lv.data();
	S.w = w.data();
	S.nlev = nlev;
	S.r = amg_r_.as<float>();
	S.z = amg_z_.as<float>();
	S.p = amg_p_.as<float>();
	S.q = amg_q_.as<float>();
	S.x0 = amg_x0_.as<float>();
	S.res = amg_res_.as<float>();
	S.delta = amg_delta_.as<float>();
	S.bl = bl.data();
	S.xl = xl.data();
	S.rl = rl.data();
	S.minv = amg_minv_.as<float>();
	S.part = amg_part_.as<double>();
	S.st = amg_st_.as<double>();
	const double f2 = (double) f * (double) f;
	if (launch_amg_start(stream_, S, d_r32_.as<float>(), f2) || launch_amg_vcycle(stream_, S, 4) ||
	    launch_amg_dir(stream_, S, 1))
		return fail(-2, "multigrid launch failed");
	auto iteration = [&]() -> int {
		if (launch_amg_matvec_update(stream_, S) || launch_amg_vcycle(stream_, S, 3) || launch_amg_dir(stream_, S, 0))
			return -1;
		return 0;
	};
	// the iterations of one batch, captured once per row layout and replayed
	if (!amg_eager_ && (!amg_exec_ || !same_rows(amg_graph_args_, x))) {
		amg_graph_drop();
		HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
		int bad = 0;
		for (int b = 0; b < kAmgBatch; b++) bad |= iteration();
		hipGraph_t g = nullptr;
		const hipError_t e = hipStreamEndCapture(stream_, &g);
		if (bad || e != hipSuccess) {
			if (g) (void) hipGraphDestroy(g);
			return fail(-2, "multigrid: graph capture failed");
		}
		amg_graph_ = g;
		HIPCHK(hipGraphInstantiate(&amg_exec_, amg_graph_, nullptr, nullptr, 0));
		amg_graph_args_ = x;
	}
	if (!amg_host_) {
		HIPCHK(hipHostMalloc((void**) &amg_host_, sizeof(double) * 2 * kAmgScalars, hipHostMallocDefault));
		for (hipEvent_t& e : amg_ev_) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
	}
	// batches run ahead of the check: batch k's scalars are read back while
	// batch k+1 runs (once the target is met every launch returns at once)
	// (a solve that stalls in FP32 stops here; the FP64 refinement then
	// restarts from the true residual)
	const int max_batches = kAmgMaxIters / kAmgBatch;
	double* last = nullptr;
	for (int k = 0; k < max_batches; k++) {
		if (amg_eager_) {
			for (int b = 0; b < kAmgBatch; b++)
				if (iteration()) return fail(-2, "multigrid launch failed");
		} else {
			HIPCHK(hipGraphLaunch(amg_exec_, stream_));
		}
		double* slot = amg_host_ + (k & 1) * kAmgScalars;
		HIPCHK(hipMemcpyAsync(slot, S.st, sizeof(double) * kAmgScalars, hipMemcpyDeviceToHost, stream_));
		HIPCHK(hipEventRecord(amg_ev_[k & 1], stream_));
		if (k == 0) continue;
		HIPCHK(hipEventSynchronize(amg_ev_[(k - 1) & 1]));
		last = amg_host_ + ((k - 1) & 1) * kAmgScalars;
		if (last[kAmgDone] != 0.0) break;
	}
	// the final scalars through the pinned slot, on the grid's stream (a
	// synchronous copy would go through the null stream)
	double* fin = amg_host_;
	HIPCHK(hipMemcpyAsync(fin, S.st, sizeof(double) * kAmgScalars, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	if (!(fin[kAmgRr] == fin[kAmgRr])) return fail(-3, "multigrid: NaN in the correction solve");
	*iters = (int) fin[kAmgIters];
	*delta = S.delta;
	return 0;
}

void Grid::amg_graph_drop()
{
	if (amg_exec_) (void) hipGraphExecDestroy(amg_exec_);
	if (amg_graph_) (void) hipGraphDestroy(amg_graph_);
	amg_exec_ = nullptr;
	amg_graph_ = nullptr;
}

}  // namespace tsg
