// Host side of the device grid (see grid.hpp).
#include "grid.hpp"
#include "prand.hpp"

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <limits>

namespace tsg {

namespace {
constexpr int kSlice = 64;
// spare row slots for atoms that turn into vacuum: one per removable node, at
// least 4096 and at most 2^20 (memory: ~20 B per entry of the removable
// nodes' widest row, ~300 MB at the cap for BCC atoms of 14 neighbours)
constexpr int kSpareMin = 4096, kSpareMax = 1 << 20;
constexpr int kSortWindow = 256;  // rows sorted by length within each window of this many
constexpr double kEpsilon0 = 8.85418781762e-12;  // grid_3d.h:285
}  // namespace

DevBuf::~DevBuf() { free(); }
void DevBuf::alloc(size_t bytes)
{
	free();
	if (bytes == 0) bytes = 8;
	if (hipMalloc(&p_, bytes) != hipSuccess) { p_ = nullptr; n_ = 0; return; }
	n_ = bytes;
}
void DevBuf::free()
{
	if (p_) (void) hipFree(p_);
	p_ = nullptr;
	n_ = 0;
}

int Grid::fail(int code, const std::string& msg)
{
	err_ = msg;
	status_ = code;
	return code;
}

int Grid::hip_check(hipError_t e, const char* where)
{
	if (e == hipSuccess) return 0;
	return fail(-2, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(x)                                      \
	do {                                               \
		if (int rc__ = hip_check((x), #x)) return rc__; \
	} while (0)

Grid::Grid(int mode, const Mesh& mesh, const int16_t* ids, const Species* species, float temperature)
    : mode_(mode), f64_(mode == 0), mesh_(mesh), temperature_(temperature)
{
	const int n = mesh_.n;
	ids_.assign(ids, ids + n);
	for (int s = 0; s < 128; s++) species_[s] = species[s];
	for (int i = 0; i < n; i++)
		if (ids_[i] < 0 || ids_[i] > 127 || !species_[ids_[i]].defined) { fail(-1, "node with an undefined species id"); return; }
	flags_.resize(n);
	for (int i = 0; i < n; i++) {
		const Species& sp = species_[ids_[i]];
		flags_[i] = (mesh_.boundary[i] ? 1 : 0) | (sp.dirichlet ? 2 : 0) | (sp.neumann ? 4 : 0);
	}
	for (int i = 0; i < n; i++)
		if (mesh_.assoc[i] < 0 || mesh_.assoc[i] >= mesh_.m() || mesh_.ptr[i + 1] == mesh_.ptr[i]) {
			fail(-4, "node " + std::to_string(i) + " belongs to no tetrahedron");
			return;
		}
	const long long nnz = (long long) mesh_.idx.size();
	row_of_entry_.resize(nnz);
	mirror_.resize(nnz);
	for (int i = 0; i < n; i++)
		for (int k = mesh_.ptr[i]; k < mesh_.ptr[i + 1]; k++) row_of_entry_[k] = i;
	for (long long k = 0; k < nnz; k++) {
		const int i = (int) row_of_entry_[k], j = mesh_.idx[k];
		const int* b = mesh_.idx.data() + mesh_.ptr[j];
		const int* e = mesh_.idx.data() + mesh_.ptr[j + 1];
		const int* f = std::lower_bound(b, e, i);
		if (f == e || *f != i) { fail(-4, "asymmetric node adjacency"); return; }
		mirror_[k] = f - mesh_.idx.data();
	}
	if (hip_check(hipGetDevice(&device_), "hipGetDevice")) return;
	if (hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate")) return;

	// ---- geometry precompute on the device
	const int m = mesh_.m();
	DevBuf d_tv, d_tn, d_assoc, d_ntp, d_nti, d_rowof, d_cc64, d_cc32;
	DevBuf d_ah32, d_ah64, d_a132, d_d32, d_d64;
	d_xyz_.alloc(sizeof(float) * 3 * (size_t) n);
	d_tv.alloc(sizeof(int) * 4 * (size_t) m);
	d_tn.alloc(sizeof(int) * 4 * (size_t) m);
	d_assoc.alloc(sizeof(int) * (size_t) n);
	d_ntp.alloc(sizeof(int) * ((size_t) n + 1));
	d_nti.alloc(sizeof(int) * mesh_.nt_idx.size());
	d_ptr_.alloc(sizeof(int) * ((size_t) n + 1));
	d_idx_.alloc(sizeof(int) * (size_t) nnz);
	d_rowof.alloc(sizeof(int) * (size_t) nnz);
	d_cc64.alloc(sizeof(double) * 3 * (size_t) m);
	d_cc32.alloc(sizeof(float) * 3 * (size_t) m);
	d_ah32.alloc(sizeof(float) * (size_t) nnz);
	d_ah64.alloc(sizeof(double) * (size_t) nnz);
	d_a132.alloc(sizeof(float) * (size_t) nnz);
	d_d32.alloc(sizeof(float) * (size_t) nnz);
	d_d64.alloc(sizeof(double) * (size_t) nnz);
	d_status_.alloc(sizeof(int) * 4);
	if (!d_xyz_.ptr() || !d_tv.ptr() || !d_tn.ptr() || !d_nti.ptr() || !d_idx_.ptr() || !d_cc64.ptr() || !d_ah64.ptr() ||
	    !d_d64.ptr() || !d_status_.ptr()) {
		fail(-2, "device allocation failed (geometry)");
		return;
	}
	std::vector<int> rowof32(nnz);
	for (long long k = 0; k < nnz; k++) rowof32[k] = (int) row_of_entry_[k];
	if (hip_check(hipMemcpy(d_xyz_.ptr(), mesh_.xyz.data(), d_xyz_.bytes(), hipMemcpyHostToDevice), "upload xyz") ||
	    hip_check(hipMemcpy(d_tv.ptr(), mesh_.tet_v.data(), d_tv.bytes(), hipMemcpyHostToDevice), "upload tets") ||
	    hip_check(hipMemcpy(d_tn.ptr(), mesh_.tet_n.data(), d_tn.bytes(), hipMemcpyHostToDevice), "upload neighbours") ||
	    hip_check(hipMemcpy(d_assoc.ptr(), mesh_.assoc.data(), d_assoc.bytes(), hipMemcpyHostToDevice), "upload assoc") ||
	    hip_check(hipMemcpy(d_ntp.ptr(), mesh_.nt_ptr.data(), d_ntp.bytes(), hipMemcpyHostToDevice), "upload nt_ptr") ||
	    hip_check(hipMemcpy(d_nti.ptr(), mesh_.nt_idx.data(), d_nti.bytes(), hipMemcpyHostToDevice), "upload nt_idx") ||
	    hip_check(hipMemcpy(d_ptr_.ptr(), mesh_.ptr.data(), d_ptr_.bytes(), hipMemcpyHostToDevice), "upload ptr") ||
	    hip_check(hipMemcpy(d_idx_.ptr(), mesh_.idx.data(), d_idx_.bytes(), hipMemcpyHostToDevice), "upload idx") ||
	    hip_check(hipMemcpy(d_rowof.ptr(), rowof32.data(), d_rowof.bytes(), hipMemcpyHostToDevice), "upload row_of") ||
	    hip_check(hipMemsetAsync(d_status_.ptr(), 0, d_status_.bytes(), stream_), "clear status"))
		return;
	// the extents of Geometry_3d::Table (setExtents, setConstrainedExtents;
	// geometry_3d.cpp:1555-1610): the coordinates' bounding box, widened by
	// the circumcentres that open Voronoi faces begin and end with
	DevBuf d_ext;
	d_ext.alloc(sizeof(int) * 6);
	const int ext0[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
	if (!d_ext.ptr() || hip_check(hipMemcpy(d_ext.ptr(), ext0, sizeof(ext0), hipMemcpyHostToDevice), "extents")) {
		fail(-2, "device allocation failed (extents)");
		return;
	}
	if (launch_circumcentres(stream_, m, d_xyz_.as<float>(), d_tv.as<int>(), d_cc64.as<double>(), d_cc32.as<float>(),
	                         d_status_.as<int>()) ||
	    launch_hull_extents(stream_, m, d_tn.as<int>(), d_cc32.as<float>(), d_ext.as<int>()) ||
	    launch_edge_areas(stream_, n, nnz, d_xyz_.as<float>(), d_tv.as<int>(), d_tn.as<int>(), d_assoc.as<int>(),
	                      d_ntp.as<int>(), d_nti.as<int>(), d_ptr_.as<int>(), d_idx_.as<int>(), d_rowof.as<int>(),
	                      d_cc64.as<double>(), d_cc32.as<float>(), d_ah32.as<float>(), d_ah64.as<double>(),
	                      d_a132.as<float>(), d_d32.as<float>(), d_d64.as<double>(), d_status_.as<int>())) {
		fail(-2, "geometry kernel launch failed");
		return;
	}
	// cell volumes of computeCharge, needed only when a species carries charge
	bool charged = false;
	for (int id = 0; id < 128; id++)
		if (species_[id].defined && species_[id].charge_density != 0.0f) charged = true;
	DevBuf d_v32, d_v64;
	if (charged) {
		d_v32.alloc(sizeof(float) * (size_t) n);
		d_v64.alloc(sizeof(double) * (size_t) n);
		if (!d_v32.ptr() || !d_v64.ptr()) {
			fail(-2, "device allocation failed (cell volumes)");
			return;
		}
		if (launch_cell_volumes(stream_, n, d_xyz_.as<float>(), d_tv.as<int>(), d_tn.as<int>(), d_assoc.as<int>(),
		                        d_ntp.as<int>(), d_nti.as<int>(), d_ptr_.as<int>(), d_idx_.as<int>(),
		                        d_cc64.as<double>(), d_cc32.as<float>(), d_d32.as<float>(), d_d64.as<double>(),
		                        d_v32.as<float>(), d_v64.as<double>(), d_status_.as<int>())) {
			fail(-2, "cell volume kernel launch failed");
			return;
		}
	}
	int gst = 0;
	a_half32_.resize(nnz); a_half64_.resize(nnz); a_one32_.resize(nnz); dist32_.resize(nnz); dist64_.resize(nnz);
	if (hip_check(hipStreamSynchronize(stream_), "geometry kernels") ||
	    hip_check(hipMemcpy(&gst, d_status_.ptr(), sizeof(int), hipMemcpyDeviceToHost), "status") ||
	    hip_check(hipMemcpy(a_half32_.data(), d_ah32.ptr(), d_ah32.bytes(), hipMemcpyDeviceToHost), "areas") ||
	    hip_check(hipMemcpy(a_half64_.data(), d_ah64.ptr(), d_ah64.bytes(), hipMemcpyDeviceToHost), "areas") ||
	    hip_check(hipMemcpy(a_one32_.data(), d_a132.ptr(), d_a132.bytes(), hipMemcpyDeviceToHost), "areas") ||
	    hip_check(hipMemcpy(dist32_.data(), d_d32.ptr(), d_d32.bytes(), hipMemcpyDeviceToHost), "distances") ||
	    hip_check(hipMemcpy(dist64_.data(), d_d64.ptr(), d_d64.bytes(), hipMemcpyDeviceToHost), "distances"))
		return;
	if (charged) {
		cell_vol32_.resize(n);
		cell_vol64_.resize(n);
		if (hip_check(hipMemcpy(cell_vol32_.data(), d_v32.ptr(), d_v32.bytes(), hipMemcpyDeviceToHost), "volumes") ||
		    hip_check(hipMemcpy(cell_vol64_.data(), d_v64.ptr(), d_v64.bytes(), hipMemcpyDeviceToHost), "volumes"))
			return;
	}
	{
		int ext[6];
		if (hip_check(hipMemcpy(ext, d_ext.ptr(), sizeof(ext), hipMemcpyDeviceToHost), "extents")) return;
		for (int k = 0; k < 3; k++) {
			float lo = mesh_.xyz[k], hi = mesh_.xyz[k];
			for (int i = 1; i < n; i++) {
				const float v = mesh_.xyz[3 * (size_t) i + k];
				if (v < lo) lo = v;
				else if (v > hi) hi = v;
			}
			extents_[k] = lo;
			extents_[3 + k] = hi;
			float clo = lo, chi = hi;
			if (ext[k] != INT_MAX) {
				const float a = float_of_key(ext[k]), b = float_of_key(ext[3 + k]);
				if (a < clo) clo = a;
				if (b > chi) chi = b;
			}
			extents_[6 + k] = clo;
			extents_[9 + k] = chi;
		}
	}
	if (gst & 1) { fail(-4, "degenerate tetrahedron (circumcentre)"); return; }
	if (gst & 2) { fail(-4, "Voronoi face walk failed"); return; }

	// ---- device state of the relaxation
	const size_t rs = f64_ ? sizeof(double) : sizeof(float);
	d_coupfull_.alloc(rs * (size_t) nnz);
	d_flags_.alloc((size_t) n);
	d_evapF_.alloc(sizeof(float) * (size_t) n);
	d_evapE_.alloc(sizeof(float) * (size_t) n);
	d_eps_.alloc(sizeof(float) * (size_t) n);
	d_phi0_.alloc(rs * (size_t) n);
	d_phi1_.alloc(rs * (size_t) n);
	d_ctl_.alloc(sizeof(RelaxCtl));
	d_sum_.alloc(sizeof(double));
	d_local_out_.alloc(sizeof(long long) * 4);
	if (!d_coupfull_.ptr() || !d_phi0_.ptr() || !d_phi1_.ptr()) { fail(-2, "device allocation failed (grid)"); return; }
	if (sync()) return;
	// initial potentials: the species presets (Process::initialization, process.cpp:65-80)
	std::vector<double> phi(n);
	for (int i = 0; i < n; i++) phi[i] = species_[ids_[i]].phi;
	if (set_phi(phi.data(), phi.data(), 0, 0)) return;
	// per-node species data of the probability models
	std::vector<float> evf(n), eve(n), eps(n);
	for (int i = 0; i < n; i++) {
		evf[i] = species_[ids_[i]].evap_field;
		eve[i] = species_[ids_[i]].evap_energy;
		eps[i] = species_[ids_[i]].epsilon;
	}
	if (hip_check(hipMemcpy(d_evapF_.ptr(), evf.data(), d_evapF_.bytes(), hipMemcpyHostToDevice), "upload evapField") ||
	    hip_check(hipMemcpy(d_evapE_.ptr(), eve.data(), d_evapE_.bytes(), hipMemcpyHostToDevice), "upload evapEnergy") ||
	    hip_check(hipMemcpy(d_eps_.ptr(), eps.data(), d_eps_.bytes(), hipMemcpyHostToDevice), "upload epsilon"))
		return;
}

Grid::~Grid()
{
	// every stream drains before any event, graph, pinned buffer or stream
	// goes away: work queued on one stream may wait on an event of another
	// (the flight waits for the main stream's snapshot, the normaliser for
	// the surface values) or copy into pinned memory freed below
	for (hipStream_t s : {stream_, fstream_, sstream_})
		if (s) (void) hipStreamSynchronize(s);
	amg_graph_drop();
	for (FlightSlot& f : fslot_) {
		if (f.ready) (void) hipEventDestroy(f.ready);
		if (f.done) (void) hipEventDestroy(f.done);
		if (f.host) (void) hipHostFree(f.host);
		if (f.phases_host) (void) hipHostFree(f.phases_host);
		f.ready = f.done = nullptr;
		f.host = nullptr;
		f.phases_host = nullptr;
	}
	for (hipEvent_t e : {ev_start_, ev_stop_, sum_src_, sum_done_, amg_ev_[0], amg_ev_[1], sst_ev_, norm_src_, norm_done_,
	                     patch_ev_})
		if (e) (void) hipEventDestroy(e);
	for (void* p : {h_pinned_, warm_host_, (void*) sum_host_, (void*) amg_host_, (void*) sst_host_, (void*) choice_host_})
		if (p) (void) hipHostFree(p);
	for (hipStream_t s : {fstream_, sstream_, stream_})
		if (s) (void) hipStreamDestroy(s);
}

// Grid_3d::Table::computeCoupling (grid_3d.cpp:1392-1409) from the
// precomputed directed-face area of entry (n1 -> n2)
double Grid::compute_coupling(int n1, int n2, long long entry) const
{
	if (!(flags_[n1] & 4) && (flags_[n2] & 4)) return 0.0;
	const Species& s1 = species_[ids_[n1]];
	const Species& s2 = species_[ids_[n2]];
	double value = 2.0;
	if (f64_) {
		value *= (double) s1.epsilon * (double) s2.epsilon;
		value /= (double) s1.epsilon + (double) s2.epsilon;
		value *= a_half64_[entry];
		value /= dist64_[entry];
	} else {
		value *= s1.epsilon * s2.epsilon;   // float product, as the reference
		value /= s1.epsilon + s2.epsilon;
		value *= a_half32_[entry];
		value /= dist32_[entry];
		value = (double) (float) value;     // stored as float
	}
	return value;
}

// Grid_3d::Table::computeCharge (grid_3d.cpp:1373-1390): cell volume (closed
// at half the mean neighbour distance) times the species' charge density
// over epsilon0; stored as the grid's element type
int Grid::get_charges(double* out) const
{
	std::copy(charge_.begin(), charge_.end(), out);
	return 0;
}

double Grid::charge_of(int node) const
{
	const float rho = species_[ids_[node]].charge_density;
	if (rho == 0.0f) return 0.0;  // the reference's product is exactly 0 for the finite volume
	double value = f64_ ? cell_vol64_[node] : (double) cell_vol32_[node];
	value *= (double) rho;
	value /= 8.85418781762e-12;  // grid_3d.h:285
	return f64_ ? value : (double) (float) value;
}

int Grid::sync()
{
	const int n = mesh_.n;
	const long long nnz = (long long) mesh_.idx.size();
	coup_.assign(nnz, std::numeric_limits<double>::quiet_NaN());
	charge_.assign(n, 0.0);
	for (int i = 0; i < n; i++) {
		charge_[i] = charge_of(i);
		if (charge_[i] != charge_[i]) return fail(-3, "Grid_3d::Table::computeCharge(): NaN");
	}
	// couplings once per pair as computeCoupling(lower, higher), mirrored
	for (int i = 0; i < n; i++)
		for (int k = mesh_.ptr[i]; k < mesh_.ptr[i + 1]; k++) {
			if (coup_[k] == coup_[k]) continue;
			const long long kk = mirror_[k];
			if (coup_[kk] != coup_[kk]) {
				const double c = compute_coupling(i, mesh_.idx[k], k);
				if (c != c) return fail(-3, "Grid_3d::Table::computeCoupling(): NaN");
				coup_[k] = c;
				coup_[kk] = c;
			} else
				coup_[k] = coup_[kk];
		}
	// upload full-CSR couplings and flags
	if (f64_) {
		HIPCHK(hipMemcpy(d_coupfull_.ptr(), coup_.data(), sizeof(double) * nnz, hipMemcpyHostToDevice));
	} else {
		std::vector<float> c32(nnz);
		for (long long k = 0; k < nnz; k++) c32[k] = (float) coup_[k];
		HIPCHK(hipMemcpy(d_coupfull_.ptr(), c32.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
	}
	HIPCHK(hipMemcpy(d_flags_.ptr(), flags_.data(), (size_t) n, hipMemcpyHostToDevice));
	return build_rows();
}

bool Grid::participates(int row, int col) const
{
	if (flags_[row] & 4) return true;       // Neumann rows use every neighbour
	return !(flags_[col] & 4);              // others skip Neumann neighbours
}

// the entries of a row in ascending neighbour order and their coupling sum,
// accumulated in double in that order (computePotential's weightSum)
void Grid::row_entries(int node, std::vector<int>* cols, std::vector<double>* coups, double* ws) const
{
	cols->clear();
	coups->clear();
	double sum = 0.0;
	for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++) {
		const int j = mesh_.idx[k];
		if (!participates(node, j)) continue;
		cols->push_back(j);
		coups->push_back(coup_[k]);
		sum += coup_[k];
	}
	*ws = sum;
}

int Grid::build_rows()
{
	const int n = mesh_.n;
	ord_dirty_ = true;
	slot_of_.assign(n, -1);
	std::vector<int> rows;
	for (int i = 0; i < n; i++)
		if (!(flags_[i] & 2)) rows.push_back(i);
	nrows_ = (int) rows.size();
	const int nslices_main = (nrows_ + kSlice - 1) / kSlice;
	// new unknown rows come from removable atoms taken away (or purged): the
	// spare slices are as wide as the widest of their rows
	int maxdeg = 1;
	long long removable = 0;
	for (int i = 0; i < n; i++)
		if (species_[ids_[i]].removable && (flags_[i] & 2)) {
			maxdeg = std::max(maxdeg, mesh_.ptr[i + 1] - mesh_.ptr[i]);
			removable++;
		}
	spare_width_ = maxdeg;
	const long long spare_rows = std::min<long long>(kSpareMax, std::max<long long>(kSpareMin, removable + kSlice));
	const int nslices = nslices_main + (int) ((spare_rows + kSlice - 1) / kSlice);
	cap_slots_ = nslices * kSlice;
	nslots_ = nslices_main * kSlice;
	next_spare_ = nslots_;
	slot_node_.assign(cap_slots_, -1);
	slot_deg_.assign(cap_slots_, 0);
	slice_width_.assign(nslices, spare_width_);
	slice_off_.assign((size_t) nslices + 1, 0);
	std::vector<std::vector<int>> ecols(nrows_);
	std::vector<std::vector<double>> ecoup(nrows_);
	std::vector<double> ws(cap_slots_, 1.0), q(cap_slots_, 0.0);
	std::vector<double> wrow(nrows_);
	for (int r = 0; r < nrows_; r++) row_entries(rows[r], &ecols[r], &ecoup[r], &wrow[r]);
	// slot order: node order, except that each window of kSortWindow rows is
	// sorted by row length (longest first, stable), so that the 64 rows of a
	// slice have nearly equal length and a wavefront's trip count matches
	// its work; the gathered potentials keep the locality of the node order
	std::vector<int> order(nrows_);
	for (int r = 0; r < nrows_; r++) order[r] = r;
	for (int w0 = 0; w0 < nrows_; w0 += kSortWindow)
		std::stable_sort(order.begin() + w0, order.begin() + std::min(nrows_, w0 + kSortWindow),
		                 [&](int x, int y) { return ecols[x].size() > ecols[y].size(); });
	{
		std::vector<std::vector<int>> c2(nrows_);
		std::vector<std::vector<double>> k2(nrows_);
		for (int sl = 0; sl < nrows_; sl++) {
			const int r = order[sl];
			c2[sl].swap(ecols[r]);
			k2[sl].swap(ecoup[r]);
			slot_node_[sl] = rows[r];
			slot_deg_[sl] = (int) c2[sl].size();
			slot_of_[rows[r]] = sl;
			ws[sl] = wrow[r];
			q[sl] = charge_[rows[r]];
		}
		ecols.swap(c2);
		ecoup.swap(k2);
	}
	for (int s = 0; s < nslices_main; s++) {
		int w = 0;
		for (int l = 0; l < kSlice; l++) {
			const int r = s * kSlice + l;
			if (r < nrows_) w = std::max(w, slot_deg_[r]);
		}
		slice_width_[s] = w;
	}
	for (int s = 0; s < nslices; s++) slice_off_[s + 1] = slice_off_[s] + (long long) slice_width_[s] * kSlice;
	const long long nent = slice_off_[nslices];
	std::vector<int> col(nent, 0);
	std::vector<double> cp(nent, 0.0);
	for (int r = 0; r < nrows_; r++) {
		const long long base = slice_off_[r / kSlice] + (r % kSlice);
		for (int k = 0; k < slot_deg_[r]; k++) {
			col[base + (long long) kSlice * k] = ecols[r][k];
			cp[base + (long long) kSlice * k] = ecoup[r][k];
		}
	}
	const size_t rs = f64_ ? sizeof(double) : sizeof(float);
	d_slot_node_.alloc(sizeof(int) * cap_slots_);
	d_slot_deg_.alloc(sizeof(int) * cap_slots_);
	d_slice_off_.alloc(sizeof(long long) * slice_off_.size());
	d_col_.alloc(sizeof(int) * (size_t) nent);
	d_coup_.alloc(rs * (size_t) nent);
	d_ws_.alloc(sizeof(double) * cap_slots_);
	d_q_.alloc(sizeof(double) * cap_slots_);
	long_set_.clear();
	for (int sl = 0; sl < nrows_; sl++)
		if (slot_deg_[sl] > kLongRow) long_set_.insert(sl);
	long_cap_ = (int) long_set_.size() + 4096;
	d_long_.alloc(sizeof(int) * (size_t) long_cap_);
	d_long_off_.alloc(sizeof(long long) * (size_t) long_cap_);
	// the long rows' contiguous copies, with room for 4096 new long rows of 64 entries
	loff_.assign(cap_slots_, -1);
	lcap_.assign(cap_slots_, 0);
	lnext_ = 0;
	for (int sl : long_set_) {
		loff_[sl] = lnext_;
		lcap_[sl] = slot_deg_[sl];
		lnext_ += slot_deg_[sl];
	}
	lstore_cap_ = lnext_ + 4096ll * 64;
	{
		std::vector<int> lc(lstore_cap_, 0);
		std::vector<double> lk(lstore_cap_, 0.0);
		for (int sl : long_set_)
			for (int k = 0; k < slot_deg_[sl]; k++) {
				lc[loff_[sl] + k] = ecols[sl][k];
				lk[loff_[sl] + k] = ecoup[sl][k];
			}
		d_lcol_.alloc(sizeof(int) * (size_t) lstore_cap_);
		d_lcoup_.alloc(rs * (size_t) lstore_cap_);
		if (!d_lcol_.ptr() || !d_lcoup_.ptr() || !d_long_off_.ptr()) return fail(-2, "device allocation failed (long rows)");
		HIPCHK(hipMemcpy(d_lcol_.ptr(), lc.data(), sizeof(int) * lc.size(), hipMemcpyHostToDevice));
		if (f64_) {
			HIPCHK(hipMemcpy(d_lcoup_.ptr(), lk.data(), sizeof(double) * lk.size(), hipMemcpyHostToDevice));
		} else {
			std::vector<float> l32(lk.begin(), lk.end());
			HIPCHK(hipMemcpy(d_lcoup_.ptr(), l32.data(), sizeof(float) * l32.size(), hipMemcpyHostToDevice));
		}
	}
	// one partial per workgroup of the largest row launch (row_launch_blocks:
	// slot groups + long-row groups, padded to a multiple of eight)
	const size_t max_blocks = (size_t) row_launch_blocks(cap_slots_, long_cap_) + 8;
	d_block_max_.alloc(sizeof(double) * max_blocks);
	d_block_flag_.alloc(sizeof(int) * max_blocks);
	if (!d_long_.ptr() || !d_block_max_.ptr()) return fail(-2, "device allocation failed (rows)");
	long_dirty_ = true;
	if (int rc = upload_long()) return rc;
	if (!d_col_.ptr() || !d_coup_.ptr()) return fail(-2, "device allocation failed (rows)");
	HIPCHK(hipMemcpy(d_slot_node_.ptr(), slot_node_.data(), d_slot_node_.bytes(), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_slot_deg_.ptr(), slot_deg_.data(), d_slot_deg_.bytes(), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_slice_off_.ptr(), slice_off_.data(), d_slice_off_.bytes(), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_col_.ptr(), col.data(), d_col_.bytes(), hipMemcpyHostToDevice));
	if (f64_) {
		HIPCHK(hipMemcpy(d_coup_.ptr(), cp.data(), d_coup_.bytes(), hipMemcpyHostToDevice));
	} else {
		std::vector<float> c32(cp.begin(), cp.end());
		HIPCHK(hipMemcpy(d_coup_.ptr(), c32.data(), d_coup_.bytes(), hipMemcpyHostToDevice));
	}
	HIPCHK(hipMemcpy(d_ws_.ptr(), ws.data(), d_ws_.bytes(), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_q_.ptr(), q.data(), d_q_.bytes(), hipMemcpyHostToDevice));
	return 0;
}

SweepArgs Grid::sweep_args()
{
	SweepArgs a;
	a.nslots = next_spare_;
	a.slot_node = d_slot_node_.as<int>();
	a.slot_deg = d_slot_deg_.as<int>();
	a.slice_off = d_slice_off_.as<long long>();
	a.col = d_col_.as<int>();
	a.coup = d_coup_.ptr();
	a.ws = d_ws_.as<double>();
	a.q = d_q_.as<double>();
	a.phi0 = d_phi0_.ptr();
	a.phi1 = d_phi1_.ptr();
	a.ctl = d_ctl_.as<RelaxCtl>();
	a.block_max = d_block_max_.as<double>();
	a.block_flag = d_block_flag_.as<int>();
	a.threshold_mode = 1;
	a.long_slots = d_long_.as<int>();
	a.nlong = (int) long_set_.size();
	a.long_off = d_long_off_.as<long long>();
	a.lcol = d_lcol_.as<int>();
	a.lcoup = d_lcoup_.ptr();
	return a;
}

void Grid::set_long(int slot, bool on)
{
	if (on) long_dirty_ |= long_set_.insert(slot).second;
	else long_dirty_ |= long_set_.erase(slot) > 0;
}

// the long-row list to the device (ascending slots), when it changed
int Grid::upload_long()
{
	if (!long_dirty_) return 0;
	if ((int) long_set_.size() > long_cap_) return fail(-7, "long-row list full");
	std::vector<int> v(long_set_.begin(), long_set_.end());
	std::vector<long long> off(v.size());
	for (size_t k = 0; k < v.size(); k++) off[k] = loff_[v[k]];
	if (!v.empty()) {
		HIPCHK(hipMemcpyAsync(d_long_.ptr(), v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, stream_));
		HIPCHK(hipMemcpyAsync(d_long_off_.ptr(), off.data(), sizeof(long long) * off.size(), hipMemcpyHostToDevice, stream_));
	}
	HIPCHK(hipStreamSynchronize(stream_));
	long_dirty_ = false;
	return 0;
}

int Grid::arm_ctl(double threshold, int cycles, int queue)
{
	if (queue < 0 || queue > 1024) return fail(-1, "queue size must be in [0, 1024]");
	RelaxCtl c;
	std::memset(&c, 0, sizeof(c));
	c.threshold = threshold;
	c.cycle = cycles;
	c.queue_size = queue;
	HIPCHK(hipMemcpyAsync(d_ctl_.ptr(), &c, offsetof(RelaxCtl, queue), hipMemcpyHostToDevice, stream_));
	return 0;
}

// the unknown rows' slots in ascending node order: the order of the
// reference's one-thread in-place sweep (uploaded when the rows changed)
int Grid::ensure_inplace_order()
{
	if (!ord_dirty_ && d_ord_slot_.ptr()) return 0;
	std::vector<int> ord;
	ord.reserve(nrows_);
	for (int i = 0; i < mesh_.n; i++)
		if (slot_of_[i] >= 0) {
			ord.push_back(slot_of_[i]);
			if (slot_deg_[slot_of_[i]] > kInplaceCap)
				return fail(-7, "relax(cycleSize > 1): a row longer than the in-place staging area");
		}
	ord_count_ = (int) ord.size();
	if (d_ord_slot_.bytes() < sizeof(int) * ord.size()) d_ord_slot_.alloc(sizeof(int) * (ord.size() + 4096));
	if (!d_ord_slot_.ptr()) return fail(-2, "device allocation failed (in-place order)");
	HIPCHK(hipMemcpy(d_ord_slot_.ptr(), ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice));
	ord_dirty_ = false;
	return 0;
}

// the cycles of one relax(cycleSize) call starting from _phiSwitch = s
// (doRelax, grid_3d.cpp:1238-1258): the table's switch stays s for the whole
// call, so odd cycles are Jacobi (read s, write 1-s) and even cycles update
// buffer s in place; the decision kernel follows the last cycle
int Grid::launch_relax_call(const SweepArgs& a, int s, int cycles)
{
	for (int c = 1; c <= cycles; c++) {
		const bool last = c == cycles;
		if (c & 1) {
			if (launch_jacobi_cycle(stream_, a, s, f64_, last)) return fail(-2, "sweep launch failed");
		} else {
			InplaceArgs P;
			P.count = ord_count_;
			P.ord_slot = d_ord_slot_.as<int>();
			P.buf = s;
			P.last = last ? 1 : 0;
			if (launch_inplace_cycle(stream_, a, P, f64_)) return fail(-2, "in-place sweep launch failed");
		}
	}
	return 0;
}

int Grid::relax_cycles(int cycles, double* deviation)
{
	if (cycles < 1) return fail(-1, "relax(cycleSize): cycleSize must be positive");
	if (cycles > 1)
		if (int rc = ensure_inplace_order()) return rc;
	if (int rc = arm_ctl(-1.0, cycles, 0)) return rc;
	SweepArgs a = sweep_args();
	if (int rc = launch_relax_call(a, sw_, cycles)) return rc;
	RelaxCtl c;
	HIPCHK(hipMemcpyAsync(&c, d_ctl_.ptr(), offsetof(RelaxCtl, queue), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	if (c.done == 4) return fail(-7, "relax(cycleSize > 1): a row longer than the in-place staging area");
	if (c.done == 3) return fail(-3, "Grid_3d::Table::computePotential(): NaN");
	sw_ = (sw_ + cycles) & 1;
	if (deviation) *deviation = c.last_dev;
	return 0;
}

int Grid::relax_threshold(double threshold, int cycles, int queue, long long* iterations, double* fault_dev,
                          double* fault_slope, double* last_dev)
{
	if (cycles < 1) return fail(-1, "relax(threshold, cycleSize, queue): cycleSize must be positive");
	if (cycles > 1)
		if (int rc = ensure_inplace_order()) return rc;
	if (int rc = arm_ctl(threshold, cycles, queue)) return rc;
	SweepArgs a = sweep_args();
	const int sw0 = sw_;
	long long launched = 0;
	// the first batch is sized from the previous call's sweep count, later
	// batches are short so that few sweeps are launched after convergence
	// (those return at once on the device's done flag)
	// (with cycleSize > 1 a "sweep" is one relax(cycleSize) call of that many
	// cycles; its in-place cycles are long, so batches stay short)
	const long long cap = cycles > 1 ? 4 : 4096;
	int batch = (int) std::min<long long>(cap, std::max<long long>(cycles > 1 ? 1 : 8, last_sweeps_ * 7 / 8));
	RelaxCtl c;
	if (!ev_start_) {
		HIPCHK(hipEventCreate(&ev_start_));
		HIPCHK(hipEventCreate(&ev_stop_));
	}
	HIPCHK(hipEventRecord(ev_start_, stream_));
	for (;;) {
		for (int b = 0; b < batch; b++) {
			if (cycles == 1) {
				if (launch_jacobi_sweep(stream_, a, (int) ((sw0 + launched) & 1), f64_)) return fail(-2, "sweep launch failed");
			} else if (int rc = launch_relax_call(a, (int) ((sw0 + launched * cycles) & 1), cycles))
				return rc;
			launched++;
		}
		HIPCHK(hipMemcpyAsync(&c, d_ctl_.ptr(), offsetof(RelaxCtl, queue), hipMemcpyDeviceToHost, stream_));
		HIPCHK(hipStreamSynchronize(stream_));
		if (c.done) break;
		batch = (int) std::min<long long>(cycles > 1 ? 4 : 64, std::max<long long>(cycles > 1 ? 1 : 4, (last_sweeps_ - c.sweeps) / 2));
	}
	HIPCHK(hipEventRecord(ev_stop_, stream_));
	HIPCHK(hipEventSynchronize(ev_stop_));
	float ms = 0.0f;
	HIPCHK(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
	stat_sweep_ms_ += ms;
	stat_sweeps_ += c.sweeps;
	stat_launched_ += launched;
	last_sweeps_ = c.sweeps;
	sw_ = (int) ((sw0 + c.sweeps * cycles) & 1);
	if (last_dev) *last_dev = c.last_dev;
	if (c.done == 4) return fail(-7, "relax(cycleSize > 1): a row longer than the in-place staging area");
	if (c.done == 3) return fail(-3, "Grid_3d::Table::computePotential(): NaN");
	if (c.done == 2) {
		if (iterations) *iterations = -1 - c.iteration;
		if (fault_dev) *fault_dev = c.fault_dev;
		if (fault_slope) *fault_slope = c.fault_slope;
		return 0;
	}
	if (iterations) *iterations = c.iteration;
	return 0;
}

int Grid::local_relax(int node, int order, double threshold, int cycles, int queue, long long* iterations)
{
	const int n = mesh_.n;
	if (node < 0 || node >= n || order < 0 || cycles < 1) return fail(-1, "localRelax: bad arguments");
	if (queue < 0 || queue > 1024) return fail(-1, "queue size must be in [0, 1024]");
	// find_localNodes (grid_3d.cpp:1284-1315): BFS levels into an ascending
	// set (the primary node enters only as a neighbour's neighbour), then the
	// Dirichlet nodes are removed (grid_3d.cpp:713-723)
	std::vector<int> in_set;
	std::vector<int> cur(1, node), next;
	static thread_local std::vector<uint8_t> mark;
	if ((int) mark.size() < n) mark.assign(n, 0);
	for (int lvl = 0; lvl < order; lvl++) {
		next.clear();
		for (int u : cur)
			for (int k = mesh_.ptr[u]; k < mesh_.ptr[u + 1]; k++) {
				const int v = mesh_.idx[k];
				if (!mark[v]) { mark[v] = 1; next.push_back(v); in_set.push_back(v); }
			}
		cur.swap(next);
	}
	for (int v : in_set) mark[v] = 0;
	std::sort(in_set.begin(), in_set.end());
	std::vector<int> nodes, slots;
	for (int v : in_set)
		if (!(flags_[v] & 2)) { nodes.push_back(v); slots.push_back(slot_of_[v]); }
	const int count = (int) nodes.size();
	if (count == 0) {
		// an empty shell: one call of doLocalRelax toggles the switch, the
		// deviation over no nodes is 0
		lsw_ = (lsw_ + cycles) % 2;
		if (iterations) *iterations = cycles;
		return 0;
	}
	// level schedule of the ascending in-place sweep: a node comes after
	// every lower-index shell neighbour, whichever of the two reads the
	// other -- v must see a lower neighbour's new value if it reads it, and a
	// lower Neumann neighbour that reads v (while v does not read it) must see
	// v's old value; both need the neighbour's level below v's
	std::vector<int> level(count, 0), pos_of;
	int nlev = 1;
	{
		// map node -> shell position through a sorted search
		for (int p = 0; p < count; p++) {
			const int v = nodes[p];
			int lv = 0;
			for (int k = mesh_.ptr[v]; k < mesh_.ptr[v + 1]; k++) {
				const int j = mesh_.idx[k];
				if (j >= v) break;  // ascending: only lower-index neighbours precede
				const auto it = std::lower_bound(nodes.begin(), nodes.end(), j);
				if (it != nodes.end() && *it == j) lv = std::max(lv, level[it - nodes.begin()] + 1);
			}
			level[p] = lv;
			nlev = std::max(nlev, lv + 1);
		}
	}
	std::vector<int> level_ptr(nlev + 1, 0), order_v(count);
	for (int p = 0; p < count; p++) level_ptr[level[p] + 1]++;
	for (int l = 0; l < nlev; l++) level_ptr[l + 1] += level_ptr[l];
	{
		std::vector<int> fill(level_ptr.begin(), level_ptr.end() - 1);
		for (int p = 0; p < count; p++) order_v[fill[level[p]]++] = p;
	}
	if (d_shell_nodes_.bytes() < sizeof(int) * (size_t) count) {
		d_shell_nodes_.alloc(sizeof(int) * 2 * (size_t) count);
		d_shell_slots_.alloc(sizeof(int) * 2 * (size_t) count);
		d_order_.alloc(sizeof(int) * 2 * (size_t) count);
	}
	if (d_level_ptr_.bytes() < sizeof(int) * level_ptr.size()) d_level_ptr_.alloc(sizeof(int) * 2 * level_ptr.size());
	HIPCHK(hipMemcpyAsync(d_shell_nodes_.ptr(), nodes.data(), sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemcpyAsync(d_shell_slots_.ptr(), slots.data(), sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemcpyAsync(d_order_.ptr(), order_v.data(), sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemcpyAsync(d_level_ptr_.ptr(), level_ptr.data(), sizeof(int) * level_ptr.size(), hipMemcpyHostToDevice, stream_));
	LocalArgs L;
	L.count = count;
	L.nodes = d_shell_nodes_.as<int>();
	L.slots = d_shell_slots_.as<int>();
	L.level_ptr = d_level_ptr_.as<int>();
	L.order = d_order_.as<int>();
	L.nlevels = nlev;
	L.read_buf = sw_;
	L.local_switch = lsw_;
	L.cycle = cycles;
	L.queue = queue;
	L.threshold = threshold;
	L.max_iterations = 100000000ll;
	L.out = d_local_out_.as<long long>();
	SweepArgs a = sweep_args();
	if (launch_local_relax(stream_, a, L, f64_)) return fail(-2, "local relax launch failed");
	long long out[3];
	HIPCHK(hipMemcpyAsync(out, d_local_out_.ptr(), sizeof(out), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	lsw_ = (int) out[1];
	if (out[2] == 3) return fail(-3, "Grid_3d::Table::computePotential(): NaN");
	if (out[2] == 4) return fail(-5, "localRelax: iteration limit");
	if (iterations) *iterations = out[0];
	return 0;
}

int Grid::get_phi(int buffer, double* out)
{
	const int n = mesh_.n;
	const void* src = buffer ? d_phi1_.ptr() : d_phi0_.ptr();
	HIPCHK(hipStreamSynchronize(stream_));
	if (f64_) {
		HIPCHK(hipMemcpy(out, src, sizeof(double) * n, hipMemcpyDeviceToHost));
	} else {
		std::vector<float> t(n);
		HIPCHK(hipMemcpy(t.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
		for (int i = 0; i < n; i++) out[i] = t[i];
	}
	return 0;
}

int Grid::set_phi(const double* p0, const double* p1, int phi_switch, int local_switch)
{
	const int n = mesh_.n;
	HIPCHK(hipStreamSynchronize(stream_));
	if (f64_) {
		HIPCHK(hipMemcpy(d_phi0_.ptr(), p0, sizeof(double) * n, hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(d_phi1_.ptr(), p1, sizeof(double) * n, hipMemcpyHostToDevice));
	} else {
		std::vector<float> t(n);
		for (int i = 0; i < n; i++) t[i] = (float) p0[i];
		HIPCHK(hipMemcpy(d_phi0_.ptr(), t.data(), sizeof(float) * n, hipMemcpyHostToDevice));
		for (int i = 0; i < n; i++) t[i] = (float) p1[i];
		HIPCHK(hipMemcpy(d_phi1_.ptr(), t.data(), sizeof(float) * n, hipMemcpyHostToDevice));
	}
	sw_ = phi_switch & 1;
	lsw_ = local_switch & 1;
	return 0;
}

// Grid_3d::Table::potential(i) = phi[(_phiSwitch + 1) % 2] (grid_3d.h:184)
int Grid::potential(int node, double* out)
{
	if (node < 0 || node >= mesh_.n) return fail(-1, "potential: bad node");
	const int b = (sw_ + 1) % 2;
	const char* base = (const char*) (b ? d_phi1_.ptr() : d_phi0_.ptr());
	if (f64_) {
		HIPCHK(hipMemcpyAsync(out, base + sizeof(double) * node, sizeof(double), hipMemcpyDeviceToHost, stream_));
		HIPCHK(hipStreamSynchronize(stream_));
	} else {
		float v;
		HIPCHK(hipMemcpyAsync(&v, base + sizeof(float) * node, sizeof(float), hipMemcpyDeviceToHost, stream_));
		HIPCHK(hipStreamSynchronize(stream_));
		*out = v;
	}
	return 0;
}

int Grid::field_o1_staged(int count, const int* nodes, double* out, int* d_nodes, double* d_fields, int* d_status)
{
	if (count <= 0) return 0;
	for (int i = 0; i < count; i++)
		if (nodes[i] < 0 || nodes[i] >= mesh_.n) return fail(-1, "field_o1: bad node");
	HIPCHK(hipMemcpyAsync(d_nodes, nodes, sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	// the status word is cleared by a kernel (no memset node if the stream is captured)
	if (launch_fill_i32(stream_, 1, d_status, 0)) return fail(-2, "fill launch failed");
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	if (launch_node_fields(stream_, count, d_nodes, d_xyz_.as<float>(), d_ptr_.as<int>(), d_idx_.as<int>(),
	                       d_coupfull_.ptr(), d_flags_.as<unsigned char>(), phi, f64_, d_fields, d_status))
		return fail(-2, "field launch failed");
	int st = 0;
	HIPCHK(hipMemcpyAsync(out, d_fields, sizeof(double) * 3 * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(&st, d_status, sizeof(int), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	if (st) return fail(-4, "Grid_3d::Table::field_o1(): singular least-squares system");
	return 0;
}

int Grid::field_o1(int count, const int* nodes, double* out)
{
	if (count <= 0) return 0;
	for (int i = 0; i < count; i++)
		if (nodes[i] < 0 || nodes[i] >= mesh_.n) return fail(-1, "field_o1: bad node");
	if (d_nodes_.bytes() < sizeof(int) * (size_t) count || d_fields_.bytes() < sizeof(double) * 3 * (size_t) count) {
		d_nodes_.alloc(sizeof(int) * (size_t) count);
		d_fields_.alloc(sizeof(double) * 4 * (size_t) count);
		if (!d_nodes_.ptr() || !d_fields_.ptr()) return fail(-2, "device allocation failed (fields)");
	}
	HIPCHK(hipMemcpyAsync(d_nodes_.ptr(), nodes, sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemsetAsync(d_status_.ptr(), 0, sizeof(int), stream_));
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	if (launch_node_fields(stream_, count, d_nodes_.as<int>(), d_xyz_.as<float>(), d_ptr_.as<int>(), d_idx_.as<int>(),
	                       d_coupfull_.ptr(), d_flags_.as<unsigned char>(), phi, f64_, d_fields_.as<double>(),
	                       d_status_.as<int>()))
		return fail(-2, "field launch failed");
	int st = 0;
	HIPCHK(hipMemcpyAsync(out, d_fields_.ptr(), sizeof(double) * 3 * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(&st, d_status_.ptr(), sizeof(int), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	if (st) return fail(-4, "Grid_3d::Table::field_o1(): singular least-squares system");
	return 0;
}

// Grid_3d::Table::field_o2 (grid_3d.cpp:993-1199) of a node list in the
// potentials potential() reads: fields and the charge of the quadratic fit
int Grid::field_o2(int count, const int* nodes, double* out, double* charge)
{
	if (count <= 0) return 0;
	for (int i = 0; i < count; i++)
		if (nodes[i] < 0 || nodes[i] >= mesh_.n) return fail(-1, "field_o2: bad node");
	if (d_nodes_.bytes() < sizeof(int) * (size_t) count || d_fields_.bytes() < sizeof(double) * 4 * (size_t) count) {
		d_nodes_.alloc(sizeof(int) * (size_t) count);
		d_fields_.alloc(sizeof(double) * 4 * (size_t) count);
		if (!d_nodes_.ptr() || !d_fields_.ptr()) return fail(-2, "device allocation failed (fields)");
	}
	HIPCHK(hipMemcpyAsync(d_nodes_.ptr(), nodes, sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemsetAsync(d_status_.ptr(), 0, sizeof(int), stream_));
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	double* fo = d_fields_.as<double>();
	if (launch_node_fields_o2(stream_, count, d_nodes_.as<int>(), d_xyz_.as<float>(), d_ptr_.as<int>(), d_idx_.as<int>(),
	                          d_flags_.as<unsigned char>(), phi, f64_, fo, fo + 3 * (size_t) count, d_status_.as<int>()))
		return fail(-2, "field launch failed");
	int st = 0;
	HIPCHK(hipMemcpyAsync(out, fo, sizeof(double) * 3 * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(charge, fo + 3 * (size_t) count, sizeof(double) * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(&st, d_status_.ptr(), sizeof(int), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	if (st & 4) return fail(-4, "Grid_3d::Table::field_o2(): singular least-squares system");
	if (st & 2) return fail(-3, "Grid_3d::Table::field_o2(): NaN");
	return 0;
}

int Grid::get_couplings(double* out) const
{
	std::copy(coup_.begin(), coup_.end(), out);
	return 0;
}
int Grid::get_flags(unsigned char* out) const
{
	std::copy(flags_.begin(), flags_.end(), out);
	return 0;
}
int Grid::get_ids(int16_t* out) const
{
	std::copy(ids_.begin(), ids_.end(), out);
	return 0;
}
int Grid::get_areas(float* a_half32, double* a_half64, float* a_one32) const
{
	if (a_half32) std::copy(a_half32_.begin(), a_half32_.end(), a_half32);
	if (a_half64) std::copy(a_half64_.begin(), a_half64_.end(), a_half64);
	if (a_one32) std::copy(a_one32_.begin(), a_one32_.end(), a_one32);
	return 0;
}

long long Grid::entry_index(int row, int col) const
{
	const int* b = mesh_.idx.data() + mesh_.ptr[row];
	const int* e = mesh_.idx.data() + mesh_.ptr[row + 1];
	const int* f = std::lower_bound(b, e, col);
	if (f == e || *f != col) return -1;
	return f - mesh_.idx.data();
}

// room for a long row of deg entries in the contiguous store: the slot's own
// if large enough, else fresh room at the end (the offsets list is uploaded
// again); 0 or -1 when the store is full
int Grid::long_storage(int slot, int deg)
{
	if (loff_[slot] >= 0 && lcap_[slot] >= deg) return 0;
	if (lnext_ + deg > lstore_cap_) return -1;
	loff_[slot] = lnext_;
	lcap_[slot] = deg;
	lnext_ += deg;
	long_dirty_ = true;
	return 0;
}

// (re)writes a row into a slot: slot arrays (int patches to target 0/1),
// entries (col int patches target 2, coupling real patches target 0), and
// ws/q (real patches targets 1/2)
int Grid::write_slot(int slot, int node, std::vector<long long>* ai, std::vector<int>* vi, std::vector<long long>* ar,
                     std::vector<double>* vr, std::vector<long long>* aw, std::vector<double>* vw)
{
	std::vector<int> cols;
	std::vector<double> coups;
	double ws;
	row_entries(node, &cols, &coups, &ws);
	const int deg = (int) cols.size();
	if (deg > slice_width_[slot / kSlice]) return -1;
	slot_node_[slot] = node;
	slot_deg_[slot] = deg;
	slot_of_[node] = slot;
	set_long(slot, deg > kLongRow);
	// int patches are encoded as (array id << 60) | index
	ai->push_back((0ll << 60) | slot); vi->push_back(node);
	ai->push_back((1ll << 60) | slot); vi->push_back(deg);
	const long long base = slice_off_[slot / kSlice] + (slot % kSlice);
	for (int k = 0; k < deg; k++) {
		ai->push_back((2ll << 60) | (base + (long long) kSlice * k)); vi->push_back(cols[k]);
		ar->push_back(base + (long long) kSlice * k); vr->push_back(coups[k]);
	}
	if (deg > kLongRow) {  // and its contiguous copy
		if (long_storage(slot, deg)) return -2;
		for (int k = 0; k < deg; k++) {
			lpi_addr_.push_back((0ll << 60) | (loff_[slot] + k)); lpi_val_.push_back(cols[k]);
			lpr_addr_.push_back(loff_[slot] + k); lpr_val_.push_back(coups[k]);
		}
	}
	aw->push_back((1ll << 60) | slot); vw->push_back(ws);
	aw->push_back((2ll << 60) | slot); vw->push_back(charge_[node]);
	return 0;
}

int Grid::set_id_and_resync(int node, int id)
{
	if (node < 0 || node >= mesh_.n || id < 0 || id > 127 || !species_[id].defined)
		return fail(-1, "setId: bad node or species");
	ids_[node] = (int16_t) id;
	refresh_vacuum(node);  // the device surface's vacuum bit (the site set itself changes only in take_away)
	const Species& sp = species_[id];
	flags_[node] = (mesh_.boundary[node] ? 1 : 0) | (sp.dirichlet ? 2 : 0) | (sp.neumann ? 4 : 0);
	ord_dirty_ = true;  // rows may appear, move or vanish

	// couplings of the node's faces as computeCoupling(node, neighbour)
	std::vector<long long> full_addr;
	std::vector<double> full_val;
	for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++) {
		const int j = mesh_.idx[k];
		if (!(!(flags_[node] & 4) && (flags_[j] & 4)) && !species_ok(j))
			return fail(-8, "Configuration::Table::operator[]()");
		const double c = compute_coupling(node, j, k);
		if (c != c) return fail(-3, "Grid_3d::Table::computeCoupling(): NaN");
		coup_[k] = c;
		coup_[mirror_[k]] = c;
		full_addr.push_back(k); full_val.push_back(c);
		full_addr.push_back(mirror_[k]); full_val.push_back(c);
	}
	charge_[node] = charge_of(node);
	if (charge_[node] != charge_[node]) return fail(-3, "Grid_3d::Table::computeCharge(): NaN");

	// rows touched: the node itself and every neighbour that owns a row
	lpi_addr_.clear(); lpi_val_.clear(); lpr_addr_.clear(); lpr_val_.clear(); lpf_addr_.clear(); lpf_val_.clear();
	std::vector<long long> ai, ar, aw;
	std::vector<int> vi;
	std::vector<double> vr, vw;
	std::vector<int> touched;
	touched.push_back(node);
	for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++) touched.push_back(mesh_.idx[k]);
	// slots that lose their row (switched off, or moved to a spare slot) and
	// the rows that moved: the correction rows of a moved row's neighbours
	// name it by its slot, so they are rewritten too
	std::vector<int> off_slots, moved;
	for (int r : touched) {
		const bool unknown = !(flags_[r] & 2);
		int slot = slot_of_[r];
		if (!unknown) {
			if (slot >= 0) {  // a row that became Dirichlet is switched off
				slot_node_[slot] = -1;
				set_long(slot, false);
				ai.push_back((0ll << 60) | slot); vi.push_back(-1);
				slot_of_[r] = -1;
				off_slots.push_back(slot);
			}
			continue;
		}
		if (slot >= 0) {
			const int w = write_slot(slot, r, &ai, &vi, &ar, &vr, &aw, &vw);
			if (w == 0) continue;
			if (w == -2) return fail(-7, "long-row store full");
		}
		if (slot >= 0) {  // does not fit its slice any more
			slot_node_[slot] = -1;
			set_long(slot, false);
			ai.push_back((0ll << 60) | slot); vi.push_back(-1);
			off_slots.push_back(slot);
			moved.push_back(r);
		}
		if (next_spare_ >= cap_slots_) return fail(-7, "row slots exhausted (more new unknown nodes than spare slots)");
		slot = next_spare_++;
		if (const int w = write_slot(slot, r, &ai, &vi, &ar, &vr, &aw, &vw))
			return fail(-7, w == -2 ? "long-row store full" : "row wider than the spare slices");
		nrows_++;
	}
	// the FP32 correction rows of the accelerated relaxation follow the
	// same slots
	std::vector<long long> fa;
	std::vector<float> fv;
	if (accel_) {
		std::vector<int> inner_rows(touched);
		for (int r : moved)
			for (int k = mesh_.ptr[r]; k < mesh_.ptr[r + 1]; k++) inner_rows.push_back(mesh_.idx[k]);
		std::sort(inner_rows.begin(), inner_rows.end());
		inner_rows.erase(std::unique(inner_rows.begin(), inner_rows.end()), inner_rows.end());
		for (int r : inner_rows)
			if (slot_of_[r] >= 0)
				if (int rc = write_inner(slot_of_[r], r, &ai, &vi, &fa, &fv))
					return fail(-7, rc == -2 ? "inner row: an unknown neighbour without a row" : "inner row too wide");
	}
	std::vector<long long> dn_addr, res_addr;
	std::vector<float> dn_val;
	if (amg_on_) amg_touch(touched, off_slots, &dn_addr, &dn_val, &res_addr);

	// ship all patches through the pinned staging area
	const size_t nr = ar.size(), nw = aw.size(), nf = full_addr.size();
	const size_t rs = f64_ ? sizeof(double) : sizeof(float);
	const size_t need = ai.size() * 12 + nr * (8 + rs) + nw * 16 + nf * (8 + rs) + fa.size() * 12 + lpi_addr_.size() * 12 +
	                    lpr_addr_.size() * (8 + rs) + lpf_addr_.size() * 12 + dn_addr.size() * 12 + res_addr.size() * 12 +
	                    48 * 16 + 1024;
	// the staging's previous upload is done (nothing else waits for the device)
	if (patch_ev_) HIPCHK(hipEventSynchronize(patch_ev_));
	else HIPCHK(hipEventCreateWithFlags(&patch_ev_, hipEventDisableTiming));
	if (need > h_pinned_bytes_) {
		// the device copy may still be read by the previous patch kernels
		HIPCHK(hipStreamSynchronize(stream_));
		if (h_pinned_) (void) hipHostFree(h_pinned_);
		h_pinned_bytes_ = std::max(need * 2, (size_t) 1 << 20);
		HIPCHK(hipHostMalloc(&h_pinned_, h_pinned_bytes_, hipHostMallocDefault));
		d_patch_addr_.alloc(h_pinned_bytes_);
	}
	char* h = (char*) h_pinned_;
	size_t off = 0;
	auto put = [&](const void* src, size_t bytes) {
		std::memcpy(h + off, src, bytes);
		const size_t at = off;
		off += (bytes + 15) & ~(size_t) 15;
		return at;
	};
	// split int patches by target array
	std::vector<long long> ia[5];
	std::vector<int> iv[5];
	for (size_t t = 0; t < ai.size(); t++) {
		const int arr = (int) (ai[t] >> 60);
		ia[arr].push_back(ai[t] & ((1ll << 60) - 1));
		iv[arr].push_back(vi[t]);
	}
	std::vector<long long> wa[3];
	std::vector<double> wv[3];
	for (size_t t = 0; t < nw; t++) {
		const int arr = (int) (aw[t] >> 60);
		wa[arr].push_back(aw[t] & ((1ll << 60) - 1));
		wv[arr].push_back(vw[t]);
	}
	size_t o_ia[5], o_iv[5], o_wa[3], o_wv[3];
	for (int t = 0; t < 5; t++) {
		o_ia[t] = put(ia[t].data(), sizeof(long long) * ia[t].size());
		o_iv[t] = put(iv[t].data(), sizeof(int) * iv[t].size());
	}
	for (int t = 0; t < 3; t++) {
		o_wa[t] = put(wa[t].data(), sizeof(long long) * wa[t].size());
		o_wv[t] = put(wv[t].data(), sizeof(double) * wv[t].size());
	}
	const size_t o_ia3 = put(fa.data(), sizeof(long long) * fa.size());
	const size_t o_iw3 = put(fv.data(), sizeof(float) * fv.size());
	// the long rows' contiguous copies: int patches split by target
	std::vector<long long> la[2];
	std::vector<int> lv[2];
	for (size_t t = 0; t < lpi_addr_.size(); t++) {
		const int arr = (int) (lpi_addr_[t] >> 60);
		la[arr].push_back(lpi_addr_[t] & ((1ll << 60) - 1));
		lv[arr].push_back(lpi_val_[t]);
	}
	size_t o_la[2], o_lv[2];
	for (int t = 0; t < 2; t++) {
		o_la[t] = put(la[t].data(), sizeof(long long) * la[t].size());
		o_lv[t] = put(lv[t].data(), sizeof(int) * lv[t].size());
	}
	const size_t o_lra = put(lpr_addr_.data(), sizeof(long long) * lpr_addr_.size());
	size_t o_lrv;
	if (f64_) {
		o_lrv = put(lpr_val_.data(), sizeof(double) * lpr_val_.size());
	} else {
		std::vector<float> l32(lpr_val_.begin(), lpr_val_.end());
		o_lrv = put(l32.data(), sizeof(float) * l32.size());
	}
	const size_t o_lfa = put(lpf_addr_.data(), sizeof(long long) * lpf_addr_.size());
	const size_t o_lfv = put(lpf_val_.data(), sizeof(float) * lpf_val_.size());
	const size_t o_ra = put(ar.data(), sizeof(long long) * nr);
	const size_t o_fa = put(full_addr.data(), sizeof(long long) * nf);
	size_t o_rv, o_fv;
	if (f64_) {
		o_rv = put(vr.data(), sizeof(double) * nr);
		o_fv = put(full_val.data(), sizeof(double) * nf);
	} else {
		std::vector<float> a32(vr.begin(), vr.end()), b32(full_val.begin(), full_val.end());
		o_rv = put(a32.data(), sizeof(float) * nr);
		o_fv = put(b32.data(), sizeof(float) * nf);
	}
	const size_t o_dna = put(dn_addr.data(), sizeof(long long) * dn_addr.size());
	const size_t o_dnv = put(dn_val.data(), sizeof(float) * dn_val.size());
	const size_t o_rsa = put(res_addr.data(), sizeof(long long) * res_addr.size());
	const std::vector<float> res_zero(res_addr.size(), 0.0f);
	const size_t o_rsv = put(res_zero.data(), sizeof(float) * res_zero.size());
	char* d = (char*) d_patch_addr_.ptr();
	HIPCHK(hipMemcpyAsync(d, h, off, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipEventRecord(patch_ev_, stream_));
	// every patch below in one launch: the multigrid's level-0 row sums of the
	// touched rows and a zero residual in the slots that lost their row, the
	// slot tables, the row sums and charges, the couplings, the inner and long
	// rows, and the node's species values
	PatchBatch pb;
	auto job = [&](size_t n, int width, size_t o_addr, size_t o_val, void* dst) {
		if (n == 0) return;
		PatchJob& jb = pb.job[pb.count++];
		jb.n = (int) n;
		jb.width = width;
		jb.addr = (const long long*) (d + o_addr);
		jb.val = d + o_val;
		jb.dst = dst;
	};
	const int rw = f64_ ? 8 : 4;
	job(dn_addr.size(), 4, o_dna, o_dnv, amg_dn_.ptr());
	job(res_addr.size(), 4, o_rsa, o_rsv, amg_res_.ptr());
	void* iarr[5] = {d_slot_node_.ptr(), d_slot_deg_.ptr(), d_col_.ptr(), d_inner_deg_.ptr(), d_icol_.ptr()};
	for (int t = 0; t < 5; t++) job(ia[t].size(), 4, o_ia[t], o_iv[t], iarr[t]);
	void* warr[3] = {nullptr, d_ws_.ptr(), d_q_.ptr()};
	for (int t = 1; t < 3; t++) job(wa[t].size(), 8, o_wa[t], o_wv[t], warr[t]);
	job(nr, rw, o_ra, o_rv, d_coup_.ptr());
	job(nf, rw, o_fa, o_fv, d_coupfull_.ptr());
	job(fa.size(), 4, o_ia3, o_iw3, d_iw_.ptr());
	void* larr[2] = {d_lcol_.ptr(), d_licol_.ptr()};
	for (int t = 0; t < 2; t++) job(la[t].size(), 4, o_la[t], o_lv[t], larr[t]);
	job(lpr_addr_.size(), rw, o_lra, o_lrv, d_lcoup_.ptr());
	job(lpf_addr_.size(), 4, o_lfa, o_lfv, d_liw_.ptr());
	NodeSpecies& ns = pb.node;
	ns.node = node;
	ns.flags = flags_[node];
	ns.evap_field = species_[id].evap_field;
	ns.evap_energy = species_[id].evap_energy;
	ns.epsilon = species_[id].epsilon;
	ns.d_flags = d_flags_.as<unsigned char>();
	ns.d_evap_field = d_evapF_.as<float>();
	ns.d_evap_energy = d_evapE_.as<float>();
	ns.d_epsilon = d_eps_.as<float>();
	if (launch_patch_batch(stream_, pb)) return fail(-2, "patch launch failed");
	// everything above is queued on the stream: the host goes on
	return upload_long();
}

// ---------------------------------------------------------------- surface

// Surface_3d::Table::init (surface_3d.cpp:50-81): a site is a removable
// non-vacuum node with a vacuum neighbour across a face of positive area
int Grid::surface_init(int vacuum_id)
{
	const int n = mesh_.n;
	if (normal_pending_) {  // it reads the bytes rewritten below
		HIPCHK(hipEventSynchronize(norm_done_));
		normal_pending_ = false;
	}
	vacuum_id_ = vacuum_id;
	sstate_.assign(n, 0);
	surf_count_ = 0;
	for (int i = 0; i < n; i++) {
		if (ids_[i] == vacuum_id) {
			sstate_[i] = kSurfVacuum;
			continue;
		}
		if (!species_[ids_[i]].removable) continue;
		for (int k = mesh_.ptr[i]; k < mesh_.ptr[i + 1]; k++)
			if (ids_[mesh_.idx[k]] == vacuum_id && a_one32_[k] > 0.0f) {
				sstate_[i] = kSurfSite;
				surf_count_++;
				break;
			}
	}
	const size_t nb = (size_t) surface_compact_blocks(n);
	if (d_sstate_.bytes() < (size_t) n) {
		d_sstate_.alloc(((size_t) n + 15) / 16 * 16);
		d_surf_.alloc(sizeof(int) * ((size_t) n + 1));
		d_surf_blk_.alloc(sizeof(int) * (nb + 1));
		d_marks_.alloc(sizeof(int) * (size_t) n);
		d_nshell_.alloc(surface_normal_workspace(n));
		d_choice_.alloc(sizeof(ChoiceRec));
		if (!d_sstate_.ptr() || !d_surf_.ptr() || !d_surf_blk_.ptr() || !d_marks_.ptr() || !d_nshell_.ptr() ||
		    !d_choice_.ptr())
			return fail(-2, "device allocation failed (surface)");
		// every entry of the list and the shell buffer holds a node index
		// from the start, the marks no call's stamp
		if (launch_fill_i32(stream_, n, d_marks_.as<int>(), 0) || launch_fill_i32(stream_, n + 1, d_surf_.as<int>(), 0) ||
		    launch_fill_i32(stream_, (int) (surface_normal_workspace(n) / sizeof(int)), d_nshell_.as<int>(), 0))
			return fail(-2, "fill launch failed");
		normal_stamp_ = 0;
	}
	if (!choice_host_) {
		HIPCHK(hipHostMalloc((void**) &choice_host_, sizeof(ChoiceRec), hipHostMallocDefault));
		std::memset(choice_host_, 0, sizeof(ChoiceRec));
	}
	HIPCHK(hipMemcpyAsync(d_sstate_.ptr(), sstate_.data(), (size_t) n, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	sst_pending_.clear();
	list_dirty_ = true;
	choice_valid_ = false;
	return surface_flush();
}

void Grid::set_site(int node, bool on)
{
	const uint8_t b = (uint8_t) ((sstate_[node] & ~kSurfSite) | (on ? kSurfSite : 0));
	if (b == sstate_[node]) return;
	surf_count_ += on ? 1 : -1;
	sstate_[node] = b;
	sst_pending_.push_back(node);
	list_dirty_ = true;
}

void Grid::refresh_vacuum(int node)
{
	if (sstate_.empty()) return;
	const uint8_t b = (uint8_t) ((sstate_[node] & ~kSurfVacuum) | (ids_[node] == vacuum_id_ ? kSurfVacuum : 0));
	if (b == sstate_[node]) return;
	sstate_[node] = b;
	sst_pending_.push_back(node);
}

// the changed state bytes to the device (pinned staging, one scatter) and,
// if a site bit changed, the site list rebuilt there; nothing waits for it
int Grid::surface_flush()
{
	if (!sst_pending_.empty()) {
		std::sort(sst_pending_.begin(), sst_pending_.end());
		sst_pending_.erase(std::unique(sst_pending_.begin(), sst_pending_.end()), sst_pending_.end());
		const size_t cnt = sst_pending_.size();
		const size_t bytes = cnt * (sizeof(int) + 1);
		if (!sst_ev_) HIPCHK(hipEventCreateWithFlags(&sst_ev_, hipEventDisableTiming));
		else HIPCHK(hipEventSynchronize(sst_ev_));  // the staging's previous upload is done
		if (sst_host_bytes_ < bytes) {
			if (sst_host_) HIPCHK(hipHostFree(sst_host_));
			sst_host_ = nullptr;
			sst_host_bytes_ = std::max(bytes * 2, (size_t) 64 << 10);
			HIPCHK(hipHostMalloc((void**) &sst_host_, sst_host_bytes_, hipHostMallocDefault));
		}
		if (d_sst_upd_.bytes() < bytes) {
			// the previous scatter reads the old buffer
			HIPCHK(hipStreamSynchronize(stream_));
			d_sst_upd_.alloc(sst_host_bytes_);
			if (!d_sst_upd_.ptr()) return fail(-2, "device allocation failed (surface)");
		}
		int* hn = (int*) sst_host_;
		unsigned char* hv = sst_host_ + cnt * sizeof(int);
		for (size_t i = 0; i < cnt; i++) {
			hn[i] = sst_pending_[i];
			hv[i] = sstate_[sst_pending_[i]];
		}
		HIPCHK(hipMemcpyAsync(d_sst_upd_.ptr(), sst_host_, bytes, hipMemcpyHostToDevice, stream_));
		HIPCHK(hipEventRecord(sst_ev_, stream_));
		// a normal in flight reads the bytes the scatter changes
		if (normal_pending_) HIPCHK(hipStreamWaitEvent(stream_, norm_done_, 0));
		if (launch_sstate_scatter(stream_, (int) cnt, d_sst_upd_.as<int>(), d_sst_upd_.as<unsigned char>() + cnt * sizeof(int),
		                          d_sstate_.as<unsigned char>()))
			return fail(-2, "surface scatter launch failed");
		sst_pending_.clear();
	}
	if (list_dirty_) {
		const int nb = surface_compact_blocks(mesh_.n);
		if (launch_surface_compact(stream_, mesh_.n, d_sstate_.as<unsigned char>(), d_surf_blk_.as<int>(), d_surf_.as<int>(),
		                           d_surf_blk_.as<int>() + nb))
			return fail(-2, "surface compaction launch failed");
		list_dirty_ = false;
		host_list_stale_ = true;
		eval_stale_ = true;
		choice_valid_ = false;
	}
	return 0;
}

// the device site list copied to surface_vec_
int Grid::host_list()
{
	if (int rc = surface_flush()) return rc;
	if (!host_list_stale_) return 0;
	surface_vec_.resize(surf_count_);
	if (surf_count_ > 0)
		HIPCHK(hipMemcpyAsync(surface_vec_.data(), d_surf_.ptr(), sizeof(int) * (size_t) surf_count_, hipMemcpyDeviceToHost,
		                      stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	host_list_stale_ = false;
	return 0;
}

int Grid::surface_nodes(int* out)
{
	if (sstate_.empty()) return 0;
	if (int rc = host_list()) return rc;
	std::copy(surface_vec_.begin(), surface_vec_.end(), out);
	return 0;
}

int Grid::surface_normal(int node, float out[3])
{
	if (node < 0 || node >= mesh_.n) return fail(-1, "Surface_3d::normal: bad node index");
	if (sstate_.empty()) {
		out[0] = out[1] = out[2] = 0.0f;
		return 0;
	}
	if (!tsg::surface_normal(mesh_, sstate_.data(), node, out)) return fail(-3, "Geometry_3d::find_best_plane(): NaN");
	return 0;
}

int Grid::ensure_side_stream()
{
	if (sstream_) return 0;
	HIPCHK(hipStreamCreateWithFlags(&sstream_, hipStreamNonBlocking));
	HIPCHK(hipEventCreateWithFlags(&sum_src_, hipEventDisableTiming));
	HIPCHK(hipEventCreateWithFlags(&sum_done_, hipEventDisableTiming));
	HIPCHK(hipEventCreateWithFlags(&norm_src_, hipEventDisableTiming));
	HIPCHK(hipEventCreateWithFlags(&norm_done_, hipEventDisableTiming));
	HIPCHK(hipHostMalloc((void**) &sum_host_, sizeof(double), hipHostMallocDefault));
	return 0;
}

// Surface_3d::Table::normal of the event's site on the device, on the side
// stream (it overlaps the take-away and the relaxation); the state bytes it
// reads change only after it (the next flush waits for it), and its record
// is read back with the next surface evaluation
int Grid::launch_normal(int node)
{
	if (int rc = surface_flush()) return rc;
	if (int rc = ensure_side_stream()) return rc;
	if (++normal_stamp_ <= 0) {
		if (launch_fill_i32(stream_, mesh_.n, d_marks_.as<int>(), 0)) return fail(-2, "fill launch failed");
		normal_stamp_ = 1;
	}
	HIPCHK(hipEventRecord(norm_src_, stream_));
	HIPCHK(hipStreamWaitEvent(sstream_, norm_src_, 0));
	NormalArgs a;
	a.node = node;
	a.n = mesh_.n;
	a.ptr = d_ptr_.as<int>();
	a.idx = d_idx_.as<int>();
	a.xyz = d_xyz_.as<float>();
	a.sstate = d_sstate_.as<unsigned char>();
	a.marks = d_marks_.as<int>();
	a.stamp = normal_stamp_;
	char* ws = d_nshell_.as<char>();
	a.ctl = (NormalCtl*) ws;
	a.shell = (int*) (ws + ((sizeof(NormalCtl) + 63) / 64) * 64);
	a.sorted = a.shell + kNormalShellCap;
	a.blk = a.sorted + kNormalShellCap;
	a.rec = d_choice_.as<ChoiceRec>();
	if (launch_surface_normal(sstream_, a)) return fail(-2, "normal launch failed");
	HIPCHK(hipEventRecord(norm_done_, sstream_));
	normal_pending_ = true;
	return 0;
}

int Grid::surface_evaluate(bool exact_sum, double* scaling_ref, int mode, int prev)
{
	if (mode < 0 || mode > 3) return fail(-1, "unknown probability mode");
	if (mode == 1 && !(temperature_ > 0.0f)) return fail(-1, "Surface_3d::Table::probBoltzmann(): temperature must be positive");
	if (prev >= mesh_.n) return fail(-1, "bad node index");
	if (sstate_.empty()) return 0;
	// the previous evaluation's normaliser still reads the surface values
	if (sum_pending_) {
		HIPCHK(hipEventSynchronize(sum_done_));
		sum_pending_ = false;
	}
	if (int rc = surface_flush()) return rc;
	const int cnt = surf_count_;
	if (coord_done_) {  // ids may have left the species table (update_coordination)
		if (int rc = host_list()) return rc;
		for (int v : surface_vec_)
			if (!species_ok(v)) return fail(-8, "Configuration::Table::operator[]()");
	}
	prob_mode_ = mode;
	if (cnt == 0 && prev < 0) {
		choice_valid_ = false;
		eval_stale_ = false;
		return 0;
	}
	const size_t rs = f64_ ? sizeof(double) : sizeof(float);
	if (d_surf_val_.bytes() < rs * (size_t) cnt) {
		// the normaliser (side stream) is done: it was waited for above
		d_surf_val_.alloc(rs * ((size_t) cnt + cnt / 2 + 1024));
		if (!d_surf_val_.ptr()) return fail(-2, "device allocation failed (surface)");
	}
	if (mode != 0 && d_surf_ef_.bytes() < rs * (size_t) cnt) {
		d_surf_ef_.alloc(rs * (size_t) std::max(cnt, 1024) * 2);
		if (!d_surf_ef_.ptr()) return fail(-2, "device allocation failed (surface)");
	}
	if (launch_fill_i32(stream_, 1, d_status_.as<int>(), 0)) return fail(-2, "fill launch failed");
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	FieldArgs f;
	f.count = cnt;
	f.nodes = d_surf_.as<int>();
	f.xyz = d_xyz_.as<float>();
	f.ptr = d_ptr_.as<int>();
	f.idx = d_idx_.as<int>();
	f.coup = d_coupfull_.ptr();
	f.flags = d_flags_.as<unsigned char>();
	f.phi = phi;
	f.evap_field = d_evapF_.as<float>();
	f.value = d_surf_val_.ptr();
	f.status = d_status_.as<int>();
	f.mode = mode;
	f.temperature = temperature_;
	f.evap_energy = d_evapE_.as<float>();
	f.epsilon = d_eps_.as<float>();
	f.value_ef = mode != 0 ? d_surf_ef_.ptr() : nullptr;
	if (launch_surface_fields(stream_, f, f64_)) return fail(-2, "field launch failed");
	// the first maximum, the apex, max |E|/F and the fields of the chosen
	// site and of prev in one record
	ChooseArgs c;
	c.count = cnt;
	c.list = d_surf_.as<int>();
	c.value = d_surf_val_.ptr();
	c.value_ef = f.value_ef;
	c.xyz = f.xyz;
	c.ptr = f.ptr;
	c.idx = f.idx;
	c.coup = f.coup;
	c.flags = f.flags;
	c.phi = phi;
	c.status = d_status_.as<int>();
	c.dev_count = d_surf_blk_.as<int>() + surface_compact_blocks(mesh_.n);
	c.prev = prev;
	c.rec = d_choice_.as<ChoiceRec>();
	// the record shares its memory with the normal of the event
	if (normal_pending_) {
		HIPCHK(hipStreamWaitEvent(stream_, norm_done_, 0));
		normal_pending_ = false;
	}
	if (launch_surface_choose(stream_, c, f64_)) return fail(-2, "choice launch failed");
	if (exact_sum && cnt > 0) {
		// the sequential normaliser (one lane, ~1 ms on a large surface) runs on
		// the side stream; it is read with the event's record and waited for
		// before the surface values change again
		if (int rc = ensure_side_stream()) return rc;
		HIPCHK(hipEventRecord(sum_src_, stream_));
		HIPCHK(hipStreamWaitEvent(sstream_, sum_src_, 0));
		if (launch_surface_sum(sstream_, cnt, d_surf_val_.ptr(), f64_, d_sum_.as<double>()))
			return fail(-2, "sum launch failed");
		HIPCHK(hipMemcpyAsync(sum_host_, d_sum_.ptr(), sizeof(double), hipMemcpyDeviceToHost, sstream_));
		HIPCHK(hipEventRecord(sum_done_, sstream_));
		sum_pending_ = true;
	}
	HIPCHK(hipMemcpyAsync(choice_host_, d_choice_.ptr(), sizeof(ChoiceRec), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	const ChoiceRec& r = *choice_host_;
	if (r.count != cnt) return fail(-2, "surface bookkeeping: the device site list differs from the host's count");
	if (r.status & 17) return fail(-4, "Grid_3d::Table::field_o1(): singular least-squares system");
	if (r.status & 4) return fail(-4, "Grid_3d::Table::field_o2(): singular least-squares system");
	if (r.status & 2) return fail(-3, "Grid_3d::Table::field_o1(): NaN");
	eval_stale_ = false;
	choice_valid_ = cnt > 0;
	if (cnt > 0) {
		scaling_ref_ = r.scaling_ref;
		if (scaling_ref) *scaling_ref = scaling_ref_;
	}
	return 0;
}

int Grid::surface_values(double* out)
{
	// the values belong to the evaluated site list; after a take-away the
	// caller's site count (surface_size) no longer describes them
	if (eval_stale_) return fail(-6, "surface changed since its last evaluation");
	const int cnt = surf_count_;
	if (cnt == 0) return 0;
	HIPCHK(hipStreamSynchronize(stream_));
	if (f64_) {
		HIPCHK(hipMemcpy(out, d_surf_val_.ptr(), sizeof(double) * cnt, hipMemcpyDeviceToHost));
	} else {
		std::vector<float> t(cnt);
		HIPCHK(hipMemcpy(t.data(), d_surf_val_.ptr(), sizeof(float) * cnt, hipMemcpyDeviceToHost));
		for (int i = 0; i < cnt; i++) out[i] = t[i];
	}
	return 0;
}

// probLinear_field normalisation (surface_3d.cpp:276-296): sequential sum in
// node order, then value / sum in the element type
int Grid::surface_probabilities(float* out)
{
	const int cnt = surf_count_;
	std::vector<double> v(cnt);
	if (int rc = surface_values(v.data())) return rc;
	if (f64_) {
		double sum = 0.0;
		for (int i = 0; i < cnt; i++) sum += v[i];
		for (int i = 0; i < cnt; i++) out[i] = (float) (v[i] / sum);
	} else {
		float sum = 0.0f;
		for (int i = 0; i < cnt; i++) sum += (float) v[i];
		for (int i = 0; i < cnt; i++) out[i] = (float) v[i] / sum;
	}
	return 0;
}

int Grid::resolve_candidate(int* node, int* near_tie, int* slot)
{
	if (!choice_valid_) return fail(-6, "surface changed since its last evaluation");
	int w = choice_host_->slot;
	*near_tie = choice_host_->near_tie;
	*node = choice_host_->node;
	if (choice_host_->near_tie) {
		// exact rule: first maximum of the normalised values
		const int cnt = surf_count_;
		std::vector<double> v(cnt);
		if (int rc = surface_values(v.data())) return rc;
		if (f64_) {
			double sum = 0.0;
			for (int i = 0; i < cnt; i++) sum += v[i];
			int pick = 0;
			for (int i = 0; i < cnt; i++)
				if (v[pick] / sum < v[i] / sum) pick = i;
			w = pick;
		} else {
			float sum = 0.0f;
			for (int i = 0; i < cnt; i++) sum += (float) v[i];
			int pick = 0;
			for (int i = 0; i < cnt; i++)
				if ((float) v[pick] / sum < (float) v[i] / sum) pick = i;
			w = pick;
		}
		if (int rc = host_list()) return rc;
		*node = surface_vec_[w];
	}
	if (slot) *slot = w;
	return 0;
}

// evapMonteCarlo (surface_3d.cpp:433-450): threshold rand()/RAND_MAX on the
// running float sum of the normalised probabilities in node order; draws
// from the process's own rand() sequence like the reference
int Grid::mc_candidate(int* node, int* slot)
{
	const int cnt = surf_count_;
	if (cnt == 0) return fail(-6, "empty surface");
	std::vector<float> p(cnt);
	if (int rc = surface_probabilities(p.data())) return rc;
	if (int rc = host_list()) return rc;
	double thr = prand();
	thr /= RAND_MAX;
	int k = 0;
	double acc = p[0];
	while (thr > acc && k + 1 < cnt) {
		k++;
		acc += p[k];
	}
	*node = surface_vec_[k];
	if (slot) *slot = k;
	return 0;
}

int Grid::surface_candidate(int* node, int* near_tie, int* slot)
{
	if (surf_count_ == 0) return fail(-6, "empty surface");
	int nt;
	const int rc = resolve_candidate(node, &nt, slot);
	if (near_tie) *near_tie = nt;
	return rc;
}

// Surface_3d::Table::update (surface_3d.cpp:83-119)
static bool has_vacuum_face(const Mesh& m, const std::vector<int16_t>& ids, const std::vector<float>& a1, int vac, int i)
{
	for (int k = m.ptr[i]; k < m.ptr[i + 1]; k++)
		if (ids[m.idx[k]] == vac && a1[k] > 0.0f) return true;
	return false;
}

// the fork's update_coordination (surface_3d.cpp:568-593): in node order,
// every id >= 10 grows by the number of neighbours whose id (already updated
// for lower indices) is >= 10; the id is a signed char (configuration.h:62)
// and wraps.  Only the ids change, as in the fork (no flags or couplings are
// recomputed); the per-node species data of the device follow them.
// splitmix64 (Steele, Lea and Flood 2014), the mixing function of one
// counter-based draw
static unsigned long long splitmix64(unsigned long long x)
{
	unsigned long long z = x + 0x9e3779b97f4a7c15ull;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

int ion_pick(unsigned int seed, long long event_index, const IonTable& t)
{
	const unsigned long long key = ((unsigned long long) seed << 32) ^ (unsigned long long) event_index;
	const double u = (double) (splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
	double total = 0.0;
	for (int k = 0; k < t.count; k++) total += (double) t.weight[k];
	const double target = u * total;
	double cum = 0.0;
	for (int k = 0; k < t.count; k++) {
		cum += (double) t.weight[k];
		if (target < cum) return k;
	}
	return t.count - 1;
}

int Grid::set_ion_table(int species, int count, const float* mass, const int* charge, const float* weight)
{
	if (species < 0 || species >= 128 || count < 0 || count > kIonTableMax || (count > 0 && (!mass || !charge || !weight)))
		return fail(-1, "ion table: species, count or arrays");
	for (int k = 0; k < count; k++)
		if (!(mass[k] > 0.0f) || !(weight[k] >= 0.0f)) return fail(-1, "ion table: mass or weight");
	IonTable& t = ions_[species];
	t.count = count;
	for (int k = 0; k < count; k++) {
		t.mass[k] = mass[k];
		t.charge[k] = charge[k];
		t.weight[k] = weight[k];
	}
	return 0;
}

int Grid::update_coordination()
{
	const int n = mesh_.n;
	bool changed = false;
	for (int i = 0; i < n; i++) {
		if (ids_[i] < 10) continue;
		int count = 0;
		for (int k = mesh_.ptr[i]; k < mesh_.ptr[i + 1]; k++)
			if (ids_[mesh_.idx[k]] >= 10) count++;
		if (count == 0) continue;
		const bool was_vacuum = ids_[i] == vacuum_id_;
		ids_[i] = (int16_t) (signed char) (ids_[i] + count);
		if (was_vacuum != (ids_[i] == vacuum_id_)) refresh_vacuum(i);
		changed = true;
	}
	if (!changed) return 0;
	std::vector<float> spec(3 * (size_t) n);
	for (int i = 0; i < n; i++) {
		const bool ok = species_ok(i);
		spec[i] = ok ? species_[ids_[i]].evap_field : 0.0f;
		spec[n + (size_t) i] = ok ? species_[ids_[i]].evap_energy : 0.0f;
		spec[2 * (size_t) n + i] = ok ? species_[ids_[i]].epsilon : 0.0f;
	}
	HIPCHK(hipMemcpy(d_evapF_.ptr(), spec.data(), sizeof(float) * n, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_evapE_.ptr(), spec.data() + n, sizeof(float) * n, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_eps_.ptr(), spec.data() + 2 * (size_t) n, sizeof(float) * n, hipMemcpyHostToDevice));
	return 0;
}

int Grid::take_away(int node)
{
	if (vacuum_id_ < 0) return fail(-6, "surface not initialised");
	auto update = [&](int v) -> int {
		if (!(sstate_[v] & kSurfSite)) return fail(-6, "Surface_3d::Table::update(): node not on the surface");
		set_site(v, false);
		for (int k = mesh_.ptr[v]; k < mesh_.ptr[v + 1]; k++) {
			const int j = mesh_.idx[k];
			if (ids_[j] == vacuum_id_) continue;
			if (!species_ok(j)) return fail(-8, "Configuration::Table::operator[]()");
			if (!species_[ids_[j]].removable) continue;
			if (has_vacuum_face(mesh_, ids_, a_one32_, vacuum_id_, j)) set_site(j, true);
		}
		return 0;
	};
	if (int rc = set_id_and_resync(node, vacuum_id_)) return rc;
	if (int rc = update(node)) return rc;
	// detached surface atoms (surface_3d.cpp:521-562)
	std::vector<int> lonely;
	for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++) {
		const int j = mesh_.idx[k];
		if (ids_[j] == vacuum_id_) continue;
		bool flag = true;
		for (int kk = mesh_.ptr[j]; kk < mesh_.ptr[j + 1]; kk++)
			if (ids_[mesh_.idx[kk]] != vacuum_id_) { flag = false; break; }
		if (flag) lonely.push_back(j);
	}
	for (int j : lonely) {
		if (int rc = set_id_and_resync(j, vacuum_id_)) return rc;
		if (int rc = update(j)) return rc;
	}
	// the state bytes and the site list follow on the device at the next
	// flush (the relaxation does not read them)
	return 0;
}

// evap_initialVelocity (surface_3d.cpp:472-502): kinetic energy kT split
// over the axes by two uniform draws, a random sign per axis
void Grid::thermal_velocity(const Species& sp, float v[3]) const
{
	double max = 2.0;
	max *= 1.3806504e-23;
	max *= temperature_;
	max /= sp.mass;
	max /= 1.660538782e-27;
	v[0] = (float) prand();
	v[0] /= RAND_MAX;
	v[0] = (float) (v[0] * max);
	v[1] = (float) prand();
	v[1] /= RAND_MAX;
	v[1] = (float) (v[1] * (max - v[0]));
	v[2] = (float) max;
	v[2] -= v[0];
	v[2] -= v[1];
	for (int i = 0; i < 3; i++) {
		v[i] = std::sqrt(v[i]);
		if (prand() % 2) v[i] *= -1.0f;
	}
}

// Steps V-VI of one event on the flight stream: the field snapshot and a
// copy of the potentials go into the event's slot on the main stream, then
// the flight kernel (thermal start velocity, trajectory, exit potential on
// the rand() stream) waits for them on the flight stream.
int Grid::flight_start(const EvapOptions& o, int ev, int node, const Species& sp, float flen, float pot_before,
                       float pot_after, float scaling)
{
	if (int rc = ensure_point_location()) return rc;
	const int n = mesh_.n;
	const size_t rs = f64_ ? sizeof(double) : sizeof(float);
	if (!fstream_) HIPCHK(hipStreamCreateWithFlags(&fstream_, hipStreamNonBlocking));
	FlightSlot& f = fslot_[flight_seq_ % 2];
	if (f.active)
		if (int rc = flight_finish(f, nullptr, o)) return rc;
	if (!f.host) {
		HIPCHK(hipHostMalloc((void**) &f.host, 32 * sizeof(float)));
		HIPCHK(hipEventCreateWithFlags(&f.ready, hipEventDisableTiming));
		HIPCHK(hipEventCreateWithFlags(&f.done, hipEventDisableTiming));
		f.nfield.alloc(sizeof(float) * 3 * (size_t) n);
		f.phi.alloc(rs * (size_t) n);
		f.io.alloc(32 * sizeof(float));
		if (!f.nfield.ptr() || !f.phi.ptr() || !f.io.ptr()) return fail(-2, "device allocation failed (flight slot)");
	}
	int* io = f.io.as<int>();
	// main stream: the event's field snapshot and potentials
	HIPCHK(hipMemsetAsync(io + 23, 0, sizeof(int), stream_));
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	if (launch_field_snapshot(stream_, n, d_xyz_.as<float>(), d_ptr_.as<int>(), d_idx_.as<int>(), d_coupfull_.ptr(),
	                          d_flags_.as<unsigned char>(), phi, f64_, f.nfield.as<float>(), io + 23))
		return fail(-2, "field snapshot launch failed");
	HIPCHK(hipMemcpyAsync(f.phi.ptr(), phi, rs * (size_t) n, hipMemcpyDeviceToDevice, stream_));
	HIPCHK(hipEventRecord(f.ready, stream_));
	// flight stream
	if (!rng_on_device_) {
		HIPCHK(hipMemcpyAsync(d_rng_.ptr(), &tsg::process_rand(), sizeof(RandState), hipMemcpyHostToDevice, fstream_));
		HIPCHK(hipStreamSynchronize(fstream_));
		rng_on_device_ = true;
	}
	// the ion: the species' own mass and charge (the reference), or an entry
	// of its ion table drawn for this event
	float ion_m = sp.mass;
	int ion_q = sp.evap_charge;
	const EventRecord& rec = flight_out_[ev];
	const IonTable& it = ions_[rec.id];
	if (o.ion_table && it.count > 0) {
		const int k = ion_pick(o.ion_seed, rec.event_index, it);
		ion_m = it.mass[k];
		ion_q = it.charge[k];
	}
	flight_out_[ev].ion_mass = ion_m;
	flight_out_[ev].ion_charge = ion_q;
	float charge = (float) ion_q;
	charge *= 1.60217733e-19f;
	float mass = ion_m;
	mass *= 1.660538782e-27f;
	int* hi = (int*) f.host;
	hi[0] = node;
	hi[1] = 0;
	f.host[2] = charge;
	f.host[3] = mass;
	HIPCHK(hipMemcpyAsync(io, f.host, 4 * sizeof(float), hipMemcpyHostToDevice, fstream_));
	HIPCHK(hipStreamWaitEvent(fstream_, f.ready, 0));
	FlightArgs a;
	std::memset(&a, 0, sizeof(a));
	TrajArgs& t = a.t;
	t.count = 1;
	t.nodes = io;
	t.snapshot = io + 1;
	t.charge = (const float*) (io + 2);
	t.mass = (const float*) (io + 3);
	t.integrator = o.integrator;
	t.stepper = o.stepper;
	t.dt0 = o.dt0;
	t.dt_min = o.dt_min;
	t.err_limit = o.err_limit;
	t.zmax = o.detector_z > 0.0f ? o.detector_z : zmax_;
	t.max_steps = 1 << 24;
	t.max_calls = 1ll << 22;
	t.n = n;
	t.m = mesh_.m();
	t.xyz = d_xyz_.as<float>();
	t.tv = d_tv_.as<int>();
	t.tn = d_tn_.as<int>();
	t.assoc = d_assoc_.as<int>();
	t.gbnd = d_gbnd_.as<unsigned char>();
	t.nfield = f.nfield.as<float>();
	t.cap = 0;
	t.len = io + 4;
	t.status = io + 5;
	t.summary = (float*) (io + 6);
	t.rng = d_rng_.as<int>();
	t.err_out = (float*) (io + 24);
	if (rec_cap_ > 0) {
		// the phases of the flight, for the caller's trajectory record
		if (f.phases_cap < rec_cap_) {
			if (f.phases_host) (void) hipHostFree(f.phases_host);
			f.phases_host = nullptr;
			f.phases_cap = 0;
			f.phases.alloc(sizeof(float) * 8 * (size_t) rec_cap_);
			HIPCHK(hipHostMalloc((void**) &f.phases_host, sizeof(float) * 8 * (size_t) rec_cap_));
			if (!f.phases.ptr()) return fail(-2, "device allocation failed (flight phases)");
			f.phases_cap = rec_cap_;
		}
		t.out = f.phases.as<float>();
		t.cap = rec_cap_;
	}
	a.thermal = o.thermal_velocity;
	a.temperature = temperature_;
	a.mass_amu = sp.mass;
	a.v0 = (float*) (io + 20);
	// the reference locates the exit point twice (voltage, time scale), and
	// only when the field before was not zero
	a.exit_walks = flen != 0.0f ? 2 : 0;
	a.phi = f.phi.ptr();
	a.f64 = f64_;
	a.exit = (float*) (io + 18);
	if (launch_flight(fstream_, a)) return fail(-2, "flight launch failed");
	HIPCHK(hipMemcpyAsync(f.host + 4, io + 4, 26 * sizeof(float), hipMemcpyDeviceToHost, fstream_));
	if (rec_cap_ > 0)
		HIPCHK(hipMemcpyAsync(f.phases_host, f.phases.ptr(), sizeof(float) * 8 * (size_t) rec_cap_, hipMemcpyDeviceToHost,
		                      fstream_));
	HIPCHK(hipEventRecord(f.done, fstream_));
	f.active = true;
	f.seq = flight_seq_++;
	f.ev = ev;
	f.flen = flen;
	f.pot_before = pot_before;
	f.pot_after = pot_after;
	f.evap_field = sp.evap_field;
	f.scaling = scaling;
	return 0;
}

// collects one flight: the record's flight fields, and the voltage running
// mean and time scale of process.cpp:983-1037 in event order
int Grid::flight_finish(FlightSlot& f, EventRecord* out, const EvapOptions& o)
{
	HIPCHK(hipEventSynchronize(f.done));
	f.active = false;
	EventRecord* recs = out ? out : flight_out_;
	EventRecord& r = recs[f.ev];
	const int* hi = (const int*) f.host;
	if (hi[23]) return fail(-4, "Grid_3d::Table::field_o1(): singular least-squares system");
	r.traj_len = hi[4];
	r.traj_status = hi[5];
	for (int c = 0; c < 3; c++) {
		r.err_pos[c] = f.host[24 + c];
		r.err_vel[c] = f.host[27 + c];
	}
	if (rec_phases_ && rec_cap_ > 0) {
		const int keep = std::min(r.traj_len, rec_cap_);
		std::memcpy(rec_phases_ + 8 * (size_t) rec_cap_ * (size_t) f.ev, f.phases_host, sizeof(float) * 8 * (size_t) keep);
		rec_len_[f.ev] = r.traj_len;
	}
	const float* sm = f.host + 6;
	for (int c = 0; c < 3; c++) {
		r.start[c] = sm[c];
		r.stop[c] = sm[3 + c];
	}
	const float t_last = sm[6];
	const bool have_back = sm[11] != 0.0f;
	if (have_back && f.flen != 0.0f) {
		const float pexit = f.host[18];
		float v = f.pot_before;
		v -= pexit;
		v *= f.evap_field;
		v /= f.flen;
		v /= f.scaling;
		vq_.push_back(v);
		vq_sum_ += v;
		while (vq_.size() - vq_head_ > (size_t) o.voltage_queue) vq_sum_ -= vq_[vq_head_++];
		if (vq_head_ > 4096) {
			vq_.erase(vq_.begin(), vq_.begin() + (long) vq_head_);
			vq_head_ = 0;
		}
		r.voltage = vq_sum_;
		r.voltage /= (float) (vq_.size() - vq_head_);
		float ts = f.pot_after;
		ts -= pexit;
		ts /= r.voltage;
		r.time_scale = std::sqrt(ts);
	} else if (f.flen == 0.0f) {
		r.voltage = 1.0f;
		r.time_scale = 1.0f;
	} else {
		r.voltage = 0.0f;
		r.time_scale = 1.0f;
	}
	r.tof = t_last * r.time_scale;
	return 0;
}

// collects every flight in the air, oldest first, and hands the rand()
// stream back to the host
int Grid::flight_drain(EventRecord* out, const EvapOptions& o)
{
	int rc = 0;
	for (;;) {
		FlightSlot* oldest = nullptr;
		for (FlightSlot& f : fslot_)
			if (f.active && (!oldest || f.seq < oldest->seq)) oldest = &f;
		if (!oldest) break;
		const int r = flight_finish(*oldest, out, o);
		if (r && !rc) rc = r;
	}
	if (rng_on_device_) {
		HIPCHK(hipMemcpyAsync(&tsg::process_rand(), d_rng_.ptr(), sizeof(RandState), hipMemcpyDeviceToHost, fstream_));
		HIPCHK(hipStreamSynchronize(fstream_));
		rng_on_device_ = false;
	}
	return rc;
}

int Grid::evaporate(const EvapOptions& o, int k, EventRecord* out, int* done, int phase_cap, float* phases,
                    int* phase_len)
{
	if (phase_cap < 0 || (phase_cap > 0 && (!phases || !phase_len))) return fail(-1, "evaporate: trajectory record");
	flight_out_ = out;
	rec_cap_ = phase_cap;
	rec_phases_ = phases;
	rec_len_ = phase_len;
	if (phase_len)
		for (int e = 0; e < k; e++) phase_len[e] = 0;
	const int rc = evaporate_events(o, k, out, done);
	// collect the flights still in the air (their records are in out)
	const int rd = flight_drain(out, o);
	flight_out_ = nullptr;
	rec_cap_ = 0;
	rec_phases_ = nullptr;
	rec_len_ = nullptr;
	return rc ? rc : rd;
}

int Grid::evaporate_events(const EvapOptions& o, int k, EventRecord* out, int* done)
{
	*done = 0;
	if (o.prob_mode < 0 || o.prob_mode > 3 || o.evap_mode < 0 || o.evap_mode > 1)
		return fail(-1, "unknown probability model or choice");
	if (o.method < 0 || o.method > 2) return fail(-1, "unknown relaxation method");
	if (o.method == 2 && !(net_ && net_predict_))
		return fail(-1, "method 2 needs a warm-start network attached with prediction on");
	if (o.accel_solver < 0 || o.accel_solver > 2) return fail(-1, "unknown correction solver");
	if (o.method >= 1 && !accel_)
		if (int rc = accel_setup(0.9)) return rc;
	amg_select(o.accel_solver >= 1, o.accel_solver == 2);
	if (vacuum_id_ != o.vacuum_id || surf_count_ == 0)
		if (int rc = surface_init(o.vacuum_id)) return rc;
	if (int rc = surface_evaluate(o.exact_probability != 0, nullptr, o.prob_mode)) return rc;
	if (o.trajectory && (o.voltage_queue < 0 || o.voltage_queue > 1 << 20)) return fail(-1, "bad voltage queue length");
	if (!apex0_set_ && surf_count_ > 0) {
		apex0_ = choice_host_->apex[2];
		apex0_set_ = true;
	}
	// the fork applies update_coordination once after the initial surface
	// evaluation (process.cpp:645-647) and after every take-away (:770)
	if (o.update_coordination && !coord_done_) {
		if (int rc = update_coordination()) return rc;
		coord_done_ = true;
	}
	using clk = std::chrono::steady_clock;
	auto t_last = clk::now();
	// host wall time since the previous mark into phase p (the phases end
	// with a stream synchronisation or a blocking readback, except where noted)
	auto mark = [&](int p) {
		const auto t = clk::now();
		stat_phase_ms_[p] += std::chrono::duration<double, std::milli>(t - t_last).count();
		t_last = t;
	};
	if (o.init_event_cnt < 0) return fail(-1, "negative initial event count");
	if (event_cnt_ < 0) event_cnt_ = o.init_event_cnt;
	for (int ev = 0; ev < k; ev++) {
		if (surf_count_ == 0) break;
		EventRecord& r = out[ev];
		std::memset(&r, 0, sizeof(r));
		r.event_index = ++event_cnt_;
		stat_phase_events_++;
		const float scaling_reference = (float) scaling_ref_;
		// the last evaluation's record: the apex, the first maximum and its
		// potential and field (read back with that evaluation)
		const ChoiceRec rec = *choice_host_;
		for (int c = 0; c < 3; c++) r.apex[c] = rec.apex[c];
		int node, nt = 0, slot = 0;
		if (o.evap_mode == 0) {
			if (int rc = surface_candidate(&node, &nt, &slot)) return rc;
		} else {
			// the choice draws from the rand() stream the flights advance
			if (int rc = flight_drain(out, o)) return rc;
			if (int rc = mc_candidate(&node, &slot)) return rc;
		}
		r.node = node;
		r.id = ids_[node];
		r.near_tie = nt;
		// the normal on the device, on the surface the site was chosen from
		if (int rc = launch_normal(node)) return rc;
		// the chosen site's value now; its probability (value / normaliser)
		// once the normaliser is in, before the surface is evaluated again
		const bool in_rec = node == rec.node;
		double vw = 0.0;
		float v32 = 0.0f;
		if (o.exact_probability) {
			if (in_rec) {
				vw = rec.value;
				v32 = (float) rec.value;
			} else {
				const size_t rs = f64_ ? sizeof(double) : sizeof(float);
				if (f64_)
					HIPCHK(hipMemcpyAsync(&vw, d_surf_val_.as<char>() + rs * slot, sizeof(double), hipMemcpyDeviceToHost, stream_));
				else
					HIPCHK(hipMemcpyAsync(&v32, d_surf_val_.as<char>() + rs * slot, sizeof(float), hipMemcpyDeviceToHost, stream_));
				HIPCHK(hipStreamSynchronize(stream_));
			}
		}
		if (in_rec) {
			r.pot_before = rec.pot;
			for (int c = 0; c < 3; c++) r.field_before[c] = rec.field[c];
		} else {
			if (int rc = potential(node, &r.pot_before)) return rc;
			if (int rc = field_o1(1, &node, r.field_before)) return rc;
		}
		mark(kPhaseChoice);
		if (int rc = take_away(node)) return rc;
		if (o.update_coordination)
			if (int rc = update_coordination()) return rc;
		mark(kPhaseTakeAway);
		if (o.method >= 1) {
			const bool warm = net_ && (net_record_ || o.method == 2);
			warm_.n = 0;
			if (warm)
				if (int rc = warm_begin(node, r.pot_before, o.method == 2, &r)) return rc;
			mark(kPhaseWarm);
			int outer;
			double fr;
			if (int rc = accel_relax(o.accel_tol, &outer, &r.global_iter, &fr)) return rc;
			r.local_iter = outer;
			mark(kPhaseRelax);
			if (warm)
				if (int rc = warm_end()) return rc;
			mark(kPhaseWarm);
		} else {
			if (int rc = local_relax(node, o.relax_shell, o.relax_threshold, o.relax_cycle, o.relax_queue, &r.local_iter)) return rc;
			for (int c = 0; c < o.relax_global_cycles; c++)
				if (int rc = relax_cycles(1, nullptr)) return rc;
			// the reference adds initEventCnt to the already offset eventIndex
			// (process.cpp:876); the sum is an unsigned int there
			if (o.refresh_interval > 0 &&
			    (unsigned int) ((unsigned int) o.init_event_cnt + (unsigned int) r.event_index) %
			            (unsigned int) o.refresh_interval == 0) {
				double fd, fs, ld;
				if (int rc = relax_threshold(o.refresh_threshold, o.refresh_cycle, o.refresh_queue, &r.global_iter, &fd, &fs, &ld)) return rc;
			}
		}
		mark(kPhaseRelax);
		if (o.exact_probability) {
			if (!sum_pending_) return fail(-2, "surface normaliser missing");
			HIPCHK(hipEventSynchronize(sum_done_));
			const double sum = *sum_host_;
			if (f64_)
				r.probability = (float) (vw / sum);
			else
				r.probability = v32 / (float) sum;
		}
		// the new surface's evaluation and choice, with the potential and
		// field of the evaporated node after the event and the event's normal
		if (int rc = surface_evaluate(o.exact_probability != 0, &r.scaling_ref, o.prob_mode, node)) return rc;
		r.surface_size = surf_count_;
		{
			const ChoiceRec& nr = *choice_host_;
			if (nr.normal_status == 1) return fail(-3, "Geometry_3d::find_best_plane(): NaN");
			if (nr.normal_status == 2) return fail(-2, "surface normal: three-hop shell beyond the device buffer");
			for (int c = 0; c < 3; c++) r.normal[c] = nr.normal[c];
			r.pot_after = nr.pot_prev;
			for (int c = 0; c < 3; c++) r.field_after[c] = nr.field_prev[c];
		}
		// V) the ion's flight in the relaxed field (process.cpp:940-975) and
		// VI) voltage and time scale (process.cpp:983-1037)
		if (o.trajectory && !(r.id >= 0 && r.id < 128 && species_[r.id].defined))
			return fail(-8, "Configuration::Table::operator[]()");
		mark(kPhaseSurface);
		float flen = 0.0f;
		for (int c = 0; c < 3; c++) flen += (float) r.field_before[c] * (float) r.field_before[c];
		flen = std::sqrt(flen);
		if (o.trajectory) {
			// on the flight stream, overlapping the next event; the flight's
			// part of the record is filled when it is collected, in event order
			if (int rc = flight_start(o, ev, node, species_[r.id], flen, (float) r.pot_before, (float) r.pot_after,
			                          scaling_reference))
				return rc;
		} else if (flen == 0.0f) {
			r.voltage = 1.0f;
			r.time_scale = 1.0f;
		} else {
			r.voltage = 0.0f;
			r.time_scale = 1.0f;
		}
		mark(kPhaseFlight);
		*done = ev + 1;
		if (apex0_ - r.apex[2] >= o.shrink_limit) break;
	}
	return 0;
}

}  // namespace tsg

namespace tsg {

// algorithmic bytes of one FP64 (or float) Jacobi sweep over the active rows,
// B = (4 + R) nnz + (4 + 2R) rows + R referenced (R = 8 for FP64: 12 nnz +
// 20 rows + 8 referenced): per entry the column index and the coupling, per
// row its offset, the coupling sum (or charge term) and the written
// potential, every referenced potential read once
int Grid::sweep_stats(double* bytes, long long* rows, long long* entries, long long* referenced, double* ms,
                      long long* sweeps, long long* launched) const
{
	const int rsz = f64_ ? 8 : 4;
	long long nr = 0, ne = 0;
	std::vector<uint8_t> ref(mesh_.n, 0);
	for (int s = 0; s < next_spare_; s++) {
		const int node = slot_node_[s];
		if (node < 0) continue;
		nr++;
		ref[node] = 1;
		for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++)
			if (participates(node, mesh_.idx[k])) { ne++; ref[mesh_.idx[k]] = 1; }
	}
	long long nref = 0;
	for (uint8_t v : ref) nref += v;
	if (bytes) *bytes = (double) ne * (4 + rsz) + (double) nr * (4 + 2 * rsz) + (double) nref * rsz;
	if (rows) *rows = nr;
	if (entries) *entries = ne;
	if (referenced) *referenced = nref;
	if (ms) *ms = stat_sweep_ms_;
	if (sweeps) *sweeps = stat_sweeps_;
	if (launched) *launched = stat_launched_;
	return 0;
}

}  // namespace tsg

namespace tsg {

AccelArgs Grid::accel_args() const
{
	AccelArgs x;
	x.nslots = next_spare_;
	x.slot_node = d_slot_node_.as<int>();
	x.inner_deg = d_inner_deg_.as<int>();
	x.slice_off = d_slice_off_.as<long long>();
	x.col = d_icol_.as<int>();
	x.w = d_iw_.as<float>();
	x.slot_deg = d_slot_deg_.as<int>();
	x.long_slots = d_long_.as<int>();
	x.nlong = (int) long_set_.size();
	x.long_off = d_long_off_.as<long long>();
	x.licol = d_licol_.as<int>();
	x.liw = d_liw_.as<float>();
	return x;
}

// the FP32 correction row of a slot: unknown participating neighbours (by
// their slots: the correction vectors are slot-indexed) with weight
// c_ij / (row coupling sum); int patches target arrays 3 (inner_deg) and 4
// (inner columns), float patches the weights
int Grid::write_inner(int slot, int node, std::vector<long long>* ai, std::vector<int>* vi, std::vector<long long>* af,
                      std::vector<float>* vf)
{
	std::vector<int> cols;
	std::vector<double> coups;
	double ws;
	row_entries(node, &cols, &coups, &ws);
	const long long base = slice_off_[slot / kSlice] + (slot % kSlice);
	int deg = 0;
	for (size_t k = 0; k < cols.size(); k++) {
		if (flags_[cols[k]] & 2) continue;
		const int js = slot_of_[cols[k]];
		if (js < 0) return -2;  // an unknown neighbour without a row
		const long long e = base + (long long) kSlice * deg;
		ai->push_back((4ll << 60) | e); vi->push_back(js);
		af->push_back(e); vf->push_back((float) (coups[k] / ws));
		deg++;
	}
	if (deg > slice_width_[slot / kSlice]) return -1;
	inner_deg_[slot] = deg;
	ai->push_back((3ll << 60) | slot); vi->push_back(deg);
	if (slot_deg_[slot] > kLongRow) {  // the long row's contiguous copy (room made by write_slot)
		if (loff_[slot] < 0 || lcap_[slot] < deg) return -1;
		int k = 0;
		for (size_t t = 0; t < cols.size(); t++) {
			if (flags_[cols[t]] & 2) continue;
			lpi_addr_.push_back((1ll << 60) | (loff_[slot] + k)); lpi_val_.push_back(slot_of_[cols[t]]);
			lpf_addr_.push_back(loff_[slot] + k); lpf_val_.push_back((float) (coups[t] / ws));
			k++;
		}
	}
	return 0;
}

int Grid::accel_setup(double safety)
{
	if (!f64_) return fail(-1, "the accelerated relaxation runs on a TSG_PLAIN_F64 grid");
	const int n = mesh_.n;
	const long long nent = slice_off_.back();
	inner_deg_.assign(cap_slots_, 0);
	std::vector<int> icol(nent, 0);
	std::vector<float> iw(nent, 0.0f);
	lpi_addr_.clear(); lpi_val_.clear(); lpf_addr_.clear(); lpf_val_.clear();
	for (int slot = 0; slot < next_spare_; slot++) {
		const int node = slot_node_[slot];
		if (node < 0) continue;
		std::vector<long long> ai, af;
		std::vector<int> vi;
		std::vector<float> vf;
		if (int rc = write_inner(slot, node, &ai, &vi, &af, &vf))
			return fail(-7, rc == -2 ? "inner row: an unknown neighbour without a row" : "inner row too wide");
		for (size_t k = 0; k < af.size(); k++) iw[af[k]] = vf[k];
		for (size_t k = 0; k < ai.size(); k++)
			if ((ai[k] >> 60) == 4) icol[ai[k] & ((1ll << 60) - 1)] = vi[k];
	}
	const int nv = cap_slots_;  // correction vectors: one entry per row slot
	d_inner_deg_.alloc(sizeof(int) * cap_slots_);
	d_icol_.alloc(sizeof(int) * nent);
	d_iw_.alloc(sizeof(float) * nent);
	d_d0_.alloc(sizeof(float) * nv);
	d_d1_.alloc(sizeof(float) * nv);
	d_r32_.alloc(sizeof(float) * nv);
	d_redout_.alloc(sizeof(double) * 2);
	if (!d_icol_.ptr() || !d_iw_.ptr() || !d_d0_.ptr()) return fail(-2, "device allocation failed (accelerated rows)");
	HIPCHK(hipMemcpy(d_inner_deg_.ptr(), inner_deg_.data(), sizeof(int) * cap_slots_, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_icol_.ptr(), icol.data(), sizeof(int) * nent, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_iw_.ptr(), iw.data(), sizeof(float) * nent, hipMemcpyHostToDevice));
	{
		// the long rows' correction entries, contiguous
		std::vector<int> lic(lstore_cap_, 0);
		std::vector<float> liw(lstore_cap_, 0.0f);
		for (size_t t = 0; t < lpi_addr_.size(); t++)
			if ((lpi_addr_[t] >> 60) == 1) lic[lpi_addr_[t] & ((1ll << 60) - 1)] = lpi_val_[t];
		for (size_t t = 0; t < lpf_addr_.size(); t++) liw[lpf_addr_[t]] = lpf_val_[t];
		d_licol_.alloc(sizeof(int) * (size_t) lstore_cap_);
		d_liw_.alloc(sizeof(float) * (size_t) lstore_cap_);
		if (!d_licol_.ptr() || !d_liw_.ptr()) return fail(-2, "device allocation failed (long correction rows)");
		HIPCHK(hipMemcpy(d_licol_.ptr(), lic.data(), sizeof(int) * lic.size(), hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(d_liw_.ptr(), liw.data(), sizeof(float) * liw.size(), hipMemcpyHostToDevice));
		lpi_addr_.clear(); lpi_val_.clear(); lpf_addr_.clear(); lpf_val_.clear();
	}
	HIPCHK(hipMemset(d_d0_.ptr(), 0, sizeof(float) * nv));
	HIPCHK(hipMemset(d_d1_.ptr(), 0, sizeof(float) * nv));
	HIPCHK(hipMemset(d_r32_.ptr(), 0, sizeof(float) * nv));
	accel_ = true;

	// spectral radius of G by power iteration with a D-weighted Rayleigh
	// quotient (G = D^-1 C is self-adjoint in that inner product)
	std::vector<float> x(nv, 0.0f), y(nv, 0.0f);
	std::vector<double> d(nv, 0.0);
	for (int slot = 0; slot < next_spare_; slot++) {
		const int node = slot_node_[slot];
		if (node < 0) continue;
		x[slot] = 1.0f;
		double ws = 0.0;
		for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++)
			if (participates(node, mesh_.idx[k])) ws += coup_[k];
		d[slot] = ws;
	}
	AccelArgs xa = accel_args();
	double lambda = 0.0;
	for (int chunk = 0; chunk < 8; chunk++) {
		HIPCHK(hipMemcpy(d_d0_.ptr(), x.data(), sizeof(float) * nv, hipMemcpyHostToDevice));
		float* in = d_d0_.as<float>();
		float* out = d_d1_.as<float>();
		for (int it = 0; it < 150; it++) {
			if (launch_apply_g(stream_, xa, in, out)) return fail(-2, "power iteration launch failed");
			std::swap(in, out);
		}
		if (launch_apply_g(stream_, xa, in, out)) return fail(-2, "power iteration launch failed");
		HIPCHK(hipStreamSynchronize(stream_));
		HIPCHK(hipMemcpy(x.data(), in, sizeof(float) * nv, hipMemcpyDeviceToHost));
		HIPCHK(hipMemcpy(y.data(), out, sizeof(float) * nv, hipMemcpyDeviceToHost));
		double num = 0.0, den = 0.0, mx = 0.0;
		for (int i = 0; i < nv; i++) {
			num += d[i] * (double) x[i] * (double) y[i];
			den += d[i] * (double) x[i] * (double) x[i];
			mx = std::max(mx, (double) std::fabs(y[i]));
		}
		if (den <= 0.0 || mx <= 0.0) return fail(-4, "power iteration lost its vector");
		lambda = num / den;
		for (int i = 0; i < nv; i++) x[i] = (float) (y[i] / mx);
	}
	rho_ = 1.0 - (1.0 - lambda) * safety;
	if (!(rho_ > 0.0 && rho_ < 1.0)) return fail(-4, "spectral radius estimate out of (0,1)");
	HIPCHK(hipMemset(d_d0_.ptr(), 0, sizeof(float) * nv));
	HIPCHK(hipMemset(d_d1_.ptr(), 0, sizeof(float) * nv));
	return 0;
}

int Grid::accel_relax(double tol, int* outer, long long* inner, double* final_residual)
{
	if (!accel_)
		if (int rc = accel_setup(0.9)) return rc;
	if (amg_want_ && !amg_on_)
		if (int rc = amg_setup()) return rc;
	if (!amg_want_) amg_on_ = false;
	const int n = mesh_.n;
	double* phi = (sw_ ? d_phi1_ : d_phi0_).as<double>();
	double* other = (sw_ ? d_phi0_ : d_phi1_).as<double>();
	SweepArgs a = sweep_args();
	AccelArgs x = accel_args();
	double sigma = rho_ / (1.0 + std::sqrt(1.0 - rho_ * rho_));
	long long steps = 0;
	int o = 0;
	double rmax = 0.0, rprev = 0.0;
	if (!ev_start_) {
		HIPCHK(hipEventCreate(&ev_start_));
		HIPCHK(hipEventCreate(&ev_stop_));
	}
	HIPCHK(hipEventRecord(ev_start_, stream_));
	for (;; o++) {
		double red[2];
		if (o == 0 && r32_ready_) {  // formed by the warm start's check on these potentials
			red[0] = r32_rmax_;
			red[1] = 0.0;
		} else {
			if (launch_resid64(stream_, a, phi, d_r32_.as<float>(), d_redout_.as<double>()))
				return fail(-2, "residual launch failed");
			HIPCHK(hipMemcpyAsync(red, d_redout_.ptr(), sizeof(red), hipMemcpyDeviceToHost, stream_));
			HIPCHK(hipStreamSynchronize(stream_));
			stat_resid_++;
		}
		r32_ready_ = false;
		rmax = red[0];
		if (red[1] != 0.0) return fail(-3, "accelerated relaxation: NaN residual");
		if (rmax <= tol) break;
		if (o >= 64) return fail(-5, "accelerated relaxation did not converge in 64 refinement steps");
		// a refinement that did not halve the residual means the spectral
		// bound is too low for the current operator: move it halfway to 1
		if (o > 0 && rmax > 0.5 * rprev && !amg_on_) {
			rho_ = 1.0 - 0.5 * (1.0 - rho_);
			sigma = rho_ / (1.0 + std::sqrt(1.0 - rho_ * rho_));
			stat_rho_raised_++;
		}
		rprev = rmax;
		// the correction solve for a reduction by f (the FP32 correction carries ~6 digits)
		const double f = std::max(1e-5, 0.5 * tol / rmax);
		if (amg_on_) {
			int it = 0;
			float* dl = nullptr;
			if (int rc = amg_solve(x, (float) f, &it, &dl)) return rc;
			steps += it;
			if (launch_apply64(stream_, x, dl, phi, nullptr)) return fail(-2, "update launch failed");
			continue;
		}
		int m = (int) std::ceil(std::log(2.0 / f) / std::log(1.0 / sigma));
		m = std::max(2, std::min(m, 200000));
		float* dprev = d_d0_.as<float>();
		float* dcur = d_d1_.as<float>();
		if (launch_cheb_init(stream_, x, d_r32_.as<float>(), dprev, dcur)) return fail(-2, "chebyshev launch failed");
		double omega = 1.0;
		for (int k = 2; k <= m; k++) {
			omega = (k == 2) ? 1.0 / (1.0 - 0.5 * rho_ * rho_) : 1.0 / (1.0 - 0.25 * rho_ * rho_ * omega);
			if (launch_cheb_step(stream_, x, d_r32_.as<float>(), dcur, dprev, (float) omega))
				return fail(-2, "chebyshev launch failed");
			std::swap(dprev, dcur);
		}
		steps += m;
		if (launch_apply64(stream_, x, dcur, phi, nullptr)) return fail(-2, "update launch failed");
	}
	// both buffers carry the converged potentials, so that potential()
	// (the buffer not written last) reads them
	HIPCHK(hipMemcpyAsync(other, phi, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
	HIPCHK(hipEventRecord(ev_stop_, stream_));
	HIPCHK(hipEventSynchronize(ev_stop_));
	float ms = 0.0f;
	HIPCHK(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
	stat_accel_ms_ += ms;
	stat_cheb_ += steps;
	if (outer) *outer = o;
	if (inner) *inner = steps;
	if (final_residual) *final_residual = rmax;
	return 0;
}

}  // namespace tsg

namespace tsg {

// average device time of one launch of a relaxation kernel on the grid's
// stream (HIP events around reps back-to-back launches): kind 0 = one FP64
// (or float) Jacobi cycle with its decision kernel, kind 1 = one FP32
// Chebyshev correction step.  The potentials are left as they were.
int Grid::time_kernel(int kind, int reps, double* avg_ms, double* bytes)
{
	if (reps <= 0) return fail(-1, "time_kernel: reps");
	if (!ev_start_) {
		HIPCHK(hipEventCreate(&ev_start_));
		HIPCHK(hipEventCreate(&ev_stop_));
	}
	const int n = mesh_.n;
	const size_t rs = f64_ ? sizeof(double) : sizeof(float);
	DevBuf save0, save1;
	save0.alloc(rs * n);
	save1.alloc(rs * n);
	HIPCHK(hipMemcpyAsync(save0.ptr(), d_phi0_.ptr(), rs * n, hipMemcpyDeviceToDevice, stream_));
	HIPCHK(hipMemcpyAsync(save1.ptr(), d_phi1_.ptr(), rs * n, hipMemcpyDeviceToDevice, stream_));
	float ms = 0.0f;
	if (kind == 0) {
		if (int rc = arm_ctl(-1.0, 1, 0)) return rc;
		SweepArgs a = sweep_args();
		if (launch_jacobi_sweep(stream_, a, 0, f64_)) return fail(-2, "sweep launch failed");
		HIPCHK(hipEventRecord(ev_start_, stream_));
		for (int r = 0; r < reps; r++)
			if (launch_jacobi_sweep(stream_, a, (r + 1) & 1, f64_)) return fail(-2, "sweep launch failed");
		HIPCHK(hipEventRecord(ev_stop_, stream_));
		HIPCHK(hipEventSynchronize(ev_stop_));
		HIPCHK(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
		if (bytes) sweep_stats(bytes, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
	} else if (kind == 1) {
		if (!accel_)
			if (int rc = accel_setup(0.9)) return rc;
		AccelArgs x = accel_args();
		float* d0 = d_d0_.as<float>();
		float* d1 = d_d1_.as<float>();
		if (launch_cheb_step(stream_, x, d_r32_.as<float>(), d0, d1, 1.0f)) return fail(-2, "chebyshev launch failed");
		HIPCHK(hipEventRecord(ev_start_, stream_));
		for (int r = 0; r < reps; r++) {
			if (launch_cheb_step(stream_, x, d_r32_.as<float>(), d0, d1, 1.5f)) return fail(-2, "chebyshev launch failed");
			std::swap(d0, d1);
		}
		HIPCHK(hipEventRecord(ev_stop_, stream_));
		HIPCHK(hipEventSynchronize(ev_stop_));
		HIPCHK(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
		HIPCHK(hipMemsetAsync(d_d0_.ptr(), 0, sizeof(float) * cap_slots_, stream_));
		HIPCHK(hipMemsetAsync(d_d1_.ptr(), 0, sizeof(float) * cap_slots_, stream_));
		if (bytes) {
			// the FP32 sweep's 8 nnz + 12 rows + 4 referenced: per entry column +
			// weight, per row its offset, r and the written iterate, every
			// referenced iterate once
			long long nr = 0, ne = 0;
			std::vector<uint8_t> ref(n, 0);
			for (int sl = 0; sl < next_spare_; sl++) {
				const int node = slot_node_[sl];
				if (node < 0) continue;
				nr++;
				ne += inner_deg_[sl];
				ref[node] = 1;
				for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++)
					if (participates(node, mesh_.idx[k]) && !(flags_[mesh_.idx[k]] & 2)) ref[mesh_.idx[k]] = 1;
			}
			long long nref = 0;
			for (uint8_t v : ref) nref += v;
			*bytes = (double) ne * 8 + (double) nr * 12 + (double) nref * 4;
		}
	} else if (kind == 2) {
		if (!accel_)
			if (int rc = accel_setup(0.9)) return rc;
		if (!amg_on_)
			if (int rc = amg_setup()) return rc;
		AmgSolve S{};
		S.x = accel_args();
		S.dn = amg_dn_.as<float>();
		S.p = amg_p_.as<float>();
		S.q = amg_q_.as<float>();
		S.part = amg_part_.as<double>();
		S.st = amg_st_.as<double>();
		// a live solve state: done cleared, p = the current FP32 residual
		const double zero = 0.0;
		HIPCHK(hipMemcpyAsync(S.st + kAmgDone, &zero, sizeof(double), hipMemcpyHostToDevice, stream_));
		HIPCHK(hipMemcpyAsync(S.p, d_r32_.ptr(), sizeof(float) * cap_slots_, hipMemcpyDeviceToDevice, stream_));
		if (launch_amg_fine_matvec(stream_, S)) return fail(-2, "multigrid launch failed");
		HIPCHK(hipEventRecord(ev_start_, stream_));
		for (int r = 0; r < reps; r++)
			if (launch_amg_fine_matvec(stream_, S)) return fail(-2, "multigrid launch failed");
		HIPCHK(hipEventRecord(ev_stop_, stream_));
		HIPCHK(hipEventSynchronize(ev_stop_));
		HIPCHK(hipEventElapsedTime(&ms, ev_start_, ev_stop_));
		if (bytes) {
			// the FP32 sweep's 8 nnz + 12 rows + 4 referenced: per entry column +
			// weight, per row its offset, the row sum and q written, every
			// referenced p once
			long long nr = 0, ne = 0;
			std::vector<uint8_t> ref(n, 0);
			for (int sl = 0; sl < next_spare_; sl++) {
				const int node = slot_node_[sl];
				if (node < 0) continue;
				nr++;
				ne += inner_deg_[sl];
				ref[node] = 1;
				for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++)
					if (participates(node, mesh_.idx[k]) && !(flags_[mesh_.idx[k]] & 2)) ref[mesh_.idx[k]] = 1;
			}
			long long nref = 0;
			for (uint8_t v : ref) nref += v;
			*bytes = (double) ne * 8 + (double) nr * 12 + (double) nref * 4;
		}
	} else
		return fail(-1, "time_kernel: kind");
	HIPCHK(hipMemcpyAsync(d_phi0_.ptr(), save0.ptr(), rs * n, hipMemcpyDeviceToDevice, stream_));
	HIPCHK(hipMemcpyAsync(d_phi1_.ptr(), save1.ptr(), rs * n, hipMemcpyDeviceToDevice, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	*avg_ms = ms / reps;
	return 0;
}

}  // namespace tsg

// ---------------------------------------------------------------- learned warm start

#include "mpnn.hpp"

namespace tsg {

int Grid::attach_mpnn(MpnnNet* net, bool record, bool predict)
{
	if (net && !f64_) return fail(-1, "the learned warm start runs on a TSG_PLAIN_F64 grid");
	net_ = net;
	net_record_ = net && record;
	net_predict_ = net && predict;
	return 0;
}

// The 3-hop shell of the evaporated atom (every node within three grid
// hops, Dirichlet ones included as context) becomes one graph of the
// network's store: edges between shell nodes weighted c_ij / sqrt(W_i W_j)
// (W = coupling sum of the node over all its neighbours), the features and
// the pre-relaxation potentials computed on the device.  With predict the
// network's output (times the shell's max residual) is added to the shell's
// unknown rows, and taken back if it raises the grid's max residual.
int Grid::warm_begin(int node, double phi_ref, bool predict, EventRecord* rec)
{
	const int n = mesh_.n;
	// the shell: every node within three hops (BFS), in ascending node order,
	// with its hop; warm_local_ maps a node to its shell index while the
	// shell is built (-1 elsewhere; cleared again below)
	if ((int) warm_local_.size() != n) warm_local_.assign(n, -1);
	std::vector<int> shell;
	std::vector<signed char> hop;
	{
		std::vector<std::pair<int, signed char>> found{{node, 0}};
		warm_local_[node] = 0;
		size_t lo = 0;
		for (int h = 1; h <= 3; h++) {
			const size_t hi = found.size();
			for (size_t f = lo; f < hi; f++) {
				const int u = found[f].first;
				for (int k = mesh_.ptr[u]; k < mesh_.ptr[u + 1]; k++) {
					const int j = mesh_.idx[k];
					if (warm_local_[j] >= 0) continue;
					warm_local_[j] = 0;
					found.push_back({j, (signed char) h});
				}
			}
			lo = hi;
		}
		std::sort(found.begin(), found.end());
		shell.resize(found.size());
		hop.resize(found.size());
		for (size_t i = 0; i < found.size(); i++) {
			shell[i] = found[i].first;
			hop[i] = found[i].second;
			warm_local_[shell[i]] = (int) i;
		}
	}
	const int ns = (int) shell.size();
	auto local = [&](int g) -> int { return warm_local_[g]; };
	struct Clear {
		std::vector<int>& m;
		const std::vector<int>& s;
		~Clear() { for (int g : s) m[g] = -1; }
	} clear_marks{warm_local_, shell};
	std::vector<double> wsum(ns, 0.0);
	for (int i = 0; i < ns; i++)
		for (int k = mesh_.ptr[shell[i]]; k < mesh_.ptr[shell[i] + 1]; k++) wsum[i] += coup_[k];
	std::vector<int> lptr(ns + 1, 0), lcol, slots(ns);
	std::vector<float> la, wv(ns, 0.0f);
	const int lv = local(node);
	int unknown = 0;
	for (int i = 0; i < ns; i++) {
		const int g = shell[i];
		slots[i] = (flags_[g] & 2) ? -1 : slot_of_[g];
		if (slots[i] >= 0) unknown++;
		for (int k = mesh_.ptr[g]; k < mesh_.ptr[g + 1]; k++) {
			const int lj = local(mesh_.idx[k]);
			if (lj < 0) continue;
			const double den = std::sqrt(wsum[i] * wsum[lj]);
			const float w = den > 0.0 ? (float) (coup_[k] / den) : 0.0f;
			lcol.push_back(lj);
			la.push_back(w);
			if (lj == lv) wv[i] = w;
		}
		lptr[i + 1] = (int) lcol.size();
	}
	const long long ne = (long long) lcol.size();
	MpnnNet& net = *net_;
	if (net.reserve(ns, ne)) {  // store full: stop recording, no prediction this event
		net_record_ = false;
		rec->warm_start = 0;
		warm_.n = 0;
		return 0;
	}
	const long long R = net.rows(), E = net.edges();
	for (int i = 0; i <= ns; i++) lptr[i] = (int) (E + lptr[i]);
	for (auto& c : lcol) c = (int) (R + c);
	if (ns > warm_cap_) {
		warm_cap_ = std::max(ns, 4096);
		d_w_nodes_.alloc(sizeof(int) * warm_cap_);
		d_w_slots_.alloc(sizeof(int) * warm_cap_);
		d_w_hop_.alloc(sizeof(signed char) * warm_cap_);
		d_w_wv_.alloc(sizeof(float) * warm_cap_);
		d_w_pre_.alloc(sizeof(double) * warm_cap_);
		d_w_res_.alloc(sizeof(double) * warm_cap_);
		d_w_scale_.alloc(sizeof(double));
		if (!d_w_nodes_.ptr() || !d_w_res_.ptr() || !d_w_scale_.ptr()) return fail(-2, "device allocation failed (warm start)");
	}
	// the uploads go through pinned staging (the stream is synchronised
	// before this function returns, so the staging is free again)
	const size_t need = 64 + 16 * 7 + sizeof(int) * (ns + 1) + (sizeof(int) + sizeof(float)) * (size_t) ne +
	                    (sizeof(int) * 2 + 1 + sizeof(float)) * (size_t) ns + 2 * sizeof(double) * 4;
	if (need > warm_host_bytes_) {
		HIPCHK(hipStreamSynchronize(stream_));
		if (warm_host_) (void) hipHostFree(warm_host_);
		warm_host_ = nullptr;
		warm_host_bytes_ = std::max(need * 2, (size_t) 1 << 20);
		HIPCHK(hipHostMalloc(&warm_host_, warm_host_bytes_, hipHostMallocDefault));
	}
	char* hb = (char*) warm_host_;
	size_t hoff = 0;
	auto stage = [&](void* dst, const void* src, size_t bytes) -> int {
		if (!bytes) return 0;
		std::memcpy(hb + hoff, src, bytes);
		const hipError_t e = hipMemcpyAsync(dst, hb + hoff, bytes, hipMemcpyHostToDevice, stream_);
		hoff += (bytes + 15) & ~(size_t) 15;
		return e == hipSuccess ? 0 : -1;
	};
	if (stage(net.ptr_at(R), lptr.data(), sizeof(int) * (ns + 1)) || stage(net.col_at(E), lcol.data(), sizeof(int) * ne) ||
	    stage(net.a_at(E), la.data(), sizeof(float) * ne) || stage(d_w_nodes_.ptr(), shell.data(), sizeof(int) * ns) ||
	    stage(d_w_slots_.ptr(), slots.data(), sizeof(int) * ns) ||
	    stage(d_w_hop_.ptr(), hop.data(), sizeof(signed char) * ns) || stage(d_w_wv_.ptr(), wv.data(), sizeof(float) * ns))
		return fail(-2, "warm start: upload failed");
	double* hscal = (double*) (hb + hoff);  // pinned scalars read back below

	WarmArgs& w = warm_;
	w.n = ns;
	w.nodes = d_w_nodes_.as<int>();
	w.slots = d_w_slots_.as<int>();
	w.hop = d_w_hop_.as<signed char>();
	w.wv = d_w_wv_.as<float>();
	w.xyz = d_xyz_.as<float>();
	double ell = 0.0;
	int cnt = 0;
	for (int k = mesh_.ptr[node]; k < mesh_.ptr[node + 1]; k++) {
		const int j = mesh_.idx[k];
		double d2 = 0.0;
		for (int c = 0; c < 3; c++) {
			const double d = (double) mesh_.xyz[3 * (size_t) j + c] - mesh_.xyz[3 * (size_t) node + c];
			d2 += d * d;
		}
		ell += std::sqrt(d2);
		cnt++;
	}
	for (int c = 0; c < 3; c++) w.v[c] = mesh_.xyz[3 * (size_t) node + c];
	w.ell = cnt ? (float) (ell / cnt) : 1.0f;
	w.phi_ref = phi_ref;
	w.phi = (sw_ ? d_phi1_ : d_phi0_).as<double>();
	w.phi_pre = d_w_pre_.as<double>();
	w.resid = d_w_res_.as<double>();
	w.scale = d_w_scale_.as<double>();
	w.x = net.x_at(R);
	w.t = net.t_at(R);
	w.mask = net.mask_at(R);
	warm_row0_ = R;
	warm_edges_ = ne;
	warm_unknown_ = unknown;
	if (launch_warm_features(stream_, w, sweep_args())) return fail(-2, "warm-start feature launch failed");
	HIPCHK(hipMemcpyAsync(hscal, w.scale, sizeof(double), hipMemcpyDeviceToHost, stream_));
	rec->warm_start = 3;
	if (predict) {
		MpnnBatch b;
		b.n = ns;
		b.row0 = (int) R;
		b.msum = (float) std::max(1, unknown);
		if (int rc = net.forward(b, stream_)) return fail(rc, "warm start: " + net.error());
		if (launch_warm_apply(stream_, w, net.y())) return fail(-2, "warm-start apply launch failed");
		// the grid's max residual after the prediction
		SweepArgs a = sweep_args();
		double* red = hscal + 2;
		if (launch_resid64(stream_, a, w.phi, d_r32_.as<float>(), d_redout_.as<double>())) return fail(-2, "residual launch failed");
		HIPCHK(hipMemcpyAsync(red, d_redout_.ptr(), 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
		HIPCHK(hipStreamSynchronize(stream_));
		const double s = hscal[0];
		rec->residual_warm = red[0];
		rec->warm_start = 1;
		if (!(red[0] <= s)) {  // worse than no prediction (or NaN): start from the removal state
			if (launch_warm_apply(stream_, w, nullptr)) return fail(-2, "warm-start apply launch failed");
			rec->warm_start = 2;
		} else if (red[1] == 0.0) {  // the relaxation's first residual is this one
			r32_ready_ = true;
			r32_rmax_ = red[0];
		}
	} else
		HIPCHK(hipStreamSynchronize(stream_));
	rec->residual_cold = hscal[0];
	return 0;
}

// the converged potentials give the graph's targets; the graph is kept
int Grid::warm_end()
{
	if (!net_record_ || warm_.n == 0) return 0;
	if (launch_warm_targets(stream_, warm_)) return fail(-2, "warm-start target launch failed");
	HIPCHK(hipStreamSynchronize(stream_));
	net_->commit_graph(warm_.n, warm_edges_, (float) std::max(1, warm_unknown_));
	return 0;
}

}  // namespace tsg

// ---------------------------------------------------------------- trajectories

namespace tsg {

// Trajectory_3d::init + integrate(delta, errorLimit) + extrapolate(Z, zmax)
// (trajectory_3d.cpp:99-332) for a batch of ions, each starting at rest on
// its node, in the field of the current potentials (the buffer potential()
// reads); one device thread per ion
int Grid::trajectories(int count, const int* nodes, const float* charge, const float* mass, int integrator, int stepper,
                       float dt0, float dt_min, float err_limit, int cap, float* out, int* len, int* status, float* summary,
                       const float* v0, bool process_rand, float* error_estimate)
{
	if (count <= 0) return 0;
	if (cap < 0 || !nodes || !charge || !mass || !len || !status || (cap > 0 && !out)) return fail(-1, "trajectory: bad arguments");
	if (process_rand && count != 1) return fail(-1, "trajectory: the rand() stream is continued for one ion at a time");
	if (integrator < 0 || integrator > 2 || stepper < 0 || stepper > 3) return fail(-1, "trajectory: unknown integrator or stepper");
	for (int i = 0; i < count; i++)
		if (nodes[i] < 0 || nodes[i] >= mesh_.n) return fail(-1, "trajectory: bad node");
	const int n = mesh_.n, m = mesh_.m();
	if (int rc = ensure_point_location()) return rc;
	HIPCHK(hipMemsetAsync(d_status_.ptr(), 0, sizeof(int), stream_));
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	if (launch_field_snapshot(stream_, n, d_xyz_.as<float>(), d_ptr_.as<int>(), d_idx_.as<int>(), d_coupfull_.ptr(),
	                          d_flags_.as<unsigned char>(), phi, f64_, d_nfield_.as<float>(), d_status_.as<int>()))
		return fail(-2, "field snapshot launch failed");
	// per ion: node, snapshot, charge, mass, start velocity in; len, status,
	// 12-float summary and the 6-float error estimate out
	const size_t in_bytes = 28 * (size_t) count, res_bytes = 80 * (size_t) count;
	const size_t out_bytes = sizeof(float) * 8 * (size_t) count * (size_t) std::max(cap, 1);
	if (d_traj_in_.bytes() < in_bytes) d_traj_in_.alloc(in_bytes * 2);
	if (d_traj_res_.bytes() < res_bytes) d_traj_res_.alloc(res_bytes * 2);
	if (cap > 0 && d_traj_out_.bytes() < out_bytes) d_traj_out_.alloc(out_bytes);
	if (!d_traj_in_.ptr() || !d_traj_res_.ptr() || (cap > 0 && !d_traj_out_.ptr()))
		return fail(-2, "device allocation failed (trajectory batch)");
	std::vector<int> snap(count, 0);
	char* din = d_traj_in_.as<char>();
	HIPCHK(hipMemcpyAsync(din, nodes, sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemcpyAsync(din + sizeof(int) * count, snap.data(), sizeof(int) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemcpyAsync(din + sizeof(int) * 2 * count, charge, sizeof(float) * count, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemcpyAsync(din + sizeof(int) * 2 * count + sizeof(float) * count, mass, sizeof(float) * count,
	                      hipMemcpyHostToDevice, stream_));
	float* dv0 = (float*) (din + 16 * (size_t) count);
	if (v0) HIPCHK(hipMemcpyAsync(dv0, v0, sizeof(float) * 3 * count, hipMemcpyHostToDevice, stream_));
	RandState& prs = tsg::process_rand();
	if (process_rand) HIPCHK(hipMemcpyAsync(d_rng_.ptr(), &prs, sizeof(RandState), hipMemcpyHostToDevice, stream_));
	TrajArgs a;
	std::memset(&a, 0, sizeof(a));
	a.count = count;
	a.nodes = (const int*) din;
	a.snapshot = (const int*) (din + sizeof(int) * count);
	a.charge = (const float*) (din + sizeof(int) * 2 * count);
	a.mass = (const float*) (din + sizeof(int) * 2 * count + sizeof(float) * count);
	a.integrator = integrator;
	a.stepper = stepper;
	a.dt0 = dt0;
	a.dt_min = dt_min;
	a.err_limit = err_limit;
	a.zmax = zmax_;
	a.max_steps = 1 << 24;
	a.max_calls = 1ll << 22;
	a.n = n;
	a.m = m;
	a.xyz = d_xyz_.as<float>();
	a.tv = d_tv_.as<int>();
	a.tn = d_tn_.as<int>();
	a.assoc = d_assoc_.as<int>();
	a.gbnd = d_gbnd_.as<unsigned char>();
	a.nfield = d_nfield_.as<float>();
	a.out = cap > 0 ? d_traj_out_.as<float>() : nullptr;
	a.cap = cap;
	int* res = d_traj_res_.as<int>();
	a.len = res;
	a.status = res + count;
	a.summary = summary ? (float*) (res + 2 * count) : nullptr;
	a.v0 = v0 ? dv0 : nullptr;
	a.rng = process_rand ? d_rng_.as<int>() : nullptr;
	a.err_out = error_estimate ? (float*) (res + 14 * count) : nullptr;
	if (launch_trajectories(stream_, a)) return fail(-2, "trajectory launch failed");
	int st = 0;
	if (cap > 0) HIPCHK(hipMemcpyAsync(out, d_traj_out_.ptr(), out_bytes, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(len, res, sizeof(int) * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(status, res + count, sizeof(int) * count, hipMemcpyDeviceToHost, stream_));
	if (summary) HIPCHK(hipMemcpyAsync(summary, res + 2 * count, sizeof(float) * 12 * count, hipMemcpyDeviceToHost, stream_));
	if (error_estimate)
		HIPCHK(hipMemcpyAsync(error_estimate, res + 14 * count, sizeof(float) * 6 * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(&st, d_status_.ptr(), sizeof(int), hipMemcpyDeviceToHost, stream_));
	if (process_rand) HIPCHK(hipMemcpyAsync(&prs, d_rng_.ptr(), sizeof(RandState), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	if (st) return fail(-4, "Grid_3d::Table::field_o1(): singular least-squares system");
	return 0;
}

int Grid::fieldlines(int count, const float* start, float distance, long long iterations, long long stepper_limit,
                     bool process_rand, int cap, float* out, int* len, int* status)
{
	if (count <= 0) return 0;
	if (!start || !len || !status || cap < 0 || (cap > 0 && !out) || iterations < 0 || stepper_limit < 0)
		return fail(-1, "fieldlines: bad arguments");
	const int n = mesh_.n, m = mesh_.m();
	if (int rc = ensure_point_location()) return rc;
	HIPCHK(hipMemsetAsync(d_status_.ptr(), 0, sizeof(int), stream_));
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	if (launch_field_snapshot(stream_, n, d_xyz_.as<float>(), d_ptr_.as<int>(), d_idx_.as<int>(), d_coupfull_.ptr(),
	                          d_flags_.as<unsigned char>(), phi, f64_, d_nfield_.as<float>(), d_status_.as<int>()))
		return fail(-2, "field snapshot launch failed");
	DevBuf din, dout;
	const size_t in_bytes = sizeof(float) * 3 * (size_t) count + sizeof(int) * 2 * (size_t) count;
	const size_t out_bytes = sizeof(float) * 7 * (size_t) count * (size_t) std::max(cap, 1);
	din.alloc(in_bytes);
	if (cap > 0) dout.alloc(out_bytes);
	if (!din.ptr() || (cap > 0 && !dout.ptr())) return fail(-2, "device allocation failed (field lines)");
	HIPCHK(hipMemcpyAsync(din.ptr(), start, sizeof(float) * 3 * count, hipMemcpyHostToDevice, stream_));
	int* res = (int*) (din.as<char>() + sizeof(float) * 3 * (size_t) count);
	RandState& prs = tsg::process_rand();
	if (process_rand) HIPCHK(hipMemcpyAsync(d_rng_.ptr(), &prs, sizeof(RandState), hipMemcpyHostToDevice, stream_));
	FieldlineArgs a;
	std::memset(&a, 0, sizeof(a));
	a.count = count;
	a.start = din.as<float>();
	a.distance = distance;
	a.iterations = iterations;
	a.stepper_limit = stepper_limit;
	a.max_calls = 1ll << 22;
	a.guess_samples = std::max(1, (int) std::pow((double) m, 0.25));
	a.m = m;
	a.xyz = d_xyz_.as<float>();
	a.tv = d_tv_.as<int>();
	a.tn = d_tn_.as<int>();
	a.gbnd = d_gbnd_.as<unsigned char>();
	a.nfield = d_nfield_.as<float>();
	a.cap = cap;
	a.out = cap > 0 ? dout.as<float>() : nullptr;
	a.len = res;
	a.status = res + count;
	a.rng = process_rand ? d_rng_.as<int>() : nullptr;
	if (launch_fieldlines(stream_, a)) return fail(-2, "field line launch failed");
	int st = 0;
	if (cap > 0) HIPCHK(hipMemcpyAsync(out, dout.ptr(), out_bytes, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(len, res, sizeof(int) * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(status, res + count, sizeof(int) * count, hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipMemcpyAsync(&st, d_status_.ptr(), sizeof(int), hipMemcpyDeviceToHost, stream_));
	if (process_rand) HIPCHK(hipMemcpyAsync(&prs, d_rng_.ptr(), sizeof(RandState), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	if (st) return fail(-4, "Grid_3d::Table::field_o1(): singular least-squares system");
	return 0;
}

// the mesh arrays of the point location, kept on the device from the first use on
int Grid::ensure_point_location()
{
	if (d_tv_.ptr()) return 0;
	const int n = mesh_.n, m = mesh_.m();
	d_tv_.alloc(sizeof(int) * 4 * (size_t) m);
	d_tn_.alloc(sizeof(int) * 4 * (size_t) m);
	d_assoc_.alloc(sizeof(int) * (size_t) n);
	d_gbnd_.alloc((size_t) n);
	d_nfield_.alloc(sizeof(float) * 3 * (size_t) n);
	d_ppos_.alloc(sizeof(float) * 4 + sizeof(int) * 4);
	d_rng_.alloc(sizeof(RandState));
	if (!d_tv_.ptr() || !d_tn_.ptr() || !d_assoc_.ptr() || !d_gbnd_.ptr() || !d_nfield_.ptr() || !d_ppos_.ptr() ||
	    !d_rng_.ptr()) {
		d_tv_.free();
		return fail(-2, "device allocation failed (point location)");
	}
	HIPCHK(hipMemcpy(d_tv_.ptr(), mesh_.tet_v.data(), d_tv_.bytes(), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_tn_.ptr(), mesh_.tet_n.data(), d_tn_.bytes(), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_assoc_.ptr(), mesh_.assoc.data(), d_assoc_.bytes(), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_gbnd_.ptr(), mesh_.boundary.data(), d_gbnd_.bytes(), hipMemcpyHostToDevice));
	zmax_ = mesh_.xyz[2];
	for (int i = 1; i < n; i++) zmax_ = std::max(zmax_, mesh_.xyz[3 * (size_t) i + 2]);
	return 0;
}

// Grid_3d::potential(pos, guess) (grid_3d.cpp:1585-1602) on the potential() buffer
int Grid::point_potential(const float pos[3], int guess, float* out, bool process_rand, int walks)
{
	if (walks < 1) return fail(-1, "point potential: walks must be positive");
	if (int rc = ensure_point_location()) return rc;
	RandState& prs = tsg::process_rand();
	if (process_rand) HIPCHK(hipMemcpyAsync(d_rng_.ptr(), &prs, sizeof(RandState), hipMemcpyHostToDevice, stream_));
	float* dp = d_ppos_.as<float>();
	int* dg = (int*) (dp + 3);
	int g = guess;
	float in[3] = {pos[0], pos[1], pos[2]};
	HIPCHK(hipMemcpyAsync(dp, in, sizeof(float) * 3, hipMemcpyHostToDevice, stream_));
	HIPCHK(hipMemcpyAsync(dg, &g, sizeof(int), hipMemcpyHostToDevice, stream_));
	const void* phi = ((sw_ + 1) % 2) ? d_phi1_.ptr() : d_phi0_.ptr();
	float* dout = (float*) (dg + 1);
	if (launch_point_potential(stream_, 1, dp, dg, d_xyz_.as<float>(), d_tv_.as<int>(), d_tn_.as<int>(), mesh_.m(), phi,
	                           f64_, dout, process_rand ? d_rng_.as<int>() : nullptr, walks))
		return fail(-2, "point potential launch failed");
	HIPCHK(hipMemcpyAsync(out, dout, sizeof(float), hipMemcpyDeviceToHost, stream_));
	if (process_rand) HIPCHK(hipMemcpyAsync(&prs, d_rng_.ptr(), sizeof(RandState), hipMemcpyDeviceToHost, stream_));
	HIPCHK(hipStreamSynchronize(stream_));
	return 0;
}

}  // namespace tsg

// ---------------------------------------------------------------- initial potentials

namespace tsg {

namespace {

template <typename R>
R length3(const R* a)
{
	R v = 0;
	for (int k = 0; k < 3; k++) v += a[k] * a[k];
	return std::sqrt(v);
}

}  // namespace

// Grid_3d::Table::reset (grid_3d.cpp:797-809): every node takes its species preset
int Grid::reset_potentials()
{
	std::vector<double> phi(mesh_.n);
	for (int i = 0; i < mesh_.n; i++) phi[i] = species_[ids_[i]].phi;
	return set_phi(phi.data(), phi.data(), sw_, lsw_);
}

// Grid_3d::Table::randomReset (grid_3d.cpp:811-848), in the grid's element
// type, drawing from the process's rand()
template <typename R>
static void random_potentials(const std::vector<int16_t>& ids, const std::vector<uint8_t>& flags, const Species* sp,
                              std::vector<double>* out)
{
	int first = -1;
	for (int id = 0; id < 128; id++)
		if (sp[id].defined) {
			first = id;
			break;
		}
	if (first < 0) return;
	R minPhi = sp[first].phi, maxPhi = sp[first].phi;
	for (int id = first; id < 128; id++) {
		if (!sp[id].defined) continue;
		const R v = sp[id].phi;
		if (v < minPhi) minPhi = v;
		else if (v > maxPhi) maxPhi = v;
	}
	for (size_t i = 0; i < ids.size(); i++) {
		R phi;
		if (flags[i] & 2) phi = sp[ids[i]].phi;
		else {
			phi = (R) prand();
			phi /= RAND_MAX;
			phi *= maxPhi - minPhi;
			phi += minPhi;
		}
		(*out)[i] = phi;
	}
}

int Grid::random_potentials()
{
	std::vector<double> phi(mesh_.n);
	if (f64_) tsg::random_potentials<double>(ids_, flags_, species_, &phi);
	else tsg::random_potentials<float>(ids_, flags_, species_, &phi);
	return set_phi(phi.data(), phi.data(), sw_, lsw_);
}

// preset_sphericalPotential (system_3d.cpp:24-57): a radial ramp from pmax
// at the centre, its slope set by the distance of the bounding box's lower
// corner (the reference uses that corner for both extreme distances)
template <typename R>
static void spherical_potentials(const Mesh& mesh, const std::vector<int16_t>& ids, const std::vector<uint8_t>& flags,
                                 const Species* sp, const float* center, float pmin, float pmax, std::vector<double>* out)
{
	R lo[3], c[3], d[3];
	for (int k = 0; k < 3; k++) {
		lo[k] = mesh.xyz[k];
		c[k] = center[k];
	}
	for (int i = 1; i < mesh.n; i++)
		for (int k = 0; k < 3; k++)
			if (mesh.xyz[3 * (size_t) i + k] < lo[k]) lo[k] = mesh.xyz[3 * (size_t) i + k];
	for (int k = 0; k < 3; k++) d[k] = lo[k] - c[k];
	const R minDistance = length3(d), maxDistance = length3(d);
	R phiSlope = pmax;
	phiSlope -= pmin;
	if (minDistance > maxDistance) phiSlope /= minDistance;
	else phiSlope /= maxDistance;
	for (int i = 0; i < mesh.n; i++) {
		R phi;
		if (flags[i] & 2) phi = sp[ids[i]].phi;
		else {
			for (int k = 0; k < 3; k++) d[k] = (R) mesh.xyz[3 * (size_t) i + k] - c[k];
			phi = pmax;
			phi -= phiSlope * length3(d);
		}
		(*out)[i] = phi;
	}
}

int Grid::spherical_potentials(const float center[3], float pmin, float pmax)
{
	std::vector<double> phi(mesh_.n);
	if (f64_) tsg::spherical_potentials<double>(mesh_, ids_, flags_, species_, center, pmin, pmax, &phi);
	else tsg::spherical_potentials<float>(mesh_, ids_, flags_, species_, center, pmin, pmax, &phi);
	return set_phi(phi.data(), phi.data(), sw_, lsw_);
}

}  // namespace tsg
