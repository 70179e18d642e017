// Warm-start network: parameters, graph store, forward / backward / Adam.
#include "mpnn.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

namespace tsg {

#define MCHK(x)                                        \
	do {                                               \
		if (int rc__ = hip_check((x), #x)) return rc__; \
	} while (0)
#define LCHK(x)                                                            \
	do {                                                                   \
		if ((x) != 0) return fail(-2, std::string("launch failed: ") + #x); \
	} while (0)

int MpnnNet::fail(int code, const std::string& msg)
{
	err_ = msg;
	status_ = code;
	return code;
}

int MpnnNet::hip_check(hipError_t e, const char* what)
{
	if (e == hipSuccess) return 0;
	return fail(-2, std::string(what) + ": " + hipGetErrorString(e));
}

MpnnNet::MpnnNet(int max_batch_rows, long long data_rows, long long data_edges, unsigned seed)
{
	if (max_batch_rows <= 0 || data_rows <= 0 || data_edges <= 0) {
		fail(-1, "network: sizes must be positive");
		return;
	}
	if (hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate")) return;
	cap_rows_ = data_rows;
	cap_edges_ = data_edges;
	d_p_.alloc(sizeof(float) * kMpnnNParam);
	d_g_.alloc(sizeof(float) * kMpnnNParam);
	d_m_.alloc(sizeof(float) * kMpnnNParam);
	d_v_.alloc(sizeof(float) * kMpnnNParam);
	d_win_t_.alloc(sizeof(bf16_t) * kMpnnIn * kMpnnHid);
	d_wl_t_.alloc(sizeof(bf16_t) * kMpnnLayers * kMpnnCat * kMpnnHid);
	d_wl_.alloc(sizeof(bf16_t) * kMpnnLayers * kMpnnCat * kMpnnHid);
	d_wout_.alloc(sizeof(bf16_t) * kMpnnHid);
	d_ptr_.alloc(sizeof(int) * (size_t) (data_rows + 1));
	d_col_.alloc(sizeof(int) * (size_t) data_edges);
	d_a_.alloc(sizeof(float) * (size_t) data_edges);
	d_x_.alloc(sizeof(bf16_t) * kMpnnIn * (size_t) data_rows);
	d_t_.alloc(sizeof(float) * (size_t) data_rows);
	d_mask_.alloc(sizeof(float) * (size_t) data_rows);
	d_losses_.alloc(sizeof(float) * 65536);
	if (!d_p_.ptr() || !d_ptr_.ptr() || !d_col_.ptr() || !d_x_.ptr() || !d_losses_.ptr()) {
		fail(-2, "network: device allocation failed");
		return;
	}
	if (hip_check(hipMemset(d_ptr_.ptr(), 0, sizeof(int)), "hipMemset")) return;
	if (ensure_work(max_batch_rows)) return;

	// initial parameters: uniform with the fan-in scale of each layer; the
	// residual layers start small so that the stack begins near identity
	std::vector<float> p(kMpnnNParam, 0.0f);
	std::mt19937 rng(seed);
	auto fill = [&](int off, int count, float bound) {
		std::uniform_real_distribution<float> u(-bound, bound);
		for (int i = 0; i < count; i++) p[off + i] = u(rng);
	};
	fill(kMpnnWIn, kMpnnIn * kMpnnHid, std::sqrt(6.0f / kMpnnIn));
	for (int o = 0; o < kMpnnHid; o++) p[kMpnnBIn + o] = 0.01f;
	for (int l = 0; l < kMpnnLayers; l++) fill(kMpnnLayer0 + l * kMpnnLayerStride, kMpnnCat * kMpnnHid, 0.3f * std::sqrt(6.0f / kMpnnCat));
	fill(kMpnnWOut, kMpnnHid, std::sqrt(6.0f / kMpnnHid) * 0.1f);
	set_params(p.data());
}

MpnnNet::~MpnnNet()
{
	if (stream_) (void) hipStreamDestroy(stream_);
}

MpnnWeights MpnnNet::weights() const
{
	MpnnWeights w;
	w.win_t = d_win_t_.as<bf16_t>();
	w.wl_t = d_wl_t_.as<bf16_t>();
	w.wl = d_wl_.as<bf16_t>();
	w.wout = d_wout_.as<bf16_t>();
	return w;
}

int MpnnNet::ensure_work(int n)
{
	if (n <= max_rows_) return 0;
	const size_t r = (size_t) n;
	const size_t nb = (r + 255) / 256;
	d_z_.alloc(sizeof(float) * (kMpnnLayers + 1) * kMpnnHid * r);
	d_h_.alloc(sizeof(float) * (kMpnnLayers + 1) * kMpnnHid * r);
	d_cat_.alloc(sizeof(bf16_t) * (kMpnnLayers + 1) * kMpnnCat * r);
	d_dh_.alloc(sizeof(float) * kMpnnHid * r);
	d_dz_.alloc(sizeof(float) * kMpnnHid * r);
	d_dzb_.alloc(sizeof(bf16_t) * kMpnnHid * r);
	d_dcat_.alloc(sizeof(float) * kMpnnCat * r);
	d_y_.alloc(sizeof(float) * r);
	d_dy_.alloc(sizeof(float) * r);
	d_part_.alloc(sizeof(float) * nb * kMpnnCat * kMpnnHid);
	d_lpart_.alloc(sizeof(float) * 2 * nb);
	if (!d_z_.ptr() || !d_h_.ptr() || !d_cat_.ptr() || !d_dcat_.ptr() || !d_part_.ptr())
		return fail(-2, "network: work buffer allocation failed");
	max_rows_ = n;
	return 0;
}

int MpnnNet::set_params(const float* host)
{
	MCHK(hipMemcpyAsync(d_p_.ptr(), host, sizeof(float) * kMpnnNParam, hipMemcpyHostToDevice, stream_));
	MCHK(hipMemsetAsync(d_m_.ptr(), 0, sizeof(float) * kMpnnNParam, stream_));
	MCHK(hipMemsetAsync(d_v_.ptr(), 0, sizeof(float) * kMpnnNParam, stream_));
	LCHK(launch_mpnn_pack(stream_, d_p_.as<float>(), weights()));
	MCHK(hipStreamSynchronize(stream_));
	step_ = 0;
	return 0;
}

int MpnnNet::get_params(float* host)
{
	MCHK(hipMemcpyAsync(host, d_p_.ptr(), sizeof(float) * kMpnnNParam, hipMemcpyDeviceToHost, stream_));
	MCHK(hipStreamSynchronize(stream_));
	return 0;
}

int MpnnNet::read_grad(float* host)
{
	MCHK(hipMemcpyAsync(host, d_g_.ptr(), sizeof(float) * kMpnnNParam, hipMemcpyDeviceToHost, stream_));
	MCHK(hipStreamSynchronize(stream_));
	return 0;
}

int MpnnNet::clear()
{
	rows_ = edges_ = 0;
	g_row0_.clear();
	g_edge0_.clear();
	g_n_.clear();
	g_msum_.clear();
	return 0;
}

int MpnnNet::reserve(int n, long long ne)
{
	if (n <= 0) return fail(-1, "network: empty graph");
	if (rows_ + n > cap_rows_ || edges_ + ne > cap_edges_) return fail(-7, "network: graph store full");
	if (n > max_rows_)
		if (int rc = ensure_work(n)) return rc;
	return 0;
}

int MpnnNet::commit_graph(int n, long long ne, float msum)
{
	g_row0_.push_back(rows_);
	g_edge0_.push_back(edges_);
	g_n_.push_back(n);
	g_msum_.push_back(msum);
	rows_ += n;
	edges_ += ne;
	return 0;
}

int MpnnNet::upload_graph(int n, const int* ptr, const int* col, const float* a, const float* x, const float* t,
                          const float* mask, bool commit)
{
	const long long ne = ptr[n];
	if (int rc = reserve(n, ne)) return rc;
	std::vector<int> gp(n + 1), gc(ne);
	for (int i = 0; i <= n; i++) gp[i] = (int) (edges_ + ptr[i]);
	for (long long e = 0; e < ne; e++) {
		if (col[e] < 0 || col[e] >= n) return fail(-1, "network: edge column out of range");
		gc[e] = (int) (rows_ + col[e]);
	}
	std::vector<bf16_t> xb((size_t) n * kMpnnIn);
	for (size_t i = 0; i < xb.size(); i++) {
		uint32_t u;
		std::memcpy(&u, &x[i], 4);
		u += 0x7fffu + ((u >> 16) & 1u);
		xb[i] = (bf16_t) (u >> 16);
	}
	float msum = 0.0f;
	for (int i = 0; i < n; i++) msum += mask[i];
	MCHK(hipMemcpy(ptr_at(rows_), gp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
	if (ne) {
		MCHK(hipMemcpy(col_at(edges_), gc.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
		MCHK(hipMemcpy(a_at(edges_), a, sizeof(float) * ne, hipMemcpyHostToDevice));
	}
	MCHK(hipMemcpy(x_at(rows_), xb.data(), sizeof(bf16_t) * xb.size(), hipMemcpyHostToDevice));
	MCHK(hipMemcpy(t_at(rows_), t, sizeof(float) * n, hipMemcpyHostToDevice));
	MCHK(hipMemcpy(mask_at(rows_), mask, sizeof(float) * n, hipMemcpyHostToDevice));
	if (commit) commit_graph(n, ne, msum);
	return 0;
}

MpnnBatch MpnnNet::batch(int g0, int g1) const
{
	MpnnBatch b;
	b.row0 = (int) g_row0_[g0];
	b.n = (int) (g_row0_[g1 - 1] + g_n_[g1 - 1] - g_row0_[g0]);
	float m = 0.0f;
	for (int g = g0; g < g1; g++) m += g_msum_[g];
	b.msum = std::max(1.0f, m);
	return b;
}

int MpnnNet::forward(const MpnnBatch& b, hipStream_t on)
{
	hipStream_t st = on ? on : stream_;
	const int n = b.n;
	if (n > max_rows_) return fail(-1, "network: batch larger than the work buffers");
	const size_t hs = (size_t) max_rows_ * kMpnnHid, cs = (size_t) max_rows_ * kMpnnCat;
	float* Z = d_z_.as<float>();
	float* H = d_h_.as<float>();
	bf16_t* C = d_cat_.as<bf16_t>();
	const float* p = d_p_.as<float>();
	const int* ptr = d_ptr_.as<int>();
	const int* col = d_col_.as<int>();
	const float* a = d_a_.as<float>();
	// layer 0 writes bf16(h0) into the left half of layer 1's operand
	LCHK(launch_mpnn_gemm(st, n, kMpnnIn, kMpnnHid, x_at(b.row0), d_win_t_.as<bf16_t>(), nullptr, p + kMpnnBIn,
	                      nullptr, Z, H, C + cs, kMpnnCat));
	for (int l = 1; l <= kMpnnLayers; l++) {
		LCHK(launch_mpnn_aggregate(st, n, b.row0, ptr, col, a, H + (l - 1) * hs, C + l * cs));
		const float* bias = p + kMpnnLayer0 + (l - 1) * kMpnnLayerStride + kMpnnCat * kMpnnHid;
		bf16_t* next = l < kMpnnLayers ? C + (l + 1) * cs : nullptr;
		LCHK(launch_mpnn_gemm(st, n, kMpnnCat, kMpnnHid, C + l * cs,
		                      d_wl_t_.as<bf16_t>() + (size_t) (l - 1) * kMpnnCat * kMpnnHid, nullptr, bias,
		                      H + (l - 1) * hs, Z + l * hs, H + l * hs, next, kMpnnCat));
	}
	LCHK(launch_mpnn_out(st, n, H + kMpnnLayers * hs, d_wout_.as<bf16_t>(), p + kMpnnBOut, d_y_.as<float>()));
	return 0;
}

int MpnnNet::backward(const MpnnBatch& b, int slot)
{
	const int n = b.n;
	const size_t hs = (size_t) max_rows_ * kMpnnHid, cs = (size_t) max_rows_ * kMpnnCat;
	float* Z = d_z_.as<float>();
	float* H = d_h_.as<float>();
	bf16_t* C = d_cat_.as<bf16_t>();
	float* g = d_g_.as<float>();
	float* dh = d_dh_.as<float>();
	float* dz = d_dz_.as<float>();
	bf16_t* dzb = d_dzb_.as<bf16_t>();
	float* dcat = d_dcat_.as<float>();
	float* part = d_part_.as<float>();
	float* lpart = d_lpart_.as<float>();
	const int* ptr = d_ptr_.as<int>();
	const int* col = d_col_.as<int>();
	const float* a = d_a_.as<float>();
	int nb = 0;
	LCHK(launch_mpnn_loss(stream_, n, d_y_.as<float>(), t_at(b.row0), mask_at(b.row0), b.msum, d_dy_.as<float>(), lpart, &nb));
	// lpart holds (loss, sum dy) pairs per block
	LCHK(launch_mpnn_reduce(stream_, nb, 2, lpart, part, 0));
	MCHK(hipMemcpyAsync(d_losses_.as<float>() + slot, part, sizeof(float), hipMemcpyDeviceToDevice, stream_));
	MCHK(hipMemcpyAsync(g + kMpnnBOut, part + 1, sizeof(float), hipMemcpyDeviceToDevice, stream_));
	// output layer
	LCHK(launch_mpnn_colsum(stream_, n, H + kMpnnLayers * hs, d_dy_.as<float>(), 1, part, &nb));
	LCHK(launch_mpnn_reduce(stream_, nb, kMpnnHid, part, g + kMpnnWOut, 0));
	LCHK(launch_mpnn_out_bwd(stream_, n, d_dy_.as<float>(), d_wout_.as<bf16_t>(), dh));
	for (int l = kMpnnLayers; l >= 1; l--) {
		float* gW = g + kMpnnLayer0 + (l - 1) * kMpnnLayerStride;
		LCHK(launch_mpnn_relu_bwd(stream_, n, dh, Z + l * hs, dz, dzb));
		LCHK(launch_mpnn_colsum(stream_, n, dz, nullptr, 0, part, &nb));
		LCHK(launch_mpnn_reduce(stream_, nb, kMpnnHid, part, gW + kMpnnCat * kMpnnHid, 0));
		LCHK(launch_mpnn_wgrad(stream_, n, kMpnnCat, C + l * cs, dzb, part, &nb));
		LCHK(launch_mpnn_reduce(stream_, nb, kMpnnCat * kMpnnHid, part, gW, 0));
		LCHK(launch_mpnn_gemm(stream_, n, kMpnnHid, kMpnnCat, dzb, d_wl_.as<bf16_t>() + (size_t) (l - 1) * kMpnnCat * kMpnnHid,
		                      dcat, nullptr, nullptr, nullptr, nullptr, nullptr, 0));
		LCHK(launch_mpnn_aggregate_bwd(stream_, n, b.row0, ptr, col, a, dh, dcat, dh));
	}
	LCHK(launch_mpnn_relu_bwd(stream_, n, dh, Z, dz, dzb));
	LCHK(launch_mpnn_colsum(stream_, n, dz, nullptr, 0, part, &nb));
	LCHK(launch_mpnn_reduce(stream_, nb, kMpnnHid, part, g + kMpnnBIn, 0));
	LCHK(launch_mpnn_wgrad(stream_, n, kMpnnIn, x_at(b.row0), dzb, part, &nb));
	LCHK(launch_mpnn_reduce(stream_, nb, kMpnnIn * kMpnnHid, part, g + kMpnnWIn, 0));
	return 0;
}

int MpnnNet::adam(float lr)
{
	step_++;
	LCHK(launch_mpnn_adam(stream_, d_p_.as<float>(), d_g_.as<float>(), d_m_.as<float>(), d_v_.as<float>(), lr, 0.9f,
	                      0.999f, 1e-8f, step_));
	LCHK(launch_mpnn_pack(stream_, d_p_.as<float>(), weights()));
	return 0;
}

int MpnnNet::read_loss(int slot, double* loss)
{
	float v;
	MCHK(hipMemcpyAsync(&v, d_losses_.as<float>() + slot, sizeof(float), hipMemcpyDeviceToHost, stream_));
	MCHK(hipStreamSynchronize(stream_));
	*loss = v;
	return 0;
}

// epochs over graphs [g_begin, g_end) in batches of consecutive graphs
// (at most batch_graphs graphs and the work-buffer rows), batch order
// shuffled every epoch; first / last = mean batch loss of the first and last epoch
int MpnnNet::train(int g_begin, int g_end, int epochs, int batch_graphs, float lr, unsigned seed, double* first,
                   double* last)
{
	if (g_begin < 0 || g_end > graphs() || g_begin >= g_end || epochs <= 0 || batch_graphs <= 0)
		return fail(-1, "network: bad training range");
	std::vector<std::pair<int, int>> batches;
	for (int g = g_begin; g < g_end;) {
		int h = g;
		long long rows = 0;
		while (h < g_end && h - g < batch_graphs && rows + g_n_[h] <= max_rows_) rows += g_n_[h++];
		if (h == g) return fail(-1, "network: a graph exceeds the work buffers");
		batches.push_back({g, h});
		g = h;
	}
	const int nbatch = (int) batches.size();
	if (nbatch > 65536 / 2) return fail(-1, "network: too many batches per epoch");
	std::mt19937 rng(seed);
	std::vector<float> lh(nbatch);
	for (int ep = 0; ep < epochs; ep++) {
		std::shuffle(batches.begin(), batches.end(), rng);
		// cosine decay of the learning rate over the epochs (from lr down to
		// lr / 2 (1 + cos(pi (epochs - 1) / epochs)) in the last)
		const float lr_ep = (float) (0.5 * lr * (1.0 + std::cos(3.14159265358979323846 * ep / epochs)));
		for (int k = 0; k < nbatch; k++) {
			const MpnnBatch b = batch(batches[k].first, batches[k].second);
			if (int rc = forward(b)) return rc;
			if (int rc = backward(b, k)) return rc;
			if (int rc = adam(lr_ep)) return rc;
		}
		if (ep == 0 || ep == epochs - 1) {
			MCHK(hipMemcpyAsync(lh.data(), d_losses_.ptr(), sizeof(float) * nbatch, hipMemcpyDeviceToHost, stream_));
			MCHK(hipStreamSynchronize(stream_));
			const double mean = std::accumulate(lh.begin(), lh.end(), 0.0) / nbatch;
			if (ep == 0 && first) *first = mean;
			if (ep == epochs - 1 && last) *last = mean;
		}
	}
	MCHK(hipStreamSynchronize(stream_));
	return 0;
}

int MpnnNet::evaluate(int g_begin, int g_end, double* loss)
{
	if (g_begin < 0 || g_end > graphs() || g_begin >= g_end) return fail(-1, "network: bad evaluation range");
	double acc = 0.0, wsum = 0.0;
	for (int g = g_begin; g < g_end;) {
		int h = g;
		long long rows = 0;
		while (h < g_end && rows + g_n_[h] <= max_rows_) rows += g_n_[h++];
		if (h == g) return fail(-1, "network: a graph exceeds the work buffers");
		const MpnnBatch b = batch(g, h);
		if (int rc = forward(b)) return rc;
		int nb = 0;
		LCHK(launch_mpnn_loss(stream_, b.n, d_y_.as<float>(), t_at(b.row0), mask_at(b.row0), b.msum, d_dy_.as<float>(),
		                      d_lpart_.as<float>(), &nb));
		LCHK(launch_mpnn_reduce(stream_, nb, 2, d_lpart_.as<float>(), d_part_.as<float>(), 0));
		float v;
		MCHK(hipMemcpyAsync(&v, d_part_.ptr(), sizeof(float), hipMemcpyDeviceToHost, stream_));
		MCHK(hipStreamSynchronize(stream_));
		acc += (double) v * b.msum;
		wsum += b.msum;
		g = h;
	}
	*loss = acc / std::max(1.0, wsum);
	return 0;
}

}  // namespace tsg
