// Host side of the learned warm start: the network's parameters and Adam
// state on the device, its training set of shell graphs (one per recorded
// evaporation event) in HBM, and the forward / backward / training passes
// built from the kernels in kernels/mpnn.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../kernels/mpnn.hpp"
#include "grid.hpp"

namespace tsg {

// rows [row0, row0 + n) of the graph store (ptr/col use absolute rows and edges)
struct MpnnBatch {
	int n = 0, row0 = 0;
	float msum = 1.0f;
};

class MpnnNet {
public:
	MpnnNet(int max_batch_rows, long long data_rows, long long data_edges, unsigned seed);
	~MpnnNet();
	int status() const { return status_; }
	const std::string& error() const { return err_; }
	hipStream_t stream() const { return stream_; }

	int set_params(const float* host);
	int get_params(float* host);

	// the store: graphs appended contiguously; the graph being built sits at
	// rows [rows_, ...) and is committed by commit_graph
	int graphs() const { return (int) g_row0_.size(); }
	long long rows() const { return rows_; }
	long long edges() const { return edges_; }
	int clear();
	// device pointers of the next graph's rows / edges (capacity checked)
	int reserve(int n, long long ne);
	int* ptr_at(long long row) const { return d_ptr_.as<int>() + row; }
	int* col_at(long long e) const { return d_col_.as<int>() + e; }
	float* a_at(long long e) const { return d_a_.as<float>() + e; }
	bf16_t* x_at(long long row) const { return d_x_.as<bf16_t>() + row * kMpnnIn; }
	float* t_at(long long row) const { return d_t_.as<float>() + row; }
	float* mask_at(long long row) const { return d_mask_.as<float>() + row; }
	int commit_graph(int n, long long ne, float msum);
	// a host graph (local CSR, fp32 features) copied into the store at its end
	int upload_graph(int n, const int* ptr, const int* col, const float* a, const float* x, const float* t,
	                 const float* mask, bool commit);

	// forward over a batch (on the network's stream, or on `on`); y (device,
	// n values) stays in the work buffers
	int forward(const MpnnBatch& b, hipStream_t on = nullptr);
	const float* y() const { return d_y_.as<float>(); }
	// loss gradient of the last forward (into the device gradient) and the
	// batch loss into d_losses_[slot]
	int backward(const MpnnBatch& b, int slot);
	int adam(float lr);
	int read_loss(int slot, double* loss);
	int read_grad(float* host);

	int train(int g_begin, int g_end, int epochs, int batch_graphs, float lr, unsigned seed, double* first, double* last);
	int evaluate(int g_begin, int g_end, double* loss);
	// batch covering graphs [g0, g1)
	MpnnBatch batch(int g0, int g1) const;

private:
	int fail(int code, const std::string& msg);
	int hip_check(hipError_t e, const char* what);
	int ensure_work(int n);

	int status_ = 0;
	std::string err_;
	hipStream_t stream_ = nullptr;
	int max_rows_ = 0;
	long long cap_rows_ = 0, cap_edges_ = 0;
	long long rows_ = 0, edges_ = 0;
	std::vector<long long> g_row0_, g_edge0_;
	std::vector<int> g_n_;
	std::vector<float> g_msum_;
	int step_ = 0;

	DevBuf d_p_, d_g_, d_m_, d_v_;                  // fp32 parameters, gradient, Adam moments
	DevBuf d_win_t_, d_wl_t_, d_wl_, d_wout_;       // bf16 weight copies
	DevBuf d_ptr_, d_col_, d_a_, d_x_, d_t_, d_mask_;  // graph store
	// work buffers for one batch
	DevBuf d_z_, d_h_, d_cat_, d_dh_, d_dz_, d_dzb_, d_dcat_, d_y_, d_dy_, d_part_, d_lpart_, d_losses_;
	MpnnWeights weights() const;
};

}  // namespace tsg
