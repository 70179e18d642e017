// C ABI over the warm-start network (include/tapsim_gpu.h, TSG_MPNN_*).
#include <memory>
#include <string>
#include <vector>

#include "../../../include/tapsim_gpu.h"
#include "../host/mpnn.hpp"
#include "errors.hpp"

struct tsg_mpnn {
	std::unique_ptr<tsg::MpnnNet> net;
};

namespace {
static_assert(TSG_MPNN_NPARAM == tsg::kMpnnNParam, "parameter layout");
static_assert(TSG_MPNN_IN == tsg::kMpnnIn && TSG_MPNN_HID == tsg::kMpnnHid, "network sizes");

int check(tsg_mpnn* h, int rc)
{
	if (rc) tsg::set_error(h->net->error());
	return rc;
}
}  // namespace

extern "C" {

int tsg_mpnn_create(int max_batch_rows, long long data_rows, long long data_edges, unsigned seed, tsg_mpnn** out)
{
	if (!out || max_batch_rows <= 0 || data_rows <= 0 || data_edges <= 0)
		return tsg::fail(TSG_E_ARG, "tsg_mpnn_create: bad arguments");
	int dev = 0;
	if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) return tsg::fail(TSG_E_HIP, "tsg_mpnn_create: no HIP device");
	try {
		auto* h = new tsg_mpnn();
		h->net.reset(new tsg::MpnnNet(max_batch_rows, data_rows, data_edges, seed));
		if (h->net->status()) {
			const int rc = h->net->status();
			tsg::set_error(h->net->error());
			delete h;
			return rc;
		}
		*out = h;
		return TSG_OK;
	} catch (const std::exception& e) {
		return tsg::fail(TSG_E_HIP, std::string("tsg_mpnn_create: ") + e.what());
	}
}

void tsg_mpnn_destroy(tsg_mpnn* h) { delete h; }

int tsg_mpnn_set_params(tsg_mpnn* h, const float* params)
{
	if (!h || !params) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->net->set_params(params));
}

int tsg_mpnn_get_params(tsg_mpnn* h, float* params)
{
	if (!h || !params) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->net->get_params(params));
}

int tsg_mpnn_eval(tsg_mpnn* h, int n, const int* ptr, const int* col, const float* a, const float* x, const float* t,
                  const float* mask, double* loss, float* y, float* grad)
{
	if (!h || n <= 0 || !ptr || !x || !t || !mask || (ptr[n] > 0 && (!col || !a)))
		return tsg::fail(TSG_E_ARG, "tsg_mpnn_eval: bad arguments");
	tsg::MpnnNet& net = *h->net;
	const long long row0 = net.rows();
	if (int rc = check(h, net.upload_graph(n, ptr, col, a, x, t, mask, false))) return rc;
	float msum = 0.0f;
	for (int i = 0; i < n; i++) msum += mask[i];
	tsg::MpnnBatch b;
	b.n = n;
	b.row0 = (int) row0;
	b.msum = msum > 1.0f ? msum : 1.0f;
	if (int rc = check(h, net.forward(b))) return rc;
	if (loss || grad) {
		if (int rc = check(h, net.backward(b, 0))) return rc;
		if (loss && check(h, net.read_loss(0, loss))) return net.status();
		if (grad && check(h, net.read_grad(grad))) return net.status();
	}
	if (y) {
		if (hipMemcpy(y, net.y(), sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
			return tsg::fail(TSG_E_HIP, "tsg_mpnn_eval: copy of y failed");
	}
	return TSG_OK;
}

int tsg_mpnn_add_graph(tsg_mpnn* h, int n, const int* ptr, const int* col, const float* a, const float* x, const float* t,
                       const float* mask)
{
	if (!h || n <= 0 || !ptr || !x || !t || !mask || (ptr[n] > 0 && (!col || !a)))
		return tsg::fail(TSG_E_ARG, "tsg_mpnn_add_graph: bad arguments");
	return check(h, h->net->upload_graph(n, ptr, col, a, x, t, mask, true));
}

int tsg_mpnn_dataset(const tsg_mpnn* h, int* graphs, long long* rows, long long* edges)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null argument");
	if (graphs) *graphs = h->net->graphs();
	if (rows) *rows = h->net->rows();
	if (edges) *edges = h->net->edges();
	return TSG_OK;
}

int tsg_mpnn_clear(tsg_mpnn* h)
{
	if (!h) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->net->clear());
}

int tsg_mpnn_train(tsg_mpnn* h, int g_begin, int g_end, int epochs, int batch_graphs, float lr, unsigned seed,
                   double* first_loss, double* last_loss)
{
	if (!h || !(lr > 0.0f)) return tsg::fail(TSG_E_ARG, "tsg_mpnn_train: bad arguments");
	return check(h, h->net->train(g_begin, g_end, epochs, batch_graphs, lr, seed, first_loss, last_loss));
}

int tsg_grid_attach_mpnn(tsg_grid* g, tsg_mpnn* h, int record, int predict)
{
	if (!g) return tsg::fail(TSG_E_ARG, "null grid");
	tsg::Grid& grid = *tsg_grid_unwrap(g);
	const int rc = grid.attach_mpnn(h ? h->net.get() : nullptr, record != 0, predict != 0);
	if (rc) tsg::set_error(grid.error());
	return rc;
}

int tsg_mpnn_evaluate(tsg_mpnn* h, int g_begin, int g_end, double* loss)
{
	if (!h || !loss) return tsg::fail(TSG_E_ARG, "null argument");
	return check(h, h->net->evaluate(g_begin, g_end, loss));
}

}  // extern "C"
