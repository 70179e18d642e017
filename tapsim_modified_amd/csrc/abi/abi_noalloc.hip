// The C ABI's entry point that allocates no device memory: field_o1 of a
// node list with its staging (node indices, fields, status word) in a
// file-scope device array instead of the grid's own buffers.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents; calls that overlap on one GPU (two host threads, two grids on
// the same device) race on it and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../../include/tapsim_gpu.h"
#include "../host/grid.hpp"
#include "errors.hpp"

namespace {

// staging of one call: the status word, then count node indices, then 3
// count doubles, each region 256-byte aligned
constexpr size_t kStatusBytes = 256;
constexpr size_t kNodesBytes = sizeof(int) * (size_t) TSG_NOALLOC_MAX_NODES;
constexpr size_t kFieldsBytes = sizeof(double) * 3 * (size_t) TSG_NOALLOC_MAX_NODES;
constexpr size_t kScratchBytes = kStatusBytes + kNodesBytes + kFieldsBytes;
static_assert(kNodesBytes % 256 == 0, "node region alignment");

// host code never uses this name as an address: the device address of each
// GPU's copy comes from hipGetSymbolAddress below
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// the device address of `scratch` on `device`, looked up once per device
// with that device current (the caller's current device is restored)
int scratch_on(int device, unsigned char** out)
{
	if (device < 0 || device >= kMaxDevices) return tsg::fail(TSG_E_ARG, "tsg_field_o1_noalloc: device index out of range");
	std::lock_guard<std::mutex> lock(g_mu);
	if (!g_ptr[device]) {
		int cur = 0;
		hipError_t e = hipGetDevice(&cur);
		if (e != hipSuccess) return tsg::fail(TSG_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
		if (cur != device && (e = hipSetDevice(device)) != hipSuccess)
			return tsg::fail(TSG_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
		void* p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (cur != device) (void) hipSetDevice(cur);
		if (e != hipSuccess || !p)
			return tsg::fail(TSG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
		g_ptr[device] = p;
	}
	*out = (unsigned char*) g_ptr[device];
	return TSG_OK;
}

}  // namespace

extern "C" int tsg_field_o1_noalloc(tsg_grid* h, int count, const int* nodes, double* out)
{
	if (!h || (count > 0 && (!nodes || !out))) return tsg::fail(TSG_E_ARG, "tsg_field_o1_noalloc: null argument");
	if (count > TSG_NOALLOC_MAX_NODES)
		return tsg::fail(TSG_E_ARG, "tsg_field_o1_noalloc: " + std::to_string(count) + " nodes exceed the staging limit of " +
		                                std::to_string(TSG_NOALLOC_MAX_NODES));
	if (count <= 0) return TSG_OK;
	tsg::Grid* g = tsg_grid_unwrap(h);
	unsigned char* base = nullptr;
	if (int rc = scratch_on(g->device(), &base)) return rc;
	int* d_status = (int*) base;
	int* d_nodes = (int*) (base + kStatusBytes);
	double* d_fields = (double*) (base + kStatusBytes + kNodesBytes);
	const int rc = g->field_o1_staged(count, nodes, out, d_nodes, d_fields, d_status);
	if (rc) tsg::set_error(g->error());
	return rc;
}
