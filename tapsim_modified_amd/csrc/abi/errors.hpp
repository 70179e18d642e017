// Shared pieces of the C ABI translation units: thread-local error text and
// the opaque mesh handle.
#pragma once
#include <string>

#include "../host/mesh.hpp"

struct tsg_mesh {
	tsg::Mesh m;
};

namespace tsg {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
}  // namespace tsg
