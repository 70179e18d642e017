// Shared pieces of the C ABI translation units: thread-local error text and
// the opaque mesh handle.
#pragma once
#include <string>

#include "../host/mesh.hpp"

struct tsg_mesh {
	tsg::Mesh m;
};

namespace tsg {
class Grid;
}
struct tsg_grid;
// the Grid behind an opaque grid handle (defined with the grid ABI)
tsg::Grid* tsg_grid_unwrap(tsg_grid*);

namespace tsg {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
}  // namespace tsg
