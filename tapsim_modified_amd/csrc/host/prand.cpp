// The process-wide rand() stream of the library (see prand.hpp).
#include "prand.hpp"

namespace tsg {

RandState& process_rand()
{
	// unseeded, the C library's rand() behaves as after srand(1)
	static RandState st = [] {
		RandState s;
		rand_seed(&s, 1u);
		return s;
	}();
	return st;
}

}  // namespace tsg
