// The process's rand() stream, owned by the library so that the device can
// continue it.
//
// The reference draws from the C library's rand() for the Monte Carlo choice
// (surface_3d.cpp:433-450), randomReset (grid_3d.cpp:811-848), the thermal
// start velocity (surface_3d.cpp:472-502), tetGuess (geometry_3d.cpp:1712-1748)
// and the first face of every point-location walk (geometry_3d.cpp:1017),
// seeded with 99 (tapsim.cpp:235).  The walks run on the GPU here, inside the
// trajectory kernels, so the generator's state has to travel with them.
//
// The generator is glibc's default rand() (random_r.c, TYPE_3): a table of 31
// words, seeded by the Park-Miller minimal standard generator (Schrage's
// factorisation), with 310 outputs discarded; each draw adds the word three
// places ahead into the current one (mod 2^32) and returns it shifted right by
// one.  For equal seeds this class and glibc's rand() return the same
// sequence (checked in the tests).
#pragma once

#include <cstdint>

namespace tsg {

struct RandState {
	int32_t s[31];
	int32_t r;  // index of the word the next draw reads and replaces is (r + 3) % 31
};

inline void rand_seed(RandState* st, unsigned int seed)
{
	int32_t word = (int32_t) (seed == 0 ? 1u : seed);
	st->s[0] = word;
	for (int i = 1; i < 31; i++) {
		const int32_t hi = word / 127773, lo = word % 127773;
		word = 16807 * lo - 2836 * hi;
		if (word < 0) word += 2147483647;
		st->s[i] = word;
	}
	st->r = 0;
	for (int i = 0; i < 310; i++) {
		const int f = (st->r + 3) % 31;
		st->s[f] = (int32_t) ((uint32_t) st->s[f] + (uint32_t) st->s[st->r]);
		st->r = (st->r + 1) % 31;
	}
}

inline int rand_next(RandState* st)
{
	const int f = (st->r + 3) % 31;
	const uint32_t v = (uint32_t) st->s[f] + (uint32_t) st->s[st->r];
	st->s[f] = (int32_t) v;
	st->r = (st->r + 1) % 31;
	return (int) (v >> 1);
}

// the process-wide stream (like the C library's, one per process)
RandState& process_rand();
inline int prand() { return rand_next(&process_rand()); }
constexpr int kRandMax = 2147483647;

}  // namespace tsg
