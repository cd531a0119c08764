// Block-shape enumeration for the exhaustive sweep (hip/primes.h:5-8).
#pragma once
#include <array>
#include <vector>

namespace tcsg {
// Divisors of n in the sweep order: mixed-radix product over the prime
// powers of n in ascending prime order, earlier primes varying fastest
// (hip/primes.cpp:26-58).
std::vector<unsigned> divisors_in_sweep_order(unsigned n);
// (x, n/x) for x in divisors_in_sweep_order(n) (hip/primes.cpp:61-71).
std::vector<std::array<unsigned, 2>> shapes2d(unsigned n);
// (x, y, n/(x*y)) for x in divisors(n), y in divisors(n/x) (hip/primes.cpp:74-88).
std::vector<std::array<unsigned, 3>> shapes3d(unsigned n);
}
