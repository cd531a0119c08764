// Test-infrastructure shim: exposes the reference's own block-shape
// enumeration (hip/primes.cpp, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/) through a C ABI so the tests can compare
// the oracle's sweep order with it.  Nothing in the product links this.
#include "primes.h"

extern "C" int ref_prime_combinations(unsigned number, unsigned *out, int cap)
{
    std::vector<unsigned int> v = primeCombinations(number);
    int n = 0;
    for (unsigned d : v) { if (n < cap) out[n] = d; ++n; }
    return n;
}

extern "C" int ref_prime_combinations3d(unsigned number, unsigned *out, int cap)
{
    auto v = primeCombinations3D(number);
    int n = 0;
    for (auto &t : v) {
        if (n < cap) { out[3 * n] = std::get<0>(t); out[3 * n + 1] = std::get<1>(t); out[3 * n + 2] = std::get<2>(t); }
        ++n;
    }
    return n;
}

extern "C" int ref_prime_combinations2d(unsigned number, unsigned *out, int cap)
{
    auto v = primeCombinations2D(number);
    int n = 0;
    for (auto &p : v) {
        if (n < cap) { out[2 * n] = p.first; out[2 * n + 1] = p.second; }
        ++n;
    }
    return n;
}
