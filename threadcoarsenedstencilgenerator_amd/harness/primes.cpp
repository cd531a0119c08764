#include "primes.h"

namespace tcsg {

std::vector<unsigned> divisors_in_sweep_order(unsigned n)
{
    std::vector<unsigned> out{1};
    unsigned rest = n;
    for (unsigned p = 2; p * p <= rest || (p <= rest && rest > 1); ++p) {
        if (p * p > rest) p = rest;  // remaining factor is prime
        unsigned mult = 0;
        while (rest % p == 0) { rest /= p; ++mult; }
        if (!mult) continue;
        std::vector<unsigned> next;
        next.reserve(out.size() * (mult + 1));
        unsigned f = 1;
        for (unsigned e = 0; e <= mult; ++e, f *= p)
            for (unsigned d : out) next.push_back(d * f);
        out.swap(next);
        if (rest == 1) break;
    }
    return out;
}

std::vector<std::array<unsigned, 2>> shapes2d(unsigned n)
{
    std::vector<std::array<unsigned, 2>> s;
    for (unsigned x : divisors_in_sweep_order(n)) s.push_back({x, n / x});
    return s;
}

std::vector<std::array<unsigned, 3>> shapes3d(unsigned n)
{
    std::vector<std::array<unsigned, 3>> s;
    for (unsigned x : divisors_in_sweep_order(n))
        for (unsigned y : divisors_in_sweep_order(n / x)) s.push_back({x, y, n / x / y});
    return s;
}

} // namespace tcsg
