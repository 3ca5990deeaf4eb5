// halton_tables.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.c): the Halton
// sampler's digit permutations as LiAR's Halton::initialize draws them
// (src/liar/samplers/halton.cpp:102-118): the 284 primes after 13, in order,
// each 0 .. p-1 shuffled by std::shuffle with one default-seeded
// std::mt19937.  (C++ because the permutations are those of the C++
// standard library's mt19937 and shuffle.)
#include <algorithm>
#include <cstdint>
#include <random>
#include <vector>

extern "C" {

// off: 285 offsets (off[k + 1] - off[k] = prime k); perm: off[284] entries
// (pass perm == NULL to get the offsets only).  Returns the entry count.
uint32_t or_halton_tables(uint16_t* perm, uint32_t* off) {
    std::vector<uint32_t> primes;
    for (uint32_t n = 17; primes.size() < 284; ++n) {
        bool prime = true;
        for (uint32_t d = 2; d * d <= n; ++d)
            if (n % d == 0) {
                prime = false;
                break;
            }
        if (prime) primes.push_back(n);
    }
    std::mt19937 rng;
    std::vector<size_t> v;
    uint32_t at = 0;
    if (off) off[0] = 0;
    for (size_t k = 0; k < primes.size(); ++k) {
        const uint32_t p = primes[k];
        v.resize(p);
        for (uint32_t i = 0; i < p; ++i) v[i] = i;
        std::shuffle(v.begin(), v.end(), rng);
        for (uint32_t i = 0; i < p; ++i)
            if (perm) perm[at + i] = (uint16_t)v[i];
        at += p;
        if (off) off[k + 1] = at;
    }
    return at;
}

}  // extern "C"
