// The one-reduction sums of products of the field header (dot4p, dot2)
// against 128-bit arithmetic and against the two-product form, on random
// and edge-value operands (host code; the device runs the same functions).
#include "../../virgo_amd/csrc/vg_field.hpp"
#include <cstdio>
#include <random>
using namespace vg;
static uint64_t ref_mod(unsigned __int128 x) { return (uint64_t)(x % P); }
int main()
{
    std::mt19937_64 g(12345);
    const uint64_t edge[] = {0, 1, 2, P - 1, P - 2, (1ull << 31) - 1, 1ull << 31, 1ull << 60, P - (1ull << 31),
                             0x7FFFFFFFull | (0x3FFFFFFFull << 31), (0x3FFFFFFFull << 31)};
    const int ne = sizeof edge / sizeof edge[0];
    long bad = 0, n = 0;
    for (long it = 0; it < 2000000; ++it) {
        uint64_t v[8];
        for (int k = 0; k < 8; ++k) v[k] = (it < 200000 || g() % 3 == 0) ? edge[g() % ne] : g() % P;
        unsigned __int128 s = 0;
        for (int k = 0; k < 4; ++k) s += (unsigned __int128)v[2 * k] * v[2 * k + 1];
        uint64_t r = dot4p(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
        if (r != ref_mod(s) || r >= P) ++bad;
        fe x = mk(v[0], v[1]), y = mk(v[2], v[3]), z = mk(v[4], v[5]), w = mk(v[6], v[7]);
        if (!feq(dot2(x, y, z, w), add(mul(x, y), mul(z, w)))) ++bad;
        // a 2^k = rotation, for canonical a (the products above also take p
        // itself as an operand, as mul does for p - 0)
        const uint64_t c = v[0] % P;
        for (int k = 0; k < 61; k += 7)
            if (rotp(c, k) != ref_mod((unsigned __int128)c << k)) ++bad;
        fe xc = mk(c, v[1] % P);
        if (!feq(halve(add(xc, xc)), xc)) ++bad;
        ++n;
    }
    const uint64_t m = P - 1;
    if (dot4p(m, m, m, m, m, m, m, m) != ref_mod(4 * ((unsigned __int128)m * m))) ++bad;
    printf("cases %ld mismatches %ld\n", n, bad);
    return bad != 0;
}
