// Compares qla_amd/csrc/qla_hostmath.hpp (compiled for the host) with the
// host C library: acosf on [-1, 1], cbrtf on all floats, sincosf on
// |x| < 120, every STRIDE-th float bit pattern plus boundary values.
// Prints "acosf N cbrtf N sincosf N checked M" (N = mismatches).
#include "qla_hostmath.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>

using namespace qla;

static long mis_a = 0, mis_c = 0, mis_s = 0, checked = 0;

static void check(float x, bool sc)
{
    checked++;
    if (std::fabs(x) <= 1.0f && hm_bits(hm_acosf(x)) != hm_bits(acosf(x))) {
        if (mis_a++ < 5) printf("acosf(%a): %a, library %a\n", x, hm_acosf(x), acosf(x));
    }
    if (hm_bits(hm_cbrtf(x)) != hm_bits(cbrtf(x))) {
        if (mis_c++ < 5) printf("cbrtf(%a): %a, library %a\n", x, hm_cbrtf(x), cbrtf(x));
    }
    float s, c, ls, lc;
    if (sc && hm_sincosf(x, &s, &c)) {
        sincosf(x, &ls, &lc);
        if (hm_bits(s) != hm_bits(ls) || hm_bits(c) != hm_bits(lc)) {
            if (mis_s++ < 5) printf("sincosf(%a): %a %a, library %a %a\n", x, s, c, ls, lc);
        }
    }
}

int main(int argc, char **argv)
{
    const unsigned long stride = argc > 1 ? strtoul(argv[1], nullptr, 10) : 61;
    const bool sc = argc > 2 ? atoi(argv[2]) != 0 : true;
    for (unsigned long u = 0; u < 0x100000000ul; u += stride) {
        const float x = hm_float((uint32_t)u);
        if (!std::isnan(x)) check(x, sc);
    }
    const float special[] = { 0.0f, -0.0f, 1.0f, -1.0f, 0.5f, -0.5f, 0x1.fffffep-2f, -0x1.fffffep-2f,
                              0x1p-57f, 0x1p-58f, 0x1.921fb6p-1f, 0x1.921fb4p-1f, -0x1.921fb6p-1f, 0x1p-12f,
                              0x1.fffffep-13f, 119.99999f, 120.0f, -119.99999f, 0x1p-126f, 0x1p-149f,
                              3.0f, 8.0f, 27.0f, -27.0f, 1e-40f, 3.4e38f, INFINITY, -INFINITY };
    for (float x : special) check(x, sc);
    printf("acosf %ld cbrtf %ld sincosf %ld checked %ld\n", mis_a, mis_c, mis_s, checked);
    return 0;
}
