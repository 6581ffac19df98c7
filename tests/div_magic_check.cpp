// Checks div_magic (mcx_amd/csrc/mcx_kernel.h) against integer division:
// every divisor up to 20000 and a spread of large ones, each against the
// edge numerators and random ones.  Built and run by test_div_magic.py.
#include <cstdint>
#include <cstdio>
#include <random>
#include "mcx_kernel.h"

static uint32_t divide(uint32_t n, uint32_t m, uint32_t sh) {
    const uint32_t t = (uint32_t)(((uint64_t)n * m) >> 32);
    return (t + ((n - t) >> (sh & 0xffu))) >> (sh >> 8);
}

int main() {
    std::mt19937_64 g(12345);
    unsigned long long n = 0, bad = 0;
    auto check = [&](uint32_t d) {
        uint32_t m, sh;
        div_magic(d, &m, &sh);
        const uint32_t edge[] = {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu};
        for (uint32_t x : edge) n++, bad += divide(x, m, sh) != x / d;
        for (int k = 0; k < 200; k++) {
            const uint32_t x = (uint32_t)g();
            n++, bad += divide(x, m, sh) != x / d;
            const uint32_t y = (x / d) * d;
            if (y) n++, bad += divide(y - 1, m, sh) != (y - 1) / d;
        }
    };
    for (uint32_t d = 1; d <= 20000; d++) check(d);
    for (int k = 0; k < 5000; k++) check(((uint32_t)g() >> (k % 32)) | 1u);
    const uint32_t special[] = {65536u, 65537u, 1u << 24, 0x80000000u, 0x80000001u, 0xfffffffeu, 0xffffffffu};
    for (uint32_t d : special) check(d);
    printf("%llu %llu\n", n, bad);
    return bad != 0;
}
