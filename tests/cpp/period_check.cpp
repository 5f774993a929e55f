// Host-side check of leco::period_quotient (leco_period.h), the period index
// filter_range_close's scan computes by one FP64 multiply and a remainder
// correction, against integer division n / base.  Cases: every q in
// [0, 2^16) at n = q*base - 1, q*base, q*base + 1 for the edge bases 1, 2, 3,
// 2^31 - 2 and 2^31 - 1, the same around exact multiples for seeded random
// bases, and seeded random n below 2^16 * base.  Prints one summary line;
// exit status 0 iff no mismatch.  Built with hipcc (host code only runs).
#include <cstdio>
#include <cstdlib>
#include <random>

#include "leco_period.h"

static uint64_t g_cases = 0, g_bad = 0;
static const uint64_t kMaxPeriods = 1ull << 16;

static void check(uint64_t n, uint64_t base) {
    if (n >= kMaxPeriods * base) return;  // outside the domain the scan uses
    ++g_cases;
    const uint64_t want = n / base;
    const uint32_t got = leco::period_quotient(n, base, leco::period_inverse(base));
    if (got != want) {
        if (g_bad < 10) std::printf("mismatch n=%llu base=%llu want=%llu got=%u\n", (unsigned long long)n,
                                    (unsigned long long)base, (unsigned long long)want, got);
        ++g_bad;
    }
}

static void around_multiples(uint64_t base, uint64_t qstep) {
    for (uint64_t q = 0; q < kMaxPeriods; q += qstep) {
        const uint64_t n = q * base;
        if (n) check(n - 1, base);
        check(n, base);
        check(n + 1, base);
    }
}

int main(int argc, char **argv) {
    const uint64_t randoms = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 2000000;
    for (uint64_t base : {1ull, 2ull, 3ull, (1ull << 31) - 2, (1ull << 31) - 1}) around_multiples(base, 1);
    std::mt19937_64 rng(20240611);
    std::uniform_int_distribution<uint64_t> bdist(1, (1ull << 31) - 1);
    for (int k = 0; k < 200; ++k) around_multiples(bdist(rng), 61);
    for (uint64_t k = 0; k < randoms; ++k) {
        const uint64_t base = k & 1 ? bdist(rng) : 1 + (rng() & 0xFFF);
        check(rng() % (kMaxPeriods * base), base);
    }
    std::printf("cases=%llu bad=%llu\n", (unsigned long long)g_cases, (unsigned long long)g_bad);
    return g_bad ? 1 : 0;
}
