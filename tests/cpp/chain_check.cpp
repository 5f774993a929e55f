// Host-side check of the segmented evaluation of the decoder's additive
// prediction chain (leco_chain.h) against the chain itself, p_{k+1} = p_k + t
// in IEEE double, on random and adversarial (p0, t, m): ties at half an ulp,
// binade boundaries, zero crossings, tiny and huge steps, non-finite values.
// Every value of every block must match bit for bit.  Prints one summary line;
// exit status 0 iff no mismatch.  Built with hipcc (host code only runs).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "leco_chain.h"

using leco::ChainSeg;

static bool same(double a, double b) {
    if (std::isnan(a) && std::isnan(b)) return true;
    return leco::chain_bits(a) == leco::chain_bits(b);
}

static uint64_t g_blocks = 0, g_values = 0, g_segments = 0, g_bad = 0;

static void check(double p0, double t, uint32_t m) {
    std::vector<double> ref(m);
    double p = p0;
    for (uint32_t k = 0; k < m; ++k) {
        ref[k] = p;
        p = p + t;
    }
    uint32_t s = 0;
    double q = p0;
    std::vector<uint32_t> probes;  // segment edges, for chain_at below
    while (s < m) {
        const ChainSeg g = leco::chain_segment(q, t, s, m - s);
        if (g.len < 1 || g.s != s || s + g.len > m) {
            if (g_bad++ < 10) std::printf("bad segment p0=%a t=%a m=%u s=%u len=%u\n", p0, t, m, s, g.len);
            return;
        }
        for (uint32_t j = s; j < s + g.len; ++j) {
            const double v = leco::chain_value(g, j);
            if (!same(v, ref[j])) {
                if (g_bad++ < 10)
                    std::printf("mismatch p0=%a t=%a m=%u j=%u seg(s=%u len=%u p=%a r=%a): %a vs %a\n", p0, t, m, j,
                                g.s, g.len, g.p, g.r, v, ref[j]);
                return;
            }
        }
        q = leco::chain_after(g, t);
        if (probes.size() < 64) {
            probes.push_back(s + g.len - 1);
            if (s + g.len < m) probes.push_back(s + g.len);
        }
        s += g.len;
        ++g_segments;
    }
    // point queries through chain_at: segment edges and random positions
    for (int z = 0; z < 8; ++z) probes.push_back((uint32_t)((uint64_t)rand() * m / ((uint64_t)RAND_MAX + 1)));
    for (const uint32_t j : probes) {
        const double v = leco::chain_at(p0, t, j, m);
        if (!same(v, ref[j])) {
            if (g_bad++ < 10) std::printf("chain_at mismatch p0=%a t=%a m=%u j=%u\n", p0, t, m, j);
            return;
        }
    }
    ++g_blocks;
    g_values += m;
}

int main(int argc, char **argv) {
    const int rounds = argc > 1 ? std::atoi(argv[1]) : 20000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto rnd_float = [&](int elo, int ehi) -> double {  // a float of random magnitude 2^[elo, ehi) and sign
        const int e = elo + (int)(rng() % (uint64_t)(ehi - elo));
        float f = (float)std::ldexp(1.0 + U(rng), e);
        if (rng() & 1) f = -f;
        return (double)f;
    };
    auto rnd_m = [&]() -> uint32_t {
        const uint64_t c = rng() % 10;
        if (c < 5) return 1 + (uint32_t)(rng() % 300);
        if (c < 9) return 1 + (uint32_t)(rng() % 5000);
        return 1 + (uint32_t)(rng() % 70000);
    };
    for (int it = 0; it < rounds; ++it) {
        const uint32_t m = rnd_m();
        switch (it % 8) {
        case 0: {  // predictions of uint32 data: |p| up to 2^33, slopes of any size
            const double p0 = std::ldexp(U(rng), 1 + (int)(rng() % 33)) * ((rng() & 3) ? 1 : -1);
            check(p0, rnd_float(-30, 33), m);
            break;
        }
        case 1: {  // tie steps: t/u = a + 1/2 in p0's binade
            const int e = (int)(rng() % 40);
            const double u = std::ldexp(1.0, e - 52);
            const uint64_t Q = (1ull << 52) | (rng() & ((1ull << 52) - 1));
            const double p0 = (double)Q * u * ((rng() & 1) ? 1 : -1);
            const uint64_t a = rng() % (1u << 20);
            const double t = (double)(float)((2.0 * (double)a + 1.0) * (u / 2)) * ((rng() & 1) ? 1 : -1);
            check(p0, t, m);
            break;
        }
        case 2: {  // starts just below / above a power of two, moving across it
            const int e = (int)(rng() % 34);
            const double t = rnd_float(-25, e > 3 ? e - 3 : 1);
            const double p0 = std::ldexp(1.0, e) - t * (double)(rng() % (m + 1));
            check(p0, t, m);
            break;
        }
        case 3: {  // zero crossings
            const double t = std::fabs(rnd_float(-20, 10));
            const double p0 = -t * ((double)(rng() % (m + 1)) + U(rng));
            check(p0, (rng() & 1) ? t : -t, m);
            check(-p0, (rng() & 1) ? t : -t, m);
            break;
        }
        case 4: {  // tiny steps (often rounded away) and zero
            const double p0 = std::ldexp(U(rng) + 0.5, (int)(rng() % 40)) * ((rng() & 1) ? 1 : -1);
            check(p0, (rng() % 5) ? rnd_float(-149, -20) : 0.0, m);
            break;
        }
        case 5: {  // huge steps relative to p
            check(rnd_float(-10, 5), rnd_float(20, 33), m);
            break;
        }
        case 6: {  // exact-power-of-two p with steps near the ulp
            const int e = (int)(rng() % 40);
            const double u = std::ldexp(1.0, e - 52);
            const double t = (double)(float)(u * (double)(1 + rng() % 7) / 4.0) * ((rng() & 1) ? 1 : -1);
            check(std::ldexp(1.0, e) * ((rng() & 1) ? 1 : -1), t, m);
            break;
        }
        default: {  // special values
            const double sp[] = {0.0, -0.0, NAN, INFINITY, -INFINITY, 4.9e-324, -4.9e-324, 2.2250738585072014e-308};
            check(sp[rng() % 8], rnd_float(-30, 30), m);
            break;
        }
        }
    }
    std::printf("chain_check blocks=%llu values=%llu segments=%llu bad=%llu\n", (unsigned long long)g_blocks,
                (unsigned long long)g_values, (unsigned long long)g_segments, (unsigned long long)g_bad);
    return g_bad ? 1 : 0;
}
