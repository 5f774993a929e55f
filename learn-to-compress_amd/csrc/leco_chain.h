// The additive prediction chain of piecewise_fix_op_max's decoder
// (bit_read.h:559-604 read_all_bit_fix_add, and the l = 0 loop of
// piecewise_fix_op_minimize_maxerror.h:158-163):
//     p_0 = theta0,  p_{k+1} = fl(p_k + t)        (IEEE double, round to nearest even)
// evaluated at any index without replaying the k additions one by one.
//
// Inside one binade [2^e, 2^(e+1)) every double is a multiple of u = 2^(e-52),
// so while the exact sum p_k + t stays in that binade, fl(p_k + t) = p_k + R*u
// with R = t/u rounded to an integer.  R is the same for every k unless t/u is
// an exact tie; then it depends on the parity of p_k/u, and once one step has
// made p_k/u even (R even) it stays the same as well.  So the chain is a
// union of segments p_{s+k} = p_s + k*r, each one exact (every term is a
// multiple of u of at most 53 bits, so fma(k, r, p_s) has no rounding), and
// a segment ends only where the chain leaves a binade (a few times per block
// at most) or where a tie step is not yet settled (one step).
//
// chain_segment() returns the segment starting at p_s; it proves the segment
// with integer arithmetic on p_s's significand (Q = p_s/u, 2^52 <= Q < 2^53)
// and falls back to a one-value segment wherever it cannot (zero, subnormals,
// steps larger than the binade), which is always correct.  A non-finite p
// stays non-finite, so one constant segment covers the rest.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace leco {

constexpr int64_t kChainLo = 1ll << 52, kChainHi = (1ll << 53) - 1;

struct ChainSeg {
    double p;      // value at index s
    double r;      // step inside the segment
    uint32_t s;    // first index of the segment
    uint32_t len;  // values p + k*r, 0 <= k < len (exact)
};

__host__ __device__ inline uint64_t chain_bits(double d) {
    uint64_t b;
    __builtin_memcpy(&b, &d, 8);
    return b;
}

// What chain_segment needs to know about the steps from p: r = fl(p + t) - p
// and, when the lattice argument applies (ok), A_0 = Q + floor(T) and the
// step R in units of u (see above).
struct ChainProbe {
    double r;
    int64_t a0, ri;
    bool ok;
};

__host__ __device__ inline ChainProbe chain_probe(double p, double t) {
    ChainProbe c{0.0, 0, 0, false};
    const double p1 = p + t, p2 = p1 + t;
    const double r = p1 - p;
    c.r = r;
    if (!(r == p2 - p1) || p == 0.0) return c;
    // in the frame where p > 0 (RNE is symmetric: the chain of (-p, -t) is the negated chain)
    const bool neg = p < 0.0;
    const double ap = neg ? -p : p, at = neg ? -t : t, ar = neg ? -r : r;
    const uint64_t pb = chain_bits(ap);
    const int E = (int)(pb >> 52);
    if (E == 0) return c;                         // subnormal p
    const int sc = 1075 - E;                      // u = 2^-sc
    const double T = __builtin_ldexp(at, sc);     // t / u
    const double R = __builtin_ldexp(ar, sc);     // r / u
    if (!(__builtin_fabs(T) < 0x1p53) || __builtin_ldexp(T, -sc) != at || __builtin_ldexp(R, -sc) != ar ||
        R != __builtin_trunc(R))
        return c;
    // step k stays in the binade iff A_k = Q + k*R + floor(T) lies in [2^52, 2^53 - 1]
    // (then p_{s+k} + t = (A_k + frac(T)) * u with A_k + frac(T) in [2^52, 2^53))
    const int64_t Q = (int64_t)((pb & 0xFFFFFFFFFFFFFull) | (1ull << 52));
    c.a0 = Q + (int64_t)__builtin_floor(T);
    c.ri = (int64_t)R;
    c.ok = c.a0 >= kChainLo && c.a0 <= kChainHi;
    return c;
}

// Valid steps k = 0 .. steps-1 of an ok probe (at least 1; saturates at `cap`).
__host__ __device__ inline uint64_t chain_steps(const ChainProbe &c, uint64_t cap) {
    if (c.ri > 0) return (uint64_t)(kChainHi - c.a0) / (uint64_t)c.ri + 1;
    if (c.ri < 0) return (uint64_t)(c.a0 - kChainLo) / (uint64_t)(-c.ri) + 1;
    return cap;
}

// Does an ok probe's segment reach k steps (k >= 1), i.e. is step k-1 valid?
// (A_k is monotone in k, so the valid steps are a prefix.)  Exact without a
// division while |R| * k stays far from 2^63.
__host__ __device__ inline bool chain_reaches(const ChainProbe &c, uint32_t k) {
    const int64_t mag = c.ri < 0 ? -c.ri : c.ri;
    if (mag < (1ll << 40)) {
        const int64_t a = c.a0 + (int64_t)(k - 1) * c.ri;
        return a >= kChainLo && a <= kChainHi;
    }
    return (uint64_t)k <= chain_steps(c, ~0ull);
}

// The segment that starts at index s with value p; rem >= 1 values are left
// (indices s .. s+rem-1).
__host__ __device__ inline ChainSeg chain_segment(double p, double t, uint32_t s, uint32_t rem) {
    ChainSeg g{p, 0.0, s, 1};
    if (rem <= 1) return g;
    if (!(__builtin_fabs(p) <= 1.7976931348623157e308)) {  // NaN or infinity: p + t == p (up to NaN payload)
        g.len = rem;
        return g;
    }
    const ChainProbe c = chain_probe(p, t);
    g.r = c.r;
    if (!c.ok) return g;
    const uint64_t steps = chain_steps(c, rem);
    g.len = steps + 1 < (uint64_t)rem ? (uint32_t)(steps + 1) : rem;
    return g;
}

__host__ __device__ inline double chain_value(const ChainSeg &g, uint32_t j) {
    return j == g.s ? g.p : __builtin_fma((double)(j - g.s), g.r, g.p);  // (p_s itself keeps a -0 sign)
}

// Value at the index right after the segment: one step of the chain from its last value.
__host__ __device__ inline double chain_after(const ChainSeg &g, double t) {
    return chain_value(g, g.s + g.len - 1) + t;
}

// p_j of the chain p_0 = p0 over a block of m values (j < m).  Usually one
// probe and one fma: a division only where the chain leaves a binade before j.
__host__ __device__ inline double chain_at(double p0, double t, uint32_t j, uint32_t m) {
    uint32_t s = 0;
    double p = p0;
    for (;;) {
        if (j == s) return p;
        if (!(__builtin_fabs(p) <= 1.7976931348623157e308)) return p;  // non-finite: constant (up to NaN payload)
        const ChainProbe c = chain_probe(p, t);
        const uint32_t k = j - s;
        if (c.ok && chain_reaches(c, k)) return __builtin_fma((double)k, c.r, p);
        // leave the segment: its last value, then one step of the chain
        uint32_t len = 1;
        if (c.ok) {
            const uint64_t steps = chain_steps(c, m - s);
            len = steps + 1 < (uint64_t)(m - s) ? (uint32_t)(steps + 1) : m - s;
        }
        p = (len > 1 ? __builtin_fma((double)(len - 1), c.r, p) : p) + t;
        s += len;
    }
}

}  // namespace leco
