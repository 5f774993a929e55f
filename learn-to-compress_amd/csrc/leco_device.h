// Device helpers shared by the gfx950 LeCo_fix kernels.
//
// Numerics: every kernel is compiled with -ffp-contract=off, so the only fused
// multiply-adds are the explicit __builtin_fma calls below.  They sit where
// the reference's -O3 -march=native x86 build fuses (lr.h:72-83,
// piecewise_fix_integer_template.h:40,54, bit_read.h:247,1187), which makes
// theta and every prediction bit-identical to the reference CPU path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace leco {

constexpr uint64_t kX86DefaultNaN = 0xFFF8000000000000ull;
constexpr int kWave = 64;

// x86 CVTTSD2SI (r64): truncation toward zero; NaN or out of [-2^63, 2^63)
// gives INT64_MIN.  The reference converts predictions with a plain
// (int64_t)/(long long) cast, which is this instruction on x86.
// Fast path: for an integer t with |t| < 2^51, t + 1.5*2^52 is exact and its
// low mantissa bits hold t + 2^51, so bits(t + 1.5*2^52) - 0x4338000000000000
// is t (two's complement).  That is one add instead of the ~8 ops of the
// generic f64 -> i64 lowering; NaN and large values take the generic path.
__device__ __forceinline__ int64_t cvtt64(double x) {
    const double t = __builtin_trunc(x);
    if (__builtin_expect(__builtin_fabs(t) < 0x1p51, 1))
        return (int64_t)((uint64_t)__double_as_longlong(t + 0x1.8p52) - 0x4338000000000000ull);
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) return INT64_MIN;
    return (int64_t)x;
}

// p(i) = cvtt64(theta0 + theta1 * i) with one rounding (fma).
__device__ __forceinline__ int64_t predict(double t0, double t1, uint32_t i) {
    return cvtt64(__builtin_fma(t1, (double)i, t0));
}

// predict() without the range branch: equal to it when
// |trunc(fma(t1, i, t0))| < 2^51 (see near_range).
__device__ __forceinline__ int64_t predict_near(double t0, double t1, uint32_t i) {
    const double t = __builtin_trunc(__builtin_fma(t1, (double)i, t0));
    return (int64_t)((uint64_t)__double_as_longlong(t + 0x1.8p52) - 0x4338000000000000ull);
}

// True when predict_near(t0, t1, i) == predict(t0, t1, i) for every i < m
// (m >= 1).  The exact t1*i + t0 is monotone in i, and rounding to nearest
// and truncation are monotone, so trunc(fma(t1, i, t0)) lies between its
// values at i = 0 and i = m - 1; NaN or infinite thetas fail the test.
__device__ __forceinline__ bool near_range(double t0, double t1, uint32_t m) {
    const double a = __builtin_trunc(__builtin_fma(t1, 0.0, t0));
    const double b = __builtin_trunc(__builtin_fma(t1, (double)(m - 1), t0));
    return __builtin_fabs(a) < 0x1p51 && __builtin_fabs(b) < 0x1p51;
}

// Low 32 bits of predict(): the fast path's bits(t + 1.5*2^52) already hold
// them (the subtracted 0x4338000000000000 has zero low word).
__device__ __forceinline__ uint32_t predict_lo32(double t0, double t1, uint32_t i) {
    const double x = __builtin_fma(t1, (double)i, t0);
    const double t = __builtin_trunc(x);
    if (__builtin_expect(__builtin_fabs(t) < 0x1p51, 1)) return (uint32_t)__double_as_longlong(t + 0x1.8p52);
    return (uint32_t)cvtt64(x);
}

// Bit length (lr.h:116-130: 64 - clzll(v)).
__device__ __forceinline__ int bitlen64(uint64_t v) { return v ? 64 - __clzll((long long)v) : 0; }

// Sign/magnitude residual of y against prediction p
// (piecewise_fix_integer_template.h:55-64): the comparison is signed int64
// for 32-bit values and unsigned for 64-bit values.
template <int W>
__device__ __forceinline__ uint64_t residual_field(uint64_t y, int64_t p, int l, uint64_t *mag_out) {
    bool pos = (W == 64) ? (y > (uint64_t)p) : ((int64_t)y > p);
    uint64_t mag = pos ? y - (uint64_t)p : (uint64_t)p - y;
    if (W == 32) mag &= 0xFFFFFFFFull;
    *mag_out = mag;
    if (l <= 0) return 0;
    return (mag & ((1ull << (l - 1)) - 1)) | ((uint64_t)pos << (l - 1));
}

// Sign (true: y > p) and magnitude of the residual, as residual_field.
template <int W>
__device__ __forceinline__ bool residual_sign(uint64_t y, int64_t p, uint64_t *mag_out) {
    const bool pos = (W == 64) ? (y > (uint64_t)p) : ((int64_t)y > p);
    uint64_t mag = pos ? y - (uint64_t)p : (uint64_t)p - y;
    if (W == 32) mag &= 0xFFFFFFFFull;
    *mag_out = mag;
    return pos;
}

// Value from a prediction and an l-bit field (bit_read.h:238-254).
__device__ __forceinline__ uint64_t apply_field(int64_t p, uint64_t field, int l) {
    uint64_t mag = field & ((1ull << (l - 1)) - 1);
    return ((field >> (l - 1)) & 1) ? (uint64_t)p + mag : (uint64_t)p - mag;
}

__device__ __forceinline__ uint64_t low_mask(int l) { return l >= 64 ? ~0ull : ((1ull << l) - 1); }

// Little-endian 8-byte read at any byte address, from two aligned loads.
// Reads up to 15 bytes past addr+8; callers keep that inside the buffer.
__device__ __forceinline__ uint64_t load_u64_any(const uint8_t *p) {
    uintptr_t a = (uintptr_t)p;
    const uint64_t *q = (const uint64_t *)(a & ~(uintptr_t)7);
    int s = (int)(a & 7) * 8;
    uint64_t w0 = q[0];
    if (!s) return w0;
    uint64_t w1 = q[1];
    return (w0 >> s) | (w1 << (64 - s));
}

// `bits` bits starting at bit `bit` of the bytes at base (l <= 64).
__device__ __forceinline__ uint64_t load_bits(const uint8_t *base, uint64_t bit, int bits) {
    const uint8_t *p = base + (bit >> 3);
    uintptr_t a = (uintptr_t)p;
    const uint64_t *q = (const uint64_t *)(a & ~(uintptr_t)7);
    int s = (int)(a & 7) * 8 + (int)(bit & 7);
    uint64_t w0 = q[0], w1 = q[1];
    uint64_t v = s ? (w0 >> s) | (w1 << (64 - s)) : w0;
    return v & low_mask(bits);
}

__device__ __forceinline__ double bits_to_double(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ uint64_t double_to_bits(double d) { return (uint64_t)__double_as_longlong(d); }

// ---- wave / workgroup reductions (wave64) ----------------------------------

// A value every lane holds alike (an LDS broadcast, say), marked uniform for
// the compiler so that branches on it stay scalar.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// 64-bit lane permutation within rows of 16 lanes (DPP; every source lane
// exists for the patterns used here).
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)v, (int)(uint32_t)v, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(v >> 32), (int)(uint32_t)(v >> 32), CTRL, 0xf, 0xf,
                                               false);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

// Lane L receives lane L-1's v (DPP wave_shr:1); lane 0 receives 0.
__device__ __forceinline__ uint64_t wave_prev_u64(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), 0x138, 0xf, 0xf, false);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

// Reduction of a commutative, associative op over each row of 16 lanes;
// every lane of the row ends with the row's result.  Lane exchanges: xor 1
// and xor 2 (quad permutations), then the 8-lane and 16-lane mirrors, which
// pair each quad with the other quad of its half-row and each half-row with
// the other half; every step combines two disjoint sets of lanes.
template <class Op>
__device__ __forceinline__ uint64_t row_reduce_u64(uint64_t v, Op op) {
    v = op(v, dpp_u64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = op(v, dpp_u64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = op(v, dpp_u64<0x141>(v));  // row_half_mirror
    v = op(v, dpp_u64<0x140>(v));  // row_mirror
    return v;
}

// Sum over the wave (all 64 lanes active), returned to every lane: row sums
// by DPP, then the four rows' totals read from lanes 0, 16, 32 and 48.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    v = row_reduce_u64(v, [](uint64_t a, uint64_t c) { return a + c; });
    uint64_t s = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, 16 * r);
        const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 16 * r);
        s += (uint64_t)lo | ((uint64_t)hi << 32);
    }
    return s;
}

// Inclusive prefix sums over the wave's 64 lanes (all active), modulo 2^32 or
// 2^64: Hillis-Steele steps inside each row of 16 lanes by DPP row shifts of
// 1, 2, 4 and 8 (a lane whose source lies before its row adds 0), then row
// broadcasts carry the rows' totals on: lane 15's into rows 1 and 3 (lane 47's
// for row 3: the broadcast reads the row before), and lane 31's into rows 2
// and 3.  The pieces of a 64-bit value move alike, and the adds carry.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_or_zero_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_or_zero_u64(uint64_t v) {
    return (uint64_t)dpp_or_zero_u32<CTRL, ROWS>((uint32_t)v) |
           ((uint64_t)dpp_or_zero_u32<CTRL, ROWS>((uint32_t)(v >> 32)) << 32);
}
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lanes");
    if constexpr (sizeof(T) == 4) {
        v += (T)dpp_or_zero_u32<0x111, 0xf>((uint32_t)v);  // row_shr:1
        v += (T)dpp_or_zero_u32<0x112, 0xf>((uint32_t)v);  // row_shr:2
        v += (T)dpp_or_zero_u32<0x114, 0xf>((uint32_t)v);  // row_shr:4
        v += (T)dpp_or_zero_u32<0x118, 0xf>((uint32_t)v);  // row_shr:8
        v += (T)dpp_or_zero_u32<0x142, 0xa>((uint32_t)v);  // row_bcast:15 into rows 1, 3
        v += (T)dpp_or_zero_u32<0x143, 0xc>((uint32_t)v);  // row_bcast:31 into rows 2, 3
    } else {
        v += (T)dpp_or_zero_u64<0x111, 0xf>((uint64_t)v);
        v += (T)dpp_or_zero_u64<0x112, 0xf>((uint64_t)v);
        v += (T)dpp_or_zero_u64<0x114, 0xf>((uint64_t)v);
        v += (T)dpp_or_zero_u64<0x118, 0xf>((uint64_t)v);
        v += (T)dpp_or_zero_u64<0x142, 0xa>((uint64_t)v);
        v += (T)dpp_or_zero_u64<0x143, 0xc>((uint64_t)v);
    }
    return v;
}

// Lane 63's value, in every lane (scalar reads).
template <typename T>
__device__ __forceinline__ T wave_last(T v) {
    if constexpr (sizeof(T) == 4) return (T)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    else
        return (T)((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63) |
                   ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), 63) << 32));
}

// Reduction over the wave (all 64 lanes active), returned to every lane: row
// results by DPP (VALU, no LDS queue), then the four rows' results read from
// lanes 0, 16, 32 and 48 (scalar).
template <class Op>
__device__ __forceinline__ uint64_t wave_reduce_u64(uint64_t v, Op op) {
    v = row_reduce_u64(v, op);
    uint64_t r = 0;
#pragma unroll
    for (int row = 0; row < 4; ++row) {
        const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, 16 * row);
        const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 16 * row);
        const uint64_t x = (uint64_t)lo | ((uint64_t)hi << 32);
        r = row ? op(r, x) : x;
    }
    return r;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
    return wave_reduce_u64(v, [](uint64_t a, uint64_t b) { return a > b ? a : b; });
}
__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
    return (int64_t)wave_reduce_u64((uint64_t)v, [](uint64_t a, uint64_t b) { return (int64_t)a > (int64_t)b ? a : b; });
}
__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
    return (int64_t)wave_reduce_u64((uint64_t)v, [](uint64_t a, uint64_t b) { return (int64_t)a < (int64_t)b ? a : b; });
}

}  // namespace leco
