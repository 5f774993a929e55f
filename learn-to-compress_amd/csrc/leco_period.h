// The period index of filter_range_close's scan: n / base by one FP64
// multiply and one correction, host and device (the host build is what
// tests/cpp/period_check.cpp compares with integer division).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace leco {

// 1 / base as the scan keeps it (base >= 1).
__host__ __device__ inline double period_inverse(uint64_t base) { return 1.0 / (double)base; }

// n / base for n < K * base, 1 <= base < 2^31, K <= 2^16 (so n < 2^47 is an
// exact double and the quotient is below 2^16).  n * (1/base) carries a
// relative error of a few 2^-53, far less than 1/base, the least distance
// from a non-integer n / base to the next integer; so its truncation is the
// quotient, or one less when n / base is an integer that the rounding left
// just below.  One remainder test corrects that case.
__host__ __device__ inline uint32_t period_quotient(uint64_t n, uint64_t base, double inv_base) {
    uint32_t q = (uint32_t)((double)n * inv_base);
    if (n - (uint64_t)q * base >= base) ++q;
    return q;
}

}  // namespace leco
