// Host-side launchers for the gfx950 kernels (defined in leco_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "leco_fix.h"

namespace leco {

// Largest block size the fit supports: sum(i) and sum(i*i) over a block are
// exact doubles up to here, so they are formed in closed form.
constexpr uint32_t kMaxBlockSize = 1u << 18;

__host__ __device__ inline uint64_t payload_bytes(uint32_t l, uint32_t n) {
    uint64_t e = ((uint64_t)l * n + 7) / 8;
    return 8 * (e / 8) + 9;
}
__host__ __device__ inline uint64_t block_bytes(uint32_t l, uint32_t w, uint32_t n) {
    if (l >= w) return 1 + (uint64_t)n * (w / 8);
    return 17 + (l ? payload_bytes(l, n) : 0);
}

// Each launcher enqueues on `st` and returns the hipGetLastError() of the launch.
template <typename T>
hipError_t launch_encode(const T *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                         leco_fix_block_param *prm, hipStream_t st);
template <typename T>
hipError_t launch_decode(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                         uint64_t lo, uint64_t hi, T *out, hipStream_t st);
template <typename T>
hipError_t launch_lookup(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                         const uint64_t *idx, uint64_t nq, T *out, hipStream_t st);
hipError_t launch_build_params(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                               uint32_t w, leco_fix_block_param *prm, hipStream_t st);

}  // namespace leco
