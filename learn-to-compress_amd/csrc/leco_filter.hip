// Predicate pushdown on compressed columns: Leco_int<T>::filter_range and
// filter_range_close (headers/piecewise_fix_integer_template.h:206-303 of the
// reference, with read_all_bit_fix_range / read_all_bit_fix_range_close,
// headers/bit_read.h:387-497) over every block of a column in one launch.
//
// One workgroup per tile of filter_tile_blocks() consecutive blocks, in
// ticket order.  Pass 1 counts the tile's matches; a decoupled look-back over
// the tiles' status words (caller workspace) gives the tile's first output
// slot; pass 2 evaluates the same windows again and writes the positions,
// compacted in order with wave ballots.  Output order is the reference's:
// block by block, period by period (filter_range_close), position by
// position.  Only the windows the reference scans are read: positions from
// the model-derived threshold on.
#include "leco_kernels.h"

namespace leco {

namespace {

// x86 CVTTSD2SI with a 32-bit destination: the reference's `int thre = ...`
// conversions (NaN or out of range -> INT_MIN).
__device__ __forceinline__ int32_t cvtt32(double x) {
    if (!(x >= -2147483648.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}
__device__ __forceinline__ double std_max(double a, double b) { return a < b ? b : a; }  // std::max(a, b)
__device__ __forceinline__ double std_min(double a, double b) { return b < a ? b : a; }  // std::min(a, b)

// The decoded value as the filter loops compare it: int64 bits, unsigned
// compare for 64-bit columns, signed int64 for 32-bit columns
// (bit_read.h:422, 478; piecewise_fix_integer_template.h:238, 285).
template <int W>
__device__ __forceinline__ bool above(uint64_t v, uint64_t f) {
    return W == 64 ? v > f : (int64_t)v > (int64_t)f;
}
template <int W>
__device__ __forceinline__ bool below(uint64_t v, uint64_t f) {
    return W == 64 ? v < f : (int64_t)v < (int64_t)f;
}

struct FilterShared {
    uint64_t wsum[kThreads / kWave];
    uint32_t wcnt[kThreads / kWave];
    uint64_t base;
    uint64_t tile;
};

// Arguments of one filter launch.
struct FilterArgs {
    const uint8_t *stream;
    const uint64_t *offs;
    uint64_t n;
    uint32_t bs;
    uint32_t kb;            // blocks per tile
    uint64_t tiles;
    uint64_t filter;        // filter_range
    const void *f1, *f2;    // filter_range_close: per-block starting filters (T)
    int32_t base;           // filter_range_close: period
    uint32_t pos_base;      // added to every reported position (uint32 arithmetic)
    uint32_t *out;
    uint64_t cap;
    uint64_t *result;       // [0] = count, [1] = flags
    uint64_t *work;         // [0] = ticket, [1 + t] = tile t's look-back status
};

// Visits the windows of block b: visit(lo, hi, raw, l, t0, t1, f1, f2) for
// each window of positions [lo, hi) with its predicate bounds (filter_range:
// value > f1; close: f1 < value < f2; raw blocks: see the raw test below).
// Returns false if the block needed more than kFilterMaxPeriods periods.
template <typename T, int W, bool CLOSE, class Visit>
__device__ __forceinline__ bool block_windows(const FilterArgs &a, uint64_t b, uint32_t m, Visit visit) {
    constexpr uint64_t wmask = W == 32 ? 0xFFFFFFFFull : ~0ull;
    const uint8_t *blk = a.stream + a.offs[b];
    const int l = blk[0];
    uint64_t f1 = CLOSE ? (uint64_t)((const T *)a.f1)[b] : (a.filter & wmask);
    uint64_t f2 = CLOSE ? (uint64_t)((const T *)a.f2)[b] : 0;
    if (l >= W) {
        // raw blocks (and the one-/two-value blocks 127/126): every stored value
        const uint32_t mm = l == 127 ? (m < 1 ? m : 1) : l == 126 ? (m < 2 ? m : 2) : m;
        visit(0u, mm, true, l, 0.0, 0.0, f1, f2);
        return true;
    }
    const double t0 = bits_to_double(load_u64_any(blk + 1)), t1 = bits_to_double(load_u64_any(blk + 9));
    const int64_t delta = l ? (int64_t)1 << (l - 1) : 0;
    if (!CLOSE) {
        // :225-229
        double x;
        if (W == 64) x = (double)(f1 + 1 - (uint64_t)delta) - t0;
        else x = (double)((int64_t)(uint32_t)(f1 + 1) - delta) - t0;
        const int32_t thre = cvtt32(std_max(x / t1, 0.0));
        if ((uint64_t)(int64_t)thre >= m) return true;
        visit((uint32_t)thre, m, false, l, t0, t1, f1, f2);
        return true;
    }
    // :269-300: windows [thre1, thre2) (l > 0) or [thre1, thre2] (l == 0),
    // filters advanced by `base` after each
    const uint64_t step = (uint64_t)(int64_t)a.base & wmask;
    for (uint32_t periods = 0;; ++periods) {
        double x1, x2;
        if (W == 64) {
            x1 = (double)(f1 - (uint64_t)delta) - t0;
            x2 = (double)(f2 + (uint64_t)delta) - t0;
        } else {
            x1 = (double)((int64_t)f1 - delta) - t0;
            x2 = (double)((int64_t)f2 + delta) - t0;
        }
        const int32_t thre1 = cvtt32(std_max(x1 / t1 - 1, 0.0));
        const int32_t thre2 = cvtt32(std_min(x2 / t1 + 1, (double)m));
        if (!((uint64_t)(int64_t)thre1 < m && thre2 > 0)) return true;
        if (periods >= kFilterMaxPeriods) return false;
        const uint32_t hi = l ? (uint32_t)thre2 : (uint32_t)thre2 + 1;
        if ((uint32_t)thre1 < hi) visit((uint32_t)thre1, hi, false, l, t0, t1, f1, f2);
        f1 = (f1 + step) & wmask;
        f2 = (f2 + step) & wmask;
    }
}

// Does position i of block b (header at blk) pass?
template <int W, bool CLOSE>
__device__ __forceinline__ bool passes(const uint8_t *blk, uint32_t i, bool raw, int l, double t0, double t1,
                                       uint64_t f1, uint64_t f2, int32_t base) {
    if (raw) {
        const uint64_t v = load_bits(blk + 1, (uint64_t)i * W, W);
        if (!CLOSE) return v > f1;
        // raw deviation: v in (f1 + k*base, f2 + k*base) for some k >= 0
        if (v <= f1) return false;
        const uint64_t d = v - f1, span = f2 > f1 ? f2 - f1 : 0;
        if (base <= 0) return v < f2;
        return (d - 1) % (uint64_t)base + 1 < span;
    }
    const int64_t p = predict(t0, t1, i);
    const uint64_t v = l ? apply_field(p, load_bits(blk + 17, (uint64_t)i * l, l), l) : (uint64_t)p;
    return CLOSE ? (above<W>(v, f1) && below<W>(v, f2)) : above<W>(v, f1);
}

template <typename T, int W, bool CLOSE>
__global__ __launch_bounds__(kThreads) void k_filter(FilterArgs a) {
    __shared__ FilterShared sh;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    if (tid == 0) sh.tile = atomicAdd((unsigned long long *)a.work, 1ull);  // work[0] is the ticket
    __syncthreads();
    const uint64_t t = sh.tile;
    const uint64_t nb = (a.n + a.bs - 1) / a.bs;
    const uint64_t b0 = t * a.kb;
    const uint32_t kbt = (uint32_t)(nb - b0 < a.kb ? nb - b0 : a.kb);
    auto block_len = [&](uint64_t b) -> uint32_t {
        const uint64_t s = b * a.bs;
        return (uint32_t)(a.n - s < a.bs ? a.n - s : a.bs);
    };

    // pass 1: count
    uint64_t mine = 0;
    bool ok = true;
    for (uint32_t k = 0; k < kbt; ++k) {
        const uint64_t b = b0 + k;
        const uint8_t *blk = a.stream + a.offs[b];
        ok &= block_windows<T, W, CLOSE>(a, b, block_len(b),
                                         [&](uint32_t lo, uint32_t hi, bool raw, int l, double t0, double t1,
                                             uint64_t f1, uint64_t f2) {
                                             for (uint32_t i = lo + tid; i < hi; i += kThreads)
                                                 mine += passes<W, CLOSE>(blk, i, raw, l, t0, t1, f1, f2, a.base);
                                         });
    }
    for (int o = kWave / 2; o > 0; o >>= 1) mine += (uint64_t)__shfl_xor((long long)mine, o, kWave);
    if (lane == 0) sh.wsum[wid] = mine;
    __syncthreads();
    if (wid == 0) {
        uint64_t total = 0;
#pragma unroll
        for (int w = 0; w < kThreads / kWave; ++w) total += sh.wsum[w];
        const uint64_t excl = look_back(t, total, a.work);
        if (lane == 0) {
            sh.base = excl;
            if (t + 1 == a.tiles) a.result[0] = excl + total;
            if (!ok) atomicOr((unsigned long long *)&a.result[1], 1ull);
        }
    }
    __syncthreads();

    // pass 2: the same windows, positions written in order
    uint64_t pos = sh.base;
    for (uint32_t k = 0; k < kbt; ++k) {
        const uint64_t b = b0 + k;
        const uint8_t *blk = a.stream + a.offs[b];
        const uint32_t bstart = a.pos_base + (uint32_t)(b * a.bs);
        block_windows<T, W, CLOSE>(a, b, block_len(b),
                                   [&](uint32_t lo, uint32_t hi, bool raw, int l, double t0, double t1, uint64_t f1,
                                       uint64_t f2) {
                                       for (uint32_t r = lo; r < hi; r += kThreads) {
                                           const uint32_t i = r + tid;
                                           const bool hit =
                                               i < hi && passes<W, CLOSE>(blk, i, raw, l, t0, t1, f1, f2, a.base);
                                           const uint64_t bal = __ballot(hit);
                                           if (lane == 0) sh.wcnt[wid] = (uint32_t)__popcll(bal);
                                           __syncthreads();
                                           uint64_t at = pos;
                                           uint32_t round = 0;
#pragma unroll
                                           for (int w = 0; w < kThreads / kWave; ++w) {
                                               if (w < wid) at += sh.wcnt[w];
                                               round += sh.wcnt[w];
                                           }
                                           at += (uint64_t)__popcll(bal & ((1ull << lane) - 1));
                                           if (hit && at < a.cap) a.out[at] = bstart + i;
                                           pos += round;
                                           __syncthreads();
                                       }
                                   });
    }
}

__global__ __launch_bounds__(kThreads) void k_filter_clear(uint64_t *__restrict__ work, uint64_t words,
                                                           uint64_t *__restrict__ result) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < words) work[i] = 0;
    if (i == 0) result[0] = result[1] = 0;
}

}  // namespace

template <typename T>
hipError_t launch_filter(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, bool close,
                         uint64_t filter, const T *f1, const T *f2, int32_t base, uint32_t pos_base, uint32_t *out,
                         uint64_t cap, uint64_t *result, uint64_t *work, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    FilterArgs a{stream, offs, n, bs, filter_tile_blocks(bs), filter_tiles(n, bs), filter, f1, f2, base, pos_base,
                 out, cap, result, work};
    const uint64_t words = a.tiles + 1;
    k_filter_clear<<<dim3((unsigned)((words + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(work, words,
                                                                                                    result);
    if (close) k_filter<T, W, true><<<dim3((unsigned)a.tiles), dim3(kThreads), 0, st>>>(a);
    else k_filter<T, W, false><<<dim3((unsigned)a.tiles), dim3(kThreads), 0, st>>>(a);
    return hipGetLastError();
}

template hipError_t launch_filter<uint32_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, bool, uint64_t,
                                            const uint32_t *, const uint32_t *, int32_t, uint32_t, uint32_t *,
                                            uint64_t, uint64_t *, uint64_t *, hipStream_t);
template hipError_t launch_filter<uint64_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, bool, uint64_t,
                                            const uint64_t *, const uint64_t *, int32_t, uint32_t, uint32_t *,
                                            uint64_t, uint64_t *, uint64_t *, hipStream_t);

}  // namespace leco
