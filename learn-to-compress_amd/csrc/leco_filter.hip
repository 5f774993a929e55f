// Predicate pushdown on compressed columns: Leco_int<T>::filter_range and
// filter_range_close (headers/piecewise_fix_integer_template.h:206-303 of the
// reference, with read_all_bit_fix_range / read_all_bit_fix_range_close,
// headers/bit_read.h:387-497) over every block of a column.
//
// Both filters run as a mark launch and an emit launch over tiles with no
// serial dependence between them.  mark counts each tile's matches (and,
// where it can, keeps them as a match bitmap), adding the count to the sum of
// its group of tiles; emit places the tile's output after the groups and
// tiles before it and writes its positions, from the bitmap or by evaluating
// the tile's windows again, compacted in order with wave ballots.  Output
// order is the reference's: block by block, period by period
// (filter_range_close), position by position.  Only the windows the reference
// scans are read: positions from the model-derived threshold on.
#include <atomic>

#include "leco_kernels.h"
#include "leco_period.h"

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

constexpr int kWaves = kThreads / kWave;
constexpr int kEmitPer = 8;  // rounds of kThreads candidates per compaction step when writing positions
constexpr uint32_t kCloseRounds = 64;  // scan rounds of a tile whose ballots pass 1 parks for pass 2
constexpr uint32_t kCloseWin = 2048;   // periods whose windows pass 1 keeps for the scan
constexpr uint32_t kCloseStage = 8192; // bytes of a block's fields the scan sweep stages through LDS
constexpr uint32_t kCloseTableLen = 32768;  // blocks shorter than this keep the window table

struct FilterShared {
    uint64_t wsum[kWaves];
    uint32_t wcnt[kWaves];
    uint32_t jc[kEmitPer][kWaves];
    uint32_t wfail[kWaves];
    uint32_t stop_passing;  // the period the close loop stopped at still passed its test (limit cut)
    uint64_t base;
    uint64_t tile;
    // per block of the tile: the close loop's period count, how pass 2 reports it (0 walk,
    // 1 scan evaluated again, 2 scan from the ballots pass 1 parked) and its first parked round
    uint32_t bK[256];
    uint8_t bscan[256];
    uint8_t bl[256];  // the block's header byte
    uint16_t bround[256];
    uint64_t rb[kCloseRounds * kWaves];  // parked scan ballots of (round, wave)
    // thre1, thre2 of each of the current block's first kCloseWin periods (blocks shorter than
    // kCloseTableLen), as int16: values below -32768 are kept as -32768, which compares with
    // every position below 2^15 as they do (as uint32 both are above every position)
    int16_t pw1[kCloseWin], pw2[kCloseWin];
    union {
        struct {  // one chunk of kThreads periods of the close walk
            uint64_t pf1[kThreads], pf2[kThreads];
            uint32_t plo[kThreads];
            uint32_t ppre[kThreads + 1];
        };
        // the scan's staged fields of one block; at the end of mark, the tile's bitmap
        alignas(16) uint64_t stage[kCloseStage / 8];
    };
};

// Arguments of one filter launch.
struct FilterArgs {
    const uint8_t *stream;
    const uint64_t *offs;
    uint64_t n;
    uint32_t bs;
    uint32_t kb;            // blocks per tile
    uint64_t tiles;
    uint64_t filter;        // filter_range; FOR filter_range_close: filter1
    uint64_t filter2;       // FOR filter_range_close: filter2
    const void *f1, *f2;    // LeCo filter_range_close: per-block starting filters (T)
    int32_t base;           // filter_range_close: period
    uint32_t pos_base;      // added to every reported position (uint32 arithmetic)
    uint32_t *out;
    uint64_t cap;
    uint64_t *result;       // [0] = count, [1] = flags
    uint64_t *work;         // the group sums, the tiles' counts, then the tiles' bitmaps
};

// Does position i of the block at blk pass?
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

// Workgroup exclusive scan of one value per thread (through sh.wcnt);
// returns this thread's prefix and the total in *total (every thread).
template <class Shared>
__device__ __forceinline__ uint32_t wg_exclusive_scan(Shared &sh, uint32_t v, uint32_t *total);

// Candidates q in [0, total) in order; cand(q, &position) says whether q
// passes and where it is.  Pass 1 (WRITE = false) adds this thread's hits to
// acc (any order); pass 2 takes kEmitPer rounds of kThreads candidates per
// step (round j: q = r + j*kThreads + tid, so loads and stores stay
// coalesced), keeps each wave's ballot per round, exchanges the per-(round,
// wave) counts through LDS once, and writes every hit at acc + the number of
// hits before it in (round, wave, lane) order -- candidate order.  acc is the
// workgroup's running output slot (the same in every thread).
template <bool WRITE, class Cand>
__device__ __forceinline__ void emit_ordered(const FilterArgs &a, FilterShared &sh, uint32_t total, Cand cand,
                                             uint64_t &acc) {
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    if (!WRITE) {
        uint32_t pos;
        for (uint32_t q = tid; q < total; q += kThreads) acc += cand(q, &pos);
        return;
    }
    const uint64_t below = (1ull << lane) - 1;
    for (uint32_t r = 0; r < total; r += kThreads * kEmitPer) {
        uint64_t bal[kEmitPer];
        uint32_t where[kEmitPer];
#pragma unroll
        for (int j = 0; j < kEmitPer; ++j) {
            const uint32_t q = r + j * kThreads + tid;
            where[j] = 0;
            bal[j] = __ballot(q < total && cand(q, &where[j]));
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < kEmitPer; ++j) sh.jc[j][wid] = (uint32_t)__popcll(bal[j]);
        }
        __syncthreads();
        uint64_t run = acc;
#pragma unroll
        for (int j = 0; j < kEmitPer; ++j) {
            uint32_t before = 0, row = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = sh.jc[j][w];
                before += w < wid ? c : 0;
                row += c;
            }
            if ((bal[j] >> lane) & 1) {
                const uint64_t at = run + before + (uint64_t)__popcll(bal[j] & below);
                if (at < a.cap) a.out[at] = where[j];
            }
            run += row;
        }
        acc = run;
        __syncthreads();
    }
}

// Positions [lo, hi) of one window against one predicate.
template <int W, bool CLOSE, bool WRITE>
__device__ __forceinline__ void scan_window(const FilterArgs &a, FilterShared &sh, const uint8_t *blk, uint32_t lo,
                                            uint32_t hi, bool raw, int l, double t0, double t1, uint64_t f1,
                                            uint64_t f2, uint32_t bstart, uint64_t &acc) {
    emit_ordered<WRITE>(a, sh, hi > lo ? hi - lo : 0,
                        [&](uint32_t q, uint32_t *where) {
                            *where = bstart + lo + q;
                            return passes<W, CLOSE>(blk, lo + q, raw, l, t0, t1, f1, f2, a.base);
                        },
                        acc);
}

template <class Shared>
__device__ __forceinline__ uint32_t wg_exclusive_scan(Shared &sh, uint32_t v, uint32_t *total) {
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    uint32_t inc = v;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)inc, o, kWave);
        if (lane >= o) inc += t;
    }
    if (lane == kWave - 1) sh.wcnt[wid] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        if (w < wid) before += sh.wcnt[w];
        all += sh.wcnt[w];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// filter_range (:206-249): the first position of block b's window, from the
// model threshold (positions before it are not scanned).  m when the window is empty.
template <int W>
__device__ __forceinline__ uint32_t gt_window(uint64_t f, int l, double t0, double t1, uint32_t m) {
    const int64_t delta = l ? (int64_t)1 << (l - 1) : 0;
    double x;
    if (W == 64) x = (double)(f + 1 - (uint64_t)delta) - t0;
    else x = (double)((int64_t)(uint32_t)(f + 1) - delta) - t0;
    const int32_t thre = cvtt32(std_max(x / t1, 0.0));
    return (uint64_t)(int64_t)thre >= m ? m : (uint32_t)thre;
}

// ---- filter_range_close (:251-303) ----------------------------------------
// The reference's loop visits periods k = 0, 1, ... of block b with filters
// (f1 + k*base, f2 + k*base) (T wrap; the same values as adding base k
// times) and window [thre1_k, thre2_k) ([.., thre2_k] when l == 0), and stops
// at the first period whose window test fails (K periods).  Two exact ways to
// produce its output:
//   walk:  the periods in order, kThreads at a time: window sizes scanned into
//          one flat list of (period, position) candidates, walked in order;
//   scan:  each position once.  When the periods' intervals are disjoint
//          (f2 - f1 <= base) and do not wrap before period K, a value can pass
//          only in its own period k*(v); when the block's values also do not
//          decrease, k* does not decrease with the position, so period-major
//          order is position order.  Position i is reported iff k*(v_i) < K,
//          i lies in that period's window and the value passes.
// Pass 1 finds K, picks scan when its conditions hold (the value order is
// checked during the same sweep) and remembers the choice for pass 2.

// Thresholds of period k (its filters f1, f2): thre1, thre2 as the reference's int conversions.
template <int W>
__device__ __forceinline__ void period_window(uint64_t f1, uint64_t f2, int64_t delta, double t0, double t1,
                                              uint32_t m, int32_t *thre1, int32_t *thre2) {
    double x1, x2;
    if (W == 64) {
        x1 = (double)(f1 - (uint64_t)delta) - t0;
        x2 = (double)(f2 + (uint64_t)delta) - t0;
    } else {
        x1 = (double)((int64_t)f1 - delta) - t0;
        x2 = (double)((int64_t)f2 + delta) - t0;
    }
    *thre1 = cvtt32(std_max(x1 / t1 - 1, 0.0));
    *thre2 = cvtt32(std_min(x2 / t1 + 1, (double)m));
}

struct CloseBlock {
    const uint8_t *blk;
    int l;
    double t0, t1;
    int64_t delta;
    uint64_t f1, f2, step;  // period 0's filters, base as T
    uint32_t m;
    double inv_base;        // 1 / base (scan mode: base >= 1)
};

// Periods whose windows close_periods keeps in the table.
__device__ __forceinline__ uint32_t close_nwin(const CloseBlock &c, uint32_t K) {
    return c.m < kCloseTableLen ? (K < kCloseWin ? K : kCloseWin) : 0;
}

// K: the periods the reference's loop runs (at most kFilterMaxPeriods);
// *cut: the loop was stopped by the limit with the next period still passing.
template <int W>
__device__ __forceinline__ uint32_t close_periods(const CloseBlock &c, FilterShared &sh, bool *cut) {
    constexpr uint64_t wmask = W == 32 ? 0xFFFFFFFFull : ~0ull;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    for (uint32_t k0 = 0;; k0 += kThreads) {
        const uint32_t k = k0 + tid;
        int32_t thre1, thre2;
        period_window<W>((c.f1 + (uint64_t)k * c.step) & wmask, (c.f2 + (uint64_t)k * c.step) & wmask, c.delta,
                         c.t0, c.t1, c.m, &thre1, &thre2);
        if (k < kCloseWin && c.m < kCloseTableLen) {
            sh.pw1[k] = (int16_t)(thre1 < -32768 ? -32768 : thre1);
            sh.pw2[k] = (int16_t)(thre2 < -32768 ? -32768 : thre2);
        }
        const bool cond = (uint64_t)(int64_t)thre1 < c.m && thre2 > 0;
        const uint64_t sb = __ballot(!cond || k >= kFilterMaxPeriods);
        if (lane == 0) sh.wfail[wid] = sb ? (uint32_t)(wid * kWave + __ffsll((long long)sb) - 1) : kThreads;
        __syncthreads();
        uint32_t first = kThreads;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) first = sh.wfail[w] < first ? sh.wfail[w] : first;
        if ((uint32_t)tid == first) sh.stop_passing = cond ? 1u : 0u;
        __syncthreads();
        if (first < (uint32_t)kThreads) {
            *cut = sh.stop_passing != 0;
            __syncthreads();
            return k0 + first;
        }
    }
}

// Walk: periods [0, K) in order, kThreads at a time.
template <int W, bool WRITE>
__device__ __forceinline__ void close_walk(const FilterArgs &a, const CloseBlock &c, FilterShared &sh, uint32_t K,
                                           uint32_t bstart, uint64_t &acc) {
    constexpr uint64_t wmask = W == 32 ? 0xFFFFFFFFull : ~0ull;
    const int tid = threadIdx.x;
    for (uint32_t k0 = 0; k0 < K; k0 += kThreads) {
        const uint32_t k = k0 + tid;
        const uint32_t live = K - k0 < (uint32_t)kThreads ? K - k0 : kThreads;
        const uint64_t f1 = (c.f1 + (uint64_t)k * c.step) & wmask, f2 = (c.f2 + (uint64_t)k * c.step) & wmask;
        int32_t thre1, thre2;
        period_window<W>(f1, f2, c.delta, c.t0, c.t1, c.m, &thre1, &thre2);
        const uint32_t hi = c.l ? (uint32_t)thre2 : (uint32_t)thre2 + 1;
        const bool on = (uint32_t)tid < live;
        const uint32_t sz = (on && (uint32_t)thre1 < hi) ? hi - (uint32_t)thre1 : 0;
        uint32_t total;
        const uint32_t pre = wg_exclusive_scan(sh, sz, &total);
        sh.pf1[tid] = f1;
        sh.pf2[tid] = f2;
        sh.plo[tid] = on ? (uint32_t)thre1 : 0;
        sh.ppre[tid] = pre;
        __syncthreads();
        emit_ordered<WRITE>(a, sh, total,
                            [&](uint32_t q, uint32_t *where) {
                                uint32_t lo = 0, hi2 = live;  // the last period j < live with ppre[j] <= q
                                while (hi2 - lo > 1) {
                                    const uint32_t mid = (lo + hi2) / 2;
                                    if (sh.ppre[mid] <= q) lo = mid;
                                    else hi2 = mid;
                                }
                                const uint32_t i = sh.plo[lo] + (q - sh.ppre[lo]);
                                *where = bstart + i;
                                return passes<W, true>(c.blk, i, false, c.l, c.t0, c.t1, sh.pf1[lo], sh.pf2[lo],
                                                       a.base);
                            },
                            acc);
        __syncthreads();
    }
}

// The value the close loop compares at position i (int64 bits); i == m is
// the l == 0 loop's extra position (model only).
__device__ __forceinline__ uint64_t close_value(const CloseBlock &c, uint32_t i) {
    const int64_t p = predict(c.t0, c.t1, i);
    return c.l ? apply_field(p, load_bits(c.blk + 17, (uint64_t)i * c.l, c.l), c.l) : (uint64_t)p;
}

// Scan: does position i (value v) pass in its own period?  The windows of
// periods below nwin come from the table close_periods filled.
template <int W>
__device__ __forceinline__ bool close_scan_hit(const CloseBlock &c, uint32_t K, uint32_t i, uint64_t v,
                                               int32_t base, const FilterShared &sh, uint32_t nwin) {
    uint64_t d;
    if (W == 64) {
        if (v <= c.f1) return false;
        d = v - c.f1;
    } else {
        if ((int64_t)v <= (int64_t)c.f1) return false;
        d = (uint64_t)((int64_t)v - (int64_t)c.f1);
    }
    // k: the last period whose f1 lies below v (periods from K on never pass)
    if (d - 1 >= (uint64_t)K * (uint64_t)base) return false;
    const uint64_t k = period_quotient(d - 1, (uint64_t)base, c.inv_base);
    const uint64_t f1 = c.f1 + k * (uint64_t)base, f2 = c.f2 + k * (uint64_t)base;
    if (!(W == 64 ? v < f2 : (int64_t)v < (int64_t)f2)) return false;
    int32_t thre1, thre2;
    if (k < nwin) {
        thre1 = sh.pw1[k];
        thre2 = sh.pw2[k];
    } else {
        period_window<W>(f1, f2, c.delta, c.t0, c.t1, c.m, &thre1, &thre2);
    }
    const uint32_t hi = c.l ? (uint32_t)thre2 : (uint32_t)thre2 + 1;
    return i >= (uint32_t)thre1 && i < hi;
}

template <typename T, int W, bool WRITE>
__device__ __forceinline__ bool close_block(const FilterArgs &a, FilterShared &sh, uint64_t *__restrict__ stage,
                                            uint64_t b, uint32_t m, uint32_t slot, uint32_t bstart, uint64_t &acc,
                                            uint32_t &rounds_used) {
    constexpr uint64_t wmask = W == 32 ? 0xFFFFFFFFull : ~0ull;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    CloseBlock c;
    c.blk = a.stream + a.offs[b];
    c.l = c.blk[0];
    if (!WRITE && tid == 0) sh.bl[slot] = (uint8_t)c.l;
    c.f1 = (uint64_t)((const T *)a.f1)[b];
    c.f2 = (uint64_t)((const T *)a.f2)[b];
    c.m = m;
    if (c.l >= W) {
        const uint32_t mm = c.l == 127 ? (m < 1 ? m : 1) : c.l == 126 ? (m < 2 ? m : 2) : m;
        scan_window<W, true, WRITE>(a, sh, c.blk, 0, mm, true, c.l, 0.0, 0.0, c.f1, c.f2, bstart, acc);
        return true;
    }
    c.t0 = bits_to_double(load_u64_any(c.blk + 1));
    c.t1 = bits_to_double(load_u64_any(c.blk + 9));
    c.delta = c.l ? (int64_t)1 << (c.l - 1) : 0;
    c.step = (uint64_t)(int64_t)a.base & wmask;
    c.inv_base = a.base > 0 ? period_inverse((uint64_t)a.base) : 0.0;
    bool cut = false;
    uint32_t K;
    bool scan;
    uint32_t mode = 0;
    if (!WRITE) {
        K = close_periods<W>(c, sh, &cut);
        // scan needs disjoint, non-wrapping intervals up to period K
        const uint64_t reach = K ? (uint64_t)(K - 1) * c.step : 0;
        scan = K > 0 && a.base > 0 && (c.f2 <= c.f1 || c.f2 - c.f1 <= (uint64_t)a.base) &&
               c.f1 <= wmask - reach && c.f2 <= wmask - reach;
    } else {
        K = sh.bK[slot];
        mode = sh.bscan[slot];
        scan = mode != 0;
    }
    if (K == 0) {
        if (!WRITE && tid == 0) { sh.bK[slot] = 0; sh.bscan[slot] = 0; }
        return !cut;
    }
    const uint32_t mp = c.l ? m : m + 1;  // positions the windows can reach
    const uint32_t nr = (mp + kThreads - 1) / kThreads;  // scan rounds of kThreads positions
    if (!WRITE && scan) {
        // one sweep, each value decoded once: the scan's count, whether the values never
        // decrease (the next value from the next lane; the last lane decodes it), and,
        // while the tile has room, the round ballots for pass 2
        const bool park = rounds_used + nr <= kCloseRounds;
        // the block's fields, [offs[b] + 17, + ceil(m*l/8)), through LDS when they fit (with
        // one word of over-read); model-only blocks (l = 0) have none
        const uint64_t g0 = (uint64_t)(c.blk - a.stream) + 17, a0 = g0 & ~15ull;
        const uint32_t n16 = (uint32_t)((g0 + (((uint64_t)m * c.l + 7) >> 3) - a0 + 15) / 16);
        const bool staged = c.l > 0 && (uint64_t)n16 * 16 + 16 <= kCloseStage;
        if (staged)
            for (uint32_t u = tid; u < n16; u += kThreads)
                *(uint4 *)(stage + 2 * u) = *(const uint4 *)(a.stream + a0 + 16ull * u);
        __syncthreads();
        const uint32_t fbit = (uint32_t)(8 * (g0 - a0));
        uint64_t cnt = 0;
        bool sorted = true;
        // the sweep, instantiated once per field source (each keeps only its own registers)
        auto sweep = [&](auto value) {
            for (uint32_t r = 0; r < nr; ++r) {
                const uint32_t i = r * kThreads + tid;
                uint64_t v = 0;
                bool h = false;
                if (i < mp) {
                    v = value(i);
                    h = close_scan_hit<W>(c, K, i, v, a.base, sh, close_nwin(c, K));
                }
                uint64_t u = (uint64_t)__shfl_down((long long)v, 1, kWave);
                if (lane == kWave - 1 && i + 1 < mp) u = value(i + 1);
                if (i + 1 < mp) sorted &= W == 64 ? v <= u : (int64_t)v <= (int64_t)u;
                cnt += h;
                if (park) {
                    const uint64_t bal = __ballot(h);
                    if (lane == 0) sh.rb[(rounds_used + r) * kWaves + wid] = bal;
                }
            }
        };
        // The common case, without branches: fields of at most 32 bits in LDS (one funnel shift
        // per field), every prediction inside cvtt64's fast range, a block shorter than
        // kCloseTableLen.  The order check takes the previous lane's value by DPP (the
        // boundaries between waves and rounds are checked after the sweep); every lane runs
        // the period test with a clamped quotient and compares the remainder with the
        // period's span, (v - f1 - 1) mod base < f2 - f1 - 1, which is v < f2 + k*base.  The
        // same hits as close_scan_hit.
        auto fast_sweep = [&](auto value, auto quot_rem) {
            const uint64_t base_v = (uint64_t)(uint32_t)a.base, kb = (uint64_t)K * base_v;
            const uint32_t nwin = close_nwin(c, K);
            const uint32_t span = (W == 64 ? c.f2 > c.f1 : (int64_t)c.f2 > (int64_t)c.f1)
                                      ? (uint32_t)(c.f2 - c.f1 - 1) : 0u;
            uint64_t bad = 0, hits = 0;  // hits: the wave's count, wave-uniform (from the ballots)
            // one round of kThreads positions; FULL: every position of the round is below mp
            auto round = [&](uint32_t r, auto full) {
                constexpr bool FULL = decltype(full)::value;
                const uint32_t i = r * kThreads + tid;
                const bool on = FULL || i < mp;
                const uint64_t v = value(FULL || on ? i : mp - 1);
                const uint64_t pv = wave_prev_u64(v);
                bad |= __ballot(on && (W == 64 ? pv > v : (int64_t)pv > (int64_t)v));
                // k: the last period whose f1 lies below v; periods from K on never pass
                const bool gt = W == 64 ? v > c.f1 : (int64_t)v > (int64_t)c.f1;
                const uint64_t d = (W == 64 ? v - c.f1 : (uint64_t)((int64_t)v - (int64_t)c.f1)) - 1;
                const bool in_k = on && gt && d < kb;
                uint32_t rem;
                const uint32_t k = quot_rem(in_k ? d : 0, &rem);
                int32_t thre1, thre2;
                if (k < nwin) {
                    thre1 = sh.pw1[k];
                    thre2 = sh.pw2[k];
                } else {
                    period_window<W>(c.f1 + (uint64_t)k * base_v, c.f2 + (uint64_t)k * base_v, c.delta, c.t0,
                                     c.t1, c.m, &thre1, &thre2);
                }
                const bool h = in_k && rem < span && i >= (uint32_t)thre1 && i < (uint32_t)thre2;
                const uint64_t bal = __ballot(h);
                hits += (uint64_t)__popcll(bal);
                if (park && lane == 0) sh.rb[(rounds_used + r) * kWaves + wid] = bal;
            };
            const uint32_t nfull = mp / kThreads;
            for (uint32_t r = 0; r < nfull; ++r) round(r, std::true_type{});
            if (nfull < nr) round(nfull, std::false_type{});
            if (lane == 0) cnt += hits;  // the workgroup sums every thread's count
            sorted = (bad & ~1ull) == 0;  // lane 0 compared nothing
            for (uint32_t j = tid; 64 * (j + 1) < mp; j += kThreads) {
                const uint64_t x = value(64 * j + 63), y = value(64 * j + 64);
                sorted &= W == 64 ? x <= y : (int64_t)x <= (int64_t)y;
            }
        };
        if (staged && c.l <= 32 && m < kCloseTableLen && near_range(c.t0, c.t1, mp)) {
            const uint32_t *b32 = (const uint32_t *)stage;
            const uint32_t lw = (uint32_t)c.l;
            const uint32_t fmask = lw >= 32 ? ~0u : (1u << lw) - 1, mmask = (1u << (lw - 1)) - 1;
            auto value = [&](uint32_t i) -> uint64_t {
                const uint32_t q = fbit + __umul24(i, lw);  // i < 2^15: the full-rate 24-bit multiply
                const uint32_t f = __builtin_amdgcn_alignbit(b32[(q >> 5) + 1], b32[q >> 5], q & 31) & fmask;
                const uint32_t mag = f & mmask;
                const int32_t res = (f >> (lw - 1)) ? (int32_t)mag : -(int32_t)mag;
                return (uint64_t)(predict_near(c.t0, c.t1, i) + (int64_t)res);
            };
            const uint32_t bu = (uint32_t)a.base;
            if ((bu & (bu - 1)) == 0) {
                const int s = __builtin_ctz(bu);
                const uint32_t bm = bu - 1;
                fast_sweep(value, [=](uint64_t d, uint32_t *rem) -> uint32_t {
                    *rem = (uint32_t)d & bm;
                    // d >> s by a 32-bit funnel shift (s <= 30; the quotient is below 2^16 for d < K*base
                    // and 0 for the d = 0 the caller passes out of range)
                    return __builtin_amdgcn_alignbit((uint32_t)(d >> 32), (uint32_t)d, (uint32_t)s);
                });
            } else {
                const double inv = c.inv_base;
                fast_sweep(value, [=](uint64_t d, uint32_t *rem) -> uint32_t {
                    const uint32_t q = period_quotient(d, (uint64_t)bu, inv);
                    *rem = (uint32_t)d - q * bu;
                    return q;
                });
            }
        } else if (staged) {
            sweep([&](uint32_t i) -> uint64_t {
                return apply_field(predict(c.t0, c.t1, i), lds_field<W>(stage, fbit + i * (uint32_t)c.l, c.l),
                                   c.l);
            });
        } else {
            sweep([&](uint32_t i) -> uint64_t { return close_value(c, i); });
        }
        const uint64_t bad = __ballot(!sorted);
        if (lane == 0) sh.wfail[wid] = bad ? 1u : 0u;
        __syncthreads();
        bool all_sorted = true;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) all_sorted &= sh.wfail[w] == 0;
        __syncthreads();
        scan = all_sorted;
        if (scan) acc += cnt;
        mode = scan ? (park ? 2u : 1u) : 0u;
    }
    if (!WRITE && tid == 0) {
        sh.bK[slot] = K;
        sh.bscan[slot] = (uint8_t)mode;
        sh.bround[slot] = (uint16_t)rounds_used;
    }
    if (!WRITE && mode == 2) rounds_used += nr;
    if (scan) {
        if (WRITE && mode == 2) {
            // the parked ballots, in (round, wave, lane) order: position order
            const uint32_t r0 = sh.bround[slot];
            const uint64_t below = (1ull << lane) - 1;
            for (uint32_t r = 0; r < nr; ++r) {
                uint32_t before = 0, all = 0;
#pragma unroll
                for (int w = 0; w < kWaves; ++w) {
                    const uint32_t cw = (uint32_t)__popcll(sh.rb[(r0 + r) * kWaves + w]);
                    before += w < wid ? cw : 0;
                    all += cw;
                }
                const uint64_t bal = sh.rb[(r0 + r) * kWaves + wid];
                if ((bal >> lane) & 1) {
                    const uint64_t at = acc + before + (uint64_t)__popcll(bal & below);
                    if (at < a.cap) a.out[at] = bstart + r * kThreads + (uint32_t)tid;
                }
                acc += all;
            }
        } else if (WRITE) {
            emit_ordered<true>(a, sh, mp,
                               [&](uint32_t i, uint32_t *where) {
                                   *where = bstart + i;
                                   return close_scan_hit<W>(c, K, i, close_value(c, i), a.base, sh, 0);
                               },
                               acc);
        }
    } else {
        close_walk<W, WRITE>(a, c, sh, K, bstart, acc);
    }
    return !cut;
}

// FOR_int<T>::filter_range (:111-158) and filter_range_close (:160-216) on
// block b: every position, value = min + field (or the raw value); a packed
// block whose max <= filter reports nothing (filter_range).  Blocks of width
// w-1 are packed, not raw (the reference's `maxerror >= w-1` raw branch reads
// them as raw values).
template <typename T, int W, bool CLOSE, bool WRITE>
__device__ __forceinline__ void for_block(const FilterArgs &a, FilterShared &sh, uint64_t b, uint32_t m,
                                          uint32_t bstart, uint64_t &acc) {
    constexpr uint64_t wmask = W == 32 ? 0xFFFFFFFFull : ~0ull;
    constexpr uint32_t tb = W / 8;
    const uint8_t *blk = a.stream + a.offs[b];
    const int l = blk[0];
    const bool raw = l >= W;
    const uint64_t f1 = a.filter & wmask, f2 = a.filter2 & wmask;
    const uint64_t base_v = (uint64_t)(int64_t)a.base & wmask;
    uint64_t lo = 0;
    if (!raw) {
        lo = load_u64_any(blk + 1) & wmask;
        const uint64_t hi = load_u64_any(blk + 1 + tb) & wmask;
        if (!CLOSE && hi <= f1) return;
    }
    emit_ordered<WRITE>(a, sh, m,
                        [&](uint32_t i, uint32_t *where) {
                            *where = bstart + i;
                            uint64_t v;
                            if (raw) v = load_bits(blk + 1, (uint64_t)i * W, W);
                            else v = l ? (lo + load_bits(blk + 1 + 2 * tb, (uint64_t)i * l, l)) & wmask : lo;
                            if (!CLOSE) return v > f1;
                            const uint64_t r = v % base_v;
                            return r > f1 && r < f2;
                        },
                        acc);
}

// Match bitmaps of the filters' tiles (kFilterTile positions at most) and the
// group sums that place each tile's output.
constexpr uint32_t kGtWords = kFilterTile / 32;  // 32-bit words of a tile's bitmap
static_assert(kGtWords == kThreads, "one bitmap word per thread");
constexpr uint32_t kGtGroup = kGtGroupTiles;     // tiles per group sum
__host__ __device__ inline uint64_t gt_groups(uint64_t tiles) { return (tiles + kGtGroup - 1) / kGtGroup; }

// Positions of one tile's bitmap written to out[base ..] in position order.
// Thread tid holds bitmap word tid (positions start + 32*tid ..) and `mine`,
// the hits in the words before it; `total` is the tile's hit count.  Each
// wave expands its own 64 words two at a time (64 positions per step: word
// pair and prefix by lane reads, the slot by mbcnt) into LDS (buf: the whole
// tile when its hits fit kEmitBuf entries, else each half of the words in
// turn, at most kEmitBuf positions each), and the buffer is then copied out
// with 8-byte stores of two positions (not 16-byte ones: on gfx950 a 16-byte
// store whose data registers an LDS load overwrites right after it, as the
// copy loop's next iteration would, can write the load's data as its upper
// half).  Only out[.. cap) is written.
constexpr uint32_t kEmitBuf = kFilterTile / 2;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void emit_bitmap(const FilterArgs &a, uint32_t word, uint32_t mine, uint32_t total,
                                            uint64_t base, uint32_t start, uint32_t *buf, uint32_t *half) {
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    if (tid == (int)(kGtWords / 2)) *half = mine;  // hits in the first half of the words
    __syncthreads();
    const uint32_t split = *half;
    const bool whole = total <= kEmitBuf;
    for (int h = 0; h < (whole ? 1 : 2); ++h) {
        const uint32_t h0 = whole || h == 0 ? 0u : split;     // hits before this pass
        const uint32_t h1 = whole || h == 1 ? total : split;  // ... and up to its end
        // waves 0-1 hold the first half of the words, waves 2-3 the second
        if (whole || (h == 0) == (wid < kWaves / 2)) {
            const uint32_t p0 = start + (uint32_t)wid * (kWave * 32);
#pragma unroll 4
            for (int j = 0; j < kWave / 2; ++j) {
                const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)word, 2 * j);
                const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)word, 2 * j + 1);
                if ((lo | hi) == 0) continue;
                const uint32_t before = (uint32_t)__builtin_amdgcn_readlane((int)mine, 2 * j) - h0;
                if (((lane < 32 ? lo >> lane : hi >> (lane - 32)) & 1) != 0)
                    buf[before + __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u))] =
                        p0 + (uint32_t)j * kWave + (uint32_t)lane;
            }
        }
        __syncthreads();
        const uint64_t dst = base + h0;
        uint32_t n = h1 - h0;
        if (dst >= a.cap) n = 0;
        else if (a.cap - dst < n) n = (uint32_t)(a.cap - dst);
        uint32_t *o = a.out + dst;
        const uint32_t head = (((uintptr_t)o >> 2) & 1) && n ? 1u : 0u;  // an entry before an 8-byte boundary
        if (tid == 0 && head) o[0] = buf[0];
        const uint32_t pairs = (n - head) / 2;
        u32x2 *o2 = (u32x2 *)(o + head);
        for (uint32_t q = tid; q < pairs; q += kThreads) {
            const uint32_t i = head + 2 * q;
            __builtin_nontemporal_store(u32x2{buf[i], buf[i + 1]}, &o2[q]);
        }
        const uint32_t done = head + 2 * pairs;
        if ((uint32_t)tid < n - done) o[done + tid] = buf[done + tid];
        __syncthreads();  // buf is read before the next pass refills it
    }
}

// The close filters, MODE 1 (LeCo filter_range_close) and 3 (FOR
// filter_range_close), in two launches with no serial dependence between
// tiles (a tile is filter_tile_blocks() whole blocks):
//   mark: pass 1 of each tile, in workgroup index order -- the tile's count
//         goes to the workspace and to the sum of its group of kGtGroupTiles
//         tiles.  A tile whose blocks were all swept in scan mode with their
//         ballots parked (LeCo blocks with fields, sorted values, disjoint
//         periods: the reference's period order is position order there) and
//         that spans at most kFilterTile positions also writes its match
//         bitmap, built from those ballots.
//   emit: a tile's offset is the sum of the groups before it plus the counts
//         before it in its group.  A bitmap tile expands its bitmap; any
//         other tile runs pass 1 again (LeCo: the blocks' period counts and
//         modes) and pass 2, the reference's order, at that offset.
// Workspace: the group sums, the tiles' counts (bit 63: the tile has a
// bitmap), then a kGtWords-word bitmap per tile.
constexpr uint64_t kBitmapTile = 1ull << 63;

template <typename T, int W, int MODE>
__global__ __launch_bounds__(kThreads) void k_filter_close_mark(FilterArgs a) {
    __shared__ FilterShared sh;
    uint64_t *const stage = sh.stage;  // the LeCo close scan's field staging, then the tile's bitmap
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t t = blockIdx.x;
    const uint64_t nb = (a.n + a.bs - 1) / a.bs;
    const uint64_t b0 = t * a.kb;
    const uint32_t kbt = (uint32_t)(nb - b0 < a.kb ? nb - b0 : a.kb);
    auto block_len = [&](uint64_t b) -> uint32_t {
        const uint64_t s = b * a.bs;
        return (uint32_t)(a.n - s < a.bs ? a.n - s : a.bs);
    };
    uint64_t mine = 0;
    bool ok = true;
    uint32_t rounds_used = 0;
    for (uint32_t k = 0; k < kbt; ++k) {
        const uint64_t b = b0 + k;
        if (MODE == 1) ok &= close_block<T, W, false>(a, sh, stage, b, block_len(b), k, 0, mine, rounds_used);
        else for_block<T, W, true, false>(a, sh, b, block_len(b), 0, mine);
    }
    for (int o = kWave / 2; o > 0; o >>= 1) mine += (uint64_t)__shfl_xor((long long)mine, o, kWave);
    if (lane == 0) sh.wsum[wid] = mine;
    __syncthreads();
    uint64_t total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) total += sh.wsum[w];
    // a bitmap tile: every block has fields (0 < l < W) and either has no periods or was
    // swept in scan mode with its ballots parked
    bool bm = MODE == 1 && (uint64_t)kbt * a.bs <= kFilterTile;
    for (uint32_t k = 0; k < kbt && bm; ++k) {
        const int l = sh.bl[k];
        bm = l > 0 && l < W && (sh.bK[k] == 0 || sh.bscan[k] == 2);
    }
    const uint64_t groups = gt_groups(a.tiles);
    if constexpr (MODE == 1) {
      if (bm) {
        uint32_t *const dst = (uint32_t *)(a.work + groups + a.tiles) + t * kGtWords;
        if (a.bs % kWave == 0) {
            // every parked ballot covers 64 positions from a multiple of 64 of the tile, so the
            // bitmap's 64-bit word j (positions 64j ..) is one ballot, or zero
            if (tid < (int)(kGtWords / 2)) {
                const uint32_t q = (uint32_t)tid * kWave, k = q / a.bs, off = q - k * a.bs;
                uint64_t word = 0;
                if (k < kbt && sh.bK[k] != 0 && off < block_len(b0 + k))
                    word = sh.rb[(sh.bround[k] + off / kThreads) * kWaves + (off % kThreads) / kWave];
                ((uint2 *)dst)[tid] = make_uint2((uint32_t)word, (uint32_t)(word >> 32));
            }
        } else {
            uint32_t *bits = (uint32_t *)stage;  // pass 1 is done with the staged fields
            bits[tid] = 0;
            __syncthreads();
            for (uint32_t k = 0; k < kbt; ++k) {
                if (sh.bK[k] == 0) continue;
                const uint32_t nr = (block_len(b0 + k) + kThreads - 1) / kThreads, r0 = sh.bround[k];
                for (uint32_t r = 0; r < nr; ++r) {
                    // the wave's 64 positions start at tile bit q: lanes 0-2 OR the ballot, shifted
                    // by q % 32, into words q/32 .. q/32+2 (nonzero parts stay inside the tile)
                    const uint64_t bal = sh.rb[(r0 + r) * kWaves + wid];
                    const uint32_t q = k * a.bs + r * kThreads + (uint32_t)wid * kWave, sft = q & 31;
                    if (bal && lane < 3) {
                        const uint32_t part = lane == 0   ? (uint32_t)(bal << sft)
                                              : lane == 1 ? (uint32_t)((bal << sft) >> 32)
                                                          : (sft ? (uint32_t)(bal >> (64 - sft)) : 0u);
                        if (part) atomicOr(&bits[(q >> 5) + lane], part);
                    }
                }
            }
            __syncthreads();
            dst[tid] = bits[tid];
        }
      }
    }
    if (tid == 0) {
        a.work[groups + t] = total | (bm ? kBitmapTile : 0);
        if (total) atomicAdd((unsigned long long *)&a.work[t / kGtGroup], (unsigned long long)total);
        if (!ok) atomicOr((unsigned long long *)&a.result[1], 1ull);
    }
}

template <typename T, int W, int MODE>
__global__ __launch_bounds__(kThreads) void k_filter_close_emit(FilterArgs a) {
    __shared__ FilterShared sh;
    uint64_t *const stage = sh.stage;  // the LeCo close scan's field staging
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t t = blockIdx.x;
    const uint64_t nb = (a.n + a.bs - 1) / a.bs;
    const uint64_t b0 = t * a.kb;
    const uint32_t kbt = (uint32_t)(nb - b0 < a.kb ? nb - b0 : a.kb);
    auto block_len = [&](uint64_t b) -> uint32_t {
        const uint64_t s = b * a.bs;
        return (uint32_t)(a.n - s < a.bs ? a.n - s : a.bs);
    };
    const uint64_t groups = gt_groups(a.tiles);
    // offset: the groups before t's group, then the tiles before t in its group
    const uint64_t gi = t / kGtGroup;
    uint64_t part = 0;
    for (uint64_t i = tid; i < gi; i += kThreads) part += a.work[i];
    if (gi * kGtGroup + tid < t) part += a.work[groups + gi * kGtGroup + tid] & ~kBitmapTile;
    for (int o = kWave / 2; o > 0; o >>= 1) part += (uint64_t)__shfl_xor((long long)part, o, kWave);
    if (lane == 0) sh.wsum[wid] = part;
    __syncthreads();
    uint64_t base = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) base += sh.wsum[w];
    const uint64_t cw = a.work[groups + t], count = cw & ~kBitmapTile;
    if (t + 1 == a.tiles && tid == 0) a.result[0] = base + count;
    __syncthreads();  // wsum is read; pass 1 below reuses the shared state
    if (!count) return;
    if constexpr (MODE == 1) {
      if (cw & kBitmapTile) {
        const uint32_t word = ((const uint32_t *)(a.work + groups + a.tiles))[t * kGtWords + tid];
        uint32_t all;
        const uint32_t mine = wg_exclusive_scan(sh, (uint32_t)__popc(word), &all);  // ends with a barrier
        static_assert(sizeof(FilterShared) >= kEmitBuf * 4 + 4, "the shared state holds the emit buffer");
        uint32_t *const buf = (uint32_t *)&sh;  // the shared state is free from here on
        emit_bitmap(a, word, mine, all, base, a.pos_base + (uint32_t)(b0 * a.bs), buf, buf + kEmitBuf);
        return;
      }
    }
    uint64_t pos = base;
    uint32_t rounds_used = 0;
    if (MODE == 1) {  // the blocks' period counts and sweep modes, as mark found them
        uint64_t cnt = 0;
        for (uint32_t k = 0; k < kbt; ++k)
            (void)close_block<T, W, false>(a, sh, stage, b0 + k, block_len(b0 + k), k, 0, cnt, rounds_used);
    }
    for (uint32_t k = 0; k < kbt; ++k) {
        const uint64_t b = b0 + k;
        const uint32_t bstart = a.pos_base + (uint32_t)(b * a.bs);
        if (MODE == 1) close_block<T, W, true>(a, sh, stage, b, block_len(b), k, bstart, pos, rounds_used);
        else for_block<T, W, true, true>(a, sh, b, block_len(b), bstart, pos);
    }
}

// ---- filter_range of LeCo (:206-249) and FOR (:111-158) columns ------------
// filter_range in three launches, with no serial dependence between tiles.
// A tile is gt_tile_blocks() whole blocks (blocks of up to kFilterTile
// values) or a kFilterTile-value chunk of one larger block: at most
// kFilterTile consecutive positions.
//   mark:  the tile's field bytes are staged through LDS when they fit; block
//          by block (the header is workgroup-uniform) the scanned window is
//          tested once, in rounds of kThreads consecutive positions
//          (neighbouring lanes read neighbouring fields); each wave ORs its
//          round ballot into the tile's LDS bitmap at the positions' offset
//          from the tile start.  The bitmap (kFilterTile bits) and the tile's
//          count go to the workspace, and the count is added to the sum of
//          its group of kGtGroup tiles.
//   emit:  a tile's output offset is the sum of the groups before its group
//          plus the counts of the tiles before it in its group, read in
//          parallel; every hit is written at that offset + the hits before it
//          in the bitmap -- position order -- with coalesced stores.  The last
//          tile writes the total to result[0].
// The workspace holds the group sums (zeroed first), the tiles' counts, then
// the tiles' bitmaps.
constexpr uint32_t kGtStage = 16384;

// The tile's first position and length: blocks [b0, b0 + kbt) whole, or
// values [c0, c1) of block b0.
struct GtTile {
    uint64_t b0;
    uint32_t kbt, c0, c1;
    bool chunked;
};
__device__ __forceinline__ GtTile gt_tile(const FilterArgs &a, uint64_t t) {
    GtTile g{0, 0, 0, 0, a.bs > kFilterTile};
    const uint64_t nb = (a.n + a.bs - 1) / a.bs;
    if (g.chunked) {
        const uint32_t cpb = (a.bs + kFilterTile - 1) / kFilterTile;
        g.b0 = t / cpb;
        g.kbt = 1;
        const uint64_t s = g.b0 * a.bs;
        const uint32_t m = (uint32_t)(a.n - s < a.bs ? a.n - s : a.bs);
        g.c0 = (uint32_t)(t - g.b0 * cpb) * kFilterTile;
        g.c1 = g.c0 + kFilterTile < m ? g.c0 + kFilterTile : m;
    } else {
        g.b0 = t * a.kb;
        g.kbt = (uint32_t)(nb - g.b0 < a.kb ? nb - g.b0 : a.kb);
    }
    return g;
}

struct GtShared {
    uint64_t a0;
    uint32_t words;
    uint32_t wcnt[kWaves];
    uint64_t wsum[kWaves];
    uint32_t bits[kGtWords];
};



template <typename T, int W, bool FOR>
__global__ __launch_bounds__(kThreads) void k_filter_gt_mark(FilterArgs a) {
    __shared__ __align__(16) uint32_t lds[kGtStage / 4];
    __shared__ GtShared sh;
    constexpr uint64_t wmask = W == 32 ? 0xFFFFFFFFull : ~0ull;
    constexpr uint32_t tb = W / 8;
    constexpr uint32_t hdr = FOR ? 1 + 2 * tb : 17;  // payload offset of a packed block
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t t = blockIdx.x;
    const GtTile g = gt_tile(a, t);
    const uint64_t tile_start = g.b0 * a.bs + g.c0;
    sh.bits[tid] = 0;
    if (tid == 0) {
        uint64_t g0 = a.offs[g.b0], g1 = a.offs[g.b0 + g.kbt];
        if (g.chunked) {  // the chunk's fields only
            const uint8_t *blk = a.stream + g0;
            const int l = blk[0];
            const uint32_t lf = l >= W ? W : (uint32_t)l;
            const uint64_t pay = g0 + (l >= W ? 1 : hdr);
            g1 = pay + (((uint64_t)g.c1 * lf + 7) >> 3);
            g0 = pay + (((uint64_t)g.c0 * lf) >> 3);
        }
        const uint64_t a0 = g0 & ~15ull, a1 = ((g1 + 15) & ~15ull) + 16;
        sh.a0 = a0;
        sh.words = a1 - a0 <= kGtStage ? (uint32_t)((a1 - a0) / 4) : 0u;
    }
    __syncthreads();
    const uint64_t a0 = sh.a0;
    const uint32_t words = sh.words;
    for (uint32_t w = tid; w < words / 4; w += kThreads)
        ((uint4 *)lds)[w] = *(const uint4 *)(a.stream + a0 + 16ull * w);
    __syncthreads();

    const uint64_t f = a.filter & wmask;
    for (uint32_t kk = 0; kk < g.kbt; ++kk) {
        const uint64_t b = g.b0 + kk;
        const uint64_t off = a.offs[b];
        const uint8_t *blk = a.stream + off;
        const uint64_t s = b * a.bs;
        const uint32_t m = (uint32_t)(a.n - s < a.bs ? a.n - s : a.bs);
        const int l = __builtin_amdgcn_readfirstlane(blk[0]);
        const bool raw = l >= W;
        const uint32_t lf = raw ? W : (uint32_t)l;
        const uint64_t pay = raw ? 1 : hdr;  // payload byte offset in the block
        // window [lo, hi) of positions the reference scans
        uint32_t lo = g.chunked ? g.c0 : 0, hi = g.chunked ? g.c1 : m;
        double t0 = 0.0, t1 = 0.0;
        uint64_t base_v = 0;
        if (raw) {
            if (!FOR) {  // LeCo's one- and two-value raw blocks (127 / 126)
                const uint32_t mm = l == 127 ? (m < 1 ? m : 1) : l == 126 ? (m < 2 ? m : 2) : m;
                hi = hi < mm ? hi : mm;
            }
        } else if (FOR) {
            base_v = load_u64_any(blk + 1) & wmask;
            if ((load_u64_any(blk + 1 + tb) & wmask) <= f) hi = lo;  // max <= filter: nothing
        } else {
            t0 = bits_to_double(load_u64_any(blk + 1));
            t1 = bits_to_double(load_u64_any(blk + 9));
            const uint32_t w0 = gt_window<W>(f, l, t0, t1, m);
            lo = lo > w0 ? lo : w0;
        }
        // bit offset of field 0 of this block in LDS (staged) or in the block
        const int64_t fbit = words ? ((int64_t)(off + pay) - (int64_t)a0) * 8 : (int64_t)pay * 8;
        // `staged`: std::true_type (fields from the LDS copy) or std::false_type
        // (from global memory), so that each round loop holds only its own reads
        auto field = [&](uint32_t j, auto staged) -> uint64_t {
            if constexpr (!decltype(staged)::value) return load_bits(blk, (uint64_t)fbit + (uint64_t)j * lf, (int)lf);
            const uint32_t q = (uint32_t)(fbit + (int64_t)j * lf), w = q >> 5, sft = q & 31;
            const uint32_t x0 = lds[w], x1 = lds[w + 1];
            uint64_t v = __builtin_amdgcn_alignbit(x1, x0, sft);
            if (W == 64) v |= (uint64_t)__builtin_amdgcn_alignbit(lds[w + 2], x1, sft) << 32;
            return v & low_mask((int)lf);
        };
        const uint32_t rel = (uint32_t)(s - tile_start);  // tile bit of block position 0
        auto round = [&](uint32_t p, auto staged) {
            const uint32_t j = p + tid;
            bool h = false;
            if (j < hi) {
                uint64_t v;
                if (raw) v = field(j, staged);
                else if (FOR) v = l ? (base_v + field(j, staged)) & wmask : base_v;
                else {
                    const int64_t pr = predict(t0, t1, j);
                    v = l ? apply_field(pr, field(j, staged), l) : (uint64_t)pr;
                }
                h = (raw || FOR) ? v > f : above<W>(v, f);
            }
            // the wave's 64 positions start at tile bit q: lanes 0-2 OR the ballot, shifted
            // by q % 32, into words q/32 .. q/32+2.  Bits of positions past the window are 0,
            // so a nonzero part never falls past the tile's kGtWords words.
            const uint64_t bal = __ballot(h);
            const uint32_t q = rel + p + (uint32_t)wid * kWave, sft = q & 31;
            if (bal && lane < 3) {
                const uint32_t part = lane == 0   ? (uint32_t)(bal << sft)
                                      : lane == 1 ? (uint32_t)((bal << sft) >> 32)
                                                  : (sft ? (uint32_t)(bal >> (64 - sft)) : 0u);
                if (part) atomicOr(&sh.bits[(q >> 5) + lane], part);
            }
        };
        // The ORs are done before any later load: a queued LDS atomic can read its
        // registers after a later load has refilled them.  Fields from LDS: one wait
        // after the block (before the next block's header loads); fields from
        // global memory: a wait after every round (separate loops, so that no
        // path leads from an OR to a load without one).
        if (words) {
            for (uint32_t p = lo; p < hi; p += kThreads) round(p, std::true_type());
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        } else {
            for (uint32_t p = lo; p < hi; p += kThreads) {
                round(p, std::false_type());
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            }
        }
    }
    __syncthreads();
    const uint32_t word = sh.bits[tid];
    uint32_t total;
    (void)wg_exclusive_scan(sh, (uint32_t)__popc(word), &total);
    const uint64_t groups = gt_groups(a.tiles);
    uint32_t *bm = (uint32_t *)(a.work + groups + a.tiles) + t * kGtWords;
    bm[tid] = word;
    if (tid == 0) {
        a.work[groups + t] = total;
        if (total) atomicAdd((unsigned long long *)&a.work[t / kGtGroup], (unsigned long long)total);
    }
}

// The positions of tile t's bitmap, at its offset (position order).
__global__ __launch_bounds__(kThreads) void k_filter_gt_emit(FilterArgs a) {
    __shared__ GtShared sh;
    __shared__ uint32_t obuf[kEmitBuf];
    __shared__ uint32_t half;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t t = blockIdx.x;
    const GtTile g = gt_tile(a, t);
    const uint32_t start = a.pos_base + (uint32_t)(g.b0 * a.bs + g.c0);  // uint32 position arithmetic
    const uint64_t groups = gt_groups(a.tiles);
    // offset: the groups before t's group, then the tiles before t in its group
    const uint64_t gi = t / kGtGroup;
    uint64_t part = 0;
    for (uint64_t i = tid; i < gi; i += kThreads) part += a.work[i];
    if (gi * kGtGroup + tid < t) part += a.work[groups + gi * kGtGroup + tid];
    for (int o = kWave / 2; o > 0; o >>= 1) part += (uint64_t)__shfl_xor((long long)part, o, kWave);
    if (lane == 0) sh.wsum[wid] = part;
    const uint32_t word = ((const uint32_t *)(a.work + groups + a.tiles))[t * kGtWords + tid];
    uint32_t total;
    const uint32_t mine = wg_exclusive_scan(sh, (uint32_t)__popc(word), &total);  // (its barrier orders wsum)
    uint64_t base = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) base += sh.wsum[w];
    if (t + 1 == a.tiles && tid == 0) a.result[0] = base + total;
    if (!total) return;
    emit_bitmap(a, word, mine, total, base, start, obuf, &half);
}

__global__ __launch_bounds__(kThreads) void k_filter_clear(uint64_t *__restrict__ work, uint64_t words,
                                                           uint64_t *__restrict__ result) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < words) work[i] = 0;
    if (i == 0) result[0] = result[1] = 0;
}

template <typename T, int W, bool FOR>
void launch_gt(const FilterArgs &a, hipStream_t st) {
    const uint64_t words = gt_groups(a.tiles);  // the group sums start at 0
    k_filter_clear<<<dim3((unsigned)((words + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(a.work, words,
                                                                                                    a.result);
    k_filter_gt_mark<T, W, FOR><<<dim3((unsigned)a.tiles), dim3(kThreads), 0, st>>>(a);
    k_filter_gt_emit<<<dim3((unsigned)a.tiles), dim3(kThreads), 0, st>>>(a);
}

template <typename T, int W, int MODE>
void launch_close(const FilterArgs &a, hipStream_t st) {
    const uint64_t words = gt_groups(a.tiles);  // the group sums start at 0
    k_filter_clear<<<dim3((unsigned)((words + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(a.work, words,
                                                                                                    a.result);
    k_filter_close_mark<T, W, MODE><<<dim3((unsigned)a.tiles), dim3(kThreads), 0, st>>>(a);
    k_filter_close_emit<T, W, MODE><<<dim3((unsigned)a.tiles), dim3(kThreads), 0, st>>>(a);
}

}  // namespace

template <typename T>
hipError_t launch_filter(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, bool close,
                         uint64_t filter, const T *f1, const T *f2, int32_t base, uint32_t pos_base, uint32_t *out,
                         uint64_t cap, uint64_t *result, uint64_t *work, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    FilterArgs a{stream, offs, n, bs, filter_tile_blocks(bs), close ? filter_tiles(n, bs) : gt_tiles(n, bs), filter,
                 0, f1, f2, base, pos_base, out, cap, result, work};
    if (close) {
        launch_close<T, W, 1>(a, st);
    } else {
        launch_gt<T, W, false>(a, st);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_for_filter(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, bool close,
                             uint64_t f1, uint64_t f2, int32_t base, uint32_t pos_base, uint32_t *out, uint64_t cap,
                             uint64_t *result, uint64_t *work, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    FilterArgs a{stream, offs, n, bs, filter_tile_blocks(bs), close ? filter_tiles(n, bs) : gt_tiles(n, bs), f1, f2,
                 nullptr, nullptr, base, pos_base, out, cap, result, work};
    if (close) {
        launch_close<T, W, 3>(a, st);
    } else {
        launch_gt<T, W, true>(a, st);
    }
    return hipGetLastError();
}

template hipError_t launch_for_filter<uint32_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, bool, uint64_t,
                                                uint64_t, int32_t, uint32_t, uint32_t *, uint64_t, uint64_t *,
                                                uint64_t *, hipStream_t);
template hipError_t launch_for_filter<uint64_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, bool, uint64_t,
                                                uint64_t, int32_t, uint32_t, uint32_t *, uint64_t, uint64_t *,
                                                uint64_t *, hipStream_t);

template hipError_t launch_filter<uint32_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, bool, uint64_t,
                                            const uint32_t *, const uint32_t *, int32_t, uint32_t, uint32_t *,
                                            uint64_t, uint64_t *, uint64_t *, hipStream_t);
template hipError_t launch_filter<uint64_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, bool, uint64_t,
                                            const uint64_t *, const uint64_t *, int32_t, uint32_t, uint32_t *,
                                            uint64_t, uint64_t *, uint64_t *, hipStream_t);

}  // namespace leco

// ---- filter_range with the library's static scratch (allocates nothing) ----
// The workspace of filter_range (group sums, tile counts, match bitmaps) lives
// in `scratch`, a file-scope device array that every GPU the library runs on
// holds until the process ends.  Host code never passes `scratch` itself (its
// host name is only a handle for the symbol): the device address comes from
// hipGetSymbolAddress, once per device, with that device current, and is kept
// per device.  Every call on one GPU shares that GPU's array; the clear kernel
// at the start of each call zeroes what must start at zero (the group sums and
// the result), so calls in a row need nothing else, but calls that overlap on
// one GPU race on it.
namespace {

constexpr int kScratchMaxDevices = 64;
__device__ __align__(256) uint64_t scratch[LECO_FIX_NOALLOC_SCRATCH_BYTES / 8];
std::atomic<void *> g_scratch_ptr[kScratchMaxDevices];

// The current device's address of `scratch`, or null if the runtime fails.
void *scratch_address(int *rc) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { *rc = LECO_FIX_EHIP; return nullptr; }
    if (dev < 0 || dev >= kScratchMaxDevices) { *rc = LECO_FIX_EINVAL; return nullptr; }
    void *p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (p) return p;
    // the same address whichever thread asks first, so a race only repeats the call
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) { *rc = LECO_FIX_EHIP; return nullptr; }
    g_scratch_ptr[dev].store(p, std::memory_order_release);
    return p;
}

}  // namespace

extern "C" int leco_fix_filter_range_noalloc(uint32_t width_bits, const uint8_t *d_stream, const uint64_t *d_offsets,
                                             uint64_t n, uint32_t block_size, uint64_t filter, uint32_t *d_out,
                                             uint64_t out_capacity, uint64_t *d_result, void *stream) {
    // every argument check comes before the first HIP call
    if ((width_bits != 32 && width_bits != 64) || n == 0 || n > (1ull << 32) || block_size == 0 ||
        block_size > leco::kMaxBlockSize || !d_stream || !d_offsets || !d_result || (out_capacity && !d_out) ||
        (width_bits == 32 && filter > 0xFFFFFFFFull))
        return LECO_FIX_EINVAL;
    const uint64_t need = leco::filter_workspace_for(leco::gt_tiles(n, block_size));
    if (need > LECO_FIX_NOALLOC_SCRATCH_BYTES) return LECO_FIX_ESCRATCH;
    int rc = LECO_FIX_OK;
    uint64_t *work = (uint64_t *)scratch_address(&rc);
    if (!work) return rc;
    const hipStream_t st = (hipStream_t)stream;
    const hipError_t e =
        width_bits == 64
            ? leco::launch_filter<uint64_t>(d_stream, d_offsets, n, block_size, false, filter, nullptr, nullptr, 0, 0,
                                            d_out, out_capacity, d_result, work, st)
            : leco::launch_filter<uint32_t>(d_stream, d_offsets, n, block_size, false, filter, nullptr, nullptr, 0, 0,
                                            d_out, out_capacity, d_result, work, st);
    return e == hipSuccess ? LECO_FIX_OK : LECO_FIX_EHIP;
}
