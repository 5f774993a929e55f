// Types shared by host code and kernels, and the kernel launchers
// (defined in leco_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <new>
#include <stdexcept>

#include "leco_fix.h"

namespace leco {

// C ABI entry points whose host code allocates (std containers) run their
// bodies through this: a C++ exception must not unwind into a C caller.
// Allocation failures (and sizes past a container's max_size) are ENOMEM.
template <class F>
inline int abi_guard(F &&body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return LECO_FIX_ENOMEM;
    } catch (const std::length_error &) {
        return LECO_FIX_ENOMEM;
    } catch (...) {
        return LECO_FIX_EINVAL;
    }
}

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
// Largest byte size of a block of n values (raw or l = w-1).
__host__ __device__ inline uint64_t block_bound(uint32_t w, uint32_t n) {
    uint64_t a = block_bytes(w - 1, w, n), b = block_bytes(w, w, n);
    return a > b ? a : b;
}

// Blocks of at most kWaveBlock values are encoded one per wave, a workgroup
// taking encode_unit_blocks() consecutive blocks of a column per ticket (about
// 4 * kWaveBlock values, at most kUnitMaxBlocks blocks); larger blocks one per
// workgroup.
constexpr uint32_t kWaveBlock = 1024;
constexpr uint32_t kUnitMaxBlocks = 64;
__host__ __device__ inline uint32_t encode_unit_blocks(uint32_t bs) {
    if (bs > kWaveBlock) return 1;
    const uint32_t per = 4 * (kWaveBlock / bs);
    return per < kUnitMaxBlocks ? per : kUnitMaxBlocks;
}
__host__ __device__ inline uint64_t encode_units(uint64_t nb, uint32_t bs) {
    const uint32_t per = encode_unit_blocks(bs);
    return (nb + per - 1) / per;
}

// Fixed-partition baselines (FOR_integer_template.h, delta_integer_template.h):
// block sizes of write_FOR_int_T (E+1 bytes, then (32 - l)/8 zero bytes in C
// integer division) and write_delta_int_T (n fields: E+1 bytes, then 8 - E%8).
constexpr int kCodecLeco = 0;  // the multi-block decode body's codec parameter: LeCo blocks
constexpr int kCodecFor = 1;
constexpr int kCodecDelta = 2;
__host__ __device__ inline uint64_t for_block_bytes(uint32_t l, uint32_t w, uint32_t n) {
    if (l >= w) return 1 + (uint64_t)n * (w / 8);
    if (l == 0) return 1 + 2 * (w / 8);
    const uint64_t e = ((uint64_t)l * n + 7) / 8;
    const int pad = (32 - (int)l) / 8;
    return 1 + 2 * (w / 8) + e + 1 + (pad > 0 ? pad : 0);
}
__host__ __device__ inline uint64_t delta_block_bytes(uint32_t l, uint32_t w, uint32_t n) {
    if (l >= w) return 1 + (uint64_t)n * (w / 8);
    if (l == 0) return 1 + w / 8;
    const uint64_t e = ((uint64_t)l * n + 7) / 8;
    return 1 + w / 8 + e + 1 + (8 - e % 8);
}

// piecewise_fix_op_max (uint32): block = [u8 tmp_bit][f64 theta0][f32 theta1]
// then nothing (tmp_bit 0), n raw uint32 (tmp_bit >= 31) or n tmp_bit-bit
// fields, E+1 bytes and (32 - l)/8 zero bytes (write_delta_T, bit_write.h:186-247).
// Encoded one block per wave, opmax_unit_blocks() consecutive blocks per
// workgroup and look-back unit; decoded kOpmaxRun consecutive values per thread.
constexpr uint32_t kOpmaxHdr = 13;
constexpr uint32_t kOpmaxRun = 16;
constexpr uint32_t kFixTile = 8192;  // fix-up list compaction: flags per workgroup
__host__ __device__ inline uint64_t opmax_block_bytes(uint32_t l, uint32_t n) {
    if (l == 0) return kOpmaxHdr;
    if (l >= 31) return kOpmaxHdr + 4ull * n;
    return kOpmaxHdr + ((uint64_t)l * n + 7) / 8 + 1 + (32 - l) / 8;
}
__host__ __device__ inline uint64_t opmax_block_bound(uint32_t n) {
    const uint64_t a = opmax_block_bytes(30, n), b = opmax_block_bytes(32, n);
    return a > b ? a : b;
}
__host__ __device__ inline uint32_t opmax_unit_blocks(uint32_t bs) {
    const uint32_t per = bs > kWaveBlock ? 4 : 4 * (kWaveBlock / bs);
    return per < kUnitMaxBlocks ? per : kUnitMaxBlocks;
}

// One column of a batch, as the device sees it.
struct BatchColumn {
    const void *in;   // values to encode (device)
    void *out;        // decoded values (device)
    uint64_t n;
    uint32_t bs;
    uint32_t width;   // 32 or 64
    uint32_t kb;      // blocks per multi-block decode tile (blocks below kStagedMinBlock values)
    uint32_t reserved;
};

// A resident column of the request server.
struct ServeColumn {
    const uint8_t *stream;
    const uint64_t *offsets;
    const leco_fix_block_param *params;
    uint64_t n;
    uint32_t bs;
    uint32_t width;
};

enum : uint32_t { kServeLookup = 0, kServeRange = 1, kServeRangeValues = 2 };

// One workgroup's share of a served batch: a chunk of a point-lookup request,
// one block of a range-decode request (staged columns), or kValueTile
// positions of a range-decode request (value-parallel columns: positions
// [lo, lo + count) go to out[0..count)).
struct ServeItem {
    uint32_t kind;
    uint32_t column;
    uint32_t count;        // lookup: indices in this chunk
    uint32_t reserved;
    const uint64_t *idx;   // lookup: indices
    void *out;             // lookup: results; range: where value `lo` goes
    uint64_t block;        // range: block id
    uint64_t lo, hi;       // range: [lo, hi) in column positions
};

// Decode shape for a block size: full decodes of blocks of at least this many
// values are staged one workgroup per block through LDS; smaller blocks are
// decoded multi_blocks() at a time per workgroup.
constexpr uint32_t kStagedMinBlock = 2048;
__host__ __device__ inline bool staged_decode(uint32_t bs) { return bs >= kStagedMinBlock; }
// Positions per value-parallel tile (one workgroup).
constexpr uint32_t kValueTile = 4096;
// Small blocks: as many consecutive blocks per workgroup as keep the tile
// within kMultiPositions positions, at most kMultiMax.  Their bytes are staged
// through kMultiLds bytes of LDS, whole blocks at a time (one block of fewer
// than kStagedMinBlock values always fits).  kMultiMax = 128 keeps the
// workgroup's LDS (staging plus block tables) near 20 KB, so that 8
// workgroups fit a CU's 160 KB; it only shortens the tiles of blocks of
// fewer than 64 values.
constexpr uint32_t kMultiPositions = 8192;
constexpr uint32_t kMultiMax = 128;
constexpr uint32_t kMultiLds = 2048 * 8 + 32;
static_assert(kMultiLds >= 1 + (kStagedMinBlock - 1) * 8 + 15, "a small block fits the staging buffer");
__host__ __device__ inline uint32_t multi_blocks(uint32_t bs, uint32_t /*w*/, uint32_t positions = kMultiPositions) {
    uint32_t k = (positions ? positions : kMultiPositions) / bs;
    if (k > kMultiMax) k = kMultiMax;
    return k ? k : 1;
}

// Filters (filter_range / filter_range_close) work on tiles of
// filter_tile_blocks() consecutive blocks (about kFilterTile positions), one
// workgroup each; a block stops after kFilterMaxPeriods periods of
// filter_range_close (flag 1 of the result).
constexpr uint32_t kFilterTile = 8192;
constexpr uint32_t kFilterMaxPeriods = 1u << 16;
__host__ __device__ inline uint32_t filter_tile_blocks(uint32_t bs) {
    const uint32_t k = kFilterTile / bs;
    return k < 1 ? 1 : (k > 256 ? 256 : k);
}
__host__ __device__ inline uint64_t filter_tiles(uint64_t n, uint32_t bs) {
    const uint64_t nb = (n + bs - 1) / bs, kb = filter_tile_blocks(bs);
    return (nb + kb - 1) / kb;
}
// filter_range (value > filter) of LeCo and FOR columns: the same tiles of
// whole blocks, except that a block of more than kFilterTile values is split
// into tiles of kFilterTile values.
__host__ __device__ inline uint64_t gt_tiles(uint64_t n, uint32_t bs) {
    if (bs <= kFilterTile) return filter_tiles(n, bs);
    return (n + bs - 1) / bs * ((bs + kFilterTile - 1) / kFilterTile);
}
// A filter launch over `tiles` tiles uses a sum per kGtGroupTiles tiles, then
// a count and a kFilterTile-bit match bitmap per tile.
constexpr uint32_t kGtGroupTiles = 256;
__host__ __device__ inline uint64_t filter_workspace_for(uint64_t tiles) {
    return 8 * ((tiles + kGtGroupTiles - 1) / kGtGroupTiles) + (8 + kFilterTile / 8) * tiles;
}

// Look-back test hooks, one per translation unit with an encoder (see
// LookbackHook in leco_kernels.h).
hipError_t set_lookback_hook_leco(uint32_t spin_limit, uint64_t withhold);
hipError_t set_lookback_hook_base(uint32_t spin_limit, uint64_t withhold);
hipError_t set_lookback_hook_opmax(uint32_t spin_limit, uint64_t withhold);

// Each launcher enqueues on `st` and returns the hipGetLastError() of the launch.
template <typename T>
hipError_t launch_encode(const T *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                         leco_fix_block_param *prm, hipStream_t st);
template <typename T>
hipError_t launch_decode(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                         uint64_t lo, uint64_t hi, T *out, hipStream_t st);
template <typename T>
hipError_t launch_decode_values(const uint8_t *stream, const leco_fix_block_param *prm, uint32_t bs, uint64_t lo,
                                uint64_t hi, T *out, hipStream_t st);
template <typename T>
hipError_t launch_lookup(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                         const uint64_t *idx, uint64_t nq, T *out, hipStream_t st);

// Grouped lookups: queries grouped by the range of 2^sh values they fall in
// (bins groups, at most kGroupMaxBins; 2^sh >= 2^20 values), then looked up
// group by group.  The workspace holds the group totals and cursors and one
// 8-byte pair (index << 31 | query) per query of a pass; a pass takes at most
// kGroupMaxQueries queries, and indices must fit 33 bits (n <= 2^33).
constexpr uint32_t kGroupChunk = 32768;  // queries per workgroup in the counting passes
constexpr uint32_t kXcds = 8;            // MI355X: accelerator dies; workgroup i runs on die i mod 8
constexpr uint32_t kGroupMaxBins = 4096;
constexpr uint64_t kGroupMaxQueries = 1ull << 31;
constexpr uint64_t kGroupQueryMask = kGroupMaxQueries - 1;
constexpr uint64_t kGroupMaxValues = 1ull << 33;
struct LookupGroups {
    uint32_t sh, bins;
};
__host__ __device__ inline LookupGroups lookup_groups(uint64_t n) {
    uint32_t sh = 20;
    while (((n + (1ull << sh) - 1) >> sh) > kGroupMaxBins) ++sh;
    return {sh, (uint32_t)((n + (1ull << sh) - 1) >> sh)};
}
inline uint64_t lookup_grouped_workspace(uint64_t nq) {
    return 2ull * kGroupMaxBins * sizeof(uint32_t) + 8ull * (nq < kGroupMaxQueries ? nq : kGroupMaxQueries);
}
template <typename T>
hipError_t launch_lookup_grouped(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                                 const uint64_t *idx, uint64_t nq, T *out, void *work, hipStream_t st);
hipError_t launch_build_params(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                               uint32_t w, leco_fix_block_param *prm, hipStream_t st);

// Batched encode of every block of ncols columns into one arena: offs holds
// total+1 arena offsets (offs[first[c] + b] for block b of column c), prm the
// total per-block params, widths total uint32 of scratch.  first[c] is the
// global index of column c's block 0; first[ncols] = total.
// fit_first[c] = index of column c's first 64-block fit group (fit_first[ncols]
// = fit_groups).
// unit_first[c] = column c's first encode unit (encode_units() per column;
// unit_first[ncols] = units), unit_col[u] = unit u's column (see
// launch_tile_columns).  widths[b]: width of global block b (see launch_batch_blocks).
hipError_t launch_batch_encode(const BatchColumn *cols, const uint64_t *first, const uint64_t *fit_first,
                               const uint64_t *unit_first, const uint32_t *unit_col, const uint32_t *widths,
                               uint32_t ncols, uint64_t total, uint64_t fit_groups, uint64_t units, uint8_t *arena,
                               uint64_t *offs, leco_fix_block_param *prm, hipStream_t st);
// Fills widths for every block of a batch (once, at creation).
hipError_t launch_batch_blocks(const BatchColumn *cols, const uint64_t *first, uint32_t ncols, uint64_t total,
                               uint32_t *widths, hipStream_t st);
// Batched decode of the tile table (tile_first[c]: column c's first tile; tile_col[t]: column of
// tile t).  Tiles [0, staged_tiles) hold one large block each, tiles [staged_tiles, tiles) runs of
// small blocks; each range is one launch of at most `grid` workgroups (0: one per tile), which
// stride over the range when there are fewer.
hipError_t launch_batch_decode(const BatchColumn *cols, const uint64_t *first, const uint64_t *tile_first,
                               const uint32_t *tile_col, uint64_t staged_tiles, uint64_t tiles, const uint8_t *arena,
                               const uint64_t *offs, uint32_t grid, hipStream_t st);
// tile_col[t] for tiles numbered in the column order ord (null: 0, 1, ...), ord_first[k] being
// the first tile of column ord[k] (monotone, ord_first[ncols] = tiles).
hipError_t launch_tile_columns(const uint64_t *ord_first, const uint32_t *ord, uint32_t ncols, uint64_t tiles,
                               uint32_t *tile_col, hipStream_t st);
// Filter over a column: filter_range (close = false: value > filter) or
// filter_range_close (per-block starting filters f1[b], f2[b], period base).
// Positions pos_base + global index (uint32) go to out[0..cap) in the
// reference's order; result[0] = count, result[1] = flags; work holds
// filter_tiles(n, bs) + 1 words.
template <typename T>
hipError_t launch_filter(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, bool close,
                         uint64_t filter, const T *f1, const T *f2, int32_t base, uint32_t pos_base, uint32_t *out,
                         uint64_t cap, uint64_t *result, uint64_t *work, hipStream_t st);
// FOR / Delta columns (codec kCodecFor or kCodecDelta): offs holds nb+1 words
// (offs[nb] = stream size after encode).
template <typename T>
hipError_t launch_base_encode(int codec, const T *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                              hipStream_t st);
// FOR_int<T> columns of blocks below kStagedMinBlock values: the LeCo
// multi-block decode body with FOR headers (tiles of multi_blocks() blocks).
template <typename T>
hipError_t launch_for_decode_multi(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, T *out,
                                   hipStream_t st);
template <typename T>
hipError_t launch_base_decode(int codec, const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                              T *out, hipStream_t st);
template <typename T>
hipError_t launch_base_lookup(int codec, const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                              const uint64_t *idx, uint64_t nq, T *out, hipStream_t st);
// FOR_int filters: filter_range (close = false: values > f1) or
// filter_range_close (v % base in (f1, f2)) over every block.
template <typename T>
hipError_t launch_for_filter(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, bool close,
                             uint64_t f1, uint64_t f2, int32_t base, uint32_t pos_base, uint32_t *out, uint64_t cap,
                             uint64_t *result, uint64_t *work, hipStream_t st);
// Served batch: `grid` workgroups stride over nitems items.
hipError_t launch_serve(const ServeColumn *cols, const ServeItem *items, uint32_t nitems, uint32_t grid,
                        hipStream_t st);

// K1 of the LeCo encoder alone: per-block fit (theta0, theta1, length) into
// prm, look-back status words offs[1..nb] and the ticket offs[0] cleared.
hipError_t launch_fit(const void *in, uint64_t n, uint32_t bs, uint32_t width, leco_fix_block_param *prm,
                      uint64_t *offs, hipStream_t st);
// piecewise_fix_op_max (leco_opmax.hip).  fix: n int8 (+1 / -1 / 0) or null.
hipError_t launch_opmax_encode(const uint32_t *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                               leco_fix_block_param *prm, int8_t *fix, hipStream_t st);
hipError_t launch_opmax_decode(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                               const int8_t *fix, uint32_t *out, hipStream_t st);
// The reference's lists applied to decoded values, raw blocks skipped.
hipError_t launch_opmax_apply_exact(uint32_t *values, uint64_t n, uint32_t bs, const uint8_t *stream,
                                    const uint64_t *offs, const int32_t *plus, uint64_t np, const int32_t *minus,
                                    uint64_t nm, hipStream_t st);
hipError_t launch_opmax_lookup(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                               const uint64_t *idx, uint64_t nq, uint32_t *out, hipStream_t st);
hipError_t launch_opmax_params(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                               leco_fix_block_param *prm, hipStream_t st);
// work: 2 * ceil(n / kFixTile) uint64; counts[0..1] = list lengths.
hipError_t launch_fix_lists(const int8_t *fix, uint64_t n, int32_t *plus, int32_t *minus, uint64_t cap,
                            uint64_t *counts, uint64_t *work, hipStream_t st);
hipError_t launch_apply_lists(uint32_t *values, uint64_t n, const int32_t *plus, uint64_t np, const int32_t *minus,
                              uint64_t nm, hipStream_t st);

}  // namespace leco
