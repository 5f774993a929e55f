// Batched encode/decode of many columns (C ABI in include/leco_fix.h) and
// the block-size autotuner.
//
// A batch keeps a device block table (column descriptors + the global index
// of each column's first block), one arena for every column's stream, and
// the arena offsets / param tables.  Encode and decode are single launches
// (plus a fit and a fix-up kernel for encode) over all blocks of all columns;
// both can be replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "leco_internal.h"

using leco::BatchColumn;

struct leco_fix_batch {
    uint32_t ncols = 0;
    uint64_t total = 0;
    std::vector<BatchColumn> h_cols;
    std::vector<uint64_t> h_first;
    std::vector<uint64_t> h_fit_first;
    uint64_t fit_groups = 0;
    BatchColumn *d_cols = nullptr;
    uint64_t *d_first = nullptr;
    uint64_t *d_fit_first = nullptr;
    uint64_t tiles = 0;            // decode tiles: one per large block, or a run of small blocks
    uint64_t staged_tiles = 0;     // tiles [0, staged_tiles) are the large blocks' (all columns)
    uint64_t *d_tile_first = nullptr;
    uint32_t *d_tile_col = nullptr;
    uint8_t *d_arena = nullptr;
    uint64_t arena_cap = 0;
    uint64_t *d_offs = nullptr;
    leco_fix_block_param *d_prm = nullptr;
    uint32_t *d_widths = nullptr;
    uint64_t units = 0;            // encode units (one ticket each)
    uint64_t *d_unit_first = nullptr;
    uint32_t *d_unit_col = nullptr;
    uint32_t decode_grid = 0;
    uint8_t *d_scratch = nullptr;  // per-column encode: stream, offsets, params of the largest column
    uint64_t scratch_cap = 0;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
};

namespace {

int hip_status(hipError_t e) { return e == hipSuccess ? LECO_FIX_OK : LECO_FIX_EHIP; }

void free_batch(leco_fix_batch *b) {
    for (auto &g : b->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    (void)hipFree(b->d_cols);
    (void)hipFree(b->d_first);
    (void)hipFree(b->d_fit_first);
    (void)hipFree(b->d_tile_first);
    (void)hipFree(b->d_tile_col);
    (void)hipFree(b->d_arena);
    (void)hipFree(b->d_offs);
    (void)hipFree(b->d_prm);
    (void)hipFree(b->d_widths);
    (void)hipFree(b->d_unit_first);
    (void)hipFree(b->d_unit_col);
    (void)hipFree(b->d_scratch);
    delete b;
}

bool valid_desc(const leco_fix_column_desc &c) {
    return c.n > 0 && c.block_size >= 1 && c.block_size <= leco::kMaxBlockSize &&
           (c.width_bits == 32 || c.width_bits == 64);
}

int enqueue(leco_fix_batch *b, int op, hipStream_t st) {
    if (op == 0)
        return hip_status(leco::launch_batch_encode(b->d_cols, b->d_first, b->d_fit_first, b->d_unit_first,
                                                    b->d_unit_col, b->d_widths, b->ncols, b->total, b->fit_groups,
                                                    b->units, b->d_arena, b->d_offs, b->d_prm, st));
    return hip_status(leco::launch_batch_decode(b->d_cols, b->d_first, b->d_tile_first, b->d_tile_col,
                                                b->staged_tiles, b->tiles, b->d_arena, b->d_offs, b->decode_grid,
                                                st));
}

}  // namespace

extern "C" {

int leco_fix_batch_create(const leco_fix_column_desc *cols, uint32_t ncols, leco_fix_batch **out) {
    return leco::abi_guard([&]() -> int {
        if (!cols || !out || ncols == 0) return LECO_FIX_EINVAL;
        for (uint32_t c = 0; c < ncols; ++c)
            if (!valid_desc(cols[c])) return LECO_FIX_EINVAL;
        leco_fix_batch *b = new (std::nothrow) leco_fix_batch;
        if (!b) return LECO_FIX_ENOMEM;
        b->ncols = ncols;
        // host tables (a failed allocation frees the batch)
        std::vector<uint64_t> tile_first, unit_first;
        std::vector<uint32_t> ord;
        std::vector<uint64_t> ord_first;
        uint64_t arena = 0;
        try {
            b->h_cols.resize(ncols);
            b->h_first.resize(ncols + 1);
            b->h_fit_first.resize(ncols + 1);
            tile_first.resize(ncols + 1);
            unit_first.resize(ncols + 1);
            for (uint32_t c = 0; c < ncols; ++c) {
                const leco_fix_column_desc &d = cols[c];
                b->h_cols[c] = BatchColumn{d.d_in, d.d_out, d.n, d.block_size, d.width_bits,
                                           leco::multi_blocks(d.block_size, d.width_bits, d.tile_positions), 0};
                b->h_first[c] = b->total;
                b->h_fit_first[c] = b->fit_groups;
                const uint64_t nb = (d.n + d.block_size - 1) / d.block_size;
                b->fit_groups += (nb + 63) / 64;
                unit_first[c] = b->units;
                b->units += leco::encode_units(nb, d.block_size);
                const uint64_t last = d.n - (nb - 1) * d.block_size;
                arena += (nb - 1) * leco::block_bound(d.width_bits, d.block_size) +
                         leco::block_bound(d.width_bits, (uint32_t)last);
                b->total += nb;
            }
            b->h_first[ncols] = b->total;
            b->h_fit_first[ncols] = b->fit_groups;
            unit_first[ncols] = b->units;
            // decode tiles: every staged column's first, then every multi column's (column order within each)
            ord.reserve(ncols);
            ord_first.reserve(ncols + 1);
            for (int staged = 1; staged >= 0; --staged) {
                for (uint32_t c = 0; c < ncols; ++c) {
                    const leco_fix_column_desc &d = cols[c];
                    if (leco::staged_decode(d.block_size) != (staged == 1)) continue;
                    const uint64_t nb = (d.n + d.block_size - 1) / d.block_size;
                    const uint32_t kb = b->h_cols[c].kb;
                    ord.push_back(c);
                    ord_first.push_back(b->tiles);
                    tile_first[c] = b->tiles;
                    b->tiles += staged ? nb : (nb + kb - 1) / kb;
                }
                if (staged) b->staged_tiles = b->tiles;
            }
            ord_first.push_back(b->tiles);
            tile_first[ncols] = b->tiles;
        } catch (...) {
            free_batch(b);
            return LECO_FIX_ENOMEM;
        }
        b->arena_cap = arena + LECO_FIX_TAIL_PAD;
        uint32_t *d_ord = nullptr;
        uint64_t *d_ord_first = nullptr;
        bool ok = hipMalloc(&b->d_cols, ncols * sizeof(BatchColumn)) == hipSuccess &&
                  hipMalloc(&b->d_first, (ncols + 1) * sizeof(uint64_t)) == hipSuccess &&
                  hipMalloc(&b->d_fit_first, (ncols + 1) * sizeof(uint64_t)) == hipSuccess &&
                  hipMalloc(&b->d_tile_first, (ncols + 1) * sizeof(uint64_t)) == hipSuccess &&
                  hipMalloc(&b->d_tile_col, b->tiles * sizeof(uint32_t)) == hipSuccess &&
                  hipMalloc(&b->d_arena, b->arena_cap) == hipSuccess &&
                  hipMalloc(&b->d_offs, (b->total + 1) * sizeof(uint64_t)) == hipSuccess &&
                  hipMalloc(&b->d_prm, b->total * sizeof(leco_fix_block_param)) == hipSuccess &&
                  hipMalloc(&b->d_widths, b->total * sizeof(uint32_t)) == hipSuccess &&
                  hipMalloc(&b->d_unit_first, (ncols + 1) * sizeof(uint64_t)) == hipSuccess &&
                  hipMalloc(&b->d_unit_col, b->units * sizeof(uint32_t)) == hipSuccess &&
                  hipMalloc(&d_ord, ncols * sizeof(uint32_t)) == hipSuccess &&
                  hipMalloc(&d_ord_first, (ncols + 1) * sizeof(uint64_t)) == hipSuccess;
        auto free_ord = [&]() {
            (void)hipFree(d_ord);
            (void)hipFree(d_ord_first);
        };
        if (!ok) {
            free_ord();
            free_batch(b);
            return LECO_FIX_ENOMEM;
        }
        if (hipMemcpy(b->d_cols, b->h_cols.data(), ncols * sizeof(BatchColumn), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->d_first, b->h_first.data(), (ncols + 1) * sizeof(uint64_t), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(b->d_fit_first, b->h_fit_first.data(), (ncols + 1) * sizeof(uint64_t), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(b->d_tile_first, tile_first.data(), (ncols + 1) * sizeof(uint64_t), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(b->d_unit_first, unit_first.data(), (ncols + 1) * sizeof(uint64_t), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(d_ord, ord.data(), ncols * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_ord_first, ord_first.data(), (ncols + 1) * sizeof(uint64_t), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemset(b->d_arena + arena, 0, LECO_FIX_TAIL_PAD) != hipSuccess ||
            leco::launch_tile_columns(d_ord_first, d_ord, ncols, b->tiles, b->d_tile_col, nullptr) != hipSuccess ||
            leco::launch_tile_columns(b->d_unit_first, nullptr, ncols, b->units, b->d_unit_col, nullptr) !=
                hipSuccess ||
            leco::launch_batch_blocks(b->d_cols, b->d_first, ncols, b->total, b->d_widths, nullptr) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            free_ord();
            free_batch(b);
            return LECO_FIX_EHIP;
        }
        free_ord();
        *out = b;
        return LECO_FIX_OK;
    });
}

int leco_fix_batch_destroy(leco_fix_batch *b) {
    return leco::abi_guard([&]() -> int {
        if (!b) return LECO_FIX_EINVAL;
        free_batch(b);
        return LECO_FIX_OK;
    });
}

int leco_fix_batch_encode(leco_fix_batch *b, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!b) return LECO_FIX_EINVAL;
        for (const BatchColumn &c : b->h_cols)
            if (!c.in) return LECO_FIX_EINVAL;
        return enqueue(b, 0, (hipStream_t)stream);
    });
}

int leco_fix_batch_decode(leco_fix_batch *b, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!b) return LECO_FIX_EINVAL;
        for (const BatchColumn &c : b->h_cols)
            if (!c.out) return LECO_FIX_EINVAL;
        return enqueue(b, 1, (hipStream_t)stream);
    });
}

int leco_fix_batch_graph_launch(leco_fix_batch *b, int op, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!b || (op != 0 && op != 1)) return LECO_FIX_EINVAL;
        for (const BatchColumn &c : b->h_cols)
            if (op == 0 ? !c.in : !c.out) return LECO_FIX_EINVAL;
        if (!b->graph[op]) {
            if (!b->cap_stream && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess)
                return LECO_FIX_EHIP;
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeRelaxed) != hipSuccess) return LECO_FIX_EHIP;
            int rc = enqueue(b, op, b->cap_stream);
            hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
            if (rc != LECO_FIX_OK || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return LECO_FIX_EHIP;
            }
            e = hipGraphInstantiate(&b->graph[op], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->graph[op] = nullptr;
                return LECO_FIX_EHIP;
            }
        }
        return hip_status(hipGraphLaunch(b->graph[op], (hipStream_t)stream));
    });
}

int leco_fix_batch_decode_per_column(leco_fix_batch *b, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!b) return LECO_FIX_EINVAL;
        hipStream_t st = (hipStream_t)stream;
        for (uint32_t c = 0; c < b->ncols; ++c) {
            const BatchColumn &col = b->h_cols[c];
            if (!col.out) return LECO_FIX_EINVAL;
            const uint64_t *offs = b->d_offs + b->h_first[c];
            const leco_fix_block_param *prm = b->d_prm + b->h_first[c];
            hipError_t e;
            (void)prm;
            e = col.width == 64 ? leco::launch_decode<uint64_t>(b->d_arena, offs, col.n, col.bs, 0, col.n,
                                                                 (uint64_t *)col.out, st)
                                : leco::launch_decode<uint32_t>(b->d_arena, offs, col.n, col.bs, 0, col.n,
                                                                 (uint32_t *)col.out, st);
            if (e != hipSuccess) return LECO_FIX_EHIP;
        }
        return LECO_FIX_OK;
    });
}

int leco_fix_batch_encode_per_column(leco_fix_batch *b, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!b) return LECO_FIX_EINVAL;
        uint64_t need = 0, nbmax = 0;
        for (const BatchColumn &c : b->h_cols) {
            if (!c.in) return LECO_FIX_EINVAL;
            need = std::max(need, leco_fix_stream_bound(c.n, c.bs, c.width));
            nbmax = std::max(nbmax, (c.n + c.bs - 1) / c.bs);
        }
        const uint64_t offs_at = (need + 255) / 256 * 256, prm_at = offs_at + ((nbmax + 1) * 8 + 255) / 256 * 256;
        const uint64_t total = prm_at + nbmax * sizeof(leco_fix_block_param);
        if (total > b->scratch_cap) {
            (void)hipFree(b->d_scratch);
            b->d_scratch = nullptr;
            b->scratch_cap = 0;
            if (hipMalloc(&b->d_scratch, total) != hipSuccess) return LECO_FIX_ENOMEM;
            b->scratch_cap = total;
        }
        hipStream_t st = (hipStream_t)stream;
        uint64_t *offs = (uint64_t *)(b->d_scratch + offs_at);
        leco_fix_block_param *prm = (leco_fix_block_param *)(b->d_scratch + prm_at);
        for (const BatchColumn &c : b->h_cols) {
            hipError_t e = c.width == 64
                               ? leco::launch_encode<uint64_t>((const uint64_t *)c.in, c.n, c.bs, b->d_scratch, offs,
                                                               prm, st)
                               : leco::launch_encode<uint32_t>((const uint32_t *)c.in, c.n, c.bs, b->d_scratch, offs,
                                                               prm, st);
            if (e != hipSuccess) return LECO_FIX_EHIP;
        }
        return LECO_FIX_OK;
    });
}

int leco_fix_batch_set_output(leco_fix_batch *b, uint32_t c, void *d_out) {
    return leco::abi_guard([&]() -> int {
        if (!b || c >= b->ncols) return LECO_FIX_EINVAL;
        b->h_cols[c].out = d_out;
        return hip_status(hipMemcpy(&b->d_cols[c], &b->h_cols[c], sizeof(BatchColumn), hipMemcpyHostToDevice));
    });
}

int leco_fix_batch_set_decode_grid(leco_fix_batch *b, uint32_t wgs) {
    return leco::abi_guard([&]() -> int {
        if (!b) return LECO_FIX_EINVAL;
        b->decode_grid = wgs;  // 0: one workgroup per tile
        if (b->graph[1]) {  // the captured launch has the old grid baked in
            (void)hipGraphExecDestroy(b->graph[1]);
            b->graph[1] = nullptr;
        }
        return LECO_FIX_OK;
    });
}

int leco_fix_batch_info(const leco_fix_batch *b, uint64_t *total, uint64_t *cap, uint8_t **arena, uint64_t **offs,
                        leco_fix_block_param **prm) {
    return leco::abi_guard([&]() -> int {
        if (!b) return LECO_FIX_EINVAL;
        if (total) *total = b->total;
        if (cap) *cap = b->arena_cap;
        if (arena) *arena = b->d_arena;
        if (offs) *offs = b->d_offs;
        if (prm) *prm = b->d_prm;
        return LECO_FIX_OK;
    });
}

int leco_fix_batch_column(const leco_fix_batch *b, uint32_t c, uint64_t *first, uint64_t *nb) {
    return leco::abi_guard([&]() -> int {
        if (!b || c >= b->ncols) return LECO_FIX_EINVAL;
        if (first) *first = b->h_first[c];
        if (nb) *nb = b->h_first[c + 1] - b->h_first[c];
        return LECO_FIX_OK;
    });
}

int leco_fix_autotune_block_sizes(const leco_fix_column_desc *cols, uint32_t ncols, const uint32_t *cand,
                                  uint32_t ncand, uint32_t windows, uint32_t window, double tol, uint32_t *out_bs,
                                  uint64_t *out_bytes, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!cols || !cand || !out_bs || ncols == 0 || ncand == 0 || windows == 0 || window == 0)
            return LECO_FIX_EINVAL;
        for (uint32_t j = 0; j < ncand; ++j)
            if (cand[j] == 0 || cand[j] > leco::kMaxBlockSize || window % cand[j]) return LECO_FIX_EINVAL;
        // tuning batch: column c, window k, candidate j -> one entry
        std::vector<leco_fix_column_desc> tune;
        std::vector<uint32_t> owner;
        for (uint32_t c = 0; c < ncols; ++c) {
            const leco_fix_column_desc &d = cols[c];
            if (!d.d_in || !valid_desc(d)) return LECO_FIX_EINVAL;
            const uint64_t slots = d.n / window;  // whole windows, aligned to every candidate
            const uint32_t nw = slots == 0 ? 1 : (uint32_t)std::min<uint64_t>(windows, slots);
            for (uint32_t k = 0; k < nw; ++k) {
                const uint64_t start = slots == 0 ? 0 : (slots * k / nw) * window;
                const uint64_t len = slots == 0 ? d.n : window;
                for (uint32_t j = 0; j < ncand; ++j) {
                    tune.push_back({(const uint8_t *)d.d_in + start * (d.width_bits / 8), nullptr, len, cand[j],
                                    d.width_bits});
                    owner.push_back(c * ncand + j);
                }
            }
        }
        leco_fix_batch *b = nullptr;
        int rc = leco_fix_batch_create(tune.data(), (uint32_t)tune.size(), &b);
        if (rc) return rc;
        const std::unique_ptr<leco_fix_batch, void (*)(leco_fix_batch *)> hold(b, free_batch);  // freed on any exit
        hipStream_t st = (hipStream_t)stream;
        rc = leco_fix_batch_encode(b, st);
        std::vector<uint64_t> offs(b->total + 1);
        if (rc == LECO_FIX_OK &&
            (hipMemcpyAsync(offs.data(), b->d_offs, offs.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st) !=
                 hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess))
            rc = LECO_FIX_EHIP;
        if (rc == LECO_FIX_OK && offs[b->total] == LECO_FIX_OFFSETS_POISON) rc = LECO_FIX_EDEVICE;
        if (rc == LECO_FIX_OK) {
            std::vector<uint64_t> bytes((size_t)ncols * ncand, 0);
            for (size_t e = 0; e < tune.size(); ++e) bytes[owner[e]] += offs[b->h_first[e + 1]] - offs[b->h_first[e]];
            for (uint32_t c = 0; c < ncols; ++c) {
                uint64_t best = UINT64_MAX;
                for (uint32_t j = 0; j < ncand; ++j) best = std::min(best, bytes[(size_t)c * ncand + j]);
                uint32_t pick = 0;
                for (uint32_t j = 0; j < ncand; ++j)
                    if ((double)bytes[(size_t)c * ncand + j] <= (double)best * (1.0 + tol) && cand[j] > pick)
                        pick = cand[j];
                out_bs[c] = pick;
                if (out_bytes)
                    for (uint32_t j = 0; j < ncand; ++j)
                        out_bytes[(size_t)c * ncand + j] = bytes[(size_t)c * ncand + j];
            }
        }
        return rc;
    });
}

}  // extern "C"

// ---- autotuner with refinement and a measured decode-cost term -----------
namespace {

// Sample windows of column d for block size bs: up to `windows` windows of
// floor(window / bs) * bs values at evenly spaced multiples of `window` (one
// window of the whole column if it is shorter than `window`).
void sample_windows(const leco_fix_column_desc &d, uint32_t bs, uint32_t windows, uint32_t window,
                    std::vector<std::pair<uint64_t, uint64_t>> *out) {
    out->clear();
    const uint64_t slots = d.n / window;
    if (slots == 0) {
        out->push_back({0, d.n});
        return;
    }
    const uint64_t len = std::max<uint64_t>(bs, (uint64_t)(window / bs) * bs);
    const uint32_t nw = (uint32_t)std::min<uint64_t>(windows, slots);
    for (uint32_t k = 0; k < nw; ++k) {
        const uint64_t start = (slots * k / nw) * window;
        out->push_back({start, std::min<uint64_t>(len, d.n - start)});
    }
}

// Encoded bytes per value of column c at each of its candidates cand[c],
// from one batched encode of every column's sample windows.
int sample_bytes(const leco_fix_column_desc *cols, uint32_t ncols, const std::vector<std::vector<uint32_t>> &cand,
                 uint32_t windows, uint32_t window, hipStream_t st, std::vector<std::vector<double>> *bpv) {
    std::vector<leco_fix_column_desc> tune;
    std::vector<std::pair<uint32_t, uint32_t>> owner;
    std::vector<std::pair<uint64_t, uint64_t>> win;
    for (uint32_t c = 0; c < ncols; ++c)
        for (uint32_t j = 0; j < cand[c].size(); ++j) {
            sample_windows(cols[c], cand[c][j], windows, window, &win);
            for (const auto &w : win) {
                tune.push_back({(const uint8_t *)cols[c].d_in + w.first * (cols[c].width_bits / 8), nullptr, w.second,
                                cand[c][j], cols[c].width_bits, 0, 0});
                owner.push_back({c, j});
            }
        }
    leco_fix_batch *b = nullptr;
    int rc = leco_fix_batch_create(tune.data(), (uint32_t)tune.size(), &b);
    if (rc) return rc;
    const std::unique_ptr<leco_fix_batch, void (*)(leco_fix_batch *)> hold(b, free_batch);
    rc = leco_fix_batch_encode(b, st);
    std::vector<uint64_t> offs(b->total + 1);
    if (rc == LECO_FIX_OK &&
        (hipMemcpyAsync(offs.data(), b->d_offs, offs.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st) !=
             hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
        rc = LECO_FIX_EHIP;
    if (rc == LECO_FIX_OK && offs[b->total] == LECO_FIX_OFFSETS_POISON) rc = LECO_FIX_EDEVICE;
    if (rc) return rc;
    std::vector<std::vector<uint64_t>> bytes(ncols), values(ncols);
    for (uint32_t c = 0; c < ncols; ++c) {
        bytes[c].assign(cand[c].size(), 0);
        values[c].assign(cand[c].size(), 0);
    }
    for (size_t e = 0; e < tune.size(); ++e) {
        bytes[owner[e].first][owner[e].second] += offs[b->h_first[e + 1]] - offs[b->h_first[e]];
        values[owner[e].first][owner[e].second] += tune[e].n;
    }
    bpv->assign(ncols, {});
    for (uint32_t c = 0; c < ncols; ++c)
        for (uint32_t j = 0; j < cand[c].size(); ++j) (*bpv)[c].push_back((double)bytes[c][j] / values[c][j]);
    return LECO_FIX_OK;
}

// Decode nanoseconds per value of every column of width w (`members`) at
// block size bs and tile shape `tile` (0 = staged), timed over three batched
// decodes of their sample windows into `scratch`.
int sample_decode_ns(const leco_fix_column_desc *cols, const std::vector<uint32_t> &members, uint32_t bs, uint32_t tile,
                     uint32_t windows, uint32_t window, uint8_t *scratch, hipStream_t st, double *ns_per_value) {
    std::vector<leco_fix_column_desc> descs;
    std::vector<std::pair<uint64_t, uint64_t>> win;
    uint64_t at = 0, values = 0;
    for (uint32_t c : members) {
        sample_windows(cols[c], bs, windows, window, &win);
        const uint32_t vb = cols[c].width_bits / 8;
        for (const auto &w : win) {
            descs.push_back({(const uint8_t *)cols[c].d_in + w.first * vb, scratch + at, w.second, bs,
                             cols[c].width_bits, tile, 0});
            at += (w.second * vb + 255) / 256 * 256;
            values += w.second;
        }
    }
    leco_fix_batch *b = nullptr;
    int rc = leco_fix_batch_create(descs.data(), (uint32_t)descs.size(), &b);
    if (rc) return rc;
    const std::unique_ptr<leco_fix_batch, void (*)(leco_fix_batch *)> hold(b, free_batch);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return LECO_FIX_EHIP;
    }
    constexpr int kReps = 5;
    rc = leco_fix_batch_encode(b, st);
    if (rc == LECO_FIX_OK) rc = leco_fix_batch_decode(b, st);  // warm-up
    if (rc == LECO_FIX_OK && hipEventRecord(e0, st) != hipSuccess) rc = LECO_FIX_EHIP;
    for (int r = 0; r < kReps && rc == LECO_FIX_OK; ++r) rc = leco_fix_batch_decode(b, st);
    float ms = 0;
    if (rc == LECO_FIX_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = LECO_FIX_EHIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == LECO_FIX_OK) *ns_per_value = (double)ms * 1e6 / kReps / (double)values;
    return rc;
}

}  // namespace

extern "C" int leco_fix_autotune_columns(const leco_fix_column_desc *cols, uint32_t ncols, const uint32_t *cand,
                                         uint32_t ncand, uint32_t windows, uint32_t window, double tol,
                                         leco_fix_autotune_choice *out, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!cols || !cand || !out || ncols == 0 || ncand == 0 || windows == 0 || window == 0 || tol < 0)
            return LECO_FIX_EINVAL;
        for (uint32_t j = 0; j < ncand; ++j)
            if (cand[j] == 0 || cand[j] > leco::kMaxBlockSize || cand[j] > window) return LECO_FIX_EINVAL;
        for (uint32_t c = 0; c < ncols; ++c)
            if (!cols[c].d_in || !valid_desc(cols[c])) return LECO_FIX_EINVAL;
        hipStream_t st = (hipStream_t)stream;
        // 1. coarse: every candidate for every column
        std::vector<std::vector<uint32_t>> sizes(ncols, std::vector<uint32_t>(cand, cand + ncand));
        std::vector<std::vector<double>> bpv;
        int rc = sample_bytes(cols, ncols, sizes, windows, window, st, &bpv);
        if (rc) return rc;
        // 2. refinement around each column's best size (search_blocksize.h:68-110 refines by +-100 around
        //    its best of 200..10000; here 3/4, 5/4 and 3/2 of the best power of two)
        std::vector<std::vector<uint32_t>> refine(ncols);
        for (uint32_t c = 0; c < ncols; ++c) {
            const uint32_t jb = (uint32_t)(std::min_element(bpv[c].begin(), bpv[c].end()) - bpv[c].begin());
            const uint32_t best = sizes[c][jb];
            for (uint32_t r : {best - best / 4, best + best / 4, best + best / 2})
                if (r >= 1 && r <= leco::kMaxBlockSize && r <= window &&
                    std::find(sizes[c].begin(), sizes[c].end(), r) == sizes[c].end() &&
                    std::find(refine[c].begin(), refine[c].end(), r) == refine[c].end())
                    refine[c].push_back(r);
        }
        std::vector<std::vector<double>> rbpv;
        rc = sample_bytes(cols, ncols, refine, windows, window, st, &rbpv);
        if (rc) return rc;
        for (uint32_t c = 0; c < ncols; ++c) {
            sizes[c].insert(sizes[c].end(), refine[c].begin(), refine[c].end());
            bpv[c].insert(bpv[c].end(), rbpv[c].begin(), rbpv[c].end());
        }
        // 3. feasible sizes: within tol of the column's fewest bytes, each with its decode shapes
        static const uint32_t kTiles[] = {8192, 16384};
        struct Shape { uint32_t width, bs, tile; };
        auto key = [](const Shape &s) { return ((uint64_t)s.width << 56) | ((uint64_t)s.bs << 24) | s.tile; };
        std::map<uint64_t, double> cost;  // shape -> ns per value
        std::vector<std::vector<std::pair<uint32_t, Shape>>> feasible(ncols);
        for (uint32_t c = 0; c < ncols; ++c) {
            const double best = *std::min_element(bpv[c].begin(), bpv[c].end());
            for (uint32_t j = 0; j < sizes[c].size(); ++j) {
                if (bpv[c][j] > best * (1.0 + tol)) continue;
                const uint32_t bs = sizes[c][j], r = bs;
                if (leco::staged_decode(bs)) {
                    feasible[c].push_back({j, Shape{cols[c].width_bits, r, 0}});
                } else {
                    for (uint32_t t : kTiles) feasible[c].push_back({j, Shape{cols[c].width_bits, r, t}});
                }
                for (const auto &f : feasible[c]) cost.emplace(key(f.second), 0.0);
            }
        }
        // 4. one timed batched decode per shape, over the columns of its width
        uint64_t scratch_bytes = 0;
        std::vector<std::vector<uint32_t>> of_width(2);
        for (uint32_t c = 0; c < ncols; ++c) {
            of_width[cols[c].width_bits == 64].push_back(c);
            scratch_bytes += (uint64_t)windows * (window + 256) * (cols[c].width_bits / 8) + cols[c].n % 256 + 256;
        }
        uint8_t *scratch = nullptr;
        if (hipMalloc(&scratch, scratch_bytes) != hipSuccess) return LECO_FIX_ENOMEM;
        const std::unique_ptr<uint8_t, hipError_t (*)(void *)> hold(scratch, hipFree);
        for (auto &kv : cost) {
            const uint32_t width = (uint32_t)(kv.first >> 56), bs = (uint32_t)((kv.first >> 24) & 0xFFFFFFFF),
                           tile = (uint32_t)(kv.first & 0xFFFFFF);
            rc = sample_decode_ns(cols, of_width[width == 64], bs, tile, windows, window, scratch, st, &kv.second);
            if (rc) return rc;
        }
        // 5. per column: start from the fewest bytes; another feasible size or shape replaces it only
        //    when its measured decode time per value is lower by more than kGain (timing noise on the
        //    sample must not trade bytes for nothing)
        constexpr double kGain = 0.03;
        for (uint32_t c = 0; c < ncols; ++c) {
            size_t bi = 0;
            for (size_t i = 1; i < feasible[c].size(); ++i) {
                const auto &f = feasible[c][i], &b = feasible[c][bi];
                if (bpv[c][f.first] < bpv[c][b.first] ||
                    (bpv[c][f.first] == bpv[c][b.first] && cost[key(f.second)] < cost[key(b.second)]))
                    bi = i;
            }
            double bc = cost[key(feasible[c][bi].second)];
            size_t pick = bi;
            for (size_t i = 0; i < feasible[c].size(); ++i) {
                const double k = cost[key(feasible[c][i].second)];
                if (k < bc * (1 - kGain) && k < cost[key(feasible[c][pick].second)]) pick = i;
            }
            const auto &f = feasible[c][pick];
            out[c] = leco_fix_autotune_choice{sizes[c][f.first], f.second.tile, bpv[c][f.first],
                                              cost[key(f.second)]};
        }
        return LECO_FIX_OK;
    });
}
