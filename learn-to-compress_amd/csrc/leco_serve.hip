// Request server over resident compressed columns (C ABI in include/leco_fix.h).
//
// Continuous batching: a host loop admits arriving point-lookup and
// range-decode requests into shared launches of k_serve.  Up to `slots`
// batches are in flight at once, one stream each; a request that arrives
// while others are on the GPU joins the next batch as soon as a slot and
// result memory are free, so batches overlap on the device.  Results live in
// a ring-allocated pool; columns inserted while traffic runs are encoded on a
// separate stream into a first-fit column pool and evicted once no admitted
// request uses them.  The pools are carved once from free device memory minus
// a reserve, which is the memory admission.  A batch's work items and lookup
// indices are written into its slot's pinned, coherent host staging, which
// the kernel reads over the host link: the loop enqueues no host-to-device
// copy per batch (such copies occasionally blocked the host call for ~10 ms).
//
// The plain mode serves the same trace with one launch and one synchronize
// per request (the per-request launch path), for comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "leco_internal.h"

namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kLookupChunk = 1024;   // indices per served work item
constexpr uint64_t kAlign = 256;
constexpr uint32_t kMaxColumns = 1u << 16;
constexpr uint64_t kIdxCap = 1ull << 22;  // staged lookup indices per batch
constexpr uint64_t kItemCap = 1ull << 18; // work items per batch
constexpr uint32_t kServeGrid = 4096;     // workgroups per served launch at most
constexpr unsigned kCopyHelpers = 7;      // threads beside the server loop for index copies (at most)

uint64_t up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Host copies of a batch's lookup indices into its pinned staging buffer.  A
// batch of 4M indices is 32 MB; one thread copies ~10 GB/s, which caps
// lookup-heavy serving well below what the GPU sustains, so a batch's copies
// above kSplitBytes are cut into parts of at least kPartBytes that the caller
// and up to `helpers` threads claim one at a time.  The caller copies every
// part nobody has claimed and then waits only for parts already being copied,
// so a helper that is slow to wake delays nothing.  A claim ticket carries the
// generation (one per run()), so a helper that wakes after its generation is
// over claims nothing.  The helpers sleep on a condition variable between
// batches.
class CopyPool {
public:
    static constexpr size_t kSplitBytes = 256 << 10;
    static constexpr size_t kPartBytes = 64 << 10;
    static constexpr uint32_t kMaxParts = 32;

    bool start(unsigned helpers) {
        try {
            for (unsigned h = 0; h < helpers; ++h) threads_.emplace_back([this] { work(); });
        } catch (...) {
            stop();
            return false;
        }
        return true;
    }
    void stop() {
        {
            std::lock_guard<std::mutex> g(m_);
            quit_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : threads_) t.join();
        threads_.clear();
    }
    void add(void *dst, const void *src, size_t bytes) {
        jobs_.push_back(Job{(uint8_t *)dst, (const uint8_t *)src, total_});
        total_ += bytes;
    }
    // Performs every recorded copy and returns when all are done.
    void run() {
        if (jobs_.empty()) return;
        if (threads_.empty() || total_ < kSplitBytes) {
            copy_range(0, total_);
        } else {
            const uint32_t parts = (uint32_t)std::min<size_t>(kMaxParts, total_ / kPartBytes);
            uint64_t gen;
            {
                std::lock_guard<std::mutex> g(m_);
                gen = ++gen_;
                parts_ = parts;
                done_.store(0, std::memory_order_relaxed);
                ticket_.store(gen << 32, std::memory_order_release);
            }
            cv_.notify_all();
            copy_parts(gen, parts);
            while (done_.load(std::memory_order_acquire) != parts) std::this_thread::yield();
        }
        jobs_.clear();
        total_ = 0;
    }

private:
    struct Job {
        uint8_t *dst;
        const uint8_t *src;
        size_t at;  // offset of this copy in the batch's concatenated bytes
    };
    // Bytes [lo, hi) of the concatenated copies.
    void copy_range(size_t lo, size_t hi) {
        size_t j = std::upper_bound(jobs_.begin(), jobs_.end(), lo, [](size_t v, const Job &x) { return v < x.at; }) -
                   jobs_.begin() - 1;
        for (; j < jobs_.size() && jobs_[j].at < hi; ++j) {
            const size_t end = j + 1 < jobs_.size() ? jobs_[j + 1].at : total_;
            const size_t a = std::max(lo, jobs_[j].at), b = std::min(hi, end);
            if (b > a) std::memcpy(jobs_[j].dst + (a - jobs_[j].at), jobs_[j].src + (a - jobs_[j].at), b - a);
        }
    }
    // Claims and copies parts of generation gen until none is left.
    void copy_parts(uint64_t gen, uint32_t parts) {
        uint64_t v = ticket_.load(std::memory_order_acquire);
        for (;;) {
            if ((v >> 32) != gen || (uint32_t)v >= parts) return;
            if (!ticket_.compare_exchange_weak(v, v + 1, std::memory_order_acq_rel, std::memory_order_acquire))
                continue;
            const uint32_t p = (uint32_t)v;
            copy_range(total_ * p / parts, total_ * (p + 1) / parts);
            done_.fetch_add(1, std::memory_order_release);
            v = ticket_.load(std::memory_order_acquire);
        }
    }
    void work() {
        uint64_t seen = 0;
        for (;;) {
            uint32_t parts;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
                parts = parts_;
            }
            copy_parts(seen, parts);
        }
    }

    std::vector<std::thread> threads_;
    std::vector<Job> jobs_;
    size_t total_ = 0;
    std::mutex m_;
    std::condition_variable cv_;
    uint64_t gen_ = 0;
    uint32_t parts_ = 1;
    bool quit_ = false;
    std::atomic<uint64_t> ticket_{0};  // generation << 32 | next part
    std::atomic<uint32_t> done_{0};    // parts of the current generation copied
};

// Ring over [0, cap): allocations are made in order and reclaimed once every
// older allocation has been released too.
class Ring {
public:
    void init(uint64_t cap) { cap_ = cap; }
    bool alloc(uint64_t n, uint64_t *off, uint64_t *id) {
        n = up(n ? n : 1, kAlign);
        if (n > cap_) return false;
        uint64_t b;
        if (q_.empty()) {
            b = 0;
        } else {
            const uint64_t tail = q_.front().b;
            if (head_ > tail) {  // used [tail, head): free [head, cap) and [0, tail)
                if (head_ + n <= cap_) b = head_;
                else if (n <= tail) b = 0;
                else return false;
            } else {             // wrapped, used [tail, cap) + [0, head): free [head, tail)
                if (head_ + n <= tail) b = head_;
                else return false;
            }
        }
        head_ = b + n;
        used_ += n;
        q_.push_back({b, n, false, next_});
        *off = b;
        *id = next_++;
        return true;
    }
    void release(uint64_t id) {
        for (auto &a : q_)
            if (a.id == id) { a.freed = true; break; }
        while (!q_.empty() && q_.front().freed) {
            used_ -= q_.front().n;
            q_.pop_front();
        }
        if (q_.empty()) head_ = 0;
    }
    uint64_t used() const { return used_; }

private:
    struct A { uint64_t b, n; bool freed; uint64_t id; };
    std::deque<A> q_;
    uint64_t cap_ = 0, head_ = 0, next_ = 0, used_ = 0;
};

// First-fit allocator with coalescing, for inserted columns.
class FreeList {
public:
    void init(uint64_t cap) { free_.clear(); if (cap) free_[0] = cap; used_ = 0; }
    bool alloc(uint64_t n, uint64_t *off) {
        n = up(n, kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += n;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = up(n, kAlign);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
    uint64_t used() const { return used_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t used_ = 0;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    uint64_t batch = 0;
    std::vector<uint64_t> reqs;
    std::vector<uint64_t> ring_ids;
    uint64_t *h_idx = nullptr, *d_idx = nullptr;
    leco::ServeItem *h_items = nullptr;
};

struct ColumnSlot {
    leco::ServeColumn col{};
    bool present = false, ready = false, evicting = false, owned = false;
    uint64_t pool_off = 0, pool_len = 0, stream_bytes = 0, prefix = 0;
    uint32_t refs = 0;
    uint64_t insert_req = UINT64_MAX, evict_req = UINT64_MAX;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint64_t *h_total = nullptr;  // pinned: stream bytes, copied after encode
};

double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

// The LECO_FIX_SERVER_TRACE slowest iterations of a run's admission loop,
// kept in a fixed array (no allocation inside the loop).
struct IterationTrace {
    leco_fix_server_iteration slow[LECO_FIX_SERVER_TRACE];
    uint32_t count = 0;
    void reset() { count = 0; }
    void add(const leco_fix_server_iteration &it) {
        auto total = [](const leco_fix_server_iteration &x) {
            return x.arrivals_us + x.completions_us + x.admission_us + x.copies_us + x.launch_us + x.rest_us;
        };
        if (count < LECO_FIX_SERVER_TRACE) {
            slow[count++] = it;
            return;
        }
        uint32_t lo = 0;
        for (uint32_t k = 1; k < count; ++k)
            if (total(slow[k]) < total(slow[lo])) lo = k;
        if (total(it) > total(slow[lo])) slow[lo] = it;
    }
};

struct leco_fix_server {
    leco_fix_server_config cfg{};
    IterationTrace trace;
    std::vector<ColumnSlot> cols;
    leco::ServeColumn *d_cols = nullptr;
    uint8_t *result_pool = nullptr;
    uint64_t result_cap = 0;
    Ring ring;
    uint8_t *column_pool = nullptr;
    uint64_t column_cap = 0;
    FreeList colfree;
    std::vector<Slot> slots;
    hipStream_t enc = nullptr;
    CopyPool copies;
};

namespace {

int upload_column(leco_fix_server *s, uint32_t c, hipStream_t st) {
    return hipMemcpyAsync(&s->d_cols[c], &s->cols[c].col, sizeof(leco::ServeColumn), hipMemcpyHostToDevice, st) ==
                   hipSuccess
               ? LECO_FIX_OK
               : LECO_FIX_EHIP;
}

void destroy_server(leco_fix_server *s) {
    s->copies.stop();
    (void)hipDeviceSynchronize();
    for (Slot &sl : s->slots) {
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        if (sl.h_idx) (void)hipHostFree(sl.h_idx);
        if (sl.h_items) (void)hipHostFree(sl.h_items);
        (void)hipFree(sl.d_idx);
    }
    for (ColumnSlot &c : s->cols) {
        if (c.e0) (void)hipEventDestroy(c.e0);
        if (c.e1) (void)hipEventDestroy(c.e1);
        if (c.h_total) (void)hipHostFree(c.h_total);
    }
    if (s->enc) (void)hipStreamDestroy(s->enc);
    (void)hipFree(s->d_cols);
    (void)hipFree(s->result_pool);
    (void)hipFree(s->column_pool);
    delete s;
}

uint64_t result_bytes(const leco_fix_server *s, const leco_fix_request &r) {
    return r.count * (s->cols[r.column].col.width / 8);
}

// Storage of an inserted column: [params][offsets][stream + tail pad].
struct InsertLayout {
    uint64_t nb, prm, offs, stream, total;
};
InsertLayout insert_layout(const leco_fix_request &r) {
    InsertLayout L;
    L.nb = (r.count + r.block_size - 1) / r.block_size;
    L.prm = 0;
    L.offs = up(L.nb * sizeof(leco_fix_block_param), kAlign);
    L.stream = L.offs + up((L.nb + 1) * sizeof(uint64_t), kAlign);
    L.total = L.stream + leco_fix_stream_bound(r.count, r.block_size, r.width_bits);
    return L;
}

bool valid_insert(const leco_fix_request &r) {
    return r.d_values && r.count > 0 && r.block_size >= 1 && r.block_size <= leco::kMaxBlockSize &&
           (r.width_bits == 32 || r.width_bits == 64);
}

// Starts encoding an inserted column; false if the column pool is full.
bool start_insert(leco_fix_server *s, const leco_fix_request &r, uint64_t ri) {
    ColumnSlot &c = s->cols[r.column];
    const InsertLayout L = insert_layout(r);
    uint64_t off;
    if (!s->colfree.alloc(L.total, &off)) return false;
    uint8_t *base = s->column_pool + off;
    c.col = leco::ServeColumn{base + L.stream, (const uint64_t *)(base + L.offs),
                              (const leco_fix_block_param *)(base + L.prm), r.count, r.block_size, r.width_bits};
    c.present = true;
    c.ready = false;
    c.evicting = false;
    c.owned = true;
    c.pool_off = off;
    c.pool_len = L.total;
    c.prefix = L.stream;
    c.insert_req = ri;
    upload_column(s, r.column, s->enc);
    (void)hipEventRecord(c.e0, s->enc);
    uint8_t *st = base + L.stream;
    uint64_t *offs = (uint64_t *)(base + L.offs);
    leco_fix_block_param *prm = (leco_fix_block_param *)(base + L.prm);
    const uint64_t cap = L.total - L.stream;
    if (r.width_bits == 64)
        (void)leco::launch_encode<uint64_t>((const uint64_t *)r.d_values, r.count, r.block_size, st, offs, prm, s->enc);
    else
        (void)leco::launch_encode<uint32_t>((const uint32_t *)r.d_values, r.count, r.block_size, st, offs, prm, s->enc);
    (void)cap;
    (void)hipMemcpyAsync(c.h_total, offs + L.nb, sizeof(uint64_t), hipMemcpyDeviceToHost, s->enc);
    (void)hipEventRecord(c.e1, s->enc);
    return true;
}

void free_column(leco_fix_server *s, uint32_t ci) {
    ColumnSlot &c = s->cols[ci];
    if (c.owned) s->colfree.release(c.pool_off, c.pool_len);
    c.present = c.ready = c.evicting = c.owned = false;
    c.col = leco::ServeColumn{};
    c.pool_len = 0;
}

}  // namespace

extern "C" {

int leco_fix_server_create(const leco_fix_server_config *cfg_in, leco_fix_server **out) {
    return leco::abi_guard([&]() -> int {
        if (!out) return LECO_FIX_EINVAL;
        leco_fix_server *s = new (std::nothrow) leco_fix_server;
        if (!s) return LECO_FIX_ENOMEM;
        if (cfg_in) s->cfg = *cfg_in;
        leco_fix_server_config &cfg = s->cfg;
        if (!cfg.slots) cfg.slots = 4;
        if (!cfg.max_batch_requests) cfg.max_batch_requests = 8192;
        if (!cfg.max_batch_values) cfg.max_batch_values = 1ull << 26;
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) { destroy_server(s); return LECO_FIX_EHIP; }
        const uint64_t staging =
            (uint64_t)cfg.slots * (kIdxCap * 8 + kItemCap * sizeof(leco::ServeItem)) + (64ull << 20);
        if (freeb < cfg.reserve_bytes + staging) { destroy_server(s); return LECO_FIX_ENOMEM; }
        const uint64_t avail = freeb - cfg.reserve_bytes - staging;
        if (!cfg.result_pool_bytes)
            cfg.result_pool_bytes = cfg.column_pool_bytes ? avail - std::min(avail, cfg.column_pool_bytes) : avail / 2;
        if (!cfg.column_pool_bytes) cfg.column_pool_bytes = avail - std::min(avail, cfg.result_pool_bytes);
        if (cfg.result_pool_bytes + cfg.column_pool_bytes > avail) { destroy_server(s); return LECO_FIX_ENOMEM; }
        s->result_cap = cfg.result_pool_bytes;
        s->column_cap = cfg.column_pool_bytes;
        try {
            s->cols.resize(kMaxColumns);
            s->slots.resize(cfg.slots);
        } catch (...) {
            destroy_server(s);
            return LECO_FIX_ENOMEM;
        }
        bool ok = hipMalloc(&s->d_cols, kMaxColumns * sizeof(leco::ServeColumn)) == hipSuccess &&
                  (s->result_cap == 0 || hipMalloc(&s->result_pool, s->result_cap) == hipSuccess) &&
                  (s->column_cap == 0 || hipMalloc(&s->column_pool, s->column_cap) == hipSuccess) &&
                  hipStreamCreateWithFlags(&s->enc, hipStreamNonBlocking) == hipSuccess;
        for (Slot &sl : s->slots) {
            ok = ok && hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
                 hipHostMalloc((void **)&sl.h_idx, kIdxCap * sizeof(uint64_t), hipHostMallocCoherent) == hipSuccess &&
                 hipHostMalloc((void **)&sl.h_items, kItemCap * sizeof(leco::ServeItem), hipHostMallocCoherent) ==
                     hipSuccess &&
                 hipMalloc(&sl.d_idx, kIdxCap * sizeof(uint64_t)) == hipSuccess;
        }
        if (!ok) { destroy_server(s); return LECO_FIX_ENOMEM; }
        const unsigned hw = std::thread::hardware_concurrency();
        if (!s->copies.start(hw > 1 ? std::min(kCopyHelpers, hw - 1) : 0)) {
            destroy_server(s);
            return LECO_FIX_ENOMEM;
        }
        s->ring.init(s->result_cap);
        s->colfree.init(s->column_cap);
        *out = s;
        return LECO_FIX_OK;
    });
}

int leco_fix_server_destroy(leco_fix_server *s) {
    return leco::abi_guard([&]() -> int {
        if (!s) return LECO_FIX_EINVAL;
        destroy_server(s);
        return LECO_FIX_OK;
    });
}

int leco_fix_server_attach_column(leco_fix_server *s, uint32_t c, const uint8_t *st, const uint64_t *offs,
                                  const leco_fix_block_param *prm, uint64_t n, uint32_t bs, uint32_t w) {
    return leco::abi_guard([&]() -> int {
        if (!s || c >= kMaxColumns || !st || !offs || !prm || n == 0 || bs == 0 || bs > leco::kMaxBlockSize ||
            (w != 32 && w != 64) || s->cols[c].present)
            return LECO_FIX_EINVAL;
        ColumnSlot &cs = s->cols[c];
        cs.col = leco::ServeColumn{st, offs, prm, n, bs, w};
        cs.present = cs.ready = true;
        cs.owned = false;
        int rc = upload_column(s, c, s->enc);
        if (rc) return rc;
        return hipStreamSynchronize(s->enc) == hipSuccess ? LECO_FIX_OK : LECO_FIX_EHIP;
    });
}

int leco_fix_server_slowest_iterations(const leco_fix_server *s, leco_fix_server_iteration *out, uint32_t cap,
                                       uint32_t *count) {
    if (!s || !count || (cap && !out)) return LECO_FIX_EINVAL;
    leco_fix_server_iteration sorted[LECO_FIX_SERVER_TRACE];
    const uint32_t n = s->trace.count;
    std::copy(s->trace.slow, s->trace.slow + n, sorted);
    auto total = [](const leco_fix_server_iteration &x) {
        return x.arrivals_us + x.completions_us + x.admission_us + x.copies_us + x.launch_us + x.rest_us;
    };
    std::sort(sorted, sorted + n, [&](const auto &a, const auto &b) { return total(a) > total(b); });
    const uint32_t k = std::min(n, cap);
    std::copy(sorted, sorted + k, out);
    *count = k;
    return LECO_FIX_OK;
}

int leco_fix_server_column(const leco_fix_server *s, uint32_t c, const uint8_t **st, const uint64_t **offs,
                           const leco_fix_block_param **prm, uint64_t *n, uint32_t *bs, uint32_t *w,
                           uint64_t *bytes) {
    return leco::abi_guard([&]() -> int {
        if (!s || c >= kMaxColumns || !s->cols[c].present || !s->cols[c].ready) return LECO_FIX_EINVAL;
        const ColumnSlot &cs = s->cols[c];
        if (st) *st = cs.col.stream;
        if (offs) *offs = cs.col.offsets;
        if (prm) *prm = cs.col.params;
        if (n) *n = cs.col.n;
        if (bs) *bs = cs.col.bs;
        if (w) *w = cs.col.width;
        if (bytes) *bytes = cs.stream_bytes;
        return LECO_FIX_OK;
    });
}

int leco_fix_server_run(leco_fix_server *s, const leco_fix_request *reqs, uint64_t nreq, int mode,
                        leco_fix_request_result *res, leco_fix_server_stats *stats) {
    return leco::abi_guard([&]() -> int {
        if (!s || (!reqs && nreq) || !res || (mode != LECO_FIX_SERVE_BATCHED && mode != LECO_FIX_SERVE_PLAIN))
            return LECO_FIX_EINVAL;
        for (uint64_t i = 0; i < nreq; ++i) {
            const leco_fix_request &r = reqs[i];
            if (r.column >= kMaxColumns || r.kind > LECO_FIX_REQ_EVICT || (i && r.arrival_s < reqs[i - 1].arrival_s))
                return LECO_FIX_EINVAL;
            if (r.kind == LECO_FIX_REQ_LOOKUP && (!r.h_idx || r.count == 0 || r.count > kIdxCap))
                return LECO_FIX_EINVAL;
            if (r.kind == LECO_FIX_REQ_RANGE && (r.count == 0 || r.count > s->cfg.max_batch_values))
                return LECO_FIX_EINVAL;
            if (r.kind == LECO_FIX_REQ_INSERT && !valid_insert(r)) return LECO_FIX_EINVAL;
            res[i] = leco_fix_request_result{0, 0, 0, -1, 0};
        }
        leco_fix_server_stats st{};
        const Clock::time_point t0 = Clock::now();
        auto now_s = [&]() { return secs(t0, Clock::now()); };
        uint64_t done = 0, next = 0, batch_id = 0;
        double last_done = 0;

        auto complete = [&](uint64_t ri, double t, int status) {
            res[ri].latency_s = t - reqs[ri].arrival_s;
            res[ri].status = status;
            if (status == LECO_FIX_OK) st.completed++;
            else st.failed++;
            last_done = std::max(last_done, t);
            ++done;
        };
        auto ensure_events = [&](ColumnSlot &c) -> bool {
            if (!c.e0 && hipEventCreate(&c.e0) != hipSuccess) return false;
            if (!c.e1 && hipEventCreate(&c.e1) != hipSuccess) return false;
            if (!c.h_total && hipHostMalloc((void **)&c.h_total, sizeof(uint64_t), hipHostMallocDefault) != hipSuccess)
                return false;
            return true;
        };
        auto finish_insert = [&](uint32_t ci, double t) {
            ColumnSlot &c = s->cols[ci];
            float ms = 0;
            (void)hipEventElapsedTime(&ms, c.e0, c.e1);
            st.encode_gpu_s += ms * 1e-3;
            if (*c.h_total == LECO_FIX_OFFSETS_POISON) {  // the encoder found a broken invariant
                const uint64_t ri = c.insert_req, ev = c.evict_req;
                free_column(s, ci);
                c.insert_req = c.evict_req = UINT64_MAX;
                complete(ri, t, LECO_FIX_EDEVICE);
                if (ev != UINT64_MAX) complete(ev, t, LECO_FIX_OK);  // an evict waiting for the insert: nothing left
                return;
            }
            st.encoded_values += c.col.n;
            st.encoded_bytes += c.col.n * (c.col.width / 8);
            c.stream_bytes = *c.h_total;
            // shrink the column's storage to what the stream needs
            const uint64_t keep = up(c.prefix + c.stream_bytes + LECO_FIX_TAIL_PAD, kAlign);
            if (keep < c.pool_len) {
                s->colfree.release(c.pool_off + keep, c.pool_len - keep);
                c.pool_len = keep;
            }
            c.ready = true;
            complete(c.insert_req, t, LECO_FIX_OK);
            c.insert_req = UINT64_MAX;
        };
        auto try_evict = [&](uint32_t ci, double t) {
            ColumnSlot &c = s->cols[ci];
            if (c.evicting && c.refs == 0 && c.ready) {
                const uint64_t ri = c.evict_req;
                free_column(s, ci);
                c.evict_req = UINT64_MAX;
                complete(ri, t, LECO_FIX_OK);
            }
        };

        if (mode == LECO_FIX_SERVE_PLAIN) {
            Slot &sl = s->slots[0];
            for (uint64_t i = 0; i < nreq; ++i) {
                const leco_fix_request &r = reqs[i];
                while (now_s() < r.arrival_s) {
                }
                const double ts = now_s();
                res[i].queued_s = ts - r.arrival_s;
                res[i].batch = batch_id++;
                ColumnSlot &c = s->cols[r.column];
                int status = LECO_FIX_OK;
                if (r.kind == LECO_FIX_REQ_INSERT) {
                    if (c.present || !ensure_events(c) || !start_insert(s, r, i)) {
                        complete(i, now_s(), LECO_FIX_ENOMEM);
                        continue;
                    }
                    (void)hipStreamSynchronize(s->enc);
                    finish_insert(r.column, now_s());
                    continue;
                }
                if (r.kind == LECO_FIX_REQ_EVICT) {
                    if (!c.present) { complete(i, now_s(), LECO_FIX_EINVAL); continue; }
                    c.evicting = true;
                    c.evict_req = i;
                    try_evict(r.column, now_s());
                    continue;
                }
                if (!c.present || !c.ready || c.evicting) { complete(i, now_s(), LECO_FIX_EINVAL); continue; }
                const uint64_t bytes = result_bytes(s, r);
                if (!r.d_out && bytes > s->result_cap) { complete(i, now_s(), LECO_FIX_ENOMEM); continue; }
                void *out = r.d_out ? r.d_out : s->result_pool;
                hipError_t e;
                if (r.kind == LECO_FIX_REQ_LOOKUP) {
                    std::memcpy(sl.h_idx, r.h_idx, r.count * sizeof(uint64_t));
                    e = hipMemcpyAsync(sl.d_idx, sl.h_idx, r.count * sizeof(uint64_t), hipMemcpyHostToDevice, sl.st);
                    if (e == hipSuccess)
                        e = c.col.width == 64
                                ? leco::launch_lookup<uint64_t>(c.col.stream, c.col.params, c.col.n, c.col.bs, sl.d_idx,
                                                                r.count, (uint64_t *)out, sl.st)
                                : leco::launch_lookup<uint32_t>(c.col.stream, c.col.params, c.col.n, c.col.bs, sl.d_idx,
                                                                r.count, (uint32_t *)out, sl.st);
                } else {
                    if (r.begin + r.count > c.col.n) { complete(i, now_s(), LECO_FIX_EINVAL); continue; }
                    if (leco::staged_decode(c.col.bs))
                        e = c.col.width == 64
                                ? leco::launch_decode<uint64_t>(c.col.stream, c.col.offsets, c.col.n, c.col.bs, r.begin,
                                                                r.begin + r.count, (uint64_t *)out, sl.st)
                                : leco::launch_decode<uint32_t>(c.col.stream, c.col.offsets, c.col.n, c.col.bs, r.begin,
                                                                r.begin + r.count, (uint32_t *)out, sl.st);
                    else
                        e = c.col.width == 64
                                ? leco::launch_decode_values<uint64_t>(c.col.stream, c.col.params, c.col.bs, r.begin,
                                                                       r.begin + r.count, (uint64_t *)out, sl.st)
                                : leco::launch_decode_values<uint32_t>(c.col.stream, c.col.params, c.col.bs, r.begin,
                                                                       r.begin + r.count, (uint32_t *)out, sl.st);
                }
                if (e == hipSuccess && r.h_out) e = hipMemcpyAsync(r.h_out, out, bytes, hipMemcpyDeviceToHost, sl.st);
                if (e == hipSuccess) e = hipStreamSynchronize(sl.st);
                if (e != hipSuccess) status = LECO_FIX_EHIP;
                st.values_out += r.count;
                st.bytes_out += bytes;
                st.batches++;
                st.max_batch_requests = 1;
                complete(i, now_s(), status);
            }
        } else {
            std::deque<uint64_t> queue;  // arrived, not yet admitted, in arrival order
            const size_t window = 4 * (size_t)s->cfg.max_batch_requests;  // queue entries examined per pass
            size_t controls = 0;  // insert / evict requests in the queue
            std::vector<uint32_t> pending_inserts;
            // per-iteration lists, kept across iterations so the loop does not allocate once they have grown
            std::vector<uint64_t> keep, admitted, ids;
            keep.reserve(window);
            admitted.reserve(s->cfg.max_batch_requests);
            ids.reserve(2 * (size_t)s->cfg.max_batch_requests);
            s->trace.reset();
            Clock::time_point prev_end = Clock::now();
            while (done < nreq) {
                bool progressed = false;
                // phase marks of this iteration (the trace keeps the slowest ones)
                const Clock::time_point p0 = Clock::now();
                Clock::time_point p1, p2, p3 = p0, p4 = p0, p5 = p0;
                leco_fix_server_iteration it{};
                it.start_s = secs(t0, p0);
                it.between_us = (float)(secs(prev_end, p0) * 1e6);
                it.queued = (uint32_t)queue.size();
                double t = it.start_s;
                for (; next < nreq && reqs[next].arrival_s <= t; ++next) {
                    queue.push_back(next);
                    controls += reqs[next].kind > LECO_FIX_REQ_RANGE;
                }
                p1 = Clock::now();

                // completions (a query that fails for another reason than "not
                // ready" fails the slot's requests, so the loop still ends)
                for (Slot &sl : s->slots) {
                    if (!sl.busy) continue;
                    const hipError_t q = hipEventQuery(sl.done);
                    if (q == hipErrorNotReady) continue;
                    t = now_s();
                    for (uint64_t ri : sl.reqs) {
                        s->cols[reqs[ri].column].refs--;
                        complete(ri, t, q == hipSuccess ? LECO_FIX_OK : LECO_FIX_EHIP);
                    }
                    for (uint64_t id : sl.ring_ids) s->ring.release(id);
                    for (uint64_t ri : sl.reqs) try_evict(reqs[ri].column, t);
                    sl.reqs.clear();
                    sl.ring_ids.clear();
                    sl.busy = false;
                    progressed = true;
                }
                for (size_t k = 0; k < pending_inserts.size();) {
                    ColumnSlot &c = s->cols[pending_inserts[k]];
                    const hipError_t q = hipEventQuery(c.e1);
                    if (q != hipErrorNotReady) {
                        t = now_s();
                        if (q == hipSuccess) {
                            finish_insert(pending_inserts[k], t);
                        } else {  // the encode's stream failed: the insert fails, its column is dropped
                            const uint64_t ri = c.insert_req, ev = c.evict_req;
                            free_column(s, pending_inserts[k]);
                            c.insert_req = c.evict_req = UINT64_MAX;
                            complete(ri, t, LECO_FIX_EHIP);
                            if (ev != UINT64_MAX) complete(ev, t, LECO_FIX_OK);
                        }
                        try_evict(pending_inserts[k], t);
                        pending_inserts[k] = pending_inserts.back();
                        pending_inserts.pop_back();
                        progressed = true;
                    } else {
                        ++k;
                    }
                }

                p2 = Clock::now();
                p3 = p4 = p5 = p2;
                // admission: control requests, then one batch into a free slot
                Slot *free_slot = nullptr;
                for (Slot &sl : s->slots)
                    if (!sl.busy) { free_slot = &sl; break; }
                if (!queue.empty()) {
                    keep.clear();
                    const size_t scan = std::min(queue.size(), window);
                    uint64_t nitems = 0, nidx = 0, nvals = 0;
                    bool full = free_slot == nullptr;
                    admitted.clear();
                    // Once the batch is full, every data request is held; the scan
                    // goes on only for insert / evict requests further back.
                    size_t ctl_ahead = controls, ctl_kept = 0, qi = 0;
                    for (; qi < scan && !(full && ctl_ahead == 0); ++qi) {
                        const uint64_t ri = queue[qi];
                        const leco_fix_request &r = reqs[ri];
                        ColumnSlot &c = s->cols[r.column];
                        t = now_s();
                        if (r.kind > LECO_FIX_REQ_RANGE) ctl_ahead--;
                        if (r.kind == LECO_FIX_REQ_INSERT) {
                            if (c.present || !ensure_events(c)) {
                                complete(ri, t, LECO_FIX_EINVAL);
                                progressed = true;
                                continue;
                            }
                            if (insert_layout(r).total > s->column_cap) {
                                complete(ri, t, LECO_FIX_ENOMEM);
                                progressed = true;
                                continue;
                            }
                            if (start_insert(s, r, ri)) {
                                res[ri].queued_s = t - r.arrival_s;
                                pending_inserts.push_back(r.column);
                                progressed = true;
                            } else {
                                keep.push_back(ri);  // waits for column memory
                                ctl_kept++;
                            }
                            continue;
                        }
                        if (r.kind == LECO_FIX_REQ_EVICT) {
                            bool earlier = false;
                            for (uint64_t k : keep)
                                if (reqs[k].column == r.column && reqs[k].kind <= LECO_FIX_REQ_RANGE) earlier = true;
                            if (!c.present) { complete(ri, t, LECO_FIX_EINVAL); progressed = true; continue; }
                            if (earlier || c.evicting) { keep.push_back(ri); ctl_kept++; continue; }
                            c.evicting = true;
                            c.evict_req = ri;
                            res[ri].queued_s = t - r.arrival_s;
                            try_evict(r.column, t);
                            progressed = true;
                            continue;
                        }
                        if (!c.present || c.evicting) { complete(ri, t, LECO_FIX_EINVAL); progressed = true; continue; }
                        if (full || !c.ready) { keep.push_back(ri); continue; }
                        if ((r.kind == LECO_FIX_REQ_RANGE && r.begin + r.count > c.col.n) ||
                            (!r.d_out && up(result_bytes(s, r), kAlign) > s->result_cap)) {
                            const bool past = r.kind == LECO_FIX_REQ_RANGE && r.begin + r.count > c.col.n;
                            complete(ri, t, past ? LECO_FIX_EINVAL : LECO_FIX_ENOMEM);
                            progressed = true;
                            continue;
                        }
                        const bool staged = leco::staged_decode(c.col.bs);
                        const uint64_t items = r.kind == LECO_FIX_REQ_LOOKUP
                                                   ? (r.count + kLookupChunk - 1) / kLookupChunk
                                               : staged ? (r.begin + r.count - 1) / c.col.bs - r.begin / c.col.bs + 1
                                                        : (r.count + leco::kValueTile - 1) / leco::kValueTile;
                        const uint64_t idx = r.kind == LECO_FIX_REQ_LOOKUP ? r.count : 0;
                        if (admitted.size() >= s->cfg.max_batch_requests || nitems + items > kItemCap ||
                            nidx + idx > kIdxCap || (nvals + r.count > s->cfg.max_batch_values && !admitted.empty())) {
                            full = true;
                            keep.push_back(ri);
                            continue;
                        }
                        uint64_t off = 0, id = 0;
                        if (!r.d_out && !s->ring.alloc(result_bytes(s, r), &off, &id)) {  // memory admission
                            full = true;
                            keep.push_back(ri);
                            continue;
                        }
                        // stage the request's work items
                        Slot &sl = *free_slot;
                        uint8_t *out = r.d_out ? (uint8_t *)r.d_out : s->result_pool + off;
                        const uint32_t bw = c.col.width / 8;
                        if (r.kind == LECO_FIX_REQ_LOOKUP) {
                            s->copies.add(sl.h_idx + nidx, r.h_idx, r.count * sizeof(uint64_t));
                            for (uint64_t q0 = 0; q0 < r.count; q0 += kLookupChunk) {
                                leco::ServeItem &w = sl.h_items[nitems++];
                                w = leco::ServeItem{leco::kServeLookup, r.column,
                                                    (uint32_t)std::min<uint64_t>(kLookupChunk, r.count - q0), 0,
                                                    sl.h_idx + nidx + q0, out + q0 * bw, 0, 0, 0};
                            }
                            nidx += r.count;
                        } else if (staged) {
                            const uint64_t hi = r.begin + r.count;
                            for (uint64_t b = r.begin / c.col.bs; b <= (hi - 1) / c.col.bs; ++b) {
                                leco::ServeItem &w = sl.h_items[nitems++];
                                w = leco::ServeItem{leco::kServeRange, r.column, 0, 0, nullptr, out, b, r.begin, hi};
                            }
                        } else {
                            for (uint64_t q0 = 0; q0 < r.count; q0 += leco::kValueTile) {
                                leco::ServeItem &w = sl.h_items[nitems++];
                                w = leco::ServeItem{leco::kServeRangeValues, r.column,
                                                    (uint32_t)std::min<uint64_t>(leco::kValueTile, r.count - q0), 0,
                                                    nullptr, out + q0 * bw, 0, r.begin + q0, 0};
                            }
                        }
                        nvals += r.count;
                        c.refs++;
                        admitted.push_back(ri);
                        if (!r.d_out) sl.ring_ids.push_back(id);
                        res[ri].queued_s = t - r.arrival_s;
                        res[ri].batch = batch_id;
                        // where the host copy (h_out), recorded now and issued after the launch, reads from
                        sl.reqs.push_back(ri);
                        sl.reqs.push_back(r.h_out ? (uint64_t)(uintptr_t)out : 0);
                    }
                    controls = ctl_ahead + ctl_kept;  // those not reached, and those held
                    for (size_t k = 0; k < qi; ++k) queue.pop_front();
                    for (size_t k = keep.size(); k-- > 0;) queue.push_front(keep[k]);
                    p3 = Clock::now();
                    s->copies.run();
                    p4 = p5 = Clock::now();
                    it.admitted = (uint32_t)admitted.size();
                    if (!admitted.empty()) {
                        Slot &sl = *free_slot;
                        // the kernel reads the work items and lookup indices straight from the
                        // slot's pinned (fine-grained, coherent) staging over the host link: no
                        // host-to-device copy is enqueued per batch
                        hipError_t e = leco::launch_serve(s->d_cols, sl.h_items, (uint32_t)nitems,
                                                          (uint32_t)std::min<uint64_t>(nitems, kServeGrid), sl.st);
                        const Clock::time_point q2 = Clock::now();
                        it.kernel_us = (float)(secs(p4, q2) * 1e6);
                        ids.clear();
                        for (size_t k = 0; k < sl.reqs.size(); k += 2) {
                            const uint64_t ri = sl.reqs[k], src = sl.reqs[k + 1];
                            ids.push_back(ri);
                            const uint64_t bytes = result_bytes(s, reqs[ri]);
                            st.values_out += reqs[ri].count;
                            st.bytes_out += bytes;
                            if (src && e == hipSuccess)
                                e = hipMemcpyAsync(reqs[ri].h_out, (const void *)(uintptr_t)src, bytes,
                                                   hipMemcpyDeviceToHost, sl.st);
                        }
                        sl.reqs.swap(ids);
                        if (e == hipSuccess) e = hipEventRecord(sl.done, sl.st);
                        if (e != hipSuccess) return LECO_FIX_EHIP;
                        sl.busy = true;
                        sl.batch = batch_id++;
                        st.batches++;
                        st.max_batch_requests = std::max<uint64_t>(st.max_batch_requests, admitted.size());
                        progressed = true;
                        p5 = Clock::now();
                        it.event_us = (float)(secs(q2, p5) * 1e6);
                    }
                    st.peak_result_bytes = std::max(st.peak_result_bytes, s->ring.used());
                    st.peak_column_bytes = std::max(st.peak_column_bytes, s->colfree.used());
                }
                // Nothing in flight, nothing arriving, nothing admitted: what is
                // queued waits for memory that nothing will free, so it fails.
                if (!progressed && next == nreq && pending_inserts.empty() && !queue.empty()) {
                    bool busy = false;
                    for (Slot &sl : s->slots) busy = busy || sl.busy;
                    if (!busy) {
                        t = now_s();
                        // every queued request completes: one that waits for memory fails with ENOMEM,
                        // an evict of a column another evict is already dropping with EINVAL
                        for (uint64_t ri : queue) {
                            const bool dup_evict =
                                reqs[ri].kind == LECO_FIX_REQ_EVICT && s->cols[reqs[ri].column].evicting;
                            complete(ri, t, dup_evict ? LECO_FIX_EINVAL : LECO_FIX_ENOMEM);
                        }
                        queue.clear();
                    }
                }
                prev_end = Clock::now();
                it.arrivals_us = (float)(secs(p0, p1) * 1e6);
                it.completions_us = (float)(secs(p1, p2) * 1e6);
                it.admission_us = (float)(secs(p2, p3) * 1e6);
                it.copies_us = (float)(secs(p3, p4) * 1e6);
                it.launch_us = (float)(secs(p4, p5) * 1e6);
                it.rest_us = (float)(secs(p5, prev_end) * 1e6);
                s->trace.add(it);
            }
        }
        st.wall_s = nreq ? last_done - reqs[0].arrival_s : 0;
        st.peak_column_bytes = std::max(st.peak_column_bytes, s->colfree.used());
        if (stats) *stats = st;
        return LECO_FIX_OK;
    });
}

}  // extern "C"
