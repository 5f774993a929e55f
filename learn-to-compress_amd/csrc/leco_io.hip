// Column files and the reference's binary input format (C ABI in include/leco_fix.h).
//
// SOSD binary (experiments/leco_int_fix_template.cpp:21-40 of the reference,
// load_data_binary): a little-endian uint64 count, then count values of T.
// Loaded straight into HBM: the file is read in kChunk pieces into two
// pinned host buffers (hipHostMalloc) while the previous piece is copied to
// the device on the caller's stream.
//
// Column file: the compressed column as it lives in HBM -- what the
// reference keeps in block_start_vec (:108,125) plus the stream -- so a
// column is encoded once and reloaded without re-encoding:
//   [0, 64)      header: "LECOCOL1", version 1, width, n, block size, flags,
//                number of blocks, stream bytes, FNV-1a of bytes [0, 48)
//   [64, ...)    block offsets, (num_blocks + 1) little-endian uint64, from 0
//   64-aligned   the stream: the blocks' bytes back to back
// The param table is not stored; loading rebuilds it on the GPU from the
// block headers.
#include <errno.h>
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <vector>

#include "leco_internal.h"

namespace {

constexpr size_t kChunk = 64ull << 20;  // bytes per staged piece
constexpr uint32_t kFileVersion = 1;
constexpr char kMagic[8] = {'L', 'E', 'C', 'O', 'C', 'O', 'L', '1'};

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t width_bits;
    uint64_t n;
    uint32_t block_size;
    uint32_t flags;
    uint64_t num_blocks;
    uint64_t stream_bytes;
    uint64_t check;     // FNV-1a 64 of bytes [0, 48)
    uint64_t reserved;
};
static_assert(sizeof(FileHeader) == 64, "the column file header is 64 bytes");

uint64_t fnv1a(const void *p, size_t len) {
    const uint8_t *b = (const uint8_t *)p;
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < len; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

uint64_t stream_start(uint64_t num_blocks) { return (64 + 8 * (num_blocks + 1) + 63) & ~63ull; }

int hip_status(hipError_t e) { return e == hipSuccess ? LECO_FIX_OK : LECO_FIX_EHIP; }

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) close(fd);
    }
};

bool read_full(int fd, void *dst, size_t len, uint64_t off) {
    uint8_t *p = (uint8_t *)dst;
    while (len) {
        const ssize_t r = pread(fd, p, len, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        off += (uint64_t)r;
        len -= (size_t)r;
    }
    return true;
}

bool write_full(int fd, const void *src, size_t len, uint64_t off) {
    const uint8_t *p = (const uint8_t *)src;
    while (len) {
        const ssize_t r = pwrite(fd, p, len, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        off += (uint64_t)r;
        len -= (size_t)r;
    }
    return true;
}

// Two pinned staging buffers and their events, kept per host thread.
struct PinnedPair {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ready = false;
    ~PinnedPair() {
        for (int i = 0; i < 2; ++i) {
            if (buf[i]) (void)hipHostFree(buf[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
        }
    }
    int init() {
        if (ready) return LECO_FIX_OK;
        for (int i = 0; i < 2; ++i) {
            if (hipHostMalloc(&buf[i], kChunk, hipHostMallocDefault) != hipSuccess) return LECO_FIX_ENOMEM;
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return LECO_FIX_EHIP;
        }
        ready = true;
        return LECO_FIX_OK;
    }
};
thread_local PinnedPair g_pinned;

// Checks the bytes of each piece while it is in host memory, before it is
// copied: check(piece, pos, len) sees file bytes [off + pos, off + pos + len).
struct PieceCheck {
    virtual bool operator()(const uint8_t *piece, uint64_t pos, size_t len) = 0;
};

// Returns rc once the copies already enqueued on st have finished: an error
// return must not leave a copy still writing the caller's buffer (or reading
// a pinned buffer the next call refills).
int drain(hipStream_t st, int rc) {
    const hipError_t e = hipStreamSynchronize(st);
    return rc != LECO_FIX_OK ? rc : hip_status(e);
}

// File bytes [off, off + len) -> device dst; piece i+1 is read while piece i
// is copied.  A piece that fails `check` is not copied and the call returns
// LECO_FIX_EINVAL once the earlier copies are done.
int file_to_device(int fd, uint64_t off, uint8_t *dst, uint64_t len, hipStream_t st, PieceCheck *check = nullptr) {
    int rc = g_pinned.init();
    if (rc) return rc;
    for (uint64_t pos = 0, i = 0; pos < len; pos += kChunk, ++i) {
        const int k = (int)(i & 1);
        const size_t piece = (size_t)(len - pos < kChunk ? len - pos : kChunk);
        if (i >= 2 && hipEventSynchronize(g_pinned.ev[k]) != hipSuccess) return drain(st, LECO_FIX_EHIP);
        if (!read_full(fd, g_pinned.buf[k], piece, off + pos)) return drain(st, LECO_FIX_EINVAL);
        if (check && !(*check)((const uint8_t *)g_pinned.buf[k], pos, piece)) return drain(st, LECO_FIX_EINVAL);
        if (hipMemcpyAsync(dst + pos, g_pinned.buf[k], piece, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipEventRecord(g_pinned.ev[k], st) != hipSuccess)
            return drain(st, LECO_FIX_EHIP);
    }
    return hip_status(hipStreamSynchronize(st));
}

// Device src[0, len) -> file bytes [off, off + len); piece i is written while piece i+1 is copied.
int device_to_file(const uint8_t *src, uint64_t len, int fd, uint64_t off, hipStream_t st) {
    int rc = g_pinned.init();
    if (rc) return rc;
    const uint64_t pieces = (len + kChunk - 1) / kChunk;
    auto piece_len = [&](uint64_t i) { return (size_t)(len - i * kChunk < kChunk ? len - i * kChunk : kChunk); };
    for (uint64_t i = 0; i <= pieces; ++i) {
        if (i < pieces) {
            const int k = (int)(i & 1);
            if (hipMemcpyAsync(g_pinned.buf[k], src + i * kChunk, piece_len(i), hipMemcpyDeviceToHost, st) !=
                    hipSuccess ||
                hipEventRecord(g_pinned.ev[k], st) != hipSuccess)
                return LECO_FIX_EHIP;
        }
        if (i > 0) {
            const int k = (int)((i - 1) & 1);
            if (hipEventSynchronize(g_pinned.ev[k]) != hipSuccess) return LECO_FIX_EHIP;
            if (!write_full(fd, g_pinned.buf[k], piece_len(i - 1), off + (i - 1) * kChunk)) return LECO_FIX_EINVAL;
        }
    }
    return LECO_FIX_OK;
}

// The reference's text input (experiments/leco_int_fix_template.cpp:43-58
// load_data, experiments/fix_int.cpp:30-44): values read from a file stream
// in the C locale by
//     while (in.good()) { T v; in >> v; if (!in.good()) break; push(v); }
// restated as a state machine over the file's bytes.  One extraction skips
// whitespace, takes an optional sign and then decimal digits, and ends at the
// first other character (which stays in the stream); it fails when no digit
// follows or the magnitude exceeds T's maximum, and a '-' negates modulo
// 2^width, as unsigned extraction does.  The loop keeps a value only when the
// stream is still good after it, so a value that the end of the file ends
// (no character after it) is not kept.
class TextParser {
  public:
    explicit TextParser(uint32_t width_bits) : max_(width_bits == 32 ? 0xFFFFFFFFull : ~0ull) {}

    // Feeds the next bytes of the file; emit(v) for every value kept.
    // Returns false once the input has ended (a failed extraction).
    template <class Emit>
    bool feed(const uint8_t *p, size_t len, Emit &&emit) {
        for (size_t i = 0; i < len; ++i) {
            const uint8_t c = p[i];
            if (c >= '0' && c <= '9') {
                const uint64_t d = c - '0';
                if (state_ != kDigits) {
                    cur_ = 0;
                    if (state_ == kIdle) neg_ = false;
                    state_ = kDigits;
                }
                if (cur_ > (max_ - d) / 10) return false;  // does not fit T
                cur_ = cur_ * 10 + d;
            } else if (c == ' ' || (c >= '\t' && c <= '\r')) {
                if (state_ == kSign) return false;
                if (state_ == kDigits) keep(emit);
                state_ = kIdle;
            } else if (c == '-' || c == '+') {
                if (state_ == kSign) return false;
                if (state_ == kDigits) keep(emit);  // the sign starts the next extraction
                neg_ = c == '-';
                state_ = kSign;
            } else {
                if (state_ == kDigits) keep(emit);  // the next extraction fails on c
                return false;
            }
        }
        return true;
    }

  private:
    enum { kIdle, kSign, kDigits };
    template <class Emit>
    void keep(Emit &&emit) {
        const uint64_t v = neg_ ? (0 - cur_) & max_ : cur_;
        emit(v);
    }
    uint64_t max_, cur_ = 0;
    int state_ = kIdle;
    bool neg_ = false;
};

constexpr size_t kTextRead = 16ull << 20;  // file bytes parsed per read

// One pass over a text file: counts its values, and with d_out stores them
// (width_bits each) to device memory through the pinned pair on st.
int text_pass(const char *path, uint32_t width_bits, uint8_t *d_out, uint64_t capacity, uint64_t *count,
              hipStream_t st) {
    if (!path || !count || (width_bits != 32 && width_bits != 64)) return LECO_FIX_EINVAL;
    Fd f;
    f.fd = open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return LECO_FIX_EINVAL;
    if (d_out) {
        const int rc = g_pinned.init();
        if (rc) return rc;
    }
    const size_t wb = width_bits / 8, per = kChunk / wb;
    std::vector<uint8_t> buf(kTextRead);
    TextParser parser(width_bits);
    uint64_t n = 0, copied = 0, pieces = 0;
    size_t fill = 0;
    int k = 0, rc = LECO_FIX_OK;
    auto flush = [&]() {
        if (!fill || rc) return;
        if (hipMemcpyAsync(d_out + copied, g_pinned.buf[k], fill * wb, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipEventRecord(g_pinned.ev[k], st) != hipSuccess) {
            rc = LECO_FIX_EHIP;
            return;
        }
        copied += fill * wb;
        fill = 0;
        k ^= 1;
        if (++pieces >= 2 && hipEventSynchronize(g_pinned.ev[k]) != hipSuccess) rc = LECO_FIX_EHIP;
    };
    auto emit = [&](uint64_t v) {
        if (d_out && !rc) {
            if (n >= capacity) {
                rc = LECO_FIX_EINVAL;
            } else {
                if (wb == 4)
                    ((uint32_t *)g_pinned.buf[k])[fill] = (uint32_t)v;
                else
                    ((uint64_t *)g_pinned.buf[k])[fill] = v;
                if (++fill == per) flush();
            }
        }
        ++n;
    };
    for (uint64_t off = 0;;) {
        const ssize_t r = pread(f.fd, buf.data(), buf.size(), (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r < 0) {
            rc = rc != LECO_FIX_OK ? rc : LECO_FIX_EINVAL;
            break;
        }
        if (r == 0 || !parser.feed(buf.data(), (size_t)r, emit)) break;
        off += (uint64_t)r;
    }
    if (d_out) {
        flush();
        rc = drain(st, rc);  // copies already enqueued finish before any return
    }
    if (rc == LECO_FIX_OK) *count = n;
    return rc;
}

int read_header(int fd, FileHeader *h) {
    if (!read_full(fd, h, sizeof(*h), 0)) return LECO_FIX_EINVAL;
    if (memcmp(h->magic, kMagic, 8) != 0 || h->version != kFileVersion) return LECO_FIX_EINVAL;
    if (fnv1a(h, 48) != h->check) return LECO_FIX_EINVAL;
    if ((h->width_bits != 32 && h->width_bits != 64) || h->n == 0 || h->block_size == 0 ||
        h->block_size > leco::kMaxBlockSize || h->num_blocks != (h->n + h->block_size - 1) / h->block_size)
        return LECO_FIX_EINVAL;
    struct stat sb;
    if (fstat(fd, &sb) != 0 || (uint64_t)sb.st_size < stream_start(h->num_blocks) + h->stream_bytes)
        return LECO_FIX_EINVAL;
    return LECO_FIX_OK;
}

uint32_t block_len(const FileHeader &h, uint64_t b) {
    const uint64_t s = b * h.block_size;
    return (uint32_t)(h.n - s < h.block_size ? h.n - s : h.block_size);
}

// Bytes of a block of m values whose header byte is l, or 0 if no decoder
// accepts l: 0..W (W = raw), and the one- and two-value forms 127 and 126 of
// decodeArray8 (piecewise_fix_integer_template.h:152-170).
uint64_t bytes_for_header(uint32_t l, uint32_t w, uint32_t m) {
    if (l <= w) return leco::block_bytes(l, w, m);
    if (l == 127 && m == 1) return 1 + w / 8;
    if (l == 126 && m == 2) return 1 + 2 * (w / 8);
    return 0;
}

bool span_possible(uint64_t span, uint32_t w, uint32_t m) {
    for (uint32_t l = 0; l <= w; ++l)
        if (leco::block_bytes(l, w, m) == span) return true;
    return bytes_for_header(126, w, m) == span || bytes_for_header(127, w, m) == span;
}

// Header byte of each block against its span, piece by piece (offsets are
// sorted, so a cursor walks the blocks whose first byte lies in the piece).
struct HeaderCheck : PieceCheck {
    const std::vector<uint64_t> &offs;
    const FileHeader &h;
    uint64_t next = 0;
    HeaderCheck(const std::vector<uint64_t> &o, const FileHeader &hd) : offs(o), h(hd) {}
    bool operator()(const uint8_t *piece, uint64_t pos, size_t len) override {
        const uint64_t nb = h.num_blocks;
        for (; next < nb && offs[next] < pos + len; ++next) {
            if (offs[next] < pos) return false;
            const uint64_t span = offs[next + 1] - offs[next];
            if (bytes_for_header(piece[offs[next] - pos], h.width_bits, block_len(h, next)) != span) return false;
        }
        return true;
    }
};

}  // namespace

extern "C" {

int leco_fix_sosd_info(const char *path, uint32_t width_bits, uint64_t *count) {
    return leco::abi_guard([&]() -> int {
        if (!path || !count || (width_bits != 32 && width_bits != 64)) return LECO_FIX_EINVAL;
        Fd f;
        f.fd = open(path, O_RDONLY | O_CLOEXEC);
        if (f.fd < 0) return LECO_FIX_EINVAL;
        uint64_t c = 0;
        struct stat sb;
        if (!read_full(f.fd, &c, 8, 0) || fstat(f.fd, &sb) != 0) return LECO_FIX_EINVAL;
        if ((uint64_t)sb.st_size < 8 || ((uint64_t)sb.st_size - 8) / (width_bits / 8) < c) return LECO_FIX_EINVAL;
        *count = c;
        return LECO_FIX_OK;
    });
}

int leco_fix_sosd_load(const char *path, uint32_t width_bits, void *d_out, uint64_t capacity, uint64_t *count,
                       void *stream) {
    return leco::abi_guard([&]() -> int {
        uint64_t c = 0;
        int rc = leco_fix_sosd_info(path, width_bits, &c);
        if (rc) return rc;
        if (!d_out || c > capacity) return LECO_FIX_EINVAL;
        Fd f;
        f.fd = open(path, O_RDONLY | O_CLOEXEC);
        if (f.fd < 0) return LECO_FIX_EINVAL;
        rc = file_to_device(f.fd, 8, (uint8_t *)d_out, c * (width_bits / 8), (hipStream_t)stream);
        if (rc == LECO_FIX_OK && count) *count = c;
        return rc;
    });
}

int leco_fix_sosd_save(const char *path, uint32_t width_bits, const void *d_in, uint64_t count, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!path || (width_bits != 32 && width_bits != 64) || (count && !d_in)) return LECO_FIX_EINVAL;
        Fd f;
        f.fd = open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
        if (f.fd < 0) return LECO_FIX_EINVAL;
        if (!write_full(f.fd, &count, 8, 0)) return LECO_FIX_EINVAL;
        return device_to_file((const uint8_t *)d_in, count * (width_bits / 8), f.fd, 8, (hipStream_t)stream);
    });
}

int leco_fix_text_info(const char *path, uint32_t width_bits, uint64_t *count) {
    return leco::abi_guard([&]() -> int {
        return text_pass(path, width_bits, nullptr, 0, count, nullptr);
    });
}

int leco_fix_text_load(const char *path, uint32_t width_bits, void *d_out, uint64_t capacity, uint64_t *count,
                       void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!d_out || !count) return LECO_FIX_EINVAL;
        return text_pass(path, width_bits, (uint8_t *)d_out, capacity, count, (hipStream_t)stream);
    });
}

int leco_fix_column_file_info(const char *path, leco_fix_file_info *info) {
    return leco::abi_guard([&]() -> int {
        if (!path || !info) return LECO_FIX_EINVAL;
        Fd f;
        f.fd = open(path, O_RDONLY | O_CLOEXEC);
        if (f.fd < 0) return LECO_FIX_EINVAL;
        FileHeader h;
        int rc = read_header(f.fd, &h);
        if (rc) return rc;
        info->n = h.n;
        info->num_blocks = h.num_blocks;
        info->stream_bytes = h.stream_bytes;
        info->block_size = h.block_size;
        info->width_bits = h.width_bits;
        return LECO_FIX_OK;
    });
}

int leco_fix_column_save(const char *path, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                         uint32_t block_size, uint32_t width_bits, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!path || !d_stream || !d_offsets || n == 0 || block_size == 0 || block_size > leco::kMaxBlockSize ||
            (width_bits != 32 && width_bits != 64))
            return LECO_FIX_EINVAL;
        hipStream_t st = (hipStream_t)stream;
        const uint64_t nb = (n + block_size - 1) / block_size;
        std::vector<uint64_t> offs(nb + 1);
        if (hipMemcpyAsync(offs.data(), d_offsets, (nb + 1) * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return LECO_FIX_EHIP;
        // offsets may be absolute (a column inside a batch arena): store them from 0
        const uint64_t base = offs[0];
        for (uint64_t b = 0; b <= nb; ++b) {
            if (offs[b] < base || (b && offs[b] < offs[b - 1])) return LECO_FIX_EINVAL;
            offs[b] -= base;
        }
        FileHeader h{};
        memcpy(h.magic, kMagic, 8);
        h.version = kFileVersion;
        h.width_bits = width_bits;
        h.n = n;
        h.block_size = block_size;
        h.num_blocks = nb;
        h.stream_bytes = offs[nb];
        h.check = fnv1a(&h, 48);
        Fd f;
        f.fd = open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
        if (f.fd < 0) return LECO_FIX_EINVAL;
        const uint64_t start = stream_start(nb);
        std::vector<uint8_t> pad(start - 64 - 8 * (nb + 1), 0);
        if (!write_full(f.fd, &h, sizeof(h), 0) || !write_full(f.fd, offs.data(), 8 * (nb + 1), 64) ||
            (!pad.empty() && !write_full(f.fd, pad.data(), pad.size(), 64 + 8 * (nb + 1))))
            return LECO_FIX_EINVAL;
        return device_to_file(d_stream + base, h.stream_bytes, f.fd, start, st);
    });
}

int leco_fix_column_load(const char *path, uint8_t *d_stream, uint64_t stream_capacity, uint64_t *d_offsets,
                         leco_fix_block_param *d_params, void *stream) {
    return leco::abi_guard([&]() -> int {
        if (!path || !d_stream || !d_offsets) return LECO_FIX_EINVAL;
        Fd f;
        f.fd = open(path, O_RDONLY | O_CLOEXEC);
        if (f.fd < 0) return LECO_FIX_EINVAL;
        FileHeader h;
        int rc = read_header(f.fd, &h);
        if (rc) return rc;
        if (stream_capacity < h.stream_bytes + LECO_FIX_TAIL_PAD) return LECO_FIX_EINVAL;
        const uint64_t nb = h.num_blocks;
        std::vector<uint64_t> offs(nb + 1);
        if (!read_full(f.fd, offs.data(), 8 * (nb + 1), 64)) return LECO_FIX_EINVAL;
        if (offs[0] != 0 || offs[nb] != h.stream_bytes) return LECO_FIX_EINVAL;
        // every block's span must be a size its header byte can give (a shifted
        // or altered offset fails here, before any HIP call) ...
        for (uint64_t b = 0; b < nb; ++b)
            if (offs[b + 1] < offs[b] || !span_possible(offs[b + 1] - offs[b], h.width_bits, block_len(h, b)))
                return LECO_FIX_EINVAL;
        // ... and, as each piece of the stream passes through host memory, the
        // block's header byte must give exactly its span: the decoders, lookups
        // and filters read a block by its header byte, so a mismatch would make
        // them read past the stream
        HeaderCheck check(offs, h);
        hipStream_t st = (hipStream_t)stream;
        rc = file_to_device(f.fd, stream_start(nb), d_stream, h.stream_bytes, st, &check);
        if (rc) return rc;
        if (check.next != nb) return LECO_FIX_EINVAL;
        if (hipMemcpyAsync(d_offsets, offs.data(), 8 * (nb + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemsetAsync(d_stream + h.stream_bytes, 0, LECO_FIX_TAIL_PAD, st) != hipSuccess)
            return drain(st, LECO_FIX_EHIP);
        if (d_params) {
            if (leco::launch_build_params(d_stream, d_offsets, h.n, h.block_size, h.width_bits, d_params, st) !=
                hipSuccess)
                return LECO_FIX_EHIP;
            rc = hip_status(hipStreamSynchronize(st));
        }
        return rc;
    });
}

}  // extern "C"
