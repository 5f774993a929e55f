// C ABI of the gfx950 LeCo_fix codec (declared in include/leco_fix.h).
//
// Column calls validate their arguments on the host and enqueue kernels on
// the caller's stream.  The one-block host-buffer calls reproduce the
// Leco_int<T> surface (piecewise_fix_integer_template.h:27,142,305): they
// stage the block through a per-thread device workspace, run the same
// kernels and copy the result back, synchronously.
#include <hip/hip_runtime.h>
#include <string.h>

#include "leco_internal.h"

using leco::kMaxBlockSize;

namespace {

int hip_status(hipError_t e) { return e == hipSuccess ? LECO_FIX_OK : LECO_FIX_EHIP; }

bool valid_bs(uint32_t bs) { return bs >= 1 && bs <= kMaxBlockSize; }

template <typename T>
int encode_column(const T *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t cap, uint64_t *offs,
                  leco_fix_block_param *prm, void *st) {
    if (n == 0 || !valid_bs(bs) || !in || !stream || !offs || !prm) return LECO_FIX_EINVAL;
    if (cap < leco_fix_stream_bound(n, bs, 8 * sizeof(T))) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_encode<T>(in, n, bs, stream, offs, prm, (hipStream_t)st));
}

template <typename T>
int decode_range(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, uint64_t lo, uint64_t hi,
                 T *out, void *st) {
    if (!valid_bs(bs) || !stream || !offs || !out || lo >= hi || hi > n) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_decode<T>(stream, offs, n, bs, lo, hi, out, (hipStream_t)st));
}

template <typename T>
int decode_range_ex(const uint8_t *stream, const uint64_t *offs, const leco_fix_block_param *prm, uint64_t n,
                    uint32_t bs, uint64_t lo, uint64_t hi, int shape, T *out, void *st) {
    if (!valid_bs(bs) || !stream || !out || lo >= hi || hi > n || shape < 0 || shape > 2) return LECO_FIX_EINVAL;
    if (shape == LECO_FIX_DECODE_AUTO) shape = LECO_FIX_DECODE_STAGED;
    if (shape == LECO_FIX_DECODE_STAGED) {
        if (!offs) return LECO_FIX_EINVAL;
        return hip_status(leco::launch_decode<T>(stream, offs, n, bs, lo, hi, out, (hipStream_t)st));
    }
    if (!prm) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_decode_values<T>(stream, prm, bs, lo, hi, out, (hipStream_t)st));
}

template <typename T>
int lookup(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs, const uint64_t *idx,
           uint64_t nq, T *out, void *st) {
    if (!valid_bs(bs) || n == 0 || !stream || !prm) return LECO_FIX_EINVAL;
    if (nq == 0) return LECO_FIX_OK;
    if (!idx || !out) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_lookup<T>(stream, prm, n, bs, idx, nq, out, (hipStream_t)st));
}

template <typename T>
int lookup_grouped(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                   const uint64_t *idx, uint64_t nq, T *out, void *work, uint64_t work_bytes, void *st) {
    if (!valid_bs(bs) || n == 0 || n > leco::kGroupMaxValues || !stream || !prm) return LECO_FIX_EINVAL;
    if (nq == 0) return LECO_FIX_OK;
    if (!idx || !out || !work || work_bytes < leco::lookup_grouped_workspace(nq) || ((uintptr_t)work & 7))
        return LECO_FIX_EINVAL;
    return hip_status(leco::launch_lookup_grouped<T>(stream, prm, n, bs, idx, nq, out, work, (hipStream_t)st));
}

// ---- per-thread staging for the one-block host-buffer calls -------------

struct Staging {
    hipStream_t stream = nullptr;
    void *dev = nullptr;
    size_t cap = 0;
    ~Staging() {
        if (dev) (void)hipFree(dev);
        if (stream) (void)hipStreamDestroy(stream);
    }
    // Device buffer of at least `bytes`; 256-byte aligned sub-buffers are
    // carved by the caller.
    int reserve(size_t bytes) {
        if (!stream && hipStreamCreate(&stream) != hipSuccess) return LECO_FIX_EHIP;
        if (bytes <= cap) return LECO_FIX_OK;
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        cap = 0;
        if (hipMalloc(&dev, bytes) != hipSuccess) return LECO_FIX_ENOMEM;
        cap = bytes;
        return LECO_FIX_OK;
    }
};

thread_local Staging g_stage;

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T>
int encode_block_host(const T *h, uint64_t len, uint8_t *res, uint64_t rcap, uint64_t *written) {
    if (!h || !res || !written || len == 0 || len > kMaxBlockSize) return LECO_FIX_EINVAL;
    const uint32_t bs = (uint32_t)len;
    const size_t in_b = up256(len * sizeof(T));
    const size_t st_b = up256(leco_fix_stream_bound(len, bs, 8 * sizeof(T)));
    const size_t off_b = up256(2 * sizeof(uint64_t));
    const size_t prm_b = up256(sizeof(leco_fix_block_param));
    int rc = g_stage.reserve(in_b + st_b + off_b + prm_b);
    if (rc) return rc;
    uint8_t *base = (uint8_t *)g_stage.dev;
    T *d_in = (T *)base;
    uint8_t *d_st = base + in_b;
    uint64_t *d_off = (uint64_t *)(base + in_b + st_b);
    leco_fix_block_param *d_prm = (leco_fix_block_param *)(base + in_b + st_b + off_b);
    hipStream_t s = g_stage.stream;
    if (hipMemcpyAsync(d_in, h, len * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    if (leco::launch_encode<T>(d_in, len, bs, d_st, d_off, d_prm, s) != hipSuccess) return LECO_FIX_EHIP;
    uint64_t offs[2];
    if (hipMemcpyAsync(offs, d_off, sizeof(offs), hipMemcpyDeviceToHost, s) != hipSuccess) return LECO_FIX_EHIP;
    if (hipStreamSynchronize(s) != hipSuccess) return LECO_FIX_EHIP;
    if (offs[1] == LECO_FIX_OFFSETS_POISON) return LECO_FIX_EDEVICE;
    if (offs[1] > rcap) return LECO_FIX_EINVAL;
    if (hipMemcpyAsync(res, d_st, offs[1], hipMemcpyDeviceToHost, s) != hipSuccess) return LECO_FIX_EHIP;
    if (hipStreamSynchronize(s) != hipSuccess) return LECO_FIX_EHIP;
    *written = offs[1];
    return LECO_FIX_OK;
}

// Bytes of an encoded block of `len` values, from its header byte.
template <typename T>
uint64_t host_block_bytes(const uint8_t *h_in, uint64_t len) {
    constexpr uint32_t W = 8 * sizeof(T);
    const uint32_t l = h_in[0];
    if (l == 127) return 1 + sizeof(T);
    if (l == 126) return 1 + 2 * sizeof(T);
    return leco::block_bytes(l, W, (uint32_t)len);
}

template <typename T>
int decode_block_host(const uint8_t *h_in, uint64_t len, T *h_out) {
    if (!h_in || !h_out || len == 0 || len > kMaxBlockSize) return LECO_FIX_EINVAL;
    const uint64_t bytes = host_block_bytes<T>(h_in, len);
    const size_t st_b = up256(bytes + LECO_FIX_TAIL_PAD);
    const size_t off_b = up256(2 * sizeof(uint64_t));
    const size_t out_b = up256(len * sizeof(T));
    int rc = g_stage.reserve(st_b + off_b + out_b);
    if (rc) return rc;
    uint8_t *base = (uint8_t *)g_stage.dev;
    uint8_t *d_st = base;
    uint64_t *d_off = (uint64_t *)(base + st_b);
    T *d_out = (T *)(base + st_b + off_b);
    hipStream_t s = g_stage.stream;
    const uint64_t offs[2] = {0, bytes};
    if (hipMemcpyAsync(d_st, h_in, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    if (hipMemsetAsync(d_st + bytes, 0, st_b - bytes, s) != hipSuccess) return LECO_FIX_EHIP;
    if (hipMemcpyAsync(d_off, offs, sizeof(offs), hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    if (leco::launch_decode<T>(d_st, d_off, len, (uint32_t)len, 0, len, d_out, s) != hipSuccess)
        return LECO_FIX_EHIP;
    if (hipMemcpyAsync(h_out, d_out, len * sizeof(T), hipMemcpyDeviceToHost, s) != hipSuccess) return LECO_FIX_EHIP;
    return hip_status(hipStreamSynchronize(s));
}

template <typename T>
int random_decode_host(const uint8_t *h_in, uint32_t to_find, T *h_value) {
    constexpr uint32_t W = 8 * sizeof(T);
    if (!h_in || !h_value) return LECO_FIX_EINVAL;
    const uint32_t l = h_in[0];
    // Only the header and the bytes holding values [0, to_find] are read.
    const uint64_t len = (uint64_t)to_find + 1;
    uint64_t bytes;
    if (l >= W) bytes = 1 + len * sizeof(T);
    else if (l == 0) bytes = 17;
    else bytes = 17 + (len * l + 7) / 8;
    const size_t st_b = up256(bytes + LECO_FIX_TAIL_PAD);
    const size_t off_b = up256(2 * sizeof(uint64_t));
    const size_t prm_b = up256(sizeof(leco_fix_block_param));
    const size_t q_b = up256(2 * sizeof(uint64_t));
    int rc = g_stage.reserve(st_b + off_b + prm_b + q_b);
    if (rc) return rc;
    uint8_t *base = (uint8_t *)g_stage.dev;
    uint8_t *d_st = base;
    uint64_t *d_off = (uint64_t *)(base + st_b);
    leco_fix_block_param *d_prm = (leco_fix_block_param *)(base + st_b + off_b);
    uint64_t *d_q = (uint64_t *)(base + st_b + off_b + prm_b);
    T *d_v = (T *)(d_q + 1);
    hipStream_t s = g_stage.stream;
    const uint64_t offs[2] = {0, bytes};
    if (hipMemcpyAsync(d_st, h_in, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    if (hipMemsetAsync(d_st + bytes, 0, st_b - bytes, s) != hipSuccess) return LECO_FIX_EHIP;
    if (hipMemcpyAsync(d_off, offs, sizeof(offs), hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    const uint64_t q = to_find;
    if (hipMemcpyAsync(d_q, &q, sizeof(q), hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    if (leco::launch_build_params(d_st, d_off, len, (uint32_t)len, W, d_prm, s) != hipSuccess) return LECO_FIX_EHIP;
    if (leco::launch_lookup<T>(d_st, d_prm, len, (uint32_t)len, d_q, 1, d_v, s) != hipSuccess) return LECO_FIX_EHIP;
    if (hipMemcpyAsync(h_value, d_v, sizeof(T), hipMemcpyDeviceToHost, s) != hipSuccess) return LECO_FIX_EHIP;
    return hip_status(hipStreamSynchronize(s));
}

template <typename T>
int filter_column(const uint8_t *d_stream, const uint64_t *d_offs, uint64_t n, uint32_t bs, bool close,
                  uint64_t filter, const T *f1, const T *f2, int32_t base, uint32_t *d_out, uint64_t cap,
                  uint64_t *d_result, void *d_work, uint64_t work_bytes, void *st) {
    if (n == 0 || n > (1ull << 32) || !valid_bs(bs) || !d_stream || !d_offs || !d_result || !d_work ||
        (cap && !d_out) || (close && (!f1 || !f2)))
        return LECO_FIX_EINVAL;
    if (work_bytes < leco_fix_filter_workspace_bytes(n, bs)) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_filter<T>(d_stream, d_offs, n, bs, close, filter, f1, f2, base, 0, d_out, cap,
                                             d_result, (uint64_t *)d_work, (hipStream_t)st));
}

// One block in host memory through the GPU path: staged, filtered as a
// one-block column with positions from block_id * block_size, copied back.
template <typename T>
int filter_block_host(const uint8_t *h_in, uint64_t len, bool close, uint64_t filter, T f1, T f2, int32_t base,
                      uint32_t *h_out, uint64_t cap, int32_t block_id, uint32_t block_size, uint64_t *count) {
    if (!h_in || !count || len == 0 || len > kMaxBlockSize || (cap && !h_out)) return LECO_FIX_EINVAL;
    const uint64_t bytes = host_block_bytes<T>(h_in, len);
    const uint32_t pos_base = (uint32_t)block_id * block_size;
    uint64_t dcap = cap < len + 1 ? cap : len + 1;
    for (;;) {
        const size_t st_b = up256(bytes + LECO_FIX_TAIL_PAD);
        const size_t small_b = up256(6 * sizeof(uint64_t) + leco_fix_filter_workspace_bytes(len, (uint32_t)len));
        const size_t out_b = up256((dcap ? dcap : 1) * sizeof(uint32_t));
        int rc = g_stage.reserve(st_b + small_b + out_b);
        if (rc) return rc;
        uint8_t *base_p = (uint8_t *)g_stage.dev;
        uint8_t *d_st = base_p;
        uint64_t *d_small = (uint64_t *)(base_p + st_b);  // offsets[2], f1, f2, result[2], work[...]
        uint32_t *d_out = (uint32_t *)(base_p + st_b + small_b);
        hipStream_t s = g_stage.stream;
        uint64_t small[8] = {0, bytes, 0, 0, 0, 0, 0, 0};
        memcpy(&small[2], &f1, sizeof(T));
        memcpy(&small[3], &f2, sizeof(T));
        if (hipMemcpyAsync(d_st, h_in, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
        if (hipMemsetAsync(d_st + bytes, 0, st_b - bytes, s) != hipSuccess) return LECO_FIX_EHIP;
        if (hipMemcpyAsync(d_small, small, sizeof(small), hipMemcpyHostToDevice, s) != hipSuccess)
            return LECO_FIX_EHIP;
        if (leco::launch_filter<T>(d_st, d_small, len, (uint32_t)len, close, filter, (const T *)&d_small[2],
                                   (const T *)&d_small[3], base, pos_base, d_out, dcap, &d_small[4], &d_small[6],
                                   s) != hipSuccess)
            return LECO_FIX_EHIP;
        uint64_t res[2];
        if (hipMemcpyAsync(res, &d_small[4], sizeof(res), hipMemcpyDeviceToHost, s) != hipSuccess)
            return LECO_FIX_EHIP;
        if (hipStreamSynchronize(s) != hipSuccess) return LECO_FIX_EHIP;
        if (res[0] > dcap && dcap < cap) {  // more matches than staged room: once more with room for all
            dcap = res[0] < cap ? res[0] : cap;
            continue;
        }
        const uint64_t ncopy = res[0] < cap ? res[0] : cap;
        if (ncopy && hipMemcpyAsync(h_out, d_out, ncopy * sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess)
            return LECO_FIX_EHIP;
        if (hipStreamSynchronize(s) != hipSuccess) return LECO_FIX_EHIP;
        *count = res[0];
        return (res[1] & 1) ? LECO_FIX_ELIMIT : LECO_FIX_OK;
    }
}

// ---- FOR_int / Delta_int baselines ----

bool valid_codec(int codec) { return codec == leco::kCodecFor || codec == leco::kCodecDelta; }

uint64_t base_block_bound(int codec, uint32_t w, uint32_t m) {
    uint64_t worst = 0;
    for (uint32_t l = 0; l <= w; ++l) {
        const uint64_t b = codec == leco::kCodecFor ? leco::for_block_bytes(l, w, m) : leco::delta_block_bytes(l, w, m);
        worst = b > worst ? b : worst;
    }
    return worst;
}

// Bytes of a stored FOR / Delta block of `len` values, from its header byte.
template <typename T>
uint64_t base_host_block_bytes(int codec, const uint8_t *h_in, uint64_t len) {
    constexpr uint32_t W = 8 * sizeof(T);
    const uint32_t l = h_in[0];
    if (codec == leco::kCodecDelta && (l == 126 || l == 127)) return 1 + (128 - l) * sizeof(T);
    return codec == leco::kCodecFor ? leco::for_block_bytes(l, W, (uint32_t)len)
                                    : leco::delta_block_bytes(l, W, (uint32_t)len);
}

template <typename T>
int base_encode_block_host(int codec, const T *h, uint64_t len, uint8_t *res, uint64_t rcap, uint64_t *written) {
    if (!valid_codec(codec) || !h || !res || !written || len == 0 || len > kMaxBlockSize) return LECO_FIX_EINVAL;
    constexpr uint32_t W = 8 * sizeof(T);
    const size_t in_b = up256(len * sizeof(T));
    const size_t st_b = up256(base_block_bound(codec, W, (uint32_t)len) + LECO_FIX_TAIL_PAD);
    const size_t off_b = up256(2 * sizeof(uint64_t));
    int rc = g_stage.reserve(in_b + st_b + off_b);
    if (rc) return rc;
    uint8_t *base = (uint8_t *)g_stage.dev;
    T *d_in = (T *)base;
    uint8_t *d_st = base + in_b;
    uint64_t *d_off = (uint64_t *)(base + in_b + st_b);
    hipStream_t s = g_stage.stream;
    if (hipMemcpyAsync(d_in, h, len * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    if (leco::launch_base_encode<T>(codec, d_in, len, (uint32_t)len, d_st, d_off, s) != hipSuccess)
        return LECO_FIX_EHIP;
    uint64_t offs[2];
    if (hipMemcpyAsync(offs, d_off, sizeof(offs), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LECO_FIX_EHIP;
    if (offs[1] == LECO_FIX_OFFSETS_POISON) return LECO_FIX_EDEVICE;
    if (offs[1] > rcap) return LECO_FIX_EINVAL;
    if (hipMemcpyAsync(res, d_st, offs[1], hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LECO_FIX_EHIP;
    *written = offs[1];
    return LECO_FIX_OK;
}

// The block staged to the device with a 2-entry offsets table; returns the staged pointers.
template <typename T>
int stage_block(int codec, const uint8_t *h_in, uint64_t len, size_t extra, uint8_t **d_st, uint64_t **d_off,
                uint8_t **d_extra) {
    const uint64_t bytes = base_host_block_bytes<T>(codec, h_in, len);
    const size_t st_b = up256(bytes + LECO_FIX_TAIL_PAD);
    const size_t off_b = up256(2 * sizeof(uint64_t));
    int rc = g_stage.reserve(st_b + off_b + up256(extra));
    if (rc) return rc;
    uint8_t *base = (uint8_t *)g_stage.dev;
    *d_st = base;
    *d_off = (uint64_t *)(base + st_b);
    *d_extra = base + st_b + off_b;
    hipStream_t s = g_stage.stream;
    const uint64_t offs[2] = {0, bytes};
    if (hipMemcpyAsync(*d_st, h_in, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(*d_st + bytes, 0, st_b - bytes, s) != hipSuccess ||
        hipMemcpyAsync(*d_off, offs, sizeof(offs), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LECO_FIX_EHIP;
    return LECO_FIX_OK;
}

template <typename T>
int base_decode_block_host(int codec, const uint8_t *h_in, uint64_t len, T *h_out) {
    if (!valid_codec(codec) || !h_in || !h_out || len == 0 || len > kMaxBlockSize) return LECO_FIX_EINVAL;
    uint8_t *d_st, *d_x;
    uint64_t *d_off;
    int rc = stage_block<T>(codec, h_in, len, len * sizeof(T), &d_st, &d_off, &d_x);
    if (rc) return rc;
    hipStream_t s = g_stage.stream;
    if (leco::launch_base_decode<T>(codec, d_st, d_off, len, (uint32_t)len, (T *)d_x, s) != hipSuccess ||
        hipMemcpyAsync(h_out, d_x, len * sizeof(T), hipMemcpyDeviceToHost, s) != hipSuccess)
        return LECO_FIX_EHIP;
    return hip_status(hipStreamSynchronize(s));
}

template <typename T>
int base_random_decode_host(int codec, const uint8_t *h_in, uint64_t len, uint32_t to_find, T *h_value) {
    if (!valid_codec(codec) || !h_in || !h_value || len == 0 || len > kMaxBlockSize || to_find >= len)
        return LECO_FIX_EINVAL;
    uint8_t *d_st, *d_x;
    uint64_t *d_off;
    int rc = stage_block<T>(codec, h_in, len, 16, &d_st, &d_off, &d_x);
    if (rc) return rc;
    hipStream_t s = g_stage.stream;
    const uint64_t q = to_find;
    if (hipMemcpyAsync(d_x, &q, sizeof(q), hipMemcpyHostToDevice, s) != hipSuccess ||
        leco::launch_base_lookup<T>(codec, d_st, d_off, len, (uint32_t)len, (const uint64_t *)d_x, 1, (T *)(d_x + 8),
                                    s) != hipSuccess ||
        hipMemcpyAsync(h_value, d_x + 8, sizeof(T), hipMemcpyDeviceToHost, s) != hipSuccess)
        return LECO_FIX_EHIP;
    return hip_status(hipStreamSynchronize(s));
}

template <typename T>
int for_filter_block_host(const uint8_t *h_in, uint64_t len, bool close, uint64_t f1, uint64_t f2, int32_t base,
                          uint32_t *h_out, uint64_t cap, int32_t block_id, uint32_t block_size, uint64_t *count) {
    if (!h_in || !count || len == 0 || len > kMaxBlockSize || (cap && !h_out) || (close && base <= 0))
        return LECO_FIX_EINVAL;
    const uint64_t dcap = cap < len ? cap : len;
    uint8_t *d_st, *d_x;
    uint64_t *d_off;
    const size_t small_b = up256(16 + leco_fix_filter_workspace_bytes(len, (uint32_t)len));
    int rc = stage_block<T>(leco::kCodecFor, h_in, len, small_b + (dcap ? dcap : 1) * 4, &d_st, &d_off, &d_x);
    if (rc) return rc;
    hipStream_t s = g_stage.stream;
    uint64_t *d_res = (uint64_t *)d_x, *d_work = d_res + 2;
    uint32_t *d_out = (uint32_t *)(d_x + small_b);
    uint64_t res[2];
    if (leco::launch_for_filter<T>(d_st, d_off, len, (uint32_t)len, close, f1, f2, base,
                                   (uint32_t)block_id * block_size, d_out, dcap, d_res, d_work, s) != hipSuccess ||
        hipMemcpyAsync(res, d_res, sizeof(res), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LECO_FIX_EHIP;
    const uint64_t ncopy = res[0] < cap ? res[0] : cap;
    if (ncopy && (hipMemcpyAsync(h_out, d_out, ncopy * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess))
        return LECO_FIX_EHIP;
    *count = res[0];
    return LECO_FIX_OK;
}

template <typename T>
int for_filter_column(const uint8_t *d_stream, const uint64_t *d_offs, uint64_t n, uint32_t bs, bool close,
                      uint64_t f1, uint64_t f2, int32_t base, uint32_t *d_out, uint64_t cap, uint64_t *d_result,
                      void *d_work, uint64_t work_bytes, void *st) {
    if (n == 0 || n > (1ull << 32) || !valid_bs(bs) || !d_stream || !d_offs || !d_result || !d_work ||
        (cap && !d_out) || (close && base <= 0))
        return LECO_FIX_EINVAL;
    if (work_bytes < leco_fix_filter_workspace_bytes(n, bs)) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_for_filter<T>(d_stream, d_offs, n, bs, close, f1, f2, base, 0, d_out, cap,
                                                 d_result, (uint64_t *)d_work, (hipStream_t)st));
}

// ---- piecewise_fix_op_max --------------------------------------------------

int opmax_encode_block_host(const uint32_t *h, uint64_t len, uint8_t *res, uint64_t rcap, uint64_t *written,
                            int8_t *h_fix) {
    if (!h || !res || !written || len == 0 || len > kMaxBlockSize) return LECO_FIX_EINVAL;
    const size_t in_b = up256(len * 4);
    const size_t st_b = up256(leco::opmax_block_bound((uint32_t)len) + LECO_FIX_TAIL_PAD);
    const size_t off_b = up256(2 * sizeof(uint64_t));
    const size_t prm_b = up256(sizeof(leco_fix_block_param));
    const size_t fix_b = up256(len);
    int rc = g_stage.reserve(in_b + st_b + off_b + prm_b + fix_b);
    if (rc) return rc;
    uint8_t *base = (uint8_t *)g_stage.dev;
    uint32_t *d_in = (uint32_t *)base;
    uint8_t *d_st = base + in_b;
    uint64_t *d_off = (uint64_t *)(base + in_b + st_b);
    leco_fix_block_param *d_prm = (leco_fix_block_param *)(base + in_b + st_b + off_b);
    int8_t *d_fix = (int8_t *)(base + in_b + st_b + off_b + prm_b);
    hipStream_t s = g_stage.stream;
    if (hipMemcpyAsync(d_in, h, len * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LECO_FIX_EHIP;
    if (leco::launch_opmax_encode(d_in, len, (uint32_t)len, d_st, d_off, d_prm, h_fix ? d_fix : nullptr, s) !=
        hipSuccess)
        return LECO_FIX_EHIP;
    uint64_t offs[2];
    if (hipMemcpyAsync(offs, d_off, sizeof(offs), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LECO_FIX_EHIP;
    if (offs[1] == LECO_FIX_OFFSETS_POISON) return LECO_FIX_EDEVICE;
    if (offs[1] > rcap) return LECO_FIX_EINVAL;
    if (hipMemcpyAsync(res, d_st, offs[1], hipMemcpyDeviceToHost, s) != hipSuccess ||
        (h_fix && hipMemcpyAsync(h_fix, d_fix, len, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess)
        return LECO_FIX_EHIP;
    *written = offs[1];
    return LECO_FIX_OK;
}

// The first `bytes` bytes of a stored block staged (zero-padded) with a
// 2-entry offsets table, its params and `extra` bytes.
int opmax_stage_block(const uint8_t *h_in, uint64_t len, uint64_t bytes, size_t extra, uint8_t **d_st,
                      uint64_t **d_off, leco_fix_block_param **d_prm, uint8_t **d_extra) {
    const size_t st_b = up256(bytes + LECO_FIX_TAIL_PAD);
    const size_t off_b = up256(2 * sizeof(uint64_t));
    const size_t prm_b = up256(sizeof(leco_fix_block_param));
    int rc = g_stage.reserve(st_b + off_b + prm_b + up256(extra));
    if (rc) return rc;
    uint8_t *base = (uint8_t *)g_stage.dev;
    *d_st = base;
    *d_off = (uint64_t *)(base + st_b);
    *d_prm = (leco_fix_block_param *)(base + st_b + off_b);
    *d_extra = base + st_b + off_b + prm_b;
    hipStream_t s = g_stage.stream;
    const uint64_t offs[2] = {0, bytes};
    if (hipMemcpyAsync(*d_st, h_in, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(*d_st + bytes, 0, st_b - bytes, s) != hipSuccess ||
        hipMemcpyAsync(*d_off, offs, sizeof(offs), hipMemcpyHostToDevice, s) != hipSuccess ||
        leco::launch_opmax_params(*d_st, *d_off, len, (uint32_t)len, *d_prm, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LECO_FIX_EHIP;
    return LECO_FIX_OK;
}

}  // namespace

extern "C" {

uint64_t leco_fix_opmax_stream_bound(uint64_t n, uint32_t bs) {
    if (n == 0 || bs == 0) return 0;
    const uint64_t nb = (n + bs - 1) / bs, last = n - (nb - 1) * bs;
    return (nb - 1) * leco::opmax_block_bound(bs) + leco::opmax_block_bound((uint32_t)last) + LECO_FIX_TAIL_PAD;
}

int leco_fix_opmax_encode_column(const uint32_t *d_in, uint64_t n, uint32_t bs, uint8_t *d_stream, uint64_t cap,
                                 uint64_t *d_offsets, leco_fix_block_param *d_params, int8_t *d_fix, void *stream) {
    if (n == 0 || !valid_bs(bs) || !d_in || !d_stream || !d_offsets || !d_params ||
        cap < leco_fix_opmax_stream_bound(n, bs))
        return LECO_FIX_EINVAL;
    return hip_status(
        leco::launch_opmax_encode(d_in, n, bs, d_stream, d_offsets, d_params, d_fix, (hipStream_t)stream));
}

int leco_fix_opmax_decode_column(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                                 const int8_t *d_fix, uint32_t *d_out, void *stream) {
    if (n == 0 || !valid_bs(bs) || !d_stream || !d_offsets || !d_out) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_opmax_decode(d_stream, d_offsets, n, bs, d_fix, d_out, (hipStream_t)stream));
}

int leco_fix_opmax_decode_column_lists(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                                       const int32_t *d_plus, uint64_t n_plus, const int32_t *d_minus,
                                       uint64_t n_minus, uint32_t *d_out, void *stream) {
    if (n == 0 || !valid_bs(bs) || !d_stream || !d_offsets || !d_out || (n_plus && !d_plus) ||
        (n_minus && !d_minus))
        return LECO_FIX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (leco::launch_opmax_decode(d_stream, d_offsets, n, bs, nullptr, d_out, st) != hipSuccess)
        return LECO_FIX_EHIP;
    return hip_status(
        leco::launch_opmax_apply_exact(d_out, n, bs, d_stream, d_offsets, d_plus, n_plus, d_minus, n_minus, st));
}

int leco_fix_opmax_lookup(const uint8_t *d_stream, const leco_fix_block_param *d_params, uint64_t n, uint32_t bs,
                          const uint64_t *d_idx, uint64_t nq, uint32_t *d_out, void *stream) {
    if (n == 0 || !valid_bs(bs) || !d_stream || !d_params) return LECO_FIX_EINVAL;
    if (nq == 0) return LECO_FIX_OK;
    if (!d_idx || !d_out) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_opmax_lookup(d_stream, d_params, n, bs, d_idx, nq, d_out, (hipStream_t)stream));
}

int leco_fix_opmax_build_params(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                                leco_fix_block_param *d_params, void *stream) {
    if (n == 0 || !valid_bs(bs) || !d_stream || !d_offsets || !d_params) return LECO_FIX_EINVAL;
    return hip_status(leco::launch_opmax_params(d_stream, d_offsets, n, bs, d_params, (hipStream_t)stream));
}

uint64_t leco_fix_opmax_lists_workspace_bytes(uint64_t n) {
    return n ? 16 * ((n + leco::kFixTile - 1) / leco::kFixTile) : 0;
}

int leco_fix_opmax_fix_lists(const int8_t *d_fix, uint64_t n, int32_t *d_plus, int32_t *d_minus, uint64_t capacity,
                             uint64_t *d_counts, void *d_work, uint64_t work_bytes, void *stream) {
    if (n == 0 || n > 0x80000000ull || !d_fix || !d_counts || !d_work || (capacity && (!d_plus || !d_minus)) ||
        work_bytes < leco_fix_opmax_lists_workspace_bytes(n))
        return LECO_FIX_EINVAL;
    return hip_status(leco::launch_fix_lists(d_fix, n, d_plus, d_minus, capacity, d_counts, (uint64_t *)d_work,
                                             (hipStream_t)stream));
}

int leco_fix_opmax_apply_lists(uint32_t *d_values, uint64_t n, const int32_t *d_plus, uint64_t n_plus,
                               const int32_t *d_minus, uint64_t n_minus, void *stream) {
    if (!d_values || (n_plus && !d_plus) || (n_minus && !d_minus)) return LECO_FIX_EINVAL;
    return hip_status(
        leco::launch_apply_lists(d_values, n, d_plus, n_plus, d_minus, n_minus, (hipStream_t)stream));
}

int leco_fix_opmax_encode_block(const uint32_t *h_in, uint64_t length, uint8_t *h_res, uint64_t res_capacity,
                                uint64_t *written, int8_t *h_fix) {
    return opmax_encode_block_host(h_in, length, h_res, res_capacity, written, h_fix);
}

int leco_fix_opmax_decode_block(const uint8_t *h_in, uint64_t length, uint32_t *h_out) {
    if (!h_in || !h_out || length == 0 || length > kMaxBlockSize) return LECO_FIX_EINVAL;
    uint8_t *d_st, *d_x;
    uint64_t *d_off;
    leco_fix_block_param *d_prm;
    int rc = opmax_stage_block(h_in, length, leco::opmax_block_bytes(h_in[0], (uint32_t)length), length * 4, &d_st,
                               &d_off, &d_prm, &d_x);
    if (rc) return rc;
    hipStream_t s = g_stage.stream;
    if (leco::launch_opmax_decode(d_st, d_off, length, (uint32_t)length, nullptr, (uint32_t *)d_x, s) !=
            hipSuccess ||
        hipMemcpyAsync(h_out, d_x, length * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return LECO_FIX_EHIP;
    return hip_status(hipStreamSynchronize(s));
}

int leco_fix_opmax_random_decode(const uint8_t *h_in, uint64_t length, uint32_t to_find, uint32_t *h_value) {
    if (!h_in || !h_value || length == 0 || length > kMaxBlockSize || to_find >= length) return LECO_FIX_EINVAL;
    uint8_t *d_st, *d_x;
    uint64_t *d_off;
    leco_fix_block_param *d_prm;
    // only the bytes up to the value's field (randomdecodeArray8 reads no further)
    const uint32_t l = h_in[0];
    const uint64_t need = l == 0 ? leco::kOpmaxHdr
                          : l >= 31 ? leco::kOpmaxHdr + 4ull * (to_find + 1)
                                    : leco::kOpmaxHdr + ((uint64_t)l * (to_find + 1) + 7) / 8;
    int rc = opmax_stage_block(h_in, length, need, 16, &d_st, &d_off, &d_prm, &d_x);
    if (rc) return rc;
    hipStream_t s = g_stage.stream;
    const uint64_t q = to_find;
    if (hipMemcpyAsync(d_x, &q, sizeof(q), hipMemcpyHostToDevice, s) != hipSuccess ||
        leco::launch_opmax_lookup(d_st, d_prm, length, (uint32_t)length, (const uint64_t *)d_x, 1,
                                  (uint32_t *)(d_x + 8), s) != hipSuccess ||
        hipMemcpyAsync(h_value, d_x + 8, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return LECO_FIX_EHIP;
    return hip_status(hipStreamSynchronize(s));
}

uint64_t leco_fix_base_stream_bound(int codec, uint64_t n, uint32_t bs, uint32_t w) {
    if (!valid_codec(codec) || n == 0 || bs == 0 || (w != 32 && w != 64)) return 0;
    const uint64_t nb = (n + bs - 1) / bs, last = n - (nb - 1) * bs;
    return (nb - 1) * base_block_bound(codec, w, bs) + base_block_bound(codec, w, (uint32_t)last) +
           LECO_FIX_TAIL_PAD;
}

#define LECO_BASE_ABI(SFX, T)                                                                                      \
    int leco_fix_base_encode_column_##SFX(int codec, const T *d_in, uint64_t n, uint32_t bs, uint8_t *d_stream,    \
                                          uint64_t cap, uint64_t *d_offsets, void *stream) {                       \
        if (!valid_codec(codec) || n == 0 || !valid_bs(bs) || !d_in || !d_stream || !d_offsets ||                  \
            cap < leco_fix_base_stream_bound(codec, n, bs, 8 * sizeof(T)))                                         \
            return LECO_FIX_EINVAL;                                                                                \
        return hip_status(leco::launch_base_encode<T>(codec, d_in, n, bs, d_stream, d_offsets, (hipStream_t)stream)); \
    }                                                                                                              \
    int leco_fix_base_decode_column_##SFX(int codec, const uint8_t *d_stream, const uint64_t *d_offsets,           \
                                          uint64_t n, uint32_t bs, T *d_out, void *stream) {                       \
        if (!valid_codec(codec) || n == 0 || !valid_bs(bs) || !d_stream || !d_offsets || !d_out)                   \
            return LECO_FIX_EINVAL;                                                                                \
        return hip_status(leco::launch_base_decode<T>(codec, d_stream, d_offsets, n, bs, d_out, (hipStream_t)stream)); \
    }                                                                                                              \
    int leco_fix_base_lookup_##SFX(int codec, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,      \
                                   uint32_t bs, const uint64_t *d_idx, uint64_t nq, T *d_out, void *stream) {      \
        if (!valid_codec(codec) || n == 0 || !valid_bs(bs) || !d_stream || !d_offsets) return LECO_FIX_EINVAL;     \
        if (nq == 0) return LECO_FIX_OK;                                                                           \
        if (!d_idx || !d_out) return LECO_FIX_EINVAL;                                                              \
        return hip_status(                                                                                         \
            leco::launch_base_lookup<T>(codec, d_stream, d_offsets, n, bs, d_idx, nq, d_out, (hipStream_t)stream)); \
    }                                                                                                              \
    int leco_fix_for_filter_range_##SFX(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,            \
                                        uint32_t bs, T filter, uint32_t *d_out, uint64_t cap, uint64_t *d_result,  \
                                        void *d_work, uint64_t work_bytes, void *stream) {                         \
        return for_filter_column<T>(d_stream, d_offsets, n, bs, false, filter, 0, 1, d_out, cap, d_result, d_work, \
                                    work_bytes, stream);                                                           \
    }                                                                                                              \
    int leco_fix_for_filter_range_close_##SFX(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,      \
                                              uint32_t bs, T filter1, T filter2, int32_t base, uint32_t *d_out,    \
                                              uint64_t cap, uint64_t *d_result, void *d_work, uint64_t work_bytes, \
                                              void *stream) {                                                      \
        return for_filter_column<T>(d_stream, d_offsets, n, bs, true, filter1, filter2, base, d_out, cap, d_result, \
                                    d_work, work_bytes, stream);                                                   \
    }                                                                                                              \
    int leco_fix_base_encode_block_##SFX(int codec, const T *h, uint64_t len, uint8_t *res, uint64_t cap,          \
                                         uint64_t *written) {                                                      \
        return base_encode_block_host<T>(codec, h, len, res, cap, written);                                        \
    }                                                                                                              \
    int leco_fix_base_decode_block_##SFX(int codec, const uint8_t *h_in, uint64_t len, T *h_out) {                 \
        return base_decode_block_host<T>(codec, h_in, len, h_out);                                                 \
    }                                                                                                              \
    int leco_fix_base_random_decode_##SFX(int codec, const uint8_t *h_in, uint64_t len, uint32_t to_find,          \
                                          T *h_value) {                                                            \
        return base_random_decode_host<T>(codec, h_in, len, to_find, h_value);                                     \
    }                                                                                                              \
    int leco_fix_for_filter_block_##SFX(const uint8_t *h_in, uint64_t len, T filter, uint32_t *h_out, uint64_t cap, \
                                        int32_t block_id, uint32_t block_size, uint64_t *count) {                  \
        return for_filter_block_host<T>(h_in, len, false, filter, 0, 1, h_out, cap, block_id, block_size, count);  \
    }                                                                                                              \
    int leco_fix_for_filter_close_block_##SFX(const uint8_t *h_in, uint64_t len, uint32_t *h_out, uint64_t cap,    \
                                              int32_t block_id, uint32_t block_size, T filter1, T filter2,         \
                                              int32_t base, uint64_t *count) {                                     \
        return for_filter_block_host<T>(h_in, len, true, filter1, filter2, base, h_out, cap, block_id, block_size, \
                                        count);                                                                    \
    }
LECO_BASE_ABI(u32, uint32_t)
LECO_BASE_ABI(u64, uint64_t)
#undef LECO_BASE_ABI

uint64_t leco_fix_filter_workspace_bytes(uint64_t n, uint32_t block_size) {
    if (n == 0 || block_size == 0) return 0;
    // group sums, counts and match bitmaps of the larger of the two filters' tile sets
    const uint64_t close_b = leco::filter_workspace_for(leco::filter_tiles(n, block_size));
    const uint64_t gt_b = leco::filter_workspace_for(leco::gt_tiles(n, block_size));
    return close_b > gt_b ? close_b : gt_b;
}

int leco_fix_filter_range_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                              uint32_t filter, uint32_t *d_out, uint64_t cap, uint64_t *d_result, void *d_work,
                              uint64_t work_bytes, void *stream) {
    return filter_column<uint32_t>(d_stream, d_offsets, n, bs, false, filter, nullptr, nullptr, 0, d_out, cap,
                                   d_result, d_work, work_bytes, stream);
}
int leco_fix_filter_range_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                              uint64_t filter, uint32_t *d_out, uint64_t cap, uint64_t *d_result, void *d_work,
                              uint64_t work_bytes, void *stream) {
    return filter_column<uint64_t>(d_stream, d_offsets, n, bs, false, filter, nullptr, nullptr, 0, d_out, cap,
                                   d_result, d_work, work_bytes, stream);
}
int leco_fix_filter_range_close_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                                    const uint32_t *f1, const uint32_t *f2, int32_t base, uint32_t *d_out,
                                    uint64_t cap, uint64_t *d_result, void *d_work, uint64_t work_bytes,
                                    void *stream) {
    return filter_column<uint32_t>(d_stream, d_offsets, n, bs, true, 0, f1, f2, base, d_out, cap, d_result, d_work,
                                   work_bytes, stream);
}
int leco_fix_filter_range_close_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                                    const uint64_t *f1, const uint64_t *f2, int32_t base, uint32_t *d_out,
                                    uint64_t cap, uint64_t *d_result, void *d_work, uint64_t work_bytes,
                                    void *stream) {
    return filter_column<uint64_t>(d_stream, d_offsets, n, bs, true, 0, f1, f2, base, d_out, cap, d_result, d_work,
                                   work_bytes, stream);
}
int leco_fix_filter_block_u32(const uint8_t *h_in, uint64_t len, uint32_t filter, uint32_t *h_out, uint64_t cap,
                              int32_t block_id, uint32_t block_size, uint64_t *count) {
    return filter_block_host<uint32_t>(h_in, len, false, filter, 0, 0, 0, h_out, cap, block_id, block_size, count);
}
int leco_fix_filter_block_u64(const uint8_t *h_in, uint64_t len, uint64_t filter, uint32_t *h_out, uint64_t cap,
                              int32_t block_id, uint32_t block_size, uint64_t *count) {
    return filter_block_host<uint64_t>(h_in, len, false, filter, 0, 0, 0, h_out, cap, block_id, block_size, count);
}
int leco_fix_filter_close_block_u32(const uint8_t *h_in, uint64_t len, uint32_t *h_out, uint64_t cap,
                                    int32_t block_id, uint32_t block_size, uint32_t f1, uint32_t f2, int32_t base,
                                    uint64_t *count) {
    return filter_block_host<uint32_t>(h_in, len, true, 0, f1, f2, base, h_out, cap, block_id, block_size, count);
}
int leco_fix_filter_close_block_u64(const uint8_t *h_in, uint64_t len, uint32_t *h_out, uint64_t cap,
                                    int32_t block_id, uint32_t block_size, uint64_t f1, uint64_t f2, int32_t base,
                                    uint64_t *count) {
    return filter_block_host<uint64_t>(h_in, len, true, 0, f1, f2, base, h_out, cap, block_id, block_size, count);
}

int leco_fix_encode_result(const uint64_t *d_offsets, uint64_t nb, uint64_t *stream_bytes, void *stream) {
    if (!d_offsets || !stream_bytes) return LECO_FIX_EINVAL;
    uint64_t end = 0;
    const hipStream_t st = (hipStream_t)stream;
    if (hipMemcpyAsync(&end, d_offsets + nb, sizeof(end), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return LECO_FIX_EHIP;
    if (end == LECO_FIX_OFFSETS_POISON) return LECO_FIX_EDEVICE;
    *stream_bytes = end;
    return LECO_FIX_OK;
}

int leco_fix_test_lookback_hook(uint32_t spin_limit, uint64_t withhold) {
    if (leco::set_lookback_hook_leco(spin_limit, withhold) != hipSuccess ||
        leco::set_lookback_hook_base(spin_limit, withhold) != hipSuccess ||
        leco::set_lookback_hook_opmax(spin_limit, withhold) != hipSuccess)
        return LECO_FIX_EHIP;
    return hip_status(hipDeviceSynchronize());
}

const char *leco_fix_strerror(int code) {
    switch (code) {
    case LECO_FIX_OK: return "ok";
    case LECO_FIX_EINVAL: return "invalid argument";
    case LECO_FIX_EHIP: return "HIP runtime error";
    case LECO_FIX_ENOMEM: return "out of memory";
    case LECO_FIX_ELIMIT: return "a filter_range_close block needed more than 65536 periods";
    case LECO_FIX_EDEVICE: return "a kernel found a broken invariant (look-back wait gave up); the output is not valid";
    case LECO_FIX_ESCRATCH: return "the input needs more workspace than the library's static scratch holds";
    default: return "unknown error";
    }
}

uint64_t leco_fix_block_count(uint64_t n, uint32_t block_size) {
    return block_size ? (n + block_size - 1) / block_size : 0;
}

uint64_t leco_fix_block_bytes(uint32_t bits, uint32_t width_bits, uint32_t length) {
    return leco::block_bytes(bits, width_bits, length);
}

uint64_t leco_fix_stream_bound(uint64_t n, uint32_t bs, uint32_t w) {
    if (n == 0 || bs == 0 || (w != 32 && w != 64)) return 0;
    const uint64_t nb = (n + bs - 1) / bs, last = n - (nb - 1) * bs;
    auto worst = [w](uint32_t m) {
        uint64_t a = leco::block_bytes(w - 1, w, m), b = leco::block_bytes(w, w, m);
        return a > b ? a : b;
    };
    return (nb - 1) * worst(bs) + worst((uint32_t)last) + LECO_FIX_TAIL_PAD;
}

int leco_fix_encode_column_u32(const uint32_t *d_in, uint64_t n, uint32_t bs, uint8_t *d_stream, uint64_t cap,
                               uint64_t *d_offsets, leco_fix_block_param *d_params, void *stream) {
    return encode_column<uint32_t>(d_in, n, bs, d_stream, cap, d_offsets, d_params, stream);
}
int leco_fix_encode_column_u64(const uint64_t *d_in, uint64_t n, uint32_t bs, uint8_t *d_stream, uint64_t cap,
                               uint64_t *d_offsets, leco_fix_block_param *d_params, void *stream) {
    return encode_column<uint64_t>(d_in, n, bs, d_stream, cap, d_offsets, d_params, stream);
}

int leco_fix_decode_column_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                               uint32_t *d_out, void *stream) {
    return decode_range<uint32_t>(d_stream, d_offsets, n, bs, 0, n, d_out, stream);
}
int leco_fix_decode_column_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                               uint64_t *d_out, void *stream) {
    return decode_range<uint64_t>(d_stream, d_offsets, n, bs, 0, n, d_out, stream);
}

int leco_fix_decode_range_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                              uint64_t begin, uint64_t end, uint32_t *d_out, void *stream) {
    return decode_range<uint32_t>(d_stream, d_offsets, n, bs, begin, end, d_out, stream);
}
int leco_fix_decode_range_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                              uint64_t begin, uint64_t end, uint64_t *d_out, void *stream) {
    return decode_range<uint64_t>(d_stream, d_offsets, n, bs, begin, end, d_out, stream);
}

int leco_fix_decode_range_ex_u32(const uint8_t *d_stream, const uint64_t *d_offsets,
                                 const leco_fix_block_param *d_params, uint64_t n, uint32_t bs, uint64_t begin,
                                 uint64_t end, int shape, uint32_t *d_out, void *stream) {
    return decode_range_ex<uint32_t>(d_stream, d_offsets, d_params, n, bs, begin, end, shape, d_out, stream);
}
int leco_fix_decode_range_ex_u64(const uint8_t *d_stream, const uint64_t *d_offsets,
                                 const leco_fix_block_param *d_params, uint64_t n, uint32_t bs, uint64_t begin,
                                 uint64_t end, int shape, uint64_t *d_out, void *stream) {
    return decode_range_ex<uint64_t>(d_stream, d_offsets, d_params, n, bs, begin, end, shape, d_out, stream);
}

int leco_fix_lookup_u32(const uint8_t *d_stream, const leco_fix_block_param *d_params, uint64_t n, uint32_t bs,
                        const uint64_t *d_idx, uint64_t nq, uint32_t *d_out, void *stream) {
    return lookup<uint32_t>(d_stream, d_params, n, bs, d_idx, nq, d_out, stream);
}
int leco_fix_lookup_u64(const uint8_t *d_stream, const leco_fix_block_param *d_params, uint64_t n, uint32_t bs,
                        const uint64_t *d_idx, uint64_t nq, uint64_t *d_out, void *stream) {
    return lookup<uint64_t>(d_stream, d_params, n, bs, d_idx, nq, d_out, stream);
}

uint64_t leco_fix_lookup_grouped_workspace(uint64_t nq) { return leco::lookup_grouped_workspace(nq); }

int leco_fix_lookup_grouped_u32(const uint8_t *d_stream, const leco_fix_block_param *d_params, uint64_t n,
                                uint32_t bs, const uint64_t *d_idx, uint64_t nq, uint32_t *d_out, void *d_work,
                                uint64_t work_bytes, void *stream) {
    return lookup_grouped<uint32_t>(d_stream, d_params, n, bs, d_idx, nq, d_out, d_work, work_bytes, stream);
}
int leco_fix_lookup_grouped_u64(const uint8_t *d_stream, const leco_fix_block_param *d_params, uint64_t n,
                                uint32_t bs, const uint64_t *d_idx, uint64_t nq, uint64_t *d_out, void *d_work,
                                uint64_t work_bytes, void *stream) {
    return lookup_grouped<uint64_t>(d_stream, d_params, n, bs, d_idx, nq, d_out, d_work, work_bytes, stream);
}

int leco_fix_build_params(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t bs,
                          uint32_t w, leco_fix_block_param *d_params, void *stream) {
    if (n == 0 || !valid_bs(bs) || (w != 32 && w != 64) || !d_stream || !d_offsets || !d_params)
        return LECO_FIX_EINVAL;
    return hip_status(leco::launch_build_params(d_stream, d_offsets, n, bs, w, d_params, (hipStream_t)stream));
}

int leco_fix_encode_block_u32(const uint32_t *h, uint64_t len, uint8_t *res, uint64_t cap, uint64_t *written) {
    return encode_block_host<uint32_t>(h, len, res, cap, written);
}
int leco_fix_encode_block_u64(const uint64_t *h, uint64_t len, uint8_t *res, uint64_t cap, uint64_t *written) {
    return encode_block_host<uint64_t>(h, len, res, cap, written);
}
int leco_fix_decode_block_u32(const uint8_t *h_in, uint64_t len, uint32_t *h_out) {
    return decode_block_host<uint32_t>(h_in, len, h_out);
}
int leco_fix_decode_block_u64(const uint8_t *h_in, uint64_t len, uint64_t *h_out) {
    return decode_block_host<uint64_t>(h_in, len, h_out);
}
int leco_fix_random_decode_u32(const uint8_t *h_in, uint32_t to_find, uint32_t *v) {
    return random_decode_host<uint32_t>(h_in, to_find, v);
}
int leco_fix_random_decode_u64(const uint8_t *h_in, uint32_t to_find, uint64_t *v) {
    return random_decode_host<uint64_t>(h_in, to_find, v);
}

}  // extern "C"
