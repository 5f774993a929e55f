// gfx950 kernels of the fixed-partition baselines FOR_int<T> and Delta_int<T>
// (headers/FOR_integer_template.h:14-250, delta_integer_template.h:14-160 of
// the reference), on the same column layout as the LeCo codec: the blocks'
// byte strings back to back, block b at offs[b].
//
// Encode: blocks in ticket order -- block statistics (FOR: min and max;
// Delta: the largest |y[i+1] - y[i]|), bit width and size, the decoupled
// look-back for the offset, then the LDS-image packer shared with the LeCo
// encoder; one wave per block and a run of blocks per ticket up to 4096
// values, one workgroup per block above.  Decode: one workgroup per block;
// Delta decodes by a workgroup prefix sum of the signed deltas (mod 2^w, so
// any summation order gives the reference's sequential result).  Lookups: one
// lane per query (FOR) or one wave per query summing the deltas before it
// (Delta).
#include "leco_kernels.h"

namespace leco {

LECO_LOOKBACK_HOOK_SETTER(set_lookback_hook_base)

namespace {

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) { return ~wave_max_u64(~v); }

template <int W>
__device__ __forceinline__ uint64_t wmask() {
    return W == 32 ? 0xFFFFFFFFull : ~0ull;
}

// Signed delta field of y[i+1] - y[i] on l bits (write_delta_int_T, bit_write.h:258-300).
template <int W>
__device__ __forceinline__ uint64_t delta_field(uint64_t a, uint64_t b, int l) {
    const bool up = b > a;
    const uint64_t mag = up ? b - a : a - b;
    return (mag & ((1ull << (l - 1)) - 1)) | ((uint64_t)up << (l - 1));
}

template <int CODEC, typename T, int W>
__global__ __launch_bounds__(kThreads) void k_base_encode(const T *__restrict__ y_all, uint64_t n, uint32_t bs,
                                                          uint64_t *__restrict__ offs, uint8_t *__restrict__ out) {
    __shared__ EncodeShared sh;
    __shared__ uint64_t s_b;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    if (tid == 0) s_b = take_ticket(offs);  // offs[0] is the ticket
    __syncthreads();
    const uint64_t b = s_b;
    const uint64_t s = b * bs;
    const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
    const T *y = y_all + s;

    // statistics
    uint64_t lo = ~0ull, hi = 0;
    for (uint32_t j = tid; j < m; j += kThreads) {
        const uint64_t v = (uint64_t)y[j];
        if (CODEC == kCodecFor) {
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        } else if (j + 1 < m) {
            const uint64_t u = (uint64_t)y[j + 1];
            const uint64_t d = (u > v ? u - v : v - u) & wmask<W>();
            hi = d > hi ? d : hi;
        }
    }
    hi = wave_max_u64(hi);
    if (CODEC == kCodecFor) lo = wave_min_u64(lo);
    if (lane == 0) {
        sh.rmag[wid] = hi;
        sh.rmax[wid] = (int64_t)lo;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kThreads / kWave; ++w) {
        hi = sh.rmag[w] > hi ? sh.rmag[w] : hi;
        if (CODEC == kCodecFor) lo = (uint64_t)sh.rmax[w] < lo ? (uint64_t)sh.rmax[w] : lo;
    }
    int l;
    if (CODEC == kCodecFor) l = hi - lo ? bitlen64(hi - lo) : 0;
    else l = hi ? bitlen64(hi) + 1 : 0;
    if (l > W) l = W;
    const uint64_t size = CODEC == kCodecFor ? for_block_bytes((uint32_t)l, W, m) : delta_block_bytes((uint32_t)l, W, m);

    if (wid == 0) {
        const uint64_t excl = look_back(b, size, offs);
        if (lane == 0) {
            sh.off = excl;
            sh.end = excl + size;
        }
    }
    __syncthreads();

    const bool raw = l == W;
    constexpr uint32_t tb = W / 8;
    PackedBlock pb;
    pb.off = sh.off;
    pb.end = sh.end;
    pb.m = m;
    pb.l = l;
    pb.nchunks = l == 0 ? 1 : (m + kChunk - 1) / kChunk;
    if (raw) {
        pb.hdr = 1;
        pb.t0b = pb.t1b = 0;
    } else if (CODEC == kCodecFor) {
        pb.hdr = 1 + 2 * tb;
        pb.t0b = W == 64 ? lo : (lo | (hi << 32));
        pb.t1b = W == 64 ? hi : 0;
    } else {
        pb.hdr = 1 + tb;
        pb.t0b = (uint64_t)y[0];
        pb.t1b = 0;
    }
    auto sync = [] { __syncthreads(); };
    for (uint32_t c = 0; c < pb.nchunks; ++c)
        pack_chunk<kThreads, kChunk>(pb, c, tid, sh.buf, out, sync, [&](int, uint32_t i) -> uint64_t {
            const uint64_t v = (uint64_t)y[i];
            if (raw) return v;
            if (CODEC == kCodecFor) return v - lo;
            // n fields: the last one, past the reference's vector of n-1 deltas, is zero
            return i + 1 < m ? delta_field<W>(v, (uint64_t)y[i + 1], l) : 0;
        });
}

// Blocks of at most kBaseUnitBlock values are encoded one per wave, a
// workgroup taking base_unit_blocks() consecutive blocks per ticket (about
// 4 * kBaseUnitBlock values, at most kBaseUnitMaxBlocks blocks): one ticket
// per block would cap the encoder at the rate of same-address atomics (about
// one per 12 ns on MI355X, 170 GB/s of uint64 input at bs 256).
constexpr uint32_t kBaseUnitBlock = 4096;
constexpr uint32_t kBaseUnitMaxBlocks = 256;
__host__ __device__ inline uint32_t base_unit_blocks(uint32_t bs) {
    const uint32_t per = 4 * (kBaseUnitBlock / bs);
    return per < kBaseUnitMaxBlocks ? per : kBaseUnitMaxBlocks;
}

struct BaseUnitShared {
    alignas(16) uint32_t buf[kLdsBytes / 4];  // one LDS image of kWaveLdsBytes per wave
    uint64_t at[kBaseUnitMaxBlocks];          // block size, then its stream offset
    uint64_t lo[kBaseUnitMaxBlocks];          // FOR: the block's minimum; Delta: its first value
    uint64_t hi[kBaseUnitMaxBlocks];          // FOR: the block's maximum
    int l[kBaseUnitMaxBlocks];
};

// The same blocks as k_base_encode, `per` consecutive blocks per workgroup
// (bs <= kBaseUnitBlock), one wave per block.  Phase A: each wave's blocks'
// statistics, bit widths and sizes.  Then one look-back for the unit over the
// units' status words offs[(u+1)*per] (StridedSlot); the unit's inner block
// ends go straight to offs[b+1].  Phase B: each wave packs its blocks,
// kWaveChunk values at a time through its LDS image.
template <int CODEC, typename T, int W>
__global__ __launch_bounds__(kThreads) void k_base_encode_units(const T *__restrict__ y_all, uint64_t n, uint32_t bs,
                                                                uint32_t per, uint64_t *__restrict__ offs,
                                                                uint8_t *__restrict__ out) {
    __shared__ BaseUnitShared sh;
    __shared__ uint64_t s_u;
    constexpr uint32_t kWaves = kThreads / kWave;
    constexpr uint32_t tb = W / 8;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    if (tid == 0) s_u = take_ticket(offs);  // offs[0] is the ticket
    __syncthreads();
    const uint64_t u = s_u;
    const uint64_t nb = (n + bs - 1) / bs;
    const uint64_t b0 = u * per;
    if (b0 >= nb) return;
    const uint32_t count = (uint32_t)(nb - b0 < per ? nb - b0 : per);

    // phase A
    for (uint32_t k = wid; k < count; k += kWaves) {
        const uint64_t s = (b0 + k) * bs;
        const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
        const T *y = y_all + s;
        uint64_t lo = ~0ull, hi = 0;
        for (uint32_t j0 = 0; j0 < m; j0 += kWaveChunk) {
            uint64_t v[kWaveVPL], w[kWaveVPL];
#pragma unroll
            for (int q = 0; q < kWaveVPL; ++q) {
                const uint32_t j = j0 + lane + q * kWave;
                v[q] = j < m ? (uint64_t)y[j] : 0;
                if (CODEC == kCodecDelta) w[q] = j + 1 < m ? (uint64_t)y[j + 1] : 0;
            }
#pragma unroll
            for (int q = 0; q < kWaveVPL; ++q) {
                const uint32_t j = j0 + lane + q * kWave;
                if (CODEC == kCodecFor) {
                    lo = j < m && v[q] < lo ? v[q] : lo;
                    hi = j < m && v[q] > hi ? v[q] : hi;
                } else {
                    const uint64_t d = (w[q] > v[q] ? w[q] - v[q] : v[q] - w[q]) & wmask<W>();
                    hi = j + 1 < m && d > hi ? d : hi;
                }
            }
        }
        hi = wave_max_u64(hi);
        if (CODEC == kCodecFor) lo = wave_min_u64(lo);
        int l;
        if (CODEC == kCodecFor) l = hi - lo ? bitlen64(hi - lo) : 0;
        else l = hi ? bitlen64(hi) + 1 : 0;
        if (l > W) l = W;
        if (lane == 0) {
            sh.at[k] = CODEC == kCodecFor ? for_block_bytes((uint32_t)l, W, m) : delta_block_bytes((uint32_t)l, W, m);
            sh.lo[k] = CODEC == kCodecFor ? lo : (uint64_t)y[0];
            sh.hi[k] = hi;
            sh.l[k] = l;
        }
    }
    __syncthreads();

    // the unit's offset (four blocks per lane), its blocks' offsets and ends
    if (wid == 0) {
        uint64_t sz[4], run = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t k = 4 * lane + r;
            sz[r] = k < count ? sh.at[k] : 0;
            run += sz[r];
        }
        const uint64_t inc = wave_inclusive_sum(run);
        const uint64_t excl = look_back(u, wave_last(inc), offs, StridedSlot{per, nb});
        uint64_t o = excl + inc - run;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t k = 4 * lane + r;
            if (k < count) {
                sh.at[k] = o;
                o += sz[r];
                if (k + 1 < count) offs[b0 + k + 1] = o;  // the last block's end is the unit's status word
            }
        }
    }
    __syncthreads();

    // phase B
    uint32_t *buf = sh.buf + wid * (kWaveLdsBytes / 4);
    for (uint32_t k = wid; k < count; k += kWaves) {
        const uint64_t s = (b0 + k) * bs;
        const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
        const T *y = y_all + s;
        const int l = __builtin_amdgcn_readfirstlane(sh.l[k]);
        const uint64_t lo = sh.lo[k], hi = sh.hi[k], off = sh.at[k];
        const bool raw = l == W;
        PackedBlock pb;
        pb.off = off;
        pb.end = off + (CODEC == kCodecFor ? for_block_bytes((uint32_t)l, W, m) : delta_block_bytes((uint32_t)l, W, m));
        pb.m = m;
        pb.l = l;
        pb.nchunks = l == 0 ? 1 : (m + kWaveChunk - 1) / kWaveChunk;
        if (raw) {
            pb.hdr = 1;
            pb.t0b = pb.t1b = 0;
        } else if (CODEC == kCodecFor) {
            pb.hdr = 1 + 2 * tb;
            pb.t0b = W == 64 ? lo : (lo | (hi << 32));
            pb.t1b = W == 64 ? hi : 0;
        } else {
            pb.hdr = 1 + tb;
            pb.t0b = lo;
            pb.t1b = 0;
        }
        auto sync = [] { wave_lds_sync(); };
        for (uint32_t c = 0; c < pb.nchunks; ++c) {
            uint64_t f[kWaveVPL];
#pragma unroll
            for (int q = 0; q < kWaveVPL; ++q) {
                const uint32_t i = c * kWaveChunk + lane + q * kWave;
                const uint64_t v = i < m ? (uint64_t)y[i] : 0;
                if (raw || l == 0) f[q] = v;
                else if (CODEC == kCodecFor) f[q] = v - lo;
                // n fields: the last one, past the reference's vector of n-1 deltas, is zero
                else f[q] = i + 1 < m ? delta_field<W>(v, (uint64_t)y[i + 1], l) : 0;
            }
            pack_chunk<kWave, kWaveChunk>(pb, c, lane, buf, out, sync,
                                                [&](int q, uint32_t) -> uint64_t { return f[q]; });
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_base_clear(uint64_t *__restrict__ offs, uint64_t words) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < words) offs[i] = 0;
}

// Block end offsets from the look-back words (flags cleared), the ticket back
// to 0; offs[nb] is poisoned if a look-back wait gave up.
__global__ __launch_bounds__(kThreads) void k_base_fixup(uint64_t *__restrict__ offs, uint64_t nb) {
    const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b >= nb) return;
    uint64_t end = offs[b + 1] & kStatusValue;
    if (b + 1 == nb) {
        if (offs[0] & kTicketError) end = LECO_FIX_OFFSETS_POISON;
        offs[0] = 0;
    }
    offs[b + 1] = end;
}

// FOR values [j0, j0 + jn) of a run staged in LDS, field j at bit
// lbit + (j - j0) * lw (lw = 0: every value is base; raw blocks: base 0,
// lw = W); streaming stores.
template <typename T, int W>
__device__ __forceinline__ void for_run(const uint64_t *__restrict__ buf, uint32_t lbit, int lw, uint64_t base,
                                        uint32_t j0, uint32_t jn, T *__restrict__ o) {
    const int tid = threadIdx.x;
    auto one = [&](uint32_t j) {
        const uint64_t f = lds_field<W>(buf, lbit + (j - j0) * (uint32_t)lw, lw);
        __builtin_nontemporal_store((T)(base + f), &o[j]);
    };
    if (jn == kChunk) {
#pragma unroll
        for (int kk = 0; kk < (int)(kChunk / kThreads); ++kk) one(j0 + tid + kk * kThreads);
    } else {
        for (uint32_t j = j0 + tid; j < j0 + jn; j += kThreads) one(j);
    }
}

// Bytes [g0, g1) of the stream into buf from 16-byte aligned address a0 =
// g0 & ~15; returns the LDS bit of byte g0.  Callers synchronize.
__device__ __forceinline__ uint32_t stage_bytes(const uint8_t *__restrict__ st, uint64_t g0, uint64_t g1,
                                                uint64_t *__restrict__ buf) {
    const uint64_t a0 = g0 & ~15ull;
    const uint32_t n16 = (uint32_t)((g1 - a0 + 15) / 16);
    for (uint32_t u = threadIdx.x; u < n16; u += kThreads)
        *(uint4 *)(buf + 2 * u) = *(const uint4 *)(st + a0 + 16ull * u);
    return (uint32_t)(8 * (g0 - a0));
}

// FOR_int<T>::decodeArray8 (:79-108) / Delta_int<T>::decodeArray8 (:85-126),
// one workgroup per block, fields staged through LDS.
// FOR: a block whose bytes fit the buffer is staged whole, header included
// (no dependent header read); a larger one stages kChunk values at a time.
// Delta: kChunk deltas at a time; each thread sums 8 consecutive deltas, a
// workgroup scan places the sums, and the values go back through the buffer
// (index i at i + i/8, which spreads a wave's 8-apart writes over the banks)
// so that the stores are lane-contiguous.
template <int CODEC, typename T, int W>
__global__ __launch_bounds__(kThreads) void k_base_decode(const uint8_t *__restrict__ st,
                                                          const uint64_t *__restrict__ offs, uint64_t n, uint32_t bs,
                                                          T *__restrict__ out) {
    constexpr uint32_t K = 8;                             // Delta: consecutive deltas per thread per step
    constexpr uint32_t kSlots = kChunk + kChunk / K;      // Delta: padded value slots of one step
    constexpr uint32_t kWords = (kLdsBytes / 8 > kSlots ? kLdsBytes / 8 : kSlots);
    __shared__ __align__(16) uint64_t buf[kWords];
    __shared__ uint64_t wsum[kThreads / kWave];
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t b = blockIdx.x;
    const uint64_t s = b * bs;
    const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
    const uint64_t off = offs[b], end = offs[b + 1];
    T *o = out + s;
    constexpr uint32_t tb = W / 8;
    if (CODEC == kCodecFor) {
        const uint64_t b0 = off & ~15ull;
        const uint32_t nb16 = (uint32_t)((end - b0 + 15) / 16);
        if (end >= off && (uint64_t)nb16 * 16 + 16 <= sizeof(buf)) {
            const uint32_t hbit = stage_bytes(st, off, end, buf);
            __syncthreads();
            const uint32_t hb = hbit / 8;
            const int l = ((const uint8_t *)buf)[hb];
            const bool raw = l >= W;
            const int lw = raw ? W : l;
            const uint64_t base = raw ? 0 : lds_u64_at(buf, hb + 1) & wmask<W>();
            const uint32_t lbit = 8 * (hb + (raw ? 1 : 1 + 2 * tb));
            for (uint32_t j0 = 0; j0 < m; j0 += kChunk)
                for_run<T, W>(buf, lbit + j0 * (uint32_t)lw, lw, base, j0, m - j0 < kChunk ? m - j0 : kChunk, o);
            return;
        }
        const int l = st[off];
        const bool raw = l >= W;
        const int lw = raw ? W : l;
        const uint64_t base = raw ? 0 : load_u64_any(st + off + 1) & wmask<W>();
        if (lw == 0) {
            for (uint32_t j = tid; j < m; j += kThreads) __builtin_nontemporal_store((T)base, &o[j]);
            return;
        }
        const uint64_t pay = off + (raw ? 1 : 1 + 2 * tb);
        for (uint32_t j0 = 0; j0 < m; j0 += kChunk) {
            const uint32_t jn = m - j0 < kChunk ? m - j0 : kChunk;
            const uint64_t g0 = pay + (uint64_t)j0 / 8 * lw;
            const uint32_t lbit = stage_bytes(st, g0, g0 + ((uint64_t)jn * lw + 7) / 8, buf);
            __syncthreads();
            for_run<T, W>(buf, lbit, lw, base, j0, jn, o);
            __syncthreads();
        }
        return;
    }
    const int l = st[off];
    if (l == 127 || l == 126) {  // :98-113, one- and two-value raw blocks
        const uint32_t cnt = l == 127 ? 1 : 2;
        if ((uint32_t)tid < cnt && (uint32_t)tid < m) o[tid] = (T)load_bits(st + off + 1, (uint64_t)tid * W, W);
        return;
    }
    if (l >= W) {
        for (uint32_t j0 = 0; j0 < m; j0 += kChunk) {
            const uint32_t jn = m - j0 < kChunk ? m - j0 : kChunk;
            const uint64_t g0 = off + 1 + (uint64_t)j0 * tb;
            const uint32_t lbit = stage_bytes(st, g0, g0 + (uint64_t)jn * tb, buf);
            __syncthreads();
            for_run<T, W>(buf, lbit, W, 0, j0, jn, o);
            __syncthreads();
        }
        return;
    }
    // Delta: o[0] = y_0, o[i] = y_0 + (signed deltas 0 .. i-1)
    const uint64_t y0 = load_u64_any(st + off + 1) & wmask<W>();
    if (tid == 0) o[0] = (T)y0;
    if (!l) {
        for (uint32_t j = 1 + tid; j < m; j += kThreads) __builtin_nontemporal_store((T)y0, &o[j]);
        return;
    }
    const uint64_t pay = off + 1 + tb;
    uint64_t carry = y0;
    for (uint32_t c0 = 0; c0 + 1 < m; c0 += kChunk) {
        const uint32_t cn = m - 1 - c0 < kChunk ? m - 1 - c0 : kChunk;  // deltas c0 .. c0+cn-1 (values c0+1 ..)
        const uint64_t g0 = pay + (uint64_t)c0 / 8 * l;
        const uint32_t lbit = stage_bytes(st, g0, g0 + ((uint64_t)cn * l + 7) / 8, buf);
        __syncthreads();
        const uint32_t d0 = tid * K;
        uint64_t d[K], run = 0;
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            d[k] = 0;
            if (d0 + k < cn) {
                const uint64_t f = lds_field<W>(buf, lbit + (d0 + k) * (uint32_t)l, l);
                const uint64_t mag = f & ((1ull << (l - 1)) - 1);
                d[k] = ((f >> (l - 1)) & 1) ? mag : 0 - mag;
            }
            run += d[k];
        }
        // exclusive prefix of the threads' sums (wrapping uint64)
        uint64_t inc = run;
        for (int sft = 1; sft < kWave; sft <<= 1) {
            const uint64_t t = (uint64_t)__shfl_up((long long)inc, sft, kWave);
            if (lane >= sft) inc += t;
        }
        if (lane == kWave - 1) wsum[wid] = inc;
        __syncthreads();  // also: every field of this step is in registers
        uint64_t before = 0, all = 0;
#pragma unroll
        for (int w = 0; w < kThreads / kWave; ++w) {
            before += w < wid ? wsum[w] : 0;
            all += wsum[w];
        }
        uint64_t v = carry + before + inc - run;
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            v += d[k];
            const uint32_t i = d0 + k;
            buf[i + i / K] = v;
        }
        __syncthreads();
        for (uint32_t i = tid; i < cn; i += kThreads) __builtin_nontemporal_store((T)buf[i + i / K], &o[c0 + 1 + i]);
        carry += all;
        __syncthreads();  // the buffer and wsum are reused by the next step
    }
}

// Delta_int<T>::decodeArray8 for blocks below kStagedMinBlock values: tiles
// of kb consecutive blocks, their bytes staged into LDS whole (as many blocks
// at a time as fit kMultiLds), one wave per block.  A wave takes 16 bytes of
// consecutive values per lane per step (4 uint32 or 2 uint64): the lane sums
// its deltas in order, and one inclusive wave scan of the lane sums plus the
// running value places them.  Against one delta per lane and a scan per 64
// values, this is 1.4x faster at block sizes 256-1000.  The scan is DPP row
// shifts and broadcasts (VALU), not LDS shuffles.  The step's values then
// cross lanes (wave shuffles) so that each store instruction writes 8 bytes
// per lane (4 for a uint32 block that starts at an odd word), contiguous
// across the wave: a 16-byte store of the lane's own values can read its
// upper half from the VGPRs only after the next LDS field read has refilled
// them.
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_delta_decode_multi(const uint8_t *__restrict__ st,
                                                                 const uint64_t *__restrict__ offs, uint64_t n,
                                                                 uint32_t bs, uint32_t kb, T *__restrict__ out) {
    __shared__ MultiShared sm;
    constexpr uint32_t tb = W / 8;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t b0 = (uint64_t)blockIdx.x * kb;
    const uint64_t nb = (n + bs - 1) / bs;
    if (b0 + kb > nb) kb = (uint32_t)(nb - b0);
    for (uint32_t k = tid; k <= kb; k += kThreads) sm.offs[k] = offs[b0 + k];
    __syncthreads();
    const uint8_t *bytes = (const uint8_t *)sm.buf;
    for (uint32_t ks = 0; ks < kb;) {
        const uint64_t a0 = sm.offs[ks] & ~15ull;
        // largest ke in (ks, kb] whose bytes [a0, offs[ke]) fit (one block always does)
        uint32_t ke = kb;
        if (sm.offs[kb] - a0 > kMultiLds) {
            uint32_t lo = ks + 1, hi = kb;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) / 2;
                if (sm.offs[mid] - a0 <= kMultiLds) lo = mid;
                else hi = mid - 1;
            }
            ke = lo;
        }
        const uint32_t n16 = (uint32_t)((sm.offs[ke] - a0 + 15) / 16);
        for (uint32_t u = tid; u < n16; u += kThreads)
            *(uint4 *)(sm.buf + 2 * u) = *(const uint4 *)(st + a0 + 16ull * u);
        __syncthreads();
        for (uint32_t k = ks + wid; k < ke; k += kThreads / kWave) {
            const uint64_t s = (b0 + k) * bs;
            const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
            const uint32_t hb = (uint32_t)(sm.offs[k] - a0);
            const uint32_t l = bytes[hb];
            T *o = out + s;
            if (l == 127 || l == 126) {  // :98-113, one- and two-value raw blocks
                if ((uint32_t)lane < (l == 127 ? 1u : 2u) && (uint32_t)lane < m)
                    o[lane] = (T)(lds_u64_at(sm.buf, hb + 1 + lane * tb) & wmask<W>());
                continue;
            }
            if (l >= (uint32_t)W) {
                for (uint32_t j = lane; j < m; j += kWave)
                    __builtin_nontemporal_store((T)lds_field<W>(sm.buf, 8 * (hb + 1) + j * W, W), &o[j]);
                continue;
            }
            const uint64_t y0 = lds_u64_at(sm.buf, hb + 1) & wmask<W>();
            if (l == 0) {
                for (uint32_t j = lane; j < m; j += kWave) __builtin_nontemporal_store((T)y0, &o[j]);
                continue;
            }
            const uint32_t pbit = 8 * (hb + 1 + tb);
            const T low = (T)((1ull << (l - 1)) - 1);
            // Value p is y0 plus deltas 0 .. p-1 (delta d is field d), sums
            // modulo 2^W in T.  Each lane takes R consecutive values p0 ..
            // p0+R-1 of a step of R*kWave (R*sizeof(T) = 16 bytes): its deltas
            // summed in order, then one wave scan of the lane totals.
            T carry = (T)y0;
            constexpr uint32_t R = 16 / sizeof(T);
            const bool pairs = ((uintptr_t)o & 7) == 0;  // uint32: 8-byte stores of value pairs
            for (uint32_t i0 = 0; i0 < m; i0 += R * kWave) {
                const uint32_t p0 = i0 + R * lane;
                T v[R];
                T run = 0;
#pragma unroll
                for (uint32_t k = 0; k < R; ++k) {
                    const uint32_t p = p0 + k;
                    if (p >= 1 && p < m) {
                        const T f = (T)lds_field<W>(sm.buf, pbit + (p - 1) * l, (int)l);
                        run += ((f >> (l - 1)) & 1) ? (f & low) : (T)(0 - (f & low));
                    }
                    v[k] = run;
                }
                const T tot = wave_inclusive_sum(run);
                const T base = carry + (tot - run);
#pragma unroll
                for (uint32_t k = 0; k < R; ++k) v[k] += base;
                const bool odd = lane & 1;
                if constexpr (W == 64) {
                    // store h writes values i0 + 64h + lane: value 2s + c of the
                    // step sits in lane s, v[c]
#pragma unroll
                    for (uint32_t h = 0; h < 2; ++h) {
                        const int src = (int)(32 * h) + (lane >> 1);
                        const T a = (T)__shfl((long long)v[0], src, kWave);
                        const T b = (T)__shfl((long long)v[1], src, kWave);
                        const uint32_t p = i0 + 64 * h + lane;
                        if (p < m) __builtin_nontemporal_store(odd ? b : a, &o[p]);
                    }
                } else if (pairs) {
                    // store h writes values i0 + 128h + 2 lane and the next:
                    // value 4s + c of the step sits in lane s, v[c]
                    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
#pragma unroll
                    for (uint32_t h = 0; h < 2; ++h) {
                        const int src = (int)(32 * h) + (lane >> 1);
                        const uint32_t a0 = (uint32_t)__shfl((int)v[0], src, kWave);
                        const uint32_t a1 = (uint32_t)__shfl((int)v[1], src, kWave);
                        const uint32_t a2 = (uint32_t)__shfl((int)v[2], src, kWave);
                        const uint32_t a3 = (uint32_t)__shfl((int)v[3], src, kWave);
                        const uint32_t x = odd ? a2 : a0, y = odd ? a3 : a1;
                        const uint32_t q = i0 + 128 * h + 2 * lane;
                        if (q + 1 < m) __builtin_nontemporal_store(u32x2{x, y}, (u32x2 *)(o + q));
                        else if (q < m) __builtin_nontemporal_store((T)x, &o[q]);
                    }
                } else {
                    // uint32 output at an odd word: store h writes values
                    // i0 + 64h + lane, value 4s + c sitting in lane s, v[c]
#pragma unroll
                    for (uint32_t h = 0; h < 4; ++h) {
                        const int src = (int)(16 * h) + (lane >> 2);
                        const uint32_t a0 = (uint32_t)__shfl((int)v[0], src, kWave);
                        const uint32_t a1 = (uint32_t)__shfl((int)v[1], src, kWave);
                        const uint32_t a2 = (uint32_t)__shfl((int)v[2], src, kWave);
                        const uint32_t a3 = (uint32_t)__shfl((int)v[3], src, kWave);
                        const uint32_t c = lane & 3;
                        const uint32_t x = c == 0 ? a0 : c == 1 ? a1 : c == 2 ? a2 : a3;
                        const uint32_t p = i0 + 64 * h + lane;
                        if (p < m) __builtin_nontemporal_store((T)x, &o[p]);
                    }
                }
                carry += wave_last(tot);
            }
        }
        __syncthreads();
        ks = ke;
    }
}

// FOR_int<T>::randomdecodeArray8 (:219-246): one lane per query.
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_for_lookup(const uint8_t *__restrict__ st,
                                                         const uint64_t *__restrict__ offs, uint64_t n, uint32_t bs,
                                                         const uint64_t *__restrict__ idx, uint64_t nq,
                                                         T *__restrict__ out) {
    const uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= nq) return;
    const uint64_t i = idx[q];
    if (i >= n) {
        out[q] = 0;
        return;
    }
    const uint64_t b = i / bs;
    const uint32_t j = (uint32_t)(i - b * bs);
    const uint8_t *blk = st + offs[b];
    const int l = blk[0];
    if (l == W) {
        out[q] = (T)load_bits(blk + 1, (uint64_t)j * W, W);
        return;
    }
    const uint64_t base = load_u64_any(blk + 1) & wmask<W>();
    out[q] = (T)(l ? base + load_bits(blk + 1 + 2 * (W / 8), (uint64_t)j * l, l) : base);
}

// Delta_int<T>::randomdecodeArray8 (:131-153) with read_Delta_int: one wave per
// query sums the signed deltas before position j.
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_delta_lookup(const uint8_t *__restrict__ st,
                                                           const uint64_t *__restrict__ offs, uint64_t n,
                                                           uint32_t bs, const uint64_t *__restrict__ idx, uint64_t nq,
                                                           T *__restrict__ out) {
    const int lane = threadIdx.x % kWave;
    const uint64_t q = (uint64_t)blockIdx.x * (kThreads / kWave) + threadIdx.x / kWave;
    if (q >= nq) return;
    const uint64_t i = idx[q];
    if (i >= n) {
        if (lane == 0) out[q] = 0;
        return;
    }
    const uint64_t b = i / bs;
    const uint32_t j = (uint32_t)(i - b * bs);
    const uint8_t *blk = st + offs[b];
    const int l = blk[0];
    if (l == W) {
        if (lane == 0) out[q] = (T)load_bits(blk + 1, (uint64_t)j * W, W);
        return;
    }
    const uint64_t base = load_u64_any(blk + 1) & wmask<W>();
    uint64_t sum = 0;
    if (l)
        for (uint32_t k = lane; k < j; k += kWave) {
            const uint64_t f = load_bits(blk + 1 + W / 8, (uint64_t)k * l, l);
            const uint64_t mag = f & ((1ull << (l - 1)) - 1);
            sum += ((f >> (l - 1)) & 1) ? mag : 0 - mag;
        }
    for (int o = kWave / 2; o > 0; o >>= 1) sum += (uint64_t)__shfl_xor((long long)sum, o, kWave);
    if (lane == 0) out[q] = (T)(base + sum);
}

}  // namespace

template <typename T>
hipError_t launch_base_encode(int codec, const T *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                              hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const uint64_t nb = (n + bs - 1) / bs;
    k_base_clear<<<dim3((unsigned)((nb + 1 + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(offs, nb + 1);
    if (bs <= kBaseUnitBlock) {
        const uint32_t per = base_unit_blocks(bs);
        const unsigned units = (unsigned)((nb + per - 1) / per);
        if (codec == kCodecFor)
            k_base_encode_units<kCodecFor, T, W><<<dim3(units), dim3(kThreads), 0, st>>>(in, n, bs, per, offs, stream);
        else
            k_base_encode_units<kCodecDelta, T, W><<<dim3(units), dim3(kThreads), 0, st>>>(in, n, bs, per, offs,
                                                                                           stream);
    } else if (codec == kCodecFor)
        k_base_encode<kCodecFor, T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(in, n, bs, offs, stream);
    else
        k_base_encode<kCodecDelta, T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(in, n, bs, offs, stream);
    k_base_fixup<<<dim3((unsigned)((nb + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(offs, nb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_base_decode(int codec, const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                              T *out, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const uint64_t nb = (n + bs - 1) / bs;
    if (codec == kCodecFor && !staged_decode(bs)) return launch_for_decode_multi<T>(stream, offs, n, bs, out, st);
    if (!staged_decode(bs)) {
        const uint32_t kb = multi_blocks(bs, W);
        k_delta_decode_multi<T, W><<<dim3((unsigned)((nb + kb - 1) / kb)), dim3(kThreads), 0, st>>>(stream, offs, n,
                                                                                                     bs, kb, out);
        return hipGetLastError();
    }
    if (codec == kCodecFor)
        k_base_decode<kCodecFor, T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(stream, offs, n, bs, out);
    else
        k_base_decode<kCodecDelta, T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(stream, offs, n, bs, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_base_lookup(int codec, const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                              const uint64_t *idx, uint64_t nq, T *out, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    if (codec == kCodecFor)
        k_for_lookup<T, W><<<dim3((unsigned)((nq + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(
            stream, offs, n, bs, idx, nq, out);
    else
        k_delta_lookup<T, W><<<dim3((unsigned)((nq + kThreads / kWave - 1) / (kThreads / kWave))), dim3(kThreads), 0,
                               st>>>(stream, offs, n, bs, idx, nq, out);
    return hipGetLastError();
}

#define LECO_BASE_INST(T)                                                                                          \
    template hipError_t launch_base_encode<T>(int, const T *, uint64_t, uint32_t, uint8_t *, uint64_t *,           \
                                              hipStream_t);                                                        \
    template hipError_t launch_base_decode<T>(int, const uint8_t *, const uint64_t *, uint64_t, uint32_t, T *,     \
                                              hipStream_t);                                                        \
    template hipError_t launch_base_lookup<T>(int, const uint8_t *, const uint64_t *, uint64_t, uint32_t,          \
                                              const uint64_t *, uint64_t, T *, hipStream_t);
LECO_BASE_INST(uint32_t)
LECO_BASE_INST(uint64_t)
#undef LECO_BASE_INST

}  // namespace leco
