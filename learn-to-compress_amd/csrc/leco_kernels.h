// Device-side constants, shared-memory layouts and block addressing used by
// the kernels in leco_kernels.hip.
#pragma once

#include <type_traits>

#include "leco_device.h"
#include "leco_internal.h"

namespace leco {

constexpr int kThreads = 256;                       // workgroup of the per-block kernels (4 waves)
constexpr uint32_t kChunk = 2048;                   // values staged through LDS at a time
constexpr int kLdsBytes = kChunk * 8 + 64;          // one chunk of raw uint64 + alignment slack
constexpr uint32_t kValTile = kValueTile;           // positions per workgroup in value-parallel decode
constexpr int kFitRows = 64;                        // blocks per fit workgroup (one lane of wave 0 each)
constexpr int kFitTK = 128;                         // values per block row in one fit tile
constexpr int kFitThreads = 256;                    // fit workgroup: four waves load, wave 0 also fits
constexpr int kFitPer = kFitRows * kFitTK / kFitThreads;  // tile values loaded per thread
constexpr int kWaveVPL = 8;                         // values per lane register-cached by a wave-encoded block
constexpr uint32_t kWaveChunk = kWaveVPL * kWave;   // values a wave packs through LDS at a time
constexpr int kWaveLdsBytes = kLdsBytes / (kThreads / kWave);  // one wave's share of the LDS image
static_assert(kMultiLds + 16 <= kLdsBytes, "staged small-block bytes plus one word of field over-read fit the LDS buffer");
constexpr int kMultiUnroll = 4;                     // positions per thread in flight in small-block decode
constexpr uint32_t kWaveDecodeMin = 128;            // small blocks of at least this many values: one wave per block
constexpr int kLookPer = 4;                         // predecessors read per lane per look-back step
constexpr int kLookWin = kWave * kLookPer;          // predecessors per look-back step (wave 0)

// Look-back status word (offs[b+1] while the encoder runs; the fix-up clears
// the flags and leaves the block's end offset): 2 flag bits + 62-bit count.  The word carries its own
// value, so relaxed agent-scope atomics (which bypass the non-coherent L1)
// are enough; nothing else is published through it.
constexpr uint64_t kStatusAggregate = 1ull << 62;   // value = this block's size
constexpr uint64_t kStatusInclusive = 2ull << 62;   // value = bytes up to and including this block
constexpr uint64_t kStatusMask = 3ull << 62;
constexpr uint64_t kStatusValue = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;              // look-back wait bound (rounds of s_sleep 1 + reload)
// Bit 63 of the ticket word offs[0]: a look-back wait of this launch reached
// its bound (a broken invariant: the unit it waited for never published), so
// the offsets it produced are wrong.  The fix-up turns it into
// offs[nb] = LECO_FIX_OFFSETS_POISON; tickets mask it off.
constexpr uint64_t kTicketError = 1ull << 63;

// Test hook of the look-back wait, one copy per translation unit (set through
// the setter that LECO_LOOKBACK_HOOK_SETTER defines there): the wait bound,
// and a unit that never publishes its status, so that a test can drive the
// bound's error path with every kernel still exiting.  The defaults leave the
// encoders as they are.
struct LookbackHook {
    uint32_t spin_limit;
    uint32_t reserved;
    uint64_t withhold;  // unit index, or ~0
};
static __device__ LookbackHook g_lookback_hook = {kSpinLimit, 0, ~0ull};
#define LECO_LOOKBACK_HOOK_SETTER(fn)                                                  \
    hipError_t fn(uint32_t spin_limit, uint64_t withhold) {                             \
        const LookbackHook h{spin_limit ? spin_limit : kSpinLimit, 0, withhold};        \
        return hipMemcpyToSymbol(HIP_SYMBOL(g_lookback_hook), &h, sizeof(h));            \
    }

__device__ __forceinline__ uint64_t take_ticket(uint64_t *offs) {
    return atomicAdd((unsigned long long *)offs, 1ull) & ~kTicketError;
}

__device__ __forceinline__ void store_status(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t load_status(const uint64_t *p) {
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kEncodeMaxWaves = kThreads / kWave;  // reduction slots of the encode workgroups
constexpr uint32_t kLargeEncodeBlock = 4096;          // larger blocks: k_encode_large (one chunk cached)

struct EncodeShared {
    alignas(16) uint32_t buf[kLdsBytes / 4];
    int64_t rmax[kEncodeMaxWaves], rmin[kEncodeMaxWaves];
    uint64_t rmag[kEncodeMaxWaves];
    // block-wide residual max / min and magnitude max, reduced by LDS atomics
    // (set by init_reductions() before the kernel's first barrier)
    int64_t amax, amin;
    uint64_t amag;
    // workgroup look-back: per wave, the sum of its window part and flags
    // (bit 0: an inclusive prefix was found, bit 1: a wait gave up)
    uint64_t lbs[kEncodeMaxWaves];
    uint32_t lbf[kEncodeMaxWaves];
    __device__ void init_reductions() {
        amax = INT64_MIN;
        amin = INT64_MAX;
        amag = 0;
    }
    uint64_t off, end;
    double t0;
    int l;
    // wave-encoded unit: per block thetas, width and byte size (then offset)
    double ut0[kUnitMaxBlocks], ut1[kUnitMaxBlocks];
    uint64_t usz[kUnitMaxBlocks];
    int ul[kUnitMaxBlocks];
};

// Where unit p's look-back status word lives.  NextSlot: offs[p+1], one word
// per unit.  StridedSlot: offs[(p+1)*stride], the last unit's at offs[last]
// -- the units are runs of `stride` blocks, and the slots in between hold
// the end offsets of the blocks inside a unit, written once its prefix is
// known and never read as statuses.
struct NextSlot {
    __device__ uint64_t operator()(int64_t p) const { return (uint64_t)(p + 1); }
};
struct StridedSlot {
    uint64_t stride, last;
    __device__ uint64_t operator()(int64_t p) const {
        const uint64_t s = (uint64_t)(p + 1) * stride;
        return s < last ? s : last;
    }
};

// One wave's part of a look-back window: PER chunks of 64 predecessors'
// status words (lane l of chunk k holds unit wtop - 64k - l; units below 0
// read as an inclusive prefix of 0), nearest first.  The part that counts
// ends at the nearest inclusive prefix (or covers all PER chunks); every
// word in it must be ready.  Words that are not are loaded again, all of
// them in flight together, so that each wait costs one round trip however
// many chunks it spans; the nearest inclusive prefix is looked for afresh
// after each round.  Returns the sum of the counted words' values (every
// lane); `found`: an inclusive prefix ends the part; `gave_up`: the wait
// bound was reached with words of the part still not ready (they count 0).
template <int PER, class Slot = NextSlot>
__device__ __forceinline__ uint64_t resolve_window(uint64_t (&stw)[PER], int64_t wtop,
                                                   const uint64_t *__restrict__ offs, uint32_t spin_limit,
                                                   bool &found, bool &gave_up, Slot slot = Slot()) {
    const int lane = threadIdx.x % kWave;
    int ks = PER;               // chunk of the nearest inclusive prefix
    uint64_t stop_mask = ~0ull; // lanes of chunk ks up to and including it
    bool ready = true;
    for (uint32_t round = 0;; ++round) {
        ks = PER;
        stop_mask = ~0ull;
        ready = true;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (k < ks) {
                const uint64_t incl = __ballot((stw[k] & kStatusMask) == kStatusInclusive);
                const uint64_t pending = __ballot((stw[k] & kStatusMask) == 0);
                if (incl) {
                    const int stop = __ffsll((long long)incl) - 1;
                    stop_mask = stop == kWave - 1 ? ~0ull : ((2ull << stop) - 1);
                    ks = k;
                    if (pending & stop_mask) ready = false;
                } else if (pending) {
                    ready = false;
                }
            }
        }
        if (ready || round + 1 >= spin_limit) break;
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int64_t p = wtop - (k * kWave + lane);
            if (k <= ks && (stw[k] & kStatusMask) == 0) stw[k] = load_status(&offs[slot(p)]);
        }
    }
    found = ks < PER;
    gave_up = !ready;
    uint64_t part = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const bool counted = k < ks || (k == ks && ((stop_mask >> lane) & 1));
        part += counted ? stw[k] & kStatusValue : 0;
    }
    return wave_sum_u64(part);
}

// Decoupled look-back of unit b (one wave; every lane returns the same
// exclusive prefix) over the status words offs[p+1] of the units before it
// (offs[0] is the ticket that orders the units).  Publishes the unit's size,
// then loads kWave*PER predecessors' words per step (PER per lane, all in
// flight together) and resolves them (resolve_window) up to the nearest one
// that carries an inclusive prefix; the prefix is its value plus the sizes
// of the units in between; finally publishes the unit's inclusive prefix.
// Every lower unit belongs to a wave that is already running and publishes
// its size before it waits, so the waits end; the spin bound (~seconds) only
// keeps a broken invariant from hanging the GPU.  A wait that reaches the
// bound counts the predecessor as empty and sets kTicketError in offs[0],
// which poisons the launch's offsets.
template <int PER = kLookPer, class Slot = NextSlot>
__device__ __forceinline__ uint64_t look_back(uint64_t b, uint64_t size, uint64_t *__restrict__ offs,
                                              Slot slot = Slot()) {
    constexpr int kWin = kWave * PER;
    const int lane = threadIdx.x % kWave;
    const LookbackHook hook = g_lookback_hook;
    const bool publish = b != hook.withhold;
    if (b == 0) {
        if (lane == 0 && publish) store_status(&offs[slot(0)], kStatusInclusive | size);
        return 0;
    }
    if (lane == 0 && publish) store_status(&offs[slot((int64_t)b)], kStatusAggregate | size);
    uint64_t excl = 0;
    bool gave_up = false;
    for (int64_t top = (int64_t)b - 1;; top -= kWin) {
        uint64_t stw[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int64_t p = top - (k * kWave + lane);
            stw[k] = p >= 0 ? load_status(&offs[slot(p)]) : kStatusInclusive;
        }
        bool found, gu;
        excl += resolve_window<PER>(stw, top, offs, hook.spin_limit, found, gu, slot);
        gave_up = gave_up || gu;
        if (found) break;
    }
    if (lane == 0 && publish) store_status(&offs[slot((int64_t)b)], kStatusInclusive | (excl + size));
    if (gave_up && lane == 0) atomicOr((unsigned long long *)offs, (unsigned long long)kTicketError);
    return excl;
}

// The workgroup look-back in two steps, so that a workgroup can publish a
// unit's size, do other work, and resolve the unit's prefix later.
// publish_size: the unit's status word (an inclusive prefix for unit 0, an
// aggregate otherwise), by thread 0.
__device__ __forceinline__ void publish_size(uint64_t b, uint64_t size, uint64_t *__restrict__ offs) {
    if (threadIdx.x == 0 && b != g_lookback_hook.withhold)
        store_status(&offs[b + 1], (b == 0 ? kStatusInclusive : kStatusAggregate) | size);
}

// resolve_wg: the exclusive prefix of unit b (every thread of the NT-thread
// workgroup returns it), then its inclusive prefix published.  Each wave
// resolves its own kWave*PER predecessors of a window of NT*PER, nearest
// first by wave, so that one round trip of status loads covers NT/kWave
// times the predecessors of a one-wave look-back.  The waves' sums and flags
// meet in LDS (lbs, lbf: one slot per wave); the window ends at the first
// wave that saw an inclusive prefix.  The first window's words come from
// lookback_issue(), called earlier so that their round trip overlaps other
// work.  Waiting and giving up as look_back().
template <int NT, int PER>
__device__ __forceinline__ void lookback_issue(uint64_t b, const uint64_t *__restrict__ offs, uint64_t (&stw)[PER]) {
    constexpr int kWin = kWave * PER;
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t wtop = (int64_t)b - 1 - (int64_t)wid * kWin;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int64_t p = wtop - (k * kWave + lane);
        stw[k] = p >= 0 ? load_status(&offs[p + 1]) : kStatusInclusive;
    }
}

template <int NT, int PER>
__device__ __forceinline__ uint64_t resolve_wg(uint64_t b, uint64_t size, uint64_t *__restrict__ offs,
                                               uint64_t *lbs, uint32_t *lbf, uint64_t (&first)[PER]) {
    constexpr int kWaves = NT / kWave;
    constexpr int kWin = kWave * PER;
    if (b == 0) return 0;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const LookbackHook hook = g_lookback_hook;
    uint64_t excl = 0;
    bool gave_up = false;
    for (int64_t top = (int64_t)b - 1;; top -= (int64_t)kWin * kWaves) {
        const int64_t wtop = top - (int64_t)wid * kWin;
        uint64_t stw[PER];
        if (top == (int64_t)b - 1) {
#pragma unroll
            for (int k = 0; k < PER; ++k) stw[k] = first[k];
        } else {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int64_t p = wtop - (k * kWave + lane);
                stw[k] = p >= 0 ? load_status(&offs[p + 1]) : kStatusInclusive;
            }
        }
        bool done, wgu;
        const uint64_t wsum = resolve_window<PER>(stw, wtop, offs, hook.spin_limit, done, wgu);
        if (lane == 0) {
            lbs[wid] = wsum;
            lbf[wid] = (done ? 1u : 0u) | (wgu ? 2u : 0u);
        }
        __syncthreads();
        bool found = false;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            if (!found) {
                const uint32_t f = lbf[w];
                excl += lbs[w];
                gave_up = gave_up || (f & 2u);
                found = (f & 1u) != 0;
            }
        }
        __syncthreads();  // the slots are rewritten by the next window
        if (found) break;
    }
    if (tid == 0) {
        if (b != hook.withhold) store_status(&offs[b + 1], kStatusInclusive | (excl + size));
        if (gave_up) atomicOr((unsigned long long *)offs, (unsigned long long)kTicketError);
    }
    return excl;
}

// Orders one wave's LDS accesses around the point (LDS executes a wave's
// instructions in issue order; the fences keep the compiler from moving
// accesses across it).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Parameter-table fields, written one by one with relaxed atomic stores: the
// compiler never merges those into one 12- or 16-byte store, and a wide store
// can read its upper registers after a later LDS load has refilled them.
__device__ __forceinline__ void put_param_u64(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void put_param(leco_fix_block_param *p, double t0, bool zero_t1, uint32_t bits,
                                          bool with_offset, uint64_t payload_offset) {
    put_param_u64((uint64_t *)&p->theta0, double_to_bits(t0));
    if (zero_t1) put_param_u64((uint64_t *)&p->theta1, 0);
    __hip_atomic_store(&p->bits, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (with_offset) put_param_u64(&p->payload_offset, payload_offset);
}

// Where a block's bytes go and what its header holds.
struct PackedBlock {
    uint64_t off, end;     // byte range of the block in the stream
    uint64_t t0b, t1b;     // header theta bits
    uint32_t m;            // values
    int l;                 // bit width
    uint32_t hdr;          // header bytes (1 raw, 17 otherwise)
    uint32_t nchunks;
    int lbyte = -1;        // header byte 0 when it differs from the field width l
};

// Image stores are 8 bytes wide: a lane's 8 bytes are read from the VGPRs as
// the store issues.  A 16-byte store can read its upper 8 bytes later, after
// a load that the compiler placed behind it has already refilled the same
// registers, and then writes the load's data.
typedef uint32_t img_u32x2 __attribute__((ext_vector_type(2)));

// Bytes [g0, g1) of the output from the LDS image `buf` of the 16-byte
// aligned range [a0, a0 + span): lane-contiguous 8-byte stores, byte stores
// at the ragged ends (the neighbouring blocks' bytes are not touched).
template <int NT>
__device__ __forceinline__ void store_image(const uint32_t *buf, uint64_t a0, uint64_t g0, uint64_t g1, uint32_t span,
                                            int tid, uint8_t *__restrict__ out) {
    const uint8_t *bb = (const uint8_t *)buf;
    for (uint32_t u = tid; u < span / 8; u += NT) {
        const uint64_t ga = a0 + 8ull * u;
        if (ga >= g0 && ga + 8 <= g1) {
            *(img_u32x2 *)(out + ga) = img_u32x2{buf[2 * u], buf[2 * u + 1]};
        } else {
            for (int k = 0; k < 8; ++k) {
                const uint64_t gb = ga + k;
                if (gb >= g0 && gb < g1) out[gb] = bb[8 * u + k];
            }
        }
    }
}

// Bytes [0, size) of the LDS image `img` (a block packed at lead 0) to
// out[off, off + size): the 8-byte aligned middle with lane-contiguous 8-byte
// stores, each built from three image words funnel-shifted by the same byte
// count; the ragged head and tail bytes (fewer than 8 each, shared with
// neighbouring blocks' segments) with byte stores, one thread each.  Reads
// image words up to (size + 11) / 4; the image has that room for the blocks
// stored here.
template <int NT>
__device__ __forceinline__ void store_image_shifted(const uint32_t *img, uint64_t off, uint64_t size, int tid,
                                                    uint8_t *__restrict__ out) {
    static_assert(NT >= 16, "head and tail bytes take 16 threads");
    const uint64_t end = off + size;
    const uint64_t a0 = (off + 7) & ~7ull;  // first 8-byte boundary at or after off
    const uint64_t hend = a0 < end ? a0 : end;
    const uint64_t a1 = a0 < end ? end & ~7ull : a0;  // the middle is [a0, a1)
    auto byte_at = [&](uint64_t g) -> uint32_t {
        const uint32_t i = (uint32_t)(g - off);
        return (img[i >> 2] >> (8 * (i & 3))) & 0xFF;
    };
    if (tid < 8) {
        const uint64_t g = off + tid;
        if (g < hend) out[g] = (uint8_t)byte_at(g);
    } else if (tid < 16) {
        const uint64_t g = a1 + (tid - 8);
        if (g >= a1 && g < end && a0 < end) out[g] = (uint8_t)byte_at(g);
    }
    const uint32_t i0 = (uint32_t)(a0 - off);  // image byte of the middle's first byte
    const uint32_t r = i0 & 3;
    const uint32_t units = (uint32_t)((a1 - a0) / 8);
    for (uint32_t u = tid; u < units; u += NT) {
        const uint32_t q = (i0 + 8 * u) >> 2;
        const uint32_t w0 = img[q], w1 = img[q + 1], w2 = img[q + 2];
        *(img_u32x2 *)(out + a0 + 8ull * u) =
            img_u32x2{__builtin_amdgcn_alignbyte(w1, w0, r), __builtin_amdgcn_alignbyte(w2, w1, r)};
    }
}

// Chunk c (CH values) of a block, packed by NT threads (thread `tid`) through
// the LDS image `buf`; sync() orders the group's LDS accesses.  value(kk, i)
// is the field of block value i = chunk position tid + kk*NT (kk a
// compile-time index, so register-cached values stay in registers).
//
// Every field is computed (any global loads in value() included) before the
// chunk's first LDS write: an LDS instruction still queued behind others
// (the image ORs are atomics that queue deeply under contention) can read
// its address and data registers after a later global load has refilled
// them, and then ORs the wrong bits at the wrong place.  No load is issued
// between the chunk's first LDS write and the wait for all of them at the end.
template <int NT, uint32_t CH, class Sync, class Value>
__device__ __forceinline__ void pack_chunk(const PackedBlock &pb, uint32_t c, int tid, uint32_t *buf,
                                           uint8_t *__restrict__ out, Sync sync, Value value) {
    constexpr int K = (int)(CH / NT);
    const int l = pb.l;
    const uint64_t cbytes = (uint64_t)(CH / 8) * l;
    const uint64_t g0 = c == 0 ? pb.off : pb.off + pb.hdr + c * cbytes;
    const uint64_t g1 = c + 1 == pb.nchunks ? pb.end : pb.off + pb.hdr + (c + 1) * cbytes;
    const uint64_t a0 = g0 & ~15ull;
    const uint32_t span = (uint32_t)(((g1 + 15) & ~15ull) - a0);
    const uint32_t lead = (uint32_t)(g0 - a0);
    const uint32_t j0 = c * CH;
    const uint32_t jn = pb.m - j0 < CH ? pb.m - j0 : CH;
    uint64_t field[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
        const uint32_t j = tid + kk * NT;  // position inside the chunk
        field[kk] = l > 0 && j < jn ? value(kk, j0 + j) : 0;
    }
    for (uint32_t w = tid; w < span / 4; w += NT) buf[w] = 0;
    sync();
    if (c == 0 && (uint32_t)tid < pb.hdr) {
        uint32_t byte;
        if (tid == 0) byte = (uint32_t)(pb.lbyte >= 0 ? pb.lbyte : l);
        else if (tid < 9) byte = (uint32_t)(pb.t0b >> (8 * (tid - 1))) & 0xFF;
        else byte = (uint32_t)(pb.t1b >> (8 * (tid - 9))) & 0xFF;
        const uint32_t pos = lead + tid;
        atomicOr(&buf[pos >> 2], byte << (8 * (pos & 3)));
    }
    if (l > 0) {
        const uint64_t pbit = 8ull * (lead + (c == 0 ? pb.hdr : 0));
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const uint32_t j = tid + kk * NT;
            if (j < jn) {
                const uint64_t q = pbit + (uint64_t)j * l;
                const uint32_t dw = (uint32_t)(q >> 5);
                const int shb = (int)(q & 31);
                const uint32_t d0 = (uint32_t)(field[kk] << shb);
                const uint32_t d1 = (uint32_t)(shb ? field[kk] >> (32 - shb) : field[kk] >> 32);
                const uint32_t d2 = shb ? (uint32_t)(field[kk] >> (64 - shb)) : 0u;
                if (d0) atomicOr(&buf[dw], d0);
                if (shb + l > 32 && d1) atomicOr(&buf[dw + 1], d1);
                if (shb + l > 64 && d2) atomicOr(&buf[dw + 2], d2);
            }
        }
    }
    sync();
    store_image<NT>(buf, a0, g0, g1, span, tid, out);
    sync();
    // every LDS operation of the chunk done before the caller's next loads
    // (s_waitcnt lgkmcnt(0); a wave-scope sync does not wait)
    __builtin_amdgcn_s_waitcnt(0xC07F);
}

// Small-block decode staging: the bytes of whole blocks, their offsets and
// per-block header tables (width | first-field bit << 7, thetas).
// Little-endian 8 bytes at LDS byte offset `byte` (two aligned words).
__device__ __forceinline__ uint64_t lds_u64_at(const uint64_t *__restrict__ buf, uint32_t byte) {
    const uint32_t w = byte >> 3;
    const int sh = (int)(byte & 7) * 8;
    return (buf[w] >> sh) | ((buf[w + 1] << 1) << (63 - sh));
}

// lw-bit field (lw <= W) at LDS bit q, from two aligned words: 32-bit words
// and one funnel shift for W = 32, 64-bit words otherwise.  Reads the word
// after the field's first one.
template <int W>
__device__ __forceinline__ uint64_t lds_field(const uint64_t *__restrict__ buf, uint32_t q, int lw) {
    if constexpr (W == 32) {
        const uint32_t *b32 = (const uint32_t *)buf;
        const uint32_t w = __builtin_amdgcn_alignbit(b32[(q >> 5) + 1], b32[q >> 5], q & 31);
        return lw >= 32 ? w : (w & ((1u << lw) - 1));
    } else {
        const uint32_t w = q >> 6;
        const int sh = (int)(q & 63);
        return ((buf[w] >> sh) | ((buf[w + 1] << 1) << (63 - sh))) & low_mask(lw);
    }
}

struct MultiShared {
    alignas(16) uint64_t buf[kLdsBytes / 8];
    uint64_t offs[kMultiMax + 1];
    double2 ht[kMultiMax];
    uint32_t hlb[kMultiMax];
};

// One ticket's worth of encode work: `count` consecutive blocks of one
// column, encoded one per wave (`wave`) or a single block by the whole
// workgroup.
struct EncodeUnit {
    const uint8_t *base;  // the column's values
    uint64_t n;
    uint32_t bs;
    uint32_t width;
    uint64_t cb0;         // first block of the unit, within the column
    uint64_t gb0;         // its index in the param table / offsets
    uint32_t count;       // blocks in the unit (0 past the last unit)
    bool wave;
};

// Blocks of one column.
struct PlainRows {
    const uint8_t *in;
    uint64_t n;
    uint32_t bs;
    uint32_t width;
    uint64_t nb;
    __device__ __forceinline__ EncodeUnit unit(uint64_t u) const {
        const uint32_t per = encode_unit_blocks(bs);
        const uint64_t cb0 = u * per;
        const uint32_t count = cb0 < nb ? (uint32_t)(nb - cb0 < per ? nb - cb0 : per) : 0;
        return {in, n, bs, width, cb0, cb0, count, bs <= kWaveBlock};
    }
};

// 64 consecutive blocks of one column, the unit of the fit kernel.
struct FitGroup {
    const uint8_t *base;  // the column's values
    uint64_t n;
    uint32_t bs;
    uint32_t width;
    uint64_t b0;          // first block of the group, within the column
    uint64_t gb0;         // its index in the param table / offsets
};

// Fit groups g0, g0 + 1, ... of one column (a launch may cover a slice of
// the column's groups); the launch with g0 = 0 also clears the ticket.
struct PlainFit {
    const uint8_t *in;
    uint64_t n;
    uint32_t bs;
    uint32_t width;
    uint64_t g0;
    __device__ __forceinline__ FitGroup group(uint32_t wg) const {
        const uint64_t b0 = (g0 + wg) * kFitRows;
        return {in, n, bs, width, b0, b0};
    }
    __device__ __forceinline__ bool clears_ticket(uint32_t wg) const { return g0 == 0 && wg == 0; }
};

// Largest c with first[c] <= it (first[0] = 0 <= it < first[ncols]).
__device__ __forceinline__ uint32_t find_column(const uint64_t *__restrict__ first, uint32_t ncols, uint64_t it) {
    uint32_t lo = 0, hi = ncols;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (first[mid] <= it) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Fit groups of many columns: fit_first[c] is column c's first group.
struct BatchFit {
    const BatchColumn *cols;
    const uint64_t *first;
    const uint64_t *fit_first;
    uint32_t ncols;
    __device__ __forceinline__ bool clears_ticket(uint32_t wg) const { return wg == 0; }
    __device__ __forceinline__ FitGroup group(uint32_t wg) const {
        const uint32_t c = find_column(fit_first, ncols, wg);
        const BatchColumn col = cols[c];
        const uint64_t b0 = ((uint64_t)wg - fit_first[c]) * kFitRows;
        return {(const uint8_t *)col.in, col.n, col.bs, col.width, b0, first[c] + b0};
    }
};

// Encode units of many columns, in column order: unit_first[c] is column c's
// first unit, unit_col[u] unit u's column, first[c] column c's first block.
struct BatchRows {
    const BatchColumn *cols;
    const uint64_t *first;
    const uint64_t *unit_first;
    const uint32_t *unit_col;
    uint64_t units;
    __device__ __forceinline__ EncodeUnit unit(uint64_t u) const {
        if (u >= units) return {nullptr, 0, 1, 64, 0, 0, 0, false};
        const uint32_t c = unit_col[u];
        const BatchColumn col = cols[c];
        const uint32_t per = encode_unit_blocks(col.bs);
        const uint64_t nb = first[c + 1] - first[c];
        const uint64_t cb0 = (u - unit_first[c]) * per;
        const uint32_t count = (uint32_t)(nb - cb0 < per ? nb - cb0 : per);
        return {(const uint8_t *)col.in, col.n, col.bs, col.width, cb0, first[c] + cb0, count, col.bs <= kWaveBlock};
    }
};

}  // namespace leco
