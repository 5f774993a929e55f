// gfx950 kernels of piecewise_fix_op_max, the uint32 IntegerCODEC of the
// reference (headers/piecewise_fix_op_minimize_maxerror.h:45-197), on the
// LeCo column layout: the blocks' byte strings back to back, block b at
// offs[b], plus a leco_fix_block_param per block for point lookups.
//
// The codec predicts with a float slope.  Its encoder stores residuals
// against pm_i = (long long)(theta0 + theta1*i) (one fused rounding), while
// its decoder rebuilds values from the additive chain p_0 = theta0,
// p_{i+1} = p_i + theta1 (read_all_bit_fix_add, bit_read.h:559-604); where
// the two truncate differently the encoder records the position in
// mul_add_diff_set (+1) or mul_add_diff_set_minus (-1) and the harness
// corrects the decoded value (experiments/fix_int.cpp:110-117).  Here those
// lists are an int8 per value (fix), compacted into the reference's sorted
// lists on request; the decoder can apply fix itself.  The chain is evaluated
// exactly at any index by leco_chain.h, so every thread starts its run of
// values where it likes.
//
// Encode = the LeCo fit kernel (lr_int::caltheta, lr.h:89-111, is the same
// recipe as lr_int_T) -> one wave per block, opmax_unit_blocks() blocks per
// workgroup: residual range, centring, bit width, size; one decoupled
// look-back per unit; pack through the shared LDS-image packer; fix flags ->
// offsets from the param table.  Decode = kOpmaxRun consecutive values per
// thread, fields staged through LDS.  Lookup = one lane per query.
#include "leco_chain.h"
#include "leco_kernels.h"

namespace leco {

LECO_LOOKBACK_HOOK_SETTER(set_lookback_hook_opmax)

namespace {

// (float)d as x86 CVTSD2SS: round to nearest even; a NaN keeps its sign and
// the top 22 bits of its payload, quieted.
__device__ __forceinline__ float narrow_x86(double d) {
    if (__builtin_isnan(d)) {
        const uint64_t b = double_to_bits(d);
        const uint32_t f = (uint32_t)(b >> 32 & 0x80000000u) | 0x7FC00000u | (uint32_t)(b >> 29 & 0x3FFFFFu);
        return __uint_as_float(f);
    }
    return (float)d;
}

// (long long)(theta0 + theta1 * (double)i): the encoder's and random access's
// prediction (piecewise_fix_op_minimize_maxerror.h:58,72; bit_read.h:1127).
__device__ __forceinline__ int64_t opmax_pm(double t0, double t1, uint32_t i) {
    return cvtt64(__builtin_fma(t1, (double)i, t0));
}

// (int)(y - pm): the reference's int residual (wraps mod 2^32).
__device__ __forceinline__ int32_t opmax_resid(uint32_t y, int64_t pm) {
    return (int32_t)(uint32_t)((uint64_t)y - (uint64_t)pm);
}

// write_delta_T field of residual d (bit_write.h:213-223): sign 1 iff d > 0.
__device__ __forceinline__ uint64_t opmax_field(int32_t d, int l) {
    const uint32_t mag = d > 0 ? (uint32_t)d : 0u - (uint32_t)d;
    return (uint64_t)(mag & ((1u << (l - 1)) - 1)) | ((uint64_t)(d > 0) << (l - 1));
}

// 32-bit wave reductions on the 64-bit DPP reduction (sign-extended for the
// signed ones): VALU and scalar reads, no LDS queue.
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) { return (int32_t)wave_max_i64((int64_t)v); }
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) { return (int32_t)wave_min_i64((int64_t)v); }
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) { return (uint32_t)wave_max_u64((uint64_t)v); }

// ---------------------------------------------------------------------------
// Encode: encodeArray8 (:45-131) per block, one wave per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_opmax_encode(const uint32_t *__restrict__ in, uint64_t n, uint32_t bs,
                                                           leco_fix_block_param *__restrict__ prm,
                                                           uint64_t *__restrict__ offs, uint8_t *__restrict__ out,
                                                           int8_t *__restrict__ fix) {
    __shared__ EncodeShared sh;
    __shared__ uint32_t s_t1b[kUnitMaxBlocks];  // theta1 as stored: float bits
    __shared__ uint64_t s_u;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    constexpr uint32_t kWaves = kThreads / kWave;
    if (tid == 0) s_u = take_ticket(offs);  // offs[0] is the ticket
    __syncthreads();
    const uint64_t u = s_u;
    const uint64_t nb = (n + bs - 1) / bs;
    const uint32_t per = opmax_unit_blocks(bs);
    const uint64_t b0 = u * per;
    const uint32_t count = (uint32_t)(nb - b0 < per ? nb - b0 : per);

    // phase A: model, residual range, centring, bit width, size
    for (uint32_t k = wid; k < count; k += kWaves) {
        const uint64_t b = b0 + k, s = b * bs;
        const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
        const uint32_t *y = in + s;
        double t0 = prm[b].theta0;
        const float t1f = narrow_x86(prm[b].theta1);
        const double t1 = (double)t1f;
        uint32_t v[kWaveVPL];
#pragma unroll
        for (int q = 0; q < kWaveVPL; ++q) {
            const uint32_t j = lane + q * kWave;
            v[q] = j < m ? y[j] : 0u;
        }
        // :53-65 residual range under the fitted line
        int32_t mx = INT32_MIN, mn = INT32_MAX;
#pragma unroll
        for (int q = 0; q < kWaveVPL; ++q) {
            const uint32_t j = lane + q * kWave;
            const int32_t d = opmax_resid(v[q], opmax_pm(t0, t1, j));
            mx = (j < m && d > mx) ? d : mx;
            mn = (j < m && d < mn) ? d : mn;
        }
        for (uint32_t j = lane + kWaveChunk; j < m; j += kWave) {
            const int32_t d = opmax_resid(y[j], opmax_pm(t0, t1, j));
            mx = d > mx ? d : mx;
            mn = d < mn ? d : mn;
        }
        mx = wave_max_i32(mx);
        mn = wave_min_i32(mn);
        // :66 theta0 += (max + min) / 2.0 (int sum wraps); a NaN theta0 keeps its bits
        if (!__builtin_isnan(t0)) t0 = t0 + (double)(int32_t)((uint32_t)mx + (uint32_t)mn) / 2.0;
        // :69-90 largest |residual| (abs(INT_MIN) is negative there and never the maximum)
        uint32_t mg = 0;
#pragma unroll
        for (int q = 0; q < kWaveVPL; ++q) {
            const uint32_t j = lane + q * kWave;
            const int32_t d = opmax_resid(v[q], opmax_pm(t0, t1, j));
            const uint32_t a = d == INT32_MIN ? 0u : (uint32_t)(d < 0 ? -d : d);
            mg = (j < m && a > mg) ? a : mg;
        }
        for (uint32_t j = lane + kWaveChunk; j < m; j += kWave) {
            const int32_t d = opmax_resid(y[j], opmax_pm(t0, t1, j));
            const uint32_t a = d == INT32_MIN ? 0u : (uint32_t)(d < 0 ? -d : d);
            mg = a > mg ? a : mg;
        }
        mg = wave_max_u32(mg);
        const int l = mg ? (32 - __clz((int)mg)) + 1 : 0;  // :95-98 bits(max_error) + 1
        if (lane == 0) {
            sh.ut0[k] = t0;
            sh.ut1[k] = t1;
            s_t1b[k] = __float_as_uint(t1f);
            sh.ul[k] = l;
            sh.usz[k] = opmax_block_bytes((uint32_t)l, m);
        }
    }
    __syncthreads();

    // unit offset: exclusive scan of the sizes (one per lane) + look-back
    if (wid == 0) {
        const uint64_t mine = (uint32_t)lane < count ? sh.usz[lane] : 0;
        const uint64_t inc = wave_inclusive_sum(mine);
        const uint64_t total = wave_last(inc);
        const uint64_t excl = look_back(u, total, offs);
        if ((uint32_t)lane < count) sh.usz[lane] = excl + inc - mine;  // now: the block's offset
    }
    __syncthreads();

    // phase B: pack (:100-126), then the fix-up flags (:71-84)
    uint32_t *buf = sh.buf + wid * (kWaveLdsBytes / 4);
    for (uint32_t k = wid; k < count; k += kWaves) {
        const uint64_t b = b0 + k, s = b * bs;
        const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
        const uint32_t *y = in + s;
        const int l = __builtin_amdgcn_readfirstlane(sh.ul[k]);
        const double t0 = sh.ut0[k], t1 = sh.ut1[k];
        const uint64_t off = sh.usz[k];
        const bool raw = l >= 31;  // write_delta_default: the values themselves, 4 bytes each
        const int lf = raw ? 32 : l;
        if (lane == 0) {
            put_param(&prm[b], t0, false, (uint32_t)l, true, off + kOpmaxHdr);
            put_param_u64((uint64_t *)&prm[b].theta1, double_to_bits(t1));
        }
        PackedBlock pb{off, off + opmax_block_bytes((uint32_t)l, m), double_to_bits(t0), (uint64_t)s_t1b[k], m, lf,
                       kOpmaxHdr,
                       lf == 0 ? 1 : (m + kWaveChunk - 1) / kWaveChunk};
        pb.lbyte = l;
        auto sync = [] { wave_lds_sync(); };
        for (uint32_t c = 0; c < pb.nchunks; ++c) {
            uint64_t f[kWaveVPL];
#pragma unroll
            for (int q = 0; q < kWaveVPL; ++q) {
                const uint32_t i = c * kWaveChunk + lane + q * kWave;
                const uint32_t v = i < m ? y[i] : 0u;
                f[q] = (raw || lf == 0) ? (uint64_t)v : opmax_field(opmax_resid(v, opmax_pm(t0, t1, i)), lf);
            }
            pack_chunk<kWave, kWaveChunk>(pb, c, lane, buf, out, sync,
                                                [&](int q, uint32_t) -> uint64_t { return f[q]; });
        }
        if (fix) {
            // runs of 8 values per lane (one 8-byte store): p_j of the decoder's chain against pm_j
            constexpr uint32_t kRun = 8;
            for (uint32_t r0 = lane * kRun; r0 < m; r0 += kWave * kRun) {
                double p = chain_at(t0, t1, r0, m);
                uint64_t word = 0;
#pragma unroll
                for (uint32_t q = 0; q < kRun; ++q) {
                    const uint32_t j = r0 + q;
                    const int64_t pa = cvtt64(p), pm = opmax_pm(t0, t1, j);
                    const int8_t fv = pm > pa ? 1 : (pm < pa ? -1 : 0);
                    word |= (uint64_t)(uint8_t)fv << (8 * q);
                    p = p + t1;
                }
                int8_t *dst = fix + s + r0;
                if (r0 + kRun <= m && ((uintptr_t)dst & 7) == 0) {
                    *(uint64_t *)dst = word;
                } else {
                    for (uint32_t q = 0; q < kRun && r0 + q < m; ++q) dst[q] = (int8_t)(word >> (8 * q));
                }
            }
        }
    }
}

// Block end offsets offs[b+1] from the param table; the ticket back to 0;
// offs[nb] is poisoned if a look-back wait gave up.
__global__ __launch_bounds__(kThreads) void k_opmax_fixup(const leco_fix_block_param *__restrict__ prm,
                                                          uint64_t *__restrict__ offs, uint64_t nb) {
    const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b >= nb) return;
    const leco_fix_block_param p = prm[b];
    uint64_t end = p.payload_offset - kOpmaxHdr + opmax_block_bytes(p.bits, p.length);
    if (b + 1 == nb) {
        if (offs[0] & kTicketError) end = LECO_FIX_OFFSETS_POISON;
        offs[0] = 0;
    }
    offs[b + 1] = end;
}

// ---------------------------------------------------------------------------
// Decode: decodeArray8 (:134-169).  Thread t owns a run of kOpmaxRun
// consecutive values of one block (runs start at multiples of kOpmaxRun
// inside their block; a block's last run may be shorter), a workgroup
// kThreads consecutive runs.  The stream bytes of the workgroup's fields (a
// contiguous range across its blocks) are staged through LDS with 16-byte
// loads; each thread places p at its first value with chain_at, then steps
// the chain exactly as the reference does, reading each field as two LDS
// words and one funnel shift.  A workgroup whose bytes exceed the staging
// buffer (tiny blocks: mostly headers) reads fields from global memory.  FIX:
// add the fix-up flag of every value of a packed or model-only block (raw
// blocks hold the values themselves).
// ---------------------------------------------------------------------------
constexpr uint32_t kOpmaxTile = kThreads * kOpmaxRun;
constexpr uint32_t kOpmaxStage = kOpmaxTile * 4 + 1024;  // raw fields of a tile + headers + slack

struct OpmaxBlock {
    uint64_t off;  // stream offset of the block
    double t0, t1;
    uint32_t m;
    int l;
};

__device__ __forceinline__ OpmaxBlock opmax_block(const uint8_t *__restrict__ st, const uint64_t *__restrict__ offs,
                                                  uint64_t n, uint32_t bs, uint64_t b) {
    OpmaxBlock k;
    k.off = offs[b];
    const uint8_t *blk = st + k.off;
    k.l = blk[0];
    k.t0 = bits_to_double(load_u64_any(blk + 1));
    k.t1 = (double)__uint_as_float((uint32_t)load_u64_any(blk + 9));
    const uint64_t s = b * bs;
    k.m = (uint32_t)(n - s < bs ? n - s : bs);
    return k;
}

// Field width as stored: raw blocks (l >= 31) hold 32-bit values.
__device__ __forceinline__ uint32_t opmax_lf(int l) { return l >= 31 ? 32u : (uint32_t)l; }

// Low 32 bits of cvtt64(x), the only bits a uint32 output keeps (fast path:
// the low word of t + 1.5*2^52 is t mod 2^32 for |t| < 2^51).
__device__ __forceinline__ uint32_t cvtt_lo32(double x) {
    const double t = __builtin_trunc(x);
    if (__builtin_expect(__builtin_fabs(t) < 0x1p51, 1)) return (uint32_t)__double_as_longlong(t + 0x1.8p52);
    return (uint32_t)cvtt64(x);
}

// Value from a field f (l bits, 2 <= l <= 30) and the prediction's low word.
__device__ __forceinline__ uint32_t opmax_apply(uint32_t pa, uint32_t f, int l) {
    const uint32_t mag = f & ((1u << (l - 1)) - 1);
    return (f >> (l - 1)) & 1 ? pa + mag : pa - mag;
}

// Runs of a column: each block splits into ceil(m / kOpmaxRun) runs that
// start at multiples of kOpmaxRun inside the block, so no run crosses a block.
__host__ __device__ inline uint32_t opmax_runs_per_block(uint32_t bs) { return (bs + kOpmaxRun - 1) / kOpmaxRun; }
__host__ __device__ inline uint64_t opmax_runs(uint64_t n, uint32_t bs) {
    const uint64_t nb = (n + bs - 1) / bs;
    const uint32_t last = (uint32_t)(n - (nb - 1) * bs);
    return (nb - 1) * opmax_runs_per_block(bs) + (last + kOpmaxRun - 1) / kOpmaxRun;
}

// Byte just past the field of value j in a block at stream offset `off` (raw: 4-byte values).
__device__ __forceinline__ uint64_t opmax_field_end(uint64_t off, int l, uint32_t j) {
    const uint32_t lf = opmax_lf(l);
    return off + kOpmaxHdr + (((uint64_t)(j + 1) * lf + 7) >> 3);
}

template <bool FIX>
__global__ __launch_bounds__(kThreads) void k_opmax_decode(const uint8_t *__restrict__ st,
                                                           const uint64_t *__restrict__ offs, uint64_t n, uint32_t bs,
                                                           uint64_t runs, const int8_t *__restrict__ fix,
                                                           uint32_t *__restrict__ out) {
    __shared__ uint32_t lds[kOpmaxStage / 4];
    __shared__ uint64_t s_a0;
    __shared__ uint32_t s_words, s_blocks;
    const int tid = threadIdx.x;
    const uint32_t rpb = opmax_runs_per_block(bs);
    const uint64_t g0 = (uint64_t)blockIdx.x * kThreads;            // first run of the workgroup
    const uint64_t b0 = g0 / rpb;
    const uint32_t rr0 = (uint32_t)(g0 - b0 * rpb);
    const uint64_t g = g0 + tid;
    const bool live = g < runs;
    const uint32_t rq = rr0 + tid;
    const uint64_t b = b0 + rq / rpb;
    const uint32_t j0 = (rq % rpb) * kOpmaxRun;
    uint64_t boff = live ? offs[b] : 0;  // in flight while the range is found
    if (tid == 0) {
        // the workgroup's whole blocks (headers included) when they fit, else
        // just the stream bytes of its fields
        const uint64_t gl = (g0 + kThreads < runs ? g0 + kThreads : runs) - 1;
        const uint64_t bl = gl / rpb;
        const uint64_t o0 = offs[b0], oe = offs[bl + 1];
        uint64_t a0 = o0 & ~15ull, a1 = ((oe + 15) & ~15ull) + 16;
        s_blocks = a1 - a0 <= kOpmaxStage;
        if (!s_blocks) {
            const uint64_t sl = bl * bs;
            const uint32_t ml = (uint32_t)(n - sl < bs ? n - sl : bs);
            const uint32_t jl0 = (uint32_t)(gl - bl * rpb) * kOpmaxRun;
            const uint32_t jl = (jl0 + kOpmaxRun < ml ? jl0 + kOpmaxRun : ml) - 1;
            const uint64_t ol = offs[bl];
            const uint64_t ga = o0 + kOpmaxHdr + (((uint64_t)rr0 * kOpmaxRun * opmax_lf(st[o0])) >> 3);
            const uint64_t gb = opmax_field_end(ol, st[ol], jl);
            a0 = ga & ~15ull;
            a1 = ((gb + 15) & ~15ull) + 16;
        }
        s_a0 = a0;
        s_words = a1 - a0 <= kOpmaxStage ? (uint32_t)((a1 - a0) / 4) : 0u;
    }
    __syncthreads();
    const uint64_t a0 = s_a0;
    const uint32_t words = s_words;
    for (uint32_t w = tid; w < words / 4; w += kThreads) ((uint4 *)lds)[w] = *(const uint4 *)(st + a0 + 16ull * w);
    __syncthreads();

    uint32_t o[kOpmaxRun];
    uint32_t len = 0;
    uint64_t i0 = 0;
    if (live) {
    OpmaxBlock k;
    if (s_blocks) {
        // header from the staged block: [l][theta0][theta1] at LDS byte boff - a0
        const uint32_t x = (uint32_t)(boff - a0);
        auto u32_at = [&](uint32_t byte) -> uint32_t {
            return __builtin_amdgcn_alignbit(lds[(byte >> 2) + 1], lds[byte >> 2], 8 * (byte & 3));
        };
        k.off = boff;
        k.l = (int)(u32_at(x) & 0xFF);
        k.t0 = bits_to_double((uint64_t)u32_at(x + 1) | ((uint64_t)u32_at(x + 5) << 32));
        k.t1 = (double)__uint_as_float(u32_at(x + 9));
        const uint64_t sb = b * bs;
        k.m = (uint32_t)(n - sb < bs ? n - sb : bs);
    } else {
        k = opmax_block(st, offs, n, bs, b);
    }
    len = k.m - j0 < kOpmaxRun ? k.m - j0 : kOpmaxRun;
    i0 = b * bs + j0;
    const bool whole = len == kOpmaxRun;
    uint64_t fw[kOpmaxRun / 8] = {};
    if (FIX) {
        if (whole && ((uintptr_t)(fix + i0) & 7) == 0) {
#pragma unroll
            for (uint32_t h = 0; h < kOpmaxRun / 8; ++h) fw[h] = *(const uint64_t *)(fix + i0 + 8 * h);
        } else {
            for (uint32_t r = 0; r < len; ++r) fw[r / 8] |= (uint64_t)(uint8_t)fix[i0 + r] << (8 * (r % 8));
        }
    }
    auto flag = [&](uint32_t r) -> uint32_t { return (uint32_t)(int32_t)(int8_t)(fw[r / 8] >> (8 * (r % 8))); };
    const uint32_t lf = opmax_lf(k.l);
    const uint32_t bit0 = words ? (uint32_t)((int64_t)(k.off + kOpmaxHdr - a0) * 8) + j0 * lf : 0u;
    const uint8_t *pay = st + k.off + kOpmaxHdr;
    // field r of the run: from LDS when staged, else from global memory
    auto field = [&](uint32_t r) -> uint32_t {
        if (words) {
            const uint32_t qb = bit0 + r * lf;
            const uint32_t v = __builtin_amdgcn_alignbit(lds[(qb >> 5) + 1], lds[qb >> 5], qb & 31);
            return lf == 32 ? v : v & ((1u << lf) - 1);
        }
        return (uint32_t)load_bits(pay, (uint64_t)(j0 + r) * lf, (int)lf);
    };
    if (k.l >= 31) {
#pragma unroll
        for (uint32_t r = 0; r < kOpmaxRun; ++r) {
            o[r] = 0;
            if (r < len) o[r] = field(r);
        }
    } else {
        double p = chain_at(k.t0, k.t1, j0, k.m);
        const int l = k.l;
        // the run's predictions lie between p and p + kOpmaxRun*t (monotone chain), so
        // one test decides whether every truncation takes cvtt_lo32's fast path
        const double pe = __builtin_fma((double)kOpmaxRun, k.t1, p);
        const bool fast = whole && words && __builtin_fabs(p) < 0x1p50 && __builtin_fabs(pe) < 0x1p50;
        if (fast && l) {
            const uint32_t fm = (2u << (l - 1)) - 1;
            uint32_t f[kOpmaxRun];
#pragma unroll
            for (uint32_t r = 0; r < kOpmaxRun; ++r) {  // all LDS reads in flight together
                const uint32_t qb = bit0 + r * lf;
                f[r] = __builtin_amdgcn_alignbit(lds[(qb >> 5) + 1], lds[qb >> 5], qb & 31) & fm;
            }
#pragma unroll
            for (uint32_t r = 0; r < kOpmaxRun; ++r) {
                const uint32_t pa = (uint32_t)__double_as_longlong(__builtin_trunc(p) + 0x1.8p52);
                o[r] = opmax_apply(pa, f[r], l) + (FIX ? flag(r) : 0u);
                p = p + k.t1;
            }
        } else if (fast) {
#pragma unroll
            for (uint32_t r = 0; r < kOpmaxRun; ++r) {
                o[r] = (uint32_t)__double_as_longlong(__builtin_trunc(p) + 0x1.8p52) + (FIX ? flag(r) : 0u);
                p = p + k.t1;
            }
        } else {
            // partial runs (a block's last), unstaged workgroups, predictions beyond 2^50
            for (uint32_t r = 0; r < kOpmaxRun; ++r) {
                o[r] = 0;
                if (r < len) {
                    const uint32_t pa = cvtt_lo32(p);
                    o[r] = (l ? opmax_apply(pa, field(r), l) : pa) + (FIX ? flag(r) : 0u);
                    p = p + k.t1;
                }
            }
        }
    }
    }

    // The workgroup's runs cover consecutive positions [p0, p1): transpose the
    // values through LDS (free again once every thread has read its fields)
    // and write the range with lane-contiguous 8-byte streaming stores (a
    // 16-byte store can read its upper half from the VGPRs only after the
    // next LDS load has refilled them).
    const uint64_t p0 = b0 * bs + (uint64_t)rr0 * kOpmaxRun;
    const uint64_t gl = (g0 + kThreads < runs ? g0 + kThreads : runs) - 1;
    const uint32_t rl = rr0 + (uint32_t)(gl - g0);
    const uint64_t bl = b0 + rl / rpb;
    const uint32_t jl = (rl % rpb) * kOpmaxRun;
    const uint64_t sl = bl * bs;
    const uint32_t ml = (uint32_t)(n - sl < bs ? n - sl : bs);
    const uint32_t cnt = (uint32_t)(sl + (jl + kOpmaxRun < ml ? jl + kOpmaxRun : ml) - p0);
    // value q sits at LDS word q + q/16 (one pad word per run keeps the threads' writes on distinct banks)
    auto at = [](uint32_t q) { return q + (q >> 4); };
    static_assert(kOpmaxTile + kOpmaxTile / 16 <= kOpmaxStage / 4, "a padded tile fits the staging buffer");
    __syncthreads();
    for (uint32_t r = 0; r < len; ++r) lds[at((uint32_t)(i0 - p0) + r)] = o[r];
    __syncthreads();
    uint32_t *dst = out + p0;
    const uint32_t head = ((uintptr_t)dst & 4) && cnt ? 1u : 0u;  // a value before an 8-byte boundary
    if ((uint32_t)tid < head) dst[0] = lds[at(0)];
    const uint32_t nvec = (cnt - head) / 2;
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    for (uint32_t v = tid; v < nvec; v += kThreads) {
        const uint32_t q = head + 2 * v;
        __builtin_nontemporal_store(u32x2{lds[at(q)], lds[at(q + 1)]}, (u32x2 *)(dst + q));
    }
    for (uint32_t q = head + 2 * nvec + tid; q < cnt; q += kThreads) dst[q] = lds[at(q)];
}

// The reference's fix-up lists applied to decoded values, skipping raw blocks
// (their values are stored, not predicted): the input, bit for bit.
__global__ __launch_bounds__(kThreads) void k_opmax_apply_exact(uint32_t *__restrict__ values, uint64_t n,
                                                                uint32_t bs, const uint8_t *__restrict__ st,
                                                                const uint64_t *__restrict__ offs,
                                                                const int32_t *__restrict__ plus, uint64_t np,
                                                                const int32_t *__restrict__ minus, uint64_t nm) {
    const uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    for (int h = 0; h < 2; ++h) {
        if (q >= (h ? nm : np)) continue;
        const int32_t p = (h ? minus : plus)[q];
        if (p < 0 || (uint64_t)p >= n || st[offs[(uint64_t)p / bs]] >= 31) continue;
        atomicAdd(&values[p], h ? 0xFFFFFFFFu : 1u);
    }
}

// ---------------------------------------------------------------------------
// Lookup: randomdecodeArray8 (:171-197) with read_bit_fix_T (bit_read.h:1107-1132):
// value + (long long)(theta0 + j*theta1), exact without any fix-up.  Indices
// >= n give 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_opmax_lookup(const uint8_t *__restrict__ st,
                                                           const leco_fix_block_param *__restrict__ prm, uint64_t n,
                                                           uint32_t bs, const uint64_t *__restrict__ idx, uint64_t nq,
                                                           uint32_t *__restrict__ out) {
    const uint64_t qi = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (qi >= nq) return;
    const uint64_t i = idx[qi];
    if (i >= n) {
        out[qi] = 0;
        return;
    }
    const uint64_t b = i / bs;
    const uint32_t j = (uint32_t)(i - b * bs);
    const leco_fix_block_param P = prm[b];
    const uint8_t *pay = st + P.payload_offset;
    const int l = (int)P.bits;
    uint32_t v;
    if (l >= 31) {
        v = (uint32_t)load_bits(pay, (uint64_t)j * 32, 32);
    } else {
        const int64_t pm = opmax_pm(P.theta0, P.theta1, j);
        if (l) {
            const uint64_t f = load_bits(pay, (uint64_t)j * l, l);
            const uint64_t mag = f & ((1ull << (l - 1)) - 1);
            v = (uint32_t)((f >> (l - 1)) & 1 ? (uint64_t)pm + mag : (uint64_t)pm - mag);
        } else {
            v = (uint32_t)pm;
        }
    }
    out[qi] = v;
}

// Param table of a stored column (its headers; length from n and bs).
__global__ __launch_bounds__(kThreads) void k_opmax_params(const uint8_t *__restrict__ st,
                                                           const uint64_t *__restrict__ offs, uint64_t n, uint32_t bs,
                                                           leco_fix_block_param *__restrict__ prm) {
    const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const uint64_t nb = (n + bs - 1) / bs;
    if (b >= nb) return;
    const OpmaxBlock k = opmax_block(st, offs, n, bs, b);
    leco_fix_block_param P;
    P.theta0 = k.t0;
    P.theta1 = k.t1;
    P.payload_offset = offs[b] + kOpmaxHdr;
    P.bits = (uint32_t)k.l;
    P.length = k.m;
    prm[b] = P;
}

// ---------------------------------------------------------------------------
// Fix-up lists: the positions of +1 flags (mul_add_diff_set) and -1 flags
// (mul_add_diff_set_minus) in ascending order, as int32 like the reference's
// std::vector<int>.  Count per tile -> scan -> ordered emission.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_fix_count(const int8_t *__restrict__ fix, uint64_t n,
                                                        uint64_t *__restrict__ work) {
    __shared__ uint32_t wsum[2][kThreads / kWave];
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t base = (uint64_t)blockIdx.x * kFixTile;
    uint32_t cp = 0, cm = 0;
    for (uint32_t r = tid; r < kFixTile; r += kThreads) {
        const uint64_t i = base + r;
        const int8_t f = i < n ? fix[i] : 0;
        cp += f > 0;
        cm += f < 0;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        cp += (uint32_t)__shfl_xor((int)cp, o, kWave);
        cm += (uint32_t)__shfl_xor((int)cm, o, kWave);
    }
    if (lane == 0) {
        wsum[0][wid] = cp;
        wsum[1][wid] = cm;
    }
    __syncthreads();
    if (tid == 0) {
        uint64_t a = 0, c = 0;
        for (int w = 0; w < kThreads / kWave; ++w) {
            a += wsum[0][w];
            c += wsum[1][w];
        }
        work[2 * blockIdx.x] = a;
        work[2 * blockIdx.x + 1] = c;
    }
}

// One workgroup: exclusive scan of the tile counts in place, totals to counts[0..1].
__global__ __launch_bounds__(kThreads) void k_fix_scan(uint64_t *__restrict__ work, uint64_t tiles,
                                                       uint64_t *__restrict__ counts) {
    __shared__ uint64_t wsum[2][kThreads / kWave];
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    uint64_t carry[2] = {0, 0};
    for (uint64_t c0 = 0; c0 < tiles; c0 += kThreads) {
        const uint64_t t = c0 + tid;
        uint64_t v[2], inc[2];
        for (int h = 0; h < 2; ++h) {
            v[h] = t < tiles ? work[2 * t + h] : 0;
            inc[h] = v[h];
            for (int o = 1; o < kWave; o <<= 1) {
                const uint64_t x = (uint64_t)__shfl_up((long long)inc[h], o, kWave);
                if (lane >= o) inc[h] += x;
            }
            if (lane == kWave - 1) wsum[h][wid] = inc[h];
        }
        __syncthreads();
        for (int h = 0; h < 2; ++h) {
            uint64_t before = 0, all = 0;
            for (int w = 0; w < kThreads / kWave; ++w) {
                before += w < wid ? wsum[h][w] : 0;
                all += wsum[h][w];
            }
            if (t < tiles) work[2 * t + h] = carry[h] + before + inc[h] - v[h];
            carry[h] += all;
        }
        __syncthreads();
    }
    if (tid == 0) {
        counts[0] = carry[0];
        counts[1] = carry[1];
    }
}

__global__ __launch_bounds__(kThreads) void k_fix_emit(const int8_t *__restrict__ fix, uint64_t n,
                                                       const uint64_t *__restrict__ work, int32_t *__restrict__ plus,
                                                       int32_t *__restrict__ minus, uint64_t cap) {
    __shared__ uint32_t wsum[2][kThreads / kWave];
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t base = (uint64_t)blockIdx.x * kFixTile;
    uint64_t at[2] = {work[2 * blockIdx.x], work[2 * blockIdx.x + 1]};
    const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
    for (uint32_t r0 = 0; r0 < kFixTile; r0 += kThreads) {
        const uint64_t i = base + r0 + tid;
        const int8_t f = i < n ? fix[i] : 0;
        const uint64_t bal[2] = {__ballot(f > 0), __ballot(f < 0)};
        if (lane == 0) {
            wsum[0][wid] = (uint32_t)__popcll(bal[0]);
            wsum[1][wid] = (uint32_t)__popcll(bal[1]);
        }
        __syncthreads();
        for (int h = 0; h < 2; ++h) {
            uint64_t before = 0, all = 0;
            for (int w = 0; w < kThreads / kWave; ++w) {
                before += w < wid ? wsum[h][w] : 0;
                all += wsum[h][w];
            }
            const bool mine = h == 0 ? f > 0 : f < 0;
            const uint64_t dst = at[h] + before + (uint64_t)__popcll(bal[h] & below);
            if (mine && dst < cap) (h == 0 ? plus : minus)[dst] = (int32_t)i;
            at[h] += all;
        }
        __syncthreads();
    }
}

// The harness's fix-up loops: values[p] += 1 for p in plus, -= 1 for p in
// minus (entries outside [0, n) are skipped).
__global__ __launch_bounds__(kThreads) void k_apply_lists(uint32_t *__restrict__ values, uint64_t n,
                                                          const int32_t *__restrict__ plus, uint64_t np,
                                                          const int32_t *__restrict__ minus, uint64_t nm) {
    const uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q < np) {
        const int32_t p = plus[q];
        if (p >= 0 && (uint64_t)p < n) atomicAdd(&values[p], 1u);
    }
    if (q < nm) {
        const int32_t p = minus[q];
        if (p >= 0 && (uint64_t)p < n) atomicAdd(&values[p], 0xFFFFFFFFu);
    }
}

}  // namespace

hipError_t launch_opmax_encode(const uint32_t *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                               leco_fix_block_param *prm, int8_t *fix, hipStream_t st) {
    const uint64_t nb = (n + bs - 1) / bs;
    const uint32_t per = opmax_unit_blocks(bs);
    const uint64_t units = (nb + per - 1) / per;
    hipError_t e = launch_fit(in, n, bs, 32, prm, offs, st);
    if (e != hipSuccess) return e;
    k_opmax_encode<<<dim3((unsigned)units), dim3(kThreads), 0, st>>>(in, n, bs, prm, offs, stream, fix);
    k_opmax_fixup<<<dim3((unsigned)((nb + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(prm, offs, nb);
    return hipGetLastError();
}

hipError_t launch_opmax_decode(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                               const int8_t *fix, uint32_t *out, hipStream_t st) {
    const uint64_t runs = opmax_runs(n, bs);
    const dim3 grid((unsigned)((runs + kThreads - 1) / kThreads));
    if (fix) k_opmax_decode<true><<<grid, dim3(kThreads), 0, st>>>(stream, offs, n, bs, runs, fix, out);
    else k_opmax_decode<false><<<grid, dim3(kThreads), 0, st>>>(stream, offs, n, bs, runs, nullptr, out);
    return hipGetLastError();
}

hipError_t launch_opmax_apply_exact(uint32_t *values, uint64_t n, uint32_t bs, const uint8_t *stream,
                                    const uint64_t *offs, const int32_t *plus, uint64_t np, const int32_t *minus,
                                    uint64_t nm, hipStream_t st) {
    const uint64_t q = np > nm ? np : nm;
    if (q == 0) return hipSuccess;
    k_opmax_apply_exact<<<dim3((unsigned)((q + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(
        values, n, bs, stream, offs, plus, np, minus, nm);
    return hipGetLastError();
}

hipError_t launch_opmax_lookup(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                               const uint64_t *idx, uint64_t nq, uint32_t *out, hipStream_t st) {
    k_opmax_lookup<<<dim3((unsigned)((nq + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(stream, prm, n, bs,
                                                                                                 idx, nq, out);
    return hipGetLastError();
}

hipError_t launch_opmax_params(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs,
                               leco_fix_block_param *prm, hipStream_t st) {
    const uint64_t nb = (n + bs - 1) / bs;
    k_opmax_params<<<dim3((unsigned)((nb + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(stream, offs, n, bs,
                                                                                                 prm);
    return hipGetLastError();
}

hipError_t launch_fix_lists(const int8_t *fix, uint64_t n, int32_t *plus, int32_t *minus, uint64_t cap,
                            uint64_t *counts, uint64_t *work, hipStream_t st) {
    const uint64_t tiles = (n + kFixTile - 1) / kFixTile;
    k_fix_count<<<dim3((unsigned)tiles), dim3(kThreads), 0, st>>>(fix, n, work);
    k_fix_scan<<<dim3(1), dim3(kThreads), 0, st>>>(work, tiles, counts);
    k_fix_emit<<<dim3((unsigned)tiles), dim3(kThreads), 0, st>>>(fix, n, work, plus, minus, cap);
    return hipGetLastError();
}

hipError_t launch_apply_lists(uint32_t *values, uint64_t n, const int32_t *plus, uint64_t np, const int32_t *minus,
                              uint64_t nm, hipStream_t st) {
    const uint64_t q = np > nm ? np : nm;
    if (q == 0) return hipSuccess;
    k_apply_lists<<<dim3((unsigned)((q + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(values, n, plus, np,
                                                                                               minus, nm);
    return hipGetLastError();
}

}  // namespace leco
