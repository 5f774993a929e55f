// gfx950 kernels of the LeCo fixed-partition codec (Codecset::Leco_int<T>,
// headers/piecewise_fix_integer_template.h:14-338 of the reference).
//
// Column layout in HBM: the blocks' byte strings back to back (block b at
// offs[b]), exactly the bytes Leco_int<T>::encodeArray8_int writes per block,
// plus a 32-byte leco_fix_block_param per block for point lookups.
//
// Encode = K1 fit (one lane per block) -> K2 residuals + bit width + offset
// (decoupled look-back) + pack, one workgroup per large block or one wave per
// small block, values kept in registers -> K3 fix-up of payload offsets.  A
// uint64 column of 4096-value blocks takes the persistent stream encoder
// instead of K2 (k_encode_stream, k_encode_deferred, k_encode_fixup_status).
// Decode = one workgroup per block; lookup = one lane per query.  The
// batched variants run the same per-block bodies over a device table of
// blocks from many columns.
#include "leco_kernels.h"

namespace leco {

LECO_LOOKBACK_HOOK_SETTER(set_lookback_hook_leco)

// ---------------------------------------------------------------------------
// K1 fit: lr_int_T<T>::caltheta (lr.h:66-85).  The sums over y and i*y are
// sequential FP64 chains in the reference, so each block's chain is replayed
// in index order by one lane: a workgroup takes 64 consecutive blocks of one
// column and its wave 0 owns one block per lane.  Tiles of 64 blocks x kFitTK
// values are loaded by all four waves (16 rows each, 64 consecutive values
// per instruction, no branches around the loads), converted and parked in
// LDS, and read back one block per lane; the next tile's loads are in flight
// while the current tile's chains run.  sum(i) and sum(i*i) are exact and
// formed in closed form.  Also clears the look-back status words (offs[b+1])
// and, in workgroup 0, the ticket counter (offs[0]).
// ---------------------------------------------------------------------------
template <typename T, class Src>
__device__ __forceinline__ void fit_group(const FitGroup &g, leco_fix_block_param *__restrict__ prm,
                                          uint64_t *__restrict__ offs, double (*tile)[kFitTK + 1]) {
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const T *y = (const T *)g.base;
    const uint64_t nb = (g.n + g.bs - 1) / g.bs;
    auto row_len = [&](uint64_t r) -> uint32_t {
        const uint64_t b = g.b0 + r;
        if (b >= nb) return 0;
        const uint64_t s = b * g.bs;
        return (uint32_t)(g.n - s < g.bs ? g.n - s : g.bs);
    };
    const uint32_t m = row_len(lane);                 // wave 0: this lane's block
    const uint32_t maxm = row_len(0);                 // the first row is never shorter than the others

    // this thread's rows: wave w loads rows 16w .. 16w+15, two halves each
    constexpr int kRowsPerWave = kFitRows / (kFitThreads / kWave);
    uint64_t reg[kFitPer];
    auto load = [&](uint32_t k0) {
#pragma unroll
        for (int it = 0; it < kFitPer; ++it) {
            const int r = wid * kRowsPerWave + it / (kFitTK / kWave);
            const uint32_t k = k0 + (uint32_t)((it % (kFitTK / kWave)) * kWave + lane);
            const uint32_t mr = row_len(r);
            const T *rp = y + (mr ? (g.b0 + r) * g.bs : 0);
            const uint32_t kk = mr ? (k < mr ? k : mr - 1) : 0;
            const uint64_t v = (uint64_t)rp[kk];
            reg[it] = k < mr ? v : 0;
        }
    };

    double sy = 0.0, sxy = 0.0;
    if (maxm) load(0);
    for (uint32_t k0 = 0; k0 < maxm; k0 += kFitTK) {
#pragma unroll
        for (int it = 0; it < kFitPer; ++it)
            tile[wid * kRowsPerWave + it / (kFitTK / kWave)][(it % (kFitTK / kWave)) * kWave + lane] =
                (double)reg[it];
        __syncthreads();
        if (k0 + kFitTK < maxm) load(k0 + kFitTK);  // in flight while wave 0 runs the chains
        if (wid == 0) {
            if (k0 + kFitTK <= m) {
#pragma unroll 16
                for (int c = 0; c < kFitTK; ++c) {
                    const double yi = tile[lane][c];
                    sy = sy + yi;
                    sxy = __builtin_fma((double)(k0 + c), yi, sxy);
                }
            } else if (k0 < m) {
                for (uint32_t c = 0; c < m - k0; ++c) {
                    const double yi = tile[lane][c];
                    sy = sy + yi;
                    sxy = __builtin_fma((double)(k0 + c), yi, sxy);
                }
            }
        }
        __syncthreads();
    }
    if (wid != 0 || !m) return;
    const uint64_t mm = m;
    const double md = (double)m;
    const double sx = (double)(mm * (mm - 1) / 2);
    const double sxx = (double)((mm - 1) * mm * (2 * mm - 1) / 6);
    const double ccc = __builtin_fma(sxy, md, -(sx * sy));
    const double xxx = __builtin_fma(sxx, md, -(sx * sx));
    double t0, t1;
    if (xxx == 0.0) {
        // one-value block: 0/0, and x86 leaves its default NaN in both thetas
        t0 = t1 = bits_to_double(kX86DefaultNaN);
    } else {
        t1 = ccc / xxx;
        t0 = __builtin_fma(-t1, sx, sy) / md;
    }
    const uint64_t gb = g.gb0 + lane;
    leco_fix_block_param *p = prm + gb;
    p->theta0 = t0;
    p->theta1 = t1;
    p->length = m;
    p->payload_offset = 0;  // the stream encoder's ticket counters start here
    offs[gb + 1] = 0;       // look-back status: not ready
}

template <class Src>
__global__ __launch_bounds__(kFitThreads) void k_fit(Src src, leco_fix_block_param *__restrict__ prm,
                                                     uint64_t *__restrict__ offs) {
    __shared__ double tile[kFitRows][kFitTK + 1];
    if (src.clears_ticket(blockIdx.x) && threadIdx.x == 0) offs[0] = 0;  // ticket
    const FitGroup g = src.group(blockIdx.x);
    if (g.width == 64) fit_group<uint64_t, Src>(g, prm, offs, tile);
    else fit_group<uint32_t, Src>(g, prm, offs, tile);
}

// ---------------------------------------------------------------------------
// K2 encode: piecewise_fix_integer_template.h:36-137.
//   pass 1: min/max int64 residual under the fitted line; theta0 += (max+min)/2
//   pass 2: largest sign/magnitude residual -> bit width l (lr.h:116-130)
//   offset: decoupled look-back over the preceding blocks' sizes
//   pack:   header [l][theta0][theta1] (or [l] + raw values when l == W), then
//           field k = (mag & (2^(l-1)-1)) | sign<<(l-1) LSB-first, then zero
//           padding (bit_write.h:257-323).  A chunk of values at a time is
//           OR-ed into an LDS image aligned like its HBM destination and
//           written with 16-byte stores; byte stores only at the block's
//           ragged ends.
// A block of more than kWaveBlock values is encoded by a whole workgroup;
// smaller blocks by one wave each, four blocks side by side per workgroup.
// Either way the block's values stay in registers across all passes.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Stream encoder pieces for full blocks of VPT*NT values (all register-cached,
// thread t holding values 2(t + k*NT) and 2(t + k*NT) + 1 in v[2k], v[2k+1])
// whose predictions stay within +-2^51 (near_range).
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ uint32_t near_pos(int tid, int k) {
    return 2 * ((uint32_t)tid + (uint32_t)(k >> 1) * NT) + (uint32_t)(k & 1);
}

// Passes 1 and 2 (piecewise_fix_integer_template.h:36-64) over the cached
// values: pass 1 predicts without the range branch, pass 2 too unless the
// recentred theta0 leaves the range.  Block-wide max / min are row
// reductions followed by LDS atomics on sh.amax / sh.amin / sh.amag (set by
// init_reductions() and ordered by barriers before the atomics); amax and
// amin are set again here once read, amag is left to the caller.  On return
// v[] holds the magnitudes, `signs` bit k the sign of v[k] (1: above the
// line), t0 the recentred theta0 and the return value the bit width l.
template <int NT, int VPT, int W>
__device__ __forceinline__ int near_passes(int tid, uint64_t (&v)[VPT], double &t0, double t1, uint32_t &signs,
                                           EncodeShared &sh) {
    constexpr uint32_t m = VPT * NT;
    const int lane = threadIdx.x % kWave;
    const auto max_i64 = [](uint64_t a, uint64_t c) { return (int64_t)a > (int64_t)c ? a : c; };
    const auto min_i64 = [](uint64_t a, uint64_t c) { return (int64_t)a < (int64_t)c ? a : c; };
    const auto max_u64 = [](uint64_t a, uint64_t c) { return a > c ? a : c; };

    int64_t mx = INT64_MIN, mn = INT64_MAX;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const int64_t d = (int64_t)(v[k] - (uint64_t)predict_near(t0, t1, near_pos<NT>(tid, k)));
        mx = d > mx ? d : mx;
        mn = d < mn ? d : mn;
    }
    mx = (int64_t)row_reduce_u64((uint64_t)mx, max_i64);
    mn = (int64_t)row_reduce_u64((uint64_t)mn, min_i64);
    if ((lane & 15) == 0) {
        __hip_atomic_fetch_max(&sh.amax, mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_min(&sh.amin, mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    mx = (int64_t)uniform_u64((uint64_t)sh.amax);  // uniform: the branches below stay scalar
    mn = (int64_t)uniform_u64((uint64_t)sh.amin);
    // :46 theta0 += (max + min) / 2 (t0 is finite here)
    t0 = t0 + (double)(int64_t)((uint64_t)mx + (uint64_t)mn) * 0.5;

    const bool near2 = near_range(t0, t1, m);
    uint64_t mg = 0;
    signs = 0;
    // positions recomputed (an opaque copy of tid): holding pass 1's 16
    // converted positions would take 32 VGPRs.  (`tid` itself is the caller's
    // opaque copy, so that nothing position-dependent leaves its loop either.)
    int tid2 = tid;
    asm volatile("" : "+v"(tid2));
    auto pass2 = [&](auto near) {
#pragma unroll
        for (int k = VPT - 1; k >= 0; --k) {  // signs built by shifts: bit k ends at position k
            const uint32_t j = near_pos<NT>(tid2, k);
            uint64_t mag;
            const bool pos =
                residual_sign<W>(v[k], decltype(near)::value ? predict_near(t0, t1, j) : predict(t0, t1, j), &mag);
            v[k] = mag;
            signs = (signs << 1) | (uint32_t)pos;
            mg = mag > mg ? mag : mg;
        }
    };
    if (near2) pass2(std::true_type{});
    else pass2(std::false_type{});
    mg = row_reduce_u64(mg, max_u64);
    if ((lane & 15) == 0) __hip_atomic_fetch_max(&sh.amag, mg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    if (tid == 0) {  // every thread has read amax / amin
        sh.amax = INT64_MIN;
        sh.amin = INT64_MAX;
    }
    mg = uniform_u64(sh.amag);
    int l = mg ? bitlen64(mg) + 1 : 0;
    return l > W ? W : l;
}

// Field k of a block with a width 1 <= l <= 32 (not raw): magnitude |
// sign << (l-1) (the magnitude is below 2^(l-1)).
template <int VPT>
__device__ __forceinline__ void near_fields(const uint64_t (&v)[VPT], uint32_t signs, int l, uint32_t (&f)[VPT]) {
#pragma unroll
    for (int k = 0; k < VPT; ++k) f[k] = (uint32_t)v[k] | (((signs >> k) & 1u) << (l - 1));
}

// The block's bytes with a width l <= 32 (not raw), ORed into a cleared LDS
// image at byte `lead`: header [l][theta0][theta1], then field j (near_fields)
// at bit 8*(lead + 17) + j*l, two 32-bit ORs per field (bit_write.h:257-323).
template <int NT, int VPT>
__device__ __forceinline__ void near_pack(int tid, const uint32_t (&f)[VPT], int l, double t0, double t1,
                                          uint32_t lead, uint32_t *img) {
    static_assert((uint64_t)VPT * NT * 32 / 8 + 17 + 2 * 16 <= kLdsBytes, "a block of 32-bit fields fits the image");
    if (tid < 17) {
        uint32_t byte;
        if (tid == 0) byte = (uint32_t)l;
        else if (tid < 9) byte = (uint32_t)(double_to_bits(t0) >> (8 * (tid - 1))) & 0xFF;
        else byte = (uint32_t)(double_to_bits(t1) >> (8 * (tid - 9))) & 0xFF;
        const uint32_t pos = lead + tid;
        __hip_atomic_fetch_or(&img[pos >> 2], byte << (8 * (pos & 3)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (l == 0) return;
    const uint32_t q0 = 8 * (lead + 17) + 2 * (uint32_t)tid * (uint32_t)l;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const uint32_t q = q0 + (uint32_t)((k >> 1) * 2 * NT + (k & 1)) * (uint32_t)l;
        const uint64_t s = (uint64_t)f[k] << (q & 31);
        __hip_atomic_fetch_or(&img[q >> 5], (uint32_t)s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_or(&img[(q >> 5) + 1], (uint32_t)(s >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

template <int NT>
__device__ __forceinline__ void clear_image(uint32_t *img) {
    uint4 *b4 = (uint4 *)img;
    for (int i = threadIdx.x; i < kLdsBytes / 16; i += NT) b4[i] = make_uint4(0, 0, 0, 0);
}

// Workgroup body (NT threads): a block of any size up to kMaxBlockSize; the
// values of the first CC pack chunks (VPT per thread) are register-cached,
// the rest re-read in the passes (mostly from L2: the block was just loaded).
template <int NT, int CC, typename T, int W>
__device__ __forceinline__ void encode_block_body(const T *__restrict__ y, uint32_t m, leco_fix_block_param *prm,
                                                  uint64_t b, uint64_t u, uint64_t *__restrict__ offs,
                                                  uint8_t *__restrict__ out, EncodeShared &sh) {
    constexpr int VPT = CC * (int)(kChunk / NT);
    static_assert(NT / kWave <= kEncodeMaxWaves, "the reduction slots hold one value per wave");
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    double t0 = prm[b].theta0;
    const double t1 = prm[b].theta1;

    uint64_t v[VPT];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const uint32_t j = tid + k * NT;
        v[k] = j < m ? (uint64_t)y[j] : 0;
    }

    // pass 1 (lanes past the block's end are masked by selects, not branches)
    int64_t mx = INT64_MIN, mn = INT64_MAX;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const uint32_t j = tid + k * NT;
        const int64_t d = (int64_t)(v[k] - (uint64_t)predict(t0, t1, j));
        mx = (j < m && d > mx) ? d : mx;
        mn = (j < m && d < mn) ? d : mn;
    }
    for (uint32_t j = tid + VPT * NT; j < m; j += NT) {
        const int64_t d = (int64_t)((uint64_t)y[j] - (uint64_t)predict(t0, t1, j));
        mx = d > mx ? d : mx;
        mn = d < mn ? d : mn;
    }
    mx = wave_max_i64(mx);
    mn = wave_min_i64(mn);
    if (lane == 0) { sh.rmax[wid] = mx; sh.rmin[wid] = mn; }
    __syncthreads();
    mx = sh.rmax[0];
    mn = sh.rmin[0];
#pragma unroll
    for (int w = 1; w < NT / kWave; ++w) {
        mx = sh.rmax[w] > mx ? sh.rmax[w] : mx;
        mn = sh.rmin[w] < mn ? sh.rmin[w] : mn;
    }
    // :46 theta0 += (max + min) / 2.  (int64 sum wraps; x * 0.5 is exact)
    if (!__builtin_isnan(t0)) t0 = t0 + (double)(int64_t)((uint64_t)mx + (uint64_t)mn) * 0.5;

    // pass 2: the register-cached values are replaced by their magnitudes
    // and signs, which the pack phase turns into fields without predicting
    // again (raw blocks re-read their values)
    uint64_t mg = 0, mag;
    uint32_t signs = 0;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
        const uint32_t j = tid + k * NT;
        const bool pos = residual_sign<W>(v[k], predict(t0, t1, j), &mag);
        v[k] = mag;
        signs |= (uint32_t)pos << k;
        mg = (j < m && mag > mg) ? mag : mg;
    }
    for (uint32_t j = tid + VPT * NT; j < m; j += NT) {
        residual_field<W>((uint64_t)y[j], predict(t0, t1, j), 0, &mag);
        mg = mag > mg ? mag : mg;
    }
    mg = wave_max_u64(mg);
    if (lane == 0) sh.rmag[wid] = mg;
    __syncthreads();

    // bit width and size (every thread, from the four wave maxima)
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) mg = sh.rmag[w] > mg ? sh.rmag[w] : mg;
    int lb = mg ? bitlen64(mg) + 1 : 0;
    if (lb > W) lb = W;
    const uint64_t size = block_bytes((uint32_t)lb, W, m);

    if (wid == 0) {
        const uint64_t excl = look_back(u, size, offs);
        if (lane == 0) {
            sh.off = excl;
            sh.end = excl + size;
            sh.l = lb;
            sh.t0 = lb == W ? 0.0 : t0;
            put_param(&prm[b], sh.t0, lb == W, (uint32_t)lb, true, excl + (lb == W ? 1 : 17));
        }
    }
    __syncthreads();
    const int l = sh.l;
    t0 = sh.t0;
    const bool raw = (l == W);
    const PackedBlock pb{sh.off, sh.end, double_to_bits(t0), double_to_bits(t1), m, l, raw ? 1u : 17u,
                         l == 0 ? 1 : (m + kChunk - 1) / kChunk};

    const uint64_t low = l > 0 ? (1ull << (l - 1)) - 1 : 0;
    auto cached = [&](int k, uint32_t i) -> uint64_t {
        if (raw) return (uint64_t)y[i];
        return (v[k] & low) | ((uint64_t)((signs >> k) & 1) << (l - 1));
    };
    auto fresh = [&](uint32_t i) -> uint64_t {
        if (raw) return (uint64_t)y[i];
        uint64_t mg2;
        return residual_field<W>((uint64_t)y[i], predict(t0, t1, i), l, &mg2);
    };
    auto sync = [] { __syncthreads(); };
    pack_chunk<NT, kChunk>(pb, 0, tid, sh.buf, out, sync,
                                       [&](int kk, uint32_t i) { return cached(kk, i); });
    if (CC > 1 && pb.nchunks > 1)
        pack_chunk<NT, kChunk>(pb, 1, tid, sh.buf, out, sync,
                                     [&](int kk, uint32_t i) { return cached((CC > 1 ? kChunk / NT : 0) + kk, i); });
    for (uint32_t c = CC; c < pb.nchunks; ++c)
        pack_chunk<NT, kChunk>(pb, c, tid, sh.buf, out, sync, [&](int, uint32_t i) { return fresh(i); });
}

// Persistent, software-pipelined encoder for a plain column of 4096-value
// blocks.  Each workgroup takes tickets in order and, per block: passes 1-2,
// publishes the block's size, ORs its bytes into a free LDS image (lead 0:
// the offset is not known yet), starts the loads of its next block, and only
// then resolves the previous block's look-back and stores that block's image
// at its offset (store_image_shifted).  The look-back's round trips thus
// overlap the next block's work instead of idling the workgroup.  Blocks
// that do not fit this path (not full, input not 2*sizeof(T)-aligned,
// predictions beyond +-2^51, raw or fields over 32 bits) first drain the
// pending block, then take the look-back at once (the general body with
// one cached chunk, or near passes plus chunked packing).  No block waits
// for a look-back while its own workgroup holds an unresolved lower ticket,
// so every wait ends (each ticket's size is published without waiting).
// Bit of leco_fix_block_param::bits marking a block whose bytes the stream
// encoder left to k_encode_deferred (its size, thetas, width and offset are
// final); k_encode_deferred clears it.
constexpr uint32_t kDeferred = 1u << 31;

// Width and recentred theta0 of a block the stream encoder does not cache
// (any length, any predictions): passes 1-2 of the general body
// (piecewise_fix_integer_template.h:36-64) over values read from memory,
// one workgroup reduction per pass.  Raw blocks (l == W) get theta0 = 0.
struct BlockWidth {
    double t0;
    int l;
};
template <typename T, int W>
__device__ __attribute__((noinline)) BlockWidth stream_block_width(const T *__restrict__ y, uint32_t m, double t0,
                                                                   double t1, EncodeShared &sh) {
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    int64_t mx = INT64_MIN, mn = INT64_MAX;
    for (uint32_t j = tid; j < m; j += kThreads) {
        const int64_t d = (int64_t)((uint64_t)y[j] - (uint64_t)predict(t0, t1, j));
        mx = d > mx ? d : mx;
        mn = d < mn ? d : mn;
    }
    mx = wave_max_i64(mx);
    mn = wave_min_i64(mn);
    if (lane == 0) {
        sh.rmax[wid] = mx;
        sh.rmin[wid] = mn;
    }
    __syncthreads();
    mx = sh.rmax[0];
    mn = sh.rmin[0];
#pragma unroll
    for (int w = 1; w < kThreads / kWave; ++w) {
        mx = sh.rmax[w] > mx ? sh.rmax[w] : mx;
        mn = sh.rmin[w] < mn ? sh.rmin[w] : mn;
    }
    // :46 theta0 += (max + min) / 2.  (int64 sum wraps; x * 0.5 is exact)
    if (!__builtin_isnan(t0)) t0 = t0 + (double)(int64_t)((uint64_t)mx + (uint64_t)mn) * 0.5;
    uint64_t mg = 0;
    for (uint32_t j = tid; j < m; j += kThreads) {
        uint64_t mag;
        residual_sign<W>((uint64_t)y[j], predict(t0, t1, j), &mag);
        mg = mag > mg ? mag : mg;
    }
    mg = wave_max_u64(mg);
    if (lane == 0) sh.rmag[wid] = mg;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kThreads / kWave; ++w) mg = sh.rmag[w] > mg ? sh.rmag[w] : mg;
    int l = mg ? bitlen64(mg) + 1 : 0;
    if (l > W) l = W;
    return {l == W ? 0.0 : t0, l};
}

// Persistent, software-pipelined encoder for a plain column of 4096-value
// blocks.  Each workgroup takes tickets in order (stream_counter); per block
// it runs passes
// 1-2, publishes the block's size, ORs its bytes into a free LDS image at
// lead 0 (the offset is not known yet), and starts the loads of its next
// block; the previous block's look-back, whose first status loads were
// issued at the start of the iteration, is resolved only then and that
// block's image stored at its offset (store_image_shifted).  The look-back
// round trips thus overlap the next block's work instead of idling the
// workgroup.  A block outside the cached path (not full, input not
// 2*sizeof(T)-aligned, predictions beyond +-2^51) gets its width from
// stream_block_width; such a block, and one whose fields exceed 32 bits or
// are raw values, is marked kDeferred: it takes part in the look-back like
// any other, and k_encode_deferred writes its bytes afterwards.  No block
// waits for a look-back while its workgroup holds an unpublished lower
// ticket (every size is published without waiting), so every wait ends.
constexpr int kStreamVPT = 2 * kChunk / kThreads;
constexpr uint32_t kStreamBlock = kStreamVPT * kThreads;
// Ticket counters: same-address device atomics complete about one per 12 ns
// (a single counter would cap the encoder near 80 blocks/us), and counters
// in different 128-byte lines proceed in parallel.  Counter c lives in
// prm[4c].payload_offset (zeroed by k_fit; the fix-up writes the real
// payload offsets from the look-back status words); workgroup w draws from
// counter w % C, and its k-th ticket is block k*C + w % C.
constexpr uint32_t kStreamCounters = 8;
__device__ __forceinline__ unsigned long long *stream_counter(leco_fix_block_param *prm, uint32_t c) {
    return (unsigned long long *)&prm[4 * c].payload_offset;
}
inline uint32_t stream_counters(uint64_t nb) {
    const uint64_t c = (nb + 3) / 4;
    return (uint32_t)(c < kStreamCounters ? c : kStreamCounters);
}

struct StreamShared {
    EncodeShared e;                            // e.buf: image 0
    alignas(16) uint32_t img1[kLdsBytes / 4];  // image 1
    uint64_t next_ticket, next_ticket2;
};

template <typename T, int W>
__global__ __launch_bounds__(kThreads, 4) void k_encode_stream(PlainRows rows, leco_fix_block_param *__restrict__ prm,
                                                               uint64_t *__restrict__ offs, uint8_t *__restrict__ out,
                                                               uint32_t nctr) {
    constexpr int VPT = kStreamVPT;
    constexpr uint32_t M = kStreamBlock;
    __shared__ StreamShared ss;
    EncodeShared &sh = ss.e;
    const int tid = threadIdx.x;
    const uint32_t myc = blockIdx.x % nctr;
    unsigned long long *ctr = stream_counter(prm, myc);
    auto take = [&]() -> uint64_t { return (uint64_t)atomicAdd(ctr, 1ull) * nctr + myc; };
    if (tid == 0) {
        ss.next_ticket = take();
        ss.next_ticket2 = take();
        sh.init_reductions();
    }
    clear_image<kThreads>(sh.buf);
    clear_image<kThreads>(ss.img1);
    __syncthreads();
    uint64_t cur = uniform_u64(ss.next_ticket);

    // the current block: unit, values (when `fast`), thetas
    const T *y = nullptr;
    uint32_t m = 0;
    uint64_t b = 0;
    bool fast = false;
    uint64_t v[VPT];
    double t0 = 0.0, t1 = 0.0;
    auto fetch = [&](uint64_t u) -> bool {
        int tl = tid;
        asm volatile("" : "+v"(tl));
        const EncodeUnit un = rows.unit(u);
        if (!un.count) return false;
        const uint64_t s = un.cb0 * un.bs;
        m = (uint32_t)(un.n - s < un.bs ? un.n - s : un.bs);
        y = (const T *)un.base + s;
        b = un.gb0;
        fast = m == M && ((uintptr_t)y & (2 * sizeof(T) - 1)) == 0;
        // thetas by vector loads, issued before the values: their wait then
        // does not hold up LDS waits (scalar loads share lgkmcnt with LDS)
        int z = 0;
        asm volatile("" : "+v"(z));
        const double2 th = *(const double2 *)&prm[b + (uint32_t)z];
        t0 = th.x;
        t1 = th.y;
        if (fast) {
#pragma unroll
            for (int k = 0; k < VPT; k += 2) {
                const T *p = y + near_pos<kThreads>(tl, k);
                if constexpr (sizeof(T) == 8) {
                    const ulonglong2 q = *(const ulonglong2 *)p;
                    v[k] = q.x;
                    v[k + 1] = q.y;
                } else {
                    const uint2 q = *(const uint2 *)p;
                    v[k] = q.x;
                    v[k + 1] = q.y;
                }
            }
        }
        return true;
    };

    // the pending block: size published, bytes in image p_img (or deferred)
    bool pend = false;
    uint64_t p_u = 0, p_b = 0, p_size = 0;
    int p_img = -1;
    uint64_t stw[kLookPer];  // the pending block's first look-back window
    auto image = [&](int i) -> uint32_t * { return i ? ss.img1 : sh.buf; };
    auto drain = [&] {
        const uint64_t off = uniform_u64(resolve_wg<kThreads, kLookPer>(p_u, p_size, offs, sh.lbs, sh.lbf, stw));
        if (p_img >= 0) {
            uint32_t *img = image(p_img);
            store_image_shifted<kThreads>(img, off, p_size, tid, out);
            __syncthreads();  // the image is read; clear it for its next block
            clear_image<kThreads>(img);
        }
        pend = false;
    };

    // Per iteration: the ticket two blocks ahead is taken first; the
    // current block's passes; the pending block's first status window is
    // loaded (fresh, and older than what follows: vmcnt retires in order);
    // the current block's size is published and, once its values are folded
    // into 32-bit fields, the next block's loads are issued; then the pack,
    // and the pending block's look-back and store while those loads fly.
    int ci = 0;  // the free image
    uint64_t nxt = uniform_u64(ss.next_ticket2);
    bool have = fetch(cur);
    while (have) {
        // an opaque copy of the thread index per block: position arithmetic
        // is redone per block instead of held in registers across the loop
        int tl = tid;
        asm volatile("" : "+v"(tl));
        uint64_t tk = 0;
        if (tid == 0) tk = take();  // the block after nxt; read back at the end of the iteration
        t0 = __longlong_as_double((long long)uniform_u64((uint64_t)__double_as_longlong(t0)));
        t1 = __longlong_as_double((long long)uniform_u64((uint64_t)__double_as_longlong(t1)));
        const uint64_t cu = cur, cb = b;
        const uint32_t cm = m;
        double tr = t0;
        const double t1c = t1;
        int l;
        bool packs = false;
        uint32_t signs = 0;
        const bool near = fast && near_range(t0, t1, M);
        if (near) {
            l = near_passes<kThreads, VPT, W>(tl, v, tr, t1, signs, sh);
            if (l == W) tr = 0.0;
        } else {
            const BlockWidth r = stream_block_width<T, W>(y, m, t0, t1, sh);
            l = __builtin_amdgcn_readfirstlane(r.l);
            tr = __longlong_as_double((long long)uniform_u64((uint64_t)__double_as_longlong(r.t0)));
        }
        if (pend) lookback_issue<kThreads, kLookPer>(p_u, offs, stw);
        const bool raw = (l == W);
        const uint64_t size = block_bytes((uint32_t)l, W, cm);
        publish_size(cu, size, offs);
        packs = near && !raw && l <= 32;
        if (tid == 0) {
            put_param(&prm[cb], tr, raw, (uint32_t)l | (packs ? 0u : kDeferred), false, 0);
        }
        uint32_t fld[VPT];
        if (packs && l > 0) near_fields<VPT>(v, signs, l, fld);
        cur = nxt;
        have = fetch(cur);  // the next block's loads fly during the pack and the drain
        if (packs) near_pack<kThreads, VPT>(tl, fld, l, tr, t1c, 0, image(ci));
        // the image ORs done before the drain's loads: a queued LDS atomic can
        // read its registers after a later load has refilled them
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        if (pend) drain();
        if (tid == 0) ss.next_ticket2 = tk;
        __syncthreads();
        if (tid == 0) sh.amag = 0;  // every thread has read it
        nxt = uniform_u64(ss.next_ticket2);
        pend = true;
        p_u = cu;
        p_b = cb;
        p_size = size;
        p_img = packs ? ci : -1;
        if (packs) ci ^= 1;
    }
    if (pend) {
        lookback_issue<kThreads, kLookPer>(p_u, offs, stw);
        drain();
    }
}

// The bytes of the blocks k_encode_stream marked kDeferred, at their final
// offsets (from the look-back status words, which end as inclusive
// prefixes): each workgroup scans kThreads param entries at a time and packs
// the marked ones chunk by chunk (fields recomputed from the values with
// the block's final theta0 and width), then clears the mark.
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_encode_deferred(PlainRows rows, leco_fix_block_param *__restrict__ prm,
                                                              const uint64_t *__restrict__ offs, uint8_t *__restrict__ out) {
    __shared__ EncodeShared sh;
    __shared__ uint64_t s_mask[kThreads / kWave];
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < rows.nb; base += (uint64_t)gridDim.x * kThreads) {
        const uint64_t bb = base + tid;
        const bool mark = bb < rows.nb && (prm[bb].bits & kDeferred);
        const uint64_t bal = __ballot(mark);
        if (lane == 0) s_mask[wid] = bal;
        __syncthreads();
        for (int w = 0; w < kThreads / kWave; ++w) {
            uint64_t mk = uniform_u64(s_mask[w]);
            while (mk) {
                const int bit = __ffsll((long long)mk) - 1;
                mk &= mk - 1;
                const uint64_t blk = base + (uint64_t)w * kWave + bit;
                const leco_fix_block_param p = prm[blk];
                const int l = (int)(p.bits & ~kDeferred);
                const bool raw = (l == W);
                const uint64_t s = blk * rows.bs;
                const uint32_t m = (uint32_t)(rows.n - s < rows.bs ? rows.n - s : rows.bs);
                const T *y = (const T *)rows.in + s;
                const uint32_t hdr = raw ? 1u : 17u;
                const uint64_t end = offs[blk + 1] & kStatusValue;  // the block's inclusive prefix
                const uint64_t off = end - block_bytes((uint32_t)l, W, m);
                const double t0 = p.theta0, t1 = p.theta1;
                const PackedBlock pb{off, end, double_to_bits(t0), double_to_bits(t1), m, l, hdr,
                                     l == 0 ? 1 : (m + kChunk - 1) / kChunk};
                auto sync = [] { __syncthreads(); };
                for (uint32_t c = 0; c < pb.nchunks; ++c)
                    pack_chunk<kThreads, kChunk>(pb, c, tid, sh.buf, out, sync, [&](int, uint32_t i) -> uint64_t {
                        if (raw) return (uint64_t)y[i];
                        uint64_t mg2;
                        return residual_field<W>((uint64_t)y[i], predict(t0, t1, i), l, &mg2);
                    });
                if (tid == 0) prm[blk].bits = (uint32_t)l;
            }
        }
        __syncthreads();  // s_mask is rewritten by the next round
    }
}

// Wave-encoded unit: `count` consecutive blocks of at most kWaveBlock values,
// wave w taking blocks w, w+4, w+8, ... of the unit.
//   phase A: each wave fits the sign/magnitude widths of its blocks (values
//            register-cached kWaveVPL per lane, the tail of a longer block
//            re-read) and parks l, thetas and byte size in LDS;
//   unit:    wave 0 scans the sizes, publishes the unit's total and does one
//            look-back over the unit status words;
//   phase B: each wave packs its blocks at their offsets, kWaveChunk values
//            at a time through its own kWaveLdsBytes of LDS (values re-read,
//            mostly from cache).
// Publishing one size per unit as soon as phase A ends keeps the look-back
// from waiting on blocks that a wave has not reached yet.
template <typename T, int W>
__device__ __forceinline__ void encode_unit_waves(const EncodeUnit &un, uint64_t u, leco_fix_block_param *prm,
                                                  uint64_t *__restrict__ offs, uint8_t *__restrict__ out,
                                                  EncodeShared &sh) {
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    constexpr uint32_t kWaves = kThreads / kWave;
    const T *col = (const T *)un.base;

    // phase A
    for (uint32_t k = wid; k < un.count; k += kWaves) {
        const uint64_t b = un.gb0 + k;
        const T *y = col + (un.cb0 + k) * un.bs;
        const uint64_t s = (un.cb0 + k) * un.bs;
        const uint32_t m = (uint32_t)(un.n - s < un.bs ? un.n - s : un.bs);
        double t0 = prm[b].theta0;
        const double t1 = prm[b].theta1;
        uint64_t v[kWaveVPL];
#pragma unroll
        for (int q = 0; q < kWaveVPL; ++q) {
            const uint32_t j = lane + q * kWave;
            v[q] = j < m ? (uint64_t)y[j] : 0;
        }
        int64_t mx = INT64_MIN, mn = INT64_MAX;
#pragma unroll
        for (int q = 0; q < kWaveVPL; ++q) {
            const uint32_t j = lane + q * kWave;
            const int64_t d = (int64_t)(v[q] - (uint64_t)predict(t0, t1, j));
            mx = (j < m && d > mx) ? d : mx;
            mn = (j < m && d < mn) ? d : mn;
        }
        for (uint32_t j = lane + kWaveChunk; j < m; j += kWave) {
            const int64_t d = (int64_t)((uint64_t)y[j] - (uint64_t)predict(t0, t1, j));
            mx = d > mx ? d : mx;
            mn = d < mn ? d : mn;
        }
        mx = wave_max_i64(mx);
        mn = wave_min_i64(mn);
        if (!__builtin_isnan(t0)) t0 = t0 + (double)(int64_t)((uint64_t)mx + (uint64_t)mn) * 0.5;
        uint64_t mg = 0, mag;
#pragma unroll
        for (int q = 0; q < kWaveVPL; ++q) {
            const uint32_t j = lane + q * kWave;
            residual_sign<W>(v[q], predict(t0, t1, j), &mag);
            mg = (j < m && mag > mg) ? mag : mg;
        }
        for (uint32_t j = lane + kWaveChunk; j < m; j += kWave) {
            residual_sign<W>((uint64_t)y[j], predict(t0, t1, j), &mag);
            mg = mag > mg ? mag : mg;
        }
        mg = wave_max_u64(mg);
        int l = mg ? bitlen64(mg) + 1 : 0;
        if (l > W) l = W;
        if (l == W) t0 = 0.0;
        if (lane == 0) {
            sh.ut0[k] = t0;
            sh.ut1[k] = l == W ? 0.0 : t1;
            sh.ul[k] = l;
            sh.usz[k] = block_bytes((uint32_t)l, W, m);
        }
    }
    __syncthreads();

    // unit offset: exclusive scan of the sizes (one per lane) + look-back
    if (wid == 0) {
        const uint64_t mine = (uint32_t)lane < un.count ? sh.usz[lane] : 0;
        const uint64_t inc = wave_inclusive_sum(mine);
        const uint64_t total = wave_last(inc);
        const uint64_t excl = look_back(u, total, offs);
        if ((uint32_t)lane < un.count) sh.usz[lane] = excl + inc - mine;  // now: the block's offset
    }
    __syncthreads();

    // phase B
    uint32_t *buf = sh.buf + wid * (kWaveLdsBytes / 4);
    for (uint32_t k = wid; k < un.count; k += kWaves) {
        const uint64_t b = un.gb0 + k;
        const T *y = col + (un.cb0 + k) * un.bs;
        const uint64_t s = (un.cb0 + k) * un.bs;
        const uint32_t m = (uint32_t)(un.n - s < un.bs ? un.n - s : un.bs);
        const int l = __builtin_amdgcn_readfirstlane(sh.ul[k]);
        const double t0 = sh.ut0[k], t1 = sh.ut1[k];
        const uint64_t off = sh.usz[k];
        const bool raw = (l == W);
        const uint32_t hdr = raw ? 1u : 17u;
        if (lane == 0) {
            put_param(&prm[b], t0, raw, (uint32_t)l, true, off + hdr);
        }
        const PackedBlock pb{off, off + block_bytes((uint32_t)l, W, m), double_to_bits(t0), double_to_bits(t1), m, l,
                             hdr, l == 0 ? 1 : (m + kWaveChunk - 1) / kWaveChunk};
        static_assert(kWaveBlock <= 2 * kWaveChunk, "a wave-encoded block packs in at most two chunks");
        auto sync = [] { wave_lds_sync(); };
        for (uint32_t c = 0; c < pb.nchunks; ++c) {
            uint64_t f[kWaveVPL];
#pragma unroll
            for (int q = 0; q < kWaveVPL; ++q) {
                const uint32_t i = c * kWaveChunk + lane + q * kWave;
                f[q] = i < m ? (uint64_t)y[i] : 0;
            }
            if (!raw && l > 0) {
#pragma unroll
                for (int q = 0; q < kWaveVPL; ++q) {
                    uint64_t mg2;
                    const uint32_t i = c * kWaveChunk + lane + q * kWave;
                    f[q] = residual_field<W>(f[q], predict(t0, t1, i), l, &mg2);
                }
            }
            pack_chunk<kWave, kWaveChunk>(pb, c, lane, buf, out, sync,
                                                [&](int q, uint32_t) -> uint64_t { return f[q]; });
        }
    }
}

// One workgroup per encode unit, in ticket order (the look-back needs every
// lower unit to be owned by a running workgroup): a unit is one block of
// more than kWaveBlock values (whole workgroup) or encode_unit_blocks()
// blocks of one column encoded one per wave.  The look-back runs over unit
// status words offs[u+1] (u < units <= blocks); the fix-up rebuilds the block
// offsets from the param table.
template <class Rows>
__global__ __launch_bounds__(kThreads, 4) void k_encode(Rows rows, leco_fix_block_param *__restrict__ prm,
                                                        uint64_t *__restrict__ offs, uint8_t *__restrict__ out) {
    __shared__ EncodeShared sh;
    __shared__ uint64_t s_u;
    if (threadIdx.x == 0) s_u = take_ticket(offs);  // offs[0] is the ticket
    __syncthreads();
    const uint64_t u = s_u;
    const EncodeUnit un = rows.unit(u);
    if (!un.count) return;
    if (un.wave) {
        if (un.width == 64) encode_unit_waves<uint64_t, 64>(un, u, prm, offs, out, sh);
        else encode_unit_waves<uint32_t, 32>(un, u, prm, offs, out, sh);
        return;
    }
    const uint64_t s = un.cb0 * un.bs;
    const uint32_t m = (uint32_t)(un.n - s < un.bs ? un.n - s : un.bs);
    if (un.width == 64)
        encode_block_body<kThreads, 2, uint64_t, 64>((const uint64_t *)un.base + s, m, prm, un.gb0, u, offs, out, sh);
    else
        encode_block_body<kThreads, 2, uint32_t, 32>((const uint32_t *)un.base + s, m, prm, un.gb0, u, offs, out, sh);
}

// A plain column of blocks of at most kWaveBlock values: wave-encoded units
// only (94 instead of 124 VGPRs without the workgroup body: 5 workgroups per
// CU instead of 4).
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_encode_waves(PlainRows rows, leco_fix_block_param *__restrict__ prm,
                                                           uint64_t *__restrict__ offs, uint8_t *__restrict__ out) {
    __shared__ EncodeShared sh;
    __shared__ uint64_t s_u;
    if (threadIdx.x == 0) s_u = take_ticket(offs);  // offs[0] is the ticket
    __syncthreads();
    const uint64_t u = s_u;
    const EncodeUnit un = rows.unit(u);
    if (!un.count) return;
    encode_unit_waves<T, W>(un, u, prm, offs, out, sh);
}

// A plain column of blocks of kWaveBlock < bs <= kLargeEncodeBlock values
// (other than the stream encoder's): the workgroup body alone (two cached
// chunks), without the wave-unit body that k_encode also carries for
// batches of mixed block sizes.
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_encode_block(PlainRows rows, leco_fix_block_param *__restrict__ prm,
                                                           uint64_t *__restrict__ offs, uint8_t *__restrict__ out) {
    __shared__ EncodeShared sh;
    __shared__ uint64_t s_u;
    if (threadIdx.x == 0) s_u = take_ticket(offs);  // offs[0] is the ticket
    __syncthreads();
    const uint64_t u = s_u;
    const EncodeUnit un = rows.unit(u);
    if (!un.count) return;
    const uint64_t s = un.cb0 * un.bs;
    const uint32_t m = (uint32_t)(un.n - s < un.bs ? un.n - s : un.bs);
    encode_block_body<kThreads, 2, T, W>((const T *)un.base + s, m, prm, un.gb0, u, offs, out, sh);
}

// A plain column of blocks of more than kLargeEncodeBlock values: the
// workgroup body caching one chunk (8 values per thread) instead of two.  At
// 68 instead of 124 VGPRs, 7 workgroups fit a CU instead of 4; blocks this
// large are re-read in the passes either way.
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_encode_large(PlainRows rows, leco_fix_block_param *__restrict__ prm,
                                                           uint64_t *__restrict__ offs, uint8_t *__restrict__ out) {
    __shared__ EncodeShared sh;
    __shared__ uint64_t s_u;
    if (threadIdx.x == 0) s_u = take_ticket(offs);  // offs[0] is the ticket
    __syncthreads();
    const uint64_t u = s_u;
    const EncodeUnit un = rows.unit(u);
    if (!un.count) return;
    const uint64_t s = un.cb0 * un.bs;
    const uint32_t m = (uint32_t)(un.n - s < un.bs ? un.n - s : un.bs);
    encode_block_body<kThreads, 1, T, W>((const T *)un.base + s, m, prm, un.gb0, u, offs, out, sh);
}

// K3: block end offsets offs[b+1] from the param table (the unit status
// words that shared offs during the encode are overwritten), and offs[0]
// (the ticket) back to 0.  The last block's lane also reads the ticket's
// error bit: if a look-back wait gave up, offs[nb] becomes the poison value.
__global__ __launch_bounds__(kThreads) void k_encode_fixup(const leco_fix_block_param *__restrict__ prm,
                                                           uint64_t *__restrict__ offs, uint64_t nb,
                                                           const uint32_t *__restrict__ widths, uint32_t width) {
    const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b >= nb) return;
    const uint32_t w = widths ? widths[b] : width;
    const leco_fix_block_param p = prm[b];
    uint64_t end = p.payload_offset - (p.bits >= w ? 1 : 17) + block_bytes(p.bits, w, p.length);
    if (b + 1 == nb) {
        if (offs[0] & kTicketError) end = LECO_FIX_OFFSETS_POISON;
        offs[0] = 0;
    }
    offs[b + 1] = end;
}

// The stream encoder's fix-up: the status words hold every block's
// inclusive prefix (its end offset); payload offsets follow from them and
// the block sizes (the ticket counters that used those fields are
// overwritten), then offs[b+1] = end and, as k_encode_fixup, offs[0] back to
// 0 and the poison if a look-back wait gave up.
__global__ __launch_bounds__(kThreads) void k_encode_fixup_status(leco_fix_block_param *__restrict__ prm,
                                                                  uint64_t *__restrict__ offs, uint64_t nb,
                                                                  uint32_t width) {
    const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b >= nb) return;
    const leco_fix_block_param p = prm[b];
    uint64_t end = offs[b + 1] & kStatusValue;
    prm[b].payload_offset = end - block_bytes(p.bits, width, p.length) + (p.bits >= width ? 1 : 17);
    if (b + 1 == nb) {
        if (offs[0] & kTicketError) end = LECO_FIX_OFFSETS_POISON;
        offs[0] = 0;
    }
    offs[b + 1] = end;
}

// ---------------------------------------------------------------------------
// K5 decode body: Leco_int<T>::decodeArray8 (:142-204) with read_all_bit_fix
// (bit_read.h:222-275).  A block whose bytes fit the LDS buffer is staged
// whole, header included, with aligned 16-byte loads as soon as its offsets
// are known, and its header is read from LDS; a larger one reads its header
// from global memory and stages its payload kChunk values at a time.  Each
// lane extracts its field, adds it to cvtt64(fma(i, theta1, theta0)) and
// stores.  Writes block values [jlo, jhi) to o[j].  A block with l == W-1 is
// decoded, not copied as raw (the reference's :171-177 treats it as raw and
// returns wrong values).
// ---------------------------------------------------------------------------
// Values [ja, jb) of a staged chunk that starts at value j0: field at LDS
// bit lbit + (j - j0) * lw, read as two aligned words and funnel-shifted
// ((w1 << 1) << (63 - sh) is w1 << (64 - sh) and also right for sh = 0).
// For 32-bit values a field (at most 32 bits) lies in two 32-bit words and
// only the low 32 bits of the prediction matter.  LDS bit offsets fit 32 bits.
template <typename T, bool RAW>
__device__ __forceinline__ void decode_run(const uint64_t *__restrict__ buf, uint32_t lbit, int lw, int l, double t0,
                                           double t1, uint32_t j0, uint32_t ja, uint32_t jb, T *__restrict__ o) {
    const int tid = threadIdx.x;
    auto one = [&](uint32_t j) {
        const uint32_t q = lbit + (j - j0) * (uint32_t)lw;
        T v;
        if constexpr (sizeof(T) == 4) {
            const uint32_t *b32 = (const uint32_t *)buf;
            const uint32_t w = __builtin_amdgcn_alignbit(b32[(q >> 5) + 1], b32[q >> 5], q & 31);
            if (RAW) {
                v = w;
            } else {
                const uint32_t pr = predict_lo32(t0, t1, j);
                const uint32_t mag = w & ((1u << (l - 1)) - 1);
                v = ((w >> (l - 1)) & 1) ? pr + mag : pr - mag;
            }
        } else {
            const uint32_t w = q >> 6;
            const int sh = (int)(q & 63);
            const uint64_t field = ((buf[w] >> sh) | ((buf[w + 1] << 1) << (63 - sh))) & low_mask(lw);
            v = RAW ? (T)field : (T)apply_field(predict(t0, t1, j), field, l);
        }
        // decoded values are written once and not read back here: streaming stores
        __builtin_nontemporal_store(v, &o[j]);
    };
    if (ja == j0 && jb == j0 + kChunk) {
#pragma unroll
        for (int kk = 0; kk < (int)(kChunk / kThreads); ++kk) one(j0 + tid + kk * kThreads);
    } else {
        for (uint32_t j = ja + tid; j < jb; j += kThreads) one(j);
    }
}

// Block bytes [off, end) (end: the next block's offset, an upper bound on
// its last byte).
template <typename T, int W>
__device__ __forceinline__ void decode_block_body(const uint8_t *__restrict__ st, uint64_t off, uint64_t end,
                                                  uint32_t m, uint32_t jlo, uint32_t jhi, T *__restrict__ o,
                                                  uint64_t *__restrict__ buf) {
    const int tid = threadIdx.x;
    const uint64_t b0 = off & ~15ull;
    const uint32_t nb16 = (uint32_t)((end - b0 + 15) / 16);
    // whole block: the staged bytes plus one word of field over-read fit the buffer
    if (end >= off && (uint64_t)nb16 * 16 + 16 <= (uint64_t)kLdsBytes) {
        for (uint32_t u = tid; u < nb16; u += kThreads)
            *(uint4 *)(buf + 2 * u) = *(const uint4 *)(st + b0 + 16ull * u);
        __syncthreads();
        const uint32_t hb = (uint32_t)(off - b0);
        const int l = ((const uint8_t *)buf)[hb];
        if (l == 127 || l == 126) {  // :152-170, single/double raw value blocks
            const uint32_t cnt = l == 127 ? 1 : 2;
            if ((uint32_t)tid < cnt && (uint32_t)tid >= jlo && (uint32_t)tid < jhi)
                o[tid] = (T)lds_u64_at(buf, hb + 1 + tid * (uint32_t)sizeof(T));
        } else {
            const bool raw = l >= W;
            const int lw = raw ? W : l;
            double t0 = 0, t1 = 0;
            if (!raw) {
                t0 = bits_to_double(lds_u64_at(buf, hb + 1));
                t1 = bits_to_double(lds_u64_at(buf, hb + 9));
            }
            if (lw == 0) {  // :190-195
                for (uint32_t j = jlo + tid; j < jhi; j += kThreads) o[j] = (T)predict(t0, t1, j);
            } else {
                const uint32_t lbit = 8 * (hb + (raw ? 1 : 17));
                for (uint32_t c = jlo / kChunk; c * kChunk < jhi; ++c) {
                    const uint32_t j0 = c * kChunk;
                    const uint32_t jn = m - j0 < kChunk ? m - j0 : kChunk;
                    const uint32_t ja = jlo > j0 ? jlo : j0;
                    const uint32_t jb = jhi < j0 + jn ? jhi : j0 + jn;
                    const uint32_t cbit = lbit + j0 * (uint32_t)lw;
                    if (raw) decode_run<T, true>(buf, cbit, lw, l, t0, t1, j0, ja, jb, o);
                    else decode_run<T, false>(buf, cbit, lw, l, t0, t1, j0, ja, jb, o);
                }
            }
        }
        __syncthreads();  // the buffer is reused by the caller's next block
        return;
    }
    const int l = st[off];
    if (l == 127 || l == 126) {  // :152-170, single/double raw value blocks
        const uint32_t cnt = l == 127 ? 1 : 2;
        if ((uint32_t)tid < cnt && (uint32_t)tid >= jlo && (uint32_t)tid < jhi)
            o[tid] = (T)load_u64_any(st + off + 1 + tid * sizeof(T));
        return;
    }
    const bool raw = l >= W;
    const int lw = raw ? W : l;
    double t0 = 0, t1 = 0;
    if (!raw) {
        t0 = bits_to_double(load_u64_any(st + off + 1));
        t1 = bits_to_double(load_u64_any(st + off + 9));
    }
    if (lw == 0) {  // :190-195
        for (uint32_t j = jlo + tid; j < jhi; j += kThreads) o[j] = (T)predict(t0, t1, j);
        return;
    }
    const uint64_t pay = off + (raw ? 1 : 17);
    const uint64_t cbytes = (uint64_t)(kChunk / 8) * lw;
    for (uint32_t c = jlo / kChunk; c * kChunk < jhi; ++c) {
        const uint32_t j0 = c * kChunk;
        const uint32_t jn = m - j0 < kChunk ? m - j0 : kChunk;
        const uint64_t g0 = pay + c * cbytes;
        const uint64_t g1 = g0 + ((uint64_t)jn * lw + 7) / 8;
        const uint64_t a0 = g0 & ~15ull;
        const uint32_t n16 = (uint32_t)((g1 - a0 + 15) / 16);
        for (uint32_t u = tid; u < n16; u += kThreads)
            *(uint4 *)(buf + 2 * u) = *(const uint4 *)(st + a0 + 16ull * u);
        __syncthreads();
        const uint32_t lbit = (uint32_t)(8 * (g0 - a0));
        const uint32_t ja = jlo > j0 ? jlo : j0;
        const uint32_t jb = jhi < j0 + jn ? jhi : j0 + jn;
        if (raw) decode_run<T, true>(buf, lbit, lw, l, t0, t1, j0, ja, jb, o);
        else decode_run<T, false>(buf, lbit, lw, l, t0, t1, j0, ja, jb, o);
        __syncthreads();
    }
}

// One step of decode_multi_waves: U positions per lane, pb + lane + 64u, of
// the block at tile position pk (header values uniform), up to b.
template <int U, int CODEC, typename T, int W>
__device__ __forceinline__ void decode_wave_step(const MultiShared &sm, uint32_t pb, uint32_t b, uint32_t pk,
                                                 uint32_t fb, uint32_t lw, uint32_t l, bool raw, double t0,
                                                 double t1, T *__restrict__ o) {
    const int lane = threadIdx.x % kWave;
    if constexpr (W == 32) {
        const uint32_t *b32 = (const uint32_t *)sm.buf;
        uint32_t x0[U], x1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            uint32_t p = pb + lane + u * kWave;
            p = p < b ? p : b - 1;
            const uint32_t q = fb + (p - pk) * lw;
            x0[u] = b32[q >> 5];
            x1[u] = b32[(q >> 5) + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t p = pb + lane + u * kWave;
            const uint32_t j = p - pk;
            const uint32_t q = fb + j * lw;
            const uint32_t w = __builtin_amdgcn_alignbit(x1[u], x0[u], q & 31);
            uint32_t val;
            if constexpr (CODEC == kCodecFor) {
                val = raw ? w : (uint32_t)double_to_bits(t0) + (w & ((1u << l) - 1));
            } else {
                const uint32_t pr = predict_lo32(t0, t1, j);
                const uint32_t lc = l > 0 ? l - 1 : 0;
                const uint32_t mag = w & ((1u << lc) - 1);
                const uint32_t v = ((w >> lc) & 1) ? pr + mag : pr - mag;
                val = raw ? w : (l ? v : pr);
            }
            if (p < b) __builtin_nontemporal_store((T)val, &o[p]);
        }
    } else {
        uint64_t w0[U], w1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            uint32_t p = pb + lane + u * kWave;
            p = p < b ? p : b - 1;
            const uint32_t q = fb + (p - pk) * lw;
            w0[u] = sm.buf[q >> 6];
            w1[u] = sm.buf[(q >> 6) + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t p = pb + lane + u * kWave;
            const uint32_t j = p - pk;
            const int sh = (int)((fb + j * lw) & 63);
            const uint64_t field = ((w0[u] >> sh) | ((w1[u] << 1) << (63 - sh))) & low_mask((int)lw);
            uint64_t val;
            if constexpr (CODEC == kCodecFor) {
                val = raw ? field : double_to_bits(t0) + field;
            } else {
                const int64_t pr = predict(t0, t1, j);
                const int lc = l > 0 ? (int)l : 1;
                const uint64_t mag = field & ((1ull << (lc - 1)) - 1);
                const uint64_t v = ((field >> (lc - 1)) & 1) ? (uint64_t)pr + mag : (uint64_t)pr - mag;
                val = raw ? field : (l ? v : (uint64_t)pr);
            }
            if (p < b) __builtin_nontemporal_store((T)val, &o[p]);
        }
    }
}

// The position loop of decode_multi_body for blocks of kWaveDecodeMin or more
// values, a multiple of 64: each wave takes whole blocks (block ks + wid, then
// every kThreads/kWave-th), so the block's header is wave-uniform (scalar
// registers) and each lane decodes positions lane, lane + 64, ... of it,
// kMultiUnroll at a time (LDS reads first, then the arithmetic).  Tile
// positions [p0, pend).  500M keys, against one position per thread with the
// header from the LDS tables: uint64 +4.7% at bs 256, +4.4% at 512, +2.7% at
// 1024; uint32 +9% at bs 256.  Sizes that are not a multiple of 64 (1000)
// gained nothing (uint64) or lost 4% (uint32) and keep the position loop.
template <int CODEC, typename T, int W>
__device__ __forceinline__ void decode_multi_waves(const MultiShared &sm, uint32_t ks, uint32_t ke, uint32_t bs,
                                                   uint32_t p0, uint32_t pend, T *__restrict__ o) {
    const int wid = threadIdx.x / kWave;
    for (uint32_t k = ks + wid; k < ke; k += kThreads / kWave) {
        const uint32_t lbk = __builtin_amdgcn_readfirstlane(sm.hlb[k - ks]);
        const double t0 = bits_to_double(uniform_u64(double_to_bits(sm.ht[k - ks].x)));
        const double t1 = bits_to_double(uniform_u64(double_to_bits(sm.ht[k - ks].y)));
        const uint32_t l = lbk & 127;
        const uint32_t fb = lbk >> 7;
        const bool raw = l >= (uint32_t)W;
        const uint32_t lw = raw ? (uint32_t)W : l;
        const uint32_t pk = k * bs;
        const uint32_t a = pk > p0 ? pk : p0;
        uint32_t b = pk + bs < pend ? pk + bs : pend;
        if (CODEC != kCodecFor && (l == 126 || l == 127)) {  // one- and two-value raw blocks
            const uint32_t e = pk + (128 - l);
            b = b < e ? b : e;
        }
        uint32_t pb = a;
        for (; pb + kMultiUnroll * kWave <= b; pb += kMultiUnroll * kWave)
            decode_wave_step<kMultiUnroll, CODEC, T, W>(sm, pb, b, pk, fb, lw, l, raw, t0, t1, o);
        // the rest of the block (fewer than kMultiUnroll*64 positions): two or one
        // positions per lane, so that no step runs with idle lanes when the rest is
        // a multiple of 64
        if (pb + 2 * kWave <= b) {
            decode_wave_step<2, CODEC, T, W>(sm, pb, b, pk, fb, lw, l, raw, t0, t1, o);
            pb += 2 * kWave;
        }
        for (; pb < b; pb += kWave) decode_wave_step<1, CODEC, T, W>(sm, pb, b, pk, fb, lw, l, raw, t0, t1, o);
    }
}

// Multi-block staged decode for small blocks: blocks [b0, b0 + kb) of one
// column (kb * bs <= kMultiPositions values) with their offsets.  The bytes
// are staged into LDS with aligned 16-byte loads, as many whole blocks at a
// time as fit in kMultiLds bytes (all of them unless the data is nearly
// incompressible); one thread per block then parses its header into small
// tables (width | first-field bit, thetas), and every position is decoded
// from its block's table entry and field.  Tile positions p in [plo, phi)
// are decoded to o[p]; the others are skipped.  Widths 127 and 126 are the
// reference decoder's one- and two-value raw blocks (:152-170): only their
// first one or two positions are written.  CODEC kCodecFor decodes FOR_int
// blocks instead (FOR_integer_template.h:79-108): header [l][base][max],
// value base + field (raw when l >= W, every position written).
template <int CODEC, typename T, int W>
__device__ __forceinline__ void decode_multi_body(const uint8_t *__restrict__ st, const uint64_t *__restrict__ offs,
                                                  uint64_t n, uint32_t bs, uint64_t b0, uint32_t kb, uint32_t plo,
                                                  uint32_t phi, T *__restrict__ o, MultiShared &sm) {
    const int tid = threadIdx.x;
    const uint64_t s0 = b0 * bs;
    const uint64_t nb = (n + bs - 1) / bs;
    if (b0 + kb > nb) kb = (uint32_t)(nb - b0);
    uint32_t cnt = (uint32_t)(n - s0 < (uint64_t)kb * bs ? n - s0 : (uint64_t)kb * bs);
    if (phi < cnt) cnt = phi;
    for (uint32_t k = tid; k <= kb; k += kThreads) sm.offs[k] = offs[b0 + k];
    __syncthreads();
    const bool pow2 = (bs & (bs - 1)) == 0;
    const int lsh = __ffs((int)bs) - 1;
    // p / bs for the other sizes: umulhi(p, ceil(2^32 / bs)) is exact while
    // p * (ceil(2^32 / bs) * bs - 2^32) < 2^32, and here p < kb * bs <= 2^18
    // with bs < 2^11 (the multi path takes blocks below kStagedMinBlock)
    const uint32_t inv = pow2 ? 0u : 0xFFFFFFFFu / bs + 1u;
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
        for (uint32_t k = ks + tid; k < ke; k += kThreads) {
            const uint32_t hb = (uint32_t)(sm.offs[k] - a0);
            const uint32_t l = bytes[hb];
            const bool raw = l >= (uint32_t)W;
            if constexpr (CODEC == kCodecFor) {
                sm.hlb[k - ks] = ((8 * (hb + (raw ? 1 : 1 + 2 * (W / 8)))) << 7) | (l & 127);
                const uint64_t base = raw ? 0 : lds_u64_at(sm.buf, hb + 1) & (W == 32 ? 0xFFFFFFFFull : ~0ull);
                sm.ht[k - ks] = make_double2(bits_to_double(base), 0.0);
            } else {
                sm.hlb[k - ks] = ((8 * (hb + (raw ? 1 : 17))) << 7) | l;
                if (!raw) {
                    sm.ht[k - ks] = make_double2(bits_to_double(lds_u64_at(sm.buf, hb + 1)),
                                                 bits_to_double(lds_u64_at(sm.buf, hb + 9)));
                }
            }
        }
        __syncthreads();
        const uint32_t p0 = ks * bs > plo ? ks * bs : plo;
        const uint32_t pend = ke * bs < cnt ? ke * bs : cnt;
        if (bs >= kWaveDecodeMin && bs % kWave == 0) {
            decode_multi_waves<CODEC, T, W>(sm, ks, ke, bs, p0, pend, o);
            __syncthreads();
            ks = ke;
            continue;
        }
        // kMultiUnroll positions per thread per step: table reads, then field
        // reads, then the arithmetic, so that the LDS latencies overlap.  The
        // field's LDS bit offset q fits 32 bits (LDS is 160 KB).
        for (uint32_t pb = p0; pb < pend; pb += kMultiUnroll * kThreads) {
            uint32_t lb[kMultiUnroll], jj[kMultiUnroll];
            double2 th[kMultiUnroll];
#pragma unroll
            for (int u = 0; u < kMultiUnroll; ++u) {
                uint32_t p = pb + tid + u * kThreads;
                p = p < pend ? p : pend - 1;
                const uint32_t k = pow2 ? p >> lsh : __umulhi(p, inv);
                jj[u] = p - k * bs;
                lb[u] = sm.hlb[k - ks];
                th[u] = sm.ht[k - ks];
            }
            if constexpr (W == 32) {
                // A field of at most 32 bits lies in two consecutive 32-bit
                // words: one funnel shift extracts it, and the value needs
                // only the low 32 bits of the prediction.
                const uint32_t *b32 = (const uint32_t *)sm.buf;
                uint32_t x0[kMultiUnroll], x1[kMultiUnroll];
#pragma unroll
                for (int u = 0; u < kMultiUnroll; ++u) {
                    const uint32_t l = lb[u] & 127;
                    const uint32_t q = (lb[u] >> 7) + jj[u] * (l >= 32 ? 32 : l);
                    x0[u] = b32[q >> 5];
                    x1[u] = b32[(q >> 5) + 1];
                }
#pragma unroll
                for (int u = 0; u < kMultiUnroll; ++u) {
                    const uint32_t p = pb + tid + u * kThreads;
                    const uint32_t j = jj[u];
                    const uint32_t l = lb[u] & 127;
                    const bool raw = l >= 32;
                    const uint32_t q = (lb[u] >> 7) + j * (raw ? 32 : l);
                    const uint32_t w = __builtin_amdgcn_alignbit(x1[u], x0[u], q & 31);
                    if constexpr (CODEC == kCodecFor) {
                        const uint32_t base = (uint32_t)double_to_bits(th[u].x);
                        const uint32_t val = raw ? w : base + (w & ((1u << l) - 1));
                        if (p < pend) __builtin_nontemporal_store((T)val, &o[p]);
                    } else {
                        const uint32_t pr = predict_lo32(th[u].x, th[u].y, j);
                        const uint32_t lc = l > 0 ? l - 1 : 0;
                        const uint32_t mag = w & ((1u << lc) - 1);
                        const uint32_t v = ((w >> lc) & 1) ? pr + mag : pr - mag;
                        const uint32_t val = raw ? w : (l ? v : pr);
                        const bool wr = p < pend && (l < 126 || j < 128 - l);
                        if (wr) __builtin_nontemporal_store((T)val, &o[p]);
                    }
                }
            } else {
                uint64_t w0[kMultiUnroll], w1[kMultiUnroll];
#pragma unroll
                for (int u = 0; u < kMultiUnroll; ++u) {
                    const int l = (int)(lb[u] & 127);
                    const uint32_t q = (lb[u] >> 7) + jj[u] * (uint32_t)(l >= W ? W : l);
                    w0[u] = sm.buf[q >> 6];
                    w1[u] = sm.buf[(q >> 6) + 1];
                }
#pragma unroll
                for (int u = 0; u < kMultiUnroll; ++u) {
                    const uint32_t p = pb + tid + u * kThreads;
                    const uint32_t j = jj[u];
                    const int l = (int)(lb[u] & 127);
                    const bool raw = l >= W;
                    const int lw = raw ? W : l;
                    const int sh = (int)(((lb[u] >> 7) + j * (uint32_t)lw) & 63);
                    const uint64_t field = ((w0[u] >> sh) | ((w1[u] << 1) << (63 - sh))) & low_mask(lw);
                    if constexpr (CODEC == kCodecFor) {
                        const uint64_t val = raw ? field : double_to_bits(th[u].x) + field;
                        if (p < pend) __builtin_nontemporal_store((T)val, &o[p]);
                    } else {
                        const int64_t pr = predict(th[u].x, th[u].y, j);
                        const int lc = l > 0 ? l : 1;
                        const uint64_t mag = field & ((1ull << (lc - 1)) - 1);
                        const uint64_t v = ((field >> (lc - 1)) & 1) ? (uint64_t)pr + mag : (uint64_t)pr - mag;
                        const uint64_t val = raw ? field : (l ? v : (uint64_t)pr);
                        const bool wr = p < pend && (l < 126 || j < (uint32_t)(128 - l));
                        if (wr) __builtin_nontemporal_store((T)val, &o[p]);
                    }
                }
            }
        }
        __syncthreads();
        ks = ke;
    }
}

// One workgroup per tile of kb blocks of one column, tiles starting at block
// b_first; values [lo, hi) go to out[global - lo].
template <int CODEC, typename T, int W>
__global__ __launch_bounds__(kThreads) void k_decode_multi(const uint8_t *__restrict__ st,
                                                           const uint64_t *__restrict__ offs, uint64_t n,
                                                           uint32_t bs, uint32_t kb, uint64_t b_first, uint64_t lo,
                                                           uint64_t hi, T *__restrict__ out) {
    __shared__ MultiShared sm;
    const uint64_t b0 = b_first + (uint64_t)blockIdx.x * kb;
    const uint64_t s0 = b0 * bs;
    const uint32_t plo = lo > s0 ? (uint32_t)(lo - s0) : 0;
    const uint32_t phi = hi - s0 < (uint64_t)kb * bs ? (uint32_t)(hi - s0) : kb * bs;
    decode_multi_body<CODEC, T, W>(st, offs, n, bs, b0, kb, plo, phi, out - lo + s0, sm);
}

// One workgroup per block of one column; values [lo, hi) go to out[global - lo].
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_decode(const uint8_t *__restrict__ st,
                                                     const uint64_t *__restrict__ offs, uint64_t n, uint32_t bs,
                                                     uint64_t b_first, uint64_t lo, uint64_t hi,
                                                     T *__restrict__ out) {
    __shared__ __align__(16) uint64_t buf[kLdsBytes / 8];
    const uint64_t b = b_first + blockIdx.x;
    const uint64_t s = b * bs;
    const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
    const uint32_t jlo = lo > s ? (uint32_t)(lo - s) : 0;
    const uint32_t jhi = hi - s < m ? (uint32_t)(hi - s) : m;
    decode_block_body<T, W>(st, offs[b], offs[b + 1], m, jlo, jhi, out - lo + s, buf);
}

// ---------------------------------------------------------------------------
// K6 lookup: Leco_int<T>::randomdecodeArray8 (:305-337) with
// read_bit_fix_int_wo_round (bit_read.h:1170-1199).  One lane per query: the
// block's parameters come from the param table, the field from an unaligned
// bit window read as two aligned 8-byte words.
// ---------------------------------------------------------------------------
// Value at position j of the block described by p, branch-free over the
// block kinds (lanes of one wave may sit in blocks of different widths):
// raw (l >= W), constant (l == 0) and packed.
template <int W>
__device__ __forceinline__ uint64_t decode_at(const uint8_t *__restrict__ st, const leco_fix_block_param &p, uint32_t j) {
    const int l = (int)p.bits;
    const bool raw = l >= W;
    const int lw = raw ? W : l;
    // an l == 0 block has no payload: its window lies in the next block or the tail pad, and is masked away
    const uint64_t field = load_bits(st + p.payload_offset, (uint64_t)j * lw, lw);
    const int64_t pr = predict(p.theta0, p.theta1, j);
    const int lc = l > 0 ? l : 1;
    const uint64_t mag = field & ((1ull << (lc - 1)) - 1);
    const uint64_t packed = ((field >> (lc - 1)) & 1) ? (uint64_t)pr + mag : (uint64_t)pr - mag;
    return raw ? field : (l ? packed : (uint64_t)pr);
}

template <int W>
__device__ __forceinline__ uint64_t lookup_one(const uint8_t *__restrict__ st,
                                               const leco_fix_block_param *__restrict__ prm, uint32_t bs,
                                               uint64_t i) {
    uint64_t b;
    uint32_t j;
    if (i <= 0xFFFFFFFFull) {
        const uint32_t i32 = (uint32_t)i;
        b = i32 / bs;
        j = i32 - (uint32_t)b * bs;
    } else {
        b = i / bs;
        j = (uint32_t)(i - b * bs);
    }
    return decode_at<W>(st, prm[b], j);
}

// Value-parallel decode of column positions [v0, v0 + cnt) into o[0..cnt):
// lane k decodes position v0 + k from its block's param entry (shared by the
// lanes of that block) and a bit window read as two aligned words.  Unlike
// the staged block decode its cost does not depend on the block size, which
// makes it the better shape for small blocks.
template <typename T, int W>
__device__ __forceinline__ void decode_values(const uint8_t *__restrict__ st,
                                              const leco_fix_block_param *__restrict__ prm, uint32_t bs,
                                              uint64_t v0, uint32_t cnt, T *__restrict__ o) {
    const uint64_t b0 = v0 / bs;
    const uint32_t j0 = (uint32_t)(v0 - b0 * bs);
    const bool pow2 = (bs & (bs - 1)) == 0;
    const int sh = __ffs((int)bs) - 1;
#pragma unroll 4
    for (uint32_t k = threadIdx.x; k < cnt; k += kThreads) {
        const uint32_t jj = j0 + k;
        const uint32_t db = pow2 ? jj >> sh : jj / bs;
        o[k] = (T)decode_at<W>(st, prm[b0 + db], jj - db * bs);
    }
}

// One workgroup per kValTile positions of one column; positions [lo, hi).
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_decode_values(const uint8_t *__restrict__ st,
                                                            const leco_fix_block_param *__restrict__ prm,
                                                            uint32_t bs, uint64_t lo, uint64_t hi,
                                                            T *__restrict__ out) {
    const uint64_t v0 = lo + (uint64_t)blockIdx.x * kValTile;
    const uint32_t cnt = (uint32_t)(hi - v0 < kValTile ? hi - v0 : kValTile);
    decode_values<T, W>(st, prm, bs, v0, cnt, out + (v0 - lo));
}

template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_lookup(const uint8_t *__restrict__ st,
                                                     const leco_fix_block_param *__restrict__ prm, uint64_t n,
                                                     uint32_t bs, const uint64_t *__restrict__ idx, uint64_t nq,
                                                     T *__restrict__ out) {
    const uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= nq) return;
    const uint64_t i = idx[q];
    out[q] = i < n ? (T)lookup_one<W>(st, prm, bs, i) : (T)0;
}

// Grouped lookups: a batch of random accesses (Leco_int<T>::randomdecodeArray8,
// piecewise_fix_integer_template.h:305-337, once per query) with the queries
// first grouped by the range of 2^sh values they fall in -- a counting sort of
// (index, query) pairs -- and then looked up group by group.  A group's blocks
// and parameter entries are then read from HBM about once for all of its
// queries, instead of one 128-byte line (or two) per query; consecutive
// workgroups share a group, so its bytes stay in the L2s and the Infinity
// Cache while they run.  Queries outside [0, n) give 0, as k_lookup.
//
// The LDS atomics of the counting passes use registers of loaded indices:
// each thread loads a batch of indices, issues the batch's atomics, and waits
// for the LDS queue before the next batch's loads (a queued LDS atomic can
// read its registers after a later load has refilled them).
constexpr int kGroupBatch = 16;  // indices per thread loaded before their LDS atomics

__global__ __launch_bounds__(kThreads) void k_group_clear(uint32_t *__restrict__ totals, uint32_t bins) {
    for (uint32_t b = threadIdx.x; b < bins; b += kThreads) totals[b] = 0;
}

// Workgroup histogram of the groups of chunk `blockIdx.x` into hist (LDS).
__device__ __forceinline__ void group_histogram(const uint64_t *__restrict__ idx, uint64_t q0, uint64_t q1,
                                                uint64_t n, uint32_t sh, uint32_t *hist) {
    for (uint64_t qb = q0 + threadIdx.x; qb < q1; qb += (uint64_t)kThreads * kGroupBatch) {
        uint64_t v[kGroupBatch];
#pragma unroll
        for (int k = 0; k < kGroupBatch; ++k) {
            const uint64_t q = qb + (uint64_t)k * kThreads;
            v[k] = q < q1 ? idx[q] : n;
        }
#pragma unroll
        for (int k = 0; k < kGroupBatch; ++k)
            if (v[k] < n) atomicAdd(&hist[(uint32_t)(v[k] >> sh)], 1u);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    }
}

// Pass 1: group sizes (a histogram per workgroup, added into totals).
__global__ __launch_bounds__(kThreads) void k_group_count(const uint64_t *__restrict__ idx, uint64_t nq, uint64_t n,
                                                          uint32_t sh, uint32_t bins, uint32_t *__restrict__ totals) {
    __shared__ uint32_t hist[kGroupMaxBins];
    for (uint32_t b = threadIdx.x; b < bins; b += kThreads) hist[b] = 0;
    __syncthreads();
    const uint64_t q0 = (uint64_t)blockIdx.x * kGroupChunk;
    const uint64_t q1 = nq - q0 < kGroupChunk ? nq : q0 + kGroupChunk;
    group_histogram(idx, q0, q1, n, sh, hist);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < bins; b += kThreads)
        if (hist[b]) atomicAdd(&totals[b], hist[b]);
}

// Group starts: the exclusive scan of totals into cursor (one workgroup).
__global__ __launch_bounds__(kThreads) void k_group_starts(const uint32_t *__restrict__ totals, uint32_t bins,
                                                           uint32_t *__restrict__ cursor) {
    __shared__ uint32_t wsum[kThreads / kWave];
    constexpr uint32_t kPer = kGroupMaxBins / kThreads;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    uint32_t s[kPer], run = 0;
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t b = tid * kPer + r;
        s[r] = b < bins ? totals[b] : 0;
        run += s[r];
    }
    const uint32_t inc = wave_inclusive_sum(run);
    if (lane == kWave - 1) wsum[wid] = inc;
    __syncthreads();
    uint32_t o = inc - run;
    for (int w = 0; w < wid; ++w) o += wsum[w];
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
        const uint32_t b = tid * kPer + r;
        if (b < bins) cursor[b] = o;
        o += s[r];
    }
}

// Pass 2: each workgroup claims a range of every group it holds queries of
// (one global atomic per group), then writes its pairs (index << 31 | query)
// there, ranked by LDS atomics.  Out-of-range queries get 0 at once.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_group_scatter(const uint64_t *__restrict__ idx, uint64_t nq,
                                                            uint64_t n, uint32_t sh, uint32_t bins,
                                                            uint32_t *__restrict__ cursor,
                                                            uint64_t *__restrict__ pairs, T *__restrict__ out) {
    __shared__ uint32_t hist[kGroupMaxBins];
    __shared__ uint32_t base[kGroupMaxBins];
    for (uint32_t b = threadIdx.x; b < bins; b += kThreads) hist[b] = 0;
    __syncthreads();
    const uint64_t q0 = (uint64_t)blockIdx.x * kGroupChunk;
    const uint64_t q1 = nq - q0 < kGroupChunk ? nq : q0 + kGroupChunk;
    group_histogram(idx, q0, q1, n, sh, hist);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < bins; b += kThreads) {
        const uint32_t h = hist[b];
        base[b] = h ? atomicAdd(&cursor[b], h) : 0u;
        hist[b] = 0;
    }
    __syncthreads();
    for (uint64_t qb = q0 + threadIdx.x; qb < q1; qb += (uint64_t)kThreads * kGroupBatch) {
        uint64_t v[kGroupBatch];
#pragma unroll
        for (int k = 0; k < kGroupBatch; ++k) {
            const uint64_t q = qb + (uint64_t)k * kThreads;
            v[k] = q < q1 ? idx[q] : n;
        }
#pragma unroll
        for (int k = 0; k < kGroupBatch; ++k) {
            const uint64_t q = qb + (uint64_t)k * kThreads;
            if (q < q1) {
                if (v[k] < n) {
                    const uint32_t g = (uint32_t)(v[k] >> sh);
                    const uint32_t r = atomicAdd(&hist[g], 1u);
                    pairs[base[g] + r] = (v[k] << 31) | q;
                } else {
                    out[q] = (T)0;
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    }
}

// Pass 3: one pair per thread, in group order; the pair count is the last
// group's end (cursor after pass 2).  Workgroups are dispatched to the 8 XCDs
// in turn (workgroup i to XCD i mod 8), so workgroup i takes segment
// (i mod 8) * (segments / 8) + i / 8: each XCD walks its own eighth of the
// pairs in order, and its L2 holds only the few groups it is on (block reads
// from HBM drop from 4.3 to 1.1 GB for 64M queries on 500M keys).  The pairs
// stream past the caches (non-temporal); the results are plain stores, which
// the L2 merges where it can (non-temporal 8-byte stores took 3.1 ms instead
// of 2.1).
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_group_lookup(const uint8_t *__restrict__ st,
                                                           const leco_fix_block_param *__restrict__ prm,
                                                           uint32_t bs, const uint64_t *__restrict__ pairs,
                                                           const uint32_t *__restrict__ end, T *__restrict__ out) {
    const uint32_t per_xcd = gridDim.x / kXcds;  // the grid is a multiple of kXcds
    const uint64_t seg = (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    const uint64_t p = seg * kThreads + threadIdx.x;
    if (p >= *end) return;
    const uint64_t pr = __builtin_nontemporal_load(&pairs[p]);
    out[pr & kGroupQueryMask] = (T)lookup_one<W>(st, prm, bs, pr >> 31);
}

// Param table from block headers (for streams this library did not write).
__global__ __launch_bounds__(kThreads) void k_build_params(const uint8_t *__restrict__ st,
                                                           const uint64_t *__restrict__ offs, uint64_t n,
                                                           uint32_t bs, uint64_t nb, uint32_t w,
                                                           leco_fix_block_param *__restrict__ prm) {
    const uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b >= nb) return;
    const uint64_t off = offs[b], s = b * bs;
    const uint32_t l = st[off];
    leco_fix_block_param p;
    p.bits = l;
    p.length = (uint32_t)(n - s < bs ? n - s : bs);
    if (l >= w) {
        p.theta0 = p.theta1 = 0.0;
        p.payload_offset = off + 1;
    } else {
        p.theta0 = bits_to_double(load_u64_any(st + off + 1));
        p.theta1 = bits_to_double(load_u64_any(st + off + 9));
        p.payload_offset = off + 17;
    }
    prm[b] = p;
}

// ---------------------------------------------------------------------------
// Batched kernels: blocks of many columns (mixed widths and block sizes) as
// one global list, items in column order.  Encode writes one arena; decode is
// a persistent kernel striding over the items.
// ---------------------------------------------------------------------------
// Batched decode over the tile table.  A tile is one large block of a column
// (staged) or multi_blocks() consecutive small blocks decoded together;
// tile_col[t] names its column and tile_first[c] is column c's first tile.
// The table lists every staged tile before every multi tile, and each kind
// has a kernel of its own, so that each runs with only its own body's
// registers and LDS (one kernel for both ran the staged tiles at the multi
// body's occupancy).  Both stride over [t0, t1) when the grid is smaller.
// jlo is the first value of each block to decode, 0 from the launcher: as a
// kernel argument rather than a literal, because with a literal 0 the
// compiler schedules the chunk loop with about twice the registers (88
// instead of 42 VGPRs, 5 instead of 8 waves per SIMD).
__global__ __launch_bounds__(kThreads) void k_batch_decode_staged(const BatchColumn *__restrict__ cols,
                                                                  const uint64_t *__restrict__ first,
                                                                  const uint64_t *__restrict__ tile_first,
                                                                  const uint32_t *__restrict__ tile_col, uint64_t t0,
                                                                  uint64_t t1, const uint8_t *__restrict__ arena,
                                                                  const uint64_t *__restrict__ offs, uint32_t jlo) {
    __shared__ __align__(16) uint64_t buf[kLdsBytes / 8];
    for (uint64_t t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
        const uint32_t c = tile_col[t];
        const BatchColumn col = cols[c];
        const uint64_t local = t - tile_first[c];
        const uint64_t s = local * col.bs;
        const uint32_t m = (uint32_t)(col.n - s < col.bs ? col.n - s : col.bs);
        const uint64_t off = offs[first[c] + local], end = offs[first[c] + local + 1];
        if (col.width == 64)
            decode_block_body<uint64_t, 64>(arena, off, end, m, jlo, m, (uint64_t *)col.out + s, buf);
        else
            decode_block_body<uint32_t, 32>(arena, off, end, m, jlo, m, (uint32_t *)col.out + s, buf);
    }
}

__global__ __launch_bounds__(kThreads) void k_batch_decode_multi(const BatchColumn *__restrict__ cols,
                                                                 const uint64_t *__restrict__ first,
                                                                 const uint64_t *__restrict__ tile_first,
                                                                 const uint32_t *__restrict__ tile_col, uint64_t t0,
                                                                 uint64_t t1, const uint8_t *__restrict__ arena,
                                                                 const uint64_t *__restrict__ offs) {
    __shared__ MultiShared sm;
    for (uint64_t t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
        const uint32_t c = tile_col[t];
        const BatchColumn col = cols[c];
        const uint32_t kb = col.kb;
        const uint64_t b0 = (t - tile_first[c]) * kb;
        if (col.width == 64)
            decode_multi_body<kCodecLeco, uint64_t, 64>(arena, offs + first[c], col.n, col.bs, b0, kb, 0, kb * col.bs,
                                            (uint64_t *)col.out + b0 * col.bs, sm);
        else
            decode_multi_body<kCodecLeco, uint32_t, 32>(arena, offs + first[c], col.n, col.bs, b0, kb, 0, kb * col.bs,
                                            (uint32_t *)col.out + b0 * col.bs, sm);
    }
}

// tile_col[t] = ord[k] (k itself when ord is null) for the largest k with
// ord_first[k] <= t: the tiles are numbered in the column order ord lists.
__global__ __launch_bounds__(kThreads) void k_tile_columns(const uint64_t *__restrict__ ord_first,
                                                           const uint32_t *__restrict__ ord, uint32_t ncols,
                                                           uint64_t tiles, uint32_t *__restrict__ tile_col) {
    const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t < tiles) {
        const uint32_t k = find_column(ord_first, ncols, t);
        tile_col[t] = ord ? ord[k] : k;
    }
}

// Per-block width of a batch (one binary search per block, once).
__global__ __launch_bounds__(kThreads) void k_batch_blocks(const BatchColumn *__restrict__ cols,
                                                           const uint64_t *__restrict__ first, uint32_t ncols,
                                                           uint64_t total, uint32_t *__restrict__ widths) {
    const uint64_t it = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (it >= total) return;
    widths[it] = cols[find_column(first, ncols, it)].width;
}

// ---------------------------------------------------------------------------
// Served batch: one launch for a mixed list of point-lookup chunks and
// range-decode blocks from many requests (continuous batching).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_serve(const ServeColumn *__restrict__ cols,
                                                    const ServeItem *__restrict__ items, uint32_t nitems) {
    __shared__ __align__(16) uint64_t buf[kLdsBytes / 8];
    uint32_t it = blockIdx.x;
    if (it >= nitems) return;
    ServeItem w = items[it];
    for (;;) {
        const ServeColumn col = cols[w.column];
        // the next item loads while this one is served
        const uint32_t nx = it + gridDim.x;
        const bool more = nx < nitems;
        ServeItem wn{};
        if (more) wn = items[nx];
        if (w.kind == kServeLookup) {
            for (uint32_t q = threadIdx.x; q < w.count; q += kThreads) {
                const uint64_t i = w.idx[q];
                if (col.width == 64)
                    ((uint64_t *)w.out)[q] = i < col.n ? lookup_one<64>(col.stream, col.params, col.bs, i) : 0;
                else
                    ((uint32_t *)w.out)[q] =
                        i < col.n ? (uint32_t)lookup_one<32>(col.stream, col.params, col.bs, i) : 0u;
            }
        } else if (w.kind == kServeRangeValues) {
            if (col.width == 64)
                decode_values<uint64_t, 64>(col.stream, col.params, col.bs, w.lo, w.count, (uint64_t *)w.out);
            else
                decode_values<uint32_t, 32>(col.stream, col.params, col.bs, w.lo, w.count, (uint32_t *)w.out);
        } else {
            // one block of a range [lo, hi): w.block, output base w.out = value lo
            const uint64_t s = w.block * col.bs;
            const uint32_t m = (uint32_t)(col.n - s < col.bs ? col.n - s : col.bs);
            const uint32_t jlo = w.lo > s ? (uint32_t)(w.lo - s) : 0;
            const uint32_t jhi = w.hi - s < m ? (uint32_t)(w.hi - s) : m;
            if (col.width == 64)
                decode_block_body<uint64_t, 64>(col.stream, col.offsets[w.block], col.offsets[w.block + 1], m, jlo, jhi,
                                                (uint64_t *)w.out - w.lo + s, buf);
            else
                decode_block_body<uint32_t, 32>(col.stream, col.offsets[w.block], col.offsets[w.block + 1], m, jlo, jhi,
                                                (uint32_t *)w.out - w.lo + s, buf);
        }
        __syncthreads();
        if (!more) break;
        it = nx;
        w = wn;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

// Workgroups of the persistent stream encoder: as many as the device holds
// at once (per device, computed once).  More would only start after the
// resident ones drained the tickets.
template <typename T, int W>
uint32_t stream_workgroups() {
    static uint32_t cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (!cached[dev]) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_encode_stream<T, W>, kThreads, 0) != hipSuccess ||
            per <= 0)
            per = 1;
        cached[dev] = (uint32_t)(cus * per);
    }
    return cached[dev];
}

template <typename T>
hipError_t launch_encode(const T *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                         leco_fix_block_param *prm, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const uint64_t nb = (n + bs - 1) / bs;
    PlainRows rows{(const uint8_t *)in, n, bs, W, nb};
    PlainFit fit{(const uint8_t *)in, n, bs, W, 0};
    k_fit<PlainFit><<<dim3((unsigned)((nb + kFitRows - 1) / kFitRows)), dim3(kFitThreads), 0, st>>>(fit, prm, offs);
    if constexpr (W == 64) {  // the stream encoder exists for uint64 only
        if (bs == kStreamBlock && ((uintptr_t)in & (2 * sizeof(T) - 1)) == 0) {
            const unsigned wgs = (unsigned)std::min<uint64_t>(nb, stream_workgroups<T, W>());
            k_encode_stream<T, W><<<dim3(wgs), dim3(kThreads), 0, st>>>(rows, prm, offs, stream, stream_counters(nb));
            k_encode_deferred<T, W><<<dim3((unsigned)std::min<uint64_t>((nb + kThreads - 1) / kThreads, wgs)),
                                      dim3(kThreads), 0, st>>>(rows, prm, offs, stream);
            k_encode_fixup_status<<<dim3((unsigned)((nb + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(
                prm, offs, nb, W);
            return hipGetLastError();
        }
    }
    if (bs > kLargeEncodeBlock)  // one block per unit
        k_encode_large<T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(rows, prm, offs, stream);
    else if (bs <= kWaveBlock)
        k_encode_waves<T, W><<<dim3((unsigned)encode_units(nb, bs)), dim3(kThreads), 0, st>>>(rows, prm, offs, stream);
    else
        k_encode_block<T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(rows, prm, offs, stream);
    k_encode_fixup<<<dim3((unsigned)((nb + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(prm, offs, nb,
                                                                                                nullptr, W);
    return hipGetLastError();
}

hipError_t launch_fit(const void *in, uint64_t n, uint32_t bs, uint32_t width, leco_fix_block_param *prm,
                      uint64_t *offs, hipStream_t st) {
    const uint64_t nb = (n + bs - 1) / bs;
    PlainFit fit{(const uint8_t *)in, n, bs, width, 0};
    k_fit<PlainFit><<<dim3((unsigned)((nb + kFitRows - 1) / kFitRows)), dim3(kFitThreads), 0, st>>>(fit, prm, offs);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_decode(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, uint64_t lo,
                         uint64_t hi, T *out, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const uint64_t bf = lo / bs, bl = (hi - 1) / bs;
    if (!staged_decode(bs)) {
        const uint32_t kb = multi_blocks(bs, W);
        const uint64_t tiles = (bl - bf + 1 + kb - 1) / kb;
        k_decode_multi<kCodecLeco, T, W><<<dim3((unsigned)tiles), dim3(kThreads), 0, st>>>(stream, offs, n, bs, kb,
                                                                                             bf, lo, hi, out);
        return hipGetLastError();
    }
    k_decode<T, W><<<dim3((unsigned)(bl - bf + 1)), dim3(kThreads), 0, st>>>(stream, offs, n, bs, bf, lo, hi, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_for_decode_multi(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, T *out,
                                   hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const uint32_t kb = multi_blocks(bs, W);
    const uint64_t nb = (n + bs - 1) / bs;
    k_decode_multi<kCodecFor, T, W><<<dim3((unsigned)((nb + kb - 1) / kb)), dim3(kThreads), 0, st>>>(
        stream, offs, n, bs, kb, 0, 0, n, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_decode_values(const uint8_t *stream, const leco_fix_block_param *prm, uint32_t bs, uint64_t lo,
                                uint64_t hi, T *out, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    k_decode_values<T, W><<<dim3((unsigned)((hi - lo + kValTile - 1) / kValTile)), dim3(kThreads), 0, st>>>(
        stream, prm, bs, lo, hi, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_lookup(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                         const uint64_t *idx, uint64_t nq, T *out, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    k_lookup<T, W><<<dim3((unsigned)((nq + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(stream, prm, n, bs,
                                                                                                 idx, nq, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_lookup_grouped(const uint8_t *stream, const leco_fix_block_param *prm, uint64_t n, uint32_t bs,
                                 const uint64_t *idx, uint64_t nq, T *out, void *work, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const LookupGroups gr = lookup_groups(n);
    uint32_t *totals = (uint32_t *)work;
    uint32_t *cursor = totals + kGroupMaxBins;
    uint64_t *pairs = (uint64_t *)(cursor + kGroupMaxBins);
    for (uint64_t c = 0; c < nq; c += kGroupMaxQueries) {
        const uint64_t m = nq - c < kGroupMaxQueries ? nq - c : kGroupMaxQueries;
        const unsigned chunks = (unsigned)((m + kGroupChunk - 1) / kGroupChunk);
        k_group_clear<<<dim3(1), dim3(kThreads), 0, st>>>(totals, gr.bins);
        k_group_count<<<dim3(chunks), dim3(kThreads), 0, st>>>(idx + c, m, n, gr.sh, gr.bins, totals);
        k_group_starts<<<dim3(1), dim3(kThreads), 0, st>>>(totals, gr.bins, cursor);
        k_group_scatter<T><<<dim3(chunks), dim3(kThreads), 0, st>>>(idx + c, m, n, gr.sh, gr.bins, cursor, pairs,
                                                                    out + c);
        const uint64_t segs = (m + kThreads - 1) / kThreads;
        k_group_lookup<T, W><<<dim3((unsigned)((segs + kXcds - 1) / kXcds * kXcds)), dim3(kThreads), 0, st>>>(
            stream, prm, bs, pairs, cursor + gr.bins - 1, out + c);
    }
    return hipGetLastError();
}

hipError_t launch_build_params(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, uint32_t w,
                               leco_fix_block_param *prm, hipStream_t st) {
    const uint64_t nb = (n + bs - 1) / bs;
    k_build_params<<<dim3((unsigned)((nb + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(stream, offs, n, bs,
                                                                                                 nb, w, prm);
    return hipGetLastError();
}

hipError_t launch_batch_blocks(const BatchColumn *cols, const uint64_t *first, uint32_t ncols, uint64_t total,
                               uint32_t *widths, hipStream_t st) {
    k_batch_blocks<<<dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(cols, first, ncols,
                                                                                                    total, widths);
    return hipGetLastError();
}

hipError_t launch_batch_encode(const BatchColumn *cols, const uint64_t *first, const uint64_t *fit_first,
                               const uint64_t *unit_first, const uint32_t *unit_col, const uint32_t *widths,
                               uint32_t ncols, uint64_t total, uint64_t fit_groups, uint64_t units, uint8_t *arena,
                               uint64_t *offs, leco_fix_block_param *prm, hipStream_t st) {
    BatchRows rows{cols, first, unit_first, unit_col, units};
    BatchFit fit{cols, first, fit_first, ncols};
    k_fit<BatchFit><<<dim3((unsigned)fit_groups), dim3(kFitThreads), 0, st>>>(fit, prm, offs);
    k_encode<BatchRows><<<dim3((unsigned)units), dim3(kThreads), 0, st>>>(rows, prm, offs, arena);
    k_encode_fixup<<<dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(prm, offs, total,
                                                                                                    widths, 0);
    return hipGetLastError();
}

hipError_t launch_batch_decode(const BatchColumn *cols, const uint64_t *first, const uint64_t *tile_first,
                               const uint32_t *tile_col, uint64_t staged_tiles, uint64_t tiles, const uint8_t *arena,
                               const uint64_t *offs, uint32_t grid, hipStream_t st) {
    auto wgs = [&](uint64_t count) {
        const uint64_t cap = grid ? grid : 0x7FFFFFFFull;
        return (unsigned)(count < cap ? count : cap);
    };
    if (staged_tiles)
        k_batch_decode_staged<<<dim3(wgs(staged_tiles)), dim3(kThreads), 0, st>>>(cols, first, tile_first, tile_col, 0,
                                                                                  staged_tiles, arena, offs, 0u);
    if (tiles > staged_tiles)
        k_batch_decode_multi<<<dim3(wgs(tiles - staged_tiles)), dim3(kThreads), 0, st>>>(
            cols, first, tile_first, tile_col, staged_tiles, tiles, arena, offs);
    return hipGetLastError();
}

hipError_t launch_tile_columns(const uint64_t *ord_first, const uint32_t *ord, uint32_t ncols, uint64_t tiles,
                               uint32_t *tile_col, hipStream_t st) {
    k_tile_columns<<<dim3((unsigned)((tiles + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(ord_first, ord,
                                                                                                    ncols, tiles,
                                                                                                    tile_col);
    return hipGetLastError();
}

hipError_t launch_serve(const ServeColumn *cols, const ServeItem *items, uint32_t nitems, uint32_t grid,
                        hipStream_t st) {
    k_serve<<<dim3(grid), dim3(kThreads), 0, st>>>(cols, items, nitems);
    return hipGetLastError();
}

template hipError_t launch_encode<uint32_t>(const uint32_t *, uint64_t, uint32_t, uint8_t *, uint64_t *,
                                            leco_fix_block_param *, hipStream_t);
template hipError_t launch_encode<uint64_t>(const uint64_t *, uint64_t, uint32_t, uint8_t *, uint64_t *,
                                            leco_fix_block_param *, hipStream_t);
template hipError_t launch_decode<uint32_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, uint64_t,
                                            uint64_t, uint32_t *, hipStream_t);
template hipError_t launch_decode<uint64_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t, uint64_t,
                                            uint64_t, uint64_t *, hipStream_t);
template hipError_t launch_for_decode_multi<uint32_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t,
                                                      uint32_t *, hipStream_t);
template hipError_t launch_for_decode_multi<uint64_t>(const uint8_t *, const uint64_t *, uint64_t, uint32_t,
                                                      uint64_t *, hipStream_t);
template hipError_t launch_decode_values<uint32_t>(const uint8_t *, const leco_fix_block_param *, uint32_t, uint64_t,
                                                   uint64_t, uint32_t *, hipStream_t);
template hipError_t launch_decode_values<uint64_t>(const uint8_t *, const leco_fix_block_param *, uint32_t, uint64_t,
                                                   uint64_t, uint64_t *, hipStream_t);
template hipError_t launch_lookup<uint32_t>(const uint8_t *, const leco_fix_block_param *, uint64_t, uint32_t,
                                            const uint64_t *, uint64_t, uint32_t *, hipStream_t);
template hipError_t launch_lookup<uint64_t>(const uint8_t *, const leco_fix_block_param *, uint64_t, uint32_t,
                                            const uint64_t *, uint64_t, uint64_t *, hipStream_t);
template hipError_t launch_lookup_grouped<uint32_t>(const uint8_t *, const leco_fix_block_param *, uint64_t, uint32_t,
                                                    const uint64_t *, uint64_t, uint32_t *, void *, hipStream_t);
template hipError_t launch_lookup_grouped<uint64_t>(const uint8_t *, const leco_fix_block_param *, uint64_t, uint32_t,
                                                    const uint64_t *, uint64_t, uint64_t *, void *, hipStream_t);

}  // namespace leco
