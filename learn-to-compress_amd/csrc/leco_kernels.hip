// gfx950 kernels of the LeCo fixed-partition codec (Codecset::Leco_int<T>,
// headers/piecewise_fix_integer_template.h:14-338 of the reference).
//
// Column layout in HBM: the blocks' byte strings back to back (block b at
// offs[b]), exactly the bytes Leco_int<T>::encodeArray8_int writes per block,
// plus a 32-byte leco_fix_block_param per block for point lookups.
//
// Encode = fit (one lane per block) -> residual/bit width (one workgroup per
// block) -> exclusive scan of block sizes -> pack (one workgroup per block).
// Decode = one workgroup per block; lookup = one lane per query.
#include "leco_device.h"
#include "leco_internal.h"

namespace leco {

constexpr int kThreads = 256;           // workgroup of the per-block kernels (4 waves)
constexpr int kFitTK = 32;              // values per block row in one fit tile
constexpr uint32_t kChunk = 2048;       // values staged through LDS at a time
constexpr int kLdsBytes = kChunk * 8 + 64;

// ---------------------------------------------------------------------------
// K1 fit: lr_int_T<T>::caltheta (lr.h:66-85).  The sums over y and i*y are
// sequential FP64 chains in the reference, so each lane owns one block and
// replays its chain in index order; 64 blocks per wave.  A tile of 64 blocks x
// kFitTK values is loaded row by row (coalesced) into LDS and read back one
// block per lane.  sum(i) and sum(i*i) are exact and formed in closed form.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_fit(const T *__restrict__ in, uint64_t n, uint32_t bs, uint64_t nb,
                                            leco_fix_block_param *__restrict__ prm) {
    __shared__ T tile[64][kFitTK + 1];
    const int lane = threadIdx.x;
    const uint64_t b0 = (uint64_t)blockIdx.x * 64;
    const uint64_t b = b0 + lane;
    const uint64_t rows = nb - b0 < 64 ? nb - b0 : 64;
    uint32_t m = 0;
    if (b < nb) {
        uint64_t s = b * bs;
        m = (uint32_t)(n - s < bs ? n - s : bs);
    }
    double sy = 0.0, sxy = 0.0;
    for (uint32_t k0 = 0; k0 < bs; k0 += kFitTK) {
#pragma unroll 8
        for (int it = 0; it < 64 * kFitTK / 64; ++it) {
            const int r = it * (64 / kFitTK) + lane / kFitTK;
            const int c = lane % kFitTK;
            T v = 0;
            if ((uint64_t)r < rows && k0 + c < bs) {
                uint64_t idx = (b0 + r) * bs + k0 + c;
                if (idx < n) v = in[idx];
            }
            tile[r][c] = v;
        }
        __syncthreads();
        if (k0 < m) {
            const uint32_t kend = m - k0 < (uint32_t)kFitTK ? m - k0 : (uint32_t)kFitTK;
            for (uint32_t c = 0; c < kend; ++c) {
                const double yi = (double)tile[lane][c];
                const double xi = (double)(k0 + c);
                sy = sy + yi;
                sxy = __builtin_fma(xi, yi, sxy);
            }
        }
        __syncthreads();
    }
    if (b >= nb) return;
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
    prm[b].theta0 = t0;
    prm[b].theta1 = t1;
    prm[b].length = m;
}

// ---------------------------------------------------------------------------
// K2 residuals: piecewise_fix_integer_template.h:36-117.  Pass 1 takes the
// min/max int64 residual under the fitted line and centres theta0 on it;
// pass 2 takes the largest sign/magnitude residual, which sets the bit width l.
// Writes theta0 (centred), l and the block's byte size (into offs[b+1]).
// ---------------------------------------------------------------------------
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_resid(const T *__restrict__ in, uint64_t n, uint32_t bs,
                                                    leco_fix_block_param *__restrict__ prm,
                                                    uint64_t *__restrict__ offs) {
    __shared__ int64_t rmax[kThreads / kWave], rmin[kThreads / kWave];
    __shared__ uint64_t rmag[kThreads / kWave];
    const uint64_t b = blockIdx.x;
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid % kWave;
    const uint64_t s = b * bs;
    const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
    const T *y = in + s;
    double t0 = prm[b].theta0;
    const double t1 = prm[b].theta1;

    int64_t mx = INT64_MIN, mn = INT64_MAX;
    for (uint32_t j = tid; j < m; j += kThreads) {
        const int64_t d = (int64_t)((uint64_t)y[j] - (uint64_t)predict(t0, t1, j));
        mx = d > mx ? d : mx;
        mn = d < mn ? d : mn;
    }
    mx = wave_max_i64(mx);
    mn = wave_min_i64(mn);
    if (lane == 0) { rmax[wid] = mx; rmin[wid] = mn; }
    __syncthreads();
    mx = rmax[0];
    mn = rmin[0];
    for (int w = 1; w < kThreads / kWave; ++w) {
        mx = rmax[w] > mx ? rmax[w] : mx;
        mn = rmin[w] < mn ? rmin[w] : mn;
    }
    // :46 theta0 += (max + min) / 2.  (int64 sum wraps; x * 0.5 is exact)
    if (!__builtin_isnan(t0)) t0 = t0 + (double)(int64_t)((uint64_t)mx + (uint64_t)mn) * 0.5;

    uint64_t mg = 0;
    for (uint32_t j = tid; j < m; j += kThreads) {
        uint64_t mag;
        residual_field<W>((uint64_t)y[j], predict(t0, t1, j), 0, &mag);
        mg = mag > mg ? mag : mg;
    }
    mg = wave_max_u64(mg);
    if (lane == 0) rmag[wid] = mg;
    __syncthreads();
    if (tid == 0) {
        for (int w = 0; w < kThreads / kWave; ++w) mg = rmag[w] > mg ? rmag[w] : mg;
        int l = mg ? bitlen64(mg) + 1 : 0;
        if (l > W) l = W;
        // a raw block stores no theta; its table entry carries zeros
        prm[b].theta0 = l == W ? 0.0 : t0;
        if (l == W) prm[b].theta1 = 0.0;
        prm[b].bits = (uint32_t)l;
        offs[b + 1] = block_bytes((uint32_t)l, W, m);
    }
}

// ---------------------------------------------------------------------------
// K3 scan: offs[1..nb] hold block sizes; turns offs[0..nb] into the exclusive
// prefix (offs[0] = 0, offs[nb] = stream bytes).  One workgroup.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_scan(uint64_t *__restrict__ offs, uint64_t nb) {
    constexpr int PER = 8;
    __shared__ uint64_t wsum[1024 / kWave];
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
    uint64_t carry = 0;
    if (tid == 0) offs[0] = 0;
    for (uint64_t base = 0; base < nb; base += 1024 * PER) {
        uint64_t v[PER], loc = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint64_t i = base + (uint64_t)tid * PER + k;
            v[k] = i < nb ? offs[1 + i] : 0;
            loc += v[k];
        }
        uint64_t x = loc;
        for (int o = 1; o < kWave; o <<= 1) {
            uint64_t u = (uint64_t)__shfl_up((long long)x, o, kWave);
            if (lane >= o) x += u;
        }
        if (lane == kWave - 1) wsum[wid] = x;
        __syncthreads();
        if (wid == 0) {
            uint64_t w = lane < 1024 / kWave ? wsum[lane] : 0;
            for (int o = 1; o < 1024 / kWave; o <<= 1) {
                uint64_t u = (uint64_t)__shfl_up((long long)w, o, kWave);
                if (lane >= o) w += u;
            }
            if (lane < 1024 / kWave) wsum[lane] = w;
        }
        __syncthreads();
        uint64_t run = carry + (wid ? wsum[wid - 1] : 0) + x - loc;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            run += v[k];
            const uint64_t i = base + (uint64_t)tid * PER + k;
            if (i < nb) offs[1 + i] = run;
        }
        carry += wsum[1024 / kWave - 1];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// K4 pack: the block byte string (piecewise_fix_integer_template.h:118-137,
// write_delta_int_T bit_write.h:257-323).  Header [l][theta0][theta1] (or [l]
// + raw values when l == W), then field k = (mag & (2^(l-1)-1)) | sign<<(l-1)
// LSB-first at bit 17*8 + k*l, then zero padding to offs[b+1].  kChunk values
// at a time are OR-ed into an LDS image aligned like their HBM destination,
// then written with 16-byte stores (byte stores only at the block's ragged
// ends, which neighbouring workgroups share).
// ---------------------------------------------------------------------------
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_pack(const T *__restrict__ in, uint64_t n, uint32_t bs,
                                                   leco_fix_block_param *__restrict__ prm,
                                                   const uint64_t *__restrict__ offs, uint8_t *__restrict__ out) {
    __shared__ __align__(16) uint32_t buf[kLdsBytes / 4];
    const uint64_t b = blockIdx.x;
    const int tid = threadIdx.x;
    const uint64_t s = b * bs;
    const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
    const int l = (int)prm[b].bits;
    const double t0 = prm[b].theta0, t1 = prm[b].theta1;
    const uint64_t off = offs[b], end = offs[b + 1];
    const bool raw = (l == W);
    const uint32_t hdr = raw ? 1 : 17;
    if (tid == 0) prm[b].payload_offset = off + hdr;
    const T *y = in + s;
    const uint32_t nchunks = l == 0 ? 1 : (m + kChunk - 1) / kChunk;
    const uint64_t cbytes = (uint64_t)(kChunk / 8) * l;  // payload bytes per full chunk
    const uint64_t t0b = double_to_bits(t0), t1b = double_to_bits(t1);

    for (uint32_t c = 0; c < nchunks; ++c) {
        const uint64_t g0 = c == 0 ? off : off + hdr + c * cbytes;
        const uint64_t g1 = c + 1 == nchunks ? end : off + hdr + (c + 1) * cbytes;
        const uint64_t a0 = g0 & ~15ull;
        const uint32_t span = (uint32_t)(((g1 + 15) & ~15ull) - a0);
        const uint32_t lb = (uint32_t)(g0 - a0);
        for (uint32_t w = tid; w < span / 4; w += kThreads) buf[w] = 0;
        __syncthreads();
        if (c == 0 && (uint32_t)tid < hdr) {
            uint32_t byte;
            if (tid == 0) byte = (uint32_t)l;
            else if (tid < 9) byte = (uint32_t)(t0b >> (8 * (tid - 1))) & 0xFF;
            else byte = (uint32_t)(t1b >> (8 * (tid - 9))) & 0xFF;
            const uint32_t pos = lb + tid;
            atomicOr(&buf[pos >> 2], byte << (8 * (pos & 3)));
        }
        if (l > 0) {
            const uint64_t pbit = 8ull * (lb + (c == 0 ? hdr : 0));
            const uint32_t j0 = c * kChunk;
            const uint32_t jn = m - j0 < kChunk ? m - j0 : kChunk;
            for (uint32_t j = tid; j < jn; j += kThreads) {
                const uint32_t i = j0 + j;
                const uint64_t v = (uint64_t)y[i];
                uint64_t field, mag;
                if (raw) field = v;
                else field = residual_field<W>(v, predict(t0, t1, i), l, &mag);
                const uint64_t q = pbit + (uint64_t)j * l;
                const uint32_t dw = (uint32_t)(q >> 5);
                const int sh = (int)(q & 31);
                const uint32_t d0 = (uint32_t)(field << sh);
                const uint32_t d1 = (uint32_t)(sh ? field >> (32 - sh) : field >> 32);
                const uint32_t d2 = sh ? (uint32_t)(field >> (64 - sh)) : 0u;
                if (d0) atomicOr(&buf[dw], d0);
                if (sh + l > 32 && d1) atomicOr(&buf[dw + 1], d1);
                if (sh + l > 64 && d2) atomicOr(&buf[dw + 2], d2);
            }
        }
        __syncthreads();
        const uint8_t *bb = (const uint8_t *)buf;
        for (uint32_t u = tid; u < span / 16; u += kThreads) {
            const uint64_t ga = a0 + 16ull * u;
            if (ga >= g0 && ga + 16 <= g1) {
                *(uint4 *)(out + ga) = *(const uint4 *)(bb + 16 * u);
            } else {
                for (int k = 0; k < 16; ++k) {
                    const uint64_t gb = ga + k;
                    if (gb >= g0 && gb < g1) out[gb] = bb[16 * u + k];
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// K5 decode: Leco_int<T>::decodeArray8 (:142-204) with read_all_bit_fix
// (bit_read.h:222-275).  One workgroup per block: the block's payload is
// staged through LDS kChunk values at a time with aligned 16-byte loads, then
// each lane extracts its field, adds it to cvtt64(fma(i, theta1, theta0)) and
// stores.  Writes values [lo, hi) of the column to out[global - lo].
// A block with l == W-1 is decoded, not copied as raw (the reference's
// :171-177 treats it as raw and returns wrong values).
// ---------------------------------------------------------------------------
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_decode(const uint8_t *__restrict__ st,
                                                     const uint64_t *__restrict__ offs, uint64_t n,
                                                     uint32_t bs, uint64_t b_first, uint64_t lo, uint64_t hi,
                                                     T *__restrict__ out) {
    __shared__ __align__(16) uint64_t buf[kLdsBytes / 8];
    const uint64_t b = b_first + blockIdx.x;
    const int tid = threadIdx.x;
    const uint64_t s = b * bs;
    const uint32_t m = (uint32_t)(n - s < bs ? n - s : bs);
    const uint32_t jlo = lo > s ? (uint32_t)(lo - s) : 0;
    const uint32_t jhi = hi - s < m ? (uint32_t)(hi - s) : m;
    T *o = out - lo + s;  // o[j] is column value s + j
    const uint64_t off = offs[b];
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
        const uint64_t lbit = 8ull * (g0 - a0);
        const uint32_t ja = jlo > j0 ? jlo : j0;
        const uint32_t jb = jhi < j0 + jn ? jhi : j0 + jn;
        for (uint32_t j = ja + tid; j < jb; j += kThreads) {
            const uint64_t q = lbit + (uint64_t)(j - j0) * lw;
            const uint32_t w = (uint32_t)(q >> 6);
            const int sh = (int)(q & 63);
            const uint64_t v = sh ? (buf[w] >> sh) | (buf[w + 1] << (64 - sh)) : buf[w];
            const uint64_t field = v & low_mask(lw);
            o[j] = raw ? (T)field : (T)apply_field(predict(t0, t1, j), field, l);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// K6 lookup: Leco_int<T>::randomdecodeArray8 (:305-337) with
// read_bit_fix_int_wo_round (bit_read.h:1170-1199).  One lane per query:
// the block's parameters come from the param table, the field from an
// unaligned bit window read as two aligned 8-byte words.
// ---------------------------------------------------------------------------
template <typename T, int W>
__global__ __launch_bounds__(kThreads) void k_lookup(const uint8_t *__restrict__ st,
                                                     const leco_fix_block_param *__restrict__ prm, uint64_t n,
                                                     uint32_t bs, const uint64_t *__restrict__ idx, uint64_t nq,
                                                     T *__restrict__ out) {
    const uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= nq) return;
    const uint64_t i = idx[q];
    if (i >= n) { out[q] = 0; return; }
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
    const leco_fix_block_param p = prm[b];
    const int l = (int)p.bits;
    uint64_t v;
    if (l >= W) {
        v = load_bits(st + p.payload_offset, (uint64_t)j * W, W);
    } else {
        const int64_t pr = predict(p.theta0, p.theta1, j);
        v = l ? apply_field(pr, load_bits(st + p.payload_offset, (uint64_t)j * l, l), l) : (uint64_t)pr;
    }
    out[q] = (T)v;
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
// launchers
// ---------------------------------------------------------------------------
template <typename T>
hipError_t launch_encode(const T *in, uint64_t n, uint32_t bs, uint8_t *stream, uint64_t *offs,
                         leco_fix_block_param *prm, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const uint64_t nb = (n + bs - 1) / bs;
    k_fit<T><<<dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, st>>>(in, n, bs, nb, prm);
    k_resid<T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(in, n, bs, prm, offs);
    k_scan<<<dim3(1), dim3(1024), 0, st>>>(offs, nb);
    k_pack<T, W><<<dim3((unsigned)nb), dim3(kThreads), 0, st>>>(in, n, bs, prm, offs, stream);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_decode(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, uint64_t lo,
                         uint64_t hi, T *out, hipStream_t st) {
    constexpr int W = 8 * sizeof(T);
    const uint64_t bf = lo / bs, bl = (hi - 1) / bs;
    k_decode<T, W><<<dim3((unsigned)(bl - bf + 1)), dim3(kThreads), 0, st>>>(stream, offs, n, bs, bf, lo, hi, out);
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

hipError_t launch_build_params(const uint8_t *stream, const uint64_t *offs, uint64_t n, uint32_t bs, uint32_t w,
                               leco_fix_block_param *prm, hipStream_t st) {
    const uint64_t nb = (n + bs - 1) / bs;
    k_build_params<<<dim3((unsigned)((nb + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(stream, offs, n, bs,
                                                                                                 nb, w, prm);
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
template hipError_t launch_lookup<uint32_t>(const uint8_t *, const leco_fix_block_param *, uint64_t, uint32_t,
                                            const uint64_t *, uint64_t, uint32_t *, hipStream_t);
template hipError_t launch_lookup<uint64_t>(const uint8_t *, const leco_fix_block_param *, uint64_t, uint32_t,
                                            const uint64_t *, uint64_t, uint64_t *, hipStream_t);

}  // namespace leco
