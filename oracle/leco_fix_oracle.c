/*
 * leco_fix_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of LeCo's fixed-length-partition integer codec,
 * Codecset::Leco_int<T> (headers/piecewise_fix_integer_template.h:14-338),
 * for T = uint32_t and T = uint64_t.  It is the checker for the GPU library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product path (learn-to-compress_amd/) never links, calls or falls
 * back to it.
 *
 * Every function names the reference lines it restates.  The reference itself
 * cannot be compiled here (headers/common.h:12-19,34,65 needs Boost, GMP and
 * Eigen, none of which is installed), so this file is pinned by:
 *   - the independent pure-Python restatement tests/spec_leco.py (exact
 *     rational fma, a bit-at-a-time writer/reader) and the fixtures it wrote
 *     under tests/golden/;
 *   - closed forms: exact-linear blocks give theta1 = b, theta0 = a; constant
 *     blocks give theta1 = 0 and l = 0; partition and size formulas;
 *   - invariants: decode(encode(x)) == x, lookup(i) == decode[i],
 *     every residual magnitude < 2^(l-1).
 * Parity unpinned: the exact theta bits of the fit in the rounding regime
 * (partial sums of y or i*y above 2^53).  They follow the recipe below, which
 * is what GCC's -O3 -march=native build of lr.h produces on an FMA-capable x86
 * host (GNU mode contracts a*b+c into fma); no run of the reference confirms
 * it.  Also unpinned: the x86 default-NaN bits used for one-value blocks.
 *
 * FP recipe (build with -ffp-contract=off so only the explicit fma() fuse):
 *   fit      lr.h:66-85   sx += x; sy += y; sxx = fma(x,x,sxx); sxy = fma(x,y,sxy)
 *                         ccc = fma(sxy,m,-(sx*sy)); xxx = fma(sxx,m,-(sx*sx))
 *                         t1 = ccc/xxx; t0 = fma(-t1,sx,sy)/m
 *   predict  piecewise_fix_integer_template.h:40,54; bit_read.h:247,1187
 *                         p(i) = cvtt64(fma(t1,(double)i,t0))
 * cvtt64 is x86 CVTTSD2SI r64: truncate toward zero, NaN or out-of-range
 * gives INT64_MIN.
 */
#include <emmintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_API __attribute__((visibility("default")))

static const uint64_t X86_DEFAULT_NAN = 0xFFF8000000000000ull;

static double bits_to_double(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }

/* x86 CVTTSD2SI with a 64-bit destination. */
ORACLE_API int64_t leco_oracle_cvtt64(double x)
{
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0))
        return INT64_MIN;
    return (int64_t)x;
}

static inline uint64_t load_val(const void *y, int width, uint64_t i)
{
    return width == 32 ? (uint64_t)((const uint32_t *)y)[i] : ((const uint64_t *)y)[i];
}

static inline void store_val(void *y, int width, uint64_t i, uint64_t v)
{
    if (width == 32) ((uint32_t *)y)[i] = (uint32_t)v;
    else ((uint64_t *)y)[i] = v;
}

/* bits_int_T (lr.h:116-130): 64 - clzll(v), 0 for v == 0. */
static inline int bitlen64(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

static inline int64_t predict(double t0, double t1, uint64_t i)
{
    return leco_oracle_cvtt64(fma(t1, (double)i, t0));
}

/* lr_int_T<T>::caltheta, lr.h:66-85, under the FMA recipe above. */
ORACLE_API void leco_oracle_fit(const void *y, int width, uint32_t m, double *t0, double *t1)
{
    double sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (uint32_t i = 0; i < m; i++) {
        double xi = (double)i, yi = (double)load_val(y, width, i);
        sx = sx + xi;
        sy = sy + yi;
        sxx = fma(xi, xi, sxx);
        sxy = fma(xi, yi, sxy);
    }
    double md = (double)m;
    double ccc = fma(sxy, md, -(sx * sy));
    double xxx = fma(sxx, md, -(sx * sx));
    if (xxx == 0.0) {
        /* 0/0 for a one-value block: x86 returns its default NaN, and the
           later fnmadd/div/add steps carry it into theta0 unchanged. */
        *t0 = *t1 = bits_to_double(X86_DEFAULT_NAN);
        return;
    }
    double th1 = ccc / xxx;
    *t1 = th1;
    *t0 = fma(-th1, sx, sy) / md;
}

/* Payload size of a packed block, bit_write.h:263-321: the writer emits
   E+1 bytes (E = ceil(l*n/8)) and then 8 - E%8 zero bytes. */
ORACLE_API uint64_t leco_oracle_payload_bytes(int l, uint32_t n)
{
    uint64_t e = ((uint64_t)l * n + 7) / 8;
    return 8 * (e / 8) + 9;
}

ORACLE_API uint64_t leco_oracle_block_bytes(int l, int width, uint32_t n)
{
    if (l >= width) return 1 + (uint64_t)n * (width / 8);
    return 17 + (l ? leco_oracle_payload_bytes(l, n) : 0);
}

/* Largest block a column of n values at block size bs can need. */
ORACLE_API uint64_t leco_oracle_stream_bound(uint64_t n, uint32_t bs, int width)
{
    if (n == 0) return 0;
    uint64_t nb = (n + bs - 1) / bs, last = n - (nb - 1) * bs, tot = 0;
    uint64_t full = leco_oracle_block_bytes(width - 1, width, bs);
    uint64_t fraw = leco_oracle_block_bytes(width, width, bs);
    uint64_t lb = leco_oracle_block_bytes(width - 1, width, (uint32_t)last);
    uint64_t lraw = leco_oracle_block_bytes(width, width, (uint32_t)last);
    tot = (nb - 1) * (full > fraw ? full : fraw) + (lb > lraw ? lb : lraw);
    return tot;
}

/* Leco_int<T>::encodeArray8_int, piecewise_fix_integer_template.h:27-140,
   with write_delta_int_T (bit_write.h:257-323).  Returns bytes written. */
ORACLE_API uint64_t leco_oracle_encode_block(const void *y, int width, uint32_t m, uint8_t *out)
{
    double t0, t1;
    leco_oracle_fit(y, width, m, &t0, &t1);

    /* pass 1 (:36-45): int64 residuals under the first theta; wrapping */
    int64_t mx = INT64_MIN, mn = INT64_MAX;
    for (uint32_t i = 0; i < m; i++) {
        uint64_t yi = load_val(y, width, i);
        int64_t d = (int64_t)(yi - (uint64_t)predict(t0, t1, i));
        if (d > mx) mx = d;
        if (d < mn) mn = d;
    }
    /* :46 -- the int64 sum wraps; x/2. is exact, so fused or not is equal */
    if (!isnan(t0))
        t0 = t0 + (double)(int64_t)((uint64_t)mx + (uint64_t)mn) * 0.5;

    /* pass 2 (:48-72): sign/magnitude residuals; the comparison is signed
       int64 for uint32 and unsigned for uint64 (usual conversions) */
    uint64_t wmask = width == 32 ? 0xFFFFFFFFull : ~0ull;
    uint64_t maxmag = 0;
    for (uint32_t i = 0; i < m; i++) {
        uint64_t yi = load_val(y, width, i);
        int64_t p = predict(t0, t1, i);
        int pos = width == 64 ? (yi > (uint64_t)p) : ((int64_t)yi > p);
        uint64_t mag = (pos ? yi - (uint64_t)p : (uint64_t)p - yi) & wmask;
        if (mag > maxmag) maxmag = mag;
    }
    /* :108-117 */
    int l = maxmag ? bitlen64(maxmag) + 1 : 0;
    if (l > width) l = width;
    out[0] = (uint8_t)l;
    if (l == width) { /* :120-128 raw block */
        for (uint32_t i = 0; i < m; i++) {
            uint64_t v = load_val(y, width, i);
            memcpy(out + 1 + (uint64_t)i * (width / 8), &v, width / 8);
        }
        return 1 + (uint64_t)m * (width / 8);
    }
    memcpy(out + 1, &t0, 8);
    memcpy(out + 9, &t1, 8);
    if (!l) return 17;

    /* write_delta_int_T: field k = (mag & (2^(l-1)-1)) | sign << (l-1),
       packed LSB-first at bit k*l; the remaining bytes are zero. */
    uint64_t pbytes = leco_oracle_payload_bytes(l, m);
    uint8_t *pay = out + 17;
    memset(pay, 0, pbytes);
    unsigned __int128 acc = 0;
    int nacc = 0;
    uint64_t wr = 0;
    uint64_t low = ((uint64_t)1 << (l - 1)) - 1;
    for (uint32_t i = 0; i < m; i++) {
        uint64_t yi = load_val(y, width, i);
        int64_t p = predict(t0, t1, i);
        int pos = width == 64 ? (yi > (uint64_t)p) : ((int64_t)yi > p);
        uint64_t mag = (pos ? yi - (uint64_t)p : (uint64_t)p - yi) & wmask;
        uint64_t field = (mag & low) | ((uint64_t)pos << (l - 1));
        acc |= (unsigned __int128)field << nacc;
        nacc += l;
        while (nacc >= 8) {
            pay[wr++] = (uint8_t)acc;
            acc >>= 8;
            nacc -= 8;
        }
    }
    if (nacc > 0) pay[wr++] = (uint8_t)acc;
    return 17 + pbytes;
}

static inline uint64_t read_field(const uint8_t *pay, uint64_t bit, int l)
{
    uint64_t byte = bit >> 3;
    int sh = (int)(bit & 7);
    unsigned __int128 win = 0;
    int need = (sh + l + 7) / 8;
    for (int k = 0; k < need; k++)
        win |= (unsigned __int128)pay[byte + k] << (8 * k);
    win >>= sh;
    return l == 64 ? (uint64_t)win : (uint64_t)win & (((uint64_t)1 << l) - 1);
}

static inline uint64_t apply_residual(int64_t p, uint64_t field, int l)
{
    uint64_t mag = field & (((uint64_t)1 << (l - 1)) - 1);
    int pos = (int)((field >> (l - 1)) & 1);
    return pos ? (uint64_t)p + mag : (uint64_t)p - mag;
}

/* Leco_int<T>::decodeArray8, piecewise_fix_integer_template.h:142-204, with
   read_all_bit_fix<T> (bit_read.h:222-275).  ref_compat = 1 also copies the
   reference's handling of l == width-1 as a raw block (:171-177), which
   returns garbage; ref_compat = 0 decodes that block correctly. */
ORACLE_API void leco_oracle_decode_block(const uint8_t *in, int width, uint32_t m, void *out, int ref_compat)
{
    int l = in[0];
    uint32_t bpv = width / 8;
    if (l == 127) { uint64_t v = 0; memcpy(&v, in + 1, bpv); store_val(out, width, 0, v); return; }
    if (l == 126) {
        uint64_t v = 0; memcpy(&v, in + 1, bpv); store_val(out, width, 0, v);
        v = 0; memcpy(&v, in + 1 + bpv, bpv); store_val(out, width, 1, v);
        return;
    }
    if (l >= width || (ref_compat && l == width - 1)) {
        memcpy(out, in + 1, (uint64_t)m * bpv);
        return;
    }
    double t0, t1;
    memcpy(&t0, in + 1, 8);
    memcpy(&t1, in + 9, 8);
    const uint8_t *pay = in + 17;
    for (uint32_t i = 0; i < m; i++) {
        int64_t p = predict(t0, t1, i);
        uint64_t v = l ? apply_residual(p, read_field(pay, (uint64_t)i * l, l), l) : (uint64_t)p;
        store_val(out, width, i, v);
    }
}

/* The same block decode in the loop shape of read_all_bit_fix<T>
   (bit_read.h:222-275), for timing the reference's CPU path: the packed
   region is consumed as consecutive unaligned 64-bit words shifted into a
   128-bit register, each step takes the register's low l bits as the next
   field, and the prediction is one fma and one CVTTSD2SI (the instruction
   the reference's (long long) cast compiles to on x86, with its NaN and
   out-of-range result INT64_MIN), stored as T.  One copy per T, like the
   template.  Output equals leco_oracle_decode_block(in, width, m, out, 0)
   (a l == width-1 block is decoded, as there); the tests hold the two
   against each other on the golden vectors and on blocks of every width. */
#define REFILL_BLOCK(NAME, T)                                                              \
    static void NAME(const uint8_t *in, uint32_t m, T *out)                                \
    {                                                                                      \
        const int l = in[0];                                                               \
        double t0, t1;                                                                     \
        memcpy(&t0, in + 1, 8);                                                            \
        memcpy(&t1, in + 9, 8);                                                            \
        if (l == 0) {                                                                      \
            for (uint32_t i = 0; i < m; i++)                                               \
                out[i] = (T)_mm_cvttsd_si64(_mm_set_sd(fma((double)i, t1, t0)));           \
            return;                                                                        \
        }                                                                                  \
        const uint8_t *word = in + 17;                                                     \
        const uint64_t mask = ((uint64_t)1 << l) - 1; /* l <= 63 here */                   \
        const uint64_t low = ((uint64_t)1 << (l - 1)) - 1;                                 \
        unsigned __int128 reg = 0;                                                         \
        int held = 0;                                                                      \
        for (uint32_t i = 0; i < m; i++) {                                                 \
            if (held < l) {                                                                \
                uint64_t w;                                                                \
                memcpy(&w, word, 8);                                                       \
                word += 8;                                                                 \
                reg |= (unsigned __int128)w << held;                                       \
                held += 64;                                                                \
            }                                                                              \
            const uint64_t field = (uint64_t)reg & mask;                                   \
            reg >>= l;                                                                     \
            held -= l;                                                                     \
            const int64_t p = _mm_cvttsd_si64(_mm_set_sd(fma((double)i, t1, t0)));         \
            const uint64_t mag = field & low;                                              \
            out[i] = (T)((field >> (l - 1)) ? (uint64_t)p + mag : (uint64_t)p - mag);      \
        }                                                                                  \
    }
REFILL_BLOCK(refill_block_u32, uint32_t)
REFILL_BLOCK(refill_block_u64, uint64_t)

ORACLE_API void leco_oracle_decode_block_refill(const uint8_t *in, int width, uint32_t m, void *out)
{
    int l = in[0];
    if (l >= width || l == 126 || l == 127) { leco_oracle_decode_block(in, width, m, out, 0); return; }
    if (width == 32) refill_block_u32(in, m, (uint32_t *)out);
    else refill_block_u64(in, m, (uint64_t *)out);
}

/* Leco_int<T>::randomdecodeArray8, :305-337, with read_bit_fix_int_wo_round
   (bit_read.h:1170-1199).  For l == 0 the reference converts the double
   straight to T (:334); this returns (T)cvtt64 like the full decoder, which
   is the same value for every stream the encoder writes below 2^63. */
ORACLE_API uint64_t leco_oracle_lookup_block(const uint8_t *in, int width, uint32_t j)
{
    int l = in[0];
    uint64_t wmask = width == 32 ? 0xFFFFFFFFull : ~0ull;
    if (l >= width) {
        uint64_t v = 0;
        memcpy(&v, in + 1 + (uint64_t)j * (width / 8), width / 8);
        return v;
    }
    double t0, t1;
    memcpy(&t0, in + 1, 8);
    memcpy(&t1, in + 9, 8);
    int64_t p = predict(t0, t1, j);
    if (!l) return (uint64_t)p & wmask;
    return apply_residual(p, read_field(in + 17, (uint64_t)j * l, l), l) & wmask;
}

/* ---- column level: experiments/leco_int_fix_template.cpp:98-213 ---------- */

typedef struct {
    const void *y; int width; uint64_t n; uint32_t bs;
    uint8_t *stream; const uint64_t *offsets; uint64_t *sizes;
    void *out; const uint64_t *idx; uint64_t nq;
    uint64_t lo, hi;
    int refill; /* decode: 1 = the refill loop (leco_oracle_decode_block_refill) */
} job_t;

static void *enc_worker(void *arg)
{
    job_t *j = (job_t *)arg;
    uint64_t bound = leco_oracle_block_bytes(j->width, j->width, j->bs);
    uint64_t b2 = leco_oracle_block_bytes(j->width - 1, j->width, j->bs);
    if (b2 > bound) bound = b2;
    for (uint64_t b = j->lo; b < j->hi; b++) {
        uint64_t s = b * j->bs, m = j->n - s < j->bs ? j->n - s : j->bs;
        const uint8_t *src = (const uint8_t *)j->y + s * (j->width / 8);
        j->sizes[b] = leco_oracle_encode_block(src, j->width, (uint32_t)m, j->stream + b * bound);
    }
    return NULL;
}

static void *dec_worker(void *arg)
{
    job_t *j = (job_t *)arg;
    for (uint64_t b = j->lo; b < j->hi; b++) {
        uint64_t s = b * j->bs, m = j->n - s < j->bs ? j->n - s : j->bs;
        uint8_t *o = (uint8_t *)j->out + s * (j->width / 8);
        if (j->refill) leco_oracle_decode_block_refill(j->stream + j->offsets[b], j->width, (uint32_t)m, o);
        else leco_oracle_decode_block(j->stream + j->offsets[b], j->width, (uint32_t)m, o, 0);
    }
    return NULL;
}

static void *ra_worker(void *arg)
{
    job_t *j = (job_t *)arg;
    for (uint64_t q = j->lo; q < j->hi; q++) {
        uint64_t i = j->idx[q], b = i / j->bs;
        uint64_t v = leco_oracle_lookup_block(j->stream + j->offsets[b], j->width, (uint32_t)(i - b * j->bs));
        store_val(j->out, j->width, q, v);
    }
    return NULL;
}

static void run_jobs(void *(*fn)(void *), job_t proto, uint64_t total, int threads)
{
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > total) threads = total ? (int)total : 1;
    pthread_t tid[256];
    job_t jobs[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; t++) {
        jobs[t] = proto;
        jobs[t].lo = total * t / threads;
        jobs[t].hi = total * (t + 1) / threads;
        if (threads == 1) { fn(&jobs[t]); return; }
        pthread_create(&tid[t], NULL, fn, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

/* Encodes a column block by block (the driver loop :111-127) into one
   contiguous stream: block b's bytes sit at offsets[b]..offsets[b+1].
   Returns the stream size.  threads > 1 encodes into bound-sized slots of a
   scratch buffer in parallel and compacts afterwards. */
ORACLE_API uint64_t leco_oracle_encode_column(const void *y, int width, uint64_t n, uint32_t bs,
                                              uint8_t *stream, uint64_t *offsets, int threads)
{
    uint64_t nb = (n + bs - 1) / bs;
    offsets[0] = 0;
    if (threads <= 1) {
        for (uint64_t b = 0; b < nb; b++) {
            uint64_t s = b * bs, m = n - s < bs ? n - s : bs;
            offsets[b + 1] = offsets[b] +
                leco_oracle_encode_block((const uint8_t *)y + s * (width / 8), width, (uint32_t)m,
                                         stream + offsets[b]);
        }
        return offsets[nb];
    }
    uint64_t bound = leco_oracle_block_bytes(width, width, bs);
    uint64_t b2 = leco_oracle_block_bytes(width - 1, width, bs);
    if (b2 > bound) bound = b2;
    uint8_t *scratch = (uint8_t *)malloc(nb * bound);
    uint64_t *sizes = (uint64_t *)malloc(nb * sizeof(uint64_t));
    if (!scratch || !sizes) { free(scratch); free(sizes); return 0; }
    job_t p = {0};
    p.y = y; p.width = width; p.n = n; p.bs = bs; p.stream = scratch; p.sizes = sizes;
    run_jobs(enc_worker, p, nb, threads);
    for (uint64_t b = 0; b < nb; b++) {
        memcpy(stream + offsets[b], scratch + b * bound, sizes[b]);
        offsets[b + 1] = offsets[b] + sizes[b];
    }
    free(scratch);
    free(sizes);
    return offsets[nb];
}

/* The timed decode loop :142-153 (without the in-loop verification);
   refill = 1 decodes each block with the refill loop. */
ORACLE_API void leco_oracle_decode_column(const uint8_t *stream, const uint64_t *offsets, int width,
                                          uint64_t n, uint32_t bs, void *out, int threads)
{
    job_t p = {0};
    p.width = width; p.n = n; p.bs = bs; p.stream = (uint8_t *)stream; p.offsets = offsets; p.out = out;
    run_jobs(dec_worker, p, (n + bs - 1) / bs, threads);
}

ORACLE_API void leco_oracle_decode_column_refill(const uint8_t *stream, const uint64_t *offsets, int width,
                                                 uint64_t n, uint32_t bs, void *out, int threads)
{
    job_t p = {0};
    p.width = width; p.n = n; p.bs = bs; p.stream = (uint8_t *)stream; p.offsets = offsets; p.out = out;
    p.refill = 1;
    run_jobs(dec_worker, p, (n + bs - 1) / bs, threads);
}

/* The random-access loop :187-207: block idx/bs, position idx%bs. */
ORACLE_API void leco_oracle_lookup_column(const uint8_t *stream, const uint64_t *offsets, int width,
                                          uint64_t n, uint32_t bs, const uint64_t *idx, uint64_t nq,
                                          void *out, int threads)
{
    job_t p = {0};
    p.width = width; p.n = n; p.bs = bs; p.stream = (uint8_t *)stream; p.offsets = offsets;
    p.idx = idx; p.nq = nq; p.out = out;
    run_jobs(ra_worker, p, nq, threads);
}

/* ---- predicate pushdown: filter_range / filter_range_close ----------------
 * Leco_int<T>::filter_range (:206-249) and filter_range_close (:251-303),
 * with read_all_bit_fix_range<T> / read_all_bit_fix_range_close<T>
 * (bit_read.h:387-441, 443-497).  Each returns the positions block_start + i
 * of matching values, in the reference's order, and their count.  The
 * reference's types decide the arithmetic: thresholds are computed in the
 * integer type of (filter +/- delta) -- uint64 (wrapping) for T = uint64,
 * int64 for T = uint32 (uint32 + 1 wraps first in filter_range) -- then
 * converted to double; `int thre = std::max(...)` is x86 CVTTSD2SI r32 (NaN
 * or out of range -> INT_MIN); decoded values are the int64
 * cvtt64(fma(i,t1,t0)) +/- magnitude, compared unsigned for T = uint64 and
 * signed int64 for T = uint32.
 * Deviations (the reference is undefined there): raw blocks (l >= w) are
 * filtered by their stored values, in position order, instead of reading
 * theta from the raw bytes and shifting by w; filter_range_close writes the
 * positions of every period after the previous ones (the reference passes the
 * same `out` to each period, :274, so only its count is defined) and gives up
 * with status 1 after LECO_ORACLE_MAX_PERIODS periods in one block. */

#define LECO_ORACLE_MAX_PERIODS (1u << 16)

static int32_t cvtt32(double x)
{
    if (!(x >= -2147483648.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}
static double std_max(double a, double b) { return a < b ? b : a; } /* std::max(a, b) */
static double std_min(double a, double b) { return b < a ? b : a; } /* std::min(a, b) */

static inline int pred_gt(int width, uint64_t dv, uint64_t f)
{
    return width == 64 ? dv > f : (int64_t)dv > (int64_t)f;
}
static inline int pred_lt(int width, uint64_t dv, uint64_t f)
{
    return width == 64 ? dv < f : (int64_t)dv < (int64_t)f;
}

/* Decoded value of position i of a non-raw block as the filter loops form it
   (int64 bits): bit_read.h:411-420 for l > 0, :237 / :285 for l == 0. */
static inline uint64_t filter_value(const uint8_t *in, int l, double t0, double t1, uint64_t i)
{
    int64_t p = predict(t0, t1, i);
    return l ? apply_residual(p, read_field(in + 17, i * (uint64_t)l, l), l) : (uint64_t)p;
}

static inline uint64_t emit(uint32_t *out, uint64_t cap, uint64_t cnt, uint32_t pos)
{
    if (cnt < cap) out[cnt] = pos;
    return cnt + 1;
}

/* Values the raw-block deviation looks at: m, or 1 / 2 for the one- and
   two-value blocks l = 127 / 126 of decodeArray8 (:152-170). */
static inline uint32_t raw_count(int l, uint32_t m)
{
    if (l == 127) return m < 1 ? m : 1;
    if (l == 126) return m < 2 ? m : 2;
    return m;
}

ORACLE_API uint64_t leco_oracle_filter_block(const uint8_t *in, int width, uint32_t m, uint64_t filter,
                                             uint32_t block_start, uint32_t *out, uint64_t cap)
{
    uint64_t wmask = width == 32 ? 0xFFFFFFFFull : ~0ull;
    filter &= wmask;
    int l = in[0];
    uint64_t cnt = 0;
    if (l >= width) {
        uint32_t mm = raw_count(l, m);
        for (uint32_t i = 0; i < mm; i++) {
            uint64_t v = 0;
            memcpy(&v, in + 1 + (uint64_t)i * (width / 8), width / 8);
            if (v > filter) cnt = emit(out, cap, cnt, block_start + i);
        }
        return cnt;
    }
    double t0, t1;
    memcpy(&t0, in + 1, 8);
    memcpy(&t1, in + 9, 8);
    int64_t delta = l ? (int64_t)1 << (l - 1) : 0;
    double x;
    if (width == 64) x = (double)(filter + 1 - (uint64_t)delta) - t0;
    else x = (double)((int64_t)(uint32_t)(filter + 1) - delta) - t0;
    int32_t thre = cvtt32(std_max(x / t1, 0.0));
    if ((uint64_t)(int64_t)thre >= m) return 0;
    for (uint32_t i = (uint32_t)thre; i < m; i++)
        if (pred_gt(width, filter_value(in, l, t0, t1, i), filter)) cnt = emit(out, cap, cnt, block_start + i);
    return cnt;
}

/* Raw-block deviation of filter_range_close: position order, each position
   once, if its value lies in (f1 + k*base, f2 + k*base) for some k >= 0
   (k = 0 only when base <= 0). */
static int raw_close_match(uint64_t v, uint64_t f1, uint64_t f2, int32_t base)
{
    if (v <= f1) return 0;
    uint64_t d = v - f1, span = f2 > f1 ? f2 - f1 : 0;
    if (base <= 0) return v < f2;
    uint64_t r = (d - 1) % (uint64_t)base + 1; /* d - k*base for the largest k with k*base < d */
    return r < span;
}

ORACLE_API int leco_oracle_filter_close_block(const uint8_t *in, int width, uint32_t m, uint64_t f1, uint64_t f2,
                                              int32_t base, uint32_t block_start, uint32_t *out, uint64_t cap,
                                              uint64_t *count)
{
    uint64_t wmask = width == 32 ? 0xFFFFFFFFull : ~0ull;
    f1 &= wmask;
    f2 &= wmask;
    int l = in[0];
    uint64_t cnt = 0;
    *count = 0;
    if (l >= width) {
        uint32_t mm = raw_count(l, m);
        for (uint32_t i = 0; i < mm; i++) {
            uint64_t v = 0;
            memcpy(&v, in + 1 + (uint64_t)i * (width / 8), width / 8);
            if (raw_close_match(v, f1, f2, base)) cnt = emit(out, cap, cnt, block_start + i);
        }
        *count = cnt;
        return 0;
    }
    double t0, t1;
    memcpy(&t0, in + 1, 8);
    memcpy(&t1, in + 9, 8);
    int64_t delta = l ? (int64_t)1 << (l - 1) : 0;
    uint64_t bstep = (uint64_t)(int64_t)base & wmask; /* T += int */
    uint32_t periods = 0;
    for (;;) {
        double x1, x2;
        if (width == 64) {
            x1 = (double)(f1 - (uint64_t)delta) - t0;
            x2 = (double)(f2 + (uint64_t)delta) - t0;
        } else {
            x1 = (double)((int64_t)f1 - delta) - t0;
            x2 = (double)((int64_t)f2 + delta) - t0;
        }
        int32_t thre1 = cvtt32(std_max(x1 / t1 - 1, 0.0));
        int32_t thre2 = cvtt32(std_min(x2 / t1 + 1, (double)m));
        if (!((uint64_t)(int64_t)thre1 < m && thre2 > 0)) break;
        if (++periods > LECO_ORACLE_MAX_PERIODS) { *count = cnt; return 1; }
        if (l) {
            for (int64_t i = thre1; i < thre2; i++) {
                uint64_t dv = filter_value(in, l, t0, t1, (uint64_t)i);
                if (pred_gt(width, dv, f1) && pred_lt(width, dv, f2))
                    cnt = emit(out, cap, cnt, block_start + (uint32_t)i);
            }
        } else {
            for (int64_t i = thre1; i <= thre2; i++) {
                uint64_t dv = (uint64_t)predict(t0, t1, (uint64_t)i);
                if (pred_gt(width, dv, f1) && pred_lt(width, dv, f2))
                    cnt = emit(out, cap, cnt, block_start + (uint32_t)i);
            }
        }
        f1 = (f1 + bstep) & wmask;
        f2 = (f2 + bstep) & wmask;
    }
    *count = cnt;
    return 0;
}

/* The driver's filter loops (experiments/leco_int_fix_template.cpp:238-252,
   283-312) over every block of a column, block_start = b * bs as uint32.
   Positions go to out[0..cap) in block order; *count gets the total. */
ORACLE_API void leco_oracle_filter_column(const uint8_t *stream, const uint64_t *offsets, int width, uint64_t n,
                                          uint32_t bs, uint64_t filter, uint32_t *out, uint64_t cap,
                                          uint64_t *count)
{
    uint64_t nb = (n + bs - 1) / bs, total = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t s = b * bs, m = n - s < bs ? n - s : bs;
        uint64_t room = total < cap ? cap - total : 0;
        total += leco_oracle_filter_block(stream + offsets[b], width, (uint32_t)m, filter, (uint32_t)s,
                                          out + (total < cap ? total : 0), room);
    }
    *count = total;
}

/* As above for filter_range_close with per-block starting filters f1[b],
   f2[b] (the driver derives them from a zonemap, :293-298).  Returns 0, or 1
   if a block needed more than LECO_ORACLE_MAX_PERIODS periods. */
ORACLE_API int leco_oracle_filter_close_column(const uint8_t *stream, const uint64_t *offsets, int width,
                                               uint64_t n, uint32_t bs, const uint64_t *f1, const uint64_t *f2,
                                               int32_t base, uint32_t *out, uint64_t cap, uint64_t *count)
{
    uint64_t nb = (n + bs - 1) / bs, total = 0;
    int st = 0;
    for (uint64_t b = 0; b < nb && !st; b++) {
        uint64_t s = b * bs, m = n - s < bs ? n - s : bs, c = 0;
        uint64_t room = total < cap ? cap - total : 0;
        st = leco_oracle_filter_close_block(stream + offsets[b], width, (uint32_t)m, f1[b], f2[b], base,
                                            (uint32_t)s, out + (total < cap ? total : 0), room, &c);
        total += c;
    }
    *count = total;
    return st;
}

/* ---- FOR_int<T> / Delta_int<T>: the fixed-partition baselines ------------
 * FOR_integer_template.h:26-250 and delta_integer_template.h:26-160, with
 * write_FOR_int_T / write_delta_int_T (bit_write.h:258-390) and the readers
 * read_all_bit_FOR, read_FOR_int, read_all_bit_Delta, read_Delta_int,
 * filter_read_all_bit_FOR[_close] (bit_read.h:1270-1505).
 *
 * FOR block:   [u8 l][T min][T max][fields], field i = y_i - min on l = bitlen(max - min)
 *              bits LSB-first: E+1 bytes (E = ceil(l*n/8)), then (32 - l)/8 zero bytes
 *              (C integer division, none for l > 32); [u8 l][T min][T max] alone for
 *              l = 0; [u8 w][raw values] when l = w.
 * Delta block: [u8 l][T y_0][fields], field i = the sign/magnitude of y_{i+1} - y_i on
 *              l = bitlen(max |delta|) + 1 bits; n fields are packed (the reference hands
 *              the writer `length` for a vector of length-1 deltas, delta_integer_template.h:81),
 *              the last one reads past that vector -- undefined there, zero here; E+1 bytes
 *              then 8 - E%8 zero bytes; [u8 0][T y_0] for l = 0; raw values when l = w.
 * Deviations (each tested): FOR decode and filters treat l = w-1 as packed (the reference
 * copies such a block as raw values, FOR_integer_template.h:88, 123); Delta decode handles
 * raw blocks (the reference reads a raw block as [y_0][fields], :98-126). */

static inline uint64_t wmask_of(int width) { return width == 32 ? 0xFFFFFFFFull : ~0ull; }

static void pack_fields(uint8_t *out, const uint64_t *f, uint64_t count, int l)
{
    /* bit k*l+b of the region = bit b of field k; bytes beyond the last field stay zero */
    for (uint64_t k = 0; k < count; k++) {
        uint64_t bit = k * (uint64_t)l;
        for (int b = 0; b < l; b++, bit++)
            if ((f[k] >> b) & 1) out[bit >> 3] |= (uint8_t)(1u << (bit & 7));
    }
}

ORACLE_API uint64_t leco_oracle_for_block_bytes(int l, int width, uint32_t n)
{
    uint32_t tb = width / 8;
    if (l >= width) return 1 + (uint64_t)n * tb;
    if (l == 0) return 1 + 2 * tb;
    uint64_t e = ((uint64_t)l * n + 7) / 8;
    int pad = (32 - l) / 8;
    return 1 + 2 * tb + e + 1 + (pad > 0 ? pad : 0);
}

ORACLE_API uint64_t leco_oracle_delta_block_bytes(int l, int width, uint32_t n)
{
    uint32_t tb = width / 8;
    if (l >= width) return 1 + (uint64_t)n * tb;
    if (l == 0) return 1 + tb;
    uint64_t e = ((uint64_t)l * n + 7) / 8;
    return 1 + tb + e + 1 + (8 - e % 8);
}

/* FOR_int<T>::encodeArray8_int (:26-77). */
ORACLE_API uint64_t leco_oracle_for_encode_block(const void *y, int width, uint32_t n, uint8_t *out)
{
    uint32_t tb = width / 8;
    uint64_t mn = load_val(y, width, 0), mx = mn;
    for (uint32_t i = 1; i < n; i++) {
        uint64_t v = load_val(y, width, i);
        if (v > mx) mx = v;
        if (v < mn) mn = v;
    }
    int l = (mx - mn) ? bitlen64(mx - mn) : 0;
    if (l > width) l = width;
    uint64_t size = leco_oracle_for_block_bytes(l, width, n);
    memset(out, 0, size);
    out[0] = (uint8_t)l;
    if (l == width) {
        memcpy(out + 1, y, (uint64_t)n * tb);
        return size;
    }
    memcpy(out + 1, &mn, tb);
    memcpy(out + 1 + tb, &mx, tb);
    if (l) {
        uint64_t *f = (uint64_t *)malloc((uint64_t)n * 8);
        for (uint32_t i = 0; i < n; i++) f[i] = (load_val(y, width, i) - mn) & wmask_of(width);
        pack_fields(out + 1 + 2 * tb, f, n, l);
        free(f);
    }
    return size;
}

/* FOR_int<T>::decodeArray8 (:79-108) with read_all_bit_FOR; ref_compat = 1 also copies
   the reference's l = w-1 -> raw handling (:88). */
ORACLE_API void leco_oracle_for_decode_block(const uint8_t *in, int width, uint32_t m, void *out, int ref_compat)
{
    int l = in[0];
    uint32_t tb = width / 8;
    if (l >= width || (ref_compat && l == width - 1)) {
        memcpy(out, in + 1, (uint64_t)m * tb);
        return;
    }
    uint64_t base = 0;
    memcpy(&base, in + 1, tb);
    for (uint32_t i = 0; i < m; i++) {
        uint64_t v = l ? base + read_field(in + 1 + 2 * tb, (uint64_t)i * l, l) : base;
        store_val(out, width, i, v & wmask_of(width));
    }
}

/* FOR_int<T>::randomdecodeArray8 (:219-246) with read_FOR_int. */
ORACLE_API uint64_t leco_oracle_for_lookup_block(const uint8_t *in, int width, uint32_t j)
{
    int l = in[0];
    uint32_t tb = width / 8;
    uint64_t v = 0;
    if (l == width) {
        memcpy(&v, in + 1 + (uint64_t)j * tb, tb);
        return v;
    }
    memcpy(&v, in + 1, tb);
    if (l) v += read_field(in + 1 + 2 * tb, (uint64_t)j * l, l);
    return v & wmask_of(width);
}

/* FOR_int<T>::filter_range (:111-158) with filter_read_all_bit_FOR: positions of values
   above filter (T compare); a block whose max <= filter reports nothing. */
ORACLE_API uint64_t leco_oracle_for_filter_block(const uint8_t *in, int width, uint32_t m, uint64_t filter,
                                                 uint32_t block_start, uint32_t *out, uint64_t cap)
{
    uint64_t wm = wmask_of(width), cnt = 0;
    uint32_t tb = width / 8;
    filter &= wm;
    int l = in[0];
    if (l >= width) {
        for (uint32_t i = 0; i < m; i++) {
            uint64_t v = 0;
            memcpy(&v, in + 1 + (uint64_t)i * tb, tb);
            if (v > filter) cnt = emit(out, cap, cnt, block_start + i);
        }
        return cnt;
    }
    uint64_t base = 0, mx = 0;
    memcpy(&base, in + 1, tb);
    memcpy(&mx, in + 1 + tb, tb);
    if (mx <= filter) return 0;
    for (uint32_t i = 0; i < m; i++) {
        uint64_t v = (l ? base + read_field(in + 1 + 2 * tb, (uint64_t)i * l, l) : base) & wm;
        if (v > filter) cnt = emit(out, cap, cnt, block_start + i);
    }
    return cnt;
}

/* FOR_int<T>::filter_range_close (:160-216): positions of values with
   v % base in (filter1, filter2) (T arithmetic; base > 0). */
ORACLE_API uint64_t leco_oracle_for_filter_close_block(const uint8_t *in, int width, uint32_t m, uint64_t f1,
                                                       uint64_t f2, int32_t base, uint32_t block_start,
                                                       uint32_t *out, uint64_t cap)
{
    uint64_t wm = wmask_of(width), cnt = 0, bt = (uint64_t)(int64_t)base & wm;
    uint32_t tb = width / 8;
    f1 &= wm;
    f2 &= wm;
    int l = in[0];
    uint64_t b0 = 0;
    if (l < width) memcpy(&b0, in + 1, tb);
    for (uint32_t i = 0; i < m; i++) {
        uint64_t v = 0;
        if (l >= width) memcpy(&v, in + 1 + (uint64_t)i * tb, tb);
        else v = (l ? b0 + read_field(in + 1 + 2 * tb, (uint64_t)i * l, l) : b0) & wm;
        uint64_t r = v % bt;
        if (r > f1 && r < f2) cnt = emit(out, cap, cnt, block_start + i);
    }
    return cnt;
}

/* Delta_int<T>::encodeArray8_int (:25-83) with write_delta_int_T. */
ORACLE_API uint64_t leco_oracle_delta_encode_block(const void *y, int width, uint32_t n, uint8_t *out)
{
    uint32_t tb = width / 8;
    uint64_t wm = wmask_of(width), mxd = 0;
    uint64_t *f = (uint64_t *)calloc((uint64_t)n + 1, 8);
    int *sg = (int *)calloc((uint64_t)n + 1, sizeof(int));
    for (uint32_t i = 0; i + 1 < n; i++) {
        uint64_t a = load_val(y, width, i), b = load_val(y, width, i + 1);
        uint64_t d = b > a ? b - a : a - b;
        sg[i] = b > a;
        f[i] = d & wm;
        if (f[i] > mxd) mxd = f[i];
    }
    int l = mxd ? bitlen64(mxd) + 1 : 0;
    if (l > width) l = width;
    uint64_t size = leco_oracle_delta_block_bytes(l, width, n);
    memset(out, 0, size);
    out[0] = (uint8_t)l;
    if (l == width) {
        memcpy(out + 1, y, (uint64_t)n * tb);
    } else {
        uint64_t y0 = load_val(y, width, 0);
        memcpy(out + 1, &y0, tb);
        if (l) {
            /* fields 0..n-2 are the deltas; field n-1 (past the reference's vector) is zero */
            for (uint32_t i = 0; i + 1 < n; i++)
                f[i] = (f[i] & (((uint64_t)1 << (l - 1)) - 1)) | ((uint64_t)sg[i] << (l - 1));
            f[n - 1] = 0;
            pack_fields(out + 1 + tb, f, n, l);
        }
    }
    free(f);
    free(sg);
    return size;
}

/* Delta_int<T>::decodeArray8 (:85-126) with read_all_bit_Delta; raw blocks are decoded as
   raw values (see the header comment). */
ORACLE_API void leco_oracle_delta_decode_block(const uint8_t *in, int width, uint32_t m, void *out)
{
    int l = in[0];
    uint32_t tb = width / 8;
    uint64_t wm = wmask_of(width);
    if (l == 127 || l == 126) {
        uint64_t v = 0;
        memcpy(&v, in + 1, tb);
        store_val(out, width, 0, v);
        if (l == 126) { v = 0; memcpy(&v, in + 1 + tb, tb); store_val(out, width, 1, v); }
        return;
    }
    if (l >= width) {
        memcpy(out, in + 1, (uint64_t)m * tb);
        return;
    }
    uint64_t base = 0;
    memcpy(&base, in + 1, tb);
    store_val(out, width, 0, base);
    for (uint32_t i = 1; i < m; i++) {
        if (l) {
            uint64_t fld = read_field(in + 1 + tb, (uint64_t)(i - 1) * l, l);
            uint64_t mag = fld & (((uint64_t)1 << (l - 1)) - 1);
            base = ((fld >> (l - 1)) & 1) ? base + mag : base - mag;
        }
        store_val(out, width, i, base & wm);
    }
}

/* Delta_int<T>::randomdecodeArray8 (:131-153) with read_Delta_int: y_0 plus the first
   j signed deltas. */
ORACLE_API uint64_t leco_oracle_delta_lookup_block(const uint8_t *in, int width, uint32_t j)
{
    int l = in[0];
    uint32_t tb = width / 8;
    uint64_t v = 0;
    if (l == width) {
        memcpy(&v, in + 1 + (uint64_t)j * tb, tb);
        return v;
    }
    memcpy(&v, in + 1, tb);
    if (l)
        for (uint32_t i = 0; i < j; i++) {
            uint64_t fld = read_field(in + 1 + tb, (uint64_t)i * l, l);
            uint64_t mag = fld & (((uint64_t)1 << (l - 1)) - 1);
            v = ((fld >> (l - 1)) & 1) ? v + mag : v - mag;
        }
    return v & wmask_of(width);
}

/* Columns of FOR / Delta blocks (codec 1 = FOR, 2 = Delta), block by block as the drivers
   experiments/FOR_int_fix_template.cpp and Delta_int_fix_template.cpp do. */
ORACLE_API uint64_t leco_oracle_base_encode_column(int codec, const void *y, int width, uint64_t n, uint32_t bs,
                                                   uint8_t *stream, uint64_t *offsets)
{
    uint64_t nb = (n + bs - 1) / bs;
    offsets[0] = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t s = b * bs, m = n - s < bs ? n - s : bs;
        const uint8_t *src = (const uint8_t *)y + s * (width / 8);
        offsets[b + 1] = offsets[b] + (codec == 1 ? leco_oracle_for_encode_block(src, width, (uint32_t)m, stream + offsets[b])
                                                  : leco_oracle_delta_encode_block(src, width, (uint32_t)m, stream + offsets[b]));
    }
    return offsets[nb];
}

ORACLE_API void leco_oracle_base_decode_column(int codec, const uint8_t *stream, const uint64_t *offsets, int width,
                                               uint64_t n, uint32_t bs, void *out)
{
    uint64_t nb = (n + bs - 1) / bs;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t s = b * bs, m = n - s < bs ? n - s : bs;
        uint8_t *dst = (uint8_t *)out + s * (width / 8);
        if (codec == 1) leco_oracle_for_decode_block(stream + offsets[b], width, (uint32_t)m, dst, 0);
        else leco_oracle_delta_decode_block(stream + offsets[b], width, (uint32_t)m, dst);
    }
}

ORACLE_API void leco_oracle_base_lookup_column(int codec, const uint8_t *stream, const uint64_t *offsets, int width,
                                               uint64_t n, uint32_t bs, const uint64_t *idx, uint64_t nq, void *out)
{
    for (uint64_t q = 0; q < nq; q++) {
        uint64_t i = idx[q], b = i / bs;
        uint32_t j = (uint32_t)(i - b * bs);
        uint64_t v = codec == 1 ? leco_oracle_for_lookup_block(stream + offsets[b], width, j)
                                : leco_oracle_delta_lookup_block(stream + offsets[b], width, j);
        store_val(out, width, q, v);
    }
}

/* FOR filters over a column (positions b*bs + i in block order); close = 1 is
   filter_range_close with (f1, f2, base) for every block. */
ORACLE_API uint64_t leco_oracle_for_filter_column(const uint8_t *stream, const uint64_t *offsets, int width,
                                                  uint64_t n, uint32_t bs, int close, uint64_t f1, uint64_t f2,
                                                  int32_t base, uint32_t *out, uint64_t cap)
{
    uint64_t nb = (n + bs - 1) / bs, total = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t s = b * bs, m = n - s < bs ? n - s : bs;
        uint64_t room = total < cap ? cap - total : 0;
        uint32_t *o = out + (total < cap ? total : 0);
        total += close ? leco_oracle_for_filter_close_block(stream + offsets[b], width, (uint32_t)m, f1, f2, base,
                                                            (uint32_t)s, o, room)
                       : leco_oracle_for_filter_block(stream + offsets[b], width, (uint32_t)m, f1, (uint32_t)s, o,
                                                      room);
    }
    return total;
}

/* ---- piecewise_fix_op_max: the uint32 IntegerCODEC ------------------------
 * headers/piecewise_fix_op_minimize_maxerror.h:45-186 with lr_int::caltheta
 * (lr.h:89-111, the same FMA recipe as lr_int_T), write_delta_T /
 * write_delta_default (bit_write.h:186-247, 395-410), read_all_bit_fix_add
 * (bit_read.h:559-604), read_all_default / read_bit_default (bit_read.h:26-75)
 * and read_bit_fix_T (bit_read.h:1107-1132).
 *   block: [u8 l][f64 theta0][f32 theta1]; l = 0: nothing more; l >= 31: n raw uint32;
 *          else n fields of l bits, field = (|d| & (2^(l-1)-1)) | (d > 0) << (l-1),
 *          d = (int32)(y - pm_i), E+1 bytes then (32 - l)/8 zero bytes.
 *   pm_i = cvtt64(fma((double)theta1, i, theta0))      (encoder, random access)
 *   pa_i = cvtt64(p_i), p_0 = theta0, p_{i+1} = p_i + (double)theta1   (decoder)
 * The decoder rebuilds y_i + pa_i - pm_i; the encoder records sign(pm_i - pa_i) per
 * value (the harness's mul_add_diff_set / _minus lists, experiments/fix_int.cpp:110-117),
 * here as fix[i] in {-1, 0, +1}; applying fix[i] to the decoded value is the harness's
 * fix-up. */

static inline int64_t opmax_pm(double t0, float t1, uint32_t i)
{
    return leco_oracle_cvtt64(fma((double)t1, (double)i, t0));
}

ORACLE_API uint64_t leco_oracle_opmax_block_bytes(int l, uint32_t n)
{
    if (l == 0) return 13;
    if (l >= 31) return 13 + 4 * (uint64_t)n;
    uint64_t e = ((uint64_t)l * n + 7) / 8;
    return 13 + e + 1 + (uint64_t)((32 - l) / 8);
}

/* piecewise_fix_op_max::encodeArray8 (:45-131); fix may be NULL. */
ORACLE_API uint64_t leco_oracle_opmax_encode_block(const uint32_t *y, uint32_t n, uint8_t *out, int8_t *fix)
{
    double t0, t1d;
    leco_oracle_fit(y, 32, n, &t0, &t1d);
    const float t1 = (float)t1d;
    int32_t mx = INT32_MIN, mn = INT32_MAX;
    for (uint32_t i = 0; i < n; i++) {
        int32_t d = (int32_t)(uint32_t)((int64_t)y[i] - opmax_pm(t0, t1, i));
        if (d > mx) mx = d;
        if (d < mn) mn = d;
    }
    t0 += (double)(int32_t)((uint32_t)mx + (uint32_t)mn) / 2.0;
    int32_t *d = (int32_t *)malloc((uint64_t)n * sizeof(int32_t));
    int32_t maxe = 0;
    double p = t0;
    for (uint32_t i = 0; i < n; i++) {
        int64_t pm = opmax_pm(t0, t1, i), pa = leco_oracle_cvtt64(p);
        if (fix) fix[i] = (int8_t)(pm > pa ? 1 : pm < pa ? -1 : 0);
        p += (double)t1;
        d[i] = (int32_t)(uint32_t)((int64_t)y[i] - pm);
        if (d[i] != INT32_MIN && (d[i] < 0 ? -d[i] : d[i]) > maxe) maxe = d[i] < 0 ? -d[i] : d[i];
    }
    int l = maxe ? (32 - __builtin_clz((uint32_t)maxe)) + 1 : 0;
    uint64_t size = leco_oracle_opmax_block_bytes(l, n);
    memset(out, 0, size);
    out[0] = (uint8_t)l;
    memcpy(out + 1, &t0, 8);
    memcpy(out + 9, &t1, 4);
    if (l >= 31) {
        memcpy(out + 13, y, 4 * (uint64_t)n);
    } else if (l) {
        uint64_t *f = (uint64_t *)malloc((uint64_t)n * 8);
        uint64_t low = ((uint64_t)1 << (l - 1)) - 1;
        for (uint32_t i = 0; i < n; i++) {
            int32_t v = d[i];
            int sign = v > 0;
            int32_t mag = v > 0 ? v : (int32_t)(0u - (uint32_t)v);
            f[i] = ((uint64_t)(int64_t)mag & low) | ((uint64_t)sign << (l - 1));
        }
        pack_fields(out + 13, f, n, l);
        free(f);
    }
    free(d);
    return size;
}

/* piecewise_fix_op_max::decodeArray8 (:134-169): the decoder's output, before the
   harness's fix-up. */
ORACLE_API void leco_oracle_opmax_decode_block(const uint8_t *in, uint32_t m, uint32_t *out)
{
    int l = in[0];
    double t0;
    float t1;
    memcpy(&t0, in + 1, 8);
    memcpy(&t1, in + 9, 4);
    if (l >= 31) {
        memcpy(out, in + 13, 4 * (uint64_t)m);
        return;
    }
    double p = t0;
    for (uint32_t i = 0; i < m; i++) {
        int64_t v = leco_oracle_cvtt64(p);
        if (l) {
            uint64_t fld = read_field(in + 13, (uint64_t)i * l, l);
            uint64_t mag = fld & (((uint64_t)1 << (l - 1)) - 1);
            v = ((fld >> (l - 1)) & 1) ? v + (int64_t)mag : v - (int64_t)mag;
        }
        out[i] = (uint32_t)v;
        p += (double)t1;
    }
}

/* piecewise_fix_op_max::randomdecodeArray8 (:171-197) with read_bit_fix_T. */
ORACLE_API uint32_t leco_oracle_opmax_lookup_block(const uint8_t *in, uint32_t j)
{
    int l = in[0];
    double t0;
    float t1;
    memcpy(&t0, in + 1, 8);
    memcpy(&t1, in + 9, 4);
    if (l >= 31) {
        uint32_t v;
        memcpy(&v, in + 13 + 4 * (uint64_t)j, 4);
        return v;
    }
    int64_t pm = opmax_pm(t0, t1, j);
    if (!l) return (uint32_t)pm;
    uint64_t fld = read_field(in + 13, (uint64_t)j * l, l);
    int64_t mag = (int64_t)(fld & (((uint64_t)1 << (l - 1)) - 1));
    return (uint32_t)(((fld >> (l - 1)) & 1 ? mag : -mag) + pm);
}

/* The harness's column loops (experiments/fix_int.cpp:62-79, 95-117). fix: n values. */
ORACLE_API uint64_t leco_oracle_opmax_encode_column(const uint32_t *y, uint64_t n, uint32_t bs, uint8_t *stream,
                                                    uint64_t *offsets, int8_t *fix)
{
    uint64_t nb = (n + bs - 1) / bs;
    offsets[0] = 0;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t s = b * bs, m = n - s < bs ? n - s : bs;
        offsets[b + 1] = offsets[b] + leco_oracle_opmax_encode_block(y + s, (uint32_t)m, stream + offsets[b],
                                                                     fix ? fix + s : NULL);
    }
    return offsets[nb];
}

ORACLE_API void leco_oracle_opmax_decode_column(const uint8_t *stream, const uint64_t *offsets, uint64_t n,
                                                uint32_t bs, uint32_t *out)
{
    uint64_t nb = (n + bs - 1) / bs;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t s = b * bs, m = n - s < bs ? n - s : bs;
        leco_oracle_opmax_decode_block(stream + offsets[b], (uint32_t)m, out + s);
    }
}
