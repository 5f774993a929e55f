/*
 * leco_fix.h -- C ABI of the MI355X (gfx950) LeCo fixed-partition integer codec.
 *
 * Drop-in for Codecset::Leco_int<T> (headers/piecewise_fix_integer_template.h:14-338
 * of the LeCo reference), T in {uint32_t, uint64_t}.  Every stream this
 * library writes is byte-identical, block by block, to what Leco_int<T>::
 * encodeArray8_int writes for the same block; a column stream is the
 * concatenation of its blocks, block b at byte offsets[b].
 *
 * Conventions
 *  - "d_" pointers are device (HBM) pointers, "h_" pointers are host pointers.
 *  - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *    Column calls are asynchronous on that stream; host-buffer block calls
 *    are synchronous.
 *  - Block b of a column of n values with block size bs holds values
 *    [b*bs, min((b+1)*bs, n)); there are ceil(n/bs) blocks
 *    (experiments/leco_int_fix_template.cpp:98-127).
 *  - A stream buffer must be readable for LECO_FIX_TAIL_PAD bytes past its
 *    last block: decode and lookup read whole 16-byte words.
 *  - Block offsets (d_offsets) are ceil(n/bs) + 1 entries, the last one the
 *    stream's size, as the encoders write them; offsets[b+1] bounds block b.
 *  - Every call returns LECO_FIX_OK (0) or a positive error code; nothing
 *    aborts.  leco_fix_strerror() names a code.
 */
#ifndef LECO_FIX_H_
#define LECO_FIX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LECO_FIX_OK 0
#define LECO_FIX_EINVAL 1   /* bad argument (size, width, null pointer, capacity) */
#define LECO_FIX_EHIP 2     /* a HIP runtime call or kernel launch failed */
#define LECO_FIX_ENOMEM 3   /* device or host allocation failed */
#define LECO_FIX_ELIMIT 4   /* a filter_range_close block needed more than 65536 periods */
#define LECO_FIX_ESCRATCH 5 /* the input needs more than the static scratch of leco_fix_filter_range_noalloc */
#define LECO_FIX_EDEVICE 6  /* a kernel found a broken invariant (an encoder's bounded look-back wait gave up) */

/* d_offsets[nb] after an encode whose device-side look-back wait reached its
 * bound (a unit never published its size: a broken invariant, not expected
 * in any run).  The stream and offsets of that encode are not valid and must
 * not be decoded.  The column and batch encoders are asynchronous, so they
 * cannot return LECO_FIX_EDEVICE themselves; leco_fix_encode_result() reads
 * d_offsets[nb] and does, as do the synchronous calls that read it. */
#define LECO_FIX_OFFSETS_POISON UINT64_MAX

#define LECO_FIX_TAIL_PAD 64

/* Per-block parameters, kept next to a column stream so that a point lookup
 * needs one table read and one payload read.  For a raw block (bits == width)
 * theta0/theta1 are unused and payload_offset points at the raw values. */
typedef struct leco_fix_block_param {
    double theta0;           /* centred intercept, as stored in the block header */
    double theta1;           /* slope, as stored in the block header */
    uint64_t payload_offset; /* stream byte offset of the packed residuals / raw values */
    uint32_t bits;           /* residual bit width l (byte 0 of the block) */
    uint32_t length;         /* values in the block */
} leco_fix_block_param;

const char *leco_fix_strerror(int code);

/* ---- sizes ------------------------------------------------------------- */

/* ceil(n / block_size): the `blocks` count of Leco_int<T>::init
 * (piecewise_fix_integer_template.h:22-26, leco_int_fix_template.cpp:100-105). */
uint64_t leco_fix_block_count(uint64_t n, uint32_t block_size);

/* Bytes a column stream buffer needs in the worst case (every block at the
 * largest of raw and l = width-1), plus LECO_FIX_TAIL_PAD.  width_bits is 32 or 64. */
uint64_t leco_fix_stream_bound(uint64_t n, uint32_t block_size, uint32_t width_bits);

/* Size of one encoded block with bit width l (bit_write.h:257-323 and
 * piecewise_fix_integer_template.h:118-137). */
uint64_t leco_fix_block_bytes(uint32_t bits, uint32_t width_bits, uint32_t length);

/* ---- column calls on device memory (asynchronous on `stream`) ----------- */

/* Encodes a column: per block, Leco_int<T>::encodeArray8_int
 * (piecewise_fix_integer_template.h:27-140).  Writes the concatenated block
 * streams to d_stream, block start offsets to d_offsets[0..nb] (d_offsets[nb]
 * = total bytes) and the lookup table to d_params[0..nb-1].
 * stream_capacity must be >= leco_fix_stream_bound(n, block_size, width). */
int leco_fix_encode_column_u32(const uint32_t *d_in, uint64_t n, uint32_t block_size,
                               uint8_t *d_stream, uint64_t stream_capacity,
                               uint64_t *d_offsets, leco_fix_block_param *d_params, void *stream);
int leco_fix_encode_column_u64(const uint64_t *d_in, uint64_t n, uint32_t block_size,
                               uint8_t *d_stream, uint64_t stream_capacity,
                               uint64_t *d_offsets, leco_fix_block_param *d_params, void *stream);

/* Full decode: per block, Leco_int<T>::decodeArray8
 * (piecewise_fix_integer_template.h:142-204) into d_out[0..n).
 * Deviation: a block with l == width-1 is decoded, where the reference
 * copies it as raw bytes (:171-177) and returns wrong values. */
int leco_fix_decode_column_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                               uint32_t block_size, uint32_t *d_out, void *stream);
int leco_fix_decode_column_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                               uint32_t block_size, uint64_t *d_out, void *stream);

/* Range decode: values [begin, end) of the column into d_out[0..end-begin). */
int leco_fix_decode_range_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                              uint32_t block_size, uint64_t begin, uint64_t end,
                              uint32_t *d_out, void *stream);
int leco_fix_decode_range_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                              uint32_t block_size, uint64_t begin, uint64_t end,
                              uint64_t *d_out, void *stream);

/* Range decode with the decode shape chosen explicitly: LECO_FIX_DECODE_AUTO
 * (= _STAGED), _STAGED (through LDS: one workgroup per block for block sizes
 * >= 2048, tiles of several blocks per workgroup below; d_params may be NULL)
 * or _VALUES (one lane per position from the param table; needs d_params). */
#define LECO_FIX_DECODE_AUTO 0
#define LECO_FIX_DECODE_STAGED 1
#define LECO_FIX_DECODE_VALUES 2
int leco_fix_decode_range_ex_u32(const uint8_t *d_stream, const uint64_t *d_offsets,
                                 const leco_fix_block_param *d_params, uint64_t n, uint32_t block_size,
                                 uint64_t begin, uint64_t end, int shape, uint32_t *d_out, void *stream);
int leco_fix_decode_range_ex_u64(const uint8_t *d_stream, const uint64_t *d_offsets,
                                 const leco_fix_block_param *d_params, uint64_t n, uint32_t block_size,
                                 uint64_t begin, uint64_t end, int shape, uint64_t *d_out, void *stream);

/* Point lookups: d_out[q] = value d_idx[q] of the column; per query
 * Leco_int<T>::randomdecodeArray8(block idx/bs, idx%bs)
 * (piecewise_fix_integer_template.h:305-337, leco_int_fix_template.cpp:189-193).
 * An index >= n yields 0. */
int leco_fix_lookup_u32(const uint8_t *d_stream, const leco_fix_block_param *d_params,
                        uint64_t n, uint32_t block_size, const uint64_t *d_idx, uint64_t nq,
                        uint32_t *d_out, void *stream);
int leco_fix_lookup_u64(const uint8_t *d_stream, const leco_fix_block_param *d_params,
                        uint64_t n, uint32_t block_size, const uint64_t *d_idx, uint64_t nq,
                        uint64_t *d_out, void *stream);

/* Point lookups in large batches, the same results as leco_fix_lookup_*: the
 * queries are first grouped by the range of 2^20 or more values they fall in
 * (a counting sort of index/query pairs into d_work), then looked up group by
 * group, so that each block's bytes and parameter entry are read from HBM
 * about once per batch instead of once per query (per query, the reference's
 * randomdecodeArray8, piecewise_fix_integer_template.h:305-337).  It moves
 * less than half the HBM bytes of leco_fix_lookup_* on uniformly random
 * queries, but takes longer there: each result is a scattered 8-byte write,
 * which costs a 32-byte HBM write.  No counterpart in the reference.  d_work holds
 * leco_fix_lookup_grouped_workspace(nq) bytes, n <= 2^33.  d_idx and d_out
 * must not overlap d_work.  Asynchronous on `stream`; calls that share a
 * workspace must be ordered. */
uint64_t leco_fix_lookup_grouped_workspace(uint64_t nq);
int leco_fix_lookup_grouped_u32(const uint8_t *d_stream, const leco_fix_block_param *d_params,
                                uint64_t n, uint32_t block_size, const uint64_t *d_idx, uint64_t nq,
                                uint32_t *d_out, void *d_work, uint64_t work_bytes, void *stream);
int leco_fix_lookup_grouped_u64(const uint8_t *d_stream, const leco_fix_block_param *d_params,
                                uint64_t n, uint32_t block_size, const uint64_t *d_idx, uint64_t nq,
                                uint64_t *d_out, void *d_work, uint64_t work_bytes, void *stream);

/* Waits for `stream` and reads the encoded size d_offsets[num_blocks] of an
 * encode enqueued on it (any codec and width): LECO_FIX_OK with
 * *stream_bytes set, or LECO_FIX_EDEVICE when the encoder poisoned the
 * offsets (LECO_FIX_OFFSETS_POISON).  No counterpart in the reference, whose
 * encoder is synchronous and cannot fail. */
int leco_fix_encode_result(const uint64_t *d_offsets, uint64_t num_blocks, uint64_t *stream_bytes, void *stream);

/* Test only: sets the encoders' look-back wait bound (0 = the default of 2^26
 * sleeps) and a unit (block ticket, or ~0 for none) that never publishes its
 * status, so that a test drives the poisoned-offsets path with every kernel
 * exiting normally.  The settings hold for later launches until reset with
 * (0, ~0). */
int leco_fix_test_lookback_hook(uint32_t spin_limit, uint64_t withhold_unit);

/* Rebuilds d_params from block headers, for a stream written elsewhere. */
int leco_fix_build_params(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                          uint32_t block_size, uint32_t width_bits,
                          leco_fix_block_param *d_params, void *stream);

/* ---- one block in host memory (synchronous; the Leco_int<T> surface) ---- */

/* Leco_int<T>::encodeArray8_int(data, length, res, nvalue)
 * (piecewise_fix_integer_template.h:27): encodes h_data[0..length) as one
 * block into h_res and stores the byte count in *bytes_written. */
int leco_fix_encode_block_u32(const uint32_t *h_data, uint64_t length, uint8_t *h_res,
                              uint64_t res_capacity, uint64_t *bytes_written);
int leco_fix_encode_block_u64(const uint64_t *h_data, uint64_t length, uint8_t *h_res,
                              uint64_t res_capacity, uint64_t *bytes_written);

/* Leco_int<T>::decodeArray8(in, length, out, nvalue) (:142): decodes one
 * block of `length` values.  The block's size follows from its header. */
int leco_fix_decode_block_u32(const uint8_t *h_in, uint64_t length, uint32_t *h_out);
int leco_fix_decode_block_u64(const uint8_t *h_in, uint64_t length, uint64_t *h_out);

/* Leco_int<T>::randomdecodeArray8(in, to_find, out, nvalue) (:305): value
 * to_find of one block. */
int leco_fix_random_decode_u32(const uint8_t *h_in, uint32_t to_find, uint32_t *h_value);
int leco_fix_random_decode_u64(const uint8_t *h_in, uint32_t to_find, uint64_t *h_value);

/* ---- predicate pushdown on compressed columns -------------------------- */

/* Leco_int<T>::filter_range (piecewise_fix_integer_template.h:206-249) over
 * every block of a column: per block, the positions i >= the block's model
 * threshold whose value is above `filter`, as uint32 column positions, block
 * by block.  Values compare as the reference compares them (unsigned for
 * u64, signed int64 for u32).  At most out_capacity positions go to d_out;
 * d_result[0] receives the total count and d_result[1] flags (bit 0: a block
 * of filter_range_close needed more than 65536 periods and was cut short).
 * d_work: leco_fix_filter_workspace_bytes(n, block_size) bytes of device
 * memory (a count and a match bitmap per tile of at most 8192 positions, and
 * a sum per 256 tiles: about n/8 bytes).  Calls that share one d_work must
 * not overlap.  Asynchronous on `stream`; n <= 2^32.
 * Raw blocks (bits >= width) are filtered by their values: the reference reads
 * theta from their raw bytes, which is undefined. */
uint64_t leco_fix_filter_workspace_bytes(uint64_t n, uint32_t block_size);
int leco_fix_filter_range_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t block_size,
                              uint32_t filter, uint32_t *d_out, uint64_t out_capacity, uint64_t *d_result,
                              void *d_work, uint64_t work_bytes, void *stream);
int leco_fix_filter_range_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n, uint32_t block_size,
                              uint64_t filter, uint32_t *d_out, uint64_t out_capacity, uint64_t *d_result,
                              void *d_work, uint64_t work_bytes, void *stream);
/* Leco_int<T>::filter_range_close (:251-303) over every block: block b starts
 * at the filters d_filter1[b], d_filter2[b] (the reference's driver derives
 * them from a zonemap) and reports values in (f1, f2), then (f1+base,
 * f2+base), ... while the block's model reaches them; positions period by
 * period (the reference overwrites earlier periods' positions, :274; its
 * count is this count). */
int leco_fix_filter_range_close_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                    uint32_t block_size, const uint32_t *d_filter1, const uint32_t *d_filter2,
                                    int32_t base, uint32_t *d_out, uint64_t out_capacity, uint64_t *d_result,
                                    void *d_work, uint64_t work_bytes, void *stream);
int leco_fix_filter_range_close_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                    uint32_t block_size, const uint64_t *d_filter1, const uint64_t *d_filter2,
                                    int32_t base, uint32_t *d_out, uint64_t out_capacity, uint64_t *d_result,
                                    void *d_work, uint64_t work_bytes, void *stream);
/* leco_fix_filter_range_u32 / _u64 for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as leco_fix_filter_range_u<width_bits>, except that there is no
 * d_work: the workspace is the library's static device array of
 * LECO_FIX_NOALLOC_SCRATCH_BYTES (64 MiB), held by every GPU the library runs
 * on until the process ends.  The call allocates nothing.  An input whose
 * workspace would exceed it (about 520M values at block sizes that divide
 * 8192) returns LECO_FIX_ESCRATCH before any HIP call.  All calls on one GPU
 * share that GPU's array: calls in a row are fine, calls that overlap on one
 * GPU (concurrent streams or threads) race and give wrong results. */
#define LECO_FIX_NOALLOC_SCRATCH_BYTES (64ull << 20)
int leco_fix_filter_range_noalloc(uint32_t width_bits, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                  uint32_t block_size, uint64_t filter, uint32_t *d_out, uint64_t out_capacity,
                                  uint64_t *d_result, void *stream);
/* The two member functions on one block in host memory (filter_range(in,
 * length, filter, out, block_id) and filter_range_close(in, length, out,
 * block_id, filter1, filter2, base)): positions block_id * block_size + i.
 * *count receives the number of matches; at most out_capacity are written.
 * Returns LECO_FIX_ELIMIT when the period limit cut a close filter short. */
int leco_fix_filter_block_u32(const uint8_t *h_in, uint64_t length, uint32_t filter, uint32_t *h_out,
                              uint64_t out_capacity, int32_t block_id, uint32_t block_size, uint64_t *count);
int leco_fix_filter_block_u64(const uint8_t *h_in, uint64_t length, uint64_t filter, uint32_t *h_out,
                              uint64_t out_capacity, int32_t block_id, uint32_t block_size, uint64_t *count);
int leco_fix_filter_close_block_u32(const uint8_t *h_in, uint64_t length, uint32_t *h_out, uint64_t out_capacity,
                                    int32_t block_id, uint32_t block_size, uint32_t filter1, uint32_t filter2,
                                    int32_t base, uint64_t *count);
int leco_fix_filter_close_block_u64(const uint8_t *h_in, uint64_t length, uint32_t *h_out, uint64_t out_capacity,
                                    int32_t block_id, uint32_t block_size, uint64_t filter1, uint64_t filter2,
                                    int32_t base, uint64_t *count);

/* ---- fixed-partition baselines: FOR_int<T> and Delta_int<T> ----------- */

/* The reference's FOR_int<T> (headers/FOR_integer_template.h:14-250) and
 * Delta_int<T> (headers/delta_integer_template.h:14-160): codec is
 * LECO_FIX_CODEC_FOR or LECO_FIX_CODEC_DELTA.  Columns use the LeCo layout
 * (blocks back to back, d_offsets[b] = block b's first byte, nb+1 offsets;
 * d_offsets[nb] = stream size after encode); block b's bytes are exactly what
 * the reference's encodeArray8_int writes for it, except that Delta's last
 * packed field (read past the reference's vector of n-1 deltas,
 * delta_integer_template.h:81) is zero.  Decoding differs from the reference
 * where the reference is wrong: FOR blocks of width 8*sizeof(T)-1 are packed
 * (its decodeArray8 copies them as raw, :88) and Delta raw blocks are raw (its
 * decodeArray8 reads them as packed, :98-126).  The one-block host calls stage
 * through HBM like the LeCo ones; random_decode takes the block length. */
#define LECO_FIX_CODEC_FOR 1
#define LECO_FIX_CODEC_DELTA 2
uint64_t leco_fix_base_stream_bound(int codec, uint64_t n, uint32_t block_size, uint32_t width_bits);
int leco_fix_base_encode_column_u32(int codec, const uint32_t *d_in, uint64_t n, uint32_t block_size,
                                    uint8_t *d_stream, uint64_t stream_capacity, uint64_t *d_offsets, void *stream);
int leco_fix_base_encode_column_u64(int codec, const uint64_t *d_in, uint64_t n, uint32_t block_size,
                                    uint8_t *d_stream, uint64_t stream_capacity, uint64_t *d_offsets, void *stream);
int leco_fix_base_decode_column_u32(int codec, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                    uint32_t block_size, uint32_t *d_out, void *stream);
int leco_fix_base_decode_column_u64(int codec, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                    uint32_t block_size, uint64_t *d_out, void *stream);
int leco_fix_base_lookup_u32(int codec, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                             uint32_t block_size, const uint64_t *d_idx, uint64_t nq, uint32_t *d_out, void *stream);
int leco_fix_base_lookup_u64(int codec, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                             uint32_t block_size, const uint64_t *d_idx, uint64_t nq, uint64_t *d_out, void *stream);
/* FOR_int<T>::filter_range (:111-158: every position, value > filter; a block
 * whose max <= filter is skipped) and filter_range_close (:160-216: value %
 * base in (filter1, filter2), base > 0), over every block; results and
 * workspace as leco_fix_filter_range_*. */
int leco_fix_for_filter_range_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                  uint32_t block_size, uint32_t filter, uint32_t *d_out, uint64_t out_capacity,
                                  uint64_t *d_result, void *d_work, uint64_t work_bytes, void *stream);
int leco_fix_for_filter_range_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                  uint32_t block_size, uint64_t filter, uint32_t *d_out, uint64_t out_capacity,
                                  uint64_t *d_result, void *d_work, uint64_t work_bytes, void *stream);
int leco_fix_for_filter_range_close_u32(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                        uint32_t block_size, uint32_t filter1, uint32_t filter2, int32_t base,
                                        uint32_t *d_out, uint64_t out_capacity, uint64_t *d_result, void *d_work,
                                        uint64_t work_bytes, void *stream);
int leco_fix_for_filter_range_close_u64(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                        uint32_t block_size, uint64_t filter1, uint64_t filter2, int32_t base,
                                        uint32_t *d_out, uint64_t out_capacity, uint64_t *d_result, void *d_work,
                                        uint64_t work_bytes, void *stream);
int leco_fix_base_encode_block_u32(int codec, const uint32_t *h_in, uint64_t length, uint8_t *h_res,
                                   uint64_t res_capacity, uint64_t *written);
int leco_fix_base_encode_block_u64(int codec, const uint64_t *h_in, uint64_t length, uint8_t *h_res,
                                   uint64_t res_capacity, uint64_t *written);
int leco_fix_base_decode_block_u32(int codec, const uint8_t *h_in, uint64_t length, uint32_t *h_out);
int leco_fix_base_decode_block_u64(int codec, const uint8_t *h_in, uint64_t length, uint64_t *h_out);
int leco_fix_base_random_decode_u32(int codec, const uint8_t *h_in, uint64_t length, uint32_t to_find,
                                    uint32_t *h_value);
int leco_fix_base_random_decode_u64(int codec, const uint8_t *h_in, uint64_t length, uint32_t to_find,
                                    uint64_t *h_value);
int leco_fix_for_filter_block_u32(const uint8_t *h_in, uint64_t length, uint32_t filter, uint32_t *h_out,
                                  uint64_t out_capacity, int32_t block_id, uint32_t block_size, uint64_t *count);
int leco_fix_for_filter_block_u64(const uint8_t *h_in, uint64_t length, uint64_t filter, uint32_t *h_out,
                                  uint64_t out_capacity, int32_t block_id, uint32_t block_size, uint64_t *count);
int leco_fix_for_filter_close_block_u32(const uint8_t *h_in, uint64_t length, uint32_t *h_out,
                                        uint64_t out_capacity, int32_t block_id, uint32_t block_size,
                                        uint32_t filter1, uint32_t filter2, int32_t base, uint64_t *count);
int leco_fix_for_filter_close_block_u64(const uint8_t *h_in, uint64_t length, uint32_t *h_out,
                                        uint64_t out_capacity, int32_t block_id, uint32_t block_size,
                                        uint64_t filter1, uint64_t filter2, int32_t base, uint64_t *count);

/* ---- piecewise_fix_op_max: the uint32 IntegerCODEC ----------------------- */

/* The reference's piecewise_fix_op_max (headers/piecewise_fix_op_minimize_maxerror.h:18-212),
 * the paper's uint32 LeCo_fix: a float slope, block = [u8 tmp_bit][f64 theta0]
 * [f32 theta1] + payload (nothing for tmp_bit 0, the raw values for tmp_bit >= 31,
 * else tmp_bit-bit sign/magnitude residuals, write_delta_T).  Its decoder
 * predicts with an additive chain that can truncate one away from the
 * encoder's prediction; the encoder lists those positions in
 * IntegerCODEC::mul_add_diff_set (+1) and mul_add_diff_set_minus (-1)
 * (codecs.h:30-31) and the harness adds them to the decoded values
 * (experiments/fix_int.cpp:110-117).  Here the lists are d_fix: one int8 per
 * value (+1, -1 or 0), identical to the reference's lists position for
 * position, raw blocks included.
 * Column layout and params as the LeCo calls; d_offsets[nb] = stream size;
 * params.theta1 holds the float slope as a double, params.bits = tmp_bit.
 * stream_bound includes LECO_FIX_TAIL_PAD. */
uint64_t leco_fix_opmax_stream_bound(uint64_t n, uint32_t block_size);
/* encodeArray8 (:45-131) of every block; d_fix (n int8) may be NULL. */
int leco_fix_opmax_encode_column(const uint32_t *d_in, uint64_t n, uint32_t block_size, uint8_t *d_stream,
                                 uint64_t stream_capacity, uint64_t *d_offsets, leco_fix_block_param *d_params,
                                 int8_t *d_fix, void *stream);
/* decodeArray8 (:134-169) of every block.  d_fix NULL: exactly the reference
 * decoder's output.  d_fix given: each value of a packed or model-only block
 * plus its flag (the harness's correction; the input, bit for bit).  Raw
 * blocks are never adjusted: they hold the values themselves (the reference
 * harness adjusts them too, which corrupts them). */
int leco_fix_opmax_decode_column(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                 uint32_t block_size, const int8_t *d_fix, uint32_t *d_out, void *stream);
/* decodeArray8 of every block, then the reference's lists (as
 * leco_fix_opmax_fix_lists writes them) applied to packed and model-only
 * blocks: the input, bit for bit, reading the sparse lists instead of one
 * flag byte per value. */
int leco_fix_opmax_decode_column_lists(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                       uint32_t block_size, const int32_t *d_plus, uint64_t n_plus,
                                       const int32_t *d_minus, uint64_t n_minus, uint32_t *d_out, void *stream);
/* randomdecodeArray8 (:171-197, read_bit_fix_T bit_read.h:1107-1132): exact
 * values without fix-ups.  Indices >= n give 0. */
int leco_fix_opmax_lookup(const uint8_t *d_stream, const leco_fix_block_param *d_params, uint64_t n,
                          uint32_t block_size, const uint64_t *d_idx, uint64_t nq, uint32_t *d_out, void *stream);
/* The param table of a stored column, from its block headers. */
int leco_fix_opmax_build_params(const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                                uint32_t block_size, leco_fix_block_param *d_params, void *stream);
/* The reference's two lists from d_fix, ascending, as int32 (std::vector<int>;
 * n <= 2^31): d_counts[0] = |mul_add_diff_set|, d_counts[1] =
 * |mul_add_diff_set_minus| (all counted, at most `capacity` of each written).
 * d_work: leco_fix_opmax_lists_workspace_bytes(n) bytes. */
uint64_t leco_fix_opmax_lists_workspace_bytes(uint64_t n);
int leco_fix_opmax_fix_lists(const int8_t *d_fix, uint64_t n, int32_t *d_plus, int32_t *d_minus,
                             uint64_t capacity, uint64_t *d_counts, void *d_work, uint64_t work_bytes, void *stream);
/* The harness's correction loops (experiments/fix_int.cpp:110-117):
 * d_values[p] += 1 for p in d_plus, -= 1 for p in d_minus (mod 2^32; entries
 * outside [0, n) skipped). */
int leco_fix_opmax_apply_lists(uint32_t *d_values, uint64_t n, const int32_t *d_plus, uint64_t n_plus,
                               const int32_t *d_minus, uint64_t n_minus, void *stream);
/* One block from / to host memory (staged through HBM): encodeArray8 with
 * the block's flags in h_fix (length int8, may be NULL; position i of the
 * block is list entry i + nvalue * block_size of the reference's lists),
 * decodeArray8 (the reference decoder's output) and randomdecodeArray8
 * (which reads the block only up to value to_find's bytes; length >
 * to_find is all it checks). */
int leco_fix_opmax_encode_block(const uint32_t *h_in, uint64_t length, uint8_t *h_res, uint64_t res_capacity,
                                uint64_t *written, int8_t *h_fix);
int leco_fix_opmax_decode_block(const uint8_t *h_in, uint64_t length, uint32_t *h_out);
int leco_fix_opmax_random_decode(const uint8_t *h_in, uint64_t length, uint32_t to_find, uint32_t *h_value);

/* ---- files ------------------------------------------------------------- */

/* SOSD binary, the reference's input format (experiments/leco_int_fix_template.cpp:21-40,
 * load_data_binary): a little-endian uint64 count, then count values of width_bits.
 * _info checks the file and returns the count; _load streams the values into
 * device memory (d_out holds capacity values) through pinned staging on `stream`
 * and returns when they are there; _save writes count device values.  Host
 * threads may call these concurrently (staging is per thread). */
int leco_fix_sosd_info(const char *path, uint32_t width_bits, uint64_t *count);
int leco_fix_sosd_load(const char *path, uint32_t width_bits, void *d_out, uint64_t capacity, uint64_t *count,
                       void *stream);
int leco_fix_sosd_save(const char *path, uint32_t width_bits, const void *d_in, uint64_t count, void *stream);

/* Text, the reference's other input format (experiments/leco_int_fix_template.cpp:43-58
 * load_data, experiments/fix_int.cpp:30-44): the values that
 *     while (in.good()) { T v; in >> v; if (!in.good()) break; push(v); }
 * reads from the file in the C locale, T of width_bits: whitespace-separated
 * decimal integers, an optional sign ('-' negates modulo 2^width_bits), up to
 * the first failed extraction (a character that starts no value, a value above
 * T's maximum); a value directly at the end of the file (no character after
 * it) is not kept, as in that loop.  _info parses the file and returns the
 * count (host only, no device call); _load parses it again into device memory
 * (d_out holds capacity values) through pinned staging on `stream`. */
int leco_fix_text_info(const char *path, uint32_t width_bits, uint64_t *count);
int leco_fix_text_load(const char *path, uint32_t width_bits, void *d_out, uint64_t capacity, uint64_t *count,
                       void *stream);

/* Column file: an encoded column as it lives in HBM (the reference keeps it as
 * block_start_vec, experiments/leco_int_fix_template.cpp:108,125): 64-byte header
 * ("LECOCOL1", version, width, n, block size, block count, stream bytes,
 * FNV-1a check), the block offsets from 0, and the stream.  _save accepts
 * absolute offsets (a column inside a batch arena).  _load needs
 * stream_capacity >= stream_bytes + LECO_FIX_TAIL_PAD and num_blocks + 1
 * offsets; when d_params is not NULL it rebuilds the param table on the GPU. */
typedef struct leco_fix_file_info {
    uint64_t n;
    uint64_t num_blocks;
    uint64_t stream_bytes;
    uint32_t block_size;
    uint32_t width_bits;
} leco_fix_file_info;
int leco_fix_column_file_info(const char *path, leco_fix_file_info *info);
int leco_fix_column_save(const char *path, const uint8_t *d_stream, const uint64_t *d_offsets, uint64_t n,
                         uint32_t block_size, uint32_t width_bits, void *stream);
int leco_fix_column_load(const char *path, uint8_t *d_stream, uint64_t stream_capacity, uint64_t *d_offsets,
                         leco_fix_block_param *d_params, void *stream);

/* ---- batches: many columns in one launch ------------------------------ */

/* One column of a batch. d_in is read by encode, d_out written by decode;
 * either may be NULL when that operation is not used. */
typedef struct leco_fix_column_desc {
    const void *d_in;
    void *d_out;
    uint64_t n;
    uint32_t block_size;
    uint32_t width_bits;     /* 32 or 64 */
    uint32_t tile_positions; /* blocks below 2048 values: positions one decode workgroup takes (the tile is
                                tile_positions / block_size consecutive blocks, 1..128); 0 = 8192 */
    uint32_t reserved;
} leco_fix_column_desc;

/* A batch owns a device block table over all blocks of its columns (column
 * order), one arena holding every column's stream back to back, the arena
 * offsets (total_blocks + 1) and the param table (total_blocks).  Column c's
 * blocks are global blocks [first, first + nb); its stream, offsets and
 * params are (arena, offsets + first, params + first) and work with every
 * column call above. */
typedef struct leco_fix_batch leco_fix_batch;

int leco_fix_batch_create(const leco_fix_column_desc *cols, uint32_t ncols, leco_fix_batch **out);
int leco_fix_batch_destroy(leco_fix_batch *batch);
/* Encodes every column: fit, then one workgroup per block in ticket order
 * (arena offsets by decoupled look-back), then the param fix-up. */
int leco_fix_batch_encode(leco_fix_batch *batch, void *stream);
/* Decodes every column with one launch over the batch's tile table (a tile
 * is one large block, or a run of small blocks staged together). */
int leco_fix_batch_decode(leco_fix_batch *batch, void *stream);
/* Same as encode (op 0) / decode (op 1), replayed from a hipGraph captured on
 * first use. */
int leco_fix_batch_graph_launch(leco_fix_batch *batch, int op, void *stream);
/* The plain per-column path over the same columns, for comparison: one
 * decode launch per column (leco_fix_decode_column_*), or one column encode
 * per column (leco_fix_encode_column_*) into a scratch stream the batch
 * allocates on first use; launched from a host loop, asynchronous on `stream`. */
int leco_fix_batch_decode_per_column(leco_fix_batch *batch, void *stream);
int leco_fix_batch_encode_per_column(leco_fix_batch *batch, void *stream);
/* Changes where decode writes column c (the graph picks it up). */
int leco_fix_batch_set_output(leco_fix_batch *batch, uint32_t c, void *d_out);
/* Decode grid, in workgroups: 0 = one per tile (the default); fewer makes the
 * launch persistent, each workgroup striding over the tiles. */
int leco_fix_batch_set_decode_grid(leco_fix_batch *batch, uint32_t workgroups);
int leco_fix_batch_info(const leco_fix_batch *batch, uint64_t *total_blocks, uint64_t *arena_capacity,
                        uint8_t **d_arena, uint64_t **d_offsets, leco_fix_block_param **d_params);
int leco_fix_batch_column(const leco_fix_batch *batch, uint32_t c, uint64_t *first_block, uint64_t *num_blocks);

/* Block-size autotuner (after search_bsize::search_block_size,
 * headers/search_blocksize.h:14-114): for each column, encodes `windows`
 * sample windows of `window` values at each candidate block size in one
 * batched GPU encode, and picks the size with the fewest encoded bytes;
 * sizes within `tolerance` (relative) of the best count as ties and the
 * largest of them wins (fewer blocks decode faster).  Writes out_bs[ncols].
 * Synchronous. */
int leco_fix_autotune_block_sizes(const leco_fix_column_desc *cols, uint32_t ncols, const uint32_t *candidates,
                                  uint32_t ncand, uint32_t windows, uint32_t window, double tolerance,
                                  uint32_t *out_bs, uint64_t *out_bytes, void *stream);

/* The autotuner with a decode-cost term and the tile shape.  For each
 * column: (1) every candidate block size is encoded on sample windows, as
 * above; (2) 3/4, 5/4 and 3/2 of the best candidate are tried too (a refinement
 * around the best size, after search_blocksize.h:68-110, which refines by
 * +-100 around its best); (3) the sizes whose encoded bytes are within
 * `tolerance` (relative) of the column's fewest are kept, each with its decode
 * shapes: staged for 2048 values and more, otherwise tiles of 8192 or 16384
 * positions; (4) every (width, size, shape) kept for any column is timed: one
 * batched decode of the sample windows of all columns of that width (five
 * repeats, HIP events); (5) each column starts from its fewest-bytes size and
 * shape, and takes another kept one only if its measured decode time per
 * value is more than 3% lower (the least such).  No counterpart in the
 * reference beyond the block-size search.  Synchronous; allocates. */
typedef struct leco_fix_autotune_choice {
    uint32_t block_size;
    uint32_t tile_positions;     /* for leco_fix_column_desc (0: staged decode) */
    double bytes_per_value;      /* encoded bytes per value on the sample windows */
    double decode_ns_per_value;  /* measured decode time per value of the chosen shape */
} leco_fix_autotune_choice;
int leco_fix_autotune_columns(const leco_fix_column_desc *cols, uint32_t ncols, const uint32_t *candidates,
                              uint32_t ncand, uint32_t windows, uint32_t window, double tolerance,
                              leco_fix_autotune_choice *out, void *stream);

/* ---- request server: resident columns, continuous batching ------------- */

#define LECO_FIX_REQ_LOOKUP 0 /* count indices h_idx[0..count) of `column` */
#define LECO_FIX_REQ_RANGE 1  /* values [begin, begin+count) of `column` */
#define LECO_FIX_REQ_INSERT 2 /* encode d_values (count values, block_size, width_bits) into slot `column` */
#define LECO_FIX_REQ_EVICT 3  /* drop slot `column` once no admitted request uses it */

#define LECO_FIX_SERVE_BATCHED 0 /* continuous batching into shared launches */
#define LECO_FIX_SERVE_PLAIN 1   /* one launch + synchronize per request, in arrival order */

typedef struct leco_fix_server_config {
    uint32_t slots;               /* batches in flight at once (one stream each); 0 = 4 */
    uint32_t max_batch_requests;  /* requests per batch; 0 = 8192 */
    uint64_t max_batch_values;    /* lookup indices + range values per batch; 0 = 2^26 */
    uint64_t reserve_bytes;       /* device memory kept free; the pools take the rest */
    uint64_t result_pool_bytes;   /* result pool (0 = half of what the reserve leaves) */
    uint64_t column_pool_bytes;   /* column pool for inserted columns (0 = the other half) */
} leco_fix_server_config;

typedef struct leco_fix_request {
    uint32_t kind;           /* LECO_FIX_REQ_* */
    uint32_t column;         /* column slot */
    uint64_t count;          /* indices, range length, or values to insert */
    uint64_t begin;          /* range start */
    const uint64_t *h_idx;   /* lookup indices (host) */
    void *h_out;             /* if set, results are copied here (host) before the request completes */
    union {
        const void *d_values; /* insert: values in HBM */
        void *d_out;          /* lookup / range: if set, results are written here in HBM (count values of
                                 the column's width) instead of the result pool; no copy, no pool memory */
    };
    uint32_t block_size;     /* insert */
    uint32_t width_bits;     /* insert */
    double arrival_s;        /* arrival time, seconds after the run starts (non-decreasing) */
} leco_fix_request;

typedef struct leco_fix_request_result {
    double latency_s;        /* completion - arrival */
    double queued_s;         /* launch - arrival */
    uint64_t batch;          /* batch it ran in */
    int status;              /* LECO_FIX_OK, or an error code (e.g. evicted column) */
    int reserved;
} leco_fix_request_result;

typedef struct leco_fix_server_stats {
    double wall_s;           /* first arrival to last completion */
    uint64_t batches;
    uint64_t completed;
    uint64_t failed;
    uint64_t max_batch_requests;
    uint64_t values_out;     /* lookup results + range values produced */
    uint64_t bytes_out;
    uint64_t encoded_values; /* inserted values encoded */
    uint64_t encoded_bytes;  /* their raw bytes */
    double encode_gpu_s;     /* summed encode time of inserts (HIP events) */
    uint64_t peak_result_bytes;
    uint64_t peak_column_bytes;
} leco_fix_server_stats;

typedef struct leco_fix_server leco_fix_server;

/* Carves the pools and starts up to 7 host helper threads, which copy large
 * batches of lookup indices into pinned staging.  A server is driven by one
 * thread at a time. */
int leco_fix_server_create(const leco_fix_server_config *cfg, leco_fix_server **out);
int leco_fix_server_destroy(leco_fix_server *srv);
/* Makes an encoded column (in caller memory) servable in slot `column`. */
int leco_fix_server_attach_column(leco_fix_server *srv, uint32_t column, const uint8_t *d_stream,
                                  const uint64_t *d_offsets, const leco_fix_block_param *d_params, uint64_t n,
                                  uint32_t block_size, uint32_t width_bits);
/* Serves a request trace in real time (arrivals by the host clock) and
 * fills one result per request; results[i] follows reqs[i]. */
int leco_fix_server_run(leco_fix_server *srv, const leco_fix_request *reqs, uint64_t nreq, int mode,
                        leco_fix_request_result *results, leco_fix_server_stats *stats);
/* The slowest iterations of the last batched run's admission loop (at most
 * LECO_FIX_SERVER_TRACE of them, slowest first), with the time each phase
 * took: where a latency stall comes from.  Returns the number written. */
#define LECO_FIX_SERVER_TRACE 16
typedef struct leco_fix_server_iteration {
    double start_s;          /* seconds after the run's start */
    float between_us;        /* since the previous iteration ended */
    float arrivals_us;       /* queueing the requests that arrived */
    float completions_us;    /* finished batches and inserts */
    float admission_us;      /* the admission scan: checks, result memory, work items */
    float copies_us;         /* lookup indices into pinned staging */
    float launch_us;         /* the launch, result copies to h_out, the completion event */
    float rest_us;           /* the rest of the iteration */
    float kernel_us;         /* of launch_us: the kernel launch */
    float event_us;          /* of launch_us: result copies to h_out and the completion event */
    uint32_t admitted;       /* requests admitted in this iteration */
    uint32_t queued;         /* requests waiting when it started */
    uint32_t reserved;
} leco_fix_server_iteration;
int leco_fix_server_slowest_iterations(const leco_fix_server *srv, leco_fix_server_iteration *out, uint32_t cap,
                                       uint32_t *count);
/* Device pointers of an inserted column (valid until it is evicted). */
int leco_fix_server_column(const leco_fix_server *srv, uint32_t column, const uint8_t **d_stream,
                           const uint64_t **d_offsets, const leco_fix_block_param **d_params, uint64_t *n,
                           uint32_t *block_size, uint32_t *width_bits, uint64_t *stream_bytes);

#ifdef __cplusplus
}
#endif

#endif /* LECO_FIX_H_ */
