// leco_int.hpp -- Codecset::Leco_int<T> on the MI355X codec library.
//
// Same class name, members and call signatures as the reference's
// headers/piecewise_fix_integer_template.h:14-338, so a driver such as
// experiments/leco_int_fix_template.cpp builds against this header and
// libleco_fix.so instead.  Every call runs the gfx950 kernels through the
// C ABI in leco_fix.h (host buffers in and out, synchronous).
//
// Differences from the reference:
//  - a failing call throws leco_fix_error (the reference reports nothing);
//  - decodeArray8 decodes a block of bit width 8*sizeof(T)-1 correctly (the
//    reference copies it as raw bytes, :171-177);
//  - filter_range / filter_range_close filter raw blocks by their values
//    (the reference reads theta from the raw bytes, which is undefined), and
//    filter_range_close writes every period's positions one after another;
//  - encodeArray8 and get_block_nums (names of the reference's IntegerCODEC,
//    headers/codecs.h:35,60) are provided: encodeArray8 is encodeArray8_int,
//    get_block_nums returns the block count given to init().
//
// For whole columns in HBM, use the leco_fix_*_column_* calls directly: one
// launch covers every block instead of one staging round trip per block.
#ifndef LECO_INT_HPP_
#define LECO_INT_HPP_

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "leco_fix.h"

namespace Codecset {

class leco_fix_error : public std::runtime_error {
public:
    leco_fix_error(int code, const char *what)
        : std::runtime_error(std::string(what) + ": " + leco_fix_strerror(code)), code(code) {}
    int code;
};

template <typename T>
class Leco_int {
    static_assert(std::is_same<T, uint32_t>::value || std::is_same<T, uint64_t>::value,
                  "Leco_int is instantiated for uint32_t and uint64_t");
    static constexpr uint32_t W = 8 * sizeof(T);

    static void check(int rc, const char *what) {
        if (rc != LECO_FIX_OK) throw leco_fix_error(rc, what);
    }

public:
    std::vector<std::pair<int, int>> mul_add_diff_set;  // :18, never filled (its producer is commented out)
    int blocks = 0;
    int block_size = 0;

    void init(int block, int block_s) {
        blocks = block;
        block_size = block_s;
    }

    // Encodes data[0..length) as one block into res; returns the end of the
    // written bytes.  res needs leco_fix_block_bytes(8*sizeof(T), 8*sizeof(T),
    // length) bytes in the worst case (the reference's drivers give 4x raw).
    uint8_t *encodeArray8_int(const T *data, const size_t length, uint8_t *res, size_t /*nvalue*/) {
        uint64_t cap = leco_fix_block_bytes(W, W, (uint32_t)length);
        uint64_t alt = leco_fix_block_bytes(W - 1, W, (uint32_t)length);
        if (alt > cap) cap = alt;
        uint64_t written = 0;
        int rc;
        if constexpr (W == 32)
            rc = leco_fix_encode_block_u32(data, length, res, cap, &written);
        else
            rc = leco_fix_encode_block_u64(data, length, res, cap, &written);
        check(rc, "encodeArray8_int");
        return res + written;
    }

    uint8_t *encodeArray8(const T *data, const size_t length, uint8_t *res, size_t nvalue) {
        return encodeArray8_int(data, length, res, nvalue);
    }

    T *decodeArray8(const uint8_t *in, const size_t length, T *out, size_t /*nvalue*/) {
        if constexpr (W == 32)
            check(leco_fix_decode_block_u32(in, length, out), "decodeArray8");
        else
            check(leco_fix_decode_block_u64(in, length, out), "decodeArray8");
        return out;
    }

    T randomdecodeArray8(const uint8_t *in, int to_find, uint32_t * /*out*/, size_t /*nvalue*/) {
        T v = 0;
        if constexpr (W == 32)
            check(leco_fix_random_decode_u32(in, (uint32_t)to_find, &v), "randomdecodeArray8");
        else
            check(leco_fix_random_decode_u64(in, (uint32_t)to_find, &v), "randomdecodeArray8");
        return v;
    }

    // Positions block_id*block_size + i of the block's values above `filter`
    // from its model threshold on (:206-249); out needs room for `length`.
    int filter_range(uint8_t *in, const size_t length, T filter, uint32_t *out, int block_id) {
        uint64_t count = 0;
        if constexpr (W == 32)
            check(leco_fix_filter_block_u32(in, length, filter, out, length, block_id, (uint32_t)block_size,
                                            &count),
                  "filter_range");
        else
            check(leco_fix_filter_block_u64(in, length, filter, out, length, block_id, (uint32_t)block_size,
                                            &count),
                  "filter_range");
        return (int)count;
    }

    // Values in (filter1, filter2), (filter1+base, filter2+base), ... within
    // the block's model reach (:251-303).  Every period's positions are
    // written one after another (the reference overwrites them at `out`);
    // out needs room for the returned count (at most length + 1 per period).
    int filter_range_close(uint8_t *in, const size_t length, uint32_t *out, int block_id, T filter1, T filter2,
                           int base) {
        uint64_t count = 0;
        auto call = [&](uint32_t *dst, uint64_t cap) {
            if constexpr (W == 32)
                return leco_fix_filter_close_block_u32(in, length, dst, cap, block_id, (uint32_t)block_size,
                                                       filter1, filter2, base, &count);
            else
                return leco_fix_filter_close_block_u64(in, length, dst, cap, block_id, (uint32_t)block_size,
                                                       filter1, filter2, base, &count);
        };
        check(call(out, length + 1), "filter_range_close");
        if (count > length + 1) check(call(out, count), "filter_range_close");
        return (int)count;
    }

    uint32_t get_block_nums() { return (uint32_t)blocks; }
};

}  // namespace Codecset

#endif  // LECO_INT_HPP_
