// leco_baselines.hpp -- Codecset::FOR_int<T> and Codecset::Delta_int<T> on the
// MI355X codec library.
//
// Same class names, members and call signatures as the reference's
// headers/FOR_integer_template.h:14-250 and delta_integer_template.h:14-160,
// so the drivers experiments/FOR_int_fix_template.cpp and
// Delta_int_fix_template.cpp build against this header and libleco_fix.so.
// Every call runs the gfx950 kernels through the C ABI in leco_fix.h (host
// buffers, synchronous).
//
// Differences from the reference:
//  - a failing call throws leco_fix_error;
//  - FOR blocks of width 8*sizeof(T)-1 decode and filter as packed (the
//    reference copies them as raw, FOR_integer_template.h:88,123);
//  - Delta raw blocks decode as raw (the reference reads them as packed);
//  - Delta's last packed field is zero (the reference reads it from past its
//    vector of length-1 deltas, delta_integer_template.h:81).
#ifndef LECO_BASELINES_HPP_
#define LECO_BASELINES_HPP_

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>
#include <vector>

#include "leco_int.hpp"

namespace Codecset {

template <typename T, int CODEC>
class BaseCodec_ {
    static_assert(std::is_same<T, uint32_t>::value || std::is_same<T, uint64_t>::value,
                  "instantiated for uint32_t and uint64_t");

protected:
    static constexpr uint32_t W = 8 * sizeof(T);
    static void check(int rc, const char *what) {
        if (rc != LECO_FIX_OK) throw leco_fix_error(rc, what);
    }

public:
    std::vector<std::pair<int, int>> mul_add_diff_set;
    int blocks = 0;
    int block_size = 0;

    void init(int block, int block_s) {
        blocks = block;
        block_size = block_s;
    }

    // res needs leco_fix_base_stream_bound(CODEC, length, length, 8*sizeof(T)) bytes in the worst case.
    uint8_t *encodeArray8_int(const T *data, const size_t length, uint8_t *res, size_t /*nvalue*/) {
        const uint64_t cap = leco_fix_base_stream_bound(CODEC, length, (uint32_t)length, W);
        uint64_t written = 0;
        int rc;
        if constexpr (W == 32)
            rc = leco_fix_base_encode_block_u32(CODEC, data, length, res, cap, &written);
        else
            rc = leco_fix_base_encode_block_u64(CODEC, data, length, res, cap, &written);
        check(rc, "encodeArray8_int");
        return res + written;
    }

    T *decodeArray8(const uint8_t *in, const size_t length, T *out, size_t /*nvalue*/) {
        if constexpr (W == 32)
            check(leco_fix_base_decode_block_u32(CODEC, in, length, out), "decodeArray8");
        else
            check(leco_fix_base_decode_block_u64(CODEC, in, length, out), "decodeArray8");
        return out;
    }

    // Only the block's bytes up to value to_find are read.
    T randomdecodeArray8(const uint8_t *in, int to_find, uint32_t * /*out*/, size_t /*nvalue*/) {
        T v = 0;
        if constexpr (W == 32)
            check(leco_fix_base_random_decode_u32(CODEC, in, (uint64_t)to_find + 1, (uint32_t)to_find, &v),
                  "randomdecodeArray8");
        else
            check(leco_fix_base_random_decode_u64(CODEC, in, (uint64_t)to_find + 1, (uint32_t)to_find, &v),
                  "randomdecodeArray8");
        return v;
    }
};

template <typename T>
class FOR_int : public BaseCodec_<T, LECO_FIX_CODEC_FOR> {
    using Base = BaseCodec_<T, LECO_FIX_CODEC_FOR>;

public:
    // Positions nvalue*block_size + i of values above filter; out needs room for `length`.
    int filter_range(uint8_t *in, const size_t length, T filter, uint32_t *out, size_t nvalue) {
        uint64_t count = 0;
        if constexpr (Base::W == 32)
            Base::check(leco_fix_for_filter_block_u32(in, length, filter, out, length, (int32_t)nvalue,
                                                      (uint32_t)this->block_size, &count),
                        "filter_range");
        else
            Base::check(leco_fix_for_filter_block_u64(in, length, filter, out, length, (int32_t)nvalue,
                                                      (uint32_t)this->block_size, &count),
                        "filter_range");
        return (int)count;
    }

    // Positions of values with v % base_val in (filter1, filter2); out needs room for `length`.
    int filter_range_close(uint8_t *in, const size_t length, uint32_t *out, size_t nvalue, T filter1, T filter2,
                           int base_val) {
        uint64_t count = 0;
        if constexpr (Base::W == 32)
            Base::check(leco_fix_for_filter_close_block_u32(in, length, out, length, (int32_t)nvalue,
                                                            (uint32_t)this->block_size, filter1, filter2, base_val,
                                                            &count),
                        "filter_range_close");
        else
            Base::check(leco_fix_for_filter_close_block_u64(in, length, out, length, (int32_t)nvalue,
                                                            (uint32_t)this->block_size, filter1, filter2, base_val,
                                                            &count),
                        "filter_range_close");
        return (int)count;
    }
};

template <typename T>
class Delta_int : public BaseCodec_<T, LECO_FIX_CODEC_DELTA> {};

}  // namespace Codecset

#endif  // LECO_BASELINES_HPP_
