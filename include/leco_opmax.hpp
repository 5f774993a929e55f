// leco_opmax.hpp -- Codecset::piecewise_fix_op_max on the MI355X codec library.
//
// Same class name, members and call signatures as the reference's uint32
// IntegerCODEC piecewise_fix_op_max (headers/piecewise_fix_op_minimize_maxerror.h:18-212,
// with IntegerCODEC's members mul_add_diff_set / mul_add_diff_set_minus,
// headers/codecs.h:30-31), so experiments/fix_int.cpp's flow builds against
// this header and libleco_fix.so once it constructs the codec directly instead
// of through CODECFactory.  Every call runs the gfx950 kernels through the C
// ABI in leco_fix.h (host buffers, synchronous).
//
// encodeArray8 appends the block's positions (i + nvalue * block_size) to
// mul_add_diff_set / mul_add_diff_set_minus exactly as the reference does;
// decodeArray8 returns the reference decoder's output, to which the harness
// applies those lists.  A failing call throws leco_fix_error.
#ifndef LECO_OPMAX_HPP_
#define LECO_OPMAX_HPP_

#include <cstddef>
#include <cstdint>
#include <iostream>
#include <string>
#include <vector>

#include "leco_int.hpp"

namespace Codecset {

class piecewise_fix_op_max {
    static void check(int rc, const char *what) {
        if (rc != LECO_FIX_OK) throw leco_fix_error(rc, what);
    }

public:
    std::vector<int> mul_add_diff_set;
    std::vector<int> mul_add_diff_set_minus;
    int block_num = 0;
    int block_size = 0;

    void init(int blocks, int blocksize, int /*extra*/) {
        block_num = blocks;
        block_size = blocksize;
    }

    // res needs leco_fix_opmax_stream_bound(length, length) bytes in the worst case.
    uint8_t *encodeArray8(uint32_t *in, const size_t length, uint8_t *res, size_t nvalue) {
        const uint64_t cap = leco_fix_opmax_stream_bound(length, (uint32_t)length);
        std::vector<int8_t> fix(length);
        uint64_t written = 0;
        check(leco_fix_opmax_encode_block(in, length, res, cap, &written, fix.data()), "encodeArray8");
        for (size_t i = 0; i < length; ++i) {
            if (fix[i] > 0) mul_add_diff_set.push_back((int)(i + nvalue * block_size));
            if (fix[i] < 0) mul_add_diff_set_minus.push_back((int)(i + nvalue * block_size));
        }
        return res + written;
    }

    uint32_t *decodeArray8(uint8_t *in, const size_t length, uint32_t *out, size_t /*nvalue*/) {
        check(leco_fix_opmax_decode_block(in, length, out), "decodeArray8");
        return out;
    }

    // Only the block's bytes up to value l are read.
    uint32_t randomdecodeArray8(uint8_t *in, const size_t l, uint32_t * /*out*/, size_t /*nvalue*/) {
        uint32_t v = 0;
        check(leco_fix_opmax_random_decode(in, (uint64_t)l + 1, (uint32_t)l, &v), "randomdecodeArray8");
        return v;
    }

    uint64_t summation(uint8_t * /*in*/, const size_t /*l*/, size_t /*nvalue*/) { return 0; }

    // The reference's uint32_t-stream variants are not implemented there either.
    uint32_t *encodeArray(uint32_t * /*in*/, const size_t /*length*/, uint32_t *out, size_t /*nvalue*/) {
        std::cout << "Haven't implement. Please try uint8_t one..." << std::endl;
        return out;
    }
    uint32_t *decodeArray(uint32_t * /*in*/, const size_t /*length*/, uint32_t *out, size_t /*nvalue*/) {
        std::cout << "Haven't implement. Please try uint8_t one..." << std::endl;
        return out;
    }
    uint32_t randomdecodeArray(uint32_t * /*in*/, const size_t /*l*/, uint32_t * /*out*/, size_t /*nvalue*/) {
        std::cout << "Haven't implement. Please try uint8_t one..." << std::endl;
        return 1;
    }

    uint32_t get_block_nums() { return 1; }
    std::string name() const { return "piecewise_fix_op_max"; }
    void destroy() {}
};

}  // namespace Codecset

#endif  // LECO_OPMAX_HPP_
