// The flow of the reference harness experiments/fix_int.cpp for
// piecewise_fix_op_max (block partition from a block count, encodeArray8 into
// malloc'd blocks shrunk with realloc, decodeArray8 of every block, the
// mul_add_diff_set / mul_add_diff_set_minus corrections, the check against
// the input, then random access) against include/leco_opmax.hpp.  Prints
// "ok plus=<n> minus=<n>" when every value and lookup matches the input.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "leco_opmax.hpp"

int main() {
    // blocks of keys whose float-slope chain drifts (both lists fill), then
    // noisy linear keys; every block is packed (the harness would also
    // "correct" raw blocks, whose values need no correction)
    std::vector<uint32_t> data;
    for (int i = 0; i < 100005; ++i) data.push_back((uint32_t)(3.0e9 + 0.4123457 * i + 0.5));
    for (int i = 0; i < 66670; ++i) data.push_back((uint32_t)(4.1e9 + 0.0271828 * i + 0.5));
    std::mt19937 rng(3);
    for (int i = 0; i < 33336; ++i) data.push_back(1000000u + 21u * i + rng() % 7);
    const int N = (int)data.size();
    int blocks = 6;
    const int block_size = N / blocks;
    blocks = N / block_size;
    if (blocks * block_size < N) blocks++;

    Codecset::piecewise_fix_op_max codec;
    codec.init(blocks, block_size, 32);
    std::vector<uint8_t *> block_start_vec;
    for (int i = 0; i < blocks; i++) {
        const int block_length = i == blocks - 1 ? N - (blocks - 1) * block_size : block_size;
        uint8_t *descriptor = (uint8_t *)malloc(block_length * sizeof(uint64_t) + 1000);
        uint8_t *res = codec.encodeArray8(data.data() + (i * block_size), block_length, descriptor, i);
        descriptor = (uint8_t *)realloc(descriptor, res - descriptor);
        block_start_vec.push_back(descriptor);
    }
    std::vector<uint32_t> recover(N);
    for (int i = 0; i < blocks; i++) {
        const int block_length = i == blocks - 1 ? N - (blocks - 1) * block_size : block_size;
        codec.decodeArray8(block_start_vec[i], block_length, recover.data() + i * block_size, i);
    }
    for (auto index : codec.mul_add_diff_set) recover[index] += 1;
    for (auto index : codec.mul_add_diff_set_minus) recover[index] -= 1;
    for (int j = 0; j < N; j++) {
        if (data[j] != recover[j]) {
            std::printf("decode mismatch at %d: %u vs %u\n", j, data[j], recover[j]);
            return 1;
        }
    }
    for (int q = 0; q < 3000; ++q) {
        const int idx = (int)(rng() % N);
        const uint32_t v = codec.randomdecodeArray8(block_start_vec[idx / block_size], idx % block_size, nullptr, N);
        if (v != data[idx]) {
            std::printf("random access mismatch at %d: %u vs %u\n", idx, data[idx], v);
            return 1;
        }
    }
    for (auto p : block_start_vec) free(p);
    if (codec.mul_add_diff_set.empty() || codec.mul_add_diff_set_minus.empty()) {
        std::printf("lists not exercised\n");
        return 1;
    }
    std::printf("ok plus=%zu minus=%zu\n", codec.mul_add_diff_set.size(), codec.mul_add_diff_set_minus.size());
    return 0;
}
