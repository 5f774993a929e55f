// The reference driver's flow (experiments/leco_int_fix_template.cpp:98-213)
// written against include/leco_int.hpp: partition a column into fixed blocks,
// encode each block into its own malloc'd buffer, decode every block, look up
// random positions, and verify everything bit for bit.  Prints one result
// line and exits non-zero on any mismatch.
//
//   leco_int_driver <n> <blocks> <width: 32|64>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "leco_int.hpp"

template <typename T>
static int run(size_t N, int blocks_arg) {
    std::mt19937_64 rng(42);
    std::vector<T> data(N);
    // sorted keys with linear trend + noise, a typical LeCo column
    double base = 1000.0;
    std::normal_distribution<double> noise(0.0, 64.0);
    for (size_t i = 0; i < N; i++) {
        double v = base + 37.5 * (double)i + noise(rng);
        data[i] = v < 0 ? 0 : (T)v;
    }
    for (size_t i = 1; i < N; i++)
        if (data[i] < data[i - 1]) data[i] = data[i - 1];

    Codecset::Leco_int<T> codec;
    int block_size = (int)(N / blocks_arg);
    int blocks = (int)(N / block_size);
    if ((size_t)blocks * block_size < N) blocks++;
    codec.init(blocks, block_size);

    std::vector<uint8_t *> block_start_vec;
    uint64_t totalsize = 0;
    for (int i = 0; i < blocks; i++) {
        int block_length = block_size;
        if (i == blocks - 1) block_length = (int)(N - (size_t)(blocks - 1) * block_size);
        uint8_t *descriptor = (uint8_t *)malloc((size_t)block_length * sizeof(T) * 4);
        uint8_t *res = codec.encodeArray8_int(data.data() + (size_t)i * block_size, block_length, descriptor, i);
        uint32_t segment_size = (uint32_t)(res - descriptor);
        descriptor = (uint8_t *)realloc(descriptor, segment_size);
        block_start_vec.push_back(descriptor);
        totalsize += segment_size;
    }

    std::vector<T> recover(N);
    for (int i = 0; i < blocks; i++) {
        int block_length = block_size;
        if (i == blocks - 1) block_length = (int)(N - (size_t)(blocks - 1) * block_size);
        codec.decodeArray8(block_start_vec[i], block_length, recover.data() + (size_t)i * block_size, i);
    }
    int bad = 0;
    for (size_t j = 0; j < N; j++)
        if (recover[j] != data[j]) { bad++; break; }

    std::mt19937 r2(7);
    for (int k = 0; k < 200 && !bad; k++) {
        uint32_t index = r2() % N;
        T v = codec.randomdecodeArray8(block_start_vec[index / block_size], index % block_size, NULL, N);
        if (v != data[index]) bad++;
    }
    double cr = totalsize * 100.0 / (sizeof(T) * N);
    printf("leco_int_driver width=%d N=%zu blocks=%d block_size=%d cr=%.4f%% get_block_nums=%u %s\n",
           (int)(8 * sizeof(T)), N, blocks, block_size, cr, codec.get_block_nums(), bad ? "MISMATCH" : "OK");
    for (auto p : block_start_vec) free(p);
    return bad ? 1 : 0;
}

int main(int argc, char **argv) {
    size_t n = argc > 1 ? strtoull(argv[1], nullptr, 10) : 100000;
    int blocks = argc > 2 ? atoi(argv[2]) : 25;
    int width = argc > 3 ? atoi(argv[3]) : 64;
    try {
        return width == 32 ? run<uint32_t>(n, blocks) : run<uint64_t>(n, blocks);
    } catch (const std::exception &e) {
        fprintf(stderr, "leco_int_driver: %s\n", e.what());
        return 2;
    }
}
