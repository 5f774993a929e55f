// The flow of the reference drivers experiments/FOR_int_fix_template.cpp and
// Delta_int_fix_template.cpp (block partition, encodeArray8_int into
// malloc'd blocks, decodeArray8, randomdecodeArray8, FOR filter loops) against
// include/leco_baselines.hpp.  Prints "ok" when every value, lookup and filter
// result matches the input.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "leco_baselines.hpp"

template <typename T, class Codec>
static bool run(const char *name, bool with_filters) {
    const int N = 200000, blocks = 49;
    const int block_size = N / blocks;
    const int nb = N / block_size + (N % block_size ? 1 : 0);
    std::mt19937_64 rng(7);
    std::vector<T> data(N);
    uint64_t v = 1000;
    for (int i = 0; i < N; ++i) {
        v += rng() % 97;
        data[i] = (T)(v + (i % 5000 == 0 ? rng() % 100000 : 0));
    }
    Codec codec;
    codec.init(nb, block_size);
    std::vector<uint8_t *> start(nb);
    for (int i = 0; i < nb; ++i) {
        const int len = i == nb - 1 ? N - (nb - 1) * block_size : block_size;
        uint8_t *d = (uint8_t *)malloc(len * sizeof(T) * 4);
        uint8_t *e = codec.encodeArray8_int(data.data() + (size_t)i * block_size, len, d, i);
        start[i] = (uint8_t *)realloc(d, e - d);
    }
    std::vector<T> rec(N);
    for (int i = 0; i < nb; ++i) {
        const int len = i == nb - 1 ? N - (nb - 1) * block_size : block_size;
        codec.decodeArray8(start[i], len, rec.data() + (size_t)i * block_size, i);
    }
    bool ok = rec == data;
    for (int q = 0; q < 2000 && ok; ++q) {
        const int idx = (int)(rng() % N);
        ok = codec.randomdecodeArray8(start[idx / block_size], idx % block_size, nullptr, N) == data[idx];
    }
    if (with_filters) {
        if constexpr (std::is_same<Codec, Codecset::FOR_int<T>>::value) {
            const T f = data[N / 3];
            std::vector<uint32_t> pos(N);
            int total = 0;
            for (int i = 0; i < nb; ++i) {
                const int len = i == nb - 1 ? N - (nb - 1) * block_size : block_size;
                total += codec.filter_range(start[i], len, f, pos.data() + total, i);
            }
            int want = 0;
            for (int j = 0; j < N; ++j) want += data[j] > f;
            ok = ok && total == want;
            total = 0;
            for (int i = 0; i < nb; ++i) {
                const int len = i == nb - 1 ? N - (nb - 1) * block_size : block_size;
                total += codec.filter_range_close(start[i], len, pos.data() + total, i, 10, 40, 100);
            }
            want = 0;
            for (int j = 0; j < N; ++j) want += data[j] % 100 > 10 && data[j] % 100 < 40;
            ok = ok && total == want;
        }
    }
    for (auto p : start) free(p);
    printf("%s %s\n", name, ok ? "ok" : "FAILED");
    return ok;
}

int main() {
    bool ok = run<uint32_t, Codecset::FOR_int<uint32_t>>("FOR_int<uint32_t>", true);
    ok &= run<uint64_t, Codecset::FOR_int<uint64_t>>("FOR_int<uint64_t>", true);
    ok &= run<uint32_t, Codecset::Delta_int<uint32_t>>("Delta_int<uint32_t>", false);
    ok &= run<uint64_t, Codecset::Delta_int<uint64_t>>("Delta_int<uint64_t>", false);
    return ok ? 0 : 1;
}
