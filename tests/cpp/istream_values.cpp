// Reads a text file the way the reference's drivers read their text inputs
// (experiments/leco_int_fix_template.cpp:43-58, experiments/fix_int.cpp:30-44):
// libstdc++'s stream extraction of T = uint32_t or uint64_t in a loop that
// keeps a value only while the stream stays good.  Prints the number of
// values, then the values, one per line.  Usage: istream_values <path> <32|64>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

template <class T>
static int run(const char *path) {
    std::ifstream in(path, std::ios::in);
    if (!in) return 2;
    std::vector<T> v;
    while (in.good()) {
        T next;
        in >> next;
        if (!in.good()) break;
        v.push_back(next);
    }
    std::printf("%zu\n", v.size());
    for (const T x : v) std::printf("%llu\n", (unsigned long long)x);
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    return std::atoi(argv[2]) == 32 ? run<uint32_t>(argv[1]) : run<uint64_t>(argv[1]);
}
