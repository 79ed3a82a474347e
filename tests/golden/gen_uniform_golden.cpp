// Writes the golden input-stream fixture (tests/golden/uniform_stream.json).
// The values come from libstdc++ itself, exactly as the reference harness
// draws them (test_tsqr.cu:24-28): std::default_random_engine(0U) and
// std::uniform_real_distribution<T>(-1, 1), filled in column-major order by
// std::generate.  The fixture holds the first 16 values (as hex bit patterns)
// and an FNV-1a-64 hash over the little-endian bytes of the whole array.
// Build and run:  g++ -O2 -o /tmp/gen gen_uniform_golden.cpp && /tmp/gen > uniform_stream.json
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

template <typename T>
static void emit(const char *name, size_t count, bool last) {
    std::vector<T> a(count);
    std::default_random_engine eng(0U);
    std::uniform_real_distribution<T> dis(-1.0f, 1.0f);
    std::generate(a.begin(), a.end(), [&] { return dis(eng); });
    uint64_t h = 1469598103934665603ull;
    const unsigned char *p = reinterpret_cast<const unsigned char *>(a.data());
    for (size_t i = 0; i < count * sizeof(T); ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    std::printf("  \"%s\": {\"count\": %zu, \"fnv1a64\": \"%016llx\", \"first16_hex\": [", name, count,
                (unsigned long long)h);
    for (int i = 0; i < 16; ++i) {
        uint64_t bits = 0;
        std::memcpy(&bits, &a[i], sizeof(T));
        std::printf("%s\"%0*llx\"", i ? ", " : "", (int)(2 * sizeof(T)), (unsigned long long)bits);
    }
    std::printf("], \"first16\": [");
    for (int i = 0; i < 16; ++i) std::printf("%s%.17g", i ? ", " : "", (double)a[i]);
    std::printf("]}%s\n", last ? "" : ",");
}

int main() {
    std::printf("{\n");
    std::printf("  \"source\": \"libstdc++ default_random_engine(0U) + uniform_real_distribution<T>(-1,1), column-major fill\",\n");
    emit<float>("float_27648x32", 27648u * 32u, false);
    emit<float>("float_2048x16", 2048u * 16u, false);
    emit<double>("double_27648x32", 27648u * 32u, false);
    emit<double>("double_1024x8", 1024u * 8u, true);
    std::printf("}\n");
    return 0;
}
