// Writes tests/golden/normal_stream.json: the seeded inputs of the two
// configurations that the reference harness does not define itself.
//   - the recursive-QR configuration, m = 27648, n = 128, iid N(0,1):
//     std::mt19937_64(0), pairs of 53-bit uniforms u = (x >> 11) 2^-53,
//     Box-Muller in FP64: r = sqrt(-2 log(1 - u1)), t = 2 pi u2,
//     z0 = r cos t, z1 = r sin t, filled column-major; the FP32 array holds
//     the FP64 values rounded to nearest;
//   - the batched configuration's 512 shapes, from its own
//     std::mt19937_64(0): per matrix u = (x >> 11) 2^-53,
//     m = floor(exp(log 4096 + u (log 262144 - log 4096))), then
//     n = {8, 16, 32, 64}[y >> 62] from the next draw y, and m rounded down
//     to a multiple of n (the reference's n | m, TallShinnyQR.h:92-98).
// The values come from libstdc++'s std::mt19937_64 and glibc's libm, so the
// fixture pins the oracle's restatement of both (oracle/tsqr_oracle.c,
// tsqr_amd/workloads.py).
// Build and run:  g++ -O2 -o /tmp/gen gen_normal_golden.cpp && /tmp/gen > normal_stream.json
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

static uint64_t fnv1a(const void *data, size_t bytes) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

static std::vector<double> normals(size_t count) {
    std::mt19937_64 eng(0);
    std::vector<double> z(count);
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t i = 0; i < count; i += 2) {
        const double u1 = (double)(eng() >> 11) * 0x1.0p-53;
        const double u2 = (double)(eng() >> 11) * 0x1.0p-53;
        const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
        const double t = two_pi * u2;
        z[i] = r * std::cos(t);
        if (i + 1 < count) z[i + 1] = r * std::sin(t);
    }
    return z;
}

int main() {
    const size_t m = 27648, n = 128, count = m * n;
    std::mt19937_64 raw(0);
    std::printf("{\n  \"source\": \"libstdc++ std::mt19937_64(0); Box-Muller in FP64 over (x >> 11) * 2^-53; glibc libm\",\n");
    std::printf("  \"mt19937_64_first8_hex\": [");
    for (int i = 0; i < 8; ++i) std::printf("%s\"%016llx\"", i ? ", " : "", (unsigned long long)raw());
    std::printf("],\n");
    const std::vector<double> z = normals(count);
    std::vector<float> zf(count);
    for (size_t i = 0; i < count; ++i) zf[i] = (float)z[i];
    std::printf("  \"normal_first16_hex\": [");
    for (int i = 0; i < 16; ++i) {
        uint64_t b;
        std::memcpy(&b, &z[i], 8);
        std::printf("%s\"%016llx\"", i ? ", " : "", (unsigned long long)b);
    }
    std::printf("],\n  \"normal_first16\": [");
    for (int i = 0; i < 16; ++i) std::printf("%s%.17g", i ? ", " : "", z[i]);
    std::printf("],\n");
    std::printf("  \"double_27648x128\": {\"count\": %zu, \"fnv1a64\": \"%016llx\"},\n", count,
                (unsigned long long)fnv1a(z.data(), count * 8));
    std::printf("  \"float_27648x128\": {\"count\": %zu, \"fnv1a64\": \"%016llx\"},\n", count,
                (unsigned long long)fnv1a(zf.data(), count * 4));
    std::mt19937_64 eng(0);
    const int widths[4] = {8, 16, 32, 64};
    const double lo = std::log(4096.0), hi = std::log(262144.0);
    std::printf("  \"batch_512_shapes\": [");
    for (int i = 0; i < 512; ++i) {
        const double u = (double)(eng() >> 11) * 0x1.0p-53;
        long long mm = (long long)std::floor(std::exp(lo + u * (hi - lo)));
        const int nn = widths[eng() >> 62];
        mm -= mm % nn;
        std::printf("%s[%lld, %d]", i ? ", " : "", mm, nn);
    }
    std::printf("]\n}\n");
    return 0;
}
