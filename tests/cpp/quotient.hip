/*
 * quotient.hip -- host check of gf::quot_count (parallelpopgen_amd/csrc/
 * gf_math.hpp), the drift kernel's j / N: equal, bit for bit, to the IEEE
 * float quotient (float)j / (float)N for every 0 <= j <= N <= 32768, and
 * for N sampled up to 2^22 with every j near 0, near N and near each
 * multiple of N/64, plus random ones.  Prints the number of cases checked
 * and exits 1 on the first mismatch.
 */
#include "../../parallelpopgen_amd/csrc/gf_math.hpp"

#include <cstdio>
#include <cstdint>
#include <cstring>

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint32_t next_u32() {
    rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
    return (uint32_t)(rng >> 11);
}

static int check(int N, int j, uint64_t &n) {
    const float fN = (float)N, fj = (float)j;
    const float r = 1.0f / fN;
    const float want = fj / fN;
    const float got = gf::quot_count(fj, fN, r);
    n++;
    uint32_t a, b;
    std::memcpy(&a, &want, 4);
    std::memcpy(&b, &got, 4);
    if (a != b) {
        std::printf("mismatch N=%d j=%d want %.9g got %.9g\n", N, j, want, got);
        return 1;
    }
    return 0;
}

int main() {
    uint64_t n = 0;
    for (int N = 1; N <= 32768; N++)
        for (int j = 0; j <= N; j++)
            if (check(N, j, n)) return 1;
    for (int k = 0; k < 4000; k++) {
        const int N = 32769 + (int)(next_u32() % (4194304u - 32768u));
        for (int j = 0; j <= 256; j++) if (check(N, j, n) || check(N, N - j, n)) return 1;
        for (int m = 1; m < 64; m++) {
            const int c = (int)((int64_t)N * m / 64);
            for (int d = -64; d <= 64; d++) if (check(N, c + d, n)) return 1;
        }
        for (int t = 0; t < 4096; t++) if (check(N, (int)(next_u32() % (uint32_t)(N + 1)), n)) return 1;
    }
    for (int N = 4194304 - 64; N <= 4194304; N++)
        for (int t = 0; t < 4096; t++) if (check(N, (int)(next_u32() % (uint32_t)(N + 1)), n)) return 1;
    std::printf("quot_count: %llu cases, all equal to the IEEE quotient\n", (unsigned long long)n);
    return 0;
}
