/*
 * tables.hip -- host check of gf::expneg_tab (parallelpopgen_amd/csrc/
 * gf_math.hpp), the Poisson inversion's exp(-m) segment table, against
 * double exp for every float m in [2^-12, 6] and every 4096th float below
 * (where exp(-m) is 1 - m to within an ulp): prints the largest error in
 * ulps of the exact value and the largest step up between consecutive
 * floats (the function is decreasing), and exits 1 if the error exceeds
 * 1.5 ulp, a step up exceeds one ulp, or exp(-0) is not 1.
 */
#include "../../parallelpopgen_amd/csrc/gf_math.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

int main() {
    const gf::ExpSeg *tab = gf::expneg_table();
    if (gf::expneg_tab(tab, 0.0f) != 1.0f) { std::printf("exp(-0) != 1\n"); return 1; }
    double worst = 0.0, worst_m = 0.0, up = 0.0;
    long long n = 0;
    float prev = 2.0f;
    uint32_t bits = 0;
    for (float m = 0.0f; m <= 6.0f;) {
        const float got = gf::expneg_tab(tab, m);
        const double want = std::exp(-(double)m);
        const float wf = (float)want;
        const double ulp = (double)(std::nextafter(wf, 2.0f) - wf);
        const double err = std::fabs((double)got - want) / ulp;
        if (err > worst) { worst = err; worst_m = m; }
        if (got > prev) { const double s = (double)(got - prev) / ulp; if (s > up) up = s; }
        prev = got;
        n++;
        std::memcpy(&bits, &m, 4);
        bits += m < 0x1p-12f ? 4096u : 1u;
        std::memcpy(&m, &bits, 4);
    }
    std::printf("expneg_tab: %lld floats in [0, 6], max error %.3f ulp at m = %.9g, max step up %.3f ulp\n",
                n, worst, worst_m, up);
    return (worst <= 1.5 && up <= 1.0) ? 0 : 1;
}
