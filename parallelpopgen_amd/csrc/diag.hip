/*
 * diag.hip -- gf_device_math: evaluates the device implementations of the
 * random-number and elementary functions on arrays, so their bit-exact
 * agreement with the host restatement can be checked function by function.
 */
#include "gf_math.hpp"
#include "gf_runtime.hpp"

namespace gf {

static __global__ void k_math(int op, const void *in0, const void *in1, const void *in2, const void *in3,
                              void *out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float *f0 = (const float *)in0, *f1 = (const float *)in1, *f2 = (const float *)in2;
        const uint32_t *u0 = (const uint32_t *)in0;
        float *fo = (float *)out;
        int *io = (int *)out;
        uint32_t *uo = (uint32_t *)out;
        switch (op) {
        case GF_MATH_PHILOX: {
            const uint32_t *a = u0 + 6 * i;
            u32x4 r = philox10(a[0], a[1], a[2], a[3], a[4], a[5]);
            uo[4 * i] = r.x; uo[4 * i + 1] = r.y; uo[4 * i + 2] = r.z; uo[4 * i + 3] = r.w;
            break;
        }
        case GF_MATH_U01: fo[i] = u01(u0[i]); break;
        case GF_MATH_EXP: fo[i] = exp_f(f0[i]); break;
        case GF_MATH_LOG: fo[i] = log_f(f0[i]); break;
        case GF_MATH_NORMCDFINV: fo[i] = normcdfinv_f(f0[i]); break;
        case GF_MATH_BINOM_DRAW:
            io[i] = binom_draw(u0[i], f1[i], f2[i], ((const float *)in3)[i]);
            break;
        case GF_MATH_POIS_DRAW: io[i] = pois_draw(u0[i], f1[i], f1[i], f2[i]); break;
        case GF_MATH_SQRT: fo[i] = __builtin_sqrtf(f0[i]); break;
        case GF_MATH_DIV: fo[i] = f0[i] / f1[i]; break;
        case GF_MATH_ROUND_AWAY: fo[i] = round_away(f0[i]); break;
        case GF_MATH_COPY: fo[i] = f0[i]; break;
        case GF_MATH_F2I: io[i] = f2i_sat(f0[i]); break;
        default: break;
        }
    }
}

} /* namespace gf */

extern "C" int32_t gf_device_math(int32_t op, const void *in0, const void *in1, const void *in2, const void *in3,
                                  void *out, int64_t n, int32_t device) {
    using namespace gf;
    if (n <= 0 || !in0 || !out || op < 0 || op > GF_MATH_F2I) { set_error("gf_device_math: bad arguments"); return GF_ERR_ARG; }
    GF_TRY(select_device(device));
    const size_t in_bytes = op == GF_MATH_PHILOX ? 24 : 4;
    const size_t out_bytes = op == GF_MATH_PHILOX ? 16 : 4;
    DevBuf<char> d0, d1, d2, d3, dout;
    GF_TRY(d0.alloc(n * in_bytes));
    const bool extra = op == GF_MATH_DIV || op == GF_MATH_BINOM_DRAW || op == GF_MATH_POIS_DRAW;
    GF_TRY(d1.alloc(extra ? n * 4 : 1));
    GF_TRY(d2.alloc(extra ? n * 4 : 1));
    GF_TRY(d3.alloc(extra ? n * 4 : 1));
    GF_TRY(dout.alloc(n * out_bytes));
    GF_HIP(hipMemcpy(d0.p, in0, n * in_bytes, hipMemcpyHostToDevice));
    if (in1) GF_HIP(hipMemcpy(d1.p, in1, n * 4, hipMemcpyHostToDevice));
    if (in2) GF_HIP(hipMemcpy(d2.p, in2, n * 4, hipMemcpyHostToDevice));
    if (in3) GF_HIP(hipMemcpy(d3.p, in3, n * 4, hipMemcpyHostToDevice));
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_math, dim3(grid), dim3(256), 0, 0, op, d0.p, d1.p, d2.p, d3.p, dout.p, n);
    GF_HIP(hipGetLastError());
    GF_HIP(hipMemcpy(out, dout.p, n * out_bytes, hipMemcpyDeviceToHost));
    return GF_OK;
}
