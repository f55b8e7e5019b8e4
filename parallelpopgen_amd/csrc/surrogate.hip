/*
 * surrogate.hip -- learned surrogate of the simulator's sample SFS
 * (no reference counterpart).
 *
 * An ensemble of E multilayer perceptrons, each
 *     6 inputs (padded to 64) -> 4 hidden layers of 512 (GELU) -> 100 outputs
 * (99 softmax logits for the n = 100 SFS bins 1..99, one log-total head),
 * BF16 weights and activations, FP32 accumulation, trained by this file's
 * forward, backward and Adam kernels.  Every matrix product runs on the
 * gfx950 matrix cores (v_mfma_f32_32x32x16_bf16): k_gemm's global->LDS
 * stage handles the three operand layouts of forward (X W^T), input
 * gradient (dZ W) and weight gradient (dZ^T X) at training batch sizes;
 * k_gemm_nt runs the forward products of large batches (prediction); the
 * ensemble members are the grid's z dimension.
 */
#include "gf_runtime.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace gf {
namespace sg {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int IN_PAD = 64;    /* input features padded to one K tile */
constexpr int HID = 512;
constexpr int NHID = 4;
constexpr int OUT = 100;      /* 99 bins + log-total */
constexpr int OUT_PAD = 128;
constexpr int NBINS = 99;
constexpr int TILE = 64;
constexpr int KSTEP = 32;
constexpr int LDSK = KSTEP + 8; /* padded LDS row (bf16 elements) */

/* parameter layout of one model, in floats */
struct Layout {
    size_t W[NHID + 1], b[NHID + 1];
    int in[NHID + 1], out[NHID + 1];
    size_t total;
    Layout() {
        size_t o = 0;
        for (int l = 0; l <= NHID; l++) {
            in[l] = l == 0 ? IN_PAD : HID;
            out[l] = l == NHID ? OUT_PAD : HID;
            W[l] = o; o += (size_t)in[l] * out[l];
            b[l] = o; o += out[l];
        }
        total = o;
    }
};

/* GELU, tanh form: 0.5 x (1 + tanh(u)), u = c (x + 0.044715 x^3), c =
 * sqrt(2/pi).  With sigma(v) = 1 / (1 + e^-v), 0.5 (1 + tanh(u)) =
 * sigma(2u), so gelu(x) = x sigma(2u), evaluated with the hardware exp2 and
 * reciprocal (about 1e-6 relative; the activations are then rounded to
 * bfloat16).  e^-2u = 2^(x (k1 + k2 x^2)) with k1 = -2 c log2(e) and k2 =
 * 0.044715 k1: five operations besides exp2 and the reciprocal.  e^-2u
 * overflows to inf for u << 0, where sigma -> 0. */
__device__ __forceinline__ float sigma2u(float x) {
    const float k1 = -2.302208198f, k2 = -0.1029432396f;
    const float t = x * __builtin_fmaf(k2, x * x, k1);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}

__device__ __forceinline__ float gelu(float x) { return x * sigma2u(x); }

/* d/dx x sigma(2u) = sigma + 2 x sigma (1 - sigma) u' */
__device__ __forceinline__ float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float sg = sigma2u(x);
    const float du = c * (1.0f + 3.0f * 0.044715f * x * x);
    return sg + 2.0f * x * sg * (1.0f - sg) * du;
}

enum Epi { EPI_HIDDEN = 0, EPI_OUTPUT = 1, EPI_DX = 2, EPI_DW = 3 };

struct GemmArgs {
    const bf16 *A; int64_t lda; int a_km; int64_t sA;   /* A(m,k) = a_km ? A[k*lda+m] : A[m*lda+k] */
    const bf16 *B; int64_t ldb; int b_nk; int64_t sB;   /* B(k,n) = b_nk ? B[n*ldb+k] : B[k*ldb+n] */
    int M, N, K;
    /* epilogue */
    const float *bias; int64_t sBias;
    float *Cf; int64_t ldc; int64_t sCf;                 /* fp32 output (Z, logits, dW) */
    bf16 *Cb; int64_t sCb;                               /* bf16 output (H or dZ) */
    const float *Zprev; int64_t sZ;                      /* pre-activation for GELU' */
    int epi;
};

/* C = A B (+ epilogue), 64x64 tile per workgroup of 4 waves; each wave owns
 * a 32x32 sub-tile accumulated by 32x32x16 BF16 MFMAs from LDS. */
__global__ __launch_bounds__(256) void k_gemm(const GemmArgs g) {
    __shared__ bf16 As[TILE][LDSK];   /* [m][k] */
    __shared__ bf16 Bs[TILE][LDSK];   /* [n][k] */
    const int z = blockIdx.z;
    const bf16 *A = g.A + z * g.sA;
    const bf16 *B = g.B + z * g.sB;
    const int m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    f32x16 acc = {};
    for (int k0 = 0; k0 < g.K; k0 += KSTEP) {
        /* ---- global -> LDS: 64 x 32 elements of each operand ---- */
        if (!g.a_km) {
            const int m = t >> 2, kc = (t & 3) * 8;
            const bf16x8 v = *(const bf16x8 *)(A + (int64_t)(m0 + m) * g.lda + k0 + kc);
            *(bf16x8 *)&As[m][kc] = v;
        } else {
            const int k = t >> 3, mc = (t & 7) * 8;
            const bf16x8 v = *(const bf16x8 *)(A + (int64_t)(k0 + k) * g.lda + m0 + mc);
#pragma unroll
            for (int q = 0; q < 8; q++) As[mc + q][k] = v[q];
        }
        if (g.b_nk) {
            const int n = t >> 2, kc = (t & 3) * 8;
            const bf16x8 v = *(const bf16x8 *)(B + (int64_t)(n0 + n) * g.ldb + k0 + kc);
            *(bf16x8 *)&Bs[n][kc] = v;
        } else {
            const int k = t >> 3, nc = (t & 7) * 8;
            const bf16x8 v = *(const bf16x8 *)(B + (int64_t)(k0 + k) * g.ldb + n0 + nc);
#pragma unroll
            for (int q = 0; q < 8; q++) Bs[nc + q][k] = v[q];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KSTEP; kk += 16) {
            const bf16x8 a = *(const bf16x8 *)&As[wm * 32 + (lane & 31)][kk + 8 * (lane >> 5)];
            const bf16x8 b = *(const bf16x8 *)&Bs[wn * 32 + (lane & 31)][kk + 8 * (lane >> 5)];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    /* ---- epilogue: accumulator r of lane l is C[8(r/4) + 4(l/32) + r%4][l%32] ---- */
    const int n = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t o = (int64_t)m * g.ldc + n;
        const float v = acc[r];
        if (g.epi == EPI_HIDDEN) {
            const float zz = v + g.bias[z * g.sBias + n];
            g.Cf[z * g.sCf + o] = zz;
            g.Cb[z * g.sCb + o] = (bf16)gelu(zz);
        } else if (g.epi == EPI_OUTPUT) {
            g.Cf[z * g.sCf + o] = v + g.bias[z * g.sBias + n];
        } else if (g.epi == EPI_DX) {
            g.Cb[z * g.sCb + o] = (bf16)(v * gelu_grad(g.Zprev[z * g.sZ + o]));
        } else {
            g.Cf[z * g.sCf + o] = v;
        }
    }
}

/* ---- large forward products: C = A B^T with both operands K-contiguous ----
 * (the forward pass and prediction at batches of thousands of rows).  A
 * 128 x 128 tile per workgroup of 4 waves, each wave a 64 x 64 quadrant of
 * 2 x 2 blocks of v_mfma_f32_32x32x16_bf16; K in stages of 64 through two
 * LDS buffers, fed from two register sets: stage s + 2's global loads are
 * issued before stage s's MFMAs, and stage s + 1 (loaded during stage
 * s - 1) is stored to the other buffer after them, one barrier per stage.
 * The register sets and accumulators take 231 VGPRs, held to two waves per
 * SIMD.  Each output element accumulates its k = 0, 16, 32, ... products
 * in the same order as k_gemm, so both kernels give the same bits.
 *
 * Tiles are dealt XCD-aware: the hardware hands workgroup b to XCD b % 8,
 * so tile L is given to the workgroup that puts the tiles of one stretch of
 * L (the N tiles of an M tile, fastest) on one XCD, where they share the A
 * rows in that XCD's L2. */
constexpr int GT = 128;
constexpr int GK = 64;
constexpr int GLDS = GK + 8;   /* padded LDS row: conflict-free 16-byte fragment reads */
constexpr int GR = GK / 8;     /* threads per tile row in the global -> LDS copy */
constexpr int GRS = 256 / GR;  /* rows one pass of the workgroup copies */
constexpr int GP = GT / GRS;   /* passes per operand and stage */

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_nt(const GemmArgs g, int MT, int NT, int T) {
    /* operand stages; after the K loop the same LDS stages the output tile */
    __shared__ __attribute__((aligned(16))) bf16 sm[2][2][GT][GLDS];
    auto &As = sm[0];
    auto &Bs = sm[1];
    const int b = (int)blockIdx.x;
    const int xcd = b & 7, r = b >> 3;
    const int per = T >> 3, rem = T & 7;
    const int L = xcd < rem ? xcd * (per + 1) + r : rem * (per + 1) + (xcd - rem) * per + r;
    const int nt = L % NT, mt = (L / NT) % MT, z = L / (NT * MT);
    const bf16 *A = g.A + z * g.sA;
    const bf16 *B = g.B + z * g.sB;
    const int m0 = mt * GT, n0 = nt * GT;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    /* global -> LDS: each thread moves rows lr + GRS i (i < GP), 8 bf16 at column lk */
    const int lr = t / GR, lk = (t % GR) * 8;
    const bf16 *pa = A + (int64_t)(m0 + lr) * g.lda + lk;
    const bf16 *pb = B + (int64_t)(n0 + lr) * g.ldb + lk;
    const int64_t da = GRS * g.lda, db = GRS * g.ldb;
    /* stage s + 1's operands wait in registers while stage s is in LDS:
     * two register sets, loads issued two stages ahead */
    bf16x8 ra[2][GP], rb[2][GP];
    const int ns = g.K / GK;
    auto load = [&](bf16x8 (&xa)[GP], bf16x8 (&xb)[GP], int k1) {
#pragma unroll
        for (int i = 0; i < GP; i++) {
            xa[i] = *(const bf16x8 *)(pa + i * da + k1);
            xb[i] = *(const bf16x8 *)(pb + i * db + k1);
        }
    };
    auto store = [&](const bf16x8 (&xa)[GP], const bf16x8 (&xb)[GP], int buf) {
#pragma unroll
        for (int i = 0; i < GP; i++) {
            *(bf16x8 *)&As[buf][lr + GRS * i][lk] = xa[i];
            *(bf16x8 *)&Bs[buf][lr + GRS * i][lk] = xb[i];
        }
    };
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x16{};
    const int fr = lane & 31, fk = 8 * (lane >> 5);
    auto compute = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; i++) fa[i] = *(const bf16x8 *)&As[buf][wm * 64 + i * 32 + fr][kk + fk];
#pragma unroll
            for (int j = 0; j < 2; j++) fb[j] = *(const bf16x8 *)&Bs[buf][wn * 64 + j * 32 + fr][kk + fk];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
    };
    /* every load of the steady-state loop is unconditional, so that the
     * compiler's wait before a store counts only the older register set
     * (a load under a branch makes it wait for both) */
    load(ra[0], rb[0], 0);
    load(ra[1], rb[1], ns > 1 ? GK : 0);   /* one stage: a harmless reload */
    store(ra[0], rb[0], 0);
    __syncthreads();
    /* at the top of each pass: stage s in LDS buffer 0, stage s + 1 in
     * flight into register set 1, set 0 free (its stage was stored in the
     * previous pass).  A buffer is written only after the barrier that
     * ended the last stage reading it. */
    int s = 0;
    for (; s + 3 < ns; s += 2) {
        load(ra[0], rb[0], (s + 2) * GK);
        __builtin_amdgcn_sched_barrier(0);   /* keep the loads ahead of the MFMAs */
        compute(0);
        store(ra[1], rb[1], 1);
        __syncthreads();
        load(ra[1], rb[1], (s + 3) * GK);
        __builtin_amdgcn_sched_barrier(0);   /* keep the loads ahead of the MFMAs */
        compute(1);
        store(ra[0], rb[0], 0);
        __syncthreads();
    }
    const int left = ns - s;   /* 1, 2 or 3 stages */
    if (left == 3) {
        load(ra[0], rb[0], (s + 2) * GK);
        __builtin_amdgcn_sched_barrier(0);   /* keep the loads ahead of the MFMAs */
        compute(0);
        store(ra[1], rb[1], 1);
        __syncthreads();
        compute(1);
        store(ra[0], rb[0], 0);
        __syncthreads();
        compute(0);
    } else if (left == 2) {
        compute(0);
        store(ra[1], rb[1], 1);
        __syncthreads();
        compute(1);
    } else {
        compute(0);
    }
    __syncthreads();
    /* epilogue: accumulator r of lane l is C[8(r/4) + 4(l/32) + r%4][l%32]
     * of its block.  The tile goes through LDS (free after the loop's last
     * barrier) so that global stores are whole 16-byte row pieces: a lane's
     * accumulators run down a column.  Row strides of 144 bf16 / 136 fp32
     * put the two half-waves (rows 4 apart) on different LDS banks. */
    if (g.epi == EPI_HIDDEN && g.Cf) {   /* pre-activations kept as well: direct stores */
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int n = n0 + wn * 64 + j * 32 + fr;
            const float bias = g.bias[z * g.sBias + n];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int m = m0 + wm * 64 + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    const int64_t o = (int64_t)m * g.ldc + n;
                    const float v = acc[i][j][q] + bias;
                    g.Cf[z * g.sCf + o] = v;
                    g.Cb[z * g.sCb + o] = (bf16)gelu(v);
                }
        }
        return;
    }
    if (g.epi == EPI_HIDDEN) {
        constexpr int S = GT + 16;
        static_assert(GT * S * sizeof(bf16) <= sizeof(sm), "bf16 output tile fits the operand LDS");
        bf16 *cs = &sm[0][0][0][0];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int c = wn * 64 + j * 32 + fr;
            const float bias = g.bias[z * g.sBias + n0 + c];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int rr = wm * 64 + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    cs[rr * S + c] = (bf16)gelu(acc[i][j][q] + bias);
                }
        }
        __syncthreads();
        bf16 *dst = g.Cb + z * g.sCb;
        for (int p = t; p < GT * (GT / 8); p += 256) {
            const int rr = p / (GT / 8), c = (p % (GT / 8)) * 8;
            *(bf16x8 *)&dst[(int64_t)(m0 + rr) * g.ldc + n0 + c] = *(const bf16x8 *)&cs[rr * S + c];
        }
    } else {
        constexpr int S = GT + 8;
        static_assert(GT * S * sizeof(float) <= sizeof(sm), "fp32 output tile fits the operand LDS");
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        float *cs = reinterpret_cast<float *>(&sm[0][0][0][0]);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int c = wn * 64 + j * 32 + fr;
            const float bias = g.bias ? g.bias[z * g.sBias + n0 + c] : 0.0f;
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int rr = wm * 64 + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    cs[rr * S + c] = acc[i][j][q] + bias;
                }
        }
        __syncthreads();
        float *dst = g.Cf + z * g.sCf;
        for (int p = t; p < GT * (GT / 4); p += 256) {
            const int rr = p / (GT / 4), c = (p % (GT / 4)) * 4;
            *(f32x4 *)&dst[(int64_t)(m0 + rr) * g.ldc + n0 + c] = *(const f32x4 *)&cs[rr * S + c];
        }
    }
}

/* softmax cross-entropy on the 99 bins + weighted squared error on the
 * log-total head; writes dLogits (bf16, /batch) and the summed loss */
__global__ void k_loss(const float *__restrict__ logits, int64_t sL, const float *__restrict__ y,
                       const float *__restrict__ tot, int B, float w_tot, bf16 *__restrict__ dlog, int64_t sD,
                       float *__restrict__ loss) {
    const int z = blockIdx.y;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float *L = logits + z * sL + (int64_t)b * OUT_PAD;
    bf16 *D = dlog + z * sD + (int64_t)b * OUT_PAD;
    const float *Y = y + (int64_t)b * NBINS;
    float mx = L[0];
    for (int k = 1; k < NBINS; k++) mx = fmaxf(mx, L[k]);
    float s = 0.f;
    for (int k = 0; k < NBINS; k++) s += __expf(L[k] - mx);
    const float ls = logf(s) + mx;
    float ce = 0.f;
    const float invB = 1.0f / (float)B;
    for (int k = 0; k < NBINS; k++) {
        const float p = __expf(L[k] - ls);
        ce -= Y[k] * (L[k] - ls);
        D[k] = (bf16)((p - Y[k]) * invB);
    }
    const float e = L[NBINS] - tot[b];
    D[NBINS] = (bf16)(2.0f * w_tot * e * invB);
    for (int k = OUT; k < OUT_PAD; k++) D[k] = (bf16)0.0f;
    atomicAdd(&loss[z], (ce + w_tot * e * e) * invB);
}

/* column sums of a bf16 [B][N] matrix -> fp32 [N] (bias gradients), N a
 * multiple of 64.  A workgroup sums 64 columns: 32 row lanes, each of 32
 * threads reading two adjacent columns (one 4-byte load) of every 32nd
 * row, then the 32 partial sums of a column are added in a fixed order, so
 * the result does not depend on timing. */
constexpr int COLSUM_LANES = 32;
static_assert(HID % 64 == 0 && OUT_PAD % 64 == 0, "k_colsum sums 64 columns per workgroup");
__global__ __launch_bounds__(1024) void k_colsum(const bf16 *__restrict__ dz, int64_t sD, int B, int N,
                                                 float *__restrict__ out, int64_t sO) {
    __shared__ float part[COLSUM_LANES][64 + 1];
    const int z = blockIdx.y;
    const int cp = (int)threadIdx.x & 31;
    const int rl = (int)threadIdx.x >> 5;
    const int n0 = (int)blockIdx.x * 64;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(dz + z * sD + n0) + cp;
    const int64_t row = N / 2;   /* in 4-byte words */
    float s0 = 0.f, s1 = 0.f;
#pragma unroll 4
    for (int b = rl; b < B; b += COLSUM_LANES) {
        const uint32_t v = src[(int64_t)b * row];
        s0 += __builtin_bit_cast(float, v << 16);            /* column n0 + 2 cp */
        s1 += __builtin_bit_cast(float, v & 0xffff0000u);    /* column n0 + 2 cp + 1 */
    }
    part[rl][2 * cp] = s0;
    part[rl][2 * cp + 1] = s1;
    __syncthreads();
    if (threadIdx.x < 64) {
        float s = 0.f;
        for (int k = 0; k < COLSUM_LANES; k++) s += part[k][threadIdx.x];
        out[z * sO + n0 + (int)threadIdx.x] = s;
    }
}

/* Adam on fp32 master weights; refreshes the bf16 copy */
__global__ void k_adam(float *__restrict__ w, bf16 *__restrict__ wb, const float *__restrict__ gr,
                       float *__restrict__ m, float *__restrict__ v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float g = gr[i];
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        wb[i] = (bf16)wi;
    }
}

__global__ void k_to_bf16(const float *__restrict__ in, bf16 *__restrict__ out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (bf16)in[i];
}

/* rows ord[0..B) of the staged training set -> the batch buffers */
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ Y, const float *__restrict__ T,
                         const int32_t *__restrict__ ord, int B, float *__restrict__ x, float *__restrict__ y,
                         float *__restrict__ t) {
    const int b = blockIdx.x;
    const int64_t src = ord[b];
    for (int k = threadIdx.x; k < NBINS; k += blockDim.x) y[(int64_t)b * NBINS + k] = Y[src * NBINS + k];
    if (threadIdx.x < 6) x[(int64_t)b * 6 + threadIdx.x] = X[src * 6 + threadIdx.x];
    if (threadIdx.x == 0) t[b] = T[src];
}

/* inputs [nvalid][6] fp32 -> [B][64] bf16, zero padded (columns past nin,
 * rows past nvalid) */
__global__ void k_pack_inputs(const float *__restrict__ x, int B, int nvalid, int nin, bf16 *__restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * IN_PAD) return;
    const int b = (int)(i / IN_PAD), k = (int)(i % IN_PAD);
    out[i] = (bf16)(k < nin && b < nvalid ? x[(int64_t)b * nin + k] : 0.0f);
}

/* ensemble statistics of the predictions: per-model softmax, mean SFS,
 * mean log-total and the ensemble disagreement (largest relative standard
 * deviation across models over the bins with mean mass >= 1e-3).  One wave
 * per input: lane l holds bins l and l + 64, so each model's logit row is
 * read once, coalesced, and the row's max and sum are wave reductions. */
__global__ __launch_bounds__(256) void k_predict_stats(const float *__restrict__ logits, int64_t sL, int E, int B,
                                                       float *__restrict__ sfs, float *__restrict__ tot,
                                                       float *__restrict__ dis) {
    const int lane = threadIdx.x & 63;
    const int b = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (b >= B) return;   /* the whole wave */
    const int k1 = lane + 64;
    const bool h1 = k1 < NBINS;
    float mean0 = 0.f, mean1 = 0.f, sq0 = 0.f, sq1 = 0.f, tsum = 0.f;
    for (int z = 0; z < E; z++) {
        const float *L = logits + z * sL + (int64_t)b * OUT_PAD;
        const float l0 = L[lane];
        const float l1 = h1 ? L[k1] : l0;
        float mx = fmaxf(l0, l1);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
        const float e0 = __expf(l0 - mx), e1 = h1 ? __expf(l1 - mx) : 0.f;
        float sum = e0 + e1;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        const float p0 = e0 / sum, p1 = e1 / sum;
        mean0 += p0; sq0 += p0 * p0;
        mean1 += p1; sq1 += p1 * p1;
        tsum += L[NBINS];
    }
    float worst = 0.f;
    {
        const float mu = mean0 / E, var = fmaxf(sq0 / E - mu * mu, 0.f);
        sfs[(int64_t)b * NBINS + lane] = mu;
        if (mu >= 1e-3f) worst = sqrtf(var) / mu;
    }
    if (h1) {
        const float mu = mean1 / E, var = fmaxf(sq1 / E - mu * mu, 0.f);
        sfs[(int64_t)b * NBINS + k1] = mu;
        if (mu >= 1e-3f) worst = fmaxf(worst, sqrtf(var) / mu);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) worst = fmaxf(worst, __shfl_xor(worst, off));
    if (lane == 0) {
        if (tot) tot[b] = tsum / E;
        if (dis) dis[b] = worst;
    }
}

} /* namespace sg */
} /* namespace gf */

using namespace gf;
using namespace gf::sg;

struct gf_sg_model {
    int E = 4, Bmax = 0;
    Layout lay;
    OwnedStream st;   /* before the buffers: they are released while it is alive */
    OwnedStream cst;  /* prediction read-back, beside the next batch's compute on st */
    DevBuf<float> w, g, m, v;       /* [E][total] */
    DevBuf<bf16> wb;
    DevBuf<bf16> x;                 /* [B][64] */
    DevBuf<float> Z;                /* [E][NHID][B][512] */
    DevBuf<bf16> H;                 /* [E][NHID][B][512] */
    DevBuf<float> logits;           /* [E][B][128] */
    DevBuf<bf16> dlog;              /* [E][B][128] */
    DevBuf<bf16> dz0, dz1;          /* [E][B][512] */
    DevBuf<float> y, tot, loss, xin;
    DevBuf<float> psfs, ptot, pdis; /* two batches' prediction outputs */
    int64_t step = 0;
    float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
};

namespace gf {
namespace sg {

static int gemm(gf_sg_model &S, GemmArgs a) {
    if (a.M % TILE || a.N % TILE || a.K % KSTEP) { set_error("surrogate gemm: shape not a multiple of the tile"); return GF_ERR_ARG; }
    const bool nt = !a.a_km && a.b_nk && (a.epi == EPI_HIDDEN || a.epi == EPI_OUTPUT);
    if (nt && a.M >= 4096 && a.M % GT == 0 && a.N % GT == 0 && a.K % GK == 0) {
        const int MT = a.M / GT, NT = a.N / GT, T = MT * NT * S.E;
        hipLaunchKernelGGL(k_gemm_nt, dim3(T), dim3(256), 0, S.st, a, MT, NT, T);
    } else {
        if (a.epi == EPI_HIDDEN && !a.Cf) { set_error("surrogate gemm: small products keep Z"); return GF_ERR_ARG; }
        hipLaunchKernelGGL(k_gemm, dim3(a.N / TILE, a.M / TILE, S.E), dim3(256), 0, S.st, a);
    }
    GF_HIP(hipGetLastError());
    return GF_OK;
}

static int ensure_batch(gf_sg_model &S, int B) {
    if (B <= S.Bmax) return GF_OK;
    const int E = S.E;
    GF_TRY(S.x.alloc((size_t)B * IN_PAD));
    GF_TRY(S.Z.alloc((size_t)E * NHID * B * HID));
    GF_TRY(S.H.alloc((size_t)E * NHID * B * HID));
    GF_TRY(S.logits.alloc((size_t)E * B * OUT_PAD));
    GF_TRY(S.dlog.alloc((size_t)E * B * OUT_PAD));
    GF_TRY(S.dz0.alloc((size_t)E * B * HID));
    GF_TRY(S.dz1.alloc((size_t)E * B * HID));
    GF_TRY(S.y.alloc((size_t)B * NBINS));
    GF_TRY(S.tot.alloc(B));
    GF_TRY(S.xin.alloc((size_t)B * 6));
    GF_TRY(S.psfs.alloc((size_t)2 * B * NBINS));
    GF_TRY(S.ptot.alloc((size_t)2 * B));
    GF_TRY(S.pdis.alloc((size_t)2 * B));
    S.Bmax = B;
    return GF_OK;
}

/* forward pass of the whole ensemble on S.x (B rows); keep_z: also store
 * the hidden layers' pre-activations Z (the backward pass reads them) */
static int forward(gf_sg_model &S, int B, bool keep_z) {
    keep_z = keep_z || B < 4096 || B % GT != 0;   /* k_gemm always stores Z */
    const Layout &L = S.lay;
    const int64_t sW = (int64_t)L.total;
    const int64_t sAct = (int64_t)NHID * B * HID;
    for (int l = 0; l <= NHID; l++) {
        GemmArgs a{};
        if (l == 0) { a.A = S.x.p; a.lda = IN_PAD; a.sA = 0; }
        else { a.A = S.H.p + (int64_t)(l - 1) * B * HID; a.lda = HID; a.sA = sAct; }
        a.a_km = 0;
        a.B = S.wb.p + L.W[l]; a.ldb = L.in[l]; a.b_nk = 1; a.sB = sW;
        a.M = B; a.N = L.out[l]; a.K = L.in[l];
        a.bias = S.w.p + L.b[l]; a.sBias = sW;
        if (l < NHID) {
            a.epi = EPI_HIDDEN;
            a.Cf = keep_z ? S.Z.p + (int64_t)l * B * HID : nullptr; a.sCf = sAct; a.ldc = HID;
            a.Cb = S.H.p + (int64_t)l * B * HID; a.sCb = sAct;
        } else {
            a.epi = EPI_OUTPUT;
            a.Cf = S.logits.p; a.sCf = (int64_t)B * OUT_PAD; a.ldc = OUT_PAD;
        }
        GF_TRY(gemm(S, a));
    }
    return GF_OK;
}

/* loss + backward + Adam for one batch already in S.x / S.y / S.tot */
static int train_step(gf_sg_model &S, int B, float lr, float w_tot, bool update) {
    const Layout &L = S.lay;
    const int E = S.E;
    const int64_t sW = (int64_t)L.total;
    const int64_t sAct = (int64_t)NHID * B * HID;
    GF_TRY(forward(S, B, true));
    hipLaunchKernelGGL(k_loss, dim3((B + 127) / 128, E), dim3(128), 0, S.st, S.logits.p, (int64_t)B * OUT_PAD,
                       S.y.p, S.tot.p, B, w_tot, S.dlog.p, (int64_t)B * OUT_PAD, S.loss.p);
    GF_HIP(hipGetLastError());
    /* dY of layer l lives in `dcur` ([E][B][out_l] bf16) */
    const bf16 *dcur = S.dlog.p;
    int64_t sD = (int64_t)B * OUT_PAD;
    for (int l = NHID; l >= 0; l--) {
        const int O = L.out[l], I = L.in[l];
        /* dW_l = dY^T X_l : M = O, N = I, K = B */
        GemmArgs a{};
        a.A = dcur; a.lda = O; a.a_km = 1; a.sA = sD;
        if (l == 0) { a.B = S.x.p; a.ldb = IN_PAD; a.sB = 0; }
        else { a.B = S.H.p + (int64_t)(l - 1) * B * HID; a.ldb = HID; a.sB = sAct; }
        a.b_nk = 0;
        a.M = O; a.N = I; a.K = B;
        a.epi = EPI_DW; a.Cf = S.g.p + L.W[l]; a.ldc = I; a.sCf = sW;
        GF_TRY(gemm(S, a));
        hipLaunchKernelGGL(k_colsum, dim3(O / 64, E), dim3(32 * COLSUM_LANES), 0, S.st, dcur, sD, B, O,
                           S.g.p + L.b[l], sW);
        GF_HIP(hipGetLastError());
        if (l == 0) break;
        /* dZ_(l-1) = (dY W_l) * gelu'(Z_(l-1)) : M = B, N = I, K = O */
        bf16 *dnext = (dcur == S.dz0.p) ? S.dz1.p : S.dz0.p;
        GemmArgs b{};
        b.A = dcur; b.lda = O; b.a_km = 0; b.sA = sD;
        b.B = S.wb.p + L.W[l]; b.ldb = I; b.b_nk = 0; b.sB = sW;
        b.M = B; b.N = I; b.K = O;
        b.epi = EPI_DX; b.Cb = dnext; b.sCb = (int64_t)B * HID; b.ldc = HID;
        b.Zprev = S.Z.p + (int64_t)(l - 1) * B * HID; b.sZ = sAct;
        GF_TRY(gemm(S, b));
        dcur = dnext;
        sD = (int64_t)B * HID;
    }
    if (update) {
        S.step++;
        const float c1 = 1.0f - std::pow(S.b1, (float)S.step);
        const float c2 = 1.0f - std::pow(S.b2, (float)S.step);
        const int64_t n = (int64_t)E * L.total;
        hipLaunchKernelGGL(k_adam, dim3(1024), dim3(256), 0, S.st, S.w.p, S.wb.p, S.g.p, S.m.p, S.v.p, n, lr, S.b1,
                           S.b2, S.eps, c1, c2);
        GF_HIP(hipGetLastError());
    }
    return GF_OK;
}

static int load_batch(gf_sg_model &S, const float *x, const float *y, const float *t, int B) {
    GF_HIP(hipMemcpyAsync(S.xin.p, x, (size_t)B * 6 * sizeof(float), hipMemcpyHostToDevice, S.st));
    if (y) GF_HIP(hipMemcpyAsync(S.y.p, y, (size_t)B * NBINS * sizeof(float), hipMemcpyHostToDevice, S.st));
    if (t) GF_HIP(hipMemcpyAsync(S.tot.p, t, (size_t)B * sizeof(float), hipMemcpyHostToDevice, S.st));
    hipLaunchKernelGGL(k_pack_inputs, dim3((B * IN_PAD + 255) / 256), dim3(256), 0, S.st, S.xin.p, B, B, 6, S.x.p);
    GF_HIP(hipGetLastError());
    return GF_OK;
}

} /* namespace sg */
} /* namespace gf */

extern "C" {

int32_t gf_sg_create(int32_t ensemble, uint64_t seed, int32_t device, gf_sg_model **out) {
    if (!out || ensemble < 1 || ensemble > 16) { set_error("gf_sg_create: invalid arguments"); return GF_ERR_ARG; }
    *out = nullptr;
    GF_TRY(select_device(device));
    std::unique_ptr<gf_sg_model> S(new gf_sg_model);
    S->E = ensemble;
    GF_HIP(S->st.create());
    StreamScope scope(S->st);
    const size_t n = (size_t)S->E * S->lay.total;
    GF_TRY(S->w.alloc(n));
    GF_TRY(S->g.alloc(n));
    GF_TRY(S->m.alloc(n));
    GF_TRY(S->v.alloc(n));
    GF_TRY(S->wb.alloc(n));
    GF_TRY(S->loss.alloc(S->E));
    /* He-style initialisation from a seeded host generator: W ~ U(-a, a),
     * a = sqrt(6 / fan_in); biases 0; padded input columns and output rows 0 */
    std::vector<float> h(n, 0.0f);
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto next = [&st]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((st >> 40) & 0xffffff) / 16777216.0f;
    };
    for (int e = 0; e < S->E; e++) {
        for (int l = 0; l <= NHID; l++) {
            const int I = S->lay.in[l], O = S->lay.out[l];
            const int real_in = l == 0 ? 6 : I, real_out = l == NHID ? OUT : O;
            const float a = std::sqrt(6.0f / (float)real_in);
            for (int o = 0; o < O; o++)
                for (int i = 0; i < I; i++) {
                    float v = (2.0f * next() - 1.0f) * a;
                    if (i >= real_in || o >= real_out) v = 0.0f;
                    h[(size_t)e * S->lay.total + S->lay.W[l] + (size_t)o * I + i] = v;
                }
        }
    }
    GF_HIP(hipMemcpy(S->w.p, h.data(), n * sizeof(float), hipMemcpyHostToDevice));
    GF_HIP(hipMemset(S->m.p, 0, n * sizeof(float)));
    GF_HIP(hipMemset(S->v.p, 0, n * sizeof(float)));
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, S->st, S->w.p, S->wb.p, (int64_t)n);
    GF_HIP(hipGetLastError());
    GF_HIP(hipStreamSynchronize(S->st));
    *out = S.release();
    return GF_OK;
}

void gf_sg_free(gf_sg_model *S) { delete S; }

int32_t gf_sg_num_params(const gf_sg_model *S) { return S ? (int32_t)S->lay.total : GF_ERR_ARG; }

int32_t gf_sg_get_params(gf_sg_model *S, float *out) {
    if (!S || !out) return GF_ERR_ARG;
    GF_HIP(hipStreamSynchronize(S->st));
    GF_HIP(hipMemcpy(out, S->w.p, (size_t)S->E * S->lay.total * sizeof(float), hipMemcpyDeviceToHost));
    return GF_OK;
}

int32_t gf_sg_set_params(gf_sg_model *S, const float *in) {
    if (!S || !in) return GF_ERR_ARG;
    const size_t n = (size_t)S->E * S->lay.total;
    GF_HIP(hipMemcpy(S->w.p, in, n * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, S->st, S->w.p, S->wb.p, (int64_t)n);
    GF_HIP(hipGetLastError());
    GF_HIP(hipStreamSynchronize(S->st));
    return GF_OK;
}

/* one forward + backward on a batch (B a multiple of 64) without updating:
 * gradients [E][total] and per-model loss, for gradient checks */
int32_t gf_sg_gradients(gf_sg_model *S, const float *x, const float *y, const float *t, int32_t B, float w_tot,
                        float *grad_out, float *loss_out) {
    if (!S || !x || !y || !t || B <= 0 || B % TILE) { set_error("gf_sg_gradients: invalid arguments"); return GF_ERR_ARG; }
    StreamScope scope(S->st);   /* buffers (re)allocated here belong to the model's stream */
    GF_TRY(ensure_batch(*S, B));
    GF_HIP(hipMemsetAsync(S->loss.p, 0, S->E * sizeof(float), S->st));
    GF_TRY(load_batch(*S, x, y, t, B));
    GF_TRY(train_step(*S, B, 0.0f, w_tot, false));
    GF_HIP(hipStreamSynchronize(S->st));
    if (grad_out) GF_HIP(hipMemcpy(grad_out, S->g.p, (size_t)S->E * S->lay.total * sizeof(float), hipMemcpyDeviceToHost));
    if (loss_out) GF_HIP(hipMemcpy(loss_out, S->loss.p, S->E * sizeof(float), hipMemcpyDeviceToHost));
    return GF_OK;
}

/* epochs over n examples (host arrays), mini-batches of B drawn in the
 * order given by `order` (n ints per epoch, or NULL for 0..n-1); returns
 * the mean loss per epoch (averaged over models) */
int32_t gf_sg_train(gf_sg_model *S, const float *x, const float *y, const float *t, int32_t n, int32_t epochs,
                    int32_t B, float lr, float w_tot, const int32_t *order, float *epoch_loss) {
    if (!S || !x || !y || !t || n < B || B <= 0 || B % TILE || epochs <= 0) { set_error("gf_sg_train: invalid arguments"); return GF_ERR_ARG; }
    StreamScope scope(S->st);   /* buffers (re)allocated here belong to the model's stream */
    GF_TRY(ensure_batch(*S, B));
    /* stage the whole training set on the device once */
    DevBuf<float> dx, dy, dt;
    DevBuf<int32_t> dord;
    GF_TRY(dx.alloc((size_t)n * 6));
    GF_TRY(dy.alloc((size_t)n * NBINS));
    GF_TRY(dt.alloc(n));
    GF_HIP(hipMemcpyAsync(dx.p, x, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice, S->st));
    GF_HIP(hipMemcpyAsync(dy.p, y, (size_t)n * NBINS * sizeof(float), hipMemcpyHostToDevice, S->st));
    GF_HIP(hipMemcpyAsync(dt.p, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, S->st));
    const int nb = n / B;
    std::vector<float> lossv(S->E);
    if (order) {
        GF_TRY(dord.alloc((size_t)epochs * n));
        GF_HIP(hipMemcpyAsync(dord.p, order, (size_t)epochs * n * sizeof(int32_t), hipMemcpyHostToDevice, S->st));
    }
    for (int e = 0; e < epochs; e++) {
        GF_HIP(hipMemsetAsync(S->loss.p, 0, S->E * sizeof(float), S->st));
        for (int b = 0; b < nb; b++) {
            if (order) {
                hipLaunchKernelGGL(k_gather, dim3(B), dim3(128), 0, S->st, dx.p, dy.p, dt.p,
                                   dord.p + (size_t)e * n + (size_t)b * B, B, S->xin.p, S->y.p, S->tot.p);
                GF_HIP(hipGetLastError());
            } else {
                GF_HIP(hipMemcpyAsync(S->xin.p, dx.p + (size_t)b * B * 6, (size_t)B * 6 * sizeof(float), hipMemcpyDeviceToDevice, S->st));
                GF_HIP(hipMemcpyAsync(S->y.p, dy.p + (size_t)b * B * NBINS, (size_t)B * NBINS * sizeof(float), hipMemcpyDeviceToDevice, S->st));
                GF_HIP(hipMemcpyAsync(S->tot.p, dt.p + (size_t)b * B, (size_t)B * sizeof(float), hipMemcpyDeviceToDevice, S->st));
            }
            hipLaunchKernelGGL(k_pack_inputs, dim3((B * IN_PAD + 255) / 256), dim3(256), 0, S->st, S->xin.p, B, B, 6, S->x.p);
            GF_TRY(train_step(*S, B, lr, w_tot, true));
        }
        GF_HIP(hipMemcpyAsync(lossv.data(), S->loss.p, S->E * sizeof(float), hipMemcpyDeviceToHost, S->st));
        GF_HIP(hipStreamSynchronize(S->st));
        if (epoch_loss) {
            float s = 0.f;
            for (float v : lossv) s += v;
            epoch_loss[e] = s / S->E / nb;
        }
    }
    GF_HIP(hipStreamSynchronize(S->st));
    return GF_OK;
}

/* ensemble prediction for n inputs: mean normalised SFS (bins 1..99),
 * mean log-total and the ensemble disagreement per input.  Batches of up
 * to 65,536 inputs alternate between two output buffers: while batch i is
 * computed on the model's stream, batch i - 1 is read back on the copy
 * stream (each waits for the other through events). */
int32_t gf_sg_predict(gf_sg_model *S, const float *x, int32_t n, float *sfs, float *tot, float *dis) {
    if (!S || !x || n <= 0 || !sfs) { set_error("gf_sg_predict: invalid arguments"); return GF_ERR_ARG; }
    StreamScope scope(S->st);   /* buffers (re)allocated here belong to the model's stream */
    GF_HIP(S->cst.create());
    const int B = std::min(65536, ((n + GT - 1) / GT) * GT);
    GF_TRY(ensure_batch(*S, B));
    struct Events {   /* [0..1] computed, [2..3] read back, per output buffer */
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events() { for (auto v : e) if (v) (void)hipEventDestroy(v); }
    } ev;
    for (auto &v : ev.e) GF_HIP(hipEventCreateWithFlags(&v, hipEventDisableTiming));
    std::vector<float> xb[2];   /* a batch's inputs stay untouched until its compute has started */
    for (auto &v : xb) v.assign((size_t)B * 6, 0.0f);
    const int nb = (n + B - 1) / B;
    for (int i = 0; i <= nb; i++) {
        if (i < nb) {
            const int off = i * B, cnt = std::min(B, n - off), u = i & 1;
            memcpy(xb[u].data(), x + (size_t)off * 6, (size_t)cnt * 6 * sizeof(float));
            if (cnt < B) memset(xb[u].data() + (size_t)cnt * 6, 0, (size_t)(B - cnt) * 6 * sizeof(float));
            if (i >= 2) GF_HIP(hipStreamWaitEvent(S->st, ev.e[2 + u], 0));   /* buffer u read back */
            GF_TRY(load_batch(*S, xb[u].data(), nullptr, nullptr, B));
            GF_TRY(forward(*S, B, false));
            hipLaunchKernelGGL(k_predict_stats, dim3((B + 3) / 4), dim3(256), 0, S->st, S->logits.p,
                               (int64_t)B * OUT_PAD, S->E, B, S->psfs.p + (size_t)u * B * NBINS,
                               S->ptot.p + (size_t)u * B, S->pdis.p + (size_t)u * B);
            GF_HIP(hipGetLastError());
            GF_HIP(hipEventRecord(ev.e[u], S->st));
        }
        if (i >= 1) {
            const int j = i - 1, off = j * B, cnt = std::min(B, n - off), u = j & 1;
            GF_HIP(hipStreamWaitEvent(S->cst, ev.e[u], 0));
            /* the spectra (99 floats per input) through the pinned staging path */
            GF_TRY(readback(sfs + (size_t)off * NBINS, S->psfs.p + (size_t)u * B * NBINS,
                            (size_t)cnt * NBINS * sizeof(float), S->cst));
            if (tot) GF_HIP(hipMemcpyAsync(tot + off, S->ptot.p + (size_t)u * B, (size_t)cnt * sizeof(float),
                                           hipMemcpyDeviceToHost, S->cst));
            if (dis) GF_HIP(hipMemcpyAsync(dis + off, S->pdis.p + (size_t)u * B, (size_t)cnt * sizeof(float),
                                           hipMemcpyDeviceToHost, S->cst));
            GF_HIP(hipEventRecord(ev.e[2 + u], S->cst));
        }
    }
    GF_HIP(hipStreamSynchronize(S->cst));
    GF_HIP(hipStreamSynchronize(S->st));
    return GF_OK;
}

/* gf_sg_predict on device memory: x, sfs, log_total and disagreement are
 * device pointers on the model's device (the last two may be NULL); the
 * batches run back to back on the model's stream, each writing its rows of
 * the outputs, and the call returns when they are done. */
int32_t gf_sg_predict_device(gf_sg_model *S, const float *x, int32_t n, float *sfs, float *tot, float *dis) {
    if (!S || !x || n <= 0 || !sfs) { set_error("gf_sg_predict_device: invalid arguments"); return GF_ERR_ARG; }
    int dev = -1;
    GF_HIP(hipGetDevice(&dev));
    for (const void *p : {(const void *)x, (const void *)sfs, (const void *)tot, (const void *)dis}) {
        if (!p) continue;
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != dev) {
            (void)hipGetLastError();
            set_error("gf_sg_predict_device: x, sfs, log_total and disagreement must be memory of the current device");
            return GF_ERR_ARG;
        }
    }
    StreamScope scope(S->st);
    const int B = std::min(65536, ((n + GT - 1) / GT) * GT);
    GF_TRY(ensure_batch(*S, B));
    for (int off = 0; off < n; off += B) {
        const int cnt = std::min(B, n - off);
        hipLaunchKernelGGL(k_pack_inputs, dim3((B * IN_PAD + 255) / 256), dim3(256), 0, S->st, x + (size_t)off * 6, B,
                           cnt, 6, S->x.p);
        GF_HIP(hipGetLastError());
        GF_TRY(forward(*S, B, false));
        hipLaunchKernelGGL(k_predict_stats, dim3((cnt + 3) / 4), dim3(256), 0, S->st, S->logits.p, (int64_t)B * OUT_PAD,
                           S->E, cnt, sfs + (size_t)off * NBINS, tot ? tot + off : nullptr, dis ? dis + off : nullptr);
        GF_HIP(hipGetLastError());
    }
    GF_HIP(hipStreamSynchronize(S->st));
    return GF_OK;
}

/* raw ensemble logits for n inputs (n a multiple of 64): [E][n][128] */
int32_t gf_sg_forward(gf_sg_model *S, const float *x, int32_t n, float *logits) {
    if (!S || !x || !logits || n <= 0 || n % TILE) { set_error("gf_sg_forward: invalid arguments"); return GF_ERR_ARG; }
    StreamScope scope(S->st);   /* buffers (re)allocated here belong to the model's stream */
    GF_TRY(ensure_batch(*S, n));
    GF_TRY(load_batch(*S, x, nullptr, nullptr, n));
    GF_TRY(forward(*S, n, false));
    GF_HIP(hipStreamSynchronize(S->st));
    for (int e = 0; e < S->E; e++)
        GF_HIP(hipMemcpy(logits + (size_t)e * n * OUT_PAD, S->logits.p + (size_t)e * n * OUT_PAD, (size_t)n * OUT_PAD * sizeof(float), hipMemcpyDeviceToHost));
    return GF_OK;
}

} /* extern "C" */
