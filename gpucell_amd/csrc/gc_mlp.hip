// gc_mlp.hip -- the per-vertex MLP of the learned membrane substepper
// (36 -> 128 -> 128 -> 128 -> 24, tanh hidden units) on the matrix cores.
//
// Every product runs on v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulation.  Operand layout of that instruction for a 16 x 16 tile D =
// A (16 x 32) * B (32 x 16): lane l = 16 g + i holds A[i][8g .. 8g+7] and
// B[8g .. 8g+7][i] (eight BF16 in four VGPRs), and receives D[4g + r][i],
// r = 0..3.  Both operands are therefore read as 16-byte rows with the
// contraction index contiguous: weights are kept as W[out][in] (and W^T for
// the backward pass), activations as [sample][feature].
//
// Inference (k_mlp_fused): one workgroup of four waves keeps all weights in
// LDS (132 KB) and streams 64-sample tiles through the four layers; each
// wave owns 16 samples, whose activations stay in an LDS ping-pong buffer,
// so HBM sees only the 36 inputs and 24 outputs per sample.
//
// Training: the same fused forward stores the BF16 activations; the
// backward pass is a BF16 NT GEMM (k_gemm_nt) per layer for the input
// gradient (with the tanh' epilogue) and a split-K NT GEMM over the batch
// for the weight gradient (partials reduced in a fixed order, so gradients
// are deterministic); Adam updates FP32 master weights, from which the
// padded BF16 copies are regenerated.
#include "gc_internal.h"
#include "../../include/gpucell/surrogate.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

using namespace gc;

namespace {

constexpr int IN = GC_MLP_IN, INP = 64, H = GC_MLP_HIDDEN, OUT = GC_MLP_OUT, OUTP = 32;
constexpr int LDI = INP + 8;   // LDS row strides in BF16 elements (16-byte multiples,
constexpr int LDW = H + 8;     // one extra 16-byte chunk against bank conflicts)
constexpr int NPARAM = H * IN + H + 2 * (H * H + H) + OUT * H + OUT;
constexpr int MLP_BS = 256;
// fused forward: 8 waves (two per SIMD) share the weights in LDS; each wave
// keeps its 16 rows' activations in one buffer, overwritten layer by layer
constexpr int FWD_BS = 512;
constexpr size_t FUSED_LDS_W = sizeof(uint16_t) * ((size_t)H * LDI + 2 * (size_t)H * LDW +
                                                   (size_t)OUTP * LDW + (FWD_BS / 64) * 16 * (size_t)LDW);
constexpr int NBIAS = 3 * H + OUTP;   // b1, b2, b3, b4 (padded), FP32, after the BF16 blocks
constexpr size_t FUSED_LDS = FUSED_LDS_W + sizeof(float) * NBIAS;
static_assert(FUSED_LDS_W % 16 == 0, "bias block alignment");

// Offsets of the blocks of the flat parameter vector.
constexpr int OW1 = 0, OB1 = OW1 + H * IN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H * H, OW4 = OB3 + H, OB4 = OW4 + OUT * H;
static_assert(OB4 + OUT == NPARAM, "parameter layout");

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(4))) float f32x4;

// FP32 -> BF16, round to nearest even (quiet NaN kept)
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline f32x4 mfma(uint4 a, uint4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// tanh(z) = 1 - 2 / (exp(2z) + 1) with v_exp_f32 and v_rcp_f32; the result
// is rounded to BF16 right after, so ~1e-7 absolute error is immaterial.
__device__ inline float tanh_hw(float z)
{
    float e = __builtin_amdgcn_exp2f(2.8853900817779268f * z);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// LDS written by some lanes of a wave, then read (or overwritten) by others
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

struct MlpDev {
    const uint16_t *w1, *w2, *w3, *w4;   // [H][INP], [H][H], [H][H], [OUTP][H] (zero padded)
    const float *b1, *b2, *b3, *b4;      // b4 padded to OUTP
};

// rows x cols BF16 matrix (cols multiple of 8) from global to LDS with
// stride ld; eight 16-byte loads in flight per thread before their stores
// (one at a time, each would wait out a full memory latency)
__device__ inline void stage(uint16_t* dst, int ld, const uint16_t* __restrict__ src, int rows, int cols)
{
    const int chunks = cols / 8, total = rows * chunks;
    for (int c0 = threadIdx.x; c0 < total; c0 += 4 * FWD_BS) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = min(c0 + u * FWD_BS, total - 1);
            v[u] = *(const uint4*)(src + 8 * (size_t)c);   // row-major: chunk c at 8 c
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = c0 + u * FWD_BS;
            if (c < total) *(uint4*)(dst + (c / chunks) * ld + 8 * (c % chunks)) = v[u];
        }
    }
}

// One layer for the 16 samples of a wave, computed transposed: tile t is
// D = W[t*16 .. +15][K] * in[16][K]^T (weights as the A operand, samples as
// B), so lane (g, i) receives neurons t*16 + 4g .. +3 of sample i: four
// adjacent outputs of one row, written back with one 8-byte store.  All
// activation fragments are read first (the output may then overwrite `in`),
// and each k-step's weight fragments are issued before its MFMAs, so the
// matrix cores do not wait out one LDS latency per instruction.
template <int K, int NT>
__device__ inline void wave_layer(const uint16_t* in, int ldi, const uint16_t* w, int ldw, f32x4* acc)
{
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    uint4 x[K / 32];
#pragma unroll
    for (int kk = 0; kk < K / 32; ++kk) x[kk] = *(const uint4*)(in + li * ldi + 32 * kk + 8 * lg);
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < K / 32; ++kk) {
        uint4 wf[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) wf[t] = *(const uint4*)(w + (t * 16 + li) * ldw + 32 * kk + 8 * lg);
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = mfma(wf[t], x[kk], acc[t]);
    }
}

// four FP32 -> four BF16 (round to nearest even, v_cvt_pk_bf16_f32)
__device__ inline uint2 pack4(float a, float b, float c, float d)
{
    bf16x2 lo = {(__bf16)a, (__bf16)b}, hi = {(__bf16)c, (__bf16)d};
    return make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
}

// bias (LDS) + tanh, BF16 into out[16][LDW] (training: also global [row][H])
template <bool TRAIN>
__device__ inline void wave_tanh_out(const f32x4* acc, const float* bias, uint16_t* out,
                                     uint16_t* __restrict__ gout, int row0, int nrows)
{
    const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int row = row0 + li;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int n0 = t * 16 + 4 * lg;
        const float4 b = *(const float4*)(bias + n0);
        uint2 v = pack4(tanh_hw(acc[t][0] + b.x), tanh_hw(acc[t][1] + b.y), tanh_hw(acc[t][2] + b.z),
                        tanh_hw(acc[t][3] + b.w));
        *(uint2*)(out + li * LDW + n0) = v;
        if (TRAIN && row < nrows) *(uint2*)(gout + (size_t)row * H + n0) = v;
    }
}

// Fused forward.  Inference: y = out[B][ycols] (ycols = 24, ldy = 24).
// Training: also xb [Bp][INP], h1..h3 [Bp][H] (BF16) and y as [Bp][OUTP].
// Waves run their tiles independently (no block barrier after the weights
// are staged).  Weights and biases are read from LDS only.
template <bool TRAIN>
__global__ __launch_bounds__(FWD_BS) void k_mlp_fused(int B, const float* __restrict__ X, MlpDev w,
    float* __restrict__ Y, int ycols, int ldy, uint16_t* __restrict__ xb, uint16_t* __restrict__ h1,
    uint16_t* __restrict__ h2, uint16_t* __restrict__ h3)
{
    extern __shared__ __align__(16) uint16_t lds[];
    uint16_t* s1 = lds;
    uint16_t* s2 = s1 + H * LDI;
    uint16_t* s3 = s2 + H * LDW;
    uint16_t* s4 = s3 + H * LDW;
    uint16_t* sa = s4 + OUTP * LDW;
    stage(s1, LDI, w.w1, H, INP);
    stage(s2, LDW, w.w2, H, H);
    stage(s3, LDW, w.w3, H, H);
    stage(s4, LDW, w.w4, OUTP, H);
    float* sb = (float*)(lds + FUSED_LDS_W / sizeof(uint16_t));
    for (int i = threadIdx.x; i < NBIAS; i += FWD_BS)
        sb[i] = i < H ? w.b1[i] : i < 2 * H ? w.b2[i - H] : i < 3 * H ? w.b3[i - 2 * H] : w.b4[i - 3 * H];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int nwaves = FWD_BS / 64;
    uint16_t* buf = sa + wave * 16 * LDW;
    const int nrows = TRAIN ? ((B + 255) / 256) * 256 : B;   // training rows are padded
    const int nchunks = (nrows + 15) / 16;                // 16 rows per wave task
    f32x4 acc[8];
    // The 16 input rows of a task (lane = column), loaded from clamped
    // addresses so the sixteen loads issue back to back (a load under a
    // branch waits out its own latency); rows past B and lanes past IN
    // read as zero.  The next task's rows are loaded while this task runs.
    float v[16];
    auto load_rows = [&](int r0) {
        const float* xp = X + min(lane, IN - 1);
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = xp[(size_t)min(r0 + i, B - 1) * IN];
    };
    const int stride = gridDim.x * nwaves;
    int task = blockIdx.x * nwaves + wave;
    if (task < nchunks) load_rows(task * 16);
    for (; task < nchunks; task += stride) {
        const int r0 = task * 16;
        {
            const bool col = lane < IN;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const uint16_t q = __builtin_bit_cast(uint16_t, (__bf16)((col && r0 + i < B) ? v[i] : 0.0f));
                buf[i * LDW + lane] = q;
                if (TRAIN && r0 + i < nrows) xb[(size_t)(r0 + i) * INP + lane] = q;
            }
        }
        if (task + stride < nchunks) load_rows((task + stride) * 16);
        wave_lds_sync();
        wave_layer<INP, 8>(buf, LDW, s1, LDI, acc);
        wave_lds_sync();
        wave_tanh_out<TRAIN>(acc, sb, buf, h1, r0, nrows);
        wave_lds_sync();
        wave_layer<H, 8>(buf, LDW, s2, LDW, acc);
        wave_lds_sync();
        wave_tanh_out<TRAIN>(acc, sb + H, buf, h2, r0, nrows);
        wave_lds_sync();
        wave_layer<H, 8>(buf, LDW, s3, LDW, acc);
        wave_lds_sync();
        wave_tanh_out<TRAIN>(acc, sb + 2 * H, buf, h3, r0, nrows);
        wave_lds_sync();
        wave_layer<H, 2>(buf, LDW, s4, LDW, acc);
        wave_lds_sync();
        const int row = r0 + li;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int j0 = t * 16 + 4 * lg;
            if (j0 >= ycols || row >= nrows) continue;   // ycols is a multiple of 4
            const float4 b = *(const float4*)(sb + 3 * H + j0);
            float* yp = Y + (size_t)row * ldy + j0;
            yp[0] = acc[t][0] + b.x; yp[1] = acc[t][1] + b.y;
            yp[2] = acc[t][2] + b.z; yp[3] = acc[t][3] + b.w;
        }
    }
}

// C = A * B^T over the contraction range of split z: A [M][lda], B [N][ldb]
// (BF16, contraction contiguous).  M, N multiples of 16; the contraction
// length per split a multiple of 32.  Block: 4 waves x (16 rows x 64 cols).
// mode 0: FP32 partial out32[z][M][N]; mode 1: BF16 out16[M][N] =
// acc * (1 - aux^2) (aux BF16 [M][N]: the tanh derivative of the layer
// below).
__global__ __launch_bounds__(MLP_BS) void k_gemm_nt(int M, int N, int Ks, const uint16_t* __restrict__ A,
    int lda, const uint16_t* __restrict__ Bm, int ldb, int mode, float* __restrict__ out32,
    uint16_t* __restrict__ out16, const uint16_t* __restrict__ aux)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
    const int m0 = blockIdx.y * 64 + wave * 16, n0 = blockIdx.x * 64;
    if (m0 >= M) return;
    const int kb = blockIdx.z * Ks;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* ap = A + (size_t)(m0 + li) * lda + kb + 8 * lg;
    for (int k0 = 0; k0 < Ks; k0 += 32) {
        uint4 a = *(const uint4*)(ap + k0);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (n0 + t * 16 >= N) continue;
            uint4 b = *(const uint4*)(Bm + (size_t)(n0 + t * 16 + li) * ldb + kb + k0 + 8 * lg);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = n0 + t * 16 + li;
        if (n0 + t * 16 >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = m0 + 4 * lg + r;
            if (mode == 0) {
                out32[((size_t)blockIdx.z * M + m) * N + n] = acc[t][r];
            } else {
                float h = bf2f(aux[(size_t)m * N + n]);
                out16[(size_t)m * N + n] = f2bf(acc[t][r] * (1.0f - h * h));
            }
        }
    }
}

// out[c][r] = in[r][c] (BF16), 64 x 64 tiles through LDS
__global__ __launch_bounds__(MLP_BS) void k_transpose(int R, int C, const uint16_t* __restrict__ in,
                                                      uint16_t* __restrict__ out)
{
    __shared__ uint16_t t[64][65];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        int i = e / 64, j = e % 64;
        if (r0 + i < R && c0 + j < C) t[i][j] = in[(size_t)(r0 + i) * C + c0 + j];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        int j = e / 64, i = e % 64;
        if (r0 + i < R && c0 + j < C) out[(size_t)(c0 + j) * R + r0 + i] = t[i][j];
    }
}

// out[i] = sum_z part[z][i] in order z = 0, 1, ...
__global__ void k_reduce_splits(int n, int S, const float* __restrict__ part, float* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < S; ++z) s += part[(size_t)z * n + i];
    out[i] = s;
}

// db[n] = sum_b gT[n][b] (BF16 values, FP32 sum; fixed-order tree)
__global__ __launch_bounds__(MLP_BS) void k_rowsum(int Bp, const uint16_t* __restrict__ gT,
                                                   float* __restrict__ db)
{
    __shared__ float s[MLP_BS];
    const int n = blockIdx.x;
    float a = 0.f;
    for (int b = threadIdx.x; b < Bp; b += blockDim.x) a += bf2f(gT[(size_t)n * Bp + b]);
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = MLP_BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = s[0];
}

// Minibatch rows from the data set: x[Bp][IN] and t[Bp][OUT] (zero rows past B)
__global__ void k_gather_batch(int B, int Bp, const int* __restrict__ idx, const float* __restrict__ X,
                               const float* __restrict__ T, float* __restrict__ xb, float* __restrict__ tb)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Bp * (IN + OUT)) return;
    int row = e / (IN + OUT), c = e % (IN + OUT);
    float v = 0.f;
    if (row < B) {
        size_t s = (size_t)idx[row];
        v = c < IN ? X[s * IN + c] : T[s * OUT + (c - IN)];
    }
    if (c < IN) xb[(size_t)row * IN + c] = v;
    else tb[(size_t)row * OUT + (c - IN)] = v;
}

// MSE: dY = 2 (y - t) / (B * OUT) as BF16 [Bp][OUTP]; block partial sums of (y - t)^2
__global__ __launch_bounds__(MLP_BS) void k_mse(int B, int Bp, const float* __restrict__ y,
    const float* __restrict__ t, uint16_t* __restrict__ dy, float* __restrict__ part)
{
    __shared__ float s[MLP_BS];
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    float q = 0.f;
    if (e < Bp * OUTP) {
        int row = e / OUTP, o = e % OUTP;
        float g = 0.f;
        if (row < B && o < OUT) {
            float d = y[(size_t)row * OUTP + o] - t[(size_t)row * OUT + o];
            q = d * d;
            g = 2.0f * d / (float)(B * OUT);
        }
        dy[e] = f2bf(g);
    }
    s[threadIdx.x] = q;
    __syncthreads();
    for (int o = MLP_BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_sum_into(int n, const float* __restrict__ part, float scale, float* __restrict__ acc)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += part[i];
    *acc += s * scale;
}

// dy (FP32 [B][OUT], caller's) -> BF16 [Bp][OUTP]
__global__ void k_dy_in(int B, int Bp, const float* __restrict__ dyin, uint16_t* __restrict__ dy)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Bp * OUTP) return;
    int row = e / OUTP, o = e % OUTP;
    dy[e] = f2bf(row < B && o < OUT ? dyin[(size_t)row * OUT + o] : 0.0f);
}

// Padded gradients -> flat parameter order
__global__ void k_pack_grad(const float* __restrict__ gw1, const float* __restrict__ gb1,
    const float* __restrict__ gw2, const float* __restrict__ gb2, const float* __restrict__ gw3,
    const float* __restrict__ gb3, const float* __restrict__ gw4, const float* __restrict__ gb4,
    float* __restrict__ g)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float v;
    if (i < OB1) { int n = i / IN, k = i % IN; v = gw1[n * INP + k]; }
    else if (i < OW2) v = gb1[i - OB1];
    else if (i < OB2) v = gw2[i - OW2];
    else if (i < OW3) v = gb2[i - OB2];
    else if (i < OB3) v = gw3[i - OW3];
    else if (i < OW4) v = gb3[i - OB3];
    else if (i < OB4) v = gw4[i - OW4];
    else v = gb4[i - OB4];
    g[i] = v;
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float c1, float c2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float gi = g[i];
    float mi = 0.9f * m[i] + 0.1f * gi;
    float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

// Flat FP32 parameters -> padded BF16 weights, their transposes, FP32 biases
__global__ void k_pack_params(const float* __restrict__ p, uint16_t* __restrict__ w1,
    uint16_t* __restrict__ w2, uint16_t* __restrict__ w3, uint16_t* __restrict__ w4,
    uint16_t* __restrict__ w1t, uint16_t* __restrict__ w2t, uint16_t* __restrict__ w3t,
    uint16_t* __restrict__ w4t, float* __restrict__ b1, float* __restrict__ b2,
    float* __restrict__ b3, float* __restrict__ b4)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < H * INP) {  // W1 [H][INP], W1^T [INP][H]
        int n = i / INP, k = i % INP;
        uint16_t q = k < IN ? f2bf(p[OW1 + n * IN + k]) : 0;
        w1[i] = q;
        w1t[k * H + n] = q;
    }
    if (i < H * H) {
        int n = i / H, k = i % H;
        uint16_t q2 = f2bf(p[OW2 + i]), q3 = f2bf(p[OW3 + i]);
        w2[i] = q2; w3[i] = q3;
        w2t[k * H + n] = q2; w3t[k * H + n] = q3;
    }
    if (i < OUTP * H) {  // W4 [OUTP][H], W4^T [H][OUTP]
        int o = i / H, k = i % H;
        uint16_t q = o < OUT ? f2bf(p[OW4 + o * H + k]) : 0;
        w4[i] = q;
        w4t[k * OUTP + o] = q;
    }
    if (i < H) { b1[i] = p[OB1 + i]; b2[i] = p[OB2 + i]; b3[i] = p[OB3 + i]; }
    if (i < OUTP) b4[i] = i < OUT ? p[OB4 + i] : 0.0f;
}

int fused_grid()
{
    static int g = 0;
    if (!g) {
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        g = ncu > 0 ? ncu : 256;
    }
    return g;
}

}  // namespace

struct gc_mlp {
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    uint16_t *w1, *w2, *w3, *w4, *w1t, *w2t, *w3t, *w4t;
    float *b1, *b2, *b3, *b4;
    long adam_t = 0;
    std::vector<void*> allocs;
    // training workspace (grown on demand)
    int cap = 0;
    uint16_t *xb = nullptr, *hh[3] = {nullptr, nullptr, nullptr}, *g4 = nullptr, *gg[3] = {nullptr, nullptr, nullptr};
    uint16_t *tA = nullptr, *tB = nullptr;   // transposes
    float *y = nullptr, *xf = nullptr, *tf = nullptr, *part = nullptr, *lpart = nullptr;
    float *gw1, *gw2, *gw3, *gw4, *gb1, *gb2, *gb3, *gb4;
    std::vector<void*> ws;

    template <class T> int alloc(T** ptr, size_t n, std::vector<void*>& list)
    {
        if (hipMalloc((void**)ptr, sizeof(T) * (n ? n : 1)) != hipSuccess) {
            *ptr = nullptr;
            return set_error(GC_ERR_NOMEM, "mlp allocation failed");
        }
        list.push_back((void*)*ptr);
        return GC_OK;
    }
    MlpDev dev() const { return MlpDev{w1, w2, w3, w4, b1, b2, b3, b4}; }
};

namespace {

void free_list(std::vector<void*>& l)
{
    for (void* q : l) (void)hipFree(q);
    l.clear();
}

int pack(gc_mlp* m)
{
    hipStream_t s = stream();
    k_pack_params<<<blocks_for(H * H, 256), 256, 0, s>>>(m->p, m->w1, m->w2, m->w3, m->w4, m->w1t,
        m->w2t, m->w3t, m->w4t, m->b1, m->b2, m->b3, m->b4);
    GC_LAUNCH_CHECK("mlp pack");
    return GC_OK;
}

int ensure_ws(gc_mlp* m, int Bp)
{
    if (Bp <= m->cap) return GC_OK;
    (void)hipStreamSynchronize(stream());
    free_list(m->ws);
    const size_t B = (size_t)Bp;
    int st;
    if ((st = m->alloc(&m->xb, B * INP, m->ws)) || (st = m->alloc(&m->hh[0], B * H, m->ws)) ||
        (st = m->alloc(&m->hh[1], B * H, m->ws)) || (st = m->alloc(&m->hh[2], B * H, m->ws)) ||
        (st = m->alloc(&m->g4, B * OUTP, m->ws)) || (st = m->alloc(&m->gg[0], B * H, m->ws)) ||
        (st = m->alloc(&m->gg[1], B * H, m->ws)) || (st = m->alloc(&m->gg[2], B * H, m->ws)) ||
        (st = m->alloc(&m->tA, B * H, m->ws)) || (st = m->alloc(&m->tB, B * H, m->ws)) ||
        (st = m->alloc(&m->y, B * OUTP, m->ws)) || (st = m->alloc(&m->xf, B * IN, m->ws)) ||
        (st = m->alloc(&m->tf, B * OUT, m->ws)) ||
        (st = m->alloc(&m->part, (B / 256) * H * H, m->ws)) ||
        (st = m->alloc(&m->lpart, (B * OUTP + MLP_BS - 1) / MLP_BS, m->ws)) ||
        (st = m->alloc(&m->gw1, (size_t)H * INP, m->ws)) || (st = m->alloc(&m->gw2, (size_t)H * H, m->ws)) ||
        (st = m->alloc(&m->gw3, (size_t)H * H, m->ws)) || (st = m->alloc(&m->gw4, (size_t)OUTP * H, m->ws)) ||
        (st = m->alloc(&m->gb1, H, m->ws)) || (st = m->alloc(&m->gb2, H, m->ws)) ||
        (st = m->alloc(&m->gb3, H, m->ws)) || (st = m->alloc(&m->gb4, OUTP, m->ws))) {
        m->cap = 0;
        return st;
    }
    m->cap = Bp;
    return GC_OK;
}

int launch_fused(gc_mlp* m, int B, const float* x, float* y, int ycols, int ldy, bool train,
                 hipStream_t s)
{
    int nrows = train ? ((B + 255) / 256) * 256 : B;
    int grid = std::min(fused_grid(), (nrows + 16 * (FWD_BS / 64) - 1) / (16 * (FWD_BS / 64)));
    if (grid < 1) return GC_OK;
    if (train)
        k_mlp_fused<true><<<grid, FWD_BS, FUSED_LDS, s>>>(B, x, m->dev(), y, ycols, ldy, m->xb, m->hh[0],
                                                          m->hh[1], m->hh[2]);
    else
        k_mlp_fused<false><<<grid, FWD_BS, FUSED_LDS, s>>>(B, x, m->dev(), y, ycols, ldy, nullptr, nullptr,
                                                           nullptr, nullptr);
    GC_LAUNCH_CHECK("mlp forward");
    return GC_OK;
}

// dW[N][K] = G^T X over the batch: gT/xT are the transposes [N][Bp], [K][Bp]
int weight_grad(gc_mlp* m, int Bp, int N, int K, const uint16_t* g, const uint16_t* xin, float* dw, float* db)
{
    hipStream_t s = stream();
    k_transpose<<<dim3((N + 63) / 64, Bp / 64), MLP_BS, 0, s>>>(Bp, N, g, m->tA);
    k_transpose<<<dim3((K + 63) / 64, Bp / 64), MLP_BS, 0, s>>>(Bp, K, xin, m->tB);
    const int S = Bp / 256;
    k_gemm_nt<<<dim3((K + 63) / 64, (N + 63) / 64, S), MLP_BS, 0, s>>>(N, K, 256, m->tA, Bp, m->tB,
                                                                          Bp, 0, m->part, nullptr, nullptr);
    k_reduce_splits<<<blocks_for(N * K, 256), 256, 0, s>>>(N * K, S, m->part, dw);
    k_rowsum<<<N, MLP_BS, 0, s>>>(Bp, m->tA, db);
    GC_LAUNCH_CHECK("mlp weight gradient");
    return GC_OK;
}

// g_prev[Bp][K] = (g[Bp][N] * W[N][K]) (1 - h_prev^2), from W^T [K][N]
int data_grad(int Bp, int N, int K, const uint16_t* g, const uint16_t* wt, const uint16_t* hprev,
              uint16_t* gprev)
{
    k_gemm_nt<<<dim3((K + 63) / 64, Bp / 64, 1), MLP_BS, 0, stream()>>>(Bp, K, N, g, N, wt, N, 1,
                                                                         nullptr, gprev, hprev);
    GC_LAUNCH_CHECK("mlp data gradient");
    return GC_OK;
}

// Backward from g4 (BF16 [Bp][OUTP]) to the flat gradient m->g
int backward(gc_mlp* m, int Bp)
{
    int st;
    if ((st = weight_grad(m, Bp, OUTP, H, m->g4, m->hh[2], m->gw4, m->gb4)) ||
        (st = data_grad(Bp, OUTP, H, m->g4, m->w4t, m->hh[2], m->gg[2])) ||
        (st = weight_grad(m, Bp, H, H, m->gg[2], m->hh[1], m->gw3, m->gb3)) ||
        (st = data_grad(Bp, H, H, m->gg[2], m->w3t, m->hh[1], m->gg[1])) ||
        (st = weight_grad(m, Bp, H, H, m->gg[1], m->hh[0], m->gw2, m->gb2)) ||
        (st = data_grad(Bp, H, H, m->gg[1], m->w2t, m->hh[0], m->gg[0])) ||
        (st = weight_grad(m, Bp, H, INP, m->gg[0], m->xb, m->gw1, m->gb1)))
        return st;
    k_pack_grad<<<blocks_for(NPARAM, 256), 256, 0, stream()>>>(m->gw1, m->gb1, m->gw2, m->gb2, m->gw3,
                                                                m->gb3, m->gw4, m->gb4, m->g);
    GC_LAUNCH_CHECK("mlp gradient pack");
    return GC_OK;
}

bool fused_attr_set = false;

}  // namespace

extern "C" int gc_MlpParamCount(void) { return NPARAM; }

extern "C" gc_mlp* gc_MlpCreate(uint64_t seed, int* status)
{
    int dummy;
    if (!status) status = &dummy;
    *status = GC_OK;
    if (!fused_attr_set) {
        if ((*status = hip_check(hipFuncSetAttribute((const void*)k_mlp_fused<true>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)FUSED_LDS), "mlp LDS attribute")) ||
            (*status = hip_check(hipFuncSetAttribute((const void*)k_mlp_fused<false>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)FUSED_LDS), "mlp LDS attribute")))
            return nullptr;
        fused_attr_set = true;
    }
    gc_mlp* m = new gc_mlp();
    int st;
    if ((st = m->alloc(&m->p, NPARAM, m->allocs)) || (st = m->alloc(&m->m, NPARAM, m->allocs)) ||
        (st = m->alloc(&m->v, NPARAM, m->allocs)) || (st = m->alloc(&m->g, NPARAM, m->allocs)) ||
        (st = m->alloc(&m->w1, (size_t)H * INP, m->allocs)) || (st = m->alloc(&m->w2, (size_t)H * H, m->allocs)) ||
        (st = m->alloc(&m->w3, (size_t)H * H, m->allocs)) || (st = m->alloc(&m->w4, (size_t)OUTP * H, m->allocs)) ||
        (st = m->alloc(&m->w1t, (size_t)H * INP, m->allocs)) || (st = m->alloc(&m->w2t, (size_t)H * H, m->allocs)) ||
        (st = m->alloc(&m->w3t, (size_t)H * H, m->allocs)) || (st = m->alloc(&m->w4t, (size_t)OUTP * H, m->allocs)) ||
        (st = m->alloc(&m->b1, H, m->allocs)) || (st = m->alloc(&m->b2, H, m->allocs)) ||
        (st = m->alloc(&m->b3, H, m->allocs)) || (st = m->alloc(&m->b4, OUTP, m->allocs))) {
        *status = st;
        gc_MlpDestroy(m);
        return nullptr;
    }
    // Glorot-uniform weights, zero biases
    std::vector<float> p(NPARAM, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_out, int fan_in) {
        float a = sqrtf(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> U(-a, a);
        for (int i = 0; i < fan_out * fan_in; ++i) p[off + i] = U(rng);
    };
    fill(OW1, H, IN); fill(OW2, H, H); fill(OW3, H, H); fill(OW4, OUT, H);
    if ((st = gc_MlpSetParams(m, p.data(), NPARAM))) {
        *status = st;
        gc_MlpDestroy(m);
        return nullptr;
    }
    return m;
}

extern "C" void gc_MlpDestroy(gc_mlp* m)
{
    if (!m) return;
    (void)hipStreamSynchronize(stream());
    free_list(m->ws);
    free_list(m->allocs);
    delete m;
}

extern "C" int gc_MlpSetParams(gc_mlp* m, const float* params, int n)
{
    if (!m || !params || n != NPARAM) return set_error(GC_ERR_INVALID, "MlpSetParams: bad arguments");
    hipStream_t s = stream();
    std::vector<float> zero(NPARAM, 0.0f);
    int st;
    if ((st = hip_check(hipMemcpyAsync(m->p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, s), "MlpSetParams")) ||
        (st = hip_check(hipMemcpyAsync(m->m, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, s), "MlpSetParams")) ||
        (st = hip_check(hipMemcpyAsync(m->v, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, s), "MlpSetParams")))
        return st;
    m->adam_t = 0;
    if ((st = pack(m))) return st;
    return hip_check(hipStreamSynchronize(s), "MlpSetParams sync");
}

extern "C" int gc_MlpGetParams(gc_mlp* m, float* params, int n)
{
    if (!m || !params || n != NPARAM) return set_error(GC_ERR_INVALID, "MlpGetParams: bad arguments");
    int st = hip_check(hipStreamSynchronize(stream()), "MlpGetParams sync");
    if (st) return st;
    return hip_check(hipMemcpy(params, m->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost), "MlpGetParams");
}

extern "C" int gc_MlpForward(gc_mlp* m, const float* x, float* y, int batch)
{
    if (!m || batch < 0 || (batch > 0 && (!x || !y))) return set_error(GC_ERR_INVALID, "MlpForward: bad arguments");
    if (batch == 0) return GC_OK;
    return launch_fused(m, batch, x, y, OUT, OUT, false, stream());
}

namespace gc {
int mlp_forward(gc_mlp* m, const float* x, float* y, int batch, hipStream_t s)
{
    if (batch == 0) return GC_OK;
    return launch_fused(m, batch, x, y, OUT, OUT, false, s);
}
}  // namespace gc

extern "C" int gc_MlpGradient(gc_mlp* m, const float* x, const float* dy, int batch, float* grad)
{
    if (!m || !x || !dy || !grad || batch <= 0) return set_error(GC_ERR_INVALID, "MlpGradient: bad arguments");
    const int Bp = ((batch + 255) / 256) * 256;
    int st = ensure_ws(m, Bp);
    if (st) return st;
    if ((st = launch_fused(m, batch, x, m->y, OUTP, OUTP, true, stream()))) return st;
    k_dy_in<<<blocks_for(Bp * OUTP, 256), 256, 0, stream()>>>(batch, Bp, dy, m->g4);
    GC_LAUNCH_CHECK("mlp dy");
    if ((st = backward(m, Bp))) return st;
    return hip_check(hipMemcpyAsync(grad, m->g, sizeof(float) * NPARAM, hipMemcpyDeviceToDevice, stream()),
                     "MlpGradient copy");
}

extern "C" int gc_MlpTrain(gc_mlp* m, const float* x, const float* y, int nsamples, int batch, int epochs,
                           float lr, uint64_t seed, float* loss_per_epoch)
{
    if (!m || !x || !y || nsamples <= 0 || batch <= 0 || epochs < 0 || !(lr > 0))
        return set_error(GC_ERR_INVALID, "MlpTrain: bad arguments");
    batch = std::min(batch, nsamples);
    const int Bp = ((batch + 255) / 256) * 256;
    int st = ensure_ws(m, Bp);
    if (st) return st;
    hipStream_t s = stream();
    int* d_idx = nullptr;
    float* d_loss = nullptr;
    if (hipMalloc((void**)&d_idx, sizeof(int) * (size_t)nsamples) != hipSuccess ||
        hipMalloc((void**)&d_loss, sizeof(float) * (size_t)(epochs > 0 ? epochs : 1)) != hipSuccess) {
        if (d_idx) (void)hipFree(d_idx);
        return set_error(GC_ERR_NOMEM, "MlpTrain: allocation failed");
    }
    std::vector<float> zl(epochs > 0 ? epochs : 1, 0.0f);
    st = hip_check(hipMemcpyAsync(d_loss, zl.data(), sizeof(float) * zl.size(), hipMemcpyHostToDevice, s), "MlpTrain");
    std::vector<int> idx(nsamples);
    std::mt19937_64 rng(seed);
    const int nb = (nsamples + batch - 1) / batch;
    const long total = (long)epochs * nb;
    long it = 0;
    const int nlp = (Bp * OUTP + MLP_BS - 1) / MLP_BS;
    for (int ep = 0; ep < epochs && !st; ++ep) {
        std::iota(idx.begin(), idx.end(), 0);
        std::shuffle(idx.begin(), idx.end(), rng);
        if ((st = hip_check(hipMemcpyAsync(d_idx, idx.data(), sizeof(int) * (size_t)nsamples,
                                           hipMemcpyHostToDevice, s), "MlpTrain idx")))
            break;
        for (int b = 0; b < nb && !st; ++b, ++it) {
            const int B = std::min(batch, nsamples - b * batch);
            k_gather_batch<<<blocks_for(Bp * (IN + OUT), 256), 256, 0, s>>>(B, Bp, d_idx + (size_t)b * batch,
                                                                             x, y, m->xf, m->tf);
            if ((st = launch_fused(m, B, m->xf, m->y, OUTP, OUTP, true, s))) break;
            k_mse<<<nlp, MLP_BS, 0, s>>>(B, Bp, m->y, m->tf, m->g4, m->lpart);
            k_sum_into<<<1, 1, 0, s>>>(nlp, m->lpart, 1.0f / ((float)B * OUT * nb), d_loss + ep);
            if ((st = hip_check(hipGetLastError(), "MlpTrain loss"))) break;
            if ((st = backward(m, Bp))) break;
            // cosine decay from lr to lr/20
            float frac = total > 1 ? (float)it / (float)(total - 1) : 0.0f;
            float lrt = lr * (0.05f + 0.95f * 0.5f * (1.0f + cosf(3.14159265f * frac)));
            m->adam_t += 1;
            float c1 = 1.0f / (1.0f - powf(0.9f, (float)m->adam_t));
            float c2 = 1.0f / (1.0f - powf(0.999f, (float)m->adam_t));
            k_adam<<<blocks_for(NPARAM, 256), 256, 0, s>>>(m->p, m->m, m->v, m->g, lrt, c1, c2);
            if ((st = hip_check(hipGetLastError(), "MlpTrain adam"))) break;
            if ((st = pack(m))) break;
        }
    }
    if (!st && loss_per_epoch && epochs > 0)
        st = hip_check(hipMemcpyAsync(loss_per_epoch, d_loss, sizeof(float) * epochs, hipMemcpyDeviceToHost, s),
                       "MlpTrain loss copy");
    int st2 = hip_check(hipStreamSynchronize(s), "MlpTrain sync");
    (void)hipFree(d_idx);
    (void)hipFree(d_loss);
    return st ? st : st2;
}
