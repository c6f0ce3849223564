// MLP surrogate for log P(k): training (forward, backward, Adam) and inference on MI355X.
//
// Network: widths w_0 -> w_1 -> ... -> w_L, GELU after every hidden layer, linear output.
// Weights and activations are BF16 in the GEMMs, accumulation is FP32 on the matrix cores
// (v_mfma_f32_16x16x16_bf16), and the optimizer keeps FP32 master weights (Adam).  The loss is the
// mean squared error of the (standardized) outputs.
//
// The three products of a layer are one GEMM description with strides:
//   forward          Z  = X  W  + b     A = X [M][K],     B = W [K][N]
//   backward data    dX = dZ Wᵀ         A = dZ [M][N],    B(p, j) = W[j][p]
//   backward weight  dW = Xᵀ dZ         A(i, p) = X[p][i], B = dZ [M][N]   (and db = column sums of dZ)
// with fused epilogues (bias, GELU and its derivative, BF16/FP32 stores, atomic accumulation).
// Kernels by shape: sbz_mlp_gemm_t (training sizes: 16-byte loads along each operand's contiguous
// index, register prefetch, split contraction for the weight gradients), sbz_mlp_gemm_kk (large
// inference batches, 128 x 128 tiles), sbz_mlp_gemm_smallk / _smallm (the 7-wide input layer), and
// sbz_mlp_gemm (any other shape, element-wise staging).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "sbz_internal.h"

typedef unsigned short bf16_t;
typedef short sbz_short4 __attribute__((ext_vector_type(4)));
typedef float sbz_float4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN inputs are not produced by this code path)
__host__ __device__ inline bf16_t sbz_f2bf(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float sbz_bf2f(bf16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

// GELU(x) = x Φ(x) with the normal CDF Φ, and its derivative Φ(x) + x φ(x)
__device__ inline float sbz_gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float sbz_gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum { SBZ_EPI_LINEAR = 0, SBZ_EPI_GELU = 1, SBZ_EPI_GELU_BWD = 2 };

struct SbzGemm {
    const bf16_t* A;
    long long a_i, a_p;   // A(i, p) = A[i a_i + p a_p]
    const bf16_t* B;
    long long b_p, b_j;   // B(p, j) = B[p b_p + j b_j]
    int M, N, K;
    const float* bias;    // [N] or null
    int epi;              // SBZ_EPI_*
    const float* zin;     // GELU_BWD: pre-activation [M][N] of the layer below
    float* out_f;         // [M][N] FP32 result (before the activation) or null
    bf16_t* out_h;        // [M][N] BF16 result (after the activation) or null
    int accum;            // 1: add the product atomically into out_f (zeroed by the host), linear only
    float* bsum;          // [N] or null: column sums of B over the contraction, added atomically
};

#define TM 64
#define TN 64
#define TK 32

extern "C" __global__ void __launch_bounds__(256) sbz_mlp_gemm(SbzGemm g) {
    __shared__ bf16_t sA[TM][TK + 4];
    __shared__ bf16_t sB[TN][TK + 4];   // stored transposed: sB[j][p]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    sbz_float4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = sbz_float4{0.f, 0.f, 0.f, 0.f};
    for (int p0 = 0; p0 < g.K; p0 += TK) {
        // stage A[i0 .. +64][p0 .. +32] and B[p0 .. +32][j0 .. +64] (zero outside the matrices)
        for (int e = tid; e < TM * TK; e += 256) {
            const int r = e / TK, c = e % TK, i = i0 + r, p = p0 + c;
            sA[r][c] = (i < g.M && p < g.K) ? g.A[(long long)i * g.a_i + (long long)p * g.a_p] : (bf16_t)0;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            const int r = e / TK, c = e % TK, j = j0 + r, p = p0 + c;
            sB[r][c] = (j < g.N && p < g.K) ? g.B[(long long)p * g.b_p + (long long)j * g.b_j] : (bf16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            // operand layout of the 16x16x16 product: lane l supplies row (l % 16) of A and
            // column (l % 16) of B at k = 4 (l / 16) .. + 3
            const int kq = kk + 4 * (lane >> 4);
            sbz_short4 af[2], bfr[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const bf16_t* pa = &sA[wm + 16 * a + (lane & 15)][kq];
                af[a] = sbz_short4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
                const bf16_t* pb = &sB[wn + 16 * a + (lane & 15)][kq];
                bfr[a] = sbz_short4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    // result layout: lane l holds rows 4 (l / 16) .. + 3 of column (l % 16)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + wm + 16 * a + 4 * (lane >> 4) + q;
                const int j = j0 + wn + 16 * b + (lane & 15);
                if (i >= g.M || j >= g.N) continue;
                float v = acc[a][b][q];
                if (g.bias) v += g.bias[j];
                const long long o = (long long)i * g.N + j;
                if (g.out_f) g.out_f[o] = v;
                if (g.out_h) {
                    float h = v;
                    if (g.epi == SBZ_EPI_GELU) h = sbz_gelu(v);
                    else if (g.epi == SBZ_EPI_GELU_BWD) h = v * sbz_gelu_d(g.zin[o]);
                    g.out_h[o] = sbz_f2bf(h);
                }
            }
}

// Large products with both operands contiguous along the contraction index p (A(i, p) = A[i a_i + p],
// B(p, j) = B[j b_j + p]): the forward pass on the transposed weight copy and the backward-data
// product on the weights themselves.  A 256-thread workgroup computes a 128 x 128 tile, each
// wavefront a 64 x 64 quarter as 4 x 4 blocks of v_mfma_f32_16x16x32_bf16; K advances 32 at a time
// through LDS, filled with 16-byte global loads and 16-byte LDS stores (rows padded by 8 elements).
// Requires M, N multiples of 128, K a multiple of 32, a_i, b_j multiples of 8 and 16-byte aligned
// operands (the host checks and otherwise uses sbz_mlp_gemm).  Measured alternatives that were
// slower on MI355X at 65536 x 512 x 512: K steps of 64 (36 KB of LDS, fewer resident workgroups)
// and prefetching the next K tile into registers.
typedef __bf16 sbz_bf16x8 __attribute__((ext_vector_type(8)));
#define FBM 128
#define FBN 128
#define FBK 32
#define FPAD 8

extern "C" __global__ void __launch_bounds__(256) sbz_mlp_gemm_kk(SbzGemm g) {
    __shared__ __attribute__((aligned(16))) bf16_t sA[FBM][FBK + FPAD];
    __shared__ __attribute__((aligned(16))) bf16_t sB[FBN][FBK + FPAD];   // sB[j][p]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * FBM, j0 = blockIdx.x * FBN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    sbz_float4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = sbz_float4{0.f, 0.f, 0.f, 0.f};
    // staging: 128 rows x 32 columns = 512 chunks of 8 elements per operand, 2 per thread
    const int r0 = tid >> 2, c0 = (tid & 3) * 8;
    for (int p0 = 0; p0 < g.K; p0 += FBK) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = r0 + 64 * h;
            const uint4 va = *(const uint4*)(g.A + (long long)(i0 + r) * g.a_i + p0 + c0);
            const uint4 vb = *(const uint4*)(g.B + (long long)(j0 + r) * g.b_j + p0 + c0);
            *(uint4*)&sA[r][c0] = va;
            *(uint4*)&sB[r][c0] = vb;
        }
        __syncthreads();
        // operand layout of the 16x16x32 product: lane l supplies row (l % 16) of A and column
        // (l % 16) of B at k = 8 (l / 16) .. + 7
        const int kq = 8 * (lane >> 4);
        sbz_bf16x8 af[4], bfr[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            af[a] = *(const sbz_bf16x8*)&sA[wm + 16 * a + (lane & 15)][kq];
            bfr[a] = *(const sbz_bf16x8*)&sB[wn + 16 * a + (lane & 15)][kq];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // result layout: lane l holds rows 4 (l / 16) .. + 3 of column (l % 16)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + wm + 16 * a + 4 * (lane >> 4) + q;
                const int j = j0 + wn + 16 * b + (lane & 15);
                float v = acc[a][b][q];
                if (g.bias) v += g.bias[j];
                const long long o = (long long)i * g.N + j;
                if (g.out_f) g.out_f[o] = v;
                if (g.out_h) {
                    float hv = v;
                    if (g.epi == SBZ_EPI_GELU) hv = sbz_gelu(v);
                    else if (g.epi == SBZ_EPI_GELU_BWD) hv = v * sbz_gelu_d(g.zin[o]);
                    g.out_h[o] = sbz_f2bf(hv);
                }
            }
}

// Training-size products (batches of ~10³ rows, widths of ~10²): the three products of a layer
// with whole 16-byte loads whatever the operand layout.  A 256-thread workgroup computes a 64 x BN
// tile (BN = 64 or 32) with v_mfma_f32_16x16x32_bf16; K advances BK = 128 at a time through
// LDS tiles that are contiguous along k (the MFMA operand order), and the next step's operands are
// loaded into registers before the current step's products, so their latency overlaps the matrix
// work.  An operand contiguous along k is copied chunk for chunk; one contiguous along its row (i)
// or column (j) index is loaded 8 rows at a time and scattered into the k-contiguous tile.  The
// weight gradients accumulate atomically into a zeroed out_f (accum), with the contraction cut into
// grid.z ranges when the tiles alone do not fill the chip, and the first row of workgroups adds the
// column sums of B (the bias gradient, bsum) from the same tiles.  Requires the contiguous extent of
// each operand and the other stride to be multiples of 8 and 16-byte aligned operands (the host
// checks, otherwise a general kernel runs).
#define GBM 64
#define GBK 128
#define GPAD 8
template <int BN, int BK>
__global__ void __launch_bounds__(256) sbz_mlp_gemm_t(SbzGemm g, int kps) {
    constexpr int WM = (BN == 64) ? 2 : 4;                  // wavefronts along M
    constexpr int WN = 4 / WM;
    constexpr int FM = GBM / WM / 16, FN = BN / WN / 16;    // 16 x 16 blocks per wavefront
    constexpr int ACH = GBM * BK / 8 / 256;                 // chunks of 8 per thread: A
    constexpr int BCH = BN * BK / 8 / 256;                  //                         B
    constexpr int KC = BK / 8;                              // chunks per k-contiguous row
    __shared__ __attribute__((aligned(16))) bf16_t sA[GBM][BK + GPAD];
    __shared__ __attribute__((aligned(16))) bf16_t sB[BN][BK + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * GBM, j0 = blockIdx.x * BN;
    const int wm = (wave / WN) * (GBM / WM), wn = (wave % WN) * (BN / WN);
    const int pb = blockIdx.z * kps, pe = min(g.K, pb + kps);
    const bool a_k = (g.a_p == 1), b_k = (g.b_p == 1);
    const bool do_bsum = g.bsum && blockIdx.y == 0;
    sbz_float4 acc[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = sbz_float4{0.f, 0.f, 0.f, 0.f};
    float bs = 0.f;   // this thread's part of the column sum of B (column tid % BN)
    uint4 ra[ACH], rb[BCH];
    // chunk c of the A tile: k-contiguous: row c/KC, k 8(c%KC); i-contiguous: k c/8, rows 8(c%8)..+7
    auto load = [&](int p0) {
#pragma unroll
        for (int h = 0; h < ACH; ++h) {
            const int c = tid + 256 * h;
            int i, p;
            if (a_k) { i = i0 + c / KC; p = p0 + 8 * (c % KC); }
            else     { p = p0 + c / 8; i = i0 + 8 * (c % 8); }
            const bool ok = (i < g.M) && (p < pe);
            ra[h] = ok ? *(const uint4*)(g.A + (long long)i * g.a_i + (long long)p * g.a_p) : uint4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int h = 0; h < BCH; ++h) {
            const int c = tid + 256 * h;
            int j, p;
            if (b_k) { j = j0 + c / KC; p = p0 + 8 * (c % KC); }
            else     { p = p0 + c / (BN / 8); j = j0 + 8 * (c % (BN / 8)); }
            const bool ok = (j < g.N) && (p < pe);
            rb[h] = ok ? *(const uint4*)(g.B + (long long)p * g.b_p + (long long)j * g.b_j) : uint4{0u, 0u, 0u, 0u};
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int h = 0; h < ACH; ++h) {
            const int c = tid + 256 * h;
            if (a_k) {
                *(uint4*)&sA[c / KC][8 * (c % KC)] = ra[h];
            } else {
                const bf16_t* e = (const bf16_t*)&ra[h];
#pragma unroll
                for (int q = 0; q < 8; ++q) sA[8 * (c % 8) + q][c / 8] = e[q];
            }
        }
#pragma unroll
        for (int h = 0; h < BCH; ++h) {
            const int c = tid + 256 * h;
            if (b_k) {
                *(uint4*)&sB[c / KC][8 * (c % KC)] = rb[h];
            } else {
                const bf16_t* e = (const bf16_t*)&rb[h];
#pragma unroll
                for (int q = 0; q < 8; ++q) sB[8 * (c % (BN / 8)) + q][c / (BN / 8)] = e[q];
            }
        }
    };
    if (pb < pe) load(pb);
    for (int p0 = pb; p0 < pe; p0 += BK) {
        store();
        __syncthreads();
        if (p0 + BK < pe) load(p0 + BK);   // next step's operands in flight during the products
        if (do_bsum) {   // column tid % BN, k slice tid / BN of the tile (zero past the matrices)
            constexpr int SL = 256 / BN;
#pragma unroll
            for (int q = 0; q < BK / SL; ++q) bs += sbz_bf2f(sB[tid % BN][(tid / BN) * (BK / SL) + q]);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            // lane l supplies row (l % 16) of A and column (l % 16) of B at k = kk + 8 (l / 16) .. + 7
            const int kq = kk + 8 * (lane >> 4);
            sbz_bf16x8 af[FM], bfr[FN];
#pragma unroll
            for (int a = 0; a < FM; ++a) af[a] = *(const sbz_bf16x8*)&sA[wm + 16 * a + (lane & 15)][kq];
#pragma unroll
            for (int b = 0; b < FN; ++b) bfr[b] = *(const sbz_bf16x8*)&sB[wn + 16 * b + (lane & 15)][kq];
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
                for (int b = 0; b < FN; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    if (do_bsum && j0 + tid % BN < g.N) unsafeAtomicAdd(&g.bsum[j0 + tid % BN], bs);
    // result layout: lane l holds rows 4 (l / 16) .. + 3 of column (l % 16)
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + wm + 16 * a + 4 * (lane >> 4) + q;
                const int j = j0 + wn + 16 * b + (lane & 15);
                if (i >= g.M || j >= g.N) continue;
                float v = acc[a][b][q];
                const long long o = (long long)i * g.N + j;
                if (g.accum) {
                    unsafeAtomicAdd(&g.out_f[o], v);
                    continue;
                }
                if (g.bias) v += g.bias[j];
                if (g.out_f) g.out_f[o] = v;
                if (g.out_h) {
                    float hv = v;
                    if (g.epi == SBZ_EPI_GELU) hv = sbz_gelu(v);
                    else if (g.epi == SBZ_EPI_GELU_BWD) hv = v * sbz_gelu_d(g.zin[o]);
                    g.out_h[o] = sbz_f2bf(hv);
                }
            }
}

// Products with a contraction of at most 16 (the input layer's forward pass, K = 7): one thread per
// output, the K products in FP32, the general epilogue.
extern "C" __global__ void __launch_bounds__(256) sbz_mlp_gemm_smallk(SbzGemm g) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)g.M * g.N) return;
    const int i = (int)(e / g.N), j = (int)(e % g.N);
    float v = 0.f;
    for (int p = 0; p < g.K; ++p)
        v += sbz_bf2f(g.A[(long long)i * g.a_i + (long long)p * g.a_p]) * sbz_bf2f(g.B[(long long)p * g.b_p + (long long)j * g.b_j]);
    if (g.bias) v += g.bias[j];
    if (g.out_f) g.out_f[e] = v;
    if (g.out_h) {
        float hv = v;
        if (g.epi == SBZ_EPI_GELU) hv = sbz_gelu(v);
        else if (g.epi == SBZ_EPI_GELU_BWD) hv = v * sbz_gelu_d(g.zin[e]);
        g.out_h[e] = sbz_f2bf(hv);
    }
}

// Accumulated products with at most 16 output rows and a long contraction (the input layer's weight
// gradient, 7 x N over the batch): a block takes 64 columns and one of gridDim.y slices of the
// contraction, its 4 wavefronts interleave over the slice, each lane keeps the M sums of its column
// and adds them (and the column sum of B for bsum) atomically into the zeroed outputs.
extern "C" __global__ void __launch_bounds__(256) sbz_mlp_gemm_smallm(SbzGemm g) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63), part = threadIdx.x >> 6;
    if (j >= g.N) return;
    const int per = (g.K + gridDim.y - 1) / gridDim.y;
    const int pb = blockIdx.y * per, pe = min(g.K, pb + per);
    float acc[16], bs = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    for (int p = pb + part; p < pe; p += 4) {
        const float b = sbz_bf2f(g.B[(long long)p * g.b_p + (long long)j * g.b_j]);
        bs += b;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < g.M) acc[i] += sbz_bf2f(g.A[(long long)i * g.a_i + (long long)p * g.a_p]) * b;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i)
        if (i < g.M) unsafeAtomicAdd(&g.out_f[(long long)i * g.N + j], acc[i]);
    if (g.bsum) unsafeAtomicAdd(&g.bsum[j], bs);
}

// BF16 transposed copy of one weight matrix: dst [N][K] = src [K][N]
extern "C" __global__ void sbz_mlp_transpose(const bf16_t* __restrict__ src, int K, int N, bf16_t* __restrict__ dst) {
    __shared__ bf16_t t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int p = by + r, j = bx + tx;
        t[r][tx] = (p < K && j < N) ? src[(long long)p * N + j] : (bf16_t)0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int j = bx + r, p = by + tx;
        if (j < N && p < K) dst[(long long)j * K + p] = t[tx][r];
    }
}

// batch gather: rows idx[r] of the training inputs (FP32 -> BF16) and targets (FP32)
extern "C" __global__ void sbz_mlp_gather(const float* __restrict__ X, const float* __restrict__ T,
                                          const int* __restrict__ idx, int B, int nin, int nout,
                                          bf16_t* __restrict__ xb, float* __restrict__ tb) {
    const int r = blockIdx.x, t = threadIdx.x;
    if (r >= B) return;
    const long long s = idx[r];
    for (int c = t; c < nin; c += blockDim.x) xb[(long long)r * nin + c] = sbz_f2bf(X[s * nin + c]);
    for (int c = t; c < nout; c += blockDim.x) tb[(long long)r * nout + c] = T[s * nout + c];
}

// inputs FP32 -> BF16
extern "C" __global__ void sbz_mlp_to_bf16(const float* __restrict__ x, long long n, bf16_t* __restrict__ y) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        y[e] = sbz_f2bf(x[e]);
}

// loss and its gradient: d = 2 (y - t) / (B nout) (BF16 for the GEMMs), loss accumulated per block
extern "C" __global__ void sbz_mlp_mse(const float* __restrict__ y, const float* __restrict__ t, long long n,
                                       float scale, bf16_t* __restrict__ dy_h, float* __restrict__ dy_f,
                                       float* __restrict__ loss) {
    __shared__ float part[256];
    float s = 0.f;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const float r = y[e] - t[e];
        s += r * r;
        const float d = 2.0f * r * scale;
        if (dy_f) dy_f[e] = d;
        dy_h[e] = sbz_f2bf(d);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] * scale);
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters, then their BF16 working copy
extern "C" __global__ void sbz_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                        const float* __restrict__ g, long long n, float lr, float b1, float b2,
                                        float eps, float c1, float c2, bf16_t* __restrict__ wh) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const float ge = g[e];
        const float me = b1 * m[e] + (1.f - b1) * ge;
        const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
        m[e] = me;
        v[e] = ve;
        const float we = w[e] - lr * (me / c1) / (sqrtf(ve / c2) + eps);
        w[e] = we;
        if (wh) wh[e] = sbz_f2bf(we);
    }
}

// ----------------------------------------------------------------------------------- host side

#define MCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            sbz_set_error("%s failed: %s", #x, hipGetErrorString(e_));                           \
            return SBZ_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

// device buffers and events released at scope exit (error paths included)
struct MlpScoped {
    std::vector<void*> bufs;
    std::vector<hipEvent_t> evs;
    ~MlpScoped() {
        for (void* p : bufs)
            if (p) (void)hipFree(p);
        for (hipEvent_t e : evs)
            if (e) (void)hipEventDestroy(e);
    }
};

struct sbz_mlp {
    int L = 0;                 // number of layers (weight matrices)
    std::vector<int> w;        // widths w[0..L]
    int max_batch = 0;
    long long n_par = 0;       // all weights then all biases, FP32
    std::vector<long long> off_w, off_b;
    float *par = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;   // FP32 [n_par]
    bf16_t* par_h = nullptr;   // BF16 copy of the weights (biases unused)
    bf16_t* par_t = nullptr;   // BF16 transposed weights ([N][K] per layer) for sbz_mlp_gemm_kk
    bool par_t_ok = false;     // par_t matches par_h
    std::vector<float*> z;     // per layer: pre-activation [max_batch][w[l+1]] FP32
    std::vector<bf16_t*> x;    // per layer input activation [max_batch][w[l]] BF16 (x[L] = output, unused)
    bf16_t *dh = nullptr;      // gradient scratch BF16 [max_batch][max width] x 2
    float* y = nullptr;        // output FP32 [max_batch][w[L]]
    float* tb = nullptr;       // batch targets FP32
    int* idx = nullptr;        // batch indices
    float* loss = nullptr;     // [1]
    float *data_x = nullptr, *data_t = nullptr;
    long long n_data = 0;
    long long step = 0;
    hipStream_t stream = nullptr;
    int wmax = 0;
};

static void mlp_free(sbz_mlp* h) {
    for (void* p : {(void*)h->par, (void*)h->m, (void*)h->v, (void*)h->grad, (void*)h->par_h, (void*)h->par_t, (void*)h->dh,
                    (void*)h->y, (void*)h->tb, (void*)h->idx, (void*)h->loss, (void*)h->data_x,
                    (void*)h->data_t})
        if (p) (void)hipFree(p);
    for (auto p : h->z)
        if (p) (void)hipFree(p);
    for (auto p : h->x)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

// the 128 x 128 kernel applies to contraction-contiguous operands of tile-multiple shape, and pays
// only when the grid fills the chip (at least 256 tiles)
static bool kk_fits(const SbzGemm& g) {
    auto al16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
    return g.a_p == 1 && g.b_p == 1 && g.M % FBM == 0 && g.N % FBN == 0 && g.K % FBK == 0 && g.a_i % 8 == 0 &&
           g.b_j % 8 == 0 && al16(g.A) && al16(g.B) && (long long)(g.M / FBM) * (g.N / FBN) >= 256;
}

// the training-size kernel: whole 16-byte chunks along each operand's contiguous index
static bool t_fits(const SbzGemm& g) {
    auto al16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
    const bool a_ok = (g.a_p == 1) ? (g.K % 8 == 0 && g.a_i % 8 == 0) : (g.a_i == 1 && g.M % 8 == 0 && g.a_p % 8 == 0);
    const bool b_ok = (g.b_p == 1) ? (g.K % 8 == 0 && g.b_j % 8 == 0) : (g.b_j == 1 && g.N % 8 == 0 && g.b_p % 8 == 0);
    return a_ok && b_ok && al16(g.A) && al16(g.B) && g.M > 0 && g.N > 0 && g.K > 0;
}

// accumulating products (g.accum) need out_f (and bsum) zeroed by the caller; the other kernels
// write their outputs
static hipError_t gemm(sbz_mlp* h, const SbzGemm& g) {
    if (!g.accum && !g.bsum && kk_fits(g)) {
        hipLaunchKernelGGL(sbz_mlp_gemm_kk, dim3(g.N / FBN, g.M / FBM), dim3(256), 0, h->stream, g);
        return hipGetLastError();
    }
    if (t_fits(g)) {
        // 64 x 64 tiles when they fill the chip, else 64 x 32; an accumulating product with a long
        // contraction (the weight gradients) is split along it until ~256 workgroups run
        const int gm = (g.M + GBM - 1) / GBM;
        const bool wide = (long long)gm * ((g.N + 63) / 64) >= 256;
        const int bn = wide ? 64 : 32;
        const long long tiles = (long long)gm * ((g.N + bn - 1) / bn);
        // an accumulating product with a long contraction is split until ~256 workgroups run
        // (staging a whole 512-deep range at once instead of 128-deep prefetched steps measured no
        // faster: these products are bound by the launch and the per-tile latency, not by the loads)
        int split = 1;
        if (g.accum && tiles < 256) split = (int)std::min<long long>(std::max<long long>(1, 256 / tiles), std::max(1, g.K / 256));
        const int kps = ((g.K + split - 1) / split + GBK - 1) / GBK * GBK;
        split = (g.K + kps - 1) / kps;
        const dim3 grid((g.N + bn - 1) / bn, gm, split);
        if (wide) hipLaunchKernelGGL((sbz_mlp_gemm_t<64, GBK>), grid, dim3(256), 0, h->stream, g, kps);
        else hipLaunchKernelGGL((sbz_mlp_gemm_t<32, GBK>), grid, dim3(256), 0, h->stream, g, kps);
        return hipGetLastError();
    }
    if (g.K <= 16 && !g.accum && !g.bsum) {
        const long long n = (long long)g.M * g.N;
        hipLaunchKernelGGL(sbz_mlp_gemm_smallk, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, g);
        return hipGetLastError();
    }
    if (g.M <= 16 && g.accum && g.epi == SBZ_EPI_LINEAR && !g.bias && !g.out_h) {
        const int slices = std::max(1, std::min(32, g.K / 64));
        hipLaunchKernelGGL(sbz_mlp_gemm_smallm, dim3((g.N + 63) / 64, slices), dim3(256), 0, h->stream, g);
        return hipGetLastError();
    }
    if (g.accum || g.bsum) return hipErrorInvalidValue;   // no other kernel accumulates
    dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM);
    hipLaunchKernelGGL(sbz_mlp_gemm, grid, dim3(256), 0, h->stream, g);
    return hipGetLastError();
}

// refresh the transposed weight copy after the weights changed
static hipError_t refresh_transposed(sbz_mlp* h) {
    if (h->par_t_ok) return hipSuccess;
    for (int l = 0; l < h->L; ++l) {
        const int K = h->w[l], N = h->w[l + 1];
        hipLaunchKernelGGL(sbz_mlp_transpose, dim3((N + 31) / 32, (K + 31) / 32), dim3(256), 0, h->stream,
                           h->par_h + h->off_w[l], K, N, h->par_t + h->off_w[l]);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    h->par_t_ok = true;
    return hipSuccess;
}

extern "C" int sbz_mlp_create(int n_layers, const int* widths, int max_batch, unsigned long long seed,
                              sbz_mlp** out) {
    sbz_clear_error();
    if (n_layers < 1 || !widths || max_batch < 1 || !out) {
        sbz_set_error("sbz_mlp_create: invalid argument");
        return SBZ_ERR_ARG;
    }
    for (int l = 0; l <= n_layers; ++l)
        if (widths[l] < 1) {
            sbz_set_error("sbz_mlp_create: width %d is not positive", l);
            return SBZ_ERR_ARG;
        }
    sbz_mlp* h = new sbz_mlp;
    h->L = n_layers;
    h->w.assign(widths, widths + n_layers + 1);
    h->max_batch = max_batch;
    long long nw = 0, nb = 0;
    for (int l = 0; l < n_layers; ++l) {
        h->off_w.push_back(nw);
        nw += (long long)widths[l] * widths[l + 1];
        h->wmax = std::max(h->wmax, std::max(widths[l], widths[l + 1]));
    }
    for (int l = 0; l < n_layers; ++l) {
        h->off_b.push_back(nw + nb);
        nb += widths[l + 1];
    }
    h->n_par = nw + nb;
    // He-uniform initialisation of the weights, zero biases
    std::vector<float> p0(h->n_par, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < n_layers; ++l) {
        const float lim = std::sqrt(6.0f / (float)widths[l]);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (long long e = 0; e < (long long)widths[l] * widths[l + 1]; ++e) p0[h->off_w[l] + e] = U(rng);
    }
    std::vector<bf16_t> ph(h->n_par);
    for (long long e = 0; e < h->n_par; ++e) ph[e] = sbz_f2bf(p0[e]);
    auto fail = [&](hipError_t e, const char* what) {
        sbz_set_error("sbz_mlp_create: %s: %s", what, hipGetErrorString(e));
        mlp_free(h);
        delete h;
        return SBZ_ERR_HIP;
    };
    hipError_t e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) return fail(e, "stream");
    const size_t P4 = sizeof(float) * h->n_par;
    for (float** p : {&h->par, &h->m, &h->v, &h->grad})
        if ((e = hipMalloc((void**)p, P4)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->par_h, sizeof(bf16_t) * h->n_par)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc((void**)&h->par_t, sizeof(bf16_t) * h->n_par)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMemcpy(h->par, p0.data(), P4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "upload");
    if ((e = hipMemcpy(h->par_h, ph.data(), sizeof(bf16_t) * h->n_par, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "upload");
    if ((e = hipMemset(h->m, 0, P4)) != hipSuccess || (e = hipMemset(h->v, 0, P4)) != hipSuccess)
        return fail(e, "memset");
    h->z.assign(n_layers, nullptr);
    h->x.assign(n_layers, nullptr);
    const size_t B = (size_t)max_batch;
    for (int l = 0; l < n_layers; ++l) {
        if ((e = hipMalloc((void**)&h->z[l], sizeof(float) * B * widths[l + 1])) != hipSuccess) return fail(e, "z");
        if ((e = hipMalloc((void**)&h->x[l], sizeof(bf16_t) * B * widths[l])) != hipSuccess) return fail(e, "x");
    }
    if ((e = hipMalloc((void**)&h->dh, sizeof(bf16_t) * B * h->wmax * 2)) != hipSuccess) return fail(e, "dh");
    if ((e = hipMalloc((void**)&h->y, sizeof(float) * B * widths[n_layers])) != hipSuccess) return fail(e, "y");
    if ((e = hipMalloc((void**)&h->tb, sizeof(float) * B * widths[n_layers])) != hipSuccess) return fail(e, "tb");
    if ((e = hipMalloc((void**)&h->idx, sizeof(int) * B)) != hipSuccess) return fail(e, "idx");
    if ((e = hipMalloc((void**)&h->loss, sizeof(float))) != hipSuccess) return fail(e, "loss");
    *out = h;
    return SBZ_OK;
}

extern "C" int sbz_mlp_destroy(sbz_mlp* h) {
    if (!h) return SBZ_OK;
    mlp_free(h);
    delete h;
    return SBZ_OK;
}

extern "C" int sbz_mlp_n_params(sbz_mlp* h, long long* n) {
    sbz_clear_error();
    if (!h || !n) {
        sbz_set_error("sbz_mlp_n_params: invalid argument");
        return SBZ_ERR_ARG;
    }
    *n = h->n_par;
    return SBZ_OK;
}

// parameters (weights [K][N] per layer, then biases) as FP32, in or out
extern "C" int sbz_mlp_get_params(sbz_mlp* h, float* p) {
    sbz_clear_error();
    if (!h || !p) {
        sbz_set_error("sbz_mlp_get_params: invalid argument");
        return SBZ_ERR_ARG;
    }
    MCHK(hipMemcpy(p, h->par, sizeof(float) * h->n_par, hipMemcpyDeviceToHost));
    return SBZ_OK;
}
extern "C" int sbz_mlp_set_params(sbz_mlp* h, const float* p) {
    sbz_clear_error();
    if (!h || !p) {
        sbz_set_error("sbz_mlp_set_params: invalid argument");
        return SBZ_ERR_ARG;
    }
    std::vector<bf16_t> ph(h->n_par);
    for (long long e = 0; e < h->n_par; ++e) ph[e] = sbz_f2bf(p[e]);
    MCHK(hipMemcpy(h->par, p, sizeof(float) * h->n_par, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(h->par_h, ph.data(), sizeof(bf16_t) * h->n_par, hipMemcpyHostToDevice));
    h->par_t_ok = false;
    return SBZ_OK;
}

// forward over B rows already in h->x[0] (BF16); keep_z keeps every pre-activation for the
// backward pass (training) — inference skips those stores
static hipError_t forward(sbz_mlp* h, int B, float* y, bool keep_z) {
    for (int l = 0; l < h->L; ++l) {
        const bool last = (l == h->L - 1);
        SbzGemm g{};
        g.A = h->x[l];
        g.a_i = h->w[l];
        g.a_p = 1;
        g.B = h->par_h + h->off_w[l];
        g.b_p = h->w[l + 1];
        g.b_j = 1;
        {
            // the same product on the transposed weights W^T [N][K], when the large kernel applies
            SbzGemm t = g;
            t.B = h->par_t + h->off_w[l];
            t.b_p = 1;
            t.b_j = h->w[l];
            t.M = B;
            t.N = h->w[l + 1];
            t.K = h->w[l];
            if (kk_fits(t)) {
                hipError_t e = refresh_transposed(h);
                if (e != hipSuccess) return e;
                g = t;
            }
        }
        g.M = B;
        g.N = h->w[l + 1];
        g.K = h->w[l];
        g.bias = h->par + h->off_b[l];
        g.epi = last ? SBZ_EPI_LINEAR : SBZ_EPI_GELU;
        g.out_f = last ? y : (keep_z ? h->z[l] : nullptr);
        g.out_h = last ? nullptr : h->x[l + 1];
        hipError_t e = gemm(h, g);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// loss (added into the device float *d_loss, zeroed by the caller) and gradient (into h->grad) of
// the batch in h->x[0] / h->tb; nothing waits for the device
static int loss_and_grad(sbz_mlp* h, int B, float* d_loss) {
    const int L = h->L, nout = h->w[L];
    MCHK(forward(h, B, h->y, true));
    // every weight and bias gradient accumulates into the zeroed gradient vector
    MCHK(hipMemsetAsync(h->grad, 0, sizeof(float) * h->n_par, h->stream));
    bf16_t* d_cur = h->dh;                           // BF16 gradient w.r.t. the current layer's output
    bf16_t* d_nxt = h->dh + (size_t)h->max_batch * h->wmax;
    const long long n = (long long)B * nout;
    hipLaunchKernelGGL(sbz_mlp_mse, dim3(64), dim3(256), 0, h->stream, h->y, h->tb, n, 1.0f / (float)n, d_cur,
                       (float*)nullptr, d_loss);
    MCHK(hipGetLastError());
    for (int l = L - 1; l >= 0; --l) {
        const int K = h->w[l], N = h->w[l + 1];
        // dW_l = X_lᵀ dZ_l and db_l = column sums of dZ_l, from the same tiles of dZ_l
        SbzGemm gw{};
        gw.A = h->x[l];
        gw.a_i = 1;
        gw.a_p = K;
        gw.B = d_cur;
        gw.b_p = N;
        gw.b_j = 1;
        gw.M = K;
        gw.N = N;
        gw.K = B;
        gw.epi = SBZ_EPI_LINEAR;
        gw.out_f = h->grad + h->off_w[l];
        gw.accum = 1;
        gw.bsum = h->grad + h->off_b[l];
        MCHK(gemm(h, gw));
        if (l == 0) break;
        // dZ_{l-1} = (dZ_l W_lᵀ) ⊙ GELU'(Z_{l-1})
        SbzGemm gd{};
        gd.A = d_cur;
        gd.a_i = N;
        gd.a_p = 1;
        gd.B = h->par_h + h->off_w[l];
        gd.b_p = 1;
        gd.b_j = N;
        gd.M = B;
        gd.N = K;
        gd.K = N;
        gd.epi = SBZ_EPI_GELU_BWD;
        gd.zin = h->z[l - 1];
        gd.out_h = d_nxt;
        MCHK(gemm(h, gd));
        std::swap(d_cur, d_nxt);
    }
    return SBZ_OK;
}

// Gradient of the mean squared error on one batch (x [B][w0], t [B][wL] host FP32) without
// updating the parameters: for tests against a CPU restatement.
extern "C" int sbz_mlp_grad(sbz_mlp* h, const float* x, const float* t, int B, float* grad, float* loss) {
    sbz_clear_error();
    if (!h || !x || !t || !grad || B < 1 || B > h->max_batch) {
        sbz_set_error("sbz_mlp_grad: invalid argument");
        return SBZ_ERR_ARG;
    }
    const int nin = h->w[0], nout = h->w[h->L];
    std::vector<bf16_t> xb((size_t)B * nin);
    for (size_t e = 0; e < xb.size(); ++e) xb[e] = sbz_f2bf(x[e]);
    MCHK(hipMemcpy(h->x[0], xb.data(), sizeof(bf16_t) * xb.size(), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(h->tb, t, sizeof(float) * (size_t)B * nout, hipMemcpyHostToDevice));
    MCHK(hipMemsetAsync(h->loss, 0, sizeof(float), h->stream));
    int st = loss_and_grad(h, B, h->loss);
    if (st != SBZ_OK) return st;
    MCHK(hipMemcpyAsync(grad, h->grad, sizeof(float) * h->n_par, hipMemcpyDeviceToHost, h->stream));
    if (loss) MCHK(hipMemcpyAsync(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MCHK(hipStreamSynchronize(h->stream));
    return SBZ_OK;
}

// Training data resident on the device (host FP32: x [n][w0], t [n][wL]).
extern "C" int sbz_mlp_set_data(sbz_mlp* h, const float* x, const float* t, long long n) {
    sbz_clear_error();
    if (!h || !x || !t || n < 1) {
        sbz_set_error("sbz_mlp_set_data: invalid argument");
        return SBZ_ERR_ARG;
    }
    if (h->data_x) (void)hipFree(h->data_x);
    if (h->data_t) (void)hipFree(h->data_t);
    h->data_x = h->data_t = nullptr;
    MCHK(hipMalloc((void**)&h->data_x, sizeof(float) * n * h->w[0]));
    MCHK(hipMalloc((void**)&h->data_t, sizeof(float) * n * h->w[h->L]));
    MCHK(hipMemcpy(h->data_x, x, sizeof(float) * n * h->w[0], hipMemcpyHostToDevice));
    MCHK(hipMemcpy(h->data_t, t, sizeof(float) * n * h->w[h->L], hipMemcpyHostToDevice));
    h->n_data = n;
    return SBZ_OK;
}

// n_batches Adam steps over consecutive batches of the index list idx[n_batches * batch] (rows of
// the resident data); loss[n_batches] receives each batch's loss (may be null).
extern "C" int sbz_mlp_train(sbz_mlp* h, const int* idx, int n_batches, int batch, float lr, float* loss,
                             float* kernel_ms) {
    sbz_clear_error();
    if (!h || !idx || n_batches < 1 || batch < 1 || batch > h->max_batch || !h->data_x || !(lr > 0)) {
        sbz_set_error("sbz_mlp_train: invalid argument (or no data set)");
        return SBZ_ERR_ARG;
    }
    for (long long e = 0; e < (long long)n_batches * batch; ++e)
        if (idx[e] < 0 || idx[e] >= h->n_data) {
            sbz_set_error("sbz_mlp_train: index %lld out of range", e);
            return SBZ_ERR_ARG;
        }
    MlpScoped sc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MCHK(hipEventCreate(&e0));
    sc.evs.push_back(e0);
    MCHK(hipEventCreate(&e1));
    sc.evs.push_back(e1);
    // every step's batch indices and loss slot on the device: the host enqueues all steps without
    // waiting for any
    int* d_idx = nullptr;
    float* d_loss = nullptr;
    MCHK(hipMalloc((void**)&d_idx, sizeof(int) * (size_t)n_batches * batch));
    sc.bufs.push_back(d_idx);
    MCHK(hipMalloc((void**)&d_loss, sizeof(float) * (size_t)n_batches));
    sc.bufs.push_back(d_loss);
    MCHK(hipMemcpyAsync(d_idx, idx, sizeof(int) * (size_t)n_batches * batch, hipMemcpyHostToDevice, h->stream));
    MCHK(hipMemsetAsync(d_loss, 0, sizeof(float) * (size_t)n_batches, h->stream));
    MCHK(hipEventRecord(e0, h->stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s < n_batches; ++s) {
        hipLaunchKernelGGL(sbz_mlp_gather, dim3(batch), dim3(256), 0, h->stream, h->data_x, h->data_t,
                           d_idx + (size_t)s * batch, batch, h->w[0], h->w[h->L], h->x[0], h->tb);
        MCHK(hipGetLastError());
        int st = loss_and_grad(h, batch, d_loss + s);
        if (st != SBZ_OK) return st;
        ++h->step;
        const float c1 = 1.f - std::pow(b1, (float)h->step), c2 = 1.f - std::pow(b2, (float)h->step);
        hipLaunchKernelGGL(sbz_mlp_adam, dim3(1024), dim3(256), 0, h->stream, h->par, h->m, h->v, h->grad, h->n_par, lr,
                           b1, b2, eps, c1, c2, h->par_h);
        MCHK(hipGetLastError());
        h->par_t_ok = false;
    }
    MCHK(hipEventRecord(e1, h->stream));
    if (loss) MCHK(hipMemcpyAsync(loss, d_loss, sizeof(float) * (size_t)n_batches, hipMemcpyDeviceToHost, h->stream));
    MCHK(hipEventSynchronize(e1));
    MCHK(hipStreamSynchronize(h->stream));
    float ms = 0.f;
    MCHK(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms = ms;
    return SBZ_OK;
}

// Inference: y [n][wL] = net(x [n][w0]) (host FP32), in batches of max_batch.
extern "C" int sbz_mlp_forward(sbz_mlp* h, const float* x, long long n, float* y, float* kernel_ms) {
    sbz_clear_error();
    if (!h || !x || !y || n < 1) {
        sbz_set_error("sbz_mlp_forward: invalid argument");
        return SBZ_ERR_ARG;
    }
    const int nin = h->w[0], nout = h->w[h->L];
    MlpScoped sc;
    float *dx = nullptr, *dy = nullptr;
    MCHK(hipMalloc((void**)&dx, sizeof(float) * n * nin));
    sc.bufs.push_back(dx);
    MCHK(hipMalloc((void**)&dy, sizeof(float) * n * nout));
    sc.bufs.push_back(dy);
    MCHK(hipMemcpy(dx, x, sizeof(float) * n * nin, hipMemcpyHostToDevice));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MCHK(hipEventCreate(&e0));
    sc.evs.push_back(e0);
    MCHK(hipEventCreate(&e1));
    sc.evs.push_back(e1);
    MCHK(hipEventRecord(e0, h->stream));
    for (long long r0 = 0; r0 < n; r0 += h->max_batch) {
        const int B = (int)std::min<long long>(h->max_batch, n - r0);
        hipLaunchKernelGGL(sbz_mlp_to_bf16, dim3(256), dim3(256), 0, h->stream, dx + r0 * nin, (long long)B * nin,
                           h->x[0]);
        MCHK(hipGetLastError());
        MCHK(forward(h, B, dy + r0 * nout, false));
    }
    MCHK(hipEventRecord(e1, h->stream));
    MCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    MCHK(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms = ms;
    MCHK(hipMemcpy(y, dy, sizeof(float) * n * nout, hipMemcpyDeviceToHost));
    return SBZ_OK;
}
