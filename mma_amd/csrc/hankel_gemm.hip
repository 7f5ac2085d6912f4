// Real GEMM out[m][n] = sum_k A[m][k] B[n][k] on the CDNA4 matrix cores for the gridded
// Hankel contraction (A = J0(rho_m Q_l), B = the spread sources, one row per real/imaginary
// part of each frequency).  Block tile 128 x 128 with four waves of 64 x 64, each wave a 2 x 2
// grid of 32 x 32 MFMA tiles; K staged through LDS 32 at a time.  Each stage's MFMA sums
// (FP32) are added into FP64 accumulators, so FP32 rounding grows with 32 terms, not with K.
//   KIND 1: v_mfma_f32_32x32x2_f32 on the FP32 operands.
//   KIND 2: each FP32 operand split while staging into bf16 x = p0 + p1 (+ 2^-16 |x|);
//           products p0 q0 + p0 q1 + p1 q0 on v_mfma_f32_32x32x16_bf16.
//   KIND 3: three-way split x = p0 + p1 + p2 (+ 2^-24 |x|); the six products of order
//           <= 2^-16: p0 q0, p0 q1, p1 q0, p0 q2, p1 q1, p2 q0.
// MFMA operand lanes: A row = lane % 32, B column = lane % 32, K index from lane / 32;
// result element v of a lane: row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32.
#include <type_traits>

#include "hankel_internal.h"

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {
constexpr int BM = 128, BN = 128, BK = 32;
constexpr int kPadF = 1;  // float row padding: conflict-free column reads
constexpr int kPadH = 8;  // bf16 row padding: keeps 16-byte alignment of 8-element reads

template <int NP>
struct SplitTiles {
    __attribute__((aligned(16))) __bf16 a[NP][BM][BK + kPadH];
    __attribute__((aligned(16))) __bf16 b[NP][BN][BK + kPadH];
};
struct FloatTiles {
    float a[BM][BK + kPadF];
    float b[BN][BK + kPadF];
};
}

template <int KIND>
__global__ void __launch_bounds__(256)
hk_gemm_kernel(int M, int N, int K, const float *A, long long lda, const float *B, long long ldb, double *out,
               long long ldo)
{
    constexpr int NP = KIND == HK_GEMM_BF16X6 ? 3 : 2;
    constexpr bool SPLIT = KIND != HK_GEMM_F32;
    __shared__ typename std::conditional<SPLIT, SplitTiles<NP>, FloatTiles>::type T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
    double accd[2][2][16];
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int v = 0; v < 16; v++) accd[ti][tj][v] = 0.0;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const int r32 = lane & 31, kh = lane >> 5;

    for (int k0 = 0; k0 < K; k0 += BK) {
        float4 ra[4], rb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int lin = i * 256 + tid, row = lin >> 3, c4 = lin & 7;
            const long long gm = m0 + row, gn = n0 + row;
            ra[i] = gm < M ? *reinterpret_cast<const float4 *>(A + gm * lda + k0 + 4 * c4) : zero4;
            rb[i] = gn < N ? *reinterpret_cast<const float4 *>(B + gn * ldb + k0 + 4 * c4) : zero4;
        }
        __syncthreads();  // the previous tile has been consumed
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int lin = i * 256 + tid, row = lin >> 3, c = 4 * (lin & 7);
            const float av[4] = {ra[i].x, ra[i].y, ra[i].z, ra[i].w};
            const float bv[4] = {rb[i].x, rb[i].y, rb[i].z, rb[i].w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if constexpr (SPLIT) {
                    float xa = av[e], xb = bv[e];
#pragma unroll
                    for (int q = 0; q < NP; q++) {
                        const __bf16 pa = (__bf16)xa, pb = (__bf16)xb;
                        T.a[q][row][c + e] = pa;
                        T.b[q][row][c + e] = pb;
                        xa -= (float)pa;
                        xb -= (float)pb;
                    }
                } else {
                    T.a[row][c + e] = av[e];
                    T.b[row][c + e] = bv[e];
                }
            }
        }
        __syncthreads();
        f32x16 acc[2][2];
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
            for (int tj = 0; tj < 2; tj++)
#pragma unroll
                for (int v = 0; v < 16; v++) acc[ti][tj][v] = 0.0f;
        if constexpr (SPLIT) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                bf16x8 ap[NP][2], bp[NP][2];
#pragma unroll
                for (int q = 0; q < NP; q++)
#pragma unroll
                    for (int t = 0; t < 2; t++) {
                        ap[q][t] = *reinterpret_cast<const bf16x8 *>(&T.a[q][wm * 64 + t * 32 + r32][kk + 8 * kh]);
                        bp[q][t] = *reinterpret_cast<const bf16x8 *>(&T.b[q][wn * 64 + t * 32 + r32][kk + 8 * kh]);
                    }
#pragma unroll
                for (int ti = 0; ti < 2; ti++)
#pragma unroll
                    for (int tj = 0; tj < 2; tj++) {
                        f32x16 c = acc[ti][tj];
                        if constexpr (NP == 3) {  // smallest terms first
                            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[2][ti], bp[0][tj], c, 0, 0, 0);
                            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1][ti], bp[1][tj], c, 0, 0, 0);
                            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0][ti], bp[2][tj], c, 0, 0, 0);
                        }
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1][ti], bp[0][tj], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0][ti], bp[1][tj], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0][ti], bp[0][tj], c, 0, 0, 0);
                        acc[ti][tj] = c;
                    }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 2) {
                float a[2], b[2];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    a[t] = T.a[wm * 64 + t * 32 + r32][kk + kh];
                    b[t] = T.b[wn * 64 + t * 32 + r32][kk + kh];
                }
#pragma unroll
                for (int ti = 0; ti < 2; ti++)
#pragma unroll
                    for (int tj = 0; tj < 2; tj++)
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
            for (int tj = 0; tj < 2; tj++)
#pragma unroll
                for (int v = 0; v < 16; v++) accd[ti][tj][v] += (double)acc[ti][tj][v];
    }
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const long long row = m0 + wm * 64 + ti * 32 + 8 * (v >> 2) + 4 * kh + (v & 3);
                const long long col = n0 + wn * 64 + tj * 32 + r32;
                if (row < M && col < N) out[row * ldo + col] = accd[ti][tj][v];
            }
}

template __global__ void hk_gemm_kernel<HK_GEMM_F32>(int, int, int, const float *, long long, const float *, long long,
                                                     double *, long long);
template __global__ void hk_gemm_kernel<HK_GEMM_BF16X3>(int, int, int, const float *, long long, const float *,
                                                        long long, double *, long long);
template __global__ void hk_gemm_kernel<HK_GEMM_BF16X6>(int, int, int, const float *, long long, const float *,
                                                        long long, double *, long long);
