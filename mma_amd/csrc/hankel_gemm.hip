// Real GEMM out[m][n] = sum_k A[m][k] B[n][k] on the CDNA4 matrix cores for the gridded
// Hankel contraction (A = J0(rho_m Q_l), B = the spread sources, one row per real/imaginary
// part of each frequency).  Block tile 128 x 128 with four waves of 64 x 64, each wave a 2 x 2
// grid of 32 x 32 MFMA tiles; K staged through LDS 32 at a time.
//   SPLIT = false: v_mfma_f32_32x32x2_f32 on the FP32 operands.
//   SPLIT = true:  each FP32 operand is split while staging into bf16 hi = rn(x) and
//                  lo = rn(x - hi); three v_mfma_f32_32x32x16_bf16 products hi*hi + hi*lo +
//                  lo*hi accumulate in FP32 (the lo*lo term, ~2^-16 relative, is dropped).
// MFMA operand lanes: A row = lane % 32, B column = lane % 32, K index from lane / 32;
// result element v of a lane: row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32.
#include "hankel_internal.h"

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {
constexpr int BM = 128, BN = 128, BK = 32;
constexpr int kPadF = 1;  // float row padding: conflict-free column reads
constexpr int kPadH = 8;  // bf16 row padding: keeps 16-byte alignment of 8-element reads
}

template <bool SPLIT>
__global__ void __launch_bounds__(256)
hk_gemm_kernel(int M, int N, int K, const float *A, long long lda, const float *B, long long ldb, double *out,
               long long ldo)
{
    __shared__ float As[SPLIT ? 1 : BM][BK + kPadF], Bs[SPLIT ? 1 : BN][BK + kPadF];
    __shared__ __attribute__((aligned(16))) __bf16 Ah[SPLIT ? BM : 1][BK + kPadH], Al[SPLIT ? BM : 1][BK + kPadH];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[SPLIT ? BN : 1][BK + kPadH], Bl[SPLIT ? BN : 1][BK + kPadH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
    f32x16 acc[2][2];
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[ti][tj][v] = 0.0f;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);

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
                    const __bf16 ah = (__bf16)av[e], bh = (__bf16)bv[e];
                    Ah[row][c + e] = ah;
                    Al[row][c + e] = (__bf16)(av[e] - (float)ah);
                    Bh[row][c + e] = bh;
                    Bl[row][c + e] = (__bf16)(bv[e] - (float)bh);
                } else {
                    As[row][c + e] = av[e];
                    Bs[row][c + e] = bv[e];
                }
            }
        }
        __syncthreads();
        const int r32 = lane & 31, kh = lane >> 5;
        if constexpr (SPLIT) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int ra_ = wm * 64 + t * 32 + r32, rb_ = wn * 64 + t * 32 + r32;
                    ah[t] = *reinterpret_cast<const bf16x8 *>(&Ah[ra_][kk + 8 * kh]);
                    al[t] = *reinterpret_cast<const bf16x8 *>(&Al[ra_][kk + 8 * kh]);
                    bh[t] = *reinterpret_cast<const bf16x8 *>(&Bh[rb_][kk + 8 * kh]);
                    bl[t] = *reinterpret_cast<const bf16x8 *>(&Bl[rb_][kk + 8 * kh]);
                }
#pragma unroll
                for (int ti = 0; ti < 2; ti++)
#pragma unroll
                    for (int tj = 0; tj < 2; tj++) {
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[ti], bh[tj], acc[ti][tj], 0, 0, 0);
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ti], bl[tj], acc[ti][tj], 0, 0, 0);
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ti], bh[tj], acc[ti][tj], 0, 0, 0);
                    }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 2) {
                float a[2], b[2];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    a[t] = As[wm * 64 + t * 32 + r32][kk + kh];
                    b[t] = Bs[wn * 64 + t * 32 + r32][kk + kh];
                }
#pragma unroll
                for (int ti = 0; ti < 2; ti++)
#pragma unroll
                    for (int tj = 0; tj < 2; tj++)
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
            }
        }
    }
    const int r32 = lane & 31, kh = lane >> 5;
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const long long row = m0 + wm * 64 + ti * 32 + 8 * (v >> 2) + 4 * kh + (v & 3);
                const long long col = n0 + wn * 64 + tj * 32 + r32;
                if (row < M && col < N) out[row * ldo + col] = (double)acc[ti][tj][v];
            }
}

template __global__ void hk_gemm_kernel<false>(int, int, int, const float *, long long, const float *, long long,
                                               double *, long long);
template __global__ void hk_gemm_kernel<true>(int, int, int, const float *, long long, const float *, long long,
                                              double *, long long);
