// Real GEMM out[m][n] = sum_k A[m][k] B[n][k] on the CDNA4 matrix cores for the gridded
// Hankel contraction (A = J0(rho_m Q_l), B = the spread sources, one row per real/imaginary
// part of each frequency).  Block tile 128 x 128 with eight waves of 32 x 64 (two waves per
// SIMD fit the registers), each wave a 1 x 2 row of 32 x 32 MFMA tiles; K staged through LDS
// 32 at a time, the next stage's global loads in flight during the current stage's MFMAs.
// Every kFold stages the MFMA sums (FP32) are added into FP64 accumulators, so FP32 rounding
// grows with 64 terms, not with K.
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
constexpr int BM = 128, BN = 128, BK = 64;
constexpr int kWaves = 8, kThreads = 64 * kWaves;  // 4 x 2 waves of 32 x 64
constexpr int kLoads = BM * BK / 4 / kThreads;      // float4 loads per thread per operand and stage
constexpr int kVecRow = BK / 4;                     // float4 per tile row
constexpr int kFold = 1;  // K stages summed in FP32 before the FP64 accumulators
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
__global__ void __launch_bounds__(kThreads, 2)
hk_gemm_kernel(int M, int N, int K, const float *A, long long lda, const float *B, long long ldb, double *out,
               long long ldo)
{
    constexpr int NP = KIND == HK_GEMM_BF16X6 ? 3 : 2;
    constexpr bool SPLIT = KIND != HK_GEMM_F32;
    __shared__ typename std::conditional<SPLIT, SplitTiles<NP>, FloatTiles>::type T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;  // rows 32 wm .. +32, columns 64 wn .. +64
    const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
    double accd[2][16];
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
        for (int v = 0; v < 16; v++) accd[tj][v] = 0.0;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const int r32 = lane & 31, kh = lane >> 5;

    float4 ra[kLoads], rb[kLoads];
    auto load_stage = [&](int k0) {
#pragma unroll
        for (int i = 0; i < kLoads; i++) {
            const int lin = i * kThreads + tid, row = lin / kVecRow, c4 = lin % kVecRow;
            const long long gm = m0 + row, gn = n0 + row;
            ra[i] = gm < M ? *reinterpret_cast<const float4 *>(A + gm * lda + k0 + 4 * c4) : zero4;
            rb[i] = gn < N ? *reinterpret_cast<const float4 *>(B + gn * ldb + k0 + 4 * c4) : zero4;
        }
    };
    f32x16 acc[2];
    auto zero_acc = [&]() {
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[tj][v] = 0.0f;
    };
    zero_acc();
    load_stage(0);
    for (int k0 = 0, stage = 1; k0 < K; k0 += BK, stage++) {
        __syncthreads();  // the previous tile has been consumed
#pragma unroll
        for (int i = 0; i < kLoads; i++) {
            const int lin = i * kThreads + tid, row = lin / kVecRow, c = 4 * (lin % kVecRow);
            float xa[4] = {ra[i].x, ra[i].y, ra[i].z, ra[i].w};
            float xb[4] = {rb[i].x, rb[i].y, rb[i].z, rb[i].w};
            if constexpr (SPLIT) {
#pragma unroll
                for (int q = 0; q < NP; q++) {
                    union {
                        __bf16 h[4];
                        uint2 u;
                    } pa, pb;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        pa.h[e] = (__bf16)xa[e];
                        pb.h[e] = (__bf16)xb[e];
                        xa[e] -= (float)pa.h[e];
                        xb[e] -= (float)pb.h[e];
                    }
                    *reinterpret_cast<uint2 *>(&T.a[q][row][c]) = pa.u;
                    *reinterpret_cast<uint2 *>(&T.b[q][row][c]) = pb.u;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    T.a[row][c + e] = xa[e];
                    T.b[row][c + e] = xb[e];
                }
            }
        }
        __syncthreads();
        if (k0 + BK < K) load_stage(k0 + BK);  // in flight during this stage's MFMAs
        if constexpr (SPLIT) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                bf16x8 ap[NP], bp[NP][2];
#pragma unroll
                for (int q = 0; q < NP; q++) {
                    ap[q] = *reinterpret_cast<const bf16x8 *>(&T.a[q][wm * 32 + r32][kk + 8 * kh]);
#pragma unroll
                    for (int t = 0; t < 2; t++)
                        bp[q][t] = *reinterpret_cast<const bf16x8 *>(&T.b[q][wn * 64 + t * 32 + r32][kk + 8 * kh]);
                }
#pragma unroll
                for (int tj = 0; tj < 2; tj++) {
                    f32x16 c = acc[tj];
                    if constexpr (NP == 3) {  // smallest terms first
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[2], bp[0][tj], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1], bp[1][tj], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0], bp[2][tj], c, 0, 0, 0);
                    }
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1], bp[0][tj], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0], bp[1][tj], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0], bp[0][tj], c, 0, 0, 0);
                    acc[tj] = c;
                }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 2) {
                const float a = T.a[wm * 32 + r32][kk + kh];
                float b[2];
#pragma unroll
                for (int t = 0; t < 2; t++) b[t] = T.b[wn * 64 + t * 32 + r32][kk + kh];
#pragma unroll
                for (int tj = 0; tj < 2; tj++) acc[tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b[tj], acc[tj], 0, 0, 0);
            }
        }
        if (stage % kFold == 0 || k0 + BK >= K) {
#pragma unroll
            for (int tj = 0; tj < 2; tj++)
#pragma unroll
                for (int v = 0; v < 16; v++) accd[tj][v] += (double)acc[tj][v];
            zero_acc();
        }
    }
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const long long row = m0 + wm * 32 + 8 * (v >> 2) + 4 * kh + (v & 3);
            const long long col = n0 + wn * 64 + tj * 32 + r32;
            if (row < M && col < N) out[row * ldo + col] = accd[tj][v];
        }
}

template __global__ void hk_gemm_kernel<HK_GEMM_F32>(int, int, int, const float *, long long, const float *, long long,
                                                     double *, long long);
template __global__ void hk_gemm_kernel<HK_GEMM_BF16X3>(int, int, int, const float *, long long, const float *,
                                                        long long, double *, long long);
template __global__ void hk_gemm_kernel<HK_GEMM_BF16X6>(int, int, int, const float *, long long, const float *,
                                                        long long, double *, long long);

// ---- pre-split operands ---------------------------------------------------------------------
// A and B already split into NP bf16 parts by their producers, the parts of a row side by side
// (part q of row r at r * ld + q * K).  K is staged through LDS SBK at a time in two buffers:
// while the MFMAs read one stage, the next is copied from global memory straight into the
// other buffer (global_load_lds_dwordx4, no registers), so a stage costs one round trip of
// latency hidden behind the previous stage's MFMAs.  Tile rows (A rows, then B rows) are
// SBK bf16 long; the 16-byte chunk c of row r sits at slot c ^ ((r / RPB) & (CH - 1)), RPB the
// rows in 256 bytes, so the 32 rows an MFMA operand read touches spread over all banks.
// The barrier that ends a stage first waits for the wave's own LDS reads (lgkmcnt): the
// compiler may leave an operand read in flight across an asm barrier, and the copies issued
// after it overwrite that buffer.
// one 64-lane copy of 16 bytes per lane from global memory to LDS (lane i writes lds + 16 i);
// a device function of its own: the builtin exists only in the device compilation
__device__ inline void copy_to_lds(const unsigned short *g, unsigned short *lds)
{
    __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
}

template <int NP>
struct SplitShape {
    static constexpr int SBK = NP == 2 ? 64 : 32;      // K per stage
    static constexpr int CH = SBK / 8;                  // 16-byte chunks per tile row
    static constexpr int RPB = 256 / (2 * SBK);         // tile rows per 256 bytes
    static constexpr int ROWS = BM + BN;                // A rows, then B rows
    static constexpr int TILE = NP * ROWS * SBK;        // bf16 per stage buffer
    static constexpr int RPI = 64 / CH;                 // tile rows per 64-lane copy
    static constexpr int PER_WAVE = NP * ROWS / RPI / kWaves;  // copies per wave per stage
    static constexpr int FOLD = 64 / SBK;               // stages per FP64 fold (64 K terms)
};

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// F16: the two parts are fp16 (HK_F16X2): h0 h0' on one set of MFMA sums, h0 h1' + h1 h0' on a
// second one folded in with 2^-11, and the result divided by the rows' scales.
template <int NP, bool F16>
__device__ inline void split_gemm(int M, int N, int K, const unsigned short *A, long long lda, const unsigned short *B,
                                  long long ldb, double *out, long long ldo)
{
    using S = SplitShape<NP>;
    using op8 = typename std::conditional<F16, f16x8, bf16x8>::type;
    extern __shared__ __attribute__((aligned(16))) unsigned short lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;  // rows 32 wm .. +32, columns 64 wn .. +64
    const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
    const int r32 = lane & 31, kh = lane >> 5;
    const int rr = lane / S::CH, cs = lane % S::CH;  // this lane's row and slot in a copy

    auto issue = [&](int k0, int buf) {
#pragma unroll
        for (int u = 0; u < S::PER_WAVE; u++) {
            const int ins = wave * S::PER_WAVE + u;                 // uniform
            const int q = ins / (S::ROWS / S::RPI);
            const int row0 = (ins % (S::ROWS / S::RPI)) * S::RPI;  // first tile row of the copy
            const int row = row0 + rr;
            const int c = cs ^ ((row / S::RPB) & (S::CH - 1));      // the chunk that lands in slot cs
            const unsigned short *g;
            if (row0 < BM)  // uniform: a copy is all A rows or all B rows
                g = A + min(m0 + row, (long long)M - 1) * lda + (long long)q * K + k0 + 8 * c;
            else
                g = B + min(n0 + row - BM, (long long)N - 1) * ldb + (long long)q * K + k0 + 8 * c;
            copy_to_lds(g, lds + buf * S::TILE + (q * S::ROWS + row0) * S::SBK);
        }
    };
    auto operand = [&](int buf, int q, int row, int c) {
        const int slot = c ^ ((row / S::RPB) & (S::CH - 1));
        return *reinterpret_cast<const op8 *>(lds + buf * S::TILE + (q * S::ROWS + row) * S::SBK + 8 * slot);
    };

    double accd[2][16];
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
        for (int v = 0; v < 16; v++) accd[tj][v] = 0.0;
    f32x16 acc[2], acc1[2];  // acc1: the fp16 cross terms
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            acc[tj][v] = 0.0f;
            acc1[tj][v] = 0.0f;
        }

    const int stages = K / S::SBK;
    // stage s's copies are in LDS for every wave once this returns
    auto stage_ready = [&](int s) {
        if (s + 1 < stages) {
            issue((s + 1) * S::SBK, (s + 1) & 1);
            // this wave's copies of stage s are done when only stage s + 1's remain
            if constexpr (S::PER_WAVE == 8)
                asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
            else
                asm volatile("s_waitcnt vmcnt(6)" ::: "memory");
        } else {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        // every wave's copies of stage s are in LDS (a barrier with a memory clobber: no LDS
        // access moves across it)
        asm volatile("s_barrier" ::: "memory");
    };
    issue(0, 0);
    if constexpr (!F16) {
        // Two sets of FP32 sums: a fold group (S::FOLD stages, 64 K terms) accumulates in one
        // while the previous group's sums, in the other, are added to the FP64 sums between its
        // MFMAs instead of after them.  The FP64 sums get the same group sums in the same order
        // as when each group is folded at its end.
        f32x16 other[2];
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int v = 0; v < 16; v++) other[tj][v] = 0.0f;
        constexpr int KK = S::SBK / 16;                   // MFMA K steps per stage
        constexpr int PER_STEP = 32 / KK;                 // sums folded per K step (first stage)
        auto group = [&](int g, f32x16 (&use)[2], f32x16 (&old)[2]) {
#pragma unroll
            for (int f = 0; f < S::FOLD; f++) {
                const int s = g * S::FOLD + f, buf = s & 1;
                stage_ready(s);
#pragma unroll
                for (int kk = 0; kk < KK; kk++) {
                    const int c = 2 * kk + kh;
                    op8 ap[NP], bp[NP][2];
#pragma unroll
                    for (int q = 0; q < NP; q++) {
                        ap[q] = operand(buf, q, wm * 32 + r32, c);
#pragma unroll
                        for (int t = 0; t < 2; t++) bp[q][t] = operand(buf, q, BM + wn * 64 + t * 32 + r32, c);
                    }
#pragma unroll
                    for (int tj = 0; tj < 2; tj++) {
                        f32x16 v = use[tj];
                        if constexpr (NP == 3) {  // smallest terms first
                            v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[2], bp[0][tj], v, 0, 0, 0);
                            v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1], bp[1][tj], v, 0, 0, 0);
                            v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0], bp[2][tj], v, 0, 0, 0);
                        }
                        v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[1], bp[0][tj], v, 0, 0, 0);
                        v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0], bp[1][tj], v, 0, 0, 0);
                        v = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[0], bp[0][tj], v, 0, 0, 0);
                        use[tj] = v;
                    }
                    if (f == 0) {
#pragma unroll
                        for (int e = kk * PER_STEP; e < (kk + 1) * PER_STEP; e++) {
                            accd[e / 16][e % 16] += (double)old[e / 16][e % 16];
                            old[e / 16][e % 16] = 0.0f;
                        }
                    }
                }
                asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");  // `buf` free for stage s + 2
            }
        };
        const int groups = stages / S::FOLD;  // K is a multiple of 64
        for (int g = 0; g < groups; g += 2) {
            group(g, acc, other);
            if (g + 1 < groups) group(g + 1, other, acc);
        }
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int v = 0; v < 16; v++) accd[tj][v] += (double)acc[tj][v] + (double)other[tj][v];
    } else {
        // fp16 parts: the cross terms on a second set of sums; with no registers left for two
        // more sets, each group is folded at its end
        for (int s = 0; s < stages; s++) {
            const int buf = s & 1;
            stage_ready(s);
#pragma unroll
            for (int kk = 0; kk < S::SBK; kk += 16) {
                const int c = kk / 8 + kh;
                op8 ap[NP], bp[NP][2];
#pragma unroll
                for (int q = 0; q < NP; q++) {
                    ap[q] = operand(buf, q, wm * 32 + r32, c);
#pragma unroll
                    for (int t = 0; t < 2; t++) bp[q][t] = operand(buf, q, BM + wn * 64 + t * 32 + r32, c);
                }
#pragma unroll
                for (int tj = 0; tj < 2; tj++) {
                    f32x16 x = acc1[tj];
                    x = __builtin_amdgcn_mfma_f32_32x32x16_f16(ap[1], bp[0][tj], x, 0, 0, 0);
                    x = __builtin_amdgcn_mfma_f32_32x32x16_f16(ap[0], bp[1][tj], x, 0, 0, 0);
                    acc1[tj] = x;
                    acc[tj] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ap[0], bp[0][tj], acc[tj], 0, 0, 0);
                }
            }
            if ((s + 1) % S::FOLD == 0 || s + 1 == stages) {
#pragma unroll
                for (int tj = 0; tj < 2; tj++)
#pragma unroll
                    for (int v = 0; v < 16; v++) {
                        accd[tj][v] += (double)acc[tj][v];
                        acc[tj][v] = 0.0f;
                        accd[tj][v] += 0x1p-11 * (double)acc1[tj][v];
                        acc1[tj][v] = 0.0f;
                    }
            }
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");  // `buf` free for stage s + 2
        }
    }
#pragma unroll
    for (int tj = 0; tj < 2; tj++) {
        const long long col = n0 + wn * 64 + tj * 32 + r32;
        double inv_b = 1.0;
        if constexpr (F16)
            if (col < N) inv_b = 1.0 / *reinterpret_cast<const double *>(B + col * ldb + 2LL * K);
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const long long row = m0 + wm * 32 + 8 * (v >> 2) + 4 * kh + (v & 3);
            if (row < M && col < N) {
                double r = accd[tj][v];
                if constexpr (F16) r = r * inv_b / *reinterpret_cast<const double *>(A + row * lda + 2LL * K);
                out[row * ldo + col] = r;
            }
        }
    }
}

template <int NP>
__global__ void __launch_bounds__(kThreads, 1)
hk_gemm_split_kernel(int M, int N, int K, const unsigned short *A, long long lda, const unsigned short *B,
                     long long ldb, double *out, long long ldo)
{
    split_gemm<NP, false>(M, N, K, A, lda, B, ldb, out, ldo);
}

__global__ void __launch_bounds__(kThreads, 1)
hk_gemm_f16x2_kernel(int M, int N, int K, const unsigned short *A, long long lda, const unsigned short *B,
                     long long ldb, double *out, long long ldo)
{
    split_gemm<2, true>(M, N, K, A, lda, B, ldb, out, ldo);
}

template __global__ void hk_gemm_split_kernel<2>(int, int, int, const unsigned short *, long long,
                                                 const unsigned short *, long long, double *, long long);
template __global__ void hk_gemm_split_kernel<3>(int, int, int, const unsigned short *, long long,
                                                 const unsigned short *, long long, double *, long long);

size_t hk_gemm_split_lds(int np) { return np == 2 ? sizeof(unsigned short) * 2 * SplitShape<2>::TILE
                                                  : sizeof(unsigned short) * 2 * SplitShape<3>::TILE; }
