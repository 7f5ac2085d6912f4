// The contraction of the Hankel far-field integral over r and the source planes
// (Hankel_transform.py:492-514 per plane, :641-661 over planes):
//   * hk_direct_kernel: every (plane, omega, rho, r) element with J0 in FP64;
//   * hk_spread_kernel + hk_bessel_kernel: the gridded form, C_l(omega) spread from the
//     sources by Lagrange interpolation in q = K r / D and the matrix J0(rho_m Q_l) whose
//     product (hankel_gemm.hip) gives the same weighted sum over planes.
#include <type_traits>

#include "hankel_internal.h"
#include "hankel_j0.h"

namespace {

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 load_src(const double2 *s, long long i) { return s[i]; }
__device__ inline double2 load_src(const float2 *s, long long i)
{
    const float2 v = s[i];
    return make_double2((double)v.x, (double)v.y);
}

// u_j = tw_j r_j pre_k(r_j, omega) e^{-i K r_j^2 / 2D} F_k(omega, r_j): the integrand of
// Hankel_transform.py:499-502 without the Bessel factor, with the trapezoid weight folded in
template <typename S>
__device__ inline double2 source_term(const hk_planes &p, const S *plane, int k, int o, int j, double K, double D,
                                      double2 pk)
{
    const double r = p.rgrid[j];
    double2 f = load_src(plane, (long long)o * p.nr + j);
    if (p.pre_mode == 2) pk = reinterpret_cast<const double2 *>(p.pre)[((long long)k * p.no + o) * p.nr + j];
    double2 v = cmul(pk, f);
    if (p.near_field) {
        double s, c;
        sincos(-K * (r * r) / (2.0 * D), &s, &c);
        v = cmul(make_double2(c, s), v);
    }
    const double wr = p.tw[j] * r;
    return make_double2(v.x * wr, v.y * wr);
}

constexpr int kThreads = 256;
constexpr int kSpreadThreads = 1024;  // the spread kernels' workgroup
constexpr int kJTile = 512;     // radial points staged in LDS per pass
constexpr int kAnchor = 32;     // exact sincos every kAnchor rotated steps

} // namespace

// grid (no, ceil(nrho / 256)): one workgroup per frequency and 256 screen radii, all planes
template <typename S>
__global__ void __launch_bounds__(kThreads)
hk_direct_kernel(hk_planes p, int n_out, const double *weights, int per_plane, double2 *out, double inv_c,
                 int uniform_r, double r0, double dr)
{
    __shared__ double2 u[kJTile];
    __shared__ double rr[kJTile];
    const int o = blockIdx.x;
    const int m = blockIdx.y * kThreads + threadIdx.x;
    const bool live = m < p.nrho;
    const double rho = live ? p.rho[m] : 0.0;
    const double K = p.ogrid[o] * inv_c;
    double2 acc_out[4] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0),
                          make_double2(0.0, 0.0)};
    (void)r0;
    for (int k = 0; k < p.nz; k++) {
        const double D = p.distance - p.zgrid[k];
        const double a = (K / D) * rho;  // x_j = a r_j
        const S *plane = reinterpret_cast<const S *>(p.src) + (long long)k * p.plane_stride;
        double2 pk = make_double2(1.0, 0.0);
        if (p.pre_mode == 1) pk = reinterpret_cast<const double2 *>(p.pre)[(long long)k * p.no + o];
        double2 acc = make_double2(0.0, 0.0);
        for (int j0 = 0; j0 < p.nr; j0 += kJTile) {
            const int nj = min(kJTile, p.nr - j0);
            __syncthreads();
            for (int jj = threadIdx.x; jj < nj; jj += kThreads) {
                u[jj] = source_term(p, plane, k, o, j0 + jj, K, D, pk);
                rr[jj] = p.rgrid[j0 + jj];
            }
            __syncthreads();
            if (!live) continue;
            if (uniform_r) {
                // cos, sin of x_j by rotation through the angle a dr, re-anchored every kAnchor steps
                double ct, st;
                sincos(a * dr, &st, &ct);
                for (int jb = 0; jb < nj; jb += kAnchor) {
                    double sx, cx;
                    sincos(a * rr[jb], &sx, &cx);
                    const int je = min(nj, jb + kAnchor);
                    for (int jj = jb; jj < je; jj++) {
                        const double x = a * rr[jj];
                        const double J = hk::j0(x, cx, sx);
                        const double2 v = u[jj];
                        acc.x = fma(J, v.x, acc.x);
                        acc.y = fma(J, v.y, acc.y);
                        const double cn = cx * ct - sx * st;
                        sx = fma(sx, ct, cx * st);
                        cx = cn;
                    }
                }
            } else {
                for (int jj = 0; jj < nj; jj++) {
                    const double J = hk::j0(a * rr[jj]);
                    const double2 v = u[jj];
                    acc.x = fma(J, v.x, acc.x);
                    acc.y = fma(J, v.y, acc.y);
                }
            }
        }
        if (!live) continue;
        if (per_plane) {
            out[((long long)k * p.nrho + m) * p.no + o] = acc;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (q < n_out) {
                    const double wq = weights[(long long)q * p.nz + k];
                    acc_out[q].x = fma(wq, acc.x, acc_out[q].x);
                    acc_out[q].y = fma(wq, acc.y, acc_out[q].y);
                }
            }
        }
    }
    if (!live || per_plane) return;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (q < n_out) out[((long long)q * p.nrho + m) * p.no + o] = acc_out[q];
}

template __global__ void hk_direct_kernel<double2>(hk_planes, int, const double *, int, double2 *, double, int,
                                                   double, double);
template __global__ void hk_direct_kernel<float2>(hk_planes, int, const double *, int, double2 *, double, int,
                                                  double, double);

namespace {

// sin and cos of x for |x| < kSincosLimit: Cody-Waite reduction x = n pi/2 + y with pi/2 split into
// three doubles (each product n * c_i exact inside an FMA), then the Taylor series of sin y
// (through y^17) and cos y (through y^18) on |y| <= pi/4, whose first omitted terms are below
// 1e-19, and the quadrant n mod 4.  The constants are pi/2 = c1 + c2 + c3 and 1/k! rounded to
// double.  Fewer registers and instructions than the general routine, which also serves
// arguments up to 1e308.
__device__ inline void sincos_reduced(double x, double *s, double *c)
{
    constexpr double c1 = 0x1.921fb54442d18p+0, c2 = 0x1.1a62633145c07p-54, c3 = -0x1.f1976b7ed8fbcp-110;
    constexpr double two_over_pi = 0x1.45f306dc9c883p-1;
    const double n = rint(x * two_over_pi);
    double y = fma(-n, c1, x);
    y = fma(-n, c2, y);
    y = fma(-n, c3, y);
    const double z = y * y;
    // 1/k! with alternating signs
    double ps = 2.8114572543455206e-15;
    ps = fma(ps, z, -7.647163731819816e-13);
    ps = fma(ps, z, 1.6059043836821613e-10);
    ps = fma(ps, z, -2.505210838544172e-08);
    ps = fma(ps, z, 2.7557319223985893e-06);
    ps = fma(ps, z, -0.0001984126984126984);
    ps = fma(ps, z, 0.008333333333333333);
    ps = fma(ps, z, -0.16666666666666666);
    const double sy = fma(y * z, ps, y);
    double pc = -1.5619206968586225e-16;
    pc = fma(pc, z, 4.779477332387385e-14);
    pc = fma(pc, z, -1.1470745597729725e-11);
    pc = fma(pc, z, 2.08767569878681e-09);
    pc = fma(pc, z, -2.755731922398589e-07);
    pc = fma(pc, z, 2.48015873015873e-05);
    pc = fma(pc, z, -0.001388888888888889);
    pc = fma(pc, z, 0.041666666666666664);
    pc = fma(pc, z, -0.5);
    const double cy = fma(z, pc, 1.0);
    const int q = (int)n & 3;
    const double sa = (q & 1) ? cy : sy, ca = (q & 1) ? sy : cy;
    *s = (q & 2) ? -sa : sa;
    *c = ((q + 1) & 2) ? -ca : ca;
}

// Lagrange basis of HK_GRID_P nodes at offsets d_i = i - (P/2 - 1) from floor(t), at f = frac(t)
struct Lagrange {
    double L[HK_GRID_P];
    __device__ explicit Lagrange(double f)
    {
        constexpr int P = HK_GRID_P;
        double t[P];
#pragma unroll
        for (int i = 0; i < P; i++) t[i] = f - (double)(i - (P / 2 - 1));
        double pre[P], suf[P];
        pre[0] = 1.0;
#pragma unroll
        for (int i = 1; i < P; i++) pre[i] = pre[i - 1] * t[i - 1];
        suf[P - 1] = 1.0;
#pragma unroll
        for (int i = P - 2; i >= 0; i--) suf[i] = suf[i + 1] * t[i + 1];
        // 1 / prod_{j != i} (d_i - d_j) = (-1)^(P-1-i) / (i! (P-1-i)!), as a product with
        // the correctly rounded reciprocal (a division per weight costs more than the rest)
#pragma unroll
        for (int i = 0; i < P; i++) {
            double den = 1.0;
            for (int q = 2; q <= i; q++) den *= (double)q;
            for (int q = 2; q <= P - 1 - i; q++) den *= (double)q;
            const double inv = ((P - 1 - i) & 1 ? -1.0 : 1.0) / den;  // folded at compile time
            L[i] = (pre[i] * suf[i]) * inv;
        }
    }
};

// element l of an operand row starting at `row` (ld already applied): double, float, the value
// split into 2 or 3 bf16 parts v ~ p0 + p1 (+ p2) (part q at row + q nq + l), or, for HK_F16X2,
// v * scale split into two fp16 parts h0 + 2^-11 h1 (the row's scale, a power of two that puts
// its largest element in [2^13, 2^14), sits as a double at row + 2 nq)
template <int TYPE>
__device__ inline void store_operand(void *out, long long row, int l, int nq, double v, double scale = 1.0)
{
    if constexpr (TYPE == HK_F64) {
        reinterpret_cast<double *>(out)[row + l] = v;
    } else if constexpr (TYPE == HK_F32) {
        reinterpret_cast<float *>(out)[row + l] = (float)v;
    } else if constexpr (TYPE == HK_F16X2) {
        unsigned short *o = reinterpret_cast<unsigned short *>(out) + row + l;
        const double x = v * scale;
        const _Float16 h0 = (_Float16)(float)x;
        const _Float16 h1 = (_Float16)(float)((x - (double)h0) * 2048.0);
        o[0] = __builtin_bit_cast(unsigned short, h0);
        o[nq] = __builtin_bit_cast(unsigned short, h1);
    } else {
        constexpr int NP = TYPE == HK_BF16X2 ? 2 : 3;
        unsigned short *o = reinterpret_cast<unsigned short *>(out) + row + l;
        double r = v;
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const __bf16 h = (__bf16)(float)r;
            o[(long long)q * nq] = __builtin_bit_cast(unsigned short, h);
            r -= (double)(float)h;
        }
    }
}

// the power of two that puts max in [2^13, 2^14) (1 for a zero row)
__device__ inline double f16_row_scale(double max) { return max > 0.0 ? ldexp(1.0, 13 - ilogb(max)) : 1.0; }

// Adds the Lagrange-weighted item v at grid position t to the bins (real and imaginary parts
// in separate LDS arrays: fewer bank conflicts than interleaved pairs).
__device__ inline void spread_item(double *acc_re, double *acc_im, int nq, double t, double2 v)
{
    const double b = floor(t);
    const Lagrange lw(t - b);
    const int l0 = (int)b + 1;  // bin of node offset -(P/2 - 1): Q_l = (l - P/2) h
    if (l0 < 0 || l0 + HK_GRID_P > nq) return;  // the host sizes the grid so this never holds
#pragma unroll
    for (int i = 0; i < HK_GRID_P; i++) {
        atomicAdd(&acc_re[l0 + i], lw.L[i] * v.x);
        atomicAdd(&acc_im[l0 + i], lw.L[i] * v.y);
    }
}

template <int TYPE>
__device__ inline void store_bins(const double *acc_re, const double *acc_im, int nq, void *out, long long ld,
                                  long long re_row)
{
    double s_re = 1.0, s_im = 1.0;
    if constexpr (TYPE == HK_F16X2) {  // every thread of the workgroup takes part
        __shared__ double part[2][kSpreadThreads / 64];
        __shared__ double scale[2];
        double m_re = 0.0, m_im = 0.0;
        for (int l = threadIdx.x; l < nq; l += blockDim.x) {
            m_re = fmax(m_re, fabs(acc_re[l]));
            m_im = fmax(m_im, fabs(acc_im[l]));
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            m_re = fmax(m_re, __shfl_xor(m_re, d, 64));
            m_im = fmax(m_im, __shfl_xor(m_im, d, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            part[0][threadIdx.x >> 6] = m_re;
            part[1][threadIdx.x >> 6] = m_im;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, b = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
                a = fmax(a, part[0][w]);
                b = fmax(b, part[1][w]);
            }
            scale[0] = f16_row_scale(a);
            scale[1] = f16_row_scale(b);
            double *o = reinterpret_cast<double *>(reinterpret_cast<unsigned short *>(out) + re_row * ld + 2 * nq);
            o[0] = scale[0];
            o = reinterpret_cast<double *>(reinterpret_cast<unsigned short *>(out) + (re_row + 1) * ld + 2 * nq);
            o[0] = scale[1];
        }
        __syncthreads();
        s_re = scale[0];
        s_im = scale[1];
    }
    for (int l = threadIdx.x; l < nq; l += blockDim.x) {
        store_operand<TYPE>(out, re_row * ld, l, nq, acc_re[l], s_re);
        store_operand<TYPE>(out, (re_row + 1) * ld, l, nq, acc_im[l], s_im);
    }
}

} // namespace

// one workgroup per (frequency, output); the nq bins of C(., omega) accumulate in LDS (FP64
// atomics).  The (plane, radius) pairs of the planes with non-zero weight are dealt to the 1024
// threads in one flat range, advanced with 32-bit counters, so every wave has work whatever nr
// is.  The general form: per-plane factors recomputed per item (used when hk_spread_window_kernel's
// table of them does not fit the LDS beside the bins).
template <int TYPE>
__global__ void __launch_bounds__(kSpreadThreads)
hk_spread_kernel(hk_planes p, const double *weights, int nq, double h, double inv_c, void *out, long long ld,
                 long long row0)
{
    extern __shared__ double bins[];
    __shared__ int k_range[2];
    double *acc_re = bins, *acc_im = bins + nq;
    const int o = blockIdx.x;
    weights += (long long)blockIdx.y * p.nz;  // output blockIdx.y
    row0 += 2LL * p.no * blockIdx.y;
    const double K = p.ogrid[o] * inv_c;
    const double inv_h = 1.0 / h;
    for (int l = threadIdx.x; l < 2 * nq; l += blockDim.x) bins[l] = 0.0;
    if (threadIdx.x == 0) {
        k_range[0] = p.nz;
        k_range[1] = -1;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < p.nz; k += blockDim.x) {
        if (weights[k] != 0.0) {
            atomicMin(&k_range[0], k);
            atomicMax(&k_range[1], k);
        }
    }
    __syncthreads();
    const int k_end = k_range[1] + 1;
    const bool c64 = p.src_c64 != 0;
    const int step_k = blockDim.x / p.nr, step_j = blockDim.x - step_k * p.nr;
    int k = k_range[0] + threadIdx.x / p.nr, j = threadIdx.x % p.nr;
    for (; k < k_end; k += step_k, j += step_j) {
        if (j >= p.nr) {  // step_j < nr: one wrap at most
            j -= p.nr;
            if (++k >= k_end) break;
        }
        const double wk = weights[k];
        if (wk == 0.0) continue;
        const double D = p.distance - p.zgrid[k];
        double2 pk = make_double2(1.0, 0.0);
        if (p.pre_mode == 1) pk = reinterpret_cast<const double2 *>(p.pre)[(long long)k * p.no + o];
        double2 v = c64 ? source_term(p, reinterpret_cast<const float2 *>(p.src) + (long long)k * p.plane_stride, k, o,
                                      j, K, D, pk)
                        : source_term(p, reinterpret_cast<const double2 *>(p.src) + (long long)k * p.plane_stride, k,
                                      o, j, K, D, pk);
        v.x *= wk;
        v.y *= wk;
        spread_item(acc_re, acc_im, nq, (K / D) * p.rgrid[j] * inv_h, v);
    }
    __syncthreads();
    store_bins<TYPE>(acc_re, acc_im, nq, out, ld, row0 + 2LL * o);
}

template __global__ void hk_spread_kernel<HK_F64>(hk_planes, const double *, int, double, double, void *, long long,
                                                  long long);
template __global__ void hk_spread_kernel<HK_F32>(hk_planes, const double *, int, double, double, void *, long long,
                                                  long long);
template __global__ void hk_spread_kernel<HK_BF16X2>(hk_planes, const double *, int, double, double, void *, long long,
                                                     long long);
template __global__ void hk_spread_kernel<HK_BF16X3>(hk_planes, const double *, int, double, double, void *, long long,
                                                     long long);
template __global__ void hk_spread_kernel<HK_F16X2>(hk_planes, const double *, int, double, double, void *, long long,
                                                    long long);

// The same spread, accumulated in registers.  A thread owns one radius r_j (and a contiguous
// range of planes) and walks its planes in order.  Its items' bins l0 = floor(K r_j / (D_k h)) + 1
// hardly move from plane to plane (D_k changes little), so the thread keeps the 8 bins
// l0 .. l0 + 7 of its current stencil in registers, adds each item's 8 Lagrange-weighted terms
// there, and writes a bin to the LDS (FP64 atomic) only when the stencil moves past it: for
// ascending planes a radius flushes (its bin span over the planes + 8) bins, instead of 8 per
// item.  Per-plane factors (K / D / h, -K / (2 D), weight, pre-factor) sit in LDS after the
// bins (kSpreadTab doubles per plane).  Any plane order is handled (a backward or far jump
// flushes the whole stencil).
// Arithmetic of the register stencil: double, or float where the operand keeps fewer bits
// than float does (two bf16 parts hold 16, two scaled fp16 parts 22).
template <typename R>
using cplx = typename std::conditional<std::is_same<R, double>::value, double2, float2>::type;
template <typename R>
__device__ inline cplx<R> mkc(R re, R im)
{
    cplx<R> v;
    v.x = re;
    v.y = im;
    return v;
}
template <typename R>
__device__ inline cplx<R> cmulr(cplx<R> a, cplx<R> b)
{
    return mkc<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename R>
struct Window {
    R re[HK_GRID_P], im[HK_GRID_P];
    int l0;  // bin of re[0]; below 0: empty
};

template <typename R>
__device__ inline void window_flush_all(Window<R> &w, double *acc_re, double *acc_im)
{
#pragma unroll
    for (int m = 0; m < HK_GRID_P; m++) {
        atomicAdd(&acc_re[w.l0 + m], (double)w.re[m]);
        atomicAdd(&acc_im[w.l0 + m], (double)w.im[m]);
        w.re[m] = R(0);
        w.im[m] = R(0);
    }
}

// adds v times the Lagrange basis at f (the same weights as Lagrange(f)) to the stencil, each
// weight formed as it is used: the suffix products are kept, the prefix product runs
template <typename R>
__device__ inline void window_add(Window<R> &w, R f, cplx<R> v)
{
    constexpr int P = HK_GRID_P;
    R suf[P];
    suf[P - 1] = R(1);
#pragma unroll
    for (int i = P - 2; i >= 0; i--) suf[i] = suf[i + 1] * (f - R(i + 1 - (P / 2 - 1)));
    R pre = R(1);
#pragma unroll
    for (int i = 0; i < P; i++) {
        double den = 1.0;
        for (int q = 2; q <= i; q++) den *= (double)q;
        for (int q = 2; q <= P - 1 - i; q++) den *= (double)q;
        const R inv = R(((P - 1 - i) & 1 ? -1.0 : 1.0) / den);  // folded at compile time
        const R L = (pre * suf[i]) * inv;
        w.re[i] = fma(L, v.x, w.re[i]);
        w.im[i] = fma(L, v.y, w.im[i]);
        pre *= f - R(i - (P / 2 - 1));
    }
}

// moves the stencil to start at bin l0 (0 <= l0 <= nq - HK_GRID_P): a step of one bin (what
// ascending planes give) flushes the bin left behind and shifts; any other move flushes all
template <typename R>
__device__ inline void window_move(Window<R> &w, int l0, double *acc_re, double *acc_im)
{
    if (l0 == w.l0) return;
    if (l0 == w.l0 + 1 && w.l0 >= 0) {
        atomicAdd(&acc_re[w.l0], (double)w.re[0]);
        atomicAdd(&acc_im[w.l0], (double)w.im[0]);
#pragma unroll
        for (int m = 0; m < HK_GRID_P - 1; m++) {
            w.re[m] = w.re[m + 1];
            w.im[m] = w.im[m + 1];
        }
        w.re[HK_GRID_P - 1] = R(0);
        w.im[HK_GRID_P - 1] = R(0);
    } else if (w.l0 >= 0) {
        window_flush_all(w, acc_re, acc_im);
    }
    w.l0 = l0;
}

// A plane's factors for one frequency, read from the LDS table in one go.
struct PlaneTab {
    double s, c, w, pre_re, pre_im;
};
__device__ inline PlaneTab plane_tab(const double *tab, int k)
{
    const double2 *t = reinterpret_cast<const double2 *>(tab + kSpreadTab * k);  // 16-byte aligned
    const double2 a = t[0], b = t[1], c = t[2];
    return {a.x, a.y, b.x, b.y, c.x};
}

// e^{i x} for |x| <= 2^-7: Taylor series through x^7 (sin) and x^8 (cos), first omitted terms
// below 1e-24
template <typename R>
__device__ inline cplx<R> expi_small(R x)
{
    const R z = x * x;
    R ps = R(-1.0 / 5040.0);
    ps = fma(ps, z, R(1.0 / 120.0));
    ps = fma(ps, z, R(-1.0 / 6.0));
    R pc = R(1.0 / 40320.0);
    pc = fma(pc, z, R(-1.0 / 720.0));
    pc = fma(pc, z, R(1.0 / 24.0));
    pc = fma(pc, z, R(-0.5));
    return mkc<R>(fma(z, pc, R(1)), fma(x * z, ps, x));
}

constexpr int kPhaseAnchor = 32;  // planes between exact near-field phases

// S: the source element type (double2 or float2), kept as loaded while it is in flight.
// NEAR: 0 no near-field phase, 1 its sincos by sincos_reduced (|phase| < kSincosLimit / 2,
// checked by the host), 2 by the general sincos.
// R: float for the two-bf16-part and the scaled two-fp16-part operands (16 and 22 bits kept;
// the item's grid position t and the phase steps' angle stay in double), double otherwise.  For
// the fp16 parts this moves the far field's error from 3.9e-8 to 1.2e-7 of its maximum and the
// pass from 2.27 to 1.89-1.94 ms.
template <int TYPE, typename S, int NEAR>
__global__ void __launch_bounds__(kSpreadThreads)
hk_spread_window_kernel(hk_planes p, const double *weights, int nq, double h, double inv_c, void *out, long long ld,
                        long long row0)
{
    using R = typename std::conditional<TYPE == HK_BF16X2 || TYPE == HK_F16X2, float, double>::type;
    extern __shared__ double bins[];
    __shared__ int k_range[2];
    double *acc_re = bins, *acc_im = bins + nq, *tab = bins + 2 * nq;
    const int o = blockIdx.x;
    weights += (long long)blockIdx.y * p.nz;  // output blockIdx.y
    row0 += 2LL * p.no * blockIdx.y;
    const double K = p.ogrid[o] * inv_c;
    const double inv_h = 1.0 / h;
    for (int l = threadIdx.x; l < 2 * nq; l += blockDim.x) bins[l] = 0.0;
    if (threadIdx.x == 0) {
        k_range[0] = p.nz;
        k_range[1] = -1;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < p.nz; k += blockDim.x) {
        const double wk = weights[k];
        if (wk != 0.0) {
            atomicMin(&k_range[0], k);
            atomicMax(&k_range[1], k);
        }
        const double D = p.distance - p.zgrid[k];
        double2 pk = make_double2(1.0, 0.0);
        if (p.pre_mode == 1) pk = reinterpret_cast<const double2 *>(p.pre)[(long long)k * p.no + o];
        double *t = tab + kSpreadTab * k;
        t[0] = (K / D) * inv_h;
        t[1] = -K / (2.0 * D);
        t[2] = wk;
        t[3] = pk.x;
        t[4] = pk.y;
    }
    __syncthreads();
    const int k_lo = k_range[0], nk = k_range[1] + 1 - k_lo;
    // thread -> (radius, plane group): a group is a whole number of waves (so a wave's plane is
    // uniform), several groups of contiguous planes when the radii leave waves over
    const int span = (p.nr + 63) & ~63;
    const int groups = span >= (int)blockDim.x ? 1 : min((int)blockDim.x / span, max(nk, 1));
    const int g = groups == 1 ? 0 : threadIdx.x / span;
    if (nk > 0 && g < groups) {
        const int ka = k_lo + (int)((long long)nk * g / groups), kb = k_lo + (int)((long long)nk * (g + 1) / groups);
        const int first = groups == 1 ? threadIdx.x : threadIdx.x - g * span;
        const int stride = groups == 1 ? (int)blockDim.x : span;
        for (int j = first; j < p.nr; j += stride) {
            const double r = p.rgrid[j];
            const double wr = p.tw[j] * r;
            const S *src = reinterpret_cast<const S *>(p.src) + (long long)ka * p.plane_stride + (long long)o * p.nr + j;
            Window<R> w;
            w.l0 = -1;
#pragma unroll
            for (int m = 0; m < HK_GRID_P; m++) {
                w.re[m] = R(0);
                w.im[m] = R(0);
            }
            // the near-field phase e^{i c_k r^2} advances from plane to plane by the small angle
            // (c_k - c_prev) r^2 (consecutive planes: ~K r^2 dz / (2 D^2)); an exact sincos where
            // that angle is not small, and every kPhaseAnchor planes
            cplx<R> ph = mkc<R>(R(1), R(0));
            double c_prev = 0.0;
            int since = kPhaseAnchor;
            const double r2 = r * r;
            S f1 = src[0];
            for (int k = ka; k < kb; k++) {
                const S fs = f1;
                src += p.plane_stride;
                if (k + 1 < kb) f1 = src[0];  // the next plane's source in flight
                const PlaneTab tk = plane_tab(tab, k);
                if (tk.w == 0.0) continue;
                cplx<R> v = cmulr<R>(mkc<R>(R(tk.pre_re), R(tk.pre_im)), mkc<R>(R(fs.x), R(fs.y)));
                if constexpr (NEAR != 0) {
                    const double d = (tk.c - c_prev) * r2;  // in double: c_k - c_prev cancels
                    if (since < kPhaseAnchor && fabs(d) <= 0x1p-7) {
                        ph = cmulr<R>(ph, expi_small<R>(R(d)));
                        since++;
                    } else {
                        // the anchor in double for either R: a float argument would carry an
                        // absolute error of ulp(c r^2) into every phase up to the next anchor
                        double sd, cd;
                        if constexpr (NEAR == 1)
                            sincos_reduced(tk.c * r2, &sd, &cd);
                        else
                            sincos(tk.c * r2, &sd, &cd);
                        ph = mkc<R>(R(cd), R(sd));
                        since = 0;
                    }
                    c_prev = tk.c;
                    v = cmulr<R>(ph, v);
                }
                const R wv = R(wr * tk.w);
                v.x *= wv;
                v.y *= wv;
                const double t = tk.s * r;  // in double: its fraction places the item on the grid
                const double b = floor(t);
                const int l0 = (int)b + 1;  // bin of node offset -(P/2 - 1): Q_l = (l - P/2) h
                if (l0 < 0 || l0 + HK_GRID_P > nq) continue;  // the host sizes the grid so this never holds
                window_move(w, l0, acc_re, acc_im);
                window_add(w, R(t - b), v);
            }
            if (w.l0 >= 0) window_flush_all(w, acc_re, acc_im);
        }
    }
    __syncthreads();
    store_bins<TYPE>(acc_re, acc_im, nq, out, ld, row0 + 2LL * o);
}

#define HK_SPREAD_WINDOW(TYPE, S, NEAR)                                                                      \
    template __global__ void hk_spread_window_kernel<TYPE, S, NEAR>(hk_planes, const double *, int, double, double, \
                                                                    void *, long long, long long);
HK_SPREAD_WINDOW(HK_F64, double2, 0)
HK_SPREAD_WINDOW(HK_F64, double2, 1)
HK_SPREAD_WINDOW(HK_F64, double2, 2)
HK_SPREAD_WINDOW(HK_F64, float2, 0)
HK_SPREAD_WINDOW(HK_F64, float2, 1)
HK_SPREAD_WINDOW(HK_F64, float2, 2)
HK_SPREAD_WINDOW(HK_F32, double2, 0)
HK_SPREAD_WINDOW(HK_F32, double2, 1)
HK_SPREAD_WINDOW(HK_F32, double2, 2)
HK_SPREAD_WINDOW(HK_F32, float2, 0)
HK_SPREAD_WINDOW(HK_F32, float2, 1)
HK_SPREAD_WINDOW(HK_F32, float2, 2)
HK_SPREAD_WINDOW(HK_BF16X2, double2, 0)
HK_SPREAD_WINDOW(HK_BF16X2, double2, 1)
HK_SPREAD_WINDOW(HK_BF16X2, double2, 2)
HK_SPREAD_WINDOW(HK_BF16X2, float2, 0)
HK_SPREAD_WINDOW(HK_BF16X2, float2, 1)
HK_SPREAD_WINDOW(HK_BF16X2, float2, 2)
HK_SPREAD_WINDOW(HK_BF16X3, double2, 0)
HK_SPREAD_WINDOW(HK_BF16X3, double2, 1)
HK_SPREAD_WINDOW(HK_BF16X3, double2, 2)
HK_SPREAD_WINDOW(HK_BF16X3, float2, 0)
HK_SPREAD_WINDOW(HK_BF16X3, float2, 1)
HK_SPREAD_WINDOW(HK_BF16X3, float2, 2)
HK_SPREAD_WINDOW(HK_F16X2, double2, 0)
HK_SPREAD_WINDOW(HK_F16X2, double2, 1)
HK_SPREAD_WINDOW(HK_F16X2, double2, 2)
HK_SPREAD_WINDOW(HK_F16X2, float2, 0)
HK_SPREAD_WINDOW(HK_F16X2, float2, 1)
HK_SPREAD_WINDOW(HK_F16X2, float2, 2)
#undef HK_SPREAD_WINDOW

// J0(rho_m |Q_l|), Q_l = (l - P/2) h
template <int TYPE>
__global__ void hk_bessel_kernel(int nrho, const double *rho, int nq, double h, void *out, long long ld)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nrho * nq) return;
    const int m = (int)(idx / nq), l = (int)(idx % nq);
    const double q = (double)(l - HK_GRID_P / 2) * h;
    if constexpr (TYPE == HK_F16X2) {  // |J0| <= 1: one scale for every row
        const double scale = 8192.0;
        if (l == 0) *reinterpret_cast<double *>(reinterpret_cast<unsigned short *>(out) + (long long)m * ld + 2 * nq) = scale;
        store_operand<TYPE>(out, (long long)m * ld, l, nq, hk::j0(rho[m] * q), scale);
    } else {
        store_operand<TYPE>(out, (long long)m * ld, l, nq, hk::j0(rho[m] * q));
    }
}

template __global__ void hk_bessel_kernel<HK_F64>(int, const double *, int, double, void *, long long);
template __global__ void hk_bessel_kernel<HK_F32>(int, const double *, int, double, void *, long long);
template __global__ void hk_bessel_kernel<HK_BF16X2>(int, const double *, int, double, void *, long long);
template __global__ void hk_bessel_kernel<HK_BF16X3>(int, const double *, int, double, void *, long long);
template __global__ void hk_bessel_kernel<HK_F16X2>(int, const double *, int, double, void *, long long);
