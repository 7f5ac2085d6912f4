// The contraction of the Hankel far-field integral over r and the source planes
// (Hankel_transform.py:492-514 per plane, :641-661 over planes):
//   * hk_direct_kernel: every (plane, omega, rho, r) element with J0 in FP64;
//   * hk_spread_kernel + hk_bessel_kernel: the gridded form, C_l(omega) spread from the
//     sources by Lagrange interpolation in q = K r / D and the matrix J0(rho_m Q_l) whose
//     product (hankel_gemm.hip) gives the same weighted sum over planes.
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
constexpr int kSpreadThreads = 1024;  // hk_spread_kernel's workgroup
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

template <int TYPE>
__device__ inline void store_operand(void *out, long long at, double v)
{
    if constexpr (TYPE == HK_F64)
        reinterpret_cast<double *>(out)[at] = v;
    else
        reinterpret_cast<float *>(out)[at] = (float)v;
}

} // namespace

// one workgroup per (frequency, output); the nq bins of C(., omega) accumulate in LDS (FP64
// atomics, real and imaginary parts in separate arrays: fewer bank conflicts than interleaved
// pairs).
// The (plane, radius) pairs of the planes with non-zero weight are dealt to the 1024 threads
// in one flat range, advanced with 32-bit counters, so every wave has work whatever nr is.
// TAB: the per-plane factors (K / D_k / h, -K / (2 D_k), w_k, pre_k) sit in LDS after the
// bins (5 doubles per plane), so an item does no division and no pre-factor load.
template <int TYPE, bool TAB>
__global__ void __launch_bounds__(kSpreadThreads)
hk_spread_kernel(hk_planes p, const double *weights, int nq, double h, double inv_c, void *out, long long ld,
                 long long row0)
{
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
        if constexpr (TAB) {
            const double D = p.distance - p.zgrid[k];
            double2 pk = make_double2(1.0, 0.0);
            if (p.pre_mode == 1) pk = reinterpret_cast<const double2 *>(p.pre)[(long long)k * p.no + o];
            double *t = tab + 5 * k;
            t[0] = (K / D) * inv_h;
            t[1] = -K / (2.0 * D);
            t[2] = wk;
            t[3] = pk.x;
            t[4] = pk.y;
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
        double2 v;
        double t;
        if constexpr (TAB) {
            const double *tk = tab + 5 * k;
            const double wk = tk[2];
            if (wk == 0.0) continue;
            const double r = p.rgrid[j];
            const long long at = (long long)k * p.plane_stride + (long long)o * p.nr + j;
            const double2 f = c64 ? load_src(reinterpret_cast<const float2 *>(p.src), at)
                                  : load_src(reinterpret_cast<const double2 *>(p.src), at);
            v = cmul(make_double2(tk[3], tk[4]), f);
            if (p.near_field) {
                double sn, cs;
                sincos(tk[1] * (r * r), &sn, &cs);
                v = cmul(make_double2(cs, sn), v);
            }
            const double wr = (p.tw[j] * r) * wk;
            v.x *= wr;
            v.y *= wr;
            t = tk[0] * r;
        } else {
            const double wk = weights[k];
            if (wk == 0.0) continue;
            const double D = p.distance - p.zgrid[k];
            const double s = K / D;
            double2 pk = make_double2(1.0, 0.0);
            if (p.pre_mode == 1) pk = reinterpret_cast<const double2 *>(p.pre)[(long long)k * p.no + o];
            v = c64 ? source_term(p, reinterpret_cast<const float2 *>(p.src) + (long long)k * p.plane_stride, k, o, j,
                                  K, D, pk)
                    : source_term(p, reinterpret_cast<const double2 *>(p.src) + (long long)k * p.plane_stride, k, o,
                                  j, K, D, pk);
            v.x *= wk;
            v.y *= wk;
            t = s * p.rgrid[j] * inv_h;
        }
        const double b = floor(t);
        const Lagrange lw(t - b);
        const int l0 = (int)b + 1;  // bin of node offset -(P/2 - 1): Q_l = (l - P/2) h
        if (l0 < 0 || l0 + HK_GRID_P > nq) continue;  // the host sizes the grid so this never holds
#pragma unroll
        for (int i = 0; i < HK_GRID_P; i++) {
            atomicAdd(&acc_re[l0 + i], lw.L[i] * v.x);
            atomicAdd(&acc_im[l0 + i], lw.L[i] * v.y);
        }
    }
    __syncthreads();
    const long long re_row = row0 + 2LL * o, im_row = re_row + 1;
    for (int l = threadIdx.x; l < nq; l += blockDim.x) {
        store_operand<TYPE>(out, re_row * ld + l, acc_re[l]);
        store_operand<TYPE>(out, im_row * ld + l, acc_im[l]);
    }
}

template __global__ void hk_spread_kernel<HK_F64, false>(hk_planes, const double *, int, double, double, void *,
                                                         long long, long long);
template __global__ void hk_spread_kernel<HK_F32, false>(hk_planes, const double *, int, double, double, void *,
                                                         long long, long long);
template __global__ void hk_spread_kernel<HK_F64, true>(hk_planes, const double *, int, double, double, void *,
                                                        long long, long long);
template __global__ void hk_spread_kernel<HK_F32, true>(hk_planes, const double *, int, double, double, void *,
                                                        long long, long long);

// J0(rho_m |Q_l|), Q_l = (l - P/2) h
template <int TYPE>
__global__ void hk_bessel_kernel(int nrho, const double *rho, int nq, double h, void *out, long long ld)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nrho * nq) return;
    const int m = (int)(idx / nq), l = (int)(idx % nq);
    const double q = (double)(l - HK_GRID_P / 2) * h;
    store_operand<TYPE>(out, (long long)m * ld + l, hk::j0(rho[m] * q));
}

template __global__ void hk_bessel_kernel<HK_F64>(int, const double *, int, double, void *, long long);
template __global__ void hk_bessel_kernel<HK_F32>(int, const double *, int, double, void *, long long);
