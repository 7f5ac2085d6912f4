// C ABI of libmmahankel.so (include/hankel.h): argument checks, launch shapes and the error
// status; every entry point returns after the work it queued on `stream` has finished, except
// the gridded contraction's steps (spread, Bessel matrix, GEMM), which return once queued.
#include <algorithm>
#include <cstdint>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "hankel_internal.h"

namespace {

thread_local int g_status = HK_OK;
thread_local std::string g_message;

int fail(int code, const std::string &msg)
{
    g_status = code;
    g_message = msg;
    return code;
}
void clear_status()
{
    g_status = HK_OK;
    g_message.clear();
}

#define HK_HIP(call)                                                                                         \
    do {                                                                                                     \
        const hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(HK_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// shared_python/units.py:3-20 and XUV_refractive_index.py:123, in the reference's operation order
HkConst make_constants()
{
    const double hbar = 1.0545718e-34, inverse_alpha_fine = 137.035999139, c_light = 299792458.0;
    const double elcharge = 1.602176565e-19, elmass = 9.10938356e-31;
    const double r_bohr = hbar * inverse_alpha_fine / (c_light * elmass);
    const double alpha_fine = 1.0 / inverse_alpha_fine;
    const double alpha2 = alpha_fine * alpha_fine;
    const double r_electron = r_bohr * alpha2;
    HkConst K;
    K.c_light = c_light;
    K.timeau = (elmass * (r_bohr * r_bohr)) / hbar;
    K.ev_per_au = elcharge / ((elmass * alpha2) * (c_light * c_light));
    K.lam_num = 2.0 * M_PI * hbar;
    K.lam_den = (elmass * c_light) * alpha2;
    K.n_ref_re = 2.7e25 * r_electron;
    return K;
}

const double kInvC = 1.0 / 299792458.0;
// largest q grid: its bins (16 B each) and the spread kernel's few static bytes fit the 160 KiB of LDS
const int kMaxBins = 10176;

int check_planes(const hk_planes *p)
{
    if (!p) return fail(HK_ERR_ARGUMENT, "null plane description");
    if (p->nz < 1 || p->no < 1 || p->nr < 2 || p->nrho < 1)
        return fail(HK_ERR_ARGUMENT, "need nz >= 1, no >= 1, nr >= 2, nrho >= 1");
    if (!p->src || !p->zgrid || !p->ogrid || !p->rgrid || !p->tw || !p->rho)
        return fail(HK_ERR_ARGUMENT, "null device array");
    if (p->pre_mode < 0 || p->pre_mode > 2 || (p->pre_mode != 0 && !p->pre))
        return fail(HK_ERR_ARGUMENT, "pre_mode must be 0 (none), 1 (plane x omega) or 2 (plane x omega x r)");
    if (p->plane_stride < (long long)p->no * p->nr) return fail(HK_ERR_ARGUMENT, "plane_stride < no * nr");
    return HK_OK;
}

// host copies of the small grids of a plane description
struct HostGrids {
    std::vector<double> r, z, o;
};

// one grid on the host: from the caller's host copy, else from the device
int host_grid(std::vector<double> &v, const double *host, const double *dev, int n)
{
    v.resize(n);
    if (host)
        std::copy(host, host + n, v.begin());
    else
        HK_HIP(hipMemcpy(v.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost));
    return HK_OK;
}

int host_grids(const hk_planes *p, HostGrids &g)
{
    if (host_grid(g.r, p->host_rgrid, p->rgrid, p->nr) != HK_OK || host_grid(g.z, p->host_zgrid, p->zgrid, p->nz) != HK_OK ||
        host_grid(g.o, p->host_ogrid, p->ogrid, p->no) != HK_OK)
        return g_status;
    return HK_OK;
}

// largest q = (omega / c) r / (distance - z) of the sources; requires r >= 0, omega >= 0 and
// every plane in front of the screen
int source_q_max(const hk_planes *p, const HostGrids &g, double *q_max)
{
    double rmax = 0.0, omax = 0.0, dmin = INFINITY;
    for (double v : g.r) {
        if (!(v >= 0.0)) return fail(HK_ERR_ARGUMENT, "the gridded contraction needs radii >= 0");
        rmax = std::max(rmax, v);
    }
    for (double v : g.o) {
        if (!(v >= 0.0)) return fail(HK_ERR_ARGUMENT, "the gridded contraction needs frequencies >= 0");
        omax = std::max(omax, v);
    }
    for (double v : g.z) dmin = std::min(dmin, p->distance - v);
    if (!(dmin > 0.0)) return fail(HK_ERR_ARGUMENT, "every source plane must lie before the screen");
    *q_max = omax * kInvC / dmin * rmax;
    return HK_OK;
}

int operand_parts(int type) { return type == HK_BF16X2 || type == HK_F16X2 ? 2 : type == HK_BF16X3 ? 3 : 1; }
// elements of the type an operand row needs at least (HK_F16X2: the parts and the row scale)
long long operand_row(int type, int nq) { return (long long)operand_parts(type) * nq + (type == HK_F16X2 ? 4 : 0); }

// r_j = r_0 + j dr to within a few ulps of max |r| (then angles advanced by rotation stay within
// ~1e-13 rad of the exact ones)
bool uniform_radii(const std::vector<double> &r, double *dr_out)
{
    const double dr = (r.back() - r.front()) / (double)(r.size() - 1);
    double rmax = 0.0;
    for (double v : r) rmax = std::max(rmax, std::fabs(v));
    bool uniform = dr > 0;
    for (size_t j = 1; j < r.size() && uniform; j++)
        uniform = std::fabs(r[j] - (r.front() + (double)j * dr)) <= 8.0 * 2.220446049250313e-16 * rmax;
    *dr_out = dr;
    return uniform;
}

} // namespace

namespace {

// The argument checks shared by both pre-factor entry points.
int check_prefactor_args(const hk_medium *m, int nz, const double *zgrid, int no, const double *ogrid, int nr,
                         const double *rgrid, double *pre, double *renorm)
{
    if (!m || nz < 1 || no < 1 || !zgrid || !ogrid || !pre) return fail(HK_ERR_ARGUMENT, "bad pre-factor arguments");
    if (m->mode < HK_PRESSURE_SCALAR || m->mode > HK_PRESSURE_ZR) return fail(HK_ERR_ARGUMENT, "unknown pressure mode");
    if (m->mode != HK_PRESSURE_ZR && !renorm)
        return fail(HK_ERR_ARGUMENT, "renorm is required for scalar and z-tabulated pressure");
    if (m->mode != HK_PRESSURE_SCALAR && (m->np_z < 1 || !m->p_z || !m->p_value))
        return fail(HK_ERR_ARGUMENT, "empty pressure table");
    if (m->mode == HK_PRESSURE_ZR && (nr < 1 || !rgrid || m->np_r < 1 || !m->p_r))
        return fail(HK_ERR_ARGUMENT, "z-r pressure table needs the r grids");
    if ((m->include_dispersion && (m->n_f1 < 1 || !m->e_f1 || !m->f1)) ||
        (m->include_absorption && (m->n_f2 < 1 || !m->e_f2 || !m->f2)))
        return fail(HK_ERR_ARGUMENT, "missing scattering-factor table");
    return HK_OK;
}

// Queues the pre-factor kernels after clearing *flag (a fill kernel, no memset node): nothing
// waits.  exit_abs: no doubles of device scratch (z-tabulated pressure).
hipError_t enqueue_prefactor(const hk_medium *m, int nz, const double *zgrid, int no, const double *ogrid, int nr,
                             const double *rgrid, double *pre, double *renorm, int *flag, double *exit_abs,
                             hipStream_t s)
{
    hipLaunchKernelGGL(hk_fill_int_kernel, dim3(1), dim3(64), 0, s, flag, 1, 0);
    const int n = m->mode == HK_PRESSURE_ZR ? no * nr : no;
    // the scalar branch: one thread per (frequency, plane)
    const unsigned planes = m->mode == HK_PRESSURE_SCALAR ? (unsigned)std::min(nz, 65535) : 1u;
    // the z-tabulated walk reads the plane grid and the pressure table from LDS when they fit
    const size_t stage = (size_t)(nz + 2 * m->np_z) * sizeof(double);
    const int staged = m->mode == HK_PRESSURE_Z && stage <= 48 * 1024;
    if (m->mode == HK_PRESSURE_Z && m->np_z <= kWalkParallelMax && nz <= 65535)
        hipLaunchKernelGGL(hk_prefactor_z_kernel, dim3((no + 127) / 128, nz), dim3(128), 0, s, make_constants(), *m,
                           nz, zgrid, no, ogrid, reinterpret_cast<double2 *>(pre), renorm, exit_abs, flag);
    else
        hipLaunchKernelGGL(hk_prefactor_kernel, dim3((n + 127) / 128, planes), dim3(128), staged ? stage : 0, s,
                           make_constants(), *m, nz, zgrid, no, ogrid, nr, rgrid, reinterpret_cast<double2 *>(pre),
                           renorm, exit_abs, flag, staged);
    if (m->mode == HK_PRESSURE_Z) {
        const long long total = (long long)nz * no;
        hipLaunchKernelGGL(hk_prefactor_finish_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, nz, no,
                           ogrid, exit_abs, reinterpret_cast<double2 *>(pre), renorm);
    }
    return hipGetLastError();
}

} // namespace

extern "C" {

int hk_last_error(void) { return g_status; }
const char *hk_last_error_message(void) { return g_message.c_str(); }


int hk_prefactor(const hk_medium *m, int nz, const double *zgrid, int no, const double *ogrid, int nr,
                 const double *rgrid, double *pre, double *renorm, void *stream)
{
    clear_status();
    if (check_prefactor_args(m, nz, zgrid, no, ogrid, nr, rgrid, pre, renorm) != HK_OK) return g_status;
    hipStream_t s = (hipStream_t)stream;
    // scratch kept between calls (an allocation and a free per call would each synchronise the
    // device): the failure flag and, per frequency, the z-tabulated absorption at the exit plane
    static std::mutex scratch_mutex;
    static int scratch_dev = -1;
    static int *flag = nullptr;
    static double *exit_abs = nullptr;
    static int exit_n = 0;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    int dev = 0;
    HK_HIP(hipGetDevice(&dev));
    if (dev != scratch_dev || no > exit_n) {
        if (flag) (void)hipFree(flag);
        if (exit_abs) (void)hipFree(exit_abs);
        flag = nullptr;
        exit_abs = nullptr;
        scratch_dev = -1;
        exit_n = 0;
        HK_HIP(hipMalloc(&flag, sizeof(int)));
        HK_HIP(hipMalloc(&exit_abs, (size_t)no * sizeof(double)));
        scratch_dev = dev;
        exit_n = no;
    }
    int got = 0;
    hipError_t e = enqueue_prefactor(m, nz, zgrid, no, ogrid, nr, rgrid, pre, renorm, flag, exit_abs, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&got, flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(HK_ERR_DEVICE, std::string("pre-factor kernel: ") + hipGetErrorString(e));
    return hk_prefactor_status(got);
}

int hk_prefactor_enqueue(const hk_medium *m, int nz, const double *zgrid, int no, const double *ogrid, int nr,
                         const double *rgrid, double *pre, double *renorm, int *flag, double *work, void *stream)
{
    clear_status();
    if (check_prefactor_args(m, nz, zgrid, no, ogrid, nr, rgrid, pre, renorm) != HK_OK) return g_status;
    if (!flag || (m->mode == HK_PRESSURE_Z && !work))
        return fail(HK_ERR_ARGUMENT, "the deferred pre-factor needs a flag and, for a z-tabulated pressure, work");
    const hipError_t e = enqueue_prefactor(m, nz, zgrid, no, ogrid, nr, rgrid, pre, renorm, flag, work,
                                           (hipStream_t)stream);
    if (e != hipSuccess) return fail(HK_ERR_DEVICE, std::string("pre-factor kernel: ") + hipGetErrorString(e));
    return HK_OK;
}

int hk_prefactor_status(int flag)
{
    clear_status();
    if (flag == HK_FLAG_RANGE) return fail(HK_ERR_TABLE, "photon energy outside the tabulated scattering factors");
    if (flag == HK_FLAG_F1) return fail(HK_ERR_TABLE, "f1-scattering coefficient not listed for this frequency.");
    return HK_OK;
}

int hk_direct(const hk_planes *p, int n_out, const double *weights, int per_plane, double *out, void *stream)
{
    clear_status();
    if (check_planes(p) != HK_OK) return g_status;
    if (!out || (!per_plane && (n_out < 1 || n_out > 4 || !weights)))
        return fail(HK_ERR_ARGUMENT, "need out and 1..4 weighted outputs (or per_plane)");
    // a uniform radial grid lets the kernel rotate cos/sin of the Bessel argument
    std::vector<double> r;
    if (host_grid(r, p->host_rgrid, p->rgrid, p->nr) != HK_OK) return g_status;
    double dr = 0.0;
    const bool uniform = uniform_radii(r, &dr);
    hipStream_t s = (hipStream_t)stream;
    dim3 grid(p->no, (p->nrho + 255) / 256);
    if (p->src_c64)
        hipLaunchKernelGGL(hk_direct_kernel<float2>, grid, dim3(256), 0, s, *p, n_out, weights, per_plane,
                           reinterpret_cast<double2 *>(out), kInvC, (int)uniform, r.front(), dr);
    else
        hipLaunchKernelGGL(hk_direct_kernel<double2>, grid, dim3(256), 0, s, *p, n_out, weights, per_plane,
                           reinterpret_cast<double2 *>(out), kInvC, (int)uniform, r.front(), dr);
    HK_HIP(hipGetLastError());
    HK_HIP(hipStreamSynchronize(s));
    return HK_OK;
}

int hk_grid_size(double q_max, double rho_max, double tol, int *nq, double *h)
{
    clear_status();
    if (!(q_max > 0) || !(rho_max > 0) || !(tol > 0) || !nq || !h) return fail(HK_ERR_ARGUMENT, "bad grid request");
    // Lagrange interpolation of P points: |error| <= (rho h)^P max_f |prod (f - d_i)| / P!
    const int P = HK_GRID_P;
    double nodal = 1.0, fact = 1.0;
    for (int i = 0; i < P; i++) {
        nodal *= std::fabs(0.5 - (double)(i - (P / 2 - 1)));
        fact *= (double)(i + 1);
    }
    const double step = std::pow(tol / (nodal / fact), 1.0 / P);  // allowed rho_max * h
    const double need = rho_max * q_max / step + P + 1;
    const long long n = ((long long)std::ceil(need) + 63) / 64 * 64;
    if (n > kMaxBins)
        return fail(HK_ERR_UNSUPPORTED,
                    "q grid of " + std::to_string(n) + " points exceeds LDS (" + std::to_string(kMaxBins) + ")");
    *nq = (int)n;
    *h = q_max * (1.0 + 1e-12) / (double)(n - P - 1);
    return HK_OK;
}

int hk_grid_spread(const hk_planes *p, int n_out, const double *weights, int nq, double h, int type, void *out,
                   long long ld, long long row0, void *stream)
{
    clear_status();
    if (check_planes(p) != HK_OK) return g_status;
    if (n_out < 1 || n_out > 65535 || !weights || !out || nq < HK_GRID_P + 2 || nq > kMaxBins || ld < nq ||
        row0 < 0 || !(h > 0))
        return fail(HK_ERR_ARGUMENT, "bad spread arguments");
    if (p->pre_mode == 2) return fail(HK_ERR_UNSUPPORTED, "the gridded contraction needs an r-independent pre-factor");
    HostGrids g;
    double q_max = 0.0;
    if (host_grids(p, g) != HK_OK || source_q_max(p, g, &q_max) != HK_OK) return g_status;
    if (q_max / h > (double)(nq - HK_GRID_P - 1) * (1.0 + 1e-9))
        return fail(HK_ERR_ARGUMENT, "q grid too short for these sources (use hk_grid_size)");
    hipStream_t s = (hipStream_t)stream;
    if (type < HK_F64 || type > HK_F16X2) return fail(HK_ERR_ARGUMENT, "unknown operand type");
    if (ld < operand_row(type, nq) || (type == HK_F16X2 && ld % 4 != 0))
        return fail(HK_ERR_ARGUMENT, "row stride below what a row of this operand type holds");
    // bins, then (when they fit the LDS too) the per-plane factors of the register-window kernel
    const size_t lds_bins = (size_t)nq * sizeof(double2);
    const size_t lds_tab = lds_bins + (size_t)p->nz * kSpreadTab * sizeof(double);
    const bool tab = lds_tab <= (size_t)kMaxBins * sizeof(double2);
    const size_t lds = tab ? lds_tab : lds_bins;
    auto launch = [&](auto kernel) -> int {
        HK_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kernel, dim3(p->no, n_out), dim3(1024), lds, s, *p, weights, nq, h, kInvC, out, ld, row0);
        return HK_OK;
    };
    int st;
    // near-field phase K r^2 / (2 D): the reduced-range sincos covers it when below kSincosMax
    const bool c64 = p->src_c64 != 0;
    double r_max = 0.0, o_max = 0.0, d_min = INFINITY;
    for (double v : g.r) r_max = std::max(r_max, std::fabs(v));
    for (double v : g.o) o_max = std::max(o_max, std::fabs(v));
    for (double v : g.z) d_min = std::min(d_min, p->distance - v);
    const int near = !p->near_field ? 0 : (o_max * kInvC * r_max * r_max / (2.0 * d_min) < 0.5 * kSincosLimit ? 1 : 2);
    if (tab) {
        auto pick = [&](auto type_tag, auto src_tag) -> int {
            constexpr int T = decltype(type_tag)::value;
            using S = typename decltype(src_tag)::type;
            return near == 0 ? launch(hk_spread_window_kernel<T, S, 0>)
                   : near == 1 ? launch(hk_spread_window_kernel<T, S, 1>)
                               : launch(hk_spread_window_kernel<T, S, 2>);
        };
        struct D2 { using type = double2; };
        struct F2 { using type = float2; };
        auto by_src = [&](auto type_tag) { return c64 ? pick(type_tag, F2{}) : pick(type_tag, D2{}); };
        st = type == HK_F64   ? by_src(std::integral_constant<int, HK_F64>{})
             : type == HK_F32 ? by_src(std::integral_constant<int, HK_F32>{})
             : type == HK_BF16X2 ? by_src(std::integral_constant<int, HK_BF16X2>{})
             : type == HK_BF16X3 ? by_src(std::integral_constant<int, HK_BF16X3>{})
                                 : by_src(std::integral_constant<int, HK_F16X2>{});
    } else {
        st = type == HK_F64   ? launch(hk_spread_kernel<HK_F64>)
             : type == HK_F32 ? launch(hk_spread_kernel<HK_F32>)
             : type == HK_BF16X2 ? launch(hk_spread_kernel<HK_BF16X2>)
             : type == HK_BF16X3 ? launch(hk_spread_kernel<HK_BF16X3>)
                                 : launch(hk_spread_kernel<HK_F16X2>);
    }
    if (st != HK_OK) return st;
    HK_HIP(hipGetLastError());
    return HK_OK;
}

int hk_grid_bessel(int nrho, const double *rho, int nq, double h, int type, void *out, long long ld, void *stream)
{
    clear_status();
    if (nrho < 1 || !rho || nq < 1 || !out || !(h > 0) || type < HK_F64 || type > HK_F16X2 ||
        ld < operand_row(type, nq) || (type == HK_F16X2 && ld % 4 != 0))
        return fail(HK_ERR_ARGUMENT, "bad Bessel arguments");
    hipStream_t s = (hipStream_t)stream;
    const long long n = (long long)nrho * nq;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (type == HK_F64)
        hipLaunchKernelGGL(hk_bessel_kernel<HK_F64>, grid, dim3(256), 0, s, nrho, rho, nq, h, out, ld);
    else if (type == HK_F32)
        hipLaunchKernelGGL(hk_bessel_kernel<HK_F32>, grid, dim3(256), 0, s, nrho, rho, nq, h, out, ld);
    else if (type == HK_BF16X2)
        hipLaunchKernelGGL(hk_bessel_kernel<HK_BF16X2>, grid, dim3(256), 0, s, nrho, rho, nq, h, out, ld);
    else if (type == HK_BF16X3)
        hipLaunchKernelGGL(hk_bessel_kernel<HK_BF16X3>, grid, dim3(256), 0, s, nrho, rho, nq, h, out, ld);
    else
        hipLaunchKernelGGL(hk_bessel_kernel<HK_F16X2>, grid, dim3(256), 0, s, nrho, rho, nq, h, out, ld);
    HK_HIP(hipGetLastError());
    return HK_OK;
}

int hk_gemm(int kind, int m, int n, int k, const float *a, long long lda, const float *b, long long ldb, double *out,
            long long ldo, void *stream)
{
    clear_status();
    if (m < 1 || n < 1 || k < 64 || k % 64 != 0 || !a || !b || !out || lda < k || ldb < k || ldo < n || lda % 4 != 0 ||
        ldb % 4 != 0 || ((uintptr_t)a & 15) != 0 || ((uintptr_t)b & 15) != 0)
        return fail(HK_ERR_ARGUMENT, "GEMM needs k a multiple of 64, 16-byte aligned operands, strides multiple of 4");
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((n + 127) / 128), (unsigned)((m + 127) / 128));
    if (kind == HK_GEMM_F32)
        hipLaunchKernelGGL(hk_gemm_kernel<HK_GEMM_F32>, grid, dim3(512), 0, s, m, n, k, a, lda, b, ldb, out, ldo);
    else if (kind == HK_GEMM_BF16X3)
        hipLaunchKernelGGL(hk_gemm_kernel<HK_GEMM_BF16X3>, grid, dim3(512), 0, s, m, n, k, a, lda, b, ldb, out, ldo);
    else if (kind == HK_GEMM_BF16X6)
        hipLaunchKernelGGL(hk_gemm_kernel<HK_GEMM_BF16X6>, grid, dim3(512), 0, s, m, n, k, a, lda, b, ldb, out, ldo);
    else
        return fail(HK_ERR_ARGUMENT, "unknown GEMM kind");
    HK_HIP(hipGetLastError());
    return HK_OK;
}

int hk_gemm_split(int parts, int m, int n, int k, const void *a, long long lda, const void *b, long long ldb,
                  double *out, long long ldo, void *stream)
{
    clear_status();
    if ((parts != 2 && parts != 3) || m < 1 || n < 1 || k < 64 || k % 64 != 0 || !a || !b || !out ||
        lda < (long long)parts * k || ldb < (long long)parts * k || ldo < n || lda % 8 != 0 || ldb % 8 != 0 ||
        ((uintptr_t)a & 15) != 0 || ((uintptr_t)b & 15) != 0)
        return fail(HK_ERR_ARGUMENT, "split GEMM needs 2 or 3 parts, k a multiple of 64, strides >= parts x k and "
                                     "multiples of 8, 16-byte aligned operands");
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((n + 127) / 128), (unsigned)((m + 127) / 128));
    const size_t lds = hk_gemm_split_lds(parts);
    auto launch = [&](auto kernel) -> int {
        HK_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kernel, grid, dim3(512), lds, s, m, n, k, static_cast<const unsigned short *>(a), lda,
                           static_cast<const unsigned short *>(b), ldb, out, ldo);
        return HK_OK;
    };
    const int st = parts == 2 ? launch(hk_gemm_split_kernel<2>) : launch(hk_gemm_split_kernel<3>);
    if (st != HK_OK) return st;
    HK_HIP(hipGetLastError());
    return HK_OK;
}

int hk_gemm_f16x2(int m, int n, int k, const void *a, long long lda, const void *b, long long ldb, double *out,
                  long long ldo, void *stream)
{
    clear_status();
    if (m < 1 || n < 1 || k < 64 || k % 64 != 0 || !a || !b || !out || lda < 2LL * k + 4 || ldb < 2LL * k + 4 ||
        ldo < n || lda % 8 != 0 || ldb % 8 != 0 || ((uintptr_t)a & 15) != 0 || ((uintptr_t)b & 15) != 0)
        return fail(HK_ERR_ARGUMENT, "fp16 GEMM needs k a multiple of 64, strides >= 2 k + 4 and multiples of 8, "
                                     "16-byte aligned operands");
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((n + 127) / 128), (unsigned)((m + 127) / 128));
    const size_t lds = hk_gemm_split_lds(2);
    auto kernel = hk_gemm_f16x2_kernel;
    HK_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kernel, grid, dim3(512), lds, s, m, n, k, static_cast<const unsigned short *>(a), lda,
                       static_cast<const unsigned short *>(b), ldb, out, ldo);
    HK_HIP(hipGetLastError());
    return HK_OK;
}

} // extern "C"
