// Analysis entry points of the CTDSE library, on the GPU:
//   window_analysis  -- photoelectron spectrum by the resolvent (window) method,
//                       1DTDSE/sources/tools.c:206-316: two Numerov-Thomas solves per energy;
//   GaborTransform   -- Gabor (windowed Fourier) transform, tools_fftw3.c:158-213.
// Built with -ffp-contract=off like the other setup code (same operation order as the
// reference's tridiagonal solve).
#include "ctdse.h"
#include "tdse_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#pragma clang fp contract(off)

namespace {

// One lane per energy.  Scratch arrays are lane-interleaved (element j of lane l at
// j*L + l) so that the lanes of a wave access consecutive addresses.
struct PesJob {
    int num_r, n_lanes, e0;
    double dx, a, dE, Estep, E_start;
    const double *V;      // soft-Coulomb potential at x_j, j = 0..num_r+1
    const double *psi;    // 2 (num_r+1), the wavefunction analysed
    double *m2, *u, *u2, *gam;  // lane-interleaved scratch, 2 (num_r+2) per lane each
    double *pes;          // output, indexed by energy
    int num_E;
};

// Thomas solve (tridiag.c:26-70) with coefficients generated on the fly:
// sub/sup real parts (E' - V)/12 - off with E' = E -/+ dE, imaginary parts -/+ dE/12;
// diagonal 10 (E' - V)/12 - diag, imaginary -/+ 10 dE/12.
__device__ void pes_solve(const PesJob &J, double Ep, double im_off, double im_dia, const double *rhs, double *u,
                          int lane)
{
    const int L = J.n_lanes;
    const int n = J.num_r + 1;
    const double off = -0.5 / (J.dx * J.dx), dia_r = 1. / (J.dx * J.dx);
    auto A = [&](int j, double &re, double &im) { re = (Ep - J.V[j]) / 12. - off; im = im_off; };
    auto Cc = [&](int j, double &re, double &im) { re = (Ep - J.V[j + 1]) / 12. - off; im = im_off; };
    auto B = [&](int j, double &re, double &im) { re = 10 * (Ep - J.V[j]) / 12. - dia_r; im = im_dia; };
    auto at = [&](const double *p, int j, int k) { return p[(2LL * j + k) * L + lane]; };
    auto put = [&](double *p, int j, int k, double v) { p[(2LL * j + k) * L + lane] = v; };
    double b0r, b0i;
    B(0, b0r, b0i);
    double pr = b0r, pi = b0i, pn = b0r * b0r + b0i * b0i;
    put(u, 0, 0, (at(rhs, 0, 0) * b0r + at(rhs, 0, 1) * b0i) / (b0r * b0r + b0i * b0i));
    put(u, 0, 1, (at(rhs, 0, 1) * b0r - at(rhs, 0, 0) * b0i) / (b0r * b0r + b0i * b0i));
    for (int j = 1; j <= n - 1; j++) {
        double cr, ci, ar, ai, br, bi;
        Cc(j - 1, cr, ci);
        A(j - 1, ar, ai);
        B(j, br, bi);
        const double gr = (cr * pr + ci * pi) / pn;
        const double gi = (ci * pr - cr * pi) / pn;
        put(J.gam, j, 0, gr);
        put(J.gam, j, 1, gi);
        pr = br - (ar * gr - ai * gi);
        pi = bi - (ai * gr + ar * gi);
        pn = pr * pr + pi * pi;
        const double ur = at(u, j - 1, 0), ui = at(u, j - 1, 1);
        const double tre = at(rhs, j, 0) - ar * ur + ai * ui;
        const double tim = at(rhs, j, 1) - ar * ui - ai * ur;
        double vr = tre * pr;
        vr += tim * pi;
        put(u, j, 0, vr / pn);
        double vi = -tre * pi;
        vi += tim * pr;
        put(u, j, 1, vi / pn);
    }
    for (int j = n - 2; j >= 0; j--) {
        const double gr = at(J.gam, j + 1, 0), gi = at(J.gam, j + 1, 1);
        const double nr = at(u, j + 1, 0), ni = at(u, j + 1, 1);
        put(u, j, 0, at(u, j, 0) - (gr * nr - gi * ni));
        put(u, j, 1, at(u, j, 1) - (gr * ni + gi * nr));
    }
}

// M2 psi (tridiag.c:261-268); src is lane-interleaved when lane >= 0, else shared.
__device__ void pes_m2(const PesJob &J, const double *src, bool shared_src, double *dst, int lane)
{
    const int L = J.n_lanes, n = J.num_r;
    auto s = [&](int j, int k) { return shared_src ? src[2 * j + k] : src[(2LL * j + k) * L + lane]; };
    auto put = [&](int j, int k, double v) { dst[(2LL * j + k) * L + lane] = v; };
    for (int k = 0; k < 2; k++) {
        put(0, k, 10 * s(0, k) / 12. + s(1, k) / 12.);
        for (int i = 1; i < n; i++) put(i, k, 10 * s(i, k) / 12. + (s(i + 1, k) + s(i - 1, k)) / 12.);
        put(n, k, 10 * s(n, k) / 12. + s(n - 1, k) / 12.);
    }
}

__global__ void pes_kernel(PesJob J)
{
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = J.e0 + lane;
    if (lane >= J.n_lanes || e >= J.num_E) return;
    const double E = J.E_start + J.Estep * e;
    pes_m2(J, J.psi, true, J.m2, lane);
    pes_solve(J, E - J.dE, -J.dE / 12, -10 * J.dE / 12, J.m2, J.u, lane);
    pes_m2(J, J.u, false, J.m2, lane);
    pes_solve(J, E + J.dE, J.dE / 12, 10 * J.dE / 12, J.m2, J.u2, lane);
    double s = 0.;
    const int L = J.n_lanes;
    for (int i = 0; i <= J.num_r; i++) {
        const double re = J.u2[(2LL * i) * L + lane], im = J.u2[(2LL * i + 1) * L + lane];
        s = s + re * re + im * im;
    }
    J.pes[e] = s * J.dx * pow(J.dE, 4.);
}

__global__ void pes_potential_kernel(double *V, int num_r, double dx, double a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > num_r + 1) return;
    const double xmax = 0.5 * num_r * dx;
    const double x = (double)k * dx - xmax;  // k = num_r + 1 is the point the reference reads past x[num_r]
    V[k] = -1.0 / __builtin_sqrt(a * a + x * x);
}

// Gabor: windowed rows w_i[t] = exp(-(a (t dt - (t_min + i dt_G)))^2) s[t]
__global__ void gabor_window_kernel(const double *sig, int N, int N_G, double dt, double t_min, double dt_G,
                                    double inv_a, double *rows)
{
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= (long long)N * N_G) return;
    const int i = (int)(id / N), t = (int)(id % N);
    rows[id] = exp(-pow(inv_a * (((double)t) * dt - (t_min + ((double)i) * dt_G)), 2.)) * sig[t];
}

__global__ void gabor_dft_kernel(const double *rows, int N, int N_G, int N_freq, double norm, double *out)
{
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= (long long)N_G * N_freq) return;
    const int i = (int)(id / N_freq), k = (int)(id % N_freq);
    const double *r = rows + (long long)i * N;
    double rs, rc;
    sincospi(-2.0 * (double)k / (double)N, &rs, &rc);
    double sr = 0, si = 0;
    for (int t0 = 0; t0 < N; t0 += 32) {
        double s, c;
        sincospi(-2.0 * (double)(((long long)k * t0) % N) / (double)N, &s, &c);
        const int t1 = min(N, t0 + 32);
        for (int t = t0; t < t1; t++) {
            sr += r[t] * c;
            si += r[t] * s;
            const double cn = c * rc - s * rs;
            s = c * rs + s * rc;
            c = cn;
        }
    }
    out[id] = norm * sqrt(sr * sr + si * si);
}

} // namespace

extern "C" int ctdse_set_status_from_analysis(int code, const char *msg);

extern "C" double *window_analysis(inputs_def inputs, double *psi, int num_E, double dE, double Estep, double E_start)
{
    ctdse_set_status_from_analysis(CTDSE_OK, "");
    const int num_r = inputs.num_r;
    if (!psi || num_E <= 0 || num_r < 2 || !(inputs.dx > 0) || !(inputs.trg.a > 0)) {
        ctdse_set_status_from_analysis(CTDSE_ERR_ARGUMENT, "window_analysis: bad arguments");
        return nullptr;
    }
    const size_t nx = num_r + 1;
    const int lanes = std::min(num_E, 512);
    double *V = nullptr, *dpsi = nullptr, *scratch = nullptr, *pes = nullptr;
    const size_t per = 2 * (nx + 1) * (size_t)lanes;
    bool ok = hipMalloc(&V, (nx + 1) * sizeof(double)) == hipSuccess &&
              hipMalloc(&dpsi, 2 * nx * sizeof(double)) == hipSuccess &&
              hipMalloc(&scratch, 4 * per * sizeof(double)) == hipSuccess &&
              hipMalloc(&pes, num_E * sizeof(double)) == hipSuccess &&
              hipMemcpy(dpsi, psi, 2 * nx * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    double *res = nullptr;
    if (ok) {
        hipLaunchKernelGGL(pes_potential_kernel, dim3((nx + 256) / 256), dim3(256), 0, 0, V, num_r, inputs.dx, inputs.trg.a);
        PesJob J{};
        J.num_r = num_r;
        J.n_lanes = lanes;
        J.dx = inputs.dx;
        J.a = inputs.trg.a;
        J.dE = dE;
        J.Estep = Estep;
        J.E_start = E_start;
        J.V = V;
        J.psi = dpsi;
        J.m2 = scratch;
        J.u = scratch + per;
        J.u2 = scratch + 2 * per;
        J.gam = scratch + 3 * per;
        J.pes = pes;
        J.num_E = num_E;
        for (int e0 = 0; e0 < num_E; e0 += lanes) {
            J.e0 = e0;
            hipLaunchKernelGGL(pes_kernel, dim3((lanes + 63) / 64), dim3(64), 0, 0, J);
        }
        res = (double *)calloc(num_E, sizeof(double));
        ok = hipGetLastError() == hipSuccess && res &&
             hipMemcpy(res, pes, num_E * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (V) (void)hipFree(V);
    if (dpsi) (void)hipFree(dpsi);
    if (scratch) (void)hipFree(scratch);
    if (pes) (void)hipFree(pes);
    if (!ok) {
        free(res);
        ctdse_set_status_from_analysis(CTDSE_ERR_DEVICE, "window_analysis: device error");
        return nullptr;
    }
    return res;
}

extern "C" double **GaborTransform(double *signal, double dt, int N, int N_freq, int N_G, double t_min, double t_max,
                                   double a)
{
    ctdse_set_status_from_analysis(CTDSE_OK, "");
    if (!signal || N < 2 || N_freq < 1 || N_freq > N / 2 + 1 || N_G < 1 || !(a != 0.0)) {
        ctdse_set_status_from_analysis(CTDSE_ERR_ARGUMENT, "GaborTransform: bad arguments");
        return nullptr;
    }
    const double inv_a = 1.0 / a;
    const double dt_G = (t_max - t_min) / N_G;
    const double norm = dt / sqrt(2. * 3.14159265358979323846);
    double *dsig = nullptr, *rows = nullptr, *out = nullptr;
    const size_t nr = (size_t)N * N_G, no = (size_t)N_freq * N_G;
    bool ok = hipMalloc(&dsig, N * sizeof(double)) == hipSuccess && hipMalloc(&rows, nr * sizeof(double)) == hipSuccess &&
              hipMalloc(&out, no * sizeof(double)) == hipSuccess &&
              hipMemcpy(dsig, signal, N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    double **res = nullptr;
    if (ok) {
        hipLaunchKernelGGL(gabor_window_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, 0, dsig, N, N_G, dt,
                           t_min, dt_G, inv_a, rows);
        hipLaunchKernelGGL(gabor_dft_kernel, dim3((unsigned)((no + 127) / 128)), dim3(128), 0, 0, rows, N, N_G, N_freq,
                           norm, out);
        ok = hipGetLastError() == hipSuccess;
        if (ok) {
            res = (double **)calloc(N_G, sizeof(double *));
            for (int i = 0; i < N_G && ok; i++) {
                res[i] = (double *)calloc(N_freq, sizeof(double));
                ok = res[i] && hipMemcpy(res[i], out + (size_t)i * N_freq, N_freq * sizeof(double),
                                         hipMemcpyDeviceToHost) == hipSuccess;
            }
        }
    }
    if (dsig) (void)hipFree(dsig);
    if (rows) (void)hipFree(rows);
    if (out) (void)hipFree(out);
    if (!ok) {
        if (res) {
            for (int i = 0; i < N_G; i++) free(res[i]);
            free(res);
        }
        ctdse_set_status_from_analysis(CTDSE_ERR_DEVICE, "GaborTransform: device error");
        return nullptr;
    }
    return res;
}
