// Pre-factor tables of the longitudinal integration (Hankel_transform.py:85-448): the
// phase of the XUV/IR dispersion and the XUV absorption accumulated along z, from the
// tabulated scattering factors f1, f2 (XUV_refractive_index.py:125-196).  One thread per
// frequency (per frequency and radius for the z-and-r pressure table).
#include "hankel_internal.h"

namespace {

// scipy interp1d(kind='linear') on an ascending table: slope form y_lo + slope (x - x_lo)
__device__ inline double lerp_at(const double *xt, const double *yt, int i, double x)
{
    const double slope = (yt[i + 1] - yt[i]) / (xt[i + 1] - xt[i]);
    return slope * (x - xt[i]) + yt[i];
}

// index i with xt[i] <= x <= xt[i+1] (x inside the table, n >= 2)
__device__ inline int bracket(const double *xt, int n, double x)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xt[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// the default interp1d: values outside the table are an error
__device__ inline bool interp_checked(const double *xt, const double *yt, int n, double x, double &y)
{
    if (!(x >= xt[0] && x <= xt[n - 1])) return false;
    if (n == 1) {
        y = yt[0];
        return true;
    }
    y = lerp_at(xt, yt, bracket(xt, n, x), x);
    return true;
}

// interp1d(bounds_error=False, fill_value=(y[0], y[-1]))
__device__ inline double interp_edge(const double *xt, const double *yt, int n, double x)
{
    if (n == 1 || x < xt[0]) return yt[0];
    if (x > xt[n - 1]) return yt[n - 1];
    return lerp_at(xt, yt, bracket(xt, n, x), x);
}

struct Walk {
    // cumulative trapezoids (cumulative_trapezoid(..., initial=0)) of the dispersion d(p_i)
    // and of the pressure p_i over the pressure z-table, evaluated on an ascending zgrid by
    // interpolation with edge fill; i advances monotonically with z.
    int i = 0;
    double cd_i = 0.0, cd_n = 0.0, cp_i = 0.0, cp_n = 0.0;
};

} // namespace

// One thread per frequency (per (frequency, radius) for the z-r table); the scalar branch
// spreads its planes over blockIdx.y, and the z-tabulated branch leaves the exponentials to
// hk_prefactor_finish_kernel (one thread per element) once the exit-plane value is known.
// staged (z-tabulated branch): the plane grid and the pressure table are first copied to
// dynamic LDS (nz + 2 np doubles), so the serial walk of each thread reads no global memory.
__global__ void hk_prefactor_kernel(HkConst K, hk_medium m, int nz, const double *zgrid, int no, const double *ogrid,
                                    int nr, const double *rgrid, double2 *pre, double *renorm, double *exit_abs,
                                    int *flag, int staged)
{
    extern __shared__ double stage[];
    if (staged) {
        for (int i = threadIdx.x; i < nz; i += blockDim.x) stage[i] = zgrid[i];
        for (int i = threadIdx.x; i < m.np_z; i += blockDim.x) {
            stage[nz + i] = m.p_z[i];
            stage[nz + m.np_z + i] = m.p_value[i];
        }
        __syncthreads();
        zgrid = stage;
        m.p_z = stage + nz;
        m.p_value = stage + nz + m.np_z;
    }
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const bool zr = m.mode == HK_PRESSURE_ZR;
    if (tid >= (zr ? no * nr : no)) return;
    const int o = zr ? tid / nr : tid;
    const int rj = zr ? tid % nr : 0;
    const double w = ogrid[o];
    // photon conversions, mynumerics.py:313,323,326
    const double w_au = w * K.timeau;
    const double energy = w_au / K.ev_per_au;
    const double lam = K.lam_num / (w_au * K.lam_den);
    double f1 = 0.0, f2 = 0.0;
    if (m.include_dispersion) {
        if (!interp_checked(m.e_f1, m.f1, m.n_f1, energy, f1)) {
            atomicMax(flag, HK_FLAG_RANGE);
            return;
        }
        if (f1 < -9998.0) {  // XUV_refractive_index.py:156
            atomicMax(flag, HK_FLAG_F1);
            return;
        }
    }
    if (m.include_absorption && !interp_checked(m.e_f2, m.f2, m.n_f2, energy, f2)) {
        atomicMax(flag, HK_FLAG_RANGE);
        return;
    }
    const double c = K.c_light;
    const double v_ir = c / m.n_ir;
    // (1/v_IR) - (1/v_XUV(p)), XUV_refractive_index.py:152-160
    auto disp = [&](double p) {
        const double n_xuv = 1.0 - p * K.n_ref_re * ((lam * lam) * f1 / (2.0 * M_PI));
        const double v_xuv = c / n_xuv;
        return (1.0 / v_ir) - (1.0 / v_xuv);
    };
    const double beta = K.n_ref_re * ((lam * lam) * f2 / (2.0 * M_PI));  // :191-195
    const double disp_vac = (m.n_ir - 1.0) / c;                           // Hankel_transform.py:188,434

    if (m.mode == HK_PRESSURE_SCALAR) {  // :405-448
        const double p = m.pressure;
        const double dfac = m.include_dispersion ? disp(p) : disp_vac;
        const double afac = m.include_absorption ? (p / c) * beta : 0.0;
        const double absw = w * afac;
        const double zl = zgrid[nz - 1];
        for (int k = blockIdx.y; k < nz; k += gridDim.y) {
            const double re = w * ((zgrid[k] - zl) * afac);
            const double im = w * (zgrid[k] * dfac);
            double s, cs;
            sincos(im, &s, &cs);
            const double e = exp(re);
            pre[(long long)k * no + o] = make_double2(p * (e * cs), p * (e * s));
            if (renorm) renorm[(long long)k * no + o] = exp((zl - zgrid[k]) * absw);
        }
        return;
    }

    // tabulated pressure: p_i on the table's z grid (interpolated in r for the z-r table)
    const int np = m.np_z;
    const double *zp = m.p_z;
    auto p_at = [&](int i) {
        if (!zr) return m.p_value[i];
        const double *row = m.p_value + (long long)i * m.np_r;
        return interp_edge(m.p_r, row, m.np_r, rgrid[rj]);
    };
    // pass over the table and the grid together
    Walk W;
    double d_i = m.include_dispersion ? disp(p_at(0)) : 0.0, p_i = p_at(0);
    double d_n = 0.0, p_n = p_i;
    if (np > 1) {
        p_n = p_at(1);
        d_n = m.include_dispersion ? disp(p_n) : 0.0;
        W.cd_n = ((zp[1] - zp[0]) * (d_i + d_n)) / 2.0;
        W.cp_n = ((zp[1] - zp[0]) * (p_i + p_n)) / 2.0;
    }
    double last_abs = 0.0;
    for (int k = 0; k < nz; k++) {
        const double z = zgrid[k];
        double cd, cp, pm;
        if (np == 1 || z < zp[0]) {
            cd = 0.0;
            cp = 0.0;
            pm = p_at(0);
        } else {
            while (W.i < np - 2 && zp[W.i + 1] < z) {  // advance to the interval holding z
                W.i++;
                W.cd_i = W.cd_n;
                W.cp_i = W.cp_n;
                d_i = d_n;
                p_i = p_n;
                p_n = p_at(W.i + 1);
                d_n = m.include_dispersion ? disp(p_n) : 0.0;
                W.cd_n = W.cd_i + ((zp[W.i + 1] - zp[W.i]) * (d_i + d_n)) / 2.0;
                W.cp_n = W.cp_i + ((zp[W.i + 1] - zp[W.i]) * (p_i + p_n)) / 2.0;
            }
            if (z > zp[np - 1]) {
                // edge fill with the table's last values: finish the walk
                while (W.i < np - 2) {
                    W.i++;
                    W.cd_i = W.cd_n;
                    W.cp_i = W.cp_n;
                    d_i = d_n;
                    p_i = p_n;
                    p_n = p_at(W.i + 1);
                    d_n = m.include_dispersion ? disp(p_n) : 0.0;
                    W.cd_n = W.cd_i + ((zp[W.i + 1] - zp[W.i]) * (d_i + d_n)) / 2.0;
                    W.cp_n = W.cp_i + ((zp[W.i + 1] - zp[W.i]) * (p_i + p_n)) / 2.0;
                }
                cd = W.cd_n;
                cp = W.cp_n;
                pm = p_n;
            } else {
                const double h = zp[W.i + 1] - zp[W.i];
                cd = ((W.cd_n - W.cd_i) / h) * (z - zp[W.i]) + W.cd_i;
                cp = ((W.cp_n - W.cp_i) / h) * (z - zp[W.i]) + W.cp_i;
                pm = ((p_n - p_i) / h) * (z - zp[W.i]) + p_i;
            }
        }
        const double dphase = m.include_dispersion ? cd : disp_vac;
        // the reference scales the cumulative pressure integral by beta before interpolating
        const double absf = m.include_absorption ? (1.0 / c) * (beta * cp) : 0.0;
        // stash (absorption, dispersion phase, pressure) until the exit-plane value is known
        const long long at = zr ? ((long long)k * no + o) * nr + rj : (long long)k * no + o;
        pre[at] = make_double2(absf, dphase);
        if (!zr && renorm) renorm[at] = pm;
        if (zr) {
            // z-r table: no renormalisation to the exit plane (:377-392)
            double s, cs;
            sincos(w * dphase, &s, &cs);
            const double e = exp(w * absf);
            pre[at] = make_double2(pm * (e * cs), pm * (e * s));
        }
        last_abs = absf;
    }
    if (!zr) exit_abs[o] = last_abs;
}

// The z-tabulated branch with one thread per (frequency, plane) instead of one per frequency:
// each thread repeats the table walk of hk_prefactor_kernel from the table's start up to its own
// plane (the walk's state at a plane depends on that plane's z only), so the cumulative
// trapezoids are formed in the same order and the stash holds the same bits; grid (ceil(no /
// 128), nz).  For tables of moderate length (the host picks it for np <= kWalkParallelMax).
__global__ void hk_prefactor_z_kernel(HkConst K, hk_medium m, int nz, const double *zgrid, int no, const double *ogrid,
                                      double2 *pre, double *renorm, double *exit_abs, int *flag)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (o >= no) return;
    const double w = ogrid[o];
    const double w_au = w * K.timeau;
    const double energy = w_au / K.ev_per_au;
    const double lam = K.lam_num / (w_au * K.lam_den);
    double f1 = 0.0, f2 = 0.0;
    if (m.include_dispersion) {
        if (!interp_checked(m.e_f1, m.f1, m.n_f1, energy, f1)) {
            if (k == 0) atomicMax(flag, HK_FLAG_RANGE);
            return;
        }
        if (f1 < -9998.0) {  // XUV_refractive_index.py:156
            if (k == 0) atomicMax(flag, HK_FLAG_F1);
            return;
        }
    }
    if (m.include_absorption && !interp_checked(m.e_f2, m.f2, m.n_f2, energy, f2)) {
        if (k == 0) atomicMax(flag, HK_FLAG_RANGE);
        return;
    }
    const double c = K.c_light;
    const double v_ir = c / m.n_ir;
    auto disp = [&](double p) {
        const double n_xuv = 1.0 - p * K.n_ref_re * ((lam * lam) * f1 / (2.0 * M_PI));
        const double v_xuv = c / n_xuv;
        return (1.0 / v_ir) - (1.0 / v_xuv);
    };
    const double beta = K.n_ref_re * ((lam * lam) * f2 / (2.0 * M_PI));
    const double disp_vac = (m.n_ir - 1.0) / c;
    const int np = m.np_z;
    const double *zp = m.p_z;
    Walk W;
    double d_i = m.include_dispersion ? disp(m.p_value[0]) : 0.0, p_i = m.p_value[0];
    double d_n = 0.0, p_n = p_i;
    if (np > 1) {
        p_n = m.p_value[1];
        d_n = m.include_dispersion ? disp(p_n) : 0.0;
        W.cd_n = ((zp[1] - zp[0]) * (d_i + d_n)) / 2.0;
        W.cp_n = ((zp[1] - zp[0]) * (p_i + p_n)) / 2.0;
    }
    auto advance = [&]() {
        W.i++;
        W.cd_i = W.cd_n;
        W.cp_i = W.cp_n;
        d_i = d_n;
        p_i = p_n;
        p_n = m.p_value[W.i + 1];
        d_n = m.include_dispersion ? disp(p_n) : 0.0;
        W.cd_n = W.cd_i + ((zp[W.i + 1] - zp[W.i]) * (d_i + d_n)) / 2.0;
        W.cp_n = W.cp_i + ((zp[W.i + 1] - zp[W.i]) * (p_i + p_n)) / 2.0;
    };
    const double z = zgrid[k];
    double cd, cp, pm;
    if (np == 1 || z < zp[0]) {
        cd = 0.0;
        cp = 0.0;
        pm = m.p_value[0];
    } else {
        while (W.i < np - 2 && zp[W.i + 1] < z) advance();
        if (z > zp[np - 1]) {
            while (W.i < np - 2) advance();
            cd = W.cd_n;
            cp = W.cp_n;
            pm = p_n;
        } else {
            const double h = zp[W.i + 1] - zp[W.i];
            cd = ((W.cd_n - W.cd_i) / h) * (z - zp[W.i]) + W.cd_i;
            cp = ((W.cp_n - W.cp_i) / h) * (z - zp[W.i]) + W.cp_i;
            pm = ((p_n - p_i) / h) * (z - zp[W.i]) + p_i;
        }
    }
    const double dphase = m.include_dispersion ? cd : disp_vac;
    const double absf = m.include_absorption ? (1.0 / c) * (beta * cp) : 0.0;
    const long long at = (long long)k * no + o;
    pre[at] = make_double2(absf, dphase);
    renorm[at] = pm;
    if (k == nz - 1) exit_abs[o] = absf;
}

// The z-tabulated branch's exponentials from the stash (absorption, dispersion phase) in pre
// and the pressure in renorm, renormalised to the exit plane (:215, :226-228).
__global__ void hk_prefactor_finish_kernel(int nz, int no, const double *ogrid, const double *exit_abs, double2 *pre,
                                           double *renorm)
{
    const long long at = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (long long)nz * no) return;
    const int o = (int)(at % no);
    const double w = ogrid[o], last_abs = exit_abs[o];
    const double2 v = pre[at];
    const double pm = renorm[at];
    double s, cs;
    sincos(w * v.y, &s, &cs);
    const double e = exp(w * (v.x - last_abs));
    pre[at] = make_double2(pm * (e * cs), pm * (e * s));
    renorm[at] = exp(w * last_abs - w * v.x);
}

__global__ void hk_fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
