// Setup kernels of the TDSE path: ground state, step-independent factor tables,
// time grid and the spectra.  Built with -ffp-contract=off: the ground state's
// inverse iteration must reproduce the reference's CPU arithmetic exactly, because
// its stopping test |E_n - E_{n-1}| <= CV = 1e-25 only succeeds at a bitwise
// fixed point (with FMA contraction the iteration falls into a limit cycle).
#include "tdse_internal.h"

#pragma clang fp contract(off)

namespace {

__host__ __device__ inline double soft_coulomb(double x, double a) { return -1.0 / __builtin_sqrt(a * a + x * x); }

// tridiag.c:26-70, one lane: complex Thomas with the reference's operation order.
__host__ __device__ int thomas(const double *sub, const double *dia, const double *sup, const double *rhs, double *u, int n,
                      double *gam)
{
    int status = 0;
    double pr = dia[0], pi = dia[1];
    double pn = dia[0] * dia[0] + dia[1] * dia[1];
    if (dia[0] == 0. && dia[1] == 0.) status = 2;
    u[0] = (rhs[0] * dia[0] + rhs[1] * dia[1]) / (dia[0] * dia[0] + dia[1] * dia[1]);
    u[1] = (rhs[1] * dia[0] - rhs[0] * dia[1]) / (dia[0] * dia[0] + dia[1] * dia[1]);
    for (int j = 1; j <= n - 1; j++) {
        const double cr = sup[2 * (j - 1)], ci = sup[2 * (j - 1) + 1];
        const double ar = sub[2 * (j - 1)], ai = sub[2 * (j - 1) + 1];
        const double gr = (cr * pr + ci * pi) / pn;
        const double gi = (ci * pr - cr * pi) / pn;
        gam[2 * j] = gr;
        gam[2 * j + 1] = gi;
        pr = dia[2 * j] - (ar * gr - ai * gi);
        pi = dia[2 * j + 1] - (ai * gr + ar * gi);
        pn = pr * pr + pi * pi;
        if (pn == 0.) status = 2;
        const double ur = u[2 * (j - 1)], ui = u[2 * (j - 1) + 1];
        const double tre = rhs[2 * j] - ar * ur + ai * ui;
        const double tim = rhs[2 * j + 1] - ar * ui - ai * ur;
        double vr = tre * pr;
        vr += tim * pi;
        u[2 * j] = vr / pn;
        double vi = -tre * pi;
        vi += tim * pr;
        u[2 * j + 1] = vi / pn;
    }
    for (int j = n - 2; j >= 0; j--) {
        const double gr = gam[2 * (j + 1)], gi = gam[2 * (j + 1) + 1];
        const double nr = u[2 * (j + 1)], ni = u[2 * (j + 1) + 1];
        u[2 * j] -= gr * nr - gi * ni;
        u[2 * j + 1] -= gr * ni + gi * nr;
    }
    return status;
}

__host__ __device__ double sumsq(const double *v, int num_r)
{
    double s = 0.;
    for (int i = 0; i <= num_r; i++) s = s + v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1];
    return s;
}

__host__ __device__ void normalise(double *v, int num_r)
{
    const double nrm = __builtin_sqrt(sumsq(v, num_r));
    for (int i = 0; i <= num_r; i++) {
        v[2 * i] /= nrm;
        v[2 * i + 1] /= nrm;
    }
}

// <M2 psi, H psi> / <M2 psi, M2 psi>, tridiag.c:172-238; V holds soft_coulomb(x_j).
__host__ __device__ double energy(const double *psi, double c, const double *V, int n, double *h, double *m)
{
    h[0] = c * psi[0] - 0.5 * c * psi[2];
    h[0] = h[0] + ((10 / 12.) * psi[0] * (V[0]) + (1 / 12.) * psi[2] * (V[1]));
    h[1] = c * psi[1] - 0.5 * c * psi[3];
    h[1] = h[1] + ((10 / 12.) * psi[1] * (V[0]) + (1 / 12.) * psi[3] * (V[1]));
    m[0] = 10 * psi[0] / 12. + psi[2] / 12.;
    m[1] = 10 * psi[1] / 12. + psi[3] / 12.;
    for (int j = 1; j < n; j++) {
        const double vm = V[j - 1], v0 = V[j], vp = V[j + 1];
        h[2 * j] = c * psi[2 * j] - 0.5 * c * (psi[2 * (j - 1)] + psi[2 * (j + 1)]);
        h[2 * j] = h[2 * j] + ((10 / 12.) * psi[2 * j] * v0 + (1 / 12.) * psi[2 * (j - 1)] * vm + (1 / 12.) * psi[2 * (j + 1)] * vp);
        h[2 * j + 1] = c * psi[2 * j + 1] - 0.5 * c * (psi[2 * (j - 1) + 1] + psi[2 * (j + 1) + 1]);
        h[2 * j + 1] = h[2 * j + 1] + ((10 / 12.) * psi[2 * j + 1] * v0 + (1 / 12.) * psi[2 * (j - 1) + 1] * vm + (1 / 12.) * psi[2 * (j + 1) + 1] * vp);
        m[2 * j] = 10 * psi[2 * j] / 12. + (psi[2 * (j + 1)] + psi[2 * (j - 1)]) / 12.;
        m[2 * j + 1] = 10 * psi[2 * j + 1] / 12. + (psi[2 * (j + 1) + 1] + psi[2 * (j - 1) + 1]) / 12.;
    }
    const double vl = V[n - 1], vn = V[n];
    h[2 * n] = c * psi[2 * n] - 0.5 * c * psi[2 * (n - 1)];
    h[2 * n] = h[2 * n] + ((10 / 12.) * psi[2 * n] * vn + (1 / 12.) * psi[2 * (n - 1)] * vl);
    h[2 * n + 1] = c * psi[2 * n + 1] - 0.5 * c * psi[2 * (n - 1) + 1];
    h[2 * n + 1] = h[2 * n + 1] + ((10 / 12.) * psi[2 * n + 1] * vn + (1 / 12.) * psi[2 * (n - 1) + 1] * vl);
    m[2 * n] = 10 * psi[2 * n] / 12. + psi[2 * (n - 1)] / 12.;
    m[2 * n + 1] = 10 * psi[2 * n + 1] / 12. + psi[2 * (n - 1) + 1] / 12.;
    double e = 0, nrm = 0;
    for (int j = 0; j <= n; j++) {
        const double e1 = h[2 * j] * m[2 * j] + h[2 * j + 1] * m[2 * j + 1];
        const double e2 = -h[2 * j] * m[2 * j + 1] + h[2 * j + 1] * m[2 * j];
        e += e1 + e2;
        nrm += m[2 * j] * m[2 * j] + m[2 * j + 1] * m[2 * j + 1];
    }
    return e / nrm;
}

// Inverse iteration for the ground state of one job, tools.c:31-51 and tridiag.c:86-159: a
// serial chain, run by one GPU lane per job or on the host for one geometry alone.
__host__ __device__ void gs_iterate(GsJob &J)
{
    const int n = J.num_r;
    const long long sz = 2 * (long long)(n + 1);
    double *V = J.work;                 // n + 2
    double *sub = V + (n + 2);          // sz each
    double *dia = sub + sz;
    double *sup = dia + sz;
    double *res = sup + sz;
    double *w1 = res + sz;
    double *w2 = w1 + sz;
    double *gam = w2 + sz;              // sz + 2
    double *x = J.x, *psi0 = J.psi0;
    const double dx = J.dx, a = J.a, eg = J.e_guess;

    const double half_span = 0.5 * n * dx;
    for (int k = 0; k <= n; k++) x[k] = (double)k * dx - half_span;
    for (int k = 0; k <= n; k++) V[k] = soft_coulomb(x[k], a);
    const double dxl = x[n] - x[n - 1];
    V[n + 1] = soft_coulomb(x[n] + dxl, a);
    const double off = -0.5 / (dx * dx), kin = 1. / (dx * dx);
    for (int k = 0; k <= n; k++) {
        psi0[2 * k] = 1.0;
        psi0[2 * k + 1] = 0.;
    }
    for (int i = 0; i <= n; i++) {
        sub[2 * i] = off - eg / 12. + V[i] / 12.;
        sub[2 * i + 1] = 0.;
        dia[2 * i] = 10 * V[i] / 12. + kin - 10 * eg / 12.;
        dia[2 * i + 1] = 0.;
        sup[2 * i] = off - eg / 12. + V[i + 1] / 12.;
        sup[2 * i + 1] = 0.;
    }
    double e_old = eg, en = 0., test;
    int it = 0, status = 0;
    do {
        // M2 psi (tridiag.c:261-268), then the Thomas solve
        w1[0] = 10 * psi0[0] / 12. + psi0[2] / 12.;
        w1[1] = 10 * psi0[1] / 12. + psi0[3] / 12.;
        for (int i = 1; i < n; i++) {
            w1[2 * i] = 10 * psi0[2 * i] / 12. + (psi0[2 * (i + 1)] + psi0[2 * (i - 1)]) / 12.;
            w1[2 * i + 1] = 10 * psi0[2 * i + 1] / 12. + (psi0[2 * (i + 1) + 1] + psi0[2 * (i - 1) + 1]) / 12.;
        }
        w1[2 * n] = 10 * psi0[2 * n] / 12. + psi0[2 * (n - 1)] / 12.;
        w1[2 * n + 1] = 10 * psi0[2 * n + 1] / 12. + psi0[2 * (n - 1) + 1] / 12.;
        if (thomas(sub, dia, sup, w1, res, n + 1, gam) != 0) status = 2;
        normalise(res, n);
        for (long long i = 0; i < sz; i++) psi0[i] = res[i];
        en = energy(res, kin, V, n, w1, w2);
        test = __builtin_sqrt((en - e_old) * (en - e_old));
        e_old = en;
        ++it;
        if (it == J.max_iter) {
            status = 1;
            break;
        }
    } while (test > J.cv);
    normalise(psi0, n);
    J.energy = en;
    J.iters = it;
    J.status = status;
}

} // namespace

// One ground-state job per workgroup, run by its first lane (many geometries side by side,
// one per wave).
__global__ void tdse_ground_state_kernel(GsJob *jobs, int n_jobs)
{
    const int id = blockIdx.x;
    if (id >= n_jobs || threadIdx.x != 0) return;
    gs_iterate(jobs[id]);
}

// The same on the host (arrays of J in host memory, work of 16 (num_r + 2) doubles): the same
// operations in the same order without contraction, so the same bits as the GPU lane and
// the reference's CPU code.
void tdse_ground_state_host(GsJob &J) { gs_iterate(J); }

// Tables, part 1 (blocks along x, one job per blockIdx.y): g_j, V'(x_j) and psi0 in tiled order.
__global__ void tdse_tables_grid_kernel(TableJob *jobs)
{
    TableJob &J = jobs[blockIdx.y];
    const int nx = J.num_r + 1;
    const long long nchunks = (nx + TDSE_CHUNK - 1) / TDSE_CHUNK;
    const long long total = tdse_tiled_size(nchunks);
    const double gk = (-3.0 * J.dt) / (J.dx * J.dx);
    const double a2 = J.a * J.a;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        const long long tile = t / TDSE_TILE - 1, rem = t % TDSE_TILE;
        const long long c = tile * 64 + (rem % 64);
        const long long j = c * TDSE_CHUNK + rem / 64;
        if (tile >= 0 && j < nx) {
            const double x = J.x[j];
            const double r2 = a2 + x * x;
            J.g[t] = gk + (0.5 * J.dt) * soft_coulomb(x, J.a);
            J.dpot[t] = x * pow(r2, -1.5);  // tools.c:108-113 (gradpot)
            J.psi0[t] = make_double2(J.psi0_nat[2 * j], J.psi0_nat[2 * j + 1]);
        } else {
            J.g[t] = 0.;
            J.dpot[t] = 0.;
            J.psi0[t] = make_double2(0., 0.);
            J.ib[t] = make_double2(0., 0.);
        }
    }
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nchunks + 128; c += (long long)gridDim.x * blockDim.x) {
        J.chq[c] = make_double2(0., 0.);
        J.chg[c] = make_double2(0., 0.);
        J.adq[c] = make_double2(0., 0.);
        J.adg[c] = make_double2(0., 0.);
    }
}

namespace {

// Tables, part 2 (serial): pivots 1/beta_j of B' (tridiag.c:38-47 on the scaled matrix) in
// natural order into J.pivot_work, truncation decay rates, PopInt window (prop.c:86-93 with
// tools_algorithmic.c:19-43) and the psi0 support window.  J.x and J.psi0_nat are read.
__host__ __device__ void pivot_scan(TableJob &J)
{
    const int nx = J.num_r + 1;
    const double k36 = 36.0 * J.dt / (J.dx * J.dx);
    const double gk = (-3.0 * J.dt) / (J.dx * J.dx);
    double *piv = J.pivot_work;  // natural-order 1/beta (interleaved)
    auto gval = [&](int j) { return (j < 0 || j >= nx) ? 0.0 : gk + (0.5 * J.dt) * soft_coulomb(J.x[j], J.a); };
    double br = 10.0, bi = 10.0 * gval(0) + k36;
    double rho_f = 0.0, rho_b = 0.0;
    double gprev = gval(0), gcur = gval(1);
    for (int j = 0; j < nx; j++) {
        if (j > 0) {
            // gam = c'_{j-1} / beta_{j-1} with c'_{j-1} = 1 + i g_j ;  beta_j = b'_j - a'_j gam, a'_j = 1 + i g_{j-1}
            const double ibr = piv[2 * (j - 1)], ibi = piv[2 * (j - 1) + 1];
            const double cr = 1.0, ci = gcur;
            const double gr = cr * ibr - ci * ibi, gi = cr * ibi + ci * ibr;
            const double ar = 1.0, ai = gprev;
            br = 10.0 - (ar * gr - ai * gi);
            bi = (10.0 * gcur + k36) - (ai * gr + ar * gi);
            gprev = gcur;
            gcur = gval(j + 1);
        }
        const double nn = br * br + bi * bi;
        const double ir = br / nn, ii = -bi / nn;
        piv[2 * j] = ir;
        piv[2 * j + 1] = ii;
        const double inv_abs = __builtin_sqrt(ir * ir + ii * ii);
        if (j > 0) {  // |q_j| = |a'_j| |ib_j|
            const double gm = gval(j - 1);
            rho_f = fmax(rho_f, __builtin_sqrt(1.0 + gm * gm) * inv_abs);
        }
        if (j < nx - 1) {  // |h_j| = |c'_j| |ib_j|, and the adjoint's |p_j| = |1 + i g_j| |ib_j|
            const double gp = gval(j + 1), g0 = gval(j);
            rho_b = fmax(rho_b, __builtin_sqrt(1.0 + gp * gp) * inv_abs);
            rho_b = fmax(rho_b, __builtin_sqrt(1.0 + g0 * g0) * inv_abs);
        }
    }
    J.rho_fwd = rho_f;
    J.rho_bwd = rho_b;
    // PopInt window: findinterval(num_r, -x_int) and findinterval(num_r, +x_int)
    auto find_k1 = [&](double v) {
        if (v < J.x[0]) return -1;
        for (int i = 0; i < J.num_r; i++)
            if (v <= J.x[i + 1]) return i;
        return J.num_r;
    };
    const int k1 = find_k1(-J.x_int);
    const int k3 = find_k1(J.x_int);
    const int k4 = (k3 < 0) ? 0 : k3 + 1;
    J.pint_lo = k1;
    J.pint_hi = k4;
    double w = 0;
    if (k1 >= 0 && k4 <= J.num_r)
        for (int j = k1; j <= k4; j++) w = w + J.psi0_nat[2 * j] * J.psi0_nat[2 * j] + J.psi0_nat[2 * j + 1] * J.psi0_nat[2 * j + 1];
    J.pint0 = w;
    double mx = 0;
    for (int j = 0; j < nx; j++) mx = fmax(mx, J.psi0_nat[2 * j] * J.psi0_nat[2 * j] + J.psi0_nat[2 * j + 1] * J.psi0_nat[2 * j + 1]);
    // the rows of <psi0|psi> (PopTot): |psi0_j| above 1e-16 max|psi0|.  The rows left out add at
    // most 1e-16 max|psi0| sqrt(nx) |psi|_2 <= 4e-15 to <psi0|psi> (Cauchy-Schwarz; |psi|_2 = 1,
    // max|psi0| < 0.25, nx <= 32001), the rounding level of the sum itself.  Far from the atom
    // the inverse iteration leaves |psi0| on a plateau (near 5e-17 max after the 4 iterations of
    // CV = 1e-12, 5e-25 after the 6 of CV = 1e-25): a window that kept the plateau would cover
    // the whole grid and make every step read psi0 at every point.
    int lo = nx, hi = -1;
    for (int j = 0; j < nx; j++) {
        const double v = J.psi0_nat[2 * j] * J.psi0_nat[2 * j] + J.psi0_nat[2 * j + 1] * J.psi0_nat[2 * j + 1];
        if (v > 1e-32 * mx) {
            if (j < lo) lo = j;
            hi = j;
        }
    }
    J.p0_lo = lo;
    J.p0_hi = hi;
}

} // namespace

// Part 2 on the GPU: one job per workgroup (blockIdx.x), first lane.
__global__ void tdse_tables_pivot_kernel(TableJob *jobs)
{
    if (threadIdx.x != 0) return;
    pivot_scan(jobs[blockIdx.x]);
}

// Part 2 on the host (J.x, J.psi0_nat and J.pivot_work in host memory), the same bits.
void tdse_tables_pivot_host(TableJob &J) { pivot_scan(J); }

// Tables, part 2b (one thread per point, one job per blockIdx.y): the pivots in tiled order.
__global__ void tdse_tables_ib_kernel(TableJob *jobs)
{
    TableJob &J = jobs[blockIdx.y];
    const int nx = J.num_r + 1;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nx; j += gridDim.x * blockDim.x)
        J.ib[tdse_index(j)] = make_double2(J.pivot_work[2 * j], J.pivot_work[2 * j + 1]);
}

// Tables, part 3 (one thread per chunk, one job per blockIdx.y): transfer products of the
// two linear recurrences of the factorised solve over each chunk.
__global__ void tdse_tables_chunk_kernel(TableJob *jobs)
{
    TableJob &J = jobs[blockIdx.y];
    const int nx = J.num_r + 1;
    const int nchunks = (nx + TDSE_CHUNK - 1) / TDSE_CHUNK;
    const double gk = (-3.0 * J.dt) / (J.dx * J.dx);
    auto gval = [&](int j) { return (j < 0 || j >= nx) ? 0.0 : gk + (0.5 * J.dt) * soft_coulomb(J.x[j], J.a); };
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        auto ib_at = [&](int j) {
            return (j >= 0 && j < nx) ? make_double2(J.pivot_work[2 * j], J.pivot_work[2 * j + 1]) : make_double2(0., 0.);
        };
        auto mul = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
        // -(1 + i g) ib
        auto coef = [](double g, double2 ib) { return make_double2(-(ib.x - g * ib.y), -(ib.y + g * ib.x)); };
        double2 Q = make_double2(1., 0.), Hh = make_double2(1., 0.), AQ = make_double2(1., 0.), AG = make_double2(1., 0.);
        for (int i = 0; i < TDSE_CHUNK; i++) {
            const int j = c * TDSE_CHUNK + i;
            const double2 ib = ib_at(j);
            Q = mul(Q, coef(gval(j - 1), ib));       // q_j
            Hh = mul(Hh, coef(gval(j + 1), ib));     // h_j
            const double2 hm = coef(gval(j), ib_at(j - 1));  // h_{j-1} = -(1 + i g_j) ib_{j-1}
            AQ = mul(AQ, make_double2(hm.x, -hm.y));
            const double2 pj = coef(gval(j), ib);   // p_j = -(1 + i g_j) ib_j
            AG = mul(AG, make_double2(pj.x, -pj.y));
        }
        J.chq[c + 64] = Q;
        J.chg[c + 64] = Hh;
        J.adq[c + 64] = AQ;
        J.adg[c + 64] = AG;
    }
}

// One alignment of a shared lattice table set (TdseGeom::sg): shared tiled index t holds
// point t_point - shift of the widest geometry (g, dpot, ib tiled, nx points), zero outside
// [0, nx).  One thread per tiled index; npad: the shared arrays' length.
__global__ void tdse_lattice_kernel(const double *g, const double *dpot, const double2 *ib, int nx, int shift,
                                    long long npad, double *sg, double *sdp, double2 *sib)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npad;
         t += (long long)gridDim.x * blockDim.x) {
        const long long tile = t / TDSE_TILE - 1, rem = t % TDSE_TILE;
        const long long j = (tile * 64 + (rem % 64)) * TDSE_CHUNK + rem / 64 - shift;
        if (tile >= 0 && j >= 0 && j < nx) {
            const long long src = tdse_index(j);
            sg[t] = g[src];
            sdp[t] = dpot[src];
            sib[t] = ib[src];
        } else {
            sg[t] = 0.0;
            sdp[t] = 0.0;
            sib[t] = make_double2(0.0, 0.0);
        }
    }
}

// The chunk transfer maps of one lattice alignment (as tdse_tables_chunk_kernel, from the
// shared tiled g and 1/beta; rows outside the lattice are zero there): forward product of
// q_j = -(1 + i g_{j-1}) ib_j and backward product of h_j = -(1 + i g_{j+1}) ib_j over each
// chunk, chunk k at [k + 64].  One thread per chunk, nchunks chunks.
__global__ void tdse_lattice_chunk_kernel(const double *sg, const double2 *sib, int nchunks, double2 *schq,
                                          double2 *schg)
{
    auto mul = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    auto coef = [](double g, double2 ib) { return make_double2(-(ib.x - g * ib.y), -(ib.y + g * ib.x)); };
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        double2 Q = make_double2(1., 0.), Hh = make_double2(1., 0.);
        for (int i = 0; i < TDSE_CHUNK; i++) {
            const long long j = (long long)c * TDSE_CHUNK + i;
            const double2 ib = sib[tdse_index(j)];
            const double gm = j > 0 ? sg[tdse_index(j - 1)] : 0.0;  // point -1: the zero row before
            Q = mul(Q, coef(gm, ib));
            Hh = mul(Hh, coef(sg[tdse_index(j + 1)], ib));
        }
        schq[c + 64] = Q;
        schg[c + 64] = Hh;
    }
}

// x_k = k dx - num_r dx / 2 (tools.c:62-79), same rounding as the reference
__global__ void tdse_xgrid_kernel(double *x, int num_r, double dx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > num_r) return;
    const double half_span = 0.5 * num_r * dx;
    x[k] = (double)k * dx - half_span;
}

// A tiled state in natural order, interleaved re, im (a snapshot row of prop.c:248-263)
__global__ void tdse_untile_kernel(const double2 *psi, int nx, double *row)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nx) return;
    const double2 v = psi[tdse_index(j)];
    row[2 * j] = v.x;
    row[2 * j + 1] = v.y;
}

__global__ void tdse_fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// tgrid of prop.c:97,118,241: t accumulated one step at a time from tmin.
__global__ void tdse_tgrid_kernel(double *tgrid, int n, double tmin, double dt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tt = tmin;
    if (n > 0) tgrid[0] = tt;
    for (int k = 0; k + 1 < n; k++) {
        tt = tt + dt;
        tgrid[k + 1] = tt;
    }
}

// Real-input DFT X_k = sum_t s_t exp(-2 pi i k t / n), k < nbins: one thread per
// bin; the twiddle is re-anchored by sincospi every 32 samples and rotated in between.
__global__ void tdse_dft_r2c_kernel(const double *sig, int n, int nbins, double2 *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbins) return;
    double rs, rc;
    sincospi(-2.0 * (double)k / (double)n, &rs, &rc);
    double sr = 0.0, si = 0.0;
    for (int t0 = 0; t0 < n; t0 += 32) {
        const long long m = ((long long)k * t0) % n;
        double s, c;
        sincospi(-2.0 * (double)m / (double)n, &s, &c);
        const int t1 = min(n, t0 + 32);
        for (int t = t0; t < t1; t++) {
            const double v = sig[t];
            sr += v * c;
            si += v * s;
            const double cn = c * rc - s * rs;
            s = c * rs + s * rc;
            c = cn;
        }
    }
    out[k] = make_double2(sr, si);
}

// calcFFTW3 conventions (tools_fftw3.c:103-131): omega_k = k 2pi/tmax,
// F = conj(X) dt/sqrt(2pi), M2 = dt^2/(2pi) |X|^2.
__global__ void tdse_spectrum_kernel(const double2 *X, int nbins, double dt, double tmax, double *omega, double *F,
                                     double *M2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbins) return;
    const double pi = 3.14159265358979323846;
    const double dxi = 2. * pi / tmax;
    const double c1 = dt / __builtin_sqrt(2. * pi);
    const double c2 = dt * dt / (2. * pi);
    const double2 v = X[k];
    if (omega) omega[k] = ((double)k) * dxi;
    if (F) {
        F[2 * k] = c1 * v.x;
        F[2 * k + 1] = -c1 * v.y;
    }
    if (M2) M2[k] = c2 * (v.x * v.x + v.y * v.y);
}

// FourInterp's inverse half (tools_fftw3.c:46-63) for k1 >= 2: bins holds the first
// nbins = floor(n_in/2)+1 bins of the input's DFT; each is divided by n_in (the
// reference's normalisation) and, sitting below the Nyquist bin of length n_out,
// counts twice in the unnormalised c2r.
__global__ void tdse_c2r_interp_kernel(const double2 *bins, int nbins, int n_in, int n_out, double *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_out) return;
    double acc = 0.0;
    double rs, rc;
    sincospi(2.0 * (double)t / (double)n_out, &rs, &rc);
    for (int k0 = 1; k0 < nbins; k0 += 32) {
        const long long m = ((long long)k0 * t) % n_out;
        double s, c;
        sincospi(2.0 * (double)m / (double)n_out, &s, &c);
        const int k1 = min(nbins, k0 + 32);
        for (int k = k0; k < k1; k++) {
            const double2 b = bins[k];
            acc += b.x * c - b.y * s;
            const double cn = c * rc - s * rs;
            s = c * rs + s * rc;
            c = cn;
        }
    }
    out[t] = (bins[0].x + 2.0 * acc) / (double)n_in;
}
