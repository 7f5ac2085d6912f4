/*
 * ctdse_oracle.c -- CPU FP64 restatement of the 1D-TDSE (CTDSE) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check
 * and the CPU-baseline leg of bench.py call it; the product (libsingleTDSE.so
 * under mma_amd/) never links it.
 *
 * Every routine restates one reference routine with the SAME floating-point
 * operation order (association, operand order, no FMA contraction: build with
 * -ffp-contract=off).  That is what makes the ground-state inverse iteration
 * reach its exact fixed point at CV = 1e-25 (the reference's own convergence
 * test compares energies to 1e-25, i.e. bit equality), and what lets the
 * oracle reproduce the compiled reference bit for bit (tests/golden).
 *
 * Complex arrays are interleaved (re, im) doubles, as in the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_GS_NOT_CONVERGED 1
#define ORC_ERR_ZERO_PIVOT 2
#define ORC_ERR_ALLOC 3
#define ORC_ERR_ARG 4

/* Soft-Coulomb potential, 1DTDSE/sources/tools.c:91-95 (potential). */
double orc_potential(double x, double a) { return -1.0 / sqrt(a * a + x * x); }

/* Its derivative V'(x) (the reference names it gradpot), tools.c:108-113. */
double orc_gradpot(double x, double a) { return x * pow(a * a + x * x, -1.5); }

/* Grid x_k = k*dx - num_r*dx/2, tools.c:62-79 (Initialise_grid_and_D2). */
void orc_grid(int num_r, double dx, double *x)
{
    double half_span = 0.5 * num_r * dx;
    for (int k = 0; k <= num_r; k++) x[k] = (double)k * dx - half_span;
}

/* Plain-sum squared norm, tools.c:122-129 (norme): sum = sum + re^2 + im^2. */
static double orc_sumsq(const double *v, int num_r)
{
    double s = 0.;
    for (int i = 0; i <= num_r; i++) s = s + v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1];
    return s;
}

/* tools.c:137-145 (normalise). */
static void orc_normalise(double *v, int num_r)
{
    double nrm = sqrt(orc_sumsq(v, num_r));
    for (int i = 0; i <= num_r; i++) { v[2 * i] /= nrm; v[2 * i + 1] /= nrm; }
}

/*
 * Complex Thomas solve, 1DTDSE/sources/tridiag.c:26-70.  Row j (j>=1) reads the
 * sub-diagonal entry sub[j-1] and row j-1 the super-diagonal entry sup[j-1];
 * gam is caller scratch of 2*(n+1) doubles.  Returns ORC_ERR_ZERO_PIVOT when a
 * pivot vanishes (the reference prints "Error 1/2" and carries on).
 */
int orc_thomas(const double *sub, const double *dia, const double *sup, const double *rhs,
               double *u, int n, double *gam)
{
    int status = ORC_OK;
    double pr = dia[0], pi = dia[1];
    double pn = dia[0] * dia[0] + dia[1] * dia[1];
    if (dia[0] == 0. && dia[1] == 0.) status = ORC_ERR_ZERO_PIVOT;
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
        if (pn == 0.) status = ORC_ERR_ZERO_PIVOT;
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

/* M2 = tridiag(1,10,1)/12 applied to psi, then the Thomas solve:
 * tridiag.c:251-274 (Inv_Tridiagonal_Matrix_complex_Numerov). */
static int orc_numerov_solve(const double *sub, const double *dia, const double *sup,
                             const double *psi, double *res, int num_r, double *m2psi, double *gam)
{
    m2psi[0] = 10 * psi[0] / 12. + psi[2] / 12.;
    m2psi[1] = 10 * psi[1] / 12. + psi[3] / 12.;
    for (int i = 1; i < num_r; i++) {
        m2psi[2 * i] = 10 * psi[2 * i] / 12. + (psi[2 * (i + 1)] + psi[2 * (i - 1)]) / 12.;
        m2psi[2 * i + 1] = 10 * psi[2 * i + 1] / 12. + (psi[2 * (i + 1) + 1] + psi[2 * (i - 1) + 1]) / 12.;
    }
    m2psi[2 * num_r] = 10 * psi[2 * num_r] / 12. + psi[2 * (num_r - 1)] / 12.;
    m2psi[2 * num_r + 1] = 10 * psi[2 * num_r + 1] / 12. + psi[2 * (num_r - 1) + 1] / 12.;
    return orc_thomas(sub, dia, sup, m2psi, res, num_r + 1, gam);
}

/* Energy <M2 psi, H psi>/<M2 psi, M2 psi>, tridiag.c:172-238
 * (E_calculation_numerov); inv_dx2 is the kinetic diagonal 1/dx^2. */
static double orc_energy(double a, const double *psi, double inv_dx2, const double *x, int num_r,
                         double *h, double *m)
{
    const double c = inv_dx2;
    const int n = num_r;
    h[0] = c * psi[0] - 0.5 * c * psi[2];
    h[0] = h[0] + ((10 / 12.) * psi[0] * (orc_potential(x[0], a)) + (1 / 12.) * psi[2] * (orc_potential(x[1], a)));
    h[1] = c * psi[1] - 0.5 * c * psi[3];
    h[1] = h[1] + ((10 / 12.) * psi[1] * (orc_potential(x[0], a)) + (1 / 12.) * psi[3] * (orc_potential(x[1], a)));
    m[0] = 10 * psi[0] / 12. + psi[2] / 12.;
    m[1] = 10 * psi[1] / 12. + psi[3] / 12.;
    for (int j = 1; j < n; j++) {
        const double vm = orc_potential(x[j - 1], a), v0 = orc_potential(x[j], a), vp = orc_potential(x[j + 1], a);
        h[2 * j] = c * psi[2 * j] - 0.5 * c * (psi[2 * (j - 1)] + psi[2 * (j + 1)]);
        h[2 * j] = h[2 * j] + ((10 / 12.) * psi[2 * j] * v0 + (1 / 12.) * psi[2 * (j - 1)] * vm + (1 / 12.) * psi[2 * (j + 1)] * vp);
        h[2 * j + 1] = c * psi[2 * j + 1] - 0.5 * c * (psi[2 * (j - 1) + 1] + psi[2 * (j + 1) + 1]);
        h[2 * j + 1] = h[2 * j + 1] + ((10 / 12.) * psi[2 * j + 1] * v0 + (1 / 12.) * psi[2 * (j - 1) + 1] * vm + (1 / 12.) * psi[2 * (j + 1) + 1] * vp);
        m[2 * j] = 10 * psi[2 * j] / 12. + (psi[2 * (j + 1)] + psi[2 * (j - 1)]) / 12.;
        m[2 * j + 1] = 10 * psi[2 * j + 1] / 12. + (psi[2 * (j + 1) + 1] + psi[2 * (j - 1) + 1]) / 12.;
    }
    const double vl = orc_potential(x[n - 1], a), vn = orc_potential(x[n], a);
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

/*
 * Ground state by inverse iteration, tools.c:31-51 (Initialise_grid_and_ground_state)
 * and tridiag.c:86-159 (Einitialise).  psi0 starts as all ones; each iteration
 * solves (T + M2 (V - Eguess)) psi' = M2 psi, normalises, and recomputes E; it
 * stops when |E - E_old| <= CV.  The reference exits the process after 1000
 * iterations; here max_iter bounds the loop and the status says what happened.
 * x (num_r+1) and psi0 (2*(num_r+1)) are written; *iters gets the count.
 */
int orc_ground_state(int num_r, double dx, double a, double e_guess, double cv, int max_iter,
                     double *x, double *psi0, double *e_out, int *iters)
{
    if (num_r < 2 || dx <= 0.) return ORC_ERR_ARG;
    const size_t sz = 2 * (size_t)(num_r + 1);
    double *sub = calloc(sz, sizeof(double)), *dia = calloc(sz, sizeof(double)), *sup = calloc(sz, sizeof(double));
    double *res = calloc(sz, sizeof(double)), *w1 = calloc(sz, sizeof(double)), *w2 = calloc(sz, sizeof(double));
    double *gam = calloc(sz + 2, sizeof(double));
    if (!sub || !dia || !sup || !res || !w1 || !w2 || !gam) {
        free(sub); free(dia); free(sup); free(res); free(w1); free(w2); free(gam);
        return ORC_ERR_ALLOC;
    }
    orc_grid(num_r, dx, x);
    const double off = -0.5 / (dx * dx), kin = 1. / (dx * dx);
    for (int k = 0; k <= num_r; k++) { psi0[2 * k] = 1.0; psi0[2 * k + 1] = 0.; }
    for (int i = 0; i <= num_r; i++) {
        sub[2 * i] = off - e_guess / 12. + orc_potential(x[i], a) / 12.;
        sub[2 * i + 1] = 0.;
        dia[2 * i] = 10 * orc_potential(x[i], a) / 12. + kin - 10 * e_guess / 12.;
        dia[2 * i + 1] = 0.;
    }
    for (int i = 0; i < num_r; i++) {
        sup[2 * i] = off - e_guess / 12. + orc_potential(x[i + 1], a) / 12.;
        sup[2 * i + 1] = 0.;
    }
    {
        const double dxl = x[num_r] - x[num_r - 1];
        sup[2 * num_r] = off - e_guess / 12. + orc_potential(x[num_r] + dxl, a) / 12.;
        sup[2 * num_r + 1] = 0.;
    }
    double e_old = e_guess, energy = 0., test;
    int it = 0, status = ORC_OK;
    do {
        if (orc_numerov_solve(sub, dia, sup, psi0, res, num_r, w1, gam) != ORC_OK) status = ORC_ERR_ZERO_PIVOT;
        orc_normalise(res, num_r);
        memcpy(psi0, res, sz * sizeof(double));
        energy = orc_energy(a, res, kin, x, num_r, w1, w2);
        test = sqrt((energy - e_old) * (energy - e_old));
        e_old = energy;
        ++it;
        if (it == max_iter) { status = ORC_ERR_GS_NOT_CONVERGED; break; }
    } while (test > cv);
    orc_normalise(psi0, num_r);
    *e_out = energy;
    *iters = it;
    free(sub); free(dia); free(sup); free(res); free(w1); free(w2); free(gam);
    return status;
}

/* Interval search, 1DTDSE/sources/tools_algorithmic.c:19-43 (findinterval). */
void orc_findinterval(int n, double v, const double *grid, int *k1, int *k2)
{
    if (v < grid[0]) { *k1 = -1; *k2 = 0; return; }
    for (int i = 0; i < n; i++)
        if (v <= grid[i + 1]) { *k1 = i; *k2 = i + 1; return; }
    *k1 = n;
    *k2 = n + 1;
}

/*
 * Time propagation, 1DTDSE/sources/prop.c:24-280 (propagation, length gauge)
 * with the observables of prop.c:293-347 (compute_expectation_values).
 *
 * in_field must hold n_steps+1 samples (the reference reads in_field[k+1]).
 * Outputs (each n_steps+1 long, index 0 = initial state): tgrid, efield,
 * sourceterm, pop_tot, pop_int, expval.  psi_out (2*(num_r+1)) receives the
 * final state.  Snapshots follow prop.c:59-64 and :248-263 when snaps != NULL:
 * n_snap = n_steps / floor(tprint/dt) rows of 2*(num_r+1) doubles.
 */
int orc_propagate(int num_r, double dx, double a, double x_int, double dt, double tmin,
                  const double *psi0, const double *x, const double *in_field, int n_steps,
                  double *tgrid, double *efield, double *sourceterm, double *pop_tot, double *pop_int,
                  double *expval, double *psi_out, double tprint, double *snaps, int n_snap)
{
    const int n = num_r;
    const size_t sz = 2 * (size_t)(n + 1);
    double *psi = psi_out;
    double *rhs = calloc(sz, sizeof(double)), *res = calloc(sz, sizeof(double));
    double *sub = calloc(sz, sizeof(double)), *dia = calloc(sz, sizeof(double)), *sup = calloc(sz, sizeof(double));
    double *gam = calloc(sz + 2, sizeof(double));
    double *pot = calloc((size_t)n + 2, sizeof(double)), *dpot = calloc((size_t)n + 1, sizeof(double));
    if (!rhs || !res || !sub || !dia || !sup || !gam || !pot || !dpot) {
        free(rhs); free(res); free(sub); free(dia); free(sup); free(gam); free(pot); free(dpot);
        return ORC_ERR_ALLOC;
    }
    int status = ORC_OK;
    /* potential(x_j) at the grid points plus V(x_n + dx) for the unused last super entry */
    for (int j = 0; j <= n; j++) { pot[j] = orc_potential(x[j], a); dpot[j] = orc_gradpot(x[j], a); }
    pot[n + 1] = orc_potential(x[n] + dx, a);

    int k1 = 0, k2 = 0, k3 = 0, k4 = 0;
    orc_findinterval(n, -x_int, x, &k1, &k2);
    orc_findinterval(n, x_int, x, &k3, &k4);
    double win = 0;
    for (int j = k1; j <= k4; j++) win = win + psi0[2 * j] * psi0[2 * j] + psi0[2 * j + 1] * psi0[2 * j + 1];

    double tt = tmin, field = 0.;
    const double cpot = 1.;
    tgrid[0] = tt; sourceterm[0] = 0.; efield[0] = field; pop_int[0] = win; pop_tot[0] = 1.0; expval[0] = 0.0;
    memcpy(psi, psi0, sz * sizeof(double));

    const int steps_per_snap = (snaps && n_snap > 0) ? (int)floor(tprint / dt) : 0;
    int i_snap = 0;
    int gating = 1;

    /* The tridiagonal of M2 + i(dt/2)H does not depend on the step; the
     * reference rebuilds it every step (prop.c:138-159) with the same values. */
    for (int j = 0; j < n; j++) {
        sub[2 * j] = 1 / 12.;
        sub[2 * j + 1] = 0.5 * dt * (-0.5 / (dx * dx)) + 0.5 * dt * 1 / 12. * (cpot * pot[j]);
        dia[2 * j] = 10 / 12.;
        dia[2 * j + 1] = 0.5 * dt * (1. / (dx * dx)) + 0.5 * dt * 10 / 12. * (cpot * pot[j]);
        sup[2 * j] = 1 / 12.;
        sup[2 * j + 1] = 0.5 * dt * (-0.5 / (dx * dx)) + 0.5 * dt * 1 / 12. * (cpot * pot[j + 1]);
    }
    sub[2 * n] = 1 / 12.;
    sub[2 * n + 1] = 0.5 * dt * (-0.5 / (dx * dx)) + 0.5 * dt * 1 / 12. * (cpot * pot[n]);
    dia[2 * n] = 10 / 12.;
    dia[2 * n + 1] = 0.5 * dt * (1. / (dx * dx)) + 0.5 * dt * 10 / 12. * (cpot * pot[n]);
    sup[2 * n] = 1 / 12.;
    sup[2 * n + 1] = 0.5 * dt * (-0.5 / (dx * dx)) + 0.5 * dt * 1 / 12. * (cpot * pot[n + 1]);

    for (int k = 0; k < n_steps; k++) {
        tt = tt + dt;
        const double coef = 0.5 * dt / (dx * dx);
        if (gating) {
            if (in_field[k] * in_field[k + 1] <= 0.0) gating = 0;
            field = 0.;
        } else {
            field = in_field[k];
        }
        /* right-hand side (M2 - i dt/2 H) psi, prop.c:161-190 */
        rhs[0] = (10 / 12.) * psi[0] + coef * psi[1] + 1 / 12. * psi[2] - 0.5 * coef * psi[3];
        rhs[0] = rhs[0] + 0.5 * dt * ((10 / 12.) * psi[1] * (cpot * pot[0]) + (1 / 12.) * psi[3] * (cpot * pot[1]));
        rhs[1] = (10 / 12.) * psi[1] - coef * psi[0] + 1 / 12. * psi[3] + 0.5 * coef * psi[2];
        rhs[1] = rhs[1] - 0.5 * dt * ((10 / 12.) * psi[0] * (cpot * pot[0]) + (1 / 12.) * psi[2] * (cpot * pot[1]));
        for (int j = 1; j < n; j++) {
            const double *p = psi + 2 * j;
            rhs[2 * j] = (10 / 12.) * p[0] + coef * p[1] + 1 / 12. * p[2] + 1 / 12. * p[-2] - 0.5 * coef * (p[-1] + p[3]);
            rhs[2 * j] = rhs[2 * j] + 0.5 * dt * ((10 / 12.) * p[1] * (cpot * pot[j]) + (1 / 12.) * p[-1] * (cpot * pot[j - 1]) + (1 / 12.) * p[3] * (cpot * pot[j + 1]));
            rhs[2 * j + 1] = (10 / 12.) * p[1] - coef * p[0] + 1 / 12. * p[3] + 1 / 12. * p[-1] + 0.5 * coef * (p[-2] + p[2]);
            rhs[2 * j + 1] = rhs[2 * j + 1] - 0.5 * dt * ((10 / 12.) * p[0] * (cpot * pot[j]) + (1 / 12.) * p[-2] * (cpot * pot[j - 1]) + (1 / 12.) * p[2] * (cpot * pot[j + 1]));
        }
        rhs[2 * n] = (10 / 12.) * psi[2 * n] + coef * psi[2 * n + 1] + 1 / 12. * psi[2 * (n - 1)] - 0.5 * coef * psi[2 * (n - 1) + 1];
        rhs[2 * n] = rhs[2 * n] + 0.5 * dt * ((10 / 12.) * psi[2 * n + 1] * (cpot * pot[n]) + (1 / 12.) * psi[2 * (n - 1) + 1] * (cpot * pot[n - 1]));
        rhs[2 * n + 1] = (10 / 12.) * psi[2 * n + 1] - coef * psi[2 * n] + 1 / 12. * psi[2 * (n - 1) + 1] + 0.5 * coef * psi[2 * (n - 1)];
        rhs[2 * n + 1] = rhs[2 * n + 1] - 0.5 * dt * ((10 / 12.) * psi[2 * n] * (cpot * pot[n]) + (1 / 12.) * psi[2 * (n - 1)] * (cpot * pot[n - 1]));

        if (orc_thomas(sub, dia, sup, rhs, res, n + 1, gam) != ORC_OK) status = ORC_ERR_ZERO_PIVOT;

        /* length-gauge phase exp(+i F dt x), prop.c:199-204 */
        for (int j = 0; j <= n; j++) {
            const double c = cos(field * dt * x[j]), s = sin(field * dt * x[j]);
            psi[2 * j] = c * res[2 * j] - s * res[2 * j + 1];
            psi[2 * j + 1] = c * res[2 * j + 1] + s * res[2 * j];
        }
        tgrid[k + 1] = tt;
        efield[k + 1] = field;

        /* observables, prop.c:302-346 */
        double pre = 0., pim = 0.;
        for (int j = 0; j <= n; j++) {
            pre = pre + psi[2 * j] * psi0[2 * j] + psi[2 * j + 1] * psi0[2 * j + 1];
            pim = pim + psi[2 * j] * psi0[2 * j + 1] - psi[2 * j + 1] * psi0[2 * j];
        }
        double w = 0;
        for (int j = k1; j <= k4; j++) w = w + psi[2 * j] * psi[2 * j] + psi[2 * j + 1] * psi[2 * j + 1];
        double pos = 0.;
        for (int j = 0; j <= n; j++) pos = pos + (psi[2 * j] * psi[2 * j] + psi[2 * j + 1] * psi[2 * j + 1]) * x[j];
        double gp = 0.;
        for (int j = 0; j <= n; j++) gp = gp + (psi[2 * j] * psi[2 * j] + psi[2 * j + 1] * psi[2 * j + 1]) * dpot[j];
        gp = -gp + efield[k + 1];
        sourceterm[k + 1] = gp;
        pop_tot[k + 1] = pre * pre + pim * pim;
        expval[k + 1] = pos;
        pop_int[k + 1] = w;

        if (steps_per_snap > 0) {
            if (k % steps_per_snap == 0 && i_snap < n_snap - 1) {
                memcpy(snaps + (size_t)i_snap * sz, psi, sz * sizeof(double));
                i_snap++;
            }
            if (k == n_steps - 1 && i_snap == n_snap - 1) memcpy(snaps + (size_t)i_snap * sz, psi, sz * sizeof(double));
        }
    }
    free(rhs); free(res); free(sub); free(dia); free(sup); free(gam); free(pot); free(dpot);
    return status;
}

/*
 * Batch driver used as the CPU baseline: n_inst independent instances sharing one
 * grid and ground state, each driven by its own field row (field_stride doubles
 * apart, n_steps+1 samples used).  One instance per thread (OpenMP), n_threads
 * threads.  Outputs: sourceterm rows of n_steps+1 (others are computed and
 * dropped).  Returns the first non-OK status.
 */
int orc_propagate_batch(int n_inst, int num_r, double dx, double a, double x_int, double dt,
                        const double *psi0, const double *x, const double *fields, int field_stride,
                        int n_steps, double *sourceterm_out, int n_threads)
{
    int status = ORC_OK;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < n_inst; i++) {
        const size_t len = (size_t)n_steps + 1;
        double *buf = malloc(5 * len * sizeof(double));
        double *psi = malloc(2 * (size_t)(num_r + 1) * sizeof(double));
        if (!buf || !psi) {
            free(buf); free(psi);
#ifdef _OPENMP
#pragma omp atomic write
#endif
            status = ORC_ERR_ALLOC;
            continue;
        }
        int st = orc_propagate(num_r, dx, a, x_int, dt, 0., psi0, x, fields + (size_t)i * field_stride, n_steps,
                               buf, buf + len, sourceterm_out + (size_t)i * len, buf + 2 * len, buf + 3 * len,
                               buf + 4 * len, psi, 0., NULL, 0);
        if (st != ORC_OK) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            status = st;
        }
        free(buf); free(psi);
    }
    return status;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
