/*
 * ref_driver.c -- glue linked against the reference's own, unmodified CTDSE
 * sources (1DTDSE/sources/{prop,tridiag,tools,tools_algorithmic,constants,
 * structures}.c) to produce golden vectors.  TEST INFRASTRUCTURE ONLY; built by
 * oracle/Makefile into oracle/_ref/, never shipped or loaded by the product.
 *
 * singleTDSE.c (call1DTDSE) needs FFTW3, which this image lacks, so this driver
 * repeats call1DTDSE's orchestration (1DTDSE/sources/singleTDSE.c:52-105) for
 * the k1 = 1 case only: there FourInterp (tools_fftw3.c:24-68) is the identity
 * up to FFT rounding, followed by the zero tail its calloc leaves.  Spectra are
 * not produced here (they are pinned separately, see tests/golden).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "structures.h"
#include "tools.h"
#include "prop.h"

/* Ground state through Initialise_grid_and_ground_state (tools.c:31). */
int ref_ground_state(int num_r, double dx, double a, double e_guess, double cv,
                     double *x_out, double *psi0_out, double *e_out)
{
    inputs_def in;
    memset(&in, 0, sizeof in);
    in.num_r = num_r;
    in.dx = dx;
    in.trg.a = a;
    in.Eguess = e_guess;
    in.CV = cv;
    Initialise_grid_and_ground_state(&in);
    memcpy(x_out, in.x, (size_t)(num_r + 1) * sizeof(double));
    memcpy(psi0_out, in.psi0, 2 * (size_t)(num_r + 1) * sizeof(double));
    *e_out = in.Einit;
    free(in.x);
    free(in.psi0);
    return 0;
}

/*
 * One call1DTDSE-equivalent run with k1 = 1.  Returns the number of exported
 * samples (outputs.Nt = N_E + 2) or -1 if the field grid would need k1 != 1.
 * Each output buffer must hold N_E + 2 doubles; psi_final 2*(num_r+1);
 * snaps (may be NULL) n_snap rows as prop.c sizes them.
 */
int ref_call(int num_r, double dx, double a, double x_int, double dt, const double *psi0, const double *x,
             const double *tgrid_in, const double *field_in, int n_e, int writewft, double tprint,
             double *tgrid, double *efield, double *sourceterm, double *pop_tot, double *pop_int,
             double *expval, double *snaps, int n_snap)
{
    inputs_def in;
    outputs_def out;
    memset(&in, 0, sizeof in);
    memset(&out, 0, sizeof out);
    in.num_r = num_r;
    in.dx = dx;
    in.trg.a = a;
    in.x_int = x_int;
    in.dt = dt;
    in.gauge = 0;
    in.analy.writewft = writewft;
    in.analy.tprint = tprint;
    in.psi0 = (double *)psi0;
    in.x = (double *)x;
    in.Efield.Nt = n_e;

    double field_dt = tgrid_in[1] - tgrid_in[0];
    int k1 = (int)floor(field_dt / dt);
    if (k1 != 1) return -1;
    in.Efield.dt = field_dt / ((double)k1);
    const int n2 = k1 * n_e;
    double *interp = calloc((size_t)n2 + 3, sizeof(double));
    memcpy(interp, field_in, (size_t)n_e * sizeof(double));
    const int nt_local = n2 + 1;
    out.tgrid = calloc((size_t)nt_local + 2, sizeof(double));
    out.Efield = calloc((size_t)nt_local + 2, sizeof(double));
    out.sourceterm = calloc((size_t)nt_local + 2, sizeof(double));
    out.PopTot = calloc((size_t)nt_local + 2, sizeof(double));
    out.PopInt = calloc((size_t)nt_local + 2, sizeof(double));
    out.expval = calloc((size_t)nt_local + 2, sizeof(double));
    out.Nt = nt_local + 1;

    double *psi = propagation(&in, &out, interp);

    const size_t nexp = (size_t)out.Nt;
    memcpy(tgrid, out.tgrid, nexp * sizeof(double));
    memcpy(efield, out.Efield, nexp * sizeof(double));
    memcpy(sourceterm, out.sourceterm, nexp * sizeof(double));
    memcpy(pop_tot, out.PopTot, nexp * sizeof(double));
    memcpy(pop_int, out.PopInt, nexp * sizeof(double));
    memcpy(expval, out.expval, nexp * sizeof(double));
    if (writewft && out.psi) {
        for (int i = 0; i < n_snap; i++) {
            if (snaps) memcpy(snaps + (size_t)i * 2 * (num_r + 1), out.psi[i], 2 * (size_t)(num_r + 1) * sizeof(double));
            free(out.psi[i]);
        }
        free(out.psi);
    }
    free(psi);
    free(interp);
    free(out.tgrid); free(out.Efield); free(out.sourceterm);
    free(out.PopTot); free(out.PopInt); free(out.expval);
    return (int)nexp;
}
