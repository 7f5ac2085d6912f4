/*
 * ctdse.h -- C ABI of libsingleTDSE.so, the MI355X implementation of the CTDSE
 * single-atom solver.
 *
 * Part 1 repeats the reference library's interface so that existing callers
 * (1DTDSE/PythonTDSE.py, 1DTDSE/PythonCTDSE/structures.py, 1DTDSE/tests.py)
 * load this library in place of the CPU one: the same symbol names and
 * signatures, and structs laid out byte-for-byte like 1DTDSE/sources/structures.h
 * (x86-64 natural alignment).  Ownership is unchanged: the library allocates
 * every output array with malloc/calloc and the caller frees them through the
 * destructors.
 *
 * Part 2 adds entry points the reference does not have: an error status in
 * place of the reference's exit() calls, and a batched, device-resident
 * interface (many instances in one persistent launch).
 *
 * Complex arrays are interleaved (re, im) doubles, as in the reference.
 */
#ifndef MMA_AMD_CTDSE_H
#define MMA_AMD_CTDSE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Part 1: reference interface ---------------------------------------- */

/* 1DTDSE/sources/structures.h:16-33 */
typedef struct Efield_var {
    double *tgrid;
    double *Field;
    double dt;
    double omega;
    double E0;
    double phi;
    double ton;
    double toff;
    int Nt;
    int nc;
} Efield_var;

/* structures.h:39-42 */
typedef struct trg_def {
    double a;
} trg_def;

/* structures.h:48-53 */
typedef struct analy_def {
    double tprint;
    int writewft;
} analy_def;

/* structures.h:59-71 */
typedef struct output_print_def {
    int Efield;
    int FEfield;
    int sourceterm;
    int Fsourceterm;
    int FEfieldM2;
    int FsourceTermM2;
    int PopTot;
    int tgrid;
    int omegagrid;
    int PopInt;
    int expval_x;
} output_print_def;

/* structures.h:80-118 */
typedef struct inputs_def {
    struct trg_def trg;
    struct Efield_var Efield;
    double Eguess;
    double Einit;
    double tmin;
    int Nt;
    int num_t;
    double dt;
    int num_r;
    double dx;
    double *psi0;
    double *x;
    int gauge;
    double x_int;
    struct analy_def analy;
    int InterpByDTorNT;
    int Ntinterp;
    struct output_print_def Print;
    double CV;
    char precision[2];
} inputs_def;

/* structures.h:120-147 */
typedef struct outputs_def {
    double *tgrid;
    double *Efield;
    double *sourceterm;
    double *omegagrid;
    double *FEfield;
    double *Fsourceterm;
    double *FEfieldM2;
    double *FsourcetermM2;
    double *PopTot;
    double *PopInt;
    double *expval;
    int Nt;
    int Nomega;
    double **psi;
} outputs_def;

/* Replaces 1DTDSE/sources/tools.c:31 (Initialise_grid_and_ground_state): allocates
 * in->x (num_r+1) and in->psi0 (2*(num_r+1)), sets in->Einit.  Ground state by
 * inverse iteration on the GPU, bit-identical to the reference's CPU result. */
void Initialise_grid_and_ground_state(inputs_def *in);

/* Replaces 1DTDSE/sources/singleTDSE.c:28 (call1DTDSE): refines the field, propagates
 * on the GPU, computes spectra, fills *outputs (library-owned arrays). Mutates
 * inputs->Efield.dt, ->Ntinterp and ->num_t exactly as the reference does. */
void call1DTDSE(inputs_def *inputs, outputs_def *outputs);

/* Replaces tools.c:423 (set_time_and_field): copies N samples into calloc'd arrays. */
void set_time_and_field(inputs_def *in, double *time, double *field, int N);

/* Replaces structures.c:92 (Set_all_prints); returned by value. */
output_print_def Set_all_prints(void);

/* Replaces structures.c:116 (Initialise_Printing_struct). */
output_print_def Initialise_Printing_struct(void);

/* Replaces tools.c:206 (window_analysis): photoelectron spectrum by the resolvent
 * method, num_E energies; returns a malloc'd array of num_E doubles. */
double *window_analysis(inputs_def inputs, double *psi, int num_E, double dE, double Estep, double E_start);

/* Replaces 1DTDSE/sources/tools_fftw3.c:158 (GaborTransform): returns N_G malloc'd
 * rows of N_freq doubles. */
double **GaborTransform(double *signal, double dt, int N, int N_freq, int N_G, double t_min, double t_max,
                        double a);

/* Replace tools.c:408 (free_arr), structures.c:76 (free_mtrx),
 * structures.c:15-49 (outputs_destructor) and structures.c:55-68 (inputs_destructor). */
void free_arr(double *buffer);
void free_mtrx(double **buf, int N_rows);
void outputs_destructor(outputs_def *outputs);
void inputs_destructor(inputs_def *in);

/* ---- Part 2: additions (no counterpart in the reference) ----------------- */

/* Status of the calling thread's last entry-point call.  The reference prints and
 * calls exit() (tridiag.c:139-142, constants.c:74-75) or prints and carries on
 * (tridiag.c:33,49); here every such condition sets a code and the void entry
 * points return early, leaving the process alive. */
enum ctdse_status {
    CTDSE_OK = 0,
    CTDSE_ERR_GS_NOT_CONVERGED = 1, /* 1000 inverse iterations without |dE| <= CV */
    CTDSE_ERR_ZERO_PIVOT = 2,       /* a Thomas pivot vanished (tridiag.c:33,49)  */
    CTDSE_ERR_ARGUMENT = 3,         /* NULL or short arrays, num_r < 2, tprint < dt, x_int outside the box */
    CTDSE_ERR_UNSUPPORTED = 4,      /* velocity gauge (dead code in the reference), k1 <= 0 */
    CTDSE_ERR_DEVICE = 5,           /* HIP runtime error, no GPU */
    CTDSE_ERR_NONFINITE = 6         /* NaN/Inf in the field or the result */
};
int ctdse_last_error(void);
const char *ctdse_last_error_message(void);

/* Opaque per-geometry state on the device: grid, ground state and the
 * step-independent Crank-Nicolson factor tables for one (num_r, dx, dt, a, x_int). */
typedef struct ctdse_geometry ctdse_geometry;

/* Builds a geometry.  psi0 == NULL: ground state by inverse iteration from
 * (e_guess, cv), as Initialise_grid_and_ground_state; otherwise psi0 (host,
 * 2*(num_r+1) interleaved doubles) is used as given.  Returns a ctdse_status. */
int ctdse_geometry_create(int num_r, double dx, double dt, double a, double e_guess, double cv, double x_int,
                          const double *psi0, ctdse_geometry **out);

/* Ground-state energy, inverse iterations used (0 if psi0 was given) and the
 * truncation depth (grid points on each side of a segment) of the segmented solve. */
int ctdse_geometry_info(const ctdse_geometry *g, double *e_init, int *gs_iterations, int *halo_points);

/* Copies x (num_r+1) and psi0 (2*(num_r+1)) to host memory. */
int ctdse_geometry_ground_state(const ctdse_geometry *g, double *x, double *psi0);

void ctdse_geometry_destroy(ctdse_geometry *g);

/*
 * Batched propagation: n_inst independent instances in one persistent launch on
 * `stream` (a hipStream_t, NULL = default stream).  Every array below is DEVICE
 * memory.  Instance i uses geometry geom[i] (host array of pointers) and runs
 * n_steps[i] Crank-Nicolson steps (the reference's outputs.Nt).
 *   field + field_off[i]: n_steps[i]+1 samples of the refined field (in_field of
 *                          prop.c); sample n_steps[i] may be 0.
 *   out + out_off[i]:     5 rows of n_steps[i] doubles each, in the order
 *                          sourceterm, PopTot, PopInt, expval, Efield (the gated
 *                          field), i.e. the exported samples 0..n_steps[i]-1.
 * field_off and out_off are host arrays.  Returns a ctdse_status once the work
 * queued on `stream` has finished.
 */
int ctdse_propagate_batch(int n_inst, const ctdse_geometry *const *geom, const int *n_steps, const double *field,
                          const long long *field_off, double *out, const long long *out_off, void *stream);

/*
 * One instance alone, the propagation call1DTDSE runs (no counterpart of its own in the
 * reference: 1DTDSE/sources/singleTDSE.c:105 calls propagation there).  Arguments as for
 * one instance of ctdse_propagate_batch.  The segments of each step run side by side, one
 * launch per step, so one instance gets most of the GPU; both recurrences are truncated at a
 * halo on either side of a segment, so results equal the batch's to rounding (about 1e-13
 * relative on the dipole), not bitwise.  Returns once the work queued on `stream` has
 * finished.
 */
int ctdse_propagate_one(const ctdse_geometry *geom, int n_steps, const double *field, double *out, void *stream);

/*
 * ctdse_propagate_one without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results (bit for bit) as
 * ctdse_propagate_one; its scratch (two state vectors, the gated field, per-segment partial
 * sums, the descriptors) lives in a fixed device array of CTDSE_NOALLOC_SCRATCH_BYTES that
 * each GPU the library runs on holds until the process ends, and it calls no hipMalloc,
 * hipFree or allocating library routine.  A run needs about
 * 32 * ctdse_state_length(geom) + 8 * (n_steps + 1) bytes plus a few KiB: larger runs fail
 * with CTDSE_ERR_ARGUMENT; a geometry whose truncation halo leaves no room for the
 * step-parallel solve fails with CTDSE_ERR_UNSUPPORTED.  It queues the work on `stream` and
 * returns without waiting.  Every call on one GPU shares that GPU's array and leaves its
 * contents behind (each call initialises what it reads): calls that overlap on one GPU, on
 * other streams or from other threads, race; order them on one stream.
 */
#define CTDSE_NOALLOC_SCRATCH_BYTES (4 << 20)
int ctdse_propagate_one_noalloc(const ctdse_geometry *geom, int n_steps, const double *field, double *out,
                                void *stream);

/*
 * Many geometries at once (no counterpart in the reference; the task farm of
 * MPI-TDSE.c:145 computes one ground state for all its instances): geometry i has
 * (num_r[i], dx[i], dt[i], a[i], e_guess[i], cv[i], x_int[i]) and its ground state by
 * inverse iteration, as ctdse_geometry_create with psi0 = NULL.  The ground states run
 * side by side in one launch per group, then the tables.  out[i] receives the
 * geometry (NULL on failure), status[i] its ctdse_status (CTDSE_ERR_GS_NOT_CONVERGED
 * still yields a geometry).  Returns the first non-OK status, or CTDSE_OK.
 */
int ctdse_geometry_create_batch(int n, const int *num_r, const double *dx, const double *dt, const double *a,
                                const double *e_guess, const double *cv, const double *x_int, ctdse_geometry **out,
                                int *status);

/* Launch shape of the persistent propagation kernel: waves per workgroup (one segment
 * of the parallel tridiagonal solve covers 64 * waves_per_group chunks of 8 points),
 * resident workgroups per compute unit, and time steps per pass over the state (1, or 2:
 * the intermediate state stays on chip wherever no save or snapshot falls between the
 * two steps; where groups_per_cu workgroups still fit a CU's LDS with it, each segment's
 * 1/beta also stays in LDS from the first step to the second). */
typedef struct ctdse_launch_config {
    int waves_per_group;
    int groups_per_cu;
    int steps_per_pass;
} ctdse_launch_config;

/* The shape ctdse_propagate_batch uses for these geometries. */
int ctdse_launch_config_default(int n_inst, const ctdse_geometry *const *geom, ctdse_launch_config *cfg);

/*
 * Prepared batch (no counterpart in the reference).  ctdse_batch_plan_create takes the
 * arguments of ctdse_propagate_batch, the distance in doubles between an instance's five
 * output rows (out_row_stride >= n_steps[i]; 0: n_steps[i], rows packed as in
 * ctdse_propagate_batch) and a launch shape (NULL: the default); it validates them,
 * uploads the instance descriptors and allocates a private work buffer, once.
 *   ctdse_batch_plan_enqueue:      queues the batch on `stream` and returns at once; it
 *                                  allocates, copies and waits for nothing, so a caller
 *                                  may capture it into a graph.
 *   ctdse_batch_plan_graph_launch: the first call captures the same work into a HIP graph
 *                                  on a stream of the library's own; every call launches
 *                                  that graph on `stream` and returns at once.
 * Every run reads the fields and writes the outputs at the device addresses given at
 * creation, so a caller may refill the field buffer between runs.  Runs of one plan must
 * not overlap (they share the queue counter and the work buffer): order them on one
 * stream.  The geometries must outlive the plan.  ctdse_batch_plan_destroy waits for the
 * plan's last run outside a stream capture before it frees the plan's buffers.
 */
typedef struct ctdse_batch_plan ctdse_batch_plan;
int ctdse_batch_plan_create(int n_inst, const ctdse_geometry *const *geom, const int *n_steps, const double *field,
                            const long long *field_off, double *out, const long long *out_off,
                            long long out_row_stride, const ctdse_launch_config *config, ctdse_batch_plan **plan);
int ctdse_batch_plan_enqueue(ctdse_batch_plan *plan, void *stream);
int ctdse_batch_plan_graph_launch(ctdse_batch_plan *plan, void *stream);
/* Shape in use and the number of persistent workgroups of one run. */
int ctdse_batch_plan_info(const ctdse_batch_plan *plan, ctdse_launch_config *config, int *workgroups);
void ctdse_batch_plan_destroy(ctdse_batch_plan *plan);

/* Length, in complex doubles, of one device state vector of this geometry (the tiled
 * layout with its zero padding); states below are arrays of such vectors. */
long long ctdse_state_length(const ctdse_geometry *g);

/*
 * Partial forward runs for checkpointed gradients (no counterpart in the reference).
 * Instance i runs steps k0[i] .. k0[i]+n_run[i]-1 (n_run = 0: k0 .. n_steps-1) of a run of
 * n_steps[i] steps, driven by the already gated field F (field + field_off[i], indexed by
 * step), starting from psi_init + init_off[i] (a state vector; psi_init NULL: the ground
 * state, then k0 must be 0).  After every save_every-th step the state is written to
 * save + save_off[i] + m * state_length (m = 0, 1, ...; save may be NULL; the buffer must
 * start zeroed: only grid rows are written).  When an instance's save slot is its start
 * state (same address) and n_run == save_every, that state vector is stepped in place
 * instead (no copy).  out (may be NULL) receives the observables as in
 * ctdse_propagate_batch, entries k0+1 .. k0+n_run of a partial run (entry 0 too when k0 is
 * 0).  All arrays are device memory; offsets are host arrays in doubles.
 */
int ctdse_propagate_states(int n_inst, const ctdse_geometry *const *geom, const int *n_steps, const int *k0,
                           const int *n_run, const double *field, const long long *field_off, const double *psi_init,
                           const long long *init_off, double *save, const long long *save_off, int save_every,
                           double *out, const long long *out_off, void *stream);

/*
 * Adjoint sweep (no counterpart in the reference): for k = k_hi[i] down to k_lo[i]+1 it
 * updates the adjoint state kappa (a state vector, start from zero at k = n_steps-1) and
 * writes grad[k-1] = dL/dF_{k-1}, given gobs[k] = dL/d sourceterm_k, the gated field F and
 * the states psi_k at states + states_off[i] + (k - k_lo[i] - 1) * state_length.
 */
int ctdse_adjoint_batch(int n_inst, const ctdse_geometry *const *geom, const int *k_hi, const int *k_lo,
                        const double *field, const long long *field_off, const double *gobs, const long long *gobs_off,
                        const double *states, const long long *states_off, double *kappa, const long long *kappa_off,
                        double *grad, const long long *grad_off, void *stream);

/*
 * Adjoint sweep that rebuilds the states (no counterpart in the reference): for k = k_hi[i]
 * down to 1 it writes grad[k-1] = dL/dF_{k-1} (grad + grad_off[i]), given gobs[k] =
 * dL/d sourceterm_k and the gated field F, with the adjoint state starting from zero at
 * k_hi[i] (the last observed step, n_steps - 1).  psi_k comes from psi_hi + hi_off[i] (the
 * state vector psi_{k_hi}) by inverse Crank-Nicolson steps, replaced by the checkpoint
 * ckpt + ckpt_off[i] + (m - 1) * state_length whenever k = m * every (m = 1 .. n_ckpt; ckpt
 * may be NULL).  One call covers the whole backward pass; it keeps two state vectors per
 * resident workgroup and stores no states.  Device arrays; offsets are host arrays in doubles.
 */
int ctdse_adjoint_reverse_batch(int n_inst, const ctdse_geometry *const *geom, const int *k_hi, const double *field,
                                const long long *field_off, const double *gobs, const long long *gobs_off,
                                const double *psi_hi, const long long *hi_off, const double *ckpt,
                                const long long *ckpt_off, int every, int n_ckpt, double *grad,
                                const long long *grad_off, void *stream);

#ifdef __cplusplus
}
#endif

#endif
