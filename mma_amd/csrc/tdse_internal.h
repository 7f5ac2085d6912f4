// Internal device-side structures of libsingleTDSE.so (not part of the public ABI).
//
// Wavefunction/table layout in HBM ("chunk-interleaved tiles"): the grid is cut
// into chunks of TDSE_CHUNK consecutive points; 64 consecutive chunks form a
// tile stored point-major, chunk-minor:
//     index(j) = (c / 64) * TDSE_TILE + (j % TDSE_CHUNK) * 64 + (c % 64),  c = j / TDSE_CHUNK
// so the 64 lanes of a wave, each owning one chunk, read and write 64
// consecutive complex values per instruction.  One zero tile precedes tile 0 and
// one follows the last tile, so every chunk index in [-64, nchunks + 64) is
// addressable and a wave never needs a bounds test on a load; points past the
// end of the grid are padding rows (identity rows with zero data).
#pragma once
#include <hip/hip_runtime.h>

#ifndef TDSE_CHUNK
#define TDSE_CHUNK 8
#endif
#define TDSE_WAVE 64
#define TDSE_TILE (TDSE_WAVE * TDSE_CHUNK)
#define TDSE_MAX_WAVES 8
#ifndef TDSE_WAVES_PER_SIMD
#define TDSE_WAVES_PER_SIMD 3      // propagation kernel: register budget 168 VGPRs, 12 waves per CU
#endif
#define TDSE_WAVES_PER_CU (4 * TDSE_WAVES_PER_SIMD)
#define TDSE_PAIR_WAVES_PER_SIMD 2 // two-step propagation kernel: 256-VGPR budget, 8 waves per CU
#define TDSE_ADJ_WAVES_PER_CU 8    // adjoint kernel: 2 waves per SIMD (224 VGPRs)
#define TDSE_REV_WAVES_PER_SIMD 2  // reconstructing adjoint kernel: 256-VGPR budget, 8 waves per CU
#define TDSE_MAX_HALO 40           // largest truncation depth in chunks

// Instance queues of a persistent propagation launch, one per XCD: queue q holds
// order[off[q] .. off[q + 1]).  A workgroup takes instances from the queue of the XCD it runs
// on first, then from the others in turn, so each XCD's L2 mostly serves the tables of its own
// queue's instances.
#define TDSE_QUEUES 8
struct TdseQueues {
    int off[TDSE_QUEUES + 1];
};

// storage offset of chunk c (c >= -64), row i of the chunk
__host__ __device__ inline long long tdse_chunk_base(long long c) { return ((c >> 6) + 1) * TDSE_TILE + (c & 63); }

__host__ __device__ inline long long tdse_index(long long j)
{
    return tdse_chunk_base(j / TDSE_CHUNK) + (j % TDSE_CHUNK) * 64;
}

// number of elements of a tiled array for nchunks chunks (with the two zero tiles)
__host__ __device__ inline long long tdse_tiled_size(long long nchunks) { return ((nchunks + 63) / 64 + 2) * TDSE_TILE; }

// Step-independent tables of one geometry (num_r, dx, dt, a, x_int), device pointers.
// The Crank-Nicolson system of one step, scaled by 12, is
//     B' u = A' psi,  B' = M2' + i H',  A' = M2' - i H',  M2' = tridiag(1, 10, 1),
//     H' = tridiag(g_{j-1}, d_j, g_{j+1}),  g_j = -3 dt/dx^2 + dt V_j / 2,  d_j = 10 g_j + 36 dt/dx^2
// (1DTDSE/sources/prop.c:138-190 times 12).  ib_j = 1/beta_j are the pivots of
// the Thomas factorisation of B' (tridiag.c:26-70), which never change.
struct TdseGeom {
    const double *g;      // g_j, tiled
    const double2 *ib;    // 1/beta_j, tiled (0 on padding rows)
    const double *dpot;   // V'(x_j), tiled
    const double2 *psi0;  // ground state, tiled
    const double2 *chq;   // per chunk c at [c + 64]: product of q_j = -a'_j ib_j (forward transfer)
    const double2 *chg;   // per chunk c at [c + 64]: product of h_j = -c'_j ib_j (backward transfer)
    const double2 *adq;   // per chunk c at [c + 64]: product of conj(h_{j-1})  (adjoint forward transfer)
    const double2 *adg;   // per chunk c at [c + 64]: product of conj(p_j), p_j = -(1 + i g_j) ib_j
                          //                           (adjoint backward transfer)
    int nx;               // num_r + 1 grid points
    int nchunks;          // ceil(nx / TDSE_CHUNK)
    long long npad;       // tiled array length, tdse_tiled_size(nchunks)
    int halo_lanes;       // H: truncation depth in chunks at the right end of a workgroup segment
    int pint_lo, pint_hi; // PopInt window [k1, k4] (prop.c:311-318)
    int p0_lo, p0_hi;     // rows where |psi0| > 1e-16 max|psi0| (the terms of <psi0|psi>)
    double dx, dt, xmax;  // x_j = j dx - xmax
    double k36;           // 36 dt / dx^2
    double pint0;         // PopInt at t = tmin (prop.c:86-93)
    // Lattice tables shared by the geometries of a batch that differ only in their extent
    // (same dx, dt, a, parity of num_r: x_j = (j - num_r/2) dx on one lattice), or null: g,
    // V' and 1/beta of the widest geometry in tiled order, where chunk c of this geometry is
    // shared chunk c + s_off.  The forward kernels read g and V' there, and 1/beta from chunk
    // head_chunks on (before it, the pivots still feel this geometry's left wall and come from
    // ib); rows past this grid read as zero.  The per-geometry tables above stay complete.
    const double *sg, *sdp;
    const double2 *sib;
    const double2 *schq, *schg;  // the lattice's chunk transfer maps at shared chunk k + 64
    int s_off, head_chunks;
};

// One instance of a batch.
struct TdseInst {
    int geom;             // index into the geometry array
    int n_steps;          // outputs.Nt of the reference: steps run and samples exported
    const double *field;  // in_field, n_steps + 1 samples
    double *src, *ptot, *pint, *expv;  // n_steps samples each (src == null: no observables)
    double *efield;       // gated field, n_steps samples (may be null)
    double *snap;         // snapshots (natural order, interleaved), n_snap rows (may be null)
    int snap_every, n_snap;
    double *psi_final;    // final state, natural order (may be null)
    // partial runs (checkpointed gradients): steps k0 .. k0+n_run-1 starting from psi_init
    int k0, n_run;        // n_run = 0: the whole run 0 .. n_steps-1
    int pregated;         // field holds the gated F_k already (no gating test)
    const double2 *psi_init;  // tiled start state (null: psi0)
    double2 *save;        // tiled states saved after every save_every-th step (may be null)
    int save_every;
    int inplace;          // psi_init is stepped in place (no work copy, no saves)
};

// One instance of an adjoint (backward) sweep over steps k_hi .. k_lo+1, see tdse_adjoint_kernel.
struct TdseAdjInst {
    int geom;
    int k_hi, k_lo;           // produces kappa_{k-1} and dL/dF_{k-1} for k = k_hi down to k_lo + 1
    const double *field;      // gated field F_k
    const double *gobs;       // dL/d sourceterm_k
    const double2 *psi;       // tiled states psi_k for k = k_lo+1 .. k_hi at psi + (k - k_lo - 1) * npad
    double2 *kappa;           // tiled adjoint state, updated in place
    double *grad;             // dL/dF_k, written for k = k_hi-1 down to k_lo
};

// One instance of a reconstructing adjoint sweep (tdse_reverse_adjoint_kernel): psi_k is
// rebuilt backward from psi_{k_hi} by inverse Crank-Nicolson steps, re-anchored at the
// checkpoints, while kappa_k runs backward from kappa_{k_hi} = 0.
struct TdseRevInst {
    int geom;
    int k_hi;                 // sweeps k = k_hi down to 1 (k_hi <= n_steps - 1 of the forward)
    const double *field;      // gated field F_k
    const double *gobs;       // dL/d sourceterm_k, k <= k_hi
    const double2 *psi_hi;    // tiled psi_{k_hi}
    const double2 *ckpt;      // tiled psi_{(m+1) every} at ckpt + m * npad, m < n_ckpt (may be null)
    int every, n_ckpt;
    double *grad;             // dL/dF_k, written for k = k_hi - 1 down to 0
};

// Ground-state job: inverse iteration in one lane (bit-identical to the CPU reference).
struct GsJob {
    int num_r, max_iter;
    double dx, a, e_guess, cv;
    double *x;        // num_r + 1
    double *psi0;     // 2 (num_r + 1), natural order, interleaved
    double *work;     // 16 (num_r + 2) doubles of scratch
    double energy;
    int iters, status;
};

// Table-builder job for one geometry.
struct TableJob {
    int num_r;
    double dx, dt, a, x_int;
    const double *x;        // natural order grid (from the ground-state job or the caller)
    const double *psi0_nat; // natural order interleaved ground state
    double *g, *dpot;       // tiled
    double2 *ib, *psi0, *chq, *chg, *adq, *adg;  // per-chunk tables: nchunks + 128 entries, chunk c at c + 64
    double *pivot_work;     // 2 (num_r + 1) scratch (natural-order 1/beta)
    double rho_fwd, rho_bwd, pint0;
    int pint_lo, pint_hi, p0_lo, p0_hi;
};

// kernels (defined in the .hip files)
__global__ void tdse_ground_state_kernel(GsJob *jobs, int n_jobs);
__global__ void tdse_tables_grid_kernel(TableJob *job);
__global__ void tdse_tables_pivot_kernel(TableJob *job);
__global__ void tdse_tables_chunk_kernel(TableJob *job);
__global__ void tdse_lattice_kernel(const double *g, const double *dpot, const double2 *ib, int nx, int shift,
                                    long long npad, double *sg, double *sdp, double2 *sib);
__global__ void tdse_lattice_chunk_kernel(const double *sg, const double2 *sib, int nchunks, double2 *schq,
                                          double2 *schg);
__global__ void tdse_tables_ib_kernel(TableJob *jobs);
// host forms of the serial setup steps (tdse_setup.hip), bit-identical to the kernels
void tdse_ground_state_host(GsJob &J);
void tdse_tables_pivot_host(TableJob &J);
// counter: TDSE_QUEUES ints, zero at launch
__global__ void tdse_propagate_kernel(const TdseGeom *geoms, const TdseInst *insts, const int *order,
                                      TdseQueues queues, int *counter, double2 *work, long long work_stride);
__global__ void tdse_propagate_pair_kernel(const TdseGeom *geoms, const TdseInst *insts, const int *order,
                                           TdseQueues queues, int *counter, double2 *work, long long work_stride);
__global__ void tdse_propagate_pair_stash_kernel(const TdseGeom *geoms, const TdseInst *insts, const int *order,
                                                 TdseQueues queues, int *counter, double2 *work,
                                                 long long work_stride);
__global__ void tdse_untile_kernel(const double2 *psi, int nx, double *row);
__global__ void tdse_parallel_step_kernel(const TdseGeom *geom, const TdseInst *insts, int k, const double2 *psi_in,
                                          double2 *psi_out, long long state_stride, double *partial, int nseg);
__global__ void tdse_parallel_finish_kernel(const TdseInst *insts, int n_inst, const double *partial, int nseg);
__global__ void tdse_gate_kernel(const double *f, int n, double *F);
__global__ void tdse_adjoint_kernel(const TdseGeom *geoms, const TdseAdjInst *insts, const int *order, int n_inst,
                                    int *counter);
__global__ void tdse_reverse_adjoint_kernel(const TdseGeom *geoms, const TdseRevInst *insts, const int *order,
                                            int n_inst, int *counter, double2 *work, long long work_stride);
__global__ void tdse_fill_int_kernel(int *p, int n, int v);
__global__ void tdse_xgrid_kernel(double *x, int num_r, double dx);
__global__ void tdse_tgrid_kernel(double *tgrid, int n, double tmin, double dt);
__global__ void tdse_dft_r2c_kernel(const double *sig, int n, int nbins, double2 *out);
__global__ void tdse_spectrum_kernel(const double2 *dft, int nbins, double dt, double tmax, double *omega,
                                     double *F, double *M2);
__global__ void tdse_c2r_interp_kernel(const double2 *bins, int nbins, int n_in, int n_out, double *out);
