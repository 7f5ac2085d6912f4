// Persistent batched Crank-Nicolson propagator (1DTDSE/sources/prop.c:24-347) for gfx950.
//
// One workgroup of NW waves owns one instance at a time (instances are taken from
// an atomic counter in the given order, largest first) and runs all its time
// steps.  Per step, the grid is covered by segments, solved in order by all its waves:
//
//   lanes each own one chunk of TDSE_CHUNK consecutive points; a segment's window is its
//   own chunks followed by H halo chunks.  The tridiagonal solve B'u = A'psi uses the global
//   Thomas factorisation of B' (pivot table ib, step-independent), written as two
//   first-order linear recurrences
//       forward   z_j = ib_j (r_j - a'_j z_{j-1})
//       backward  u_j = z_j + h_j u_{j+1},   h_j = -c'_j ib_j
//   each evaluated in two passes around a 64-lane scan of chunk transfer maps
//   (carry-out with zero carry-in, exclusive scan with the chunk products from
//   the tables, recompute with the true carry-in).  The forward recurrence is exact: it
//   enters each segment with z at the last point of the previous segment, which that
//   segment left in LDS (segments run left to right).  The backward recurrence starts
//   from zero at the right end of the window: the inverse of B' decays by |h| <= rho < 1
//   per point, so after H*TDSE_CHUNK >= log(1e-17)/log(rho) points the truncation is
//   below 1e-17 of max|psi| and the own chunks are exact to rounding (truncated SPIKE).
//   Segments read psi of step k and overwrite it with step k+1 in place: a segment reads
//   no chunk an earlier one wrote except the point just left of its window, whose old
//   value the previous segment also left in LDS.  Waves whose lanes all lie past the
//   grid (the tail of the last segment) skip the solve and only join the barriers.
//
// Observables (prop.c:293-347) are reduced per wave, then across waves in LDS in a
// fixed order (deterministic).  HBM traffic per grid-point-step: psi read + write,
// 32 bytes; the tables are shared by all instances of a geometry and stay in L2.
#include "tdse_internal.h"
#include "tdse_wave.h"

namespace {

using namespace tdse;

// Solve one segment of one step with all waves of the workgroup.
// snap_row != nullptr: also store the new psi there (natural order).
// Each lane keeps its chunk of psi and the chunk's coefficients (g_j, 1/beta_j) in
// registers for the whole solve: all loads of a segment are issued together, once.
// edge_in / edge_out: {old psi, forward z} at the last own point of the previous / this
// segment (LDS, alternating between segments).
template <bool SAVE, bool SNAP>
__device__ void solve_segment(const TdseGeom &G, int seg, int own_chunks, int wave, int lane, int nw,
                              double2 *__restrict__ psi, double Fdt, double2 rot, Acc &acc, double *snap_row,
                              double2 (*slot_f)[2], double2 (*slot_b)[2], const double2 *edge_in, double2 *edge_out,
                              double2 *__restrict__ save_to)
{
    const int gl = wave * 64 + lane;                 // lane index in the workgroup
    const int cs = seg * own_chunks;                 // first own chunk
    // live waves: those with a lane on the grid (chunks past the grid are zero padding and
    // carry the zero map, so dropping whole waves of them changes nothing)
    const int nwl = min(nw, (G.nchunks - cs + 63) >> 6);
    if (wave >= nwl) {
        __syncthreads();  // the two barriers of block_carry
        __syncthreads();
        return;
    }
    const int nl = nwl * 64;
    const int c = cs + gl;
    // lanes past the grid read chunk nchunks instead, which is zero in every buffer and
    // table (padding row)
    const int cl = min(c, G.nchunks);
    const long long base = tdse_chunk_base(cl);

    double2 P[C], IB[C];
    double Gv[C];
#pragma unroll
    for (int i = 0; i < C; i++) {
        P[i] = ld2(psi, base + 64 * i);
        Gv[i] = ld1(G.g, base + 64 * i);
        IB[i] = ld2(G.ib, base + 64 * i);
    }
    if (gl == own_chunks - 1) edge_out[0] = P[C - 1];  // the next segment's left neighbour
    const double2 Qc = ld2(G.chq, cl + 64);
    const double2 Gc = ld2(G.chg, cl + 64);
    // neighbours across the chunk edges: shuffles inside a wave, loads across waves,
    // LDS for the point left of the window, zero past the window's right end
    const double2 zero2 = make_double2(0.0, 0.0);
    double2 left = dpp2<kWaveShr1>(P[C - 1], zero2), right = dpp2<kWaveShl1>(P[0], zero2);
    double gL = dpp<kWaveShr1>(Gv[C - 1], 0.0), gR = dpp<kWaveShl1>(Gv[0], 0.0);
    if (lane == 0) {
        const long long nb = tdse_chunk_base(min(c - 1, G.nchunks)) + 64 * (C - 1);  // chunk -1: zero padding
        left = gl > 0 ? ld2(psi, nb) : (seg > 0 ? edge_in[0] : zero2);  // gl > 0: not yet overwritten
        gL = ld1(G.g, nb);
    }
    if (lane == 63) {
        const long long nb = tdse_chunk_base(min(c + 1, G.nchunks));
        right = gl < nl - 1 ? ld2(psi, nb) : zero2;
        gR = gl < nl - 1 ? ld1(G.g, nb) : 0.0;
    }
    const double2 z_in = seg > 0 ? edge_in[1] : zero2;

    // r' = (M2' - i H') psi
    {
        double2 pm = left;
        double gm = gL;
        const double K = G.k36;
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 p0 = P[i];
            const double2 pp = (i + 1 < C) ? P[i + 1] : right;
            const double gp = (i + 1 < C) ? Gv[i + 1] : gR;
            const double d = fma(10.0, Gv[i], K);
            const double mr = fma(10.0, p0.x, pm.x + pp.x);
            const double mi = fma(10.0, p0.y, pm.y + pp.y);
            const double hr = fma(gm, pm.x, fma(d, p0.x, gp * pp.x));
            const double hi = fma(gm, pm.y, fma(d, p0.y, gp * pp.y));
            P[i] = make_double2(mr + hi, mi - hr);
            pm = p0;
            gm = Gv[i];
        }
    }

    // forward recurrence z_j = ib_j (r_j - (1 + i g_{j-1}) z_{j-1})
    {
        double2 z = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double gm = (i > 0) ? Gv[i - 1] : gL;
            const double tr = fma(gm, z.y, P[i].x - z.x);
            const double ti = fma(-gm, z.x, P[i].y - z.y);
            z = cmul(make_double2(tr, ti), IB[i]);
        }
        double2 Q = Qc;
        wave_scan<true>(Q, z, lane);
        z = block_carry<true>(Q, z, wave, lane, nwl, slot_f, z_in);
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double gm = (i > 0) ? Gv[i - 1] : gL;
            const double tr = fma(gm, z.y, P[i].x - z.x);
            const double ti = fma(-gm, z.x, P[i].y - z.y);
            z = cmul(make_double2(tr, ti), IB[i]);
            P[i] = z;
        }
        if (gl == own_chunks - 1) edge_out[1] = z;  // the next segment's forward carry
    }

    // backward recurrence u_j = z_j + h_j u_{j+1},  h_j = -(1 + i g_{j+1}) ib_j
    {
        double2 u = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double gp = (i + 1 < C) ? Gv[i + 1] : gR;
            const double hr = fma(gp, IB[i].y, -IB[i].x), hi = -fma(gp, IB[i].x, IB[i].y);
            u = make_double2(P[i].x + hr * u.x - hi * u.y, P[i].y + hr * u.y + hi * u.x);
        }
        double2 Gt = Gc;
        wave_scan<false>(Gt, u, lane);
        u = block_carry<false>(Gt, u, wave, lane, nwl, slot_b);
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double gp = (i + 1 < C) ? Gv[i + 1] : gR;
            const double hr = fma(gp, IB[i].y, -IB[i].x), hi = -fma(gp, IB[i].x, IB[i].y);
            u = make_double2(P[i].x + hr * u.x - hi * u.y, P[i].y + hr * u.y + hi * u.x);
            P[i] = u;
        }
    }

    // keep the loads of the observable tables below this point: hoisted above the passes they
    // would hold registers through the whole solve
    asm volatile("" ::: "memory");
    if (gl >= own_chunks || c >= G.nchunks) return;  // only the segment's own chunks continue

    // length-gauge phase exp(+i F dt x_j) (prop.c:199-204) by rotation from the chunk start,
    // store, and the observables of prop.c:302-333 that need every point
    const int j0 = c * C;
    const double x0 = __dsub_rn(__dmul_rn((double)j0, G.dx), G.xmax);
    double s0, c0;
    sincos(Fdt * x0, &s0, &c0);
    double2 e = make_double2(c0, s0);
    const double *dp = G.dpot + base;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const double2 v = cmul(e, P[i]);
        P[i] = v;
        st2(psi, base + 64 * i, v);
        if constexpr (SAVE) {
            if (save_to) st2(save_to, base + 64 * i, v);
        }
        const double n2 = fma(v.x, v.x, v.y * v.y);
        const double xj = __dsub_rn(__dmul_rn((double)(j0 + i), G.dx), G.xmax);
        acc.pos = fma(n2, xj, acc.pos);
        acc.gp = fma(n2, ld1(dp, 64 * i), acc.gp);
        e = cmul(e, rot);
    }
    // <psi0|psi> over the rows where psi0 is not negligible
    if (j0 + C - 1 >= G.p0_lo && j0 <= G.p0_hi) {
        const double2 *q0 = G.psi0 + base;
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 q = ld2(q0, 64 * i);
            acc.pr = fma(P[i].x, q.x, fma(P[i].y, q.y, acc.pr));
            acc.pi = fma(P[i].x, q.y, fma(-P[i].y, q.x, acc.pi));
        }
    }
    // PopInt window [k1, k4]
    if (j0 + C - 1 >= G.pint_lo && j0 <= G.pint_hi) {
#pragma unroll
        for (int i = 0; i < C; i++) {
            const int j = j0 + i;
            if (j >= G.pint_lo && j <= G.pint_hi) acc.pint += fma(P[i].x, P[i].x, P[i].y * P[i].y);
        }
    }
    if (SNAP && snap_row) {
#pragma unroll
        for (int i = 0; i < C; i++) {
            if (j0 + i < G.nx) {
                snap_row[2 * (j0 + i)] = P[i].x;
                snap_row[2 * (j0 + i) + 1] = P[i].y;
            }
        }
    }
}

} // namespace

// SAVE: some instance stores states (partial runs); SNAP: some instance takes snapshots.
// Both are off for plain batches, which then carry neither store in the step loop.
template <bool SAVE, bool SNAP>
__global__ void __launch_bounds__(64 * TDSE_MAX_WAVES, TDSE_WAVES_PER_SIMD)
tdse_propagate_kernel(const TdseGeom *geoms, const TdseInst *insts, const int *order, int n_inst, int *counter,
                      double2 *work, long long work_stride)
{
    __shared__ double part[2][TDSE_MAX_WAVES][5];
    __shared__ double2 slot_f[TDSE_MAX_WAVES][2], slot_b[TDSE_MAX_WAVES][2];
    __shared__ double2 edge[2][2];
    __shared__ int cur;
    const int nw = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    double2 *buf = work + (long long)blockIdx.x * work_stride;

    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(cur);  // uniform: keeps the instance in SGPRs
        __syncthreads();
        if (t >= n_inst) break;
        const TdseInst &I = insts[order[t]];
        const TdseGeom &G = geoms[I.geom];
        const long long npad = G.npad;
        const int own_chunks = nw * 64 - G.halo_lanes;  // host guarantees >= 64
        const int nseg = (G.nchunks + own_chunks - 1) / own_chunks;

        // start state: psi0 (t = tmin) or a saved state; padding chunks are zero in both
        const double2 *init = I.psi_init ? I.psi_init : G.psi0;
        for (long long q = threadIdx.x; q < npad; q += blockDim.x) buf[q] = init[q];
        const int k_start = I.k0;
        const int k_stop = I.n_run > 0 ? I.k0 + I.n_run : I.n_steps;
        if (threadIdx.x == 0 && I.src && k_start == 0) {
            I.src[0] = 0.0;
            I.ptot[0] = 1.0;
            I.pint[0] = G.pint0;
            I.expv[0] = 0.0;
            if (I.efield) I.efield[0] = 0.0;
        }
        __syncthreads();

        bool gating = !I.pregated;  // prop.c:125-134
        int i_snap = 0;
        for (int k = k_start; k < k_stop; k++) {
            const double fk = I.field[k];
            double F;
            if (gating) {
                if (fk * I.field[k + 1] <= 0.0) gating = false;
                F = 0.0;
            } else {
                F = fk;
            }
            double2 *save_to = nullptr;
            if (SAVE && I.save && (k - k_start + 1) % I.save_every == 0)
                save_to = I.save + (long long)((k - k_start + 1) / I.save_every - 1) * npad;
            const double Fdt = F * G.dt;
            double rs, rc;
            sincos(Fdt * G.dx, &rs, &rc);
            const double2 rot = make_double2(rc, rs);
            // snapshots, prop.c:248-255: row i_snap gets psi after step k when k % every == 0
            double *snap_row = nullptr;
            if (SNAP && I.snap && k % I.snap_every == 0 && i_snap < I.n_snap - 1) {
                snap_row = I.snap + (long long)i_snap * 2 * G.nx;
                i_snap++;
            }
            Acc acc = {0.0, 0.0, 0.0, 0.0, 0.0};
            for (int s = 0; s < nseg; s++)
                solve_segment<SAVE, SNAP>(G, s, own_chunks, wave, lane, nw, buf, Fdt, rot, acc, snap_row, slot_f,
                                          slot_b, edge[s & 1], edge[(s + 1) & 1], save_to);
            // a final snapshot that coincides with the last step is written after the loop
            const double pr = wave_sum(acc.pr), pi = wave_sum(acc.pi), pint = wave_sum(acc.pint);
            const double pos = wave_sum(acc.pos), gp = wave_sum(acc.gp);
            if (lane == 0) {
                part[k & 1][wave][0] = pr;
                part[k & 1][wave][1] = pi;
                part[k & 1][wave][2] = pint;
                part[k & 1][wave][3] = pos;
                part[k & 1][wave][4] = gp;
            }
            __syncthreads();
            if (threadIdx.x == 0 && I.src && k + 1 < I.n_steps) {
                double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
                for (int w = 0; w < nw; w++) {
                    s0 += part[k & 1][w][0];
                    s1 += part[k & 1][w][1];
                    s2 += part[k & 1][w][2];
                    s3 += part[k & 1][w][3];
                    s4 += part[k & 1][w][4];
                }
                I.src[k + 1] = -s4 + F;
                I.ptot[k + 1] = s0 * s0 + s1 * s1;
                I.pint[k + 1] = s2;
                I.expv[k + 1] = s3;
                if (I.efield) I.efield[k + 1] = F;
            }
        }
        // prop.c:256-262: after the last step, row n_snap-1 gets the final state if every
        // earlier row was used; the final state itself is optional output
        const bool last_snap = I.snap && I.n_snap > 0 && i_snap == I.n_snap - 1;
        if (last_snap || I.psi_final) {
            const double2 *fin = buf;
            double *row = last_snap ? I.snap + (long long)(I.n_snap - 1) * 2 * G.nx : nullptr;
            for (int j = threadIdx.x; j < G.nx; j += blockDim.x) {
                const double2 v = fin[tdse_index(j)];
                if (I.psi_final) {
                    I.psi_final[2 * j] = v.x;
                    I.psi_final[2 * j + 1] = v.y;
                }
                if (row) {
                    row[2 * j] = v.x;
                    row[2 * j + 1] = v.y;
                }
            }
        }
        __syncthreads();
    }
}

template __global__ void tdse_propagate_kernel<false, false>(const TdseGeom *, const TdseInst *, const int *, int, int *,
                                                             double2 *, long long);
template __global__ void tdse_propagate_kernel<true, false>(const TdseGeom *, const TdseInst *, const int *, int, int *,
                                                            double2 *, long long);
template __global__ void tdse_propagate_kernel<false, true>(const TdseGeom *, const TdseInst *, const int *, int, int *,
                                                            double2 *, long long);
template __global__ void tdse_propagate_kernel<true, true>(const TdseGeom *, const TdseInst *, const int *, int, int *,
                                                           double2 *, long long);
