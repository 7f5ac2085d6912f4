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
// Observables (prop.c:293-347) of the state after step k are taken in step k + 1 from the
// psi it loads anyway (a final pass covers the last state of a partial run); they are
// reduced per wave, then across waves in LDS in a fixed order (deterministic).  HBM traffic
// per grid-point-step: psi read + write, 32 bytes; the tables are shared by all instances of
// a geometry and stay in L2.
//
// Also here: the two-step pass (tdse_propagate_pair_kernel: segment s of step k kept in
// registers while segment s-1 of step k+1 is solved from them; the stash form keeps 1/beta in
// LDS between the two), and the step-parallel form for one instance alone
// (tdse_parallel_step_kernel: every segment a workgroup of its own, halo on both sides, one
// launch per step).
#include "tdse_internal.h"
#include "tdse_wave.h"

namespace {

using namespace tdse;

// Observables of prop.c:302-333 that need every point, over one lane's chunk of a state
// (P, the chunk's rows at tiled offset base, first point j0): sum |psi|^2 x, sum |psi|^2 V',
// <psi0|psi> over the rows where psi0 is not negligible, and the PopInt window [k1, k4]
// (dp: V' at the chunk's rows).
// Only lanes holding one of the segment's own chunks add (own).
__device__ inline void observe_chunk(const TdseGeom &G, const double2 (&P)[C], const double (&dp)[C], long long base,
                                     int j0, bool own, Acc &acc)
{
    if (!own) return;
    const bool in_pint = j0 + C - 1 >= G.pint_lo && j0 <= G.pint_hi;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const double2 v = P[i];
        const double n2 = fma(v.x, v.x, v.y * v.y);
        const double xj = grid_x(j0 + i, G.dx, G.xmax);
        acc.pos = fma(n2, xj, acc.pos);
        acc.gp = fma(n2, dp[i], acc.gp);
        const int j = j0 + i;
        acc.pint = fma(in_pint && j >= G.pint_lo && j <= G.pint_hi ? 1.0 : 0.0, n2, acc.pint);  // exact 0 or n2
    }
    if (j0 + C - 1 >= G.p0_lo && j0 <= G.p0_hi) {
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 q = ld2(G.psi0, base + 64 * i);
            acc.pr = fma(P[i].x, q.x, fma(P[i].y, q.y, acc.pr));
            acc.pi = fma(P[i].x, q.y, fma(-P[i].y, q.x, acc.pi));
        }
    }
}

// Step-independent table values of one lane's chunk in one segment.
struct SegTables {
    double2 IB[C];   // 1/beta_j
    double Gv[C];    // g_j
    double dp[C];    // V'(x_j)
    double2 Qc, Gc;  // chunk transfer maps (forward, backward)
    double gL, gR;   // g left of the chunk (lane 0) and right of it (lane 63)
};

// IB_LDS: 1/beta from `stash` (point i of lane gl at stash[i * 64 nw + gl]) instead of HBM.
// ROW_PTRS: address the C rows of a table from one opaque lane pointer (lane_ptr), each row an
// immediate offset: 1-1.6 % faster in the one-step kernel, 0.5 % slower in the two-step ones
// (3.6 % in the adjoint sweep), which keep the plain indexing.
template <bool IB_LDS = false, bool ROW_PTRS = false>
__device__ inline void load_tables(const TdseGeom &G, int cs, int wave, int lane, int nw, SegTables &T,
                                   const double2 *stash = nullptr)
{
    const int nwl = min(nw, (G.nchunks - cs + 63) >> 6);
    const int gl = wave * 64 + lane;
    const int c = cs + gl;
    const int cl = min(c, G.nchunks);
    const long long base = tdse_chunk_base(cl);
    if (G.sg) {
        // lattice tables shared with the other geometries of the batch (uniform branch): rows
        // past this grid, and the point left of it, read as zero; the first H chunks and the
        // last one keep their own pivots and chunk maps (the walls of this grid)
        const long long sb = tdse_chunk_base(cl + G.s_off);
        const int w0 = cs + wave * 64;  // this wave's first chunk
        if (w0 > 0 && w0 >= G.head_chunks && w0 + 63 < G.nchunks - 1) {
            // (uniform) a wave with no head chunk and none of the last two chunks: every row and
            // the right neighbour on the grid, lattice rows and maps only, nothing to mask
            T.Qc = ld2(G.schq, cl + G.s_off + 64);
            T.Gc = ld2(G.schg, cl + G.s_off + 64);
            if constexpr (ROW_PTRS) {
                const double *pg = lane_ptr(G.sg, sb), *pd = lane_ptr(G.sdp, sb);
                const double2 *pib = lane_ptr(G.sib, sb);
#pragma unroll
                for (int i = 0; i < C; i++) {
                    T.Gv[i] = ld1(pg, 64 * i);
                    T.dp[i] = ld1(pd, 64 * i);
                    T.IB[i] = ld2(pib, 64 * i);
                }
            } else {
#pragma unroll
                for (int i = 0; i < C; i++) {
                    T.Gv[i] = ld1(G.sg, sb + 64 * i);
                    T.dp[i] = ld1(G.sdp, sb + 64 * i);
                    if constexpr (IB_LDS)
                        T.IB[i] = stash[i * 64 * nw + gl];
                    else
                        T.IB[i] = ld2(G.sib, sb + 64 * i);
                }
            }
            T.gL = lane == 0 ? ld1(G.sg, tdse_chunk_base(c - 1 + G.s_off) + 64 * (C - 1)) : 0.0;
            T.gR = lane == 63 && gl < nwl * 64 - 1 ? ld1(G.sg, tdse_chunk_base(c + 1 + G.s_off)) : 0.0;
            return;
        }
        const bool head = c < G.head_chunks;
        const bool own_maps = head || c >= G.nchunks - 1;
        T.Qc = ld2(own_maps ? G.chq : G.schq, (own_maps ? cl : cl + G.s_off) + 64);
        T.Gc = ld2(own_maps ? G.chg : G.schg, (own_maps ? cl : cl + G.s_off) + 64);
        const double2 zero2 = make_double2(0.0, 0.0);
        if constexpr (ROW_PTRS) {
            const double *pg = lane_ptr(G.sg, sb), *pd = lane_ptr(G.sdp, sb);
            const double2 *pib = lane_ptr(head ? G.ib : G.sib, head ? base : sb);
#pragma unroll
            for (int i = 0; i < C; i++) {
                const bool live = c * C + i < G.nx;
                const double gv = ld1(pg, 64 * i), dv = ld1(pd, 64 * i);
                const double2 v = ld2(pib, 64 * i);
                T.Gv[i] = live ? gv : 0.0;
                T.dp[i] = live ? dv : 0.0;
                T.IB[i] = live ? v : zero2;
            }
        } else {
#pragma unroll
            for (int i = 0; i < C; i++) {
                const bool live = c * C + i < G.nx;
                const double gv = ld1(G.sg, sb + 64 * i), dv = ld1(G.sdp, sb + 64 * i);
                T.Gv[i] = live ? gv : 0.0;
                T.dp[i] = live ? dv : 0.0;
                if constexpr (IB_LDS) {
                    T.IB[i] = stash[i * 64 * nw + gl];
                } else {
                    const double2 v = ld2(head ? G.ib : G.sib, (head ? base : sb) + 64 * i);
                    T.IB[i] = live ? v : zero2;
                }
            }
        }
        const int jl = c * C - 1, jr = (c + 1) * C;  // the points left and right of the chunk
        T.gL = lane == 0 && jl >= 0 && jl < G.nx ? ld1(G.sg, tdse_chunk_base(c - 1 + G.s_off) + 64 * (C - 1)) : 0.0;
        T.gR = lane == 63 && gl < nwl * 64 - 1 && jr < G.nx ? ld1(G.sg, tdse_chunk_base(c + 1 + G.s_off)) : 0.0;
        return;
    }
    T.Qc = ld2(G.chq, cl + 64);
    T.Gc = ld2(G.chg, cl + 64);
    if constexpr (ROW_PTRS) {
        const double *pg = lane_ptr(G.g, base), *pd = lane_ptr(G.dpot, base);
        const double2 *pib = lane_ptr(G.ib, base);
#pragma unroll
        for (int i = 0; i < C; i++) {
            T.Gv[i] = ld1(pg, 64 * i);
            T.IB[i] = ld2(pib, 64 * i);
            T.dp[i] = ld1(pd, 64 * i);
        }
    } else {
#pragma unroll
        for (int i = 0; i < C; i++) {
            T.Gv[i] = ld1(G.g, base + 64 * i);
            if constexpr (IB_LDS)
                T.IB[i] = stash[i * 64 * nw + gl];
            else
                T.IB[i] = ld2(G.ib, base + 64 * i);
            T.dp[i] = ld1(G.dpot, base + 64 * i);
        }
    }
    T.gL = lane == 0 ? ld1(G.g, tdse_chunk_base(min(c - 1, G.nchunks)) + 64 * (C - 1)) : 0.0;  // chunk -1: zero
    T.gR = lane == 63 && gl < nwl * 64 - 1 ? ld1(G.g, tdse_chunk_base(min(c + 1, G.nchunks))) : 0.0;
}

// On-chip hand-off between the two steps of a pair pass (see tdse_propagate_kernel): the
// first H chunks of the step-k+1 state of the segment just solved (the next stage's right
// halo), and each wave's first and last point of it (neighbours across wave edges).
struct PairLds {
    double2 halo[2][TDSE_MAX_HALO][C];
    double2 bnd[2][TDSE_MAX_WAVES][2];
};

// How a segment gets its state and where its result goes:
//   kSingle: psi_k from HBM, psi_{k+1} stored to HBM (one step per pass);
//   kKeep:   psi_k from HBM, psi_{k+1} kept in registers (R) and its first H chunks and
//            wave-edge points published in LDS (first step of a pair);
//   kFromRegs: psi_{k+1} from R (own chunks) and LDS (halo chunks, wave-edge neighbours),
//            psi_{k+2} stored to HBM (second step of a pair).
//   kParallel: one step of one segment whose window carries a halo on both sides (both
//            recurrences truncated), so that all segments of a step are independent; psi_k
//            from `psi`, psi_{k+1} of the own chunks to `psi_out`.
enum SegMode { kSingle = 0, kKeep = 1, kFromRegs = 2, kParallel = 3 };

// Solve one segment of one step with all waves of the workgroup.
// observe: add the observables of the state the step starts from (psi_k, read for the
// solve anyway) to acc; the step's own result is observed by the next step.
// Each lane keeps its chunk of psi and the chunk's coefficients (g_j, 1/beta_j) in
// registers for the whole solve: all loads of a segment are issued together, once.
// edge_in / edge_out: {old psi, forward z} at the last own point of the previous / this
// segment (LDS, alternating between segments).  par: parity of the segment (PairLds
// buffers).
// STASH (pair passes): the first stage (kKeep) leaves the segment's 1/beta in `stash` (LDS),
// where the second stage (kFromRegs) of the same segment reads it instead of fetching it again.
template <int MODE, bool STASH = false>
__device__ void solve_segment(const TdseGeom &G, int seg, int own_chunks, int wave, int lane, int nw,
                              double2 *__restrict__ psi, double2 e0, double2 rot, bool observe, Acc &acc,
                              double2 (*slot_f)[2], double2 (*slot_b)[2], const double2 *edge_in, double2 *edge_out,
                              double2 (&R)[C], PairLds *pl, double2 *__restrict__ psi_out = nullptr,
                              double2 *stash = nullptr)
{
    const int gl = wave * 64 + lane;                 // lane index in the workgroup
    // first chunk of the window: the first own chunk, or H chunks before it (kParallel)
    const int cs = MODE == kParallel ? seg * own_chunks - G.halo_lanes : seg * own_chunks;
    // live waves: those with a lane on the grid (chunks past the grid are zero padding and
    // carry the zero map, so dropping whole waves of them changes nothing)
    const int nwl = min(nw, (G.nchunks - cs + 63) >> 6);
    if (wave >= nwl) {
        lds_barrier();  // the two barriers of block_carry
        lds_barrier();
        return;
    }
    const int nl = nwl * 64;
    const int c = cs + gl;
    // lanes past the grid read chunk nchunks instead, which is zero in every buffer and
    // table (padding row)
    const int cl = min(c, G.nchunks);
    const long long base = tdse_chunk_base(cl);
    const int par = seg & 1;

    // raised priority while this wave issues the segment's loads, so that waves with loads to
    // issue go ahead of waves that only compute
    __builtin_amdgcn_s_setprio(2);
    SegTables T;
    if constexpr (STASH && MODE == kFromRegs) {
        load_tables<true>(G, cs, wave, lane, nw, T, stash);
    } else {
        load_tables<false, MODE == kSingle>(G, cs, wave, lane, nw, T);
        if constexpr (STASH && MODE == kKeep) {
#pragma unroll
            for (int i = 0; i < C; i++) stash[i * 64 * nw + gl] = T.IB[i];
        }
    }
    double2 P[C];
    const double2 zero2 = make_double2(0.0, 0.0);
    double2 nb = zero2;  // lane 0: the point left of the wave's first chunk; lane 63: right of its last
    if constexpr (MODE != kFromRegs) {
#pragma unroll
        for (int i = 0; i < C; i++) P[i] = ld2(psi, base + 64 * i);
        if (lane == 0 && (gl > 0 || MODE == kParallel))
            nb = ld2(psi, tdse_chunk_base(c - 1) + 64 * (C - 1));  // not yet overwritten (chunk -1..-41: zero)
        if (lane == 63 && gl < nl - 1) nb = ld2(psi, tdse_chunk_base(min(c + 1, G.nchunks)));
    } else {
        // own chunks: the previous stage's registers; halo chunks: its successor's first H
        // chunks (LDS); past the grid: zero (the previous stage left those lanes unset)
        const bool on_grid = c < G.nchunks;
#pragma unroll
        for (int i = 0; i < C; i++)
            P[i] = !on_grid ? zero2 : (gl < own_chunks ? R[i] : pl->halo[par ^ 1][gl - own_chunks][i]);
        if (lane == 0 && gl > 0 && c - 1 < G.nchunks) {
            const int q = gl - 1;  // last chunk of the wave before
            nb = q < own_chunks ? pl->bnd[par][wave - 1][1] : pl->halo[par ^ 1][q - own_chunks][C - 1];
        }
        if (lane == 63 && gl < nl - 1 && c + 1 < G.nchunks) {
            const int q = gl + 1;  // first chunk of the wave after
            nb = q < own_chunks ? pl->bnd[par][wave + 1][0] : pl->halo[par ^ 1][q - own_chunks][0];
        }
    }
    __builtin_amdgcn_s_setprio(0);
    const bool own = MODE == kParallel
                         ? gl >= G.halo_lanes && gl < G.halo_lanes + own_chunks && c >= 0 && c < G.nchunks
                         : gl < own_chunks && c < G.nchunks;
    const int j0 = c * C;
    // observables of the state this step starts from (psi_k)
    if (observe) observe_chunk(G, P, T.dp, base, j0, own, acc);
    const double2(&IB)[C] = T.IB;
    const double(&Gv)[C] = T.Gv;
    const double2 Qc = T.Qc, Gc = T.Gc;
    double gL = T.gL, gR = T.gR;
    if (MODE != kParallel && gl == own_chunks - 1) edge_out[0] = P[C - 1];  // the next segment's left neighbour
    // neighbours across the chunk edges: shuffles inside a wave, loads across waves, LDS for
    // the point left of the window, zero past the window's right end
    {
        const double gLd = dpp_z<kWaveShr1>(Gv[C - 1]), gRd = dpp_z<kWaveShl1>(Gv[0]);
        gL = lane == 0 ? gL : gLd;
        gR = lane == 63 ? gR : gRd;
    }
    double2 left = dpp2_z<kWaveShr1>(P[C - 1]), right = dpp2_z<kWaveShl1>(P[0]);
    if (lane == 0) left = (gl > 0 || MODE == kParallel) ? nb : (seg > 0 ? edge_in[0] : zero2);
    if (lane == 63) right = nb;
    // the forward recurrence enters exactly from the previous segment, or from zero H chunks
    // before the own chunks (kParallel: truncated like the backward one)
    const double2 z_in = (MODE != kParallel && seg > 0) ? edge_in[1] : zero2;

    // r' = (M2' - i H') psi
    cn_rhs(P, Gv, left, right, gL, gR, G.k36);

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
        if (MODE != kParallel && gl == own_chunks - 1) edge_out[1] = z;  // the next segment's forward carry
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

    if (!own) return;  // only the segment's own chunks continue

    // length-gauge phase exp(+i F dt x_j) (prop.c:199-204) by rotation from the chunk start
    // (e0), then store (or keep, for the second step of a pair)
    double2 e = e0;
    if constexpr (MODE == kKeep) {
#pragma unroll
        for (int i = 0; i < C; i++) {
            R[i] = cmul(e, P[i]);
            e = cmul(e, rot);
        }
        if (gl < G.halo_lanes) {
#pragma unroll
            for (int i = 0; i < C; i++) pl->halo[par][gl][i] = R[i];
        }
        if (lane == 0) pl->bnd[par][wave][0] = R[0];
        if (lane == 63) pl->bnd[par][wave][1] = R[C - 1];
    } else {
        double2 *dst = MODE == kParallel ? psi_out : psi;
#pragma unroll
        for (int i = 0; i < C; i++) {
            st2(dst, base + 64 * i, cmul(e, P[i]));
            e = cmul(e, rot);
        }
    }
}

// Observables of a whole state in memory (psi, tiled), with the lane-to-chunk assignment
// of the solve: the final state of a partial run, which no later step observes.
__device__ inline void observe_state(const TdseGeom &G, int nseg, int own_chunks, int wave, int lane, int nw,
                                     const double2 *psi, Acc &acc)
{
    for (int s = 0; s < nseg; s++) {
        const int cs = s * own_chunks;
        if (wave >= min(nw, (G.nchunks - cs + 63) >> 6)) continue;
        const int c = cs + wave * 64 + lane;
        const long long base = tdse_chunk_base(min(c, G.nchunks));
        double2 P[C];
        double dp[C];
#pragma unroll
        for (int i = 0; i < C; i++) {
            P[i] = ld2(psi, base + 64 * i);
            dp[i] = ld1(G.dpot, base + 64 * i);
        }
        observe_chunk(G, P, dp, base, c * C, wave * 64 + lane < own_chunks && c < G.nchunks, acc);
    }
}

// Per-wave sums of the observables into part[wave] (lane 0 writes).
__device__ inline void wave_observables(const Acc &acc, int wave, int lane, double (*part)[5])
{
    const double pr = wave_sum(acc.pr), pi = wave_sum(acc.pi), pint = wave_sum(acc.pint);
    const double pos = wave_sum(acc.pos), gp = wave_sum(acc.gp);
    if (lane == 0) {
        part[wave][0] = pr;
        part[wave][1] = pi;
        part[wave][2] = pint;
        part[wave][3] = pos;
        part[wave][4] = gp;
    }
}

// Entry idx of the outputs (prop.c:302-340) from the wave sums, added in wave order; F is
// the field of the step that produced the state.  One thread.
__device__ inline void store_observables(const TdseInst &I, const double (*part)[5], int nw, int idx, double F)
{
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    for (int w = 0; w < nw; w++) {
        s0 += part[w][0];
        s1 += part[w][1];
        s2 += part[w][2];
        s3 += part[w][3];
        s4 += part[w][4];
    }
    I.src[idx] = -s4 + F;
    I.ptot[idx] = s0 * s0 + s1 * s1;
    I.pint[idx] = s2;
    I.expv[idx] = s3;
    if (I.efield) I.efield[idx] = F;
}

// Field of step k after the gating of prop.c:125-134 (F = 0 up to and including the first
// sign change of in_field); `gating` carries the state from step to step.
__device__ inline double gated_field(const TdseInst &I, int k, bool &gating)
{
    const double fk = I.field[k];
    if (gating) {
        if (fk * I.field[k + 1] <= 0.0) gating = false;
        return 0.0;
    }
    return fk;
}

// Phase factors of one step: per-point rotation, the phase at the first point of this lane's
// chunk in segment 0, and the rotation from one segment to the next.
struct StepPhase {
    double2 rot, e_lane, r_seg;
};
__device__ inline StepPhase step_phase(const TdseGeom &G, double F, int wave, int lane, int own_chunks)
{
    StepPhase p;
    const double Fdt = F * G.dt;
    sincos(Fdt * G.dx, &p.rot.y, &p.rot.x);
    const double x = grid_x((wave * 64 + lane) * C, G.dx, G.xmax);
    sincos(Fdt * x, &p.e_lane.y, &p.e_lane.x);
    sincos(Fdt * ((double)(own_chunks * C) * G.dx), &p.r_seg.y, &p.r_seg.x);
    return p;
}

// The persistent propagation loop.  PAIR: whenever two steps k, k+1 follow each other with
// no state save or snapshot between them, they run in one pass over psi: segment s of step
// k is solved and kept in registers, then segment s-1 of step k+1 is solved from those
// registers (its right halo is the first H chunks of segment s of step k+1, handed over in
// LDS) and stored.  psi then streams through HBM once per two steps.
// States saved for partial runs and snapshots are copied from the finished state between
// steps (they are rare), so the segment solve carries no extra stores.
// The XCD this wave runs on (hardware register XCC_ID, bits 3:0).
__device__ inline int xcd_id()
{
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return (int)(v & 15u);
}

// Next instance of the launch for this workgroup (thread 0 only): from the queue of this XCD
// first, then the others in turn; tried (LDS) counts the queues found drained so far.  -1 when
// every queue is drained: a queue is left only after its counter passed its length, so once
// every workgroup has left every queue, each instance was taken exactly once.
__device__ inline int next_instance(const TdseQueues &qs, int *counter, int &tried)
{
    const int q0 = xcd_id() % TDSE_QUEUES;
    for (; tried < TDSE_QUEUES; tried++) {
        const int q = (q0 + tried) % TDSE_QUEUES;
        const int len = qs.off[q + 1] - qs.off[q];
        if (len == 0) continue;
        const int t = atomicAdd(counter + q, 1);
        if (t < len) return qs.off[q] + t;
    }
    return -1;
}

template <bool PAIR, bool STASH = false>
__device__ void propagate_loop(const TdseGeom *__restrict__ geoms, const TdseInst *__restrict__ insts,
                               const int *__restrict__ order, const TdseQueues &queues, int *counter, double2 *work,
                               long long work_stride, PairLds *pl, double2 *stash = nullptr)
{
    __shared__ double part[2][TDSE_MAX_WAVES][5];
    __shared__ double2 slot_f[TDSE_MAX_WAVES][2], slot_b[TDSE_MAX_WAVES][2];
    __shared__ double2 edge[2][2], edge_b[2][2];
    __shared__ int cur, tried;  // tried: thread 0's count of drained queues
    const int nw = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    double2 *const slot = work + (long long)blockIdx.x * work_stride;
    double2 R0[C], R1[C];  // pair passes: step-k+1 state of the last two segments solved
    if (threadIdx.x == 0) tried = 0;

    for (;;) {
        if (threadIdx.x == 0) cur = next_instance(queues, counter, tried);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(cur);  // uniform: keeps the instance in SGPRs
        __syncthreads();
        if (t < 0) break;
        const TdseInst &I = insts[order[t]];
        const TdseGeom &G = geoms[I.geom];
        const long long npad = G.npad;
        const int own_chunks = nw * 64 - G.halo_lanes;  // host guarantees >= 64
        const int nseg = (G.nchunks + own_chunks - 1) / own_chunks;

        // start state: psi0 (t = tmin) or a saved state; padding chunks are zero in both.  An
        // in-place instance steps the caller's state vector itself.
        double2 *const buf = I.inplace ? const_cast<double2 *>(I.psi_init) : slot;
        if (!I.inplace) {
            const double2 *init = I.psi_init ? I.psi_init : G.psi0;
            for (long long q = threadIdx.x; q < npad; q += blockDim.x) buf[q] = init[q];
        }
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
        double F_prev = 0.0;  // field of the previous step
        // a state save or a snapshot is due after step k
        auto save_due = [&](int k) { return I.save && !I.inplace && (k - k_start + 1) % I.save_every == 0; };
        auto snap_due = [&](int k) { return I.snap && k % I.snap_every == 0 && i_snap < I.n_snap - 1; };
        for (int k = k_start; k < k_stop;) {
            const bool pair = PAIR && k + 1 < k_stop && !save_due(k) && !snap_due(k);
            if (pair) {
                const double F0 = gated_field(I, k, gating);
                const double F1 = gated_field(I, k + 1, gating);
                const StepPhase ph0 = step_phase(G, F0, wave, lane, own_chunks);
                const StepPhase ph1 = step_phase(G, F1, wave, lane, own_chunks);
                const bool observe0 = k > k_start;
                Acc acc0 = {0.0, 0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0, 0.0};
                double2 e0 = ph0.e_lane, e1 = ph1.e_lane;
                // segment s of step k into R(s&1), then segment s-1 of step k+1 from R((s-1)&1)
                // (STASH: segment s's 1/beta in stash buffer s & 1 from the first stage to the second)
                const int stash_len = 64 * nw * C;
                for (int s = 0; s <= nseg; s++) {
                    if (s < nseg) {
                        double2 *st = STASH ? stash + (s & 1) * stash_len : nullptr;
                        if (s & 1)
                            solve_segment<kKeep, STASH>(G, s, own_chunks, wave, lane, nw, buf, e0, ph0.rot, observe0,
                                                        acc0, slot_f, slot_b, edge[s & 1], edge[(s + 1) & 1], R1, pl,
                                                        nullptr, st);
                        else
                            solve_segment<kKeep, STASH>(G, s, own_chunks, wave, lane, nw, buf, e0, ph0.rot, observe0,
                                                        acc0, slot_f, slot_b, edge[s & 1], edge[(s + 1) & 1], R0, pl,
                                                        nullptr, st);
                        e0 = cmul(e0, ph0.r_seg);
                    }
                    lds_barrier();  // the halo and wave-edge points of segment s are in LDS
                    if (s > 0) {
                        const int sb = s - 1;
                        double2 *st = STASH ? stash + (sb & 1) * stash_len : nullptr;
                        if (sb & 1)
                            solve_segment<kFromRegs, STASH>(G, sb, own_chunks, wave, lane, nw, buf, e1, ph1.rot, true,
                                                            acc1, slot_f, slot_b, edge_b[sb & 1],
                                                            edge_b[(sb + 1) & 1], R1, pl, nullptr, st);
                        else
                            solve_segment<kFromRegs, STASH>(G, sb, own_chunks, wave, lane, nw, buf, e1, ph1.rot, true,
                                                            acc1, slot_f, slot_b, edge_b[sb & 1],
                                                            edge_b[(sb + 1) & 1], R0, pl, nullptr, st);
                        e1 = cmul(e1, ph1.r_seg);
                    }
                }
                wave_observables(acc0, wave, lane, part[0]);
                wave_observables(acc1, wave, lane, part[1]);
                __syncthreads();  // also: psi of step k + 2 is complete in memory
                if (threadIdx.x == 0 && I.src) {
                    if (observe0) store_observables(I, part[0], nw, k, F_prev);
                    store_observables(I, part[1], nw, k + 1, F0);
                }
                F_prev = F1;
                k += 1;  // the save/snapshot test below is for step k + 1
            } else {
                const double F = gated_field(I, k, gating);
                const StepPhase ph = step_phase(G, F, wave, lane, own_chunks);
                // observables of psi_k (entry k), except for the run's start state, whose entry is
                // the initial values (k = 0) or was written by the run that ended there
                const bool observe = k > k_start;
                Acc acc = {0.0, 0.0, 0.0, 0.0, 0.0};
                double2 e0 = ph.e_lane;
                for (int s = 0; s < nseg; s++) {
                    solve_segment<kSingle>(G, s, own_chunks, wave, lane, nw, buf, e0, ph.rot, observe, acc, slot_f,
                                           slot_b, edge[s & 1], edge[(s + 1) & 1], R0, pl);
                    e0 = cmul(e0, ph.r_seg);
                }
                wave_observables(acc, wave, lane, part[k & 1]);
                __syncthreads();  // also: psi of step k + 1 is complete in memory
                if (threadIdx.x == 0 && I.src && observe) store_observables(I, part[k & 1], nw, k, F_prev);
                F_prev = F;
            }
            // state after step k: saves (partial runs) and snapshots (prop.c:248-255: row i_snap
            // gets psi after step k when k % every == 0)
            double2 *save_to = save_due(k) ? I.save + (long long)((k - k_start + 1) / I.save_every - 1) * npad : nullptr;
            double *snap_row = nullptr;
            if (snap_due(k)) {
                snap_row = I.snap + (long long)i_snap * 2 * G.nx;
                i_snap++;
            }
            if (save_to || snap_row) {
                if (save_to)
                    for (long long q = threadIdx.x; q < npad; q += blockDim.x) save_to[q] = buf[q];
                if (snap_row)
                    for (int j = threadIdx.x; j < G.nx; j += blockDim.x) {
                        const double2 v = buf[tdse_index(j)];
                        snap_row[2 * j] = v.x;
                        snap_row[2 * j + 1] = v.y;
                    }
                __syncthreads();  // the copies have read psi before the next step overwrites it
            }
            k++;
        }
        // the final state of a partial run: entry k_stop (a full run has no entry n_steps)
        if (k_stop > k_start && k_stop < I.n_steps) {
            Acc acc = {0.0, 0.0, 0.0, 0.0, 0.0};
            observe_state(G, nseg, own_chunks, wave, lane, nw, buf, acc);
            wave_observables(acc, wave, lane, part[k_stop & 1]);
            __syncthreads();
            if (threadIdx.x == 0 && I.src) store_observables(I, part[k_stop & 1], nw, k_stop, F_prev);
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

// Entry idx of the outputs from the per-segment partial sums of a step-parallel run, added in
// segment order; F is the gated field of the step that produced the state.  One thread.
__device__ inline void store_partials(const TdseInst &I, const double *parts, int nseg, int idx, double F)
{
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    for (int q = 0; q < nseg; q++) {
        s0 += parts[5 * q + 0];
        s1 += parts[5 * q + 1];
        s2 += parts[5 * q + 2];
        s3 += parts[5 * q + 3];
        s4 += parts[5 * q + 4];
    }
    I.src[idx] = -s4 + F;
    I.ptot[idx] = s0 * s0 + s1 * s1;
    I.pint[idx] = s2;
    I.expv[idx] = s3;
    if (I.efield) I.efield[idx] = F;
}

} // namespace

// One time step per pass over psi (168-VGPR budget, 12 waves per CU).
__global__ void __launch_bounds__(64 * TDSE_MAX_WAVES, TDSE_WAVES_PER_SIMD)
tdse_propagate_kernel(const TdseGeom *__restrict__ geoms, const TdseInst *__restrict__ insts,
                      const int *__restrict__ order, TdseQueues queues, int *counter, double2 *work, long long work_stride)
{
    propagate_loop<false>(geoms, insts, order, queues, counter, work, work_stride, nullptr);
}

// Two time steps per pass wherever no save or snapshot falls between them.
__global__ void __launch_bounds__(64 * TDSE_MAX_WAVES, TDSE_PAIR_WAVES_PER_SIMD)
tdse_propagate_pair_kernel(const TdseGeom *__restrict__ geoms, const TdseInst *__restrict__ insts,
                           const int *__restrict__ order, TdseQueues queues, int *counter, double2 *work,
                           long long work_stride)
{
    __shared__ PairLds pl;
    propagate_loop<true>(geoms, insts, order, queues, counter, work, work_stride, &pl);
}

// The same, with each segment's 1/beta kept in LDS between the two stages of a pass (dynamic
// LDS of 2 * 64 * waves * 8 complex doubles): the second stage does not fetch it again.
__global__ void __launch_bounds__(64 * TDSE_MAX_WAVES, TDSE_PAIR_WAVES_PER_SIMD)
tdse_propagate_pair_stash_kernel(const TdseGeom *__restrict__ geoms, const TdseInst *__restrict__ insts,
                                 const int *__restrict__ order, TdseQueues queues, int *counter, double2 *work,
                                 long long work_stride)
{
    __shared__ PairLds pl;
    extern __shared__ double2 stash[];
    propagate_loop<true, true>(geoms, insts, order, queues, counter, work, work_stride, &pl, stash);
}

// One time step k of instance blockIdx.y, split over independent segments (blockIdx.x) whose
// windows carry a halo on both sides: psi_k from psi_in, psi_{k+1} to psi_out (the caller
// alternates the two).  The field is the gated one (I.field[k] = F_k).  The observables of
// psi_k (k >= 1) go to partial[k & 1][instance][segment]; the next launch sums those of
// entry k - 1 in segment order (tdse_parallel_finish_kernel the last).  For one instance
// alone this runs a step in the time of one segment instead of all of them in turn.
__global__ void __launch_bounds__(64 * TDSE_MAX_WAVES, TDSE_WAVES_PER_SIMD)
tdse_parallel_step_kernel(const TdseGeom *__restrict__ geom, const TdseInst *__restrict__ insts, int k,
                          const double2 *psi_in, double2 *psi_out, long long state_stride, double *partial, int nseg)
{
    __shared__ double part[TDSE_MAX_WAVES][5];
    __shared__ double2 slot_f[TDSE_MAX_WAVES][2], slot_b[TDSE_MAX_WAVES][2];
    const int nw = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int inst = blockIdx.y, seg = blockIdx.x, n_inst = gridDim.y;
    const TdseInst &I = insts[inst];
    const TdseGeom &G = *geom;
    const int own_chunks = nw * 64 - 2 * G.halo_lanes;  // host guarantees >= 64
    if (seg == 0 && threadIdx.x == 0 && I.src) {
        if (k == 0) {
            I.src[0] = 0.0;
            I.ptot[0] = 1.0;
            I.pint[0] = G.pint0;
            I.expv[0] = 0.0;
            if (I.efield) I.efield[0] = 0.0;
        } else if (k >= 2) {
            store_partials(I, partial + ((long long)((k - 1) & 1) * n_inst + inst) * nseg * 5, nseg, k - 1,
                           I.field[k - 2]);
        }
    }
    const double F = I.field[k];
    const double Fdt = F * G.dt;
    double2 rot, e0;
    sincos(Fdt * G.dx, &rot.y, &rot.x);
    const int c = seg * own_chunks - G.halo_lanes + wave * 64 + lane;  // this lane's chunk
    sincos(Fdt * grid_x(c * C, G.dx, G.xmax), &e0.y, &e0.x);
    Acc acc = {0.0, 0.0, 0.0, 0.0, 0.0};
    double2 R[C];  // unused by this mode
    solve_segment<kParallel>(G, seg, own_chunks, wave, lane, nw,
                             const_cast<double2 *>(psi_in) + (long long)inst * state_stride, e0, rot, k >= 1, acc,
                             slot_f, slot_b, nullptr, nullptr, R, nullptr,
                             psi_out + (long long)inst * state_stride);
    wave_observables(acc, wave, lane, part);
    __syncthreads();
    if (threadIdx.x == 0 && k >= 1) {
        double *dst = partial + (((long long)(k & 1) * n_inst + inst) * nseg + seg) * 5;
        for (int q = 0; q < 5; q++) {
            double sum = 0.0;
            for (int w = 0; w < nw; w++) sum += part[w][q];
            dst[q] = sum;
        }
    }
}

// The last observed entry (n_steps - 1) of a step-parallel run: one thread per instance.
__global__ void tdse_parallel_finish_kernel(const TdseInst *insts, int n_inst, const double *partial, int nseg)
{
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= n_inst) return;
    const TdseInst &I = insts[inst];
    const int last = I.n_steps - 1;
    if (!I.src || last < 1) return;
    store_partials(I, partial + ((long long)(last & 1) * n_inst + inst) * nseg * 5, nseg, last,
                   last >= 1 ? I.field[last - 1] : 0.0);
}

// The gated field of prop.c:125-134: F_k = 0 up to and including the first k with
// f_k f_{k+1} <= 0, f_k after it; n samples (the last is read only as f_{k+1}).  One block.
__global__ void tdse_gate_kernel(const double *f, int n, double *F)
{
    __shared__ int kg;
    if (threadIdx.x == 0) {
        int g = n;
        for (int k = 0; k + 1 < n; k++)
            if (f[k] * f[k + 1] <= 0.0) {
                g = k;
                break;
            }
        kg = g;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) F[k] = k <= kg ? 0.0 : f[k];
}
