// Discrete adjoint of the batched Crank-Nicolson propagation that rebuilds the states it
// needs instead of reading them (no counterpart in the reference).
//
// The forward step psi_{k+1} = D(F_k) B'^-1 A' psi_k (tdse_propagate.hip) is invertible, and
// since A' = conj(B') its inverse is the forward step on the conjugate state:
//     conj(psi_{k-1}) = B'^-1 A' phi,   phi = D(F_{k-1}) conj(psi_k) = conj(e psi_k),
//     e = exp(-i F_{k-1} dt x).
// One sweep k = k_hi .. 1 therefore carries both psi_k and the adjoint state kappa_k
// (tdse_adjoint.hip):
//     lambda_k    = A'^H kappa_k - (dL/ds_k) V' psi_k
//     dL/dF_{k-1} = dL/ds_k - 2 dt Im sum_j x_j conj(lambda_{k,j}) psi_{k,j}
//     kappa_{k-1} = B'^-H (e lambda_k)
// with two segmented tridiagonal solves per step (the adjoint's and the forward's
// recurrences, each chained exactly from segment to segment with a truncated right halo).
// The reconstruction is re-anchored at every stored checkpoint psi_{mK}, so rounding drifts
// over at most K inverse steps (the step is norm-preserving: errors do not grow).
// HBM traffic per grid-point-step: psi and kappa read and written, 64 bytes, and no stored
// states besides the checkpoints.
#include "tdse_internal.h"
#include "tdse_wave.h"

namespace {

using namespace tdse;

__device__ inline double2 conj2(double2 v) { return make_double2(v.x, -v.y); }

// One segment of reverse step k with all waves of the workgroup; kap and psi hold kappa_k and
// psi_k on entry and kappa_{k-1}, psi_{k-1} on the segment's own chunks on exit (step_psi
// false: psi is left alone).  e0 = e at the first point of this lane's chunk.
// edge_in / edge_out: {old kappa, y, phi, z} at the last own point of the previous / this
// segment.  park: this wave's LDS (C x 64 complex) holding phi during the kappa solve.
__device__ void reverse_segment(const TdseGeom &G, int seg, int own_chunks, int wave, int lane, int nw,
                                double2 *__restrict__ psi, double2 *__restrict__ kap, bool step_psi, double gobs,
                                double2 e0, double2 rot_conj, double &gacc, double2 (*slot_f)[2],
                                double2 (*slot_b)[2], const double2 *edge_in, double2 *edge_out, double2 *park)
{
    const int gl = wave * 64 + lane;
    const int cs = seg * own_chunks;
    const int nwl = min(nw, (G.nchunks - cs + 63) >> 6);  // waves with a lane on the grid
    if (wave >= nwl) {
        lds_barrier();  // the barriers of the two solves
        lds_barrier();
        if (step_psi) {
            lds_barrier();
            lds_barrier();
        }
        return;
    }
    const int nl = nwl * 64;
    const int c = cs + gl;
    const int cl = min(c, G.nchunks);  // lanes past the grid read the zero padding chunk
    const long long base = tdse_chunk_base(cl);
    const double2 zero2 = make_double2(0.0, 0.0);

    double2 Kp[C], S[C], IB[C];
    double Gv[C], dp[C];
#pragma unroll
    for (int i = 0; i < C; i++) {
        Kp[i] = ld2(kap, base + 64 * i);
        S[i] = ld2(psi, base + 64 * i);
        Gv[i] = ld1(G.g, base + 64 * i);
        IB[i] = ld2(G.ib, base + 64 * i);
        dp[i] = ld1(G.dpot, base + 64 * i);
    }
    // across the wave's edges: lane 0 the point left of its chunk, lane 63 the point right
    double2 kl = zero2, kr = zero2, sl = zero2, sr = zero2, ibL = zero2;
    double gL = 0.0, gR = 0.0;
    if (lane == 0) {
        const long long nb = tdse_chunk_base(min(c - 1, G.nchunks)) + 64 * (C - 1);  // chunk -1: zero
        if (gl > 0) {  // not yet overwritten
            kl = ld2(kap, nb);
            sl = ld2(psi, nb);
        }
        ibL = ld2(G.ib, nb);
        gL = ld1(G.g, nb);
    }
    if (lane == 63 && gl < nl - 1) {
        const long long nb = tdse_chunk_base(min(c + 1, G.nchunks));
        kr = ld2(kap, nb);
        sr = ld2(psi, nb);
        gR = ld1(G.g, nb);
    }
    const double2 Qa = ld2(G.adq, cl + 64), Ga = ld2(G.adg, cl + 64);
    const double2 Qf = ld2(G.chq, cl + 64), Gf = ld2(G.chg, cl + 64);
    {
        const double2 ibd = dpp2_z<kWaveShr1>(IB[C - 1]);  // 1/beta left of the chunk, inside the wave
        ibL = lane == 0 ? ibL : ibd;
    }
    const bool own = gl < own_chunks && c < G.nchunks;
    if (gl == own_chunks - 1) edge_out[0] = Kp[C - 1];

    // lambda = A'^H kappa - gobs V' psi, the gradient term, mu = e lambda (into Kp) and
    // phi = conj(e psi) (into S)
    double2 e_end;  // e at the point right of the chunk
    {
        double2 pm = dpp2_z<kWaveShr1>(Kp[C - 1]), kright = dpp2_z<kWaveShl1>(Kp[0]);
        if (lane == 0) pm = gl > 0 ? kl : (seg > 0 ? edge_in[0] : zero2);
        if (lane == 63) kright = kr;
        const int j0 = c * C;
        const double K36 = G.k36;
        double2 e = e0;
        double gsum = 0.0;
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 k0 = Kp[i];
            const double2 kp = (i + 1 < C) ? Kp[i + 1] : kright;
            const double g = Gv[i];
            const double d = fma(10.0, g, K36);
            const double sre = pm.x + kp.x, sim = pm.y + kp.y;
            // M2' kappa + i (g (k_{j-1} + k_{j+1}) + d k_j)
            double lr = fma(10.0, k0.x, sre) - fma(g, sim, d * k0.y);
            double li = fma(10.0, k0.y, sim) + fma(g, sre, d * k0.x);
            const double2 ps = S[i];
            const double w = gobs * dp[i];
            lr = fma(-w, ps.x, lr);
            li = fma(-w, ps.y, li);
            const double xj = grid_x(j0 + i, G.dx, G.xmax);
            gsum = fma(xj, fma(lr, ps.y, -li * ps.x), gsum);
            Kp[i] = cmul(e, make_double2(lr, li));
            S[i] = conj2(cmul(e, ps));
            e = cmul(e, rot_conj);
            pm = k0;
        }
        if (own) gacc += gsum;
        e_end = e;
    }
    if (gl == own_chunks - 1) edge_out[2] = S[C - 1];
    // phi across the chunk edges for the psi solve; phi itself waits in LDS
    double2 left = dpp2_z<kWaveShr1>(S[C - 1]), right = dpp2_z<kWaveShl1>(S[0]);
    if (lane == 0) left = gl > 0 ? conj2(cmul(cmul(e0, conj2(rot_conj)), sl)) : (seg > 0 ? edge_in[2] : zero2);
    if (lane == 63) right = conj2(cmul(e_end, sr));
    if (step_psi) {
#pragma unroll
        for (int i = 0; i < C; i++) park[64 * i + lane] = S[i];
    }

    // kappa_{k-1} = B'^-H mu: forward y_j = mu_j + conj(h_{j-1}) y_{j-1} (exact carry from the
    // previous segment), backward nu_j = conj(ib_j) y_j + conj(p_j) nu_{j+1} (truncated)
    {
        double2 y = zero2;
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 b = (i > 0) ? IB[i - 1] : ibL;
            const double g = Gv[i];
            const double hr = -(b.x - g * b.y), hi = b.y + g * b.x;  // conj(h_{j-1})
            y = make_double2(Kp[i].x + hr * y.x - hi * y.y, Kp[i].y + hr * y.y + hi * y.x);
        }
        double2 Q = Qa;
        wave_scan<true>(Q, y, lane);
        y = block_carry<true>(Q, y, wave, lane, nwl, slot_f, seg > 0 ? edge_in[1] : zero2);
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 b = (i > 0) ? IB[i - 1] : ibL;
            const double g = Gv[i];
            const double hr = -(b.x - g * b.y), hi = b.y + g * b.x;
            y = make_double2(Kp[i].x + hr * y.x - hi * y.y, Kp[i].y + hr * y.y + hi * y.x);
            Kp[i] = y;
        }
        if (gl == own_chunks - 1) edge_out[1] = y;
    }
    {
        double2 u = zero2;
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double2 ib = IB[i];
            const double g = Gv[i];
            const double pr = -(ib.x - g * ib.y), pi = ib.y + g * ib.x;  // conj(p_j)
            const double tr = ib.x * Kp[i].x + ib.y * Kp[i].y, ti = ib.x * Kp[i].y - ib.y * Kp[i].x;
            u = make_double2(tr + pr * u.x - pi * u.y, ti + pr * u.y + pi * u.x);
        }
        double2 Gt = Ga;
        wave_scan<false>(Gt, u, lane);
        u = block_carry<false>(Gt, u, wave, lane, nwl, slot_b);
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double2 ib = IB[i];
            const double g = Gv[i];
            const double pr = -(ib.x - g * ib.y), pi = ib.y + g * ib.x;
            const double tr = ib.x * Kp[i].x + ib.y * Kp[i].y, ti = ib.x * Kp[i].y - ib.y * Kp[i].x;
            u = make_double2(tr + pr * u.x - pi * u.y, ti + pr * u.y + pi * u.x);
            Kp[i] = u;
        }
    }
    if (own) {
#pragma unroll
        for (int i = 0; i < C; i++) st2(kap, base + 64 * i, Kp[i]);
    }
    if (!step_psi) return;

    // psi_{k-1} = conj(B'^-1 A' phi): the forward step's recurrences (tdse_propagate.hip)
#pragma unroll
    for (int i = 0; i < C; i++) S[i] = park[64 * i + lane];  // this lane's own writes
    {
        const double gLd = dpp_z<kWaveShr1>(Gv[C - 1]), gRd = dpp_z<kWaveShl1>(Gv[0]);
        gL = lane == 0 ? gL : gLd;
        gR = lane == 63 ? gR : gRd;
    }
    cn_rhs(S, Gv, left, right, gL, gR, G.k36);
    {
        double2 z = zero2;
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double gm = (i > 0) ? Gv[i - 1] : gL;
            const double tr = fma(gm, z.y, S[i].x - z.x);
            const double ti = fma(-gm, z.x, S[i].y - z.y);
            z = cmul(make_double2(tr, ti), IB[i]);
        }
        double2 Q = Qf;
        wave_scan<true>(Q, z, lane);
        z = block_carry<true>(Q, z, wave, lane, nwl, slot_f, seg > 0 ? edge_in[3] : zero2);
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double gm = (i > 0) ? Gv[i - 1] : gL;
            const double tr = fma(gm, z.y, S[i].x - z.x);
            const double ti = fma(-gm, z.x, S[i].y - z.y);
            z = cmul(make_double2(tr, ti), IB[i]);
            S[i] = z;
        }
        if (gl == own_chunks - 1) edge_out[3] = z;
    }
    {
        double2 u = zero2;
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double gp = (i + 1 < C) ? Gv[i + 1] : gR;
            const double hr = fma(gp, IB[i].y, -IB[i].x), hi = -fma(gp, IB[i].x, IB[i].y);
            u = make_double2(S[i].x + hr * u.x - hi * u.y, S[i].y + hr * u.y + hi * u.x);
        }
        double2 Gt = Gf;
        wave_scan<false>(Gt, u, lane);
        u = block_carry<false>(Gt, u, wave, lane, nwl, slot_b);
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double gp = (i + 1 < C) ? Gv[i + 1] : gR;
            const double hr = fma(gp, IB[i].y, -IB[i].x), hi = -fma(gp, IB[i].x, IB[i].y);
            u = make_double2(S[i].x + hr * u.x - hi * u.y, S[i].y + hr * u.y + hi * u.x);
            S[i] = u;
        }
    }
    if (own) {
#pragma unroll
        for (int i = 0; i < C; i++) st2(psi, base + 64 * i, conj2(S[i]));
    }
}

} // namespace

// One workgroup per instance at a time (atomic queue, largest first); per workgroup two
// state buffers (psi, kappa) of work_stride complex values in work.
__global__ void __launch_bounds__(64 * TDSE_MAX_WAVES, TDSE_REV_WAVES_PER_SIMD)
tdse_reverse_adjoint_kernel(const TdseGeom *__restrict__ geoms, const TdseRevInst *__restrict__ insts,
                            const int *__restrict__ order, int n_inst, int *counter, double2 *work,
                            long long work_stride)
{
    __shared__ double part[2][TDSE_MAX_WAVES];
    __shared__ double2 slot_f[TDSE_MAX_WAVES][2], slot_b[TDSE_MAX_WAVES][2];
    __shared__ double2 edge[2][4];
    __shared__ int cur;
    extern __shared__ double2 tdse_park[];  // dynamic: C x 64 complex per wave
    const int nw = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    double2 *park = tdse_park + wave * (C * 64);
    double2 *psi = work + 2 * (long long)blockIdx.x * work_stride;
    double2 *kap = psi + work_stride;

    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(cur);
        __syncthreads();
        if (t >= n_inst) break;
        const TdseRevInst &I = insts[order[t]];
        const TdseGeom &G = geoms[I.geom];
        const long long npad = G.npad;
        const int own_chunks = nw * 64 - G.halo_lanes;
        const int nseg = (G.nchunks + own_chunks - 1) / own_chunks;
        for (long long q = threadIdx.x; q < npad; q += blockDim.x) {
            psi[q] = I.psi_hi[q];
            kap[q] = make_double2(0.0, 0.0);  // kappa_{k_hi}: no observation after k_hi
        }
        __syncthreads();
        for (int k = I.k_hi; k >= 1; k--) {
            // re-anchor psi_k at a checkpoint (slot m holds psi_{(m+1) K})
            if (k < I.k_hi && I.ckpt && k % I.every == 0 && k / I.every <= I.n_ckpt) {
                const double2 *src = I.ckpt + (long long)(k / I.every - 1) * npad;
                for (long long q = threadIdx.x; q < npad; q += blockDim.x) psi[q] = src[q];
                __syncthreads();
            }
            const double F = I.field[k - 1];
            const double Fdt = F * G.dt;
            double rs, rc;
            sincos(-Fdt * G.dx, &rs, &rc);
            const double2 rot_conj = make_double2(rc, rs);
            double2 e_lane, r_seg;  // e at this lane's first chunk in segment 0; segment-to-segment rotation
            {
                const double x = grid_x((wave * 64 + lane) * C, G.dx, G.xmax);
                sincos(-Fdt * x, &e_lane.y, &e_lane.x);
                sincos(-Fdt * ((double)(own_chunks * C) * G.dx), &r_seg.y, &r_seg.x);
            }
            const double gobs = I.gobs[k];
            const bool step_psi = k > 1;  // psi_0 is not needed
            double gacc = 0.0;
            double2 e0 = e_lane;
            for (int s = 0; s < nseg; s++) {
                reverse_segment(G, s, own_chunks, wave, lane, nw, psi, kap, step_psi, gobs, e0, rot_conj, gacc,
                                slot_f, slot_b, edge[s & 1], edge[(s + 1) & 1], park);
                e0 = cmul(e0, r_seg);
            }
            const double gw = wave_sum(gacc);
            if (lane == 0) part[k & 1][wave] = gw;
            __syncthreads();  // also: psi_{k-1} and kappa_{k-1} are complete in memory
            if (threadIdx.x == 0) {
                double tot = 0.0;
                for (int w = 0; w < nw; w++) tot += part[k & 1][w];
                I.grad[k - 1] = gobs - 2.0 * G.dt * tot;
            }
        }
        __syncthreads();
    }
}
