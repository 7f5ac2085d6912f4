// Discrete adjoint of the batched Crank-Nicolson propagation (no counterpart in the
// reference): gradients of any loss of the dipole series with respect to the gated field.
//
// Forward model (tdse_propagate.hip, 1DTDSE/sources/prop.c:115-245):
//     psi_{k+1} = D(F_k) B'^-1 A' psi_k,  D(F) = diag exp(+i F dt x_j),
//     s_k = -sum_j |psi_{k,j}|^2 V'_j + F_{k-1}           (sourceterm)
// With lambda_k = dL/d conj(psi_k) and the carried state kappa_k = B'^-H D(F_k)^H lambda_{k+1}:
//     lambda_k   = A'^H kappa_k - (dL/ds_k) V' psi_k
//     dL/dF_{k-1} = dL/ds_k - 2 dt Im sum_j x_j conj(lambda_{k,j}) psi_{k,j}
//     kappa_{k-1} = B'^-H ( exp(-i F_{k-1} dt x) lambda_k )
// B'^-H reuses the forward pivots: B' = L U (Thomas), B'^H = U^H L^H, so the solve is
//     forward   y_j  = mu_j + conj(h_{j-1}) y_{j-1},     h_{j-1} = -(1 + i g_j) ib_{j-1}
//     backward  nu_j = conj(ib_j) y_j + conj(p_j) nu_{j+1}, p_j  = -(1 + i g_j) ib_j
// two first-order recurrences with the forward's decay rate, evaluated with the same
// segmented scan as the forward solve: y chained exactly from segment to segment, nu
// truncated H chunks right of each segment's own chunks.  A'^H = M2' + i H'^T with
// (H'^T v)_j = g_j (v_{j-1} + v_{j+1}) + d_j v_j.
#include "tdse_internal.h"
#include "tdse_wave.h"

namespace {

using namespace tdse;

// edge_in / edge_out: {old kappa, y} at the last own point of the previous / this segment.
__device__ void adjoint_segment(const TdseGeom &G, int seg, int own_chunks, int wave, int lane, int nw,
                                double2 *__restrict__ kap, const double2 *__restrict__ psik, double gobs, double2 e0,
                                double2 rot_conj, double &gacc, double2 (*slot_f)[2], double2 (*slot_b)[2],
                                const double2 *edge_in, double2 *edge_out)
{
    const int gl = wave * 64 + lane;
    const int c0 = seg * own_chunks;
    const int nwl = min(nw, (G.nchunks - c0 + 63) >> 6);  // waves with a lane on the grid
    if (wave >= nwl) {
        lds_barrier();  // the two barriers of block_carry
        lds_barrier();
        return;
    }
    const int nl = nwl * 64;
    const int c = c0 + gl;
    const int cl = min(c, G.nchunks);
    const long long base = tdse_chunk_base(cl);

    double2 P[C], IB[C];
    double Gv[C];
#pragma unroll
    for (int i = 0; i < C; i++) {
        P[i] = ld2(kap, base + 64 * i);
        Gv[i] = ld1(G.g, base + 64 * i);
        IB[i] = ld2(G.ib, base + 64 * i);
    }
    if (gl == own_chunks - 1) edge_out[0] = P[C - 1];
    const double2 zero2 = make_double2(0.0, 0.0);
    double2 left = dpp2_z<kWaveShr1>(P[C - 1]), right = dpp2_z<kWaveShl1>(P[0]);
    double2 ibL = dpp2_z<kWaveShr1>(IB[C - 1]);
    if (lane == 0) {
        const long long nb = tdse_chunk_base(min(c - 1, G.nchunks)) + 64 * (C - 1);  // chunk -1: zero padding
        left = gl > 0 ? ld2(kap, nb) : (seg > 0 ? edge_in[0] : zero2);  // gl > 0: not yet overwritten
        ibL = ld2(G.ib, nb);
    }
    if (lane == 63) {
        const long long nb = tdse_chunk_base(min(c + 1, G.nchunks));
        right = gl < nl - 1 ? ld2(kap, nb) : zero2;
    }
    const double2 y_in = seg > 0 ? edge_in[1] : zero2;
    const bool own = gl < own_chunks && c < G.nchunks;

    // lambda = A'^H kappa + source, the gradient term, then the conjugate phase
    {
        const int j0 = c * C;
        double2 e = e0;  // exp(-i F dt x) at the chunk's first point
        const double K = G.k36;
        double2 pm = left;
        double gsum = 0.0;
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 k0 = P[i];
            const double2 kp = (i + 1 < C) ? P[i + 1] : right;
            const double g = Gv[i];
            const double d = fma(10.0, g, K);
            const double sr = pm.x + kp.x, si = pm.y + kp.y;
            // M2' kappa + i (g (k_{j-1} + k_{j+1}) + d k_j)
            double lr = fma(10.0, k0.x, sr) - fma(g, si, d * k0.y);
            double li = fma(10.0, k0.y, si) + fma(g, sr, d * k0.x);
            const double2 ps = ld2(psik, base + 64 * i);
            const double w = gobs * ld1(G.dpot, base + 64 * i);
            lr = fma(-w, ps.x, lr);
            li = fma(-w, ps.y, li);
            const double xj = grid_x(j0 + i, G.dx, G.xmax);
            gsum = fma(xj, fma(lr, ps.y, -li * ps.x), gsum);
            P[i] = cmul(e, make_double2(lr, li));
            e = cmul(e, rot_conj);
            pm = k0;
        }
        if (own) gacc += gsum;
    }

    // forward recurrence y_j = mu_j + conj(h_{j-1}) y_{j-1}
    {
        double2 y = zero2;
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 b = (i > 0) ? IB[i - 1] : ibL;
            const double g = Gv[i];
            const double hr = -(b.x - g * b.y), hi = b.y + g * b.x;  // conj(h_{j-1})
            y = make_double2(P[i].x + hr * y.x - hi * y.y, P[i].y + hr * y.y + hi * y.x);
        }
        double2 Q = ld2(G.adq, cl + 64);
        wave_scan<true>(Q, y, lane);
        y = block_carry<true>(Q, y, wave, lane, nwl, slot_f, y_in);
#pragma unroll
        for (int i = 0; i < C; i++) {
            const double2 b = (i > 0) ? IB[i - 1] : ibL;
            const double g = Gv[i];
            const double hr = -(b.x - g * b.y), hi = b.y + g * b.x;
            y = make_double2(P[i].x + hr * y.x - hi * y.y, P[i].y + hr * y.y + hi * y.x);
            P[i] = y;
        }
        if (gl == own_chunks - 1) edge_out[1] = y;
    }

    // backward recurrence nu_j = conj(ib_j) y_j + conj(p_j) nu_{j+1}
    {
        double2 u = zero2;
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double2 ib = IB[i];
            const double g = Gv[i];
            const double pr = -(ib.x - g * ib.y), pi = ib.y + g * ib.x;  // conj(p_j)
            const double tr = ib.x * P[i].x + ib.y * P[i].y, ti = ib.x * P[i].y - ib.y * P[i].x;  // conj(ib) y
            u = make_double2(tr + pr * u.x - pi * u.y, ti + pr * u.y + pi * u.x);
        }
        double2 Gt = ld2(G.adg, cl + 64);
        wave_scan<false>(Gt, u, lane);
        u = block_carry<false>(Gt, u, wave, lane, nwl, slot_b);
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            const double2 ib = IB[i];
            const double g = Gv[i];
            const double pr = -(ib.x - g * ib.y), pi = ib.y + g * ib.x;
            const double tr = ib.x * P[i].x + ib.y * P[i].y, ti = ib.x * P[i].y - ib.y * P[i].x;
            u = make_double2(tr + pr * u.x - pi * u.y, ti + pr * u.y + pi * u.x);
            P[i] = u;
        }
    }

    if (!own) return;
#pragma unroll
    for (int i = 0; i < C; i++) st2(kap, base + 64 * i, P[i]);
}

} // namespace

// One workgroup per instance at a time (atomic queue), all waves on one segment.
__global__ void __launch_bounds__(64 * TDSE_MAX_WAVES)
tdse_adjoint_kernel(const TdseGeom *__restrict__ geoms, const TdseAdjInst *__restrict__ insts,
                    const int *__restrict__ order, int n_inst, int *counter)
{
    __shared__ double part[2][TDSE_MAX_WAVES];
    __shared__ double2 slot_f[TDSE_MAX_WAVES][2], slot_b[TDSE_MAX_WAVES][2];
    __shared__ double2 edge[2][2];
    __shared__ int cur;
    const int nw = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(cur);  // uniform: keeps the instance in SGPRs
        __syncthreads();
        if (t >= n_inst) break;
        const TdseAdjInst &I = insts[order[t]];
        const TdseGeom &G = geoms[I.geom];
        const long long npad = G.npad;
        const int own_chunks = nw * 64 - G.halo_lanes;
        const int nseg = (G.nchunks + own_chunks - 1) / own_chunks;
        for (int k = I.k_hi; k > I.k_lo; k--) {
            const double F = I.field[k - 1];
            const double Fdt = F * G.dt;
            double rs, rc;
            sincos(-Fdt * G.dx, &rs, &rc);
            const double2 rot_conj = make_double2(rc, rs);
            // the phase at the first point of this lane's chunk in segment 0, and its rotation
            // from one segment to the next
            double2 e_lane, r_seg;
            {
                const double x = grid_x((wave * 64 + lane) * C, G.dx, G.xmax);
                sincos(-Fdt * x, &e_lane.y, &e_lane.x);
                sincos(-Fdt * ((double)(own_chunks * C) * G.dx), &r_seg.y, &r_seg.x);
            }
            const double gobs = I.gobs[k];
            const double2 *psik = I.psi + (long long)(k - I.k_lo - 1) * npad;
            double gacc = 0.0;
            double2 e0 = e_lane;
            for (int s = 0; s < nseg; s++) {
                adjoint_segment(G, s, own_chunks, wave, lane, nw, I.kappa, psik, gobs, e0, rot_conj, gacc, slot_f,
                                slot_b, edge[s & 1], edge[(s + 1) & 1]);
                e0 = cmul(e0, r_seg);
            }
            const double gw = wave_sum(gacc);
            if (lane == 0) part[k & 1][wave] = gw;
            __syncthreads();
            if (threadIdx.x == 0) {
                double tot = 0.0;
                for (int w = 0; w < nw; w++) tot += part[k & 1][w];
                I.grad[k - 1] = gobs - 2.0 * G.dt * tot;
            }
        }
        __syncthreads();
    }
}
