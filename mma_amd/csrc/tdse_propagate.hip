// Persistent batched Crank-Nicolson propagator (1DTDSE/sources/prop.c:24-347) for gfx950.
//
// One workgroup of NW waves owns one instance at a time (instances are taken from
// an atomic counter in the given order, largest first) and runs all its time
// steps.  Per step, the grid is covered by segments; one wave solves one segment:
//
//   lanes 0..63 each own one chunk of TDSE_CHUNK consecutive points of the
//   segment's extended window, which is the segment's own chunks plus H halo
//   chunks on each side.  The tridiagonal solve B'u = A'psi uses the global
//   Thomas factorisation of B' (pivot table ib, step-independent), written as two
//   first-order linear recurrences
//       forward   z_j = ib_j (r_j - a'_j z_{j-1})
//       backward  u_j = z_j + h_j u_{j+1},   h_j = -c'_j ib_j
//   each evaluated in two passes around a 64-lane scan of chunk transfer maps
//   (carry-out with zero carry-in, exclusive scan with the chunk products from
//   the tables, recompute with the true carry-in).  The two recurrences are
//   started from zero at the ends of the extended window: the inverse of B' decays
//   by |q|, |h| <= rho < 1 per point, so after H*TDSE_CHUNK >= log(1e-17)/log(rho)
//   points the truncation is below 1e-17 of max|psi| and the own chunks are exact
//   to rounding (truncated-SPIKE).  Segments of one step are independent, read
//   psi of step k from one buffer and write step k+1 to the other.
//
// Observables (prop.c:293-347) are reduced per wave, then across waves in LDS in a
// fixed order (deterministic).  HBM traffic per grid-point-step: psi read + write,
// 32 bytes; the tables are shared by all instances of a geometry and stay in L2.
#include "tdse_internal.h"

namespace {

constexpr int C = TDSE_CHUNK;
constexpr int kWindow = 8;  // table loads a pass may issue ahead of their use

__device__ inline double2 shfl_up2(double2 v, int d)
{
    return make_double2(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64));
}
__device__ inline double2 shfl_down2(double2 v, int d)
{
    return make_double2(__shfl_down(v.x, d, 64), __shfl_down(v.y, d, 64));
}
__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

struct Acc {
    double pr, pi, pint, pos, gp;
};

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Solve one segment of one step.  snap_row != nullptr: also store the new psi there (natural order).
__device__ void solve_segment(const TdseGeom &G, int seg, int lane, const double2 *__restrict__ src,
                              double2 *__restrict__ dst, double Fdt, double2 rot, Acc &acc, double *snap_row)
{
    const int H = G.halo_lanes;
    const int c = seg * G.own_lanes - H + lane;  // >= -H > -64: always inside the padded layout
    const long long base = tdse_chunk_base(c);
    const double *gt = G.g + base;
    const double2 *ibt = G.ib + base;
    // The table pointers are made opaque before every pass so that each pass reloads
    // its coefficients from cache instead of keeping all of them live in registers.
    auto fresh = [&]() {
        asm volatile("" : "+v"(gt));
        asm volatile("" : "+v"(ibt));
    };

    double2 P[C];
#pragma unroll
    for (int i = 0; i < C; i++) P[i] = src[base + 64 * i];
    // neighbours across the chunk edges; zero outside the extended window (truncation zone)
    double2 left = shfl_up2(P[C - 1], 1), right = shfl_down2(P[0], 1);
    const double g_first = gt[0], g_last = gt[64 * (C - 1)];
    double gL = __shfl_up(g_last, 1, 64), gR = __shfl_down(g_first, 1, 64);
    if (lane == 0) {
        left = make_double2(0.0, 0.0);
        gL = 0.0;
    }
    if (lane == 63) {
        right = make_double2(0.0, 0.0);
        gR = 0.0;
    }

    // r' = (M2' - i H') psi
    {
        double2 pm = left;
        double gm = gL, g0 = g_first;
        const double K = G.k36;
#pragma unroll
        for (int i = 0; i < C; i++) {
            if (i % kWindow == 0) fresh();
            const double2 p0 = P[i];
            const double2 pp = (i + 1 < C) ? P[i + 1] : right;
            const double gp = (i + 1 < C) ? gt[64 * (i + 1)] : gR;
            const double d = fma(10.0, g0, K);
            const double mr = fma(10.0, p0.x, pm.x + pp.x);
            const double mi = fma(10.0, p0.y, pm.y + pp.y);
            const double hr = fma(gm, pm.x, fma(d, p0.x, gp * pp.x));
            const double hi = fma(gm, pm.y, fma(d, p0.y, gp * pp.y));
            P[i] = make_double2(mr + hi, mi - hr);
            pm = p0;
            gm = g0;
            g0 = gp;
        }
    }

    // forward recurrence z_j = ib_j (r_j - (1 + i g_{j-1}) z_{j-1})
    {
        fresh();
        double2 z = make_double2(0.0, 0.0);
        double gm = gL;
#pragma unroll
        for (int i = 0; i < C; i++) {
            if (i % kWindow == 0) fresh();
            const double2 ib = ibt[64 * i];
            const double tr = fma(gm, z.y, P[i].x - z.x);
            const double ti = fma(-gm, z.x, P[i].y - z.y);
            z = cmul(make_double2(tr, ti), ib);
            gm = gt[64 * i];
        }
        // exclusive scan of the chunk maps z_out = z + Q z_in over the lanes
        const double2 *qp = G.chq + (c + 64);
        asm volatile("" : "+v"(qp));
        double2 Q = *qp;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double2 Qo = shfl_up2(Q, d), zo = shfl_up2(z, d);
            if (lane >= d) {
                z = make_double2(z.x + Q.x * zo.x - Q.y * zo.y, z.y + Q.x * zo.y + Q.y * zo.x);
                Q = cmul(Q, Qo);
            }
        }
        z = shfl_up2(z, 1);
        if (lane == 0) z = make_double2(0.0, 0.0);
        fresh();
        gm = gL;
#pragma unroll
        for (int i = 0; i < C; i++) {
            if (i % kWindow == 0) fresh();
            const double2 ib = ibt[64 * i];
            const double tr = fma(gm, z.y, P[i].x - z.x);
            const double ti = fma(-gm, z.x, P[i].y - z.y);
            z = cmul(make_double2(tr, ti), ib);
            P[i] = z;
            gm = gt[64 * i];
        }
    }

    // backward recurrence u_j = z_j + h_j u_{j+1},  h_j = -(1 + i g_{j+1}) ib_j
    {
        fresh();
        double2 u = make_double2(0.0, 0.0);
        double gp = gR;
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            if ((C - 1 - i) % kWindow == 0) fresh();
            const double2 ib = ibt[64 * i];
            const double hr = fma(gp, ib.y, -ib.x), hi = -fma(gp, ib.x, ib.y);
            u = make_double2(P[i].x + hr * u.x - hi * u.y, P[i].y + hr * u.y + hi * u.x);
            gp = gt[64 * i];
        }
        const double2 *gq = G.chg + (c + 64);
        asm volatile("" : "+v"(gq));
        double2 Gt = *gq;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double2 Go = shfl_down2(Gt, d), uo = shfl_down2(u, d);
            if (lane + d < 64) {
                u = make_double2(u.x + Gt.x * uo.x - Gt.y * uo.y, u.y + Gt.x * uo.y + Gt.y * uo.x);
                Gt = cmul(Gt, Go);
            }
        }
        u = shfl_down2(u, 1);
        if (lane == 63) u = make_double2(0.0, 0.0);
        fresh();
        gp = gR;
#pragma unroll
        for (int i = C - 1; i >= 0; i--) {
            if ((C - 1 - i) % kWindow == 0) fresh();
            const double2 ib = ibt[64 * i];
            const double hr = fma(gp, ib.y, -ib.x), hi = -fma(gp, ib.x, ib.y);
            u = make_double2(P[i].x + hr * u.x - hi * u.y, P[i].y + hr * u.y + hi * u.x);
            P[i] = u;
            gp = gt[64 * i];
        }
    }

    if (lane < H || lane >= 64 - H || c >= G.nchunks) return;  // only the segment's own chunks continue

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
        if (i % kWindow == 0) asm volatile("" : "+v"(dp));
        const double2 v = cmul(e, P[i]);
        P[i] = v;
        dst[base + 64 * i] = v;
        const double n2 = fma(v.x, v.x, v.y * v.y);
        const double xj = __dsub_rn(__dmul_rn((double)(j0 + i), G.dx), G.xmax);
        acc.pos = fma(n2, xj, acc.pos);
        acc.gp = fma(n2, dp[64 * i], acc.gp);
        e = cmul(e, rot);
    }
    // <psi0|psi> over the rows where psi0 is not negligible
    if (j0 + C - 1 >= G.p0_lo && j0 <= G.p0_hi) {
        const double2 *q0 = G.psi0 + base;
#pragma unroll
        for (int i = 0; i < C; i++) {
            if (i % kWindow == 0) asm volatile("" : "+v"(q0));
            const double2 q = q0[64 * i];
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
    if (snap_row) {
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

__global__ void __launch_bounds__(256, 2)
tdse_propagate_kernel(const TdseGeom *geoms, const TdseInst *insts, const int *order, int n_inst, int *counter,
                      double2 *work, long long work_stride)
{
    __shared__ double part[2][TDSE_MAX_WAVES][5];
    __shared__ int cur;
    const int nw = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    double2 *bufA = work + (long long)blockIdx.x * 2 * work_stride;
    double2 *bufB = bufA + work_stride;

    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int t = cur;
        __syncthreads();
        if (t >= n_inst) break;
        const TdseInst &I = insts[order[t]];
        const TdseGeom &G = geoms[I.geom];
        const long long npad = G.npad;

        // psi(t = tmin) = psi0
        for (long long q = threadIdx.x; q < npad; q += blockDim.x) bufA[q] = G.psi0[q];
        if (threadIdx.x == 0) {
            I.src[0] = 0.0;
            I.ptot[0] = 1.0;
            I.pint[0] = G.pint0;
            I.expv[0] = 0.0;
            if (I.efield) I.efield[0] = 0.0;
        }
        __syncthreads();

        bool gating = true;  // prop.c:125-134
        int i_snap = 0;
        for (int k = 0; k < I.n_steps; k++) {
            const double fk = I.field[k], fk1 = I.field[k + 1];
            double F;
            if (gating) {
                if (fk * fk1 <= 0.0) gating = false;
                F = 0.0;
            } else {
                F = fk;
            }
            const double Fdt = F * G.dt;
            double rs, rc;
            sincos(Fdt * G.dx, &rs, &rc);
            const double2 rot = make_double2(rc, rs);
            // snapshots, prop.c:248-255: row i_snap gets psi after step k when k % every == 0
            double *snap_row = nullptr;
            if (I.snap && k % I.snap_every == 0 && i_snap < I.n_snap - 1) {
                snap_row = I.snap + (long long)i_snap * 2 * G.nx;
                i_snap++;
            }
            const double2 *src = (k & 1) ? bufB : bufA;
            double2 *dst = (k & 1) ? bufA : bufB;
            Acc acc = {0.0, 0.0, 0.0, 0.0, 0.0};
            for (int s = wave; s < G.nseg; s += nw) solve_segment(G, s, lane, src, dst, Fdt, rot, acc, snap_row);
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
            if (threadIdx.x == 0 && k + 1 < I.n_steps) {
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
            const double2 *fin = (I.n_steps & 1) ? bufB : bufA;
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
