// Wave- and workgroup-level building blocks of the segmented Crank-Nicolson solves
// (forward propagation and its adjoint): cross-lane moves in the VALU (DPP), inclusive
// scans of complex affine maps over a wave, and the carry across the waves of a
// workgroup through LDS.
#pragma once
#include "tdse_internal.h"

namespace tdse {

constexpr int C = TDSE_CHUNK;

__device__ inline double2 shfl_up2(double2 v, int d)
{
    return make_double2(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64));
}
__device__ inline double2 shfl_down2(double2 v, int d)
{
    return make_double2(__shfl_down(v.x, d, 64), __shfl_down(v.y, d, 64));
}
// Loads and stores through global-address-space pointers: the arrays live in device memory
// but arrive through structs, where the compiler would otherwise emit flat accesses (which
// wait on both the memory and the LDS counters).
#define TDSE_GLOBAL __attribute__((address_space(1)))
typedef double tdse_d2 __attribute__((ext_vector_type(2)));
__device__ inline double2 ld2(const double2 *p, long long i)
{
    const tdse_d2 v = ((const TDSE_GLOBAL tdse_d2 *)p)[i];
    return make_double2(v.x, v.y);
}
__device__ inline void st2(double2 *p, long long i, double2 v)
{
    tdse_d2 w;
    w.x = v.x;
    w.y = v.y;
    ((TDSE_GLOBAL tdse_d2 *)p)[i] = w;
}
__device__ inline double ld1(const double *p, long long i) { return ((const TDSE_GLOBAL double *)p)[i]; }

// p + i kept opaque to the compiler: the loads at constant distances from it (the C rows of a
// chunk, 64 elements apart) then take the distance as the instruction's immediate offset.
// Written as base + 64 k, whose low bits the compiler knows to be clear, the sum becomes an OR
// and every load gets its own 64-bit address (four VALU instructions each).
template <typename T>
__device__ inline const T *lane_ptr(const T *p, long long i)
{
    unsigned long long a = reinterpret_cast<unsigned long long>(p + i);
    asm("" : "+v"(a));
    return reinterpret_cast<const T *>(a);
}

// Workgroup barrier for exchanges through LDS only: waits for this wave's LDS accesses but
// not for its global loads and stores (a segment's psi stores stay in flight across the next
// segment's barriers).  Global memory written by one wave and read by another needs
// __syncthreads() in between.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// x_j = j dx - xmax in one rounding (an FMA), within an ulp of the grid's own x_j (a product
// and a difference, rounded twice)
__device__ inline double grid_x(int j, double dx, double xmax) { return fma((double)j, dx, -xmax); }

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---- cross-lane moves in the VALU (DPP), 64-bit values as two 32-bit halves ----
// Lanes without a source keep `old` (bound_ctrl off), so scans pad with the identity map.
constexpr int kRowShr = 0x110, kRowShl = 0x100, kWaveShr1 = 0x138, kWaveShl1 = 0x130;
// lane 15 of each row to the next row, lane 31 to rows 2 and 3 (with the row masks below)
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

// ROWS: the rows (of 16 lanes) that take the moved value; the others keep `old`
template <int CTRL, int ROWS = 0xf>
__device__ inline double dpp(double v, double old)
{
    const long long b = __builtin_bit_cast(long long, v), o = __builtin_bit_cast(long long, old);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROWS = 0xf>
__device__ inline double2 dpp2(double2 v, double2 old)
{
    return make_double2(dpp<CTRL, ROWS>(v.x, old.x), dpp<CTRL, ROWS>(v.y, old.y));
}
// Full-row-mask moves whose lanes without a source take 0 (bound_ctrl): no `old` to set up
// first.  dpp2_one: a complex value whose lanes without a source take 1 + 0i (only the high
// half of the real part needs an `old`).
template <int CTRL>
__device__ inline double dpp_z(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ inline double2 dpp2_z(double2 v)
{
    return make_double2(dpp_z<CTRL>(v.x), dpp_z<CTRL>(v.y));
}
template <int CTRL>
__device__ inline double2 dpp2_one(double2 v)
{
    const long long b = __builtin_bit_cast(long long, v.x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0x3ff00000, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return make_double2(__builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo), dpp_z<CTRL>(v.y));
}
__device__ inline double readlane(double v, int l)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double2 readlane2(double2 v, int l) { return make_double2(readlane(v.x, l), readlane(v.y, l)); }

// (Q, z) <- (Q, z) o (Qo, zo): apply the map (Qo, zo) first, then (Q, z)
__device__ inline void compose(double2 &Q, double2 &z, double2 Qo, double2 zo)
{
    z = make_double2(z.x + Q.x * zo.x - Q.y * zo.y, z.y + Q.x * zo.y + Q.y * zo.x);
    Q = make_double2(Q.x * Qo.x - Q.y * Qo.y, Q.x * Qo.y + Q.y * Qo.x);
}

// Inclusive scan over the 64 lanes of the affine maps x_out = z + Q x_in, in lane order
// (FWD: lane l composes lanes 0..l) or reverse order (lanes l..63).  Rows of 16 lanes
// scan with DPP shifts; rows are then joined: in lane order by two DPP row broadcasts (lane 15
// into rows 1 and 3, then lane 31 into rows 2 and 3), in reverse order through the row totals
// read with readlane.
template <bool FWD>
__device__ inline void wave_scan(double2 &Q, double2 &z, int lane)
{
    const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
#define TDSE_SCAN_LEVEL(D)                                                                                   \
    {                                                                                                        \
        const double2 Qo = FWD ? dpp2_one<kRowShr + D>(Q) : dpp2_one<kRowShl + D>(Q);                        \
        const double2 zo = FWD ? dpp2_z<kRowShr + D>(z) : dpp2_z<kRowShl + D>(z);                            \
        compose(Q, z, Qo, zo);                                                                               \
    }
    TDSE_SCAN_LEVEL(1)
    TDSE_SCAN_LEVEL(2)
    TDSE_SCAN_LEVEL(4)
    TDSE_SCAN_LEVEL(8)
#undef TDSE_SCAN_LEVEL
    if constexpr (FWD) {
        // rows 1 and 3 take the totals of rows 0 and 2; then rows 2 and 3 take lanes 0..31
        // (rows that take nothing compose with the identity map, which leaves them exact)
        {
            const double2 Qo = dpp2<kRowBcast15, 0xa>(Q, one), zo = dpp2<kRowBcast15, 0xa>(z, zero);
            compose(Q, z, Qo, zo);
        }
        {
            const double2 Qo = dpp2<kRowBcast31, 0xc>(Q, one), zo = dpp2<kRowBcast31, 0xc>(z, zero);
            compose(Q, z, Qo, zo);
        }
        return;
    }
    // reverse order: row totals are the first lane of each row; carry into this row from the
    // rows after it
    const int row = lane >> 4;
    double2 tq[4], tz[4];
#pragma unroll
    for (int r = 1; r < 4; r++) {
        tq[r] = readlane2(Q, 16 * r);
        tz[r] = readlane2(z, 16 * r);
    }
    double2 CQ = one, CZ = zero;
#pragma unroll
    for (int r = 3; r > 0; r--)
        if (r > row) {
            double2 q = tq[r], zz = tz[r];
            compose(q, zz, CQ, CZ);
            CQ = q;
            CZ = zz;
        }
    compose(Q, z, CQ, CZ);
}

// r' = (M2' - i H') psi over one lane's chunk P (in place), M2' = (1, 10, 1) and
// H' = (g_{j-1}, 10 g_j + K, g_{j+1}); left / right: the points either side of the chunk, gL /
// gR their g.  Real part summed as pm.x + gm pm.y + pp.x + gp pp.y + 10 p0.x + d p0.y (the
// imaginary part likewise): 12 instructions per point.
__device__ inline void cn_rhs(double2 (&P)[C], const double (&Gv)[C], double2 left, double2 right, double gL,
                              double gR, double K)
{
    const double ten = 10.0;
    double2 pm = left;
    double gm = gL;
#pragma unroll
    for (int i = 0; i < C; i++) {
        const double2 p0 = P[i];
        const double2 pp = (i + 1 < C) ? P[i + 1] : right;
        const double gp = (i + 1 < C) ? Gv[i + 1] : gR;
        const double d = fma(ten, Gv[i], K);
        double re = fma(gp, pp.y, fma(gm, pm.y, pm.x) + pp.x);
        double im = fma(-gp, pp.x, fma(-gm, pm.x, pm.y) + pp.y);
        re = fma(d, p0.y, fma(ten, p0.x, re));
        im = fma(-d, p0.x, fma(ten, p0.y, im));
        P[i] = make_double2(re, im);
        pm = p0;
        gm = Gv[i];
    }
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

// Cross-wave part of an exclusive scan of affine maps x_out = z + Q x_in.
// (Qi, zi): the wave-inclusive composition of lanes [0..lane] (forward) or [lane..63]
// (backward).  Returns the carry into this lane from everything before it in the
// workgroup-wide lane order over the nw live waves.  The value entering the window is
// x_in (forward: the exact value before lane 0 of wave 0; backward: zero after the last
// live wave).
template <bool FWD>
__device__ inline double2 block_carry(double2 Qi, double2 zi, int wave, int lane, int nw, double2 (*slot)[2],
                                      double2 x_in = make_double2(0.0, 0.0))
{
    // exclusive within the wave: shift the inclusive maps by one lane
    const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
    const double2 Qe = FWD ? dpp2_one<kWaveShr1>(Qi) : dpp2_one<kWaveShl1>(Qi);
    const double2 ze = FWD ? dpp2_z<kWaveShr1>(zi) : dpp2_z<kWaveShl1>(zi);
    if (lane == (FWD ? 63 : 0)) {
        slot[wave][0] = Qi;
        slot[wave][1] = zi;
    }
    lds_barrier();
    // carry into this wave: compose the totals of the waves before it
    double2 Z = FWD ? x_in : make_double2(0.0, 0.0);
    if (FWD) {
        for (int v = 0; v < wave; v++) {
            const double2 q = slot[v][0], z = slot[v][1];
            Z = make_double2(z.x + q.x * Z.x - q.y * Z.y, z.y + q.x * Z.y + q.y * Z.x);
        }
    } else {
        for (int v = nw - 1; v > wave; v--) {
            const double2 q = slot[v][0], z = slot[v][1];
            Z = make_double2(z.x + q.x * Z.x - q.y * Z.y, z.y + q.x * Z.y + q.y * Z.x);
        }
    }
    return make_double2(ze.x + Qe.x * Z.x - Qe.y * Z.y, ze.y + Qe.x * Z.y + Qe.y * Z.x);
}

} // namespace tdse
