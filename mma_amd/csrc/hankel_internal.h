// Shared declarations of the Hankel kernels and their host launcher.
#pragma once
#include <hip/hip_runtime.h>

#include "hankel.h"

// Photon/medium constants of the Python unit set (shared_python/units.py:3-20,
// XUV_refractive_index.py:123), evaluated on the host in the reference's operation order.
struct HkConst {
    double c_light;
    double timeau;      // omega_au = omega_SI * timeau
    double ev_per_au;   // E[eV] = omega_au / ev_per_au   (mynumerics.py:313,326)
    double lam_num;     // lambda_SI = lam_num / (omega_au * lam_den)  (mynumerics.py:323)
    double lam_den;
    double n_ref_re;    // N_ref * r_electron_classical
};

enum HkErrFlag { HK_FLAG_OK = 0, HK_FLAG_RANGE = 1, HK_FLAG_F1 = 2 };

__global__ void hk_prefactor_kernel(HkConst K, hk_medium m, int nz, const double *zgrid, int no, const double *ogrid,
                                    int nr, const double *rgrid, double2 *pre, double *renorm, double *exit_abs,
                                    int *flag, int staged);
// p[0 .. n) = v (flags and counters cleared by a kernel, not a memset node)
__global__ void hk_fill_int_kernel(int *p, int n, int v);
__global__ void hk_prefactor_z_kernel(HkConst K, hk_medium m, int nz, const double *zgrid, int no, const double *ogrid,
                                      double2 *pre, double *renorm, double *exit_abs, int *flag);
constexpr int kWalkParallelMax = 4096;  // pressure tables up to this length: one thread per (frequency, plane)
__global__ void hk_prefactor_finish_kernel(int nz, int no, const double *ogrid, const double *exit_abs, double2 *pre,
                                           double *renorm);

template <typename S>
__global__ void hk_direct_kernel(hk_planes p, int n_out, const double *weights, int per_plane, double2 *out,
                                 double inv_c, int uniform_r, double r0, double dr);

template <int TYPE>
__global__ void hk_spread_kernel(hk_planes p, const double *weights, int nq, double h, double inv_c, void *out,
                                 long long ld, long long row0);

// range of sincos_reduced (hankel_contract.hip): |x| below it keeps n = rint(2x/pi) an exact int
constexpr double kSincosLimit = 1e9;

// hk_spread_window_kernel: doubles of per-plane factors in LDS after the bins (5 used, padded
// to keep each plane's entry 16-byte aligned)
constexpr int kSpreadTab = 6;
template <int TYPE, typename S, int NEAR>
__global__ void hk_spread_window_kernel(hk_planes p, const double *weights, int nq, double h, double inv_c, void *out,
                                        long long ld, long long row0);

template <int TYPE>
__global__ void hk_bessel_kernel(int nrho, const double *rho, int nq, double h, void *out, long long ld);

template <int NP>
__global__ void hk_gemm_split_kernel(int M, int N, int K, const unsigned short *A, long long lda, const unsigned short *B,
                                     long long ldb, double *out, long long ldo);
size_t hk_gemm_split_lds(int np);  // dynamic LDS of hk_gemm_split_kernel<np> (and of the f16x2 kernel: np 2)
__global__ void hk_gemm_f16x2_kernel(int M, int N, int K, const unsigned short *A, long long lda, const unsigned short *B,
                                     long long ldb, double *out, long long ldo);

template <int KIND>
__global__ void hk_gemm_kernel(int m, int n, int k, const float *a, long long lda, const float *b, long long ldb,
                               double *out, long long ldo);
