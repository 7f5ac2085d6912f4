/*
 * hankel.h -- C ABI of libmmahankel.so, the MI355X implementation of the Hankel far-field
 * stage of Hankel/Hankel_transform.py.
 *
 * The reference is Python; its interface for this path is the set of functions of
 * Hankel_transform.py (FSources_provider :29-80, get_propagation_pre_factor_function :85-448,
 * HankelTransform :452-519, Hankel_long :522-687).  mma_amd/hankel.py mirrors those
 * signatures and calls the entry points below; they take DEVICE pointers (plain doubles,
 * complex values as interleaved (re, im) pairs), sizes and a hipStream_t, and return a
 * status (0 = OK) once the work queued on the stream has finished -- except the three steps of
 * the gridded contraction (hk_grid_spread, hk_grid_bessel, hk_gemm), which return once their
 * kernel is queued (launch errors are still reported): their results are for later work on
 * the same stream.
 *
 * Geometry of the far-field integral, for plane k at z_k, frequency omega_o and screen
 * radius rho_m (SI units, Hankel_transform.py:494-512):
 *     P_k(rho_m, omega_o) = sum_j tw_j pre_k(r_j, omega_o) e^{-i K r_j^2 / (2 D_k)} r_j
 *                                  F_k(omega_o, r_j) J0(K r_j rho_m / D_k),
 *     K = omega_o / c,  D_k = distance - z_k,  tw_j = trapezoid weights of the r grid.
 */
#ifndef MMA_HANKEL_H
#define MMA_HANKEL_H

#ifdef __cplusplus
extern "C" {
#endif

enum hk_status {
    HK_OK = 0,
    HK_ERR_ARGUMENT = 3,
    HK_ERR_UNSUPPORTED = 4,
    HK_ERR_DEVICE = 5,
    HK_ERR_TABLE = 7 /* photon energy outside the scattering-factor table or f1 not listed
                        (XUV_refractive_index.py:54-55,156: the reference raises ValueError) */
};

int hk_last_error(void);
const char *hk_last_error_message(void);

/* ---- pre-factor tables (replaces get_propagation_pre_factor_function, :85-448) ---------- */

enum hk_pressure_mode {
    HK_PRESSURE_SCALAR = 0, /* :405-448: pre[k][o] = p exp(omega (z_k i disp + (z_k - z_last) abs)) */
    HK_PRESSURE_Z = 1,      /* :146-239: cumulative z integrals of the tabulated pressure */
    HK_PRESSURE_ZR = 2      /* :301-399: pressure tabulated in z and r; pre[k][o][r] */
};

typedef struct {
    int mode;                       /* hk_pressure_mode */
    double pressure;                /* HK_PRESSURE_SCALAR: pressure [bar] */
    int np_z, np_r;                 /* pressure table sizes (z; r for HK_PRESSURE_ZR) */
    const double *p_z, *p_r;        /* device: pressure-table grids [m], ascending */
    const double *p_value;          /* device: np_z values, or np_z x np_r (row-major, z major) */
    int include_dispersion, include_absorption;
    double n_ir;                    /* effective IR refractive index */
    int n_f1, n_f2;                 /* scattering-factor tables of the chosen gas/tables */
    const double *e_f1, *f1;        /* device: energies [eV] and f1 (dispersion tables) */
    const double *e_f2, *f2;        /* device: energies [eV] and f2 (absorption tables) */
} hk_medium;

/* pre: device complex, nz x no (SCALAR, Z) or nz x no x nr (ZR), row-major.
 * renorm: device real nz x no (exp_renorm of the reference) or NULL (not produced for ZR).
 * zgrid (nz), ogrid (no, rad/s), rgrid (nr): device. */
int hk_prefactor(const hk_medium *m, int nz, const double *zgrid, int no, const double *ogrid, int nr,
                 const double *rgrid, double *pre, double *renorm, void *stream);

/* The same tables, queued on `stream` with nothing waiting (hk_prefactor waits for the tables to
 * report a table error at once).  flag: device int, cleared here and raised by the kernels when
 * a photon energy lies outside the scattering-factor tables or has no f1 listed; once the
 * stream has passed this work the caller reads it and hands it to hk_prefactor_status, which
 * returns what hk_prefactor would have (HK_OK or HK_ERR_TABLE: the reference's interpolation
 * out of range, or its XUV_refractive_index.py:156 error).  work: `no` doubles of device scratch for a z-tabulated
 * pressure (else unused, may be NULL).  Calls that share flag or work must be ordered on one
 * stream.  No counterpart in the reference (its pre-factor function raises at once); for
 * callers that keep a far-field pass free of host waits. */
int hk_prefactor_enqueue(const hk_medium *m, int nz, const double *zgrid, int no, const double *ogrid, int nr,
                         const double *rgrid, double *pre, double *renorm, int *flag, double *work, void *stream);
int hk_prefactor_status(int flag);

/* ---- the contraction over r and z ------------------------------------------------------- */

/* Source planes: src + k * plane_stride is plane k, laid out [o][j] (no x nr, row-major) as
 * FSources_provider yields them (:66-70); complex128 (src_c64 = 0) or complex64 (src_c64 = 1).
 * pre: NULL (factor 1), nz x no (pre_mode 1) or nz x no x nr (pre_mode 2) complex.
 * tw: trapezoid weights of rgrid (nr).  weights: n_out x nz real (plane weights of each
 * output); per_plane = 1 instead writes every plane's transform to its own output. */
typedef struct {
    int nz, no, nr, nrho;
    const void *src;
    int src_c64;
    long long plane_stride;         /* complex elements between planes */
    const double *zgrid, *ogrid, *rgrid, *tw, *rho;
    double distance;
    int near_field;
    const double *pre;
    int pre_mode;
    /* optional HOST copies of zgrid, ogrid and rgrid (NULL: the entry points that check the
     * grids on the host copy them from the device, a synchronous transfer each) */
    const double *host_zgrid, *host_ogrid, *host_rgrid;
} hk_planes;

/* Direct FP64 contraction: J0 evaluated in FP64 for every (plane, omega, rho, r) element
 * (the reference's HankelTransform, :492-514, for all planes in one launch).
 * out: device complex128, (per_plane ? nz : n_out) blocks of nrho x no (row-major: the
 * (Nr_FF, No) orientation of Hankel_long's results).  n_out <= 4 unless per_plane. */
int hk_direct(const hk_planes *p, int n_out, const double *weights, int per_plane, double *out, void *stream);

/* Gridded contraction (no counterpart in the reference; same integral, reassociated):
 * J0(s r rho) = J0(q rho) with q = K r / D_k is interpolated in q from a uniform grid
 * Q_l = (l - HK_GRID_P/2) h, l < nq, by Lagrange interpolation of HK_GRID_P points, so that
 *     sum_k w_k P_k(rho_m, omega_o) = sum_l J0(rho_m Q_l) C_l(omega_o)
 * with C spread from the sources (FP64) and the sum over l a real GEMM shared by all omega.
 * hk_grid_size picks nq (a multiple of 64) and h for a J0-interpolation error below tol. */
#define HK_GRID_P 8
int hk_grid_size(double q_max, double rho_max, double tol, int *nq, double *h);

/* Operand element types: double, float, or the value split into 2 (HK_BF16X2) or 3 (HK_BF16X3)
 * bf16 parts v ~ p0 + p1 (+ p2), stored as 16-bit words with the parts of a row side by side
 * (part q of element l of a row at row_start + q * nq + l; a row then holds parts * nq words
 * and ld >= parts * nq): the operands of hk_gemm_split. */
enum hk_operand { HK_F64 = 0, HK_F32 = 1, HK_BF16X2 = 2, HK_BF16X3 = 3, HK_F16X2 = 4 };
/* HK_F16X2: each row scaled by a power of two s (largest element in [2^13, 2^14)) and split into
 * two fp16 parts, v s ~ h0 + 2^-11 h1 (part q at row_start + q * nq + l), the double s at
 * row_start + 2 * nq (16-bit words; ld >= 2 * nq + 4, a multiple of 8): the operands of
 * hk_gemm_f16x2. */

/* C operand, "B^T" layout: for output q < n_out, row row0 + 2 (no q + o) holds the real
 * parts and the next row the imaginary parts of frequency o; nq columns, row stride ld
 * (elements of the type; >= nq, or >= parts * nq for the split types).  weights: n_out x nz
 * plane weights (zero weights skip the plane).  All outputs are spread in one launch. */
int hk_grid_spread(const hk_planes *p, int n_out, const double *weights, int nq, double h, int type, void *out,
                   long long ld, long long row0, void *stream);

/* A operand: J0(rho_m |Q_l|), nrho rows x nq (row stride ld), any hk_operand type. */
int hk_grid_bessel(int nrho, const double *rho, int nq, double h, int type, void *out, long long ld, void *stream);

enum hk_gemm_kind {
    HK_GEMM_F32 = 1,    /* FP32 MFMA (v_mfma_f32_32x32x2_f32) */
    HK_GEMM_BF16X3 = 2, /* both FP32 operands split into bf16 hi + lo on the fly; three bf16 MFMA
                           products hi*hi + hi*lo + lo*hi (v_mfma_f32_32x32x16_bf16) */
    HK_GEMM_BF16X6 = 3  /* three-way bf16 split, the six products down to 2^-16 relative */
};
/* All kinds accumulate each 32-deep K stage in FP32 on the matrix cores and the stages in FP64. */

/* out[m][n] = sum_k A[m][k] B[n][k] on the matrix cores, written as double with row stride
 * ldo.  A: m x k floats (row stride lda), B: n x k floats (ldb); k a multiple of 64, strides
 * multiples of 4. */
int hk_gemm(int kind, int m, int n, int k, const float *a, long long lda, const float *b, long long ldb, double *out,
            long long ldo, void *stream);

/* The same product from operands already split into `parts` (2 or 3) bf16 parts (HK_BF16X2,
 * HK_BF16X3 layout: row r at a + r * lda, part q of it at + q * k; 16-bit words): parts 2 gives
 * p0 q0 + p0 q1 + p1 q0, parts 3 the six products of HK_GEMM_BF16X6.  No splitting in the
 * kernel, and the operand tiles are copied global -> LDS directly.  k a multiple of 64,
 * lda, ldb >= parts * k and multiples of 8, 16-byte aligned operands. */
int hk_gemm_split(int parts, int m, int n, int k, const void *a, long long lda, const void *b, long long ldb,
                  double *out, long long ldo, void *stream);

/* The same product from HK_F16X2 operands: h0 h0' + 2^-11 (h0 h1' + h1 h0') on the fp16 matrix
 * cores (v_mfma_f32_32x32x16_f16), divided by the two row scales; 22 bits per operand against
 * 16 for two bf16 parts, at the same rate. */
int hk_gemm_f16x2(int m, int n, int k, const void *a, long long lda, const void *b, long long ldb, double *out,
                  long long ldo, void *stream);

#ifdef __cplusplus
}
#endif
#endif
