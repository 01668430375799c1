/*
 * rgf.h -- C ABI of librgf_hip.so, an MI355X (gfx950) solver for symmetric
 * positive-definite block-tridiagonal-with-arrowhead matrices: nt diagonal
 * blocks of size ns x ns, nt-1 sub-diagonal coupling blocks and nd dense arrow
 * rows/columns (n = ns*nt + nd).
 *
 * Each entry point replaces a member of rjanalik/RGF's `template<class T> class
 * RGF` (RGF.H:11-80) for T = double; the C++ wrapper with the reference's own
 * method names is RGF.hpp next to this file.  All status codes are rgf_status.
 * Handles are independent; one handle must not be used by two threads at once.
 */
#ifndef RGF_H
#define RGF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rgf_handle_s *rgf_handle_t;

typedef enum {
    RGF_OK = 0,
    RGF_EINVAL = 1,   /* bad shape, bad CSC pattern, bad argument */
    RGF_ENOTSPD = 2,  /* a pivot of the block Cholesky was not positive */
    RGF_ENOMEM = 3,   /* device allocation failed */
    RGF_EDEVICE = 4,  /* HIP runtime error */
    RGF_ESTATE = 5    /* call not valid in this state (e.g. no resident matrix) */
} rgf_status;

typedef enum {
    RGF_FP64 = 0,          /* every product in FP64 on MFMA */
    /* 1 is not used: an FP32-product mode would cap the product speed-up at the 2x of
       FP32 over FP64 MFMA peak, so only the split-BF16 mode is provided */
    RGF_MIXED_BF16X3 = 2   /* products as split-BF16 (3 terms) on MFMA, FP64 elsewhere */
} rgf_precision;

typedef enum {
    RGF_SELINV_EXACT = 0,     /* full Takahashi recursion (exact diag(A^-1) for every nd) */
    RGF_SELINV_REFERENCE = 1  /* RGF.H:476-503 as written: drops the arrow cross terms */
} rgf_selinv_mode;

typedef enum {
    RGF_CHOLESKY_AUTO = 0,        /* task graph for handles of at most 4 matrices, tile launches beyond */
    RGF_CHOLESKY_TASK_GRAPH = 1,  /* each diagonal block's Cholesky and L_ii^-1 as one launch of a tile
                                     task graph (latency: the pivot chain overlaps the other tiles' work) */
    RGF_CHOLESKY_TILE_LAUNCHES = 2 /* a launch per 64-wide column block plus GEMM launches for the
                                     trailing updates and L_ii^-1 (throughput over many matrices) */
} rgf_cholesky_schedule;

typedef struct {
    int precision;      /* rgf_precision */
    int selinv_mode;    /* rgf_selinv_mode */
    int keep_matrix;    /* 1: keep A resident beside the factor (residual, refinement, re-factorize) */
    int device;         /* HIP device ordinal */
    int refine_iters;   /* max FP64 iterative-refinement steps of a mixed-precision solve */
    double refine_tol;  /* stop refining when ||b - A x|| / ||b|| <= refine_tol */
    int graphs;         /* 1: run the factorize and selected-inversion sweeps as HIP graphs,
                           captured on first use and replayed (repeated sweeps of one handle) */
    int keep_splits;    /* mixed precision: 1 keeps the split-BF16 copy of every coupling panel
                           [B_i; D_i] that the factorization's products write, for the selected
                           inversion's W = [B_i; D_i] X_i (~4 bytes per panel entry, about a
                           quarter more device memory at ns = 2865); 0 splits them again there */
    int cholesky;       /* rgf_cholesky_schedule.  The two schedules sum the same terms in different
                           orders: results agree to rounding, and bitwise between handles that run
                           the same schedule (a matrix of a batch and the same matrix alone) */
    int keep_w;         /* 1: keep the selected inversion's W_i = [B_i; D_i] L_ii^-1 of every block
                           (nt (ns + nd) ns doubles per matrix, 6.6 GB at ns = 2865, nt = 100, nd = 6);
                           solves after a selected inversion of the current factor then chain one
                           product per block and direction, r_{i+1} = b_{i+1} - W_i r_i and
                           x_i = X_i^T y_i - W_i^T [x_{i+1}; x_d], with the X_i products of all blocks
                           in one launch each way.  The same solution to rounding.  FP64 handles
                           only (ignored with mixed precision: through its split-BF16 W_i a refined
                           solve needs a second refinement step); 0 by default */
} rgf_options;

typedef struct {
    uint64_t seed;
    uint32_t matrix_id;
    double sigma_c;     /* std-dev of the coupling blocks (0.1) */
    double sigma_a;     /* std-dev of the arrow rows (0.05) */
    double theta_d;     /* scale of the diagonal blocks (1) */
    double theta_c;     /* scale of the coupling blocks (1) */
    int theta_mode;     /* 0: theta_d / theta_c as given; 1: per instance theta_d ~ U(0.5,2),
                           theta_c ~ U(0.1,1) drawn from (seed, matrix id) */
} rgf_gen_params;

typedef struct {
    size_t ns, nt, nd, n;
    size_t device_bytes;        /* bytes this handle holds on the device */
    double flops_factor;        /* canonical algorithmic flops: factorize */
    double flops_selinv;        /* canonical algorithmic flops: selected inversion */
    double bytes_solve_per_rhs; /* canonical bytes streamed by one solve (forward + backward) */
} rgf_info_t;

void rgf_default_options(rgf_options *opts);
void rgf_default_gen_params(rgf_gen_params *p);

/* RGF<T>::RGF(size_t* ia, size_t* ja, T* a, size_t ns, size_t nt, size_t nd), RGF.H:85-146.
 * Lower-triangle CSC, 0-based, entries confined to the block-arrow pattern (checked).
 * Unlike the reference the arrays are copied to the device at construction and
 * need not outlive the handle.  Requires nd <= ns and nd <= 64. */
int rgf_create_csc(const size_t *ia, const size_t *ja, const double *a, size_t ns, size_t nt, size_t nd,
                   const rgf_options *opts, rgf_handle_t *out);

/* Extension: the matrix is generated on the device from (seed, matrix id); the
 * definition is in rgf_gen.hip (benchmark workloads). */
int rgf_create_generated(size_t ns, size_t nt, size_t nd, const rgf_gen_params *gen,
                         const rgf_options *opts, rgf_handle_t *out);

/* Extension: `count` matrices of one block structure in one handle (matrix b has id
 * gen->matrix_id + b).  factorize / selected inversion run over all of them in the
 * same launches; the single-matrix entry points address the matrix chosen by
 * rgf_select (default 0). */
int rgf_create_generated_batch(size_t ns, size_t nt, size_t nd, size_t count, const rgf_gen_params *gen,
                               const rgf_options *opts, rgf_handle_t *out);
int rgf_batch_count(rgf_handle_t h);
int rgf_select(rgf_handle_t h, size_t index);
/* Regenerate the matrices of a generated handle (e.g. after an in-place factorization). */
int rgf_regenerate(rgf_handle_t h);
/* Device bytes a handle of this shape and these options (NULL: defaults) holds after
 * factorize, solve and selected inversion (0 for an invalid shape). */
size_t rgf_footprint_bytes(size_t ns, size_t nt, size_t nd, size_t count, const rgf_options *opts);

/* RGF<T>::~RGF(), RGF.H:151-170 */
int rgf_destroy(rgf_handle_t h);

/* RGF<T>::factorize(), RGF.H:510-546.  RGF_ENOTSPD carries the 1-based column in
 * rgf_last_error(). */
int rgf_factorize(rgf_handle_t h);

/* RGF<T>::solve(T* x, T* b, size_t nrhs), RGF.H:558-578 (and solve(b, nrhs),
 * RGF.H:550-554, with x == b).  Host arrays, column-major n x nrhs, ld = n.
 * Factorizes first if needed. */
int rgf_solve(rgf_handle_t h, double *x, const double *b, size_t nrhs);

/* RGF<T>::logDet(), RGF.H:616-627. */
int rgf_logdet(rgf_handle_t h, double *out);

/* RGF<T>::RGFdiag(T* diag), RGF.H:582-612: diag(A^-1), length n, host array.
 * The factor stays valid afterwards (the reference re-factorizes, RGF.H:588). */
int rgf_selinv_diag(rgf_handle_t h, double *diag);

/* Batched forms: logdet of every matrix (count values) and diag(A^-1) of every
 * matrix (count x n, matrix b at b*n). */
int rgf_logdet_batch(rgf_handle_t h, double *out);
int rgf_selinv_diag_batch(rgf_handle_t h, double *diag);

/* RGF<T>::residualNorm(T* x, T* b), RGF.H:631-655: ||b - A x||_2 (needs keep_matrix). */
int rgf_residual_norm(rgf_handle_t h, const double *x, const double *b, double *out);

/* RGF<T>::residualNormNormalized(T* x, T* b), RGF.H:659-663. */
int rgf_residual_norm_normalized(rgf_handle_t h, const double *x, const double *b, double *out);

/* Device memory of handles comes from a per-device arena that keeps freed memory for
 * reuse (creating and destroying handles then calls neither hipMalloc nor hipFree, which
 * waits for the whole device).  Extension: the idle bytes the arena keeps per device
 * (default 4 GiB, or RGF_DEVICE_CACHE_GB in the environment); lowering it releases
 * wholly free memory beyond the new limit at once.  *previous (may be NULL) receives the
 * old limit. */
int rgf_set_device_cache_limit(size_t bytes, size_t *previous);
/* Extension: release every wholly free block of the arena of `device` (-1: all devices). */
int rgf_trim_device_cache(int device);

/* Extension (no counterpart in the reference, whose products are MAGMA calls): tile shapes
 * of the product launches, chosen per launch signature -- the transposes and, for each
 * product of the launch, m, n, k, its triangle flags and batch entries.  mode 0: fixed size
 * thresholds (default); 1: the cached choice of a signature where there is one, thresholds
 * otherwise; 2: as 1, and a launch whose signature is not cached first runs every candidate
 * tile shape on scratch outputs (its real outputs untouched; not under HIP graph capture,
 * where thresholds apply), times them with HIP events and caches the fastest.  Every tile
 * shape sums each output in the same order, so the choice never changes a result bit.
 * Process-wide; tune while one thread drives the device, so the timings are its own.
 * *previous (may be NULL) receives the old mode. */
int rgf_set_autotune(int mode, int *previous);
/* Number of cached signatures; clear the cache; save it as text, or add a saved one
 * (RGF_EINVAL for an unreadable or malformed file, the cache then unchanged). */
int rgf_autotune_entries(void);
int rgf_autotune_clear(void);
int rgf_autotune_save(const char *path);
int rgf_autotune_load(const char *path);

/* Message of the last failed call on h (or of the last failed create when h is NULL). */
const char *rgf_last_error(rgf_handle_t h);

/* ---- device-resident extensions (pointers are device memory of the handle's device) ---- */
int rgf_solve_device(rgf_handle_t h, double *x_dev, const double *b_dev, size_t nrhs);
/* No counterpart in the reference.  rgf_solve_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): it allocates nothing and keeps its solve buffers in a static 128 MiB device
 * array per GPU (held until the process ends).  Same results as rgf_solve_device, bit for
 * bit.  The handle must be factorized already (RGF_ESTATE otherwise); a solve whose
 * buffers exceed the array is rejected with RGF_EINVAL (FP64: n * 8 bytes per right-hand
 * side per pass, passes shrink to fit; refined mixed precision: 16 columns per pass, about
 * 3 * 16 * n * 8 bytes, 4 * 16 * n * 8 in place).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race. */
int rgf_solve_device_noalloc(rgf_handle_t h, double *x_dev, const double *b_dev, size_t nrhs);
/* diag(A^-1) of every matrix of the handle into device memory (count x n). */
int rgf_selinv_diag_device(rgf_handle_t h, double *diag_dev);
/* Extension (gradients of INLA-type objectives need more than the diagonal; the reference
 * forms only diag(A^-1), RGF.H:582-612): the blocks of A^-1 on the block pattern of A for the matrix
 * chosen by rgf_select, into device buffers (any may be NULL to skip it), column-major:
 *   diag_blocks   nt blocks Sigma_ii      (ns x ns each, block i at i*ns*ns)
 *   sub_blocks    nt-1 blocks Sigma_{i+1,i} (ns x ns each)
 *   arrow_blocks  nt blocks Sigma_{d,i}   (nd x ns each, block i at i*nd*ns)
 *   corner        Sigma_dd                (nd x nd)
 * Needs selinv_mode = RGF_SELINV_EXACT.  Asynchronous on the handle's stream (see
 * rgf_synchronize); also refreshes what rgf_selinv_diag returns. */
int rgf_selinv_blocks_device(rgf_handle_t h, double *diag_blocks, double *sub_blocks, double *arrow_blocks,
                             double *corner);
int rgf_synchronize(rgf_handle_t h);

/* ---- handle groups (extension; the reference handles one matrix per object, RGF.H:510-546) ----
 * Handles of uneven shapes (any ns, nt, nd, batch count and precision, one device) whose
 * factorizations or selected inversions run as one HIP graph launch: each handle's sweep is
 * a branch on its own stream, so the sweeps overlap on the device.  The graph is captured on
 * the first call and replayed by later ones (e.g. the re-factorizations of an INLA
 * hyperparameter loop after rgf_regenerate or new values); results are bitwise those of the
 * handles run alone.  The handles must outlive the group, and must not be used by another
 * thread during a group call.  Group calls are asynchronous like the handles' own; each
 * handle's stream continues after the group's work. */
typedef struct rgf_group_s *rgf_group_t;
int rgf_group_create(const rgf_handle_t *handles, size_t n, rgf_group_t *out);
int rgf_group_destroy(rgf_group_t g);
/* Factorize every handle (waits for the sweeps to check the pivots).  An in-place handle
 * (keep_matrix = 0) must hold a matrix, not its factor (RGF_ESTATE).  RGF_ENOTSPD if a
 * matrix is not positive definite: rgf_last_error of its handle names the column. */
int rgf_group_factorize(rgf_group_t g);
/* diag(A^-1) of every matrix of handle q into diag_devs[q] (device, count_q x n_q);
 * factorizes handles that are not factorized yet, first, one by one. */
int rgf_group_selinv_diag_device(rgf_group_t g, double *const *diag_devs);
int rgf_group_synchronize(rgf_group_t g);
/* Refinement of the last mixed-precision solve: steps taken and the final
 * max_c ||b_c - A x_c|| / ||b_c||. */
int rgf_refine_stats(rgf_handle_t h, int *iters, double *rel_residual);
int rgf_get_info(rgf_handle_t h, rgf_info_t *info);

/* A in the reference's MF layout (RGF.H:95 doubles; lower triangles of the diagonal
 * blocks and of the corner), for parity of the logical blocks.  Needs keep_matrix. */
int rgf_export_matrix_mf(rgf_handle_t h, double *mf);
/* The factor in MF layout; the diagonal blocks and the corner hold L_ii^-1 and L_dd^-1
 * (the representation kept on the device), the couplings B_i and arrows D_i as in L. */
int rgf_export_factor_mf(rgf_handle_t h, double *mf);

/* ---- RGF<std::complex<double>> (RGF.H:756-780 instantiates the class for CPX) ----
 * Complex-symmetric algebra, as the reference's complex block updates (transposes without
 * conjugation, RGF.H:260, 276-308): A = A^T, A = L L^T with L_jj the principal square root
 * of each pivot, logDet = 2 sum log L_jj (principal logarithms, so its imaginary part is not
 * reduced modulo 2 pi).  Complex arrays are interleaved (re, im) doubles, the layout of
 * std::complex<double>; the CSC arrays and shapes are as rgf_create_csc's.  FP64 only; the
 * matrix stays resident (residual, re-factorization); RGF_ENOTSPD reports a zero or
 * non-finite pivot (1-based column in rgf_z_last_error).  The handle is independent of the
 * FP64 handle type. */
typedef struct rgf_zhandle_s *rgf_zhandle_t;
/* RGF<CPX>::RGF(size_t* ia, size_t* ja, CPX* a, size_t ns, size_t nt, size_t nd), RGF.H:85-146 */
int rgf_z_create_csc(const size_t *ia, const size_t *ja, const double *a, size_t ns, size_t nt, size_t nd,
                     const rgf_options *opts, rgf_zhandle_t *out);
/* RGF<CPX>::~RGF(), RGF.H:151-170 */
int rgf_z_destroy(rgf_zhandle_t h);
/* RGF<CPX>::factorize(), RGF.H:510-546 */
int rgf_z_factorize(rgf_zhandle_t h);
/* RGF<CPX>::solve(CPX* x, CPX* b, size_t nrhs), RGF.H:550-578: n x nrhs, column-major, ld n */
int rgf_z_solve(rgf_zhandle_t h, double *x, const double *b, size_t nrhs);
/* CPX RGF<CPX>::logDet(), RGF.H:616-627: out[0] = real part, out[1] = imaginary part */
int rgf_z_logdet(rgf_zhandle_t h, double *out);
/* RGF<CPX>::RGFdiag(CPX* diag), RGF.H:582-612: diag(A^-1), n complex values */
int rgf_z_selinv_diag(rgf_zhandle_t h, double *diag);
/* RGF<CPX>::residualNorm(CPX* x, CPX* b), RGF.H:631-655: ||b - A x||_2 */
int rgf_z_residual_norm(rgf_zhandle_t h, const double *x, const double *b, double *out);
/* RGF<CPX>::residualNormNormalized(CPX* x, CPX* b), RGF.H:659-663 */
int rgf_z_residual_norm_normalized(rgf_zhandle_t h, const double *x, const double *b, double *out);
const char *rgf_z_last_error(rgf_zhandle_t h);

/* Per-process GEMM kernel timing with HIP events on the launching stream:
 * op_class 0 = NN, 1 = NT, 2 = TN, 3 = TT (FP64), 4..7 the same on split-BF16,
 * 8 = every split-BF16 class, 9 = every FP64 class, -1 = off.  read() synchronizes and
 * returns the summed kernel time (ms), the summed canonical flops and the launch count.
 * A measurement aid: enable it only while one thread drives the library. */
int rgf_kernel_timing_enable(int op_class);
int rgf_kernel_timing_read(double *ms, double *flops, long long *launches);

#ifdef __cplusplus
}
#endif
#endif /* RGF_H */
