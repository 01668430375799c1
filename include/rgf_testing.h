/*
 * rgf_testing.h -- diagnostic entry points of librgf_hip.so used by the test
 * suite.  They have no counterpart in rjanalik/RGF: they expose single kernels
 * of the sweep so a failure can be localised to one kernel.
 */
#ifndef RGF_TESTING_H
#define RGF_TESTING_H

#ifdef __cplusplus
extern "C" {
#endif

/* C = alpha*op(A)*op(B) + beta*C on the device FP64 MFMA GEMM kernel.  Host
 * arrays, column-major; ta/tb: 0 = N, 1 = T; flags as the kernel's GemmFlags
 * (1 lower-C stores, 2/4 op(A) lower/upper, 8/16 op(B) lower/upper); flag 256 runs the
 * split-BF16 kernel of the mixed-precision path instead of the FP64 one, flag 512
 * the split-BF16 product with operands split into bf16 hi/lo copies first; bits 12-14
 * force the FP64 tile shape (1 = 64x64, 2 = 128x64, 3 = 128x128, 4 = 32x32; 0 automatic).
 * A is (ta ? k x m : m x k) with ld lda, B is (tb ? n x k : k x n) with ld ldb. */
int rgf_debug_gemm(int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
                   const double *B, int ldb, double beta, double *C, int ldc, int flags);

/* In-place Cholesky (lower) of an n x n host matrix on the device, followed by
 * L^-1 in place when invert != 0; returns the log-determinant of the input in
 * *logdet and RGF_ENOTSPD for a non-positive pivot. */
int rgf_debug_potrf(int n, double *A, int lda, int invert, double *logdet);

/* Device arena counters since load: allocations served from existing slabs, hipMalloc
 * calls (slabs added), hipFree calls (wholly free slabs released over the cap) and the
 * free bytes held now (GiB). */
int rgf_debug_cache_stats(long long *hits, long long *mallocs, long long *frees, double *cached_gb);

/* Tuning loops: average device time (ms) of one GEMM launch of the given shape
 * (operands allocated and filled once; flags as rgf_debug_gemm), and of the blocked
 * Cholesky of an n x n block and of its 64 x 64 tile kernel alone. */
int rgf_debug_bench_gemm(int ta, int tb, int m, int n, int k, int flags, int reps, double *ms_per_launch);
int rgf_debug_bench_potrf(int n, int reps, double *ms_potrf, double *ms_tile);
/* One task-graph Cholesky of the bench_potrf block with its tasks traced: up to cap records
 * of 8 values (workgroup, kind 0 chain step / 1 TRSM / 2 UPDATE / 3 the chain's next-tile
 * update, i, j, generation, wall-clock stamps at claim, dependencies met, done) into out;
 * *nrec receives the number of tasks traced and *ticks_per_us the clock rate. */
int rgf_debug_trace_potrf(int n, long long *out, long long cap, long long *nrec, double *ticks_per_us);

/* The claim order of the task-graph Cholesky's workers for T tile rows (with_x: the L^-1 tasks
 * too; xlag, promo: the order's two parameters, as tuning keys 20 and 21 set them): task t as
 * out[5 t ..] = kind (1 TRSM, 2 UPDATE, 6 XACC), i, j, p, generation -- the same decoding the
 * kernel's workers run.  Returns the task count (at most cap tasks written); host only. */
int rgf_debug_dag_order(int T, int with_x, int xlag, int promo, int *out, int cap);

/* Process-wide kernel variant switches, for measurements of whole sweeps: key 1 = k per
 * LDS stage of the pre-split split-BF16 product (32 or 64); key 3 = its LDS row layout
 * (0 padded rows, 1 unpadded rows with XOR-swizzled 16-byte chunks); key 4 = FP64 GEMM loads of
 * full k-steps through loop-invariant lane offsets (1) or per-element masked loads (0); key 8 =
 * lower-stored products with triangular operands keep the round-robin dealing of tiles over
 * the XCDs (1) or take the XCD-contiguous tile ranges of the other launches (0); key 9 = the
 * same choice for full-tile products with triangular operands; key 11 = the FP64 selected
 * inversion launches X_i^T X_i and W^T Z as one launch of two products (1) or one after the
 * other (0); key 12 = with key 11 on, the FP64 sweep also computes W = F X of the next block
 * in the launch of this block's Z = S W (1) or in a launch of its own (0); key 13 = FP64 launches of
 * fewer 64 x 64 workgroups than the value take 32 x 32 tiles (0: never; any value >= 0); key 14 =
 * pre-split split-BF16 products without triangle flags and with at least this many 256 x 128
 * workgroups take the 256 x 128 kernel (0: every product without triangular operands, lower-
 * stored ones included; any value >= 0, 2^30 in effect never); key 15 = the mixed path's merges of L_ii^-1 on split-BF16 (1) or FP64 (0)
 * products.  Returns the previous value,
 * or -1 for an unknown key or value (nothing changed).  Not for concurrent use with
 * running handles. */
int rgf_debug_set_tuning(int key, int value);

#ifdef __cplusplus
}
#endif
#endif
