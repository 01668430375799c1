// Internal declarations shared by the HIP translation units of librgf_hip.so.
// Everything here is gfx950 (CDNA4) code; no other target is built.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <functional>

#define RGF_HIP_CHECK(expr)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) return rgf::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

namespace rgf {

int hip_fail(hipError_t e, const char* what, const char* file, int line);

// ---------------------------------------------------------------- GEMM
// C = alpha * op(A) * op(B) + beta * Cin, column-major, FP64 on MFMA.
// Flags restrict the k-range per output tile for triangular operands (the
// operand must hold explicit zeros in its other triangle; skipping is an
// optimisation, not a mask) and restrict stores to the lower triangle of C.
enum GemmFlags : int {
    G_CLOWER = 1,   // store only C[i][j] with i >= j
    G_ALOWER = 2,   // op(A)[i][k] == 0 for k > i
    G_AUPPER = 4,   // op(A)[i][k] == 0 for k < i
    G_BLOWER = 8,   // op(B)[k][j] == 0 for k < j
    G_BUPPER = 16,  // op(B)[k][j] == 0 for k > j
};

struct GemmDesc {
    const double* A;
    const double* B;
    const double* Cin;   // may alias C; ignored when beta == 0
    double* C;
    long long sA, sB, sC, sCin;   // strides between batch entries (elements)
    int m, n, k;
    int lda, ldb, ldc, ldcin;
    int flags;
    int count;                    // batch entries described by this desc (inner level)
    int count2;                   // outer batch level (matrices of a batched handle), >= 1
    long long sA2, sB2, sC2, sCin2;
    double alpha, beta;
    // filled by the launcher
    int tiles_m, tiles_n, zbase, tri_tiles;
    int deal_rr;                  // 1: tiles dealt round-robin over the XCDs (no contiguous ranges)
    int pad_;
};

// Lower-triangular tile enumeration (GemmDesc::tri_tiles): tile g of the launch is (ti, tj),
// row by row.  Launches whose k-range depends on the tile (triangular operands) have tiles
// of unequal work; with deal_rr they keep the hardware's round-robin dealing of workgroups
// over the XCDs instead of giving each XCD a contiguous tile range.  For the lower-stored
// product with an upper op(A) (k from the tile row on) contiguous ranges give one XCD all
// the long rows: 2.3x the mean work at 45 tile rows; round-robin: 1.01x.
__device__ __forceinline__ bool tri_tile(int g, int T, int& ti, int& tj)
{
    if (g >= T * (T + 1) / 2) return false;
    ti = (int)((sqrt(8.0 * g + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > g) ti--;
    while ((ti + 1) * (ti + 2) / 2 <= g) ti++;
    tj = g - ti * (ti + 1) / 2;
    return true;
}

constexpr int GEMM_MAXD = 4;
struct GemmArgs {
    int nd;
    int fast_loads;   // FP64 kernel: full k-steps loaded through loop-invariant lane offsets
    int pad_[2];
    GemmDesc d[GEMM_MAXD];
};

enum class Op { N, T };

// Launch one GEMM "call" made of up to GEMM_MAXD descriptors that share the
// transposes (all entries are independent).  Returns hipError_t as int.
// cfg selects the tile shape (0 = 64x64, 1 = 128x64, 2 = 128x128, 3 = 32x32; -1 automatic).
int gemm_f64(Op ta, Op tb, GemmDesc* descs, int ndesc, hipStream_t s, int cfg = -1);
extern int g_gemm_fast_loads;   // 1: GemmArgs::fast_loads set on FP64 launches
extern int g_selinv_fuse;   // 1: the FP64 selected inversion launches X^T X and W^T Z together
extern int g_selinv_pipe;   // 1: ... and computes W of the next block in the launch of Z
extern int g_rr_tri;     // 1: lower-stored launches with triangular operands dealt round-robin
extern int g_rr_bands;   // 1: the same for full-tile launches with triangular operands
extern int g_small_tile_wgs;   // automatic tile shape: 32x32 below this many 64x64 workgroups
// tile order of a launch: triangular enumeration for lower-stored square tilings, and
// round-robin XCD dealing where the k-range depends on the tile
inline void set_tile_order(GemmDesc& d, bool tri)
{
    d.tri_tiles = tri ? 1 : 0;
    d.deal_rr = (d.flags & (G_ALOWER | G_AUPPER | G_BLOWER | G_BUPPER)) && (tri ? g_rr_tri : g_rr_bands) ? 1 : 0;
}

// Same interface on split-BF16 MFMA (three bf16 products per FP64 product, FP32
// accumulation inside a call, FP64 alpha/beta epilogue).
int gemm_bf16x3(Op ta, Op tb, GemmDesc* descs, int ndesc, hipStream_t s);

// Split-BF16 product with the operands split once, before the product, into bf16
// hi/lo copies in `scratch` (k padded to roundup(k, 64), 16 x 32 blocks).  One descriptor
// (count x count2 entries); scratch must hold split_scratch_elems(m, n, k, entries)
// bf16 values.  Same result as gemm_bf16x3 to FP32 accumulation order.
size_t split_scratch_elems(int m, int n, int k, int entries);
int gemm_bf16x3_split(Op ta, Op tb, GemmDesc* d, __bf16* scratch, size_t scratch_elems, hipStream_t s);
// The two halves of gemm_bf16x3_split, for operands used by several products: split
// op(X) (rows x k; op(X)[r][k] = X[r + k*ld], or X[k + r*ld] when kcontig) of
// count x count2 matrices (element strides s1, s2) into dst, split_entry_elems(rows, k)
// bf16 values per matrix; then a product on split copies: As holds op(A) (m x k, kcontig
// = TA), Bs holds op(B)^T (n x k, kcontig = !TB), both split with the product's k.
size_t split_entry_elems(int rows, int k);
int split_operand(const double* X, int ld, int rows, int k, bool kcontig, int count, long long s1, int count2,
                  long long s2, __bf16* dst, hipStream_t s);
// Cs (may be null): the product also writes the bf16 hi/lo split of its stored C into Cs, as
// split_operand(C, ldc, n, m, kcontig = true, ...) would (one entry per matrix of the
// descriptor, split_entry_elems(n, m) each): C used as op(B) (N) or op(A) (T) of a later
// product.  Cr (may be null): the same for C as op(A) (N) or op(B) (T), as
// split_operand(C, ldc, m, n, kcontig = false, ...) would (split_entry_elems(m, n) each).
// rowsA / rowsB: rows the operand copies were split for, when more than m / n (the copy of
// [B; D] read for B alone); 0 otherwise.
int gemm_bf16x3_presplit(Op ta, Op tb, const GemmDesc* d, const __bf16* As, const __bf16* Bs, hipStream_t s,
                         __bf16* Cs = nullptr, __bf16* Cr = nullptr, int rowsA = 0, int rowsB = 0);
// ---------------------------------------------------------------- split-BF16 copies
// Element (r, k) of one plane of a split copy whose k pitch is 32 * nkb: 16-row x 32-k blocks
// of 1 KB (block (r / 16, k / 32) at element ((r / 16) * nkb + k / 32) * 512), inside a block
// 64 chunks of 8 k, chunk 4 (r % 16) + ((k / 8) % 4 XOR (r / 4) % 4).
__host__ __device__ __forceinline__ long long split_index(int r, int k, int nkb)
{
    return ((long long)(r >> 4) * nkb + (k >> 5)) * 512 +
           ((((r & 15) << 2) | (((k >> 3) & 3) ^ ((r >> 2) & 3))) << 3) + (k & 7);
}
// x -> hi = x rounded once to bf16 (nearest, ties to even), f = (float)x, lo = bf16(f - hi):
// x = hi + lo to about 2^-16 relative.  hi goes through x rounded to float with round-to-odd
// (truncate, then set the last bit when inexact), which a second rounding to the shorter
// bf16 leaves equal to the single rounding of x; written out so that every kernel produces
// the same bits (left to the compiler, (bf16)(float)x is folded into one rounding in some
// kernels and rounded twice in others, which differ where (float)x lies on a bf16 tie).
__device__ __forceinline__ void split_bf16(double x, __bf16& hi, __bf16& lo)
{
    const float f = (float)x;
    unsigned u = __float_as_uint(f);
    const double fd = (double)f;
    if (fd != x && !(u & 1u)) u += fabs(x) > fabs(fd) ? 1u : 0xffffffffu;
    hi = (__bf16)__uint_as_float(u);
    lo = (__bf16)(f - (float)hi);
}
// A split copy written by a kernel beside its FP64 result: entry e at p + e * stride, lo
// plane at + lo, k pitch ldk (p null: none)
struct SplitOut {
    __bf16* p;
    long long stride, lo;
    int ldk;
};
__device__ __forceinline__ void split_store(const SplitOut& o, long long e, int r, int k, double v)
{
    __bf16 h, l;
    split_bf16(v, h, l);
    const long long i = e * o.stride + split_index(r, k, o.ldk >> 5);
    o.p[i] = h;
    o.p[o.lo + i] = l;
}

// A split copy, or a block of one, as an operand or destination of a pre-split product: its
// origin at p (a multiple of 16 rows and 32 k inside the copy), lo plane at p + lo, k pitch
// ldk (a multiple of 64), rows readable from p; entry (b1, b2) of the product's descriptor
// (b1 < count, b2 < count2) at p + b1 * s1 + b2 * s2.  pad (destinations): write zeros in
// the copy's k beyond C (a whole copy, not a block of a larger one).
struct SplitView {
    __bf16* p;
    long long lo, s1, s2;
    int ldk, rows, pad;
};
// the layout split_operand writes: rows x k per entry, count x count2 entries one after another
// (count: the inner entry count of the product's descriptor, GemmDesc::count)
SplitView split_view(__bf16* p, int rows, int k, int count);
// element (r, k) of entry (b1, b2) of a view, both planes
__device__ __forceinline__ void split_store_v(const SplitView& v, long long b1, long long b2, int r, int k, double x)
{
    __bf16 h, l;
    split_bf16(x, h, l);
    const long long i = b1 * v.s1 + b2 * v.s2 + split_index(r, k, v.ldk >> 5);
    v.p[i] = h;
    v.p[v.lo + i] = l;
}
// the block of a view at row r, k c (multiples of 16 and 32), entry (b1, b2) steps s1, s2
inline SplitView split_subview(const SplitView& v, int r, int c, long long s1, long long s2)
{
    SplitView o = v;
    o.p = v.p + ((long long)(r >> 4) * (v.ldk >> 5) + (c >> 5)) * 512;
    o.rows = v.rows - r;
    o.s1 = s1;
    o.s2 = s2;
    o.pad = 0;
    return o;
}
// element offset of the block at row r, k c of a view (for entry steps of sub-blocks)
inline long long split_block_offset(const SplitView& v, int r, int c)
{
    return ((long long)(r >> 4) * (v.ldk >> 5) + (c >> 5)) * 512;
}
// Up to two products of the same transposes in one launch (the tiles of both in one grid).
// cs / cr (p null: none): the product's stored C as op(C)^T / op(C) copies (see Cs / Cr).
struct PresplitCall {
    GemmDesc d;
    SplitView a, b, cs, cr;
};
int gemm_bf16x3_presplit_n(Op ta, Op tb, const PresplitCall* calls, int ncalls, hipStream_t s);
extern int g_presplit_kb;   // k per LDS stage of gemm_bf16x3_presplit (32 or 64)
extern int g_presplit_swizzle;   // 1: its LDS rows unpadded with XOR-swizzled chunks
extern int g_wide_min_wgs;       // its 256 x 128 kernel for launches of at least this many workgroups

// Tile-shape choice of a product launch (kind 0: FP64 GEMM, 1: pre-split split-BF16): the
// cached configuration of its signature, or -1 (then the launcher's size thresholds apply).
// With tuning on and the signature not cached, every candidate (cands[0]: the threshold
// choice) is first run through `run` on descriptors whose outputs are redirected to
// scratch, timed, and the fastest cached.  Not inside HIP graph capture.
int tune_choose(int kind, Op ta, Op tb, const GemmDesc* d, int nd, const int* cands, int ncand, hipStream_t s,
                const std::function<int(int, const GemmDesc*)>& run);
bool tune_in_trial();   // a trial launch of tune_choose is being issued (no kernel timing)

// Kernel-timing hooks shared by the GEMM launchers (op class: 0..3 FP64 NN/NT/TN/TT,
// 4..7 the same transposes on split-BF16).
void* timing_begin(int op_class, const GemmDesc* d, int nd, hipStream_t s);
void timing_end(void* token, hipStream_t s);
// Changes whenever kernel timing is (re)configured: a HIP graph captured with the
// timing events of an earlier epoch must be captured again.
long long timing_epoch();

// Convenience for a single, non-batched product.
inline GemmDesc gdesc(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int m,
                      int n, int k, double alpha, double beta, int flags = 0)
{
    GemmDesc d{};
    d.A = A; d.B = B; d.C = C; d.Cin = C;
    d.lda = lda; d.ldb = ldb; d.ldc = ldc; d.ldcin = ldc;
    d.m = m; d.n = n; d.k = k; d.alpha = alpha; d.beta = beta; d.flags = flags; d.count = 1; d.count2 = 1;
    return d;
}

// Repeat a descriptor over `count` matrices at strides (A, B, C; Cin follows C).
inline void batch2(GemmDesc& d, int count, long long sA, long long sB, long long sC)
{
    d.count2 = count; d.sA2 = sA; d.sB2 = sB; d.sC2 = sC; d.sCin2 = sC;
}

// ---------------------------------------------------------------- dense block kernels
// Every launcher below works on `count` matrices at the given element strides
// (s* arguments); info is one int per matrix.
// xdiag: T = ceil(n/64) tiles of 64 x 64 per matrix receiving the inverted diagonal tiles.
// The diagonal tiles of A receive L only if store_diag_l (the factor sweep replaces
// them by their inverses from xdiag and skips the store).
// dag_flags (potrf_dag_flag_elems(n) ints per matrix, zero before the first call; each call
// leaves them zero): the factorization runs as one task-graph launch per call (unless
// store_diag_l or the batch is too large for it); a wait of that launch that expires sets
// the matrix's info to INT_MIN.
// With xbuf (potrf_dag_xbuf_elems(n) doubles per matrix at stride sXb) and x_done, the
// task-graph launch also forms X = L^-1 and leaves it in the block (as trtri_f64_inplace
// would; then *x_done is set).
// schedule: rgf_cholesky_schedule (0 auto: potrf_uses_dag(count)).
int potrf_f64(double* A, int lda, long long sA, int n, int count, double* logdet_parts, long long sP, int* info,
              int info_base, double* xdiag, long long sX, bool store_diag_l, hipStream_t s, int* dag_flags = nullptr,
              double* xbuf = nullptr, long long sXb = 0, bool* x_done = nullptr, int schedule = 0);
bool potrf_dag_schedule(int schedule, int count);
size_t potrf_dag_flag_elems(int n);
// the workers' claim order for T tile rows (5 ints per task); returns the task count
int potrf_dag_order(int T, bool with_x, int xlag, bool promo, int* out, int cap);
size_t potrf_dag_xbuf_elems(int n);
bool potrf_uses_dag(int count);   // a launch over `count` matrices takes the task-graph path
extern int g_chol_dag, g_chol_dag_wgs, g_chol_dag_min, g_chol_dag_x, g_chol_dag_xlag, g_chol_dag_promo;
extern long long* g_chol_dag_trace;   // device buffer: [0] records, then 8 longs per task (tuning)
extern long long g_chol_dag_trace_cap;
// A product launcher supplied by the caller (precision policy of a handle).
struct ProdFn {
    int (*fn)(void* ctx, Op ta, Op tb, GemmDesc* d, int nd, hipStream_t s);
    void* ctx;
};
// Split copies of X = L^-1 that the mixed path's inverse keeps up to date (see trtri below):
// xr holds X's rows (element (r, k) = X[r][k]), xc X's columns (element (j, k) = X[k][j]),
// each a whole copy per matrix (entry b at p + b * s1) with xcopy_rows(n) rows; scratch holds
// the split copies of one level's L21 and T blocks (trtri_scratch_elems(n, count) bf16).
struct TrtriCopies {
    SplitView xr, xc;
    __bf16* scratch;
    size_t scratch_elems;
};
inline int xcopy_rows(int n) { return (n + 127) / 128 * 128 + 128; }
size_t trtri_scratch_elems(int n, int count);
// In-place L^-1; xdiag (may be null) holds the diagonal tiles already inverted.  The
// merge products run through `prod` (null: FP64 GEMM).  plane != 0: complex L stored as
// two planes (imaginary part `plane` elements after the real part, for A, tmp and xdiag
// alike); the diagonal tiles of both planes are placed and `prod` multiplies complex.
// cp (mixed path): the merges of blocks of at least 128 rows run as pre-split products that
// read X11 and X22 from the copies and write each X21 into both of them (only L21 and T are
// split); the 128-row diagonal blocks, their zeros above the diagonal and the k padding
// enter the copies from a pass after the smaller merges.  Regions above the diagonal outside
// those blocks are never written (a product with triangle flags never reads them).
int trtri_f64_inplace(double* A, int lda, long long sA, int n, int count, double* tmp, long long sT,
                      const double* xdiag, long long sX, hipStream_t s, const ProdFn* prod = nullptr,
                      long long plane = 0, const TrtriCopies* cp = nullptr);
int potrf_tile_only(double* A, int lda, double* xdiag, double* logdet_slot, int* info, hipStream_t s);
// split copies of X = L^-1 (n x n at A, zeros above the diagonal) as trtri_f64_inplace's cp
// leaves them: xr (rows) and xc (columns; xc.p null: none)
int xsplit_copies(const double* A, int lda, long long sA, int n, int count, const SplitView& xr, const SplitView& xc,
                  hipStream_t s);
int copy2d(const double* src, int lds, double* dst, int ldd, int rows, int cols, double alpha,
           hipStream_t s, int count = 1, long long sS = 0, long long sD = 0);
int fill2d(double* dst, int ldd, int rows, int cols, double v, hipStream_t s, int count = 1, long long sD = 0);
// Corner: Cout = Cin - sum_i D_i D_i^T (D_i: nd x ns at Dbase + i*dstride, ld ldd),
// then in-place Cholesky of the nd x nd corner, its inverse X_dd (written to Xdd),
// log-det part and info.  nd <= 64; Cin, Ldd, Xdd are nd x nd (ld nd) per matrix.
int corner_factor(const double* Dbase, long long dstride, long long sM, int ldd, int ns, int nt, int nd,
                  const double* Cin, int ldcin, double* Ldd, double* Xdd, double* partial, double* logdet_slot,
                  long long sP, int* info, int info_base, int count, hipStream_t s);

// ---------------------------------------------------------------- skinny products (solves)
// Y = alpha * op(A) * R + beta * Y0, A is m x k (N) or k x m (T) col-major,
// R is k x nr (ld ldr), Y is m x nr (ld ldy), nr <= 16.  tri: 0 none, 1 = op(A) lower,
// 2 = op(A) upper.  Uses a split-k partial buffer `work` of size >= skinny_work(m, k, nr).
size_t skinny_work_elems(int m, int k, int nr);
// Partials of skinny_f64s over `count` matrices of one product shape (the split of k
// into slabs depends on m and k only, so each matrix is summed as it would be alone).
size_t skinny_batch_work_elems(Op ta, int m, int k, int count);
// work_elems: capacity of `work` (0: skinny_work_elems(m, k, nr), the most any call needs)
// ctr (may be null): SKINNY_CTRS zero ints, left zero; with them the split-k partials are
// reduced inside the product launch (same sums, same order), else by a second launch.
constexpr int SKINNY_CTRS = 16384;
int skinny_f64(Op ta, int tri, int m, int k, int nr, double alpha, const double* A, int lda,
               const double* R, int ldr, double beta, const double* Y0, int ldy0, double* Y, int ldy,
               double* work, hipStream_t s, size_t work_elems = 0, int* ctr = nullptr, bool wide = false);
// Partial doubles one call of this shape with nr right-hand sides actually writes (0 when
// k is not split).
size_t skinny_call_work(Op ta, int m, int k, int nr);
extern int g_skinny_target_wg;
// Same with element strides: R(kk, c) = R[kk * rsk + c * rsj], Y(i, c) = Y[i * ysi + c * ysj]
// (Y0 likewise), e.g. row-major operands as transposed views, over `count` matrices at
// strides sA, sR, sY0, sY in one launch; work holds work_elems doubles (>= count * 16 * m).
int skinny_f64s(Op ta, int tri, int m, int k, int nr, double alpha, const double* A, int lda, const double* R,
                long long rsk, long long rsj, double beta, const double* Y0, long long y0si, long long y0sj,
                double* Y, long long ysi, long long ysj, double* work, size_t work_elems, hipStream_t s,
                int count = 1, long long sA = 0, long long sR = 0, long long sY0 = 0, long long sY = 0,
                int* ctr = nullptr, const double* R2 = nullptr, int k1 = 0, double* Y2 = nullptr, int m1 = 0,
                int slab_cap = 64, bool wide = false);
// (R2, k1): rows kk >= k1 of R read from R2 (row kk - k1, R's strides); (Y2, m1): rows
// i >= m1 of Y (and of Y0) are those of Y2 (row i - m1, Y's strides).  Unbatched only.
// slab_cap: at most this many k-slabs (1: no partials -- for launches over many blocks).
// wide: the MFMA kernel (that of three or more columns) for any column count, so a column
// is summed the same way whatever the size of the pass that carries it.

// ---------------------------------------------------------------- generator / ingestion
struct GenParams {
    uint64_t seed;
    uint32_t matrix_id;
    int ns, nt, nd;
    double sigma_c, sigma_a, theta_d, theta_c;
};
// Writes A (lower diag blocks mirrored to full, couplings, arrow, lower corner mirrored)
// into panels P (panel stride pstride, ld ldp) and corner C (ld nd).  tmp: ns*ns doubles,
// rowsum: n doubles.
int generate_matrix(const GenParams& g, double* P, long long pstride, int ldp, double* Cdd,
                    double* tmp, double* rowsum, hipStream_t s);
double gen_uniform_host(uint64_t seed, uint32_t mid, uint32_t idx);
int scatter_csc(const uint64_t* ia, const uint64_t* ja, const double* a, long long nnz, int ns,
                int nt, int nd, double* P, long long pstride, int ldp, double* Cdd, hipStream_t s);
int mirror_lower(double* A, int lda, int n, hipStream_t s);

// Mirror the lower triangle of A into its upper one (and gather the diagonal); with X the
// lower triangle first receives A + X; with cp (mixed path) the split copy of the whole
// symmetric n x n result as op(A) rows (row r, k c) is written too.
int mirror_lower_diag(double* A, int lda, long long sA, int n, double* diag, long long sDg, int count, hipStream_t s,
                      const double* X = nullptr, int ldx = 0, long long sX = 0, const SplitOut* cp = nullptr);

// ---------------------------------------------------------------- selinv helpers
// Finish one Takahashi step: (with X) add the lower triangle of X into that of S, mirror
// the lower triangle of S (ns x ns, ld lds) into its upper triangle, write diag(S) to
// diag_out, and (exact mode) store S_{d,i} = -Z[zrow0:zrow0+nd, :] into S[ns:ns+nd, 0:ns]
// and its transpose.
// With cp, the split copy of S's rows 0 .. ns + nd - 1 (as op(A) rows) receives what these
// two steps write (the corner and the k padding of the copy are the caller's).
int selinv_finish(double* S, int lds, long long sS, int ns, int nd, const double* Z, int ldz, long long sZ,
                  int zrow0, int exact, double* diag_out, long long sDg, int count, hipStream_t s,
                  const double* X = nullptr, int ldx = 0, long long sX = 0, const SplitOut* cp = nullptr);
int extract_diag(const double* A, int lda, int n, double* out, hipStream_t s, int count = 1, long long sA = 0,
                 long long sO = 0);

}  // namespace rgf
