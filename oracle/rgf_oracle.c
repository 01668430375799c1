/*
 * CPU oracle for the RGF hot path.  TEST INFRASTRUCTURE ONLY: it is loaded by
 * tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
 * cpu_baseline leg.  The product (rgf_amd/) never links, loads or calls it.
 *
 * It restates, call by call on host LAPACK/BLAS (MKL libmkl_rt, LP64), the
 * block algorithm of rjanalik/RGF over the reference's packed "MF" layout:
 *
 *   layout / diag_pos ........ RGF.H:85-138, RGF.H:174-229
 *   FirstStageFactor ......... RGF.H:233-340
 *   SecondStageSolve ......... RGF.H:344-418
 *   ThirdStageRGF (compat) ... RGF.H:422-506
 *   logDet ................... RGF.H:616-627, Blas.H:1084-1092
 *   residualNorm(.Normalized)  RGF.H:631-663
 *   CSC -> MF scatter ........ CWC_utility.cu:998-1029 (getPos), :1085-1116
 *
 * Deliberate deviations from the reference, each also pinned by a dense check:
 *   - nt == 1 with nd > 0: the reference skips the arrow-row trsm
 *     (guard at RGF.H:258) and produces a wrong factor; the oracle solves the
 *     nd arrow rows against L_00 as every other block does.
 *   - rgf_oracle_selinv(..., exact=1) runs the complete Takahashi recursion
 *     (with the two cross terms D_i^T S_{d,i+1} B_i + B_i^T S_{d,i+1}^T D_i
 *     that RGF.H:476-503 omits); exact=0 reproduces RGF.H:422-506 as written.
 *
 * Pins (tests/test_oracle.py): dense numpy Cholesky equality of every block
 * (the block-arrow pattern has no fill), numpy slogdet, numpy solve and the
 * residual, diag(inv(A)) for the exact mode and for nd = 0, the reference's
 * own restated formula evaluated densely in numpy for the compat mode, the
 * logdet goldens of the reference's data/ fixtures, and Philox4x32-10
 * known-answer vectors (Salmon et al., SC'11, Random123 kat_vectors).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define RGF_TWO_PI 6.283185307179586476925286766559

/* ---- host BLAS / LAPACK (Fortran ABI, LP64) ---------------------------- */
extern void dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info);
extern void dtrtri_(const char *uplo, const char *diag, const int *n, double *a, const int *lda, int *info);
extern void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag,
                   const int *m, const int *n, const double *alpha, const double *a, const int *lda,
                   double *b, const int *ldb);
extern void dgemm_(const char *ta, const char *tb, const int *m, const int *n, const int *k,
                   const double *alpha, const double *a, const int *lda, const double *b,
                   const int *ldb, const double *beta, double *c, const int *ldc);
extern double dnrm2_(const int *n, const double *x, const int *incx);

static void gemm(char ta, char tb, long m, long n, long k, double alpha, const double *a, long lda,
                 const double *b, long ldb, double beta, double *c, long ldc)
{
    if (m <= 0 || n <= 0) return;
    int im = (int)m, in = (int)n, ik = (int)k, ila = (int)lda, ilb = (int)ldb, ilc = (int)ldc;
    if (k <= 0) { /* C := beta*C */
        for (long j = 0; j < n; j++)
            for (long i = 0; i < m; i++) c[i + j * ldc] *= beta;
        return;
    }
    dgemm_(&ta, &tb, &im, &in, &ik, &alpha, a, &ila, b, &ilb, &beta, c, &ilc);
}

static void trsm(char side, char uplo, char ta, char diag, long m, long n, double alpha,
                 const double *a, long lda, double *b, long ldb)
{
    if (m <= 0 || n <= 0) return;
    int im = (int)m, in = (int)n, ila = (int)lda, ilb = (int)ldb;
    dtrsm_(&side, &uplo, &ta, &diag, &im, &in, &alpha, a, &ila, b, &ilb);
}

static int potrf(long n, double *a, long lda)
{
    if (n <= 0) return 0;
    int in = (int)n, ila = (int)lda, info = 0;
    dpotrf_("L", &in, a, &ila, &info);
    return info;
}

static int trtri(long n, double *a, long lda)
{
    if (n <= 0) return 0;
    int in = (int)n, ila = (int)lda, info = 0;
    dtrtri_("L", "N", &in, a, &ila, &info);
    return info;
}

/* ---- MF layout (RGF.H:85-138, 174-229) --------------------------------- */
typedef struct {
    long ns, nt, nd, n, nblock, bsize;
    long *diag_pos;
} layout_t;

static int layout_init(layout_t *L, long ns, long nt, long nd)
{
    L->ns = ns; L->nt = nt; L->nd = nd;
    L->n = ns * nt + nd;                                   /* RGF.H:94 */
    L->nblock = nd > 0 ? nt + 1 : nt;                       /* RGF.H:97-100 */
    L->bsize = ns > nd ? ns : nd;                           /* create_blocks, RGF.H:667-680 */
    L->diag_pos = (long *)malloc(sizeof(long) * (size_t)L->n);
    if (!L->diag_pos) return -1;
    for (long ib = 0; ib + 1 < nt; ib++)                    /* lda = 2ns+nd, RGF.H:117-123 */
        for (long i = 0; i < ns; i++)
            L->diag_pos[ib * ns + i] = ib * ns * (2 * ns + nd) + i * (2 * ns + nd + 1);
    for (long i = 0; i < ns; i++)                            /* lda = ns+nd, RGF.H:125-129 */
        L->diag_pos[(nt - 1) * ns + i] = (nt - 1) * ns * (2 * ns + nd) + i * (ns + nd + 1);
    for (long i = 0; i < nd; i++)                            /* corner, lda = nd, RGF.H:131-138 */
        L->diag_pos[nt * ns + i] = (nt - 1) * ns * (2 * ns + nd) + ns * (ns + nd) + i * (nd + 1);
    return 0;
}

static void layout_free(layout_t *L) { free(L->diag_pos); L->diag_pos = NULL; }

/* mf_block_lda, RGF.H:182-195 (r unused, as in the reference) */
static long lda_of(const layout_t *L, long c)
{
    if (c < L->nt - 1) return 2 * L->ns + L->nd;
    if (c < L->nt) return L->ns + L->nd;
    return L->nd;
}
/* start of supernode c (its diagonal block), mf_block_index(c,c), RGF.H:174-178 */
static long snode(const layout_t *L, long c) { return L->diag_pos[c * L->bsize]; }
/* mf_dense_block_offset, RGF.H:212-221: row offset of the arrow rows in supernode i */
static long arrow_off(const layout_t *L, long i)
{
    if (i < L->nt - 1) return 2 * L->bsize;
    if (i == L->nt - 1) return L->bsize;
    return 0;
}

long rgf_oracle_mf_size(long ns, long nt, long nd)
{
    long n = ns * nt + nd;
    return 2 * (nt - 1) * ns * ns + ns * ns + n * nd;       /* RGF.H:95 */
}

int rgf_oracle_diag_pos(long ns, long nt, long nd, long *out)
{
    layout_t L;
    if (layout_init(&L, ns, nt, nd)) return -1;
    memcpy(out, L.diag_pos, sizeof(long) * (size_t)L.n);
    layout_free(&L);
    return 0;
}

/* getPos, CWC_utility.cu:998-1029: offset of lower-triangle entry (r,c) in MF */
long rgf_oracle_getpos(long r, long c, long ns, long nt, long nd)
{
    long ib = c / ns;
    long b2 = (ib < nt - 1) ? c : (nt - 1) * ns;
    long b1 = (ib < nt - 1) ? 0 : (ib == nt - 1 ? c % ns : ns);
    long bd = (c > ns * nt) ? c - ns * nt : 0;
    long pos;
    if (r < ns * nt) pos = r - ib * ns;
    else if (ib < nt - 1) pos = r - (nt - 2) * ns;
    else if (ib == nt - 1) pos = r - (nt - 1) * ns;
    else pos = r - nt * ns;
    return b2 * (2 * ns + nd) + b1 * (ns + nd) + bd * nd + pos;
}

/* Whole-matrix CSC (lower triangle, 0-based, size_t indices; main.C:50-70)
 * scattered into a zeroed MF array (init_supernode, RGF.H:684-700).
 * Returns 0, or -(k+1) for the first nonzero k outside the block-arrow
 * pattern (the reference scatters such entries to wrong places silently). */
long rgf_oracle_csc_to_mf(const uint64_t *ia, const uint64_t *ja, const double *a,
                          long ns, long nt, long nd, double *mf)
{
    long n = ns * nt + nd;
    memset(mf, 0, sizeof(double) * (size_t)rgf_oracle_mf_size(ns, nt, nd));
    for (long c = 0; c < n; c++) {
        for (uint64_t k = ia[c]; k < ia[c + 1]; k++) {
            long r = (long)ja[k];
            long bc = c / ns, br = r / ns;
            int ok;
            if (r < c || r >= n) ok = 0;
            else if (c >= ns * nt) ok = 1;                       /* corner */
            else if (r >= ns * nt) ok = 1;                       /* arrow row */
            else ok = (br == bc) || (br == bc + 1);              /* diag or coupling */
            if (!ok) return -(long)k - 1;
            mf[rgf_oracle_getpos(r, c, ns, nt, nd)] = a[k];
        }
    }
    return 0;
}

/* ---- FirstStageFactor, RGF.H:233-340 (in place on MF) ------------------ */
int rgf_oracle_factor(double *mf, long ns, long nt, long nd)
{
    layout_t L;
    if (layout_init(&L, ns, nt, nd)) return -1;
    int info = 0, bad = 0;
    double *corner = nd > 0 ? mf + L.diag_pos[nt * ns] : NULL;

    /* block 0: potrf, then trsm of [A_10; A_d0] (RGF.H:245-263) */
    double *R = mf + snode(&L, 0);
    info = potrf(ns, R, lda_of(&L, 0));
    if (info && !bad) bad = (int)info;
    if (nt > 1)
        trsm('R', 'L', 'T', 'N', ns + nd, ns, 1.0, R, lda_of(&L, 0), R + ns, lda_of(&L, 0));
    else if (nd > 0) /* deviation: RGF.H:258 skips this, leaving D_0 unsolved */
        trsm('R', 'L', 'T', 'N', nd, ns, 1.0, R, lda_of(&L, 0), R + arrow_off(&L, 0), lda_of(&L, 0));

    for (long ib = 1; ib < nt; ib++) {          /* RGF.H:266-322 */
        double *M = mf + snode(&L, ib - 1);
        R = mf + snode(&L, ib);
        long ldm = lda_of(&L, ib - 1), ldr = lda_of(&L, ib);
        /* A_ii -= B_{i-1} B_{i-1}^T (full square, RGF.H:276) */
        gemm('N', 'T', ns, ns, ns, -1.0, M + ns, ldm, M + ns, ldm, 1.0, R, ldr);
        if (nd > 0) {
            /* A_di -= D_{i-1} B_{i-1}^T (RGF.H:279), A_dd -= D_{i-1} D_{i-1}^T (RGF.H:280) */
            gemm('N', 'T', nd, ns, ns, -1.0, M + arrow_off(&L, ib - 1), ldm, M + ns, ldm, 1.0,
                 R + arrow_off(&L, ib), ldr);
            gemm('N', 'T', nd, nd, ns, -1.0, M + arrow_off(&L, ib - 1), ldm, M + arrow_off(&L, ib - 1),
                 ldm, 1.0, corner, nd);
        }
        info = potrf(ns, R, ldr);
        if (info && !bad) bad = (int)(ib * ns + info);
        if (ib < nt - 1)   /* [B_i; D_i] = [A_{i+1,i}; A_di] L_ii^-T (RGF.H:287) */
            trsm('R', 'L', 'T', 'N', ns + nd, ns, 1.0, R, ldr, R + ns, ldr);
        else if (nd > 0)   /* last block: arrow rows only (RGF.H:315-318) */
            trsm('R', 'L', 'T', 'N', nd, ns, 1.0, R, ldr, R + arrow_off(&L, ib), ldr);
    }
    if (nd > 0) {                               /* corner, RGF.H:324-339 */
        R = mf + snode(&L, nt - 1);
        gemm('N', 'T', nd, nd, ns, -1.0, R + arrow_off(&L, nt - 1), lda_of(&L, nt - 1),
             R + arrow_off(&L, nt - 1), lda_of(&L, nt - 1), 1.0, corner, nd);
        info = potrf(nd, corner, nd);
        if (info && !bad) bad = (int)(nt * ns + info);
    }
    layout_free(&L);
    return bad;   /* 0, or 1-based global column of the first non-positive pivot */
}

/* ---- logDet, RGF.H:616-627 (sequential sum of log L_kk, times 2) ------- */
double rgf_oracle_logdet(const double *mf, long ns, long nt, long nd)
{
    layout_t L;
    if (layout_init(&L, ns, nt, nd)) return NAN;
    double s = 0.0;
    for (long k = 0; k < L.n; k++) s += log(mf[L.diag_pos[k]]);
    layout_free(&L);
    return 2.0 * s;
}

/* ---- SecondStageSolve, RGF.H:344-418; rhs is n x nrhs, ld = n ---------- */
int rgf_oracle_solve(const double *mf, long ns, long nt, long nd, double *rhs, long nrhs)
{
    layout_t L;
    if (layout_init(&L, ns, nt, nd)) return -1;
    long n = L.n;
    /* forward */
    trsm('L', 'L', 'N', 'N', ns, nrhs, 1.0, mf + snode(&L, 0), lda_of(&L, 0), rhs, n);
    for (long ib = 1; ib < nt; ib++) {
        gemm('N', 'N', ns, nrhs, ns, -1.0, mf + snode(&L, ib - 1) + ns, lda_of(&L, ib - 1),
             rhs + (ib - 1) * ns, n, 1.0, rhs + ib * ns, n);
        trsm('L', 'L', 'N', 'N', ns, nrhs, 1.0, mf + snode(&L, ib), lda_of(&L, ib), rhs + ib * ns, n);
    }
    if (nd > 0) {
        for (long i = 0; i < nt; i++)
            gemm('N', 'N', nd, nrhs, ns, -1.0, mf + snode(&L, i) + arrow_off(&L, i), lda_of(&L, i),
                 rhs + i * ns, n, 1.0, rhs + nt * ns, n);
        trsm('L', 'L', 'N', 'N', nd, nrhs, 1.0, mf + L.diag_pos[nt * ns], nd, rhs + nt * ns, n);
        /* backward, arrow first (RGF.H:386-400) */
        trsm('L', 'L', 'T', 'N', nd, nrhs, 1.0, mf + L.diag_pos[nt * ns], nd, rhs + nt * ns, n);
        for (long i = 0; i < nt; i++)
            gemm('T', 'N', ns, nrhs, nd, -1.0, mf + snode(&L, i) + arrow_off(&L, i), lda_of(&L, i),
                 rhs + nt * ns, n, 1.0, rhs + i * ns, n);
    }
    for (long ib = nt - 1; ib > 0; ib--) {
        trsm('L', 'L', 'T', 'N', ns, nrhs, 1.0, mf + snode(&L, ib), lda_of(&L, ib), rhs + ib * ns, n);
        gemm('T', 'N', ns, nrhs, ns, -1.0, mf + snode(&L, ib - 1) + ns, lda_of(&L, ib - 1),
             rhs + ib * ns, n, 1.0, rhs + (ib - 1) * ns, n);
    }
    trsm('L', 'L', 'T', 'N', ns, nrhs, 1.0, mf + snode(&L, 0), lda_of(&L, 0), rhs, n);
    layout_free(&L);
    return 0;
}

/* ---- ThirdStageRGF, RGF.H:422-506, and the exact recursion --------------
 * mf holds the factor (unchanged).  diag receives diag(A^-1), length n.
 * exact = 0: the reference's formula S_ii = X^T [I + B^T S_{i+1} B + D^T S_dd D] X.
 * exact = 1: S_ii = X^T X + W^T Sb W with W = [B; D] X and the full
 *            Sb = [[S_{i+1}, S_{d,i+1}^T], [S_{d,i+1}, S_dd]]; S_{d,i} = -(Sb W)[ns:, :].
 * Returns 0, or -1 on allocation failure. */
int rgf_oracle_selinv(const double *mf, long ns, long nt, long nd, double *diag, int exact)
{
    layout_t L;
    if (layout_init(&L, ns, nt, nd)) return -1;
    long m = ns + nd;
    double *X = (double *)calloc((size_t)(ns * ns), sizeof(double));
    double *T1 = (double *)calloc((size_t)(m * m), sizeof(double));
    double *T2 = (double *)calloc((size_t)(m * m), sizeof(double));
    double *Sb = (double *)calloc((size_t)(m * m), sizeof(double));   /* S_below, ld = m */
    double *Sdd = (double *)calloc((size_t)(nd > 0 ? nd * nd : 1), sizeof(double));
    double *W = (double *)calloc((size_t)(m * ns), sizeof(double));
    double *Z = (double *)calloc((size_t)(m * ns), sizeof(double));
    if (!X || !T1 || !T2 || !Sb || !Sdd || !W || !Z) {
        free(X); free(T1); free(T2); free(Sb); free(Sdd); free(W); free(Z); layout_free(&L);
        return -1;
    }

    if (nd > 0) {   /* S_dd = L_dd^-T L_dd^-1 (RGF.H:439-442) */
        const double *C = mf + L.diag_pos[nt * ns];
        for (long j = 0; j < nd; j++)
            for (long i = 0; i < nd; i++) T1[i + j * nd] = (i >= j) ? C[i + j * nd] : 0.0;
        trtri(nd, T1, nd);
        gemm('T', 'N', nd, nd, nd, 1.0, T1, nd, T1, nd, 0.0, Sdd, nd);
        for (long i = 0; i < nd; i++) diag[nt * ns + i] = Sdd[i + i * nd];
    }

    for (long ib = nt - 1; ib >= 0; ib--) {
        const double *P = mf + snode(&L, ib);
        long ldp = lda_of(&L, ib);
        int last = (ib == nt - 1);
        /* X = L_ii^-1 (tril + trtri, RGF.H:486-487) */
        for (long j = 0; j < ns; j++)
            for (long i = 0; i < ns; i++) X[i + j * ns] = (i >= j) ? P[i + j * ldp] : 0.0;
        trtri(ns, X, ns);

        if (!exact) {
            /* T1 = I + B^T S_{i+1} B + D^T S_dd D ; S_ii = X^T T1 X (RGF.H:489-502) */
            memset(T1, 0, sizeof(double) * (size_t)(ns * ns));
            for (long i = 0; i < ns; i++) T1[i + i * ns] = 1.0;
            if (!last) {   /* S_{i+1} sits in Sb[0:ns, 0:ns] */
                gemm('T', 'N', ns, ns, ns, 1.0, P + ns, ldp, Sb, m, 0.0, T2, ns);
                gemm('N', 'N', ns, ns, ns, 1.0, T2, ns, P + ns, ldp, 1.0, T1, ns);
            }
            if (nd > 0) {
                gemm('T', 'N', ns, nd, nd, 1.0, P + arrow_off(&L, ib), ldp, Sdd, nd, 0.0, T2, ns);
                gemm('N', 'N', ns, ns, nd, 1.0, T2, ns, P + arrow_off(&L, ib), ldp, 1.0, T1, ns);
            }
            gemm('T', 'N', ns, ns, ns, 1.0, X, ns, T1, ns, 0.0, T2, ns);
            gemm('N', 'N', ns, ns, ns, 1.0, T2, ns, X, ns, 0.0, T1, ns);
            /* next S_below: [[S_ii, 0], [0, S_dd]] */
            for (long j = 0; j < ns; j++)
                for (long i = 0; i < ns; i++) Sb[i + j * m] = T1[i + j * ns];
        } else {
            /* F = [B_i; D_i] (or D_i alone for the last block); W = F X */
            long fr = last ? nd : m;                /* rows of F */
            long r0 = last ? ns : 0;                /* first row/col of Sb used */
            const double *F = last ? P + arrow_off(&L, ib) : P + ns;
            if (!last) {
                /* Sb = [[S_{i+1}, S_{d,i+1}^T], [S_{d,i+1}, S_dd]] already holds S_{i+1}, S_{d,i+1} */
            }
            for (long j = 0; j < nd; j++)
                for (long i = 0; i < nd; i++) Sb[(ns + i) + (ns + j) * m] = Sdd[i + j * nd];
            if (fr > 0) {
                gemm('N', 'N', fr, ns, ns, 1.0, F, ldp, X, ns, 0.0, W, m);
                gemm('N', 'N', fr, ns, fr, 1.0, Sb + r0 + r0 * m, m, W, m, 0.0, Z, m);
            }
            /* S_ii = X^T X + W^T Z */
            gemm('T', 'N', ns, ns, ns, 1.0, X, ns, X, ns, 0.0, T1, ns);
            if (fr > 0) gemm('T', 'N', ns, ns, fr, 1.0, W, m, Z, m, 1.0, T1, ns);
            /* next S_below: S_ii top-left, S_{d,i} = -Z[rows of D] */
            for (long j = 0; j < ns; j++)
                for (long i = 0; i < ns; i++) Sb[i + j * m] = T1[i + j * ns];
            for (long j = 0; j < ns; j++)
                for (long d = 0; d < nd; d++) {
                    double v = -Z[(fr - nd + d) + j * m];
                    Sb[(ns + d) + j * m] = v;
                    Sb[j + (ns + d) * m] = v;
                }
        }
        for (long i = 0; i < ns; i++) diag[ib * ns + i] = T1[i + i * ns];
    }
    free(X); free(T1); free(T2); free(Sb); free(Sdd); free(W); free(Z);
    layout_free(&L);
    return 0;
}

/* ---- residualNorm, RGF.H:631-655; symmetric lower-CSC matvec ----------- */
double rgf_oracle_residual_norm(const uint64_t *ia, const uint64_t *ja, const double *a, long n,
                                const double *x, const double *b)
{
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    if (!r) return NAN;
    memcpy(r, b, sizeof(double) * (size_t)n);
    for (long ic = 0; ic < n; ic++)
        for (uint64_t i = ia[ic]; i < ia[ic + 1]; i++) {
            long ir = (long)ja[i];
            r[ir] -= a[i] * x[ic];
            if (ir != ic) r[ic] -= a[i] * x[ir];
        }
    int in = (int)n, one = 1;
    double res = dnrm2_(&in, r, &one);
    free(r);
    return res;
}

/* residualNormNormalized, RGF.H:659-663 */
double rgf_oracle_residual_norm_normalized(const uint64_t *ia, const uint64_t *ja, const double *a,
                                           long n, const double *x, const double *b)
{
    int in = (int)n, one = 1;
    return rgf_oracle_residual_norm(ia, ja, a, n, x, b) / dnrm2_(&in, b, &one);
}

/* ---- synthetic generator (host restatement) ------------------------------
 * Philox4x32-10 counter RNG (Salmon, Moraes, Dror, Shaw, SC'11).  One call
 * gives four 32-bit words -> two 53-bit uniforms -> two normals by the
 * Box-Muller transform.  Stream of element e of a (kind, block) array:
 *   key = (seed lo, seed hi), counter = (e/2 lo, e/2 hi, kind<<24 | block, matrix id),
 * normal index e%2.  Kinds: 1 = G_t (ns x ns, diag block t), 2 = C_t (ns x ns,
 * coupling block (t+1,t)), 3 = arrow (nd x ns*nt, column-major over the
 * global column), 4 = H (nd x nd), 5 = per-instance scale factors.
 * Matrix (configs of the benchmark): diag block B_t = theta_d*(G_t G_t^T/ns + I),
 * coupling sigma_c*theta_c*N(0,1), arrow sigma_a*N(0,1), corner H H^T + nd I;
 * then every diagonal entry gets + (sum_{j!=k} |a_kj| over the full symmetric
 * row) + 1.  Output: A in MF layout (lower parts of diagonal blocks and corner). */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void rgf_oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double gen_normal(uint64_t seed, uint32_t matrix_id, uint32_t kind, uint32_t block, uint64_t e)
{
    uint64_t p = e >> 1;
    uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), (kind << 24) | block, matrix_id};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    rgf_oracle_philox4x32_10(ctr, key, w);
    /* 53-bit uniforms in [0,1) */
    double u1 = (double)(((uint64_t)(w[0] >> 5) << 26) | (w[1] >> 6)) * (1.0 / 9007199254740992.0);
    double u2 = (double)(((uint64_t)(w[2] >> 5) << 26) | (w[3] >> 6)) * (1.0 / 9007199254740992.0);
    double r = sqrt(-2.0 * log(1.0 - u1));   /* 1-u1 in (0,1] */
    double t = RGF_TWO_PI * u2;
    return (e & 1) ? r * sin(t) : r * cos(t);
}

/* normals of elements e0 .. e0+count-1 of one (kind, block) stream, one Philox call per pair */
static void gen_normals(uint64_t seed, uint32_t matrix_id, uint32_t kind, uint32_t block, uint64_t e0,
                        long count, double scale, double *out)
{
    long k = 0;
    while (k < count) {
        uint64_t e = e0 + (uint64_t)k;
        uint64_t p = e >> 1;
        uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), (kind << 24) | block, matrix_id};
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        uint32_t w[4];
        rgf_oracle_philox4x32_10(ctr, key, w);
        double u1 = (double)(((uint64_t)(w[0] >> 5) << 26) | (w[1] >> 6)) * (1.0 / 9007199254740992.0);
        double u2 = (double)(((uint64_t)(w[2] >> 5) << 26) | (w[3] >> 6)) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(1.0 - u1));
        double t = RGF_TWO_PI * u2;
        if ((e & 1) == 0) {
            out[k++] = scale * (r * cos(t));
            if (k < count) out[k++] = scale * (r * sin(t));
        } else {
            out[k++] = scale * (r * sin(t));
        }
    }
}

double rgf_oracle_gen_uniform(uint64_t seed, uint32_t matrix_id, uint32_t idx)
{
    uint32_t ctr[4] = {idx, 0u, (5u << 24), matrix_id};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    rgf_oracle_philox4x32_10(ctr, key, w);
    return (double)(((uint64_t)(w[0] >> 5) << 26) | (w[1] >> 6)) * (1.0 / 9007199254740992.0);
}

double rgf_oracle_gen_normal(uint64_t seed, uint32_t matrix_id, uint32_t kind, uint32_t block, uint64_t e)
{
    return gen_normal(seed, matrix_id, kind, block, e);
}

int rgf_oracle_generate(uint64_t seed, uint32_t matrix_id, long ns, long nt, long nd,
                        double sigma_c, double sigma_a, double theta_d, double theta_c, double *mf)
{
    layout_t L;
    if (layout_init(&L, ns, nt, nd)) return -1;
    long n = L.n;
    memset(mf, 0, sizeof(double) * (size_t)rgf_oracle_mf_size(ns, nt, nd));
    double *G = (double *)malloc(sizeof(double) * (size_t)(ns * ns));
    double *rs = (double *)calloc((size_t)n, sizeof(double));
    if (!G || !rs) { free(G); free(rs); layout_free(&L); return -1; }
    for (long t = 0; t < nt; t++) {
        double *P = mf + snode(&L, t);
        long ldp = lda_of(&L, t);
        gen_normals(seed, matrix_id, 1, (uint32_t)t, 0, ns * ns, 1.0, G);
        /* B_t = G G^T / ns + I, lower triangle only */
        gemm('N', 'T', ns, ns, ns, 1.0 / (double)ns, G, ns, G, ns, 0.0, P, ldp);
        for (long j = 0; j < ns; j++) {
            for (long i = 0; i < j; i++) P[i + j * ldp] = 0.0;
            P[j + j * ldp] += 1.0;
            for (long i = j; i < ns; i++) P[i + j * ldp] *= theta_d;
        }
        if (t < nt - 1) {
            gen_normals(seed, matrix_id, 2, (uint32_t)t, 0, ns * ns, theta_c * sigma_c, G);
            for (long j = 0; j < ns; j++)
                for (long i = 0; i < ns; i++) P[ns + i + j * ldp] = G[i + j * ns];
        }
        if (nd > 0) {
            gen_normals(seed, matrix_id, 3, 0, (uint64_t)(t * ns * nd), ns * nd, sigma_a, G);
            for (long j = 0; j < ns; j++)
                for (long d = 0; d < nd; d++) P[arrow_off(&L, t) + d + j * ldp] = G[d + j * nd];
        }
    }
    if (nd > 0) {
        double *C = mf + L.diag_pos[nt * ns];
        double *H = G;   /* ns >= nd */
        gen_normals(seed, matrix_id, 4, 0, 0, nd * nd, 1.0, H);
        gemm('N', 'T', nd, nd, nd, 1.0, H, nd, H, nd, 0.0, C, nd);
        for (long j = 0; j < nd; j++) {
            for (long i = 0; i < j; i++) C[i + j * nd] = 0.0;
            C[j + j * nd] += (double)nd;
        }
    }
    /* row abs-sums over the full symmetric matrix, from the stored lower part */
    for (long t = 0; t < nt; t++) {
        const double *P = mf + snode(&L, t);
        long ldp = lda_of(&L, t);
        long rows_below = (t < nt - 1) ? ns : 0;
        for (long j = 0; j < ns; j++) {
            long cg = t * ns + j;
            for (long i = j + 1; i < ns; i++) {
                double v = fabs(P[i + j * ldp]);
                rs[t * ns + i] += v; rs[cg] += v;
            }
            for (long i = 0; i < rows_below; i++) {
                double v = fabs(P[ns + i + j * ldp]);
                rs[(t + 1) * ns + i] += v; rs[cg] += v;
            }
            for (long d = 0; d < nd; d++) {
                double v = fabs(P[arrow_off(&L, t) + d + j * ldp]);
                rs[nt * ns + d] += v; rs[cg] += v;
            }
        }
    }
    if (nd > 0) {
        const double *C = mf + L.diag_pos[nt * ns];
        for (long j = 0; j < nd; j++)
            for (long i = j + 1; i < nd; i++) {
                double v = fabs(C[i + j * nd]);
                rs[nt * ns + i] += v; rs[nt * ns + j] += v;
            }
    }
    for (long k = 0; k < n; k++) mf[L.diag_pos[k]] += rs[k] + 1.0;
    free(G); free(rs);
    layout_free(&L);
    return 0;
}
