// Dense FP64 block kernels of the factor / selected-inversion sweep:
//   potrf   blocked right-looking Cholesky of one ns x ns diagonal block
//           (replaces magma_dpotrf_gpu at RGF.H:254,284,312,335); the 64-wide
//           panel is factored in LDS and its log-diagonal summed in the same
//           kernel (logDet, RGF.H:616-627), the trailing update is an MFMA GEMM.
//   trtri   in-place L^-1 by diagonal-tile inversion in LDS + recursive
//           GEMM merges X21 = -X22 L21 X11 (replaces magma_dtrtri_gpu,
//           RGF.H:441,450,471,487, and the tril/identity kernels d_tril,
//           d_init_eye at CWC_utility.cu:231-251, 792-815).
//   corner  the nd x nd arrowhead corner (RGF.H:280,308,332,335,439-442).
//   small   copy / fill / mirror / diagonal gather (d_fill, d_indexed_copy_offset,
//           CWC_utility.cu:883-901, 964-979).
#include "rgf_internal.hpp"
#include <climits>

namespace rgf {

constexpr int NB = 64;

// ------------------------------------------------------------------ potrf
// One workgroup factors the kb x kb (kb <= 64) diagonal tile at A[k0, k0] and
// inverts it.  Each of the 256 lanes owns a 4x4 register block (bi, bj) of the
// 16x16 block grid (lower blocks only do work), so both sweeps are 16 blocked
// steps of register math with one or two barriers each:
//   factor   step K: lane (K,K) factors its 4x4 block; lanes (i,K) solve
//            L_iK = A_iK L_KK^-T; lanes (i,j), j > K, update A_ij -= L_iK L_jK^T
//   inverse  step K: lanes (K,J) finish X_KJ = -X_KK T_KJ (X_KK = L_KK^-1);
//            lanes (I,J), I > K >= J, accumulate T_IJ += L_IK X_KJ
// The tile is identity-padded beyond kb.  Writes L_kk back (zeros above the
// diagonal), X_kk = L_kk^-1 to xout (64 x 64, ld 64), the tile's sum of log L_jj
// (fixed-order tree) and the first non-positive pivot.
__global__ __launch_bounds__(256) void potrf_tile_kernel(double* A, int lda, long long sA, int k0, int kb,
                                                         double* xout, long long sX, double* logdet_slot,
                                                         long long sP, int* info, int info_base)
{
    {
        const long long b = blockIdx.x;   // batch entry
        A += b * sA;
        xout += b * sX;
        logdet_slot += b * sP;
        info += b;
    }
    __shared__ double L[NB][NB + 1];
    __shared__ double X[NB][NB + 1];
    __shared__ double red[NB];
    __shared__ int bad_col;
    const int tid = threadIdx.x;
    const int bi = tid >> 4, bj = tid & 15;
    const bool lower = bi >= bj;
    if (tid == 0) bad_col = INT_MAX;
    double a[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = 4 * bi + u, c = 4 * bj + v;
            double x = 0.0;
            if (r < kb && c < kb) { if (r >= c) x = A[(long long)(k0 + c) * lda + (k0 + r)]; }
            else if (r == c) x = 1.0;
            a[u][v] = x;
        }
    __syncthreads();
    // ---- factor
    for (int K = 0; K < 16; K++) {
        if (bi == K && bj == K) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double d = a[v][v];
                const bool bad = !(d > 0.0);
                const double sd = bad ? 1.0 : sqrt(d);
                if (bad && 4 * K + v < kb && bad_col == INT_MAX) bad_col = 4 * K + v;
                a[v][v] = sd;
                const double ri = 1.0 / sd;
#pragma unroll
                for (int u = v + 1; u < 4; u++) a[u][v] *= ri;
#pragma unroll
                for (int u = v + 1; u < 4; u++)
#pragma unroll
                    for (int w = v + 1; w <= u; w++) a[u][w] -= a[u][v] * a[w][v];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) L[4 * K + u][4 * K + v] = (u >= v) ? a[u][v] : 0.0;
        }
        __syncthreads();
        if (bj == K && bi > K) {
            double lk[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v <= u; v++) lk[u][v] = L[4 * K + u][4 * K + v];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double ri = 1.0 / lk[v][v];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    double x = a[u][v];
#pragma unroll
                    for (int w = 0; w < v; w++) x -= a[u][w] * lk[v][w];
                    a[u][v] = x * ri;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) L[4 * bi + u][4 * K + v] = a[u][v];
        }
        __syncthreads();
        if (bj > K && lower) {
            double li[4][4], lj[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    li[u][w] = L[4 * bi + u][4 * K + w];
                    lj[u][w] = L[4 * bj + u][4 * K + w];
                }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    double x = a[u][v];
#pragma unroll
                    for (int w = 0; w < 4; w++) x -= li[u][w] * lj[v][w];
                    a[u][v] = x;
                }
        }
    }
    __syncthreads();
    // ---- inverse: diagonal blocks first (lanes (K,K)), T accumulators in registers
    double t[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) t[u][v] = 0.0;
    double xd[4][4];   // X_ii of this lane's block row (used by lanes with bi == K at step K)
    {
        double l[4][4];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) l[u][v] = (u >= v) ? L[4 * bi + u][4 * bi + v] : 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double x = (r == c) ? 1.0 : 0.0;
#pragma unroll
                for (int q = 0; q < r; q++) x -= l[r][q] * xd[q][c];
                xd[r][c] = (r >= c) ? x / l[r][r] : 0.0;
            }
        }
    }
    for (int K = 0; K < 16; K++) {
        if (bi == K && bj <= K) {   // finish X_KJ
            double xk[4][4];
            if (bj == K) {
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int v = 0; v < 4; v++) xk[u][v] = xd[u][v];
            } else {
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        double x = 0.0;
#pragma unroll
                        for (int w = 0; w <= u; w++) x -= xd[u][w] * t[w][v];
                        xk[u][v] = x;
                    }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) X[4 * K + u][4 * bj + v] = xk[u][v];
        }
        __syncthreads();
        if (bi > K && bj <= K) {    // T_IJ += L_IK X_KJ
            double li[4][4], xk[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    li[u][w] = L[4 * bi + u][4 * K + w];
                    xk[u][w] = X[4 * K + u][4 * bj + w];
                }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    double x = t[u][v];
#pragma unroll
                    for (int w = 0; w < 4; w++) x += li[u][w] * xk[w][v];
                    t[u][v] = x;
                }
        }
    }
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) {
        const int r = e & (NB - 1), c = e >> 6;
        xout[c * NB + r] = (r >= c) ? X[r][c] : 0.0;
        if (r < kb && c < kb) A[(long long)(k0 + c) * lda + (k0 + r)] = (r >= c) ? L[r][c] : 0.0;
    }
    if (tid < NB) red[tid] = tid < kb ? log(L[tid][tid]) : 0.0;
    __syncthreads();
    for (int off = NB / 2; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
    }
    if (tid == 0) {
        *logdet_slot = red[0];
        if (bad_col != INT_MAX) atomicMin(info, info_base + k0 + bad_col + 1);
    }
}

// Blocked right-looking Cholesky of the n x n block at A (lower): per 64-wide panel
// the tile kernel, the rows below as an in-place product with X_kk^T (the output has
// one column tile, so every workgroup reads only the rows it writes), and the
// trailing lower update on MFMA.  xdiag receives the T inverted diagonal tiles.
int potrf_f64(double* A, int lda, long long sA, int n, int count, double* logdet_parts, long long sP, int* info,
              int info_base, double* xdiag, long long sX, hipStream_t s)
{
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        const int below = n - k0 - kb;
        double* xt = xdiag + (long long)(k0 / NB) * NB * NB;
        hipLaunchKernelGGL(potrf_tile_kernel, dim3(count), dim3(256), 0, s, A, lda, sA, k0, kb, xt, sX,
                           logdet_parts + k0 / NB, sP, info, info_base);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
        if (below > 0) {
            double* Pn = A + (long long)k0 * lda + (k0 + kb);
            GemmDesc p = gdesc(Pn, lda, xt, NB, Pn, lda, below, kb, kb, 1.0, 0.0, G_BUPPER);
            batch2(p, count, sA, sX, sA);
            int rc = gemm_f64(Op::N, Op::T, &p, 1, s);
            if (rc) return rc;
            double* Cn = A + (long long)(k0 + kb) * lda + (k0 + kb);
            GemmDesc d = gdesc(Pn, lda, Pn, lda, Cn, lda, below, below, kb, -1.0, 1.0, G_CLOWER);
            batch2(d, count, sA, sA, sA);
            rc = gemm_f64(Op::N, Op::T, &d, 1, s);
            if (rc) return rc;
        }
    }
    return 0;
}

// one launch of the tile kernel on the leading tile (tuning only)
int potrf_tile_only(double* A, int lda, double* xdiag, double* logdet_slot, int* info, hipStream_t s)
{
    hipLaunchKernelGGL(potrf_tile_kernel, dim3(1), dim3(256), 0, s, A, lda, 0LL, 0, NB, xdiag, 0LL, logdet_slot, 0LL,
                       info, 0);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------------ trtri
// grid.x = T*T tiles.  (ti < tj): zero the tile above the diagonal.  (ti == tj):
// invert the diagonal tile in LDS.  (ti > tj): nothing (written by the merges).
__global__ __launch_bounds__(256) void trtri_diag_kernel(double* A, int lda, long long sA, int n, int T,
                                                         const double* xdiag, long long sX)
{
    A += (long long)blockIdx.y * sA;
    if (xdiag) xdiag += (long long)blockIdx.y * sX;
    __shared__ double Ls[NB][NB + 1];
    __shared__ double Xs[NB][NB + 1];
    const int ti = blockIdx.x % T, tj = blockIdx.x / T;
    const int tid = threadIdx.x;
    const int r0 = ti * NB, c0 = tj * NB;
    const int rb = min(NB, n - r0), cb = min(NB, n - c0);
    if (ti > tj) return;
    if (ti < tj) {
        for (int e = tid; e < NB * NB; e += 256) {
            int r = e & (NB - 1), c = e >> 6;
            if (r < rb && c < cb) A[(long long)(c0 + c) * lda + (r0 + r)] = 0.0;
        }
        return;
    }
    const int kb = rb;
    if (xdiag) {   // tiles already inverted by the Cholesky tile kernel
        const double* xt = xdiag + (long long)ti * NB * NB;
        for (int e = tid; e < kb * kb; e += 256) {
            int r = e % kb, c = e / kb;
            A[(long long)(r0 + c) * lda + (r0 + r)] = xt[c * NB + r];
        }
        return;
    }
    for (int e = tid; e < NB * NB; e += 256) {
        int r = e & (NB - 1), c = e >> 6;
        Ls[r][c] = (r < kb && c < kb && r >= c) ? A[(long long)(r0 + c) * lda + (r0 + r)] : 0.0;
        Xs[r][c] = (r == c && r < kb) ? 1.0 : 0.0;
    }
    __syncthreads();
    // forward substitution L X = I, row by row
    for (int r = 0; r < kb; r++) {
        if (tid <= r) Xs[r][tid] /= Ls[r][r];
        __syncthreads();
        const int w = kb - r - 1, cols = r + 1;
        for (int e = tid; e < w * cols; e += 256) {
            int p = r + 1 + e % w, c = e / w;
            Xs[p][c] -= Ls[p][r] * Xs[r][c];
        }
        __syncthreads();
    }
    for (int e = tid; e < kb * kb; e += 256) {
        int r = e % kb, c = e / kb;
        A[(long long)(r0 + c) * lda + (r0 + r)] = (r >= c) ? Xs[r][c] : 0.0;
    }
}

int trtri_f64_inplace(double* A, int lda, long long sA, int n, int count, double* tmp, long long sT,
                      const double* xdiag, long long sX, hipStream_t s)
{
    const int T = (n + NB - 1) / NB;
    hipLaunchKernelGGL(trtri_diag_kernel, dim3(T * T, count), dim3(256), 0, s, A, lda, sA, n, T, xdiag, sX);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    for (int sz = 1; sz < T; sz *= 2) {
        const int G = (T - sz + 2 * sz - 1) / (2 * sz);   // groups with a non-empty right half
        const int n1 = sz * NB;
        const long long gstride = (long long)(2 * sz * NB) * (1 + (long long)lda);
        const long long tstride = (long long)n1 * n1;
        // groups whose right half is complete share one descriptor
        int gfull = 0;
        for (int g = 0; g < G; g++)
            if ((g * 2 * sz + 2 * sz) * NB <= n) gfull = g + 1;
        GemmDesc d1[2], d2[2];
        int nd1 = 0;
        for (int part = 0; part < 2; part++) {
            int g0 = part == 0 ? 0 : gfull;
            int cnt = part == 0 ? gfull : G - gfull;
            if (cnt <= 0) continue;
            const int c0 = g0 * 2 * sz * NB, r0 = c0 + n1;
            const int n2 = min(n1, n - r0);
            double* L21 = A + (long long)c0 * lda + r0;
            double* X11 = A + (long long)c0 * lda + c0;
            double* X22 = A + (long long)r0 * lda + r0;
            double* Tg = tmp + (long long)g0 * tstride;
            // T = L21 X11   (X11 lower)
            GemmDesc a = gdesc(L21, lda, X11, lda, Tg, n1, n2, n1, n1, 1.0, 0.0, G_BLOWER);
            a.count = cnt; a.sA = gstride; a.sB = gstride; a.sC = tstride; a.sCin = tstride;
            batch2(a, count, sA, sA, sT);
            // X21 = -X22 T  (X22 lower)
            GemmDesc b = gdesc(X22, lda, Tg, n1, L21, lda, n2, n1, n2, -1.0, 0.0, G_ALOWER);
            b.count = cnt; b.sA = gstride; b.sB = tstride; b.sC = gstride; b.sCin = gstride;
            batch2(b, count, sA, sT, sA);
            d1[nd1] = a; d2[nd1] = b; nd1++;
        }
        int rc = gemm_f64(Op::N, Op::N, d1, nd1, s);
        if (rc) return rc;
        rc = gemm_f64(Op::N, Op::N, d2, nd1, s);
        if (rc) return rc;
    }
    return 0;
}

// ------------------------------------------------------------------ small kernels
__global__ void copy2d_kernel(const double* src, int lds, long long sS, double* dst, int ldd, long long sD,
                              int rows, int cols, double alpha)
{
    src += (long long)blockIdx.y * sS;
    dst += (long long)blockIdx.y * sD;
    const long long total = (long long)rows * cols;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        int r = (int)(e % rows), c = (int)(e / rows);
        dst[(long long)c * ldd + r] = alpha * src[(long long)c * lds + r];
    }
}

int copy2d(const double* src, int lds, double* dst, int ldd, int rows, int cols, double alpha,
           hipStream_t s, int count, long long sS, long long sD)
{
    if (rows <= 0 || cols <= 0 || count <= 0) return 0;
    long long total = (long long)rows * cols;
    int grid = (int)min((total + 255) / 256, (long long)8192);
    hipLaunchKernelGGL(copy2d_kernel, dim3(grid, count), dim3(256), 0, s, src, lds, sS, dst, ldd, sD, rows, cols,
                       alpha);
    return (int)hipGetLastError();
}

__global__ void fill2d_kernel(double* dst, int ldd, long long sD, int rows, int cols, double v)
{
    dst += (long long)blockIdx.y * sD;
    const long long total = (long long)rows * cols;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        int r = (int)(e % rows), c = (int)(e / rows);
        dst[(long long)c * ldd + r] = v;
    }
}

int fill2d(double* dst, int ldd, int rows, int cols, double v, hipStream_t s, int count, long long sD)
{
    if (rows <= 0 || cols <= 0 || count <= 0) return 0;
    long long total = (long long)rows * cols;
    int grid = (int)min((total + 255) / 256, (long long)8192);
    hipLaunchKernelGGL(fill2d_kernel, dim3(grid, count), dim3(256), 0, s, dst, ldd, sD, rows, cols, v);
    return (int)hipGetLastError();
}

// mirror the lower triangle into the upper one through 32x32 LDS tiles; optional
// diagonal gather.
__global__ __launch_bounds__(256) void mirror_kernel(double* A, int lda, long long sA, int n, int T, double* diag,
                                                     long long sDg)
{
    A += (long long)blockIdx.y * sA;
    if (diag) diag += (long long)blockIdx.y * sDg;
    __shared__ double t[32][33];
    const int bi = blockIdx.x % T, bj = blockIdx.x / T;   // destination tile (bi row, bj col), bi <= bj
    if (bi > bj) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    // source: tile (bj row-block, bi col-block) of the lower triangle
    for (int q = ty; q < 32; q += 8) {
        int r = bj * 32 + tx, c = bi * 32 + q;
        t[q][tx] = (r < n && c < n) ? A[(long long)c * lda + r] : 0.0;   // t[c_local][r_local]
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        int r = bi * 32 + tx, c = bj * 32 + q;   // destination (r, c) = source (c, r)
        if (r < n && c < n && r < c) A[(long long)c * lda + r] = t[tx][q];
        if (diag && r < n && r == c) diag[r] = A[(long long)c * lda + r];
    }
}

int mirror_lower_diag(double* A, int lda, long long sA, int n, double* diag, long long sDg, int count, hipStream_t s)
{
    if (n <= 0 || count <= 0) return 0;
    const int T = (n + 31) / 32;
    hipLaunchKernelGGL(mirror_kernel, dim3(T * T, count), dim3(256), 0, s, A, lda, sA, n, T, diag, sDg);
    return (int)hipGetLastError();
}

int mirror_lower(double* A, int lda, int n, hipStream_t s) { return mirror_lower_diag(A, lda, 0, n, nullptr, 0, 1, s); }

__global__ void extract_diag_kernel(const double* A, int lda, long long sA, int n, double* out, long long sO)
{
    A += (long long)blockIdx.y * sA;
    out += (long long)blockIdx.y * sO;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = A[(long long)i * lda + i];
}

int extract_diag(const double* A, int lda, int n, double* out, hipStream_t s, int count, long long sA, long long sO)
{
    if (n <= 0 || count <= 0) return 0;
    hipLaunchKernelGGL(extract_diag_kernel, dim3((n + 255) / 256, count), dim3(256), 0, s, A, lda, sA, n, out, sO);
    return (int)hipGetLastError();
}

// S_{d,i} = -Z[zrow0 + d, c] into S[ns + d, c] and S[c, ns + d]
__global__ void selinv_offdiag_kernel(double* S, int lds, long long sS, int ns, int nd, const double* Z, int ldz,
                                      long long sZ, int zrow0)
{
    S += (long long)blockIdx.y * sS;
    Z += (long long)blockIdx.y * sZ;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ns) return;
    for (int d = 0; d < nd; d++) {
        double v = -Z[(long long)c * ldz + zrow0 + d];
        S[(long long)c * lds + ns + d] = v;
        S[(long long)(ns + d) * lds + c] = v;
    }
}

int selinv_finish(double* S, int lds, long long sS, int ns, int nd, const double* Z, int ldz, long long sZ,
                  int zrow0, int exact, double* diag_out, long long sDg, int count, hipStream_t s)
{
    int rc = mirror_lower_diag(S, lds, sS, ns, diag_out, sDg, count, s);
    if (rc) return rc;
    if (exact && nd > 0) {
        hipLaunchKernelGGL(selinv_offdiag_kernel, dim3((ns + 255) / 256, count), dim3(256), 0, s, S, lds, sS, ns,
                           nd, Z, ldz, sZ, zrow0);
        return (int)hipGetLastError();
    }
    return 0;
}

// ------------------------------------------------------------------ corner
// partial[i] = D_i D_i^T for panel i (nd x nd, full), one workgroup per panel.
__global__ __launch_bounds__(256) void corner_partial_kernel(const double* Dbase, long long dstride,
                                                             long long sM, int ldd, int ns, int nd, double* partial,
                                                             long long sPart)
{
    const int i = blockIdx.x;
    const double* D = Dbase + (long long)blockIdx.y * sM + (long long)i * dstride;
    partial += (long long)blockIdx.y * sPart;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int npairs = nd * nd;
    for (int pq = wave; pq < npairs; pq += 4) {
        const int p = pq % nd, q = pq / nd;
        if (p < q) continue;
        double acc = 0.0;
        for (int c = lane; c < ns; c += 64) acc += D[(long long)c * ldd + p] * D[(long long)c * ldd + q];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) {
            partial[(long long)i * npairs + p + q * nd] = acc;
            partial[(long long)i * npairs + q + p * nd] = acc;
        }
    }
}

// one workgroup: S = Cin - sum_i partial[i]; Cholesky; log-det; L_dd and X_dd = L_dd^-1.
__global__ __launch_bounds__(256) void corner_finish_kernel(const double* partial, long long sPart, int nparts,
                                                            int nd, const double* Cin, int ldcin, long long sC,
                                                            double* Ldd, double* Xdd, double* logdet_slot,
                                                            long long sP, int* info, int info_base)
{
    {
        const long long b = blockIdx.x;
        partial += b * sPart;
        Cin += b * sC;
        Ldd += b * sC;
        Xdd += b * sC;
        logdet_slot += b * sP;
        info += b;
    }
    __shared__ double Ls[NB][NB + 1];
    __shared__ double Xs[NB][NB + 1];
    __shared__ int bad_col;
    const int tid = threadIdx.x;
    if (tid == 0) bad_col = INT_MAX;
    for (int e = tid; e < nd * nd; e += 256) {
        int r = e % nd, c = e / nd;
        double v = Cin[(long long)c * ldcin + r];
        for (int i = 0; i < nparts; i++) v -= partial[(long long)i * nd * nd + e];
        Ls[r][c] = (r >= c) ? v : 0.0;
        Xs[r][c] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nd; j++) {
        double d = Ls[j][j];
        __syncthreads();
        bool bad = !(d > 0.0);
        double sd = bad ? 1.0 : sqrt(d);
        if (tid == j) { Ls[j][j] = sd; if (bad) bad_col = min(bad_col, j); }
        else if (tid > j && tid < nd) Ls[tid][j] /= sd;
        __syncthreads();
        const int w = nd - j - 1;
        for (int e = tid; e < w * w; e += 256) {
            int r = j + 1 + e % w, c = j + 1 + e / w;
            if (r >= c) Ls[r][c] -= Ls[r][j] * Ls[c][j];
        }
        __syncthreads();
    }
    for (int r = 0; r < nd; r++) {
        if (tid <= r) Xs[r][tid] /= Ls[r][r];
        __syncthreads();
        const int w = nd - r - 1, cols = r + 1;
        for (int e = tid; e < w * cols; e += 256) {
            int p = r + 1 + e % w, c = e / w;
            Xs[p][c] -= Ls[p][r] * Xs[r][c];
        }
        __syncthreads();
    }
    for (int e = tid; e < nd * nd; e += 256) {
        int r = e % nd, c = e / nd;
        Ldd[(long long)c * nd + r] = (r >= c) ? Ls[r][c] : 0.0;
        Xdd[(long long)c * nd + r] = (r >= c) ? Xs[r][c] : 0.0;
    }
    if (tid == 0) {
        double s = 0.0;
        for (int j = 0; j < nd; j++) s += log(Ls[j][j]);
        *logdet_slot = s;
        if (bad_col != INT_MAX) atomicMin(info, info_base + bad_col + 1);
    }
}

int corner_factor(const double* Dbase, long long dstride, long long sM, int ldd, int ns, int nt, int nd,
                  const double* Cin, int ldcin, double* Ldd, double* Xdd, double* partial, double* logdet_slot,
                  long long sP, int* info, int info_base, int count, hipStream_t s)
{
    if (nd <= 0) return 0;
    if (nd > NB || ldcin != nd) return (int)hipErrorInvalidValue;
    const long long sPart = (long long)nt * nd * nd, sC = (long long)nd * nd;
    hipLaunchKernelGGL(corner_partial_kernel, dim3(nt, count), dim3(256), 0, s, Dbase, dstride, sM, ldd, ns, nd,
                       partial, sPart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(corner_finish_kernel, dim3(count), dim3(256), 0, s, partial, sPart, nt, nd, Cin, ldcin, sC,
                       Ldd, Xdd, logdet_slot, sP, info, info_base);
    return (int)hipGetLastError();
}

}  // namespace rgf
