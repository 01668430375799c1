// RGF<std::complex<double>> (the reference instantiates its class for CPX = std::complex<double>,
// RGF.H:756-780, Types.H:41-45): the same block sweep for complex matrices.  The reference's
// complex block updates transpose without conjugating (trsm 'R','L','T','N' at RGF.H:260, 287,
// 317; gemm 'N','T' at RGF.H:276-308), i.e. complex-symmetric algebra: A = A^T (not A^H),
// A = L L^T with L_jj the principal square root of the pivot, log det = 2 sum log L_jj.  This
// file implements that algebra throughout (MAGMA's zpotrf, which the reference calls for the
// diagonal blocks, is the Hermitian factorization, so the reference is consistent only for
// real-valued input, where both readings agree).
//
// Layout: every complex array is two FP64 planes, the imaginary part `P` doubles after the real
// part, all buffers of a handle carved from one allocation so that one offset serves every
// pointer.  A complex product is two launches of the FP64 MFMA GEMM of rgf_gemm.hip with two
// descriptors each (C_r = a A_r B_r + b C_r, C_i = a A_r B_i + b C_i; then C_r -= a A_i B_i,
// C_i += a A_i B_r); copies, the mirror of a symmetric block, the diagonal gather and the
// CSC scatter are linear and run once per plane.  The 64 x 64 diagonal tiles are factored and
// inverted in LDS by ztile_kernel.
#include "rgf_internal.hpp"
#include "../../include/rgf.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

namespace rgf {
namespace {

constexpr int ZNB = 64;

// ------------------------------------------------------------------ complex scalars
struct zd {
    double r, i;
};
__device__ __forceinline__ zd zmul(zd a, zd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ zd zsub(zd a, zd b) { return {a.r - b.r, a.i - b.i}; }
// 1 / z by Smith's scaling (no overflow of |z|^2)
__device__ __forceinline__ zd zinv(zd z)
{
    if (fabs(z.r) >= fabs(z.i)) {
        const double t = z.i / z.r, den = z.r + z.i * t;
        return {1.0 / den, -t / den};
    }
    const double t = z.r / z.i, den = z.r * t + z.i;
    return {t / den, -1.0 / den};
}
// principal square root (real part >= 0), without cancellation in either part
__device__ __forceinline__ zd zsqrt(zd z)
{
    if (z.r == 0.0 && z.i == 0.0) return {0.0, 0.0};
    const double m = hypot(z.r, z.i);
    if (z.r >= 0.0) {
        const double t = sqrt(0.5 * (m + z.r));
        return {t, z.i / (2.0 * t)};
    }
    const double t = sqrt(0.5 * (m - z.r));
    return {fabs(z.i) / (2.0 * t), copysign(t, z.i)};
}

// ------------------------------------------------------------------ diagonal tile
// One workgroup: the kb x kb (kb <= 64) lower tile at A (planes P apart) is factored as
// L L^T (right-looking, in LDS), L^-1 is written to xout (64 x 64, ld 64, zeros above the
// diagonal and beyond kb), sum_j log L_jj (principal logarithms) to part / part + P, and the
// first zero or non-finite pivot, 1-based from info_base, into *info (atomic minimum).
__global__ __launch_bounds__(256) void ztile_kernel(const double* A, int lda, long long P, int kb, double* xout,
                                                    double* part, int* info, int info_base)
{
    __shared__ double Lr[ZNB][ZNB + 1], Li[ZNB][ZNB + 1];
    __shared__ double Xr[ZNB][ZNB + 1], Xi[ZNB][ZNB + 1];
    __shared__ double vr[ZNB], vi[ZNB];   // 1 / L_jj
    __shared__ int bad_s;
    const int tid = threadIdx.x;
    if (tid == 0) bad_s = INT_MAX;
    for (int e = tid; e < ZNB * ZNB; e += 256) {
        const int r = e & (ZNB - 1), c = e >> 6;
        const bool in = r < kb && c < kb && c <= r;
        Lr[r][c] = in ? A[(long long)c * lda + r] : (r == c ? 1.0 : 0.0);
        Li[r][c] = in ? A[P + (long long)c * lda + r] : 0.0;
    }
    for (int j = 0; j < kb; j++) {
        __syncthreads();
        zd d{Lr[j][j], Li[j][j]};
        const bool bad = (d.r == 0.0 && d.i == 0.0) || !isfinite(d.r) || !isfinite(d.i);
        if (bad) {   // recorded; continue with a unit pivot so that nothing turns into NaN
            if (tid == 0 && bad_s == INT_MAX) bad_s = j;
            d = {1.0, 0.0};
        }
        const zd s = zsqrt(d), v = zinv(s);
        __syncthreads();   // every lane has read the pivot
        if (tid == 0) {
            Lr[j][j] = s.r;
            Li[j][j] = s.i;
            vr[j] = v.r;
            vi[j] = v.i;
        }
        if (tid > j && tid < kb) {
            const zd l = zmul({Lr[tid][j], Li[tid][j]}, v);
            Lr[tid][j] = l.r;
            Li[tid][j] = l.i;
        }
        __syncthreads();
        const int w = kb - j - 1;
        for (int e = tid; e < w * w; e += 256) {
            const int r = j + 1 + e % w, c = j + 1 + e / w;
            if (c <= r) {
                const zd u = zsub({Lr[r][c], Li[r][c]}, zmul({Lr[r][j], Li[r][j]}, {Lr[c][j], Li[c][j]}));
                Lr[r][c] = u.r;
                Li[r][c] = u.i;
            }
        }
    }
    __syncthreads();
    // X = L^-1, one column per lane of the first wave: X_cc = 1/L_cc, then row by row
    // X_rc = -(sum_{c <= k < r} L_rk X_kc) / L_rr
    if (tid < ZNB) {
        const int c = tid;
        for (int r = 0; r < ZNB; r++) {
            Xr[r][c] = 0.0;
            Xi[r][c] = 0.0;
        }
        if (c < kb) {
            Xr[c][c] = vr[c];
            Xi[c][c] = vi[c];
            for (int r = c + 1; r < kb; r++) {
                zd acc{0.0, 0.0};
                for (int k = c; k < r; k++) {
                    const zd p = zmul({Lr[r][k], Li[r][k]}, {Xr[k][c], Xi[k][c]});
                    acc.r += p.r;
                    acc.i += p.i;
                }
                const zd x = zmul(acc, {vr[r], vi[r]});
                Xr[r][c] = -x.r;
                Xi[r][c] = -x.i;
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < ZNB * ZNB; e += 256) {
        const int r = e & (ZNB - 1), c = e >> 6;
        xout[c * ZNB + r] = Xr[r][c];
        xout[P + c * ZNB + r] = Xi[r][c];
    }
    if (tid == 0) {   // fixed order
        double lr = 0.0, li = 0.0;
        for (int j = 0; j < kb; j++) {
            lr += log(hypot(Lr[j][j], Li[j][j]));
            li += atan2(Li[j][j], Lr[j][j]);
        }
        part[0] = lr;
        part[P] = li;
        if (bad_s != INT_MAX) atomicMin(info, info_base + bad_s + 1);
    }
}

__global__ void zset_int_kernel(int* p, int v) { *p = v; }

// sqrt(sum |z|^2) over n complex entries (planes P apart), one workgroup, fixed-order tree
__global__ __launch_bounds__(1024) void znorm2_kernel(const double* z, long long n, long long P, double* out)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 1024) s += z[i] * z[i] + z[P + i] * z[P + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sqrt(red[0]);
}

}  // namespace
}  // namespace rgf

using namespace rgf;

struct rgf_zhandle_s {
    int ns = 0, nt = 0, nd = 0;
    long long n = 0;
    rgf_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    int ldp = 0, ldf = 0, T = 0;
    long long pstride = 0, mstride = 0;
    long long P = 0;            // plane offset (doubles) of every complex array of the handle
    double* base = nullptr;     // 2 P doubles
    // real planes (the imaginary plane of each is P doubles further)
    double *Ap = nullptr, *Add = nullptr, *Lp = nullptr, *Ldd = nullptr, *Xdd = nullptr, *Ftmp = nullptr;
    double *ttmp = nullptr, *xdiag = nullptr, *ptmp = nullptr, *parts = nullptr;
    double *Sb[2] = {nullptr, nullptr}, *W = nullptr, *Z = nullptr, *diag = nullptr;
    double *Y = nullptr, *X = nullptr, *Bv = nullptr, *R = nullptr;
    double* scal = nullptr;     // 4 doubles
    int* dinfo = nullptr;
    bool factorized = false;
    std::string err;
};

namespace {

thread_local std::string g_zcreate_error;

int zset_err(rgf_zhandle_s* h, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf; else g_zcreate_error = buf;
    return code;
}

int zdev(rgf_zhandle_s* h, int hip_rc, const char* where)
{
    if (hip_rc == 0) return RGF_OK;
    return zset_err(h, hip_rc == (int)hipErrorOutOfMemory ? RGF_ENOMEM : RGF_EDEVICE, "%s: HIP error %d (%s)", where,
                    hip_rc, hipGetErrorString((hipError_t)hip_rc));
}

// C = alpha op(A) op(B) + beta Cin for complex operands (planes P apart): two launches of the
// FP64 GEMM, each carrying the real and the imaginary part of up to two complex products.
int zgemm(Op ta, Op tb, const GemmDesc* d, int nd, long long P, hipStream_t s)
{
    for (int q0 = 0; q0 < nd; q0 += 2) {
        const int nq = std::min(2, nd - q0);
        GemmDesc l1[4], l2[4];
        for (int q = 0; q < nq; q++) {
            const GemmDesc& c = d[q0 + q];
            GemmDesc rr = c, ri = c, ir = c, ii = c;
            ri.B = c.B + P; ri.C = c.C + P; ri.Cin = c.Cin + P;                        // C_i = a A_r B_i + b Cin_i
            ir.A = c.A + P; ir.B = c.B + P; ir.alpha = -c.alpha; ir.beta = 1.0;         // C_r -= a A_i B_i
            ir.Cin = c.C; ir.ldcin = c.ldc; ir.sCin = c.sC; ir.sCin2 = c.sC2;
            ii.A = c.A + P; ii.C = c.C + P; ii.beta = 1.0;                              // C_i += a A_i B_r
            ii.Cin = c.C + P; ii.ldcin = c.ldc; ii.sCin = c.sC; ii.sCin2 = c.sC2;
            l1[2 * q] = rr; l1[2 * q + 1] = ri;
            l2[2 * q] = ir; l2[2 * q + 1] = ii;
        }
        int rc = gemm_f64(ta, tb, l1, 2 * nq, s);
        if (rc) return rc;
        if ((rc = gemm_f64(ta, tb, l2, 2 * nq, s))) return rc;
    }
    return 0;
}

int zprod_cb(void* ctx, Op ta, Op tb, GemmDesc* d, int nd, hipStream_t s)
{
    return zgemm(ta, tb, d, nd, ((rgf_zhandle_s*)ctx)->P, s);
}

int zgemm1(rgf_zhandle_s* h, Op ta, Op tb, GemmDesc d) { return zgemm(ta, tb, &d, 1, h->P, h->stream); }

// copy of a complex rows x cols block (both planes)
int zcopy(rgf_zhandle_s* h, const double* src, int lds, double* dst, int ldd, int rows, int cols)
{
    return copy2d(src, lds, dst, ldd, rows, cols, 1.0, h->stream, 2, h->P, h->P);
}

// blocked right-looking complex-symmetric Cholesky of the n x n block at A (lower part
// read); xdiag receives the inverted 64 x 64 diagonal tiles, parts[p] their log sums
int zpotrf(rgf_zhandle_s* h, double* A, int lda, int n, double* parts, int info_base)
{
    const long long P = h->P;
    hipStream_t s = h->stream;
    int rc, p = 0;
    for (int k0 = 0; k0 < n; k0 += ZNB, p++) {
        const int kb = std::min(ZNB, n - k0), below = n - k0 - kb;
        double* xt = h->xdiag + (long long)p * ZNB * ZNB;
        hipLaunchKernelGGL(ztile_kernel, dim3(1), dim3(256), 0, s, A + (long long)k0 * lda + k0, lda, P, kb, xt,
                           parts + p, h->dinfo, info_base + k0);
        if ((rc = (int)hipGetLastError())) return rc;
        if (below <= 0) continue;
        double* Ab = A + (long long)k0 * lda + k0 + kb;
        if ((rc = zcopy(h, Ab, lda, h->ptmp, below, below, kb))) return rc;
        // L[below, k] = A[below, k] X_kk^T   (X_kk^T upper)
        if ((rc = zgemm1(h, Op::N, Op::T, gdesc(h->ptmp, below, xt, ZNB, Ab, lda, below, kb, kb, 1.0, 0.0, G_BUPPER))))
            return rc;
        // trailing lower part -= L[below, k] L[below, k]^T
        double* Ct = A + (long long)(k0 + kb) * lda + k0 + kb;
        if ((rc = zgemm1(h, Op::N, Op::T, gdesc(Ab, lda, Ab, lda, Ct, lda, below, below, kb, -1.0, 1.0, G_CLOWER))))
            return rc;
    }
    return 0;
}

// FirstStageFactor (RGF.H:233-340) in complex-symmetric algebra; panel layout as the FP64
// path (diagonal block -> L_ii^-1, coupling -> B_i, arrow -> D_i), corner -> X_dd (64 x 64 tile).
int zfactor_sweep(rgf_zhandle_s* h)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp, ldf = h->ldf;
    hipStream_t s = h->stream;
    int rc;
    hipLaunchKernelGGL(zset_int_kernel, dim3(1), dim3(1), 0, s, h->dinfo, INT_MAX);
    if ((rc = (int)hipGetLastError())) return rc;
    const ProdFn zp{zprod_cb, h};
    for (int i = 0; i < nt; i++) {
        double* Pi = h->Lp + (long long)i * h->pstride;
        const double* Ai = h->Ap + (long long)i * h->pstride;
        if (i == 0) {
            if ((rc = zcopy(h, Ai, ldp, Pi, ldp, ns, ns))) return rc;
            if (nd > 0 && (rc = zcopy(h, Ai + 2 * ns, ldp, h->Ftmp + ns, ldf, nd, ns))) return rc;
        } else {
            const double* Pm = h->Lp + (long long)(i - 1) * h->pstride;
            GemmDesc d = gdesc(Pm + ns, ldp, Pm + ns, ldp, Pi, ldp, ns, ns, ns, -1.0, 1.0, G_CLOWER);
            d.Cin = Ai;
            d.ldcin = ldp;
            if ((rc = zgemm1(h, Op::N, Op::T, d))) return rc;   // S_i = A_ii - B_{i-1} B_{i-1}^T
            if (nd > 0) {                                        // A_di - D_{i-1} B_{i-1}^T
                GemmDesc a = gdesc(Pm + 2 * ns, ldp, Pm + ns, ldp, h->Ftmp + ns, ldf, nd, ns, ns, -1.0, 1.0, 0);
                a.Cin = Ai + 2 * ns;
                a.ldcin = ldp;
                if ((rc = zgemm1(h, Op::N, Op::T, a))) return rc;
            }
        }
        if ((rc = zpotrf(h, Pi, ldp, ns, h->parts + (long long)i * h->T, i * ns))) return rc;
        if ((rc = trtri_f64_inplace(Pi, ldp, 0, ns, 1, h->ttmp, 0, h->xdiag, 0, s, &zp, h->P))) return rc;
        const bool has_next = i < nt - 1;
        if (has_next && (rc = zcopy(h, Ai + ns, ldp, h->Ftmp, ldf, ns, ns))) return rc;
        const int mF = (has_next ? ns : 0) + nd;
        if (mF > 0) {   // [B_i; D_i] = [A_{i+1,i}; A_di'] X_i^T
            const double* src = h->Ftmp + (has_next ? 0 : ns);
            double* dst = Pi + (has_next ? ns : 2 * ns);
            if ((rc = zgemm1(h, Op::N, Op::T, gdesc(src, ldf, Pi, ldp, dst, ldp, mF, ns, ns, 1.0, 0.0, G_BUPPER))))
                return rc;
        }
    }
    if (nd > 0) {   // S_dd = A_dd - sum_i D_i D_i^T (the arrow rows of all panels: nd x nt*ns, ld ldp)
        GemmDesc d = gdesc(h->Lp + 2 * ns, ldp, h->Lp + 2 * ns, ldp, h->Ldd, nd, nd, nd, nt * ns, -1.0, 1.0, G_CLOWER);
        d.Cin = h->Add;
        d.ldcin = nd;
        if ((rc = zgemm1(h, Op::N, Op::T, d))) return rc;
        hipLaunchKernelGGL(ztile_kernel, dim3(1), dim3(256), 0, s, h->Ldd, nd, h->P, nd, h->Xdd,
                           h->parts + (long long)nt * h->T, h->dinfo, nt * ns);
        if ((rc = (int)hipGetLastError())) return rc;
    }
    return 0;
}

int zensure_factorized(rgf_zhandle_s* h)
{
    if (h->factorized) return RGF_OK;
    hipSetDevice(h->device);
    int rc = zfactor_sweep(h);
    if (rc) return zdev(h, rc, "factorize");
    int info = INT_MAX;
    hipError_t e = hipMemcpyAsync(&info, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return zdev(h, (int)e, "factorize");
    if (info != INT_MAX)
        return zset_err(h, RGF_ENOTSPD, "zero or non-finite pivot of the complex-symmetric factorization at column %d",
                        info);
    h->factorized = true;
    return RGF_OK;
}

// SecondStageSolve (RGF.H:344-418): x (n x nr, ld n, both planes) from b (may alias x)
int zsolve_block(rgf_zhandle_s* h, double* x, const double* b, int nr)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp, n = (int)h->n;
    double* Y = h->Y;
    int rc;
    if (x != b && (rc = zcopy(h, b, n, x, n, n, nr))) return rc;
    double* xd = x + (long long)nt * ns;
    double* yd = Y + (long long)nt * ns;
    for (int i = 0; i < nt; i++) {
        const double* Pi = h->Lp + (long long)i * h->pstride;
        double* ri = x + (long long)i * ns;
        double* yi = Y + (long long)i * ns;
        if (i > 0) {
            const double* Pm = h->Lp + (long long)(i - 1) * h->pstride;
            if ((rc = zgemm1(h, Op::N, Op::N, gdesc(Pm + ns, ldp, Y + (long long)(i - 1) * ns, n, ri, n, ns, nr, ns,
                                                    -1.0, 1.0))))
                return rc;
        }
        if ((rc = zgemm1(h, Op::N, Op::N, gdesc(Pi, ldp, ri, n, yi, n, ns, nr, ns, 1.0, 0.0, G_ALOWER)))) return rc;
        if (nd > 0 && (rc = zgemm1(h, Op::N, Op::N, gdesc(Pi + 2 * ns, ldp, yi, n, xd, n, nd, nr, ns, -1.0, 1.0))))
            return rc;
    }
    if (nd > 0) {
        if ((rc = zgemm1(h, Op::N, Op::N, gdesc(h->Xdd, ZNB, xd, n, yd, n, nd, nr, nd, 1.0, 0.0, G_ALOWER)))) return rc;
        if ((rc = zgemm1(h, Op::T, Op::N, gdesc(h->Xdd, ZNB, yd, n, xd, n, nd, nr, nd, 1.0, 0.0, G_AUPPER)))) return rc;
    }
    for (int i = nt - 1; i >= 0; i--) {
        const double* Pi = h->Lp + (long long)i * h->pstride;
        double* yi = Y + (long long)i * ns;
        double* xi = x + (long long)i * ns;
        if (nd > 0 && (rc = zgemm1(h, Op::T, Op::N, gdesc(Pi + 2 * ns, ldp, xd, n, yi, n, ns, nr, nd, -1.0, 1.0))))
            return rc;
        if (i < nt - 1 && (rc = zgemm1(h, Op::T, Op::N, gdesc(Pi + ns, ldp, x + (long long)(i + 1) * ns, n, yi, n, ns,
                                                              nr, ns, -1.0, 1.0))))
            return rc;
        if ((rc = zgemm1(h, Op::T, Op::N, gdesc(Pi, ldp, yi, n, xi, n, ns, nr, ns, 1.0, 0.0, G_AUPPER)))) return rc;
    }
    return 0;
}

// ThirdStageRGF (RGF.H:422-506), G-formulation as the FP64 path: W = F X_i, Z = Sb W,
// S_ii = X_i^T X_i + W^T Z, S_{d,i} = -Z[arrow rows]; exact = 0 keeps the reference's formula.
int zselinv_sweep(rgf_zhandle_s* h)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp, lds = h->ldf, m = ns + nd;
    const long long P = h->P;
    const int exact = h->opt.selinv_mode == RGF_SELINV_EXACT;
    hipStream_t s = h->stream;
    int rc;
    for (int q = 0; q < 2; q++)
        if ((rc = fill2d(h->Sb[q], lds, m, m, 0.0, s, 2, P))) return rc;
    if (nd > 0) {
        for (int q = 0; q < 2; q++) {
            double* Sdd = h->Sb[q] + (long long)ns * lds + ns;
            if ((rc = zgemm1(h, Op::T, Op::N, gdesc(h->Xdd, ZNB, h->Xdd, ZNB, Sdd, lds, nd, nd, nd, 1.0, 0.0, 0))))
                return rc;
        }
        if ((rc = extract_diag(h->Sb[0] + (long long)ns * lds + ns, lds, nd, h->diag + (long long)nt * ns, s, 2, P, P)))
            return rc;
    }
    int cur = 0;
    for (int i = nt - 1; i >= 0; i--) {
        const double* Pi = h->Lp + (long long)i * h->pstride;
        const bool last = i == nt - 1;
        const int mF = last ? nd : m;
        const double* F = last ? Pi + 2 * ns : Pi + ns;
        const int r0 = last ? ns : 0;
        double* Sc = h->Sb[cur];
        double* Sn = h->Sb[cur ^ 1];
        // X_i^T X_i (lower part) into Sn
        if ((rc = zgemm1(h, Op::T, Op::N, gdesc(Pi, ldp, Pi, ldp, Sn, lds, ns, ns, ns, 1.0, 0.0,
                                                G_CLOWER | G_AUPPER | G_BLOWER))))
            return rc;
        if (mF > 0) {
            if ((rc = zgemm1(h, Op::N, Op::N, gdesc(F, ldp, Pi, ldp, h->W, lds, mF, ns, ns, 1.0, 0.0, G_BLOWER))))
                return rc;
            if ((rc = zgemm1(h, Op::N, Op::N, gdesc(Sc + (long long)r0 * lds + r0, lds, h->W, lds, h->Z, lds, mF, ns, mF,
                                                    1.0, 0.0, 0))))
                return rc;
            if ((rc = zgemm1(h, Op::T, Op::N, gdesc(h->W, lds, h->Z, lds, Sn, lds, ns, ns, mF, 1.0, 1.0, G_CLOWER))))
                return rc;
        }
        if ((rc = selinv_finish(Sn, lds, P, ns, nd, h->Z, lds, P, mF - nd, exact && mF > 0,
                                h->diag + (long long)i * ns, P, 2, s)))
            return rc;
        cur ^= 1;
    }
    return 0;
}

// r = b - A x with the resident complex A (full diagonal blocks and corner), one column
int zresidual(rgf_zhandle_s* h, const double* x, const double* b, double* r)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp, n = (int)h->n;
    int rc;
    if ((rc = zcopy(h, b, n, r, n, n, 1))) return rc;
    const double* xd = x + (long long)nt * ns;
    double* rd = r + (long long)nt * ns;
    for (int t = 0; t < nt; t++) {
        const double* At = h->Ap + (long long)t * h->pstride;
        const double* xt = x + (long long)t * ns;
        double* rt = r + (long long)t * ns;
        if ((rc = zgemm1(h, Op::N, Op::N, gdesc(At, ldp, xt, n, rt, n, ns, 1, ns, -1.0, 1.0)))) return rc;
        if (t < nt - 1) {
            if ((rc = zgemm1(h, Op::T, Op::N, gdesc(At + ns, ldp, xt + ns, n, rt, n, ns, 1, ns, -1.0, 1.0)))) return rc;
            if ((rc = zgemm1(h, Op::N, Op::N, gdesc(At + ns, ldp, xt, n, rt + ns, n, ns, 1, ns, -1.0, 1.0)))) return rc;
        }
        if (nd > 0) {
            if ((rc = zgemm1(h, Op::T, Op::N, gdesc(At + 2 * ns, ldp, xd, n, rt, n, ns, 1, nd, -1.0, 1.0)))) return rc;
            if ((rc = zgemm1(h, Op::N, Op::N, gdesc(At + 2 * ns, ldp, xt, n, rd, n, nd, 1, ns, -1.0, 1.0)))) return rc;
        }
    }
    if (nd > 0 && (rc = zgemm1(h, Op::N, Op::N, gdesc(h->Add, nd, xd, n, rd, n, nd, 1, nd, -1.0, 1.0)))) return rc;
    return 0;
}

// host interleaved (re, im) pairs <-> the two planes of a device array
int zupload(rgf_zhandle_s* h, const double* z, double* dst, size_t count)
{
    std::vector<double> re(count), im(count);
    for (size_t k = 0; k < count; k++) {
        re[k] = z[2 * k];
        im[k] = z[2 * k + 1];
    }
    hipError_t e = hipMemcpy(dst, re.data(), count * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dst + h->P, im.data(), count * sizeof(double), hipMemcpyHostToDevice);
    return (int)e;
}

int zdownload(rgf_zhandle_s* h, const double* src, double* z, size_t count)
{
    std::vector<double> re(count), im(count);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(re.data(), src, count * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(im.data(), src + h->P, count * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return (int)e;
    for (size_t k = 0; k < count; k++) {
        z[2 * k] = re[k];
        z[2 * k + 1] = im[k];
    }
    return 0;
}

int zresidual_norm(rgf_zhandle_s* h, const double* x, const double* b, double* out, bool normalized)
{
    int rc = RGF_OK;
    hipSetDevice(h->device);
    const size_t n = (size_t)h->n;
    if ((rc = zupload(h, x, h->X, n)) || (rc = zupload(h, b, h->Bv, n))) return zdev(h, rc, "residual upload");
    if ((rc = zresidual(h, h->X, h->Bv, h->R))) return zdev(h, rc, "residual");
    hipLaunchKernelGGL(znorm2_kernel, dim3(1), dim3(1024), 0, h->stream, h->R, h->n, h->P, h->scal);
    hipLaunchKernelGGL(znorm2_kernel, dim3(1), dim3(1024), 0, h->stream, h->Bv, h->n, h->P, h->scal + 1);
    double v[2];
    hipError_t e = hipMemcpyAsync(v, h->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return zdev(h, (int)e, "residual");
    *out = normalized ? v[0] / v[1] : v[0];
    return RGF_OK;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int rgf_z_destroy(rgf_zhandle_t h)
{
    if (!h) return RGF_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->base) hipFree(h->base);
    if (h->scal) hipFree(h->scal);
    if (h->dinfo) hipFree(h->dinfo);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return RGF_OK;
}

int rgf_z_create_csc(const size_t* ia, const size_t* ja, const double* a, size_t ns, size_t nt, size_t nd,
                     const rgf_options* opts, rgf_zhandle_t* out)
{
    if (!ia || !ja || !a || !out) return zset_err(nullptr, RGF_EINVAL, "null argument");
    *out = nullptr;
    if (ns == 0 || nt == 0) return zset_err(nullptr, RGF_EINVAL, "ns and nt must be positive (ns=%zu nt=%zu)", ns, nt);
    if (nd > ns) return zset_err(nullptr, RGF_EINVAL, "nd=%zu > ns=%zu is not supported", nd, ns);
    if (nd > 64) return zset_err(nullptr, RGF_EINVAL, "nd=%zu > 64 is not supported", nd);
    if (2 * ns + nd > 65536 || ns * nt + nd > (size_t)INT_MAX / 16)
        return zset_err(nullptr, RGF_EINVAL, "matrix too large (ns=%zu nt=%zu nd=%zu)", ns, nt, nd);
    const size_t n = ns * nt + nd, sn = ns * nt;
    if (ia[0] != 0) return zset_err(nullptr, RGF_EINVAL, "ia[0] must be 0");
    for (size_t c = 0; c < n; c++) {
        if (ia[c + 1] < ia[c]) return zset_err(nullptr, RGF_EINVAL, "ia is not non-decreasing at column %zu", c);
        for (size_t k = ia[c]; k < ia[c + 1]; k++) {
            const size_t r = ja[k];
            const bool ok = !(r < c || r >= n) && (c >= sn || r >= sn || r / ns == c / ns || r / ns == c / ns + 1);
            if (!ok)
                return zset_err(nullptr, RGF_EINVAL,
                                "CSC entry %zu (row %zu, column %zu) is outside the lower block-arrow pattern", k, r, c);
        }
    }
    auto* h = new rgf_zhandle_s();
    rgf_default_options(&h->opt);
    if (opts) h->opt = *opts;
    if (h->opt.precision != RGF_FP64) {
        const int prec = h->opt.precision;
        delete h;
        return zset_err(nullptr, RGF_EINVAL, "the complex solver runs in FP64 only (precision %d)", prec);
    }
    h->ns = (int)ns; h->nt = (int)nt; h->nd = (int)nd; h->n = (long long)n;
    h->device = h->opt.device;
    hipError_t e = hipSetDevice(h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        int rc = zset_err(nullptr, RGF_EDEVICE, "no usable HIP device %d (%s)", h->device, hipGetErrorString(e));
        rgf_z_destroy(h);
        return rc;
    }
    // one plane's layout
    h->ldp = (int)((2 * ns + nd + 7) / 8 * 8);
    h->ldf = (int)((ns + nd + 7) / 8 * 8);
    h->T = (int)((ns + 63) / 64);
    h->pstride = (long long)h->ldp * (long long)ns;
    h->mstride = h->pstride * (long long)nt;
    size_t tt = 1;
    for (int sz = 1; sz < h->T; sz *= 2) {
        const size_t G = (size_t)((h->T - sz + 2 * sz - 1) / (2 * sz));
        tt = std::max(tt, G * (size_t)(sz * 64) * (size_t)(sz * 64));
    }
    size_t off = 0;
    std::vector<std::pair<double**, size_t>> parts_list;
    auto take = [&](double** p, size_t cnt) { parts_list.push_back({p, off}); off += (std::max<size_t>(cnt, 1) + 31) / 32 * 32; };
    const size_t cnd = std::max<size_t>(nd * nd, 1), n16 = n * 16;
    take(&h->Ap, (size_t)h->mstride);
    take(&h->Add, cnd);
    take(&h->Lp, (size_t)h->mstride);
    take(&h->Ldd, cnd);
    take(&h->Xdd, 64 * 64);
    take(&h->Ftmp, (size_t)h->ldf * ns);
    take(&h->ttmp, tt);
    take(&h->xdiag, (size_t)h->T * 64 * 64);
    take(&h->ptmp, ns * 64);
    take(&h->parts, (size_t)nt * h->T + 1);
    take(&h->Sb[0], (size_t)h->ldf * (ns + nd));
    take(&h->Sb[1], (size_t)h->ldf * (ns + nd));
    take(&h->W, (size_t)h->ldf * ns);
    take(&h->Z, (size_t)h->ldf * ns);
    take(&h->diag, n);
    take(&h->Y, n16);
    take(&h->X, n16);
    take(&h->Bv, n16);
    take(&h->R, n16);
    h->P = (long long)off;
    e = hipMalloc((void**)&h->base, 2 * off * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&h->scal, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&h->dinfo, sizeof(int));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        int rc = zset_err(nullptr, RGF_ENOMEM, "device allocation of %zu bytes failed (%s)", 2 * off * sizeof(double),
                          hipGetErrorString(e));
        rgf_z_destroy(h);
        return rc;
    }
    for (auto& q : parts_list) *q.first = h->base + q.second;
    // matrix: zero, then scatter the real and the imaginary parts of the CSC values
    const size_t nnz = ia[n];
    std::vector<double> are(std::max<size_t>(nnz, 1)), aim(std::max<size_t>(nnz, 1));
    for (size_t k = 0; k < nnz; k++) {
        are[k] = a[2 * k];
        aim[k] = a[2 * k + 1];
    }
    uint64_t *dia = nullptr, *dja = nullptr;
    double* dv = nullptr;
    e = hipMalloc((void**)&dia, (n + 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&dja, std::max<size_t>(nnz, 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&dv, 2 * std::max<size_t>(nnz, 1) * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(dia, ia, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(dja, ja, nnz * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(dv, are.data(), nnz * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(dv + nnz, aim.data(), nnz * sizeof(double), hipMemcpyHostToDevice);
    int krc = 0;
    if (e == hipSuccess) {
        krc = fill2d(h->Ap, h->ldp, h->ldp, (int)(nt * ns), 0.0, h->stream, 2, h->P);
        if (!krc && nd > 0) krc = fill2d(h->Add, (int)nd, (int)nd, (int)nd, 0.0, h->stream, 2, h->P);
        for (int pl = 0; pl < 2 && !krc; pl++)
            krc = scatter_csc(dia, dja, dv + pl * nnz, (long long)nnz, (int)ns, (int)nt, (int)nd, h->Ap + pl * h->P,
                              h->pstride, h->ldp, h->Add + pl * h->P, h->stream);
        hipError_t es = hipStreamSynchronize(h->stream);
        if (!krc && es != hipSuccess) e = es;
    }
    if (dia) hipFree(dia);
    if (dja) hipFree(dja);
    if (dv) hipFree(dv);
    if (e != hipSuccess || krc) {
        int rc = zset_err(nullptr, RGF_EDEVICE, "CSC upload: HIP error %d", e != hipSuccess ? (int)e : krc);
        rgf_z_destroy(h);
        return rc;
    }
    *out = h;
    return RGF_OK;
}

int rgf_z_factorize(rgf_zhandle_t h)
{
    if (!h) return RGF_EINVAL;
    h->factorized = false;
    return zensure_factorized(h);
}

int rgf_z_solve(rgf_zhandle_t h, double* x, const double* b, size_t nrhs)
{
    if (!h || !x || !b) return RGF_EINVAL;
    if (nrhs == 0) return RGF_OK;
    int rc = zensure_factorized(h);
    if (rc) return rc;
    const size_t n = (size_t)h->n;
    for (size_t c0 = 0; c0 < nrhs; c0 += 16) {
        const int nr = (int)std::min<size_t>(16, nrhs - c0);
        if ((rc = zupload(h, b + 2 * c0 * n, h->Bv, n * nr))) return zdev(h, rc, "solve upload");
        if ((rc = zsolve_block(h, h->X, h->Bv, nr))) return zdev(h, rc, "solve");
        if ((rc = zdownload(h, h->X, x + 2 * c0 * n, n * nr))) return zdev(h, rc, "solve download");
    }
    return RGF_OK;
}

int rgf_z_logdet(rgf_zhandle_t h, double* out)
{
    if (!h || !out) return RGF_EINVAL;
    int rc = zensure_factorized(h);
    if (rc) return rc;
    const size_t np = (size_t)h->nt * h->T + 1;
    std::vector<double> re(np), im(np);
    hipError_t e = hipMemcpyAsync(re.data(), h->parts, np * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(im.data(), h->parts + h->P, np * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return zdev(h, (int)e, "logdet");
    const size_t used = (size_t)h->nt * h->T + (h->nd > 0 ? 1 : 0);
    double sr = 0.0, si = 0.0;
    for (size_t k = 0; k < used; k++) {   // fixed order
        sr += re[k];
        si += im[k];
    }
    out[0] = 2.0 * sr;
    out[1] = 2.0 * si;
    return RGF_OK;
}

int rgf_z_selinv_diag(rgf_zhandle_t h, double* diag)
{
    if (!h || !diag) return RGF_EINVAL;
    int rc = zensure_factorized(h);
    if (rc) return rc;
    if ((rc = zselinv_sweep(h))) return zdev(h, rc, "selinv");
    if ((rc = zdownload(h, h->diag, diag, (size_t)h->n))) return zdev(h, rc, "selinv download");
    return RGF_OK;
}

int rgf_z_residual_norm(rgf_zhandle_t h, const double* x, const double* b, double* out)
{
    if (!h || !x || !b || !out) return RGF_EINVAL;
    return zresidual_norm(h, x, b, out, false);
}

int rgf_z_residual_norm_normalized(rgf_zhandle_t h, const double* x, const double* b, double* out)
{
    if (!h || !x || !b || !out) return RGF_EINVAL;
    return zresidual_norm(h, x, b, out, true);
}

const char* rgf_z_last_error(rgf_zhandle_t h) { return h ? h->err.c_str() : g_zcreate_error.c_str(); }

}  // extern "C"
