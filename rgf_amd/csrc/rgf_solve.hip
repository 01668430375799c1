// Skinny FP64 products Y = alpha*op(A)*R + beta*Y0 with nr <= 16 right-hand
// sides: the HBM-bound block solves (SecondStageSolve, RGF.H:344-418, which the
// reference runs on the host with dtrsm_/dgemm_) and the block matvec of the
// residual (RGF.H:631-663).  The factor stores L_ii^-1, so every triangular
// solve is a product here and the whole solve is a stream over the factor.
//
// N: one lane per row of op(A) (coalesced column reads), the R slab staged in
//    LDS; k split into slabs across workgroups for occupancy, partials reduced
//    in a fixed order (deterministic).
// T: one wave per output column (a column of A read contiguously by 64 lanes),
//    wave-level reduction; k split into slabs only when there are few columns.
#include "rgf_internal.hpp"
#include <algorithm>

namespace rgf {

constexpr int MAXR = 16;
constexpr int KCH = 64;   // k-chunk staged in LDS

static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

struct SkinnyPlan {
    int slabs, kslab;
};

static SkinnyPlan plan_n(int m, int k)
{
    const int wgm = cdiv(m, 256);
    int slabs = std::max(1, std::min(cdiv(k, KCH), cdiv(768, wgm)));
    slabs = std::min(slabs, 64);
    int kslab = cdiv(cdiv(k, slabs), KCH) * KCH;
    slabs = cdiv(k, kslab);
    return {std::max(slabs, 1), kslab};
}

static SkinnyPlan plan_t(int m, int k)
{
    const int wgm = cdiv(m, 4);
    int slabs = std::max(1, std::min(cdiv(k, 256), cdiv(1024, wgm)));
    slabs = std::min(slabs, 64);
    int kslab = cdiv(cdiv(k, slabs), 64) * 64;
    slabs = cdiv(k, kslab);
    return {std::max(slabs, 1), kslab};
}

size_t skinny_work_elems(int m, int k, int nr)
{
    (void)k;
    (void)nr;
    return (size_t)64 * (size_t)m * (size_t)MAXR;
}

// partial layout: work[(s * nr + c) * m + i]
template <int NR>
__global__ __launch_bounds__(256) void skinny_n_kernel(int m, int k, int nr, int kslab, int tri, double alpha,
                                                       const double* __restrict__ A, int lda,
                                                       const double* __restrict__ R, int ldr, double beta,
                                                       const double* Y0, int ldy0, double* Y, int ldy,
                                                       double* work, int direct)
{
    __shared__ double Rs[KCH][NR];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 256 + tid;
    const int s = blockIdx.y;
    const int kb = s * kslab, ke = min(k, kb + kslab);
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
    const int rowmax = min(m, (int)(blockIdx.x * 256 + 256)) - 1;
    int kend = ke;
    if (tri == 1) kend = min(kend, rowmax + 1);   // op(A) lower: A[i][kk] == 0 for kk > i
    int kbeg = kb;
    if (tri == 2) kbeg = max(kbeg, (int)(blockIdx.x * 256));   // op(A) upper
    for (int k0 = kbeg; k0 < kend; k0 += KCH) {
        const int kc = min(KCH, kend - k0);
        __syncthreads();
        for (int e = tid; e < KCH * NR; e += 256) {
            int kk = e % KCH, c = e / KCH;
            Rs[kk][c] = (kk < kc && c < nr) ? R[(long long)c * ldr + k0 + kk] : 0.0;
        }
        __syncthreads();
        if (i < m) {
            const double* Ap = A + (long long)k0 * lda + i;
            int kk = 0;
            for (; kk + 4 <= kc; kk += 4) {
                double a0 = Ap[(long long)(kk + 0) * lda];
                double a1 = Ap[(long long)(kk + 1) * lda];
                double a2 = Ap[(long long)(kk + 2) * lda];
                double a3 = Ap[(long long)(kk + 3) * lda];
#pragma unroll
                for (int c = 0; c < NR; c++)
                    acc[c] += a0 * Rs[kk][c] + a1 * Rs[kk + 1][c] + a2 * Rs[kk + 2][c] + a3 * Rs[kk + 3][c];
            }
            for (; kk < kc; kk++) {
                double a0 = Ap[(long long)kk * lda];
#pragma unroll
                for (int c = 0; c < NR; c++) acc[c] += a0 * Rs[kk][c];
            }
        }
    }
    if (i >= m) return;
    if (direct) {
#pragma unroll
        for (int c = 0; c < NR; c++) {
            if (c >= nr) break;
            double v = alpha * acc[c];
            if (beta != 0.0) v += beta * Y0[(long long)c * ldy0 + i];
            Y[(long long)c * ldy + i] = v;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NR; c++) work[((long long)s * NR + c) * m + i] = acc[c];
    }
}

template <int NR>
__global__ __launch_bounds__(256) void skinny_t_kernel(int m, int k, int nr, int kslab, int tri, double alpha,
                                                       const double* __restrict__ A, int lda,
                                                       const double* __restrict__ R, int ldr, double beta,
                                                       const double* Y0, int ldy0, double* Y, int ldy,
                                                       double* work, int direct)
{
    __shared__ double Rs[NR][KCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.x * 4 + wave;   // output row (column of A)
    const int s = blockIdx.y;
    const int kb = s * kslab, ke = min(k, kb + kslab);
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
    int kbeg = kb, kend = ke;
    const int jmin = blockIdx.x * 4, jmax = min(m, jmin + 4) - 1;
    if (tri == 2) kbeg = max(kbeg, (jmin / KCH) * KCH);   // op(A) upper: A^T[j][kk] == 0 for kk < j
    if (tri == 1) kend = min(kend, jmax + 1);             // op(A) lower
    for (int k0 = kbeg; k0 < kend; k0 += KCH) {
        const int kc = min(KCH, kend - k0);
        __syncthreads();
        for (int e = tid; e < KCH * NR; e += 256) {
            int kk = e % KCH, c = e / KCH;
            Rs[c][kk] = (kk < kc && c < nr) ? R[(long long)c * ldr + k0 + kk] : 0.0;
        }
        __syncthreads();
        if (j < m && lane < kc) {
            double a = A[(long long)j * lda + k0 + lane];
#pragma unroll
            for (int c = 0; c < NR; c++) acc[c] += a * Rs[c][lane];
        }
    }
#pragma unroll
    for (int c = 0; c < NR; c++) {
        double v = acc[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        acc[c] = v;
    }
    if (j >= m || lane != 0) return;
    if (direct) {
#pragma unroll
        for (int c = 0; c < NR; c++) {
            if (c >= nr) break;
            double v = alpha * acc[c];
            if (beta != 0.0) v += beta * Y0[(long long)c * ldy0 + j];
            Y[(long long)c * ldy + j] = v;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NR; c++) work[((long long)s * NR + c) * m + j] = acc[c];
    }
}

__global__ void skinny_reduce_kernel(int m, int nr, int wstride, int slabs, double alpha, const double* work,
                                     double beta, const double* Y0, int ldy0, double* Y, int ldy)
{
    const long long total = (long long)m * nr;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % m), c = (int)(e / m);
        double v = 0.0;
        for (int s = 0; s < slabs; s++) v += work[((long long)s * wstride + c) * m + i];
        v *= alpha;
        if (beta != 0.0) v += beta * Y0[(long long)c * ldy0 + i];
        Y[(long long)c * ldy + i] = v;
    }
}

template <int NR>
static int launch_skinny(Op ta, int tri, int m, int k, int nr, double alpha, const double* A, int lda,
                         const double* R, int ldr, double beta, const double* Y0, int ldy0, double* Y,
                         int ldy, double* work, hipStream_t s)
{
    SkinnyPlan p = ta == Op::N ? plan_n(m, k) : plan_t(m, k);
    const int direct = p.slabs == 1;
    if (ta == Op::N) {
        dim3 grid(cdiv(m, 256), p.slabs);
        hipLaunchKernelGGL((skinny_n_kernel<NR>), grid, dim3(256), 0, s, m, k, nr, p.kslab, tri, alpha, A, lda,
                           R, ldr, beta, Y0, ldy0, Y, ldy, work, direct);
    } else {
        dim3 grid(cdiv(m, 4), p.slabs);
        hipLaunchKernelGGL((skinny_t_kernel<NR>), grid, dim3(256), 0, s, m, k, nr, p.kslab, tri, alpha, A, lda,
                           R, ldr, beta, Y0, ldy0, Y, ldy, work, direct);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    if (!direct) {
        long long total = (long long)m * nr;
        int grid = (int)std::min((total + 255) / 256, (long long)4096);
        hipLaunchKernelGGL(skinny_reduce_kernel, dim3(grid), dim3(256), 0, s, m, nr, NR, p.slabs, alpha, work,
                           beta, Y0, ldy0, Y, ldy);
        e = hipGetLastError();
    }
    return (int)e;
}

int skinny_f64(Op ta, int tri, int m, int k, int nr, double alpha, const double* A, int lda,
               const double* R, int ldr, double beta, const double* Y0, int ldy0, double* Y, int ldy,
               double* work, hipStream_t s)
{
    if (m <= 0 || nr <= 0) return 0;
    if (k <= 0) {   // Y = beta*Y0
        return copy2d(Y0, ldy0, Y, ldy, m, nr, beta, s);
    }
    if (nr == 1) return launch_skinny<1>(ta, tri, m, k, nr, alpha, A, lda, R, ldr, beta, Y0, ldy0, Y, ldy, work, s);
    if (nr <= 2) return launch_skinny<2>(ta, tri, m, k, nr, alpha, A, lda, R, ldr, beta, Y0, ldy0, Y, ldy, work, s);
    if (nr <= 4) return launch_skinny<4>(ta, tri, m, k, nr, alpha, A, lda, R, ldr, beta, Y0, ldy0, Y, ldy, work, s);
    if (nr <= 8) return launch_skinny<8>(ta, tri, m, k, nr, alpha, A, lda, R, ldr, beta, Y0, ldy0, Y, ldy, work, s);
    if (nr <= MAXR) return launch_skinny<16>(ta, tri, m, k, nr, alpha, A, lda, R, ldr, beta, Y0, ldy0, Y, ldy, work, s);
    return (int)hipErrorInvalidValue;
}

}  // namespace rgf
