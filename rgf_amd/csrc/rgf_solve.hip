// Skinny FP64 products Y = alpha*op(A)*R + beta*Y0 with nr <= 16 right-hand
// sides: the HBM-bound block solves (SecondStageSolve, RGF.H:344-418, which the
// reference runs on the host with dtrsm_/dgemm_) and the block matvec of the
// residual (RGF.H:631-663).  The factor stores L_ii^-1, so every triangular
// solve is a product here and the whole solve is a stream over the factor.
//
// One or two right-hand sides (FP64 VALU):
//   N: 64 rows of op(A) per workgroup, four lanes per row each over a quarter of every
//      k-chunk (coalesced column reads, 8 columns in flight), the R chunk staged in LDS;
//   T: one wave per output row (a column of A read contiguously by 64 lanes, 8 loads in
//      flight), wave-level reduction.
// Three to sixteen right-hand sides: FP64 MFMA, 64 rows of op(A) per workgroup.
// k is split into slabs across workgroups for memory-level parallelism; the partials are
// summed in a fixed order (deterministic), inside the product launch where that is a few
// partials per thread, else by a reduction launch.
#include "rgf_internal.hpp"
#include <algorithm>

namespace rgf {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int MAXR = 16;
constexpr int KCH = 128;   // k-chunk of R staged in LDS (N kernel)
constexpr int KU = 8;      // A loads in flight per lane
constexpr int NROWS = 64;  // rows of op(A) per workgroup of the N kernel
constexpr int KCT = 256;   // k-chunk of the T kernel (4 per lane x 64 lanes)
constexpr int KT1 = 8;     // A loads in flight per lane in the T kernel for one or two right-hand sides
// workgroups one product launch aims at (k split into slabs until there are about this many):
// 512 measured against 256 ... 2048 at ns = 2865 -- one-RHS solve 6.71 -> 6.10 ms (fewer
// partials to reduce, longer streams per workgroup), 16 RHS unchanged; 256 slows 16 RHS
int g_skinny_target_wg = 512;
constexpr int kMaxTargetWG = 2048;   // work buffers are sized for targets up to this

static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

struct SkinnyPlan {
    int slabs, kslab;
};

// k is split into slabs (partials reduced afterwards) until one matrix's launch has about
// `target` workgroups: the products stream A once from HBM and need many loads in flight.
// The split depends on (m, k) alone -- not on the batch count or the caller's buffer --
// so a matrix of a batch is summed in the same order as the same matrix alone (bitwise
// equal results); at most 64 slabs, which skinny_work_elems provides for one matrix.
static SkinnyPlan plan(int wgm, int k, int chunk, int target)
{
    int slabs = std::max(1, std::min({cdiv(k, chunk), cdiv(target, (long long)wgm), 64}));
    int kslab = cdiv(cdiv(k, slabs), chunk) * chunk;
    slabs = cdiv(k, kslab);
    return {std::max(slabs, 1), kslab};
}


// three or more right-hand sides: the MFMA kernel (64 rows of op(A) per workgroup)
static inline bool mfma_kernel(int nr) { return nr > 2; }

static SkinnyPlan plan_for(Op ta, int m, int k, int nr, int target = g_skinny_target_wg)
{
    if (mfma_kernel(nr)) return plan(cdiv(m, 64), k, KCT, target);
    if (ta == Op::T) return plan(cdiv(m, 4), k, KCT, target);
    return plan(cdiv(m, NROWS), k, KCH, target);
}

size_t skinny_batch_work_elems(Op ta, int m, int k, int count)
{
    // the most slabs any column count gives (the VALU kernels' partials -- at most twice the
    // slabs of two columns -- take less than the MFMA kernel's sixteen columns)
    const SkinnyPlan p = plan_for(ta, m, k, MAXR, kMaxTargetWG);
    return (size_t)count * (size_t)(p.slabs > 1 ? p.slabs : 1) * MAXR * (size_t)m;
}

size_t skinny_work_elems(int m, int k, int nr)
{
    (void)k;
    (void)nr;
    return (size_t)64 * (size_t)m * (size_t)MAXR;
}

// Operands by element strides: R(kk, c) = R[kk * rsk + c * rsj], Y(i, c) = Y[i * ysi + c * ysj]
// (column-major: rsk = ysi = 1), so a row-major R or Y is a transposed view.
struct SkinnyIO {
    const double* R;
    long long rsk, rsj;
    const double* Y0;
    long long y0si, y0sj;
    double* Y;
    long long ysi, ysj;
    long long sA, sR, sY0, sY;   // strides between the matrices of a batch (blockIdx.z)
    // split-k partials reduced in the product launch: one arrival counter per row block and
    // matrix (zero before the launch, left zero), alpha / beta of the epilogue (null: a
    // separate reduction launch)
    int* ctr;
    double alpha, beta;
    int m, nr, slabs;
    // two-segment operands (unbatched): rows kk >= k1 of R come from R2 (row kk - k1, R's
    // strides), rows i >= m1 of Y (and Y0) are those of Y2 (row i - m1, Y's strides)
    const double* R2;
    double* Y2;
    int k1, m1;
};

__device__ __forceinline__ double r_at(const SkinnyIO& io, int kk, int c)
{
    return kk < io.k1 ? io.R[(long long)kk * io.rsk + (long long)c * io.rsj]
                      : io.R2[(long long)(kk - io.k1) * io.rsk + (long long)c * io.rsj];
}
__device__ __forceinline__ double y0_at(const SkinnyIO& io, int i, int c)
{
    return i < io.m1 ? io.Y0[(long long)i * io.y0si + (long long)c * io.y0sj]
                     : io.Y2[(long long)(i - io.m1) * io.ysi + (long long)c * io.ysj];
}
__device__ __forceinline__ double* y_at(const SkinnyIO& io, int i, int c)
{
    return i < io.m1 ? io.Y + (long long)i * io.ysi + (long long)c * io.ysj
                     : io.Y2 + (long long)(i - io.m1) * io.ysi + (long long)c * io.ysj;
}

// The split-k partials of the rows [r0, r0 + rows) of matrix b are in `work` (written by
// this workgroup for slab blockIdx.y): every workgroup stores its partials with device-
// coherent stores, waits for them, and counts itself in; the last of the row block's slabs
// to arrive sums the partials of all slabs in slab order, s = 0, 1, ... (the order of the
// separate reduction kernel, so the result is the same) and writes Y.  It then clears the
// counter for the next launch.
__device__ __forceinline__ double ld_sc1(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_sc1(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int NR>
__device__ __forceinline__ void fused_reduce(const SkinnyIO& io, const double* work_b, int r0, int rows)
{
    __shared__ int s_last;
    __builtin_amdgcn_s_waitcnt(0);   // this workgroup's partial stores are complete
    __syncthreads();
    int* const ctr = io.ctr + (long long)blockIdx.z * gridDim.x + blockIdx.x;
    if (threadIdx.x == 0)
        s_last = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == io.slabs - 1;
    __syncthreads();
    if (!s_last) return;
    const int m = io.m;
    // device-coherent plain loads (the sc1 cache policy), eight slabs in flight at a time
    const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc((void*)work_b, (short)0, 0x7fffffff, 0x00020000);
    auto ld = [&](int sl, int c, int i) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rw, (int)(((sl * NR + c) * (long long)m + i) * 8), 0, 16));
    };
    for (int e = threadIdx.x; e < rows * io.nr; e += blockDim.x) {
        const int i = r0 + e % rows, c = e / rows;
        double v = 0.0;
        int sl = 0;
        for (; sl + 8 <= io.slabs; sl += 8) {
            double t[8];
#pragma unroll
            for (int q = 0; q < 8; q++) t[q] = ld(sl + q, c, i);
#pragma unroll
            for (int q = 0; q < 8; q++) v += t[q];
        }
        for (; sl < io.slabs; sl++) v += ld(sl, c, i);
        v *= io.alpha;
        if (io.beta != 0.0) v += io.beta * y0_at(io, i, c);
        *y_at(io, i, c) = v;
    }
    if (threadIdx.x == 0) __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// partial layout: work[((b * slabs + s) * NR + c) * m + i] for matrix b = blockIdx.z
// N4 (one or two right-hand sides): 64 rows of op(A) = A per workgroup, each row by four lanes (one per wave), lane q
// taking the q-th quarter of every KCH-wide k-chunk (column reads coalesced across the
// wave, KU columns in flight per lane); the four partial sums of a row are added in a
// fixed order through LDS
template <int NR>
__global__ __launch_bounds__(256) void skinny_n4_kernel(int m, int k, int nr, int kslab, int tri, double alpha,
                                                       const double* __restrict__ A, int lda, SkinnyIO io,
                                                       double beta, double* work, int direct)
{
    __shared__ double Rs[KCH][NR];
    __shared__ double red[4][NR][NROWS];
    const int tid = threadIdx.x, row = tid & (NROWS - 1), q = tid >> 6;
    const int i = blockIdx.x * NROWS + row;
    const int s = blockIdx.y;
    {
        const long long b = blockIdx.z;
        A += b * io.sA;
        io.R += b * io.sR;
        io.Y0 += b * io.sY0;
        io.Y += b * io.sY;
        work += b * (long long)gridDim.y * NR * m;
    }
    const int kb = s * kslab, ke = min(k, kb + kslab);
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
    const int rowmax = min(m, (int)(blockIdx.x * NROWS + NROWS)) - 1;
    int kend = ke;
    if (tri == 1) kend = min(kend, rowmax + 1);   // op(A) lower: A[i][kk] == 0 for kk > i
    int kbeg = kb;
    if (tri == 2) kbeg = max(kbeg, (int)(blockIdx.x * NROWS));   // op(A) upper
    constexpr int KQ = KCH / 4;   // columns of a chunk per lane
    for (int k0 = kbeg; k0 < kend; k0 += KCH) {
        const int kc = min(KCH, kend - k0);
        __syncthreads();
        for (int e = tid; e < KCH * NR; e += 256) {
            const int kk = e % KCH, c = e / KCH;
            Rs[kk][c] = (kk < kc && c < nr) ? r_at(io, k0 + kk, c) : 0.0;
        }
        __syncthreads();
        if (i < m) {
            const int q0 = q * KQ;
            const double* Ap = A + (long long)(k0 + q0) * lda + i;
            for (int kk = 0; kk < KQ; kk += KU) {
                double a[KU];
#pragma unroll
                for (int u = 0; u < KU; u++) a[u] = q0 + kk + u < kc ? Ap[(long long)(kk + u) * lda] : 0.0;
#pragma unroll
                for (int u = 0; u < KU; u++)
#pragma unroll
                    for (int c = 0; c < NR; c++) acc[c] = fma(a[u], Rs[q0 + kk + u][c], acc[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NR; c++) red[q][c][row] = acc[c];
    __syncthreads();
    if (q == 0 && i < m) {
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] = ((red[0][c][row] + red[1][c][row]) + red[2][c][row]) + red[3][c][row];
        if (direct) {
#pragma unroll
            for (int c = 0; c < NR; c++) {
                if (c >= nr) break;
                double v = alpha * acc[c];
                if (beta != 0.0) v += beta * y0_at(io, i, c);
                *y_at(io, i, c) = v;
            }
        } else if (!io.ctr) {
#pragma unroll
            for (int c = 0; c < NR; c++) work[((long long)s * NR + c) * m + i] = acc[c];
        } else {
#pragma unroll
            for (int c = 0; c < NR; c++) st_sc1(work + ((long long)s * NR + c) * m + i, acc[c]);
        }
    }
    if (!direct && io.ctr) fused_reduce<NR>(io, work, blockIdx.x * NROWS, min(NROWS, m - (int)blockIdx.x * NROWS));
}

// T for one or two right-hand sides: one wave per output row j of op(A) = A^T (column j of A
// read contiguously), lanes over k with KT1 loads of A in flight, each lane reading its own R
// entries, wave-level reduction
template <int NR>
__global__ __launch_bounds__(256) void skinny_t1_kernel(int m, int k, int nr, int kslab, int tri, double alpha,
                                                       const double* __restrict__ A, int lda, SkinnyIO io,
                                                       double beta, double* work, int direct)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.x * 4 + wave;   // output row (column of A)
    const int s = blockIdx.y;
    {
        const long long b = blockIdx.z;
        A += b * io.sA;
        io.R += b * io.sR;
        io.Y0 += b * io.sY0;
        io.Y += b * io.sY;
        work += b * (long long)gridDim.y * NR * m;
    }
    const int kb = s * kslab, ke = min(k, kb + kslab);
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) acc[c] = 0.0;
    int kbeg = kb, kend = ke;
    const int jmin = blockIdx.x * 4, jmax = min(m, jmin + 4) - 1;
    if (tri == 2) kbeg = max(kbeg, (jmin / KCT) * KCT);   // op(A) upper: A^T[j][kk] == 0 for kk < j
    if (tri == 1) kend = min(kend, jmax + 1);             // op(A) lower
    const double* Aj = A + (long long)min(j, m - 1) * lda;
    if (j < m) {
        for (int k0 = kbeg; k0 < kend; k0 += 64 * KT1) {
            double a[KT1], r[KT1][NR];
#pragma unroll
            for (int u = 0; u < KT1; u++) {
                const int kk = k0 + lane + 64 * u;
                const bool in = kk < kend;
                a[u] = in ? Aj[kk] : 0.0;
#pragma unroll
                for (int c = 0; c < NR; c++) r[u][c] = in && c < nr ? r_at(io, kk, c) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < KT1; u++)
#pragma unroll
                for (int c = 0; c < NR; c++) acc[c] = fma(a[u], r[u][c], acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < NR; c++) {
        double v = acc[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        acc[c] = v;
    }
    if (j < m && lane == 0) {
        if (direct) {
#pragma unroll
            for (int c = 0; c < NR; c++) {
                if (c >= nr) break;
                double v = alpha * acc[c];
                if (beta != 0.0) v += beta * y0_at(io, j, c);
                *y_at(io, j, c) = v;
            }
        } else if (!io.ctr) {
#pragma unroll
            for (int c = 0; c < NR; c++) work[((long long)s * NR + c) * m + j] = acc[c];
        } else {
#pragma unroll
            for (int c = 0; c < NR; c++) st_sc1(work + ((long long)s * NR + c) * m + j, acc[c]);
        }
    }
    if (!direct && io.ctr) fused_reduce<NR>(io, work, blockIdx.x * 4, min(4, m - (int)blockIdx.x * 4));
}


// Three to sixteen right-hand sides on FP64 MFMA (v_mfma_f64_16x16x4: 16 rows of op(A) x 4 k
// times 4 k x 16 right-hand sides per instruction).  A workgroup takes 64 rows of op(A) (16
// per wave) over a k-slab.  N: the A fragment (16 consecutive rows at 4 consecutive k) is
// read straight from the column-major A; T: op(A) row j is column j of A, the fragment 16
// columns at 4 consecutive k, each column's 128-byte line used by four successive steps.
// R is read per lane (rows of R stay in cache: every workgroup reads them).  Partials and
// the epilogue as in the other skinny kernels.
constexpr int MKU = 8;   // k-steps of 4 with their loads in flight
template <bool TRANS>
__global__ __launch_bounds__(256) void skinny_mfma_kernel(int m, int k, int nr, int kslab, int tri, double alpha,
                                                          const double* __restrict__ A, int lda, SkinnyIO io,
                                                          double beta, double* work, int direct)
{
    constexpr int NR = MAXR;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int s = blockIdx.y;
    {
        const long long b = blockIdx.z;
        A += b * io.sA;
        io.R += b * io.sR;
        io.Y0 += b * io.sY0;
        io.Y += b * io.sY;
        work += b * (long long)gridDim.y * NR * m;
    }
    const int r0 = blockIdx.x * 64 + 16 * wave;   // this wave's rows of op(A)
    const int kb = s * kslab, ke = min(k, kb + kslab);
    int kbeg = kb, kend = ke;
    const int bmin = blockIdx.x * 64, bmax = min(m, bmin + 64) - 1;
    if (tri == 1) kend = min(kend, bmax + 1);             // op(A) lower
    if (tri == 2) kbeg = max(kbeg, bmin & ~3);            // op(A) upper
    const int ra = min(r0 + l16, m - 1);                  // A-fragment row (clamped; rows >= m are not stored)
    const double* Ap = TRANS ? A + (long long)ra * lda : A + ra;
    const long long kst = TRANS ? 1 : lda;                // step of A along k
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = kbeg; k0 < kend; k0 += 4 * MKU) {
        double a[MKU], r[MKU];
#pragma unroll
        for (int u = 0; u < MKU; u++) {
            const int kk = k0 + 4 * u + lq;
            const bool in = kk < kend;
            a[u] = in ? Ap[(long long)kk * kst] : 0.0;
            r[u] = in && l16 < nr ? r_at(io, kk, l16) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < MKU; u++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], r[u], acc, 0, 0, 0);
    }
    // lane (l16, lq) item t: row r0 + 4 t + lq of op(A), right-hand side l16
    const int c = l16;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int i = r0 + 4 * t + lq;
        if (i >= m || c >= nr) continue;
        if (direct) {
            double v = alpha * acc[t];
            if (beta != 0.0) v += beta * y0_at(io, i, c);
            *y_at(io, i, c) = v;
        } else if (!io.ctr) {
            work[((long long)s * NR + c) * m + i] = acc[t];
        } else {
            st_sc1(work + ((long long)s * NR + c) * m + i, acc[t]);
        }
    }
    if (!direct && io.ctr) fused_reduce<NR>(io, work, blockIdx.x * 64, min(64, m - (int)blockIdx.x * 64));
}

__global__ void skinny_reduce_kernel(int m, int nr, int wstride, int slabs, int count, double alpha,
                                     const double* work, double beta, SkinnyIO io)
{
    const long long per = (long long)m * nr, total = per * count;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long b = e / per, r = e % per;
        const int i = (int)(r % m), c = (int)(r / m);
        const double* w = work + b * (long long)slabs * wstride * m;
        double v = 0.0;
        for (int s = 0; s < slabs; s++) v += w[((long long)s * wstride + c) * m + i];
        v *= alpha;
        if (i >= io.m1) {   // second output segment (unbatched)
            if (beta != 0.0) v += beta * y0_at(io, i, c);
            *y_at(io, i, c) = v;
            continue;
        }
        if (beta != 0.0) v += beta * io.Y0[b * io.sY0 + (long long)i * io.y0si + (long long)c * io.y0sj];
        io.Y[b * io.sY + (long long)i * io.ysi + (long long)c * io.ysj] = v;
    }
}

template <int NR>
static int launch_skinny(Op ta, int tri, int m, int k, int nr, double alpha, const double* A, int lda,
                         const SkinnyIO& io_in, double beta, double* work, size_t work_elems, int count, hipStream_t s,
                         int slab_cap)
{
    constexpr bool mf = NR == MAXR;   // the MFMA kernel (for any nr: 16 columns of partials)
    SkinnyPlan p = plan_for(ta, m, k, mf ? MAXR : nr);
    if (p.slabs > slab_cap) {   // fewer, longer slabs (a launch over many blocks has enough workgroups)
        const int chunk = mf || ta == Op::T ? KCT : KCH;
        p.kslab = cdiv(cdiv(k, slab_cap), chunk) * chunk;
        p.slabs = cdiv(k, p.kslab);
    }
    const int direct = p.slabs == 1;
    if (!direct && (size_t)count * p.slabs * NR * m > work_elems) return (int)hipErrorInvalidValue;
    SkinnyIO io = io_in;
    {   // reduce the partials in the product launch when the counters cover its row blocks
        // (one workgroup per row block sums its slabs: only where that is a few partials per
        // thread -- one or two right-hand sides -- else the reduction launch spreads them out)
        const int rpw = mf ? 64 : ta == Op::T ? 4 : NROWS;
        if (direct || !io.ctr || (long long)cdiv(m, rpw) * count > SKINNY_CTRS || (long long)rpw * nr > 256)
            io.ctr = nullptr;
        io.alpha = alpha;
        io.beta = beta;
        io.m = m;
        io.nr = nr;
        io.slabs = p.slabs;
    }
    if constexpr (mf) {
        dim3 grid(cdiv(m, 64), p.slabs, count);
        if (ta == Op::N)
            hipLaunchKernelGGL((skinny_mfma_kernel<false>), grid, dim3(256), 0, s, m, k, nr, p.kslab, tri, alpha, A,
                               lda, io, beta, work, direct);
        else
            hipLaunchKernelGGL((skinny_mfma_kernel<true>), grid, dim3(256), 0, s, m, k, nr, p.kslab, tri, alpha, A,
                               lda, io, beta, work, direct);
    } else if (ta == Op::N) {
        dim3 grid(cdiv(m, NROWS), p.slabs, count);
        hipLaunchKernelGGL((skinny_n4_kernel<NR>), grid, dim3(256), 0, s, m, k, nr, p.kslab, tri, alpha, A, lda, io,
                           beta, work, direct);
    } else {
        dim3 grid(cdiv(m, 4), p.slabs, count);
        hipLaunchKernelGGL((skinny_t1_kernel<NR>), grid, dim3(256), 0, s, m, k, nr, p.kslab, tri, alpha, A, lda, io,
                           beta, work, direct);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    if (!direct && !io.ctr) {
        long long total = (long long)m * nr * count;
        int grid = (int)std::min((total + 255) / 256, (long long)4096);
        hipLaunchKernelGGL(skinny_reduce_kernel, dim3(grid), dim3(256), 0, s, m, nr, NR, p.slabs, count, alpha, work,
                           beta, io);
        e = hipGetLastError();
    }
    return (int)e;
}

int skinny_f64s(Op ta, int tri, int m, int k, int nr, double alpha, const double* A, int lda, const double* R,
                long long rsk, long long rsj, double beta, const double* Y0, long long y0si, long long y0sj,
                double* Y, long long ysi, long long ysj, double* work, size_t work_elems, hipStream_t s, int count,
                long long sA, long long sR, long long sY0, long long sY, int* ctr, const double* R2, int k1,
                double* Y2, int m1, int slab_cap, bool wide)
{
    if (m <= 0 || nr <= 0 || count <= 0) return 0;
    if (nr > MAXR) return (int)hipErrorInvalidValue;
    const SkinnyIO io{R, rsk, rsj, Y0, y0si, y0sj, Y, ysi, ysj, sA, sR, sY0, sY, ctr, 0.0, 0.0, 0, 0, 0,
                      R2 ? R2 : R, Y2 ? Y2 : Y, R2 ? k1 : k, Y2 ? m1 : m};
    if ((R2 || Y2) && count != 1) return (int)hipErrorInvalidValue;
    if (k <= 0) {   // Y = beta*Y0
        if (count == 1 && ysi == 1 && y0si == 1) return copy2d(Y0, (int)y0sj, Y, (int)ysj, m, nr, beta, s);
        return (int)hipErrorInvalidValue;
    }
    slab_cap = std::max(1, slab_cap);
    if (wide) return launch_skinny<MAXR>(ta, tri, m, k, nr, alpha, A, lda, io, beta, work, work_elems, count, s,
                                         slab_cap);
    if (nr == 1) return launch_skinny<1>(ta, tri, m, k, nr, alpha, A, lda, io, beta, work, work_elems, count, s, slab_cap);
    if (nr <= 2) return launch_skinny<2>(ta, tri, m, k, nr, alpha, A, lda, io, beta, work, work_elems, count, s, slab_cap);
    return launch_skinny<MAXR>(ta, tri, m, k, nr, alpha, A, lda, io, beta, work, work_elems, count, s,
                               slab_cap);   // MFMA kernel
}

int skinny_f64(Op ta, int tri, int m, int k, int nr, double alpha, const double* A, int lda,
               const double* R, int ldr, double beta, const double* Y0, int ldy0, double* Y, int ldy,
               double* work, hipStream_t s, size_t work_elems, int* ctr, bool wide)
{
    return skinny_f64s(ta, tri, m, k, nr, alpha, A, lda, R, 1, ldr, beta, Y0, 1, ldy0, Y, 1, ldy, work,
                       work_elems ? work_elems : skinny_work_elems(m, k, nr), s, 1, 0, 0, 0, 0, ctr, nullptr, 0,
                       nullptr, 0, 64, wide);
}

size_t skinny_call_work(Op ta, int m, int k, int nr)
{
    if (m <= 0 || k <= 0 || nr <= 0) return 0;
    const SkinnyPlan p = plan_for(ta, m, k, nr);
    if (p.slabs == 1) return 0;
    const int NR = nr == 1 ? 1 : nr <= 2 ? 2 : MAXR;   // launch_skinny's buckets
    return (size_t)p.slabs * NR * (size_t)m;
}

}  // namespace rgf
