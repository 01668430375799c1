// Split-BF16 GEMM on CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16), the product
// kernel of the mixed-precision path.
//
// FP64 operands are split while they are staged into LDS: x -> f = (float)x,
// hi = bf16(f), lo = bf16(f - hi), so x = hi + lo to about 2^-16 relative.  Each
// product is hi*hi + hi*lo + lo*hi (three MFMAs; lo*lo is below the kept
// precision) accumulated in FP32 by the matrix core, and the epilogue applies
// alpha / beta in FP64, so C keeps FP64 storage and FP64 accumulation across
// calls (e.g. the Schur complement accumulated into the FP64 diagonal block).
//
// Same descriptor interface and tile order as the FP64 kernel.  BMxBN tiles of
// 64x64 (single-matrix shapes) or 128x128 (batched shapes), 4 waves in 2x2,
// K in steps of 32.  LDS holds op(A) as [row][k] and op(B) as [col][k] (k
// contiguous) for hi and lo, so a lane fetches its 8 k-values with one 16-byte
// read.  The k-assignment of the 8 values inside the MFMA is the same for both
// operands, so any lane-to-k order of the instruction gives the full sum.
#include "rgf_internal.hpp"

#include <algorithm>
#include <vector>

namespace rgf {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int MBK = 32;     // k per step
constexpr int KPAD = 8;     // bf16 padding per LDS row (16 bytes)
constexpr int LROW = MBK + KPAD;

__device__ __forceinline__ void split_bf16(double x, __bf16& hi, __bf16& lo)
{
    const float f = (float)x;
    hi = (__bf16)f;
    lo = (__bf16)(f - (float)hi);
}

__device__ __forceinline__ void select_desc_m(const GemmArgs& args, int z, GemmDesc& D)
{
    D = args.d[0];
#pragma unroll
    for (int q = 1; q < GEMM_MAXD; q++)
        if (q < args.nd && z >= args.d[q].zbase) D = args.d[q];
}

template <bool TA, bool TB, int BT>
__global__ __launch_bounds__(256, (BT == 128 ? 2 : 4)) void gemm_bf16x3_kernel(GemmArgs args)
{
    constexpr int WT = BT / 2;        // wave tile edge
    constexpr int NF = WT / 16;       // 16x16 MFMA blocks per wave edge
    constexpr int EPT = BT * MBK / 256;   // elements per thread per operand per k-step
    __shared__ __attribute__((aligned(16))) __bf16 Ah[BT][LROW];
    __shared__ __attribute__((aligned(16))) __bf16 Al[BT][LROW];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[BT][LROW];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[BT][LROW];

    const int z = blockIdx.y;
    GemmDesc D;
    select_desc_m(args, z, D);
    const int flags = D.flags;
    int g = blockIdx.x;
    {
        const int G = gridDim.x, q = G >> 3, rem = G & 7, xcd = g & 7, loc = g >> 3;
        g = xcd < rem ? xcd * (q + 1) + loc : rem * (q + 1) + (xcd - rem) * q + loc;
    }
    int ti, tj;
    if (D.tri_tiles) {
        if (g >= D.tiles_m * (D.tiles_m + 1) / 2) return;
        ti = (int)((sqrt(8.0 * g + 1.0) - 1.0) * 0.5);
        while (ti * (ti + 1) / 2 > g) ti--;
        while ((ti + 1) * (ti + 2) / 2 <= g) ti++;
        tj = g - ti * (ti + 1) / 2;
    } else {
        if (g >= D.tiles_m * D.tiles_n) return;
        const int band = g / (8 * D.tiles_n);
        const int ti0 = band * 8, gsz = min(D.tiles_m - ti0, 8);
        const int in = g - band * 8 * D.tiles_n;
        ti = ti0 + in % gsz;
        tj = in / gsz;
    }
    const int i0 = ti * BT, j0 = tj * BT;
    if ((flags & G_CLOWER) && j0 > i0 + BT - 1) return;

    const int bz = z - D.zbase;
    const long long b1 = bz % D.count, b2 = bz / D.count;
    const double* A = D.A + b1 * D.sA + b2 * D.sA2;
    const double* B = D.B + b1 * D.sB + b2 * D.sB2;
    double* C = D.C + b1 * D.sC + b2 * D.sC2;
    const double* Cin = D.Cin + b1 * D.sCin + b2 * D.sCin2;
    const int M = D.m, N = D.n, K = D.k;
    const int lda = D.lda, ldb = D.ldb;

    int klo = 0, khi = K;
    if (flags & G_ALOWER) khi = min(khi, i0 + BT);
    if (flags & G_AUPPER) klo = max(klo, i0);
    if (flags & G_BLOWER) klo = max(klo, j0);
    if (flags & G_BUPPER) khi = min(khi, j0 + BT);
    klo = (klo / MBK) * MBK;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lq = lane >> 4;

    v4f acc[NF][NF];
#pragma unroll
    for (int a = 0; a < NF; a++)
#pragma unroll
        for (int b = 0; b < NF; b++) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

    double ra[EPT], rb[EPT];
    // Row-contiguous operand (op(X)[r][k] = X[r + k*ld]): thread owns one row r and
    // EPT consecutive k.  k-contiguous operand (op(X)[r][k] = X[k + r*ld]): thread owns
    // one k and EPT rows spaced 256/MBK apart.
    auto load_op = [&](const double* X, int ld, int r0, int R, int k0, bool kcontig, double* reg) {
        if (!kcontig) {
            const int r = tid % BT, kb = (tid / BT) * EPT;
            const int gr = r0 + r;
#pragma unroll
            for (int e = 0; e < EPT; e++) {
                const int gk = k0 + kb + e;
                reg[e] = (gr < R && gk < khi) ? X[(long long)gk * ld + gr] : 0.0;
            }
        } else {
            const int kk = tid % MBK, rb0 = tid / MBK;
            const int gk = k0 + kk;
#pragma unroll
            for (int e = 0; e < EPT; e++) {
                const int gr = r0 + rb0 + e * (256 / MBK);
                reg[e] = (gr < R && gk < khi) ? X[(long long)gr * ld + gk] : 0.0;
            }
        }
    };
    auto store_op = [&](__bf16 (*H)[LROW], __bf16 (*Lo)[LROW], bool kcontig, const double* reg) {
        if (!kcontig) {
            const int r = tid % BT, kb = (tid / BT) * EPT;
#pragma unroll
            for (int e = 0; e < EPT; e += 8) {
                bf16x8 vh, vl;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    __bf16 h, l;
                    split_bf16(reg[e + q], h, l);
                    vh[q] = h;
                    vl[q] = l;
                }
                *(bf16x8*)&H[r][kb + e] = vh;
                *(bf16x8*)&Lo[r][kb + e] = vl;
            }
        } else {
            const int kk = tid % MBK, rb0 = tid / MBK;
#pragma unroll
            for (int e = 0; e < EPT; e++) {
                __bf16 h, l;
                split_bf16(reg[e], h, l);
                H[rb0 + e * (256 / MBK)][kk] = h;
                Lo[rb0 + e * (256 / MBK)][kk] = l;
            }
        }
    };

    // op(A) row-contiguous iff !TA; op(B)^T rows are columns of op(B): contiguous iff TB
    if (klo < khi) {
        load_op(A, lda, i0, M, klo, TA, ra);
        load_op(B, ldb, j0, N, klo, !TB, rb);
    }
    for (int k0 = klo; k0 < khi; k0 += MBK) {
        store_op(Ah, Al, TA, ra);
        store_op(Bh, Bl, !TB, rb);
        __syncthreads();
        if (k0 + MBK < khi) {
            load_op(A, lda, i0, M, k0 + MBK, TA, ra);
            load_op(B, ldb, j0, N, k0 + MBK, !TB, rb);
        }
        bf16x8 ah[NF], al[NF], bh[NF], bl[NF];
#pragma unroll
        for (int a = 0; a < NF; a++) {
            const int r = wm * WT + a * 16 + l16;
            ah[a] = *(const bf16x8*)&Ah[r][8 * lq];
            al[a] = *(const bf16x8*)&Al[r][8 * lq];
        }
#pragma unroll
        for (int b = 0; b < NF; b++) {
            const int c = wn * WT + b * 16 + l16;
            bh[b] = *(const bf16x8*)&Bh[c][8 * lq];
            bl[b] = *(const bf16x8*)&Bl[c][8 * lq];
        }
#pragma unroll
        for (int a = 0; a < NF; a++)
#pragma unroll
            for (int b = 0; b < NF; b++) {
                v4f t = acc[a][b];
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[b], ah[a], t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[b], al[a], t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[b], ah[a], t, 0, 0, 0);
                acc[a][b] = t;
            }
        __syncthreads();
    }

    // epilogue: the MFMA was fed (op(B) fragment, op(A) fragment), so item r of lane
    // (l16 + 16*lq) is D[4*lq + r][l16] = C[i = l16][j = 4*lq + r] of each 16x16 block
    const double alpha = D.alpha, beta = D.beta;
    const int ldc = D.ldc, ldcin = D.ldcin;
#pragma unroll
    for (int a = 0; a < NF; a++) {
        const int gi = i0 + wm * WT + a * 16 + l16;
#pragma unroll
        for (int b = 0; b < NF; b++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gj = j0 + wn * WT + b * 16 + 4 * lq + r;
                if (gi < M && gj < N && (!(flags & G_CLOWER) || gi >= gj)) {
                    double v = alpha * (double)acc[a][b][r];
                    if (beta != 0.0) v += beta * Cin[(long long)gj * ldcin + gi];
                    C[(long long)gj * ldc + gi] = v;
                }
            }
        }
    }
}

template <int BT>
static int launch_mixed(Op ta, Op tb, GemmArgs& args, dim3 grid, hipStream_t s)
{
    if (ta == Op::N && tb == Op::N) hipLaunchKernelGGL((gemm_bf16x3_kernel<false, false, BT>), grid, dim3(256), 0, s, args);
    else if (ta == Op::N && tb == Op::T) hipLaunchKernelGGL((gemm_bf16x3_kernel<false, true, BT>), grid, dim3(256), 0, s, args);
    else if (ta == Op::T && tb == Op::N) hipLaunchKernelGGL((gemm_bf16x3_kernel<true, false, BT>), grid, dim3(256), 0, s, args);
    else hipLaunchKernelGGL((gemm_bf16x3_kernel<true, true, BT>), grid, dim3(256), 0, s, args);
    return (int)hipGetLastError();
}

int gemm_bf16x3(Op ta, Op tb, GemmDesc* descs, int ndesc, hipStream_t s)
{
    if (ndesc <= 0) return 0;
    if (ndesc > GEMM_MAXD) return (int)hipErrorInvalidValue;
    // tile edge: 128 when the launch has enough 128-tiles to fill the device, else 64
    long long work128 = 0;
    for (int q = 0; q < ndesc; q++) {
        const GemmDesc& d = descs[q];
        if (d.m <= 0 || d.n <= 0 || d.count <= 0) continue;
        const long long t = (long long)((d.m + 127) / 128) * ((d.n + 127) / 128);
        work128 += t * d.count * std::max(d.count2, 1);
    }
    const int BT = work128 >= 768 ? 128 : 64;
    GemmArgs args{};
    int z = 0, maxtiles = 0, nd = 0;
    for (int q = 0; q < ndesc; q++) {
        GemmDesc d = descs[q];
        if (d.count2 < 1) d.count2 = 1;
        if (d.m <= 0 || d.n <= 0 || d.count <= 0) continue;
        d.tiles_m = (d.m + BT - 1) / BT;
        d.tiles_n = (d.n + BT - 1) / BT;
        d.tri_tiles = (d.flags & G_CLOWER) && d.tiles_m == d.tiles_n;
        d.zbase = z;
        z += d.count * d.count2;
        const int nt = d.tri_tiles ? d.tiles_m * (d.tiles_m + 1) / 2 : d.tiles_m * d.tiles_n;
        if (nt > maxtiles) maxtiles = nt;
        args.d[nd++] = d;
    }
    if (nd == 0) return 0;
    args.nd = nd;
    dim3 grid(maxtiles, z);
    void* tok = timing_begin(4 + (ta == Op::T ? 2 : 0) + (tb == Op::T ? 1 : 0), args.d, nd, s);
    int rc = BT == 128 ? launch_mixed<128>(ta, tb, args, grid, s) : launch_mixed<64>(ta, tb, args, grid, s);
    timing_end(tok, s);
    return rc;
}

}  // namespace rgf
