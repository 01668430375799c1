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
    if (!D.deal_rr) {
        const int G = gridDim.x, q = G >> 3, rem = G & 7, xcd = g & 7, loc = g >> 3;
        g = xcd < rem ? xcd * (q + 1) + loc : rem * (q + 1) + (xcd - rem) * q + loc;
    }
    int ti, tj;
    if (D.tri_tiles) {   // lower-triangular C: only the tiles on or below the diagonal
        if (!tri_tile(g, D.tiles_m, ti, tj)) return;
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
        set_tile_order(d, (d.flags & G_CLOWER) && d.tiles_m == d.tiles_n);
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

// ======================================================================== pre-split operands
// The kernel above converts every operand element once per output tile that reads it.
// Here each operand is split once into bf16 hi/lo copies of op(A) as [row][k] and
// op(B)^T as [col][k] (k padded with zeros to ldk = roundup(K, 64)), stored in 16 x 32
// blocks of one kilobyte (split_index), and the product kernels stream them into LDS
// one contiguous block per wave-wide 16-byte load.
namespace rgf {

constexpr int SPLIT_PITCH = 64;

// Layout of a split copy (one plane: hi, or lo): rows padded to a multiple of 128, k to the
// pitch ldk (a multiple of 64), stored as 16-row x 32-k blocks of 1 KB -- block (g, kc) of
// rows 16g.. and k 32kc.. at element ((g * ldk / 32) + kc) * 512 -- and inside a block as 64
// 16-byte chunks: chunk 4 (r % 16) + ((k / 8) % 4 XOR (r / 4) % 4) holds k .. k + 7 (k a
// multiple of 8) of row r.  One wave's 16-byte loads of a block are then one contiguous
// kilobyte, and a block copied chunk for chunk into LDS is a 16-row x 32-k tile whose rows
// (64 bytes) have their chunks XOR-swizzled by (row / 4) % 4: the sixteen rows of one chunk
// column fall in sixteen distinct 16-byte bank groups (split_index, rgf_internal.hpp).
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// dst[e] plane at split_index(r, k) = split(op(X)[r][k]) for r < R, k < ldk (0 beyond
// K); hi at dst, lo at dst + lo_off.  op(X)[r][k] = X[r + k*ld] (kcontig = false) or X[k + r*ld].
// grid: (ldk/64, ceil(R/32), entries).  A 32 x 64 tile is read coalesced along its
// contiguous edge into LDS (all eight loads of a lane in flight: out-of-range lanes read
// element 0 and keep zero); each lane then converts 4 consecutive k of a row and stores
// them as 8-byte hi and lo groups.
__global__ __launch_bounds__(256) void split_bf16_kernel(const double* X, int ld, int R, int K, bool kcontig,
                                                         __bf16* dst, int ldk, long long lo_off, int count,
                                                         long long sX1, long long sX2, long long sD)
{
    __shared__ double t[32][65];
    const int e = blockIdx.z;
    const long long b1 = e % count, b2 = e / count;
    X += b1 * sX1 + b2 * sX2;
    dst += (long long)e * sD;
    const int k0 = blockIdx.x * 64, r0 = blockIdx.y * 32;
    const int tid = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int idx = q * 256 + tid;
        const int r = kcontig ? idx >> 6 : idx & 31, k = kcontig ? idx & 63 : idx >> 5;
        const int gr = r0 + r, gk = k0 + k;
        const bool ok = gr < R && gk < K;
        const long long at = kcontig ? (long long)gr * ld + gk : (long long)gk * ld + gr;
        const double v = X[ok ? at : 0];
        t[r][k] = ok ? v : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int g = q * 256 + tid;
        const int r = g >> 4, kq = (g & 15) * 4;
        const int gr = r0 + r;
        if (gr < R) {
            bf16x4 vh, vl;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                __bf16 h, l;
                split_bf16(t[r][kq + j], h, l);
                vh[j] = h;
                vl[j] = l;
            }
            const long long o = split_index(gr, k0 + kq, ldk >> 5);   // 4 consecutive k of one chunk
            *(bf16x4*)(dst + o) = vh;
            *(bf16x4*)(dst + lo_off + o) = vl;
        }
    }
}

// Operands and destinations of one descriptor of a pre-split launch: op(A) rows and op(B)^T
// rows (SplitView); cs, cr (p null: none) receive the stored C from the epilogue as the split
// copies split_operand would write for C (same values, so they replace that pass): cs as
// op(C)^T (row j = column j of C, k = C's rows), for C as op(B) (N) or op(A) (T) of a later
// product; cr as op(C) (row i = row i of C, k = C's columns), for op(A) (N) or op(B) (T).
// Rows of a copy beyond C's stay unwritten; with pad the k beyond C up to the pitch is zero.
struct PresplitOperands {
    SplitView a, b, cs, cr;
};
constexpr int PRESPLIT_MAXD = 2;
struct PresplitArgs {
    PresplitOperands op[PRESPLIT_MAXD];
};
// descriptor of launch entry z and its operands (uniform selects, no indexed kernel argument)
__device__ __forceinline__ void select_desc_ops(const GemmArgs& args, const PresplitArgs& pa, int z, GemmDesc& D,
                                                PresplitOperands& O)
{
    D = args.d[0];
    O = pa.op[0];
#pragma unroll
    for (int t = 1; t < PRESPLIT_MAXD; t++)
        if (t < args.nd && z >= args.d[t].zbase) {
            D = args.d[t];
            O = pa.op[t];
        }
}

// epilogue of a product tile: C entries (gi, gj .. gj + 3) (gj a multiple of 4) of entry
// (b1, b2) with values v (0 where C is not stored) into the split copies; the op(C) copy
// takes them as one 8-byte group of each plane (four consecutive k of one chunk)
__device__ __forceinline__ void copy_out(const PresplitOperands& O, long long b1, long long b2, int gi, int gj,
                                         int M, int N, const double* v)
{
    const SplitView& cs = O.cs;
    if (cs.p && gi < (cs.pad ? cs.ldk : M)) {   // k = gi (beyond M: zero, the k padding)
        __bf16* const e = cs.p + b1 * cs.s1 + b2 * cs.s2;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (gj + r < N) {
                __bf16 h, l;
                split_bf16(v[r], h, l);
                const long long o = split_index(gj + r, gi, cs.ldk >> 5);
                e[o] = h;
                e[cs.lo + o] = l;
            }
        }
    }
    const SplitView& cr = O.cr;
    if (cr.p && gi < M && gj < (cr.pad ? cr.ldk : N)) {   // k = gj; the pitch and N are multiples of 4
        bf16x4 h4, l4;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            __bf16 h, l;
            split_bf16(v[r], h, l);   // columns beyond N: zero (the k padding of the copy)
            h4[r] = h;
            l4[r] = l;
        }
        __bf16* const e = cr.p + b1 * cr.s1 + b2 * cr.s2;
        const long long o = split_index(gi, gj, cr.ldk >> 5);
        *(bf16x4*)(e + o) = h4;
        *(bf16x4*)(e + cr.lo + o) = l4;
    }
}

// KB: k per LDS stage (32 or 64; the split copies are padded to a multiple of 64).
// SW: LDS rows unpadded with the 16-byte chunk index XOR-swizzled by the row
// (chunk ^ (row / (16 / CPR)) mod CPR): any 16 lanes that store 16 chunks of 16 / CPR
// consecutive rows, or read one chunk of 16 consecutive rows, touch 16 distinct
// 16-byte bank groups.  Without SW rows are padded by 16 bytes.
// co: the split copies of the stored C the epilogue writes (CopyOut).
template <int BT, int KB = 32, bool SW = false>
__global__ __launch_bounds__(256, (BT == 128 ? 2 : 4)) void gemm_bf16x3s_kernel(GemmArgs args, PresplitArgs pa)
{
    constexpr int WT = BT / 2;
    constexpr int NF = WT / 16;
    constexpr int CH = BT * KB / 8 / 256;    // 16-byte chunks per thread per array per k-step
    constexpr int KR = SW ? KB : KB + KPAD;  // LDS row (bf16)
    constexpr int CPR = KB / 8;              // 16-byte chunks per row
    constexpr int RPG = 16 / CPR;            // rows sharing one swizzle value
    __shared__ __attribute__((aligned(16))) __bf16 Ah[BT][KR];
    __shared__ __attribute__((aligned(16))) __bf16 Al[BT][KR];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[BT][KR];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[BT][KR];

    const int z = blockIdx.y;
    GemmDesc D;
    PresplitOperands O;
    select_desc_ops(args, pa, z, D, O);
    const long long loA = O.a.lo, loB = O.b.lo;
    const int ldka = O.a.ldk, ldkb = O.b.ldk;
    const int flags = D.flags;
    int g = blockIdx.x;
    if (!D.deal_rr) {
        const int G = gridDim.x, q = G >> 3, rem = G & 7, xcd = g & 7, loc = g >> 3;
        g = xcd < rem ? xcd * (q + 1) + loc : rem * (q + 1) + (xcd - rem) * q + loc;
    }
    int ti, tj;
    if (D.tri_tiles) {   // lower-triangular C: only the tiles on or below the diagonal
        if (!tri_tile(g, D.tiles_m, ti, tj)) return;
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

    const int bz = z - D.zbase;   // entry index inside this descriptor; split copies are per entry
    const long long b1 = bz % D.count, b2 = bz / D.count;
    const __bf16* Ahg = O.a.p + b1 * O.a.s1 + b2 * O.a.s2;   // this entry's operand copies
    const __bf16* Bhg = O.b.p + b1 * O.b.s1 + b2 * O.b.s2;
    double* C = D.C + b1 * D.sC + b2 * D.sC2;
    const double* Cin = D.Cin + b1 * D.sCin + b2 * D.sCin2;
    const int M = D.m, N = D.n, K = D.k;

    int klo = 0, khi = K;
    if (flags & G_ALOWER) khi = min(khi, i0 + BT);
    if (flags & G_AUPPER) klo = max(klo, i0);
    if (flags & G_BLOWER) klo = max(klo, j0);
    if (flags & G_BUPPER) khi = min(khi, j0 + BT);
    klo = (klo / KB) * KB;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lq = lane >> 4;

    v4f acc[NF][NF];
#pragma unroll
    for (int a = 0; a < NF; a++)
#pragma unroll
        for (int b = 0; b < NF; b++) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

    // Chunk c (0..BT*CPR-1) of a [BT][KB] stage: 1 KB block c / 64 of the stage (row group
    // g, k block kh), chunk p = c % 64 of that block -- row 16g + p / 4, logical k-chunk
    // 4kh + ((p % 4) XOR (p / 16) % 4); 64 lanes read one block, a contiguous kilobyte.
    bf16x8 rah[CH], ral[CH], rbh[CH], rbl[CH];
    constexpr int KBB = KB / 32;   // 32-k blocks per stage
    auto chunk_rk = [](int c, int& r, int& j) {
        const int b = c >> 6, p = c & 63, g = b / KBB, kh = b % KBB;
        r = 16 * g + (p >> 2);
        j = 4 * kh + ((p & 3) ^ ((p >> 4) & 3));
    };
    // bf16 offset in an LDS row of logical 16-byte chunk j of row r
    auto lcol = [](int r, int j) { return 8 * (SW ? (j ^ ((r / RPG) % CPR)) : j); };
    // Loads without masks: the split copies hold roundup(rows, 128) rows of a 64-aligned
    // pitch, so every row and k of a tile is inside its entry.  Rows beyond M (N) hold
    // unwritten values that reach only C entries that are never stored; k beyond K is
    // zero.  Lane offsets from the entry (bytes) are loop-invariant 32-bit values; a stage
    // at k0 adds k0 / 32 blocks.
    const int nkba = ldka >> 5, nkbb = ldkb >> 5;
    unsigned lofa[CH], lofb[CH];
#pragma unroll
    for (int q = 0; q < CH; q++) {
        const int c = tid + 256 * q, b = c >> 6, g = b / KBB, kh = b % KBB, p = c & 63;
        lofa[q] = 2u * (((unsigned)((i0 >> 4) + g) * (unsigned)nkba + kh) * 512u + 8u * p);
        lofb[q] = 2u * (((unsigned)((j0 >> 4) + g) * (unsigned)nkbb + kh) * 512u + 8u * p);
    }
    const char* const Ah0 = (const char*)Ahg;
    const char* const Al0 = (const char*)(Ahg + loA);
    const char* const Bh0 = (const char*)Bhg;
    const char* const Bl0 = (const char*)(Bhg + loB);
    auto load = [&](int k0) {
        const long long kb = 2LL * 512 * (k0 >> 5);
#pragma unroll
        for (int q = 0; q < CH; q++) {
            rah[q] = *(const bf16x8*)(Ah0 + kb + lofa[q]);
            ral[q] = *(const bf16x8*)(Al0 + kb + lofa[q]);
            rbh[q] = *(const bf16x8*)(Bh0 + kb + lofb[q]);
            rbl[q] = *(const bf16x8*)(Bl0 + kb + lofb[q]);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int q = 0; q < CH; q++) {
            int r, j;
            chunk_rk(tid + 256 * q, r, j);
            const int col = lcol(r, j);
            *(bf16x8*)&Ah[r][col] = rah[q];
            *(bf16x8*)&Al[r][col] = ral[q];
            *(bf16x8*)&Bh[r][col] = rbh[q];
            *(bf16x8*)&Bl[r][col] = rbl[q];
        }
    };
    // the split copies are zero beyond K up to their 64-aligned pitch, so a k-step may
    // read past khi (the values there are zero or belong to skipped triangles: both exact)
    if (klo < khi) load(klo);
    for (int k0 = klo; k0 < khi; k0 += KB) {
        store();
        __syncthreads();
        if (k0 + KB < khi) load(k0 + KB);
#pragma unroll
        for (int ks = 0; ks < KB / 32; ks++) {
            bf16x8 ah[NF], al[NF], bh[NF], bl[NF];
#pragma unroll
            for (int a = 0; a < NF; a++) {
                const int r = wm * WT + a * 16 + l16;
                ah[a] = *(const bf16x8*)&Ah[r][lcol(r, 4 * ks + lq)];
                al[a] = *(const bf16x8*)&Al[r][lcol(r, 4 * ks + lq)];
            }
#pragma unroll
            for (int b = 0; b < NF; b++) {
                const int c = wn * WT + b * 16 + l16;
                bh[b] = *(const bf16x8*)&Bh[c][lcol(c, 4 * ks + lq)];
                bl[b] = *(const bf16x8*)&Bl[c][lcol(c, 4 * ks + lq)];
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
        }
        __syncthreads();
    }

    const double alpha = D.alpha, beta = D.beta;
    const int ldc = D.ldc, ldcin = D.ldcin;
#pragma unroll
    for (int a = 0; a < NF; a++) {
        const int gi = i0 + wm * WT + a * 16 + l16;
#pragma unroll
        for (int b = 0; b < NF; b++) {
            const int gj0 = j0 + wn * WT + b * 16 + 4 * lq;
            double v[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gj = gj0 + r;
                v[r] = 0.0;
                if (gi < M && gj < N && (!(flags & G_CLOWER) || gi >= gj)) {
                    v[r] = alpha * (double)acc[a][b][r];
                    if (beta != 0.0) v[r] += beta * Cin[(long long)gj * ldcin + gi];
                    C[(long long)gj * ldc + gi] = v[r];
                }
            }
            copy_out(O, b1, b2, gi, gj0, M, N, v);
        }
    }
}

// Wide pre-split product for launches of many tiles (batches): 256 x 128 tiles, four waves
// of 128 x 64 (each op(A) fragment read from LDS feeds four MFMA columns, each op(B)
// fragment eight rows: 3/4 of the LDS bytes per MFMA of the 128 x 128 kernel), one
// workgroup per CU.  K advances in stages of 32 through three LDS buffers (48 KB each)
// filled by buffer loads straight into LDS: every wave-wide 16-byte load copies one
// kilobyte block of a split copy, whose chunk order is the LDS tile's (rows of 64 bytes,
// chunks XOR-swizzled by (row / 4) % 4).  Stage s + 3 is loaded while stage s is computed;
// the fragments of stage s + 1 are read from LDS into a second register set during the
// MFMAs of stage s.  Each output gets the MFMAs of the 128 x 128 kernel in the same order
// (k-blocks of 32 ascending, hi*lo, lo*hi, hi*hi per block), so both give the same bits.
// rowsA, rowsB: rows of the split copies (multiples of 128); tile rows beyond them read the
// last 16-row group (they feed only C entries that are never stored).  Lower-stored square
// products (tri_tiles) take, per 128-wide column, the tiles from its diagonal down, so each
// covers two 128 x 128 tiles on or below the diagonal.  An isolated full 2865^3 product,
// batch of 8: 1.21 ms against 1.50 ms on the 128 x 128 kernel, same bits.
typedef __attribute__((address_space(3))) void lds_void_t;

__global__ __launch_bounds__(256, 1) void gemm_bf16x3w_kernel(GemmArgs args, PresplitArgs pa)
{
    constexpr int BM = 256, BN = 128, KB = 32, NS = 3;
    constexpr int SA = BM * KB, SB = BN * KB, STAGE = 2 * SA + 2 * SB;   // bf16 per array / stage
    __shared__ __attribute__((aligned(16))) __bf16 lds[NS * STAGE];

    const int z = blockIdx.y;
    GemmDesc D;
    PresplitOperands O;
    select_desc_ops(args, pa, z, D, O);
    const long long loA = O.a.lo, loB = O.b.lo;
    const int ldka = O.a.ldk, ldkb = O.b.ldk, rowsA = O.a.rows, rowsB = O.b.rows;
    const int flags = D.flags;
    int g = blockIdx.x;
    if (!D.deal_rr) {
        const int G = gridDim.x, q = G >> 3, rem = G & 7, xcd = g & 7, loc = g >> 3;
        g = xcd < rem ? xcd * (q + 1) + loc : rem * (q + 1) + (xcd - rem) * q + loc;
    }
    int i0, j0;
    if (D.tri_tiles) {   // lower-stored square C: column tj's tiles start at its diagonal
        // (rows 128 tj + 256 q), so each covers two 128 x 128 tiles on or below it
        const int T = D.tiles_n;   // 128-wide tile columns (= 128-row tile rows)
        int tj = 0, rem = g;
        while (tj < T && rem >= (T - tj + 1) / 2) rem -= (T - tj++ + 1) / 2;
        if (tj >= T) return;
        i0 = BN * tj + BM * rem;
        j0 = BN * tj;
    } else {             // bands of 8 tile-rows, column-major inside a band
        if (g >= D.tiles_m * D.tiles_n) return;
        const int band = g / (8 * D.tiles_n);
        const int ti0 = band * 8, gsz = min(D.tiles_m - ti0, 8);
        const int in = g - band * 8 * D.tiles_n;
        i0 = (ti0 + in % gsz) * BM;
        j0 = (in / gsz) * BN;
        if ((flags & G_CLOWER) && j0 > i0 + BM - 1) return;
    }

    const int bz = z - D.zbase;
    const long long b1 = bz % D.count, b2 = bz / D.count;
    const __bf16* Ahg = O.a.p + b1 * O.a.s1 + b2 * O.a.s2;   // this entry's operand copies
    const __bf16* Bhg = O.b.p + b1 * O.b.s1 + b2 * O.b.s2;
    double* C = D.C + b1 * D.sC + b2 * D.sC2;
    const double* Cin = D.Cin + b1 * D.sCin + b2 * D.sCin2;
    const int M = D.m, N = D.n, K = D.k;

    int klo = 0, khi = K;
    if (flags & G_ALOWER) khi = min(khi, i0 + BM);
    if (flags & G_AUPPER) klo = max(klo, i0);
    if (flags & G_BLOWER) klo = max(klo, j0);
    if (flags & G_BUPPER) khi = min(khi, j0 + BN);
    klo = (klo / KB) * KB;
    const int nk = khi > klo ? (khi - klo + KB - 1) / KB : 0;

    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;

    // load q of wave w copies 16-row group 4q + w of the tile (k-block of the stage in soffset)
    int offa[4], offb[2];
#pragma unroll
    for (int q = 0; q < 4; q++)
        offa[q] = (min((i0 >> 4) + 4 * q + wave, (rowsA >> 4) - 1) * (ldka >> 5)) * 1024 + 16 * lane;
#pragma unroll
    for (int q = 0; q < 2; q++)
        offb[q] = (min((j0 >> 4) + 4 * q + wave, (rowsB >> 4) - 1) * (ldkb >> 5)) * 1024 + 16 * lane;
    // ranges: both planes of the rows readable from the entry's origin (bytes)
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        (void*)Ahg, (short)0, (int)(2 * (loA + (long long)rowsA * ldka)), 0x00020000);
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void*)Bhg, (short)0, (int)(2 * (loB + (long long)rowsB * ldkb)), 0x00020000);
    const int loAb = (int)(2 * loA), loBb = (int)(2 * loB);
    auto issue = [&](int s, int buf) {
        const int kb = ((klo >> 5) + s) * 1024;
        __bf16* st = lds + buf * STAGE;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void_t*)(st + (4 * q + wave) * 512), 16, offa[q], kb, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void_t*)(st + SA + (4 * q + wave) * 512), 16, offa[q],
                                                     kb + loAb, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void_t*)(st + 2 * SA + (4 * q + wave) * 512), 16, offb[q],
                                                     kb, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void_t*)(st + 2 * SA + SB + (4 * q + wave) * 512), 16,
                                                     offb[q], kb + loBb, 0, 0);
        }
    };

    v4f acc[8][4];
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

    struct Frag {
        bf16x8 ah[8], al[8], bh[4], bl[4];
    };
    Frag F0, F1;
    auto read_frags = [&](Frag& F, int buf) {
        const __bf16* sAh = lds + buf * STAGE;
        const __bf16* sAl = sAh + SA;
        const __bf16* sBh = sAh + 2 * SA;
        const __bf16* sBl = sBh + SB;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int c = wn * 64 + b * 16 + l16, o = c * KB + 8 * (lq ^ ((c >> 2) & 3));
            F.bh[b] = *(const bf16x8*)(sBh + o);
            F.bl[b] = *(const bf16x8*)(sBl + o);
        }
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const int r = wm * 128 + a * 16 + l16, o = r * KB + 8 * (lq ^ ((r >> 2) & 3));
            F.ah[a] = *(const bf16x8*)(sAh + o);
            F.al[a] = *(const bf16x8*)(sAl + o);
        }
    };
    auto mfmas = [&](const Frag& F) {
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                v4f t = acc[a][b];
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F.bl[b], F.ah[a], t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F.bh[b], F.al[a], t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F.bh[b], F.ah[a], t, 0, 0, 0);
                acc[a][b] = t;
            }
    };
    // stage s: its fragments are in Fc; those of stage s + 1 are read into Fn meanwhile
    auto step = [&](int s, Frag& Fc, Frag& Fn) {
        // this wave's reads of Fc and its loads of stage s + 1 complete (the loads of
        // stage s + 2, if any, stay in flight: 12 per stage); then every wave's
        const int later = min(NS - 2, nk - 2 - s);                 // stages loading after s + 1
        if (later >= 1) __builtin_amdgcn_s_waitcnt(0x0070 | 12);   // vmcnt(12) lgkmcnt(0)
        else __builtin_amdgcn_s_waitcnt(0x0070);                    // vmcnt(0) lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        if (s + NS < nk) issue(s + NS, s % NS);   // the buffer stage s was read from
        read_frags(Fn, (s + 1) % NS);             // unconditional: no branch between reads and MFMAs
        mfmas(Fc);
        // one LDS read after every four MFMAs
#pragma unroll
        for (int q = 0; q < 24; q++) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
        }
    };
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (s < nk) issue(s, s);
    if (nk > 0) {
        if (nk >= 3) __builtin_amdgcn_s_waitcnt(0x0F70 | (24 & 15) | ((24 >> 4) << 14));   // vmcnt(24)
        else if (nk == 2) __builtin_amdgcn_s_waitcnt(0x0F70 | 12);
        else __builtin_amdgcn_s_waitcnt(0x0F70);
        __builtin_amdgcn_s_barrier();
        read_frags(F0, 0);
    }
    int s = 0;
    for (; s + 1 < nk; s += 2) {
        step(s, F0, F1);
        step(s + 1, F1, F0);
    }
    if (s < nk) step(s, F0, F1);

    const double alpha = D.alpha, beta = D.beta;
    const int ldc = D.ldc, ldcin = D.ldcin;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const int gi = i0 + wm * 128 + a * 16 + l16;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int gj0 = j0 + wn * 64 + b * 16 + 4 * lq;
            double v[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gj = gj0 + r;
                v[r] = 0.0;
                if (gi < M && gj < N && (!(flags & G_CLOWER) || gi >= gj)) {
                    v[r] = alpha * (double)acc[a][b][r];
                    if (beta != 0.0) v[r] += beta * Cin[(long long)gj * ldcin + gi];
                    C[(long long)gj * ldc + gi] = v[r];
                }
            }
            copy_out(O, b1, b2, gi, gj0, M, N, v);
        }
    }
}

// k per LDS stage of the pre-split product: 64 (half the barriers per MFMA of 32, 72 KB
// of LDS for 128-wide tiles, still two workgroups per CU) measured 5-6 % faster over the
// whole mixed sweep at ns = 2865 (single matrix 0.314 -> 0.295 s, batch of 8 0.133 ->
// 0.126 s per matrix), same bits.  32 stays selectable for measurements.
int g_presplit_kb = 64;
// unpadded XOR-swizzled LDS rows: 2 % faster over the mixed sweep at ns = 2865 (single
// 0.293 -> 0.287 s, batch of 8 0.1235 -> 0.1219 s per matrix), same bits; padded rows
// stay selectable for measurements
int g_presplit_swizzle = 1;
// launches of at least this many 256 x 128 workgroups take the wide kernel
int g_wide_min_wgs = 1536;

size_t split_entry_elems(int rows, int k)
{
    const size_t ldk = (size_t)((k + SPLIT_PITCH - 1) / SPLIT_PITCH * SPLIT_PITCH);
    return 2 * (size_t)((rows + 127) / 128 * 128) * ldk;
}

size_t split_scratch_elems(int m, int n, int k, int entries)
{
    return (size_t)entries * (split_entry_elems(m, k) + split_entry_elems(n, k));
}

int split_operand(const double* X, int ld, int rows, int k, bool kcontig, int count, long long s1, int count2,
                  long long s2, __bf16* dst, hipStream_t s)
{
    if (rows <= 0 || k <= 0 || count <= 0) return 0;
    if (count2 < 1) count2 = 1;
    const int ldk = (k + SPLIT_PITCH - 1) / SPLIT_PITCH * SPLIT_PITCH;
    const long long rp = (rows + 127) / 128 * 128;
    dim3 grid(ldk / 64, (rows + 31) / 32, count * count2);
    hipLaunchKernelGGL(split_bf16_kernel, grid, dim3(256), 0, s, X, ld, rows, k, kcontig, dst, ldk, rp * ldk, count,
                       s1, s2, 2 * rp * ldk);
    return (int)hipGetLastError();
}

SplitView split_view(__bf16* p, int rows, int k, int count)
{
    SplitView v{};
    v.p = p;
    v.ldk = (k + SPLIT_PITCH - 1) / SPLIT_PITCH * SPLIT_PITCH;
    v.rows = (rows + 127) / 128 * 128;
    v.lo = (long long)v.rows * v.ldk;
    v.s1 = 2 * v.lo;
    v.s2 = v.s1 * count;
    v.pad = 1;
    return v;
}

// cfg: 0 = 64 x 64 tiles, 1 = 128 x 128, 2 = the 256 x 128 kernel (launches without triangular
// operands only), -1 = by size; with allowed (may be null) the configurations this launch
// may take (bit c for cfg c)
static int presplit_launch(Op ta, Op tb, const PresplitCall* calls, int ncalls, hipStream_t s, int cfg,
                           int* allowed = nullptr, int* automatic = nullptr)
{
    if (ncalls <= 0) return 0;
    if (ncalls > PRESPLIT_MAXD) return (int)hipErrorInvalidValue;
    GemmArgs args{};
    PresplitArgs pa{};
    long long work128 = 0, wide_wgs = 0;
    bool all_full = true, no_tri = true;
    int nd = 0;
    for (int q = 0; q < ncalls; q++) {
        GemmDesc d = calls[q].d;
        if (d.count2 < 1) d.count2 = 1;
        if (d.m <= 0 || d.n <= 0 || d.count <= 0) continue;
        const long long entries = (long long)d.count * d.count2;
        PresplitOperands& O = pa.op[nd];
        O.a = calls[q].a;
        O.b = calls[q].b;
        O.cs = calls[q].cs;
        O.cr = calls[q].cr;
        work128 += (long long)((d.m + 127) / 128) * ((d.n + 127) / 128) * entries;
        // the 256 x 128 kernel: lower-stored square C takes, per 128-wide column, ceil(rows on
        // or below its diagonal / 256) tiles
        const long long tw_m = (d.m + 255) / 256, tw_n = (d.n + 127) / 128;
        long long wt = tw_m * tw_n;
        if ((d.flags & G_CLOWER) && d.m == d.n) {
            wt = 0;
            for (long long tj = 0; tj < tw_n; tj++) wt += (tw_n - tj + 1) / 2;
        }
        wide_wgs += wt * entries;
        // buffer-load ranges: both planes of an entry's copy below 2 GB from its origin
        const long long span = 2 * std::max(O.a.lo + (long long)O.a.rows * O.a.ldk, O.b.lo + (long long)O.b.rows * O.b.ldk);
        // products with triangular operands never take it: its 256-row tiles would read the
        // copies' k-blocks above the 128-row diagonal blocks, which the copies of L^-1 leave unwritten
        const bool tri_op = d.flags & (G_ALOWER | G_AUPPER | G_BLOWER | G_BUPPER);
        all_full = all_full && d.flags == 0 && span < (1LL << 31);
        no_tri = no_tri && !tri_op && span < (1LL << 31);
        args.d[nd++] = d;
    }
    if (nd == 0) return 0;
    args.nd = nd;
    // By default only full products take the 256 x 128 kernel: in the c2 sweep (batch of 8) the
    // products with triangular operands (unequal tiles) and the lower-stored ones ran 2-18 %
    // slower on it than on twice as many 128 x 128 workgroups, the full product Z = S W 4 % faster.
    const bool wide_auto = (all_full || (g_wide_min_wgs == 0 && no_tri)) && wide_wgs >= g_wide_min_wgs;
    const int BT_auto = work128 >= 768 ? 128 : 64;
    if (allowed) *allowed = 3 | (no_tri ? 4 : 0);
    if (automatic) *automatic = wide_auto ? 2 : (BT_auto == 128 ? 1 : 0);
    if (allowed || automatic) return 0;   // (a query: nothing launched)
    if (cfg == 2 && !no_tri) return (int)hipErrorInvalidValue;
    const bool wide = cfg < 0 ? wide_auto : cfg == 2;
    const int BT = cfg < 0 ? BT_auto : (cfg == 1 ? 128 : 64);
    int z = 0, maxtiles = 0;
    for (int q = 0; q < nd; q++) {
        GemmDesc& d = args.d[q];
        const bool sq_lower = (d.flags & G_CLOWER) && d.m == d.n;
        int tiles;
        if (wide) {
            d.tiles_m = (d.m + 255) / 256;
            d.tiles_n = (d.n + 127) / 128;
            set_tile_order(d, sq_lower);   // round-robin dealing as the square tiles' (lower-stored launches)
            tiles = d.tiles_m * d.tiles_n;
            if (sq_lower) {
                tiles = 0;
                for (int tj = 0; tj < d.tiles_n; tj++) tiles += (d.tiles_n - tj + 1) / 2;
            }
        } else {
            d.tiles_m = (d.m + BT - 1) / BT;
            d.tiles_n = (d.n + BT - 1) / BT;
            set_tile_order(d, (d.flags & G_CLOWER) && d.tiles_m == d.tiles_n);
            tiles = d.tri_tiles ? d.tiles_m * (d.tiles_m + 1) / 2 : d.tiles_m * d.tiles_n;
        }
        d.zbase = z;
        z += d.count * d.count2;
        maxtiles = std::max(maxtiles, tiles);
    }
    const dim3 grid(maxtiles, z);
    void* tok = timing_begin(4 + (ta == Op::T ? 2 : 0) + (tb == Op::T ? 1 : 0), args.d, nd, s);
    if (wide) {
        hipLaunchKernelGGL(gemm_bf16x3w_kernel, grid, dim3(256), 0, s, args, pa);
    } else {
        const bool k64 = g_presplit_kb == 64, sw = g_presplit_swizzle != 0;
#define RGF_PRESPLIT(bt, kb, swz) \
        hipLaunchKernelGGL((gemm_bf16x3s_kernel<bt, kb, swz>), grid, dim3(256), 0, s, args, pa)
        if (BT == 128) {
            if (k64) { if (sw) RGF_PRESPLIT(128, 64, true); else RGF_PRESPLIT(128, 64, false); }
            else     { if (sw) RGF_PRESPLIT(128, 32, true); else RGF_PRESPLIT(128, 32, false); }
        } else {
            if (k64) { if (sw) RGF_PRESPLIT(64, 64, true); else RGF_PRESPLIT(64, 64, false); }
            else     { if (sw) RGF_PRESPLIT(64, 32, true); else RGF_PRESPLIT(64, 32, false); }
        }
#undef RGF_PRESPLIT
    }
    hipError_t e = hipGetLastError();
    timing_end(tok, s);
    return (int)e;
}

int gemm_bf16x3_presplit_n(Op ta, Op tb, const PresplitCall* calls, int ncalls, hipStream_t s)
{
    if (ncalls <= 0) return 0;
    if (ncalls > PRESPLIT_MAXD) return (int)hipErrorInvalidValue;
    int allowed = 0, def = 0;
    int rc = presplit_launch(ta, tb, calls, ncalls, s, -1, &allowed, &def);
    if (rc) return rc;
    int cands[3], nc = 0;
    cands[nc++] = def;
    for (int c = 0; c < 3; c++)
        if (c != def && (allowed >> c & 1)) cands[nc++] = c;
    GemmDesc d[PRESPLIT_MAXD];
    for (int q = 0; q < ncalls; q++) d[q] = calls[q].d;
    int cfg = tune_choose(1, ta, tb, d, ncalls, cands, nc, s, [&](int c, const GemmDesc* dd) {
        // trial launches: outputs redirected, no copies written
        PresplitCall t[PRESPLIT_MAXD];
        for (int q = 0; q < ncalls; q++) {
            t[q] = calls[q];
            t[q].d = dd[q];
            t[q].cs = SplitView{};
            t[q].cr = SplitView{};
        }
        return presplit_launch(ta, tb, t, ncalls, s, c);
    });
    if (cfg < 0 || !(allowed >> cfg & 1)) cfg = def;
    return presplit_launch(ta, tb, calls, ncalls, s, cfg);
}

int gemm_bf16x3_presplit(Op ta, Op tb, const GemmDesc* d, const __bf16* As, const __bf16* Bs, hipStream_t s,
                         __bf16* Cs, __bf16* Cr, int rowsA, int rowsB)
{
    const int count = std::max(d->count, 1);
    PresplitCall c{};
    c.d = *d;
    c.a = split_view((__bf16*)As, std::max(d->m, rowsA), d->k, count);
    c.b = split_view((__bf16*)Bs, std::max(d->n, rowsB), d->k, count);
    if (Cs) c.cs = split_view(Cs, d->n, d->m, count);
    if (Cr) c.cr = split_view(Cr, d->m, d->n, count);
    return gemm_bf16x3_presplit_n(ta, tb, &c, 1, s);
}

// One product on split-BF16 MFMA with operands split first into `scratch` (bf16
// elements, >= split_scratch_elems of the launch).  One descriptor per call.
int gemm_bf16x3_split(Op ta, Op tb, GemmDesc* d0, __bf16* scratch, size_t scratch_elems, hipStream_t s)
{
    GemmDesc d = *d0;
    if (d.count2 < 1) d.count2 = 1;
    if (d.m <= 0 || d.n <= 0 || d.count <= 0) return 0;
    const int entries = d.count * d.count2;
    if (split_scratch_elems(d.m, d.n, d.k, entries) > scratch_elems) return (int)hipErrorInvalidValue;
    __bf16* As = scratch;
    __bf16* Bs = scratch + (size_t)entries * split_entry_elems(d.m, d.k);
    int rc;
    // op(A)[i][k]: k contiguous iff TA; op(B)^T[j][k] = op(B)[k][j]: k contiguous iff !TB
    if ((rc = split_operand(d.A, d.lda, d.m, d.k, ta == Op::T, d.count, d.sA, d.count2, d.sA2, As, s))) return rc;
    if ((rc = split_operand(d.B, d.ldb, d.n, d.k, tb == Op::N, d.count, d.sB, d.count2, d.sB2, Bs, s))) return rc;
    return gemm_bf16x3_presplit(ta, tb, &d, As, Bs, s);
}

}  // namespace rgf
