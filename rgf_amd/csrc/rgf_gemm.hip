// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// One workgroup (4 waves, 256 lanes) computes a TM x TN tile of C (64x64, 128x64 or
// 128x128, chosen per launch from the tile count); each wave a quarter as MFMA
// 16x16 blocks.  K advances in steps of 16 through
// LDS (op(A) and op(B) stored k-major, one double of padding per row), with the
// next step's global loads issued into registers before the current step's
// MFMAs.  The MFMA is fed (op(B) fragment, op(A) fragment), so each lane's four
// accumulator values are four C columns of one C row; the 16 lanes of a lane
// group then store 16 consecutive rows of a column (a 128-byte segment) --
// coalesced for column-major C.
//
// Used for every dense product of the sweep: Schur updates (RGF.H:276-280,
// 304-308, 332), the triangular solves as products with L^-1 (RGF.H:260, 287,
// 317), the trtri merges (RGF.H:441, 450, 471, 487) and the Takahashi products
// (RGF.H:442, 454-458, 472, 491-501).
#include "rgf_internal.hpp"
#include "../../include/rgf.h"

#include <algorithm>
#include <vector>

namespace rgf {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int BK = 16, PADL = 1;

// resident workgroups per CU the register budget of each tile shape allows
constexpr int gemm_occupancy(int tm, int tn)
{
    return tm * tn >= 128 * 128 ? 2 : (tm * tn >= 128 * 64 ? 3 : (tm * tn >= 64 * 64 ? 5 : 8));
}

__device__ __forceinline__ void select_desc(const GemmArgs& args, int z, GemmDesc& D)
{
    D = args.d[0];
#pragma unroll
    for (int q = 1; q < GEMM_MAXD; q++)
        if (q < args.nd && z >= args.d[q].zbase) D = args.d[q];
}

template <bool TA, bool TB, int TM, int TN>
__global__ __launch_bounds__(256, gemm_occupancy(TM, TN)) void gemm_f64_kernel(GemmArgs args)
{
    constexpr int WM = TM / 2, WN = TN / 2;          // wave tile (waves in 2 x 2)
    constexpr int FA = WM / 16, FB = WN / 16;        // MFMA blocks per wave edge
    constexpr int LA = TM * BK / 256, LB = TN * BK / 256;   // loads per lane per k-step
    __shared__ double As[BK][TM + PADL];
    __shared__ double Bs[BK][TN + PADL];

    const int z = blockIdx.y;
    GemmDesc D;
    select_desc(args, z, D);
    const int flags = D.flags;
    // XCD-aware tile order: workgroups are dealt round-robin over the 8 XCDs, so
    // remap the launch index to give each XCD a contiguous range of tiles (their
    // A row-panels and B column-panels then meet in that XCD's L2) -- unless the
    // launch's tiles have unequal k-work (deal_rr), where the round-robin balances it.
    int g = blockIdx.x;
    if (!D.deal_rr) {
        const int G = gridDim.x, q = G >> 3, rem = G & 7, xcd = g & 7, loc = g >> 3;
        g = xcd < rem ? xcd * (q + 1) + loc : rem * (q + 1) + (xcd - rem) * q + loc;
    }
    int ti, tj;
    if (D.tri_tiles) {   // lower-triangular C: only the tiles on or below the diagonal
        if (!tri_tile(g, D.tiles_m, ti, tj)) return;
    } else {             // bands of 8 tile-rows, column-major inside a band
        if (g >= D.tiles_m * D.tiles_n) return;
        const int band = g / (8 * D.tiles_n);
        const int ti0 = band * 8, gsz = min(D.tiles_m - ti0, 8);
        const int in = g - band * 8 * D.tiles_n;
        ti = ti0 + in % gsz;
        tj = in / gsz;
    }
    const int i0 = ti * TM, j0 = tj * TN;
    if ((flags & G_CLOWER) && j0 > i0 + TM - 1) return;

    const int bz = z - D.zbase;
    const long long b1 = bz % D.count, b2 = bz / D.count;
    const double* A = D.A + b1 * D.sA + b2 * D.sA2;
    const double* B = D.B + b1 * D.sB + b2 * D.sB2;
    double* C = D.C + b1 * D.sC + b2 * D.sC2;
    const double* Cin = D.Cin + b1 * D.sCin + b2 * D.sCin2;
    const int M = D.m, N = D.n, K = D.k;
    const int lda = D.lda, ldb = D.ldb;

    int klo = 0, khi = K;
    if (flags & G_ALOWER) khi = min(khi, i0 + TM);
    if (flags & G_AUPPER) klo = max(klo, i0);
    if (flags & G_BLOWER) klo = max(klo, j0);
    if (flags & G_BUPPER) khi = min(khi, j0 + TN);
    klo = (klo / BK) * BK;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lq = lane >> 4;

    v4d acc[FA][FB];
#pragma unroll
    for (int a = 0; a < FA; a++)
#pragma unroll
        for (int b = 0; b < FB; b++) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

    double ra[LA], rb[LB];
    // lane -> (row or column, k) of the staged tiles: rows (columns) contiguous in
    // memory take (tid % TM, tid / TM + (256 / TM) q), k-contiguous ones (tid / 16 + 16 q, tid % 16)
    auto a_rk = [&](int q, int& r, int& kk) {
        if (!TA) { r = tid % TM; kk = tid / TM + (256 / TM) * q; }
        else     { kk = tid & 15; r = (tid >> 4) + 16 * q; }
    };
    auto b_ck = [&](int q, int& c, int& kk) {
        if (!TB) { kk = tid & 15; c = (tid >> 4) + 16 * q; }
        else     { c = tid % TN; kk = tid / TN + (256 / TN) * q; }
    };

    // global -> register loads for the k-step starting at k0
    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int q = 0; q < LA; q++) {
            int r, kk;
            a_rk(q, r, kk);
            const int gi = i0 + r, gk = k0 + kk;
            // op(A)[i][k] = A[i + k*lda] (N) or A[k + i*lda] (T)
            ra[q] = (gi < M && gk < khi) ? A[TA ? (long long)gi * lda + gk : (long long)gk * lda + gi] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < LB; q++) {
            int c, kk;
            b_ck(q, c, kk);
            const int gj = j0 + c, gk = k0 + kk;
            // op(B)[k][j] = B[k + j*ldb] (N) or B[j + k*ldb] (T)
            rb[q] = (gj < N && gk < khi) ? B[TB ? (long long)gk * ldb + gj : (long long)gj * ldb + gk] : 0.0;
        }
    };
    auto store_tiles = [&]() {
#pragma unroll
        for (int q = 0; q < LA; q++) {
            int r, kk;
            a_rk(q, r, kk);
            As[kk][r] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < LB; q++) {
            int c, kk;
            b_ck(q, c, kk);
            Bs[kk][c] = rb[q];
        }
    };

    // Full k-steps: each address is the tile origin moved to the k-step (uniform) plus a
    // loop-invariant 32-bit lane offset, with no per-load address arithmetic or bounds
    // branch.  Lanes on rows of op(A) beyond M (columns of op(B) beyond N) load the last
    // valid row (column) instead: they feed only C entries that are never stored.  The
    // k-step that crosses khi keeps the per-element masks (k beyond khi must read zero).
    const bool fast = args.fast_loads && (long long)(BK + TM) * lda < (1LL << 28) &&
                      (long long)(BK + TN) * ldb < (1LL << 28);
    unsigned aoff[LA], boff[LB];   // bytes (below 2^31 when fast)
#pragma unroll
    for (int q = 0; q < LA; q++) {
        int r, kk;
        a_rk(q, r, kk);
        const unsigned rr = (unsigned)min(r, M - 1 - i0);
        aoff[q] = 8u * (TA ? rr * (unsigned)lda + kk : (unsigned)kk * (unsigned)lda + rr);
    }
#pragma unroll
    for (int q = 0; q < LB; q++) {
        int c, kk;
        b_ck(q, c, kk);
        const unsigned cc = (unsigned)min(c, N - 1 - j0);
        boff[q] = 8u * (TB ? (unsigned)kk * (unsigned)ldb + cc : cc * (unsigned)ldb + kk);
    }
    const double* const At = A + (TA ? (long long)i0 * lda : (long long)i0);
    const double* const Bt = B + (TB ? (long long)j0 : (long long)j0 * ldb);
    auto load = [&](int k0) {
        if (fast && k0 + BK <= khi) {
            const char* pa = (const char*)(At + (TA ? (long long)k0 : (long long)k0 * lda));
            const char* pb = (const char*)(Bt + (TB ? (long long)k0 * ldb : (long long)k0));
#pragma unroll
            for (int q = 0; q < LA; q++) ra[q] = *(const double*)(pa + aoff[q]);
#pragma unroll
            for (int q = 0; q < LB; q++) rb[q] = *(const double*)(pb + boff[q]);
        } else {
            load_tiles(k0);
        }
    };

    if (klo < khi) load(klo);
    for (int k0 = klo; k0 < khi; k0 += BK) {
        store_tiles();
        __syncthreads();
        if (k0 + BK < khi) load(k0 + BK);
#pragma unroll
        for (int ks = 0; ks < BK / 4; ks++) {
            const int kr = ks * 4 + lq;
            double af[FA], bf[FB];
#pragma unroll
            for (int a = 0; a < FA; a++) af[a] = As[kr][wm * WM + a * 16 + l16];
#pragma unroll
            for (int b = 0; b < FB; b++) bf[b] = Bs[kr][wn * WN + b * 16 + l16];
#pragma unroll
            for (int a = 0; a < FA; a++)
#pragma unroll
                for (int b = 0; b < FB; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[b], af[a], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }

    // epilogue: accumulator item r of lane (l16 + 16*lq) is D[4*r + lq][l16] of the
    // 16x16 MFMA block (v_mfma_f64_16x16x4_f64 layout, checked on gfx950), i.e.
    // C[i = l16][j = 4*r + lq]
    const double alpha = D.alpha, beta = D.beta;
    const int ldc = D.ldc, ldcin = D.ldcin;
#pragma unroll
    for (int a = 0; a < FA; a++) {
        const int gi = i0 + wm * WM + a * 16 + l16;
#pragma unroll
        for (int b = 0; b < FB; b++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gj = j0 + wn * WN + b * 16 + 4 * r + lq;
                if (gi < M && gj < N && (!(flags & G_CLOWER) || gi >= gj)) {
                    double v = alpha * acc[a][b][r];
                    if (beta != 0.0) v += beta * Cin[(long long)gj * ldcin + gi];
                    C[(long long)gj * ldc + gi] = v;
                }
            }
        }
    }
}


// ---------------------------------------------------------------- kernel timing
// Per-process instrumentation used by the benchmark: HIP events recorded on the
// launching stream around every launch of one transpose class, so the measured
// time is the kernel's own duration on that stream.
namespace {
struct TimedLaunch {
    hipEvent_t a, b;
    double flops;
};
int g_timing_class = -1;
std::vector<TimedLaunch> g_timed;
long long g_timing_epoch = 0;   // changes whenever the timed events are destroyed

double canonical_flops(const GemmDesc& d)
{
    const int atri = d.flags & (G_ALOWER | G_AUPPER), btri = d.flags & (G_BLOWER | G_BUPPER);
    const int cl = d.flags & G_CLOWER;
    double f = 2.0 * d.m * (double)d.n * d.k * d.count * d.count2;
    if (atri && btri && cl) return f / 6.0;
    if (atri) f *= 0.5;
    if (btri) f *= 0.5;
    if (cl) f *= 0.5;
    return f;
}
}  // namespace

template <int TM, int TN>
static void launch_f64(Op ta, Op tb, const GemmArgs& args, dim3 grid, hipStream_t s)
{
    const dim3 block(256);
    if (ta == Op::N && tb == Op::N) hipLaunchKernelGGL((gemm_f64_kernel<false, false, TM, TN>), grid, block, 0, s, args);
    else if (ta == Op::N && tb == Op::T) hipLaunchKernelGGL((gemm_f64_kernel<false, true, TM, TN>), grid, block, 0, s, args);
    else if (ta == Op::T && tb == Op::N) hipLaunchKernelGGL((gemm_f64_kernel<true, false, TM, TN>), grid, block, 0, s, args);
    else hipLaunchKernelGGL((gemm_f64_kernel<true, true, TM, TN>), grid, block, 0, s, args);
}

int g_gemm_fast_loads = 1;
int g_rr_tri = 1;
int g_selinv_fuse = 1;   // used by the selected-inversion sweep
int g_selinv_pipe = 1;   // used by the selected-inversion sweep
int g_rr_bands = 0;
int g_small_tile_wgs = 1024;   // launches of fewer 64x64 workgroups than this take 32x32 tiles

// 64x64 workgroups a launch of these descriptors would have
static long long wgs64(const GemmDesc* descs, int ndesc)
{
    long long w = 0;
    for (int q = 0; q < ndesc; q++) {
        const GemmDesc& d = descs[q];
        if (d.m <= 0 || d.n <= 0 || d.count <= 0) continue;
        const long long tm = (d.m + 63) / 64, tn = (d.n + 63) / 64;
        const long long t = ((d.flags & G_CLOWER) && tm == tn) ? tm * (tm + 1) / 2 : tm * tn;
        w += t * d.count * std::max(d.count2, 1);
    }
    return w;
}

// Tile shape of a launch: cfg 0 = 64x64, 1 = 128x64, 2 = 128x128, 3 = 32x32, -1 = automatic.
// Measured on MI355X at the sweep's shapes (ns = 2865, single and 4-matrix batches):
// 64x64 at 5 workgroups per CU is the fastest or within a few percent for launches that
// fill the chip (the larger tiles lose occupancy); launches of fewer than four 64x64
// workgroups per CU (the upper merges of L^-1, the late rank-128 Cholesky updates) leave
// SIMDs with one or two waves and run at a fraction of the rate, so they take 32x32 tiles
// at up to 8 workgroups per CU (c2 FP64 step 0.489 -> 0.463 s; thresholds of 512 .. 1536
// workgroups within 1 %, 2100 slower).  Every tile shape sums each output over the same k-steps in
// the same order (tiles only skip k-steps of exact zeros), so the shape never changes a bit.
int gemm_f64(Op ta, Op tb, GemmDesc* descs, int ndesc, hipStream_t s, int cfg)
{
    if (ndesc <= 0) return 0;
    if (ndesc > GEMM_MAXD) return (int)hipErrorInvalidValue;
    if (cfg < 0) {
        const int def = wgs64(descs, ndesc) < g_small_tile_wgs ? 3 : 0;
        const int cands[4] = {def, def == 0 ? 3 : 0, 1, 2};
        cfg = tune_choose(0, ta, tb, descs, ndesc, cands, 4, s, [&](int c, const GemmDesc* dd) {
            GemmDesc t[GEMM_MAXD];
            std::copy(dd, dd + ndesc, t);
            return gemm_f64(ta, tb, t, ndesc, s, c);
        });
        if (cfg < 0 || cfg > 3) cfg = def;
    }
    const int TM = cfg == 3 ? 32 : (cfg >= 1 ? 128 : 64), TN = cfg == 3 ? 32 : (cfg == 2 ? 128 : 64);
    GemmArgs args{};
    int z = 0, maxtiles = 0, nd = 0;
    for (int q = 0; q < ndesc; q++) {
        GemmDesc d = descs[q];
        if (d.count2 < 1) d.count2 = 1;
        if (d.m <= 0 || d.n <= 0 || d.count <= 0) continue;
        d.tiles_m = (d.m + TM - 1) / TM;
        d.tiles_n = (d.n + TN - 1) / TN;
        set_tile_order(d, (d.flags & G_CLOWER) && d.tiles_m == d.tiles_n && TM == TN);
        d.zbase = z;
        z += d.count * d.count2;
        const int nt = d.tri_tiles ? d.tiles_m * (d.tiles_m + 1) / 2 : d.tiles_m * d.tiles_n;
        if (nt > maxtiles) maxtiles = nt;
        args.d[nd++] = d;
    }
    if (nd == 0) return 0;
    args.nd = nd;
    args.fast_loads = g_gemm_fast_loads;
    dim3 grid(maxtiles, z);
    const int cls = (ta == Op::T ? 2 : 0) + (tb == Op::T ? 1 : 0);
    void* tok = timing_begin(cls, args.d, nd, s);
    if (cfg == 3) launch_f64<32, 32>(ta, tb, args, grid, s);
    else if (cfg == 2) launch_f64<128, 128>(ta, tb, args, grid, s);
    else if (cfg == 1) launch_f64<128, 64>(ta, tb, args, grid, s);
    else launch_f64<64, 64>(ta, tb, args, grid, s);
    hipError_t err = hipGetLastError();
    timing_end(tok, s);
    return (int)err;
}

void* timing_begin(int op_class, const GemmDesc* d, int nd, hipStream_t s)
{
    // class 8 selects every split-BF16 class (4..7), class 9 every FP64 class (0..3)
    const bool on = !tune_in_trial() && (op_class == g_timing_class || (g_timing_class == 8 && op_class >= 4 && op_class <= 7) ||
                    (g_timing_class == 9 && op_class >= 0 && op_class <= 3));
    if (!on) return nullptr;
    auto* tl = new TimedLaunch{};
    hipEventCreate(&tl->a);
    hipEventCreate(&tl->b);
    tl->flops = 0.0;
    for (int q = 0; q < nd; q++) tl->flops += canonical_flops(d[q]);
    hipEventRecord(tl->a, s);
    return tl;
}

void timing_end(void* token, hipStream_t s)
{
    if (!token) return;
    auto* tl = (TimedLaunch*)token;
    hipEventRecord(tl->b, s);
    g_timed.push_back(*tl);
    delete tl;
}

long long timing_epoch() { return g_timing_epoch; }

}  // namespace rgf

extern "C" int rgf_kernel_timing_enable(int op_class)
{
    for (auto& t : rgf::g_timed) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
    rgf::g_timed.clear();
    rgf::g_timing_class = op_class;
    rgf::g_timing_epoch++;
    return 0;
}

extern "C" int rgf_kernel_timing_read(double* ms, double* flops, long long* launches)
{
    double tms = 0.0, tf = 0.0;
    for (auto& t : rgf::g_timed) {
        if (hipEventSynchronize(t.b) != hipSuccess) return RGF_EDEVICE;
        float e = 0.f;
        if (hipEventElapsedTime(&e, t.a, t.b) != hipSuccess) return RGF_EDEVICE;
        tms += e;
        tf += t.flops;
    }
    if (ms) *ms = tms;
    if (flops) *flops = tf;
    if (launches) *launches = (long long)rgf::g_timed.size();
    return 0;
}

