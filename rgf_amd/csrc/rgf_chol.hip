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
#include <algorithm>
#include <climits>

namespace rgf {

constexpr int NB = 64;

// ------------------------------------------------------------------ potrf
// One workgroup (4 waves) factors the kb x kb (kb <= 64) diagonal tile at A[k0, k0]
// and inverts it.  The tile is identity-padded beyond kb and held in LDS.
//   factor   four 16-column panels.  Wave 0 factors a panel with one row per lane in
//            registers: step j broadcasts the pivot of lane j (readlane), scales by
//            1/sqrt(pivot) (v_rsq_f64 + two Newton steps) and updates the rest of
//            the panel with column j (published in LDS).  Then the four waves apply
//            the rank-16 update to the trailing 16 x 16 blocks on FP64 MFMA.
//   inverse  X = L^-1 by blocks: the four 16 x 16 diagonal blocks by forward
//            substitution (one column per lane), then the merges
//            X21 = -X22 (L21 X11) at 16 -> 32 and 32 -> 64.
// Writes X_kk to xout (64 x 64, ld 64), the tile's sum of log L_jj (fixed shuffle
// tree), the first non-positive pivot and, if store_l, L_kk back into the tile (zeros
// above the diagonal).
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int LP = NB + 2;   // LDS row pitch in doubles (16-byte aligned rows)

__device__ __forceinline__ double bcast_lane(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(d) for d > 0: hardware estimate y0 and one third-order correction
// y = y0 (1 + e/2 + 3 e^2/8), e = 1 - d y0^2 (relative error O(e^3))
__device__ __forceinline__ double rsqrt_nr(double d)
{
    const double y0 = __builtin_amdgcn_rsq(d);
    const double e = fma(-d * y0, y0, 1.0);
    return fma(y0, e * fma(e, 0.375, 0.5), y0);
}

// (16 RA) x 64 x 64 MFMA product per wave on LDS operands, accumulated into the caller's
// C block: wave (wm, wn) adds sgn * X Y^T to rows 16 RA wm.., columns 32 wn.. (X, Y
// row-major with pitch LP, 64 columns).  acc[a][b] item t is row 16 RA wm + 16 a + l16,
// column 32 wn + 16 b + 4 t + lq.  RA = 2: a 64 x 64 C over the four waves; RA = 1: 32 x 64.
template <int RA>
__device__ __forceinline__ void mfma_nt(v4d (&acc)[RA][2], const double (*X)[LP], const double (*Y)[LP], double sgn,
                                        int wm, int wn, int l16, int lq)
{
#pragma unroll
    for (int ks = 0; ks < NB / 4; ks++) {
        double af[RA], bf[2];
#pragma unroll
        for (int a = 0; a < RA; a++) af[a] = sgn * X[16 * RA * wm + 16 * a + l16][4 * ks + lq];
#pragma unroll
        for (int b = 0; b < 2; b++) bf[b] = Y[32 * wn + 16 * b + l16][4 * ks + lq];
#pragma unroll
        for (int a = 0; a < RA; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[b], af[a], acc[a][b], 0, 0, 0);
    }
}

// acc += X Y^T (the layout and sign of mfma_nt with +1) for lower-triangular Y: output column
// block c0 + 16 b needs the k-steps below c0 + 16 b + 16 only (Y[c][k] = 0 for k > c); each
// output entry gets mfma_nt's MFMA sequence with the all-zero steps left out.  K0: k-steps of
// both column blocks of the wave, K1: of its second block.
template <int RA, int K0, int K1>
__device__ __forceinline__ void mfma_nt_lower_w(v4d (&acc)[RA][2], const double (*X)[LP], const double (*Y)[LP],
                                                int wm, int wn, int l16, int lq)
{
#pragma unroll
    for (int ks = 0; ks < K1; ks++) {
        double af[RA], bf[2];
#pragma unroll
        for (int a = 0; a < RA; a++) af[a] = X[16 * RA * wm + 16 * a + l16][4 * ks + lq];
#pragma unroll
        for (int b = 0; b < 2; b++) bf[b] = Y[32 * wn + 16 * b + l16][4 * ks + lq];
#pragma unroll
        for (int a = 0; a < RA; a++) {
            if (ks < K0) acc[a][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[0], af[a], acc[a][0], 0, 0, 0);
            acc[a][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[1], af[a], acc[a][1], 0, 0, 0);
        }
    }
}
template <int RA>
__device__ __forceinline__ void mfma_nt_lower(v4d (&acc)[RA][2], const double (*X)[LP], const double (*Y)[LP],
                                              int wm, int wn, int l16, int lq)
{
    if (wn == 0)
        mfma_nt_lower_w<RA, 4, 8>(acc, X, Y, wm, wn, l16, lq);
    else
        mfma_nt_lower_w<RA, 12, 16>(acc, X, Y, wm, wn, l16, lq);
}

// rows of the panel below the tile handled per workgroup pass (half a tile: the MFMA work
// of a pass, on the launch's critical path, is half that of a 64-row block)
constexpr int RB = 32;

// ---- phases of the Cholesky tile kernels (called by all 256 lanes of the workgroup)
// Ls -= P P^T on the lower quarters of the 64 x 64 tile (rows >= kb of P are zero).
__device__ __forceinline__ void tile_sub_panel(double (*Ls)[LP], const double (*Pk)[LP], int wave, int l16, int lq)
{
    const int wm = wave & 1, wn = wave >> 1;
    if (wn <= wm) {
        v4d acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int t = 0; t < 4; t++) acc[a][b][t] = Ls[32 * wm + 16 * a + l16][32 * wn + 16 * b + 4 * t + lq];
        mfma_nt<2>(acc, Pk, Pk, -1.0, wm, wn, l16, lq);
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int t = 0; t < 4; t++) Ls[32 * wm + 16 * a + l16][32 * wn + 16 * b + 4 * t + lq] = acc[a][b][t];
    }
    __syncthreads();
}

// The same over eight waves (16-row strips x 32-column halves; the two strips wholly above
// the diagonal skip).
__device__ __forceinline__ void tile_sub_panel8(double (*Ls)[LP], const double (*Pk)[LP], int wave, int l16, int lq)
{
    const int wm = wave & 3, wn = wave >> 2;
    if (16 * wm + 15 >= 32 * wn) {
        v4d acc[1][2];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[0][b][t] = Ls[16 * wm + l16][32 * wn + 16 * b + 4 * t + lq];
        mfma_nt<1>(acc, Pk, Pk, -1.0, wm, wn, l16, lq);
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) Ls[16 * wm + l16][32 * wn + 16 * b + 4 * t + lq] = acc[0][b][t];
    }
    __syncthreads();
}

// Cholesky of the tile in Ls (identity-padded beyond kb) in four 16-column panels; each
// row's 1/sqrt and sqrt of its pivot go to rinv and dgl; bad receives the first
// non-positive pivot (wave 0).
template <int NW = 4>
__device__ __forceinline__ void tile_factor(double (*Ls)[LP], double* colb, double* rinv, double* dgl, int kb,
                                            int& bad, int tid)
{
    const int lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    // ---- factor
#pragma unroll
    for (int K = 0; K < 4; K++) {
        const int o = 16 * K;
        if (wave == 0) {
            // Step j needs only uniform values on its critical path: the pivot d, and
            // a = A[j+1][j+1], b = A[j+1][j] of lane j+1 (current).  With y = 1/sqrt(d),
            // L[j+1][j] = b y and the next pivot is a - (b y)^2, the very operations
            // lane j+1 performs on its own row.  Each step first starts the next pivot's
            // 1/sqrt (region A), then does the row updates that hide its latency (region
            // B): columns j+1, j+2 with readlane values, the rest with column j read
            // from LDS, applied one step later so no LDS wait stalls the chain.
            const int r = lane;
            double p[16], cbv[16];
#pragma unroll
            for (int c = 0; c < 16; c++) p[c] = Ls[r][o + c];
            double d = bcast_lane(p[0], o);
            double b = bcast_lane(p[0], o + 1), a = bcast_lane(p[1], o + 1);
            bool ok = d > 0.0;
            double y = rsqrt_nr(ok ? d : 1.0);
            double lprev = 0.0;
            double my_rinv = 1.0, my_dgl = 1.0;   // 1/sqrt and sqrt of this lane's row's pivot
#pragma unroll
            for (int jj = 0; jj < 16; jj++) {
                const int j = o + jj;
                // region A: next pivot
                double s1 = 0.0, dn = 1.0, yn = 1.0;
                bool okn = true;
                if (jj < 15) {
                    s1 = b * y;                     // L[j+1][j]
                    dn = fma(-s1, s1, a);           // lane j+1's updated diagonal
                    okn = dn > 0.0;
                    yn = rsqrt_nr(okn ? dn : 1.0);
                }
                __builtin_amdgcn_sched_barrier(0);
                // region B: row updates
                if (jj >= 1 && jj <= 13) {
#pragma unroll
                    for (int k = jj + 2; k < 16; k++) p[k] = fma(-lprev, cbv[k], p[k]);
                }
                if (!ok && j < kb) bad = min(bad, j);
                // lane j holds p[jj] == d, so its entry is sqrt(d) = d y
                const double l = r >= j ? p[jj] * y : 0.0;
                p[jj] = l;
                if (r == j) {   // kept in registers, stored once after the panel
                    my_rinv = ok ? y : 1.0;
                    my_dgl = ok ? l : 1.0;
                }
                if (jj < 15) {
                    p[jj + 1] = fma(-l, s1, p[jj + 1]);
                    if (jj < 14) {
                        const double s2 = bcast_lane(l, j + 2);   // L[j+2][j]
                        p[jj + 2] = fma(-l, s2, p[jj + 2]);
                        b = bcast_lane(p[jj + 1], j + 2);
                        a = bcast_lane(p[jj + 2], j + 2);
                    }
                    if (jj < 13) {
                        colb[r] = l;
                        // one wave's LDS operations complete in issue order, so the reads
                        // below return every lane's write (and the next step's write comes
                        // after them); the fence only keeps the compiler from reordering
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
                        for (int k = jj + 3; k < 16; k++) cbv[k] = colb[o + k];
                        lprev = l;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                d = dn;
                y = yn;
                ok = okn;
            }
            if (r >= o) {
#pragma unroll
                for (int c = 0; c < 16; c++) Ls[r][o + c] = p[c];
            }
            if (r >= o && r < o + 16) {
                rinv[r] = my_rinv;
                dgl[r] = my_dgl;
            }
        }
        __syncthreads();
        // trailing blocks (I, J), K < J <= I: T_IJ -= L_IK L_JK^T, one MFMA block per wave
        // (accumulator item q of lane (l16, lq) is T[16 I + l16][16 J + 4 q + lq])
        const int nbr = 3 - K, nblk = nbr * (nbr + 1) / 2;
        for (int q = wave; q < nblk; q += NW) {
            int ii = 0;
            while ((ii + 1) * (ii + 2) / 2 <= q) ii++;
            const int I = K + 1 + ii, J = K + 1 + (q - ii * (ii + 1) / 2);
            v4d acc;
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = Ls[16 * I + l16][16 * J + 4 * t + lq];
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                const double xi = -Ls[16 * I + l16][o + 4 * ks + lq];
                const double yj = Ls[16 * J + l16][o + 4 * ks + lq];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(yj, xi, acc, 0, 0, 0);
            }
#pragma unroll
            for (int t = 0; t < 4; t++) Ls[16 * I + l16][16 * J + 4 * t + lq] = acc[t];
        }
        __syncthreads();
    }
}

// Sum of log L_jj of the tile (fixed shuffle tree) into *logdet_slot and the first bad
// pivot (1-based, offset info_base) into *info, by the lead workgroup.
__device__ __forceinline__ void tile_logdet(const double* dgl, int kb, int bad, bool lead, double* logdet_slot,
                                            int* info, int info_base, int tid)
{
    const int lane = tid & 63, wave = tid >> 6;
    if (wave == 0) {
        double v = lane < kb ? log(dgl[lane]) : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0 && lead) {
            *logdet_slot = v;
            if (bad != INT_MAX) atomicMin(info, info_base + bad + 1);
        }
    }

}

// X = L^-1 of the factored tile in Ls into Xs (Ts: merge scratch).
template <int NW = 4>
__device__ __forceinline__ void tile_inverse(const double (*Ls)[LP], double (*Xs)[LP],
                                             double (*Ts)[NB / 2][NB / 2 + 2], const double* rinv, int tid)
{
    const int lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    // ---- inverse, level 0: wave w inverts diagonal block w (lane c < 16 owns column c)
    if (lane < 16 && wave < 4) {
        const int o = 16 * wave, c = lane;
        double x[16];
#pragma unroll
        for (int rr = 0; rr < 16; rr++) {
            double s = rr == c ? 1.0 : 0.0;
#pragma unroll
            for (int q = 0; q < rr; q++) s = fma(-Ls[o + rr][o + q], x[q], s);
            x[rr] = s * rinv[o + rr];
        }
#pragma unroll
        for (int rr = 0; rr < 16; rr++) Xs[o + rr][o + c] = x[rr];
    }
    __syncthreads();
    // ---- level 1 (waves 0, 1: pair p = wave of 16-blocks) and level 2 (four waves, one
    // 16 x 16 block each): T = L21 X11, then X21 = -X22 T, on FP64 MFMA.  Accumulator
    // item t of lane (l16, lq) is row l16, column 4 t + lq of the 16 x 16 output block.
    if (wave < 2) {
        const int o = 32 * wave;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[o + 4 * ks + lq][o + l16], Ls[o + 16 + l16][o + 4 * ks + lq],
                                                       acc, 0, 0, 0);
#pragma unroll
        for (int t = 0; t < 4; t++) Ts[wave][l16][4 * t + lq] = acc[t];
    }
    __syncthreads();
    if (wave < 2) {
        const int o = 32 * wave;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[wave][4 * ks + lq][l16], Xs[o + 16 + l16][o + 16 + 4 * ks + lq],
                                                       acc, 0, 0, 0);
#pragma unroll
        for (int t = 0; t < 4; t++) Xs[o + 16 + l16][o + 4 * t + lq] = -acc[t];
    }
    __syncthreads();
    {
        // waves 0..3, one 16 x 16 block each (further waves only take part in the barriers)
        const int bi = 16 * ((wave >> 1) & 1), bj = 16 * (wave & 1);
        v4d acc = {0.0, 0.0, 0.0, 0.0};
        if (wave < 4) {
#pragma unroll
            for (int ks = 0; ks < 8; ks++)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[4 * ks + lq][bj + l16], Ls[32 + bi + l16][4 * ks + lq],
                                                           acc, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++) Ts[0][bi + l16][bj + 4 * t + lq] = acc[t];
        }
        __syncthreads();
        if (wave < 4) {
            acc = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < 8; ks++)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ts[0][4 * ks + lq][bj + l16],
                                                           Xs[32 + bi + l16][32 + 4 * ks + lq], acc, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++) Xs[32 + bi + l16][bj + 4 * t + lq] = -acc[t];
        }
    }
    __syncthreads();
}

// Column block k0 of the blocked Cholesky, one launch: tile factor and inverse, then the
// panel rows.  With prev_update the update of this column block by the previous panel
// (columns k0-64..k0-1), which the trailing update left out, is applied first: to the
// tile (A_kk -= P_k P_k^T) and to each row block (A_Rk -= P_R P_k^T).
__global__ __launch_bounds__(256) void potrf_tile_kernel(double* A, int lda, long long sA, int k0, int kb,
                                                         int below, bool prev_update, double* xout, long long sX,
                                                         double* logdet_slot, long long sP, int* info, int info_base,
                                                         bool store_l)
{
    {
        const long long b = blockIdx.y;   // batch entry
        A += b * sA;
        xout += b * sX;
        logdet_slot += b * sP;
        info += b;
    }
    const bool lead = blockIdx.x == 0;   // writes X_kk, the log-determinant part and info
    // One LDS pool, the per-step broadcast arrays at its start: below 64 KB their
    // addresses fit the instructions' immediate offsets (no address register set up per
    // access in the pivot chain).
    constexpr int TSZ = 2 * (NB / 2) * (NB / 2 + 2);
    __shared__ __attribute__((aligned(16))) double pool[3 * NB + 4 * NB * LP + TSZ];
    double* const colb = pool;
    double* const rinv = pool + NB;
    double* const dgl = pool + 2 * NB;
    double (*const Ls)[LP] = reinterpret_cast<double (*)[LP]>(pool + 3 * NB);
    double (*const Xs)[LP] = Ls + NB;
    double (*const Pk)[LP] = Ls + 2 * NB;   // previous panel, rows of this tile
    double (*const Pr)[LP] = Ls + 3 * NB;   // previous panel, rows of a row block
    double (*const Ts)[NB / 2][NB / 2 + 2] =
        reinterpret_cast<double (*)[NB / 2][NB / 2 + 2]>(pool + 3 * NB + 4 * NB * LP);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    int bad = INT_MAX;   // first non-positive pivot (wave 0, uniform)
    {
        // all 16 loads of a lane in flight at once (masked lanes read the tile's corner)
        double v[NB * NB / 256];
#pragma unroll
        for (int q = 0; q < NB * NB / 256; q++) {
            const int r = lane, c = 4 * q + wave;
            const bool in = r < kb && c < kb;
            v[q] = A[in ? (long long)(k0 + c) * lda + (k0 + r) : (long long)k0 * lda + k0];
        }
        double w[NB * NB / 256];
        if (prev_update) {
            const double* P = A + (long long)(k0 - NB) * lda;
#pragma unroll
            for (int q = 0; q < NB * NB / 256; q++) {
                const int r = lane, c = 4 * q + wave;
                w[q] = P[r < kb ? (long long)c * lda + (k0 + r) : (long long)k0];
            }
        }
#pragma unroll
        for (int q = 0; q < NB * NB / 256; q++) {
            const int r = lane, c = 4 * q + wave;
            Ls[r][c] = (r < kb && c < kb) ? (c <= r ? v[q] : 0.0) : (r == c ? 1.0 : 0.0);
            Xs[r][c] = 0.0;
            if (prev_update) Pk[r][c] = r < kb ? w[q] : 0.0;
        }
    }
    __syncthreads();
    if (prev_update) tile_sub_panel(Ls, Pk, wave, l16, lq);

    tile_factor(Ls, colb, rinv, dgl, kb, bad, tid);
    tile_logdet(dgl, kb, bad, lead, logdet_slot, info, info_base + k0, tid);
    tile_inverse(Ls, Xs, Ts, rinv, tid);
    for (int e = tid; e < NB * NB; e += 256) {
        const int r = e & (NB - 1), c = e >> 6;
        if (lead) {
            xout[c * NB + r] = Xs[r][c];
            if (store_l && r < kb && c < kb) A[(long long)(k0 + c) * lda + (k0 + r)] = r >= c ? Ls[r][c] : 0.0;
        }
    }

    // ---- the panel below the tile, in place: L[R][k] = A[R][k] X_kk^T for row blocks R
    // of RB = 32 rows (block rb by workgroup rb mod gridDim.x).  Each wave computes a
    // 16 x 32 part on FP64 MFMA; item t of lane (l16, lq) is row l16, column 4 t + lq of
    // each 16 x 16 block.
    const int nrb = (below + RB - 1) / RB;
    const int wm = wave & 1, wn = wave >> 1;
    for (int rb = blockIdx.x; rb < nrb; rb += gridDim.x) {
        const int r0 = k0 + kb + rb * RB, rows = min(RB, below - rb * RB);
        __syncthreads();   // Ls is free (inverse, stores or the previous block done)
        {
            // element e = q * 256 + tid: row e % RB, column e / RB (32 rows of a column in
            // a row of lanes: coalesced)
            double v[RB * NB / 256], w[RB * NB / 256];
#pragma unroll
            for (int q = 0; q < RB * NB / 256; q++) {
                const int e = q * 256 + tid, r = e % RB, c = e / RB;
                const bool in = r < rows && c < kb;
                v[q] = A[in ? (long long)(k0 + c) * lda + (r0 + r) : (long long)k0 * lda + r0];
                if (prev_update) w[q] = A[r < rows ? (long long)(k0 - NB + c) * lda + (r0 + r) : (long long)k0 * lda + r0];
            }
#pragma unroll
            for (int q = 0; q < RB * NB / 256; q++) {
                const int e = q * 256 + tid, r = e % RB, c = e / RB;
                Ls[r][c] = (r < rows && c < kb) ? v[q] : 0.0;
                if (prev_update) Pr[r][c] = r < rows ? w[q] : 0.0;
            }
        }
        __syncthreads();
        v4d acc[1][2];
        if (prev_update) {   // A_Rk -= P_R P_k^T, in place in Ls (each lane rewrites its own entries)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int t = 0; t < 4; t++) acc[0][b][t] = Ls[16 * wm + l16][32 * wn + 16 * b + 4 * t + lq];
            mfma_nt<1>(acc, Pr, Pk, -1.0, wm, wn, l16, lq);
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int t = 0; t < 4; t++) Ls[16 * wm + l16][32 * wn + 16 * b + 4 * t + lq] = acc[0][b][t];
            __syncthreads();
        }
#pragma unroll
        for (int b = 0; b < 2; b++) acc[0][b] = (v4d){0.0, 0.0, 0.0, 0.0};
        mfma_nt<1>(acc, Ls, Xs, 1.0, wm, wn, l16, lq);   // R X^T (X^T[k][j] = X[j][k])
        const int row = 16 * wm + l16;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int col = 32 * wn + 16 * b + 4 * t + lq;
                if (row < rows && col < kb) A[(long long)(k0 + col) * lda + (r0 + row)] = acc[0][b][t];
            }
    }
}

// ------------------------------------------------------------------ potrf, one launch
// The blocked Cholesky of a diagonal block as ONE launch over its 64 x 64 tiles (T tile
// rows), a task graph run by G workgroups per matrix (grid G x count):
//   workgroup 0, the chain, runs step k = 0 .. T-1 on LDS operands:
//     L_{k,k-1} = A_{k,k-1} X_{k-1}^T (stored), tile (k,k) -= L_{k,k-1} L_{k,k-1}^T,
//     factor + inverse + log-diagonal of the tile (X_k stored), then
//     A_{k+1,k} -= L_{k+1,k-1} L_{k,k-1}^T, kept in LDS for step k + 1;
//   workgroups 1 .. G-1 claim the other tasks from the matrix's ticket counter in the
//   order of generations g = 0, 1, ...:
//     TRSM(i, g), i >= g + 2: A_ig -= L_{i,g-1} L_{g,g-1}^T, then L_ig = A_ig X_g^T;
//     UPDATE(i, j, g - 1), i >= j >= g + 1 (nearest columns first): A_ij -= L_{i,g-1} L_{j,g-1}^T;
//     with an X buffer, the inverse X = L^-1 beside the factorization, by columns of tiles:
//     XACC(i, j, k = g - 2), i > k >= j: Y_ij += L_ik X_kj (Y_ij = L_ij X_jj for k = j), and
//     for k = i - 1 then X_ij = -X_ii Y_ij (X_jj: the chain's tile inverses).
// Every tile receives panels 0, 1, ... in increasing order (per-tile counters), so the
// result does not depend on which workgroup runs which task.  Dependencies are per-matrix
// flags in global memory, written after an agent-scope release and read before an
// agent-scope acquire.  Chain step g, then generation g, for g = 0, 1, ..., is a
// topological order of the tasks and the workers claim in it, so the launch progresses
// with the chain and any one worker resident.  Every wait is bounded in time: when one
// expires the matrix's info becomes INT_MIN and its workgroups leave.  The last workgroup
// of a matrix to leave clears the matrix's flags, which the next launch expects at zero.
constexpr int FL_TICKET = 0, FL_EXIT = 1, FL_ABORT = 2, FL_HDR = 4;

// flags: header, xdone[T], then T x T arrays ldone, cnt, acnt (terms summed into Y_ij), xrdy (X_ij stored)
__host__ __device__ inline int dag_flag_elems(int T) { return FL_HDR + T + 4 * T * T; }

// Tasks of generation g in claim order: ntr TRSM(i, g); nup UPDATE(i, j, g - 1) (column j
// ascending, rows i >= j ascending), without the first npo of them when they were promoted
// into generation g - 1; npi promoted UPDATE(g + 2, g + 2, g) and UPDATE(g + 3, g + 2, g) --
// the chain's needs at step g + 2 (its diagonal tile and the P3 operand A_{g+3,g+2}), one
// generation before their own, right after their inputs (TRSM(., g) and the UPDATE(., ., g - 1)
// of the same tiles) in the order; nxa XACC(i, j, k = g - 2 - xlag): the inverse's terms, off
// the chain's path, lagging xlag generations so the chain's inputs are claimed first.
struct DagGen {
    int ntr, nup, npo, npi, nxa;
    __host__ __device__ int total() const { return ntr + nup - npo + npi + nxa; }
};
__host__ __device__ inline DagGen dag_gen(int T, int g, bool with_x, int xlag, bool promo)
{
    DagGen d;
    d.ntr = T - g - 2 > 0 ? T - g - 2 : 0;
    d.nup = g >= 1 && T - g - 1 > 0 ? (T - g - 1) * (T - g) / 2 : 0;
    d.npo = promo && g >= 1 ? min(2, max(0, T - g - 1)) : 0;
    d.npi = promo ? min(2, max(0, T - g - 2)) : 0;
    const int k = g - 2 - xlag;
    d.nxa = with_x && k >= 0 && k < T - 1 ? (k + 1) * (T - 1 - k) : 0;
    return d;
}

__host__ __device__ inline int dag_gens(int T, bool with_x, int xlag) { return with_x ? T + 1 + xlag : T - 1; }

__host__ __device__ inline int dag_tasks(int T, bool with_x, int xlag, bool promo)
{
    int total = 0;
    for (int g = 0; g < dag_gens(T, with_x, xlag); g++) total += dag_gen(T, g, with_x, xlag, promo).total();
    return total;
}

// Task t of the claim order: kind 1 TRSM(i, g) (j = g), 2 UPDATE(i, j, p), 6 XACC(i, j, k = p);
// g its generation
struct DagTask {
    int kind, i, j, p, g;
};
__host__ __device__ inline DagTask dag_decode(int T, bool with_x, int xlag, bool promo, int t)
{
    int g = 0;
    DagGen dg = dag_gen(T, g, with_x, xlag, promo);
    while (t >= dg.total()) {
        t -= dg.total();
        dg = dag_gen(T, ++g, with_x, xlag, promo);
    }
    const int ntr = dg.ntr, nupg = dg.nup - dg.npo;
    if (t >= ntr + nupg + dg.npi) {   // XACC(i, j, k = g - 2 - xlag)
        t -= ntr + nupg + dg.npi;
        const int k = g - 2 - xlag;
        return {6, k + 1 + t / (k + 1), t % (k + 1), k, g};
    }
    if (t < ntr) return {1, g + 2 + t, g, g, g};   // TRSM(i, g)
    t -= ntr;
    if (t >= nupg) return {2, g + 2 + (t - nupg), g + 2, g, g};   // UPDATE promoted into g (p = g)
    t += dg.npo;   // UPDATE(i, j, g - 1), column j ascending, rows i >= j
    int j = g + 1;
    while (t >= T - j) {
        t -= T - j;
        j++;
    }
    return {2, j + t, j, g - 1, g};
}

struct DagArgs {
    double* A;
    long long sA;
    double* xdiag;
    long long sX;
    double* parts;
    long long sP;
    int* info;
    int* flags;
    long long sF;
    long long timeout;   // wall-clock ticks one wait may take
    int lda, n, T, ntasks, info_base;
    long long* trace;    // null, or: [0] record count, then records of 8 (see dag_trace)
    long long trace_cap;
    double* xbuf;        // null, or the off-diagonal tiles of X = L^-1 (ld ldx, per matrix sXb)
    long long sXb;
    int ldx;
    int xlag, promo;   // claim order (dag_gen)
};

// (thread 0, tracing) one record: workgroup, kind (0 chain step, 1 TRSM, 2 UPDATE, 3 the
// chain's A_{k+1,k} update), i, j, g and three wall-clock stamps
__device__ __forceinline__ void dag_trace(const DagArgs& a, int kind, int i, int j, int g, long long t0,
                                          long long t1, long long t2)
{
    if (!a.trace) return;
    const long long r = atomicAdd((unsigned long long*)a.trace, 1ULL);
    if (r >= a.trace_cap) return;
    long long* q = a.trace + 1 + 8 * r;
    q[0] = blockIdx.x + 65536LL * blockIdx.y;
    q[1] = kind;
    q[2] = i;
    q[3] = j;
    q[4] = g;
    q[5] = t0;
    q[6] = t1;
    q[7] = t2;
}

__device__ __forceinline__ int dag_peek(const int* f)
{
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// (thread 0) until *f >= v; false when the matrix aborted or the deadline passed
__device__ __forceinline__ bool dag_wait(const int* f, int v, const int* fl, long long deadline)
{
    while (dag_peek(f) < v) {
        if (dag_peek(fl + FL_ABORT) || wall_clock64() > deadline) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Tiles shared between workgroups are read and written only by agent-scope atomic loads
// and stores (relaxed): each such access is coherent across the XCDs by itself, so no
// L2 write-back or invalidation is needed.  A flag is stored after the workgroup's
// stores have completed (every wave waits for its memory counters, then the barrier),
// and the tiles it guards are loaded only after it has been seen.
__device__ __forceinline__ double ldc(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stc(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// All threads: thread 0 waits for up to three flags (null: none), then the workgroup
// proceeds.  false: the matrix aborted (a wait expired here or elsewhere).
__device__ __forceinline__ bool dag_acquire(const DagArgs& a, int* fl, int* info, const int* f0, int v0,
                                            const int* f1, int v1, const int* f2, int v2, int* s_ok, int tid)
{
    __syncthreads();
    if (tid == 0) {
        const long long dl = wall_clock64() + a.timeout;
        bool ok = (!f0 || dag_wait(f0, v0, fl, dl)) && (!f1 || dag_wait(f1, v1, fl, dl)) &&
                  (!f2 || dag_wait(f2, v2, fl, dl));
        if (!ok) {
            __hip_atomic_store(fl + FL_ABORT, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            atomicMin(info, INT_MIN);
        }
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok;
}

// All threads: once every store of the workgroup has completed, set *f = v.
__device__ __forceinline__ void dag_release(int* f, int v, int tid)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the same with a second flag
__device__ __forceinline__ void dag_release2(int* f, int v, int* f2, int v2, int tid)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
        __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(f2, v2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// workgroup of the task-graph launch: eight waves (two per SIMD); a 64 x 64 tile product
// gives each wave a 16 x 32 part (mfma_nt<1>)
constexpr int DNT = 512, DNW = DNT / 64, DRA = 8 / DNW;

// rows x cols of the tile at (r0, c0) into S, zero padded (lanes along rows: coalesced)
__device__ __forceinline__ void dag_load(double (*S)[LP], const double* A, int lda, int r0, int c0, int rows,
                                         int cols, int tid)
{
    double v[NB * NB / DNT];
#pragma unroll
    for (int q = 0; q < NB * NB / DNT; q++) {
        const int e = q * DNT + tid, r = e & (NB - 1), c = e >> 6;
        const bool in = r < rows && c < cols;
        v[q] = ldc(A + (in ? (long long)(c0 + c) * lda + (r0 + r) : (long long)c0 * lda + r0));
        v[q] = in ? v[q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NB * NB / DNT; q++) {
        const int e = q * DNT + tid, r = e & (NB - 1), c = e >> 6;
        S[r][c] = v[q];
    }
}

// the same, transposed: S[c][r] = tile(r, c)
__device__ __forceinline__ void dag_load_t(double (*S)[LP], const double* A, int lda, int r0, int c0, int rows,
                                           int cols, int tid)
{
    double v[NB * NB / DNT];
#pragma unroll
    for (int q = 0; q < NB * NB / DNT; q++) {
        const int e = q * DNT + tid, r = e & (NB - 1), c = e >> 6;
        const bool in = r < rows && c < cols;
        v[q] = ldc(A + (in ? (long long)(c0 + c) * lda + (r0 + r) : (long long)c0 * lda + r0));
        v[q] = in ? v[q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NB * NB / DNT; q++) {
        const int e = q * DNT + tid, r = e & (NB - 1), c = e >> 6;
        S[c][r] = v[q];
    }
}

// the wave's (16 RA) x 32 part of a 64 x 64 accumulator (mfma_nt<RA> layout) from / to LDS
template <int RA>
__device__ __forceinline__ void acc_from(v4d (&acc)[RA][2], const double (*S)[LP], int wm, int wn, int l16, int lq)
{
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[a][b][t] = S[16 * RA * wm + 16 * a + l16][32 * wn + 16 * b + 4 * t + lq];
}
template <int RA>
__device__ __forceinline__ void acc_to(const v4d (&acc)[RA][2], double (*S)[LP], int wm, int wn, int l16, int lq)
{
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) S[16 * RA * wm + 16 * a + l16][32 * wn + 16 * b + 4 * t + lq] = acc[a][b][t];
}
template <int RA>
__device__ __forceinline__ void acc_to_t(const v4d (&acc)[RA][2], double (*S)[LP], int wm, int wn, int l16, int lq)
{
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) S[32 * wn + 16 * b + 4 * t + lq][16 * RA * wm + 16 * a + l16] = acc[a][b][t];
}
// the accumulator from the tile at (r0, c0) (64 x 64 readable)
template <int RA>
__device__ __forceinline__ void acc_load(v4d (&acc)[RA][2], const double* A, int lda, int r0, int c0, int wm, int wn,
                                         int l16, int lq)
{
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++)
                acc[a][b][t] = ldc(A + (long long)(c0 + 32 * wn + 16 * b + 4 * t + lq) * lda + (r0 + 16 * RA * wm + 16 * a + l16));
}
template <int RA>
__device__ __forceinline__ void acc_zero(v4d (&acc)[RA][2])
{
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
}
// rows x cols of the accumulator into the tile at (r0, c0)
template <int RA>
__device__ __forceinline__ void acc_store(const v4d (&acc)[RA][2], double* A, int lda, int r0, int c0, int rows,
                                          int cols, int wm, int wn, int l16, int lq)
{
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int r = 16 * RA * wm + 16 * a + l16, c = 32 * wn + 16 * b + 4 * t + lq;
                if (r < rows && c < cols) stc(A + (long long)(c0 + c) * lda + (r0 + r), acc[a][b][t]);
            }
}

__global__ __launch_bounds__(DNT) void potrf_dag_kernel(DagArgs a)
{
    const long long mb = blockIdx.y;
    double* const A = a.A + mb * a.sA;
    double* const xd = a.xdiag + mb * a.sX;
    double* const parts = a.parts + mb * a.sP;
    int* const info = a.info + mb;
    int* const fl = a.flags + mb * a.sF;
    int* const xdone = fl + FL_HDR;
    int* const ldone = xdone + a.T;   // ldone[i * T + k]: L_ik stored
    int* const cnt = ldone + a.T * a.T;   // cnt[i * T + j]: panels applied to tile (i, j)
    int* const acnt = cnt + a.T * a.T;    // acnt[i * T + j]: terms summed into Y_ij
    int* const xrdy = acnt + a.T * a.T;   // xrdy[i * T + j]: X_ij stored
    const int lda = a.lda, n = a.n, T = a.T;
    const bool with_x = a.xbuf != nullptr;
    double* const xb = with_x ? a.xbuf + mb * a.sXb : nullptr;
    const int ldx = a.ldx;

    constexpr int TSZ = 2 * (NB / 2) * (NB / 2 + 2);
    __shared__ __attribute__((aligned(16))) double pool[3 * NB + 4 * NB * LP + TSZ];
    __shared__ int s_ok, s_task;
    double* const colb = pool;
    double* const rinv = pool + NB;
    double* const dgl = pool + 2 * NB;
    double (*const B0)[LP] = reinterpret_cast<double (*)[LP]>(pool + 3 * NB);
    double (*const B1)[LP] = B0 + NB;
    double (*const B2)[LP] = B0 + 2 * NB;
    double (*const B3)[LP] = B0 + 3 * NB;
    double (*const Ts)[NB / 2][NB / 2 + 2] =
        reinterpret_cast<double (*)[NB / 2][NB / 2 + 2]>(pool + 3 * NB + 4 * NB * LP);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int wm = wave % (4 / DRA), wn = wave / (4 / DRA);
    for (int e = tid; e < 4 * NB * LP; e += DNT) pool[3 * NB + e] = 0.0;   // X's upper blocks stay zero
    __syncthreads();
    v4d acc[DRA][2];

    if (blockIdx.x == 0) {
        // ---- the chain
        double (*Bl)[LP] = B0;   // tile (k, k); later L_{k+1,k-1}
        double (*Bx)[LP] = B1;   // X_{k-1}, then X_k
        double (*Ba)[LP] = B2;   // A_{k,k-1} (panels <= k-2 applied), then L_{k,k-1}
        double (*Bn)[LP] = B3;   // A_{k+1,k} being brought to panel k-1
        for (int k = 0; k < T; k++) {
            const int k0 = NB * k, kb = min(NB, n - k0);
            const long long c0 = wall_clock64();
            if (!dag_acquire(a, fl, info, k >= 2 ? cnt + k * T + k : nullptr, k - 1, nullptr, 0, nullptr, 0, &s_ok,
                             tid))
                break;
            const long long c1 = wall_clock64();
            {   // tile (k, k): lower triangle, identity beyond kb
                double v[NB * NB / DNT];
#pragma unroll
                for (int q = 0; q < NB * NB / DNT; q++) {
                    const int e = q * DNT + tid, r = e & (NB - 1), c = e >> 6;
                    const bool in = r < kb && c < kb;
                    v[q] = ldc(A + (in ? (long long)(k0 + c) * lda + (k0 + r) : (long long)k0 * lda + k0));
                }
#pragma unroll
                for (int q = 0; q < NB * NB / DNT; q++) {
                    const int e = q * DNT + tid, r = e & (NB - 1), c = e >> 6;
                    Bl[r][c] = (r < kb && c < kb) ? (c <= r ? v[q] : 0.0) : (r == c ? 1.0 : 0.0);
                }
            }
            __syncthreads();
            const long long cl = wall_clock64();
            if (k >= 1) {
                acc_zero(acc);
                mfma_nt_lower<DRA>(acc, Ba, Bx, wm, wn, l16, lq);   // L_{k,k-1} = A_{k,k-1} X_{k-1}^T (X lower)
                acc_store(acc, A, lda, k0, k0 - NB, kb, NB, wm, wn, l16, lq);
                __syncthreads();
                acc_to(acc, Ba, wm, wn, l16, lq);
                __syncthreads();
                tile_sub_panel8(Bl, Ba, wave, l16, lq);
            }
            const long long cp = wall_clock64();
            int bad = INT_MAX;
            tile_factor<DNW>(Bl, colb, rinv, dgl, kb, bad, tid);
            const long long cf = wall_clock64();
            tile_logdet(dgl, kb, bad, true, parts + k, info, a.info_base + k0, tid);
            tile_inverse<DNW>(Bl, Bx, Ts, rinv, tid);
            const long long ci = wall_clock64();
            for (int e = tid; e < NB * NB; e += DNT) {
                const int r = e & (NB - 1), c = e >> 6;
                stc(xd + (long long)k * NB * NB + c * NB + r, Bx[r][c]);
            }
            if (k >= 1)
                dag_release2(xdone + k, 1, ldone + k * T + (k - 1), 1, tid);
            else
                dag_release(xdone + k, 1, tid);
            if (tid == 0) {
                const long long ce = wall_clock64();
                dag_trace(a, 0, k, k, k, c0, c1, ce);
                dag_trace(a, 4, k, k, k, cl, cp, cf);
                dag_trace(a, 5, k, k, k, ci, ce, 0);
            }
            if (k + 1 < T) {
                const int kb1 = min(NB, n - k0 - NB);
                const long long d0 = wall_clock64();
                if (!dag_acquire(a, fl, info, k >= 1 ? ldone + (k + 1) * T + (k - 1) : nullptr, 1,
                                 k >= 2 ? cnt + (k + 1) * T + k : nullptr, k - 1, nullptr, 0, &s_ok, tid))
                    break;
                const long long d1 = wall_clock64();
                dag_load(Bn, A, lda, k0 + NB, k0, kb1, NB, tid);
                if (k >= 1) dag_load(Bl, A, lda, k0 + NB, k0 - NB, kb1, NB, tid);   // L_{k+1,k-1}
                __syncthreads();
                if (k >= 1) {
                    acc_from(acc, Bn, wm, wn, l16, lq);
                    mfma_nt<DRA>(acc, Bl, Ba, -1.0, wm, wn, l16, lq);
                    __syncthreads();
                    acc_to(acc, Bn, wm, wn, l16, lq);
                }
                double (*t)[LP] = Ba;
                Ba = Bn;
                Bn = t;
                if (tid == 0) dag_trace(a, 3, k + 1, k, k, d0, d1, wall_clock64());
            }
        }
    } else {
        // ---- workers
        for (;;) {
            __syncthreads();
            if (tid == 0)
                s_task = dag_peek(fl + FL_ABORT)
                             ? -1
                             : __hip_atomic_fetch_add(fl + FL_TICKET, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            int t = s_task;
            if (t < 0 || t >= a.ntasks) break;
            const long long w0 = wall_clock64();
            const DagTask tk = dag_decode(T, with_x, a.xlag, a.promo, t);
            const int g = tk.g;
            if (tk.kind == 6) {
                // XACC(i, j, k)
                const int k = tk.p, i = tk.i, j = tk.j;
                const int ri = min(NB, n - NB * i);
                if (!dag_acquire(a, fl, info, ldone + i * T + k, 1, k == j ? xdone + j : xrdy + k * T + j, 1,
                                 k > j ? acnt + i * T + j : nullptr, k - j, &s_ok, tid))
                    break;
                const long long w1 = wall_clock64();
                dag_load(B0, A, lda, NB * i, NB * k, ri, NB, tid);   // L_ik
                if (k == j)
                    dag_load_t(B1, xd + (long long)j * NB * NB, NB, 0, 0, NB, NB, tid);
                else
                    dag_load_t(B1, xb, ldx, NB * k, NB * j, NB, NB, tid);
                if (k == i - 1) dag_load(B2, xd + (long long)i * NB * NB, NB, 0, 0, NB, NB, tid);   // X_ii
                if (k > j)
                    acc_load(acc, xb, ldx, NB * i, NB * j, wm, wn, l16, lq);
                else
                    acc_zero(acc);
                __syncthreads();
                mfma_nt<DRA>(acc, B0, B1, 1.0, wm, wn, l16, lq);   // Y_ij += L_ik X_kj
                if (k < i - 1) {
                    acc_store(acc, xb, ldx, NB * i, NB * j, NB, NB, wm, wn, l16, lq);
                    dag_release(acnt + i * T + j, k - j + 1, tid);
                } else {
                    acc_to_t(acc, B3, wm, wn, l16, lq);
                    __syncthreads();
                    acc_zero(acc);
                    mfma_nt<DRA>(acc, B2, B3, -1.0, wm, wn, l16, lq);   // X_ij = -X_ii Y_ij
                    acc_store(acc, xb, ldx, NB * i, NB * j, NB, NB, wm, wn, l16, lq);
                    dag_release(xrdy + i * T + j, 1, tid);
                }
                if (tid == 0) dag_trace(a, 6, i, j, g, w0, w1, wall_clock64());
            } else if (tk.kind == 1) {
                // TRSM(i, g)
                const int i = tk.i, rows = min(NB, n - NB * i);
                if (!dag_acquire(a, fl, info, xdone + g, 1, g >= 1 ? ldone + i * T + (g - 1) : nullptr, 1,
                                 g >= 2 ? cnt + i * T + g : nullptr, g - 1, &s_ok, tid))
                    break;
                const long long w1 = wall_clock64();
                dag_load(B3, A, lda, NB * i, NB * g, rows, NB, tid);
                dag_load(B1, xd + (long long)g * NB * NB, NB, 0, 0, NB, NB, tid);
                if (g >= 1) {
                    dag_load(B0, A, lda, NB * i, NB * (g - 1), rows, NB, tid);
                    dag_load(B2, A, lda, NB * g, NB * (g - 1), NB, NB, tid);
                }
                __syncthreads();
                if (g >= 1) {
                    acc_from(acc, B3, wm, wn, l16, lq);
                    mfma_nt<DRA>(acc, B0, B2, -1.0, wm, wn, l16, lq);
                    __syncthreads();
                    acc_to(acc, B3, wm, wn, l16, lq);
                    __syncthreads();
                }
                acc_zero(acc);
                mfma_nt_lower<DRA>(acc, B3, B1, wm, wn, l16, lq);   // L_ig = A_ig X_g^T (X lower)
                acc_store(acc, A, lda, NB * i, NB * g, rows, NB, wm, wn, l16, lq);
                dag_release(ldone + i * T + g, 1, tid);
                if (tid == 0) dag_trace(a, 1, i, g, g, w0, w1, wall_clock64());
            } else {
                // UPDATE(i, j, p): of generation g (p = g - 1), or promoted into it (p = g)
                const int i = tk.i, j = tk.j, p = tk.p;
                const int ri = min(NB, n - NB * i), rj = min(NB, n - NB * j);
                if (!dag_acquire(a, fl, info, ldone + i * T + p, 1, j != i ? ldone + j * T + p : nullptr, 1,
                                 p >= 1 ? cnt + i * T + j : nullptr, p, &s_ok, tid))
                    break;
                const long long w1 = wall_clock64();
                dag_load(B3, A, lda, NB * i, NB * j, ri, rj, tid);
                dag_load(B0, A, lda, NB * i, NB * p, ri, NB, tid);
                if (j != i) dag_load(B2, A, lda, NB * j, NB * p, rj, NB, tid);
                __syncthreads();
                acc_from(acc, B3, wm, wn, l16, lq);
                mfma_nt<DRA>(acc, B0, j != i ? B2 : B0, -1.0, wm, wn, l16, lq);
                acc_store(acc, A, lda, NB * i, NB * j, ri, rj, wm, wn, l16, lq);
                dag_release(cnt + i * T + j, p + 1, tid);
                if (tid == 0) dag_trace(a, 2, i, j, g, w0, w1, wall_clock64());
            }
        }
    }
    // leave; the last workgroup of the matrix clears its flags
    __syncthreads();
    if (tid == 0) {
        const int old = __hip_atomic_fetch_add(fl + FL_EXIT, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = old == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (s_ok) {
        const int ne = dag_flag_elems(T);
        for (int e = tid; e < ne; e += DNT) __hip_atomic_store(fl + e, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

int g_chol_dag = 1;         // 1: diagonal-block Cholesky as one task-graph launch (where flags are given)
int g_chol_dag_wgs = 256;   // its workgroups per launch (all matrices together)
int g_chol_dag_min = 64;    // ... used only when each matrix gets at least this many
int g_chol_dag_x = 1;       // 1: the task graph also forms L^-1 (where the caller gives a buffer)
// claim order: the inverse's tasks two generations late and the chain's own updates one
// early (n = 2865 Cholesky + L^-1 1498 -> 1453 us, c2 FP64 step -1 %; same bits)
int g_chol_dag_xlag = 2;    // generations the inverse's tasks lag in the claim order
int g_chol_dag_promo = 1;   // 1: the chain's two updates per step are claimed a generation early

size_t potrf_dag_flag_elems(int n) { return (size_t)dag_flag_elems((n + NB - 1) / NB); }

int potrf_dag_order(int T, bool with_x, int xlag, bool promo, int* out, int cap)
{
    const int n = dag_tasks(T, with_x, xlag, promo);
    for (int t = 0; t < n && t < cap; t++) {
        const DagTask k = dag_decode(T, with_x, xlag, promo, t);
        const int v[5] = {k.kind, k.i, k.j, k.p, k.g};
        for (int q = 0; q < 5; q++) out[5 * t + q] = v[q];
    }
    return n;
}
size_t potrf_dag_xbuf_elems(int n)
{
    const size_t T = (size_t)(n + NB - 1) / NB;
    return T * NB * NB * std::max<size_t>(T - 1, 1);   // 64T rows x the tile columns below the last
}

// After the task-graph launch with an X buffer: X = L^-1 into the n x n block at A, tile
// (ti, tj) of a workgroup: below the diagonal from xbuf, on it the tile inverse, above zeros.
__global__ __launch_bounds__(256) void dag_finish_kernel(double* A, int lda, long long sA, int n, int T,
                                                         const double* xdiag, long long sX, const double* xbuf,
                                                         int ldx, long long sXb)
{
    const long long mb = blockIdx.y;
    A += mb * sA;
    xdiag += mb * sX;
    xbuf += mb * sXb;
    const int ti = blockIdx.x % T, tj = blockIdx.x / T;
    const int r0 = ti * NB, c0 = tj * NB, rb = min(NB, n - r0), cb = min(NB, n - c0);
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int r = e & (NB - 1), c = e >> 6;
        if (r >= rb || c >= cb) continue;
        double v = 0.0;
        if (ti == tj) v = xdiag[(long long)ti * NB * NB + c * NB + r];
        else if (ti > tj) v = xbuf[(long long)(c0 + c) * ldx + (r0 + r)];
        A[(long long)(c0 + c) * lda + (r0 + r)] = v;
    }
}

// Split copies of X = L^-1 (the n x n block at A, zeros above the diagonal) for the mixed
// path's products: xr element (r, k) = X[r][k], xc element (r, k) = X[k][r] (xc.p null:
// none), k < ldk (zeros beyond n).  grid: (row blocks of 64 x k blocks of 64, count).
__global__ __launch_bounds__(256) void xsplit_kernel(const double* A, int lda, long long sA, int n, SplitView xr,
                                                     SplitView xc)
{
    const long long e = blockIdx.y;
    A += e * sA;
    const int nkb = xr.ldk / NB;
    const int r0 = (blockIdx.x / nkb) * NB, k0 = (blockIdx.x % nkb) * NB;
    for (int t = threadIdx.x; t < NB * NB; t += 256) {
        const int a = t & (NB - 1), b = t >> 6;
        {   // xr: lanes along r (column reads of A coalesced)
            const int r = r0 + a, k = k0 + b;
            if (r < n) split_store_v(xr, e, 0, r, k, (k < n && k <= r) ? A[(long long)k * lda + r] : 0.0);
        }
        if (xc.p) {   // xc: lanes along k
            const int r = r0 + b, k = k0 + a;
            if (r < n) split_store_v(xc, e, 0, r, k, (k < n && k >= r) ? A[(long long)r * lda + k] : 0.0);
        }
    }
}

int xsplit_copies(const double* A, int lda, long long sA, int n, int count, const SplitView& xr, const SplitView& xc,
                  hipStream_t s)
{
    const int nrb = (n + NB - 1) / NB, nkb = xr.ldk / NB;
    hipLaunchKernelGGL(xsplit_kernel, dim3(nrb * nkb, count), dim3(256), 0, s, A, lda, sA, n, xr, xc);
    return (int)hipGetLastError();
}

static long long dag_timeout_ticks()
{
    static long long ticks[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (!ticks[dev]) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
        ticks[dev] = (long long)khz * 1000LL * 10;   // 10 s for one dependency
    }
    return ticks[dev];
}

// Blocked right-looking Cholesky of the n x n block at A (lower), 64-wide column blocks
// with the trailing matrix updated every second block by a rank-128 GEMM: column block
// p is one launch of the tile kernel (for odd p first the update of the block by panel
// p-1, then factor + inverse of the diagonal tile, then the rows below as an in-place
// product with X_pp^T spread over the launch's workgroups); after an odd block the
// columns beyond it are updated by panels p-1 and p together (K = 128: the same traffic
// on the trailing matrix as one rank-64 update, half the launches).  xdiag receives the
// T inverted diagonal tiles.
long long* g_chol_dag_trace = nullptr;   // tracing of the task-graph launch (tuning)
long long g_chol_dag_trace_cap = 0;

bool potrf_uses_dag(int count) { return g_chol_dag && g_chol_dag_wgs / std::max(count, 1) >= std::max(2, g_chol_dag_min); }

bool potrf_dag_schedule(int schedule, int count)
{
    return schedule == 1 || (schedule == 0 && potrf_uses_dag(count));
}

int potrf_f64(double* A, int lda, long long sA, int n, int count, double* logdet_parts, long long sP, int* info,
              int info_base, double* xdiag, long long sX, bool store_diag_l, hipStream_t s, int* dag_flags,
              double* xbuf, long long sXb, bool* x_done, int schedule)
{
    if (x_done) *x_done = false;
    if (dag_flags && !store_diag_l && potrf_dag_schedule(schedule, count)) {
        const int T = (n + NB - 1) / NB;
        const bool with_x = xbuf && x_done && g_chol_dag_x;
        DagArgs a{A, sA, xdiag, sX, logdet_parts, sP, info, dag_flags, (long long)dag_flag_elems(T),
                  dag_timeout_ticks(), lda, n, T, dag_tasks(T, with_x, g_chol_dag_xlag, g_chol_dag_promo != 0),
                  info_base, g_chol_dag_trace, g_chol_dag_trace_cap, with_x ? xbuf : nullptr, sXb, T * NB,
                  g_chol_dag_xlag, g_chol_dag_promo};
        // workers: enough to finish the tasks in the chain's time -- a task takes about a quarter of
        // a chain step (a 64 x 64 x 64 MFMA product and three tile loads against a tile factorization),
        // so about ntasks / (4 T) workgroups; the early generations hold most tasks and the chain
        // waits on them unless there are three times that (measured at ns = 2865); more only spin
        const int want = 1 + (int)std::min<long long>((long long)a.ntasks, (3LL * a.ntasks + 4 * T - 1) / (4LL * T));
        const int G = std::min(std::max(2, g_chol_dag_wgs / std::max(count, 1)), std::max(2, want));
        hipLaunchKernelGGL(potrf_dag_kernel, dim3(G, count), dim3(DNT), 0, s, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || !with_x) return (int)e;
        hipLaunchKernelGGL(dag_finish_kernel, dim3(T * T, count), dim3(256), 0, s, A, lda, sA, n, T, xdiag, sX, xbuf,
                           T * NB, sXb);
        e = hipGetLastError();
        if (e == hipSuccess) *x_done = true;
        return (int)e;
    }
    int p = 0;
    for (int k0 = 0; k0 < n; k0 += NB, p++) {
        const int kb = min(NB, n - k0);
        const int below = n - k0 - kb;
        const bool odd = p & 1;
        double* xt = xdiag + (long long)p * NB * NB;

        // every workgroup of a matrix factors the tile (in LDS) and then produces its
        // share of the RB-row blocks of the panel below
        // (with store_diag_l the lead workgroup writes L_kk over the tile that the other
        // workgroups of the launch read at their start: one workgroup per matrix then)
        const int nrb = (below + RB - 1) / RB;
        const int nf = store_diag_l ? 1 : std::max(1, std::min(nrb, std::max(1, 256 / count)));
        hipLaunchKernelGGL(potrf_tile_kernel, dim3(nf, count), dim3(256), 0, s, A, lda, sA, k0, kb, below, odd, xt,
                           sX, logdet_parts + p, sP, info, info_base, store_diag_l);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
        // after an odd block: rank-128 update by panels p-1, p; after an even block the
        // next (odd) block applies panel p to itself
        const bool update = below > 0 && odd;
        if (update) {
            const int c0 = odd ? k0 - NB : k0, kk = (odd ? NB : 0) + kb;   // panels [c0, c0 + kk)
            const int t0 = k0 + kb;
            const double* Pn = A + (long long)c0 * lda + t0;
            double* Cn = A + (long long)t0 * lda + t0;
            GemmDesc d = gdesc(Pn, lda, Pn, lda, Cn, lda, n - t0, n - t0, kk, -1.0, 1.0, G_CLOWER);
            batch2(d, count, sA, sA, sA);
            int rc = gemm_f64(Op::N, Op::T, &d, 1, s);
            if (rc) return rc;
        }
    }
    return 0;
}

// one launch of the tile kernel on the leading tile (tuning only)
int potrf_tile_only(double* A, int lda, double* xdiag, double* logdet_slot, int* info, hipStream_t s)
{
    hipLaunchKernelGGL(potrf_tile_kernel, dim3(1, 1), dim3(256), 0, s, A, lda, 0LL, 0, NB, 0, false, xdiag, 0LL,
                       logdet_slot, 0LL, info, 0, false);
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

// The 128-row diagonal blocks of X (complete once the 64-row merges are done) into both split
// copies, with their zeros above the diagonal and beyond n, and the k padding [n, ldk) of
// their rows.  grid: (ceil(n / 128), count); lanes along k.
__global__ __launch_bounds__(256) void trtri_copy_kernel(const double* A, int lda, long long sA, int n, SplitView xr,
                                                         SplitView xc)
{
    const long long e = blockIdx.y;
    A += e * sA;
    const int r0 = 128 * blockIdx.x;
    for (int t = threadIdx.x; t < 128 * 128; t += 256) {
        const int a = t & 127, b = t >> 7;
        const int r = r0 + b, k = r0 + a;   // xr: X[r][k]; xc: X[k][r]
        if (r < n && k < xr.ldk) {
            split_store_v(xr, e, 0, r, k, (k < n && r >= k) ? A[(long long)k * lda + r] : 0.0);
            split_store_v(xc, e, 0, r, k, (k < n && k >= r) ? A[(long long)r * lda + k] : 0.0);
        }
    }
    const int pad = xr.ldk - n;   // both copies have the pitch roundup(n, 64)
    for (int t = threadIdx.x; t < 128 * pad; t += 256) {
        const int r = r0 + t / pad, k = n + t % pad;
        if (r < n && k >= r0 + 128) {   // (inside the block: written above)
            split_store_v(xr, e, 0, r, k, 0.0);
            split_store_v(xc, e, 0, r, k, 0.0);
        }
    }
}

size_t trtri_scratch_elems(int n, int count)
{
    // per level: the split copies of L21 (n2 x n1) and of T (n1 x n2) for every group
    const int T = (n + NB - 1) / NB;
    size_t need = 0;
    for (int sz = 2; sz < T; sz *= 2) {
        const int n1 = sz * NB, G = (T - sz + 2 * sz - 1) / (2 * sz);
        need = std::max(need, 2 * (size_t)G * count * split_entry_elems(n1, n1));
    }
    return need;
}

int trtri_f64_inplace(double* A, int lda, long long sA, int n, int count, double* tmp, long long sT,
                      const double* xdiag, long long sX, hipStream_t s, const ProdFn* prod, long long plane,
                      const TrtriCopies* cp)
{
    auto run = [&](GemmDesc* d, int nd) {
        return prod ? prod->fn(prod->ctx, Op::N, Op::N, d, nd, s) : gemm_f64(Op::N, Op::N, d, nd, s);
    };
    const int T = (n + NB - 1) / NB;
    for (int pl = 0; pl < (plane ? 2 : 1); pl++) {   // the diagonal tiles of each plane
        const long long o = pl * plane;
        hipLaunchKernelGGL(trtri_diag_kernel, dim3(T * T, count), dim3(256), 0, s, A + o, lda, sA, n, T,
                           xdiag ? xdiag + o : nullptr, sX);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    bool copied = false;
    auto copy_blocks = [&]() {   // the 128-row diagonal blocks into the copies (once)
        if (!cp || copied) return 0;
        copied = true;
        hipLaunchKernelGGL(trtri_copy_kernel, dim3((n + 127) / 128, count), dim3(256), 0, s, A, lda, sA, n, cp->xr,
                           cp->xc);
        return (int)hipGetLastError();
    };
    for (int sz = 1; sz < T; sz *= 2) {
        const int G = (T - sz + 2 * sz - 1) / (2 * sz);   // groups with a non-empty right half
        const int n1 = sz * NB;
        const long long gstride = (long long)(2 * sz * NB) * (1 + (long long)lda);
        const long long tstride = (long long)n1 * n1;
        const bool split = cp && n1 >= 128;
        if (split) {
            int rc = copy_blocks();
            if (rc) return rc;
        }
        // groups whose right half is complete share one descriptor
        int gfull = 0;
        for (int g = 0; g < G; g++)
            if ((g * 2 * sz + 2 * sz) * NB <= n) gfull = g + 1;
        GemmDesc d1[2], d2[2];
        PresplitCall c1[2], c2[2];
        __bf16* scr = split ? cp->scratch : nullptr;
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
            d1[nd1] = a; d2[nd1] = b;
            if (split) {
                // L21 split here; X11 (columns) and X22 (rows) read from the copies; T's copy
                // written by its product; X21 written into both copies by its product
                const size_t le = (size_t)cnt * count * split_entry_elems(n2, n1);
                const size_t te = (size_t)cnt * count * split_entry_elems(n1, n2);
                if ((size_t)(scr - cp->scratch) + le + te > cp->scratch_elems) return (int)hipErrorInvalidValue;
                __bf16* const Ls = scr;
                __bf16* const Ts = scr + le;
                scr += le + te;
                int rc = split_operand(L21, lda, n2, n1, false, cnt, gstride, count, sA, Ls, s);
                if (rc) return rc;
                const int step = 2 * sz * NB;
                const long long st_r = split_block_offset(cp->xr, step, step);
                const long long st_c = split_block_offset(cp->xc, step, step);
                c1[nd1] = PresplitCall{a, split_view(Ls, n2, n1, cnt),
                                       split_subview(cp->xc, c0, c0, st_c, cp->xc.s1), split_view(Ts, n1, n2, cnt),
                                       SplitView{}};
                c2[nd1] = PresplitCall{b, split_subview(cp->xr, r0, r0, st_r, cp->xr.s1), split_view(Ts, n1, n2, cnt),
                                       split_subview(cp->xc, c0, r0, st_c, cp->xc.s1),
                                       split_subview(cp->xr, r0, c0, st_r, cp->xr.s1)};
            }
            nd1++;
        }
        int rc = split ? gemm_bf16x3_presplit_n(Op::N, Op::N, c1, nd1, s) : run(d1, nd1);
        if (rc) return rc;
        rc = split ? gemm_bf16x3_presplit_n(Op::N, Op::N, c2, nd1, s) : run(d2, nd1);
        if (rc) return rc;
    }
    return copy_blocks();   // (no merge of 128 rows or more: the copies are the diagonal blocks)
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
// With X: the lower triangle of A first receives A + X (the workgroup owning a lower tile
// reads, sums and writes it back before mirroring it; its diagonal comes from the sum).
__global__ __launch_bounds__(256) void mirror_kernel(double* A, int lda, long long sA, int n, int T, double* diag,
                                                     long long sDg, const double* X, int ldx, long long sX,
                                                     SplitOut cp)
{
    A += (long long)blockIdx.y * sA;
    if (diag) diag += (long long)blockIdx.y * sDg;
    if (X) X += (long long)blockIdx.y * sX;
    __shared__ double t[32][33];
    const int bi = blockIdx.x % T, bj = blockIdx.x / T;   // destination tile (bi row, bj col), bi <= bj
    if (bi > bj) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    // source: tile (bj row-block, bi col-block) of the lower triangle
    for (int q = ty; q < 32; q += 8) {
        int r = bj * 32 + tx, c = bi * 32 + q;
        double v = (r < n && c < n) ? A[(long long)c * lda + r] : 0.0;
        if (X && r < n && c < n && r >= c) {
            v += X[(long long)c * ldx + r];
            A[(long long)c * lda + r] = v;
        }
        t[q][tx] = v;   // t[c_local][r_local]
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        int r = bi * 32 + tx, c = bj * 32 + q;   // destination (r, c) = source (c, r)
        if (r < n && c < n && r < c) A[(long long)c * lda + r] = t[tx][q];
        if (diag && r < n && r == c) diag[r] = t[tx][q];
    }
    if (cp.p) {
        // S[bi*32 + a][bj*32 + b] = t[a][b] (bi < bj), t[min][max] on a diagonal tile.  Copy
        // rows bj*32 + q (k = bi*32 + tx) and bi*32 + q (k = bj*32 + tx): lanes along k.
        const long long e = blockIdx.y;
        for (int q = ty; q < 32; q += 8) {
            const int ra = bj * 32 + q, ka = bi * 32 + tx;   // S[ra][ka] = S[ka][ra]
            if (ra < n && ka < n) split_store(cp, e, ra, ka, bi < bj ? t[tx][q] : t[min(tx, q)][max(tx, q)]);
            const int rb = bi * 32 + q, kb = bj * 32 + tx;
            if (bi < bj && rb < n && kb < n) split_store(cp, e, rb, kb, t[q][tx]);
        }
    }
}

int mirror_lower_diag(double* A, int lda, long long sA, int n, double* diag, long long sDg, int count, hipStream_t s,
                      const double* X, int ldx, long long sX, const SplitOut* cp)
{
    if (n <= 0 || count <= 0) return 0;
    const int T = (n + 31) / 32;
    const SplitOut none{};
    hipLaunchKernelGGL(mirror_kernel, dim3(T * T, count), dim3(256), 0, s, A, lda, sA, n, T, diag, sDg, X, ldx, sX,
                       cp ? *cp : none);
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
                                      long long sZ, int zrow0, SplitOut cp)
{
    S += (long long)blockIdx.y * sS;
    Z += (long long)blockIdx.y * sZ;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ns) return;
    for (int d = 0; d < nd; d++) {
        double v = -Z[(long long)c * ldz + zrow0 + d];
        S[(long long)c * lds + ns + d] = v;
        S[(long long)(ns + d) * lds + c] = v;
        if (cp.p) {
            split_store(cp, blockIdx.y, ns + d, c, v);
            split_store(cp, blockIdx.y, c, ns + d, v);
        }
    }
}

int selinv_finish(double* S, int lds, long long sS, int ns, int nd, const double* Z, int ldz, long long sZ,
                  int zrow0, int exact, double* diag_out, long long sDg, int count, hipStream_t s, const double* X,
                  int ldx, long long sX, const SplitOut* cp)
{
    int rc = mirror_lower_diag(S, lds, sS, ns, diag_out, sDg, count, s, X, ldx, sX, cp);
    if (rc) return rc;
    if (exact && nd > 0) {
        const SplitOut none{};
        hipLaunchKernelGGL(selinv_offdiag_kernel, dim3((ns + 255) / 256, count), dim3(256), 0, s, S, lds, sS, ns,
                           nd, Z, ldz, sZ, zrow0, cp ? *cp : none);
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
