// C ABI of librgf_hip.so: handle lifetime, the FP64 block sweep (factorize,
// logdet, solve, selected inversion) and the residual, all device-resident.
//
// Device layout of one matrix (all FP64, column-major, 256-byte aligned arrays):
//   panel t (t = 0..nt-1), ld = ldp = roundup(2ns+nd, 8), ns columns:
//       rows [0, ns)        diagonal block  (A: full symmetric;  factor: L_tt^-1, lower)
//       rows [ns, 2ns)      coupling block A_{t+1,t}            (factor: B_t = L_{t+1,t})
//       rows [2ns, 2ns+nd)  arrow block A_{d,t}                 (factor: D_t = L_{d,t})
//   corner nd x nd (ld nd)  (A: full symmetric; factor: L_dd and X_dd = L_dd^-1)
// i.e. the reference's supernodes (RGF.H:114-138) with a uniform, aligned
// leading dimension, kept resident for the handle's lifetime instead of being
// streamed to host memory after every block (RGF.H:263, 290, 321, 704-728).
#include "rgf_internal.hpp"
#include "rgf_arena.hpp"
#include <cstdlib>
#include "../../include/rgf.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

namespace rgf {

static thread_local std::string g_create_error;

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file,
             line);
    g_create_error = buf;
    return (int)e;
}

static inline int r8(long long v) { return (int)((v + 7) / 8 * 8); }

}  // namespace rgf

using namespace rgf;

struct rgf_handle_s {
    int ns = 0, nt = 0, nd = 0;
    long long n = 0;
    int count = 1;              // matrices of the same block structure held by the handle
    int cur = 0;                // matrix addressed by the single-matrix entry points
    rgf_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t gx_factor = nullptr, gx_selinv = nullptr;   // captured sweeps (opt.graphs)
    long long gx_epoch_factor = -1, gx_epoch_selinv = -1;
    struct PieceGraph {   // a captured piece of a solve, keyed by what it reads and writes
        int kind, mb, nr;
        const void* x;
        const void* b;
        hipGraphExec_t gx;
        long long epoch;
    };
    std::vector<PieceGraph> gx_pieces;   // most recently used last, at most kMaxPieceGraphs
    int ldp = 0;                // panel leading dimension
    long long pstride = 0;      // elements between panels of one matrix
    long long mstride = 0;      // elements between matrices (nt panels)
    int T = 0;                  // 64-wide tiles per diagonal block
    int ldf = 0;                // ld of Ftmp / W / Z / S buffers
    long long cstride = 1;      // nd*nd (corner stride), >= 1
    long long fstride = 0;      // Ftmp / W / Z stride per matrix: ldf * ns
    long long sstride = 0;      // S_below stride per matrix: ldf * (ns + nd)
    long long tstride = 0;      // trtri scratch per matrix
    long long xstride = 0;      // inverted diagonal tiles per matrix: T * 64 * 64
    long long partstride = 0;   // log-diagonal parts per matrix: nt*T + 1
    // device buffers (all x count)
    double* Ap = nullptr;       // resident matrix panels (keep_matrix)
    double* Add = nullptr;      // resident matrix corner
    double* Lp = nullptr;       // factor panels
    double* Ldd = nullptr;      // corner L_dd
    double* Xdd = nullptr;      // corner X_dd = L_dd^-1
    double* Ftmp = nullptr;     // (ns+nd) x ns staging of the rows below a diagonal block
    double* ttmp = nullptr;     // trtri merge scratch
    double* xdiag = nullptr;    // inverted 64 x 64 diagonal tiles of the current block
    double* parts = nullptr;    // per-tile log-diagonal sums: nt*T + 1
    double* cpart = nullptr;    // corner partials nt * nd * nd
    int* dinfo = nullptr;       // first non-positive pivot per matrix (INT_MAX = none)
    int* cflags = nullptr;      // task-graph Cholesky: dependency flags per matrix (zero between launches)
    double* xbuf = nullptr;     // task-graph Cholesky: off-diagonal tiles of L_ii^-1 as they form
    // selected inversion
    double* Sb[2] = {nullptr, nullptr};
    double* W = nullptr;
    double* Z = nullptr;
    double* diag = nullptr;     // count * n
    // solve (one matrix at a time)
    double* Y = nullptr;        // n x 16
    double* X = nullptr;        // n x 16
    double* Bv = nullptr;       // n x 16
    double* R = nullptr;        // n x 16 residual of the refinement
    double* Dc = nullptr;       // n x 16 correction of the refinement
    double* work = nullptr;     // skinny split-k partials
    size_t work_elems = 0;
    int* sctr = nullptr;        // skinny products' arrival counters (SKINNY_CTRS, zero between launches)
    int last_refine_iters = 0;  // refinement steps taken by the last mixed solve
    double last_refine_res = 0.0;
    double* scal = nullptr;     // small device scalars
    __bf16* split = nullptr;    // bf16 hi/lo operand copies of the split-BF16 products
    size_t split_elems = 0;
    __bf16* split_x = nullptr;  // selected inversion: split copy of X_i kept for three products
    __bf16* split_s = nullptr;  // selected inversion: split copy of the Sigma block S_c (op(A) of Z)
    __bf16* lsplit = nullptr;   // split copies of the coupling panels [B_i; D_i] (as op(A) rows):
                                // one per block with keep_splits, else the last one written
    __bf16* xr = nullptr;       // split copy of X_i's rows (the current block's)
    __bf16* xc = nullptr;       // split copies of X_i's columns: one per block with keep_splits
    bool factorized = false;
    double* Wall = nullptr;     // keep_w: W_i of every block (slot i of matrix b at (b nt + i) fstride)
    bool w_valid = false;       // Wall holds the W_i of the current factor (a selected inversion ran)
    bool broken = false;        // in-place factorization failed: A is gone
    bool generated = false;
    rgf_gen_params gen{};
    size_t bytes = 0;
    std::string err;
    std::vector<std::pair<void*, size_t>> allocs;   // cached device blocks (pointer, bytes)
};

namespace {

int set_err(rgf_handle_s* h, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf; else g_create_error = buf;
    return code;
}

int dev_status(rgf_handle_s* h, int hip_rc, const char* where)
{
    if (hip_rc == 0) return RGF_OK;
    return set_err(h, hip_rc == (int)hipErrorOutOfMemory ? RGF_ENOMEM : RGF_EDEVICE, "%s: HIP error %d (%s)",
                   where, hip_rc, hipGetErrorString((hipError_t)hip_rc));
}

// Device memory comes from a per-device arena (rgf_arena.hpp): slabs of at least 1 GiB
// from hipMalloc, sub-allocated best-fit and coalesced on free, so creating and
// destroying handles calls neither hipMalloc nor hipFree once the arena has grown to the
// working set -- hipFree waits for the whole device, which would stall the other requests
// of a serving loop.  Free bytes are kept up to a cap (4 GiB by default, RGF_DEVICE_CACHE_GB
// or rgf_set_device_cache_limit); beyond it wholly free slabs are released, so a library
// embedded in a larger process does not sit on idle device memory after its handles are
// gone (rgf_trim_device_cache releases every wholly free slab at once).
std::atomic<size_t>& cache_cap_ref()
{
    static std::atomic<size_t> cap{[] {
        const char* v = getenv("RGF_DEVICE_CACHE_GB");
        if (v) return (size_t)(std::max(0.0, atof(v)) * (double)(1ull << 30));
        return (size_t)4 << 30;
    }()};
    return cap;
}
size_t cache_cap() { return cache_cap_ref().load(std::memory_order_relaxed); }
struct HipSlabs {
    static int grow(void** p, size_t n)
    {
        const hipError_t e = hipMalloc(p, n);
        if (e != hipSuccess) (void)hipGetLastError();
        return (int)e;
    }
    static bool oom(int rc) { return rc == (int)hipErrorOutOfMemory; }
    static void release(void* p) { hipFree(p); }
};
Arena<HipSlabs> g_arena[64];

hipError_t cache_alloc(int dev, size_t bytes, void** out, size_t* got)
{
    if (dev < 0 || dev >= 64) {
        *got = bytes;
        return hipMalloc(out, bytes);
    }
    return (hipError_t)g_arena[dev].alloc(bytes, out, got);
}

// p must no longer be in use by any stream.
void cache_free(int dev, void* p, size_t bytes)
{
    (void)bytes;
    if (!p) return;
    if (dev < 0 || dev >= 64 || !g_arena[dev].free(p, cache_cap())) hipFree(p);
}

template <class T>
int dalloc(rgf_handle_s* h, T** p, size_t count)
{
    void* q = nullptr;
    size_t got = 0;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = cache_alloc(h->device, bytes, &q, &got);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_err(h, RGF_ENOMEM, "device allocation of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
    }
    h->allocs.push_back({q, got});
    h->bytes += got;
    *p = (T*)q;
    return RGF_OK;
}

__global__ void set_int_kernel(int* p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int validate_shape(size_t ns, size_t nt, size_t nd)
{
    if (ns == 0 || nt == 0) return set_err(nullptr, RGF_EINVAL, "ns and nt must be positive (ns=%zu nt=%zu)", ns, nt);
    if (nd > ns) return set_err(nullptr, RGF_EINVAL, "nd=%zu > ns=%zu is not supported (the reference's block stride needs nd <= ns)", nd, ns);
    if (nd > 64) return set_err(nullptr, RGF_EINVAL, "nd=%zu > 64 is not supported", nd);
    if (2 * ns + nd > 65536 || ns * nt + nd > (size_t)INT_MAX / 16)
        return set_err(nullptr, RGF_EINVAL, "matrix too large (ns=%zu nt=%zu nd=%zu)", ns, nt, nd);
    return RGF_OK;
}

// bytes a handle of this shape allocates up front (admission control of the serving loop)
// bf16 elements of the operand copies of the largest split-BF16 product of a sweep
// (every product has m, n, k <= ns + nd)
size_t split_need(int ns, int nd, int count)
{
    // two operands of the largest product, plus the copy of Z = Sb W that the selected
    // inversion's product writes for W^T Z
    return split_scratch_elems(ns + nd, ns + nd, ns + nd, count) + (size_t)count * split_entry_elems(ns + nd, ns + nd);
}
// the selected inversion's copies: X_i (ns x ns) and the Sigma block S_c (ns + nd square)
size_t split_x_need(int ns, int nd, int count)
{
    return (size_t)count * (split_entry_elems(ns, ns) + split_entry_elems(ns + nd, ns + nd));
}
// split copies of the coupling panels [B_i; D_i] (ns + nd rows, k = ns): every block's, or one
size_t lsplit_need(int ns, int nt, int nd, int count, int keep_splits)
{
    return (size_t)(keep_splits ? nt : 1) * count * split_entry_elems(ns + nd, ns);
}
// the copies of X_i = L_ii^-1 the mixed factorization keeps (ns >= 128): rows (one block's)
// and columns (every block's with keep_splits)
size_t xcopy_entry(int ns) { return 2 * (size_t)xcopy_rows(ns) * ((ns + 63) / 64 * 64); }
size_t xcopy_need(int ns, int nt, int count, int keep_splits)
{
    return ns >= 128 ? (size_t)(1 + (keep_splits ? nt : 1)) * count * xcopy_entry(ns) : 0;
}
// the h->split scratch: the largest product's operands (and the copy of Z), or a level of
// the inverse's L21 and T copies
size_t split_scratch_need(int ns, int nd, int count)
{
    return std::max(split_need(ns, nd, count), trtri_scratch_elems(ns, count));
}

// doubles of the skinny products' split-k partials: the solves and the residual (one
// matrix, m <= max(n, 2ns+nd)) and the factor sweep's arrow-row update (count matrices
// of m = k = ns at once, each split as it would be alone)
size_t solve_work_elems(int ns, int nt, int nd, int count)
{
    const long long n = (long long)ns * nt + nd;
    const int mx = (int)std::max<long long>(n, 2LL * ns + nd);
    return std::max(skinny_work_elems(mx, mx, 16), skinny_batch_work_elems(Op::N, ns, ns, count));
}

size_t footprint_bytes(int ns, int nt, int nd, int count, int keep, int precision, int keep_splits, int cholesky,
                       int keep_w)
{
    const size_t ldp = (size_t)r8(2LL * ns + nd), ldf = (size_t)r8((long long)ns + nd);
    const size_t T = (size_t)(ns + 63) / 64;
    size_t tt = 1;
    for (size_t sz = 1; sz < T; sz *= 2) {
        size_t G = (T - sz + 2 * sz - 1) / (2 * sz);
        tt = std::max(tt, G * (sz * 64) * (sz * 64));
    }
    const size_t cnd = (size_t)std::max(nd * nd, 1);
    size_t per = ldp * ns * nt * (keep ? 2 : 1) + cnd * (keep ? 3 : 2) + ldf * ns + tt + T * 4096 + nt * T + 1 +
                 (size_t)std::max(nt * nd * nd, 1);
    per += 2 * ldf * (ns + nd) + 2 * ldf * ns + (size_t)ns * nt + nd;   // selected inversion
    if (keep_w && precision == RGF_FP64) per += ldf * ns * nt;              // W_i of every block
    // solve buffers (Y, X, Bv, R, Dc: n x 16 each) and the skinny products' partials
    size_t once = 5 * ((size_t)ns * nt + nd) * 16 + solve_work_elems(ns, nt, nd, count) + SKINNY_CTRS / 2;
    size_t bytes = (per * count + once) * sizeof(double) + potrf_dag_flag_elems(ns) * count * sizeof(int);
    if (precision == RGF_FP64 && potrf_dag_schedule(cholesky, count))
        bytes += potrf_dag_xbuf_elems(ns) * count * sizeof(double);
    if (precision == RGF_MIXED_BF16X3)
        bytes += (split_scratch_need(ns, nd, count) + split_x_need(ns, nd, count) +
                  lsplit_need(ns, nt, nd, count, keep_splits) + xcopy_need(ns, nt, count, keep_splits)) * 2;
    return bytes;
}

int alloc_core(rgf_handle_s* h)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, B = h->count;
    h->ldp = r8(2LL * ns + nd);
    h->pstride = (long long)h->ldp * ns;
    h->mstride = h->pstride * nt;
    h->T = (ns + 63) / 64;
    h->ldf = r8((long long)ns + nd);
    h->cstride = std::max(nd * nd, 1);
    h->fstride = (long long)h->ldf * ns;
    h->sstride = (long long)h->ldf * (ns + nd);
    h->xstride = (long long)h->T * 64 * 64;
    h->partstride = (long long)nt * h->T + 1;
    size_t tt = 1;
    for (int sz = 1; sz < h->T; sz *= 2) {
        size_t G = (size_t)((h->T - sz + 2 * sz - 1) / (2 * sz));
        tt = std::max(tt, G * (size_t)(sz * 64) * (size_t)(sz * 64));
    }
    h->tstride = (long long)tt;
    int rc;
    if (h->opt.keep_matrix) {
        if ((rc = dalloc(h, &h->Ap, (size_t)h->mstride * B))) return rc;
        if ((rc = dalloc(h, &h->Add, (size_t)h->cstride * B))) return rc;
    }
    if ((rc = dalloc(h, &h->Lp, (size_t)h->mstride * B))) return rc;
    if ((rc = dalloc(h, &h->Ldd, (size_t)h->cstride * B))) return rc;
    if ((rc = dalloc(h, &h->Xdd, (size_t)h->cstride * B))) return rc;
    if ((rc = dalloc(h, &h->Ftmp, (size_t)h->fstride * B))) return rc;
    if ((rc = dalloc(h, &h->ttmp, (size_t)h->tstride * B))) return rc;
    if ((rc = dalloc(h, &h->xdiag, (size_t)h->xstride * B))) return rc;
    if ((rc = dalloc(h, &h->parts, (size_t)h->partstride * B))) return rc;
    if ((rc = dalloc(h, &h->cpart, (size_t)std::max(nt * nd * nd, 1) * B))) return rc;
    if ((rc = dalloc(h, &h->dinfo, (size_t)B))) return rc;
    if ((rc = dalloc(h, &h->cflags, potrf_dag_flag_elems(ns) * B))) return rc;
    if ((rc = (int)hipMemsetAsync(h->cflags, 0, potrf_dag_flag_elems(ns) * B * sizeof(int), h->stream)))
        return dev_status(h, rc, "create");
    // L_ii^-1 in the same launch on the FP64 path (the mixed path's inverse merges are split-BF16
    // products, cheaper than the launch's FP64 tiles)
    if (h->opt.precision == RGF_FP64 && potrf_dag_schedule(h->opt.cholesky, B) &&
        (rc = dalloc(h, &h->xbuf, potrf_dag_xbuf_elems(ns) * B)))
        return rc;
    if ((rc = dalloc(h, &h->scal, 64))) return rc;
    if (h->opt.precision == RGF_MIXED_BF16X3) {
        h->split_elems = split_scratch_need(ns, nd, B);
        if ((rc = dalloc(h, &h->split, h->split_elems))) return rc;
        if ((rc = dalloc(h, &h->lsplit, lsplit_need(ns, nt, nd, B, h->opt.keep_splits)))) return rc;
        if (ns >= 128) {
            if ((rc = dalloc(h, &h->xr, (size_t)B * xcopy_entry(ns)))) return rc;
            if ((rc = dalloc(h, &h->xc, (size_t)(h->opt.keep_splits ? nt : 1) * B * xcopy_entry(ns)))) return rc;
        }
    }
    return RGF_OK;
}

int ensure_solve_buffers(rgf_handle_s* h)
{
    if (h->Y) return RGF_OK;
    int rc;
    const size_t n16 = (size_t)h->n * 16;
    if ((rc = dalloc(h, &h->Y, n16))) return rc;
    if ((rc = dalloc(h, &h->X, n16))) return rc;
    if ((rc = dalloc(h, &h->Bv, n16))) return rc;
    if ((rc = dalloc(h, &h->R, n16))) return rc;
    if ((rc = dalloc(h, &h->Dc, n16))) return rc;
    h->work_elems = solve_work_elems(h->ns, h->nt, h->nd, h->count);
    if ((rc = dalloc(h, &h->work, h->work_elems))) return rc;
    if ((rc = dalloc(h, &h->sctr, (size_t)SKINNY_CTRS))) return rc;
    if ((rc = (int)hipMemsetAsync(h->sctr, 0, SKINNY_CTRS * sizeof(int), h->stream))) return dev_status(h, rc, "solve buffers");
    return RGF_OK;
}

// Buffers of a solve: Y (n x 16), the refinement's residual R and correction Dc, Bv (b
// kept for an in-place refined solve), and the skinny products' split-k partials.  A
// handle's own (ensure_solve_buffers) or the no-allocation entry point's static scratch.
struct SolveBufs {
    double* Y;
    double* R;
    double* Dc;
    double* Bv;
    double* work;
    size_t work_elems;
    int* ctr;   // the skinny products' arrival counters (SKINNY_CTRS)
    bool wide = false;   // the MFMA skinny kernel for every pass (a solve of three or more columns)
};

SolveBufs handle_bufs(rgf_handle_s* h) { return {h->Y, h->R, h->Dc, h->Bv, h->work, h->work_elems, h->sctr}; }

int ensure_selinv_buffers(rgf_handle_s* h)
{
    if (h->W) return RGF_OK;
    int rc;
    const int B = h->count;
    if ((rc = dalloc(h, &h->Sb[0], (size_t)h->sstride * B))) return rc;
    if ((rc = dalloc(h, &h->Sb[1], (size_t)h->sstride * B))) return rc;
    if ((rc = dalloc(h, &h->W, (size_t)h->fstride * B))) return rc;
    if ((rc = dalloc(h, &h->Z, (size_t)h->fstride * B))) return rc;
    if ((rc = dalloc(h, &h->diag, (size_t)h->n * B))) return rc;
    // (FP64 only: through the split-BF16 W_i of the mixed path a refined solve took two
    // refinement steps instead of one, c2 1 RHS 21.0 -> 27.2 ms)
    if (h->opt.keep_w && h->opt.precision == RGF_FP64 && (rc = dalloc(h, &h->Wall, (size_t)h->fstride * h->nt * B)))
        return rc;
    if (h->opt.precision == RGF_MIXED_BF16X3) {
        if ((rc = dalloc(h, &h->split_x, split_x_need(h->ns, h->nd, B)))) return rc;
        h->split_s = h->split_x + (size_t)B * split_entry_elems(h->ns, h->ns);
    }
    return RGF_OK;
}

rgf_handle_s* new_handle(size_t ns, size_t nt, size_t nd, size_t count, const rgf_options* opts, int* rc)
{
    *rc = validate_shape(ns, nt, nd);
    if (*rc) return nullptr;
    if (count < 1 || count > 4096) {
        *rc = set_err(nullptr, RGF_EINVAL, "batch count %zu out of range [1, 4096]", count);
        return nullptr;
    }
    auto* h = new rgf_handle_s();
    rgf_default_options(&h->opt);
    if (opts) h->opt = *opts;
    h->ns = (int)ns; h->nt = (int)nt; h->nd = (int)nd;
    h->count = (int)count;
    h->n = (long long)ns * nt + nd;
    h->device = h->opt.device;
    if (h->opt.precision != RGF_FP64 && h->opt.precision != RGF_MIXED_BF16X3) {
        *rc = set_err(nullptr, RGF_EINVAL, "precision mode %d is not available in this build", h->opt.precision);
        delete h;
        return nullptr;
    }
    if (h->opt.keep_w != 0 && h->opt.keep_w != 1) {
        *rc = set_err(nullptr, RGF_EINVAL, "keep_w must be 0 or 1 (got %d)", h->opt.keep_w);
        delete h;
        return nullptr;
    }
    if (h->opt.cholesky < 0 || h->opt.cholesky > 2) {
        *rc = set_err(nullptr, RGF_EINVAL, "cholesky schedule %d is not one of 0 (auto), 1 (task graph), 2 (tile launches)",
                      h->opt.cholesky);
        delete h;
        return nullptr;
    }
    hipError_t e = hipSetDevice(h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (h->stream) hipStreamDestroy(h->stream);
        *rc = set_err(nullptr, RGF_EDEVICE, "no usable HIP device %d (%s)", h->device, hipGetErrorString(e));
        delete h;
        return nullptr;
    }
    *rc = alloc_core(h);
    if (*rc) {
        g_create_error = h->err;
        rgf_destroy(h);
        return nullptr;
    }
    return h;
}

// pointers into matrix b: the resident matrix (Ap if kept, else the factor panels)
double* a_panel(rgf_handle_s* h, int t, int b = 0)
{
    return (h->Ap ? h->Ap : h->Lp) + (long long)b * h->mstride + (long long)t * h->pstride;
}
double* a_corner(rgf_handle_s* h, int b = 0) { return (h->Add ? h->Add : h->Ldd) + (long long)b * h->cstride; }
double* l_panel(rgf_handle_s* h, int t, int b = 0) { return h->Lp + (long long)b * h->mstride + (long long)t * h->pstride; }

// Products that the mixed path moves to split-BF16 MFMA (Schur and triangular-solve
// updates, the Takahashi products); everything else stays on the FP64 kernel.
int prod(rgf_handle_s* h, Op ta, Op tb, GemmDesc* d, int nd)
{
    if (h->opt.precision != RGF_MIXED_BF16X3) return gemm_f64(ta, tb, d, nd, h->stream);
    // products with both output edges >= 128 split their operands once into bf16
    // copies; thin ones (the arrow rows) split while staging
    GemmDesc rest[GEMM_MAXD];
    int nrest = 0, rc;
    for (int q = 0; q < nd; q++) {
        if (d[q].m >= 128 && d[q].n >= 128) {
            if ((rc = gemm_bf16x3_split(ta, tb, &d[q], h->split, h->split_elems, h->stream))) return rc;
        } else {
            rest[nrest++] = d[q];
        }
    }
    return nrest ? gemm_bf16x3(ta, tb, rest, nrest, h->stream) : 0;
}

bool presplit_ok(rgf_handle_s* h, const GemmDesc& d)
{
    return h->opt.precision == RGF_MIXED_BF16X3 && d.m >= 128 && d.n >= 128;
}

// split op(X) of a product's operand (as gemm_bf16x3_split would) into dst
int split_a(rgf_handle_s* h, Op ta, const GemmDesc& d, __bf16* dst)
{
    return split_operand(d.A, d.lda, d.m, d.k, ta == Op::T, d.count, d.sA, d.count2, d.sA2, dst, h->stream);
}
// A product whose two operands have the same split copy (op(A) = op(B)^T, e.g. the
// Schur update B B^T): split once
int prod_same(rgf_handle_s* h, Op ta, Op tb, GemmDesc* d)
{
    const bool same = d->A == d->B && d->lda == d->ldb && d->m == d->n && (ta == Op::T) == (tb == Op::N) &&
                      d->sA == d->sB && d->sA2 == d->sB2;
    if (!same || !presplit_ok(h, *d)) return prod(h, ta, tb, d, 1);
    int rc = split_a(h, ta, *d, h->split);
    return rc ? rc : gemm_bf16x3_presplit(ta, tb, d, h->split, h->split, h->stream);
}

// The coupling panel [B_j; D_j] (rows ns .. 2ns + nd of panel j) is written by a pre-split
// product on the mixed path when it has at least 128 rows (and ns >= 128); that product
// also writes its split copy as op(A) rows (k = ns) into panel_copy(h, j): read by the
// Schur update of block j + 1 (B_j B_j^T) and, with keep_splits, by the selected
// inversion's W = [B_j; D_j] X_j.
int g_trtri_mixed = 1;   // mixed path: L_ii^-1 merges as split-BF16 products (1) or FP64 (0)
int panel_rows(const rgf_handle_s* h, int j) { return (j < h->nt - 1 ? h->ns : 0) + h->nd; }
bool panel_presplit(const rgf_handle_s* h, int j)
{
    return h->opt.precision == RGF_MIXED_BF16X3 && panel_rows(h, j) >= 128 && h->ns >= 128;
}
__bf16* panel_copy(rgf_handle_s* h, int j)
{
    return h->lsplit + (size_t)(h->opt.keep_splits ? j : 0) * h->count * split_entry_elems(h->ns + h->nd, h->ns);
}
// whole-copy views of X_i's rows / columns (entry = matrix of the handle)
SplitView xcopy_view(const rgf_handle_s* h, __bf16* p)
{
    SplitView v{};
    v.p = p;
    v.ldk = (h->ns + 63) / 64 * 64;
    v.rows = xcopy_rows(h->ns);
    v.lo = (long long)v.rows * v.ldk;
    v.s1 = v.s2 = 2 * v.lo;   // one entry per matrix, whichever level of the descriptor counts them
    v.pad = 1;
    return v;
}
SplitView xr_view(const rgf_handle_s* h) { return xcopy_view(h, h->xr); }
SplitView xc_view(const rgf_handle_s* h, int j)
{
    return xcopy_view(h, h->xc + (size_t)(h->opt.keep_splits ? j : 0) * h->count * xcopy_entry(h->ns));
}

// ------------------------------------------------------------------ factorize
// FirstStageFactor (RGF.H:233-340) with every block resident, over all matrices
// of the handle at once (each launch covers the whole batch):
//   S_i = A_ii - B_{i-1} B_{i-1}^T ; A_di -= D_{i-1} B_{i-1}^T      (one GEMM launch)
//   L_ii = chol(S_i)   (log-diagonal summed in the tile kernel)
//   X_i  = L_ii^-1     (in place)
//   [B_i; D_i] = [A_{i+1,i}; A_di] X_i^T                             (one GEMM launch)
//   S_dd = A_dd - sum_i D_i D_i^T ; L_dd = chol(S_dd) ; X_dd = L_dd^-1
int prod_cb(void* ctx, Op ta, Op tb, GemmDesc* d, int nd, hipStream_t)
{
    return prod((rgf_handle_s*)ctx, ta, tb, d, nd);
}

int factor_sweep(rgf_handle_s* h)
{
    const ProdFn pf{prod_cb, h};   // trtri merges: split-BF16 products on the mixed path
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp, ldf = h->ldf, B = h->count;
    const long long ms = h->mstride, fs = h->fstride;
    hipStream_t s = h->stream;
    int rc;
    hipLaunchKernelGGL(set_int_kernel, dim3((B + 255) / 256), dim3(256), 0, s, h->dinfo, B, INT_MAX);
    if ((rc = (int)hipGetLastError())) return rc;
    // F = [A_{i+1,i}; A_di - D_{i-1} B_{i-1}^T], the operand of [B_i; D_i] = F X_i^T: staged in
    // Ftmp, except where that product splits its operands first (the mixed path's pre-split
    // products): then F is assembled in its own rows of the panel, which the product
    // overwrites only after the split -- no copy of the coupling block
    auto f_in_panel = [&](int i) { return panel_presplit(h, i); };
    for (int i = 0; i < nt; i++) {
        double* P = l_panel(h, i);
        double* A = a_panel(h, i);
        const bool fp = f_in_panel(i);
        double* const Fd = fp ? P + 2 * ns : h->Ftmp + ns;   // arrow rows of F
        const int ldF = fp ? ldp : ldf;
        const long long sF = fp ? ms : fs;
        if (i == 0) {
            if (h->Ap && (rc = copy2d(A, ldp, P, ldp, ns, ns, 1.0, s, B, ms, ms))) return rc;
            if (nd > 0 && Fd != A + 2 * ns && (rc = copy2d(A + 2 * ns, ldp, Fd, ldF, nd, ns, 1.0, s, B, ms, sF)))
                return rc;
        } else {
            const double* Pm = l_panel(h, i - 1);
            GemmDesc d = gdesc(Pm + ns, ldp, Pm + ns, ldp, P, ldp, ns, ns, ns, -1.0, 1.0, G_CLOWER);
            d.Cin = A; d.ldcin = ldp;
            batch2(d, B, ms, ms, ms);
            // arrow rows, A_di - D_{i-1} B_{i-1}^T (nd x ns, k = ns): as a skinny product
            // in transposed form, Ftmp_d^T = A_d^T - B_{i-1} D_{i-1}^T (16 arrow rows per
            // launch); a GEMM launch would give it only ns/64 tiles with a long k loop
            for (int c0 = 0; c0 < nd; c0 += 16) {
                const int nc = std::min(16, nd - c0);
                rc = skinny_f64s(Op::N, 0, ns, ns, nc, -1.0, Pm + ns, ldp, Pm + 2 * ns + c0, ldp, 1, 1.0,
                                 A + 2 * ns + c0, ldp, 1, Fd + c0, ldF, 1, h->work, h->work_elems, s, B,
                                 ms, ms, ms, sF, h->sctr);
                if (rc) return rc;
            }
            if (panel_presplit(h, i - 1)) {   // B_{i-1}: the first ns rows of its panel's copy
                const int r = panel_rows(h, i - 1);
                __bf16* const c = panel_copy(h, i - 1);
                if ((rc = gemm_bf16x3_presplit(Op::N, Op::T, &d, c, c, s, nullptr, nullptr, r, r))) return rc;
            } else if ((rc = prod_same(h, Op::N, Op::T, &d))) {
                return rc;
            }
        }
        // the task-graph Cholesky forms X_i = L_ii^-1 in the same launch (x_done)
        bool x_done = false;
        if ((rc = potrf_f64(P, ldp, ms, ns, B, h->parts + (long long)i * h->T, h->partstride, h->dinfo, i * ns,
                            h->xdiag, h->xstride, false, s, h->cflags, h->xbuf, (long long)potrf_dag_xbuf_elems(ns),
                            h->xbuf ? &x_done : nullptr, h->opt.cholesky)))
            return rc;
        // mixed path: the inverse keeps split copies of X_i's rows (the product below) and
        // columns (the selected inversion's)
        TrtriCopies xcp{xr_view(h), h->xc ? xc_view(h, i) : SplitView{}, h->split, h->split_elems};
        const bool xcopies = h->xr && g_trtri_mixed;
        if (x_done) {
            if (xcopies && (rc = xsplit_copies(P, ldp, ms, ns, B, xcp.xr, xcp.xc, s))) return rc;
        } else if ((rc = trtri_f64_inplace(P, ldp, ms, ns, B, h->ttmp, h->tstride, h->xdiag, h->xstride, s,
                                           h->opt.precision == RGF_MIXED_BF16X3 && g_trtri_mixed ? &pf : nullptr, 0,
                                           xcopies ? &xcp : nullptr))) {
            return rc;
        }
        const bool has_next = i < nt - 1;
        double* const Fc = fp ? P + ns : h->Ftmp;   // coupling rows of F
        if (has_next && Fc != A + ns && (rc = copy2d(A + ns, ldp, Fc, ldF, ns, ns, 1.0, s, B, ms, sF))) return rc;
        const int mF = (has_next ? ns : 0) + nd;
        if (mF > 0) {
            const double* src = has_next ? Fc : Fd;
            double* dst = P + (has_next ? ns : 2 * ns);
            GemmDesc d = gdesc(src, ldF, P, ldp, dst, ldp, mF, ns, ns, 1.0, 0.0, G_BUPPER);
            batch2(d, B, sF, ms, ms);
            if (fp) {   // F and X_i^T split (F before the product overwrites it); the product
                        // writes the split copy of [B_i; D_i]
                __bf16* const As = h->split;
                __bf16* const Bs = h->split + (size_t)B * split_entry_elems(mF, ns);
                if ((rc = split_a(h, Op::N, d, As))) return rc;
                PresplitCall c{d, split_view(As, mF, ns, d.count), split_view(Bs, ns, ns, d.count), SplitView{},
                               split_view(panel_copy(h, i), mF, ns, d.count)};
                if (xcopies) {   // op(B)^T = X_i's rows: the inverse's copy
                    c.b = xr_view(h);
                } else if ((rc = split_operand(d.B, d.ldb, d.n, d.k, false, d.count, d.sB, d.count2, d.sB2, Bs, s))) {
                    return rc;
                }
                if ((rc = gemm_bf16x3_presplit_n(Op::N, Op::T, &c, 1, s))) return rc;
            } else if ((rc = prod(h, Op::N, Op::T, &d, 1))) {
                return rc;
            }
        }
    }
    if (nd > 0) {
        rc = corner_factor(h->Lp + 2 * ns, h->pstride, ms, ldp, ns, nt, nd, a_corner(h), nd, h->Ldd, h->Xdd, h->cpart,
                           h->parts + (long long)nt * h->T, h->partstride, h->dinfo, nt * ns, B, s);
        if (rc) return rc;
    }
    return 0;
}

// Run a sweep on h->stream, through a HIP graph when the handle asks for it: captured
// and instantiated on the first call (and again after the kernel-timing configuration
// changed, since the graph records its events), replayed afterwards.  Every device
// pointer a sweep uses belongs to the handle, so a replay does the same work.
template <class Fn>
int run_captured(rgf_handle_s* h, hipGraphExec_t& gx, long long& epoch, Fn&& fn)
{
    if (!h->opt.graphs) return fn();
    const long long ep = timing_epoch();
    if (gx && epoch != ep) {
        hipGraphExecDestroy(gx);
        gx = nullptr;
    }
    if (!gx) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return (int)e;
        const int rc = fn();
        e = hipStreamEndCapture(h->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return rc ? rc : (int)e;
        }
        e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            gx = nullptr;
            return (int)e;
        }
        epoch = ep;
    }
    return (int)hipGraphLaunch(gx, h->stream);
}

// The same for a piece of a solve, keyed by (kind, matrix, RHS count, the buffers it
// reads and writes): repeated solves on the same buffers replay their graph.
constexpr size_t kMaxPieceGraphs = 8;
template <class Fn>
int run_piece(rgf_handle_s* h, int kind, int mb, int nr, const void* x, const void* b, Fn&& fn)
{
    if (!h->opt.graphs) return fn();
    if (h->w_valid && h->Wall) kind += 100;   // a solve through the kept W_i is another sequence
    auto& v = h->gx_pieces;
    size_t i = 0;
    while (i < v.size() && !(v[i].kind == kind && v[i].mb == mb && v[i].nr == nr && v[i].x == x && v[i].b == b)) i++;
    if (i == v.size()) {
        if (v.size() >= kMaxPieceGraphs) {   // drop the least recently used
            if (v.front().gx) hipGraphExecDestroy(v.front().gx);
            v.erase(v.begin());
        }
        v.push_back({kind, mb, nr, x, b, nullptr, -1});
        i = v.size() - 1;
    } else if (i + 1 != v.size()) {
        std::rotate(v.begin() + i, v.begin() + i + 1, v.end());
        i = v.size() - 1;
    }
    return run_captured(h, v[i].gx, v[i].epoch, std::forward<Fn>(fn));
}

int check_factor_info(rgf_handle_s* h);
int do_factorize(rgf_handle_s* h)
{
    if (!h->Ap && h->factorized) return RGF_OK;   // in place: A is gone, the factor is current
    if (h->broken) return set_err(h, RGF_ESTATE, "factorization failed earlier and the matrix was not kept");
    hipSetDevice(h->device);
    int rc;
    if (h->nd > 0 && h->nt > 1 && (rc = ensure_solve_buffers(h))) return rc;   // arrow products use `work`
    h->w_valid = false;   // the kept W_i belong to the factor being replaced
    rc = run_captured(h, h->gx_factor, h->gx_epoch_factor, [h] { return factor_sweep(h); });
    if (rc) return dev_status(h, rc, "factorize");
    return check_factor_info(h);
}

// After a factor sweep on h->stream: the first non-positive pivot of each matrix (waits
// for the sweep); marks the handle factorized or failed.
int check_factor_info(rgf_handle_s* h)
{
    std::vector<int> info(h->count);
    hipError_t e = hipMemcpyAsync(info.data(), h->dinfo, sizeof(int) * h->count, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return dev_status(h, (int)e, "factorize");
    for (int b = 0; b < h->count; b++)
        if (info[b] == INT_MIN) {   // a wait of the task-graph Cholesky expired
            h->factorized = false;
            if (!h->Ap) h->broken = true;
            return set_err(h, RGF_EDEVICE, "factorize: a dependency wait of the Cholesky launch expired (matrix %d)",
                           b);
        } else if (info[b] != INT_MAX) {
            h->factorized = false;
            if (!h->Ap) h->broken = true;
            if (h->count == 1)
                return set_err(h, RGF_ENOTSPD, "matrix is not positive definite: non-positive pivot at column %d",
                               info[b]);
            return set_err(h, RGF_ENOTSPD, "matrix %d of the batch is not positive definite: non-positive pivot at column %d",
                           b, info[b]);
        }
    h->factorized = true;
    return RGF_OK;
}

int ensure_factorized(rgf_handle_s* h)
{
    if (h->factorized) return RGF_OK;
    if (h->broken) return set_err(h, RGF_ESTATE, "factorization failed earlier and the matrix was not kept");
    return do_factorize(h);
}

// ------------------------------------------------------------------ solve
// SecondStageSolve (RGF.H:344-418) as products with the stored inverses, for matrix b:
//   forward   y_i = X_i (b_i - B_{i-1} y_{i-1}),  y_d = X_dd (b_d - sum D_i y_i)
//   backward  x_d = X_dd^T y_d,  x_i = X_i^T (y_i - B_i^T x_{i+1} - D_i^T x_d)
// x, b: device, n x nr (ld n), nr <= 16; Y is scratch.
int solve_block(rgf_handle_s* h, const SolveBufs& sb, int mb, double* x, const double* b, int nr)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp;
    const int n = (int)h->n;
    hipStream_t s = h->stream;
    double* Y = sb.Y;
    double* w = sb.work;
    const size_t we = sb.work_elems;
    int* const ct = sb.ctr;
    const Op N = Op::N, T = Op::T;
    const double* Xdd = h->Xdd + (long long)mb * h->cstride;
    int rc;
    if (x != b && (rc = copy2d(b, n, x, n, n, nr, 1.0, s))) return rc;   // x holds r during the forward pass
    double* xd = x + (long long)nt * ns;
    double* yd = Y + (long long)nt * ns;
    if (h->w_valid && h->Wall) {
        // With the selected inversion's W_i = [B_i; D_i] X_i kept: r_i = b_i - B_{i-1} y_{i-1} =
        // b_i - W_{i-1} r_{i-1} (and x_d -= W_i's arrow rows r_i) is a chain of one product per
        // block; y_i = X_i r_i for all blocks in one launch; after the corner, x_i = X_i^T y_i for
        // all blocks in one launch, then x_i -= W_i^T [x_{i+1}; x_d] block by block
        const long long fs = h->fstride;
        const double* const W0 = h->Wall + (long long)mb * nt * fs;
        const int ldw = h->ldf;
        const long long ps = h->pstride;
        const double* const P0 = l_panel(h, 0, mb);
        for (int i = 1; i < nt; i++) {
            double* ri = x + (long long)i * ns;
            if ((rc = skinny_f64s(N, 0, ns + nd, ns, nr, -1.0, W0 + (long long)(i - 1) * fs, ldw,
                                  x + (long long)(i - 1) * ns, 1, n, 1.0, ri, 1, n, ri, 1, n, w, we, s, 1, 0, 0, 0, 0,
                                  ct, nullptr, 0, nd ? xd : nullptr, ns, 64, sb.wide)))
                return rc;
        }
        if (nd > 0 && (rc = skinny_f64(N, 0, nd, ns, nr, -1.0, W0 + (long long)(nt - 1) * fs, ldw,
                                       x + (long long)(nt - 1) * ns, n, 1.0, xd, n, xd, n, w, s, we, ct, sb.wide)))
            return rc;
        if ((rc = skinny_f64s(N, 1, ns, ns, nr, 1.0, P0, ldp, x, 1, n, 0.0, Y, 1, n, Y, 1, n, w, we, s, nt, ps, ns,
                              ns, ns, ct, nullptr, 0, nullptr, 0, 1, sb.wide)))
            return rc;
        if (nd > 0) {
            if ((rc = skinny_f64(N, 1, nd, nd, nr, 1.0, Xdd, nd, xd, n, 0.0, yd, n, yd, n, w, s, we, ct, sb.wide)))
                return rc;
            if ((rc = skinny_f64(T, 2, nd, nd, nr, 1.0, Xdd, nd, yd, n, 0.0, xd, n, xd, n, w, s, we, ct, sb.wide)))
                return rc;
        }
        if ((rc = skinny_f64s(T, 2, ns, ns, nr, 1.0, P0, ldp, Y, 1, n, 0.0, x, 1, n, x, 1, n, w, we, s, nt, ps, ns,
                              ns, ns, ct, nullptr, 0, nullptr, 0, 1, sb.wide)))
            return rc;
        double* const xl = x + (long long)(nt - 1) * ns;
        if (nd > 0 && (rc = skinny_f64(T, 0, ns, nd, nr, -1.0, W0 + (long long)(nt - 1) * fs, ldw, xd, n, 1.0, xl,
                                       n, xl, n, w, s, we, ct, sb.wide)))
            return rc;
        for (int i = nt - 2; i >= 0; i--) {
            double* xi = x + (long long)i * ns;
            if ((rc = skinny_f64s(T, 0, ns, ns + nd, nr, -1.0, W0 + (long long)i * fs, ldw, x + (long long)(i + 1) * ns,
                                  1, n, 1.0, xi, 1, n, xi, 1, n, w, we, s, 1, 0, 0, 0, 0, ct, nd ? xd : nullptr, ns,
                                  nullptr, 0, 64, sb.wide)))
                return rc;
        }
        return 0;
    }
    // forward: [r_i; x_d] -= [B_{i-1}; D_{i-1}] y_{i-1} (the coupling and arrow rows of panel
    // i - 1 are contiguous: one product, its rows past ns accumulate into x_d), y_i = X_i r_i
    for (int i = 0; i < nt; i++) {
        double* P = l_panel(h, i, mb);
        double* ri = x + (long long)i * ns;
        double* yi = Y + (long long)i * ns;
        if (i > 0) {
            const double* Pm = l_panel(h, i - 1, mb);
            if ((rc = skinny_f64s(N, 0, ns + nd, ns, nr, -1.0, Pm + ns, ldp, Y + (long long)(i - 1) * ns, 1, n, 1.0, ri,
                                  1, n, ri, 1, n, w, we, s, 1, 0, 0, 0, 0, ct, nullptr, 0, nd ? xd : nullptr, ns, 64,
                                  sb.wide)))
                return rc;
        }
        if ((rc = skinny_f64(N, 1, ns, ns, nr, 1.0, P, ldp, ri, n, 0.0, yi, n, yi, n, w, s, we, ct, sb.wide)))
            return rc;
    }
    if (nd > 0) {
        const double* P = l_panel(h, nt - 1, mb);   // x_d -= D_{nt-1} y_{nt-1}
        if ((rc = skinny_f64(N, 0, nd, ns, nr, -1.0, P + 2 * ns, ldp, Y + (long long)(nt - 1) * ns, n, 1.0, xd, n, xd,
                             n, w, s, we, ct, sb.wide)))
            return rc;
        if ((rc = skinny_f64(N, 1, nd, nd, nr, 1.0, Xdd, nd, xd, n, 0.0, yd, n, yd, n, w, s, we, ct, sb.wide)))
            return rc;
        if ((rc = skinny_f64(T, 2, nd, nd, nr, 1.0, Xdd, nd, yd, n, 0.0, xd, n, xd, n, w, s, we, ct, sb.wide)))
            return rc;
    }
    // backward: y_i -= [B_i; D_i]^T [x_{i+1}; x_d] (one product over both row groups of the
    // panel), x_i = X_i^T y_i
    for (int i = nt - 1; i >= 0; i--) {
        double* P = l_panel(h, i, mb);
        double* yi = Y + (long long)i * ns;
        double* xi = x + (long long)i * ns;
        if (i < nt - 1) {
            if ((rc = skinny_f64s(T, 0, ns, ns + nd, nr, -1.0, P + ns, ldp, x + (long long)(i + 1) * ns, 1, n, 1.0,
                                  yi, 1, n, yi, 1, n, w, we, s, 1, 0, 0, 0, 0, ct, nd ? xd : nullptr, ns, nullptr, 0,
                                  64, sb.wide)))
                return rc;
        } else if (nd > 0 && (rc = skinny_f64(T, 0, ns, nd, nr, -1.0, P + 2 * ns, ldp, xd, n, 1.0, yi, n, yi, n, w, s,
                                              we, ct, sb.wide))) {
            return rc;
        }
        if ((rc = skinny_f64(T, 2, ns, ns, nr, 1.0, P, ldp, yi, n, 0.0, xi, n, xi, n, w, s, we, ct, sb.wide)))
            return rc;
    }
    return 0;
}

int block_matvec_residual(rgf_handle_s* h, const SolveBufs& sb, int mb, const double* x, const double* b, double* r,
                          int nr);

// column 2-norms of an n x nr array (ld n), one workgroup per column, fixed-order tree
__global__ __launch_bounds__(1024) void colnorm2_kernel(const double* r, long long n, double* out)
{
    __shared__ double red[1024];
    r += (long long)blockIdx.x * n;
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 1024) s += r[i] * r[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(red[0]);
}

__global__ void axpy_kernel(double* y, const double* x, long long total)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x)
        y[i] += x[i];
}

// Solve with the (mixed-precision) factor, then FP64 iterative refinement against
// the resident FP64 matrix: r = b - A x; x += M^-1 r, until max_c ||r_c|| / ||b_c||
// <= refine_tol or refine_iters steps.  b must not alias x.  eager: launch every piece
// directly (no graph capture, which allocates).
int solve_refined(rgf_handle_s* h, const SolveBufs& sb, int mb, double* x, const double* b, int nr, bool eager)
{
    const long long n = h->n;
    hipStream_t s = h->stream;
    auto piece = [&](int kind, const void* px, const void* pb, auto&& fn) {
        return eager ? fn() : run_piece(h, kind, mb, nr, px, pb, fn);
    };
    int rc = piece(0, x, b, [&] { return solve_block(h, sb, mb, x, b, nr); });
    if (rc) return rc;
    h->last_refine_iters = 0;
    h->last_refine_res = 0.0;
    if (!h->Ap || h->opt.refine_iters <= 0) return 0;
    hipLaunchKernelGGL(colnorm2_kernel, dim3(nr), dim3(1024), 0, s, b, n, h->scal + 32);
    for (int it = 0; it <= h->opt.refine_iters; it++) {
        rc = piece(1, x, b, [&] {
            const int r = block_matvec_residual(h, sb, mb, x, b, sb.R, nr);
            if (r) return r;
            hipLaunchKernelGGL(colnorm2_kernel, dim3(nr), dim3(1024), 0, s, sb.R, n, h->scal + 16);
            return (int)hipGetLastError();
        });
        if (rc) return rc;
        double v[32];
        hipError_t e = hipMemcpyAsync(v, h->scal + 16, 32 * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return (int)e;
        double worst = 0.0;
        for (int c = 0; c < nr; c++) worst = std::max(worst, v[16 + c] > 0 ? v[c] / v[16 + c] : v[c]);
        h->last_refine_res = worst;
        if (worst <= h->opt.refine_tol || it == h->opt.refine_iters) break;
        rc = piece(2, x, nullptr, [&] {
            const int r = solve_block(h, sb, mb, sb.Dc, sb.R, nr);
            if (r) return r;
            const long long total = n * nr;
            hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)std::min<long long>((total + 255) / 256, 4096)), dim3(256), 0,
                               s, x, sb.Dc, total);
            return (int)hipGetLastError();
        });
        if (rc) return rc;
        h->last_refine_iters = it + 1;
    }
    return 0;
}

int do_solve_device(rgf_handle_s* h, double* x, const double* b, size_t nrhs)
{
    int rc = ensure_factorized(h);
    if (rc) return rc;
    if ((rc = ensure_solve_buffers(h))) return rc;
    const long long n = h->n;
    const bool refine = h->opt.precision != RGF_FP64;
    SolveBufs sb = handle_bufs(h);
    sb.wide = nrhs > 2;   // every pass on the MFMA kernel: a column's sums do not depend on its pass
    for (size_t c0 = 0; c0 < nrhs; c0 += 16) {
        const int nr = (int)std::min<size_t>(16, nrhs - c0);
        if (refine) {
            const double* bb = b + c0 * n;
            if (bb == x + c0 * n) {   // in place: keep b for the residual
                if ((rc = copy2d(bb, (int)n, sb.Bv, (int)n, (int)n, nr, 1.0, h->stream))) return dev_status(h, rc, "solve");
                bb = sb.Bv;
            }
            rc = solve_refined(h, sb, h->cur, x + c0 * n, bb, nr, false);
        } else {
            double* xc = x + c0 * n;
            const double* bc = b + c0 * n;
            rc = run_piece(h, 0, h->cur, nr, xc, bc, [&] { return solve_block(h, sb, h->cur, xc, bc, nr); });
        }
        if (rc) return dev_status(h, rc, "solve");
    }
    return RGF_OK;
}

// ------------------------------------------------------------------ selected inversion
// ThirdStageRGF (RGF.H:422-506) in the G-formulation with X_i = L_ii^-1 resident,
// over all matrices of the handle at once:
//   W = F_i X_i, F_i = [B_i; D_i]            Z = Sb W   (Sb = [[S_{i+1}, S_{d,i+1}^T], [S_{d,i+1}, S_dd]])
//   S_ii = X_i^T X_i + W^T Z                 S_{d,i} = -Z[D rows]
// exact = 0 keeps Sb block-diagonal, which is the reference's formula.
// diag_out: count x n (matrix b at b*n).
// Blocks of A^-1 on the pattern of A for matrix mb, written as the sweep forms them
// (device buffers, any may be null): Sigma_ii (nt x ns x ns), Sigma_{i+1,i} = -Z[0:ns]
// ((nt-1) x ns x ns), Sigma_{d,i} (nt x nd x ns, formed by the exact recursion) and
// Sigma_dd (nd x nd); column-major blocks.
struct SigmaOut {
    double* diag;
    double* sub;
    double* arrow;
    double* corner;
    int mb;
};

int selinv_sweep(rgf_handle_s* h, double* diag_out, const SigmaOut* so = nullptr)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp, lds = h->ldf, B = h->count;
    const long long ms = h->mstride, ss = h->sstride, fs = h->fstride, n = h->n;
    const int m = ns + nd;
    const int exact = h->opt.selinv_mode == RGF_SELINV_EXACT;
    hipStream_t s = h->stream;
    int rc;
    if ((rc = fill2d(h->Sb[0], lds, m, m, 0.0, s, B, ss))) return rc;
    if ((rc = fill2d(h->Sb[1], lds, m, m, 0.0, s, B, ss))) return rc;
    if (nd > 0) {
        for (int q = 0; q < 2; q++) {
            double* Sdd = h->Sb[q] + (long long)ns * lds + ns;
            GemmDesc d = gdesc(h->Xdd, nd, h->Xdd, nd, Sdd, lds, nd, nd, nd, 1.0, 0.0, 0);
            batch2(d, B, h->cstride, h->cstride, ss);
            if ((rc = gemm_f64(Op::T, Op::N, &d, 1, s))) return rc;
        }
        if ((rc = extract_diag(h->Sb[0] + (long long)ns * lds + ns, lds, nd, diag_out + (long long)nt * ns, s, B, ss, n)))
            return rc;
    }
    // Mixed path: the split copy of S_c, op(A) of Z = S_c W, is written by the finishing
    // kernels of the step before (mirror, off-diagonal), its corner and k padding once here
    // (from Sb[0]: zero but for the corner S_dd, which both S buffers hold throughout)
    const bool sc_copy = h->opt.precision == RGF_MIXED_BF16X3 && m >= 128 && ns >= 128;
    SplitOut scp{};
    if (sc_copy) {
        scp.p = h->split_s;
        scp.stride = (long long)split_entry_elems(m, m);
        scp.lo = scp.stride / 2;
        scp.ldk = (m + 63) / 64 * 64;
        if ((rc = split_operand(h->Sb[0], lds, m, m, false, B, ss, 1, 0, h->split_s, s))) return rc;
    }
    if (nd > 0) {
        if (so && so->corner &&
            (rc = copy2d(h->Sb[0] + so->mb * ss + (long long)ns * lds + ns, lds, so->corner, nd, nd, nd, 1.0, s)))
            return rc;
    }
    // Split-BF16 path: X_i is split once (k-contiguous columns) for W = F X, X^T X as
    // both operands; W once for Z = Sb W and W^T Z.  Regions: h->split_x holds X, the
    // second half of h->split holds W, the first half the operand used once.
    __bf16* const rT = h->split;
    __bf16* const rW = h->split ? h->split + (size_t)B * split_entry_elems(m, m) : nullptr;
    __bf16* const rZ = h->split ? h->split + 2 * (size_t)B * split_entry_elems(m, m) : nullptr;
    int cur = 0;
    // FP64 pipeline: W_j = F_j X_j needs only the factor, so W of the next block is
    // computed in the launch of this block's Z = Sc W (two descriptors), alternating
    // between h->W and Ftmp (idle outside the factor sweep)
    const bool pipe = h->opt.precision == RGF_FP64 && g_selinv_fuse && g_selinv_pipe;
    double* const wb[2] = {h->W, h->Ftmp};
    // keep_w: W_j in its own slot of Wall (kept for the solves), matrices nt slots apart
    const bool keepw = h->Wall != nullptr;
    const long long wst = keepw ? (long long)nt * fs : fs;
    auto wptr = [&](int j) { return keepw ? h->Wall + (long long)j * fs : pipe ? wb[j & 1] : h->W; };
    auto wdesc = [&](int j) {
        const double* Pj = l_panel(h, j);
        const bool lj = j == nt - 1;
        GemmDesc d = gdesc(lj ? Pj + 2 * ns : Pj + ns, ldp, Pj, ldp, wptr(j), lds, lj ? nd : m, ns, ns, 1.0, 0.0,
                           G_BLOWER);
        batch2(d, B, ms, ms, wst);
        return d;
    };
    if (pipe && nd > 0) {
        GemmDesc d = wdesc(nt - 1);
        if ((rc = prod(h, Op::N, Op::N, &d, 1))) return rc;
    }
    for (int i = nt - 1; i >= 0; i--) {
        const double* P = l_panel(h, i);
        const bool last = i == nt - 1;
        const int mF = last ? nd : m;
        const double* F = last ? P + 2 * ns : P + ns;
        const int r0 = last ? ns : 0;
        double* Sc = h->Sb[cur];
        double* Sn = h->Sb[cur ^ 1];
        const double* xsum = nullptr;   // X^T X launched with W^T Z: added by the finishing pass
        GemmDesc dx = gdesc(P, ldp, P, ldp, Sn, lds, ns, ns, ns, 1.0, 0.0, G_CLOWER | G_AUPPER | G_BLOWER);
        batch2(dx, B, ms, ms, ss);
        const bool px = presplit_ok(h, dx);
        // X_i's columns: the factorization's copy (keep_splits), else split here
        const bool kept_x = px && h->opt.keep_splits && h->xc && g_trtri_mixed;
        const SplitView xv = kept_x ? xc_view(h, i) : split_view(h->split_x, ns, ns, dx.count);
        if (px && !kept_x && (rc = split_a(h, Op::T, dx, h->split_x))) return rc;
        auto xtx = [&]() {   // X^T X from the copies of X
            const PresplitCall c{dx, xv, xv, SplitView{}, SplitView{}};
            return gemm_bf16x3_presplit_n(Op::T, Op::N, &c, 1, s);
        };
        double* const Wi = wptr(i);
        if (pipe) {
            GemmDesc zz[2];
            int nz = 0;
            if (mF > 0) {
                zz[nz] = gdesc(Sc + (long long)r0 * lds + r0, lds, Wi, lds, h->Z, lds, mF, ns, mF, 1.0, 0.0, 0);
                batch2(zz[nz++], B, ss, wst, fs);
            }
            if (i > 0) zz[nz++] = wdesc(i - 1);   // into wb[(i - 1) & 1], free since W^T Z of block i + 1
            if (nz && (rc = prod(h, Op::N, Op::N, zz, nz))) return rc;
            if (mF > 0) {
                GemmDesc d = gdesc(Wi, lds, h->Z, lds, Sn, lds, ns, ns, mF, 1.0, 0.0, G_CLOWER);
                batch2(d, B, wst, fs, ss);
                GemmDesc two[2] = {dx, d};
                two[0].C = Sc;
                two[0].Cin = Sc;
                if ((rc = prod(h, Op::T, Op::N, two, 2))) return rc;
                xsum = Sc;
            } else if ((rc = prod(h, Op::T, Op::N, &dx, 1))) {
                return rc;
            }
        } else if (mF > 0) {
            GemmDesc d = gdesc(F, ldp, P, ldp, Wi, lds, mF, ns, ns, 1.0, 0.0, G_BLOWER);
            batch2(d, B, ms, ms, wst);
            const bool pw = px && presplit_ok(h, d);
            if (pw) {   // op(A) = [B_i; D_i]: the factorization's copy (keep_splits) or split now;
                        // op(B) = X: the split of X^T X's op(A); the product writes the split of W
                const bool kept = h->opt.keep_splits && panel_presplit(h, i);
                __bf16* const Fs = kept ? panel_copy(h, i) : rT;
                const PresplitCall ww{d, split_view(Fs, mF, ns, d.count), xv, split_view(rW, ns, mF, d.count),
                                      SplitView{}};
                if ((!kept && (rc = split_a(h, Op::N, d, rT))) ||
                    (rc = gemm_bf16x3_presplit_n(Op::N, Op::N, &ww, 1, s)))
                    return rc;
            } else if ((rc = prod(h, Op::N, Op::N, &d, 1))) {
                return rc;
            }
            d = gdesc(Sc + (long long)r0 * lds + r0, lds, Wi, lds, h->Z, lds, mF, ns, mF, 1.0, 0.0, 0);
            batch2(d, B, ss, wst, fs);
            if (pw) {   // op(A) = S_c: its copy from the step before (or split now); op(B) = W from
                        // the W product; the product writes the split of Z
                const bool kept = sc_copy && r0 == 0 && mF == m;
                if ((!kept && (rc = split_a(h, Op::N, d, rT))) ||
                    (rc = gemm_bf16x3_presplit(Op::N, Op::N, &d, kept ? h->split_s : rT, rW, s, rZ)))
                    return rc;
            } else if ((rc = prod(h, Op::N, Op::N, &d, 1))) {
                return rc;
            }
            // FP64: X^T X and W^T Z in one launch (two descriptors, twice the tiles of either);
            // X^T X goes to the top-left ns x ns of Sc, dead once Z = Sc W is done, and the
            // finishing pass adds it into Sn (the same sum as accumulating it, same bits)
            const bool fuse = h->opt.precision == RGF_FP64 && g_selinv_fuse;
            if (px) {
                if ((rc = xtx())) return rc;
            } else if (!fuse && (rc = prod(h, Op::T, Op::N, &dx, 1))) {
                return rc;
            }
            d = gdesc(Wi, lds, h->Z, lds, Sn, lds, ns, ns, mF, 1.0, fuse ? 0.0 : 1.0, G_CLOWER);
            batch2(d, B, wst, fs, ss);
            if (pw) {   // op(A) = W^T: the split of Z = Sb W's op(B); op(B) = Z from the Z product
                if ((rc = gemm_bf16x3_presplit(Op::T, Op::N, &d, rW, rZ, s))) return rc;
            } else if (fuse) {
                GemmDesc two[2] = {dx, d};
                two[0].C = Sc;
                two[0].Cin = Sc;
                if ((rc = prod(h, Op::T, Op::N, two, 2))) return rc;
                xsum = Sc;
            } else if ((rc = prod(h, Op::T, Op::N, &d, 1))) {
                return rc;
            }
        } else if (px) {
            if ((rc = xtx())) return rc;
        } else if ((rc = prod(h, Op::T, Op::N, &dx, 1))) {
            return rc;
        }
        if ((rc = selinv_finish(Sn, lds, ss, ns, nd, h->Z, lds, fs, mF - nd, exact && mF > 0,
                                diag_out + (long long)i * ns, n, B, s, xsum, lds, ss, sc_copy ? &scp : nullptr)))
            return rc;
        if (so) {
            const double* Sm = Sn + so->mb * ss;
            const long long bs = (long long)ns * ns;
            if (so->diag && (rc = copy2d(Sm, lds, so->diag + i * bs, ns, ns, ns, 1.0, s))) return rc;
            if (so->sub && !last && (rc = copy2d(h->Z + so->mb * fs, lds, so->sub + i * bs, ns, ns, ns, -1.0, s)))
                return rc;
            if (so->arrow && nd > 0 &&
                (rc = copy2d(Sm + ns, lds, so->arrow + (long long)i * nd * ns, nd, nd, ns, 1.0, s)))
                return rc;
        }
        cur ^= 1;
    }
    return 0;
}

int do_selinv_device(rgf_handle_s* h, double* diag_dev)
{
    int rc = ensure_factorized(h);
    if (rc) return rc;
    if ((rc = ensure_selinv_buffers(h))) return rc;
    if (!h->opt.graphs) {
        rc = selinv_sweep(h, diag_dev);
    } else {   // the captured sweep writes the handle's buffer
        rc = run_captured(h, h->gx_selinv, h->gx_epoch_selinv, [h] { return selinv_sweep(h, h->diag); });
        if (!rc && diag_dev != h->diag)
            rc = (int)hipMemcpyAsync(diag_dev, h->diag, sizeof(double) * h->n * h->count, hipMemcpyDeviceToDevice,
                                     h->stream);
    }
    if (rc) return dev_status(h, rc, "selinv");
    h->w_valid = h->Wall != nullptr;
    return RGF_OK;
}

// ------------------------------------------------------------------ residual
// r = b - A x with the resident A (full symmetric diagonal blocks), then ||r||_2.
__global__ __launch_bounds__(1024) void norm2_kernel(const double* r, long long n, double* out)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 1024) s += r[i] * r[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sqrt(red[0]);
}

int block_matvec_residual(rgf_handle_s* h, const SolveBufs& sb, int mb, const double* x, const double* b, double* r,
                          int nr)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp, n = (int)h->n;
    hipStream_t s = h->stream;
    double* w = sb.work;
    const size_t we = sb.work_elems;
    int* const ct = sb.ctr;
    const Op N = Op::N, T = Op::T;
    int rc;
    // The products of r = b - A x are independent: each block-wide product runs over all
    // blocks in one launch (batch index = block, panels at pstride), without k-slabs -- the
    // blocks give the launch its workgroups -- and they accumulate into r in a fixed order:
    // r_t = b_t - A_tt x_t; r_{t+1} -= A_{t+1,t} x_t; r_t -= A_{t+1,t}^T x_{t+1};
    // r_t -= A_{d,t}^T x_d; r_d = b_d - A_d x (the arrow rows of all panels are one nd x nt ns
    // matrix: panels are contiguous) - A_dd x_d.
    const double* A0 = h->Ap + (long long)mb * h->mstride;
    const long long ps = h->pstride;
    const double* xd = x + (long long)nt * ns;
    double* rd = r + (long long)nt * ns;
    if ((rc = skinny_f64s(N, 0, ns, ns, nr, -1.0, A0, ldp, x, 1, n, 1.0, b, 1, n, r, 1, n, w, we, s, nt, ps, ns, ns, ns,
                          ct, nullptr, 0, nullptr, 0, 1, sb.wide)))
        return rc;
    if (nt > 1) {
        if ((rc = skinny_f64s(N, 0, ns, ns, nr, -1.0, A0 + ns, ldp, x, 1, n, 1.0, r + ns, 1, n, r + ns, 1, n, w, we, s,
                              nt - 1, ps, ns, ns, ns, ct, nullptr, 0, nullptr, 0, 1, sb.wide)))
            return rc;
        if ((rc = skinny_f64s(T, 0, ns, ns, nr, -1.0, A0 + ns, ldp, x + ns, 1, n, 1.0, r, 1, n, r, 1, n, w, we, s,
                              nt - 1, ps, ns, ns, ns, ct, nullptr, 0, nullptr, 0, 1, sb.wide)))
            return rc;
    }
    if (nd > 0) {
        if ((rc = skinny_f64s(T, 0, ns, nd, nr, -1.0, A0 + 2 * ns, ldp, xd, 1, n, 1.0, r, 1, n, r, 1, n, w, we, s, nt,
                              ps, 0, ns, ns, ct, nullptr, 0, nullptr, 0, 1, sb.wide)))
            return rc;
        if ((rc = skinny_f64(N, 0, nd, nt * ns, nr, -1.0, A0 + 2 * ns, ldp, x, n, 1.0, b + (long long)nt * ns, n, rd,
                             n, w, s, we, ct, sb.wide)))
            return rc;
    }
    if (nd > 0 &&
        (rc = skinny_f64(N, 0, nd, nd, nr, -1.0, h->Add + (long long)mb * h->cstride, nd, xd, n, 1.0, rd, n, rd,
                         n, w, s, we, ct, sb.wide)))
        return rc;
    return 0;
}

int residual_norm_host(rgf_handle_s* h, const double* x, const double* b, double* out, bool normalized)
{
    if (!h->Ap) return set_err(h, RGF_ESTATE, "residual needs the matrix: create the handle with keep_matrix=1");
    int rc = ensure_solve_buffers(h);
    if (rc) return rc;
    hipSetDevice(h->device);
    const size_t bytes = (size_t)h->n * sizeof(double);
    hipStream_t s = h->stream;
    hipError_t e = hipMemcpyAsync(h->X, x, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h->Bv, b, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return dev_status(h, (int)e, "residual upload");
    rc = block_matvec_residual(h, handle_bufs(h), h->cur, h->X, h->Bv, h->Y, 1);
    if (rc) return dev_status(h, rc, "residual");
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, s, h->Y, h->n, h->scal);
    hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(1024), 0, s, h->Bv, h->n, h->scal + 1);
    double v[2];
    e = hipMemcpyAsync(v, h->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return dev_status(h, (int)e, "residual");
    *out = normalized ? v[0] / v[1] : v[0];
    return RGF_OK;
}

// ------------------------------------------------------------------ MF export
// Packs the panels of one matrix into the reference's MF layout (RGF.H:95, 114-138).
int export_mf(rgf_handle_s* h, const double* panels, const double* corner, double* mf)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd, ldp = h->ldp;
    std::vector<double> buf((size_t)h->pstride);
    size_t off = 0;
    for (int t = 0; t < nt; t++) {
        hipError_t e = hipMemcpy(buf.data(), panels + (long long)t * h->pstride, buf.size() * sizeof(double),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) return dev_status(h, (int)e, "export");
        const bool last = t == nt - 1;
        const int lda = last ? ns + nd : 2 * ns + nd;
        for (int c = 0; c < ns; c++) {
            double* col = mf + off + (size_t)c * lda;
            const double* src = buf.data() + (size_t)c * ldp;
            for (int r = 0; r < ns; r++) col[r] = (r < c) ? 0.0 : src[r];
            if (!last)
                for (int r = 0; r < ns; r++) col[ns + r] = src[ns + r];
            for (int d = 0; d < nd; d++) col[(last ? ns : 2 * ns) + d] = src[2 * ns + d];
        }
        off += (size_t)lda * ns;
    }
    if (nd > 0) {
        std::vector<double> cb((size_t)nd * nd);
        hipError_t e = hipMemcpy(cb.data(), corner, cb.size() * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return dev_status(h, (int)e, "export");
        for (int c = 0; c < nd; c++)
            for (int r = 0; r < nd; r++) mf[off + (size_t)c * nd + r] = (r < c) ? 0.0 : cb[(size_t)c * nd + r];
    }
    return RGF_OK;
}

// ------------------------------------------------------------------ generation
int generate_all(rgf_handle_s* h)
{
    const int ns = h->ns, nt = h->nt, nd = h->nd;
    double *tmp = nullptr, *rows = nullptr;
    size_t gt = 0, gr = 0;
    hipError_t e = cache_alloc(h->device, (size_t)ns * ns * sizeof(double), (void**)&tmp, &gt);
    if (e == hipSuccess) e = cache_alloc(h->device, (size_t)h->n * sizeof(double), (void**)&rows, &gr);
    int krc = 0;
    for (int b = 0; b < h->count && e == hipSuccess && !krc; b++) {
        const rgf_gen_params& gp = h->gen;
        GenParams g{gp.seed, gp.matrix_id + (uint32_t)b, ns, nt, nd, gp.sigma_c, gp.sigma_a, gp.theta_d, gp.theta_c};
        if (gp.theta_mode == 1) {   // per-instance scales theta_d ~ U(0.5, 2), theta_c ~ U(0.1, 1)
            g.theta_d = 0.5 + 1.5 * gen_uniform_host(gp.seed, g.matrix_id, 0);
            g.theta_c = 0.1 + 0.9 * gen_uniform_host(gp.seed, g.matrix_id, 1);
        }
        krc = generate_matrix(g, a_panel(h, 0, b), h->pstride, h->ldp, a_corner(h, b), tmp, rows, h->stream);
    }
    hipError_t es = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = es;
    cache_free(h->device, tmp, gt);
    cache_free(h->device, rows, gr);
    if (e != hipSuccess || krc) return dev_status(h, e != hipSuccess ? (int)e : krc, "generate");
    h->factorized = false;
    h->w_valid = false;
    h->broken = false;
    return RGF_OK;
}

int create_generated(size_t ns, size_t nt, size_t nd, size_t count, const rgf_gen_params* gen, const rgf_options* opts,
                     rgf_handle_t* out)
{
    if (!out) return set_err(nullptr, RGF_EINVAL, "null argument");
    *out = nullptr;
    int rc;
    rgf_handle_s* h = new_handle(ns, nt, nd, count, opts, &rc);
    if (!h) return rc;
    rgf_default_gen_params(&h->gen);
    if (gen) h->gen = *gen;
    h->generated = true;
    rc = generate_all(h);
    if (rc) {
        g_create_error = h->err;
        rgf_destroy(h);
        return rc;
    }
    *out = h;
    return RGF_OK;
}

// ------------------------------------------------------------------ no-allocation solve
// For callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc): the solve buffers live in one static
// device array per GPU instead of the handle.  Every call on a GPU shares that GPU's
// array (held until the process ends), so calls that overlap on one GPU race; nothing in
// it needs to start at zero (every buffer is written before it is read).
constexpr size_t kScratchBytes = (size_t)128 << 20;
constexpr size_t kScratchElems = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchElems];

std::mutex g_scratch_mu;
double* g_scratch_ptr[64] = {};   // device address of `scratch` on each device

// The array's device address on the current device (hipGetSymbolAddress once per device;
// the array's name is a host-side stand-in and must never reach a kernel).
int scratch_on_device(int dev, double** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
        }
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Partials the solve (and, for the refinement, the block matvec) needs with nr columns:
// the largest over the skinny products of one matrix's sweep.
size_t solve_call_work(const rgf_handle_s* h, int nr)
{
    const int ns = h->ns, nd = h->nd;
    size_t w = 0;
    const Op N = Op::N, T = Op::T;
    for (auto c : {std::make_tuple(N, ns, ns), std::make_tuple(N, nd, ns), std::make_tuple(N, nd, nd),
                   std::make_tuple(T, nd, nd), std::make_tuple(T, ns, nd), std::make_tuple(T, ns, ns),
                   std::make_tuple(N, ns + nd, ns), std::make_tuple(T, ns, ns + nd), std::make_tuple(N, 2 * ns + nd, ns)})
        w = std::max(w, skinny_call_work(std::get<0>(c), std::get<1>(c), std::get<2>(c), nr));
    return w;
}

inline size_t r32(size_t v) { return (v + 31) / 32 * 32; }   // 256-byte steps

// the skinny products' arrival counters at the start of the scratch (zero when the array
// is loaded; every launch leaves them zero)
constexpr size_t kScratchCtr = SKINNY_CTRS / 2;   // doubles

// scratch doubles of a chunk of nr columns (refined: R, Dc and, in place, Bv besides Y)
size_t noalloc_need(const rgf_handle_s* h, int nr, bool refine, bool in_place)
{
    const size_t col = r32((size_t)h->n * nr);
    return kScratchCtr + col * (1 + (refine ? 2 : 0) + (refine && in_place ? 1 : 0)) + r32(solve_call_work(h, nr));
}

int solve_device_noalloc(rgf_handle_s* h, double* x, const double* b, size_t nrhs)
{
    const long long n = h->n;
    const bool refine = h->opt.precision != RGF_FP64 && h->Ap && h->opt.refine_iters > 0;
    const bool in_place = x == b;
    // FP64 columns are independent of their chunk (same products, same order: a call of three
    // or more columns runs every chunk on the MFMA kernel, as rgf_solve_device does), so the
    // chunk shrinks to what the scratch holds; a refined chunk keeps the 16 columns of
    // rgf_solve_device, whose refinement stops on the chunk's worst column
    int nr = (int)std::min<size_t>(16, nrhs);
    if (!refine)
        while (nr > 1 && noalloc_need(h, nr, false, false) > kScratchElems) nr /= 2;
    const size_t need = noalloc_need(h, nr, refine, in_place);
    if (need > kScratchElems)
        return set_err(h, RGF_EINVAL,
                       "solve of n = %lld with %d right-hand sides per pass needs %zu bytes of scratch; the static "
                       "scratch of the no-allocation entry point holds %zu bytes (use rgf_solve_device)",
                       n, nr, need * sizeof(double), kScratchBytes);
    if (!h->factorized)
        return set_err(h, RGF_ESTATE, "the no-allocation solve needs a factorized handle (call rgf_factorize first)");
    double* base = nullptr;
    int rc = scratch_on_device(h->device, &base);
    if (rc) return dev_status(h, rc, "scratch address");
    const size_t col = r32((size_t)n * nr);
    SolveBufs sb{};
    sb.wide = nrhs > 2;   // the kernels of rgf_solve_device for this call, whatever the chunk
    sb.ctr = (int*)base;
    sb.Y = base + kScratchCtr;
    size_t off = kScratchCtr + col;
    if (refine) {
        sb.R = base + off; off += col;
        sb.Dc = base + off; off += col;
        if (in_place) { sb.Bv = base + off; off += col; }
    }
    sb.work = base + off;
    sb.work_elems = kScratchElems - off;
    h->last_refine_iters = 0;
    h->last_refine_res = 0.0;
    for (size_t c0 = 0; c0 < nrhs; c0 += (size_t)nr) {
        const int nc = (int)std::min<size_t>((size_t)nr, nrhs - c0);
        double* xc = x + c0 * n;
        const double* bc = b + c0 * n;
        if (refine) {
            if (in_place) {
                if ((rc = copy2d(bc, (int)n, sb.Bv, (int)n, (int)n, nc, 1.0, h->stream))) break;
                bc = sb.Bv;
            }
            rc = solve_refined(h, sb, h->cur, xc, bc, nc, true);
        } else {
            rc = solve_block(h, sb, h->cur, xc, bc, nc);
        }
        if (rc) break;
    }
    return rc ? dev_status(h, rc, "solve (no allocation)") : RGF_OK;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

void rgf_default_options(rgf_options* o)
{
    o->precision = RGF_FP64;
    o->selinv_mode = RGF_SELINV_EXACT;
    o->keep_matrix = 1;
    o->device = 0;
    o->refine_iters = 10;
    o->refine_tol = 1e-14;
    o->graphs = 0;
    o->keep_splits = 1;
    o->keep_w = 0;
    o->cholesky = RGF_CHOLESKY_AUTO;
}

void rgf_default_gen_params(rgf_gen_params* p)
{
    p->seed = 0;
    p->matrix_id = 0;
    p->sigma_c = 0.1;
    p->sigma_a = 0.05;
    p->theta_d = 1.0;
    p->theta_c = 1.0;
    p->theta_mode = 0;
}

int rgf_create_csc(const size_t* ia, const size_t* ja, const double* a, size_t ns, size_t nt, size_t nd,
                   const rgf_options* opts, rgf_handle_t* out)
{
    if (!ia || !ja || !a || !out) return set_err(nullptr, RGF_EINVAL, "null argument");
    *out = nullptr;
    int rc = validate_shape(ns, nt, nd);
    if (rc) return rc;
    const size_t n = ns * nt + nd, sn = ns * nt;
    if (ia[0] != 0) return set_err(nullptr, RGF_EINVAL, "ia[0] must be 0");
    for (size_t c = 0; c < n; c++) {
        if (ia[c + 1] < ia[c]) return set_err(nullptr, RGF_EINVAL, "ia is not non-decreasing at column %zu", c);
        for (size_t k = ia[c]; k < ia[c + 1]; k++) {
            const size_t r = ja[k];
            bool ok;
            if (r < c || r >= n) ok = false;
            else if (c >= sn || r >= sn) ok = true;
            else ok = (r / ns == c / ns) || (r / ns == c / ns + 1);
            if (!ok)
                return set_err(nullptr, RGF_EINVAL,
                               "CSC entry %zu (row %zu, column %zu) is outside the lower block-arrow pattern", k, r, c);
        }
    }
    rgf_handle_s* h = new_handle(ns, nt, nd, 1, opts, &rc);
    if (!h) return rc;
    const size_t nnz = ia[n];
    uint64_t *dia = nullptr, *dja = nullptr;
    double* da = nullptr;
    size_t g1 = 0, g2 = 0, g3 = 0;
    hipError_t e = cache_alloc(h->device, (n + 1) * sizeof(uint64_t), (void**)&dia, &g1);
    if (e == hipSuccess) e = cache_alloc(h->device, std::max<size_t>(nnz, 1) * sizeof(uint64_t), (void**)&dja, &g2);
    if (e == hipSuccess) e = cache_alloc(h->device, std::max<size_t>(nnz, 1) * sizeof(double), (void**)&da, &g3);
    if (e == hipSuccess) e = hipMemcpy(dia, ia, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(dja, ja, nnz * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz) e = hipMemcpy(da, a, nnz * sizeof(double), hipMemcpyHostToDevice);
    int krc = 0;
    if (e == hipSuccess) {
        double* P = a_panel(h, 0);
        krc = fill2d(P, h->ldp, h->ldp, (int)(nt * ns), 0.0, h->stream);   // panels are contiguous
        if (!krc) krc = fill2d(a_corner(h), (int)std::max<size_t>(nd, 1), (int)nd, (int)nd, 0.0, h->stream);
        if (!krc)
            krc = scatter_csc(dia, dja, da, (long long)nnz, (int)ns, (int)nt, (int)nd, P, h->pstride, h->ldp,
                              a_corner(h), h->stream);
        if (!krc) e = hipStreamSynchronize(h->stream);
    }
    hipStreamSynchronize(h->stream);
    cache_free(h->device, dia, g1);
    cache_free(h->device, dja, g2);
    cache_free(h->device, da, g3);
    if (e != hipSuccess || krc) {
        rc = dev_status(h, e != hipSuccess ? (int)e : krc, "CSC upload");
        g_create_error = h->err;
        rgf_destroy(h);
        return rc;
    }
    *out = h;
    return RGF_OK;
}

int rgf_create_generated(size_t ns, size_t nt, size_t nd, const rgf_gen_params* gen, const rgf_options* opts,
                         rgf_handle_t* out)
{
    return create_generated(ns, nt, nd, 1, gen, opts, out);
}

int rgf_create_generated_batch(size_t ns, size_t nt, size_t nd, size_t count, const rgf_gen_params* gen,
                               const rgf_options* opts, rgf_handle_t* out)
{
    return create_generated(ns, nt, nd, count, gen, opts, out);
}

int rgf_regenerate(rgf_handle_t h)
{
    if (!h) return RGF_EINVAL;
    if (!h->generated) return set_err(h, RGF_ESTATE, "handle was not created by the generator");
    hipSetDevice(h->device);
    return generate_all(h);
}

size_t rgf_footprint_bytes(size_t ns, size_t nt, size_t nd, size_t count, const rgf_options* opts)
{
    if (validate_shape(ns, nt, nd)) return 0;
    rgf_options o;
    rgf_default_options(&o);
    if (opts) o = *opts;
    return footprint_bytes((int)ns, (int)nt, (int)nd, (int)count, o.keep_matrix, o.precision, o.keep_splits,
                           o.cholesky, o.keep_w);
}

int rgf_destroy(rgf_handle_t h)
{
    if (!h) return RGF_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->gx_factor) hipGraphExecDestroy(h->gx_factor);
    if (h->gx_selinv) hipGraphExecDestroy(h->gx_selinv);
    for (auto& g : h->gx_pieces)
        if (g.gx) hipGraphExecDestroy(g.gx);
    for (auto& a : h->allocs) cache_free(h->device, a.first, a.second);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return RGF_OK;
}

int rgf_set_device_cache_limit(size_t bytes, size_t* previous)
{
    const size_t old = cache_cap_ref().exchange(bytes);
    if (previous) *previous = old;
    for (int d = 0; d < 64; d++)
        if (g_arena[d].free_bytes() > bytes) {
            hipSetDevice(d);
            g_arena[d].trim_to(bytes);
        }
    return RGF_OK;
}

int rgf_trim_device_cache(int device)
{
    if (device < -1 || device >= 64) return set_err(nullptr, RGF_EINVAL, "device %d out of range", device);
    for (int d = 0; d < 64; d++)
        if ((device < 0 || d == device) && g_arena[d].free_bytes() > 0) {
            hipSetDevice(d);
            g_arena[d].trim_to(0);
        }
    return RGF_OK;
}

const char* rgf_last_error(rgf_handle_t h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int rgf_batch_count(rgf_handle_t h) { return h ? h->count : 0; }

int rgf_select(rgf_handle_t h, size_t index)
{
    if (!h) return RGF_EINVAL;
    if (index >= (size_t)h->count) return set_err(h, RGF_EINVAL, "index %zu >= batch count %d", index, h->count);
    h->cur = (int)index;
    return RGF_OK;
}

int rgf_factorize(rgf_handle_t h)
{
    if (!h) return RGF_EINVAL;
    if (!h->Ap && h->factorized) return RGF_OK;
    h->factorized = false;
    return do_factorize(h);
}

int rgf_logdet_batch(rgf_handle_t h, double* out)
{
    if (!h || !out) return RGF_EINVAL;
    int rc = ensure_factorized(h);
    if (rc) return rc;
    const size_t np = (size_t)h->partstride * h->count;
    std::vector<double> parts(np);
    hipError_t e = hipMemcpyAsync(parts.data(), h->parts, np * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return dev_status(h, (int)e, "logdet");
    const size_t used = (size_t)h->nt * h->T + (h->nd > 0 ? 1 : 0);
    for (int b = 0; b < h->count; b++) {
        double s = 0.0;
        for (size_t k = 0; k < used; k++) s += parts[(size_t)b * h->partstride + k];   // fixed order
        out[b] = 2.0 * s;
    }
    return RGF_OK;
}

int rgf_logdet(rgf_handle_t h, double* out)
{
    if (!h || !out) return RGF_EINVAL;
    std::vector<double> all(h->count);
    int rc = rgf_logdet_batch(h, all.data());
    if (rc) return rc;
    *out = all[h->cur];
    return RGF_OK;
}

int rgf_solve_device(rgf_handle_t h, double* x_dev, const double* b_dev, size_t nrhs)
{
    if (!h || !x_dev || !b_dev) return RGF_EINVAL;
    hipSetDevice(h->device);
    return do_solve_device(h, x_dev, b_dev, nrhs);
}

int rgf_solve_device_noalloc(rgf_handle_t h, double* x_dev, const double* b_dev, size_t nrhs)
{
    if (!h || !x_dev || !b_dev) return RGF_EINVAL;
    if (nrhs == 0) return RGF_OK;
    hipSetDevice(h->device);
    return solve_device_noalloc(h, x_dev, b_dev, nrhs);
}

int rgf_solve(rgf_handle_t h, double* x, const double* b, size_t nrhs)
{
    if (!h || !x || !b) return RGF_EINVAL;
    if (nrhs == 0) return RGF_OK;
    hipSetDevice(h->device);
    int rc = ensure_factorized(h);
    if (rc) return rc;
    if ((rc = ensure_solve_buffers(h))) return rc;
    const long long n = h->n;
    for (size_t c0 = 0; c0 < nrhs; c0 += 16) {
        const int nr = (int)std::min<size_t>(16, nrhs - c0);
        const size_t bytes = (size_t)n * nr * sizeof(double);
        hipError_t e = hipMemcpyAsync(h->Bv, b + c0 * n, bytes, hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) return dev_status(h, (int)e, "solve upload");
        SolveBufs sb = handle_bufs(h);
        sb.wide = nrhs > 2;
        rc = h->opt.precision != RGF_FP64
                 ? solve_refined(h, sb, h->cur, h->X, h->Bv, nr, false)
                 : run_piece(h, 0, h->cur, nr, h->X, h->Bv, [&] { return solve_block(h, sb, h->cur, h->X, h->Bv, nr); });
        if (rc) return dev_status(h, rc, "solve");
        e = hipMemcpyAsync(x + c0 * n, h->X, bytes, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return dev_status(h, (int)e, "solve download");
    }
    return RGF_OK;
}

int rgf_selinv_diag_device(rgf_handle_t h, double* diag_dev)
{
    if (!h || !diag_dev) return RGF_EINVAL;
    hipSetDevice(h->device);
    return do_selinv_device(h, diag_dev);
}

int rgf_selinv_blocks_device(rgf_handle_t h, double* diag_blocks, double* sub_blocks, double* arrow_blocks,
                             double* corner)
{
    if (!h) return RGF_EINVAL;
    if (h->opt.selinv_mode != RGF_SELINV_EXACT)
        return set_err(h, RGF_EINVAL, "the blocks of A^-1 need the exact selected inversion (selinv_mode 0)");
    hipSetDevice(h->device);
    int rc = ensure_factorized(h);
    if (rc) return rc;
    if ((rc = ensure_selinv_buffers(h))) return rc;
    const SigmaOut so{diag_blocks, sub_blocks, arrow_blocks, corner, h->cur};
    rc = selinv_sweep(h, h->diag, &so);   // eager: the output buffers vary between calls
    if (rc) return dev_status(h, rc, "selinv blocks");
    h->w_valid = h->Wall != nullptr;
    return RGF_OK;
}

int rgf_selinv_diag_batch(rgf_handle_t h, double* diag)
{
    if (!h || !diag) return RGF_EINVAL;
    hipSetDevice(h->device);
    int rc = ensure_factorized(h);
    if (rc) return rc;
    if ((rc = ensure_selinv_buffers(h))) return rc;
    rc = do_selinv_device(h, h->diag);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(diag, h->diag, (size_t)h->n * h->count * sizeof(double), hipMemcpyDeviceToHost,
                                  h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return dev_status(h, (int)e, "selinv download");
    return RGF_OK;
}

int rgf_selinv_diag(rgf_handle_t h, double* diag)
{
    if (!h || !diag) return RGF_EINVAL;
    hipSetDevice(h->device);
    int rc = ensure_factorized(h);
    if (rc) return rc;
    if ((rc = ensure_selinv_buffers(h))) return rc;
    rc = do_selinv_device(h, h->diag);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(diag, h->diag + (long long)h->cur * h->n, (size_t)h->n * sizeof(double),
                                  hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return dev_status(h, (int)e, "selinv download");
    return RGF_OK;
}

int rgf_residual_norm(rgf_handle_t h, const double* x, const double* b, double* out)
{
    if (!h || !x || !b || !out) return RGF_EINVAL;
    return residual_norm_host(h, x, b, out, false);
}

int rgf_residual_norm_normalized(rgf_handle_t h, const double* x, const double* b, double* out)
{
    if (!h || !x || !b || !out) return RGF_EINVAL;
    return residual_norm_host(h, x, b, out, true);
}

int rgf_refine_stats(rgf_handle_t h, int* iters, double* rel_residual)
{
    if (!h) return RGF_EINVAL;
    if (iters) *iters = h->last_refine_iters;
    if (rel_residual) *rel_residual = h->last_refine_res;
    return RGF_OK;
}

int rgf_synchronize(rgf_handle_t h)
{
    if (!h) return RGF_EINVAL;
    hipError_t e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? RGF_OK : dev_status(h, (int)e, "synchronize");
}

int rgf_get_info(rgf_handle_t h, rgf_info_t* info)
{
    if (!h || !info) return RGF_EINVAL;
    const double ns = h->ns, nt = h->nt, nd = h->nd;
    const double P = ns * ns * ns / 3.0, Pd = nd * nd * nd / 3.0;
    info->ns = h->ns; info->nt = h->nt; info->nd = h->nd; info->n = (size_t)h->n;
    info->device_bytes = h->bytes;
    info->flops_factor = nt * P + (nt - 1) * 2 * ns * ns * ns + nd * (nt * ns * ns + 2 * (nt - 1) * ns * ns + nt * nd * ns) + Pd;
    info->flops_selinv = nt * 2 * P + (nt - 1) * 4 * ns * ns * ns + nd * (nt * ns * ns + 4 * (nt - 1) * ns * ns) + 2 * Pd;
    info->bytes_solve_per_rhs = 2.0 * 8.0 * (nt * ns * (ns + 1) / 2 + (nt - 1) * ns * ns + nt * ns * nd + nd * (nd + 1) / 2);
    return RGF_OK;
}

int rgf_export_matrix_mf(rgf_handle_t h, double* mf)
{
    if (!h || !mf) return RGF_EINVAL;
    if (!h->Ap) return set_err(h, RGF_ESTATE, "matrix not kept (keep_matrix=0)");
    hipSetDevice(h->device);
    hipStreamSynchronize(h->stream);
    return export_mf(h, h->Ap + (long long)h->cur * h->mstride, h->Add + (long long)h->cur * h->cstride, mf);
}

int rgf_export_factor_mf(rgf_handle_t h, double* mf)
{
    if (!h || !mf) return RGF_EINVAL;
    hipSetDevice(h->device);
    int rc = ensure_factorized(h);
    if (rc) return rc;
    hipStreamSynchronize(h->stream);
    return export_mf(h, h->Lp + (long long)h->cur * h->mstride, h->Xdd + (long long)h->cur * h->cstride, mf);
}

}  // extern "C"

// ====================================================================== diagnostics
#include "../../include/rgf_testing.h"

namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};
}  // namespace

extern "C" int rgf_debug_gemm(int ta, int tb, int m, int n, int k, double alpha, const double* A, int lda,
                              const double* B, int ldb, double beta, double* C, int ldc, int flags)
{
    if (m <= 0 || n <= 0 || k < 0) return RGF_EINVAL;
    const size_t na = (size_t)lda * (ta ? m : k), nb = (size_t)ldb * (tb ? k : n), nc = (size_t)ldc * n;
    DevBuf a, b, c;
    hipStream_t s = nullptr;
    hipError_t e = hipMalloc(&a.p, std::max<size_t>(na, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&b.p, std::max<size_t>(nb, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&c.p, nc * 8);
    if (e == hipSuccess && na) e = hipMemcpy(a.p, A, na * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && nb) e = hipMemcpy(b.p, B, nb * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c.p, C, nc * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_err(nullptr, RGF_EDEVICE, "debug_gemm: %s", hipGetErrorString(e));
    GemmDesc d = gdesc((const double*)a.p, lda, (const double*)b.p, ldb, (double*)c.p, ldc, m, n, k, alpha, beta,
                       flags & 0xff);
    DevBuf sc;
    int rc;
    if (flags & 0x200) {
        const size_t sce = split_scratch_elems(m, n, k, 1);
        if (hipMalloc(&sc.p, sce * 2) != hipSuccess) return set_err(nullptr, RGF_ENOMEM, "debug_gemm: scratch");
        rc = gemm_bf16x3_split(ta ? Op::T : Op::N, tb ? Op::T : Op::N, &d, (__bf16*)sc.p, sce, s);
    } else {
        rc = (flags & 0x100) ? gemm_bf16x3(ta ? Op::T : Op::N, tb ? Op::T : Op::N, &d, 1, s)
                             : gemm_f64(ta ? Op::T : Op::N, tb ? Op::T : Op::N, &d, 1, s, ((flags >> 12) & 7) - 1);
    }
    if (rc) return dev_status(nullptr, rc, "debug_gemm");
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(C, c.p, nc * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_err(nullptr, RGF_EDEVICE, "debug_gemm: %s", hipGetErrorString(e));
    return RGF_OK;
}

extern "C" int rgf_debug_potrf(int n, double* A, int lda, int invert, double* logdet)
{
    if (n <= 0 || lda < n) return RGF_EINVAL;
    const int T = (n + 63) / 64;
    size_t tt = 1;
    for (int sz = 1; sz < T; sz *= 2) {
        size_t G = (size_t)((T - sz + 2 * sz - 1) / (2 * sz));
        tt = std::max(tt, G * (size_t)(sz * 64) * (size_t)(sz * 64));
    }
    DevBuf a, parts, info, tmp, xd, fl, xb;
    hipError_t e = hipMalloc(&a.p, (size_t)lda * n * 8);
    if (e == hipSuccess) e = hipMalloc(&xb.p, potrf_dag_xbuf_elems(n) * 8);
    if (e == hipSuccess) e = hipMalloc(&fl.p, potrf_dag_flag_elems(n) * sizeof(int));
    if (e == hipSuccess) e = hipMemset(fl.p, 0, potrf_dag_flag_elems(n) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&xd.p, (size_t)T * 64 * 64 * 8);
    if (e == hipSuccess) e = hipMalloc(&parts.p, (size_t)T * 8);
    if (e == hipSuccess) e = hipMalloc(&info.p, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&tmp.p, tt * 8);
    if (e == hipSuccess) e = hipMemcpy(a.p, A, (size_t)lda * n * 8, hipMemcpyHostToDevice);
    int init = INT_MAX;
    if (e == hipSuccess) e = hipMemcpy(info.p, &init, sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_err(nullptr, RGF_EDEVICE, "debug_potrf: %s", hipGetErrorString(e));
    // L in place (diagonal tiles included), or L^-1: the inverse takes the diagonal tiles from
    // xdiag, so that run factors as the sweeps do (task-graph launch)
    bool x_done = false;
    int rc = potrf_f64((double*)a.p, lda, 0, n, 1, (double*)parts.p, 0, (int*)info.p, 0, (double*)xd.p, 0, !invert,
                       nullptr, (int*)fl.p, (double*)xb.p, 0, &x_done);
    if (!rc && invert && !x_done) rc = trtri_f64_inplace((double*)a.p, lda, 0, n, 1, (double*)tmp.p, 0, (const double*)xd.p, 0, nullptr);
    if (rc) return dev_status(nullptr, rc, "debug_potrf");
    std::vector<double> hp(T);
    int hinfo = 0;
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(A, a.p, (size_t)lda * n * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hp.data(), parts.p, (size_t)T * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_err(nullptr, RGF_EDEVICE, "debug_potrf: %s", hipGetErrorString(e));
    double sum = 0.0;
    for (double v : hp) sum += v;
    if (logdet) *logdet = 2.0 * sum;
    if (hinfo == INT_MIN) return set_err(nullptr, RGF_EDEVICE, "debug_potrf: a dependency wait expired");
    if (hinfo != INT_MAX) return set_err(nullptr, RGF_ENOTSPD, "non-positive pivot at column %d", hinfo);
    return RGF_OK;
}

// Device-timed loops over single kernels, for tuning (buffers allocated once, filled with
// a deterministic pattern; results are discarded).
__global__ void bench_fill_kernel(double* p, long long n, double scale)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = scale * (double)((i * 2654435761LL) % 1000) / 1000.0;
}

__global__ void diag_shift_kernel(double* A, int n, double v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(long long)i * n + i] += v;
}

extern "C" int rgf_debug_bench_gemm(int ta, int tb, int m, int n, int k, int flags, int reps, double* ms_per_launch)
{
    if (m <= 0 || n <= 0 || k <= 0 || reps <= 0 || !ms_per_launch) return RGF_EINVAL;
    const int lda = ta ? k : m, ldb = tb ? n : k, ldc = m;
    const size_t na = (size_t)lda * (ta ? m : k), nb = (size_t)ldb * (tb ? k : n), nc = (size_t)ldc * n;
    DevBuf a, b, c;
    hipError_t e = hipMalloc(&a.p, na * 8);
    if (e == hipSuccess) e = hipMalloc(&b.p, nb * 8);
    if (e == hipSuccess) e = hipMalloc(&c.p, nc * 8);
    if (e != hipSuccess) return set_err(nullptr, RGF_ENOMEM, "bench_gemm: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(bench_fill_kernel, dim3(1024), dim3(256), 0, nullptr, (double*)a.p, (long long)na, 1.0);
    hipLaunchKernelGGL(bench_fill_kernel, dim3(1024), dim3(256), 0, nullptr, (double*)b.p, (long long)nb, 1.0);
    hipLaunchKernelGGL(bench_fill_kernel, dim3(1024), dim3(256), 0, nullptr, (double*)c.p, (long long)nc, 1.0);
    GemmDesc d = gdesc((const double*)a.p, lda, (const double*)b.p, ldb, (double*)c.p, ldc, m, n, k, 1.0, 1.0, flags & 0xff);
    DevBuf sc;
    size_t sce = 0;
    if (flags & 0x200) {
        sce = split_scratch_elems(m, n, k, 1);
        if (hipMalloc(&sc.p, sce * 2) != hipSuccess) return set_err(nullptr, RGF_ENOMEM, "bench_gemm: scratch");
    }
    auto launch = [&]() {
        const Op oa = ta ? Op::T : Op::N, ob = tb ? Op::T : Op::N;
        if (flags & 0x200) return gemm_bf16x3_split(oa, ob, &d, (__bf16*)sc.p, sce, nullptr);
        // bits 12-14: FP64 tile shape + 1 (0 = automatic)
        return (flags & 0x100) ? gemm_bf16x3(oa, ob, &d, 1, nullptr)
                               : gemm_f64(oa, ob, &d, 1, nullptr, ((flags >> 12) & 7) - 1);
    };
    int rc = launch();
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && !rc; r++) rc = launch();
    hipEventRecord(e1, nullptr);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return dev_status(nullptr, rc, "bench_gemm");
    *ms_per_launch = ms / reps;
    return RGF_OK;
}

extern "C" int rgf_debug_dag_order(int T, int with_x, int xlag, int promo, int* out, int cap)
{
    if (T < 1 || T > 4096 || xlag < 0 || xlag > 16 || cap < 0 || (cap > 0 && !out)) return -1;
    return potrf_dag_order(T, with_x != 0, xlag, promo != 0, out, cap);
}

extern "C" int rgf_debug_set_tuning(int key, int value)
{
    if (key == 1 && (value == 32 || value == 64)) {
        const int old = g_presplit_kb;
        g_presplit_kb = value;
        return old;
    }
    if (key == 3 && (value == 0 || value == 1)) {
        const int old = g_presplit_swizzle;
        g_presplit_swizzle = value;
        return old;
    }
    if (key == 4 && (value == 0 || value == 1)) {
        const int old = g_gemm_fast_loads;
        g_gemm_fast_loads = value;
        return old;
    }
    if (key == 8 && (value == 0 || value == 1)) {
        const int old = g_rr_tri;
        g_rr_tri = value;
        return old;
    }
    if (key == 12 && (value == 0 || value == 1)) {
        const int old = g_selinv_pipe;
        g_selinv_pipe = value;
        return old;
    }
    if (key == 11 && (value == 0 || value == 1)) {
        const int old = g_selinv_fuse;
        g_selinv_fuse = value;
        return old;
    }
    if (key == 13 && value >= 0) {
        const int old = g_small_tile_wgs;
        g_small_tile_wgs = value;
        return old;
    }
    if (key == 15 && (value == 0 || value == 1)) {
        const int old = g_trtri_mixed;
        g_trtri_mixed = value;
        return old;
    }
    if (key == 14 && value >= 0) {
        const int old = g_wide_min_wgs;
        g_wide_min_wgs = value;
        return old;
    }
    if (key == 16 && (value == 0 || value == 1)) {
        const int old = g_chol_dag;
        g_chol_dag = value;
        return old;
    }
    if (key == 17 && value >= 2) {
        const int old = g_chol_dag_wgs;
        g_chol_dag_wgs = value;
        return old;
    }
    if (key == 18 && value >= 2) {
        const int old = g_chol_dag_min;
        g_chol_dag_min = value;
        return old;
    }
    if (key == 19 && (value == 0 || value == 1)) {
        const int old = g_chol_dag_x;
        g_chol_dag_x = value;
        return old;
    }
    if (key == 20 && value >= 0 && value <= 16) {
        const int old = g_chol_dag_xlag;
        g_chol_dag_xlag = value;
        return old;
    }
    if (key == 21 && (value == 0 || value == 1)) {
        const int old = g_chol_dag_promo;
        g_chol_dag_promo = value;
        return old;
    }
    if (key == 22 && value >= 64 && value <= 2048) {   // the work buffers are sized for targets up to 2048
        const int old = g_skinny_target_wg;
        g_skinny_target_wg = value;
        return old;
    }
    if (key == 9 && (value == 0 || value == 1)) {
        const int old = g_rr_bands;
        g_rr_bands = value;
        return old;
    }
    return -1;
}

extern "C" int rgf_debug_cache_stats(long long* hits, long long* mallocs, long long* frees, double* cached_gb)
{
    long long h = 0, m = 0, f = 0;
    size_t t = 0;
    for (const auto& A : g_arena) {
        const auto st = A.stats();
        h += st.hits;
        m += st.grows;
        f += st.releases;
        t += A.free_bytes();
    }
    if (hits) *hits = h;
    if (mallocs) *mallocs = m;
    if (frees) *frees = f;
    if (cached_gb) *cached_gb = (double)t / (double)(1ull << 30);
    return RGF_OK;
}

extern "C" int rgf_debug_bench_potrf(int n, int reps, double* ms_potrf, double* ms_tile)
{
    if (n <= 0 || reps <= 0) return RGF_EINVAL;
    const int T = (n + 63) / 64;
    DevBuf a, a0, parts, info, xd, fl, xb, tmp;
    const size_t na = (size_t)n * n;
    hipError_t e = hipMalloc(&a.p, na * 8);
    if (e == hipSuccess) e = hipMalloc(&xb.p, potrf_dag_xbuf_elems(n) * 8);
    if (e == hipSuccess) e = hipMalloc(&tmp.p, na * 8);
    if (e == hipSuccess) e = hipMalloc(&fl.p, potrf_dag_flag_elems(n) * sizeof(int));
    if (e == hipSuccess) e = hipMemset(fl.p, 0, potrf_dag_flag_elems(n) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&a0.p, na * 8);
    if (e == hipSuccess) e = hipMalloc(&parts.p, (size_t)T * 8);
    if (e == hipSuccess) e = hipMalloc(&info.p, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&xd.p, (size_t)T * 64 * 64 * 8);
    if (e != hipSuccess) return set_err(nullptr, RGF_ENOMEM, "bench_potrf: %s", hipGetErrorString(e));
    // SPD test block: small symmetric noise + n on the diagonal
    hipLaunchKernelGGL(bench_fill_kernel, dim3(1024), dim3(256), 0, nullptr, (double*)a0.p, (long long)na, 0.001);
    int rc = mirror_lower((double*)a0.p, n, n, nullptr);
    hipLaunchKernelGGL(diag_shift_kernel, dim3((n + 255) / 256), dim3(256), 0, nullptr, (double*)a0.p, n, (double)n);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float tot = 0.f;
    for (int r = 0; r < reps && !rc; r++) {
        rc = copy2d((const double*)a0.p, n, (double*)a.p, n, n, n, 1.0, nullptr);
        hipEventRecord(e0, nullptr);
        bool x_done = false;   // Cholesky and L^-1, as a factor sweep forms them
        if (!rc) rc = potrf_f64((double*)a.p, n, 0, n, 1, (double*)parts.p, 0, (int*)info.p, 0, (double*)xd.p, 0, false, nullptr,
                                (int*)fl.p, (double*)xb.p, 0, &x_done);
        if (!rc && !x_done) rc = trtri_f64_inplace((double*)a.p, n, 0, n, 1, (double*)tmp.p, 0, (const double*)xd.p, 0, nullptr);
        hipEventRecord(e1, nullptr);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        tot += ms;
    }
    // the tile kernel alone, repeated on the first tile
    float mst = 0.f;
    if (!rc) {
        hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps * 8 && !rc; r++) rc = potrf_tile_only((double*)a.p, n, (double*)xd.p, (double*)parts.p, (int*)info.p, nullptr);
        hipEventRecord(e1, nullptr);
        hipEventSynchronize(e1);
        hipEventElapsedTime(&mst, e0, e1);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return dev_status(nullptr, rc, "bench_potrf");
    if (ms_potrf) *ms_potrf = tot / reps;
    if (ms_tile) *ms_tile = mst / (reps * 8);
    return RGF_OK;
}

extern "C" int rgf_debug_trace_potrf(int n, long long* out, long long cap, long long* nrec, double* ticks_per_us)
{
    if (n <= 0 || cap < 0 || !out || !nrec) return RGF_EINVAL;
    const int T = (n + 63) / 64;
    DevBuf a, parts, info, xd, fl, tr, xb;
    const size_t na = (size_t)n * n;
    hipError_t e = hipMalloc(&a.p, na * 8);
    if (e == hipSuccess) e = hipMalloc(&xb.p, potrf_dag_xbuf_elems(n) * 8);
    if (e == hipSuccess) e = hipMalloc(&fl.p, potrf_dag_flag_elems(n) * sizeof(int));
    if (e == hipSuccess) e = hipMemset(fl.p, 0, potrf_dag_flag_elems(n) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&parts.p, (size_t)T * 8);
    if (e == hipSuccess) e = hipMalloc(&info.p, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&xd.p, (size_t)T * 64 * 64 * 8);
    if (e == hipSuccess) e = hipMalloc(&tr.p, (size_t)(1 + 8 * cap) * 8);
    if (e == hipSuccess) e = hipMemset(tr.p, 0, 8);
    if (e != hipSuccess) return set_err(nullptr, RGF_ENOMEM, "trace_potrf: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(bench_fill_kernel, dim3(1024), dim3(256), 0, nullptr, (double*)a.p, (long long)na, 0.001);
    int rc = mirror_lower((double*)a.p, n, n, nullptr);
    hipLaunchKernelGGL(diag_shift_kernel, dim3((n + 255) / 256), dim3(256), 0, nullptr, (double*)a.p, n, (double)n);
    const int old = g_chol_dag;
    g_chol_dag = 1;
    g_chol_dag_trace = (long long*)tr.p;
    g_chol_dag_trace_cap = cap;
    bool x_done = false;
    if (!rc) rc = potrf_f64((double*)a.p, n, 0, n, 1, (double*)parts.p, 0, (int*)info.p, 0, (double*)xd.p, 0, false,
                            nullptr, (int*)fl.p, (double*)xb.p, 0, &x_done);
    g_chol_dag_trace = nullptr;
    g_chol_dag_trace_cap = 0;
    g_chol_dag = old;
    if (rc) return dev_status(nullptr, rc, "trace_potrf");
    long long cntr = 0;
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(&cntr, tr.p, 8, hipMemcpyDeviceToHost);
    const long long got = std::min(cntr, cap);
    if (e == hipSuccess && got > 0) e = hipMemcpy(out, (long long*)tr.p + 1, (size_t)got * 64, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_err(nullptr, RGF_EDEVICE, "trace_potrf: %s", hipGetErrorString(e));
    *nrec = cntr;
    int khz = 100000;
    hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0);
    if (ticks_per_us) *ticks_per_us = khz / 1000.0;
    return RGF_OK;
}

// ------------------------------------------------------------------ handle groups
// Handles of any shapes, batch counts and precisions on one device whose sweeps run as one
// HIP graph: each handle's sweep is a branch on its own stream, forked from and joined to
// the group's stream, so the sweeps of uneven matrices overlap on the device and the host
// issues one launch per group.  Captured on the first call (and again when the kernel-timing
// configuration or the selected inversion's outputs change), replayed afterwards: the
// handles' device buffers never move.
struct rgf_group_s {
    std::vector<rgf_handle_s*> hs;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, after = nullptr;
    std::vector<hipEvent_t> before, joins;   // one per handle
    hipGraphExec_t gx_factor = nullptr, gx_selinv = nullptr;
    long long ep_factor = -1, ep_selinv = -1;
    std::vector<double*> sel_out;            // outputs the captured selected inversion writes
};

namespace {
// a failed group call: its message (set on the handle concerned) also becomes rgf_last_error(NULL)
int group_fail(rgf_handle_s* h, int rc)
{
    if (rc != RGF_OK && h) g_create_error = h->err;
    return rc;
}

template <class Fn>
int group_run(rgf_group_s* g, hipGraphExec_t& gx, long long& epoch, Fn&& branch)
{
    const long long ep = timing_epoch();
    if (gx && epoch != ep) {
        hipGraphExecDestroy(gx);
        gx = nullptr;
    }
    const size_t n = g->hs.size();
    hipError_t e = hipSuccess;
    // the group starts after the work already queued on every handle's stream
    for (size_t q = 0; q < n && e == hipSuccess; q++) {
        e = hipEventRecord(g->before[q], g->hs[q]->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(g->stream, g->before[q], 0);
    }
    if (e != hipSuccess) return (int)e;
    if (!gx) {
        hipGraph_t graph = nullptr;
        if ((e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) return (int)e;
        int rc = (int)hipEventRecord(g->fork, g->stream);
        for (size_t q = 0; q < n && rc == 0; q++) {
            rgf_handle_s* h = g->hs[q];
            rc = (int)hipStreamWaitEvent(h->stream, g->fork, 0);
            if (rc == 0) rc = branch(h, q);
            // every stream that joined the capture is joined back, also after a failure
            const hipError_t j = hipEventRecord(g->joins[q], h->stream);
            const hipError_t w = j == hipSuccess ? hipStreamWaitEvent(g->stream, g->joins[q], 0) : j;
            if (rc == 0) rc = (int)w;
        }
        e = hipStreamEndCapture(g->stream, &graph);
        if (rc || e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return rc ? rc : (int)e;
        }
        e = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) {
            gx = nullptr;
            return (int)e;
        }
        epoch = ep;
    }
    if ((e = hipGraphLaunch(gx, g->stream)) != hipSuccess) return (int)e;
    // every handle's stream continues after the group
    if ((e = hipEventRecord(g->after, g->stream)) != hipSuccess) return (int)e;
    for (size_t q = 0; q < n && e == hipSuccess; q++) e = hipStreamWaitEvent(g->hs[q]->stream, g->after, 0);
    return (int)e;
}
}  // namespace

extern "C" int rgf_group_create(const rgf_handle_t* handles, size_t n, rgf_group_t* out)
{
    if (!out) return set_err(nullptr, RGF_EINVAL, "group: null output");
    *out = nullptr;
    if (!handles || n == 0) return set_err(nullptr, RGF_EINVAL, "group: no handles");
    std::vector<rgf_handle_s*> hs(handles, handles + n);
    for (size_t q = 0; q < n; q++) {
        if (!hs[q]) return set_err(nullptr, RGF_EINVAL, "group: handle %zu is null", q);
        if (hs[q]->device != hs[0]->device) return set_err(nullptr, RGF_EINVAL, "group: handles on different devices");
        for (size_t p = 0; p < q; p++)
            if (hs[p] == hs[q]) return set_err(nullptr, RGF_EINVAL, "group: handle %zu given twice", q);
    }
    auto* g = new rgf_group_s();
    g->hs = hs;
    g->device = hs[0]->device;
    g->before.assign(n, nullptr);
    g->joins.assign(n, nullptr);
    hipSetDevice(g->device);
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->after, hipEventDisableTiming);
    for (size_t q = 0; q < n && e == hipSuccess; q++) {
        e = hipEventCreateWithFlags(&g->before[q], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&g->joins[q], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        rgf_group_destroy(g);
        return set_err(nullptr, RGF_EDEVICE, "group: %s", hipGetErrorString(e));
    }
    *out = g;
    return RGF_OK;
}

extern "C" int rgf_group_destroy(rgf_group_t g)
{
    if (!g) return RGF_OK;
    hipSetDevice(g->device);
    if (g->stream) hipStreamSynchronize(g->stream);
    if (g->gx_factor) hipGraphExecDestroy(g->gx_factor);
    if (g->gx_selinv) hipGraphExecDestroy(g->gx_selinv);
    for (hipEvent_t ev : g->before)
        if (ev) hipEventDestroy(ev);
    for (hipEvent_t ev : g->joins)
        if (ev) hipEventDestroy(ev);
    if (g->fork) hipEventDestroy(g->fork);
    if (g->after) hipEventDestroy(g->after);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return RGF_OK;
}

extern "C" int rgf_group_factorize(rgf_group_t g)
{
    if (!g) return set_err(nullptr, RGF_EINVAL, "group: null group");
    hipSetDevice(g->device);
    for (rgf_handle_s* h : g->hs) {
        if (h->broken)
            return group_fail(h, set_err(h, RGF_ESTATE, "factorization failed earlier and the matrix was not kept"));
        if (!h->Ap && h->factorized)
            return group_fail(h, set_err(h, RGF_ESTATE,
                                         "factorized in place already: regenerate the matrix before a group factorize"));
        int rc;
        if (h->nd > 0 && h->nt > 1 && (rc = ensure_solve_buffers(h))) return group_fail(h, rc);   // arrow products use `work`
        h->w_valid = false;
    }
    int rc = group_run(g, g->gx_factor, g->ep_factor, [](rgf_handle_s* h, size_t) { return factor_sweep(h); });
    if (rc) return group_fail(g->hs[0], dev_status(g->hs[0], rc, "group factorize"));
    int first = RGF_OK;
    for (rgf_handle_s* h : g->hs) {
        const int r = check_factor_info(h);
        if (r != RGF_OK && first == RGF_OK) first = group_fail(h, r);
    }
    return first;
}

extern "C" int rgf_group_selinv_diag_device(rgf_group_t g, double* const* diag_devs)
{
    if (!g || !diag_devs) return set_err(nullptr, RGF_EINVAL, "group: null argument");
    hipSetDevice(g->device);
    for (size_t q = 0; q < g->hs.size(); q++) {
        rgf_handle_s* h = g->hs[q];
        if (!diag_devs[q]) return set_err(nullptr, RGF_EINVAL, "group: null output for handle %zu", q);
        int rc = ensure_factorized(h);
        if (!rc) rc = ensure_selinv_buffers(h);
        if (rc) return group_fail(h, rc);
    }
    std::vector<double*> out(diag_devs, diag_devs + g->hs.size());
    if (g->gx_selinv && out != g->sel_out) {
        hipGraphExecDestroy(g->gx_selinv);
        g->gx_selinv = nullptr;
    }
    g->sel_out = out;
    int rc = group_run(g, g->gx_selinv, g->ep_selinv,
                       [&out](rgf_handle_s* h, size_t q) { return selinv_sweep(h, out[q]); });
    if (rc) return group_fail(g->hs[0], dev_status(g->hs[0], rc, "group selinv"));
    for (rgf_handle_s* h : g->hs) h->w_valid = h->Wall != nullptr;
    return RGF_OK;
}

extern "C" int rgf_group_synchronize(rgf_group_t g)
{
    if (!g) return set_err(nullptr, RGF_EINVAL, "group: null group");
    hipSetDevice(g->device);
    const hipError_t e = hipStreamSynchronize(g->stream);
    return e == hipSuccess ? RGF_OK : group_fail(g->hs[0], dev_status(g->hs[0], (int)e, "group synchronize"));
}
