// Tile-shape autotuning of the product launches (FP64 and pre-split split-BF16 GEMMs).
//
// A launch's signature is what its cost depends on: the transposes and, for each of its
// descriptors, m, n, k, the triangle flags and the batch entries.  The launchers pass
// their candidate tile shapes (the size-threshold choice first); with tuning on, a
// signature seen for the first time has every candidate timed with HIP events on scratch
// outputs -- the real operands are only read, the real outputs untouched -- and the
// fastest is cached for every later launch of that signature.  Every tile shape of one
// kernel family sums each output over the same k-blocks in the same order, so the choice
// changes timings only, never a result bit.
#include "rgf_internal.hpp"
#include "../../include/rgf.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace rgf {

namespace {
std::mutex g_tune_mu;
int g_tune_mode = 0;   // 0 thresholds, 1 cached choices, 2 cached choices + tune misses
struct Choice {
    int cfg;
    double ms;
};
std::map<std::vector<long long>, Choice> g_tune;
thread_local bool t_trial = false;   // inside a trial launch: no tuning, no kernel timing
constexpr size_t TUNE_SCRATCH_MAX = size_t(4) << 30;   // bytes of scratch outputs per tuning
constexpr int TUNE_REPS = 3;

std::vector<long long> signature(int kind, Op ta, Op tb, const GemmDesc* d, int nd)
{
    std::vector<long long> v{kind, ta == Op::T, tb == Op::T, 0};
    int used = 0;
    for (int q = 0; q < nd; q++) {
        if (d[q].m <= 0 || d[q].n <= 0 || d[q].count <= 0) continue;
        used++;
        v.insert(v.end(), {(long long)d[q].m, (long long)d[q].n, (long long)d[q].k, (long long)d[q].flags,
                           (long long)d[q].count * std::max(d[q].count2, 1)});
    }
    v[3] = used;
    return v;
}
}  // namespace

bool tune_in_trial() { return t_trial; }

int tune_choose(int kind, Op ta, Op tb, const GemmDesc* d, int nd, const int* cands, int ncand, hipStream_t s,
                const std::function<int(int, const GemmDesc*)>& run)
{
    if (t_trial) return -1;
    std::vector<long long> key;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_tune_mode == 0) return -1;
        key = signature(kind, ta, tb, d, nd);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return it->second.cfg;
        if (g_tune_mode != 2 || ncand < 2) return -1;
    }
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return -1;

    // compact scratch outputs: entry e of descriptor q at m * n * e (ld m); Cin copied
    // there first when the product reads it (it may be the real C)
    GemmDesc td[GEMM_MAXD];
    size_t elems = 0;
    for (int q = 0; q < nd; q++) {
        const long long entries = (long long)std::max(d[q].count, 0) * std::max(d[q].count2, 1);
        const size_t one = (size_t)std::max(d[q].m, 0) * (size_t)std::max(d[q].n, 0) * (size_t)entries;
        elems += one * (d[q].beta != 0.0 ? 2 : 1);
    }
    if (elems == 0 || elems * sizeof(double) > TUNE_SCRATCH_MAX) return -1;
    double* scratch = nullptr;
    if (hipMalloc(&scratch, elems * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    size_t off = 0;
    int rc = 0;
    for (int q = 0; q < nd && rc == 0; q++) {
        td[q] = d[q];
        GemmDesc& t = td[q];
        if (t.count2 < 1) t.count2 = 1;
        if (t.m <= 0 || t.n <= 0 || t.count <= 0) continue;
        const long long mn = (long long)t.m * t.n;
        t.C = scratch + off;
        t.ldc = t.m;
        t.sC = mn;
        t.sC2 = mn * t.count;
        off += (size_t)mn * t.count * t.count2;
        if (d[q].beta != 0.0) {
            double* cin = scratch + off;
            off += (size_t)mn * t.count * t.count2;
            for (int b2 = 0; b2 < t.count2 && rc == 0; b2++)
                rc = copy2d(d[q].Cin + (long long)b2 * d[q].sCin2, d[q].ldcin, cin + (long long)b2 * mn * t.count, t.m,
                            t.m, t.n, 1.0, s, t.count, d[q].sCin, mn);
            t.Cin = cin;
        } else {
            t.Cin = t.C;
        }
        t.ldcin = t.m;
        t.sCin = mn;
        t.sCin2 = mn * t.count;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = -1;
    int best = -1;
    double best_ms = 0.0, first_ms = 0.0;
    t_trial = true;
    for (int c = 0; c < ncand && rc == 0; c++) {
        rc = run(cands[c], td);   // warm-up (first touch of the scratch, code load)
        if (rc) break;
        if ((rc = (int)hipEventRecord(e0, s))) break;
        for (int r = 0; r < TUNE_REPS && rc == 0; r++) rc = run(cands[c], td);
        if (rc || (rc = (int)hipEventRecord(e1, s)) || (rc = (int)hipEventSynchronize(e1))) break;
        float ms = 0.f;
        if ((rc = (int)hipEventElapsedTime(&ms, e0, e1))) break;
        const double t = ms / TUNE_REPS;
        if (c == 0) first_ms = t;
        if (best < 0 || t < best_ms) {
            best = cands[c];
            best_ms = t;
        }
    }
    t_trial = false;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipFree(scratch);
    if (rc || best < 0) {
        (void)hipGetLastError();
        return -1;
    }
    // the threshold choice stays unless another shape is at least 2 % faster
    if (best != cands[0] && best_ms > 0.98 * first_ms) {
        best = cands[0];
        best_ms = first_ms;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = Choice{best, best_ms};
    return best;
}

}  // namespace rgf

extern "C" int rgf_set_autotune(int mode, int* previous)
{
    if (mode < 0 || mode > 2) return RGF_EINVAL;
    std::lock_guard<std::mutex> lk(rgf::g_tune_mu);
    if (previous) *previous = rgf::g_tune_mode;
    rgf::g_tune_mode = mode;
    return RGF_OK;
}

extern "C" int rgf_autotune_entries(void)
{
    std::lock_guard<std::mutex> lk(rgf::g_tune_mu);
    return (int)rgf::g_tune.size();
}

extern "C" int rgf_autotune_clear(void)
{
    std::lock_guard<std::mutex> lk(rgf::g_tune_mu);
    rgf::g_tune.clear();
    return RGF_OK;
}

// Text file: a header line, then one line per signature: its length, the signature
// values, the chosen configuration and the time of one launch with it (ms).
extern "C" int rgf_autotune_save(const char* path)
{
    if (!path) return RGF_EINVAL;
    std::lock_guard<std::mutex> lk(rgf::g_tune_mu);
    FILE* f = fopen(path, "w");
    if (!f) return RGF_EINVAL;
    fprintf(f, "rgf-autotune 1 gfx950\n");
    for (const auto& kv : rgf::g_tune) {
        fprintf(f, "%zu", kv.first.size());
        for (long long v : kv.first) fprintf(f, " %lld", v);
        fprintf(f, " %d %.6f\n", kv.second.cfg, kv.second.ms);
    }
    const bool ok = fclose(f) == 0;
    return ok ? RGF_OK : RGF_EINVAL;
}

extern "C" int rgf_autotune_load(const char* path)
{
    if (!path) return RGF_EINVAL;
    FILE* f = fopen(path, "r");
    if (!f) return RGF_EINVAL;
    char head[64] = {0};
    int version = 0;
    if (fscanf(f, "%63s %d %*s", head, &version) != 2 || strcmp(head, "rgf-autotune") != 0 || version != 1) {
        fclose(f);
        return RGF_EINVAL;
    }
    std::map<std::vector<long long>, rgf::Choice> got;
    size_t len = 0;
    int rc = RGF_OK;
    while (fscanf(f, "%zu", &len) == 1) {
        if (len < 4 || len > 4 + 5 * rgf::GEMM_MAXD) {
            rc = RGF_EINVAL;
            break;
        }
        std::vector<long long> key(len);
        for (size_t i = 0; i < len && rc == RGF_OK; i++)
            if (fscanf(f, "%lld", &key[i]) != 1) rc = RGF_EINVAL;
        rgf::Choice c{};
        if (rc == RGF_OK && fscanf(f, "%d %lf", &c.cfg, &c.ms) != 2) rc = RGF_EINVAL;
        // configurations each launcher accepts: FP64 0..3, split-BF16 0..2
        if (rc == RGF_OK && (key[0] < 0 || key[0] > 1 || c.cfg < 0 || c.cfg > (key[0] == 0 ? 3 : 2) ||
                             key[3] * 5 + 4 != (long long)len))
            rc = RGF_EINVAL;
        if (rc != RGF_OK) break;
        got[key] = c;
    }
    if (rc == RGF_OK && !feof(f)) rc = RGF_EINVAL;
    fclose(f);
    if (rc != RGF_OK) return rc;
    std::lock_guard<std::mutex> lk(rgf::g_tune_mu);
    for (const auto& kv : got) rgf::g_tune[kv.first] = kv.second;
    return RGF_OK;
}
