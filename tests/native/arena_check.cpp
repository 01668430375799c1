// CPU check of the device arena's bookkeeping (rgf_amd/csrc/rgf_arena.hpp) with a host
// backend: random allocate / free sequences against an independent record of the live
// ranges.  Exit status 0 when every check holds; the first failure is printed.
#include "rgf_arena.hpp"
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <vector>

namespace {
size_t g_limit = 0;          // bytes the backend may hold (0: unlimited)
size_t g_held = 0;
std::map<void*, size_t> g_slabs;
long long g_grow_calls = 0, g_release_calls = 0;

struct HostSlabs {
    static int grow(void** p, size_t n)
    {
        g_grow_calls++;
        if (g_limit && g_held + n > g_limit) return 2;   // out of memory
        void* q = std::aligned_alloc(4096, n);   // device slabs are at least 256-byte aligned
        if (!q) return 2;
        *p = q;
        g_slabs[q] = n;
        g_held += n;
        return 0;
    }
    static bool oom(int rc) { return rc == 2; }
    static void release(void* p)
    {
        g_release_calls++;
        auto it = g_slabs.find(p);
        if (it == g_slabs.end()) { std::printf("release of an unknown slab\n"); std::exit(1); }
        g_held -= it->second;
        g_slabs.erase(it);
        std::free(p);
    }
};

#define CHECK(c, ...)                                                     \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("check failed (%s:%d): ", __FILE__, __LINE__);    \
            std::printf(__VA_ARGS__);                                     \
            std::printf("\n");                                            \
            std::exit(1);                                                 \
        }                                                                 \
    } while (0)

using A = rgf::Arena<HostSlabs>;
constexpr size_t MiB = size_t(1) << 20, GiB = size_t(1) << 30;

// every live range lies inside one backend slab and overlaps no other
void check_live(const std::map<uintptr_t, size_t>& live)
{
    uintptr_t end = 0;
    for (auto& kv : live) {
        CHECK(kv.first % A::kAlign == 0, "misaligned block %zx", (size_t)kv.first);
        CHECK(kv.first >= end, "overlap at %zx", (size_t)kv.first);
        end = kv.first + kv.second;
        bool inside = false;
        for (auto& s : g_slabs) {
            const uintptr_t b = (uintptr_t)s.first;
            if (kv.first >= b && end <= b + s.second) inside = true;
        }
        CHECK(inside, "block outside every slab");
    }
}
}  // namespace

int main()
{
    // 1. small blocks share one slab; freeing them all coalesces it into one free range
    {
        A a;
        std::vector<void*> ps;
        std::map<uintptr_t, size_t> live;
        for (int i = 0; i < 200; i++) {
            void* p;
            size_t got;
            CHECK(a.alloc(1 + 37 * i, &p, &got) == 0, "alloc");
            CHECK(got >= 1 + 37 * (size_t)i && got % A::kAlign == 0, "rounded size");
            ps.push_back(p);
            live[(uintptr_t)p] = got;
        }
        check_live(live);
        CHECK(a.slab_count() == 1, "one slab for small blocks, got %zu", a.slab_count());
        std::shuffle(ps.begin(), ps.end(), std::mt19937(3));
        for (void* p : ps) CHECK(a.free(p, ~size_t(0)), "free");
        CHECK(a.free_ranges() == 1 && a.free_bytes() == a.slab_bytes(), "not coalesced: %zu ranges", a.free_ranges());
        int dummy;
        CHECK(!a.free(&dummy, ~size_t(0)), "foreign pointer accepted");
    }
    // 2. random mixed sizes (bytes .. a few GiB) with interleaved frees; reuse without growth
    //    once the working set is held; cap trimming releases only wholly free slabs
    {
        const size_t held0 = g_held;   // slabs of the first arena (an arena keeps its slabs until exit)
        A a;
        std::mt19937_64 rng(7);
        std::map<uintptr_t, size_t> live;
        std::vector<std::pair<void*, size_t>> held;
        for (int it = 0; it < 20000; it++) {
            const bool do_free = !held.empty() && (held.size() > 60 || rng() % 2);
            if (do_free) {
                const size_t k = rng() % held.size();
                CHECK(a.free(held[k].first, 8 * GiB), "free");
                live.erase((uintptr_t)held[k].first);
                held[k] = held.back();
                held.pop_back();
            } else {
                const int cls = rng() % 10;
                const size_t n = cls < 6 ? 1 + rng() % (4 * MiB) : cls < 9 ? MiB + rng() % (300 * MiB)
                                                                          : GiB + rng() % (2 * GiB);
                void* p;
                size_t got;
                CHECK(a.alloc(n, &p, &got) == 0, "alloc %zu", n);
                live[(uintptr_t)p] = got;
                held.push_back({p, got});
            }
            if (it % 500 == 0) check_live(live);
            CHECK(a.free_bytes() + [&] { size_t t = 0; for (auto& kv : live) t += kv.second; return t; }() ==
                      a.slab_bytes(), "free + live != slab bytes");
        }
        check_live(live);
        const long long grows = a.stats().grows;
        // the same working set again: served from the slabs already held
        std::vector<std::pair<void*, size_t>> again;
        for (auto& h : held) { CHECK(a.free(h.first, ~size_t(0)), "free"); }
        live.clear();
        for (auto& h : held) {
            void* p;
            size_t got;
            CHECK(a.alloc(h.second, &p, &got) == 0, "realloc");
            live[(uintptr_t)p] = got;
            again.push_back({p, got});
        }
        check_live(live);
        CHECK(a.stats().grows == grows, "grew while re-allocating a working set it held");
        for (auto& h : again) CHECK(a.free(h.first, 0), "free");
        CHECK(a.slab_count() == 0 && a.free_bytes() == 0 && g_held == held0, "cap 0 keeps %zu slabs, %zu bytes",
              a.slab_count(), g_held - held0);
    }
    // 3. out of memory: wholly free slabs are given back and the request retried
    {
        const size_t held0 = g_held;
        g_limit = held0 + 3 * GiB;
        A a;
        void *p1, *p2, *p3;
        size_t got;
        CHECK(a.alloc(GiB + MiB, &p1, &got) == 0, "first");        // slab 1: 1 GiB + 2 MiB
        CHECK(a.free(p1, ~size_t(0)), "free");                       // kept (no cap)
        CHECK(a.alloc(2 * GiB, &p2, &got) == 0, "needs the kept slab back");
        CHECK(a.stats().releases == 1, "the free slab was released for the retry");
        CHECK(a.alloc(2 * GiB, &p3, &got) != 0, "over the limit must fail");
        CHECK(a.free(p2, 0), "free");
        CHECK(g_held == held0, "everything released at cap 0");
        g_limit = 0;
    }
    std::printf("arena checks passed (backend grows %lld, releases %lld)\n", g_grow_calls, g_release_calls);
    return 0;
}
