// Per-device memory arena of librgf_hip.so (host code only; no HIP types, so the
// bookkeeping is tested on the CPU with a host backend).
//
// Slabs obtained from the backend (at least kSlabMin bytes, or the size of a larger
// request, rounded to kSlabGranule) are sub-allocated best-fit at kAlign alignment;
// freed ranges are merged with their free neighbours.  Free bytes are kept up to a cap;
// beyond it slabs that are wholly free are released, largest first.  When the backend
// reports out of memory every wholly free slab is released and the request retried once.
//
// Backend (static members):
//   int  grow(void** p, size_t n)   0 on success, else an error code
//   bool oom(int rc)                rc means out of memory
//   void release(void* p)           give a slab back (may synchronize the device: it is
//                                   called with the arena's lock released)
#pragma once
#include <algorithm>
#include <cstddef>
#include <map>
#include <mutex>
#include <vector>

namespace rgf {

template <class Backend>
class Arena {
public:
    static constexpr size_t kAlign = 256;
    static constexpr size_t kSlabMin = size_t(1) << 30;
    static constexpr size_t kSlabGranule = size_t(2) << 20;
    struct Stats {
        long long hits = 0;      // allocations served without growing
        long long grows = 0;     // backend grow calls
        long long releases = 0;  // slabs given back
    };

    ~Arena()
    {
        for (Slab* s : slabs_) delete s;
    }

    // `bytes` (>= 1) at kAlign alignment; *got = bytes rounded up to kAlign.
    int alloc(size_t bytes, void** out, size_t* got)
    {
        const size_t want = (std::max<size_t>(bytes, 1) + kAlign - 1) / kAlign * kAlign;
        *got = want;
        std::unique_lock<std::mutex> lk(mu_);
        if (char* p = take(want)) {
            stats_.hits++;
            *out = p;
            return 0;
        }
        const size_t ssz = (std::max(want, kSlabMin) + kSlabGranule - 1) / kSlabGranule * kSlabGranule;
        void* base = nullptr;
        stats_.grows++;
        int rc = Backend::grow(&base, ssz);
        if (rc != 0 && Backend::oom(rc)) {
            std::vector<void*> drop;
            trim(0, drop);
            lk.unlock();
            for (void* q : drop) Backend::release(q);
            lk.lock();
            if (char* p = take(want)) {   // another thread may have freed meanwhile
                stats_.hits++;
                *out = p;
                return 0;
            }
            stats_.grows++;
            rc = Backend::grow(&base, ssz);
        }
        if (rc != 0) return rc;
        Slab* s = new Slab{(char*)base, ssz, 0, {}};
        slabs_.push_back(s);
        put(s, 0, ssz);
        *out = take(want);
        return 0;
    }

    // p from alloc and no longer in use; returns false (and does nothing) for a pointer
    // the arena did not hand out
    bool free(void* p, size_t cap)
    {
        std::vector<void*> drop;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = live_.find((char*)p);
            if (it == live_.end()) return false;
            Slab* s = it->second.first;
            const size_t n = it->second.second;
            live_.erase(it);
            put(s, (size_t)((char*)p - s->base), n);
            if (free_total_ > cap) trim(cap, drop);
        }
        for (void* q : drop) Backend::release(q);
        return true;
    }

    // release wholly free slabs, largest first, while more than `keep` bytes are free
    void trim_to(size_t keep)
    {
        std::vector<void*> drop;
        {
            std::lock_guard<std::mutex> lk(mu_);
            trim(keep, drop);
        }
        for (void* q : drop) Backend::release(q);
    }

    size_t free_bytes() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return free_total_;
    }
    size_t slab_bytes() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        size_t t = 0;
        for (const Slab* s : slabs_) t += s->size;
        return t;
    }
    size_t slab_count() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return slabs_.size();
    }
    // free ranges over all slabs (1 per wholly free slab once everything is freed)
    size_t free_ranges() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return by_size_.size();
    }
    Stats stats() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return stats_;
    }

private:
    struct Slab {
        char* base;
        size_t size;
        size_t free_bytes;
        std::map<size_t, size_t> free_off;   // offset -> length of each free range
    };

    void unindex(Slab* s, size_t off, size_t n)
    {
        auto range = by_size_.equal_range(n);
        for (auto it = range.first; it != range.second; ++it)
            if (it->second.first == s && it->second.second == off) {
                by_size_.erase(it);
                return;
            }
    }

    // return [off, off + n) of s to the free lists, merged with free neighbours
    void put(Slab* s, size_t off, size_t n)
    {
        s->free_bytes += n;
        free_total_ += n;
        auto next = s->free_off.lower_bound(off);
        if (next != s->free_off.end() && next->first == off + n) {
            unindex(s, next->first, next->second);
            n += next->second;
            next = s->free_off.erase(next);
        }
        if (next != s->free_off.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                unindex(s, prev->first, prev->second);
                off = prev->first;
                n += prev->second;
                s->free_off.erase(prev);
            }
        }
        s->free_off[off] = n;
        by_size_.emplace(n, std::make_pair(s, off));
    }

    // best-fit range of `want` bytes, or null
    char* take(size_t want)
    {
        auto it = by_size_.lower_bound(want);
        if (it == by_size_.end()) return nullptr;
        Slab* s = it->second.first;
        const size_t off = it->second.second, n = it->first;
        by_size_.erase(it);
        s->free_off.erase(off);
        if (n > want) {   // the remainder's neighbours are allocated (free ones were merged)
            s->free_off[off + want] = n - want;
            by_size_.emplace(n - want, std::make_pair(s, off + want));
        }
        s->free_bytes -= want;
        free_total_ -= want;
        char* p = s->base + off;
        live_[p] = std::make_pair(s, want);
        return p;
    }

    // detach wholly free slabs, largest first, while more than `keep` bytes are free
    void trim(size_t keep, std::vector<void*>& drop)
    {
        std::vector<Slab*> empty;
        for (Slab* s : slabs_)
            if (s->free_bytes == s->size) empty.push_back(s);
        std::sort(empty.begin(), empty.end(), [](const Slab* x, const Slab* y) { return x->size > y->size; });
        for (Slab* s : empty) {
            if (free_total_ <= keep) break;
            unindex(s, 0, s->size);
            free_total_ -= s->size;
            slabs_.erase(std::find(slabs_.begin(), slabs_.end(), s));
            drop.push_back(s->base);
            stats_.releases++;
            delete s;
        }
    }

    mutable std::mutex mu_;
    std::vector<Slab*> slabs_;
    std::multimap<size_t, std::pair<Slab*, size_t>> by_size_;   // free: length -> (slab, offset)
    std::map<char*, std::pair<Slab*, size_t>> live_;            // allocation -> (slab, length)
    size_t free_total_ = 0;
    Stats stats_;
};

}  // namespace rgf
