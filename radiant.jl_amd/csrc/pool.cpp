// pool.cpp -- cache of device allocations for problems that come and go (served requests,
// batches, one-shot solves).
//
// hipFree synchronises the whole device and hipMalloc of large blocks maps pages: with several
// requests in flight on their own streams, every problem created or destroyed stalled the others
// (on the 3D request stream, 78 allocations and frees per request took a fifth of the serving
// threads' time).  Blocks released here stay mapped and are handed to the next allocation of a
// fitting size.  A caller releases a block only once no device work that uses it is pending (a
// problem synchronises its stream before it is destroyed), so a cached block can be reused on any
// stream at once.
//
// Sizes are rounded to classes (powers of two below 1 MiB; above, steps of 1/8 of the power of two
// below, at least 2 MiB), and an allocation takes the smallest cached block of at least its class
// and at most twice it.  Cached bytes above a cap (RADIANT_HIP_POOL_GIB, default 64 GiB per device)
// are freed, least recently cached first, and when hipMalloc fails every cached block is freed
// before the one retry.  RADIANT_HIP_POOL=0 turns the cache off (plain hipMalloc / hipFree).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

#include "engine.hpp"

namespace rt {

namespace {

struct DevCache {
    std::mutex mu;
    std::multimap<size_t, std::pair<uint64_t, void*>> free_;  // cached blocks by size: (release tick, block)
    std::unordered_map<void*, size_t> live;                   // allocated blocks and their sizes
    size_t cached = 0;
    uint64_t tick = 0;
};

DevCache& cache_of(int dev) {
    static std::mutex mu;
    static std::map<int, DevCache*> all;
    std::lock_guard<std::mutex> lk(mu);
    DevCache*& c = all[dev];
    if (!c) c = new DevCache();  // lives until the process ends
    return *c;
}

bool pool_on() {
    static const bool on = [] {
        const char* e = std::getenv("RADIANT_HIP_POOL");
        return !(e && e[0] == '0');
    }();
    return on;
}

size_t pool_cap() {
    static const size_t cap = [] {
        const char* e = std::getenv("RADIANT_HIP_POOL_GIB");
        const double g = e ? std::atof(e) : 64.0;
        return (size_t)((g > 0 ? g : 0.0) * 1073741824.0);
    }();
    return cap;
}

size_t size_class(size_t b) {
    if (b < 256) b = 256;
    if (b < (1u << 20)) {
        size_t c = 256;
        while (c < b) c <<= 1;
        return c;
    }
    size_t p = 1u << 20;
    while (p * 2 <= b) p <<= 1;
    size_t step = p / 8;
    if (step < (2u << 20)) step = 2u << 20;
    return (b + step - 1) / step * step;
}

int current_device() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

// frees cached blocks, least recently released first, until at most `keep` bytes stay cached
// (lock held)
void shrink(DevCache& c, size_t keep) {
    while (c.cached > keep && !c.free_.empty()) {
        auto old = c.free_.begin();
        for (auto it = c.free_.begin(); it != c.free_.end(); ++it)
            if (it->second.first < old->second.first) old = it;
        c.cached -= old->first;
        (void)hipFree(old->second.second);
        c.free_.erase(old);
    }
}

}  // namespace

void* dev_alloc(size_t bytes) {
    if (!pool_on()) {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(RT_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        }
        return p;
    }
    DevCache& c = cache_of(current_device());
    const size_t cls = size_class(bytes);
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.free_.lower_bound(cls);
    if (it != c.free_.end() && it->first <= 2 * cls) {
        void* p = it->second.second;
        c.live[p] = it->first;
        c.cached -= it->first;
        c.free_.erase(it);
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        shrink(c, 0);
        e = hipMalloc(&p, cls);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(RT_ENOMEM, "hipMalloc of " + std::to_string(cls) + " bytes failed: " + hipGetErrorString(e));
        }
    }
    c.live[p] = cls;
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    if (!pool_on()) {
        (void)hipFree(p);
        return;
    }
    DevCache& c = cache_of(current_device());
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) {  // not from the cache
        (void)hipFree(p);
        return;
    }
    const size_t sz = it->second;
    c.live.erase(it);
    c.free_.emplace(sz, std::make_pair(++c.tick, p));
    c.cached += sz;
    if (c.cached > pool_cap()) shrink(c, pool_cap());
}

size_t dev_cached_bytes() {
    DevCache& c = cache_of(current_device());
    std::lock_guard<std::mutex> lk(c.mu);
    return c.cached;
}

void dev_trim() {
    DevCache& c = cache_of(current_device());
    std::lock_guard<std::mutex> lk(c.mu);
    shrink(c, 0);
}

}  // namespace rt
