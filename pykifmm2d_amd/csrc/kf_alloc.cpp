// kf_alloc.cpp -- device memory of the library's internal buffers (DevBuf), with a cache of
// freed blocks per device.
//
// A tree and plan build allocates and frees about 140 buffers (work lists, sort keys, counts,
// scan tiles), and hipMalloc / hipFree took about half of the 1M-point build.  Freed blocks of
// up to 64 MiB are kept per device in size classes (a quarter of a power of two, at least
// 256 bytes, so a block wastes at most a fifth of itself) and handed out again for requests of
// the same class.  A freed block may still be read or written by kernels queued before its
// release, so the first reuse after any release synchronises the device (what hipFree itself
// does on every call).  The cache holds at most 512 MiB per device; beyond that, and for larger
// blocks, memory goes back with hipFree.  A failing hipMalloc empties the cache and tries once
// more.  The cache lives until the process ends.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "kf_internal.h"

namespace kf {
namespace {

constexpr size_t KF_CACHE_MAX_BLOCK = size_t(64) << 20;
constexpr size_t KF_CACHE_MAX_BYTES = size_t(512) << 20;

struct DeviceCache {
    std::map<size_t, std::vector<void*>> free;  // class bytes -> blocks
    size_t held = 0;                            // bytes in `free`
    bool dirty = false;                         // a block was released since the last synchronisation
};

struct Cache {
    std::mutex mu;
    std::unordered_map<int, DeviceCache> dev;
    std::unordered_map<void*, std::pair<int, size_t>> live;  // cached-class block -> (device, class)
};

Cache& cache() {
    static Cache* c = new Cache();  // never destroyed: blocks outlive static destructors
    return *c;
}

size_t class_of(size_t bytes) {
    size_t b = bytes < 256 ? 256 : bytes;
    size_t top = 1;
    while ((top << 1) <= b) top <<= 1;
    const size_t step = top / 4 < 256 ? 256 : top / 4;
    return (b + step - 1) / step * step;
}

// Returns every cached block of device d to HIP (the caller holds the lock).
void flush(int d, DeviceCache& dc) {
    if (dc.held == 0) return;
    (void)hipDeviceSynchronize();
    for (auto& kv : dc.free)
        for (void* p : kv.second) (void)hipFree(p);
    dc.free.clear();
    dc.held = 0;
    dc.dirty = false;
    (void)d;
}

}  // namespace

void* dev_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    int d = 0;
    KF_HIP_CHECK(hipGetDevice(&d));
    Cache& c = cache();
    if (bytes > KF_CACHE_MAX_BLOCK) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            std::lock_guard<std::mutex> g(c.mu);
            flush(d, c.dev[d]);
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) throw KfError{KF_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e)};
        return p;
    }
    const size_t cls = class_of(bytes);
    std::lock_guard<std::mutex> g(c.mu);
    DeviceCache& dc = c.dev[d];
    auto it = dc.free.find(cls);
    if (it != dc.free.end() && !it->second.empty()) {
        if (dc.dirty) {  // kernels queued before a release may still use the block
            KF_HIP_CHECK(hipDeviceSynchronize());
            dc.dirty = false;
        }
        void* p = it->second.back();
        it->second.pop_back();
        dc.held -= cls;
        c.live[p] = {d, cls};
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        flush(d, dc);
        e = hipMalloc(&p, cls);
    }
    if (e != hipSuccess) throw KfError{KF_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e)};
    c.live[p] = {d, cls};
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    Cache& c = cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) {  // a block above the cached sizes
        (void)hipFree(p);
        return;
    }
    const int d = it->second.first;
    const size_t cls = it->second.second;
    c.live.erase(it);
    DeviceCache& dc = c.dev[d];
    if (dc.held + cls > KF_CACHE_MAX_BYTES) {
        (void)hipFree(p);
        return;
    }
    dc.free[cls].push_back(p);
    dc.held += cls;
    dc.dirty = true;
}

}  // namespace kf
