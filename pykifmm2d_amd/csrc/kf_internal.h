// kf_internal.h -- internal structures of libkifmm2d (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kf_common.h"
#include "tree_host.h"

namespace kf {

void set_error(const std::string& msg);

struct KfError {
    int code;
    std::string msg;
};

#define KF_HIP_CHECK(expr)                                                               \
    do {                                                                                 \
        hipError_t kf_e_ = (expr);                                                       \
        if (kf_e_ != hipSuccess)                                                         \
            throw ::kf::KfError{KF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(kf_e_)}; \
    } while (0)

#define KF_REQUIRE(cond, code, msg)                                \
    do {                                                           \
        if (!(cond)) throw ::kf::KfError{(code), std::string(msg)}; \
    } while (0)

// Device memory of DevBuf (kf_alloc.cpp): blocks of up to 64 MiB come from and return to a
// per-device cache of freed blocks; the first reuse after a release synchronises the device.
void* dev_alloc(size_t bytes);
void dev_free(void* p);

// Owning device buffer (library-internal allocations only; caller buffers are raw pointers).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool own = true;  // false: a view of memory another object owns (view())
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && own) dev_free(p);
        p = nullptr;
        n = 0;
        own = true;
    }
    // Refers to count elements at ptr without owning them.
    void view(T* ptr, size_t count) {
        release();
        p = count ? ptr : nullptr;
        n = count;
        own = false;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        p = static_cast<T*>(dev_alloc(count * sizeof(T)));
        n = count;
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        if (!h.empty())
            KF_HIP_CHECK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s) {
        if (n) KF_HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
};

// Runs a C-ABI body, converting exceptions into a return code + thread-local message.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return KF_OK;
    } catch (const KfError& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return KF_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_error(e.what());
        return KF_ERR_ARG;
    }
}

}  // namespace kf

// A tabulated kernel (kf_ktable_create): the coefficients on the host, and on the device once
// a direct sum or kernel form has needed them (uploaded by the first such call, under `once`).
struct kf_ktable {
    std::vector<double> c;
    mutable kf::DevBuf<double> dc;
    mutable std::once_flag once;
    int nseg = 0, deg = 0, s0 = 0;  // segments [j du, (j + 1) du), j = s0 .. s0 + nseg - 1
    int ld = 0;                     // row length: deg + 1 rounded up to even, zero padded
    double du = 1.0, g0 = 0.0;
    void upload() const {
        std::call_once(once, [this] {
            dc.alloc(c.size());
            KF_HIP_CHECK(hipMemcpy(dc.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
        });
    }
    KfTable view(bool device) const {
        KfTable t;
        t.c = device ? dc.p : c.data();
        t.nseg = nseg;
        t.deg = deg;
        t.ld = ld;
        t.s0 = s0;
        t.rdu = 1.0 / du;
        t.g0 = g0;
        return t;
    }
};

namespace kf {
// The evaluation kernels' form of a kernel (KerT, kf_common.h).
template <int KIND>
inline KerT<KIND> make_ker(const KfKernel& k) {
    KerT<KIND> r{k.k, k.scale};
    if constexpr (KIND == KF_TABLE) r.t = k.td;
    return r;
}

// Device points of one tree, or of every tree of a batch build (concatenated in tree order).
struct TreePoints {
    DevBuf<double> xs, ys;      // points in tree order
    DevBuf<int32_t> perm;       // position -> index into its own tree's input (reference Tree.ordv)
    DevBuf<int32_t> gperm;      // batch builds: position -> index into the concatenated input
    int n_trees = 1;
};
}  // namespace kf

namespace kf {
struct DevForest;
}

struct kf_tree {
    // host box structure: built by kf_tree_from_sorted_keys, or copied from `dev` by host()
    mutable kf::HostTree h;
    mutable std::once_flag mirrored;
    std::shared_ptr<kf::DevForest> dev;   // device box structure of a GPU build (shared by a batch)
    std::shared_ptr<kf::TreePoints> pts;  // shared by the trees of one batch build
    int slot = 0;                         // index of the tree in its batch
    int64_t off = 0;                      // first point of the tree in pts
    const double* xs = nullptr;           // = pts->xs + off
    const double* ys = nullptr;
    const int32_t* perm = nullptr;        // = pts->perm + off
    double ymax = 0.0;
    const kf::HostTree& host() const;     // h, copied from the device structure on first use
    int levels() const;
    int64_t n_boxes() const;
    int64_t level_size(int l) const;
};
