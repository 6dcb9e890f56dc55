// direct.hip -- FP64 direct sums and dense kernel matrices on the GPU.
//
// kf_direct is the levels <= 2 fallback of on_the_fly_fmm (reference fmm.py:229-231) and
// Kernel_Apply (fmm.py:170-188); its compensated (Neumaier) accumulation also makes it the
// device-side direct-sum reference used to state FMM errors.
#include <algorithm>
#include <mutex>
#include <string>

#include "kf_internal.h"

namespace kf {

static constexpr int DT = 256;
static constexpr int64_t DCHUNK = 16384;  // sources per partial sum

// Partial compensated sums: block (bx, by) sums sources [by*DCHUNK, (by+1)*DCHUNK) for its
// 256 targets and writes (sum, correction) to part[by][j].
template <int KIND>
__global__ void __launch_bounds__(DT) k_direct_part(KerT<KIND> ker, const double* __restrict__ sx,
                                                    const double* __restrict__ sy, const double* __restrict__ tau,
                                                    int64_t ns, const double* __restrict__ tx,
                                                    const double* __restrict__ ty, const int64_t* __restrict__ tidx,
                                                    int64_t nt, int self_mode, double* __restrict__ part) {
    __shared__ double bx[DT], by[DT], bt[DT];
    const int64_t j = blockIdx.x * (int64_t)DT + threadIdx.x;
    const bool act = j < nt;
    const double x = act ? tx[j] : 0.0, y = act ? ty[j] : 0.0;
    const int64_t self = act ? (tidx ? tidx[j] : j) : -1;
    const int64_t lo = (int64_t)blockIdx.y * DCHUNK, hi = min(ns, lo + DCHUNK);
    double s = 0.0, c = 0.0;
    for (int64_t i0 = lo; i0 < hi; i0 += DT) {
        const int m = (int)min<int64_t>(DT, hi - i0);
        __syncthreads();
        if ((int)threadIdx.x < m) {
            bx[threadIdx.x] = sx[i0 + threadIdx.x];
            by[threadIdx.x] = sy[i0 + threadIdx.x];
            bt[threadIdx.x] = tau[i0 + threadIdx.x];
        }
        __syncthreads();
        if (!act) continue;
        for (int k = 0; k < m; ++k) {
            if (self_mode == 1 && i0 + k == self) continue;
            double dx = x - bx[k], dy = y - by[k];
            if (self_mode == 2 && dx == 0.0 && dy == 0.0) continue;
            double v = ker.g(fma(dx, dx, dy * dy)) * bt[k];
            double t = s + v;
            c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
            s = t;
        }
    }
    if (act) {
        part[(size_t)blockIdx.y * 2 * nt + j] = s;
        part[(size_t)blockIdx.y * 2 * nt + nt + j] = c;
    }
}

// Deterministic compensated reduction of the partials in chunk order.
__global__ void k_direct_reduce(const double* __restrict__ part, int nchunk, int64_t nt, double* __restrict__ out) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= nt) return;
    double s = 0.0, c = 0.0;
    for (int k = 0; k < nchunk; ++k) {
        double v = part[(size_t)k * 2 * nt + j];
        double t = s + v;
        c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
        s = t;
        c += part[(size_t)k * 2 * nt + nt + j];
    }
    out[j] = s + c;
}

__global__ void k_kernel_form(KfKernel ker, const double* __restrict__ sx, const double* __restrict__ sy,
                              int64_t ns, const double* __restrict__ tx, const double* __restrict__ ty,
                              int64_t nt, double* __restrict__ out) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= ns * nt) return;
    int64_t j = e / ns, i = e - j * ns;
    out[e] = ker.eval(sx[i], sy[i], tx[j], ty[j]);
}

static KfKernel to_kernel(const kf_kernel_desc* kd) {
    KF_REQUIRE(kd, KF_ERR_ARG, "null kernel");
    KF_REQUIRE(kd->kind == KF_LAPLACE || kd->kind == KF_MODHELM || kd->kind == KF_TABLE, KF_ERR_ARG,
               "unknown kernel kind");
    KF_REQUIRE(kd->kind != KF_TABLE || kd->table, KF_ERR_ARG, "a tabulated kernel needs its table");
    KF_REQUIRE(!kd->table_im, KF_ERR_ARG,
               "direct sums and kernel forms take a complex kernel's two parts as two real kernels");
    KfKernel k{kd->kind, kd->k, kd->scale};
    if (kd->kind == KF_TABLE) {  // the caller's table, alive while the call's kernels run
        kd->table->upload();
        k.th = kd->table->view(false);
        k.td = kd->table->view(true);
    }
    return k;
}

// Argument checks shared by the direct-sum entry points; returns the number of source chunks.
static int check_direct(int64_t ns, int64_t nt, int self_mode, const double* sx, const double* sy,
                        const double* tau, const double* tx, const double* ty, const double* out) {
    KF_REQUIRE(ns >= 0 && nt >= 0, KF_ERR_ARG, "negative size");
    KF_REQUIRE(self_mode >= 0 && self_mode <= 2, KF_ERR_ARG, "self_mode must be 0, 1 or 2");
    if (nt == 0) return 0;
    KF_REQUIRE(tx && ty && out && (ns == 0 || (sx && sy && tau)), KF_ERR_ARG, "null pointer");
    const int64_t nchunk = std::max<int64_t>(1, (ns + DCHUNK - 1) / DCHUNK);
    KF_REQUIRE(nchunk <= 65535, KF_ERR_ARG, "too many sources for one direct sum");
    return (int)nchunk;
}

// The partial sums into part[nchunk][2][nt], then their reduction into out.
static void launch_direct(const KfKernel& ker, const double* sx, const double* sy, const double* tau, int64_t ns,
                          const double* tx, const double* ty, const int64_t* tidx, int64_t nt, int self_mode,
                          double* out, int nchunk, double* part, hipStream_t s) {
    dim3 grid((unsigned)((nt + DT - 1) / DT), (unsigned)nchunk);
    if (ker.kind == KF_LAPLACE)
        k_direct_part<KF_LAPLACE><<<grid, DT, 0, s>>>(make_ker<KF_LAPLACE>(ker), sx, sy, tau, ns, tx, ty, tidx, nt,
                                                       self_mode, part);
    else if (ker.kind == KF_TABLE)
        k_direct_part<KF_TABLE><<<grid, DT, 0, s>>>(make_ker<KF_TABLE>(ker), sx, sy, tau, ns, tx, ty, tidx, nt,
                                                     self_mode, part);
    else
        k_direct_part<KF_MODHELM><<<grid, DT, 0, s>>>(make_ker<KF_MODHELM>(ker), sx, sy, tau, ns, tx, ty, tidx, nt,
                                                       self_mode, part);
    k_direct_reduce<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(part, nchunk, nt, out);
    KF_HIP_CHECK(hipGetLastError());
}

// Scratch of kf_direct_noalloc: the partial sums of one call.  Every GPU the library runs on
// holds its own copy until the process ends; every element a call reads it has written
// earlier in the same call, so nothing is cleared.  In host code this name stands for the
// array, not for its device address: scratch_address() gives that address for the current
// device (hipGetSymbolAddress, once per device).
static __device__ __align__(256) double scratch[KF_DIRECT_NOALLOC_MAX_DOUBLES];

static double* scratch_address() {
    static std::mutex mu;
    static double* addr[256] = {};
    int dev = 0;
    KF_HIP_CHECK(hipGetDevice(&dev));
    KF_REQUIRE(dev >= 0 && dev < 256, KF_ERR_HIP, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        KF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        KF_REQUIRE(p, KF_ERR_HIP, "hipGetSymbolAddress returned no address for the direct-sum scratch");
        addr[dev] = (double*)p;
    }
    return addr[dev];
}

}  // namespace kf

using namespace kf;

extern "C" {

int kf_direct(const kf_kernel_desc* kernel, const double* d_sx, const double* d_sy, const double* d_tau,
              int64_t ns, const double* d_tx, const double* d_ty, const int64_t* d_tidx, int64_t nt,
              int32_t self_mode, double* d_out, void* stream) {
    return guarded([&] {
        KfKernel ker = to_kernel(kernel);
        const int nchunk = check_direct(ns, nt, self_mode, d_sx, d_sy, d_tau, d_tx, d_ty, d_out);
        if (nt == 0) return;
        hipStream_t s = (hipStream_t)stream;
        DevBuf<double> part;
        part.alloc((size_t)nchunk * 2 * nt);
        launch_direct(ker, d_sx, d_sy, d_tau, ns, d_tx, d_ty, d_tidx, nt, self_mode, d_out, nchunk, part.p, s);
        // the partials buffer is released when this call returns
        KF_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int kf_direct_noalloc(const kf_kernel_desc* kernel, const double* d_sx, const double* d_sy, const double* d_tau,
                      int64_t ns, const double* d_tx, const double* d_ty, const int64_t* d_tidx, int64_t nt,
                      int32_t self_mode, double* d_out, void* stream) {
    return guarded([&] {
        KF_REQUIRE(kernel, KF_ERR_ARG, "null kernel");
        KF_REQUIRE(kernel->kind != KF_TABLE || (kernel->table && kernel->table->dc.p), KF_ERR_STATE,
                   "kf_direct_noalloc: a tabulated kernel's device copy is made by its first kf_direct or "
                   "kf_kernel_form call, which this form does not make");
        KfKernel ker = to_kernel(kernel);
        const int nchunk = check_direct(ns, nt, self_mode, d_sx, d_sy, d_tau, d_tx, d_ty, d_out);
        if (nt == 0) return;
        KF_REQUIRE((uint64_t)nchunk * 2 * (uint64_t)nt <= (uint64_t)KF_DIRECT_NOALLOC_MAX_DOUBLES, KF_ERR_ARG,
                   "kf_direct_noalloc: 2 nt ceil(ns / 16384) = " + std::to_string((uint64_t)nchunk * 2 * nt) +
                       " partial sums exceed its scratch of " + std::to_string(KF_DIRECT_NOALLOC_MAX_DOUBLES) +
                       " doubles");
        launch_direct(ker, d_sx, d_sy, d_tau, ns, d_tx, d_ty, d_tidx, nt, self_mode, d_out, nchunk,
                      scratch_address(), (hipStream_t)stream);
    });
}

int kf_kernel_form(const kf_kernel_desc* kernel, const double* d_sx, const double* d_sy, int64_t ns,
                   const double* d_tx, const double* d_ty, int64_t nt, double* d_out, void* stream) {
    return guarded([&] {
        KfKernel ker = to_kernel(kernel);
        KF_REQUIRE(ns >= 0 && nt >= 0, KF_ERR_ARG, "negative size");
        if (ns == 0 || nt == 0) return;
        KF_REQUIRE(d_sx && d_sy && d_tx && d_ty && d_out, KF_ERR_ARG, "null pointer");
        int64_t tot = ns * nt;
        k_kernel_form<<<(unsigned)((tot + 255) / 256), 256, 0, (hipStream_t)stream>>>(ker, d_sx, d_sy, ns, d_tx,
                                                                                      d_ty, nt, d_out);
        KF_HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"
