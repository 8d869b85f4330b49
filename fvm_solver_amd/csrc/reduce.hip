// Global fp64 reductions of the step loop for gfx950: conservation totals and the CFL
// time step.  Both are two-pass and deterministic: a fixed grid of 256 blocks writes
// per-block partials, and one block combines them in block order.
#include "fvm_common.hpp"

#include <mutex>

namespace fvm {

constexpr int kRedBlocks = 256;
constexpr int kRedThreads = 256;

// Scratch of the allocation-free conservation entry point: the per-block partials
// [kRedBlocks][C] for C <= kScratchMaxC (16 KiB).  Every GPU the library runs on holds
// its own copy for the life of the process.  Host code never passes the array's name to
// a kernel (in host code it is a stand-in, not a device address): it passes the address
// hipGetSymbolAddress returned for the current device.
constexpr int kScratchMaxC = 8;
static __device__ __align__(256) double scratch[kRedBlocks * kScratchMaxC];

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The current device's scratch address, looked up once per device (with that device
// current) on the first call that needs it; nullptr (error set) on failure.
static double *device_scratch(const char *who) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("%s: no current HIP device (or device id beyond %d)", who, kMaxDevices - 1);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("%s: hipGetSymbolAddress(scratch) failed on device %d: %s", who, dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<double *>(ptr);
    }
    return g_scratch_ptr[dev];
}

__global__ void __launch_bounds__(kRedThreads)
k_cons_partial(const float *__restrict__ U, const double *__restrict__ vol, int64_t N, int C, double *__restrict__ part) {
    __shared__ double s[kRedThreads];
    for (int j = 0; j < C; ++j) {
        double acc = 0.0;
        for (int64_t c = static_cast<int64_t>(blockIdx.x) * kRedThreads + threadIdx.x; c < N;
             c += static_cast<int64_t>(gridDim.x) * kRedThreads)
            acc += vol[c] * static_cast<double>(U[c * C + j]);
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int w = kRedThreads / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x * C + j] = s[0];
        __syncthreads();
    }
}

__global__ void k_cons_final(const double *__restrict__ part, int nblk, int C, double *__restrict__ out) {
    const int j = threadIdx.x;
    if (j >= C) return;
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += part[b * C + j];
    out[j] = acc;
}

__global__ void __launch_bounds__(kRedThreads)
k_cfl_partial(const float *__restrict__ prim, const double *__restrict__ h, int64_t N, int dim, double gR,
              double *__restrict__ part) {
    __shared__ double s[kRedThreads];
    const int C = dim + 2;
    double best = 1.0e300;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * kRedThreads + threadIdx.x; c < N;
         c += static_cast<int64_t>(gridDim.x) * kRedThreads) {
        const float *P = prim + c * C;
        double v2 = 0.0;
        for (int i = 0; i < dim; ++i) v2 += static_cast<double>(P[i]) * P[i];
        const double cs = sqrt(gR * static_cast<double>(P[dim + 1]));
        best = fmin(best, h[c] / (sqrt(v2) + cs));
    }
    s[threadIdx.x] = best;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_cfl_final(const double *__restrict__ part, int nblk, double cfl, double *__restrict__ out) {
    if (threadIdx.x != 0) return;
    double m = 1.0e300;
    for (int b = 0; b < nblk; ++b) m = fmin(m, part[b]);
    *out = cfl * m;
}

}  // namespace fvm

using namespace fvm;

extern "C" {

int fvm_conservation_f64(const float *state, const double *vol, int64_t N, int32_t C, double *out, double *partial,
                         void *stream) {
    if (!state || !vol || !out || !partial || N < 1 || C < 1 || C > 64) {
        set_error("fvm_conservation_f64: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_cons_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, s, state, vol, N, C, partial);
    hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, s, partial, kRedBlocks, C, out);
    return check_launch("fvm_conservation_f64");
}

int fvm_conservation_f64_noalloc(const float *state, const double *vol, int64_t N, int32_t C, double *out,
                                 void *stream) {
    if (!state || !vol || !out || N < 1 || C < 1) {
        set_error("fvm_conservation_f64_noalloc: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (C > kScratchMaxC) {
        set_error("fvm_conservation_f64_noalloc: C = %d exceeds the scratch limit of %d variables", C, kScratchMaxC);
        return FVM_ERR_ARG;
    }
    double *part = device_scratch("fvm_conservation_f64_noalloc");
    if (!part) return FVM_ERR_HIP;
    // every block overwrites all C of its partials, so nothing needs clearing
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_cons_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, s, state, vol, N, C, part);
    hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, s, part, kRedBlocks, C, out);
    return check_launch("fvm_conservation_f64_noalloc");
}

int fvm_cfl_dt(const float *prim, const double *h, int64_t N, int32_t dim, double gR, double cfl, double *out,
               double *partial, void *stream) {
    if (!prim || !h || !out || !partial || N < 1 || (dim != 2 && dim != 3)) {
        set_error("fvm_cfl_dt: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_cfl_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, s, prim, h, N, dim, gR, partial);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(64), 0, s, partial, kRedBlocks, cfl, out);
    return check_launch("fvm_cfl_dt");
}

}  // extern "C"
