// ELL sparse products and CSR -> ELL packing for gfx950.
//
// One row per lane: a wavefront covers 64 consecutive rows, so with the slot-major
// layout (value k of row r at k * M + r) every slot load of a wave is one contiguous
// 256-byte segment.  The B right-hand columns of a row are accumulated in registers
// (B <= 8 per launch row-tile), so each gathered x row is read once per nonzero.
// Replaces the Triton kernels of utils/ell_kernel.py:24-210 and csr_to_ell (:225-256).
#include <stdarg.h>

#include "fvm_common.hpp"

namespace fvm {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

template <int BT, typename CT>
__global__ void __launch_bounds__(256)
k_ell_spmm(const float *__restrict__ vals, int64_t vsm, int64_t vsk, const CT *__restrict__ cols, int64_t csm,
           int64_t csk, int64_t M, int K, const float *__restrict__ x, int64_t xsn, int64_t xsb, int B,
           const float *__restrict__ bias, float *__restrict__ y, int64_t ysm, int64_t ysb) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= M) return;
    const int b0 = blockIdx.y * BT;
    const int nb = min(BT, B - b0);
    float acc[BT];
#pragma unroll
    for (int j = 0; j < BT; ++j) acc[j] = 0.0f;
    for (int k = 0; k < K; ++k) {
        const float v = vals[r * vsm + k * vsk];
        const int64_t c = static_cast<int64_t>(cols[r * csm + k * csk]);
        const float *xr = x + c * xsn + static_cast<int64_t>(b0) * xsb;
#pragma unroll
        for (int j = 0; j < BT; ++j)
            if (j < nb) acc[j] += v * xr[j * xsb];
    }
    const float bb = bias ? bias[r] : 0.0f;
#pragma unroll
    for (int j = 0; j < BT; ++j)
        if (j < nb) y[r * ysm + static_cast<int64_t>(b0 + j) * ysb] = bias ? acc[j] + bb : acc[j];
}

// Fast path for the operators of the solver (K <= 8 slots, slot-major values and 32-bit
// columns, contiguous right-hand side and output): the K values and columns of the row
// are loaded first, then the K gathered rows of x, then the products in slot order, so
// a row costs two memory round trips instead of one per slot; with B = 4 the x rows and
// the output row move as 16-byte vectors.  Same products and summation order as
// k_ell_spmm.  Launch shape (the autotuner's choice): RPL rows per lane, the lanes of a
// block on consecutive rows in each of RPL row groups (so every slot load of a wave stays
// one contiguous segment), blockDim.x threads per block.
template <int B, int KT, int RPL>
__global__ void __launch_bounds__(512)
k_ell_spmm_fast(const float *__restrict__ vals, const int32_t *__restrict__ cols, int64_t vsk, int64_t M,
                const float *__restrict__ x, const float *__restrict__ bias, float *__restrict__ y) {
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * blockDim.x * RPL + threadIdx.x;
    float v[RPL][KT];
    int32_t c[RPL][KT];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        const int64_t r = r0 + static_cast<int64_t>(q) * blockDim.x;
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            v[q][k] = r < M ? vals[r + k * vsk] : 0.0f;
            c[q][k] = r < M ? cols[r + k * vsk] : 0;
        }
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        const int64_t r = r0 + static_cast<int64_t>(q) * blockDim.x;
        if (r >= M) break;
        float acc[B];
#pragma unroll
        for (int j = 0; j < B; ++j) acc[j] = 0.0f;
        if constexpr (B == 4) {
            float4 xr[KT];
#pragma unroll
            for (int k = 0; k < KT; ++k) xr[k] = reinterpret_cast<const float4 *>(x)[c[q][k]];
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                acc[0] += v[q][k] * xr[k].x;
                acc[1] += v[q][k] * xr[k].y;
                acc[2] += v[q][k] * xr[k].z;
                acc[3] += v[q][k] * xr[k].w;
            }
        } else {
            float xr[KT][B];
#pragma unroll
            for (int k = 0; k < KT; ++k)
#pragma unroll
                for (int j = 0; j < B; ++j) xr[k][j] = x[static_cast<int64_t>(c[q][k]) * B + j];
#pragma unroll
            for (int k = 0; k < KT; ++k)
#pragma unroll
                for (int j = 0; j < B; ++j) acc[j] += v[q][k] * xr[k][j];
        }
        if (bias) {
            const float bb = bias[r];
#pragma unroll
            for (int j = 0; j < B; ++j) acc[j] = acc[j] + bb;
        }
        if constexpr (B == 4) {
            reinterpret_cast<float4 *>(y)[r] = make_float4(acc[0], acc[1], acc[2], acc[3]);
        } else {
#pragma unroll
            for (int j = 0; j < B; ++j) y[r * B + j] = acc[j];
        }
    }
}

template <int B, int RPL>
static bool launch_spmm_fast_rpl(const float *vals, const int32_t *cols, int64_t vsk, int64_t M, int K,
                                 const float *x, const float *bias, float *y, int block, hipStream_t s) {
    const dim3 grid(blocks_for(M, block * RPL)), tb(block);
    switch (K) {
        case 1: hipLaunchKernelGGL((k_ell_spmm_fast<B, 1, RPL>), grid, tb, 0, s, vals, cols, vsk, M, x, bias, y); return true;
        case 2: hipLaunchKernelGGL((k_ell_spmm_fast<B, 2, RPL>), grid, tb, 0, s, vals, cols, vsk, M, x, bias, y); return true;
        case 3: hipLaunchKernelGGL((k_ell_spmm_fast<B, 3, RPL>), grid, tb, 0, s, vals, cols, vsk, M, x, bias, y); return true;
        case 4: hipLaunchKernelGGL((k_ell_spmm_fast<B, 4, RPL>), grid, tb, 0, s, vals, cols, vsk, M, x, bias, y); return true;
        case 5: hipLaunchKernelGGL((k_ell_spmm_fast<B, 5, RPL>), grid, tb, 0, s, vals, cols, vsk, M, x, bias, y); return true;
        case 6: hipLaunchKernelGGL((k_ell_spmm_fast<B, 6, RPL>), grid, tb, 0, s, vals, cols, vsk, M, x, bias, y); return true;
        case 8: hipLaunchKernelGGL((k_ell_spmm_fast<B, 8, RPL>), grid, tb, 0, s, vals, cols, vsk, M, x, bias, y); return true;
        default: return false;
    }
}

template <int B>
static bool launch_spmm_fast(const float *vals, const int32_t *cols, int64_t vsk, int64_t M, int K, const float *x,
                             const float *bias, float *y, int rpl, int block, hipStream_t s) {
    switch (rpl) {
        case 2: return launch_spmm_fast_rpl<B, 2>(vals, cols, vsk, M, K, x, bias, y, block, s);
        case 4: return launch_spmm_fast_rpl<B, 4>(vals, cols, vsk, M, K, x, bias, y, block, s);
        default: return launch_spmm_fast_rpl<B, 1>(vals, cols, vsk, M, K, x, bias, y, block, s);
    }
}

template <int BT, typename CT>
static void launch_spmm(const float *vals, int64_t vsm, int64_t vsk, const CT *cols, int64_t csm, int64_t csk,
                        int64_t M, int K, const float *x, int64_t xsn, int64_t xsb, int B, const float *bias,
                        float *y, int64_t ysm, int64_t ysb, hipStream_t s) {
    dim3 grid(blocks_for(M, 256), (B + BT - 1) / BT);
    hipLaunchKernelGGL((k_ell_spmm<BT, CT>), grid, dim3(256), 0, s, vals, vsm, vsk, cols, csm, csk, M, K, x, xsn,
                       xsb, B, bias, y, ysm, ysb);
}

template <typename CT>
static void dispatch_spmm(const float *vals, int64_t vsm, int64_t vsk, const CT *cols, int64_t csm, int64_t csk,
                          int64_t M, int K, const float *x, int64_t xsn, int64_t xsb, int B, const float *bias,
                          float *y, int64_t ysm, int64_t ysb, hipStream_t s) {
    switch (B) {
        case 1: launch_spmm<1>(vals, vsm, vsk, cols, csm, csk, M, K, x, xsn, xsb, B, bias, y, ysm, ysb, s); break;
        case 2: launch_spmm<2>(vals, vsm, vsk, cols, csm, csk, M, K, x, xsn, xsb, B, bias, y, ysm, ysb, s); break;
        case 3: launch_spmm<3>(vals, vsm, vsk, cols, csm, csk, M, K, x, xsn, xsb, B, bias, y, ysm, ysb, s); break;
        case 4: launch_spmm<4>(vals, vsm, vsk, cols, csm, csk, M, K, x, xsn, xsb, B, bias, y, ysm, ysb, s); break;
        case 5: launch_spmm<5>(vals, vsm, vsk, cols, csm, csk, M, K, x, xsn, xsb, B, bias, y, ysm, ysb, s); break;
        default: launch_spmm<8>(vals, vsm, vsk, cols, csm, csk, M, K, x, xsn, xsb, B, bias, y, ysm, ysb, s); break;
    }
}

__global__ void k_max_row_nnz(const int64_t *__restrict__ crow, int64_t M, int *out) {
    __shared__ int smax[256];
    int best = 0;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < M;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x)
        best = max(best, static_cast<int>(crow[r + 1] - crow[r]));
    smax[threadIdx.x] = best;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, smax[0]);
}

__global__ void __launch_bounds__(256)
k_ell_pack(const int64_t *__restrict__ crow, const int64_t *__restrict__ col, const float *__restrict__ val,
           int64_t M, int K, float *__restrict__ ev, int *__restrict__ ec, int64_t stride, int *overflow) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= M) return;
    const int64_t s = crow[r];
    const int64_t n = crow[r + 1] - s;
    if (n > K) atomicOr(overflow, 1);
    for (int k = 0; k < K; ++k) {
        const bool in = k < n;
        ev[k * stride + r] = in ? val[s + k] : 0.0f;
        ec[k * stride + r] = in ? static_cast<int>(col[s + k]) : 0;
    }
}

}  // namespace fvm

using namespace fvm;

extern "C" {

const char *fvm_last_error(void) { return g_err; }

int fvm_abi_version(void) { return 18; }

int fvm_ell_spmm_f32_shape(const float *vals, int64_t vsm, int64_t vsk, const void *cols, int32_t cols_is64,
                           int64_t csm, int64_t csk, int64_t M, int32_t K, const float *x, int64_t xsn, int64_t xsb,
                           int32_t B, const float *bias, float *y, int64_t ysm, int64_t ysb, int32_t rows_per_lane,
                           int32_t block_threads, void *stream) {
    if (M < 0 || K < 0 || B < 1 || (M > 0 && (!y || (K > 0 && (!vals || !cols || !x))))) {
        set_error("fvm_ell_spmm_f32: invalid arguments (M=%lld K=%d B=%d)", (long long)M, K, B);
        return FVM_ERR_ARG;
    }
    if ((rows_per_lane != 1 && rows_per_lane != 2 && rows_per_lane != 4) ||
        (block_threads != 64 && block_threads != 128 && block_threads != 256 && block_threads != 512)) {
        set_error("fvm_ell_spmm_f32_shape: rows_per_lane must be 1, 2 or 4 and block_threads 64, 128, 256 or 512");
        return FVM_ERR_ARG;
    }
    if (M == 0) return FVM_OK;
    hipStream_t s = as_stream(stream);
    // fast path: slot-major int32 ELL (value k of row r at r + k vsk, the same strides for
    // the columns), contiguous x [n][B] and y [M][B], B in {1, 4} (16-byte aligned for 4)
    const bool aligned4 = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15u) == 0;
    if (!cols_is64 && K >= 1 && vsm == 1 && csm == 1 && csk == vsk && xsb == 1 && ysb == 1 && xsn == B &&
        ysm == B && (B == 1 || (B == 4 && aligned4))) {
        const int32_t *c32 = static_cast<const int32_t *>(cols);
        const bool done =
            B == 4 ? launch_spmm_fast<4>(vals, c32, vsk, M, K, x, bias, y, rows_per_lane, block_threads, s)
                   : launch_spmm_fast<1>(vals, c32, vsk, M, K, x, bias, y, rows_per_lane, block_threads, s);
        if (done) return check_launch("fvm_ell_spmm_f32");
    }
    if (cols_is64)
        dispatch_spmm(vals, vsm, vsk, static_cast<const int64_t *>(cols), csm, csk, M, K, x, xsn, xsb, B, bias, y,
                      ysm, ysb, s);
    else
        dispatch_spmm(vals, vsm, vsk, static_cast<const int32_t *>(cols), csm, csk, M, K, x, xsn, xsb, B, bias, y,
                      ysm, ysb, s);
    return check_launch("fvm_ell_spmm_f32");
}

int fvm_ell_spmm_f32(const float *vals, int64_t vsm, int64_t vsk, const void *cols, int32_t cols_is64, int64_t csm,
                     int64_t csk, int64_t M, int32_t K, const float *x, int64_t xsn, int64_t xsb, int32_t B,
                     const float *bias, float *y, int64_t ysm, int64_t ysb, void *stream) {
    return fvm_ell_spmm_f32_shape(vals, vsm, vsk, cols, cols_is64, csm, csk, M, K, x, xsn, xsb, B, bias, y, ysm, ysb,
                                  1, 256, stream);
}

int fvm_ell_fast_path(int64_t vsm, int64_t vsk, int32_t cols_is64, int64_t csm, int64_t csk, int32_t K,
                      int64_t xsn, int64_t xsb, int32_t B, int64_t ysm, int64_t ysb, const void *x, const void *y) {
    const bool aligned4 = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15u) == 0;
    return !cols_is64 && K >= 1 && K <= 8 && K != 7 && vsm == 1 && csm == 1 && csk == vsk && xsb == 1 && ysb == 1 &&
           xsn == B && ysm == B && (B == 1 || (B == 4 && aligned4));
}

int fvm_csr_max_row_nnz(const int64_t *crow, int64_t M, int32_t *out_max, void *stream) {
    if (M < 0 || !crow || !out_max) {
        set_error("fvm_csr_max_row_nnz: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (M == 0) return FVM_OK;
    const unsigned int nb = min(blocks_for(M, 256), 1024u);
    hipLaunchKernelGGL(k_max_row_nnz, dim3(nb), dim3(256), 0, as_stream(stream), crow, M, out_max);
    return check_launch("fvm_csr_max_row_nnz");
}

int fvm_ell_pack_f32(const int64_t *crow, const int64_t *col, const float *val, int64_t M, int32_t K, float *ev,
                     int32_t *ec, int64_t stride, int32_t *overflow, void *stream) {
    if (M < 0 || K < 0 || stride < M || !crow || !overflow || (M > 0 && K > 0 && (!ev || !ec))) {
        set_error("fvm_ell_pack_f32: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (M == 0) return FVM_OK;
    hipLaunchKernelGGL(k_ell_pack, dim3(blocks_for(M, 256)), dim3(256), 0, as_stream(stream), crow, col, val, M, K,
                       ev, ec, stride, overflow);
    return check_launch("fvm_ell_pack_f32");
}

}  // extern "C"
