// Message-passing corrector on the mesh graph: BF16 MFMA GEMMs with FP32 accumulation,
// neighbour aggregation forward/backward, activations, loss and Adam, for gfx950.
//
// The network (the corrector of the learned path) is
//   h0 = relu(x W_in + b_in),  h_{l+1} = h_l + relu([h_l | mean_nb h_l] W_l + b_l)  (l < 3),
//   y  = h3 W_out + b_out,
// with hidden width H = 128 and the mean over a cell's interior face neighbours.  The
// mesh adjacency is symmetric, so the aggregation's backward is a gather as well
// (deterministic, no atomics).
#include "fvm_common.hpp"

namespace fvm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32;

// C(i, j) = sum_k A(i, k) B(k, j) over k in this split's range, with A(i, k) at
// A[i * a_rs + k * a_cs] and B(k, j) at B[k * b_rs + j * b_cs] (any bf16 layout: tiles
// are staged through LDS).  One 256-thread workgroup computes a 64 x 64 tile; wave w
// owns rows 16w..16w+15 and issues four 16x16x32 MFMAs per 32-wide K step.
// Split-K (gridDim.z > 1) writes fp32 partials part[z][M][N]; otherwise the epilogue
// adds the bias and writes the pre-activation z (fp32, Cf, row stride c_rs) and/or
// res + relu(z) (bf16, Cb; relu and res optional).
template <bool AV, bool BV>
__global__ void __launch_bounds__(256)
k_gemm_bf16(const __bf16 *__restrict__ A, int64_t a_rs, int64_t a_cs, const __bf16 *__restrict__ B, int64_t b_rs,
            int64_t b_cs, int64_t M, int64_t N, int64_t K, int64_t k_chunk, const float *__restrict__ bias,
            float *__restrict__ Cf, int64_t c_rs, __bf16 *__restrict__ Cb, int relu, const __bf16 *__restrict__ res,
            float *__restrict__ part) {
    // rows padded to 40 elements (80 bytes): 16-byte fragment reads stay aligned and
    // consecutive rows start in different banks
    __shared__ __attribute__((aligned(16))) __bf16 As[BM][BK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[BN][BK + 8];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t m0 = static_cast<int64_t>(blockIdx.x) * BM, n0 = static_cast<int64_t>(blockIdx.y) * BN;
    const int64_t kb = static_cast<int64_t>(blockIdx.z) * k_chunk;
    const int64_t ke = kb + k_chunk < K ? kb + k_chunk : K;
    f32x4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    const int lr = tid >> 2, lk = (tid & 3) * 8;
    for (int64_t k0 = kb; k0 < ke; k0 += BK) {
        {
            // a K-contiguous operand stages 8 elements with one 16-byte load
            const int64_t gi = m0 + lr;
            if (AV && gi < M && k0 + lk + 8 <= ke) {
                *reinterpret_cast<bf16x8_t *>(&As[lr][lk]) =
                    *reinterpret_cast<const bf16x8_t *>(A + gi * a_rs + k0 + lk);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int64_t gk = k0 + lk + e;
                    As[lr][lk + e] = (gi < M && gk < ke) ? A[gi * a_rs + gk * a_cs] : __bf16(0.0f);
                }
            }
            const int64_t gj = n0 + lr;
            if (BV && gj < N && k0 + lk + 8 <= ke) {
                *reinterpret_cast<bf16x8_t *>(&Bs[lr][lk]) =
                    *reinterpret_cast<const bf16x8_t *>(B + gj * b_cs + k0 + lk);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int64_t gk = k0 + lk + e;
                    Bs[lr][lk + e] = (gj < N && gk < ke) ? B[gk * b_rs + gj * b_cs] : __bf16(0.0f);
                }
            }
        }
        __syncthreads();
        const bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(&As[16 * wave + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t *>(&Bs[16 * cb + (lane & 15)][8 * (lane >> 4)]);
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[cb], 0, 0, 0);
        }
        __syncthreads();
    }
    const bool split = gridDim.z > 1;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int64_t j = n0 + 16 * cb + (lane & 15);
        if (j >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t i = m0 + 16 * wave + 4 * (lane >> 4) + r;
            if (i >= M) continue;
            float v = acc[cb][r];
            if (split) {
                part[(static_cast<int64_t>(blockIdx.z) * M + i) * N + j] = v;
                continue;
            }
            if (bias) v += bias[j];
            if (Cf) Cf[i * c_rs + j] = v;
            if (relu) v = fmaxf(v, 0.0f);
            if (res) v += static_cast<float>(res[i * N + j]);
            if (Cb) Cb[i * N + j] = __bf16(v);
        }
    }
}

__global__ void __launch_bounds__(256)
k_splitk_reduce(const float *__restrict__ part, int S, int64_t M, int64_t N, const float *__restrict__ bias,
                float *__restrict__ Cf, int64_t c_rs, __bf16 *__restrict__ Cb, int relu,
                const __bf16 *__restrict__ res) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= M * N) return;
    const int64_t i = t / N, j = t % N;
    float v = 0.0f;
    for (int z = 0; z < S; ++z) v += part[z * M * N + t];
    if (bias) v += bias[j];
    if (Cf) Cf[i * c_rs + j] = v;
    if (relu) v = fmaxf(v, 0.0f);
    if (res) v += static_cast<float>(res[t]);
    if (Cb) Cb[t] = __bf16(v);
}

// A[c] = [h[c] | mean of h over the interior face neighbours of c]   (bf16, [N][2H])
__global__ void __launch_bounds__(256)
k_mp_aggregate(const __bf16 *__restrict__ h, int H, const int32_t *__restrict__ nbr, int64_t N, int NL,
               __bf16 *__restrict__ A) {
    const int chunks = H / 8;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= N * chunks) return;
    const int64_t c = t / chunks;
    const int q = static_cast<int>(t % chunks);
    const bf16x8_t own = *reinterpret_cast<const bf16x8_t *>(h + c * H + 8 * q);
    *reinterpret_cast<bf16x8_t *>(A + c * 2 * H + 8 * q) = own;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int cnt = 0;
    for (int k = 0; k < NL; ++k) {
        const int64_t nb = nbr[k * N + c];
        if (nb < N) {
            const bf16x8_t v = *reinterpret_cast<const bf16x8_t *>(h + nb * H + 8 * q);
#pragma unroll
            for (int e = 0; e < 8; ++e) s[e] += static_cast<float>(v[e]);
            ++cnt;
        }
    }
    const float inv = cnt ? 1.0f / static_cast<float>(cnt) : 0.0f;
    bf16x8_t m;
#pragma unroll
    for (int e = 0; e < 8; ++e) m[e] = __bf16(s[e] * inv);
    *reinterpret_cast<bf16x8_t *>(A + c * 2 * H + H + 8 * q) = m;
}

// dh[c] = dA[c][:H] + sum_{nb interior neighbour of c} dA[nb][H:] / n_nb  (+ dres[c])
__global__ void __launch_bounds__(256)
k_mp_aggregate_bwd(const float *__restrict__ dA, int H, const int32_t *__restrict__ nbr, int64_t N, int NL,
                   const float *__restrict__ dres, float *__restrict__ dh) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= N * H) return;
    const int64_t c = t / H;
    const int e = static_cast<int>(t % H);
    float v = dA[c * 2 * H + e];
    for (int k = 0; k < NL; ++k) {
        const int64_t nb = nbr[k * N + c];
        if (nb < N) {
            int cnt = 0;
            for (int kk = 0; kk < NL; ++kk) cnt += nbr[kk * N + nb] < N;
            v += dA[nb * 2 * H + H + e] / static_cast<float>(cnt);
        }
    }
    if (dres) v += dres[t];
    dh[t] = v;
}

// h_out = h_in + relu(z) (bf16 out); with h_in == NULL: h_out = relu(z)
__global__ void __launch_bounds__(256)
k_relu_residual(const float *__restrict__ z, const __bf16 *__restrict__ h_in, __bf16 *__restrict__ h_out, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float r = fmaxf(z[t], 0.0f);
    h_out[t] = __bf16(h_in ? static_cast<float>(h_in[t]) + r : r);
}

// dz = dh * (z > 0), written as bf16 (GEMM operand) and fp32
__global__ void __launch_bounds__(256)
k_relu_bwd(const float *__restrict__ dh, const float *__restrict__ z, __bf16 *__restrict__ dz_b,
           float *__restrict__ dz_f, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float g = z[t] > 0.0f ? dh[t] : 0.0f;
    dz_b[t] = __bf16(g);
    if (dz_f) dz_f[t] = g;
}

// Column sums of a [M][N] fp32 matrix: per-block partials [gridDim.x][N], summed in block
// order by k_colsum_final.
__global__ void __launch_bounds__(256)
k_colsum_partial(const float *__restrict__ X, int64_t M, int N, float *__restrict__ part) {
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        float s = 0.0f;
        for (int64_t i = blockIdx.x; i < M; i += gridDim.x) s += X[i * N + j];
        part[static_cast<int64_t>(blockIdx.x) * N + j] = s;
    }
}

__global__ void k_colsum_final(const float *__restrict__ part, int nblk, int N, float *__restrict__ out) {
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        float s = 0.0f;
        for (int b = 0; b < nblk; ++b) s += part[static_cast<int64_t>(b) * N + j];
        out[j] = s;
    }
}

// Masked MSE over the first n_out of `ld` columns: loss partials (fp64) and
// dy = scale * 2 (y - t), scale = 1 / (rows * n_out).
__global__ void __launch_bounds__(256)
k_mse(const float *__restrict__ y, int64_t ld_y, const float *__restrict__ tgt, int64_t M, int n_out,
      float *__restrict__ dy, double *__restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    const float scale = 1.0f / static_cast<float>(M * n_out);
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < M * ld_y;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t i = t / ld_y;
        const int j = static_cast<int>(t % ld_y);
        float g = 0.0f;
        if (j < n_out) {
            const float d = y[t] - tgt[i * n_out + j];
            s += static_cast<double>(d) * d;
            g = 2.0f * d * scale;
        }
        dy[t] = g;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam with bias correction on fp32 master weights; refreshes the bf16 copy.
__global__ void __launch_bounds__(256)
k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
       __bf16 *__restrict__ wb, int64_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    const float wt = w[t] - lr * (mt / bc1) / (sqrtf(vt / bc2) + eps);
    w[t] = wt;
    if (wb) wb[t] = __bf16(wt);
}

__global__ void __launch_bounds__(256) k_cast_f32_bf16(const float *__restrict__ x, __bf16 *__restrict__ y, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) y[t] = __bf16(x[t]);
}

// Network input of one window of a 2-D batch (C = 4, KX = 32), per cell c:
//   [0, 4)   (prim - mu_p) / sd_p
//   [4, 8)   W f0 / s1                 W = window length of the cell's mesh
//   [8, 12)  W (f0 - fm) / s2          fm = f half a window earlier
//   [12, 14) geometry features
//   [14, 22) h d/dx prim / sd_p, then h d/dy prim / sd_p
//   [22, 30) h W d/dx f0 / s1, then h W d/dy f0 / s1
//   [30, 32) 0
// s1, s2: [n_meshes][C] scales of the cell's mesh.  The gradients are least-squares
// weighted sums over the interior face neighbours, sum_k wg[k][i][c] (v_nb(k) - v_c)
// (a boundary face contributes nothing); wg is [NL][2][N].
template <int NL>
__global__ void __launch_bounds__(256)
k_mp_window_features(const float4 *__restrict__ prim, const float4 *__restrict__ f0, const float4 *__restrict__ fm,
                     const float *__restrict__ win, const float *__restrict__ geo, const float *__restrict__ hc,
                     const int32_t *__restrict__ nbr, const float *__restrict__ wg, int64_t N,
                     const float *__restrict__ mu_p, const float *__restrict__ sd_p, const float *__restrict__ s1,
                     const float *__restrict__ s2, const int32_t *__restrict__ cell_mesh, __bf16 *__restrict__ X) {
    constexpr int C = 4, KX = 32;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int32_t nb[NL];
    float wx[NL], wy[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        nb[k] = nbr[k * N + c];
        wx[k] = wg[(2 * k) * N + c];
        wy[k] = wg[(2 * k + 1) * N + c];
    }
    const float4 pc = prim[c], fc = f0[c], gc = fm[c];
    float4 pn[NL], fn[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const bool in = nb[k] < N;           // boundary faces: the cell's own values (zero difference)
        pn[k] = in ? prim[nb[k]] : pc;
        fn[k] = in ? f0[nb[k]] : fc;
    }
    const float p[C] = {pc.x, pc.y, pc.z, pc.w}, f[C] = {fc.x, fc.y, fc.z, fc.w}, g[C] = {gc.x, gc.y, gc.z, gc.w};
    float dp[2][C], df[2][C];
#pragma unroll
    for (int j = 0; j < C; ++j) dp[0][j] = dp[1][j] = df[0][j] = df[1][j] = 0.0f;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const float q[C] = {pn[k].x, pn[k].y, pn[k].z, pn[k].w}, r[C] = {fn[k].x, fn[k].y, fn[k].z, fn[k].w};
#pragma unroll
        for (int j = 0; j < C; ++j) {
            dp[0][j] += wx[k] * (q[j] - p[j]);
            dp[1][j] += wy[k] * (q[j] - p[j]);
            df[0][j] += wx[k] * (r[j] - f[j]);
            df[1][j] += wy[k] * (r[j] - f[j]);
        }
    }
    const int m = cell_mesh ? cell_mesh[c] : 0;
    const float w = win[c], h = hc[c];
    float x[KX];
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const float isp = 1.0f / sd_p[j], i1 = w / s1[m * C + j], i2 = w / s2[m * C + j];
        x[j] = (p[j] - mu_p[j]) * isp;
        x[C + j] = f[j] * i1;
        x[2 * C + j] = (f[j] - g[j]) * i2;
        x[14 + j] = h * dp[0][j] * isp;
        x[14 + C + j] = h * dp[1][j] * isp;
        x[22 + j] = h * df[0][j] * i1;
        x[22 + C + j] = h * df[1][j] * i1;
    }
    x[12] = geo[2 * c];
    x[13] = geo[2 * c + 1];
    x[30] = 0.0f;
    x[31] = 0.0f;
    // the row leaves as 16-byte vectors
#pragma unroll
    for (int q = 0; q < KX; q += 8) {
        bf16x8_t v;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = __bf16(x[q + e]);
        *reinterpret_cast<bf16x8_t *>(X + c * KX + q) = v;
    }
}

// U_out = U + win (a f0 + b fm) + scale[m(c)] y   (fm and y optional)
__global__ void __launch_bounds__(256)
k_mp_window_apply(const float *__restrict__ U, const float *__restrict__ f0, const float *__restrict__ fm, float a,
                  float b, const float *__restrict__ win, const float *__restrict__ y, int64_t ld_y,
                  const float *__restrict__ scale, const int32_t *__restrict__ cell_mesh, int64_t N, int C,
                  float *__restrict__ U_out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= N * C) return;
    const int64_t c = t / C;
    const int j = static_cast<int>(t % C);
    float inc = a * f0[t];
    if (fm) inc += b * fm[t];
    float v = U[t] + win[c] * inc;
    if (y) v += scale[(cell_mesh ? cell_mesh[c] : 0) * C + j] * y[c * ld_y + j];
    U_out[t] = v;
}

// Both predicted states of a window in one pass:
//   Um = U + win (am f0 + bm fm) + scale y[:, 0:C],  U1 = U + win (ae f0 + be fm) + scale y[:, C:2C]
// (y fp32 [N][ld_y], optional).
__global__ void __launch_bounds__(256)
k_mp_window_predict(const float *__restrict__ U, const float *__restrict__ f0, const float *__restrict__ fm,
                    float am, float bm, float ae, float be, const float *__restrict__ win, const float *__restrict__ y,
                    int64_t ld_y, const float *__restrict__ scale, const int32_t *__restrict__ cell_mesh, int64_t N,
                    int C, float *__restrict__ Um, float *__restrict__ U1) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= N * C) return;
    const int64_t c = t / C;
    const int j = static_cast<int>(t % C);
    const float u = U[t], a = f0[t], b = fm[t], w = win[c];
    float vm = u + w * (am * a + bm * b), v1 = u + w * (ae * a + be * b);
    if (y) {
        const float sc = scale[(cell_mesh ? cell_mesh[c] : 0) * C + j];
        vm += sc * y[c * ld_y + j];
        v1 += sc * y[c * ld_y + C + j];
    }
    Um[t] = vm;
    U1[t] = v1;
}

// The per-mesh reductions below take the variable count as a template argument CT (4 or
// 5; 0 = runtime C, up to 8) so that a thread's fp64 partial sums are registers: indexed by
// a runtime count they live in scratch memory, a load and a store per update.
// grid (64, n_meshes): per mesh q, fp64 partial sums of (w a_j)^2 and (w (a_j - b_j))^2
template <int CT>
__global__ void __launch_bounds__(256)
k_mesh_sumsq2_partial(const float *__restrict__ A, const float *__restrict__ B, const float *__restrict__ win,
                      const int64_t *__restrict__ off, int C_rt, double *__restrict__ part) {
    __shared__ double red[256];
    constexpr int CM = CT ? CT : 8;
    const int C = CT ? CT : C_rt;
    const int q = blockIdx.y;
    const int64_t a = off[q], b = off[q + 1];
    double s[2 * CM];
#pragma unroll
    for (int j = 0; j < 2 * CM; ++j) s[j] = 0.0;
    for (int64_t c = a + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < b;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float w = win[c];
#pragma unroll
        for (int j = 0; j < CM; ++j) {
            if (j >= C) continue;
            const float x = A[c * C + j];
            const float v = w * x, d = w * (x - B[c * C + j]);
            s[j] += static_cast<double>(v) * v;
            s[C + j] += static_cast<double>(d) * d;
        }
    }
#pragma unroll
    for (int j = 0; j < 2 * CM; ++j) {
        if (j >= 2 * C) continue;
        red[threadIdx.x] = s[j];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(static_cast<int64_t>(q) * gridDim.x + blockIdx.x) * 2 * C + j] = red[0];
        __syncthreads();
    }
}

// out1[q][j] = max(floor, sqrt(sum_b part[q][b][j] / n_q)), out2 from columns C + j
__global__ void k_mesh_rms2_final(const double *__restrict__ part, int nblk, int n_meshes, int C,
                                  const int64_t *__restrict__ off, float floor_abs, float *__restrict__ out1,
                                  float *__restrict__ out2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_meshes * 2 * C) return;
    const int q = t / (2 * C), j = t % (2 * C);
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(static_cast<int64_t>(q) * nblk + b) * 2 * C + j];
    const double n = static_cast<double>(off[q + 1] - off[q]);
    const float r = static_cast<float>(sqrt(s / (n > 0 ? n : 1.0)));
    (j < C ? out1 : out2)[q * C + (j % C)] = r > floor_abs ? r : floor_abs;
}

// grid (64, n_meshes): per mesh q, fp64 partial sums of (w x_j)^2 over its cells
template <int CT>
__global__ void __launch_bounds__(256)
k_mesh_sumsq_partial(const float *__restrict__ X, const float *__restrict__ win, const int64_t *__restrict__ off,
                     int C_rt, double *__restrict__ part) {
    __shared__ double red[256];
    constexpr int CM = CT ? CT : 8;
    const int C = CT ? CT : C_rt;
    const int q = blockIdx.y;
    const int64_t a = off[q], b = off[q + 1];
    double s[CM];
#pragma unroll
    for (int j = 0; j < CM; ++j) s[j] = 0.0;
    for (int64_t c = a + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < b;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float w = win ? win[c] : 1.0f;
#pragma unroll
        for (int j = 0; j < CM; ++j) {
            if (j >= C) continue;
            const float v = w * X[c * C + j];
            s[j] += static_cast<double>(v) * v;
        }
    }
#pragma unroll
    for (int j = 0; j < CM; ++j) {
        if (j >= C) continue;
        red[threadIdx.x] = s[j];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(static_cast<int64_t>(q) * gridDim.x + blockIdx.x) * C + j] = red[0];
        __syncthreads();
    }
}

__global__ void k_mesh_rms_final(const double *__restrict__ part, int nblk, int n_meshes, int C,
                                 const int64_t *__restrict__ off, float floor_abs, float *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_meshes * C) return;
    const int q = t / C, j = t % C;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(static_cast<int64_t>(q) * nblk + b) * C + j];
    const double n = static_cast<double>(off[q + 1] - off[q]);
    const float r = static_cast<float>(sqrt(s / (n > 0 ? n : 1.0)));
    out[t] = r > floor_abs ? r : floor_abs;
}

// grid (64, n_meshes): block b of mesh q sums cells off[q] + b, + 64 * 256 ... strided
template <int CT>
__global__ void __launch_bounds__(256)
k_mp_defect_partial(const float *__restrict__ U0, const float *__restrict__ U1, const float *__restrict__ f0,
                    const float *__restrict__ f1, const float *__restrict__ win, const int64_t *__restrict__ off,
                    int C_rt, double *__restrict__ part) {
    __shared__ double red[256];
    constexpr int CM = CT ? CT : 8;
    const int C = CT ? CT : C_rt;
    const int q = blockIdx.y;
    const int64_t a = off[q], b = off[q + 1];
    double s[2 * CM];
#pragma unroll
    for (int j = 0; j < 2 * CM; ++j) s[j] = 0.0;
    for (int64_t c = a + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < b;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float hw = 0.5f * win[c];
#pragma unroll
        for (int j = 0; j < CM; ++j) {
            if (j >= C) continue;
            const int64_t o = c * C + j;
            const float d = (U1[o] - U0[o]) - hw * (f0[o] + f1[o]);
            s[j] += static_cast<double>(d) * d;
            s[C + j] += static_cast<double>(U1[o]) * U1[o];
        }
    }
#pragma unroll
    for (int j = 0; j < 2 * CM; ++j) {
        if (j >= 2 * C) continue;
        red[threadIdx.x] = s[j];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(static_cast<int64_t>(q) * gridDim.x + blockIdx.x) * 2 * C + j] = red[0];
        __syncthreads();
    }
}

// per mesh, the ncol columns of the block partials summed in block order
__global__ void k_mp_defect_final(const double *__restrict__ part, int nblk, int n_meshes, int ncol,
                                  double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_meshes * ncol) return;
    const int q = t / ncol, j = t % ncol;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(static_cast<int64_t>(q) * nblk + b) * ncol + j];
    out[t] = s;
}

// One message-passing layer, aggregation fused into the GEMM's operand loads:
//   h_out = h_in + relu([h_in | mean_nb h_in] W + b)       (H = 128, K = 2H = 256)
// Persistent 512-thread workgroups stage W^T (bf16 [128][256], 66 KB) in LDS once; each
// wave then takes 16-cell tiles.  The MFMA computes features x cells (W^T is the A
// operand, read from LDS), so a lane's B operand is 8 consecutive k of one cell's row:
// it is loaded straight from h into registers -- the cell's own row for k < 128, the
// fp32 mean of its neighbours' chunks for k >= 128 -- with every load of the tile
// issued before the first MFMA, and each lane ends with 4 consecutive output features
// of one cell per column block (8-byte epilogue loads and stores).  8 K-steps x 8
// feature blocks of v_mfma_f32_16x16x32_bf16, in the generic GEMM's K-step order.
// Z (pre-activation, fp32) and the operand A (bf16 [N][2H]) are written when non-NULL
// (training).
constexpr int kLayerThreads = 512;

template <int NLT>
__global__ void __launch_bounds__(kLayerThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_mp_layer(const __bf16 *__restrict__ h, const int32_t *__restrict__ nbr, int64_t N, const __bf16 *__restrict__ WT,
           const float *__restrict__ bias, __bf16 *__restrict__ h_out, float *__restrict__ Z,
           __bf16 *__restrict__ A_out) {
    constexpr int H = 128, K = 2 * H, KP = K + 8;
    __shared__ __attribute__((aligned(16))) __bf16 Ws[H][KP];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int t = tid; t < H * (K / 8); t += kLayerThreads) {
        const int j = t / (K / 8), kc = (t % (K / 8)) * 8;
        *reinterpret_cast<bf16x8_t *>(&Ws[j][kc]) = *reinterpret_cast<const bf16x8_t *>(WT + j * K + kc);
    }
    __syncthreads();
    constexpr int nwaves = kLayerThreads / 64;
    const int64_t ntiles = (N + 15) / 16;
    const int kq = 8 * (lane >> 4);
    const int col = lane & 15;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * nwaves + wave; t < ntiles;
         t += static_cast<int64_t>(gridDim.x) * nwaves) {
        const int64_t r0 = t * 16;
        const int64_t row = r0 + col;
        const bool rlive = row < N;
        const int64_t rs = rlive ? row : 0;
        // the B fragments do not depend on the tile; an opaque per-tile offset keeps the
        // compiler from hoisting all 64 of them out of the loop (256 registers)
        int wofs = 0;
        asm volatile("" : "+v"(wofs));
        const __bf16 *wsb = &Ws[col][kq] + wofs;
        int64_t nb[NLT];
        bool val[NLT];
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < NLT; ++k) {
            const int64_t v = rlive ? static_cast<int64_t>(nbr[k * N + row]) : N;
            val[k] = v < N;
            cnt += val[k];
            nb[k] = val[k] ? v : rs;
        }
        bf16x8_t own[4], nv[NLT][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            own[i] = *reinterpret_cast<const bf16x8_t *>(h + rs * H + 32 * i + kq);
#pragma unroll
            for (int k = 0; k < NLT; ++k) nv[k][i] = *reinterpret_cast<const bf16x8_t *>(h + nb[k] * H + 32 * i + kq);
        }
        const float inv = cnt ? 1.0f / static_cast<float>(cnt) : 0.0f;
        f32x4_t acc[8];
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) acc[cb] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!rlive) {
#pragma unroll
                for (int e = 0; e < 8; ++e) own[i][e] = __bf16(0.0f);
            }
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const bf16x8_t wf = *reinterpret_cast<const bf16x8_t *>(wsb + 16 * cb * KP + 32 * i);
                acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, own[i], acc[cb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float sacc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < NLT; ++k) {
                if (val[k]) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) sacc[e] += static_cast<float>(nv[k][i][e]);
                }
            }
            bf16x8_t mv;
#pragma unroll
            for (int e = 0; e < 8; ++e) mv[e] = __bf16(sacc[e] * inv);
            if (A_out && rlive) {
                *reinterpret_cast<bf16x8_t *>(A_out + row * K + 32 * i + kq) = own[i];
                *reinterpret_cast<bf16x8_t *>(A_out + row * K + H + 32 * i + kq) = mv;
            }
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const bf16x8_t wf = *reinterpret_cast<const bf16x8_t *>(wsb + 16 * cb * KP + H + 32 * i);
                acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, mv, acc[cb], 0, 0, 0);
            }
        }
        // W^T is the MFMA's A operand and the cell rows its B operand, so the 16 x 16
        // output block is features x cells: this lane holds features 16cb + 4g .. +3 of
        // cell `row` -- 4 consecutive values that load and store as one 8-byte access
        if (rlive) {
            const int g4 = 4 * (lane >> 4);
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const int j0 = 16 * cb + g4;
                const float4 bb = *reinterpret_cast<const float4 *>(bias + j0);
                const float z0 = acc[cb][0] + bb.x, z1 = acc[cb][1] + bb.y, z2 = acc[cb][2] + bb.z,
                            z3 = acc[cb][3] + bb.w;
                if (Z) *reinterpret_cast<float4 *>(Z + row * H + j0) = make_float4(z0, z1, z2, z3);
                typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
                const bf16x4_t hr = *reinterpret_cast<const bf16x4_t *>(h + row * H + j0);
                bf16x4_t o;
                o[0] = __bf16(static_cast<float>(hr[0]) + fmaxf(z0, 0.0f));
                o[1] = __bf16(static_cast<float>(hr[1]) + fmaxf(z1, 0.0f));
                o[2] = __bf16(static_cast<float>(hr[2]) + fmaxf(z2, 0.0f));
                o[3] = __bf16(static_cast<float>(hr[3]) + fmaxf(z3, 0.0f));
                *reinterpret_cast<bf16x4_t *>(h_out + row * H + j0) = o;
            }
        }
    }
}

// Inference encoder: h0 = relu(X W_in + b_in) for KX = 32 inputs and H = 128 features.
// One wave per 16-cell tile (grid-stride over tiles, four tiles per pass): the X rows are the B operand (lane
// g|col holds inputs 8g..8g+7 of cell col), W_in^T the A operand, kept in registers with
// the lane's bias values for the whole kernel; eight v_mfma_f32_16x16x32_bf16 give
// features x cells, stored as 16-byte bf16 chunks (the rows of W_in^T permuted so that a
// lane's outputs of two blocks are 8 consecutive features).
__global__ void __launch_bounds__(256)
k_mp_encode(const __bf16 *__restrict__ X, int64_t N, const __bf16 *__restrict__ WT, const float *__restrict__ bias,
            __bf16 *__restrict__ h0) {
    constexpr int H = 128, KX = 32;
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    // row m = 4g' + r' of MFMA block cb computes feature 32(cb/2) + 8g' + 4(cb%2) + r', so
    // the lane holding rows 4g .. 4g+3 of blocks 2q and 2q+1 holds features 32q + 8g .. +7
    // of its cell and stores them as one 16-byte chunk
    auto feat = [](int cb, int m) { return 32 * (cb >> 1) + 8 * (m >> 2) + 4 * (cb & 1) + (m & 3); };
    bf16x8_t wf[8];
    float bb[8][4];
#pragma unroll
    for (int cb = 0; cb < 8; ++cb) {
        wf[cb] = *reinterpret_cast<const bf16x8_t *>(WT + feat(cb, col) * KX + 8 * g);
#pragma unroll
        for (int r = 0; r < 4; ++r) bb[cb][r] = bias[feat(cb, 4 * g + r)];
    }
    const int64_t ntiles = (N + 15) / 16;
    const int64_t nw = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
    // four tiles per pass: their X rows are loaded together before the MFMAs
    constexpr int U = 4;
    for (int64_t t0 = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); t0 < ntiles;
         t0 += U * nw) {
        bf16x8_t xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = (t0 + u * nw) * 16 + col;
            if (t0 + u * nw < ntiles && row < N) {
                xv[u] = *reinterpret_cast<const bf16x8_t *>(X + row * KX + 8 * g);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) xv[u][e] = __bf16(0.0f);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = (t0 + u * nw) * 16 + col;
            const bool live = t0 + u * nw < ntiles && row < N;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const f32x4_t a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[2 * q], xv[u],
                                                                             f32x4_t{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
                const f32x4_t a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[2 * q + 1], xv[u],
                                                                             f32x4_t{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
                if (live) {
                    bf16x8_t o;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        o[r] = __bf16(fmaxf(a0[r] + bb[2 * q][r], 0.0f));
                        o[4 + r] = __bf16(fmaxf(a1[r] + bb[2 * q + 1][r], 0.0f));
                    }
                    *reinterpret_cast<bf16x8_t *>(h0 + row * H + 32 * q + 8 * g) = o;
                }
            }
        }
    }
}

// Inference message layer: h_out = h + relu([h | mean_nb h] W + b) as k_mp_layer (same
// operand layout and K-step order), arranged for registers and latency:
//  * the rows of W^T are staged in LDS in the order p(16cb + 4g + r) = 32(cb/2) + 8g +
//    4(cb%2) + r, so the output features a lane holds (rows 4g..4g+3 of each 16-row
//    block cb) are features it already holds as its own B operand (32i + 8g + e): the
//    residual comes from registers and no row is loaded twice;
//  * the neighbour ids of the next tile are fetched while the current tile computes;
//  * the bias sits in LDS.
// LAST: the decoder is applied in the epilogue instead of storing h_out,
//   y[c][j] = b_out[j] + sum_f bf16(h_out[c][f]) W_out[f][j]     (j < 8, fp32),
// each lane summing its 32 features and the four lanes of a cell reducing by shuffles.
template <int NLT, bool LAST>
__global__ void __launch_bounds__(kLayerThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_mp_layer_inf(const __bf16 *__restrict__ h, const int32_t *__restrict__ nbr, int64_t N,
               const __bf16 *__restrict__ WT, const float *__restrict__ bias, __bf16 *__restrict__ h_out,
               const __bf16 *__restrict__ WoT, const float *__restrict__ b_out, float *__restrict__ y) {
    constexpr int H = 128, K = 2 * H, KP = K + 8, NO = 8;
    __shared__ __attribute__((aligned(16))) __bf16 Ws[H][KP];
    __shared__ __attribute__((aligned(16))) float bs[H];
    __shared__ __attribute__((aligned(16))) float wo[LAST ? H : 1][NO];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // LDS row m = 16cb + 4g + r holds output feature p(m)
    auto perm = [](int m) { const int cb = m >> 4, gg = (m >> 2) & 3, r = m & 3; return 32 * (cb >> 1) + 8 * gg + 4 * (cb & 1) + r; };
    for (int t = tid; t < H * (K / 8); t += kLayerThreads) {
        const int j = t / (K / 8), kc = (t % (K / 8)) * 8;
        *reinterpret_cast<bf16x8_t *>(&Ws[j][kc]) = *reinterpret_cast<const bf16x8_t *>(WT + perm(j) * K + kc);
    }
    for (int t = tid; t < H; t += kLayerThreads) bs[t] = bias[perm(t)];
    if constexpr (LAST) {
        for (int t = tid; t < H * NO; t += kLayerThreads) wo[t % H][t / H] = static_cast<float>(WoT[t]);
    }
    __syncthreads();
    constexpr int nwaves = kLayerThreads / 64;
    const int64_t ntiles = (N + 15) / 16;
    const int g = lane >> 4, kq = 8 * g, col = lane & 15, g4 = 4 * g;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * nwaves;
    int64_t t = static_cast<int64_t>(blockIdx.x) * nwaves + wave;
    int32_t nbn[NLT];
    {
        const int64_t row = t * 16 + col;
#pragma unroll
        for (int k = 0; k < NLT; ++k) nbn[k] = (t < ntiles && row < N) ? nbr[k * N + row] : static_cast<int32_t>(N);
    }
    for (; t < ntiles; t += stride) {
        const int64_t row = t * 16 + col;
        const bool rlive = row < N;
        const int64_t rs = rlive ? row : 0;
        // the LDS fragments do not depend on the tile; an opaque per-tile offset keeps the
        // compiler from hoisting them out of the loop (they would not fit in registers)
        int wofs = 0;
        asm volatile("" : "+v"(wofs));
        const __bf16 *wsb = &Ws[col][kq] + wofs;
        const float *wob = &wo[0][0] + wofs;
        int64_t nb[NLT];
        bool val[NLT];
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < NLT; ++k) {
            val[k] = nbn[k] < N;
            cnt += val[k];
            nb[k] = val[k] ? static_cast<int64_t>(nbn[k]) : rs;
        }
        typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
        bf16x8_t own[4], nv[NLT][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            own[i] = *reinterpret_cast<const bf16x8_t *>(h + rs * H + 32 * i + kq);
#pragma unroll
            for (int k = 0; k < NLT; ++k) nv[k][i] = *reinterpret_cast<const bf16x8_t *>(h + nb[k] * H + 32 * i + kq);
        }
        {
            const int64_t tn = t + stride, rn = tn * 16 + col;
#pragma unroll
            for (int k = 0; k < NLT; ++k) nbn[k] = (tn < ntiles && rn < N) ? nbr[k * N + rn] : static_cast<int32_t>(N);
        }
        const float inv = cnt ? 1.0f / static_cast<float>(cnt) : 0.0f;
        f32x4_t acc[8];
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) acc[cb] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!rlive) {
#pragma unroll
                for (int e = 0; e < 8; ++e) own[i][e] = __bf16(0.0f);
            }
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const bf16x8_t wf = *reinterpret_cast<const bf16x8_t *>(wsb + 16 * cb * KP + 32 * i);
                acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, own[i], acc[cb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float sacc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < NLT; ++k) {
                if (val[k]) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) sacc[e] += static_cast<float>(nv[k][i][e]);
                }
            }
            bf16x8_t mv;
#pragma unroll
            for (int e = 0; e < 8; ++e) mv[e] = __bf16(sacc[e] * inv);
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const bf16x8_t wf = *reinterpret_cast<const bf16x8_t *>(wsb + 16 * cb * KP + H + 32 * i);
                acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, mv, acc[cb], 0, 0, 0);
            }
        }
        // this lane holds LDS rows 16cb + 4g .. +3 of cell `row`: features f0 .. f0 + 3,
        // f0 = 32(cb/2) + 8g + 4(cb%2), whose inputs are own[cb/2][4(cb%2) .. +3]
        float yp[NO];
        if constexpr (LAST) {
#pragma unroll
            for (int j = 0; j < NO; ++j) yp[j] = 0.0f;
        }
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) {
            const int j0 = 32 * (cb >> 1) + kq + 4 * (cb & 1);
            const float4 bb = *reinterpret_cast<const float4 *>(&bs[16 * cb + g4]);
            const float z[4] = {acc[cb][0] + bb.x, acc[cb][1] + bb.y, acc[cb][2] + bb.z, acc[cb][3] + bb.w};
            bf16x4_t o;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                o[r] = __bf16(static_cast<float>(own[cb >> 1][4 * (cb & 1) + r]) + fmaxf(z[r], 0.0f));
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float hv = static_cast<float>(o[r]);
                    const float4 w0 = *reinterpret_cast<const float4 *>(wob + (j0 + r) * NO);
                    const float4 w1 = *reinterpret_cast<const float4 *>(wob + (j0 + r) * NO + 4);
                    yp[0] += hv * w0.x;
                    yp[1] += hv * w0.y;
                    yp[2] += hv * w0.z;
                    yp[3] += hv * w0.w;
                    yp[4] += hv * w1.x;
                    yp[5] += hv * w1.y;
                    yp[6] += hv * w1.z;
                    yp[7] += hv * w1.w;
                }
            } else {
                if (rlive) *reinterpret_cast<bf16x4_t *>(h_out + row * H + j0) = o;
            }
        }
        if constexpr (LAST) {
#pragma unroll
            for (int j = 0; j < NO; ++j) {
                yp[j] += __shfl_xor(yp[j], 16);
                yp[j] += __shfl_xor(yp[j], 32);
            }
            if (rlive && g == 0) {
                *reinterpret_cast<float4 *>(y + row * NO) =
                    make_float4(yp[0] + b_out[0], yp[1] + b_out[1], yp[2] + b_out[2], yp[3] + b_out[3]);
                *reinterpret_cast<float4 *>(y + row * NO + 4) =
                    make_float4(yp[4] + b_out[4], yp[5] + b_out[5], yp[6] + b_out[6], yp[7] + b_out[7]);
            }
        }
    }
}

__global__ void __launch_bounds__(256)
k_transpose_bf16(const float *__restrict__ W, int64_t nrow, int64_t ncol, __bf16 *__restrict__ WT) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nrow * ncol) return;
    const int64_t r = t / ncol, c = t % ncol;
    WT[c * nrow + r] = __bf16(W[t]);
}

// grid (64, n_meshes): per mesh, fp64 partial sums of the Simpson defect of the window
// d = U1 - U0 - w/6 (f0 + 4 fm + f1), of the Hermite midpoint defect
// e = Um - (U0 + U1)/2 - w/8 (f0 - f1), and of U1^2, per variable
template <int CT>
__global__ void __launch_bounds__(256)
k_mp_defect3_partial(const float *__restrict__ U0, const float *__restrict__ Um, const float *__restrict__ U1,
                     const float *__restrict__ f0, const float *__restrict__ fm, const float *__restrict__ f1,
                     const float *__restrict__ win, const int64_t *__restrict__ off, int C_rt,
                     double *__restrict__ part) {
    __shared__ double red[256];
    constexpr int CM = CT ? CT : 8;
    const int C = CT ? CT : C_rt;
    const int q = blockIdx.y;
    const int64_t a = off[q], b = off[q + 1];
    double s[3 * CM];
#pragma unroll
    for (int j = 0; j < 3 * CM; ++j) s[j] = 0.0;
    for (int64_t c = a + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < b;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float w = win[c];
#pragma unroll
        for (int j = 0; j < CM; ++j) {
            if (j >= C) continue;
            const int64_t o = c * C + j;
            const float d = (U1[o] - U0[o]) - (w / 6.0f) * ((f0[o] + 4.0f * fm[o]) + f1[o]);
            const float e = (Um[o] - 0.5f * (U0[o] + U1[o])) - (w * 0.125f) * (f0[o] - f1[o]);
            s[j] += static_cast<double>(d) * d;
            s[C + j] += static_cast<double>(e) * e;
            s[2 * C + j] += static_cast<double>(U1[o]) * U1[o];
        }
    }
#pragma unroll
    for (int j = 0; j < 3 * CM; ++j) {
        if (j >= 3 * C) continue;
        red[threadIdx.x] = s[j];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(static_cast<int64_t>(q) * gridDim.x + blockIdx.x) * 3 * C + j] = red[0];
        __syncthreads();
    }
}

}  // namespace fvm

using namespace fvm;

extern "C" {

int fvm_gemm_bf16(const void *A, int64_t a_rs, int64_t a_cs, const void *B, int64_t b_rs, int64_t b_cs, int64_t M,
                  int64_t N, int64_t K, int32_t splits, const float *bias, float *Cf, int64_t c_rs, void *Cb,
                  int32_t relu, const void *res, float *part, void *stream) {
    if (!A || !B || M < 0 || N < 0 || K < 0 || splits < 1 || splits > 65535 || (!Cf && !Cb) ||
        (splits > 1 && !part) || M > (int64_t(1) << 36) || N > (int64_t(65535) * BN)) {
        set_error("fvm_gemm_bf16: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (M == 0 || N == 0) return FVM_OK;
    int64_t k_chunk = (K + splits - 1) / splits;
    k_chunk = ((k_chunk + BK - 1) / BK) * BK;
    const int S = static_cast<int>((K + k_chunk - 1) / k_chunk > 0 ? (K + k_chunk - 1) / k_chunk : 1);
    hipStream_t s = as_stream(stream);
    dim3 grid(static_cast<unsigned int>((M + BM - 1) / BM), static_cast<unsigned int>((N + BN - 1) / BN), S);
    const __bf16 *r = static_cast<const __bf16 *>(res);
    const __bf16 *a = static_cast<const __bf16 *>(A), *b = static_cast<const __bf16 *>(B);
    __bf16 *cb = static_cast<__bf16 *>(Cb);
    const bool av = a_cs == 1 && a_rs % 8 == 0 && (reinterpret_cast<uintptr_t>(A) & 15u) == 0;
    const bool bv = b_rs == 1 && b_cs % 8 == 0 && (reinterpret_cast<uintptr_t>(B) & 15u) == 0;
    if (av && bv)
        hipLaunchKernelGGL((k_gemm_bf16<true, true>), grid, dim3(256), 0, s, a, a_rs, a_cs, b, b_rs, b_cs, M, N, K,
                           k_chunk, bias, Cf, c_rs, cb, relu, r, part);
    else if (av)
        hipLaunchKernelGGL((k_gemm_bf16<true, false>), grid, dim3(256), 0, s, a, a_rs, a_cs, b, b_rs, b_cs, M, N, K,
                           k_chunk, bias, Cf, c_rs, cb, relu, r, part);
    else if (bv)
        hipLaunchKernelGGL((k_gemm_bf16<false, true>), grid, dim3(256), 0, s, a, a_rs, a_cs, b, b_rs, b_cs, M, N, K,
                           k_chunk, bias, Cf, c_rs, cb, relu, r, part);
    else
        hipLaunchKernelGGL((k_gemm_bf16<false, false>), grid, dim3(256), 0, s, a, a_rs, a_cs, b, b_rs, b_cs, M, N, K,
                           k_chunk, bias, Cf, c_rs, cb, relu, r, part);
    if (S > 1)
        hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks_for(M * N, 256)), dim3(256), 0, s, part, S, M, N, bias, Cf,
                           c_rs, static_cast<__bf16 *>(Cb), relu, r);
    return check_launch("fvm_gemm_bf16");
}

int fvm_mp_aggregate(const void *h, int32_t H, const int32_t *nbr, int64_t N, int32_t NL, void *A, void *stream) {
    if (!h || !nbr || !A || H % 8 || N < 0 || NL < 1) {
        set_error("fvm_mp_aggregate: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    hipLaunchKernelGGL(k_mp_aggregate, dim3(blocks_for(N * (H / 8), 256)), dim3(256), 0, as_stream(stream),
                       static_cast<const __bf16 *>(h), H, nbr, N, NL, static_cast<__bf16 *>(A));
    return check_launch("fvm_mp_aggregate");
}

int fvm_mp_aggregate_bwd(const float *dA, int32_t H, const int32_t *nbr, int64_t N, int32_t NL, const float *dres,
                         float *dh, void *stream) {
    if (!dA || !nbr || !dh || N < 0 || NL < 1 || H < 1) {
        set_error("fvm_mp_aggregate_bwd: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    hipLaunchKernelGGL(k_mp_aggregate_bwd, dim3(blocks_for(N * H, 256)), dim3(256), 0, as_stream(stream), dA, H, nbr,
                       N, NL, dres, dh);
    return check_launch("fvm_mp_aggregate_bwd");
}

int fvm_relu_residual(const float *z, const void *h_in, void *h_out, int64_t n, void *stream) {
    if (!z || !h_out || n < 0) {
        set_error("fvm_relu_residual: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (n == 0) return FVM_OK;
    hipLaunchKernelGGL(k_relu_residual, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), z,
                       static_cast<const __bf16 *>(h_in), static_cast<__bf16 *>(h_out), n);
    return check_launch("fvm_relu_residual");
}

int fvm_relu_bwd(const float *dh, const float *z, void *dz_b, float *dz_f, int64_t n, void *stream) {
    if (!dh || !z || !dz_b || n < 0) {
        set_error("fvm_relu_bwd: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (n == 0) return FVM_OK;
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), dh, z,
                       static_cast<__bf16 *>(dz_b), dz_f, n);
    return check_launch("fvm_relu_bwd");
}

int fvm_colsum_f32(const float *X, int64_t M, int32_t N, float *out, float *part /* [256][N] */, void *stream) {
    if (!X || !out || !part || M < 0 || N < 1) {
        set_error("fvm_colsum_f32: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_colsum_partial, dim3(256), dim3(256), 0, s, X, M, N, part);
    hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(256), 0, s, part, 256, N, out);
    return check_launch("fvm_colsum_f32");
}

int fvm_mse_loss(const float *y, int64_t ld_y, const float *tgt, int64_t M, int32_t n_out, float *dy,
                 double *part /* [256] */, void *stream) {
    if (!y || !tgt || !dy || !part || M < 1 || n_out < 1 || ld_y < n_out) {
        set_error("fvm_mse_loss: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, as_stream(stream), y, ld_y, tgt, M, n_out, dy, part);
    return check_launch("fvm_mse_loss");
}

int fvm_adam_f32(float *w, const float *g, float *m, float *v, void *wb, int64_t n, float lr, float b1, float b2,
                 float eps, int32_t step, void *stream) {
    if (!w || !g || !m || !v || n < 0 || step < 1) {
        set_error("fvm_adam_f32: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (n == 0) return FVM_OK;
    const float bc1 = 1.0f - powf(b1, static_cast<float>(step));
    const float bc2 = 1.0f - powf(b2, static_cast<float>(step));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), w, g, m, v,
                       static_cast<__bf16 *>(wb), n, lr, b1, b2, eps, bc1, bc2);
    return check_launch("fvm_adam_f32");
}

int fvm_cast_f32_bf16(const float *x, void *y, int64_t n, void *stream) {
    if (!x || !y || n < 0) {
        set_error("fvm_cast_f32_bf16: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (n == 0) return FVM_OK;
    hipLaunchKernelGGL(k_cast_f32_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), x,
                       static_cast<__bf16 *>(y), n);
    return check_launch("fvm_cast_f32_bf16");
}

int fvm_mp_window_features(const float *prim, const float *f0, const float *fm, const float *win, const float *geo,
                           const float *hc, const int32_t *nbr, int32_t NL, const float *wg, int64_t N, int32_t C,
                           const float *mu_p, const float *sd_p, const float *s1, const float *s2,
                           const int32_t *cell_mesh, void *X, int32_t KX, void *stream) {
    if (!prim || !f0 || !fm || !win || !geo || !hc || !nbr || !wg || NL != 3 || C != 4 || !mu_p || !sd_p || !s1 ||
        !s2 || !X || N < 0 || KX != 32 || (reinterpret_cast<uintptr_t>(X) & 15u) ||
        ((reinterpret_cast<uintptr_t>(prim) | reinterpret_cast<uintptr_t>(f0) | reinterpret_cast<uintptr_t>(fm)) &
         15u)) {
        set_error("fvm_mp_window_features: invalid arguments (this build takes 2-D meshes: NL = 3, C = 4, KX = 32, "
                  "16-byte aligned rows)");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    hipLaunchKernelGGL((k_mp_window_features<3>), dim3(blocks_for(N, 256)), dim3(256), 0, as_stream(stream),
                       reinterpret_cast<const float4 *>(prim), reinterpret_cast<const float4 *>(f0),
                       reinterpret_cast<const float4 *>(fm), win, geo, hc, nbr, wg, N, mu_p, sd_p, s1, s2, cell_mesh,
                       static_cast<__bf16 *>(X));
    return check_launch("fvm_mp_window_features");
}

int fvm_mp_window_apply(const float *U, const float *f0, const float *fm, float a, float b, const float *win,
                        const float *y, int64_t ld_y, const float *scale, const int32_t *cell_mesh, int64_t N,
                        int32_t C, float *U_out, void *stream) {
    if (!U || !f0 || !win || (y && (!scale || ld_y < C)) || !U_out || N < 0 || C < 1) {
        set_error("fvm_mp_window_apply: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    hipLaunchKernelGGL(k_mp_window_apply, dim3(blocks_for(N * C, 256)), dim3(256), 0, as_stream(stream), U, f0, fm, a,
                       b, win, y, ld_y, scale, cell_mesh, N, C, U_out);
    return check_launch("fvm_mp_window_apply");
}

int fvm_mp_window_predict(const float *U, const float *f0, const float *fm, float am, float bm, float ae, float be,
                          const float *win, const float *y, int64_t ld_y, const float *scale,
                          const int32_t *cell_mesh, int64_t N, int32_t C, float *Um, float *U1, void *stream) {
    if (!U || !f0 || !fm || !win || (y && (!scale || ld_y < 2 * C)) || !Um || !U1 || N < 0 || C < 1) {
        set_error("fvm_mp_window_predict: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    hipLaunchKernelGGL(k_mp_window_predict, dim3(blocks_for(N * C, 256)), dim3(256), 0, as_stream(stream), U, f0, fm,
                       am, bm, ae, be, win, y, ld_y, scale, cell_mesh, N, C, Um, U1);
    return check_launch("fvm_mp_window_predict");
}

int fvm_mesh_rms2(const float *A, const float *B, const float *win, const int64_t *off, int32_t n_meshes, int32_t C,
                  float floor_abs, float *out1, float *out2, double *part, void *stream) {
    if (!A || !B || !win || !off || !out1 || !out2 || !part || n_meshes < 1 || n_meshes > 65535 || C < 1 || C > 8) {
        set_error("fvm_mesh_rms2: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipStream_t s = as_stream(stream);
    if (C == 4)
        hipLaunchKernelGGL(k_mesh_sumsq2_partial<4>, dim3(64, n_meshes), dim3(256), 0, s, A, B, win, off, C, part);
    else if (C == 5)
        hipLaunchKernelGGL(k_mesh_sumsq2_partial<5>, dim3(64, n_meshes), dim3(256), 0, s, A, B, win, off, C, part);
    else
        hipLaunchKernelGGL(k_mesh_sumsq2_partial<0>, dim3(64, n_meshes), dim3(256), 0, s, A, B, win, off, C, part);
    hipLaunchKernelGGL(k_mesh_rms2_final, dim3(blocks_for(static_cast<int64_t>(n_meshes) * 2 * C, 256)), dim3(256), 0,
                       s, part, 64, n_meshes, C, off, floor_abs, out1, out2);
    return check_launch("fvm_mesh_rms2");
}

int fvm_mesh_rms(const float *X, const float *win, const int64_t *off, int32_t n_meshes, int32_t C, float floor_abs,
                 float *out, double *part, void *stream) {
    if (!X || !off || !out || !part || n_meshes < 1 || n_meshes > 65535 || C < 1 || C > 8) {
        set_error("fvm_mesh_rms: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipStream_t s = as_stream(stream);
    if (C == 4)
        hipLaunchKernelGGL(k_mesh_sumsq_partial<4>, dim3(64, n_meshes), dim3(256), 0, s, X, win, off, C, part);
    else if (C == 5)
        hipLaunchKernelGGL(k_mesh_sumsq_partial<5>, dim3(64, n_meshes), dim3(256), 0, s, X, win, off, C, part);
    else
        hipLaunchKernelGGL(k_mesh_sumsq_partial<0>, dim3(64, n_meshes), dim3(256), 0, s, X, win, off, C, part);
    hipLaunchKernelGGL(k_mesh_rms_final, dim3(blocks_for(static_cast<int64_t>(n_meshes) * C, 256)), dim3(256), 0, s,
                       part, 64, n_meshes, C, off, floor_abs, out);
    return check_launch("fvm_mesh_rms");
}

int fvm_mp_defect(const float *U0, const float *U1, const float *f0, const float *f1, const float *win,
                  const int64_t *off, int32_t n_meshes, int32_t C, double *out, double *part, void *stream) {
    if (!U0 || !U1 || !f0 || !f1 || !win || !off || !out || !part || n_meshes < 1 || n_meshes > 65535 || C < 1 ||
        C > 8) {
        set_error("fvm_mp_defect: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipStream_t s = as_stream(stream);
    if (C == 4)
        hipLaunchKernelGGL(k_mp_defect_partial<4>, dim3(64, n_meshes), dim3(256), 0, s, U0, U1, f0, f1, win, off, C, part);
    else if (C == 5)
        hipLaunchKernelGGL(k_mp_defect_partial<5>, dim3(64, n_meshes), dim3(256), 0, s, U0, U1, f0, f1, win, off, C, part);
    else
        hipLaunchKernelGGL(k_mp_defect_partial<0>, dim3(64, n_meshes), dim3(256), 0, s, U0, U1, f0, f1, win, off, C, part);
    hipLaunchKernelGGL(k_mp_defect_final, dim3(blocks_for(static_cast<int64_t>(n_meshes) * 2 * C, 256)), dim3(256), 0,
                       s, part, 64, n_meshes, 2 * C, out);
    return check_launch("fvm_mp_defect");
}

int fvm_mp_layer(const void *h_in, int32_t H, const int32_t *nbr, int64_t N, int32_t NL, const void *WT,
                 const float *bias, void *h_out, float *Z, void *A_out, void *stream) {
    if (!h_in || !nbr || !WT || !bias || !h_out || H != 128 || N < 0 || (NL != 3 && NL != 4) || h_in == h_out) {
        set_error("fvm_mp_layer: invalid arguments (H must be 128, NL 3 or 4, h_out distinct from h_in)");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    // persistent: two workgroups per CU (the LDS copy of W^T is 66 KB) or fewer if
    // there are fewer tiles
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n_cu = v;
        else
            n_cu = 256;
    }
    const int64_t wave_tiles = (N + 15) / 16;
    const int64_t blocks = (wave_tiles + (kLayerThreads / 64) - 1) / (kLayerThreads / 64);
    const dim3 grid(static_cast<unsigned int>(blocks < 2 * n_cu ? blocks : 2 * n_cu));
    const __bf16 *hi = static_cast<const __bf16 *>(h_in);
    const __bf16 *wt = static_cast<const __bf16 *>(WT);
    __bf16 *ho = static_cast<__bf16 *>(h_out), *ao = static_cast<__bf16 *>(A_out);
    if (NL == 3)
        hipLaunchKernelGGL(k_mp_layer<3>, grid, dim3(kLayerThreads), 0, as_stream(stream), hi, nbr, N, wt, bias, ho, Z,
                           ao);
    else
        hipLaunchKernelGGL(k_mp_layer<4>, grid, dim3(kLayerThreads), 0, as_stream(stream), hi, nbr, N, wt, bias, ho, Z,
                           ao);
    return check_launch("fvm_mp_layer");
}

static int layer_grid_cap() {
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n_cu = v;
        else
            n_cu = 256;
    }
    return n_cu;
}

int fvm_mp_encode(const void *X, int64_t N, int32_t KX, const void *WT, const float *bias, int32_t H, void *h0,
                  void *stream) {
    if (!X || !WT || !bias || !h0 || KX != 32 || H != 128 || N < 0 || (reinterpret_cast<uintptr_t>(X) & 15u) ||
        (reinterpret_cast<uintptr_t>(WT) & 15u) || (reinterpret_cast<uintptr_t>(h0) & 7u)) {
        set_error("fvm_mp_encode: invalid arguments (KX must be 32, H 128, aligned rows)");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    const int64_t tiles = (N + 15) / 16;
    const int64_t blocks = (tiles + 3) / 4;
    const int64_t cap = 8LL * layer_grid_cap();
    hipLaunchKernelGGL(k_mp_encode, dim3(static_cast<unsigned int>(blocks < cap ? blocks : cap)), dim3(256), 0,
                       as_stream(stream), static_cast<const __bf16 *>(X), N, static_cast<const __bf16 *>(WT), bias,
                       static_cast<__bf16 *>(h0));
    return check_launch("fvm_mp_encode");
}

int fvm_mp_layer_inf(const void *h_in, int32_t H, const int32_t *nbr, int64_t N, int32_t NL, const void *WT,
                     const float *bias, void *h_out, const void *WoT, const float *b_out, int32_t n_out, float *y,
                     void *stream) {
    const bool last = y != nullptr;
    if (!h_in || !nbr || !WT || !bias || H != 128 || N < 0 || (NL != 3 && NL != 4) ||
        (!last && (!h_out || h_out == h_in)) || (last && (!WoT || !b_out || n_out != 8)) ||
        (reinterpret_cast<uintptr_t>(h_in) & 15u) || (reinterpret_cast<uintptr_t>(WT) & 15u) ||
        (y && (reinterpret_cast<uintptr_t>(y) & 15u))) {
        set_error("fvm_mp_layer_inf: invalid arguments (H must be 128, NL 3 or 4; with y: WoT, b_out and n_out = 8; "
                  "without: h_out distinct from h_in)");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    const int64_t wave_tiles = (N + 15) / 16;
    const int64_t blocks = (wave_tiles + (kLayerThreads / 64) - 1) / (kLayerThreads / 64);
    const int64_t cap = 2LL * layer_grid_cap();
    const dim3 grid(static_cast<unsigned int>(blocks < cap ? blocks : cap));
    const __bf16 *hi = static_cast<const __bf16 *>(h_in);
    const __bf16 *wt = static_cast<const __bf16 *>(WT);
    const __bf16 *wo = static_cast<const __bf16 *>(WoT);
    __bf16 *ho = static_cast<__bf16 *>(h_out);
    hipStream_t s = as_stream(stream);
    if (NL == 3) {
        if (last)
            hipLaunchKernelGGL((k_mp_layer_inf<3, true>), grid, dim3(kLayerThreads), 0, s, hi, nbr, N, wt, bias, ho,
                               wo, b_out, y);
        else
            hipLaunchKernelGGL((k_mp_layer_inf<3, false>), grid, dim3(kLayerThreads), 0, s, hi, nbr, N, wt, bias, ho,
                               wo, b_out, y);
    } else {
        if (last)
            hipLaunchKernelGGL((k_mp_layer_inf<4, true>), grid, dim3(kLayerThreads), 0, s, hi, nbr, N, wt, bias, ho,
                               wo, b_out, y);
        else
            hipLaunchKernelGGL((k_mp_layer_inf<4, false>), grid, dim3(kLayerThreads), 0, s, hi, nbr, N, wt, bias, ho,
                               wo, b_out, y);
    }
    return check_launch("fvm_mp_layer_inf");
}

int fvm_transpose_bf16(const float *W, int64_t nrow, int64_t ncol, void *WT, void *stream) {
    if (!W || !WT || nrow < 0 || ncol < 0) {
        set_error("fvm_transpose_bf16: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (nrow * ncol == 0) return FVM_OK;
    hipLaunchKernelGGL(k_transpose_bf16, dim3(blocks_for(nrow * ncol, 256)), dim3(256), 0, as_stream(stream), W, nrow,
                       ncol, static_cast<__bf16 *>(WT));
    return check_launch("fvm_transpose_bf16");
}

int fvm_mp_defect3(const float *U0, const float *Um, const float *U1, const float *f0, const float *fm, const float *f1,
                   const float *win, const int64_t *off, int32_t n_meshes, int32_t C, double *out, double *part,
                   void *stream) {
    if (!U0 || !Um || !U1 || !f0 || !fm || !f1 || !win || !off || !out || !part || n_meshes < 1 ||
        n_meshes > 65535 || C < 1 || C > 8) {
        set_error("fvm_mp_defect3: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipStream_t s = as_stream(stream);
    if (C == 4)
        hipLaunchKernelGGL(k_mp_defect3_partial<4>, dim3(64, n_meshes), dim3(256), 0, s, U0, Um, U1, f0, fm, f1, win, off,
                           C, part);
    else if (C == 5)
        hipLaunchKernelGGL(k_mp_defect3_partial<5>, dim3(64, n_meshes), dim3(256), 0, s, U0, Um, U1, f0, fm, f1, win, off,
                           C, part);
    else
        hipLaunchKernelGGL(k_mp_defect3_partial<0>, dim3(64, n_meshes), dim3(256), 0, s, U0, Um, U1, f0, fm, f1, win, off,
                           C, part);
    hipLaunchKernelGGL(k_mp_defect_final, dim3(blocks_for(static_cast<int64_t>(n_meshes) * 3 * C, 256)), dim3(256), 0,
                       s, part, 64, n_meshes, 3 * C, out);
    return check_launch("fvm_mp_defect3");
}

}  // extern "C"
