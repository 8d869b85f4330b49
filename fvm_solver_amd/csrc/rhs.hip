// Fused right-hand side of the compressible Navier-Stokes FVM step for gfx950.
//
// One RHS evaluation (FVMEquation.forward, fvm_equation.py:406-435) is three launches
// (plus, only when the boundary means are not already known, a small pre-pass):
//   k_bc_mean    (pre-pass) per characteristic-BC group, partial sums of the speed of
//                sound of the boundary cells (for the group-wide mean of
//                facet_boundary.py:427) in fixed per-tile slots; k_grad reduces them in
//                a fixed order, so the mean is deterministic
//   k_grad       per cell: the values of its boundary facets -- A_bc.x + b_bc, the
//                Neumann non-orthogonal correction with the previous call's limited
//                gradients, the characteristic BC (facet_boundary.py:124-168, 418-512)
//                -- then the LSQ gradient G.[U; U_bc] and the slope limiter
//                (facet_process.py:387-418, limiter.py); stores the unlimited gradient
//                and the limiter factor phi, from which the facet kernel rebuilds the
//                limited face values and gradients with the reference's operation order
//   k_facet_flux per facet: face reconstruction, normal-gradient correction, stress,
//                advection, pressure, viscosity, heating, KT dissipation
//                (facet_process.py:332-385, fvm_equation.py:62-368) -> one flux record
//   k_div        per cell: the divergence (flux_mat row) of its facets' fluxes, then an
//                epilogue: write dU/dt, or apply an explicit Runge-Kutta stage update,
//                emit the next stage's primitives and the boundary means of those
//                primitives (so the next stage needs no pre-pass).
// The facet flux costs ~500 VALU operations; evaluating it once per facet and passing it
// through HBM (16-20 bytes per facet, written once, read twice) is cheaper than
// evaluating it from both sides.
//
// This file is compiled with -ffp-contract=off: products and sums are rounded
// separately, in the order the reference's ATen op sequence performs them, so that the
// branchy limiter sees the same inputs as the CPU restatement.  Two places trade exact
// rounding for VALU work (the stage is VALU-bound): the facet flux uses explicit fused
// multiply-adds, hardware square roots and reciprocals and takes (T/T0)^1.5 as
// x sqrt(x), and the limiter factors divide through the hardware reciprocal.  Each
// costs about one ulp per operation; the 1000-step parity bar is unaffected.
#include "fvm_common.hpp"

namespace fvm {

constexpr int kThreads = 256;
constexpr int kMaxGroups = 4;
// Meshes of a batch whose physics and time step a tile of the fused stage keeps in LDS
// (cells are ordered by mesh: a 256-cell tile spans its first mesh and the next ones;
// facets and cells of further meshes read them from global memory).
constexpr int kTileMeshes = 4;

template <int D>
struct Dims {
    static constexpr int C = D + 2;
    static constexpr int NL = D + 1;
    static constexpr int REC = (D + 1) * (D + 2);  // gp record: g[D][C] then phi[C]
    static constexpr int GEO = D == 2 ? 12 : 20;  // n, n_hat, |n|, 1/(n.d), X, d_slot0, d_slot1
    static constexpr int CH = 2 * D + 3;
};

// Quotient through the hardware reciprocal (1 ulp) and one product: the limiter factors
// are smooth in their inputs (continuous across their branch), and an IEEE division
// costs ~10 VALU operations; ~12 of them per cell dominated the limiter.  The default
// limiter (p = 4) also fuses its multiply-adds.
__device__ __forceinline__ float quot(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Square root by the hardware instruction (1 ulp; the correctly rounded sequence costs
// ~14 VALU operations).  Arguments here are temperatures and squared speeds: normal
// floats, never denormal.
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int lp>
__device__ __forceinline__ float lim_phi(float delta, float dU, float eps) {
    if constexpr (lp == 4) {
        const float a = fabsf(delta), b = fabsf(dU);
        const float b2 = 2.0f * b;
        const float a2 = a * a;
        const float a_eps = fmaf(a2, a2, eps);
        const float S = b2 * fmaf(b2, b2 - a, a2);
        const float d3 = a2 * delta;        // (delta delta) delta: a2 = |delta|^2 is delta^2 exactly
        const float phi = quot(fmaf(a, S, a_eps), fmaf(b, d3 + S, a_eps));
        return a < b2 ? phi : 1.0f;
    }
    if constexpr (lp == 1) {
        const float s = 2.0f * ((dU > 0.0f ? 1.0f : 0.0f) - 0.5f);
        const float r = quot(delta, s * (fabsf(dU) + 1e-8f));
        return fminf(fmaxf(r, 0.0f), 1.0f);
    }
    if constexpr (lp == 2) {
        const float d2 = delta * delta;
        return quot((d2 + eps) + (2.0f * delta) * dU, ((d2 + 2.0f * (dU * dU)) + delta * dU) + eps);
    }
    if constexpr (lp == 3) {
        const float a = fabsf(delta), b = fabsf(dU);
        const float a_eps = (a * a) * a + eps;
        const float S = 4.0f * (b * b);
        const float phi = quot(a_eps + a * S, a_eps + b * (delta * delta + S));
        return a < 2.0f * b ? phi : 1.0f;
    }
    if constexpr (lp == 5) {
        const float a = fabsf(delta), b = fabsf(dU);
        const float a2 = a * a, d2 = delta * delta;
        const float a_eps = (a2 * a2) * a + eps;
        const float S = (8.0f * (b * b)) * (a2 - (2.0f * b) * (a - b));
        const float phi = quot(a_eps + a * S, a_eps + b * (d2 * d2 + S));
        return a < 2.0f * b ? phi : 1.0f;
    }
    static_assert(lp >= 1 && lp <= 5, "limiter order 1..5");
    return 1.0f;
}

// X_orthog = d / (n_hat.d) - n_hat (facet_process.py:88-95) is orthogonal to n_hat
// (X.n_hat = 1 - |n_hat|^2 = 0), so the records hold its D - 1 coordinates in a tangent
// basis of n_hat: in 2-D t = (-n_hat_y, n_hat_x); in 3-D the basis of Duff et al.,
// "Building an Orthonormal Basis, Revisited" (JCGT 6(1), 2017), with sign(+-0) = +1.
template <int D>
__device__ __forceinline__ void tangent_X(const float (&nh)[D], const float (&xt)[D - 1], float (&X)[D]) {
    if constexpr (D == 2) {
        X[0] = -nh[1] * xt[0];
        X[1] = nh[0] * xt[0];
    } else {
        const float sg = nh[2] >= 0.0f ? 1.0f : -1.0f;
        const float a = -__builtin_amdgcn_rcpf(sg + nh[2]);
        const float b = nh[0] * nh[1] * a;
        const float t1[3] = {fmaf(sg * nh[0] * nh[0], a, 1.0f), sg * b, -sg * nh[0]};
        const float t2[3] = {b, fmaf(nh[1] * nh[1], a, sg), -nh[1]};
#pragma unroll
        for (int i = 0; i < 3; ++i) X[i] = fmaf(xt[0], t1[i], xt[1] * t2[i]);
    }
}

// A facet's geometry record as the flux takes it, [n (D), n_hat (D), |n|, 1/(n_hat.d),
// X_orthog (D)], from n, 1/(n_hat.d) and X_orthog's tangent coordinates: |n| and n_hat
// are derived from the fp32 normal as the reference's FacetCalc derives them (facet_size
// = |n|, normals_hat = n / facet_size, facet_process.py:84-86), through the hardware
// square root and reciprocal (1 ulp each).  Both pipelines call this, so they stay
// bitwise alike.
template <int D>
__device__ __forceinline__ void facet_geo_full(const float (&n)[D], float inv_d, const float (&xt)[D - 1],
                                               float (&geo)[3 * D + 2]) {
    float ss = n[0] * n[0];
#pragma unroll
    for (int i = 1; i < D; ++i) ss = fmaf(n[i], n[i], ss);
    const float fs = sqrt_hw(ss);
    const float ir = __builtin_amdgcn_rcpf(fs);
    float nh[D], X[D];
#pragma unroll
    for (int i = 0; i < D; ++i) nh[i] = n[i] * ir;
    tangent_X<D>(nh, xt, X);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        geo[i] = n[i];
        geo[D + i] = nh[i];
        geo[2 * D + 2 + i] = X[i];
    }
    geo[2 * D] = fs;
    geo[2 * D + 1] = inv_d;
}

// Workgroups are dealt round-robin to the 8 XCDs (workgroup w runs on XCD w % 8), each
// with its own L2.  Give each XCD a contiguous range of tiles instead, so that the
// neighbour gathers of Morton-adjacent tiles hit the L2 that already holds them.
// Bijective on [0, T) for virtual ids v < T; a persistent grid must be a multiple of 8.
__device__ __forceinline__ int64_t xcd_tile(int64_t v, int64_t T) {
    const int64_t x = v & 7, local = v >> 3, q = T >> 3, r = T & 7;
    return x * q + (x < r ? x : r) + local;
}


// Per-cell records of C floats; 2-D records (C = 4) are 16-byte aligned (the host
// wrappers check the base pointers) and move as one dwordx4.
template <int C>
__device__ __forceinline__ void load_vec(const float *__restrict__ src, float (&u)[C]) {
    if constexpr (C == 4) {
        const float4 q = *reinterpret_cast<const float4 *>(src);
        u[0] = q.x;
        u[1] = q.y;
        u[2] = q.z;
        u[3] = q.w;
    } else {
#pragma unroll
        for (int j = 0; j < C; ++j) u[j] = src[j];
    }
}

template <int C>
__device__ __forceinline__ void store_vec(float *__restrict__ dst, const float (&u)[C]) {
    if constexpr (C == 4) {
        *reinterpret_cast<float4 *>(dst) = make_float4(u[0], u[1], u[2], u[3]);
    } else {
#pragma unroll
        for (int j = 0; j < C; ++j) dst[j] = u[j];
    }
}

template <int D>
__device__ __forceinline__ void load_prim(const float *__restrict__ src, float (&u)[D + 2]) {
    load_vec<D + 2>(src, u);
}

// FluidConstitution.state_to_primative (fvm_equation.py:35-44) with one IEEE division
// (1/rho, then products: one extra rounding per component against the reference's
// quotients).
template <int D>
__device__ __forceinline__ void state_to_prim(const float (&U)[D + 2], float inv_cv, float (&P)[D + 2]) {
    const float rho = U[D];
    const float ir = 1.0f / rho;
    float sq = 0.0f;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        P[i] = U[i] * ir;
        sq = (i == 0) ? P[i] * P[i] : sq + P[i] * P[i];
    }
    P[D] = rho;
    P[D + 1] = inv_cv * (U[D + 1] * ir - 0.5f * sq);
}

// ------------------------------------------------------------------------------------
// Boundary face values
// ------------------------------------------------------------------------------------
// Values of boundary facet b from its interior cell's primitives P and gradient record
// gprev of the previous call (facet_boundary.py:124-168, characteristic BC :436-512).
template <int D>
__device__ __forceinline__ void bc_value(const FvmMeshDev &m, const FvmPhys &p, int64_t b, const float (&P)[D + 2],
                                         const float *__restrict__ gprev, int has_prev, const float *gmean,
                                         int gbase, float (&u)[D + 2]) {
    using Dm = Dims<D>;
    constexpr int C = Dm::C, CH = Dm::CH, REC = Dm::REC;
    // every row this facet may need is loaded up front, in one round trip: the value and
    // Neumann rows, the cell's previous gradient record (used under has_prev) and the
    // characteristic row (used by group facets; every boundary facet has one)
    const int flags = m.bc_flags[b];
    const int grp = m.bc_group[b];
    float bv[C], l[D], gr[REC], ch[CH];
#pragma unroll
    for (int j = 0; j < C; ++j) bv[j] = m.bc_val[b * C + j];
#pragma unroll
    for (int i = 0; i < D; ++i) l[i] = m.bc_l[b * D + i];
#pragma unroll
    for (int q = 0; q < REC; ++q) gr[q] = gprev[q];
#pragma unroll
    for (int q = 0; q < CH; ++q) ch[q] = m.bc_char[b * CH + q];
#pragma unroll
    for (int j = 0; j < C; ++j) u[j] = ((flags >> j) & 1) ? P[j] + bv[j] : bv[j];
    if (has_prev && flags) {
#pragma unroll
        for (int j = 0; j < C; ++j) {
            if ((flags >> j) & 1) {
                const float ph = gr[D * C + j];
                float s = l[0] * (gr[j] * ph);
#pragma unroll
                for (int i = 1; i < D; ++i) s = s + l[i] * (gr[i * C + j] * ph);
                u[j] = u[j] + s;
            }
        }
    }
    if (grp < 0) return;
    const float cmean = gmean[grp - gbase];
    float n[D], vt[D];
    float vn = 0.0f;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        n[i] = ch[i];
        vn = (i == 0) ? P[i] * n[i] : vn + P[i] * n[i];
    }
#pragma unroll
    for (int i = 0; i < D; ++i) vt[i] = P[i] - vn * n[i];
    const float rho = P[D], T = P[D + 1];
    const float pr = (p.R * rho) * T;
    const float c = sqrtf(p.gR * T);
    const float drho = ch[D] - rho;
    const float dvn = ch[D + 1] - vn;
    const float dp = ch[D + 2] - pr;
    const float inv_c = 1.0f / c;
    const float inv_c2 = inv_c * inv_c;
    const float h_dp = (0.5f * inv_c2) * dp;
    const float h_dvn = ((0.5f * rho) * inv_c) * dvn;
    float x0 = -h_dvn + h_dp;
    float x1 = drho - inv_c2 * dp;
    float x2 = h_dvn + h_dp;
    const float scale = 10.0f / cmean;
    const float g0 = 0.5f * (1.0f - tanhf(scale * (vn - c)));
    const float g1 = 0.5f * (1.0f - tanhf(scale * vn));
    const float g2 = 0.5f * (1.0f - tanhf(scale * (vn + c)));
    x0 = g0 * x0;
    x1 = g1 * x1;
    x2 = g2 * x2;
    const float c2 = c * c;
    const float cor = c / rho;
    const float rho_b = rho + ((x0 + x1) + x2);
    const float vn_b = vn + ((-cor) * x0 + cor * x2);
    const float p_b = pr + c2 * (x0 + x2);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const float vtb = g1 * vt[i] + (1.0f - g1) * ch[D + 3 + i];
        u[i] = vn_b * n[i] + vtb;
    }
    u[D] = rho_b;
    u[D + 1] = p_b / (p.R * rho_b);
}

// Block-level partial sums of the speed of sound over the characteristic-BC facets:
// every thread of the block calls it (uniform control flow).  A slot is a 256-cell tile
// holding interior cells of group facets; the groups of the meshes its cells belong to
// are slot_g0 .. slot_g0 + slot_ng - 1 (<= 4).  Thread contributions go through a fixed
// tree into part[slot][q].  The consumer (the gradient phase of the next evaluation) sums
// the partials of a group in a fixed order, so no grid-wide fence or counter is needed
// and the kernel boundary orders the writes before the reads.  A stage epilogue writes
// into work->bc_part_next, never into the work->bc_part its own launch reads: tiles of
// one launch run in no fixed order, and a tile that starts after another tile's epilogue
// must still see the partials of the stage's input.  red holds one float per thread of
// the workgroup; the tree sums the first kThreads (a workgroup of 2 kThreads threads has
// its tile cells on those; the others are never live).
template <int D>
__device__ void bc_mean_block(const FvmMeshDev &m, const FvmPhys &p, float *__restrict__ part, int64_t c, bool live,
                              float T, int slot, float *red) {
    constexpr int NL = D + 1;
    const int64_t N = m.n_cells;
    const int g0 = m.slot_g0[slot], ng = m.slot_ng[slot];
    float cs[kMaxGroups] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (live) {
        const float cc = sqrtf(p.gR * T);
        for (int k = 0; k < NL; ++k) {
            const int64_t nb = m.cell_nbr[k * N + c];
            if (nb >= N) {
                const int q0 = m.bc_group[nb - N] - g0;
#pragma unroll
                for (int q = 0; q < kMaxGroups; ++q)
                    if (q0 == q) cs[q] += cc;
            }
        }
    }
    for (int q = 0; q < ng; ++q) {
        float v = cs[0];
#pragma unroll
        for (int r = 1; r < kMaxGroups; ++r) v = (q == r) ? cs[r] : v;
        red[threadIdx.x] = v;
        __syncthreads();
        for (int h = kThreads / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[slot * kMaxGroups + q] = red[0];
        __syncthreads();
    }
}

// Group means (block-uniform call): mean[q] for groups g0 + q, q < ng, each the sum of
// its mesh's slot partials in a fixed strided-then-tree order (the same bits in every
// block) divided by the group's facet count (facet_boundary.py:427).
template <int NT = kThreads>
__device__ void block_group_means_g(const FvmMeshDev &m, const FvmWork &w, int g0, int ng, float *gmean,
                                    float *red) {
    __syncthreads();                       // the previous tile's readers are done
    for (int q = 0; q < ng; ++q) {
        const int g = g0 + q;
        float v = 0.0f;
        // the first kThreads threads of a larger workgroup sum, in the same order
        for (int i = m.group_s0[g] + threadIdx.x; (NT == kThreads || threadIdx.x < kThreads) && i < m.group_s1[g];
             i += kThreads)
            v += w.bc_part[i * kMaxGroups + (g - m.slot_g0[i])];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int h = kThreads / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) gmean[q] = red[0] / m.group_count[g];
        __syncthreads();
    }
}

// Group means of a tile's boundary-mean slot: the groups of the meshes it touches.
__device__ void block_group_means(const FvmMeshDev &m, const FvmWork &w, int slot, float *gmean, float *red) {
    block_group_means_g(m, w, m.slot_g0[slot], m.slot_ng[slot], gmean, red);
}

// Pre-pass: boundary means of `prim` when no previous stage produced them.
template <int D>
__global__ void __launch_bounds__(kThreads) k_bc_mean(FvmMeshDev m, FvmPhys p, FvmWork w,
                                                      const float *__restrict__ prim) {
    const int64_t c = static_cast<int64_t>(m.slot_blk[blockIdx.x]) * kThreads + threadIdx.x;
    const bool live = c < m.n_cells;
    const float T = live ? prim[c * (D + 2) + D + 1] : 0.0f;
    __shared__ float red[kThreads];
    bc_mean_block<D>(m, p, w.bc_part, c, live, T, blockIdx.x, red);
}

// ------------------------------------------------------------------------------------
// LSQ gradient + slope limiter
// ------------------------------------------------------------------------------------
// Gradient G.[U; U_bc] of one cell (its G_mats row, stored in slot order: the value of
// each facet slot's neighbour or boundary column, then the cell's own column) and the
// limiter factor phi = min over the cell's facets (facet_process.py:387-418,
// limiter.py).  P: the cell's primitives; Un: the values behind its facets (neighbour
// primitives, or boundary values).  The row is summed in slot order with fused
// multiply-adds: fixed register operands, where the reference's ascending column order
// would select each column's source per cell.
// With FACE, also the limited face values at the cell's facets, Uf[k][j] = P[j] + phi[j]
// dU[k][j], from the same reconstruction dU the limiter used (side_values' formula).
template <int D, int LP, bool FACE = false>
__device__ __forceinline__ void cell_gradient(const float (&P)[D + 2], const float (&Un)[D + 1][D + 2],
                                              const float (&gv)[D + 2][D], const float (&disp)[D + 1][D], float eps,
                                              float (&g)[D][D + 2], float (&phi)[D + 2],
                                              float (*Uf)[D + 2] = nullptr) {
    constexpr int C = D + 2, NL = D + 1;
#pragma unroll
    for (int jj = 0; jj <= NL; ++jj) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
#pragma unroll
            for (int j = 0; j < C; ++j) {
                const float u = jj < NL ? Un[jj < NL ? jj : 0][j] : P[j];
                g[i][j] = (jj == 0) ? gv[jj][i] * u : fmaf(gv[jj][i], u, g[i][j]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < C; ++j) {
        float mx = P[j], mn = P[j];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            mx = fmaxf(mx, Un[k][j]);
            mn = fminf(mn, Un[k][j]);
        }
        const float upper = mx - P[j], lower = mn - P[j];
        float ph = 0.0f;
        float dUk[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            float dU = g[0][j] * disp[k][0];
#pragma unroll
            for (int i = 1; i < D; ++i) dU = fmaf(g[i][j], disp[k][i], dU);
            dUk[k] = dU;
            const float lk = lim_phi<LP>(dU > 0.0f ? upper : lower, dU, eps);
            ph = (k == 0) ? lk : fminf(ph, lk);
        }
        phi[j] = ph;
        if constexpr (FACE) {
#pragma unroll
            for (int k = 0; k < NL; ++k) Uf[k][j] = fmaf(dUk[k], ph, P[j]);
        }
    }
}

template <int D, int LP>
__global__ void __launch_bounds__(kThreads)
k_grad(FvmMeshDev m, FvmPhys p, FvmWork w, const float *__restrict__ prim, int has_prev) {
    using Dm = Dims<D>;
    constexpr int C = Dm::C, NL = Dm::NL;
    const int64_t N = m.n_cells;
    const int64_t ntiles = (N + kThreads - 1) / kThreads;
    float *__restrict__ gp = w.gp;
    // the previous call's records: a separate buffer, or (in place) gp itself -- each
    // cell reads its own record before overwriting it
    const float *gprev = w.gp_prev ? w.gp_prev : w.gp;
    __shared__ float gmean[kMaxGroups];
    __shared__ float red[kThreads];
    for (int64_t v = blockIdx.x; v < ntiles; v += gridDim.x) {
    const int64_t tile = xcd_tile(v, ntiles);
    const int64_t c = tile * kThreads + threadIdx.x;
    const int slot = m.n_groups > 0 ? m.blk_slot[tile] : -1;
    if (slot >= 0) block_group_means(m, w, slot, gmean, red);
    const int g0 = slot >= 0 ? m.slot_g0[slot] : 0;
    if (c >= N) continue;
    // every per-cell load is issued before the neighbour gathers that depend on them,
    // and the gathers are branch-free (a boundary slot gathers the cell itself and is
    // overwritten below), so a cell costs two memory round trips
    float P[C];
    load_prim<D>(prim + c * C, P);
    int64_t nbk[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) nbk[k] = m.cell_nbr[k * N + c];
    float gv[NL + 1][D];
#pragma unroll
    for (int jj = 0; jj <= NL; ++jj)
#pragma unroll
        for (int i = 0; i < D; ++i) gv[jj][i] = m.cell_gval[(jj * D + i) * N + c];
    float disp[NL][D];
#pragma unroll
    for (int k = 0; k < NL; ++k)
#pragma unroll
        for (int i = 0; i < D; ++i) disp[k][i] = m.cell_disp[(k * D + i) * N + c];
    const float eps = m.cell_eps[c];
    float Un[NL][C];
#pragma unroll
    for (int k = 0; k < NL; ++k) load_prim<D>(prim + (nbk[k] < N ? nbk[k] : c) * C, Un[k]);
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        if (nbk[k] >= N) {
            // this cell is the boundary facet's only cell: it computes the facet's values
            // (reading its own previous gradient record before overwriting it below)
            const int64_t b = nbk[k] - N;
            bc_value<D>(m, p, b, P, gprev + c * Dm::REC, has_prev, gmean, g0, Un[k]);
            store_vec<C>(w.ubc + b * C, Un[k]);
        }
    }
    float g[D][C], phi[C];
    cell_gradient<D, LP>(P, Un, gv, disp, eps, g, phi);
    float *o = gp + c * Dm::REC;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < C; ++j) o[i * C + j] = g[i][j];
#pragma unroll
    for (int j = 0; j < C; ++j) o[D * C + j] = phi[j];
    }
}

// ------------------------------------------------------------------------------------
// Facet flux (canonical slot order) -- identical for both neighbours of a facet
// ------------------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void facet_flux(const FvmPhys &p, float mu0, float mub0, float kappa,
                                           const float (&U0)[D + 2], const float (&U1)[D + 2],
                                           const float (&gl0)[D][D + 1], const float (&gl1)[D][D + 1],
                                           const float (&gdn)[D + 1], const float (&geo)[3 * D + 2],
                                           float (&F)[D + 2]) {
    // Multiply-adds are written as explicit fused operations (fmaf): a fused product-sum
    // is rounded once instead of twice, a difference from the reference's separately
    // rounded ATen operations at fp32 round-off, and ~190 fewer VALU operations per
    // facet.  Explicit (not compiler-contracted) so every kernel that inlines the flux
    // rounds identically.
    float n[D], nh[D], X[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        n[i] = geo[i];
        nh[i] = geo[D + i];
        X[i] = geo[2 * D + 2 + i];
    }
    const float fs = geo[2 * D];
    // Face gradient: mean of both sides' limited gradients, normal part replaced by the
    // two-point difference with non-orthogonal correction (facet_process.py:362-385).
    float gls[D][D + 1];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j <= D; ++j) gls[i][j] = (gl0[i][j] + gl1[i][j]) * 0.5f;
    float dFdn[D + 1], gdotn[D + 1];
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        float sx = gls[0][j] * X[0], sn = gls[0][j] * nh[0];
#pragma unroll
        for (int i = 1; i < D; ++i) {
            sx = fmaf(gls[i][j], X[i], sx);
            sn = fmaf(gls[i][j], nh[i], sn);
        }
        dFdn[j] = gdn[j] - sx;
        gdotn[j] = sn;
    }
    float gV[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) gV[i][j] = fmaf(dFdn[j] - gdotn[j], nh[i], gls[i][j]);
    const float gTn = dFdn[D];

    // Conserved face values and normal velocities (facet_process.py:353-359)
    float mom0[D], mom1[D], sq0 = 0.0f, sq1 = 0.0f, phi0 = 0.0f, phi1 = 0.0f;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        mom0[i] = U0[i] * U0[D];
        mom1[i] = U1[i] * U1[D];
        sq0 = (i == 0) ? U0[i] * U0[i] : fmaf(U0[i], U0[i], sq0);
        sq1 = (i == 0) ? U1[i] * U1[i] : fmaf(U1[i], U1[i], sq1);
        phi0 = (i == 0) ? U0[i] * n[i] : fmaf(U0[i], n[i], phi0);
        phi1 = (i == 0) ? U1[i] * n[i] : fmaf(U1[i], n[i], phi1);
    }
    const float r0 = U0[D], r1 = U1[D], T0 = U0[D + 1], T1 = U1[D + 1];
    const float Q0 = r0 * fmaf(p.cv, T0, 0.5f * sq0);
    const float Q1 = r1 * fmaf(p.cv, T1, 0.5f * sq1);

    // Stress tensor (FluidConstitution{2D,3D}._tau; 3-D divergence uses D00 + D11 only)
    const float Tm = (T0 + T1) * 0.5f;
    float Dm_[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Dm_[i][j] = 0.5f * (gV[i][j] + gV[j][i]);
    const float I1 = Dm_[0][0] + Dm_[1][1];
    // Sutherland: (T/T0)^1.5 as x sqrt(x) (a general powf costs ~50 operations)
    const float x = quot(Tm, p.T0);
    const float mu = quot((mu0 * (x * sqrt_hw(x))) * p.T0pS, Tm + p.S);
    const float mub = quot(mub0 * (Tm * Tm), p.T0sq);
    const float a0 = (-mub) * I1;
    const float a1 = -2.0f * mu;
    float tau[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) tau[i][j] = i == j ? fmaf(a1, Dm_[i][j], a0) : a1 * Dm_[i][j];

    const float P0 = (p.R * r0) * T0, P1 = (p.R * r1) * T1;
    const float c0 = sqrt_hw(p.gR * T0), c1 = sqrt_hw(p.gR * T1);

    // Advection (fvm_equation.py:210-231) and pressure (:309-321)
    const float Pm = (P0 + P1) * 0.5f;
#pragma unroll
    for (int i = 0; i < D; ++i) F[i] = fmaf(Pm, n[i], fmaf(mom0[i], phi0, mom1[i] * phi1) * 0.5f);
    F[D] = fmaf(r0, phi0, r1 * phi1) * 0.5f;
    F[D + 1] = fmaf(Q0 + P0, phi0, (Q1 + P1) * phi1) * 0.5f;
    // Viscosity (:246-257)
#pragma unroll
    for (int j = 0; j < D; ++j) {
        float v = tau[0][j] * n[0];
#pragma unroll
        for (int i = 1; i < D; ++i) v = fmaf(tau[i][j], n[i], v);
        F[j] = F[j] + v;
    }
    // Viscous heating and conduction (:272-294)
    float Vb[D];
#pragma unroll
    for (int j = 0; j < D; ++j) Vb[j] = (U0[j] + U1[j]) * 0.5f;
    float heat = 0.0f;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const float tv = tau[i][j] * Vb[j];
            heat = (i == 0 && j == 0) ? tv * n[i] : fmaf(tv, n[i], heat);
        }
    heat = fmaf((-kappa) * gTn, fs, heat);
    F[D + 1] = F[D + 1] + heat;
    // Kurganov-Tadmor dissipation (:337-368)
    const float Vn0 = sqrt_hw(sq0), Vn1 = sqrt_hw(sq1);
    const float Vmax = fmaxf(Vn0, Vn1);
    const float cm = fmaxf(c0, c1);
    const float M = fabsf(quot(Vmax, cm + 1e-8f));
    const float vf = fminf(fmaxf(M, p.v_factor), 1.0f);
    const float a_vel = fmaf(vf, cm, Vmax);
    const float a_oth = cm + Vmax;
    const float hv = 0.5f * a_vel, ho = 0.5f * a_oth;
#pragma unroll
    for (int i = 0; i < D; ++i) F[i] = fmaf(hv * (mom0[i] - mom1[i]), fs, F[i]);
    F[D] = fmaf(ho * (r0 - r1), fs, F[D]);
    F[D + 1] = fmaf(ho * (Q0 - Q1), fs, F[D + 1]);
}

// Limited face value and limited gradient (grad components V..., T) of one cell side.
template <int D>
__device__ __forceinline__ void side_values(const float (&P)[D + 2], const float *__restrict__ g,
                                            const float (&disp)[D], float (&U)[D + 2], float (&gl)[D][D + 1]) {
    constexpr int C = D + 2;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const float ph = g[D * C + j];
        float dU = g[j] * disp[0];
#pragma unroll
        for (int i = 1; i < D; ++i) dU = fmaf(g[i * C + j], disp[i], dU);
        U[j] = fmaf(dU, ph, P[j]);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int j = 0; j < D; ++j) gl[i][j] = g[i * C + j] * g[D * C + j];
        gl[i][D] = g[i * C + D + 1] * g[D * C + D + 1];
    }
}

// Per facet (device facet order): both sides' limited face values and gradients, the
// two-point normal gradient, and the complete facet flux -> fflux[f][C].  Each facet is
// evaluated once; the divergence kernel adds it to one cell and subtracts it from the
// other, so the two contributions cancel exactly.
template <int D>
__global__ void __launch_bounds__(kThreads)
k_facet_flux(FvmMeshDev m, FvmPhys p, FvmWork w, const float *__restrict__ prim) {
    const float *__restrict__ gp = w.gp;
    using Dm = Dims<D>;
    constexpr int C = Dm::C;
    const int64_t N = m.n_cells, NF = m.n_facets;
    const int64_t ntiles = (NF + kThreads - 1) / kThreads;
    for (int64_t v = blockIdx.x; v < ntiles; v += gridDim.x) {
        const int64_t f = xcd_tile(v, ntiles) * kThreads + threadIdx.x;
        if (f >= NF) continue;
        const int64_t c0 = m.fac_cell[f], c1 = m.fac_cell[NF + f];
        const float *geo = m.fac_geo + f * Dm::GEO;
        const bool outer0 = c0 >= N;                 // slot 0 is a boundary facet's outer side
        const int64_t own = outer0 ? c1 : c0;
        float mu0 = p.mu0, mub0 = p.mub0, kappa = p.kappa;
        if (m.mesh_phys) {
            const int mesh = m.cell_mesh[own];
            mu0 = m.mesh_phys[4 * mesh];
            mub0 = m.mesh_phys[4 * mesh + 1];
            kappa = m.mesh_phys[4 * mesh + 2];
        }
        float d0[D], d1[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            d0[i] = geo[3 * D + 2 + i];
            d1[i] = geo[4 * D + 2 + i];
        }
        // both sides' records are gathered unconditionally (a boundary facet gathers its
        // interior cell twice) so the loads of a facet overlap
        const int64_t ca = outer0 ? c1 : c0, cb = c1 < N ? c1 : c0;
        float P0[C], P1[C], ga[Dm::REC], gb[Dm::REC];
        load_prim<D>(prim + ca * C, P0);
        load_prim<D>(prim + cb * C, P1);
#pragma unroll
        for (int q = 0; q < Dm::REC; q += C) {
            float t0[C], t1[C];
            load_vec<C>(gp + ca * Dm::REC + q, t0);
            load_vec<C>(gp + cb * Dm::REC + q, t1);
#pragma unroll
            for (int j = 0; j < C; ++j) {
                ga[q + j] = t0[j];
                gb[q + j] = t1[j];
            }
        }
        float U0s[C], U1s[C], g0s[D][D + 1], g1s[D][D + 1], gdn[D + 1];
        if (c0 < N && c1 < N) {
            side_values<D>(P0, ga, d0, U0s, g0s);
            side_values<D>(P1, gb, d1, U1s, g1s);
            // A_face_grad row: (-1/d, +1/d) on (slot-0, slot-1) cells, ascending columns
            const float inv_d = geo[2 * D + 1];
            const bool left_first = c0 < c1;
#pragma unroll
            for (int jg = 0; jg <= D; ++jg) {
                const int j = jg < D ? jg : D + 1;
                const float tl = (-inv_d) * P0[j], tr = inv_d * P1[j];
                gdn[jg] = (left_first ? tl + tr : tr + tl) + 0.0f;
            }
        } else {
            // boundary facet: the outer side takes the boundary values and the interior
            // cell's limited gradient; the normal gradient is A_face_grad.x + b_face_grad
            const int64_t b = (outer0 ? c0 : c1) - N;
            float Uo[C], glo[D][D + 1], Ux[C];
            side_values<D>(P0, ga, outer0 ? d1 : d0, Uo, glo);
            load_prim<D>(w.ubc + b * C, Ux);
#pragma unroll
            for (int j = 0; j < C; ++j) {
                U0s[j] = outer0 ? Ux[j] : Uo[j];
                U1s[j] = outer0 ? Uo[j] : Ux[j];
            }
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int j = 0; j <= D; ++j) {
                    g0s[i][j] = glo[i][j];
                    g1s[i][j] = glo[i][j];
                }
#pragma unroll
            for (int jg = 0; jg <= D; ++jg) {
                const int j = jg < D ? jg : D + 1;
                gdn[jg] = m.bc_ga[b * C + j] * P0[j] + m.bc_gb[b * C + j];
            }
        }
        float gr[3 * D + 2], gn[D], gx[D - 1];
#pragma unroll
        for (int i = 0; i < D; ++i) gn[i] = geo[i];
#pragma unroll
        for (int i = 0; i < D - 1; ++i) gx[i] = geo[2 * D + 2 + i];
        facet_geo_full<D>(gn, geo[2 * D + 1], gx, gr);
        float Fx[C];
        facet_flux<D>(p, mu0, mub0, kappa, U0s, U1s, g0s, g1s, gdn, gr, Fx);
        float *out = w.fflux + f * C;
        if constexpr (C == 4) {
            *reinterpret_cast<float4 *>(out) = make_float4(Fx[0], Fx[1], Fx[2], Fx[3]);
        } else {
#pragma unroll
            for (int j = 0; j < C; ++j) out[j] = Fx[j];
        }
    }
}

enum { EPI_DUDT = 0, EPI_STAGE = 1 };

// The stage epilogue's per-cell inputs: U0 and the RK accumulators, or (Adams-Bashforth)
// the two previous derivatives.  Issued before the divergence so that they are in
// flight together with the flux gathers.
template <int D, int EPI>
__device__ __forceinline__ void stage_inputs(const FvmStage &st, int64_t c, bool live, float (&U0v)[D + 2],
                                             float (&ap)[D + 2], float (&a2p)[D + 2]) {
    constexpr int C = D + 2;
    if (EPI == EPI_STAGE && live) {
        const int64_t o = c * C;
        load_vec<C>(st.U0 + o, U0v);
        if (st.mode == 1) {
            if (st.n_hist >= 1) load_vec<C>(st.f_m1 + o, ap);
            if (st.n_hist >= 2) load_vec<C>(st.f_m2 + o, a2p);
        } else {
            if (st.stage != 0) load_vec<C>(st.acc + o, ap);
            if (st.stage != 0 && st.acc2 != nullptr) load_vec<C>(st.acc2 + o, a2p);
        }
    }
}

// Epilogue of cell c with dU/dt = acc.  Block-uniform call (every thread of the block,
// live or not): writes dU/dt, or applies the explicit stage update and emits the next
// stage's primitives, their boundary-mean partials (slot of 256-cell tile `tile`) and,
// on the last stage of an embedded tableau, the fp64 column-norm partials of `tile`.
// dtm: the time step of the cell's mesh (read by the caller; unused for dU/dt).
template <int D, int EPI>
__device__ __forceinline__ void stage_epilogue(const FvmMeshDev &m, const FvmPhys &p, const FvmWork &w,
                                               float *__restrict__ dudt, const FvmStage &st, int64_t tile,
                                               int64_t c, bool live, float dtm, const float (&acc)[D + 2],
                                               const float (&U0v)[D + 2], const float (&ap)[D + 2],
                                               const float (&a2p)[D + 2], float *red, double *sred) {
    constexpr int C = D + 2;
    const int64_t N = m.n_cells;
    if (EPI == EPI_DUDT) {
        if (live) store_vec<C>(dudt + c * C, acc);
        return;
    }
    const bool first = st.stage == 0;
    const bool has2 = st.acc2 != nullptr;
    // Stage flags are uniform; every per-cell load is issued before the first use so
    // that a cell's memory round trips overlap.  nrm: the fp32 vectors whose squares the
    // last stage of an embedded tableau sums in fp64 (acc - acc2, acc, U0); set and read
    // only on that path.
    float nrm[3][C];
    float Pn[C];
#pragma unroll
    for (int j = 0; j < C; ++j) Pn[j] = 0.0f;
    if (live && st.mode == 1) {
        // Adams-Bashforth 3 (fixed dt, one RHS per step); ap / a2p hold f_m1 / f_m2
        const int64_t o = c * C;
        const float(&fm1)[C] = ap;
        const float(&fm2)[C] = a2p;
        const float c12 = dtm / 12.0f;
        float U1[C];
#pragma unroll
        for (int j = 0; j < C; ++j) {
            const float f0 = acc[j];
            const float g1 = st.n_hist >= 1 ? fm1[j] : f0;
            const float g2 = st.n_hist >= 2 ? fm2[j] : g1;
            const float comb = (23.0f * f0 - 16.0f * g1) + 5.0f * g2;
            U1[j] = U0v[j] + c12 * comb;
        }
        store_vec<C>(st.f_out + o, acc);
        store_vec<C>(st.U_out + o, U1);
        state_to_prim<D>(U1, p.inv_cv, Pn);
        store_vec<C>(st.prim_next + o, Pn);
    } else if (live) {
        const int64_t o = c * C;
        const float dt = dtm;
        float kk[C], a[C], a2[C];
#pragma unroll
        for (int j = 0; j < C; ++j) kk[j] = dt * acc[j];
        if (first) {
#pragma unroll
            for (int j = 0; j < C; ++j) a[j] = st.b * kk[j];
        } else {
#pragma unroll
            for (int j = 0; j < C; ++j) a[j] = ap[j] + st.b * kk[j];
        }
        if (has2) {
#pragma unroll
            for (int j = 0; j < C; ++j) a2[j] = first ? st.b2 * kk[j] : a2p[j] + st.b2 * kk[j];
        }
        if (st.store_k) store_vec<C>(st.kbuf + static_cast<int64_t>(st.stage) * N * C + o, kk);
        float Un[C];
        if (!st.last) {
            float inc[C];
#pragma unroll
            for (int j = 0; j < C; ++j) inc[j] = 0.0f;
            for (int t = 0; t < st.n_inc; ++t) {
                const int sidx = st.inc_idx[t];
                const float cf = st.inc_coef[t];
                float kv[C];
                if (sidx == st.stage) {
#pragma unroll
                    for (int j = 0; j < C; ++j) kv[j] = kk[j];
                } else {
                    load_vec<C>(st.kbuf + static_cast<int64_t>(sidx) * N * C + o, kv);
                }
#pragma unroll
                for (int j = 0; j < C; ++j) {
                    const float term = cf * kv[j];
                    inc[j] = t == 0 ? term : inc[j] + term;
                }
            }
#pragma unroll
            for (int j = 0; j < C; ++j) Un[j] = U0v[j] + inc[j];
            store_vec<C>(st.acc + o, a);
            if (has2) store_vec<C>(st.acc2 + o, a2);
        } else {
#pragma unroll
            for (int j = 0; j < C; ++j) Un[j] = U0v[j] + a[j];
            store_vec<C>(st.U_out + o, Un);
            if (has2) {
#pragma unroll
                for (int j = 0; j < C; ++j) {
                    nrm[0][j] = a[j] - a2[j];
                    nrm[1][j] = a[j];
                    nrm[2][j] = U0v[j];
                }
            }
        }
        state_to_prim<D>(Un, p.inv_cv, Pn);
        store_vec<C>(st.prim_next + o, Pn);
    }
    // boundary means of the next stage's primitives (block-uniform branch)
    if (m.n_groups > 0) {
        const int slot = m.blk_slot[tile];
        if (slot >= 0) bc_mean_block<D>(m, p, w.bc_part_next, c, live, Pn[D + 1], slot, red);
    }
    if (st.last && st.acc2) {
        // fp64 block partials of the three column norms (Adaptive.update_stepsize)
        double sq[3 * C];
#pragma unroll
        for (int q = 0; q < 3 * C; ++q) {
            const double x = live ? static_cast<double>(nrm[q / C][q % C]) : 0.0;
            sq[q] = x * x;
        }
#pragma unroll
        for (int q = 0; q < 3 * C; ++q) {
            sred[threadIdx.x] = sq[q];
            __syncthreads();
            for (int h = kThreads / 2; h > 0; h >>= 1) {
                if (threadIdx.x < h) sred[threadIdx.x] += sred[threadIdx.x + h];
                __syncthreads();
            }
            if (threadIdx.x == 0) w.red[tile * 3 * C + q] = sred[0];
            __syncthreads();
        }
    }
}

// Per cell: the divergence (flux_mat row: +-1/vol on its facets in ascending reference
// facet order) of the facet fluxes, then the epilogue -- dU/dt, or an explicit stage
// update with the next stage's primitives and their boundary means.
template <int D, int EPI>
__global__ void __launch_bounds__(kThreads)
k_div(FvmMeshDev m, FvmPhys p, FvmWork w, float *__restrict__ dudt, FvmStage st) {
    using Dm = Dims<D>;
    constexpr int C = Dm::C, NL = Dm::NL;
    const int64_t N = m.n_cells;
    const int64_t ntiles = (N + kThreads - 1) / kThreads;
    __shared__ float red[kThreads];
    __shared__ double sred[kThreads];
    for (int64_t v = blockIdx.x; v < ntiles; v += gridDim.x) {
        const int64_t tile = xcd_tile(v, ntiles);
        const int64_t c = tile * kThreads + threadIdx.x;
        const bool live = c < N;
        const float dtm = (EPI == EPI_STAGE && live) ? st.dt[m.cell_mesh ? m.cell_mesh[c] : 0] : 0.0f;
        float U0v[C], ap[C], a2p[C];
        stage_inputs<D, EPI>(st, c, live, U0v, ap, a2p);
        float acc[C];
#pragma unroll
        for (int j = 0; j < C; ++j) acc[j] = 0.0f;
        if (live) {
            const int info = m.cell_info[c];
            const float inv_vol = m.cell_inv_vol[c];
            float Fk[NL][C];
#pragma unroll
            for (int k = 0; k < NL; ++k)
                load_vec<C>(w.fflux + static_cast<int64_t>(m.cell_facet[k * N + c]) * C, Fk[k]);
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const float vk = ((info >> k) & 1) ? inv_vol : -inv_vol;
#pragma unroll
                for (int j = 0; j < C; ++j) acc[j] = acc[j] + vk * Fk[k][j];
            }
        }
        stage_epilogue<D, EPI>(m, p, w, dudt, st, tile, c, live, dtm, acc, U0v, ap, a2p, red, sred);
    }
}

// ------------------------------------------------------------------------------------
// Tile-fused stage: one 256-cell tile per workgroup (2-D and 3-D)
// ------------------------------------------------------------------------------------
// The workgroup stages in LDS the primitives of its tile, of the tile's one-ring halo H1
// (cells outside the tile sharing a facet with it) and of H1's other neighbours H2;
// computes the boundary values, gradient and limiter of tile + H1 cells, keeping their
// limited gradients and their face values in LDS; evaluates the fluxes of the tile's
// owned and foreign facets; then every tile cell sums its facets' fluxes and runs the
// stage epilogue.  The arithmetic is that of k_grad, k_facet_flux and k_div, through the
// same device functions in the same order, so the results are bitwise those of the
// three-kernel pipeline; only the traffic differs (gradient records and facet fluxes
// never leave the CU).  Gradient records are written back only for cells with a
// boundary facet (the next call's Neumann correction reads them) into w.gp, which must
// not alias w.gp_prev: other tiles still read gp_prev for their halo cells.
//
// Face values are kept per facet of the tile's list and side, FF[q][side][C]: each is
// written once, by the cell on that side (a boundary facet's outer value by its interior
// cell), and the facet's flux overwrites FF[q][0] in place (the thread evaluating facet
// q is the only one reading FF[q] in that phase).  Per cell slot storage would hold the
// face values of H1 cells' facets outside the tile too, and a separate flux array
// would add nfl C floats: this layout needs 29.5 KB on the 1M-cell 2-D mesh (5
// workgroups per CU) and ~67 KB for 256 tetrahedra.
constexpr int kFzTab = 12;
constexpr unsigned kFzNone = 0xFFFFu;
constexpr unsigned kFzSide1 = 0x8000u;   // H1 facet table: the H1 cell is the slot-1 side

// compact facet record of the fused stage: n [D], 1/(n_hat.d), X_orthog's tangent
// coordinates [D - 1]
template <int D>
__host__ __device__ constexpr int fz_geo_n() { return 2 * D; }
// H1 static record: LSQ values of the neighbour columns [NL][D], displacements of the
// first NL - 1 slots [NL-1][D], eps; stored component-major, [HN][sum n1] (a lane per H1
// cell reads one coalesced row per component)
template <int D>
__host__ __device__ constexpr int fz_hrec_n() { return (D + 1) * D + D * D + 1; }

// A cell's last slot displacement: minus the fp32 sum of the others (the facet midpoints
// average to the centroid; the mesh layout stores exactly this value, so the pipelines
// that load it see the same floats).
template <int D>
__device__ __forceinline__ void last_disp(float (&disp)[D + 1][D]) {
#pragma unroll
    for (int ii = 0; ii < D; ++ii) {
        float acc = disp[0][ii];
#pragma unroll
        for (int k = 1; k < D; ++k) acc = acc + disp[k][ii];
        disp[D][ii] = -acc;
    }
}
template <int D>
__host__ __device__ constexpr int fz_fpt() { return D == 2 ? 2 : 3; }   // facets per thread

__host__ __device__ inline int64_t fz_r4(int64_t x) { return (x + 3) & ~int64_t(3); }

// LDS layout of k_fused in floats, each block 16-byte aligned: Pl [ncl][C], GL
// [ngl][D (D+1)] (a cell's limited V and T gradients in one row: the flux phase reads a
// side's gradients with two or three vector loads from one address), FF [nfl][2][C].
template <int D>
__host__ __device__ inline int64_t fz_gl_offset(int ncl) {
    return fz_r4(static_cast<int64_t>(ncl) * (D + 2));
}
template <int D>
__host__ __device__ inline int64_t fz_ff_offset(int ncl, int ngl) {
    return fz_gl_offset<D>(ncl) + fz_r4(static_cast<int64_t>(ngl) * D * (D + 1));
}

// A cell's row of GL: its limited gradients gl[i][j] (j < D: velocity components, j = D:
// temperature) at gl_row[i (D+1) + j]; 8-byte aligned in 2-D (6 floats), 16-byte in 3-D.
template <int D>
__device__ __forceinline__ void gl_store(float *__restrict__ row, const float (&v)[D * (D + 1)]) {
    if constexpr (D == 2) {
        float2 *r = reinterpret_cast<float2 *>(row);
        r[0] = make_float2(v[0], v[1]);
        r[1] = make_float2(v[2], v[3]);
        r[2] = make_float2(v[4], v[5]);
    } else {
        float4 *r = reinterpret_cast<float4 *>(row);
        r[0] = make_float4(v[0], v[1], v[2], v[3]);
        r[1] = make_float4(v[4], v[5], v[6], v[7]);
        r[2] = make_float4(v[8], v[9], v[10], v[11]);
    }
}

template <int D>
__device__ __forceinline__ void gl_load(const float *__restrict__ row, float (&g)[D][D + 1]) {
    if constexpr (D == 2) {
        const float2 *r = reinterpret_cast<const float2 *>(row);
        const float2 a = r[0], b = r[1], c = r[2];
        g[0][0] = a.x; g[0][1] = a.y; g[0][2] = b.x;
        g[1][0] = b.y; g[1][1] = c.x; g[1][2] = c.y;
    } else {
        const float4 *r = reinterpret_cast<const float4 *>(row);
        const float4 a = r[0], b = r[1], c = r[2];
        const float t[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j <= D; ++j) g[i][j] = t[i * (D + 1) + j];
    }
}

// One gradient cell of the fused kernel (tile cell or H1 cell, local id l, global gc):
// boundary values, gradient and limiter; the limited V, T gradients go to GL, the face
// value at each facet of the tile's list to FF[q][side] (fq[k] = q | side << 15, or
// kFzNone for a facet outside the list).  Tile cells also write their boundary values
// (to w.ubc and, as the facet's outer value, to FF) and, when they have a boundary facet,
// their gradient record (the next call's Neumann correction).
template <int D, int LP>
__device__ __forceinline__ void fused_grad_cell(const FvmMeshDev &m, const FvmPhys &p, const FvmWork &w,
                                                const float *Pl, float *GL, float *FF, int ngl, int l, int64_t gc,
                                                bool own, const float (&gv)[D + 2][D], const float (&disp)[D + 1][D],
                                                float eps, const ushort4 nrow, int has_prev, const float *gmean,
                                                int g0, const ushort4 fq4) {
    using Dm = Dims<D>;
    constexpr int C = Dm::C, NL = Dm::NL;
    const int64_t N = m.n_cells;
    const unsigned nl[4] = {nrow.x, nrow.y, nrow.z, nrow.w};
    const unsigned fq[4] = {fq4.x, fq4.y, fq4.z, fq4.w};
    float P[C];
    load_vec<C>(Pl + l * C, P);
    float Un[NL][C];
#pragma unroll
    for (int k = 0; k < NL; ++k) load_vec<C>(Pl + (nl[k] != kFzNone ? nl[k] : l) * C, Un[k]);
    bool bnd = false;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        if (nl[k] == kFzNone) {
            bnd = true;
            const int64_t b = static_cast<int64_t>(m.cell_nbr[k * N + gc]) - N;
            bc_value<D>(m, p, b, P, w.gp_prev + gc * Dm::REC, has_prev, gmean, g0, Un[k]);
            if (own) {
                store_vec<C>(w.ubc + b * C, Un[k]);
                // the boundary facet's outer value: the side opposite this cell's
                const unsigned q = fq[k] & ~kFzSide1, side = (fq[k] & kFzSide1) ? 1u : 0u;
                store_vec<C>(FF + (q * 2 + (1u - side)) * C, Un[k]);
            }
        }
    }
    float g[D][C], phi[C], Uf[NL][C];
    cell_gradient<D, LP, true>(P, Un, gv, disp, eps, g, phi, Uf);
    if (own && bnd) {
        float *o = w.gp + gc * Dm::REC;
#pragma unroll
        for (int ii = 0; ii < D; ++ii)
#pragma unroll
            for (int j = 0; j < C; ++j) o[ii * C + j] = g[ii][j];
#pragma unroll
        for (int j = 0; j < C; ++j) o[D * C + j] = phi[j];
    }
    // limited face values (side_values' operation order) at the facets in the list
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        if (fq[k] == kFzNone) continue;
        const unsigned q = fq[k] & ~kFzSide1, side = (fq[k] & kFzSide1) ? 1u : 0u;
        store_vec<C>(FF + (q * 2 + side) * C, Uf[k]);
    }
    float glv[D * (D + 1)];
#pragma unroll
    for (int ii = 0; ii < D; ++ii) {
#pragma unroll
        for (int j = 0; j < D; ++j) glv[ii * (D + 1) + j] = g[ii][j] * phi[j];
        glv[ii * (D + 1) + D] = g[ii][D + 1] * phi[D + 1];
    }
    gl_store<D>(GL + l * (D * (D + 1)), glv);
}

// Facet q of the tile's list (descriptor d, compact geometry geo): both sides' face
// values from FF[q], limited gradients from GL, the normal gradient, and the flux into
// FF[q][0].
template <int D>
__device__ __forceinline__ void fused_facet(const FvmMeshDev &m, const FvmPhys &p, const float *Pl, const float *GL,
                                            int ngl, float *FF, int q, int64_t f, unsigned d,
                                            const float (&gc)[2 * D], const unsigned short *meshT, int mt0,
                                            const float *physT) {
    using Dm = Dims<D>;
    constexpr int C = Dm::C;
    const int64_t N = m.n_cells, NF = m.n_facets;
    const unsigned l0 = d & 0xFFFu, l1 = (d >> 14) & 0xFFFu;
    const bool left_first = (d >> 28) & 1u, bnd = (d >> 29) & 1u, outer0 = (d >> 30) & 1u;
    float mu0 = p.mu0, mub0 = p.mub0, kappa = p.kappa;
    if (m.mesh_phys) {
        // one side of every facet of the tile's list is a tile cell (the interior cell of
        // a boundary facet; the lowest cell of an owned facet; the higher one of a foreign
        // facet), whose mesh id the staging phase left in LDS, with the physics of the
        // tile's first kTileMeshes meshes (mt0, mt0 + 1, ...)
        const unsigned lt = bnd ? (outer0 ? l1 : l0) : (l0 < kThreads ? l0 : l1);
        const int mesh = meshT[lt];
        const int rel = mesh - mt0;
        if (rel >= 0 && rel < kTileMeshes) {
            mu0 = physT[4 * rel];
            mub0 = physT[4 * rel + 1];
            kappa = physT[4 * rel + 2];
        } else {
            mu0 = m.mesh_phys[4 * mesh];
            mub0 = m.mesh_phys[4 * mesh + 1];
            kappa = m.mesh_phys[4 * mesh + 2];
        }
    }
    float geo[3 * D + 2];
    {
        float n[D], xt[D - 1];
#pragma unroll
        for (int i = 0; i < D; ++i) n[i] = gc[i];
#pragma unroll
        for (int i = 0; i < D - 1; ++i) xt[i] = gc[D + 1 + i];
        facet_geo_full<D>(n, gc[D], xt, geo);
    }
    float U0s[C], U1s[C], g0s[D][D + 1], g1s[D][D + 1], gdn[D + 1];
    load_vec<C>(FF + (2 * q) * C, U0s);
    load_vec<C>(FF + (2 * q + 1) * C, U1s);
    if (!bnd) {
        float P0[C], P1[C];
        load_vec<C>(Pl + l0 * C, P0);
        load_vec<C>(Pl + l1 * C, P1);
        gl_load<D>(GL + l0 * (D * (D + 1)), g0s);
        gl_load<D>(GL + l1 * (D * (D + 1)), g1s);
        // A_face_grad row: (-1/d, +1/d) on (slot-0, slot-1) cells, ascending columns
        const float inv_d = geo[2 * D + 1];
#pragma unroll
        for (int jg = 0; jg <= D; ++jg) {
            const int j = jg < D ? jg : D + 1;
            const float tl = (-inv_d) * P0[j], tr = inv_d * P1[j];
            gdn[jg] = (left_first ? tl + tr : tr + tl) + 0.0f;
        }
    } else {
        // boundary facet: its interior cell is a tile cell (the facet's lowest cell); the
        // outer side's value was computed by that cell in the gradient phase (the same
        // bc_value on the same inputs as the three-kernel pipeline)
        const unsigned li = outer0 ? l1 : l0;
        float Pi[C];
        load_vec<C>(Pl + li * C, Pi);
        const int64_t b = static_cast<int64_t>(m.fac_cell[outer0 ? f : NF + f]) - N;
        gl_load<D>(GL + li * (D * (D + 1)), g0s);
#pragma unroll
        for (int ii = 0; ii < D; ++ii)
#pragma unroll
            for (int j = 0; j <= D; ++j) g1s[ii][j] = g0s[ii][j];
#pragma unroll
        for (int jg = 0; jg <= D; ++jg) {
            const int j = jg < D ? jg : D + 1;
            gdn[jg] = m.bc_ga[b * C + j] * Pi[j] + m.bc_gb[b * C + j];
        }
    }
    float Fx[C];
    facet_flux<D>(p, mu0, mub0, kappa, U0s, U1s, g0s, g1s, gdn, geo, Fx);
    store_vec<C>(FF + (2 * q) * C, Fx);
}

// Descriptor and compact geometry of facet q of the tile's list (owned run, then the
// foreign list).
template <int D>
__device__ __forceinline__ void fused_facet_load(const FvmMeshDev &m, int q, int fo0, int nfo, int foff, int64_t &f,
                                                 unsigned &d, float (&geo)[2 * D]) {
    constexpr int FG = fz_geo_n<D>();
    const bool owned = q < nfo;
    f = owned ? static_cast<int64_t>(fo0) + q : static_cast<int64_t>(m.fz_ffac[foff + q - nfo]);
    d = owned ? m.fz_fdesc[f] : m.fz_fdesc_x[foff + q - nfo];
    const float *src = owned ? m.fz_geo + f * FG : m.fz_geo_x + static_cast<int64_t>(foff + q - nfo) * FG;
    // rows of 2D floats: 16-byte aligned in 2-D (one load), 8-byte aligned in 3-D
    if constexpr (D == 2) {
        const float4 v = *reinterpret_cast<const float4 *>(src);
        geo[0] = v.x;
        geo[1] = v.y;
        geo[2] = v.z;
        geo[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < FG; k += 2) {
            const float2 v = *reinterpret_cast<const float2 *>(src + k);
            geo[k] = v.x;
            geo[k + 1] = v.y;
        }
    }
}

// One gradient cell's register inputs in the fused stage: LSQ row values, displacements,
// limiter epsilon, local neighbour row and face-value positions (q | side << 15 per slot).
template <int D>
struct FzGradIn {
    float gv[D + 2][D];
    float disp[D + 1][D];
    float eps;
    ushort4 nr, fq;
};

// Slot words of the fused stage (one 32-bit word per cell slot): bits 0-11 the local id
// of the neighbour behind the slot (0xFFF: boundary facet), 12-21 the facet's position in
// the tile's list (0x3FF: not in it), 22 the cell is the facet's slot-1 side, 23-24 the
// slot's position in the reference's local order of the cell's LSQ row.  Decoded into the
// neighbour row, the face-value positions (q | kFzSide1 on the slot-1 side, kFzNone when
// not in the list), and the order positions.
template <int D>
__device__ __forceinline__ void fz_decode_words(const unsigned (&wd)[D + 1], FzGradIn<D> &g,
                                                unsigned (&ref)[D + 1]) {
    constexpr int NL = D + 1;
    unsigned nl[4] = {kFzNone, kFzNone, kFzNone, kFzNone}, fq[4] = {kFzNone, kFzNone, kFzNone, kFzNone};
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const unsigned w = wd[k], l = w & 0xFFFu, q = (w >> 12) & 0x3FFu;
        nl[k] = l == 0xFFFu ? kFzNone : l;
        fq[k] = q == 0x3FFu ? kFzNone : (q | (((w >> 22) & 1u) ? kFzSide1 : 0u));
        ref[k] = (w >> 23) & 3u;
    }
    g.nr = make_ushort4(nl[0], nl[1], nl[2], nl[3]);
    g.fq = make_ushort4(fq[0], fq[1], fq[2], fq[3]);
}

// The cell's own LSQ column, which is not stored: minus the fp32 sum of the neighbour
// columns in the reference's local order, as build_sparse_gradient_matrix forms the
// G_mats diagonal (fvm_mesh.py:5-42) -- bitwise that stored value.
template <int D>
__device__ __forceinline__ void fz_own_column(FzGradIn<D> &g, const unsigned (&ref)[D + 1]) {
    constexpr int NL = D + 1;
#pragma unroll
    for (int ii = 0; ii < D; ++ii) {
        float acc = 0.0f;
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            float t = g.gv[0][ii];
#pragma unroll
            for (int k = 1; k < NL; ++k) t = ref[k] == static_cast<unsigned>(r) ? g.gv[k][ii] : t;
            acc = r == 0 ? t : acc + t;
        }
        g.gv[NL][ii] = -acc;
    }
}

// Inputs of tile cell c: LSQ values, displacements, eps and slot words; info gets the
// cell's slot bits (bit k: slot-1 side of facet k) and cf its facets' list positions.
template <int D>
__device__ __forceinline__ void fz_load_tile_cell(const FvmMeshDev &m, int64_t c, FzGradIn<D> &g, int &info,
                                                  ushort4 &cf) {
    constexpr int NL = D + 1;
    const int64_t N = m.n_cells;
#pragma unroll
    for (int jj = 0; jj < NL; ++jj)
#pragma unroll
        for (int ii = 0; ii < D; ++ii) g.gv[jj][ii] = m.cell_gval[(jj * D + ii) * N + c];
#pragma unroll
    for (int k = 0; k < NL - 1; ++k)
#pragma unroll
        for (int ii = 0; ii < D; ++ii) g.disp[k][ii] = m.cell_disp[(k * D + ii) * N + c];
    last_disp<D>(g.disp);
    g.eps = m.cell_eps[c];
    unsigned wd[NL], ref[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) wd[k] = m.fz_cw[k * N + c];
    fz_decode_words<D>(wd, g, ref);
    fz_own_column<D>(g, ref);
    unsigned q[4] = {0u, 0u, 0u, 0u};
    info = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        q[k] = (wd[k] >> 12) & 0x3FFu;
        info |= static_cast<int>((wd[k] >> 22) & 1u) << k;
    }
    cf = make_ushort4(q[0], q[1], q[2], q[3]);
}

// Inputs of the tile's H1 entry h (global id returned): its component-major static record
// and slot words.
template <int D>
__device__ __forceinline__ int64_t fz_load_h1_cell(const FvmMeshDev &m, int h, int halo_off, int h1_off,
                                                   FzGradIn<D> &g) {
    constexpr int NL = D + 1, HREC = fz_hrec_n<D>();
    const int64_t gc = m.fz_halo[halo_off + h];
    const float *r = m.fz_hrec + h1_off + h;
    const int64_t hs = m.fz_nh1;             // component stride of the H1 records and words
    float rv[HREC];
#pragma unroll
    for (int q = 0; q < HREC; ++q) rv[q] = r[q * hs];
    unsigned wd[NL], ref[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) wd[k] = m.fz_hw[k * hs + h1_off + h];
#pragma unroll
    for (int jj = 0; jj < NL; ++jj)
#pragma unroll
        for (int ii = 0; ii < D; ++ii) g.gv[jj][ii] = rv[jj * D + ii];
#pragma unroll
    for (int k = 0; k < NL - 1; ++k)
#pragma unroll
        for (int ii = 0; ii < D; ++ii) g.disp[k][ii] = rv[NL * D + k * D + ii];
    last_disp<D>(g.disp);
    g.eps = rv[NL * D + (NL - 1) * D];
    fz_decode_words<D>(wd, g, ref);
    fz_own_column<D>(g, ref);
    return gc;
}

// Workgroup shapes (NT threads, one 256-cell tile either way):
//   NT = kThreads      thread i owns tile cell i, H1 entry i (while i < n1) and facets
//                      i, i + 256, ... of the tile's list (up to fz_fpt<D>());
//   NT = 320 .. 512    threads 0..255 own the tile cells, threads 256.. the H1 entries
//                      (local id = thread id; the host guarantees n1 <= NT - 256), and
//                      facets i, i + NT, ...: more waves per tile for the same LDS, each
//                      doing one gradient and fewer facets.
// Both compute the same numbers (each cell, facet and reduction is evaluated by the same
// code in the same order).  Each phase's global loads are issued before the previous
// phase's work (they land in registers while it computes), so a tile costs about one
// memory round trip per phase boundary rather than one per dependent load.  The host
// guarantees at most 256 H1 cells and fz_fpt<D>() x 256 facets per tile.
template <int D, int LP, int EPI, int NT>
__global__ void __launch_bounds__(NT)
k_fused(FvmMeshDev m, FvmPhys p, FvmWork w, const float *__restrict__ prim, float *__restrict__ dudt, FvmStage st,
        int has_prev) {
    using Dm = Dims<D>;
    constexpr int C = Dm::C, NL = Dm::NL;
    constexpr bool SPLIT = NT > kThreads;
    constexpr int FS = (fz_fpt<D>() * kThreads + NT - 1) / NT;   // facet slots per thread
    static_assert(NT % 64 == 0 && NT >= kThreads && NT <= 2 * kThreads, "fused stage: 256 to 512 threads");
    static_assert(FS >= 1 && FS <= 3, "fused stage: 1 to 3 facets per thread");
    extern __shared__ float4 fz_lds[];     // 16-byte aligned: 2-D records move as ds_*_b128
    float *lds = reinterpret_cast<float *>(fz_lds);
    const int64_t N = m.n_cells;
    const int64_t ntiles = (N + kThreads - 1) / kThreads;
    int64_t tile = xcd_tile(blockIdx.x, ntiles);
    if constexpr (EPI == EPI_STAGE) {
        if (st.tile_list) {                      // grid = the caller's tile list
            tile = st.tile_list[xcd_tile(blockIdx.x, st.n_tile_list)];
            if (tile < 0 || tile >= ntiles) return;
        }
    }
    const int *tb = m.fz_tile + tile * kFzTab;
    const int halo_off = tb[0], n1 = tb[1], n2 = tb[2], fo0 = tb[4], nfo = tb[5];
    const int foff = tb[6], nfor = tb[7], g0 = tb[8], ng = tb[9], h1_off = tb[10];
    float *Pl = lds;
    float *GL = lds + fz_gl_offset<D>(m.fz_ncl);
    float *FF = lds + fz_ff_offset<D>(m.fz_ncl, m.fz_ngl);
    // batches only (after FF): the mesh id of each tile cell, then mu0, mub0, kappa and dt
    // of the tile's first kTileMeshes meshes (cells are ordered by mesh)
    unsigned short *meshT = reinterpret_cast<unsigned short *>(FF + static_cast<int64_t>(m.fz_nfl) * 2 * C);
    float *physT = reinterpret_cast<float *>(meshT + kThreads);
    float *red = GL;                                   // reductions reuse the gradient rows
    double *sred = reinterpret_cast<double *>(GL);     // (before and after the phases using them)
    const int i = threadIdx.x;
    const int h = SPLIT ? i - kThreads : i;            // this thread's H1 entry
    const int64_t c0 = tile * kThreads;
    const int nown = static_cast<int>(N - c0 < kThreads ? N - c0 : kThreads);
    const int nft = nfo + nfor;
    const bool has_own = i < nown, has_h1 = h >= 0 && h < n1;
    __shared__ float gmean[kMaxGroups];
    if (ng > 0) block_group_means_g<NT>(m, w, g0, ng, gmean, red);
    int mt0 = 0;
    if (m.mesh_phys) {
        mt0 = m.cell_mesh[c0];
        const int mi = mt0 + i;
        if (i < kTileMeshes && mi < m.n_meshes) {
            physT[4 * i] = m.mesh_phys[4 * mi];
            physT[4 * i + 1] = m.mesh_phys[4 * mi + 1];
            physT[4 * i + 2] = m.mesh_phys[4 * mi + 2];
            physT[4 * i + 3] = EPI == EPI_STAGE ? st.dt[mi] : 0.0f;
        }
    }

    // ---- phase 0: primitives of tile, H1 and H2 to LDS; the gradient inputs of this
    // thread's tile cell and H1 cell to registers ----
    int oinfo = 0;
    ushort4 cf = make_ushort4(0, 0, 0, 0);
    int64_t fA = 0;
    unsigned dA = 0;
    float geoA[2 * D];
    if constexpr (!SPLIT) {
        FzGradIn<D> og, hg;
        og.nr = make_ushort4(0, 0, 0, 0);
        og.eps = 0.0f;
        if (has_own) {
            float u[C];
            load_prim<D>(prim + (c0 + i) * C, u);
            store_vec<C>(Pl + i * C, u);
        }
        if (has_own) {
            fz_load_tile_cell<D>(m, c0 + i, og, oinfo, cf);
            if (m.mesh_phys) meshT[i] = static_cast<unsigned short>(m.cell_mesh[c0 + i]);
        }
        hg.eps = 0.0f;
        hg.nr = make_ushort4(0, 0, 0, 0);
        hg.fq = make_ushort4(kFzNone, kFzNone, kFzNone, kFzNone);
        int64_t hgc = 0;
        if (has_h1) hgc = fz_load_h1_cell<D>(m, h, halo_off, h1_off, hg);
        for (int q = i; q < n1 + n2; q += NT) {
            float u[C];
            load_prim<D>(prim + static_cast<int64_t>(m.fz_halo[halo_off + q]) * C, u);
            store_vec<C>(Pl + (kThreads + q) * C, u);
        }
        __syncthreads();

        // ---- phase 1: gradients (the first facet's inputs in flight during the H1
        // cell's).  The first facet's loads are issued after the tile cell's gradient,
        // not before: its registers are then not live through that gradient, which keeps
        // the 2-D kernel at 93 VGPRs (5 waves per SIMD, and 5 workgroups per CU with its
        // 29.5 KB of LDS) instead of 106 (4 waves per SIMD): 206.9 -> 197.2 us per RK4
        // step at 1M cells.
        if (has_own)
            fused_grad_cell<D, LP>(m, p, w, Pl, GL, FF, m.fz_ngl, i, c0 + i, true, og.gv, og.disp, og.eps, og.nr,
                                   has_prev, gmean, g0, og.fq);
        if (i < nft) fused_facet_load<D>(m, i, fo0, nfo, foff, fA, dA, geoA);
        if (has_h1)
            fused_grad_cell<D, LP>(m, p, w, Pl, GL, FF, m.fz_ngl, kThreads + i, hgc, false, hg.gv, hg.disp, hg.eps,
                                   hg.nr, has_prev, gmean, g0, hg.fq);
    } else {
        // one gradient cell per thread: tile cell i, or the H1 cell of local id i (its
        // inputs in the same registers)
        FzGradIn<D> og;
        int64_t ogc = c0 + i;
        if (has_own) {
            float u[C];
            load_prim<D>(prim + (c0 + i) * C, u);
            store_vec<C>(Pl + i * C, u);
            fz_load_tile_cell<D>(m, c0 + i, og, oinfo, cf);
            if (m.mesh_phys) meshT[i] = static_cast<unsigned short>(m.cell_mesh[c0 + i]);
        } else if (has_h1) {
            ogc = fz_load_h1_cell<D>(m, h, halo_off, h1_off, og);
        }
        for (int q = i; q < n1 + n2; q += NT) {
            float u[C];
            load_prim<D>(prim + static_cast<int64_t>(m.fz_halo[halo_off + q]) * C, u);
            store_vec<C>(Pl + (kThreads + q) * C, u);
        }
        __syncthreads();

        // ---- phase 1: gradients, then the first facet's loads ----
        if (has_own || has_h1)
            fused_grad_cell<D, LP>(m, p, w, Pl, GL, FF, m.fz_ngl, i, ogc, has_own, og.gv, og.disp, og.eps, og.nr,
                                   has_prev, gmean, g0, og.fq);
        if (i < nft) fused_facet_load<D>(m, i, fo0, nfo, foff, fA, dA, geoA);
    }
    __syncthreads();

    // ---- phase 2: facet fluxes (the next facet's and the epilogue's inputs in flight) ----
    int64_t fB = 0;
    unsigned dB = 0;
    float geoB[2 * D];
    const bool twoB = FS >= 2 && i + NT < nft;
    if (twoB) fused_facet_load<D>(m, i + NT, fo0, nfo, foff, fB, dB, geoB);
    const int64_t c = c0 + i;
    float dtm = 0.0f;
    if (EPI == EPI_STAGE && has_own) {
        if (m.mesh_phys) {
            const int mesh = meshT[i], rel = mesh - mt0;
            dtm = rel >= 0 && rel < kTileMeshes ? physT[4 * rel + 3] : st.dt[mesh];
        } else {
            dtm = st.dt[m.cell_mesh ? m.cell_mesh[c] : 0];
        }
    }
    float U0v[C], ap[C], a2p[C];
    stage_inputs<D, EPI>(st, c, has_own, U0v, ap, a2p);
    float inv_vol = 0.0f;
    if (has_own) inv_vol = m.cell_inv_vol[c];
    if (i < nft) fused_facet<D>(m, p, Pl, GL, m.fz_ngl, FF, i, fA, dA, geoA, meshT, mt0, physT);
    if constexpr (FS > 2) {
        // third facet of the thread: its inputs in flight during the second's flux
        int64_t fC = 0;
        unsigned dC = 0;
        float geoC[2 * D];
        const bool threeC = i + 2 * NT < nft;
        if (threeC) fused_facet_load<D>(m, i + 2 * NT, fo0, nfo, foff, fC, dC, geoC);
        if (twoB) fused_facet<D>(m, p, Pl, GL, m.fz_ngl, FF, i + NT, fB, dB, geoB, meshT, mt0, physT);
        if (threeC) fused_facet<D>(m, p, Pl, GL, m.fz_ngl, FF, i + 2 * NT, fC, dC, geoC, meshT, mt0, physT);
    } else if constexpr (FS == 2) {
        if (twoB) fused_facet<D>(m, p, Pl, GL, m.fz_ngl, FF, i + NT, fB, dB, geoB, meshT, mt0, physT);
    }
    __syncthreads();

    // ---- phase 3: divergence and stage epilogue of the tile's cells ----
    float acc[C];
#pragma unroll
    for (int j = 0; j < C; ++j) acc[j] = 0.0f;
    if (has_own) {
        const unsigned tf[4] = {cf.x, cf.y, cf.z, cf.w};
        float Fk[NL][C];
#pragma unroll
        for (int k = 0; k < NL; ++k) load_vec<C>(FF + (2 * tf[k]) * C, Fk[k]);
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const float vk = ((oinfo >> k) & 1) ? inv_vol : -inv_vol;
#pragma unroll
            for (int j = 0; j < C; ++j) acc[j] = acc[j] + vk * Fk[k][j];
        }
    }
    stage_epilogue<D, EPI>(m, p, w, dudt, st, tile, c, has_own, dtm, acc, U0v, ap, a2p, red, sred);
}

__global__ void k_update_stepsize(const double *__restrict__ red, int C, int nblk, float atol, float rtol, float mtol,
                                  float order, float a_gt, float a_le, float *dt) {
    // one wave: lane q < 3C sums its column over the blocks in a fixed order
    __shared__ double tot[64];
    const int q = threadIdx.x;
    double s = 0.0;
    if (q < 3 * C)
        for (int b = 0; b < nblk; ++b) s += red[static_cast<int64_t>(b) * 3 * C + q];
    tot[q] = s;
    __syncthreads();
    if (q == 0) {
        float E = 0.0f;
        for (int j = 0; j < C; ++j) {
            const float nd = static_cast<float>(sqrt(tot[j]));
            const float nh = static_cast<float>(sqrt(tot[C + j]));
            const float nu = static_cast<float>(sqrt(tot[2 * C + j]));
            const float e = nd / ((atol + rtol * nh) + mtol * nu);
            E = j == 0 ? e : E + e;
        }
        E = E / static_cast<float>(C);
        const float factor = 0.9f * powf(E, -1.0f / order);
        const float alpha = E > 1.0f ? a_gt : a_le;
        *dt = *dt * (alpha + (1.0f - alpha) * factor);
    }
}

__global__ void __launch_bounds__(kThreads) k_state_to_prim2(const float *__restrict__ s, float *__restrict__ p,
                                                             int64_t N, float inv_cv) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= N) return;
    float U[4], P[4];
    load_prim<2>(s + c * 4, U);
    state_to_prim<2>(U, inv_cv, P);
#pragma unroll
    for (int j = 0; j < 4; ++j) p[c * 4 + j] = P[j];
}

__global__ void __launch_bounds__(kThreads) k_state_to_prim3(const float *__restrict__ s, float *__restrict__ p,
                                                             int64_t N, float inv_cv) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (c >= N) return;
    float U[5], P[5];
    load_prim<3>(s + c * 5, U);
    state_to_prim<3>(U, inv_cv, P);
#pragma unroll
    for (int j = 0; j < 5; ++j) p[c * 5 + j] = P[j];
}

static int check_mesh(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const char *who) {
    if (!m || !p || !w || (m->dim != 2 && m->dim != 3) || m->n_cells <= 0 || m->n_groups < 0 ||
        m->n_groups > (1 << 16) || m->n_bc < 0 || !w->gp || (m->n_bc > 0 && !w->ubc) ||
        !m->fac_cell || !m->fac_geo || !w->fflux || m->n_facets < 1 || m->n_facets >= (int64_t(1) << 31) ||
        (m->n_groups > 0 && (!w->bc_part || m->n_slots < 1 || !m->blk_slot ||
                             !m->slot_blk || !m->slot_g0 || !m->slot_ng || !m->group_s0 || !m->group_s1 ||
                             !m->group_count)) ||
        (m->mesh_phys && !m->cell_mesh) || m->n_meshes < 1 ||
        p->lim_p < 1 || p->lim_p > 5 || m->n_cells > (int64_t(1) << 30)) {
        set_error("%s: invalid mesh/work/phys arguments", who);
        return FVM_ERR_ARG;
    }
    return FVM_OK;
}

// Dynamic LDS of k_fused in bytes (fused_tiles.py fused_lds_bytes is the same formula).
template <int D>
static size_t fused_lds_bytes(const FvmMeshDev &m) {
    // + the 16-bit mesh ids of the tile's cells and the tile's mesh table, for a batch
    return static_cast<size_t>(fz_ff_offset<D>(m.fz_ncl, m.fz_ngl) + static_cast<int64_t>(m.fz_nfl) * 2 * (D + 2)) *
               sizeof(float) +
           (m.mesh_phys ? kThreads * sizeof(unsigned short) + kTileMeshes * 4 * sizeof(float) : 0);
}

// LDS bytes one workgroup may use on the current device (160 KiB on gfx950), looked up
// once per device.
static size_t max_lds_per_workgroup() {
    constexpr int kMaxDev = 64;
    static int cached[kMaxDev] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return 64 * 1024;
    if (cached[dev] <= 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        cached[dev] = v;
    }
    return static_cast<size_t>(cached[dev]);
}

template <int D>
static int check_fused_d(const FvmMeshDev &m, const char *who) {
    if (!m.fz_tile || !m.fz_halo || !m.fz_hrec || !m.fz_cw || !m.fz_hw || !m.fz_fdesc || !m.fz_ffac ||
        !m.fz_fdesc_x || !m.fz_geo || !m.fz_geo_x || m.fz_ncl < kThreads || m.fz_ngl < kThreads ||
        m.fz_ngl > 2 * kThreads || m.fz_ngl > m.fz_ncl || m.fz_ncl > 0xFFF || m.fz_nfl < 1 ||
        m.fz_nfl > fz_fpt<D>() * kThreads || fused_lds_bytes<D>(m) > max_lds_per_workgroup() ||
        (m.mesh_phys && m.n_meshes > 0xFFFF) ||
        (m.fz_threads != 0 && m.fz_threads != kThreads && m.fz_threads != 320 && m.fz_threads != 384 &&
         m.fz_threads != 512) ||
        (m.fz_threads > kThreads && m.fz_ngl > m.fz_threads)) {
        set_error("%s: invalid tile tables for the tile-fused stage (or %zu B of LDS beyond the device's %zu)", who,
                  fused_lds_bytes<D>(m), max_lds_per_workgroup());
        return FVM_ERR_ARG;
    }
    return FVM_OK;
}

static int check_fused(const FvmMeshDev &m, const FvmWork &w, const char *who) {
    const int e = m.dim == 2 ? check_fused_d<2>(m, who) : m.dim == 3 ? check_fused_d<3>(m, who) : FVM_ERR_ARG;
    if (e) {
        if (m.dim != 2 && m.dim != 3) set_error("%s: invalid dimension", who);
        return e;
    }
    if (!w.gp_prev || w.gp_prev == w.gp) {
        set_error("%s: the tile-fused stage needs work->gp_prev as a buffer separate from work->gp", who);
        return FVM_ERR_ARG;
    }
    return FVM_OK;
}

template <typename K>
static void launch_fused_kernel(K kernel, int threads, dim3 grid, size_t lds, hipStream_t s, const FvmMeshDev &m,
                                const FvmPhys &p, const FvmWork &w, const float *prim, float *dudt,
                                const FvmStage &st, int has_prev) {
    if (lds > 64 * 1024)   // dynamic LDS beyond 64 KiB is requested explicitly
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL(kernel, grid, dim3(threads), lds, s, m, p, w, prim, dudt, st, has_prev);
}

template <int D, int EPI, int NT>
static void launch_fused_nt(const FvmMeshDev &m, const FvmPhys &p, const FvmWork &w, const float *prim, float *dudt,
                            const FvmStage &st, hipStream_t s, int has_prev) {
    const dim3 grid(EPI == EPI_STAGE && st.tile_list ? static_cast<unsigned int>(st.n_tile_list)
                                                     : blocks_for(m.n_cells, kThreads));
    const size_t lds = fused_lds_bytes<D>(m);
    switch (p.lim_p) {
        case 1: launch_fused_kernel(k_fused<D, 1, EPI, NT>, NT, grid, lds, s, m, p, w, prim, dudt, st, has_prev); break;
        case 2: launch_fused_kernel(k_fused<D, 2, EPI, NT>, NT, grid, lds, s, m, p, w, prim, dudt, st, has_prev); break;
        case 3: launch_fused_kernel(k_fused<D, 3, EPI, NT>, NT, grid, lds, s, m, p, w, prim, dudt, st, has_prev); break;
        case 5: launch_fused_kernel(k_fused<D, 5, EPI, NT>, NT, grid, lds, s, m, p, w, prim, dudt, st, has_prev); break;
        default: launch_fused_kernel(k_fused<D, 4, EPI, NT>, NT, grid, lds, s, m, p, w, prim, dudt, st, has_prev); break;
    }
}

// One workgroup per 256-cell tile, of m.fz_threads threads (0 = 256; 320, 384, 512).
template <int D, int EPI>
static void launch_fused(const FvmMeshDev &m, const FvmPhys &p, const FvmWork &w, const float *prim, float *dudt,
                         const FvmStage &st, hipStream_t s, int has_prev) {
    switch (m.fz_threads) {
        case 320: launch_fused_nt<D, EPI, 320>(m, p, w, prim, dudt, st, s, has_prev); break;
        case 384: launch_fused_nt<D, EPI, 384>(m, p, w, prim, dudt, st, s, has_prev); break;
        case 512: launch_fused_nt<D, EPI, 512>(m, p, w, prim, dudt, st, s, has_prev); break;
        default: launch_fused_nt<D, EPI, kThreads>(m, p, w, prim, dudt, st, s, has_prev); break;
    }
}

// parts: 1 = boundary-mean pre-pass, 2 = gradient + limiter (+ boundary values),
// 4 = facet fluxes, 8 = divergence (+ stage update), 16 = tile-fused stage (the caller
// has checked the tile tables).  The pre-pass runs only when `need_mean`.
template <int D>
static int launch_rhs(const FvmMeshDev &m, const FvmPhys &p, const FvmWork &w, const float *prim, float *dudt,
                      const FvmStage *st, hipStream_t s, int has_prev, int need_mean, int parts = 15) {
    if ((parts & 1) && need_mean && m.n_groups > 0)
        hipLaunchKernelGGL(k_bc_mean<D>, dim3(m.n_slots), dim3(kThreads), 0, s, m, p, w, prim);
    if (parts & 16) {
        if (st) {
            launch_fused<D, EPI_STAGE>(m, p, w, prim, nullptr, *st, s, has_prev);
        } else {
            FvmStage none{};
            launch_fused<D, EPI_DUDT>(m, p, w, prim, dudt, none, s, has_prev);
        }
        return check_launch(st ? "fvm_rk_stage_f32" : "fvm_rhs_f32");
    }
    const unsigned int ntiles = blocks_for(m.n_cells, kThreads);
    const dim3 grid(m.persist_blocks > 0 && static_cast<unsigned int>(m.persist_blocks) < ntiles
                        ? static_cast<unsigned int>(m.persist_blocks) : ntiles);
    if (parts & 2) {
        switch (p.lim_p) {
            case 1: hipLaunchKernelGGL((k_grad<D, 1>), grid, dim3(kThreads), 0, s, m, p, w, prim, has_prev); break;
            case 2: hipLaunchKernelGGL((k_grad<D, 2>), grid, dim3(kThreads), 0, s, m, p, w, prim, has_prev); break;
            case 3: hipLaunchKernelGGL((k_grad<D, 3>), grid, dim3(kThreads), 0, s, m, p, w, prim, has_prev); break;
            case 5: hipLaunchKernelGGL((k_grad<D, 5>), grid, dim3(kThreads), 0, s, m, p, w, prim, has_prev); break;
            default: hipLaunchKernelGGL((k_grad<D, 4>), grid, dim3(kThreads), 0, s, m, p, w, prim, has_prev); break;
        }
    }
    if (parts & 4) {
        const unsigned int ftiles = blocks_for(m.n_facets, kThreads);
        const dim3 fgrid(m.persist_blocks > 0 && static_cast<unsigned int>(m.persist_blocks) < ftiles
                             ? static_cast<unsigned int>(m.persist_blocks) : ftiles);
        hipLaunchKernelGGL(k_facet_flux<D>, fgrid, dim3(kThreads), 0, s, m, p, w, prim);
    }
    if (parts & 8) {
        if (st) {
            hipLaunchKernelGGL((k_div<D, EPI_STAGE>), grid, dim3(kThreads), 0, s, m, p, w, nullptr, *st);
        } else {
            FvmStage none{};
            hipLaunchKernelGGL((k_div<D, EPI_DUDT>), grid, dim3(kThreads), 0, s, m, p, w, dudt, none);
        }
    }
    return check_launch(st ? "fvm_rk_stage_f32" : "fvm_rhs_f32");
}

// 2-D per-cell records move as 16-byte vectors: their base pointers must be aligned.
static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

static int check_align(const FvmMeshDev *m, const FvmWork *w, const FvmStage *st, const float *dudt,
                       const float *prim, const char *who) {
    if (m->dim != 2) return FVM_OK;
    bool ok = aligned16(prim) && aligned16(w->ubc) && aligned16(w->fflux) && aligned16(dudt);
    if (st)
        ok = ok && aligned16(st->U0) && aligned16(st->acc) && aligned16(st->acc2) && aligned16(st->kbuf) &&
             aligned16(st->prim_next) && aligned16(st->U_out) && aligned16(st->f_m1) && aligned16(st->f_m2) &&
             aligned16(st->f_out);
    if (!ok) {
        set_error("%s: 2-D state/primitive/flux buffers must be 16-byte aligned", who);
        return FVM_ERR_ARG;
    }
    return FVM_OK;
}

static int check_stage(const FvmMeshDev *m, const FvmWork *w, const FvmStage *st, const char *who,
                       bool m_has_tiles = true) {
    if (m->n_groups > 0 && (!w->bc_part_next || w->bc_part_next == w->bc_part)) {
        set_error("%s: a stage needs work->bc_part_next as a buffer separate from work->bc_part", who);
        return FVM_ERR_ARG;
    }
    if (st && st->tile_list && !m_has_tiles) {
        set_error("%s: a tile list needs the tile-fused stage (mesh with tile tables)", who);
        return FVM_ERR_ARG;
    }
    if (st && st->tile_list && (st->n_tile_list < 1 || st->n_tile_list > (m->n_cells + kThreads - 1) / kThreads)) {
        set_error("%s: n_tile_list %d outside [1, %lld]", who, st->n_tile_list,
                  static_cast<long long>((m->n_cells + kThreads - 1) / kThreads));
        return FVM_ERR_ARG;
    }
    if (st && st->mode == 1) {
        if (!st->prim || !st->U0 || !st->prim_next || !st->dt || !st->U_out || !st->f_out || st->n_hist < 0 ||
            (st->n_hist >= 1 && !st->f_m1) || (st->n_hist >= 2 && !st->f_m2) || st->acc2) {
            set_error("%s: invalid Adams-Bashforth stage arguments", who);
            return FVM_ERR_ARG;
        }
        return FVM_OK;
    }
    if (!st || st->mode != 0 || !st->prim || !st->U0 || !st->acc || !st->prim_next || !st->dt || (st->last && !st->U_out) ||
        st->n_inc < 0 || st->n_inc > 10 || (st->store_k && !st->kbuf) || (st->last && st->acc2 && !w->red)) {
        set_error("%s: invalid stage arguments", who);
        return FVM_ERR_ARG;
    }
    for (int t = 0; t < st->n_inc; ++t)
        if (st->inc_idx[t] < 0 || st->inc_idx[t] > st->stage || (st->inc_idx[t] != st->stage && !st->kbuf)) {
            set_error("%s: increment refers to a stage that is not stored", who);
            return FVM_ERR_ARG;
        }
    return FVM_OK;
}

}  // namespace fvm

using namespace fvm;

extern "C" {

int fvm_rhs_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const float *prim, float *dudt,
                int32_t has_prev, void *stream) {
    int e = check_mesh(m, p, w, "fvm_rhs_f32");
    if (e) return e;
    if (!prim || !dudt) {
        set_error("fvm_rhs_f32: null prim/dudt");
        return FVM_ERR_ARG;
    }
    e = check_align(m, w, nullptr, dudt, prim, "fvm_rhs_f32");
    if (!e && m->fz_tile) e = check_fused(*m, *w, "fvm_rhs_f32");
    if (e) return e;
    if (m->fz_tile)
        return m->dim == 2 ? launch_rhs<2>(*m, *p, *w, prim, dudt, nullptr, as_stream(stream), has_prev, 1, 1 | 16)
                           : launch_rhs<3>(*m, *p, *w, prim, dudt, nullptr, as_stream(stream), has_prev, 1, 1 | 16);
    return m->dim == 2 ? launch_rhs<2>(*m, *p, *w, prim, dudt, nullptr, as_stream(stream), has_prev, 1)
                       : launch_rhs<3>(*m, *p, *w, prim, dudt, nullptr, as_stream(stream), has_prev, 1);
}

int fvm_grad_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const float *prim, int32_t has_prev,
                 void *stream) {
    int e = check_mesh(m, p, w, "fvm_grad_f32");
    if (e) return e;
    if (!prim || (w->gp_prev && w->gp_prev == w->gp)) {
        set_error("fvm_grad_f32: null prim, or gp_prev given equal to gp");
        return FVM_ERR_ARG;
    }
    e = check_align(m, w, nullptr, nullptr, prim, "fvm_grad_f32");
    if (e) return e;
    return m->dim == 2 ? launch_rhs<2>(*m, *p, *w, prim, nullptr, nullptr, as_stream(stream), has_prev, 1, 1 | 2)
                       : launch_rhs<3>(*m, *p, *w, prim, nullptr, nullptr, as_stream(stream), has_prev, 1, 1 | 2);
}

int fvm_rk_stage_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const FvmStage *st, void *stream) {
    int e = check_mesh(m, p, w, "fvm_rk_stage_f32");
    if (!e) e = check_stage(m, w, st, "fvm_rk_stage_f32", m->fz_tile != nullptr);
    if (!e) e = check_align(m, w, st, nullptr, st->prim, "fvm_rk_stage_f32");
    if (!e && m->fz_tile) e = check_fused(*m, *w, "fvm_rk_stage_f32");
    if (e) return e;
    if (m->fz_tile)
        return m->dim == 2
                   ? launch_rhs<2>(*m, *p, *w, st->prim, nullptr, st, as_stream(stream), st->has_prev, !st->bc_ready,
                                   1 | 16)
                   : launch_rhs<3>(*m, *p, *w, st->prim, nullptr, st, as_stream(stream), st->has_prev, !st->bc_ready,
                                   1 | 16);
    return m->dim == 2 ? launch_rhs<2>(*m, *p, *w, st->prim, nullptr, st, as_stream(stream), st->has_prev, !st->bc_ready)
                       : launch_rhs<3>(*m, *p, *w, st->prim, nullptr, st, as_stream(stream), st->has_prev, !st->bc_ready);
}

int fvm_rk_stage_part_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const FvmStage *st,
                          int32_t parts, void *stream) {
    if (parts < 1 || parts > 31 || ((parts & 16) && (parts & 14))) {
        set_error("fvm_rk_stage_part_f32: parts must be a nonzero mask of 1|2|4|8, or 1|16");
        return FVM_ERR_ARG;
    }
    int e = check_mesh(m, p, w, "fvm_rk_stage_part_f32");
    if (!e) e = check_stage(m, w, st, "fvm_rk_stage_part_f32", m->fz_tile != nullptr && !(parts & 14));
    if (!e) e = check_align(m, w, st, nullptr, st->prim, "fvm_rk_stage_part_f32");
    if (!e && (parts & 16)) e = check_fused(*m, *w, "fvm_rk_stage_part_f32");
    if (e) return e;
    return m->dim == 2
               ? launch_rhs<2>(*m, *p, *w, st->prim, nullptr, st, as_stream(stream), st->has_prev, !st->bc_ready, parts)
               : launch_rhs<3>(*m, *p, *w, st->prim, nullptr, st, as_stream(stream), st->has_prev, !st->bc_ready, parts);
}

int fvm_max_lds_per_workgroup(int32_t *out) {
    if (!out) {
        set_error("fvm_max_lds_per_workgroup: null out");
        return FVM_ERR_ARG;
    }
    *out = static_cast<int32_t>(max_lds_per_workgroup());
    return FVM_OK;
}

int fvm_update_stepsize(const FvmWork *w, int32_t C, int32_t nblk, float atol, float rtol, float mtol, float order,
                        float a_gt, float a_le, float *dt, void *stream) {
    if (!w || !w->red || !dt || C < 1 || 3 * C > 64 || nblk < 1) {
        set_error("fvm_update_stepsize: invalid arguments");
        return FVM_ERR_ARG;
    }
    hipLaunchKernelGGL(k_update_stepsize, dim3(1), dim3(64), 0, as_stream(stream), w->red, C, nblk, atol, rtol, mtol,
                       order, a_gt, a_le, dt);
    return check_launch("fvm_update_stepsize");
}

int fvm_state_to_prim_f32(const float *state, float *prim, int64_t N, int32_t dim, float inv_cv, void *stream) {
    if (!state || !prim || N < 0 || (dim != 2 && dim != 3) || (dim == 2 && !(aligned16(state) && aligned16(prim)))) {
        set_error("fvm_state_to_prim_f32: invalid arguments (2-D buffers must be 16-byte aligned)");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    if (dim == 2)
        hipLaunchKernelGGL(k_state_to_prim2, dim3(blocks_for(N, kThreads)), dim3(kThreads), 0, as_stream(stream),
                           state, prim, N, inv_cv);
    else
        hipLaunchKernelGGL(k_state_to_prim3, dim3(blocks_for(N, kThreads)), dim3(kThreads), 0, as_stream(stream),
                           state, prim, N, inv_cv);
    return check_launch("fvm_state_to_prim_f32");
}

}  // extern "C"
