// FP64 geometric precompute on the GPU: facet normals and midpoints, cell volumes and
// centroids, centroid->facet displacements with the facet orientation signs, and the
// weighted least-squares gradient matrices (FVMMesh2D/3D, fvm_mesh.py:131-261, 263-344,
// 367-582).  Everything is fp64: the LSQ normal matrices of sliver cells are badly
// conditioned, and the operators are rounded to fp32 only when the device layout is
// packed.  The integer topology (facet lookup, left/right ordering) stays on the host.
#include "fvm_common.hpp"

namespace fvm {

// 2-D: n = (dy, -dx) of v1 - v0, |n| = edge length.  3-D: n = (v1 - v0) x (v2 - v0),
// |n| = twice the face area (the reference's convention).  mid = vertex mean.
template <int D>
__global__ void __launch_bounds__(256)
k_geom_facets(const double *__restrict__ V, const int64_t *__restrict__ Fv, int64_t F, double *__restrict__ n,
              double *__restrict__ nh, double *__restrict__ mid) {
    const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double p[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int i = 0; i < D; ++i) p[a][i] = V[Fv[f * D + a] * D + i];
    double nn[D];
    if constexpr (D == 2) {
        nn[0] = p[1][1] - p[0][1];
        nn[1] = -(p[1][0] - p[0][0]);
    } else {
        const double e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const double e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
        nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
        nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) s += nn[i] * nn[i];
    const double len = sqrt(s);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        n[f * D + i] = nn[i];
        nh[f * D + i] = nn[i] / len;
        double m = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) m += p[a][i];
        mid[f * D + i] = m / D;
    }
}

// 2-D volume = |cross(b - a, c - a)| / 2, 3-D = |det(b - a, c - a, d - a)| / 6;
// centroid = vertex mean.
template <int D>
__global__ void __launch_bounds__(256)
k_geom_cells(const double *__restrict__ V, const int64_t *__restrict__ Cv, int64_t N, double *__restrict__ vol,
             double *__restrict__ cen) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    double p[D + 1][D];
#pragma unroll
    for (int a = 0; a <= D; ++a)
#pragma unroll
        for (int i = 0; i < D; ++i) p[a][i] = V[Cv[c * (D + 1) + a] * D + i];
    if constexpr (D == 2) {
        const double ab0 = p[1][0] - p[0][0], ab1 = p[1][1] - p[0][1];
        const double ac0 = p[2][0] - p[0][0], ac1 = p[2][1] - p[0][1];
        vol[c] = 0.5 * fabs(ab0 * ac1 - ab1 * ac0);
    } else {
        double b[3], cc[3], d[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            b[i] = p[1][i] - p[0][i];
            cc[i] = p[2][i] - p[0][i];
            d[i] = p[3][i] - p[0][i];
        }
        const double x0 = cc[1] * d[2] - cc[2] * d[1], x1 = cc[2] * d[0] - cc[0] * d[2], x2 = cc[0] * d[1] - cc[1] * d[0];
        vol[c] = fabs(b[0] * x0 + b[1] * x1 + b[2] * x2) / 6.0;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double m = 0.0;
#pragma unroll
        for (int a = 0; a <= D; ++a) m += p[a][i];
        cen[c * D + i] = m / (D + 1);
    }
}

// disp[c][k] = mid[f] - cen[c] for the cell's k-th facet f = c2f[c][k];
// sign = sign(n_hat[f] . disp) (+1 when the normal points out of the cell)
template <int D>
__global__ void __launch_bounds__(256)
k_geom_disp_signs(const double *__restrict__ cen, const double *__restrict__ mid, const double *__restrict__ nh,
                  const int64_t *__restrict__ c2f, int64_t N, double *__restrict__ disp, int64_t *__restrict__ sgn) {
    constexpr int NL = D + 1;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= N * NL) return;
    const int64_t c = t / NL;
    const int64_t f = c2f[t];
    double dot = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const double dv = mid[f * D + i] - cen[c * D + i];
        disp[t * D + i] = dv;
        dot += nh[f * D + i] * dv;
    }
    sgn[t] = (dot > 0.0) - (dot < 0.0);
}

// Weighted least squares (fvm_mesh.py:263-344): offsets d_k to the neighbour centroid, or
// to the facet midpoint on boundary facets; w_k = |d_k|^(-1/4), x 7.5 on boundary facets;
// A = sum_k w_k^2 d_k d_k^T; G = A^{-1} [d_k w_k^2]_k  -> G[c][i][k].
template <int D>
__global__ void __launch_bounds__(256)
k_geom_lsq(const double *__restrict__ cen, const double *__restrict__ mid, const int64_t *__restrict__ c2f,
           const int64_t *__restrict__ other, const uint8_t *__restrict__ is_bc, int64_t N, double *__restrict__ G) {
    constexpr int NL = D + 1;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    double dv[NL][D], w2[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int64_t t = c * NL + k;
        const bool bc = is_bc[t] != 0;
        const double *q = bc ? mid + c2f[t] * D : cen + other[t] * D;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            dv[k][i] = q[i] - cen[c * D + i];
            s += dv[k][i] * dv[k][i];
        }
        double w = pow(sqrt(s), -0.25);
        if (bc) w *= 7.5;
        w2[k] = w * w;
    }
    double A[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NL; ++k) s += w2[k] * dv[k][i] * dv[k][j];
            A[i][j] = s;
        }
    double Ai[D][D];
    if constexpr (D == 2) {
        const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        Ai[0][0] = A[1][1] / det;
        Ai[0][1] = -A[0][1] / det;
        Ai[1][0] = -A[1][0] / det;
        Ai[1][1] = A[0][0] / det;
    } else {
        const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
        const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
        const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
        const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
        Ai[0][0] = c00 / det;
        Ai[1][0] = c01 / det;
        Ai[2][0] = c02 / det;
        Ai[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
        Ai[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
        Ai[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
        Ai[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
        Ai[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
        Ai[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) s += Ai[i][j] * dv[k][j];
            G[(c * D + i) * NL + k] = s * w2[k];
        }
}

}  // namespace fvm

using namespace fvm;

extern "C" {

int fvm_geom_facets_f64(const double *V, int32_t dim, const int64_t *facets, int64_t F, double *normals,
                        double *normals_hat, double *midpoints, void *stream) {
    if (!V || !facets || !normals || !normals_hat || !midpoints || F < 0 || (dim != 2 && dim != 3)) {
        set_error("fvm_geom_facets_f64: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (F == 0) return FVM_OK;
    if (dim == 2)
        hipLaunchKernelGGL(k_geom_facets<2>, dim3(blocks_for(F, 256)), dim3(256), 0, as_stream(stream), V, facets, F,
                           normals, normals_hat, midpoints);
    else
        hipLaunchKernelGGL(k_geom_facets<3>, dim3(blocks_for(F, 256)), dim3(256), 0, as_stream(stream), V, facets, F,
                           normals, normals_hat, midpoints);
    return check_launch("fvm_geom_facets_f64");
}

int fvm_geom_cells_f64(const double *V, int32_t dim, const int64_t *cells, int64_t N, double *volumes,
                       double *centroids, void *stream) {
    if (!V || !cells || !volumes || !centroids || N < 0 || (dim != 2 && dim != 3)) {
        set_error("fvm_geom_cells_f64: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    if (dim == 2)
        hipLaunchKernelGGL(k_geom_cells<2>, dim3(blocks_for(N, 256)), dim3(256), 0, as_stream(stream), V, cells, N,
                           volumes, centroids);
    else
        hipLaunchKernelGGL(k_geom_cells<3>, dim3(blocks_for(N, 256)), dim3(256), 0, as_stream(stream), V, cells, N,
                           volumes, centroids);
    return check_launch("fvm_geom_cells_f64");
}

int fvm_geom_disp_signs_f64(const double *centroids, const double *midpoints, const double *normals_hat, int32_t dim,
                            const int64_t *cell_to_facet, int64_t N, double *disp, int64_t *signs, void *stream) {
    if (!centroids || !midpoints || !normals_hat || !cell_to_facet || !disp || !signs || N < 0 ||
        (dim != 2 && dim != 3)) {
        set_error("fvm_geom_disp_signs_f64: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    const int64_t T = N * (dim + 1);
    if (dim == 2)
        hipLaunchKernelGGL(k_geom_disp_signs<2>, dim3(blocks_for(T, 256)), dim3(256), 0, as_stream(stream), centroids,
                           midpoints, normals_hat, cell_to_facet, N, disp, signs);
    else
        hipLaunchKernelGGL(k_geom_disp_signs<3>, dim3(blocks_for(T, 256)), dim3(256), 0, as_stream(stream), centroids,
                           midpoints, normals_hat, cell_to_facet, N, disp, signs);
    return check_launch("fvm_geom_disp_signs_f64");
}

int fvm_geom_lsq_f64(const double *centroids, const double *midpoints, int32_t dim, const int64_t *cell_to_facet,
                     const int64_t *other, const uint8_t *is_bc, int64_t N, double *G, void *stream) {
    if (!centroids || !midpoints || !cell_to_facet || !other || !is_bc || !G || N < 0 || (dim != 2 && dim != 3)) {
        set_error("fvm_geom_lsq_f64: invalid arguments");
        return FVM_ERR_ARG;
    }
    if (N == 0) return FVM_OK;
    if (dim == 2)
        hipLaunchKernelGGL(k_geom_lsq<2>, dim3(blocks_for(N, 256)), dim3(256), 0, as_stream(stream), centroids,
                           midpoints, cell_to_facet, other, is_bc, N, G);
    else
        hipLaunchKernelGGL(k_geom_lsq<3>, dim3(blocks_for(N, 256)), dim3(256), 0, as_stream(stream), centroids,
                           midpoints, cell_to_facet, other, is_bc, N, G);
    return check_launch("fvm_geom_lsq_f64");
}

}  // extern "C"
