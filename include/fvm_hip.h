/*
 * fvm_hip.h -- C ABI of the MI355X (gfx950) FVM step-loop library (libfvm_hip.so).
 *
 * Every pointer is a caller-owned device buffer unless stated otherwise; nothing on the
 * hot path allocates.  Every call takes the HIP stream to launch on as `void *stream`
 * (a hipStream_t; NULL is the default stream), launches asynchronously and returns 0 on
 * success or a nonzero FVM_ERR_* code; fvm_last_error() then describes the failure
 * (thread-local).  The struct arguments (FvmMeshDev, FvmPhys, FvmWork, FvmStage) live in
 * host memory and carry device pointers; the library keeps no reference to them after
 * the call returns.
 *
 * Layout conventions (N cells, F facets, Fb boundary facets, d = 2|3, C = d + 2,
 * NL = d + 1 facets per cell; "[a][b]" = a-major, b contiguous):
 *   state, prim, dudt : [N][C] fp32, component order as the reference
 *                       (momentum/velocity..., rho, E/T) -- fvm_equation.py:35-55
 *   per-cell slot arrays are slot-major, [slot][N] (one coalesced load per slot).
 */
#ifndef FVM_HIP_H
#define FVM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVM_OK 0
#define FVM_ERR_ARG 1          /* invalid argument (shape, size, null pointer)            */
#define FVM_ERR_ROW_OVERFLOW 2 /* csr_to_ell: a row holds more nonzeros than K            */
#define FVM_ERR_HIP 3          /* a HIP runtime call or kernel launch failed              */

/* Description of the last failure on the calling thread ("" if none). */
const char *fvm_last_error(void);
/* Library ABI version (bumped on any struct, signature or array-layout change). */
int fvm_abi_version(void);

/* ---------------------------------------------------------------------------------
 * ELL sparse products.  Replace the Triton kernels of utils/ell_kernel.py:24-210
 * (ell_spmm_kernel, _ell_spmv_kernel, ell_spmv, ell_spmm).
 *   y[r, j] = sum_{k<K} vals[r, k] * x[cols[r, k], j]  (+ bias[r])
 * vals/cols are addressed through explicit strides (element units), so both the
 * reference's row-major [M, K] layout and this library's slot-major layout
 * (stride_m = 1, stride_k = M) are accepted; cols are int32 (cols_is64 = 0) or int64.
 * Accumulation is fp32 in slot order k = 0..K-1; bias may be NULL.
 * ------------------------------------------------------------------------------- */
int fvm_ell_spmm_f32(const float *vals, int64_t vals_stride_m, int64_t vals_stride_k,
                     const void *cols, int32_t cols_is64, int64_t cols_stride_m, int64_t cols_stride_k,
                     int64_t M, int32_t K,
                     const float *x, int64_t x_stride_n, int64_t x_stride_b, int32_t B,
                     const float *bias,
                     float *y, int64_t y_stride_m, int64_t y_stride_b, void *stream);

/* fvm_ell_spmm_f32 with an explicit launch shape for the fast path (slot-major int32
 * ELL, contiguous x and y, B in {1, 4}, K <= 8 except 7): rows_per_lane in {1, 2, 4},
 * block_threads in {64, 128, 256, 512}.  Every shape computes bitwise the same y; the
 * autotuner times them (the reference autotunes its Triton kernels over block shape,
 * warps and stages, utils/ell_kernel.py:8-23, 83-98, keyed on M, B, K and the bias).
 * fvm_ell_spmm_f32 is this call with (1, 256).  Other layouts ignore the shape. */
int fvm_ell_spmm_f32_shape(const float *vals, int64_t vals_stride_m, int64_t vals_stride_k,
                           const void *cols, int32_t cols_is64, int64_t cols_stride_m, int64_t cols_stride_k,
                           int64_t M, int32_t K,
                           const float *x, int64_t x_stride_n, int64_t x_stride_b, int32_t B,
                           const float *bias,
                           float *y, int64_t y_stride_m, int64_t y_stride_b,
                           int32_t rows_per_lane, int32_t block_threads, void *stream);

/* 1 if a product with these strides, K, B and base pointers takes the fast path (the
 * launch shape then matters), else 0.  No device access. */
int fvm_ell_fast_path(int64_t vals_stride_m, int64_t vals_stride_k, int32_t cols_is64, int64_t cols_stride_m,
                      int64_t cols_stride_k, int32_t K, int64_t x_stride_n, int64_t x_stride_b, int32_t B,
                      int64_t y_stride_m, int64_t y_stride_b, const void *x, const void *y);

/* ---------------------------------------------------------------------------------
 * CSR -> ELL packing.  Replaces csr_to_ell (utils/ell_kernel.py:225-256): slot k of
 * row r holds the k-th stored entry of the row, padding is col 0 / val 0.
 * fvm_csr_max_row_nnz writes max_r(crow[r+1]-crow[r]) to *out_max (device int32; the
 * caller zeroes it first).  fvm_ell_pack_f32 writes slot-major vals/cols
 * (index k * slot_stride + r) and sets *overflow (device int32, caller-zeroed) to 1 if
 * a row holds more than K entries -- the reference raises ValueError there.
 * ------------------------------------------------------------------------------- */
int fvm_csr_max_row_nnz(const int64_t *crow, int64_t M, int32_t *out_max, void *stream);
int fvm_ell_pack_f32(const int64_t *crow, const int64_t *col, const float *val, int64_t M, int32_t K,
                     float *ell_vals, int32_t *ell_cols, int64_t slot_stride, int32_t *overflow,
                     void *stream);

/* ---------------------------------------------------------------------------------
 * Fused RHS.  Device-resident mesh produced once at set-up from the reference's
 * MeshCache / BoundarySetter / BC arrays (facet_process.py:58-226,
 * facet_boundary.py:32-385).  Per-cell slot arrays list a cell's facets in ascending
 * facet index (the column order of flux_mat's ELL rows).
 * ------------------------------------------------------------------------------- */
typedef struct FvmMeshDev {
    int32_t dim;          /* 2 or 3                                                   */
    int32_t n_groups;     /* characteristic BC groups (<= 4)                          */
    int64_t n_cells;      /* N                                                        */
    int64_t n_facets;     /* F                                                        */
    int64_t n_bc;         /* Fb                                                       */
    /* per cell -------------------------------------------------------------------- */
    const int32_t *cell_facet;  /* [NL][N] facet ids, ascending                       */
    const int32_t *cell_nbr;    /* [NL][N] neighbour cell, or N + boundary index      */
    const int32_t *cell_info;   /* [N] bit k: this cell's slot in facet k;
                                   bits 8+2k..9+2k: facet k's position in the
                                   neighbour's list                                   */
    const float *cell_disp;     /* [NL][d][N] centroid -> facet midpoint; slot NL-1
                                   holds -(fp32 sum of slots 0..NL-2 in slot order),
                                   which the fused stage derives instead of reading  */
    const float *cell_gval;     /* [NL+1][d][N] LSQ gradient row values (G_mats rows
                                   d*c+i) in slot order: the column of neighbour k
                                   (or boundary value) for k < NL, then the cell's    */
    const float *cell_inv_vol;  /* [N] fp32(1/vol): |flux_mat| entries                */
    const float *cell_eps;      /* [N] limiter epsilon (K sqrt(vol))^(p+1)            */
    /* per facet ------------------------------------------------------------------- */
    /* Facets are stored in a device order of their own (by lowest adjacent cell, the
       boundary facets of each 256-cell run of lowest cells last), and cell_facet holds
       positions in that order.                                                         */
    const float *fac_geo;       /* [F][G] G = 12 (2-D) / 20 (3-D): n[d], n_hat[d],
                                   |n|, 1/(n_hat.d), X_orthog's d-1 coordinates in
                                   the tangent basis of n_hat and a 0 (X_orthog is
                                   orthogonal to n_hat; 0 on the boundary),
                                   centroid->midpoint of the slot-0, slot-1 cells  */
    const int32_t *fac_cell;    /* [2][F] slot-0 and slot-1 cell; N + boundary index
                                   for the outer side of a boundary facet             */
    /* per boundary facet (boundary-local index b) ---------------------------------- */
    const int32_t *bc_cell;     /* [Fb] interior cell                                 */
    const int32_t *bc_group;    /* [Fb] characteristic group, or -1                   */
    const int32_t *bc_flags;    /* [Fb] bit j: Neumann comp j (A_bc entry + A_corr)   */
    const float *bc_val;        /* [Fb][C] b_bc (Dirichlet value / dUdn * dist)       */
    const float *bc_l;          /* [Fb][d] tangential centroid->facet vector l        */
    const float *bc_ga;         /* [Fb][C] A_face_grad coefficient (-1/dist or 0)     */
    const float *bc_gb;         /* [Fb][C] b_face_grad                                */
    const float *bc_char;       /* [Fb][2d+3]: n_out[d], rho_inf, v_n_inf, p_inf,
                                   v_t_inf[d]                                         */
    /* characteristic-BC group means (facet_boundary.py:427); a group = one BC mode of
       one mesh ----------------------------------------------------------------------- */
    int32_t n_slots;            /* 256-cell tiles holding cells of group facets       */
    int32_t n_meshes;           /* meshes concatenated block-diagonally (>= 1)        */
    const int32_t *blk_slot;    /* [ceil(N/256)] partial slot of a tile, or -1        */
    const int32_t *slot_blk;    /* [n_slots] tile of each slot                        */
    const int32_t *slot_g0;     /* [n_slots] first group of the slot's mesh           */
    const int32_t *slot_ng;     /* [n_slots] groups of that mesh (<= 4)               */
    const int32_t *group_s0;    /* [n_groups] first slot of the group's mesh          */
    const int32_t *group_s1;    /* [n_groups] one past its last slot                  */
    const float *group_count;   /* [n_groups] facets per group                        */
    /* batches of meshes (NULL for one mesh) ----------------------------------------- */
    const int32_t *cell_mesh;   /* [N] mesh of each cell; tiles never straddle meshes */
    const float *mesh_phys;     /* [n_meshes][4] mu0, mub0, kappa, 0 per mesh         */
    int32_t persist_blocks;     /* workgroups of the per-cell kernels (0 = one per
                                   256-cell tile); tiles are walked grid-stride       */
    int32_t fz_threads;         /* threads per tile of the tile-fused stage: 0 or 256
                                   (a thread per tile cell, which also takes an H1
                                   cell and up to 2 (2-D) / 3 (3-D) facets), or 320,
                                   384, 512 (tile cells on threads 0..255, H1 cells on
                                   the others: needs fz_ngl <= fz_threads; fewer facets
                                   per thread); same results either way              */
    /* tile-fused stage (2-D and 3-D; fz_tile NULL = the three-kernel pipeline; the
       tables must fit fvm_max_lds_per_workgroup, fz_ncl <= 4095, fz_ngl <= 512 and
       fz_nfl <= 512 in 2-D, 768 in 3-D).  Tile t holds
       cells 256t..256t+255 (local ids 0..255), its one-ring halo H1 (local ids 256..,
       ascending global id) and H1's other neighbours H2 (after H1).  Its facets are
       the owned run [fo0, fo0+nfo) of the device facet order (lowest adjacent cell in
       the tile), then nfor foreign facets (interior facets of a tile cell whose lowest
       cell is in an earlier tile).  Facet descriptor bits: 0-11 local id of the slot-0
       cell, 12-13 its slot index k, 14-25 / 26-27 the same for the slot-1 cell, 28
       slot-0 cell has the lower global id, 29 boundary facet, 30 the outer side is
       slot 0 (outer local id 0xFFF).                                                 */
    const int32_t *fz_tile;     /* [T][12] halo_off, n1, n2, (unused), fo0, nfo, foff,
                                   nfor, g0, ng (group means the tile needs), h1_off  */
    const int32_t *fz_halo;     /* H1 then H2 global ids of each tile, at halo_off    */
    const float *fz_hrec;       /* [HN][fz_nh1] H1 records, component-major (record of
                                   H1 entry h1_off + i in column h1_off + i): LSQ
                                   values of the NL neighbour columns [NL][d] (the
                                   own column is derived, below), displacements
                                   of slots 0..NL-2 [NL-1][d] (the last is minus
                                   their sum), eps; HN = 11 (2-D) / 22 (3-D)          */
    const uint32_t *fz_cw;      /* [NL][N] slot words of every cell (slot-major):
                                   bits 0-11 local id of the neighbour behind the
                                   slot in the cell's tile (0xFFF: boundary facet),
                                   12-21 the facet's position in the tile's list,
                                   22 the cell is the facet's slot-1 side, 23-24 the
                                   slot's position in the reference's local order of
                                   the LSQ row: the kernel sums the neighbour columns
                                   in that order and negates for the cell's own
                                   column, bitwise the stored G_mats diagonal        */
    const uint32_t *fz_hw;      /* [NL][fz_nh1] the same for H1 entries (local ids in
                                   the tile; position 0x3FF: facet not in the list)  */
    const uint32_t *fz_fdesc;   /* [F] descriptor of each facet in its owner tile     */
    const int32_t *fz_ffac;     /* foreign facet ids of each tile, at foff            */
    const uint32_t *fz_fdesc_x; /* their descriptors                                  */
    const float *fz_geo;        /* [F][2d] n, 1/(n_hat.d), X_orthog's tangent
                                   coordinates [d-1] (device order); |n| and n_hat
                                   are derived from n as the reference derives them
                                   (facet_process.py:84-86), X_orthog from n_hat and
                                   the coordinates (2-D tangent (-n_hat_y, n_hat_x);
                                   3-D the basis of Duff et al., JCGT 6(1), 2017)    */
    const float *fz_geo_x;      /* [sum nfor][2d] the same for the foreign lists      */
    int32_t fz_ncl;             /* LDS rows: max staged cells (256 + n1 + n2)         */
    int32_t fz_ngl;             /*           max gradient cells (256 + n1)            */
    int32_t fz_nfl;             /*           max facets (nfo + nfor)                  */
    int32_t fz_nh1;             /* sum n1: H1 entries of all tiles (fz_hrec stride)   */
} FvmMeshDev;

typedef struct FvmPhys {      /* scalars of FluidConstitution / SlopeLimiter / KT      */
    float inv_cv;     /* fp32(1 / C_v)                                                */
    float cv;         /* C_v                                                          */
    float R;          /* C_v (gamma - 1)                                              */
    float gR;         /* fp32(gamma * R)                                              */
    float mu0, T0, T0pS, S, mub0, T0sq; /* Sutherland / bulk viscosity constants       */
    float kappa;      /* thermal conductivity                                         */
    float v_factor;   /* KT velocity clamp                                            */
    int32_t lim_p;    /* limiter order 1..5                                           */
} FvmPhys;

typedef struct FvmWork {      /* device scratch owned by the caller                    */
    float *ubc;       /* [Fb][C] boundary face values                                 */
    float *gp;        /* [N][(d+1)C] unlimited LSQ gradient [d][C] then limiter phi[C];
                         persists between calls (the next call's Neumann correction)  */
    float *bc_part;   /* [n_slots][4] partial sums of c per slot and group, of the
                         primitives the next gradient launch reads                    */
    double *red;      /* [ceil(N/256)][3C] fp64 partial column norms (adaptive dt)    */
    float *fflux;     /* [F][C] facet fluxes of the current evaluation (device order) */
    const float *gp_prev; /* records of the previous call, read by the Neumann
                         correction: gp itself (NULL) for the three-kernel pipeline;
                         a second buffer for the tile-fused stage, which writes the
                         records of boundary cells only, into gp (callers swap the two
                         buffers between calls)                                        */
    float *bc_part_next; /* [n_slots][4] where a stage epilogue writes the partial sums
                         of c of the next stage's primitives (required by the stage
                         entry points when n_groups > 0, separate from bc_part: tiles
                         of one launch still read bc_part; callers swap the two after
                         each stage)                                                   */
} FvmWork;

/* One RHS evaluation dU/dt = FVMEquation.forward(prim) (fvm_equation.py:406-435).
 * has_prev != 0 applies the Neumann non-orthogonal correction with the gradients the
 * previous call left (work->gp_prev, or work->gp) (facet_boundary.py:135-136).  With
 * tile tables (m->fz_tile) one tile-fused launch, else up to four: boundary-mean
 * pre-pass, gradient + limiter (+ boundary values), facet fluxes, divergence. */
int fvm_rhs_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const float *prim,
                float *dudt, int32_t has_prev, void *stream);

/* One explicit Runge-Kutta stage with the RHS fused into the stage update
 * (Butcher / ButcherAdapt, integrators.py:305-393).  With k_i = dt * f(prim_i):
 *   acc  (+)= b_i k_i          acc2 (+)= b2_i k_i  (if acc2)      k store if store_k
 *   not last: prim_next = prim(U0 + sum_j inc_coef[j] k_{inc_idx[j]})
 *   last:     U_out = U0 + acc, prim_next = prim(U_out); with acc2 also fp64 partial
 *             column sums of squares of (acc-acc2), acc, U0 into work->red.           */
typedef struct FvmStage {
    const float *prim;     /* [N][C] stage input primitives                           */
    const float *U0;       /* [N][C] state at the start of the step                   */
    float *acc;            /* [N][C] sum b_j k_j                                      */
    float *acc2;           /* [N][C] sum b2_j k_j or NULL                             */
    float *kbuf;           /* [s][N][C] stored stages or NULL                         */
    float *prim_next;      /* [N][C]                                                  */
    float *U_out;          /* [N][C] (last stage)                                     */
    const float *dt;       /* device time step; [n_meshes] for a batch                */
    int32_t stage, last, store_k, has_prev;
    int32_t bc_ready;      /* work->bc_part already holds the boundary sums of prim
                              (the previous stage's epilogue wrote them, into what
                              was then work->bc_part_next)                            */
    float b, b2;
    int32_t n_inc;
    int32_t inc_idx[10];
    float inc_coef[10];
    /* mode 1 = fixed-dt Adams-Bashforth 3 (one RHS per step, build-defined):
       U_out = U0 + dt/12 (23 f0 - 16 f_m1 + 5 f_m2), f0 written to f_out;
       n_hist < 2 substitutes f0 for the missing history (f_m1, then f_m2) */
    int32_t mode;
    int32_t n_hist;
    const float *f_m1;     /* [N][C]                                                  */
    const float *f_m2;     /* [N][C]                                                  */
    float *f_out;          /* [N][C]                                                  */
    /* [n_tile_list] distinct ids of 256-cell tiles, or NULL: with the tile-fused stage
       only these tiles are stepped, one workgroup each (the other tiles are neither read
       nor written, and the listed tiles spread over every XCD whatever their ids); a
       batch re-steps only the meshes whose cells lie in listed tiles.  Cells of other
       meshes that share a listed tile are computed from stale inputs and must be
       overwritten by the caller.  1 <= n_tile_list <= ceil(N / 256); ids outside
       [0, ceil(N / 256)) are skipped.  No reference counterpart. */
    const int32_t *tile_list;
    int32_t n_tile_list;
} FvmStage;

int fvm_rk_stage_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const FvmStage *st,
                     void *stream);

/* The same stage split into its launches, for per-kernel timing with stream events:
 * parts is a mask of 1 = boundary-mean pre-pass (only if !bc_ready), 2 = gradient +
 * limiter + boundary values, 4 = facet fluxes, 8 = divergence + stage update (the
 * three-kernel pipeline, also on a mesh with tile tables), or 16 = the tile-fused stage
 * (needs m->fz_tile).  Launching the masks 1, 2, 4, 8 in turn, or 1, 16, equals
 * fvm_rk_stage_f32.  No reference counterpart (measurement hook). */
int fvm_rk_stage_part_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const FvmStage *st,
                          int32_t parts, void *stream);

/* Boundary values, gradient and limiter of every cell of prim into work->gp (and
 * work->ubc), reading the previous records from work->gp_prev (or gp): the full
 * cell_grads state the reference keeps (facet_process.py:387-418), for callers of the
 * tile-fused stage, which keeps only boundary cells' records. */
int fvm_grad_f32(const FvmMeshDev *m, const FvmPhys *p, const FvmWork *w, const float *prim, int32_t has_prev,
                 void *stream);

/* LDS bytes one workgroup may use on the current device (the tile tables of the fused
 * stage must fit it).  No reference counterpart. */
int fvm_max_lds_per_workgroup(int32_t *out);

/* Adaptive.update_stepsize (integrators.py:42-62) from the partial norms the last
 * adaptive stage left in work->red: dt *= alpha + (1 - alpha) 0.9 E^(-1/order). */
int fvm_update_stepsize(const FvmWork *w, int32_t C, int32_t n_blocks, float atol, float rtol,
                        float mtol, float order, float alpha_gt1, float alpha_le1, float *dt,
                        void *stream);

/* state -> primitives (FluidConstitution.state_to_primative, fvm_equation.py:35-44). */
int fvm_state_to_prim_f32(const float *state, float *prim, int64_t N, int32_t dim, float inv_cv,
                          void *stream);

/* sum_c vol_c U_c per conserved variable, accumulated in fp64 (no reference
 * counterpart).  out[C] (device, fp64) is overwritten; deterministic. */
int fvm_conservation_f64(const float *state, const double *vol, int64_t N, int32_t C, double *out,
                         double *partial /* [256*C] */, void *stream);

/* fvm_conservation_f64 without the caller's workspace: the per-block partials live in
 * the library's own device array (16 KiB per GPU, held until the process ends), so the
 * call allocates nothing -- for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  C <= 8, else FVM_ERR_ARG.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other host threads) race. */
int fvm_conservation_f64_noalloc(const float *state, const double *vol, int64_t N, int32_t C, double *out,
                                 void *stream);

/* CFL time step (no reference counterpart; the reference takes a fixed cfg.dt):
 * dt = cfl * min_c h_c / (|V_c| + c_c), h_c = d vol_c / sum_f |n_f|.  out (device fp64). */
int fvm_cfl_dt(const float *prim, const double *h, int64_t N, int32_t dim, double gR, double cfl,
               double *out, double *partial /* [256] */, void *stream);

/* ---------------------------------------------------------------------------------
 * FP64 geometric precompute (FVMMesh2D/3D, fvm_mesh.py:131-261, 263-344, 367-582).
 * Vertices V [nv][dim] fp64; cells [N][dim+1] and facets [F][dim] int64 vertex ids.
 * --------------------------------------------------------------------------------- */

/* normals (2-D (dy, -dx) of v1 - v0; 3-D (v1 - v0) x (v2 - v0), |n| = 2 x area),
 * unit normals and midpoints, [F][dim] each. */
int fvm_geom_facets_f64(const double *V, int32_t dim, const int64_t *facets, int64_t F, double *normals,
                        double *normals_hat, double *midpoints, void *stream);

/* cell volumes [N] and centroids [N][dim] (vertex mean). */
int fvm_geom_cells_f64(const double *V, int32_t dim, const int64_t *cells, int64_t N, double *volumes,
                       double *centroids, void *stream);

/* disp[c][k] = midpoint(f) - centroid(c), signs[c][k] = sign(n_hat(f) . disp) for
 * f = cell_to_facet[c][k] ([N][dim+1]). */
int fvm_geom_disp_signs_f64(const double *centroids, const double *midpoints, const double *normals_hat,
                            int32_t dim, const int64_t *cell_to_facet, int64_t N, double *disp, int64_t *signs,
                            void *stream);

/* Weighted LSQ gradient matrices G [N][dim][dim+1]: offsets to the neighbour centroid
 * (other[c][k]) or, where is_bc[c][k], to the facet midpoint; weights |d|^(-1/4), x 7.5
 * on boundary facets; G = (sum w^2 d d^T)^(-1) [w^2 d_k]. */
int fvm_geom_lsq_f64(const double *centroids, const double *midpoints, int32_t dim, const int64_t *cell_to_facet,
                     const int64_t *other, const uint8_t *is_bc, int64_t N, double *G, void *stream);

/* ---------------------------------------------------------------------------------
 * Learned corrector on the mesh graph (no reference counterpart: the reference has no
 * neural network).  A message-passing network predicts the state a window of k solver
 * steps ahead; these are the kernels of its forward pass, backward pass and Adam update.
 * bf16 operands are passed as void* (2-byte bfloat16), fp32 accumulation throughout.
 * --------------------------------------------------------------------------------- */

/* z = A.B + bias[N] with A(i,k) = A[i*a_rs + k*a_cs], B(k,j) = B[k*b_rs + j*b_cs]
 * (any strides, so transposed operands need no copy); 64x64 tiles on
 * v_mfma_f32_16x16x32_bf16.  Cf (fp32, row stride c_rs) receives z; Cb (bf16, row stride
 * N) receives res + relu(z) (relu if relu != 0, res bf16 [M][N] if non-NULL).
 * splits > 1 splits K over workgroups into fp32 partials part[splits][M][N] that a
 * second launch sums in split order. */
int fvm_gemm_bf16(const void *A, int64_t a_rs, int64_t a_cs, const void *B, int64_t b_rs, int64_t b_cs,
                  int64_t M, int64_t N, int64_t K, int32_t splits, const float *bias, float *Cf, int64_t c_rs,
                  void *Cb, int32_t relu, const void *res, float *part, void *stream);

/* A[c] = [h[c] | mean of h over c's interior face neighbours] (bf16 [N][2H]); nbr is
 * FvmMeshDev.cell_nbr ([NL][N], >= N marks a boundary facet). H % 8 == 0. */
int fvm_mp_aggregate(const void *h, int32_t H, const int32_t *nbr, int64_t N, int32_t NL, void *A,
                     void *stream);

/* Backward of fvm_mp_aggregate (symmetric adjacency, so a gather):
 * dh[c] = dA[c][:H] + sum_{n in nbr(c)} dA[n][H:] / |nbr(n)|  (+ dres[c] if non-NULL). */
int fvm_mp_aggregate_bwd(const float *dA, int32_t H, const int32_t *nbr, int64_t N, int32_t NL,
                         const float *dres, float *dh, void *stream);

/* One message-passing layer with the aggregation fused into the GEMM:
 * h_out = h_in + relu([h_in | mean_nb h_in] W + bias), H = 128, NL = 3 or 4.  WT is W
 * transposed, bf16 [H][2H]; Z (fp32 [N][H], pre-activation) and A_out (bf16 [N][2H],
 * the aggregated operand) are written when non-NULL (training).  h_out != h_in. */
int fvm_mp_layer(const void *h_in, int32_t H, const int32_t *nbr, int64_t N, int32_t NL, const void *WT,
                 const float *bias, void *h_out, float *Z, void *A_out, void *stream);

/* WT[c][r] = bf16(W[r][c]) for fp32 W [nrow][ncol]. */
int fvm_transpose_bf16(const float *W, int64_t nrow, int64_t ncol, void *WT, void *stream);

/* h_out = h_in + relu(z) (bf16), or relu(z) when h_in is NULL. */
int fvm_relu_residual(const float *z, const void *h_in, void *h_out, int64_t n, void *stream);

/* dz = dh * (z > 0) as bf16 (dz_b) and, if dz_f is non-NULL, fp32. */
int fvm_relu_bwd(const float *dh, const float *z, void *dz_b, float *dz_f, int64_t n, void *stream);

/* out[j] = sum_i X[i][j] for X fp32 [M][N]; part is [256][N] scratch; fixed order. */
int fvm_colsum_f32(const float *X, int64_t M, int32_t N, float *out, float *part, void *stream);

/* Mean squared error over the first n_out columns of y ([M][ld_y]) against tgt
 * ([M][n_out]): dy = 2 (y - tgt) / (M n_out) (0 in the padding columns), and fp64
 * per-block sums of squares in part[256] (loss = sum(part) / (M n_out)). */
int fvm_mse_loss(const float *y, int64_t ld_y, const float *tgt, int64_t M, int32_t n_out, float *dy,
                 double *part, void *stream);

/* Adam with bias correction (step >= 1) on fp32 weights w; wb (bf16 copy) optional. */
int fvm_adam_f32(float *w, const float *g, float *m, float *v, void *wb, int64_t n, float lr, float b1,
                 float b2, float eps, int32_t step, void *stream);

int fvm_cast_f32_bf16(const float *x, void *y, int64_t n, void *stream);

/* Inference encoder of the corrector network: h0 = relu(X W_in + b_in), X bf16 [N][KX],
 * WT = W_in^T bf16 [H][KX], h0 bf16 [N][H]; KX = 32, H = 128.  No reference counterpart. */
int fvm_mp_encode(const void *X, int64_t N, int32_t KX, const void *WT, const float *bias, int32_t H, void *h0,
                  void *stream);

/* Inference message layer h_out = h_in + relu([h_in | mean_nb h_in] W + b) (as
 * fvm_mp_layer, without the training outputs).  With y non-NULL the layer is the last one
 * and its decoder is fused: y[c][j] = b_out[j] + sum_f bf16(h_out[c][f]) W_out[f][j]
 * (fp32 [N][n_out], n_out = 8, WoT = W_out^T bf16 [n_out][H]) and h_out is not written.
 * No reference counterpart. */
int fvm_mp_layer_inf(const void *h_in, int32_t H, const int32_t *nbr, int64_t N, int32_t NL, const void *WT,
                     const float *bias, void *h_out, const void *WoT, const float *b_out, int32_t n_out, float *y,
                     void *stream);

/* Network input of one window of the learned corrector (2-D batches: NL = 3, C = 4,
 * KX = 32), per cell c, bf16 row X[c]:
 *   [(prim - mu_p) / sd_p, W f0 / s1, W (f0 - fm) / s2, geo[c][0..2),
 *    h d/dx prim / sd_p, h d/dy prim / sd_p, h W d/dx f0 / s1, h W d/dy f0 / s1, 0, 0]
 * with W = win[c] (the window length of the cell's mesh), h = hc[c] (cell size), f0 the
 * RHS at the window start, fm the RHS half a window earlier, s1/s2 [n_meshes][C] scales
 * of mesh cell_mesh[c] (0 when cell_mesh is NULL), and gradients
 * sum_k wg[k][i][c] (v[nbr[k][c]] - v[c]) over the interior face neighbours (nbr >= N:
 * boundary, no term; wg [NL][2][N]).  prim, f0, fm and X rows must be 16-byte aligned.
 * No reference counterpart (the reference has no learned path). */
int fvm_mp_window_features(const float *prim, const float *f0, const float *fm, const float *win, const float *geo,
                           const float *hc, const int32_t *nbr, int32_t NL, const float *wg, int64_t N, int32_t C,
                           const float *mu_p, const float *sd_p, const float *s1, const float *s2,
                           const int32_t *cell_mesh, void *X, int32_t KX, void *stream);

/* Window update: U_out = U + win * (a f0 + b fm) + scale[m(c)] * y[:, :C] (fm and y may
 * be NULL; y fp32 with row stride ld_y, scale [n_meshes][C]). */
int fvm_mp_window_apply(const float *U, const float *f0, const float *fm, float a, float b, const float *win,
                        const float *y, int64_t ld_y, const float *scale, const int32_t *cell_mesh, int64_t N,
                        int32_t C, float *U_out, void *stream);

/* Both predicted states of a window: Um = U + win (am f0 + bm fm) + scale[m(c)] y[:, 0:C],
 * U1 = U + win (ae f0 + be fm) + scale[m(c)] y[:, C:2C] (y optional, fp32, row stride
 * ld_y >= 2C).  No reference counterpart. */
int fvm_mp_window_predict(const float *U, const float *f0, const float *fm, float am, float bm, float ae, float be,
                          const float *win, const float *y, int64_t ld_y, const float *scale,
                          const int32_t *cell_mesh, int64_t N, int32_t C, float *Um, float *U1, void *stream);

/* Per mesh and column: out1[mesh][j] = RMS over the mesh's cells of win[c] A[c][j] and
 * out2[mesh][j] = RMS of win[c] (A[c][j] - B[c][j]), each at least floor_abs (fp64 sums;
 * part is [n_meshes][64][2C] scratch).  No reference counterpart. */
int fvm_mesh_rms2(const float *A, const float *B, const float *win, const int64_t *off, int32_t n_meshes, int32_t C,
                  float floor_abs, float *out1, float *out2, double *part, void *stream);

/* Per mesh and column, out[mesh][j] = max(floor_abs, sqrt(mean over the mesh's cells of
 * (win[c] X[c][j])^2)) (win NULL = 1), fp64 sums; part is [n_meshes][64][C] scratch. */
int fvm_mesh_rms(const float *X, const float *win, const int64_t *off, int32_t n_meshes, int32_t C,
                 float floor_abs, float *out, double *part, void *stream);

/* Trapezoid defect of a predicted window, per mesh:
 * d = U1 - U0 - win/2 (f0 + f1); out[mesh][j] = sum d_j^2, out[mesh][C + j] = sum U1_j^2
 * (fp64, cells of each mesh in [off[mesh], off[mesh+1])); part is
 * [n_meshes][64][2C] fp64 scratch. */
int fvm_mp_defect(const float *U0, const float *U1, const float *f0, const float *f1, const float *win,
                  const int64_t *off, int32_t n_meshes, int32_t C, double *out, double *part, void *stream);

/* Simpson and Hermite-midpoint defects of a window predicted at its midpoint and end, per
 * mesh: d = U1 - U0 - win/6 (f0 + 4 fm + f1), e = Um - (U0 + U1)/2 - win/8 (f0 - f1);
 * out[mesh][j] = sum d_j^2, out[mesh][C + j] = sum e_j^2, out[mesh][2C + j] = sum U1_j^2
 * (fp64); part is [n_meshes][64][3C] scratch. */
int fvm_mp_defect3(const float *U0, const float *Um, const float *U1, const float *f0, const float *fm,
                   const float *f1, const float *win, const int64_t *off, int32_t n_meshes, int32_t C,
                   double *out, double *part, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FVM_HIP_H */
