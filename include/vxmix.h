/*
 * vxmix.h — C ABI of libvxmix.so, the gfx950 (MI355X) implementation of the
 * veros per-timestep mixing hot path:
 *   - the kbot-masked batched tridiagonal solve behind solve_implicit, and
 *   - the isoneutral triad-tensor stencils isoneutral_diffusion_pre /
 *     isoneutral_diffusion.
 *
 * Conventions shared by every entry point
 *   - All array pointers are DEVICE pointers (hipMalloc'd or equivalent).
 *   - 3-D fields are (x, y, z) row-major with z contiguous (stride 1); the
 *     x and y strides are given in ELEMENTS. Level k = 0 is the deepest level.
 *   - kbot is int32, 1-based index of the deepest wet level; 0 means land.
 *     A column's wet cells are the suffix k in [kbot-1, nz).
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Work is enqueued asynchronously; no entry point synchronises the host.
 *   - Return value: VX_OK (0) or a VX_ERR_* code; vx_last_error() returns a
 *     message for the calling thread's most recent failure. No entry point
 *     exits the process or throws across the ABI.
 *   - Entry points allocate nothing unless stated; workspaces are
 *     caller-provided.
 */
#ifndef VXMIX_H
#define VXMIX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VX_OK = 0,
    VX_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad shape/stride, kbot out of range */
    VX_ERR_UNSUPPORTED = 2,      /* e.g. nz above the compiled maximum */
    VX_ERR_LAUNCH = 3,           /* the HIP runtime reported an error */
};

/* Library version and limits. */
const char* vx_version(void);
int32_t vx_max_nz(void);
/* Message for this thread's most recent non-zero return. */
const char* vx_last_error(void);

/*
 * Batched kbot-masked tridiagonal solve, float64 throughout.
 *
 * Replaces solve_implicit(a, b, c, d, water_mask, edge_mask, b_edge, d_edge)
 * (veros/core/utilities.py:51-59) with its solver
 * solve_tridiagonal_numpy (veros/core/operators.py:60-77) /
 * solve_tridiagonal_jax + the tdma custom call (veros/core/operators.py:102-156,
 * veros/core/special/tdma_.py:49-68), i.e. the XLA targets tdma_cython_double
 * (veros/core/special/tdma_cython_.pyx:49-83) and tdma_cuda_double
 * (veros/core/special/cuda_tdma_kernels.cu:81-101).
 *
 * Per column (i, j) with kb = kbot[i*kbot_sx + j*kbot_sy] > 0, the rows
 * k = kb-1 .. nz-1 form one system: the sub-diagonal is taken as 0 at
 * k = kb-1 (the edge), the super-diagonal as 0 at k = nz-1; at the edge b and
 * d are replaced by b_edge / d_edge when those are non-NULL. x is written on
 * every level of every column, 0 outside the wet suffix. a, b, c, d, b_edge,
 * d_edge and x share the strides (sx, sy, 1) and shape (nx, ny, nz). x may
 * alias d (in-place solve).
 */
int vx_tdma_f64(void* stream,
                const double* a, const double* b, const double* c, const double* d,
                const double* b_edge, const double* d_edge,
                const int32_t* kbot, int64_t kbot_sx, int64_t kbot_sy,
                double* x,
                int64_t nx, int64_t ny, int32_t nz, int64_t sx, int64_t sy);

/*
 * Same system and layout as vx_tdma_f64, solved by float32 Thomas sweeps with
 * `refine_steps` (0..2) steps of iterative refinement whose residual
 * r = d - A x is formed in float64 and whose correction is accumulated into a
 * float64 solution. Inputs and output are float64.
 */
int vx_tdma_mixed(void* stream,
                  const double* a, const double* b, const double* c, const double* d,
                  const double* b_edge, const double* d_edge,
                  const int32_t* kbot, int64_t kbot_sx, int64_t kbot_sy,
                  double* x,
                  int64_t nx, int64_t ny, int32_t nz, int64_t sx, int64_t sy,
                  int32_t refine_steps);

/*
 * Reverse-mode derivative of vx_tdma_f64 (the pullback of
 * veros/core/utilities.py:51-59 solve_implicit, which the reference gets from
 * JAX's autodiff of its tridiagonal solve). Given the forward inputs, the
 * forward solution x and gx = dL/dx, solves the transposed system
 * A^T lam = gx per wet column and writes
 *   gd = lam, gb = -lam x, ga[k] = -lam[k] x[k-1], gc[k] = -lam[k] x[k+1]
 * on the wet rows (zero on dry rows, ga on the edge row and gc on the top
 * row). When b_edge (d_edge) was given, the edge row's derivative goes to
 * gb_edge (gd_edge) and gb (gd) is zero there; gb_edge/gd_edge must be
 * non-NULL exactly when b_edge/d_edge are, and are zero off the edge row.
 * All arrays share the layout of vx_tdma_f64. Deterministic (no atomics).
 */
int vx_tdma_adjoint_f64(void* stream,
                        const double* a, const double* b, const double* c,
                        const double* b_edge, const double* d_edge,
                        const int32_t* kbot, int64_t kbot_sx, int64_t kbot_sy,
                        const double* x, const double* gx,
                        double* ga, double* gb, double* gc, double* gd,
                        double* gb_edge, double* gd_edge,
                        int64_t nx, int64_t ny, int32_t nz, int64_t sx, int64_t sy);

/* ------------------------------------------------------------------------
 * Batched TDMA: many independent grids of different shapes and depths
 * (each with its own kbot), solved by one persistent launch per 16-level nz
 * bucket. The plan is built once (it allocates a small device table and keeps
 * the systems' pointers); run / graph_launch allocate nothing and may be
 * called every time step. No reference counterpart: veros solves one grid
 * per call (veros/core/utilities.py:51-59); this serves many grids or many
 * independent sub-domains per launch.
 * ------------------------------------------------------------------------ */
typedef struct vx_tdma_system {
    const double *a, *b, *c, *d;        /* (nx, ny, nz), strides (sx, sy, 1) */
    const double *b_edge, *d_edge;      /* nullable, same layout */
    const int32_t* kbot;                /* (nx, ny), strides (kbot_sx, kbot_sy) */
    int64_t kbot_sx, kbot_sy;
    double* x;                          /* output, same layout as a */
    int64_t nx, ny;
    int32_t nz;
    int32_t _pad;
    int64_t sx, sy;
} vx_tdma_system;

typedef struct vx_tdma_plan vx_tdma_plan;

/* mode 0: the float64 segmented solve of vx_tdma_f64; mode 1: the mixed
 * solve of vx_tdma_mixed with one float64 refinement step. Either way one
 * persistent launch per nz bucket. Validation errors leave *out NULL. */
int vx_tdma_plan_create(const vx_tdma_system* systems, int32_t nsys, int32_t mode, vx_tdma_plan** out);
/* Solve every system: one launch per non-empty nz bucket, on `stream`. */
int vx_tdma_plan_run(vx_tdma_plan* plan, void* stream);
/* Autotune: time (block size, blocks per CU) candidates per bucket on the
 * plan's data, keep the fastest; synchronises the stream. Choices are cached
 * process-wide for later plans of the same bucket, mode and size. Tuning
 * solves every system many times and writes x; a plan in which some x
 * overlaps an input (in-place solve) is rejected with
 * VX_ERR_INVALID_ARGUMENT, since its right-hand side would be overwritten. */
int vx_tdma_plan_tune(vx_tdma_plan* plan, void* stream);
/* First call captures vx_tdma_plan_run into a hipGraph; every call launches
 * that graph on `stream`. */
int vx_tdma_plan_graph_launch(vx_tdma_plan* plan, void* stream);
/* Buckets, total columns, and per bucket (nz bucket, block size, blocks per
 * CU or 0 for one block per 256 columns) into choices[3*q..3*q+2]. */
int vx_tdma_plan_info(const vx_tdma_plan* plan, int32_t* nbuckets, int64_t* total_columns,
                      int32_t* choices, int32_t max_choices);
void vx_tdma_plan_destroy(vx_tdma_plan* plan);

/* ------------------------------------------------------------------------
 * Isoneutral stencils. Arrays are full veros state arrays including the
 * 2-cell halos: X = nx + 4, Y = ny + 4, each C-contiguous.
 *   temp/salt/tracer/int_drhodX : (X, Y, nz, 3), time level innermost
 *       (grid->time_layout = VX_LEVELS_INNER, the reference's layout,
 *       veros/variables.py:338), or the three levels stored one after another
 *       (VX_LEVELS_OUTER: level l of cell (i, j, k) at l*X*Y*nz +
 *       (i*Y + j)*nz + k; the same arrays seen as the (X, Y, nz, 3) view with
 *       strides (Y*nz, nz, 1, X*Y*nz)). One layout for every time-level
 *       array of a call; the velocities u, v of the vertical friction and
 *       the energies tke, eke, E_iw follow it too.
 *   K_iso, K_11, K_22, K_33, dtr_iso, P_diss : (X, Y, nz)
 *   Ai_ez, Ai_nz, Ai_bx, Ai_by   : (X, Y, nz, 2, 2)
 *   kbot : (X, Y) int32;  dxt, dxu : (X);  dyt, dyu, cost, cosu : (Y);
 *   dzt, dzw : (nz)
 * Masks are not passed: maskT/U/V/W are derived from kbot in the kernels the
 * way calc_topo_kernel builds them (veros/core/numerics.py:195-221), which
 * requires the halo columns of kbot to be consistent (cyclic copy or 0).
 * ------------------------------------------------------------------------ */

#define VX_LEVELS_INNER 0
#define VX_LEVELS_OUTER 1

typedef struct vx_grid {
    const int32_t* kbot;
    const double *dxt, *dxu, *dyt, *dyu, *cost, *cosu, *dzt, *dzw;
    int64_t nx, ny;
    int32_t nz;
    int32_t time_layout;  /* VX_LEVELS_INNER (0) or VX_LEVELS_OUTER (1) */
} vx_grid;

typedef struct vx_iso_params {
    double drhodT, drhodS;   /* linear-EOS derivatives (eq_of_state_type 1) */
    double K_iso_steep, iso_slopec, iso_dslope;
    double dt_tracer, grav, rho_0;
} vx_iso_params;

/*
 * isoneutral_diffusion_pre (veros/core/isoneutral/isoneutral.py:18-229).
 * Reads temp/salt at time level `tau` and K_iso; writes Ai_ez, Ai_nz, Ai_bx,
 * Ai_by, K_11, K_22, K_33 in place over exactly the index ranges the
 * reference writes (all other entries keep their values), including
 * K_33[..., nz-1] = 0 over the whole plane.
 */
int vx_iso_pre_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                   const double* temp, const double* salt, int32_t tau,
                   const double* K_iso,
                   double* Ai_ez, double* Ai_nz, double* Ai_bx, double* Ai_by,
                   double* K_11, double* K_22, double* K_33);

/*
 * isoneutral_diffusion for one tracer (veros/core/isoneutral/diffusion.py:172-295,
 * iso=True), in place: tracer[..., taup1] receives the explicit flux
 * divergence times dt_tracer followed by the implicit K_33 vertical solve,
 * dtr_iso accumulates both tendencies, and, when P_diss is non-NULL
 * (enable_conserve_energy), P_diss accumulates the dissipation of dynamic
 * enthalpy computed from int_drhodX[..., tau] (veros/core/diffusion.py:9-62).
 * `workspace` must hold vx_iso_workspace_bytes(grid, 1) bytes (the face
 * fluxes and the wet-cell plan built from kbot); calls that share a workspace
 * must be ordered on one stream.
 */
int64_t vx_iso_workspace_bytes(const vx_grid* grid, int32_t ntracers);
int vx_iso_diffusion_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                         double* tracer, int32_t tau, int32_t taup1,
                         const double* K_iso,
                         const double* Ai_ez, const double* Ai_nz,
                         const double* Ai_bx, const double* Ai_by,
                         const double* K_11, const double* K_22, const double* K_33,
                         double* dtr_iso,
                         const double* int_drhodX, double* P_diss,
                         void* workspace);

/*
 * isoneutral_diffusion(temp) followed by isoneutral_diffusion(salt), as
 * thermodynamics calls them back to back (veros/core/thermodynamics.py:431-432),
 * sharing every read of the isoneutral tensor. Results equal two calls of
 * vx_iso_diffusion_f64 (temperature first), including the order in which
 * P_diss accumulates the two tracers' terms. `workspace` must hold
 * vx_iso_workspace_bytes(grid, 2) bytes.
 */
int vx_iso_diffusion_ts_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                            double* temp, double* salt, int32_t tau, int32_t taup1,
                            const double* K_iso,
                            const double* Ai_ez, const double* Ai_nz,
                            const double* Ai_bx, const double* Ai_by,
                            const double* K_11, const double* K_22, const double* K_33,
                            double* dtemp_iso, double* dsalt_iso,
                            const double* int_drhodT, const double* int_drhodS, double* P_diss,
                            void* workspace);

/*
 * The sequence veros/core/thermodynamics.py:430-432 runs each step,
 *   isoneutral_diffusion_pre, isoneutral_diffusion(temp), isoneutral_diffusion(salt),
 * in one call: the outputs of vx_iso_pre_f64 followed by
 * vx_iso_diffusion_ts_f64 on the same arrays. The face fluxes of temperature
 * and salinity are formed from the isoneutral tensor while it is computed,
 * instead of reading it back, and the tracer step runs over the wet cells
 * only. `workspace` holds vx_iso_pre_diffusion_ts_workspace_bytes(grid)
 * bytes (the two tracers' face fluxes and the wet-cell plan built from kbot;
 * the same size as vx_iso_workspace_bytes(grid, 2)).
 */
int64_t vx_iso_pre_diffusion_ts_workspace_bytes(const vx_grid* grid);

int vx_iso_pre_diffusion_ts_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                                double* temp, double* salt, int32_t tau, int32_t taup1,
                                const double* K_iso,
                                double* Ai_ez, double* Ai_nz, double* Ai_bx, double* Ai_by,
                                double* K_11, double* K_22, double* K_33,
                                double* dtemp_iso, double* dsalt_iso,
                                const double* int_drhodT, const double* int_drhodS, double* P_diss,
                                void* workspace);

/*
 * vx_iso_pre_diffusion_ts_f64 without a workspace argument. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The face fluxes and the wet-cell plan live in a fixed device
 * array of the library, 192 MiB per GPU, held from the library's load on
 * that GPU until the process ends; grids whose
 * vx_iso_pre_diffusion_ts_workspace_bytes exceed it are rejected with
 * VX_ERR_UNSUPPORTED. The call allocates nothing (no hipMalloc, no helper
 * stream: every launch is on `stream`). Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller. Results equal
 * vx_iso_pre_diffusion_ts_f64's.
 */
int vx_iso_pre_diffusion_ts_f64_noalloc(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                                        double* temp, double* salt, int32_t tau, int32_t taup1,
                                        const double* K_iso,
                                        double* Ai_ez, double* Ai_nz, double* Ai_bx, double* Ai_by,
                                        double* K_11, double* K_22, double* K_33,
                                        double* dtemp_iso, double* dsalt_iso,
                                        const double* int_drhodT, const double* int_drhodS, double* P_diss);

/* ------------------------------------------------------------------------
 * Implicit vertical integration of TKE, EKE and internal-wave energy: the
 * remaining solve_implicit callers (veros/core/tke.py:227, eke.py:157,
 * idemix.py:167), each with its coefficient build, dissipation diagnostic
 * and lateral diffusion in one call. The advection branches of the
 * reference kernels are not covered (callers reject them). Energies are
 * time-level arrays (grid->time_layout); 3-D fields (X, Y, nz); surface and
 * bottom fluxes (X, Y). The forcing is given as terms summed left to right,
 * forcing = sign[0]*arr[0] + sign[1]*arr[1] + ..., in the order the
 * reference's settings branches add them.
 * ------------------------------------------------------------------------ */
typedef struct vx_terms {
    const double* arr[12];
    double sign[12];
    int32_t n;
    int32_t _pad;
} vx_terms;

typedef struct vx_tke_params {
    double dt_mom, alpha_tke, c_eps, K_h_tke;
    int32_t hor_diffusion;  /* enable_tke_hor_diffusion */
    int32_t _pad;
} vx_tke_params;

/* integrate_tke_kernel (veros/core/tke.py:123-321): tke[..., taup1] on the
 * wet cells of the interior columns, then the surface correction
 * tke_surf_corr and max(0, .) of the top level (every interior column),
 * then the lateral diffusion (hor_diffusion); tke_diss over the whole array
 * from the solved, unclamped tke. */
int vx_integrate_tke_f64(void* stream, const vx_grid* grid, const vx_tke_params* prm, double* tke,
                         int32_t tau, int32_t taup1, const double* kappaM, const double* sqrttke,
                         const double* mxl, const double* forc_tke_surface, const vx_terms* forc,
                         double* tke_surf_corr, double* tke_diss);

typedef struct vx_eke_params {
    double dt_tracer, alpha_eke, eke_c_eps;
} vx_eke_params;

/* integrate_eke_kernel (veros/core/eke.py:93-255): eke[..., taup1], then
 * the lateral diffusion with 0.5*max(500, K_gm + K_gm'); eke_diss_iw =
 * c_int * eke[taup1] and eke_diss_tke = 0 over the whole array. */
int vx_integrate_eke_f64(void* stream, const vx_grid* grid, const vx_eke_params* prm, double* eke,
                         int32_t tau, int32_t taup1, const double* kappaM, const double* sqrteke,
                         const double* eke_len, const double* K_gm, const vx_terms* forc,
                         double* eke_diss_iw, double* eke_diss_tke);

typedef struct vx_idemix_params {
    double dt_tracer, tau_v, tau_h, eke_diss_surfbot_frac;
    int32_t hor_diffusion;   /* enable_idemix_hor_diffusion */
    int32_t eke_diss_mode;   /* 0: forc as given; 1: its column integral at the
                                bottom cell (enable_eke_diss_bottom); 2: split
                                between bottom and surface (enable_eke_diss_surfbot) */
} vx_idemix_params;

/* integrate_idemix_kernel (veros/core/idemix.py:47-265): E_iw[..., taup1]
 * with the edge right-hand side d_edge = d + dt*forc_iw_bottom/dzw, then the
 * lateral diffusion (hor_diffusion); iw_diss = alpha_c * max(0, E_iw[tau]) *
 * E_iw[taup1] over the whole array. `forc` is moved per eke_diss_mode, then
 * `forc_post` (nullable: bottom friction) is added. */
int vx_integrate_idemix_f64(void* stream, const vx_grid* grid, const vx_idemix_params* prm, double* E_iw,
                            int32_t tau, int32_t taup1, const double* c0, const double* v0,
                            const double* alpha_c, const double* forc_iw_bottom,
                            const double* forc_iw_surface, const vx_terms* forc,
                            const vx_terms* forc_post, double* iw_diss);

/*
 * enforce_boundaries (veros/core/utilities.py:7-21) on one process: with
 * enable_cyclic_x, arr[-2:] = arr[2:4] and arr[:2] = arr[-4:-2] along the
 * first axis (X = nx + 4 >= 8); element r of x row i at
 * arr[i*row_stride + r*elem_stride] for r < row_elems (e.g. a contiguous
 * (X, Y, nz) array: row_elems = row_stride = Y*nz, elem_stride = 1; level l
 * of a levels-innermost (X, Y, nz, 3) array: pass arr + l, row_stride =
 * 3*Y*nz, elem_stride = 3). Without cyclic x there is nothing to exchange.
 */
int vx_enforce_boundaries_f64(void* stream, double* arr, int64_t X, int64_t row_elems, int64_t row_stride,
                              int64_t elem_stride, int32_t enable_cyclic_x);

/*
 * isoneutral_diag_streamfunction_kernel (veros/core/isoneutral/isoneutral.py:232-255):
 * B2_gm[1:-2, 2:-2] = 0.25 * diffloc_e * sum(Ai_ez) and
 * B1_gm[2:-2, 1:-2] = -0.25 * diffloc_n * sum(Ai_nz), diffloc the 4-point
 * average of K_gm (edge-padded in z) around the face; other entries keep
 * their values. All arrays (X, Y, nz[, 2, 2]).
 */
int vx_iso_diag_streamfunction_f64(void* stream, const vx_grid* grid, const double* K_gm, const double* Ai_ez,
                                   const double* Ai_nz, double* B1_gm, double* B2_gm);

/*
 * vertmix_tempsalt (veros/core/thermodynamics.py:248-300): implicit vertical
 * mixing of temperature and salinity with diffusivity kappaH (X, Y, nz) and
 * the surface forcing forc_temp_surface / forc_salt_surface (X, Y) on the top
 * level, in place on temp/salt[..., taup1] over the wet cells of the interior
 * columns (grid->kbot, dzt, dzw are read; the other grid arrays may be NULL).
 * dtemp_vmix / dsalt_vmix (X, Y, nz) are overwritten everywhere with
 * (new - old) / dt_tracer (zero off the wet cells). With enable_cyclic_x the
 * x halos of level taup1 are then copied as enforce_boundaries does on one
 * process (veros/core/utilities.py:7-21; needs nx >= 4).
 */
int vx_vertmix_tempsalt_f64(void* stream, const vx_grid* grid, double dt_tracer,
                            double* temp, double* salt, int32_t taup1,
                            const double* kappaH,
                            const double* forc_temp_surface, const double* forc_salt_surface,
                            double* dtemp_vmix, double* dsalt_vmix, int32_t enable_cyclic_x);

/*
 * implicit_vert_friction (veros/core/friction.py:90-205): implicit vertical
 * friction of u and v (X, Y, nz, 3) with viscosity kappaM (X, Y, nz) averaged
 * onto the faces. Over the face columns [1, X-2) x [1, Y-2): the solve of
 * level tau goes in place into u/v[..., taup1] on the levels k >= max(kbot,
 * kbot_next) - 1 (a face with one dry side keeps tau's value there), and
 * du_mix / dv_mix are set to (taup1 - tau) / dt_mom on every level; outside
 * that range they keep their values. The dissipation of the solve,
 * interpolated onto the T grid with dxu / dxt (u) and area_v / area_t (v)
 * (veros/core/numerics.py:313-336), is added to K_diss_v. grid->kbot, dzt,
 * dzw, dxt and dxu are read; the other grid arrays may be NULL. tau and taup1
 * must differ. The fastest path wants the levels outermost, nz even and
 * every array 16-byte aligned.
 */
int vx_implicit_vert_friction_f64(void* stream, const vx_grid* grid, double dt_mom,
                                  double* u, double* v, int32_t tau, int32_t taup1,
                                  const double* kappaM, const double* area_t, const double* area_v,
                                  double* du_mix, double* dv_mix, double* K_diss_v);

/*
 * isoneutral_skew_diffusion for one tracer (veros/core/isoneutral/diffusion.py:172-207,
 * 210-283, 298-307, iso=False): the same flux stencil with K_iso = 0 and
 * K_skew = K_gm, the explicit update of tracer[..., taup1] and dtr_iso, no
 * implicit part; P_diss_skew (when non-NULL, with int_drhodX) accumulates the
 * lateral dissipation and the explicit vertical term. `workspace` as for
 * vx_iso_diffusion_f64.
 */
int vx_iso_skew_diffusion_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                              double* tracer, int32_t tau, int32_t taup1,
                              const double* K_gm,
                              const double* Ai_ez, const double* Ai_nz,
                              const double* Ai_bx, const double* Ai_by,
                              const double* K_11, const double* K_22,
                              double* dtr_iso,
                              const double* int_drhodX, double* P_diss_skew,
                              void* workspace);

/*
 * Reverse-mode derivative of one vx_iso_diffusion_f64 step (the JAX pullback
 * of veros/core/isoneutral/diffusion.py:172-207 with iso=True and the energy
 * branch off) with respect to the tracer and K_iso, holding the isoneutral
 * tensor (Ai_*, K_11, K_22, K_33) fixed. The forward step maps
 * (T[tau], T[taup1], K_iso) to T_new = tracer[..., taup1] after the call.
 * `tracer` holds the forward INPUT (level tau is read). Given g_new = dL/dT_new
 * (X, Y, nz), writes dL/dT[tau] -> g_tau, dL/dT[taup1] -> g_taup1 and
 * dL/dK_iso -> g_K_iso, each (X, Y, nz) and fully overwritten.
 * `workspace` must hold vx_iso_adjoint_workspace_bytes(grid) bytes. The
 * scatter into g_tau and g_K_iso uses float64 atomics: results agree run to
 * run to rounding, not bitwise.
 */
int64_t vx_iso_adjoint_workspace_bytes(const vx_grid* grid);
int vx_iso_diffusion_adjoint_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                                 const double* tracer, int32_t tau,
                                 const double* K_iso,
                                 const double* Ai_ez, const double* Ai_nz,
                                 const double* Ai_bx, const double* Ai_by,
                                 const double* K_11, const double* K_22, const double* K_33,
                                 const double* g_new,
                                 double* g_tau, double* g_taup1, double* g_K_iso,
                                 void* workspace);

#ifdef __cplusplus
}
#endif

#endif /* VXMIX_H */
