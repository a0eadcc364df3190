/*
 * swm_gpu.h -- C ABI of the MI355X (gfx950) shallow-water dynamical-core path.
 *
 * The reference model (vvshashkin/cubedsphere_sbp_swm, Fortran 2008 + MPI) has no C
 * interface; its hot path is reached through Fortran type-bound procedures.  Each entry
 * point below replaces one of them (file:line relative to the reference's src/) and is
 * bound from Fortran with ISO_C_BINDING (cubedsphere_sbp_swm_amd/fortran/swm_gpu_mod.f90)
 * or from Python with ctypes (cubedsphere_sbp_swm_amd/__init__.py).
 *
 * Conventions
 *  - Every function returns SWM_OK (0) or a negative status; swm_last_error() gives the
 *    message of the last failure on the calling thread (the reference aborts through
 *    parcomm_global%abort; the Fortran binding turns a negative status into that abort).
 *  - Device state vectors are opaque device pointers holding h, u, v of all six panels:
 *    3 fields x field_elems elements of the context precision (8 or 4 bytes).
 *  - Host arrays use the reference tile layout p(1-hw:n+hw, 1-hw:n+hw) per panel (i
 *    fastest), panels 1..6 consecutive, FP64, n = N+1 Ah points per panel side.
 *  - Host pointers are borrowed for the duration of a call; device buffers allocated by
 *    swm_state_alloc are owned by the caller until swm_state_free.
 *  - One HIP stream per context; calls on one context are not thread-safe.
 */
#ifndef SWM_GPU_H
#define SWM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define SWM_OK 0
#define SWM_EARG (-1)    /* invalid argument or shape */
#define SWM_ENAME (-2)   /* unknown or unsupported operator name */
#define SWM_EHIP (-3)    /* HIP runtime error */
#define SWM_ESTATE (-4)  /* object in the wrong state (e.g. wrong context) */
#define SWM_EIO (-5)     /* output file cannot be opened or written */

typedef struct swm_ctx swm_ctx;

const char *swm_last_error(void);

/* create_domain + create_swm_operator (+ create_timescheme 'rk4'):
 *   domain/domain_factory_mod.f90:17-136, models/shallow_water/operator/operator_swm_factory_mod.f90:43-131,
 *   time_schemes/timescheme_factory_mod.f90:42-59.
 * Operator names as in the &shallow_water_model namelist (config_swm_mod.f90:70-79):
 *   div  "divergence_ah42_sbp" | "divergence_ah43_sbp" | "divergence_ah63_sbp" | "divergence_ah2"
 *   grad "gradient_ah42_sbp_ecs" | "gradient_ah43_sbp_ecs" | "gradient_ah63_sbp_ecs" | "gradient_ah21_sbp_ecs"
 *   curl "curl_" + div name
 *   quadrature "SBP_Ah21_quadrature" | "SBP_Ah42_quadrature" | "SBP_Ah63_quadrature"
 * scale = sphere radius, omega = rotation rate (axis z), grav = gravity.
 * precision_bits = 64 (FP64 state and stencils) or 32 (FP32 state and stencils, metric
 * precomputed in FP64, compensated FP64 integrals). */
int swm_ctx_create(swm_ctx **ctx, int N, const char *div_op_name, const char *grad_op_name,
                   const char *curl_op_name, const char *quadrature_name, double scale,
                   double omega, double grav, int precision_bits);
int swm_ctx_destroy(swm_ctx *ctx);
/* n = N+1, row pitch, elements per field, bytes per element */
int swm_ctx_info(const swm_ctx *ctx, int *n, int *pitch, long long *field_elems, int *elem_bytes);
/* run subsequent work of this context on a caller-provided hipStream_t (NULL = own stream) */
int swm_ctx_set_stream(swm_ctx *ctx, void *hip_stream);
int swm_ctx_synchronize(swm_ctx *ctx);

/* create_stvec_swm (models/shallow_water/stvec/stvec_swm_factory_mod.f90:7-28) and
 * stvec_t%create_similar (stvec_swm_mod.f90:23-36): device h,u,v */
int swm_state_alloc(swm_ctx *ctx, void **state);
int swm_state_free(swm_ctx *ctx, void *state);
/* host <-> device copies of h,u,v (reference tile layout with halo width halo_width) */
int swm_state_upload(swm_ctx *ctx, void *state, const double *h, const double *u,
                     const double *v, int halo_width);
int swm_state_download(swm_ctx *ctx, const void *state, double *h, double *u, double *v,
                       int halo_width);

/* operator_t%apply (operator_mod.f90:11-25; operator_swm_mod.f90:61-113): vout = RHS(vin) */
int swm_rhs(swm_ctx *ctx, const void *vin, void *vout);

/* stvec_t%assign_s1v1s2v2 / update_s1v1s2v2 (stvec_swm_mod.f90:38-80; grid_field_mod.f90:430-485):
 * out = s1*v1 + s2*v2  /  out = out + s1*v1 + s2*v2, mesh points of all three fields */
int swm_assign(swm_ctx *ctx, void *out, double s1, const void *v1, double s2, const void *v2);
int swm_update(swm_ctx *ctx, void *out, double s1, const void *v1, double s2, const void *v2);

/* timescheme_t%step for rk4 (time_schemes/rk4_mod.f90:20-42), nsteps times, in place,
 * one fused kernel per RK stage (4 launches per step, captured in a HIP graph) */
int swm_rk4_steps(swm_ctx *ctx, void *state, double dt, int nsteps);

/* operator_t%get_diagnostics (operator_swm_mod.f90:115-260):
 * out = {hmin, hmax, mass, TE, KE, PE, enstrophy}; integrals use the context's quadrature
 * with compensated (Neumaier) FP64 summation */
int swm_diagnostics(swm_ctx *ctx, const void *state, double out[7]);
/* No counterpart in the reference: swm_diagnostics for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops).  Same kernels and the same seven values, written to
 * the device array `out` (7 doubles) in stream order on the context stream, with no allocation
 * and no synchronisation: the partial sums live in a library-owned device scratch of 128.1 KiB
 * per GPU (at most 1024 blocks of partial sums; a context asking for more is rejected) that
 * every call on that GPU shares, so calls that overlap on one GPU race. */
int swm_diagnostics_noalloc(swm_ctx *ctx, const void *state, double *out);

/* orography h_surf (operator_swm_factory_mod.f90:115-116); NULL removes it */
int swm_set_orography(swm_ctx *ctx, const double *hs, int halo_width);

/* Seeded spherical-harmonic initial state (the reference has no seeded generator; this is
 * the one used for the benchmark configurations): h = h0 + h' with max|h'| = h_amp,
 * V = r x grad(psi) with max|V| = v_max, degrees 1..lmax with N(0,1)*l^-2 coefficients;
 * covariant u = V.a1, v = V.a2 as set_vector_test_field (test_fields/test_fields_mod.f90:397-433).
 * Output in the host reference layout with halo width halo_width (halo left zero). */
int swm_ic_spherical_harmonics(swm_ctx *ctx, unsigned long long seed, int lmax, double h0,
                               double h_amp, double v_max, double *h, double *u, double *v,
                               int halo_width);

/* ---- Hyperdiffusion (FP64; N+1 >= 12 for the sbp42 family, >= 18 for sbp63) ----
 * swm_laplace: laplace_Ah_sbp21_narrow_t%calc_laplace (differential_operators/horizontal/
 * laplace/laplace_Ah_sbp21_narrow_mod.f90:22-56), laplace_Ah_sbp42_narrow_t%calc_laplace
 * (laplace_Ah_sbp42_narrow_mod.f90:23-58) or laplace_Ah_sbp63_narrow_t%calc_laplace
 * (laplace_Ah_sbp63_narrow_mod.f90:494-529), picked by the name create_laplace_operator takes
 * (laplace_factory_mod.f90:11-33): "laplace_ah_sbp21_narrow", "laplace_ah_sbp42_narrow" or
 * "laplace_ah_sbp63_narrow" (other names return SWM_ENAME).  out = narrow-stencil SBP Laplacian of the field f (one field
 * of the context layout: field_elems elements; out distinct from f), edge-synced.
 * swm_diff_apply: operator_swm_diff_t%apply (models/shallow_water/operator/
 * operator_swm_diff_mod.f90:25-46) as built by create_swm_diff_operator
 * (operator_swm_diff_factory_mod.f90:11-34): vout = hyperdiffusion tendency of vin (distinct
 * state vectors).  Names as in the &shallow_water_model namelist (hordiff_factory_mod.f90:
 * 11-51): hordiff_uv_name "hordiff_vec_xyz_Ah_sbp_{21,42,63}_narrow", hordiff_h_name
 * "hordiff_scalar_Ah_sbp_{21,42,63}_narrow", each independently;
 * other names return SWM_ENAME.
 * swm_diff_step: timescheme_diff%step with diff_time_scheme "explicit_Eul1"
 * (time_schemes/explicit_Eul1_mod.f90:20-30) on that operator: state += dt * diff(state),
 * as the drivers call it after each RK4 step (ts2_mod.f90:145-146). */
int swm_laplace(swm_ctx *ctx, const char *laplace_name, const void *f, void *out);
int swm_diff_apply(swm_ctx *ctx, const void *vin, void *vout, const char *hordiff_uv_name,
                   double uv_diff_coeff, const char *hordiff_h_name, double h_diff_coeff);
int swm_diff_step(swm_ctx *ctx, void *state, double dt, const char *hordiff_uv_name,
                  double uv_diff_coeff, const char *hordiff_h_name, double h_diff_coeff);

/* ---- Output in the reference's on-disk formats ----
 * Records are REAL*4, written as direct-access record rec_num (1-based) of a file that is
 * created if absent and otherwise kept (other records unchanged); a record of W values starts at
 * byte (rec_num-1) * 4 * W (record lengths in 4-byte words, as the reference's Intel Fortran
 * build writes them).  Field pointers are one field of the context layout (field_elems
 * elements), e.g. the h, u or v plane of a state vector.
 * swm_latlon_outputer_create: create_latlon_outputer(outputer, Nlat, Nlon, scalar_grid_type,
 * domain) (outputer/outputer_factory_mod.f90:45-136) with scalar_grid_type "Ah" (others return
 * SWM_ENAME): cubic regrid to the Nlat x Nlon lat-lon grid (lat from -pi/2 to pi/2 inclusive,
 * lon from 0 in steps of 2 pi/Nlon; regrid_factory_mod.f90:9-60,183-301).
 * swm_latlon_vec_outputer_create: create_latlon_vec_outputer(outputer, Nlat, Nlon, staggering,
 * components_type, domain) (outputer_factory_mod.f90:138-190) with staggering "Ah" and
 * components_type "covariant".
 * swm_latlon_write: latlon_outputer_t%write(f, domain, file_name, rec_num)
 * (latlon_outputer_mod.f90:44-98): one record of Nlon x Nlat values, lon fastest.
 * swm_latlon_vec_write: latlon_vec_outputer_t%write(u, v, domain, file_name_u, file_name_v,
 * rec_num) (latlon_outputer_mod.f90:100-190): the zonal and meridional components.
 * swm_paneled_write: master_paneled_outputer_t%write(f, domain, file_name, rec_num) of an
 * outputer made by create_master_paneled_outputer(outputer, "xy", domain)
 * (master_paneled_outputer_mod.f90:24-77): one record of (N+1) x (N+1) x 6 values, i fastest,
 * then j, then panel. */
typedef struct swm_outputer swm_outputer;
int swm_latlon_outputer_create(swm_ctx *ctx, int Nlat, int Nlon, const char *scalar_grid_type,
                               swm_outputer **out);
int swm_latlon_vec_outputer_create(swm_ctx *ctx, int Nlat, int Nlon, const char *staggering,
                                   const char *components_type, swm_outputer **out);
int swm_outputer_destroy(swm_outputer *outputer);
int swm_latlon_write(swm_outputer *outputer, const void *f, const char *file_name, int rec_num);
int swm_latlon_vec_write(swm_outputer *outputer, const void *u, const void *v, const char *file_name_u,
                         const char *file_name_v, int rec_num);
int swm_paneled_write(swm_ctx *ctx, const void *f, const char *file_name, int rec_num);

/* Tile configuration of the fused stage kernel (no counterpart in the reference): the tile shape
 * in points, tile_cols x tile_rows = 64 x 8, 32 x 16, 16 x 32 (512 threads) or 32 x 8 (256
 * threads; the row count must be below N+1; the sbp21 kernel uses 64 x 16 tiles for 16 rows and
 * 64 x 8 otherwise), and, for FP64, whether the metric terms J, Qi are read from memory (0) or
 * evaluated on the fly from tan(alpha) (1; identical values).  swm_ctx_set_tiling /
 * swm_ctx_get_tiling name the 512-thread shapes by their row count (8, 16, 32).
 * swm_ctx_autotune times `reps` RK4 steps on `state` (the state is left unchanged) for every
 * configuration, in two interleaved rounds, keeps the fastest (minimum over the rounds) and
 * returns its time per step. */
int swm_ctx_set_tile_shape(swm_ctx *ctx, int tile_cols, int tile_rows, int metric_on_the_fly);
int swm_ctx_get_tile_shape(const swm_ctx *ctx, int *tile_cols, int *tile_rows, int *metric_on_the_fly);
int swm_ctx_set_tiling(swm_ctx *ctx, int tile_rows, int metric_on_the_fly);
int swm_ctx_get_tiling(const swm_ctx *ctx, int *tile_rows, int *metric_on_the_fly);
int swm_ctx_autotune(swm_ctx *ctx, void *state, int reps, double *best_us);

/* Measurement helper (no counterpart in the reference): `reps` back-to-back launches of one
 * fused RK stage kernel (stage_id 1, 2 or 4) on the context stream, with the buffers the RK4 step
 * gives it (stage 2: y = a work copy of `state`, v0 = `state`); outputs go to internal work
 * buffers, `state` is unchanged. */
int swm_bench_stage(swm_ctx *ctx, void *state, int stage_id, int reps);

/* ---- Adjoint form (no counterpart in the reference, which has no adjoint) ----
 * FP64 contexts with the SBP families 42/43/63, N+1 >= 15 (42, 43) or >= 23 (63).
 *
 * swm_rhs_vjp: out = (dF/dy)^T mu, the transpose of the linearised right-hand side of
 * operator_swm_t%apply (operator_swm_mod.f90:61-113) at state y, applied to mu (all three are
 * state vectors of `ctx`, out distinct from y and mu).
 *
 * swm_adj_*: reverse-mode differentiation of `nsteps` RK4 steps (rk4_mod.f90:20-42) of size dt.
 * swm_adj_forward runs the steps from y0, keeps the state every `checkpoint_every` steps and
 * writes the final state to yT; swm_adj_backward takes lamT = dJ/dy(T) and writes
 * lam0 = dJ/dy(0), recomputing each segment of states from its checkpoint and each step's
 * stage inputs.  The forward records the last segment's stage inputs as it runs them, so the
 * first backward after it recomputes one segment fewer; backward may be called again with
 * another lamT.  The tape owns its buffers until swm_adj_destroy: nsteps/checkpoint_every
 * checkpoints plus, for one segment, checkpoint_every x (4 state vectors + 4 fields). */
typedef struct swm_adj swm_adj;
int swm_rhs_vjp(swm_ctx *ctx, const void *y, const void *mu, void *out);
int swm_adj_create(swm_ctx *ctx, int nsteps, double dt, int checkpoint_every, swm_adj **tape);
int swm_adj_destroy(swm_adj *tape);
int swm_adj_forward(swm_adj *tape, const void *y0, void *yT);
int swm_adj_backward(swm_adj *tape, const void *lamT, void *lam0);
/* Launch configuration of the transposed RHS (no counterpart in the reference): tile_x = 16
 * or 32 runs it as one fused launch per evaluation on tiles of tile_x x 16 points (16 is the
 * default); 0 runs it as two launches (seeds and edge syncs, then D^T).  All give the same
 * bits. */
int swm_adj_set_tiling(swm_ctx *ctx, int tile_x);

/* ---- Batched ensemble form (no counterpart in the reference) ----
 * Cubes of different resolutions, one context and one state vector each (same precision and
 * SBP family 42/43/63, N >= 16), advanced together: every RK4 stage of the whole ensemble is
 * one kernel launch over all members' tiles, and one ensemble step (4 launches) and a run of
 * steps are captured once in hipGraphs.  dt holds one time step per member.  The ensemble runs
 * on the first member's stream; the states stay owned by the caller.  The launch's blocks walk
 * the task list of all members: blocks_per_slot = 1..64 sizes a persistent grid of that many
 * blocks per resident block slot, 0 launches one block per task (the hardware dispatcher
 * hands out the tasks).  swm_ens_autotune times the stage-2 launch over the tile shapes of
 * swm_ctx_set_tile_shape, the metric source (FP64: read or on the fly) and the grid (1 or 4
 * blocks per resident slot, or one block per task) and keeps the fastest; swm_ens_bench_stage
 * runs `reps` stage-2 launches (measurement helper). */
typedef struct swm_ens swm_ens;
int swm_ens_create(swm_ctx *const *members, void *const *states, int nmembers, swm_ens **ens);
int swm_ens_destroy(swm_ens *ens);
int swm_ens_rk4_steps(swm_ens *ens, const double *dt, int nsteps);
int swm_ens_set_tile_shape(swm_ens *ens, int tile_cols, int tile_rows, int metric_on_the_fly, int blocks_per_slot);
int swm_ens_get_tile_shape(const swm_ens *ens, int *tile_cols, int *tile_rows, int *metric_on_the_fly,
                           int *blocks_per_slot);
int swm_ens_set_tiling(swm_ens *ens, int tile_rows, int metric_on_the_fly, int blocks_per_slot);
int swm_ens_get_tiling(const swm_ens *ens, int *tile_rows, int *metric_on_the_fly, int *blocks_per_slot);
int swm_ens_autotune(swm_ens *ens, int reps, double *best_us);
int swm_ens_bench_stage(swm_ens *ens, int reps);

/* ECS cubic halo on Ah points: ecs_halo_xy_t%get_halo_scalar (equiang_cs/ecs_halo_xy_mod.f90:68-219)
 * with the weights of init_halo_xy_interp (equiang_cs/ecs_halo_factory_mod.f90:228-269).  `field`
 * is one field of the context (e.g. a state vector's h, u or v plane: state + k*field_elems
 * elements); the panel extended by halo_width (1..8, N >= halo_width + 3) rows of virtual points
 * beyond every edge is written to `out` as 6 x (n+2hw) x (n+2hw) elements of the context
 * precision, [panel][j][i] with 1-based i, j from 1-hw to n+hw, including the averaged
 * duplicated edge lines and cube corners.  swm_ecs_halo_xy writes device memory,
 * swm_ecs_halo_xy_host FP64 host memory. */
int swm_ecs_halo_xy(swm_ctx *ctx, const void *field, int halo_width, void *out);
int swm_ecs_halo_xy_host(swm_ctx *ctx, const void *field, int halo_width, double *out);

/* ECS cubic halo on O points (A-grid cell centres): ecs_halo_t%get_halo_scalar(f, domain,
 * halo_width) of the halo procedure create_halo_procedure(halo, domain, halo_width, "ECS_O")
 * (equiang_cs/ecs_halo_mod.f90:71-166, ecs_halo_factory_mod.f90:24-44,116-226), including its
 * width-8 'full' O-point exchange.  field_o: N x N x 6 cell values of the context precision
 * (i fastest, then j, then panel); out: (N+2hw) x (N+2hw) x 6, i, j from 1-hw to N+hw.
 * 1 <= halo_width <= 8, N >= 8.  Edge strips and the cube-corner special points are computed as
 * the reference computes them (weights as for a halo object created with width max(hw, 2); the
 * corner step reads the second row).  The other points of the halo corner squares, which the
 * reference leaves as they were (only for hw >= 3), are NaN.  swm_ecs_halo_o takes and writes
 * device memory; swm_ecs_halo_o_host FP64 host memory. */
int swm_ecs_halo_o(swm_ctx *ctx, const void *field_o, int halo_width, void *out);
int swm_ecs_halo_o_host(swm_ctx *ctx, const double *field_o, int halo_width, double *out);

/* ECS vector halo on O points (A grid, contravariant components): halo_vec_t%get_halo_vector(u, v,
 * domain, halo_width) of create_ecs_A_vec_halo_procedure (equiang_cs/ecs_halo_vec_a_factory_mod.f90:
 * 20-175, ecs_halo_vec_a_mod.f90:53-192), with its width-8 O-point vector exchange.  u_o, v_o and
 * out_u, out_v laid out as field_o / out of swm_ecs_halo_o, the same widths and NaN points.
 * Currently switched off on the device: returns SWM_ESTATE (an unresolved GPU memory fault in its
 * first run); the algorithm is restated and tested in the oracle. */
int swm_ecs_halo_vec_a(swm_ctx *ctx, const void *u_o, const void *v_o, int halo_width, void *out_u, void *out_v);
int swm_ecs_halo_vec_a_host(swm_ctx *ctx, const double *u_o, const double *v_o, int halo_width, double *out_u,
                            double *out_v);

/* Device-to-device copies between a state vector and three dense device arrays of shape
 * (6, n, n) (panel, j, i) in the context precision (e.g. the storage of framework tensors); no
 * counterpart in the reference. */
int swm_state_copy_from_device(swm_ctx *ctx, void *state, const void *h, const void *u, const void *v);
int swm_state_copy_to_device(swm_ctx *ctx, const void *state, void *h, void *u, void *v);

/* Diagnostics of the set-up (not part of the reference interface): device copies of the
 * metric planes (which: 0 J, 1 Qi11, 2 Qi12, 3 Qi22, 4 f) in host layout, and the cube-edge
 * neighbour table (6*4*n*2*4 ints, see geometry.h). */
int swm_debug_geometry(swm_ctx *ctx, int which, double *out, int halo_width);
int swm_debug_halo_table(swm_ctx *ctx, int *table);

/* ---- C grid (staggering 'C'), SBP42 vector-invariant family, FP64 ----
 * h at cell centres (N x N per panel), u at x faces (N+1 x N), v at y faces (N x N+1).  Host
 * arrays use the layout of the Ah grid of the same N: (6, N+1+2hw, N+1+2hw), point (i, j) of any
 * type at [p][j-1+hw][i-1+hw]; points outside a type's range are ignored on upload and zero on
 * download.  Device state vectors hold h, u, v as 3 fields of field_elems doubles. */
typedef struct swm_cctx swm_cctx;

/* create_domain(staggering 'C') + create_swm_operator with the &shallow_water_model_c names
 * (namelist_swm:80-102; create_vector_invariant_swm_operator,
 * models/shallow_water/operator/operator_swm_factory_mod.f90:43-120):
 * div "divergence_c_sbp42" (div_factory_mod.f90:67-87), grad "gradient_c_sbp42"
 * (grad_factory_mod.f90:73-87), curl "curl_c_sbp42" (curl_factory_mod.f90:57-88), coriolis
 * "coriolis_Cgrid_sbp42" (coriolis_factory_mod.f90:45-84), KE "KE_Cgrid_sbp42"
 * (KE_factory_mod.f90:33-50), co2contra "co2contra_c_sbp42" (co2contra_factory_mod.f90:39-65),
 * massflux "massflux_c_sbp42" (massflux_factory_mod.f90:56-61), quadrature "SBP_C42_quadrature";
 * other names return SWM_ENAME.  8 <= N <= 8192. */
int swm_c_create(swm_cctx **out, int N, const char *div_op_name, const char *grad_op_name,
                 const char *curl_op_name, const char *coriolis_op_name, const char *KE_op_name,
                 const char *co2contra_op_name, const char *massflux_op_name, const char *quadrature_name,
                 double scale, double omega, double grav);
int swm_c_destroy(swm_cctx *ctx);
/* N, halo width of the stored fields, rows and pitch of one panel, elements of one field */
int swm_c_info(const swm_cctx *ctx, int *N, int *halo, int *rows, int *pitch, long long *field_elems);
int swm_c_synchronize(swm_cctx *ctx);
/* stvec_swm_t allocation / host copies (stvec_swm_factory_mod.f90 on mesh_p, mesh_u, mesh_v) */
int swm_c_state_alloc(swm_cctx *ctx, void **state);
int swm_c_state_free(swm_cctx *ctx, void *state);
int swm_c_state_upload(swm_cctx *ctx, void *state, const double *h, const double *u, const double *v,
                       int halo_width);
int swm_c_state_download(swm_cctx *ctx, const void *state, double *h, double *u, double *v, int halo_width);
/* operator_swm_t%h_surf at o points (operator_swm_mod.f90:85); NULL removes it */
int swm_c_set_orography(swm_cctx *ctx, const double *hs, int halo_width);
/* operator_swm_t%apply (operator_swm_mod.f90:61-113) with the C-grid operators: vout = RHS(vin),
 * enqueued on the context stream */
int swm_c_rhs(swm_cctx *ctx, const void *vin, void *vout);
/* nsteps of rk4_t%step (time_schemes/rk4_mod.f90:20-42) in place, HIP-graph captured */
int swm_c_rk4_steps(swm_cctx *ctx, void *state, double dt, int nsteps);
/* operator_t%get_diagnostics (operator_swm_mod.f90:115-236) with SBP_C42_quadrature
 * (quadrature_factory_mod.f90:88-175): out = {hmin, hmax, mass, TE, KE, PE, enstrophy} with
 * compensated FP64 sums; enstrophy is NaN (calc_enstrophy, :238-260, divides the curl at corner
 * points by h at the same indices and is marked in the reference as not working for the C grid) */
int swm_c_diagnostics(swm_cctx *ctx, const void *state, double out[7]);

/* ECS vector halo of the state's (u, v): create_vector_halo_procedure(halo, domain, halo_width, name)
 * + get_halo_vector(u, v, domain, halo_width) with name "ecs_C_vec" (contravariant components,
 * ecs_halo_vec_c_t) or "ecs_C_vec_covariant" (ecs_halo_vec_c_cov_t) (halo_factory_mod.f90,
 * equiang_cs/ecs_halo_vec_c_mod.f90:76-595, ecs_halo_vec_c_factory_mod.f90:18-449).  Device
 * outputs out_u (6, N+2hw, N+1+2hw) and out_v (6, N+1+2hw, N+2hw) doubles, i fastest, for i, j from
 * 1-hw: the panel's points (the normal component's edge faces averaged with the neighbour's
 * copies for "ecs_C_vec"), the halo strips of width hw beyond the four edges, NaN in the corner
 * squares (not defined by the procedure).  1 <= halo_width <= 5; other names return SWM_ENAME.
 * Enqueued on the context stream. */
int swm_c_ecs_halo_vec(swm_cctx *ctx, const char *halo_procedure_name, const void *state, int halo_width,
                       void *out_u, void *out_v);
/* create_grad_operator(domain, "gradient_c2_ecs") + calc_grad(gx, gy, f, domain)
 * (grad_c2_ecs_mod.f90:33-51,98-139): the ECS_O halo of width 1 of f (o points), then
 * (f(i) - f(i-1)) / (hx scale) at x points i = 1..N+1, j = 0..N+1 (gx) and likewise at y points
 * i = 0..N+1, j = 1..N+1 (gy).  f, gx, gy: device fields of field_elems doubles in the stored
 * layout (the h, u, v slots of a state vector will do); other names return SWM_ENAME. */
int swm_c_grad(swm_cctx *ctx, const char *grad_op_name, const void *f, void *gx, void *gy);
/* create_massflux_operator(domain, "massflux_c_up4") + calc_massflux(fx, fy, f, u, v, domain)
 * (massflux_Cgrid_mod.f90:87-104,166-207): the ECS_O halo of width 3 of f, then the 4-point
 * upwind-biased face value times u at x points (fx) and v at y points (fy).  Device fields in the
 * stored layout; other names return SWM_ENAME. */
int swm_c_massflux(swm_cctx *ctx, const char *massflux_op_name, const void *f, const void *u, const void *v,
                   void *fx, void *fy);

#ifdef __cplusplus
}
#endif
#endif
