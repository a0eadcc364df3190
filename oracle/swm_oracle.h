/*
 * swm_oracle.h -- CPU FP64 restatement of the reference shallow-water hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (cubedsphere_sbp_swm_amd)
 * never links, imports or calls it.
 *
 * Every function restates one reference routine; the citation of the file and
 * lines it follows is given next to each definition in swm_oracle.c.  Paths are
 * relative to the reference's src/ directory.
 *
 * Storage: one field = 6 panels x (n+2*hw) x (n+2*hw) doubles, n = N+1 Ah
 * points per panel side, halo width hw; 1-based panel-local indices
 * i,j in [1-hw, n+hw] as in the reference's tile arrays p(i,j,k).
 */
#ifndef SWM_ORACLE_H
#define SWM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_grid ora_grid;

/* grad_order in {21,42,43,63}; div_kind in {2,42,43,63} (2 = divergence_ah2) */
ora_grid *ora_create(int N, int grad_order, int div_kind, double scale,
                     double omega, double grav);
void ora_destroy(ora_grid *g);

int ora_n(const ora_grid *g);      /* points per panel side (N+1) */
int ora_num_threads(void);         /* OpenMP threads the point loops use */
int ora_hw(const ora_grid *g);     /* halo width of stored fields */
int ora_ld(const ora_grid *g);     /* leading dimension n+2*hw */
long ora_field_len(const ora_grid *g); /* 6*ld*ld */

/* geometry accessors: which = 0 J, 1 Qi11, 2 Qi12, 3 Qi22, 4 f (Coriolis),
 * 5..7 r, 8..10 a1, 11..13 a2, 14..16 b1, 17..19 b2, 20..22 Q11,Q12,Q22 */
void ora_geometry(const ora_grid *g, int which, double *out);
/* q weights of the covariant Ah vector sync: which 0 qb, 1 qt, 2 ql, 3 qr; n values */
void ora_sync_weights(const ora_grid *g, int which, double *out);
/* SBP quadrature edge weights A(1..n) of the given order (21/42/63) */
void ora_quad_weights(const ora_grid *g, int order, double *out);

/* halo exchange map of the xy-point exchange (width 2):
 * out[panel][j][i] over the stored extent: encoded source or -1.
 * code = ((src_panel*ld + (js-1+hw))*ld + (is-1+hw)); vec = swap | (su<0)<<1 | (sv<0)<<2 */
void ora_halo_map(const ora_grid *g, int *src, int *vec);

/* fill halos of stored fields (xy-point exchange semantics) */
void ora_exchange(const ora_grid *g, double *f, int width);
void ora_exchange_vec(const ora_grid *g, double *u, double *v, int width);

/* operators (all fields in stored layout; halos of inputs are overwritten) */
void ora_sbp_apply(const ora_grid *g, int order, char dir, const double *f, double *df);
void ora_div(const ora_grid *g, double *u, double *v, double *div);
void ora_grad(const ora_grid *g, double *f, double *gx, double *gy);
void ora_curl(const ora_grid *g, double *u, double *v, double *curl);
void ora_scalar_sync(const ora_grid *g, double *f);
void ora_vec_sync_covariant(const ora_grid *g, double *gx, double *gy);
/* narrow-stencil SBP Laplacians (order 42 or 63) and the hyperdiffusion built on them
 * (N+1 >= ora_narrow_min_n(order): 12 for sbp42, 18 for sbp63) */
int ora_narrow_min_n(int order);
void ora_laplace_narrow(const ora_grid *g, int order, const double *f, double *out);
void ora_hordiff_scalar_narrow(const ora_grid *g, int order, double hordiff_coeff, const double *f,
                               double *tend);
void ora_hordiff_vec_narrow(const ora_grid *g, int order, double hordiff_coeff, const double *u,
                            const double *v, double *ut, double *vt);
void ora_diff_step_narrow(const ora_grid *g, int uv_order, int h_order, double *h, double *u, double *v,
                          double dt, double uv_diff_coeff, double h_diff_coeff);

/* cubic lat-lon regrid of an Ah scalar (fv == NULL) or covariant vector (f, fv) onto Nlat x Nlon
 * points, lon fastest; optional target table (panel, x0, y0: 0-based; w: wx[4], wy[4] per point) */
void ora_latlon_regrid(const ora_grid *g, int Nlat, int Nlon, const double *f, const double *fv, double *out_a,
                       double *out_b, int *panel, int *x0, int *y0, double *w);

/* full RHS: out = RHS(in); in/out are [h,u,v] each a stored field; hs may be NULL */
void ora_rhs(const ora_grid *g, double *in_h, double *in_u, double *in_v,
             const double *hs, double *dh, double *du, double *dv);
/* classical RK4 steps in place on (h,u,v) */
void ora_rk4(const ora_grid *g, double *h, double *u, double *v, const double *hs,
             double dt, int nsteps);
/* diagnostics: hmin, hmax, mass, TE, KE, PE, enstrophy */
void ora_diagnostics(const ora_grid *g, int quad_order, double *h, double *u,
                     double *v, const double *hs, double out[7]);
/* quadrature mass of a field, and l2norm / maxabs of the reference's vec_math */
double ora_quad_mass(const ora_grid *g, int quad_order, const double *f);
double ora_l2norm(const ora_grid *g, const double *f);
double ora_maxabs(const ora_grid *g, const double *f);

/* test-field projection: cartesian vectors (vx,vy,vz) sampled at r -> covariant
 * (kind 0) or contravariant (kind 1) components; points incl. halo */
void ora_project_vector(const ora_grid *g, int kind, const double *vx,
                        const double *vy, const double *vz, double *u, double *v);

/* ECS cubic halo on xy points (ecs_halo_xy_mod.f90, ecs_halo_factory_mod.f90:228-269):
 * in place on a stored field, halo rows 1..hw (hw + 1 <= stored halo width); weights
 * w[((j-1)*(n+2hw) + i-1+hw)*4 + s+1] and base indices for i = 1-hw..n+hw, j = 1..hw */
void ora_ecs_xy_weights(const ora_grid *g, int hw, double *w, int *ind);
int ora_ecs_halo_xy(const ora_grid *g, double *f, int hw);
/* ECS cubic halo of an O-point (cell-centre) field f[6][N][N] into out[6][N+2hw][N+2hw]
 * (1 <= hw <= 8, N >= 8); weights for rows 1..rows, positions -1..N+2 */
void ora_ecs_o_weights(const ora_grid *g, int rows, double *w, int *ind);
int ora_ecs_halo_o(const ora_grid *g, const double *f, int hw, double *out);
/* ECS vector halo on O points (A grid, contravariant u, v): transform matrices for rows
 * 1..rows (edges 1 j<1, 2 j>N, 3 i<1, 4 i>N; tm[(((edge-1)*rows + r-1)*N + t-1)*4 + k]) and the
 * halo of (u, v) [6][N][N] into out_u, out_v [6][N+2hw][N+2hw] */
void ora_ecs_vec_a_tm(const ora_grid *g, int rows, double *tm);
int ora_ecs_halo_vec_a(const ora_grid *g, const double *u, const double *v, int hw, double *out_u, double *out_v);

/* generic classical RK4 on a diagonal complex linear ODE df/dt = w f (the
 * reference's iomega mock model); returns f after one step */
void ora_rk4_iomega(int n, const double *w_re, const double *w_im,
                    double *f_re, double *f_im, double dt);

/* ---- C grid, SBP42 vector-invariant family (divergence_c_sbp42, gradient_c_sbp42,
 * curl_c_sbp42, coriolis_Cgrid_sbp42, KE_Cgrid_sbp42, co2contra_c_sbp42, massflux_c_sbp42).
 * Point types t: 0 o (cell centres, N x N), 1 x (u, N+1 x N), 2 y (v, N x N+1), 3 q (corners,
 * N+1 x N+1).  Every field uses the storage of the Ah grid of the same N (ora_cgrid_grid),
 * index (i, j) of any type at the same place. ---- */
typedef struct ora_cgrid ora_cgrid;
ora_cgrid *ora_cgrid_create(int N, double scale, double omega, double grav);
void ora_cgrid_destroy(ora_cgrid *c);
ora_grid *ora_cgrid_grid(ora_cgrid *c);
/* which: 0 J, 1 Qi11, 2 Qi12, 3 Qi22, 4 f, 5 r, 6 a1, 7 a2, 8 b1, 9 b2 (comp 0..2 for vectors) */
int ora_cgrid_geometry(const ora_cgrid *c, int t, int which, int comp, double *out);
void ora_c_co2contra(const ora_cgrid *c, const double *u, const double *v, double *ut, double *vt);
void ora_c_massflux(const ora_cgrid *c, const double *h, const double *ut, const double *vt, double *hu, double *hv);
void ora_c_KE(const ora_cgrid *c, const double *u, const double *v, const double *ut, const double *vt, double *KE);
/* f's O halo (depth 3) is written */
void ora_c_grad(const ora_cgrid *c, double *f, double *gx, double *gy);
/* the C vector halo of u, v (depth 4) is written; curl at q points, corner copies averaged */
void ora_c_curl(const ora_cgrid *c, double *u, double *v, double *curl);
void ora_c_coriolis(const ora_cgrid *c, const double *hu, const double *hv, const double *h, const double *curl,
                    double *cor_u, double *cor_v);
void ora_c_div(const ora_cgrid *c, const double *u, const double *v, double *div);
void ora_c_exchange_o(const ora_cgrid *c, double *f, int depth);
void ora_c_exchange_vec(const ora_cgrid *c, double *u, double *v, int depth);
/* RHS and RK4 of the C-grid SWM: h at o points, u at x points, v at y points; hs may be NULL */
void ora_c_rhs(const ora_cgrid *c, double *h, double *u, double *v, const double *hs,
               double *dh, double *du, double *dv);
void ora_c_rk4(const ora_cgrid *c, double *h, double *u, double *v, const double *hs, double dt, int nsteps);
double ora_c_l2norm(const ora_cgrid *c, int t, const double *f);
/* SBP_C42_quadrature mass on the mesh of point type t, and the diagnostics hmin, hmax, mass, TE,
 * KE, PE, enstrophy (NaN: the reference's enstrophy does not apply to the C grid) */
double ora_c_quad_mass(const ora_cgrid *c, int t, const double *f);
void ora_c_diagnostics(const ora_cgrid *c, double *h, double *u, double *v, const double *hs, double out[7]);
double ora_c_maxabs(const ora_cgrid *c, int t, const double *f);
/* ECS vector halo on the C grid (the reference's "ecs_C_vec" / "ecs_C_vec_covariant" halo procedures),
 * 1 <= hw <= 5, N >= 8: interpolation tables (see swm_oracle.c) and the halo of u (x points), v (y
 * points) into out_u (6, N+2hw, N+1+2hw), out_v (6, N+1+2hw, N+2hw) with i, j from 1-hw; NaN in
 * the corner squares, which the procedure does not define */
void ora_c_ecs_vec_tables(const ora_cgrid *c, int hw, int covariant, double *w0, int *i0, double *w1, int *i1,
                          double *w2, int *i2, double *q1, double *q2, double *ct);
int ora_c_ecs_halo_vec(const ora_cgrid *c, const double *u, const double *v, int hw, int covariant, double *out_u,
                       double *out_v);
/* gradient_c2_ecs (gx at x points j = 0..N+1, gy at y points i = 0..N+1) and massflux_c_up4 */
int ora_c_grad_c2_ecs(const ora_cgrid *c, const double *f, double *gx, double *gy);
int ora_c_massflux_up4(const ora_cgrid *c, const double *f, const double *u, const double *v, double *fx,
                       double *fy);

#ifdef __cplusplus
}
#endif
#endif
