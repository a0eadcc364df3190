// Host-side FP64 set-up of the equiangular cubed sphere for the device path:
// metric terms, Coriolis parameter, edge-sync weights, quadrature weights,
// cube-edge neighbour tables and the seeded spherical-harmonic initial state.
// Compiled by the host compiler with -ffp-contract=off.
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

// Layout of every per-point array: [panel 0..5][j 0..n-1][i 0..pitch-1], n = N+1.

// Metric terms at Ah points of the unit sphere (ecs_metric_mod.f90:303-478) and the
// Coriolis parameter f = 2*omega*(axis.r)/|r| (coriolis_factory_mod.f90:151-172).
// Any output pointer may be NULL.
void swmh_geometry(int N, int pitch, double omega, const double axis[3],
                   double *J, double *Qi11, double *Qi12, double *Qi22, double *f,
                   double *r /*3 planes*/, double *a1 /*3 planes*/, double *a2 /*3 planes*/);

// Contravariant basis vectors b1, b2 (3 planes each) of every Ah point, device layout
// (ecs_metric_mod.f90:367-395 on the prototype face, rotated to panel p).
void swmh_basis_b(int N, int pitch, double *b1, double *b2);

// Covariant Ah vector-sync weights qb, qt, ql, qr (n values each)
// (ecs_Ah_vec_sync_factory_mod.f90:101-149).
void swmh_sync_weights(int N, double *qb, double *qt, double *ql, double *qr);

// SBP quadrature weights A(1..n) for order 21/42/63 (quadrature_factory_mod.f90:158-177).
int swmh_quad_weights(int N, int order, double *A);

// Cube-edge neighbour table.  For panel p, edge e (0: i=1 side, 1: i=n side, 2: j=1 side,
// 3: j=n side), position k (0..n-1 along the edge) and halo depth d (0: the neighbour's copy
// of the shared edge point, 1: its first interior line), entry
//   tab[(((p*4+e)*n+k)*ndepth+d)*4 + {0,1,2,3}] = {q, iq, jq, code}
// with 0-based (iq,jq) on panel q and code bit 0 = components swapped, bit 1 = first
// component negated, bit 2 = second component negated (vector halo convention).
void swmh_halo_table(int N, int ndepth, int *tab);

// Edge syncs as sparse matrices over the 6*(4n-4) panel-edge points (slot numbering: j = 0:
// i; j = n-1: n+i; i = 0: 2n+j-1; i = n-1: 3n+j-3; global slot p*(4n-4)+slot).  vector = 0:
// the scalar sync (halo_Ah_scalar_sync_mod.f90:52-93), entries over slots; vector = 1: the
// covariant vector sync (ecs_Ah_vec_sync_mod.f90:77-124), entries over c*6*(4n-4)+slot with
// component c = 0 (x) or 1 (y).  Writes at most `cap` (row, col, weight) triplets and
// returns the number of entries.
int swmh_edge_sync_matrix(int N, int vector, int *rows, int *cols, double *w, int cap);

// ECS cubic halo on xy points: weights w[((j-1)*(n+2hw) + i-1+hw)*4 + s+1] (s = -1..2) and
// 1-based stencil base indices ind[(j-1)*(n+2hw) + i-1+hw] for i = 1-hw .. n+hw, halo row
// j = 1..hw (init_halo_xy_interp, ecs_halo_factory_mod.f90:228-269).  Returns 0 on success.
int swmh_ecs_xy_weights(int N, int hw, double *w, int *ind);

// ECS cubic halo on O points (cell centres): weights w[((r-1)*(N+4) + t+1)*4 + s+1] (s = -1..2)
// and 1-based stencil base indices ind[(r-1)*(N+4) + t+1] for positions t = -1..N+2 along an
// edge and halo rows r = 1..rows (init_halo_interp, ecs_halo_factory_mod.f90:193-226).
void swmh_ecs_o_weights(int N, int rows, double *w, int *ind);

// ECS vector halo on O points (A grid): transform matrices of init_transform_matrix
// (ecs_halo_vec_a_factory_mod.f90:114-175) from the exchanged neighbour components to this panel's
// contravariant components, tm[(((edge-1)*rows + r-1)*N + t-1)*4 + k] for edge 1 (below j = 1),
// 2 (above j = N), 3 (left of i = 1), 4 (right of i = N), halo row r, position t = 1..N.
void swmh_ecs_vec_a_tm(int N, int rows, double *tm);

// Cubic lat-lon regrid of Ah fields (create_latlon_interp, regridders/regrid_factory_mod.f90:
// 183-301, scalar_grid_type "Ah", interp_type "cubic"; covariant components as in
// create_latlon_vector_regrid :80-181).  Target q = (j-1)*Nlon + (i-1), lat = -pi/2 +
// (j-1) pi/(Nlat-1), lon = (i-1) 2 pi/Nlon.  panel[q] (0-based), x0[q], y0[q] (0-based first
// stencil point, the 4 x 4 stencil inside the panel's mesh points), wx[4q..4q+3], wy[..] (cubic
// Lagrange weights) and, when rot is not NULL, rot[4q..4q+3] = (u2u, u2v, v2u, v2v): the
// lat-lon components of covariant (u, v) are (u2u u + v2u v, u2v u + v2v v).
void swmh_latlon_table(int N, int Nlat, int Nlon, int *panel, int *x0, int *y0, double *wx, double *wy,
                       double *rot);

// Seeded initial state: h = h0 + h' (max|h'| = h_amp), V = r x grad(psi) (max|V| = v_max),
// h', psi random real spherical-harmonic series of degree 1..lmax with N(0,1)*l^-2
// coefficients drawn from a splitmix64/Box-Muller stream of `seed`; covariant components
// u = V.a1, v = V.a2 (test_fields_mod.f90:397-433).  Output arrays in device layout.
int swmh_ic_sh(int N, int pitch, unsigned long long seed, int lmax, double h0, double h_amp,
               double v_max, double *h, double *u, double *v, int nthreads);

// C grid: metric terms and Coriolis parameter at the points of type t (0 o, 1 x, 2 y, 3 q) in
// the stored C layout [panel][j-1+H (rows)][i-1+H (pitch)], i, j = 1..N+1; any output may be NULL.
void swmh_c_geometry(int N, int t, int H, int rows, int pitch, double omega, const double axis[3],
                     double *J, double *Qi11, double *Qi12, double *Qi22, double *f);

// C grid cube-edge gather lists (element offsets in the stored layout, the second component of a
// vector pair at +fs): tangential vector halo to depth_vec (u beyond the y edges, v beyond the x
// edges, 6*4*(N+1)*depth_vec entries), the face halo index 0 / nx+1 of the normal component
// (6*4*N entries), the O-point cross halo to depth_o (6*4*N*depth_o) and, for each q point on
// a panel edge, its own offset and the offset of the neighbour's copy across that edge
// (6*4*(N+1) entries, edges in the order i=1, i=N+1, j=1, j=N+1).  counts = {vec, face, o, q}.
// Returns 0, or a negative value when a neighbour point is not of the expected type.
int swmh_c_halo_lists(int N, int H, int rows, int pitch, long long fs, int depth_vec, int depth_o,
                      long long *vec_dst, long long *vec_src, signed char *vec_sign,
                      long long *face_dst, long long *face_src, signed char *face_sign,
                      long long *o_dst, long long *o_src, long long *q_own, long long *q_nb,
                      int *counts);

// C grid ECS vector halo (create_ecs_C_vec_halo_procedure / ..._covariant_..., ecs_halo_vec_c_factory_mod.f90:
// 18-270), one table set for every edge: w0/i0 tangential interpolation (positions 1..N+1), w1/i1
// normal interpolation (1..N), w2/i2 interpolation of the other component to the transformed
// component's positions (1..N+1); q1 = the transformed component's own coefficient (qtt or qnn),
// q2 = the other's (qtn or qnt), ct the covariant tangential stretch.  Entry (i, j) of position i
// and halo row j at (i-1)*hw + j-1 (weights x4).  1 <= hw <= 5, N >= 8; returns 0.
int swmh_c_ecs_vec_tables(int N, int hw, int covariant, double *w0, int *i0, double *w1, int *i1, double *w2,
                          int *i2, double *q1, double *q2, double *ct);

// Sources of the strips beyond each panel edge to `depth` rows, as element offsets in the stored C
// layout from the u field (v at +fs) or the o field: normal component at face depth D (0 = the
// neighbour's copy of the shared face) [p][e][D][k < N], tangential component at centre depth D
// [p][e][D][k < N+1] with signs, cell centres at depth D [p][e][D][k < N]; edges e = 0 (i < 1),
// 1 (i > N+1 / N), 2 (j < 1), 3 (j > ...).  Returns 0, or a negative value on a topology mismatch.
int swmh_c_ecs_strip_sources(int N, int H, int rows, int pitch, long long fs, int depth, long long *nsrc,
                             signed char *nsign, long long *tsrc, signed char *tsign, long long *osrc);

#ifdef __cplusplus
}
#endif
