/*
 * FP64 CPU oracle of sparrowpy's DirectionalRadiosityFast hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference algorithm
 * (sparrowpy 1.0.1) in plain C so that tests, the package's smoke check and
 * the benchmark's CPU-baseline leg can compare the HIP kernels against it.
 * The product (sparrowpy_amd) never links, loads or calls it.
 *
 * Every function names the reference function and file:line it follows.
 * Arrays are C-contiguous, float64 unless stated otherwise; booleans are
 * uint8 (0/1); polygon corner arrays have NPT = 4 corners of 3 doubles.
 */
#ifndef SPARROW_ORACLE_H
#define SPARROW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* geometry.py:841-909 _basic_visibility (polygon with npts corners) */
int orc_basic_visibility(const double *vis_point, const double *eval_point,
                         const double *surf_points, int npts,
                         const double *surf_normal);

/* geometry.py:614-686 _point_in_polygon */
int orc_point_in_polygon(const double *point3d, const double *polygon3d,
                         int npts, const double *plane_normal);

/* geometry.py:750-797 _check_patch2patch_visibility: out (P,P) upper
 * triangle, surfaces (NS, npts, 3) with normals (NS, 3). */
void orc_p2p_visibility(int64_t P, const double *centers, int64_t NS,
                        const double *surf_normals, const double *surf_pts,
                        int npts, uint8_t *out, int nthreads);

/* geometry.py:750-797 for a list of pairs (i < j each, int64 [n,2]):
 * out[q] = visibility of pair q (sampled rows of large scenes). */
void orc_p2p_visibility_pairs(const double *centers, int64_t NS,
                              const double *surf_normals,
                              const double *surf_pts, int npts, int64_t n,
                              const int64_t *pairs, uint8_t *out,
                              int nthreads);

/* geometry.py:799-839 _check_point2patch_visibility: out (P,) */
void orc_point_visibility(const double *point, int64_t P,
                          const double *centers, int64_t NS,
                          const double *surf_normals, const double *surf_pts,
                          int npts, uint8_t *out);

/* integration.py:38-114 stokes_integration (4-corner patches) */
double orc_stokes(const double *patch_i, const double *patch_j,
                  double patch_i_area);

/* integration.py:232-289 nusselt_integration (regular grid) */
double orc_nusselt(const double *patch_i, const double *patch_j,
                   const double *normal_i, const double *normal_j,
                   int nsamples);

/* universal.py:54-96 universal_form_factor */
double orc_universal_ff(const double *src_pts, const double *src_normal,
                        double src_area, const double *rcv_pts,
                        const double *rcv_normal);

/* universal.py:12-52 patch2patch_ff_universal: out (P,P), zero-filled here */
void orc_form_factors(int64_t P, const double *pts, const double *normals,
                      const double *areas, int64_t V, const int32_t *pairs,
                      double *out, int nthreads);

/* integration.py:295-344 pt_solution; mode_receiver: 0 source, 1 receiver */
double orc_pt_solution(const double *point, const double *patch_points,
                       int npts, int mode_receiver);

/* universal.py:98-147 _source2patch_energy_universal.
 * air may be NULL; energy (P,B), distance (P,) */
void orc_source_energy(const double *source, int64_t P,
                       const double *centers, const double *pts,
                       const uint8_t *vis, const double *air, int64_t B,
                       double *energy, double *distance);

/* universal.py:149-160 _patch2receiver_energy_universal: out (P,) */
void orc_receiver_energy(const double *receiver, int64_t P,
                         const double *pts, const uint8_t *vis, double *out);

/* RadiosityFast.py:1277-1312 get_scattering_data_receiver_index.
 * pos_i (n,3), pos_j (3), receivers (W,D,3), wall_ids (n) -> out (n) */
void orc_receiver_index(int64_t n, const double *pos_i, const double *pos_j,
                        int64_t D, const double *receivers,
                        const int64_t *wall_ids, int64_t *out);

/* RadiosityFast.py:1359-1390 get_scattering_data_source: index of the
 * nearest incoming direction of wall w for direction pos_h - pos_i. */
int64_t orc_source_index(const double *pos_h, const double *pos_i,
                         int64_t S, const double *sources, int64_t wall);

/* RadiosityFast.py:1234-1272 _form_factors_with_directivity_dim.
 * vis (P,P) u8, ff (P,P), air (B) or NULL, wall_ids (P),
 * brdf (NBRDF,S,D,B) or NULL, brdf_index (W), sources (W,S,3),
 * out (P,P,D,B) zero-filled here (D = 1 when brdf is NULL). */
void orc_fft(int64_t P, const uint8_t *vis, const double *ff, int64_t B,
             const double *centers, const double *areas, const double *air,
             const int64_t *wall_ids, int64_t S, int64_t D,
             const double *brdf, const int64_t *brdf_index,
             const double *sources, double *out, int nthreads);

/* RadiosityFast.py:1234-1272 for a list of directed visible pairs
 * (i -> j, int64 [n,2]); ff_upper[q] = ff[min(i,j), max(i,j)].
 * out (n,D,B). */
void orc_fft_pairs(int64_t n, const int64_t *pairs, const double *ff_upper,
                   int64_t B, const double *centers, const double *areas,
                   const double *air, const int64_t *wall_ids, int64_t S,
                   int64_t D, const double *brdf, const int64_t *brdf_index,
                   const double *sources, double *out);

/* RadiosityFast.py:988-1034 _add_directional: out (P,D,B) */
void orc_add_directional(int64_t P, int64_t B, const double *energy0,
                         const double *source, const double *centers,
                         const int64_t *wall_ids, int64_t S, int64_t D,
                         const double *sources, const double *brdf,
                         const int64_t *brdf_index, double *out);

/* RadiosityFast.py:1037-1070 _energy_exchange_init_energy.
 * out (P,D,B,T) zero-filled here; returns -1 if a delay is >= T. */
int orc_init_energy(int64_t T, int64_t P, int64_t D, int64_t B,
                    const double *e0dir, const double *d0, double c,
                    double dt, double *out);

/* RadiosityFast.py:1073-1145 _energy_exchange.
 * dist (P,P), fft (P,P,D,B), p2o (P,P) int64, pairs (V,2) int32,
 * out (P,D,B,T).  work: 2*P*D*B*T doubles of scratch.  nthreads <= 1 runs
 * the reference's pair-major loop; > 1 runs the same per-destination
 * accumulation order in parallel over destinations (bit-identical).
 * Returns -1 if an init delay is >= T. */
int orc_exchange(int64_t T, int64_t P, int64_t D, int64_t B,
                 const double *e0dir, const double *d0, const double *dist,
                 const double *fft, const int64_t *p2o, double c, double dt,
                 int64_t K, int64_t V, const int32_t *pairs, double *out,
                 double *work, int nthreads);

/* One order of the exchange on dense arrays (vis (P,P) upper triangle,
 * dist, fft, p2o as orc_exchange): cur (P,D,B,T) overwritten. */
void orc_exchange_step(int64_t T, int64_t P, int64_t D, int64_t B,
                       const double *prev, double *cur, const double *dist,
                       const double *fft, const int64_t *p2o, double c,
                       double dt, const uint8_t *vis, int nthreads);

/* Adjoint of one order (transpose of RadiosityFast.py:1133-1143):
 * mu_out = lam + A^T mu_next; dfft (P,P,D,B) += per-pair coefficient
 * gradients against the forward order ek.  No reference counterpart. */
void orc_exchange_adjoint_step(int64_t T, int64_t P, int64_t D, int64_t B,
                               const double *lam, const double *mu_next,
                               const double *ek, const double *dist,
                               const double *fft, const int64_t *p2o,
                               double c, double dt, const uint8_t *vis,
                               double *mu_out, double *dfft, int nthreads);

/* RadiosityFast.py:1359-1390 source index for directed pairs (n,2). */
void orc_source_index_pairs(int64_t n, const int64_t *pairs,
                            const double *centers, int64_t S,
                            const double *sources, const int64_t *wall_ids,
                            int64_t *out);

/* One reflection order of the exchange restricted to destinations
 * [j_begin, j_end) of the per-destination lists (CPU timing sample).
 * cur is overwritten on those destinations. */
void orc_exchange_order(int64_t T, int64_t P, int64_t D, int64_t B,
                        const double *prev, double *cur,
                        const int64_t *row_ptr, const int32_t *row_src,
                        const int32_t *row_lag, const int32_t *row_dir,
                        const double *fft, int64_t j_begin, int64_t j_end,
                        int nthreads);

/* One order (RadiosityFast.py:1121-1143) for n_dest destinations with
 * compact source rows: prev_rows (n_rows,B,T), per entry row_src (index
 * into prev_rows), row_lag, coef (e,D,B); cur (n_dest,D,B,T). */
void orc_exchange_rows(int64_t T, int64_t D, int64_t B, int64_t n_dest,
                       const int64_t *row_ptr, const int64_t *row_src,
                       const int64_t *row_lag, const double *coef,
                       const double *prev_rows, double *cur, int nthreads);

/* RadiosityFast.py:1148-1185 _collect_receiver_energy:
 * E (P,B,T), dist (P), air (B) -> out (P,B,T) */
void orc_collect_receiver_energy(int64_t P, int64_t B, int64_t T,
                                 const double *E, const double *dist,
                                 double c, double dt, const double *air,
                                 double *out);

/* int(d / c / dt) and ceil variant, as the reference computes delays */
int64_t orc_delay_floor(double d, double c, double dt);
int64_t orc_delay_ceil(double d, double c, double dt);

#ifdef __cplusplus
}
#endif
#endif
