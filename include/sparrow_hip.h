/*
 * sparrow_hip: MI355X (gfx950) kernels for sparrowpy's directional
 * radiosity hot path, behind a plain C ABI.
 *
 * Conventions
 *  - Every pointer argument is a DEVICE pointer unless named *_host.
 *  - Arrays are C-contiguous with the shapes given in brackets.
 *  - Geometry is float64; visibility flags are uint8 (0/1); patch and
 *    direction indices are int32 or int64 as stated; energies are float32.
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *    asynchronous unless documented otherwise; buffers are caller-owned and
 *    no entry point allocates device memory.
 *  - Return value: SPR_OK (0) or a negative SPR_ERR_* code; the message is
 *    available from spr_last_error() on the calling thread.
 *
 * Each entry point names the reference function it replaces
 * (sparrowpy 1.0.1 file:line).
 */
#ifndef SPARROW_HIP_H
#define SPARROW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPR_OK 0
#define SPR_ERR_ARG (-1)         /* bad shape, size or pointer */
#define SPR_ERR_HIP (-2)         /* HIP runtime error */
#define SPR_ERR_RANGE (-3)       /* a delay falls outside the histogram */
#define SPR_ERR_UNSUPPORTED (-4) /* e.g. more directions than supported */
#define SPR_ERR_WORKSPACE (-5)   /* workspace too small */

/* Message of the last failing call on this thread ("" if none). */
const char *spr_last_error(void);

/* Library version string. */
const char *spr_version(void);

/* ---------------------------------------------------------------------
 * Bake: visibility
 * ------------------------------------------------------------------- */

/* geometry._check_patch2patch_visibility (sparrowpy/geometry.py:750-797),
 * called by DirectionalRadiosityFast.bake_geometry (RadiosityFast.py:374).
 * centers[P,3]; surfaces (possible blockers) surf_normals[NS,3],
 * surf_points[NS,4,3].  vis[P,P]: vis[i,j] for i<j is 1 when no surface
 * blocks the segment between the centres, 0 otherwise; the diagonal and the
 * lower triangle are 0.  Bit-exact with the reference predicate. */
int spr_p2p_visibility(int64_t P, const double *centers, int64_t NS,
                       const double *surf_normals, const double *surf_points,
                       uint8_t *vis, void *workspace, int64_t workspace_bytes,
                       void *stream);

/* Workspace bytes of spr_p2p_visibility / spr_point_visibility for NS
 * blocking surfaces (one precomputed record per surface).
 * Sizing helper; no counterpart in the reference. */
int64_t spr_visibility_workspace(int64_t NS);

/* geometry._check_point2patch_visibility (sparrowpy/geometry.py:799-839),
 * called by init_source_energy (RadiosityFast.py:482) and
 * _collect_energy_patches (RadiosityFast.py:714), batched over NQ query
 * points.  points[NQ,3], centers[P,3], surf_normals[NS,3],
 * surf_points[NS,4,3] -> vis[NQ,P]. */
int spr_point_visibility(int64_t NQ, const double *points, int64_t P,
                         const double *centers, int64_t NS,
                         const double *surf_normals,
                         const double *surf_points, uint8_t *vis,
                         void *workspace, int64_t workspace_bytes,
                         void *stream);

/* Visible-pair list, the loop at RadiosityFast.py:377-387.
 * Pass 1 (pairs == NULL): writes row_offsets[P+1] (exclusive scan of the
 * per-row counts of vis[i, j>i]); row_offsets[P] is the pair count V.
 * Pass 2 (pairs != NULL, row_offsets from pass 1): writes pairs[V,2]
 * (int32, row-major i<j, the reference order). */
int spr_visible_pairs(int64_t P, const uint8_t *vis, int64_t *row_offsets,
                      int32_t *pairs, void *stream);

/* ---------------------------------------------------------------------
 * Bake: form factors and direction maps
 * ------------------------------------------------------------------- */

/* form_factor.patch2patch_ff_universal (sparrowpy/form_factor/
 * universal.py:12-52): for each pair (i,j) in pairs[V,2], ff[i,j] =
 * universal_form_factor(patch i -> patch j) (Nusselt analogue with 64
 * regular samples when the patches share a vertex, Stokes contour
 * integration otherwise).  points[P,4,3], normals[P,3], areas[P].
 * ff[P,P] is zero-filled first. */
int spr_form_factors(int64_t P, const double *points, const double *normals,
                     const double *areas, int64_t V, const int32_t *pairs,
                     double *ff, void *stream);

/* get_scattering_data_receiver_index (RadiosityFast.py:1277-1312) as used
 * to build the patch-to-outgoing-direction map (RadiosityFast.py:403-414):
 * p2o[i,j] = argmin_d |out_dirs[wall_i, d] - unit(c_j - c_i)|^2 when
 * (i,j) is visible in either orientation, D otherwise.  vis[P,P] (upper
 * triangle), centers[P,3], out_dirs[W,D,3], wall_ids[P] (int64) ->
 * p2o[P,P] (int64). */
int spr_direction_map(int64_t P, const uint8_t *vis, const double *centers,
                      int64_t D, const double *out_dirs,
                      const int64_t *wall_ids, int64_t *p2o, void *stream);

/* get_scattering_data_receiver_index for one target point (receiver
 * collection, RadiosityFast.py:728-730): idx[k] = argmin_d
 * |out_dirs[wall_k, d] - unit(point - c_k)|^2 for every patch k.
 * points[NQ,3] -> idx[NQ,P] (int64). */
int spr_point_direction_index(int64_t NQ, const double *points, int64_t P,
                              const double *centers, int64_t D,
                              const double *out_dirs,
                              const int64_t *wall_ids, int64_t *idx,
                              void *stream);

/* _form_factors_with_directivity_dim (RadiosityFast.py:1234-1272), dense:
 * fft[i,j,d,b] = F~_ij * exp(-air_b * |u_ij|) * brdf[bi(w_i), s_ij, d, b]
 * for visible (i,j), 0 elsewhere, with F~_ij = ff[i,j] (i<j) or
 * ff[j,i]*A_j/A_i, u_ij = unit(c_i - c_j) and s_ij the nearest incoming
 * direction of wall w_i to u_ij.  air may be NULL (no air factor); brdf may
 * be NULL (then D = 1 and no BRDF factor).  vis[P,P], ff[P,P],
 * centers[P,3], areas[P], air[B], wall_ids[P], brdf[NBRDF,S,D,B],
 * brdf_index[W] (int64), in_dirs[W,S,3] -> fft[P,P,D,B] (float64). */
int spr_fft_dense(int64_t P, const uint8_t *vis, const double *ff, int64_t B,
                  const double *centers, const double *areas,
                  const double *air, const int64_t *wall_ids, int64_t S,
                  int64_t D, const double *brdf, const int64_t *brdf_index,
                  const double *in_dirs, double *fft, void *stream);

/* ---------------------------------------------------------------------
 * Source injection
 * ------------------------------------------------------------------- */

/* _source2patch_energy_universal (universal.py:98-147) followed by
 * _add_directional (RadiosityFast.py:988-1034), the body of
 * init_source_energy (RadiosityFast.py:482-497).  source[3], centers[P,3],
 * points[P,4,3], vis[P] (source visibility), air[B] (may be NULL),
 * wall_ids[P], brdf[NBRDF,S,D,B], brdf_index[W], in_dirs[W,S,3] ->
 * energy0[P,B] (float64, before the BRDF), e0dir[P,D,B] (float64),
 * distance[P] (float64; 0 for patches the source does not see). */
int spr_source_energy(const double *source, int64_t P, const double *centers,
                      const double *points, const uint8_t *vis,
                      const double *air, int64_t B, const int64_t *wall_ids,
                      int64_t S, int64_t D, const double *brdf,
                      const int64_t *brdf_index, const double *in_dirs,
                      double *energy0, double *e0dir, double *distance,
                      void *stream);

/* ---------------------------------------------------------------------
 * Energy exchange (the hot loop)
 *
 * The exchange runs on a destination-major "plan": for every destination
 * patch j, the list of source patches i that see it (ascending i, the
 * reference's accumulation order), each with its delay int(d_ij/c/dt),
 * outgoing direction p2o[i,j] and coefficient row fft[i,j,:,b].
 * Plan arrays (caller-owned): row_ptr[P+1] (int64), src[Vd], lag[Vd],
 * dir[Vd] (int32), coef[Vd,B,DP] (float64, DP = spr_padded_directions(D),
 * zero-padded).  Vd = 2V.
 * State (caller-owned, float32, [P,D,B,T] each): prev, cur (order
 * histograms) and tot_hi, tot_lo (running total as an unevaluated sum
 * hi + lo, i.e. compensated accumulation).
 * ------------------------------------------------------------------- */

/* Directions padded to the kernel's register tile (1,2,4,...,64); 0 when
 * D is unsupported (D > 64).
 * Sizing helper; no counterpart in the reference. */
int64_t spr_padded_directions(int64_t D);

/* Plan structure from the visibility matrix: row_ptr and src.
 * Pass 1 (src == NULL) writes row_ptr (row_ptr[P] = Vd); pass 2 fills src.
 * The rows list, for destination j, every i != j with vis[min(i,j),
 * max(i,j)] = 1 in ascending order.
 * Replaces the visible-pair walk of _energy_exchange
 * (RadiosityFast.py:1124-1143). */
int spr_plan_rows(int64_t P, const uint8_t *vis, int64_t *row_ptr,
                  int32_t *src, void *stream);

/* Plan delays, directions and coefficients from the reference's dense
 * arrays (the inputs of _energy_exchange, RadiosityFast.py:1073-1076):
 * dist[P,P], p2o[P,P] (int64), fft[P,P,D,B] (float64). */
int spr_plan_from_dense(int64_t P, int64_t D, int64_t B, double c, double dt,
                        int64_t Vd, const int64_t *row_ptr,
                        const int32_t *src, const double *dist,
                        const int64_t *p2o, const double *fft, int32_t *lag,
                        int32_t *dir, double *coef, void *stream);

/* Plan delays, directions and coefficients computed directly from the
 * geometry (fuses RadiosityFast.py:538-543 pair distances, the p2o map and
 * _form_factors_with_directivity_dim without materialising the dense
 * [P,P,D,B] tensor).  Arguments as spr_fft_dense plus p2o[P,P]; the
 * coefficients are bit-identical to spr_fft_dense's values. */
int spr_plan_from_geometry(int64_t P, int64_t D, int64_t B, double c,
                           double dt, int64_t Vd, const int64_t *row_ptr,
                           const int32_t *src, const double *centers,
                           const double *areas, const double *ff,
                           const int64_t *p2o, const double *air,
                           const int64_t *wall_ids, int64_t S,
                           const double *brdf, const int64_t *brdf_index,
                           const double *in_dirs, int32_t *lag, int32_t *dir,
                           double *coef, void *stream);

/* _energy_exchange_init_energy (RadiosityFast.py:1037-1070): order 0.
 * e0dir[P,D,B] (float64), d0[P] -> prev = tot_hi = histogram with e0dir at
 * bin int(d0/c/dt), tot_lo = residual.  Delays >= T set *err_flag to 1
 * (device int, caller clears it) and are skipped. */
int spr_exchange_init(int64_t P, int64_t D, int64_t B, int64_t T,
                      const double *e0dir, const double *d0, double c,
                      double dt, float *prev, float *tot_hi, float *tot_lo,
                      int *err_flag, void *stream);

/* One reflection order of _energy_exchange (RadiosityFast.py:1121-1144):
 * cur[j,:,b,t] = sum_i coef[e,b,:] * prev[i, dir[e], b, t - lag[e]]
 * (per-order sums accumulated in float64), then tot += cur. */
int spr_exchange_step(int64_t P, int64_t D, int64_t B, int64_t T,
                      const int64_t *row_ptr, const int32_t *src,
                      const int32_t *lag, const int32_t *dir,
                      const double *coef, const float *prev, float *cur,
                      float *tot_hi, float *tot_lo, void *stream);

/* Grouped plan.  The reference builds every coefficient row as
 * F~[i,j,:,b] = (F~_ij * exp(-air_b |u_ij|)) * brdf[bi(w_i), s_ij, :, b]
 * (RadiosityFast.py:1255-1270), so sources of one destination that share
 * the BRDF row (bi(w_i), s_ij) can be summed with their scalar weights first
 * and spread over the D directions once per group.
 * Pass 1: src[Vd] (plan rows, in/out) is reordered within each row by the
 * row key bi(w_i)*S + s_ij (stable), key[Vd] receives the keys and
 * grp_ptr[P+1] the exclusive scan of the group counts (grp_ptr[P] = G).
 * max_row = longest row (row_ptr[j+1]-row_ptr[j]); rows up to 8192 entries
 * are supported (SPR_ERR_UNSUPPORTED otherwise).  in_dirs may be NULL (no
 * BRDF: one group per row). */
int spr_plan_grouped_sort(int64_t P, int64_t max_row, const int64_t *row_ptr,
                          int32_t *src, const double *centers,
                          const int64_t *wall_ids, int64_t S,
                          const int64_t *brdf_index, const double *in_dirs,
                          int32_t *key, int64_t *grp_ptr, void *stream);

/* Pass 2: grp_start[G+1] (int64, first entry of each group, grp_start[G] =
 * Vd), grp_key[G] (int32), rowid[Vd] (int32, src*D + p2o[src,j]),
 * lag[Vd] (int32: the delay, with bit 30 set on the last entry of each
 * group), scal[B,Vd] (float64, band-major, F~_ij * exp(-air_b |u_ij|)) and
 * rtab[NK,B,DP] (float64, the BRDF rows with NK = NBRDF*S; brdf may be NULL
 * with D = NK = 1).
 * Replaces, with spr_plan_grouped_sort, the per-pair terms of one order of
 * _energy_exchange, RadiosityFast.py:1121-1144. */
int spr_plan_grouped_fill(int64_t P, int64_t D, int64_t B, double c,
                          double dt, int64_t Vd, const int64_t *row_ptr,
                          const int32_t *src, const int32_t *key,
                          const int64_t *grp_ptr, const double *centers,
                          const double *areas, const double *ff,
                          const int64_t *p2o, const double *air, int64_t NK,
                          const double *brdf, int64_t *grp_start,
                          int32_t *grp_key, int32_t *rowid, int32_t *lag,
                          double *scal, double *rtab, void *stream);

/* The two passes of spr_exchange_step_grouped, separately:
 * gsum[g,b,t] = sum_{e in g} scal[b,e] * prev[rowid[e], b, t - lag[e]].
 * Replaces, with a spread, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144. */
int spr_exchange_gather_grouped(int64_t P, int64_t B, int64_t T, int64_t Vd,
                                const int64_t *row_ptr,
                                const int64_t *grp_ptr, const int32_t *rowid,
                                const int32_t *lag, const double *scal,
                                const float *prev, float *gsum, void *stream);

/* cur[j,:,b,t] = sum_{g of j} rtab[grp_key[g], b, :] * gsum[g,b,t] and the
 * compensated running total tot_hi + tot_lo += cur.
 * Replaces, with a gather, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144. */
int spr_exchange_spread_grouped(int64_t P, int64_t D, int64_t B, int64_t T,
                                const int64_t *grp_ptr,
                                const int32_t *grp_key, const double *rtab,
                                const float *gsum, float *cur, float *tot_hi,
                                float *tot_lo, void *stream);

/* As spr_exchange_spread_grouped, adding two orders to the running total
 * at once: tot_hi + tot_lo += prev + cur, where prev (the previous order,
 * not yet in the total; its own spread ran without the total) must not
 * alias cur.  Reading and writing the total every other order instead of
 * every order saves 12 bytes per state element per two orders.
 * Replaces, with a gather, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144 (two orders' totals). */
int spr_exchange_spread_grouped_pair(int64_t P, int64_t D, int64_t B,
                                     int64_t T, const int64_t *grp_ptr,
                                     const int32_t *grp_key,
                                     const double *rtab, const float *gsum,
                                     const float *prev, float *cur,
                                     float *tot_hi, float *tot_lo,
                                     void *stream);

/* Spread pass of the diffuse + specular plan: the same cur (and total) as
 * spr_exchange_spread_grouped on that plan, bit for bit, formed as
 * cur[j,d] = gsum[diffuse group of j] + gsum[group of j with key d]
 * instead of D multiply-adds per group.  prev, tot_hi, tot_lo as in the
 * two grouped spreads (all NULL: histogram only; prev NULL: add cur to the
 * total; else add prev + cur).
 * Replaces, with spr_exchange_gather_ds, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144. */
int spr_exchange_spread_ds(int64_t P, int64_t D, int64_t B, int64_t T,
                           const int64_t *grp_ptr, const int32_t *grp_key,
                           const float *gsum, const float *prev, float *cur,
                           float *tot_hi, float *tot_lo, void *stream);

/* tot_hi + tot_lo += x[n] (compensated, float64 arithmetic): adds an order
 * left out of the total by the pairing above.
 * The running total of _energy_exchange (RadiosityFast.py:1144). */
int spr_exchange_total_add(int64_t n, const float *x, float *tot_hi,
                           float *tot_lo, void *stream);

/* One reflection order on a grouped plan (same result as
 * spr_exchange_step on the equivalent generic plan): a gather pass writes
 * the per-group sums gsum[G,B,T] (float32 workspace), a spread pass turns
 * them into cur and the running total.
 * Replaces one order of _energy_exchange, RadiosityFast.py:1121-1144. */
int spr_exchange_step_grouped(int64_t P, int64_t D, int64_t B, int64_t T,
                              int64_t Vd, const int64_t *row_ptr,
                              const int64_t *grp_ptr, const int32_t *grp_key,
                              const int32_t *rowid, const int32_t *lag,
                              const double *scal, const double *rtab,
                              float *gsum, const float *prev, float *cur,
                              float *tot_hi, float *tot_lo, void *stream);

/* Diffuse + specular plan.  create_from_scattering (brdf.py:8-130) makes
 * every BRDF row a constant s/pi (1-alpha) on all outgoing directions plus
 * one specular entry, so F~[i,j,:,b] = w_a * 1 + w_c * e_m with one
 * specular direction m per row: each entry then costs two multiply-adds
 * instead of D, whatever the number of distinct BRDFs.
 * Analysis: for the NK = NBRDF*S rows of brdf[NBRDF,S,D,B] (NULL: D = 1,
 * no BRDF), a[NK,B] and c[NK,B] (float64) and spec_dir[NK] (int32, D when
 * the row has no specular part); *ok (int32, device) is 1 if every row has
 * this form, else 0 (then use the grouped or generic plan). */
int spr_brdf_diffuse_specular(int64_t NK, int64_t D, int64_t B,
                              const double *brdf, double *a, double *c,
                              int32_t *spec_dir, int32_t *ok, void *stream);

/* Pass 1 of the plan: like spr_plan_grouped_sort with the entry key
 * spec_dir[bi(w_i)*S + s_ij]; entries without a specular part (key D)
 * form no group, and each non-empty row gets one more group, its diffuse
 * group, last in the row.
 * Replaces, with spr_plan_ds_fill, the per-pair terms of one order of
 * _energy_exchange, RadiosityFast.py:1121-1144. */
int spr_plan_ds_sort(int64_t P, int64_t D, int64_t max_row,
                     const int64_t *row_ptr, int32_t *src,
                     const double *centers, const int64_t *wall_ids,
                     int64_t S, const int64_t *brdf_index,
                     const double *in_dirs, const int32_t *spec_dir,
                     int32_t *key, int64_t *grp_ptr, void *stream);

/* Pass 2: as spr_plan_grouped_fill, with per-band weights scal[B,Vd] =
 * F~_ij exp(-air_b |u|) c[k,b] (specular) and dscal[B,Vd] = ... a[k,b]
 * (diffuse), group keys m (specular direction) or D+1 (the diffuse group,
 * grp_start = row end), and rtab[D+2,B,DP] (unit vectors, a zero row and a
 * row of ones), to be used with spr_exchange_gather_ds and
 * spr_exchange_spread_grouped.  Optional (may be NULL): krow[Vd] (int32,
 * the BRDF row k of each entry) and wgt[B,Vd] (float64, the pair factor
 * F~_ij exp(-air_b |u|) without the BRDF), which the adjoint uses.
 * Replaces, with spr_plan_ds_sort, the per-pair terms of one order of
 * _energy_exchange, RadiosityFast.py:1121-1144. */
int spr_plan_ds_fill(int64_t P, int64_t D, int64_t B, double c, double dt,
                     int64_t Vd, const int64_t *row_ptr, const int32_t *src,
                     const int32_t *key, const int64_t *grp_ptr,
                     const double *centers, const double *areas,
                     const double *ff, const int64_t *p2o, const double *air,
                     const int64_t *wall_ids, int64_t S,
                     const int64_t *brdf_index, const double *in_dirs,
                     const double *a, const double *cs, int64_t *grp_start,
                     int32_t *grp_key, int32_t *rowid, int32_t *lag,
                     double *scal, double *dscal, double *rtab,
                     int32_t *krow, double *wgt, void *stream);

/* Gather pass of the diffuse + specular plan: gsum rows of the specular
 * groups as spr_exchange_gather_grouped with scal, and the row's diffuse
 * group = sum over all its entries with dscal.
 * Replaces, with spr_exchange_spread_ds, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144. */
int spr_exchange_gather_ds(int64_t P, int64_t B, int64_t T, int64_t Vd,
                           const int64_t *row_ptr, const int64_t *grp_ptr,
                           const int32_t *rowid, const int32_t *lag,
                           const double *scal, const double *dscal,
                           const float *prev, float *gsum, void *stream);

/* Destination tiles of the grouped plan (the _energy_exchange loop,
 * RadiosityFast.py:1073-1145, as grouped above).  A tile is JT (2..16)
 * consecutive rows of a plan (a caller may permute and pad the rows so
 * that a tile is a block of neighbouring patches); its entries are merged
 * into distinct (key, source row) entries, in key then source-row order,
 * each with one delay word and one weight per slot (row of the tile), so
 * the JT gathers of one source row share one load.
 *
 * Pass 1: skey[Vd], sorted[Vd], payload[Vd] (int64 workspace), first[Vd+1]
 * (int64: exclusive count of distinct entries before each merged position;
 * first[Vd] = M, the number of merged entries) and tptr[ceil(P/JT)+1]
 * (merged CSR of the tiles).  skey[e] = (key << 32) | rowid[e] is computed
 * from the groups (grp_start, grp_key), or, when both are NULL, given by
 * the caller (the diffuse + specular plan keys entries by their own
 * specular direction). */
int spr_plan_tile_order(int64_t P, int64_t JT, int64_t Vd, int64_t G,
                        const int64_t *row_ptr, const int64_t *grp_start,
                        const int32_t *grp_key, const int32_t *rowid,
                        int64_t *skey, int64_t *sorted, int64_t *payload,
                        int64_t *first, int64_t *tptr, void *stream);

/* Pass 2: mrow[M] (int32 source row), mlag[M,JT] (int32 delay per slot,
 * 0x3fffffff for a slot without that entry, bit 30 set where the slot's
 * group ends: on the last merged entry of the key run) and mw[B,M,JT]
 * (float64 weight per band and slot, 0 for a slot without the entry).
 * Replaces, with spr_plan_tile_order, the per-pair terms of one order of
 * _energy_exchange, RadiosityFast.py:1121-1144. */
int spr_plan_tile_fill(int64_t P, int64_t JT, int64_t B, int64_t Vd,
                       int64_t M, const int64_t *row_ptr,
                       const int32_t *rowid, const int32_t *lag,
                       const double *scal, const int64_t *sorted,
                       const int64_t *payload, const int64_t *first,
                       int32_t *mrow, int32_t *mlag, double *mw, void *stream);

/* Pass 3 (16-row tiles): mspan[M,4] (int32) = {source row, largest and
 * smallest delay of the merged entry's slots, 0}, the window each merged
 * entry's source row is staged over by the 16-row gathers.
 * Replaces, with spr_plan_tile_order, the per-pair terms of one order of
 * _energy_exchange, RadiosityFast.py:1121-1144. */
int spr_plan_tile_span(int64_t M, int64_t JT, const int32_t *mrow,
                       const int32_t *mlag, int32_t *mspan, void *stream);

/* Gather pass on the tiles: gsum bit-identical to
 * spr_exchange_gather_grouped (each slot sums its groups in the same
 * order), for use with spr_exchange_spread_grouped.  JT = 4 or 8: the
 * slots' loads of a merged source row are issued together and share cache
 * lines (16-row tiles: spr_exchange_gather_tiled_shape).  grp_ptr[slot] =
 * index of the slot's first group (the plan's grp_ptr for consecutive
 * rows; for permuted rows, that of the slot's destination); P = number of
 * slots.
 * Replaces, with a spread, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144. */
int spr_exchange_gather_tiled(int64_t P, int64_t JT, int64_t B, int64_t T,
                              int64_t M, const int64_t *tptr,
                              const int64_t *grp_ptr, const int32_t *mrow,
                              const int32_t *mlag, const double *mw,
                              const float *prev, float *gsum, void *stream);

/* spr_exchange_gather_tiled on 16-row tiles: one wave per slot, each
 * merged entry's window of its source row (mspan, spr_plan_tile_span)
 * staged once in LDS with the batch's delay words and weights.  R = 16
 * bins per lane (1,024-bin blocks) or 8 (512-bin blocks); the same group
 * sums, bit for bit.
 * Replaces, with a spread, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144. */
int spr_exchange_gather_tiled_shape(int64_t P, int64_t JT, int64_t R,
                                    int64_t B, int64_t T, int64_t M,
                                    const int64_t *tptr,
                                    const int64_t *grp_ptr,
                                    const int32_t *mspan, const int32_t *mlag,
                                    const double *mw, const float *prev,
                                    float *gsum, void *stream);

/* Gather pass of the diffuse + specular plan on 16-row tiles (merged by
 * each entry's specular direction): gsum bit-identical to
 * spr_exchange_gather_ds, for use with spr_exchange_spread_ds.  mw and
 * mdw[B,M,16]: specular and diffuse weights per slot (two
 * spr_plan_tile_fill passes, with scal and with dscal); gstart[slot],
 * gdiff[slot] (int64): the slot's first group and its diffuse group (-1:
 * the slot's row has no entries); mspan from spr_plan_tile_span.
 * Replaces, with spr_exchange_spread_ds, one order of _energy_exchange,
 * RadiosityFast.py:1121-1144. */
int spr_exchange_gather_tiled_ds(int64_t P, int64_t JT, int64_t B, int64_t T,
                                 int64_t M, const int64_t *tptr,
                                 const int64_t *gstart, const int64_t *gdiff,
                                 const int32_t *mspan, const int32_t *mlag,
                                 const double *mw, const double *mdw,
                                 const float *prev, float *gsum,
                                 void *stream);

/* etc[P,D,B,T] (float32) = tot_hi + tot_lo (may alias tot_hi).
 * The total histogram _energy_exchange returns (RadiosityFast.py:1114-1145). */
int spr_exchange_finalize(int64_t n, const float *tot_hi,
                          const float *tot_lo, float *etc, void *stream);

/* Bytes of workspace spr_energy_exchange needs.
 * Sizing helper for spr_energy_exchange; no counterpart in the reference. */
int64_t spr_energy_exchange_workspace(int64_t P, int64_t D, int64_t B,
                                      int64_t T, int64_t V);

/* _energy_exchange (RadiosityFast.py:1073-1145) with the reference's
 * argument list: T, e0dir[P,D,B], d0[P], dist[P,P], fft[P,P,D,B],
 * p2o[P,P] (int64), c, dt, K, pairs[V,2] (int32) -> etc[P,D,B,T]
 * (float32).  workspace: spr_energy_exchange_workspace(...) bytes.
 * When each destination's F~ rows are, band by band, multiples of a few
 * vectors (entries that share a BRDF row; checked entry by entry to 2^-28
 * of the row's largest element) in at most min(ceil(V/2), 2*P*D) groups, it
 * runs the grouped two-pass exchange (group sums, then a spread by the
 * group's row), else the per-pair plan; spr_energy_exchange_groups tells
 * which.  Synchronous: returns SPR_ERR_RANGE if a source delay is >= T
 * (the reference raises IndexError there). */
int spr_energy_exchange(int64_t T, int64_t P, int64_t D, int64_t B,
                        const double *e0dir, const double *d0,
                        const double *dist, const double *fft,
                        const int64_t *p2o, double c, double dt, int64_t K,
                        int64_t V, const int32_t *pairs, float *etc,
                        void *workspace, int64_t workspace_bytes,
                        void *stream);

/* hkey[e] (uint64, Vd entries) = hash of entry e's F~ row fft[src[e], j]
 * (destination j of row_ptr's row holding e) with every band divided by its
 * first element of largest magnitude and rounded to multiples of 2^-30:
 * rows that are band by band multiples of one another hash alike (up to
 * that rounding), the matching that spr_energy_exchange groups by.  A
 * grouped plan built on it must still check every entry against its
 * group's row.  Helper of the grouped plan from a saved dense F~; no
 * counterpart in the reference. */
int spr_plan_dense_hash(int64_t P, int64_t D, int64_t B,
                        const int64_t *row_ptr, const int32_t *src,
                        const double *fft, uint64_t *hkey, void *stream);

/* Groups of the grouped plan the last spr_energy_exchange(_noalloc) call
 * on this thread ran, 0 if it ran the per-pair plan.
 * Diagnostic of spr_energy_exchange; no counterpart in the reference. */
int64_t spr_energy_exchange_groups(void);

/* Scratch built into the library for spr_energy_exchange_noalloc, per GPU
 * (64 MiB: enough for the 148-patch shoebox with 16 directions, one band and
 * 1,000 bins, which needs about 49 MB). */
#define SPR_NOALLOC_SCRATCH_BYTES (64ll << 20)

/* No counterpart in the reference.  spr_energy_exchange for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same arguments minus
 * the workspace, the same kernels and results.  Its workspace is a
 * file-scope device array of SPR_NOALLOC_SCRATCH_BYTES, one per GPU, held
 * until the process ends; the call allocates and frees nothing.  Returns
 * SPR_ERR_WORKSPACE when spr_energy_exchange_workspace(P, D, B, T, V)
 * exceeds SPR_NOALLOC_SCRATCH_BYTES.  Every call on one GPU uses that GPU's
 * array, so calls that overlap on one GPU (other threads or streams) race;
 * serialise them.  Synchronous, like spr_energy_exchange. */
int spr_energy_exchange_noalloc(int64_t T, int64_t P, int64_t D, int64_t B,
                                const double *e0dir, const double *d0,
                                const double *dist, const double *fft,
                                const int64_t *p2o, double c, double dt,
                                int64_t K, int64_t V, const int32_t *pairs,
                                float *etc, void *stream);

/* ---------------------------------------------------------------------
 * Receiver collection
 * ------------------------------------------------------------------- */

/* _patch2receiver_energy_universal (universal.py:149-160), batched:
 * points[NQ,3], patch points[P,4,3], vis[NQ,P] -> factor[NQ,P] (float64). */
int spr_receiver_energy(int64_t NQ, const double *points, int64_t P,
                        const double *patch_points, const uint8_t *vis,
                        double *factor, void *stream);

/* _collect_energy_patches + _collect_receiver_energy
 * (RadiosityFast.py:686-752, 1148-1185) and the patch sum of
 * collect_energy_receiver_mono (RadiosityFast.py:591-592):
 * out[r,(k,)b,t] = sum_k etc[k, idx[r,k], b, (t - n_rk) mod T] *
 *                  factor[r,k] * exp(-air_b * d_rk),
 * n_rk = ceil(d_rk/c/dt), d_rk = |c_k - point_r|.
 * etc[P,D,B,T] (float32), points[NR,3], centers[P,3], idx[NR,P] (int64),
 * factor[NR,P], air[B].  patchwise = 0: out[NR,B,T] (float32, patch sum in
 * float64); patchwise = 1: out[NR,P,B,T] (float32). */
int spr_collect(int64_t NR, int64_t P, int64_t D, int64_t B, int64_t T,
                const float *etc, const double *points,
                const double *centers, const int64_t *idx,
                const double *factor, const double *air, double c, double dt,
                int patchwise, float *out, void *stream);

/* ---------------------------------------------------------------------
 * Adjoint of the exchange (no counterpart in the reference, which has no
 * gradients; the forward it differentiates is _energy_exchange,
 * RadiosityFast.py:1073-1145, on a diffuse + specular plan).  Tables
 * a[NK,B] (diffuse constant) and cs[NK,B] (specular part) are the BRDF
 * rows in the form spr_brdf_diffuse_specular returns.
 * ------------------------------------------------------------------- */

/* Plan weights from tables: scal[b,e] = wgt[b,e] cs[krow_e, b],
 * dscal[b,e] = wgt[b,e] a[krow_e, b] (krow, wgt from spr_plan_ds_fill).
 * Adjoint support; no counterpart in the reference. */
int spr_ds_weights(int64_t B, int64_t Vd, const int32_t *krow,
                   const double *wgt, const double *a, const double *cs,
                   double *scal, double *dscal, void *stream);

/* Injection from tables (_add_directional, RadiosityFast.py:988-1034):
 * k0[P] (int32) = BRDF row of each patch for the direction of the source,
 * e0dir[P,D,B] = energy0[P,B] (a + cs e_m)[k0, b] with m = spec_dir[k0].
 * energy0 is spr_source_energy's output before the BRDF. */
int spr_ds_injection(int64_t P, int64_t D, int64_t B, const double *source,
                     const double *centers, const int64_t *wall_ids,
                     int64_t S, const int64_t *brdf_index,
                     const double *in_dirs, const double *energy0,
                     const double *a, const double *cs,
                     const int32_t *spec_dir, int32_t *k0, double *e0dir,
                     void *stream);

/* Source-major view of a diffuse + specular plan.  Pass 1: trow_ptr[P+1]
 * (int64) = exclusive scan of the entries per source (rowid / D).
 * Adjoint support; no counterpart in the reference. */
int spr_plan_ds_transpose_rows(int64_t P, int64_t D, const int64_t *row_ptr,
                               const int32_t *rowid, int64_t *trow_ptr,
                               void *stream);

/* Pass 2 (max_trow = longest source row, at most 8192): for transposed
 * entry q of source i, sorted by (direction of the source row, destination):
 * te[q] (forward entry index), tj[q] (destination), tlag[q] (delay, bit 30
 * on the last entry of a direction group), tm[q] (specular direction of
 * the entry, 0 if none), tgptr[P*D+1] (int64: first entry of source i with
 * direction >= d at i*D + d), tkey[Vd] (int32: direction << 25 |
 * destination).  cursor[P] (int64) is workspace; key[Vd] is the plan's
 * entry key (specular direction or D).
 * Adjoint support; no counterpart in the reference. */
int spr_plan_ds_transpose_fill(int64_t P, int64_t D, int64_t Vd,
                               int64_t max_trow, const int64_t *row_ptr,
                               const int32_t *rowid, const int32_t *lag,
                               const int32_t *key, const int64_t *trow_ptr,
                               int64_t *cursor, int32_t *te, int32_t *tkey,
                               int32_t *tj, int32_t *tlag, int32_t *tm,
                               int64_t *tgptr, void *stream);

/* Transposed weights: twa[b,q] = dscal[b, te[q]], twc[b,q] = scal[b,te[q]].
 * Adjoint support; no counterpart in the reference. */
int spr_ds_transposed_weights(int64_t B, int64_t Vd, const int32_t *te,
                              const double *scal, const double *dscal,
                              double *twa, double *twc, void *stream);

/* Adjoint of spr_collect (mono): lam[P,D,B,T] (float32, zero-filled here)
 * = dL/d(total histogram) for grad[NR,B,T] = dL/d(receiver histograms);
 * the other arguments as for spr_collect.
 * Adjoint of RadiosityFast.py:1148-1185; no counterpart in the reference. */
int spr_collect_adjoint(int64_t NR, int64_t P, int64_t D, int64_t B,
                        int64_t T, const float *grad, const double *points,
                        const double *centers, const int64_t *idx,
                        const double *factor, const double *air, double c,
                        double dt, float *lam, void *stream);

/* musum[P,B,T] = sum over d of mu[P,D,B,T] (float64 sum).
 * Adjoint support; no counterpart in the reference. */
int spr_direction_sum(int64_t P, int64_t D, int64_t B, int64_t T,
                      const float *mu, float *musum, void *stream);

/* One backward order: muout = lam + A^T munext (the transposed exchange;
 * muout is first set to lam, then each direction row with entries gets its
 * sum added), with musum = spr_direction_sum(munext) and ek the forward
 * order-k histogram; adds each transposed entry's table-gradient dot
 * products to ga[2,B,Vd] (float64): ga[0,b,q] += sum_t musum[j,b,t+n]
 * ek[i,d,b,t], ga[1,b,q] += sum_t munext[j,m,b,t+n] ek[i,d,b,t].  tkey is
 * spr_plan_ds_transpose_fill's tkey (direction << 25 | destination), kept
 * after the fill.  muout must not alias an input.  T < 2^29, P < 2^25.
 * Adjoint of RadiosityFast.py:1121-1144; no counterpart in the reference. */
int spr_exchange_adjoint_ds(int64_t P, int64_t D, int64_t B, int64_t T,
                            int64_t Vd, const int64_t *tgptr,
                            const int32_t *tkey, const int32_t *tlag,
                            const int32_t *tm, const double *twa,
                            const double *twc, const float *lam,
                            const float *musum, const float *munext,
                            const float *ek, float *muout, double *ga,
                            void *stream);

/* Table gradients from ga: gA[k,b] += wgt[b,e] ga[0,b,q] and gC[k,b] +=
 * wgt[b,e] ga[1,b,q] for e = te[q], k = krow[e] (float64 atomics).
 * Adjoint support; no counterpart in the reference. */
int spr_ds_table_grad(int64_t B, int64_t Vd, const int32_t *te,
                      const int32_t *krow, const double *wgt,
                      const double *ga, double *gA, double *gC,
                      void *stream);

/* Injection part of the table gradients, from mu0 = dL/dE_0[P,D,B,T]:
 * gA[k0_p,b] += energy0[p,b] sum_d mu0[p,d,b,n_p], gC[k0_p,b] +=
 * energy0[p,b] mu0[p,m,b,n_p] with n_p = int(d0[p]/c/dt).
 * Adjoint of RadiosityFast.py:988-1034; no counterpart in the reference. */
int spr_ds_injection_grad(int64_t P, int64_t D, int64_t B, int64_t T,
                          const double *energy0, const double *d0, double c,
                          double dt, const int32_t *k0,
                          const int32_t *spec_dir, const float *mu0,
                          double *gA, double *gC, void *stream);

/* Adjoint of the grouped plan (BRDF rows of any form; rtab[k,b,d] =
 * table[k,d,b]).  nu[G,B,T] (float32) = the transposed spread of mu =
 * dL/dE_{k+1}[P,D,B,T]: nu[g,b,t] = sum_d rtab[key_g,b,d] mu[j_g,d,b,t]
 * (float64 sum), j_g the destination of group g (grp_ptr).  T < 2^29.
 * Adjoint of the spread half of RadiosityFast.py:1121-1144 (as grouped by
 * spr_plan_grouped_fill); no counterpart in the reference. */
int spr_grouped_spread_adjoint(int64_t P, int64_t D, int64_t B, int64_t T,
                               const int64_t *grp_ptr, const int32_t *grp_key,
                               const double *rtab, const float *mu, float *nu,
                               void *stream);

/* Per-group table gradient of one order: gg[g,b,d] += sum_t mu[j_g,d,b,t]
 * gsum[g,b,t] (gg float64 [G,B,DP], DP = spr_padded_directions(D);
 * accumulated over orders, each element owned by one lane), with gsum the
 * forward order's group sums (spr_exchange_gather_grouped of E_k) and mu =
 * dL/dE_{k+1}.  T < 2^29.
 * Adjoint support; no counterpart in the reference. */
int spr_grouped_table_grad(int64_t P, int64_t D, int64_t B, int64_t T,
                           const int64_t *grp_ptr, const float *mu,
                           const float *gsum, double *gg, void *stream);

/* Transposed gather: muout[r,b,t] = lam[r,b,t] + sum_{q in T(r)}
 * tscal[b,q] nu[tgrp_q,b,t+tlag_q] (float64 sum in q order; bins past T
 * add nothing), r = i*D + d a source row, T(r) = [trow_ptr[r],
 * trow_ptr[r+1]) its plan entries in entry order with their group tgrp,
 * delay tlag (no flag) and weights tscal[B,Vd].  muout must not alias lam.
 * Adjoint of the gather half of RadiosityFast.py:1121-1144; no
 * counterpart in the reference. */
int spr_grouped_gather_adjoint(int64_t P, int64_t D, int64_t B, int64_t T,
                               int64_t Vd, const int64_t *trow_ptr,
                               const int32_t *tgrp, const int32_t *tlag,
                               const double *tscal, const float *nu,
                               const float *lam, float *muout, void *stream);

/* gt[key_g,d,b] += gg[g,b,d] over the G groups (float64 atomics): the
 * table gradient [NK,D,B] from spr_grouped_table_grad's sums.
 * Adjoint support; no counterpart in the reference. */
int spr_grouped_key_grad(int64_t G, int64_t D, int64_t B,
                         const int32_t *grp_key, const double *gg, double *gt,
                         void *stream);

/* Injection part of the table gradient [NK,D,B] for E_0dir[p,d,b] =
 * energy0[p,b] table[k0_p,d,b]: gt[k0_p,d,b] += energy0[p,b]
 * mu0[p,d,b,n_p], n_p = int(d0[p]/c/dt) (float64 atomics).
 * Adjoint of RadiosityFast.py:988-1034; no counterpart in the reference. */
int spr_injection_table_grad(int64_t P, int64_t D, int64_t B, int64_t T,
                             const double *energy0, const double *d0,
                             double c, double dt, const int32_t *k0,
                             const float *mu0, double *gt, void *stream);

#ifdef __cplusplus
}
#endif
#endif
