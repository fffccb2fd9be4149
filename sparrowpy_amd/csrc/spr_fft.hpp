// Directional form factor F~[i,j,:,:] of one directed patch pair, shared
// by the dense bake (spr_fft_dense) and the exchange plan builder
// (spr_plan_from_geometry) so both produce identical bits.
//
// Reference: _form_factors_with_directivity_dim (RadiosityFast.py:1251-1270)
// with get_scattering_data_source (RadiosityFast.py:1386-1390).  The BRDF
// and incoming-direction set are those of the SENDING patch's wall and the
// air factor uses the length of the unit vector (not the pair distance),
// as in the reference.
#pragma once

#include "spr_common.hpp"

namespace spr {

struct PairFactor {
  double f;           // F~_ij (reciprocity-scaled form factor)
  double distance;    // |unit(c_i - c_j)|, the air-attenuation length used
  const double *row;  // brdf[bi(w_i), s_ij, :, :] (D*B values) or nullptr
};

__device__ __forceinline__ PairFactor pair_factor(
    int64_t i, int64_t j, int64_t P, const double *centers,
    const double *areas, const double *ff, const int64_t *wall_ids,
    int64_t S, int64_t D, int64_t B, const double *brdf,
    const int64_t *brdf_index, const double *in_dirs) {
  d3 diff = load3(centers + 3 * i) - load3(centers + 3 * j);
  diff = div(diff, norm(diff));
  PairFactor pf;
  pf.f = (i < j) ? ff[i * P + j] : (ff[j * P + i] * areas[j]) / areas[i];
  pf.distance = norm(diff);
  pf.row = nullptr;
  if (brdf) {
    int64_t w = wall_ids[i];
    int64_t s = nearest_direction(in_dirs + w * S * 3, S, diff);
    pf.row = brdf + (brdf_index[w] * S + s) * D * B;
  }
  return pf;
}

// F~[i,j,d,b] from a PairFactor; air_b = exp(-air[b] * distance) or
// ignored when has_air is false.
__device__ __forceinline__ double pair_coef(const PairFactor &pf,
                                            double air_b, bool has_air,
                                            int64_t d, int64_t b,
                                            int64_t B) {
  double v = pf.f;
  if (has_air) v = v * air_b;
  if (pf.row) v = v * pf.row[d * B + b];
  return v;
}

}  // namespace spr
