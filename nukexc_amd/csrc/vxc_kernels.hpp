// Launchers of the batched V_xc kernels (kernels_vxc.hip).
#pragma once

#include "nkx_internal.hpp"

namespace nkx {

void vk_gather_p(hipStream_t s, int nbf, int nbf_pad, const int *fperm, const double *P,
                 double *Pd);
void vk_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int narr, int gga,
               double *tiles);
void vk_density(hipStream_t s, int nspin, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                const double *Pd, int narr, int gga, double *tiles, const XcList &xl,
                double *coef, double *blockEN);
void vk_zoverlap(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int narr,
                 double *tiles);
// nz Z arrays (the last nz of the narr per slot), partials for Z array k at
// vpart + k * nsplit * npairs * 64 * 64
void vk_vgemm(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int nsplit,
              int narr, int nz, const double *tiles, double *vpart);
void vk_vreduce(hipStream_t s, const Plan &pl, int nsplit, const double *vpart, double *V);
void vk_en_reduce(hipStream_t s, int64_t nb, const double *en, double *out);

// mixed form (kernels_vxc_mixed.hip): pre-split bf16 planes in MFMA fragment
// order; store holds mixed_slot_bytes(gga) per live slot
int64_t mixed_slot_bytes(int gga);
void vkm_gather_p(hipStream_t s, int nbf, int nfb, const int *fperm, const double *P,
                  uint16_t *Ps);
void vkm_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int gga, uint16_t *store);
void vkm_density(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                 const uint16_t *Ps, int gga, uint16_t *store, const XcList &xl, double *coef,
                 double *blockEN);
void vkm_vgemm(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
               int nsplit, int gga, const uint16_t *store, double *vpart);

}  // namespace nkx
