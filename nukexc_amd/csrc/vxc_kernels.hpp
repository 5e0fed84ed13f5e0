// Launchers of the batched V_xc kernels (kernels_vxc.hip).
#pragma once

#include "nkx_internal.hpp"

namespace nkx {

void vk_gather_p(hipStream_t s, int nbf, int nbf_pad, const int *fperm, const double *P,
                 double *Pd);
void vk_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int narr, int gga,
               double *tiles);
// group: batches per workgroup (1, 2 or 4); epoint (may be NULL): w e_xc
// per sorted point; rhopt (may be NULL, nspin 1): rho and grad rho per sorted
// point
void vk_density(hipStream_t s, int nspin, int group, int64_t b0, int64_t b1, int64_t s0,
                const Plan &pl, const double *Pd, int narr, int gga, double *tiles,
                const XcList &xl, double *coef, double *blockEN, double *epoint,
                double *rhopt = nullptr);
// screened-grid set-up (kernels_sparse.hip): Morton order of the points
// (AoS; w, owner may be NULL) in place and the box of each run of per points
// (6 doubles per box); the CSR list of functions (host centres and radii)
// meeting each box, offsets always, neighbors when not NULL; returns its length
void sparse_bounding_boxes(hipStream_t s, int64_t npts, double *pts, double *w, int *owner,
                           int per, double *boxes);
int64_t sparse_neighbor_list(hipStream_t s, int64_t nbox, const double *boxes, int nbf,
                             const double *ctr_h, const double *rad_h, int *offsets,
                             int *neighbors);
// P = C diag(f) C^T (C column-major nbf x nocc)
void vk_density_matrix(hipStream_t s, int nbf, int nocc, const double *C, const double *f,
                       double *P);
// rho, grad rho per sorted point (4 per padded point) to the grid's order
void vk_scatter_density(hipStream_t s, const Plan &pl, const double *r4, double *rho, double *gx,
                        double *gy, double *gz, double *sigma);
void vk_zoverlap(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int narr,
                 double *tiles);
// nz Z arrays (the last nz of the narr per slot), partials for Z array k at
// vpart + k * nsplit * npairs * 64 * 64
void vk_vgemm(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int nsplit,
              int narr, int nz, const double *tiles, double *vpart);
// V of molecule m is Vmol_dev[m] (device array of device pointers), its
// function count mol_nbf_dev[m]
void vk_vreduce(hipStream_t s, const Plan &pl, int nsplit, const double *vpart,
                double *const *Vmol_dev, const int *mol_nbf_dev);
// out[3 m + c] = (E, N_alpha, N_beta) of molecule m
void vk_en_reduce(hipStream_t s, const Plan &pl, const double *en, double *out);
void vk_fill0(hipStream_t s, double *x, int64_t n);

// mixed form (kernels_vxc_i8.hip): int8 slices with per-row exponents,
// exact integer MFMA; the store holds mixed_slot_bytes(gga) per live slot and
// the gathered P mixed_p_bytes(nfb)
int64_t mixed_slot_bytes(int gga);
int64_t mixed_p_bytes(int nfb);
void vkm_gather_p(hipStream_t s, int nbf, int nfb, const int *fperm, const double *P,
                  signed char *Ps);
void vkm_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int gga,
                signed char *store);
// Pm sections of a batch's exponent groups re-sliced to shared point exponents
void vkm_pm_group(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int gga,
                  signed char *store);
// density shape codes of the int8 form (other values run as 65)
bool mixed_density_shape_ok(int shape);
void vkm_density(hipStream_t s, int shape, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                 const signed char *Ps, int gga, signed char *store, const XcList &xl,
                 double *coef, double *blockEN, double *part);
void vkm_density_finish(hipStream_t s, int shape, int64_t b0, int64_t b1, int64_t s0,
                        const Plan &pl, int gga, signed char *store, const XcList &xl,
                        double *coef, double *blockEN, const double *part);
// per-batch reduction of part and the functional only (no Z slices);
// epoint (may be NULL): w e_xc per sorted point
void vkm_density_functional(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                            int gga, const XcList &xl, double *coef, double *blockEN,
                            const double *part, double *epoint);
// zstore_b (may be NULL): Z sections of the second spin, mixed_zsection_bytes()
// per slot, instead of the slots' own
void vkm_vgemm(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int nsplit,
               int gga, const signed char *store, double *vpart,
               const signed char *zstore_b = nullptr);
// spin-polarised functional and Z slices of both spins (split density form):
// Z_alpha into the slots, Z_beta into zstore_b
// (store NULL: the functional only; epoint, may be NULL: w e_xc per point)
void vkm_density_finish_pol(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                            int gga, signed char *store, signed char *zstore_b,
                            const XcList &xl, double *coef_a, double *coef_b,
                            double *blockEN, const double *part_a, const double *part_b,
                            double *epoint = nullptr);
int64_t mixed_zsection_bytes();

// split-bf16 mixed form (kernels_vxc_bf16.hip): three bf16 planes per value,
// six products; bf16_slot_bytes(gga) per slot, 3 nbf_pad^2 bf16 for P
int64_t bf16_slot_bytes(int gga);
void vkb_gather_p(hipStream_t s, int nbf, int nfb, const int *fperm, const double *P,
                  uint16_t *Ps);
void vkb_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int gga, uint16_t *store);
void vkb_density(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                 const uint16_t *Ps, int gga, uint16_t *store, const XcList &xl, double *coef,
                 double *blockEN);
void vkb_vgemm(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
               int nsplit, int gga, const uint16_t *store, double *vpart);

// nuclear gradient (kernels_grad.hip)
// grad_basis: per batch, per-function partial sums of the basis-centre
// derivative over the batch (cpart[(slot - s0) * 64 * 3 + mu * 3 + k]) and the
// per-point sum tpt[3 * gp + k] with the sign of the point-motion term.
// coef of the spin of Pd with a stride of cstride doubles per point; accum
// adds into tpt (second spin) instead of writing it
void vkg_basis(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
               const double *Pd, int narr, int gga, const double *tiles, const double *coef,
               int cstride, bool accum, double *cpart, double *tpt);
// mixed form (kernels_grad_i8.hip), on the int8 slot store after vkm_colloc,
// vkm_density (split shape) and vkm_density_functional: Y = 2 a Phi + b . grad
// Phi slices per slot (rows = points), then X = P Phi and W = P Y on integer
// MFMA with the same per-function and per-point sums as vkg_basis
// ystore_b (may be NULL): Y sections of the second spin, SEC bytes per slot,
// instead of the slots' Z sections; accum as vkg_basis
void vkgm_y_slots(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int gga,
                  const double *coef, signed char *store, signed char *ystore_b = nullptr);
void vkgm_basis(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                const signed char *Ps, int gga, const signed char *store, const double *coef,
                const signed char *ystore_b, bool accum, double *cpart, double *tpt);
// fold the per-slot partials of batches [b0, b1) into gacc[seg][nbf_pad * 3]
void vkg_fold(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int nseg,
              const double *cpart, double *gacc);
// pown[sp] = owner of sorted point sp (-1 for padding)
void vkg_owners(hipStream_t s, const Plan &pl, const int *owner, int *pown);
// point-motion and Becke weight-derivative terms per atom; partition 0 = fixed
// weights, 1 = capped Becke, 2 = Becke.  wpart[block][3 natoms], nblk blocks;
// ptab: 2 natoms^2 doubles of scratch for the per-pair table.
void vkg_weights(hipStream_t s, const Plan &pl, int natoms, const double *centers,
                 const int *pown, const double *epoint, const double *tpt, int partition,
                 int nblk, double *ptab, double *wpart);
// out[k] = sum over rows r < nrow of in[r * ncol + k], fixed order
void vkg_colsum(hipStream_t s, int nrow, int ncol, const double *in, double *out);

}  // namespace nkx
