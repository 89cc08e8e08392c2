// sph_kernels.hpp -- host launchers of the gfx950 kernels (sph_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "sph_common.hpp"

namespace sph {

// state masks for the box state filter of the passes
constexpr int MASK_STEP = (1 << BOX_INIT) | (1 << BOX_RUN);
constexpr int MASK_RUN = (1 << BOX_RUN);
constexpr int MASK_INIT = (1 << BOX_INIT);
constexpr int MASK_ALL = (1 << BOX_INIT) | (1 << BOX_RUN) | (1 << BOX_DONE);

// build stage: slot owners (on a layout change), then plan, keys, sort,
// permutation, cell starts and lists of the building boxes
void launch_slot_box(const ArenaView &A, hipStream_t st);
void launch_build_plan(const ArenaView &A, int force_mask, hipStream_t st);
void launch_keys(const ArenaView &A, int force_mask, hipStream_t st);
// cells != 0: also the cell starts from the sorted keys (radix-sort path)
void launch_permute(const ArenaView &A, int force_mask, int cells, hipStream_t st);
// arenas whose boxes all fit box_sort_fits: per-box counting sort in place of
// keys + radix sort + cell starts (same order)
void launch_box_sort(const ArenaView &A, int force_mask, int nbox, hipStream_t st);
bool box_sort_fits(int ncells, int n);
void launch_build_lists(const ArenaView &A, int force_mask, int fp32, hipStream_t st);
void launch_density(const ArenaView &A, int fp32, int state_mask, int unroll, hipStream_t st);
void launch_grad(const ArenaView &A, int fp32, int state_mask, int diag, int unroll, hipStream_t st);
// step != 0: the stepping launch, which also takes each running slot's
// leapfrog step into the second state buffers (and carries the other live
// slots over); the runtime swaps the buffers after it
void launch_rates(const ArenaView &A, int fp32, int state_mask, int unroll, int step, hipStream_t st);
// the step pass as a launch of its own, after launch_rates(step = 0)
void launch_leapfrog(const ArenaView &A, int fp32, hipStream_t st);
void launch_box_step(const ArenaView &A, int nbox, hipStream_t st);
void launch_upload(const ArenaView &A, int off, int n, int field, const double *src, hipStream_t st);
void launch_init_slots(const ArenaView &A, int off, int n, hipStream_t st);
void launch_download(const ArenaView &A, int box, int n, int field, double *dst, hipStream_t st);
void launch_inverse(const int *ids, int off, int n, int *inv, hipStream_t st);
void launch_pairs(const ArenaView &A, int box, int n, const int *inv, unsigned *count,
                  const unsigned *offs, int *jbuf, int pass, hipStream_t st);
void launch_pairs_emit(int n, const unsigned *offs, const unsigned *count, const int *jbuf,
                       long long *iap, long long cap, hipStream_t st);
void launch_separations(const ArenaView &A, int box, const int *inv, long long np,
                        const long long *iap, double *drij, double *rij, double *rsq, double *dv,
                        hipStream_t st);
void launch_lucy(long long np, const double *r, const double *dx, double h, int dim, double *w,
                 double *dwdx, hipStream_t st);
void launch_min_image(long long np, double *dr, double Lx, double Ly, hipStream_t st);
void launch_set_u(const ArenaView &A, int box, int n, const double *T, double T0, hipStream_t st);
void launch_energy_now(const ArenaView &A, int box, double *out3, hipStream_t st);
void launch_fp32_accumulate(long long n, const double *x, int unroll, double *out2, hipStream_t st);
void launch_splash(long long n, const double *r, const double *m, const double *rho, const double *a,
                   const double *h, const SplashGrid &G, double4 *rec, double4 *rec_sorted, unsigned *keys,
                   unsigned *k_out, unsigned *v_out, unsigned *k_alt, unsigned *v_alt, void *sort_tmp, int key_bits,
                   int *cs, int nx, int ny, const double *x, const double *y, double *Z, hipStream_t st);

}  // namespace sph
