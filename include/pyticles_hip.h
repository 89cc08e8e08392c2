/*
 * pyticles_hip.h -- C ABI of the MI355X (gfx950) smooth-particle library.
 *
 * Plain C types only: pointers, sizes, a POD parameter struct and int status
 * codes (0 = OK).  Every call that fails sets a message readable through
 * sph_last_error().  Host arrays are C-order float64 (int64 for pair
 * indices), borrowed only for the duration of the call.  All particle
 * arrays are 2-D: r[n][2], v[n][2], gradv[n][2][2], P[n][2][2], jq[n][2].
 *
 * The reference (andrewcharles/pyticles) has no C interface: its operator
 * API is a set of Python objects sharing NumPy arrays.  Each entry point
 * below names the reference function whose work it replaces; the Python
 * package pyticles_amd re-exposes them under the reference's names.
 */
#ifndef PYTICLES_HIP_H
#define PYTICLES_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SPH_OK 0
#define SPH_ERR_ARG 1      /* invalid argument or size                         */
#define SPH_ERR_HIP 2      /* HIP runtime failure                              */
#define SPH_ERR_NOMEM 3    /* device allocation or arena budget exhausted     */
#define SPH_ERR_CAP 4      /* caller-supplied capacity too small (pairs)      */
#define SPH_ERR_STATE 5    /* call not valid in the current state             */
#define SPH_ERR_NUMERIC 6  /* non-finite value or EOS pole (rho*b >= 1)       */

/* rate terms (bit mask) */
#define SPH_PRESSURE 1u    /* repulsive vdW pressure  (forces.py:321-368)      */
#define SPH_COHESION 2u    /* cohesive pressure       (forces.py:371-405)      */
#define SPH_VISCOSITY 4u   /* viscous stress          (spam_complete_force.py:158-165) */
#define SPH_CONDUCTION 8u  /* heat conduction         (c_forces.pyx:185-239)   */
#define SPH_ALL_TERMS 15u

/* precision forms */
#define SPH_FP64 0         /* FP64 everywhere                                  */
#define SPH_FP32C 1        /* FP32 pair arithmetic, Kahan-compensated sums,    */
                           /* FP64 integration, energy equation and totals    */

typedef struct sph_ctx sph_ctx;
typedef struct sph_box sph_box;
typedef struct sph_arena sph_arena;

/* Physical and numerical parameters of one periodic 2-D box.  Names follow
 * the reference: h/hl = p.h/p.hlr (particles.py:331-341), rc/tol = VerletList
 * cutoff/tolerance (neighbour_list.py:153-157), a/b/kB = ADASH/BDASH/KBDASH
 * (properties.py:18-20), eta/zeta (spam_complete_force.py:43-44), kcond =
 * thermalk (f_properties.py:14). */
typedef struct {
    double Lx, Ly;
    double h, hl;
    double rc, tol;
    double a, b, kB;
    double eta, zeta, kcond;
    double dt;
    double skin;       /* neighbour-list skin: lists cover r_cut + skin and are
                          rebuilt before any particle moves skin/2 (0 = every
                          step); r_cut = min(sqrt(rc^2 + tol^2), max(h, hl))  */
    unsigned terms;    /* SPH_* rate terms used by sph_step                    */
    int precision;     /* SPH_FP64 or SPH_FP32C                                */
} sph_params;

/* ---- context ------------------------------------------------------------ */
/* One context owns one device and one HIP stream.  Not shared across threads
 * without external locking (the reference is single-threaded CPython). */
int sph_ctx_create(int device, sph_ctx **out);
int sph_ctx_destroy(sph_ctx *ctx);
/* Message of the last failure on this thread (never NULL). */
const char *sph_last_error(void);
/* Library version string. */
const char *sph_version(void);
/* Per-pass GPU timing with HIP events (off by default; on disables graph
 * replay so every pass can be bracketed). */
int sph_ctx_set_timing(sph_ctx *ctx, int on);

/* ---- one box ------------------------------------------------------------ */
/* Device-resident state of n particles.  Replaces the arrays of
 * SmoothParticleSystem (particles.py:325-344) for the hot path. */
int sph_box_create(sph_ctx *ctx, int64_t n, const sph_params *prm, sph_box **out);
int sph_box_destroy(sph_box *box);
int sph_box_set_params(sph_box *box, const sph_params *prm);
/* Upload state (any pointer may be NULL to keep the current value).
 * Replaces scatter_state (particles.py:512-526). */
int sph_box_upload(sph_box *box, const double *r, const double *v,
                   const double *m, const double *u);
/* Download state and derived fields in the caller's particle order (any
 * pointer may be NULL).  rho..P hold the values of the last density /
 * gradient / rate evaluation.  A download of r alone leaves an open run
 * (sph_box_advance) open; any other field closes it first.  Replaces the write-backs of
 * spam_complete_force.py:171-183 and f_properties.py:138-147. */
int sph_box_download(sph_box *box, double *r, double *v, double *u,
                     double *rho, double *rho_lr, double *T, double *p,
                     double *pco, double *acc, double *udot, double *gradv,
                     double *jq, double *P);

/* ---- (a) neighbour list ------------------------------------------------- */
/* Periodic cell-list build of every unordered pair with
 * rsq < rc^2 + tol^2, i<j, written lexicographically into iap[cap][2].
 * Replaces VerletList.build (neighbour_list.py:160-189).  *nip_out gets the
 * true count; returns SPH_ERR_CAP (after filling cap pairs) if it exceeds
 * cap. */
int sph_nlist_build(sph_box *box, int64_t cap, int64_t *iap, int64_t *nip_out);
/* Per-pair dr = r_j - r_i (minimum image), r, r^2, dv = v_j - v_i for the
 * given list.  Replaces NeighbourList.separations (neighbour_list.py:63-83)
 * and pairsep (pairsep.pyx:25-81).  Outputs may be NULL. */
int sph_separations(sph_box *box, int64_t nip, const int64_t *iap,
                    double *drij, double *rij, double *rsq, double *dv);
/* Minimum image of np displacements in place (dr[np][2]).  Replaces
 * NeighbourList.minimum_image (neighbour_list.py:105-123). */
int sph_minimum_image(sph_ctx *ctx, int64_t np, double *dr, double Lx, double Ly);
/* The FP32C passes' compensated accumulator on n terms x[n] (n <= 2^24):
 * the terms rounded to FP32, summed in FP32 rounds of `unroll` terms, each
 * round added to a Kahan total (*kahan); *plain gets the plain FP32 running
 * sum of the same terms.  No counterpart in the reference: it lets a test
 * compare the compensated scheme with plain FP32 accumulation. */
int sph_fp32_accumulate(sph_ctx *ctx, int64_t n, const double *x, int unroll, double *kahan, double *plain);

/* ---- (b) kernel, density, EOS ------------------------------------------- */
/* Lucy kernel W and dW/dx for np pairs in 1-3 dimensions (dx[np][dim]).
 * Replaces spkernel.lucy_kernel (spkernel.py:86-118). */
int sph_kernel_lucy(sph_ctx *ctx, int64_t np, const double *r, const double *dx,
                    double h, int dim, double *w, double *dwdx);
/* The reference's C-level 3-D Lucy kernel for one separation, with its
 * signature (c_properties.pxd:1, c_properties.pyx:50-80), so Cython code
 * that cimports it links against this library: w = W(|r|, h), dwdx[3] =
 * dW/dr dx / r (0 at r = 0 and for |r| > h), q = 105 / (16 pi h^3), in FP32
 * as the reference's inline.  A scalar host helper for callers that loop
 * over pairs themselves; batches of separations go to sph_kernel_lucy. */
void lucy_kernel_3d(float r, const float *dx, float h, float *w, float *dwdx);
/* Summation density (short and long h) and EOS (T, p, pco).  Replaces
 * properties.spam_properties (properties.py:63-120) and the density / EOS
 * part of f_properties.spam_properties (f_properties.py:100-120). */
int sph_density(sph_box *box);
/* Velocity gradient and heat flux (needs sph_density first).  Replaces
 * splib.calc_grad_v / calc_heat_flux_3d (f_properties.py:111,130). */
int sph_gradients(sph_box *box);
/* Rates of change for the given terms at the current state (density and,
 * if needed, gradients included); results in acc / udot.  Replaces
 * c_forces.SpamForce / CohesiveSpamForce / SpamConduction .apply
 * (c_forces.pyx:62-239) and sphforce3d.calc_sphforce3d
 * (spam_complete_force.py:158-165). */
int sph_rates(sph_box *box, unsigned terms);
/* u = kB*T - a*rho from a density pass (properties.py:44-46,119).  T may be
 * NULL to use T0 for every particle. */
int sph_init_thermal(sph_box *box, const double *T, double T0);

/* ---- time stepping ------------------------------------------------------ */
/* nsteps leapfrog steps (kick-drift-kick, consecutive half kicks fused) of
 * the full hot path: cell build, density+EOS, gradients, rates, kick/drift
 * with periodic wrap.  Replaces SmoothParticleSystem.update
 * (particles.py:459-494) for the periodic case. */
int sph_step(sph_box *box, int64_t nsteps);
/* Advance nsteps leapfrog steps and leave the run open: r at the new integer
 * time, v and u at the half step, one rate evaluation per step.  Successive
 * calls continue one run; k steps taken in any number of calls give the
 * bits of sph_step(box, k) once the run is closed.  A run is closed by
 * sph_box_sync and by any call that reads or replaces the state (downloads,
 * uploads, single passes, energies, pair export, snapshots); sph_step
 * continues an open run by its nsteps and closes it.  Without an open run
 * the first step starts one (k = 0, half kick).  Replaces the per-tick
 * SmoothParticleSystem.update (particles.py:459-494) as the reference's
 * driver calls it, once per step (toybox.py:101-103). */
int sph_box_advance(sph_box *box, int64_t nsteps);
/* Close an open run: the rate evaluation at the current step and the closing
 * half kick, so v and u are synchronised with r (no-op without an open run).
 * Reports a numerical failure of the run as SPH_ERR_NUMERIC. */
int sph_box_sync(sph_box *box);
/* Counters of the box's current or last run: steps taken, neighbour-list
 * builds, scheduled builds postponed by the list review, and open = 1 while
 * a run is open (any pointer may be NULL).  No counterpart in the
 * reference: the bench reports the builds its timed steps held. */
int sph_box_run_stats(sph_box *box, int64_t *steps, int64_t *builds, int64_t *extends, int *open);
/* Totals at the current (synchronised) state: sum m v^2/2, sum m u, sum. */
int sph_energy(sph_box *box, double *ke, double *ue, double *etot);
/* Accumulated GPU time per pass in ms: [0] cell build (bin, sort, lists),
 * [1] density, [2] gradients, [3] rates, [4] pair export, [5] number of
 * steps, [6] kick/drift.  n = number of entries to read (<= 7).  reset != 0
 * clears them.  Recorded only while sph_ctx_set_timing is on.
 * Mirrors the p.timing dictionary (particles.py:164-169,359). */
int sph_box_timing(sph_box *box, double *out, int n, int reset);
/* Average duration (microseconds) of one launch of a pair pass (0 density
 * + EOS, 1 gradients, 2 rates) at the box's current state: HIP events on
 * the context stream around reps back-to-back launches.  The passes write
 * only derived fields (densities, EOS, gradients, rates), never r, v, u or
 * m; the next sph_step re-derives them at its first evaluation, so timing
 * does not change later steps.  Pass 3 is the whole list build stage (cell
 * sort, permutation, lists): it leaves every particle's values but may
 * reorder the slots, and the next run builds at its first step.  The box
 * must have been stepped at least once (its lists and derived fields
 * exist).  No counterpart in the reference: the measurement behind the
 * dominant kernel's roofline figure. */
int sph_box_time_pass(sph_box *box, int pass, int reps, double *us_per_launch);

/* ---- rendering ------------------------------------------------------------ */
/* SPH interpolation of a particle field onto grid points (x[i], y[j]):
 * Z[i][j] = sum_p a_p m_p / rho_p W(|dr|, h_p) over the particles with
 * |dr| < cutoff, dr = (x_p - x[i], y[j] - y_p) under the minimum image with
 * periods (bounds[1], bounds[3]) (r[n][2]; bounds = xmin, xmax, ymin, ymax).
 * Replaces interpolate.splash_map_2d (interpolate.py:39-88) as a per-pixel
 * gather over a cell list (no atomics); equal to the reference's
 * per-particle scatter when cutoff >= max h + pixel width, since its
 * half-open pixel ranges then drop only pixels of zero weight. */
int sph_splash_map_2d(sph_ctx *ctx, int64_t n, const double *r, const double *m, const double *rho,
                      const double *a, const double *h, int nx, const double *x, int ny, const double *y,
                      const double *bounds, double cutoff, double *Z);
/* The same rendering without any allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  Every array is a
 * DEVICE pointer (r[n][2], m, rho, a, h, x[nx], y[ny] in; Z[nx][ny] out);
 * bounds[4] and cutoff stay on the host.  The particle records, sort keys,
 * cell starts and sort counters live in a fixed scratch area of the library
 * (5.6 MB of device memory per GPU, held until the process ends): at most
 * 65,536 particles and 65,536 grid cells (cells one cutoff wide); larger
 * calls return SPH_ERR_ARG before launching anything.  The work is enqueued
 * on the context's stream and the call returns without waiting.  All calls
 * on one GPU share its scratch: calls that overlap on one GPU (from several
 * contexts, streams or threads) race.  It calls no hipMalloc, hipFree or
 * allocating library routine. */
int sph_splash_map_2d_noalloc(sph_ctx *ctx, int64_t n, const double *r, const double *m, const double *rho,
                              const double *a, const double *h, int nx, const double *x, int ny, const double *y,
                              const double *bounds, double cutoff, double *Z);

/* ---- trajectory snapshots ----------------------------------------------- */
/* Pinned host memory (hipHostMalloc) for asynchronous snapshots. */
int sph_host_alloc(int64_t bytes, void **out);
int sph_host_free(void *p);
/* Start an asynchronous snapshot of r[n][2], v[n][2], u[n], m[n] in the
 * caller's particle order (any pointer may be NULL): the state is copied on
 * the device in stream order, then to the host buffers on a copy stream,
 * so steps enqueued afterwards overlap the transfer.  Buffers from
 * sph_host_alloc make the copy asynchronous.  A snapshot still in flight is
 * waited for first.  Replaces the state gather of spam_nc.write_step
 * (spam_nc.py:121-145). */
int sph_box_snapshot(sph_box *box, double *r, double *v, double *u, double *m);
/* Wait until the last snapshot has landed in its host buffers. */
int sph_box_snapshot_wait(sph_box *box);

/* ---- batching and serving ----------------------------------------------- */
/* An arena holds many independent boxes in one device-resident particle
 * pool; one step launch advances every running box by one step.  No
 * counterpart in the reference (it runs one system per process). */
int sph_arena_create(sph_ctx *ctx, int64_t particle_capacity, int precision,
                     sph_arena **out);
int sph_arena_destroy(sph_arena *ar);
/* Admit a box of n particles (r[n][2], v[n][2], m[n]) at temperature T0
 * (u = kB*T0 - a*rho is set on device) that will run nsteps steps.
 * Returns SPH_ERR_NOMEM if it does not fit; *box_id receives a handle. */
int sph_arena_add(sph_arena *ar, int64_t n, const sph_params *prm,
                  const double *r, const double *v, const double *m, double T0,
                  int64_t nsteps, int *box_id);
/* Advance every running box by nsteps arena steps (replays the captured
 * step graph).  Boxes that finish stay frozen until removed. */
int sph_arena_step(sph_arena *ar, int64_t nsteps);
/* Per-box progress: steps done (or -1 while initialising) and done flag. */
int sph_arena_poll(sph_arena *ar, int box_id, int64_t *steps_done, int *done);
/* States (0 free, 1 initialising, 2 running, 3 done), step index and health
 * bits of box ids 0..nmax-1 with one device read (ids never used or removed
 * read as free). */
int sph_arena_poll_all(sph_arena *ar, int nmax, int *state, int64_t *k, int *health);
/* Copy a box's state out (same layout as sph_box_download; any NULL). */
int sph_arena_result(sph_arena *ar, int box_id, double *r, double *v, double *u,
                     double *rho, double *e0, double *e1);
/* Release a box's slots. */
int sph_arena_remove(sph_arena *ar, int box_id);
/* Occupancy: particles in use, capacity, running boxes, total boxes. */
int sph_arena_stats(sph_arena *ar, int64_t *used, int64_t *capacity,
                    int *running, int *boxes);
/* Per-pass GPU time of the arena's steps, as sph_box_timing. */
int sph_arena_timing(sph_arena *ar, double *out, int n, int reset);
/* Replay the step as a captured HIP graph (default on) or launch directly. */
int sph_arena_set_graph(sph_arena *ar, int on);
int sph_box_set_graph(sph_box *box, int on);
/* Launch schedule of the pair passes: list entries whose loads are in
 * flight per lane (1, 2 or 4; 0 = the precision's default).  Chosen by the
 * autotuner (pyticles_amd/autotune.py); any value gives the same results up
 * to summation order.  No counterpart in the reference. */
int sph_arena_set_schedule(sph_arena *ar, int unroll);
int sph_box_set_schedule(sph_box *box, int unroll);
/* The same for the rate pass alone (0 = as the other passes): its walk keeps
 * five records per entry in flight, so its best depth differs (two at one
 * large FP64 box, where four cost occupancy; four for batched boxes).
 * Tuned per workload class.  No counterpart in the reference. */
int sph_arena_set_rates_unroll(sph_arena *ar, int unroll);
int sph_box_set_rates_unroll(sph_box *box, int unroll);
/* Work-items per block of the three pair passes: 128, 256 or 512 (0 = the
 * default 256).  A launch schedule the autotuner picks per workload class;
 * results do not depend on it.  No counterpart in the reference. */
int sph_arena_set_pass_block(sph_arena *ar, int block);
int sph_box_set_pass_block(sph_box *box, int block);
/* Entry width of the neighbour lists the boxes' next builds write: 16
 * (default: slot offsets, half the list bytes of 32; a box whose offsets do
 * not fit 16 bits holds the step of that build and continues with 32-bit
 * entries by itself) or 32 (slots).  Results are the same bits either way.
 * sph_*_list_bits reports the width of a box's lists in memory (after its
 * first build).  No counterpart in the reference (its lists are Python
 * arrays, neighbour_list.py:32-35). */
/* Where the leapfrog step runs: inside the rate pass's launch (1, the
 * default: no re-read of the rates, one launch fewer per step) or as a
 * launch of its own after it (0); the same results.  A launch schedule the
 * autotuner picks per workload class.  No counterpart in the reference. */
int sph_arena_set_fused_step(sph_arena *ar, int on);
int sph_box_set_fused_step(sph_box *box, int on);
int sph_arena_set_list_bits(sph_arena *ar, int bits);
int sph_box_set_list_bits(sph_box *box, int bits);
int sph_arena_list_bits(sph_arena *ar, int box_id, int *bits);
int sph_box_list_bits(sph_box *box, int *bits);
/* Device bytes an arena of particle_capacity slots allocates (the serving
 * loop's HBM-budget admission).  No counterpart in the reference. */
int sph_arena_footprint(int64_t particle_capacity, int precision, int64_t *bytes);

#ifdef __cplusplus
}
#endif

#endif /* PYTICLES_HIP_H */
