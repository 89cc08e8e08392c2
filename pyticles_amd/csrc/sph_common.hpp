// sph_common.hpp -- types shared by the host runtime and the gfx950 kernels.
//
// Device layout (structure of arrays, one slot per particle; slots are
// sorted by (box, cell) at every neighbour-list build and keep that order
// until the next build; all arrays live in one arena shared by every box):
//   state            r, v (double2), u, m (double), id (int); v and u are
//                    half-step values between steps (leapfrog)
//   predicted        vp, up: velocity / energy predicted to the current
//                    integer time from the last rates; every pass reads these
//   second buffers   r_nx, vp_nx (and FP32C rf_nx, rlo_nx): the step pass
//                    writes the stepped positions, predicted velocities and
//                    records there; the two sets swap roles every step
//   last rates       acc (double2), udot
//   sorted scratch   *_s copies written by the permutation of a build
//   neighbour list   tiles of 64 slots in groups of 4 entries, short entries
//                    (r < h + skin at build) at P = 0.., long entries
//                    (h + skin <= r < r_cut + skin) at P = lcap-1 down;
//                    16-bit slot offsets (halfword ((s/64)*(lcap/4)*2 +
//                    P/4)*256 + (s%64)*4 + P%4) or, for a box whose offsets
//                    do not fit, 32-bit slots (word ((s/64)*(lcap/4) +
//                    P/4)*256 + (s%64)*4 + P%4); cnt[s] = (short, long)
//                    counts, bit 31 of the short count set when an offset is
//                    taken modulo the box's slot count
//   derived          rho (double2: rho, rho_lr), T, S (double4: Sxx, Sxy,
//                    Syy = (p I + Pi)/rho^2, and m), Q (double2: q/rho^2),
//                    vt (double2: m/rho, T), aco / uco (cohesive rates)
//   FP32 records     rf: the high 32 bits of x, y as 48-bit box-relative
//                    fixed point (2^48 units per side) and float v; rlo:
//                    their low 16 bits; S32 (float4: S, m), g32 (rlo,
//                    m/rho, T), q32 (rlo, q/rho^2)
// A box is a contiguous slot range [part_off, part_off + n) plus a key range
// [key_base, key_base + ncx*ncy) of cells (key_base + ncx*ncy, the next
// segment's first key, ends its cell_start range); free slot ranges own one
// key each.  Key ranges increase with the slot offset, so a stable sort of
// the building slots' keys never moves a particle across boxes.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace sph {

enum BoxState : int { BOX_FREE = 0, BOX_INIT = 1, BOX_RUN = 2, BOX_DONE = 3 };

enum Health : int {
    HEALTH_NONFINITE = 1,   // NaN/Inf in density, rates or state
    HEALTH_EOS_POLE = 2,    // rho * b >= 1 (vdW pole, properties.py:41)
    HEALTH_SKIN = 4,        // displacement since the last list build exceeded skin/2
    HEALTH_LISTCAP = 8,     // a particle had more neighbours than the list holds
};

// Largest box side whose FP32C long-range walk takes separations from the
// high 32 bits of the positions alone (resolution side / 2^32, 2.4e-7 at
// 1024, about one FP32 rounding of a separation); wider boxes read the low
// bits too, as the short-range terms always do.
constexpr double FP32C_HI_ONLY_MAX_L = 1024.0;

// Static description of one box (host-owned, uploaded when it changes).
struct BoxDesc {
    double Lx, Ly, hLx, hLy;       // sides and half sides (minimum image)
    double cwx, cwy, icwx, icwy;   // cell widths and their inverses
    double h, hl, ih, ihl;         // smoothing lengths and inverses
    double qh, qhl, fh, fhl;       // Lucy norms 5/(pi h^2) and -12 q / h^2
    double cut2, h2, hl2, R2list;  // candidate cut, short, long, list radius^2
    double a, b, kB;               // vdW constants
    double eta, zeta, kcond;       // transport coefficients
    double dt, T0;                 // time step, initial temperature
    double skin, rnl2, hsk2, hskin2;  // list skin, (r_cut+skin)^2, (h+skin)^2, (skin/2)^2
    double munif;                  // the common mass when all masses are equal, else 0
    double rin;                    // interior margin 2 (r_cut + skin): a particle at least this far
                                   // from every side has no listed neighbour across a side
    int ncx, ncy, sx, sy;          // cell grid and stencil half widths (0/1)
    int key_base, part_off, n, terms;
    int precision;
    int lolong;                    // FP32C: the long-range walk reads the low position bits too
                                   // (sides over FP32C_HI_ONLY_MAX_L)
    int build16;                   // builds write 16-bit list entries (0: 32-bit slots)
    int pad2;
};

// Dynamic per-box state (device-owned; the host writes single entries
// between launches and reads them back for polling).
struct BoxDyn {
    int state;      // BoxState
    int health;     // Health bits
    int64_t k;      // current step index in [0, nsteps]
    int64_t nsteps; // steps requested for this run
    double E0, E1;  // total energy at the first and last synchronised state
    int64_t next_rebuild;   // step index of the next neighbour-list build
    unsigned long long vmax2, amax2;  // max |v|^2, |a|^2 at the last build (double bits)
    int64_t builds;         // list builds so far
    double vlast, alast;    // max |v|, |a| measured at the last build
    int64_t built_at;       // step index of the last build
    int64_t extends;        // scheduled builds postponed on the measured displacement
    unsigned long long dmax2;  // largest squared displacement of the box's slots in
                               // waves shared with other boxes, gathered on its review
                               // step (double bits; cleared when read)
    int list16;     // the lists in memory hold 16-bit entries (set by each build)
    int relist;     // a build met a neighbour offset beyond 16 bits: the box holds
                    // (no pass, integration or bookkeeping) until the runtime has
                    // switched it to 32-bit entries and cleared this
};

// Segment of the slot space: a box or a free range, sorted by offset.
struct SegDesc {
    int off, len;   // slot range
    int box;        // box index or -1 when free
    int key;        // key base (box) or the segment's single key (free)
};

// FP32C records that a pass gathers per neighbour in one 16-byte load: the
// gradient pass's (the neighbour's low position bits, m / rho and T) and the
// rate pass's conduction record (low position bits and q / rho^2).  The
// position bits are those of rlo at the time the record is written, which
// no launch between the writer and the reader changes.
struct alignas(16) GRec32 {
    unsigned lo;
    float v;   // m / rho
    double T;
};
struct alignas(16) QRec32 {
    unsigned lo;
    float qx, qy;  // q / rho^2
    float pad;
};

struct ArenaView {
    // state (slot order of the last build)
    double2 *r, *v, *vp, *acc;
    // the second buffers of what other slots read during the step pass:
    // the leapfrog step writes the new position, predicted velocity and
    // FP32C record there (the runtime swaps them with r, vp, rf, rlo after
    // every step)
    double2 *r_nx, *vp_nx;
    float4 *rf_nx;
    unsigned *rlo_nx;
    double *u, *m, *up, *udot;
    int *id;
    // sorted scratch written by the permutation of a build
    double2 *r_s, *v_s, *vp_s, *acc_s;
    double *u_s, *m_s, *up_s, *udot_s;
    int *id_s;
    // build bookkeeping
    int *pcell;          // cell of the slot at the last build
    double2 *r_ref;      // position at the last build (skin check)
    int *slot_box;       // box of the slot (-1 = free)
    unsigned *nbr;       // neighbour lists: 64-slot tiles, 4-entry groups (list_base)
    int2 *cnt;           // (short, long) entries per slot
    // derived fields
    double2 *rho;
    double *T;
    double4 *S;
    double2 *Q;
    double2 *vt;         // (m / rho, T): neighbour record of the gradient pass
    double2 *aco;        // cohesive acceleration from the density pass
    double *uco;         // cohesive du/dt from the density pass
    double4 *gradv;
    double2 *jq;
    double *e;
    // FP32C neighbour records (see fix_rec)
    float4 *rf;          // high 32 bits of the 48-bit box-relative x, y; predicted v
    unsigned *rlo;       // low 16 bits of x (bits 0-15) and y (bits 16-31)
    float4 *S32;         // (Sxx, Sxy, Syy, m): the rate pass's stress record
    GRec32 *g32;         // the gradient pass's record (rlo, m / rho, T), written by the density pass
    QRec32 *q32;         // the rate pass's conduction record (rlo, q / rho^2), written by the gradient pass
    // cells and sort
    int *cell_start;
    unsigned *keys;      // sorted keys (final buffer of the radix sort)
    unsigned *vals;      // source slot of each sorted slot
    unsigned *keys_in;   // unsorted keys written by the key pass
    // build plan: plan_n[1] = building slots M; box_coff[b] = compact
    // position of building box b's first slot (its slots numbered on);
    // tile_list[0 .. plan_n[2]) = the 64-slot tiles holding building slots,
    // in slot order, each once
    int *plan_n;
    int *box_coff;
    int *tile_list;
    double *wdisp;       // per 64-slot wave wholly inside one box: its largest squared
                         // displacement since the build, stored on the box's review step
    double2 *wbuild;     // per 64-slot wave wholly inside one building box: its largest
                         // squared predicted speed and acceleration, stored at the build
    double *wenergy;     // per 64-slot wave wholly inside one box: the sum of its slots'
                         // energies, stored on the box's first and last step
    // boxes
    const BoxDesc *boxes;
    BoxDyn *dyn;
    BoxDyn *dyn_host;    // single boxes: host-coherent pinned memory where k_box_step leaves
                         // every box's state at the end of each step (null: not kept)
    const SegDesc *segs;
    int nseg;
    int n;               // slots in use (high-water mark)
    int nkeys;           // total keys
    int cap;             // slot capacity
    int lcap;            // list entries per slot
    int pass_block;      // work-items per block of the pair passes (128, 256 or 512)
};

// Grid of the splash rendering: particle cells at least one cutoff wide.
struct SplashGrid {
    double Lx, Ly, cwx, cwy, cutoff;
    int ncx, ncy;
};

}  // namespace sph
