// sph_common.hpp -- types shared by the host runtime and the gfx950 kernels.
//
// Device layout (structure of arrays, one slot per particle, sorted by cell
// after every build; all arrays live in one arena shared by every box):
//   canonical state  r, v (double2), u, m (double), id (int)  -- end of step
//   sorted copies    r_s, v_s, u_s, m_s, id_s, pcell          -- read by passes
//                    vp_s, up_s: velocity / energy predicted to the current
//                    integer time from the last rates (leapfrog predictor)
//   derived          rho (double2: rho, rho_lr), T, S (double4: Sxx, Sxy,
//                    Syy, Sco = P/rho^2 and pco/rho_lr^2), Q (double2: q/rho^2)
//   FP32 mirrors     rf (float4: x, y relative to the cell corner, m, cell
//                    column|row), vf (float2), rho32 (float2), T32, S32, Q32
// A box is a contiguous slot range [part_off, part_off + n) plus a key range
// [key_base, key_base + ncx*ncy] of cells; free slot ranges own one key each
// so a stable sort over the whole arena never moves a particle across boxes.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace sph {

enum BoxState : int { BOX_FREE = 0, BOX_INIT = 1, BOX_RUN = 2, BOX_DONE = 3 };

enum Health : int {
    HEALTH_NONFINITE = 1,   // NaN/Inf in density, rates or state
    HEALTH_EOS_POLE = 2,    // rho * b >= 1 (vdW pole, properties.py:41)
    HEALTH_SKIN = 4,        // displacement since the last cell build exceeded the skin
};

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
    int ncx, ncy, sx, sy;          // cell grid and stencil half widths (0/1)
    int key_base, part_off, n, terms;
    int precision, pad0, pad1, pad2;
};

// Dynamic per-box state (device-owned; the host writes single entries
// between launches and reads them back for polling).
struct BoxDyn {
    int state;      // BoxState
    int health;     // Health bits
    int64_t k;      // current step index in [0, nsteps]
    int64_t nsteps; // steps requested for this run
    double E0, E1;  // total energy at the first and last synchronised state
};

// Segment of the slot space: a box or a free range, sorted by offset.
struct SegDesc {
    int off, len;   // slot range
    int box;        // box index or -1 when free
    int key;        // key base (box) or the segment's single key (free)
};

struct ArenaView {
    // canonical state
    double2 *r, *v;
    double *u, *m;
    int *id;
    // sorted copies
    double2 *r_s, *v_s;
    double *u_s, *m_s;
    int *id_s, *pcell;
    double2 *vp_s;
    double *up_s;
    double2 *acc_s;
    double *udot_s;
    // per-slot box index (-1 = free)
    int *slot_box;
    // derived fields
    double2 *rho;
    double *T;
    double4 *S;
    double2 *Q;
    double2 *acc;
    double *udot;
    double4 *gradv;
    double2 *jq;
    double *e;
    // FP32 mirrors
    float4 *rf;
    float2 *vf;
    float2 *rho32;
    float *T32;
    float4 *S32;
    float2 *Q32;
    // cells and sort
    int *cell_start;
    unsigned *keys;      // sorted keys (final buffer of the radix sort)
    unsigned *vals;      // source slot of each sorted slot
    unsigned *keys_in;   // unsorted keys written by the key pass
    unsigned *vals_in;
    // boxes
    const BoxDesc *boxes;
    BoxDyn *dyn;
    const SegDesc *segs;
    int nseg;
    int n;               // slots in use (high-water mark)
    int nkeys;           // total keys
    int pad;
};

}  // namespace sph
