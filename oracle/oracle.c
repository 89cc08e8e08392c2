/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the pyticles smooth-particle path (neighbour list,
 * Lucy kernel, summation density, van der Waals EOS, pair forces and rates,
 * leapfrog).  It is the checker for the HIP library: only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it.  The product
 * path never calls it.
 *
 * Loop structure follows the reference's pair-list code: an (i<j) pair list in
 * lexicographic order, every pair scattered to both ends, exactly as the Cython
 * and Python loops do (neighbour_list.py:160-189, properties.py:84-98,
 * c_forces.pyx:96-111, c_forces.pyx:224-237).  The float32 arithmetic and the
 * known bugs of the Cython path are not reproduced (pairsep.pyx:41-70 works in
 * float; c_properties.pyx:150 never resets drk; c_properties.pyx:166-188 uses
 * rho_j for particle i).
 *
 * Everything is 2-D (the configurations this path serves are 2-D; the
 * reference's 2-D branch of the Lucy kernel is spkernel.py:97-98).
 *
 * Parity status per function (what pins it; the tests hold the checks):
 *   orc_min_image, orc_pairs_*   pinned: brute force, lattice tie count 24*N,
 *                                reference known answers (rij == 1.0)
 *   orc_lucy                     pinned: closed forms W(0,h)=5/(pi h^2),
 *                                unit integral, W(h)=W'(h)=0, finite differences
 *   orc_density                  pinned: lattice closed sums, brute long-double
 *   orc_eos                      pinned: vdw_temp(vdw_energy(T)) == T, p=2,pco=-2
 *   orc_gradients                pinned by lattice closed forms (the reference's
 *                                Fortran splib.calc_grad_v / calc_heat_flux_3d is
 *                                absent): linear and quadratic T and v profiles
 *                                give grad T, grad v and q = -k grad T exactly up
 *                                to one lattice factor c computed independently
 *   orc_rates                    pressure/cohesion pinned by two-particle
 *                                analytic case, momentum and energy balance;
 *                                viscosity (eta and zeta), conduction magnitude
 *                                and sign, viscous force assembly and heating
 *                                pinned by lattice closed forms: quadratic T
 *                                (udot = 2 k beta c^2/rho), uniform dilation
 *                                (Pi = -2 zeta eps c I), quadratic shear
 *                                (a_x = 2 eta gamma c^2/rho and its heating);
 *                                planted mistakes (factor 2, sign, transposed
 *                                operand) fail them
 *   orc_run                      leapfrog defined here (the reference has none);
 *                                checked by energy drift and step-size order
 *   orc_splash_map_2d            pinned: lattice rendering at particle sites is
 *                                exactly 1 (sum_j W / rho = rho / rho), linear
 *                                in a, brute-force pixel sums on tiny cases
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#if defined(_OPENMP)
#include <omp.h>
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

/* Two builds of this file: liboracle.so (no OpenMP) is the checker and sums */
/* every pair loop serially in pair order; liboracle_omp.so (-fopenmp) is    */
/* only the all-cores CPU baseline: pair loops split over threads into      */
/* private partial sums added in thread order, so its results differ from   */
/* the checker's by rounding only.                                          */

typedef struct {
    double Lx, Ly;      /* periodic box sides (box.xmax, box.ymax)                 */
    double h, hl;       /* short / long smoothing lengths (p.h, p.hlr)             */
    double rc, tol;     /* VerletList cutoff and tolerance, neighbour_list.py:153  */
    double a, b, kB;    /* vdW constants ADASH, BDASH, KBDASH, properties.py:18-20 */
    double eta, zeta;   /* shear / bulk viscosity, spam_complete_force.py:43-44    */
    double kcond;       /* thermal conductivity (thermalk, f_properties.py:14)     */
    double dt;          /* time step                                               */
} orc_params;

/* math.pi as used by spkernel.py:15,98 */
#define ORC_PI 3.14159265358979323846

enum { ORC_PRESSURE = 1, ORC_COHESION = 2, ORC_VISCOSITY = 4, ORC_CONDUCTION = 8 };

/* ------------------------------------------------------------------------ */
/* a1  minimum image: neighbour_list.py:105-123.  All '>' tests first, then  */
/* all '<' tests, both strict, per component.                                */
void orc_min_image(double *dx, double *dy, double Lx, double Ly)
{
    double x = *dx, y = *dy;
    if (x > Lx / 2.) x = x - Lx;
    if (y > Ly / 2.) y = y - Ly;
    if (x < -Lx / 2.) x = x + Lx;
    if (y < -Ly / 2.) y = y + Ly;
    *dx = x;
    *dy = y;
}

/* squared separation with the reference's expression order (neighbour_list.py:177).
 * The build recipe compiles with -ffp-contract=off so this is never an fma. */
static inline double pair_rsq(double dx, double dy)
{
    return dx * dx + dy * dy;
}

/* ------------------------------------------------------------------------ */
/* a2  VerletList.build: neighbour_list.py:160-189.  All i<j, keep iff       */
/* rsq < rc^2 + tol^2 (strict).  Pairs are written lexicographically.        */
/* Returns the true pair count; writes at most cap pairs.                    */
int64_t orc_pairs_brute(int64_t n, const double *r, double Lx, double Ly,
                        double R2, int64_t cap, int64_t *iap)
{
    int64_t nip = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = i + 1; j < n; ++j) {
            double dx = r[2 * j] - r[2 * i];
            double dy = r[2 * j + 1] - r[2 * i + 1];
            orc_min_image(&dx, &dy, Lx, Ly);
            double rsq = pair_rsq(dx, dy);
            if (rsq < R2) {
                if (nip < cap) {
                    iap[2 * nip] = i;
                    iap[2 * nip + 1] = j;
                }
                ++nip;
            }
        }
    }
    return nip;
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

typedef struct {
    const double *r;
    double Lx, Ly, R2;
    int ncx, ncy;
    const int64_t *head, *order, *cellof;
} cell_grid;

/* Partners j > i of particle i with r^2 < R2 from the 3x3 cell stencil,    */
/* ascending in j; returns their count.                                      */
static int64_t cell_partners(const cell_grid *g, int64_t i, int64_t *js)
{
    const double *r = g->r;
    int ncx = g->ncx, ncy = g->ncy;
    int cx = (int)(g->cellof[i] % ncx), cy = (int)(g->cellof[i] / ncx);
    int64_t nj = 0;
    int dxs = ncx >= 3 ? 1 : 0, dys = ncy >= 3 ? 1 : 0;
    for (int oy = -dys; oy <= dys; ++oy) {
        int yy = (cy + oy + ncy) % ncy;
        for (int ox = -dxs; ox <= dxs; ++ox) {
            int xx = (cx + ox + ncx) % ncx;
            int64_t c = (int64_t)yy * ncx + xx;
            for (int64_t t = g->head[c]; t < g->head[c + 1]; ++t) {
                int64_t j = g->order[t];
                if (j <= i) continue;
                double dx = r[2 * j] - r[2 * i];
                double dy = r[2 * j + 1] - r[2 * i + 1];
                orc_min_image(&dx, &dy, g->Lx, g->Ly);
                if (pair_rsq(dx, dy) < g->R2) js[nj++] = j;
            }
        }
    }
    qsort(js, (size_t)nj, sizeof(int64_t), cmp_i64);
    return nj;
}

/* Same pair set via an O(N) periodic cell list; identical output order.     */
/* The pair test is the same expression as orc_pairs_brute; only the         */
/* candidate enumeration differs.                                            */
int64_t orc_pairs_cell(int64_t n, const double *r, double Lx, double Ly,
                       double R2, int64_t cap, int64_t *iap)
{
    double R = sqrt(R2) * (1.0 + 1e-9);
    int ncx = (int)floor(Lx / R), ncy = (int)floor(Ly / R);
    if (ncx < 3) ncx = 1;
    if (ncy < 3) ncy = 1;
    int64_t nc = (int64_t)ncx * ncy;
    int64_t *head = (int64_t *)malloc(sizeof(int64_t) * (nc + 1));
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t *cellof = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    memset(head, 0, sizeof(int64_t) * (nc + 1));
    for (int64_t i = 0; i < n; ++i) {
        int cx = (int)(r[2 * i] / Lx * ncx), cy = (int)(r[2 * i + 1] / Ly * ncy);
        if (cx < 0) cx = 0;
        if (cx >= ncx) cx = ncx - 1;
        if (cy < 0) cy = 0;
        if (cy >= ncy) cy = ncy - 1;
        cellof[i] = (int64_t)cy * ncx + cx;
        head[cellof[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c) head[c + 1] += head[c];
    {
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (nc > 0 ? nc : 1));
        memcpy(fill, head, sizeof(int64_t) * nc);
        for (int64_t i = 0; i < n; ++i) order[fill[cellof[i]]++] = i;
        free(fill);
    }
    cell_grid g = {r, Lx, Ly, R2, ncx, ncy, head, order, cellof};
    int64_t nip = 0;
#if defined(_OPENMP)
    /* count pass, exclusive scan, then each particle writes its own rows    */
    int64_t *at = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
#pragma omp parallel
    {
        int64_t *js = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) at[i + 1] = cell_partners(&g, i, js);
#pragma omp single
        {
            at[0] = 0;
            for (int64_t i = 0; i < n; ++i) at[i + 1] += at[i];
        }
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            if (at[i] >= cap) continue;
            int64_t nj = cell_partners(&g, i, js);
            for (int64_t k = 0; k < nj && at[i] + k < cap; ++k) {
                iap[2 * (at[i] + k)] = i;
                iap[2 * (at[i] + k) + 1] = js[k];
            }
        }
        free(js);
    }
    nip = at[n];
    free(at);
#else
    int64_t *js = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        int64_t nj = cell_partners(&g, i, js);
        for (int64_t k = 0; k < nj; ++k) {
            if (nip < cap) {
                iap[2 * nip] = i;
                iap[2 * nip + 1] = js[k];
            }
            ++nip;
        }
    }
    free(js);
#endif
    free(head);
    free(order);
    free(cellof);
    return nip;
}

/* a4  NeighbourList.separations: neighbour_list.py:63-83 (FP64 semantics:   */
/* min-image applied before rij).                                            */
void orc_separations(const double *r, const double *v, double Lx, double Ly,
                     int64_t nip, const int64_t *iap,
                     double *drij, double *rij, double *rsq, double *dv)
{
    OMP_FOR
    for (int64_t k = 0; k < nip; ++k) {
        int64_t i = iap[2 * k], j = iap[2 * k + 1];
        double dx = r[2 * j] - r[2 * i];
        double dy = r[2 * j + 1] - r[2 * i + 1];
        orc_min_image(&dx, &dy, Lx, Ly);
        double q = pair_rsq(dx, dy);
        if (drij) { drij[2 * k] = dx; drij[2 * k + 1] = dy; }
        if (rsq) rsq[k] = q;
        if (rij) rij[k] = sqrt(q);
        if (dv) {
            dv[2 * k] = v[2 * j] - v[2 * i];
            dv[2 * k + 1] = v[2 * j + 1] - v[2 * i + 1];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* b1  spkernel.lucy_kernel: spkernel.py:86-118, literal form.  dim in 1..3. */
static double lucy_norm(double h, int dim)
{
    if (dim == 1) return 5. / (4. * h);                 /* spkernel.py:96  */
    if (dim == 2) return 5. / (ORC_PI * h * h);           /* spkernel.py:98  */
    return 105. / (ORC_PI * 16. * (h * h * h));           /* spkernel.py:100 */
}

void orc_lucy(int64_t np, const double *r, const double *dx, double h, int dim,
              double *w, double *dwdx)
{
    double q = lucy_norm(h, dim);
    for (int64_t k = 0; k < np; ++k) {
        double rr = fabs(r[k]);
        double t = 1. - rr / h;
        if (rr < h) {
            w[k] = q * (1. + 3. * rr / h) * (t * t * t);
            for (int c = 0; c < dim; ++c) {
                if (rr == 0.)
                    dwdx[dim * k + c] = 0.;
                else
                    dwdx[dim * k + c] = q * ((-12. / (h * h * h * h)) * (rr * rr * rr)
                                             + (24. / (h * h * h)) * (rr * rr)
                                             - (12. * rr / (h * h))) * dx[dim * k + c] / rr;
            }
        } else {
            w[k] = 0.;
            for (int c = 0; c < dim; ++c) dwdx[dim * k + c] = 0.;
        }
    }
}

/* 2-D Lucy W and gradient factor for the pair loops.  dwdx = F * (r_j - r_i)  */
/* with F = (dW/dr)/r = -12 q (1-r/h)^2 / h^2, the reference's polynomial     */
/* (spkernel.py:112-113) factored; F*dx is 0 at r=0 like spkernel.py:108-109. */
static inline void lucy2(double r, double h, double *W, double *F)
{
    if (r < h) {
        double q = 5. / (ORC_PI * h * h);
        double s = r / h, t = 1. - s;
        *W = q * (1. + 3. * s) * (t * t * t);
        *F = -12. * q * (t * t) / (h * h);
    } else {
        *W = 0.;
        *F = 0.;
    }
}

/* ------------------------------------------------------------------------ */
/* Pair scatter loops.  body(ctx, k0, k1, out) adds pairs [k0, k1) into the  */
/* nf arrays out[f] (len[f] doubles each).  Serially that is one call over  */
/* all pairs on the caller's arrays; with OpenMP each thread sums its block */
/* of pairs into zeroed private arrays, which are then added to out[] in    */
/* thread order.                                                             */
typedef void (*pair_body)(const void *ctx, int64_t k0, int64_t k1, double *const *out);

static void scatter_pairs(pair_body body, const void *ctx, int64_t nip, int nf,
                          const int64_t *len, double *const *out)
{
#if defined(_OPENMP)
    int nt = omp_get_max_threads();
    int64_t width = 0;
    for (int f = 0; f < nf; ++f) width += len[f];
    double *priv = nt > 1 ? (double *)calloc((size_t)nt * (size_t)width, sizeof(double)) : NULL;
    if (priv) {
#pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num(), nth = omp_get_num_threads();
            double *mine[4] = {NULL, NULL, NULL, NULL};
            int64_t off = (int64_t)t * width;
            for (int f = 0; f < nf; ++f) {
                mine[f] = priv + off;
                off += len[f];
            }
            body(ctx, nip * t / nth, nip * (t + 1) / nth, mine);
#pragma omp barrier
            int64_t base = 0;
            for (int f = 0; f < nf; ++f) {
#pragma omp for schedule(static)
                for (int64_t x = 0; x < len[f]; ++x) {
                    double s = 0.;
                    for (int q = 0; q < nth; ++q) s += priv[(int64_t)q * width + base + x];
                    out[f][x] += s;
                }
                base += len[f];
            }
        }
        free(priv);
        return;
    }
#endif
    body(ctx, 0, nip, out);
}

/* ------------------------------------------------------------------------ */
/* b3  density: properties.py:77-91 / spdensity.py:40-56, long range as      */
/* spam_properties_ls properties.py:138-161.  Self term W(0,h) is not        */
/* multiplied by m_i (properties.py:77-78).                                  */
typedef struct {
    const double *m, *rij;
    const int64_t *iap;
    double h, hl;
} density_ctx;

static void density_body(const void *cv, int64_t k0, int64_t k1, double *const *out)
{
    const density_ctx *c = (const density_ctx *)cv;
    const double *m = c->m;
    double *rho = out[0], *rho_lr = out[1];
    for (int64_t k = k0; k < k1; ++k) {
        int64_t i = c->iap[2 * k], j = c->iap[2 * k + 1];
        double W, Wl, f;
        lucy2(c->rij[k], c->h, &W, &f);
        lucy2(c->rij[k], c->hl, &Wl, &f);
        rho[i] += W * m[j];
        rho[j] += W * m[i];
        rho_lr[i] += Wl * m[j];
        rho_lr[j] += Wl * m[i];
    }
}

void orc_density(int64_t n, const double *m, int64_t nip, const int64_t *iap,
                 const double *rij, double h, double hl,
                 double *rho, double *rho_lr)
{
    double w0, w0l, f;
    lucy2(0., h, &w0, &f);
    lucy2(0., hl, &w0l, &f);
    for (int64_t i = 0; i < n; ++i) { rho[i] = w0; rho_lr[i] = w0l; }
    density_ctx c = {m, rij, iap, h, hl};
    int64_t len[2] = {n, n};
    double *out[2] = {rho, rho_lr};
    scatter_pairs(density_body, &c, nip, 2, len, out);
}

/* b4  EOS: vdw properties.py:38-41 (p from short rho, pco = -a rho_lr^2 as */
/* f_properties.py:118-119), vdw_temp properties.py:48-49 with the T >= 0    */
/* clamp of spam_complete_force.py:152.                                      */
void orc_eos(int64_t n, const double *rho, const double *rho_lr, const double *u,
             double a, double b, double kB, double *T, double *p, double *pco)
{
    OMP_FOR
    for (int64_t i = 0; i < n; ++i) {
        double t = (u[i] + a * rho[i]) / kB;
        if (t < 0.) t = 0.;
        T[i] = t;
        p[i] = (rho[i] * kB * t) / (1. - rho[i] * b);
        pco[i] = -a * rho_lr[i] * rho_lr[i];
    }
}

/* b5 / c3 gradients.  grad_i W_ij = -dwdx_ij with dwdx built from r_j - r_i */
/*   gradv_i[a][b] = sum_j (m_j/rho_j) (v_j - v_i)_a (grad_i W_ij)_b          */
/*   gradT_i[b]    = sum_j (m_j/rho_j) (T_j - T_i)   (grad_i W_ij)_b          */
/* The reference's splib.calc_grad_v / calc_heat_flux_3d (f_properties.py: */
/* 111,130) are absent; lattice closed forms pin these sums (linear and     */
/* quadratic profiles).  Layout gradv[n][2][2], gradT[n][2].                */
typedef struct {
    const double *m, *rho, *v, *T, *rij, *drij;
    const int64_t *iap;
    double h;
} gradients_ctx;

static void gradients_body(const void *cv, int64_t k0, int64_t k1, double *const *out)
{
    const gradients_ctx *c = (const gradients_ctx *)cv;
    const double *m = c->m, *rho = c->rho, *v = c->v, *T = c->T, *drij = c->drij;
    double *gradv = out[0], *gradT = out[1];
    for (int64_t k = k0; k < k1; ++k) {
        int64_t i = c->iap[2 * k], j = c->iap[2 * k + 1];
        double W, F;
        lucy2(c->rij[k], c->h, &W, &F);
        if (F == 0.) continue;
        double gx = -F * drij[2 * k], gy = -F * drij[2 * k + 1];  /* grad_i W */
        double dvx = v[2 * j] - v[2 * i], dvy = v[2 * j + 1] - v[2 * i + 1];
        double dT = T[j] - T[i];
        double vj = m[j] / rho[j], vi = m[i] / rho[i];
        gradv[4 * i + 0] += vj * dvx * gx;
        gradv[4 * i + 1] += vj * dvx * gy;
        gradv[4 * i + 2] += vj * dvy * gx;
        gradv[4 * i + 3] += vj * dvy * gy;
        gradv[4 * j + 0] += vi * dvx * gx;
        gradv[4 * j + 1] += vi * dvx * gy;
        gradv[4 * j + 2] += vi * dvy * gx;
        gradv[4 * j + 3] += vi * dvy * gy;
        gradT[2 * i + 0] += vj * dT * gx;
        gradT[2 * i + 1] += vj * dT * gy;
        gradT[2 * j + 0] += vi * dT * gx;
        gradT[2 * j + 1] += vi * dT * gy;
    }
}

void orc_gradients(int64_t n, const double *m, const double *rho, const double *v,
                   const double *T, int64_t nip, const int64_t *iap,
                   const double *rij, const double *drij, double h,
                   double *gradv, double *gradT)
{
    memset(gradv, 0, sizeof(double) * 4 * n);
    memset(gradT, 0, sizeof(double) * 2 * n);
    gradients_ctx c = {m, rho, v, T, rij, drij, iap, h};
    int64_t len[2] = {4 * n, 2 * n};
    double *out[2] = {gradv, gradT};
    scatter_pairs(gradients_body, &c, nip, 2, len, out);
}

/* c4 per-particle tensors.  Short-range pressure tensor                      */
/*   P_i = p_i I + Pi_i,  Pi_i = -eta (G + G^T - tr G I) - zeta tr G I  (d=2)  */
/* stored divided by rho_i^2 as S = (Sxx, Sxy, Syy); cohesive pco_i/rho_lr^2; */
/* heat flux q_i = -k gradT_i stored as Q = q_i / rho_i^2.                    */
/* The composition P = p_rev + p_rev_lr + pi_irr is spam_complete_force.py:178. */
void orc_tensors(int64_t n, const double *rho, const double *rho_lr, const double *p,
                 const double *pco, const double *gradv, const double *gradT,
                 double eta, double zeta, double kcond, unsigned terms,
                 double *S, double *Sco, double *Q, double *P, double *jq)
{
    OMP_FOR
    for (int64_t i = 0; i < n; ++i) {
        double pi_xx = 0., pi_xy = 0., pi_yy = 0.;
        if ((terms & ORC_VISCOSITY) && gradv) {
            double gxx = gradv[4 * i], gxy = gradv[4 * i + 1];
            double gyx = gradv[4 * i + 2], gyy = gradv[4 * i + 3];
            double tr = gxx + gyy;
            pi_xx = -eta * (2. * gxx - tr) - zeta * tr;
            pi_yy = -eta * (2. * gyy - tr) - zeta * tr;
            pi_xy = -eta * (gxy + gyx);
        }
        double pr = (terms & ORC_PRESSURE) ? p[i] : 0.;
        double r2 = rho[i] * rho[i];
        S[3 * i + 0] = (pr + pi_xx) / r2;
        S[3 * i + 1] = pi_xy / r2;
        S[3 * i + 2] = (pr + pi_yy) / r2;
        Sco[i] = (terms & ORC_COHESION) ? pco[i] / (rho_lr[i] * rho_lr[i]) : 0.;
        double qx = 0., qy = 0.;
        if ((terms & ORC_CONDUCTION) && gradT) {
            qx = -kcond * gradT[2 * i];
            qy = -kcond * gradT[2 * i + 1];
        }
        Q[2 * i] = qx / r2;
        Q[2 * i + 1] = qy / r2;
        if (P) {
            double pc = (terms & ORC_COHESION) ? pco[i] : 0.;
            P[4 * i + 0] = pr + pc + pi_xx;
            P[4 * i + 1] = pi_xy;
            P[4 * i + 2] = pi_xy;
            P[4 * i + 3] = pr + pc + pi_yy;
        }
        if (jq) { jq[2 * i] = qx; jq[2 * i + 1] = qy; }
    }
}

/* c1-c4 rates, scattered per pair (forces.py:338-405, c_forces.pyx:96-111,  */
/* 163-178, 224-237):                                                        */
/*   f   = (S_i + S_j) . dwdx_h + (Sco_i + Sco_j) dwdx_hl                    */
/*   a_i += m_j f,  a_j -= m_i f                                             */
/*   e   = 0.5 dv . f;   c = (Q_i + Q_j) . dwdx_h                            */
/*   u_i += m_j (e + c),  u_j += m_i (e - c)                                 */
/* The reference's SpamForce omits m_j in vdot (forces.py:357-362); with    */
/* m = 1 (every configuration) the two are bitwise identical.               */
typedef struct {
    const double *m, *v, *rij, *drij, *S, *Sco, *Q;
    const int64_t *iap;
    double h, hl;
} rates_ctx;

static void rates_body(const void *cv, int64_t k0, int64_t k1, double *const *out)
{
    const rates_ctx *cx = (const rates_ctx *)cv;
    const double *m = cx->m, *v = cx->v, *drij = cx->drij, *S = cx->S, *Sco = cx->Sco, *Q = cx->Q;
    double *acc = out[0], *udot = out[1];
    for (int64_t k = k0; k < k1; ++k) {
        int64_t i = cx->iap[2 * k], j = cx->iap[2 * k + 1];
        double W, F, Wl, Fl;
        lucy2(cx->rij[k], cx->h, &W, &F);
        lucy2(cx->rij[k], cx->hl, &Wl, &Fl);
        double dx = drij[2 * k], dy = drij[2 * k + 1];
        double wx = F * dx, wy = F * dy, wlx = Fl * dx, wly = Fl * dy;
        double axx = S[3 * i] + S[3 * j], axy = S[3 * i + 1] + S[3 * j + 1];
        double ayy = S[3 * i + 2] + S[3 * j + 2];
        double cs = Sco[i] + Sco[j];
        double fx = axx * wx + axy * wy + cs * wlx;
        double fy = axy * wx + ayy * wy + cs * wly;
        double dvx = v[2 * j] - v[2 * i], dvy = v[2 * j + 1] - v[2 * i + 1];
        double e = 0.5 * (dvx * fx + dvy * fy);
        double c = (Q[2 * i] + Q[2 * j]) * wx + (Q[2 * i + 1] + Q[2 * j + 1]) * wy;
        acc[2 * i] += m[j] * fx;
        acc[2 * i + 1] += m[j] * fy;
        acc[2 * j] -= m[i] * fx;
        acc[2 * j + 1] -= m[i] * fy;
        udot[i] += m[j] * (e + c);
        udot[j] += m[i] * (e - c);
    }
}

void orc_rates(int64_t n, const double *m, const double *v, int64_t nip,
               const int64_t *iap, const double *rij, const double *drij,
               double h, double hl, const double *S, const double *Sco,
               const double *Q, double *acc, double *udot)
{
    memset(acc, 0, sizeof(double) * 2 * n);
    memset(udot, 0, sizeof(double) * n);
    rates_ctx c = {m, v, rij, drij, S, Sco, Q, iap, h, hl};
    int64_t len[2] = {2 * n, n};
    double *out[2] = {acc, udot};
    scatter_pairs(rates_body, &c, nip, 2, len, out);
}

/* ------------------------------------------------------------------------ */
/* Whole derivative evaluation (particles.py:544-570 composition):           */
/* pairs -> separations -> density -> EOS -> gradients -> tensors -> rates.  */
typedef struct {
    int64_t cap;
    int64_t *iap;
    double *rij, *drij;
    double *rho, *rho_lr, *T, *p, *pco, *gradv, *gradT, *S, *Sco, *Q;
} orc_work;

static int work_ensure(orc_work *w, int64_t n, int64_t nip)
{
    if (nip > w->cap) {
        free(w->iap); free(w->rij); free(w->drij);
        w->cap = nip + nip / 4 + 16;
        w->iap = (int64_t *)malloc(sizeof(int64_t) * 2 * w->cap);
        w->rij = (double *)malloc(sizeof(double) * w->cap);
        w->drij = (double *)malloc(sizeof(double) * 2 * w->cap);
        if (!w->iap || !w->rij || !w->drij) return -1;
    }
    if (!w->T) {
        w->rho = (double *)malloc(sizeof(double) * n);
        w->rho_lr = (double *)malloc(sizeof(double) * n);
        w->T = (double *)malloc(sizeof(double) * n);
        w->p = (double *)malloc(sizeof(double) * n);
        w->pco = (double *)malloc(sizeof(double) * n);
        w->gradv = (double *)malloc(sizeof(double) * 4 * n);
        w->gradT = (double *)malloc(sizeof(double) * 2 * n);
        w->S = (double *)malloc(sizeof(double) * 3 * n);
        w->Sco = (double *)malloc(sizeof(double) * n);
        w->Q = (double *)malloc(sizeof(double) * 2 * n);
    }
    return 0;
}

static void work_free(orc_work *w)
{
    free(w->iap); free(w->rij); free(w->drij); free(w->rho); free(w->rho_lr); free(w->T); free(w->p); free(w->pco);
    free(w->gradv); free(w->gradT); free(w->S); free(w->Sco); free(w->Q);
}

static int64_t build_pairs(orc_work *w, int64_t n, const double *r, const orc_params *P)
{
    double R2 = P->rc * P->rc + P->tol * P->tol;
    int64_t nip = orc_pairs_cell(n, r, P->Lx, P->Ly, R2, w->cap, w->iap);
    if (nip > w->cap) {
        if (work_ensure(w, n, nip) != 0) return -1;
        nip = orc_pairs_cell(n, r, P->Lx, P->Ly, R2, w->cap, w->iap);
    }
    orc_separations(r, NULL, P->Lx, P->Ly, nip, w->iap, w->drij, w->rij, NULL, NULL);
    return nip;
}

/* Outputs (any may be NULL): rho, rho_lr, T, acc[n,2], udot, gradv[n,2,2],  */
/* jq[n,2], P[n,2,2].  Returns nip, or -1 on allocation failure.             */
static int64_t derivs_w(orc_work *w, int64_t n, const double *r, const double *v,
                        const double *m, const double *u, const orc_params *P,
                        unsigned terms, double *rho, double *rho_lr, double *T,
                        double *acc, double *udot, double *gradv_out, double *jq_out,
                        double *P_out)
{
    if (work_ensure(w, n, 30 * n + 64) != 0) return -1;
    int64_t nip = build_pairs(w, n, r, P);
    if (nip < 0) return -1;
    orc_density(n, m, nip, w->iap, w->rij, P->h, P->hl, w->rho, w->rho_lr);
    if (rho) memcpy(rho, w->rho, sizeof(double) * n);
    if (rho_lr) memcpy(rho_lr, w->rho_lr, sizeof(double) * n);
    orc_eos(n, w->rho, w->rho_lr, u, P->a, P->b, P->kB, w->T, w->p, w->pco);
    if (T) memcpy(T, w->T, sizeof(double) * n);
    int need_grad = (terms & (ORC_VISCOSITY | ORC_CONDUCTION)) || gradv_out;
    if (need_grad)
        orc_gradients(n, m, w->rho, v, w->T, nip, w->iap, w->rij, w->drij, P->h,
                      w->gradv, w->gradT);
    else {
        memset(w->gradv, 0, sizeof(double) * 4 * n);
        memset(w->gradT, 0, sizeof(double) * 2 * n);
    }
    if (gradv_out) memcpy(gradv_out, w->gradv, sizeof(double) * 4 * n);
    orc_tensors(n, w->rho, w->rho_lr, w->p, w->pco, w->gradv, w->gradT, P->eta, P->zeta,
                P->kcond, terms, w->S, w->Sco, w->Q, P_out, jq_out);
    if (acc && udot)
        orc_rates(n, m, v, nip, w->iap, w->rij, w->drij, P->h, P->hl, w->S, w->Sco, w->Q,
                  acc, udot);
    return nip;
}

int64_t orc_derivatives(int64_t n, const double *r, const double *v, const double *m,
                        const double *u, const orc_params *P, unsigned terms,
                        double *rho, double *rho_lr, double *T, double *acc,
                        double *udot, double *gradv_out, double *jq_out, double *P_out)
{
    orc_work w;
    memset(&w, 0, sizeof(w));
    double *a = acc, *ud = udot;
    if ((acc != NULL) != (udot != NULL)) {
        if (!a) a = (double *)malloc(sizeof(double) * 2 * n);
        if (!ud) ud = (double *)malloc(sizeof(double) * n);
    }
    int64_t nip = derivs_w(&w, n, r, v, m, u, P, terms, rho, rho_lr, T, a, ud,
                           gradv_out, jq_out, P_out);
    if (a != acc) free(a);
    if (ud != udot) free(ud);
    work_free(&w);
    return nip;
}

/* Periodic wrap after a drift (replaces PeriodicBox.apply, box.py:33-47,    */
/* which sets an escaped coordinate to 0 / xmax instead of translating it). */
static inline double wrap1(double x, double L)
{
    if (x >= L) x -= L;
    else if (x < 0.) x += L;
    return x;
}

/* Leapfrog (kick-drift-kick with the two half kicks of consecutive steps    */
/* fused).  State between steps: r_k, v_{k-1/2}, u_{k-1/2}.  The rates at   */
/* step k take the velocity and energy predicted to time k from the last    */
/* rates, v^p = v_{k-1/2} + dt/2 a_{k-1} (and u likewise), so the           */
/* velocity-dependent terms (pressure work, viscosity) stay second order:   */
/*   step 0 : rates(r_0, v_0, u_0);       v += dt/2 a; u += dt/2 udot; drift */
/*   step k : rates(r_k, v^p_k, u^p_k);   v += dt a;   u += dt udot;   drift */
/*   finish : rates(r_K, v^p_K, u^p_K);   v += dt/2 a; u += dt/2 udot        */
/* drift: r += dt v, then periodic wrap.  nsteps steps cost nsteps+1 rate   */
/* evaluations and end synchronised.  energy (optional, nsteps+1 values):   */
/* sum m (v^2/2 + u) at each integer time, v_k = v_{k-1/2} + dt/2 a_k.      */
int orc_run(int64_t n, double *r, double *v, const double *m, double *u,
            const orc_params *P, unsigned terms, int64_t nsteps, double *energy)
{
    double *acc = (double *)malloc(sizeof(double) * 2 * n);
    double *udot = (double *)malloc(sizeof(double) * n);
    double *vp = (double *)malloc(sizeof(double) * 2 * n);
    double *up = (double *)malloc(sizeof(double) * n);
    orc_work w;
    memset(&w, 0, sizeof(w));
    if (!acc || !udot || !vp || !up) return -1;
    const double hdt = 0.5 * P->dt;
    for (int64_t k = 0; k <= nsteps; ++k) {
        OMP_FOR
        for (int64_t i = 0; i < n; ++i) {
            if (k == 0) {
                vp[2 * i] = v[2 * i];
                vp[2 * i + 1] = v[2 * i + 1];
                up[i] = u[i];
            } else {
                vp[2 * i] = v[2 * i] + hdt * acc[2 * i];
                vp[2 * i + 1] = v[2 * i + 1] + hdt * acc[2 * i + 1];
                up[i] = u[i] + hdt * udot[i];
            }
        }
        if (derivs_w(&w, n, r, vp, m, up, P, terms, NULL, NULL, NULL, acc, udot,
                     NULL, NULL, NULL) < 0) {
            free(acc); free(udot); free(vp); free(up); work_free(&w);
            return -1;
        }
        double kick = (k == 0 || k == nsteps) ? hdt : P->dt;
        if (energy) {
            double e = 0.;
            for (int64_t i = 0; i < n; ++i) {
                double hx = v[2 * i] + (k == 0 ? 0. : hdt * acc[2 * i]);
                double hy = v[2 * i + 1] + (k == 0 ? 0. : hdt * acc[2 * i + 1]);
                double uu = u[i] + (k == 0 ? 0. : hdt * udot[i]);
                e += m[i] * (0.5 * (hx * hx + hy * hy) + uu);
            }
            energy[k] = e;
        }
        OMP_FOR
        for (int64_t i = 0; i < n; ++i) {
            v[2 * i] += kick * acc[2 * i];
            v[2 * i + 1] += kick * acc[2 * i + 1];
            u[i] += kick * udot[i];
            if (k < nsteps) {
                r[2 * i] = wrap1(r[2 * i] + P->dt * v[2 * i], P->Lx);
                r[2 * i + 1] = wrap1(r[2 * i + 1] + P->dt * v[2 * i + 1], P->Ly);
            }
        }
    }
    free(acc);
    free(udot);
    free(vp);
    free(up);
    work_free(&w);
    return 0;
}

/* Initial internal energy u = kB T - a rho (properties.py:44-46, as used at */
/* initialisation, properties.py:119).                                       */
void orc_init_u(int64_t n, const double *r, const double *m, const double *T,
                const orc_params *P, double *u)
{
    double *rho = (double *)malloc(sizeof(double) * n);
    double *rl = (double *)malloc(sizeof(double) * n);
    double *zero = (double *)calloc((size_t)(2 * n), sizeof(double));
    double *uu = (double *)calloc((size_t)n, sizeof(double));
    orc_derivatives(n, r, zero, m, uu, P, 0, rho, rl, NULL, NULL, NULL, NULL, NULL, NULL);
    for (int64_t i = 0; i < n; ++i) u[i] = P->kB * T[i] - P->a * rho[i];
    free(rho); free(rl); free(zero); free(uu);
}

int orc_num_threads(void)
{
#if defined(_OPENMP)
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* SPH scatter to a grid: interpolate.splash_map_2d (interpolate.py:39-88),  */
/* restated loop by loop.  Every particle adds a_i m_i / rho_i W(s, h_i) to  */
/* the pixels of its half-open pixel ranges [floor((r - c - min)/res),       */
/* floor((r + c - min)/res)) (wrapped into the grid) whose pixel coordinate  */
/* lies within the cutoff under the minimum image with periods (xmax, ymax); */
/* Z is [nx][ny] (Z[xpx][ypx]).                                             */
void orc_splash_map_2d(int64_t n, const double *r, const double *m, const double *rho,
                       const double *a, const double *h, int nx, const double *x, int ny,
                       const double *y, const double *bounds, double cutoff, double *Z)
{
    const double xmin = bounds[0], xmax = bounds[1], ymin = bounds[2], ymax = bounds[3];
    const double res = x[1] - x[0];
    memset(Z, 0, sizeof(double) * (size_t)nx * ny);
    for (int64_t i = 0; i < n; ++i) {
        const int xlo = (int)floor((r[2 * i] - cutoff - xmin) / res);
        const int ylo = (int)floor((r[2 * i + 1] - cutoff - ymin) / res);
        const int xhi = (int)floor((r[2 * i] + cutoff - xmin) / res);
        const int yhi = (int)floor((r[2 * i + 1] + cutoff - ymin) / res);
        for (int yp0 = ylo; yp0 < yhi; ++yp0) {
            for (int xp0 = xlo; xp0 < xhi; ++xp0) {
                int yp = yp0, xp = xp0;
                if (yp >= ny) yp -= ny;
                if (yp < 0) yp += ny;
                if (xp >= nx) xp -= nx;
                if (xp < 0) xp += nx;
                double dx = r[2 * i] - x[xp], dy = y[yp] - r[2 * i + 1];
                orc_min_image(&dx, &dy, xmax, ymax);
                const double rsq = pair_rsq(dx, dy);
                if (rsq < cutoff * cutoff) {
                    const double s = sqrt(rsq);
                    double w = 0.;
                    if (s < h[i]) {  /* spkernel.lucy_w, 2-D (spkernel.py:54-63) */
                        const double t = 1. - s / h[i];
                        w = lucy_norm(h[i], 2) * (1. + 3. * s / h[i]) * (t * t * t);
                    }
                    Z[(size_t)xp * ny + yp] += a[i] * m[i] / rho[i] * w;
                }
            }
        }
    }
}
