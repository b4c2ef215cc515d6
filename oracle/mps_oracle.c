/* CPU restatement of the mps3d reference multigrid path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
 * this file, as the checker; the product path never calls it.
 *
 * Every routine follows the reference's floating-point evaluation order so
 * that, compiled without FMA contraction, it reproduces the reference's
 * results bit for bit on one rank (checked against the reference binary's
 * golden vectors in tests/test_oracle_golden.py). Citations are to files
 * under the reference tree.
 */
#include "mps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(L, i, j, k) ((size_t)(i) + (size_t)((L)->nx + 2) * ((size_t)(j) + (size_t)((L)->ny + 2) * (size_t)(k)))
#define IIX(L, i, j, k) ((size_t)((i) - 1) + (size_t)(L)->nx * ((size_t)((j) - 1) + (size_t)(L)->ny * (size_t)((k) - 1)))

enum { LEFT = 0, RIGHT, BOTTOM, TOP, BACK, FRONT };

/* One multigrid level on the global grid. */
typedef struct {
    int nx, ny, nz;       /* global cells on this level */
    int lx, ly, lz;       /* per-rank cells on this level */
    int glx, gly, glz;    /* per-rank cells on the finest level (g_lx..) */
    double dx, dy, dz;
    int bottom;           /* coarsest level (gathered to the root) */
    int bc[6];
    double *bcv[6];
    double *D, *vx, *vy, *vz, *reac; /* ghosted coefficient arrays */
    int timederiv;
    double dt;
    int own;              /* coefficient arrays owned by this level */
} lev_t;

static double gv(const double *a, size_t ix) { return a ? a[ix] : 0.0; }

/* findcellcontributions, mgridlib/discretepde.f90:620-670 */
static void cell_contrib(double ul, double ur, double kl, double kr, double vl, double vr,
                         double abv, double d, double *dl, double *dr, double *ol, double *orr)
{
    double dl_ = 0.0, dr_ = 0.0, ol_ = 0.0, or_ = 0.0;
    double kh = 0.5 * (kl + kr);
    dl_ = dl_ + (kh / d) * abv;
    dr_ = dr_ + (kh / d) * abv;
    ol_ = ol_ - ur * (kh / d) * abv;
    or_ = or_ - ul * (kh / d) * abv;
    double vh = 0.5 * (vl + vr);
    if (vh > 0.0) {
        dl_ = dl_ + vh * abv;
        or_ = or_ - vh * ul * abv;
    } else {
        dr_ = dr_ - vh * abv;
        ol_ = ol_ + vh * ur * abv;
    }
    *dl = dl_; *dr = dr_; *ol = ol_; *orr = or_;
}

/* diag and offdiag of one cell, in the accumulation order of find_AX
 * (mgridlib/discretepde.f90:754-1106): x faces, x BCs, y faces, y BCs,
 * z faces, z BCs, reaction, 1/dt. */
static void cell_row(const lev_t *L, const double *u, int i, int j, int k, double *pd, double *po)
{
    double d = 0.0, o = 0.0, dl, dr, ol, orr;
    size_t c = IDX(L, i, j, k);
    const double *D = L->D;
    /* ---- x ---- */
    double abv = 1.0 / L->dx;
    if (i > 1) {
        size_t m = IDX(L, i - 1, j, k);
        cell_contrib(u[m], u[c], gv(D, m), gv(D, c), gv(L->vx, m), gv(L->vx, c), abv, L->dx, &dl, &dr, &ol, &orr);
        d = d + dr; o = o + orr;
    }
    if (i < L->nx) {
        size_t p = IDX(L, i + 1, j, k);
        cell_contrib(u[c], u[p], gv(D, c), gv(D, p), gv(L->vx, c), gv(L->vx, p), abv, L->dx, &dl, &dr, &ol, &orr);
        d = d + dl; o = o + ol;
    }
    if (i == 1) {
        double vb = gv(L->vx, c);
        if (L->bc[LEFT] == ORC_BC_DIRC) {
            d = d + 2.0 * gv(D, c) / (L->dx * L->dx);
            if (vb < 0.0) d = d - vb / L->dx;
        } else if (L->bc[LEFT] == ORC_BC_ZGRD) {
            d = d - vb / L->dx;
        }
    }
    if (i == L->nx) {
        double vb = gv(L->vx, c);
        if (L->bc[RIGHT] == ORC_BC_DIRC) {
            d = d + 2.0 * gv(D, c) / (L->dx * L->dx);
            if (vb > 0.0) d = d + vb / L->dx;
        } else if (L->bc[RIGHT] == ORC_BC_ZGRD) {
            d = d + vb / L->dx;
        }
    }
    /* ---- y ---- */
    abv = 1.0 / L->dy;
    if (j > 1) {
        size_t m = IDX(L, i, j - 1, k);
        cell_contrib(u[m], u[c], gv(D, m), gv(D, c), gv(L->vy, m), gv(L->vy, c), abv, L->dy, &dl, &dr, &ol, &orr);
        d = d + dr; o = o + orr;
    }
    if (j < L->ny) {
        size_t p = IDX(L, i, j + 1, k);
        cell_contrib(u[c], u[p], gv(D, c), gv(D, p), gv(L->vy, c), gv(L->vy, p), abv, L->dy, &dl, &dr, &ol, &orr);
        d = d + dl; o = o + ol;
    }
    if (j == 1) {
        double vb = gv(L->vy, c);
        if (L->bc[BOTTOM] == ORC_BC_DIRC) {
            d = d + 2.0 * gv(D, c) / (L->dy * L->dy);
            if (vb < 0.0) d = d - vb / L->dy;
        } else if (L->bc[BOTTOM] == ORC_BC_ZGRD) {
            d = d - vb / L->dy;
        }
    }
    if (j == L->ny) {
        double vb = gv(L->vy, c);
        if (L->bc[TOP] == ORC_BC_DIRC) {
            d = d + 2.0 * gv(D, c) / (L->dy * L->dy);
            if (vb > 0.0) d = d + vb / L->dy;
        } else if (L->bc[TOP] == ORC_BC_ZGRD) {
            d = d + vb / L->dy;
        }
    }
    /* ---- z ---- */
    abv = 1.0 / L->dz;
    if (k > 1) {
        size_t m = IDX(L, i, j, k - 1);
        cell_contrib(u[m], u[c], gv(D, m), gv(D, c), gv(L->vz, m), gv(L->vz, c), abv, L->dz, &dl, &dr, &ol, &orr);
        d = d + dr; o = o + orr;
    }
    if (k < L->nz) {
        size_t p = IDX(L, i, j, k + 1);
        cell_contrib(u[c], u[p], gv(D, c), gv(D, p), gv(L->vz, c), gv(L->vz, p), abv, L->dz, &dl, &dr, &ol, &orr);
        d = d + dl; o = o + ol;
    }
    if (k == 1) {
        double vb = gv(L->vz, c);
        if (L->bc[BACK] == ORC_BC_DIRC) {
            d = d + 2.0 * gv(D, c) / (L->dz * L->dz);
            if (vb < 0.0) d = d - vb / L->dz;
        } else if (L->bc[BACK] == ORC_BC_ZGRD) {
            d = d - vb / L->dz;
        }
    }
    if (k == L->nz) {
        double vb = gv(L->vz, c);
        if (L->bc[FRONT] == ORC_BC_DIRC) {
            d = d + 2.0 * gv(D, c) / (L->dz * L->dz);
            if (vb > 0.0) d = d + vb / L->dz;
        } else if (L->bc[FRONT] == ORC_BC_ZGRD) {
            d = d + vb / L->dz;
        }
    }
    d = d - gv(L->reac, c);
    if (L->timederiv) d = d + 1.0 / L->dt;
    *pd = d;
    *po = o;
}

/* find_AX, mgridlib/discretepde.f90:672-1119 */
static void level_apply(const lev_t *L, const double *u, double *ax, double *diag, double *off)
{
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) {
                double d, o;
                cell_row(L, u, i, j, k, &d, &o);
                size_t n = IIX(L, i, j, k);
                if (ax) ax[n] = d * u[IDX(L, i, j, k)] + o;
                if (diag) diag[n] = d;
                if (off) off[n] = o;
            }
}

/* find_bvec, mgridlib/discretepde.f90:382-618 */
static void level_bvec(const lev_t *L, const double *sterm, const double *source, int applybc,
                       double *b)
{
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) {
                size_t c = IDX(L, i, j, k);
                double v = 0.0;
                if (applybc) {
                    if (i == 1) {
                        double g = L->bcv[LEFT][(j - 1) + (size_t)L->ny * (k - 1)];
                        if (L->bc[LEFT] == ORC_BC_DIRC) {
                            double vb = gv(L->vx, c);
                            v = v + 2.0 * gv(L->D, c) * g / (L->dx * L->dx);
                            if (vb > 0.0) v = v + vb * g / L->dx;
                        } else if (L->bc[LEFT] == ORC_BC_FLUX) {
                            v = v - g / L->dx;
                        }
                    }
                    if (i == L->nx) {
                        double g = L->bcv[RIGHT][(j - 1) + (size_t)L->ny * (k - 1)];
                        if (L->bc[RIGHT] == ORC_BC_DIRC) {
                            double vb = gv(L->vx, c);
                            v = v + 2.0 * gv(L->D, c) * g / (L->dx * L->dx);
                            if (vb < 0.0) v = v - vb * g / L->dx;
                        } else if (L->bc[RIGHT] == ORC_BC_FLUX) {
                            v = v - g / L->dx;
                        }
                    }
                    if (j == 1) {
                        double g = L->bcv[BOTTOM][(i - 1) + (size_t)L->nx * (k - 1)];
                        if (L->bc[BOTTOM] == ORC_BC_DIRC) {
                            double vb = gv(L->vy, c);
                            v = v + 2.0 * gv(L->D, c) * g / (L->dy * L->dy);
                            if (vb > 0.0) v = v + vb * g / L->dy;
                        } else if (L->bc[BOTTOM] == ORC_BC_FLUX) {
                            v = v - g / L->dy;
                        }
                    }
                    if (j == L->ny) {
                        double g = L->bcv[TOP][(i - 1) + (size_t)L->nx * (k - 1)];
                        if (L->bc[TOP] == ORC_BC_DIRC) {
                            double vb = gv(L->vy, c);
                            v = v + 2.0 * gv(L->D, c) * g / (L->dy * L->dy);
                            if (vb < 0.0) v = v - vb * g / L->dy;
                        } else if (L->bc[TOP] == ORC_BC_FLUX) {
                            v = v - g / L->dy;
                        }
                    }
                    if (k == 1) {
                        double g = L->bcv[BACK][(i - 1) + (size_t)L->nx * (j - 1)];
                        if (L->bc[BACK] == ORC_BC_DIRC) {
                            double vb = gv(L->vz, c);
                            v = v + 2.0 * gv(L->D, c) * g / (L->dz * L->dz);
                            if (vb > 0.0) v = v + vb * g / L->dz;
                        } else if (L->bc[BACK] == ORC_BC_FLUX) {
                            v = v - g / L->dz;
                        }
                    }
                    if (k == L->nz) {
                        double g = L->bcv[FRONT][(i - 1) + (size_t)L->nx * (j - 1)];
                        if (L->bc[FRONT] == ORC_BC_DIRC) {
                            double vb = gv(L->vz, c);
                            v = v + 2.0 * gv(L->D, c) * g / (L->dz * L->dz);
                            if (vb < 0.0) v = v - vb * g / L->dz;
                        } else if (L->bc[FRONT] == ORC_BC_FLUX) {
                            v = v - g / L->dz;
                        }
                    }
                }
                size_t n = IIX(L, i, j, k);
                v = v + (sterm ? sterm[n] : 0.0);
                if (applybc) v = v + gv(source, c);
                b[n] = v;
            }
}

/* Red/black colour of a cell: find_redblack_indices (discretepde.f90:13-63)
 * with local_to_global_index (globalvars.f90:121-134), whose offsets use the
 * finest per-rank size g_l at every level. Returns 1 for red (odd). */
static int is_red(const lev_t *L, int i, int j, int k)
{
    long gi, gj, gk;
    if (L->bottom) {
        gi = i; gj = j; gk = k; /* root rank, serial indices */
    } else {
        gi = (long)((i - 1) / L->lx) * L->glx + (i - 1) % L->lx + 1;
        gj = (long)((j - 1) / L->ly) * L->gly + (j - 1) % L->ly + 1;
        gk = (long)((k - 1) / L->lz) * L->glz + (k - 1) % L->lz + 1;
    }
    return ((gi + gj + gk) % 2) != 0;
}

/* one colour pass: u = (b - off)/diag on cells of one colour, with the
 * operator evaluated from the state before the pass */
static void colour_pass(const lev_t *L, double *u, const double *b, int red)
{
    /* offdiag of a cell of one colour only involves the other colour, so
     * evaluating it on the fly equals the reference's full find_AX */
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) {
                if (is_red(L, i, j, k) != red) continue;
                double d, o;
                cell_row(L, u, i, j, k, &d, &o);
                u[IDX(L, i, j, k)] = (b[IIX(L, i, j, k)] - o) / d;
            }
}

/* perform_gseidel_smoothing, discretepde.f90:132-236 */
static void level_smooth(const lev_t *L, double *u, const double *b, int niter, double *res)
{
    size_t n = (size_t)L->nx * L->ny * L->nz;
    double *ax = malloc(n * sizeof(double));
    for (int it = 0; it < niter; ++it) {
        colour_pass(L, u, b, 1);
        colour_pass(L, u, b, 0);
    }
    level_apply(L, u, ax, NULL, NULL);
    for (size_t q = 0; q < n; ++q) res[q] = b[q] - ax[q];
    free(ax);
}

/* restriction, mgridsteps.f90:11-53 */
void orc_restrict(int nx, int ny, int nz, const double *f, double *c)
{
    int cx = nx / 2, cy = ny / 2, cz = nz / 2;
#define F(i, j, k) f[(size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1))]
    for (int k = 1; k <= cz; ++k)
        for (int j = 1; j <= cy; ++j)
            for (int i = 1; i <= cx; ++i) {
                double s = 0.0;
                s = s + F(2 * i - 1, 2 * j - 1, 2 * k - 1) + F(2 * i, 2 * j - 1, 2 * k - 1);
                s = s + F(2 * i - 1, 2 * j, 2 * k - 1) + F(2 * i, 2 * j, 2 * k - 1);
                s = s + F(2 * i - 1, 2 * j - 1, 2 * k) + F(2 * i, 2 * j - 1, 2 * k);
                s = s + F(2 * i - 1, 2 * j, 2 * k) + F(2 * i, 2 * j, 2 * k);
                c[(size_t)(i - 1) + (size_t)cx * ((size_t)(j - 1) + (size_t)cy * (size_t)(k - 1))] = (1.0 / 8.0) * s;
            }
#undef F
}

/* restriction of a ghosted array's interior into a ghosted coarse array */
static void restrict_ghosted(const lev_t *F, const lev_t *C, const double *f, double *c)
{
    size_t n = (size_t)F->nx * F->ny * F->nz, nc = (size_t)C->nx * C->ny * C->nz;
    double *fi = malloc(n * sizeof(double)), *ci = malloc(nc * sizeof(double));
    for (int k = 1; k <= F->nz; ++k)
        for (int j = 1; j <= F->ny; ++j)
            for (int i = 1; i <= F->nx; ++i) fi[IIX(F, i, j, k)] = f[IDX(F, i, j, k)];
    orc_restrict(F->nx, F->ny, F->nz, fi, ci);
    for (int k = 1; k <= C->nz; ++k)
        for (int j = 1; j <= C->ny; ++j)
            for (int i = 1; i <= C->nx; ++i) c[IDX(C, i, j, k)] = ci[IIX(C, i, j, k)];
    free(fi);
    free(ci);
}

/* restrict_boundary for one face array, mgridsteps.f90:55-128 */
void orc_restrict_face(int a, int b, const double *f, double *c)
{
    int ca = a / 2, cb = b / 2;
    for (int q = 1; q <= cb; ++q)
        for (int p = 1; p <= ca; ++p)
            c[(p - 1) + (size_t)ca * (q - 1)] = 0.25 * (f[(2 * p - 2) + (size_t)a * (2 * q - 2)] +
                                                         f[(2 * p - 1) + (size_t)a * (2 * q - 2)] +
                                                         f[(2 * p - 2) + (size_t)a * (2 * q - 1)] +
                                                         f[(2 * p - 1) + (size_t)a * (2 * q - 1)]);
}

/* prolongation1, mgridsteps.f90:170-216 */
void orc_prolong1(int cx, int cy, int cz, const double *e, double *fine)
{
    int fx = 2 * cx, fy = 2 * cy;
#define E(i, j, k) e[(size_t)(i) + (size_t)(cx + 2) * ((size_t)(j) + (size_t)(cy + 2) * (size_t)(k))]
#define FI(i, j, k) fine[(size_t)((i) - 1) + (size_t)fx * ((size_t)((j) - 1) + (size_t)fy * (size_t)((k) - 1))]
    for (int kc = 1; kc <= cz; ++kc)
        for (int jc = 1; jc <= cy; ++jc)
            for (int ic = 1; ic <= cx; ++ic) {
                double ec = E(ic, jc, kc);
                double sx = 0.5 * (E(ic + 1, jc, kc) - E(ic - 1, jc, kc));
                double sy = 0.5 * (E(ic, jc + 1, kc) - E(ic, jc - 1, kc));
                double sz = 0.5 * (E(ic, jc, kc + 1) - E(ic, jc, kc - 1));
                FI(2 * ic - 1, 2 * jc - 1, 2 * kc - 1) = ec + 0.25 * (-sx - sy - sz);
                FI(2 * ic, 2 * jc - 1, 2 * kc - 1) = ec + 0.25 * (+sx - sy - sz);
                FI(2 * ic - 1, 2 * jc, 2 * kc - 1) = ec + 0.25 * (-sx + sy - sz);
                FI(2 * ic, 2 * jc, 2 * kc - 1) = ec + 0.25 * (+sx + sy - sz);
                FI(2 * ic - 1, 2 * jc - 1, 2 * kc) = ec + 0.25 * (-sx - sy + sz);
                FI(2 * ic, 2 * jc - 1, 2 * kc) = ec + 0.25 * (+sx - sy + sz);
                FI(2 * ic - 1, 2 * jc, 2 * kc) = ec + 0.25 * (-sx + sy + sz);
                FI(2 * ic, 2 * jc, 2 * kc) = ec + 0.25 * (+sx + sy + sz);
            }
#undef E
#undef FI
}

/* update_boundary_ghostvalues, discretepde.f90:1121-1267 */
static void level_ghosts(const lev_t *L, double *u)
{
    int nx = L->nx, ny = L->ny, nz = L->nz;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            size_t f = (j - 1) + (size_t)ny * (k - 1);
            u[IDX(L, 0, j, k)] = L->bc[LEFT] == ORC_BC_DIRC ? 2.0 * L->bcv[LEFT][f] - u[IDX(L, 1, j, k)]
                                                             : u[IDX(L, 1, j, k)];
            u[IDX(L, nx + 1, j, k)] = L->bc[RIGHT] == ORC_BC_DIRC ? 2.0 * L->bcv[RIGHT][f] - u[IDX(L, nx, j, k)]
                                                                   : u[IDX(L, nx, j, k)];
        }
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
            size_t f = (i - 1) + (size_t)nx * (k - 1);
            u[IDX(L, i, 0, k)] = L->bc[BOTTOM] == ORC_BC_DIRC ? 2.0 * L->bcv[BOTTOM][f] - u[IDX(L, i, 1, k)]
                                                               : u[IDX(L, i, 1, k)];
            u[IDX(L, i, ny + 1, k)] = L->bc[TOP] == ORC_BC_DIRC ? 2.0 * L->bcv[TOP][f] - u[IDX(L, i, ny, k)]
                                                                 : u[IDX(L, i, ny, k)];
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            size_t f = (i - 1) + (size_t)nx * (j - 1);
            u[IDX(L, i, j, 0)] = L->bc[BACK] == ORC_BC_DIRC ? 2.0 * L->bcv[BACK][f] - u[IDX(L, i, j, 1)]
                                                             : u[IDX(L, i, j, 1)];
            u[IDX(L, i, j, nz + 1)] = L->bc[FRONT] == ORC_BC_DIRC ? 2.0 * L->bcv[FRONT][f] - u[IDX(L, i, j, nz)]
                                                                   : u[IDX(L, i, j, nz)];
        }
}

double orc_rms(int n, const double *r)
{
    double s = 0.0;
    for (int q = 0; q < n; ++q) s = s + r[q] * r[q];
    return sqrt(s / (double)n);
}

/* findnorm over a whole ghosted array (vectorfunctions.f90:6-22) */
static double gnorm(size_t n, const double *v)
{
    double s = 0.0;
    for (size_t q = 0; q < n; ++q) s = s + v[q] * v[q];
    return sqrt(s);
}

/* gseidel_precond with niter = 1, discretepde.f90:238-337 */
static void precond(const lev_t *L, double *minvx, const double *x)
{
    size_t ng = (size_t)(L->nx + 2) * (L->ny + 2) * (L->nz + 2);
    memset(minvx, 0, ng * sizeof(double));
    size_t n = (size_t)L->nx * L->ny * L->nz;
    double *xi = malloc(n * sizeof(double));
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) xi[IIX(L, i, j, k)] = x[IDX(L, i, j, k)];
    if (n == 1) {
        double d, o;
        cell_row(L, minvx, 1, 1, 1, &d, &o);
        minvx[IDX(L, 1, 1, 1)] = xi[0] / d;
    } else {
        colour_pass(L, minvx, xi, 1);
        colour_pass(L, minvx, xi, 0);
    }
    free(xi);
}

/* performgmres, gmres.f90:197-434, with find_AX and gseidel_precond */
static int level_gmres(const lev_t *L, const double *b, double *x, int maxk, int nrestarts,
                       double tol)
{
    size_t ng = (size_t)(L->nx + 2) * (L->ny + 2) * (L->nz + 2);
    size_t n = (size_t)L->nx * L->ny * L->nz;
    double *bg = calloc(ng, sizeof(double)), *minvb = calloc(ng, sizeof(double));
    double *x0 = calloc(ng, sizeof(double)), *r0 = calloc(ng, sizeof(double));
    double *r = calloc(ng, sizeof(double)), *w = calloc(ng, sizeof(double));
    double *av = calloc(ng, sizeof(double)), *ax = malloc(n * sizeof(double));
    double *V = calloc(ng * (size_t)(maxk + 1), sizeof(double));
    double *H = calloc((size_t)(maxk + 1) * maxk, sizeof(double));
    double *cs = calloc(maxk, sizeof(double)), *sn = calloc(maxk, sizeof(double));
    double *be = calloc(maxk + 1, sizeof(double)), *y = calloc(maxk, sizeof(double));
    int status = 0;
#define HM(i, j) H[(size_t)((i) - 1) + (size_t)(maxk + 1) * (size_t)((j) - 1)]
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) bg[IDX(L, i, j, k)] = b[IIX(L, i, j, k)];
    for (int rs = 1; rs <= nrestarts; ++rs) {
        memset(V, 0, ng * (size_t)(maxk + 1) * sizeof(double));
        memset(H, 0, (size_t)(maxk + 1) * maxk * sizeof(double));
        memset(cs, 0, maxk * sizeof(double));
        memset(sn, 0, maxk * sizeof(double));
        memset(be, 0, (maxk + 1) * sizeof(double));
        memset(y, 0, maxk * sizeof(double));
        memset(r0, 0, ng * sizeof(double));
        precond(L, minvb, bg);
        double bnorm = gnorm(ng, minvb);
        if (bnorm == 0.0) {
            memset(x, 0, ng * sizeof(double));
            break;
        }
        memcpy(x0, x, ng * sizeof(double));
        level_apply(L, x0, ax, NULL, NULL);
        for (int k = 1; k <= L->nz; ++k)
            for (int j = 1; j <= L->ny; ++j)
                for (int i = 1; i <= L->nx; ++i)
                    r0[IDX(L, i, j, k)] = b[IIX(L, i, j, k)] - ax[IIX(L, i, j, k)];
        precond(L, r, r0);
        double beta = gnorm(ng, r);
        for (size_t q = 0; q < ng; ++q) V[q] = r[q] / beta;
        be[0] = beta;
        beta = gnorm(ng, r);
        if (fabs(beta) < tol) break;
        memcpy(x0, x, ng * sizeof(double));
        for (size_t q = 0; q < ng; ++q) V[q] = r[q] / beta;
        int optk = maxk, lucky = 0;
        double residnorm = 0.0;
        for (int kd = 1; kd <= maxk; ++kd) {
            memcpy(x, x0, ng * sizeof(double));
            /* arnoldialgorithm, gmres.f90:9-122 */
            const double *vj = V + ng * (size_t)(kd - 1);
            memset(av, 0, ng * sizeof(double));
            level_apply(L, vj, ax, NULL, NULL);
            for (int k = 1; k <= L->nz; ++k)
                for (int j = 1; j <= L->ny; ++j)
                    for (int i = 1; i <= L->nx; ++i) av[IDX(L, i, j, k)] = ax[IIX(L, i, j, k)];
            precond(L, w, av); /* w holds MinvAvj */
            double *wj = malloc(ng * sizeof(double));
            memcpy(wj, w, ng * sizeof(double));
            for (int ii = 1; ii <= kd; ++ii) {
                const double *vi = V + ng * (size_t)(ii - 1);
                double h = 0.0;
                for (int k = 1; k <= L->nz; ++k)
                    for (int j = 1; j <= L->ny; ++j)
                        for (int i = 1; i <= L->nx; ++i) h = h + w[IDX(L, i, j, k)] * vi[IDX(L, i, j, k)];
                HM(ii, kd) = h;
                for (size_t q = 0; q < ng; ++q) wj[q] = wj[q] - h * vi[q];
            }
            double hn = gnorm(ng, wj);
            HM(kd + 1, kd) = hn;
            int nanflag = 0;
            if (hn > 1e-25) {
                double *vn = V + ng * (size_t)kd;
                for (size_t q = 0; q < ng; ++q) vn[q] = wj[q] / hn;
            } else if (!(hn != hn) && !((hn / hn != 1.0) && hn != 0.0)) {
                lucky = 1;
            } else {
                nanflag = 1;
            }
            free(wj);
            if (nanflag) { status = -1; goto done; }
            /* triangularize, gmres.f90:159-195 */
            for (int ii = 1; ii <= kd - 1; ++ii) {
                double t = cs[ii - 1] * HM(ii, kd) + sn[ii - 1] * HM(ii + 1, kd);
                HM(ii + 1, kd) = -sn[ii - 1] * HM(ii, kd) + cs[ii - 1] * HM(ii + 1, kd);
                HM(ii, kd) = t;
            }
            double rr = HM(kd, kd), ss = HM(kd + 1, kd), c, s;
            if (rr == 0.0) { c = 0.0; s = 1.0; }
            else { c = rr / sqrt(rr * rr + ss * ss); s = ss / sqrt(rr * rr + ss * ss); }
            HM(kd, kd) = c * HM(kd, kd) + s * HM(kd + 1, kd);
            HM(kd + 1, kd) = 0.0;
            cs[kd - 1] = c;
            sn[kd - 1] = s;
            be[kd] = -s * be[kd - 1];
            be[kd - 1] = c * be[kd - 1];
            residnorm = fabs(be[kd]);
            if (lucky) { optk = kd; break; }
        }
        /* uppertrisolve, gmres.f90:124-142 */
        for (int ii = optk; ii >= 1; --ii) {
            y[ii - 1] = be[ii - 1];
            for (int jj = ii + 1; jj <= optk; ++jj) y[ii - 1] = y[ii - 1] - HM(ii, jj) * y[jj - 1];
            y[ii - 1] = y[ii - 1] / HM(ii, ii);
        }
        for (int jj = 1; jj <= optk; ++jj) {
            const double *v = V + ng * (size_t)(jj - 1);
            for (size_t q = 0; q < ng; ++q) x[q] = x[q] + y[jj - 1] * v[q];
        }
        if (residnorm / bnorm <= tol || lucky) break;
    }
done:
#undef HM
    free(bg); free(minvb); free(x0); free(r0); free(r); free(w); free(av); free(ax);
    free(V); free(H); free(cs); free(sn); free(be); free(y);
    return status;
}

/* ------------------------------------------------------------------ */
/* level hierarchy                                                      */

static int idiv_ok(int n, int p) { return p > 0 && n % p == 0; }

/* Builds the level list; only_finest stops after level 0. Returns NULL when
 * a non-bottom level has an odd per-rank size (the reference truncates
 * those silently in restriction, mgridsteps.f90:21-23). */
static lev_t *build_levels(const orc_problem *p, int *nlev_out, int only_finest)
{
    int glx = p->nx / p->px, gly = p->ny / p->py, glz = p->nz / p->pz;
    double gdx = p->lenx / p->nx, gdy = p->leny / p->ny, gdz = p->lenz / p->nz;
    double llx = glx * gdx, lly = gly * gdy, llz = glz * gdz; /* g_llen, par_decompose.f90:40-42 */
    int cap = 64, n = 0;
    lev_t *Ls = calloc(cap, sizeof(lev_t));
    int lx = glx, ly = gly, lz = glz;
    for (;;) {
        lev_t *L = &Ls[n];
        L->lx = lx; L->ly = ly; L->lz = lz;
        L->glx = glx; L->gly = gly; L->glz = glz;
        L->nx = lx * p->px; L->ny = ly * p->py; L->nz = lz * p->pz;
        L->timederiv = p->timederiv;
        L->dt = p->dt;
        for (int f = 0; f < 6; ++f) L->bc[f] = p->bc[f];
        L->bottom = (lx <= 1 || ly <= 1 || lz <= 1);
        if (L->bottom) {
            /* bottomsolve: serial lengths and sizes, mgridsteps.f90:301-336 */
            L->dx = (llx * p->px) / L->nx;
            L->dy = (lly * p->py) / L->ny;
            L->dz = (llz * p->pz) / L->nz;
        } else {
            L->dx = llx / lx;
            L->dy = lly / ly;
            L->dz = llz / lz;
        }
        ++n;
        if (L->bottom || only_finest) break;
        if ((lx & 1) || (ly & 1) || (lz & 1)) {
            free(Ls);
            return NULL;
        }
        lx /= 2; ly /= 2; lz /= 2;
    }
    /* coefficients and BC faces: finest borrows, coarser levels restrict */
    Ls[0].D = (double *)p->D; Ls[0].vx = (double *)p->vx; Ls[0].vy = (double *)p->vy;
    Ls[0].vz = (double *)p->vz; Ls[0].reac = (double *)p->reac;
    for (int f = 0; f < 6; ++f) Ls[0].bcv[f] = (double *)p->bcv[f];
    for (int l = 1; l < n; ++l) {
        lev_t *F = &Ls[l - 1], *C = &Ls[l];
        size_t ngc = (size_t)(C->nx + 2) * (C->ny + 2) * (C->nz + 2);
        C->own = 1;
        double **dst[5] = {&C->D, &C->vx, &C->vy, &C->vz, &C->reac};
        double *src[5] = {F->D, F->vx, F->vy, F->vz, F->reac};
        for (int q = 0; q < 5; ++q) {
            *dst[q] = calloc(ngc, sizeof(double));
            if (src[q]) restrict_ghosted(F, C, src[q], *dst[q]);
        }
        int fa[6] = {F->ny, F->ny, F->nx, F->nx, F->nx, F->nx};
        int fb[6] = {F->nz, F->nz, F->nz, F->nz, F->ny, F->ny};
        for (int f = 0; f < 6; ++f) {
            C->bcv[f] = calloc((size_t)(fa[f] / 2 > 0 ? fa[f] / 2 : 1) * (fb[f] / 2 > 0 ? fb[f] / 2 : 1), sizeof(double));
            orc_restrict_face(fa[f], fb[f], F->bcv[f], C->bcv[f]);
        }
    }
    *nlev_out = n;
    return Ls;
}

static void free_levels(lev_t *Ls, int n)
{
    for (int l = 1; l < n; ++l) {
        free(Ls[l].D); free(Ls[l].vx); free(Ls[l].vy); free(Ls[l].vz); free(Ls[l].reac);
        for (int f = 0; f < 6; ++f) free(Ls[l].bcv[f]);
    }
    free(Ls);
}

/* bottomsolve, mgridsteps.f90:218-382, on the gathered coarse grid */
static void level_bottom(const lev_t *L, double *u, const double *sterm, int applybc,
                         const double *source)
{
    size_t ng = (size_t)(L->nx + 2) * (L->ny + 2) * (L->nz + 2);
    size_t n = (size_t)L->nx * L->ny * L->nz;
    double *x = calloc(ng, sizeof(double)), *b = malloc(n * sizeof(double));
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) x[IDX(L, i, j, k)] = u[IDX(L, i, j, k)];
    /* serial_bcvals are allocated but never filled (mgridsteps.f90:514) */
    lev_t Z = *L;
    double *zf = calloc((size_t)(L->nx + L->ny + L->nz) * (L->nx + L->ny + L->nz) + 1, sizeof(double));
    for (int f = 0; f < 6; ++f) Z.bcv[f] = zf;
    level_bvec(&Z, sterm, source, applybc, b);
    free(zf);
    /* err_tol = ERRTOLR*0.0001 with a default-real 0.0001 literal */
    double tol = 1e-11 * (double)0.0001f;
    level_gmres(L, b, x, L->nx * L->ny * L->nz, 1, tol);
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) u[IDX(L, i, j, k)] = x[IDX(L, i, j, k)];
    free(x);
    free(b);
}

/* perform_vcycle, mgridsteps.f90:384-579 */
static double vcycle_rec(lev_t *Ls, int l, double *u, const double *sterm, const double *source)
{
    lev_t *L = &Ls[l];
    int applybc = (l == 0);
    if (L->bottom) {
        level_bottom(L, u, sterm, applybc, source);
        return 0.0; /* final_res is never assigned by performgmres (gmres.f90:425-432) */
    }
    lev_t *C = &Ls[l + 1];
    size_t n = (size_t)L->nx * L->ny * L->nz, nc = (size_t)C->nx * C->ny * C->nz;
    size_t ngc = (size_t)(C->nx + 2) * (C->ny + 2) * (C->nz + 2);
    double *b = malloc(n * sizeof(double)), *res = malloc(n * sizeof(double));
    double *uold = calloc(ngc, sizeof(double)), *unew = malloc(ngc * sizeof(double));
    double *rc = malloc(nc * sizeof(double)), *tau = malloc(nc * sizeof(double));
    double *ax = malloc(nc * sizeof(double)), *ef = malloc(n * sizeof(double));
    level_bvec(L, sterm, source, applybc, b);
    level_smooth(L, u, b, 2, res);
    restrict_ghosted(L, C, u, uold);
    orc_restrict(L->nx, L->ny, L->nz, res, rc);
    level_apply(C, uold, ax, NULL, NULL);
    for (size_t q = 0; q < nc; ++q) tau[q] = ax[q] + rc[q];
    level_ghosts(C, uold);
    memcpy(unew, uold, ngc * sizeof(double));
    vcycle_rec(Ls, l + 1, unew, tau, NULL);
    level_ghosts(C, unew);
    for (size_t q = 0; q < ngc; ++q) unew[q] = unew[q] - uold[q]; /* error_2h */
    orc_prolong1(C->nx, C->ny, C->nz, unew, ef);
    for (int k = 1; k <= L->nz; ++k)
        for (int j = 1; j <= L->ny; ++j)
            for (int i = 1; i <= L->nx; ++i) u[IDX(L, i, j, k)] = u[IDX(L, i, j, k)] + ef[IIX(L, i, j, k)];
    level_smooth(L, u, b, 2, res);
    double rn = orc_rms((int)n, res);
    free(b); free(res); free(uold); free(unew); free(rc); free(tau); free(ax); free(ef);
    return rn;
}

static int check_problem(const orc_problem *p)
{
    return p->nx > 0 && p->ny > 0 && p->nz > 0 && idiv_ok(p->nx, p->px) && idiv_ok(p->ny, p->py) &&
           idiv_ok(p->nz, p->pz);
}

double orc_vcycle(const orc_problem *p, double *u, const double *sterm)
{
    if (!check_problem(p)) return -1.0;
    int nl;
    lev_t *Ls = build_levels(p, &nl, 0);
    if (!Ls) return -1.0;
    double r = vcycle_rec(Ls, 0, u, sterm, p->source);
    free_levels(Ls, nl);
    return r;
}

int orc_solve(const orc_problem *p, double *u, int max_cycles, double tol, double *history)
{
    if (!check_problem(p)) return -1;
    int nl, it;
    lev_t *Ls = build_levels(p, &nl, 0);
    if (!Ls) return -1;
    for (it = 0; it < max_cycles; ++it) {
        double r = vcycle_rec(Ls, 0, u, NULL, p->source);
        if (history) history[it] = r;
        if (r <= tol) { ++it; break; }
    }
    free_levels(Ls, nl);
    return it;
}

static lev_t finest(const orc_problem *p)
{
    int nl;
    lev_t *Ls = build_levels(p, &nl, 1);
    lev_t L = Ls[0];
    free(Ls);
    L.bottom = 0;
    return L;
}

void orc_apply(const orc_problem *p, const double *u, double *ax, double *diag, double *off)
{
    lev_t L = finest(p);
    level_apply(&L, u, ax, diag, off);
}

void orc_bvec(const orc_problem *p, const double *sterm, int applybc, double *b)
{
    lev_t L = finest(p);
    level_bvec(&L, sterm, p->source, applybc, b);
}

void orc_smooth(const orc_problem *p, double *u, const double *sterm, int niter, double *res)
{
    lev_t L = finest(p);
    size_t n = (size_t)L.nx * L.ny * L.nz;
    double *b = malloc(n * sizeof(double));
    level_bvec(&L, sterm, p->source, 1, b);
    level_smooth(&L, u, b, niter, res);
    free(b);
}

void orc_ghosts(const orc_problem *p, double *u)
{
    lev_t L = finest(p);
    level_ghosts(&L, u);
}

int orc_gmres(const orc_problem *p, const double *b, double *x, int maxkspdim, int nrestarts,
              double tol)
{
    lev_t L = finest(p);
    L.bottom = 1; /* single rank: serial colouring */
    return level_gmres(&L, b, x, maxkspdim, nrestarts, tol);
}

/* minmodlimiter, plasma_pde/poissoneq.f90:549-563 */
static double minmod_phi(double ntr, double dtr)
{
    if (dtr != 0.0) {
        double r = ntr / dtr;
        double m = r < 1.0 ? r : 1.0;
        return m > 0.0 ? m : 0.0;
    }
    return 1.0;
}

/* update_efield, plasma_pde/poissoneq.f90:565-680: the ghost values of phi
 * from the boundary conditions (update_boundary_ghostvalues, in place), the
 * central difference 0.5 (phi(+1) - phi(-1)) / d per direction (:592-606),
 * the minmod limiter when asked (:613-656), E = -E * scale (:660), and the
 * zero-gradient ghost faces of each component (:662-678; ghost edges and
 * corners are left as they were). efield: three ghosted components one after
 * the other, as efield(:,:,:,NDIM). */
void orc_efield(const orc_problem *p, double *phi, double *efield, double scale, int limiter)
{
    lev_t L = finest(p);
    level_ghosts(&L, phi);
    size_t ng = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    double *E[3] = {efield, efield + ng, efield + 2 * ng};
    const size_t sy = (size_t)(L.nx + 2), sz = sy * (size_t)(L.ny + 2);
    const size_t st[3] = {1, sy, sz};
    const double d[3] = {L.dx, L.dy, L.dz};
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i) {
                size_t c = IDX(&L, i, j, k);
                for (int q = 0; q < 3; ++q) {
                    double e = 0.5 * (phi[c + st[q]] - phi[c - st[q]]) / d[q];
                    if (limiter) e = minmod_phi(phi[c] - phi[c - st[q]], phi[c + st[q]] - phi[c]) * e;
                    E[q][c] = -e * scale;
                }
            }
    lev_t Z = L;
    for (int f = 0; f < 6; ++f) Z.bc[f] = ORC_BC_ZGRD;
    for (int q = 0; q < 3; ++q) level_ghosts(&Z, E[q]);
}

/* ---------------------------------------------------------------------
 * Transport assembly, plasma_pde/{edeneq,ionseq,eenergyeq,poissoneq}.f90,
 * with the table chemistry of plasma_pde/chemistry/ion_elec_chem.f90. Each
 * routine follows its reference loop by loop: zero the outputs, fill the
 * interior, scale the whole arrays, add soln/dt over the whole arrays.
 * Constants: plasma_pde/fundconstants.f90:3-9.
 * --------------------------------------------------------------------- */
#define ORC_ECHARGE 1.602176e-19
#define ORC_EPS0 8.854188e-12
#define ORC_KBMAN 1.380650e-23

/* real**integer (m >= 0) by repeated squaring */
static double orc_powi(double x, int m)
{
    double r = (m & 1) ? x : 1.0;
    while (m >>= 1) {
        x = x * x;
        if (m & 1) r = r * x;
    }
    return r;
}

/* getarrhrate, ion_elec_chem.f90:260-274 */
static double orc_arrh(const double *k, double T)
{
    double A = k[0], alpha = k[1], Ea = k[2];
    return A * pow(T, alpha) * exp(-Ea / T);
}

/* getcustomrate of O2_chem.f90:413-448 (kind 1): rateconst = c0; expterm =
 * c1/T + c2/(T**2) + c3/(T**3) + c4/(T**4) left to right (the reference
 * writes the signs into the expression: a - b is a + (-b) exactly);
 * rateconst*exp(expterm). Of H2_chem.f90:423-437 (kind 2):
 * alpha*abs(efield)*exp(-E0/abs(efield)). */
/* getcustomrate: kind 1 the fits of O2_chem.f90:413-448,
 * q0 exp(q1/T + q2/T**2 + q3/T**3 + q4/T**4) (the fit's negative terms as
 * negative coefficients); kind 2 H2_chem.f90:423-437,
 * alpha |E| exp(-E0/|E|) with alpha = q0, E0 = q1 */
static double orc_customrate(const double *q, int kind, double T, double efield)
{
    if (kind == 1) {
        /* the powers as gfortran -O3 forms them from the fit's T**n: a chain
         * of products, T**4 = ((T*T)*T)*T (the reference build's rates match
         * this bit for bit and (T*T)*(T*T) in about two cases of three) */
        double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
        double expterm = q[1] / T + q[2] / T2 + q[3] / T3 + q[4] / T4;
        return q[0] * exp(expterm);
    }
    return q[0] * fabs(efield) * exp(-q[1] / fabs(efield));
}

/* the rate of reaction r as getspecproduction chooses it (isratearrh) */
static double orc_rate(const orc_chem *ch, int r, double Te, double efield)
{
    int kind = ch->rate_kind ? ch->rate_kind[r] : 0;
    if (kind == 0) return orc_arrh(ch->k_arrh + 3 * r, Te);
    return orc_customrate(ch->rate_coef + 5 * r, kind, Te, efield);
}

/* efield_mag of the callers (edeneq.f90:405-409): sqrt(Ex**2 + Ey**2 + Ez**2) */
static double orc_efield_mag(const double *efield, size_t ng, size_t c)
{
    double m = efield[c] * efield[c];
    m = m + efield[ng + c] * efield[ng + c];
    m = m + efield[2 * ng + c] * efield[2 * ng + c];
    return sqrt(m);
}

/* specmult of reaction r: prod_j specden(j)**reactants(r,j) */
static double orc_specmult(const orc_chem *ch, int r, const double *specden)
{
    double m = 1.0;
    for (int j = 0; j < ch->nspecies; ++j) m = m * orc_powi(specden[j], ch->reactants[r + ch->nreac * j]);
    return m;
}

/* getspecproduction, ion_elec_chem.f90:111-156, O2_chem.f90:246-291
 * (specnum 1-based) */
static double orc_specproduction(const orc_chem *ch, int specnum, double Te, const double *specden, double efield)
{
    double specprod = 0.0;
    for (int r = 0; r < ch->nreac; ++r) {
        int s = specnum - 1;
        int nmoles = ch->products[r + ch->nreac * s] - ch->reactants[r + ch->nreac * s];
        if (nmoles != 0) {
            double k = orc_rate(ch, r, Te, efield);
            double specmult = orc_specmult(ch, r, specden);
            specprod = specprod + nmoles * k * specmult;
        }
    }
    return specprod;
}

/* getelectroninelasticterm, ion_elec_chem.f90:158-199, O2_chem.f90:293-334 */
static double orc_inelterm(const orc_chem *ch, double Te, const double *specden, double efield)
{
    double inelterm = 0.0;
    for (int r = 0; r < ch->nreac; ++r) {
        if (ch->elecenergy[r] != 0) {
            double k = orc_rate(ch, r, Te, efield);
            double specmult = orc_specmult(ch, r, specden);
            inelterm = inelterm + k * specmult * ch->elecenergy[r];
        }
    }
    return inelterm;
}

/* getspecmobility (H2_O2_chem.f90:600-672; ion_elec_chem.f90:222-242):
 * kind 1 (electrons) mobility / sum_j specarray(j)/exp(a L^2 + b L + c),
 * L = log(Te/eVtoK); kind 2 (ions) mobility*Natm / sum_j specarray(j)/K_j,
 * Natm = Patm/k_BMAN/Troom; kind 0 the constant. Specnum 1-based. */
static double orc_spec_mobility(const orc_chem *ch, int specnum, const double *sa, double Te)
{
    int s = specnum - 1, ns = ch->nspecies;
    int kind = ch->mob_kind ? ch->mob_kind[s] : 0;
    if (kind == 0) return ch->mobility[s];
    double Te_in_eV = Te / 11604.5;
    double log_Te = log(Te_in_eV);
    double Natm = 101325.0 / ORC_KBMAN / 300.0;
    double mob_inv = 0.0;
    int first = 1;
    for (int j = 0; j < ns; ++j) {
        if (!ch->mob_terms[j + ns * s]) continue;
        const double *co = ch->mob_coef + 3 * (j + ns * s);
        double inv = kind == 1 ? sa[j] / exp(co[0] * (log_Te * log_Te) + co[1] * log_Te + co[2]) : sa[j] / co[0];
        mob_inv = first ? inv : mob_inv + inv;
        first = 0;
    }
    return kind == 2 ? ch->mobility[s] * Natm / mob_inv : ch->mobility[s] / mob_inv;
}

/* getspecdcoeff: kind 1 k_BMAN*Te*mob/(Z*ECHARGE), kind 2 the same at Tg,
 * kind 0 the constant */
static double orc_spec_dcoeff(const orc_chem *ch, int specnum, const double *sa, double Te, double Tg)
{
    int s = specnum - 1;
    int kind = ch->dcoeff_kind ? ch->dcoeff_kind[s] : 0;
    if (kind == 0) return ch->dcoeff[s];
    double mob = orc_spec_mobility(ch, specnum, sa, Te);
    double T = kind == 1 ? Te : Tg;
    return ORC_KBMAN * T * mob / (ch->charge[s] * ORC_ECHARGE);
}

static void orc_specarray(const orc_chem *ch, const double *numden, size_t ng, size_t c, double *sa)
{
    for (int s = 0; s < ch->nspecies; ++s) sa[s] = numden[(size_t)s * ng + c];
}

/* the scaling and time-term tail shared by the *_update_transport routines */
static void orc_transport_tail(size_t ng, int timederiv, double dt, const double *scaling, const double *soln,
                               double *vel, double *dcoeff, double *reac, double *source)
{
    for (size_t c = 0; c < 3 * ng; ++c) vel[c] = vel[c] * scaling[0];
    for (size_t c = 0; c < ng; ++c) {
        dcoeff[c] = dcoeff[c] * scaling[1];
        reac[c] = reac[c] * scaling[2];
        source[c] = source[c] * scaling[3];
    }
    if (timederiv)
        for (size_t c = 0; c < ng; ++c) source[c] = source[c] + soln[c] / dt;
}

void orc_species_transport(const orc_problem *p, const orc_chem *ch, int specnum, double dt, int timederiv,
                           const double *efield, const double *electemp, const double *numden, double Tg,
                           const double *scaling, const double *soln, orc_transport_out out)
{
    lev_t L = finest(p);
    size_t ng = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    for (size_t c = 0; c < 3 * ng; ++c) out.vel[c] = 0.0;
    for (size_t c = 0; c < ng; ++c) out.dcoeff[c] = out.reac[c] = out.source[c] = 0.0;
    double sa[64];
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i) {
                size_t c = IDX(&L, i, j, k);
                orc_specarray(ch, numden, ng, c, sa);
                double mu = orc_spec_mobility(ch, specnum, sa, electemp[c]);
                for (int q = 0; q < 3; ++q) out.vel[q * ng + c] = mu * efield[q * ng + c];
                out.dcoeff[c] = orc_spec_dcoeff(ch, specnum, sa, electemp[c], Tg);
                double specprod = orc_specproduction(ch, specnum, electemp[c], sa, orc_efield_mag(efield, ng, c));
                out.source[c] = out.source[c] + specprod;
                if (out.prod) out.prod[c] = out.source[c];
            }
    orc_transport_tail(ng, timederiv, dt, scaling, soln, out.vel, out.dcoeff, out.reac, out.source);
}

/* findjheating_eachdir, eenergyeq.f90:461-555, for cell c along stride st */
static double orc_jheat_dir(const double *ne, const double *ef, size_t c, size_t st, double mu_e, double D_e,
                            double dx)
{
    double grad_ne = (ne[c] - ne[c - st]) / dx;
    double elec_field_left = (ef[c] + ef[c - st]) * 0.5;
    double e_den = (mu_e * elec_field_left > 0) ? ne[c - st] : ne[c];
    double Gamma_left = (mu_e * e_den * elec_field_left - D_e * grad_ne);
    grad_ne = (ne[c + st] - ne[c]) / dx;
    double elec_field_right = (ef[c] + ef[c + st]) * 0.5;
    e_den = (mu_e * elec_field_right > 0) ? ne[c] : ne[c + st];
    double Gamma_right = (mu_e * e_den * elec_field_right - D_e * grad_ne);
    double jheat = -(ORC_ECHARGE * 0.5 * (elec_field_left + elec_field_right) * 0.5 * (Gamma_left + Gamma_right));
    if (jheat < 0.0) jheat = 0.0;
    return jheat;
}

void orc_eenergy_transport(const orc_problem *p, const orc_chem *ch, double dt, int timederiv, double lscale,
                           const double *efield, const double *electemp, const double *numden, double Tg,
                           const double *scaling, const double *soln, orc_transport_out out)
{
    lev_t L = finest(p);
    size_t ng = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    size_t n = (size_t)L.nx * L.ny * L.nz;
    const size_t st[3] = {1, (size_t)(L.nx + 2), (size_t)(L.nx + 2) * (L.ny + 2)};
    const double dxl[3] = {L.dx * lscale, L.dy * lscale, L.dz * lscale};
    int e = ch->especnum - 1;
    const double *elecden = numden + (size_t)e * ng;
    double *jheat = calloc(n, sizeof(double)), *elastic = calloc(n, sizeof(double));
    double *inelastic = calloc(n, sizeof(double));
    double sa[64];
    size_t t;
    /* jouleheatingterm (:394-459) */
    t = 0;
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i, ++t) {
                size_t c = IDX(&L, i, j, k);
                orc_specarray(ch, numden, ng, c, sa);
                double mu_e = orc_spec_mobility(ch, e + 1, sa, electemp[c]);
                double D_e = orc_spec_dcoeff(ch, e + 1, sa, electemp[c], Tg);
                double jx = orc_jheat_dir(elecden, efield, c, st[0], mu_e, D_e, dxl[0]);
                double jy = orc_jheat_dir(elecden, efield + ng, c, st[1], mu_e, D_e, dxl[1]);
                double jz = orc_jheat_dir(elecden, efield + 2 * ng, c, st[2], mu_e, D_e, dxl[2]);
                jheat[t] = jx + jy + jz;
            }
    /* inelastic_colterm (:614-677) */
    t = 0;
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i, ++t) {
                size_t c = IDX(&L, i, j, k);
                orc_specarray(ch, numden, ng, c, sa);
                inelastic[t] = orc_inelterm(ch, electemp[c], sa, orc_efield_mag(efield, ng, c)) * ORC_ECHARGE;
            }
    /* elastic_colterm (:557-612), getelectroncollisionfrequency
     * (ion_elec_chem.f90:244-258) */
    double m_e = ch->molmass[e], m_bg = ch->molmass[ch->bgspecnum - 1];
    t = 0;
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i, ++t) {
                size_t c = IDX(&L, i, j, k);
                orc_specarray(ch, numden, ng, c, sa);
                double nu = -(ORC_ECHARGE / m_e / orc_spec_mobility(ch, e + 1, sa, electemp[c]));
                elastic[t] = 3.0 * ORC_KBMAN * elecden[c] * (electemp[c] - Tg) * (m_e / m_bg) * nu;
            }
    for (size_t c = 0; c < 3 * ng; ++c) out.vel[c] = 0.0;
    for (size_t c = 0; c < ng; ++c) out.dcoeff[c] = out.reac[c] = out.source[c] = 0.0;
    const double five_by_three = 5.0 / 3.0;
    t = 0;
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i, ++t) {
                size_t c = IDX(&L, i, j, k);
                orc_specarray(ch, numden, ng, c, sa);
                double mu_e = orc_spec_mobility(ch, e + 1, sa, electemp[c]);
                for (int q = 0; q < 3; ++q) out.vel[q * ng + c] = five_by_three * mu_e * efield[q * ng + c];
                out.dcoeff[c] = five_by_three * orc_spec_dcoeff(ch, e + 1, sa, electemp[c], Tg);
                out.source[c] = jheat[t] - elastic[t] + inelastic[t];
            }
    free(jheat);
    free(elastic);
    free(inelastic);
    orc_transport_tail(ng, timederiv, dt, scaling, soln, out.vel, out.dcoeff, out.reac, out.source);
}

void orc_poisson_transport(const orc_problem *p, const orc_chem *ch, double dt, int timederiv,
                           const double *elecden, const double *iondens, const double *scaling,
                           const double *soln, orc_transport_out out)
{
    lev_t L = finest(p);
    size_t ng = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    for (size_t c = 0; c < 3 * ng; ++c) out.vel[c] = 0.0;
    for (size_t c = 0; c < ng; ++c) out.dcoeff[c] = out.reac[c] = out.source[c] = 0.0;
    int nions = ch->ionspecmax - ch->ionspecmin + 1;
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i) {
                size_t c = IDX(&L, i, j, k);
                double eden = elecden[c], total_ionden = 0.0;
                for (int s = 0; s < nions; ++s)
                    total_ionden = total_ionden + ch->charge[ch->ionspecmin - 1 + s] * iondens[(size_t)s * ng + c];
                out.source[c] = (total_ionden - eden);
            }
    for (size_t c = 0; c < ng; ++c) {
        out.source[c] = out.source[c] * ORC_ECHARGE / ORC_EPS0;
        out.dcoeff[c] = 1.0 * scaling[1];
        out.source[c] = out.source[c] * scaling[3];
    }
    if (timederiv)
        for (size_t c = 0; c < ng; ++c) out.source[c] = out.source[c] + soln[c] / dt;
}

void orc_electron_temp(const orc_problem *p, double *eenergysoln, const double *elecden, double eenergyscale,
                       double etempmin, double *electemp)
{
    lev_t L = finest(p);
    const double three_by_two = 1.5;
    for (int k = 1; k <= L.nz; ++k)
        for (int j = 1; j <= L.ny; ++j)
            for (int i = 1; i <= L.nx; ++i) {
                size_t c = IDX(&L, i, j, k);
                electemp[c] = eenergysoln[c] * eenergyscale / (three_by_two * elecden[c] * ORC_KBMAN);
                if (electemp[c] < etempmin) {
                    electemp[c] = etempmin;
                    eenergysoln[c] = three_by_two * elecden[c] * ORC_KBMAN * etempmin / eenergyscale;
                }
            }
}

/* find_sg_exponential_flux, mps1d/source/convdiff.f90:10-64 */
void orc_sg_flux(double cL, double cR, double DL, double DR, double uL, double uR, double dx,
                 double *wL, double *wR, double *flux)
{
    const double eps = 1e-10;
    double wl = 0.0, wr = 0.0;
    double ch = 0.5 * (cL + cR);
    double Dh = 2.0 * DL * DR / (DL + DR + eps);
    double sgn = ch <= 0.0 ? -1.0 : 1.0;
    if (fabs(Dh) < eps) {
        wl = cL * 0.5 * (1.0 + sgn);
        wr = cR * 0.5 * (1.0 - sgn);
    } else {
        double Pe = ch * dx / Dh;
        if (fabs(Pe) < eps) {
            wl = Dh / dx;
            wr = -Dh / dx;
        } else if (Pe > 0.0) {
            wl = ch / (1.0 - exp(-Pe));
            wr = -ch * exp(-Pe) / (1.0 - exp(-Pe));
        } else {
            wl = ch * exp(Pe) / (exp(Pe) - 1.0);
            wr = -ch / (exp(Pe) - 1.0);
        }
    }
    *wL = wl;
    *wR = wr;
    *flux = wl * uL + wr * uR;
}
