/*
 * cf_oracle.c -- CPU restatement of Corrfunc's theory pair counters in FP64.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the GPU engine: it may
 * be built and called only by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py. The product library never links or calls it.
 *
 * It restates the reference algorithm (Corrfunc v2.5.3) step by step, in the
 * reference's FP64 arithmetic order, so that pair counts are bit-identical:
 *
 *   oracle_setup_bins     utils/utils.c:62-106 (+ getnumlines utils.c:715-745)
 *   get_max_min           utils/gridlink_utils.c.src:52-70
 *   get_binsize           utils/gridlink_utils.c.src:31-49
 *   gridlink              utils/gridlink_impl.c.src:65-436 (serial; cells sorted on z)
 *   gen_cell_pairs        utils/gridlink_impl.c.src:439-625, duplicate check
 *                         utils/gridlink_utils.h.src:46-72
 *   kernel_3d (DD, xi)    theory/DD/countpairs_kernels.c.src:877-999,
 *                         theory/xi/xi_kernels.c.src:851-...
 *   kernel_rppi / wp      theory/DDrppi/countpairs_rp_pi_kernels.c.src:968-1123,
 *                         theory/wp/wp_kernels.c.src:1353-1534
 *   kernel_smu            theory/DDsmu/countpairs_s_mu_kernels.c.src:1018-1190
 *   drivers               theory/DD/countpairs_impl.c.src:136-707,
 *                         theory/DDrppi/countpairs_rp_pi_impl.c.src:135-709,
 *                         theory/DDsmu/countpairs_s_mu_impl.c.src:136-717,
 *                         theory/xi/countpairs_xi_impl.c.src:135-643,
 *                         theory/wp/countpairs_wp_impl.c.src:148-676
 *
 * The pruning (z-window, per-particle minimum separation, cell-pair culling) is
 * the fallback kernels' pruning for every instruction set. The r^2 / s^2 / rp^2
 * rounding order follows `instruction_set` (reference per-ISA kernels):
 *   DD     : fallback/SSE/AVX  (dx*dx+dy*dy)+dz*dz     AVX512F fma(dz,dz,fma(dy,dy,dx*dx))
 *   xi     : fallback/SSE      (dx*dx+dy*dy)+dz*dz     AVX/AVX2 dx*dx+(dy*dy+dz*dz)
 *                                                      AVX512F fma(dz,dz,fma(dy,dy,dx*dx))
 *   rp_pi  : fallback/SSE/AVX  dx*dx+dy*dy             AVX512F fma(dy,dy,dx*dx)
 *   wp     : fallback/SSE/AVX  dx*dx+dy*dy             AVX2/AVX512F fma(dy,dy,dx*dx)
 *   s_mu   : fallback/SSE/AVX  (dx*dx+dy*dy)+dz*dz     AVX512F fma(dx,dx,fma(dy,dy,dz*dz))
 * and for AVX512F the (rp,pi) / (s,mu) flattened bin index is truncated from the
 * FP64 sum kbin*(n2+1) + frac (countpairs_rp_pi_kernels.c.src:249-256,
 * countpairs_s_mu_kernels.c.src:268-277). Build with -ffp-contract=off.
 *
 * float32 inputs are up-cast to float64 and run through the FP64 path.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include <float.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/corrfunc_theory.h"
#include "../include/corrfunc_mocks.h"

#define BOOST_CELL_THRESH 10
#define BOOST_NUMPART_THRESH 250
#define BOOST_BIN_REF 1

enum { MODE_DD = 0, MODE_RPPI = 1, MODE_SMU = 2, MODE_XI = 3, MODE_WP = 4, MODE_RPPI_MOCKS = 5, MODE_SMU_MOCKS = 6 };
/* rounding-order ids */
enum { R2_SUM_XY_Z = 0, R2_FMA_DZ_DY_DX = 1, R2_X_SUM_YZ = 2, R2_FMA_DX_DY_DZ = 3,
       RP_SUM = 10, RP_FMA = 11 };

/* Statistics of the last oracle call (cell grid and canonical work). */
typedef struct {
    int nmesh[3];
    int refine[3];
    int64_t num_cell_pairs;
    double tested_pairs;  /* sum over cell pairs of N1*N2, N(N-1)/2 for same_cell */
} oracle_stats;
oracle_stats oracle_last_stats;

/* ------------------------------------------------------------------ bins */
static int64_t count_bin_lines(const char *fname)
{
    FILE *fp = fopen(fname, "rt");
    if (!fp) return -1;
    char line[10000];
    int64_t n = 0;
    while (fgets(line, sizeof line, fp)) {
        char *c = line;
        while (*c != '\0' && isspace((unsigned char)*c)) c++;
        if (*c != '\0' && *c != '#') n++;
    }
    fclose(fp);
    return n;
}

/* utils/utils.c:62-106: nbin = lines+1 edges; rupp[0] = first low edge,
 * rupp[k] = high edge of line k; rupp[nbin-1] = rupp[nbin] = rmax. */
int oracle_setup_bins(const char *fname, double *rmin, double *rmax, int *nbin, double **rupp)
{
    int64_t nl = count_bin_lines(fname);
    if (nl < 0) { fprintf(stderr, "oracle: could not open bin file %s\n", fname); return EXIT_FAILURE; }
    *nbin = (int)nl + 1;
    *rupp = calloc((size_t)*nbin + 1, sizeof(double));
    FILE *fp = fopen(fname, "r");
    if (!fp) { free(*rupp); return EXIT_FAILURE; }
    char buf[1000];
    int index = 1;
    double lo, hi;
    while (fgets(buf, sizeof buf, fp)) {
        if (sscanf(buf, "%lf %lf", &lo, &hi) == 2) {
            if (index == 1) { *rmin = lo; (*rupp)[0] = lo; }
            (*rupp)[index] = hi;
            index++;
        }
    }
    *rmax = (*rupp)[index - 1];
    fclose(fp);
    (*rupp)[*nbin] = *rmax;
    (*rupp)[*nbin - 1] = *rmax;
    return EXIT_SUCCESS;
}

/* ------------------------------------------------------------- gridding */
typedef struct {
    int64_t n;
    double *x, *y, *z, *w;     /* views into the gridded SoA arrays */
    double xb[2], yb[2], zb[2];
} ocell;

typedef struct {
    int64_t c1, c2;
    double xwrap, ywrap, zwrap;
    double min_dx, min_dy, min_dz;
    double cx, cy, cz;
    int same;
} ocellpair;

typedef struct {
    ocell *cells;
    double *x, *y, *z, *w;
    int nx, ny, nz;
} olattice;

static void free_lattice(olattice *L)
{
    if (!L) return;
    free(L->cells); free(L->x); free(L->y); free(L->z); free(L->w);
    L->cells = NULL; L->x = L->y = L->z = L->w = NULL;
}

static void get_max_min(int64_t n, const double *X, const double *Y, const double *Z,
                        double *xmin, double *ymin, double *zmin, double *xmax, double *ymax, double *zmax)
{
    for (int64_t i = 0; i < n; i++) {
        if (X[i] < *xmin) *xmin = X[i];
        if (Y[i] < *ymin) *ymin = Y[i];
        if (Z[i] < *zmin) *zmin = Z[i];
        if (X[i] > *xmax) *xmax = X[i];
        if (Y[i] > *ymax) *ymax = Y[i];
        if (Z[i] > *zmax) *zmax = Z[i];
    }
}

static int get_binsize(double xdiff, double xwrap, double rmax, int refine, int max_ncells,
                       double *binsize, int *nlattice)
{
    int nmesh = (int)(refine * xdiff / rmax);
    nmesh = nmesh < 1 ? 1 : nmesh;
    if (xwrap > 0 && rmax >= xwrap / 2) {
        fprintf(stderr, "oracle> ERROR: rmax=%f must be less than half of periodic boxize=%f to avoid double-counting particles\n", rmax, xwrap);
        return EXIT_FAILURE;
    }
    if (nmesh > max_ncells) nmesh = max_ncells;
    if (nmesh < 2) nmesh = 2;
    *binsize = xdiff / nmesh;
    *nlattice = nmesh;
    return EXIT_SUCCESS;
}

static const double *g_sort_z;
static int cmp_z_then_index(const void *a, const void *b)
{
    int64_t ia = *(const int64_t *)a, ib = *(const int64_t *)b;
    double za = g_sort_z[ia], zb = g_sort_z[ib];
    if (za < zb) return -1;
    if (za > zb) return 1;
    return ia < ib ? -1 : (ia > ib);
}

static int gridlink(int64_t np, const double *X, const double *Y, const double *Z, const double *W,
                    double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
                    double max_x, double max_y, double max_z,
                    double xwrap, double ywrap, double zwrap,
                    int rx, int ry, int rz, int max_cells, olattice *L)
{
    int nx, ny, nz;
    double xbs, ybs, zbs;
    if (get_binsize(xmax - xmin, xwrap, max_x, rx, max_cells, &xbs, &nx) != EXIT_SUCCESS ||
        get_binsize(ymax - ymin, ywrap, max_y, ry, max_cells, &ybs, &ny) != EXIT_SUCCESS ||
        get_binsize(zmax - zmin, zwrap, max_z, rz, max_cells, &zbs, &nz) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    const int64_t ncells = (int64_t)nx * ny * nz;
    const double xinv = xbs > 0 ? 1.0 / xbs : 0.;
    const double yinv = ybs > 0 ? 1.0 / ybs : 0.;
    const double zinv = zbs > 0 ? 1.0 / zbs : 0.;
    int64_t *cidx = malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    int64_t *count = calloc((size_t)ncells + 1, sizeof(int64_t));
    int64_t oob = 0;
    for (int64_t i = 0; i < np; i++) {
        int ix = (int)((X[i] - xmin) * xinv);
        int iy = (int)((Y[i] - ymin) * yinv);
        int iz = (int)((Z[i] - zmin) * zinv);
        if (ix > nx - 1) ix--;
        if (iy > ny - 1) iy--;
        if (iz > nz - 1) iz--;
        oob += ix < 0 || ix >= nx || iy < 0 || iy >= ny || iz < 0 || iz >= nz;
        cidx[i] = (int64_t)ix * ny * nz + (int64_t)iy * nz + iz;
    }
    if (oob) {
        fprintf(stderr, "oracle> %lld particles are out of bounds. Check periodic wrapping?\n", (long long)oob);
        free(cidx); free(count);
        return EXIT_FAILURE;
    }
    for (int64_t i = 0; i < np; i++) count[cidx[i] + 1]++;
    for (int64_t c = 0; c < ncells; c++) count[c + 1] += count[c];
    int64_t *order = malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    int64_t *fill = calloc((size_t)ncells, sizeof(int64_t));
    for (int64_t i = 0; i < np; i++) order[count[cidx[i]] + fill[cidx[i]]++] = i;
    /* sort each cell on z (ties kept in input order; the counts do not depend on it) */
    g_sort_z = Z;
    for (int64_t c = 0; c < ncells; c++) {
        int64_t n = count[c + 1] - count[c];
        if (n > 1) qsort(order + count[c], (size_t)n, sizeof(int64_t), cmp_z_then_index);
    }
    L->nx = nx; L->ny = ny; L->nz = nz;
    L->cells = calloc((size_t)ncells, sizeof(ocell));
    size_t nb = sizeof(double) * (size_t)(np > 0 ? np : 1);
    L->x = malloc(nb); L->y = malloc(nb); L->z = malloc(nb);
    L->w = W ? malloc(nb) : NULL;
    for (int64_t k = 0; k < np; k++) {
        int64_t i = order[k];
        L->x[k] = X[i]; L->y[k] = Y[i]; L->z[k] = Z[i];
        if (W) L->w[k] = W[i];
    }
    for (int64_t c = 0; c < ncells; c++) {
        ocell *cell = &L->cells[c];
        int64_t s = count[c];
        cell->n = count[c + 1] - s;
        cell->x = L->x + s; cell->y = L->y + s; cell->z = L->z + s;
        cell->w = W ? L->w + s : NULL;
        cell->xb[0] = cell->yb[0] = cell->zb[0] = DBL_MAX;
        cell->xb[1] = cell->yb[1] = cell->zb[1] = -DBL_MAX;
        for (int64_t k = 0; k < cell->n; k++) {
            double xv = cell->x[k], yv = cell->y[k], zv = cell->z[k];
            cell->xb[0] = xv < cell->xb[0] ? xv : cell->xb[0];
            cell->yb[0] = yv < cell->yb[0] ? yv : cell->yb[0];
            cell->zb[0] = zv < cell->zb[0] ? zv : cell->zb[0];
            cell->xb[1] = xv > cell->xb[1] ? xv : cell->xb[1];
            cell->yb[1] = yv > cell->yb[1] ? yv : cell->yb[1];
            cell->zb[1] = zv > cell->zb[1] ? zv : cell->zb[1];
        }
    }
    free(cidx); free(count); free(order); free(fill);
    return EXIT_SUCCESS;
}

/* utils/gridlink_impl.c.src:439-625 */
static ocellpair *gen_cell_pairs(const olattice *L1, const olattice *L2, int64_t *ncp,
                                 int rx, int ry, int rz,
                                 double xwrap, double ywrap, double zwrap,
                                 double max3, double max2, double max1,
                                 int min_sep_opt, int autocorr, int px, int py, int pz)
{
    const int nx = L1->nx, ny = L1->ny, nz = L1->nz;
    const int64_t ncells = (int64_t)nx * ny * nz;
    const int64_t maxngb = (int64_t)(2 * rx + 1) * (2 * ry + 1) * (2 * rz + 1);
    ocellpair *cp = malloc(sizeof(ocellpair) * (size_t)(ncells * maxngb > 0 ? ncells * maxngb : 1));
    int64_t n = 0;
    const int anyp = px || py || pz;
    const int dupcheck = anyp && (nx < 2 * rx + 1 || ny < 2 * ry + 1 || nz < 2 * rz + 1);
    for (int64_t icell = 0; icell < ncells; icell++) {
        const ocell *first = &L1->cells[icell];
        if (first->n == 0) continue;
        const int iz = (int)(icell % nz);
        const int ix = (int)(icell / ((int64_t)ny * nz));
        const int iy = (int)((icell - iz - (int64_t)ix * nz * ny) / nz);
        int64_t nthis = 0;
        for (int iix = -rx; iix <= rx; iix++) {
            const int iiix = px ? (ix + iix + nx) % nx : ix + iix;
            if (iiix < 0 || iiix >= nx) continue;
            const double offx = (ix + iix >= 0 && ix + iix < nx) ? 0.0 : (ix + iix < 0 ? xwrap : -xwrap);
            for (int iiy = -ry; iiy <= ry; iiy++) {
                const int iiiy = py ? (iy + iiy + ny) % ny : iy + iiy;
                if (iiiy < 0 || iiiy >= ny) continue;
                const double offy = (iy + iiy >= 0 && iy + iiy < ny) ? 0.0 : (iy + iiy < 0 ? ywrap : -ywrap);
                for (int iiz = -rz; iiz <= rz; iiz++) {
                    const int iiiz = pz ? (iz + iiz + nz) % nz : iz + iiz;
                    if (iiiz < 0 || iiiz >= nz) continue;
                    const double offz = (iz + iiz >= 0 && iz + iiz < nz) ? 0.0 : (iz + iiz < 0 ? zwrap : -zwrap);
                    const int64_t icell2 = iiiz + (int64_t)nz * iiiy + (int64_t)nz * ny * iiix;
                    if ((autocorr == 1 && icell2 > icell) || L2->cells[icell2].n == 0) continue;
                    if (dupcheck) {
                        int dup = 0;
                        for (int64_t jj = 0; jj < nthis; jj++) {
                            const ocellpair *p = &cp[n - jj - 1];
                            if (p->c2 == icell2 && p->xwrap == offx && p->ywrap == offy && p->zwrap == offz) { dup = 1; break; }
                        }
                        if (dup) continue;
                    }
                    const ocell *second = &L2->cells[icell2];
                    double cx = 0, cy = 0, cz = 0, mdx = 0, mdy = 0, mdz = 0;
                    if (min_sep_opt) {
                        const double xl = first->xb[0] + offx, xh = first->xb[1] + offx;
                        const double yl = first->yb[0] + offy, yh = first->yb[1] + offy;
                        const double zl = first->zb[0] + offz, zh = first->zb[1] + offz;
                        cx = iix < 0 ? xl : (iix > 0 ? xh : 0.0);
                        cy = iiy < 0 ? yl : (iiy > 0 ? yh : 0.0);
                        cz = iiz < 0 ? zl : (iiz > 0 ? zh : 0.0);
                        mdx = iix != 0 ? ((iix < 0 ? xl : xh) - (iix < 0 ? second->xb[1] : second->xb[0])) : 0.0;
                        mdy = iiy != 0 ? ((iiy < 0 ? yl : yh) - (iiy < 0 ? second->yb[1] : second->yb[0])) : 0.0;
                        mdz = iiz != 0 ? ((iiz < 0 ? zl : zh) - (iiz < 0 ? second->zb[1] : second->zb[0])) : 0.0;
                        if (max3 > 0) {
                            const double s = mdx * mdx + mdy * mdy + mdz * mdz;
                            if (s >= max3 * max3) continue;
                        }
                        if (max2 > 0) {
                            const double s = mdx * mdx + mdy * mdy;
                            if (s >= max2 * max2) continue;
                        }
                        if (max1 > 0 && iiz != 0) {
                            if (mdz * mdz >= max1 * max1) continue;
                        }
                    }
                    ocellpair *p = &cp[n];
                    p->c1 = icell; p->c2 = icell2;
                    p->xwrap = offx; p->ywrap = offy; p->zwrap = offz;
                    p->min_dx = mdx; p->min_dy = mdy; p->min_dz = mdz;
                    p->cx = cx; p->cy = cy; p->cz = cz;
                    p->same = (autocorr == 1 && icell2 == icell) ? 1 : 0;
                    n++;
                    nthis++;
                }
            }
        }
    }
    *ncp = n;
    return cp;
}

/* ---------------------------------------------------------------- kernels */
typedef struct {
    int mode;
    int r2form;         /* rounding order id */
    int float_index;    /* AVX512F flattened-bin rule for rp_pi / s_mu */
    int nbin;           /* number of edges */
    const double *rupp_sqr;
    double sqr_rmax, sqr_rmin;
    double pimax;       /* DD/xi: rmax; rp_pi/wp: pimax; s_mu: smax*mu_max */
    int n2;             /* npibin or nmu_bins */
    double inv_d2;      /* 1/dpi or 1/dmu */
    double sqr_mu_max;
    double sqr_rmax_bins;   /* on-sky rp-pi: rpmax^2 (sqr_rmax holds the 3-D limit) */
    int need_avg, need_w;
} okparams;

typedef struct { uint64_t *np; double *avg; double *wavg; } oacc;

static inline double r2_3d(int f, double dx, double dy, double dz)
{
    switch (f) {
    case R2_FMA_DZ_DY_DX: return fma(dz, dz, fma(dy, dy, dx * dx));
    case R2_X_SUM_YZ:     return dx * dx + (dy * dy + dz * dz);
    case R2_FMA_DX_DY_DZ: return fma(dx, dx, fma(dy, dy, dz * dz));
    default:              return (dx * dx + dy * dy) + dz * dz;
    }
}
static inline double rp2(int f, double dx, double dy)
{
    return f == RP_FMA ? fma(dy, dy, dx * dx) : dx * dx + dy * dy;
}

/* bin search shared by all kernels: first k from the top with r2 >= rupp_sqr[k-1] */
static inline int find_kbin(const okparams *P, double r2)
{
    for (int k = P->nbin - 1; k >= 1; k--)
        if (r2 >= P->rupp_sqr[k - 1]) return k;
    return 0;
}

static void kernel_cellpair(const okparams *P, const ocell *c0, const ocell *c1, const ocellpair *cp, oacc *A)
{
    const int64_t N0 = c0->n, N1 = c1->n;
    if (N0 == 0 || N1 == 0) return;
    const double *x0 = c0->x, *y0 = c0->y, *z0 = c0->z, *w0 = c0->w;
    const double *x1 = c1->x, *y1 = c1->y, *z1s = c1->z, *w1 = c1->w;
    const double offx = cp->xwrap, offy = cp->ywrap, offz = cp->zwrap;
    const double mxd = cp->min_dx, myd = cp->min_dy, mzd = cp->min_dz;
    const int mode = P->mode;
    const int is2d = (mode == MODE_RPPI || mode == MODE_WP);
    const double pimax = P->pimax;
    double max_all_dz;
    if (is2d) max_all_dz = pimax;
    else if (mode == MODE_SMU) {
        max_all_dz = sqrt(P->sqr_rmax - mxd * mxd - myd * myd);
        max_all_dz = pimax < max_all_dz ? pimax : max_all_dz;
    } else {
        /* DD: SQRT(rpmax*rpmax - ...) (rpmax*rpmax == sqr_rpmax); xi: SQRT(sqr_rmax - ...) */
        max_all_dz = sqrt(P->sqr_rmax - mxd * mxd - myd * myd);
    }
    int64_t z1 = 0;  /* running index into the secondary cell */
    for (int64_t i = 0; i < N0; i++) {
        const double xpos = x0[i] + offx, ypos = y0[i] + offy, zpos = z0[i] + offz;
        double max_dz = max_all_dz;
        const double this_dz = z1s[z1] - zpos;
        if (this_dz >= max_all_dz) continue;
        if (cp->same == 1) {
            z1++;
        } else {
            const double mdx = mxd > 0 ? mxd + fabs(xpos - cp->cx) : mxd;
            const double mdy = myd > 0 ? myd + fabs(ypos - cp->cy) : myd;
            const double mdz = mzd > 0 ? (this_dz > 0 ? this_dz : mzd + fabs(zpos - cp->cz)) : mzd;
            if (is2d) {
                const double s = mdx * mdx + mdy * mdy;
                if (s >= P->sqr_rmax || mdz >= pimax) continue;
                const double target = zpos - pimax;
                while (z1 < N1 && z1s[z1] <= target) z1++;
            } else if (mode == MODE_SMU) {
                const double s = mdx * mdx + mdy * mdy + mdz * mdz;
                if (s >= P->sqr_rmax || mdz >= pimax) continue;
                const double sqr_max_dz = P->sqr_rmax - mdx * mdx - mdy * mdy;
                max_dz = sqr_max_dz < pimax * pimax ? sqrt(sqr_max_dz) : pimax;
                const double target = zpos - max_all_dz;
                while (z1 < N1 && z1s[z1] <= target) z1++;
            } else {
                const double s = mdx * mdx + mdy * mdy + mdz * mdz;
                if (s >= P->sqr_rmax) continue;
                max_dz = sqrt(P->sqr_rmax - mdx * mdx - mdy * mdy);
                const double target = zpos - max_all_dz;
                while (z1 < N1 && z1s[z1] <= target) z1++;
            }
        }
        if (z1 == N1) break;
        int64_t j0 = z1;
        if (!is2d) {
            const double target = zpos - max_dz;
            while (j0 != N1 && z1s[j0] <= target) j0++;
        }
        const double wi = w0 ? w0[i] : 0.0;
        for (int64_t j = j0; j < N1; j++) {
            const double dx = x1[j] - xpos;
            const double dy = y1[j] - ypos;
            double dz = z1s[j] - zpos;
            if (is2d) {
                if (dz <= -pimax) continue;
                if (dz >= pimax) break;
                dz = fabs(dz);
                const double r2 = rp2(P->r2form, dx, dy);
                if (r2 >= P->sqr_rmax || r2 < P->sqr_rmin) continue;
                const int k = find_kbin(P, r2);
                if (k == 0) continue;
                int ibin;
                if (mode == MODE_WP) {
                    ibin = k;
                } else if (P->float_index) {
                    ibin = (int)((double)k * (double)(P->n2 + 1) + dz * P->inv_d2);
                } else {
                    int pb = (int)(dz * P->inv_d2);
                    pb = pb > P->n2 ? P->n2 : pb;
                    ibin = k * (P->n2 + 1) + pb;
                }
                A->np[ibin]++;
                if (P->need_avg) A->avg[ibin] += sqrt(r2);
                if (P->need_w) A->wavg[ibin] += wi * w1[j];
            } else if (mode == MODE_SMU) {
                if (dz >= max_dz) break;
                double s2, sqr_dz = dz * dz;
                if (P->r2form == R2_FMA_DX_DY_DZ) s2 = fma(dx, dx, fma(dy, dy, sqr_dz));
                else s2 = (dx * dx + dy * dy) + sqr_dz;
                if (s2 >= P->sqr_rmax || s2 < P->sqr_rmin) continue;
                if (sqr_dz >= s2 * P->sqr_mu_max) continue;
                const double mu = sqrt(sqr_dz / s2);
                const int k = find_kbin(P, s2);
                if (k == 0) continue;
                int ibin;
                if (P->float_index) {
                    ibin = (int)((double)k * (double)(P->n2 + 1) + mu * P->inv_d2);
                } else {
                    int mb = (int)(mu * P->inv_d2);
                    mb = mb > P->n2 ? P->n2 : mb;
                    ibin = k * (P->n2 + 1) + mb;
                }
                A->np[ibin]++;
                if (P->need_avg) A->avg[ibin] += sqrt(s2);
                if (P->need_w) A->wavg[ibin] += wi * w1[j];
            } else {
                if (dz >= max_dz) break;
                const double r2 = r2_3d(P->r2form, dx, dy, dz);
                if (r2 >= P->sqr_rmax || r2 < P->sqr_rmin) continue;
                const int k = find_kbin(P, r2);
                if (k == 0) continue;
                A->np[k]++;
                if (P->need_avg) A->avg[k] += sqrt(r2);
                if (P->need_w) A->wavg[k] += wi * w1[j];
            }
        }
    }
}

/* On-sky pairs (mocks/DDrppi_mocks/countpairs_rp_pi_mocks_kernels.c.src:1124-1297,
 * mocks/DDsmu_mocks/countpairs_s_mu_mocks_kernels.c.src fallback): the pruning of
 * the 3-D kernel with the maximum 3-D separation (P->sqr_rmax holds
 * rpmax^2 + pimax^2 for rp-pi, smax^2 for s-mu), then the line-of-sight split
 * of s along l = x_i + x_j. P->pimax holds pimax (rp-pi) or mu_max (s-mu). */
static void kernel_cellpair_mocks(const okparams *P, const ocell *c0, const ocell *c1, const ocellpair *cp, oacc *A)
{
    const int64_t N0 = c0->n, N1 = c1->n;
    if (N0 == 0 || N1 == 0) return;
    const double *x0 = c0->x, *y0 = c0->y, *z0 = c0->z, *w0 = c0->w;
    const double *x1 = c1->x, *y1 = c1->y, *z1s = c1->z, *w1 = c1->w;
    const double mxd = cp->min_dx, myd = cp->min_dy, mzd = cp->min_dz;
    const int rppi = P->mode == MODE_RPPI_MOCKS;
    const double sqr_max_sep = P->sqr_rmax;
    const double sqr_lim = P->pimax * P->pimax;   /* pimax^2 or mu_max^2 */
    const double max_all_dz = sqrt(sqr_max_sep - mxd * mxd - myd * myd);
    int64_t z1 = 0;
    for (int64_t i = 0; i < N0; i++) {
        const double xpos = x0[i], ypos = y0[i], zpos = z0[i];
        double max_dz = max_all_dz;
        const double this_dz = z1s[z1] - zpos;
        if (this_dz >= max_all_dz) continue;
        if (cp->same == 1) {
            z1++;
        } else {
            const double mdx = mxd > 0 ? mxd + fabs(xpos - cp->cx) : mxd;
            const double mdy = myd > 0 ? myd + fabs(ypos - cp->cy) : myd;
            const double mdz = mzd > 0 ? (this_dz > 0 ? this_dz : mzd + fabs(zpos - cp->cz)) : mzd;
            if (mdx * mdx + mdy * mdy + mdz * mdz >= sqr_max_sep) continue;
            max_dz = sqrt(sqr_max_sep - mdx * mdx - mdy * mdy);
            const double target = zpos - max_all_dz;
            while (z1 != N1 && z1s[z1] <= target) z1++;
        }
        if (z1 == N1) break;
        int64_t j0 = z1;
        const double target = zpos - max_dz;
        while (j0 != N1 && z1s[j0] <= target) j0++;
        const double wi = w0 ? w0[i] : 0.0;
        for (int64_t j = j0; j < N1; j++) {
            const double parx = xpos + x1[j], pary = ypos + y1[j], parz = zpos + z1s[j];
            const double perpx = x1[j] - xpos, perpy = y1[j] - ypos, perpz = z1s[j] - zpos;
            if (perpz >= max_dz) break;
            const double sqr_s = perpx * perpx + perpy * perpy + perpz * perpz;
            const double s_dot_l = parx * perpx + pary * perpy + parz * perpz;
            const double sqr_l = parx * parx + pary * pary + parz * parz;
            const double sqr_s_dot_l = s_dot_l * s_dot_l;
            int k, ibin;
            double sep;
            if (rppi) {
                if (sqr_s >= sqr_max_sep) continue;
                if (sqr_s_dot_l >= sqr_l * sqr_lim) continue;
                const double sqr_par = sqr_s_dot_l / sqr_l;
                const int pb = sqr_par >= sqr_lim ? P->n2 : (int)(sqrt(sqr_par) * P->inv_d2);
                const double sqr_perp = sqr_s - sqr_par;
                if (sqr_perp >= P->sqr_rmax_bins || sqr_perp < P->sqr_rmin) continue;
                k = find_kbin(P, sqr_perp);
                if (k == 0) continue;
                ibin = k * (P->n2 + 1) + pb;
                sep = P->need_avg ? sqrt(sqr_perp) : 0.0;
            } else {
                if (sqr_s >= sqr_max_sep || sqr_s < P->sqr_rmin) continue;
                const double sqr_mu = sqr_s_dot_l / (sqr_l * sqr_s);
                const int mb = sqr_mu >= sqr_lim ? P->n2 : (int)(sqrt(sqr_mu) * P->inv_d2);
                k = find_kbin(P, sqr_s);
                if (k == 0) continue;
                ibin = k * (P->n2 + 1) + mb;
                sep = P->need_avg ? sqrt(sqr_s) : 0.0;
            }
            A->np[ibin]++;
            if (P->need_avg) A->avg[ibin] += sep;
            if (P->need_w) A->wavg[ibin] += wi * w1[j];
        }
    }
}

/* ------------------------------------------------------------ helpers */
static int resolve_isa(int isa_in)
{
    if (isa_in >= 0) return isa_in;
#if defined(__x86_64__)
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f")) return AVX512F;
    if (__builtin_cpu_supports("avx2")) return AVX2;
    if (__builtin_cpu_supports("avx")) return AVX;
    if (__builtin_cpu_supports("sse4.2")) return SSE42;
#endif
    return FALLBACK;
}

static int r2form_for(int mode, int isa_id)
{
    switch (mode) {
    case MODE_DD:   return isa_id == AVX512F ? R2_FMA_DZ_DY_DX : R2_SUM_XY_Z;
    case MODE_XI:   return isa_id == AVX512F ? R2_FMA_DZ_DY_DX : ((isa_id == AVX || isa_id == AVX2) ? R2_X_SUM_YZ : R2_SUM_XY_Z);
    case MODE_RPPI: return isa_id == AVX512F ? RP_FMA : RP_SUM;
    case MODE_WP:   return (isa_id == AVX512F || isa_id == AVX2) ? RP_FMA : RP_SUM;
    case MODE_SMU:  return isa_id == AVX512F ? R2_FMA_DX_DY_DZ : R2_SUM_XY_Z;
    }
    return R2_SUM_XY_Z;
}

static double *upcast(const void *p, int64_t n, size_t ftype)
{
    if (p == NULL) return NULL;
    double *d = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (ftype == sizeof(float)) {
        const float *f = p;
        for (int64_t i = 0; i < n; i++) d[i] = (double)f[i];
    } else {
        memcpy(d, p, sizeof(double) * (size_t)n);
    }
    return d;
}

static int check_common(const struct config_options *o)
{
    if (!(o->float_type == sizeof(float) || o->float_type == sizeof(double))) {
        fprintf(stderr, "ERROR: oracle> Can only handle doubles or floats. Got an array of size = %zu\n", o->float_type);
        return EXIT_FAILURE;
    }
    if (strncmp(o->version, CORRFUNC_API_VERSION, sizeof(o->version) - 1) != 0 &&
        strncmp(o->version, "2.5.3", sizeof(o->version) - 1) != 0) {
        fprintf(stderr, "Error: Do not know this API version = `%s'. Expected version = `%s'\n", o->version, CORRFUNC_API_VERSION);
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

static double now_s(void)
{
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec + 1e-6 * t.tv_usec;
}

/* Core shared by the five modes: gridding, cell pairs, the cell-pair loop.
 * Returns per-slot sums (not yet doubled / averaged). */
typedef struct {
    int mode;
    int autocorr;
    int64_t N1, N2;
    const double *X1, *Y1, *Z1, *W1, *X2, *Y2, *Z2, *W2;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double xwrap, ywrap, zwrap;
    int px, py, pz;
    double gx, gy, gz;            /* max sizes handed to gridlink */
    double max3, max2, max1;      /* cell-pair culling */
    int boost_kind;               /* 0: DD-style (+1 on x,y), 1: DDsmu-style (x BOOST_BIN_REF),
                                     2: on-sky style (+1 on x,y, density of the larger set) */
    int numthreads;
    okparams kp;
    int64_t totbins;
} ocore;

static int run_core(ocore *C, struct config_options *o, uint64_t *np, double *avg, double *wavg)
{
    olattice L1 = {0}, L2 = {0};
    int rx = o->bin_refine_factors[0], ry = o->bin_refine_factors[1], rz = o->bin_refine_factors[2];
    if (gridlink(C->N1, C->X1, C->Y1, C->Z1, C->W1, C->xmin, C->xmax, C->ymin, C->ymax, C->zmin, C->zmax,
                 C->gx, C->gy, C->gz, C->xwrap, C->ywrap, C->zwrap, rx, ry, rz, o->max_cells_per_dim, &L1) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    if (C->boost_kind == 0 || C->boost_kind == 2) {
        /* on-sky engines take the larger of the two sets' densities
           (countpairs_rp_pi_mocks_impl.c.src:454-457) */
        const double n_eff = C->boost_kind == 2 && C->N2 > C->N1 ? (double)C->N2 : (double)C->N1;
        const double avg_np = n_eff / (L1.nx * L1.ny * L1.nz);
        const int max_nmesh = (int)fmax(L1.nx, fmax(L1.ny, L1.nz));
        if ((max_nmesh <= BOOST_CELL_THRESH || avg_np >= BOOST_NUMPART_THRESH) && max_nmesh < o->max_cells_per_dim) {
            if (get_bin_refine_scheme(o) == BINNING_DFL) {
                free_lattice(&L1);
                for (int i = 0; i < 2; i++) o->bin_refine_factors[i] += BOOST_BIN_REF;
                rx = o->bin_refine_factors[0]; ry = o->bin_refine_factors[1];
                if (gridlink(C->N1, C->X1, C->Y1, C->Z1, C->W1, C->xmin, C->xmax, C->ymin, C->ymax, C->zmin, C->zmax,
                             C->gx, C->gy, C->gz, C->xwrap, C->ywrap, C->zwrap, rx, ry, rz, o->max_cells_per_dim, &L1) != EXIT_SUCCESS)
                    return EXIT_FAILURE;
            }
        }
    } else {
        if (L1.nx <= BOOST_CELL_THRESH && L1.ny <= BOOST_CELL_THRESH && L1.nz <= BOOST_CELL_THRESH &&
            o->max_cells_per_dim >= BOOST_BIN_REF * BOOST_CELL_THRESH) {
            if (get_bin_refine_scheme(o) == BINNING_DFL) {
                free_lattice(&L1);
                for (int i = 0; i < 3; i++) o->bin_refine_factors[i] *= BOOST_BIN_REF;
                rx = o->bin_refine_factors[0]; ry = o->bin_refine_factors[1]; rz = o->bin_refine_factors[2];
                if (gridlink(C->N1, C->X1, C->Y1, C->Z1, C->W1, C->xmin, C->xmax, C->ymin, C->ymax, C->zmin, C->zmax,
                             C->gx, C->gy, C->gz, C->xwrap, C->ywrap, C->zwrap, rx, ry, rz, o->max_cells_per_dim, &L1) != EXIT_SUCCESS)
                    return EXIT_FAILURE;
            }
        }
    }
    const olattice *Lb = &L1;
    if (!C->autocorr) {
        if (gridlink(C->N2, C->X2, C->Y2, C->Z2, C->W2, C->xmin, C->xmax, C->ymin, C->ymax, C->zmin, C->zmax,
                     C->gx, C->gy, C->gz, C->xwrap, C->ywrap, C->zwrap, rx, ry, rz, o->max_cells_per_dim, &L2) != EXIT_SUCCESS) {
            free_lattice(&L1);
            return EXIT_FAILURE;
        }
        if (!(L1.nx == L2.nx && L1.ny == L2.ny && L1.nz == L2.nz)) {
            fprintf(stderr, "Error: The two sets of 3-D lattices do not have identical bins.\n");
            free_lattice(&L1); free_lattice(&L2);
            return EXIT_FAILURE;
        }
        Lb = &L2;
    }
    int64_t ncp = 0;
    ocellpair *cps = gen_cell_pairs(&L1, Lb, &ncp, rx, ry, rz, C->xwrap, C->ywrap, C->zwrap,
                                    C->max3, C->max2, C->max1, o->enable_min_sep_opt, C->autocorr, C->px, C->py, C->pz);
    oracle_last_stats.nmesh[0] = L1.nx; oracle_last_stats.nmesh[1] = L1.ny; oracle_last_stats.nmesh[2] = L1.nz;
    oracle_last_stats.refine[0] = rx; oracle_last_stats.refine[1] = ry; oracle_last_stats.refine[2] = rz;
    oracle_last_stats.num_cell_pairs = ncp;
    double tested = 0;
    for (int64_t k = 0; k < ncp; k++) {
        const double a = (double)L1.cells[cps[k].c1].n, b = (double)Lb->cells[cps[k].c2].n;
        tested += cps[k].same ? a * (a - 1) / 2 : a * b;
    }
    oracle_last_stats.tested_pairs = tested;

    const int64_t nb = C->totbins;
    int nth = C->numthreads > 0 ? C->numthreads : 1;
#ifdef _OPENMP
    if (nth > omp_get_max_threads()) nth = omp_get_max_threads();
#else
    nth = 1;
#endif
    uint64_t *tnp = calloc((size_t)(nb * nth), sizeof(uint64_t));
    double *tavg = calloc((size_t)(nb * nth), sizeof(double));
    double *twavg = calloc((size_t)(nb * nth), sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        oacc A = {tnp + nb * tid, tavg + nb * tid, twavg + nb * tid};
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
        for (int64_t k = 0; k < ncp; k++)
            if (C->kp.mode >= MODE_RPPI_MOCKS)
                kernel_cellpair_mocks(&C->kp, &L1.cells[cps[k].c1], &Lb->cells[cps[k].c2], &cps[k], &A);
            else
                kernel_cellpair(&C->kp, &L1.cells[cps[k].c1], &Lb->cells[cps[k].c2], &cps[k], &A);
    }
    for (int t = 0; t < nth; t++)
        for (int64_t b = 0; b < nb; b++) {
            np[b] += tnp[nb * t + b];
            avg[b] += tavg[nb * t + b];
            wavg[b] += twavg[nb * t + b];
        }
    free(tnp); free(tavg); free(twavg); free(cps);
    free_lattice(&L1);
    if (!C->autocorr) free_lattice(&L2);
    return EXIT_SUCCESS;
}

/* validation of the per-call options, common to the 5 drivers */
static void fix_options(struct config_options *o, int reset_maxcells_first)
{
    (void)reset_maxcells_first;
    if (o->max_cells_per_dim == 0) {
        fprintf(stderr, "Warning: Max. cells per dimension is set to 0 - resetting to `NLATMAX' = %d\n", NLATMAX);
        o->max_cells_per_dim = NLATMAX;
    }
    for (int i = 0; i < 3; i++) {
        if (o->bin_refine_factors[i] < 1) {
            fprintf(stderr, "Warning: bin refine factor along axis = %d *must* be >=1. Instead found bin refine factor =%d\n",
                    i, o->bin_refine_factors[i]);
            reset_bin_refine_factors(o);
            break;
        }
    }
    o->sort_on_z = 1;
}

/* boxsize resolution for DD / DDrppi / DDsmu (countpairs_impl.c.src:231-251) */
static int resolve_box(const struct config_options *o, double xmin, double xmax, double ymin, double ymax,
                       double zmin, double zmax, int *px, int *py, int *pz, double *xw, double *yw, double *zw)
{
    if (o->periodic && o->boxsize == BOXSIZE_NOTGIVEN) {
        fprintf(stderr, "boxsize = %g must be specified with periodic wrap.\n", o->boxsize);
        return EXIT_FAILURE;
    }
    const double by = o->boxsize_y == BOXSIZE_NOTGIVEN ? o->boxsize : o->boxsize_y;
    const double bz = o->boxsize_z == BOXSIZE_NOTGIVEN ? o->boxsize : o->boxsize_z;
    *px = o->periodic && o->boxsize_x >= 0;
    *py = o->periodic && by >= 0;
    *pz = o->periodic && bz >= 0;
    *xw = *px ? (o->boxsize_x > 0 ? o->boxsize_x : (xmax - xmin)) : 0.;
    *yw = *py ? (by > 0 ? by : (ymax - ymin)) : 0.;
    *zw = *pz ? (bz > 0 ? bz : (zmax - zmin)) : 0.;
    return EXIT_SUCCESS;
}

/* autocorr doubling, self pairs, averages (countpairs_impl.c.src:608-664) */
static void finalize_sums(int autocorr, double rlow0, int64_t ND1, const double *W1, int need_avg, int need_w,
                          int self_slot, int64_t nb, uint64_t *np, double *avg, double *wavg)
{
    if (autocorr == 1) {
        for (int64_t i = 0; i < nb; i++) {
            np[i] *= 2;
            if (need_avg) avg[i] *= 2.0;
            if (need_w) wavg[i] *= 2.0;
        }
        if (rlow0 <= 0.0) {
            np[self_slot] += (uint64_t)ND1;
            if (need_w)
                for (int64_t j = 0; j < ND1; j++) wavg[1] += W1[j] * W1[j];
        }
    }
    for (int64_t i = 0; i < nb; i++) {
        if (np[i] > 0) {
            if (need_avg) avg[i] /= (double)np[i];
            if (need_w) wavg[i] /= (double)np[i];
        }
    }
}

/* ============================================================== DD ===== */
int oracle_countpairs(const int64_t ND1, void *X1, void *Y1, void *Z1,
                      const int64_t ND2, void *X2, void *Y2, void *Z2,
                      const int numthreads, const int autocorr, const char *binfile,
                      results_countpairs *results, struct config_options *options,
                      struct extra_options *extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    struct extra_options dummy;
    if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
    const int need_w = extra->weight_method != NONE;
    const double t0 = now_s();
    fix_options(options, 0);
    double *rupp = NULL, rmin = 0, rmax = 0;
    int nbin = 0;
    if (oracle_setup_bins(binfile, &rmin, &rmax, &nbin, &rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!(rmin >= 0.0 && rmax > 0.0 && rmin < rmax && nbin > 0)) {
        fprintf(stderr, "Error: Could not setup with R bins correctly. (rmin = %lf, rmax = %lf, with nbins = %d).\n", rmin, rmax, nbin);
        free(rupp);
        return EXIT_FAILURE;
    }
    const size_t ft = options->float_type;
    double *x1 = upcast(X1, ND1, ft), *y1 = upcast(Y1, ND1, ft), *z1 = upcast(Z1, ND1, ft);
    double *w1 = need_w ? upcast(extra->weights0.weights[0], ND1, ft) : NULL;
    double *x2 = NULL, *y2 = NULL, *z2 = NULL, *w2 = NULL;
    if (!autocorr) {
        x2 = upcast(X2, ND2, ft); y2 = upcast(Y2, ND2, ft); z2 = upcast(Z2, ND2, ft);
        w2 = need_w ? upcast(extra->weights1.weights[0], ND2, ft) : NULL;
    }
    ocore C;
    memset(&C, 0, sizeof C);
    C.mode = MODE_DD; C.autocorr = autocorr; C.N1 = ND1; C.N2 = ND2;
    C.X1 = x1; C.Y1 = y1; C.Z1 = z1; C.W1 = w1; C.X2 = x2; C.Y2 = y2; C.Z2 = z2; C.W2 = w2;
    C.xmin = C.ymin = C.zmin = DBL_MAX; C.xmax = C.ymax = C.zmax = -DBL_MAX;
    get_max_min(ND1, x1, y1, z1, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
    if (!autocorr) get_max_min(ND2, x2, y2, z2, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
    int status = EXIT_FAILURE;
    uint64_t *np = NULL; double *avg = NULL, *wavg = NULL;
    double *rupp_sqr = malloc(sizeof(double) * nbin);
    if (resolve_box(options, C.xmin, C.xmax, C.ymin, C.ymax, C.zmin, C.zmax, &C.px, &C.py, &C.pz, &C.xwrap, &C.ywrap, &C.zwrap) != EXIT_SUCCESS)
        goto out;
    if (get_bin_refine_scheme(options) == BINNING_DFL) {
        if (rmax < 0.05 * C.xwrap) options->bin_refine_factors[0] = 1;
        if (rmax < 0.05 * C.ywrap) options->bin_refine_factors[1] = 1;
        if (rmax < 0.05 * C.zwrap) options->bin_refine_factors[2] = 1;
    }
    C.gx = C.gy = C.gz = rmax;
    C.max3 = rmax; C.max2 = -1.0; C.max1 = -1.0;
    C.boost_kind = 0;
    C.numthreads = numthreads;
    for (int i = 0; i < nbin; i++) rupp_sqr[i] = rupp[i] * rupp[i];
    const int isa_id = resolve_isa(options->instruction_set);
    C.kp.mode = MODE_DD; C.kp.r2form = r2form_for(MODE_DD, isa_id); C.kp.nbin = nbin;
    C.kp.rupp_sqr = rupp_sqr; C.kp.sqr_rmax = rupp_sqr[nbin - 1]; C.kp.sqr_rmin = rupp_sqr[0];
    C.kp.pimax = rmax; C.kp.need_avg = options->need_avg_sep; C.kp.need_w = need_w;
    C.totbins = nbin;
    np = calloc((size_t)nbin, sizeof(uint64_t)); avg = calloc((size_t)nbin, sizeof(double)); wavg = calloc((size_t)nbin, sizeof(double));
    if (run_core(&C, options, np, avg, wavg) != EXIT_SUCCESS) goto out;
    finalize_sums(autocorr, rupp[0], ND1, w1, options->need_avg_sep, need_w, 1, nbin, np, avg, wavg);
    results->nbin = nbin;
    results->npairs = malloc(sizeof(uint64_t) * nbin);
    results->rupp = malloc(sizeof(double) * nbin);
    results->rpavg = calloc(nbin, sizeof(double));
    results->weightavg = calloc(nbin, sizeof(double));
    for (int i = 0; i < nbin; i++) {
        results->npairs[i] = np[i];
        results->rupp[i] = rupp[i];
        results->rpavg[i] = options->need_avg_sep ? avg[i] : 0.0;
        results->weightavg[i] = need_w ? wavg[i] : 0.0;
    }
    reset_bin_refine_factors(options);
    if (options->c_api_timer) options->c_api_time = now_s() - t0;
    status = EXIT_SUCCESS;
out:
    free(np); free(avg); free(wavg); free(rupp); free(rupp_sqr);
    free(x1); free(y1); free(z1); free(w1); free(x2); free(y2); free(z2); free(w2);
    return status;
}

/* =========================================================== DDrppi ==== */
int oracle_countpairs_rp_pi(const int64_t ND1, void *X1, void *Y1, void *Z1,
                            const int64_t ND2, void *X2, void *Y2, void *Z2,
                            const int numthreads, const int autocorr, const char *binfile,
                            const double pimax_in, results_countpairs_rp_pi *results,
                            struct config_options *options, struct extra_options *extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    struct extra_options dummy;
    if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
    const int need_w = extra->weight_method != NONE;
    const double t0 = now_s();
    fix_options(options, 0);
    const int npibin = (int)pimax_in;
    double *rupp = NULL, rmin = 0, rmax = 0;
    int nbin = 0;
    if (oracle_setup_bins(binfile, &rmin, &rmax, &nbin, &rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!(rmin >= 0.0 && rmax > 0.0 && rmin < rmax && nbin > 0)) {
        fprintf(stderr, "Error: Could not setup with R bins correctly.\n");
        free(rupp);
        return EXIT_FAILURE;
    }
    const int64_t totnbins = (int64_t)(npibin + 1) * (nbin + 1);
    double *rupp_sqr = malloc(sizeof(double) * nbin);
    for (int i = 0; i < nbin; i++) rupp_sqr[i] = rupp[i] * rupp[i];
    const size_t ft = options->float_type;
    double *x1 = upcast(X1, ND1, ft), *y1 = upcast(Y1, ND1, ft), *z1 = upcast(Z1, ND1, ft);
    double *w1 = need_w ? upcast(extra->weights0.weights[0], ND1, ft) : NULL;
    double *x2 = NULL, *y2 = NULL, *z2 = NULL, *w2 = NULL;
    if (!autocorr) {
        x2 = upcast(X2, ND2, ft); y2 = upcast(Y2, ND2, ft); z2 = upcast(Z2, ND2, ft);
        w2 = need_w ? upcast(extra->weights1.weights[0], ND2, ft) : NULL;
    }
    const double pimax = pimax_in;
    ocore C;
    memset(&C, 0, sizeof C);
    C.mode = MODE_RPPI; C.autocorr = autocorr; C.N1 = ND1; C.N2 = ND2;
    C.X1 = x1; C.Y1 = y1; C.Z1 = z1; C.W1 = w1; C.X2 = x2; C.Y2 = y2; C.Z2 = z2; C.W2 = w2;
    C.xmin = C.ymin = C.zmin = DBL_MAX; C.xmax = C.ymax = C.zmax = -DBL_MAX;
    get_max_min(ND1, x1, y1, z1, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
    if (!autocorr) get_max_min(ND2, x2, y2, z2, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
    int status = EXIT_FAILURE;
    uint64_t *np = NULL; double *avg = NULL, *wavg = NULL;
    if (resolve_box(options, C.xmin, C.xmax, C.ymin, C.ymax, C.zmin, C.zmax, &C.px, &C.py, &C.pz, &C.xwrap, &C.ywrap, &C.zwrap) != EXIT_SUCCESS)
        goto out;
    if (get_bin_refine_scheme(options) == BINNING_DFL) {
        if (rmax < 0.05 * C.xwrap) options->bin_refine_factors[0] = 1;
        if (rmax < 0.05 * C.ywrap) options->bin_refine_factors[1] = 1;
        if (pimax < 0.05 * C.zwrap) options->bin_refine_factors[2] = 1;
    }
    C.gx = C.gy = rmax; C.gz = pimax;
    C.max3 = -1.0; C.max2 = rmax; C.max1 = pimax;
    C.boost_kind = 0;
    C.numthreads = numthreads;
    const int isa_id = resolve_isa(options->instruction_set);
    C.kp.mode = MODE_RPPI; C.kp.r2form = r2form_for(MODE_RPPI, isa_id); C.kp.float_index = isa_id == AVX512F;
    C.kp.nbin = nbin; C.kp.rupp_sqr = rupp_sqr; C.kp.sqr_rmax = rupp_sqr[nbin - 1]; C.kp.sqr_rmin = rupp_sqr[0];
    C.kp.pimax = pimax; C.kp.n2 = npibin;
    {
        const double dpi = pimax / npibin;
        C.kp.inv_d2 = 1.0 / dpi;
    }
    C.kp.need_avg = options->need_avg_sep; C.kp.need_w = need_w;
    C.totbins = totnbins;
    np = calloc((size_t)totnbins, sizeof(uint64_t)); avg = calloc((size_t)totnbins, sizeof(double)); wavg = calloc((size_t)totnbins, sizeof(double));
    if (run_core(&C, options, np, avg, wavg) != EXIT_SUCCESS) goto out;
    finalize_sums(autocorr, rupp[0], ND1, w1, options->need_avg_sep, need_w, npibin + 1, totnbins, np, avg, wavg);
    results->nbin = nbin;
    results->npibin = npibin;
    results->pimax = pimax;
    results->npairs = calloc((size_t)totnbins, sizeof(uint64_t));
    results->rupp = malloc(sizeof(double) * nbin);
    results->rpavg = calloc((size_t)totnbins, sizeof(double));
    results->weightavg = calloc((size_t)totnbins, sizeof(double));
    for (int i = 0; i < nbin; i++) {
        results->rupp[i] = rupp[i];
        for (int j = 0; j < npibin; j++) {
            const int64_t idx = (int64_t)i * (npibin + 1) + j;
            results->npairs[idx] = np[idx];
            results->rpavg[idx] = options->need_avg_sep ? avg[idx] : 0.0;
            results->weightavg[idx] = need_w ? wavg[idx] : 0.0;
        }
    }
    reset_bin_refine_factors(options);
    if (options->c_api_timer) options->c_api_time = now_s() - t0;
    status = EXIT_SUCCESS;
out:
    free(np); free(avg); free(wavg); free(rupp); free(rupp_sqr);
    free(x1); free(y1); free(z1); free(w1); free(x2); free(y2); free(z2); free(w2);
    return status;
}

/* ============================================================ DDsmu ==== */
int oracle_countpairs_s_mu(const int64_t ND1, void *X1, void *Y1, void *Z1,
                           const int64_t ND2, void *X2, void *Y2, void *Z2,
                           const int numthreads, const int autocorr, const char *sbinfile,
                           const double max_mu, const int nmu_bins,
                           results_countpairs_s_mu *results, struct config_options *options,
                           struct extra_options *extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    struct extra_options dummy;
    if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
    const int need_w = extra->weight_method != NONE;
    const double t0 = now_s();
    options->sort_on_z = 1;
    for (int i = 0; i < 3; i++)
        if (options->bin_refine_factors[i] < 1) { reset_bin_refine_factors(options); break; }
    if (options->max_cells_per_dim == 0) options->max_cells_per_dim = NLATMAX;
    if (options->fast_divide_and_NR_steps >= MAX_FAST_DIVIDE_NR_STEPS) options->fast_divide_and_NR_steps = 0;
    double *supp = NULL, smin = 0, smax = 0;
    int nsbin = 0;
    if (oracle_setup_bins(sbinfile, &smin, &smax, &nsbin, &supp) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!(smin >= 0.0 && smax > 0.0 && smin < smax && nsbin > 0)) {
        fprintf(stderr, "Error: Could not setup with S bins correctly.\n");
        free(supp);
        return EXIT_FAILURE;
    }
    if (max_mu <= 0.0 || max_mu > 1.0) {
        fprintf(stderr, "Error: max_mu must be > 0 and <= 1.0 (got %lf)\n", max_mu);
        free(supp);
        return EXIT_FAILURE;
    }
    if (nmu_bins < 1) {
        fprintf(stderr, "Error: nmu_bins must be >= 1 (got %d)\n", nmu_bins);
        free(supp);
        return EXIT_FAILURE;
    }
    const int64_t totnbins = (int64_t)(nmu_bins + 1) * (nsbin + 1);
    double *supp_sqr = malloc(sizeof(double) * nsbin);
    for (int i = 0; i < nsbin; i++) supp_sqr[i] = supp[i] * supp[i];
    const double mu_max = max_mu;
    const double pimax = smax * mu_max;
    const size_t ft = options->float_type;
    double *x1 = upcast(X1, ND1, ft), *y1 = upcast(Y1, ND1, ft), *z1 = upcast(Z1, ND1, ft);
    double *w1 = need_w ? upcast(extra->weights0.weights[0], ND1, ft) : NULL;
    double *x2 = NULL, *y2 = NULL, *z2 = NULL, *w2 = NULL;
    if (!autocorr) {
        x2 = upcast(X2, ND2, ft); y2 = upcast(Y2, ND2, ft); z2 = upcast(Z2, ND2, ft);
        w2 = need_w ? upcast(extra->weights1.weights[0], ND2, ft) : NULL;
    }
    ocore C;
    memset(&C, 0, sizeof C);
    C.mode = MODE_SMU; C.autocorr = autocorr; C.N1 = ND1; C.N2 = ND2;
    C.X1 = x1; C.Y1 = y1; C.Z1 = z1; C.W1 = w1; C.X2 = x2; C.Y2 = y2; C.Z2 = z2; C.W2 = w2;
    C.xmin = C.ymin = C.zmin = DBL_MAX; C.xmax = C.ymax = C.zmax = -DBL_MAX;
    get_max_min(ND1, x1, y1, z1, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
    if (!autocorr) get_max_min(ND2, x2, y2, z2, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
    int status = EXIT_FAILURE;
    uint64_t *np = NULL; double *avg = NULL, *wavg = NULL;
    if (resolve_box(options, C.xmin, C.xmax, C.ymin, C.ymax, C.zmin, C.zmax, &C.px, &C.py, &C.pz, &C.xwrap, &C.ywrap, &C.zwrap) != EXIT_SUCCESS)
        goto out;
    C.gx = C.gy = smax; C.gz = pimax;
    C.max3 = smax; C.max2 = -1.0; C.max1 = pimax;
    C.boost_kind = 1;
    C.numthreads = numthreads;
    const int isa_id = resolve_isa(options->instruction_set);
    C.kp.mode = MODE_SMU; C.kp.r2form = r2form_for(MODE_SMU, isa_id); C.kp.float_index = isa_id == AVX512F;
    C.kp.nbin = nsbin; C.kp.rupp_sqr = supp_sqr; C.kp.sqr_rmax = supp_sqr[nsbin - 1]; C.kp.sqr_rmin = supp_sqr[0];
    C.kp.pimax = pimax; C.kp.n2 = nmu_bins;
    {
        const double dmu = mu_max / nmu_bins;
        C.kp.inv_d2 = 1.0 / dmu;
    }
    C.kp.sqr_mu_max = mu_max * mu_max;
    C.kp.need_avg = options->need_avg_sep; C.kp.need_w = need_w;
    C.totbins = totnbins;
    np = calloc((size_t)totnbins, sizeof(uint64_t)); avg = calloc((size_t)totnbins, sizeof(double)); wavg = calloc((size_t)totnbins, sizeof(double));
    if (run_core(&C, options, np, avg, wavg) != EXIT_SUCCESS) goto out;
    finalize_sums(autocorr, supp[0], ND1, w1, options->need_avg_sep, need_w, nmu_bins + 1, totnbins, np, avg, wavg);
    results->nsbin = nsbin;
    results->nmu_bins = nmu_bins;
    results->mu_max = max_mu;
    results->mu_min = 0.0;
    results->npairs = calloc((size_t)totnbins, sizeof(uint64_t));
    results->supp = malloc(sizeof(double) * nsbin);
    results->savg = calloc((size_t)totnbins, sizeof(double));
    results->weightavg = calloc((size_t)totnbins, sizeof(double));
    for (int i = 0; i < nsbin; i++) {
        results->supp[i] = supp[i];
        for (int j = 0; j < nmu_bins; j++) {
            const int64_t idx = (int64_t)i * (nmu_bins + 1) + j;
            results->npairs[idx] = np[idx];
            results->savg[idx] = options->need_avg_sep ? avg[idx] : 0.0;
            results->weightavg[idx] = need_w ? wavg[idx] : 0.0;
        }
    }
    reset_bin_refine_factors(options);
    if (options->c_api_timer) options->c_api_time = now_s() - t0;
    status = EXIT_SUCCESS;
out:
    free(np); free(avg); free(wavg); free(supp); free(supp_sqr);
    free(x1); free(y1); free(z1); free(w1); free(x2); free(y2); free(z2); free(w2);
    return status;
}

/* ============================================================ xi / wp == */
static int xi_wp_common(int mode, const int64_t ND, void *X, void *Y, void *Z, const double boxsize,
                        const int numthreads, const char *binfile, const double pimax_in,
                        uint64_t **np_out, double **avg_out, double **wavg_out, double **rupp_out, int *nbin_out,
                        double **w_out, struct config_options *options, struct extra_options *extra, int *need_w_out)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    struct extra_options dummy;
    if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
    const int need_w = extra->weight_method != NONE;
    *need_w_out = need_w;
    options->periodic = 1;
    options->autocorr = 1;
    options->sort_on_z = 1;
    if (mode == MODE_XI && options->max_cells_per_dim == 0) options->max_cells_per_dim = NLATMAX;
    for (int i = 0; i < 3; i++)
        if (options->bin_refine_factors[i] < 1) { reset_bin_refine_factors(options); break; }
    if (options->max_cells_per_dim == 0) options->max_cells_per_dim = NLATMAX;
    double *rupp = NULL, rmin = 0, rmax = 0;
    int nbin = 0;
    if (oracle_setup_bins(binfile, &rmin, &rmax, &nbin, &rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!(rmin >= 0.0 && rmax > 0.0 && rmin < rmax && nbin > 0)) {
        fprintf(stderr, "Error: Could not setup with R bins correctly.\n");
        free(rupp);
        return EXIT_FAILURE;
    }
    if (get_bin_refine_scheme(options) == BINNING_DFL) {
        if (mode == MODE_XI) {
            if (rmax < 0.05 * boxsize)
                for (int i = 0; i < 3; i++) options->bin_refine_factors[i] = 1;
        } else {
            if (rmax < 0.05 * boxsize)
                for (int i = 0; i < 2; i++) options->bin_refine_factors[i] = 1;
            if (pimax_in < 0.05 * boxsize) options->bin_refine_factors[2] = 1;
        }
    }
    const size_t ft = options->float_type;
    double *x = upcast(X, ND, ft), *y = upcast(Y, ND, ft), *z = upcast(Z, ND, ft);
    double *w = need_w ? upcast(extra->weights0.weights[0], ND, ft) : NULL;
    double *rupp_sqr = malloc(sizeof(double) * nbin);
    for (int i = 0; i < nbin; i++) rupp_sqr[i] = rupp[i] * rupp[i];
    ocore C;
    memset(&C, 0, sizeof C);
    C.mode = mode; C.autocorr = 1; C.N1 = ND;
    C.X1 = x; C.Y1 = y; C.Z1 = z; C.W1 = w;
    C.xmin = C.ymin = C.zmin = 0.0; C.xmax = C.ymax = C.zmax = boxsize;
    C.xwrap = C.ywrap = C.zwrap = boxsize;
    C.px = C.py = C.pz = 1;
    C.boost_kind = 0;
    C.numthreads = numthreads;
    const int isa_id = resolve_isa(options->instruction_set);
    C.kp.mode = mode; C.kp.r2form = r2form_for(mode, isa_id);
    C.kp.nbin = nbin; C.kp.rupp_sqr = rupp_sqr; C.kp.sqr_rmax = rupp_sqr[nbin - 1]; C.kp.sqr_rmin = rupp_sqr[0];
    C.kp.need_avg = options->need_avg_sep; C.kp.need_w = need_w;
    if (mode == MODE_XI) {
        C.gx = C.gy = C.gz = rmax;
        C.max3 = rmax; C.max2 = -1.0; C.max1 = -1.0;
        C.kp.pimax = rmax;
    } else {
        C.gx = C.gy = rmax; C.gz = pimax_in;
        C.max3 = -1.0; C.max2 = rmax; C.max1 = pimax_in;
        C.kp.pimax = pimax_in;
    }
    C.totbins = nbin;
    uint64_t *np = calloc((size_t)nbin, sizeof(uint64_t));
    double *avg = calloc((size_t)nbin, sizeof(double)), *wavg = calloc((size_t)nbin, sizeof(double));
    int st = run_core(&C, options, np, avg, wavg);
    free(rupp_sqr); free(x); free(y); free(z);
    if (st != EXIT_SUCCESS) { free(np); free(avg); free(wavg); free(rupp); free(w); return EXIT_FAILURE; }
    finalize_sums(1, rupp[0], ND, w, options->need_avg_sep, need_w, 1, nbin, np, avg, wavg);
    *np_out = np; *avg_out = avg; *wavg_out = wavg; *rupp_out = rupp; *nbin_out = nbin; *w_out = w;
    return EXIT_SUCCESS;
}

int oracle_countpairs_xi(const int64_t ND, void *X, void *Y, void *Z, const double boxsize,
                         const int numthreads, const char *binfile, results_countpairs_xi *results,
                         struct config_options *options, struct extra_options *extra)
{
    const double t0 = now_s();
    uint64_t *np; double *avg, *wavg, *rupp, *w; int nbin, need_w;
    if (xi_wp_common(MODE_XI, ND, X, Y, Z, boxsize, numthreads, binfile, 0.0, &np, &avg, &wavg, &rupp, &nbin, &w,
                     options, extra, &need_w) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    results->nbin = nbin;
    results->npairs = malloc(sizeof(uint64_t) * nbin);
    results->xi = malloc(sizeof(double) * nbin);
    results->rupp = malloc(sizeof(double) * nbin);
    results->ravg = malloc(sizeof(double) * nbin);
    results->weightavg = calloc(nbin, sizeof(double));
    double wsum = (double)ND, wsq = (double)ND;
    if (need_w) {
        wsum = 0.;
        for (int64_t j = 0; j < ND; j++) { wsum += w[j]; wsq += w[j] * w[j]; }
    }
    const double prefac = wsum * (wsum - wsum / ND) / (boxsize * boxsize * boxsize);
    double rlow = 0.0;
    for (int i = 0; i < nbin; i++) {
        results->npairs[i] = np[i];
        results->rupp[i] = rupp[i];
        results->ravg[i] = options->need_avg_sep ? avg[i] : 0.0;
        results->weightavg[i] = need_w ? wavg[i] : 0.0;
        double weight0 = (double)results->npairs[i];
        if (need_w) weight0 *= results->weightavg[i];
        const double vol = 4.0 / 3.0 * M_PI * (rupp[i] * rupp[i] * rupp[i] - rlow * rlow * rlow);
        if (vol > 0.0) {
            double wr = prefac * vol;
            if (rlow <= 0.) wr += wsq;
            results->xi[i] = weight0 / wr - 1.0;
        } else {
            results->xi[i] = -2.0;
        }
        rlow = results->rupp[i];
    }
    free(np); free(avg); free(wavg); free(rupp); free(w);
    reset_bin_refine_factors(options);
    if (options->c_api_timer) options->c_api_time = now_s() - t0;
    return EXIT_SUCCESS;
}

int oracle_countpairs_wp(const int64_t ND, void *X, void *Y, void *Z, const double boxsize,
                         const int numthreads, const char *binfile, const double pimax,
                         results_countpairs_wp *results, struct config_options *options,
                         struct extra_options *extra)
{
    const double t0 = now_s();
    uint64_t *np; double *avg, *wavg, *rupp, *w; int nbin, need_w;
    if (xi_wp_common(MODE_WP, ND, X, Y, Z, boxsize, numthreads, binfile, pimax, &np, &avg, &wavg, &rupp, &nbin, &w,
                     options, extra, &need_w) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    results->nbin = nbin;
    results->pimax = pimax;
    results->npairs = malloc(sizeof(uint64_t) * nbin);
    results->wp = malloc(sizeof(double) * nbin);
    results->rupp = malloc(sizeof(double) * nbin);
    results->rpavg = malloc(sizeof(double) * nbin);
    results->weightavg = calloc(nbin, sizeof(double));
    double wsum = (double)ND, wsq = (double)ND;
    if (need_w) {
        wsum = 0;
        for (int64_t j = 0; j < ND; j++) { wsum += w[j]; wsq += w[j] * w[j]; }
    }
    const double prefac = wsum * (wsum - wsum / ND) / (boxsize * boxsize * boxsize);
    double rlow = 0.0;
    const double twice_pimax = 2.0 * pimax;
    for (int i = 0; i < nbin; i++) {
        results->npairs[i] = np[i];
        results->rupp[i] = rupp[i];
        results->rpavg[i] = options->need_avg_sep ? avg[i] : 0.0;
        results->weightavg[i] = need_w ? wavg[i] : 0.0;
        double weight0 = (double)results->npairs[i];
        if (need_w) weight0 *= results->weightavg[i];
        const double vol = M_PI * (results->rupp[i] * results->rupp[i] - rlow * rlow) * twice_pimax;
        if (vol > 0.0) {
            double wr = prefac * vol;
            if (rlow <= 0.) wr += wsq;
            results->wp[i] = (weight0 / wr - 1) * twice_pimax;
        } else {
            results->wp[i] = -2.0 * twice_pimax;
        }
        rlow = results->rupp[i];
    }
    free(np); free(avg); free(wavg); free(rupp); free(w);
    reset_bin_refine_factors(options);
    if (options->c_api_timer) options->c_api_time = now_s() - t0;
    return EXIT_SUCCESS;
}

void oracle_free_results(results_countpairs *r) { free(r->npairs); free(r->rupp); free(r->rpavg); free(r->weightavg); memset(r, 0, sizeof *r); }
void oracle_free_results_rp_pi(results_countpairs_rp_pi *r) { free(r->npairs); free(r->rupp); free(r->rpavg); free(r->weightavg); memset(r, 0, sizeof *r); }
void oracle_free_results_s_mu(results_countpairs_s_mu *r) { free(r->npairs); free(r->supp); free(r->savg); free(r->weightavg); memset(r, 0, sizeof *r); }
void oracle_free_results_xi(results_countpairs_xi *r) { free(r->npairs); free(r->xi); free(r->rupp); free(r->ravg); free(r->weightavg); memset(r, 0, sizeof *r); }
void oracle_free_results_wp(results_countpairs_wp *r) { free(r->npairs); free(r->wp); free(r->rupp); free(r->rpavg); free(r->weightavg); memset(r, 0, sizeof *r); }

/* ======================================================= on-sky (mocks) ==== */
#define MOCKS_SPEED_OF_LIGHT 299800.0   /* utils/set_cosmo_dist.h:19 */
#define MOCKS_PI_OVER_180 0.017453292519943295769236907684886127134428718885417254560971

/* check_ra_dec_cz (countpairs_rp_pi_mocks_impl.c.src:43-101): fixes the caller's
 * arrays in place, in their own precision. */
static int mocks_check_ra_dec_cz(int64_t N, void *phi_v, void *theta_v, void *cz_v, size_t ft, int is_comoving)
{
    if (N == 0) return EXIT_SUCCESS;
    if (phi_v == NULL || theta_v == NULL || cz_v == NULL) {
        fprintf(stderr, "Input arrays can not be NULL.\n");
        return EXIT_FAILURE;
    }
    int fix_cz = 0, fix_ra = 0, fix_dec = 0;
    double max_cz = 0.0;
    for (int64_t i = 0; i < N; i++) {
        const double ph = ft == 4 ? ((float *)phi_v)[i] : ((double *)phi_v)[i];
        const double th = ft == 4 ? ((float *)theta_v)[i] : ((double *)theta_v)[i];
        const double c = ft == 4 ? ((float *)cz_v)[i] : ((double *)cz_v)[i];
        if (c > max_cz) max_cz = c;
        if (ph < 0.0) fix_ra = 1;
        if (th > 90.0) fix_dec = 1;
        if (th > 180) {
            fprintf(stderr, "theta[%lld] = %g should be less than 180 deg\n", (long long)i, th);
            return EXIT_FAILURE;
        }
    }
    if (max_cz < 10.0 && is_comoving == 0) fix_cz = 1;
    for (int64_t i = 0; i < N && (fix_cz || fix_ra || fix_dec); i++) {
        if (ft == 4) {
            float *ph = phi_v, *th = theta_v, *c = cz_v;
            if (fix_ra) ph[i] += (float)180.0;
            if (fix_dec) th[i] -= (float)90.0;
            if (fix_cz) c[i] *= (float)MOCKS_SPEED_OF_LIGHT;
        } else {
            double *ph = phi_v, *th = theta_v, *c = cz_v;
            if (fix_ra) ph[i] += 180.0;
            if (fix_dec) th[i] -= 90.0;
            if (fix_cz) c[i] *= MOCKS_SPEED_OF_LIGHT;
        }
    }
    return EXIT_SUCCESS;
}

/* cosmology_params.c:26-56 */
static int mocks_cosmology(int which, double *om, double *ol)
{
    switch (which) {
    case 1: *om = 0.25; break;
    case 2: *om = 0.302; break;
    default:
        fprintf(stderr, "ERROR: Cosmology=%d not implemented\n", which);
        return EXIT_FAILURE;
    }
    *ol = 1.0 - *om;
    return EXIT_SUCCESS;
}

/* set_cosmo_dist.c:25-71: comoving distance (Mpc/h) tabulated every 1/max_size in
 * z by Simpson's rule on steps of 1e-2/max_size. Returns the table length. */
static int mocks_cosmo_table(double zmax, int max_size, double *zc, double *dc, double om, double ol)
{
    int i = 0;
    const double ok = 1.0 - om - ol;
    const double Dh = MOCKS_SPEED_OF_LIGHT * 0.01 / 1.0;
    const double Deltaz = 1.0 / max_size;
    const double dz = 1e-2 * Deltaz;
    double Eint = 0.0, E2 = 1.0, z2 = Deltaz;
    for (double z = 2.0 * dz; z < zmax; z += 2.0 * dz) {
        const double E0 = E2;
        const double a = 1 + z - dz, b = 1 + z;
        const double E1 = 1.0 / sqrt(om * (a * a * a) + ok * a + ol);
        E2 = 1.0 / sqrt(om * (b * b * b) + ok * b + ol);
        Eint += dz * (E0 + 4. * E1 + E2) / 3.;
        if (z > (z2 - 1e-10) && z < (z2 + 1e-10)) {
            if (i >= max_size) break;
            zc[i] = z;
            dc[i] = Eint * Dh;
            z2 += Deltaz;
            i++;
        }
    }
    return i;
}

/* Linear interpolation in a table (the published textbook rule: bracket by
 * bisection, y = y_lo + (x - x_lo) / (x_hi - x_lo) * (y_hi - y_lo)); x outside
 * the table is an error. */
static int mocks_interp(const double *xa, const double *ya, int n, double x, double *y)
{
    if (n < 2 || !(x >= xa[0] && x <= xa[n - 1])) return EXIT_FAILURE;
    size_t lo = 0, hi = (size_t)n - 1;
    while (hi > lo + 1) {
        const size_t m = (hi + lo) / 2;
        if (xa[m] > x) hi = m; else lo = m;
    }
    const double dx = xa[lo + 1] - xa[lo];
    *y = dx > 0.0 ? ya[lo] + (x - xa[lo]) / dx * (ya[lo + 1] - ya[lo]) : 0.0;
    return EXIT_SUCCESS;
}

/* (RA, DEC, cz or distance) -> Cartesian (countpairs_rp_pi_mocks_impl.c.src:309-390) */
static int mocks_positions(int64_t N1, const double *ra1, const double *dec1, const double *cz1,
                           int64_t N2, const double *ra2, const double *dec2, const double *cz2, int autocorr,
                           int is_comoving, int cosmology, double **X1, double **Y1, double **Z1,
                           double **X2, double **Y2, double **Z2)
{
    double om, ol;
    if (mocks_cosmology(cosmology, &om, &ol) != EXIT_SUCCESS) return EXIT_FAILURE;
    double *D1 = malloc(sizeof(double) * (size_t)(N1 > 0 ? N1 : 1));
    double *D2 = autocorr ? D1 : malloc(sizeof(double) * (size_t)(N2 > 0 ? N2 : 1));
    if (is_comoving) {
        memcpy(D1, cz1, sizeof(double) * (size_t)N1);
        if (!autocorr) memcpy(D2, cz2, sizeof(double) * (size_t)N2);
    } else {
        double czmax = 0.0;
        const double inv_c = 1.0 / MOCKS_SPEED_OF_LIGHT;
        for (int64_t i = 0; i < N1; i++) if (cz1[i] > czmax) czmax = cz1[i];
        if (!autocorr) for (int64_t i = 0; i < N2; i++) if (cz2[i] > czmax) czmax = cz2[i];
        const double zmax = czmax * inv_c + 0.01;
        const int ws = 10000;
        double *zc = calloc(ws, sizeof(double)), *dc = calloc(ws, sizeof(double));
        const int nz = mocks_cosmo_table(zmax, ws, zc, dc, om, ol);
        int bad = 0;
        for (int64_t i = 0; i < N1 && !bad; i++) bad = mocks_interp(zc, dc, nz, cz1[i] * inv_c, &D1[i]);
        if (!autocorr)
            for (int64_t i = 0; i < N2 && !bad; i++) bad = mocks_interp(zc, dc, nz, cz2[i] * inv_c, &D2[i]);
        free(zc); free(dc);
        if (bad) {
            fprintf(stderr, "oracle> cz outside the tabulated redshift range\n");
            free(D1); if (!autocorr) free(D2);
            return EXIT_FAILURE;
        }
    }
    const int64_t ns[2] = {N1, autocorr ? 0 : N2};
    const double *ra[2] = {ra1, ra2}, *dec[2] = {dec1, dec2}, *D[2] = {D1, D2};
    double **out[2][3] = {{X1, Y1, Z1}, {X2, Y2, Z2}};
    for (int s = 0; s < 2; s++) {
        if (s == 1 && autocorr) { *X2 = *Y2 = *Z2 = NULL; break; }
        const int64_t n = ns[s];
        double *x = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double *y = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double *z = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        for (int64_t i = 0; i < n; i++) {
            const double cd = cos(dec[s][i] * MOCKS_PI_OVER_180);
            x[i] = D[s][i] * cd * cos(ra[s][i] * MOCKS_PI_OVER_180);
            y[i] = D[s][i] * cd * sin(ra[s][i] * MOCKS_PI_OVER_180);
            z[i] = D[s][i] * sin(dec[s][i] * MOCKS_PI_OVER_180);
        }
        *out[s][0] = x; *out[s][1] = y; *out[s][2] = z;
    }
    free(D1);
    if (!autocorr) free(D2);
    return EXIT_SUCCESS;
}

/* Shared driver of the two on-sky modes (countpairs_rp_pi_mocks_impl.c.src:200-847,
 * countpairs_s_mu_mocks_impl.c.src:200-850). lim = pimax or mu_max; n2 = npibin
 * or nmu_bins. Fills per-slot sums already doubled and averaged. */
static int mocks_common(int mode, int64_t ND1, void *RA1, void *DEC1, void *CZ1, int64_t ND2, void *RA2,
                        void *DEC2, void *CZ2, int numthreads, int autocorr, const char *binfile, double lim,
                        int n2, int cosmology, struct config_options *options, struct extra_options *extra,
                        double **rupp_out, int *nbin_out, uint64_t **np_out, double **avg_out, double **wavg_out)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    struct extra_options dummy;
    if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
    const int need_w = extra->weight_method != NONE;
    fix_options(options, 0);
    const size_t ft = options->float_type;
    if (mocks_check_ra_dec_cz(ND1, RA1, DEC1, CZ1, ft, options->is_comoving_dist) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!autocorr && mocks_check_ra_dec_cz(ND2, RA2, DEC2, CZ2, ft, options->is_comoving_dist) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    double om, ol;
    if (mocks_cosmology(cosmology, &om, &ol) != EXIT_SUCCESS) return EXIT_FAILURE;
    double *rupp = NULL, rmin = 0, rmax = 0;
    int nbin = 0;
    if (oracle_setup_bins(binfile, &rmin, &rmax, &nbin, &rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!(rmin > 0.0 && rmax > 0.0 && rmin < rmax && nbin > 0)) {
        fprintf(stderr, "Error: Could not setup with bins correctly. (min = %lf, max = %lf, with nbins = %d).\n",
                rmin, rmax, nbin);
        free(rupp);
        return EXIT_FAILURE;
    }
    if (mode == MODE_SMU_MOCKS) {
        if (lim <= 0.0 || lim > 1.0) {
            fprintf(stderr, "Error: max_mu must be greater than 0 and at most 1. The passed value is max_mu = %lf.\n", lim);
            free(rupp);
            return EXIT_FAILURE;
        }
        if (n2 < 1) {
            fprintf(stderr, "Error: Number of mu bins = %d must be at least 1\n", n2);
            free(rupp);
            return EXIT_FAILURE;
        }
    }
    const int64_t totbins = (int64_t)(n2 + 1) * (nbin + 1);
    uint64_t *np = calloc((size_t)totbins, sizeof(uint64_t));
    double *avg = calloc((size_t)totbins, sizeof(double)), *wavg = calloc((size_t)totbins, sizeof(double));
    *rupp_out = rupp; *nbin_out = nbin; *np_out = np; *avg_out = avg; *wavg_out = wavg;
    if (ND1 == 0 || (autocorr == 0 && ND2 == 0)) return EXIT_SUCCESS;

    double *ra1 = upcast(RA1, ND1, ft), *dec1 = upcast(DEC1, ND1, ft), *cz1 = upcast(CZ1, ND1, ft);
    double *w1 = need_w ? upcast(extra->weights0.weights[0], ND1, ft) : NULL;
    double *ra2 = NULL, *dec2 = NULL, *cz2 = NULL, *w2 = NULL;
    if (!autocorr) {
        ra2 = upcast(RA2, ND2, ft); dec2 = upcast(DEC2, ND2, ft); cz2 = upcast(CZ2, ND2, ft);
        w2 = need_w ? upcast(extra->weights1.weights[0], ND2, ft) : NULL;
    }
    double *X1 = NULL, *Y1 = NULL, *Z1 = NULL, *X2 = NULL, *Y2 = NULL, *Z2 = NULL;
    double *rupp_sqr = malloc(sizeof(double) * nbin);
    for (int i = 0; i < nbin; i++) rupp_sqr[i] = rupp[i] * rupp[i];
    int status = EXIT_FAILURE;
    if (mocks_positions(ND1, ra1, dec1, cz1, ND2, ra2, dec2, cz2, autocorr, options->is_comoving_dist, cosmology,
                        &X1, &Y1, &Z1, &X2, &Y2, &Z2) != EXIT_SUCCESS)
        goto out;
    {
        const double sqr_max_sep = mode == MODE_RPPI_MOCKS ? rmax * rmax + lim * lim : rmax * rmax;
        const double max_sep = sqrt(sqr_max_sep);
        ocore C;
        memset(&C, 0, sizeof C);
        C.mode = mode; C.autocorr = autocorr; C.N1 = ND1; C.N2 = ND2;
        C.X1 = X1; C.Y1 = Y1; C.Z1 = Z1; C.W1 = w1;
        C.X2 = autocorr ? NULL : X2; C.Y2 = autocorr ? NULL : Y2; C.Z2 = autocorr ? NULL : Z2; C.W2 = w2;
        C.xmin = C.ymin = C.zmin = DBL_MAX; C.xmax = C.ymax = C.zmax = -DBL_MAX;
        get_max_min(ND1, X1, Y1, Z1, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
        if (!autocorr) get_max_min(ND2, X2, Y2, Z2, &C.xmin, &C.ymin, &C.zmin, &C.xmax, &C.ymax, &C.zmax);
        if (get_bin_refine_scheme(options) == BINNING_DFL) {
            if (max_sep < 0.05 * (C.xmax - C.xmin)) options->bin_refine_factors[0] = 1;
            if (max_sep < 0.05 * (C.ymax - C.ymin)) options->bin_refine_factors[1] = 1;
            if (max_sep < 0.05 * (C.zmax - C.zmin)) options->bin_refine_factors[2] = 1;
        }
        C.gx = C.gy = C.gz = max_sep;
        C.max3 = max_sep; C.max2 = -1.0; C.max1 = -1.0;
        C.boost_kind = 2;
        C.numthreads = numthreads;
        C.kp.mode = mode; C.kp.r2form = R2_SUM_XY_Z; C.kp.float_index = 0;
        C.kp.nbin = nbin; C.kp.rupp_sqr = rupp_sqr;
        C.kp.sqr_rmax = sqr_max_sep; C.kp.sqr_rmax_bins = rmax * rmax; C.kp.sqr_rmin = rmin * rmin;
        C.kp.pimax = lim; C.kp.n2 = n2;
        {
            const double d2 = mode == MODE_RPPI_MOCKS ? lim / n2 : lim / (double)n2;
            C.kp.inv_d2 = 1.0 / d2;
        }
        C.kp.need_avg = options->need_avg_sep; C.kp.need_w = need_w;
        C.totbins = totbins;
        if (run_core(&C, options, np, avg, wavg) != EXIT_SUCCESS) goto out;
        finalize_sums(autocorr, rupp[0], ND1, w1, options->need_avg_sep, need_w, n2 + 1, totbins, np, avg, wavg);
    }
    status = EXIT_SUCCESS;
out:
    free(ra1); free(dec1); free(cz1); free(w1); free(ra2); free(dec2); free(cz2); free(w2);
    free(X1); free(Y1); free(Z1); free(X2); free(Y2); free(Z2); free(rupp_sqr);
    reset_bin_refine_factors(options);
    return status;
}

int oracle_countpairs_mocks(const int64_t ND1, void *RA1, void *DEC1, void *CZ1,
                            const int64_t ND2, void *RA2, void *DEC2, void *CZ2,
                            const int numthreads, const int autocorr, const char *binfile, const double pimax,
                            const int cosmology, results_countpairs_mocks *results,
                            struct config_options *options, struct extra_options *extra)
{
    const double t0 = now_s();
    const int npibin = (int)pimax;
    double *rupp = NULL, *avg = NULL, *wavg = NULL;
    uint64_t *np = NULL;
    int nbin = 0;
    if (mocks_common(MODE_RPPI_MOCKS, ND1, RA1, DEC1, CZ1, ND2, RA2, DEC2, CZ2, numthreads, autocorr, binfile,
                     pimax, npibin, cosmology, options, extra, &rupp, &nbin, &np, &avg, &wavg) != EXIT_SUCCESS) {
        free(rupp); free(np); free(avg); free(wavg);
        return EXIT_FAILURE;
    }
    const int64_t tot = (int64_t)(npibin + 1) * (nbin + 1);
    const int need_w = extra != NULL && extra->weight_method != NONE;
    results->nbin = nbin; results->npibin = npibin; results->pimax = pimax;
    results->npairs = calloc((size_t)tot, sizeof(uint64_t));
    results->rupp = malloc(sizeof(double) * nbin);
    results->rpavg = calloc((size_t)tot, sizeof(double));
    results->weightavg = calloc((size_t)tot, sizeof(double));
    for (int i = 0; i < nbin; i++) {
        results->rupp[i] = rupp[i];
        for (int j = 0; j < npibin; j++) {
            const int64_t k = (int64_t)i * (npibin + 1) + j;
            results->npairs[k] = np[k];
            results->rpavg[k] = options->need_avg_sep ? avg[k] : 0.0;
            results->weightavg[k] = need_w ? wavg[k] : 0.0;
        }
    }
    free(rupp); free(np); free(avg); free(wavg);
    if (options->c_api_timer) options->c_api_time = now_s() - t0;
    return EXIT_SUCCESS;
}

int oracle_countpairs_mocks_s_mu(const int64_t ND1, void *RA1, void *DEC1, void *CZ1,
                                 const int64_t ND2, void *RA2, void *DEC2, void *CZ2,
                                 const int numthreads, const int autocorr, const char *sbinfile, const double mu_max,
                                 const int nmu_bins, const int cosmology, results_countpairs_mocks_s_mu *results,
                                 struct config_options *options, struct extra_options *extra)
{
    const double t0 = now_s();
    double *supp = NULL, *avg = NULL, *wavg = NULL;
    uint64_t *np = NULL;
    int nbin = 0;
    if (mocks_common(MODE_SMU_MOCKS, ND1, RA1, DEC1, CZ1, ND2, RA2, DEC2, CZ2, numthreads, autocorr, sbinfile,
                     mu_max, nmu_bins, cosmology, options, extra, &supp, &nbin, &np, &avg, &wavg) != EXIT_SUCCESS) {
        free(supp); free(np); free(avg); free(wavg);
        return EXIT_FAILURE;
    }
    const int64_t tot = (int64_t)(nmu_bins + 1) * (nbin + 1);
    const int need_w = extra != NULL && extra->weight_method != NONE;
    results->nsbin = nbin; results->nmu_bins = nmu_bins; results->mu_max = mu_max; results->mu_min = 0.0;
    results->npairs = calloc((size_t)tot, sizeof(uint64_t));
    results->supp = malloc(sizeof(double) * nbin);
    results->savg = calloc((size_t)tot, sizeof(double));
    results->weightavg = calloc((size_t)tot, sizeof(double));
    for (int i = 0; i < nbin; i++) {
        results->supp[i] = supp[i];
        for (int j = 0; j < nmu_bins; j++) {
            const int64_t k = (int64_t)i * (nmu_bins + 1) + j;
            results->npairs[k] = np[k];
            results->savg[k] = options->need_avg_sep ? avg[k] : 0.0;
            results->weightavg[k] = need_w ? wavg[k] : 0.0;
        }
    }
    free(supp); free(np); free(avg); free(wavg);
    if (options->c_api_timer) options->c_api_time = now_s() - t0;
    return EXIT_SUCCESS;
}

void oracle_free_results_mocks(results_countpairs_mocks *r) { free(r->npairs); free(r->rupp); free(r->rpavg); free(r->weightavg); memset(r, 0, sizeof *r); }
void oracle_free_results_mocks_s_mu(results_countpairs_mocks_s_mu *r) { free(r->npairs); free(r->supp); free(r->savg); free(r->weightavg); memset(r, 0, sizeof *r); }
