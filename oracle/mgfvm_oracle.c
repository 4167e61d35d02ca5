/*
 * mgfvm_oracle.c -- CPU restatement of the reference solver's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against.  Only tests/, the smoke() entry and the cpu_baseline leg
 * of bench.py load it; the product library never links or calls it.
 *
 * What it restates: the FAS multigrid V-cycle of SIMPLE relaxations on a
 * staggered 2-D grid in Version-6-26-04.c (the reference "V6"), compiled
 * without OpenACC (the CPU build).  Every routine below cites the V6 lines it
 * follows.  The floating-point expression trees are kept exactly as C
 * evaluates the reference's expressions (left-to-right association, no FMA:
 * build with -ffp-contract=off), so that the results are bit-identical to the
 * reference CPU build.
 *
 * Pinning (how this restatement is itself checked):
 *   - end-to-end: fields and residual history bit-identical to the real
 *     reference built from its own sources (tests/golden fixtures written by
 *     tests/golden/make_golden.py via oracle/ref), for Jacobi and red-black,
 *     with obstacles, discrete inlets/outlets, ibc 2/3/4 and dx != dy;
 *   - per routine: known-answer states produced by calling the reference's
 *     own functions (harness B in oracle/ref) on staged states;
 *   - mathematics: closed forms and invariants in tests/test_oracle_math.py.
 *
 * Deliberate definitions where the reference is undefined:
 *   - levels ig >= ig_obs (the reference leaves their obstacle index arrays
 *     uninitialised and reads them; V6:471, 939-963, 1153-1177): a cell is
 *     an obstacle when all of its children are; the restriction onto such a
 *     level zeroes the fine residual at penalised points (penalty_mask) and
 *     the prolongation from it drops corrections at its penalised points.
 *     Equal to the reference whenever ig_obs == ngrid.  For levels >= ig_obs
 *     this is pinned only by invariants (a converged state is a V-cycle
 *     fixed point), not by the reference.
 *   - an optional per-row inlet u profile (level-0 ghost face) replaces the
 *     discrete-inlet format for long profiles; applied after init and before
 *     the coarse-geometry restriction so coarse inlets follow V6:1440-1446.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mgfvm_oracle.h"

/* ------------------------------------------------------------------ */
/* error reporting                                                      */
/* ------------------------------------------------------------------ */

static char g_err[512];

const char *orc_last_error(void) { return g_err; }

static int fail(int code, const char *msg)
{
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

/* ------------------------------------------------------------------ */
/* state                                                                */
/* ------------------------------------------------------------------ */

typedef struct {
    int npx, npy;
    double dx, dy, amu;
    int has_obs;                     /* obstacles act on this level */
    int nob;                         /* obstacle rectangles on this level */
    int ext;                         /* level >= ig_obs: derived (coarsened) obstacles */
    int *isobs, *ieobs, *jsobs, *jeobs;
    int *jsinlt, *jeinlt, *jsoutlt, *jeoutlt;
    int *tag;
    double *areax, *areay;
    double *ae, *aw, *an, *as, *ap;
    double *u, *v, *p, *pp, *sc, *su, *rs;
    double *apu, *apv, *spu, *spv, *spsc;
    double *resu, *resv, *ressc;
    double *c;
    double *cx, *cy;
} lvl_t;

struct orc_solver {
    orc_config cfg;
    lvl_t *L;
    double *hist;                    /* nvcycle entries */
    int hist_len;
    int verbose;
    int gs_lex;                      /* isolve 2 sweeps in lexicographic order (variant) */
};

/* names of the double arrays, in a fixed order (used by get/set) */
static const char *k_dnames[] = {
    "areax", "areay", "ae", "aw", "an", "as", "ap", "u", "v", "p", "pp",
    "sc", "su", "rs", "apu", "apv", "spu", "spv", "spsc", "resu", "resv",
    "ressc", "c", "cx", "cy", NULL};

static double **darr(lvl_t *l, const char *name)
{
    double **tab[] = {&l->areax, &l->areay, &l->ae, &l->aw, &l->an, &l->as,
                      &l->ap, &l->u, &l->v, &l->p, &l->pp, &l->sc, &l->su,
                      &l->rs, &l->apu, &l->apv, &l->spu, &l->spv, &l->spsc,
                      &l->resu, &l->resv, &l->ressc, &l->c, &l->cx, &l->cy};
    for (int k = 0; k_dnames[k]; ++k)
        if (strcmp(k_dnames[k], name) == 0) return tab[k];
    return NULL;
}

/* ------------------------------------------------------------------ */
/* input parsing: V6:280-401 (readInput), V6:2367-2400 (read_obstacle) */
/* ------------------------------------------------------------------ */

static int read_pattern(const char *path, orc_config *c)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(-2, "cannot open obstacle pattern file");
    if (fscanf(f, "%d", &c->nobs) != 1) {
        fclose(f);
        return fail(-3, "failed reading obstacle count");
    }
    if (c->nobs < 0) c->nobs = 0;
    c->obs = c->nobs ? (int *)malloc(sizeof(int) * 4 * (size_t)c->nobs) : NULL;
    for (int k = 0; k < c->nobs; ++k) {
        int *o = c->obs + 4 * k;
        if (fscanf(f, " %d %d %d %d", &o[0], &o[1], &o[2], &o[3]) != 4) {
            fclose(f);
            return fail(-3, "failed reading obstacle coordinates");
        }
    }
    fclose(f);
    return 0;
}

int orc_parse_files(const char *inputdata, const char *pattern, orc_config *c)
{
    memset(c, 0, sizeof *c);
    FILE *f = fopen(inputdata, "r");
    if (!f) return fail(-2, "cannot open InputData");
    int rc = 0;
#define RD(n, ...)                                                        \
    do {                                                                  \
        if (fscanf(f, __VA_ARGS__) != (n)) {                              \
            fclose(f);                                                    \
            return fail(-3, "failed reading InputData");                  \
        }                                                                 \
    } while (0)
    RD(13, "%d %d %d %d %d %d %d %d %d %d %d %lf %lf", &c->nfx, &c->nfy,
       &c->ngrid, &c->ig_obs, &c->ibeg_obs, &c->ig_inlt, &c->ig_outlt,
       &c->nvcycle, &c->nswpm, &c->nswpp, &c->isolve, &c->omega,
       &c->convergence);
    RD(6, "%lf %lf %lf %lf %lf %lf", &c->xl, &c->yl, &c->amuf, &c->prsc,
       &c->relxm, &c->relxp);
    RD(4, "%d %d %d %d", &c->ibc[0], &c->ibc[1], &c->ibc[2], &c->ibc[3]);
    for (int s = 0; s < 4; ++s)
        RD(4, "%lf %lf %lf %lf", &c->bc[s][0], &c->bc[s][1], &c->bc[s][2],
           &c->bc[s][3]);
    RD(4, "%lf %lf %lf %lf", &c->guess[0], &c->guess[1], &c->guess[2],
       &c->guess[3]);
    /* V6:357 reads the pattern file between the guess line and sc_obs */
    if ((rc = read_pattern(pattern, c)) != 0) {
        fclose(f);
        return rc;
    }
    RD(1, "%lf", &c->sc_obs);
    RD(1, "%d", &c->ninlt);
    if (c->ninlt < 0) c->ninlt = 0;
    c->inlt_j = c->ninlt ? (int *)calloc(2 * (size_t)c->ninlt, sizeof(int)) : NULL;
    c->inlt_v = c->ninlt ? (double *)calloc(3 * (size_t)c->ninlt, sizeof(double)) : NULL;
    for (int k = 0; k < c->ninlt; ++k)   /* return value unchecked: V6:373 */
        (void)!fscanf(f, " %d %d %lf %lf %lf", &c->inlt_j[2 * k],
                      &c->inlt_j[2 * k + 1], &c->inlt_v[3 * k],
                      &c->inlt_v[3 * k + 1], &c->inlt_v[3 * k + 2]);
    RD(1, "%d", &c->noutlt);
    if (c->noutlt < 0) c->noutlt = 0;
    c->outlt_j = c->noutlt ? (int *)calloc(2 * (size_t)c->noutlt, sizeof(int)) : NULL;
    c->outlt_v = c->noutlt ? (double *)calloc(3 * (size_t)c->noutlt, sizeof(double)) : NULL;
    for (int k = 0; k < c->noutlt; ++k)  /* V6:389 */
        (void)!fscanf(f, " %d %d %lf %lf %lf", &c->outlt_j[2 * k],
                      &c->outlt_j[2 * k + 1], &c->outlt_v[3 * k],
                      &c->outlt_v[3 * k + 1], &c->outlt_v[3 * k + 2]);
#undef RD
    fclose(f);
    c->dpdxm = 0.0;   /* never assigned in the reference: V6:33 */
    c->dpdym = 0.0;
    return 0;
}

void orc_config_free(orc_config *c)
{
    free(c->obs);
    free(c->inlt_j);
    free(c->inlt_v);
    free(c->outlt_j);
    free(c->outlt_v);
    free(c->inlet_u);
    memset(c, 0, sizeof *c);
}

/* ------------------------------------------------------------------ */
/* hierarchy: V6:402-506 (AllocateMemorySoA)                            */
/* ------------------------------------------------------------------ */

static double *dalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

static int build_levels(orc_solver *s)
{
    orc_config *c = &s->cfg;
    if (c->ngrid < 1) return fail(-4, "ngrid must be >= 1");
    if (c->nobs > 0 && (c->ig_obs < 1 || c->ig_obs > c->ngrid))
        return fail(-4, "ig_obs must be in [1, ngrid] when obstacles exist");
    if (c->ig_inlt > c->ngrid || c->ig_outlt > c->ngrid)
        return fail(-4, "ig_inlt/ig_outlt must be <= ngrid");
    s->L = (lvl_t *)calloc((size_t)c->ngrid, sizeof(lvl_t));
    for (int ig = 0; ig < c->ngrid; ++ig) {
        lvl_t *l = &s->L[ig];
        int fac = 1 << ig;
        l->npx = c->nfx / fac + 2;
        l->npy = c->nfy / fac + 2;
        size_t n = (size_t)l->npx * (size_t)l->npy;
        l->dx = c->xl / (l->npx - 2);
        l->dy = c->yl / (l->npy - 2);
        l->amu = c->amuf;
        for (int k = 0; k_dnames[k]; ++k) *darr(l, k_dnames[k]) = dalloc(n);
        l->tag = (int *)calloc(n, sizeof(int));
        int no = c->nobs > 0 ? c->nobs : 1;
        if (c->nobs > 0 && ig >= c->ig_obs) no = (l->npx - 2) * (l->npy - 2);
        l->isobs = (int *)calloc((size_t)no, sizeof(int));
        l->ieobs = (int *)calloc((size_t)no, sizeof(int));
        l->jsobs = (int *)calloc((size_t)no, sizeof(int));
        l->jeobs = (int *)calloc((size_t)no, sizeof(int));
        int ni = c->ninlt > 0 ? c->ninlt : 1, nq = c->noutlt > 0 ? c->noutlt : 1;
        l->jsinlt = (int *)calloc((size_t)ni, sizeof(int));
        l->jeinlt = (int *)calloc((size_t)ni, sizeof(int));
        l->jsoutlt = (int *)calloc((size_t)nq, sizeof(int));
        l->jeoutlt = (int *)calloc((size_t)nq, sizeof(int));
    }
    /* obstacle index scaling, V6:468-479: only levels ig < ig_obs */
    for (int k = 0; k < c->nobs; ++k) {
        const int *o = c->obs + 4 * k;
        for (int ig = 0; ig < c->ig_obs; ++ig) {
            int fac = 1 << (c->ig_obs - 1 - ig);
            lvl_t *l = &s->L[ig];
            l->isobs[k] = (o[0] + c->ibeg_obs - 1) * fac + 1;
            l->ieobs[k] = (o[1] + c->ibeg_obs) * fac;
            l->jsobs[k] = (o[2] - 1) * fac + 1;
            l->jeobs[k] = o[3] * fac;
            if (l->isobs[k] < 1 || l->ieobs[k] > l->npx - 2 ||
                l->jsobs[k] < 1 || l->jeobs[k] > l->npy - 2)
                return fail(-5, "obstacle outside the interior");
        }
    }
    for (int ig = 0; ig < c->ngrid && ig < c->ig_obs; ++ig) s->L[ig].nob = c->nobs;
    /* Levels below the obstacle level (ig >= ig_obs), where the reference is
     * undefined: a cell is an obstacle when all of its 2x2 children on the
     * next finer level are, listed one rectangle per cell, i-major then j
     * ascending.  The coarse obstacle region is thus a subset of the finer
     * one (blocking a coarse cell that is partly open would cut flow paths
     * the finer levels have and can leave inlet-fed dead ends, whose p'
     * equation is singular). */
    for (int ig = c->ig_obs; c->nobs > 0 && ig < c->ngrid; ++ig) {
        lvl_t *f = &s->L[ig - 1], *l = &s->L[ig];
        int nxf = f->npx, nyf = f->npy;
        unsigned char *blk = (unsigned char *)calloc((size_t)nxf * nyf, 1);
        for (int k = 0; k < f->nob; ++k)
            for (int j = f->jsobs[k]; j <= f->jeobs[k]; ++j)
                for (int i = f->isobs[k]; i <= f->ieobs[k]; ++i) blk[j * nxf + i] = 1;
        int n = 0;
        for (int ic = 1; ic <= l->npx - 2; ++ic)
            for (int jc = 1; jc <= l->npy - 2; ++jc) {
                int all = 1;
                for (int dj = -1; dj <= 0; ++dj)
                    for (int di = -1; di <= 0; ++di) {
                        int fi = 2 * ic + di, fj = 2 * jc + dj;
                        all &= (fi >= 1 && fi <= nxf - 2 && fj >= 1 && fj <= nyf - 2) &&
                               blk[fj * nxf + fi];
                    }
                if (all) {
                    l->isobs[n] = ic; l->ieobs[n] = ic;
                    l->jsobs[n] = jc; l->jeobs[n] = jc;
                    ++n;
                }
            }
        l->nob = n;
        l->ext = 1;
        free(blk);
    }
    for (int ig = 0; ig < c->ngrid; ++ig) s->L[ig].has_obs = s->L[ig].nob > 0;
    /* discrete inlets/outlets, V6:481-502 */
    for (int k = 0; k < c->ninlt; ++k)
        for (int ig = 0; ig < c->ig_inlt; ++ig) {
            int fac = 1 << (c->ig_inlt - 1 - ig);
            s->L[ig].jsinlt[k] = (c->inlt_j[2 * k] - 1) * fac + 1;
            s->L[ig].jeinlt[k] = c->inlt_j[2 * k + 1] * fac;
        }
    for (int k = 0; k < c->noutlt; ++k)
        for (int ig = 0; ig < c->ig_outlt; ++ig) {
            int fac = 1 << (c->ig_outlt - 1 - ig);
            s->L[ig].jsoutlt[k] = (c->outlt_j[2 * k] - 1) * fac + 1;
            s->L[ig].jeoutlt[k] = c->outlt_j[2 * k + 1] * fac;
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* initial / boundary state: V6:546-676 (init)                          */
/* ------------------------------------------------------------------ */

static void init_state(orc_solver *s)
{
    orc_config *c = &s->cfg;
    for (int ig = 0; ig < c->ngrid; ++ig) {
        lvl_t *l = &s->L[ig];
        for (int j = 0; j < l->npy; ++j)
            for (int i = 0; i < l->npx; ++i) {
                int ij = j * l->npx + i;
                l->u[ij] = c->guess[0];
                l->v[ij] = c->guess[1];
                l->sc[ij] = c->guess[3];
                l->p[ij] = c->guess[2];
                l->pp[ij] = 0.0; l->su[ij] = 0.0; l->c[ij] = 0.0;
                l->ap[ij] = 0.0; l->resu[ij] = 0.0; l->resv[ij] = 0.0;
                l->ressc[ij] = 0.0; l->rs[ij] = 0.0;
                l->tag[ij] = 1;
                l->apu[ij] = 1.0e30; l->apv[ij] = 1.0e30;
                l->spu[ij] = 0.0; l->spv[ij] = 0.0; l->spsc[ij] = 0.0;
                l->areax[ij] = l->dy; l->areay[ij] = l->dx;
                l->cx[ij] = 0.0; l->cy[ij] = 0.0;
            }
    }
    /* ghost values, V6:582-613 (xm, xp, ym, yp in this order) */
    for (int ig = 0; ig < c->ngrid; ++ig) {
        lvl_t *l = &s->L[ig];
        int nx = l->npx, ny = l->npy;
        for (int j = 0; j < ny; ++j) {
            int ij = j * nx;
            l->u[ij] = c->bc[0][0]; l->v[ij] = c->bc[0][1];
            l->p[ij] = c->bc[0][2]; l->sc[ij] = c->bc[0][3];
        }
        for (int j = 0; j < ny; ++j) {
            int ij = j * nx + nx - 1;
            l->u[ij - 1] = c->bc[1][0];   /* east boundary face, V6:593 */
            l->v[ij] = c->bc[1][1];
            l->p[ij] = c->bc[1][2]; l->sc[ij] = c->bc[1][3];
        }
        for (int i = 0; i < nx; ++i) {
            l->u[i] = c->bc[2][0]; l->v[i] = c->bc[2][1];
            l->p[i] = c->bc[2][2]; l->sc[i] = c->bc[2][3];
        }
        for (int i = 0; i < nx; ++i) {
            int ij = (ny - 1) * nx + i;
            l->u[ij] = c->bc[3][0]; l->v[ij] = c->bc[3][1];
            l->p[ij] = c->bc[3][2]; l->sc[ij] = c->bc[3][3];
        }
    }
    /* obstacles, V6:617-641 */
    for (int ig = 0; ig < c->ngrid; ++ig) {
        lvl_t *l = &s->L[ig];
        if (!l->has_obs) continue;
        int nx = l->npx;
        for (int k = 0; k < l->nob; ++k)
            for (int j = l->jsobs[k]; j <= l->jeobs[k]; ++j)
                for (int i = l->isobs[k]; i <= l->ieobs[k]; ++i) {
                    int ij = j * nx + i;
                    l->u[ij] = 0.0; l->v[ij] = 0.0;
                    l->u[ij - 1] = 0.0; l->v[ij - nx] = 0.0;
                    l->spu[ij] = 1.0e30; l->spv[ij] = 1.0e30;
                    l->spu[ij - 1] = 1.0e30; l->spv[ij - nx] = 1.0e30;
                    l->spsc[ij] = 1.0e30;
                    l->tag[ij] = 0;
                    l->sc[ij] = c->sc_obs;
                }
    }
    /* discrete inlets, V6:645-658 */
    for (int ig = 0; ig < c->ig_inlt; ++ig) {
        lvl_t *l = &s->L[ig];
        for (int k = 0; k < c->ninlt; ++k)
            for (int j = l->jsinlt[k]; j <= l->jeinlt[k]; ++j) {
                int ij = j * l->npx;
                l->u[ij] = c->inlt_v[3 * k];
                l->v[ij] = c->inlt_v[3 * k + 1];
                l->sc[ij] = c->inlt_v[3 * k + 2];
            }
    }
    /* discrete outlets, V6:661-674 */
    for (int ig = 0; ig < c->ig_outlt; ++ig) {
        lvl_t *l = &s->L[ig];
        for (int k = 0; k < c->noutlt; ++k)
            for (int j = l->jsoutlt[k]; j <= l->jeoutlt[k]; ++j) {
                int ij = j * l->npx + l->npx - 2;
                l->u[ij] = c->outlt_v[3 * k];
                l->v[ij] = c->outlt_v[3 * k + 1];
                l->sc[ij + 1] = c->outlt_v[3 * k + 2];
            }
    }
    /* optional level-0 inlet profile (extension; rows 1..nfy) */
    if (c->inlet_u) {
        lvl_t *l = &s->L[0];
        for (int j = 1; j <= l->npy - 2; ++j) l->u[j * l->npx] = c->inlet_u[j - 1];
    }
}

/* ------------------------------------------------------------------ */
/* coarse geometry + coarse boundary values: V6:1391-1473 (rest_geom)  */
/* ------------------------------------------------------------------ */

static void rest_geom(orc_solver *s)
{
    int ng = s->cfg.ngrid;
    for (int ig = 1; ig < ng; ++ig) {
        lvl_t *f = &s->L[ig - 1], *c = &s->L[ig];
        int nxf = f->npx, nxc = c->npx, nyc = c->npy;
        for (int jc = 1; jc < nyc - 1; ++jc)
            for (int ic = 0; ic < nxc - 1; ++ic) {
                int ij = (2 * jc) * nxf + 2 * ic;
                c->areax[jc * nxc + ic] = f->areax[ij] + f->areax[ij - nxf];
            }
        for (int jc = 0; jc < nyc - 1; ++jc)
            for (int ic = 1; ic < nxc - 1; ++ic) {
                int ij = (2 * jc) * nxf + 2 * ic;
                c->areay[jc * nxc + ic] = f->areay[ij] + f->areay[ij - 1];
            }
    }
    for (int ig = 1; ig < ng; ++ig) {            /* inlet column, V6:1431-1447 */
        lvl_t *f = &s->L[ig - 1], *c = &s->L[ig];
        int nxf = f->npx, nxc = c->npx, nyc = c->npy;
        for (int jc = 1; jc < nyc - 1; ++jc) {
            int ijc = jc * nxc, ij = (2 * jc) * nxf;
            c->u[ijc] = 0.5 * (f->u[ij] + f->u[ij - nxf]);
            c->v[ijc] = 0.5 * (f->v[ij] + f->v[ij - 1]);  /* wraps to row end */
        }
    }
    for (int ig = 1; ig < ng; ++ig) {            /* outlet face, V6:1452-1471 */
        lvl_t *f = &s->L[ig - 1], *c = &s->L[ig];
        int nxf = f->npx, nxc = c->npx, nyc = c->npy;
        for (int jc = 1; jc < nyc - 1; ++jc) {
            int ijc = jc * nxc + (nxc - 2);
            int ij = (2 * jc) * nxf + 2 * (nxc - 2);
            c->u[ijc] = 0.5 * (f->u[ij] + f->u[ij - nxf]);
            c->v[ijc] = 0.5 * (f->v[ij] + f->v[ij - 1]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* SIMPLE relaxation pieces                                             */
/* ------------------------------------------------------------------ */

/* V6:678-694 */
static void fluxes(lvl_t *l)
{
    int nx = l->npx, ny = l->npy;
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            int ij = j * nx + i;
            l->cx[ij] = (0.5 * l->u[ij]) * l->areax[ij];
            l->cy[ij] = (0.5 * l->v[ij]) * l->areay[ij];
        }
}

/* V6:815-860 */
static void coefu(lvl_t *l)
{
    int nx = l->npx, ny = l->npy;
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 2; ++i) {
            int ij = j * nx + i, ijs = ij - nx;
            double cxm = 0.5 * (l->cx[ij] + l->cx[ij - 1]);
            double cxp = 0.5 * (l->cx[ij] + l->cx[ij + 1]);
            double cym = 0.5 * (l->cy[ijs] + l->cy[ijs + 1]);
            double cyp = 0.5 * (l->cy[ij] + l->cy[ij + 1]);
            double dxm = l->amu * (l->areax[ij] / l->dx);
            double dym = l->amu * (l->areay[ij] / l->dy);
            l->aw[ij] = fmax(fabs(cxm), dxm) + cxm;
            l->ae[ij] = fmax(fabs(cxp), dxm) - cxp;
            l->as[ij] = fmax(fabs(cym), dym) + cym;
            l->an[ij] = fmax(fabs(cyp), dym) - cyp;
            l->ap[ij] = ((l->aw[ij] + l->ae[ij]) + l->as[ij]) + l->an[ij];
        }
    for (int i = 1; i < nx - 2; ++i) {           /* no-slip walls, V6:849-858 */
        int ijn = (ny - 2) * nx + i, ijs = nx + i;
        double add = l->amu * (l->areay[ijn] / l->dy);
        l->an[ijn] += add; l->ap[ijn] += add;
        l->as[ijs] += add; l->ap[ijs] += add;
    }
}

/* V6:741-767 */
static void umom(lvl_t *l, const orc_config *c)
{
    double relxm1 = 1.0 - c->relxm;
    int nx = l->npx, ny = l->npy;
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 2; ++i) {
            int ij = j * nx + i;
            l->su[ij] = l->areax[ij] * (l->p[ij] - l->p[ij + 1]);
            l->ap[ij] = l->ap[ij] / c->relxm;
            l->su[ij] = ((l->su[ij] + (relxm1 * l->ap[ij]) * l->u[ij]) + l->resu[ij]) +
                        (c->dpdxm * l->areax[ij]) * l->areay[ij];
            l->ap[ij] = l->ap[ij] + l->spu[ij];
            l->apu[ij] = l->ap[ij];
        }
}

/* V6:966-1016 */
static void coefv(lvl_t *l)
{
    int nx = l->npx, ny = l->npy;
    for (int j = 1; j < ny - 2; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = j * nx + i, ijn = ij + nx, ijs = ij - nx;
            double cxm = 0.5 * (l->cx[ij - 1] + l->cx[ijn - 1]);
            double cxp = 0.5 * (l->cx[ij] + l->cx[ijn]);
            double cym = 0.5 * (l->cy[ij] + l->cy[ijs]);
            double cyp = 0.5 * (l->cy[ij] + l->cy[ijn]);
            double dxm = l->amu * (l->areax[ij] / l->dx);
            double dym = l->amu * (l->areay[ij] / l->dy);
            double dxp = dxm, dyp = dym;  /* same expressions in V6:989-990 */
            l->aw[ij] = fmax(fabs(cxm), dxm) + cxm;
            l->ae[ij] = fmax(fabs(cxp), dxp) - cxp;
            l->as[ij] = fmax(fabs(cym), dym) + cym;
            l->an[ij] = fmax(fabs(cyp), dyp) - cyp;
            l->ap[ij] = ((l->aw[ij] + l->ae[ij]) + l->as[ij]) + l->an[ij];
        }
    for (int j = 1; j < ny - 2; ++j) {           /* side walls, V6:1004-1015 */
        int ije = j * nx + (nx - 2), ijw = j * nx + 1;
        double add = l->amu * (l->areay[ije] / l->dy);
        l->ae[ije] += add; l->ap[ije] += add;
        l->aw[ijw] += add; l->ap[ijw] += add;
    }
}

/* V6:769-793 */
static void vmom(lvl_t *l, const orc_config *c)
{
    double relxm1 = 1.0 - c->relxm;
    int nx = l->npx, ny = l->npy;
    for (int j = 1; j < ny - 2; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = j * nx + i;
            l->su[ij] = l->areay[ij] * (l->p[ij] - l->p[ij + nx]);
            l->ap[ij] = l->ap[ij] / c->relxm;
            l->su[ij] = ((l->su[ij] + (relxm1 * l->ap[ij]) * l->v[ij]) + l->resv[ij]) +
                        (c->dpdym * l->areax[ij]) * l->areay[ij];
            l->ap[ij] = l->ap[ij] + l->spv[ij];
            l->apv[ij] = l->ap[ij];
        }
}

/* V6:861-964 including the host-side obstacle adds V6:939-963 */
static void coefsc(lvl_t *l, const orc_config *c)
{
    int nx = l->npx, ny = l->npy;
    double amusc = c->amuf / c->prsc;
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = j * nx + i, ijs = ij - nx;
            double cxm = l->cx[ij - 1], cxp = l->cx[ij];
            double cym = l->cy[ijs], cyp = l->cy[ij];
            double dxm = amusc * (l->areax[ij - 1] / l->dx);
            double dym = amusc * (l->areay[ijs] / l->dy);
            double dxp = amusc * (l->areax[ij] / l->dx);
            double dyp = amusc * (l->areay[ij] / l->dy);
            l->aw[ij] = fmax(fabs(cxm), dxm) + cxm;
            l->ae[ij] = fmax(fabs(cxp), dxp) - cxp;
            l->as[ij] = fmax(fabs(cym), dym) + cym;
            l->an[ij] = fmax(fabs(cyp), dyp) - cyp;
            l->ap[ij] = ((l->aw[ij] + l->ae[ij]) + l->as[ij]) + l->an[ij];
        }
    if (!l->has_obs) return;
    for (int k = 0; k < l->nob; ++k)             /* i outer, j inner */
        for (int i = l->isobs[k]; i <= l->ieobs[k]; ++i)
            for (int j = l->jsobs[k]; j <= l->jeobs[k]; ++j) {
                int ij = j * nx + i;
                double add = amusc * (l->areax[ij] / l->dx);
                l->ae[ij - 1] += add; l->ap[ij - 1] += add;
                l->aw[ij + 1] += add; l->ap[ij + 1] += add;
                add = amusc * (l->areay[ij] / l->dy);
                l->an[ij - nx] += add; l->ap[ij - nx] += add;
                l->as[ij + nx] += add; l->ap[ij + nx] += add;
            }
}

/* V6:795-812 */
static void scalar(lvl_t *l, const orc_config *c)
{
    int nx = l->npx, ny = l->npy;
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = j * nx + i;
            l->su[ij] = l->spsc[ij] * c->sc_obs + l->ressc[ij];
            l->ap[ij] = l->ap[ij] + l->spsc[ij];
        }
}

/* V6:1018-1092 */
static void coefp(lvl_t *l, const orc_config *c)
{
    int nx = l->npx, ny = l->npy;
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = j * nx + i, ijs = ij - nx;
            l->aw[ij] = (l->areax[ij - 1] * l->areax[ij - 1]) / l->apu[ij - 1];
            l->ae[ij] = (l->areax[ij] * l->areax[ij]) / l->apu[ij];
            l->as[ij] = (l->areay[ijs] * l->areay[ijs]) / l->apv[ijs];
            l->an[ij] = (l->areay[ij] * l->areay[ij]) / l->apv[ij];
            l->ap[ij] = (((l->aw[ij] + l->ae[ij]) + l->as[ij]) + l->an[ij]) +
                        (double)(1 - l->tag[ij]) * 1.0e30;
        }
    const double diag = 1.0e30;
    if (c->ibc[0] == 2)
        for (int j = 1; j < ny - 1; ++j) { int ij = j * nx + 1; l->ap[ij] = l->ap[ij] + diag; l->su[ij] = 0.0; }
    if (c->ibc[1] == 2)
        for (int j = 1; j < ny - 1; ++j) { int ij = j * nx + nx - 2; l->ap[ij] = l->ap[ij] + diag; l->su[ij] = 0.0; }
    if (c->ibc[2] == 2)
        for (int i = 1; i < nx - 1; ++i) { int ij = nx + i; l->ap[ij] = l->ap[ij] + diag; l->su[ij] = 0.0; }
    if (c->ibc[3] == 2)
        for (int i = 1; i < nx - 1; ++i) { int ij = (ny - 2) * nx + i; l->ap[ij] = l->ap[ij] + diag; l->su[ij] = 0.0; }
}

/* V6:1096-1125; returns the level's L1 continuity defect */
static double masserror(lvl_t *l)
{
    int nx = l->npx, ny = l->npy;
    double err = 0.0;
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = j * nx + i, ijs = ij - nx;
            double div = (((l->areax[ij] * l->u[ij] - l->areax[ij] * l->u[ij - 1]) +
                           l->areay[ij] * l->v[ij]) - l->areay[ij] * l->v[ijs]);
            l->su[ij] = -div;
            l->pp[ij] = 0.0;
            err += fabs(div);
        }
    return err;
}

/* V6:1127-1178, including the host-side obstacle-edge p copies */
static void update(lvl_t *l, const orc_config *c)
{
    int nx = l->npx, ny = l->npy;
    const double ppref = 0.0;
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = j * nx + i;
            l->u[ij] = l->u[ij] + (l->areax[ij] * (l->pp[ij] - l->pp[ij + 1])) / l->apu[ij];
            l->v[ij] = l->v[ij] + (l->areay[ij] * (l->pp[ij] - l->pp[ij + nx])) / l->apv[ij];
            l->p[ij] = l->p[ij] + c->relxp * (l->pp[ij] - ppref);
        }
    if (!l->has_obs) return;
    for (int k = 0; k < l->nob; ++k) {
        int j = l->jsobs[k];
        for (int i = l->isobs[k]; i <= l->ieobs[k]; ++i) l->p[j * nx + i] = l->p[j * nx + i - nx];
        j = l->jeobs[k];
        for (int i = l->isobs[k]; i <= l->ieobs[k]; ++i) l->p[j * nx + i] = l->p[j * nx + i + nx];
        int i = l->isobs[k];
        for (int jj = l->jsobs[k]; jj <= l->jeobs[k]; ++jj) l->p[jj * nx + i] = l->p[jj * nx + i - 1];
        i = l->ieobs[k];
        for (int jj = l->jsobs[k]; jj <= l->jeobs[k]; ++jj) l->p[jj * nx + i] = l->p[jj * nx + i + 1];
    }
}

/* V6:1179-1210 (ibc == 3; the y-branch index expression is kept as written) */
static void pbound(lvl_t *l, const orc_config *c)
{
    int nx = l->npx, ny = l->npy;
    if (c->ibc[0] == 3)
        for (int j = 1; j < ny - 1; ++j) {
            int ij = j * nx, ij1 = ij + nx - 3;
            l->u[ij] = l->u[ij1]; l->u[ij1 + 1] = l->u[ij + 1];
            l->v[ij] = l->v[ij1]; l->v[ij1 + 1] = l->v[ij + 1];
        }
    if (c->ibc[2] == 3)
        for (int i = 1; i < nx - 1; ++i) {
            int ij = i, ij1 = ij + ny - 3;
            l->u[ij] = l->u[ij1]; l->u[ij1 + nx] = l->u[ij + nx];
            l->v[ij] = l->v[ij1]; l->v[ij1 + nx] = l->v[ij + nx];
        }
}

/* V6:1211-1228 (ibc_xp == 4) */
static void outflow(lvl_t *l, const orc_config *c)
{
    int nx = l->npx, ny = l->npy;
    if (c->ibc[1] != 4) return;
    for (int j = 0; j < ny; ++j) {
        int ij = j * nx + nx - 1;
        l->u[ij - 1] = l->u[ij - 2];
        l->v[ij] = l->v[ij - 1];
        l->p[ij] = l->p[ij - 1];
        l->sc[ij] = l->sc[ij - 1];
    }
}

/* Jacobi, V6:1231-1296.  wj is 1 whatever the field (the weight expression
 * at V6:1241 is a discarded statement). */
static void jacobi(lvl_t *l, double *phi, int nx, int ny, int nsw)
{
    int npx = l->npx;
    size_t n = (size_t)l->npx * (size_t)l->npy;
    const double wj = 1.0;
    double *phi_new = (double *)malloc(n * sizeof(double));
    memcpy(phi_new, phi, n * sizeof(double));
    double *cur = phi, *next = phi_new;
    for (int it = 0; it < nsw; ++it) {
        for (int j = 1; j < ny; ++j)
            for (int i = 1; i < nx; ++i) {
                int ij = j * npx + i;
                double sum_nb = ((l->aw[ij] * cur[ij - 1] + l->ae[ij] * cur[ij + 1]) +
                                 l->as[ij] * cur[ij - npx]) + l->an[ij] * cur[ij + npx];
                double g = (sum_nb + l->su[ij]) / (l->ap[ij] + 1.0e-30);
                next[ij] = (1.0 - wj) * cur[ij] + wj * g;
            }
        double *t = cur; cur = next; next = t;
    }
    if (cur != phi) memcpy(phi, phi_new, n * sizeof(double));
    free(phi_new);
}

/* red-black Gauss-Seidel, V6:1297-1351 */
static void rbgs(lvl_t *l, double *phi, int nx, int ny, int nsw)
{
    int npx = l->npx;
    for (int it = 0; it < nsw; ++it)
        for (int colour = 0; colour < 2; ++colour)
            for (int j = 1; j < ny; ++j)
                for (int i = 1; i < nx; ++i) {
                    if ((i + j) % 2 != colour) continue;
                    int ij = j * npx + i;
                    double sum_nb = ((l->aw[ij] * phi[ij - 1] + l->ae[ij] * phi[ij + 1]) +
                                     l->as[ij] * phi[ij - npx]) + l->an[ij] * phi[ij + npx];
                    phi[ij] = (sum_nb + l->su[ij]) / l->ap[ij];
                }
}

/* Lexicographic Gauss-Seidel: NOT a routine of the reference.  The labelled
 * variant of isolve 2 the GPU library offers: solvergs's loops (V6:1308-1350)
 * without the colour tests, i.e. one in-place pass per sweep with j outer
 * and i inner, each point updated with solvergs's expression.  Pinned by the
 * mathematics only (tests/test_oracle_math.py: fixed point = the dense
 * solve; the last point's equation holds after a sweep). */
static void lexgs(lvl_t *l, double *phi, int nx, int ny, int nsw)
{
    int npx = l->npx;
    for (int it = 0; it < nsw; ++it)
        for (int j = 1; j < ny; ++j)
            for (int i = 1; i < nx; ++i) {
                int ij = j * npx + i;
                double sum_nb = ((l->aw[ij] * phi[ij - 1] + l->ae[ij] * phi[ij + 1]) +
                                 l->as[ij] * phi[ij - npx]) + l->an[ij] * phi[ij + npx];
                phi[ij] = (sum_nb + l->su[ij]) / l->ap[ij];
            }
}

static void smooth(orc_solver *s, lvl_t *l, double *phi, int nx, int ny)
{
    int nsw = (phi == l->pp) ? s->cfg.nswpp : s->cfg.nswpm;   /* V6:1238 */
    if (s->cfg.isolve == 1) jacobi(l, phi, nx, ny, nsw);
    if (s->cfg.isolve == 2) {
        if (s->gs_lex)
            lexgs(l, phi, nx, ny, nsw);
        else
            rbgs(l, phi, nx, ny, nsw);
    }
}

void orc_set_gs_lex(orc_solver *s, int lex) { s->gs_lex = lex ? 1 : 0; }

/* V6:696-737 */
static void solve_u(orc_solver *s, lvl_t *l)
{ coefu(l); umom(l, &s->cfg); smooth(s, l, l->u, l->npx - 2, l->npy - 1); }

static void solve_v(orc_solver *s, lvl_t *l)
{ coefv(l); vmom(l, &s->cfg); smooth(s, l, l->v, l->npx - 1, l->npy - 2); }

static double solve_pprime(orc_solver *s, lvl_t *l)
{
    double err = masserror(l);
    coefp(l, &s->cfg);
    smooth(s, l, l->pp, l->npx - 1, l->npy - 1);
    update(l, &s->cfg);
    return err;
}

static void solve_sc(orc_solver *s, lvl_t *l)
{ coefsc(l, &s->cfg); scalar(l, &s->cfg); smooth(s, l, l->sc, l->npx - 1, l->npy - 1); }

/* ------------------------------------------------------------------ */
/* restriction: V6:1354-1389 (restv), 1474-1740 (restru/restrsc/restrv) */
/* ------------------------------------------------------------------ */

static void restv(lvl_t *f, lvl_t *c)
{
    int nxf = f->npx, nxc = c->npx, nyc = c->npy;
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) {
            int ijc = jc * nxc + ic;
            int ij = (2 * jc) * nxf + 2 * ic, ijw = ij - 1, ijs = ij - nxf;
            c->u[ijc] = (f->u[ij] * f->areax[ij] + f->u[ijs] * f->areax[ijs]) /
                        (c->areax[ijc] + 1.0e-30);
            c->v[ijc] = (f->v[ij] * f->areay[ij] + f->v[ijw] * f->areay[ijw]) /
                        (c->areay[ijc] + 1.0e-30);
            c->p[ijc] = 0.25 * (((f->p[ij] + f->p[ijw]) + f->p[ijs]) + f->p[ijs - 1]);
            c->sc[ijc] = 0.25 * (((f->sc[ij] + f->sc[ijw]) + f->sc[ijs]) + f->sc[ijs - 1]);
        }
}

static void zero(double *a, size_t n) { for (size_t k = 0; k < n; ++k) a[k] = 0.0; }

/* Restriction onto a level below the obstacle level (ig >= ig_obs, where
 * the reference is undefined): the fine residual is evaluated with
 * un-penalised coefficients (V6:1500-1508), so at penalised faces/cells it
 * never vanishes -- there it is balanced by the 1e30 penalty.  It is zeroed
 * at penalised points before it enters the coarse tau term, so a converged
 * fine state is a V-cycle fixed point.  Restrictions between levels above
 * ig_obs (all of them when ig_obs == ngrid) keep the reference's
 * arithmetic exactly. */
static int penalty_mask(const lvl_t *f, const lvl_t *c) { (void)f; return c->ext; }

static void restru(lvl_t *f, lvl_t *c, const orc_config *cf)
{
    const int mask = penalty_mask(f, c);
    int nxf = f->npx, nyf = f->npy, nxc = c->npx, nyc = c->npy;
    coefu(f);
    zero(f->rs, (size_t)nxf * nyf);
    for (int j = 1; j < nyf - 1; ++j)
        for (int i = 1; i < nxf - 2; ++i) {
            int ij = j * nxf + i;
            f->rs[ij] = ((((((f->aw[ij] * f->u[ij - 1] + f->ae[ij] * f->u[ij + 1]) +
                             f->as[ij] * f->u[ij - nxf]) + f->an[ij] * f->u[ij + nxf]) -
                           f->ap[ij] * f->u[ij]) +
                          f->areax[ij] * (f->p[ij] - f->p[ij + 1])) + f->resu[ij]) +
                        (cf->dpdxm * f->areax[ij]) * f->areay[ij];
            if (f->areax[ij] == 0.0) f->rs[ij] = 0.0;
            if (mask && f->spu[ij] > 1.0e10) f->rs[ij] = 0.0;
        }
    coefu(c);
    zero(c->rs, (size_t)nxc * nyc);
    for (int j = 1; j < nyc - 1; ++j)
        for (int i = 1; i < nxc - 2; ++i) {
            int ij = j * nxc + i;
            c->rs[ij] = (((((c->aw[ij] * c->u[ij - 1] + c->ae[ij] * c->u[ij + 1]) +
                            c->as[ij] * c->u[ij - nxc]) + c->an[ij] * c->u[ij + nxc]) -
                          c->ap[ij] * c->u[ij]) +
                         c->areax[ij] * (c->p[ij] - c->p[ij + 1])) +
                        (cf->dpdxm * c->areax[ij]) * c->areay[ij];
        }
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 2; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            c->resu[ijc] = ((f->rs[ij] + f->rs[ij - nxf]) +
                            0.5 * (((f->rs[ij - 1] + f->rs[ij - 1 - nxf]) + f->rs[ij + 1]) +
                                   f->rs[ij + 1 - nxf])) - c->rs[ijc];
        }
}

static void restrv(lvl_t *f, lvl_t *c, const orc_config *cf)
{
    const int mask = penalty_mask(f, c);
    int nxf = f->npx, nyf = f->npy, nxc = c->npx, nyc = c->npy;
    coefv(f);
    zero(f->rs, (size_t)nxf * nyf);
    for (int j = 1; j < nyf - 2; ++j)
        for (int i = 1; i < nxf - 1; ++i) {
            int ij = j * nxf + i;
            f->rs[ij] = ((((((f->aw[ij] * f->v[ij - 1] + f->ae[ij] * f->v[ij + 1]) +
                             f->as[ij] * f->v[ij - nxf]) + f->an[ij] * f->v[ij + nxf]) -
                           f->ap[ij] * f->v[ij]) +
                          f->areay[ij] * (f->p[ij] - f->p[ij + nxf])) + f->resv[ij]) +
                        (cf->dpdym * f->areax[ij]) * f->areay[ij];
            if (f->areay[ij] == 0.0) f->rs[ij] = 0.0;
            if (mask && f->spv[ij] > 1.0e10) f->rs[ij] = 0.0;
        }
    coefv(c);
    zero(c->rs, (size_t)nxc * nyc);
    for (int j = 1; j < nyc - 2; ++j)
        for (int i = 1; i < nxc - 1; ++i) {
            int ij = j * nxc + i;
            c->rs[ij] = (((((c->aw[ij] * c->v[ij - 1] + c->ae[ij] * c->v[ij + 1]) +
                            c->as[ij] * c->v[ij - nxc]) + c->an[ij] * c->v[ij + nxc]) -
                          c->ap[ij] * c->v[ij]) +
                         c->areay[ij] * (c->p[ij] - c->p[ij + nxc])) +
                        (cf->dpdym * c->areax[ij]) * c->areay[ij];
        }
    for (int jc = 1; jc < nyc - 2; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            c->resv[ijc] = ((f->rs[ij] + f->rs[ij - 1]) +
                            0.5 * (((f->rs[ij - nxf] + f->rs[ij - nxf - 1]) + f->rs[ij + nxf]) +
                                   f->rs[ij + nxf - 1])) - c->rs[ijc];
        }
}

static void restrsc(lvl_t *f, lvl_t *c, const orc_config *cf)
{
    const int mask = penalty_mask(f, c);
    int nxf = f->npx, nyf = f->npy, nxc = c->npx, nyc = c->npy;
    coefsc(f, cf);
    zero(f->rs, (size_t)nxf * nyf);
    for (int j = 1; j < nyf - 1; ++j)
        for (int i = 1; i < nxf - 1; ++i) {
            int ij = j * nxf + i;
            f->rs[ij] = (((((f->aw[ij] * f->sc[ij - 1] + f->ae[ij] * f->sc[ij + 1]) +
                            f->as[ij] * f->sc[ij - nxf]) + f->an[ij] * f->sc[ij + nxf]) -
                          f->ap[ij] * f->sc[ij]) + f->ressc[ij]);
            if (mask && f->spsc[ij] > 1.0e10) f->rs[ij] = 0.0;
        }
    coefsc(c, cf);
    zero(c->rs, (size_t)nxc * nyc);
    for (int j = 1; j < nyc - 1; ++j)
        for (int i = 1; i < nxc - 1; ++i) {
            int ij = j * nxc + i;
            c->rs[ij] = ((((c->aw[ij] * c->sc[ij - 1] + c->ae[ij] * c->sc[ij + 1]) +
                           c->as[ij] * c->sc[ij - nxc]) + c->an[ij] * c->sc[ij + nxc]) -
                         c->ap[ij] * c->sc[ij]);
        }
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            c->ressc[ijc] = (((f->rs[ij] + f->rs[ij - nxf]) + f->rs[ij - 1]) +
                             f->rs[ij - 1 - nxf]) - c->rs[ijc];
        }
}

/* ------------------------------------------------------------------ */
/* prolongation: V6:1745-2028                                           */
/* ------------------------------------------------------------------ */

static void prolu(lvl_t *f, lvl_t *c)
{
    int nxf = f->npx, nxc = c->npx, nyc = c->npy;
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) c->c[jc * nxc + ic] = 0.0;
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 2; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            double Ru = (f->u[ij] * f->areax[ij] + f->u[ij - nxf] * f->areax[ij - nxf]) /
                        (c->areax[ijc] + 1.0E-30);
            c->c[ijc] = c->u[ijc] - Ru;
            if (c->ext && c->spu[ijc] > 1.0e10) c->c[ijc] = 0.0;
        }
    const double f1 = 3. / 4., f2 = 1. / 4.;
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 2; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            double corr1 = f1 * c->c[ijc] + f2 * c->c[ijc + nxc];
            double corr2 = f1 * c->c[ijc] + f2 * c->c[ijc - nxc];
            if (f->spu[ij] > 1.0e10) { corr1 = 0.0; corr2 = 0.0; }
            f->u[ij] += corr1;
            f->u[ij - nxf] += corr2;
            f->u[ij - 1] += 0.5 * corr1;
            f->u[ij - nxf - 1] += 0.5 * corr2;
            f->u[ij + 1] += 0.5 * corr1;
            f->u[ij - nxf + 1] += 0.5 * corr2;
        }
}

static void prolv(lvl_t *f, lvl_t *c)
{
    int nxf = f->npx, nxc = c->npx, nyc = c->npy;
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) c->c[jc * nxc + ic] = 0.0;
    for (int jc = 1; jc < nyc - 2; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            double Rv = (f->v[ij] * f->areay[ij] + f->v[ij - 1] * f->areay[ij - 1]) /
                        (c->areay[ijc] + 1.0E-30);
            c->c[ijc] = c->v[ijc] - Rv;
            if (c->ext && c->spv[ijc] > 1.0e10) c->c[ijc] = 0.0;
        }
    const double f1 = 3. / 4., f2 = 1. / 4.;
    for (int jc = 1; jc < nyc - 2; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            double corr1 = f1 * c->c[ijc] + f2 * c->c[ijc + 1];
            double corr2 = f1 * c->c[ijc] + f2 * c->c[ijc - 1];
            if (f->spv[ij] > 1.0e10) { corr1 = 0.0; corr2 = 0.0; }
            f->v[ij] += corr1;
            f->v[ij - 1] += corr2;
            f->v[ij - nxf] += 0.5 * corr1;
            f->v[ij - 1 - nxf] += 0.5 * corr2;
            f->v[ij + nxf] += 0.5 * corr1;
            f->v[ij - 1 + nxf] += 0.5 * corr2;
        }
}

static void prol_cell(lvl_t *f, lvl_t *c, double *ff, const double *fc, int mask)
{
    int nxf = f->npx, nxc = c->npx, nyc = c->npy;
    for (int k = 0; k < nxc * nyc; ++k) c->c[k] = 0.0;
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic, ijs = ij - nxf;
            double R = 0.25 * (((ff[ij] + ff[ij - 1]) + ff[ijs]) + ff[ijs - 1]);
            c->c[ijc] = fc[ijc] - R;
            if (mask && f->spu[ij] > 1.0e10) c->c[ijc] = 0.0;
            if (c->ext && c->spsc[ijc] > 1.0e10) c->c[ijc] = 0.0;
        }
    const double f1 = 9. / 16., f2 = 3. / 16., f3 = 3. / 16., f4 = 1. / 16.;
    const double *cc = c->c;
    for (int jc = 1; jc < nyc - 1; ++jc)
        for (int ic = 1; ic < nxc - 1; ++ic) {
            int ijc = jc * nxc + ic, ij = (2 * jc) * nxf + 2 * ic;
            ff[ij] += ((f1 * cc[ijc] + f2 * cc[ijc + 1]) + f3 * cc[ijc + nxc]) + f4 * cc[ijc + 1 + nxc];
            ff[ij - 1] += ((f1 * cc[ijc] + f2 * cc[ijc - 1]) + f3 * cc[ijc + nxc]) + f4 * cc[ijc - 1 + nxc];
            ff[ij - nxf] += ((f1 * cc[ijc] + f2 * cc[ijc + 1]) + f3 * cc[ijc - nxc]) + f4 * cc[ijc + 1 - nxc];
            ff[ij - nxf - 1] += ((f1 * cc[ijc] + f2 * cc[ijc - 1]) + f3 * cc[ijc - nxc]) + f4 * cc[ijc - 1 - nxc];
        }
}

/* V6:1894-1961 */
static void prolp(lvl_t *f, lvl_t *c) { prol_cell(f, c, f->p, c->p, 1); }
/* V6:1962-2028 */
static void prolsc(lvl_t *f, lvl_t *c) { prol_cell(f, c, f->sc, c->sc, 0); }

/* ------------------------------------------------------------------ */
/* public entry points                                                  */
/* ------------------------------------------------------------------ */

int orc_create(const orc_config *cfg, orc_solver **out)
{
    *out = NULL;
    orc_solver *s = (orc_solver *)calloc(1, sizeof *s);
    s->cfg = *cfg;
    /* deep-copy the arrays so the caller keeps ownership of cfg */
    orc_config *c = &s->cfg;
#define DUP(p, n, T)                                                     \
    do {                                                                 \
        if (cfg->p) {                                                    \
            c->p = (T *)malloc(sizeof(T) * (size_t)(n));                 \
            memcpy(c->p, cfg->p, sizeof(T) * (size_t)(n));               \
        }                                                                \
    } while (0)
    DUP(obs, 4 * cfg->nobs, int);
    DUP(inlt_j, 2 * cfg->ninlt, int);
    DUP(inlt_v, 3 * cfg->ninlt, double);
    DUP(outlt_j, 2 * cfg->noutlt, int);
    DUP(outlt_v, 3 * cfg->noutlt, double);
    DUP(inlet_u, cfg->nfy, double);
#undef DUP
    int rc = build_levels(s);
    if (rc) { orc_destroy(s); return rc; }
    init_state(s);
    rest_geom(s);
    s->hist_len = c->nvcycle > 0 ? c->nvcycle : 1;
    s->hist = dalloc((size_t)s->hist_len);
    *out = s;
    return 0;
}

int orc_create_from_files(const char *inputdata, const char *pattern, orc_solver **out)
{
    orc_config c;
    int rc = orc_parse_files(inputdata, pattern, &c);
    if (rc) { orc_config_free(&c); *out = NULL; return rc; }
    rc = orc_create(&c, out);
    orc_config_free(&c);
    return rc;
}

/* the same, with a level-0 inlet u profile (nfy values, rows 1..nfy) set
 * after init and before rest_geom carries it to the coarse levels */
int orc_create_from_files_inlet(const char *inputdata, const char *pattern,
                                const double *inlet_u, orc_solver **out)
{
    orc_config c;
    int rc = orc_parse_files(inputdata, pattern, &c);
    if (rc) { orc_config_free(&c); *out = NULL; return rc; }
    if (inlet_u) {
        c.inlet_u = (double *)malloc(sizeof(double) * (size_t)c.nfy);
        if (!c.inlet_u) { orc_config_free(&c); *out = NULL; return -1; }
        memcpy(c.inlet_u, inlet_u, sizeof(double) * (size_t)c.nfy);
    }
    rc = orc_create(&c, out);
    orc_config_free(&c);
    return rc;
}

void orc_destroy(orc_solver *s)
{
    if (!s) return;
    if (s->L) {
        for (int ig = 0; ig < s->cfg.ngrid; ++ig) {
            lvl_t *l = &s->L[ig];
            for (int k = 0; k_dnames[k]; ++k) free(*darr(l, k_dnames[k]));
            free(l->tag); free(l->isobs); free(l->ieobs); free(l->jsobs); free(l->jeobs);
            free(l->jsinlt); free(l->jeinlt); free(l->jsoutlt); free(l->jeoutlt);
        }
        free(s->L);
    }
    free(s->hist);
    orc_config_free(&s->cfg);
    free(s);
}

void orc_set_verbose(orc_solver *s, int v) { s->verbose = v; }
int orc_num_levels(const orc_solver *s) { return s->cfg.ngrid; }

int orc_level_dims(const orc_solver *s, int ig, int *npx, int *npy)
{
    if (ig < 0 || ig >= s->cfg.ngrid) return fail(-1, "level out of range");
    *npx = s->L[ig].npx;
    *npy = s->L[ig].npy;
    return 0;
}

int orc_get(orc_solver *s, int ig, const char *name, double *out)
{
    if (ig < 0 || ig >= s->cfg.ngrid) return fail(-1, "level out of range");
    lvl_t *l = &s->L[ig];
    size_t n = (size_t)l->npx * l->npy;
    if (strcmp(name, "tag") == 0) {
        for (size_t k = 0; k < n; ++k) out[k] = (double)l->tag[k];
        return 0;
    }
    double **a = darr(l, name);
    if (!a) return fail(-1, "unknown array name");
    memcpy(out, *a, n * sizeof(double));
    return 0;
}

int orc_set(orc_solver *s, int ig, const char *name, const double *in)
{
    if (ig < 0 || ig >= s->cfg.ngrid) return fail(-1, "level out of range");
    lvl_t *l = &s->L[ig];
    size_t n = (size_t)l->npx * l->npy;
    if (strcmp(name, "tag") == 0) {
        for (size_t k = 0; k < n; ++k) l->tag[k] = (int)in[k];
        return 0;
    }
    double **a = darr(l, name);
    if (!a) return fail(-1, "unknown array name");
    memcpy(*a, in, n * sizeof(double));
    return 0;
}

/* Relaxation of one level.  leg 0 = down-leg order (V6:201-207: pbound,
 * outflow, solve_sc); leg 1 = up-leg order (V6:237-243: pbound, solve_sc,
 * outflow).  Returns the level's continuity defect. */
double orc_relax(orc_solver *s, int ig, int leg)
{
    lvl_t *l = &s->L[ig];
    fluxes(l);
    solve_u(s, l);
    solve_v(s, l);
    double err = solve_pprime(s, l);
    pbound(l, &s->cfg);
    if (leg == 0) { outflow(l, &s->cfg); solve_sc(s, l); }
    else { solve_sc(s, l); outflow(l, &s->cfg); }
    return err;
}

/* V6:211-220 */
void orc_restrict(orc_solver *s, int ig)
{
    lvl_t *f = &s->L[ig], *c = &s->L[ig + 1];
    restv(f, c);
    fluxes(c);
    fluxes(f);
    restru(f, c, &s->cfg);
    restrv(f, c, &s->cfg);
    restrsc(f, c, &s->cfg);
}

/* V6:230-233: coarse level igc onto igc-1 */
void orc_prolong(orc_solver *s, int igc)
{
    lvl_t *f = &s->L[igc - 1], *c = &s->L[igc];
    prolu(f, c);
    prolv(f, c);
    prolp(f, c);
    prolsc(f, c);
}

/* One V-cycle, V6:197-246.  Returns the level-0 defect recorded at V6:723. */
double orc_vcycle(orc_solver *s, int icyc)
{
    int ng = s->cfg.ngrid;
    double rec = 0.0;
    for (int ig = 0; ig < ng; ++ig) {
        double e = orc_relax(s, ig, 0);
        if (ig == 0) {
            rec = e;
            if (s->verbose) printf("icycle = %d error = %e\n", icyc, e);
        }
        if (ig != ng - 1) orc_restrict(s, ig);
    }
    for (int ig = ng - 1; ig > 0; --ig) {
        orc_prolong(s, ig);
        if (ig != 1) (void)orc_relax(s, ig - 1, 1);
    }
    return rec;
}

/* V6:193-249.  hist must hold nvcycle entries.  Returns 0. */
int orc_solve(orc_solver *s, int nvcycle, double tol, double *hist, int *ncycles, int *converged)
{
    int icyc, conv = 0, done = 0;
    for (icyc = 0; icyc < nvcycle; ++icyc) {
        double e = orc_vcycle(s, icyc);
        if (hist) hist[icyc] = e;
        ++done;
        if (e < tol) { conv = 1; break; }
    }
    if (ncycles) *ncycles = done;
    if (converged) *converged = conv;
    return 0;
}

/* Single-routine entry points (for known-answer tests against the
 * reference's own functions). */
int orc_call(orc_solver *s, const char *fn, int ig, double *ret)
{
    if (ig < 0 || ig >= s->cfg.ngrid) return fail(-1, "level out of range");
    lvl_t *l = &s->L[ig], *c = (ig + 1 < s->cfg.ngrid) ? &s->L[ig + 1] : NULL;
    const orc_config *cf = &s->cfg;
    double r = 0.0;
    if (!strcmp(fn, "fluxes")) fluxes(l);
    else if (!strcmp(fn, "coefu")) coefu(l);
    else if (!strcmp(fn, "umom")) umom(l, cf);
    else if (!strcmp(fn, "coefv")) coefv(l);
    else if (!strcmp(fn, "vmom")) vmom(l, cf);
    else if (!strcmp(fn, "coefsc")) coefsc(l, cf);
    else if (!strcmp(fn, "scalar")) scalar(l, cf);
    else if (!strcmp(fn, "coefp")) coefp(l, cf);
    else if (!strcmp(fn, "masserror")) r = masserror(l);
    else if (!strcmp(fn, "update")) update(l, cf);
    else if (!strcmp(fn, "pbound")) pbound(l, cf);
    else if (!strcmp(fn, "outflow")) outflow(l, cf);
    else if (!strcmp(fn, "solve_u")) solve_u(s, l);
    else if (!strcmp(fn, "solve_v")) solve_v(s, l);
    else if (!strcmp(fn, "solve_pprime")) r = solve_pprime(s, l);
    else if (!strcmp(fn, "solve_sc")) solve_sc(s, l);
    else if (!strcmp(fn, "jacobi_u")) jacobi(l, l->u, l->npx - 2, l->npy - 1, cf->nswpm);
    else if (!strcmp(fn, "jacobi_pp")) jacobi(l, l->pp, l->npx - 1, l->npy - 1, cf->nswpp);
    else if (!strcmp(fn, "rbgs_u")) rbgs(l, l->u, l->npx - 2, l->npy - 1, cf->nswpm);
    else if (!strcmp(fn, "rbgs_pp")) rbgs(l, l->pp, l->npx - 1, l->npy - 1, cf->nswpp);
    else if (!strcmp(fn, "lexgs_pp")) lexgs(l, l->pp, l->npx - 1, l->npy - 1, cf->nswpp);
    else if (c && !strcmp(fn, "restv")) restv(l, c);
    else if (c && !strcmp(fn, "restru")) restru(l, c, cf);
    else if (c && !strcmp(fn, "restrv")) restrv(l, c, cf);
    else if (c && !strcmp(fn, "restrsc")) restrsc(l, c, cf);
    else if (c && !strcmp(fn, "prolu")) prolu(l, c);
    else if (c && !strcmp(fn, "prolv")) prolv(l, c);
    else if (c && !strcmp(fn, "prolp")) prolp(l, c);
    else if (c && !strcmp(fn, "prolsc")) prolsc(l, c);
    else return fail(-1, "unknown routine or no coarser level");
    if (ret) *ret = r;
    return 0;
}
