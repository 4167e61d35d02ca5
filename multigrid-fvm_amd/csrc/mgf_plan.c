/*
 * mgf_plan.c -- host set-up of the level hierarchy and the per-point data
 * that lets the device kernels apply the reference's sequential obstacle
 * loops in parallel with identical arithmetic.
 *
 *   level sizes, metrics, obstacle/inlet index scaling: AllocateMemorySoA,
 *       Version-6-26-04.c:402-506
 *   initial and boundary state: init, Version-6-26-04.c:546-676
 *   coarse face areas and coarse inlet/outlet values: rest_geom,
 *       Version-6-26-04.c:1391-1473
 *
 * Levels ig >= ig_obs, on which the reference reads uninitialised obstacle
 * index arrays (V6:471, 618, 939-963, 1153-1177): a cell there is an obstacle
 * when all of its children are (mgf_plan_level.ext marks these levels; the
 * restriction onto and prolongation from them mask penalised points, see
 * the kernels).  With ig_obs == ngrid nothing here differs from the
 * reference.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mgf_plan.h"

static int perr(char *err, size_t n, int code, const char *msg)
{
    if (err && n) snprintf(err, n, "%s", msg);
    return code;
}

void mgf_plan_free(mgf_plan *pl)
{
    if (!pl) return;
    if (pl->lv) {
        for (int ig = 0; ig < pl->ngrid; ++ig) {
            mgf_plan_level *l = &pl->lv[ig];
            free(l->u); free(l->v); free(l->p); free(l->sc);
            free(l->flags); free(l->sccode); free(l->pslot); free(l->porig);
        }
        free(pl->lv);
    }
    mgf_problem_release(&pl->prob);
    free(pl);
}

/* ---- validation ---------------------------------------------------------- */

static int validate(const mgf_problem *p, int strict, char *err, size_t n)
{
    if (p->nfx < 1 || p->nfy < 1) return perr(err, n, MGF_ERR_CONFIG, "nfx and nfy must be positive");
    if (p->ngrid < 1 || p->ngrid > 24) return perr(err, n, MGF_ERR_CONFIG, "ngrid must be in [1, 24]");
    if ((p->nfx >> (p->ngrid - 1)) < 1 || (p->nfy >> (p->ngrid - 1)) < 1)
        return perr(err, n, MGF_ERR_CONFIG, "the coarsest level has no interior cells");
    if (p->nobs > 0 && (p->ig_obs < 1 || p->ig_obs > p->ngrid))
        return perr(err, n, MGF_ERR_CONFIG, "ig_obs must be in [1, ngrid] when obstacles are given");
    if (p->ninlt > 0 && p->ig_inlt > p->ngrid)
        return perr(err, n, MGF_ERR_CONFIG, "ig_inlt must be <= ngrid");
    if (p->noutlt > 0 && p->ig_outlt > p->ngrid)
        return perr(err, n, MGF_ERR_CONFIG, "ig_outlt must be <= ngrid");
    if (p->nobs > 0 && !p->obs) return perr(err, n, MGF_ERR_ARG, "obstacle array missing");
    if (p->ninlt > 0 && (!p->inlt_j || !p->inlt_v)) return perr(err, n, MGF_ERR_ARG, "inlet arrays missing");
    if (p->noutlt > 0 && (!p->outlt_j || !p->outlt_v)) return perr(err, n, MGF_ERR_ARG, "outlet arrays missing");
    if (p->relxm == 0.0) return perr(err, n, MGF_ERR_CONFIG, "relxm must be nonzero");
    if (p->prsc == 0.0) return perr(err, n, MGF_ERR_CONFIG, "prsc must be nonzero");
    if ((double)(p->nfx + 2) * (double)(p->nfy + 2) > 1.5e9)
        return perr(err, n, MGF_ERR_LIMIT, "grid too large for 32-bit point indices");
    if (strict) {
        if (p->nobs > 1000) return perr(err, n, MGF_ERR_LIMIT, "nobs > 1000 (reference limit, V6:29)");
        if (p->ninlt > 10 || p->noutlt > 10)
            return perr(err, n, MGF_ERR_LIMIT, "more than 10 inlets/outlets (reference limit, V6:29)");
        if (p->nvcycle > 3000)
            return perr(err, n, MGF_ERR_LIMIT, "nvcycle > 3000 (reference history limit, V6:39)");
    }
    return MGF_OK;
}

/* ---- build ----------------------------------------------------------------- */

int mgf_plan_build(const mgf_problem *p, int strict, mgf_plan **out, char *err, size_t errlen)
{
    *out = NULL;
    int rc = validate(p, strict, err, errlen);
    if (rc) return rc;
    mgf_plan *pl = (mgf_plan *)calloc(1, sizeof *pl);
    mgf_problem_copy(p, &pl->prob);
    const mgf_problem *c = &pl->prob;
    int ng = c->ngrid;
    pl->ngrid = ng;
    pl->lv = (mgf_plan_level *)calloc((size_t)ng, sizeof(mgf_plan_level));

    /* per-level index ranges of obstacles / inlets / outlets (V6:468-502) */
    int **ob = (int **)calloc((size_t)ng, sizeof(int *));   /* nob[ig] x {is, ie, js, je} */
    int *nob = (int *)calloc((size_t)ng, sizeof(int));
    int *inl = (int *)calloc((size_t)ng * 2 * (size_t)(c->ninlt > 0 ? c->ninlt : 1), sizeof(int));
    int *outl = (int *)calloc((size_t)ng * 2 * (size_t)(c->noutlt > 0 ? c->noutlt : 1), sizeof(int));
    double **areax = (double **)calloc((size_t)ng, sizeof(double *));
    double **areay = (double **)calloc((size_t)ng, sizeof(double *));

    for (int ig = 0; ig < ng; ++ig) {
        mgf_plan_level *l = &pl->lv[ig];
        int fac = 1 << ig;
        l->npx = c->nfx / fac + 2;
        l->npy = c->nfy / fac + 2;
        l->dx = c->xl / (l->npx - 2);
        l->dy = c->yl / (l->npy - 2);
        l->ax = l->dy;
        l->ay = l->dx;
        size_t n = (size_t)l->npx * (size_t)l->npy;
        l->u = (double *)malloc(n * sizeof(double));
        l->v = (double *)malloc(n * sizeof(double));
        l->p = (double *)malloc(n * sizeof(double));
        l->sc = (double *)malloc(n * sizeof(double));
        l->flags = (uint8_t *)calloc(n, 1);
        l->sccode = (uint64_t *)calloc(n, sizeof(uint64_t));
        l->pslot = (int *)malloc(n * sizeof(int));
        for (size_t k = 0; k < n; ++k) l->pslot[k] = -1;
        areax[ig] = (double *)malloc(n * sizeof(double));
        areay[ig] = (double *)malloc(n * sizeof(double));
    }
    for (int ig = 0; ig < ng; ++ig) {
        mgf_plan_level *l = &pl->lv[ig];
        size_t cap = (c->nobs > 0 && ig >= c->ig_obs) ? (size_t)(l->npx - 2) * (size_t)(l->npy - 2)
                                                       : (size_t)(c->nobs > 0 ? c->nobs : 1);
        ob[ig] = (int *)calloc(4 * cap, sizeof(int));
        if (c->nobs > 0 && ig < c->ig_obs) nob[ig] = c->nobs;
    }
    for (int k = 0; k < c->nobs; ++k) {
        const int *o = c->obs + 4 * k;
        for (int ig = 0; ig < c->ig_obs; ++ig) {
            int fac = 1 << (c->ig_obs - 1 - ig);
            int *r = ob[ig] + (size_t)k * 4;
            r[0] = (o[0] + c->ibeg_obs - 1) * fac + 1;
            r[1] = (o[1] + c->ibeg_obs) * fac;
            r[2] = (o[2] - 1) * fac + 1;
            r[3] = o[3] * fac;
            mgf_plan_level *l = &pl->lv[ig];
            if (r[0] < 1 || r[1] > l->npx - 2 || r[2] < 1 || r[3] > l->npy - 2) {
                rc = perr(err, errlen, MGF_ERR_GEOMETRY, "obstacle rectangle outside the interior");
                goto done;
            }
        }
    }
    /* Levels ig >= ig_obs, where the reference is undefined: a cell is an
     * obstacle when all of its 2x2 children are, one rectangle per cell,
     * i-major then j ascending (a subset of the finer obstacle region:
     * blocking partly open cells would cut flow paths and can leave
     * inlet-fed dead ends whose p' equation is singular). */
    for (int ig = c->ig_obs; c->nobs > 0 && ig < ng; ++ig) {
        mgf_plan_level *f = &pl->lv[ig - 1], *l = &pl->lv[ig];
        int nxf = f->npx, nyf = f->npy;
        unsigned char *blk = (unsigned char *)calloc((size_t)nxf * nyf, 1);
        for (int k = 0; k < nob[ig - 1]; ++k) {
            const int *r = ob[ig - 1] + (size_t)k * 4;
            for (int j = r[2]; j <= r[3]; ++j)
                for (int i = r[0]; i <= r[1]; ++i) blk[j * nxf + i] = 1;
        }
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
                    int *r = ob[ig] + (size_t)n * 4;
                    r[0] = ic; r[1] = ic; r[2] = jc; r[3] = jc;
                    ++n;
                }
            }
        nob[ig] = n;
        l->ext = 1;
        free(blk);
    }
    for (int ig = 0; ig < ng; ++ig) pl->lv[ig].has_obs = nob[ig] > 0;

    for (int k = 0; k < c->ninlt; ++k)
        for (int ig = 0; ig < c->ig_inlt; ++ig) {
            int fac = 1 << (c->ig_inlt - 1 - ig);
            int *r = inl + ((size_t)ig * (size_t)c->ninlt + (size_t)k) * 2;
            r[0] = (c->inlt_j[2 * k] - 1) * fac + 1;
            r[1] = c->inlt_j[2 * k + 1] * fac;
            if (r[0] < 0 || r[1] > pl->lv[ig].npy - 1) {
                rc = perr(err, errlen, MGF_ERR_GEOMETRY, "inlet rows outside the grid");
                goto done;
            }
        }
    for (int k = 0; k < c->noutlt; ++k)
        for (int ig = 0; ig < c->ig_outlt; ++ig) {
            int fac = 1 << (c->ig_outlt - 1 - ig);
            int *r = outl + ((size_t)ig * (size_t)c->noutlt + (size_t)k) * 2;
            r[0] = (c->outlt_j[2 * k] - 1) * fac + 1;
            r[1] = c->outlt_j[2 * k + 1] * fac;
            if (r[0] < 0 || r[1] > pl->lv[ig].npy - 1) {
                rc = perr(err, errlen, MGF_ERR_GEOMETRY, "outlet rows outside the grid");
                goto done;
            }
        }

    /* ---- init (V6:546-676) ---- */
    for (int ig = 0; ig < ng; ++ig) {
        mgf_plan_level *l = &pl->lv[ig];
        int nx = l->npx, ny = l->npy;
        size_t n = (size_t)nx * ny;
        for (size_t k = 0; k < n; ++k) {
            l->u[k] = c->guess[0];
            l->v[k] = c->guess[1];
            l->sc[k] = c->guess[3];
            l->p[k] = c->guess[2];
            areax[ig][k] = l->dy;
            areay[ig][k] = l->dx;
        }
        for (int j = 0; j < ny; ++j) {                     /* xm ghosts */
            int ij = j * nx;
            l->u[ij] = c->bc[0][0]; l->v[ij] = c->bc[0][1];
            l->p[ij] = c->bc[0][2]; l->sc[ij] = c->bc[0][3];
        }
        for (int j = 0; j < ny; ++j) {                     /* xp: u on face npx-2 */
            int ij = j * nx + nx - 1;
            l->u[ij - 1] = c->bc[1][0]; l->v[ij] = c->bc[1][1];
            l->p[ij] = c->bc[1][2]; l->sc[ij] = c->bc[1][3];
        }
        for (int i = 0; i < nx; ++i) {                     /* ym */
            l->u[i] = c->bc[2][0]; l->v[i] = c->bc[2][1];
            l->p[i] = c->bc[2][2]; l->sc[i] = c->bc[2][3];
        }
        for (int i = 0; i < nx; ++i) {                     /* yp */
            int ij = (ny - 1) * nx + i;
            l->u[ij] = c->bc[3][0]; l->v[ij] = c->bc[3][1];
            l->p[ij] = c->bc[3][2]; l->sc[ij] = c->bc[3][3];
        }
    }
    for (int ig = 0; ig < ng; ++ig) {                      /* obstacles */
        mgf_plan_level *l = &pl->lv[ig];
        if (!l->has_obs) continue;
        int nx = l->npx;
        for (int k = 0; k < nob[ig]; ++k) {
            const int *r = ob[ig] + (size_t)k * 4;
            for (int j = r[2]; j <= r[3]; ++j)
                for (int i = r[0]; i <= r[1]; ++i) {
                    int ij = j * nx + i;
                    l->u[ij] = 0.0; l->v[ij] = 0.0;
                    l->u[ij - 1] = 0.0; l->v[ij - nx] = 0.0;
                    l->flags[ij] |= MGF_F_SPU | MGF_F_SPV | MGF_F_SPSC | MGF_F_OBS;
                    l->flags[ij - 1] |= MGF_F_SPU;
                    l->flags[ij - nx] |= MGF_F_SPV;
                    l->sc[ij] = c->sc_obs;
                }
        }
    }
    for (int ig = 0; ig < c->ig_inlt && c->ninlt > 0; ++ig) {   /* discrete inlets */
        mgf_plan_level *l = &pl->lv[ig];
        for (int k = 0; k < c->ninlt; ++k) {
            const int *r = inl + ((size_t)ig * (size_t)c->ninlt + (size_t)k) * 2;
            for (int j = r[0]; j <= r[1]; ++j) {
                int ij = j * l->npx;
                l->u[ij] = c->inlt_v[3 * k];
                l->v[ij] = c->inlt_v[3 * k + 1];
                l->sc[ij] = c->inlt_v[3 * k + 2];
            }
        }
    }
    for (int ig = 0; ig < c->ig_outlt && c->noutlt > 0; ++ig) { /* discrete outlets */
        mgf_plan_level *l = &pl->lv[ig];
        for (int k = 0; k < c->noutlt; ++k) {
            const int *r = outl + ((size_t)ig * (size_t)c->noutlt + (size_t)k) * 2;
            for (int j = r[0]; j <= r[1]; ++j) {
                int ij = j * l->npx + l->npx - 2;
                l->u[ij] = c->outlt_v[3 * k];
                l->v[ij] = c->outlt_v[3 * k + 1];
                l->sc[ij + 1] = c->outlt_v[3 * k + 2];
            }
        }
    }
    if (c->inlet_u) {                                      /* inlet profile extension */
        mgf_plan_level *l = &pl->lv[0];
        for (int j = 1; j <= l->npy - 2; ++j) l->u[j * l->npx] = c->inlet_u[j - 1];
    }

    /* ---- rest_geom (V6:1391-1473) ---- */
    for (int ig = 1; ig < ng; ++ig) {
        mgf_plan_level *f = &pl->lv[ig - 1], *cl = &pl->lv[ig];
        int nxf = f->npx, nxc = cl->npx, nyc = cl->npy;
        for (int jc = 1; jc < nyc - 1; ++jc)
            for (int ic = 0; ic < nxc - 1; ++ic) {
                int ij = (2 * jc) * nxf + 2 * ic;
                areax[ig][jc * nxc + ic] = areax[ig - 1][ij] + areax[ig - 1][ij - nxf];
            }
        for (int jc = 0; jc < nyc - 1; ++jc)
            for (int ic = 1; ic < nxc - 1; ++ic) {
                int ij = (2 * jc) * nxf + 2 * ic;
                areay[ig][jc * nxc + ic] = areay[ig - 1][ij] + areay[ig - 1][ij - 1];
            }
    }
    for (int ig = 1; ig < ng; ++ig) {
        mgf_plan_level *f = &pl->lv[ig - 1], *cl = &pl->lv[ig];
        int nxf = f->npx, nxc = cl->npx, nyc = cl->npy;
        for (int jc = 1; jc < nyc - 1; ++jc) {
            int ijc = jc * nxc, ij = (2 * jc) * nxf;
            cl->u[ijc] = 0.5 * (f->u[ij] + f->u[ij - nxf]);
            cl->v[ijc] = 0.5 * (f->v[ij] + f->v[ij - 1]);   /* previous row's last point */
        }
    }
    for (int ig = 1; ig < ng; ++ig) {
        mgf_plan_level *f = &pl->lv[ig - 1], *cl = &pl->lv[ig];
        int nxf = f->npx, nxc = cl->npx, nyc = cl->npy;
        for (int jc = 1; jc < nyc - 1; ++jc) {
            int ijc = jc * nxc + (nxc - 2);
            int ij = (2 * jc) * nxf + 2 * (nxc - 2);
            cl->u[ijc] = 0.5 * (f->u[ij] + f->u[ij - nxf]);
            cl->v[ijc] = 0.5 * (f->v[ij] + f->v[ij - 1]);
        }
    }
    /* the kernels use one face area per direction and level: every entry
     * must equal it bit for bit (true whenever the levels nest 2:1) */
    for (int ig = 0; ig < ng; ++ig) {
        mgf_plan_level *l = &pl->lv[ig];
        size_t n = (size_t)l->npx * l->npy;
        for (size_t k = 0; k < n; ++k)
            if (memcmp(&areax[ig][k], &l->ax, sizeof(double)) != 0 ||
                memcmp(&areay[ig][k], &l->ay, sizeof(double)) != 0) {
                rc = perr(err, errlen, MGF_ERR_GEOMETRY,
                          "levels do not nest 2:1 (nfx and nfy must be divisible by 2^(ngrid-1))");
                goto done;
            }
    }

    /* ---- obstacle diffusion adds of coefsc (V6:939-963), per target ---- */
    for (int ig = 0; ig < ng; ++ig) {
        mgf_plan_level *l = &pl->lv[ig];
        if (!l->has_obs) continue;
        int nx = l->npx, ny = l->npy;
        for (int k = 0; k < nob[ig]; ++k) {
            const int *r = ob[ig] + (size_t)k * 4;
            for (int i = r[0]; i <= r[1]; ++i)
                for (int j = r[2]; j <= r[3]; ++j) {
                    int ij = j * nx + i;
                    /* (target, coefficient slot 0 ae 1 aw 2 an 3 as, y-type) */
                    int tg[4] = {ij - 1, ij + 1, ij - nx, ij + nx};
                    for (int t = 0; t < 4; ++t) {
                        int x = tg[t];
                        int ti = x % nx, tj = x / nx;
                        if (ti < 1 || ti > nx - 2 || tj < 1 || tj > ny - 2) continue;
                        uint64_t code = l->sccode[x];
                        unsigned cnt = (unsigned)((code >> (5 * t)) & 31u);
                        unsigned nap = (unsigned)((code >> 20) & 63u);
                        if (cnt >= MGF_SC_MAXCNT || nap >= MGF_SC_MAXAP) {
                            rc = perr(err, errlen, MGF_ERR_LIMIT,
                                      "too many overlapping obstacle diffusion adds on one cell");
                            goto done;
                        }
                        code += (uint64_t)1 << (5 * t);
                        if (t >= 2) code |= (uint64_t)1 << (26 + nap);
                        code = (code & ~((uint64_t)63 << 20)) | ((uint64_t)(nap + 1) << 20);
                        l->sccode[x] = code;
                        l->flags[x] |= MGF_F_SCADD;
                    }
                }
        }
    }

    /* ---- obstacle-edge pressure copies of update (V6:1153-1177) ----
     * Symbolic execution of p[d] = p[s] in list order gives, for every
     * destination, the cell whose freshly updated value it ends with. */
    for (int ig = 0; ig < ng; ++ig) {
        mgf_plan_level *l = &pl->lv[ig];
        if (!l->has_obs) continue;
        int nx = l->npx;
        size_t n = (size_t)nx * l->npy;
        int *src = (int *)malloc(n * sizeof(int));
        for (size_t k = 0; k < n; ++k) src[k] = (int)k;
        for (int k = 0; k < nob[ig]; ++k) {
            const int *r = ob[ig] + (size_t)k * 4;
            int j = r[2];
            for (int i = r[0]; i <= r[1]; ++i) src[j * nx + i] = src[j * nx + i - nx];
            j = r[3];
            for (int i = r[0]; i <= r[1]; ++i) src[j * nx + i] = src[j * nx + i + nx];
            int i = r[0];
            for (int jj = r[2]; jj <= r[3]; ++jj) src[jj * nx + i] = src[jj * nx + i - 1];
            i = r[1];
            for (int jj = r[2]; jj <= r[3]; ++jj) src[jj * nx + i] = src[jj * nx + i + 1];
        }
        int cnt = 0;
        for (size_t k = 0; k < n; ++k)
            if (src[k] != (int)k) ++cnt;
        l->npcopy = cnt;
        l->porig = (int *)malloc((size_t)(cnt > 0 ? cnt : 1) * sizeof(int));
        cnt = 0;
        for (size_t k = 0; k < n; ++k)
            if (src[k] != (int)k) {
                l->pslot[k] = cnt;
                l->porig[cnt++] = src[k];
                l->flags[k] |= MGF_F_PCOPY;
                l->flags[src[k]] |= MGF_F_PORIG;
            }
        free(src);
    }

done:
    for (int ig = 0; ig < ng; ++ig) {
        free(areax[ig]);
        free(areay[ig]);
    }
    for (int ig = 0; ig < ng; ++ig) free(ob[ig]);
    free(areax); free(areay); free(ob); free(nob); free(inl); free(outl);
    if (rc) {
        mgf_plan_free(pl);
        return rc;
    }
    *out = pl;
    return MGF_OK;
}
