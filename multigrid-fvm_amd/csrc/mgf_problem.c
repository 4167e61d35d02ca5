/*
 * mgf_problem.c -- reading InputData / pattern.txt into an mgf_problem.
 *
 * Grammar and reading order follow readInput (Version-6-26-04.c:280-401)
 * and read_obstacle_data (Version-6-26-04.c:2367-2400): whitespace
 * separated fields, the pattern file is read between the initial-guess line
 * and sc_obs (V6:357), and the discrete inlet/outlet lines are read without
 * checking the scan result (V6:373, 389).  Where the reference calls
 * exit(), this returns a status and leaves a message.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mgf_plan.h"

static __thread char g_errbuf[512];

void mgf_set_error_string(const char *msg)
{
    snprintf(g_errbuf, sizeof g_errbuf, "%s", msg ? msg : "");
}

const char *mgf_error_string(void) { return g_errbuf; }

int mgf_abi_version(void) { return MGF_ABI_VERSION; }

void mgf_options_default(mgf_options *o)
{
    memset(o, 0, sizeof *o);
    o->precision = MGF_PREC_FP64;
    o->use_graph = 1;
    o->device = -1;
    o->strict_limits = 0;
    o->sweeps_per_launch = 0;
    o->sweep_threads = 0;
    o->split_assembly = 0;
    o->gs_order = 0;
    o->compact = 1;
    o->lanes = 0;
}

static int err(int code, const char *fmt, const char *what)
{
    char b[512];
    snprintf(b, sizeof b, fmt, what);
    mgf_set_error_string(b);
    return code;
}

/* the struct handed out by mgf_problem_read owns its arrays */
typedef struct {
    mgf_problem p;
    int *obs, *inlt_j, *outlt_j;
    double *inlt_v, *outlt_v, *inlet_u;
} owned_problem;

void mgf_problem_release(mgf_problem *p)
{
    if (!p) return;
    free((void *)p->obs);
    free((void *)p->inlt_j);
    free((void *)p->outlt_j);
    free((void *)p->inlt_v);
    free((void *)p->outlt_v);
    free((void *)p->inlet_u);
    memset(p, 0, sizeof *p);
}

void mgf_problem_free(mgf_problem *p)
{
    if (!p) return;
    mgf_problem_release(p);
    free(p);
}

static void *dupmem(const void *src, size_t n)
{
    if (!src || n == 0) return NULL;
    void *d = malloc(n);
    if (d) memcpy(d, src, n);
    return d;
}

int mgf_problem_copy(const mgf_problem *src, mgf_problem *dst)
{
    *dst = *src;
    dst->obs = (const int *)dupmem(src->obs, sizeof(int) * 4 * (size_t)(src->nobs > 0 ? src->nobs : 0));
    dst->inlt_j = (const int *)dupmem(src->inlt_j, sizeof(int) * 2 * (size_t)(src->ninlt > 0 ? src->ninlt : 0));
    dst->outlt_j = (const int *)dupmem(src->outlt_j, sizeof(int) * 2 * (size_t)(src->noutlt > 0 ? src->noutlt : 0));
    dst->inlt_v = (const double *)dupmem(src->inlt_v, sizeof(double) * 3 * (size_t)(src->ninlt > 0 ? src->ninlt : 0));
    dst->outlt_v = (const double *)dupmem(src->outlt_v, sizeof(double) * 3 * (size_t)(src->noutlt > 0 ? src->noutlt : 0));
    dst->inlet_u = (const double *)dupmem(src->inlet_u, sizeof(double) * (size_t)(src->nfy > 0 ? src->nfy : 0));
    return MGF_OK;
}

static int read_pattern(const char *path, mgf_problem *p)
{
    FILE *f = fopen(path, "r");
    if (!f) return err(MGF_ERR_IO, "cannot open obstacle pattern file %s", path);
    int nobs = 0;
    if (fscanf(f, "%d", &nobs) != 1) {
        fclose(f);
        return err(MGF_ERR_PARSE, "failed reading the obstacle count of %s", path);
    }
    if (nobs < 0) nobs = 0;
    int *o = nobs ? (int *)malloc(sizeof(int) * 4 * (size_t)nobs) : NULL;
    for (int k = 0; k < nobs; ++k) {
        if (fscanf(f, " %d %d %d %d", &o[4 * k], &o[4 * k + 1], &o[4 * k + 2], &o[4 * k + 3]) != 4) {
            fclose(f);
            free(o);
            return err(MGF_ERR_PARSE, "failed reading obstacle coordinates in %s", path);
        }
    }
    fclose(f);
    p->nobs = nobs;
    p->obs = o;
    return MGF_OK;
}

int mgf_problem_read(const char *inputdata, const char *pattern, mgf_problem **out)
{
    if (!inputdata || !pattern || !out) return err(MGF_ERR_ARG, "%s", "null argument");
    *out = NULL;
    FILE *f = fopen(inputdata, "r");
    if (!f) return err(MGF_ERR_IO, "cannot open %s", inputdata);
    mgf_problem *p = (mgf_problem *)calloc(1, sizeof *p);
    int rc = MGF_OK;
    static const char *line_names[] = {"first", "second", "third", "fourth", "fifth",
                                       "sixth", "seventh", "eighth"};
#define RD(n, line, ...)                                                        \
    do {                                                                        \
        if (fscanf(f, __VA_ARGS__) != (n)) {                                    \
            rc = err(MGF_ERR_PARSE, "failed reading the %s line of InputData",  \
                     line_names[line]);                                         \
            goto fail;                                                          \
        }                                                                       \
    } while (0)
    RD(13, 0, "%d %d %d %d %d %d %d %d %d %d %d %lf %lf", &p->nfx, &p->nfy, &p->ngrid,
       &p->ig_obs, &p->ibeg_obs, &p->ig_inlt, &p->ig_outlt, &p->nvcycle, &p->nswpm,
       &p->nswpp, &p->isolve, &p->omega, &p->convergence);
    RD(6, 1, "%lf %lf %lf %lf %lf %lf", &p->xl, &p->yl, &p->amuf, &p->prsc, &p->relxm, &p->relxp);
    RD(4, 2, "%d %d %d %d", &p->ibc[0], &p->ibc[1], &p->ibc[2], &p->ibc[3]);
    for (int s = 0; s < 4; ++s)
        RD(4, 3 + s, "%lf %lf %lf %lf", &p->bc[s][0], &p->bc[s][1], &p->bc[s][2], &p->bc[s][3]);
    RD(4, 7, "%lf %lf %lf %lf", &p->guess[0], &p->guess[1], &p->guess[2], &p->guess[3]);
    if ((rc = read_pattern(pattern, p)) != MGF_OK) goto fail;
    if (fscanf(f, "%lf", &p->sc_obs) != 1) {
        rc = err(MGF_ERR_PARSE, "%s", "failed reading the obstacle scalar line of InputData");
        goto fail;
    }
    if (fscanf(f, "%d", &p->ninlt) != 1) {
        rc = err(MGF_ERR_PARSE, "%s", "failed reading the inlet count of InputData");
        goto fail;
    }
    if (p->ninlt < 0) p->ninlt = 0;
    {
        int *jj = p->ninlt ? (int *)calloc(2 * (size_t)p->ninlt, sizeof(int)) : NULL;
        double *vv = p->ninlt ? (double *)calloc(3 * (size_t)p->ninlt, sizeof(double)) : NULL;
        for (int k = 0; k < p->ninlt; ++k)
            (void)!fscanf(f, " %d %d %lf %lf %lf", &jj[2 * k], &jj[2 * k + 1], &vv[3 * k],
                          &vv[3 * k + 1], &vv[3 * k + 2]);
        p->inlt_j = jj;
        p->inlt_v = vv;
    }
    if (fscanf(f, "%d", &p->noutlt) != 1) {
        rc = err(MGF_ERR_PARSE, "%s", "failed reading the outlet count of InputData");
        goto fail;
    }
    if (p->noutlt < 0) p->noutlt = 0;
    {
        int *jj = p->noutlt ? (int *)calloc(2 * (size_t)p->noutlt, sizeof(int)) : NULL;
        double *vv = p->noutlt ? (double *)calloc(3 * (size_t)p->noutlt, sizeof(double)) : NULL;
        for (int k = 0; k < p->noutlt; ++k)
            (void)!fscanf(f, " %d %d %lf %lf %lf", &jj[2 * k], &jj[2 * k + 1], &vv[3 * k],
                          &vv[3 * k + 1], &vv[3 * k + 2]);
        p->outlt_j = jj;
        p->outlt_v = vv;
    }
#undef RD
    fclose(f);
    p->dpdxm = 0.0;
    p->dpdym = 0.0;
    *out = p;
    return MGF_OK;
fail:
    fclose(f);
    mgf_problem_free(p);
    return rc;
}
