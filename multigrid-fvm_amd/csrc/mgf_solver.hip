/*
 * mgf_solver.hip -- solver objects behind the C ABI of include/mgfvm.h.
 *
 * One solver = one problem on one device: a single HBM allocation (the
 * "arena") holding every level's arrays, one HIP stream, a device control
 * block, the residual history, and (once a solve starts) one captured
 * V-cycle graph that is replayed until the device-side stopping test fires.
 *
 * The schedule is the reference's, Version-6-26-04.c:195-249 (relax every
 * level on the way down with FAS restriction in between, prolongate on the
 * way up with post-relaxation of every level except 0, stop when the
 * level-0 sum of |div| recorded in this cycle is below the tolerance).
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mgf_dev.h"
#include "mgf_plan.h"
#include "mgfvm.h"

struct mgf_solver {
    mgf_plan *plan;
    mgf_options opt;
    int device;
    int ng;
    hipStream_t st;
    void *arena;
    size_t arena_bytes;
    DevLevel lv[MGF_MAX_LEVELS];
    double *init_state[MGF_MAX_LEVELS];   /* u, v, p, sc as built at creation */
    DevPhys phys;
    int isolve, nswpm, nswpp;
    DevCtl *ctl;            /* device */
    DevCtl *ctl_host;       /* pinned mirror */
    double *hist;           /* device, hist_cap entries */
    int hist_cap;
    hipGraphExec_t gexec;
    int graph_ready;
    char err[512];
};

static int set_err(mgf_solver *s, int code, const char *fmt, ...)
    __attribute__((format(printf, 3, 4)));

static int set_err(mgf_solver *s, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (s) snprintf(s->err, sizeof s->err, "%s", buf);
    mgf_set_error_string(buf);
    return code;
}

#define HIPCHK(s, call)                                                                 \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_err((s), MGF_ERR_DEVICE, "%s failed: %s", #call,                 \
                           hipGetErrorString(e_));                                      \
    } while (0)

static int check_launch(mgf_solver *s)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_err(s, MGF_ERR_DEVICE, "kernel launch failed: %s", hipGetErrorString(e));
    return MGF_OK;
}

/* ---------------------------------------------------------------------- */
/* creation                                                                 */
/* ---------------------------------------------------------------------- */

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static int alloc_arena(mgf_solver *s)
{
    const mgf_plan *pl = s->plan;
    size_t total = 0;
    for (int ig = 0; ig < s->ng; ++ig) {
        const mgf_plan_level *l = &pl->lv[ig];
        size_t n = (size_t)l->npx * l->npy;
        total += 27 * align256(n * sizeof(double));
        total += align256((size_t)mgf_partial_blocks(l->npx, l->npy) * sizeof(double));
        total += align256(n);                                  /* flags */
        total += align256(n * sizeof(uint64_t));               /* sccode */
        total += align256(n * sizeof(int));                    /* pslot */
        total += align256((size_t)(l->npcopy + 1) * sizeof(int));
        total += align256((size_t)(l->npcopy + 1) * sizeof(double));
    }
    s->arena_bytes = total;
    HIPCHK(s, hipMalloc(&s->arena, total));
    HIPCHK(s, hipMemsetAsync(s->arena, 0, total, s->st));

    char *cur = (char *)s->arena;
    auto take = [&](size_t bytes) {
        char *p = cur;
        cur += align256(bytes);
        return (void *)p;
    };
    const mgf_problem *c = &pl->prob;
    const double amusc = c->amuf / c->prsc;
    for (int ig = 0; ig < s->ng; ++ig) {
        const mgf_plan_level *l = &pl->lv[ig];
        DevLevel &L = s->lv[ig];
        memset(&L, 0, sizeof L);
        L.npx = l->npx;
        L.npy = l->npy;
        L.npcopy = l->npcopy;
        L.has_obs = l->has_obs;
        L.ext = l->ext;
        L.ax = l->ax;
        L.ay = l->ay;
        /* the reference evaluates these per point from areax/areay arrays
         * that hold exactly ax/ay; same expressions, same bits */
        L.du_x = c->amuf * (l->ax / l->dx);
        L.du_y = c->amuf * (l->ay / l->dy);
        L.ds_x = amusc * (l->ax / l->dx);
        L.ds_y = amusc * (l->ay / l->dy);
        L.axax = l->ax * l->ax;
        L.ayay = l->ay * l->ay;
        L.bodyx = (c->dpdxm * l->ax) * l->ay;
        L.bodyy = (c->dpdym * l->ax) * l->ay;
        size_t n = (size_t)l->npx * l->npy, nb = n * sizeof(double);
        double **arrs[23] = {&L.u, &L.v, &L.p, &L.pp, &L.sc, &L.resu, &L.resv, &L.ressc,
                             &L.apu, &L.apv, &L.cx, &L.cy, &L.aw, &L.ae, &L.as, &L.an,
                             &L.ap, &L.su, &L.tmp, &L.cu, &L.cv, &L.cp, &L.csc};
        for (int k = 0; k < 23; ++k) *arrs[k] = (double *)take(nb);
        s->init_state[ig] = (double *)take(4 * nb);
        L.partial = (double *)take((size_t)mgf_partial_blocks(l->npx, l->npy) * sizeof(double));
        L.flags = (const uint8_t *)take(n);
        L.sccode = (const uint64_t *)take(n * sizeof(uint64_t));
        L.pslot = (const int *)take(n * sizeof(int));
        L.porig = (const int *)take((size_t)(l->npcopy + 1) * sizeof(int));
        L.pstash = (double *)take((size_t)(l->npcopy + 1) * sizeof(double));

        HIPCHK(s, hipMemcpyAsync(L.u, l->u, nb, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(L.v, l->v, nb, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(L.p, l->p, nb, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(L.sc, l->sc, nb, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync((void *)L.flags, l->flags, n, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync((void *)L.sccode, l->sccode, n * sizeof(uint64_t),
                                 hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync((void *)L.pslot, l->pslot, n * sizeof(int),
                                 hipMemcpyHostToDevice, s->st));
        if (l->npcopy > 0)
            HIPCHK(s, hipMemcpyAsync((void *)L.porig, l->porig, (size_t)l->npcopy * sizeof(int),
                                     hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(s->init_state[ig], l->u, nb, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(s->init_state[ig] + n, l->v, nb, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(s->init_state[ig] + 2 * n, l->p, nb, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(s->init_state[ig] + 3 * n, l->sc, nb, hipMemcpyHostToDevice, s->st));
        /* apu = apv = 1e30 everywhere (V6:568-569); their ghost entries are
         * never overwritten and make the p' boundary coefficients ~0 */
        mgf_launch_fill(L.apu, 1.0e30, n, s->st);
        mgf_launch_fill(L.apv, 1.0e30, n, s->st);
        int rc = check_launch(s);
        if (rc) return rc;
    }
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

static int ensure_hist(mgf_solver *s, int need)
{
    if (need <= s->hist_cap) return MGF_OK;
    int cap = need > 4096 ? need : 4096;
    double *h = NULL;
    HIPCHK(s, hipMalloc(&h, (size_t)cap * sizeof(double)));
    HIPCHK(s, hipMemsetAsync(h, 0, (size_t)cap * sizeof(double), s->st));
    if (s->hist) {
        HIPCHK(s, hipStreamSynchronize(s->st));
        (void)hipFree(s->hist);
    }
    s->hist = h;
    s->hist_cap = cap;
    if (s->graph_ready) {            /* the captured graph points at the old buffer */
        (void)hipGraphExecDestroy(s->gexec);
        s->graph_ready = 0;
    }
    return MGF_OK;
}

extern "C" int mgf_create(const mgf_problem *p, const mgf_options *opt, mgf_solver **out)
{
    if (!p || !out) return set_err(NULL, MGF_ERR_ARG, "null argument");
    *out = NULL;
    mgf_options o;
    if (opt)
        o = *opt;
    else
        mgf_options_default(&o);
    if (o.precision != MGF_PREC_FP64)
        return set_err(NULL, MGF_ERR_UNSUPPORTED, "precision mode %d not available", o.precision);
    mgf_solver *s = (mgf_solver *)calloc(1, sizeof *s);
    s->opt = o;
    char e[512] = {0};
    int rc = mgf_plan_build(p, o.strict_limits, &s->plan, e, sizeof e);
    if (rc) {
        set_err(NULL, rc, "%s", e);
        free(s);
        return rc;
    }
    s->ng = s->plan->ngrid;
    if (s->ng > MGF_MAX_LEVELS) {
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_LIMIT, "more than %d levels", MGF_MAX_LEVELS);
    }
    const mgf_problem *c = &s->plan->prob;
    s->isolve = c->isolve;
    s->nswpm = c->nswpm;
    s->nswpp = c->nswpp;
    s->phys.relxm = c->relxm;
    s->phys.relxm1 = 1.0 - c->relxm;
    s->phys.relxp = c->relxp;
    s->phys.sc_obs = c->sc_obs;
    for (int k = 0; k < 4; ++k) s->phys.ibc[k] = c->ibc[k];

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_DEVICE, "no HIP device available");
    }
    if (o.device >= 0) {
        if (hipSetDevice(o.device) != hipSuccess) {
            mgf_destroy(s);
            return set_err(NULL, MGF_ERR_DEVICE, "cannot select device %d", o.device);
        }
    }
    (void)hipGetDevice(&s->device);
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_DEVICE, "stream creation failed");
    }
    if ((rc = alloc_arena(s)) != MGF_OK) {
        char keep[512];
        snprintf(keep, sizeof keep, "%s", s->err);
        mgf_destroy(s);
        return set_err(NULL, rc, "%s", keep);
    }
    if (hipMalloc(&s->ctl, sizeof(DevCtl)) != hipSuccess ||
        hipHostMalloc(&s->ctl_host, sizeof(DevCtl), 0) != hipSuccess) {
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_DEVICE, "control block allocation failed");
    }
    memset(s->ctl_host, 0, sizeof(DevCtl));
    (void)hipMemcpy(s->ctl, s->ctl_host, sizeof(DevCtl), hipMemcpyHostToDevice);
    if ((rc = ensure_hist(s, c->nvcycle > 0 ? c->nvcycle : 1)) != MGF_OK) {
        char keep[512];
        snprintf(keep, sizeof keep, "%s", s->err);
        mgf_destroy(s);
        return set_err(NULL, rc, "%s", keep);
    }
    (void)hipStreamSynchronize(s->st);
    *out = s;
    return MGF_OK;
}

extern "C" void mgf_destroy(mgf_solver *s)
{
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    if (s->graph_ready) (void)hipGraphExecDestroy(s->gexec);
    if (s->arena) (void)hipFree(s->arena);
    if (s->hist) (void)hipFree(s->hist);
    if (s->ctl) (void)hipFree(s->ctl);
    if (s->ctl_host) (void)hipHostFree(s->ctl_host);
    if (s->st) (void)hipStreamDestroy(s->st);
    mgf_plan_free(s->plan);
    free(s);
}

extern "C" const char *mgf_last_error(const mgf_solver *s) { return s ? s->err : mgf_error_string(); }
extern "C" int mgf_num_levels(const mgf_solver *s) { return s ? s->ng : 0; }
extern "C" void *mgf_stream(mgf_solver *s) { return s ? (void *)s->st : NULL; }

extern "C" int mgf_level_dims(const mgf_solver *s, int ig, int *npx, int *npy)
{
    if (!s || ig < 0 || ig >= s->ng) return MGF_ERR_ARG;
    if (npx) *npx = s->lv[ig].npx;
    if (npy) *npy = s->lv[ig].npy;
    return MGF_OK;
}

/* ---------------------------------------------------------------------- */
/* schedule                                                                 */
/* ---------------------------------------------------------------------- */

/* V6:201-207 (leg 0) and V6:237-243 (leg 1) */
static void enqueue_relax(mgf_solver *s, int ig, int leg, int record)
{
    const DevLevel &L = s->lv[ig];
    hipStream_t st = s->st;
    mgf_launch_fluxes(L, s->ctl, st);
    mgf_launch_asm_u(L, s->phys, s->ctl, st);
    mgf_launch_smooth(L, 0, s->isolve, s->nswpm, s->ctl, st);
    mgf_launch_asm_v(L, s->phys, s->ctl, st);
    mgf_launch_smooth(L, 1, s->isolve, s->nswpm, s->ctl, st);
    mgf_launch_asm_p(L, s->phys, s->ctl, ig, record, s->hist, st);
    mgf_launch_smooth(L, 2, s->isolve, s->nswpp, s->ctl, st);
    mgf_launch_update(L, s->phys, s->ctl, st);
    mgf_launch_pbound(L, s->phys, s->ctl, st);
    if (leg == 0) {
        mgf_launch_outflow(L, s->phys, s->ctl, st);
        mgf_launch_asm_sc(L, s->phys, s->ctl, st);
        mgf_launch_smooth(L, 3, s->isolve, s->nswpm, s->ctl, st);
    } else {
        mgf_launch_asm_sc(L, s->phys, s->ctl, st);
        mgf_launch_smooth(L, 3, s->isolve, s->nswpm, s->ctl, st);
        mgf_launch_outflow(L, s->phys, s->ctl, st);
    }
}

static void enqueue_vcycle(mgf_solver *s)
{
    const int ng = s->ng;
    for (int ig = 0; ig < ng; ++ig) {
        enqueue_relax(s, ig, 0, ig == 0);
        if (ig != ng - 1) mgf_launch_restrict(s->lv[ig], s->lv[ig + 1], s->phys, s->ctl, s->st);
    }
    for (int ig = ng - 1; ig > 0; --ig) {
        mgf_launch_prolong(s->lv[ig - 1], s->lv[ig], s->ctl, s->st);
        if (ig != 1) enqueue_relax(s, ig - 1, 1, 0);
    }
}

static int reset_ctl(mgf_solver *s, int cyc, double tol)
{
    memset(s->ctl_host, 0, sizeof(DevCtl));
    s->ctl_host->cyc = cyc;
    s->ctl_host->tol = tol;
    HIPCHK(s, hipMemcpyAsync(s->ctl, s->ctl_host, sizeof(DevCtl), hipMemcpyHostToDevice, s->st));
    return MGF_OK;
}

static int read_ctl(mgf_solver *s)
{
    HIPCHK(s, hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, s->st));
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

/* Restore the state mgf_create built, on the device (no host transfer).
 * Everything a solve reads before writing is restored: u, v, p, sc (initial
 * and boundary values), pp and the tau terms (zero), and the momentum
 * diagonals (1e30; their ghost entries are never rewritten).  The other
 * arrays are scratch that every step rewrites before it reads them. */
extern "C" int mgf_reset(mgf_solver *s)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    for (int ig = 0; ig < s->ng; ++ig) {
        DevLevel &L = s->lv[ig];
        size_t n = (size_t)L.npx * L.npy, nb = n * sizeof(double);
        HIPCHK(s, hipMemcpyAsync(L.u, s->init_state[ig], nb, hipMemcpyDeviceToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(L.v, s->init_state[ig] + n, nb, hipMemcpyDeviceToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(L.p, s->init_state[ig] + 2 * n, nb, hipMemcpyDeviceToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync(L.sc, s->init_state[ig] + 3 * n, nb, hipMemcpyDeviceToDevice, s->st));
        mgf_launch_fill(L.pp, 0.0, n, s->st);
        mgf_launch_fill(L.resu, 0.0, n, s->st);
        mgf_launch_fill(L.resv, 0.0, n, s->st);
        mgf_launch_fill(L.ressc, 0.0, n, s->st);
        mgf_launch_fill(L.apu, 1.0e30, n, s->st);
        mgf_launch_fill(L.apv, 1.0e30, n, s->st);
    }
    return check_launch(s);
}

extern "C" int mgf_synchronize(mgf_solver *s)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

extern "C" int mgf_relax(mgf_solver *s, int ig, int leg, double *mass_err)
{
    if (!s || ig < 0 || ig >= s->ng || (leg != 0 && leg != 1))
        return set_err(s, MGF_ERR_ARG, "bad level or leg");
    int rc = reset_ctl(s, 0, 0.0);
    if (rc) return rc;
    enqueue_relax(s, ig, leg, 0);
    if ((rc = check_launch(s))) return rc;
    if ((rc = read_ctl(s))) return rc;
    if (mass_err) *mass_err = s->ctl_host->err[ig];
    return MGF_OK;
}

extern "C" int mgf_restrict(mgf_solver *s, int ig)
{
    if (!s || ig < 0 || ig + 1 >= s->ng) return set_err(s, MGF_ERR_ARG, "bad fine level");
    int rc = reset_ctl(s, 0, 0.0);
    if (rc) return rc;
    mgf_launch_restrict(s->lv[ig], s->lv[ig + 1], s->phys, s->ctl, s->st);
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

extern "C" int mgf_prolong(mgf_solver *s, int igc)
{
    if (!s || igc < 1 || igc >= s->ng) return set_err(s, MGF_ERR_ARG, "bad coarse level");
    int rc = reset_ctl(s, 0, 0.0);
    if (rc) return rc;
    mgf_launch_prolong(s->lv[igc - 1], s->lv[igc], s->ctl, s->st);
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

extern "C" int mgf_vcycle(mgf_solver *s, int icyc, double *mass_err)
{
    if (!s || icyc < 0) return set_err(s, MGF_ERR_ARG, "bad cycle index");
    int rc = ensure_hist(s, icyc + 1);
    if (rc) return rc;
    if ((rc = reset_ctl(s, icyc, 0.0))) return rc;
    enqueue_vcycle(s);
    if ((rc = check_launch(s))) return rc;
    double h = 0.0;
    HIPCHK(s, hipMemcpyAsync(&h, s->hist + icyc, sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(s, hipStreamSynchronize(s->st));
    if (mass_err) *mass_err = h;
    return MGF_OK;
}

static int capture_graph(mgf_solver *s)
{
    if (s->graph_ready) return MGF_OK;
    hipGraph_t g;
    HIPCHK(s, hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    enqueue_vcycle(s);
    mgf_launch_end_cycle(s->ctl, s->hist, s->st);
    hipError_t e = hipStreamEndCapture(s->st, &g);
    if (e != hipSuccess) return set_err(s, MGF_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&s->gexec, g, NULL, NULL, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return set_err(s, MGF_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
    s->graph_ready = 1;
    return MGF_OK;
}

extern "C" int mgf_solve(mgf_solver *s, int nvcycle, double tol, double *history, int *ncycles,
                         int *converged)
{
    if (!s || nvcycle < 0) return set_err(s, MGF_ERR_ARG, "bad arguments");
    int rc = ensure_hist(s, nvcycle > 0 ? nvcycle : 1);
    if (rc) return rc;
    if ((rc = reset_ctl(s, 0, tol))) return rc;
    if (s->opt.use_graph && (rc = capture_graph(s))) return rc;
    int launched = 0, chunk = 1;
    while (launched < nvcycle) {
        int k = chunk < nvcycle - launched ? chunk : nvcycle - launched;
        for (int t = 0; t < k; ++t) {
            if (s->opt.use_graph) {
                HIPCHK(s, hipGraphLaunch(s->gexec, s->st));
            } else {
                enqueue_vcycle(s);
                mgf_launch_end_cycle(s->ctl, s->hist, s->st);
            }
        }
        if ((rc = check_launch(s))) return rc;
        launched += k;
        if ((rc = read_ctl(s))) return rc;
        if (s->ctl_host->done) break;
        if (chunk < 32) chunk *= 2;
    }
    if ((rc = read_ctl(s))) return rc;
    int nc = s->ctl_host->cyc;
    if (history && nc > 0)
        HIPCHK(s, hipMemcpy(history, s->hist, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost));
    if (ncycles) *ncycles = nc;
    if (converged) *converged = s->ctl_host->done;
    if (s->ctl_host->nonfinite)
        return set_err(s, MGF_ERR_NONFINITE, "non-finite residual recorded");
    return MGF_OK;
}

/* ---------------------------------------------------------------------- */
/* data access                                                              */
/* ---------------------------------------------------------------------- */

static double *field_ptr(mgf_solver *s, int ig, const char *name)
{
    DevLevel &L = s->lv[ig];
    struct {
        const char *n;
        double *p;
    } tab[] = {{"u", L.u},       {"v", L.v},       {"p", L.p},         {"pp", L.pp},
               {"sc", L.sc},     {"resu", L.resu}, {"resv", L.resv},   {"ressc", L.ressc},
               {"apu", L.apu},   {"apv", L.apv}};
    for (auto &t : tab)
        if (strcmp(t.n, name) == 0) return t.p;
    return NULL;
}

extern "C" int mgf_get_field(mgf_solver *s, int ig, const char *name, double *out)
{
    if (!s || ig < 0 || ig >= s->ng || !name || !out) return set_err(s, MGF_ERR_ARG, "bad arguments");
    double *p = field_ptr(s, ig, name);
    if (!p) return set_err(s, MGF_ERR_ARG, "unknown field %s", name);
    size_t n = (size_t)s->lv[ig].npx * s->lv[ig].npy;
    HIPCHK(s, hipMemcpyAsync(out, p, n * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

extern "C" int mgf_set_field(mgf_solver *s, int ig, const char *name, const double *in)
{
    if (!s || ig < 0 || ig >= s->ng || !name || !in) return set_err(s, MGF_ERR_ARG, "bad arguments");
    double *p = field_ptr(s, ig, name);
    if (!p) return set_err(s, MGF_ERR_ARG, "unknown field %s", name);
    size_t n = (size_t)s->lv[ig].npx * s->lv[ig].npy;
    HIPCHK(s, hipMemcpyAsync(p, in, n * sizeof(double), hipMemcpyHostToDevice, s->st));
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

extern "C" int mgf_get_fields(mgf_solver *s, int ig, double *u, double *v, double *p, double *pp,
                              double *sc)
{
    const char *names[5] = {"u", "v", "p", "pp", "sc"};
    double *outs[5] = {u, v, p, pp, sc};
    for (int k = 0; k < 5; ++k)
        if (outs[k]) {
            int rc = mgf_get_field(s, ig, names[k], outs[k]);
            if (rc) return rc;
        }
    return MGF_OK;
}

/* ---------------------------------------------------------------------- */
/* measurement                                                              */
/* ---------------------------------------------------------------------- */

extern "C" int mgf_time_kernel(mgf_solver *s, int kernel, int ig, int reps, double *mean_us,
                               double *bytes_per_launch)
{
    if (!s || ig < 0 || ig >= s->ng || reps < 1) return set_err(s, MGF_ERR_ARG, "bad arguments");
    if (kernel != MGF_KERNEL_SWEEP_PP && kernel != MGF_KERNEL_SWEEP_U)
        return set_err(s, MGF_ERR_ARG, "unknown kernel id");
    DevLevel &L = s->lv[ig];
    const int field = kernel == MGF_KERNEL_SWEEP_PP ? 2 : 0;
    double *phi = field == 2 ? L.pp : L.u;
    size_t n = (size_t)L.npx * L.npy;
    double *save = NULL;
    HIPCHK(s, hipMalloc(&save, n * sizeof(double)));
    HIPCHK(s, hipMemcpyAsync(save, phi, n * sizeof(double), hipMemcpyDeviceToDevice, s->st));
    int rc = reset_ctl(s, 0, 0.0);
    if (rc) { (void)hipFree(save); return rc; }
    const int isolve = s->isolve == 1 ? 1 : 2;
    /* warm up, then time */
    mgf_launch_smooth(L, field, isolve, 1, s->ctl, s->st);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s->st);
    mgf_launch_smooth(L, field, isolve, reps, s->ctl, s->st);
    (void)hipEventRecord(e1, s->st);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipMemcpyAsync(phi, save, n * sizeof(double), hipMemcpyDeviceToDevice, s->st);
    (void)hipStreamSynchronize(s->st);
    (void)hipFree(save);
    if ((rc = check_launch(s))) return rc;
    int nx = field == 2 ? L.npx - 1 : L.npx - 2, ny = L.npy - 1;
    double nunk = (double)(nx - 1) * (double)(ny - 1);
    /* one launch per sweep (Jacobi sweep, or both red-black colours) */
    int per_sweep = 1;
    if (mean_us) *mean_us = 1000.0 * (double)ms / ((double)reps * per_sweep);
    if (bytes_per_launch) *bytes_per_launch = 8.0 * 8.0 * nunk / per_sweep;
    return MGF_OK;
}
