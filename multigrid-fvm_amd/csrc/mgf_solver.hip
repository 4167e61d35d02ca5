/*
 * mgf_solver.hip -- solver objects behind the C ABI of include/mgfvm.h.
 *
 * A solver holds one or more independent problems ("instances") on one
 * device: one HBM allocation (the "arena") with every level of every
 * instance, one HIP stream, one device control block and residual history
 * per instance, and (once a solve starts) one captured V-cycle graph that is
 * replayed until every instance's device-side stopping test has fired.
 *
 * The V-cycle is the reference's schedule, Version-6-26-04.c:195-249, run
 * for all instances at once: levels are indexed by their depth above the
 * coarsest level, every launch covers the same-depth, same-shape level of
 * every instance that has it (blockIdx.z picks the instance), and the order
 * of steps within each instance is exactly the reference's:
 *   down: relax(d) then restrict(d -> d-1), d from the finest to 0;
 *   up:   prolong(d-1 -> d), then relax(d) unless d is the instance's finest
 *         level (level 0 gets no post-relaxation, V6:235);
 *   end:  stop an instance when its level-0 sum |div| recorded in this cycle
 *         is below its tolerance (V6:247-248) or its cycle limit is reached.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "mgf_dev.h"
#include "mgf_plan.h"
#include "mgfvm.h"
#include "mgfvm_tangent.h"

struct Inst {
    mgf_plan *plan;
    int ng;
    DevLevel lv[MGF_MAX_LEVELS];          /* host copies of the descriptors */
    real_t *init_state[MGF_MAX_LEVELS];   /* u, v, p, sc as built at creation */
    int nvcycle;
    double tol;
};

enum OpKind { OP_RELAX_DOWN, OP_RELAX_UP, OP_RESTRICT, OP_PROLONG };
struct Op {
    OpKind kind;
    int g;                                /* group or pair index */
};

/* one V-cycle schedule over a set of instances: its launch groups and
 * transfer pairs (descriptor arrays on the device) and, once captured, its
 * graph */
struct Program {
    std::vector<int> members;             /* instance indices */
    int *d_members = nullptr;             /* the same on the device (stopping test) */
    DevLevel *d_desc = nullptr;           /* all group descriptor arrays */
    std::vector<DevGroup> groups;
    std::vector<DevPair> pairs;
    std::vector<int> group_pair;          /* down-leg group -> its restriction (or -1) */
    std::vector<Op> ops;                  /* one V-cycle */
    hipGraphExec_t gexec = nullptr;
    int graph_ready = 0;
};

/* one line of the tuning table: the merged relaxation's shapes on levels
 * of npx x npy points */
struct TuneEntry {
    int npx, npy, uvs, pp, nt;
};

struct mgf_solver {
    int nb = 0;
    std::vector<Inst> inst;
    mgf_options opt{};
    int device = 0;
    hipStream_t st = nullptr;
    int isolve = 0, nswpm = 0, nswpp = 0;
    int smoother = 0;                     /* isolve, or 3: the lexicographic variant */
    void *arena = nullptr;
    size_t arena_bytes = 0;
    Program full;                         /* every instance */
    std::vector<Program> lanes;           /* the running instances, one program per lane */
    std::vector<hipStream_t> lane_st;     /* lane l's stream (lane 0: st) */
    std::vector<hipEvent_t> lane_ev;      /* lane l's end of chunk, waited for on st */
    hipStream_t est = nullptr;            /* the stream the enqueue_* calls launch on */
    /* single-instance step API: group of level ig, pair with fine level ig */
    std::vector<int> level_group;
    std::vector<int> level_pair;
    DevCtl *d_ctl = nullptr;              /* nb control blocks */
    DevCtl *h_ctl = nullptr;              /* pinned mirror */
    double *d_hist = nullptr;             /* nb x hist_cap */
    int hist_cap = 0;
    /* a solver of mgf_solve_noalloc: every device array is carved out of
     * the library's static scratch (scr_cur .. scr_end), the control-block
     * mirror is ordinary host memory, and the stream is the caller's */
    int noalloc = 0;
    char *scr_cur = nullptr, *scr_end = nullptr;
    int own_stream = 1;
    void *adj = nullptr;                  /* adjoint workspace (mgf_adjoint.hip) */
    std::vector<TuneEntry> tune;          /* launch shapes per level shape */
    int tune_lanes = 0;                   /* lane count of the tuning table (0: none) */
    char err[512] = {0};
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

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

/* Device memory of a solver: hipMalloc, or for a no-alloc solver the next
 * 256-byte aligned piece of the static scratch (never given back: such a
 * solver lives for one call). */
static int dev_alloc(mgf_solver *s, void **p, size_t bytes)
{
    if (s->noalloc) {
        const size_t b = align256(bytes);
        if ((size_t)(s->scr_end - s->scr_cur) < b)
            return set_err(s, MGF_ERR_LIMIT,
                           "the problem needs more device memory than the %zu-byte no-alloc "
                           "scratch holds",
                           (size_t)MGF_NOALLOC_SCRATCH_BYTES);
        *p = s->scr_cur;
        s->scr_cur += b;
        return MGF_OK;
    }
    HIPCHK(s, hipMalloc(p, bytes));
    return MGF_OK;
}

static void dev_free(const mgf_solver *s, void *p)
{
    if (p && !s->noalloc) (void)hipFree(p);
}

/* ---------------------------------------------------------------------- */
/* arena                                                                    */
/* ---------------------------------------------------------------------- */

static size_t level_bytes(const mgf_plan_level *l, int mixed)
{
    size_t n = (size_t)l->npx * l->npy;
    /* 23 arrays + 4 initial-state copies + uc, vc, scn, pnew */
    size_t t = 31 * align256(n * sizeof(real_t));
    if (mixed) t += 11 * align256(n * sizeof(float));
    t += align256((size_t)mgf_partial_blocks(l->npx, l->npy) * sizeof(double));
    t += align256(n);                               /* flags */
    t += align256(n * sizeof(uint64_t));            /* sccode */
    t += align256(n * sizeof(int));                 /* pslot */
    t += align256((size_t)(l->npcopy + 1) * sizeof(int));
    t += align256((size_t)(l->npcopy + 1) * sizeof(real_t));
    return t;
}

/* host FP64 values <-> device arrays of the path's scalar: a plain copy when
 * that scalar is double; for dual numbers the values travel with zero
 * derivative and come back without it */
static hipError_t to_device(double *dst, const double *src, size_t n, hipStream_t st)
{
    return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, st);
}
static hipError_t to_device(mgf_dual *dst, const double *src, size_t n, hipStream_t st)
{
    std::vector<mgf_dual> h(src, src + n);
    hipError_t e = hipMemcpyAsync(dst, h.data(), n * sizeof(mgf_dual), hipMemcpyHostToDevice, st);
    return e != hipSuccess ? e : hipStreamSynchronize(st);
}
static hipError_t to_host(double *dst, const double *src, size_t n, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, st);
    return e != hipSuccess ? e : hipStreamSynchronize(st);
}
static hipError_t to_host(double *dst, const mgf_dual *src, size_t n, hipStream_t st)
{
    std::vector<mgf_dual> h(n);
    hipError_t e = hipMemcpyAsync(h.data(), src, n * sizeof(mgf_dual), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (size_t k = 0; k < n; ++k) dst[k] = h[k].v;
    return e;
}

/* a value and (in the sensitivity build) its derivative seed */
static inline void set_real(double &x, double v, double) { x = v; }
static inline void set_real(mgf_dual &x, double v, double d) { x = mgf_dual(v, d); }

/* derivative seed of u on every level: the inlet column of level 0, carried
 * to the coarse levels as rest_geom carries the values, and the east
 * boundary u (rows 1..npy-2 of every level, as init sets it and rest_geom
 * keeps it) */
static void inlet_seed(const mgf_plan *pl, const double *d_inlet, double d_u_xp,
                       std::vector<std::vector<double>> &du)
{
    du.assign((size_t)pl->ngrid, std::vector<double>());
    for (int ig = 0; ig < pl->ngrid; ++ig) {
        const mgf_plan_level *l = &pl->lv[ig];
        du[ig].assign((size_t)l->npx * l->npy, 0.0);
        for (int j = 1; j <= l->npy - 2; ++j) du[ig][(size_t)j * l->npx + l->npx - 2] = d_u_xp;
    }
    if (!d_inlet) return;
    const mgf_plan_level *l0 = &pl->lv[0];
    for (int j = 1; j <= l0->npy - 2; ++j) du[0][(size_t)j * l0->npx] = d_inlet[j - 1];
    for (int ig = 1; ig < pl->ngrid; ++ig) {
        const int nxf = pl->lv[ig - 1].npx, nxc = pl->lv[ig].npx, nyc = pl->lv[ig].npy;
        for (int jc = 1; jc < nyc - 1; ++jc) {
            const size_t ij = (size_t)(2 * jc) * nxf;
            du[ig][(size_t)jc * nxc] = 0.5 * (du[ig - 1][ij] + du[ig - 1][ij - nxf]);
        }
    }
}

static int upload_instance(mgf_solver *s, Inst &I, char *&cur, DevCtl *ctl, const mgf_seed *seed)
{
    std::vector<std::vector<double>> du;
    inlet_seed(I.plan, seed ? seed->d_inlet_u : NULL, seed ? seed->d_u_xp : 0.0, du);
    const mgf_plan *pl = I.plan;
    const mgf_problem *c = &pl->prob;
    auto take = [&](size_t bytes) {
        char *p = cur;
        cur += align256(bytes);
        return (void *)p;
    };
    const double amusc = c->amuf / c->prsc;
    DevPhys P;
    P.relxm = c->relxm;
    P.relxm1 = 1.0 - c->relxm;
    P.irelxm = 1.0 / c->relxm;
    P.relxp = c->relxp;
    P.sc_obs = c->sc_obs;
    for (int k = 0; k < 4; ++k) P.ibc[k] = c->ibc[k];
    for (int ig = 0; ig < I.ng; ++ig) {
        const mgf_plan_level *l = &pl->lv[ig];
        DevLevel &L = I.lv[ig];
        memset(&L, 0, sizeof L);
        L.npx = l->npx;
        L.npy = l->npy;
        L.ig = ig;
        L.npcopy = l->npcopy;
        L.has_obs = l->has_obs;
        L.ext = l->ext;
        L.nblk = mgf_partial_blocks(l->npx, l->npy);
        L.ax = l->ax;
        L.ay = l->ay;
        /* the reference evaluates these per point from areax/areay arrays
         * that hold exactly ax/ay; same expressions, same bits */
        /* d(amusc) = d(amuf) / prsc - amuf d(prsc) / prsc^2 */
        const double dmu = seed ? seed->d_amuf : 0.0;
        const double dmusc = seed ? seed->d_amuf / c->prsc -
                                        c->amuf * seed->d_prsc / (c->prsc * c->prsc)
                                  : 0.0;
        set_real(L.du_x, c->amuf * (l->ax / l->dx), dmu * (l->ax / l->dx));
        set_real(L.du_y, c->amuf * (l->ay / l->dy), dmu * (l->ay / l->dy));
        set_real(L.ds_x, amusc * (l->ax / l->dx), dmusc * (l->ax / l->dx));
        set_real(L.ds_y, amusc * (l->ay / l->dy), dmusc * (l->ay / l->dy));
        L.axax = l->ax * l->ax;
        L.ayay = l->ay * l->ay;
        L.bodyx = (c->dpdxm * l->ax) * l->ay;
        L.bodyy = (c->dpdym * l->ax) * l->ay;
        L.P = P;
        L.ctl = ctl;
        size_t n = (size_t)l->npx * l->npy, nb = n * sizeof(real_t);
        real_t **arrs[23] = {&L.u, &L.v, &L.p, &L.pp, &L.sc, &L.resu, &L.resv, &L.ressc,
                             &L.apu, &L.apv, &L.cx, &L.cy, &L.aw, &L.ae, &L.as, &L.an,
                             &L.ap, &L.su, &L.tmp, &L.cu, &L.cv, &L.cp, &L.csc};
        for (int k = 0; k < 23; ++k) *arrs[k] = (real_t *)take(nb);
        I.init_state[ig] = (real_t *)take(4 * nb);
        L.partial = (double *)take((size_t)L.nblk * sizeof(double));
        L.isolve = c->isolve;
        L.pset = c->nswpm >= 1;
        if (s->opt.precision == MGF_PREC_MIXED) {
            float **a32[11] = {&L.aw32, &L.ae32, &L.as32, &L.an32, &L.iap32, &L.r32,
                               &L.da32, &L.db32, &L.du32, &L.dv32, &L.ds32};
            for (int k = 0; k < 11; ++k) *a32[k] = (float *)take(n * sizeof(float));
        }
        L.uc = (real_t *)take(nb);
        L.vc = (real_t *)take(nb);
        L.scn = (real_t *)take(nb);
        L.pnew = (real_t *)take(nb);
        L.flags = (const uint8_t *)take(n);
        L.sccode = (const uint64_t *)take(n * sizeof(uint64_t));
        L.pslot = (const int *)take(n * sizeof(int));
        L.porig = (const int *)take((size_t)(l->npcopy + 1) * sizeof(int));
        L.pstash = (real_t *)take((size_t)(l->npcopy + 1) * sizeof(real_t));

        const double *src4[4] = {l->u, l->v, l->p, l->sc};
        real_t *dst4[4] = {L.u, L.v, L.p, L.sc};
        for (int k = 0; k < 4; ++k) {
            if (k == 0 && seed && (seed->d_inlet_u || seed->d_u_xp != 0.0)) {
                std::vector<real_t> h(n);
                for (size_t q = 0; q < n; ++q) set_real(h[q], l->u[q], du[ig][q]);
                HIPCHK(s, hipMemcpy(dst4[0], h.data(), nb, hipMemcpyHostToDevice));
                HIPCHK(s, hipMemcpy(I.init_state[ig], h.data(), nb, hipMemcpyHostToDevice));
                continue;
            }
            HIPCHK(s, to_device(dst4[k], src4[k], n, s->st));
            HIPCHK(s, to_device(I.init_state[ig] + k * n, src4[k], n, s->st));
        }
        HIPCHK(s, hipMemcpyAsync((void *)L.flags, l->flags, n, hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync((void *)L.sccode, l->sccode, n * sizeof(uint64_t),
                                 hipMemcpyHostToDevice, s->st));
        HIPCHK(s, hipMemcpyAsync((void *)L.pslot, l->pslot, n * sizeof(int),
                                 hipMemcpyHostToDevice, s->st));
        if (l->npcopy > 0)
            HIPCHK(s, hipMemcpyAsync((void *)L.porig, l->porig, (size_t)l->npcopy * sizeof(int),
                                     hipMemcpyHostToDevice, s->st));
        /* apu = apv = 1e30 everywhere (V6:568-569); their ghost entries are
         * never overwritten and make the p' boundary coefficients ~0 */
        mgf_launch_fill(L.apu, 1.0e30, n, s->st);
        mgf_launch_fill(L.apv, 1.0e30, n, s->st);
        int rc = check_launch(s);
        if (rc) return rc;
    }
    /* the pageable host plan arrays must stay alive until the copies finish */
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

/* ---------------------------------------------------------------------- */
/* groups and the V-cycle program                                           */
/* ---------------------------------------------------------------------- */

static void fill_group_meta(const mgf_solver *s, DevGroup &G, const std::vector<DevLevel> &m)
{
    G.count = (int)m.size();
    G.npx = m[0].npx;
    G.npy = m[0].npy;
    G.uvs_shape = G.pp_shape = G.nt_asm = 0;
    for (const TuneEntry &t : s->tune)
        if (t.npx == G.npx && t.npy == G.npy) {
            G.uvs_shape = t.uvs;
            G.pp_shape = t.pp;
            G.nt_asm = t.nt;
        }
    G.max_npcopy = 0;
    G.any_pbound_x = G.any_pbound_y = G.any_outflow = 0;
    G.any_finest = 0;
    for (const DevLevel &L : m) {
        if (L.ig == 0) G.any_finest = 1;
        if (L.npcopy > G.max_npcopy) G.max_npcopy = L.npcopy;
        if (L.P.ibc[0] == 3) G.any_pbound_x = 1;
        if (L.P.ibc[2] == 3) G.any_pbound_y = 1;
        if (L.P.ibc[1] == 4) G.any_outflow = 1;
    }
}

/* Groups: at every depth d above the coarsest level, the levels of all
 * instances that reach that depth, split by shape (kind 0: every such level,
 * relaxed on the way down; kind 1: those that are not their instance's
 * finest level, relaxed again on the way up).  Pairs: the same-shape levels
 * at depth d >= 1 with their next coarser levels, for the transfers. */
static int build_program(mgf_solver *s, Program &P, const std::vector<int> &members)
{
    int D = 0;
    for (int b : members) D = s->inst[b].ng - 1 > D ? s->inst[b].ng - 1 : D;
    struct Key {
        int d, nx, ny, kind;
        bool operator==(const Key &o) const {
            return d == o.d && nx == o.nx && ny == o.ny && kind == o.kind;
        }
    };
    std::vector<std::vector<DevLevel>> gm;              /* group members */
    std::vector<std::vector<DevLevel>> pf, pc;          /* pair members */
    std::vector<std::pair<Key, int>> gkey, pkey;
    auto find = [](const std::vector<std::pair<Key, int>> &v, const Key &k) {
        for (const auto &e : v)
            if (e.first == k) return e.second;
        return -1;
    };
    std::vector<std::vector<int>> down_g(D + 1), up_g(D + 1), pr(D + 1);
    for (int d = D; d >= 0; --d) {
        for (int b : members) {
            const Inst &I = s->inst[b];
            const int ig = I.ng - 1 - d;
            if (ig < 0) continue;
            const DevLevel &L = I.lv[ig];
            for (int kind = 0; kind < 2; ++kind) {
                if (kind == 1 && ig == 0) continue;
                Key k{d, L.npx, L.npy, kind};
                int g = find(gkey, k);
                if (g < 0) {
                    g = (int)gm.size();
                    gm.emplace_back();
                    gkey.push_back({k, g});
                    (kind == 0 ? down_g : up_g)[d].push_back(g);
                }
                gm[g].push_back(L);
            }
            if (d >= 1) {
                Key k{d, L.npx, L.npy, 2};
                int p = find(pkey, k);
                if (p < 0) {
                    p = (int)pf.size();
                    pf.emplace_back();
                    pc.emplace_back();
                    pkey.push_back({k, p});
                    pr[d].push_back(p);
                }
                pf[p].push_back(L);
                pc[p].push_back(I.lv[ig + 1]);
            }
        }
    }
    std::vector<Op> ops;
    for (int d = D; d >= 0; --d) {
        for (int g : down_g[d]) ops.push_back({OP_RELAX_DOWN, g});
        if (d >= 1)
            for (int p : pr[d]) ops.push_back({OP_RESTRICT, p});
    }
    for (int d = 1; d <= D; ++d) {
        for (int p : pr[d]) ops.push_back({OP_PROLONG, p});
        for (int g : up_g[d]) ops.push_back({OP_RELAX_UP, g});
    }
    size_t ndesc = 0;
    for (auto &g : gm) ndesc += g.size();
    for (auto &p : pf) ndesc += 2 * p.size();
    P.members = members;
    int arc;
    if ((arc = dev_alloc(s, (void **)&P.d_members, (members.size() + 1) * sizeof(int)))) return arc;
    HIPCHK(s, hipMemcpy(P.d_members, members.data(), members.size() * sizeof(int),
                        hipMemcpyHostToDevice));
    if ((arc = dev_alloc(s, (void **)&P.d_desc, ndesc * sizeof(DevLevel)))) return arc;
    std::vector<DevLevel> host;
    host.reserve(ndesc);
    P.groups.assign(gm.size(), DevGroup());
    P.pairs.assign(pf.size(), DevPair());
    for (const auto &e : gkey) {
        P.groups[e.second].depth = e.first.d;
        P.groups[e.second].up = e.first.kind;
    }
    for (size_t g = 0; g < gm.size(); ++g) {
        DevGroup &G = P.groups[g];
        G.d = P.d_desc + host.size();
        fill_group_meta(s, G, gm[g]);
        host.insert(host.end(), gm[g].begin(), gm[g].end());
    }
    for (size_t p = 0; p < pf.size(); ++p) {
        DevPair &R = P.pairs[p];
        R.count = (int)pf[p].size();
        R.npxf = pf[p][0].npx;
        R.npyf = pf[p][0].npy;
        R.npxc = pc[p][0].npx;
        R.npyc = pc[p][0].npy;
        for (const DevLevel &c : pc[p])
            if (c.npx != R.npxc || c.npy != R.npyc)
                return set_err(s, MGF_ERR_GEOMETRY, "coarse shapes differ within a transfer group");
        R.f = P.d_desc + host.size();
        host.insert(host.end(), pf[p].begin(), pf[p].end());
        R.c = P.d_desc + host.size();
        host.insert(host.end(), pc[p].begin(), pc[p].end());
    }
    HIPCHK(s, hipMemcpy(P.d_desc, host.data(), host.size() * sizeof(DevLevel),
                        hipMemcpyHostToDevice));
    P.ops = ops;
    /* the restriction of each down-leg group: the pair with its key (same
     * levels, same order) */
    P.group_pair.assign(gm.size(), -1);
    for (const auto &e : gkey)
        if (e.first.kind == 0 && e.first.d >= 1)
            P.group_pair[e.second] = find(pkey, Key{e.first.d, e.first.nx, e.first.ny, 2});
    if (&P == &s->full && s->nb == 1) {
        const Inst &I = s->inst[0];
        s->level_group.assign(I.ng, -1);
        s->level_pair.assign(I.ng, -1);
        for (int ig = 0; ig < I.ng; ++ig) {
            const int d = I.ng - 1 - ig;
            s->level_group[ig] = down_g[d][0];
            if (d >= 1) s->level_pair[ig] = pr[d][0];
        }
    }
    return MGF_OK;
}

/* ---------------------------------------------------------------------- */
/* creation                                                                 */
/* ---------------------------------------------------------------------- */

static int ensure_hist(mgf_solver *s, int need)
{
    if (need <= s->hist_cap) return MGF_OK;
    const int cap = need > 4096 || s->noalloc ? need : 4096;
    double *h = NULL;
    HIPCHK(s, hipStreamSynchronize(s->st));
    int rc = dev_alloc(s, (void **)&h, (size_t)cap * s->nb * sizeof(double));
    if (rc) return rc;
    HIPCHK(s, hipMemsetAsync(h, 0, (size_t)cap * s->nb * sizeof(double), s->st));
    HIPCHK(s, hipStreamSynchronize(s->st));
    dev_free(s, s->d_hist);
    s->d_hist = h;
    s->hist_cap = cap;
    return MGF_OK;
}

extern "C" int mgf_batch_create(const mgf_problem *const *probs, int n, const mgf_options *opt,
                                mgf_solver **out)
{
    return mgf_batch_create_seeded(probs, n, opt, out, NULL);
}

/* the derivative seeds exist only where the scalar carries derivatives */
static bool seeds_supported(const double &) { return false; }
static bool seeds_supported(const mgf_dual &) { return true; }

/* the tuning table's text (mgf_autotune's format): "lanes N" and
 * "level NPX NPY UVS PP THREADS" lines, '#' comments */
static int parse_tuning(mgf_solver *s, const char *text)
{
    s->tune.clear();
    s->tune_lanes = 0;
    const char *p = text;
    int line = 0;
    while (*p) {
        const char *e = strchr(p, '\n');
        const size_t len = e ? (size_t)(e - p) : strlen(p);
        char buf[256];
        const size_t n = len < sizeof buf - 1 ? len : sizeof buf - 1;
        memcpy(buf, p, n);
        buf[n] = 0;
        ++line;
        char *hash = strchr(buf, '#');
        if (hash) *hash = 0;
        TuneEntry t;
        int nl;
        char word[16] = "";
        if (sscanf(buf, "%15s", word) == 1) {
            if (!strcmp(word, "lanes") && sscanf(buf, "%*s %d", &nl) == 1 && nl >= 1 &&
                nl <= 4) {
                s->tune_lanes = nl;
            } else if (!strcmp(word, "level") &&
                       sscanf(buf, "%*s %d %d %d %d %d", &t.npx, &t.npy, &t.uvs, &t.pp, &t.nt) ==
                           5 &&
                       t.uvs >= 0 && t.uvs <= 2 && t.pp >= 0 && t.pp <= 2 &&
                       (t.nt == 0 || t.nt == 256 || t.nt == 512 || t.nt == 1024)) {
                s->tune.push_back(t);
            } else {
                return set_err(s, MGF_ERR_PARSE, "tuning table line %d not understood", line);
            }
        }
        p = e ? e + 1 : p + len;
    }
    return MGF_OK;
}

/* scratch (or null): build a no-alloc solver whose device arrays are carved
 * out of [scratch, scratch + scratch_bytes) and which launches on the
 * caller's stream cst */
static int create_solver(const mgf_problem *const *probs, int n, const mgf_options *opt,
                         mgf_solver **out, const mgf_seed *seeds, char *scratch,
                         size_t scratch_bytes, hipStream_t cst);

int mgf_batch_create_seeded(const mgf_problem *const *probs, int n, const mgf_options *opt,
                            mgf_solver **out, const mgf_seed *seeds)
{
    return create_solver(probs, n, opt, out, seeds, nullptr, 0, nullptr);
}

static int create_solver(const mgf_problem *const *probs, int n, const mgf_options *opt,
                         mgf_solver **out, const mgf_seed *seeds, char *scratch,
                         size_t scratch_bytes, hipStream_t cst)
{
    if (!probs || n < 1 || !out) return set_err(NULL, MGF_ERR_ARG, "bad arguments");
    const real_t probe_scalar = 0.0;
    if (seeds && !seeds_supported(probe_scalar))
        return set_err(NULL, MGF_ERR_UNSUPPORTED, "derivative seeds need the sensitivity library");
    *out = NULL;
    mgf_options o;
    if (opt)
        o = *opt;
    else
        mgf_options_default(&o);
    if (o.precision != MGF_PREC_FP64 && o.precision != MGF_PREC_MIXED)
        return set_err(NULL, MGF_ERR_UNSUPPORTED, "precision mode %d not available", o.precision);
    if (o.precision == MGF_PREC_MIXED && seeds_supported(probe_scalar))
        return set_err(NULL, MGF_ERR_UNSUPPORTED,
                       "the sensitivity library runs the FP64 path only");
    mgf_solver *s = new mgf_solver();
    s->opt = o;
    s->nb = n;
    if (scratch) {
        s->noalloc = 1;
        s->scr_cur = scratch;
        s->scr_end = scratch + scratch_bytes;
    }
    s->inst.resize(n);
    for (Inst &I : s->inst) {
        I.plan = NULL;
        I.ng = 0;
    }
    char e[512] = {0};
    int rc = MGF_OK;
    for (int b = 0; b < n; ++b) {
        if (!probs[b]) {
            mgf_destroy(s);
            return set_err(NULL, MGF_ERR_ARG, "problem %d is null", b);
        }
        rc = mgf_plan_build(probs[b], o.strict_limits, &s->inst[b].plan, e, sizeof e);
        if (rc) {
            mgf_destroy(s);
            if (n > 1) return set_err(NULL, rc, "problem %d: %s", b, e);
            return set_err(NULL, rc, "%s", e);
        }
        Inst &I = s->inst[b];
        I.ng = I.plan->ngrid;
        I.nvcycle = I.plan->prob.nvcycle;
        I.tol = I.plan->prob.convergence;
        if (I.ng > MGF_MAX_LEVELS) {
            mgf_destroy(s);
            return set_err(NULL, MGF_ERR_LIMIT, "more than %d levels", MGF_MAX_LEVELS);
        }
        const mgf_problem *c = &I.plan->prob;
        if (b == 0) {
            s->isolve = c->isolve;
            s->nswpm = c->nswpm;
            s->nswpp = c->nswpp;
        } else if (c->isolve != s->isolve || c->nswpm != s->nswpm || c->nswpp != s->nswpp) {
            mgf_destroy(s);
            return set_err(NULL, MGF_ERR_CONFIG,
                           "batched problems must share isolve, nswpm and nswpp");
        }
    }
    s->smoother = s->isolve;
    if (o.gs_order != 0 && o.gs_order != 1) {
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_ARG, "gs_order must be 0 (red-black) or 1 (lexicographic)");
    }
    if (o.gs_order == 1 && s->isolve == 2) {
        if (o.precision != MGF_PREC_FP64) {
            mgf_destroy(s);
            return set_err(NULL, MGF_ERR_UNSUPPORTED,
                           "the lexicographic Gauss-Seidel variant runs in FP64 only");
        }
        for (const Inst &I : s->inst)
            if (I.plan->prob.nfy > LEX_MAX_ROWS) {
                mgf_destroy(s);
                return set_err(NULL, MGF_ERR_UNSUPPORTED,
                               "the lexicographic Gauss-Seidel variant needs nfy <= %d",
                               LEX_MAX_ROWS);
            }
        s->smoother = 3;
    }
    if (o.tuning && (rc = parse_tuning(s, o.tuning))) {
        char keep[512];
        snprintf(keep, sizeof keep, "%s", s->err);
        mgf_destroy(s);
        return set_err(NULL, rc, "%s", keep);
    }
    s->opt.tuning = nullptr;              /* the caller's text need not outlive the call */
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_DEVICE, "no HIP device available");
    }
    if (o.device >= 0 && hipSetDevice(o.device) != hipSuccess) {
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_DEVICE, "cannot select device %d", o.device);
    }
    (void)hipGetDevice(&s->device);
    if (s->noalloc) {
        s->st = cst;
        s->own_stream = 0;
    } else if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
        s->st = nullptr;
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_DEVICE, "stream creation failed");
    }
    s->est = s->st;
    s->lane_st.assign(1, s->st);
    s->lane_ev.assign(1, nullptr);
    size_t total = 0;
    for (Inst &I : s->inst)
        for (int ig = 0; ig < I.ng; ++ig)
            total += level_bytes(&I.plan->lv[ig], o.precision == MGF_PREC_MIXED);
    s->arena_bytes = total;
    if (s->noalloc) {
        /* everything a solve of these problems carves, checked before any
         * device work: the arena, the control blocks, the histories and
         * (bounded by) the launch descriptors of every level in a group and
         * a transfer pair */
        size_t ndesc = 0;
        int maxnv0 = 1;
        for (Inst &I : s->inst) {
            ndesc += 3 * (size_t)I.ng;
            maxnv0 = I.nvcycle > maxnv0 ? I.nvcycle : maxnv0;
        }
        const size_t need = align256(total) + align256((size_t)n * sizeof(DevCtl)) +
                            align256((size_t)maxnv0 * n * sizeof(double)) +
                            align256(ndesc * sizeof(DevLevel)) + align256((n + 1) * sizeof(int));
        if (need > scratch_bytes) {
            mgf_destroy(s);
            return set_err(NULL, MGF_ERR_LIMIT,
                           "the problem needs %zu bytes of device memory; the no-alloc scratch "
                           "holds %zu",
                           need, scratch_bytes);
        }
    }
    if ((rc = dev_alloc(s, &s->arena, total)) ||
        (rc = dev_alloc(s, (void **)&s->d_ctl, (size_t)n * sizeof(DevCtl)))) {
        char keep[512];
        snprintf(keep, sizeof keep, "%s", s->err);
        mgf_destroy(s);
        return set_err(NULL, rc, "%s", keep);
    }
    if (s->noalloc)
        s->h_ctl = new DevCtl[n];
    else if (hipHostMalloc(&s->h_ctl, (size_t)n * sizeof(DevCtl), 0) != hipSuccess) {
        s->h_ctl = nullptr;
        mgf_destroy(s);
        return set_err(NULL, MGF_ERR_DEVICE, "control block allocation failed");
    }
    (void)hipMemsetAsync(s->arena, 0, total, s->st);
    int maxnv = 1;
    for (Inst &I : s->inst) maxnv = I.nvcycle > maxnv ? I.nvcycle : maxnv;
    if ((rc = ensure_hist(s, maxnv)) == MGF_OK) {
        char *cur = (char *)s->arena;
        for (int b = 0; b < n && !rc; ++b)
            rc = upload_instance(s, s->inst[b], cur, s->d_ctl + b, seeds ? &seeds[b] : NULL);
    }
    if (!rc) {
        std::vector<int> all(n);
        for (int b = 0; b < n; ++b) all[b] = b;
        rc = build_program(s, s->full, all);
    }
    if (!rc) {
        memset(s->h_ctl, 0, (size_t)n * sizeof(DevCtl));
        for (int b = 0; b < n; ++b) s->h_ctl[b].hist = s->d_hist + (size_t)b * s->hist_cap;
        if (hipMemcpy(s->d_ctl, s->h_ctl, (size_t)n * sizeof(DevCtl), hipMemcpyHostToDevice) !=
            hipSuccess)
            rc = set_err(s, MGF_ERR_DEVICE, "control block upload failed");
    }
    if (rc) {
        char keep[512];
        snprintf(keep, sizeof keep, "%s", s->err);
        mgf_destroy(s);
        return set_err(NULL, rc, "%s", keep);
    }
    *out = s;
    return MGF_OK;
}

extern "C" int mgf_create(const mgf_problem *p, const mgf_options *opt, mgf_solver **out)
{
    if (!p || !out) return set_err(NULL, MGF_ERR_ARG, "null argument");
    const mgf_problem *one[1] = {p};
    return mgf_batch_create(one, 1, opt, out);
}

/* ---------------------------------------------------------------------- */
/* the solve that allocates nothing                                         */
/* ---------------------------------------------------------------------- */

/* The device memory of mgf_solve_noalloc: one static array per device,
 * held until the process ends.  In host code `scratch` names a host-side
 * stand-in, not the device array: its device address comes from
 * hipGetSymbolAddress, once per device with that device current, and only
 * that address is ever handed to a kernel or a copy. */
static __device__ __align__(256) unsigned char scratch[MGF_NOALLOC_SCRATCH_BYTES];

#define MGF_NOALLOC_MAX_DEVICES 64
static void *g_scratch_addr[MGF_NOALLOC_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

static int scratch_address(int dev, char **out)
{
    if (dev < 0 || dev >= MGF_NOALLOC_MAX_DEVICES)
        return set_err(NULL, MGF_ERR_LIMIT, "device %d beyond the no-alloc device table", dev);
    pthread_mutex_lock(&g_scratch_mu);
    int rc = MGF_OK;
    if (!g_scratch_addr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            rc = set_err(NULL, MGF_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                         hipGetErrorString(e));
        else
            g_scratch_addr[dev] = ptr;
    }
    *out = (char *)g_scratch_addr[dev];
    pthread_mutex_unlock(&g_scratch_mu);
    return rc;
}

extern "C" int mgf_solve_noalloc(const mgf_problem *p, const mgf_options *opt, void *stream,
                                 double *history, int *ncycles, int *converged, double *u,
                                 double *v, double *pr, double *sc)
{
    if (!p) return set_err(NULL, MGF_ERR_ARG, "null problem");
    mgf_options o;
    if (opt)
        o = *opt;
    else
        mgf_options_default(&o);
    /* eager V-cycles on one stream: capturing and instantiating a graph, and
     * lane streams, would allocate */
    o.use_graph = 0;
    o.lanes = 1;
    o.compact = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(NULL, MGF_ERR_DEVICE, "no HIP device available");
    if (o.device >= 0 && hipSetDevice(o.device) != hipSuccess)
        return set_err(NULL, MGF_ERR_DEVICE, "cannot select device %d", o.device);
    int dev = 0;
    (void)hipGetDevice(&dev);
    o.device = dev;
    char *base = nullptr;
    int rc = scratch_address(dev, &base);
    if (rc) return rc;
    const mgf_problem *one[1] = {p};
    mgf_solver *s = nullptr;
    rc = create_solver(one, 1, &o, &s, nullptr, base, MGF_NOALLOC_SCRATCH_BYTES,
                       (hipStream_t)stream);
    if (rc) return rc;
    rc = mgf_solve(s, p->nvcycle, p->convergence, history, ncycles, converged);
    if (rc == MGF_OK || rc == MGF_ERR_NONFINITE) {
        const char *names[4] = {"u", "v", "p", "sc"};
        double *outs[4] = {u, v, pr, sc};
        for (int k = 0; k < 4; ++k) {
            if (!outs[k]) continue;
            const int r2 = mgf_get_field(s, 0, names[k], outs[k]);
            if (r2) {
                rc = r2;
                break;
            }
        }
    }
    if (rc) {
        char keep[512];
        snprintf(keep, sizeof keep, "%s", s->err);
        (void)hipStreamSynchronize(s->st);
        mgf_destroy(s);
        return set_err(NULL, rc, "%s", keep);
    }
    (void)hipStreamSynchronize(s->st);
    mgf_destroy(s);
    return MGF_OK;
}

/* the caller has synchronised the stream the program's graph ran on */
static void release_program(const mgf_solver *s, Program &P)
{
    if (P.graph_ready) (void)hipGraphExecDestroy(P.gexec);
    dev_free(s, P.d_desc);
    dev_free(s, P.d_members);
    P = Program();
}

extern "C" void mgf_destroy(mgf_solver *s)
{
    if (!s) return;
    for (hipStream_t l : s->lane_st)
        if (l) (void)hipStreamSynchronize(l);
    for (Program &P : s->lanes) release_program(s, P);
    release_program(s, s->full);
    for (size_t l = 1; l < s->lane_st.size(); ++l) (void)hipStreamDestroy(s->lane_st[l]);
    for (hipEvent_t e : s->lane_ev)
        if (e) (void)hipEventDestroy(e);
    dev_free(s, s->arena);
    dev_free(s, s->d_hist);
    dev_free(s, s->d_ctl);
    if (s->h_ctl) {
        if (s->noalloc)
            delete[] s->h_ctl;
        else
            (void)hipHostFree(s->h_ctl);
    }
#ifndef MGF_TANGENT
    if (s->adj) mgf_adjoint_release(s->adj);
#endif
    if (s->st && s->own_stream) (void)hipStreamDestroy(s->st);
    for (Inst &I : s->inst) mgf_plan_free(I.plan);
    delete s;
}

extern "C" const char *mgf_last_error(const mgf_solver *s) { return s ? s->err : mgf_error_string(); }
extern "C" int mgf_batch_size(const mgf_solver *s) { return s ? s->nb : 0; }
extern "C" void *mgf_stream(mgf_solver *s) { return s ? (void *)s->st : NULL; }

extern "C" int mgf_batch_num_levels(const mgf_solver *s, int b)
{
    return (s && b >= 0 && b < s->nb) ? s->inst[b].ng : 0;
}

extern "C" int mgf_num_levels(const mgf_solver *s) { return mgf_batch_num_levels(s, 0); }

extern "C" int mgf_batch_level_dims(const mgf_solver *s, int b, int ig, int *npx, int *npy)
{
    if (!s || b < 0 || b >= s->nb || ig < 0 || ig >= s->inst[b].ng) return MGF_ERR_ARG;
    if (npx) *npx = s->inst[b].lv[ig].npx;
    if (npy) *npy = s->inst[b].lv[ig].npy;
    return MGF_OK;
}

extern "C" int mgf_level_dims(const mgf_solver *s, int ig, int *npx, int *npy)
{
    return mgf_batch_level_dims(s, 0, ig, npx, npy);
}

/* ---------------------------------------------------------------------- */
/* schedule                                                                 */
/* ---------------------------------------------------------------------- */

/* V6:201-207 (leg 0) and V6:237-243 (leg 1) for a group.  err: whether the
 * level's sum |div| is reduced (1) and recorded as the history entry of its
 * level-0 members (2); masserror computes it on every call, but only the
 * value of the finest level is recorded and read (V6:716-723, 248), so a
 * V-cycle reduces it where it is recorded and the step call where it
 * returns it. */
/* mixed precision: the same steps with FP64 assembly + residual, FP32
 * correction sweeps and FP64 apply (mgf_kernels_mixed.hip) */
static int enqueue_relax_mixed(mgf_solver *s, const DevGroup &G, int leg, int err,
                               const DevPair *R)
{
    hipStream_t st = s->est;
    const SweepCfg cfg = mgf_sweep_cfg(s->opt.sweeps_per_launch, s->opt.sweep_threads,
                                       s->opt.split_assembly);
    /* u, v and sc in one launch where the configuration allows: sc's
     * equation reads only the fluxes (those of the u and v the relaxation
     * starts from), sc itself and its penalty / tau terms, none of which
     * the update or the periodic copies write; its correction is applied by
     * the update */
    const int merged = mgf_launch_relax_merged_mixed(G, s->isolve, s->nswpm, s->nswpp, cfg, err,
                                                     leg == 0 ? R : nullptr, st);
    if (merged) {
        mgf_launch_pbound(G, st);
        return merged == 2;
    }
    mgf_launch_fluxes(G, st);
    mgf_launch_relax_field_mixed(G, 0, s->isolve, s->nswpm, cfg, 0, st);
    mgf_launch_relax_field_mixed(G, 1, s->isolve, s->nswpm, cfg, 0, st);
    mgf_launch_relax_field_mixed(G, 2, s->isolve, s->nswpp, cfg, err, st);
    mgf_launch_update(G, 1, st);
    mgf_launch_pbound(G, st);
    if (leg == 0) {
        mgf_launch_outflow(G, st);
        mgf_launch_relax_field_mixed(G, 3, s->isolve, s->nswpm, cfg, 0, st);
    } else {
        mgf_launch_relax_field_mixed(G, 3, s->isolve, s->nswpm, cfg, 0, st);
        mgf_launch_outflow(G, st);
    }
    return 0;
}

/* R: the restriction that follows a down-leg relaxation (or null); returns
 * 1 when the relaxation did its restv as well */
static int enqueue_relax(mgf_solver *s, const DevGroup &G, int leg, int err,
                         const DevPair *R = nullptr)
{
    if (s->opt.precision == MGF_PREC_MIXED) return enqueue_relax_mixed(s, G, leg, err, R);
    hipStream_t st = s->est;
    /* pend: u and v each in one launch whose result stays in uc, vc for the
     * p' assembly and the update (no copy-back launches); 2: also sc, before
     * p' (its equation reads the fluxes of the u, v the relaxation starts
     * from, sc and its own terms, none of which u, v or p' change), and the
     * fluxes formed inside the assemblies */
    int pend = mgf_relax64_uv_pending(G, s->smoother, s->nswpm);
    if (pend && !G.any_outflow) pend = 2;
    if (pend != 2) mgf_launch_fluxes(G, st);
    if (pend == 2 && mgf_launch_relax64_uvs(G, s->smoother, s->nswpm, st)) {
        /* u, v and sc in one launch */
    } else {
        if (!mgf_launch_relax64(G, 0, s->smoother, s->nswpm, st, pend)) {
            mgf_launch_asm_u(G, st);
            mgf_launch_smooth(G, 0, s->smoother, s->nswpm, st);
        }
        if (!mgf_launch_relax64(G, 1, s->smoother, s->nswpm, st, pend)) {
            mgf_launch_asm_v(G, st);
            mgf_launch_smooth(G, 1, s->smoother, s->nswpm, st);
        }
        if (pend == 2) (void)mgf_launch_relax64(G, 3, s->smoother, s->nswpm, st, pend);
    }
    int stashed = 0;
    if (!mgf_launch_relax64_p(G, s->smoother, s->nswpp, err, pend != 0, st,
                              pend == 2 ? &stashed : nullptr)) {
        mgf_launch_asm_p(G, err, st, pend != 0);
        mgf_launch_smooth(G, 2, s->smoother, s->nswpp, st);
    }
    const int upd = pend == 2 ? (stashed ? 5 : 4) : pend ? 3 : 0;
    if (pend == 2 && R && !G.any_pbound_x && !G.any_pbound_y) {
        /* nothing runs between the update and restv (no periodic copies, no
         * outflow ghosts, sc already relaxed): one launch does both */
        mgf_launch_update_restv(G, *R, upd, st);
        return 1;
    }
    mgf_launch_update(G, upd, st);
    mgf_launch_pbound(G, st);
    if (leg == 0) mgf_launch_outflow(G, st);
    if (pend != 2 && !mgf_launch_relax64(G, 3, s->smoother, s->nswpm, st)) {
        mgf_launch_asm_sc(G, st);
        mgf_launch_smooth(G, 3, s->smoother, s->nswpm, st);
    }
    if (leg == 1) mgf_launch_outflow(G, st);
    return 0;
}

static void enqueue_vcycle(mgf_solver *s, const Program &P)
{
    /* pairs whose restv the preceding relaxation's update already did */
    std::vector<char> restv_done(P.pairs.size(), 0);
    for (const Op &o : P.ops) {
        switch (o.kind) {
        case OP_RELAX_DOWN: {
            const DevGroup &G = P.groups[o.g];
            const int pr = P.group_pair[o.g];
            if (enqueue_relax(s, G, 0, G.any_finest ? 2 : 0, pr >= 0 ? &P.pairs[pr] : nullptr))
                restv_done[pr] = 1;
            break;
        }
        case OP_RELAX_UP: enqueue_relax(s, P.groups[o.g], 1, 0); break;
        case OP_RESTRICT: mgf_launch_restrict(P.pairs[o.g], !restv_done[o.g], s->est); break;
        case OP_PROLONG: mgf_launch_prolong(P.pairs[o.g], 1, s->est); break;
        }
    }
}

/* control blocks: every instance starts counting at cycle `cyc`, with its
 * tolerance and cycle limit (a limit <= cyc marks it finished) */
static int reset_ctl(mgf_solver *s, int cyc, const double *tol, const int *limit)
{
    for (int b = 0; b < s->nb; ++b) {
        DevCtl &c = s->h_ctl[b];
        memset(&c, 0, sizeof c);
        c.cyc = cyc;
        c.tol = tol ? tol[b] : 0.0;
        c.maxcyc = limit ? limit[b] : 0x7fffffff;
        c.done = c.maxcyc <= cyc;
        c.hist = s->d_hist + (size_t)b * s->hist_cap;
    }
    HIPCHK(s, hipMemcpyAsync(s->d_ctl, s->h_ctl, (size_t)s->nb * sizeof(DevCtl),
                             hipMemcpyHostToDevice, s->st));
    return MGF_OK;
}

static int read_ctl(mgf_solver *s)
{
    HIPCHK(s, hipMemcpyAsync(s->h_ctl, s->d_ctl, (size_t)s->nb * sizeof(DevCtl),
                             hipMemcpyDeviceToHost, s->st));
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

/* Restore the state mgf_create built, on the device (no host transfer).
 * Everything a solve reads before writing is restored: u, v, p, sc (initial
 * and boundary values), pp and the tau terms (zero), and the momentum
 * diagonals (1e30; their ghost entries are never rewritten).  The other
 * arrays are scratch that every step rewrites before it reads them. */
static int reset_instance(mgf_solver *s, Inst &I)
{
    for (int ig = 0; ig < I.ng; ++ig) {
        DevLevel &L = I.lv[ig];
        size_t n = (size_t)L.npx * L.npy, nb = n * sizeof(real_t);
        real_t *dst4[4] = {L.u, L.v, L.p, L.sc};
        for (int k = 0; k < 4; ++k)
            HIPCHK(s, hipMemcpyAsync(dst4[k], I.init_state[ig] + k * n, nb,
                                     hipMemcpyDeviceToDevice, s->st));
        mgf_launch_fill(L.pp, 0.0, n, s->st);
        mgf_launch_fill(L.resu, 0.0, n, s->st);
        mgf_launch_fill(L.resv, 0.0, n, s->st);
        mgf_launch_fill(L.ressc, 0.0, n, s->st);
        mgf_launch_fill(L.apu, 1.0e30, n, s->st);
        mgf_launch_fill(L.apv, 1.0e30, n, s->st);
    }
    return MGF_OK;
}

extern "C" int mgf_reset(mgf_solver *s)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    for (Inst &I : s->inst) {
        int rc = reset_instance(s, I);
        if (rc) return rc;
    }
    return check_launch(s);
}

/* A new inlet profile for problem b: its initial state is built again from
 * its problem with inlet_u replaced (init and rest_geom, V6:546-676,
 * 1391-1473: the level-0 west face and the coarse inlet values), and the
 * problem is reset to it, as if the solver had been created with it. */
extern "C" int mgf_set_inlet_profile(mgf_solver *s, int b, const double *inlet_u)
{
    if (!s || b < 0 || b >= s->nb) return set_err(s, MGF_ERR_ARG, "bad solver or problem index");
#ifdef MGF_TANGENT
    (void)inlet_u;
    return set_err(s, MGF_ERR_UNSUPPORTED,
                   "the sensitivity library takes the inlet profile (with its seed) at creation");
#else
    Inst &I = s->inst[b];
    mgf_problem q;
    mgf_problem_copy(&I.plan->prob, &q);
    free((void *)q.inlet_u);
    q.inlet_u = NULL;
    if (inlet_u) {
        double *c = (double *)malloc(sizeof(double) * (size_t)q.nfy);
        if (!c) {
            mgf_problem_release(&q);
            return set_err(s, MGF_ERR_LIMIT, "out of host memory");
        }
        memcpy(c, inlet_u, sizeof(double) * (size_t)q.nfy);
        q.inlet_u = c;
    }
    mgf_plan *np = NULL;
    char e[256] = "";
    int rc = mgf_plan_build(&q, s->opt.strict_limits, &np, e, sizeof e);
    mgf_problem_release(&q);
    if (rc) return set_err(s, rc, "%s", e);
    for (int ig = 0; ig < I.ng; ++ig) {
        const mgf_plan_level *l = &np->lv[ig];
        const size_t n = (size_t)l->npx * l->npy;
        const double *src4[4] = {l->u, l->v, l->p, l->sc};
        for (int k = 0; k < 4; ++k) {
            const hipError_t he = to_device(I.init_state[ig] + k * n, src4[k], n, s->st);
            if (he != hipSuccess) {
                (void)hipStreamSynchronize(s->st);
                mgf_plan_free(np);
                return set_err(s, MGF_ERR_DEVICE, "%s", hipGetErrorString(he));
            }
        }
    }
    /* the pageable host arrays must stay alive until the copies finish */
    const hipError_t he = hipStreamSynchronize(s->st);
    if (he != hipSuccess) {
        mgf_plan_free(np);
        return set_err(s, MGF_ERR_DEVICE, "%s", hipGetErrorString(he));
    }
    mgf_plan_free(I.plan);
    I.plan = np;
    if ((rc = reset_instance(s, I))) return rc;
    return check_launch(s);
#endif
}

extern "C" int mgf_synchronize(mgf_solver *s)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

static int single_only(mgf_solver *s)
{
    if (s->nb != 1)
        return set_err(s, MGF_ERR_UNSUPPORTED, "step-level calls need a single-problem solver");
    return MGF_OK;
}

extern "C" int mgf_relax(mgf_solver *s, int ig, int leg, double *mass_err)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    int rc = single_only(s);
    if (rc) return rc;
    if (ig < 0 || ig >= s->inst[0].ng || (leg != 0 && leg != 1))
        return set_err(s, MGF_ERR_ARG, "bad level or leg");
    if ((rc = reset_ctl(s, 0, NULL, NULL))) return rc;
    /* a leg-0 relaxation of level 0 also writes history slot 0 (unused here) */
    enqueue_relax(s, s->full.groups[s->level_group[ig]], leg, leg == 0 ? 2 : 1);
    if ((rc = check_launch(s))) return rc;
    if ((rc = read_ctl(s))) return rc;
    if (mass_err) *mass_err = s->h_ctl[0].err[ig];
    return MGF_OK;
}

extern "C" int mgf_restrict(mgf_solver *s, int ig)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    int rc = single_only(s);
    if (rc) return rc;
    if (ig < 0 || ig + 1 >= s->inst[0].ng) return set_err(s, MGF_ERR_ARG, "bad fine level");
    if ((rc = reset_ctl(s, 0, NULL, NULL))) return rc;
    mgf_launch_restrict(s->full.pairs[s->level_pair[ig]], 1, s->st);
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

extern "C" int mgf_prolong(mgf_solver *s, int igc)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    int rc = single_only(s);
    if (rc) return rc;
    if (igc < 1 || igc >= s->inst[0].ng) return set_err(s, MGF_ERR_ARG, "bad coarse level");
    if ((rc = reset_ctl(s, 0, NULL, NULL))) return rc;
    mgf_launch_prolong(s->full.pairs[s->level_pair[igc - 1]], 0, s->st);
    if ((rc = check_launch(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->st));
    return MGF_OK;
}

extern "C" int mgf_vcycle(mgf_solver *s, int icyc, double *mass_err)
{
    if (!s || icyc < 0) return set_err(s, MGF_ERR_ARG, "bad cycle index");
    int rc = single_only(s);
    if (rc) return rc;
    if ((rc = ensure_hist(s, icyc + 1))) return rc;
    if ((rc = reset_ctl(s, icyc, NULL, NULL))) return rc;
    enqueue_vcycle(s, s->full);
    if ((rc = check_launch(s))) return rc;
    double h = 0.0;
    HIPCHK(s, hipMemcpyAsync(&h, s->d_hist + icyc, sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(s, hipStreamSynchronize(s->st));
    if (mass_err) *mass_err = h;
    return MGF_OK;
}

/* the graph holds no history or control pointers of its own: those live in
 * the control blocks, so it stays valid when the history buffer grows */
static int capture_graph(mgf_solver *s, Program &P)
{
    if (P.graph_ready) return MGF_OK;
    hipGraph_t g;
    HIPCHK(s, hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    s->est = s->st;
    enqueue_vcycle(s, P);
    mgf_launch_end_cycle(s->d_ctl, P.d_members, (int)P.members.size(), s->st);
    hipError_t e = hipStreamEndCapture(s->st, &g);
    if (e != hipSuccess)
        return set_err(s, MGF_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&P.gexec, g, NULL, NULL, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return set_err(s, MGF_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
    P.graph_ready = 1;
    return MGF_OK;
}

/* the lane count for n running instances: opt.lanes (0: MGF_AUTO_LANES),
 * at most MGF_MAX_LANES and n.  Each lane's stream takes a hardware queue
 * while it runs; HIP gives a process 4 by default, and with 4 lanes plus
 * the caller's own stream two lanes end up sharing one (4 lanes measured
 * slower than 1 on the batch benchmark, 3 fastest). */
#define MGF_MAX_LANES 4
#define MGF_AUTO_LANES 3
static int lane_count(const mgf_solver *s, size_t n)
{
    int nl = s->opt.lanes > 0 ? s->opt.lanes : (s->tune_lanes > 0 ? s->tune_lanes : MGF_AUTO_LANES);
    if (nl > MGF_MAX_LANES) nl = MGF_MAX_LANES;
    if ((size_t)nl > n) nl = (int)n;
    return nl < 1 ? 1 : nl;
}

/* streams and events for nl lanes (lane 0 is the solver's stream) */
static int ensure_lanes(mgf_solver *s, int nl)
{
    while ((int)s->lane_st.size() < nl) {
        hipStream_t st;
        HIPCHK(s, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->lane_st.push_back(st);
        s->lane_ev.push_back(nullptr);
    }
    for (int l = 0; l < nl; ++l)
        if (!s->lane_ev[l]) HIPCHK(s, hipEventCreateWithFlags(&s->lane_ev[l], hipEventDisableTiming));
    return MGF_OK;
}

/* the running instances split into nl programs of about equal cell count
 * (largest first, each to the lightest lane), captured when graphs are on */
static int build_lanes(mgf_solver *s, const std::vector<int> &alive, int nl)
{
    for (Program &P : s->lanes) release_program(s, P);
    s->lanes.clear();
    std::vector<int> order(alive);
    auto cells = [&](int b) { return (double)s->inst[b].lv[0].npx * s->inst[b].lv[0].npy; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    std::vector<std::vector<int>> part(nl);
    std::vector<double> load(nl, 0.0);
    for (int b : order) {
        int best = 0;
        for (int l = 1; l < nl; ++l)
            if (load[l] < load[best]) best = l;
        part[best].push_back(b);
        load[best] += cells(b);
    }
    int rc;
    s->lanes.resize(nl);
    for (int l = 0; l < nl; ++l) {
        std::sort(part[l].begin(), part[l].end());
        if ((rc = build_program(s, s->lanes[l], part[l]))) return rc;
        if (s->opt.use_graph && (rc = capture_graph(s, s->lanes[l]))) return rc;
    }
    return MGF_OK;
}

/* Run V-cycles until every instance is done: its recorded residual fell
 * below its tolerance, or it ran its cycle limit (both decided on the
 * device).  The host enqueues cycles in growing chunks and polls the control
 * blocks between chunks; kernels of finished instances return at once, so
 * cycles enqueued past an instance's end change nothing.
 *
 * Lanes: a single problem's V-cycle is a chain of dependent launches, many
 * of them small (coarse levels) or in their tails, but different problems
 * are independent.  A batch's running problems are split into lanes of
 * about equal cell count, each replaying its own V-cycle graph on its own
 * stream, so the GPU overlaps one lane's latency-bound phases with the
 * others' work.
 *
 * Compaction: launches over stopped instances still cost their dispatch
 * (and a group whose members have all stopped, its launch latency), which
 * in a batch whose slowest members run far longer than the rest is a
 * sizeable share of the tail.  When at most 7/8 of the instances of the
 * running schedule are still live, the lanes are rebuilt over the live ones
 * alone and replace the old ones from the next chunk on (7/8: the benchmark
 * batch's step 1% shorter than with 3/4 or 9/10, 2% shorter than with 1/2).
 *
 * Every instance's arithmetic is its own, so neither changes the results. */
static int run_cycles(mgf_solver *s, const int *limit, const double *tol)
{
    int maxnv = 0;
    for (int b = 0; b < s->nb; ++b) maxnv = limit[b] > maxnv ? limit[b] : maxnv;
    int rc = ensure_hist(s, maxnv > 0 ? maxnv : 1);
    if (rc) return rc;
    if ((rc = reset_ctl(s, 0, tol, limit))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->st));
    for (Program &P : s->lanes) release_program(s, P);
    s->lanes.clear();
    std::vector<int> alive;
    for (int b = 0; b < s->nb; ++b)
        if (!s->h_ctl[b].done) alive.push_back(b);
    int nl = lane_count(s, alive.size());
    /* the running programs: the full one while it is the only lane */
    std::vector<Program *> run;
    if (nl <= 1) {
        if (s->opt.use_graph && (rc = capture_graph(s, s->full))) return rc;
        run.push_back(&s->full);
    } else {
        if ((rc = ensure_lanes(s, nl)) || (rc = build_lanes(s, alive, nl))) return rc;
        for (Program &P : s->lanes) run.push_back(&P);
    }
    size_t scheduled = s->full.members.size();
    if (nl > 1) scheduled = alive.size();
    int launched = 0, chunk = 1;
    while (launched < maxnv) {
        const int k = chunk < maxnv - launched ? chunk : maxnv - launched;
        const int nr = (int)run.size();
        /* the other lanes start after what the solver's stream holds */
        if (nr > 1) {
            HIPCHK(s, hipEventRecord(s->lane_ev[0], s->st));
            for (int l = 1; l < nr; ++l)
                HIPCHK(s, hipStreamWaitEvent(s->lane_st[l], s->lane_ev[0], 0));
        }
        for (int t = 0; t < k; ++t)
            for (int l = 0; l < nr; ++l) {
                Program &P = *run[l];
                hipStream_t st = s->lane_st[l];
                if (s->opt.use_graph) {
                    HIPCHK(s, hipGraphLaunch(P.gexec, st));
                } else {
                    s->est = st;
                    enqueue_vcycle(s, P);
                    mgf_launch_end_cycle(s->d_ctl, P.d_members, (int)P.members.size(), st);
                    s->est = s->st;
                }
            }
        for (int l = 1; l < nr; ++l) {
            HIPCHK(s, hipEventRecord(s->lane_ev[l], s->lane_st[l]));
            HIPCHK(s, hipStreamWaitEvent(s->st, s->lane_ev[l], 0));
        }
        if ((rc = check_launch(s))) return rc;
        launched += k;
        if ((rc = read_ctl(s))) return rc;       /* every lane is idle from here */
        alive.clear();
        for (const Program *P : run)
            for (int b : P->members)
                if (!s->h_ctl[b].done) alive.push_back(b);
        if (alive.empty()) break;
        if (s->opt.use_graph && s->opt.compact && 8 * alive.size() <= 7 * scheduled) {
            std::sort(alive.begin(), alive.end());
            nl = lane_count(s, alive.size());
            if ((rc = ensure_lanes(s, nl)) || (rc = build_lanes(s, alive, nl))) return rc;
            run.clear();
            for (Program &P : s->lanes) run.push_back(&P);
            scheduled = alive.size();
        }
        /* chunks of up to 32 cycles between polls: the benchmark batch's step
         * was 0.2-2% longer with caps of 8, 16, 64 or 128 */
        if (chunk < 32) chunk *= 2;
    }
    return read_ctl(s);
}

static int last_value(mgf_solver *s, int b, double *out)
{
    const int nc = s->h_ctl[b].cyc;
    *out = 0.0;
    if (nc > 0)
        HIPCHK(s, hipMemcpy(out, s->d_hist + (size_t)b * s->hist_cap + nc - 1, sizeof(double),
                            hipMemcpyDeviceToHost));
    return MGF_OK;
}

/* 1 when problem b's last solve met its tolerance with finite residuals,
 * 0 when it did not (or none ran), negative on a device error */
int mgf_instance_converged(mgf_solver *s, int b)
{
    if (!s || b < 0 || b >= s->nb) return MGF_ERR_ARG;
    int rc = read_ctl(s);
    if (rc) return rc;
    double last = 0.0;
    if ((rc = last_value(s, b, &last))) return rc;
    const DevCtl &c = s->h_ctl[b];
    return (c.cyc > 0 && last < c.tol && !c.nonfinite) ? 1 : 0;
}

extern "C" int mgf_solve(mgf_solver *s, int nvcycle, double tol, double *history, int *ncycles,
                         int *converged)
{
    if (!s || nvcycle < 0) return set_err(s, MGF_ERR_ARG, "bad arguments");
    int rc = single_only(s);
    if (rc) return rc;
    if ((rc = run_cycles(s, &nvcycle, &tol))) return rc;
    const int nc = s->h_ctl[0].cyc;
    if (history && nc > 0)
        HIPCHK(s, hipMemcpy(history, s->d_hist, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost));
    if (ncycles) *ncycles = nc;
    if (converged) {
        double last;
        if ((rc = last_value(s, 0, &last))) return rc;
        *converged = (nc > 0 && last < tol) ? 1 : 0;
    }
    if (s->h_ctl[0].nonfinite) return set_err(s, MGF_ERR_NONFINITE, "non-finite residual recorded");
    return MGF_OK;
}

extern "C" int mgf_batch_solve(mgf_solver *s, int *ncycles, int *converged)
{
    if (!s) return set_err(s, MGF_ERR_ARG, "null solver");
    std::vector<int> lim(s->nb);
    std::vector<double> tol(s->nb);
    for (int b = 0; b < s->nb; ++b) {
        lim[b] = s->inst[b].nvcycle;
        tol[b] = s->inst[b].tol;
    }
    int rc = run_cycles(s, lim.data(), tol.data());
    if (rc) return rc;
    int nonfinite = 0;
    for (int b = 0; b < s->nb; ++b) {
        if (ncycles) ncycles[b] = s->h_ctl[b].cyc;
        if (converged) {
            double last;
            if ((rc = last_value(s, b, &last))) return rc;
            converged[b] = (s->h_ctl[b].cyc > 0 && last < tol[b]) ? 1 : 0;
        }
        nonfinite |= s->h_ctl[b].nonfinite;
    }
    if (nonfinite) return set_err(s, MGF_ERR_NONFINITE, "non-finite residual recorded");
    return MGF_OK;
}

extern "C" int mgf_batch_history(mgf_solver *s, int b, double *history, int cap)
{
    if (!s || b < 0 || b >= s->nb) return set_err(s, MGF_ERR_ARG, "bad instance");
    int rc = read_ctl(s);
    if (rc) return rc;
    const int nc = s->h_ctl[b].cyc;
    const int n = nc < cap ? nc : cap;
    if (history && n > 0)
        HIPCHK(s, hipMemcpy(history, s->d_hist + (size_t)b * s->hist_cap,
                            (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return nc;
}

/* ---------------------------------------------------------------------- */
/* data access                                                              */
/* ---------------------------------------------------------------------- */

static real_t *field_ptr(DevLevel &L, const char *name)
{
    struct {
        const char *n;
        real_t *p;
    } tab[] = {{"u", L.u},       {"v", L.v},       {"p", L.p},         {"pp", L.pp},
               {"sc", L.sc},     {"resu", L.resu}, {"resv", L.resv},   {"ressc", L.ressc},
               {"apu", L.apu},   {"apv", L.apv}};
    for (auto &t : tab)
        if (strcmp(t.n, name) == 0) return t.p;
    return NULL;
}

/* device array of a field of problem b, level ig, and its length (for the
 * sensitivity calls) */
int mgf_field_array(mgf_solver *s, int b, int ig, const char *name, real_t **ptr, size_t *n,
                    hipStream_t *st)
{
    if (!s || b < 0 || b >= s->nb || ig < 0 || ig >= s->inst[b].ng || !name)
        return set_err(s, MGF_ERR_ARG, "bad arguments");
    DevLevel &L = s->inst[b].lv[ig];
    *ptr = field_ptr(L, name);
    if (!*ptr) return set_err(s, MGF_ERR_ARG, "unknown field %s", name);
    *n = (size_t)L.npx * L.npy;
    *st = s->st;
    return MGF_OK;
}

/* host copies of problem b's level descriptors (for the adjoint) */
const DevLevel *mgf_instance_levels(mgf_solver *s, int b, int *ng)
{
    if (!s || b < 0 || b >= s->nb) return NULL;
    *ng = s->inst[b].ng;
    return s->inst[b].lv;
}

void **mgf_adjoint_slot(mgf_solver *s) { return &s->adj; }

const mgf_problem *mgf_instance_problem(mgf_solver *s, int b)
{
    return (s && b >= 0 && b < s->nb) ? &s->inst[b].plan->prob : NULL;
}

const mgf_options *mgf_solver_options(mgf_solver *s)
{
    return s ? &s->opt : NULL;
}

int mgf_set_error(mgf_solver *s, int code, const char *msg)
{
    return set_err(s, code, "%s", msg);
}

extern "C" int mgf_batch_get_field(mgf_solver *s, int b, int ig, const char *name, double *out)
{
    if (!s || b < 0 || b >= s->nb || ig < 0 || ig >= s->inst[b].ng || !name || !out)
        return set_err(s, MGF_ERR_ARG, "bad arguments");
    DevLevel &L = s->inst[b].lv[ig];
    real_t *p = field_ptr(L, name);
    if (!p) return set_err(s, MGF_ERR_ARG, "unknown field %s", name);
    const size_t n = (size_t)L.npx * L.npy;
    HIPCHK(s, to_host(out, p, n, s->st));
    return MGF_OK;
}

extern "C" int mgf_get_field(mgf_solver *s, int ig, const char *name, double *out)
{
    return mgf_batch_get_field(s, 0, ig, name, out);
}

extern "C" int mgf_set_field(mgf_solver *s, int ig, const char *name, const double *in)
{
    if (!s || ig < 0 || ig >= s->inst[0].ng || !name || !in)
        return set_err(s, MGF_ERR_ARG, "bad arguments");
    DevLevel &L = s->inst[0].lv[ig];
    real_t *p = field_ptr(L, name);
    if (!p) return set_err(s, MGF_ERR_ARG, "unknown field %s", name);
    const size_t n = (size_t)L.npx * L.npy;
    HIPCHK(s, to_device(p, in, n, s->st));
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
    if (!s || reps < 1) return set_err(s, MGF_ERR_ARG, "bad arguments");
    if (ig < 0 || ig >= s->inst[0].ng) return set_err(s, MGF_ERR_ARG, "bad level");
    if (kernel < MGF_KERNEL_SWEEP_PP || kernel > MGF_KERNEL_RELAX_UVS)
        return set_err(s, MGF_ERR_ARG, "unknown kernel id");
    const int mixed = s->opt.precision == MGF_PREC_MIXED;
    const int uvs = kernel == MGF_KERNEL_RELAX_UVS;
    const int fused = kernel == MGF_KERNEL_RELAX_U || kernel == MGF_KERNEL_RELAX_PP || uvs;
    if (fused && (!mixed || s->isolve != 2 || s->opt.split_assembly))
        return set_err(s, MGF_ERR_UNSUPPORTED,
                       "fused relaxation launches exist for mixed-precision red-black only");
    int rc;
    if (!mixed && (rc = single_only(s))) return rc;
    /* the group of every instance's level at instance 0's depth of level ig */
    const int d = s->inst[0].ng - 1 - ig;
    const DevLevel &L0 = s->inst[0].lv[ig];
    int g = -1;
    for (size_t k = 0; k < s->full.groups.size(); ++k) {
        const DevGroup &G = s->full.groups[k];
        if (!G.up && G.depth == d && G.npx == L0.npx && G.npy == L0.npy) g = (int)k;
    }
    if (g < 0) return set_err(s, MGF_ERR_ARG, "no group for level %d", ig);
    const DevGroup &G = s->full.groups[g];
    const int field = (kernel == MGF_KERNEL_SWEEP_PP || kernel == MGF_KERNEL_RELAX_PP) ? 2 : 0;
    const int isolve = s->isolve == 1 ? 1 : 2;
    const int fp64_smoother = s->smoother == 3 ? 3 : isolve;
    const int even = reps + (reps & 1);    /* an even count ends without a copy-back */
    real_t *save = NULL, *phi = NULL;
    const size_t n = (size_t)L0.npx * L0.npy;
    if (!mixed) {
        /* FP64 sweeps work in place: keep the field */
        phi = field == 2 ? L0.pp : L0.u;
        HIPCHK(s, hipMalloc(&save, n * sizeof(real_t)));
        HIPCHK(s, hipMemcpyAsync(save, phi, n * sizeof(real_t), hipMemcpyDeviceToDevice, s->st));
    }
    /* the sweeps check the stop flags: clear them for the measurement */
    for (int b = 0; b < s->nb; ++b) s->h_ctl[b].done = 0;
    (void)hipMemcpyAsync(s->d_ctl, s->h_ctl, (size_t)s->nb * sizeof(DevCtl), hipMemcpyHostToDevice,
                         s->st);
    /* mixed red-black: launches as a relaxation issues them, each of
     * spl = min(sweeps per launch, sweeps of that field) sweeps */
    const int nsw_field = field == 2 ? s->nswpp : s->nswpm;
    SweepCfg cfg = mgf_sweep_cfg(s->opt.sweeps_per_launch, s->opt.sweep_threads, 1);
    if (nsw_field >= 1 && cfg.spl > nsw_field) cfg.spl = nsw_field;
    const int spl = (mixed && isolve == 2) ? cfg.spl : 1;
    const SweepCfg fcfg = mgf_sweep_cfg(s->opt.sweeps_per_launch, s->opt.sweep_threads, 0);
    double fused_bpu = 0.0;
    int uvs_ok = 1;
    auto launch = [&](int nl) {
        if (uvs) {
            for (int k = 0; k < nl; ++k) uvs_ok &= mgf_launch_uvs_mixed(G, isolve, s->nswpm, fcfg, s->st);
        } else if (fused) {
            int nf = 0;
            for (int k = 0; k < nl; ++k)
                fused_bpu = mgf_launch_relax_first_mixed(G, field, nsw_field, fcfg, &nf, s->st);
        } else if (mixed) {
            mgf_launch_sweeps_mixed(G, field, isolve, nl * spl, cfg, s->st);
        } else {
            mgf_launch_smooth(G, field, fp64_smoother, nl, s->st);
        }
    };
    launch(2);                                            /* warm up */
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s->st);
    launch(even);
    (void)hipEventRecord(e1, s->st);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (save) {
        (void)hipMemcpyAsync(phi, save, n * sizeof(real_t), hipMemcpyDeviceToDevice, s->st);
        (void)hipStreamSynchronize(s->st);
        (void)hipFree(save);
    }
    if ((rc = check_launch(s))) return rc;
    if (!uvs_ok)
        return set_err(s, MGF_ERR_UNSUPPORTED,
                       "no merged u, v, sc launch for this level or configuration");
    if (uvs) {
        const double ncell = (double)(G.npx - 2) * (double)(G.npy - 2) * (double)G.count;
        if (mean_us) *mean_us = 1000.0 * (double)ms / (double)even;
        /* u, v, p, sc, resu, resv, ressc and the flag byte read; u + du,
         * v + dv (FP64), the sc correction (FP32), apu and apv written */
        if (bytes_per_launch) *bytes_per_launch = (7 * 8.0 + 1.0 + 2 * 8.0 + 4.0 + 2 * 8.0) * ncell;
        return MGF_OK;
    }
    int nx = field == 2 ? G.npx - 1 : G.npx - 2, ny = G.npy - 1;
    const double nunk = (double)(nx - 1) * (double)(ny - 1) * (double)G.count;
    /* algorithmic words per unknown and sweep: aw ae as an ap su and phi read
     * + write (FP64), or aw ae as an 1/ap r and d read + write (FP32); a
     * launch counts its sweeps */
    if (mean_us) *mean_us = 1000.0 * (double)ms / (double)even;
    if (bytes_per_launch)
        *bytes_per_launch = fused ? fused_bpu * nunk : (double)spl * 8.0 * (mixed ? 4.0 : 8.0) * nunk;
    return MGF_OK;
}

/* ---------------------------------------------------------------------- */
/* autotuner                                                                */
/* ---------------------------------------------------------------------- */

/* the shapes of the tuning table on an existing program's groups (its
 * captured graph, made with the old shapes, is dropped) */
static void apply_tuning(mgf_solver *s, Program &P)
{
    for (DevGroup &G : P.groups) {
        G.uvs_shape = G.pp_shape = G.nt_asm = 0;
        for (const TuneEntry &t : s->tune)
            if (t.npx == G.npx && t.npy == G.npy) {
                G.uvs_shape = t.uvs;
                G.pp_shape = t.pp;
                G.nt_asm = t.nt;
            }
    }
    if (P.graph_ready) {
        (void)hipGraphExecDestroy(P.gexec);
        P.graph_ready = 0;
    }
}

extern "C" int mgf_autotune(mgf_solver *s, int reps, char *table, size_t cap)
{
    if (!s || reps < 1) return set_err(s, MGF_ERR_ARG, "bad arguments");
#ifdef MGF_TANGENT
    (void)table;
    (void)cap;
    return set_err(s, MGF_ERR_UNSUPPORTED, "the sensitivity library has no alternative shapes");
#else
    if (s->opt.precision != MGF_PREC_MIXED || s->isolve != 2 || s->opt.split_assembly)
        return set_err(s, MGF_ERR_UNSUPPORTED,
                       "launch shapes are tuned for mixed-precision red-black with fused assembly");
    int rc;
    /* a representative state: three V-cycles from the initial one */
    if ((rc = mgf_reset(s))) return rc;
    {
        std::vector<int> lim(s->nb, 3);
        std::vector<double> tol(s->nb, 0.0);
        if ((rc = run_cycles(s, lim.data(), tol.data()))) return rc;
    }
    for (int b = 0; b < s->nb; ++b) s->h_ctl[b].done = 0;
    HIPCHK(s, hipMemcpyAsync(s->d_ctl, s->h_ctl, (size_t)s->nb * sizeof(DevCtl),
                             hipMemcpyHostToDevice, s->st));
    const SweepCfg fcfg = mgf_sweep_cfg(s->opt.sweeps_per_launch, s->opt.sweep_threads, 0);
    hipEvent_t e0, e1;
    HIPCHK(s, hipEventCreate(&e0));
    HIPCHK(s, hipEventCreate(&e1));
    auto timed = [&](auto &&launch) {
        launch();
        launch();
        (void)hipEventRecord(e0, s->st);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(e1, s->st);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return (double)ms / reps;
    };
    std::vector<TuneEntry> tab;
    for (const DevGroup &G : s->full.groups) {
        if (G.up) continue;
        bool seen = false;
        for (const TuneEntry &t : tab) seen |= t.npx == G.npx && t.npy == G.npy;
        if (seen) continue;
        DevGroup Gc = G;
        Gc.uvs_shape = 1;
        Gc.nt_asm = 0;
        if (!mgf_launch_uvs_mixed(Gc, s->isolve, s->nswpm, fcfg, s->st)) continue;
        /* the u, v, sc launch: 16-row / 512, 32-row / 256, 32-row / 512 */
        const int cu[3][2] = {{1, 512}, {2, 256}, {2, 512}};
        int bu = 0;
        double tu = 1e30;
        for (int k = 0; k < 3; ++k) {
            Gc.uvs_shape = cu[k][0];
            Gc.nt_asm = cu[k][1];
            const double t = timed([&] { (void)mgf_launch_uvs_mixed(Gc, s->isolve, s->nswpm, fcfg, s->st); });
            if (t < tu) {
                tu = t;
                bu = k;
            }
        }
        Gc.uvs_shape = cu[bu][0];
        Gc.nt_asm = cu[bu][1];
        int bp = 1;
        double tp = 1e30;
        for (int k = 1; k <= 2 && s->nswpp >= 1 && s->nswpp <= MGF_MAX_SWEEPS_PER_LAUNCH; ++k) {
            Gc.pp_shape = k;
            const double t = timed([&] { (void)mgf_launch_pp_merged_mixed(Gc, s->nswpp, s->st); });
            if (t < tp) {
                tp = t;
                bp = k;
            }
        }
        tab.push_back(TuneEntry{G.npx, G.npy, cu[bu][0], bp, cu[bu][1]});
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if ((rc = check_launch(s))) return rc;
    s->tune = tab;
    apply_tuning(s, s->full);
    /* lanes: the whole solve of the batch (each problem to its own stopping
     * test) with each count; the lanes matter most in the tail, where only
     * the slowest problems still run, so nothing shorter measures them (the
     * caller's explicit choice stands) */
    int best_lanes = 0;
    if (s->opt.lanes == 0 && s->nb > 1) {
        double best = 1e30;
        std::vector<int> lim(s->nb);
        std::vector<double> tol(s->nb);
        for (int b = 0; b < s->nb; ++b) {
            lim[b] = s->inst[b].nvcycle;
            tol[b] = s->inst[b].tol;
        }
        for (int nl = 1; nl <= MGF_MAX_LANES && nl <= s->nb; ++nl) {
            s->tune_lanes = nl;
            double t = 1e30;
            for (int rep = 0; rep < 1; ++rep) {
                if ((rc = mgf_reset(s))) return rc;
                HIPCHK(s, hipStreamSynchronize(s->st));
                const auto t0 = std::chrono::steady_clock::now();
                if ((rc = run_cycles(s, lim.data(), tol.data()))) return rc;
                const double dt =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                t = dt < t ? dt : t;
            }
            if (t < best) {
                best = t;
                best_lanes = nl;
            }
        }
    }
    s->tune_lanes = best_lanes;
    if ((rc = mgf_reset(s))) return rc;
    HIPCHK(s, hipStreamSynchronize(s->st));
    std::string text = "# mgf_autotune: merged mixed relaxation shapes per level shape\n";
    char line[128];
    if (best_lanes) {
        snprintf(line, sizeof line, "lanes %d\n", best_lanes);
        text += line;
    }
    for (const TuneEntry &t : tab) {
        snprintf(line, sizeof line, "level %d %d %d %d %d\n", t.npx, t.npy, t.uvs, t.pp, t.nt);
        text += line;
    }
    if (table && cap > 0) {
        const size_t n = text.size() < cap - 1 ? text.size() : cap - 1;
        memcpy(table, text.data(), n);
        table[n] = 0;
    }
    return (int)text.size();
#endif
}
