/*
 * mgf_dev.h -- device-side data of one solve group (internal, HIP C++).
 *
 * Layout in HBM: per instance one allocation holding, per level, the
 * reference's row-major npx*npy arrays (ij = j*npx + i) in FP64, 256-byte
 * aligned.  The face areas are uniform per level (checked at set-up), so they
 * and every metric derived from them are scalars here; the 1e30 penalty
 * arrays and the obstacle tag are one flag byte per point.
 *
 * Every kernel works on a group of levels with the same shape -- one level
 * of one instance, or the same-depth level of many instances of a batch --
 * and reads the descriptor of its level from a device array indexed by
 * blockIdx.z.
 */
#ifndef MGF_DEV_H
#define MGF_DEV_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "mgf_dual.h"

#define MGF_MAX_LEVELS 24

/* device-resident control block of one instance's solve */
struct DevCtl {
    int cyc;                 /* V-cycles finished */
    int done;                /* stopping test fired: later kernels return at once */
    int nonfinite;           /* a recorded residual was NaN or infinite */
    int maxcyc;              /* cycle limit of this solve */
    double tol;
    double *hist;            /* residual history of this instance */
    double err[MGF_MAX_LEVELS];   /* latest sum |div| per level, from the relaxations
                                     that reduce it (the step calls always; a V-cycle
                                     only on the recorded level-0 relaxation) */
};

struct DevPhys {
    double relxm, relxm1, relxp, sc_obs;
    double irelxm;        /* 1 / relxm, for the mixed path's assembly */
    int ibc[4];
};

/* The level descriptor.  Its body is written once and instantiated twice:
 * DevLevel with generic pointers (host side, and what the launches pass),
 * and DevLevelG, the same layout with every array pointer in the global
 * address space, which the kernels read the descriptor through.  Pointers
 * loaded from memory are otherwise generic, and the compiler then emits
 * flat loads and stores, which also count against the LDS wait counter: a
 * kernel that mixes them with LDS traffic waits for its global loads at
 * every LDS wait. */
#ifdef MGF_TANGENT
/* dual numbers are structs, and clang does not copy a struct out of
 * global-address-space storage into a generic object: generic pointers in
 * the sensitivity build */
#define MGF_G
#else
#define MGF_G __attribute__((address_space(1)))
#endif
#define MGF_LEVEL_BODY(GQ)                                                              \
    int npx, npy;                                                                     \
    int ig;               /* level index within its instance (0 = finest) */          \
    int npcopy;                                                                       \
    int has_obs;                                                                      \
    int ext;              /* level below the obstacle level: coarsened obstacles */   \
    int nblk;             /* workgroups of the residual kernel (partials) */          \
    int isolve;           /* 1 Jacobi, 2 red-black */                                 \
    int pset;             /* mixed: u, v, sc are swept (nswpm >= 1), so penalised     \
                             points are set, not corrected (pending_value) */         \
    /* metrics, each computed on the host with the reference's expression */         \
    real_t ax, ay;        /* areax (= dy), areay (= dx) */                            \
    real_t du_x, du_y;    /* amu*(areax/dx), amu*(areay/dy): coefu, coefv, walls */   \
    real_t ds_x, ds_y;    /* amusc*(areax/dx), amusc*(areay/dy): coefsc, its adds */  \
    real_t axax, ayay;    /* areax*areax, areay*areay: coefp */                       \
    real_t bodyx, bodyy;  /* (dpdxm*areax)*areay, (dpdym*areax)*areay */              \
    DevPhys P;            /* the instance's physics / boundary types */               \
    GQ DevCtl *ctl;       /* the instance's control block */                          \
    /* state */                                                                       \
    GQ real_t *u, *v, *p, *pp, *sc;                                                   \
    GQ real_t *resu, *resv, *ressc;  /* FAS tau terms (zero on level 0) */            \
    GQ real_t *apu, *apv;            /* momentum diagonals (persist; ghosts 1e30) */  \
    /* scratch shared by the equations, as in the reference */                        \
    GQ real_t *cx, *cy;                                                               \
    GQ real_t *aw, *ae, *as, *an, *ap, *su;                                           \
    GQ real_t *tmp;                  /* smoother ping-pong */                         \
    GQ real_t *cu, *cv, *cp, *csc;   /* coarse-grid corrections (ghosts stay 0) */    \
    GQ double *partial;              /* per-block partial sums of |div| */            \
    /* mixed precision (null in FP64 mode): FP32 coefficients, 1/diagonal,           \
     * right-hand side (the FP64 residual for u, v, sc; -div for p'), the two        \
     * buffers of the FP32 correction sweeps, and the corrections of u and v,        \
     * which stay pending (u + du, v + dv) until the update writes u and v, and      \
     * (when u, v and sc relax in one launch) that of sc, applied there too */        \
    GQ float *aw32, *ae32, *as32, *an32, *iap32, *r32, *da32, *db32, *du32, *dv32;    \
    GQ float *ds32;                                                                   \
    /* the merged relaxation's u and v with their corrections applied (FP64, over    \
     * the swept ranges), read by its p' assembly and update instead of u + du;      \
     * FP64: the relaxed u, v (and sc in scn) kept there until the update, and      \
     * in pnew the updated p of the copy origins, written by the last p' launch */    \
    GQ real_t *uc, *vc, *scn, *pnew;                                                  \
    /* obstacle data */                                                               \
    GQ const uint8_t *flags;                                                          \
    GQ const uint64_t *sccode;                                                        \
    GQ const int *pslot;                                                              \
    GQ const int *porig;                                                              \
    GQ real_t *pstash;

struct DevLevel {
    MGF_LEVEL_BODY()
};
struct DevLevelG {
    MGF_LEVEL_BODY(MGF_G)
};
static_assert(sizeof(DevLevel) == sizeof(DevLevelG), "descriptor views differ in layout");
static_assert(offsetof(DevLevel, pstash) == offsetof(DevLevelG, pstash),
              "descriptor views differ in layout");
/* the FP64 path's scalar in the global address space */
typedef MGF_G real_t gdbl;
typedef MGF_G float gflt;

/* a set of same-shape levels launched together */
struct DevGroup {
    const DevLevel *d;       /* device array of descriptors */
    int count;
    int npx, npy;
    int max_npcopy;
    int any_pbound_x, any_pbound_y, any_outflow;
    int any_finest;          /* some member is its instance's level 0 */
    int depth;               /* levels above the coarsest */
    int up;                  /* 1: the up-leg subset (no finest levels) */
    /* launch shapes of the merged mixed relaxation (result-neutral), from the
     * solver's tuning table: 0 the built-in size rule, 1 the small-level
     * shape (16-row tiles), 2 the large-level shape (32-row tiles); nt_asm:
     * threads of the large u, v, sc launch (0: the options') */
    int uvs_shape, pp_shape, nt_asm;
};

/* a restriction / prolongation between two same-length groups */
struct DevPair {
    const DevLevel *f, *c;   /* device arrays, element k of f is the finer level of c[k] */
    int count;
    int npxf, npyf, npxc, npyc;
};

/* solver internals shared with the sensitivity calls (mgf_solver.hip) */
struct mgf_seed;
struct mgf_problem;
struct mgf_options;
struct mgf_solver;
int mgf_batch_create_seeded(const mgf_problem *const *probs, int n, const mgf_options *opt,
                            mgf_solver **out, const mgf_seed *seeds);
int mgf_field_array(mgf_solver *s, int b, int ig, const char *name, real_t **ptr, size_t *n,
                    hipStream_t *st);
const mgf_problem *mgf_instance_problem(mgf_solver *s, int b);
const mgf_options *mgf_solver_options(mgf_solver *s);
int mgf_set_error(mgf_solver *s, int code, const char *msg);
const DevLevel *mgf_instance_levels(mgf_solver *s, int b, int *ng);
void **mgf_adjoint_slot(mgf_solver *s);
void mgf_adjoint_release(void *work);
int mgf_instance_converged(mgf_solver *s, int b);

/* launch wrappers (mgf_kernels.hip) */
void mgf_launch_fluxes(const DevGroup &G, hipStream_t st);
void mgf_launch_asm_u(const DevGroup &G, hipStream_t st);
void mgf_launch_asm_v(const DevGroup &G, hipStream_t st);
/* err: reduce the sum |div| (1), and record it as the history entry of
 * level-0 members (2), or neither (0) */
/* pend: u and v relaxed with pend (their results in uc, vc) */
void mgf_launch_asm_p(const DevGroup &G, int err, hipStream_t st, int pend = 0);
void mgf_launch_asm_sc(const DevGroup &G, hipStream_t st);
/* field: 0 u, 1 v, 2 pp, 3 sc; isolve 1 Jacobi, 2 red-black, 3 lexicographic
 * Gauss-Seidel (the labelled variant; rows <= LEX_MAX_ROWS) */
#define LEX_MAX_ROWS 1024
void mgf_launch_smooth(const DevGroup &G, int field, int isolve, int nsw, hipStream_t st);
/* FP64 red-black of u (0), v (1) or sc (3) with nsw <= 3: assembly and
 * sweeps in one launch (replaces mgf_launch_asm_* + mgf_launch_smooth);
 * returns 0, launching nothing, when not applicable.  pend (u, v): the
 * result stays in uc / vc until the update (mgf_launch_update 3), which
 * needs both u and v relaxed so: mgf_relax64_uv_pending says when */
int mgf_relax64_uv_pending(const DevGroup &G, int isolve, int nsw);
/* pend 2's u, v and sc relaxations as one launch (0 when not applicable) */
int mgf_launch_relax64_uvs(const DevGroup &G, int isolve, int nsw, hipStream_t st);
/* pend 2 (no outflow boundary, so nothing writes sc's ghosts between the
 * update and the sc solve): the fluxes are formed from u, v inside each
 * assembly (no k_fluxes), sc is relaxed before p' with its result in scn,
 * and the update (4) applies u, v and sc */
int mgf_launch_relax64(const DevGroup &G, int field, int isolve, int nsw, hipStream_t st,
                       int pend = 0);
/* the same for p' (any nsw; err as mgf_launch_asm_p; replaces
 * mgf_launch_asm_p + mgf_launch_smooth; pend as mgf_launch_asm_p) */
int mgf_launch_relax64_p(const DevGroup &G, int isolve, int nsw, int err, int pend,
                         hipStream_t st, int *stashed = nullptr);
/* pending: mixed precision, u and v corrections still in du32 / dv32 (1);
 * the merged relaxation's u + du, v + dv in uc, vc, the sc correction in
 * ds32 and the updated p of the copy origins in tmp (2); FP64, the relaxed
 * u and v in uc, vc (3), and sc in scn (4), and the copy origins' updated p
 * in pnew (5: mgf_launch_relax64_p's *stashed) */
void mgf_launch_update(const DevGroup &G, int pending, hipStream_t st);
void mgf_launch_pbound(const DevGroup &G, hipStream_t st);
void mgf_launch_outflow(const DevGroup &G, hipStream_t st);
/* restv (unless an update_restv launch did it) and the tau terms */
void mgf_launch_restrict(const DevPair &R, int restv, hipStream_t st);
/* the update of a down-leg group fused with restv onto its coarser levels
 * (R: the transfer pair of the same levels, members in the same order) */
void mgf_launch_update_restv(const DevGroup &G, const DevPair &R, int pending, hipStream_t st);
/* stored: c = x_c - R x_f with R x_f as the restriction stored it (a
 * V-cycle, whose fine level has not changed since); 0: R x_f recomputed from
 * the fine level (the step-level call) */
void mgf_launch_prolong(const DevPair &R, int stored, hipStream_t st);
/* the stopping test for the n instances listed in members (device array) */
void mgf_launch_end_cycle(DevCtl *ctls, const int *members, int n, hipStream_t st);
/* mixed precision: FP64 assembly + residual, FP32 correction sweeps, FP64 apply */
void mgf_launch_asm_mixed(const DevGroup &G, int field, int err, hipStream_t st);
/* shape of the temporally blocked red-black sweeps: sweeps per launch,
 * threads per workgroup of the launches on the stored equation (nt) and of
 * the assembling launches (nt_asm) -- each workgroup takes a tile of 64
 * columns -- and whether the assembly is fused into the first launch */
struct SweepCfg {
    int spl, nt, nt_asm, fuse;
};
#ifndef MGF_SWEEP_TILE_ROWS
#define MGF_SWEEP_TILE_ROWS 16
#endif
#define MGF_AUTO_SWEEPS_PER_LAUNCH 3
#define MGF_MAX_SWEEPS_PER_LAUNCH 6
/* sweeps of a launch that also assembles the equation */
#define MGF_MAX_FUSED_SWEEPS 3
/* threads of the launches on the stored equation: 1024 measured 7% faster
 * than 512 for three p' sweeps over 84 members of 512 x 512 (256: 47%
 * slower) */
#define MGF_AUTO_SWEEP_THREADS 1024
#ifndef MGF_AUTO_ASM_THREADS
#define MGF_AUTO_ASM_THREADS 512
#endif
SweepCfg mgf_sweep_cfg(int sweeps_per_launch, int threads, int split_assembly);
/* the first (assembling) launch of a fused relaxation alone, for timing;
 * returns compulsory bytes per unknown */
double mgf_launch_relax_first_mixed(const DevGroup &G, int field, int nsw, SweepCfg cfg,
                                    int *n_out, hipStream_t st);
/* assembly + sweeps + FP64 update of one field in a relaxation */
void mgf_launch_relax_field_mixed(const DevGroup &G, int field, int isolve, int nsw,
                                  SweepCfg cfg, int err, hipStream_t st);
/* A mixed relaxation up to its update with u, v and sc in one launch: that
 * launch, p' (whose last launch also stashes the updated p of the copy
 * origins), the sum |div| as `err` says, and the update, which applies sc's
 * pending correction too.  0 (nothing launched) when the group or the
 * configuration does not allow it: red-black with fused assembly, nswm
 * within one launch, nswp >= 1, and no outflow boundary (which rewrites
 * sc's ghost between the update and the sc solve).  R (or null): the
 * restriction that follows this down-leg relaxation; without periodic
 * boundaries the update then does its restv too, and 2 is returned. */
int mgf_launch_relax_merged_mixed(const DevGroup &G, int isolve, int nswm, int nswp,
                                  SweepCfg cfg, int err, const DevPair *R, hipStream_t st);
/* its first launch alone (u, v and sc; for timing: it rewrites what it
 * wrote before, given unchanged inputs); 0 when not applicable */
int mgf_launch_uvs_mixed(const DevGroup &G, int isolve, int nswm, SweepCfg cfg, hipStream_t st);
/* its one-launch p' (for timing, as above); 0 when not applicable */
int mgf_launch_pp_merged_mixed(const DevGroup &G, int nswp, hipStream_t st);
void mgf_launch_smooth_mixed(const DevGroup &G, int field, int isolve, int nsw, SweepCfg cfg,
                             hipStream_t st);
/* the FP32 sweeps alone (they only touch the correction buffers); returns
 * the number of launches */
int mgf_launch_sweeps_mixed(const DevGroup &G, int field, int isolve, int nsw, SweepCfg cfg,
                            hipStream_t st);
void mgf_launch_fill(real_t *a, double v, size_t n, hipStream_t st);
int mgf_partial_blocks(int npx, int npy);

#endif
