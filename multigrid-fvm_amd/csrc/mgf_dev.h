/*
 * mgf_dev.h -- device-side data of one solver (internal, HIP C++).
 *
 * Layout in HBM: one allocation per solver holding, per level, the
 * reference's row-major npx*npy arrays (ij = j*npx + i) in FP64, 256-byte
 * aligned.  The face areas are uniform per level (checked at set-up), so they
 * and every metric derived from them are scalars here; the 1e30 penalty
 * arrays and the obstacle tag are one flag byte per point.
 */
#ifndef MGF_DEV_H
#define MGF_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MGF_MAX_LEVELS 24

struct DevLevel {
    int npx, npy;
    int npcopy;
    int has_obs;
    int ext;              /* level below the obstacle level: coarsened obstacles */
    /* metrics, each computed on the host with the reference's expression */
    double ax, ay;        /* areax (= dy), areay (= dx) */
    double du_x, du_y;    /* amu*(areax/dx), amu*(areay/dy): coefu, coefv, wall adds */
    double ds_x, ds_y;    /* amusc*(areax/dx), amusc*(areay/dy): coefsc and its adds */
    double axax, ayay;    /* areax*areax, areay*areay: coefp */
    double bodyx, bodyy;  /* (dpdxm*areax)*areay, (dpdym*areax)*areay */
    /* state */
    double *u, *v, *p, *pp, *sc;
    double *resu, *resv, *ressc;     /* FAS tau terms (zero on level 0) */
    double *apu, *apv;               /* momentum diagonals (persist; ghosts 1e30) */
    /* scratch shared by the equations, as in the reference */
    double *cx, *cy;
    double *aw, *ae, *as, *an, *ap, *su;
    double *tmp;                     /* Jacobi ping-pong */
    double *cu, *cv, *cp, *csc;      /* coarse-grid corrections (ghosts stay 0) */
    double *partial;                 /* per-block partial sums of |div| */
    /* obstacle data */
    const uint8_t *flags;
    const uint64_t *sccode;
    const int *pslot;
    const int *porig;
    double *pstash;
};

struct DevPhys {
    double relxm, relxm1, relxp, sc_obs;
    int ibc[4];
};

/* device-resident control block of a solve */
struct DevCtl {
    int cyc;                 /* V-cycles finished */
    int done;                /* stopping test fired: later kernels return at once */
    int nonfinite;           /* a recorded residual was NaN or infinite */
    int pad;
    double tol;
    unsigned int counter[MGF_MAX_LEVELS];   /* last-block reduction tickets */
    double err[MGF_MAX_LEVELS];             /* latest sum |div| per level */
};

/* launch wrappers (mgf_kernels.hip) */
void mgf_launch_fluxes(const DevLevel &L, const DevCtl *ctl, hipStream_t st);
void mgf_launch_asm_u(const DevLevel &L, const DevPhys &P, const DevCtl *ctl, hipStream_t st);
void mgf_launch_asm_v(const DevLevel &L, const DevPhys &P, const DevCtl *ctl, hipStream_t st);
void mgf_launch_asm_p(const DevLevel &L, const DevPhys &P, DevCtl *ctl, int lev, int record,
                      double *hist, hipStream_t st);
void mgf_launch_asm_sc(const DevLevel &L, const DevPhys &P, const DevCtl *ctl, hipStream_t st);
/* field: 0 u, 1 v, 2 pp, 3 sc */
void mgf_launch_smooth(const DevLevel &L, int field, int isolve, int nsw, const DevCtl *ctl,
                       hipStream_t st);
void mgf_launch_update(const DevLevel &L, const DevPhys &P, const DevCtl *ctl, hipStream_t st);
void mgf_launch_pbound(const DevLevel &L, const DevPhys &P, const DevCtl *ctl, hipStream_t st);
void mgf_launch_outflow(const DevLevel &L, const DevPhys &P, const DevCtl *ctl, hipStream_t st);
void mgf_launch_restrict(const DevLevel &F, const DevLevel &C, const DevPhys &P,
                         const DevCtl *ctl, hipStream_t st);
void mgf_launch_prolong(const DevLevel &F, const DevLevel &C, const DevCtl *ctl, hipStream_t st);
void mgf_launch_end_cycle(DevCtl *ctl, const double *hist, hipStream_t st);
void mgf_launch_fill(double *a, double v, size_t n, hipStream_t st);
int mgf_partial_blocks(int npx, int npy);

#endif
