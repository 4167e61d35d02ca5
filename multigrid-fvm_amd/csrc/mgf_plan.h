/*
 * mgf_plan.h -- host-side set-up of the level hierarchy (internal).
 *
 * The plan is everything the device needs that depends only on the input:
 * level sizes and uniform metrics, the initial/boundary state of every level
 * after init + rest_geom, and the obstacle effects re-expressed as per-cell
 * data a parallel kernel can apply without the reference's sequential
 * loops:
 *   flags    one byte per point: which penalties are active there
 *            (spu/spv/spsc = 1e30 in V6:629-633), obstacle cell (tag = 0),
 *            and whether the point has scalar-coefficient adds or an
 *            obstacle-edge pressure copy;
 *   sccode   per point, the ordered sequence of obstacle diffusion adds
 *            coefsc applies on the host (V6:939-963);
 *   pcopy    the obstacle-edge pressure copies of update (V6:1153-1177)
 *            resolved to one origin per destination.
 */
#ifndef MGF_PLAN_H
#define MGF_PLAN_H

#include <stddef.h>
#include <stdint.h>

#include "mgfvm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* per-point flag bits */
#define MGF_F_SPU    0x01u   /* spu = 1e30 on this u face */
#define MGF_F_SPV    0x02u   /* spv = 1e30 on this v face */
#define MGF_F_SPSC   0x04u   /* spsc = 1e30 in this cell */
#define MGF_F_OBS    0x08u   /* obstacle cell: tag = 0 */
#define MGF_F_SCADD  0x10u   /* sccode[ij] holds obstacle diffusion adds */
#define MGF_F_PCOPY  0x20u   /* p[ij] is overwritten by an obstacle-edge copy */
#define MGF_F_PORIG  0x40u   /* some obstacle-edge copy ends with this cell's updated p */

/* sccode layout: counts of x-type adds to ae (bits 0-4) and aw (5-9), of
 * y-type adds to an (10-14) and as (15-19), number of adds to ap (20-25),
 * then the ap add sequence, one bit per add in application order
 * (0 = x-type value, 1 = y-type value), from bit 26 upward. */
#define MGF_SC_MAXCNT   31
#define MGF_SC_MAXAP    38

typedef struct mgf_plan_level {
    int npx, npy;
    double dx, dy;
    double ax, ay;             /* uniform face areas: areax = dy, areay = dx */
    int has_obs;               /* obstacles act on this level */
    int ext;                   /* ig >= ig_obs: obstacles coarsened from the finer level */
    /* initial state after init + rest_geom, npx*npy each */
    double *u, *v, *p, *sc;
    uint8_t *flags;
    uint64_t *sccode;          /* zero where MGF_F_SCADD is clear */
    int *pslot;                /* -1, or the copy slot of this destination */
    int npcopy;                /* copy slots */
    int *porig;                /* npcopy origins (cell index) */
} mgf_plan_level;

typedef struct mgf_plan {
    mgf_problem prob;          /* deep copy; arrays owned by the plan */
    int ngrid;
    mgf_plan_level *lv;
} mgf_plan;

int mgf_plan_build(const mgf_problem *p, int strict_limits, mgf_plan **out,
                   char *err, size_t errlen);
void mgf_plan_free(mgf_plan *pl);
/* deep copy of a problem (arrays duplicated) */
int mgf_problem_copy(const mgf_problem *src, mgf_problem *dst);
void mgf_problem_release(mgf_problem *p);
void mgf_set_error_string(const char *msg);

#ifdef __cplusplus
}
#endif
#endif
