/*
 * mgfvm_tangent.h -- forward sensitivities of the converged steady solution
 * (libmgfvm_tangent.so).
 *
 * libmgfvm_tangent.so is the FP64 solver of libmgfvm.so built with a dual
 * number (value, derivative) as its scalar: every kernel of the V-cycle
 * carries, alongside each value, its derivative along one seeded direction
 * of the inputs.  Run to convergence, the derivative fields converge to the
 * derivative of the discrete steady solution (the iteration's fixed point)
 * and the value fields are bit-identical to libmgfvm.so's.  It exports the
 * whole mgfvm.h ABI (FP64 only; MGF_PREC_MIXED is rejected) plus the calls
 * below.  A batch of n problems with n seeds gives n directional
 * derivatives in one solve.
 *
 * The reference has no derivative capability; these calls serve the
 * gradient workloads of the benchmark (dJ/d(inlet profile, amuf, prsc)).
 * This is forward mode: the cost grows with the number of directions.
 */
#ifndef MGFVM_TANGENT_H
#define MGFVM_TANGENT_H

#include "mgfvm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One input direction.  amuf enters every diffusion coefficient (momentum
 * amuf * area / spacing, scalar amuf / prsc * area / spacing, and the wall
 * adds); the inlet profile is the level-0 u on the west boundary face, rows
 * 1..nfy, carried to the coarse levels as rest_geom carries the values
 * (V6:1391-1473); u_xp is the east boundary u (rows 1..nfy of every level,
 * V6:593).  With a fixed east u the discrete problem has a steady solution
 * only while inflow and outflow balance, so a direction that changes the
 * inlet profile must change u_xp by the mean of that change. */
typedef struct mgf_seed {
    double d_amuf;
    double d_prsc;
    const double *d_inlet_u;   /* nfy values, or NULL for none */
    double d_u_xp;
} mgf_seed;

/* mgf_batch_create with one seed per problem (seeds[b] seeds problem b). */
int mgf_tangent_batch_create(const mgf_problem *const *probs, const mgf_seed *seeds, int n,
                             const mgf_options *opt, mgf_solver **out);
/* value and derivative of a field of problem b, level ig (either may be
 * NULL); fields as for mgf_batch_get_field */
int mgf_tangent_get_field(mgf_solver *s, int b, int ig, const char *name, double *value,
                          double *derivative);
/* The two objectives of problem b and their derivatives along its seed:
 *   J[0] = mean p over the first interior column - mean p over the last
 *          (pressure drop, cells i = 1 and i = nfx),
 *   J[1] = sum over rows of areax * u * sc at the east boundary face, with
 *          sc of the last interior cell (outlet scalar flux, upwind). */
int mgf_tangent_objectives(mgf_solver *s, int b, double J[2], double dJ[2]);

/* The whole gradient of J with respect to the inlet profile (rows 1..nfy),
 * amuf and prsc: the nfy + 2 directions of problem p (its inlet profile as
 * given in p->inlet_u), solved as batches of at most max_batch problems
 * (0: all in one batch).  An inlet-row direction also raises the east
 * boundary u by 1/nfy (see mgf_seed), so it is the derivative along a
 * mass-consistent change.  Outputs: J[2]; g_inlet[2 * nfy] with
 * dJ[k]/d(inlet row r) at k * nfy + (r - 1); g_amuf[2]; g_prsc[2];
 * *converged = 1 when every direction's solve met its tolerance (any output
 * may be NULL).  Forward mode: nfy + 2 solves batched on the device; the
 * reverse-mode gradient of one objective is mgf_adjoint (mgfvm_adjoint.h,
 * libmgfvm.so), which the tests check against this. */
int mgf_tangent_gradient(const mgf_problem *p, const mgf_options *opt, int max_batch,
                         double J[2], double *g_inlet, double g_amuf[2], double g_prsc[2],
                         int *converged);

#ifdef __cplusplus
}
#endif
#endif
