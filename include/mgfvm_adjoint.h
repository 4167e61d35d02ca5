/*
 * mgfvm_adjoint.h -- reverse-mode discrete adjoint of the converged steady
 * solution (libmgfvm.so).
 *
 * At a converged state the V-cycle's fixed point is a zero of the level-0
 * steady residual R(x; theta): the unrelaxed u and v momentum of
 * coefu/umom and coefv/vmom, the continuity defect of masserror and the
 * scalar transport of coefsc/scalar (Version-6-26-04.c:741-860, 861-1016,
 * 1096-1125), with x the interior u, v, p and sc of level 0 and theta the
 * inlet u profile (level-0 west face), the east boundary u, amuf and prsc.
 * For an objective J(x, theta) the adjoint state lambda solves
 *     K^T lambda = dJ/dx,   K = dR/dx at the converged state,
 * and the gradient is dJ/dtheta = dJ/dtheta|_x - lambda^T dR/dtheta: one
 * linear solve per objective, whatever the number of parameters.
 *
 * K is the exact linearisation (the hybrid scheme's branches frozen where
 * the converged state takes them), applied in transposed form by a kernel
 * that recomputes every coefficient and its derivatives from the stored
 * converged state.  The solve is GMRES on the finest level, preconditioned
 * by V-cycles of the same multigrid hierarchy: on every level an adjoint
 * SIMPLE relaxation (red-black sweeps of the transposed scalar and momentum
 * operators, and the pressure-correction equation for the adjoint
 * continuity multiplier), the reference's restriction weights for the
 * residuals and its prolongation weights for the corrections.  All in FP64,
 * for every problem of a batch at once (one launch per same-shape level).
 *
 * The reference has no derivative capability; these calls replace none of
 * its routines.  Penalised unknowns (faces and cells of obstacles, 1e30 on
 * their diagonal) are held fixed, which moves the gradient by O(1e-30)
 * relative.  Every ghost value is held fixed too, which is exact for
 * Dirichlet ghosts (ibc 1) and discrete inlets and outlets; a solver with
 * any other boundary type (p' Dirichlet rows, periodic, outflow) is refused
 * with MGF_ERR_CONFIG by mgf_adjoint, mgf_batch_adjoint and
 * mgf_adjoint_apply.
 */
#ifndef MGFVM_ADJOINT_H
#define MGFVM_ADJOINT_H

#include "mgfvm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MGF_OBJ_PRESSURE_DROP 0   /* J1 = mean p of the first interior column - of the last */
#define MGF_OBJ_OUTLET_FLUX   1   /* J2 = sum over rows of areax * u_east * sc of the last cell */

typedef struct mgf_adjoint_opts {
    double rtol;       /* stop when ||dJ/dx - K^T lambda|| <= rtol ||dJ/dx|| (default 1e-11) */
    int max_iter;      /* GMRES iterations in all (default 600) */
    int restart;       /* Krylov vectors kept before a restart (default 60; the
                          workspace is restart + 4 level-0 vectors per problem) */
    int only_converged;/* 1 (default): problems whose last solve did not meet its
                          tolerance are skipped (gradients NaN, iters -1): the
                          adjoint is that of the steady solution */
    int single_colour; /* 1: the preconditioner's red-black sweeps as one launch per
                          colour and sweep instead of several sweeps per launch
                          (default 0); the same result bit for bit, for checking */
    int precond_cycles;/* V-cycles per use of the preconditioner, 1..4; 0 (default):
                          chosen by the batch size */
} mgf_adjoint_opts;

void mgf_adjoint_opts_default(mgf_adjoint_opts *o);

/* The two objectives of problem b at the solver's current level-0 state
 * (J[0]: pressure drop, J[1]: outlet scalar flux). */
int mgf_objectives(mgf_solver *s, int b, double J[2]);

/* The gradient of one objective for problem 0 of s at its current state
 * (solve it first): dJ/d(inlet row r) in g_inlet[r - 1] (r = 1..nfy; a row
 * direction also raises the east boundary u by 1/nfy, the mass-consistent
 * change), dJ/d(amuf) in *g_amuf, dJ/d(prsc) in *g_prsc.  Any output may
 * be NULL.  The adjoint call of the design's boundary: one reverse solve
 * per objective.  Works on FP64 and mixed-precision solvers (the state is
 * FP64 in both; the adjoint runs in FP64).  MGF_ERR_CONFIG if the solver's
 * last solve did not converge; MGF_ERR_NONFINITE if the adjoint solve did
 * not reach its tolerance within max_iter (outputs then hold the last
 * iterate's gradient). */
int mgf_adjoint(mgf_solver *s, int objective, double *g_inlet, double *g_amuf, double *g_prsc);

/* Every problem of a batch at once.  Outputs (any may be NULL), problem b
 * at index b or, for g_inlet, at offset sum_{k<b} nfy_k: J[b],
 * g_inlet, g_amuf[b], g_prsc[b], iters[b] (GMRES iterations, -1 for a
 * skipped problem), relres[b] (final ||dJ/dx - K^T lambda|| / ||dJ/dx||).
 * MGF_ERR_NONFINITE if a problem that was not skipped missed its
 * tolerance. */
int mgf_batch_adjoint(mgf_solver *s, int objective, const mgf_adjoint_opts *opt, double *J,
                      double *g_inlet, double *g_amuf, double *g_prsc, int *iters,
                      double *relres);

/* The adjoint state of problem b after the last mgf_adjoint /
 * mgf_batch_adjoint call: the multipliers of the u, v momentum, continuity
 * and scalar equations on level 0 (npx*npy each, zero at ghost and
 * penalised points); any may be NULL. */
int mgf_adjoint_get_state(mgf_solver *s, int b, double *lam_u, double *lam_v, double *lam_c,
                          double *lam_s);

/* Verification: one stage of the solve applied to problem b alone, at
 * the current state, on caller vectors of 4 * npx * npy doubles (level 0:
 * the u, v, continuity and scalar parts, each npx*npy).  op 0: out =
 * K^T in (the exact transposed Jacobian); op 1: out = the
 * preconditioner (the default number of cycles of the adjoint V-cycle
 * iteration from zero) applied to the right-hand side in. */
int mgf_adjoint_apply(mgf_solver *s, int b, int op, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
