/*
 * mgfvm_oracle.h -- CPU restatement of the reference solver (test
 * infrastructure only; never linked into the product library).
 */
#ifndef MGFVM_ORACLE_H
#define MGFVM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Everything InputData + pattern.txt carry (Version-6-26-04.c:27-39). */
typedef struct orc_config {
    int nfx, nfy, ngrid, ig_obs, ibeg_obs, ig_inlt, ig_outlt;
    int nvcycle, nswpm, nswpp, isolve;
    double omega, convergence;
    double xl, yl, amuf, prsc, relxm, relxp;
    int ibc[4];          /* xm, xp, ym, yp */
    double bc[4][4];     /* [side][u v p sc] */
    double guess[4];     /* u v p sc */
    double sc_obs;
    int nobs;
    int *obs;            /* nobs x {if, il, jf, jl} */
    int ninlt, noutlt;
    int *inlt_j, *outlt_j;     /* n x {jf, jl} */
    double *inlt_v, *outlt_v;  /* n x {u, v, sc} */
    double dpdxm, dpdym;
    double *inlet_u;     /* optional: nfy level-0 inlet face values */
} orc_config;

typedef struct orc_solver orc_solver;

const char *orc_last_error(void);
int orc_parse_files(const char *inputdata, const char *pattern, orc_config *c);
void orc_config_free(orc_config *c);
int orc_create(const orc_config *cfg, orc_solver **out);
int orc_create_from_files(const char *inputdata, const char *pattern, orc_solver **out);
int orc_create_from_files_inlet(const char *inputdata, const char *pattern,
                                const double *inlet_u, orc_solver **out);
void orc_destroy(orc_solver *s);
void orc_set_verbose(orc_solver *s, int v);
/* isolve 2 sweeps in lexicographic order instead of red-black (a variant
 * the GPU library defines; the reference has only red-black) */
void orc_set_gs_lex(orc_solver *s, int lex);
int orc_num_levels(const orc_solver *s);
int orc_level_dims(const orc_solver *s, int ig, int *npx, int *npy);
int orc_get(orc_solver *s, int ig, const char *name, double *out);
int orc_set(orc_solver *s, int ig, const char *name, const double *in);
double orc_relax(orc_solver *s, int ig, int leg);
void orc_restrict(orc_solver *s, int ig);
void orc_prolong(orc_solver *s, int igc);
double orc_vcycle(orc_solver *s, int icyc);
int orc_solve(orc_solver *s, int nvcycle, double tol, double *hist, int *ncycles, int *converged);
int orc_call(orc_solver *s, const char *fn, int ig, double *ret);

#ifdef __cplusplus
}
#endif
#endif
