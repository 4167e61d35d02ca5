/*
 * mgf_main.c -- command-line solver, a drop-in for the reference program's
 * main() (Version-6-26-04.c:180-276): reads InputData and pattern.txt from
 * the working directory (or the paths given), echoes the inputs as
 * readInput does (V6:293-397), solves on the GPU, prints one
 * "icycle = %d error = %e" line per V-cycle (the line GUI.py:527 parses),
 * the elapsed time, and writes vtk_acc.vtk and convergence.csv: stdout and
 * both files are the reference program's, byte for byte, except the timing
 * line's value (and the OBS block of a problem without obstacles, which the
 * reference leaves without data).  The converged flag and cycle count go to
 * stderr.
 *
 * usage: mgfvm_run [InputData [pattern.txt]] [--no-graph] [--full-history]
 * Exit status: 0 solved (converged or not), 1 input error, 2 device error,
 * 3 non-finite residual.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mgf_output.h"
#include "mgfvm.h"

static void echo_inputs(const mgf_problem *p)
{
    printf("nx_fine = %d, ny_fine = %d, ngrid = %d, nvcycle = %d, "
           "nswpm = %d, nswpp = %d, isolve = %d,omega = %lf, convergence = %e\n",
           p->nfx, p->nfy, p->ngrid, p->nvcycle, p->nswpm, p->nswpp, p->isolve, p->omega,
           p->convergence);
    printf("xlength = %lf, ylength = %lf, viscosity = %lf, scalar Prandtl number = %lf, "
           "relax_momentum = %lf, relax_press = %lf\n",
           p->xl, p->yl, p->amuf, p->prsc, p->relxm, p->relxp);
    printf("ibc_xm = %d, ibc_xp = %d, ibc_ym = %d, ibc_yp = %d\n", p->ibc[0], p->ibc[1],
           p->ibc[2], p->ibc[3]);
    static const char *side[4] = {"xm", "xp", "ym", "yp"};
    for (int s = 0; s < 4; ++s)
        printf("u_%s = %lf, v_%s = %lf, p_%s = %lf, sc_%s = %lf\n", side[s], p->bc[s][0],
               side[s], p->bc[s][1], side[s], p->bc[s][2], side[s], p->bc[s][3]);
    printf("u_guess = %lf, v_guess = %lf, p_guess = %lf, sc_guess = %lf\n", p->guess[0],
           p->guess[1], p->guess[2], p->guess[3]);
    printf("nobs %d \n", p->nobs);
    printf("obs_sc %lf\n", p->sc_obs);
    printf("ninlt %d\n", p->ninlt);
    for (int k = 0; k < p->ninlt; ++k)
        printf(" jfinlt = %d, jlinlt = %d, uinlt = %lf, vinlt = %lf,tinlt = %lf\n",
               p->inlt_j[2 * k], p->inlt_j[2 * k + 1], p->inlt_v[3 * k], p->inlt_v[3 * k + 1],
               p->inlt_v[3 * k + 2]);
    printf("noutlt %d \n", p->noutlt);
    for (int k = 0; k < p->noutlt; ++k)
        printf(" jfoutlt = %d, jloutlt = %d, uoutlt = %lf, voutlt = %lf,scoutlt = %lf\n",
               p->outlt_j[2 * k], p->outlt_j[2 * k + 1], p->outlt_v[3 * k],
               p->outlt_v[3 * k + 1], p->outlt_v[3 * k + 2]);
}

int main(int argc, char **argv)
{
    const char *fin = "InputData", *fpat = "pattern.txt";
    int use_graph = 1, full_hist = 0, npos = 0;
    for (int a = 1; a < argc; ++a) {
        if (!strcmp(argv[a], "--no-graph")) use_graph = 0;
        else if (!strcmp(argv[a], "--full-history")) full_hist = 1;
        else if (npos == 0) { fin = argv[a]; ++npos; }
        else if (npos == 1) { fpat = argv[a]; ++npos; }
    }
    mgf_problem *p = NULL;
    int rc = mgf_problem_read(fin, fpat, &p);
    if (rc) {
        fprintf(stderr, "ERROR: %s\n", mgf_error_string());
        return 1;
    }
    echo_inputs(p);
    mgf_options o;
    mgf_options_default(&o);
    o.use_graph = use_graph;
    mgf_solver *s = NULL;
    rc = mgf_create(p, &o, &s);
    if (rc) {
        fprintf(stderr, "ERROR: %s\n", mgf_error_string());
        mgf_problem_free(p);
        return rc == MGF_ERR_DEVICE ? 2 : 1;
    }
    int nv = p->nvcycle > 0 ? p->nvcycle : 0;
    double *hist = (double *)calloc((size_t)(nv > 0 ? nv : 1), sizeof(double));
    int ncyc = 0, conv = 0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    rc = mgf_solve(s, nv, p->convergence, hist, &ncyc, &conv);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (rc && rc != MGF_ERR_NONFINITE) {
        fprintf(stderr, "ERROR: %s\n", mgf_last_error(s));
        mgf_destroy(s);
        mgf_problem_free(p);
        free(hist);
        return 2;
    }
    for (int k = 0; k < ncyc; ++k) printf("icycle = %d error = %e\n", k, hist[k]);
    double secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    printf("simple solution took %f seconds to execute \n", secs);
    /* stdout carries only what the reference prints; the status goes to stderr */
    fprintf(stderr, "converged = %d after %d V-cycles\n", conv, ncyc);

    int npx, npy;
    mgf_level_dims(s, 0, &npx, &npy);
    size_t n = (size_t)npx * npy;
    double *u = (double *)malloc(n * sizeof(double)), *v = (double *)malloc(n * sizeof(double));
    double *pr = (double *)malloc(n * sizeof(double)), *sc = (double *)malloc(n * sizeof(double));
    unsigned char *obs = (unsigned char *)calloc(n, 1);
    mgf_get_fields(s, 0, u, v, pr, NULL, sc);
    mgf_obstacle_mask(p, npx, npy, obs);
    /* The reference writes its outputs when convergence_data[icyc] < tol
     * after the loop, which also holds when the cycle limit was reached
     * (the entry one past the last cycle is still 0.0, V6:265); it writes
     * one CSV row fewer than the cycles run (V6:2552).  Both kept. */
    mgf_write_vtk("vtk_acc.vtk", npx, npy, p->xl / (npx - 2), p->yl / (npy - 2), u, v, pr, sc,
                  obs, 0, 1);
    int rows = full_hist ? ncyc : (ncyc > 0 ? ncyc - 1 : 0);
    mgf_write_convergence("convergence.csv", hist, rows, p->nvcycle);
    free(u); free(v); free(pr); free(sc); free(obs); free(hist);
    mgf_destroy(s);
    mgf_problem_free(p);
    return rc == MGF_ERR_NONFINITE ? 3 : 0;
}
