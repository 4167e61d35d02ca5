/*
 * mgf_batch_main.c -- mgfvm_batch, the batched replacement of
 * run_patterns_and_save_vtk.sh: for every Reynolds number given and every
 * pattern_<n>.txt of a directory, the InputData template with amuf = 1/Re,
 * all solved together as one batch on the GPU (instead of one process per
 * pattern), with the script's outputs:
 *   <out>/Re_<re>/vtk_files/vtk_<n>.vtk
 *   <out>/Re_<re>/convergence_files/convergence_<n>.csv
 * The script writes 1/Re through `bc` with scale=6, which truncates to six
 * decimals (run_patterns_and_save_vtk.sh:23); that is kept unless
 * --exact-re is given.
 *
 *   mgfvm_batch --re 20[,50,...] [--input InputData] [--patterns patterns]
 *               [--out solution] [--mixed] [--max-batch N] [--no-graph]
 *               [--exact-re] [--full-history]
 * Exit status: 0 done, 1 input error, 2 device error.
 */
#include <dirent.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "mgf_output.h"
#include "mgfvm.h"

struct job {
    double re;
    int number;          /* <n> of pattern_<n>.txt */
    mgf_problem *prob;
};

static int mkdirs(const char *path)
{
    char tmp[4096];
    snprintf(tmp, sizeof tmp, "%s", path);
    for (char *p = tmp + 1; *p; ++p) {
        if (*p != '/') continue;
        *p = 0;
        if (mkdir(tmp, 0777) && errno != EEXIST) return -1;
        *p = '/';
    }
    return (mkdir(tmp, 0777) && errno != EEXIST) ? -1 : 0;
}

static int cmp_int(const void *a, const void *b)
{
    return *(const int *)a - *(const int *)b;
}

static const char *re_tag(double re, char *buf, size_t n)
{
    snprintf(buf, n, "%g", re);
    return buf;
}

int main(int argc, char **argv)
{
    const char *input = "InputData", *pdir = "patterns", *out = "solution", *res = NULL;
    int mixed = 0, max_batch = 0, use_graph = 1, exact_re = 0, full_hist = 0;
    for (int a = 1; a < argc; ++a) {
        const char *o = argv[a];
        const char *v = a + 1 < argc ? argv[a + 1] : NULL;
        if (!strcmp(o, "--mixed")) mixed = 1;
        else if (!strcmp(o, "--no-graph")) use_graph = 0;
        else if (!strcmp(o, "--exact-re")) exact_re = 1;
        else if (!strcmp(o, "--full-history")) full_hist = 1;
        else if (!v) { fprintf(stderr, "ERROR: %s needs a value\n", o); return 1; }
        else if (!strcmp(o, "--re")) { res = v; ++a; }
        else if (!strcmp(o, "--input")) { input = v; ++a; }
        else if (!strcmp(o, "--patterns")) { pdir = v; ++a; }
        else if (!strcmp(o, "--out")) { out = v; ++a; }
        else if (!strcmp(o, "--max-batch")) { max_batch = atoi(v); ++a; }
        else { fprintf(stderr, "ERROR: unknown option %s\n", o); return 1; }
    }
    if (!res) {
        fprintf(stderr, "usage: mgfvm_batch --re R1[,R2...] [--input InputData] "
                        "[--patterns DIR] [--out DIR] [--mixed] [--max-batch N] [--no-graph] "
                        "[--exact-re] [--full-history]\n");
        return 1;
    }
    /* Reynolds numbers */
    double re[256];
    int nre = 0;
    for (const char *p = res; *p && nre < 256;) {
        char *end;
        double r = strtod(p, &end);
        if (end == p || !(r > 0.0)) { fprintf(stderr, "ERROR: bad Reynolds list %s\n", res); return 1; }
        re[nre++] = r;
        p = *end == ',' ? end + 1 : end;
    }
    /* pattern_<n>.txt files, in number order */
    DIR *d = opendir(pdir);
    if (!d) { perror(pdir); return 1; }
    int *nums = NULL, nn = 0, cap = 0;
    for (struct dirent *e; (e = readdir(d));) {
        int num;
        char tail[8];
        if (sscanf(e->d_name, "pattern_%d.%7s", &num, tail) == 2 && !strcmp(tail, "txt")) {
            if (nn == cap) {
                cap = cap ? 2 * cap : 64;
                nums = (int *)realloc(nums, sizeof(int) * (size_t)cap);
            }
            nums[nn++] = num;
        }
    }
    closedir(d);
    if (nn == 0) { fprintf(stderr, "ERROR: no pattern_*.txt in %s\n", pdir); return 1; }
    qsort(nums, (size_t)nn, sizeof(int), cmp_int);

    const int njob = nre * nn;
    struct job *jobs = (struct job *)calloc((size_t)njob, sizeof(struct job));
    for (int r = 0; r < nre; ++r) {
        /* the script's 1/Re: bc with scale=6 truncates */
        const double amuf = exact_re ? 1.0 / re[r] : floor(1.0e6 / re[r]) / 1.0e6;
        for (int k = 0; k < nn; ++k) {
            struct job *J = &jobs[r * nn + k];
            char ppath[4096];
            snprintf(ppath, sizeof ppath, "%s/pattern_%d.txt", pdir, nums[k]);
            J->re = re[r];
            J->number = nums[k];
            if (mgf_problem_read(input, ppath, &J->prob)) {
                fprintf(stderr, "ERROR: %s: %s\n", ppath, mgf_error_string());
                return 1;
            }
            J->prob->amuf = amuf;
        }
    }
    mgf_options opt;
    mgf_options_default(&opt);
    opt.precision = mixed ? MGF_PREC_MIXED : MGF_PREC_FP64;
    opt.use_graph = use_graph;
    const int chunk = max_batch > 0 ? max_batch : njob;
    int status = 0;
    for (int j0 = 0; j0 < njob && !status; j0 += chunk) {
        const int nb = njob - j0 < chunk ? njob - j0 : chunk;
        const mgf_problem **probs = (const mgf_problem **)malloc(sizeof(void *) * (size_t)nb);
        for (int b = 0; b < nb; ++b) probs[b] = jobs[j0 + b].prob;
        mgf_solver *s = NULL;
        int rc = mgf_batch_create(probs, nb, &opt, &s);
        free(probs);
        if (rc) {
            fprintf(stderr, "ERROR: %s\n", mgf_error_string());
            status = rc == MGF_ERR_DEVICE ? 2 : 1;
            break;
        }
        int *ncyc = (int *)calloc((size_t)nb, sizeof(int)), *conv = (int *)calloc((size_t)nb, sizeof(int));
        rc = mgf_batch_solve(s, ncyc, conv);
        if (rc && rc != MGF_ERR_NONFINITE) {
            fprintf(stderr, "ERROR: %s\n", mgf_last_error(s));
            status = 2;
        }
        for (int b = 0; b < nb && !status; ++b) {
            const struct job *J = &jobs[j0 + b];
            const mgf_problem *p = J->prob;
            char tag[64], dv[4096], dc[4096], path[4200];
            re_tag(J->re, tag, sizeof tag);
            snprintf(dv, sizeof dv, "%s/Re_%s/vtk_files", out, tag);
            snprintf(dc, sizeof dc, "%s/Re_%s/convergence_files", out, tag);
            if (mkdirs(dv) || mkdirs(dc)) { perror(out); status = 1; break; }
            int npx, npy;
            mgf_batch_level_dims(s, b, 0, &npx, &npy);
            const size_t n = (size_t)npx * npy;
            double *u = (double *)malloc(n * sizeof(double)), *v = (double *)malloc(n * sizeof(double));
            double *pr = (double *)malloc(n * sizeof(double)), *sc = (double *)malloc(n * sizeof(double));
            unsigned char *obs = (unsigned char *)malloc(n);
            const int cap_h = p->nvcycle > 0 ? p->nvcycle : 1;
            double *hist = (double *)calloc((size_t)cap_h, sizeof(double));
            mgf_batch_get_field(s, b, 0, "u", u);
            mgf_batch_get_field(s, b, 0, "v", v);
            mgf_batch_get_field(s, b, 0, "p", pr);
            mgf_batch_get_field(s, b, 0, "sc", sc);
            const int nh = mgf_batch_history(s, b, hist, cap_h);
            mgf_obstacle_mask(p, npx, npy, obs);
            printf("  -> pattern_%d.txt for Re = %s: %d V-cycles, %s, final error %e\n", J->number,
                   tag, ncyc[b], conv[b] ? "converged" : "not converged",
                   nh > 0 ? hist[nh - 1] : 0.0);
            snprintf(path, sizeof path, "%s/vtk_%d.vtk", dv, J->number);
            mgf_write_vtk(path, npx, npy, p->xl / (npx - 2), p->yl / (npy - 2), u, v, pr, sc, obs,
                          0, 0);
            snprintf(path, sizeof path, "%s/convergence_%d.csv", dc, J->number);
            const int rows = full_hist ? nh : (nh > 0 ? nh - 1 : 0);
            mgf_write_convergence(path, hist, rows, p->nvcycle);
            free(u); free(v); free(pr); free(sc); free(obs); free(hist);
        }
        free(ncyc);
        free(conv);
        mgf_destroy(s);
    }
    for (int j = 0; j < njob; ++j) mgf_problem_free(jobs[j].prob);
    free(jobs);
    free(nums);
    if (!status) printf("All processing completed successfully.\n");
    return status;
}
