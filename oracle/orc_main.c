/*
 * orc_main.c -- command-line front end of the CPU restatement (test
 * infrastructure).  Runs a whole solve on InputData + pattern.txt and writes
 * the final state of every level in the dump format shared with the
 * reference harness (oracle/ref/harness_a.c):
 *
 *   "MGFD" int32 version=1, int32 ngrid, int32 nhist
 *   per level: int32 npx, int32 npy, then NDUMP arrays of npx*npy float64
 *              in the order of orc_dump_names, then tag as float64
 *   nhist float64 residual-history entries
 *
 * usage: orc_run InputData pattern.txt out.bin [nvcycle_override]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mgfvm_oracle.h"

static const char *orc_dump_names[] = {
    "u", "v", "p", "pp", "sc", "su", "rs", "resu", "resv", "ressc", "c",
    "ae", "aw", "an", "as", "ap", "apu", "apv", "spu", "spv", "spsc",
    "cx", "cy", "areax", "areay", NULL};

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s InputData pattern.txt out.bin [nvcycle]\n", argv[0]);
        return 2;
    }
    orc_config cfg;
    if (orc_parse_files(argv[1], argv[2], &cfg)) {
        fprintf(stderr, "parse: %s\n", orc_last_error());
        return 1;
    }
    if (argc > 4) cfg.nvcycle = atoi(argv[4]);
    orc_solver *s = NULL;
    if (orc_create(&cfg, &s)) {
        fprintf(stderr, "create: %s\n", orc_last_error());
        return 1;
    }
    int nv = cfg.nvcycle;
    double *hist = (double *)calloc(nv > 0 ? (size_t)nv : 1, sizeof(double));
    int ncyc = 0, conv = 0;
    orc_set_verbose(s, getenv("ORC_VERBOSE") != NULL);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    orc_solve(s, nv, cfg.convergence, hist, &ncyc, &conv);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double secs = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    fprintf(stderr, "cycles %d converged %d last %.17g seconds %.6f\n", ncyc, conv,
            ncyc ? hist[ncyc - 1] : 0.0, secs);

    FILE *f = fopen(argv[3], "wb");
    if (!f) { perror("open dump"); return 1; }
    int ng = orc_num_levels(s), ver = 1;
    fwrite("MGFD", 1, 4, f);
    fwrite(&ver, 4, 1, f);
    fwrite(&ng, 4, 1, f);
    fwrite(&ncyc, 4, 1, f);
    for (int ig = 0; ig < ng; ++ig) {
        int npx, npy;
        orc_level_dims(s, ig, &npx, &npy);
        fwrite(&npx, 4, 1, f);
        fwrite(&npy, 4, 1, f);
        double *buf = (double *)malloc(sizeof(double) * (size_t)npx * npy);
        for (int k = 0; orc_dump_names[k]; ++k) {
            orc_get(s, ig, orc_dump_names[k], buf);
            fwrite(buf, sizeof(double), (size_t)npx * npy, f);
        }
        orc_get(s, ig, "tag", buf);
        fwrite(buf, sizeof(double), (size_t)npx * npy, f);
        free(buf);
    }
    fwrite(hist, sizeof(double), (size_t)ncyc, f);
    fclose(f);
    free(hist);
    orc_destroy(s);
    orc_config_free(&cfg);
    return 0;
}
