/*
 * mgf_patterns_main.c -- mgfvm_patterns, the seeded replacement of the
 * reference's interactive pattern generators (create_patterns_staggered.c,
 * create_patterns_random.c): writes pattern_<k>.txt (the pattern.txt
 * grammar, V6:2367-2400) and pattern_<k>.pgm (the generators' P2 image,
 * black = blocked) for k = 1..count into a directory, plus manifest.txt with
 * one line per pattern: file, kind, parameters, seed, rectangles, blocked
 * fraction.  Pattern k draws from seed + k.
 *
 *   mgfvm_patterns staggered --n-base N --spacing S [--spacing-y S]
 *                  --shift quarter|half|full --count K [--seed S] [--out DIR]
 *       pattern_1 inline, pattern_2 staggered, pattern_3.. perturbed
 *       (the reference generator's numbering)
 *   mgfvm_patterns particles --n N --ratio R --pct P1,P2,P3 --count K ...
 *   mgfvm_patterns bernoulli --n-base N --porosity P --count K ...
 * common: --keep-edges (allow obstacles in the first and last i column),
 *         --strict (fail when a pattern exceeds the reference's 1000
 *         rectangles, V6:31)
 * Exit status: 0 written, 1 bad arguments or limit exceeded, 2 I/O error.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "mgfvm.h"

static int usage(void)
{
    fprintf(stderr,
            "usage: mgfvm_patterns staggered|particles|bernoulli [--n-base N | --n N] "
            "[--spacing S] [--spacing-y S] [--shift quarter|half|full] [--ratio R] "
            "[--pct P1,P2,P3] [--porosity P] --count K [--seed S] [--out DIR] "
            "[--keep-edges] [--strict]\n");
    return 1;
}

static int write_pattern(const char *dir, int k, const int *rects, int cnt, int n)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/pattern_%d.txt", dir, k);
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "%d\n", cnt);
    for (int r = 0; r < cnt; ++r)
        fprintf(f, "%d %d %d %d\n", rects[4 * r], rects[4 * r + 1], rects[4 * r + 2],
                rects[4 * r + 3]);
    fclose(f);
    unsigned char *g = (unsigned char *)calloc((size_t)n * n, 1);
    if (!g) return -1;
    for (int r = 0; r < cnt; ++r)
        for (int i = rects[4 * r] - 1; i < rects[4 * r + 1]; ++i)
            for (int j = rects[4 * r + 2] - 1; j < rects[4 * r + 3]; ++j)
                if (i >= 0 && i < n && j >= 0 && j < n) g[(size_t)i * n + j] = 1;
    snprintf(path, sizeof path, "%s/pattern_%d.pgm", dir, k);
    f = fopen(path, "w");
    if (!f) {
        free(g);
        return -1;
    }
    fprintf(f, "P2\n%d %d\n255\n", n, n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) fprintf(f, "%d ", g[(size_t)i * n + j] ? 0 : 255);
        fprintf(f, "\n");
    }
    fclose(f);
    free(g);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return usage();
    const char *kind = argv[1], *out = "patterns", *shift = "full";
    int n = 16, sx = 2, sy = -1, count = 0, keep_edges = 0, strict = 0;
    double ratio = 0.3, porosity = 0.8, pct[3] = {50.0, 30.0, 20.0};
    unsigned long long seed = 1;
    for (int a = 2; a < argc; ++a) {
        const char *o = argv[a];
        const char *v = a + 1 < argc ? argv[a + 1] : NULL;
        if (!strcmp(o, "--keep-edges")) keep_edges = 1;
        else if (!strcmp(o, "--strict")) strict = 1;
        else if (!v) return usage();
        else if (!strcmp(o, "--n-base") || !strcmp(o, "--n")) { n = atoi(v); ++a; }
        else if (!strcmp(o, "--spacing")) { sx = atoi(v); ++a; }
        else if (!strcmp(o, "--spacing-y")) { sy = atoi(v); ++a; }
        else if (!strcmp(o, "--shift")) { shift = v; ++a; }
        else if (!strcmp(o, "--ratio")) { ratio = atof(v); ++a; }
        else if (!strcmp(o, "--porosity")) { porosity = atof(v); ++a; }
        else if (!strcmp(o, "--pct")) {
            if (sscanf(v, "%lf,%lf,%lf", &pct[0], &pct[1], &pct[2]) != 3) return usage();
            ++a;
        }
        else if (!strcmp(o, "--count")) { count = atoi(v); ++a; }
        else if (!strcmp(o, "--seed")) { seed = strtoull(v, NULL, 10); ++a; }
        else if (!strcmp(o, "--out")) { out = v; ++a; }
        else return usage();
    }
    if (sy < 0) sy = sx;
    int scale = 1;
    if (!strcmp(shift, "quarter")) scale = 4;
    else if (!strcmp(shift, "half")) scale = 2;
    else if (strcmp(shift, "full")) return usage();
    const int is_stag = !strcmp(kind, "staggered"), is_part = !strcmp(kind, "particles"),
              is_bern = !strcmp(kind, "bernoulli");
    if (!(is_stag || is_part || is_bern) || count < 1 || n < 1) return usage();
    if (mkdir(out, 0777) && errno != EEXIST) {
        perror(out);
        return 2;
    }
    char mpath[4096];
    snprintf(mpath, sizeof mpath, "%s/manifest.txt", out);
    FILE *man = fopen(mpath, "w");
    if (!man) {
        perror(mpath);
        return 2;
    }
    fprintf(man, "# file kind parameters seed rectangles blocked_fraction\n");
    const int ng = is_stag ? n * scale : n;
    int status = 0;
    for (int k = 1; k <= count; ++k) {
        const uint64_t s = (uint64_t)seed + (uint64_t)k;
        int cnt;
        char params[256];
        if (is_stag) {
            cnt = mgf_pattern_staggered(n, sx, sy, scale, k, s, !keep_edges, NULL, 0);
            snprintf(params, sizeof params, "n_base=%d spacing=%d,%d scale=%d variant=%d", n, sx,
                     sy, scale, k);
        } else if (is_part) {
            cnt = mgf_pattern_particles(n, ratio, pct[0], pct[1], pct[2], s, !keep_edges, NULL, 0);
            snprintf(params, sizeof params, "n=%d ratio=%g pct=%g,%g,%g", n, ratio, pct[0],
                     pct[1], pct[2]);
        } else {
            cnt = mgf_pattern_bernoulli(n, porosity, s, !keep_edges, NULL, 0);
            snprintf(params, sizeof params, "n_base=%d porosity=%g", n, porosity);
        }
        if (cnt < 0) {
            fprintf(stderr, "ERROR: pattern %d: %s\n", k, mgf_error_string());
            fclose(man);
            return 1;
        }
        int *rects = (int *)malloc(sizeof(int) * 4 * (size_t)(cnt > 0 ? cnt : 1));
        if (!rects) return 2;
        if (is_stag) mgf_pattern_staggered(n, sx, sy, scale, k, s, !keep_edges, rects, cnt);
        else if (is_part) mgf_pattern_particles(n, ratio, pct[0], pct[1], pct[2], s, !keep_edges, rects, cnt);
        else mgf_pattern_bernoulli(n, porosity, s, !keep_edges, rects, cnt);
        long cells = 0;
        for (int r = 0; r < cnt; ++r)
            cells += (long)(rects[4 * r + 1] - rects[4 * r] + 1) * (rects[4 * r + 3] - rects[4 * r + 2] + 1);
        if (write_pattern(out, k, rects, cnt, ng)) {
            perror("pattern");
            free(rects);
            fclose(man);
            return 2;
        }
        fprintf(man, "pattern_%d.txt %s %s seed=%llu %d %.6f\n", k, kind, params,
                (unsigned long long)s, cnt, (double)cells / ((double)ng * ng));
        if (cnt > 1000) {
            fprintf(stderr, "%s pattern_%d: %d rectangles, above the reference's limit of 1000\n",
                    strict ? "ERROR:" : "warning:", k, cnt);
            if (strict) status = 1;
        }
        free(rects);
    }
    fclose(man);
    return status;
}
