/*
 * harness_b.c -- per-routine known-answer states from the reference's own
 * functions (test infrastructure; built by oracle/ref/Makefile from the
 * reference's source compiled with -Dmain=v6_reference_main, output only
 * under oracle/_ref/).
 *
 * usage (run inside a directory holding InputData and pattern.txt):
 *   ref_b NSTAGE out.bin
 * Sets the problem up with the reference's readInput / AllocateMemorySoA /
 * init / rest_geom, runs NSTAGE V-cycles of the reference schedule
 * (Version-6-26-04.c:197-246) to reach a non-trivial state (the "stage"),
 * writes the stage, then for every routine in the table restores the stage,
 * calls the reference routine once and writes the resulting state.
 *
 * Record format (little endian):
 *   char name[32], int32 level, float64 return value, then for every level:
 *   int32 npx, int32 npy, 25 arrays of npx*npy float64 in the order of
 *   oracle/orc_main.c, then tag as float64.
 * The first record is the stage, named "stage".
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct SoA {
    int npx, npy;
    int *isobs, *ieobs, *jsobs, *jeobs, *tag;
    int *jsinlt, *jeinlt, *jsoutlt, *jeoutlt;
    double dx, dy, amu;
    double *areax, *areay;
    double *ae, *aw, *an, *as, *ap;
    double *u, *v, *p, *pp, *sc, *su, *rs;
    double *apu, *apv, *spu, *spv, *spsc;
    double *resu, *resv, *ressc;
    double *c;
    double *cx, *cy;
};

extern int ngrid, nswpm, nswpp, isolve;

void readInput(void);
struct SoA *AllocateMemorySoA(void);
void init(struct SoA *);
void rest_geom(struct SoA *);
void fluxes(struct SoA *);
void solve_u(struct SoA *);
void solve_v(struct SoA *);
void solve_pprime(struct SoA *, int, int);
void solve_sc(struct SoA *);
void coefu(struct SoA *);
void umom(struct SoA *);
void coefv(struct SoA *);
void vmom(struct SoA *);
void coefp(struct SoA *);
double masserror(struct SoA *);
void update(struct SoA *);
void pbound(struct SoA *);
void outflow(struct SoA *);
void coefsc(struct SoA *);
void scalar(struct SoA *);
void solver(struct SoA *, double *, int, int);
void solvergs(struct SoA *, double *, int, int);
void restv(struct SoA *, struct SoA *);
void restru(struct SoA *, struct SoA *);
void restrv(struct SoA *, struct SoA *);
void restrsc(struct SoA *, struct SoA *);
void prolu(struct SoA *, struct SoA *);
void prolv(struct SoA *, struct SoA *);
void prolp(struct SoA *, struct SoA *);
void prolsc(struct SoA *, struct SoA *);

#define NARR 25
static double **arrays(struct SoA *l, double **out[NARR])
{
    double **t[NARR] = {&l->u, &l->v, &l->p, &l->pp, &l->sc, &l->su, &l->rs,
                        &l->resu, &l->resv, &l->ressc, &l->c, &l->ae, &l->aw,
                        &l->an, &l->as, &l->ap, &l->apu, &l->apv, &l->spu,
                        &l->spv, &l->spsc, &l->cx, &l->cy, &l->areax, &l->areay};
    memcpy(out, t, sizeof t);
    return NULL;
}

static double **g_saved;   /* [ngrid * NARR] */

static void save_stage(struct SoA *L)
{
    g_saved = (double **)calloc((size_t)ngrid * NARR, sizeof(double *));
    for (int ig = 0; ig < ngrid; ++ig) {
        double **a[NARR];
        arrays(&L[ig], a);
        size_t n = (size_t)L[ig].npx * L[ig].npy;
        for (int k = 0; k < NARR; ++k) {
            g_saved[ig * NARR + k] = (double *)malloc(n * sizeof(double));
            memcpy(g_saved[ig * NARR + k], *a[k], n * sizeof(double));
        }
    }
}

static void restore_stage(struct SoA *L)
{
    for (int ig = 0; ig < ngrid; ++ig) {
        double **a[NARR];
        arrays(&L[ig], a);
        size_t n = (size_t)L[ig].npx * L[ig].npy;
        for (int k = 0; k < NARR; ++k)
            memcpy(*a[k], g_saved[ig * NARR + k], n * sizeof(double));
    }
}

static void write_state(FILE *f, const char *name, int level, double ret, struct SoA *L)
{
    char nm[32] = {0};
    strncpy(nm, name, sizeof nm - 1);
    fwrite(nm, 1, 32, f);
    fwrite(&level, 4, 1, f);
    fwrite(&ret, 8, 1, f);
    for (int ig = 0; ig < ngrid; ++ig) {
        struct SoA *l = &L[ig];
        size_t n = (size_t)l->npx * l->npy;
        fwrite(&l->npx, 4, 1, f);
        fwrite(&l->npy, 4, 1, f);
        double **a[NARR];
        arrays(l, a);
        for (int k = 0; k < NARR; ++k) fwrite(*a[k], sizeof(double), n, f);
        double *t = (double *)malloc(n * sizeof(double));
        for (size_t k = 0; k < n; ++k) t[k] = (double)l->tag[k];
        fwrite(t, sizeof(double), n, f);
        free(t);
    }
}

static void vcycle(struct SoA *L, int icyc)
{
    for (int ig = 0; ig < ngrid; ++ig) {
        struct SoA *l = &L[ig];
        fluxes(l); solve_u(l); solve_v(l); solve_pprime(l, icyc, ig);
        pbound(l); outflow(l); solve_sc(l);
        if (ig != ngrid - 1) {
            struct SoA *c = &L[ig + 1];
            restv(l, c); fluxes(c); fluxes(l);
            restru(l, c); restrv(l, c); restrsc(l, c);
        }
    }
    for (int ig = ngrid - 1; ig > 0; ig--) {
        struct SoA *c = &L[ig], *f = &L[ig - 1];
        prolu(f, c); prolv(f, c); prolp(f, c); prolsc(f, c);
        if (ig != 1) {
            fluxes(f); solve_u(f); solve_v(f); solve_pprime(f, icyc, ig);
            pbound(f); solve_sc(f); outflow(f);
        }
    }
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s NSTAGE out.bin\n", argv[0]);
        return 2;
    }
    int nstage = atoi(argv[1]);
    readInput();
    struct SoA *L = AllocateMemorySoA();
    init(L);
    rest_geom(L);
    for (int icyc = 0; icyc < nstage; ++icyc) vcycle(L, icyc);
    save_stage(L);
    FILE *f = fopen(argv[2], "wb");
    if (!f) { perror("open"); return 1; }
    write_state(f, "stage", -1, 0.0, L);

    static const char *names[] = {
        "fluxes", "coefu", "umom", "coefv", "vmom", "coefsc", "scalar",
        "coefp", "masserror", "update", "pbound", "outflow", "jacobi_u",
        "jacobi_pp", "rbgs_u", "rbgs_pp", "solve_u", "solve_v",
        "solve_pprime", "solve_sc", "restv", "restru", "restrv", "restrsc",
        "prolu", "prolv", "prolp", "prolsc", NULL};
    for (int ig = 0; ig < ngrid; ++ig) {
        struct SoA *l = &L[ig], *c = (ig + 1 < ngrid) ? &L[ig + 1] : NULL;
        for (int k = 0; names[k]; ++k) {
            const char *n = names[k];
            int needs_c = !strncmp(n, "rest", 4) || !strncmp(n, "prol", 4);
            if (needs_c && !c) continue;
            restore_stage(L);
            double ret = 0.0;
            if (!strcmp(n, "fluxes")) fluxes(l);
            else if (!strcmp(n, "coefu")) coefu(l);
            else if (!strcmp(n, "umom")) umom(l);
            else if (!strcmp(n, "coefv")) coefv(l);
            else if (!strcmp(n, "vmom")) vmom(l);
            else if (!strcmp(n, "coefsc")) coefsc(l);
            else if (!strcmp(n, "scalar")) scalar(l);
            else if (!strcmp(n, "coefp")) coefp(l);
            else if (!strcmp(n, "masserror")) ret = masserror(l);
            else if (!strcmp(n, "update")) update(l);
            else if (!strcmp(n, "pbound")) pbound(l);
            else if (!strcmp(n, "outflow")) outflow(l);
            else if (!strcmp(n, "jacobi_u")) solver(l, l->u, l->npx - 2, l->npy - 1);
            else if (!strcmp(n, "jacobi_pp")) solver(l, l->pp, l->npx - 1, l->npy - 1);
            else if (!strcmp(n, "rbgs_u")) solvergs(l, l->u, l->npx - 2, l->npy - 1);
            else if (!strcmp(n, "rbgs_pp")) solvergs(l, l->pp, l->npx - 1, l->npy - 1);
            else if (!strcmp(n, "solve_u")) solve_u(l);
            else if (!strcmp(n, "solve_v")) solve_v(l);
            else if (!strcmp(n, "solve_pprime")) solve_pprime(l, 0, 1);
            else if (!strcmp(n, "solve_sc")) solve_sc(l);
            else if (!strcmp(n, "restv")) restv(l, c);
            else if (!strcmp(n, "restru")) restru(l, c);
            else if (!strcmp(n, "restrv")) restrv(l, c);
            else if (!strcmp(n, "restrsc")) restrsc(l, c);
            else if (!strcmp(n, "prolu")) prolu(l, c);
            else if (!strcmp(n, "prolv")) prolv(l, c);
            else if (!strcmp(n, "prolp")) prolp(l, c);
            else if (!strcmp(n, "prolsc")) prolsc(l, c);
            write_state(f, n, ig, ret, L);
        }
    }
    fclose(f);
    return 0;
}
