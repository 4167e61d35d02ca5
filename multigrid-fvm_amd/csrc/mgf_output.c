/*
 * mgf_output.c -- result files in the reference's formats.
 *
 *   mgf_write_vtk: ASCII structured-grid VTK of cell-centred velocity,
 *       pressure, scalar and obstacle mask, as write_vtk_2d_cc
 *       (Version-6-26-04.c:2261-2363), including its averages and
 *       tortuosity printout.  When there are no obstacles the OBS block is
 *       filled with zeros (the reference writes its header without data,
 *       which VTK readers reject).
 *   mgf_write_convergence: "Iteration,convergence_data" CSV as
 *       write_convergence_data_file (V6:2544-2558).
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "mgf_output.h"
#include "mgfvm.h"

void mgf_obstacle_mask(const mgf_problem *p, int npx, int npy, unsigned char *obs)
{
    memset(obs, 0, (size_t)npx * npy);
    if (p->nobs <= 0 || p->ig_obs < 1) return;
    const int fac = 1 << (p->ig_obs - 1);
    for (int k = 0; k < p->nobs; ++k) {
        const int *ob = p->obs + 4 * k;
        const int is = (ob[0] + p->ibeg_obs - 1) * fac + 1, ie = (ob[1] + p->ibeg_obs) * fac;
        const int js = (ob[2] - 1) * fac + 1, je = ob[3] * fac;
        for (int j = js; j <= je; ++j)
            for (int i = is; i <= ie; ++i)
                if (i >= 0 && i < npx && j >= 0 && j < npy) obs[(size_t)j * npx + i] = 1;
    }
}

int mgf_write_vtk(const char *fname, int npx, int npy, double dx, double dy, const double *u,
                  const double *v, const double *p, const double *sc, const unsigned char *obs,
                  int iter, int print_stats)
{
    int I = npx - 2, J = npy - 2, N = I * J;
    double sum_vx = 0.0, sum_mag = 0.0;
    FILE *fp = fopen(fname, "w");
    if (!fp) {
        perror("fopen");
        return -1;
    }
    fprintf(fp, "# vtk DataFile Version 3.0\n");
    fprintf(fp, "2D SIMPLE output iter %d\n", iter);
    fprintf(fp, "ASCII\n");
    fprintf(fp, "DATASET STRUCTURED_GRID\n");
    fprintf(fp, "DIMENSIONS %d %d 1\n", I, J);
    fprintf(fp, "POINTS %d float\n", N);
    for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) fprintf(fp, "%f %f 0.0\n", (i - 0.5) * dx, (j - 0.5) * dy);
    fprintf(fp, "\nPOINT_DATA %d\n", N);
    fprintf(fp, "VECTORS Velocity float\n");
    for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) {
            int ij = j * npx + i;
            double ucc = 0.5 * (u[ij] + u[ij - 1]);
            double vcc = 0.5 * (v[ij] + v[ij - npx]);
            sum_vx += ucc;
            sum_mag += sqrt(ucc * ucc + vcc * vcc);
            fprintf(fp, "%f %f 0.0\n", ucc, vcc);
        }
    fprintf(fp, "\nSCALARS P float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) fprintf(fp, "%f\n", p[j * npx + i]);
    fprintf(fp, "\nSCALARS SC float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) fprintf(fp, "%f\n", sc[j * npx + i]);
    fprintf(fp, "\nSCALARS OBS float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int j = 1; j <= J; ++j)
        for (int i = 1; i <= I; ++i) fprintf(fp, "%f\n", (obs && obs[j * npx + i]) ? 1.0 : 0.0);
    fclose(fp);
    if (print_stats) {
        printf("Average velocity magnitude at iter %d: %e\n", iter, sum_mag / N);
        printf("Average x-velocity at iter %d: %e\n", iter, sum_vx / N);
        printf("Tortuosity at iter %d: %e\n", iter, sum_mag / sum_vx);
    }
    return 0;
}

int mgf_write_convergence(const char *fname, const double *hist, int nrows, int nvcycle)
{
    FILE *fp = fopen(fname, "w");
    if (!fp) {
        fprintf(stderr, "ERROR: Could not open convergence_data file %s for writing.\n", fname);
        return -1;
    }
    fprintf(fp, "Iteration,convergence_data\n");
    for (int i = 0; i < nrows; i++) fprintf(fp, "%d,%e\n", i + 1, hist[i]);
    fclose(fp);
    printf("Successfully wrote %d convergence_data records to %s\n", nvcycle, fname);
    return 0;
}
