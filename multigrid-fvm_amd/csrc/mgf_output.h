#ifndef MGF_OUTPUT_H
#define MGF_OUTPUT_H

#ifdef __cplusplus
extern "C" {
#endif

int mgf_write_vtk(const char *fname, int npx, int npy, double dx, double dy, const double *u,
                  const double *v, const double *p, const double *sc, const unsigned char *obs,
                  int iter, int print_stats);
int mgf_write_convergence(const char *fname, const double *hist, int nrows, int nvcycle);
struct mgf_problem;
/* level-0 obstacle mask (1 in blocked cells) of a problem, npx*npy bytes,
 * from the obstacle rectangles scaled as AllocateMemorySoA does (V6:471-478) */
void mgf_obstacle_mask(const struct mgf_problem *p, int npx, int npy, unsigned char *obs);

#ifdef __cplusplus
}
#endif
#endif
