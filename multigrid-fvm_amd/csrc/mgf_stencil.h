/*
 * mgf_stencil.h -- device-side building blocks shared by the FP64 and the
 * mixed-precision kernels (internal): workgroup shape, the per-point
 * coefficient assembly of the reference's coefu / coefv / coefsc with their
 * boundary and obstacle adds, and the level-descriptor prologue.
 */
#ifndef MGF_STENCIL_H
#define MGF_STENCIL_H

#include <hip/hip_runtime.h>

#include "mgf_dev.h"
#include "mgf_plan.h"

#define BX 64
#define BY 4

struct Coef {
    double aw, ae, as, an, ap;
};

/* face fluxes from the materialised arrays (relaxation: cx/cy are computed
 * once at its start, V6:201, and the u/v/sc assemblies read those) */
struct FluxArr {
    const double *cx, *cy;
    __device__ double X(int k) const { return cx[k]; }
    __device__ double Y(int k) const { return cy[k]; }
};
/* face fluxes recomputed from the current velocities (restriction: the
 * reference recomputes them right before, V6:215-216, with nothing changing
 * u/v in between), same expression as fluxes (V6:690-691) */
struct FluxInl {
    const double *u, *v;
    double ax, ay;
    __device__ double X(int k) const { return (0.5 * u[k]) * ax; }
    __device__ double Y(int k) const { return (0.5 * v[k]) * ay; }
};

/* descriptor of this workgroup's level, and the early exit once its
 * instance's stopping test has fired */
#define LEVEL(Ls)                                 \
    const DevLevel &L = (Ls)[blockIdx.z];         \
    if (L.ctl->done) return;
#define PAIR(Fs, Cs)                              \
    const DevLevel &F = (Fs)[blockIdx.z];         \
    const DevLevel &C = (Cs)[blockIdx.z];         \
    if (C.ctl->done) return;

/* coefu + no-slip wall doubling at one u face (V6:825-858) */
template <class FX>
__device__ __forceinline__ Coef coefu_at(const DevLevel &L, const FX &f, int i, int j)
{
    const int nx = L.npx;
    const int ij = j * nx + i, ijs = ij - nx;
    double cxm = 0.5 * (f.X(ij) + f.X(ij - 1));
    double cxp = 0.5 * (f.X(ij) + f.X(ij + 1));
    double cym = 0.5 * (f.Y(ijs) + f.Y(ijs + 1));
    double cyp = 0.5 * (f.Y(ij) + f.Y(ij + 1));
    Coef c;
    c.aw = fmax(fabs(cxm), L.du_x) + cxm;
    c.ae = fmax(fabs(cxp), L.du_x) - cxp;
    c.as = fmax(fabs(cym), L.du_y) + cym;
    c.an = fmax(fabs(cyp), L.du_y) - cyp;
    c.ap = ((c.aw + c.ae) + c.as) + c.an;
    if (j == L.npy - 2) { c.an += L.du_y; c.ap += L.du_y; }
    if (j == 1) { c.as += L.du_y; c.ap += L.du_y; }
    return c;
}

/* coefv + side-wall adds at one v face (V6:976-1015) */
template <class FX>
__device__ __forceinline__ Coef coefv_at(const DevLevel &L, const FX &f, int i, int j)
{
    const int nx = L.npx;
    const int ij = j * nx + i, ijn = ij + nx, ijs = ij - nx;
    double cxm = 0.5 * (f.X(ij - 1) + f.X(ijn - 1));
    double cxp = 0.5 * (f.X(ij) + f.X(ijn));
    double cym = 0.5 * (f.Y(ij) + f.Y(ijs));
    double cyp = 0.5 * (f.Y(ij) + f.Y(ijn));
    Coef c;
    c.aw = fmax(fabs(cxm), L.du_x) + cxm;
    c.ae = fmax(fabs(cxp), L.du_x) - cxp;
    c.as = fmax(fabs(cym), L.du_y) + cym;
    c.an = fmax(fabs(cyp), L.du_y) - cyp;
    c.ap = ((c.aw + c.ae) + c.as) + c.an;
    if (i == nx - 2) { c.ae += L.du_y; c.ap += L.du_y; }   /* V6:1010 uses areay/dy */
    if (i == 1) { c.aw += L.du_y; c.ap += L.du_y; }
    return c;
}

/* coefsc at one cell including the obstacle diffusion adds that the
 * reference applies on the host in obstacle-list order (V6:872-963) */
template <class FX>
__device__ __forceinline__ Coef coefsc_at(const DevLevel &L, const FX &f, int i, int j,
                                          uint8_t fl)
{
    const int nx = L.npx;
    const int ij = j * nx + i, ijs = ij - nx;
    double cxm = f.X(ij - 1), cxp = f.X(ij), cym = f.Y(ijs), cyp = f.Y(ij);
    Coef c;
    c.aw = fmax(fabs(cxm), L.ds_x) + cxm;
    c.ae = fmax(fabs(cxp), L.ds_x) - cxp;
    c.as = fmax(fabs(cym), L.ds_y) + cym;
    c.an = fmax(fabs(cyp), L.ds_y) - cyp;
    c.ap = ((c.aw + c.ae) + c.as) + c.an;
    if (fl & MGF_F_SCADD) {
        uint64_t code = L.sccode[ij];
        int n_ae = (int)(code & 31u), n_aw = (int)((code >> 5) & 31u);
        int n_an = (int)((code >> 10) & 31u), n_as = (int)((code >> 15) & 31u);
        int n_ap = (int)((code >> 20) & 63u);
        for (int k = 0; k < n_ae; ++k) c.ae += L.ds_x;
        for (int k = 0; k < n_aw; ++k) c.aw += L.ds_x;
        for (int k = 0; k < n_an; ++k) c.an += L.ds_y;
        for (int k = 0; k < n_as; ++k) c.as += L.ds_y;
        uint64_t seq = code >> 26;
        for (int k = 0; k < n_ap; ++k) c.ap += ((seq >> k) & 1u) ? L.ds_y : L.ds_x;
    }
    return c;
}

#define RANGE2D(i0, i1, j0, j1)                                   \
    const int i = (i0) + (int)(blockIdx.x * BX + threadIdx.x);    \
    const int j = (j0) + (int)(blockIdx.y * BY + threadIdx.y);    \
    if (i > (i1) || j > (j1)) return;

static inline dim3 grid2d(int ni, int nj)
{
    return dim3((unsigned)((ni + BX - 1) / BX), (unsigned)((nj + BY - 1) / BY));
}


__device__ __forceinline__ double block_sum(double v, double *lds)
{
    const int tid = threadIdx.y * BX + threadIdx.x;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((tid & 63) == 0) lds[tid >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (tid == 0) t = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    __syncthreads();
    return t;
}

/* the field a smoother works on and the reference's solver extents
 * (V6:701-736: u (npx-2, npy-1), v (npx-1, npy-2), pp and sc (npx-1, npy-1)) */
__device__ __forceinline__ double *field_of(const DevLevel &L, int field, int &nx, int &ny)
{
    switch (field) {
    case 0: nx = L.npx - 2; ny = L.npy - 1; return L.u;
    case 1: nx = L.npx - 1; ny = L.npy - 2; return L.v;
    case 2: nx = L.npx - 1; ny = L.npy - 1; return L.pp;
    default: nx = L.npx - 1; ny = L.npy - 1; return L.sc;
    }
}

#endif
