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
    real_t aw, ae, as, an, ap;
};

/* face fluxes from the materialised arrays (relaxation: cx/cy are computed
 * once at its start, V6:201, and the u/v/sc assemblies read those) */
struct FluxArr {
    const gdbl *cx, *cy;
    __device__ real_t X(int k) const { return cx[k]; }
    __device__ real_t Y(int k) const { return cy[k]; }
};
/* face fluxes recomputed from the current velocities (restriction: the
 * reference recomputes them right before, V6:215-216, with nothing changing
 * u/v in between), same expression as fluxes (V6:690-691) */
struct FluxInl {
    const gdbl *u, *v;
    real_t ax, ay;
    __device__ real_t X(int k) const { return (0.5 * u[k]) * ax; }
    __device__ real_t Y(int k) const { return (0.5 * v[k]) * ay; }
};

/* Workgroups of a launch go to the 8 XCDs round-robin by linear id, and
 * each XCD has its own L2.  2-D launches pad their rows of tiles to a
 * multiple of 8 (tile_grid: gridDim.y = 8 b), and tile_id() hands XCD k the
 * rows [k b, (k+1) b) of every group member: tiles whose halos overlap (same
 * row, or neighbouring rows of one band) share an L2, and every member is
 * spread over all XCDs, so members that are still iterating after others
 * have stopped keep the whole chip busy.  Padding rows lie past the level's
 * extent, where the kernels' range checks skip them.  Grids whose gridDim.y
 * is not a multiple of 8 keep the hardware order. */
#define MGF_NXCD 8u
struct TileId {
    int x, y, z;
};
__device__ __forceinline__ TileId tile_id()
{
    if (gridDim.y % MGF_NXCD) return TileId{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
    const unsigned gx = gridDim.x, b = gridDim.y / MGF_NXCD;
    const unsigned lin = blockIdx.x + gx * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned k = lin % MGF_NXCD, m = lin / MGF_NXCD, B = b * gx;
    const unsigned z = m / B, o = m - z * B, row = o / gx;
    return TileId{(int)(o - row * gx), (int)(k * b + row), (int)z};
}

/* launch grid for ni x nj points per member in tiles of BX x by, rows of
 * tiles padded to a multiple of 8 */
static inline dim3 tile_grid(int ni, int nj, int nz, int by = BY)
{
    const unsigned gy = (unsigned)((nj + by - 1) / by);
    return dim3((unsigned)((ni + BX - 1) / BX), (gy + MGF_NXCD - 1) / MGF_NXCD * MGF_NXCD,
                (unsigned)nz);
}

/* descriptor of this workgroup's level (T = its logical tile), and the
 * early exit once its instance's stopping test has fired */
#define LEVEL(Ls)                                 \
    const TileId T = tile_id();                   \
    const DevLevelG &L = reinterpret_cast<const DevLevelG *>(Ls)[T.z]; \
    if (L.ctl->done) return;
/* the same for the 1-D boundary kernels (no tiles, no renumbering) */
#define LEVEL1D(Ls)                               \
    const DevLevelG &L = reinterpret_cast<const DevLevelG *>(Ls)[blockIdx.z]; \
    if (L.ctl->done) return;
#define PAIR(Fs, Cs)                              \
    const TileId T = tile_id();                   \
    const DevLevelG &F = reinterpret_cast<const DevLevelG *>(Fs)[T.z]; \
    const DevLevelG &C = reinterpret_cast<const DevLevelG *>(Cs)[T.z]; \
    if (C.ctl->done) return;

/* coefu + no-slip wall doubling at one u face in row j (V6:825-858) from
 * the face fluxes around it: cx at (i-1, i, i+1, row j), cy at (i, i+1) of
 * rows j-1 and j */
__device__ __forceinline__ Coef coefu_vals(const DevLevelG &L, int j, real_t xw, real_t x0,
                                           real_t xe, real_t ys0, real_t ys1, real_t y0, real_t y1)
{
    real_t cxm = 0.5 * (x0 + xw);
    real_t cxp = 0.5 * (x0 + xe);
    real_t cym = 0.5 * (ys0 + ys1);
    real_t cyp = 0.5 * (y0 + y1);
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

template <class FX>
__device__ __forceinline__ Coef coefu_at(const DevLevelG &L, const FX &f, int i, int j)
{
    const int nx = L.npx;
    const int ij = j * nx + i, ijs = ij - nx;
    return coefu_vals(L, j, f.X(ij - 1), f.X(ij), f.X(ij + 1), f.Y(ijs), f.Y(ijs + 1), f.Y(ij),
                      f.Y(ij + 1));
}

/* coefv + side-wall adds at one v face in column i (V6:976-1015) from cx
 * at (i-1, i) of rows j and j+1 and cy at rows j-1, j, j+1 of column i */
__device__ __forceinline__ Coef coefv_vals(const DevLevelG &L, int i, real_t xw, real_t xwn,
                                           real_t x0, real_t xn, real_t ys, real_t y0, real_t yn)
{
    const int nx = L.npx;
    real_t cxm = 0.5 * (xw + xwn);
    real_t cxp = 0.5 * (x0 + xn);
    real_t cym = 0.5 * (y0 + ys);
    real_t cyp = 0.5 * (y0 + yn);
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

template <class FX>
__device__ __forceinline__ Coef coefv_at(const DevLevelG &L, const FX &f, int i, int j)
{
    const int nx = L.npx;
    const int ij = j * nx + i, ijn = ij + nx, ijs = ij - nx;
    return coefv_vals(L, i, f.X(ij - 1), f.X(ijn - 1), f.X(ij), f.X(ijn), f.Y(ijs), f.Y(ij),
                      f.Y(ijn));
}

/* coefsc at cell ij including the obstacle diffusion adds that the
 * reference applies on the host in obstacle-list order (V6:872-963), from
 * the fluxes through its four faces */
__device__ __forceinline__ Coef coefsc_vals(const DevLevelG &L, int ij, uint8_t fl, real_t cxm,
                                            real_t cxp, real_t cym, real_t cyp)
{
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

template <class FX>
__device__ __forceinline__ Coef coefsc_at(const DevLevelG &L, const FX &f, int i, int j,
                                          uint8_t fl)
{
    const int nx = L.npx;
    const int ij = j * nx + i, ijs = ij - nx;
    return coefsc_vals(L, ij, fl, f.X(ij - 1), f.X(ij), f.Y(ijs), f.Y(ij));
}

/* Mixed precision: the value an FP32 correction d leaves at a point of its
 * swept range.  A penalised point (1e30 added to its diagonal: flag pen) is
 * set rather than corrected.  Its FP64 Gauss-Seidel value is
 * (su + sum a_nb phi_nb) / (ap + 1e30), which is `target` to ~1e-30
 * absolute: 0 for u and v faces, sc_obs for sc cells.  phi + d would keep
 * the FP32 rounding of d ~ -phi, ~1e-7 |phi|; that decays only by 1e-7 per
 * relaxation, and a fluid cell whose four faces are all penalised turns its
 * divergence into p' through 1/ap ~ 1e33. */
__device__ __forceinline__ real_t pending_value(real_t phi, float d, bool pen, real_t target)
{
    return pen ? target : phi + (real_t)d;
}

#define RANGE2D(i0, i1, j0, j1)                                   \
    const int i = (i0) + (int)(T.x * BX + threadIdx.x);           \
    const int j = (j0) + (int)(T.y * BY + threadIdx.y);           \
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
__device__ __forceinline__ gdbl *field_of(const DevLevelG &L, int field, int &nx, int &ny)
{
    switch (field) {
    case 0: nx = L.npx - 2; ny = L.npy - 1; return L.u;
    case 1: nx = L.npx - 1; ny = L.npy - 2; return L.v;
    case 2: nx = L.npx - 1; ny = L.npy - 1; return L.pp;
    default: nx = L.npx - 1; ny = L.npy - 1; return L.sc;
    }
}

#endif
