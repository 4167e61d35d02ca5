/*
 * mgf_kernels.hip -- FP64 kernels of the SIMPLE relaxation, smoothers and
 * FAS grid transfers for gfx950 (MI355X).
 *
 * Arithmetic: every value is computed with the reference's expression tree
 * (C left-to-right association), IEEE add/mul/div/fmax/fabs only, built with
 * -ffp-contract=off so no FMA is formed.  The results are therefore
 * bit-identical to the reference CPU build (Version-6-26-04.c compiled with
 * -ffp-contract=off); only the order of the sum of |div| differs (a fixed
 * tree here, so it is run-to-run deterministic).
 *
 * Mapping: one thread per grid point, 64x4 workgroups (one wavefront per
 * 64-point row segment, so every array access of a wavefront is one
 * contiguous 512-byte run of a row).  The work is HBM/MALL-bandwidth bound
 * stencil code (about 0.16 flop per byte), so no matrix cores are used.
 *
 * Early exit: every kernel returns immediately once the device control
 * block says the stopping test has fired, so V-cycles enqueued past
 * convergence change nothing and the cycle count equals the reference's.
 */
#include <hip/hip_runtime.h>

#include "mgf_stencil.h"

/* ---------------------------------------------------------------------- */
/* fluxes, V6:678-694: j in [0,npy-2], i in [0,npx-2]                       */
/* ---------------------------------------------------------------------- */
__global__ void __launch_bounds__(BX * BY) k_fluxes(const DevLevel *Ls)
{
    LEVEL(Ls);
    RANGE2D(0, L.npx - 2, 0, L.npy - 2);
    const int ij = j * L.npx + i;
    L.cx[ij] = (0.5 * L.u[ij]) * L.ax;
    L.cy[ij] = (0.5 * L.v[ij]) * L.ay;
}

/* ---------------------------------------------------------------------- */
/* coefu + walls + umom, V6:741-767, 815-860: j in [1,npy-2], i in [1,npx-3] */
/* ---------------------------------------------------------------------- */
/* the stored FP64 equation of one point: coefficients, diagonal, source */
struct Pt {
    real_t aw, ae, as, an, ap, su;
};

/* face fluxes of the relaxation: the k_fluxes arrays, or (INL) formed from
 * u, v with the same expression while u, v still hold the values the
 * relaxation started from */
template <bool INL>
__device__ __forceinline__ auto relax_flux(const DevLevelG &L)
{
    if constexpr (INL)
        return FluxInl{L.u, L.v, L.ax, L.ay};
    else
        return FluxArr{L.cx, L.cy};
}

/* coefu + walls + umom at u face (i, j), V6:741-767, 815-860 */
template <bool INL = false>
__device__ __forceinline__ Pt asm64_u(const DevLevelG &L, int i, int j)
{
    const DevPhys &P = L.P;
    const int ij = j * L.npx + i;
    Coef c = coefu_at(L, relax_flux<INL>(L), i, j);
    real_t su = L.ax * (L.p[ij] - L.p[ij + 1]);
    real_t ap = c.ap / P.relxm;
    su = ((su + (P.relxm1 * ap) * L.u[ij]) + L.resu[ij]) + L.bodyx;
    ap = ap + ((L.flags[ij] & MGF_F_SPU) ? 1.0e30 : 0.0);
    return Pt{c.aw, c.ae, c.as, c.an, ap, su};
}

/* coefv + side walls + vmom at v face (i, j), V6:769-793, 966-1016 */
template <bool INL = false>
__device__ __forceinline__ Pt asm64_v(const DevLevelG &L, int i, int j)
{
    const DevPhys &P = L.P;
    const int ij = j * L.npx + i;
    Coef c = coefv_at(L, relax_flux<INL>(L), i, j);
    real_t su = L.ay * (L.p[ij] - L.p[ij + L.npx]);
    real_t ap = c.ap / P.relxm;
    su = ((su + (P.relxm1 * ap) * L.v[ij]) + L.resv[ij]) + L.bodyy;
    ap = ap + ((L.flags[ij] & MGF_F_SPV) ? 1.0e30 : 0.0);
    return Pt{c.aw, c.ae, c.as, c.an, ap, su};
}

/* coefsc + obstacle adds + scalar at cell (i, j), V6:795-812, 861-964 */
template <bool INL = false>
__device__ __forceinline__ Pt asm64_sc(const DevLevelG &L, int i, int j)
{
    const DevPhys &P = L.P;
    const int ij = j * L.npx + i;
    const uint8_t fl = L.flags[ij];
    Coef c = coefsc_at(L, relax_flux<INL>(L), i, j, fl);
    const double spsc = (fl & MGF_F_SPSC) ? 1.0e30 : 0.0;
    return Pt{c.aw, c.ae, c.as, c.an, c.ap + spsc, spsc * P.sc_obs + L.ressc[ij]};
}

/* masserror + coefp (+ p' Dirichlet rows) at cell (i, j), V6:1018-1125;
 * mag = |div|.  PEND: the u and v relaxations left their results over their
 * swept ranges (u: i <= npx-3, v: j <= npy-3) in uc, vc for the update */
template <bool PEND = false>
__device__ __forceinline__ Pt asm64_p(const DevLevelG &L, int i, int j, double &mag)
{
    const DevPhys &P = L.P;
    const int nx = L.npx;
    const int ij = j * nx + i, ijs = ij - nx;
    const auto *ue = (PEND && i <= nx - 3) ? L.uc : L.u;
    const auto *uw = (PEND && i >= 2) ? L.uc : L.u;
    const auto *vn = (PEND && j <= L.npy - 3) ? L.vc : L.v;
    const auto *vs = (PEND && j >= 2) ? L.vc : L.v;
    real_t div = ((L.ax * ue[ij] - L.ax * uw[ij - 1]) + L.ay * vn[ij]) - L.ay * vs[ijs];
    real_t su = -div;
    mag = val(fabs(div));
    real_t aw = L.axax / L.apu[ij - 1];
    real_t ae = L.axax / L.apu[ij];
    real_t as = L.ayay / L.apv[ijs];
    real_t an = L.ayay / L.apv[ij];
    real_t ap = (((aw + ae) + as) + an) + ((L.flags[ij] & MGF_F_OBS) ? 1.0e30 : 0.0);
    const double diag = 1.0e30;
    if (P.ibc[0] == 2 && i == 1) { ap = ap + diag; su = 0.0; }
    if (P.ibc[1] == 2 && i == nx - 2) { ap = ap + diag; su = 0.0; }
    if (P.ibc[2] == 2 && j == 1) { ap = ap + diag; su = 0.0; }
    if (P.ibc[3] == 2 && j == L.npy - 2) { ap = ap + diag; su = 0.0; }
    return Pt{aw, ae, as, an, ap, su};
}

__device__ __forceinline__ void store_eq(const DevLevelG &L, int ij, const Pt &q)
{
    L.aw[ij] = q.aw; L.ae[ij] = q.ae; L.as[ij] = q.as; L.an[ij] = q.an;
    L.ap[ij] = q.ap;
    L.su[ij] = q.su;
}

/* j in [1,npy-2], i in [1,npx-3] */
__global__ void __launch_bounds__(BX * BY) k_asm_u(const DevLevel *Ls)
{
    LEVEL(Ls);
    RANGE2D(1, L.npx - 3, 1, L.npy - 2);
    const int ij = j * L.npx + i;
    const Pt q = asm64_u(L, i, j);
    store_eq(L, ij, q);
    L.apu[ij] = q.ap;
}

/* j in [1,npy-3], i in [1,npx-2] */
__global__ void __launch_bounds__(BX * BY) k_asm_v(const DevLevel *Ls)
{
    LEVEL(Ls);
    RANGE2D(1, L.npx - 2, 1, L.npy - 3);
    const int ij = j * L.npx + i;
    const Pt q = asm64_v(L, i, j);
    store_eq(L, ij, q);
    L.apv[ij] = q.ap;
}

/* [1,npy-2]x[1,npx-2] */
__global__ void __launch_bounds__(BX * BY) k_asm_sc(const DevLevel *Ls)
{
    LEVEL(Ls);
    RANGE2D(1, L.npx - 2, 1, L.npy - 2);
    store_eq(L, j * L.npx + i, asm64_sc(L, i, j));
}

/* ---------------------------------------------------------------------- */
/* masserror + coefp (+ p' Dirichlet rows), V6:1018-1125                    */
/* also the sum of |div| with a deterministic two-level tree               */
/* ---------------------------------------------------------------------- */
template <bool PEND>
__global__ void __launch_bounds__(BX * BY) k_asm_p(const DevLevel *Ls)
{
    LEVEL(Ls);
    const DevPhys &P = L.P;
    __shared__ double lds[BY];
    const int i = 1 + (int)(T.x * BX + threadIdx.x);
    const int j = 1 + (int)(T.y * BY + threadIdx.y);
    const int nx = L.npx;
    double mag = 0.0;
    if (i <= nx - 2 && j <= L.npy - 2) {
        const int ij = j * nx + i;
        store_eq(L, ij, asm64_p<PEND>(L, i, j, mag));
        L.pp[ij] = 0.0;
    }
    double bs = block_sum(mag, lds);
    const int pb = T.y * (int)gridDim.x + T.x;     /* padding rows have none */
    if (threadIdx.x == 0 && threadIdx.y == 0 && pb < L.nblk) L.partial[pb] = bs;
}

/* sum of the per-block partials in a fixed order (one workgroup), so the
 * level's sum |div| is the same bits on every run */
__global__ void __launch_bounds__(BX * BY) k_reduce_err(const DevLevel *Ls, int record)
{
    LEVEL(Ls);
    __shared__ double lds[BY];
    const int tid = threadIdx.y * BX + threadIdx.x;
    double t = 0.0;
    for (int k = tid; k < L.nblk; k += BX * BY) t += L.partial[k];
    double tot = block_sum(t, lds);
    if (tid == 0) {
        L.ctl->err[L.ig] = tot;
        /* the level-0 value of the down-leg is the instance's history entry
         * (V6:721-724) */
        if (record && L.ig == 0) L.ctl->hist[L.ctl->cyc] = tot;
    }
}

/* ---------------------------------------------------------------------- */
/* smoothers.  (nx, ny) are the reference's solver arguments: points        */
/* j in [1,ny-1], i in [1,nx-1] (V6:1265-1266, 1314-1315)                   */
/* ---------------------------------------------------------------------- */

/* one Jacobi sweep src -> dst; neighbours outside the swept range are read
 * from `base` (the field itself), which is what the reference's full copy
 * of phi into phi_new provides (V6:1254).  wj = 1 (V6:1241). */
__global__ void __launch_bounds__(BX * BY) k_jacobi(const DevLevel *Ls, int field, int parity)
{
    LEVEL(Ls);
    int nx, ny;
    gdbl *base = field_of(L, field, nx, ny);
    const gdbl *src = parity ? L.tmp : base;
    gdbl *dst = parity ? base : L.tmp;
    RANGE2D(1, nx - 1, 1, ny - 1);
    const int npx = L.npx;
    const int ij = j * npx + i;
    const real_t W = (i - 1 >= 1) ? src[ij - 1] : base[ij - 1];
    const real_t E = (i + 1 <= nx - 1) ? src[ij + 1] : base[ij + 1];
    const real_t S = (j - 1 >= 1) ? src[ij - npx] : base[ij - npx];
    const real_t N = (j + 1 <= ny - 1) ? src[ij + npx] : base[ij + npx];
    const real_t sum_nb = ((L.aw[ij] * W + L.ae[ij] * E) + L.as[ij] * S) + L.an[ij] * N;
    const real_t g = (sum_nb + L.su[ij]) / (L.ap[ij] + 1.0e-30);
    const double wj = 1.0;
    dst[ij] = (1.0 - wj) * src[ij] + wj * g;
}

/* after an odd number of ping-pong sweeps the result is in tmp */
__global__ void __launch_bounds__(BX * BY) k_copy_range(const DevLevel *Ls, int field)
{
    LEVEL(Ls);
    int nx, ny;
    gdbl *dst = field_of(L, field, nx, ny);
    RANGE2D(1, nx - 1, 1, ny - 1);
    dst[j * L.npx + i] = L.tmp[j * L.npx + i];
}

/* One whole red-black sweep (both colours, V6:1308-1349) per launch.
 * A 64 x RB_TY tile is staged in LDS with a 2-point halo read from `src`
 * (points outside the swept range from `base`, the field itself); the red
 * points of the tile and of a 1-point ring around it are updated first (the
 * ring values are recomputed by each neighbouring tile from the same old
 * values, so they are the same bits), then the black points of the tile, and
 * the tile is written to `dst`.  Reading `src` and writing `dst` (ping-pong)
 * keeps every tile's inputs the sweep's old values, exactly as the
 * reference's colour-by-colour in-place passes see them.
 * Each thread owns one column and RB_TY/BY rows of the tile and loads the
 * six coefficients of its points into registers together with the halo
 * staging, so the sweep pays one memory latency, not one per colour. */
#define RB_TY 16
#define RB_Q (RB_TY / BY)
#define RB_RING (2 * (BX + 2) + 2 * RB_TY)

__device__ __forceinline__ void rb_ring_pos(int k, int &r, int &c)
{
    /* ring of the tile in LDS coordinates (tile occupies rows/cols 2..) */
    if (k < BX + 2) { r = 1; c = 1 + k; }
    else if (k < 2 * (BX + 2)) { r = RB_TY + 2; c = 1 + (k - (BX + 2)); }
    else if (k < 2 * (BX + 2) + RB_TY) { r = 2 + (k - 2 * (BX + 2)); c = 1; }
    else { r = 2 + (k - 2 * (BX + 2) - RB_TY); c = BX + 2; }
}

__global__ void __launch_bounds__(BX * BY) k_rb_sweep(const DevLevel *Ls, int field, int parity)
{
    LEVEL(Ls);
    int nx, ny;
    const gdbl *base = field_of(L, field, nx, ny);
    const gdbl *src = parity ? L.tmp : base;
    gdbl *dst = parity ? (gdbl *)base : L.tmp;
    __shared__ real_t t[RB_TY + 4][BX + 4];
    const int npx = L.npx, npy = L.npy;
    const int i0 = 1 + T.x * BX, j0 = 1 + T.y * RB_TY;
    const int tid = threadIdx.y * BX + threadIdx.x;
    constexpr int NST = ((RB_TY + 4) * (BX + 4) + BX * BY - 1) / (BX * BY);

    /* 1. issue every global load: halo staging, own points, one ring point */
    real_t sv[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int k = tid + q * BX * BY;
        real_t val = 0.0;
        if (k < (RB_TY + 4) * (BX + 4)) {
            const int r = k / (BX + 4), c = k % (BX + 4);
            const int gi = i0 - 2 + c, gj = j0 - 2 + r;
            if (gi >= 0 && gi < npx && gj >= 0 && gj < npy) {
                const int g = gj * npx + gi;
                const bool in = (gi >= 1 && gi <= nx - 1 && gj >= 1 && gj <= ny - 1);
                val = in ? src[g] : base[g];
            }
        }
        sv[q] = val;
    }
    const int gi = i0 + threadIdx.x;
    Pt own[RB_Q];
    bool live[RB_Q];
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int gj = j0 + threadIdx.y + BY * q;
        live[q] = (gi <= nx - 1 && gj <= ny - 1);
        if (live[q]) {
            const int g = gj * npx + gi;
            own[q] = Pt{L.aw[g], L.ae[g], L.as[g], L.an[g], L.ap[g], L.su[g]};
        } else {
            own[q] = Pt{0.0, 0.0, 0.0, 0.0, 1.0, 0.0};
        }
    }
    int rr = 0, rc = 0;
    bool ring = false;
    Pt rp = Pt{0.0, 0.0, 0.0, 0.0, 1.0, 0.0};
    if (tid < RB_RING) {
        rb_ring_pos(tid, rr, rc);
        const int ri = i0 - 2 + rc, rj = j0 - 2 + rr;
        ring = (((ri + rj) & 1) == 0) && ri >= 1 && ri <= nx - 1 && rj >= 1 && rj <= ny - 1;
        if (ring) {
            const int g = rj * npx + ri;
            rp = Pt{L.aw[g], L.ae[g], L.as[g], L.an[g], L.ap[g], L.su[g]};
        }
    }
    /* 2. staging to LDS */
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int k = tid + q * BX * BY;
        if (k < (RB_TY + 4) * (BX + 4)) t[k / (BX + 4)][k % (BX + 4)] = sv[q];
    }
    __syncthreads();
    /* 3. red points: own and ring */
    real_t res[RB_Q];
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int r = 2 + threadIdx.y + BY * q, c = 2 + threadIdx.x;
        const int gj = j0 + threadIdx.y + BY * q;
        res[q] = t[r][c];
        if (live[q] && ((gi + gj) & 1) == 0) {
            const Pt &o = own[q];
            const real_t sum_nb = ((o.aw * t[r][c - 1] + o.ae * t[r][c + 1]) + o.as * t[r - 1][c]) +
                                  o.an * t[r + 1][c];
            res[q] = (sum_nb + o.su) / o.ap;
        }
    }
    real_t rv = 0.0;
    if (ring) {
        const real_t sum_nb = ((rp.aw * t[rr][rc - 1] + rp.ae * t[rr][rc + 1]) +
                               rp.as * t[rr - 1][rc]) + rp.an * t[rr + 1][rc];
        rv = (sum_nb + rp.su) / rp.ap;
    }
    /* red updates read only black neighbours, so red results can be stored
     * without a barrier in between */
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int gj = j0 + threadIdx.y + BY * q;
        if (live[q] && ((gi + gj) & 1) == 0) t[2 + threadIdx.y + BY * q][2 + threadIdx.x] = res[q];
    }
    if (ring) t[rr][rc] = rv;
    __syncthreads();
    /* 4. black points, then write the tile */
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int r = 2 + threadIdx.y + BY * q, c = 2 + threadIdx.x;
        const int gj = j0 + threadIdx.y + BY * q;
        if (!live[q]) continue;
        real_t val = res[q];
        if (((gi + gj) & 1) == 1) {
            const Pt &o = own[q];
            const real_t sum_nb = ((o.aw * t[r][c - 1] + o.ae * t[r][c + 1]) + o.as * t[r - 1][c]) +
                                  o.an * t[r + 1][c];
            val = (sum_nb + o.su) / o.ap;
        }
        dst[gj * npx + gi] = val;
    }
}

/* ---------------------------------------------------------------------- */
/* update, V6:1127-1178                                                     */
/* ---------------------------------------------------------------------- */

/* freshly updated p of every copy origin, taken before the update overwrites
 * anything (the reference copies after its update loop, in list order) */
__global__ void k_pstash(const DevLevel *Ls)
{
    LEVEL1D(Ls);
    const DevPhys &P = L.P;
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= L.npcopy) return;
    const int o = L.porig[s];
    const int oi = o % L.npx, oj = o / L.npx;
    const bool in = (oi >= 1 && oi <= L.npx - 2 && oj >= 1 && oj <= L.npy - 2);
    L.pstash[s] = in ? L.p[o] + P.relxp * (L.pp[o] - 0.0) : L.p[o];
}

/* PENDING >= 1: mixed precision, where the corrections of the u and v
 * solves are still in du32 / dv32 over their swept ranges (the value of u
 * there is u + du, formed exactly as the solve's own update would, and 0 on
 * penalised faces: pending_value); 2: the
 * sc solve ran in the same launch as u and v (which left u + du, v + dv in
 * uc, vc), and its correction in ds32 is
 * applied here (sc += d, as k_apply_mx would; the update does not read sc),
 * and the last p' launch left the updated p of every copy origin in tmp
 * (the value k_pstash would stash); 3: FP64, the u and v relaxations left
 * their results over their swept ranges in uc, vc (the fields hold the old
 * values there); 4: as 3, and sc's relaxation left its result in scn; 5: as
 * 4, and the last p' launch left the copy origins' updated p in pnew */
struct Cell4 {
    real_t u, v, p, sc;
};

/* the update of cell (i, j), stored; returns the new u, v, p and sc */
template <int PENDING>
__device__ __forceinline__ Cell4 update_cell(const DevLevelG &L, int i, int j)
{
    const DevPhys &P = L.P;
    const int nx = L.npx;
    const int ij = j * nx + i;
    const real_t pp = L.pp[ij];
    const uint8_t fl = L.flags[ij];
    real_t u, v;
    if (PENDING >= 2) {
        /* the merged relaxation formed u + du, v + dv in uc, vc (2); the FP64
         * u and v relaxations left their results there (3) */
        u = (i <= nx - 3 ? L.uc : L.u)[ij];
        v = (j <= L.npy - 3 ? L.vc : L.v)[ij];
    } else {
        u = L.u[ij];
        v = L.v[ij];
        if (PENDING) {
            const float du = L.du32[ij], dv = L.dv32[ij];
            u = i <= nx - 3 ? pending_value(u, du, L.pset && (fl & MGF_F_SPU), 0.0) : u;
            v = j <= L.npy - 3 ? pending_value(v, dv, L.pset && (fl & MGF_F_SPV), 0.0) : v;
        }
    }
    Cell4 c;
    c.u = u + (L.ax * (pp - L.pp[ij + 1])) / L.apu[ij];
    c.v = v + (L.ay * (pp - L.pp[ij + nx])) / L.apv[ij];
    c.sc = PENDING >= 4 ? L.scn[ij] : L.sc[ij];
    if (PENDING == 2) c.sc = pending_value(c.sc, L.ds32[ij], L.pset && (fl & MGF_F_SPSC), P.sc_obs);
    if (fl & MGF_F_PCOPY) {
        if (PENDING == 2 || PENDING == 5) {
            /* the origin's updated p, stashed by the last p' launch (in tmp,
             * or in pnew for FP64); an origin outside the updated range keeps
             * its p */
            const int o = L.porig[L.pslot[ij]];
            const int oi = o % nx, oj = o / nx;
            const bool in = (oi >= 1 && oi <= nx - 2 && oj >= 1 && oj <= L.npy - 2);
            c.p = in ? (PENDING == 2 ? L.tmp : L.pnew)[o] : L.p[o];
        } else {
            c.p = L.pstash[L.pslot[ij]];
        }
    } else {
        c.p = L.p[ij] + P.relxp * (pp - 0.0);
    }
    L.u[ij] = c.u;
    L.v[ij] = c.v;
    L.p[ij] = c.p;
    if (PENDING == 2 || PENDING >= 4) L.sc[ij] = c.sc;
    return c;
}

template <int PENDING>
__global__ void __launch_bounds__(BX * BY) k_update(const DevLevel *Ls)
{
    LEVEL(Ls);
    RANGE2D(1, L.npx - 2, 1, L.npy - 2);
    (void)update_cell<PENDING>(L, i, j);
}

/* The update of a level followed at once by restv onto its next coarser
 * level (nothing runs between them in a down-leg relaxation without
 * periodic boundaries).  Threads take fine cells as k_update does; a 64 x 4
 * workgroup holds whole 2 x 2 blocks (its rows start at an odd j, its
 * columns at an odd i), so after an exchange through LDS the thread of each
 * block's (2ic, 2jc) cell restricts the block with restv's expressions
 * (V6:1370-1388).  The fine interior is exactly the blocks of the coarse
 * interior. */
template <int PENDING>
__global__ void __launch_bounds__(BX * BY) k_update_restv(const DevLevel *Fs, const DevLevel *Cs)
{
    PAIR(Fs, Cs);
    __shared__ real_t X[4][BY][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = 1 + (int)(T.x * BX + tx), j = 1 + (int)(T.y * BY + ty);
    const bool in = i <= F.npx - 2 && j <= F.npy - 2;
    if (T.y * BY + 1 > F.npy - 2) return;           /* padding rows: the whole workgroup */
    Cell4 c{};
    if (in) c = update_cell<PENDING>(F, i, j);
    X[0][ty][tx] = c.u;
    X[1][ty][tx] = c.v;
    X[2][ty][tx] = c.p;
    X[3][ty][tx] = c.sc;
    __syncthreads();
    if (!in || !(tx & 1) || !(ty & 1)) return;
    /* this thread's cell is (2ic, 2jc): ij; its block's ij-1, ij-nxf, ij-nxf-1 */
    const int ic = i / 2, jc = j / 2;
    const int ijc = jc * C.npx + ic;
    const real_t ru = (c.u * F.ax + X[0][ty - 1][tx] * F.ax) / (C.ax + 1.0e-30);
    const real_t rv = (c.v * F.ay + X[1][ty][tx - 1] * F.ay) / (C.ay + 1.0e-30);
    const real_t rp = 0.25 * (((c.p + X[2][ty][tx - 1]) + X[2][ty - 1][tx]) + X[2][ty - 1][tx - 1]);
    const real_t rs =
        0.25 * (((c.sc + X[3][ty][tx - 1]) + X[3][ty - 1][tx]) + X[3][ty - 1][tx - 1]);
    C.u[ijc] = ru;
    C.v[ijc] = rv;
    C.p[ijc] = rp;
    C.sc[ijc] = rs;
    /* R x_f, kept for the V-cycle's prolongation (k_prol_s) */
    C.cu[ijc] = ru;
    C.cv[ijc] = rv;
    C.cp[ijc] = rp;
    C.csc[ijc] = rs;
}

/* ---------------------------------------------------------------------- */
/* periodic / outflow boundaries, V6:1179-1228                              */
/* ---------------------------------------------------------------------- */
__global__ void k_pbound_x(const DevLevel *Ls)
{
    LEVEL1D(Ls);
    if (L.P.ibc[0] != 3) return;
    const int j = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j > L.npy - 2) return;
    const int ij = j * L.npx, ij1 = ij + L.npx - 3;
    L.u[ij] = L.u[ij1];
    L.u[ij1 + 1] = L.u[ij + 1];
    L.v[ij] = L.v[ij1];
    L.v[ij1 + 1] = L.v[ij + 1];
}

/* the y-periodic branch indexes ij + ny - 3 (V6:1203), so its iterations
 * depend on each other; one thread runs them in the reference's order */
__global__ void k_pbound_y(const DevLevel *Ls)
{
    LEVEL1D(Ls);
    if (L.P.ibc[2] != 3) return;
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int nx = L.npx, ny = L.npy;
    for (int i = 1; i < nx - 1; ++i) {
        const int ij = i, ij1 = ij + ny - 3;
        L.u[ij] = L.u[ij1];
        L.u[ij1 + nx] = L.u[ij + nx];
        L.v[ij] = L.v[ij1];
        L.v[ij1 + nx] = L.v[ij + nx];
    }
}

__global__ void k_outflow(const DevLevel *Ls)
{
    LEVEL1D(Ls);
    if (L.P.ibc[1] != 4) return;
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j > L.npy - 1) return;
    const int ij = j * L.npx + L.npx - 1;
    L.u[ij - 1] = L.u[ij - 2];
    L.v[ij] = L.v[ij - 1];
    L.p[ij] = L.p[ij - 1];
    L.sc[ij] = L.sc[ij - 1];
}

/* ---------------------------------------------------------------------- */
/* restriction                                                              */
/* ---------------------------------------------------------------------- */

/* restv, V6:1354-1389: coarse [1,nyc-2]x[1,nxc-2].  The restricted values
 * R x_f are also kept in cu, cv, cp, csc: the fine level does not change
 * again before the V-cycle prolongs onto it, and the prolongation's R x_f
 * (V6:1762-1995) is the same expression on the same values, so the stored
 * copy is that R x_f bit for bit (k_prol_s reads it instead of
 * recomputing it). */
__global__ void __launch_bounds__(BX * BY) k_restv(const DevLevel *Fs, const DevLevel *Cs)
{
    PAIR(Fs, Cs);
    const int ic = 1 + (int)(T.x * BX + threadIdx.x);
    const int jc = 1 + (int)(T.y * BY + threadIdx.y);
    if (ic > C.npx - 2 || jc > C.npy - 2) return;
    const int nxf = F.npx;
    const int ijc = jc * C.npx + ic;
    const int ij = (2 * jc) * nxf + 2 * ic, ijw = ij - 1, ijs = ij - nxf;
    const real_t ru = (F.u[ij] * F.ax + F.u[ijs] * F.ax) / (C.ax + 1.0e-30);
    const real_t rv = (F.v[ij] * F.ay + F.v[ijw] * F.ay) / (C.ay + 1.0e-30);
    const real_t rp = 0.25 * (((F.p[ij] + F.p[ijw]) + F.p[ijs]) + F.p[ijs - 1]);
    const real_t rs = 0.25 * (((F.sc[ij] + F.sc[ijw]) + F.sc[ijs]) + F.sc[ijs - 1]);
    C.u[ijc] = ru;
    C.v[ijc] = rv;
    C.p[ijc] = rp;
    C.sc[ijc] = rs;
    C.cu[ijc] = ru;
    C.cv[ijc] = rv;
    C.cp[ijc] = rp;
    C.csc[ijc] = rs;
}

/* where the residuals take u and v from, relative to their point (di, dj):
 * the level's arrays, or a workgroup's LDS copy of them (row stride w) */
struct StG {
    const gdbl *u, *v;
    int ij, nx;
    __device__ real_t U(int di, int dj) const { return u[ij + dj * nx + di]; }
    __device__ real_t V(int di, int dj) const { return v[ij + dj * nx + di]; }
};
struct StL {
    const real_t *u, *v;
    int o, w;
    __device__ real_t U(int di, int dj) const { return u[o + dj * w + di]; }
    __device__ real_t V(int di, int dj) const { return v[o + dj * w + di]; }
};
__device__ __forceinline__ StG st_g(const DevLevelG &L, int i, int j)
{
    return StG{L.u, L.v, j * L.npx + i, L.npx};
}

/* momentum residual at a u face with un-relaxed, un-penalised coefficients
 * (restru, V6:1496-1510 fine with resu, 1532-1545 coarse without); the
 * fluxes are fluxes' expression on the current u, v (coefu_at with
 * FluxInl) */
template <class S>
__device__ __forceinline__ real_t rs_u(const DevLevelG &L, int i, int j, const S &s, real_t body,
                                       bool fine, bool mask = false)
{
    const int nx = L.npx, ij = j * nx + i;
    const real_t ax = L.ax, ay = L.ay;
    Coef c = coefu_vals(L, j, (0.5 * s.U(-1, 0)) * ax, (0.5 * s.U(0, 0)) * ax,
                        (0.5 * s.U(1, 0)) * ax, (0.5 * s.V(0, -1)) * ay, (0.5 * s.V(1, -1)) * ay,
                        (0.5 * s.V(0, 0)) * ay, (0.5 * s.V(1, 0)) * ay);
    real_t r = ((((c.aw * s.U(-1, 0) + c.ae * s.U(1, 0)) + c.as * s.U(0, -1)) +
                 c.an * s.U(0, 1)) - c.ap * s.U(0, 0)) + L.ax * (L.p[ij] - L.p[ij + 1]);
    if (fine) r = r + L.resu[ij];
    r = r + body;
    if (fine && L.ax == 0.0) r = 0.0;
    if (mask && (L.flags[ij] & MGF_F_SPU)) r = 0.0;
    return r;
}

template <class S>
__device__ __forceinline__ real_t rs_v(const DevLevelG &L, int i, int j, const S &s, real_t body,
                                       bool fine, bool mask = false)
{
    const int nx = L.npx, ij = j * nx + i;
    const real_t ax = L.ax, ay = L.ay;
    Coef c = coefv_vals(L, i, (0.5 * s.U(-1, 0)) * ax, (0.5 * s.U(-1, 1)) * ax,
                        (0.5 * s.U(0, 0)) * ax, (0.5 * s.U(0, 1)) * ax, (0.5 * s.V(0, -1)) * ay,
                        (0.5 * s.V(0, 0)) * ay, (0.5 * s.V(0, 1)) * ay);
    real_t r = ((((c.aw * s.V(-1, 0) + c.ae * s.V(1, 0)) + c.as * s.V(0, -1)) +
                 c.an * s.V(0, 1)) - c.ap * s.V(0, 0)) + L.ay * (L.p[ij] - L.p[ij + nx]);
    if (fine) r = r + L.resv[ij];
    r = r + body;
    if (fine && L.ay == 0.0) r = 0.0;
    if (mask && (L.flags[ij] & MGF_F_SPV)) r = 0.0;
    return r;
}

template <class S>
__device__ __forceinline__ real_t rs_sc(const DevLevelG &L, int i, int j, const S &st, bool fine,
                                        bool mask = false)
{
    const int nx = L.npx, ij = j * nx + i;
    const uint8_t fl = L.flags[ij];
    Coef c = coefsc_vals(L, ij, fl, (0.5 * st.U(-1, 0)) * L.ax, (0.5 * st.U(0, 0)) * L.ax,
                         (0.5 * st.V(0, -1)) * L.ay, (0.5 * st.V(0, 0)) * L.ay);
    const gdbl *s = L.sc;
    real_t r = (((c.aw * s[ij - 1] + c.ae * s[ij + 1]) + c.as * s[ij - nx]) + c.an * s[ij + nx]) -
               c.ap * s[ij];
    if (fine) r = r + L.ressc[ij];
    if (mask && (fl & MGF_F_SPSC)) r = 0.0;
    return r;
}

/* FAS tau terms resu/resv/ressc on the coarse level (V6:1547-1563,
 * 1632-1647, 1723-1739).  A workgroup covers 64 x 4 coarse points: it first
 * evaluates every fine residual they need once, at unit stride, into LDS
 * (the reference's rs scratch array, V6:1496-1530, restricted to the
 * workgroup's footprint), then each coarse point combines its fine
 * residuals in the reference's order and subtracts its coarse residual.
 * The fine u and v the residual stencils reach (fine columns fi0-1 ..
 * fi0+2BX+1, rows fj0-1 .. fj0+2BY+1) are first copied to LDS at unit
 * stride, so each is loaded once per workgroup instead of by every stencil
 * that reaches it (FP64 only: dual numbers would not fit). */
#define TAU_UW (2 * BX + 1)
#define TAU_W (2 * BX)
#define TAU_H (2 * BY)
/* 1: stage u, v in LDS (3 workgroups per CU); 0: read them from the level
 * (60 VGPRs either way, 4 workgroups per CU): staging measured 1.6% faster
 * on the batch */
#ifndef MGF_TAU_STAGE
#define MGF_TAU_STAGE 1
#endif
#define TAU_SW (2 * BX + 3)
#define TAU_SH (2 * BY + 3)
/* threads of a k_tau workgroup: more than the 64 x 4 coarse points, so each
 * thread evaluates fewer fine residuals one after another */
#define TAU_NT (2 * BX * BY)
__global__ void __launch_bounds__(TAU_NT) k_tau(const DevLevel *Fs, const DevLevel *Cs)
{
    PAIR(Fs, Cs);
    constexpr bool STAGE = MGF_TAU_STAGE && sizeof(real_t) == sizeof(double);
    constexpr int NS = STAGE ? TAU_SH * TAU_SW : 1;
    __shared__ real_t U[TAU_H][TAU_UW];
    __shared__ real_t V[TAU_H + 1][TAU_W];
    __shared__ real_t Q[TAU_H][TAU_W];
    __shared__ real_t Su[NS], Sv[NS];
    const int nxc = C.npx, nyc = C.npy;
    const int ic0 = 1 + T.x * BX, jc0 = 1 + T.y * BY;
    if (ic0 > nxc - 2 || jc0 > nyc - 2) return;          /* padding tiles */
    const int fi0 = 2 * ic0 - 1, fj0 = 2 * jc0 - 1;
    const int tid = threadIdx.y * BX + threadIdx.x;
    /* the coarse residuals first: their loads overlap the fine phase's
     * instead of waiting behind its barrier */
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int ic = ic0 + tx, jc = jc0 + ty;
    const bool own = ty < BY && ic <= nxc - 2 && jc <= nyc - 2;
    real_t rcu = 0.0, rcv = 0.0, rcs = 0.0;
    if (own) {
        if (ic <= nxc - 3) rcu = rs_u(C, ic, jc, st_g(C, ic, jc), C.bodyx, false);
        if (jc <= nyc - 3) rcv = rs_v(C, ic, jc, st_g(C, ic, jc), C.bodyy, false);
        rcs = rs_sc(C, ic, jc, st_g(C, ic, jc), false);
    }
    if (STAGE) {
        const int nxf = F.npx, nyf = F.npy;
        for (int e = tid; e < TAU_SH * TAU_SW; e += TAU_NT) {
            const int r = e / TAU_SW, c = e - r * TAU_SW;
            const int gi = fi0 - 1 + c, gj = fj0 - 1 + r;
            const bool ok = gi <= nxf - 1 && gj <= nyf - 1;
            Su[e] = ok ? F.u[gj * nxf + gi] : 0.0;
            Sv[e] = ok ? F.v[gj * nxf + gi] : 0.0;
        }
        __syncthreads();
    }
    /* the fine point (fi0 + c, fj0 + r)'s access to u and v */
    auto fst = [&](int c, int r) {
        if constexpr (STAGE)
            return StL{Su, Sv, (r + 1) * TAU_SW + (c + 1), TAU_SW};
        else
            return st_g(F, fi0 + c, fj0 + r);
    };
    /* onto a level below the obstacle level the penalised fine residuals
     * are dropped (their nonzero value is balanced by the 1e30 penalty,
     * which the residual leaves out); never the case when ig_obs == ngrid */
    const bool m = C.ext != 0;
    /* last coarse point of the workgroup for each equation, and the fine
     * columns / rows its residual stencils reach (relative to fi0, fj0) */
    const int icu = min(ic0 + BX - 1, nxc - 3), icc = min(ic0 + BX - 1, nxc - 2);
    const int jcc = min(jc0 + BY - 1, nyc - 2), jcv = min(jc0 + BY - 1, nyc - 3);
    const int cu = icu >= ic0 ? 2 * (icu - ic0) + 2 : -1, ru = 2 * (jcc - jc0) + 1;
    const int cv = 2 * (icc - ic0) + 1, rv = jcv >= jc0 ? 2 * (jcv - jc0) + 2 : -1;
    const int cq = 2 * (icc - ic0) + 1, rq = 2 * (jcc - jc0) + 1;
    for (int e = tid; e < TAU_H * TAU_UW; e += TAU_NT) {
        const int r = e / TAU_UW, c = e - r * TAU_UW;
        U[r][c] = (c <= cu && r <= ru) ? rs_u(F, fi0 + c, fj0 + r, fst(c, r), F.bodyx, true, m)
                                       : 0.0;
    }
    for (int e = tid; e < (TAU_H + 1) * TAU_W; e += TAU_NT) {
        const int r = e / TAU_W, c = e - r * TAU_W;
        V[r][c] = (c <= cv && r <= rv) ? rs_v(F, fi0 + c, fj0 + r, fst(c, r), F.bodyy, true, m)
                                       : 0.0;
    }
    for (int e = tid; e < TAU_H * TAU_W; e += TAU_NT) {
        const int r = e / TAU_W, c = e - r * TAU_W;
        Q[r][c] = (c <= cq && r <= rq) ? rs_sc(F, fi0 + c, fj0 + r, fst(c, r), true, m) : 0.0;
    }
    __syncthreads();
    if (!own) return;
    const int ijc = jc * nxc + ic;
    const int c1 = 2 * tx + 1, r1 = 2 * ty + 1;     /* fine (2ic, 2jc) in the tables */
    if (ic <= nxc - 3) {
        real_t a = U[r1][c1], b = U[r1 - 1][c1];
        real_t w0 = U[r1][c1 - 1], w1 = U[r1 - 1][c1 - 1];
        real_t e0 = U[r1][c1 + 1], e1 = U[r1 - 1][c1 + 1];
        C.resu[ijc] = ((a + b) + 0.5 * (((w0 + w1) + e0) + e1)) - rcu;
    }
    if (jc <= nyc - 3) {
        real_t a = V[r1][c1], b = V[r1][c1 - 1];
        real_t s0 = V[r1 - 1][c1], s1 = V[r1 - 1][c1 - 1];
        real_t n0 = V[r1 + 1][c1], n1 = V[r1 + 1][c1 - 1];
        C.resv[ijc] = ((a + b) + 0.5 * (((s0 + s1) + n0) + n1)) - rcv;
    }
    {
        real_t a = Q[r1][c1], b = Q[r1 - 1][c1];
        real_t c2 = Q[r1][c1 - 1], d = Q[r1 - 1][c1 - 1];
        C.ressc[ijc] = (((a + b) + c2) + d) - rcs;
    }
}

/* ---------------------------------------------------------------------- */
/* prolongation, V6:1745-2028, in gather form                               */
/* ---------------------------------------------------------------------- */

/* coarse corrections c = x_c - R x_f for u, v, p, sc (prolu 1762-1779,
 * prolv 1838-1855, prolp 1911-1928, prolsc 1979-1995).  All four read only
 * pre-prolongation fine data, so one pass serves them all.  The entries the
 * reference leaves at zero (ghosts; cu of the last face column, cv of the
 * last face row) are never read by the prolongation. */
__global__ void __launch_bounds__(BX * BY) k_prol_c(const DevLevel *Fs, const DevLevel *Cs)
{
    PAIR(Fs, Cs);
    const int ic = 1 + (int)(T.x * BX + threadIdx.x);
    const int jc = 1 + (int)(T.y * BY + threadIdx.y);
    const int nxc = C.npx, nyc = C.npy, nxf = F.npx;
    if (ic > nxc - 2 || jc > nyc - 2) return;
    const int ijc = jc * nxc + ic;
    const int ij = (2 * jc) * nxf + 2 * ic, ijs = ij - nxf;
    /* corrections from penalised points of a level below the obstacle
     * level carry no information about the finer level and are dropped */
    const uint8_t cfl = C.ext ? C.flags[ijc] : (uint8_t)0;
    if (ic <= nxc - 3) {
        real_t Ru = (F.u[ij] * F.ax + F.u[ijs] * F.ax) / (C.ax + 1.0E-30);
        real_t cu = C.u[ijc] - Ru;
        if (cfl & MGF_F_SPU) cu = 0.0;
        C.cu[ijc] = cu;
    }
    if (jc <= nyc - 3) {
        real_t Rv = (F.v[ij] * F.ay + F.v[ij - 1] * F.ay) / (C.ay + 1.0E-30);
        real_t cv = C.v[ijc] - Rv;
        if (cfl & MGF_F_SPV) cv = 0.0;
        C.cv[ijc] = cv;
    }
    {
        real_t Rp = 0.25 * (((F.p[ij] + F.p[ij - 1]) + F.p[ijs]) + F.p[ijs - 1]);
        real_t cp = C.p[ijc] - Rp;
        if (F.flags[ij] & MGF_F_SPU) cp = 0.0;
        if (cfl & MGF_F_SPSC) cp = 0.0;
        C.cp[ijc] = cp;
        real_t Rs = 0.25 * (((F.sc[ij] + F.sc[ij - 1]) + F.sc[ijs]) + F.sc[ijs - 1]);
        real_t cs = C.sc[ijc] - Rs;
        if (cfl & MGF_F_SPSC) cs = 0.0;
        C.csc[ijc] = cs;
    }
}

/* where the application takes the coarse corrections from: the level's
 * cu, cv, cp, csc as k_prol_c left them, with the fine face flags read from
 * the fine level ... */
struct CorrG {
    const DevLevelG &F, &C;
    __device__ real_t u(int ic, int jc) const { return C.cu[jc * C.npx + ic]; }
    __device__ real_t v(int ic, int jc) const { return C.cv[jc * C.npx + ic]; }
    __device__ real_t p(int ic, int jc) const { return C.cp[jc * C.npx + ic]; }
    __device__ real_t s(int ic, int jc) const { return C.csc[jc * C.npx + ic]; }
    /* flags of the fine point (2ic, 2jc) */
    __device__ uint8_t ff(int ic, int jc) const { return F.flags[(2 * jc) * F.npx + 2 * ic]; }
};

/* ... or a workgroup's LDS tile of them, coarse points (ic0 - 1 .. ic0 + BX,
 * jc0 - 1 .. jc0 + BY), formed by k_prol_s */
#define PROL_W (BX + 2)
#define PROL_H (BY + 2)
struct CorrL {
    const real_t (*cu)[PROL_W], (*cv)[PROL_W], (*cp)[PROL_W], (*cs)[PROL_W];
    const uint8_t (*fl)[PROL_W];
    int ic0, jc0;
    __device__ real_t u(int ic, int jc) const { return cu[jc - jc0 + 1][ic - ic0 + 1]; }
    __device__ real_t v(int ic, int jc) const { return cv[jc - jc0 + 1][ic - ic0 + 1]; }
    __device__ real_t p(int ic, int jc) const { return cp[jc - jc0 + 1][ic - ic0 + 1]; }
    __device__ real_t s(int ic, int jc) const { return cs[jc - jc0 + 1][ic - ic0 + 1]; }
    __device__ uint8_t ff(int ic, int jc) const { return fl[jc - jc0 + 1][ic - ic0 + 1]; }
};

/* face correction of prolu/prolv for coarse face k along the face direction,
 * mixing in the neighbour across it (3/4, 1/4), zero where the fine face at
 * (2ic, 2jc) is penalised (V6:1794-1798, 1870-1874) */
template <class A>
__device__ __forceinline__ real_t corr_u(const A &a, int ic, int jc, bool upper)
{
    const double f1 = 3. / 4., f2 = 1. / 4.;
    real_t c = upper ? f1 * a.u(ic, jc) + f2 * a.u(ic, jc + 1) : f1 * a.u(ic, jc) + f2 * a.u(ic, jc - 1);
    if (a.ff(ic, jc) & MGF_F_SPU) c = 0.0;
    return c;
}

template <class A>
__device__ __forceinline__ real_t corr_v(const A &a, int ic, int jc, bool right)
{
    const double f1 = 3. / 4., f2 = 1. / 4.;
    real_t c = right ? f1 * a.v(ic, jc) + f2 * a.v(ic + 1, jc) : f1 * a.v(ic, jc) + f2 * a.v(ic - 1, jc);
    if (a.ff(ic, jc) & MGF_F_SPV) c = 0.0;
    return c;
}

/* bilinear 9/16, 3/16, 3/16, 1/16 correction (V6:1945-1955) of the fine
 * cell on side (di, dj) of a coarse cell, from the coarse correction's 3 x 3
 * neighbourhood n[dj+1][di+1] */
__device__ __forceinline__ real_t corr_cell_n(const real_t n[3][3], int di, int dj)
{
    const double f1 = 9. / 16., f2 = 3. / 16., f3 = 3. / 16., f4 = 1. / 16.;
    return ((f1 * n[1][1] + f2 * n[1][1 + di]) + f3 * n[1 + dj][1]) + f4 * n[1 + dj][1 + di];
}

/* Every fine point gathers its corrections; a thread takes the four fine
 * cells (and their faces) of coarse cell (ic, jc), so the coarse values are
 * read once for four points.  Odd faces, which the reference updates twice
 * (atomics in its GPU build), take the two halves in its CPU loop order
 * (smaller coarse index first). */
struct Fine16 {
    real_t u[4], v[4], p[4], s[4];
};

/* the fine values the thread of coarse cell (ic, jc) updates (each fine
 * point belongs to exactly one coarse cell, so they can be read early) */
__device__ __forceinline__ Fine16 prol_load(const DevLevelG &F, int ic, int jc)
{
    const int nxf = F.npx;
    const int i1 = 2 * ic - 1, i2 = 2 * ic;
    const int r1 = (2 * jc - 1) * nxf, r2 = (2 * jc) * nxf;
    Fine16 x;
    x.u[0] = F.u[r2 + i1]; x.u[1] = F.u[r1 + i1]; x.u[2] = F.u[r2 + i2]; x.u[3] = F.u[r1 + i2];
    x.v[0] = F.v[r1 + i2]; x.v[1] = F.v[r1 + i1]; x.v[2] = F.v[r2 + i2]; x.v[3] = F.v[r2 + i1];
    x.p[0] = F.p[r1 + i1]; x.p[1] = F.p[r1 + i2]; x.p[2] = F.p[r2 + i1]; x.p[3] = F.p[r2 + i2];
    x.s[0] = F.sc[r1 + i1]; x.s[1] = F.sc[r1 + i2]; x.s[2] = F.sc[r2 + i1]; x.s[3] = F.sc[r2 + i2];
    return x;
}

template <class A>
__device__ __forceinline__ void prol_apply_at(const DevLevelG &F, const DevLevelG &C, int ic,
                                              int jc, const A &a, const Fine16 &x)
{
    const int nxf = F.npx, nxc = C.npx, nyc = C.npy;
    const int i1 = 2 * ic - 1, i2 = 2 * ic;
    const int r1 = (2 * jc - 1) * nxf, r2 = (2 * jc) * nxf;
    /* u faces i1 (odd: halves from coarse faces ic-1 and ic) and i2 = 2ic
     * (coarse face ic, if it is an interior face); rows 2jc (upper) and
     * 2jc-1 */
    {
        const bool hr = ic <= nxc - 3, hl = ic >= 2;
        real_t crU = 0.0, crL = 0.0, clU = 0.0, clL = 0.0;
        if (hr) { crU = corr_u(a, ic, jc, true); crL = corr_u(a, ic, jc, false); }
        if (hl) { clU = corr_u(a, ic - 1, jc, true); clL = corr_u(a, ic - 1, jc, false); }
        real_t u = x.u[0];
        if (hl) u = u + 0.5 * clU;
        if (hr) u = u + 0.5 * crU;
        F.u[r2 + i1] = u;
        u = x.u[1];
        if (hl) u = u + 0.5 * clL;
        if (hr) u = u + 0.5 * crL;
        F.u[r1 + i1] = u;
        if (hr) {
            F.u[r2 + i2] = x.u[2] + crU;
            F.u[r1 + i2] = x.u[3] + crL;
        }
    }
    /* v faces: row 2jc-1 (odd: halves from coarse faces jc-1 and jc) and row
     * 2jc (coarse face jc, if interior); columns 2ic (right) and 2ic-1 */
    {
        const bool hu = jc <= nyc - 3, hd = jc >= 2;
        real_t cuR = 0.0, cuL = 0.0, cdR = 0.0, cdL = 0.0;
        if (hu) { cuR = corr_v(a, ic, jc, true); cuL = corr_v(a, ic, jc, false); }
        if (hd) { cdR = corr_v(a, ic, jc - 1, true); cdL = corr_v(a, ic, jc - 1, false); }
        real_t v = x.v[0];
        if (hd) v = v + 0.5 * cdR;
        if (hu) v = v + 0.5 * cuR;
        F.v[r1 + i2] = v;
        v = x.v[1];
        if (hd) v = v + 0.5 * cdL;
        if (hu) v = v + 0.5 * cuL;
        F.v[r1 + i1] = v;
        if (hu) {
            F.v[r2 + i2] = x.v[2] + cuR;
            F.v[r2 + i1] = x.v[3] + cuL;
        }
    }
    /* cells */
    {
        real_t np[3][3], ns[3][3];
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
            for (int di = -1; di <= 1; ++di) {
                np[dj + 1][di + 1] = a.p(ic + di, jc + dj);
                ns[dj + 1][di + 1] = a.s(ic + di, jc + dj);
            }
        F.p[r1 + i1] = x.p[0] + corr_cell_n(np, -1, -1);
        F.p[r1 + i2] = x.p[1] + corr_cell_n(np, 1, -1);
        F.p[r2 + i1] = x.p[2] + corr_cell_n(np, -1, 1);
        F.p[r2 + i2] = x.p[3] + corr_cell_n(np, 1, 1);
        F.sc[r1 + i1] = x.s[0] + corr_cell_n(ns, -1, -1);
        F.sc[r1 + i2] = x.s[1] + corr_cell_n(ns, 1, -1);
        F.sc[r2 + i1] = x.s[2] + corr_cell_n(ns, -1, 1);
        F.sc[r2 + i2] = x.s[3] + corr_cell_n(ns, 1, 1);
    }
}

__global__ void __launch_bounds__(BX * BY) k_prol_apply(const DevLevel *Fs, const DevLevel *Cs)
{
    PAIR(Fs, Cs);
    const int ic = 1 + (int)(T.x * BX + threadIdx.x);
    const int jc = 1 + (int)(T.y * BY + threadIdx.y);
    if (ic > C.npx - 2 || jc > C.npy - 2) return;
    prol_apply_at(F, C, ic, jc, CorrG{F, C}, prol_load(F, ic, jc));
}

/* The V-cycle's prolongation in one launch.  cu, cv, cp, csc hold R x_f as
 * the restriction stored it (k_restv, k_update_restv), which is the R x_f
 * the reference recomputes here, so c = x_c - R x_f is formed per
 * workgroup, with k_prol_c's masks and zero where k_prol_c leaves zero, for
 * the workgroup's coarse points and a one-point ring around them, in LDS;
 * then applied as by k_prol_apply.  Not for a prolongation whose fine level
 * changed after the restriction (the step-level call uses k_prol_c). */
__global__ void __launch_bounds__(BX * BY) k_prol_s(const DevLevel *Fs, const DevLevel *Cs)
{
    PAIR(Fs, Cs);
    __shared__ real_t Su[PROL_H][PROL_W], Sv[PROL_H][PROL_W];
    __shared__ real_t Sp[PROL_H][PROL_W], Ss[PROL_H][PROL_W];
    __shared__ uint8_t Sf[PROL_H][PROL_W];
    const int nxc = C.npx, nyc = C.npy, nxf = F.npx;
    const int ic0 = 1 + T.x * BX, jc0 = 1 + T.y * BY;
    if (ic0 > nxc - 2 || jc0 > nyc - 2) return;          /* padding tiles */
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int ic = ic0 + (int)threadIdx.x, jc = jc0 + (int)threadIdx.y;
    const bool own = ic <= nxc - 2 && jc <= nyc - 2;
    /* the thread's fine values first, in flight across the barrier */
    Fine16 x{};
    if (own) x = prol_load(F, ic, jc);
    for (int e = tid; e < PROL_H * PROL_W; e += BX * BY) {
        const int r = e / PROL_W, q = e - r * PROL_W;
        const int ic = ic0 - 1 + q, jc = jc0 - 1 + r;
        real_t cu = 0.0, cv = 0.0, cp = 0.0, cs = 0.0;
        uint8_t ffl = 0;
        if (ic >= 1 && ic <= nxc - 2 && jc >= 1 && jc <= nyc - 2) {
            const int ijc = jc * nxc + ic;
            const uint8_t cfl = C.ext ? C.flags[ijc] : (uint8_t)0;
            ffl = F.flags[(2 * jc) * nxf + 2 * ic];
            if (ic <= nxc - 3) {
                cu = C.u[ijc] - C.cu[ijc];
                if (cfl & MGF_F_SPU) cu = 0.0;
            }
            if (jc <= nyc - 3) {
                cv = C.v[ijc] - C.cv[ijc];
                if (cfl & MGF_F_SPV) cv = 0.0;
            }
            cp = C.p[ijc] - C.cp[ijc];
            if (ffl & MGF_F_SPU) cp = 0.0;
            if (cfl & MGF_F_SPSC) cp = 0.0;
            cs = C.sc[ijc] - C.csc[ijc];
            if (cfl & MGF_F_SPSC) cs = 0.0;
        }
        Su[r][q] = cu;
        Sv[r][q] = cv;
        Sp[r][q] = cp;
        Ss[r][q] = cs;
        Sf[r][q] = ffl;
    }
    __syncthreads();
    if (!own) return;
    prol_apply_at(F, C, ic, jc, CorrL{Su, Sv, Sp, Ss, Sf, ic0, jc0}, x);
}

/* ---------------------------------------------------------------------- */
/* end of a V-cycle: stopping test of V6:247-248                            */
/* ---------------------------------------------------------------------- */
__global__ void k_end_cycle(DevCtl *ctls, const int *members, int n)
{
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    DevCtl *ctl = &ctls[members[k]];
    if (ctl->done) return;
    const double h = ctl->hist[ctl->cyc];
    if (!(h == h) || h == __builtin_huge_val() || h == -__builtin_huge_val()) ctl->nonfinite = 1;
    if (h < ctl->tol) ctl->done = 1;
    ctl->cyc = ctl->cyc + 1;
    if (ctl->cyc >= ctl->maxcyc) ctl->done = 1;
}

__global__ void k_fill(real_t *a, double v, size_t n)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}

/* ====================================================================== */
/* launch wrappers: one launch per group, blockIdx.z = member of the group */
/* ====================================================================== */

int mgf_partial_blocks(int npx, int npy)
{
    dim3 g = grid2d(npx - 2, npy - 2);
    return (int)(g.x * g.y);
}


void mgf_launch_fluxes(const DevGroup &G, hipStream_t st)
{
    hipLaunchKernelGGL(k_fluxes, tile_grid(G.npx - 1, G.npy - 1, G.count), dim3(BX, BY), 0, st, G.d);
}

void mgf_launch_asm_u(const DevGroup &G, hipStream_t st)
{
    if (G.npx - 3 < 1) return;
    hipLaunchKernelGGL(k_asm_u, tile_grid(G.npx - 3, G.npy - 2, G.count), dim3(BX, BY), 0, st, G.d);
}

void mgf_launch_asm_v(const DevGroup &G, hipStream_t st)
{
    if (G.npy - 3 < 1) return;
    hipLaunchKernelGGL(k_asm_v, tile_grid(G.npx - 2, G.npy - 3, G.count), dim3(BX, BY), 0, st, G.d);
}

void mgf_launch_reduce_err(const DevGroup &G, int record, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_err, dim3(1, 1, (unsigned)G.count), dim3(BX, BY), 0, st, G.d,
                       record);
}

void mgf_launch_asm_p(const DevGroup &G, int err, hipStream_t st, int pend)
{
    if (pend)
        hipLaunchKernelGGL(k_asm_p<true>, tile_grid(G.npx - 2, G.npy - 2, G.count), dim3(BX, BY),
                           0, st, G.d);
    else
        hipLaunchKernelGGL(k_asm_p<false>, tile_grid(G.npx - 2, G.npy - 2, G.count), dim3(BX, BY),
                           0, st, G.d);
    if (err) mgf_launch_reduce_err(G, err == 2, st);
}

void mgf_launch_asm_sc(const DevGroup &G, hipStream_t st)
{
    hipLaunchKernelGGL(k_asm_sc, tile_grid(G.npx - 2, G.npy - 2, G.count), dim3(BX, BY), 0, st, G.d);
}

/* Lexicographic Gauss-Seidel: the labelled variant of isolve 2 that this
 * library adds (the reference has only red-black, V6:1297-1351): nsw
 * in-place passes with j outer and i inner, each point updated with
 * solvergs's expression.  Run as anti-diagonal wavefronts, one workgroup per
 * problem: thread k owns row r = k + 1 and updates column c = t - k + 1 at
 * step t, so each point sees exactly the values the sequential loop gives
 * it -- west and south already updated this sweep (its own previous result,
 * and the row below's result of the previous step, passed through LDS), east
 * and north not yet (read from the field before their owners reach them:
 * the next column's operands are loaded one step ahead, and a workgroup's
 * accesses are ordered across its barriers).  Rows <= LEX_MAX_ROWS. */
struct LexOp {
    real_t aw, ae, as, an, ap, su, e, n;
};

__device__ __forceinline__ LexOp lex_op(const DevLevelG &L, const gdbl *phi, int ij, int npx)
{
    return LexOp{L.aw[ij], L.ae[ij], L.as[ij], L.an[ij], L.ap[ij], L.su[ij], phi[ij + 1],
                 phi[ij + npx]};
}

__global__ void __launch_bounds__(LEX_MAX_ROWS) k_lexgs(const DevLevel *Ls, int field, int nsw)
{
    const DevLevelG &L = reinterpret_cast<const DevLevelG *>(Ls)[blockIdx.z];
    if (L.ctl->done) return;
    int nx, ny;
    gdbl *phi = field_of(L, field, nx, ny);
    const int NC = nx - 1, NR = ny - 1, npx = L.npx;
    __shared__ real_t sbuf[2][LEX_MAX_ROWS];
    const int k = threadIdx.x, r = k + 1;
    const bool row = r <= NR;
    const int base = r * npx;
    for (int it = 0; it < nsw; ++it) {
        real_t west = 0.0;
        LexOp op{};
        if (row) {
            west = phi[base];
            op = lex_op(L, phi, base + 1, npx);
        }
        for (int t = 0; t < NC + NR - 1; ++t) {
            const int c = t - k + 1;
            if (row && c >= 1 && c <= NC) {
                const int ij = base + c;
                const real_t south = k == 0 ? (real_t)phi[ij - npx] : sbuf[(t - 1) & 1][k - 1];
                const real_t s =
                    ((op.aw * west + op.ae * op.e) + op.as * south) + op.an * op.n;
                const real_t x = (s + op.su) / op.ap;
                phi[ij] = x;
                west = x;
                sbuf[t & 1][k] = x;
                if (c < NC) op = lex_op(L, phi, ij + 1, npx);
            }
            __syncthreads();
        }
    }
}

/* S red-black sweeps per launch in FP64 (a solve's nsw sweeps in
 * ceil(nsw / 3) launches instead of nsw), temporally blocked as the mixed
 * path's FP32 sweeps are: a workgroup stages its 64 x RB_TY tile with a halo
 * of H = 2S points from `src` (points outside the swept range from `base`,
 * the field) and the stored FP64 equation of its points, and colour step m
 * updates the points at distance >= m from the region's edge, with
 * k_rb_sweep's expression; after 2S steps the tile holds what S single
 * sweeps give, bit for bit, and is written to `dst` (ping-pong, so every
 * workgroup's halo is the launch's old values). */
#define RBX_NT 512
/* relaxations of at least this many sweeps take the blocked launches (two
 * sweeps as one launch and a copy-back already beat two single sweeps: the
 * equation is read once) */
#ifndef MGF_RBX_MIN
#define MGF_RBX_MIN 2
#endif
template <int S>
struct RbxShape {
    static constexpr int H = 2 * S, RW = BX + 2 * H, RH = RB_TY + 2 * H, NR = RW * RH;
};

/* one workgroup's tile: t (NR), and for p' mg (RB_TY x BX) and rows (RB_TY)
 * in LDS */
template <int S, int ASMF, bool STORE, int PEND, bool STASH = false>
__device__ __forceinline__ void rb_sweeps_tile(const DevLevelG &L, const TileId &T, int field,
                                               int parity, real_t *t, double *mg, double *rows)
{
    constexpr int H = 2 * S, RW = BX + 2 * H, RH = RB_TY + 2 * H, NR = RW * RH;
    constexpr int NPT = (NR + RBX_NT - 1) / RBX_NT;
    /* p': the tile's |div| per point, then per tile row, for the partial sums */
    constexpr bool PF = ASMF == 2;
    int nx, ny;
    const gdbl *base = field_of(L, field, nx, ny);
    const gdbl *src = parity ? L.tmp : base;
    gdbl *dst = parity ? (gdbl *)base : L.tmp;
    /* PEND u, v (and sc with PEND 2): the result stays in uc / vc / scn for
     * the p' assembly and the update (the field keeps its old values, so no
     * copy-back); PEND 2: fluxes formed inline (relax_flux) */
    constexpr bool INL = PEND == 2;
    if constexpr (PEND && ASMF == 0) dst = L.uc;
    if constexpr (PEND && ASMF == 1) dst = L.vc;
    if constexpr (PEND == 2 && ASMF == 3) dst = L.scn;
    const int npx = L.npx, npy = L.npy;
    const int ri0 = 1 + T.x * BX - H, rj0 = 1 + T.y * RB_TY - H;
    const int tid = threadIdx.y * BX + threadIdx.x;
    if constexpr (PF) {
        for (int k = tid; k < RB_TY * BX; k += RBX_NT) mg[k] = 0.0;
        __syncthreads();
    }
    Pt own[NPT];
    /* per slot: distance to the region's edge (bits 0-7), colour (bit 8),
     * inside the swept range (bit 9) */
    unsigned meta[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        /* own[k] is set only for points inside the swept range (meta bit 9)
         * and read only for those: no initial value, which the compiler would
         * otherwise materialise in every branch */
        const int p = tid + k * RBX_NT;
        unsigned m = 0;
        if (p < NR) {
            const int r = p / RW, c = p - r * RW;
            const int gi = ri0 + c, gj = rj0 + r;
            real_t val = 0.0;
            if (gi >= 0 && gi < npx && gj >= 0 && gj < npy) {
                const int g = gj * npx + gi;
                const bool in = gi >= 1 && gi <= nx - 1 && gj >= 1 && gj <= ny - 1;
                const int d = min(min(c, RW - 1 - c), min(r, RH - 1 - r));
                /* p' starts every relaxation from zero (masserror, V6:1116) */
                val = in ? (PF ? real_t(0.0) : src[g]) : base[g];
                if (in) {
                    if constexpr (ASMF == 0) {
                        own[k] = asm64_u<INL>(L, gi, gj);
                    } else if constexpr (ASMF == 1) {
                        own[k] = asm64_v<INL>(L, gi, gj);
                    } else if constexpr (ASMF == 2) {
                        double mag;
                        own[k] = asm64_p<(PEND != 0)>(L, gi, gj, mag);
                        if (d >= H) mg[(r - H) * BX + (c - H)] = mag;
                    } else if constexpr (ASMF == 3) {
                        own[k] = asm64_sc<INL>(L, gi, gj);
                    } else {
                        own[k] = Pt{L.aw[g], L.ae[g], L.as[g], L.an[g], L.ap[g], L.su[g]};
                    }
                    m = (unsigned)d | ((unsigned)((gi + gj) & 1) << 8) | (1u << 9);
                }
            }
            t[p] = val;
        }
        meta[k] = m;
    }
    __syncthreads();
#pragma unroll
    for (int step = 1; step <= 2 * S; ++step) {
        const unsigned want = (1u << 9) | ((unsigned)((step - 1) & 1) << 8);
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const unsigned m = meta[k];
            if ((m & 0x300u) == want && (int)(m & 0xffu) >= step) {
                const int p = tid + k * RBX_NT;
                const Pt &o = own[k];
                const real_t sum_nb = ((o.aw * t[p - 1] + o.ae * t[p + 1]) + o.as * t[p - RW]) +
                                      o.an * t[p + RW];
                t[p] = (sum_nb + o.su) / o.ap;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const unsigned m = meta[k];
        if ((m & (1u << 9)) && (int)(m & 0xffu) >= H) {
            const int p = tid + k * RBX_NT;
            const int r = p / RW, c = p - r * RW;
            const int g = (rj0 + r) * npx + (ri0 + c);
            dst[g] = t[p];
            if constexpr (ASMF == 0) L.apu[g] = own[k].ap;
            if constexpr (ASMF == 1) L.apv[g] = own[k].ap;
            if constexpr (STORE) store_eq(L, g, own[k]);
            /* p' (last launch): the copy origins' updated p, as k_pstash */
            if constexpr (STASH)
                if (L.flags[g] & MGF_F_PORIG) L.pnew[g] = L.p[g] + L.P.relxp * (t[p] - 0.0);
        }
    }
    if constexpr (PF) {
        /* the |div| partial sums of k_asm_p, bit for bit: per 64 x 4 block,
         * each row summed by one wave with block_sum's shuffle tree, then the
         * block's four rows in order */
        const int w = threadIdx.y, lane = threadIdx.x;
        for (int row = w; row < RB_TY; row += RBX_NT / BX) {
            double v = mg[row * BX + lane];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) rows[row] = v;
        }
        __syncthreads();
        if (tid < RB_TY / BY) {
            const int q = tid;
            const double bs = ((rows[q * BY] + rows[q * BY + 1]) + rows[q * BY + 2]) +
                              rows[q * BY + 3];
            const int gxp = (npx - 2 + BX - 1) / BX;
            const int pb = (T.y * (RB_TY / BY) + q) * gxp + T.x;
            if (pb < L.nblk) L.partial[pb] = bs;
        }
    }
}

template <int S, int ASMF = -1, bool STORE = false, int PEND = 0, bool STASH = false>
__global__ void __launch_bounds__(RBX_NT) k_rb_sweeps(const DevLevel *Ls, int field, int parity)
{
    LEVEL(Ls);
    constexpr bool PF = ASMF == 2;
    __shared__ real_t t[RbxShape<S>::NR];
    __shared__ double mg[PF ? RB_TY * BX : 1];
    __shared__ double rows[PF ? RB_TY : 1];
    rb_sweeps_tile<S, ASMF, STORE, PEND, STASH>(L, T, field, parity, t, mg, rows);
}

/* The u, v and sc relaxations of a group in one launch (FP64, the pending
 * order without an outflow boundary: fluxes formed from u, v, results left
 * in uc, vc, scn, so none of the three reads what another writes).  The grid
 * is the sc tile grid with three logical z-slices per member (field = z %
 * 3); tiles past u's or v's smaller range hold no unknowns and write
 * nothing. */
template <int S>
__global__ void __launch_bounds__(RBX_NT) k_rb_sweeps_uvs(const DevLevel *Ls)
{
    const TileId T3 = tile_id();
    const TileId T{T3.x, T3.y, T3.z / 3};
    const int f = T3.z - 3 * T.z;
    const DevLevelG &L = reinterpret_cast<const DevLevelG *>(Ls)[T.z];
    if (L.ctl->done) return;
    __shared__ real_t t[RbxShape<S>::NR];
    if (f == 0)
        rb_sweeps_tile<S, 0, false, 2, false>(L, T, 0, 0, t, nullptr, nullptr);
    else if (f == 1)
        rb_sweeps_tile<S, 1, false, 2, false>(L, T, 1, 0, t, nullptr, nullptr);
    else
        rb_sweeps_tile<S, 3, false, 2, false>(L, T, 3, 0, t, nullptr, nullptr);
}

/* A whole FP64 red-black solve of u, v or sc with at most three sweeps in
 * one launch that also assembles the equation (k_asm_u / k_asm_v /
 * k_asm_sc's expressions, at every point of the region: the halo's points
 * are assembled again by each workgroup that needs them, from inputs the
 * launch does not write), then the copy-back; the equation never crosses
 * HBM.  PEND (u, v): no copy-back, the result stays in uc / vc.  0 when not
 * applicable (another smoother or more sweeps). */
template <int F, int PEND>
static void launch_relax64_f(const DevGroup &G, int nsw, const dim3 &g, hipStream_t st)
{
    const dim3 b(BX, RBX_NT / BX);
    if (nsw == 1)
        hipLaunchKernelGGL((k_rb_sweeps<1, F, false, PEND>), g, b, 0, st, G.d, F, 0);
    else if (nsw == 2)
        hipLaunchKernelGGL((k_rb_sweeps<2, F, false, PEND>), g, b, 0, st, G.d, F, 0);
    else
        hipLaunchKernelGGL((k_rb_sweeps<3, F, false, PEND>), g, b, 0, st, G.d, F, 0);
}

/* The FP64 p' relaxation (red-black): its first launch assembles (k_asm_p's
 * expressions, with the |div| partial sums), runs up to three sweeps from
 * zero and stores the tile's equation when launches on the stored equation
 * follow; err as mgf_launch_asm_p.  0 when not applicable. */
template <bool PEND, bool STASH>
static void launch_relax64_p0(const DevGroup &G, int n0, int store, const dim3 &g,
                              hipStream_t st)
{
    constexpr int PD = PEND ? 1 : 0;
    const dim3 b(BX, RBX_NT / BX);
    if (n0 == 3) {
        if (store) hipLaunchKernelGGL((k_rb_sweeps<3, 2, true, PD>), g, b, 0, st, G.d, 2, 0);
        else hipLaunchKernelGGL((k_rb_sweeps<3, 2, false, PD, STASH>), g, b, 0, st, G.d, 2, 0);
    } else if (n0 == 2) {
        hipLaunchKernelGGL((k_rb_sweeps<2, 2, false, PD, STASH>), g, b, 0, st, G.d, 2, 0);
    } else {
        hipLaunchKernelGGL((k_rb_sweeps<1, 2, false, PD, STASH>), g, b, 0, st, G.d, 2, 0);
    }
}

int mgf_launch_relax64_p(const DevGroup &G, int isolve, int nsw, int err, int pend,
                         hipStream_t st, int *stashed)
{
#ifdef MGF_TANGENT
    (void)G; (void)isolve; (void)nsw; (void)err; (void)pend; (void)st; (void)stashed;
    return 0;
#else
    const int nx = G.npx - 1, ny = G.npy - 1;
    if (isolve != 2 || nsw < 1 || nx - 1 < 1 || ny - 1 < 1) return 0;
    const dim3 g = tile_grid(nx - 1, ny - 1, G.count, RB_TY), b(BX, RBX_NT / BX);
    const int n0 = nsw < 3 ? nsw : 3;
    /* stash (in pnew) when asked, copies exist and the last launch is a
     * blocked one (not the single-sweep kernel of a one-sweep remainder) */
    const int nlast = nsw <= 3 ? nsw : (nsw - 4) % 3 + 1;
    const bool stash = stashed && G.max_npcopy > 0 && (nsw <= 3 || nlast >= 2);
    if (stashed) *stashed = stash;
    if (pend && stash && nsw <= 3)
        launch_relax64_p0<true, true>(G, n0, 0, g, st);
    else if (pend)
        launch_relax64_p0<true, false>(G, n0, nsw > 3, g, st);
    else
        launch_relax64_p0<false, false>(G, n0, nsw > 3, g, st);
    if (err) mgf_launch_reduce_err(G, err == 2, st);
    int launches = 1;
    for (int done = n0; done < nsw; ++launches) {
        const int n = nsw - done < 3 ? nsw - done : 3;
        const bool last = done + n == nsw;
        if (n == 3 && last && stash)
            hipLaunchKernelGGL((k_rb_sweeps<3, -1, false, 0, true>), g, b, 0, st, G.d, 2,
                               launches & 1);
        else if (n == 3)
            hipLaunchKernelGGL(k_rb_sweeps<3>, g, b, 0, st, G.d, 2, launches & 1);
        else if (n == 2 && last && stash)
            hipLaunchKernelGGL((k_rb_sweeps<2, -1, false, 0, true>), g, b, 0, st, G.d, 2,
                               launches & 1);
        else if (n == 2)
            hipLaunchKernelGGL(k_rb_sweeps<2>, g, b, 0, st, G.d, 2, launches & 1);
        else
            hipLaunchKernelGGL(k_rb_sweep, g, dim3(BX, BY), 0, st, G.d, 2, launches & 1);
        done += n;
    }
    if (launches & 1)
        hipLaunchKernelGGL(k_copy_range, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st,
                           G.d, 2);
    return 1;
#endif
}

/* u, v and sc with pend 2 in one launch (k_rb_sweeps_uvs); 0 when not
 * applicable (launching nothing) */
int mgf_launch_relax64_uvs(const DevGroup &G, int isolve, int nsw, hipStream_t st)
{
#ifdef MGF_TANGENT
    (void)G; (void)isolve; (void)nsw; (void)st;
    return 0;
#else
    if (!mgf_relax64_uv_pending(G, isolve, nsw) || G.any_outflow) return 0;
    const dim3 g0 = tile_grid(G.npx - 2, G.npy - 2, G.count, RB_TY);
    const dim3 g(g0.x, g0.y, 3 * g0.z), b(BX, RBX_NT / BX);
    if (nsw == 1)
        hipLaunchKernelGGL(k_rb_sweeps_uvs<1>, g, b, 0, st, G.d);
    else if (nsw == 2)
        hipLaunchKernelGGL(k_rb_sweeps_uvs<2>, g, b, 0, st, G.d);
    else
        hipLaunchKernelGGL(k_rb_sweeps_uvs<3>, g, b, 0, st, G.d);
    return 1;
#endif
}

int mgf_relax64_uv_pending(const DevGroup &G, int isolve, int nsw)
{
#ifdef MGF_TANGENT
    (void)G; (void)isolve; (void)nsw;
    return 0;
#else
    /* both u (npx-3 x npy-2 points) and v (npx-2 x npy-3) take one launch */
    return isolve == 2 && nsw >= 1 && nsw <= 3 && G.npx - 3 >= 1 && G.npy - 3 >= 1;
#endif
}

int mgf_launch_relax64(const DevGroup &G, int field, int isolve, int nsw, hipStream_t st,
                       int pend)
{
#ifdef MGF_TANGENT
    (void)G; (void)field; (void)isolve; (void)nsw; (void)st; (void)pend;
    return 0;
#else
    if (isolve != 2 || nsw < 1 || nsw > 3 || field == 2) return 0;
    const int nx = field == 0 ? G.npx - 2 : G.npx - 1;
    const int ny = field == 1 ? G.npy - 2 : G.npy - 1;
    if (nx - 1 < 1 || ny - 1 < 1) return 0;
    const dim3 g = tile_grid(nx - 1, ny - 1, G.count, RB_TY);
    if (field == 3 && pend != 2) pend = 0;
    if (field == 0 && pend == 2)
        launch_relax64_f<0, 2>(G, nsw, g, st);
    else if (field == 0 && pend)
        launch_relax64_f<0, 1>(G, nsw, g, st);
    else if (field == 0)
        launch_relax64_f<0, 0>(G, nsw, g, st);
    else if (field == 1 && pend == 2)
        launch_relax64_f<1, 2>(G, nsw, g, st);
    else if (field == 1 && pend)
        launch_relax64_f<1, 1>(G, nsw, g, st);
    else if (field == 1)
        launch_relax64_f<1, 0>(G, nsw, g, st);
    else if (pend)
        launch_relax64_f<3, 2>(G, nsw, g, st);
    else
        launch_relax64_f<3, 0>(G, nsw, g, st);
    if (!pend)
        hipLaunchKernelGGL(k_copy_range, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st,
                           G.d, field);
    return 1;
#endif
}

void mgf_launch_smooth(const DevGroup &G, int field, int isolve, int nsw, hipStream_t st)
{
    int nx, ny;
    switch (field) {
    case 0: nx = G.npx - 2; ny = G.npy - 1; break;
    case 1: nx = G.npx - 1; ny = G.npy - 2; break;
    default: nx = G.npx - 1; ny = G.npy - 1; break;
    }
    if (nx - 1 < 1 || ny - 1 < 1 || nsw <= 0 || isolve < 1 || isolve > 3) return;
    if (isolve == 3) {
        const unsigned nt = (unsigned)((G.npy - 2 + 63) / 64 * 64);
        hipLaunchKernelGGL(k_lexgs, dim3(1, 1, (unsigned)G.count), dim3(nt), 0, st, G.d, field,
                           nsw);
        return;
    }
#ifndef MGF_TANGENT
    if (isolve == 2 && nsw >= MGF_RBX_MIN) {
        /* three sweeps per launch (the last launch the rest); the launch
         * count's parity says where the result is */
        const dim3 g = tile_grid(nx - 1, ny - 1, G.count, RB_TY);
        const dim3 b(BX, RBX_NT / BX);
        int launches = 0;
        for (int done = 0; done < nsw; ++launches) {
            const int n = nsw - done < 3 ? nsw - done : 3;
            if (n == 3)
                hipLaunchKernelGGL(k_rb_sweeps<3>, g, b, 0, st, G.d, field, launches & 1);
            else if (n == 2)
                hipLaunchKernelGGL(k_rb_sweeps<2>, g, b, 0, st, G.d, field, launches & 1);
            else
                hipLaunchKernelGGL(k_rb_sweep, g, dim3(BX, BY), 0, st, G.d, field, launches & 1);
            done += n;
        }
        if (launches & 1)
            hipLaunchKernelGGL(k_copy_range, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0,
                               st, G.d, field);
        return;
    }
#endif
    for (int it = 0; it < nsw; ++it) {
        if (isolve == 1)
            hipLaunchKernelGGL(k_jacobi, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st, G.d,
                               field, it & 1);
        else
            hipLaunchKernelGGL(k_rb_sweep, tile_grid(nx - 1, ny - 1, G.count, RB_TY), dim3(BX, BY),
                               0, st, G.d, field, it & 1);
    }
    if (nsw & 1)
        hipLaunchKernelGGL(k_copy_range, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st,
                           G.d, field);
}

void mgf_launch_update(const DevGroup &G, int pending, hipStream_t st)
{
    if (G.max_npcopy > 0 && pending != 2 && pending != 5)
        hipLaunchKernelGGL(k_pstash, dim3((unsigned)((G.max_npcopy + 255) / 256), 1,
                                          (unsigned)G.count),
                           dim3(256), 0, st, G.d);
    const dim3 g = tile_grid(G.npx - 2, G.npy - 2, G.count);
    if (pending == 5)
        hipLaunchKernelGGL(k_update<5>, g, dim3(BX, BY), 0, st, G.d);
    else if (pending == 4)
        hipLaunchKernelGGL(k_update<4>, g, dim3(BX, BY), 0, st, G.d);
    else if (pending == 3)
        hipLaunchKernelGGL(k_update<3>, g, dim3(BX, BY), 0, st, G.d);
    else if (pending == 2)
        hipLaunchKernelGGL(k_update<2>, g, dim3(BX, BY), 0, st, G.d);
    else if (pending)
        hipLaunchKernelGGL(k_update<1>, g, dim3(BX, BY), 0, st, G.d);
    else
        hipLaunchKernelGGL(k_update<0>, g, dim3(BX, BY), 0, st, G.d);
}

void mgf_launch_pbound(const DevGroup &G, hipStream_t st)
{
    if (G.any_pbound_x)
        hipLaunchKernelGGL(k_pbound_x, dim3((unsigned)((G.npy + 255) / 256), 1, (unsigned)G.count),
                           dim3(256), 0, st, G.d);
    if (G.any_pbound_y)
        hipLaunchKernelGGL(k_pbound_y, dim3(1, 1, (unsigned)G.count), dim3(64), 0, st, G.d);
}

void mgf_launch_outflow(const DevGroup &G, hipStream_t st)
{
    if (G.any_outflow)
        hipLaunchKernelGGL(k_outflow, dim3((unsigned)((G.npy + 255) / 256), 1, (unsigned)G.count),
                           dim3(256), 0, st, G.d);
}

void mgf_launch_restrict(const DevPair &R, int restv, hipStream_t st)
{
    dim3 g = tile_grid(R.npxc - 2, R.npyc - 2, R.count);
    if (restv) hipLaunchKernelGGL(k_restv, g, dim3(BX, BY), 0, st, R.f, R.c);
    hipLaunchKernelGGL(k_tau, g, dim3(BX, TAU_NT / BX), 0, st, R.f, R.c);
}

void mgf_launch_update_restv(const DevGroup &G, const DevPair &R, int pending, hipStream_t st)
{
    if (G.max_npcopy > 0 && pending != 2 && pending != 5)
        hipLaunchKernelGGL(k_pstash, dim3((unsigned)((G.max_npcopy + 255) / 256), 1,
                                          (unsigned)G.count),
                           dim3(256), 0, st, G.d);
    const dim3 g = tile_grid(G.npx - 2, G.npy - 2, G.count);
    if (pending == 5)
        hipLaunchKernelGGL(k_update_restv<5>, g, dim3(BX, BY), 0, st, G.d, R.c);
    else if (pending == 4)
        hipLaunchKernelGGL(k_update_restv<4>, g, dim3(BX, BY), 0, st, G.d, R.c);
    else if (pending == 2)
        hipLaunchKernelGGL(k_update_restv<2>, g, dim3(BX, BY), 0, st, G.d, R.c);
    else if (pending)
        hipLaunchKernelGGL(k_update_restv<1>, g, dim3(BX, BY), 0, st, G.d, R.c);
    else
        hipLaunchKernelGGL(k_update_restv<0>, g, dim3(BX, BY), 0, st, G.d, R.c);
}

void mgf_launch_prolong(const DevPair &R, int stored, hipStream_t st)
{
    const dim3 g = tile_grid(R.npxc - 2, R.npyc - 2, R.count);
    if (stored) {
        hipLaunchKernelGGL(k_prol_s, g, dim3(BX, BY), 0, st, R.f, R.c);
        return;
    }
    hipLaunchKernelGGL(k_prol_c, g, dim3(BX, BY), 0, st, R.f, R.c);
    hipLaunchKernelGGL(k_prol_apply, g, dim3(BX, BY), 0, st, R.f, R.c);
}

void mgf_launch_end_cycle(DevCtl *ctls, const int *members, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_end_cycle, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, ctls,
                       members, n);
}

void mgf_launch_fill(real_t *a, double v, size_t n, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, v, n);
}
