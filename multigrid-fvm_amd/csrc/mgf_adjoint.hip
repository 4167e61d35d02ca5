/*
 * mgf_adjoint.hip -- the reverse-mode discrete adjoint of include/mgfvm_adjoint.h.
 *
 * Unknowns and equations (level 0 of each problem): the interior u faces,
 * v faces, cells (continuity / p) and cells (scalar) that carry no 1e30
 * penalty.  The steady residual of each row is restated here, templated on
 * the scalar type, from the reference's expressions:
 *   R_u = ap u_P - aw u_W - ae u_E - as u_S - an u_N - areax (p_P - p_E)
 *         with coefu's hybrid coefficients and wall adds
 *         (Version-6-26-04.c:815-860; umom's under-relaxation, V6:741-767,
 *         cancels at the fixed point);
 *   R_v   the same with coefv and its side-wall adds (V6:966-1016, 769-793);
 *   R_c = areax (u_P - u_W) + areay (v_P - v_S)          (masserror, V6:1096-1125);
 *   R_s = ap s_P - sum a_nb s_nb with coefsc and its obstacle adds
 *         (V6:861-964, scalar V6:795-812).
 * Evaluated with a dual number seeded on one unknown (or one parameter),
 * a row gives one entry of the Jacobian K (or of dR/dtheta) exactly, with
 * fmax / fabs branches frozen at the state's values.  K^T lambda is formed
 * in gather form: each unknown sums lambda_r dR_r/dx over the rows whose
 * stencil holds it (13 rows for a u or v face, 4 for p, 5 for sc), all
 * recomputed from the stored converged state (no coefficient storage).
 *
 * Solver: right-preconditioned restarted GMRES per problem (lock step over
 * a batch; Arnoldi with classical Gram-Schmidt and one reorthogonalisation,
 * the small least-squares problems on the host), the preconditioner one
 * V-cycle of the same level hierarchy as the primal solve:
 *   relax(level):  the scalar multiplier by red-black sweeps of the
 *                  transposed scalar operator; the u, v multipliers by
 *                  red-black sweeps of the transposed momentum operators
 *                  (diagonal ap / relxm as the reference's umom/vmom, the
 *                  continuity multiplier's gradient and the scalar coupling
 *                  on the right-hand side); the continuity defect, the
 *                  pressure-correction equation of coefp (areas^2 / ap) for
 *                  the continuity multiplier's correction, and the update
 *                  (V6:1127-1151's form);
 *   down:          relax, residual of the level's Picard operator, restrict
 *                  (restru/restrv's face weights, V6:1547-1560; cell sums);
 *   coarsest:      relaxed four times;
 *   up:            prolongate (prolu/prolv/prolp weights, V6:1782-1813,
 *                  1932-1955) and relax.
 * The transposed operator of a convection-diffusion stencil at a point has
 * the point's own diagonal and, as off-diagonals, its neighbours'
 * coefficients towards it: for the hybrid scheme these are the point's own
 * face coefficients with the face flux reversed (ae(W) = h-(cxm_P) etc.), so
 * they too come from the point's own fluxes.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>

#include <cmath>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mgf_dev.h"
#include "mgf_plan.h"
#include "mgfvm.h"
#include "mgfvm_adjoint.h"

#define ABX 64
#define ABY 4

/* ---------------------------------------------------------------------- */
/* descriptors                                                              */
/* ---------------------------------------------------------------------- */

/* one level of one problem: the primal descriptor (state, flags, metrics)
 * and the adjoint arrays, each npx*npy doubles: correction e (u v c s),
 * right-hand side f, residual / sweep right-hand side r, the continuity
 * multiplier's correction dp, and the recomputed momentum diagonals
 * apu = ap_u / relxm, apv = ap_v / relxm (1e30 where the face is fixed) */
struct AdjLevel {
    DevLevel P;
    int inst;
    const int *live;
    double *e[4], *f[4], *r[4];
    double *dp, *apu, *apv;
    double *tmp;        /* the other buffer of the blocked sweeps */
};

struct AdjGroup {
    AdjLevel *d;
    int count, npx, npy, depth;
};
struct AdjPair {
    AdjLevel *f, *c;
    int count, npxf, npyf, npxc, npyc;
};

/* derivative seeds of the metrics for the parameter gradients */
struct MetSeed {
    double du_x, du_y, ds_x, ds_y;
};

/* per problem: level-0 vectors (4 arrays of npx*npy, contiguous) */
struct AdjVec {
    double *V;           /* restart + 1 Krylov vectors */
    double *X, *G, *W;   /* solution, right-hand side dJ/dx, work */
    long n4;             /* 4 * npx * npy */
};

/* ---------------------------------------------------------------------- */
/* scalar helpers                                                           */
/* ---------------------------------------------------------------------- */

__device__ __forceinline__ double dfabs(double a) { return fabs(a); }
__device__ __forceinline__ double dfmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ mgf_dual dfabs(mgf_dual a) { return fabs(a); }
__device__ __forceinline__ mgf_dual dfmax(mgf_dual a, mgf_dual b) { return fmax(a, b); }

template <class S> __device__ __forceinline__ S hplus(S c, S d) { return dfmax(dfabs(c), d) + c; }
template <class S> __device__ __forceinline__ S hminus(S c, S d) { return dfmax(dfabs(c), d) - c; }

template <class S> __device__ __forceinline__ S mk(double v, bool seed);
template <> __device__ __forceinline__ double mk<double>(double v, bool) { return v; }
template <> __device__ __forceinline__ mgf_dual mk<mgf_dual>(double v, bool seed)
{
    return mgf_dual(v, seed ? 1.0 : 0.0);
}

/* a seed: field (0 u, 1 v, 2 p, 3 sc) and flat index, or the whole east
 * boundary column of u (east = 1), or none (f < 0) */
struct Seed {
    int f, ij, east;
};

/* the state read by the residual rows, with a derivative seed */
template <class S>
struct State {
    const double *u, *v, *p, *s;
    int nx;
    Seed sd;
    __device__ S U(int k) const
    {
        bool hit = (sd.f == 0 && k == sd.ij) ||
                   (sd.east && (k % nx) == nx - 2);
        return mk<S>(u[k], hit);
    }
    __device__ S V(int k) const { return mk<S>(v[k], sd.f == 1 && k == sd.ij); }
    __device__ S P(int k) const { return mk<S>(p[k], sd.f == 2 && k == sd.ij); }
    __device__ S Sc(int k) const { return mk<S>(s[k], sd.f == 3 && k == sd.ij); }
};

template <class S>
struct Met {
    S du_x, du_y, ds_x, ds_y;
};

template <class S>
__device__ __forceinline__ Met<S> metrics(const DevLevel &L, const MetSeed *ms)
{
    Met<S> m;
    m.du_x = mk<S>(L.du_x, false);
    m.du_y = mk<S>(L.du_y, false);
    m.ds_x = mk<S>(L.ds_x, false);
    m.ds_y = mk<S>(L.ds_y, false);
    return m;
}
template <>
__device__ __forceinline__ Met<mgf_dual> metrics<mgf_dual>(const DevLevel &L, const MetSeed *ms)
{
    Met<mgf_dual> m;
    m.du_x = mgf_dual(L.du_x, ms ? ms->du_x : 0.0);
    m.du_y = mgf_dual(L.du_y, ms ? ms->du_y : 0.0);
    m.ds_x = mgf_dual(L.ds_x, ms ? ms->ds_x : 0.0);
    m.ds_y = mgf_dual(L.ds_y, ms ? ms->ds_y : 0.0);
    return m;
}

/* ---------------------------------------------------------------------- */
/* the steady residual rows                                                 */
/* ---------------------------------------------------------------------- */

__device__ __forceinline__ bool urow(const DevLevel &L, int i, int j)
{
    return i >= 1 && i <= L.npx - 3 && j >= 1 && j <= L.npy - 2 &&
           !(L.flags[j * L.npx + i] & MGF_F_SPU);
}
__device__ __forceinline__ bool vrow(const DevLevel &L, int i, int j)
{
    return i >= 1 && i <= L.npx - 2 && j >= 1 && j <= L.npy - 3 &&
           !(L.flags[j * L.npx + i] & MGF_F_SPV);
}
__device__ __forceinline__ bool crow(const DevLevel &L, int i, int j)
{
    return i >= 1 && i <= L.npx - 2 && j >= 1 && j <= L.npy - 2 &&
           !(L.flags[j * L.npx + i] & MGF_F_OBS);
}
__device__ __forceinline__ bool srow(const DevLevel &L, int i, int j)
{
    return i >= 1 && i <= L.npx - 2 && j >= 1 && j <= L.npy - 2 &&
           !(L.flags[j * L.npx + i] & MGF_F_SPSC);
}

/* R_u at face (i, j): coefu (V6:825-858) + umom without relaxation */
template <class S>
__device__ S res_u(const DevLevel &L, const Met<S> &m, const State<S> &x, int i, int j)
{
    const int nx = L.npx, ij = j * nx + i, ijs = ij - nx;
    const double ax = L.ax, ay = L.ay;
    const S uP = x.U(ij), uW = x.U(ij - 1), uE = x.U(ij + 1), uS = x.U(ijs), uN = x.U(ij + nx);
    const S cxm = 0.5 * ((0.5 * uP) * ax + (0.5 * uW) * ax);
    const S cxp = 0.5 * ((0.5 * uP) * ax + (0.5 * uE) * ax);
    const S cym = 0.5 * ((0.5 * x.V(ijs)) * ay + (0.5 * x.V(ijs + 1)) * ay);
    const S cyp = 0.5 * ((0.5 * x.V(ij)) * ay + (0.5 * x.V(ij + 1)) * ay);
    S aw = hplus(cxm, m.du_x), ae = hminus(cxp, m.du_x);
    S as = hplus(cym, m.du_y), an = hminus(cyp, m.du_y);
    if (j == L.npy - 2) an = an + m.du_y;
    if (j == 1) as = as + m.du_y;
    const S ap = ((aw + ae) + as) + an;
    return ap * uP - aw * uW - ae * uE - as * uS - an * uN - ax * (x.P(ij) - x.P(ij + 1));
}

/* R_v at face (i, j): coefv (V6:976-1015) + vmom without relaxation */
template <class S>
__device__ S res_v(const DevLevel &L, const Met<S> &m, const State<S> &x, int i, int j)
{
    const int nx = L.npx, ij = j * nx + i, ijn = ij + nx, ijs = ij - nx;
    const double ax = L.ax, ay = L.ay;
    const S vP = x.V(ij), vW = x.V(ij - 1), vE = x.V(ij + 1), vS = x.V(ijs), vN = x.V(ijn);
    const S cxm = 0.5 * ((0.5 * x.U(ij - 1)) * ax + (0.5 * x.U(ijn - 1)) * ax);
    const S cxp = 0.5 * ((0.5 * x.U(ij)) * ax + (0.5 * x.U(ijn)) * ax);
    const S cym = 0.5 * ((0.5 * vP) * ay + (0.5 * vS) * ay);
    const S cyp = 0.5 * ((0.5 * vP) * ay + (0.5 * vN) * ay);
    S aw = hplus(cxm, m.du_x), ae = hminus(cxp, m.du_x);
    const S as = hplus(cym, m.du_y), an = hminus(cyp, m.du_y);
    if (i == nx - 2) ae = ae + m.du_y;   /* V6:1010: amu * areay / dy */
    if (i == 1) aw = aw + m.du_y;
    const S ap = ((aw + ae) + as) + an;
    return ap * vP - aw * vW - ae * vE - as * vS - an * vN - ay * (x.P(ij) - x.P(ijn));
}

/* R_c at cell (i, j): masserror's divergence (V6:1113-1114) */
template <class S>
__device__ S res_c(const DevLevel &L, const State<S> &x, int i, int j)
{
    const int ij = j * L.npx + i;
    return L.ax * x.U(ij) - L.ax * x.U(ij - 1) + L.ay * x.V(ij) - L.ay * x.V(ij - L.npx);
}

/* R_s at cell (i, j): coefsc with its obstacle adds (V6:872-963) + scalar */
template <class S>
__device__ S res_s(const DevLevel &L, const Met<S> &m, const State<S> &x, int i, int j)
{
    const int nx = L.npx, ij = j * nx + i, ijs = ij - nx;
    const double ax = L.ax, ay = L.ay;
    const S cxm = (0.5 * x.U(ij - 1)) * ax, cxp = (0.5 * x.U(ij)) * ax;
    const S cym = (0.5 * x.V(ijs)) * ay, cyp = (0.5 * x.V(ij)) * ay;
    S aw = hplus(cxm, m.ds_x), ae = hminus(cxp, m.ds_x);
    S as = hplus(cym, m.ds_y), an = hminus(cyp, m.ds_y);
    S ap = ((aw + ae) + as) + an;
    if (L.flags[ij] & MGF_F_SCADD) {
        const uint64_t code = L.sccode[ij];
        const int n_ae = (int)(code & 31u), n_aw = (int)((code >> 5) & 31u);
        const int n_an = (int)((code >> 10) & 31u), n_as = (int)((code >> 15) & 31u);
        const int n_ap = (int)((code >> 20) & 63u);
        for (int k = 0; k < n_ae; ++k) ae = ae + m.ds_x;
        for (int k = 0; k < n_aw; ++k) aw = aw + m.ds_x;
        for (int k = 0; k < n_an; ++k) an = an + m.ds_y;
        for (int k = 0; k < n_as; ++k) as = as + m.ds_y;
        const uint64_t seq = code >> 26;
        for (int k = 0; k < n_ap; ++k) ap = ap + (((seq >> k) & 1u) ? m.ds_y : m.ds_x);
    }
    const S sP = x.Sc(ij);
    return ap * sP - aw * x.Sc(ij - 1) - ae * x.Sc(ij + 1) - as * x.Sc(ijs) - an * x.Sc(ij + nx);
}

/* d(row)/d(seed) * lambda_row for the row (t, i, j) if it is an equation */
__device__ __forceinline__ double drow(const DevLevel &L, const Met<mgf_dual> &m,
                                       const State<mgf_dual> &x, const double *const lam[4],
                                       int t, int i, int j)
{
    if (i < 0 || j < 0 || i >= L.npx || j >= L.npy) return 0.0;
    const int ij = j * L.npx + i;
    switch (t) {
    case 0:
        if (!urow(L, i, j) || lam[0][ij] == 0.0) return 0.0;
        return lam[0][ij] * res_u(L, m, x, i, j).d;
    case 1:
        if (!vrow(L, i, j) || lam[1][ij] == 0.0) return 0.0;
        return lam[1][ij] * res_v(L, m, x, i, j).d;
    case 2:
        if (!crow(L, i, j) || lam[2][ij] == 0.0) return 0.0;
        return lam[2][ij] * res_c(L, x, i, j).d;
    default:
        if (!srow(L, i, j) || lam[3][ij] == 0.0) return 0.0;
        return lam[3][ij] * res_s(L, m, x, i, j).d;
    }
}

/* ---------------------------------------------------------------------- */
/* K^T lambda on level 0 (exact)                                            */
/* ---------------------------------------------------------------------- */

/* y = K^T lam: lam and y are 4 arrays (u, v, c, s) of npx*npy each */
__global__ void k_adj_kt(const AdjLevel *As, const double *const *lamv, double *const *yv)
{
    const AdjLevel &A = As[blockIdx.z];
    if (!A.live[A.inst]) return;
    const DevLevel &L = A.P;
    const int i = (int)(blockIdx.x * ABX + threadIdx.x), j = (int)(blockIdx.y * ABY + threadIdx.y);
    const int nx = L.npx, ny = L.npy;
    if (i >= nx || j >= ny) return;
    const int ij = j * nx + i;
    const double *lam[4];
    double *y[4];
    for (int k = 0; k < 4; ++k) {
        lam[k] = lamv[4 * A.inst + k];
        y[k] = yv[4 * A.inst + k];
    }
    const Met<mgf_dual> m = metrics<mgf_dual>(L, nullptr);
    State<mgf_dual> x{L.u, L.v, L.p, L.sc, nx, Seed{-1, 0, 0}};
    /* u column */
    double acc = 0.0;
    if (urow(L, i, j)) {
        x.sd = Seed{0, ij, 0};
        acc = drow(L, m, x, lam, 0, i, j) + drow(L, m, x, lam, 0, i - 1, j) +
              drow(L, m, x, lam, 0, i + 1, j) + drow(L, m, x, lam, 0, i, j - 1) +
              drow(L, m, x, lam, 0, i, j + 1) + drow(L, m, x, lam, 1, i, j) +
              drow(L, m, x, lam, 1, i, j - 1) + drow(L, m, x, lam, 1, i + 1, j) +
              drow(L, m, x, lam, 1, i + 1, j - 1) + drow(L, m, x, lam, 2, i, j) +
              drow(L, m, x, lam, 2, i + 1, j) + drow(L, m, x, lam, 3, i, j) +
              drow(L, m, x, lam, 3, i + 1, j);
    }
    y[0][ij] = acc;
    acc = 0.0;
    if (vrow(L, i, j)) {
        x.sd = Seed{1, ij, 0};
        acc = drow(L, m, x, lam, 1, i, j) + drow(L, m, x, lam, 1, i - 1, j) +
              drow(L, m, x, lam, 1, i + 1, j) + drow(L, m, x, lam, 1, i, j - 1) +
              drow(L, m, x, lam, 1, i, j + 1) + drow(L, m, x, lam, 0, i, j) +
              drow(L, m, x, lam, 0, i - 1, j) + drow(L, m, x, lam, 0, i, j + 1) +
              drow(L, m, x, lam, 0, i - 1, j + 1) + drow(L, m, x, lam, 2, i, j) +
              drow(L, m, x, lam, 2, i, j + 1) + drow(L, m, x, lam, 3, i, j) +
              drow(L, m, x, lam, 3, i, j + 1);
    }
    y[1][ij] = acc;
    acc = 0.0;
    if (crow(L, i, j)) {
        x.sd = Seed{2, ij, 0};
        acc = drow(L, m, x, lam, 0, i, j) + drow(L, m, x, lam, 0, i - 1, j) +
              drow(L, m, x, lam, 1, i, j) + drow(L, m, x, lam, 1, i, j - 1);
    }
    y[2][ij] = acc;
    acc = 0.0;
    if (srow(L, i, j)) {
        x.sd = Seed{3, ij, 0};
        acc = drow(L, m, x, lam, 3, i, j) + drow(L, m, x, lam, 3, i - 1, j) +
              drow(L, m, x, lam, 3, i + 1, j) + drow(L, m, x, lam, 3, i, j - 1) +
              drow(L, m, x, lam, 3, i, j + 1);
    }
    y[3][ij] = acc;
}

/* ---------------------------------------------------------------------- */
/* the preconditioner V-cycle (Picard transposed operators)                 */
/* ---------------------------------------------------------------------- */

struct Tco {
    double w, e, s, n;
};

/* the transposed hybrid coefficients of the u equation at face (i, j):
 * the neighbours' coefficients towards it, from its own face fluxes with
 * the flux reversed; and its own (unrelaxed) diagonal with the wall adds */
__device__ __forceinline__ Tco tco_u(const DevLevel &L, int i, int j, double *ap)
{
    const int nx = L.npx, ij = j * nx + i, ijs = ij - nx;
    const double ax = L.ax, ay = L.ay;
    const double cxm = 0.5 * ((0.5 * L.u[ij]) * ax + (0.5 * L.u[ij - 1]) * ax);
    const double cxp = 0.5 * ((0.5 * L.u[ij]) * ax + (0.5 * L.u[ij + 1]) * ax);
    const double cym = 0.5 * ((0.5 * L.v[ijs]) * ay + (0.5 * L.v[ijs + 1]) * ay);
    const double cyp = 0.5 * ((0.5 * L.v[ij]) * ay + (0.5 * L.v[ij + 1]) * ay);
    double aw = hplus(cxm, (double)L.du_x), ae = hminus(cxp, (double)L.du_x);
    double as = hplus(cym, (double)L.du_y), an = hminus(cyp, (double)L.du_y);
    if (j == L.npy - 2) an += L.du_y;
    if (j == 1) as += L.du_y;
    if (ap) *ap = ((aw + ae) + as) + an;
    return Tco{hminus(cxm, (double)L.du_x), hplus(cxp, (double)L.du_x),
               hminus(cym, (double)L.du_y), hplus(cyp, (double)L.du_y)};
}

__device__ __forceinline__ Tco tco_v(const DevLevel &L, int i, int j, double *ap)
{
    const int nx = L.npx, ij = j * nx + i, ijn = ij + nx, ijs = ij - nx;
    const double ax = L.ax, ay = L.ay;
    const double cxm = 0.5 * ((0.5 * L.u[ij - 1]) * ax + (0.5 * L.u[ijn - 1]) * ax);
    const double cxp = 0.5 * ((0.5 * L.u[ij]) * ax + (0.5 * L.u[ijn]) * ax);
    const double cym = 0.5 * ((0.5 * L.v[ij]) * ay + (0.5 * L.v[ijs]) * ay);
    const double cyp = 0.5 * ((0.5 * L.v[ij]) * ay + (0.5 * L.v[ijn]) * ay);
    double aw = hplus(cxm, (double)L.du_x), ae = hminus(cxp, (double)L.du_x);
    const double as = hplus(cym, (double)L.du_y), an = hminus(cyp, (double)L.du_y);
    if (i == nx - 2) ae += L.du_y;
    if (i == 1) aw += L.du_y;
    if (ap) *ap = ((aw + ae) + as) + an;
    return Tco{hminus(cxm, (double)L.du_x), hplus(cxp, (double)L.du_x),
               hminus(cym, (double)L.du_y), hplus(cyp, (double)L.du_y)};
}

__device__ __forceinline__ Tco tco_s(const DevLevel &L, int i, int j, double *ap)
{
    const int nx = L.npx, ij = j * nx + i, ijs = ij - nx;
    const double ax = L.ax, ay = L.ay;
    const double cxm = (0.5 * L.u[ij - 1]) * ax, cxp = (0.5 * L.u[ij]) * ax;
    const double cym = (0.5 * L.v[ijs]) * ay, cyp = (0.5 * L.v[ij]) * ay;
    if (ap) {
        const Met<double> m = metrics<double>(L, nullptr);
        double aw = hplus(cxm, m.ds_x), ae = hminus(cxp, m.ds_x);
        double as = hplus(cym, m.ds_y), an = hminus(cyp, m.ds_y);
        double a = ((aw + ae) + as) + an;
        if (L.flags[ij] & MGF_F_SCADD) {
            const uint64_t code = L.sccode[ij];
            const int n_ap = (int)((code >> 20) & 63u);
            const uint64_t seq = code >> 26;
            for (int k = 0; k < n_ap; ++k) a += ((seq >> k) & 1u) ? m.ds_y : m.ds_x;
        }
        *ap = a;
    }
    return Tco{hminus(cxm, (double)L.ds_x), hplus(cxp, (double)L.ds_x),
               hminus(cym, (double)L.ds_y), hplus(cyp, (double)L.ds_y)};
}

/* d h-(c)/dc * (area / 2) and d h+(c)/dc * (area / 2) of the scalar's face
 * coefficient through a face with flux c (the scalar coupling S^T) */
__device__ __forceinline__ void dsc(double c, double d, double half_area, double *dm, double *dp)
{
    const double sg = (fabs(c) >= d) ? (c < 0.0 ? -1.0 : 1.0) : 0.0;
    *dm = (sg - 1.0) * half_area;
    *dp = (sg + 1.0) * half_area;
}

#define ADJ_LEVEL                                                       \
    const AdjLevel &A = As[blockIdx.z];                                 \
    if (!A.live[A.inst]) return;                                        \
    const DevLevel &L = A.P;                                            \
    const int nx = L.npx, ny = L.npy;                                   \
    const int i = (int)(blockIdx.x * ABX + threadIdx.x);                \
    const int j = (int)(blockIdx.y * ABY + threadIdx.y);                \
    if (i >= nx || j >= ny) return;                                     \
    const int ij = j * nx + i;

/* apu = ap_u / relxm, apv = ap_v / relxm at the active faces, 1e30 at the
 * others (as the reference's apu/apv ghosts, V6:568-569) */
__global__ void k_adj_diag(const AdjLevel *As)
{
    ADJ_LEVEL
    double ap;
    if (urow(L, i, j)) {
        (void)tco_u(L, i, j, &ap);
        A.apu[ij] = ap / L.P.relxm;
    } else {
        A.apu[ij] = 1.0e30;
    }
    if (vrow(L, i, j)) {
        (void)tco_v(L, i, j, &ap);
        A.apv[ij] = ap / L.P.relxm;
    } else {
        A.apv[ij] = 1.0e30;
    }
}

/* one colour of red-black on the transposed scalar operator: phi = e_s,
 * right-hand side f_s */
__global__ void k_adj_rb_s(const AdjLevel *As, int colour)
{
    ADJ_LEVEL
    if (((i + j) & 1) != colour || !srow(L, i, j)) return;
    double ap;
    const Tco t = tco_s(L, i, j, &ap);
    const double *e = A.e[3];
    const double s = t.w * e[ij - 1] + t.e * e[ij + 1] + t.s * e[ij - nx] + t.n * e[ij + nx];
    A.e[3][ij] = (s + A.f[3][ij]) / ap;
}

/* right-hand sides of the u and v sweeps: f - D^T e_c - S^T e_s + the
 * under-relaxation term (1 - relxm) (ap / relxm) e (umom, V6:759-761) */
__global__ void k_adj_rhs_uv(const AdjLevel *As)
{
    ADJ_LEVEL
    const double relxm1 = L.P.relxm1;
    const double *ec = A.e[2], *es = A.e[3];
    if (urow(L, i, j)) {
        double dm, dp_;
        dsc((0.5 * L.u[ij]) * L.ax, L.ds_x, 0.5 * L.ax, &dm, &dp_);
        const double sP = L.sc[ij], sE = L.sc[ij + 1];
        const double st = (sP - sE) * (dm * es[ij]) + (sE - sP) * (dp_ * es[ij + 1]);
        const double dt = L.ax * (ec[ij] - ec[ij + 1]);
        A.r[0][ij] = A.f[0][ij] - dt - st + relxm1 * A.apu[ij] * A.e[0][ij];
    } else {
        A.r[0][ij] = 0.0;
    }
    if (vrow(L, i, j)) {
        double dm, dp_;
        dsc((0.5 * L.v[ij]) * L.ay, L.ds_y, 0.5 * L.ay, &dm, &dp_);
        const double sP = L.sc[ij], sN = L.sc[ij + nx];
        const double st = (sP - sN) * (dm * es[ij]) + (sN - sP) * (dp_ * es[ij + nx]);
        const double dt = L.ay * (ec[ij] - ec[ij + nx]);
        A.r[1][ij] = A.f[1][ij] - dt - st + relxm1 * A.apv[ij] * A.e[1][ij];
    } else {
        A.r[1][ij] = 0.0;
    }
}

/* one colour of red-black on the transposed u (field 0) or v (1) operator,
 * diagonal ap / relxm, right-hand side r */
__global__ void k_adj_rb_uv(const AdjLevel *As, int field, int colour)
{
    ADJ_LEVEL
    if (((i + j) & 1) != colour) return;
    if (field == 0 ? !urow(L, i, j) : !vrow(L, i, j)) return;
    const Tco t = field == 0 ? tco_u(L, i, j, nullptr) : tco_v(L, i, j, nullptr);
    const double *e = A.e[field];
    const double s = t.w * e[ij - 1] + t.e * e[ij + 1] + t.s * e[ij - nx] + t.n * e[ij + nx];
    A.e[field][ij] = (s + A.r[field][ij]) / (field == 0 ? A.apu[ij] : A.apv[ij]);
}

/* the continuity defect f_c - G^T e (G^T e = -(areax (e_u,P - e_u,W) +
 * areay (e_v,P - e_v,S))), and dp = 0 */
__global__ void k_adj_defect(const AdjLevel *As)
{
    ADJ_LEVEL
    A.dp[ij] = 0.0;
    if (!crow(L, i, j)) {
        A.r[2][ij] = 0.0;
        return;
    }
    const double *eu = A.e[0], *ev = A.e[1];
    A.r[2][ij] = A.f[2][ij] + (L.ax * (eu[ij] - eu[ij - 1]) + L.ay * (ev[ij] - ev[ij - nx]));
}

/* one colour of red-black on the pressure-correction operator of coefp
 * (V6:1030-1040: areax^2 / apu, areay^2 / apv, obstacle cells 1e30).  A
 * fluid cell all of whose faces are fixed (a pocket enclosed by obstacles,
 * on this level) has a diagonal of ~1e-33: its multiplier is not
 * determined by the equations (no active face carries it), and a sweep
 * there would amplify the defect by 1e33, so it keeps dp = 0. */
#define ADJ_FIXED_FACE 1.0e29
__global__ void k_adj_rb_p(const AdjLevel *As, int colour)
{
    ADJ_LEVEL
    if (((i + j) & 1) != colour || !crow(L, i, j)) return;
    if (A.apu[ij - 1] >= ADJ_FIXED_FACE && A.apu[ij] >= ADJ_FIXED_FACE &&
        A.apv[ij - nx] >= ADJ_FIXED_FACE && A.apv[ij] >= ADJ_FIXED_FACE)
        return;
    const double aw = L.axax / A.apu[ij - 1], ae = L.axax / A.apu[ij];
    const double as = L.ayay / A.apv[ij - nx], an = L.ayay / A.apv[ij];
    const double ap = ((aw + ae) + as) + an;
    const double *d = A.dp;
    const double s = aw * d[ij - 1] + ae * d[ij + 1] + as * d[ij - nx] + an * d[ij + nx];
    A.dp[ij] = (s + A.r[2][ij]) / ap;
}

/* S red-black sweeps of one multiplier (FIELD 0 u, 1 v, 2 the continuity
 * correction dp, 3 sc) in one launch, temporally blocked: a workgroup takes
 * an ATB_X x ATB_Y tile and a halo of H = 2S points into LDS, forms each
 * updatable point's transposed coefficients, right-hand side and diagonal
 * once, and colour step m (colour (m - 1) & 1, as the single-colour
 * launches run them) updates the points at distance >= m from the region's
 * edge; after 2S steps the tile holds what S sweeps of k_adj_rb_s /
 * k_adj_rb_uv / k_adj_rb_p give, bit for bit (same expressions).  Points
 * outside the level's arrays hold 0 and are only read by points that are
 * never updated.  The halo is other workgroups' tiles, so the launch reads
 * one buffer and writes the other (dir 0: the multiplier's array to tmp,
 * 1: back), every point of its tile, updated or not: the launches of a
 * relaxation come in pairs, and the second leaves the result in place. */
#define ATB_X 64
#define ATB_Y 16
#define ATB_NT 512
template <int FIELD, int S>
__global__ void __launch_bounds__(ATB_NT) k_adj_rbT(const AdjLevel *As, int dir)
{
    constexpr int H = 2 * S, RW = ATB_X + 2 * H, RH = ATB_Y + 2 * H, NR = RW * RH;
    constexpr int NPT = (NR + ATB_NT - 1) / ATB_NT;
    const AdjLevel &A = As[blockIdx.z];
    if (!A.live[A.inst]) return;
    const DevLevel &L = A.P;
    const int nx = L.npx, ny = L.npy;
    __shared__ double t[NR];
    const int ri0 = (int)blockIdx.x * ATB_X - H, rj0 = (int)blockIdx.y * ATB_Y - H;
    const int tid = (int)(threadIdx.y * ATB_X + threadIdx.x);
    double *const xf = FIELD == 2 ? A.dp : A.e[FIELD];
    const double *src = dir ? A.tmp : xf;
    double *dst = dir ? xf : A.tmp;
    const double *rhs = FIELD == 3 ? A.f[3] : A.r[FIELD];
    double cw[NPT], ce[NPT], cs[NPT], cn[NPT], rr[NPT], dg[NPT];
    unsigned meta[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int p = tid + k * ATB_NT;
        double val = 0.0;
        unsigned m = 0;
        cw[k] = ce[k] = cs[k] = cn[k] = rr[k] = 0.0;
        dg[k] = 1.0;
        if (p < NR) {
            const int r = p / RW, c = p - r * RW;
            const int gi = ri0 + c, gj = rj0 + r;
            if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
                const int g = gj * nx + gi;
                val = src[g];
                bool upd;
                if (FIELD == 0) upd = urow(L, gi, gj);
                else if (FIELD == 1) upd = vrow(L, gi, gj);
                else if (FIELD == 3) upd = srow(L, gi, gj);
                else
                    upd = crow(L, gi, gj) &&
                          !(A.apu[g - 1] >= ADJ_FIXED_FACE && A.apu[g] >= ADJ_FIXED_FACE &&
                            A.apv[g - nx] >= ADJ_FIXED_FACE && A.apv[g] >= ADJ_FIXED_FACE);
                if (upd) {
                    if (FIELD == 2) {
                        cw[k] = L.axax / A.apu[g - 1];
                        ce[k] = L.axax / A.apu[g];
                        cs[k] = L.ayay / A.apv[g - nx];
                        cn[k] = L.ayay / A.apv[g];
                        dg[k] = ((cw[k] + ce[k]) + cs[k]) + cn[k];
                    } else {
                        double ap;
                        const Tco tc = FIELD == 0 ? tco_u(L, gi, gj, nullptr)
                                                  : (FIELD == 1 ? tco_v(L, gi, gj, nullptr)
                                                                : tco_s(L, gi, gj, &ap));
                        cw[k] = tc.w;
                        ce[k] = tc.e;
                        cs[k] = tc.s;
                        cn[k] = tc.n;
                        dg[k] = FIELD == 0 ? A.apu[g] : (FIELD == 1 ? A.apv[g] : ap);
                    }
                    rr[k] = rhs[g];
                    m = (unsigned)((gi + gj) & 1) << 8 | (1u << 9);
                }
                /* bit 10: inside the level's arrays, written back */
                m |= (unsigned)min(min(c, RW - 1 - c), min(r, RH - 1 - r)) | (1u << 10);
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
                const int p = tid + k * ATB_NT;
                const double sum = cw[k] * t[p - 1] + ce[k] * t[p + 1] + cs[k] * t[p - RW] +
                                   cn[k] * t[p + RW];
                t[p] = (sum + rr[k]) / dg[k];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const unsigned m = meta[k];
        if ((m & (1u << 10)) && (int)(m & 0xffu) >= H) {
            const int p = tid + k * ATB_NT;
            const int r = p / RW, c = p - r * RW;
            dst[(rj0 + r) * nx + (ri0 + c)] = t[p];
        }
    }
}

/* the update: e_u -= areax (dp_P - dp_E) / apu, e_v likewise, e_c += relxp dp */
__global__ void k_adj_update(const AdjLevel *As)
{
    ADJ_LEVEL
    const double *d = A.dp;
    if (urow(L, i, j)) A.e[0][ij] -= L.ax * (d[ij] - d[ij + 1]) / A.apu[ij];
    if (vrow(L, i, j)) A.e[1][ij] -= L.ay * (d[ij] - d[ij + nx]) / A.apv[ij];
    if (crow(L, i, j)) A.e[2][ij] += L.P.relxp * d[ij];
}

/* r = f - L e with L the level's Picard operator (transposed hybrid
 * coefficients, the multipliers' couplings D^T, G^T and S^T) */
__global__ void k_adj_resid(const AdjLevel *As)
{
    ADJ_LEVEL
    const double *eu = A.e[0], *ev = A.e[1], *ec = A.e[2], *es = A.e[3];
    double ap;
    if (urow(L, i, j)) {
        const Tco t = tco_u(L, i, j, &ap);
        double dm, dp_;
        dsc((0.5 * L.u[ij]) * L.ax, L.ds_x, 0.5 * L.ax, &dm, &dp_);
        const double sP = L.sc[ij], sE = L.sc[ij + 1];
        const double lu = ap * eu[ij] - t.w * eu[ij - 1] - t.e * eu[ij + 1] - t.s * eu[ij - nx] -
                          t.n * eu[ij + nx] + L.ax * (ec[ij] - ec[ij + 1]) +
                          ((sP - sE) * (dm * es[ij]) + (sE - sP) * (dp_ * es[ij + 1]));
        A.r[0][ij] = A.f[0][ij] - lu;
    } else {
        A.r[0][ij] = 0.0;
    }
    if (vrow(L, i, j)) {
        const Tco t = tco_v(L, i, j, &ap);
        double dm, dp_;
        dsc((0.5 * L.v[ij]) * L.ay, L.ds_y, 0.5 * L.ay, &dm, &dp_);
        const double sP = L.sc[ij], sN = L.sc[ij + nx];
        const double lv = ap * ev[ij] - t.w * ev[ij - 1] - t.e * ev[ij + 1] - t.s * ev[ij - nx] -
                          t.n * ev[ij + nx] + L.ay * (ec[ij] - ec[ij + nx]) +
                          ((sP - sN) * (dm * es[ij]) + (sN - sP) * (dp_ * es[ij + nx]));
        A.r[1][ij] = A.f[1][ij] - lv;
    } else {
        A.r[1][ij] = 0.0;
    }
    if (crow(L, i, j)) {
        const double gt = -(L.ax * (eu[ij] - eu[ij - 1])) - L.ay * (ev[ij] - ev[ij - nx]);
        A.r[2][ij] = A.f[2][ij] - gt;
    } else {
        A.r[2][ij] = 0.0;
    }
    if (srow(L, i, j)) {
        const Tco t = tco_s(L, i, j, &ap);
        const double ls = ap * es[ij] - t.w * es[ij - 1] - t.e * es[ij + 1] - t.s * es[ij - nx] -
                          t.n * es[ij + nx];
        A.r[3][ij] = A.f[3][ij] - ls;
    } else {
        A.r[3][ij] = 0.0;
    }
}

/* coarse right-hand sides from the fine residuals (restru / restrv's
 * weights, V6:1547-1560 and 1723-1736; cell sums for continuity and the
 * scalar) and a zero coarse correction */
__global__ void k_adj_restrict(const AdjLevel *Fs, const AdjLevel *Cs)
{
    const AdjLevel &F = Fs[blockIdx.z], &C = Cs[blockIdx.z];
    if (!C.live[C.inst]) return;
    const DevLevel &Lc = C.P;
    const int nxc = Lc.npx, nyc = Lc.npy, nxf = F.P.npx;
    const int ic = (int)(blockIdx.x * ABX + threadIdx.x), jc = (int)(blockIdx.y * ABY + threadIdx.y);
    if (ic >= nxc || jc >= nyc) return;
    const int ijc = jc * nxc + ic;
    for (int k = 0; k < 4; ++k) C.e[k][ijc] = 0.0;
    double fu = 0.0, fv = 0.0, fc = 0.0, fs = 0.0;
    if (ic >= 1 && ic <= nxc - 2 && jc >= 1 && jc <= nyc - 2) {
        const int ij = (2 * jc) * nxf + 2 * ic;
        const double *ru = F.r[0], *rv = F.r[1], *rc = F.r[2], *rs = F.r[3];
        if (urow(Lc, ic, jc))
            fu = (ru[ij] + ru[ij - nxf]) +
                 0.5 * (ru[ij - 1] + ru[ij - 1 - nxf] + ru[ij + 1] + ru[ij + 1 - nxf]);
        if (vrow(Lc, ic, jc))
            fv = (rv[ij] + rv[ij - 1]) +
                 0.5 * (rv[ij - nxf] + rv[ij - nxf - 1] + rv[ij + nxf] + rv[ij + nxf - 1]);
        if (crow(Lc, ic, jc)) fc = ((rc[ij] + rc[ij - 1]) + rc[ij - nxf]) + rc[ij - nxf - 1];
        if (srow(Lc, ic, jc)) fs = ((rs[ij] + rs[ij - 1]) + rs[ij - nxf]) + rs[ij - nxf - 1];
    }
    C.f[0][ijc] = fu;
    C.f[1][ijc] = fv;
    C.f[2][ijc] = fc;
    C.f[3][ijc] = fs;
}

/* fine corrections += the interpolated coarse ones (gather form of prolu,
 * prolv: 3/4 - 1/4 along the face, halves on the odd faces; prolp / prolsc:
 * 9/16, 3/16, 3/16, 1/16), at the fine level's active points */
__global__ void k_adj_prolong(const AdjLevel *Fs, const AdjLevel *Cs)
{
    const AdjLevel &F = Fs[blockIdx.z], &C = Cs[blockIdx.z];
    if (!C.live[C.inst]) return;
    const DevLevel &Lf = F.P;
    const int nxf = Lf.npx, nyf = Lf.npy, nxc = C.P.npx, nyc = C.P.npy;
    const int i = (int)(blockIdx.x * ABX + threadIdx.x), j = (int)(blockIdx.y * ABY + threadIdx.y);
    if (i >= nxf || j >= nyf) return;
    const int ij = j * nxf + i;
    /* coarse value with zero outside the coarse interior ranges */
    auto cu = [&](int a, int b) {
        return (a >= 1 && a <= nxc - 3 && b >= 1 && b <= nyc - 2) ? C.e[0][b * nxc + a] : 0.0;
    };
    auto cv = [&](int a, int b) {
        return (a >= 1 && a <= nxc - 2 && b >= 1 && b <= nyc - 3) ? C.e[1][b * nxc + a] : 0.0;
    };
    auto cc = [&](int k, int a, int b) {
        return (a >= 1 && a <= nxc - 2 && b >= 1 && b <= nyc - 2) ? C.e[k][b * nxc + a] : 0.0;
    };
    const double f1 = 0.75, f2 = 0.25;
    if (urow(Lf, i, j)) {
        /* fine face i = 2 ic (+-1 for the odd faces), row j = 2 jc or 2 jc - 1 */
        const int jc = (j + 1) / 2;
        const bool upper = (j & 1) == 0;       /* j = 2 jc */
        auto corr = [&](int ic) {
            return upper ? f1 * cu(ic, jc) + f2 * cu(ic, jc + 1)
                         : f1 * cu(ic, jc) + f2 * cu(ic, jc - 1);
        };
        double add;
        if ((i & 1) == 0)
            add = corr(i / 2);
        else
            add = 0.5 * corr((i - 1) / 2) + 0.5 * corr((i + 1) / 2);
        F.e[0][ij] += add;
    }
    if (vrow(Lf, i, j)) {
        const int ic = (i + 1) / 2;
        const bool right = (i & 1) == 0;       /* i = 2 ic */
        auto corr = [&](int jc) {
            return right ? f1 * cv(ic, jc) + f2 * cv(ic + 1, jc)
                         : f1 * cv(ic, jc) + f2 * cv(ic - 1, jc);
        };
        double add;
        if ((j & 1) == 0)
            add = corr(j / 2);
        else
            add = 0.5 * corr((j - 1) / 2) + 0.5 * corr((j + 1) / 2);
        F.e[1][ij] += add;
    }
    if (i >= 1 && i <= nxf - 2 && j >= 1 && j <= nyf - 2) {
        const int ic = (i + 1) / 2, jc = (j + 1) / 2;
        const int di = (i & 1) == 0 ? 1 : -1, dj = (j & 1) == 0 ? 1 : -1;
        for (int k = 2; k < 4; ++k) {
            if (k == 2 ? !crow(Lf, i, j) : !srow(Lf, i, j)) continue;
            F.e[k][ij] += (9.0 / 16.0) * cc(k, ic, jc) + (3.0 / 16.0) * cc(k, ic + di, jc) +
                          (3.0 / 16.0) * cc(k, ic, jc + dj) +
                          (1.0 / 16.0) * cc(k, ic + di, jc + dj);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* level-0 vectors, right-hand side, gradient                               */
/* ---------------------------------------------------------------------- */

/* dJ/dx of the objective into G (layout u, v, c, s) and J */
__global__ void k_adj_objective(const AdjLevel *As, const AdjVec *Vs, int objective,
                                double *J)
{
    const AdjLevel &A = As[blockIdx.x];
    const AdjVec &W = Vs[A.inst];
    const DevLevel &L = A.P;
    const int nx = L.npx, ny = L.npy, n = nx * ny, nfy = ny - 2;
    double *G = W.G;
    for (long k = threadIdx.x; k < W.n4; k += blockDim.x) G[k] = 0.0;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double j1a = 0.0, j1b = 0.0, j2 = 0.0;
    for (int j = 1; j <= ny - 2; ++j) {
        const int i0 = j * nx + 1, i1 = j * nx + nx - 2;
        j1a += L.p[i0];
        j1b += L.p[i1];
        j2 += (L.ax * L.u[i1]) * L.sc[i1];
        if (objective == 0) {
            if (crow(L, 1, j)) G[2 * n + i0] += 1.0 / (double)nfy;
            if (crow(L, nx - 2, j)) G[2 * n + i1] -= 1.0 / (double)nfy;
        } else if (srow(L, nx - 2, j)) {
            G[3 * n + i1] = L.ax * L.u[i1];
        }
    }
    J[2 * A.inst] = j1a / (double)nfy - j1b / (double)nfy;
    J[2 * A.inst + 1] = j2;
}

/* partial sums of -lambda^T dR/d(amuf) and -lambda^T dR/d(prsc) per block */
__global__ void k_adj_grad_par(const AdjLevel *As, const MetSeed *seeds, const double *const *lamv,
                               double *partial, int nblk_per)
{
    const AdjLevel &A = As[blockIdx.z];
    const DevLevel &L = A.P;
    const int nx = L.npx, ny = L.npy;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    const int i = (int)(blockIdx.x * ABX + threadIdx.x), j = (int)(blockIdx.y * ABY + threadIdx.y);
    double ga = 0.0, gp = 0.0;
    if (i < nx && j < ny) {
        const double *lam[4];
        for (int k = 0; k < 4; ++k) lam[k] = lamv[4 * A.inst + k];
        State<mgf_dual> x{L.u, L.v, L.p, L.sc, nx, Seed{-1, 0, 0}};
        const Met<mgf_dual> ma = metrics<mgf_dual>(L, &seeds[2 * A.inst]);
        const Met<mgf_dual> mp = metrics<mgf_dual>(L, &seeds[2 * A.inst + 1]);
        const int ij = j * nx + i;
        if (urow(L, i, j)) ga -= lam[0][ij] * res_u(L, ma, x, i, j).d;
        if (vrow(L, i, j)) ga -= lam[1][ij] * res_v(L, ma, x, i, j).d;
        if (srow(L, i, j)) {
            ga -= lam[3][ij] * res_s(L, ma, x, i, j).d;
            gp -= lam[3][ij] * res_s(L, mp, x, i, j).d;
        }
    }
    __shared__ double red[2][ABX * ABY];
    const int t = threadIdx.y * ABX + threadIdx.x;
    red[0][t] = ga;
    red[1][t] = gp;
    __syncthreads();
    for (int s = ABX * ABY / 2; s > 0; s >>= 1) {
        if (t < s) {
            red[0][t] += red[0][t + s];
            red[1][t] += red[1][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        partial[((size_t)A.inst * nblk_per + blk) * 2] = red[0][0];
        partial[((size_t)A.inst * nblk_per + blk) * 2 + 1] = red[1][0];
    }
}

/* -lambda^T dR/d(inlet row j) for every row j (one thread per row), and
 * the east u's -lambda^T dR/d(u_xp) + dJ/d(u_xp) (thread of row 0) */
__global__ void k_adj_grad_bc(const AdjLevel *As, const double *const *lamv, int objective,
                              double *const *gin, double *guxp)
{
    const AdjLevel &A = As[blockIdx.z];
    const DevLevel &L = A.P;
    const int nx = L.npx, ny = L.npy;
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j > ny - 2) return;
    const double *lam[4];
    for (int k = 0; k < 4; ++k) lam[k] = lamv[4 * A.inst + k];
    const Met<mgf_dual> m = metrics<mgf_dual>(L, nullptr);
    State<mgf_dual> x{L.u, L.v, L.p, L.sc, nx, Seed{-1, 0, 0}};
    if (j >= 1) {
        x.sd = Seed{0, j * nx, 0};
        const double g = drow(L, m, x, lam, 0, 1, j) + drow(L, m, x, lam, 1, 1, j) +
                         drow(L, m, x, lam, 1, 1, j - 1) + drow(L, m, x, lam, 2, 1, j) +
                         drow(L, m, x, lam, 3, 1, j);
        gin[A.inst][j - 1] = -g;
        return;
    }
    /* j == 0: the east boundary u, every row at once */
    x.sd = Seed{-1, 0, 1};
    double g = 0.0, dj = 0.0;
    for (int r = 1; r <= ny - 2; ++r) {
        g += drow(L, m, x, lam, 0, nx - 3, r) + drow(L, m, x, lam, 1, nx - 2, r) +
             drow(L, m, x, lam, 2, nx - 2, r) + drow(L, m, x, lam, 3, nx - 2, r);
        if (objective == 1) dj += L.ax * L.sc[r * nx + nx - 2];
    }
    guxp[A.inst] = dj - g;
}

/* ---- Krylov vector kernels (per problem: blockIdx.z) ------------------ */

/* Dot products of long vectors: every workgroup of DOT_T threads takes a
 * chunk of DOT_T * DOT_PER entries (thread t: entries t, t + DOT_T, ...),
 * writes one partial sum per product, and k_dots_sum adds a product's
 * partials.  Every sum runs in a fixed order (a thread's entries in turn, a
 * wave's shuffle tree, the four waves in turn, then the same for the
 * partials), so the results do not depend on scheduling. */
#define DOT_T 256
#define DOT_PER 8
/* Krylov vectors per workgroup of k_dots: W is read once per DOT_KB of them */
#define DOT_KB 16

/* the sums over the workgroup of v[0..NV-1], in thread i < NV of the
 * returned LDS row out (red: 4 x NV doubles) */
template <int NV>
__device__ __forceinline__ void wg_sums(double (&v)[NV], double *red)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i)
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_down(v[i], off, 64);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) red[wv * NV + i] = v[i];
    __syncthreads();
}

/* partials of <V_k, W> for k in [blockIdx.y * DOT_KB, +DOT_KB) and k < nk,
 * and (wnorm, blockIdx.y 0) of <W, W> from the same read of W; partial
 * layout ((b * nk + k) * nblk + chunk), <W, W> in wpart (b * nblk + chunk) */
__global__ void __launch_bounds__(DOT_T) k_dots(const AdjVec *Vs, const int *live, int nk,
                                                int wnorm, double *partial, double *wpart,
                                                int nblk)
{
    const AdjVec &W = Vs[blockIdx.z];
    if (!live[blockIdx.z]) return;
    __shared__ double red[4 * (DOT_KB + 1)];
    const long base = (long)blockIdx.x * DOT_T * DOT_PER + threadIdx.x;
    const int k0 = (int)blockIdx.y * DOT_KB;
    double w[DOT_PER];
#pragma unroll
    for (int t = 0; t < DOT_PER; ++t) {
        const long q = base + (long)t * DOT_T;
        w[t] = q < W.n4 ? W.W[q] : 0.0;
    }
    double acc[DOT_KB + 1];
#pragma unroll
    for (int i = 0; i <= DOT_KB; ++i) acc[i] = 0.0;
#pragma unroll
    for (int kk = 0; kk < DOT_KB; ++kk) {
        if (k0 + kk < nk) {
            const double *a = W.V + (size_t)(k0 + kk) * W.n4;
#pragma unroll
            for (int t = 0; t < DOT_PER; ++t) {
                const long q = base + (long)t * DOT_T;
                if (q < W.n4) acc[kk] += a[q] * w[t];
            }
        }
    }
    const bool wn = wnorm && blockIdx.y == 0;
    if (wn)
#pragma unroll
        for (int t = 0; t < DOT_PER; ++t) acc[DOT_KB] += w[t] * w[t];
    wg_sums<DOT_KB + 1>(acc, red);
    const int i = threadIdx.x;
    if (i < DOT_KB && k0 + i < nk) {
        const double v = ((red[i] + red[(DOT_KB + 1) + i]) + red[2 * (DOT_KB + 1) + i]) +
                         red[3 * (DOT_KB + 1) + i];
        partial[((size_t)blockIdx.z * nk + k0 + i) * nblk + blockIdx.x] = v;
    }
    if (wn && i == DOT_KB) {
        const double v = ((red[i] + red[(DOT_KB + 1) + i]) + red[2 * (DOT_KB + 1) + i]) +
                         red[3 * (DOT_KB + 1) + i];
        wpart[(size_t)blockIdx.z * nblk + blockIdx.x] = v;
    }
}

/* partials of <X, X> (which 0: W, 1: G), layout (b * nblk + chunk) */
__global__ void __launch_bounds__(DOT_T) k_norm2(const AdjVec *Vs, const int *live, int which,
                                                 double *partial, int nblk)
{
    const AdjVec &W = Vs[blockIdx.z];
    if (!live[blockIdx.z]) return;
    __shared__ double red[4];
    const double *x = which == 0 ? W.W : W.G;
    const long base = (long)blockIdx.x * DOT_T * DOT_PER + threadIdx.x;
    double acc[1] = {0.0};
#pragma unroll
    for (int t = 0; t < DOT_PER; ++t) {
        const long q = base + (long)t * DOT_T;
        if (q < W.n4) acc[0] += x[q] * x[q];
    }
    wg_sums<1>(acc, red);
    if (threadIdx.x == 0)
        partial[(size_t)blockIdx.z * nblk + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

/* out[b * nk + k] = the sum of product k's partials of problem b (grid
 * (nk, nb)); 0 for a problem that is not live */
__global__ void __launch_bounds__(DOT_T) k_dots_sum(const double *partial, const int *live, int nk,
                                                    int nblk, double *out)
{
    const int k = blockIdx.x, b = blockIdx.y;
    if (!live[b]) {
        if (threadIdx.x == 0) out[(size_t)b * nk + k] = 0.0;
        return;
    }
    __shared__ double red[4];
    const double *p = partial + ((size_t)b * nk + k) * nblk;
    double s[1] = {0.0};
    for (int q = threadIdx.x; q < nblk; q += DOT_T) s[0] += p[q];
    wg_sums<1>(s, red);
    if (threadIdx.x == 0) out[(size_t)b * nk + k] = ((red[0] + red[1]) + red[2]) + red[3];
}

/* W -= sum_k h[b, k] V_k (k < nk), and the partials of <W, W> of the result
 * (layout b * nblk + chunk) from the same pass */
__global__ void __launch_bounds__(DOT_T) k_orth(const AdjVec *Vs, const int *live, int nk,
                                                const double *h, int hstride, double *wpart,
                                                int nblk)
{
    const AdjVec &W = Vs[blockIdx.z];
    if (!live[blockIdx.z]) return;
    __shared__ double red[4];
    const double *hb = h + (size_t)blockIdx.z * hstride;
    const long base = (long)blockIdx.x * DOT_T * DOT_PER + threadIdx.x;
    /* the thread's entries in registers, one coefficient and one load per
     * entry for each Krylov vector (each entry's subtractions in k order) */
    double x[DOT_PER];
#pragma unroll
    for (int t = 0; t < DOT_PER; ++t) {
        const long q = base + (long)t * DOT_T;
        x[t] = q < W.n4 ? W.W[q] : 0.0;
    }
    for (int k = 0; k < nk; ++k) {
        const double hk = hb[k];
        const double *v = W.V + (size_t)k * W.n4;
#pragma unroll
        for (int t = 0; t < DOT_PER; ++t) {
            const long q = base + (long)t * DOT_T;
            if (q < W.n4) x[t] -= hk * v[q];
        }
    }
    double acc[1] = {0.0};
#pragma unroll
    for (int t = 0; t < DOT_PER; ++t) {
        const long q = base + (long)t * DOT_T;
        if (q < W.n4) {
            W.W[q] = x[t];
            acc[0] += x[t] * x[t];
        }
    }
    wg_sums<1>(acc, red);
    if (threadIdx.x == 0)
        wpart[(size_t)blockIdx.z * nblk + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

/* dst = c[b] * src with src/dst picked by sel: 0 V_j = W / h, 1 W = G - W (residual) */
__global__ void k_vec_op(const AdjVec *Vs, const int *live, int op, int jv, const double *c)
{
    const AdjVec &W = Vs[blockIdx.z];
    if (!live[blockIdx.z]) return;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= W.n4) return;
    if (op == 0) W.V[(size_t)jv * W.n4 + q] = W.W[q] * c[blockIdx.z];
    else if (op == 1) W.W[q] = W.G[q] - W.W[q];
    else if (op == 2) W.X[q] += W.W[q];
}

/* the level-0 right-hand side f of problem b's V-cycle: V_j (op 0) or
 * sum_k y[b, k] V_k (op 1), and the result of its V-cycle e to W (op 2) */
__global__ void k_f0(const AdjLevel *As, const AdjVec *Vs, int op, int jv, int nk, const double *y,
                     int ystride)
{
    const AdjLevel &A = As[blockIdx.z];
    if (!A.live[A.inst]) return;
    const AdjVec &W = Vs[A.inst];
    const long n = (long)A.P.npx * A.P.npy;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= W.n4) return;
    const int k = (int)(q / n);
    const long o = q - (long)k * n;
    if (op == 0) {
        A.f[k][o] = W.V[(size_t)jv * W.n4 + q];
    } else if (op == 1) {
        double acc = 0.0;
        for (int t = 0; t < nk; ++t) acc += y[(size_t)A.inst * ystride + t] * W.V[(size_t)t * W.n4 + q];
        A.f[k][o] = acc;
    } else {
        W.W[q] = A.e[k][o];
    }
}

/* a zero correction on the finest levels before a V-cycle */
__global__ void k_adj_zero_e(const AdjLevel *As)
{
    ADJ_LEVEL
    for (int k = 0; k < 4; ++k) A.e[k][ij] = 0.0;
}

/* ---------------------------------------------------------------------- */
/* host side                                                                */
/* ---------------------------------------------------------------------- */

/* internal accessors of the solver (mgf_solver.hip) */
const DevLevel *mgf_instance_levels(mgf_solver *s, int b, int *ng);
void **mgf_adjoint_slot(mgf_solver *s);

struct AdjWork {
    int nb = 0, restart = 0;
    int single_colour = 0;                          /* sweeps as single-colour launches */
    int pc_cycles = 1;                              /* V-cycles per preconditioner use */
    std::vector<int> ng;
    std::vector<std::vector<AdjLevel>> lv;          /* host copies per problem and level */
    std::vector<AdjVec> vec;
    std::vector<double *> blocks;                   /* device allocations (freed at release) */
    std::vector<AdjGroup> groups;                   /* per depth (may hold several shapes) */
    std::vector<std::vector<int>> depth_groups;     /* depth -> group indices */
    std::vector<AdjPair> pairs;
    std::vector<std::vector<int>> depth_pairs;      /* depth (of the fine level) -> pairs */
    AdjLevel *d_desc = nullptr;
    AdjVec *d_vec = nullptr;
    int *d_live = nullptr;
    std::vector<int> h_live;
    double **d_lam = nullptr;                       /* 4 per problem: level-0 lambda = X */
    double **d_kin = nullptr, **d_kout = nullptr;   /* K^T operands: 4 per problem */
    double *d_small = nullptr;                      /* h, y, J, seeds scratch */
    double *d_partial = nullptr;
    size_t partial_cap = 0;
    double **d_gin = nullptr;
    MetSeed *d_seeds = nullptr;
    int D = 0;
    AdjGroup g0;                                    /* all level-0 members (for K^T, vectors) */
    std::vector<AdjGroup> g0s;                      /* level-0 members grouped by shape */
    std::vector<int> last_iters;
    std::vector<double> last_relres;
    int maxn4 = 0;
};

void mgf_adjoint_release(void *p)
{
    AdjWork *w = (AdjWork *)p;
    if (!w) return;
    for (double *b : w->blocks) (void)hipFree(b);
    delete w;
}

#define AHIP(s, call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            char m_[256];                                                              \
            snprintf(m_, sizeof m_, "%s failed: %s", #call, hipGetErrorString(e_));    \
            return mgf_set_error((s), MGF_ERR_DEVICE, m_);                             \
        }                                                                              \
    } while (0)

static double *dalloc(AdjWork *w, size_t n)
{
    double *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
    w->blocks.push_back(p);
    return p;
}

static int build_work(mgf_solver *s, int restart, AdjWork **out)
{
    AdjWork *w = new AdjWork();
    const int nb = mgf_batch_size(s);
    w->nb = nb;
    w->restart = restart;
    w->ng.resize(nb);
    w->lv.resize(nb);
    w->vec.resize(nb);
    int D = 0;
    hipStream_t st = (hipStream_t)mgf_stream(s);
    for (int b = 0; b < nb; ++b) {
        int ng = 0;
        const DevLevel *lv = mgf_instance_levels(s, b, &ng);
        w->ng[b] = ng;
        D = std::max(D, ng - 1);
        w->lv[b].resize(ng);
        for (int ig = 0; ig < ng; ++ig) {
            AdjLevel &A = w->lv[b][ig];
            memset(&A, 0, sizeof A);
            A.P = lv[ig];
            A.inst = b;
            const size_t n = (size_t)lv[ig].npx * lv[ig].npy;
            double *blk = dalloc(w, 16 * n);
            if (!blk) {
                mgf_adjoint_release(w);
                return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint workspace allocation failed");
            }
            if (hipMemsetAsync(blk, 0, 16 * n * sizeof(double), st) != hipSuccess) {
                mgf_adjoint_release(w);
                return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint workspace clear failed");
            }
            for (int k = 0; k < 4; ++k) {
                A.e[k] = blk + k * n;
                A.f[k] = blk + (4 + k) * n;
                A.r[k] = blk + (8 + k) * n;
            }
            A.dp = blk + 12 * n;
            A.apu = blk + 13 * n;
            A.apv = blk + 14 * n;
            A.tmp = blk + 15 * n;
        }
        const long n4 = 4L * lv[0].npx * lv[0].npy;
        w->maxn4 = std::max<long>(w->maxn4, n4);
        double *vb = dalloc(w, (size_t)(restart + 4) * n4);
        if (!vb) {
            mgf_adjoint_release(w);
            return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint Krylov space allocation failed");
        }
        AdjVec &V = w->vec[b];
        V.V = vb;
        V.X = vb + (size_t)(restart + 1) * n4;
        V.G = vb + (size_t)(restart + 2) * n4;
        V.W = vb + (size_t)(restart + 3) * n4;
        V.n4 = n4;
    }
    w->D = D;
    if (hipMalloc(&w->d_live, nb * sizeof(int)) != hipSuccess) {
        mgf_adjoint_release(w);
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint flag allocation failed");
    }
    w->blocks.push_back((double *)w->d_live);
    for (int b = 0; b < nb; ++b)
        for (AdjLevel &A : w->lv[b]) A.live = w->d_live;
    /* groups: same depth and shape; pairs: a level at depth d >= 1 with its
     * coarser level */
    std::vector<std::vector<AdjLevel>> gm, pf, pc;
    std::vector<int> gdepth, pdepth;
    w->depth_groups.assign(D + 1, {});
    w->depth_pairs.assign(D + 1, {});
    for (int d = D; d >= 0; --d)
        for (int b = 0; b < nb; ++b) {
            const int ig = w->ng[b] - 1 - d;
            if (ig < 0) continue;
            const AdjLevel &A = w->lv[b][ig];
            int g = -1;
            for (int q : w->depth_groups[d])
                if (gm[q][0].P.npx == A.P.npx && gm[q][0].P.npy == A.P.npy) g = q;
            if (g < 0) {
                g = (int)gm.size();
                gm.emplace_back();
                gdepth.push_back(d);
                w->depth_groups[d].push_back(g);
            }
            gm[g].push_back(A);
            if (d >= 1) {
                int p = -1;
                for (int q : w->depth_pairs[d])
                    if (pf[q][0].P.npx == A.P.npx && pf[q][0].P.npy == A.P.npy) p = q;
                if (p < 0) {
                    p = (int)pf.size();
                    pf.emplace_back();
                    pc.emplace_back();
                    w->depth_pairs[d].push_back(p);
                }
                pf[p].push_back(A);
                pc[p].push_back(w->lv[b][ig + 1]);
            }
        }
    /* level-0 members by shape */
    std::vector<std::vector<AdjLevel>> z;
    for (int b = 0; b < nb; ++b) {
        const AdjLevel &A = w->lv[b][0];
        int g = -1;
        for (size_t q = 0; q < z.size(); ++q)
            if (z[q][0].P.npx == A.P.npx && z[q][0].P.npy == A.P.npy) g = (int)q;
        if (g < 0) {
            g = (int)z.size();
            z.emplace_back();
        }
        z[g].push_back(A);
    }
    size_t nd = 0;
    for (auto &g : gm) nd += g.size();
    for (auto &p : pf) nd += 2 * p.size();
    for (auto &g : z) nd += g.size();
    if (hipMalloc(&w->d_desc, nd * sizeof(AdjLevel)) != hipSuccess) {
        mgf_adjoint_release(w);
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint descriptor allocation failed");
    }
    w->blocks.push_back((double *)w->d_desc);
    std::vector<AdjLevel> host;
    host.reserve(nd);
    for (size_t g = 0; g < gm.size(); ++g) {
        AdjGroup G;
        G.d = w->d_desc + host.size();
        G.count = (int)gm[g].size();
        G.npx = gm[g][0].P.npx;
        G.npy = gm[g][0].P.npy;
        G.depth = gdepth[g];
        host.insert(host.end(), gm[g].begin(), gm[g].end());
        w->groups.push_back(G);
    }
    for (size_t p = 0; p < pf.size(); ++p) {
        AdjPair R;
        R.count = (int)pf[p].size();
        R.npxf = pf[p][0].P.npx;
        R.npyf = pf[p][0].P.npy;
        R.npxc = pc[p][0].P.npx;
        R.npyc = pc[p][0].P.npy;
        R.f = w->d_desc + host.size();
        host.insert(host.end(), pf[p].begin(), pf[p].end());
        R.c = w->d_desc + host.size();
        host.insert(host.end(), pc[p].begin(), pc[p].end());
        for (const AdjLevel &c : pc[p])
            if (c.P.npx != R.npxc || c.P.npy != R.npyc) {
                mgf_adjoint_release(w);
                return mgf_set_error(s, MGF_ERR_GEOMETRY, "coarse shapes differ within a pair");
            }
        w->pairs.push_back(R);
    }
    for (auto &g : z) {
        AdjGroup G;
        G.d = w->d_desc + host.size();
        G.count = (int)g.size();
        G.npx = g[0].P.npx;
        G.npy = g[0].P.npy;
        G.depth = -1;
        host.insert(host.end(), g.begin(), g.end());
        w->g0s.push_back(G);
    }
    if (hipMemcpy(w->d_desc, host.data(), host.size() * sizeof(AdjLevel), hipMemcpyHostToDevice) !=
        hipSuccess) {
        mgf_adjoint_release(w);
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint descriptor upload failed");
    }
    /* device arrays of pointers and vectors */
    size_t bytes = sizeof(AdjVec) * nb;
    if (hipMalloc(&w->d_vec, bytes) != hipSuccess ||
        hipMemcpy(w->d_vec, w->vec.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        mgf_adjoint_release(w);
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint vector table failed");
    }
    w->blocks.push_back((double *)w->d_vec);
    std::vector<double *> lam(4 * nb);
    for (int b = 0; b < nb; ++b) {
        const long n = w->vec[b].n4 / 4;
        for (int k = 0; k < 4; ++k) lam[4 * b + k] = w->vec[b].X + k * n;
    }
    double ***pp[3] = {&w->d_lam, &w->d_kin, &w->d_kout};
    for (int q = 0; q < 3; ++q) {
        if (hipMalloc(pp[q], 4 * nb * sizeof(double *)) != hipSuccess) {
            mgf_adjoint_release(w);
            return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint pointer table failed");
        }
        w->blocks.push_back((double *)*pp[q]);
    }
    if (hipMemcpy(w->d_lam, lam.data(), 4 * nb * sizeof(double *), hipMemcpyHostToDevice) !=
        hipSuccess) {
        mgf_adjoint_release(w);
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint pointer upload failed");
    }
    const size_t small = (size_t)nb * (3 * (restart + 2) + 12) + 16;
    w->d_small = dalloc(w, small);
    if (hipMalloc(&w->d_gin, nb * sizeof(double *)) != hipSuccess ||
        hipMalloc(&w->d_seeds, 2 * nb * sizeof(MetSeed)) != hipSuccess || !w->d_small) {
        mgf_adjoint_release(w);
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint scratch allocation failed");
    }
    w->blocks.push_back((double *)w->d_gin);
    w->blocks.push_back((double *)w->d_seeds);
    w->h_live.assign(nb, 1);
    *out = w;
    return MGF_OK;
}

static int ensure_partial(mgf_solver *s, AdjWork *w, size_t n)
{
    if (n <= w->partial_cap) return MGF_OK;
    if (w->d_partial) {
        for (auto it = w->blocks.begin(); it != w->blocks.end(); ++it)
            if (*it == w->d_partial) {
                w->blocks.erase(it);
                break;
            }
        (void)hipFree(w->d_partial);
        w->d_partial = nullptr;
    }
    w->d_partial = dalloc(w, n);
    if (!w->d_partial) return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint partial sums allocation failed");
    w->partial_cap = n;
    return MGF_OK;
}

static dim3 agrid(int nx, int ny, int nz)
{
    return dim3((unsigned)((nx + ABX - 1) / ABX), (unsigned)((ny + ABY - 1) / ABY), (unsigned)nz);
}
static const dim3 ABLK(ABX, ABY);

/* one adjoint SIMPLE relaxation of a group */
/* nsw red-black sweeps of one multiplier: an even number of launches of up
 * to three sweeps each (k_adj_rbT, to tmp and back), or single colours in
 * place (one sweep, or single) */
template <int FIELD>
static void adj_sweeps(const AdjGroup &G, int nsw, bool single, hipStream_t st)
{
    const dim3 g = agrid(G.npx, G.npy, G.count);
    if (single || nsw < 2) {
        for (int k = 0; k < nsw; ++k)
            for (int colour = 0; colour < 2; ++colour) {
                if (FIELD == 3) hipLaunchKernelGGL(k_adj_rb_s, g, ABLK, 0, st, G.d, colour);
                else if (FIELD == 2) hipLaunchKernelGGL(k_adj_rb_p, g, ABLK, 0, st, G.d, colour);
                else hipLaunchKernelGGL(k_adj_rb_uv, g, ABLK, 0, st, G.d, FIELD, colour);
            }
        return;
    }
    const dim3 gt((unsigned)((G.npx + ATB_X - 1) / ATB_X), (unsigned)((G.npy + ATB_Y - 1) / ATB_Y),
                  (unsigned)G.count);
    const dim3 bt(ATB_X, ATB_NT / ATB_X);
    const int nl = 2 * ((nsw + 5) / 6);
    for (int k = 0; k < nl; ++k) {
        const int n = nsw / nl + (k < nsw % nl ? 1 : 0);
        const int dir = k & 1;
        if (n == 3) hipLaunchKernelGGL((k_adj_rbT<FIELD, 3>), gt, bt, 0, st, G.d, dir);
        else if (n == 2) hipLaunchKernelGGL((k_adj_rbT<FIELD, 2>), gt, bt, 0, st, G.d, dir);
        else hipLaunchKernelGGL((k_adj_rbT<FIELD, 1>), gt, bt, 0, st, G.d, dir);
    }
}

static void relax_group(const AdjGroup &G, int nswm, int nswp, bool single, hipStream_t st)
{
    const dim3 g = agrid(G.npx, G.npy, G.count);
    adj_sweeps<3>(G, nswm, single, st);
    hipLaunchKernelGGL(k_adj_rhs_uv, g, ABLK, 0, st, G.d);
    adj_sweeps<0>(G, nswm, single, st);
    adj_sweeps<1>(G, nswm, single, st);
    hipLaunchKernelGGL(k_adj_defect, g, ABLK, 0, st, G.d);
    adj_sweeps<2>(G, nswp, single, st);
    hipLaunchKernelGGL(k_adj_update, g, ABLK, 0, st, G.d);
}

#define ADJ_COARSE_RELAX 4

/* e0 = M f0: one V-cycle of every live problem (level-0 f already set) */
static void vcycle(AdjWork *w, int nswm, int nswp, hipStream_t st, bool first = true)
{
    const int D = w->D;
    /* level-0 (and every finest) correction starts at zero: restriction
     * zeroes the coarse ones; the finest of each problem is cleared here,
     * unless this is a further cycle of the same iteration */
    if (first)
        for (const AdjGroup &G : w->g0s) {
            const dim3 g = agrid(G.npx, G.npy, G.count);
            hipLaunchKernelGGL(k_adj_zero_e, g, ABLK, 0, st, G.d);
        }
    for (int d = D; d >= 0; --d) {
        for (int q : w->depth_groups[d]) {
            const AdjGroup &G = w->groups[q];
            const int reps = d == 0 ? ADJ_COARSE_RELAX : 1;
            for (int r = 0; r < reps; ++r) relax_group(G, nswm, nswp, w->single_colour, st);
            if (d >= 1) hipLaunchKernelGGL(k_adj_resid, agrid(G.npx, G.npy, G.count), ABLK, 0, st, G.d);
        }
        if (d >= 1)
            for (int p : w->depth_pairs[d]) {
                const AdjPair &R = w->pairs[p];
                hipLaunchKernelGGL(k_adj_restrict, agrid(R.npxc, R.npyc, R.count), ABLK, 0, st, R.f,
                                   R.c);
            }
    }
    for (int d = 1; d <= D; ++d) {
        for (int p : w->depth_pairs[d]) {
            const AdjPair &R = w->pairs[p];
            hipLaunchKernelGGL(k_adj_prolong, agrid(R.npxf, R.npyf, R.count), ABLK, 0, st, R.f, R.c);
        }
        for (int q : w->depth_groups[d])
            relax_group(w->groups[q], nswm, nswp, w->single_colour, st);
    }
}

/* precond_cycles 0: the V-cycles per preconditioner use for a batch of nb
 * problems.  A second cycle cuts the GMRES iterations by about a third and
 * costs a V-cycle per iteration; it pays where the batch holds problems that
 * need many iterations, whose Gram-Schmidt cost grows with the iterations
 * squared.  Measured (512^2 adjoint-workload problems, restart 60; one
 * cycle against two): 2, 4, 8 problems 0.062/0.089, 0.110/0.124,
 * 0.151/0.170 s; 16, 32, 64 problems 0.491/0.466, 0.798/0.745, 1.44/1.32 s;
 * the 256^2 single problem 0.062/0.077 s. */
#ifndef ADJ_AUTO_CYCLES
#define ADJ_AUTO_CYCLES(nb) ((nb) >= 16 ? 2 : 1)
#endif
/* the preconditioner: w->pc_cycles cycles of the V-cycle iteration from
 * e = 0 (a fixed linear operator on f) */
static void precondition(AdjWork *w, int nswm, int nswp, hipStream_t st)
{
    for (int c = 0; c < w->pc_cycles; ++c) vcycle(w, nswm, nswp, st, c == 0);
}

/* ---- pieces of the solve ---------------------------------------------- */

static void set_f0_from(AdjWork *w, int op, int jv, int nk, const double *y, int ystride,
                        hipStream_t st)
{
    for (const AdjGroup &G : w->g0s) {
        const long n4 = 4L * G.npx * G.npy;
        const dim3 g((unsigned)((n4 + 255) / 256), 1, (unsigned)G.count);
        hipLaunchKernelGGL(k_f0, g, dim3(256), 0, st, G.d, w->d_vec, op, jv, nk, y, ystride);
    }
}

/* W = K^T e0 (the V-cycle's level-0 result) or W = K^T X (from_x) */
static void apply_kt(AdjWork *w, bool from_x, hipStream_t st)
{
    double **in = from_x ? w->d_lam : w->d_kin;
    for (const AdjGroup &G : w->g0s)
        hipLaunchKernelGGL(k_adj_kt, agrid(G.npx, G.npy, G.count), ABLK, 0, st, G.d,
                           (const double *const *)in, w->d_kout);
}

static int dot_chunks(const AdjWork *w)
{
    return (int)((w->maxn4 + DOT_T * DOT_PER - 1) / (DOT_T * DOT_PER));
}

/* d_out[b] = <X, X> per live problem (which 0: W, 1: G), 0 for the others */
static int norms(mgf_solver *s, AdjWork *w, int which, double *d_out, hipStream_t st)
{
    const int nblk = dot_chunks(w), nb = w->nb;
    int rc = ensure_partial(s, w, (size_t)nb * nblk);
    if (rc) return rc;
    hipLaunchKernelGGL(k_norm2, dim3((unsigned)nblk, 1, (unsigned)nb), dim3(DOT_T), 0, st, w->d_vec,
                       w->d_live, which, w->d_partial, nblk);
    hipLaunchKernelGGL(k_dots_sum, dim3(1, (unsigned)nb), dim3(DOT_T), 0, st, w->d_partial,
                       w->d_live, 1, nblk, d_out);
    return MGF_OK;
}

/* One classical Gram-Schmidt pass on W of every live problem against
 * V_0..V_{nk-1}: d_hk[b * nk + k] = <V_k, W>, then W -= sum_k h_k V_k;
 * d_wn0[b] = <W, W> before the pass (when not null, from the same read of
 * W) and d_wn1[b] after it (from the subtraction's own pass). */
static int cgs_pass(mgf_solver *s, AdjWork *w, int nk, double *d_hk, double *d_wn0, double *d_wn1,
                    hipStream_t st)
{
    const int nblk = dot_chunks(w), nb = w->nb;
    int rc = ensure_partial(s, w, (size_t)nb * (nk + 1) * nblk);
    if (rc) return rc;
    double *part = w->d_partial, *wpart = part + (size_t)nb * nk * nblk;
    const unsigned gy = (unsigned)((nk + DOT_KB - 1) / DOT_KB);
    hipLaunchKernelGGL(k_dots, dim3((unsigned)nblk, gy, (unsigned)nb), dim3(DOT_T), 0, st,
                       w->d_vec, w->d_live, nk, d_wn0 != nullptr, part, wpart, nblk);
    hipLaunchKernelGGL(k_dots_sum, dim3((unsigned)nk, (unsigned)nb), dim3(DOT_T), 0, st, part,
                       w->d_live, nk, nblk, d_hk);
    if (d_wn0)
        hipLaunchKernelGGL(k_dots_sum, dim3(1, (unsigned)nb), dim3(DOT_T), 0, st, wpart, w->d_live,
                           1, nblk, d_wn0);
    hipLaunchKernelGGL(k_orth, dim3((unsigned)nblk, 1, (unsigned)nb), dim3(DOT_T), 0, st, w->d_vec,
                       w->d_live, nk, d_hk, nk, wpart, nblk);
    hipLaunchKernelGGL(k_dots_sum, dim3(1, (unsigned)nb), dim3(DOT_T), 0, st, wpart, w->d_live, 1,
                       nblk, d_wn1);
    return MGF_OK;
}

static void vec_launch(AdjWork *w, int op, int jv, const double *c, hipStream_t st)
{
    const dim3 g((unsigned)((w->maxn4 + 255) / 256), 1, (unsigned)w->nb);
    hipLaunchKernelGGL(k_vec_op, g, dim3(256), 0, st, w->d_vec, w->d_live, op, jv, c);
}

static int upload_live(mgf_solver *s, AdjWork *w, hipStream_t st)
{
    AHIP(s, hipMemcpyAsync(w->d_live, w->h_live.data(), w->nb * sizeof(int),
                           hipMemcpyHostToDevice, st));
    return MGF_OK;
}

static int sweeps_of(mgf_solver *s, int *nswm, int *nswp)
{
    const mgf_problem *p = mgf_instance_problem(s, 0);
    *nswm = p->nswpm > 0 ? p->nswpm : 2;
    *nswp = p->nswpp > 0 ? p->nswpp : 6;
    return MGF_OK;
}

/* The residual rows treat every ghost value as fixed, which is the steady
 * problem with Dirichlet ghosts (ibc 1) on every side; discrete inlets and
 * outlets set fixed ghosts too.  Periodic (3) and outflow (4) ghosts are
 * copies of unknowns, and p' Dirichlet rows (2) leave their cells' mass
 * balance unenforced: the converged state is then not a zero of these rows,
 * so such problems are refused. */
static int check_boundaries(mgf_solver *s)
{
    const int n = mgf_batch_size(s);
    for (int b = 0; b < n; ++b) {
        const mgf_problem *p = mgf_instance_problem(s, b);
        for (int k = 0; k < 4; ++k)
            if (p->ibc[k] != 1) {
                char msg[160];
                snprintf(msg, sizeof msg,
                         "problem %d: the adjoint needs ibc 1 on every side (side %d has %d)", b,
                         k, p->ibc[k]);
                return mgf_set_error(s, MGF_ERR_CONFIG, msg);
            }
    }
    return MGF_OK;
}

static int prepare(mgf_solver *s, int restart, AdjWork **out)
{
    int brc = check_boundaries(s);
    if (brc) return brc;
    void **slot = mgf_adjoint_slot(s);
    AdjWork *w = (AdjWork *)*slot;
    if (w && w->restart != restart) {
        mgf_adjoint_release(w);
        *slot = nullptr;
        w = nullptr;
    }
    if (!w) {
        int rc = build_work(s, restart, &w);
        if (rc) return rc;
        *slot = w;
    }
    hipStream_t st = (hipStream_t)mgf_stream(s);
    w->h_live.assign(w->nb, 1);
    int rc = upload_live(s, w, st);
    if (rc) return rc;
    /* the momentum diagonals of every level, from the current state */
    for (const AdjGroup &G : w->groups)
        hipLaunchKernelGGL(k_adj_diag, agrid(G.npx, G.npy, G.count), ABLK, 0, st, G.d);
    /* the K^T operand tables: e of level 0 in, W out */
    std::vector<double *> kin(4 * w->nb), kout(4 * w->nb);
    for (int b = 0; b < w->nb; ++b) {
        const long n = w->vec[b].n4 / 4;
        for (int k = 0; k < 4; ++k) {
            kin[4 * b + k] = w->lv[b][0].e[k];
            kout[4 * b + k] = w->vec[b].W + k * n;
        }
    }
    AHIP(s, hipMemcpyAsync(w->d_kin, kin.data(), kin.size() * sizeof(double *),
                           hipMemcpyHostToDevice, st));
    AHIP(s, hipMemcpyAsync(w->d_kout, kout.data(), kout.size() * sizeof(double *),
                           hipMemcpyHostToDevice, st));
    AHIP(s, hipStreamSynchronize(st));
    *out = w;
    return MGF_OK;
}

extern "C" void mgf_adjoint_opts_default(mgf_adjoint_opts *o)
{
    if (!o) return;
    o->rtol = 1e-11;
    o->max_iter = 600;
    o->restart = 60;
    o->only_converged = 1;
    o->single_colour = 0;
    o->precond_cycles = 0;
}

/* Givens rotation of (a, b) */
static void givens(double a, double b, double *c, double *s)
{
    if (b == 0.0) {
        *c = 1.0;
        *s = 0.0;
    } else {
        const double r = hypot(a, b);
        *c = a / r;
        *s = b / r;
    }
}

extern "C" int mgf_batch_adjoint(mgf_solver *s, int objective, const mgf_adjoint_opts *opt,
                                 double *J, double *g_inlet, double *g_amuf, double *g_prsc,
                                 int *iters, double *relres)
{
    if (!s) return mgf_set_error(s, MGF_ERR_ARG, "null solver");
    if (objective != 0 && objective != 1)
        return mgf_set_error(s, MGF_ERR_ARG, "objective must be 0 (pressure drop) or 1 (outlet flux)");
    mgf_adjoint_opts o;
    mgf_adjoint_opts_default(&o);
    if (opt) o = *opt;
    if (o.restart < 1 || o.restart > 200 || o.max_iter < 1 || !(o.rtol > 0.0) ||
        o.precond_cycles < 0 || o.precond_cycles > 4)
        return mgf_set_error(s, MGF_ERR_ARG, "bad adjoint options");
    AdjWork *w = nullptr;
    int rc = prepare(s, o.restart, &w);
    if (rc) return rc;
    w->single_colour = o.single_colour != 0;
    w->pc_cycles = o.precond_cycles > 0 ? o.precond_cycles : ADJ_AUTO_CYCLES(w->nb);
    hipStream_t st = (hipStream_t)mgf_stream(s);
    const int nb = w->nb, m = o.restart;
    int nswm, nswp;
    sweeps_of(s, &nswm, &nswp);
    /* small device scratch: h, h2 (nb x (m + 2) each), y (nb x m), c (nb), J (2 nb), the
     * norms n0, n1, n2 (nb each) */
    double *d_h = w->d_small, *d_y = d_h + (size_t)nb * (m + 2), *d_c = d_y + (size_t)nb * m;
    double *d_J = d_c + nb, *d_h2 = d_J + 2 * nb, *d_n0 = d_h2 + (size_t)nb * (m + 2);
    double *d_n1 = d_n0 + nb, *d_n2 = d_n1 + nb;
    /* the objective's dJ/dx and J */
    for (const AdjGroup &G : w->g0s)
        hipLaunchKernelGGL(k_adj_objective, dim3((unsigned)G.count), dim3(256), 0, st, G.d,
                           w->d_vec, objective, d_J);
    std::vector<double> hJ(2 * nb), gnorm(nb), beta(nb);
    AHIP(s, hipMemcpyAsync(hJ.data(), d_J, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, st));
    if ((rc = norms(s, w, 1, d_h, st))) return rc;
    std::vector<double> hh((size_t)nb * (m + 2));
    AHIP(s, hipMemcpyAsync(hh.data(), d_h, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    AHIP(s, hipStreamSynchronize(st));
    for (int b = 0; b < nb; ++b) gnorm[b] = sqrt(hh[b]);
    /* X = 0 */
    {
        std::vector<double> zero(nb, 0.0);
        AHIP(s, hipMemcpyAsync(d_c, zero.data(), nb * sizeof(double), hipMemcpyHostToDevice, st));
        for (int b = 0; b < nb; ++b)
            AHIP(s, hipMemsetAsync(w->vec[b].X, 0, w->vec[b].n4 * sizeof(double), st));
    }
    std::vector<int> done(nb, 0), total(nb, 0), skip(nb, 0);
    std::vector<double> rel(nb, 1.0);
    for (int b = 0; b < nb; ++b) {
        if (o.only_converged) {
            const int cv = mgf_instance_converged(s, b);
            if (cv < 0) return cv;
            skip[b] = !cv;
        }
        if (skip[b] || !(gnorm[b] > 0.0)) {
            done[b] = 1;      /* skipped, or dJ/dx = 0: lambda = 0 */
            rel[b] = 0.0;
        }
    }
    std::vector<double> H((size_t)nb * (m + 1) * m), cs((size_t)nb * m), sn((size_t)nb * m);
    std::vector<double> gv((size_t)nb * (m + 1)), yv((size_t)nb * m);
    std::vector<int> jcount(nb), inner(nb);
    std::vector<double> hc(nb);
    int nonfinite = 0;
    while (true) {
        bool any = false;
        for (int b = 0; b < nb; ++b) {
            w->h_live[b] = !done[b] && total[b] < o.max_iter;
            any |= w->h_live[b] != 0;
        }
        if (!any) break;
        if ((rc = upload_live(s, w, st))) return rc;
        /* W = G - K^T X, beta = ||W||, V_0 = W / beta */
        apply_kt(w, true, st);
        vec_launch(w, 1, 0, d_c, st);
        if ((rc = norms(s, w, 0, d_h, st))) return rc;
        AHIP(s, hipMemcpyAsync(hh.data(), d_h, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        AHIP(s, hipStreamSynchronize(st));
        for (int b = 0; b < nb; ++b) {
            beta[b] = sqrt(hh[b]);
            if (!w->h_live[b]) continue;
            rel[b] = beta[b] / gnorm[b];
            if (!std::isfinite(beta[b])) {
                nonfinite = 1;
                done[b] = 1;
                w->h_live[b] = 0;
            } else if (rel[b] <= o.rtol) {
                done[b] = 1;
                w->h_live[b] = 0;
            }
            hc[b] = w->h_live[b] ? 1.0 / beta[b] : 0.0;
            jcount[b] = 0;
            inner[b] = w->h_live[b];
            std::fill(gv.begin() + (size_t)b * (m + 1), gv.begin() + (size_t)(b + 1) * (m + 1), 0.0);
            gv[(size_t)b * (m + 1)] = beta[b];
        }
        if ((rc = upload_live(s, w, st))) return rc;
        AHIP(s, hipMemcpyAsync(d_c, hc.data(), nb * sizeof(double), hipMemcpyHostToDevice, st));
        vec_launch(w, 0, 0, d_c, st);
        for (int j = 0; j < m; ++j) {
            bool act = false;
            for (int b = 0; b < nb; ++b) {
                w->h_live[b] = inner[b] && total[b] < o.max_iter;
                act |= w->h_live[b] != 0;
            }
            if (!act) break;
            if ((rc = upload_live(s, w, st))) return rc;
            /* W = K^T M V_j */
            set_f0_from(w, 0, j, 0, nullptr, 0, st);
            precondition(w, nswm, nswp, st);
            apply_kt(w, false, st);
            /* classical Gram-Schmidt, and a second pass for the problems whose W
             * kept less than 1/sqrt(2) of its norm through the first (the
             * reorthogonalisation criterion of Daniel, Gragg, Kaufman and
             * Stewart, 1976): after it W is orthogonal to the V_k to working
             * precision, and most steps need one pass */
            const size_t nh = (size_t)nb * (j + 1);
            std::vector<double> hsum(nh), n0(nb), n1(nb);
            if ((rc = cgs_pass(s, w, j + 1, d_h, d_n0, d_n1, st))) return rc;
            AHIP(s, hipMemcpyAsync(hsum.data(), d_h, nh * sizeof(double), hipMemcpyDeviceToHost,
                                   st));
            AHIP(s, hipMemcpyAsync(n0.data(), d_n0, nb * sizeof(double), hipMemcpyDeviceToHost, st));
            AHIP(s, hipMemcpyAsync(n1.data(), d_n1, nb * sizeof(double), hipMemcpyDeviceToHost, st));
            AHIP(s, hipStreamSynchronize(st));
            {
                const std::vector<int> live_step = w->h_live;
                bool again = false;
                for (int b = 0; b < nb; ++b) {
                    w->h_live[b] = live_step[b] && !(n1[b] >= 0.5 * n0[b]);
                    again |= w->h_live[b] != 0;
                }
                if (again) {
                    std::vector<double> h2(nh), n2(nb);
                    if ((rc = upload_live(s, w, st))) return rc;
                    if ((rc = cgs_pass(s, w, j + 1, d_h2, nullptr, d_n2, st))) return rc;
                    AHIP(s, hipMemcpyAsync(h2.data(), d_h2, nh * sizeof(double),
                                           hipMemcpyDeviceToHost, st));
                    AHIP(s, hipMemcpyAsync(n2.data(), d_n2, nb * sizeof(double),
                                           hipMemcpyDeviceToHost, st));
                    AHIP(s, hipStreamSynchronize(st));
                    for (int b = 0; b < nb; ++b) {
                        if (!w->h_live[b]) continue;
                        for (int k = 0; k <= j; ++k) hsum[(size_t)b * (j + 1) + k] += h2[(size_t)b * (j + 1) + k];
                        n1[b] = n2[b];
                    }
                }
                w->h_live = live_step;
                if (again && (rc = upload_live(s, w, st))) return rc;
            }
            for (int b = 0; b < nb; ++b) {
                hc[b] = 0.0;
                if (!w->h_live[b]) continue;
                double *Hb = &H[(size_t)b * (m + 1) * m];
                const double hn = sqrt(n1[b]);
                for (int k = 0; k <= j; ++k) Hb[k * m + j] = hsum[(size_t)b * (j + 1) + k];
                Hb[(j + 1) * m + j] = hn;
                /* apply the previous rotations, then a new one */
                for (int k = 0; k < j; ++k) {
                    const double c = cs[(size_t)b * m + k], sv = sn[(size_t)b * m + k];
                    const double a0 = Hb[k * m + j], a1 = Hb[(k + 1) * m + j];
                    Hb[k * m + j] = c * a0 + sv * a1;
                    Hb[(k + 1) * m + j] = -sv * a0 + c * a1;
                }
                double c, sv;
                givens(Hb[j * m + j], Hb[(j + 1) * m + j], &c, &sv);
                cs[(size_t)b * m + j] = c;
                sn[(size_t)b * m + j] = sv;
                Hb[j * m + j] = c * Hb[j * m + j] + sv * Hb[(j + 1) * m + j];
                Hb[(j + 1) * m + j] = 0.0;
                double *g = &gv[(size_t)b * (m + 1)];
                g[j + 1] = -sv * g[j];
                g[j] = c * g[j];
                jcount[b] = j + 1;
                total[b] += 1;
                const double est = fabs(g[j + 1]) / gnorm[b];
                if (!std::isfinite(est) || !std::isfinite(hn)) {
                    nonfinite = 1;
                    inner[b] = 0;
                    done[b] = 1;
                } else if (est <= o.rtol || hn <= 1e-300 * beta[b]) {
                    inner[b] = 0;
                } else {
                    hc[b] = 1.0 / hn;
                }
            }
            AHIP(s, hipMemcpyAsync(d_c, hc.data(), nb * sizeof(double), hipMemcpyHostToDevice, st));
            vec_launch(w, 0, j + 1, d_c, st);
        }
        /* X += M (V y) for every problem that iterated */
        bool upd = false;
        for (int b = 0; b < nb; ++b) {
            double *y = &yv[(size_t)b * m];
            std::fill(y, y + m, 0.0);
            w->h_live[b] = !done[b] && jcount[b] > 0;
            if (!w->h_live[b]) continue;
            upd = true;
            const double *Hb = &H[(size_t)b * (m + 1) * m];
            const double *g = &gv[(size_t)b * (m + 1)];
            for (int k = jcount[b] - 1; k >= 0; --k) {
                double t = g[k];
                for (int q = k + 1; q < jcount[b]; ++q) t -= Hb[k * m + q] * y[q];
                y[k] = t / Hb[k * m + k];
            }
        }
        if (!upd) break;
        if ((rc = upload_live(s, w, st))) return rc;
        AHIP(s, hipMemcpyAsync(d_y, yv.data(), (size_t)nb * m * sizeof(double),
                               hipMemcpyHostToDevice, st));
        set_f0_from(w, 1, 0, m, d_y, m, st);
        precondition(w, nswm, nswp, st);
        set_f0_from(w, 2, 0, 0, nullptr, 0, st);
        vec_launch(w, 2, 0, d_c, st);
        AHIP(s, hipStreamSynchronize(st));
        if (hipGetLastError() != hipSuccess)
            return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint kernel launch failed");
        /* the loop head checks the true residual of the updated X */
    }
    /* final relative residuals */
    for (int b = 0; b < nb; ++b) w->h_live[b] = 1;
    if ((rc = upload_live(s, w, st))) return rc;
    apply_kt(w, true, st);
    vec_launch(w, 1, 0, d_c, st);
    if ((rc = norms(s, w, 0, d_h, st))) return rc;
    AHIP(s, hipMemcpyAsync(hh.data(), d_h, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    AHIP(s, hipStreamSynchronize(st));
    int unconverged = 0;
    for (int b = 0; b < nb; ++b) {
        if (skip[b]) {
            rel[b] = NAN;
            continue;
        }
        rel[b] = gnorm[b] > 0.0 ? sqrt(hh[b]) / gnorm[b] : 0.0;
        if (!std::isfinite(rel[b])) nonfinite = 1;
        if (!(rel[b] <= 10.0 * o.rtol)) unconverged = 1;
    }
    /* gradients */
    std::vector<MetSeed> seeds(2 * nb);
    int maxny = 0;
    size_t nfy_total = 0;
    std::vector<size_t> off(nb);
    for (int b = 0; b < nb; ++b) {
        const mgf_problem *p = mgf_instance_problem(s, b);
        const DevLevel &L = w->lv[b][0].P;
        const double dx = p->xl / (double)(L.npx - 2), dy = p->yl / (double)(L.npy - 2);
        const double rx = L.ax / dx, ry = L.ay / dy;
        seeds[2 * b] = MetSeed{rx, ry, rx / p->prsc, ry / p->prsc};
        const double q = -p->amuf / (p->prsc * p->prsc);
        seeds[2 * b + 1] = MetSeed{0.0, 0.0, q * rx, q * ry};
        maxny = std::max(maxny, L.npy);
        off[b] = nfy_total;
        nfy_total += (size_t)(L.npy - 2);
    }
    int maxblk = 0;
    for (const AdjGroup &G : w->g0s) {
        const dim3 g = agrid(G.npx, G.npy, 1);
        maxblk = std::max(maxblk, (int)(g.x * g.y));
    }
    if ((rc = ensure_partial(s, w, std::max((size_t)nb * maxblk * 2, nfy_total + nb)))) return rc;
    double *d_gin = w->d_partial;                    /* reused after the partial sums are read */
    AHIP(s, hipMemcpyAsync(w->d_seeds, seeds.data(), seeds.size() * sizeof(MetSeed),
                           hipMemcpyHostToDevice, st));
    AHIP(s, hipMemsetAsync(w->d_partial, 0, (size_t)nb * maxblk * 2 * sizeof(double), st));
    for (const AdjGroup &G : w->g0s)
        hipLaunchKernelGGL(k_adj_grad_par, agrid(G.npx, G.npy, G.count), ABLK, 0, st, G.d,
                           w->d_seeds, (const double *const *)w->d_lam, w->d_partial, maxblk);
    std::vector<double> part((size_t)nb * maxblk * 2);
    AHIP(s, hipMemcpyAsync(part.data(), w->d_partial, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    AHIP(s, hipStreamSynchronize(st));
    std::vector<double *> gptr(nb);
    for (int b = 0; b < nb; ++b) gptr[b] = d_gin + off[b];
    double *d_guxp = d_gin + nfy_total;
    AHIP(s, hipMemcpyAsync(w->d_gin, gptr.data(), nb * sizeof(double *), hipMemcpyHostToDevice, st));
    for (const AdjGroup &G : w->g0s)
        hipLaunchKernelGGL(k_adj_grad_bc, dim3((unsigned)((G.npy - 1 + 63) / 64), 1, (unsigned)G.count),
                           dim3(64), 0, st, G.d, (const double *const *)w->d_lam, objective,
                           w->d_gin, d_guxp);
    std::vector<double> gin(nfy_total + nb);
    AHIP(s, hipMemcpyAsync(gin.data(), d_gin, gin.size() * sizeof(double), hipMemcpyDeviceToHost,
                           st));
    AHIP(s, hipStreamSynchronize(st));
    if (hipGetLastError() != hipSuccess)
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint gradient launch failed");
    w->last_iters = total;
    w->last_relres = rel;
    for (int b = 0; b < nb; ++b) {
        const int nfy = w->lv[b][0].P.npy - 2;
        const double guxp = gin[nfy_total + b];
        double ga = 0.0, gp = 0.0;
        for (int q = 0; q < maxblk; ++q) {
            ga += part[((size_t)b * maxblk + q) * 2];
            gp += part[((size_t)b * maxblk + q) * 2 + 1];
        }
        const double nan = skip[b] ? NAN : 0.0;
        if (J) J[b] = hJ[2 * b + objective];
        if (g_amuf) g_amuf[b] = ga + nan;
        if (g_prsc) g_prsc[b] = gp + nan;
        if (g_inlet)
            for (int r = 0; r < nfy; ++r)
                g_inlet[off[b] + r] = gin[off[b] + r] + guxp / (double)nfy + nan;
        if (iters) iters[b] = skip[b] ? -1 : total[b];
        if (relres) relres[b] = rel[b];
    }
    if (nonfinite) return mgf_set_error(s, MGF_ERR_NONFINITE, "non-finite adjoint residual");
    if (unconverged)
        return mgf_set_error(s, MGF_ERR_NONFINITE,
                             "adjoint solve did not reach its tolerance within max_iter");
    return MGF_OK;
}

extern "C" int mgf_adjoint(mgf_solver *s, int objective, double *g_inlet, double *g_amuf,
                           double *g_prsc)
{
    if (!s) return mgf_set_error(s, MGF_ERR_ARG, "null solver");
    if (mgf_batch_size(s) != 1)
        return mgf_set_error(s, MGF_ERR_UNSUPPORTED, "mgf_adjoint needs a single-problem solver");
    const int cv = mgf_instance_converged(s, 0);
    if (cv < 0) return cv;
    if (!cv)
        return mgf_set_error(s, MGF_ERR_CONFIG,
                             "the solver's last solve did not converge: the adjoint is that of "
                             "the steady solution");
    return mgf_batch_adjoint(s, objective, nullptr, nullptr, g_inlet, g_amuf, g_prsc, nullptr,
                             nullptr);
}

extern "C" int mgf_objectives(mgf_solver *s, int b, double J[2])
{
    if (!s || !J || b < 0 || b >= mgf_batch_size(s)) return mgf_set_error(s, MGF_ERR_ARG, "bad arguments");
    int npx = 0, npy = 0;
    int rc = mgf_batch_level_dims(s, b, 0, &npx, &npy);
    if (rc) return rc;
    const size_t n = (size_t)npx * npy;
    std::vector<double> u(n), p(n), sc(n);
    if ((rc = mgf_batch_get_field(s, b, 0, "u", u.data())) ||
        (rc = mgf_batch_get_field(s, b, 0, "p", p.data())) ||
        (rc = mgf_batch_get_field(s, b, 0, "sc", sc.data())))
        return rc;
    const mgf_problem *q = mgf_instance_problem(s, b);
    const double ax = q->yl / (double)(npy - 2);
    double a = 0.0, c = 0.0, j2 = 0.0;
    for (int j = 1; j <= npy - 2; ++j) {
        a += p[(size_t)j * npx + 1];
        c += p[(size_t)j * npx + npx - 2];
        const size_t ij = (size_t)j * npx + npx - 2;
        j2 += (ax * u[ij]) * sc[ij];
    }
    J[0] = a / (double)(npy - 2) - c / (double)(npy - 2);
    J[1] = j2;
    return MGF_OK;
}

extern "C" int mgf_adjoint_get_state(mgf_solver *s, int b, double *lam_u, double *lam_v,
                                     double *lam_c, double *lam_s)
{
    if (!s || b < 0 || b >= mgf_batch_size(s)) return mgf_set_error(s, MGF_ERR_ARG, "bad arguments");
    AdjWork *w = (AdjWork *)*mgf_adjoint_slot(s);
    if (!w) return mgf_set_error(s, MGF_ERR_ARG, "no adjoint solve on this solver yet");
    const long n = w->vec[b].n4 / 4;
    double *outs[4] = {lam_u, lam_v, lam_c, lam_s};
    hipStream_t st = (hipStream_t)mgf_stream(s);
    for (int k = 0; k < 4; ++k)
        if (outs[k])
            AHIP(s, hipMemcpyAsync(outs[k], w->vec[b].X + k * n, n * sizeof(double),
                                   hipMemcpyDeviceToHost, st));
    AHIP(s, hipStreamSynchronize(st));
    return MGF_OK;
}

extern "C" int mgf_adjoint_apply(mgf_solver *s, int b, int op, const double *in, double *out)
{
    if (!s || b < 0 || b >= mgf_batch_size(s) || !in || !out || (op != 0 && op != 1))
        return mgf_set_error(s, MGF_ERR_ARG, "bad arguments");
    mgf_adjoint_opts o;
    mgf_adjoint_opts_default(&o);
    AdjWork *w = (AdjWork *)*mgf_adjoint_slot(s);
    int rc = prepare(s, w ? w->restart : o.restart, &w);
    if (rc) return rc;
    w->single_colour = 0;
    w->pc_cycles = ADJ_AUTO_CYCLES(w->nb);
    hipStream_t st = (hipStream_t)mgf_stream(s);
    for (int q = 0; q < w->nb; ++q) w->h_live[q] = q == b;
    if ((rc = upload_live(s, w, st))) return rc;
    const long n4 = w->vec[b].n4;
    int nswm, nswp;
    sweeps_of(s, &nswm, &nswp);
    if (op == 0) {
        AHIP(s, hipMemcpyAsync(w->vec[b].X, in, n4 * sizeof(double), hipMemcpyHostToDevice, st));
        apply_kt(w, true, st);
    } else {
        AHIP(s, hipMemcpyAsync(w->vec[b].V, in, n4 * sizeof(double), hipMemcpyHostToDevice, st));
        set_f0_from(w, 0, 0, 0, nullptr, 0, st);
        precondition(w, nswm, nswp, st);
        set_f0_from(w, 2, 0, 0, nullptr, 0, st);
    }
    AHIP(s, hipMemcpyAsync(out, w->vec[b].W, n4 * sizeof(double), hipMemcpyDeviceToHost, st));
    AHIP(s, hipStreamSynchronize(st));
    if (hipGetLastError() != hipSuccess)
        return mgf_set_error(s, MGF_ERR_DEVICE, "adjoint kernel launch failed");
    return MGF_OK;
}
