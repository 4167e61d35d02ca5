/*
 * mgf_kernels_mixed.hip -- mixed-precision smoothing for gfx950: the
 * equations are assembled in FP64 exactly as in the FP64 path, the smoother
 * sweeps run in FP32 on a correction, and the correction is added in FP64.
 *
 * u, v, sc (defect correction): the assembly kernel evaluates the FP64
 * residual r = su + sum(a_nb phi_nb) - ap phi of the reference's equation
 * and stores the coefficients, 1/ap and r in FP32; nsw FP32 sweeps of the
 * same smoother (red-black or Jacobi) on A d = r from d = 0, with d = 0
 * outside the swept range; then phi += d in FP64.  In exact arithmetic this
 * is the same iteration as nsw sweeps on A phi = b from phi (the smoothers
 * are linear stationary iterations), and at convergence r -> 0, so the FP32
 * rounding of d does not limit the converged FP64 state.
 * p': already a correction that starts from zero every relaxation
 * (masserror sets pp = 0, V6:1116), smoothed directly in FP32 with the FP64
 * mass defect -div as right-hand side; pp is then widened to FP64.
 * The level-0 sum |div|, the update of u, v, p and every grid transfer stay
 * FP64.
 */
#include <hip/hip_runtime.h>

#include "mgf_stencil.h"

/* FP64 assembly of one equation at one point plus its FP64 residual, stored
 * in FP32 (field 0 u, 1 v, 3 sc; p' has its own kernel below) */
__global__ void __launch_bounds__(BX * BY) k_asm_mx(const DevLevel *Ls, int field)
{
    LEVEL(Ls);
    const DevPhys &P = L.P;
    const int nx = L.npx;
    int i0 = 1, i1 = nx - 2, j1 = L.npy - 2;
    if (field == 0) i1 = nx - 3;
    if (field == 1) j1 = L.npy - 3;
    RANGE2D(i0, i1, 1, j1);
    const int ij = j * nx + i;
    Coef c;
    double su, ap;
    const double *phi;
    if (field == 0) {
        c = coefu_at(L, FluxArr{L.cx, L.cy}, i, j);
        su = L.ax * (L.p[ij] - L.p[ij + 1]);
        ap = c.ap / P.relxm;
        su = ((su + (P.relxm1 * ap) * L.u[ij]) + L.resu[ij]) + L.bodyx;
        ap = ap + ((L.flags[ij] & MGF_F_SPU) ? 1.0e30 : 0.0);
        L.apu[ij] = ap;
        phi = L.u;
    } else if (field == 1) {
        c = coefv_at(L, FluxArr{L.cx, L.cy}, i, j);
        su = L.ay * (L.p[ij] - L.p[ij + nx]);
        ap = c.ap / P.relxm;
        su = ((su + (P.relxm1 * ap) * L.v[ij]) + L.resv[ij]) + L.bodyy;
        ap = ap + ((L.flags[ij] & MGF_F_SPV) ? 1.0e30 : 0.0);
        L.apv[ij] = ap;
        phi = L.v;
    } else {
        const uint8_t fl = L.flags[ij];
        c = coefsc_at(L, FluxArr{L.cx, L.cy}, i, j, fl);
        const double spsc = (fl & MGF_F_SPSC) ? 1.0e30 : 0.0;
        su = spsc * P.sc_obs + L.ressc[ij];
        ap = c.ap + spsc;
        phi = L.sc;
    }
    const double r = ((((su + c.aw * phi[ij - 1]) + c.ae * phi[ij + 1]) + c.as * phi[ij - nx]) +
                      c.an * phi[ij + nx]) - ap * phi[ij];
    L.aw32[ij] = (float)c.aw;
    L.ae32[ij] = (float)c.ae;
    L.as32[ij] = (float)c.as;
    L.an32[ij] = (float)c.an;
    L.iap32[ij] = (float)(L.isolve == 1 ? 1.0 / (ap + 1.0e-30) : 1.0 / ap);
    L.r32[ij] = (float)r;
}

/* masserror + coefp in FP64 (V6:1018-1125), stored in FP32; partial sums of
 * |div| in FP64 for the deterministic reduction */
__global__ void __launch_bounds__(BX * BY) k_asm_p_mx(const DevLevel *Ls)
{
    LEVEL(Ls);
    const DevPhys &P = L.P;
    __shared__ double lds[BY];
    const int i = 1 + (int)(blockIdx.x * BX + threadIdx.x);
    const int j = 1 + (int)(blockIdx.y * BY + threadIdx.y);
    const int nx = L.npx;
    double mag = 0.0;
    if (i <= nx - 2 && j <= L.npy - 2) {
        const int ij = j * nx + i, ijs = ij - nx;
        double div = ((L.ax * L.u[ij] - L.ax * L.u[ij - 1]) + L.ay * L.v[ij]) - L.ay * L.v[ijs];
        double su = -div;
        mag = fabs(div);
        double aw = L.axax / L.apu[ij - 1];
        double ae = L.axax / L.apu[ij];
        double as = L.ayay / L.apv[ijs];
        double an = L.ayay / L.apv[ij];
        double ap = (((aw + ae) + as) + an) + ((L.flags[ij] & MGF_F_OBS) ? 1.0e30 : 0.0);
        const double diag = 1.0e30;
        if (P.ibc[0] == 2 && i == 1) { ap = ap + diag; su = 0.0; }
        if (P.ibc[1] == 2 && i == nx - 2) { ap = ap + diag; su = 0.0; }
        if (P.ibc[2] == 2 && j == 1) { ap = ap + diag; su = 0.0; }
        if (P.ibc[3] == 2 && j == L.npy - 2) { ap = ap + diag; su = 0.0; }
        L.aw32[ij] = (float)aw;
        L.ae32[ij] = (float)ae;
        L.as32[ij] = (float)as;
        L.an32[ij] = (float)an;
        L.iap32[ij] = (float)(L.isolve == 1 ? 1.0 / (ap + 1.0e-30) : 1.0 / ap);
        L.r32[ij] = (float)su;
    }
    double bs = block_sum(mag, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) L.partial[blockIdx.y * gridDim.x + blockIdx.x] = bs;
}

/* extents of the swept range of a field (the reference's solver arguments) */
__device__ __forceinline__ void extents(const DevLevel &L, int field, int &nx, int &ny)
{
    (void)field_of(L, field, nx, ny);
}

/* One FP32 red-black sweep (both colours) of A d = r on a 64 x RB_TY tile,
 * d read from buffer it&1 (zero on the first sweep and outside the swept
 * range) and written to buffer (it+1)&1.  Same tiling as the FP64 sweep. */
#define RB_TY 16
#define RB_Q (RB_TY / BY)
#define RB_RING (2 * (BX + 2) + 2 * RB_TY)

struct Pt32 {
    float aw, ae, as, an, iap, r;
};

__device__ __forceinline__ void ring_pos(int k, int &r, int &c)
{
    if (k < BX + 2) { r = 1; c = 1 + k; }
    else if (k < 2 * (BX + 2)) { r = RB_TY + 2; c = 1 + (k - (BX + 2)); }
    else if (k < 2 * (BX + 2) + RB_TY) { r = 2 + (k - 2 * (BX + 2)); c = 1; }
    else { r = 2 + (k - 2 * (BX + 2) - RB_TY); c = BX + 2; }
}

__device__ __forceinline__ Pt32 load_pt(const DevLevel &L, int g)
{
    return Pt32{L.aw32[g], L.ae32[g], L.as32[g], L.an32[g], L.iap32[g], L.r32[g]};
}

__global__ void __launch_bounds__(BX * BY) k_rb32(const DevLevel *Ls, int field, int it)
{
    LEVEL(Ls);
    __shared__ float t[RB_TY + 4][BX + 4];
    int nx, ny;
    extents(L, field, nx, ny);
    const float *src = it == 0 ? nullptr : ((it & 1) ? L.db32 : L.da32);
    float *dst = (it & 1) ? L.da32 : L.db32;
    const int npx = L.npx;
    const int i0 = 1 + (int)blockIdx.x * BX, j0 = 1 + (int)blockIdx.y * RB_TY;
    const int tid = threadIdx.y * BX + threadIdx.x;
    constexpr int NST = ((RB_TY + 4) * (BX + 4) + BX * BY - 1) / (BX * BY);
    float sv[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int k = tid + q * BX * BY;
        float val = 0.f;
        if (src && k < (RB_TY + 4) * (BX + 4)) {
            const int r = k / (BX + 4), c = k % (BX + 4);
            const int gi = i0 - 2 + c, gj = j0 - 2 + r;
            if (gi >= 1 && gi <= nx - 1 && gj >= 1 && gj <= ny - 1) val = src[gj * npx + gi];
        }
        sv[q] = val;
    }
    const int gi = i0 + threadIdx.x;
    Pt32 own[RB_Q];
    bool live[RB_Q];
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int gj = j0 + threadIdx.y + BY * q;
        live[q] = (gi <= nx - 1 && gj <= ny - 1);
        own[q] = live[q] ? load_pt(L, gj * npx + gi) : Pt32{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    }
    int rr = 0, rc = 0;
    bool ring = false;
    Pt32 rp = Pt32{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (tid < RB_RING) {
        ring_pos(tid, rr, rc);
        const int ri = i0 - 2 + rc, rj = j0 - 2 + rr;
        ring = (((ri + rj) & 1) == 0) && ri >= 1 && ri <= nx - 1 && rj >= 1 && rj <= ny - 1;
        if (ring) rp = load_pt(L, rj * npx + ri);
    }
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int k = tid + q * BX * BY;
        if (k < (RB_TY + 4) * (BX + 4)) t[k / (BX + 4)][k % (BX + 4)] = sv[q];
    }
    __syncthreads();
    float res[RB_Q];
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int r = 2 + threadIdx.y + BY * q, c = 2 + threadIdx.x;
        const int gj = j0 + threadIdx.y + BY * q;
        res[q] = t[r][c];
        if (live[q] && ((gi + gj) & 1) == 0) {
            const Pt32 &o = own[q];
            const float s = ((o.aw * t[r][c - 1] + o.ae * t[r][c + 1]) + o.as * t[r - 1][c]) +
                            o.an * t[r + 1][c];
            res[q] = (s + o.r) * o.iap;
        }
    }
    float rv = 0.f;
    if (ring) {
        const float s = ((rp.aw * t[rr][rc - 1] + rp.ae * t[rr][rc + 1]) + rp.as * t[rr - 1][rc]) +
                        rp.an * t[rr + 1][rc];
        rv = (s + rp.r) * rp.iap;
    }
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int gj = j0 + threadIdx.y + BY * q;
        if (live[q] && ((gi + gj) & 1) == 0) t[2 + threadIdx.y + BY * q][2 + threadIdx.x] = res[q];
    }
    if (ring) t[rr][rc] = rv;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RB_Q; ++q) {
        const int r = 2 + threadIdx.y + BY * q, c = 2 + threadIdx.x;
        const int gj = j0 + threadIdx.y + BY * q;
        if (!live[q]) continue;
        float val = res[q];
        if (((gi + gj) & 1) == 1) {
            const Pt32 &o = own[q];
            const float s = ((o.aw * t[r][c - 1] + o.ae * t[r][c + 1]) + o.as * t[r - 1][c]) +
                            o.an * t[r + 1][c];
            val = (s + o.r) * o.iap;
        }
        dst[gj * npx + gi] = val;
    }
}

/* one FP32 Jacobi sweep of A d = r (wj = 1, V6:1241) */
__global__ void __launch_bounds__(BX * BY) k_jac32(const DevLevel *Ls, int field, int it)
{
    LEVEL(Ls);
    int nx, ny;
    extents(L, field, nx, ny);
    RANGE2D(1, nx - 1, 1, ny - 1);
    const float *src = it == 0 ? nullptr : ((it & 1) ? L.db32 : L.da32);
    float *dst = (it & 1) ? L.da32 : L.db32;
    const int npx = L.npx, ij = j * npx + i;
    float W = 0.f, E = 0.f, S = 0.f, N = 0.f;
    if (src) {
        W = (i - 1 >= 1) ? src[ij - 1] : 0.f;
        E = (i + 1 <= nx - 1) ? src[ij + 1] : 0.f;
        S = (j - 1 >= 1) ? src[ij - npx] : 0.f;
        N = (j + 1 <= ny - 1) ? src[ij + npx] : 0.f;
    }
    const float s = ((L.aw32[ij] * W + L.ae32[ij] * E) + L.as32[ij] * S) + L.an32[ij] * N;
    dst[ij] = (s + L.r32[ij]) * L.iap32[ij];
}

/* phi += d (u, v, sc) or pp = d, in FP64, over the swept range */
__global__ void __launch_bounds__(BX * BY) k_apply_mx(const DevLevel *Ls, int field, int nsw)
{
    LEVEL(Ls);
    int nx, ny;
    double *phi = field_of(L, field, nx, ny);
    RANGE2D(1, nx - 1, 1, ny - 1);
    const float *d = (nsw & 1) ? L.db32 : L.da32;
    const int ij = j * L.npx + i;
    if (field == 2)
        phi[ij] = nsw > 0 ? (double)d[ij] : 0.0;
    else
        phi[ij] = phi[ij] + (double)d[ij];
}

/* ====================================================================== */

void mgf_launch_reduce_err(const DevGroup &G, int record, hipStream_t st);

static inline dim3 grid3(int ni, int nj, int nz, int by = BY)
{
    return dim3((unsigned)((ni + BX - 1) / BX), (unsigned)((nj + by - 1) / by), (unsigned)nz);
}

void mgf_launch_asm_mixed(const DevGroup &G, int field, int record, hipStream_t st)
{
    if (field == 2) {
        hipLaunchKernelGGL(k_asm_p_mx, grid3(G.npx - 2, G.npy - 2, G.count), dim3(BX, BY), 0, st,
                           G.d);
        mgf_launch_reduce_err(G, record, st);
        return;
    }
    int ni = G.npx - 2, nj = G.npy - 2;
    if (field == 0) ni = G.npx - 3;
    if (field == 1) nj = G.npy - 3;
    if (ni < 1 || nj < 1) return;
    hipLaunchKernelGGL(k_asm_mx, grid3(ni, nj, G.count), dim3(BX, BY), 0, st, G.d, field);
}

void mgf_launch_sweeps_mixed(const DevGroup &G, int field, int isolve, int nsw, hipStream_t st)
{
    int nx, ny;
    switch (field) {
    case 0: nx = G.npx - 2; ny = G.npy - 1; break;
    case 1: nx = G.npx - 1; ny = G.npy - 2; break;
    default: nx = G.npx - 1; ny = G.npy - 1; break;
    }
    if (nx - 1 < 1 || ny - 1 < 1 || (isolve != 1 && isolve != 2)) return;
    for (int it = 0; it < nsw; ++it) {
        if (isolve == 2)
            hipLaunchKernelGGL(k_rb32, grid3(nx - 1, ny - 1, G.count, RB_TY), dim3(BX, BY), 0, st,
                               G.d, field, it);
        else
            hipLaunchKernelGGL(k_jac32, grid3(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st, G.d,
                               field, it);
    }
}

void mgf_launch_smooth_mixed(const DevGroup &G, int field, int isolve, int nsw, hipStream_t st)
{
    int nx, ny;
    switch (field) {
    case 0: nx = G.npx - 2; ny = G.npy - 1; break;
    case 1: nx = G.npx - 1; ny = G.npy - 2; break;
    default: nx = G.npx - 1; ny = G.npy - 1; break;
    }
    if (nx - 1 < 1 || ny - 1 < 1 || (isolve != 1 && isolve != 2)) return;
    for (int it = 0; it < nsw; ++it) {
        if (isolve == 2)
            hipLaunchKernelGGL(k_rb32, grid3(nx - 1, ny - 1, G.count, RB_TY), dim3(BX, BY), 0, st,
                               G.d, field, it);
        else
            hipLaunchKernelGGL(k_jac32, grid3(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st, G.d,
                               field, it);
    }
    /* without sweeps the correction is zero: u, v, sc stay, pp becomes 0 */
    if (nsw > 0 || field == 2)
        hipLaunchKernelGGL(k_apply_mx, grid3(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st, G.d,
                           field, nsw);
}
