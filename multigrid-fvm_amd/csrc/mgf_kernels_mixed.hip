/*
 * mgf_kernels_mixed.hip -- mixed-precision smoothing for gfx950: the
 * equations are assembled in FP64 exactly as in the FP64 path, the smoother
 * sweeps run in FP32 on a correction, and the correction is added in FP64.
 *
 * u, v, sc (defect correction): the assembly kernel evaluates the FP64
 * residual r = su + sum(a_nb phi_nb) - ap phi of the reference's equation
 * and stores the coefficients, 1/ap (an FP32 reciprocal) and r in FP32; the
 * under-relaxed diagonal is ap * (1/relxm), whose rounding the relaxation
 * terms cancel at the fixed point; nsw FP32 sweeps of the
 * same smoother (red-black or Jacobi) on A d = r from d = 0, with d = 0
 * outside the swept range; then phi += d in FP64.  In exact arithmetic this
 * is the same iteration as nsw sweeps on A phi = b from phi (the smoothers
 * are linear stationary iterations), and at convergence r -> 0, so the FP32
 * rounding of d does not limit the converged FP64 state.
 * p': already a correction that starts from zero every relaxation
 * (masserror sets pp = 0, V6:1116), smoothed directly in FP32 with the FP64
 * mass defect -div as right-hand side; pp is then widened to FP64.  Its
 * coefficients area^2 / ap (coefp) feed nothing but the FP32 sweeps and are
 * formed in FP32; at the fixed point the defect, and with it p', is zero
 * whatever they are.
 * The level-0 sum |div|, the update of u, v, p and every grid transfer stay
 * FP64.
 *
 * The launch sequence of a relaxation: k_rb32x_uvs (u, v, sc together, when
 * the configuration allows; see mgf_launch_relax_merged_mixed) and one p'
 * launch, or field by field through mgf_launch_relax_field_mixed (split
 * assembly, Jacobi, outflow boundaries).
 */
#include <hip/hip_runtime.h>

#include "mgf_stencil.h"

struct Pt32 {
    float aw, ae, as, an, iap, r;
};

__device__ __forceinline__ Pt32 load_pt(const DevLevelG &L, int g)
{
    return Pt32{L.aw32[g], L.ae32[g], L.as32[g], L.an32[g], L.iap32[g], L.r32[g]};
}

__device__ __forceinline__ void store_pt(const DevLevelG &L, int g, const Pt32 &q)
{
    L.aw32[g] = q.aw;
    L.ae32[g] = q.ae;
    L.as32[g] = q.as;
    L.an32[g] = q.an;
    L.iap32[g] = q.iap;
    L.r32[g] = q.r;
}

/* 1 / a in FP32: v_rcp_f32 (1 ulp) and one Newton step, r + r (1 - a r),
 * which leaves it within about half an ulp; 0 and infinity keep the
 * reciprocal's own inf and 0 (the Newton step would turn them into NaN). */
__device__ __forceinline__ float rcp32(float a)
{
    const float r = __builtin_amdgcn_rcpf(a);
    const float q = __builtin_fmaf(r, __builtin_fmaf(-a, r, 1.0f), r);
    return __builtin_isfinite(q) ? q : r;
}

/* The FP32 reciprocal of a diagonal for the correction sweeps (isolve 1
 * keeps the Jacobi solver's 1e-30 guard).  Only the FP32 correction
 * operator uses it: the FP64 residual, not these digits, sets the converged
 * state. */
__device__ __forceinline__ float rcp_diag(float ap, int isolve)
{
    return rcp32(isolve == 1 ? ap + 1.0e-30f : ap);
}

/* FP64 assembly of one equation at one point plus its FP64 residual, rounded
 * to FP32 (field 0 u, 1 v, 3 sc); ap_out = the FP64 diagonal, which the
 * caller stores as apu / apv for u / v.  No global stores here, so a caller
 * assembling many points can keep all their loads in flight.  INL: the face
 * fluxes are formed from u and v (FluxInl) instead of read from cx / cy;
 * the same values while u and v are those cx / cy were computed from. */
template <bool INL>
__device__ __forceinline__ Coef coef_pt(const DevLevelG &L, int field, int i, int j, uint8_t fl)
{
    if (INL) {
        const FluxInl f{L.u, L.v, L.ax, L.ay};
        return field == 0 ? coefu_at(L, f, i, j)
                          : (field == 1 ? coefv_at(L, f, i, j) : coefsc_at(L, f, i, j, fl));
    }
    const FluxArr f{L.cx, L.cy};
    return field == 0 ? coefu_at(L, f, i, j)
                      : (field == 1 ? coefv_at(L, f, i, j) : coefsc_at(L, f, i, j, fl));
}

template <bool INL = false>
__device__ __forceinline__ Pt32 asm_pt(const DevLevelG &L, int field, int i, int j, double &ap_out)
{
    const DevPhys &P = L.P;
    const int nx = L.npx;
    const int ij = j * nx + i;
    Coef c;
    double su, ap;
    const gdbl *phi;
    if (field == 0) {
        c = coef_pt<INL>(L, 0, i, j, 0);
        su = L.ax * (L.p[ij] - L.p[ij + 1]);
        ap = c.ap * P.irelxm;
        su = ((su + (P.relxm1 * ap) * L.u[ij]) + L.resu[ij]) + L.bodyx;
        ap = ap + ((L.flags[ij] & MGF_F_SPU) ? 1.0e30 : 0.0);
        phi = L.u;
    } else if (field == 1) {
        c = coef_pt<INL>(L, 1, i, j, 0);
        su = L.ay * (L.p[ij] - L.p[ij + nx]);
        ap = c.ap * P.irelxm;
        su = ((su + (P.relxm1 * ap) * L.v[ij]) + L.resv[ij]) + L.bodyy;
        ap = ap + ((L.flags[ij] & MGF_F_SPV) ? 1.0e30 : 0.0);
        phi = L.v;
    } else {
        const uint8_t fl = L.flags[ij];
        c = coef_pt<INL>(L, 3, i, j, fl);
        const double spsc = (fl & MGF_F_SPSC) ? 1.0e30 : 0.0;
        su = spsc * P.sc_obs + L.ressc[ij];
        ap = c.ap + spsc;
        phi = L.sc;
    }
    const double r = ((((su + c.aw * phi[ij - 1]) + c.ae * phi[ij + 1]) + c.as * phi[ij - nx]) +
                      c.an * phi[ij + nx]) - ap * phi[ij];
    ap_out = ap;
    return Pt32{(float)c.aw, (float)c.ae, (float)c.as, (float)c.an, rcp_diag((float)ap, L.isolve),
                (float)r};
}

/* masserror + coefp in FP64 (V6:1018-1125) at one cell, rounded to FP32;
 * mag = |div|.  u and v carry their pending corrections over their swept
 * ranges (u faces 1..npx-3, v faces 1..npy-3): the values are u + du,
 * v + dv (pending_value: 0 on penalised faces), formed as the solves' own
 * update would form them -- here (CORR false), or already formed by the
 * merged relaxation in uc, vc (CORR). */
template <bool CORR = false>
__device__ __forceinline__ Pt32 asm_pt_p(const DevLevelG &L, int i, int j, double &mag)
{
    const DevPhys &P = L.P;
    const int nx = L.npx;
    const int ij = j * nx + i, ijs = ij - nx;
    double ue, uw, vn, vs;
    if (CORR) {
        ue = (i <= nx - 3 ? L.uc : L.u)[ij];
        uw = (i - 1 >= 1 ? L.uc : L.u)[ij - 1];
        vn = (j <= L.npy - 3 ? L.vc : L.v)[ij];
        vs = (j - 1 >= 1 ? L.vc : L.v)[ijs];
    } else {
        const double u0 = L.u[ij], u1 = L.u[ij - 1], v0 = L.v[ij], v1 = L.v[ijs];
        const float du0 = L.du32[ij], du1 = L.du32[ij - 1], dv0 = L.dv32[ij], dv1 = L.dv32[ijs];
        const uint8_t f0 = L.flags[ij], fw = L.flags[ij - 1], fs = L.flags[ijs];
        ue = i <= nx - 3 ? pending_value(u0, du0, L.pset && (f0 & MGF_F_SPU), 0.0) : u0;
        uw = i - 1 >= 1 ? pending_value(u1, du1, L.pset && (fw & MGF_F_SPU), 0.0) : u1;
        vn = j <= L.npy - 3 ? pending_value(v0, dv0, L.pset && (f0 & MGF_F_SPV), 0.0) : v0;
        vs = j - 1 >= 1 ? pending_value(v1, dv1, L.pset && (fs & MGF_F_SPV), 0.0) : v1;
    }
    double div = ((L.ax * ue - L.ax * uw) + L.ay * vn) - L.ay * vs;
    double su = -div;
    mag = fabs(div);
    /* the p' equation is only ever solved in FP32, so its coefficients
     * (coefp's area^2 / ap of the faces) are formed in FP32 */
    const float axax = (float)L.axax, ayay = (float)L.ayay;
    const float aw = axax * rcp32((float)L.apu[ij - 1]);
    const float ae = axax * rcp32((float)L.apu[ij]);
    const float as = ayay * rcp32((float)L.apv[ijs]);
    const float an = ayay * rcp32((float)L.apv[ij]);
    float ap = (((aw + ae) + as) + an) + ((L.flags[ij] & MGF_F_OBS) ? 1.0e30f : 0.0f);
    const float diag = 1.0e30f;
    if (P.ibc[0] == 2 && i == 1) { ap = ap + diag; su = 0.0; }
    if (P.ibc[1] == 2 && i == nx - 2) { ap = ap + diag; su = 0.0; }
    if (P.ibc[2] == 2 && j == 1) { ap = ap + diag; su = 0.0; }
    if (P.ibc[3] == 2 && j == L.npy - 2) { ap = ap + diag; su = 0.0; }
    return Pt32{aw, ae, as, an, rcp_diag(ap, L.isolve), (float)su};
}

/* assembly of u, v or sc over the swept range, stored for the sweeps */
__global__ void __launch_bounds__(BX * BY) k_asm_mx(const DevLevel *Ls, int field)
{
    LEVEL(Ls);
    const int nx = L.npx;
    int i0 = 1, i1 = nx - 2, j1 = L.npy - 2;
    if (field == 0) i1 = nx - 3;
    if (field == 1) j1 = L.npy - 3;
    RANGE2D(i0, i1, 1, j1);
    const int ij = j * nx + i;
    double ap;
    const Pt32 q = asm_pt(L, field, i, j, ap);
    if (field == 0) L.apu[ij] = ap;
    if (field == 1) L.apv[ij] = ap;
    store_pt(L, ij, q);
}

/* assembly of p' over the cells, stored for the sweeps; partial sums of
 * |div| in FP64 per 64 x 4 block for the deterministic reduction */
__global__ void __launch_bounds__(BX * BY) k_asm_p_mx(const DevLevel *Ls)
{
    LEVEL(Ls);
    __shared__ double lds[BY];
    const int i = 1 + (int)(T.x * BX + threadIdx.x);
    const int j = 1 + (int)(T.y * BY + threadIdx.y);
    const int nx = L.npx;
    double mag = 0.0;
    if (i <= nx - 2 && j <= L.npy - 2) store_pt(L, j * nx + i, asm_pt_p(L, i, j, mag));
    double bs = block_sum(mag, lds);
    const int pb = T.y * (int)gridDim.x + T.x;     /* padding rows have none */
    if (threadIdx.x == 0 && threadIdx.y == 0 && pb < L.nblk) L.partial[pb] = bs;
}

/* extents of the swept range of a field (the reference's solver arguments) */
__device__ __forceinline__ void extents(const DevLevelG &L, int field, int &nx, int &ny)
{
    (void)field_of(L, field, nx, ny);
}

/* S FP32 red-black sweeps (2S colour steps) of A d = r in one launch,
 * temporally blocked.  The workgroup takes its 64 x TY tile with a halo of
 * H = 2S points on every side (equation in registers, d in LDS); colour step
 * m updates the points at distance >= m from the edge of that region, so
 * after 2S steps the tile itself holds exactly what S single sweeps give
 * (same expressions, same order: the result is bit-identical).  d crosses
 * HBM once per launch instead of once per sweep.  d is read from buffer
 * launch&1 (zero for launch 0 and outside the swept range) and written to
 * the other buffer, or, for p' on the last launch, widened into pp.
 *
 * MODE_LOAD: the equation comes from the FP32 arrays of k_asm_mx /
 * k_asm_p_mx.  MODE_ASM (launch 0 only): the kernel assembles it itself in
 * FP64 from the state, for the halo too (recomputed by each workgroup that
 * needs it), so the FP32 equation never crosses HBM; the tile's apu / apv
 * and, for p', the |div| partial sums (per 64 x 4 block, in block_sum's
 * order, so they equal k_asm_p_mx's bit for bit) are written as by the
 * assembly kernels.  MODE_ASM_STORE: MODE_ASM that also stores the tile's
 * FP32 equation for the launches that follow.  MODE_ASM_INL: MODE_ASM with
 * the face fluxes formed from u and v (valid while u and v are still the
 * values the relaxation started from), and the sc correction left pending in
 * ds32.  Nothing the assembly reads is written by the launch, so halo
 * recomputation is race-free. */
#define MODE_LOAD 0
#define MODE_ASM 1
#define MODE_ASM_STORE 2
#define MODE_ASM_INL 3
/* MODE_ASM for the merged relaxation's p': u + du, v + dv read from uc, vc */
#define MODE_ASM_CORR 4
/* tile rows of the merged u, v, sc launch: taller than the other launches'
 * tiles, as its halo is assembled three times over (once per field) */
#ifndef MGF_UVS_TILE_ROWS
#define MGF_UVS_TILE_ROWS 32
#endif
/* tile rows and threads of the merged launch on small levels (at most
 * MGF_UVS_SMALL_MAX): 16-row tiles (4 slots per thread instead of 6) shorten
 * each workgroup's chain of dependent loads, which is what such a launch
 * waits for (1024 threads, or 32 rows with 1024, were slower; 8 rows no
 * better) */
#ifndef MGF_UVS_SMALL_ROWS
#define MGF_UVS_SMALL_ROWS 16
#endif
#ifndef MGF_UVS_SMALL_THREADS
#define MGF_UVS_SMALL_THREADS 512
#endif
/* tile rows of the launches on a stored equation: 32 (halo 1.63x the tile
 * at three sweeps, against 2.08x on 16 rows) measured 2% faster on 84
 * members of 512 x 512 than 16 with 512 threads, and 24 and 48 slower; with
 * 1024 threads 48 and 64 rows were slower as well */
#ifndef MGF_LOAD_TILE_ROWS
#define MGF_LOAD_TILE_ROWS 32
#endif
/* tile rows of the other assembling launches */
#ifndef MGF_ASM_TILE_ROWS
#define MGF_ASM_TILE_ROWS MGF_SWEEP_TILE_ROWS
#endif
/* p' relaxations of up to this many sweeps run as one assembling launch on
 * tiles of MGF_PP_TILE_ROWS rows (0: the assemble-and-store launch plus
 * launches on the stored equation).  Six sweeps in one launch assemble a
 * 12-point halo (2.4x the tile on 32 rows) but never store and reload the
 * FP32 equation; measured 2% faster on the batch than 3 + 3. */
#ifndef MGF_PP_ONE_LAUNCH
#define MGF_PP_ONE_LAUNCH 6
#endif
#ifndef MGF_PP_TILE_ROWS
#define MGF_PP_TILE_ROWS 32
#endif
/* tile rows of the merged relaxation's p' launch on small levels (16: one
 * 64 x 64 tile per level, assembled once, was slower -- such a launch is
 * latency-bound and fewer, longer workgroups lengthen it) */
#ifndef MGF_SMALL_TILE_ROWS
#define MGF_SMALL_TILE_ROWS MGF_SWEEP_TILE_ROWS
#endif
/* threads of the merged relaxation's p' launch (512 or 1024) */
#ifndef MGF_PP_THREADS
#define MGF_PP_THREADS 512
#endif
template <int S, int TYR = MGF_SWEEP_TILE_ROWS>
struct RbShape {
    static constexpr int TY = TYR;
    static constexpr int H = 2 * S, RW = BX + 2 * H, RH = TY + 2 * H, NR = RW * RH;
};

/* the work of one workgroup on tile T of level L; t: NR floats of LDS, mg
 * and rows: the |div| rows of an assembling p' launch (TY x BX and TY
 * doubles; unused otherwise) */
template <int S, int NT, int MODE, int FIELD, int TYR = MGF_SWEEP_TILE_ROWS>
__device__ __forceinline__ void rb32x_tile(const DevLevelG &L, const TileId T, int field_rt,
                                           int launch, int last, float *t, double *mg,
                                           double *rows)
{
    /* FIELD >= 0: the field is fixed at compile time (assembling launches);
     * -1: taken from field_rt */
    const int field = FIELD >= 0 ? FIELD : field_rt;
    using Sh = RbShape<S, TYR>;
    constexpr int TY = Sh::TY;
    constexpr int H = Sh::H, RW = Sh::RW, RH = Sh::RH;
    constexpr int NR = Sh::NR;
    constexpr int NPT = (NR + NT - 1) / NT, NW = NT / BX;
    constexpr bool ASM = MODE != MODE_LOAD;
    constexpr bool INL = MODE == MODE_ASM_INL;
    constexpr bool CORR = MODE == MODE_ASM_CORR;
    constexpr int MR = ASM ? TY : 1;
    int nx, ny;
    extents(L, field, nx, ny);
    const gflt *src = launch == 0 ? nullptr : ((launch & 1) ? L.db32 : L.da32);
    gflt *dst = (launch & 1) ? L.da32 : L.db32;
    /* the last launch of u / v leaves the correction pending in du32 / dv32
     * (the merged launch: u + du, v + dv in uc, vc, written below, and sc's
     * correction in ds32 for the update to apply) */
    if (last && field == 0) dst = L.du32;
    if (last && field == 1) dst = L.dv32;
    if (INL && field == 3) dst = L.ds32;
    const bool pfield = field == 2;
    const bool stash = pfield && (last & 2) != 0;
    const int npx = L.npx;
    const int ri0 = 1 + T.x * BX - H, rj0 = 1 + T.y * TY - H;
    const int tid = threadIdx.y * BX + threadIdx.x;
    if (ASM && pfield) {
        for (int k = tid; k < MR * BX; k += NT) mg[k] = 0.0;
        __syncthreads();
    }
    Pt32 pt[NPT];
    /* the tile's FP64 diagonals of u / v (stored as apu / apv at the end:
     * no global store between the assembly loads) */
    double apk[ASM ? NPT : 1];
    /* per slot: distance to the region's edge (bits 0-7), colour (bit 8),
     * inside the swept range (bit 9), a copy origin whose updated p is
     * stashed (bit 10), a penalised u / v face of the merged launch
     * (bit 11) */
    unsigned meta[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        /* pt[k] and apk[k] are set only for points inside the swept range
         * (meta bit 9) and read only for those: no initial value, which the
         * compiler would otherwise materialise in every branch */
        const int p = tid + k * NT;
        float val = 0.f;
        unsigned m = 0;
        const int r = p / RW, c = p - r * RW;
        const int gi = ri0 + c, gj = rj0 + r;
        {
            if (p < NR && gi >= 1 && gi <= nx - 1 && gj >= 1 && gj <= ny - 1) {
                const int g = gj * npx + gi;
                const int d = min(min(c, RW - 1 - c), min(r, RH - 1 - r));
                if (ASM) {
                    if (pfield) {
                        double mag;
                        pt[k] = asm_pt_p<CORR>(L, gi, gj, mag);
                        if (d >= H) mg[(r - H) * BX + (c - H)] = mag;
                    } else {
                        pt[k] = asm_pt<INL>(L, field, gi, gj, apk[ASM ? k : 0]);
                        /* the merged launch: a penalised u / v face (the
                         * assembly has loaded its flag already) */
                        if (INL && field != 3)
                            m |= (unsigned)((L.flags[g] &
                                             (field == 0 ? MGF_F_SPU : MGF_F_SPV)) != 0) << 11;
                    }
                } else {
                    pt[k] = load_pt(L, g);
                }
                if (src) val = src[g];
                m |= (unsigned)d | ((unsigned)((gi + gj) & 1) << 8) | (1u << 9);
                if (stash && d >= H && (L.flags[g] & MGF_F_PORIG)) m |= 1u << 10;
            }
            if (p < NR) t[p] = val;
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
                const int p = tid + k * NT;
                const Pt32 &o = pt[k];
                const float s = ((o.aw * t[p - 1] + o.ae * t[p + 1]) + o.as * t[p - RW]) +
                                o.an * t[p + RW];
                t[p] = (s + o.r) * o.iap;
            }
        }
        __syncthreads();
    }
    const bool to_pp = last && pfield;
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const unsigned m = meta[k];
        if ((m & (1u << 9)) && (int)(m & 0xffu) >= H) {
            const int p = tid + k * NT;
            const int r = p / RW, c = p - r * RW;
            const int g = (rj0 + r) * npx + (ri0 + c);
            if (to_pp) {
                const double ppv = (double)t[p];
                L.pp[g] = ppv;
                /* the freshly updated p of a copy origin, for the update's
                 * obstacle-edge copies (what k_pstash would stash) */
                if (m & (1u << 10))
                    L.tmp[g] = L.p[g] + L.P.relxp * (ppv - 0.0);
            } else if (INL && field == 0) {
                /* the merged relaxation: u + du formed here, once */
                L.uc[g] = pending_value(L.u[g], t[p], L.pset && (m & (1u << 11)), 0.0);
            } else if (INL && field == 1) {
                L.vc[g] = pending_value(L.v[g], t[p], L.pset && (m & (1u << 11)), 0.0);
            } else {
                dst[g] = t[p];
            }
            if (ASM && field == 0) L.apu[g] = apk[ASM ? k : 0];
            if (ASM && field == 1) L.apv[g] = apk[ASM ? k : 0];
            if (MODE == MODE_ASM_STORE) store_pt(L, g, pt[k]);
        }
    }
    if (ASM && pfield) {
        /* per 64 x 4 block of the tile: each row of the tile is summed by
         * one wave with block_sum's shuffle tree, then the four rows of a
         * block in order */
        const int w = threadIdx.y, lane = threadIdx.x;
#pragma unroll
        for (int row = w; row < TY; row += NW) {
            double v = mg[row * BX + lane];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) rows[row] = v;
        }
        __syncthreads();
        if (tid < TY / BY) {
            const int q = tid;
            const double bs = ((rows[q * BY] + rows[q * BY + 1]) + rows[q * BY + 2]) +
                              rows[q * BY + 3];
            const int gxp = (npx - 2 + BX - 1) / BX;
            const int pb = (T.y * (TY / BY) + q) * gxp + T.x;
            if (pb < L.nblk) L.partial[pb] = bs;
        }
    }
}

template <int S, int NT, int MODE, int FIELD, int TYR = MGF_SWEEP_TILE_ROWS>
__global__ void __launch_bounds__(NT) k_rb32x(const DevLevel *Ls, int field_rt, int launch,
                                              int last)
{
    LEVEL(Ls);
    constexpr int MR = MODE != MODE_LOAD ? TYR : 1;
    __shared__ float t[RbShape<S, TYR>::NR];
    __shared__ double mg[MR * BX];
    __shared__ double rows[MR];
    rb32x_tile<S, NT, MODE, FIELD, TYR>(L, T, field_rt, launch, last, t, mg, rows);
}

/* The u, v and sc relaxations of a group in one launch (S sweeps each, the
 * only launch of each): in mixed precision u and v keep their corrections
 * pending until the update, so none of the three equations reads what
 * another one changes, and the fluxes are formed from u and v.  The sc
 * correction stays pending in ds32 until the update.  The grid is the sc
 * tile grid with three logical z-slices per member (field = z % 3); tiles
 * past u's or v's smaller range hold no unknowns and write nothing. */
template <int S, int NT, int TYU = MGF_UVS_TILE_ROWS>
__global__ void __launch_bounds__(NT) k_rb32x_uvs(const DevLevel *Ls)
{
    const TileId T3 = tile_id();
    const TileId T{T3.x, T3.y, T3.z / 3};
    const int f = T3.z - 3 * T.z;
    const DevLevelG &L = reinterpret_cast<const DevLevelG *>(Ls)[T.z];
    if (L.ctl->done) return;
    __shared__ float t[RbShape<S, TYU>::NR];
    __shared__ double mg[BX];
    __shared__ double rows[1];
    if (f == 0)
        rb32x_tile<S, NT, MODE_ASM_INL, 0, TYU>(L, T, 0, 0, 1, t, mg, rows);
    else if (f == 1)
        rb32x_tile<S, NT, MODE_ASM_INL, 1, TYU>(L, T, 1, 0, 1, t, mg, rows);
    else
        rb32x_tile<S, NT, MODE_ASM_INL, 3, TYU>(L, T, 3, 0, 1, t, mg, rows);
}

/* one FP32 Jacobi sweep of A d = r (wj = 1, V6:1241) */
__global__ void __launch_bounds__(BX * BY) k_jac32(const DevLevel *Ls, int field, int it)
{
    LEVEL(Ls);
    int nx, ny;
    extents(L, field, nx, ny);
    RANGE2D(1, nx - 1, 1, ny - 1);
    const gflt *src = it == 0 ? nullptr : ((it & 1) ? L.db32 : L.da32);
    gflt *dst = (it & 1) ? L.da32 : L.db32;
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

/* the result of nl sweep launches (its parity names the buffer holding d)
 * over the swept range: sc += d in FP64 (penalised cells set to sc_obs:
 * pending_value; nothing without sweeps); pp = d in FP64; for u and v the
 * correction is left pending in du32 / dv32 (zero without sweeps) */
__global__ void __launch_bounds__(BX * BY) k_apply_mx(const DevLevel *Ls, int field, int nl)
{
    LEVEL(Ls);
    int nx, ny;
    gdbl *phi = field_of(L, field, nx, ny);
    RANGE2D(1, nx - 1, 1, ny - 1);
    const gflt *d = (nl & 1) ? L.db32 : L.da32;
    const int ij = j * L.npx + i;
    if (field == 0)
        L.du32[ij] = nl > 0 ? d[ij] : 0.f;
    else if (field == 1)
        L.dv32[ij] = nl > 0 ? d[ij] : 0.f;
    else if (field == 2)
        phi[ij] = nl > 0 ? (double)d[ij] : 0.0;
    else
        phi[ij] = nl > 0 ? pending_value(phi[ij], d[ij], L.pset && (L.flags[ij] & MGF_F_SPSC), L.P.sc_obs)
                         : phi[ij];
}

/* ====================================================================== */

void mgf_launch_reduce_err(const DevGroup &G, int record, hipStream_t st);


void mgf_launch_asm_mixed(const DevGroup &G, int field, int err, hipStream_t st)
{
    if (field == 2) {
        hipLaunchKernelGGL(k_asm_p_mx, tile_grid(G.npx - 2, G.npy - 2, G.count), dim3(BX, BY), 0, st,
                           G.d);
        if (err) mgf_launch_reduce_err(G, err == 2, st);
        return;
    }
    int ni = G.npx - 2, nj = G.npy - 2;
    if (field == 0) ni = G.npx - 3;
    if (field == 1) nj = G.npy - 3;
    if (ni < 1 || nj < 1) return;
    hipLaunchKernelGGL(k_asm_mx, tile_grid(ni, nj, G.count), dim3(BX, BY), 0, st, G.d, field);
}

template <int S, int MODE, int FIELD>
static void launch_rbx_f(const DevGroup &G, int field, int launch, int last, int nx, int ny,
                         int nt, hipStream_t st)
{
    /* assembling launches may take taller tiles (their halo costs an
     * assembly per point); the launches on the stored equation keep theirs */
    constexpr int TYR = MODE == MODE_LOAD ? MGF_LOAD_TILE_ROWS : MGF_ASM_TILE_ROWS;
    const dim3 g = tile_grid(nx - 1, ny - 1, G.count, TYR);
    if (nt == 1024)
        hipLaunchKernelGGL((k_rb32x<S, 1024, MODE, FIELD, TYR>), g, dim3(BX, 16), 0, st, G.d,
                           field, launch, last);
    else if (nt == 512)
        hipLaunchKernelGGL((k_rb32x<S, 512, MODE, FIELD, TYR>), g, dim3(BX, 8), 0, st, G.d,
                           field, launch, last);
    else
        hipLaunchKernelGGL((k_rb32x<S, 256, MODE, FIELD, TYR>), g, dim3(BX, 4), 0, st, G.d,
                           field, launch, last);
}

/* sweeps on the stored equation (any field) */
static void launch_load(const DevGroup &G, int field, int launch, int last, int n, int nx, int ny,
                        int nt, hipStream_t st)
{
    switch (n) {
    case 1: launch_rbx_f<1, MODE_LOAD, -1>(G, field, launch, last, nx, ny, nt, st); break;
    case 2: launch_rbx_f<2, MODE_LOAD, -1>(G, field, launch, last, nx, ny, nt, st); break;
    case 3: launch_rbx_f<3, MODE_LOAD, -1>(G, field, launch, last, nx, ny, nt, st); break;
    case 4: launch_rbx_f<4, MODE_LOAD, -1>(G, field, launch, last, nx, ny, nt, st); break;
    case 5: launch_rbx_f<5, MODE_LOAD, -1>(G, field, launch, last, nx, ny, nt, st); break;
    default: launch_rbx_f<6, MODE_LOAD, -1>(G, field, launch, last, nx, ny, nt, st); break;
    }
}

template <int S, int MODE>
static void launch_asm_s(const DevGroup &G, int field, int last, int nx, int ny, int nt,
                         hipStream_t st)
{
    switch (field) {
    case 0: launch_rbx_f<S, MODE, 0>(G, field, 0, last, nx, ny, nt, st); break;
    case 1: launch_rbx_f<S, MODE, 1>(G, field, 0, last, nx, ny, nt, st); break;
    case 2: launch_rbx_f<S, MODE, 2>(G, field, 0, last, nx, ny, nt, st); break;
    default: launch_rbx_f<S, MODE, 3>(G, field, 0, last, nx, ny, nt, st); break;
    }
}

/* the first, assembling launch of a relaxation: n <= MGF_MAX_FUSED_SWEEPS
 * sweeps; store: later launches follow and need the stored equation */
static void launch_asm(const DevGroup &G, int field, int last, bool store, int n, int nx, int ny,
                       int nt, hipStream_t st)
{
    if (store) {
        switch (n) {
        case 1: launch_asm_s<1, MODE_ASM_STORE>(G, field, last, nx, ny, nt, st); break;
        case 2: launch_asm_s<2, MODE_ASM_STORE>(G, field, last, nx, ny, nt, st); break;
        default: launch_asm_s<3, MODE_ASM_STORE>(G, field, last, nx, ny, nt, st); break;
        }
    } else {
        switch (n) {
        case 1: launch_asm_s<1, MODE_ASM>(G, field, last, nx, ny, nt, st); break;
        case 2: launch_asm_s<2, MODE_ASM>(G, field, last, nx, ny, nt, st); break;
        default: launch_asm_s<3, MODE_ASM>(G, field, last, nx, ny, nt, st); break;
        }
    }
}

static void sweep_extents(const DevGroup &G, int field, int &nx, int &ny)
{
    switch (field) {
    case 0: nx = G.npx - 2; ny = G.npy - 1; break;
    case 1: nx = G.npx - 1; ny = G.npy - 2; break;
    default: nx = G.npx - 1; ny = G.npy - 1; break;
    }
}

SweepCfg mgf_sweep_cfg(int sweeps_per_launch, int threads, int split_assembly)
{
    SweepCfg c;
    c.spl = sweeps_per_launch <= 0 ? MGF_AUTO_SWEEPS_PER_LAUNCH : sweeps_per_launch;
    if (c.spl > MGF_MAX_SWEEPS_PER_LAUNCH) c.spl = MGF_MAX_SWEEPS_PER_LAUNCH;
    const bool given = threads == 256 || threads == 512 || threads == 1024;
    c.nt = given ? threads : MGF_AUTO_SWEEP_THREADS;
    c.nt_asm = given ? threads : MGF_AUTO_ASM_THREADS;
    c.fuse = split_assembly ? 0 : 1;
    return c;
}

/* nsw sweeps on the stored FP32 equation, as launches of at most spl sweeps
 * (red-black) or one sweep each (Jacobi); returns the number of launches,
 * whose parity says which buffer holds the result */
int mgf_launch_sweeps_mixed(const DevGroup &G, int field, int isolve, int nsw, SweepCfg cfg,
                            hipStream_t st)
{
    int nx, ny;
    sweep_extents(G, field, nx, ny);
    if (nx - 1 < 1 || ny - 1 < 1 || (isolve != 1 && isolve != 2)) return 0;
    int launch = 0;
    if (isolve == 1) {
        for (; launch < nsw; ++launch)
            hipLaunchKernelGGL(k_jac32, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st,
                               G.d, field, launch);
        return launch;
    }
    for (int done = 0; done < nsw; ++launch) {
        const int n = nsw - done < cfg.spl ? nsw - done : cfg.spl;
        launch_load(G, field, launch, 0, n, nx, ny, cfg.nt, st);
        done += n;
    }
    return launch;
}

/* A level whose interior fits in one tile column and four tile rows (the
 * coarse levels of a hierarchy): its launches are latency-bound, so they
 * take 1024-thread workgroups and all sweeps of a relaxation in the one
 * assembling launch (its halo then spans most of the level anyway). */
static bool small_group(const DevGroup &G)
{
    return G.npx - 2 <= BX && G.npy - 2 <= 4 * MGF_SWEEP_TILE_ROWS;
}

/* levels of at most n x n interior points */
static bool level_at_most(const DevGroup &G, int n)
{
    return G.npx - 2 <= n && G.npy - 2 <= n;
}
/* the merged relaxation takes the small-level shapes (16-row tiles; for p'
 * 1024 threads) up to these sizes: on 128 x 128 levels too they shorten the
 * latency-bound workgroups (2.5% on the batch against 64; 256 was slower) */
#ifndef MGF_UVS_SMALL_MAX
#define MGF_UVS_SMALL_MAX 128
#endif
#ifndef MGF_PP_SMALL_MAX
#define MGF_PP_SMALL_MAX 128
#endif

template <int S>
static void launch_small_s(const DevGroup &G, int field, int last, int nx, int ny,
                           hipStream_t st)
{
    const dim3 g = tile_grid(nx - 1, ny - 1, G.count, MGF_SWEEP_TILE_ROWS);
    const dim3 b(BX, 16);
    switch (field) {
    case 0: hipLaunchKernelGGL((k_rb32x<S, 1024, MODE_ASM, 0>), g, b, 0, st, G.d, 0, 0, last); break;
    case 1: hipLaunchKernelGGL((k_rb32x<S, 1024, MODE_ASM, 1>), g, b, 0, st, G.d, 1, 0, last); break;
    case 2: hipLaunchKernelGGL((k_rb32x<S, 1024, MODE_ASM, 2>), g, b, 0, st, G.d, 2, 0, last); break;
    default: hipLaunchKernelGGL((k_rb32x<S, 1024, MODE_ASM, 3>), g, b, 0, st, G.d, 3, 0, last); break;
    }
}

/* the launches of a fused red-black relaxation of one field (the first one
 * assembles); lastbits: the `last` argument of the final launch (bit 0:
 * final; bit 1, p' only: also stash the updated p of copy origins in tmp).
 * Returns the number of launches. */
template <int S>
static void launch_pp_one_s(const DevGroup &G, int last, int nx, int ny, int nt, hipStream_t st)
{
    const dim3 g = tile_grid(nx - 1, ny - 1, G.count, MGF_PP_TILE_ROWS);
    if (nt == 1024)
        hipLaunchKernelGGL((k_rb32x<S, 1024, MODE_ASM, 2, MGF_PP_TILE_ROWS>), g, dim3(BX, 16), 0,
                           st, G.d, 2, 0, last);
    else
        hipLaunchKernelGGL((k_rb32x<S, 512, MODE_ASM, 2, MGF_PP_TILE_ROWS>), g, dim3(BX, 8), 0,
                           st, G.d, 2, 0, last);
}

static int launch_fused(const DevGroup &G, int field, int nsw, SweepCfg cfg, int lastbits,
                        hipStream_t st)
{
    int nx, ny;
    sweep_extents(G, field, nx, ny);
    if (field == 2 && nsw <= MGF_PP_ONE_LAUNCH && nsw > MGF_MAX_FUSED_SWEEPS &&
        !small_group(G)) {
        const int nt = cfg.nt_asm == 1024 ? 1024 : 512;
        switch (nsw) {
        case 4: launch_pp_one_s<4>(G, lastbits, nx, ny, nt, st); break;
        case 5: launch_pp_one_s<5>(G, lastbits, nx, ny, nt, st); break;
        default: launch_pp_one_s<6>(G, lastbits, nx, ny, nt, st); break;
        }
        return 1;
    }
    if (small_group(G) && nsw <= MGF_MAX_SWEEPS_PER_LAUNCH) {
        switch (nsw) {
        case 1: launch_small_s<1>(G, field, lastbits, nx, ny, st); break;
        case 2: launch_small_s<2>(G, field, lastbits, nx, ny, st); break;
        case 3: launch_small_s<3>(G, field, lastbits, nx, ny, st); break;
        case 4: launch_small_s<4>(G, field, lastbits, nx, ny, st); break;
        case 5: launch_small_s<5>(G, field, lastbits, nx, ny, st); break;
        default: launch_small_s<6>(G, field, lastbits, nx, ny, st); break;
        }
        return 1;
    }
    int launch = 0;
    for (int done = 0; done < nsw; ++launch) {
        int lim = cfg.spl;
        if (launch == 0 && lim > MGF_MAX_FUSED_SWEEPS) lim = MGF_MAX_FUSED_SWEEPS;
        const int n = nsw - done < lim ? nsw - done : lim;
        const int last = done + n == nsw ? lastbits : 0;
        if (launch > 0)
            launch_load(G, field, launch, last, n, nx, ny, cfg.nt, st);
        else
            launch_asm(G, field, last, !last, n, nx, ny, cfg.nt_asm, st);
        done += n;
    }
    return launch;
}

/* smoothing of one field in a relaxation: assembly, nsw sweeps and the
 * update of the FP64 unknown (field 0 u, 1 v, 2 p', 3 sc).  Red-black with
 * cfg.fuse: the first sweep launch assembles the equation itself, and for p'
 * the last launch writes pp; otherwise (Jacobi, no sweeps, or split
 * assembly) the assembly kernel stores the FP32 equation first.  err: the
 * level's sum |div| is reduced after the p' assembly (1), and is also the
 * history entry of level-0 members (2), or is not needed (0). */
void mgf_launch_relax_field_mixed(const DevGroup &G, int field, int isolve, int nsw,
                                  SweepCfg cfg, int err, hipStream_t st)
{
    int nx, ny;
    sweep_extents(G, field, nx, ny);
    if (isolve != 2 || nsw < 1 || !cfg.fuse || nx - 1 < 1 || ny - 1 < 1) {
        mgf_launch_asm_mixed(G, field, err, st);
        mgf_launch_smooth_mixed(G, field, isolve, nsw, cfg, st);
        return;
    }
    const int launch = launch_fused(G, field, nsw, cfg, 1, st);
    /* p': the last launch wrote pp; u, v: the last launch left the
     * correction pending in du32 / dv32; sc: apply it */
    if (field == 2) {
        if (err) mgf_launch_reduce_err(G, err == 2, st);
    } else if (field == 3) {
        hipLaunchKernelGGL(k_apply_mx, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st,
                           G.d, field, launch);
    }
}

template <int S>
static void launch_uvs_s(const DevGroup &G, int nt, hipStream_t st)
{
    if (G.nt_asm) nt = G.nt_asm;
    const bool small = G.uvs_shape ? G.uvs_shape == 1 : level_at_most(G, MGF_UVS_SMALL_MAX);
    if (small) {
        /* small levels: latency-bound, shaped by their own tile rows and
         * threads */
        const dim3 g0 = tile_grid(G.npx - 2, G.npy - 2, G.count, MGF_UVS_SMALL_ROWS);
        const dim3 g(g0.x, g0.y, 3 * g0.z);
        hipLaunchKernelGGL((k_rb32x_uvs<S, MGF_UVS_SMALL_THREADS, MGF_UVS_SMALL_ROWS>), g,
                           dim3(BX, MGF_UVS_SMALL_THREADS / BX), 0, st, G.d);
        return;
    }
    const dim3 g0 = tile_grid(G.npx - 2, G.npy - 2, G.count, MGF_UVS_TILE_ROWS);
    const dim3 g(g0.x, g0.y, 3 * g0.z);
    if (nt == 1024)
        hipLaunchKernelGGL((k_rb32x_uvs<S, 1024>), g, dim3(BX, 16), 0, st, G.d);
    else if (nt == 512)
        hipLaunchKernelGGL((k_rb32x_uvs<S, 512>), g, dim3(BX, 8), 0, st, G.d);
    else
        hipLaunchKernelGGL((k_rb32x_uvs<S, 256>), g, dim3(BX, 4), 0, st, G.d);
}

int mgf_launch_uvs_mixed(const DevGroup &G, int isolve, int nswm, SweepCfg cfg, hipStream_t st)
{
    const int lim = cfg.spl < MGF_MAX_FUSED_SWEEPS ? cfg.spl : MGF_MAX_FUSED_SWEEPS;
    if (isolve != 2 || !cfg.fuse || nswm < 1 || nswm > lim || G.any_outflow || G.npx - 3 < 1 ||
        G.npy - 3 < 1)
        return 0;
    switch (nswm) {
    case 1: launch_uvs_s<1>(G, cfg.nt_asm, st); break;
    case 2: launch_uvs_s<2>(G, cfg.nt_asm, st); break;
    default: launch_uvs_s<3>(G, cfg.nt_asm, st); break;
    }
    return 1;
}

/* the merged relaxation's p': one assembling launch with all S sweeps,
 * reading u + du, v + dv from uc, vc, writing pp and the stash of the copy
 * origins (last = 3) */
template <int S>
static void launch_pp_merged_s(const DevGroup &G, hipStream_t st)
{
    int nx, ny;
    sweep_extents(G, 2, nx, ny);
    const bool small = G.pp_shape ? G.pp_shape == 1 : level_at_most(G, MGF_PP_SMALL_MAX);
    if (small) {
        const dim3 g = tile_grid(nx - 1, ny - 1, G.count, MGF_SMALL_TILE_ROWS);
        hipLaunchKernelGGL((k_rb32x<S, 1024, MODE_ASM_CORR, 2, MGF_SMALL_TILE_ROWS>), g,
                           dim3(BX, 16), 0, st, G.d, 2, 0, 3);
    } else {
        const dim3 g = tile_grid(nx - 1, ny - 1, G.count, MGF_PP_TILE_ROWS);
        hipLaunchKernelGGL((k_rb32x<S, MGF_PP_THREADS, MODE_ASM_CORR, 2, MGF_PP_TILE_ROWS>), g,
                           dim3(BX, MGF_PP_THREADS / BX), 0, st, G.d, 2, 0, 3);
    }
}

int mgf_launch_pp_merged_mixed(const DevGroup &G, int nswp, hipStream_t st)
{
    if (nswp < 1 || nswp > MGF_MAX_SWEEPS_PER_LAUNCH) return 0;
    switch (nswp) {
    case 1: launch_pp_merged_s<1>(G, st); break;
    case 2: launch_pp_merged_s<2>(G, st); break;
    case 3: launch_pp_merged_s<3>(G, st); break;
    case 4: launch_pp_merged_s<4>(G, st); break;
    case 5: launch_pp_merged_s<5>(G, st); break;
    default: launch_pp_merged_s<6>(G, st); break;
    }
    return 1;
}

int mgf_launch_relax_merged_mixed(const DevGroup &G, int isolve, int nswm, int nswp,
                                  SweepCfg cfg, int err, const DevPair *R, hipStream_t st)
{
    if (nswp < 1 || nswp > MGF_MAX_SWEEPS_PER_LAUNCH ||
        !mgf_launch_uvs_mixed(G, isolve, nswm, cfg, st))
        return 0;
    (void)mgf_launch_pp_merged_mixed(G, nswp, st);
    if (err) mgf_launch_reduce_err(G, err == 2, st);
    if (R && !G.any_pbound_x && !G.any_pbound_y) {
        mgf_launch_update_restv(G, *R, 2, st);
        return 2;
    }
    mgf_launch_update(G, 2, st);
    return 1;
}

/* the first launch of a fused red-black relaxation of a field on its own
 * (for timing: it only rewrites what it wrote before, given unchanged
 * inputs); *n_out = its sweeps.  Returns the compulsory HBM bytes per
 * unknown: the FP64 inputs of the assembly read once (u: cx, cy, p, u, resu;
 * v likewise; sc: cx, cy, sc, ressc; p': u, v, apu, apv), the flag byte, d
 * written in FP32 (p' alone: pp in FP64), apu / apv written for u / v, and
 * the FP32 equation (6 words) when later launches follow. */
double mgf_launch_relax_first_mixed(const DevGroup &G, int field, int nsw, SweepCfg cfg,
                                    int *n_out, hipStream_t st)
{
    int nx, ny;
    sweep_extents(G, field, nx, ny);
    int lim = cfg.spl < MGF_MAX_FUSED_SWEEPS ? cfg.spl : MGF_MAX_FUSED_SWEEPS;
    const bool one = field == 2 && nsw <= MGF_PP_ONE_LAUNCH && nsw > MGF_MAX_FUSED_SWEEPS &&
                     !small_group(G);
    const int n = one ? nsw : (nsw < lim ? nsw : lim);
    const int last = n == nsw;
    *n_out = n;
    if (n < 1 || nx - 1 < 1 || ny - 1 < 1) return 0.0;
    /* as launch_fused issues it (p' written to pp, not to the stash) */
    if (one)
        (void)launch_fused(G, field, nsw, cfg, 1, st);
    else
        launch_asm(G, field, last, !last, n, nx, ny, cfg.nt_asm, st);
    double b = 1.0 + (last ? 0.0 : 24.0);
    switch (field) {
    case 0:
    case 1: b += 5 * 8.0 + 4.0 + 8.0; break;
    case 2: b += 4 * 8.0 + (last ? 8.0 : 4.0); break;
    default: b += 4 * 8.0 + 4.0; break;
    }
    return b;
}

void mgf_launch_smooth_mixed(const DevGroup &G, int field, int isolve, int nsw, SweepCfg cfg,
                             hipStream_t st)
{
    int nx, ny;
    sweep_extents(G, field, nx, ny);
    if (nx - 1 < 1 || ny - 1 < 1 || (isolve != 1 && isolve != 2)) return;
    const int nl = mgf_launch_sweeps_mixed(G, field, isolve, nsw, cfg, st);
    /* without sweeps the correction is zero: sc stays, pp becomes 0, du / dv
     * become 0 */
    if (nl > 0 || field != 3)
        hipLaunchKernelGGL(k_apply_mx, tile_grid(nx - 1, ny - 1, G.count), dim3(BX, BY), 0, st,
                           G.d, field, nl);
}
