/*
 * mgf_tangent.hip -- the calls of include/mgfvm_tangent.h, built only into
 * libmgfvm_tangent.so (the FP64 path compiled with real_t = mgf_dual).
 *
 * The sensitivity library runs the FP64 path only: mgf_batch_create rejects
 * MGF_PREC_MIXED there, so the mixed-precision launchers that the V-cycle
 * code in mgf_solver.hip names are never reached; the definitions at the end of this file stand in for
 * them and report the misuse if one is reached anyway.
 */
#include <hip/hip_runtime.h>

#include <stdio.h>

#include <algorithm>
#include <vector>

#include "mgf_dev.h"
#include "mgfvm_tangent.h"

extern "C" int mgf_tangent_batch_create(const mgf_problem *const *probs, const mgf_seed *seeds,
                                        int n, const mgf_options *opt, mgf_solver **out)
{
    if (!seeds) {
        mgf_set_error(NULL, MGF_ERR_ARG, "seeds missing");
        return MGF_ERR_ARG;
    }
    return mgf_batch_create_seeded(probs, n, opt, out, seeds);
}

static int fetch(mgf_solver *s, int b, int ig, const char *name, std::vector<mgf_dual> &h,
                 int *npx, int *npy)
{
    real_t *p;
    size_t n;
    hipStream_t st;
    int rc = mgf_field_array(s, b, ig, name, &p, &n, &st);
    if (rc) return rc;
    h.resize(n);
    if (hipMemcpyAsync(h.data(), p, n * sizeof(mgf_dual), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return mgf_set_error(s, MGF_ERR_DEVICE, "field download failed");
    if (npx) mgf_batch_level_dims(s, b, ig, npx, npy);
    return MGF_OK;
}

extern "C" int mgf_tangent_get_field(mgf_solver *s, int b, int ig, const char *name,
                                     double *value, double *derivative)
{
    std::vector<mgf_dual> h;
    int rc = fetch(s, b, ig, name, h, NULL, NULL);
    if (rc) return rc;
    for (size_t k = 0; k < h.size(); ++k) {
        if (value) value[k] = h[k].v;
        if (derivative) derivative[k] = h[k].d;
    }
    return MGF_OK;
}

extern "C" int mgf_tangent_objectives(mgf_solver *s, int b, double J[2], double dJ[2])
{
    if (!J || !dJ) return mgf_set_error(s, MGF_ERR_ARG, "bad arguments");
    std::vector<mgf_dual> u, p, sc;
    int npx = 0, npy = 0, rc;
    if ((rc = fetch(s, b, 0, "u", u, &npx, &npy)) || (rc = fetch(s, b, 0, "p", p, NULL, NULL)) ||
        (rc = fetch(s, b, 0, "sc", sc, NULL, NULL)))
        return rc;
    const mgf_problem *c = mgf_instance_problem(s, b);
    const double ax = c->yl / (double)(npy - 2);
    /* J1: mean p of the first interior column minus that of the last */
    mgf_dual pin(0.0), pout(0.0);
    for (int j = 1; j <= npy - 2; ++j) {
        pin += p[(size_t)j * npx + 1];
        pout += p[(size_t)j * npx + npx - 2];
    }
    const mgf_dual j1 = pin / (double)(npy - 2) - pout / (double)(npy - 2);
    /* J2: areax * u * sc through the east boundary face (upwind sc) */
    mgf_dual j2(0.0);
    for (int j = 1; j <= npy - 2; ++j) {
        const size_t ij = (size_t)j * npx + npx - 2;
        j2 += (ax * u[ij]) * sc[ij];
    }
    J[0] = j1.v;
    J[1] = j2.v;
    dJ[0] = j1.d;
    dJ[1] = j2.d;
    return MGF_OK;
}

extern "C" int mgf_tangent_gradient(const mgf_problem *p, const mgf_options *opt, int max_batch,
                                    double J[2], double *g_inlet, double g_amuf[2],
                                    double g_prsc[2], int *converged)
{
    if (!p || max_batch < 0 || p->nfy < 1) {
        mgf_set_error(NULL, MGF_ERR_ARG, "bad arguments");
        return MGF_ERR_ARG;
    }
    const int nfy = p->nfy, nd = nfy + 2;
    const int chunk = max_batch > 0 && max_batch < nd ? max_batch : nd;
    std::vector<double> unit((size_t)chunk * nfy);
    std::vector<mgf_seed> seeds(chunk);
    std::vector<const mgf_problem *> probs(chunk, p);
    std::vector<int> ncyc(chunk), conv(chunk);
    int all_conv = 1, rc = MGF_OK;
    /* directions 0 .. nfy-1: inlet rows 1 .. nfy; nfy: amuf; nfy + 1: prsc */
    for (int d0 = 0; d0 < nd && rc == MGF_OK; d0 += chunk) {
        const int n = nd - d0 < chunk ? nd - d0 : chunk;
        std::fill(unit.begin(), unit.end(), 0.0);
        for (int b = 0; b < n; ++b) {
            const int d = d0 + b;
            mgf_seed &sd = seeds[b];
            sd = mgf_seed{0.0, 0.0, NULL, 0.0};
            if (d < nfy) {
                unit[(size_t)b * nfy + d] = 1.0;
                sd.d_inlet_u = &unit[(size_t)b * nfy];
                sd.d_u_xp = 1.0 / (double)nfy;
            } else if (d == nfy) {
                sd.d_amuf = 1.0;
            } else {
                sd.d_prsc = 1.0;
            }
        }
        mgf_solver *s = NULL;
        if ((rc = mgf_tangent_batch_create(probs.data(), seeds.data(), n, opt, &s))) break;
        rc = mgf_batch_solve(s, ncyc.data(), conv.data());
        for (int b = 0; b < n && rc == MGF_OK; ++b) {
            double Jb[2], dJ[2];
            if ((rc = mgf_tangent_objectives(s, b, Jb, dJ))) break;
            const int d = d0 + b;
            if (d == 0 && J) { J[0] = Jb[0]; J[1] = Jb[1]; }
            if (d < nfy && g_inlet) { g_inlet[d] = dJ[0]; g_inlet[nfy + d] = dJ[1]; }
            if (d == nfy && g_amuf) { g_amuf[0] = dJ[0]; g_amuf[1] = dJ[1]; }
            if (d == nfy + 1 && g_prsc) { g_prsc[0] = dJ[0]; g_prsc[1] = dJ[1]; }
            all_conv &= conv[b];
        }
        if (rc) {
            /* keep the solver's message for mgf_error_string */
            mgf_set_error(NULL, rc, mgf_last_error(s));
        }
        mgf_destroy(s);
    }
    if (converged) *converged = rc == MGF_OK && all_conv;
    return rc;
}

/* ---- mixed-precision launchers: not part of this library --------------- */

static void not_here(const char *what)
{
    fprintf(stderr, "mgfvm (sensitivity library): %s called; mixed precision is not built here\n",
            what);
}

SweepCfg mgf_sweep_cfg(int sweeps_per_launch, int threads, int split_assembly)
{
    (void)sweeps_per_launch;
    (void)threads;
    (void)split_assembly;
    SweepCfg c;
    c.spl = 1;
    c.nt = 256;
    c.nt_asm = 256;
    c.fuse = 0;
    return c;
}

void mgf_launch_relax_field_mixed(const DevGroup &, int, int, int, SweepCfg, int, hipStream_t)
{
    not_here("mgf_launch_relax_field_mixed");
}

int mgf_launch_relax_merged_mixed(const DevGroup &, int, int, int, SweepCfg, int, const DevPair *,
                                  hipStream_t)
{
    not_here("mgf_launch_relax_merged_mixed");
    return 0;
}

int mgf_launch_uvs_mixed(const DevGroup &, int, int, SweepCfg, hipStream_t)
{
    not_here("mgf_launch_uvs_mixed");
    return 0;
}

int mgf_launch_pp_merged_mixed(const DevGroup &, int, hipStream_t)
{
    not_here("mgf_launch_pp_merged_mixed");
    return 0;
}

int mgf_launch_sweeps_mixed(const DevGroup &, int, int, int, SweepCfg, hipStream_t)
{
    not_here("mgf_launch_sweeps_mixed");
    return 0;
}

double mgf_launch_relax_first_mixed(const DevGroup &, int, int, SweepCfg, int *n_out, hipStream_t)
{
    not_here("mgf_launch_relax_first_mixed");
    *n_out = 0;
    return 0.0;
}
