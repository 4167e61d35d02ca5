/*
 * mgf_tangent.hip -- the calls of include/mgfvm_tangent.h, built only into
 * libmgfvm_tangent.so (the FP64 path compiled with real_t = mgf_dual).
 *
 * The sensitivity library runs the FP64 path only: mgf_batch_create rejects
 * MGF_PREC_MIXED there, so the mixed-precision launchers the driver names
 * are never reached; the definitions at the end of this file stand in for
 * them and report the misuse if one is reached anyway.
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
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
