/*
 * mgf_patterns.c -- seeded obstacle patterns in the reference's pattern.txt
 * convention (1-based `if il jf jl` rectangles on the obstacle base level).
 *
 * The algorithms are those of the reference's interactive generators,
 * create_patterns_staggered.c:71-188 (inline, staggered and perturbed
 * arrays) and create_patterns_random.c:58-138 (3x3 / 2x2 / 1x1 particles
 * placed without overlap), plus the Bernoulli mask the benchmark's random
 * configurations use.  The reference seeds rand() with time(NULL); here every
 * draw comes from splitmix64-seeded xoshiro256**, so a pattern is a function
 * of its parameters and seed.  The generators' first grid index is the
 * solver's i (x), as in the reference's output line "r+1 r+1 c+1 c+1", and
 * rectangles are emitted one per blocked cell, i-major then j ascending --
 * the order is part of the input's meaning (the obstacle-edge pressure
 * copies run in list order, Version-6-26-04.c:1153-1177).  Optionally the
 * first and last columns stay free, so the plug inlet and outlet keep
 * matching fluxes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mgf_plan.h"

struct rng {
    uint64_t s[4];
};

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void rng_seed(struct rng *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int k = 0; k < 4; ++k) {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        r->s[k] = z ^ (z >> 31);
    }
}

static uint64_t rng_next(struct rng *r)
{
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform in [0, 1) from 53 random bits */
static double rng_uniform(struct rng *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static int rng_below(struct rng *r, int n) { return (int)(rng_next(r) % (uint64_t)n); }

/* blocked[i*n + j] -> rectangles, i-major then j ascending; returns count */
static int emit(const unsigned char *blocked, int n, int free_edge, int *rects, int cap)
{
    int cnt = 0;
    for (int i = 0; i < n; ++i) {
        if (free_edge && (i == 0 || i == n - 1)) continue;
        for (int j = 0; j < n; ++j) {
            if (!blocked[(size_t)i * n + j]) continue;
            if (rects && cnt < cap) {
                int *q = rects + 4 * cnt;
                q[0] = i + 1; q[1] = i + 1; q[2] = j + 1; q[3] = j + 1;
            }
            ++cnt;
        }
    }
    return cnt;
}

int mgf_pattern_bernoulli(int n_base, double porosity, uint64_t seed, int free_edge_columns,
                          int *rects, int cap)
{
    if (n_base < 1 || !(porosity >= 0.0 && porosity <= 1.0) || cap < 0) {
        mgf_set_error_string("bernoulli pattern: need n_base >= 1, 0 <= porosity <= 1");
        return MGF_ERR_ARG;
    }
    unsigned char *b = (unsigned char *)calloc((size_t)n_base * n_base, 1);
    if (!b) {
        mgf_set_error_string("out of memory");
        return MGF_ERR_LIMIT;
    }
    struct rng r;
    rng_seed(&r, seed);
    /* one draw per cell in i-major order, edge columns included */
    for (int i = 0; i < n_base; ++i)
        for (int j = 0; j < n_base; ++j) b[(size_t)i * n_base + j] = rng_uniform(&r) < (1.0 - porosity);
    int cnt = emit(b, n_base, free_edge_columns, rects, cap);
    free(b);
    return cnt;
}

int mgf_pattern_staggered(int n_base, int spacing_x, int spacing_y, int scale, int variant,
                          uint64_t seed, int free_edge_columns, int *rects, int cap)
{
    if (n_base < 1 || spacing_x < 0 || spacing_y < 0 || (scale != 1 && scale != 2 && scale != 4) ||
        variant < 1 || cap < 0) {
        mgf_set_error_string("staggered pattern: need n_base >= 1, spacings >= 0, scale 1, 2 "
                             "or 4, variant >= 1");
        return MGF_ERR_ARG;
    }
    /* create_patterns_staggered.c:97-151: sizes in sub-cell units */
    const int n = n_base * scale, obj = scale, pad = 5 * scale;
    const int step_x = obj + spacing_x * scale, step_y = obj + spacing_y * scale;
    const int npad = n + 2 * pad, center = obj / 2;
    const int nr = (npad + center) / step_y + 1, nc = (npad + center) / step_x + 1;
    int *ob = (int *)malloc(sizeof(int) * 4 * (size_t)nr * nc);
    unsigned char *g = (unsigned char *)calloc((size_t)n * n, 1);
    if (!ob || !g) {
        free(ob);
        free(g);
        return MGF_ERR_LIMIT;
    }
    int no = 0;
    for (int r = -center; r <= npad; r += step_y)
        for (int c = -center; c <= npad; c += step_x) {
            int *o = ob + 4 * no++;
            o[0] = r - pad; o[1] = r + obj - 1 - pad; o[2] = c - pad; o[3] = c + obj - 1 - pad;
        }
    if (variant >= 2) {              /* staggered: every other column moved half a step */
        const int voff = step_y / 2;
        for (int k = 0; k < no; ++k) {
            int *o = ob + 4 * k;
            if (((o[2] + pad + center) / step_x) % 2 == 1) { o[0] += voff; o[1] += voff; }
        }
    }
    if (variant >= 3) {              /* perturbed: each obstacle stays or moves one unit */
        struct rng rg;
        rng_seed(&rg, seed);
        for (int k = 0; k < no; ++k) {
            const int mv = rng_below(&rg, 5);
            const int dr = mv == 1 ? -1 : (mv == 2 ? 1 : 0), dc = mv == 3 ? -1 : (mv == 4 ? 1 : 0);
            int *o = ob + 4 * k;
            o[0] += dr; o[1] += dr; o[2] += dc; o[3] += dc;
        }
    }
    for (int k = 0; k < no; ++k) {
        const int *o = ob + 4 * k;
        for (int r = o[0]; r <= o[1]; ++r)
            for (int c = o[2]; c <= o[3]; ++c)
                if (r >= 0 && r < n && c >= 0 && c < n) g[(size_t)r * n + c] = 1;
    }
    int cnt = emit(g, n, free_edge_columns, rects, cap);
    free(ob);
    free(g);
    return cnt;
}

int mgf_pattern_particles(int n_grid, double ratio, double pct1, double pct2, double pct3,
                          uint64_t seed, int free_edge_columns, int *rects, int cap)
{
    if (n_grid < 7 || !(ratio >= 0.0 && ratio <= 1.0) || pct1 < 0 || pct2 < 0 || pct3 < 0 ||
        fabs(pct1 + pct2 + pct3 - 100.0) > 0.1 || cap < 0) {
        /* create_patterns_random.c:27-30, 47-50 */
        mgf_set_error_string("particle pattern: need n >= 7, 0 <= ratio <= 1 and size "
                             "percentages summing to 100");
        return MGF_ERR_ARG;
    }
    /* create_patterns_random.c:58-70: particle counts for the target area */
    const double p1 = pct1 / 100.0, p2 = pct2 / 100.0, p3 = pct3 / 100.0;
    const double target_area = (double)n_grid * n_grid * ratio;
    const double avg = p1 * 1.0 + p2 * 4.0 + p3 * 9.0;
    const int total = avg > 0.0 ? (int)(target_area / avg + 0.5) : 0;
    const int want[3] = {(int)(total * p3 + 0.5), (int)(total * p2 + 0.5), (int)(total * p1 + 0.5)};
    const int sizes[3] = {3, 2, 1};
    /* rows (the solver's i) available to particles: all, or all but the edges */
    const int i0 = free_edge_columns ? 1 : 0, ni = free_edge_columns ? n_grid - 2 : n_grid;
    unsigned char *g = (unsigned char *)calloc((size_t)n_grid * n_grid, 1);
    if (!g) return MGF_ERR_LIMIT;
    struct rng r;
    rng_seed(&r, seed);
    int cnt = 0;
    /* create_patterns_random.c:91-135: largest particles first, up to 10000
     * placement attempts each, a particle is dropped when none fits; each
     * placed particle is one rectangle, in placement order */
    for (int s = 0; s < 3; ++s) {
        const int size = sizes[s];
        if (ni - size + 1 < 1) continue;
        for (int k = 0; k < want[s]; ++k) {
            for (int attempt = 0; attempt < 10000; ++attempt) {
                const int ri = i0 + rng_below(&r, ni - size + 1);
                const int cj = rng_below(&r, n_grid - size + 1);
                int overlap = 0;
                for (int a = 0; a < size && !overlap; ++a)
                    for (int b = 0; b < size; ++b)
                        if (g[(size_t)(ri + a) * n_grid + (cj + b)]) { overlap = 1; break; }
                if (overlap) continue;
                for (int a = 0; a < size; ++a)
                    for (int b = 0; b < size; ++b) g[(size_t)(ri + a) * n_grid + (cj + b)] = 1;
                if (rects && cnt < cap) {
                    int *q = rects + 4 * cnt;
                    q[0] = ri + 1; q[1] = ri + size; q[2] = cj + 1; q[3] = cj + size;
                }
                ++cnt;
                break;
            }
        }
    }
    free(g);
    return cnt;
}
