/*
 * mgfvm.h -- C ABI of the MI355X-native SIMPLE / FAS-multigrid steady-state
 * solver for incompressible channel flow with a passive scalar on a
 * staggered structured 2-D grid with obstacle masks.
 *
 * Drop-in boundary.  The reference (akashu4295/Multigrid-FVM,
 * Version-6-26-04.c, "V6" below) has no library interface: its hot path is a
 * set of internal C routines taking `struct SoA*` levels and mutable
 * file-scope globals, driven by main().  Each entry point here replaces one
 * group of those call sites; the citation says which.  Conventions:
 *   - plain C types only; status 0 = success, negative = error;
 *   - the library never calls exit(); the message of the last error is
 *     available from mgf_last_error() (per solver) or mgf_error_string()
 *     (calls without a solver, per thread);
 *   - the library owns device memory, the caller owns host buffers;
 *   - one solver per host thread at a time (externally synchronised); each
 *     solver owns one HIP stream on the device it was created on.
 *
 * Grid conventions (V6 Appendix-A of the reference's own indexing): level ig
 * has npx = nfx/2^ig + 2 by npy = nfy/2^ig + 2 points, row-major
 * ij = j*npx + i.  u[ij] lives on the east face of cell (i,j), v[ij] on its
 * north face; p, pp (pressure correction) and sc are cell centred; ghost
 * layers are i = 0, npx-1 and j = 0, npy-1.
 */
#ifndef MGFVM_H
#define MGFVM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGF_ABI_VERSION 1

/* status codes */
#define MGF_OK               0
#define MGF_ERR_ARG         -1   /* bad argument (null pointer, level out of range, unknown name) */
#define MGF_ERR_IO          -2   /* file cannot be opened (V6:282-286, 2369-2372 exit there) */
#define MGF_ERR_PARSE       -3   /* malformed InputData / pattern.txt (V6:288-392, 2374-2390) */
#define MGF_ERR_CONFIG      -4   /* inconsistent configuration (e.g. ig_obs > ngrid, V6:471 UB) */
#define MGF_ERR_GEOMETRY    -5   /* obstacle/inlet outside the interior, or grid not 2:1 nested */
#define MGF_ERR_LIMIT       -6   /* a size limit of this implementation is exceeded */
#define MGF_ERR_DEVICE      -7   /* HIP runtime error */
#define MGF_ERR_NONFINITE   -8   /* non-finite residual (the reference does not detect it) */
#define MGF_ERR_UNSUPPORTED -9   /* option or configuration this library does not support */

/* Everything InputData and pattern.txt carry (V6:27-39 globals, grammar
 * README.md:7-78).  Arrays are owned by whoever built the struct; the
 * solver copies what it needs at mgf_create(). */
typedef struct mgf_problem {
    int nfx, nfy;              /* fine-grid cells in x and y */
    int ngrid;                 /* multigrid levels */
    int ig_obs;                /* 1-based level on which obstacles are given */
    int ibeg_obs;              /* i offset of the obstacle pattern */
    int ig_inlt, ig_outlt;     /* 1-based levels of discrete inlets/outlets (0: none) */
    int nvcycle;               /* V-cycle limit */
    int nswpm, nswpp;          /* smoother sweeps: momentum+scalar, pressure correction */
    int isolve;                /* 1 Jacobi, 2 red-black Gauss-Seidel (V6:701-736) */
    double omega;              /* read but unused by the reference (V6:1241) */
    double convergence;        /* stop when the level-0 sum |div| falls below this */
    double xl, yl;             /* domain size */
    double amuf;               /* viscosity (1/Re) */
    double prsc;               /* scalar Prandtl/Schmidt number */
    double relxm, relxp;       /* under-relaxation: momentum, pressure */
    int ibc[4];                /* boundary types xm, xp, ym, yp: 1 Dirichlet ghosts,
                                  2 p' Dirichlet row, 3 periodic, 4 (xp) outflow */
    double bc[4][4];           /* ghost values [xm,xp,ym,yp][u,v,p,sc] */
    double guess[4];           /* initial u, v, p, sc */
    double sc_obs;             /* scalar value imposed in obstacles */
    int nobs;                  /* obstacle rectangles */
    const int *obs;            /* nobs x {if, il, jf, jl}, 1-based on level ig_obs-1 */
    int ninlt, noutlt;         /* discrete inlets/outlets */
    const int *inlt_j, *outlt_j;        /* n x {jf, jl} */
    const double *inlt_v, *outlt_v;     /* n x {u, v, sc} */
    double dpdxm, dpdym;       /* body force; never assigned by the reference (0) */
    const double *inlet_u;     /* optional: nfy level-0 inlet u values (rows 1..nfy),
                                  applied after the initial state (extension) */
} mgf_problem;

/* precision modes */
#define MGF_PREC_FP64   0      /* bit-exact with the reference CPU build */
#define MGF_PREC_MIXED  1      /* FP32 smoothers/transfers, FP64 outer residuals */

typedef struct mgf_options {
    int precision;             /* MGF_PREC_* */
    int use_graph;             /* capture one V-cycle in a hipGraph and replay it */
    int device;                /* HIP device ordinal (-1: current) */
    int strict_limits;         /* reject inputs beyond the reference's static limits
                                  (nobs <= 1000, ninlt/noutlt <= 10, nvcycle <= 3000) */
    int sweeps_per_launch;     /* MGF_PREC_MIXED red-black: sweeps fused into one
                                  temporally blocked launch (1-6; 0: automatic).  The
                                  result does not depend on it. */
    int sweep_threads;         /* threads per workgroup of those launches (256, 512 or
                                  1024; 0: automatic); result-neutral as well */
    int split_assembly;        /* MGF_PREC_MIXED red-black: 1 = assemble each equation in
                                  a kernel of its own that stores it in FP32; 0 (default)
                                  = the first sweep launch assembles it on the fly.
                                  Result-neutral. */
    int gs_order;              /* isolve 2 sweeps: 0 (default) red-black, the reference's
                                  solvergs (V6:1297-1351); 1 lexicographic Gauss-Seidel,
                                  a labelled variant the reference does not have: its
                                  loops without the colour tests (j outer, i inner), run
                                  as anti-diagonal wavefronts.  FP64 only, nfy <= 1024. */
    int compact;               /* batched solves with use_graph: once enough instances
                                  have stopped, capture the V-cycle again over the live
                                  ones only (1, default), instead of launching over all
                                  of them with the stopped ones returning at once (0).
                                  Result-neutral. */
    int lanes;                 /* batched solves: the running problems are split into
                                  this many groups ("lanes", balanced by cell count), each
                                  replaying its own V-cycle on its own stream, so one
                                  lane's latency-bound coarse levels and launch tails
                                  overlap another lane's work (1-4; 0: automatic, 3, or
                                  the tuning table's).  Result-neutral. */
    const char *tuning;        /* launch-shape table written by mgf_autotune (text; NULL:
                                  the built-in rules).  Result-neutral. */
} mgf_options;

typedef struct mgf_solver mgf_solver;

/* ---- problem description (host only, no device needed) ---------------- */

/* Read InputData + pattern.txt.  Replaces readInput (V6:280-401) and
 * read_obstacle_data (V6:2367-2400); errors return a status instead of
 * exit().  Free with mgf_problem_free. */
int mgf_problem_read(const char *inputdata_path, const char *pattern_path,
                     mgf_problem **out);
void mgf_problem_free(mgf_problem *p);
void mgf_options_default(mgf_options *o);
const char *mgf_error_string(void);
int mgf_abi_version(void);

/* ---- seeded obstacle patterns (host only) ------------------------------ */

/* Each returns the number of rectangles of the pattern and writes the first
 * `cap` of them to rects (4 ints each: the 1-based `if il jf jl` of a
 * pattern.txt line, on the obstacle base level), or a negative status.
 * Rectangles are one per blocked cell (or particle), in the order the
 * reference's generators write them; the first index is the solver's i (x).
 * free_edge_columns: keep the first and last i column free, so a plug inlet
 * and outlet keep matching fluxes.  The reference seeds rand() with the
 * time; these draw from splitmix64-seeded xoshiro256**. */

/* Every base cell blocked with probability 1 - porosity (one draw per cell,
 * i-major).  No counterpart in the reference; the random configurations of
 * the benchmark are defined this way. */
int mgf_pattern_bernoulli(int n_base, double porosity, uint64_t seed, int free_edge_columns,
                          int *rects, int cap);
/* create_patterns_staggered.c:71-188: 1x1 base obstacles with spacing_x /
 * spacing_y free cells between them on a grid refined by `scale` (4, 2, 1:
 * quarter, half, full cell shifts); variant 1 the inline array (its
 * pattern_1), 2 the staggered array (pattern_2), >= 3 the staggered array
 * with every obstacle staying or moving one sub-cell unit (pattern_3...).
 * The rectangles live on the (n_base * scale)^2 grid. */
int mgf_pattern_staggered(int n_base, int spacing_x, int spacing_y, int scale, int variant,
                          uint64_t seed, int free_edge_columns, int *rects, int cap);
/* create_patterns_random.c:58-138: 3x3, 2x2 and 1x1 particles (pct3, pct2,
 * pct1 percent of the count that fills `ratio` of the n_grid^2 cells),
 * largest first, each placed at a random free spot (up to 10000 attempts,
 * dropped when none fits); one rectangle per particle. */
int mgf_pattern_particles(int n_grid, double ratio, double pct1, double pct2, double pct3,
                          uint64_t seed, int free_edge_columns, int *rects, int cap);

/* ---- solver lifetime -------------------------------------------------- */

/* Build the level hierarchy, initial/boundary state and coarse geometry on
 * the host exactly as AllocateMemorySoA (V6:402-506), init (V6:546-676) and
 * rest_geom (V6:1391-1473) do, derive the device masks, and upload.
 * Replaces V6:184-192 (allocation, setup and the OpenACC enter-data). */
int mgf_create(const mgf_problem *p, const mgf_options *opt, mgf_solver **out);
/* Replaces V6:270-273 (exit-data and free). */
void mgf_destroy(mgf_solver *s);
const char *mgf_last_error(const mgf_solver *s);
int mgf_num_levels(const mgf_solver *s);
int mgf_level_dims(const mgf_solver *s, int ig, int *npx, int *npy);
/* the solver's HIP stream (a hipStream_t), for callers that order work on it */
void *mgf_stream(mgf_solver *s);

/* Restore the initial state built by mgf_create, on the device, without a
 * host transfer (a fresh solve from inputs already resident in HBM).  The
 * reference has no counterpart: it solves once per process.  Asynchronous
 * on the solver's stream. */
int mgf_reset(mgf_solver *s);
/* Wait for all work on the solver's stream. */
int mgf_synchronize(mgf_solver *s);

/* ---- a solve that allocates nothing ------------------------------------ */

/* Size of the static device scratch of mgf_solve_noalloc: one array per
 * device the library runs on, held until the process ends.  It holds every
 * level of one problem of up to 256 x 256 cells (FP64 or mixed). */
#define MGF_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code around graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mgf_create + mgf_solve(p->nvcycle, p->convergence) +
 * mgf_get_fields(level 0) + mgf_destroy in one call that calls no hipMalloc,
 * hipFree or other allocating routine.  Every device array of the solve
 * (all levels' state and scratch, control block, history, launch
 * descriptors) is carved out of the library's static scratch of
 * MGF_NOALLOC_SCRATCH_BYTES; a problem that needs more fails with
 * MGF_ERR_LIMIT before any device work.  V-cycles run eagerly (no graph, one
 * lane) on `stream` (a hipStream_t; NULL: the default stream) on the current
 * device (or opt->device).  Outputs are caller-owned host buffers: history
 * (p->nvcycle entries), the level-0 u, v, p, sc (npx*npy each); any may be
 * NULL.  Results are bit-identical to the allocating calls.  Every call on
 * one device shares that device's scratch and overwrites it (it is cleared
 * where a solve needs zeros): calls that overlap on one device race. */
int mgf_solve_noalloc(const mgf_problem *p, const mgf_options *opt, void *stream,
                      double *history, int *ncycles, int *converged, double *u, double *v,
                      double *pr, double *sc);

/* ---- reference-granular steps (parity and debugging) ------------------ */

/* One SIMPLE relaxation of level ig: fluxes, solve_u, solve_v,
 * solve_pprime, pbound, then outflow+solve_sc (leg 0, down-leg order
 * V6:201-207) or solve_sc+outflow (leg 1, up-leg order V6:237-243).
 * *mass_err (may be NULL) receives the level's sum |div| (V6:1096-1125). */
int mgf_relax(mgf_solver *s, int ig, int leg, double *mass_err);
/* FAS restriction ig -> ig+1: restv, fluxes x2, restru, restrv, restrsc
 * (V6:211-220). */
int mgf_restrict(mgf_solver *s, int ig_fine);
/* Prolongation ig_coarse -> ig_coarse-1: prolu, prolv, prolp, prolsc
 * (V6:230-233). */
int mgf_prolong(mgf_solver *s, int ig_coarse);

/* ---- whole cycles ------------------------------------------------------ */

/* One V-cycle (V6:197-246); *mass_err = level-0 sum |div| recorded in it. */
int mgf_vcycle(mgf_solver *s, int icyc, double *mass_err);
/* V-cycles until the recorded residual is below tol or nvcycle cycles ran
 * (V6:193-249).  history (may be NULL) receives nvcycle entries of which
 * *ncycles are valid; *converged = 1 if the stopping test fired.  Returns
 * MGF_ERR_NONFINITE (after finishing) if a recorded residual was not finite. */
int mgf_solve(mgf_solver *s, int nvcycle, double tol, double *history,
              int *ncycles, int *converged);

/* ---- batches of independent problems ----------------------------------- */

/* One solver for n independent problems (obstacle patterns, Reynolds
 * numbers, grid sizes) that share isolve, nswpm and nswpp.  Every launch of
 * the V-cycle covers the same-depth level (counted from the coarsest) of
 * every problem that has it; each problem keeps the reference's schedule,
 * its own tolerance (convergence) and cycle limit (nvcycle) and stops on its
 * own.  The reference runs one problem per process
 * (run_patterns_and_save_vtk.sh:43-77 loops over patterns); this replaces
 * that loop.  mgf_create() is the n = 1 case; the step-level calls above
 * need n = 1. */
int mgf_batch_create(const mgf_problem *const *problems, int n, const mgf_options *opt,
                     mgf_solver **out);
int mgf_batch_size(const mgf_solver *s);
int mgf_batch_num_levels(const mgf_solver *s, int b);
int mgf_batch_level_dims(const mgf_solver *s, int b, int ig, int *npx, int *npy);
/* Solve every problem; ncycles[n] and converged[n] (either may be NULL). */
int mgf_batch_solve(mgf_solver *s, int *ncycles, int *converged);
/* Copy up to cap history entries of problem b; returns the cycles run. */
int mgf_batch_history(mgf_solver *s, int b, double *history, int cap);
int mgf_batch_get_field(mgf_solver *s, int b, int ig, const char *name, double *out);

/* ---- data ----------------------------------------------------------------- */

/* Copy a level array to/from a caller buffer of npx*npy doubles.  Names:
 * u v p pp sc resu resv ressc apu apv.  Replaces UpdateHostForOutput
 * (V6:172-176) and gives the state access the reference's globals gave. */
int mgf_get_field(mgf_solver *s, int ig, const char *name, double *out);
int mgf_set_field(mgf_solver *s, int ig, const char *name, const double *in);
int mgf_get_fields(mgf_solver *s, int ig, double *u, double *v, double *p,
                   double *pp, double *sc);
/* Replace problem b's inlet u profile (nfy values on the level-0 west face,
 * rows 1..nfy, as mgf_problem.inlet_u; NULL: the InputData inlet) and reset
 * problem b to the initial state built with it, as mgf_create would have
 * built it (init and rest_geom's inlet values, V6:546-676, 1440-1446).
 * b = 0 for a single-problem solver.  With a plug east boundary the caller
 * keeps the profile's mean equal to the east u.  The sensitivity library
 * takes the profile at creation only (MGF_ERR_UNSUPPORTED here). */
int mgf_set_inlet_profile(mgf_solver *s, int b, const double *inlet_u);

/* ---- measurement ------------------------------------------------------------ */

/* Project-written autotuner (no counterpart in the reference).  For a
 * mixed-precision red-black solver with the fused relaxation, measures on
 * the device, with HIP events, the result-neutral launch shapes of the
 * merged relaxation on every level shape of its problems (the u, v, sc
 * launch: 16-row tiles with 512 threads, or 32-row tiles with 256 / 512
 * threads; the one-launch p': 16-row tiles with 1024 threads, or 32-row
 * tiles with 512), each candidate `reps` times over all the problems'
 * levels of that shape, and the lane count (1-4, unless the options fix it)
 * by timing the whole solve of the batch with each (the lanes matter in
 * the tail, where only the slowest problems still run); applies the
 * fastest and writes the table as text
 * into table (cap bytes, NUL-terminated; may be NULL), in the format
 * mgf_options.tuning reads:
 *     lanes <n>
 *     level <npx> <npy> <uvs 1|2> <p' 1|2> <threads>
 * Returns the table's length, or a negative status.  The problems are
 * reset to their initial state afterwards.  MGF_ERR_UNSUPPORTED for other
 * solvers (whose launches have no alternative shapes). */
int mgf_autotune(mgf_solver *s, int reps, char *table, size_t cap);

/* Kernel ids for mgf_time_kernel */
#define MGF_KERNEL_SWEEP_PP  0   /* one full smoother sweep of the p' equation (FP64
                                    red-black: timed as relaxations issue them, three
                                    sweeps per launch) */
#define MGF_KERNEL_SWEEP_U   1   /* one full smoother sweep of the u equation */
#define MGF_KERNEL_RELAX_U   2   /* mixed red-black: the first launch of the u
                                    relaxation (assembly + sweeps in one kernel) */
#define MGF_KERNEL_RELAX_PP  3   /* the same for p' */
#define MGF_KERNEL_RELAX_UVS 4   /* mixed red-black: the launch that relaxes u, v and
                                    sc together (each assembled, nswpm sweeps), as
                                    a V-cycle issues it */
/* Time `reps` smoother launches of level ig (for a batch: the launch
 * covering every problem's level at the depth of problem 0's level ig) with
 * HIP events on the solver's stream, on the current coefficients.  A launch
 * is one sweep, except for mixed-precision red-black, where it is the first
 * launch of that field's relaxation: min(sweeps_per_launch, nswpm or nswpp)
 * sweeps.  FP64 mode works in place on a single problem and restores the
 * field; mixed mode sweeps only its correction buffers.  Returns the mean
 * duration of one launch in microseconds and the algorithmic bytes of its
 * sweeps (8 words per unknown and sweep: 4 neighbour coefficients,
 * diagonal, right-hand side, the unknown read and written; 8 bytes each in
 * FP64, 4 in FP32).  MGF_KERNEL_RELAX_*: the launch that assembles the
 * equation and runs the first min(sweeps_per_launch, nswpm or nswpp)
 * sweeps; its bytes are the compulsory ones per unknown (the FP64 state the
 * assembly reads, the flag byte, the FP32 correction or pp written, apu /
 * apv, and the FP32 equation when later launches reuse it).
 * MGF_KERNEL_RELAX_UVS: bytes per interior cell are the union of what its
 * three equations must move: u, v, p, sc, resu, resv, ressc read (FP64),
 * the flag byte; u + du and v + dv (FP64), the sc correction (FP32), apu and
 * apv written (93 bytes). */
int mgf_time_kernel(mgf_solver *s, int kernel, int ig, int reps,
                    double *mean_us, double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
