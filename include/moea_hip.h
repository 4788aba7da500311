/*
 * moea_hip.h -- C ABI of the MI355X-native MOEA engine (libmoea_hip.so).
 *
 * One NSGA-III / RVEA generation on the GPU: constrained non-dominated sorting,
 * ideal point and FP64 hyperplane normalization, perpendicular-distance
 * association, niching, RVEA angle-penalized selection, tournament / random
 * mating, SBX + polynomial mutation, the built-in objectives and a native run
 * loop that chains them without host synchronization.
 *
 * Conventions (those of the reference, python/csrc/native/include/cuda_moea/
 * core/types.cuh:21-51):
 *   - variables  (n, d)  row-major float32
 *   - objectives (m, ld) objective-major float32 (column i = individual i)
 *   - constraints (n,)   float32, feasible means exactly 0.0f
 *   - bounds (2d,)       interleaved [lb0, ub0, lb1, ub1, ...]
 *   - directions (m, k)  objective-major, unit length
 * All pointers are device pointers unless stated; `stream` is a hipStream_t.
 * The caller owns every buffer, including the workspace whose size the
 * *_workspace_bytes functions report; the compute entry points allocate no
 * device memory and make no host synchronization unless stated, so their
 * launches can be captured into a HIP graph (the batched engine does that
 * itself: moea_batch_set_graph).
 * Every function returns MOEA_OK (0) or a negative MOEA_E_* code; the message
 * of the last failure on the calling thread is moea_last_error().
 */
#ifndef MOEA_HIP_H
#define MOEA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOEA_OK 0
#define MOEA_E_INVALID (-1)
#define MOEA_E_HIP (-2)
#define MOEA_E_WORKSPACE (-3)
#define MOEA_E_UNSUPPORTED (-4)
#define MOEA_E_DEVICE (-5)

#define MOEA_MAX_OBJECTIVES 10

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

/* ABI version of this header (bumped on any signature change). */
int moea_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* moea_last_error(void);

/* ------------------------------------------------------------------------ */
/* Draws                                                                     */
/* ------------------------------------------------------------------------ */

/* Injected draws for one generation.  Replaces the cuRAND-backed draw kernels
 * the reference launches per operator (core/cuda/cuda_random.cu:1393-1702),
 * whose stream layout depends on device occupancy and cannot be reproduced.
 * Any NULL array is generated in-kernel from (seed, generation) instead:
 * Philox4x32-10 for mating / SBX / PM, and the reference's own SplitMix64
 * counter hash (cuda_random.cu:1606-1644) for the niching tie-breaks.
 *   mate_a, mate_b : (n) tournament candidates (mate_b unused by random mating)
 *   sbx_u, sbx_p, sbx_copy : (n/2 * d) floats in [0,1), index pair*d + j
 *   sbx_sign       : (n/2 * d) int32 0/1
 *   pm_gate, pm_delta : (n * d) floats in [0,1), index row*d + j
 *   niche_a, niche_b  : (2n) uint32, consumed by sorted position
 */
typedef struct moea_draws {
    const int32_t* mate_a;
    const int32_t* mate_b;
    const float* sbx_u;
    const float* sbx_p;
    const int32_t* sbx_sign;
    const float* sbx_copy;
    const float* pm_gate;
    const float* pm_delta;
    const uint32_t* niche_a;
    const uint32_t* niche_b;
} moea_draws;

/* ------------------------------------------------------------------------ */
/* Non-dominated sorting                                                     */
/* ------------------------------------------------------------------------ */

/* Workspace bytes for moea_nds_rank on n individuals. */
size_t moea_nds_workspace_bytes(int n);

/* Constrained non-dominated sorting of n individuals.  Replaces run_ndsort /
 * run_ndsort_init (selection/nsga3/kernels/non_dominated_sort.cuh, defined at
 * non_dominated_sort.cu:1333-1574).  Fronts are peeled until at least stop_at
 * individuals are assigned; the others get front -1.  Writes fronts (n,) and
 * max_front (1,) = index of the last peeled front, both on the device.
 * max_front = -2 reports a device failure found after the call returned (the
 * peeling launch gave up waiting at its grid barrier; fronts are invalid).
 * f: (m, ldf) objectives, cv: (n,) constraint violations. */
int moea_nds_rank(const float* f, int ldf, const float* cv, int n, int m, int stop_at,
                  int32_t* fronts, int32_t* max_front, void* workspace, size_t workspace_bytes,
                  void* stream);

/* moea_nds_rank without a workspace argument and without any allocation.  It
 * has no counterpart in the reference: it is for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  The workspace is a 16 MiB device array inside the
 * library, one per GPU, held until the process ends; n is limited to
 * moea_nds_static_max_n() (8192) rows and larger n returns MOEA_E_INVALID.
 * Results equal moea_nds_rank's on the same inputs.  Every call on a GPU uses
 * that GPU's one array: calls that overlap on one GPU (different streams or
 * threads without ordering between them) race. */
int moea_nds_rank_static(const float* f, int ldf, const float* cv, int n, int m, int stop_at,
                         int32_t* fronts, int32_t* max_front, void* stream);

/* Largest n moea_nds_rank_static accepts. */
int moea_nds_static_max_n(void);

/* Test hook, no counterpart in the reference: the number of polls after which
 * a grid barrier of the front-peeling launches gives up (0 restores the
 * default, about 2^24).  A small value makes the bounded waits give up, so
 * tests can check that the failure is reported (max_front = -2, the stats
 * device-error word, MOEA_E_DEVICE from moea_batch_check / moea_serve_check)
 * while every launch still ends.  Process-wide; not for production use. */
void moea_debug_set_spin_limit(uint32_t polls);

/* ------------------------------------------------------------------------ */
/* Reference directions (host)                                               */
/* ------------------------------------------------------------------------ */

/* Das-Dennis directions.  kind 0: sized from requested_count
 * (reference/nsga3_reference_generator.cu:62-179); kind 1: from partitions
 * (:182-221); kind 2: RVEA variant, values clamped to >= 1e-6 before
 * normalizing (reference/rvea_reference_generator.cu:185-217).
 * Returns K; when out_host is non-NULL it receives (m, K) unit directions and
 * raw_host (optional) the unnormalized lattice points. */
int moea_das_dennis(int requested_count, int m, int partitions, int kind, float* out_host,
                    float* raw_host);

/* ------------------------------------------------------------------------ */
/* NSGA-III environmental selection                                          */
/* ------------------------------------------------------------------------ */

typedef struct moea_population {
    float* x;        /* (n, d) */
    float* f;        /* (m, n) */
    float* cv;       /* (n,)   */
} moea_population;

/* NSGA-III CV quantization (selection/nsga3/kernels/cv_quantization.cu):
 * before non-dominated sorting, each violation is replaced by its bin, 0 when
 * feasible (cv == 0, or cv <= feasibility_epsilon when that is positive) and
 * otherwise 1 + floor(B x) clamped to [1, B], where x = min(max(cv, 0),
 * clip_upper) / clip_upper, or log1p(log_alpha u) / log1p(log_alpha
 * clip_upper) when log_alpha > 0.  Only the sort sees the bins; survivors
 * keep their violations.  bins = 0 turns it off. */
typedef struct moea_cv_quantization {
    int bins;
    float clip_upper, log_alpha, feasibility_epsilon;
} moea_cv_quantization;

/* The bins of cv (n) into out (n); MOEA_E_INVALID for parameters
 * the reference rejects (validate_cv_quant_params). */
int moea_quantize_cv(const float* cv, float* out, int n, const moea_cv_quantization* q, void* stream);

/* Niching path of an NSGA-III selection.  Every path selects the same
 * survivors; they differ in speed and workspace use.  All zero: automatic.
 * The reference has one path (niching.cu:250-400); this choice has no
 * counterpart there. */
typedef struct moea_niching {
    int mode;      /* 0 automatic, 1 one workgroup over only the last front's
                      members, grouped by direction (a niche of more than 128
                      members falls back to that workgroup's radix sorts),
                      2 three device-wide sorts of all 2n slots, 3 the one
                      workgroup's radix sorts only.  The batched and served
                      engines take one workgroup per instance in mode 1, and
                      in mode 0 while every instance has 2n <= 8192. */
    int lds_cap;   /* one-workgroup paths: last-front members kept in LDS at
                      most (0: as many as the path's LDS holds, about 9500 for
                      the grouping at K = 32640, 5888 for the radix sorts);
                      the others go to the workspace */
} moea_niching;

typedef struct moea_nsga3_select_args {
    int n, m, d, k;
    moea_population parents;    /* input */
    moea_population offspring;  /* input */
    moea_population next;       /* output; must not alias parents/offspring */
    int32_t* next_fronts;       /* (n,) output: fronts of the survivors */
    const float* directions;    /* (m, k) */
    float* ideal;               /* (m,) running ideal point, updated */
    const moea_draws* draws;    /* NULL: all in-kernel */
    uint64_t seed;
    int generation;
    int directions_ready;       /* nonzero: an earlier call on this workspace with these
                                   same directions left their cone tables in it (the
                                   workspace not used for anything else since); 0 builds
                                   them */
    int32_t* stats;             /* optional (8,) device: max_front, n_prior, n_last,
                                   fallback flag, n_rem, -, device error (sticky:
                                   set to 1 when the sort's front peeling gave up
                                   waiting at its grid barrier; that selection
                                   then keeps the parents; only the caller clears
                                   it), - */
    void* workspace;
    size_t workspace_bytes;
    void* stream;
    moea_cv_quantization cv_quant;   /* bins = 0: the sort compares raw violations */
    /* optional normalization outputs, written when the selection normalizes
     * (n_rem > 0; otherwise left unchanged); NULL skips them: */
    int32_t* out_extremes;      /* (m,) merged-row index of each axis's extreme
                                   point, -1 when no row is finite there */
    double* out_intercepts;     /* (m,) hyperplane intercepts a (FP64 solve, or
                                   the row-max fallback when stats[3] = 1)
                                   before their float rounding and sanitizing */
    moea_niching niching;       /* niching path; all zero: automatic */
} moea_nsga3_select_args;

size_t moea_nsga3_workspace_bytes(int n, int m, int k);

/* One NSGA-III selection over the 2n merged individuals.  Replaces
 * IEnvironmentSelector::select for NSGA3EnvironmentSelector
 * (selection/nsga3/nsga3_environment_selector.cu:289-538): constrained NDS,
 * prior/last front split, ideal point, extreme points with the M x M
 * hyperplane solve in FP64 and its row-max fallback, association, niching
 * and the survivor gather in ascending merged-index order. */
int moea_nsga3_select(const moea_nsga3_select_args* args);



/* Perpendicular-distance association of n normalized individuals: for
 * individual i, x = fn[:, i] * a (a: (m,) intercepts) and ref[i] = the
 * direction with the smallest d2 (dot by an fmaf chain, r = fmaf(-dot, w, x),
 * d2 by an fmaf chain; smallest index on ties; non-finite d2 never chosen,
 * ref -1 and dist +inf when none is finite), dist[i] = sqrtf(d2).  Replaces
 * execute_association (selection/nsga3/kernels/association.cu:586-608).
 * method 0: automatic (pruning for M <= 3 and K >= 2048); 1: every
 * direction; 2: cone pruning (directions in blocks of 64 with a bounding cone
 * each; blocks that cannot hold the minimum are skipped), same result as 1 bit
 * for bit.  fn: (m, ld); workspace from
 * moea_associate_workspace_bytes(ld, m, k). */
size_t moea_associate_workspace_bytes(int ld, int m, int k);
int moea_associate(const float* fn, int ld, int n, int m, const float* a, const float* directions, int k,
                   float* dist, int32_t* ref, int method, void* workspace, size_t workspace_bytes,
                   void* stream);

/* Running ideal point over one population (ideal_point.cu:97-133); used by
 * NSGA3EnvironmentSelector::prepare on the initial parents. */
int moea_ideal_point(const float* f, int ldf, int n, int m, float* ideal, void* stream);

/* ------------------------------------------------------------------------ */
/* Reproduction                                                              */
/* ------------------------------------------------------------------------ */

#define MOEA_MATING_TOURNAMENT 0
#define MOEA_MATING_RANDOM 1
/* The caller's mating: parent_indices is an input holding n indices in
 * [0, n), read by SBX as they stand (PythonMating, adapters.cpp:196-227).
 * An index outside [0, n) is clamped into it so that no read leaves the
 * parents; the host adapter rejects such indices before the call. */
#define MOEA_MATING_GIVEN 2
#define MOEA_MUTATION_NONE 0
#define MOEA_MUTATION_POLYNOMIAL 1

typedef struct moea_reproduce_args {
    int n, d;
    int active;                  /* mating draws candidates from [0, active) */
    const int32_t* active_device;/* optional (1,) device count overriding `active` */
    int mating;                  /* MOEA_MATING_* */
    const float* parents_x;      /* (n, d) */
    const int32_t* parents_rank; /* (n,) tournament only */
    const float* parents_cv;     /* (n,) tournament only */
    const float* bounds;         /* (2d,) */
    float sbx_eta, sbx_prob, sbx_copy_prob;   /* annealed eta for this generation */
    int mutation;                /* MOEA_MUTATION_* */
    float pm_eta, pm_prob;
    const moea_draws* draws;     /* NULL: in-kernel */
    uint64_t seed;
    int generation;
    float* offspring_x;          /* (n, d) output */
    int32_t* parent_indices;     /* (n,) output: mating winners */
    void* stream;
} moea_reproduce_args;

/* Mating -> SBX -> PM.  Replaces ReproductionPipeline::generate
 * (reproduction/reproduction_pipeline.cu:37-56) with TournamentMating /
 * RandomMating (mating_selection.cu:11-108), SimulatedBinaryCrossover
 * (sbx_crossover.cu:20-190) and PolynomialMutation (polynomial_mutation.cu:20-124). */
int moea_reproduce(const moea_reproduce_args* args);

/* Linear eta annealing (sbx_crossover.cu:12-18), host helper. */
float moea_annealed(float begin, float end, int generation, int total);

/* Uniform initial population lb + u * (ub - lb) from the in-kernel generator
 * (Population::initialize_random, core/population.cu:97-115). */
int moea_init_population(float* x, int n, int d, const float* bounds, uint64_t seed, void* stream);

/* ------------------------------------------------------------------------ */
/* Built-in objectives (IProblemEvaluator::evaluate,                          */
/* problem/problem_evaluator.cuh:10-31)                                       */
/* ------------------------------------------------------------------------ */

/* f_c(x) = sum_j (x_j - centers[c, j])^2, c < m; cv = 0.  f is (m, ldf). */
int moea_eval_quadratic(const float* x, int n, int d, const float* centers, int m, float* f,
                        int ldf, float* cv, void* stream);

/* DTLZ1-DTLZ7 (problem 1-7) or ConvexDTLZ2 (problem 8), as the reference's
 * built-in kernels (problem/dtlz/dtlz_kernels.cu:28-717): DTLZ1 clamps x to
 * [0, 1], the spherical problems use max(0, cos), DTLZ4 maps x to x^100 and
 * DTLZ5 clamps its theta to [0, 1]; cv = 0.  Problems 9-14 are the
 * constrained variants C1-DTLZ1, C1-DTLZ3, C2-DTLZ2, C2-ConvexDTLZ2, C3-DTLZ1
 * and C3-DTLZ4 (dtlz_kernels.cu:824-1165): their base problem's objectives
 * and the violation in cv, constraints always active.  Problem 15 is CSDP,
 * the car side-impact design problem (d = 7, m = 3; dtlz_kernels.cu:719-738,
 * 1046-1085).  C1-DTLZ3 and C2-DTLZ2
 * take m in {2, 3, 5, 8, 10}, C2-ConvexDTLZ2 m in {3, 5, 8, 10}.  Needs
 * 2 <= m <= d. */
int moea_eval_dtlz(int problem, const float* x, int n, int d, int m, float* f, int ldf, float* cv,
                   void* stream);

/* Seeded MLP objective: h = tanh(W1 x) (hidden units), y = W2 h (m outputs),
 * f_c = y_c^2 + 0.01 * ||x - 0.5||^2; cv = 0.  w1: (hidden, d), w2: (m, hidden).
 * hwork: (n, hidden) scratch. */
int moea_eval_mlp(const float* x, int n, int d, const float* w1, int hidden, const float* w2,
                  int m, float* f, int ldf, float* cv, float* hwork, void* stream);

/* ------------------------------------------------------------------------ */
/* RVEA                                                                      */
/* ------------------------------------------------------------------------ */

typedef struct moea_rvea_select_args {
    int n, m, d, k;
    int active;                 /* live parent slots [0, active) */
    const int32_t* active_device;/* optional (1,) device count overriding `active` */
    moea_population parents, offspring, next;
    const float* directions;    /* (m, k) current (adapted) directions */
    const float* gamma;         /* (k,) */
    float penalty;              /* m * (t / G)^alpha, host-computed */
    int32_t* out_active;        /* (1,) device: new active count */
    void* workspace;
    size_t workspace_bytes;
    void* stream;
} moea_rvea_select_args;

size_t moea_rvea_workspace_bytes(int n, int m, int k);

/* RVEA environmental selection.  Replaces RVEAEnvironmentSelector::select
 * (selection/rvea/rvea_environment_selector.cu:163-287): z-min translation,
 * cosine partition, APD, one elitist winner per direction, truncation, and
 * the gather with NaN / sentinel dead rows. */
int moea_rvea_select(const moea_rvea_select_args* args);

/* APD penalty scale m * (t / G)^alpha (angle_penalized_distance.cu:159-184). */
float moea_rvea_penalty(int m, int t, int max_generations, float alpha);

/* gamma_j = acos(max_{k != j} v_j . v_k) (angle_penalized_distance.cu:387-417). */
int moea_rvea_gamma(const float* directions, int m, int k, float* gamma, void* stream);

/* V = normalize(V0 * (zmax - zmin)) over the first n_active columns of f
 * (rvea_reference_adaptation.cu:279-317).  n_active is read on the device;
 * scratch holds at least m floats. */
int moea_rvea_adapt(const float* v0, const float* f, int ldf, const int32_t* n_active, int m,
                    int k, float* v_out, float* scratch, void* stream);

/* ------------------------------------------------------------------------ */
/* Native run loop                                                           */
/* ------------------------------------------------------------------------ */

#define MOEA_ALGORITHM_NSGA3 0
#define MOEA_ALGORITHM_RVEA 1

#define MOEA_PROBLEM_QUADRATIC 0
#define MOEA_PROBLEM_MLP 1
#define MOEA_PROBLEM_CALLBACK 2
#define MOEA_PROBLEM_DTLZ1 3
#define MOEA_PROBLEM_DTLZ2 4
#define MOEA_PROBLEM_DTLZ3 5
#define MOEA_PROBLEM_DTLZ4 6
#define MOEA_PROBLEM_DTLZ5 7
#define MOEA_PROBLEM_DTLZ6 8
#define MOEA_PROBLEM_DTLZ7 9
#define MOEA_PROBLEM_CONVEX_DTLZ2 10
#define MOEA_PROBLEM_C1_DTLZ1 11
#define MOEA_PROBLEM_C1_DTLZ3 12
#define MOEA_PROBLEM_C2_DTLZ2 13
#define MOEA_PROBLEM_C2_CONVEX_DTLZ2 14
#define MOEA_PROBLEM_C3_DTLZ1 15
#define MOEA_PROBLEM_C3_DTLZ4 16
#define MOEA_PROBLEM_CSDP 17

/* Problem evaluator callback: the C form of IProblemEvaluator::evaluate
 * (problem/problem_evaluator.cuh:10-31).  Must write f (m, n) and cv (n) for
 * the n rows of x (n, d), enqueued on `stream`; generation is -1 for the
 * initial population (algorithm.cu:180-184).  Returns MOEA_OK or an error. */
typedef int (*moea_evaluate_fn)(void* user, const float* x, int n, int d, int m, float* f,
                                float* cv, int generation, int max_generations, void* stream);

/* Optional companion of a callback problem: IProblemEvaluator::prepare_parent,
 * which EvolutionaryAlgorithm::step calls before reproduction
 * (algorithm.cu:230-241; the Python form in adapters.cpp:175-186).  It may
 * rewrite the parents' f and cv, enqueued on `stream`.  Returns 1 when the
 * parents changed (the selector then prepares again: NSGA-III re-sorts the
 * parent fronts), 0 when not, or a negative MOEA_E_* code. */
typedef int (*moea_prepare_parent_fn)(void* user, moea_population* parents, int n, int d, int m,
                                      int generation, int max_generations, void* stream);

typedef struct moea_problem_desc {
    int kind;                 /* MOEA_PROBLEM_* */
    const float* centers;     /* quadratic: (m, d) */
    const float* w1;          /* mlp: (hidden, d) */
    const float* w2;          /* mlp: (m, hidden) */
    int hidden;
    float* hwork;             /* mlp: (n, hidden) scratch */
    moea_evaluate_fn callback;
    void* user;
    /* Constrained DTLZ variants (MOEA_PROBLEM_C*): constraints count from the
     * generation g whose (g + 1) / max_generations reaches this ratio; before
     * it cv = 0, and at that generation the parents' cv is recomputed and the
     * NSGA-III parent fronts re-sorted (dtlz_kernels.cu:1295-1394).  <= 0:
     * always on. */
    float cv_activation_ratio;
    moea_prepare_parent_fn prepare_parent;   /* callback problems: optional (NULL: none) */
} moea_problem_desc;

/* Strategy hooks: the C form of the reference's strategy plugins
 * (IMatingSelector, ICrossoverOperator, IMutationOperator,
 * IReferenceDirectionProvider; Python adapters in adapters.cpp:196-338).
 * A NULL hook keeps the built-in strategy.  Every hook enqueues its work on
 * `stream` and returns MOEA_OK or an error, which moea_run_step returns. */

/* Mating: write n parent indices (pairs 2i, 2i+1 are crossed).  rank is the
 * parents' NSGA-III front ranks (NULL under RVEA); active_device the RVEA
 * active count (NULL under NSGA-III: all n). */
typedef int (*moea_mate_fn)(void* user, const moea_population* parents, int n, int d, int m,
                            const int32_t* rank, const int32_t* active_device, int32_t* parent_indices,
                            int generation, int max_generations, void* stream);
/* Crossover: write the (n, d) offspring variables from the parents and the
 * mating's parent_indices. */
typedef int (*moea_crossover_fn)(void* user, const moea_population* parents, int n, int d, int m,
                                 const int32_t* parent_indices, float* offspring_x, int generation,
                                 int max_generations, void* stream);
/* Mutation: update offspring->x (n, d) in place, before evaluation. */
typedef int (*moea_mutate_fn)(void* user, const moea_population* offspring, int n, int d, int m,
                              int generation, int max_generations, void* stream);
/* Reference-direction update after each selection (the survivors `next`):
 * may rewrite directions (m, k) in place.  Returns 1 if it did, 0 if not, a
 * negative error otherwise.  Replaces RVEA's built-in adaptation. */
typedef int (*moea_directions_fn)(void* user, const moea_population* next, int n, int d, int m,
                                  const int32_t* active_device, float* directions, int k, int generation,
                                  int max_generations, void* stream);

/* Environmental selection (IEnvironmentSelector::select): write the n
 * survivors into next (x, f, cv) from parents and offspring.  The hook keeps
 * the selector's state: before the next generation it writes the survivors'
 * ranks into args.fronts[args.fronts_current] (read by TournamentMating) and
 * the active count into args.active_device (read by the mating draws). */
typedef int (*moea_select_fn)(void* user, const moea_population* parents, const moea_population* offspring,
                              const moea_population* next, int n, int d, int m, int generation,
                              int max_generations, void* stream);

/* Re-preparation of a caller's selector (IEnvironmentSelector::prepare) on
 * the parents, when a problem changes their evaluation during the run: the
 * constrained DTLZ variants at the generation their constraints switch on. */
typedef int (*moea_prepare_fn)(void* user, const moea_population* parents, int n, int d, int m, int generation,
                               int max_generations, void* stream);

typedef struct moea_strategy_hooks {
    void* user;
    moea_mate_fn mate;                     /* NULL: args.mating */
    moea_crossover_fn crossover;           /* NULL: SBX */
    moea_mutate_fn mutate;                 /* NULL: args.mutation */
    moea_directions_fn update_directions;  /* NULL: RVEA adaptation / fixed directions */
    moea_select_fn select;                 /* NULL: args.algorithm's selection */
    moea_prepare_fn prepare;               /* with select: re-prepare the caller's selector */
} moea_strategy_hooks;

/* State of one run.  The caller owns every buffer; moea_run_* update the
 * integer fields (current, fronts_current, generation). */
typedef struct moea_run_args {
    int algorithm;             /* MOEA_ALGORITHM_* */
    int n, m, d, k, max_generations;
    moea_population pops[3];   /* rotating parents / offspring / next */
    int current;               /* index of the parents in pops */
    int32_t* fronts[2];        /* NSGA-III parent fronts, ping-pong */
    int fronts_current;
    int32_t* parent_indices;   /* (n) mating winners */
    const float* bounds;       /* (2d) */
    const float* directions;   /* (m, k); RVEA adapts it in place */
    const float* directions0;  /* RVEA initial directions (m, k) */
    float* gamma;              /* RVEA (k) */
    float* ideal;              /* NSGA-III (m) */
    int32_t* active_device;    /* RVEA (1) */
    float* adapt_scratch;      /* RVEA (m) */
    int32_t* stats;            /* NSGA-III optional (8) per-generation selection stats */
    int32_t* stats_init;       /* NSGA-III (8) */
    int mating, mutation;
    float sbx_eta[2], sbx_prob, sbx_copy_prob;
    float pm_eta[2], pm_prob;
    float rvea_alpha, rvea_frequency;
    uint64_t seed;
    int initial_x_given;       /* 1: pops[current].x already holds the initial population */
    moea_problem_desc problem;
    int generation;
    int directions_ready;      /* NSGA-III: cone tables of `directions` in the workspace;
                                  cleared by moea_run_init, set after the first selection */
    void* workspace;
    size_t workspace_bytes;
    void* stream;
    const moea_strategy_hooks* hooks;  /* NULL: every strategy built in */
    moea_cv_quantization cv_quant;     /* NSGA-III; bins = 0: off */
    moea_niching niching;              /* NSGA-III niching path; all zero: automatic */
} moea_run_args;

size_t moea_run_workspace_bytes(int algorithm, int n, int m, int k);

/* EvolutionaryAlgorithm::initialize (algorithm.cu:104-217): initial population
 * (unless given), evaluation at generation -1, NSGA-III parent fronts and ideal
 * point or RVEA gamma and active count. */
int moea_run_init(moea_run_args* args);

/* Up to `generations` calls of EvolutionaryAlgorithm::step (algorithm.cu:219-286),
 * stopping at max_generations; no host synchronization. */
int moea_run_step(moea_run_args* args, int generations);

/* ------------------------------------------------------------------------ */
/* Batched NSGA-III                                                          */
/* ------------------------------------------------------------------------ */

/* One instance of a batch: an NSGA-III run with tournament mating, SBX and
 * polynomial mutation (the moea_run_* operators) on a built-in objective:
 * quadratic-center (problem MOEA_PROBLEM_QUADRATIC, centers) or the seeded MLP
 * (MOEA_PROBLEM_MLP, w1 / w2 / hidden).  All instances of a moea_batch run the
 * same number of generations; a moea_serve instance runs its own. */
typedef struct moea_batch_problem {
    int n, m, d, k;            /* population (even), objectives (<= 10), variables, directions */
    int max_generations;
    uint64_t seed;
    int problem;               /* MOEA_PROBLEM_QUADRATIC or MOEA_PROBLEM_MLP */
    const float* centers;      /* quadratic: (m, d) device */
    const float* w1;           /* mlp: (hidden, d) device */
    const float* w2;           /* mlp: (m, hidden) device */
    int hidden;
    const float* bounds;       /* (2d) device */
    const float* directions;   /* (m, k) device, unit length */
    float sbx_eta[2], sbx_prob, sbx_copy_prob;
    float pm_eta[2], pm_prob;
    moea_niching niching;      /* batch: the same in every problem of the batch;
                                  serve: ignored (moea_serve_config's applies) */
} moea_batch_problem;

typedef struct moea_batch moea_batch;   /* host handle */

/* The reference has no batched engine: each EvolutionaryAlgorithm
 * (algorithm/algorithm.cu:104-329) runs one population.  This advances many
 * independent instances of uneven n / m / d with one launch sequence per
 * generation; instance i computes exactly what moea_run_init / moea_run_step
 * compute for it alone.  Every device buffer lives in the caller's workspace
 * (size from moea_batch_workspace_bytes); the handle is host memory only. */
size_t moea_batch_workspace_bytes(const moea_batch_problem* problems, int count);
int moea_batch_create(const moea_batch_problem* problems, int count, void* workspace,
                      size_t workspace_bytes, void* stream, moea_batch** out);
/* Initial populations, evaluation, parent fronts and ideal points. */
int moea_batch_init(moea_batch* batch);
/* Up to `generations` generations of every instance; no host synchronization. */
int moea_batch_step(moea_batch* batch, int generations);
int moea_batch_generation(const moea_batch* batch);
/* Device views of instance `index`'s current population (x (n, d), f (m, n),
 * cv (n)) and its fronts (n); valid until the next moea_batch_step. */
int moea_batch_view(const moea_batch* batch, int index, moea_population* out, int32_t** fronts);
/* Waits for the batch's queued work and returns MOEA_E_DEVICE if a
 * front-peeling launch of any earlier generation gave up waiting at its grid
 * barrier (the fronts of that generation were invalid; the affected
 * instances kept their parents), else MOEA_OK. */
int moea_batch_check(const moea_batch* batch);
/* Generation graphs (no reference counterpart; the batched engine is this
 * library's own).  With enable != 0, moea_batch_step replays one HIP graph
 * per generation, captured from the batch's own launch sequence the first
 * time it is needed and again whenever the set of running instances changed.
 * Every per-generation quantity lives in the device instance table, so the
 * replays are the same launches as the direct path and give the same bits.
 * The captured graph holds kernel nodes only (a capture with any other node
 * fails with MOEA_E_UNSUPPORTED).  0 switches back to direct launches. */
int moea_batch_set_graph(moea_batch* batch, int enable);
/* Nodes of the current generation graph (0 before the first capture); the
 * kernel and memset node counts and the number of captures so far. */
int moea_batch_graph_info(const moea_batch* batch, int* kernel_nodes, int* memset_nodes, int* captures);
/* The batched front-peeling plan: workgroups used = the peel kernel's
 * co-resident capacity / capacity_divisor, one workgroup per
 * rows_per_workgroup 64-row blocks of an instance.  The plan decides only
 * which workgroup peels which rows; the fronts do not depend on it. */
int moea_batch_peel_plan(const moea_batch* batch, int* capacity_divisor, int* rows_per_workgroup);
int moea_batch_set_peel_plan(moea_batch* batch, int capacity_divisor, int rows_per_workgroup);
/* Autotuner: runs the batch's next generations under each candidate peel
 * plan ((1, 16), (1, 32), (2, 16), (2, 32)), twice each (forward, then
 * backward order), `trials` >= 2 generations per visit with the first of each
 * visit untimed, and keeps the fastest.  The generations are the run's own:
 * the results equal an untuned run.  Writes the chosen candidate index and
 * up to max_candidates mean ms per generation (-1 for a candidate that could
 * not hold the batch or did not run); returns the generations it ran. */
int moea_batch_autotune(moea_batch* batch, int trials, int* chosen, double* ms_per_generation, int max_candidates);
void moea_batch_destroy(moea_batch* batch);

/* ------------------------------------------------------------------------ */
/* Served NSGA-III (continuous batching)                                     */
/* ------------------------------------------------------------------------ */

/* No counterpart in the reference, which runs one population per
 * EvolutionaryAlgorithm.  A server keeps up to max_instances runs in slots and
 * advances every running one by a generation per moea_serve_step, with the
 * launch sequence of moea_batch_step; runs are admitted and released between
 * steps.  Each run's buffers (moea_serve_instance_bytes) come from an arena:
 * the part of the caller's workspace beyond moea_serve_reserved_bytes, so the
 * workspace size is the HBM budget.  A run computes exactly what moea_run_init
 * and moea_run_step compute for it alone, whatever else shares the steps. */
typedef struct moea_serve_config {
    int max_instances;         /* slots */
    long long max_rows;        /* bound on the sum of 2n over live runs (sort scratch) */
    moea_niching niching;      /* niching path of every served run; all zero: automatic */
} moea_serve_config;

typedef struct moea_serve moea_serve;   /* host handle */

size_t moea_serve_reserved_bytes(const moea_serve_config* config);
/* Arena bytes one run takes (0 for an invalid problem). */
size_t moea_serve_instance_bytes(const moea_batch_problem* problem);
int moea_serve_create(const moea_serve_config* config, void* workspace, size_t workspace_bytes,
                      void* stream, moea_serve** out);
/* Places a run in a free slot (MOEA_E_WORKSPACE when no slot, rows or arena
 * space is free); its initialization runs at the start of the next step.  The
 * problem's device arrays must outlive the run. */
int moea_serve_admit(moea_serve* server, const moea_batch_problem* problem, int* slot);
/* Initializes newly admitted runs, then advances every running run by one
 * generation.  Slots of runs that reached max_generations in this step are
 * written to `finished` (up to max_finished; n_finished gets the full count).
 * Synchronizes the stream only when the set of runs changed. */
int moea_serve_step(moea_serve* server, int32_t* finished, int max_finished, int* n_finished);
/* Generations completed by the run in `slot`. */
int moea_serve_generation(const moea_serve* server, int slot);
/* Device views of the run in `slot` (as moea_batch_view); valid until the next
 * step or the run's release. */
int moea_serve_view(const moea_serve* server, int slot, moea_population* out, int32_t** fronts);
/* Frees the slot and its arena space (synchronizes the stream first). */
int moea_serve_release(moea_serve* server, int slot);
/* As moea_batch_check, for a server.  moea_serve_step and
 * moea_serve_release also return MOEA_E_DEVICE at the synchronizations they
 * already make. */
int moea_serve_check(const moea_serve* server);
/* As moea_batch_set_graph, for a server: steps between changes of the set
 * of runs replay one captured generation. */
int moea_serve_set_graph(moea_serve* server, int enable);
void moea_serve_destroy(moea_serve* server);

/* ------------------------------------------------------------------------ */
/* Native problem plugins                                                    */
/* ------------------------------------------------------------------------ */

/* A plugin is a shared library built with hipcc against
 * include/cuda_moea/native/plugin.hpp: it exports
 * `const cuda_moea::native::Plugin* cuda_moea_problem_v1()` returning
 * {abi = 1, build_id, name, parameter schema, create, destroy}
 * (python/csrc/native/include/cuda_moea/native/plugin.cuh:35-58).  These
 * functions are the reference's loader (python/csrc/native_problem.cpp:13-170)
 * behind the C ABI.  A created problem plugs into a run as a callback
 * problem: moea_problem_desc{kind = MOEA_PROBLEM_CALLBACK,
 * callback = moea_plugin_evaluate, prepare_parent = moea_plugin_prepare_parent,
 * user = the moea_plugin_problem*}. */
typedef struct moea_plugin moea_plugin;
typedef struct moea_plugin_problem moea_plugin_problem;

/* Parameter types, numbered as the schema's variant alternatives. */
#define MOEA_PARAM_INT 0
#define MOEA_PARAM_FLOAT 1
#define MOEA_PARAM_BOOL 2
#define MOEA_PARAM_STR 3
#define MOEA_PARAM_FLOAT_LIST 4

typedef struct moea_param_value {
    const char* name;
    int type;                 /* MOEA_PARAM_* */
    int list_length;          /* float_list */
    long long int_value;      /* int; bool as 0 / 1 */
    double float_value;       /* float */
    const char* str_value;    /* str */
    const double* list;       /* float_list */
} moea_param_value;

/* The SDK build id this library accepts (CUDA_MOEA_NATIVE_BUILD_ID). */
const char* moea_native_build_id(void);

/* dlopens `path` and checks its registration, in the reference's order
 * (native_problem.cpp:16-34): the entry symbol ("Missing
 * cuda_moea_problem_v1 registration"), abi == 1 and an exact build_id ("SDK
 * mismatch"; nothing past those two fields is read before), a complete record,
 * unique non-null parameter names.  MOEA_E_INVALID with the message on
 * failure (the library is closed again). */
int moea_plugin_open(const char* path, moea_plugin** out);
/* Releases the caller's handle; the library stays loaded while problems
 * created from it live. */
void moea_plugin_close(moea_plugin* plugin);
const char* moea_plugin_name(const moea_plugin* plugin);
const char* moea_plugin_build_id(const moea_plugin* plugin);
/* The parameter schema (native_problem.cpp:155-170): count, then per index
 * the name, type and default (pointers valid while the plugin is open) and
 * the description. */
int moea_plugin_param_count(const moea_plugin* plugin);
int moea_plugin_param(const moea_plugin* plugin, int index, moea_param_value* out, const char** description);

/* Builds a problem (native_problem.cpp:75-109): every schema parameter takes
 * the given value or its default; unknown names are rejected; values are
 * checked against the default's type (an int is accepted for a float; floats
 * finite; bool and int strict); the factory's exceptions become
 * MOEA_E_INVALID "Native problem construction failed: ..."; the instance's
 * ProblemInfo must equal the requested dimension, objectives and bounds. */
int moea_plugin_create(moea_plugin* plugin, int dimension, int objectives, const float* lower_bounds,
                       const float* upper_bounds, const moea_param_value* params, int param_count,
                       moea_plugin_problem** out);
int moea_plugin_problem_constraints(const moea_plugin_problem* problem);
/* IProblemEvaluator::initialize with a CudaContext on `stream` (the run's
 * stream) with `seed`; `bounds` is the run's device bounds array
 * (interleaved lb, ub), handed to evaluate in PopulationView::bounds. */
int moea_plugin_problem_initialize(moea_plugin_problem* problem, void* stream, uint64_t seed,
                                   const float* bounds);
/* Waits for the problem's stream, destroys the instance, and unloads the
 * library when no handle or problem still uses it (native_problem.cpp:110-126). */
void moea_plugin_problem_destroy(moea_plugin_problem* problem);
/* moea_evaluate_fn with user = the moea_plugin_problem*.  Violations of a
 * problem with no constraints are written as zero before evaluate runs. */
int moea_plugin_evaluate(void* user, const float* x, int n, int d, int m, float* f, float* cv, int generation,
                         int max_generations, void* stream);
/* moea_prepare_parent_fn with user = the moea_plugin_problem*. */
int moea_plugin_prepare_parent(void* user, moea_population* parents, int n, int d, int m, int generation,
                               int max_generations, void* stream);
int moea_plugin_reset(moea_plugin_problem* problem);

/* ------------------------------------------------------------------------ */
/* Profiling                                                                 */
/* ------------------------------------------------------------------------ */

/* Opt-in per-stage timing with HIP events on the launch stream, the role of
 * the reference's ASAP_NDSORT_FRONT_PROP_TIMING breakdown
 * (non_dominated_sort.cu:101-107).  Off by default. */
int moea_profile_enable(int enable);

/* Per-stage total milliseconds and launch counts since the last read; returns
 * the number of stages.  Synchronizes on the recorded events. */
int moea_profile_read(double* ms_out, long long* count_out, int max_stages);

const char* moea_profile_stage_name(int stage);

#ifdef __cplusplus
}
#endif

#endif /* MOEA_HIP_H */
