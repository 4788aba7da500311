// Native generation loop: the role of EvolutionaryAlgorithm::initialize / step /
// run_device (NS/algorithm/algorithm.cu:104-329) for the built-in operators.
//
// One generation is reproduce -> evaluate -> select (-> RVEA direction
// adaptation), a fixed sequence of stream-ordered launches with no host
// synchronization and no allocation, so it can also be captured into a graph.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

__global__ void set_i32(int32_t* p, int v) { p[0] = v; }

__global__ void zero_f32(float* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

bool constrained_dtlz(int kind) { return kind >= MOEA_PROBLEM_C1_DTLZ1 && kind <= MOEA_PROBLEM_CSDP; }

// MOEAStdTestEvaluator::is_cv_active (dtlz_kernels.cu:1300-1309): constraints of
// an evaluation at generation g count once (g + 1) / max_generations reaches
// the ratio; the initial population (g = -1) is evaluated without them.
bool cv_active(const moea_run_args& a, int g)
{
    const float ratio = a.problem.cv_activation_ratio;
    if (!constrained_dtlz(a.problem.kind) || !(ratio > 0.0f)) return true;
    if (g < 0) return false;
    const int total = a.max_generations > 0 ? a.max_generations : 1;
    return static_cast<float>(g + 1) / static_cast<float>(total) >= ratio;
}

// is_cv_activation_switch_node (dtlz_kernels.cu:1311-1322): generation 0 when
// the constraints are on from the start (ratio <= 0 included), else the
// generation at which they switch on
bool cv_switch(const moea_run_args& a, int g)
{
    if (!constrained_dtlz(a.problem.kind)) return false;
    return g == 0 ? cv_active(a, 0) : (!cv_active(a, g - 1) && cv_active(a, g));
}

// RVEA rows past the live count keep their 1e15 objectives and violations
// (and NaN variables) when a switch re-evaluates the parents.
__global__ void rvea_dead_sentinels(float* __restrict__ f, float* __restrict__ cv, int n, int m,
                                    const int32_t* __restrict__ active)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || i < active[0]) return;
    for (int k = 0; k < m; ++k) f[static_cast<size_t>(k) * n + i] = 1e15f;
    cv[i] = 1e15f;
}

int evaluate_problem(const moea_run_args& a, const moea_population& p, hipStream_t st, int generation);

int evaluate(const moea_run_args& a, const moea_population& p, hipStream_t st, int generation)
{
    MOEA_CHECK(evaluate_problem(a, p, st, generation));
    if (!cv_active(a, generation)) {
        // MOEAStdTestEvaluator::evaluate: violations count only once active
        hipLaunchKernelGGL(zero_f32, dim3(ceil_div(a.n, 256)), dim3(256), 0, st, p.cv, a.n);
        MOEA_HIP(hipGetLastError());
    }
    return kOk;
}

int evaluate_problem(const moea_run_args& a, const moea_population& p, hipStream_t st, int generation)
{
    const moea_problem_desc& pr = a.problem;
    switch (pr.kind) {
        case MOEA_PROBLEM_QUADRATIC:
            return eval_quadratic(p.x, a.n, a.d, pr.centers, a.m, p.f, a.n, p.cv, st);
        case MOEA_PROBLEM_MLP:
            return eval_mlp(p.x, a.n, a.d, pr.w1, pr.hidden, pr.w2, a.m, p.f, a.n, p.cv, pr.hwork, st);
        case MOEA_PROBLEM_DTLZ1:
            return eval_dtlz(1, p.x, a.n, a.d, a.m, p.f, a.n, p.cv, st);
        case MOEA_PROBLEM_DTLZ2: case MOEA_PROBLEM_DTLZ3: case MOEA_PROBLEM_DTLZ4: case MOEA_PROBLEM_DTLZ5:
        case MOEA_PROBLEM_DTLZ6: case MOEA_PROBLEM_DTLZ7: case MOEA_PROBLEM_CONVEX_DTLZ2:
            // problem kinds 4..10 are DTLZ2..DTLZ7 and ConvexDTLZ2 (kernel kinds 2..8)
            return eval_dtlz(pr.kind - MOEA_PROBLEM_DTLZ2 + 2, p.x, a.n, a.d, a.m, p.f, a.n, p.cv, st);
        case MOEA_PROBLEM_C1_DTLZ1: case MOEA_PROBLEM_C1_DTLZ3: case MOEA_PROBLEM_C2_DTLZ2:
        case MOEA_PROBLEM_C2_CONVEX_DTLZ2: case MOEA_PROBLEM_C3_DTLZ1: case MOEA_PROBLEM_C3_DTLZ4:
            // problem kinds 11..16 are the constrained variants (kernel kinds 9..14)
            return eval_dtlz(pr.kind - MOEA_PROBLEM_C1_DTLZ1 + 9, p.x, a.n, a.d, a.m, p.f, a.n, p.cv, st);
        case MOEA_PROBLEM_CSDP:
            return eval_dtlz(15, p.x, a.n, a.d, a.m, p.f, a.n, p.cv, st);
        case MOEA_PROBLEM_CALLBACK:
            if (!pr.callback) return fail(kInvalidArgument, "moea_run: callback problem without callback");
            return pr.callback(pr.user, p.x, a.n, a.d, a.m, p.f, p.cv, generation, a.max_generations, st);
        default:
            return fail(kInvalidArgument, "moea_run: unknown problem kind");
    }
}

int check(const moea_run_args& a)
{
    if (a.n <= 0 || a.m <= 0 || a.m > kMaxM || a.d <= 0 || a.k <= 0 || a.max_generations <= 0)
        return fail(kInvalidArgument, "Algorithm requires positive population_size and max_generations");
    if (a.algorithm != MOEA_ALGORITHM_NSGA3 && a.algorithm != MOEA_ALGORITHM_RVEA)
        return fail(kInvalidArgument, "algorithm must be NSGA3 or RVEA");
    if (a.algorithm == MOEA_ALGORITHM_NSGA3 && (a.n % 2) != 0)
        return fail(kInvalidArgument, "The current NSGA-III backend requires an even population size");
    for (int i = 0; i < 3; ++i)
        if (!a.pops[i].x || !a.pops[i].f || !a.pops[i].cv)
            return fail(kInvalidArgument, "moea_run: population buffers missing");
    if (!a.bounds || !a.directions || !a.parent_indices || !a.workspace || a.current < 0 || a.current > 2)
        return fail(kInvalidArgument, "moea_run: missing buffers");
    if (a.algorithm == MOEA_ALGORITHM_NSGA3 && (!a.fronts[0] || !a.fronts[1] || !a.ideal))
        return fail(kInvalidArgument, "moea_run: NSGA-III needs fronts and ideal buffers");
    if (a.algorithm == MOEA_ALGORITHM_RVEA &&
        (!a.gamma || !a.directions0 || !a.active_device || !a.adapt_scratch))
        return fail(kInvalidArgument, "moea_run: RVEA needs gamma, initial directions and active buffers");
    if (a.hooks && a.hooks->select && (!a.active_device || !a.fronts[0] || !a.fronts[1]))
        return fail(kInvalidArgument, "moea_run: a selection hook needs fronts and active buffers");
    return kOk;
}

// NSGA3EnvironmentSelector::prepare's sort: the parents' fronts over n with
// stop at n (nsga3_environment_selector.cu:236-275), on quantized violations
// when CV quantization is on (the bins sit in the workspace past the sort's
// own part).
int parent_fronts(moea_run_args& a, const moea_population& par, hipStream_t st)
{
    const float* cv_sort = par.cv;
    if (a.cv_quant.bins > 0) {
        const size_t nds_bytes = (nds_workspace_bytes(a.n) + 255) / 256 * 256;
        if (nds_bytes + sizeof(float) * static_cast<size_t>(a.n) > a.workspace_bytes)
            return fail(kWorkspaceTooSmall, "moea_run: workspace too small");
        float* q = reinterpret_cast<float*>(static_cast<char*>(a.workspace) + nds_bytes);
        MOEA_CHECK(quantize_cv(par.cv, q, a.n, a.cv_quant, st));
        cv_sort = q;
    }
    // a failed sort is also flagged in the sticky device-error word of stats
    return nds_run(par.f, a.n, cv_sort, a.n, a.m, a.n, a.fronts[a.fronts_current], a.stats_init, a.workspace,
                   a.workspace_bytes, st, a.stats ? reinterpret_cast<uint32_t*>(a.stats + 6) : nullptr);
}

}  // namespace

size_t run_workspace_bytes(int algorithm, int n, int m, int k)
{
    if (algorithm == MOEA_ALGORITHM_NSGA3) {
        const size_t a = nsga3_workspace_bytes(n, m, k);
        // the parents' sort, then room for their quantized violations
        const size_t b = (nds_workspace_bytes(n) + 255) / 256 * 256 + sizeof(float) * static_cast<size_t>(n);
        return a > b ? a : b;
    }
    return rvea_workspace_bytes(n, m, k);
}

int run_init(moea_run_args& a)
{
    MOEA_CHECK(check(a));
    hipStream_t st = static_cast<hipStream_t>(a.stream);
    moea_population& par = a.pops[a.current];
    if (!a.initial_x_given) MOEA_CHECK(init_population(par.x, a.n, a.d, a.bounds, a.seed, st));
    MOEA_CHECK(evaluate(a, par, st, -1));
    if (a.algorithm == MOEA_ALGORITHM_NSGA3) {
        // NSGA3EnvironmentSelector::prepare: parent fronts over n with stop at n,
        // and the initial ideal point (nsga3_environment_selector.cu:236-287)
        MOEA_CHECK(parent_fronts(a, par, st));
        MOEA_CHECK(ideal_point(par.f, a.n, a.n, a.m, a.ideal, st));
    } else {
        MOEA_CHECK(rvea_gamma(a.directions, a.m, a.k, a.gamma, st));
        // active = n (RVEAEnvironmentSelector::reset / allocate)
        hipLaunchKernelGGL(set_i32, dim3(1), dim3(1), 0, st, a.active_device, a.n);
        MOEA_HIP(hipGetLastError());
    }
    a.generation = 0;
    a.directions_ready = 0;   // run_init's sort used the workspace
    return kOk;
}

namespace {

// AdaptiveRVEADirections adapts after generation g when g + 1 is a multiple of
// round(frequency * max_generations), at least 1 (reference_directions.cu:132-140).
bool adapt_now(const moea_run_args& a, int g)
{
    int interval = static_cast<int>(std::lround(static_cast<double>(a.rvea_frequency) * a.max_generations));
    if (interval < 1) interval = 1;
    return (g + 1) % interval == 0;
}

}  // namespace

int run_step(moea_run_args& a, int generations)
{
    MOEA_CHECK(check(a));
    hipStream_t st = static_cast<hipStream_t>(a.stream);
    for (int it = 0; it < generations && a.generation < a.max_generations; ++it) {
        const int g = a.generation;
        const int off = (a.current + 1) % 3, nxt = (a.current + 2) % 3;
        moea_population& par = a.pops[a.current];
        moea_population& ofs = a.pops[off];
        moea_population& nx = a.pops[nxt];
        bool parents_changed = false;
        if (cv_switch(a, g)) {
            // DTLZProblem::prepare_parent (dtlz_kernels.cu:1363-1394, algorithm.cu:233-241): the
            // constraints switch on at g, so the parents' violations are recomputed (a fresh
            // evaluation gives the same objectives)
            MOEA_CHECK(evaluate_problem(a, par, st, g));
            if (a.algorithm == MOEA_ALGORITHM_RVEA && a.active_device) {
                hipLaunchKernelGGL(rvea_dead_sentinels, dim3(ceil_div(a.n, 256)), dim3(256), 0, st, par.f, par.cv,
                                   a.n, a.m, a.active_device);
                MOEA_HIP(hipGetLastError());
            }
            parents_changed = true;
        } else if (a.problem.kind == MOEA_PROBLEM_CALLBACK && a.problem.prepare_parent) {
            // IProblemEvaluator::prepare_parent of a callback problem (algorithm.cu:230-241)
            const int rc = a.problem.prepare_parent(a.problem.user, &par, a.n, a.d, a.m, g, a.max_generations,
                                                    a.stream);
            if (rc < 0) return rc;
            parents_changed = rc > 0;
        }
        if (parents_changed) {
            // IEnvironmentSelector::prepare: NSGA-III re-sorts its parent fronts; the built-in
            // RVEA selector keeps its state (rvea_environment_selector.cu:155-161)
            if (a.hooks && a.hooks->select) {
                if (a.hooks->prepare)
                    MOEA_CHECK(a.hooks->prepare(a.hooks->user, &par, a.n, a.d, a.m, g, a.max_generations,
                                                a.stream));
            } else if (a.algorithm == MOEA_ALGORITHM_NSGA3) {
                MOEA_CHECK(parent_fronts(a, par, st));
                a.directions_ready = 0;   // the sort used the workspace: the cone tables are rebuilt
            }
        }
        moea_reproduce_args r{};
        r.n = a.n; r.d = a.d;
        r.active = a.n;
        const moea_strategy_hooks* H = a.hooks;
        const bool hook_select = H && H->select;
        // a caller's selector supplies the mating state (rank and active count)
        r.active_device = (a.algorithm == MOEA_ALGORITHM_RVEA || hook_select) ? a.active_device : nullptr;
        r.mating = a.mating;
        r.parents_x = par.x;
        r.parents_rank = (a.algorithm == MOEA_ALGORITHM_NSGA3 || hook_select) ? a.fronts[a.fronts_current] : nullptr;
        r.parents_cv = par.cv;
        r.bounds = a.bounds;
        r.sbx_eta = annealed(a.sbx_eta[0], a.sbx_eta[1], g, a.max_generations);
        r.sbx_prob = a.sbx_prob;
        r.sbx_copy_prob = a.sbx_copy_prob;
        r.mutation = a.mutation;
        r.pm_eta = annealed(a.pm_eta[0], a.pm_eta[1], g, a.max_generations);
        r.pm_prob = a.pm_prob;
        r.draws = nullptr;
        r.seed = a.seed;
        r.generation = g;
        r.offspring_x = ofs.x;
        r.parent_indices = a.parent_indices;
        r.stream = a.stream;
        if (r.mating == MOEA_MATING_TOURNAMENT && !r.parents_rank)
            return fail(kInvalidArgument, "TournamentMating requires a rank array for every parent");
        stage_begin(kStReproduce, st);
        if (!H || (!H->mate && !H->crossover && !H->mutate)) {
            MOEA_CHECK(reproduce(r));
        } else {
            // ReproductionPipeline::generate (reproduction_pipeline.cu:37-56)
            // with the caller's strategies in place of the built-in ones
            if (H->mate) {
                MOEA_CHECK(H->mate(H->user, &par, a.n, a.d, a.m, r.parents_rank, r.active_device,
                                   a.parent_indices, g, a.max_generations, a.stream));
                r.mating = MOEA_MATING_GIVEN;
            }
            if (H->crossover) {
                if (!H->mate) MOEA_CHECK(mate(r));
                MOEA_CHECK(H->crossover(H->user, &par, a.n, a.d, a.m, a.parent_indices, ofs.x, g,
                                        a.max_generations, a.stream));
                if (!H->mutate && a.mutation == MOEA_MUTATION_POLYNOMIAL)
                    MOEA_CHECK(mutate_rows(ofs.x, ofs.x, a.n, a.d, a.bounds, r.pm_eta, r.pm_prob, a.seed, g, st));
            } else {
                if (H->mutate) r.mutation = MOEA_MUTATION_NONE;
                MOEA_CHECK(reproduce(r));
            }
            if (H->mutate)
                MOEA_CHECK(H->mutate(H->user, &ofs, a.n, a.d, a.m, g, a.max_generations, a.stream));
        }
        stage_end(kStReproduce, st);
        stage_begin(kStEvaluate, st);
        MOEA_CHECK(evaluate(a, ofs, st, g));
        stage_end(kStEvaluate, st);
        if (hook_select) {
            MOEA_CHECK(H->select(H->user, &par, &ofs, &nx, a.n, a.d, a.m, g, a.max_generations, a.stream));
            if (H->update_directions) {
                const int rc = H->update_directions(H->user, &nx, a.n, a.d, a.m, a.active_device,
                                                    const_cast<float*>(a.directions), a.k, g, a.max_generations,
                                                    a.stream);
                if (rc < 0) return rc;
                if (rc > 0) a.directions_ready = 0;
            } else if (a.algorithm == MOEA_ALGORITHM_RVEA && adapt_now(a, g)) {
                MOEA_CHECK(rvea_adapt(a.directions0, nx.f, a.n, a.active_device, a.m, a.k,
                                      const_cast<float*>(a.directions), a.adapt_scratch, st));
                MOEA_CHECK(rvea_gamma(a.directions, a.m, a.k, a.gamma, st));
            }
        } else if (a.algorithm == MOEA_ALGORITHM_NSGA3) {
            moea_nsga3_select_args s{};
            s.n = a.n; s.m = a.m; s.d = a.d; s.k = a.k;
            s.parents = par; s.offspring = ofs; s.next = nx;
            s.next_fronts = a.fronts[1 - a.fronts_current];
            s.directions = a.directions;
            s.ideal = a.ideal;
            s.draws = nullptr;
            s.seed = a.seed;
            s.generation = g;
            s.directions_ready = a.directions_ready;
            s.cv_quant = a.cv_quant;
            s.niching = a.niching;
            s.stats = a.stats;
            s.workspace = a.workspace; s.workspace_bytes = a.workspace_bytes;
            s.stream = a.stream;
            MOEA_CHECK(nsga3_select(s));
            a.directions_ready = 1;
            a.fronts_current = 1 - a.fronts_current;
            if (H && H->update_directions) {
                const int rc = H->update_directions(H->user, &nx, a.n, a.d, a.m, nullptr,
                                                    const_cast<float*>(a.directions), a.k, g, a.max_generations,
                                                    a.stream);
                if (rc < 0) return rc;
                if (rc > 0) a.directions_ready = 0;     // cone tables follow the new directions
            }
        } else {
            moea_rvea_select_args s{};
            s.n = a.n; s.m = a.m; s.d = a.d; s.k = a.k;
            s.active = a.n;
            s.active_device = a.active_device;
            s.parents = par; s.offspring = ofs; s.next = nx;
            s.directions = a.directions;
            s.gamma = a.gamma;
            // EvoX-style t = generation + 1 (rvea_environment_selector.cu:236-239)
            float total = static_cast<float>(a.max_generations);
            float progress = static_cast<float>(g + 1) / (total < 1.0f ? 1.0f : total);
            progress = progress < 0.0f ? 0.0f : (progress > 1.0f ? 1.0f : progress);
            s.penalty = static_cast<float>(a.m) *
                        static_cast<float>(std::pow(static_cast<double>(progress), static_cast<double>(a.rvea_alpha)));
            s.out_active = a.active_device;
            s.workspace = a.workspace; s.workspace_bytes = a.workspace_bytes;
            s.stream = a.stream;
            stage_begin(kStRveaSelect, st);
            MOEA_CHECK(rvea_select(s));
            stage_end(kStRveaSelect, st);
            if (H && H->update_directions) {
                const int rc = H->update_directions(H->user, &nx, a.n, a.d, a.m, a.active_device,
                                                    const_cast<float*>(a.directions), a.k, g, a.max_generations,
                                                    a.stream);
                if (rc < 0) return rc;
                if (rc > 0) MOEA_CHECK(rvea_gamma(a.directions, a.m, a.k, a.gamma, st));
            } else if (adapt_now(a, g)) {
                // AdaptiveRVEADirections::update (reference_directions.cu:132-162)
                stage_begin(kStRveaAdapt, st);
                MOEA_CHECK(rvea_adapt(a.directions0, nx.f, a.n, a.active_device, a.m, a.k,
                                      const_cast<float*>(a.directions), a.adapt_scratch, st));
                MOEA_CHECK(rvea_gamma(a.directions, a.m, a.k, a.gamma, st));
                stage_end(kStRveaAdapt, st);
            }
        }
        a.current = nxt;
        a.generation = g + 1;
    }
    return kOk;
}

}  // namespace moea

extern "C" {

size_t moea_run_workspace_bytes(int algorithm, int n, int m, int k)
{
    if (n <= 0 || m <= 0 || k <= 0) return 0;
    return moea::run_workspace_bytes(algorithm, n, m, k);
}

int moea_run_init(moea_run_args* args)
{
    if (!args) return moea::fail(moea::kInvalidArgument, "moea_run_init: null args");
    return moea::run_init(*args);
}

int moea_run_step(moea_run_args* args, int generations)
{
    if (!args) return moea::fail(moea::kInvalidArgument, "moea_run_step: null args");
    return moea::run_step(*args, generations);
}

}  // extern "C"
