// Batched and served NSGA-III: many independent populations of uneven N / M /
// D advanced together, one launch sequence per generation for all of them.
//
// The reference has no batching: each EvolutionaryAlgorithm (NS/algorithm/
// algorithm.cu:104-329) owns its stream and runs one population.  Here each
// instance's buffers live in caller-provided device memory, a device table of
// BInst records describes them, and every stage is one launch over the
// instances (batch_internal.hpp).  Instance i computes exactly what
// moea_run_init / moea_run_step compute for it alone: the same kernel bodies
// and draws keyed by its own seed and generation.
//
// Two front ends share the core:
//   moea_batch_*  a fixed set of instances carved from one workspace, all
//                 running the same number of generations;
//   moea_serve_*  continuous batching: instances are admitted into slots while
//                 others run, their buffers allocated from an arena (the HBM
//                 budget) and freed on release; each runs its own generations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "batch_internal.hpp"
#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

constexpr int kAssocChunkB = 1024;   // association chunk (directions per LDS stage)
constexpr int kStateFree = 0, kStateRunning = 1, kStateAdmitted = 2, kStateFinished = 3;

int bitlen(long long v)
{
    int b = 0;
    while (v > 0) { ++b; v >>= 1; }
    return b < 1 ? 1 : b;
}

struct InstSizes {
    int n, m, d, k, nm, W, Wc, npad, nchunks, nb, hidden;
};

InstSizes sizes_of(const moea_batch_problem& p)
{
    InstSizes z{};
    z.n = p.n; z.m = p.m; z.d = p.d; z.k = p.k;
    z.nm = 2 * p.n;
    z.W = ceil_div(z.nm, 64);
    z.Wc = ceil_div(z.W, 8);
    z.npad = z.W * 64;
    z.nchunks = ceil_div(p.k, kAssocChunkB);
    z.nb = ceil_div(z.nm, 256);
    z.hidden = (p.problem == MOEA_PROBLEM_MLP) ? p.hidden : 0;
    return z;
}

// Carves one instance's buffers (c.base may be null to size them).
void carve_instance(Carve& c, const moea_batch_problem& p, BInst& I)
{
    const InstSizes z = sizes_of(p);
    for (int s = 0; s < 3; ++s) {
        I.x[s] = c.take<float>(static_cast<size_t>(z.n) * z.d);
        I.f[s] = c.take<float>(static_cast<size_t>(z.m) * z.n);
        I.cv[s] = c.take<float>(z.n);
    }
    I.fronts[0] = c.take<int32_t>(z.n);
    I.fronts[1] = c.take<int32_t>(z.n);
    I.winners = c.take<int32_t>(z.n);
    I.ideal = c.take<float>(kMaxM);
    I.hwork = z.hidden ? c.take<float>(static_cast<size_t>(z.n) * z.hidden) : nullptr;
    I.fm = c.take<float>(static_cast<size_t>(z.m) * z.nm);
    I.cvm = c.take<float>(z.nm);
    I.fronts_mix = c.take<int32_t>(z.nm);
    I.meta = c.take<int32_t>(8);
    I.gidx = c.take<int32_t>(z.nm);
    I.fn = c.take<float>(static_cast<size_t>(z.m) * z.nm);
    I.asf_val = c.take<float>(static_cast<size_t>(z.nb) * z.m);
    I.asf_idx = c.take<int32_t>(static_cast<size_t>(z.nb) * z.m);
    I.row_max = c.take<float>(static_cast<size_t>(z.nb) * z.m);
    I.a = c.take<float>(kMaxM);
    I.part_d2 = c.take<float>(static_cast<size_t>(z.nchunks) * z.nm);
    I.part_idx = c.take<int32_t>(static_cast<size_t>(z.nchunks) * z.nm);
    I.dist = c.take<float>(z.nm);
    I.ref = c.take<int32_t>(z.nm);
    I.rho = c.take<int32_t>(z.k);
    I.flag = c.take<uint8_t>(z.nm);
    I.chosen = c.take<int32_t>(z.n);
    I.B = c.take<uint64_t>(static_cast<size_t>(z.W) * z.Wc * 512);
    I.fs = c.take<float>(static_cast<size_t>(z.m) * z.npad);
    I.cvs = c.take<float>(z.npad);
    I.perm = c.take<int32_t>(z.nm);
    I.hi = c.take<int32_t>(z.W);
    I.kind = c.take<uint8_t>(z.W);
    I.rend = c.take<int32_t>(z.nm);
    I.A = c.take<uint64_t>(2 * static_cast<size_t>(z.W));
    I.pstate = c.take<uint32_t>(kBatchPeelStateWords);
    I.wp = nullptr; I.cperm = nullptr; I.cones = nullptr; I.cone_buf = nullptr;
    if (assoc_prunable(z.k, z.m)) {
        I.cone_buf = c.take_bytes(cone_bytes(z.m, z.k));
        if (I.cone_buf) cone_views(I.cone_buf, z.m, z.k, &I.wp, &I.cperm, &I.cones);
    }
}

size_t instance_bytes(const moea_batch_problem& p)
{
    Carve c(nullptr, ~size_t{0});
    BInst I{};
    carve_instance(c, p, I);
    return c.used;
}

void fill_record(const moea_batch_problem& p, BInst& I)
{
    const InstSizes z = sizes_of(p);
    I.state = kStateAdmitted;
    I.cur = 0; I.fcur = 0; I.gen = 0;
    I.n = z.n; I.m = z.m; I.d = z.d; I.k = z.k;
    I.nm = z.nm; I.W = z.W; I.Wc = z.Wc; I.npad = z.npad;
    I.max_gen = p.max_generations;
    I.nchunks = z.nchunks;
    I.ddiv = fast_div_make(static_cast<uint32_t>(z.d));
    I.seed = p.seed;
    I.sbx_eta0 = p.sbx_eta[0]; I.sbx_eta1 = p.sbx_eta[1];
    I.sbx_prob = p.sbx_prob; I.sbx_copy = p.sbx_copy_prob;
    I.pm_eta0 = p.pm_eta[0]; I.pm_eta1 = p.pm_eta[1]; I.pm_prob = p.pm_prob;
    I.problem = p.problem;
    I.w1 = p.w1; I.w2 = p.w2; I.hidden = z.hidden;
    I.centers = p.centers; I.bounds = p.bounds; I.dirs = p.directions;
}

int validate_problem(const moea_batch_problem& p)
{
    if (p.n <= 0 || (p.n & 1) || p.m <= 0 || p.m > kMaxM || p.d <= 0 || p.k <= 0 || p.max_generations <= 0)
        return fail(kInvalidArgument, "moea_batch: invalid sizes (n must be even and positive, 1 <= m <= 10)");
    if (2 * p.n > nds_max_n()) return fail(kUnsupported, "moea_batch: population too large");
    if (!p.bounds || !p.directions) return fail(kInvalidArgument, "moea_batch: bounds and directions are required");
    if (p.problem == MOEA_PROBLEM_QUADRATIC) {
        if (!p.centers) return fail(kInvalidArgument, "moea_batch: quadratic problem without centers");
    } else if (p.problem == MOEA_PROBLEM_MLP) {
        if (!p.w1 || !p.w2 || p.hidden <= 0) return fail(kInvalidArgument, "moea_batch: MLP problem needs w1, w2, hidden");
    } else {
        return fail(kInvalidArgument, "moea_batch: problem must be MOEA_PROBLEM_QUADRATIC or MOEA_PROBLEM_MLP");
    }
    return kOk;
}

// Grids of one state (the instances a set of launches serves).
struct Grids {
    BGrid init256{}, rows256{}, sbx256{};
    BEvalGrids eval{};
    BSelGrids sel{};
    BNdsGrids nds0{}, nds1{};
};

// Grid offsets are kept as table indices until the table is on the device.
struct TableBuilder {
    std::vector<int> t;
    // appends `v` and returns its offset as a placeholder pointer (fixed after upload)
    const int* raw(const std::vector<int>& v)
    {
        const size_t off = t.size();
        t.insert(t.end(), v.begin(), v.end());
        return reinterpret_cast<const int*>(off);
    }
    BGrid add(const std::vector<long long>& counts)
    {
        BGrid g{};
        const size_t off = t.size();
        long long acc = 0;
        for (long long c : counts) {
            t.push_back(static_cast<int>(acc));
            acc += c;
        }
        t.push_back(static_cast<int>(acc));
        g.map.start = reinterpret_cast<const int*>(off);
        g.map.count = static_cast<int>(counts.size());
        g.blocks = static_cast<int>(acc);
        return g;
    }
};

void fix(BGrid& g, const int* base) { g.map.start = base + reinterpret_cast<size_t>(g.map.start); }
void fix(const int*& p, const int* base) { p = base + reinterpret_cast<size_t>(p); }

void fix_grids(Grids& G, const int* base)
{
    fix(G.init256, base); fix(G.rows256, base); fix(G.sbx256, base);
    for (int m = 1; m <= kMaxM; ++m) {
        fix(G.eval.quad_rows4[m], base);
        fix(G.eval.mlp_rows4[m], base);
        fix(G.sel.norm_prep[m], base);
        fix(G.sel.assoc[m], base);
        fix(G.sel.assoc_pruned[m], base);
        fix(G.nds0.dom[m], base);
        fix(G.nds1.dom[m], base);
    }
    fix(G.eval.mlp_tiles, base);
    fix(G.sel.merged256, base); fix(G.sel.rows4, base); fix(G.sel.objectives, base);
    fix(G.nds0.rows256, base); fix(G.nds0.npad256, base); fix(G.nds0.blocks64, base);
    fix(G.nds1.rows256, base); fix(G.nds1.npad256, base); fix(G.nds1.blocks64, base);
    fix(G.nds0.peel_first, base); fix(G.nds0.peel_jobs, base);
    fix(G.nds1.peel_first, base); fix(G.nds1.peel_jobs, base);
}

// Builds the grids for the instances in `state` and assigns their sort-key
// offsets (contiguous in slot order).  Returns the rows sorted (sum of 2n),
// or a negative status when the peel cannot be planned.
long long build_grids(std::vector<BInst>& host, int state, int peel_capacity, int peel_rows, TableBuilder& T,
                      Grids& G)
{
    const int count = static_cast<int>(host.size());
    std::vector<long long> init256(count, 0), rows256(count, 0), sbx256(count, 0), merged256(count, 0),
        objectives(count, 0), rows4(count, 0), mlp_tiles(count, 0), n0_rows(count, 0), n0_npad(count, 0),
        n0_w(count, 0), n1_rows(count, 0), n1_npad(count, 0), n1_w(count, 0);
    std::vector<std::vector<long long>> q4(kMaxM + 1, std::vector<long long>(count, 0)),
        h4(kMaxM + 1, std::vector<long long>(count, 0)), prep(kMaxM + 1, std::vector<long long>(count, 0)),
        assoc(kMaxM + 1, std::vector<long long>(count, 0)), dom0(kMaxM + 1, std::vector<long long>(count, 0)),
        dom1(kMaxM + 1, std::vector<long long>(count, 0)), apr(kMaxM + 1, std::vector<long long>(count, 0));
    int max_k = 1, max_nm = 1;
    std::vector<int> peel_inst, peel_W0, peel_W1;
    long long total_merged = 0, total_n = 0;
    for (int i = 0; i < count; ++i) {
        BInst& I = host[i];
        if (I.state != state) continue;
        const int n = I.n, m = I.m, nm = I.nm;
        const int W0 = ceil_div(n, 64), Wc0 = ceil_div(W0, 8);
        const int nb = ceil_div(nm, 256);
        init256[i] = ceil_div(static_cast<long long>(n) * I.d, 256);
        rows256[i] = ceil_div(n, 256);
        sbx256[i] = ceil_div(static_cast<long long>(n / 2) * I.d + ((n & 1) ? I.d : 0), 256);
        merged256[i] = nb;
        objectives[i] = m;
        rows4[i] = ceil_div(n, 4);
        if (I.problem == MOEA_PROBLEM_MLP) {
            mlp_tiles[i] = mlp_tile_count(n, I.hidden);
            h4[m][i] = ceil_div(n, 4);
        } else {
            q4[m][i] = ceil_div(n, 4);
        }
        prep[m][i] = nb;
        if (I.cones) apr[m][i] = ceil_div(nm, 4);
        else assoc[m][i] = static_cast<long long>(nb) * I.nchunks;
        n0_rows[i] = ceil_div(n, 256);
        n0_npad[i] = ceil_div(W0 * 64, 256);
        n0_w[i] = W0;
        dom0[m][i] = ceil_div(4ll * Wc0 * (Wc0 + 1), 4);
        n1_rows[i] = ceil_div(nm, 256);
        n1_npad[i] = ceil_div(I.npad, 256);
        n1_w[i] = I.W;
        dom1[m][i] = ceil_div(4ll * I.Wc * (I.Wc + 1), 4);
        max_k = std::max(max_k, I.k);
        max_nm = std::max(max_nm, nm);
        peel_inst.push_back(i);
        peel_W0.push_back(W0);
        peel_W1.push_back(I.W);
        I.key_off = total_merged;
        I.key_off_n = total_n;
        total_merged += nm;
        total_n += n;
    }
    G.init256 = T.add(init256);
    G.rows256 = T.add(rows256);
    G.sbx256 = T.add(sbx256);
    for (int m = 1; m <= kMaxM; ++m) {
        G.eval.quad_rows4[m] = T.add(q4[m]);
        G.eval.mlp_rows4[m] = T.add(h4[m]);
    }
    G.eval.mlp_tiles = T.add(mlp_tiles);
    G.sel.merged256 = T.add(merged256);
    G.sel.rows4 = T.add(rows4);
    G.sel.objectives = T.add(objectives);
    for (int m = 1; m <= kMaxM; ++m) {
        G.sel.norm_prep[m] = T.add(prep[m]);
        G.sel.assoc[m] = T.add(assoc[m]);
        G.sel.assoc_pruned[m] = T.add(apr[m]);
    }
    G.sel.ref_bits = bitlen(max_k + 1);
    G.sel.rank_bits = bitlen(max_nm);
    G.sel.total_merged = total_merged;
    G.sel.max_nm = static_cast<int>(max_nm);
    G.nds0.rows256 = T.add(n0_rows);
    G.nds0.npad256 = T.add(n0_npad);
    G.nds0.blocks64 = T.add(n0_w);
    G.nds1.rows256 = T.add(n1_rows);
    G.nds1.npad256 = T.add(n1_npad);
    G.nds1.blocks64 = T.add(n1_w);
    for (int m = 1; m <= kMaxM; ++m) {
        G.nds0.dom[m] = T.add(dom0[m]);
        G.nds1.dom[m] = T.add(dom1[m]);
    }
    for (int mode = 0; mode < 2; ++mode) {
        BNdsGrids& N = mode ? G.nds1 : G.nds0;
        std::vector<int> first, jobs;
        size_t lds = 0;
        const int st = batch_peel_plan(peel_inst, mode ? peel_W1 : peel_W0, peel_capacity, peel_rows, first, jobs, lds);
        if (st != kOk) return st;
        N.peel_wgs = static_cast<int>(first.size()) - 1;
        N.peel_lds = lds;
        N.peel_first = T.raw(first);
        N.peel_jobs = T.raw(jobs);
    }
    G.nds0.total_rows = total_n;
    G.nds1.total_rows = total_merged;
    return total_merged;
}

// Prefix tables per state: (1 + 3 + 2 kMaxM + 1 + 3 + 3 kMaxM + 3 + 3 + 2 kMaxM) grids.
constexpr int kGridsPerState = 14 + 7 * kMaxM;

}  // namespace

// Shared state of a batch or a server.
struct BatchCore {
    int slots = 0;
    std::vector<BInst> host;                  // host mirror of the device table
    std::vector<moea_batch_problem> probs;
    BInst* dev_insts = nullptr;
    int* dev_tables = nullptr;
    size_t table_ints = 0;
    BNdsScratch scratch{};
    long long max_rows = 0;                   // capacity of the sort scratch (sum of 2n)
    hipStream_t stream{};
    Grids run{}, init{};
    bool dirty = true;                        // membership changed since the last upload
    int peel_capacity = 0;                    // co-resident blocks of the batched peel
    bool error_ready = false;                 // the device error word has been cleared
    moea_niching niching{};                   // niching path of every selection
    // front-peeling plan: workgroups used = co-resident capacity / peel_cap_div,
    // one workgroup per peel_rows row blocks (chosen by moea_batch_autotune)
    int peel_cap_div = 1, peel_rows = 16;
    // generation graph: one generation's launches captured from the stream and
    // replayed while the set of running instances is unchanged
    bool graph_on = false, graph_valid = false;
    hipGraph_t graph{};
    hipGraphExec_t graph_exec{};
    hipStream_t capture_stream{};             // own stream to capture on when `stream` is the null stream
    int graph_nodes = 0, graph_kernels = 0, graph_memsets = 0, graph_captures = 0;

    BatchCore() = default;
    BatchCore(const BatchCore&) = delete;
    BatchCore& operator=(const BatchCore&) = delete;
    ~BatchCore()
    {
        if (graph_exec || graph) (void)hipStreamSynchronize(stream);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
    }
};

namespace {

size_t core_reserved(int slots, long long max_rows, Carve& c, BatchCore* C)
{
    BInst* insts = c.take<BInst>(slots);
    // grid prefixes of both states, plus the peel plans of both states and modes
    // (at most kBarMaxBlocks + 1 block starts and 256 jobs per instance each)
    const size_t table_ints = static_cast<size_t>(2) * kGridsPerState * (slots + 1) +
                              4 * (static_cast<size_t>(kBarMaxBlocks) + 1 + 3 * 256 * static_cast<size_t>(slots));
    int* tables = c.take<int>(table_ints);
    uint64_t* k_in = c.take<uint64_t>(max_rows);
    uint64_t* k_out = c.take<uint64_t>(max_rows);
    int32_t* v_in = c.take<int32_t>(max_rows);
    int32_t* v_out = c.take<int32_t>(max_rows);
    const int ib = bitlen(slots - 1);
    size_t sort_bytes = 0;
    for (int w = 34 + ib; w <= 64; ++w)
        sort_bytes = std::max(sort_bytes, sort_pairs_temp_bytes(static_cast<int>(max_rows), w));
    void* tmp = c.take_bytes(sort_bytes);
    uint32_t* error = c.take<uint32_t>(1);
    if (C) {
        C->dev_insts = insts;
        C->dev_tables = tables;
        C->table_ints = table_ints;
        C->scratch = BNdsScratch{k_in, k_out, v_in, v_out, tmp, sort_bytes, ib, error};
    }
    return c.used;
}

// The sticky device error word (set when a peel's bounded wait gave up),
// read at a point where the stream is already synchronized.
int core_device_error(const BatchCore& C)
{
    if (!C.error_ready) return kOk;           // no device work queued yet
    uint32_t err = 0;
    MOEA_HIP(hipMemcpy(&err, C.scratch.error, sizeof(err), hipMemcpyDeviceToHost));
    if (err != 0)
        return fail(kDeviceFault, "moea_batch: a front-peeling launch gave up waiting at its grid barrier; "
                                  "fronts of the affected generation are invalid");
    return kOk;
}

// Clears the error word before the core's first device work (not at
// creation, which touches no device memory).
int core_clear_error(BatchCore& C)
{
    if (C.error_ready) return kOk;
    const uint32_t zero = 0;
    MOEA_HIP(hipMemcpy(C.scratch.error, &zero, sizeof(zero), hipMemcpyHostToDevice));
    C.error_ready = true;
    return kOk;
}

// Rebuilds both grid sets and uploads the tables and the instance table
// (synchronizing the stream first: the device table may be in use).
int core_upload(BatchCore& C)
{
    if (C.peel_capacity == 0) C.peel_capacity = batch_peel_capacity();
    if (C.peel_capacity <= 0) return fail(kUnsupported, "moea_batch: the peel kernel cannot be resident");
    C.graph_valid = false;                    // the grids and tables change: recapture
    const int cap = std::max(1, C.peel_capacity / std::max(1, C.peel_cap_div));
    TableBuilder T;
    const long long run_rows = build_grids(C.host, kStateRunning, cap, C.peel_rows, T, C.run);
    if (run_rows < 0) return static_cast<int>(run_rows);
    const long long init_rows = build_grids(C.host, kStateAdmitted, cap, C.peel_rows, T, C.init);
    if (init_rows < 0) return static_cast<int>(init_rows);
    if (run_rows > C.max_rows || init_rows > C.max_rows)
        return fail(kWorkspaceTooSmall, "moea_batch: more rows than the sort scratch holds");
    if (T.t.size() > C.table_ints) return fail(kInvalidArgument, "moea_batch: table capacity exceeded");
    MOEA_HIP(hipStreamSynchronize(C.stream));
    MOEA_CHECK(core_device_error(C));
    MOEA_CHECK(core_clear_error(C));
    MOEA_HIP(hipMemcpy(C.dev_tables, T.t.data(), T.t.size() * sizeof(int), hipMemcpyHostToDevice));
    MOEA_HIP(hipMemcpy(C.dev_insts, C.host.data(), C.host.size() * sizeof(BInst), hipMemcpyHostToDevice));
    fix_grids(C.run, C.dev_tables);
    fix_grids(C.init, C.dev_tables);
    C.dirty = false;
    return kOk;
}

bool any_state(const BatchCore& C, int state)
{
    for (const BInst& I : C.host)
        if (I.state == state) return true;
    return false;
}

// Initializes the admitted instances (state 2 -> 1): initial population,
// evaluation at generation -1, parent fronts and ideal point.
int core_init_admitted(BatchCore& C)
{
    if (!any_state(C, kStateAdmitted)) return kOk;
    if (C.dirty) MOEA_CHECK(core_upload(C));
    BatchLaunch L{C.dev_insts, C.slots, C.stream, C.niching};
    for (const BInst& I : C.host)
        if (I.state == kStateAdmitted && I.cone_buf) MOEA_CHECK(cones_build(I.dirs, I.k, I.m, I.cone_buf, C.stream));
    MOEA_CHECK(batch_init_population(L, C.init.init256.map, C.init.init256.blocks));
    MOEA_CHECK(batch_evaluate(L, C.init.eval, 0));
    MOEA_CHECK(batch_nds(L, C.init.nds0, C.scratch, 0, kStateAdmitted));
    MOEA_CHECK(batch_ideal(L, C.init.sel, 0));
    MOEA_CHECK(batch_advance(L, kStateAdmitted, kStateRunning, 0));
    for (BInst& I : C.host)
        if (I.state == kStateAdmitted) I.state = kStateRunning;
    C.dirty = true;
    return kOk;
}

// The launches of one generation of every running instance.  They take
// every per-generation quantity (generation, buffer rotation, annealed
// operator parameters, draw keys) from the device instance table, which the
// last launch advances, so the same sequence serves every generation until
// the set of running instances changes.
int core_step_launch(BatchCore& C, hipStream_t st)
{
    BatchLaunch L{C.dev_insts, C.slots, st, C.niching};
    stage_begin(kStReproduce, st);
    MOEA_CHECK(batch_reproduce(L, C.run.rows256.map, C.run.rows256.blocks, C.run.sbx256.map, C.run.sbx256.blocks));
    stage_end(kStReproduce, st);
    stage_begin(kStEvaluate, st);
    MOEA_CHECK(batch_evaluate(L, C.run.eval, 1));
    stage_end(kStEvaluate, st);
    MOEA_CHECK(batch_select(L, C.run.sel, C.run.nds1, C.scratch));
    MOEA_CHECK(batch_advance(L, kStateRunning, kStateRunning, 1));
    return kOk;
}

void graph_release(BatchCore& C)
{
    if (C.graph_exec) (void)hipGraphExecDestroy(C.graph_exec);
    if (C.graph) (void)hipGraphDestroy(C.graph);
    C.graph_exec = nullptr;
    C.graph = nullptr;
    C.graph_valid = false;
}

// Captures core_step_launch into a graph.  Only kernel nodes are accepted: a
// memset, copy or host node would need its own replay semantics.  The null
// stream cannot be captured, so a batch launched on it captures on a stream
// of its own (capture runs nothing) and replays on its stream.
int core_capture(BatchCore& C)
{
    MOEA_HIP(hipStreamSynchronize(C.stream));   // the old executable may be in flight
    graph_release(C);
    hipStream_t cs = C.stream;
    if (!cs) {
        if (!C.capture_stream) MOEA_HIP(hipStreamCreateWithFlags(&C.capture_stream, hipStreamNonBlocking));
        cs = C.capture_stream;
    }
    MOEA_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const int st = core_step_launch(C, cs);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    if (st != kOk) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (ec != hipSuccess || !g) return hip_fail(ec != hipSuccess ? ec : hipErrorUnknown, "hipStreamEndCapture");
    size_t count = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &count);
    std::vector<hipGraphNode_t> nodes(count);
    if (e == hipSuccess && count) e = hipGraphGetNodes(g, nodes.data(), &count);
    int kernels = 0, memsets = 0, other = 0;
    for (size_t i = 0; e == hipSuccess && i < count; ++i) {
        hipGraphNodeType t{};
        e = hipGraphNodeGetType(nodes[i], &t);
        if (t == hipGraphNodeTypeKernel) ++kernels;
        else if (t == hipGraphNodeTypeMemset) ++memsets;
        else ++other;
    }
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphGetNodes");
    }
    C.graph_nodes = static_cast<int>(count);
    C.graph_kernels = kernels;
    C.graph_memsets = memsets;
    if (memsets || other) {
        (void)hipGraphDestroy(g);
        return fail(kUnsupported, "moea_batch: the captured generation holds non-kernel nodes");
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate");
    }
    C.graph = g;
    C.graph_exec = x;
    C.graph_valid = true;
    C.graph_captures += 1;
    return kOk;
}

// One generation of every running instance; instances reaching their
// generation count become finished.  Appends finished slots to `done`.
int core_step(BatchCore& C, std::vector<int>* done)
{
    if (!any_state(C, kStateRunning)) return kOk;
    if (C.dirty) MOEA_CHECK(core_upload(C));
    if (C.graph_on) {
        if (!C.graph_valid) MOEA_CHECK(core_capture(C));
        MOEA_HIP(hipGraphLaunch(C.graph_exec, C.stream));
    } else {
        MOEA_CHECK(core_step_launch(C, C.stream));
    }
    for (int i = 0; i < C.slots; ++i) {
        BInst& I = C.host[i];
        if (I.state != kStateRunning) continue;
        I.cur = (I.cur + 2) % 3;
        I.fcur = 1 - I.fcur;
        I.gen += 1;
        if (I.gen >= I.max_gen) {
            I.state = kStateFinished;
            C.dirty = true;
            if (done) done->push_back(i);
        }
    }
    return kOk;
}

int core_view(const BatchCore& C, int slot, moea_population* out, int32_t** fronts)
{
    if (slot < 0 || slot >= C.slots || !out) return fail(kInvalidArgument, "moea_batch: slot out of range");
    const BInst& I = C.host[slot];
    if (I.state == kStateFree) return fail(kInvalidArgument, "moea_batch: slot is free");
    out->x = I.x[I.cur];
    out->f = I.f[I.cur];
    out->cv = I.cv[I.cur];
    if (fronts) *fronts = I.fronts[I.fcur];
    return kOk;
}

}  // namespace

// ------------------------------------------------------------------ batch --
struct BatchState {
    BatchCore core;
    int generation = 0, max_generations = 0;
};

namespace {

int validate_batch(const moea_batch_problem* probs, int count)
{
    if (!probs || count <= 0) return fail(kInvalidArgument, "moea_batch: no problems");
    for (int i = 0; i < count; ++i) {
        MOEA_CHECK(validate_problem(probs[i]));
        if (probs[i].max_generations != probs[0].max_generations)
            return fail(kInvalidArgument, "moea_batch: every instance of a batch runs the same number of generations");
        if (probs[i].niching.mode != probs[0].niching.mode || probs[i].niching.lds_cap != probs[0].niching.lds_cap)
            return fail(kInvalidArgument, "moea_batch: every instance of a batch uses the same niching path");
    }
    if (!niching_valid(probs[0].niching)) return fail(kInvalidArgument, "moea_batch: invalid niching mode or capacity");
    return kOk;
}

size_t batch_layout(const moea_batch_problem* probs, int count, char* base, BatchCore* C)
{
    long long rows = 0;
    for (int i = 0; i < count; ++i) rows += 2ll * probs[i].n;
    Carve c(base, ~size_t{0});
    core_reserved(count, rows, c, C);
    for (int i = 0; i < count; ++i) {
        BInst I{};
        carve_instance(c, probs[i], I);
        if (C) {
            fill_record(probs[i], I);
            C->host[i] = I;
        }
    }
    return c.used;
}

}  // namespace

size_t batch_workspace_bytes(const moea_batch_problem* probs, int count)
{
    if (validate_batch(probs, count) != kOk) return 0;
    return batch_layout(probs, count, nullptr, nullptr);
}

int batch_create(const moea_batch_problem* probs, int count, void* ws, size_t ws_bytes, hipStream_t stream,
                 BatchState** out)
{
    MOEA_CHECK(validate_batch(probs, count));
    if (!out) return fail(kInvalidArgument, "moea_batch_create: null output handle");
    const size_t need = batch_layout(probs, count, nullptr, nullptr);
    if (!ws || ws_bytes < need) return fail(kWorkspaceTooSmall, "moea_batch_create: workspace too small");
    BatchState* B = new (std::nothrow) BatchState();
    if (!B) return fail(kInvalidArgument, "moea_batch_create: out of host memory");
    BatchCore& C = B->core;
    C.slots = count;
    C.host.assign(count, BInst{});
    C.probs.assign(probs, probs + count);
    C.stream = stream;
    for (int i = 0; i < count; ++i) C.max_rows += 2ll * probs[i].n;
    batch_layout(probs, count, static_cast<char*>(ws), &C);
    C.niching = probs[0].niching;
    B->max_generations = probs[0].max_generations;
    *out = B;
    return kOk;
}

int batch_init(BatchState& B)
{
    BatchCore& C = B.core;
    for (int i = 0; i < C.slots; ++i) {
        BInst& I = C.host[i];
        I.state = kStateAdmitted;
        I.cur = 0; I.fcur = 0; I.gen = 0;
    }
    C.dirty = true;
    MOEA_CHECK(core_init_admitted(C));
    B.generation = 0;
    return kOk;
}

int batch_step(BatchState& B, int generations)
{
    if (!any_state(B.core, kStateRunning) && B.generation == 0)
        return fail(kInvalidArgument, "moea_batch_step: call moea_batch_init first");
    for (int it = 0; it < generations && B.generation < B.max_generations; ++it) {
        MOEA_CHECK(core_step(B.core, nullptr));
        B.generation += 1;
    }
    return kOk;
}

// Peel plans the tuner tries: (capacity divisor, row blocks per workgroup).
constexpr int kTuneCandidates = 4;
constexpr int kTunePlan[kTuneCandidates][2] = {{1, 16}, {1, 32}, {2, 16}, {2, 32}};

// Times each peel plan on the batch's own next generations (two passes over
// the candidates, forward then backward, `trials` generations per visit,
// the first of each visit untimed: it uploads the plan and recaptures) and
// keeps the fastest.  The plan only decides which workgroup peels which rows,
// so the generations it runs are the run's own.  Returns the generations run.
int batch_autotune(BatchState& B, int trials, int* chosen, double* ms_out, int max_out)
{
    BatchCore& C = B.core;
    if (trials < 2) return fail(kInvalidArgument, "moea_batch_autotune: trials must be at least 2");
    if (!any_state(C, kStateRunning)) return fail(kInvalidArgument, "moea_batch_autotune: call moea_batch_init first");
    double total[kTuneCandidates] = {};
    int visits[kTuneCandidates] = {};
    bool valid[kTuneCandidates];
    const int keep_div = C.peel_cap_div, keep_rows = C.peel_rows;
    for (int c = 0; c < kTuneCandidates; ++c) {
        // a plan that cannot hold the largest instance is skipped
        C.peel_cap_div = kTunePlan[c][0];
        C.peel_rows = kTunePlan[c][1];
        C.dirty = true;
        valid[c] = core_upload(C) == kOk;
    }
    C.peel_cap_div = keep_div;
    C.peel_rows = keep_rows;
    C.dirty = true;
    hipEvent_t a = nullptr, b = nullptr;
    MOEA_HIP(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(kHipError, "moea_batch_autotune: hipEventCreate failed");
    }
    int ran = 0, status = kOk;
    for (int pass = 0; pass < 2 && status == kOk; ++pass) {
        for (int v = 0; v < kTuneCandidates && status == kOk; ++v) {
            const int c = pass == 0 ? v : kTuneCandidates - 1 - v;
            if (!valid[c] || B.generation + trials > B.max_generations) continue;
            C.peel_cap_div = kTunePlan[c][0];
            C.peel_rows = kTunePlan[c][1];
            C.dirty = true;
            status = core_step(C, nullptr);
            if (status != kOk) break;
            B.generation += 1;
            if (hipEventRecord(a, C.stream) != hipSuccess) { status = fail(kHipError, "hipEventRecord"); break; }
            for (int t = 1; t < trials && status == kOk; ++t) {
                status = core_step(C, nullptr);
                B.generation += 1;
            }
            if (status != kOk) break;
            float ms = 0.0f;
            if (hipEventRecord(b, C.stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                status = fail(kHipError, "moea_batch_autotune: event timing failed");
                break;
            }
            total[c] += ms / (trials - 1);
            visits[c] += 1;
            ran += trials;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (status != kOk) return status;
    int best = -1;
    for (int c = 0; c < kTuneCandidates; ++c) {
        const double ms = visits[c] ? total[c] / visits[c] : -1.0;
        if (ms_out && c < max_out) ms_out[c] = ms;
        if (visits[c] && (best < 0 || ms < total[best] / visits[best])) best = c;
    }
    if (best >= 0) {
        C.peel_cap_div = kTunePlan[best][0];
        C.peel_rows = kTunePlan[best][1];
        C.dirty = true;
    }
    if (chosen) *chosen = best;
    return ran;
}

// ------------------------------------------------------------------ serve --
struct ServeState {
    BatchCore core;
    size_t arena_off = 0, arena_bytes = 0;    // instance arena inside the workspace
    char* base = nullptr;
    std::vector<std::pair<size_t, size_t>> free_list;   // (offset, bytes), sorted by offset
    std::vector<std::pair<size_t, size_t>> owned;       // per slot: (offset, bytes)
};

namespace {

constexpr size_t kArenaAlign = 256;

bool arena_alloc(ServeState& S, size_t bytes, size_t& off)
{
    bytes = align_up(bytes, kArenaAlign);
    for (size_t i = 0; i < S.free_list.size(); ++i) {
        auto& r = S.free_list[i];
        if (r.second >= bytes) {
            off = r.first;
            r.first += bytes;
            r.second -= bytes;
            if (r.second == 0) S.free_list.erase(S.free_list.begin() + static_cast<long>(i));
            return true;
        }
    }
    return false;
}

void arena_free(ServeState& S, size_t off, size_t bytes)
{
    bytes = align_up(bytes, kArenaAlign);
    auto it = std::lower_bound(S.free_list.begin(), S.free_list.end(), std::make_pair(off, size_t{0}));
    it = S.free_list.insert(it, {off, bytes});
    // coalesce with the next and the previous range
    auto nx = it + 1;
    if (nx != S.free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        S.free_list.erase(nx);
    }
    if (it != S.free_list.begin()) {
        auto pv = it - 1;
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            S.free_list.erase(it);
        }
    }
}

}  // namespace

size_t serve_reserved_bytes(int max_instances, long long max_rows)
{
    if (max_instances <= 0 || max_rows <= 0) return 0;
    Carve c(nullptr, ~size_t{0});
    return align_up(core_reserved(max_instances, max_rows, c, nullptr), kArenaAlign);
}

int serve_create(const moea_serve_config& cfg, void* ws, size_t ws_bytes, hipStream_t stream, ServeState** out)
{
    if (!out) return fail(kInvalidArgument, "moea_serve_create: null output handle");
    if (cfg.max_instances <= 0 || cfg.max_rows <= 0 || cfg.max_rows > (1ll << 30) || !niching_valid(cfg.niching))
        return fail(kInvalidArgument, "moea_serve_create: invalid configuration");
    const size_t reserved = serve_reserved_bytes(cfg.max_instances, cfg.max_rows);
    if (!ws || ws_bytes <= reserved) return fail(kWorkspaceTooSmall, "moea_serve_create: workspace too small");
    ServeState* S = new (std::nothrow) ServeState();
    if (!S) return fail(kInvalidArgument, "moea_serve_create: out of host memory");
    BatchCore& C = S->core;
    C.slots = cfg.max_instances;
    C.host.assign(C.slots, BInst{});
    C.probs.assign(C.slots, moea_batch_problem{});
    C.stream = stream;
    C.max_rows = cfg.max_rows;
    C.niching = cfg.niching;
    Carve c(ws, ws_bytes);
    core_reserved(C.slots, C.max_rows, c, &C);
    S->base = static_cast<char*>(ws);
    S->arena_off = reserved;
    S->arena_bytes = ws_bytes - reserved;
    S->free_list.push_back({reserved, S->arena_bytes});
    S->owned.assign(C.slots, {0, 0});
    *out = S;
    return kOk;
}

int serve_admit(ServeState& S, const moea_batch_problem& p, int* slot_out)
{
    MOEA_CHECK(validate_problem(p));
    BatchCore& C = S.core;
    int slot = -1;
    for (int i = 0; i < C.slots; ++i)
        if (C.host[i].state == kStateFree) { slot = i; break; }
    if (slot < 0) return fail(kWorkspaceTooSmall, "moea_serve_admit: no free slot");
    long long rows_run = 0, rows_init = 0;
    for (const BInst& I : C.host) {
        if (I.state == kStateRunning || I.state == kStateAdmitted) rows_run += I.nm;
        if (I.state == kStateAdmitted) rows_init += I.n;
    }
    if (rows_run + 2ll * p.n > C.max_rows) return fail(kWorkspaceTooSmall, "moea_serve_admit: row capacity full");
    const size_t bytes = instance_bytes(p);
    size_t off = 0;
    if (!arena_alloc(S, bytes, off)) return fail(kWorkspaceTooSmall, "moea_serve_admit: HBM budget full");
    Carve c(S.base + off, bytes);
    BInst I{};
    carve_instance(c, p, I);
    fill_record(p, I);
    C.host[slot] = I;
    C.probs[slot] = p;
    S.owned[slot] = {off, bytes};
    C.dirty = true;
    (void)rows_init;
    *slot_out = slot;
    return kOk;
}

int serve_step(ServeState& S, std::vector<int>& done)
{
    MOEA_CHECK(core_init_admitted(S.core));
    return core_step(S.core, &done);
}

int serve_release(ServeState& S, int slot)
{
    BatchCore& C = S.core;
    if (slot < 0 || slot >= C.slots || C.host[slot].state == kStateFree)
        return fail(kInvalidArgument, "moea_serve_release: slot is not in use");
    // the device may still use the buffers: order the release after queued work
    MOEA_HIP(hipStreamSynchronize(C.stream));
    MOEA_CHECK(core_device_error(C));
    arena_free(S, S.owned[slot].first, S.owned[slot].second);
    S.owned[slot] = {0, 0};
    C.host[slot] = BInst{};
    C.dirty = true;
    return kOk;
}

}  // namespace moea

using namespace moea;

struct moea_batch {
    BatchState* state;
};

struct moea_serve {
    ServeState* state;
};

extern "C" {

size_t moea_batch_workspace_bytes(const moea_batch_problem* problems, int count)
{
    return batch_workspace_bytes(problems, count);
}

int moea_batch_create(const moea_batch_problem* problems, int count, void* workspace, size_t workspace_bytes,
                      void* stream, moea_batch** out)
{
    if (!out) return fail(kInvalidArgument, "moea_batch_create: null output handle");
    *out = nullptr;
    BatchState* st = nullptr;
    MOEA_CHECK(batch_create(problems, count, workspace, workspace_bytes, static_cast<hipStream_t>(stream), &st));
    moea_batch* b = new (std::nothrow) moea_batch{st};
    if (!b) {
        delete st;
        return fail(kInvalidArgument, "moea_batch_create: out of host memory");
    }
    *out = b;
    return kOk;
}

int moea_batch_init(moea_batch* batch)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_init: null batch");
    return batch_init(*batch->state);
}

int moea_batch_step(moea_batch* batch, int generations)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_step: null batch");
    return batch_step(*batch->state, generations);
}

int moea_batch_generation(const moea_batch* batch)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_generation: null batch");
    return batch->state->generation;
}

int moea_batch_view(const moea_batch* batch, int index, moea_population* out, int32_t** fronts)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_view: null batch");
    return core_view(batch->state->core, index, out, fronts);
}

int moea_batch_check(const moea_batch* batch)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_check: null batch");
    MOEA_HIP(hipStreamSynchronize(batch->state->core.stream));
    return core_device_error(batch->state->core);
}

int moea_batch_set_graph(moea_batch* batch, int enable)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_set_graph: null batch");
    batch->state->core.graph_on = enable != 0;
    return kOk;
}

int moea_batch_graph_info(const moea_batch* batch, int* kernel_nodes, int* memset_nodes, int* captures)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_graph_info: null batch");
    const BatchCore& C = batch->state->core;
    if (kernel_nodes) *kernel_nodes = C.graph_kernels;
    if (memset_nodes) *memset_nodes = C.graph_memsets;
    if (captures) *captures = C.graph_captures;
    return C.graph_nodes;
}

int moea_batch_autotune(moea_batch* batch, int trials, int* chosen, double* ms_per_generation, int max_candidates)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_autotune: null batch");
    return batch_autotune(*batch->state, trials, chosen, ms_per_generation, max_candidates);
}

int moea_batch_peel_plan(const moea_batch* batch, int* capacity_divisor, int* rows_per_workgroup)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_peel_plan: null batch");
    const BatchCore& C = batch->state->core;
    if (capacity_divisor) *capacity_divisor = C.peel_cap_div;
    if (rows_per_workgroup) *rows_per_workgroup = C.peel_rows;
    return kOk;
}

int moea_batch_set_peel_plan(moea_batch* batch, int capacity_divisor, int rows_per_workgroup)
{
    if (!batch || !batch->state) return fail(kInvalidArgument, "moea_batch_set_peel_plan: null batch");
    if (capacity_divisor < 1 || capacity_divisor > 64 || rows_per_workgroup < 1 || rows_per_workgroup > 32)
        return fail(kInvalidArgument, "moea_batch_set_peel_plan: divisor in [1, 64], rows in [1, 32]");
    BatchCore& C = batch->state->core;
    C.peel_cap_div = capacity_divisor;
    C.peel_rows = rows_per_workgroup;
    C.dirty = true;
    return kOk;
}

void moea_batch_destroy(moea_batch* batch)
{
    if (!batch) return;
    delete batch->state;
    delete batch;
}

size_t moea_serve_reserved_bytes(const moea_serve_config* config)
{
    if (!config) return 0;
    return serve_reserved_bytes(config->max_instances, config->max_rows);
}

size_t moea_serve_instance_bytes(const moea_batch_problem* problem)
{
    if (!problem || validate_problem(*problem) != kOk) return 0;
    return align_up(instance_bytes(*problem), 256);
}

int moea_serve_create(const moea_serve_config* config, void* workspace, size_t workspace_bytes, void* stream,
                      moea_serve** out)
{
    if (!config || !out) return fail(kInvalidArgument, "moea_serve_create: null argument");
    *out = nullptr;
    ServeState* st = nullptr;
    MOEA_CHECK(serve_create(*config, workspace, workspace_bytes, static_cast<hipStream_t>(stream), &st));
    moea_serve* s = new (std::nothrow) moea_serve{st};
    if (!s) {
        delete st;
        return fail(kInvalidArgument, "moea_serve_create: out of host memory");
    }
    *out = s;
    return kOk;
}

int moea_serve_admit(moea_serve* server, const moea_batch_problem* problem, int* slot)
{
    if (!server || !server->state || !problem || !slot) return fail(kInvalidArgument, "moea_serve_admit: null argument");
    return serve_admit(*server->state, *problem, slot);
}

int moea_serve_step(moea_serve* server, int32_t* finished, int max_finished, int* n_finished)
{
    if (!server || !server->state || !n_finished) return fail(kInvalidArgument, "moea_serve_step: null argument");
    std::vector<int> done;
    MOEA_CHECK(serve_step(*server->state, done));
    const int n = static_cast<int>(done.size());
    *n_finished = n;
    if (finished)
        for (int i = 0; i < n && i < max_finished; ++i) finished[i] = done[i];
    return kOk;
}

int moea_serve_view(const moea_serve* server, int slot, moea_population* out, int32_t** fronts)
{
    if (!server || !server->state) return fail(kInvalidArgument, "moea_serve_view: null server");
    return core_view(server->state->core, slot, out, fronts);
}

int moea_serve_generation(const moea_serve* server, int slot)
{
    if (!server || !server->state) return fail(kInvalidArgument, "moea_serve_generation: null server");
    const BatchCore& C = server->state->core;
    if (slot < 0 || slot >= C.slots || C.host[slot].state == 0)
        return fail(kInvalidArgument, "moea_serve_generation: slot is not in use");
    return C.host[slot].gen;
}

int moea_serve_release(moea_serve* server, int slot)
{
    if (!server || !server->state) return fail(kInvalidArgument, "moea_serve_release: null server");
    return serve_release(*server->state, slot);
}

int moea_serve_check(const moea_serve* server)
{
    if (!server || !server->state) return fail(kInvalidArgument, "moea_serve_check: null server");
    MOEA_HIP(hipStreamSynchronize(server->state->core.stream));
    return core_device_error(server->state->core);
}

int moea_serve_set_graph(moea_serve* server, int enable)
{
    if (!server || !server->state) return fail(kInvalidArgument, "moea_serve_set_graph: null server");
    server->state->core.graph_on = enable != 0;
    return kOk;
}

void moea_serve_destroy(moea_serve* server)
{
    if (!server) return;
    delete server->state;
    delete server;
}

}  // extern "C"
