// Batched NSGA-III engine: many independent populations of uneven N / M / D
// advanced together, one launch sequence per generation for the whole batch.
//
// The reference has no batching: each EvolutionaryAlgorithm (NS/algorithm/
// algorithm.cu:104-329) owns its CUDA context and runs one population.  Here
// every instance's buffers are carved out of one caller-provided workspace and
// each stage is one launch over all instances (see batch_internal.hpp), so the
// cost of a generation is one set of launches however many runs it carries.
// Each instance computes exactly what moea_run_init / moea_run_step compute
// for it alone (same kernels bodies, same draws keyed by the instance's seed).
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

int bitlen(long long v)
{
    int b = 0;
    while (v > 0) { ++b; v >>= 1; }
    return b < 1 ? 1 : b;
}

}  // namespace

struct BatchState {
    std::vector<moea_batch_problem> probs;
    std::vector<BInst> host;          // host copy of the instance table
    BInst* dev_insts = nullptr;
    int count = 0;
    hipStream_t stream{};
    // grids
    BGrid init256{}, rows256{}, sbx256{}, rows4_m[kMaxM + 1]{};
    BSelGrids sel{};
    BNdsGrids nds0{}, nds1{};
    BNdsScratch scratch{};
    // host copies of the prefix tables, copied to the device at init
    std::vector<int> tables;
    int* dev_tables = nullptr;
    // run state
    int current = 0, fronts_current = 0, generation = 0, max_generations = 0;
    bool initialized = false;
};

namespace {

// Per-instance block counts -> prefix table appended to `tables`; returns the
// grid with its table offset in the map pointer (fixed up after the copy).
BGrid add_grid(std::vector<int>& tables, const std::vector<long long>& counts)
{
    BGrid g{};
    const size_t off = tables.size();
    long long acc = 0;
    for (long long c : counts) {
        tables.push_back(static_cast<int>(acc));
        acc += c;
    }
    tables.push_back(static_cast<int>(acc));
    g.map.start = reinterpret_cast<const int*>(off);   // offset until fixed up
    g.map.count = static_cast<int>(counts.size());
    g.blocks = static_cast<int>(acc);
    return g;
}

void fix(BGrid& g, int* base)
{
    g.map.start = base + reinterpret_cast<size_t>(g.map.start);
}

struct InstSizes {
    int n, m, d, k, nm, W, Wc, npad, nchunks, nb;
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
    return z;
}

// Lays out everything in the workspace (base may be null to size it).
size_t layout(BatchState& B, char* base)
{
    Carve c(base, ~size_t{0});
    const int count = B.count;
    B.dev_insts = c.take<BInst>(count);
    B.dev_tables = c.take<int>(B.tables.size());
    long long total_merged = 0, total_n = 0;
    for (int i = 0; i < count; ++i) {
        const InstSizes z = sizes_of(B.probs[i]);
        BInst& I = B.host[i];
        for (int s = 0; s < 3; ++s) {
            I.x[s] = c.take<float>(static_cast<size_t>(z.n) * z.d);
            I.f[s] = c.take<float>(static_cast<size_t>(z.m) * z.n);
            I.cv[s] = c.take<float>(z.n);
        }
        I.fronts[0] = c.take<int32_t>(z.n);
        I.fronts[1] = c.take<int32_t>(z.n);
        I.winners = c.take<int32_t>(z.n);
        I.ideal = c.take<float>(kMaxM);
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
        I.cursor = c.take<int32_t>(z.nm);
        I.A = nullptr;
        I.key_off = total_merged;
        I.key_off_n = total_n;
        total_merged += z.nm;
        total_n += z.n;
    }
    BNdsScratch& S = B.scratch;
    S.keys_in = c.take<uint64_t>(total_merged);
    S.keys_out = c.take<uint64_t>(total_merged);
    S.vals_in = c.take<int32_t>(total_merged);
    S.vals_out = c.take<int32_t>(total_merged);
    // temporary storage for the widest of the key widths the batch sorts with
    const int ib = S.inst_bits;
    const int widths[] = {34 + ib, 32 + B.sel.ref_bits + ib, 32 + B.sel.rank_bits + ib, 64};
    S.sort_bytes = 0;
    for (int w : widths)
        S.sort_bytes = std::max(S.sort_bytes, sort_pairs_temp_bytes(static_cast<int>(total_merged), w));
    S.sort_temp = c.take_bytes(S.sort_bytes);
    return c.used;
}

// Fills the host instance records (sizes, problem, parameters) and grids.
int plan(BatchState& B)
{
    const int count = B.count;
    B.host.assign(count, BInst{});
    B.tables.clear();
    std::vector<long long> init256, rows256, sbx256, merged256, objectives, rows4;
    std::vector<std::vector<long long>> rows4_m(kMaxM + 1, std::vector<long long>(count, 0));
    std::vector<std::vector<long long>> prep_m(kMaxM + 1, std::vector<long long>(count, 0));
    std::vector<std::vector<long long>> assoc_m(kMaxM + 1, std::vector<long long>(count, 0));
    std::vector<long long> n0_rows, n0_npad, n0_w, n1_rows, n1_npad, n1_w;
    std::vector<std::vector<long long>> dom0(kMaxM + 1, std::vector<long long>(count, 0));
    std::vector<std::vector<long long>> dom1(kMaxM + 1, std::vector<long long>(count, 0));
    int max_k = 1, max_nm = 1, max_W0 = 1, max_W1 = 1;
    long long total_merged = 0, total_n = 0;
    for (int i = 0; i < count; ++i) {
        const moea_batch_problem& p = B.probs[i];
        const InstSizes z = sizes_of(p);
        BInst& I = B.host[i];
        I.n = z.n; I.m = z.m; I.d = z.d; I.k = z.k;
        I.nm = z.nm; I.W = z.W; I.Wc = z.Wc; I.npad = z.npad;
        I.max_gen = p.max_generations;
        I.nchunks = z.nchunks;
        I.seed = p.seed;
        I.sbx_eta0 = p.sbx_eta[0]; I.sbx_eta1 = p.sbx_eta[1];
        I.sbx_prob = p.sbx_prob; I.sbx_copy = p.sbx_copy_prob;
        I.pm_eta0 = p.pm_eta[0]; I.pm_eta1 = p.pm_eta[1]; I.pm_prob = p.pm_prob;
        I.centers = p.centers; I.bounds = p.bounds; I.dirs = p.directions;
        init256.push_back(ceil_div(static_cast<long long>(z.n) * z.d, 256));
        rows256.push_back(ceil_div(z.n, 256));
        sbx256.push_back(ceil_div(static_cast<long long>(z.n / 2) * z.d + ((z.n & 1) ? z.d : 0), 256));
        merged256.push_back(z.nb);
        objectives.push_back(z.m);
        rows4.push_back(ceil_div(z.n, 4));
        rows4_m[z.m][i] = ceil_div(z.n, 4);
        prep_m[z.m][i] = z.nb;
        assoc_m[z.m][i] = static_cast<long long>(z.nb) * z.nchunks;
        const int W0 = ceil_div(z.n, 64), Wc0 = ceil_div(W0, 8);
        n0_rows.push_back(ceil_div(z.n, 256));
        n0_npad.push_back(ceil_div(W0 * 64, 256));
        n0_w.push_back(W0);
        dom0[z.m][i] = ceil_div(4ll * Wc0 * (Wc0 + 1), 4);
        n1_rows.push_back(ceil_div(z.nm, 256));
        n1_npad.push_back(ceil_div(z.npad, 256));
        n1_w.push_back(z.W);
        dom1[z.m][i] = ceil_div(4ll * z.Wc * (z.Wc + 1), 4);
        max_k = std::max(max_k, z.k);
        max_nm = std::max(max_nm, z.nm);
        max_W0 = std::max(max_W0, W0);
        max_W1 = std::max(max_W1, z.W);
        total_merged += z.nm;
        total_n += z.n;
    }
    std::vector<int>& T = B.tables;
    B.init256 = add_grid(T, init256);
    B.rows256 = add_grid(T, rows256);
    B.sbx256 = add_grid(T, sbx256);
    for (int m = 1; m <= kMaxM; ++m) B.rows4_m[m] = add_grid(T, rows4_m[m]);
    BSelGrids& G = B.sel;
    G.merged256 = add_grid(T, merged256);
    G.rows4 = add_grid(T, rows4);
    G.objectives = add_grid(T, objectives);
    for (int m = 1; m <= kMaxM; ++m) {
        G.norm_prep[m] = add_grid(T, prep_m[m]);
        G.assoc[m] = add_grid(T, assoc_m[m]);
    }
    G.ref_bits = bitlen(max_k + 1);
    G.rank_bits = bitlen(max_nm);
    G.total_merged = total_merged;
    B.nds0.rows256 = add_grid(T, n0_rows);
    B.nds0.npad256 = add_grid(T, n0_npad);
    B.nds0.blocks64 = add_grid(T, n0_w);
    B.nds1.rows256 = add_grid(T, n1_rows);
    B.nds1.npad256 = add_grid(T, n1_npad);
    B.nds1.blocks64 = add_grid(T, n1_w);
    for (int m = 1; m <= kMaxM; ++m) {
        B.nds0.dom[m] = add_grid(T, dom0[m]);
        B.nds1.dom[m] = add_grid(T, dom1[m]);
    }
    B.nds0.max_W = max_W0;
    B.nds1.max_W = max_W1;
    B.nds0.total_rows = total_n;
    B.nds1.total_rows = total_merged;
    B.scratch.inst_bits = bitlen(count - 1);
    // the widest sort key: instance bits above 32 + rank bits (niching) or 34 (NDS)
    if (32 + std::max(G.ref_bits, G.rank_bits) + B.scratch.inst_bits > 64 || 34 + B.scratch.inst_bits > 64)
        return fail(kUnsupported, "moea_batch: too many instances for the sort key width");
    return kOk;
}

void fix_all(BatchState& B)
{
    int* base = B.dev_tables;
    fix(B.init256, base);
    fix(B.rows256, base);
    fix(B.sbx256, base);
    for (int m = 1; m <= kMaxM; ++m) fix(B.rows4_m[m], base);
    fix(B.sel.merged256, base);
    fix(B.sel.rows4, base);
    fix(B.sel.objectives, base);
    for (int m = 1; m <= kMaxM; ++m) {
        fix(B.sel.norm_prep[m], base);
        fix(B.sel.assoc[m], base);
        fix(B.nds0.dom[m], base);
        fix(B.nds1.dom[m], base);
    }
    fix(B.nds0.rows256, base);
    fix(B.nds0.npad256, base);
    fix(B.nds0.blocks64, base);
    fix(B.nds1.rows256, base);
    fix(B.nds1.npad256, base);
    fix(B.nds1.blocks64, base);
}

int validate(const moea_batch_problem* probs, int count)
{
    if (!probs || count <= 0) return fail(kInvalidArgument, "moea_batch: no problems");
    for (int i = 0; i < count; ++i) {
        const moea_batch_problem& p = probs[i];
        if (p.n <= 0 || (p.n & 1) || p.m <= 0 || p.m > kMaxM || p.d <= 0 || p.k <= 0 || p.max_generations <= 0)
            return fail(kInvalidArgument, "moea_batch: invalid sizes (n must be even and positive, 1 <= m <= 10)");
        if (2 * p.n > nds_max_n()) return fail(kUnsupported, "moea_batch: population too large");
        if (!p.centers || !p.bounds || !p.directions)
            return fail(kInvalidArgument, "moea_batch: centers, bounds and directions are required");
        if (p.max_generations != probs[0].max_generations)
            return fail(kInvalidArgument, "moea_batch: every instance of a batch runs the same number of generations");
    }
    return kOk;
}

}  // namespace

size_t batch_workspace_bytes(const moea_batch_problem* probs, int count)
{
    if (validate(probs, count) != kOk) return 0;
    BatchState B;
    B.count = count;
    B.probs.assign(probs, probs + count);
    if (plan(B) != kOk) return 0;
    return layout(B, nullptr);
}

int batch_create(const moea_batch_problem* probs, int count, void* ws, size_t ws_bytes, hipStream_t stream,
                 BatchState** out)
{
    MOEA_CHECK(validate(probs, count));
    if (!out) return fail(kInvalidArgument, "moea_batch_create: null output handle");
    BatchState* B = new (std::nothrow) BatchState();
    if (!B) return fail(kInvalidArgument, "moea_batch_create: out of host memory");
    B->count = count;
    B->probs.assign(probs, probs + count);
    B->stream = stream;
    B->max_generations = probs[0].max_generations;
    int st = plan(*B);
    if (st != kOk) { delete B; return st; }
    const size_t need = layout(*B, nullptr);
    if (!ws || ws_bytes < need) {
        delete B;
        return fail(kWorkspaceTooSmall, "moea_batch_create: workspace too small");
    }
    layout(*B, static_cast<char*>(ws));
    // tables, then the instance records that point into the workspace
    hipError_t e = hipMemcpyAsync(B->dev_tables, B->tables.data(), B->tables.size() * sizeof(int),
                                  hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(B->dev_insts, B->host.data(), B->host.size() * sizeof(BInst), hipMemcpyHostToDevice,
                           stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        delete B;
        return hip_fail(e, "moea_batch_create: table upload");
    }
    fix_all(*B);
    *out = B;
    return kOk;
}

int batch_init(BatchState& B)
{
    BatchLaunch L{B.dev_insts, B.count, B.stream};
    B.current = 0;
    B.fronts_current = 0;
    MOEA_CHECK(batch_init_population(L, B.init256.map, B.init256.blocks, B.current));
    for (int m = 1; m <= kMaxM; ++m)
        MOEA_CHECK(batch_eval_quadratic(L, m, B.rows4_m[m].map, B.rows4_m[m].blocks, B.current));
    MOEA_CHECK(batch_nds(L, B.nds0, B.scratch, 0, B.current, B.fronts_current));
    MOEA_CHECK(batch_ideal(L, B.sel, 0, B.current));
    B.generation = 0;
    B.initialized = true;
    return kOk;
}

int batch_step(BatchState& B, int generations)
{
    if (!B.initialized) return fail(kInvalidArgument, "moea_batch_step: call moea_batch_init first");
    BatchLaunch L{B.dev_insts, B.count, B.stream};
    for (int it = 0; it < generations && B.generation < B.max_generations; ++it) {
        const int g = B.generation;
        const int cur = B.current, off = (cur + 1) % 3, nxt = (cur + 2) % 3;
        stage_begin(kStReproduce, B.stream);
        MOEA_CHECK(batch_reproduce(L, B.rows256.map, B.rows256.blocks, B.sbx256.map, B.sbx256.blocks, cur, off,
                                   B.fronts_current, g));
        stage_end(kStReproduce, B.stream);
        stage_begin(kStEvaluate, B.stream);
        for (int m = 1; m <= kMaxM; ++m)
            MOEA_CHECK(batch_eval_quadratic(L, m, B.rows4_m[m].map, B.rows4_m[m].blocks, off));
        stage_end(kStEvaluate, B.stream);
        MOEA_CHECK(batch_select(L, B.sel, B.nds1, B.scratch, cur, off, nxt, 1 - B.fronts_current, g));
        B.fronts_current = 1 - B.fronts_current;
        B.current = nxt;
        B.generation = g + 1;
    }
    return kOk;
}

}  // namespace moea

using namespace moea;

struct moea_batch {
    BatchState* state;
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
    if (!batch || !batch->state || !out) return fail(kInvalidArgument, "moea_batch_view: null argument");
    const BatchState& B = *batch->state;
    if (index < 0 || index >= B.count) return fail(kInvalidArgument, "moea_batch_view: index out of range");
    const BInst& I = B.host[index];
    out->x = I.x[B.current];
    out->f = I.f[B.current];
    out->cv = I.cv[B.current];
    if (fronts) *fronts = I.fronts[B.fronts_current];
    return kOk;
}

void moea_batch_destroy(moea_batch* batch)
{
    if (!batch) return;
    delete batch->state;
    delete batch;
}

}  // extern "C"
