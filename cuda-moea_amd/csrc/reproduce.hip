// Mating -> SBX -> polynomial mutation, fused into two coalesced kernels.
//
// Replaces ReproductionPipeline::generate (NS/reproduction/reproduction_pipeline.cu:37-56):
// TournamentMating / RandomMating (mating_selection.cu:11-108), SBX
// (sbx_crossover.cu:20-190) and PM (polynomial_mutation.cu:20-124).  The
// reference materializes every draw array with separate cuRAND kernels; here
// draws are either injected or generated in registers (Philox4x32-10), and the
// crossover and mutation of a pair happen in one pass over its two rows.
#include "batch_internal.hpp"
#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

// Mating winner of slot t (candidates from [0, active)).
__device__ __forceinline__ int mating_one(int t, int n, int active, int mating, const int32_t* __restrict__ rank,
                                 const float* __restrict__ cv, const int32_t* __restrict__ inj_a,
                                 const int32_t* __restrict__ inj_b, uint64_t seed, uint32_t gen)
{
    active = active < 1 ? 1 : (active > n ? n : active);
    int a, b;
    if (inj_a) {
        // injected candidates are clamped into [0, n) so a bad draw cannot read out of bounds
        a = min(max(inj_a[t], 0), n - 1);
        b = inj_b ? min(max(inj_b[t], 0), n - 1) : a;
    } else {
        const U4 r = draw4(seed, static_cast<uint64_t>(t), gen, kStMate);
        a = static_cast<int>(r.x % static_cast<uint32_t>(active));
        b = static_cast<int>(r.y % static_cast<uint32_t>(active));
    }
    if (mating == MOEA_MATING_RANDOM) return a;
    // tournament_winner_kernel (mating_selection.cu:11-35)
    const float ca = cv[a], cb = cv[b];
    if (ca == 0.0f && cb != 0.0f) return a;
    if (ca != 0.0f && cb == 0.0f) return b;
    if (ca != 0.0f && cb != 0.0f) return (ca <= cb) ? a : b;
    return (rank[a] <= rank[b]) ? a : b;
}

__global__ void mating_kernel(int n, int active_host, const int32_t* __restrict__ active_dev, int mating,
                              const int32_t* __restrict__ rank,
                              const float* __restrict__ cv, const int32_t* __restrict__ inj_a,
                              const int32_t* __restrict__ inj_b, uint64_t seed, uint32_t gen,
                              int32_t* __restrict__ winners)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int active = active_dev ? active_dev[0] : active_host;
    winners[t] = mating_one(t, n, active, mating, rank, cv, inj_a, inj_b, seed, gen);
}

// SBX of one (pair, variable) element.  In midpoint +/- delta both operands are
// products; the NVPTX combiner fuses the first (0.5f * (x1 + x2), an exact
// product), so the sums evaluate unfused as written.
// power = 1 / (1 + eta), formed once per launch (set_derived).
__device__ __forceinline__ void sbx_one(float x1, float x2, float ub, float up, int sign, float uc, float lo,
                               float hi, float power, float prob, float copy_prob, float& ca, float& cb)
{
    if (up >= prob || uc < copy_prob) { ca = x1; cb = x2; return; }
    // one powf_root on the selected base (a wave holds lanes of both branches)
    const float base = (ub < 0.5f) ? 2.0f * ub : 1.0f / (2.0f - 2.0f * ub);
    float beta = powf_root(base, power);
    beta = beta * static_cast<float>(2 * sign - 1);
    const float mid = 0.5f * (x1 + x2);
    const float delta = 0.5f * beta * (x1 - x2);
    ca = fminf(fmaxf(mid + delta, lo), hi);
    cb = fminf(fmaxf(mid - delta, lo), hi);
}

// Deb's polynomial mutation of one element; value + delta * range is fused.
// exponent = eta + 1 and inv_exponent = 1 / exponent, formed once per launch.
__device__ __forceinline__ float pm_one(float value, float ug, float ud, float lo, float hi, float exponent,
                                        float inv_exponent, float rate)
{
    if (ug >= rate) return value;
    const float range = hi - lo;
    const float normalized = (value - lo) / range;
    // both branches' operands selected first, then one power and one root:
    //   ud <= 0.5: vm = 2 ud + (1 - 2 ud) (1 - x)^e,          delta = vm^(1/e) - 1
    //   else:      vm = 2 (1 - ud) + 2 (ud - 0.5) x^e,        delta = 1 - vm^(1/e)
    // (x^e with e > 1 can reach far below 2^-126: evaluated in double)
    const bool lower = ud <= 0.5f;
    const float xy = lower ? 1.0f - normalized : normalized;
    const float a = lower ? 2.0f * ud : 2.0f * (1.0f - ud);
    const float b = lower ? 1.0f - 2.0f * ud : 2.0f * (ud - 0.5f);
    const float vm = a + b * powf_det(xy, exponent);
    const float pw = powf_root(vm, inv_exponent);
    const float delta = lower ? pw - 1.0f : 1.0f - pw;
    return fminf(fmaxf(fmaf(delta, range, value), lo), hi);
}

struct RepArgs {
    int n, d, pairs;
    FastDiv dd;          // division by d
    int small;           // pairs * d + d < 2^31: element indices fit the fast division
    const float* px;
    const int32_t* winners;
    const float* bounds;
    float eta_c, prob_c, copy_c;
    int mutate;
    float eta_m, rate_m;
    float power_c, exp_m, inv_exp_m;   // 1 / (1 + eta_c), eta_m + 1, 1 / (eta_m + 1)
    const float* sbx_u; const float* sbx_p; const int32_t* sbx_sign; const float* sbx_copy;
    const float* pm_gate; const float* pm_delta;
    uint64_t seed;
    uint32_t gen;
    float* ox;
};

// The launch-wide terms of the SBX and mutation formulas, each one IEEE float
// operation (correctly rounded on the host and on the device alike).
__host__ __device__ __forceinline__ void set_derived(RepArgs& R)
{
    R.power_c = 1.0f / (1.0f + R.eta_c);
    R.exp_m = R.eta_m + 1.0f;
    R.inv_exp_m = 1.0f / R.exp_m;
}

// Parent row crossed into offspring row `row`.  kWinners 0: the winners as
// the mating kernel wrote them; 1: a caller's winners, clamped into [0, n);
// 2: the row itself (mutation alone).
template <int kWinners>
__device__ __forceinline__ int winner_of(const RepArgs& R, int row)
{
    if constexpr (kWinners == 2) return row;
    const int w = R.winners[row];
    if constexpr (kWinners == 1) return min(max(w, 0), R.n - 1);
    return w;
}

// Element e of the (pair, variable) space; the odd tail row (n odd) copies its
// parent (copy_last_parent_kernel, sbx_crossover.cu:72-83) and is then mutated.
// kSmall: every element index is below 2^31 (R.small), so the draw counter's
// high word is the constant 0 and the first Philox rounds are wave-uniform.
template <int kWinners, bool kSmall = false>
__device__ __forceinline__ void sbx_pm_elem(const RepArgs& R, long long e)
{
    const long long total_pair = static_cast<long long>(R.pairs) * R.d;
    const bool tail = (R.n & 1) != 0;
    if (e >= total_pair + (tail ? R.d : 0)) return;
    const uint64_t di = kSmall ? static_cast<uint64_t>(static_cast<uint32_t>(e)) : static_cast<uint64_t>(e);
    if (e < total_pair) {
        int pair, j;
        if (R.small) {
            const uint32_t e32 = static_cast<uint32_t>(e);
            pair = static_cast<int>(fast_div(e32, R.dd));
            j = static_cast<int>(e32 - static_cast<uint32_t>(pair) * static_cast<uint32_t>(R.d));
        } else {
            pair = static_cast<int>(e / R.d);
            j = static_cast<int>(e - static_cast<long long>(pair) * R.d);
        }
        const int w0 = winner_of<kWinners>(R, 2 * pair), w1 = winner_of<kWinners>(R, 2 * pair + 1);
        const float x1 = R.px[static_cast<size_t>(w0) * R.d + j];
        const float x2 = R.px[static_cast<size_t>(w1) * R.d + j];
        const float lo = R.bounds[2 * j], hi = R.bounds[2 * j + 1];
        float ub, up, uc;
        int sg;
        U4 rs{0u, 0u, 0u, 0u};
        const bool draw_s = !R.sbx_u || (R.mutate && !R.pm_gate);
        if (draw_s) rs = draw4(R.seed, di, R.gen, kStSbx);
        if (R.sbx_u) {
            ub = R.sbx_u[e]; up = R.sbx_p[e]; sg = R.sbx_sign[e]; uc = R.sbx_copy[e];
        } else {
            ub = u01(rs.x); up = u01(rs.y); uc = u01(rs.z); sg = sbx_direction(rs);
        }
        float ca, cb;
        sbx_one(x1, x2, ub, up, sg, uc, lo, hi, R.power_c, R.prob_c, R.copy_c, ca, cb);
        if (R.mutate) {
            const size_t ra = static_cast<size_t>(2 * pair) * R.d + j;
            const size_t rb = ra + R.d;
            float ga, da = 0.0f, gb, db = 0.0f;
            if (R.pm_gate) {
                ga = R.pm_gate[ra]; da = R.pm_delta[ra]; gb = R.pm_gate[rb]; db = R.pm_delta[rb];
            } else {
                ga = u01(rs.w);
                gb = gate_b(rs);
                if (ga < R.rate_m || gb < R.rate_m) {
                    const U4 rp = draw4(R.seed, di, R.gen, kStPm);
                    da = u01(rp.x);
                    db = u01(rp.y);
                }
            }
            ca = pm_one(ca, ga, da, lo, hi, R.exp_m, R.inv_exp_m, R.rate_m);
            cb = pm_one(cb, gb, db, lo, hi, R.exp_m, R.inv_exp_m, R.rate_m);
        }
        R.ox[static_cast<size_t>(2 * pair) * R.d + j] = ca;
        R.ox[static_cast<size_t>(2 * pair + 1) * R.d + j] = cb;
    } else {
        const int j = static_cast<int>(e - total_pair);
        const int row = R.n - 1;
        const float lo = R.bounds[2 * j], hi = R.bounds[2 * j + 1];
        float v = R.px[static_cast<size_t>(winner_of<kWinners>(R, row)) * R.d + j];
        if (R.mutate) {
            const size_t ra = static_cast<size_t>(row) * R.d + j;
            float g, dd = 0.0f;
            if (R.pm_gate) {
                g = R.pm_gate[ra]; dd = R.pm_delta[ra];
            } else {
                const U4 rs = draw4(R.seed, di, R.gen, kStSbx);
                g = u01(rs.w);
                if (g < R.rate_m) dd = u01(draw4(R.seed, di, R.gen, kStPm).x);
            }
            v = pm_one(v, g, dd, lo, hi, R.exp_m, R.inv_exp_m, R.rate_m);
        }
        R.ox[static_cast<size_t>(row) * R.d + j] = v;
    }
}

// One thread per (pair, variable).
template <int kWinners, bool kSmall>
__global__ __launch_bounds__(256) void sbx_pm_kernel(RepArgs R)
{
    sbx_pm_elem<kWinners, kSmall>(R, static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// Population::initialize_random (core/population.cu:97-115):
// lower + u * (upper - lower), contracted to one fma by nvcc.
__device__ __forceinline__ void init_elem(float* __restrict__ x, long long e, int d, const float* __restrict__ bounds,
                                 uint64_t seed)
{
    const int j = static_cast<int>(e % d);
    const U4 r = draw4(seed, static_cast<uint64_t>(e), 0u, kStInit);
    const float lo = bounds[2 * j], hi = bounds[2 * j + 1];
    x[e] = fmaf(u01(r.x), hi - lo, lo);
}

__global__ void init_kernel(float* __restrict__ x, long long count, int d,
                            const float* __restrict__ bounds, uint64_t seed)
{
    const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < count) init_elem(x, e, d, bounds, seed);
}

// ---------------------------------------------------------------- batched --
// The same element work per instance of a batch (tournament mating, SBX and
// polynomial mutation with the instance's annealed etas, in-kernel draws).

__host__ __device__ __forceinline__ float annealed_hd(float begin, float end, int generation, int total)
{
    float progress = 1.0f;
    if (total > 0) {
        progress = static_cast<float>(generation) / static_cast<float>(total);
        if (progress < 0.0f) progress = 0.0f;
        if (progress > 1.0f) progress = 1.0f;
    }
    return begin + (end - begin) * progress;
}

__global__ void b_init_kernel(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const long long e = static_cast<long long>(local) * blockDim.x + threadIdx.x;
    if (e < static_cast<long long>(I.n) * I.d) init_elem(I.x[I.cur], e, I.d, I.bounds, I.seed);
}

__global__ void b_mating_kernel(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int t = local * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (t >= I.n) return;
    I.winners[t] = mating_one(t, I.n, I.n, MOEA_MATING_TOURNAMENT, I.fronts[I.fcur], I.cv[I.cur], nullptr,
                              nullptr, I.seed, static_cast<uint32_t>(I.gen));
}

__global__ __launch_bounds__(256) void b_sbx_pm_kernel(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int cur = I.cur, off = (I.cur + 1) % 3, gen = I.gen;
    RepArgs R{};
    R.n = I.n; R.d = I.d; R.pairs = I.n / 2;
    R.dd = I.ddiv;
    R.small = (static_cast<long long>(R.pairs) + 1) * R.d < (1ll << 31);
    R.px = I.x[cur]; R.winners = I.winners; R.bounds = I.bounds;
    R.eta_c = annealed_hd(I.sbx_eta0, I.sbx_eta1, gen, I.max_gen);
    R.prob_c = I.sbx_prob; R.copy_c = I.sbx_copy;
    R.mutate = 1;
    R.eta_m = annealed_hd(I.pm_eta0, I.pm_eta1, gen, I.max_gen);
    R.rate_m = I.pm_prob / static_cast<float>(I.d);
    R.seed = I.seed; R.gen = static_cast<uint32_t>(gen); R.ox = I.x[off];
    set_derived(R);
    sbx_pm_elem<0>(R, static_cast<long long>(local) * blockDim.x + threadIdx.x);
}

}  // namespace

float annealed(float begin, float end, int generation, int total)
{
    return annealed_hd(begin, end, generation, total);
}

int mate(const moea_reproduce_args& A)
{
    if (A.mating != MOEA_MATING_TOURNAMENT && A.mating != MOEA_MATING_RANDOM)
        return fail(kInvalidArgument, "mate: unknown mating kind");
    if (A.mating == MOEA_MATING_TOURNAMENT && (!A.parents_rank || !A.parents_cv))
        return fail(kInvalidArgument, "TournamentMating requires a rank array for every parent");
    const moea_draws* dr = A.draws;
    hipLaunchKernelGGL(mating_kernel, dim3(ceil_div(A.n, 256)), dim3(256), 0, static_cast<hipStream_t>(A.stream),
                       A.n, A.active, A.active_device, A.mating, A.parents_rank, A.parents_cv,
                       dr ? dr->mate_a : nullptr, dr ? dr->mate_b : nullptr, A.seed,
                       static_cast<uint32_t>(A.generation), A.parent_indices);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int mutate_rows(const float* x_in, float* x_out, int n, int d, const float* bounds, float pm_eta, float pm_prob,
                uint64_t seed, int generation, hipStream_t st)
{
    // The element pass of reproduce with every row its own parent and crossover
    // switched off (probability 0 copies), so each row gets the mutation it
    // would get after SBX in the same generation.  x_in may equal x_out: every
    // element is read and written by one thread.
    if (n <= 0 || d <= 0 || !x_in || !x_out || !bounds)
        return fail(kInvalidArgument, "mutate_rows: invalid arguments");
    RepArgs R{};
    R.n = n; R.d = d; R.pairs = n / 2;
    R.dd = fast_div_make(static_cast<uint32_t>(d));
    R.small = (static_cast<long long>(R.pairs) + 1) * d < (1ll << 31);
    R.px = x_in; R.winners = nullptr; R.bounds = bounds;
    R.eta_c = 1.0f; R.prob_c = 0.0f; R.copy_c = 0.0f;
    R.mutate = 1; R.eta_m = pm_eta; R.rate_m = pm_prob / static_cast<float>(d);
    R.seed = seed; R.gen = static_cast<uint32_t>(generation); R.ox = x_out;
    set_derived(R);
    const long long total = static_cast<long long>(R.pairs) * d + ((n & 1) ? d : 0);
    const dim3 grid(static_cast<unsigned>((total + 255) / 256));
    if (R.small) hipLaunchKernelGGL((sbx_pm_kernel<2, true>), grid, dim3(256), 0, st, R);
    else hipLaunchKernelGGL((sbx_pm_kernel<2, false>), grid, dim3(256), 0, st, R);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int reproduce(const moea_reproduce_args& A)
{
    if (A.n <= 0 || A.d <= 0 || A.active <= 0 || A.active > A.n)
        return fail(kInvalidArgument, "reproduce: invalid sizes or active count");
    if (!A.parents_x || !A.bounds || !A.offspring_x || !A.parent_indices)
        return fail(kInvalidArgument, "reproduce: null pointer");
    if (A.mating == MOEA_MATING_TOURNAMENT && (!A.parents_rank || !A.parents_cv))
        return fail(kInvalidArgument, "TournamentMating requires a rank array for every parent");
    if (A.mating != MOEA_MATING_TOURNAMENT && A.mating != MOEA_MATING_RANDOM && A.mating != MOEA_MATING_GIVEN)
        return fail(kInvalidArgument, "reproduce: unknown mating kind");
    hipStream_t st = static_cast<hipStream_t>(A.stream);
    const moea_draws* dr = A.draws;
    const uint32_t gen = static_cast<uint32_t>(A.generation);
    if (A.mating != MOEA_MATING_GIVEN) MOEA_CHECK(mate(A));
    RepArgs R{};
    R.n = A.n; R.d = A.d; R.pairs = A.n / 2;
    R.dd = fast_div_make(static_cast<uint32_t>(A.d));
    R.small = (static_cast<long long>(R.pairs) + 1) * R.d < (1ll << 31);
    R.px = A.parents_x; R.winners = A.parent_indices; R.bounds = A.bounds;
    R.eta_c = A.sbx_eta; R.prob_c = A.sbx_prob; R.copy_c = A.sbx_copy_prob;
    R.mutate = (A.mutation == MOEA_MUTATION_POLYNOMIAL) ? 1 : 0;
    R.eta_m = A.pm_eta; R.rate_m = A.pm_prob / static_cast<float>(A.d);
    set_derived(R);
    if (dr) {
        R.sbx_u = dr->sbx_u; R.sbx_p = dr->sbx_p; R.sbx_sign = dr->sbx_sign; R.sbx_copy = dr->sbx_copy;
        R.pm_gate = dr->pm_gate; R.pm_delta = dr->pm_delta;
        if ((R.sbx_u == nullptr) != (R.sbx_p == nullptr) || (R.sbx_u == nullptr) != (R.sbx_sign == nullptr) ||
            (R.sbx_u == nullptr) != (R.sbx_copy == nullptr) || (R.pm_gate == nullptr) != (R.pm_delta == nullptr))
            return fail(kInvalidArgument, "reproduce: partially injected draw set");
    }
    R.seed = A.seed; R.gen = gen; R.ox = A.offspring_x;
    const long long total = static_cast<long long>(R.pairs) * A.d + ((A.n & 1) ? A.d : 0);
    const dim3 grid(static_cast<unsigned>((total + 255) / 256));
    if (A.mating == MOEA_MATING_GIVEN) {
        if (R.small) hipLaunchKernelGGL((sbx_pm_kernel<1, true>), grid, dim3(256), 0, st, R);
        else hipLaunchKernelGGL((sbx_pm_kernel<1, false>), grid, dim3(256), 0, st, R);
    } else {
        if (R.small) hipLaunchKernelGGL((sbx_pm_kernel<0, true>), grid, dim3(256), 0, st, R);
        else hipLaunchKernelGGL((sbx_pm_kernel<0, false>), grid, dim3(256), 0, st, R);
    }
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int init_population(float* x, int n, int d, const float* bounds, uint64_t seed, hipStream_t stream)
{
    if (!x || !bounds || n <= 0 || d <= 0) return fail(kInvalidArgument, "init_population: invalid arguments");
    const long long count = static_cast<long long>(n) * d;
    hipLaunchKernelGGL(init_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, stream,
                       x, count, d, bounds, seed);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int batch_init_population(const BatchLaunch& L, const BMap& mp, int nblocks)
{
    if (nblocks <= 0) return kOk;
    hipLaunchKernelGGL(b_init_kernel, dim3(nblocks), dim3(256), 0, L.stream, L.insts, mp);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int batch_reproduce(const BatchLaunch& L, const BMap& rows256, int nrows_blocks, const BMap& sbx256,
                    int nsbx_blocks)
{
    if (nrows_blocks <= 0) return kOk;
    hipLaunchKernelGGL(b_mating_kernel, dim3(nrows_blocks), dim3(256), 0, L.stream, L.insts, rows256);
    MOEA_HIP(hipGetLastError());
    hipLaunchKernelGGL(b_sbx_pm_kernel, dim3(nsbx_blocks), dim3(256), 0, L.stream, L.insts, sbx256);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

}  // namespace moea
