// moea_bench: the NSGA-III generation benchmark driven through the C ABI only
// (no Python, no torch), for profilers that want a plain native process.
//
//   moea_bench [--n N] [--m M] [--d D] [--gens K] [--warmup W] [--seed S] [--stages]
//              [--copy-mib C]
//   moea_bench --mix R [--gens K] [--warmup W] [--seed S] [--stages]
//   moea_bench --mlp --n N --m M --d D [--hidden H] [--gens K] [--warmup W]
//
// Workload: NSGA-III with quadratic-center objectives f_k(x) = sum_j (x_j - c_kj)^2,
// x in [0, 1]^D, centers c drawn like torch.rand((M, D)) after manual_seed(seed)
// (MT19937, 24-bit floats), Das-Dennis directions for N, SBX eta 30 p 1,
// polynomial mutation eta 20 p 1, tournament mating -- the cuda_moea.NSGA3
// defaults.  Prints one JSON line with gens/s over the K timed generations.
// --copy-mib C first runs one plain copy kernel over C MiB (a dispatch of known
// traffic, C MiB read and C MiB written, for calibrating HBM byte counters).
//
// --mix R serves R concurrent runs through moea_serve_* instead: N log-uniform
// in [64, 8192] (even), M in {3, 5, 8, 10}, D log-uniform in [10, 4096],
// quadratic-center or seeded-MLP (hidden 256, weights N(0, 1/fan_in) from a host
// normal generator) objectives 50/50, directions sized from N, each run long
// enough not to finish; W warmup steps, then K timed steps (ms per step).
//
// --mlp times the seeded-MLP objective alone (moea_eval_mlp: the FP32 MFMA
// hidden layer and the head) on an (N, D) population: K timed calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/moea_hip.h"

namespace {

#define CHECK_HIP(x)                                                                   \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                    \
            std::exit(2);                                                              \
        }                                                                              \
    } while (0)

#define CHECK_MOEA(x)                                                                  \
    do {                                                                               \
        int s_ = (x);                                                                  \
        if (s_ != 0) {                                                                 \
            std::fprintf(stderr, "libmoea_hip error %d: %s\n", s_, moea_last_error()); \
            std::exit(3);                                                              \
        }                                                                              \
    } while (0)

template <typename T>
T* dalloc(size_t count)
{
    void* p = nullptr;
    CHECK_HIP(hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 1));
    return static_cast<T*>(p);
}

__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t count)
{
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

float* upload(const std::vector<float>& h)
{
    float* d = dalloc<float>(h.size());
    CHECK_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return d;
}

std::string stage_summary(int gens)
{
    std::string out;
    double sms[32];
    long long cnt[32];
    const int ns = moea_profile_read(sms, cnt, 32);
    for (int i = 0; i < ns && i < 32; ++i) {
        if (!cnt[i]) continue;
        char buf[128];
        std::snprintf(buf, sizeof(buf), "%s\"%s\": %.4f", out.empty() ? "" : ", ", moea_profile_stage_name(i),
                      sms[i] / gens);
        out += buf;
    }
    return out;
}

int serve_mix(int runs, int gens, int warmup, unsigned long long seed, bool stages, hipStream_t stream)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<float> Nrm(0.0f, 1.0f);
    const int ms_choice[4] = {3, 5, 8, 10};
    std::vector<moea_batch_problem> probs(runs);
    long long rows = 0, n_sum = 0, mlp = 0;
    for (int i = 0; i < runs; ++i) {
        int n = static_cast<int>(std::lround(std::exp(std::log(64.0) + U(rng) * (std::log(8192.0) - std::log(64.0)))));
        n += n & 1;
        const int m = ms_choice[rng() % 4];
        const int d = static_cast<int>(std::lround(std::exp(std::log(10.0) + U(rng) * (std::log(4096.0) - std::log(10.0)))));
        const bool is_mlp = (rng() & 1) != 0;
        moea_batch_problem& p = probs[i];
        std::memset(&p, 0, sizeof(p));
        p.n = n; p.m = m; p.d = d;
        const int k = moea_das_dennis(n, m, 0, 0, nullptr, nullptr);
        if (k <= 0) { std::fprintf(stderr, "das_dennis failed: %s\n", moea_last_error()); return 3; }
        std::vector<float> dirs(static_cast<size_t>(m) * k), raw(static_cast<size_t>(m) * k);
        CHECK_MOEA(moea_das_dennis(n, m, 0, 0, dirs.data(), raw.data()) < 0 ? -1 : 0);
        p.k = k;
        p.max_generations = warmup + 2 * gens + 10;
        p.seed = seed + i;
        std::vector<float> bounds(2 * static_cast<size_t>(d));
        for (int j = 0; j < d; ++j) { bounds[2 * j] = 0.0f; bounds[2 * j + 1] = 1.0f; }
        p.bounds = upload(bounds);
        p.directions = upload(dirs);
        if (is_mlp) {
            const int hidden = 256;
            std::vector<float> w1(static_cast<size_t>(hidden) * d), w2(static_cast<size_t>(m) * hidden);
            for (auto& w : w1) w = Nrm(rng) / std::sqrt(static_cast<float>(d));
            for (auto& w : w2) w = Nrm(rng) / std::sqrt(static_cast<float>(hidden));
            p.problem = MOEA_PROBLEM_MLP;
            p.w1 = upload(w1);
            p.w2 = upload(w2);
            p.hidden = hidden;
            ++mlp;
        } else {
            std::vector<float> c(static_cast<size_t>(m) * d);
            for (auto& v : c) v = static_cast<float>(U(rng));
            p.problem = MOEA_PROBLEM_QUADRATIC;
            p.centers = upload(c);
        }
        p.sbx_eta[0] = p.sbx_eta[1] = 30.0f;
        p.sbx_prob = 1.0f;
        p.pm_eta[0] = p.pm_eta[1] = 20.0f;
        p.pm_prob = 1.0f;
        rows += 2ll * n;
        n_sum += n;
    }
    moea_serve_config cfg{runs, rows};
    size_t bytes = moea_serve_reserved_bytes(&cfg);
    for (auto& p : probs) bytes += moea_serve_instance_bytes(&p);
    void* ws = nullptr;
    CHECK_HIP(hipMalloc(&ws, bytes));
    moea_serve* srv = nullptr;
    CHECK_MOEA(moea_serve_create(&cfg, ws, bytes, stream, &srv));
    for (auto& p : probs) {
        int slot = -1;
        CHECK_MOEA(moea_serve_admit(srv, &p, &slot));
    }
    int nf = 0;
    for (int i = 0; i < warmup + 1; ++i) CHECK_MOEA(moea_serve_step(srv, nullptr, 0, &nf));
    CHECK_HIP(hipStreamSynchronize(stream));
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_HIP(hipEventRecord(e0, stream));
    for (int i = 0; i < gens; ++i) CHECK_MOEA(moea_serve_step(srv, nullptr, 0, &nf));
    CHECK_HIP(hipEventRecord(e1, stream));
    CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    std::string stage_json;
    if (stages) {
        CHECK_MOEA(moea_profile_enable(1));
        for (int i = 0; i < gens; ++i) CHECK_MOEA(moea_serve_step(srv, nullptr, 0, &nf));
        CHECK_HIP(hipStreamSynchronize(stream));
        stage_json = stage_summary(gens);
        CHECK_MOEA(moea_profile_enable(0));
    }
    std::printf("{\"workload\": \"served NSGA-III mix\", \"runs\": %d, \"mlp_runs\": %lld, \"individuals\": %lld, "
                "\"steps\": %d, \"warmup\": %d, \"ms_per_step\": %.4f, \"workspace_bytes\": %zu%s%s%s}\n",
                runs, mlp, n_sum, gens, warmup, ms / gens, bytes, stages ? ", \"stage_ms_per_step\": {" : "",
                stage_json.c_str(), stages ? "}" : "");
    std::fflush(stdout);
    moea_serve_destroy(srv);
    CHECK_HIP(hipFree(ws));
    return 0;
}

int mlp_only(int n, int m, int d, int hidden, int gens, int warmup, unsigned long long seed, hipStream_t stream)
{
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> Nrm(0.0f, 1.0f);
    std::vector<float> w1(static_cast<size_t>(hidden) * d), w2(static_cast<size_t>(m) * hidden);
    for (auto& w : w1) w = Nrm(rng) / std::sqrt(static_cast<float>(d));
    for (auto& w : w2) w = Nrm(rng) / std::sqrt(static_cast<float>(hidden));
    float* dw1 = upload(w1);
    float* dw2 = upload(w2);
    float* x = dalloc<float>(static_cast<size_t>(n) * d);
    float* f = dalloc<float>(static_cast<size_t>(m) * n);
    float* cv = dalloc<float>(n);
    float* hw = dalloc<float>(static_cast<size_t>(n) * hidden);
    std::vector<float> bounds(2 * static_cast<size_t>(d));
    for (int j = 0; j < d; ++j) { bounds[2 * j] = 0.0f; bounds[2 * j + 1] = 1.0f; }
    float* db = upload(bounds);
    CHECK_MOEA(moea_init_population(x, n, d, db, seed, stream));
    for (int i = 0; i < warmup; ++i) CHECK_MOEA(moea_eval_mlp(x, n, d, dw1, hidden, dw2, m, f, n, cv, hw, stream));
    CHECK_HIP(hipStreamSynchronize(stream));
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_HIP(hipEventRecord(e0, stream));
    for (int i = 0; i < gens; ++i) CHECK_MOEA(moea_eval_mlp(x, n, d, dw1, hidden, dw2, m, f, n, cv, hw, stream));
    CHECK_HIP(hipEventRecord(e1, stream));
    CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    const double flop = 2.0 * n * static_cast<double>(d) * hidden;
    std::printf("{\"workload\": \"seeded MLP objective\", \"N\": %d, \"M\": %d, \"D\": %d, \"hidden\": %d, "
                "\"calls\": %d, \"ms_per_call\": %.4f, \"hidden_layer_flop\": %.4g, \"tflops\": %.2f}\n",
                n, m, d, hidden, gens, ms / gens, flop, flop / (ms / gens * 1e-3) / 1e12);
    std::fflush(stdout);
    CHECK_HIP(hipFree(x));
    CHECK_HIP(hipFree(f));
    CHECK_HIP(hipFree(cv));
    CHECK_HIP(hipFree(hw));
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    int n = 32768, m = 3, d = 1024, gens = 50, warmup = 5;
    unsigned long long seed = 0;
    bool stages = false;
    int copy_mib = 0;
    int mix = 0;
    bool mlp = false;
    int hidden = 256;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", a.c_str()); std::exit(1); }
            return argv[++i];
        };
        if (a == "--n") n = std::atoi(next());
        else if (a == "--m") m = std::atoi(next());
        else if (a == "--d") d = std::atoi(next());
        else if (a == "--gens") gens = std::atoi(next());
        else if (a == "--warmup") warmup = std::atoi(next());
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--stages") stages = true;
        else if (a == "--copy-mib") copy_mib = std::atoi(next());
        else if (a == "--mix") mix = std::atoi(next());
        else if (a == "--mlp") mlp = true;
        else if (a == "--hidden") hidden = std::atoi(next());
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 1; }
    }
    if (n <= 0 || n % 2 || m <= 0 || d <= 0 || gens <= 0 || warmup < 0) {
        std::fprintf(stderr, "invalid sizes\n");
        return 1;
    }
    CHECK_HIP(hipSetDevice(0));
    hipStream_t stream;
    CHECK_HIP(hipStreamCreate(&stream));
    if (copy_mib > 0) {
        const size_t count = static_cast<size_t>(copy_mib) * (1u << 20) / sizeof(float4);
        float4* src = dalloc<float4>(count);
        float4* dst = dalloc<float4>(count);
        CHECK_HIP(hipMemsetAsync(src, 0, count * sizeof(float4), stream));
        hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, stream, src, dst, count);
        CHECK_HIP(hipGetLastError());
        CHECK_HIP(hipStreamSynchronize(stream));
        CHECK_HIP(hipFree(src));
        CHECK_HIP(hipFree(dst));
        std::printf("{\"copy_kernel_mib\": %d}\n", copy_mib);
    }

    if (mix > 0) return serve_mix(mix, gens, warmup, seed, stages, stream);
    if (mlp) {
        if (hidden <= 0 || m > MOEA_MAX_OBJECTIVES) { std::fprintf(stderr, "invalid sizes\n"); return 1; }
        return mlp_only(n, m, d, hidden, gens, warmup, seed, stream);
    }

    // directions (host Das-Dennis, as the Python API does)
    const int k = moea_das_dennis(n, m, 0, 0, nullptr, nullptr);
    if (k <= 0) { std::fprintf(stderr, "das_dennis failed: %s\n", moea_last_error()); return 3; }
    std::vector<float> dirs(static_cast<size_t>(m) * k), raw(static_cast<size_t>(m) * k);
    CHECK_MOEA(moea_das_dennis(n, m, 0, 0, dirs.data(), raw.data()) < 0 ? -1 : 0);

    // centers: torch.rand((m, d)) after manual_seed(seed)
    std::mt19937 mt(static_cast<uint32_t>(seed));
    std::vector<float> centers(static_cast<size_t>(m) * d);
    for (auto& c : centers) c = static_cast<float>(mt() & 0xFFFFFFu) * 5.9604644775390625e-8f;
    std::vector<float> bounds(2 * static_cast<size_t>(d));
    for (int j = 0; j < d; ++j) { bounds[2 * j] = 0.0f; bounds[2 * j + 1] = 1.0f; }

    const int total = warmup + 2 * gens;
    moea_run_args a;
    std::memset(&a, 0, sizeof(a));
    a.algorithm = MOEA_ALGORITHM_NSGA3;
    a.n = n; a.m = m; a.d = d; a.k = k; a.max_generations = total;
    for (int i = 0; i < 3; ++i) {
        a.pops[i].x = dalloc<float>(static_cast<size_t>(n) * d);
        a.pops[i].f = dalloc<float>(static_cast<size_t>(m) * n);
        a.pops[i].cv = dalloc<float>(n);
    }
    a.current = 0;
    a.fronts[0] = dalloc<int32_t>(n);
    a.fronts[1] = dalloc<int32_t>(n);
    a.fronts_current = 0;
    a.parent_indices = dalloc<int32_t>(n);
    float* d_bounds = dalloc<float>(bounds.size());
    CHECK_HIP(hipMemcpy(d_bounds, bounds.data(), bounds.size() * sizeof(float), hipMemcpyHostToDevice));
    a.bounds = d_bounds;
    float* d_dirs = dalloc<float>(dirs.size());
    float* d_dirs0 = dalloc<float>(dirs.size());
    CHECK_HIP(hipMemcpy(d_dirs, dirs.data(), dirs.size() * sizeof(float), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_dirs0, dirs.data(), dirs.size() * sizeof(float), hipMemcpyHostToDevice));
    a.directions = d_dirs;
    a.directions0 = d_dirs0;
    a.gamma = dalloc<float>(k);
    a.ideal = dalloc<float>(m);
    int32_t* d_active = dalloc<int32_t>(1);
    CHECK_HIP(hipMemcpy(d_active, &n, sizeof(int), hipMemcpyHostToDevice));
    a.active_device = d_active;
    a.adapt_scratch = dalloc<float>(m > 16 ? m : 16);
    a.stats = dalloc<int32_t>(8);
    a.stats_init = dalloc<int32_t>(8);
    CHECK_HIP(hipMemset(a.stats, 0, 8 * sizeof(int32_t)));
    CHECK_HIP(hipMemset(a.stats_init, 0, 8 * sizeof(int32_t)));
    a.mating = MOEA_MATING_TOURNAMENT;
    a.mutation = MOEA_MUTATION_POLYNOMIAL;
    a.sbx_eta[0] = a.sbx_eta[1] = 30.0f;
    a.sbx_prob = 1.0f;
    a.sbx_copy_prob = 0.0f;
    a.pm_eta[0] = a.pm_eta[1] = 20.0f;
    a.pm_prob = 1.0f;
    a.seed = seed;
    a.initial_x_given = 0;
    float* d_centers = dalloc<float>(centers.size());
    CHECK_HIP(hipMemcpy(d_centers, centers.data(), centers.size() * sizeof(float), hipMemcpyHostToDevice));
    a.problem.kind = MOEA_PROBLEM_QUADRATIC;
    a.problem.centers = d_centers;
    a.workspace_bytes = moea_run_workspace_bytes(MOEA_ALGORITHM_NSGA3, n, m, k);
    CHECK_HIP(hipMalloc(&a.workspace, a.workspace_bytes));
    a.stream = stream;

    CHECK_MOEA(moea_run_init(&a));
    if (warmup) CHECK_MOEA(moea_run_step(&a, warmup));
    CHECK_HIP(hipStreamSynchronize(stream));
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_HIP(hipEventRecord(e0, stream));
    CHECK_MOEA(moea_run_step(&a, gens));
    CHECK_HIP(hipEventRecord(e1, stream));
    CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));

    std::string stage_json;
    if (stages) {
        CHECK_MOEA(moea_profile_enable(1));
        CHECK_MOEA(moea_run_step(&a, gens));
        CHECK_HIP(hipStreamSynchronize(stream));
        double sms[32];
        long long cnt[32];
        const int ns = moea_profile_read(sms, cnt, 32);
        CHECK_MOEA(moea_profile_enable(0));
        for (int i = 0; i < ns && i < 32; ++i) {
            if (!cnt[i]) continue;
            char buf[128];
            std::snprintf(buf, sizeof(buf), "%s\"%s\": %.4f", stage_json.empty() ? "" : ", ",
                          moea_profile_stage_name(i), sms[i] / gens);
            stage_json += buf;
        }
    }
    int32_t stats[8];
    CHECK_HIP(hipMemcpy(stats, a.stats, sizeof(stats), hipMemcpyDeviceToHost));
    std::printf("{\"workload\": \"NSGA-III quadratic-center\", \"N\": %d, \"M\": %d, \"D\": %d, \"K\": %d, "
                "\"gens\": %d, \"warmup\": %d, \"ms_per_gen\": %.4f, \"gens_per_s\": %.3f, "
                "\"first_center\": %.9g, \"last_front\": %d%s%s%s}\n",
                n, m, d, k, gens, warmup, ms / gens, 1000.0 * gens / ms, centers[0], stats[0],
                stages ? ", \"stage_ms_per_gen\": {" : "", stage_json.c_str(), stages ? "}" : "");
    std::fflush(stdout);

    for (int i = 0; i < 3; ++i) {
        CHECK_HIP(hipFree(a.pops[i].x));
        CHECK_HIP(hipFree(a.pops[i].f));
        CHECK_HIP(hipFree(a.pops[i].cv));
    }
    CHECK_HIP(hipFree(a.workspace));
    CHECK_HIP(hipStreamDestroy(stream));
    return 0;
}
