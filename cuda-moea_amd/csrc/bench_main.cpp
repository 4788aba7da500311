// moea_bench: the NSGA-III generation benchmark driven through the C ABI only
// (no Python, no torch), for profilers that want a plain native process.
//
//   moea_bench [--n N] [--m M] [--d D] [--gens K] [--warmup W] [--seed S] [--stages]
//              [--copy-mib C]
//
// Workload: NSGA-III with quadratic-center objectives f_k(x) = sum_j (x_j - c_kj)^2,
// x in [0, 1]^D, centers c drawn like torch.rand((M, D)) after manual_seed(seed)
// (MT19937, 24-bit floats), Das-Dennis directions for N, SBX eta 30 p 1,
// polynomial mutation eta 20 p 1, tournament mating -- the cuda_moea.NSGA3
// defaults.  Prints one JSON line with gens/s over the K timed generations.
// --copy-mib C first runs one plain copy kernel over C MiB (a dispatch of known
// traffic, C MiB read and C MiB written, for calibrating HBM byte counters).
#include <hip/hip_runtime.h>

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

}  // namespace

int main(int argc, char** argv)
{
    int n = 32768, m = 3, d = 1024, gens = 50, warmup = 5;
    unsigned long long seed = 0;
    bool stages = false;
    int copy_mib = 0;
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
