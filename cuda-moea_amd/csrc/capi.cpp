// extern "C" entry points of libmoea_hip.so (declared in include/moea_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg)
{
    set_last_error(msg);
    return status;
}

int hip_fail(hipError_t err, const char* what)
{
    set_last_error(std::string(what) + ": " + hipGetErrorString(err));
    return kHipError;
}

namespace {

long long comb_ll(int n, int k)
{
    if (k > n) return 0;
    if (k == 0 || k == n) return 1;
    if (k > n - k) k = n - k;
    long long r = 1;
    for (int i = 1; i <= k; ++i) { r *= (n - k + i); r /= i; }
    return r;
}

// One Das-Dennis layer in std::prev_permutation order of a mask with the first
// m-1 bits set (nsga3_reference_generator.cu:28-60): the set-bit positions run
// through all (m-1)-combinations in lexicographic order.
void layer(int h, int m, float* out, int ld, int col0)
{
    const int len = h + m - 1;
    const long long count = comb_ll(len, m - 1);
    std::vector<int> pos(m > 1 ? m - 1 : 1);
    for (int i = 0; i < m - 1; ++i) pos[i] = i;
    for (long long p = 0; p < count; ++p) {
        int prev = -1;
        for (int dim = 0; dim < m; ++dim) {
            const int next = (dim < m - 1) ? pos[dim] : len;
            out[static_cast<size_t>(dim) * ld + col0 + p] =
                static_cast<float>(next - prev - 1) / static_cast<float>(h);
            prev = next;
        }
        int i = m - 2;
        while (i >= 0 && pos[i] == len - (m - 1) + i) --i;
        if (i < 0) break;
        pos[i] += 1;
        for (int j = i + 1; j < m - 1; ++j) pos[j] = pos[j - 1] + 1;
    }
}

}  // namespace

}  // namespace moea

using namespace moea;

extern "C" {

int moea_abi_version(void) { return 4; }

const char* moea_last_error(void) { return g_last_error.c_str(); }

size_t moea_nds_workspace_bytes(int n) { return n > 0 ? nds_workspace_bytes(n) : 0; }

int moea_nds_rank(const float* f, int ldf, const float* cv, int n, int m, int stop_at,
                  int32_t* fronts, int32_t* max_front, void* workspace, size_t workspace_bytes,
                  void* stream)
{
    if (ldf < n) return fail(kInvalidArgument, "moea_nds_rank: ldf < n");
    return nds_run(f, ldf, cv, n, m, stop_at, fronts, max_front, workspace, workspace_bytes,
                   static_cast<hipStream_t>(stream));
}

int moea_nds_rank_static(const float* f, int ldf, const float* cv, int n, int m, int stop_at,
                         int32_t* fronts, int32_t* max_front, void* stream)
{
    if (ldf < n) return fail(kInvalidArgument, "moea_nds_rank_static: ldf < n");
    return nds_run_static(f, ldf, cv, n, m, stop_at, fronts, max_front, static_cast<hipStream_t>(stream));
}

int moea_nds_static_max_n(void) { return nds_static_max_n(); }

void moea_debug_set_spin_limit(uint32_t polls) { set_peel_spin_limit(polls); }

int moea_das_dennis(int requested_count, int m, int partitions, int kind, float* out_host,
                    float* raw_host)
{
    if (m <= 0 || m > kMaxM) return fail(kInvalidArgument, "moea_das_dennis: invalid objective count");
    int h1, h2 = 0, n1, n2 = 0;
    if (kind == 1) {
        if (partitions <= 0) return fail(kInvalidArgument, "Reference partitions and objective count must be positive");
        h1 = partitions;
        n1 = static_cast<int>(comb_ll(h1 + m - 1, m - 1));
    } else if (kind == 0 || kind == 2) {
        h1 = 0;
        while (comb_ll(h1 + m - 1, m - 1) <= requested_count) ++h1;
        --h1;
        if (h1 <= 0) return fail(kInvalidArgument, "moea_das_dennis: requested count too small");
        n1 = static_cast<int>(comb_ll(h1 + m - 1, m - 1));
        if (h1 < m) {
            h2 = 1;
            while (comb_ll(h1 + m - 1, m - 1) + comb_ll(h2 + m - 1, m - 1) <= requested_count) ++h2;
            --h2;
            if (h2 > 0) n2 = static_cast<int>(comb_ll(h2 + m - 1, m - 1));
        }
    } else {
        return fail(kInvalidArgument, "moea_das_dennis: unknown kind");
    }
    const int k = n1 + n2;
    if (!out_host && !raw_host) return k;
    std::vector<float> pts(static_cast<size_t>(m) * k);
    layer(h1, m, pts.data(), k, 0);
    if (n2 > 0) {
        std::vector<float> tmp(static_cast<size_t>(m) * n2);
        layer(h2, m, tmp.data(), n2, 0);
        const float centroid = 1.0f / static_cast<float>(m);
        for (int p = 0; p < n2; ++p)
            for (int dd = 0; dd < m; ++dd)
                pts[static_cast<size_t>(dd) * k + n1 + p] = 0.5f * (tmp[static_cast<size_t>(dd) * n2 + p] + centroid);
    }
    if (kind == 2)
        for (auto& v : pts) if (v < 1e-6f) v = 1e-6f;
    if (raw_host) for (size_t i = 0; i < pts.size(); ++i) raw_host[i] = pts[i];
    if (out_host) {
        for (int p = 0; p < k; ++p) {
            float ss = 0.0f;
            for (int dd = 0; dd < m; ++dd) { const float v = pts[static_cast<size_t>(dd) * k + p]; ss = ss + v * v; }
            if (kind == 1) {
                const float inv = 1.0f / std::sqrt(ss);
                for (int dd = 0; dd < m; ++dd) out_host[static_cast<size_t>(dd) * k + p] = pts[static_cast<size_t>(dd) * k + p] * inv;
            } else {
                const float nrm = std::sqrt(ss);
                for (int dd = 0; dd < m; ++dd) out_host[static_cast<size_t>(dd) * k + p] = pts[static_cast<size_t>(dd) * k + p] / nrm;
            }
        }
    }
    return k;
}

size_t moea_nsga3_workspace_bytes(int n, int m, int k)
{
    if (n <= 0 || m <= 0 || k <= 0) return 0;
    return nsga3_workspace_bytes(n, m, k);
}

int moea_quantize_cv(const float* cv, float* out, int n, const moea_cv_quantization* q, void* stream)
{
    if (!q) return fail(kInvalidArgument, "moea_quantize_cv: null parameters");
    return quantize_cv(cv, out, n, *q, static_cast<hipStream_t>(stream));
}

int moea_nsga3_select(const moea_nsga3_select_args* args)
{
    if (!args) return fail(kInvalidArgument, "moea_nsga3_select: null args");
    return nsga3_select(*args);
}

size_t moea_associate_workspace_bytes(int ld, int m, int k) { return associate_workspace_bytes(ld, m, k); }

int moea_associate(const float* fn, int ld, int n, int m, const float* a, const float* directions, int k,
                   float* dist, int32_t* ref, int method, void* workspace, size_t workspace_bytes, void* stream)
{
    return associate(fn, ld, n, m, a, directions, k, dist, ref, method, workspace, workspace_bytes,
                     static_cast<hipStream_t>(stream));
}

int moea_ideal_point(const float* f, int ldf, int n, int m, float* ideal, void* stream)
{
    return ideal_point(f, ldf, n, m, ideal, static_cast<hipStream_t>(stream));
}

int moea_reproduce(const moea_reproduce_args* args)
{
    if (!args) return fail(kInvalidArgument, "moea_reproduce: null args");
    return reproduce(*args);
}

float moea_annealed(float begin, float end, int generation, int total)
{
    return annealed(begin, end, generation, total);
}

int moea_init_population(float* x, int n, int d, const float* bounds, uint64_t seed, void* stream)
{
    return init_population(x, n, d, bounds, seed, static_cast<hipStream_t>(stream));
}

int moea_eval_quadratic(const float* x, int n, int d, const float* centers, int m, float* f, int ldf,
                        float* cv, void* stream)
{
    return eval_quadratic(x, n, d, centers, m, f, ldf, cv, static_cast<hipStream_t>(stream));
}

int moea_eval_dtlz(int problem, const float* x, int n, int d, int m, float* f, int ldf, float* cv,
                   void* stream)
{
    return eval_dtlz(problem, x, n, d, m, f, ldf, cv, static_cast<hipStream_t>(stream));
}

int moea_eval_mlp(const float* x, int n, int d, const float* w1, int hidden, const float* w2, int m,
                  float* f, int ldf, float* cv, float* hwork, void* stream)
{
    return eval_mlp(x, n, d, w1, hidden, w2, m, f, ldf, cv, hwork, static_cast<hipStream_t>(stream));
}

size_t moea_rvea_workspace_bytes(int n, int m, int k)
{
    if (n <= 0 || m <= 0 || k <= 0) return 0;
    return rvea_workspace_bytes(n, m, k);
}

int moea_rvea_select(const moea_rvea_select_args* args)
{
    if (!args) return fail(kInvalidArgument, "moea_rvea_select: null args");
    return rvea_select(*args);
}

float moea_rvea_penalty(int m, int t, int max_generations, float alpha)
{
    float total = static_cast<float>(max_generations);
    if (total < 1.0f) total = 1.0f;
    float progress = static_cast<float>(t) / total;
    if (progress < 0.0f) progress = 0.0f;
    if (progress > 1.0f) progress = 1.0f;
    return static_cast<float>(m) * static_cast<float>(std::pow(static_cast<double>(progress), static_cast<double>(alpha)));
}

int moea_rvea_gamma(const float* directions, int m, int k, float* gamma, void* stream)
{
    return rvea_gamma(directions, m, k, gamma, static_cast<hipStream_t>(stream));
}

int moea_rvea_adapt(const float* v0, const float* f, int ldf, const int32_t* n_active, int m, int k,
                    float* v_out, float* scratch, void* stream)
{
    return rvea_adapt(v0, f, ldf, n_active, m, k, v_out, scratch, static_cast<hipStream_t>(stream));
}

}  // extern "C"
