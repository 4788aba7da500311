// Internal interfaces between the HIP modules of libmoea_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/moea_hip.h"

namespace moea {

// nds.hip
size_t nds_workspace_bytes(int n);
int nds_max_n();
// max_front receives -2 when the peel gave up waiting at a grid barrier (the
// fronts are then invalid); `error` (optional) is a sticky flag also set then.
// keys_ready: the caller has already written each row's key (nds_key_from)
// and index into the buffers nds_key_buffers reports for this workspace.
int nds_run(const float* f, int ldf, const float* cv, int n, int m, int stop_at, int32_t* fronts,
            int32_t* max_front, void* ws, size_t ws_bytes, hipStream_t stream, uint32_t* error = nullptr,
            bool keys_ready = false);
void nds_key_buffers(void* ws, int n, uint64_t** keys, int32_t** vals);
int nds_run_static(const float* f, int ldf, const float* cv, int n, int m, int stop_at, int32_t* fronts,
                   int32_t* max_front, hipStream_t stream);
int nds_static_max_n();
void set_peel_spin_limit(uint32_t limit);

// nsga3.hip
size_t nsga3_workspace_bytes(int n, int m, int k);
int nsga3_select(const moea_nsga3_select_args& a);
bool niching_valid(const moea_niching& c);
bool cv_quant_valid(const moea_cv_quantization& q);
int quantize_cv(const float* cv, float* out, int n, const moea_cv_quantization& q, hipStream_t stream);
// Cone tables of the pruned association, built once per set of directions.
bool assoc_prunable(int k, int m);
size_t cone_bytes(int m, int k);
void cone_views(void* buf, int m, int k, float** wp, int32_t** perm, float** cones);
int cones_build(const float* w, int k, int m, void* buf, hipStream_t st);
int ideal_point(const float* f, int ldf, int n, int m, float* ideal, hipStream_t stream);
size_t associate_workspace_bytes(int ld, int m, int k);
int associate(const float* fn, int ld, int n, int m, const float* a, const float* w, int k, float* dist,
              int32_t* ref, int method, void* ws, size_t ws_bytes, hipStream_t st);

// reproduce.hip
int reproduce(const moea_reproduce_args& a);
int mate(const moea_reproduce_args& a);          // the mating kernel alone
int mutate_rows(const float* x_in, float* x_out, int n, int d, const float* bounds, float pm_eta, float pm_prob,
                uint64_t seed, int generation, hipStream_t stream);
int init_population(float* x, int n, int d, const float* bounds, uint64_t seed, hipStream_t stream);

// evaluate.hip
int eval_quadratic(const float* x, int n, int d, const float* centers, int m, float* f, int ldf,
                   float* cv, hipStream_t stream);
int eval_dtlz(int kind, const float* x, int n, int d, int m, float* f, int ldf, float* cv, hipStream_t stream);
int eval_mlp(const float* x, int n, int d, const float* w1, int hidden, const float* w2, int m,
             float* f, int ldf, float* cv, float* hwork, hipStream_t stream);

// rvea.hip
size_t rvea_workspace_bytes(int n, int m, int k);
int rvea_select(const moea_rvea_select_args& a);
int rvea_gamma(const float* v, int m, int k, float* gamma, hipStream_t stream);
int rvea_adapt(const float* v0, const float* f, int ldf, const int32_t* n_active, int m, int k,
               float* v_out, float* scratch, hipStream_t stream);

// sort.hip: stable LSD radix sort of (u64 key, i32 value) pairs, n fixed on the host
size_t sort_pairs_temp_bytes(int n, int end_bit = 64);
int sort_pairs(void* temp, size_t temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
               const int32_t* vals_in, int32_t* vals_out, int n, hipStream_t stream,
               int end_bit = 64);

// host helpers
float annealed(float begin, float end, int generation, int total);

// Bump allocator over a caller-provided workspace.
struct Carve {
    char* base;
    size_t used = 0;
    size_t cap;
    Carve(void* p, size_t c) : base(static_cast<char*>(p)), cap(c) {}
    template <typename T>
    T* take(size_t count)
    {
        const size_t bytes = (count * sizeof(T) + 255) / 256 * 256;
        T* out = reinterpret_cast<T*>(base ? base + used : nullptr);
        used += bytes;
        return out;
    }
    void* take_bytes(size_t bytes)
    {
        void* out = base ? base + used : nullptr;
        used += (bytes + 255) / 256 * 256;
        return out;
    }
    bool ok() const { return used <= cap; }
};

}  // namespace moea

namespace moea {
size_t run_workspace_bytes(int algorithm, int n, int m, int k);
int run_init(moea_run_args& a);
int run_step(moea_run_args& a, int generations);
}  // namespace moea

namespace moea {
// profile.cpp: opt-in stage timing
enum Stage : int {
    kStReproduce = 0, kStEvaluate, kStDominance, kStPeel, kStSelectPrep, kStAssociation,
    kStNiching, kStGather, kStRveaSelect, kStRveaAdapt, kStNdsSort, kNumStages
};
void stage_begin(int stage, hipStream_t stream);
void stage_end(int stage, hipStream_t stream);
}  // namespace moea
