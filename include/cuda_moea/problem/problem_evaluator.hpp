// Problem-evaluator interface for native (C++/HIP) problems of the MI355X
// engine.  The class and struct shapes are those of the reference's SDK
// (python/csrc/native/include/cuda_moea/problem/problem_evaluator.cuh:10-31,
// core/types.cuh:31-60, core/contexts.cuh:13-17, core/cuda_context.cuh:22-49)
// with hipStream_t where the reference has cudaStream_t, so a plugin source
// written against that interface compiles here with hipcc.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace cuda_moea {

// A population on the device.  variables (size, dimension) row-major,
// objectives (objective_count, size) objective-major, constraints (size,),
// bounds interleaved [lb0, ub0, lb1, ub1, ...].
struct PopulationView {
    float* variables = nullptr;
    float* objectives = nullptr;
    float* constraints = nullptr;
    const float* bounds = nullptr;
    int size = 0;
    int dimension = 0;
    int objective_count = 0;
};

struct ProblemInfo {
    std::string name;
    int dimension = 0;
    int objective_count = 0;
    int constraint_count = 0;
    std::vector<float> lower_bounds;
    std::vector<float> upper_bounds;
};

// The streams of a run.  Evaluation and execution are the same stream in this
// engine: every launch of a generation is ordered on the caller's stream.
class CudaContext {
public:
    CudaContext(hipStream_t stream, int device, unsigned long long seed)
        : stream_(stream), device_(device), seed_(seed) {}
    hipStream_t evaluation_stream() const noexcept { return stream_; }
    hipStream_t execution_stream() const noexcept { return stream_; }
    int device() const noexcept { return device_; }
    unsigned long long random_seed() const noexcept { return seed_; }

private:
    hipStream_t stream_;
    int device_;
    unsigned long long seed_;
};

struct EvaluationContext {
    int generation = 0;            // -1 for the initial population
    int max_generations = 0;
    CudaContext& cuda;
};

class IProblemEvaluator {
public:
    virtual ~IProblemEvaluator() = default;

    virtual const ProblemInfo& info() const noexcept = 0;
    virtual void initialize(CudaContext&) {}

    // Writes objectives and constraints of every row, enqueued on `stream`.
    virtual void evaluate(PopulationView population, const EvaluationContext& context, hipStream_t stream) = 0;

    // Called before each reproduction; may rewrite the parents' objectives and
    // constraints (enqueued on `stream`).  Returns true when it changed them.
    virtual bool prepare_parent(PopulationView, const EvaluationContext&, hipStream_t) { return false; }

    virtual void reset() {}
};

}  // namespace cuda_moea
