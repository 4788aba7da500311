// BiSphere: a two-objective native problem registered as a plugin.
//
//   f1(x) = sum_j x_j^2
//   f2(x) = sum_j (x_j - offset)^2
//   cv(x) = max(f1(x) - radius^2, 0)   when constrained, else 0
//
// The problem of the reference's plugin example
// (examples/native/bi_sphere/problem.cu:7-60, registration.cpp:3-10): same
// registered name, parameters, defaults and construction checks.  One
// 64-lane wavefront per row: lanes stride the row's variables and the two
// sums are reduced across the wave with cross-lane shuffles, four rows per
// 256-thread workgroup.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <stdexcept>

#include "cuda_moea/native/plugin.hpp"

namespace {

constexpr int kWave = 64;
constexpr int kRowsPerBlock = 4;

__global__ __launch_bounds__(kWave * kRowsPerBlock) void bi_sphere_rows(const float* __restrict__ x,
                                                                         float* __restrict__ f,
                                                                         float* __restrict__ cv, int n, int d,
                                                                         float offset, float radius_sq,
                                                                         int constrained)
{
    const int row = blockIdx.x * kRowsPerBlock + static_cast<int>(threadIdx.x) / kWave;
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    if (row >= n) return;                       // whole waves leave together
    const float* xr = x + static_cast<size_t>(row) * d;
    float near = 0.0f, far = 0.0f;
    for (int j = lane; j < d; j += kWave) {
        const float v = xr[j];
        const float w = v - offset;
        near = fmaf(v, v, near);
        far = fmaf(w, w, far);
    }
    for (int s = kWave / 2; s > 0; s >>= 1) {
        near += __shfl_xor(near, s, kWave);
        far += __shfl_xor(far, s, kWave);
    }
    if (lane == 0) {
        f[row] = near;
        f[static_cast<size_t>(n) + row] = far;
        cv[row] = constrained ? fmaxf(near - radius_sq, 0.0f) : 0.0f;
    }
}

class BiSphere final : public cuda_moea::IProblemEvaluator {
public:
    explicit BiSphere(const cuda_moea::native::ProblemConfig& c)
        : offset_(static_cast<float>(c.get<double>("offset"))),
          radius_(static_cast<float>(c.get<double>("radius"))),
          constrained_(c.get<bool>("constrained"))
    {
        if (c.objectives != 2) throw std::invalid_argument("BiSphere requires objectives=2");
        if (!std::isfinite(offset_) || !std::isfinite(radius_) || radius_ < 0.0f ||
            radius_ > std::sqrt(std::numeric_limits<float>::max()))
            throw std::invalid_argument("BiSphere requires a finite offset and a nonnegative radius whose square is a finite float");
        info_.name = "BiSphere";
        info_.dimension = c.dimension;
        info_.objective_count = 2;
        info_.constraint_count = constrained_ ? 1 : 0;
        info_.lower_bounds = c.lower_bounds;
        info_.upper_bounds = c.upper_bounds;
    }

    const cuda_moea::ProblemInfo& info() const noexcept override { return info_; }

    void evaluate(cuda_moea::PopulationView p, const cuda_moea::EvaluationContext&, hipStream_t stream) override
    {
        if (p.size <= 0) return;
        const int blocks = (p.size + kRowsPerBlock - 1) / kRowsPerBlock;
        hipLaunchKernelGGL(bi_sphere_rows, dim3(blocks), dim3(kWave * kRowsPerBlock), 0, stream, p.variables,
                           p.objectives, p.constraints, p.size, p.dimension, offset_, radius_ * radius_,
                           constrained_ ? 1 : 0);
    }

private:
    cuda_moea::ProblemInfo info_;
    float offset_, radius_;
    bool constrained_;
};

const std::vector<cuda_moea::native::ParameterSpec> kParameters{
    {"offset", 2.0, "Center of the second sphere"},
    {"radius", 3.0, "Feasible radius around the origin"},
    {"constrained", false, "Enable sum(x*x) <= radius*radius"},
};

}  // namespace

CUDA_MOEA_REGISTER_PROBLEM(BiSphere, "BiSphere", kParameters)
