// Native problem plugins: a shared library built with hipcc registers one
// IProblemEvaluator type under the entry symbol cuda_moea_problem_v1, with a
// parameter schema and factory functions.  The registration record has the
// reference's layout (python/csrc/native/include/cuda_moea/native/plugin.cuh:
// 14-58): the loader reads its leading primitive fields (abi, build_id)
// before it touches any C++ member, and accepts only abi == 1 and an exact
// build_id match.  This is an exact-build extension interface, not a
// cross-version stable ABI.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <variant>
#include <vector>

#include "cuda_moea/native/build_id.h"
#include "cuda_moea/problem/problem_evaluator.hpp"

namespace cuda_moea::native {

// Parameter types, in this order: int, float, bool, str, float_list.
using Parameter = std::variant<std::int64_t, double, bool, std::string, std::vector<double>>;

struct ParameterSpec {
    const char* name;
    Parameter default_value;
    const char* description;
};

struct ProblemConfig {
    int dimension;
    int objectives;
    std::vector<float> lower_bounds;
    std::vector<float> upper_bounds;
    std::unordered_map<std::string, Parameter> parameters;

    template <class T>
    const T& get(const std::string& name) const
    {
        return std::get<T>(parameters.at(name));
    }
};

struct Plugin {
    std::uint32_t abi;
    const char* build_id;
    const char* name;
    const std::vector<ParameterSpec>* parameters;
    IProblemEvaluator* (*create)(const ProblemConfig&);
    void (*destroy)(IProblemEvaluator*) noexcept;
};
using Entry = const Plugin* (*)();

}  // namespace cuda_moea::native

// In exactly one translation unit of the plugin: Type is constructed from a
// ProblemConfig, Name is the registered problem name, Schema a
// std::vector<ParameterSpec> with static storage.
#define CUDA_MOEA_REGISTER_PROBLEM(Type, Name, Schema)                                        \
    extern "C" __attribute__((visibility("default"))) const ::cuda_moea::native::Plugin*     \
    cuda_moea_problem_v1()                                                                   \
    {                                                                                        \
        static const ::cuda_moea::native::Plugin record{                                     \
            1u, CUDA_MOEA_NATIVE_BUILD_ID, Name, &(Schema),                                  \
            [](const ::cuda_moea::native::ProblemConfig& c) -> ::cuda_moea::IProblemEvaluator* { \
                return new Type(c);                                                          \
            },                                                                               \
            [](::cuda_moea::IProblemEvaluator* p) noexcept { delete p; }};                   \
        return &record;                                                                      \
    }
