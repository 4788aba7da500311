// Native problem plugins: opening a plugin library, checking its
// registration, building a problem from a parameter set, and the C entry
// points through which the run loop evaluates it.  Replaces the reference's
// loader (python/csrc/native_problem.cpp:13-170) behind the C ABI; the checks
// and their order follow it (entry symbol, abi, exact build id, complete
// registration, unique parameter names, typed parameters with defaults,
// unknown names rejected, ProblemInfo matching the request).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <set>
#include <string>
#include <vector>

#include "common.hpp"
#include "cuda_moea/native/plugin.hpp"
#include "moea_internal.hpp"

namespace native = cuda_moea::native;

struct moea_plugin {
    void* handle = nullptr;
    const native::Plugin* record = nullptr;
    std::atomic<int> refs{1};           // the open handle plus one per live problem
};

struct moea_plugin_problem {
    moea_plugin* plugin = nullptr;
    cuda_moea::IProblemEvaluator* instance = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    unsigned long long seed = 0;
    const float* bounds = nullptr;      // device, interleaved (set by initialize)
    cuda_moea::CudaContext* cuda = nullptr;
};

namespace moea {
namespace {

void release(moea_plugin* p)
{
    if (p && p->refs.fetch_sub(1) == 1) {
        if (p->handle) dlclose(p->handle);
        delete p;
    }
}

// Parameter type index of the schema's default (variant order: int, float,
// bool, str, float_list); the MOEA_PARAM_* numbering is the same.
int type_of(const native::Parameter& v) { return static_cast<int>(v.index()); }

// native_problem.cpp:41-72: a value must have the type of the schema's
// default; a float accepts an integer; floats must be finite.
int convert(const moea_param_value& in, const native::ParameterSpec& spec, native::Parameter& out)
{
    const std::string bad = std::string("Invalid type/value for native parameter '") + spec.name + "'";
    switch (type_of(spec.default_value)) {
        case MOEA_PARAM_INT:
            if (in.type != MOEA_PARAM_INT) return fail(kInvalidArgument, bad);
            out = static_cast<std::int64_t>(in.int_value);
            return kOk;
        case MOEA_PARAM_FLOAT: {
            double v;
            if (in.type == MOEA_PARAM_FLOAT) v = in.float_value;
            else if (in.type == MOEA_PARAM_INT) v = static_cast<double>(in.int_value);
            else return fail(kInvalidArgument, bad);
            if (!std::isfinite(v)) return fail(kInvalidArgument, bad);
            out = v;
            return kOk;
        }
        case MOEA_PARAM_BOOL:
            if (in.type != MOEA_PARAM_BOOL) return fail(kInvalidArgument, bad);
            out = in.int_value != 0;
            return kOk;
        case MOEA_PARAM_STR:
            if (in.type != MOEA_PARAM_STR || !in.str_value) return fail(kInvalidArgument, bad);
            out = std::string(in.str_value);
            return kOk;
        case MOEA_PARAM_FLOAT_LIST: {
            if (in.type != MOEA_PARAM_FLOAT_LIST || in.list_length < 0 || (in.list_length > 0 && !in.list))
                return fail(kInvalidArgument, bad);
            std::vector<double> v(in.list, in.list + in.list_length);
            for (double x : v)
                if (!std::isfinite(x)) return fail(kInvalidArgument, bad);
            out = std::move(v);
            return kOk;
        }
        default:
            return fail(kInvalidArgument, bad);
    }
}

__global__ void zero_rows(float* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace
}  // namespace moea

using namespace moea;

extern "C" {

const char* moea_native_build_id(void) { return CUDA_MOEA_NATIVE_BUILD_ID; }

int moea_plugin_open(const char* path, moea_plugin** out)
{
    if (!path || !out) return fail(kInvalidArgument, "moea_plugin_open: null argument");
    *out = nullptr;
    void* handle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!handle) {
        const char* why = dlerror();
        return fail(kInvalidArgument, std::string("Cannot load native problem: ") + (why ? why : path));
    }
    auto reject = [&](const std::string& msg) {
        dlclose(handle);
        return fail(kInvalidArgument, msg);
    };
    const auto entry = reinterpret_cast<native::Entry>(dlsym(handle, "cuda_moea_problem_v1"));
    if (!entry) return reject("Missing cuda_moea_problem_v1 registration");
    const native::Plugin* rec = nullptr;
    try {
        rec = entry();
    } catch (...) {
        return reject("cuda_moea_problem_v1 raised an exception");
    }
    // the primitive prefix first: nothing else of the record is read before
    // the build is known to be this one
    if (!rec || rec->abi != 1u || !rec->build_id || std::strcmp(rec->build_id, CUDA_MOEA_NATIVE_BUILD_ID) != 0)
        return reject("Native problem SDK mismatch; rebuild the plugin against this library's SDK (build id " +
                      std::string(CUDA_MOEA_NATIVE_BUILD_ID) + ")");
    if (!rec->name || !rec->parameters || !rec->create || !rec->destroy)
        return reject("Incomplete native problem registration");
    std::set<std::string> names;
    for (const auto& spec : *rec->parameters)
        if (!spec.name || !names.insert(spec.name).second)
            return reject("Invalid or duplicate native parameter name");
    moea_plugin* p = new moea_plugin();
    p->handle = handle;
    p->record = rec;
    *out = p;
    return kOk;
}

void moea_plugin_close(moea_plugin* plugin) { release(plugin); }

const char* moea_plugin_name(const moea_plugin* plugin)
{
    return (plugin && plugin->record) ? plugin->record->name : "";
}

const char* moea_plugin_build_id(const moea_plugin* plugin)
{
    return (plugin && plugin->record) ? plugin->record->build_id : "";
}

int moea_plugin_param_count(const moea_plugin* plugin)
{
    if (!plugin || !plugin->record) return fail(kInvalidArgument, "moea_plugin_param_count: null plugin");
    return static_cast<int>(plugin->record->parameters->size());
}

int moea_plugin_param(const moea_plugin* plugin, int index, moea_param_value* out, const char** description)
{
    if (!plugin || !plugin->record || !out) return fail(kInvalidArgument, "moea_plugin_param: null argument");
    const auto& specs = *plugin->record->parameters;
    if (index < 0 || index >= static_cast<int>(specs.size()))
        return fail(kInvalidArgument, "moea_plugin_param: index out of range");
    const native::ParameterSpec& s = specs[static_cast<size_t>(index)];
    std::memset(out, 0, sizeof(*out));
    out->name = s.name;
    out->type = type_of(s.default_value);
    switch (out->type) {
        case MOEA_PARAM_INT: out->int_value = std::get<std::int64_t>(s.default_value); break;
        case MOEA_PARAM_FLOAT: out->float_value = std::get<double>(s.default_value); break;
        case MOEA_PARAM_BOOL: out->int_value = std::get<bool>(s.default_value) ? 1 : 0; break;
        case MOEA_PARAM_STR: out->str_value = std::get<std::string>(s.default_value).c_str(); break;
        case MOEA_PARAM_FLOAT_LIST: {
            const auto& v = std::get<std::vector<double>>(s.default_value);
            out->list = v.data();
            out->list_length = static_cast<int>(v.size());
            break;
        }
        default: break;
    }
    if (description) *description = s.description ? s.description : "";
    return kOk;
}

int moea_plugin_create(moea_plugin* plugin, int dimension, int objectives, const float* lower_bounds,
                       const float* upper_bounds, const moea_param_value* params, int param_count,
                       moea_plugin_problem** out)
{
    if (!plugin || !plugin->record || !out || !lower_bounds || !upper_bounds || param_count < 0 ||
        (param_count > 0 && !params))
        return fail(kInvalidArgument, "moea_plugin_create: null argument");
    *out = nullptr;
    if (dimension <= 0 || objectives <= 0) return fail(kInvalidArgument, "dimension and objectives must be positive");
    const native::Plugin& rec = *plugin->record;
    native::ProblemConfig config{dimension, objectives, std::vector<float>(lower_bounds, lower_bounds + dimension),
                                 std::vector<float>(upper_bounds, upper_bounds + dimension), {}};
    for (int i = 0; i < param_count; ++i) {
        if (!params[i].name) return fail(kInvalidArgument, "native problem parameter without a name");
        bool known = false;
        for (const auto& spec : *rec.parameters) known = known || std::strcmp(spec.name, params[i].name) == 0;
        if (!known) return fail(kInvalidArgument, std::string("Unknown native problem parameter: ") + params[i].name);
    }
    for (const auto& spec : *rec.parameters) {
        native::Parameter value = spec.default_value;
        for (int i = 0; i < param_count; ++i)
            if (std::strcmp(spec.name, params[i].name) == 0) MOEA_CHECK(convert(params[i], spec, value));
        config.parameters.emplace(spec.name, std::move(value));
    }
    cuda_moea::IProblemEvaluator* inst = nullptr;
    try {
        inst = rec.create(config);
    } catch (const std::exception& e) {
        return fail(kInvalidArgument, std::string("Native problem construction failed: ") + e.what());
    } catch (...) {
        return fail(kHipError, "Native problem construction failed with a nonstandard exception");
    }
    if (!inst) return fail(kHipError, "Native problem factory returned null");
    const cuda_moea::ProblemInfo& info = inst->info();
    if (info.dimension != dimension || info.objective_count != objectives || info.lower_bounds != config.lower_bounds ||
        info.upper_bounds != config.upper_bounds || info.constraint_count < 0) {
        rec.destroy(inst);
        return fail(kInvalidArgument, "Native ProblemInfo must match the requested dimensions and bounds");
    }
    moea_plugin_problem* p = new moea_plugin_problem();
    p->plugin = plugin;
    p->instance = inst;
    plugin->refs.fetch_add(1);
    *out = p;
    return kOk;
}

int moea_plugin_problem_constraints(const moea_plugin_problem* p)
{
    if (!p || !p->instance) return fail(kInvalidArgument, "moea_plugin_problem_constraints: null problem");
    return p->instance->info().constraint_count;
}

int moea_plugin_problem_initialize(moea_plugin_problem* p, void* stream, uint64_t seed, const float* bounds)
{
    if (!p || !p->instance) return fail(kInvalidArgument, "moea_plugin_problem_initialize: null problem");
    MOEA_HIP(hipGetDevice(&p->device));
    p->stream = static_cast<hipStream_t>(stream);
    p->seed = seed;
    p->bounds = bounds;
    delete p->cuda;
    p->cuda = new cuda_moea::CudaContext(p->stream, p->device, p->seed);
    try {
        p->instance->initialize(*p->cuda);
    } catch (const std::exception& e) {
        return fail(kInvalidArgument, std::string("Native problem initialization failed: ") + e.what());
    } catch (...) {
        return fail(kHipError, "Native problem initialization failed with a nonstandard exception");
    }
    return kOk;
}

void moea_plugin_problem_destroy(moea_plugin_problem* p)
{
    if (!p) return;
    // work the instance enqueued must finish before its code is unloaded
    if (p->cuda) (void)hipStreamSynchronize(p->stream);
    if (p->instance) p->plugin->record->destroy(p->instance);
    if (p->cuda) (void)hipStreamSynchronize(p->stream);
    delete p->cuda;
    release(p->plugin);
    delete p;
}

int moea_plugin_evaluate(void* user, const float* x, int n, int d, int m, float* f, float* cv, int generation,
                         int max_generations, void* stream)
{
    moea_plugin_problem* p = static_cast<moea_plugin_problem*>(user);
    if (!p || !p->instance || !p->cuda) return fail(kInvalidArgument, "moea_plugin_evaluate: problem not initialized");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (p->instance->info().constraint_count == 0 && n > 0) {
        // an unconstrained problem need not write its violations
        hipLaunchKernelGGL(zero_rows, dim3(ceil_div(n, 256)), dim3(256), 0, st, cv, n);
        MOEA_HIP(hipGetLastError());
    }
    cuda_moea::PopulationView view{const_cast<float*>(x), f, cv, p->bounds, n, d, m};
    cuda_moea::EvaluationContext ctx{generation, max_generations, *p->cuda};
    try {
        p->instance->evaluate(view, ctx, st);
    } catch (const std::exception& e) {
        return fail(kInvalidArgument, std::string("Native problem evaluation failed: ") + e.what());
    } catch (...) {
        return fail(kHipError, "Native problem evaluation failed with a nonstandard exception");
    }
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int moea_plugin_prepare_parent(void* user, moea_population* parents, int n, int d, int m, int generation,
                               int max_generations, void* stream)
{
    moea_plugin_problem* p = static_cast<moea_plugin_problem*>(user);
    if (!p || !p->instance || !p->cuda || !parents)
        return fail(kInvalidArgument, "moea_plugin_prepare_parent: problem not initialized");
    cuda_moea::PopulationView view{parents->x, parents->f, parents->cv, p->bounds, n, d, m};
    cuda_moea::EvaluationContext ctx{generation, max_generations, *p->cuda};
    bool changed = false;
    try {
        changed = p->instance->prepare_parent(view, ctx, static_cast<hipStream_t>(stream));
    } catch (const std::exception& e) {
        return fail(kInvalidArgument, std::string("Native problem prepare_parent failed: ") + e.what());
    } catch (...) {
        return fail(kHipError, "Native problem prepare_parent failed with a nonstandard exception");
    }
    MOEA_HIP(hipGetLastError());
    return changed ? 1 : 0;
}

int moea_plugin_reset(moea_plugin_problem* p)
{
    if (!p || !p->instance) return fail(kInvalidArgument, "moea_plugin_reset: null problem");
    try {
        p->instance->reset();
    } catch (const std::exception& e) {
        return fail(kInvalidArgument, std::string("Native problem reset failed: ") + e.what());
    }
    return kOk;
}

}  // extern "C"
