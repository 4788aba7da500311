"""Native problem plugins without a GPU: the loader's registration checks
(missing entry symbol, ABI and build-id mismatch read before anything else of
the record, incomplete record, duplicate parameter names), the BiSphere
plugin's schema, and parameter validation at NativeProblem and at algorithm
construction (python/csrc/native_problem.cpp:13-109;
tests/python/test_native_problem.py:153-231 of the reference)."""
import ctypes
import subprocess
from pathlib import Path

import pytest

import cuda_moea as cm
from cuda_moea import _lib
from cuda_moea.native import NativeProblem, native_problem_schema

ROOT = Path(__file__).resolve().parent.parent
PLUGIN = ROOT / "cuda-moea_amd" / "lib" / "plugins" / "libbi_sphere.so"
INCLUDE = ROOT / "include"


@pytest.fixture(scope="module")
def plugin():
    if not PLUGIN.exists():
        subprocess.run(["make", "-s", "-j8", "-C", str(ROOT / "cuda-moea_amd")], check=True)
    return PLUGIN


def _problem(plugin, **kw):
    opts = dict(library=plugin, name="BiSphere", dimension=7, objectives=2, lower_bounds=-5.0, upper_bounds=5.0)
    opts.update(kw)
    return NativeProblem(**opts)


def _compile(tmp_path, name, source, sdk=False):
    """A host-only shared library from `source` (g++; with the SDK headers
    and the HIP headers' host declarations when `sdk`)."""
    src = tmp_path / f"{name}.cpp"
    src.write_text(source)
    out = tmp_path / f"lib{name}.so"
    cmd = ["g++", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(out)]
    if sdk:
        cmd[1:1] = ["-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", "-I/opt/rocm/include"]
    subprocess.run(cmd, check=True)
    return out


def test_schema_of_bi_sphere(plugin):
    s = native_problem_schema(plugin)
    assert s["name"] == "BiSphere"
    assert s["build_id"] == _lib.lib.moea_native_build_id().decode()
    assert s["parameters"] == {
        "offset": {"type": "float", "default": 2.0, "description": "Center of the second sphere"},
        "radius": {"type": "float", "default": 3.0, "description": "Feasible radius around the origin"},
        "constrained": {"type": "bool", "default": False, "description": "Enable sum(x*x) <= radius*radius"},
    }


def test_registration_and_algorithm_construction(plugin):
    p = _problem(plugin, parameters={"offset": 4.0, "constrained": True})
    assert p.schema["parameters"]["radius"]["default"] == 3.0
    assert p._spec()["type"] == "NativeProblem"
    for algorithm in (cm.NSGA3, cm.RVEA):
        inst = algorithm(problem=p, population_size=32, max_generations=2, print_progress=False)
        assert not inst.initialized
        assert inst._native.constraints == 1
        del inst
    assert _problem(plugin)._instantiate().constraints == 0


def test_spec_is_a_copy(plugin):
    p = _problem(plugin, parameters={"offset": 1.0})
    spec = p._spec()
    spec["parameters"]["offset"] = 99
    assert p._spec()["parameters"]["offset"] == 1.0


def test_invalid_parameters_names_and_dimensions(plugin):
    for params in ({"offest": 2}, {"offset": True}, {"offset": float("nan")}, {"constrained": 1}):
        with pytest.raises(ValueError):
            _problem(plugin, parameters=params)
    with pytest.raises(ValueError):
        _problem(plugin, name="unknown")
    with pytest.raises(ValueError):
        _problem(plugin, lower_bounds=5.0)
    with pytest.raises(ValueError, match="objectives=2"):
        cm.NSGA3(problem=_problem(plugin, objectives=3))
    with pytest.raises(ValueError, match="exactly one"):
        NativeProblem(name="BiSphere", dimension=7, objectives=2)
    with pytest.raises(FileNotFoundError):
        _problem(plugin, library=plugin.parent / "absent.so")


def test_loader_revalidates_parameters(plugin):
    """The loader checks the values again against the schema's types, also
    when the Python-side spec was changed after its checks."""
    p = _problem(plugin)
    p._config["parameters"] = {"offset": True}
    with pytest.raises(ValueError, match="Invalid type/value"):
        cm.NSGA3(problem=p)
    p._config["parameters"] = {"radius": -1.0}            # the factory's own check
    with pytest.raises(ValueError, match="construction failed"):
        cm.NSGA3(problem=p)
    p._config["parameters"] = {"offset": 3}               # an int is accepted for a float
    assert cm.NSGA3(problem=p)._native.constraints == 0
    p._config["parameters"] = {"nope": 1.0}
    with pytest.raises(ValueError, match="Unknown native problem parameter"):
        cm.NSGA3(problem=p)


def test_missing_entry_symbol(tmp_path):
    lib = _compile(tmp_path, "nosym", 'extern "C" int something_else() { return 0; }')
    with pytest.raises(ValueError, match="Missing cuda_moea_problem_v1"):
        native_problem_schema(lib)


def test_unloadable_library(tmp_path):
    bad = tmp_path / "notalib.so"
    bad.write_bytes(b"not an ELF file")
    with pytest.raises(ValueError, match="Cannot load native problem"):
        native_problem_schema(bad)


@pytest.mark.parametrize("abi,build_id", [(1, '"wrong"'), (2, None), (1, "nullptr")])
def test_sdk_mismatch_rejected_before_the_record_is_read(tmp_path, abi, build_id):
    """A record whose abi or build id differs is refused after reading only
    those two leading fields: this stub has nothing past them."""
    bid = build_id if build_id is not None else '"' + _lib.lib.moea_native_build_id().decode() + '"'
    src = ('extern "C" { struct Stub { unsigned abi; const char* id; }; '
           f'const Stub* cuda_moea_problem_v1() {{ static Stub p{{{abi}u, {bid}}}; return &p; }} }}')
    lib = _compile(tmp_path, f"mismatch{abi}", src)
    with pytest.raises(ValueError, match="SDK mismatch"):
        native_problem_schema(lib)


_SDK_STUB = r'''
#include "cuda_moea/native/plugin.hpp"
#include <stdexcept>
struct Stub final : cuda_moea::IProblemEvaluator {
    cuda_moea::ProblemInfo i;
    explicit Stub(const cuda_moea::native::ProblemConfig& c) { i.dimension = c.dimension; }
    const cuda_moea::ProblemInfo& info() const noexcept override { return i; }
    void evaluate(cuda_moea::PopulationView, const cuda_moea::EvaluationContext&, hipStream_t) override {}
};
static const std::vector<cuda_moea::native::ParameterSpec> kSchema{SCHEMA};
extern "C" __attribute__((visibility("default"))) const cuda_moea::native::Plugin* cuda_moea_problem_v1()
{
    static const cuda_moea::native::Plugin rec{1u, CUDA_MOEA_NATIVE_BUILD_ID, NAME, &kSchema, CREATE,
        [](cuda_moea::IProblemEvaluator* p) noexcept { delete p; }};
    return &rec;
}
'''
_CREATE = "[](const cuda_moea::native::ProblemConfig& c) -> cuda_moea::IProblemEvaluator* { return new Stub(c); }"


def _sdk_stub(tmp_path, name, schema, reg_name='"Stub"', create=_CREATE):
    src = _SDK_STUB.replace("SCHEMA", schema).replace("NAME", reg_name).replace("CREATE", create)
    return _compile(tmp_path, name, src, sdk=True)


def test_duplicate_parameter_name(tmp_path):
    lib = _sdk_stub(tmp_path, "dup", '{"a", 1.0, ""}, {"b", true, ""}, {"a", 2.0, ""}')
    with pytest.raises(ValueError, match="duplicate native parameter"):
        native_problem_schema(lib)


def test_incomplete_registration(tmp_path):
    lib = _sdk_stub(tmp_path, "noname", '{"a", 1.0, ""}', reg_name="nullptr")
    with pytest.raises(ValueError, match="Incomplete native problem registration"):
        native_problem_schema(lib)
    lib = _sdk_stub(tmp_path, "nocreate", '{"a", 1.0, ""}', create="nullptr")
    with pytest.raises(ValueError, match="Incomplete native problem registration"):
        native_problem_schema(lib)


def test_sdk_stub_schema_all_types_and_info_mismatch(tmp_path):
    """Every parameter type's default comes back through the C schema; a
    problem whose ProblemInfo disagrees with the request is refused."""
    lib = _sdk_stub(tmp_path, "types", '{"i", std::int64_t(-7), "an int"}, {"f", 0.25, ""}, {"b", true, ""}, '
                                       '{"s", std::string("abc"), ""}, {"l", std::vector<double>{1.5, -2.0}, ""}',
                    reg_name='"Types"')
    s = native_problem_schema(lib)
    assert {k: (v["type"], v["default"]) for k, v in s["parameters"].items()} == {
        "i": ("int", -7), "f": ("float", 0.25), "b": ("bool", True), "s": ("str", "abc"),
        "l": ("float_list", [1.5, -2.0])}
    assert s["parameters"]["i"]["description"] == "an int"
    p = NativeProblem(library=lib, name="Types", dimension=3, objectives=2,
                      parameters={"i": 5, "f": 1, "b": False, "s": "x", "l": [1, 2.5]})
    with pytest.raises(ValueError, match="ProblemInfo must match"):
        p._instantiate()                      # the stub reports 0 objectives
    for bad in ({"i": 1.5}, {"l": [1.0, float("inf")]}, {"s": 3}, {"b": 0}):
        with pytest.raises(ValueError):
            NativeProblem(library=lib, name="Types", dimension=3, objectives=2, parameters=bad)


def test_c_abi_rejects_null_and_range(plugin):
    h = ctypes.c_void_p()
    assert _lib.lib.moea_plugin_open(None, ctypes.byref(h)) == -1
    assert _lib.lib.moea_plugin_open(str(plugin).encode(), ctypes.byref(h)) == 0
    try:
        assert _lib.lib.moea_plugin_param_count(h) == 3
        v, desc = _lib.ParamValue(), ctypes.c_char_p()
        assert _lib.lib.moea_plugin_param(h, 3, ctypes.byref(v), ctypes.byref(desc)) == -1
        assert _lib.lib.moea_plugin_param(h, -1, ctypes.byref(v), ctypes.byref(desc)) == -1
        lb = (ctypes.c_float * 4)(0, 0, 0, 0)
        ub = (ctypes.c_float * 4)(1, 1, 1, 1)
        out = ctypes.c_void_p()
        assert _lib.lib.moea_plugin_create(h, 0, 2, lb, ub, None, 0, ctypes.byref(out)) == -1
        assert _lib.lib.moea_plugin_create(h, 4, 2, lb, ub, None, 0, ctypes.byref(out)) == 0
        _lib.lib.moea_plugin_problem_destroy(out)
    finally:
        _lib.lib.moea_plugin_close(h)
