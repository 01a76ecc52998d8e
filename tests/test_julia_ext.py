"""The Julia package extension (pmf-glm_amd/ext/MHDKernelsHIPExt.jl) cannot run here
(no Julia), so its binding to the C ABI is checked statically: every ccall names a
function that include/pmfmhd.h declares, with the return type and one argument type per
parameter that the prototype fixes; its PmfParams mirrors pmf_params field for field; its
flag and integrator constants equal the header's; and it specialises every seam the
reference's production driver reaches (integrator_cube.jl:467-540,
integrator_ref.jl:210-258, timestep.jl:17-23)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pmfmhd.h")
EXT = os.path.join(ROOT, "pmf-glm_amd", "ext", "MHDKernelsHIPExt.jl")

# C parameter type (names removed, spaces normalised) -> Julia types that pass it
C_TO_JULIA = {
    "pmf_state*": {"Ptr{Cvoid}"},
    "pmf_surrogate*": {"Ptr{Cvoid}"},
    "pmf_ensemble*": {"Ptr{Cvoid}"},
    "void*": {"Ptr{Cvoid}"},
    "const void*": {"Ptr{Cvoid}"},
    "const void* const*": {"Ptr{Ptr{Cvoid}}"},
    "void* const*": {"Ptr{Ptr{Cvoid}}"},
    "pmf_state**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "void**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "pmf_surrogate**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "pmf_ensemble**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "const int32_t*": {"Ptr{Int32}", "Ref{Int32}"},
    "const char*": {"Cstring", "Ptr{UInt8}"},
    "const pmf_params*": {"Ref{PmfParams}", "Ptr{PmfParams}"},
    "pmf_params*": {"Ref{PmfParams}", "Ptr{PmfParams}"},
    "double": {"Cdouble", "Float64"},
    "int32_t": {"Int32", "Cint"},
    "int64_t": {"Int64"},
    "uint64_t": {"UInt64"},
    "double*": {"Ptr{Cdouble}", "Ref{Cdouble}", "Ref{Float64}", "Ptr{Float64}"},
    "const double*": {"Ptr{Cdouble}", "Ref{Cdouble}", "Ptr{Float64}"},
    "int64_t*": {"Ptr{Int64}", "Ref{Int64}"},
    "int32_t*": {"Ptr{Int32}", "Ref{Int32}"},
    "float*": {"Ptr{Cfloat}", "Ptr{Float32}"},
    "const float*": {"Ptr{Cfloat}", "Ptr{Float32}"},
}
RET_TO_JULIA = {"int": {"Cint", "Int32"}, "const char*": {"Cstring", "Ptr{UInt8}"}}


def _norm(t):
    t = re.sub(r"\s*\*", "*", t.strip())
    t = re.sub(r"\*(?=\w)", "* ", t)
    return re.sub(r"\s+", " ", t)


def header_prototypes():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    src = re.sub(r"//[^\n]*", " ", src)
    out = {}
    for ret, name, args in re.findall(r"([A-Za-z_][\w\s\*]*?)\s*\b(pmf_\w+)\s*\(([^()]*)\)\s*;", src):
        params = []
        a = args.strip()
        if a and a != "void":
            for p in a.split(","):
                p = p.strip()
                m = re.match(r"(.*?)(\w+)\s*\[\s*\d*\s*\]$", p)  # array parameter -> pointer
                if m:
                    params.append(_norm(m.group(1) + "*"))
                    continue
                m = re.match(r"(.*[\s\*])(\w+)$", p)
                params.append(_norm(m.group(1)))
        out[name] = (_norm(ret), params)
    return out


def ext_ccalls():
    src = open(EXT).read()
    src = re.sub(r"#[^\n]*", " ", src)
    calls = re.findall(r"ccall\(\(:(\w+),\s*LIB\),\s*([\w{}]+),\s*\(([^()]*)\)", src)
    out = []
    for name, ret, args in calls:
        types = [t.strip() for t in args.split(",") if t.strip()]
        out.append((name, ret, types))
    return out


def test_header_parses():
    protos = header_prototypes()
    assert protos["pmf_step_into"] == ("int", ["pmf_state*", "const void* const*", "void* const*", "void*",
                                               "double", "double", "double", "int32_t"])
    assert protos["pmf_conserved_totals_fp64"][1] == ["pmf_state*", "double*"]
    assert protos["pmf_last_error"] == ("const char*", [])
    for _, params in protos.values():
        for p in params:
            assert p in C_TO_JULIA, p


def test_every_ccall_matches_its_prototype():
    protos = header_prototypes()
    calls = ext_ccalls()
    assert len(calls) >= 8
    for name, ret, types in calls:
        assert name in protos, f"{name} is not declared in include/pmfmhd.h"
        cret, cparams = protos[name]
        assert ret in RET_TO_JULIA[cret], (name, ret, cret)
        assert len(types) == len(cparams), (name, types, cparams)
        for k, (jt, ct) in enumerate(zip(types, cparams)):
            assert jt in C_TO_JULIA[ct], (name, k, jt, ct)


def test_params_struct_and_constants_mirror_the_header():
    h = open(HEADER).read()
    body = re.search(r"typedef struct pmf_params \{(.*?)\} pmf_params;", h, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", " ", body, flags=re.S)
    cfields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, names = decl.split(None, 1)
        for n in names.split(","):
            n = n.strip()
            m = re.match(r"(\w+)\[(\d+)\]", n)
            jt = {"int32_t": "Int32", "double": "Float64"}[ctype]
            cfields.append((m.group(1), f"NTuple{{{m.group(2)},{jt}}}") if m else (n, jt))
    jl = open(EXT).read()
    jbody = re.search(r"struct PmfParams\n(.*?)\nend", jl, re.S).group(1)
    jfields = [tuple(x.strip().split("::")) for x in jbody.strip().splitlines()]
    assert jfields == cfields
    enums = dict((k, int(v)) for k, v in re.findall(r"(PMF_\w+)\s*=\s*(\d+)", h))
    for k, v in re.findall(r"const (\w+) = Int32\((\d+)\)", jl):
        assert enums["PMF_" + k] == int(v), k


@pytest.mark.parametrize("method", ["_cube_launch!", "_cube_drag_launch!", "_cube_radiation_launch!",
                                    "max_wavespeed!", "step_ref!", "step_ref_drag!", "step_ref_radiation!",
                                    "device_zeros"])
def test_extension_specialises_the_production_seams(method):
    jl = open(EXT).read()
    assert re.search(r"MHDKernels\." + re.escape(method) + r"\(", jl), method
