"""The Julia binding (gaussmle.jl_amd/julia/GaussMLEHIP.jl) against the C ABI.

Julia cannot run in this image, so the binding is pinned textually: every
`ccall` in it names a function that include/gaussmle.h declares, with the same
number of arguments, Julia argument and return types that match the C types,
and as many argument expressions as declared types.  The `Request` struct
mirrors `gmle_request` field by field, and the PSF-type methods map
GaussMLE.jl's four PSF types (src/psf_models.jl:39, 64, 91, 141) to the enum
gmle_model ids and the 8-float PSF layout the Python binding uses.
"""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "gaussmle.h"
SHIM = ROOT / "gaussmle.jl_amd" / "julia" / "GaussMLEHIP.jl"

# C type (const and parameter names dropped, pointer stars attached) -> Julia spellings
C_TO_JULIA = {
    "int": {"Cint", "Int32"},
    "int64_t": {"Int64"},
    "uint64_t": {"UInt64"},
    "size_t": {"Csize_t", "UInt"},
    "float": {"Cfloat", "Float32"},
    "double": {"Cdouble", "Float64"},
    "void": {"Cvoid", "Nothing"},
    "float*": {"Ptr{Float32}", "Ptr{Cfloat}"},
    "double*": {"Ptr{Float64}", "Ptr{Cdouble}"},
    "int32_t*": {"Ptr{Int32}", "Ref{Int32}"},
    "int8_t*": {"Ptr{Int8}"},
    "int64_t*": {"Ptr{Int64}", "Ref{Int64}"},
    "uint16_t*": {"Ptr{UInt16}"},
    "uint32_t*": {"Ptr{UInt32}"},
    "char*": {"Ptr{UInt8}", "Cstring", "Ptr{Cchar}"},
    "void*": {"Ptr{Cvoid}", "Ptr{Nothing}"},
    "gmle_service*": {"Ptr{Cvoid}"},
    "gmle_service**": {"Ptr{Ptr{Cvoid}}", "Ref{Ptr{Cvoid}}"},
    "gmle_request*": {"Ref{Request}", "Ptr{Request}"},
}


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _c_type(decl: str) -> str:
    """'const float* d_data' -> 'float*'; 'int model' -> 'int'."""
    decl = decl.replace("const ", " ").strip()
    stars = decl.count("*")
    words = decl.replace("*", " ").split()
    return words[0] + "*" * stars


def header_prototypes() -> dict:
    text = _strip_comments(HEADER.read_text())
    protos = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(gmle_\w+)\s*\(([^;{]*?)\)\s*;", text):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        ret = ret.split()[-1] if "*" not in ret else ret
        args = [] if params in ("", "void") else [_c_type(p) for p in params.split(",")]
        protos[name] = (_c_type(m.group(1)), args)
    return protos


def _split_top(s: str) -> list:
    """Split on commas at nesting depth 0 (parentheses and braces)."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        out.append("".join(cur).strip())
    return out


def julia_ccalls() -> list:
    text = "\n".join(line.split("#", 1)[0] for line in SHIM.read_text().splitlines())
    calls = []
    for m in re.finditer(r"\bccall\(", text):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        parts = _split_top(text[m.end():i - 1])
        name = re.match(r"\(:(\w+),", parts[0]).group(1)
        tup = parts[2].strip()
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = _split_top(tup[1:-1])
        calls.append((name, parts[1].strip(), types, parts[3:]))
    return calls


def test_header_parses():
    protos = header_prototypes()
    for name in ("gmle_fit", "gmle_fit_device", "gmle_fit_batch", "gmle_service_submit", "gmle_last_error"):
        assert name in protos
    assert len(protos["gmle_fit"][1]) == 22
    assert protos["gmle_fit"][1][1] == "float*" and protos["gmle_fit"][1][3] == "int64_t"


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 12
    for name, ret, types, args in calls:
        assert name in protos, f"{name} is not declared in include/gaussmle.h"
        cret, cargs = protos[name]
        assert len(types) == len(cargs), f"{name}: {len(types)} Julia types, {len(cargs)} C parameters"
        assert len(args) == len(types), f"{name}: {len(args)} arguments for {len(types)} types"
        assert ret in C_TO_JULIA[cret], f"{name}: returns {cret}, ccall says {ret}"
        for k, (jt, ct) in enumerate(zip(types, cargs)):
            assert jt in C_TO_JULIA[ct], f"{name} argument {k + 1}: C {ct}, Julia {jt}"


def test_request_struct_mirrors_gmle_request():
    text = _strip_comments(HEADER.read_text())
    body = re.search(r"typedef struct gmle_request \{(.*?)\} gmle_request;", text, re.S).group(1)
    cfields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            cfields.append((decl.replace("*", " ").split()[-1], _c_type(decl)))
    jl = SHIM.read_text()
    jbody = re.search(r"struct Request\n(.*?)\nend", jl, re.S).group(1)
    jfields = [tuple(x.strip() for x in line.split("::")) for line in jbody.strip().splitlines()]
    assert [f for f, _ in jfields] == [f for f, _ in cfields]
    for (fj, tj), (fc, tc) in zip(jfields, cfields):
        assert tj in C_TO_JULIA[tc], f"Request.{fj}: C {tc}, Julia {tj}"


def test_psf_type_methods_match_the_python_binding():
    """model_id / psf_block for each of GaussMLE.jl's PSF types: the enum ids
    and the 8-float layout of models.py psf_array()."""
    from gaussmle_amd import models
    jl = SHIM.read_text()
    ids = dict(re.findall(r"model_id\(::_GMLE\.(\w+)\) = Cint\((\d)\)", jl))
    assert ids == {"GaussianXYNB": "0", "GaussianXYNBS": "1", "GaussianXYNBSXSY": "2", "AstigmaticXYZNB": "3"}
    for cls in (models.GaussianXYNB, models.GaussianXYNBS, models.GaussianXYNBSXSY, models.AstigmaticXYZNB):
        assert int(ids[cls.__name__]) == cls.model_id
    blocks = dict(re.findall(r"psf_block\((?:p)?::_GMLE\.(\w+)\) = (.*)", jl))
    assert blocks["GaussianXYNBS"].strip() == "zeros(Float32, 8)"
    assert blocks["GaussianXYNBSXSY"].strip() == "zeros(Float32, 8)"
    xy = re.match(r"Float32\[(.*)\]", blocks["GaussianXYNB"].strip()).group(1)
    assert [t.strip() for t in xy.split(",")] == ["p.σ"] + ["0"] * 7
    ast = re.match(r"Float32\[(.*)\]", blocks["AstigmaticXYZNB"].strip()).group(1)
    fields = [t.strip()[2:] for t in ast.split(",")]
    # the reference struct's field order (src/psf_models.jl:141-154) is the C layout
    assert fields == ["σx₀", "σy₀", "Ax", "Ay", "Bx", "By", "γ", "d"]
    a = models.AstigmaticXYZNB(1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0).psf_array()
    np.testing.assert_array_equal(a, np.arange(1, 9, dtype=np.float32))
    x = models.GaussianXYNB(1.5).psf_array()
    np.testing.assert_array_equal(x, np.array([1.5, 0, 0, 0, 0, 0, 0, 0], np.float32))


def test_integration_snippet_calls_what_exists():
    """The GPU branch the integration notes at the repository root show calls
    hip_fit! with the PSF struct and constraints, a method the shim defines."""
    doc = (ROOT / "INTEGRATION.md").read_text()
    snippet = re.search(r"```julia\n(.*?)```", doc, re.S).group(1)
    assert "GaussMLEHIP.hip_fit!(" in snippet
    call = snippet[snippet.index("GaussMLEHIP.hip_fit!(") + len("GaussMLEHIP.hip_fit!("):]
    depth, i = 1, 0
    while depth:
        depth += {"(": 1, ")": -1}.get(call[i], 0)
        i += 1
    args = _split_top(call[:i - 1])
    jl = SHIM.read_text()
    sig = re.search(r"function hip_fit!\(([^;]*?psf_pixels[^;]*?); kwargs\.\.\.\)", jl, re.S).group(1)
    assert len(args) == len(_split_top(sig)) == 11
    for name in re.findall(r"\b(model_id|psf_block)\(", snippet):
        assert re.search(rf"\b{name}\(::_GMLE\.", jl), name
