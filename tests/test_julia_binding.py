"""The Julia binding under julia/ against the C-ABI header.

Julia is not installed here, so the extension cannot be run; this checks what
can be checked without it: every ccall in julia/ext/*.jl names a function that
include/splinegrids_hip.h declares and the library exports, with as many
arguments as the prototype and each argument type the C type's Julia
counterpart; the descriptor structs have the C field order; every extension
file is listed in the Project.toml fragment.
"""
from __future__ import annotations

import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "splinegrids_hip.h")
EXT = os.path.join(ROOT, "julia", "ext")

C_TO_JULIA = {
    "int32_t": {"Int32"}, "int64_t": {"Int64"}, "int": {"Cint"},
    "void*": {"Ptr{Cvoid}"}, "const void*": {"Ptr{Cvoid}"},
    "const int32_t*": {"Ptr{Int32}"}, "int32_t*": {"Ptr{Int32}", "Ptr{Cvoid}"},
    "int64_t*": {"Ptr{Int64}"}, "double*": {"Ptr{Float64}"},
    "const sg_grid_desc*": {"Ptr{SgGridDesc}"}, "sg_plan*": {"Ptr{Cvoid}"},
    "const sg_plan*": {"Ptr{Cvoid}"}, "sg_plan**": {"Ptr{Ptr{Cvoid}}"},
    "sg_batch*": {"Ptr{Cvoid}"}, "const sg_batch*": {"Ptr{Cvoid}"},
    "sg_batch**": {"Ptr{Ptr{Cvoid}}"}, "void* const*": {"Ptr{Ptr{Cvoid}}"},
    "const void* const*": {"Ptr{Ptr{Cvoid}}"}, "char*": {"Ptr{UInt8}"},
    "const char*": {"Cstring"}, "const sg_refmat*": {"Ptr{SgRefmat}"},
}


def _norm(t: str) -> str:
    return re.sub(r"\s*\*\s*", "*", " ".join(t.split())).strip()


def _prototypes():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    protos = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(sg_\w+)\s*\(([^;{]*?)\)\s*;", text):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
        types = []
        if params and params != "void":
            for prm in params.split(","):
                prm = re.sub(r"\s*\*\s*", "* ", " ".join(prm.split())).strip()
                types.append(_norm(prm.rsplit(" ", 1)[0]))
        protos[name] = (ret, types)
    return protos


def _ccalls(path):
    text = open(path).read()
    out = []
    for m in re.finditer(r"ccall\(FN\[:(\w+)\],\s*(\w+),\s*\(", text):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        inner = text[m.end():i - 1]
        types, cur, depth = [], "", 0
        for ch in inner:
            if ch == "," and depth == 0:
                types.append(cur.strip())
                cur = ""
                continue
            depth += {"{": 1, "}": -1}.get(ch, 0)
            cur += ch
        if cur.strip():
            types.append(cur.strip())
        out.append((m.group(1), m.group(2), types))
    return out


def test_every_ccall_matches_the_header(sg):
    protos = _prototypes()
    assert "sg_evaluate" in protos and "sg_batch_create" in protos
    _lib = sg._lib

    calls = []
    for f in sorted(os.listdir(EXT)):
        calls += _ccalls(os.path.join(EXT, f))
    assert len(calls) >= 12
    for name, ret, types in calls:
        assert name in protos, name
        cret, ctypes_ = protos[name]
        assert name in _lib.EXPORTS, name
        want_ret = "Cstring" if cret.endswith("char*") else "Cint"
        assert ret == want_ret, (name, ret, cret)
        assert len(types) == len(ctypes_), (name, types, ctypes_)
        for jt, ct in zip(types, ctypes_):
            assert jt in {_norm(k): v for k, v in C_TO_JULIA.items()}[ct], (name, jt, ct)


def test_symbols_resolved_at_init_are_exported(sg):
    _lib = sg._lib

    text = open(os.path.join(EXT, "SplineGridsHIPExt.jl")).read()
    block = re.search(r"const SYMBOLS = \((.*?)\)\n", text, re.S).group(1)
    names = re.findall(r":(\w+)", block)
    used = {n for f in os.listdir(EXT) for n, _, _ in _ccalls(os.path.join(EXT, f))}
    assert used <= set(names), used - set(names)
    assert set(names) <= set(_lib.EXPORTS), set(names) - set(_lib.EXPORTS)


def test_descriptor_structs_follow_the_header():
    text = open(os.path.join(EXT, "SplineGridsHIPExt.jl")).read()
    dim = re.search(r"struct SgDimDesc(.*?)end", text, re.S).group(1).split()
    assert dim == ["degree::Int32", "n_basis::Int32", "n_samples::Int32",
                   "max_derivative_order::Int32", "knots_all::Ptr{Float64}",
                   "sample_points::Ptr{Float64}"]
    grid = re.search(r"struct SgGridDesc(.*?)end", text, re.S).group(1).split()
    assert grid == ["n_in::Int32", "n_out::Int32", "is_nurbs::Int32", "device::Int32",
                    "dims::Ptr{SgDimDesc}"]
    h = open(HEADER).read()
    assert re.search(r"int32_t degree;\s*int32_t n_basis;\s*int32_t n_samples;\s*"
                     r"int32_t max_derivative_order;\s*const double\* knots_all;.*?"
                     r"const double\* sample_points;", h, re.S)
    assert re.search(r"int32_t n_in;.*?int32_t n_out;.*?int32_t is_nurbs;.*?"
                     r"int32_t device;.*?const sg_dim_desc\* dims;", h, re.S)
    refmat = re.search(r"struct SgRefmat(.*?)end", text, re.S).group(1).split()
    assert refmat == ["m::Int32", "n::Int32", "nnz::Int32", "row_pointer::Ptr{Int32}",
                      "column_start::Ptr{Int32}", "nzval::Ptr{Cvoid}"]
    assert re.search(r"int32_t m, n, nnz;\s*const int32_t\* row_pointer;\s*"
                     r"const int32_t\* column_start;\s*const void\* nzval;", h, re.S)


def test_project_fragment_lists_every_extension():
    frag = open(os.path.join(ROOT, "julia", "Project.toml.ext")).read()
    for f in os.listdir(EXT):
        mod = os.path.splitext(f)[0]
        assert re.search(rf"^{mod} = ", frag, re.M), mod
        assert f"module {mod}" in open(os.path.join(EXT, f)).read()
