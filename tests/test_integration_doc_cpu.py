"""The Julia ccall binding printed in the integration guide, checked as text against
include/hprqp.h (Julia is not installed here, so the binding is never run):

- CParams, CResult, SaveState, WsInfo, StepParams and KKT list the fields
  of hprqp_params, hprqp_result, hprqp_save_state, hprqp_ws_info,
  hprqp_step_params and hprqp_kkt_out in order, each with a Julia type of the
  C field's size;
- every field the binding assigns on a CParams exists;
- every ccall names an entry point the header declares, with as many
  argument types as the C prototype has parameters, each type compatible
  with the parameter's C type, and as many arguments as types.

The ctypes mirrors in hprqp._lib are themselves checked against the header
by gcc (test_host_cpu.test_struct_layouts_match_the_header), so they serve as
the list of C fields here."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DOC = ROOT / "INTEGRATION.md"
HEADER = ROOT / "include" / "hprqp.h"


def _julia_code():
    blocks = re.findall(r"```julia\n(.*?)```", DOC.read_text(), re.S)
    assert blocks, "INTEGRATION.md has no julia block"
    return "\n".join(blocks)


def _struct_fields(code, name):
    m = re.search(rf"^(?:mutable\s+)?struct\s+{name}\b(.*?)^end$", code, re.S | re.M)
    assert m, f"struct {name} not found"
    body = m.group(1).split("function")[0]
    return re.findall(r"(\w+)::(Float64|Int64|Int32|Ptr\{\w+\})", body)


def _ctypes_kind(t):
    if t is C.c_double:
        return "Float64"
    if t is C.c_int64:
        return "Int64"
    if t is C.c_int32:
        return "Int32"
    return "Ptr"  # c_void_p, POINTER(...), CFUNCTYPE(...)


@pytest.mark.parametrize("jl,cname", [("CParams", "hprqp_params"), ("CResult", "hprqp_result"),
                                      ("SaveState", "hprqp_save_state"), ("WsInfo", "hprqp_ws_info"),
                                      ("StepParams", "hprqp_step_params"), ("KKT", "hprqp_kkt_out")])
def test_julia_structs_mirror_the_header(jl, cname):
    from hprqp import _lib

    cls = getattr(_lib, cname)
    got = [(f, "Ptr" if t.startswith("Ptr") else t) for f, t in _struct_fields(_julia_code(), jl)]
    want = [(f, _ctypes_kind(t)) for f, t in cls._fields_]
    assert got == want


def test_julia_assigns_only_existing_params_fields():
    code = _julia_code()
    fields = {f for f, _ in _struct_fields(code, "CParams")}
    assigned = set(re.findall(r"\bp\.(\w+)\s*=(?!=)", code))
    assert assigned, "the binding sets no parameter"
    assert assigned <= fields, assigned - fields
    # CParams is built by its default constructor, not positionally
    assert re.search(r"\bp\s*=\s*CParams\(\)", code)
    assert not re.search(r"CParams\([^)\s]", code)


def _split_top(s):
    """Split s at commas outside (), {} and []."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _closing(s, i):
    """Index of the parenthesis closing the one at s[i]."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j
    raise AssertionError("unbalanced parentheses")


def _ccalls(code):
    calls = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*lib\)", code):
        close = _closing(code, m.start() + len("ccall"))
        parts = _split_top(code[m.start() + len("ccall("):close])
        name, ret, types = m.group(1), parts[1], parts[2]
        assert types.startswith("(") and types.endswith(")"), (name, types)
        calls.append((name, ret, _split_top(types[1:-1]), parts[3:]))
    return calls


def _prototypes():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(hprqp_\w+)\s*\(([^;{]*?)\)\s*;", text):
        params = [p.strip() for p in m.group(2).split(",")]
        protos[m.group(1)] = [] if params == ["void"] else params
    return protos


def _c_kind(param):
    if "*" in param:
        return "ptr"
    t = param.rsplit(None, 1)[0].replace("const", "").strip()
    return {"int32_t": "i32", "int": "i32", "int64_t": "i64", "double": "f64", "size_t": "usize",
            "uint64_t": "u64", "float": "f32"}[t]


def _jl_kind(t):
    if t.startswith(("Ptr{", "Ref{")) or t == "Cstring":
        return "ptr"
    return {"Int32": "i32", "Cint": "i32", "Int64": "i64", "Float64": "f64", "Cdouble": "f64",
            "Csize_t": "usize", "UInt64": "u64", "Float32": "f32", "Cfloat": "f32"}[t]


def test_julia_ccalls_match_the_prototypes():
    protos = _prototypes()
    calls = _ccalls(_julia_code())
    assert len(calls) >= 10
    for name, _ret, types, args in calls:
        assert name in protos, f"{name} is not declared in include/hprqp.h"
        params = protos[name]
        assert len(types) == len(params), (name, types, params)
        assert [_jl_kind(t) for t in types] == [_c_kind(p) for p in params], (name, types, params)
        assert len(args) == len(types), (name, args)
