"""The Julia HIPBackend shim cannot run here (no julia), so its bindings are
checked as text against the C header: every ccall names a declared and
exported entry point and passes as many argument types as the C declaration
has parameters."""
import os
import re

import pytest

from volumefluxes import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sinfvm_hip.h")
SHIM = os.path.join(ROOT, "volumefluxes.jl_amd", "julia", "HIPBackend.jl")


def header_arity():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(sfv_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        p = params.strip()
        out[name] = 0 if p in ("", "void") else len(p.split(","))
    return out


def shim_ccalls():
    text = open(SHIM).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(sfv_\w+),\s*libsinfvm\),\s*([\w{}]+),\s*\(", text):
        # the argument-type tuple: up to its matching parenthesis
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        types = [t for t in (x.strip() for x in text[m.end():i - 1].split(",")) if t]
        calls.append((m.group(1), len(types)))
    return calls


def test_shim_calls_match_the_header():
    arity = header_arity()
    calls = shim_ccalls()
    assert len(calls) >= 20
    for name, n in calls:
        assert name in arity, f"{name} is not declared in include/sinfvm_hip.h"
        assert n == arity[name], f"{name}: {n} argument types in the shim, {arity[name]} parameters in C"


def test_shim_symbols_are_exported():
    L = C.lib()
    missing = sorted({name for name, _ in shim_ccalls() if not hasattr(L, name)})
    assert not missing, missing
