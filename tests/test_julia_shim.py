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


def _code(text):
    """The shim without comments and docstrings."""
    text = re.sub(r'"""[\s\S]*?"""', "", text)
    text = re.sub(r'"[^"\n]*"', '""', text)
    return "\n".join(line.split("#", 1)[0] for line in text.splitlines())


def test_shim_binds_the_reference_module():
    # The reference's module is VolumeFluxes (src/SinFVM.jl:21); every reference
    # test does `using VolumeFluxes`.
    code = _code(open(SHIM).read())
    assert re.search(r"^\s*using VolumeFluxes\s*$", code, flags=re.M)
    assert "SinFVM" not in code
    assert re.search(r"<:\s*VolumeFluxes\.Backend", code)
    for name in ("perform_step!", "simulate_to_time", "current_state", "set_current_state!"):
        assert re.search(rf"function VolumeFluxes\.{re.escape(name)}\(|VolumeFluxes\.{re.escape(name)}\(sim::", code), name


def test_shim_keys_handles_by_identity():
    # sim.t is an MVector mutated every step (simulator.jl:205): a table that
    # hashes it by content loses the handle after the first step and merges
    # simulators at the same t.  The registry must key on identity.
    code = _code(open(SHIM).read())
    assert "WeakKeyDict" not in code
    decl = re.search(r"const HANDLES\s*=\s*(.+)", code).group(1)
    assert "IdDict" in decl or "UInt" in decl, decl
    stores = re.findall(r"HANDLES\[(.+?)\]\s*=", code)
    assert stores and all("objectid" in k for k in stores), stores
    assert re.search(r"get\(HANDLES,\s*objectid\(", code)
    assert re.search(r"finalizer\(.*release!", code)      # the handle dies with sim.t


def test_shim_runs_1d_systems_through_the_embedding():
    code = _code(open(SHIM).read())
    assert "ShallowWaterEquations1D" in code
    assert re.search(r"c\.embedded_1d\s*=\s*one_d\s*\?\s*1\s*:\s*0", code)
    assert "2.0^60" in code                        # dy of the embedding (volumefluxes.model.embed_1d)
    from volumefluxes.model import EMBED_ROWS
    assert re.search(rf"const EMBED_ROWS\s*=\s*{EMBED_ROWS}\b", code)


def test_shim_config_struct_matches_the_c_layout():
    # field names and order of `mutable struct SfvConfig` against ctypes' sfv_config
    text = open(SHIM).read()
    body = re.search(r"mutable struct SfvConfig\n(.*?)\n\s*SfvConfig\(\) = new\(\)", text, flags=re.S).group(1)
    names = []
    for line in body.splitlines():
        for part in line.split(";"):
            m = re.match(r"\s*(\w+)::", part)
            if m:
                names.append(m.group(1))
    assert names == [f[0] for f in C.SfvConfig._fields_]
