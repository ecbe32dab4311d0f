"""Per-phase VALU instruction budget of the march kernels' row loop.

Reads the AMD assembly of this build with line tables (the `isa-budget`
target of the Makefile: hipcc -gline-tables-only --cuda-device-only -S of
csrc/swe_kernels_march.hip, same flags as the library), finds the row loops
of k_march<double, 1, true> and k_march<double, 2, true> (the kernels of
the f64 terrain workload), and attributes every instruction of each loop to
the source function or row phase its line-table entry names.  The counts
are static: one loop iteration is one row of a wave, and both sides of a
branch are counted (the general face path, for example, runs only in waves
whose inputs left the ranged band).

usage: python3 isa_budget.py march.s [csrc_dir]
"""
from __future__ import annotations

import collections
import os
import re
import sys

KERNELS = {
    "stage 1": "_ZN3sfv7k_marchIdLi1ELb1E",
    "stage 2": "_ZN3sfv7k_marchIdLi2ELb1E",
}

# source function -> phase
PHASE = {
    "reconstruct_cell": "reconstruction (minmod slopes, face values)",
    "reconstruct_cell_nb": "reconstruction (minmod slopes, face values)",
    "minmod": "reconstruction (minmod slopes, face values)",
    "central_upwind_rng": "faces: central-upwind flux, ranged quotients",
    "phys_flux_rng": "faces: central-upwind flux, ranged quotients",
    "recip_rng": "faces: central-upwind flux, ranged quotients",
    "div_rng": "faces: central-upwind flux, ranged quotients",
    "sqrt_rng": "faces: central-upwind flux, ranged quotients",
    "central_upwind": "faces: general path (waves out of band)",
    "phys_flux": "faces: general path (waves out of band)",
    "desingularize": "desingularisation (faces and friction)",
    "friction_scalar": "friction (scalar and implicit update)",
    "friction_law": "friction (scalar and implicit update)",
    "apply_friction": "friction (scalar and implicit update)",
    "stage1_cell": "U1 rebuild: RK stage 1 and dry clamp",
    "in_band": "band checks and votes",
    "in_band5": "band checks and votes",
    "set_issue_priority": "issue priority (rows left)",
    "dpp_shift": "lane shifts (DPP)",
    "shfl_dn1": "lane shifts (DPP)",
    "shfl_up1": "lane shifts (DPP)",
    "next_col": "lane shifts (DPP)",
    "row_rsrc": "row loads and stores (descriptors, offsets)",
    "rsrc_ld": "row loads and stores (descriptors, offsets)",
    "row_ld": "row loads and stores (descriptors, offsets)",
    "row_st": "row loads and stores (descriptors, offsets)",
    "Bload_row": "row loads and stores (descriptors, offsets)",
    "fetch": "row loads and stores (descriptors, offsets)",
    "fetch_b": "row loads and stores (descriptors, offsets)",
    "jl_max": "Julia max/min (NaN-propagating selects)",
    "jl_min": "Julia max/min (NaN-propagating selects)",
    "nz_max": "Julia max/min (NaN-propagating selects)",
    "nz_min": "Julia max/min (NaN-propagating selects)",
    "sdiv": "x flux difference and bottom source",
    "bc4": "cell bottoms",
    "bcell4": "cell bottoms",
    "face_y": "faces: dispatch and wave-speed maxima",
    "face_x": "faces: dispatch and wave-speed maxima",
    "sources": "sources (rain)",
    "apply_sources": "sources (rain)",
}
# row phases inside march_tile's `body` lambda, by the comment that opens them
BODY_MARKERS = [
    ("// 1. y reconstruction", "row glue: reconstruction inputs and face bottoms"),
    ("// 2. the faces", "faces: dispatch and wave-speed maxima"),
    ("// 3. x part", "x flux difference and bottom source"),
    ("// 4. finish row", "y flux difference, RK2 combination, output clamp"),
    ("// shift the windows", "window shifts (register moves)"),
]
DEF_RE = re.compile(r"(?:__device__[^(;]*?\b(\w+)\s*\(|auto\s+(\w+)\s*=\s*\[)")


def definitions(path):
    """(start, end, name) of every function and lambda defined in a source file."""
    lines = open(path).read().split("\n")
    out = []
    for i, line in enumerate(lines):
        m = DEF_RE.search(line)
        if not m:
            continue
        name = m.group(1) or m.group(2)
        depth, opened, end = 0, False, None
        for j in range(i, len(lines)):
            code = lines[j].split("//")[0]
            for ch in code:
                if ch == "{":
                    depth += 1
                    opened = True
                elif ch == "}":
                    depth -= 1
            if opened and depth <= 0:
                end = j
                break
            if not opened and ";" in code:
                break
        if end is not None:
            out.append((i + 1, end + 1, name))
    return out, lines


def phase_of(defs, lines, line):
    if line == 0:
        return "no source line (compiler-made moves and merges)"
    inner = None
    for s, e, name in defs:
        if s <= line <= e and (inner is None or e - s < inner[1] - inner[0]):
            inner = (s, e, name)
    if inner is None:
        return "other"
    name = inner[2]
    if name == "body":
        label = "row glue: loop control"
        for k in range(inner[0], line + 1):
            for marker, lab in BODY_MARKERS:
                if marker in lines[k - 1]:
                    label = lab
        return label
    if name in PHASE:
        return PHASE[name]
    if name == "march_tile":
        return "march_tile setup and loop control"
    return name


# functions whose phase absorbs the helpers they call (desingularisation,
# square roots, quotients, NaN-propagating selects): a helper's instructions
# count under the innermost of these on its inlining chain
MAJOR = {"reconstruct_cell", "reconstruct_cell_nb", "central_upwind_rng", "phys_flux_rng", "central_upwind",
         "phys_flux", "friction_scalar", "friction_law", "apply_friction", "stage1_cell"}
FRAME_RE = re.compile(r"([\w./-]+\.(?:hpp|hip|h)):(\d+):\d+")


def frames_of(line):
    """The inlining chain in a .loc comment, innermost first: [(file base name, line)]."""
    if ";" not in line:
        return []
    return [(os.path.basename(f), int(n)) for f, n in FRAME_RE.findall(line.split(";", 1)[1])]


def main(asm_path, csrc):
    text = open(asm_path).read().split("\n")
    files = {}
    for line in text:
        m = re.match(r'\s*\.file\s+(\d+)\s+"([^"]*)"\s+"([^"]*)"', line)
        if m:
            files[int(m.group(1))] = os.path.join(m.group(2), m.group(3))
    src_defs, by_name = {}, {}
    for idx, path in files.items():
        base = os.path.basename(path)
        local = os.path.join(csrc, base)
        if os.path.exists(local):
            if base not in by_name:
                by_name[base] = definitions(local)
            src_defs[idx] = by_name[base]

    def func_at(base, line):
        if base not in by_name or line == 0:
            return None
        inner = None
        for s_, e_, name in by_name[base][0]:
            if s_ <= line <= e_ and (inner is None or e_ - s_ < inner[1] - inner[0]):
                inner = (s_, e_, name)
        return inner[2] if inner else None
    for stage, prefix in KERNELS.items():
        start = next(i for i, l in enumerate(text) if l.startswith(prefix) and ":" in l.split(";")[0])
        # the function runs to its end label: cold blocks follow the first s_endpgm
        end = next(i for i in range(start, len(text)) if text[i].startswith(".Lfunc_end"))
        # blocks of each outermost loop, from the compiler's loop annotations
        # ("Loop Header: Depth=1", "in Loop: Header=BBx_y Depth=1"); a block of
        # a nested loop belongs to the outermost loop around it
        parent, members = {}, collections.defaultdict(list)
        cur = None
        loc_at, chain_at = {}, {}
        loc, chain = (None, 0), []
        for i in range(start, end):
            line = text[i]
            m = re.match(r"\s*\.loc\s+(\d+)\s+(\d+)", line)
            if m:                          # the line-table entry in force, in layout order
                loc, chain = (int(m.group(1)), int(m.group(2))), frames_of(line)
            loc_at[i], chain_at[i] = loc, chain
            m = re.match(r"^(?:\.L(BB\w+):|; %bb\.(\d+):)(.*)", line)
            if m:
                name = m.group(1) or None
                note = m.group(3)
                h = re.search(r"Loop Header: Depth=(\d+)", note)
                inl = re.search(r"in Loop: Header=(BB\w+) Depth=(\d+)", note)
                if h and name:
                    d = int(h.group(1))
                    outer = cur if d > 1 else None
                    parent[name] = parent.get(outer, outer) if outer else name
                    cur = parent[name]
                elif inl:
                    cur = parent.get(inl.group(1), inl.group(1))
                else:
                    cur = None
            if cur:
                members[cur].append(i)
        for li, (hdr, idx) in enumerate(sorted(members.items(), key=lambda kv: kv[1][0])):
            if sum(1 for i in idx if re.match(r"\s+v_", text[i])) < 300:
                continue
            a, b = idx[0], idx[-1]
            by_phase = collections.Counter()
            by_helper = collections.Counter()
            cls = collections.Counter()
            ops = collections.Counter()
            for i in idx:
                line = text[i]
                loc, chain = loc_at[i], chain_at[i]
                m = re.match(r"\s+([sv]_\w+|buffer_\w+|global_\w+|scratch_\w+|ds_\w+)", line)
                if not m:
                    continue
                op = m.group(1)
                kind = "VALU" if op.startswith("v_") else "SALU" if op.startswith("s_") else "VMEM"
                if op.startswith("ds_"):
                    kind = "LDS"
                cls[kind] += 1
                if kind != "VALU":
                    continue
                ops[op] += 1
                if loc[0] in src_defs:
                    defs, src_lines = src_defs[loc[0]]
                    ph = phase_of(defs, src_lines, loc[1])
                elif "math" in os.path.basename(files.get(loc[0], "")):
                    ph = "library sqrt"
                else:
                    ph = "HIP runtime helpers (votes, lane ids)"
                by_helper[ph] += 1
                # fold a helper into the innermost major function that inlined it
                for base, ln in chain[1:] if chain else []:
                    f = func_at(base, ln)
                    if f in MAJOR:
                        ph = PHASE[f]
                        break
                by_phase[ph] += 1
            total = cls["VALU"]
            print(f"{stage}, loop at {hdr} (assembly lines {a + 1}..{b + 1}): VALU {total}, SALU {cls['SALU']}, "
                  f"VMEM {cls['VMEM']}, LDS {cls['LDS']}")
            print("  by phase (helpers counted in the phase that calls them):")
            for ph, c in by_phase.most_common():
                print(f"  {c:6d}  {100.0 * c / total:5.1f}%  {ph}")
            print("  by innermost source function:")
            for ph, c in by_helper.most_common():
                print(f"  {c:6d}  {100.0 * c / total:5.1f}%  {ph}")
            fp = ["v_div_scale_f64", "v_div_fmas_f64", "v_div_fixup_f64", "v_rcp_f64_e32", "v_rsq_f64_e32",
                  "v_sqrt_f64_e32", "v_mov_b64_e32", "v_cndmask_b32_e32", "v_cndmask_b32_e64", "v_mov_b32_dpp"]
            print("  selected opcodes: " + ", ".join(f"{o} {ops[o]}" for o in fp if ops[o]))
            print()


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else os.path.join(here, "csrc"))
