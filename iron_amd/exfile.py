"""Reader and writer for the CMISS exnode/exelem text formats that iron's
Fields_NodesExport / Fields_ElementsExport write (field_IO_routines.F90), and a
reader for the analytic-analysis (.opanal) reports of
analytic_analysis_routines.F90.

Only what regular Lagrange meshes produce is handled: node-based components with
derivatives and versions, and element "grid based" components (element-constant
fields such as an element pressure), which an exelem file stores as one row of
grid-point values per component after each "Element:" line.
"""
from __future__ import annotations

import re
from pathlib import Path

_FIELD_RE = re.compile(r"^\s*\d+\)\s*([^,]+),")
_NCOMP_RE = re.compile(r"#Components=(\d+)")
_NDERIV_RE = re.compile(r"#Derivatives=\s*(\d+)")
_NVER_RE = re.compile(r"#Versions=\s*(\d+)")
_GRID_RE = re.compile(r"#xi1=(\d+)(?:,\s*#xi2=(\d+))?(?:,\s*#xi3=(\d+))?")


def read_exnode(path) -> dict:
    """{node: {field name: [[values of component 1], [component 2], ...]}}.

    Each component's list holds (1 + #Derivatives) * #Versions values."""
    lines = Path(path).read_text().splitlines()
    out, fields, i = {}, [], 0
    while i < len(lines):
        s = lines[i].strip()
        if s.startswith("#Fields="):
            nf = int(s.split("=")[1])
            fields, i = [], i + 1
            for _ in range(nf):
                hdr = lines[i]
                i += 1
                name = _FIELD_RE.match(hdr).group(1).strip()
                nc = int(_NCOMP_RE.search(hdr).group(1))
                sizes = []
                for _ in range(nc):
                    cl = lines[i]
                    i += 1
                    nd = int(_NDERIV_RE.search(cl).group(1))
                    mv = _NVER_RE.search(cl)
                    sizes.append((nd + 1) * (int(mv.group(1)) if mv else 1))
                fields.append((name, sizes))
            continue
        if s.startswith("Node:"):
            node = int(s.split()[1])
            i += 1
            total = sum(sum(sz) for _, sz in fields)
            vals = []
            while len(vals) < total:
                vals += [float(v) for v in lines[i].split()]
                i += 1
            rec, k = {}, 0
            for name, sizes in fields:
                rec[name] = []
                for sz in sizes:
                    rec[name].append(vals[k:k + sz])
                    k += sz
            out[node] = rec
            continue
        i += 1
    return out


def read_exelem(path) -> dict:
    """{element: {"nodes": [...], "grid": {(field name, component): [grid values]}}}.

    Component numbers are 1-based as in the file."""
    lines = Path(path).read_text().splitlines()
    out, grid_comps, i = {}, [], 0
    field = None
    while i < len(lines):
        s = lines[i].strip()
        if s.startswith("#Fields="):
            grid_comps = []
            i += 1
            continue
        m = _FIELD_RE.match(lines[i])
        if m and "#Components=" in lines[i]:
            field = m.group(1).strip()
            i += 1
            continue
        if "grid based" in s and field is not None:
            comp = int(s.split(".")[0])
            g = _GRID_RE.search(lines[i + 1])
            npts = 1
            for v in g.groups():
                if v is not None:
                    npts *= int(v) + 1
            grid_comps.append(((field, comp), npts))
            i += 2
            continue
        if s.startswith("Element:"):
            elem = int(s.split()[1])
            rec = {"nodes": [], "grid": {}}
            i += 1
            while i < len(lines) and not lines[i].strip().startswith("Element:") and not lines[i].strip().startswith("Shape."):
                t = lines[i].strip()
                if t == "Values:":
                    i += 1
                    vals = []
                    total = sum(n for _, n in grid_comps)
                    while len(vals) < total:
                        vals += [float(v) for v in lines[i].split()]
                        i += 1
                    k = 0
                    for key, n in grid_comps:
                        rec["grid"][key] = vals[k:k + n]
                        k += n
                    continue
                if t == "Nodes:":
                    i += 1
                    while i < len(lines) and re.match(r"^\s*\d+(\s+\d+)*\s*$", lines[i]):
                        rec["nodes"] += [int(v) for v in lines[i].split()]
                        i += 1
                    continue
                i += 1
            out[elem] = rec
            continue
        i += 1
    return out


def write_exnode(path, group: str, fields: list, values: dict):
    """Write node values in the exnode layout read_exnode accepts.

    fields: [(name, kind, [component labels])] with kind e.g. "coordinate, rectangular
    cartesian"; values: {node: [[comp values] per component, over all fields]}."""
    lines = [f" Group name: {group}", f" #Fields={len(fields)}"]
    idx = 1
    for f, (name, kind, labels) in enumerate(fields):
        lines.append(f" {f + 1}) {name}, {kind}, #Components={len(labels)}")
        for lab in labels:
            lines.append(f"   {lab}.  Value index= {idx}, #Derivatives= 0")
            idx += 1
    for node in sorted(values):
        lines.append(f" Node: {node:12d}")
        for comp in values[node]:
            for v in comp:
                lines.append(f"  {v: .16E}")
    Path(path).write_text("\n".join(lines) + "\n")


def read_opanal(path) -> dict:
    """Nodal blocks of an analytic-analysis report:
    {variable label: {node: [numerical, analytic]}} (the "Nodal errors" table that
    AnalyticAnalysis_Output writes per field variable, one row per node:
    node, derivative, numerical, analytic, errors...)."""
    out, var = {}, None
    in_nodal = False
    for line in Path(path).read_text().splitlines():
        m = re.match(r"\s*Variable\s+\d+\s*:\s*(.+)$", line)
        if m:
            var = m.group(1).strip()
            out[var] = {}
            in_nodal = False
            continue
        if line.strip().startswith("Nodal errors"):
            in_nodal = True
            continue
        if in_nodal and var is not None:
            parts = line.split()
            if len(parts) >= 4 and parts[0].isdigit():
                out[var][int(parts[0])] = [float(parts[2]), float(parts[3])]
            elif line.strip().startswith("RMS"):
                in_nodal = False
    return out
