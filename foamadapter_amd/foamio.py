"""OpenFOAM ASCII field files: write (FoamAdapter::write, src/auxiliary/writers.cpp:9-135)
and read back, for checkpoint and restart of cases.

write_field writes <case>/<time>/<name> as a volScalarField (values [N]) or
volVectorField ([N, 3]) with dimensions [0 0 0 0 0 0 0] (dimless, as the
reference's writers), the internal field as a nonuniform list and every patch as
`calculated` with its boundary values -- boundary values are flat over the
non-empty patches in patch order (the NeoN layout); empty patches are written as
`empty`.  read_field returns (internal, {patch: values}).  Values are written
with 17 significant digits, so a write/read round trip is exact in FP64.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from .mesh import Patch, _field_values, _strip_header

_HEADER = """/*--------------------------------*- C++ -*----------------------------------*\\
  =========                 |
  \\\\      /  F ield         | written by foamadapter_amd
   \\\\    /   O peration     |
    \\\\  /    A nd           |
     \\\\/     M anipulation  |
\\*---------------------------------------------------------------------------*/
FoamFile
{{
    version     2.0;
    format      ascii;
    class       {cls};
    location    "{loc}";
    object      {name};
}}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

"""


def _fmt(v: float) -> str:
    return repr(float(v)) if np.isfinite(v) else ("nan" if np.isnan(v) else ("inf" if v > 0 else "-inf"))


def _list(values: np.ndarray, ncomp: int) -> str:
    kind = "scalar" if ncomp == 1 else "vector"
    rows = values.reshape(-1) if ncomp == 1 else values.reshape(-1, ncomp)
    out = [f"nonuniform List<{kind}> \n{len(rows)}\n("]
    if ncomp == 1:
        out += [_fmt(v) for v in rows]
    else:
        out += ["(" + " ".join(_fmt(x) for x in r) + ")" for r in rows]
    out.append(")")
    return "\n".join(out)


def write_field(case_dir, time: str, name: str, values, patches: list[Patch],
                boundary=None) -> Path:
    """Write a cell field as <case_dir>/<time>/<name> (volScalarField or volVectorField)."""
    v = np.asarray(values, dtype=np.float64)
    ncomp = 1 if v.ndim == 1 else v.shape[1]
    if ncomp not in (1, 3):
        raise ValueError("scalar [N] or vector [N, 3] fields only")
    live = [p for p in patches if not p.empty]
    nB = sum(p.size for p in live)
    b = None if boundary is None else np.asarray(boundary, dtype=np.float64).reshape(nB, ncomp)
    cls = "volScalarField" if ncomp == 1 else "volVectorField"
    text = [_HEADER.format(cls=cls, loc=time, name=name), "dimensions      [0 0 0 0 0 0 0];\n\n",
            "internalField   " + _list(v, ncomp) + ";\n\n", "boundaryField\n{\n"]
    off = 0
    for p in patches:
        text.append(f"    {p.name}\n    {{\n")
        if p.empty:
            text.append("        type            empty;\n")
        else:
            text.append("        type            calculated;\n")
            if b is not None and p.size:
                text.append("        value           " + _list(b[off:off + p.size], ncomp) + ";\n")
            elif p.size:
                zero = "0" if ncomp == 1 else "(0 0 0)"
                text.append(f"        value           uniform {zero};\n")
            off += p.size
        text.append("    }\n")
    text.append("}\n\n\n// ************************************************************************* //\n")
    d = Path(case_dir) / time
    d.mkdir(parents=True, exist_ok=True)
    path = d / name
    path.write_text("".join(text))
    return path


def _subdicts(text: str, where) -> dict:
    """{keyword: body} of the top-level `keyword { body }` entries of a dictionary
    body, bracketed by matching braces (keywords compared exactly, quotes
    removed), so a patch whose name is part of another's is never confused with
    it.  Entries that are not sub-dictionaries are skipped."""
    out, depth, start, key = {}, 0, 0, None
    tok = re.compile(r'[{};]|"[^"]*"|[^\s{};"]+')
    last = None
    for m in tok.finditer(text):
        t = m.group(0)
        if t == "{":
            if depth == 0:
                if last is None:
                    raise ValueError(f"sub-dictionary without a keyword in {where}")
                key, start = last.strip('"'), m.end()
            depth += 1
        elif t == "}":
            depth -= 1
            if depth < 0:
                raise ValueError(f"unbalanced braces in {where}")
            if depth == 0:
                out[key] = text[start:m.start()]
            last = None
        elif depth == 0:
            last = None if t == ";" else t
    if depth:
        raise ValueError(f"unbalanced braces in {where}")
    return out


def read_field(path, patches: list[Patch], ncomp: int = 1):
    """(internal [N] or [N, 3], {patch name: boundary values}) of an ASCII field file."""
    text = _strip_header(Path(path).read_text())
    internal = _field_values(text, "internalField", None, ncomp)
    top = _subdicts(text[text.index("boundaryField"):], path)
    if "boundaryField" not in top:
        raise ValueError(f"no boundaryField in {path}")
    bf = _subdicts(top["boundaryField"], path)
    out = {}
    for p in patches:
        if p.empty:
            continue
        if p.name not in bf:
            raise ValueError(f"patch {p.name} missing in {path}")
        vals = _field_values(bf[p.name], "value", p.size, ncomp)
        out[p.name] = None if vals is None else vals.reshape(-1, ncomp) if ncomp > 1 else vals
    return (internal.reshape(-1, ncomp) if ncomp > 1 else internal), out


def read_boundary_conditions(path, patches: list[Patch], ncomp: int = 1) -> dict:
    """{patch: (kind, boundary values [size] / [size, 3])} of an ASCII field file,
    mapped as FoamAdapter's readVolBoundaryConditions maps them
    (include/FoamAdapter/auxiliary/readers.hpp:43-95): fixedValue keeps its value,
    noSlip is fixedValue 0, zeroGradient is fixedGradient 0, fixedGradient keeps
    its gradient, calculated / extrapolatedCalculated are calculated (value or
    0).  Empty patches are skipped."""
    text = _strip_header(Path(path).read_text())
    top = _subdicts(text[text.index("boundaryField"):], path)
    bf = _subdicts(top["boundaryField"], path)
    out = {}
    for p in patches:
        if p.empty:
            continue
        if p.name not in bf:
            raise ValueError(f"patch {p.name} missing in {path}")
        body = bf[p.name]
        m = re.search(r"\btype\s+([A-Za-z]+)\s*;", body)
        if not m:
            raise ValueError(f"patch {p.name} of {path} has no type")
        t = m.group(1)
        shape = (p.size,) if ncomp == 1 else (p.size, ncomp)
        if t == "noSlip":
            kind, vals = "fixedValue", np.zeros(shape)
        elif t == "zeroGradient":
            kind, vals = "fixedGradient", np.zeros(shape)
        elif t == "fixedGradient":
            kind, vals = t, _field_values(body, "gradient", p.size, ncomp)
        elif t in ("fixedValue", "calculated", "extrapolatedCalculated"):
            kind = "fixedValue" if t == "fixedValue" else "calculated"
            vals = _field_values(body, "value", p.size, ncomp)
            vals = np.zeros(shape) if vals is None else vals
        else:
            raise ValueError(f"boundary condition {t} of patch {p.name} is not supported")
        out[p.name] = (kind, np.asarray(vals, np.float64).reshape(shape))
    return out
