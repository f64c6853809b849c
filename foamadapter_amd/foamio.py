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


def read_field(path, patches: list[Patch], ncomp: int = 1):
    """(internal [N] or [N, 3], {patch name: boundary values}) of an ASCII field file."""
    text = _strip_header(Path(path).read_text())
    internal = _field_values(text, "internalField", None, ncomp)
    bf = text[text.index("boundaryField"):]
    out = {}
    for p in patches:
        if p.empty:
            continue
        i = bf.find(p.name)
        if i < 0:
            raise ValueError(f"patch {p.name} missing in {path}")
        seg = bf[i:]
        seg = seg[:seg.index("}") + 1] if "nonuniform" not in seg.split("}")[0] else seg
        vals = _field_values(seg, "value", p.size, ncomp)
        out[p.name] = None if vals is None else vals.reshape(-1, ncomp) if ncomp > 1 else vals
    return (internal.reshape(-1, ncomp) if ncomp > 1 else internal), out
