"""The single-block cases of the reference's tutorials (tutorials/cavity,
tutorials/heatTranfer2D: system/blockMeshDict with one `hex` block and
`simpleGrading (1 1 1)`) meshed without OpenFOAM's blockMesh.

The block must be an axis-aligned box whose vertices follow blockMesh's hex
ordering (0-3 the bottom face counter-clockwise from the minimum corner, 4-7
above them).  The mesh is the structured hex block of PolyMesh.hex (cells
numbered x fastest, then y, then z, as blockMesh numbers one block), scaled by `scale` /
`convertToMeters` and placed at vertex 0, and each boundary patch of the
dictionary collects the block sides whose vertex quads it lists, in its order;
sides no patch lists form `defaultFaces` (empty, as blockMesh's default).
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from .mesh import Patch, PolyMesh

# block sides by the hex-local vertices of their quads (blockMesh's face
# numbering), in the order PolyMesh.hex lays out its boundary patches
_SIDES = {"x0": {0, 4, 7, 3}, "x1": {1, 2, 6, 5}, "y0": {0, 1, 5, 4},
          "y1": {3, 7, 6, 2}, "z0": {0, 3, 2, 1}, "z1": {4, 5, 6, 7}}
_ORDER = ["x0", "x1", "y0", "y1", "z0", "z1"]


def _strip(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _section(text: str, key: str) -> str:
    """The body of `key ( ... );` with balanced parentheses."""
    m = re.search(r"\b" + re.escape(key) + r"\s*\(", text)
    if not m:
        raise ValueError(f"blockMeshDict: no {key} list")
    i, depth = m.end(), 1
    start = i
    while depth:
        if i >= len(text):
            raise ValueError(f"blockMeshDict: unbalanced {key} list")
        depth += {"(": 1, ")": -1}.get(text[i], 0)
        i += 1
    return text[start:i - 1]


def _ints(s: str) -> list[int]:
    return [int(x) for x in s.split()]


def read_block_mesh_dict(path) -> PolyMesh:
    text = _strip(Path(path).read_text())
    m = re.search(r"\b(?:scale|convertToMeters)\s+([^;]+);", text)
    scale = float(m.group(1)) if m else 1.0
    verts = np.array([[float(x) for x in t.split()]
                      for t in re.findall(r"\(([^()]*)\)", _section(text, "vertices"))])
    blocks = _section(text, "blocks")
    bm = re.findall(r"hex\s*\(([^)]*)\)\s*\(([^)]*)\)\s*simpleGrading\s*\(([^)]*)\)", blocks)
    if len(bm) != 1:
        raise ValueError("blockMeshDict: exactly one hex block with simpleGrading is supported")
    hv, cells, grading = _ints(bm[0][0]), _ints(bm[0][1]), [float(x) for x in bm[0][2].split()]
    if len(hv) != 8 or len(cells) != 3 or any(g != 1.0 for g in grading):
        raise ValueError("blockMeshDict: a hex of 8 vertices, 3 cell counts and unit grading")
    v = verts[hv]
    lo, hi = v[0], v[6]
    box = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
                    [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                    [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]])
    if not np.allclose(v, box) or np.any(hi <= lo):
        raise ValueError("blockMeshDict: the block must be an axis-aligned box in hex order")
    local = {g: i for i, g in enumerate(hv)}          # global vertex -> hex-local index
    # boundary: name { type T; faces ( (a b c d) ... ); }
    bnd = _section(text, "boundary")
    named = []
    for name, body in re.findall(r"(\w+)\s*\{(.*?)\}", bnd, flags=re.S):
        tm = re.search(r"\btype\s+(\w+)\s*;", body)
        quads = re.findall(r"\(([^()]*)\)", _section(body, "faces"))
        sides = []
        for q in quads:
            ids = set(local[int(x)] for x in q.split())
            side = [k for k, s in _SIDES.items() if s == ids]
            if len(side) != 1:
                raise ValueError(f"blockMeshDict: face ({q}) of {name} is not a side of the block")
            sides.append(side[0])
        named.append((name, tm.group(1) if tm else "patch", sides))
    used = [s for _, _, ss in named for s in ss]
    if len(used) != len(set(used)):
        raise ValueError("blockMeshDict: a block side is in two patches")
    rest = [s for s in _ORDER if s not in used]
    if rest:
        named.append(("defaultFaces", "empty", rest))
    base = PolyMesh.hex(tuple(cells), lengths=tuple((hi - lo) * scale))
    return _regroup(base, named, lo * scale)


def _regroup(base: PolyMesh, named, origin) -> PolyMesh:
    """base (patches x0 ... z1) with its boundary faces regrouped into the
    named patches, in their order, and its points moved to `origin`."""
    nI = base.n_internal_faces
    side = {p.name: p for p in base.patches}
    foff, fpts = base.face_offsets, base.face_points
    order = list(range(nI))
    patches = []
    for name, ptype, sides in named:
        start = len(order)
        for s in sides:
            p = side[s]
            order.extend(range(p.start, p.start + p.size))
        patches.append(Patch(name, ptype, start, len(order) - start))
    order = np.array(order, np.int64)
    counts = np.diff(foff)[order]
    new_off = np.zeros(len(order) + 1, np.int32)
    new_off[1:] = np.cumsum(counts)
    starts = foff[order].astype(np.int64)
    idx = np.repeat(starts - new_off[:-1], counts) + np.arange(int(new_off[-1]))
    new_pts = fpts[idx].astype(np.int32)
    return PolyMesh(base.points + np.asarray(origin), new_off, new_pts,
                    base.owner[order].astype(np.int32), base.neighbour.copy(), patches)
