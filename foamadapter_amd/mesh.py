"""Meshes: OpenFOAM polyMesh primitives on the host, device meshes in libfvhip.

* ``PolyMesh`` -- points, faces, owner, neighbour, patches (what an OpenFOAM
  ``constant/polyMesh`` directory holds); read from ASCII case files
  (``read_polymesh``) or generated as a structured hex block (``PolyMesh.hex``,
  the blockMesh cases of the reference: benchmarks/benchmarkSuite/templates/
  3DCube/system/blockMeshDict).
* ``Mesh`` -- the device mesh (NeoN UnstructuredMesh + BoundaryMesh,
  src/datastructures/meshAdapter.cpp:59-136) built by libfvhip, either from a
  PolyMesh (geometry computed on the GPU) or from NeoN-layout geometry arrays.
"""
from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import call, lib
from .runtime import Context, prec_of


def _ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


@dataclass
class Patch:
    name: str
    type: str
    start: int
    size: int

    @property
    def empty(self) -> bool:
        return self.type == "empty"


@dataclass
class PolyMesh:
    points: np.ndarray          # (nP, 3) float64
    face_offsets: np.ndarray    # (nF+1,) int32
    face_points: np.ndarray     # (.,) int32
    owner: np.ndarray           # (nF,) int32
    neighbour: np.ndarray       # (nI,) int32
    patches: list[Patch] = field(default_factory=list)

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_internal_faces(self) -> int:
        return int(self.neighbour.shape[0])

    @property
    def n_cells(self) -> int:
        return int(max(self.owner.max(initial=-1), self.neighbour.max(initial=-1)) + 1)

    def fv_patches(self) -> list[Patch]:
        """Patches that carry finite-volume boundary faces (empty patches carry none)."""
        return [p for p in self.patches if not p.empty]

    def n_boundary_faces(self) -> int:
        return sum(p.size for p in self.fv_patches())

    def desc(self) -> tuple[_lib.FvPolyMeshDesc, list]:
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return C.c_void_p(a.ctypes.data) if a.size else None

        d = _lib.FvPolyMeshDesc()
        d.n_points = self.n_points
        d.n_faces = self.n_faces
        d.n_internal_faces = self.n_internal_faces
        d.n_cells = self.n_cells
        d.n_patches = len(self.patches)
        d.h_points = arr(self.points, np.float64)
        d.h_face_offsets = arr(self.face_offsets, np.int32)
        d.h_face_points = arr(self.face_points, np.int32)
        d.h_owner = arr(self.owner, np.int32)
        d.h_neighbour = arr(self.neighbour, np.int32)
        d.h_patch_start = arr([p.start for p in self.patches], np.int32)
        d.h_patch_size = arr([p.size for p in self.patches], np.int32)
        d.h_patch_empty = arr([1 if p.empty else 0 for p in self.patches], np.int32)
        return d, keep

    def renumbered(self, method: int = 0) -> tuple["PolyMesh", np.ndarray, np.ndarray]:
        """Reverse Cuthill-McKee renumbering (fv_polymesh_renumber).

        Returns (mesh, cell_new_to_old, face_new_to_old); face entries are ~old
        for faces whose orientation flipped."""
        d, keep = self.desc()
        nF, N = self.n_faces, self.n_cells
        foff = np.empty(nF + 1, np.int32)
        fpts = np.empty_like(self.face_points)
        own = np.empty(nF, np.int32)
        nbr = np.empty(self.n_internal_faces, np.int32)
        cmap = np.empty(N, np.int32)
        fmap = np.empty(nF, np.int32)
        call("fv_polymesh_renumber", C.byref(d), method, _ptr(foff), _ptr(fpts), _ptr(own),
             _ptr(nbr), _ptr(cmap), _ptr(fmap))
        del keep
        return PolyMesh(self.points, foff, fpts, own, nbr, list(self.patches)), cmap, fmap

    @staticmethod
    def hex(n: int | tuple[int, int, int], lengths=(1.0, 1.0, 1.0), jitter: float = 0.0,
            jitter_seed: int = 0, perm_seed: int = 0) -> "PolyMesh":
        """Structured hex block with patches x0, x1, y0, y1, z0, z1 (all walls)."""
        nx, ny, nz = (n, n, n) if isinstance(n, int) else n
        sizes = np.zeros(6, np.int64)
        call("fv_hexmesh_sizes", nx, ny, nz, _ptr(sizes))
        nP, nC, nI, nB, nPatch, nFP = (int(s) for s in sizes)
        nF = nI + nB
        pts = np.empty((nP, 3), np.float64)
        foff = np.empty(nF + 1, np.int32)
        fpts = np.empty(nFP, np.int32)
        own = np.empty(nF, np.int32)
        nbr = np.empty(nI, np.int32)
        pstart = np.empty(nPatch, np.int32)
        psize = np.empty(nPatch, np.int32)
        call("fv_hexmesh_generate", nx, ny, nz, float(lengths[0]), float(lengths[1]),
             float(lengths[2]), float(jitter), int(jitter_seed), int(perm_seed), _ptr(pts),
             _ptr(foff), _ptr(fpts), _ptr(own), _ptr(nbr), _ptr(pstart), _ptr(psize))
        names = ["x0", "x1", "y0", "y1", "z0", "z1"]
        patches = [Patch(names[i], "wall", int(pstart[i]), int(psize[i])) for i in range(nPatch)]
        return PolyMesh(pts, foff, fpts, own, nbr, patches)


# --------------------------------------------------------------------------
# OpenFOAM ASCII readers
# --------------------------------------------------------------------------

def _strip_header(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    i = text.find("FoamFile")
    if i >= 0:
        j = text.find("}", i)
        text = text[j + 1:]
    return text


def _list_body(text: str) -> tuple[int, str]:
    m = re.search(r"(\d+)\s*\(", text)
    if not m:
        raise ValueError("no list found")
    n = int(m.group(1))
    start = m.end()
    depth, i = 1, start
    while depth:
        ch = text[i]
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        i += 1
    return n, text[start:i - 1]


def read_label_list(path: Path) -> np.ndarray:
    n, body = _list_body(_strip_header(Path(path).read_text()))
    a = np.array(body.split(), dtype=np.int64)
    assert a.size == n, (path, a.size, n)
    return a.astype(np.int32)


def read_points(path: Path) -> np.ndarray:
    n, body = _list_body(_strip_header(Path(path).read_text()))
    a = np.array(body.replace("(", " ").replace(")", " ").split(), dtype=np.float64)
    assert a.size == 3 * n
    return a.reshape(n, 3)


def read_faces(path: Path) -> tuple[np.ndarray, np.ndarray]:
    n, body = _list_body(_strip_header(Path(path).read_text()))
    faces = re.findall(r"(\d+)\s*\(([^)]*)\)", body)
    assert len(faces) == n
    off = np.zeros(n + 1, np.int32)
    pts = []
    for i, (k, lst) in enumerate(faces):
        v = [int(x) for x in lst.split()]
        assert len(v) == int(k)
        off[i + 1] = off[i] + len(v)
        pts.extend(v)
    return off, np.array(pts, np.int32)


def read_boundary(path: Path) -> list[Patch]:
    text = _strip_header(Path(path).read_text())
    out = []
    for m in re.finditer(r"(\w+)\s*\{([^}]*)\}", text):
        body = m.group(2)
        typ = re.search(r"type\s+(\w+)\s*;", body).group(1)
        nf = int(re.search(r"nFaces\s+(\d+)\s*;", body).group(1))
        sf = int(re.search(r"startFace\s+(\d+)\s*;", body).group(1))
        out.append(Patch(m.group(1), typ, sf, nf))
    return out


def read_polymesh(case_dir: str | Path) -> PolyMesh:
    """Read ``constant/polyMesh`` of an OpenFOAM case in ASCII format."""
    d = Path(case_dir) / "constant" / "polyMesh"
    pts = read_points(d / "points")
    off, fp = read_faces(d / "faces")
    own = read_label_list(d / "owner")
    nbr = read_label_list(d / "neighbour")
    return PolyMesh(pts, off, fp, own, nbr, read_boundary(d / "boundary"))


def _field_values(text: str, key: str, n_expected: int | None, ncomp: int):
    m = re.search(key + r"\s+(uniform|nonuniform)", text)
    if not m:
        return None
    rest = text[m.end():]
    if m.group(1) == "uniform":
        v = re.match(r"\s*(\([^)]*\)|[-+0-9.eE]+)", rest).group(1)
        vals = np.array(v.strip("()").split(), dtype=np.float64)
        return np.tile(vals, (n_expected or 1, 1)).reshape(-1) if n_expected else vals
    n, body = _list_body(rest)
    a = np.array(body.replace("(", " ").replace(")", " ").split(), dtype=np.float64)
    assert a.size == n * ncomp
    return a


def read_internal_field(path: str | Path, ncomp: int = 1) -> np.ndarray:
    """internalField of an OpenFOAM ASCII vol/surface field (nonuniform list)."""
    text = _strip_header(Path(path).read_text())
    a = _field_values(text, "internalField", None, ncomp)
    return a.reshape(-1, ncomp) if ncomp > 1 else a


# --------------------------------------------------------------------------
# device mesh
# --------------------------------------------------------------------------

class Mesh:
    """Device mesh handle (fv_mesh)."""

    def __init__(self, ctx: Context, handle: C.c_void_p, patches: list[Patch] | None = None):
        self.ctx = ctx
        self.handle = handle
        s = np.zeros(4, np.int32)
        call("fv_mesh_sizes", handle, _ptr(s))
        self.n_cells, self.n_internal_faces, self.n_boundary_faces, self.n_patches = (int(x) for x in s)
        self.patches = patches or []
        ctx.adopt(self)

    cell_map = None   # device int32 [N]: device cell i holds caller cell cell_map[i]
    face_map = None   # device int32 [F+B]: device face j holds caller face (~ if flipped)

    RENUMBER = {"rcm": 0, "morton": 1}

    @classmethod
    def from_poly(cls, ctx: Context, poly: PolyMesh, renumber: bool | str = False) -> "Mesh":
        """Device mesh from polyMesh primitives (geometry computed on the GPU).

        renumber=True (Z-order, 'morton') or 'rcm' (reverse Cuthill-McKee) stores
        the mesh renumbered, as OpenFOAM's renumberMesh does to a case; fields
        then cross the boundary through to_device()/to_caller(), which apply the
        permutation."""
        cmap = fmap = None
        if renumber:
            method = self_method = cls.RENUMBER["morton" if renumber is True else renumber]
            poly, cmap, fmap = poly.renumbered(self_method)
            del method
        d, keep = poly.desc()
        h = C.c_void_p()
        call("fv_mesh_create_poly", ctx.handle, C.byref(d), C.byref(h))
        del keep
        m = cls(ctx, h, poly.fv_patches() if poly.patches else None)
        m.all_patches = list(poly.patches)   # including empty ones (field files list them)
        if renumber:
            F = m.n_internal_faces
            bidx = np.arange(F, F + m.n_boundary_faces, dtype=np.int32)
            m.cell_map = ctx.upload(cmap)
            m.face_map = ctx.upload(np.concatenate([fmap[:F], bidx]))
            m.poly = poly
        return m

    def _permute(self, src, mode: int, kind: str, oriented: bool):
        if self.cell_map is None:
            return src
        mp = self.cell_map if kind == "cell" else self.face_map
        ncomp = int(np.prod(src.shape[1:], dtype=np.int64)) if len(src.shape) > 1 else 1
        dst = self.ctx.empty(src.shape, src.dtype)
        call("fv_permute", self.ctx.handle, prec_of(src.dtype), src.shape[0], ncomp, mode,
             1 if oriented else 0, mp.ptr, src.ptr, dst.ptr)
        return dst

    def set_gather(self, mode: int):
        """0: LDS-staged face blocks when possible, each operator's default workgroup
        size (default); 1: per-cell stencil walk; 2/3/4: staged with 256/512/1024
        cells per workgroup."""
        call("fv_mesh_set_gather", self.handle, int(mode))

    def set_gather_run(self, run: int):
        """Consecutive bricks each XCD takes in the staged gather (default 8; 1: the
        natural order).  The results are the same bits for any run."""
        call("fv_mesh_set_gather_run", self.handle, int(run))

    @property
    def gather_kernel(self) -> str:
        m = C.c_int()
        call("fv_mesh_get_gather", self.handle, C.byref(m))
        return {1: "stencil-walk", 2: "staged-256", 3: "staged", 4: "staged-1024"}[m.value]

    def gather_stats(self, cells: int) -> tuple[int, int]:
        """(workgroups, workgroups taking the per-cell path) of the staged gather
        with `cells` cells per workgroup on this mesh."""
        c = np.zeros(2, np.int64)
        call("fv_mesh_gather_stats", self.handle, int(cells), _ptr(c))
        return int(c[0]), int(c[1])

    def to_device(self, src, kind: str = "cell", oriented: bool = False):
        """Caller-numbered device array -> this mesh's numbering (identity if not renumbered).
        kind: 'cell' ([N,...]) or 'face' ([F+B], oriented=True for fluxes)."""
        return self._permute(src, 0, kind, oriented)

    def to_caller(self, src, kind: str = "cell", oriented: bool = False):
        """This mesh's numbering -> the caller's numbering."""
        return self._permute(src, 1, kind, oriented)

    @classmethod
    def from_geometry(cls, ctx: Context, g: dict) -> "Mesh":
        """g: dict of NeoN-layout host arrays keyed like ``get`` names."""
        keep = []

        def arr(name, dt):
            a = np.ascontiguousarray(g[name], dtype=dt)
            keep.append(a)
            return C.c_void_p(a.ctypes.data) if a.size else None

        d = _lib.FvMeshDesc()
        d.n_cells = len(g["V"])
        d.n_internal_faces = len(g["owner"])
        d.n_boundary_faces = len(g["faceCells"])
        d.n_patches = len(g["patchOffsets"]) - 1
        f64, i32 = np.float64, np.int32
        d.h_V, d.h_C = arr("V", f64), arr("C", f64)
        d.h_owner, d.h_neighbour = arr("owner", i32), arr("neighbour", i32)
        d.h_Sf, d.h_Cf, d.h_magSf = arr("Sf", f64), arr("Cf", f64), arr("magSf", f64)
        d.h_weights, d.h_delta_coeffs = arr("weights", f64), arr("deltaCoeffs", f64)
        d.h_nonorth_delta_coeffs = arr("nonOrthDeltaCoeffs", f64)
        d.h_face_cells = arr("faceCells", i32)
        d.h_Sf_b, d.h_Cf_b, d.h_magSf_b = arr("Sf_b", f64), arr("Cf_b", f64), arr("magSf_b", f64)
        d.h_nf_b, d.h_delta_b, d.h_weights_b = arr("nf_b", f64), arr("delta_b", f64), arr("weights_b", f64)
        d.h_delta_coeffs_b = arr("deltaCoeffs_b", f64)
        d.h_nonorth_delta_coeffs_b = arr("nonOrthDeltaCoeffs_b", f64)
        d.h_patch_offsets = arr("patchOffsets", i32)
        h = C.c_void_p()
        call("fv_mesh_create", ctx.handle, C.byref(d), C.byref(h))
        del keep
        return cls(ctx, h)

    _SHAPES = {
        "V": ("N", 1, np.float64), "C": ("N", 3, np.float64),
        "owner": ("F", 1, np.int32), "neighbour": ("F", 1, np.int32),
        "Sf": ("F", 3, np.float64), "Cf": ("F", 3, np.float64), "magSf": ("F", 1, np.float64),
        "weights": ("F", 1, np.float64), "deltaCoeffs": ("F", 1, np.float64),
        "nonOrthDeltaCoeffs": ("F", 1, np.float64), "faceCells": ("B", 1, np.int32),
        "Sf_b": ("B", 3, np.float64), "Cf_b": ("B", 3, np.float64), "magSf_b": ("B", 1, np.float64),
        "nf_b": ("B", 3, np.float64), "delta_b": ("B", 3, np.float64),
        "weights_b": ("B", 1, np.float64), "deltaCoeffs_b": ("B", 1, np.float64),
        "nonOrthDeltaCoeffs_b": ("B", 1, np.float64), "patchOffsets": ("P1", 1, np.int32),
        "cellFaceOffsets": ("N1", 1, np.int32), "cellFaces": ("S", 1, np.int32),
    }

    def get(self, name: str) -> np.ndarray:
        kind, comp, dt = self._SHAPES[name]
        n = {"N": self.n_cells, "F": self.n_internal_faces, "B": self.n_boundary_faces,
             "P1": self.n_patches + 1, "N1": self.n_cells + 1,
             "S": 2 * self.n_internal_faces + self.n_boundary_faces}[kind]
        out = np.empty(n * comp, dt)
        call("fv_mesh_get", self.handle, _lib.GEOM_IDS[name], _ptr(out))
        return out.reshape(n, comp) if comp > 1 else out

    def geometry(self) -> dict:
        return {k: self.get(k) for k in self._SHAPES if k not in ("cellFaceOffsets", "cellFaces")}

    def close(self):
        if self.handle:
            lib.fv_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BoundaryConditions:
    """Per-patch boundary kinds of a field (fvcc::VolumeBoundary list,
    include/FoamAdapter/auxiliary/readers.hpp:31-108)."""

    KINDS = {"fixedValue": 0, "fixedGradient": 1, "zeroGradient": 1, "calculated": 2,
             "extrapolatedCalculated": 2, "noSlip": 0}

    def __init__(self, mesh: Mesh, kinds):
        ks = np.array([self.KINDS[k] if isinstance(k, str) else int(k) for k in kinds], np.int32)
        if ks.size != mesh.n_patches:
            raise ValueError(f"need {mesh.n_patches} patch kinds, got {ks.size}")
        self.mesh = mesh
        self.kinds = ks
        h = C.c_void_p()
        call("fv_bc_create", mesh.handle, _ptr(ks), C.byref(h))
        self.handle = h
        mesh.ctx.adopt(self)

    def close(self):
        if self.handle:
            lib.fv_bc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
