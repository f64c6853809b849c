"""OpenFOAM case directories, as the reference's tutorials lay them out, run by
the reference's applications on the GPU path:

* load_icofoam(ctx, case): examples/neoIcoFoam (tutorials/cavity);
* load_heat_transfer(ctx, case): examples/heatTransfer (tutorials/heatTranfer2D).

The mesh comes from constant/polyMesh, or from a single-block
system/blockMeshDict (blockmesh.read_block_mesh_dict); the fields from the
start time directory (0, else 0.orig) with their boundary conditions mapped as
FoamAdapter's readers map them; the physical constants from
constant/transportProperties; the time step from system/controlDict; the
schemes and solver entries from system/fvSchemes and system/fvSolution.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import dsl, fvcc
from .piso import IcoFoam


def _number(v) -> float:
    """The value of `nu 0.01;` or of the dimensioned `nu [0 2 -1 0 0 0 0] 0.01;`."""
    if isinstance(v, list):
        v = v[-1]
    return float(v)


def read_case_mesh(case_dir):
    from .blockmesh import read_block_mesh_dict
    from .mesh import read_polymesh
    case = Path(case_dir)
    if (case / "constant" / "polyMesh" / "owner").exists():
        return read_polymesh(case)
    return read_block_mesh_dict(case / "system" / "blockMeshDict")


def volume_field(ctx, mesh, poly, path, ncomp, name) -> fvcc.VolumeField:
    """VolumeField of a case field file: internal values, boundary conditions
    mapped as readers.hpp maps them, fixedGradient gradients, and the boundary
    values those conditions give."""
    from .foamio import read_boundary_conditions, read_field
    from .mesh import BoundaryConditions
    internal, _ = read_field(path, poly.patches, ncomp)
    N = mesh.n_cells
    internal = np.asarray(internal, np.float64)
    if internal.size == ncomp:                       # uniform
        internal = np.tile(internal.reshape(-1), (N, 1)).reshape((N,) if ncomp == 1 else (N, ncomp))
    bcs = read_boundary_conditions(path, poly.patches, ncomp)
    # one kind per patch of the mesh; empty patches carry no boundary faces
    kinds = ["calculated" if p.empty else bcs[p.name][0] for p in poly.patches]
    shape = (mesh.n_boundary_faces,) if ncomp == 1 else (mesh.n_boundary_faces, ncomp)
    vals, grads = np.zeros(shape), np.zeros(shape)
    off = 0
    for p in poly.fv_patches():
        k, v = bcs[p.name]
        if k == "fixedGradient":
            grads[off:off + p.size] = v
        else:
            vals[off:off + p.size] = v
        off += p.size
    bc = BoundaryConditions(mesh, kinds)
    g = ctx.upload(grads) if np.any(grads) else None
    f = fvcc.VolumeField(mesh, ctx.upload(internal), ctx.upload(vals), bc, name, g)
    f.correct_boundary_conditions()                  # fixedGradient faces from their cells
    return f


def _start_dir(case: Path, control: dict, field: str) -> Path:
    start = str(control.get("startTime", 0))
    for d in (start, "0", "0.orig"):
        if (case / d / field).exists():
            return case / d
    raise ValueError(f"{case}: no {field} field in the start time directory")


def _dicts(case: Path):
    from .foamdict import read_foam_dict
    return (read_foam_dict(case / "system" / "controlDict"),
            read_foam_dict(case / "system" / "fvSchemes"),
            read_foam_dict(case / "system" / "fvSolution"),
            read_foam_dict(case / "constant" / "transportProperties"))


def load_icofoam(ctx, case_dir, overrides: dict | None = None,
                 solver_entries: dict | None = None) -> IcoFoam:
    """IcoFoam on a case as neoIcoFoam reads one (examples/neoIcoFoam/
    neoIcoFoam.cpp:30-78): U and p, nu, deltaT / adjustTimeStep / maxCo /
    maxDeltaT, the scheme of div(phi,U), the U and p solvers and the PISO
    controls.  solver_entries names the fvSolution entries used for U and p
    (default {"U": "U", "p": "p"}; the reference's cavity tutorial also holds
    Ginkgo-form entries nfU / nfP); overrides replaces entries of the result
    ({"p": {...}, "U": {...}, "PISO": {...}})."""
    from .mesh import Mesh
    case = Path(case_dir)
    poly = read_case_mesh(case)
    mesh = Mesh.from_poly(ctx, poly)
    control, schemes, solution, transport = _dicts(case)
    nu = _number(transport["nu"])
    tdir = _start_dir(case, control, "U")
    U = volume_field(ctx, mesh, poly, tdir / "U", 3, "U")
    p = volume_field(ctx, mesh, poly, tdir / "p", 1, "p")
    phi = fvcc.flux(ctx, mesh, U)
    nFB = mesh.n_internal_faces + mesh.n_boundary_faces
    nuf = fvcc.SurfaceField(mesh, ctx.upload(np.full(nFB, nu)), "nu")
    div = schemes.get("divSchemes", {})
    conv = div.get("div(phi,U)", div.get("default"))
    fvSchemes = {"div": " ".join(conv) if isinstance(conv, list) else str(conv)} if conv else None
    sols = solution.get("solvers", {})
    names = {"U": "U", "p": "p", **(solver_entries or {})}
    fvSolution = {"U": dict(sols.get(names["U"], {})), "p": dict(sols.get(names["p"], {})),
                  "PISO": dict(solution.get("PISO", {}))}
    for k, v in (overrides or {}).items():
        fvSolution[k] = {**fvSolution.get(k, {}), **v}
    yes = ("yes", "on", "true", 1)
    solver = IcoFoam(U, p, phi, nuf, float(control["deltaT"]),
                     t0=float(control.get("startTime", 0)), fvSchemes=fvSchemes,
                     fvSolution=fvSolution,
                     adjustTimeStep=control.get("adjustTimeStep", "no") in yes,
                     maxCo=float(control.get("maxCo", 1.0)),
                     maxDeltaT=float(control.get("maxDeltaT", np.inf)))
    solver.poly, solver.nu = poly, nu
    solver.endTime = float(control.get("endTime", 0))
    return solver


class HeatTransfer:
    """examples/heatTransfer/heatTransfer.cpp:60-86: each time step solves
    ddt(T) - laplacian(kappa, T) = 0 with dsl::solve (Euler, the kappa surface
    field, the solver entry of fvSolution); the old time is the field of the
    step before."""

    def __init__(self, T: fvcc.VolumeField, kappa: fvcc.SurfaceField, deltaT: float,
                 t0: float = 0.0, fvSchemes: dict | None = None, solver: dict | None = None):
        self.T, self.kappa, self.deltaT, self.t = T, kappa, float(deltaT), float(t0)
        self.fvSchemes, self.solver = fvSchemes, solver
        if T.old_time is None:
            from .dsl import _copy
            T.old_time = fvcc.VolumeField(T.mesh, _copy(T.internal), None)

    def step(self):
        self.t += self.deltaT
        eqn = dsl.imp.ddt(self.T) - dsl.imp.laplacian(self.kappa, self.T)
        return dsl.solve(eqn, self.T, self.t, self.deltaT, self.fvSchemes, self.solver)

    def run(self, nSteps: int):
        return [self.step() for _ in range(nSteps)]


def load_heat_transfer(ctx, case_dir, solver_entry: str = "nfT",
                       overrides: dict | None = None) -> HeatTransfer:
    """HeatTransfer on a case as heatTransfer.cpp reads one: T from the start
    time directory, kappa from transportProperties, deltaT from controlDict,
    the solver entry `solver_entry` of fvSolution (the reference reads nfT)."""
    from .mesh import Mesh
    case = Path(case_dir)
    poly = read_case_mesh(case)
    mesh = Mesh.from_poly(ctx, poly)
    control, schemes, solution, transport = _dicts(case)
    kappa = _number(transport["kappa"])
    T = volume_field(ctx, mesh, poly, _start_dir(case, control, "T") / "T", 1, "T")
    nFB = mesh.n_internal_faces + mesh.n_boundary_faces
    k = fvcc.SurfaceField(mesh, ctx.upload(np.full(nFB, kappa)), "kappa")
    solver = {**dict(solution.get("solvers", {}).get(solver_entry, {})), **(overrides or {})}
    h = HeatTransfer(T, k, float(control["deltaT"]), float(control.get("startTime", 0)),
                     None, solver)
    h.poly, h.kappa_value = poly, kappa
    h.endTime = float(control.get("endTime", 0))
    return h
