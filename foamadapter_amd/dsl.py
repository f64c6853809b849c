"""Expressions and solve with NeoN's dsl names, plus FoamAdapter's PDESolver.

    eqn = dsl.Expression(dsl.imp.ddt(T) - dsl.imp.laplacian(kappa, T))
    stats = dsl.solve(eqn, T, t, dt, solver={"relTol": 1e-10, "maxIter": 1000})

(examples/heatTransfer/heatTransfer.cpp:67-76).  Terms are assembled in order
into the mesh's LinearSystem (implicit terms through fv_assemble_*, explicit
terms evaluated by the face operators and moved to the right-hand side), then
solved by Jacobi-PCG on the GPU.  Euler time stepping: the ddt term reads the
field's old-time copy, which solve() refreshes after each solve.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._lib import FV_LINEAR, FV_UPWIND, call
from . import fvcc
from .la import LinearSystem, SparsityPattern, SolverStats, solve_cg
from .runtime import DeviceArray


@dataclass
class Term:
    kind: str              # ddt | laplacian | div | source | exp_div | exp_laplacian
    implicit: bool
    args: tuple
    coeff: float = 1.0
    scheme: str = "linear"

    def __mul__(self, c):
        return Term(self.kind, self.implicit, self.args, self.coeff * float(c), self.scheme)

    __rmul__ = __mul__

    def __neg__(self):
        return self * -1.0

    def __add__(self, other):
        return Expression([self]) + other

    def __sub__(self, other):
        return Expression([self]) - other


@dataclass
class Expression:
    terms: list = field(default_factory=list)

    def __init__(self, terms=None):
        if isinstance(terms, Expression):
            terms = terms.terms
        elif isinstance(terms, Term):
            terms = [terms]
        self.terms = list(terms or [])

    def __add__(self, other):
        o = other.terms if isinstance(other, Expression) else [other]
        return Expression(self.terms + o)

    def __sub__(self, other):
        o = other.terms if isinstance(other, Expression) else [other]
        return Expression(self.terms + [-t for t in o])

    def __neg__(self):
        return Expression([-t for t in self.terms])

    def read(self, fvSchemes: dict | None):
        """Pick div schemes from an fvSchemes-like dict {'div': 'Gauss upwind'}."""
        if fvSchemes and "div" in fvSchemes:
            s = fvSchemes["div"].split()[-1]
            for t in self.terms:
                if t.kind in ("div", "exp_div"):
                    t.scheme = s
        return self


class imp:
    """Implicit operators (dsl::imp)."""

    @staticmethod
    def ddt(psi):
        return Term("ddt", True, (psi,))

    @staticmethod
    def laplacian(gamma, psi):
        return Term("laplacian", True, (gamma, psi))

    @staticmethod
    def div(phi, psi, scheme="linear"):
        return Term("div", True, (phi, psi), scheme=scheme)

    @staticmethod
    def source(coeff_field, psi):
        return Term("source", True, (coeff_field, psi))


class exp:
    """Explicit operators (dsl::exp)."""

    @staticmethod
    def div(phi, psi=None, scheme="linear"):
        return Term("exp_div", False, (phi, psi), scheme=scheme)

    @staticmethod
    def laplacian(gamma, psi):
        return Term("exp_laplacian", False, (gamma, psi))


def _ptr_or_none(a):
    return a.ptr if a is not None else None


def _linsys(mesh) -> LinearSystem:
    ls = getattr(mesh, "_linsys", None)
    if ls is None or ls.handle is None:
        ls = LinearSystem(SparsityPattern.readOrCreate(mesh))
        mesh._linsys = ls
    return ls


def assemble(expr: Expression, psi: fvcc.VolumeField, t: float, dt: float,
             ls: LinearSystem | None = None) -> LinearSystem:
    """Expression::assemble(t, dt, sparsity, ls) (expression.hpp:69-73)."""
    mesh = psi.mesh
    ctx = mesh.ctx
    ls = ls or _linsys(mesh)
    ls.zero()
    tmp = None
    for term in expr.terms:
        k, c = term.kind, term.coeff
        if k == "ddt":
            if psi.old_time is None:
                psi.old_time = fvcc.VolumeField(mesh, _copy(psi.internal), None)
            call("fv_assemble_ddt_euler", ls.handle, psi.old_time.internal.ptr, float(dt), c)
        elif k == "laplacian":
            gamma, f = term.args
            call("fv_assemble_laplacian", ls.handle, gamma.values.ptr, f.bcs.handle,
                 f.boundary.ptr, _ptr_or_none(f.boundary_gradient), c)
        elif k == "div":
            phi, f = term.args
            sch = FV_UPWIND if term.scheme == "upwind" else FV_LINEAR
            call("fv_assemble_div", ls.handle, sch, phi.values.ptr, f.bcs.handle, f.boundary.ptr,
                 _ptr_or_none(f.boundary_gradient), c)
        elif k == "source":
            cf, f = term.args
            arr = cf.internal if isinstance(cf, fvcc.VolumeField) else cf
            call("fv_assemble_source", ls.handle, arr.ptr, c)
        elif k in ("exp_div", "exp_laplacian"):
            if tmp is None:
                tmp = ctx.empty((mesh.n_cells,))
            if k == "exp_div":
                phi, f = term.args
                fvcc.GaussGreenDiv(ctx, mesh, [term.scheme]).div(tmp, phi, f, 1.0)
            else:
                gamma, f = term.args
                fvcc.GaussGreenLaplacian(ctx, mesh).laplacian(tmp, gamma, f, 1.0)
            call("fv_assemble_explicit", ls.handle, tmp.ptr, c)
        else:
            raise ValueError(f"unknown term {k}")
    return ls


def _copy(a: DeviceArray) -> DeviceArray:
    d = a.ctx.empty(a.shape, a.dtype)
    call("fv_memcpy_d2d", a.ctx.handle, d.ptr, a.ptr, a.nbytes)
    return d


def solve(expr: Expression, psi: fvcc.VolumeField, t: float = 0.0, dt: float = 1.0,
          fvSchemes: dict | None = None, solver: dict | None = None,
          post_assembly=()) -> SolverStats:
    """dsl::solve(expr, psi, t, dt, fvSchemes, solverDict) (heatTransfer.cpp:69-76)."""
    expr = Expression(expr).read(fvSchemes)
    ls = assemble(expr, psi, t, dt)
    for fn in post_assembly:
        fn(ls)
    stats = solve_cg(ls, psi.internal, solver)
    psi.correct_boundary_conditions()
    if any(tm.kind == "ddt" for tm in expr.terms):
        call("fv_memcpy_d2d", psi.mesh.ctx.handle, psi.old_time.internal.ptr, psi.internal.ptr,
             psi.internal.nbytes)
    return stats


class PDESolver:
    """FoamAdapter::PDESolver (include/FoamAdapter/datastructures/expression.hpp:23-179)."""

    def __init__(self, expr: Expression, psi: fvcc.VolumeField, t: float = 0.0, dt: float = 1.0,
                 fvSchemes: dict | None = None, solver: dict | None = None):
        self.expr = Expression(expr).read(fvSchemes)
        self.psi, self.t, self.dt, self.solver = psi, t, dt, solver
        self.sparsityPattern = SparsityPattern.readOrCreate(psi.mesh)
        self.ls = LinearSystem(self.sparsityPattern)
        self._ref = None

    def assemble(self) -> LinearSystem:
        return assemble(self.expr, self.psi, self.t, self.dt, self.ls)

    def setReference(self, cell: int, value: float):
        self._ref = (int(cell), float(value))

    def linearSystem(self) -> LinearSystem:
        return self.ls

    def solve(self) -> SolverStats:
        self.assemble()
        if self._ref is not None:
            call("fv_set_reference", self.ls.handle, self._ref[0], self._ref[1])
        stats = solve_cg(self.ls, self.psi.internal, self.solver)
        self.psi.correct_boundary_conditions()
        print(f"[fvhip] Solving for {self.psi.name}: Initial residual: {stats.initResNorm} "
              f"Final residual: {stats.finalResNorm} No Iterations: {stats.numIter}")
        return stats


def applyOperator(ls: LinearSystem, psi: fvcc.VolumeField) -> DeviceArray:
    """FoamAdapter::applyOperator: A.psi - b (expression.hpp:204-220)."""
    return ls.residual(psi.internal)
