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

from ._lib import FV_LINEAR, FV_UPWIND, call, lib
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


def _on_component(expr: Expression, psi, psi_k) -> Expression:
    """The expression with the field psi replaced by its component psi_k."""
    def swap(a):
        return psi_k if a is psi else a
    return Expression([Term(t.kind, t.implicit, tuple(swap(a) for a in t.args), t.coeff, t.scheme)
                       for t in expr.terms])


def assemble(expr: Expression, psi: fvcc.VolumeField, t: float, dt: float,
             ls: LinearSystem | None = None) -> LinearSystem:
    """Expression::assemble(t, dt, sparsity, ls) (expression.hpp:69-73).

    A Vec3 field (PDESolver<Vec3>, the momentum predictor of
    examples/neoIcoFoam/neoIcoFoam.cpp:94-96) is assembled component by
    component: the matrix, which the boundary-condition kinds shared by the
    components make the same for all three, is that of component 0, and the
    three right-hand sides are kept as ``ls.rhs3`` ([N, 3] device array); the
    system's own right-hand side holds component 0."""
    if psi.ncomp > 1:
        return _assemble_vec(expr, psi, t, dt, ls)
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


def _assemble_vec(expr, psi, t, dt, ls):
    mesh = psi.mesh
    ctx = mesh.ctx
    ls = ls or _linsys(mesh)
    if psi.old_time is None and any(tm.kind == "ddt" for tm in expr.terms):
        psi.old_time = fvcc.VolumeField(mesh, _copy(psi.internal), None)
    rhs3 = ctx.empty((mesh.n_cells, psi.ncomp))
    rhs = ctx.empty((mesh.n_cells,))
    for k in reversed(range(psi.ncomp)):          # component 0 last: its system stays in ls
        pk = psi.component(k)
        assemble(_on_component(expr, psi, pk), pk, t, dt, ls)
        call("fv_memcpy_d2d", ctx.handle, rhs.ptr, lib.fv_linsys_rhs(ls.handle), rhs.nbytes)
        fvcc.set_component(rhs3, k, rhs)
    ls.rhs3 = rhs3
    return ls


def _solve_vec(ls: LinearSystem, psi: fvcc.VolumeField, solver) -> list[SolverStats]:
    """Each component of psi solved on the shared matrix with its right-hand side."""
    ctx = psi.mesh.ctx
    stats = []
    for k in range(psi.ncomp):
        b = fvcc.component(ls.rhs3, k)
        call("fv_memcpy_d2d", ctx.handle, lib.fv_linsys_rhs(ls.handle), b.ptr, b.nbytes)
        x = fvcc.component(psi.internal, k)
        stats.append(solve_cg(ls, x, solver))
        fvcc.set_component(psi.internal, k, x)
    return stats


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
    if psi.ncomp > 1:
        stats = _solve_vec(ls, psi, solver)
    else:
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
            if self.psi.ncomp > 1:
                raise ValueError("setReference applies to scalar equations")
            call("fv_set_reference", self.ls.handle, self._ref[0], self._ref[1])
        if self.psi.ncomp > 1:
            all_stats = _solve_vec(self.ls, self.psi, self.solver)
            self.psi.correct_boundary_conditions()
            print(f"[fvhip] Solving for {self.psi.name}: Final residuals "
                  f"{[s.finalResNorm for s in all_stats]} No Iterations "
                  f"{[s.numIter for s in all_stats]}")
            return all_stats
        stats = solve_cg(self.ls, self.psi.internal, self.solver)
        self.psi.correct_boundary_conditions()
        print(f"[fvhip] Solving for {self.psi.name}: Initial residual: {stats.initResNorm} "
              f"Final residual: {stats.finalResNorm} No Iterations: {stats.numIter}")
        return stats


def applyOperator(ls: LinearSystem, psi: fvcc.VolumeField) -> DeviceArray:
    """FoamAdapter::applyOperator: A.psi - b (expression.hpp:204-220)."""
    return ls.residual(psi.internal)
