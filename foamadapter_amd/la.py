"""Linear algebra handles with NeoN's la names: SparsityPattern, LinearSystem,
computeResidual, the fvSolution mapping and the Jacobi-PCG / Jacobi-BiCGStab
solves (libfvhip fv_sparsity_*, fv_linsys_*, fv_solve_cg, fv_solve_bicgstab)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import call, lib
from .mesh import Mesh
from .runtime import DeviceArray


class SparsityPattern:
    """la::SparsityPattern (expression.hpp:35; test_sparsityPattern.cpp:36-62)."""

    def __init__(self, mesh: Mesh):
        h = C.c_void_p()
        call("fv_sparsity_create", mesh.handle, C.byref(h))
        self.handle = h
        self.mesh = mesh
        s = np.zeros(2, np.int64)
        call("fv_sparsity_sizes", h, C.c_void_p(s.ctypes.data))
        self.n_rows, self.nnz = int(s[0]), int(s[1])
        mesh.ctx.adopt(self)

    @classmethod
    def readOrCreate(cls, mesh: Mesh) -> "SparsityPattern":
        sp = getattr(mesh, "_sparsity", None)
        if sp is None or sp.handle is None:
            sp = cls(mesh)
            mesh._sparsity = sp
        return sp

    def _get(self, which: int, n: int) -> np.ndarray:
        out = np.empty(n, np.int32)
        call("fv_sparsity_get", self.handle, which, C.c_void_p(out.ctypes.data))
        return out

    def rowOffs(self):
        return self._get(0, self.n_rows + 1)

    def colIdxs(self):
        return self._get(1, self.nnz)

    def diagOffset(self):
        return self._get(2, self.n_rows)

    def ownerOffset(self):
        return self._get(3, self.mesh.n_internal_faces)

    def neighbourOffset(self):
        return self._get(4, self.mesh.n_internal_faces)

    def close(self):
        if self.handle:
            lib.fv_sparsity_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LinearSystem:
    """la::LinearSystem<scalar, localIdx> in FP64 (createEmptyLinearSystem)."""

    def __init__(self, sp: SparsityPattern):
        h = C.c_void_p()
        call("fv_linsys_create", sp.handle, C.byref(h))
        self.handle = h
        self.sp = sp
        self.mesh = sp.mesh
        sp.mesh.ctx.adopt(self)

    def zero(self):
        call("fv_linsys_zero", self.handle)

    def values(self) -> np.ndarray:
        out = np.empty(self.sp.nnz)
        call("fv_memcpy_d2h", self.mesh.ctx.handle, C.c_void_p(out.ctypes.data),
             lib.fv_linsys_values(self.handle), out.nbytes)
        return out

    def rhs(self) -> np.ndarray:
        out = np.empty(self.sp.n_rows)
        call("fv_memcpy_d2h", self.mesh.ctx.handle, C.c_void_p(out.ctypes.data),
             lib.fv_linsys_rhs(self.handle), out.nbytes)
        return out

    def diag(self) -> np.ndarray:
        """FoamAdapter::diag (expression.hpp:181-201)."""
        return self.values()[self.sp.rowOffs()[:-1] + self.sp.diagOffset()]

    def scipy(self):
        import scipy.sparse as sps
        n = self.sp.n_rows
        return sps.csr_matrix((self.values(), self.sp.colIdxs(), self.sp.rowOffs()), shape=(n, n))

    def residual(self, x: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
        """la::computeResidual: out = A.x - b (expression.hpp:204-220)."""
        if out is None:
            out = self.mesh.ctx.empty((self.sp.n_rows,))
        call("fv_residual", self.handle, x.ptr, out.ptr)
        return out

    def close(self):
        if self.handle:
            lib.fv_linsys_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class SolverStats:
    """la::SolverStats (expression.hpp:143-158) plus the refinement count and
    the explicit FP64 residual at exit."""
    initResNorm: float
    finalResNorm: float
    numIter: int
    outerIter: int
    trueResNorm: float
    status: int


# OpenFOAM solver / preconditioner names -> (Ginkgo type, solver run here)
# (src/compatibility/fvSolution.cpp:19-48: updateSolver)
SOLVER_MAP = {
    "PCG": ("solver::Cg", "cg"),
    "PBiCG": ("solver::Bicg", None),
    "PBiCGStab": ("solver::Bicgstab", "bicgstab"),
    "smoothSolver": ("solver::Bicgstab", "bicgstab"),
    "GAMG": ("solver::Multigrid", None),
}
# fvSolution.cpp:50-104: updatePreconditioner
PRECONDITIONER_MAP = {
    "DIC": {"type": "preconditioner::Jacobi", "max_block_size": 1},
    "DILU": {"type": "preconditioner::Ilu", "reverse_apply": False,
             "factorization": {"type": "factorization::ParIlu"}},
}


def mapFvSolution(solver: dict) -> dict:
    """FoamAdapter::mapFvSolution (src/compatibility/fvSolution.cpp:148-159): map an
    OpenFOAM fvSolution solver dictionary to the NeoN/Ginkgo form (solver "Ginkgo",
    type, preconditioner dictionary, criteria {iteration, relative_residual_norm,
    absolute_residual_norm}); a dictionary with configFile is returned unchanged."""
    d = dict(solver)
    if "configFile" in d:
        return d
    name = d.get("solver")
    if name in SOLVER_MAP:
        d["solver"] = "Ginkgo"
        d["type"] = SOLVER_MAP[name][0]
    if "preconditioner" not in d:
        d["preconditioner"] = dict(PRECONDITIONER_MAP["DIC"])
    d.pop("smoother", None)
    pre = d["preconditioner"]
    if isinstance(pre, dict):
        if isinstance(pre.get("type"), dict):
            raise ValueError("a multigrid (GAMG) preconditioner cannot be mapped: choose DIC "
                             "(point Jacobi) for this solver")
    elif pre in PRECONDITIONER_MAP:
        d["preconditioner"] = dict(PRECONDITIONER_MAP[pre])
    crit = dict(d.get("criteria", {}))
    crit["iteration"] = 1000
    if "relTol" in d:
        crit["relative_residual_norm"] = d.pop("relTol")
    if "maxIter" in d:
        crit["iteration"] = d.pop("maxIter")
    if "tolerance" in d:
        crit["absolute_residual_norm"] = d.pop("tolerance")
    d["criteria"] = crit
    return d


def _method(d: dict) -> str:
    """The solver that runs for a (mapped or plain) dictionary: 'cg' or 'bicgstab'."""
    t = d.get("type")
    name = d.get("solver", "PCG")
    if t is None and name in SOLVER_MAP:
        t = SOLVER_MAP[name][0]
    if t in (None, "solver::Cg") or name == "CG":
        return "cg"
    if t == "solver::Bicgstab" or name in ("BiCGStab", "Bicgstab"):
        return "bicgstab"
    raise ValueError(f"solver {t or name!r} is not supported (CG and BiCGStab only)")


def solver_opts(solver: dict | None) -> _lib.FvSolveOpts:
    """Map an fvSolution-style solver dictionary (src/compatibility/fvSolution.cpp:106-159:
    relTol -> relative_residual_norm, maxIter -> iteration, default 1000, tolerance ->
    absolute_residual_norm), or its mapped Ginkgo form, to solver options.
    'precision': 'mixed' selects FP32 CG inside FP64 iterative refinement;
    'graph': False launches a repeated CG solve kernel by kernel instead of
    replaying it as a captured hipGraph (the same bits)."""
    d = dict(solver or {})
    crit = d.get("criteria", {})
    o = _lib.FvSolveOpts()
    o.rel_tol = float(d.get("relTol", crit.get("relative_residual_norm", d.get("rel_tol", 1e-6))))
    o.max_iter = int(d.get("maxIter", crit.get("iteration", d.get("maxIters", d.get("max_iter", 1000)))))
    o.abs_tol = float(d.get("tolerance", crit.get("absolute_residual_norm", 0.0)))
    mode = d.get("precision", "double")
    if mode not in ("double", "mixed"):
        raise ValueError(f"unknown precision {mode!r}")
    o.mode = _lib.FV_SOLVE_F32_IR if mode == "mixed" else _lib.FV_SOLVE_F64
    # mixed precision: an FP64 residual replacement ends each FP32 chunk of at
    # most checkEvery iterations, or earlier once the FP32 residual fell by
    # innerRelTol; FP64: the host polls the stop flag every checkEvery iterations
    o.inner_rel_tol = float(d.get("innerRelTol", 0.0))  # 0: the library's default
    o.max_outer = int(d.get("maxOuter", 0))
    o.check_every = int(d.get("checkEvery", 32))
    # a repeated CG solve is replayed as a captured hipGraph unless 'graph' is False
    o.eager = 0 if d.get("graph", True) else 1
    _method(d)
    pre = d.get("preconditioner", "DIC")
    if isinstance(pre, dict):
        pre = pre.get("type")
    if pre not in ("DIC", "Jacobi", "diagonal", "preconditioner::Jacobi"):
        raise ValueError(f"preconditioner {pre!r} is not supported (point Jacobi only)")
    return o


class Batch:
    """Many independent systems solved in one launch sequence (fv_batch_*)."""

    def __init__(self, ctx, precision: str = "mixed", tile_rows: int = 1024, chunk: int = 16):
        mode = _lib.FV_SOLVE_F32_IR if precision == "mixed" else _lib.FV_SOLVE_F64
        h = C.c_void_p()
        call("fv_batch_create", ctx.handle, mode, int(tile_rows), int(chunk), C.byref(h))
        self.handle = h
        self.ctx = ctx
        self.members = []
        ctx.adopt(self)

    def add(self, ls: LinearSystem, x: DeviceArray, xold: DeviceArray | None = None) -> int:
        i = C.c_int()
        call("fv_batch_add", self.handle, ls.handle, x.ptr, xold.ptr if xold is not None else None,
             C.byref(i))
        self.members.append((ls, x, xold))
        return i.value

    def clear(self):
        call("fv_batch_clear", self.handle)
        self.members = []

    def set_shape(self, tile_rows: int, chunk: int):
        call("fv_batch_set_shape", self.handle, int(tile_rows), int(chunk))

    def timing(self) -> dict:
        """GPU milliseconds and counts of the chunks and batched starts so far
        (HIP events on the batch's stream)."""
        ms, n = np.zeros(2), np.zeros(2, np.int64)
        call("fv_batch_timing", self.handle, C.c_void_p(ms.ctypes.data), C.c_void_p(n.ctypes.data))
        return {"chunk_ms": float(ms[0]), "start_ms": float(ms[1]), "chunks": int(n[0]),
                "starts": int(n[1])}

    def autotune(self, solver: dict | None = None, cache_path: str | None = None) -> tuple[int, int]:
        """Pick tile rows and chunk length by timing solves (cached per size bucket)."""
        o = solver_opts(solver)
        t, c = C.c_int(), C.c_int()
        call("fv_batch_autotune", self.handle, C.byref(o),
             cache_path.encode() if cache_path else None, C.byref(t), C.byref(c))
        return t.value, c.value

    def solve(self, solver: dict | None = None, use_graph: bool = True,
              raise_on_noconverge: bool = True) -> list[SolverStats]:
        o = solver_opts(solver)
        n = len(self.members)
        st = (_lib.FvSolverStats * max(n, 1))()
        rc = lib.fv_batch_solve(self.handle, C.byref(o), 1 if use_graph else 0, st)
        if rc not in (_lib.FV_OK, _lib.FV_ENOCONVERGE) or (rc != _lib.FV_OK and raise_on_noconverge):
            raise _lib.FvError(rc, _lib.last_error())
        return [SolverStats(s.init_res, s.final_res, s.iters, s.outer_iters, s.true_res, rc)
                for s in st[:n]]

    def close(self):
        if self.handle:
            lib.fv_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HeatAssembly:
    """ddt(T) + coeff*laplacian(gamma, T) (Euler implicit) of many systems in one
    launch (fv_assemble_ddt_laplacian_many), each bit-identical to its own
    fv_assemble_ddt_laplacian.  items: (LinearSystem, T_old, dt, gamma, bc, T_b,
    T_grad or None) per system; the argument tables are built once, so a time
    step costs one call."""

    def __init__(self, items):
        n = len(items)
        self.n = n
        self._keep = list(items)        # the device arrays stay alive with the tables
        vp = lambda xs: (C.c_void_p * max(n, 1))(*[C.c_void_p(_ptr(x)) for x in xs])  # noqa: E731
        self.ls = vp([it[0].handle for it in items])
        self.Told = vp([it[1].ptr for it in items])
        self.dt = (C.c_double * max(n, 1))(*[float(it[2]) for it in items])
        self.gamma = vp([it[3].ptr for it in items])
        self.bc = vp([it[4].handle for it in items])
        self.Tb = vp([it[5].ptr for it in items])
        self.Tg = vp([it[6].ptr if it[6] is not None else None for it in items])

    def assemble(self, coeff: float = -1.0):
        call("fv_assemble_ddt_laplacian_many", self.n, self.ls, self.Told, self.dt, self.gamma,
             self.bc, self.Tb, self.Tg, float(coeff))


def _ptr(x):
    if x is None:
        return None
    return x.value if isinstance(x, C.c_void_p) else int(x)


def solve_cg(ls: LinearSystem, x: DeviceArray, solver: dict | None = None,
             raise_on_noconverge: bool = True) -> SolverStats:
    """Solve with the method the dictionary names (Jacobi-PCG by default, Jacobi-
    BiCGStab for PBiCGStab / smoothSolver / solver::Bicgstab)."""
    o = solver_opts(solver)
    st = _lib.FvSolverStats()
    fn = lib.fv_solve_bicgstab if _method(dict(solver or {})) == "bicgstab" else lib.fv_solve_cg
    rc = fn(ls.handle, x.ptr, C.byref(o), C.byref(st))
    if rc not in (_lib.FV_OK, _lib.FV_ENOCONVERGE) or (rc != _lib.FV_OK and raise_on_noconverge):
        raise _lib.FvError(rc, _lib.last_error())
    return SolverStats(st.init_res, st.final_res, st.iters, st.outer_iters, st.true_res, rc)
