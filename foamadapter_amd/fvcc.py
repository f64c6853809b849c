"""Finite-volume fields and explicit operators with NeoN's fvcc names.

Mirrors the calls FoamAdapter's tests and benchmarks make
(test/test_operators.cpp, benchmarks/bench_explicitOperators.cpp):

    fvcc.GaussGreenGrad(exec, mesh).grad(T, coeff, out)
    fvcc.GaussGreenDiv(exec, mesh, ["linear"]).div(out, phi, T, coeff)
    fvcc.GaussGreenLaplacian(exec, mesh, ["linear", "uncorrected"]).laplacian(out, gamma, T, coeff)
    fvcc.SurfaceInterpolation(exec, mesh, ["linear"]).interpolate(T, Tf)
    fvcc.FaceNormalGradient(exec, mesh, ["uncorrected"]).faceNormalGrad(T, out)

``exec`` is a ``Context``; every call runs the HIP kernels of libfvhip.
"""
from __future__ import annotations

import numpy as np

from ._lib import FV_LINEAR, FV_UPWIND, call
from .mesh import BoundaryConditions, Mesh
from .runtime import Context, DeviceArray, prec_of


class VolumeField:
    """Cell field: internal values [N] (or [N,3]) plus boundary values [B]
    (fvcc::VolumeField: internalVector() / boundaryData().value())."""

    def __init__(self, mesh: Mesh, internal: DeviceArray, boundary: DeviceArray | None,
                 bcs: BoundaryConditions | None = None, name: str = "",
                 boundary_gradient: DeviceArray | None = None):
        self.mesh = mesh
        self.internal = internal
        self.boundary = boundary
        self.bcs = bcs
        self.name = name
        self.boundary_gradient = boundary_gradient
        self.old_time: VolumeField | None = None

    @property
    def dtype(self):
        return self.internal.dtype

    @classmethod
    def from_numpy(cls, mesh: Mesh, values: np.ndarray, boundary: np.ndarray | None = None,
                   bcs: BoundaryConditions | None = None, name: str = "", dtype=np.float64):
        ctx = mesh.ctx
        internal = ctx.upload(np.asarray(values, dtype))
        if boundary is None:
            bshape = (mesh.n_boundary_faces,) + tuple(np.shape(values)[1:])
            b = ctx.zeros(bshape, dtype)
        else:
            b = ctx.upload(np.asarray(boundary, dtype))
        return cls(mesh, internal, b, bcs, name)

    @property
    def ncomp(self) -> int:
        """1 for a scalar field, 3 for a Vec3 field ([N, 3], xyz interleaved)."""
        return 1 if len(self.internal.shape) == 1 else int(self.internal.shape[1])

    def component(self, k: int) -> "VolumeField":
        """Scalar copy of component k (internal, boundary, fixed gradient and
        old-time values) with the same boundary conditions."""
        def comp(a):
            return None if a is None else component(a, k)
        f = VolumeField(self.mesh, comp(self.internal), comp(self.boundary), self.bcs,
                        f"{self.name}[{k}]", comp(self.boundary_gradient))
        if self.old_time is not None:
            f.old_time = VolumeField(self.mesh, comp(self.old_time.internal), None)
        return f

    def correct_boundary_conditions(self):
        """fixedGradient/zeroGradient faces: Tb = T_P + g/deltaCoeffs (each
        component of a Vec3 field)."""
        if self.bcs is None or self.mesh.n_boundary_faces == 0:
            return
        if self.ncomp > 1:
            for k in range(self.ncomp):
                c = self.component(k)
                c.correct_boundary_conditions()
                set_component(self.boundary, k, c.boundary)
            return
        call("fv_correct_boundary", self.mesh.ctx.handle, self.mesh.handle, prec_of(self.dtype),
             self.internal.ptr, self.bcs.handle,
             self.boundary_gradient.ptr if self.boundary_gradient is not None else None,
             self.boundary.ptr)

    # NeoN spelling
    correctBoundaryConditions = correct_boundary_conditions

    def numpy(self) -> np.ndarray:
        return self.internal.numpy()


def component(a: DeviceArray, k: int) -> DeviceArray:
    """Component k of an interleaved [n, ncomp] device array, as a new [n] array."""
    n, nc = a.shape[0], a.shape[1]
    out = a.ctx.empty((n,), a.dtype)
    call("fv_component", a.ctx.handle, prec_of(a.dtype), n, nc, k, 0, a.ptr, out.ptr)
    return out


def set_component(a: DeviceArray, k: int, src: DeviceArray):
    """Component k of the interleaved [n, ncomp] device array a = src [n]."""
    call("fv_component", a.ctx.handle, prec_of(a.dtype), a.shape[0], a.shape[1], k, 1, src.ptr,
         a.ptr)


class SurfaceField:
    """Face field of length F+B: internal faces then boundary faces
    (readers.hpp:175-217 constructSurfaceField)."""

    def __init__(self, mesh: Mesh, values: DeviceArray, name: str = ""):
        self.mesh = mesh
        self.values = values
        self.name = name

    @property
    def dtype(self):
        return self.values.dtype

    @classmethod
    def from_numpy(cls, mesh: Mesh, values: np.ndarray, name: str = "", dtype=np.float64):
        v = np.asarray(values, dtype)
        n = mesh.n_internal_faces + mesh.n_boundary_faces
        if v.shape[0] != n:
            raise ValueError(f"surface field needs {n} values, got {v.shape[0]}")
        return cls(mesh, mesh.ctx.upload(v), name)

    def numpy(self) -> np.ndarray:
        return self.values.numpy()


def _scheme(tokens) -> int:
    toks = [tokens] if isinstance(tokens, str) else list(tokens)
    toks = [t for t in toks if t != "Gauss"]
    name = toks[0] if toks else "linear"
    if name == "linear":
        return FV_LINEAR
    if name == "upwind":
        return FV_UPWIND
    raise ValueError(f"unsupported interpolation scheme {name!r}")


def _check_corr(tokens):
    toks = [tokens] if isinstance(tokens, str) else list(tokens)
    if toks and toks[-1] not in ("uncorrected", "linear", "Gauss"):
        raise ValueError(f"unsupported snGrad scheme {toks[-1]!r} (only 'uncorrected')")


class SurfaceInterpolation:
    def __init__(self, exec: Context, mesh: Mesh, scheme=("linear",)):
        self.exec, self.mesh, self.scheme = exec, mesh, _scheme(scheme)

    def interpolate(self, T: VolumeField, Tf: SurfaceField, phi: SurfaceField | None = None):
        if self.scheme == FV_UPWIND and phi is None:
            raise ValueError("upwind interpolation needs the face flux phi")
        call("fv_interpolate", self.exec.handle, self.mesh.handle, self.scheme, prec_of(T.dtype),
             phi.values.ptr if phi is not None else None, T.internal.ptr, T.boundary.ptr,
             Tf.values.ptr)
        return Tf


class GaussGreenGrad:
    def __init__(self, exec: Context, mesh: Mesh):
        self.exec, self.mesh = exec, mesh

    def grad(self, T: VolumeField, coeff: float = 1.0, out: DeviceArray | None = None):
        if out is None:
            out = self.exec.empty((self.mesh.n_cells, 3), T.dtype)
        call("fv_grad_gauss", self.exec.handle, self.mesh.handle, prec_of(T.dtype), float(coeff),
             T.internal.ptr, T.boundary.ptr, out.ptr)
        return out


class GaussGreenDiv:
    def __init__(self, exec: Context, mesh: Mesh, scheme=("linear",)):
        self.exec, self.mesh, self.scheme = exec, mesh, _scheme(scheme)

    def div(self, out: DeviceArray | None, phi: SurfaceField, T: VolumeField | None = None,
            coeff: float = 1.0):
        """out = coeff * div(phi*T_f); T None gives surfaceIntegrate(phi)."""
        if out is None:
            out = self.exec.empty((self.mesh.n_cells,), phi.dtype)
        call("fv_div_gauss", self.exec.handle, self.mesh.handle, self.scheme, prec_of(phi.dtype),
             float(coeff), phi.values.ptr, T.internal.ptr if T is not None else None,
             T.boundary.ptr if T is not None else None, out.ptr)
        return out


class GaussGreenLaplacian:
    def __init__(self, exec: Context, mesh: Mesh, scheme=("linear", "uncorrected")):
        _check_corr(scheme)
        self.exec, self.mesh = exec, mesh

    def laplacian(self, out: DeviceArray | None, gamma: SurfaceField, T: VolumeField,
                  coeff: float = 1.0):
        if T.bcs is None:
            raise ValueError("laplacian needs the field's boundary conditions")
        if out is None:
            out = self.exec.empty((self.mesh.n_cells,), T.dtype)
        call("fv_laplacian_uncorrected", self.exec.handle, self.mesh.handle, prec_of(T.dtype),
             float(coeff), gamma.values.ptr, T.internal.ptr, T.bcs.handle, T.boundary.ptr,
             T.boundary_gradient.ptr if T.boundary_gradient is not None else None, out.ptr)
        return out


class FaceNormalGradient:
    def __init__(self, exec: Context, mesh: Mesh, scheme=("uncorrected",)):
        _check_corr(scheme)
        self.exec, self.mesh = exec, mesh

    def faceNormalGrad(self, T: VolumeField, out: SurfaceField | None = None):
        if T.bcs is None:
            raise ValueError("snGrad needs the field's boundary conditions")
        n = self.mesh.n_internal_faces + self.mesh.n_boundary_faces
        if out is None:
            out = SurfaceField(self.mesh, self.exec.empty((n,), T.dtype))
        call("fv_sngrad_uncorrected", self.exec.handle, self.mesh.handle, prec_of(T.dtype),
             T.internal.ptr, T.bcs.handle, T.boundary.ptr,
             T.boundary_gradient.ptr if T.boundary_gradient is not None else None,
             out.values.ptr)
        return out


def flux(exec: Context, mesh: Mesh, U: VolumeField, out: SurfaceField | None = None) -> SurfaceField:
    """phi = Sf . interpolate(U) (src/algorithms/pressureVelocityCoupling.cpp:242-264)."""
    n = mesh.n_internal_faces + mesh.n_boundary_faces
    if out is None:
        out = SurfaceField(mesh, exec.empty((n,), U.dtype))
    call("fv_flux", exec.handle, mesh.handle, prec_of(U.dtype), U.internal.ptr, U.boundary.ptr,
         out.values.ptr)
    return out


def computeCoNum(phi: SurfaceField, dt: float) -> float:
    """Courant number 0.5 * max_c(sum_faces |phi| / V) * dt (NeoN computeCoNum,
    examples/scalarAdvection/scalarAdvection.cpp:69; OpenFOAM CourantNo.H)."""
    import ctypes as C
    m = phi.mesh
    co = C.c_double()
    call("fv_courant_number", m.ctx.handle, m.handle, phi.values.ptr, float(dt), C.byref(co))
    return co.value


def setDeltaT(dt: float, coNum: float, maxCo: float, maxDeltaT: float) -> float:
    """New time step after setDeltaT (src/auxiliary/setup.cpp:13-22)."""
    import ctypes as C
    out = C.c_double()
    call("fv_set_delta_t", float(dt), float(coNum), float(maxCo), float(maxDeltaT), C.byref(out))
    return out.value


def write(field: VolumeField, case_dir, time: str, name: str | None = None):
    """FoamAdapter::write(volField, mesh, fieldName) (src/auxiliary/writers.cpp:66-135):
    the field as an OpenFOAM ASCII file <case_dir>/<time>/<name>, internal values in
    the caller's cell numbering (a renumbered mesh is mapped back; boundary faces
    keep their positions under renumbering)."""
    from .foamio import write_field
    m = field.mesh
    internal = field.internal
    mapped = m.to_caller(internal, "cell") if m.cell_map is not None else internal
    vals = mapped.numpy()
    if mapped is not internal:
        mapped.close()
    b = field.boundary.numpy() if field.boundary is not None else None
    patches = getattr(m, "all_patches", None) or m.patches
    return write_field(case_dir, time, name or field.name, vals, patches, b)
