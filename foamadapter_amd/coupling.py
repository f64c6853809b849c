"""Pressure-velocity coupling helpers with FoamAdapter's names
(src/algorithms/pressureVelocityCoupling.cpp:14-267) over libfvhip's
fv_compute_rau, fv_compute_hbya, fv_update_face_velocity, fv_update_velocity
and fv_constrain_hbya.

The reference passes a PDESolver<Vec3> (assembled momentum matrix with a Vec3
right-hand side).  Here the momentum matrix is a scalar LinearSystem (the
reference reads component 0 of its coefficients) and the Vec3 right-hand side
is passed alongside as a [N, 3] device array.
"""
from __future__ import annotations

from . import fvcc
from ._lib import call
from .la import LinearSystem
from .runtime import DeviceArray


def computeRAU(ls: LinearSystem) -> fvcc.VolumeField:
    """rAU = V / diag(A), boundary values extrapolated (:38-61)."""
    m = ls.mesh
    ctx = m.ctx
    rau = ctx.empty((m.n_cells,))
    raub = ctx.empty((m.n_boundary_faces,))
    call("fv_compute_rau", ls.handle, rau.ptr, raub.ptr)
    return fvcc.VolumeField(m, rau, raub, None, "rAU")


def computeRAUandHByA(ls: LinearSystem, b3: DeviceArray, U: fvcc.VolumeField
                      ) -> tuple[fvcc.VolumeField, fvcc.VolumeField]:
    """(rAU, HbyA) with HbyA = (b - sum_nb A(c,nb) U_nb) rAU / V (:63-119); the
    sum runs in ascending face order (no atomics), boundary values extrapolated."""
    m = ls.mesh
    ctx = m.ctx
    rAU = computeRAU(ls)
    h = ctx.empty((m.n_cells, 3))
    hb = ctx.empty((m.n_boundary_faces, 3))
    call("fv_compute_hbya", ls.handle, b3.ptr, U.internal.ptr, rAU.internal.ptr, h.ptr, hb.ptr)
    return rAU, fvcc.VolumeField(m, h, hb, None, "HbyA")


def constrainHbyA(U: fvcc.VolumeField, p: fvcc.VolumeField | None, hByA: fvcc.VolumeField):
    """HbyA_b = U_b on patches whose velocity condition is fixedValue (:14-36)."""
    call("fv_constrain_hbya", U.bcs.handle, U.boundary.ptr, hByA.boundary.ptr)


def updateFaceVelocity(predictedPhi: fvcc.SurfaceField, p_ls: LinearSystem, p: fvcc.VolumeField,
                       phi: fvcc.SurfaceField):
    """phi = predictedPhi - (A(N,P) p_N - A(P,N) p_P) inside, and
    predictedPhi_b - (bRhs - bMat p_P) on boundary faces (:122-195)."""
    call("fv_update_face_velocity", p_ls.handle, predictedPhi.values.ptr, p.internal.ptr,
         phi.values.ptr)


def updateVelocity(hByA: fvcc.VolumeField, rAU: fvcc.VolumeField, p: fvcc.VolumeField,
                   U: fvcc.VolumeField):
    """U = HbyA - rAU grad(p) (:197-211)."""
    m = U.mesh
    gradP = fvcc.GaussGreenGrad(m.ctx, m).grad(p)
    call("fv_update_velocity", m.ctx.handle, m.handle, hByA.internal.ptr, rAU.internal.ptr,
         gradP.ptr, U.internal.ptr)
    gradP.close()


def flux(U: fvcc.VolumeField, out: fvcc.SurfaceField | None = None) -> fvcc.SurfaceField:
    """phi = Sf . interpolate(U) with boundary Sf_b . U_b (:213-267), into out
    when given."""
    return fvcc.flux(U.mesh.ctx, U.mesh, U, out)
