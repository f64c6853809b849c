"""FoamAdapter's neoIcoFoam time loop (examples/neoIcoFoam/neoIcoFoam.cpp:80-180):
incompressible laminar flow by the PISO algorithm on the GPU path.

Each time step, as the reference's loop:

* oldTime(U) = U; Co = computeCoNum(phi, dt), and dt from setDeltaT when the
  time step is adjusted (:82-91);
* UEqn = ddt(U) + div(phi, U) - laplacian(nu, U) as a PDESolver<Vec3>, solved
  when the momentum predictor is on, else only assembled (:94-107);
* nCorrectors times (:110-175): rAU, HbyA = computeRAUandHByA(UEqn);
  constrainHbyA(U, p, HbyA); rAUf = interpolate(rAU) (linear);
  phiHbyA = flux(HbyA); nNonOrthogonalCorrectors + 1 times
  pEqn = laplacian(rAUf, p) - div(phiHbyA), setReference when p needs one,
  solve, p.correctBoundaryConditions(), and on the last of them
  phi = phiHbyA - pEqn.flux() (updateFaceVelocity); then U = HbyA - rAU grad p
  (updateVelocity) and U.correctBoundaryConditions().

The reference's loop leaves out OpenFOAM's ddtCorr, adjustPhi and
constrainPressure (its TODOs at :124-136); so does this one.  Every operator and
solve is a libfvhip call; the equations keep their linear systems between time
steps (no allocation per step beyond the coupling helpers' results).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import coupling, dsl, fvcc
from ._lib import call


@dataclass
class PisoControls:
    """The PISO dictionary of fvSolution (OpenFOAM pisoControl): corrector
    counts, the momentum predictor switch and the pressure reference."""
    nCorrectors: int = 2
    nNonOrthogonalCorrectors: int = 0
    momentumPredictor: bool = True
    pRefCell: int = 0
    pRefValue: float = 0.0

    @classmethod
    def from_dict(cls, d: dict | None) -> "PisoControls":
        d = d or {}
        yes = ("yes", "on", "true", "1")
        return cls(int(d.get("nCorrectors", 2)), int(d.get("nNonOrthogonalCorrectors", 0)),
                   str(d.get("momentumPredictor", "yes")).lower() in yes,
                   int(d.get("pRefCell", 0)), float(d.get("pRefValue", 0.0)))


@dataclass
class StepReport:
    """What one time step did: time, step size, Courant number before the step,
    the momentum solves (one per component), the pressure solves (one per
    corrector), and the continuity error after the step (OpenFOAM
    continuityErrs: dt times the volume-weighted mean of |div phi|)."""
    time: float
    deltaT: float
    coNum: float
    U: list = field(default_factory=list)
    p: list = field(default_factory=list)
    sumLocalContErr: float = 0.0


class IcoFoam:
    """neoIcoFoam on a device mesh.

    U: Vec3 VolumeField with boundary conditions (its boundary values hold the
    fixedValue data); p: scalar VolumeField with boundary conditions; phi:
    SurfaceField (the face flux, updated in place); nu: SurfaceField
    (kinematic viscosity on every face, a calculated field as in :69-73).
    fvSchemes selects the convection scheme ({"div": "Gauss upwind"} or
    "Gauss linear"); fvSolution holds the solver dictionaries {"U": {...},
    "p": {...}} and the PISO controls {"PISO": {...}}."""

    def __init__(self, U: fvcc.VolumeField, p: fvcc.VolumeField, phi: fvcc.SurfaceField,
                 nu: fvcc.SurfaceField, deltaT: float, t0: float = 0.0,
                 fvSchemes: dict | None = None, fvSolution: dict | None = None,
                 adjustTimeStep: bool = False, maxCo: float = 1.0, maxDeltaT: float = np.inf):
        if U.ncomp != 3 or p.ncomp != 1:
            raise ValueError("IcoFoam needs a Vec3 velocity and a scalar pressure")
        sol = fvSolution or {}
        self.U, self.p, self.phi, self.nu = U, p, phi, nu
        self.mesh = m = U.mesh
        self.ctx = m.ctx
        self.t, self.deltaT = float(t0), float(deltaT)
        self.adjustTimeStep, self.maxCo, self.maxDeltaT = adjustTimeStep, maxCo, maxDeltaT
        self.piso = PisoControls.from_dict(sol.get("PISO"))
        self.solverU = sol.get("U", {"solver": "PBiCGStab", "relTol": 1e-8, "maxIter": 1000})
        self.solverP = sol.get("p", {"solver": "PCG", "relTol": 1e-8, "maxIter": 1000})
        nFB = m.n_internal_faces + m.n_boundary_faces
        if U.old_time is None:
            U.old_time = fvcc.VolumeField(m, self.ctx.empty(U.internal.shape), None)
        self.UEqn = dsl.PDESolver(dsl.imp.ddt(U) + dsl.imp.div(phi, U) - dsl.imp.laplacian(nu, U),
                                  U, self.t, self.deltaT, fvSchemes, self.solverU)
        # the pressure equation's fields, overwritten by every corrector
        self.rAUf = fvcc.SurfaceField(m, self.ctx.empty((nFB,)), "rAUf")
        self.phiHbyA = fvcc.SurfaceField(m, self.ctx.empty((nFB,)), "phiHbyA")
        self.pEqn = dsl.PDESolver(dsl.imp.laplacian(self.rAUf, p) - dsl.exp.div(self.phiHbyA),
                                  p, self.t, self.deltaT, fvSchemes, self.solverP)
        # a pressure with no fixedValue patch is fixed up to a constant: reference it
        kinds = p.bcs.kinds if p.bcs is not None else []
        self.needReference = not any(int(k) == 0 for k in kinds)
        if self.needReference and self.piso.pRefCell >= 0:
            self.pEqn.setReference(self.piso.pRefCell, self.piso.pRefValue)
        self._interp = fvcc.SurfaceInterpolation(self.ctx, m, ["linear"])
        self._div = fvcc.GaussGreenDiv(self.ctx, m)

    def step(self) -> StepReport:
        U, p, phi, ctx, m = self.U, self.p, self.phi, self.ctx, self.mesh
        call("fv_memcpy_d2d", ctx.handle, U.old_time.internal.ptr, U.internal.ptr,
             U.internal.nbytes)
        # the time advances by the step size in force, then (as the reference's
        # loop does after runTime.loop()) the Courant number of the current flux
        # sets the step size this step's equations use
        self.t += self.deltaT
        coNum = fvcc.computeCoNum(phi, self.deltaT)
        if self.adjustTimeStep:
            self.deltaT = fvcc.setDeltaT(self.deltaT, coNum, self.maxCo, self.maxDeltaT)
        rep = StepReport(self.t, self.deltaT, coNum)
        for eq in (self.UEqn, self.pEqn):
            eq.t, eq.dt = self.t, self.deltaT
        if self.piso.momentumPredictor:
            rep.U = self.UEqn.solve()
        else:
            self.UEqn.assemble()
        uls = self.UEqn.linearSystem()
        for _ in range(self.piso.nCorrectors):
            rAU, HbyA = coupling.computeRAUandHByA(uls, uls.rhs3, U)
            coupling.constrainHbyA(U, p, HbyA)
            self._interp.interpolate(rAU, self.rAUf)
            coupling.flux(HbyA, out=self.phiHbyA)
            for k in range(self.piso.nNonOrthogonalCorrectors + 1):
                rep.p.append(self.pEqn.solve())
                p.correct_boundary_conditions()
                if k == self.piso.nNonOrthogonalCorrectors:
                    coupling.updateFaceVelocity(self.phiHbyA, self.pEqn.linearSystem(), p, phi)
            coupling.updateVelocity(HbyA, rAU, p, U)
            U.correct_boundary_conditions()
            for f in (rAU, HbyA):
                f.internal.close()
                f.boundary.close()
        rep.sumLocalContErr = self.continuity_error()
        return rep

    def continuity_error(self) -> float:
        """dt * sum_c V_c |div(phi)_c| / sum_c V_c (OpenFOAM continuityErrs.H)."""
        d = self._div.div(None, self.phi).numpy()
        V = self.mesh.get("V")
        return float(self.deltaT * np.sum(V * np.abs(d)) / np.sum(V))

    def run(self, nSteps: int) -> list[StepReport]:
        return [self.step() for _ in range(nSteps)]

