"""neoIcoFoam (examples/neoIcoFoam/neoIcoFoam.cpp:80-180) on the GPU path against
the same loop composed from the oracle: a lid-driven cavity on a jittered 8^3
mesh, four PISO time steps (momentum predictor, two pressure correctors).

Every GPU operator of the loop is bit-identical to its oracle counterpart on
the same inputs (tests/test_gpu_coupling.py, test_gpu_implicit.py); here the
solves (BiCGStab for the momentum components, CG for the pressure, both to a
1e-13 relative residual) stand against scipy's direct solves, so U, p and phi
are compared within 1e-8 after every step.  Invariants: the corrected flux is
divergence-free to the solver tolerance (continuity error), the lid drives the
flow, and the walls hold the no-slip values.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle_cpu as orc

pytestmark = pytest.mark.gpu

N_EDGE, DT, NU, STEPS, NCORR = 8, 0.05, 0.01, 4, 2
LID = 3        # patch y1 of PolyMesh.hex: the lid moves with U = (1, 0, 0)


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def _cavity(g):
    B = len(g["faceCells"])
    po = g["patchOffsets"]
    Ub = np.zeros((B, 3))
    Ub[po[LID]:po[LID + 1], 0] = 1.0
    return Ub


def _oracle_loop(g, Ub, steps=STEPS, ncorr=NCORR, dt=DT, nu_=NU, scheme=1, predictor=True,
                 uk=None, pk=None, U0=None):
    """The neoIcoFoam step restated with the oracle: Euler ddt, convection
    (scheme 1 upwind, 0 linear), laplacian (walls fixedValue unless uk says
    otherwise), per-component direct solves on the shared momentum matrix (or,
    without the momentum predictor, the matrix only), then ncorr times
    rAU/HbyA, constrainHbyA, rAUf, phiHbyA, the pressure equation (pk, by
    default zeroGradient walls; reference cell 0 at 0) solved directly,
    updateFaceVelocity and updateVelocity."""
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    nP = len(g["patchOffsets"]) - 1
    sp = orc.sparsity(g)
    uk = [0] * nP if uk is None else uk
    pk = [1] * nP if pk is None else pk
    nu = np.full(F + B, nu_)
    U = np.zeros((N, 3)) if U0 is None else U0.copy()
    p, pb = np.zeros(N), np.zeros(B)
    phi = orc.flux(g, U, Ub)
    out = []
    for _ in range(steps):
        Uold = U.copy()
        rhs = []
        for k in range(3):
            ls = orc.LinSys(g, sp)
            ls.ddt(Uold[:, k].copy(), dt)
            ls.div(phi, Ub[:, k].copy(), uk, None, scheme)
            ls.laplacian(nu, Ub[:, k].copy(), uk, None, -1.0)
            rhs.append(ls.rhs)
            vals = ls.values
        if predictor:
            A = ls.scipy().tocsc()
            U = np.stack([spla.spsolve(A, r) for r in rhs], axis=1)
        b3 = np.stack(rhs, axis=1)
        for _ in range(ncorr):
            rau, raub = orc.rau(g, sp, vals)
            h, hb = orc.hbya(g, sp, vals, b3, U, rau)
            hb = orc.constrain_hbya(g, uk, Ub, hb)
            rauf = orc.interpolate(g, rau, raub, 0)
            phiH = orc.flux(g, h, hb)
            pl = orc.LinSys(g, sp)
            pl.laplacian(rauf, pb, pk)
            pl.rhs += orc.div(g, phiH, np.ones(N), np.ones(B)) * g["V"]
            pl.set_reference(0, 0.0)
            p = spla.spsolve(pl.scipy().tocsc(), pl.rhs)
            pb = orc.correct_boundary(g, p, pb, pk)
            phi = orc.update_face_velocity(g, sp, pl.values, pl.bMat, pl.bRhs, phiH, p)
            U = orc.update_velocity(h, rau, orc.grad(g, p, pb))
        out.append((U.copy(), p.copy(), phi.copy()))
    return out


def test_icofoam_cavity_matches_oracle_loop(ctx):
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    from foamadapter_amd.piso import IcoFoam
    print(f"[case] icoFoam cavity {N_EDGE}^3 jittered, {STEPS} steps", flush=True)
    poly = fa.PolyMesh.hex(N_EDGE, jitter=0.15, jitter_seed=91, perm_seed=92)
    m = fa.Mesh.from_poly(ctx, poly)
    g = orc.geometry(poly)
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    Ub = _cavity(g)
    U = fvcc.VolumeField(m, ctx.zeros((N, 3)), ctx.upload(Ub),
                         fa.BoundaryConditions(m, ["fixedValue"] * 6), "U")
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)),
                         fa.BoundaryConditions(m, ["zeroGradient"] * 6), "p")
    phi = fvcc.flux(ctx, m, U)
    nu = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, NU)), "nu")
    solver = IcoFoam(U, p, phi, nu, DT, fvSchemes={"div": "Gauss upwind"},
                     fvSolution={"U": {"solver": "PBiCGStab", "relTol": 1e-13, "maxIter": 1000},
                                 "p": {"solver": "PCG", "relTol": 1e-13, "maxIter": 2000},
                                 "PISO": {"nCorrectors": NCORR, "pRefCell": 0, "pRefValue": 0}})
    assert solver.needReference
    ref = _oracle_loop(g, Ub)
    for step, (Ur, pr, phir) in enumerate(ref):
        rep = solver.step()
        print(f"[case] step {step}: Co {rep.coNum:.3f} contErr {rep.sumLocalContErr:.2e} "
              f"p iters {[s.numIter for s in rep.p]}", flush=True)
        assert abs(rep.time - DT * (step + 1)) < 1e-12 and rep.deltaT == DT
        assert len(rep.U) == 3 and len(rep.p) == NCORR
        assert relmax(U.numpy(), Ur) < 1e-8, (step, relmax(U.numpy(), Ur))
        assert relmax(p.numpy(), pr) < 1e-8, (step, relmax(p.numpy(), pr))
        assert relmax(phi.numpy(), phir) < 1e-8, (step, relmax(phi.numpy(), phir))
        assert rep.sumLocalContErr < 1e-10
        assert np.array_equal(U.boundary.numpy(), Ub)          # no-slip walls and the lid
    assert np.max(U.numpy()[:, 0]) > 0.05                      # the lid drives the flow
    assert 0.0 < rep.coNum < 1.0


def test_icofoam_adjusts_the_time_step(ctx):
    """adjustTimeStep: Co from the current flux sets the next step size by
    setDeltaT (src/auxiliary/setup.cpp:13-22); the time advances by the step in
    force when the step begins, as runTime.loop() does."""
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    from foamadapter_amd.piso import IcoFoam
    poly = fa.PolyMesh.hex(6, jitter=0.1, jitter_seed=93, perm_seed=94)
    m = fa.Mesh.from_poly(ctx, poly)
    g = orc.geometry(poly)
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    U = fvcc.VolumeField(m, ctx.zeros((N, 3)), ctx.upload(_cavity(g)),
                         fa.BoundaryConditions(m, ["fixedValue"] * 6), "U")
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)),
                         fa.BoundaryConditions(m, ["zeroGradient"] * 6), "p")
    phi = fvcc.flux(ctx, m, U)
    nu = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, NU)), "nu")
    solver = IcoFoam(U, p, phi, nu, 0.01, adjustTimeStep=True, maxCo=0.5, maxDeltaT=0.2,
                     fvSchemes={"div": "Gauss upwind"})
    t, dt = 0.0, 0.01
    for _ in range(4):
        co = orc.courant(g, phi.numpy(), dt)
        rep = solver.step()
        t += dt
        want = fvcc.setDeltaT(dt, co, 0.5, 0.2)
        assert abs(rep.coNum - co) <= 1e-14 * max(co, 1.0)
        assert abs(rep.time - t) < 1e-14 and rep.deltaT == want
        dt = want
    assert solver.deltaT > 0.01                                 # the step grew (Co < 0.5)
