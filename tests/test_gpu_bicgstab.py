"""Jacobi-preconditioned BiCGStab on non-symmetric systems (implicit advection,
test/test_advection.cpp:171-178 runs the same solver on ddt(T) + div(phi, T)).

Bars: the explicit FP64 residual ||b - A x|| / ||b|| at most the tolerance;
the solution within cond-bounded error of scipy's direct solve; repeated
solves bit-identical; iteration limit reported as FV_ENOCONVERGE.
"""
import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _advection(ctx, n=12, seed=41, scheme="upwind", dt=0.05):
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc
    poly = fa.PolyMesh.hex(n, jitter=0.15, jitter_seed=seed, perm_seed=seed)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    # solid-body rotation about the z axis through the cube centre, plus a drift
    C = m.get("C").reshape(N, 3)
    U = np.stack([-(C[:, 1] - 0.5) + 0.3, C[:, 0] - 0.5, 0.1 * np.ones(N)], axis=1)
    Cb = m.get("Cf_b").reshape(B, 3)
    Ub = np.stack([-(Cb[:, 1] - 0.5) + 0.3, Cb[:, 0] - 0.5, 0.1 * np.ones(B)], axis=1)
    phi = fvcc.flux(ctx, m, fvcc.VolumeField(m, ctx.upload(U), ctx.upload(Ub)))
    bc = fa.BoundaryConditions(m, ["zeroGradient"] * m.n_patches)
    T0 = np.exp(-40 * ((C[:, 0] - 0.3) ** 2 + (C[:, 1] - 0.5) ** 2))
    T = fvcc.VolumeField(m, ctx.upload(T0), ctx.zeros((B,)), bc, "T")
    T.correct_boundary_conditions()
    eqn = dsl.imp.ddt(T) + dsl.imp.div(phi, T, scheme=scheme)
    return m, T, eqn, dt


@pytest.mark.parametrize("scheme", ["upwind", "linear"])
def test_bicgstab_advection_steps(ctx, scheme):
    import scipy.sparse.linalg as spla
    from foamadapter_amd import dsl, la
    m, T, eqn, dt = _advection(ctx, scheme=scheme)
    solver = la.mapFvSolution({"solver": "PBiCGStab", "preconditioner": "DIC", "relTol": 1e-10,
                               "maxIter": 500})
    for step in range(3):
        print(f"[case] advection {scheme} step {step}", flush=True)
        ls = dsl.assemble(eqn, T, step * dt, dt)
        A, b = ls.scipy(), ls.rhs()
        assert abs(A - A.T).max() > 1e-6                   # non-symmetric
        x0 = T.internal.numpy()
        st = la.solve_cg(ls, T.internal, solver)
        x = T.internal.numpy()
        assert st.finalResNorm <= 1e-10 and st.trueResNorm <= 1e-9, st
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-9
        ref = spla.spsolve(A.tocsc(), b)
        assert np.max(np.abs(x - ref)) <= 1e-8 * np.max(np.abs(ref))
        # the same solve again from the same start gives the same bits
        from foamadapter_amd import _lib
        xx = m.ctx.upload(x0)
        st2 = la.solve_cg(ls, xx, solver)
        assert np.array_equal(xx.numpy(), x) and st2.numIter == st.numIter
        xx.close()
        T.correct_boundary_conditions()
        _lib.call("fv_memcpy_d2d", m.ctx.handle, T.old_time.internal.ptr, T.internal.ptr,
                  T.internal.nbytes)


def test_bicgstab_iteration_limit(ctx):
    from foamadapter_amd import FvError, _lib, dsl, la
    m, T, eqn, dt = _advection(ctx, n=10)
    ls = dsl.assemble(eqn, T, 0.0, dt)
    with pytest.raises(FvError) as e:
        la.solve_cg(ls, T.internal, {"solver": "PBiCGStab", "relTol": 1e-14, "maxIter": 2})
    assert e.value.code == _lib.FV_ENOCONVERGE
    st = la.solve_cg(ls, T.internal, {"solver": "PBiCGStab", "relTol": 1e-14, "maxIter": 2},
                     raise_on_noconverge=False)
    assert st.numIter == 2 and st.status == _lib.FV_ENOCONVERGE


def test_cg_absolute_tolerance(ctx):
    """criteria absolute_residual_norm: CG stops once ||r|| <= tolerance."""
    from foamadapter_amd import dsl, fvcc, la
    m, T, eqn, dt = _advection(ctx, n=10)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    kap = fvcc.SurfaceField(m, m.ctx.upload(np.full(F + B, 1e-2)))
    ls = dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1e-3)
    b = ls.rhs()
    tol = 1e-6 * np.linalg.norm(b)
    x = m.ctx.upload(T.internal.numpy())
    st = la.solve_cg(ls, x, la.mapFvSolution({"solver": "PCG", "tolerance": tol, "relTol": 0}))
    r = np.linalg.norm(b - ls.scipy() @ x.numpy())
    assert st.finalResNorm * np.linalg.norm(b) <= tol * 1.0001 and r <= 1.01 * tol


def _set_delta_t(dt, co, max_co, max_dt):
    """setDeltaT (src/auxiliary/setup.cpp:13-22) restated for the oracle's loop."""
    f = max_co / (co + 1e-15)
    return min(min(min(f, 1.0 + 0.1 * f), 1.2) * dt, max_dt)


@pytest.mark.parametrize("scheme,steps", [("upwind", 40), ("linear", 12)])
def test_forward_euler_advection_march(ctx, scheme, steps):
    """Scalar advection with 'forwardEuler' (test/test_advection.cpp:116-170):
    each step solves imp::ddt(T) + exp::div(phi, T) -- a diagonal system, so
    T_new = T_old - dt div(phi, T_old) -- with the reference's BiCGStab
    dictionary, phi = phi0 cos(pi (t + dt/2) / endTime), and the time step
    adapted to the Courant number as examples/scalarAdvection/scalarAdvection.cpp:
    51-93 does (computeCoNum, then setDeltaT; the step solved uses the dt read
    before the adjustment, the clock advances by the adjusted one).  The march
    matches the oracle's explicit update step by step (the reference's own bar
    is 1e-10 at the end, test_advection.cpp:168), and the sequence of time steps
    is the oracle's."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc
    n = 32
    poly = fa.PolyMesh.hex((n, n, 1), lengths=(1.0, 1.0, 1.0 / n))
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    g = orc.geometry(m.poly)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    C, Cb = g["C"], g["Cf_b"]
    pi = np.pi

    def vortex(X):  # initFields of test_advection.cpp:21-46
        x, y = X[:, 0], X[:, 1]
        return np.stack([-np.sin(2 * pi * y) * np.sin(pi * x) ** 2,
                         np.sin(2 * pi * x) * np.sin(pi * y) ** 2, np.zeros(len(x))], axis=1)

    U, Ub = vortex(C), vortex(Cb)
    T0 = np.exp(-0.5 * (((C[:, 0] - 0.5) / 0.05) ** 2 + ((C[:, 1] - 0.75) / 0.05) ** 2))
    phi0 = orc.flux(g, U, Ub)
    bc = fa.BoundaryConditions(m, ["zeroGradient"] * m.n_patches)
    T = fvcc.VolumeField(m, ctx.upload(T0), ctx.zeros((B,)), bc, "T")
    T.correct_boundary_conditions()
    solver = {"solver": "Ginkgo", "type": "solver::Bicgstab",
              "criteria": {"iteration": 500, "relative_residual_norm": 1e-14}}
    end_time, max_co, max_dt = 0.5, 0.4, 1.0
    fc = g["faceCells"]
    # GPU march
    t, dt, dts = 0.0, 1e-3, []
    phiF = fvcc.SurfaceField(m, ctx.zeros((F + B,)))
    for step in range(steps):
        c = np.cos(pi * (t + 0.5 * dt) / end_time)
        phiF = fvcc.SurfaceField(m, ctx.upload(phi0 * c))
        co = fvcc.computeCoNum(phiF, dt)
        new_dt = fvcc.setDeltaT(dt, co, max_co, max_dt)
        dsl.solve(dsl.imp.ddt(T) + dsl.exp.div(phiF, T, scheme=scheme), T, t, dt, solver=solver)
        dts.append((dt, co))
        t, dt = t + new_dt, new_dt
    got = T.numpy()
    # oracle march: explicit update with the oracle's divergence and Courant number
    Tr = T0.copy()
    t, dt, odts = 0.0, 1e-3, []
    for step in range(steps):
        c = np.cos(pi * (t + 0.5 * dt) / end_time)
        phi = phi0 * c
        co = orc.courant(g, phi, dt)
        new_dt = _set_delta_t(dt, co, max_co, max_dt)
        Tr = Tr - dt * orc.div(g, phi, Tr, Tr[fc], 1 if scheme == "upwind" else 0)
        odts.append((dt, co))
        t, dt = t + new_dt, new_dt
    assert [d for d, _ in dts] == [d for d, _ in odts]          # same time steps, bit for bit
    assert dts[-1][0] > 5 * dts[0][0]                            # the step grew (factor 1.2)
    if steps >= 40:
        assert max(co for _, co in dts) > 0.3                    # ... up to the Courant limit
    assert relmax(got, Tr) <= 1e-12, relmax(got, Tr)
    assert np.max(np.abs(got - T0)) > 1e-2                       # the field really moved
    bc.close()
