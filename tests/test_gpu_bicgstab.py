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
