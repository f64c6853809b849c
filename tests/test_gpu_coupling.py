"""GPU parity of the pressure-velocity coupling helpers
(src/algorithms/pressureVelocityCoupling.cpp:14-211) against the CPU oracle.

Bars: rAU, HbyA, the boundary coefficients of the laplacian assembly, the
updated face flux and the updated velocity bit-identical to the oracle (same
serial face order, no FMA); constrainHbyA exact; and, through the whole GPU path
(assemble, solve, update), a corrected flux that is divergence-free to the
solver tolerance.
"""
import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(ctx):
    import foamadapter_amd as fa
    poly = fa.PolyMesh.hex(9, jitter=0.2, jitter_seed=31, perm_seed=32)
    print("[case] hex9 jittered, permuted", flush=True)
    return poly, fa.Mesh.from_poly(ctx, poly), orc.geometry(poly)


def _momentum(ctx, m, g, seed=5):
    """ddt + upwind div (asymmetric) - laplacian, on the GPU and in the oracle."""
    import foamadapter_amd as fa
    from foamadapter_amd import _lib, la
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    kinds = [(0, 1, 2)[i % 3] for i in range(m.n_patches)]
    d = dict(Told=orc.uniform(N, -1, 1, seed, 0), phi=orc.uniform(F + B, -1, 1, seed, 1),
             gamma=orc.uniform(F + B, 0.5, 2, seed, 2), Tb=orc.uniform(B, -1, 1, seed, 3),
             Tg=orc.uniform(B, -1, 1, seed, 4))
    D = {k: ctx.upload(v) for k, v in d.items()}
    bc = fa.BoundaryConditions(m, kinds)
    ls = la.LinearSystem(la.SparsityPattern.readOrCreate(m))
    _lib.call("fv_assemble_ddt_euler", ls.handle, D["Told"].ptr, 1e-2, 1.0)
    _lib.call("fv_assemble_div", ls.handle, 1, D["phi"].ptr, bc.handle, D["Tb"].ptr, D["Tg"].ptr,
              1.0)
    _lib.call("fv_assemble_laplacian", ls.handle, D["gamma"].ptr, bc.handle, D["Tb"].ptr,
              D["Tg"].ptr, -1.0)
    ref = orc.LinSys(g, orc.sparsity(g))
    ref.ddt(d["Told"], 1e-2)
    ref.div(d["phi"], d["Tb"], kinds, d["Tg"], 1, 1.0)
    ref.laplacian(d["gamma"], d["Tb"], kinds, d["Tg"], -1.0)
    return ls, ref, bc, kinds, D


def _bcoeffs(ctx, ls, B):
    import ctypes as C
    from foamadapter_amd import _lib
    pm, pr = C.c_void_p(), C.c_void_p()
    _lib.call("fv_linsys_boundary_coeffs", ls.handle, C.byref(pm), C.byref(pr))
    out = []
    for ptr in (pm, pr):
        h = np.empty(B)
        _lib.call("fv_memcpy_d2h", ctx.handle, C.c_void_p(h.ctypes.data), ptr, h.nbytes)
        out.append(h)
    return out


def test_boundary_coefficients_and_rau_hbya_bitwise(ctx, case):
    from foamadapter_amd import coupling, fvcc
    poly, m, g = case
    N, B = len(g["V"]), len(g["faceCells"])
    ls, ref, bc, kinds, D = _momentum(ctx, m, g)
    assert np.array_equal(ls.values(), ref.values) and np.array_equal(ls.rhs(), ref.rhs)
    bm, br = _bcoeffs(ctx, ls, B)
    assert np.array_equal(bm, ref.bMat) and np.array_equal(br, ref.bRhs)
    U = orc.uniform(3 * N, -1, 1, 8, 0).reshape(N, 3)
    b3 = orc.uniform(3 * N, -1, 1, 8, 1).reshape(N, 3)
    Uf = fvcc.VolumeField(m, ctx.upload(U), ctx.zeros((B, 3)))
    rAU, HbyA = coupling.computeRAUandHByA(ls, ctx.upload(b3), Uf)
    rau_ref, raub_ref = orc.rau(g, orc.sparsity(g), ref.values)
    h_ref, hb_ref = orc.hbya(g, orc.sparsity(g), ref.values, b3, U, rau_ref)
    assert np.array_equal(rAU.internal.numpy(), rau_ref)
    assert np.array_equal(rAU.boundary.numpy(), raub_ref)
    assert np.array_equal(HbyA.internal.numpy(), h_ref)
    assert np.array_equal(HbyA.boundary.numpy(), hb_ref)
    # constrainHbyA: fixedValue patches take U_b
    Ub = orc.uniform(3 * B, -1, 1, 8, 2).reshape(B, 3)
    import foamadapter_amd as fa
    ubc = fa.BoundaryConditions(m, kinds)
    Uf2 = fvcc.VolumeField(m, ctx.upload(U), ctx.upload(Ub), ubc)
    coupling.constrainHbyA(Uf2, None, HbyA)
    assert np.array_equal(HbyA.boundary.numpy(), orc.constrain_hbya(g, kinds, Ub, hb_ref))
    # updateVelocity: U = HbyA - rAU grad(p)
    p = orc.uniform(N, -1, 1, 8, 3)
    pb = orc.uniform(B, -1, 1, 8, 4)
    pf = fvcc.VolumeField(m, ctx.upload(p), ctx.upload(pb))
    coupling.updateVelocity(HbyA, rAU, pf, Uf)
    u_ref = orc.update_velocity(h_ref, rau_ref, orc.grad(g, p, pb))
    assert np.array_equal(Uf.internal.numpy(), u_ref)
    ls.close()


def test_corrected_flux_divergence_free_on_gpu(ctx, case):
    """phiHbyA -> solve laplacian(rAU_f, p) = div(phiHbyA) -> updateFaceVelocity:
    the corrected flux is divergence-free (GPU assembly, FP64 CG, GPU update),
    and the update is bitwise the oracle's for the same p."""
    import foamadapter_amd as fa
    from foamadapter_amd import coupling, dsl, fvcc
    poly, m, g = case
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    fc = g["faceCells"]
    rau = orc.uniform(N, 0.5, 2.0, 9, 0)
    rauf = orc.interpolate(g, rau, rau[fc])
    U = orc.uniform(3 * N, -1, 1, 9, 1).reshape(N, 3)
    Ub = orc.uniform(3 * B, -1, 1, 9, 2).reshape(B, 3)
    Uf = fvcc.VolumeField(m, ctx.upload(U), ctx.upload(Ub))
    phiH = coupling.flux(Uf)
    kinds = ["fixedValue"] + ["zeroGradient"] * (m.n_patches - 1)
    pbc = fa.BoundaryConditions(m, kinds)
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)), pbc, "p")
    gam = fvcc.SurfaceField(m, ctx.upload(rauf))
    pe = dsl.PDESolver(dsl.imp.laplacian(gam, p) - dsl.exp.div(phiH), p, 0.0, 1.0,
                       solver={"relTol": 1e-12, "maxIter": 20000})
    st = pe.solve()
    assert st.finalResNorm <= 1e-12
    phi = fvcc.SurfaceField(m, ctx.empty((F + B,)))
    coupling.updateFaceVelocity(phiH, pe.linearSystem(), p, phi)
    ph = phi.values.numpy()
    ones_c, ones_b = np.ones(N), np.ones(B)
    d = orc.div(g, phiH.values.numpy(), ones_c, ones_b)
    cont = orc.div(g, ph, ones_c, ones_b)
    assert np.max(np.abs(cont)) <= 1e-8 * np.max(np.abs(d)), np.max(np.abs(cont))
    ls = pe.linearSystem()
    bm, br = _bcoeffs(ctx, ls, B)
    ref = orc.update_face_velocity(g, orc.sparsity(g), ls.values(), bm, br,
                                   phiH.values.numpy(), p.internal.numpy())
    assert np.array_equal(ph, ref)


def test_courant_number_matches_oracle(ctx, case):
    from foamadapter_amd import fvcc
    poly, m, g = case
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    phi = orc.uniform(F + B, -1, 1, 12, 0)
    co = fvcc.computeCoNum(fvcc.SurfaceField(m, ctx.upload(phi)), 1e-3)
    assert co == orc.courant(g, phi, 1e-3)
