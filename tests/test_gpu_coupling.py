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


# ---- the reference's pressure-velocity coupling case (test/setup_pressureVelocityCoupling:
# 3x3x3 unit cube, walls noSlip, inlet U = (1 0 0), outlet zeroGradient U and p = 0.1) ----

@pytest.fixture(scope="module")
def pv_case(ctx, fixtures_dir):
    import foamadapter_amd as fa
    from foamadapter_amd.foamio import read_boundary_conditions
    from foamadapter_amd.mesh import read_polymesh
    case = fixtures_dir / "setup_pressureVelocityCoupling"
    poly = read_polymesh(case)
    print("[case] setup_pressureVelocityCoupling 3x3x3", flush=True)
    m = fa.Mesh.from_poly(ctx, poly)
    ubc = read_boundary_conditions(case / "0" / "ofU", poly.patches, 3)
    pbc = read_boundary_conditions(case / "0" / "ofp", poly.patches, 1)
    return poly, m, orc.geometry(poly), ubc, pbc


def _velocity(ctx, m, g, ubc, scale=1.0):
    """U random per cell (test/common.hpp randomVectorField, seeded here), its
    boundary values from the case's conditions, old time zero, phi = flux(U),
    and the momentum equation ddt(U) + div(phi, U) - laplacian(nu, U) of
    test_pressureVelocityCoupling.cpp:110-115 (upwind, nu = 0.01, dt = 1)."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    names = list(ubc)
    kinds = [ubc[n][0] for n in names]
    Ub = np.concatenate([ubc[n][1] for n in names])
    U = scale * orc.uniform(3 * N, -1.0, 1.0, 77, 0).reshape(N, 3)
    fc = g["faceCells"]
    for i, n in enumerate(names):                      # zeroGradient: U_b = U_P
        if kinds[i] == "fixedGradient":
            sl = slice(g["patchOffsets"][i], g["patchOffsets"][i + 1])
            Ub[sl] = U[fc[sl]]
    bc = fa.BoundaryConditions(m, kinds)
    Uf = fvcc.VolumeField(m, ctx.upload(U), ctx.upload(Ub), bc, "U")
    Uf.old_time = fvcc.VolumeField(m, ctx.zeros((N, 3)), None)
    phi = fvcc.flux(ctx, m, Uf)
    nu = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, 0.01)))
    eqn = dsl.imp.ddt(Uf) + dsl.imp.div(phi, Uf, scheme="upwind") - dsl.imp.laplacian(nu, Uf)
    return Uf, U, Ub, phi, nu, eqn, [fa.BoundaryConditions.KINDS[k] for k in kinds]


def _oracle_momentum(g, phi, Ub, kinds):
    """The same equation in the oracle, one right-hand side per component."""
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    sp = orc.sparsity(g)
    rhs, vals = [], None
    for k in range(3):
        ls = orc.LinSys(g, sp)
        ls.ddt(np.zeros(N), 1.0)
        ls.div(phi, Ub[:, k].copy(), kinds, None, 1)
        ls.laplacian(np.full(F + B, 0.01), Ub[:, k].copy(), kinds, None, -1.0)
        rhs.append(ls.rhs)
        if vals is None:
            vals = ls.values
        assert np.array_equal(vals, ls.values)       # the components share the matrix
    return sp, vals, np.stack(rhs, axis=1)


@pytest.mark.parametrize("scale", [1.0, 2.5])
def test_pv_case_rau_hbya_constrain(ctx, pv_case, scale):
    """Sections rAU, 'rAU modified U', HbyA, constrainHbyA and 'HbyA modified U'
    of test/test_pressureVelocityCoupling.cpp:110-205 on the reference's case:
    the Vec3 momentum system (matrix and three right-hand sides), rAU, HbyA and
    the constrained boundary values bit-identical to the oracle."""
    from foamadapter_amd import coupling, dsl
    poly, m, g, ubc, pbc = pv_case
    Uf, U, Ub, phi, nu, eqn, kinds = _velocity(ctx, m, g, ubc, scale)
    pe = dsl.PDESolver(eqn, Uf, 0.0, 1.0)
    ls = pe.assemble()
    sp, vals, b3 = _oracle_momentum(g, phi.values.numpy(), Ub, kinds)
    assert np.array_equal(ls.values(), vals)
    assert np.array_equal(ls.rhs3.numpy(), b3)
    assert abs(ls.scipy() - ls.scipy().T).max() > 1e-3        # upwind: not symmetric
    rAU, HbyA = coupling.computeRAUandHByA(ls, ls.rhs3, Uf)
    rau, raub = orc.rau(g, sp, vals)
    h, hb = orc.hbya(g, sp, vals, b3, U, rau)
    assert np.array_equal(rAU.internal.numpy(), rau) and np.array_equal(rAU.boundary.numpy(), raub)
    assert np.array_equal(HbyA.internal.numpy(), h) and np.array_equal(HbyA.boundary.numpy(), hb)
    coupling.constrainHbyA(Uf, None, HbyA)
    want = orc.constrain_hbya(g, kinds, Ub, hb)
    assert np.array_equal(HbyA.boundary.numpy(), want)
    po = g["patchOffsets"]
    assert np.array_equal(want[po[1]:po[2]], np.tile([1.0, 0.0, 0.0], (po[2] - po[1], 1)))  # inlet


def test_pv_case_matrix_flux(ctx, pv_case):
    """Section 'matrix flux' (test_pressureVelocityCoupling.cpp:207-262):
    pEqn = laplacian(rAUf = 0.1, p) - div(phi), then updateFaceVelocity(phi,
    pEqn, phi0) = phi - pEqn.flux(), bit-identical to the oracle on every
    internal and boundary face."""
    import foamadapter_amd as fa
    from foamadapter_amd import coupling, dsl, fvcc
    poly, m, g, ubc, pbc = pv_case
    Uf, U, Ub, phi, nu, eqn, _ = _velocity(ctx, m, g, ubc)
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    names = list(pbc)
    pk = [pbc[n][0] for n in names]
    pb = np.concatenate([pbc[n][1] for n in names])
    p0 = orc.uniform(N, -1.0, 1.0, 78, 0)
    p = fvcc.VolumeField(m, ctx.upload(p0), ctx.upload(pb), fa.BoundaryConditions(m, pk), "p")
    p.correct_boundary_conditions()
    rauf = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, 0.1)))
    pe = dsl.PDESolver(dsl.imp.laplacian(rauf, p) - dsl.exp.div(phi), p, 0.0, 1.0)
    ls = pe.assemble()
    out = fvcc.SurfaceField(m, ctx.zeros((F + B,)))
    coupling.updateFaceVelocity(phi, ls, p, out)
    kinds = [fa.BoundaryConditions.KINDS[k] for k in pk]
    ref = orc.LinSys(g, orc.sparsity(g))
    ref.laplacian(np.full(F + B, 0.1), p.boundary.numpy(), kinds)
    ref.rhs += orc.div(g, phi.values.numpy(), np.ones(N), np.ones(B)) * g["V"]
    assert np.array_equal(ls.values(), ref.values) and np.array_equal(ls.rhs(), ref.rhs)
    want = orc.update_face_velocity(g, orc.sparsity(g), ref.values, ref.bMat, ref.bRhs,
                                    phi.values.numpy(), p0)
    assert np.array_equal(out.values.numpy(), want)


def test_pv_case_vector_momentum_solve(ctx, pv_case):
    """PDESolver<Vec3>: the momentum predictor solved component by component on
    the shared (non-symmetric) matrix with BiCGStab; each component agrees with
    scipy's direct solve of that component's system."""
    import scipy.sparse.linalg as spla
    from foamadapter_amd import dsl
    poly, m, g, ubc, pbc = pv_case
    Uf, U, Ub, phi, nu, eqn, kinds = _velocity(ctx, m, g, ubc)
    pe = dsl.PDESolver(eqn, Uf, 0.0, 1.0, solver={"solver": "PBiCGStab", "relTol": 1e-13,
                                                   "maxIter": 500})
    stats = pe.solve()
    assert len(stats) == 3 and all(s.finalResNorm <= 1e-13 for s in stats)
    A = pe.linearSystem().scipy().tocsc()
    b3 = pe.linearSystem().rhs3.numpy()
    x = Uf.internal.numpy()
    ref = np.stack([spla.spsolve(A, b3[:, k]) for k in range(3)], axis=1)
    assert np.max(np.abs(ref[:, 0])) > 0.1                 # the inlet drives the x component
    for k in range(3):                                     # (y, z: zero right-hand sides)
        assert np.max(np.abs(x[:, k] - ref[:, k])) <= 1e-10 * np.max(np.abs(ref)), k
    # outlet (zeroGradient) boundary values follow the solution
    po, fc = g["patchOffsets"], g["faceCells"]
    assert np.array_equal(Uf.boundary.numpy()[po[2]:po[3]], x[fc[po[2]:po[3]]])
