"""GPU parity of the implicit path: sparsity pattern, assembly (ddt, laplacian,
div, source, explicit terms, SetReference), A.x - b, and Jacobi-PCG (FP64 and
FP32-inner/FP64-refined) against the CPU oracle and scipy.

Bars: sparsity arrays exact; assembled values/rhs and residuals bit-identical to
the oracle (same serial face order, no FMA); CG solutions within the solver
tolerance (x differences bounded by cond(A)*tol; checked against the oracle CG and
scipy's direct solve); refined mixed-precision solves reach the FP64 residual
tolerance and agree with the FP64 solution to 1e-6 relative.
"""
import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def meshes(ctx, fixtures_dir):
    import foamadapter_amd as fa
    from skewed_meshes import refined_block
    polys = {
        "setup_stencil3D": fa.read_polymesh(fixtures_dir / "setup_stencil3D"),
        "setup_operator": fa.read_polymesh(fixtures_dir / "setup_operator"),
        "hex10_jit_perm": fa.PolyMesh.hex(10, jitter=0.2, jitter_seed=21, perm_seed=22),
        "refined_block": refined_block(3),   # polyhedral, ten-entry rows
    }
    out = {}
    for k, p in polys.items():
        print(f"[case] {k}", flush=True)
        out[k] = (p, fa.Mesh.from_poly(ctx, p), orc.geometry(p))
    return out


def test_sparsity_matches_oracle(meshes):
    from foamadapter_amd import la
    for name, (p, m, g) in meshes.items():
        sp = la.SparsityPattern(m)
        ref = orc.sparsity(g)
        assert np.array_equal(sp.rowOffs(), ref["rowOffs"]), name
        assert np.array_equal(sp.colIdxs(), ref["colIdxs"]), name
        assert np.array_equal(sp.diagOffset(), ref["diagOffset"]), name
        assert np.array_equal(sp.ownerOffset(), ref["ownerOffset"]), name
        assert np.array_equal(sp.neighbourOffset(), ref["neighbourOffset"]), name
        sp.close()


def test_wide_rows_solve(ctx, meshes):
    """Rows wider than the SpMV's group of eight entries loaded together (the
    refined block's coarse cells have ten): FP64 CG, mixed CG and BiCGStab agree
    with scipy's direct solve."""
    import scipy.sparse.linalg as spla
    from foamadapter_amd import dsl, la
    p, m, g = meshes["refined_block"]
    T, kap, T0, Tb = _heat(ctx, m, g, dt=1e-2, kappa=1.0)
    ls = dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1e-2)
    A, b = ls.scipy(), ls.rhs()
    assert np.diff(A.indptr).max() == 10
    ref = spla.spsolve(A.tocsc(), b)
    for solver in ({"relTol": 1e-12, "maxIter": 500},
                   {"relTol": 1e-12, "maxIter": 500, "precision": "mixed"},
                   la.mapFvSolution({"solver": "PBiCGStab", "preconditioner": "DIC",
                                     "relTol": 1e-12, "maxIter": 500})):
        x = ctx.upload(T0)
        st = la.solve_cg(ls, x, solver)
        assert st.finalResNorm <= 1e-12, (solver, st)
        assert relmax(x.numpy(), ref) < 1e-10, (solver, relmax(x.numpy(), ref))
        x.close()


def _inputs(g, seed=3):
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    return dict(Told=orc.uniform(N, 300, 400, seed, 0), gamma=orc.uniform(F + B, 0.5, 2, seed, 1),
                phi=orc.uniform(F + B, -1, 1, seed, 2), Tb=orc.uniform(B, 300, 400, seed, 3),
                Tg=orc.uniform(B, -1, 1, seed, 4), cf=orc.uniform(N, 1, 2, seed, 5),
                E=orc.uniform(N, -1, 1, seed, 6), x=orc.uniform(N, 300, 400, seed, 7))


def test_assembly_bitwise(ctx, meshes):
    import foamadapter_amd as fa
    from foamadapter_amd import _lib, la
    for name, (p, m, g) in meshes.items():
        d = _inputs(g)
        kinds = [(0, 1, 2)[i % 3] for i in range(m.n_patches)]
        bc = fa.BoundaryConditions(m, kinds)
        sp = la.SparsityPattern.readOrCreate(m)
        ls = la.LinearSystem(sp)
        D = {k: ctx.upload(v) for k, v in d.items()}
        # the heat-transfer sum plus every other operator, in this order
        _lib.call("fv_assemble_ddt_euler", ls.handle, D["Told"].ptr, 1e-3, 1.0)
        _lib.call("fv_assemble_laplacian", ls.handle, D["gamma"].ptr, bc.handle, D["Tb"].ptr,
                  D["Tg"].ptr, -1.0)
        _lib.call("fv_assemble_div", ls.handle, 0, D["phi"].ptr, bc.handle, D["Tb"].ptr,
                  D["Tg"].ptr, 0.5)
        _lib.call("fv_assemble_div", ls.handle, 1, D["phi"].ptr, bc.handle, D["Tb"].ptr,
                  D["Tg"].ptr, 0.25)
        _lib.call("fv_assemble_source", ls.handle, D["cf"].ptr, 2.0)
        _lib.call("fv_assemble_explicit", ls.handle, D["E"].ptr, 1.5)
        _lib.call("fv_set_reference", ls.handle, 1, 3.5)
        ref = orc.LinSys(g, orc.sparsity(g))
        ref.ddt(d["Told"], 1e-3, 1.0)
        ref.laplacian(d["gamma"], d["Tb"], kinds, d["Tg"], -1.0)
        ref.div(d["phi"], d["Tb"], kinds, d["Tg"], 0, 0.5)
        ref.div(d["phi"], d["Tb"], kinds, d["Tg"], 1, 0.25)
        ref.source(d["cf"], 2.0)
        ref.rhs -= 1.5 * (d["E"] * g["V"])
        ref.set_reference(1, 3.5)
        assert np.array_equal(ls.values(), ref.values), (name, relmax(ls.values(), ref.values))
        assert np.array_equal(ls.rhs(), ref.rhs), (name, relmax(ls.rhs(), ref.rhs))
        res = ls.residual(D["x"]).numpy()
        assert np.array_equal(res, ref.residual(d["x"])), name
        ls.close()
        bc.close()


def _heat(ctx, m, g, dt=1e-3, kappa=1e-2, seed=1001):
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    T0 = orc.uniform(N, 300.0, 400.0, seed, 0)
    po = m.get("patchOffsets")
    Tb = np.full(B, 350.0)
    Tb[po[0]:po[1]] = 300.0
    Tb[po[1]:po[2]] = 400.0
    bc = fa.BoundaryConditions(m, ["fixedValue"] * m.n_patches)
    T = fvcc.VolumeField(m, ctx.upload(T0), ctx.upload(Tb), bc, "T")
    kap = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, kappa)))
    return T, kap, T0, Tb


@pytest.mark.parametrize("n,precision", [(16, "double"), (32, "double"), (32, "mixed")])
def test_heat_transfer_steps_match_oracle(ctx, n, precision):
    """Implicit heat equation ddt(T) - laplacian(kappa,T), Euler, Jacobi-CG to 1e-10
    (examples/heatTransfer/heatTransfer.cpp:60-86), 10 steps, on a 16^3 cube and
    at the heat-transfer case's own size (32^3), in FP64 and with the FP32 CG
    inside FP64 refinement."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl
    print(f"[case] heat {n}^3 {precision}", flush=True)
    poly = fa.PolyMesh.hex(n)
    m = fa.Mesh.from_poly(ctx, poly)
    g = orc.geometry(poly)
    T, kap, T0, Tb = _heat(ctx, m, g)
    dt, kappa = 1e-3, 1e-2
    sp = orc.sparsity(g)
    Tref = T0.copy()
    gam = np.full(m.n_internal_faces + m.n_boundary_faces, kappa)
    fc = g["faceCells"]
    for step in range(10):
        before = T.numpy()
        st = dsl.solve(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, step * dt, dt,
                       solver={"relTol": 1e-10, "maxIter": 1000, "precision": precision})
        # energy: heat gained by the cells = net flux through the walls (to the
        # solver tolerance; internal fluxes cancel)
        after = T.numpy()
        gained = np.sum(g["V"] * (after - before)) / dt
        wall = kappa * np.sum(g["magSf_b"] * g["deltaCoeffs_b"] * (Tb - after[fc]))
        assert abs(gained - wall) <= 1e-6 * max(abs(gained), abs(wall)), (step, gained, wall)
        ls = orc.LinSys(g, sp)
        ls.ddt(Tref, dt)
        ls.laplacian(gam, Tb, [0] * 6, coeff=-1.0)
        Tref, ost = ls.cg(Tref, 1e-10, 1000)
        assert st.finalResNorm <= 1e-10, (st, ost)
        if precision == "double":                 # same recurrence as the oracle's CG
            assert abs(st.numIter - ost["iters"]) <= 2, (st, ost)
    got = T.numpy()
    assert relmax(got, Tref) < 1e-9
    assert got.min() >= 300.0 - 1e-9 and got.max() <= 400.0 + 1e-9   # discrete maximum principle


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_repeated_solves_replay_as_graph(ctx, precision):
    """A repeated solve launches the previous solve's schedule at once, and once
    the same schedule has come five times in a row replays it as a captured
    hipGraph.  A 12-step heat march (32^3, 5 CG iterations per step: captured at
    step 6, replayed at steps 7-12) gives the same bits
    and iteration counts at every step with the graph as launched kernel by
    kernel, with the march's own eager work (assembly, old-time copy, boundary
    update, the reads here) between the replays."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl
    poly = fa.PolyMesh.hex(32)
    g = orc.geometry(poly)
    out = {}
    for graph in (True, False):
        print(f"[case] heat 32^3 {precision} march, graph={graph}", flush=True)
        m = fa.Mesh.from_poly(ctx, poly)          # its own linear system and workspace
        T, kap, T0, Tb = _heat(ctx, m, g)
        steps = []
        for step in range(12):
            st = dsl.solve(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, step * 1e-3, 1e-3,
                           solver={"relTol": 1e-10, "maxIter": 1000, "precision": precision,
                                   "graph": graph})
            assert st.status == 0 and st.finalResNorm <= 1e-10
            steps.append((T.numpy(), st.numIter, st.trueResNorm))
        out[graph] = steps
    for k, (a, b) in enumerate(zip(out[True], out[False])):
        assert a[1] == b[1] and a[2] == b[2] and np.array_equal(a[0], b[0]), k


def test_cg_against_scipy_and_determinism(ctx, meshes):
    import scipy.sparse.linalg as spla
    from foamadapter_amd import dsl, la
    p, m, g = meshes["hex10_jit_perm"]
    T, kap, T0, Tb = _heat(ctx, m, g, dt=1e-2, kappa=1.0)
    ls = dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1e-2)
    A, b = ls.scipy(), ls.rhs()
    ref = spla.spsolve(A.tocsc(), b)
    xs = []
    for _ in range(2):
        x = ctx.upload(T0)
        st = la.solve_cg(ls, x, {"relTol": 1e-12, "maxIter": 2000})
        xs.append(x.numpy())
        assert st.finalResNorm <= 1e-12 and st.trueResNorm < 1e-11
    assert np.array_equal(xs[0], xs[1])          # deterministic reductions
    assert relmax(xs[0], ref) < 1e-10


def test_cg_large_system_both_iteration_forms(ctx):
    """A system above 2^20 rows runs the FP64 CG as three kernels per iteration
    (below it the direction is fused into the SpMV): both forms reach the
    tolerance on the explicit residual (checked on the host), and the mixed
    solver (always fused) agrees with the FP64 solution."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc, la
    n = 104                                   # 1,124,864 cells
    print(f"[case] {n}^3 heat system, FP64 and mixed CG", flush=True)
    m = fa.Mesh.from_poly(ctx, fa.PolyMesh.hex(n), renumber=True)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    assert N >= 1 << 20
    bc = fa.BoundaryConditions(m, ["fixedValue"] * 6)
    T = fvcc.VolumeField(m, ctx.uniform(N, 300.0, 400.0, 41, 0),
                         ctx.uniform(B, 300.0, 400.0, 41, 1), bc, "T")
    kap = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, 1e-1)))
    ls = dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1e-3)
    A, b = ls.scipy(), ls.rhs()
    xs = {}
    for prec in ("double", "mixed"):
        x = ctx.upload(T.internal.numpy())
        st = la.solve_cg(ls, x, {"relTol": 1e-9, "maxIter": 5000, "precision": prec})
        xs[prec] = x.numpy()
        r = np.linalg.norm(b - A @ xs[prec]) / np.linalg.norm(b)
        assert st.finalResNorm <= 1e-9 and r <= 1.01e-9, (prec, st, r)
    assert relmax(xs["mixed"], xs["double"]) < 1e-7


def test_mixed_precision_refinement(ctx):
    """Pressure-Poisson shape: laplacian(1,p) with p=0 on x0, zeroGradient
    elsewhere, b = div(phi) of a random velocity; FP32 CG inside FP64
    refinement to 1e-8 vs plain FP64 CG."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc, la
    poly = fa.PolyMesh.hex(20, jitter=0.2, jitter_seed=31, perm_seed=32)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    kinds = ["fixedValue"] + ["zeroGradient"] * (m.n_patches - 1)
    bc = fa.BoundaryConditions(m, kinds)
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)), bc, "p")
    U = fvcc.VolumeField(m, ctx.uniform(3 * N, -1, 1, 33, 0), ctx.uniform(3 * B, -1, 1, 33, 1))
    phi = fvcc.flux(ctx, m, U)
    one = fvcc.SurfaceField(m, ctx.upload(np.ones(F + B)))
    eqn = dsl.imp.laplacian(one, p) - dsl.exp.div(phi)
    ls = dsl.assemble(eqn, p, 0.0, 1.0)
    x64 = ctx.zeros((N,))
    s64 = la.solve_cg(ls, x64, {"relTol": 1e-8, "maxIter": 5000})
    x32 = ctx.zeros((N,))
    s32 = la.solve_cg(ls, x32, {"relTol": 1e-8, "maxIter": 5000, "precision": "mixed"})
    assert s64.finalResNorm <= 1e-8 and s32.finalResNorm <= 1e-8 and s32.trueResNorm <= 1e-8
    assert s32.outerIter >= 2
    a, b = x32.numpy(), x64.numpy()
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    # and against a direct solve of the same system
    import scipy.sparse.linalg as spla
    ref = spla.spsolve(ls.scipy().tocsc(), ls.rhs())
    assert np.linalg.norm(a - ref) / np.linalg.norm(ref) < 1e-6


def test_no_convergence_reported(ctx, meshes):
    from foamadapter_amd import FvError, _lib, dsl, la
    p, m, g = meshes["hex10_jit_perm"]
    T, kap, T0, Tb = _heat(ctx, m, g, dt=1.0, kappa=1.0)
    ls = dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1.0)
    x = ctx.upload(T0)
    with pytest.raises(FvError) as e:
        la.solve_cg(ls, x, {"relTol": 1e-14, "maxIter": 2})
    assert e.value.code == _lib.FV_ENOCONVERGE
    st = la.solve_cg(ls, ctx.upload(T0), {"relTol": 1e-14, "maxIter": 2}, raise_on_noconverge=False)
    assert st.numIter == 2 and st.status == _lib.FV_ENOCONVERGE


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_mixed_sign_diagonal_fails_loudly(ctx, meshes, precision):
    """The CG runs Jacobi as a symmetric scaling by the square roots of the
    diagonal, which needs a diagonal of one sign: one cell whose implicit
    source flips its diagonal makes the solve fail with FV_ENOCONVERGE and a
    non-finite residual, never a silent answer of another method."""
    from foamadapter_amd import FvError, _lib, dsl, la
    p, m, g = meshes["hex10_jit_perm"]
    T, kap, T0, Tb = _heat(ctx, m, g, dt=1.0, kappa=1.0)
    N = m.n_cells
    c = np.zeros(N)
    c[N // 2] = 1e6
    ls = dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T) - dsl.imp.source(ctx.upload(c), T),
                      T, 0.0, 1.0)
    d = ls.scipy().diagonal()
    assert (d < 0).sum() == 1 and (d > 0).sum() == N - 1     # the forced condition
    print(f"[case] mixed-sign diagonal, {precision}", flush=True)
    with pytest.raises(FvError) as e:
        la.solve_cg(ls, ctx.upload(T0), {"relTol": 1e-10, "maxIter": 50, "precision": precision})
    assert e.value.code == _lib.FV_ENOCONVERGE and "not finite" in str(e.value)
    st = la.solve_cg(ls, ctx.upload(T0), {"relTol": 1e-10, "maxIter": 50, "precision": precision},
                     raise_on_noconverge=False)
    assert st.status == _lib.FV_ENOCONVERGE and not np.isfinite(st.finalResNorm)


def test_pdesolver_set_reference(ctx, meshes):
    """PDESolver with setReference on an all-zeroGradient Poisson problem
    (neoIcoFoam.cpp:144-160 pattern): solvable, and pinned by the reference cell."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc
    p_, m, g = meshes["hex10_jit_perm"]
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    bc = fa.BoundaryConditions(m, ["zeroGradient"] * m.n_patches)
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)), bc, "p")
    # internal fluxes random, boundary fluxes zero: sum of div(phi) V vanishes,
    # so the pure-Neumann problem is compatible and the reference cell pins it
    phi = np.concatenate([orc.uniform(F, -1, 1, 5, 5), np.zeros(B)])
    one = fvcc.SurfaceField(m, ctx.upload(np.ones(F + B)))
    solver = dsl.PDESolver(dsl.imp.laplacian(one, p) - dsl.imp.source(ctx.upload(np.zeros(N)), p)
                           - dsl.exp.div(fvcc.SurfaceField(m, ctx.upload(phi))),
                           p, solver={"relTol": 1e-10, "maxIter": 5000})
    solver.setReference(0, 0.0)
    st = solver.solve()
    assert st.finalResNorm <= 1e-10 and st.numIter > 5
    # the solution satisfies the assembled (referenced) system and the reference value
    ls = solver.linearSystem()
    r = ls.residual(p.internal).numpy()
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(ls.rhs())
    div = orc.div(g, phi, np.ones(N), np.ones(B))
    lap = orc.laplacian(g, np.ones(F + B), p.numpy(), p.boundary.numpy(), [1] * m.n_patches)
    # away from the reference cell lap(p) = div(phi)
    assert np.max(np.abs((lap - div)[1:])) <= 1e-6 * np.max(np.abs(div))
