"""Batched solves: many heat systems of uneven sizes in one launch sequence.

Bars: every system reaches its FP64 relative-residual tolerance (explicit
||b - A x|| / ||b|| checked on the host with scipy), agrees with the direct
solve within cond-bounded error, and its bits do not depend on the other
members of the batch (alone == in the mix), for FP64 and mixed precision, with
and without hipGraph replay.
"""
import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu


def _heat_system(ctx, n, kappa, seed, jitter=0.0):
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc, la
    poly = fa.PolyMesh.hex(n, jitter=jitter, jitter_seed=seed, perm_seed=seed)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    walls = orc.uniform(6, 300.0, 400.0, seed, 9)
    po = m.get("patchOffsets")
    Tb = np.concatenate([np.full(po[i + 1] - po[i], walls[i]) for i in range(6)])
    bc = fa.BoundaryConditions(m, ["fixedValue"] * 6)
    T = fvcc.VolumeField(m, ctx.uniform(N, 300.0, 400.0, seed, 0), ctx.upload(Tb), bc, "T")
    kap = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, kappa)))
    ls = la.LinearSystem(la.SparsityPattern(m))
    dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1e-3, ls)
    return {"mesh": m, "T": T, "ls": ls, "x0": T.numpy(), "bc": bc}


@pytest.fixture(scope="module")
def systems(ctx):
    sizes = [(5, 1e-1), (12, 1e-2), (9, 1e-4), (20, 3e-2), (16, 1e-1), (7, 5e-3)]
    out = []
    for i, (n, k) in enumerate(sizes):
        print(f"[case] batch member {i}: {n}^3 kappa={k}", flush=True)
        out.append(_heat_system(ctx, n, k, 100 + i, jitter=0.15))
    return out


def _run(ctx, systems, idx, precision, graph, tile_rows=256, chunk=8):
    from foamadapter_amd import la
    b = la.Batch(ctx, precision, tile_rows, chunk)
    xs = []
    for i in idx:
        x = ctx.upload(systems[i]["x0"])
        b.add(systems[i]["ls"], x)
        xs.append(x)
    stats = b.solve({"relTol": 1e-8, "maxIter": 2000, "innerRelTol": 1e-3}, use_graph=graph)
    res = [x.numpy() for x in xs]
    b.close()
    return res, stats


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_converges_and_matches_direct(ctx, systems, precision, graph):
    import scipy.sparse.linalg as spla
    idx = list(range(len(systems)))
    xs, stats = _run(ctx, systems, idx, precision, graph)
    for i, x, st in zip(idx, xs, stats):
        ls = systems[i]["ls"]
        A, b = ls.scipy(), ls.rhs()
        r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert r <= 1e-8 * 1.01, (i, r, st)
        assert st.finalResNorm <= 1e-8 and abs(st.finalResNorm - r) <= 1e-10, (i, st.finalResNorm, r)
        ref = spla.spsolve(A.tocsc(), b)
        assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) < 1e-7, i


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_batch_wide_rows(ctx, systems, precision):
    """A polyhedral member with ten-entry rows (non-conforming refinement) in a
    batch with a hex member: converges, matches the direct solve, and gives the
    same bits alone and in the batch."""
    import scipy.sparse.linalg as spla
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc, la
    from skewed_meshes import refined_block
    m = fa.Mesh.from_poly(ctx, refined_block(3))
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    bc = fa.BoundaryConditions(m, ["fixedValue"] * m.n_patches)
    Tb = orc.uniform(B, 300.0, 400.0, 7, 1)
    T = fvcc.VolumeField(m, ctx.uniform(N, 300.0, 400.0, 7, 0), ctx.upload(Tb), bc, "T")
    kap = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, 1.0)))
    ls = la.LinearSystem(la.SparsityPattern(m))
    dsl.assemble(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1e-2, ls)
    A, b = ls.scipy(), ls.rhs()
    assert np.diff(A.indptr).max() == 10
    ref = spla.spsolve(A.tocsc(), b)
    x0 = T.numpy()
    out = []
    for members in ([ls], [systems[1]["ls"], ls]):
        bt = la.Batch(ctx, precision, 64, 8)
        xs = [ctx.upload(x0 if s is ls else systems[1]["x0"]) for s in members]
        for s, x in zip(members, xs):
            bt.add(s, x)
        st = bt.solve({"relTol": 1e-10, "maxIter": 500}, use_graph=True)
        out.append(xs[-1].numpy())
        assert st[-1].finalResNorm <= 1e-10, st[-1]
        bt.close()
    assert np.max(np.abs(out[0] - ref)) / np.max(np.abs(ref)) < 1e-9
    assert np.array_equal(out[0], out[1])
    ls.close()
    bc.close()
    m.close()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_alone_equals_in_mix(ctx, systems, precision):
    """A system's bits do not depend on the rest of the batch."""
    mix, _ = _run(ctx, systems, [0, 1, 2, 3, 4, 5], precision, True)
    for i in (1, 3, 5):
        alone, _ = _run(ctx, systems, [i], precision, True)
        assert np.array_equal(alone[0], mix[i]), i
    # another mix and order
    other, _ = _run(ctx, systems, [3, 1], precision, False)
    assert np.array_equal(other[0], mix[3]) and np.array_equal(other[1], mix[1])


def test_graph_and_stream_agree(ctx, systems):
    a, _ = _run(ctx, systems, [0, 2, 4], "mixed", True)
    b, _ = _run(ctx, systems, [0, 2, 4], "mixed", False)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_iteration_limit_reported(ctx, systems):
    from foamadapter_amd import FvError, _lib, la
    b = la.Batch(ctx, "mixed", 256, 4)
    x = ctx.upload(systems[3]["x0"])
    b.add(systems[3]["ls"], x)
    with pytest.raises(FvError) as e:
        b.solve({"relTol": 1e-14, "maxIter": 3})
    assert e.value.code == _lib.FV_ENOCONVERGE
    b.close()


def test_fused_heat_assembly_bitwise(ctx, systems):
    """fv_assemble_ddt_laplacian == zero + ddt + laplacian, bit for bit."""
    from foamadapter_amd import _lib, la
    s = systems[3]
    m, T = s["mesh"], s["T"]
    F, B = m.n_internal_faces, m.n_boundary_faces
    kap = ctx.upload(np.full(F + B, 3e-2))
    a = la.LinearSystem(la.SparsityPattern.readOrCreate(m))
    b = la.LinearSystem(la.SparsityPattern.readOrCreate(m))
    b.zero()
    _lib.call("fv_assemble_ddt_euler", b.handle, T.internal.ptr, 1e-3, 1.0)
    _lib.call("fv_assemble_laplacian", b.handle, kap.ptr, T.bcs.handle, T.boundary.ptr, None, -1.0)
    _lib.call("fv_assemble_ddt_laplacian", a.handle, T.internal.ptr, 1e-3, kap.ptr, T.bcs.handle,
              T.boundary.ptr, None, -1.0)
    assert np.array_equal(a.values(), b.values()) and np.array_equal(a.rhs(), b.rhs())


def test_heat_assembly_many_bitwise(ctx, systems):
    """One fv_assemble_ddt_laplacian_many launch over meshes of several sizes ==
    each system's own fv_assemble_ddt_laplacian, bit for bit (matrix, rhs and
    boundary coefficients), with and without a gradient field."""
    import ctypes as C
    from foamadapter_amd import _lib, la

    def bcoeffs(ls, B):
        pm, pr = C.c_void_p(), C.c_void_p()
        _lib.call("fv_linsys_boundary_coeffs", ls.handle, C.byref(pm), C.byref(pr))
        out = []
        for ptr in (pm, pr):
            h = np.empty(B)
            _lib.call("fv_memcpy_d2h", ctx.handle, C.c_void_p(h.ctypes.data), ptr, h.nbytes)
            out.append(h)
        return out

    items, singles = [], []
    for i, s in enumerate(systems):
        m, T = s["mesh"], s["T"]
        F, B = m.n_internal_faces, m.n_boundary_faces
        gam = ctx.upload(orc.uniform(F + B, 1e-3, 1e-1, 300 + i, 0))
        tg = ctx.upload(orc.uniform(B, -1.0, 1.0, 300 + i, 1)) if i % 2 else None
        dt = 1e-3 * (1 + i)
        sp = la.SparsityPattern.readOrCreate(m)
        many, one = la.LinearSystem(sp), la.LinearSystem(sp)
        items.append((many, T.internal, dt, gam, T.bcs, T.boundary, tg))
        _lib.call("fv_assemble_ddt_laplacian", one.handle, T.internal.ptr, dt, gam.ptr,
                  T.bcs.handle, T.boundary.ptr, tg.ptr if tg is not None else None, -1.0)
        singles.append((one, B))
    la.HeatAssembly(items).assemble(-1.0)
    for (many, *_), (one, B) in zip(items, singles):
        assert np.array_equal(many.values(), one.values())
        assert np.array_equal(many.rhs(), one.rhs())
        for a, b in zip(bcoeffs(many, B), bcoeffs(one, B)):
            assert np.array_equal(a, b)


def test_batch_copies_old_solution(ctx, systems):
    """fv_batch_solve leaves x_old = x for members added with a copy target."""
    from foamadapter_amd import la
    b = la.Batch(ctx, "mixed", 256, 8)
    xs, xos = [], []
    for i in (0, 3, 5):
        x = ctx.upload(systems[i]["x0"])
        xo = ctx.zeros((len(systems[i]["x0"]),))
        b.add(systems[i]["ls"], x, xo if i != 3 else None)
        xs.append(x)
        xos.append(xo)
    b.solve({"relTol": 1e-8, "maxIter": 2000})
    assert np.array_equal(xos[0].numpy(), xs[0].numpy())
    assert not np.any(xos[1].numpy())          # no copy target: untouched
    assert np.array_equal(xos[2].numpy(), xs[2].numpy())
    b.close()


def test_autotune_picks_and_caches(ctx, systems, tmp_path):
    from foamadapter_amd import la
    b = la.Batch(ctx, "mixed", 1024, 16)
    xs = [ctx.upload(s["x0"]) for s in systems]
    for s, x in zip(systems, xs):
        b.add(s["ls"], x)
    before = [x.numpy() for x in xs]
    cache = str(tmp_path / "tune.json")
    t, c = b.autotune({"relTol": 1e-8, "maxIter": 2000}, cache)
    assert t in (256, 512, 1024, 2048, 4096) and c in (8, 16, 32)
    for x, x0 in zip(xs, before):          # trials restore the initial guess
        assert np.array_equal(x.numpy(), x0)
    import json
    entries = json.load(open(cache))
    assert len(entries) == 1 and list(entries.values())[0]["tile_rows"] == t
    # a second batch of the same shape takes the cached shape
    b2 = la.Batch(ctx, "mixed", 1024, 16)
    for s in systems:
        b2.add(s["ls"], ctx.upload(s["x0"]))
    assert b2.autotune({"relTol": 1e-8, "maxIter": 2000}, cache) == (t, c)
    stats = b.solve({"relTol": 1e-8, "maxIter": 2000})
    assert all(st.finalResNorm <= 1e-8 for st in stats)
    b.close()
    b2.close()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_batch_absolute_tolerance(ctx, systems, precision):
    """An fvSolution 'relTol 0; tolerance X' dictionary (absolute_residual_norm,
    src/compatibility/fvSolution.cpp:139-144) stops each member of a batch on
    ||b - A x|| <= X, as fv_solve_cg does, and returns OK."""
    from foamadapter_amd import la
    idx = [0, 1, 3]
    b = la.Batch(ctx, precision, 256, 8)
    xs = []
    for i in idx:
        x = ctx.upload(systems[i]["x0"])
        b.add(systems[i]["ls"], x)
        xs.append(x)
    bn = [np.linalg.norm(systems[i]["ls"].rhs()) for i in idx]
    atol = 1e-9 * min(bn)
    stats = b.solve({"relTol": 0.0, "tolerance": atol, "maxIter": 3000})
    b.close()
    for i, x, st, nb in zip(idx, xs, stats, bn):
        ls = systems[i]["ls"]
        r = np.linalg.norm(ls.rhs() - ls.scipy() @ x.numpy())
        assert st.status == 0 and st.finalResNorm * nb <= atol * 1.0001, (i, st)
        assert r <= atol * 1.01, (i, r, atol)
