"""The request server: continuous batching of heat and Poisson requests on
resident meshes, HBM-budget admission, failure handling, and results that are
the same alone and in the mix.

Bars: every completed request's final FP64 relative residual <= its tolerance;
heat results agree with the CPU oracle's Euler march (same steps, CG to 1e-10)
to 1e-6 relative; Poisson results satisfy the assembled system (checked with
scipy) to the tolerance; a request's result is bit-identical alone and in a
mix; failure paths are forced from the host (tiny budget, tiny iteration
limit, non-finite coefficient, cancellation).
"""
import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu

MESHES = [(6, 201), (9, 202), (12, 203)]


def _polys():
    import foamadapter_amd as fa
    return [fa.PolyMesh.hex(n, jitter=0.2, jitter_seed=s, perm_seed=s) for n, s in MESHES]


@pytest.fixture(scope="module")
def polys():
    return _polys()


def _server(polys, **kw):
    from foamadapter_amd.serve import Server
    srv = Server(0, **kw)
    ids = [srv.add_mesh(p, renumber=True) for p in polys]
    return srv, ids


def _oracle_heat(poly, kappa, dt, steps, walls, T0):
    g = orc.geometry(poly)
    sp = orc.sparsity(g)
    po = g["patchOffsets"]
    Tb = np.concatenate([np.full(po[i + 1] - po[i], walls[i]) for i in range(6)])
    gam = np.full(len(g["owner"]) + len(Tb), kappa)
    T = T0.copy()
    for _ in range(steps):
        ls = orc.LinSys(g, sp)
        ls.ddt(T, dt)
        ls.laplacian(gam, Tb, [0] * 6, coeff=-1.0)
        T, st = ls.cg(T, 1e-12, 5000)
    return T


@pytest.mark.parametrize("batched,precision,solo", [(True, "mixed", -1), (True, "double", -1),
                                                    (False, "double", 0), (True, "mixed", 500)])
def test_heat_requests_match_oracle(polys, batched, precision, solo):
    """(solo 500: the 9^3 and 12^3 requests of the batched server run alone
    through the single-system solver, the 6^3 one in the batch)"""
    srv, ids = _server(polys, batched=batched, precision=precision, tile_rows=256, chunk=8,
                       solo_rows=solo)
    reqs = []
    for k, (mid, poly) in enumerate(zip(ids, polys)):
        N = poly.n_cells
        walls = orc.uniform(6, 300, 400, 300 + k, 0)
        T0 = orc.uniform(N, 300, 400, 300 + k, 1)
        kappa = [1e-1, 1e-2, 3e-2][k]
        rid = srv.submit_heat(mid, kappa=kappa, dt=1e-3, steps=3, walls=walls, T0=T0, rel_tol=1e-10)
        reqs.append((rid, mid, poly, kappa, walls, T0))
    for rid, mid, poly, kappa, walls, T0 in reqs:
        T, st = srv.wait(rid, mid)
        assert st.status == 0 and st.final_res <= 1e-10, st
        ref = _oracle_heat(poly, kappa, 1e-3, 3, walls, T0)
        assert np.max(np.abs(T - ref)) / np.max(np.abs(ref)) < 1e-6
    s = srv.stats()
    assert s["completed"] == 3 and s["failed"] == 0 and s["p99_ms"] >= s["p50_ms"] > 0
    assert s["solo_runs"] == (2 if solo == 500 else 0)
    srv.close()


def test_solo_lane_matches_plain_and_closes_cleanly(polys):
    """Requests on meshes of at least solo_rows cells run on the solo lane (a
    second context and worker thread, concurrent with the batch): their results
    are bit-identical to the same requests through a plain server (the same
    fv_solve_cg), while the batch serves the small mesh; a queued solo request
    can be cancelled, and one never collected is released when the server
    closes with both workers running."""
    work = [("poisson", 2, 11), ("heat", 1, 12), ("heat", 0, 13), ("poisson", 1, 14),
            ("heat", 2, 15), ("poisson", 0, 16)]

    def submit(srv, ids, kind, k, seed):
        if kind == "poisson":
            return srv.submit_poisson(ids[k], seed=seed, rel_tol=1e-8)
        return srv.submit_heat(ids[k], kappa=2e-2, steps=2, seed=seed, rel_tol=1e-8)

    print("[case] solo lane vs plain", flush=True)
    plain, pids = _server(polys, batched=False, precision="mixed", chunk=8)
    ref = [plain.wait(submit(plain, pids, *w), pids[w[1]]) for w in work]
    plain.close()
    srv, ids = _server(polys, batched=True, precision="mixed", tile_rows=256, chunk=8,
                       solo_rows=500)
    rids = [submit(srv, ids, *w) for w in work]
    extra = [srv.submit_poisson(ids[2], seed=s) for s in (31, 32, 33)]
    srv.cancel(extra[2])
    for (kind, k, seed), rid, (x0, st0) in zip(work, rids, ref):
        x, st = srv.wait(rid, ids[k])
        assert st.status == 0 and st0.status == 0, (kind, k)
        if k > 0:                                    # the 9^3 and 12^3 meshes: solo lane
            assert np.array_equal(x, x0), (kind, k)
        else:
            assert np.linalg.norm(x - x0) <= 1e-6 * np.linalg.norm(x0), (kind, k)
    assert srv.wait(extra[0], ids[2])[1].status == 0
    st2 = srv.wait(extra[2])[1].status
    assert st2 in (0, 6)                             # cancelled unless it already ran
    s = srv.stats()
    assert s["solo_runs"] >= 5 and s["solo_ms"] > 0
    srv.close()                                      # extra[1] never collected


def test_poisson_requests_solve_their_system(ctx, polys):
    """Poisson request result satisfies the assembled system (rebuilt here on the
    caller's numbering with the same seeded velocity) to the tolerance."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc
    srv, ids = _server(polys, batched=True, precision="mixed", tile_rows=256, chunk=8)
    rid = srv.submit_poisson(ids[2], seed=77, rel_tol=1e-8)
    p_srv, st = srv.wait(rid, ids[2])
    assert st.status == 0 and st.final_res <= 1e-8
    srv.close()
    # same problem through the library on the renumbered mesh, mapped back
    m = fa.Mesh.from_poly(ctx, polys[2], renumber=True)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    bc = fa.BoundaryConditions(m, ["fixedValue"] + ["zeroGradient"] * 5)
    p = fvcc.VolumeField(m, ctx.zeros((N,)), ctx.zeros((B,)), bc, "p")
    U = fvcc.VolumeField(m, ctx.uniform(3 * N, -1, 1, 77, 1), ctx.uniform(3 * B, -1, 1, 77, 2))
    phi = fvcc.flux(ctx, m, U)
    one = fvcc.SurfaceField(m, ctx.upload(np.ones(F + B)))
    ls = dsl.assemble(dsl.imp.laplacian(one, p) - dsl.exp.div(phi), p, 0.0, 1.0)
    A, b = ls.scipy(), ls.rhs()
    x_dev = p_srv[m.cell_map.numpy()]          # caller numbering -> device numbering
    assert np.linalg.norm(b - A @ x_dev) / np.linalg.norm(b) <= 1.01e-8


def test_result_same_alone_and_in_mix(polys):
    from foamadapter_amd.serve import Server  # noqa: F401
    srv, ids = _server(polys, batched=True, precision="mixed", tile_rows=256, chunk=8)
    rid = srv.submit_heat(ids[1], kappa=2e-2, steps=2, seed=5)
    alone, _ = srv.wait(rid, ids[1])
    srv.close()
    srv, ids = _server(polys, batched=True, precision="mixed", tile_rows=256, chunk=8)
    others = [srv.submit_poisson(ids[2], seed=9), srv.submit_heat(ids[0], kappa=1e-1, seed=3)]
    rid = srv.submit_heat(ids[1], kappa=2e-2, steps=2, seed=5)
    others.append(srv.submit_heat(ids[2], kappa=1e-3, seed=4))
    mix, st = srv.wait(rid, ids[1])
    for o in others:
        srv.wait(o)
    assert st.status == 0
    assert np.array_equal(alone, mix)
    srv.close()


def test_results_through_pinned_pool(polys):
    """Results wait in pinned host buffers reused from a pool: repeats of one
    request, interleaved with requests on meshes of other sizes and collected
    newest first, return the first run's bits; requests never collected are
    released when the server closes."""
    srv, ids = _server(polys, batched=True, precision="mixed", tile_rows=256, chunk=8)
    ref, st = srv.wait(srv.submit_heat(ids[1], kappa=2e-2, steps=2, seed=5), ids[1])
    assert st.status == 0
    repeats, others = [], []
    for k in range(3):
        repeats.append(srv.submit_heat(ids[1], kappa=2e-2, steps=2, seed=5))
        others.append((srv.submit_poisson(ids[k], seed=20 + k), ids[k]))
    srv.submit_heat(ids[0], kappa=1e-2, seed=40)      # never collected
    srv.submit_poisson(ids[2], seed=41)                # never collected
    for r in reversed(repeats):
        res, st = srv.wait(r, ids[1])
        assert st.status == 0 and np.array_equal(res, ref)
    for r, mid in others:
        res, st = srv.wait(r, mid)
        assert st.status == 0 and res.shape == (polys[mid].n_cells,) and np.all(np.isfinite(res))
    srv.close()


def test_budget_admission_and_rejection(polys):
    from foamadapter_amd import FvError
    from foamadapter_amd.serve import HEAT
    # size the budget from the server's own figures: the resident meshes plus
    # one and a half heat requests on the largest mesh -> one request at a time
    probe, ids = _server(polys, batched=True)
    resident = probe.stats()["resident_bytes"]
    est = probe.request_bytes(ids[2], HEAT)
    probe.close()
    assert resident > 0 and est > 8 * polys[2].n_cells
    srv, ids = _server(polys, batched=True, hbm_budget=resident + est * 3 // 2, tile_rows=256,
                       chunk=8)
    rids = [srv.submit_heat(ids[2], kappa=1e-2, steps=1, seed=i) for i in range(4)]
    for r in rids:
        _, st = srv.wait(r)
        assert st.status == 0
    s = srv.stats()
    assert s["completed"] == 4 and s["peak_active"] == 1
    srv.close()
    # room for two at a time
    srv, ids = _server(polys, batched=True, hbm_budget=resident + est * 2, tile_rows=256, chunk=8)
    rids = [srv.submit_heat(ids[2], kappa=1e-2, steps=1, seed=i) for i in range(4)]
    for r in rids:
        assert srv.wait(r)[1].status == 0
    assert srv.stats()["peak_active"] <= 2
    srv.close()
    # a request that can never fit is rejected at submission
    srv, ids = _server(polys, batched=True, hbm_budget=resident + est // 2)
    with pytest.raises(FvError) as e:
        srv.submit_heat(ids[2], kappa=1e-2)
    assert e.value.code == 2   # FV_ENOMEM_BUDGET
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_failures_and_cancellation(polys):
    srv, ids = _server(polys, batched=True, tile_rows=256, chunk=4)
    bad_iter = srv.submit_poisson(ids[2], seed=1, rel_tol=1e-14, max_iter=5)
    bad_nan = srv.submit_heat(ids[0], kappa=float("nan"), steps=1, seed=2)
    good = srv.submit_heat(ids[1], kappa=1e-2, steps=1, seed=3)
    _, s1 = srv.wait(bad_iter)
    _, s2 = srv.wait(bad_nan)
    _, s3 = srv.wait(good)
    assert s1.status == 3 and s2.status == 3 and s3.status == 0   # FV_ENOCONVERGE
    srv.close()
    # one at a time: a queued request is cancelled before it runs
    srv, ids = _server(polys, batched=False)
    first = srv.submit_poisson(ids[2], seed=1)
    queued = [srv.submit_poisson(ids[2], seed=s) for s in (2, 3)]
    srv.cancel(queued[-1])
    _, st0 = srv.wait(first)
    _, st1 = srv.wait(queued[0])
    _, st2 = srv.wait(queued[1])
    assert st0.status == 0 and st1.status == 0 and st2.status == 6   # FV_ECANCELLED
    assert srv.stats()["cancelled"] == 1
    srv.close()
