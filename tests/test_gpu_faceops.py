"""GPU parity of the mesh geometry and the explicit face operators against the
CPU oracle (and the OpenFOAM golden fields) through the C ABI.

Bar: FP64 within 1e-12 relative (max-norm) of the oracle -- in practice the
kernels follow the oracle's serial operation order without FMA contraction and
agree bit for bit, which test_*_bitwise checks; FP32 within 1e-5 of the FP64
oracle and bitwise equal to the FP32 oracle.
"""
import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def relmax(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _meshes(fixtures_dir):
    from foamadapter_amd.mesh import PolyMesh, read_polymesh
    from skewed_meshes import SHEAR_A, affine_hex, refined_block, tapered_hex
    return {
        "setup_operator": read_polymesh(fixtures_dir / "setup_operator"),
        "setup_stencil3D": read_polymesh(fixtures_dir / "setup_stencil3D"),
        "setup_unstructuredMesh": read_polymesh(fixtures_dir / "setup_unstructuredMesh"),
        "hex8": PolyMesh.hex(8),
        "hex12_jit_perm": PolyMesh.hex(12, jitter=0.2, jitter_seed=1002, perm_seed=1002),
        "hex7x5x3_jit": PolyMesh.hex((7, 5, 3), lengths=(0.7, 0.5, 0.3), jitter=0.3,
                                     jitter_seed=5),
        "hex40_jit_perm": PolyMesh.hex(40, jitter=0.2, jitter_seed=9, perm_seed=10),
        # skewed planar faces with closed-form geometry (pinned in test_oracle_cpu)
        "hex6_affine_perm": affine_hex(6, perm_seed=3)[1],
        # x- and y-faces at ~88 degrees to their deltas: clamped nonOrthDeltaCoeffs
        "hex6_shear_perm": affine_hex(6, perm_seed=3, A=SHEAR_A)[1],
        "hex4_tapered": tapered_hex(4)[1],
        # polyhedral cells: hanging points, ten-entry rows, a cell with nine lower
        # faces in one workgroup (more than the staged gather's list holds)
        "refined_block": refined_block(3),
    }


@pytest.fixture(scope="module")
def cases(ctx, fixtures_dir):
    from foamadapter_amd import Mesh
    out = {}
    for name, poly in _meshes(fixtures_dir).items():
        print(f"[case] {name}: building device mesh", flush=True)
        m = Mesh.from_poly(ctx, poly)
        out[name] = (poly, m, orc.geometry(poly))
    return out


GEOM_KEYS = ["V", "C", "owner", "neighbour", "Sf", "Cf", "magSf", "weights", "deltaCoeffs",
             "nonOrthDeltaCoeffs", "faceCells", "Sf_b", "Cf_b", "magSf_b", "nf_b", "delta_b",
             "weights_b", "deltaCoeffs_b", "nonOrthDeltaCoeffs_b", "patchOffsets"]


def test_geometry_matches_oracle(cases):
    for name, (poly, m, g) in cases.items():
        for k in GEOM_KEYS:
            got = m.get(k)
            assert got.shape == np.shape(g[k]), (name, k)
            if got.dtype.kind == "i":
                assert np.array_equal(got, g[k]), (name, k)
            else:
                assert relmax(got, g[k]) <= TOL64, (name, k, relmax(got, g[k]))


def test_geometry_bitwise(cases):
    for name, (poly, m, g) in cases.items():
        for k in GEOM_KEYS:
            assert np.array_equal(m.get(k), g[k]), (name, k)


def test_cell_face_stencil(cases):
    """CellToFaceStencil: each cell's faces, ascending (test_stencils.cpp:34-54)."""
    for name, (poly, m, g) in cases.items():
        off, faces = m.get("cellFaceOffsets"), m.get("cellFaces")
        F = len(g["owner"])
        N = len(g["V"])
        expect = [[] for _ in range(N)]
        for f, (o, n) in enumerate(zip(g["owner"], g["neighbour"])):
            expect[o].append(f)
            expect[n].append(f)
        for b, c in enumerate(g["faceCells"]):
            expect[c].append(F + b)
        for c in range(N):
            assert list(faces[off[c]:off[c + 1]]) == sorted(expect[c]), (name, c)


def _fields(g, seed):
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    T = orc.uniform(N, -1.0, 1.0, seed, 0)
    Tb = orc.uniform(B, -1.0, 1.0, seed, 1)
    phi = orc.uniform(F + B, -1.0, 1.0, seed, 2)
    gamma = orc.uniform(F + B, 0.5, 2.0, seed, 3)
    Tg = orc.uniform(B, -1.0, 1.0, seed, 4)
    return T, Tb, phi, gamma, Tg


def _kinds(m, g):
    # alternate BC kinds over the patches: fixedValue, fixedGradient, calculated
    return [(0, 1, 2)[p % 3] for p in range(m.n_patches)]


def _run_all(ctx, m, g, dt, seed=7):
    from foamadapter_amd import BoundaryConditions, fvcc
    T, Tb, phi, gamma, Tg = _fields(g, seed)
    kinds = _kinds(m, g)
    bc = BoundaryConditions(m, kinds)
    Tf = fvcc.VolumeField(m, ctx.upload(T.astype(dt)), ctx.upload(Tb.astype(dt)), bc,
                          boundary_gradient=ctx.upload(Tg.astype(dt)))
    phiF = fvcc.SurfaceField(m, ctx.upload(phi.astype(dt)))
    gamF = fvcc.SurfaceField(m, ctx.upload(gamma.astype(dt)))
    n = m.n_internal_faces + m.n_boundary_faces
    res = {}
    out = fvcc.SurfaceField(m, ctx.empty((n,), dt))
    fvcc.SurfaceInterpolation(ctx, m, ["linear"]).interpolate(Tf, out)
    res["interp_linear"] = out.numpy()
    fvcc.SurfaceInterpolation(ctx, m, ["upwind"]).interpolate(Tf, out, phiF)
    res["interp_upwind"] = out.numpy()
    res["grad"] = fvcc.GaussGreenGrad(ctx, m).grad(Tf).numpy()
    res["div_linear"] = fvcc.GaussGreenDiv(ctx, m, ["Gauss", "linear"]).div(None, phiF, Tf).numpy()
    res["div_upwind"] = fvcc.GaussGreenDiv(ctx, m, ["Gauss", "upwind"]).div(None, phiF, Tf).numpy()
    res["surface_integrate"] = fvcc.GaussGreenDiv(ctx, m).div(None, phiF).numpy()
    res["laplacian"] = fvcc.GaussGreenLaplacian(ctx, m).laplacian(None, gamF, Tf).numpy()
    res["sngrad"] = fvcc.FaceNormalGradient(ctx, m).faceNormalGrad(Tf).numpy()
    ref = {
        "interp_linear": orc.interpolate(g, T, Tb, 0, None, dt),
        "interp_upwind": orc.interpolate(g, T, Tb, 1, phi, dt),
        "grad": orc.grad(g, T, Tb, dt),
        "div_linear": orc.div(g, phi, T, Tb, 0, dt),
        "div_upwind": orc.div(g, phi, T, Tb, 1, dt),
        "surface_integrate": orc.div(g, phi, np.ones(len(g["V"])), np.ones(len(g["faceCells"])), 0, dt),
        "laplacian": orc.laplacian(g, gamma, T, Tb, kinds, Tg, dt),
        "sngrad": orc.sngrad(g, T, Tb, kinds, Tg, dt),
    }
    bc.close()
    return res, ref


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_face_ops_parity(ctx, cases, dt):
    for name, (poly, m, g) in cases.items():
        print(f"[case] {name} {np.dtype(dt).name}", flush=True)
        res, ref = _run_all(ctx, m, g, dt)
        for op in res:
            err = relmax(res[op], ref[op])
            tol = TOL64 if dt == np.float64 else 1e-6
            assert err <= tol, (name, op, err)
            if dt == np.float32:
                ref64 = _run_ref64(g, op)
                assert relmax(res[op], ref64) < 1e-4, (name, op)


def _run_ref64(g, op):
    # FP64 oracle of the same op for the FP32 accuracy check
    T, Tb, phi, gamma, Tg = _fields(g, 7)
    kinds = [(0, 1, 2)[p % 3] for p in range(len(g["patchOffsets"]) - 1)]
    return {
        "interp_linear": lambda: orc.interpolate(g, T, Tb, 0),
        "interp_upwind": lambda: orc.interpolate(g, T, Tb, 1, phi),
        "grad": lambda: orc.grad(g, T, Tb),
        "div_linear": lambda: orc.div(g, phi, T, Tb, 0),
        "div_upwind": lambda: orc.div(g, phi, T, Tb, 1),
        "surface_integrate": lambda: orc.div(g, phi, np.ones(len(g["V"])), np.ones(len(g["faceCells"]))),
        "laplacian": lambda: orc.laplacian(g, gamma, T, Tb, kinds, Tg),
        "sngrad": lambda: orc.sngrad(g, T, Tb, kinds, Tg),
    }[op]()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_face_ops_bitwise(ctx, cases, dt):
    for name, (poly, m, g) in cases.items():
        res, ref = _run_all(ctx, m, g, dt)
        for op in res:
            assert np.array_equal(res[op], ref[op]), (name, op, np.mean(res[op] == ref[op]))


def test_golden_through_gpu(ctx, cases, fixtures_dir):
    """OpenFOAM goldens (test_operators.cpp:72-80,109-113) through the HIP path."""
    from foamadapter_amd import fvcc
    from foamadapter_amd.mesh import read_internal_field
    poly, m, g = cases["setup_operator"]
    case = fixtures_dir / "setup_operator"
    T = read_internal_field(case / "0" / "T")
    Tb = T[g["faceCells"]]
    phi = np.concatenate([read_internal_field(case / "0" / "phi"), np.zeros(m.n_boundary_faces)])
    Tf = fvcc.VolumeField(m, ctx.upload(T), ctx.upload(Tb))
    grad = fvcc.GaussGreenGrad(ctx, m).grad(Tf).numpy()
    div = fvcc.GaussGreenDiv(ctx, m, ["linear"]).div(None, fvcc.SurfaceField(m, ctx.upload(phi)), Tf).numpy()
    of_g = read_internal_field(case / "0" / "ofGradT", 3)
    of_d = read_internal_field(case / "0" / "ofDivT")
    assert relmax(grad[:, :2], of_g[:, :2]) < 1e-15
    assert relmax(div, of_d) < 1e-15


def test_flux_and_correct_boundary(ctx, cases):
    from foamadapter_amd import BoundaryConditions, fvcc
    for name, (poly, m, g) in cases.items():
        N, B = len(g["V"]), len(g["faceCells"])
        U = orc.uniform(3 * N, -1, 1, 3, 5).reshape(N, 3)
        Ub = orc.uniform(3 * B, -1, 1, 3, 6).reshape(B, 3)
        phi = fvcc.flux(ctx, m, fvcc.VolumeField(m, ctx.upload(U), ctx.upload(Ub))).numpy()
        ref = orc.flux(g, U, Ub)
        assert np.array_equal(phi, ref), (name, relmax(phi, ref))
        kinds = _kinds(m, g)
        T = orc.uniform(N, 0, 1, 3, 7)
        Tb0 = orc.uniform(B, 0, 1, 3, 8)
        Tg = orc.uniform(B, -1, 1, 3, 9)
        bc = BoundaryConditions(m, kinds)
        f = fvcc.VolumeField(m, ctx.upload(T), ctx.upload(Tb0), bc, boundary_gradient=ctx.upload(Tg))
        f.correct_boundary_conditions()
        assert np.array_equal(f.boundary.numpy(), orc.correct_boundary(g, T, Tb0, kinds, Tg)), name


def test_deterministic_rerun(ctx, cases):
    poly, m, g = cases["hex40_jit_perm"]
    a, _ = _run_all(ctx, m, g, np.float64, seed=11)
    b, _ = _run_all(ctx, m, g, np.float64, seed=11)
    for op in a:
        assert np.array_equal(a[op], b[op]), op


def test_mesh_from_geometry_arrays(ctx, cases):
    """fv_mesh_create (NeoN-layout arrays) gives the same operators as the poly path."""
    from foamadapter_amd import Mesh
    poly, m, g = cases["hex12_jit_perm"]
    m2 = Mesh.from_geometry(ctx, g)
    a, _ = _run_all(ctx, m, g, np.float64)
    b, _ = _run_all(ctx, m2, g, np.float64)
    for op in a:
        assert np.array_equal(a[op], b[op]), op
    m2.close()


def test_bad_arguments_raise(ctx, cases):
    from foamadapter_amd import FvError, _lib
    poly, m, g = cases["hex8"]
    with pytest.raises(FvError):
        _lib.call("fv_grad_gauss", ctx.handle, m.handle, 5, 1.0, None, None, None)
    with pytest.raises(FvError):
        _lib.call("fv_laplacian_uncorrected", ctx.handle, m.handle, 0, 1.0, None, None, None,
                  None, None, None)


def test_renumbered_mesh_roundtrip(ctx):
    """A mesh renumbered at import (reverse Cuthill-McKee) gives the operators of
    the caller's mesh: fields mapped in, results mapped back, within 1e-12; and
    bit-identical to the oracle run on the renumbered polyMesh."""
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    poly = fa.PolyMesh.hex(14, jitter=0.2, jitter_seed=41, perm_seed=42)
    g = orc.geometry(poly)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    g_ren = orc.geometry(m.poly)
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    T, Tb, phi, gamma, Tg = _fields(g, 43)
    kinds = [0, 1, 2, 0, 1, 2]
    bc = fa.BoundaryConditions(m, kinds)
    Td = m.to_device(ctx.upload(T), "cell")
    phid = m.to_device(ctx.upload(phi), "face", oriented=True)
    gamd = m.to_device(ctx.upload(gamma), "face")
    Tf = fvcc.VolumeField(m, Td, ctx.upload(Tb), bc, boundary_gradient=ctx.upload(Tg))
    grad = m.to_caller(fvcc.GaussGreenGrad(ctx, m).grad(Tf), "cell").numpy()
    div = m.to_caller(fvcc.GaussGreenDiv(ctx, m).div(None, fvcc.SurfaceField(m, phid), Tf), "cell").numpy()
    lap = m.to_caller(fvcc.GaussGreenLaplacian(ctx, m).laplacian(None, fvcc.SurfaceField(m, gamd), Tf), "cell").numpy()
    sng = m.to_caller(fvcc.FaceNormalGradient(ctx, m).faceNormalGrad(Tf).values, "face", oriented=True).numpy()
    assert relmax(grad, orc.grad(g, T, Tb)) <= TOL64
    assert relmax(div, orc.div(g, phi, T, Tb)) <= TOL64
    assert relmax(lap, orc.laplacian(g, gamma, T, Tb, kinds, Tg)) <= TOL64
    assert relmax(sng, orc.sngrad(g, T, Tb, kinds, Tg)) <= TOL64
    # bitwise against the oracle on the renumbered mesh itself
    cm = m.cell_map.numpy()
    assert np.array_equal(fvcc.GaussGreenGrad(ctx, m).grad(Tf).numpy(), orc.grad(g_ren, T[cm], Tb))
    # the renumbering is a permutation and brings face neighbours close in index
    # (Z-order bounds the mean owner-neighbour distance, not the maximum)
    assert sorted(cm) == list(range(N))
    d0 = np.mean(np.abs(g["owner"] - g["neighbour"]))
    d1 = np.mean(np.abs(g_ren["owner"] - g_ren["neighbour"]))
    assert d1 < d0 / 4


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_staged_and_walk_kernels_agree(ctx, cases, dt):
    """The LDS-staged face-block gather and the per-cell stencil walk give the
    same bits (same arithmetic, same face order)."""
    for name, (poly, m, g) in cases.items():
        assert m.gather_kernel == "staged", name
        staged, _ = _run_all(ctx, m, g, dt)
        m.set_gather(1)
        assert m.gather_kernel == "stencil-walk"
        walk, ref = _run_all(ctx, m, g, dt)
        for op in staged:
            assert np.array_equal(staged[op], walk[op]), (name, op)
            assert np.array_equal(walk[op], ref[op]), (name, op)
        # larger workgroups (partial last blocks on these small meshes)
        for mode, label in ((2, "staged-256"), (4, "staged-1024")):
            m.set_gather(mode)
            assert m.gather_kernel == label
            big, _ = _run_all(ctx, m, g, dt)
            for op in staged:
                assert np.array_equal(big[op], walk[op]), (name, label, op)
        m.set_gather(0)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_staged_gather_box_with_one_cell_tail(ctx, dt):
    """A 17 x 23 x 55 box (21505 cells), jittered, permuted, Z-order renumbered.
    Its sides are not powers of two, so the bricks are ragged: nine in ten still
    fit the staged gather's lists at every workgroup size (the Z-order uses one
    quantisation step on all axes, and the external list holds a record per
    thread).  With 256, 512 and 1024 cells per workgroup the last workgroup holds
    one cell, the highest-numbered, which owns no face, so that workgroup takes
    only its list faces.  Every size is bit-identical to the oracle on the
    renumbered mesh, the few workgroups that do not fit included."""
    import foamadapter_amd as fa
    dims = (17, 23, 55)
    poly = fa.PolyMesh.hex(dims, lengths=tuple(0.1 * d for d in dims), jitter=0.2,
                           jitter_seed=71, perm_seed=72)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    g = orc.geometry(m.poly)
    N = len(g["V"])
    assert N == 21505 and all(N % c == 1 for c in (256, 512, 1024))
    assert not np.any(g["owner"] == N - 1)   # the tail cell owns no internal face
    for mode, cells in ((2, 256), (3, 512), (4, 1024)):
        nb, over = m.gather_stats(cells)
        print(f"[case] {cells}-cell workgroups: {over} of {nb} take the per-cell path", flush=True)
        assert nb == (N + cells - 1) // cells and over <= nb // 10, (cells, nb, over)
        m.set_gather(mode)
        res, ref = _run_all(ctx, m, g, dt, seed=73)
        for op in ("grad", "div_linear", "div_upwind", "surface_integrate", "laplacian"):
            assert np.array_equal(res[op], ref[op]), (cells, op)
    m.set_gather(0)
    m.close()


@pytest.mark.parametrize("method", ["morton", "rcm"])
def test_renumbered_methods_bitwise(ctx, method):
    """Both renumberings: operators bit-identical to the oracle on the renumbered
    polyMesh, and within 1e-12 of the caller's mesh after mapping back."""
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    poly = fa.PolyMesh.hex(16, jitter=0.2, jitter_seed=51, perm_seed=52)
    m = fa.Mesh.from_poly(ctx, poly, renumber=method)
    g_ren = orc.geometry(m.poly)
    g = orc.geometry(poly)
    T, Tb, phi, gamma, Tg = _fields(g, 53)
    cm = m.cell_map.numpy()
    Td = ctx.upload(T[cm])
    bc = fa.BoundaryConditions(m, ["fixedValue"] * 6)
    Tf = fvcc.VolumeField(m, Td, ctx.upload(Tb), bc)
    phid = m.to_device(ctx.upload(phi), "face", oriented=True)
    grad = fvcc.GaussGreenGrad(ctx, m).grad(Tf)
    div = fvcc.GaussGreenDiv(ctx, m).div(None, fvcc.SurfaceField(m, phid), Tf)
    assert np.array_equal(grad.numpy(), orc.grad(g_ren, T[cm], Tb))
    assert np.array_equal(div.numpy(), orc.div(g_ren, phid.numpy(), T[cm], Tb))
    assert relmax(m.to_caller(grad, "cell").numpy(), orc.grad(g, T, Tb)) <= TOL64
    assert relmax(m.to_caller(div, "cell").numpy(), orc.div(g, phi, T, Tb)) <= TOL64


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_unaligned_caller_arrays(ctx, dt):
    """Face arrays the caller passes at an address that is not 16-byte aligned
    (an offset view): the staged gathers and the scalar interpolate/snGrad
    kernels (taken when the result is unaligned) give the same bits as the
    aligned call."""
    import ctypes as C
    import foamadapter_amd as fa
    from foamadapter_amd import _lib
    poly = fa.PolyMesh.hex(14, jitter=0.2, jitter_seed=61, perm_seed=62)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    g = orc.geometry(m.poly)
    T, Tb, phi, gamma, _ = _fields(g, 63)
    N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
    es = np.dtype(dt).itemsize
    prec = _lib.FV_F64 if dt == np.float64 else _lib.FV_F32
    bc = fa.BoundaryConditions(m, ["fixedValue", "zeroGradient"] * 3)
    Td, Tbd = ctx.upload(T.astype(dt)), ctx.upload(Tb.astype(dt))

    def shifted(a):                 # device copy of a, one element past a 16-byte boundary
        buf = ctx.upload(np.concatenate([[0.0], a]).astype(dt))
        return buf, C.c_void_p(buf.ptr.value + es)

    pa, pu = ctx.upload(phi.astype(dt)), shifted(phi)
    ga, gu = ctx.upload(gamma.astype(dt)), shifted(gamma)

    def run(phi_p, gam_p, outs):
        o = {k: ctx.empty(sh, dt) for k, sh in outs.items()}
        o_ptr = {k: v.ptr for k, v in o.items()}
        _lib.call("fv_div_gauss", ctx.handle, m.handle, 0, prec, 1.0, phi_p, Td.ptr, Tbd.ptr,
                  o_ptr["div"])
        _lib.call("fv_div_gauss", ctx.handle, m.handle, 1, prec, 1.0, phi_p, Td.ptr, Tbd.ptr,
                  o_ptr["up"])
        _lib.call("fv_laplacian_uncorrected", ctx.handle, m.handle, prec, 1.0, gam_p, Td.ptr,
                  bc.handle, Tbd.ptr, None, o_ptr["lap"])
        return {k: v.numpy() for k, v in o.items()}

    shapes = {"div": (N,), "up": (N,), "lap": (N,)}
    a = run(pa.ptr, ga.ptr, shapes)
    u = run(pu[1], gu[1], shapes)
    for k in a:
        assert np.array_equal(a[k], u[k]), k
    # an unaligned result of interpolate / snGrad
    obuf = ctx.empty((F + B + 1,), dt)
    optr = C.c_void_p(obuf.ptr.value + es)
    ref = ctx.empty((F + B,), dt)
    _lib.call("fv_interpolate", ctx.handle, m.handle, 0, prec, None, Td.ptr, Tbd.ptr, ref.ptr)
    _lib.call("fv_interpolate", ctx.handle, m.handle, 0, prec, None, Td.ptr, Tbd.ptr, optr)
    assert np.array_equal(obuf.numpy()[1:], ref.numpy())
    _lib.call("fv_sngrad_uncorrected", ctx.handle, m.handle, prec, Td.ptr, bc.handle, Tbd.ptr,
              None, ref.ptr)
    _lib.call("fv_sngrad_uncorrected", ctx.handle, m.handle, prec, Td.ptr, bc.handle, Tbd.ptr,
              None, optr)
    assert np.array_equal(obuf.numpy()[1:], ref.numpy())


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_tiny_meshes_bitwise(ctx, dt):
    """Edge sizes: one cell (no internal face), two cells, a 2x2x2 block and a
    3x1x1 row, as given and renumbered: every operator bit-identical to the
    oracle, and one implicit heat step solved to the oracle's dense solution."""
    import foamadapter_amd as fa
    from foamadapter_amd import dsl, fvcc
    for n in (1, (2, 1, 1), 2, (3, 1, 1)):
        for renumber in (False, True):
            poly = fa.PolyMesh.hex(n)
            print(f"[case] hex {n} renumber={renumber} {np.dtype(dt).name}", flush=True)
            m = fa.Mesh.from_poly(ctx, poly, renumber=renumber)
            g = orc.geometry(m.poly if renumber else poly)
            res, ref = _run_all(ctx, m, g, dt)
            for op in res:
                assert np.array_equal(res[op], ref[op]), (n, renumber, op)
            if dt != np.float64:
                continue
            N, F, B = m.n_cells, m.n_internal_faces, m.n_boundary_faces
            bc = fa.BoundaryConditions(m, ["fixedValue"] * m.n_patches)
            T0 = orc.uniform(N, 300.0, 400.0, 3, 0)
            Tb = orc.uniform(B, 300.0, 400.0, 3, 1)
            T = fvcc.VolumeField(m, ctx.upload(T0), ctx.upload(Tb), bc, "T")
            kap = fvcc.SurfaceField(m, ctx.upload(np.full(F + B, 0.5)))
            st = dsl.solve(dsl.imp.ddt(T) - dsl.imp.laplacian(kap, T), T, 0.0, 1e-2,
                           solver={"relTol": 1e-12, "maxIter": 100})
            ls = orc.LinSys(g, orc.sparsity(g))
            ls.ddt(T0, 1e-2)
            ls.laplacian(np.full(F + B, 0.5), Tb, [0] * m.n_patches, coeff=-1.0)
            sp = orc.sparsity(g)
            A = np.zeros((N, N))
            for r in range(N):
                A[r, sp["colIdxs"][sp["rowOffs"][r]:sp["rowOffs"][r + 1]]] = \
                    ls.values[sp["rowOffs"][r]:sp["rowOffs"][r + 1]]
            assert st.finalResNorm <= 1e-12
            assert relmax(T.numpy(), np.linalg.solve(A, ls.rhs)) < 1e-11, (n, renumber)
