"""Parity at production shapes: the staged-gather fast path at full-brick scale.

The sweep's configuration -- a jittered, randomly numbered hex mesh renumbered
at import along a Z-order curve, the default workgroup size of each operator,
16-byte face streams -- runs here on a 64^3 mesh (262,144 cells, 1,024
workgroups of 256 cells), where every workgroup is a full brick and none takes
the per-cell path (checked through fv_mesh_gather_stats).  Every operator of
the sweep (test/test_operators.cpp:52-131, benchmarks/bench_explicitOperators.cpp)
is compared bit for bit with the oracle: FP64 with the multi-core restatement
(whose bits are the serial oracle's), FP32 with the serial FP32 oracle.
"""
import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu

N_EDGE = 64


@pytest.fixture(scope="module")
def big(ctx):
    import foamadapter_amd as fa
    print(f"[case] {N_EDGE}^3 jittered, permuted, Z-order renumbered mesh", flush=True)
    poly = fa.PolyMesh.hex(N_EDGE, jitter=0.2, jitter_seed=1002, perm_seed=1002)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    g = orc.geometry(m.poly)
    yield m, g
    m.close()


def test_geometry_bitwise_at_scale(big):
    m, g = big
    for k in ("V", "C", "Sf", "Cf", "magSf", "weights", "deltaCoeffs", "nonOrthDeltaCoeffs",
              "Sf_b", "nf_b", "delta_b", "deltaCoeffs_b", "nonOrthDeltaCoeffs_b"):
        assert np.array_equal(m.get(k), g[k]), k


def test_no_workgroup_takes_the_per_cell_path(big):
    """Every workgroup of every staged-gather size fits its LDS lists: the fast
    path is what the parity tests below exercise."""
    m, g = big
    N = len(g["V"])
    assert m.gather_kernel == "staged"
    for cells in (256, 512, 1024):
        nb, over = m.gather_stats(cells)
        assert nb == (N + cells - 1) // cells and over == 0, (cells, nb, over)


def _sweep_fields(g, seed=1002):
    """The sweep's inputs: T ~ U[-1,1], zeroGradient walls (T_b = T_P), phi =
    Sf . interpolate(U) with U_i ~ U[-1,1] (U_b = U_P), gamma ~ U[0.5,2]."""
    N, F, B = len(g["V"]), len(g["owner"]), len(g["faceCells"])
    fc = g["faceCells"]
    T = orc.uniform(N, -1.0, 1.0, seed, 0)
    U = orc.uniform(3 * N, -1.0, 1.0, seed, 1).reshape(N, 3)
    gamma = orc.uniform(F + B, 0.5, 2.0, seed, 2)
    phi = orc.flux(g, U, np.ascontiguousarray(U[fc]))
    return T, T[fc].copy(), phi, gamma


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_sweep_operators_bitwise_at_scale(ctx, big, dt):
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    m, g = big
    name = np.dtype(dt).name
    T, Tb, phi, gamma = _sweep_fields(g)
    kinds = [1] * m.n_patches                       # zeroGradient
    bc = fa.BoundaryConditions(m, ["zeroGradient"] * m.n_patches)
    Tf = fvcc.VolumeField(m, ctx.upload(T.astype(dt)), ctx.upload(Tb.astype(dt)), bc)
    phiF = fvcc.SurfaceField(m, ctx.upload(phi.astype(dt)))
    gamF = fvcc.SurfaceField(m, ctx.upload(gamma.astype(dt)))
    nFB = m.n_internal_faces + m.n_boundary_faces
    got = {}
    out = fvcc.SurfaceField(m, ctx.empty((nFB,), dt))
    fvcc.SurfaceInterpolation(ctx, m, ["linear"]).interpolate(Tf, out)
    got["interp_linear"] = out.numpy()
    got["grad"] = fvcc.GaussGreenGrad(ctx, m).grad(Tf).numpy()
    got["div_linear"] = fvcc.GaussGreenDiv(ctx, m, ["Gauss", "linear"]).div(None, phiF, Tf).numpy()
    got["div_upwind"] = fvcc.GaussGreenDiv(ctx, m, ["Gauss", "upwind"]).div(None, phiF, Tf).numpy()
    got["laplacian"] = fvcc.GaussGreenLaplacian(ctx, m).laplacian(None, gamF, Tf).numpy()
    got["sngrad"] = fvcc.FaceNormalGradient(ctx, m).faceNormalGrad(Tf).numpy()
    if dt == np.float64:
        omp = orc.OmpOps(g)
        ref = {"interp_linear": omp.interpolate(T, Tb), "grad": omp.grad(T, Tb),
               "div_linear": omp.div(phi, T, Tb, 0), "div_upwind": omp.div(phi, T, Tb, 1),
               "laplacian": omp.laplacian(gamma, T, Tb, kinds), "sngrad": omp.sngrad(T, Tb, kinds)}
        omp.close()
    else:
        ref = {"interp_linear": orc.interpolate(g, T, Tb, 0, None, dt),
               "grad": orc.grad(g, T, Tb, dt), "div_linear": orc.div(g, phi, T, Tb, 0, dt),
               "div_upwind": orc.div(g, phi, T, Tb, 1, dt),
               "laplacian": orc.laplacian(g, gamma, T, Tb, kinds, None, dt),
               "sngrad": orc.sngrad(g, T, Tb, kinds, None, dt)}
    for op in ref:
        print(f"[case] {N_EDGE}^3 {op} {name}", flush=True)
        assert got[op].dtype == dt
        assert np.array_equal(got[op], ref[op]), (op, name, np.mean(got[op] == ref[op]))
    bc.close()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_mixed_boundary_kinds_bitwise_at_scale(ctx, big, dt):
    """Every operator with fixedValue / fixedGradient / calculated patches and
    random boundary values, interpolate upwind and surfaceIntegrate included."""
    from test_gpu_faceops import _run_all
    m, g = big
    res, ref = _run_all(ctx, m, g, dt, seed=17)
    for op in res:
        assert np.array_equal(res[op], ref[op]), (op, np.dtype(dt).name)


def test_brick_order_does_not_change_bits(ctx, big):
    """The XCD-chunked order of the workgroups (runs of 1, 3 and 8 bricks per
    XCD; 3 leaves padding workgroups that exit at once) gives the same bits."""
    import foamadapter_amd as fa
    from foamadapter_amd import fvcc
    m, g = big
    T, Tb, phi, gamma = _sweep_fields(g, seed=1003)
    bc = fa.BoundaryConditions(m, ["zeroGradient"] * m.n_patches)
    Tf = fvcc.VolumeField(m, ctx.upload(T), ctx.upload(Tb), bc)
    phiF = fvcc.SurfaceField(m, ctx.upload(phi))
    gamF = fvcc.SurfaceField(m, ctx.upload(gamma))
    res = {}
    for run in (1, 3, 8):
        print(f"[case] {N_EDGE}^3 gather run {run}", flush=True)
        m.set_gather_run(run)
        res[run] = [fvcc.GaussGreenGrad(ctx, m).grad(Tf).numpy(),
                    fvcc.GaussGreenDiv(ctx, m, ["Gauss", "linear"]).div(None, phiF, Tf).numpy(),
                    fvcc.GaussGreenLaplacian(ctx, m).laplacian(None, gamF, Tf).numpy()]
    m.set_gather_run(8)
    for run in (1, 3):
        for a, b in zip(res[run], res[8]):
            assert np.array_equal(a, b), run
    bc.close()
