"""fv_courant_number_noalloc: the Courant number for callers that cannot
allocate, through the C ABI, against the allocating fv_courant_number (and the
oracle) on two calls in a row, the second on a smaller mesh, so the shared
scratch array still holds the first call's larger contents.  Bar: bitwise
(a maximum and one product, the same expression on both paths)."""
import ctypes as C

import numpy as np
import pytest

import oracle_cpu as orc

pytestmark = pytest.mark.gpu


def _case(ctx, n, seed):
    import foamadapter_amd as fa
    poly = fa.PolyMesh.hex(n, jitter=0.2, jitter_seed=seed, perm_seed=seed)
    m = fa.Mesh.from_poly(ctx, poly, renumber=True)
    phi = ctx.uniform(m.n_internal_faces + m.n_boundary_faces, -1.0, 1.0, seed, 3)
    return m, phi


def test_noalloc_courant_matches_allocating_entry(ctx):
    from foamadapter_amd import _lib
    out = ctx.zeros((2,))
    for i, (n, seed, dt) in enumerate([(40, 3, 1e-3), (12, 4, 2.5e-3)]):
        print(f"[case] noalloc courant call {i}: {n}^3", flush=True)
        m, phi = _case(ctx, n, seed)
        ref = C.c_double()
        _lib.call("fv_courant_number", ctx.handle, m.handle, phi.ptr, dt, C.byref(ref))
        ctx.sync()
        used = ctx.bytes_in_use
        _lib.call("fv_courant_number_noalloc", ctx.handle, m.handle, phi.ptr, dt,
                  C.c_void_p(out.ptr.value + 8 * i))
        ctx.sync()
        assert ctx.bytes_in_use == used                  # nothing taken from the pool
        got = float(out.numpy()[i])
        assert got == ref.value, (i, got, ref.value)
        assert got == orc.courant(orc.geometry(m.poly), phi.numpy(), dt), i
        phi.close()
        m.close()
    assert out.numpy()[0] != out.numpy()[1]


def test_noalloc_courant_rejects_bad_arguments(ctx):
    from foamadapter_amd import FvError, _lib
    m, phi = _case(ctx, 4, 5)
    with pytest.raises(FvError):
        _lib.call("fv_courant_number_noalloc", ctx.handle, m.handle, phi.ptr, 1e-3, None)
    m.close()
