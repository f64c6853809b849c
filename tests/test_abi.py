"""The C-ABI library loads and exports every function include/fvhip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "fvhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fv_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    from foamadapter_amd import _lib
    names = header_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(_lib.declared_symbols()), set(names) - set(_lib.declared_symbols())


def test_error_path_without_gpu():
    from foamadapter_amd import _lib
    rc = _lib.lib.fv_ctx_create(0, 0, None)
    assert rc == _lib.FV_EINVAL
    assert "NULL" in _lib.last_error()


def test_version():
    from foamadapter_amd import _lib
    assert b"gfx950" in _lib.lib.fv_version()


def test_hexmesh_host_generator():
    """The blockMesh-style generator is host code: counts, ordering, orientation."""
    from foamadapter_amd.mesh import PolyMesh
    n = 5
    p = PolyMesh.hex(n, jitter=0.2, jitter_seed=1, perm_seed=3)
    assert p.n_cells == n ** 3 and p.n_internal_faces == 3 * n * n * (n - 1)
    assert p.n_faces == p.n_internal_faces + 6 * n * n
    own, nbr = p.owner[: p.n_internal_faces], p.neighbour
    assert np.all(own < nbr)
    key = own.astype(np.int64) * (n ** 3) + nbr
    assert np.all(np.diff(key) > 0)           # upper-triangular order
    assert sorted(set(own) | set(nbr)) == list(range(n ** 3))
    # same seeds -> same mesh; other permutation seed -> other numbering
    q = PolyMesh.hex(n, jitter=0.2, jitter_seed=1, perm_seed=3)
    assert np.array_equal(p.points, q.points) and np.array_equal(p.face_points, q.face_points)
    r = PolyMesh.hex(n, jitter=0.2, jitter_seed=1, perm_seed=4)
    assert not np.array_equal(p.owner, r.owner)
    # jitter moves interior points only
    u = PolyMesh.hex(n)
    moved = np.any(p.points != u.points, axis=1)
    inner = np.all((u.points > 0) & (u.points < 1), axis=1)
    assert np.array_equal(moved, inner)
